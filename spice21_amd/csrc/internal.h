// Host-side internal interfaces of the engine (not part of the C ABI).
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/spice21_mi355x.h"
#include "models.h"

namespace s21 {

// Errors carry an S21_* code and the reference's SpError text.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& s) : std::runtime_error(s), code(c) {}
};
[[noreturn]] inline void fail(int code, const std::string& s) { throw Error(code, s); }
void set_last_error(const std::string& s);

// ---- parameter derivation (derive.cpp) ----
struct Mos1Derived {
    Mos1Core core;
    Mos1Caps caps;
    bool has_caps;     // any gate capacitance (Meyer or overlap) is nonzero
    bool has_rd, has_rs;  // model carries rd/rsh resp. rs/rsh (internal nodes exist)
};
Mos1Derived derive_mos1(const s21_params& model, const s21_params& inst, const s21_options& o);
DiodeP derive_diode(const s21_params& model, const s21_params& inst, const s21_options& o);
bool diode_model_has_rs(const s21_params& model);

// ---- flattened component after analysis-specific augmentation ----
struct Comp {
    int type;
    int node[6];
    int model, inst;
    double value, value_tran, acm;  // value_tran differs from value only for the IC forcing elements
};

// ---- per-topology structure (topology.cpp) ----
struct Topology {
    int n = 0;
    std::vector<std::string> names;
    std::vector<int> kinds;
    std::vector<Comp> comps;
    bool tran = false;  // built for a transient (capacitor and charge entries live)

    // matrix elements in Matrix::make creation order (external coordinates)
    std::vector<int> erow, ecol;

    // pivot order of the reference's first factorisation
    std::vector<int> row_i2e, row_e2i, col_i2e, col_e2i;

    // L+U pattern in internal coordinates, CSR by row, columns increasing
    int nnz = 0;
    std::vector<int> rp, ci, diag;
    std::vector<int> elem_pos;       // element -> LU position
    std::vector<uint8_t> is_a;       // LU position holds a matrix element (not pure fill)
    std::vector<int> upd_ptr;        // per LU position (L entries only non-empty)
    std::vector<int> upd_src, upd_dst;
    std::vector<int> lvl_ptr, lvl_rows;    // row levels of the factorisation / forward sweep
    std::vector<int> blvl_ptr, blvl_rows;  // row levels of the backward sweep
    // Pipelined factorisation / forward schedule: level l holds chunks (row, first L position, end L
    // position); a chunk runs once the pivot rows of its L entries are complete, so a long row advances
    // together with the rows it depends on instead of waiting for all of them.
    std::vector<int> clvl_ptr, ch_row, ch_q0, ch_q1;
    std::vector<int> res_ptr, res_pos, res_xcol;  // residual rows (internal) over A entries

    // devices by type, in component order within each type
    std::vector<int> m1_comp, m0_comp, d_comp, lin_comp;
    std::vector<int> m1_nodes;   // [6 per device]
    std::vector<uint32_t> m1_mask;
    std::vector<int> m1_base;
    std::vector<uint8_t> m1_caps;  // any instance has gate capacitance
    std::vector<int> m0_nodes;   // [4 per device]
    std::vector<uint32_t> m0_mask;
    std::vector<int> m0_base;
    std::vector<double> m0_p;
    std::vector<int> d_nodes;    // [3 per device]
    std::vector<uint32_t> d_mask;
    std::vector<int> d_base;
    std::vector<int> lin_type, lin_nodes;  // [3 per device]
    std::vector<uint32_t> lin_mask;
    std::vector<int> lin_base;

    // Kernel records (16-bit fields, packed in fours or pairs), built by topo_records:
    //   gt {slot, k0, k1, part}   gather tasks over g_idx[k0, k1); part = 0xFFFF writes the slot itself
    //   gc {slot, p0, p1, 0}      sums of the partials of a split slot
    //   rt {row, e0, e1, part}    residual tasks over re[e0, e1); re {LU position, x column}
    //   rc {row, p0, p1, 0}       sums of the partials of a split residual row
    //   fch {row, l0, l1, u0}     factorisation chunks, level order (clvl_ptr)
    //   fl {q, pivot position, k, n_updates}; fu {src, dst}
    //   bro {row, diag, e0, e1}   backward-sweep rows, level order (blvl_ptr); bu {LU position, column}
    std::vector<uint16_t> gt, gc, rt, re, rc, fch, fl, fu, bro, bu;
    int n_part = 0;  // partial-sum slots needed by the split gather / residual tasks

    int n_contrib = 0;
    int sym_status = S21_OK;  // failure of the pivot search, reported when an instance needs a solve
    std::string sym_error;
    std::vector<int> g_ptr, g_idx;  // gather lists for nnz LU slots followed by n RHS slots
    std::vector<int> contrib_comp;  // component of each contribution (introspection)
    int64_t flops_lu = 0;
};

// Builds a topology from a flattened netlist (already IC-augmented when tran); `first_values` are the
// matrix values of the reference's first Newton iteration, in element order, used for pivoting.
void topo_elements(Topology& t);
void topo_pivot(Topology& t, const std::vector<double>& first_values);
void topo_pivot_search(Topology& t, const std::vector<double>& first_values);
void topo_schedule(Topology& t, const std::vector<uint8_t>& mos1_has_caps);
void topo_records(Topology& t);

// Reference-order OP assembly on the host (analysis.rs:153-168) at voltages x for the pivot search.
struct HostInstance {
    std::vector<Mos1Derived> m1;
    std::vector<DiodeP> dio;
    std::vector<double> lin_value;  // per linear component, operating-point value
};
void host_assemble_op(const Topology& t, const HostInstance& h, const s21_options& o, const std::vector<double>& x,
                      std::vector<double>& vals, std::vector<double>& rhs);

}  // namespace s21
