// Host-side internal interfaces of the engine (not part of the C ABI).
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/spice21_mi355x.h"
#include "bsim4.h"
#include "layout.h"
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
constexpr int HUB_CONTRIB = 4;
struct Topology {
    int n = 0;
    std::vector<std::string> names;
    std::vector<int> kinds;
    std::vector<Comp> comps;
    bool tran = false;  // built for a transient (capacitor and charge entries live)
    bool ac = false;    // built for AC: small-signal stamps (capacitance entries live, no Mos1 RHS)

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

    // Devices in kernel order: classes Mos1, Mos0, Diode, linear (R/C/V/I); within a class by colour,
    // then component order.  Devices of one colour share no matrix / RHS slot other than hub slots, so
    // the kernel adds a colour's stamps straight into A and the RHS; a hub slot (one with more than
    // HUB_CONTRIB contributing components) keeps every contribution apart and sums them in component
    // order afterwards.
    std::vector<int> dev_comp;            // component of each device position
    int cls_start[N_CLS + 1] = {0};       // device positions of each class
    std::vector<int> col_ptr, col_cls;    // colours: device-position ranges [col_ptr[c], col_ptr[c+1]) and class
    std::vector<uint16_t> dev_rec;        // per device: node + 1 for six terminals, flags, 0
    std::vector<uint16_t> codes;          // per device entry: slot in A|RHS, 0x8000 | hub index, or 0xFFFF
    int code_base[N_CLS] = {0};           // first code of each class (stride code_stride(class))
    int n_hub = 0;                        // hub contributions
    std::vector<int> hub_slot;            // hub slots, ascending
    std::vector<int> g_ptr, g_idx;        // per hub slot: its hub contributions in component-then-entry order

    // Kernel records (16-bit fields, packed in fours or pairs), built by topo_records:
    //   gt {slot, k0, k1, part}   hub gather tasks over g_idx[k0, k1); part = 0xFFFF writes the slot
    //   gc {slot, p0, p1, 0}      sums of the partials of a split hub slot
    //   rt {row, e0, e1, part}    residual tasks over re[e0, e1); re {LU position, x column}
    //   rc {row, p0, p1, 0}       sums of the partials of a split residual row
    //   fch {row, l0, l1, u0}     factorisation chunks, level order (clvl_ptr)
    //   fl {q, pivot position, k, n_updates}; fu {src, dst}
    //   bro {row, diag, e0, e1}   backward-sweep rows, level order (blvl_ptr); bu {LU position, column}
    std::vector<uint16_t> gt, gc, rt, re, rc, fch, fl, fu, bro, bu;
    int n_part = 0;  // partial-sum slots needed by the split gather / residual tasks

    int sym_status = S21_OK;  // failure of the pivot search, reported when an instance needs a solve
    std::string sym_error;
    int64_t flops_lu = 0;
};

// Components of a flattened netlist for one analysis; a transient appends its initial-condition forcing
// (a 1 S resistor and a source per IC, then 1e-9 S and 0 V after the operating point).
void build_comps(int n_vars, const std::vector<std::string>& names, const std::vector<int>& kinds,
                 const std::vector<s21_comp>& comps, int n_m1_models, int n_m1_insts, int n_d_models, int n_d_insts,
                 const s21_tran_options* tr, Topology& t);
// Builds a topology from a flattened netlist (already IC-augmented when tran); `first_values` are the
// matrix values of the reference's first Newton iteration, in element order, used for pivoting.
void topo_elements(Topology& t);
void topo_pivot(Topology& t, const std::vector<double>& first_values);
void topo_pivot_search(Topology& t, const std::vector<double>& first_values);
// A given pivot order (internal index -> external row / column) and its fill pattern.
// The reference's searches of several factorisations in turn: the first on first_values, each later one
// on the next value set, from the arrangement and fill the previous search left; the last order is kept.
void topo_pivot_sequence(Topology& t, const std::vector<double>& first_values,
                         const std::vector<std::vector<double>>& later_values);
void topo_pivot_forced(Topology& t, const std::vector<int>& row_i2e, const std::vector<int>& col_i2e);
// The same on the complex matrix of the first AC iteration (Matrix<Complex<f64>>: magnitudes are norms).
void topo_pivot_cx(Topology& t, const std::vector<double>& re, const std::vector<double>& im);
// hub_contrib: slots with more contributing components are hubs (1: every shared slot, so each device
// class evaluates in one colour round -- for a latency-bound launch of few instances).
void topo_schedule(Topology& t, const std::vector<uint8_t>& mos1_has_caps, int hub_contrib = HUB_CONTRIB);
void topo_records(Topology& t);

// Reference-order OP assembly on the host (analysis.rs:153-168) at voltages x for the pivot search.
struct HostInstance {
    std::vector<Mos1Derived> m1;
    std::vector<DiodeP> dio;
    std::vector<b4::B4Par> b4;      // per BSIM4 in component order
    std::vector<double> lin_value;  // per linear component, operating-point value
    std::vector<double> dio_vd;     // per diode, the guess its limiting starts from (empty: 0, as at x = 0)
};
void host_assemble_op(const Topology& t, const HostInstance& h, const s21_options& o, const std::vector<double>& x,
                      std::vector<double>& vals, std::vector<double>& rhs);
// Reference-order AC assembly (analysis.rs:237-251) at angular frequency omega for the pivot search:
// acop[k] / grd[k] / grs[k] per Mos1 in component order, lin_ac per linear component (R: g, C: c,
// V: acm).
void host_assemble_ac(const Topology& t, const std::vector<Mos1AcOp>& acop, const std::vector<double>& grd,
                      const std::vector<double>& grs, const std::vector<double>& lin_ac, double omega,
                      std::vector<double>& re, std::vector<double>& im);

}  // namespace s21
