// Host symbolic phase: the frozen structure the GPU Newton kernel runs on.
//
//  * topo_elements   -- matrix elements in the order Component::create_matrix_elems makes them
//                       (comps/mod.rs:127-132, 190-195, 291-298; comps/mos.rs:896-903, 1044-1050;
//                       comps/diode.rs:248-256; comps/bsim4/bsim4solver.rs:28-116), deduplicated as
//                       Matrix::make does (sparse21/mod.rs:265).
//  * topo_pivot      -- the reference's pivot choice on the first Newton iteration's matrix: Markowitz
//                       diagonal search with relative threshold 1e-3 and the x5 tie cut-off, the
//                       column-n submatrix search, then the global maximum; the last index is never
//                       searched (sparse21/mod.rs:647-919).  Restated over permutation arrays instead of
//                       orthogonal linked lists, visiting candidates in the same order.
//  * topo_schedule   -- L+U pattern in internal order, IKJ update lists, row levels for the
//                       factorisation and both sweeps, device colours, per-entry slot codes and hub gather lists.
#include <algorithm>
#include <array>
#include <cmath>
#include <numeric>
#include <unordered_map>

#include "bsim4_eval.h"
#include "internal.h"

namespace s21 {
namespace {

inline int64_t key(int r, int c, int n) { return (int64_t)r * (int64_t)(n + 1) + c; }

// Terminal-role -> variable of one component (ground = S21_GND).
inline int mos1_role_node(const Comp& c, int role) { return c.node[role]; }

}  // namespace

void topo_elements(Topology& t) {
    std::unordered_map<int64_t, int> made;
    t.erow.clear();
    t.ecol.clear();
    auto make = [&](int r, int c) {
        if (r < 0 || c < 0) return;
        int64_t k = key(r, c, t.n);
        if (made.count(k)) return;
        made[k] = (int)t.erow.size();
        t.erow.push_back(r);
        t.ecol.push_back(c);
    };
    for (const Comp& c : t.comps) {
        const int* nd = c.node;
        switch (c.type) {
        case S21_COMP_V:
            make(nd[0], nd[2]);
            make(nd[2], nd[0]);
            make(nd[1], nd[2]);
            make(nd[2], nd[1]);
            break;
        case S21_COMP_C:
            make(nd[0], nd[0]);
            make(nd[0], nd[1]);
            make(nd[1], nd[0]);
            make(nd[1], nd[1]);
            break;
        case S21_COMP_R:
            make(nd[0], nd[0]);
            make(nd[0], nd[1]);
            make(nd[1], nd[0]);
            make(nd[1], nd[1]);
            break;
        case S21_COMP_I: break;
        case S21_COMP_MOS1: {
            static const int order[6] = {RG, RD, RS, RB, RDP, RSP};
            for (int a : order)
                for (int b : order) make(mos1_role_node(c, a), mos1_role_node(c, b));
            break;
        }
        case S21_COMP_MOS0: {
            static const int pr[6][2] = {{RD, RD}, {RS, RS}, {RD, RS}, {RS, RD}, {RD, RG}, {RS, RG}};
            for (auto& q : pr) make(nd[q[0]], nd[q[1]]);
            break;
        }
        case S21_COMP_DIODE:
            make(nd[0], nd[0]);
            make(nd[0], nd[2]);
            make(nd[2], nd[0]);
            make(nd[2], nd[2]);
            make(nd[1], nd[2]);
            make(nd[2], nd[1]);
            make(nd[1], nd[1]);
            break;
        case S21_COMP_BSIM4: {  // create_matps with every internal node on its terminal, no charge node
            enum { D = 0, G = 1, S = 2, B = 3 };
            static const int order[22][2] = {{D, B}, {G, B}, {S, B}, {B, D}, {B, G}, {B, S}, {B, B}, {D, D},
                                             {G, G}, {S, S}, {D, D}, {S, S}, {D, D}, {G, D}, {G, S}, {S, S},
                                             {D, S}, {D, D}, {D, G}, {S, G}, {S, S}, {S, D}};
            for (auto& q : order) make(nd[q[0]], nd[q[1]]);
            break;
        }
        default: fail(S21_INVALID, "Invalid Component");
        }
    }
}

// Flattened components of one analysis, IC forcing appended for a transient (Tran::ic,
// analysis.rs:510-525).
void build_comps(int n_vars, const std::vector<std::string>& names, const std::vector<int>& kinds,
                 const std::vector<s21_comp>& comps, int n_m1_models, int n_m1_insts, int n_d_models, int n_d_insts,
                 const s21_tran_options* tr, Topology& t) {
    t.n = n_vars;
    t.names = names;
    t.kinds = kinds;
    t.comps.clear();
    t.tran = tr != nullptr;
    for (const s21_comp& c : comps) {
        Comp k{};
        k.type = c.type;
        for (int j = 0; j < 6; j++) k.node[j] = c.node[j];
        k.model = c.model;
        k.inst = c.inst;
        k.value = c.value;
        k.value_tran = c.value;
        k.acm = c.acm;
        if (c.type == S21_COMP_MOS1) {
            if (c.model < 0 || c.model >= n_m1_models || c.inst < 0 || c.inst >= n_m1_insts)
                fail(S21_ERR, "Parameters not defined");
        } else if (c.type == S21_COMP_DIODE) {
            if (c.model < 0 || c.model >= n_d_models || c.inst < 0 || c.inst >= n_d_insts)
                fail(S21_ERR, "Parameters not defined");
        }
        for (int j = 0; j < 6; j++)
            if (k.node[j] >= n_vars || k.node[j] < S21_GND) fail(S21_INVALID, "component node out of range");
        t.comps.push_back(k);
    }
    if (tr) {
        for (int k = 0; k < tr->n_ic; k++) {
            const int node = tr->ic_var[k];
            if (node < 0 || node >= n_vars) fail(S21_INVALID, "initial condition node out of range");
            std::string nm = tr->ic_name ? tr->ic_name[k] : names[node];
            const int fnode = t.n++, ivar = t.n++;
            t.names.push_back("." + nm + ".vic");
            t.kinds.push_back(S21_VAR_V);
            t.names.push_back("." + nm + ".iic");
            t.kinds.push_back(S21_VAR_I);
            Comp r{};
            r.type = S21_COMP_R;
            r.node[0] = fnode;
            r.node[1] = node;
            r.node[2] = r.node[3] = r.node[4] = r.node[5] = S21_GND;
            r.value = 1.0;
            r.value_tran = 1e-9;  // analysis.rs:547-549
            t.comps.push_back(r);
            Comp v{};
            v.type = S21_COMP_V;
            v.node[0] = fnode;
            v.node[1] = S21_GND;
            v.node[2] = ivar;
            v.node[3] = v.node[4] = v.node[5] = S21_GND;
            v.value = tr->ic_val[k];
            v.value_tran = 0.0;  // analysis.rs:544-546
            t.comps.push_back(v);
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Host assembly in the reference's order (analysis.rs:153-168): components in order, stamps in order.
// ------------------------------------------------------------------------------------------------
void host_assemble_op(const Topology& t, const HostInstance& h, const s21_options& o, const std::vector<double>& x,
                      std::vector<double>& vals, std::vector<double>& rhs) {
    std::unordered_map<int64_t, int> elem;
    for (size_t k = 0; k < t.erow.size(); k++) elem[key(t.erow[k], t.ecol[k], t.n)] = (int)k;
    vals.assign(t.erow.size(), 0.0);
    rhs.assign(t.n, 0.0);
    auto X = [&](int v) { return v < 0 ? 0.0 : x[v]; };
    auto addm = [&](int r, int c, double v) {
        if (r < 0 || c < 0) return;
        vals[elem.at(key(r, c, t.n))] += v;
    };
    auto addb = [&](int r, double v) {
        if (r >= 0) rhs[r] += v;
    };
    int im1 = 0, id = 0, il = 0, ib4 = 0;
    for (const Comp& c : t.comps) {
        const int* nd = c.node;
        switch (c.type) {
        case S21_COMP_BSIM4: {  // the first load of a fresh device: its guess is all zero
            const b4::B4Par& P = h.b4[ib4++];
            b4::B4Op N;
            b4::op_eval(P, b4::B4Op{}, X(nd[0]), X(nd[1]), X(nd[2]), X(nd[3]), o.gmin, N);
            b4::stamp(P, N, b4::B4Tran{}, [&](int k, double v) {
                const int r = nd[b4::b4_row(k)];
                if (b4::b4_col(k) < 0)
                    addb(r, v);
                else
                    addm(r, nd[b4::b4_col(k)], v);
            });
            break;
        }
        case S21_COMP_MOS1: {
            const Mos1Derived& D = h.m1[im1++];
            mos1_load<false>(D.core, nullptr, nullptr, X(nd[RD]), X(nd[RG]), X(nd[RS]), X(nd[RB]), o.gmin, 0.0,
                             [&](int e, double v) {
                                 const int r = nd[mos1_entry_row(e)];
                                 if (mos1_entry_col(e) < 0)
                                     addb(r, v);
                                 else
                                     addm(r, nd[mos1_entry_col(e)], v);
                             },
                             nullptr);
            break;
        }
        case S21_COMP_MOS0: {
            const double p = c.model == 1 ? -1.0 : 1.0;
            mos0_load(p, X(nd[RD]), X(nd[RG]), X(nd[RS]), o.gmin, [&](int e, double v) {
                const int r = nd[mos0_entry_row(e)];
                if (mos0_entry_col(e) < 0)
                    addb(r, v);
                else
                    addm(r, nd[mos0_entry_col(e)], v);
            });
            break;
        }
        case S21_COMP_DIODE: {
            const double vd_guess = h.dio_vd.empty() ? 0.0 : h.dio_vd[id];
            diode_load<false>(h.dio[id++], X(nd[DR_]), X(nd[DN_]), o.gmin, 0.0, vd_guess, 0.0,
                              [&](int e, double v) {
                                  const int r = nd[diode_entry_row(e)];
                                  if (diode_entry_col(e) < 0)
                                      addb(r, v);
                                  else
                                      addm(r, nd[diode_entry_col(e)], v);
                              },
                              nullptr);
            break;
        }
        case S21_COMP_R: {
            const double g = h.lin_value[il++];
            addm(nd[0], nd[0], g);
            addm(nd[1], nd[1], g);
            addm(nd[0], nd[1], -g);
            addm(nd[1], nd[0], -g);
            break;
        }
        case S21_COMP_C: il++; break;  // nothing in the operating point (comps/mod.rs:205-209)
        case S21_COMP_V: {
            const double vv = h.lin_value[il++];
            addm(nd[0], nd[2], 1.0);
            addm(nd[2], nd[0], 1.0);
            addm(nd[1], nd[2], -1.0);
            addm(nd[2], nd[1], -1.0);
            addb(nd[2], vv);
            break;
        }
        case S21_COMP_I: {
            const double i = h.lin_value[il++];
            addb(nd[0], i);
            addb(nd[1], -i);
            break;
        }
        }
    }
}

void host_assemble_ac(const Topology& t, const std::vector<Mos1AcOp>& acop, const std::vector<double>& grd,
                      const std::vector<double>& grs, const std::vector<double>& lin_ac, double omega,
                      std::vector<double>& re, std::vector<double>& im) {
    std::unordered_map<int64_t, int> elem;
    for (size_t k = 0; k < t.erow.size(); k++) elem[key(t.erow[k], t.ecol[k], t.n)] = (int)k;
    re.assign(t.erow.size(), 0.0);
    im.assign(t.erow.size(), 0.0);
    auto addm = [&](int r, int c, double vr, double vi) {
        if (r < 0 || c < 0) return;
        const int e = elem.at(key(r, c, t.n));
        re[e] += vr;
        im[e] += vi;
    };
    int im1 = 0, il = 0;
    for (const Comp& c : t.comps) {
        const int* nd = c.node;
        switch (c.type) {
        case S21_COMP_MOS1: {
            const int k = im1++;
            mos1_load_ac(acop[k], grd[k], grs[k], omega, [&](int e, double vr, double vi) {
                addm(nd[mos1_entry_row(e)], nd[mos1_entry_col(e)], vr, vi);
            });
            break;
        }
        case S21_COMP_R: {  // comps/mod.rs:311-322
            const double g = lin_ac[il++];
            addm(nd[0], nd[0], g, 0.0);
            addm(nd[1], nd[1], g, 0.0);
            addm(nd[0], nd[1], -g, 0.0);
            addm(nd[1], nd[0], -g, 0.0);
            break;
        }
        case S21_COMP_C: {  // comps/mod.rs:223-238
            const double wc = omega * lin_ac[il++];
            addm(nd[0], nd[0], 0.0, wc);
            addm(nd[1], nd[1], 0.0, wc);
            addm(nd[0], nd[1], 0.0, -wc);
            addm(nd[1], nd[0], 0.0, -wc);
            break;
        }
        case S21_COMP_V:  // comps/mod.rs:139-150 (the RHS does not enter the pivot search)
            il++;
            addm(nd[0], nd[2], 1.0, 0.0);
            addm(nd[2], nd[0], 1.0, 0.0);
            addm(nd[1], nd[2], -1.0, 0.0);
            addm(nd[2], nd[1], -1.0, 0.0);
            break;
        default:
            fail(S21_ERR, "AC Not Implemented For This Component!");
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Pivot search (restated from sparse21/mod.rs:647-919).
// ------------------------------------------------------------------------------------------------
namespace {
// Complex values of the AC matrix, with the arithmetic of num-complex (naive division) that the
// reference's Matrix<Complex<f64>> uses; magnitudes are the norm (hypot).
struct HCx {
    double re = 0.0, im = 0.0;
};
inline HCx operator+(HCx a, HCx b) { return {a.re + b.re, a.im + b.im}; }
inline HCx operator-(HCx a, HCx b) { return {a.re - b.re, a.im - b.im}; }
inline HCx operator*(HCx a, HCx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline HCx operator/(HCx a, HCx b) {
    const double ns = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / ns, (a.im * b.re - a.re * b.im) / ns};
}
inline double mag(double v) { return std::fabs(v); }
inline double mag(HCx v) { return std::hypot(v.re, v.im); }
inline bool is_zero(double v) { return v == 0.0; }
inline bool is_zero(HCx v) { return v.re == 0.0 && v.im == 0.0; }

template <class T>
struct PivotState {
    int n;
    std::vector<int> orow, ocol;     // element -> external (original) row / col
    std::vector<T> val;
    std::vector<std::vector<int>> rows, cols;  // external row / col -> element ids
    std::unordered_map<int64_t, int> at;       // (external row, external col) -> element
    std::vector<int> r_i2e, r_e2i, c_i2e, c_e2i;
    std::vector<long> mr, mc;  // Markowitz counts by internal index

    int lookup(int orr, int occ) const {
        auto it = at.find(key(orr, occ, n));
        return it == at.end() ? -1 : it->second;
    }
    int irow(int e) const { return r_e2i[orow[e]]; }
    int icol(int e) const { return c_e2i[ocol[e]]; }
    int add(int orr, int occ, T v) {
        int e = (int)orow.size();
        orow.push_back(orr);
        ocol.push_back(occ);
        val.push_back(v);
        rows[orr].push_back(e);
        cols[occ].push_back(e);
        at[key(orr, occ, n)] = e;
        return e;
    }
    // Elements of internal column k with internal row >= from, in increasing internal row order.
    void col_from(int k, int from, std::vector<int>& out) const {
        out.clear();
        for (int e : cols[c_i2e[k]])
            if (irow(e) >= from) out.push_back(e);
        std::sort(out.begin(), out.end(), [&](int a, int b) { return irow(a) < irow(b); });
    }
    void row_from(int k, int from, std::vector<int>& out) const {
        out.clear();
        for (int e : rows[r_i2e[k]])
            if (icol(e) >= from) out.push_back(e);
        std::sort(out.begin(), out.end(), [&](int a, int b) { return icol(a) < icol(b); });
    }
    long mark(int e) const {
        long a = mr[irow(e)], b = mc[icol(e)];
        if (a <= 0 || b <= 0) fail(S21_PIVOT, "Pivot Search Fail");
        return (a - 1) * (b - 1);
    }
    int max_of(const std::vector<int>& lst) const {  // first element with the largest |value|
        int best = lst[0];
        double bv = mag(val[best]);
        for (size_t i = 1; i < lst.size(); i++)
            if (mag(val[lst[i]]) > bv) {
                best = lst[i];
                bv = mag(val[lst[i]]);
            }
        return best;
    }
    int search_diagonal(int n0, std::vector<int>& tmp) const {  // sparse21/mod.rs:735-783
        int best = -1;
        long best_mark = -1;  // "usize::MAX"
        double best_ratio = 0.0;
        long ties = 0;
        for (int k = n0; k < n; k++) {
            int d = lookup(r_i2e[k], c_i2e[k]);
            if (d < 0) continue;
            col_from(k, n0, tmp);
            if (tmp.empty()) continue;
            int mx = max_of(tmp);
            double threshold = 1e-3 * mag(val[mx]) + 0.0;
            if (mag(val[d]) < threshold) continue;
            long m = mark(d);
            if (best_mark < 0 || m < best_mark) {
                ties = 0;
                best = d;
                best_mark = m;
                best_ratio = mag(val[d] / val[mx]);
            } else if (m == best_mark) {
                ties += 1;
                double ratio = mag(val[d] / val[mx]);
                if (ratio > best_ratio) {
                    best = d;
                    best_mark = m;
                    best_ratio = ratio;
                }
                if (ties >= best_mark * 5) return best;
            }
        }
        return best;
    }
    int search_submatrix(int n0, std::vector<int>& tmp) const {  // sparse21/mod.rs:785-834 (column n0 only)
        col_from(n0, n0, tmp);
        if (tmp.empty()) return -1;
        int mx = max_of(tmp);
        int best = -1;
        long best_mark = -1;
        double best_ratio = 0.0;
        for (int e : tmp) {
            long m = mark(e);
            if (best_mark < 0 || m < best_mark) {
                best = e;
                best_mark = m;
                best_ratio = mag(val[e] / val[mx]);
            } else if (m == best_mark) {
                double ratio = mag(val[e] / val[mx]);
                if (ratio > best_ratio) {
                    best = e;
                    best_mark = m;
                    best_ratio = ratio;
                }
            }
        }
        return best;
    }
    int find_max(int n0, std::vector<int>& tmp) const {  // sparse21/mod.rs:837-863
        int best = -1;
        double bv = 0.0;
        for (int k = n0; k < n; k++) {
            col_from(k, n0, tmp);
            for (int e : tmp)
                if (mag(val[e]) > bv) {
                    best = e;
                    bv = mag(val[e]);
                }
        }
        return best;
    }
    void swap_rows(int a, int b) {
        if (a == b) return;
        std::swap(r_i2e[a], r_i2e[b]);
        r_e2i[r_i2e[a]] = a;
        r_e2i[r_i2e[b]] = b;
        std::swap(mr[a], mr[b]);
    }
    void swap_cols(int a, int b) {
        if (a == b) return;
        std::swap(c_i2e[a], c_i2e[b]);
        c_e2i[c_i2e[a]] = a;
        c_e2i[c_i2e[b]] = b;
        std::swap(mc[a], mc[b]);
    }
    // `numeric` false: the fill pattern and Markowitz counts only (an order given in advance).
    void eliminate(int pivot, int n0, std::vector<int>& L, std::vector<int>& U, bool numeric = true) {  // sparse21/mod.rs:865-919
        const T pv = val[pivot];
        if (numeric && is_zero(pv)) fail(S21_PIVOT, "Assert Neq Failed: zero pivot");
        col_from(n0, n0 + 1, L);
        row_from(n0, n0 + 1, U);
        if (numeric)
            for (int l : L) val[l] = val[l] / pv;
        for (int u : U) {
            const int uc = ocol[u];
            for (int l : L) {
                int t = lookup(orow[l], uc);
                if (t < 0) {
                    t = add(orow[l], uc, T{});
                    mr[irow(t)] += 1;  // fill-in while factoring bumps both counts (sparse21/mod.rs:350-353)
                    mc[icol(t)] += 1;
                }
                if (numeric) {
                    const T v = val[u] * val[l];
                    val[t] = val[t] - v;
                }
            }
            mc[icol(u)] -= 1;
        }
        mr[n0] -= 1;
        mc[n0] -= 1;
        for (int l : L) mr[irow(l)] -= 1;
    }
};
}  // namespace

// Unpivoted pattern used when the pivot search fails: the residual and convergence test still run (the
// reference may converge before it ever factors), and the failure is reported when a solve is needed.
static void topo_identity_pattern(Topology& t) {
    const int n = t.n;
    t.row_i2e.resize(n);
    std::iota(t.row_i2e.begin(), t.row_i2e.end(), 0);
    t.row_e2i = t.col_i2e = t.col_e2i = t.row_i2e;
    std::vector<std::vector<std::pair<int, int>>> byrow(n);
    for (size_t e = 0; e < t.erow.size(); e++) byrow[t.erow[e]].push_back({t.ecol[e], (int)e});
    t.rp.assign(n + 1, 0);
    t.ci.clear();
    t.diag.assign(n, -1);
    t.elem_pos.assign(t.erow.size(), -1);
    t.is_a.clear();
    for (int i = 0; i < n; i++) {
        std::sort(byrow[i].begin(), byrow[i].end());
        for (auto& pr : byrow[i]) {
            int pos = (int)t.ci.size();
            if (pr.first == i) t.diag[i] = pos;
            t.ci.push_back(pr.first);
            t.is_a.push_back(1);
            t.elem_pos[pr.second] = pos;
        }
        t.rp[i + 1] = (int)t.ci.size();
    }
    t.nnz = (int)t.ci.size();
}

template <class T>
static double pivot_search(Topology& t, const std::vector<T>& first_values, const std::vector<int>* order_r = nullptr,
                           const std::vector<int>* order_c = nullptr,
                           const std::vector<std::vector<T>>* later_values = nullptr);

void topo_pivot(Topology& t, const std::vector<double>& first_values) {
    t.sym_status = S21_OK;
    t.sym_error.clear();
    try {
        pivot_search<double>(t, first_values);
    } catch (Error& e) {
        t.sym_status = e.code;
        t.sym_error = e.what();
        topo_identity_pattern(t);
    }
}

void topo_pivot_cx(Topology& t, const std::vector<double>& re, const std::vector<double>& im) {
    t.sym_status = S21_OK;
    t.sym_error.clear();
    std::vector<HCx> v(re.size());
    for (size_t k = 0; k < v.size(); k++) v[k] = HCx{re[k], im[k]};
    try {
        pivot_search<HCx>(t, v);
    } catch (Error& e) {
        t.sym_status = e.code;
        t.sym_error = e.what();
        topo_identity_pattern(t);
    }
}

void topo_pivot_search(Topology& t, const std::vector<double>& first_values) { pivot_search<double>(t, first_values); }

void topo_pivot_sequence(Topology& t, const std::vector<double>& first_values,
                         const std::vector<std::vector<double>>& later_values) {
    t.sym_status = S21_OK;
    t.sym_error.clear();
    // The sequence stands in for the reference's searches at every factorisation by the first one and the
    // given later ones, so its arrangement can differ from the reference's, which factored this last matrix.
    // From that arrangement the search can meet a zero pivot (a column whose candidates are all zero after
    // the earlier eliminations), or leave a last pivot -- never searched -- that is zero or subnormal, which
    // the kernel's factorisation cannot keep either; the resumed solve would stop at the same iteration
    // again.  The last matrix is then searched afresh instead, as a new solver's first factorisation would.
    try {
        const double last = pivot_search<double>(t, first_values, nullptr, nullptr, &later_values);
        if (!later_values.empty() && std::fabs(last) < 2.2250738585072014e-308)
            fail(S21_PIVOT, "Assert Neq Failed: zero pivot");
        return;
    } catch (Error& e) {
        t.sym_status = e.code;
        t.sym_error = e.what();
    }
    if (!later_values.empty()) {
        try {
            pivot_search<double>(t, later_values.back());
            t.sym_status = S21_OK;
            t.sym_error.clear();
            return;
        } catch (Error& e) {
            t.sym_status = e.code;
            t.sym_error = e.what();
        }
    }
    topo_identity_pattern(t);
}

void topo_pivot_forced(Topology& t, const std::vector<int>& row_i2e, const std::vector<int>& col_i2e) {
    if ((int)row_i2e.size() != t.n || (int)col_i2e.size() != t.n) fail(S21_INVALID, "pivot order of the wrong size");
    t.sym_status = S21_OK;
    t.sym_error.clear();
    std::vector<double> ones(t.erow.size(), 1.0);
    pivot_search<double>(t, ones, &row_i2e, &col_i2e);
}

// With order_r / order_c (internal index -> external row / column), the pivots are those entries instead
// of the search's choice and no values are computed: the fill pattern of a given order.  With
// later_values, each set is searched in turn the way the reference's next factorisation searches: from
// the arrangement and the elements (fill included) the previous one left, element values reset to zero
// and reloaded (Matrix::reset / update / lu_factorize, sparse21/mod.rs:272-282, 647-674), Markowitz
// counts restarting from each row's and column's element count (setup_factoring, :191-196).
// Returns the magnitude of the last pivot after the elimination (1 for a given order, whose values are not
// computed).
template <class T>
static double pivot_search(Topology& t, const std::vector<T>& first_values, const std::vector<int>* order_r,
                           const std::vector<int>* order_c, const std::vector<std::vector<T>>* later_values) {
    const int n = t.n;
    PivotState<T> s;
    s.n = n;
    s.rows.resize(n);
    s.cols.resize(n);
    s.r_i2e.resize(n);
    s.c_i2e.resize(n);
    std::iota(s.r_i2e.begin(), s.r_i2e.end(), 0);
    std::iota(s.c_i2e.begin(), s.c_i2e.end(), 0);
    s.r_e2i = s.r_i2e;
    s.c_e2i = s.c_i2e;
    s.mr.assign(n, 0);
    s.mc.assign(n, 0);
    for (size_t k = 0; k < t.erow.size(); k++) {
        s.add(t.erow[k], t.ecol[k], first_values[k]);
        s.mr[t.erow[k]] += 1;
        s.mc[t.ecol[k]] += 1;
    }
    if (n == 0) fail(S21_ERR, "Assert Gt Failed: 0 <= 0");
    for (int k = 0; k < n; k++)
        if (s.rows[k].empty() || s.cols[k].empty()) fail(S21_SINGULAR, "Singular Matrix");
    std::vector<int> tmp, L, U;
    const bool forced = order_r != nullptr;
    const size_t passes = 1 + (later_values && !forced ? later_values->size() : 0);
    for (size_t pass = 0; pass < passes; pass++) {
        if (pass > 0) {
            const std::vector<T>& v = (*later_values)[pass - 1];
            for (size_t e = 0; e < s.val.size(); e++) s.val[e] = e < v.size() ? v[e] : T{};
            for (int k = 0; k < n; k++) {
                s.mr[k] = (long)s.rows[s.r_i2e[k]].size();
                s.mc[k] = (long)s.cols[s.c_i2e[k]].size();
            }
        }
        for (int n0 = 0; n0 < n - 1; n0++) {
            int pv;
            if (forced) {
                pv = s.lookup((*order_r)[n0], (*order_c)[n0]);
                if (pv < 0) fail(S21_INVALID, "given pivot order names an absent entry");
            } else {
                pv = s.search_diagonal(n0, tmp);
                if (pv < 0) pv = s.search_submatrix(n0, tmp);
                if (pv < 0) pv = s.find_max(n0, tmp);
                if (pv < 0) fail(S21_PIVOT, "Pivot Search Fail");
            }
            s.swap_rows(s.irow(pv), n0);
            s.swap_cols(s.icol(pv), n0);
            s.eliminate(pv, n0, L, U, !forced);
        }
    }
    t.row_i2e = s.r_i2e;
    t.row_e2i = s.r_e2i;
    t.col_i2e = s.c_i2e;
    t.col_e2i = s.c_e2i;
    const int last = s.lookup(s.r_i2e[n - 1], s.c_i2e[n - 1]);
    if (last < 0) fail(S21_SINGULAR, "Singular Matrix");
    const double last_mag = forced ? 1.0 : (double)mag(s.val[last]);

    // Pattern in internal coordinates; the first erow.size() entries are the matrix elements.
    const int ne = (int)s.orow.size();
    std::vector<std::vector<std::pair<int, int>>> byrow(n);  // internal col, element id
    for (int e = 0; e < ne; e++) byrow[s.irow(e)].push_back({s.icol(e), e});
    t.rp.assign(n + 1, 0);
    t.ci.clear();
    t.diag.assign(n, -1);
    t.elem_pos.assign(t.erow.size(), -1);
    t.is_a.clear();
    for (int i = 0; i < n; i++) {
        std::sort(byrow[i].begin(), byrow[i].end());
        for (auto& pr : byrow[i]) {
            int pos = (int)t.ci.size();
            if (pr.first == i) t.diag[i] = pos;
            t.ci.push_back(pr.first);
            bool is_elem = pr.second < (int)t.erow.size();
            t.is_a.push_back(is_elem ? 1 : 0);
            if (is_elem) t.elem_pos[pr.second] = pos;
        }
        t.rp[i + 1] = (int)t.ci.size();
        if (t.diag[i] < 0) fail(S21_SINGULAR, "Singular Matrix");
    }
    t.nnz = (int)t.ci.size();
    return last_mag;
}

void topo_schedule(Topology& t, const std::vector<uint8_t>& m1_caps, int hub_contrib) {
    const int n = t.n;
    // IKJ update lists: for L entry (i,k), subtract l_ik * u_kj from (i,j) for each U entry (k,j).
    // upd_ptr is indexed by LU position; only L positions own updates.
    std::vector<int> posmap(n, -1);
    t.upd_ptr.assign(t.nnz + 1, 0);
    t.upd_src.clear();
    t.upd_dst.clear();
    t.flops_lu = 0;
    for (int i = 0; i < n; i++) {
        for (int q = t.rp[i]; q < t.rp[i + 1]; q++) posmap[t.ci[q]] = q;
        for (int q = t.rp[i]; q < t.rp[i + 1]; q++) {
            t.upd_ptr[q] = (int)t.upd_src.size();
            if (q >= t.diag[i] || t.diag[i] < 0) continue;
            const int k = t.ci[q];
            for (int u = t.diag[k] + 1; u < t.rp[k + 1]; u++) {
                int dst = posmap[t.ci[u]];
                if (dst < 0) {
                    // the unpivoted pattern of a failed search has no fill; it is never factored (the
                    // kernel reports the search's failure when an instance needs a solve)
                    if (t.sym_status != S21_OK) continue;
                    fail(S21_ERR, "internal: fill pattern not closed");
                }
                t.upd_src.push_back(u);
                t.upd_dst.push_back(dst);
            }
            t.flops_lu += 1 + 2 * (t.rp[k + 1] - t.diag[k] - 1);
        }
        for (int q = t.rp[i]; q < t.rp[i + 1]; q++) posmap[t.ci[q]] = -1;
    }
    t.upd_ptr[t.nnz] = (int)t.upd_src.size();
    // Row levels: factorisation / forward sweep depend on L rows, backward on U rows.
    std::vector<int> lf(n, 0), lb(n, 0);
    int maxf = 0, maxb = 0;
    for (int i = 0; i < n; i++) {
        int l = 0;
        for (int q = t.rp[i]; q < t.diag[i]; q++) l = std::max(l, lf[t.ci[q]] + 1);
        lf[i] = l;
        maxf = std::max(maxf, l);
    }
    for (int i = n - 1; i >= 0; i--) {
        int l = 0;
        for (int q = t.diag[i] + 1; q < t.rp[i + 1]; q++) l = std::max(l, lb[t.ci[q]] + 1);
        lb[i] = l;
        maxb = std::max(maxb, l);
    }
    auto bucket = [&](const std::vector<int>& lev, int mx, std::vector<int>& ptr, std::vector<int>& rows) {
        ptr.assign(mx + 2, 0);
        for (int i = 0; i < n; i++) ptr[lev[i] + 1]++;
        for (int l = 0; l <= mx; l++) ptr[l + 1] += ptr[l];
        rows.assign(n, 0);
        std::vector<int> fillp(ptr.begin(), ptr.end() - 1);
        for (int i = 0; i < n; i++) rows[fillp[lev[i]]++] = i;
    };
    bucket(lf, maxf, t.lvl_ptr, t.lvl_rows);
    bucket(lb, maxb, t.blvl_ptr, t.blvl_rows);

    // Chunked schedule over L entries, rows in internal order (pivot rows come first).
    {
        std::vector<int> complete(n, 0);  // first level at which the row's factor is final
        std::vector<std::array<int, 4>> chunks;  // level, row, q0, q1
        int maxl = -1;
        for (int i = 0; i < n; i++) {
            int cur = 0, last = -1;
            for (int q = t.rp[i]; t.diag[i] >= 0 && q < t.diag[i]; q++) {
                const int lvl = std::max(cur, complete[t.ci[q]]);
                if (last >= 0 && chunks[last][0] == lvl)
                    chunks[last][3] = q + 1;
                else {
                    chunks.push_back({lvl, i, q, q + 1});
                    last = (int)chunks.size() - 1;
                }
                cur = lvl;
            }
            complete[i] = last >= 0 ? cur + 1 : 0;
            maxl = std::max(maxl, last >= 0 ? cur : -1);
        }
        const int nl = maxl + 1;
        t.clvl_ptr.assign(nl + 1, 0);
        for (auto& c : chunks) t.clvl_ptr[c[0] + 1]++;
        for (int l = 0; l < nl; l++) t.clvl_ptr[l + 1] += t.clvl_ptr[l];
        std::vector<int> fillp(t.clvl_ptr.begin(), t.clvl_ptr.end() - 1);
        t.ch_row.assign(chunks.size(), 0);
        t.ch_q0.assign(chunks.size(), 0);
        t.ch_q1.assign(chunks.size(), 0);
        for (auto& c : chunks) {
            const int at = fillp[c[0]]++;
            t.ch_row[at] = c[1];
            t.ch_q0[at] = c[2];
            t.ch_q1[at] = c[3];
        }
    }

    // Residual rows over A entries (internal row i = external row row_i2e[i]).
    t.res_ptr.assign(n + 1, 0);
    t.res_pos.clear();
    t.res_xcol.clear();
    for (int i = 0; i < n; i++) {
        for (int q = t.rp[i]; q < t.rp[i + 1]; q++)
            if (t.is_a[q]) {
                t.res_pos.push_back(q);
                t.res_xcol.push_back(t.col_i2e[t.ci[q]]);
            }
        t.res_ptr[i + 1] = (int)t.res_pos.size();
    }

    // Devices: the live (entry, slot) pairs of every component, in entry order.  A slot is an LU position
    // of A (0 .. nnz-1) or an RHS row (nnz + row); grounded entries are dropped.
    std::unordered_map<int64_t, int> elem;
    for (size_t k = 0; k < t.erow.size(); k++) elem[key(t.erow[k], t.ecol[k], t.n)] = (int)k;
    auto target = [&](int r, int c) -> int {
        if (r < 0) return -1;
        if (c == -2) return t.nnz + r;  // RHS
        if (c < 0) return -1;
        return t.elem_pos[elem.at(key(r, c, t.n))];
    };
    const int ncomp = (int)t.comps.size();
    std::vector<std::vector<std::pair<int, int>>> live(ncomp);
    std::vector<int> cls_of(ncomp, CLS_LIN);
    std::vector<uint16_t> flags(ncomp, 0);
    int im1 = 0;
    for (int ci = 0; ci < ncomp; ci++) {
        const Comp& c = t.comps[ci];
        const int* nd = c.node;
        auto add = [&](int e, int r, int col) {
            const int sl = target(r, col);
            if (sl >= 0) live[ci].push_back({e, sl});
        };
        switch (c.type) {
        case S21_COMP_MOS1: {
            const bool caps = m1_caps[im1++] != 0;
            const bool rdrs = nd[RDP] != nd[RD] || nd[RSP] != nd[RS];
            const bool tied = !rdrs && nd[RS] == nd[RB];
            for (int e = 0; e < MOS1_NE; e++) {
                if (mos1_entry_tr_only(e) && !(caps && (t.tran || t.ac))) continue;  // exactly zero
                if (mos1_entry_r_only(e) && !rdrs) continue;                        // grd = grs = 0
                if (tied && mos1t_merged(e)) continue;                              // merged in registers
                if (t.ac && mos1_entry_col(e) < 0) continue;                        // no AC RHS
                const int r = nd[mos1_entry_row(e)];
                const int col = mos1_entry_col(e) < 0 ? -2 : nd[mos1_entry_col(e)];
                if (mos1_entry_col(e) >= 0 && col < 0) continue;
                add(e, r, col);
            }
            cls_of[ci] = tied ? CLS_MOS1T : CLS_MOS1;
            flags[ci] = caps && t.tran ? DEV_CAPS : 0;
            break;
        }
        case S21_COMP_MOS0:
            if (t.ac) fail(S21_ERR, "AC Not Implemented For This Component!");  // comps/mod.rs:86-88
            for (int e = 0; e < MOS0_NE; e++) {
                const int r = nd[mos0_entry_row(e)];
                const int col = mos0_entry_col(e) < 0 ? -2 : nd[mos0_entry_col(e)];
                if (mos0_entry_col(e) >= 0 && col < 0) continue;
                add(e, r, col);
            }
            cls_of[ci] = CLS_MOS0;
            flags[ci] = c.model == 1 ? DEV_PMOS : 0;
            break;
        case S21_COMP_DIODE:
            if (t.ac) fail(S21_ERR, "AC Not Implemented For This Component!");
            for (int e = 0; e < DIODE_NE; e++) {
                const int r = nd[diode_entry_row(e)];
                const int col = diode_entry_col(e) < 0 ? -2 : nd[diode_entry_col(e)];
                if (diode_entry_col(e) >= 0 && col < 0) continue;
                add(e, r, col);
            }
            cls_of[ci] = CLS_DIODE;
            break;
        case S21_COMP_BSIM4:  // no load_ac in the reference's BSIM4: Component's default (comps/mod.rs:86-88)
            if (t.ac) fail(S21_ERR, "AC Not Implemented For This Component!");
            for (int e = 0; e < b4::B4_NS; e++) {
                const int r = nd[b4::b4_row(e)];
                const int col = b4::b4_col(e) < 0 ? -2 : nd[b4::b4_col(e)];
                if (b4::b4_col(e) >= 0 && col < 0) continue;
                add(e, r, col);
            }
            cls_of[ci] = CLS_BSIM4;
            break;
        default: {
            const int p = nd[0], q = nd[1], iv = nd[2];
            if (c.type == S21_COMP_I && t.ac) fail(S21_ERR, "AC Not Implemented For This Component!");
            if (c.type == S21_COMP_R || (c.type == S21_COMP_C && (t.tran || t.ac))) {
                add(0, p, p < 0 ? -1 : p);
                add(1, q, q < 0 ? -1 : q);
                add(2, p, q < 0 ? -1 : q);
                add(3, q, p < 0 ? -1 : p);
                if (c.type == S21_COMP_C && !t.ac) {
                    add(4, p, -2);
                    add(5, q, -2);
                }
            } else if (c.type == S21_COMP_V) {
                add(0, p, iv);
                add(1, iv, p < 0 ? -1 : p);
                add(2, q, iv);
                add(3, iv, q < 0 ? -1 : q);
                add(4, iv, -2);
            } else if (c.type == S21_COMP_I) {
                add(0, p, -2);
                add(1, q, -2);
            }
            cls_of[ci] = CLS_LIN;
            flags[ci] = (uint16_t)(c.type << 4);
        }
        }
    }
    const int nslot = t.nnz + n;
    if (nslot >= 0x8000) fail(S21_INVALID, "circuit too large for the batched kernel (slot index)");
    // Hub slots: more than hub_contrib distinct contributing components ...
    std::vector<int> ncontrib(nslot, 0), lastc(nslot, -1);
    for (int ci = 0; ci < ncomp; ci++)
        for (auto& es : live[ci])
            if (lastc[es.second] != ci) {
                lastc[es.second] = ci;
                ncontrib[es.second]++;
            }
    std::vector<uint8_t> hub(nslot, 0);
    for (int sl = 0; sl < nslot; sl++) hub[sl] = ncontrib[sl] > hub_contrib;
    // ... and every slot one component stamps twice (the kernel applies a device's stamps as a batch)
    std::fill(lastc.begin(), lastc.end(), -1);
    for (int ci = 0; ci < ncomp; ci++)
        for (auto& es : live[ci]) {
            if (lastc[es.second] == ci) hub[es.second] = 1;
            lastc[es.second] = ci;
        }
    // Greedy colouring, class by class, components in order: the first colour of the class in which no
    // earlier member shares a non-hub slot with the component.
    std::vector<std::vector<int>> slot_cols(nslot);
    std::vector<std::vector<int>> cols;  // members per colour (component order)
    std::vector<int> colcls;
    std::vector<uint8_t> used;
    for (int k = 0; k < N_CLS; k++) {
        const size_t first = cols.size();
        for (int ci = 0; ci < ncomp; ci++) {
            if (cls_of[ci] != k) continue;
            used.assign(cols.size() + 1, 0);
            for (auto& es : live[ci])
                if (!hub[es.second])
                    for (int cc : slot_cols[es.second]) used[cc] = 1;
            size_t cc = first;
            while (cc < cols.size() && used[cc]) cc++;
            if (cc == cols.size()) {
                cols.emplace_back();
                colcls.push_back(k);
            }
            cols[cc].push_back(ci);
            for (auto& es : live[ci])
                if (!hub[es.second]) {
                    std::vector<int>& v = slot_cols[es.second];
                    if (std::find(v.begin(), v.end(), (int)cc) == v.end()) v.push_back((int)cc);
                }
        }
    }
    // Device order, records and codes.
    t.dev_comp.clear();
    t.col_ptr.assign(1, 0);
    t.col_cls.clear();
    for (int k = 0; k < N_CLS; k++) t.cls_start[k] = -1;
    for (size_t cc = 0; cc < cols.size(); cc++) {
        if (t.cls_start[colcls[cc]] < 0) t.cls_start[colcls[cc]] = (int)t.dev_comp.size();
        for (int ci : cols[cc]) t.dev_comp.push_back(ci);
        t.col_ptr.push_back((int)t.dev_comp.size());
        t.col_cls.push_back(colcls[cc]);
    }
    const int ndev = (int)t.dev_comp.size();
    t.cls_start[N_CLS] = ndev;
    for (int k = N_CLS - 1; k >= 0; k--)
        if (t.cls_start[k] < 0) t.cls_start[k] = t.cls_start[k + 1];
    std::vector<int> pos_of(ncomp, -1);
    for (int p = 0; p < ndev; p++) pos_of[t.dev_comp[p]] = p;
    int nc = 0;
    for (int k = 0; k < N_CLS; k++) {
        t.code_base[k] = nc;
        nc += (t.cls_start[k + 1] - t.cls_start[k]) * code_stride(k);
    }
    if (nc > 65535 || ndev > 65535) fail(S21_INVALID, "circuit too large for the batched kernel (devices)");
    t.codes.assign(nc, 0xFFFF);
    t.dev_rec.assign((size_t)8 * ndev, 0);
    for (int p = 0; p < ndev; p++) {
        const int ci = t.dev_comp[p];
        for (int j = 0; j < 6; j++) t.dev_rec[8 * p + j] = (uint16_t)(t.comps[ci].node[j] + 1);
        t.dev_rec[8 * p + 6] = flags[ci];
    }
    auto code_at = [&](int ci, int e) -> uint16_t& {
        const int k = cls_of[ci], p = pos_of[ci];
        return t.codes[t.code_base[k] + (p - t.cls_start[k]) * code_stride(k) + e];
    };
    // Hub contributions numbered in component-then-entry order: each hub slot's list is then the order
    // Solver::update adds its stamps.
    std::vector<std::vector<int>> hub_list(nslot);
    t.n_hub = 0;
    for (int ci = 0; ci < ncomp; ci++)
        for (auto& es : live[ci]) {
            if (hub[es.second]) {
                if (t.n_hub >= 0x7FFF) fail(S21_INVALID, "circuit too large for the batched kernel (hub)");
                code_at(ci, es.first) = (uint16_t)(0x8000 | t.n_hub);
                hub_list[es.second].push_back(t.n_hub++);
            } else {
                code_at(ci, es.first) = (uint16_t)es.second;
            }
        }
    t.hub_slot.clear();
    t.g_ptr.assign(1, 0);
    t.g_idx.clear();
    for (int sl = 0; sl < nslot; sl++) {
        if (!hub[sl]) continue;
        t.hub_slot.push_back(sl);
        for (int h : hub_list[sl]) t.g_idx.push_back(h);
        t.g_ptr.push_back((int)t.g_idx.size());
    }
}

}  // namespace s21

namespace s21 {

// Packs the schedule into the records the kernel walks (see Topology).  Long gather lists and residual
// rows are split into tasks of at most CHUNK elements whose partial sums a second pass adds in order.
void topo_records(Topology& t) {
    constexpr int CHUNK = 8;
    auto u = [](int v) -> uint16_t {
        if (v < 0 || v > 65535) fail(S21_INVALID, "circuit too large for the batched kernel (16-bit index)");
        return (uint16_t)v;
    };
    const int n = t.n;
    int part = 0, part_r = 0;
    t.gt.clear();
    t.gc.clear();
    for (size_t h = 0; h < t.hub_slot.size(); h++) {
        const int s = t.hub_slot[h];
        const int k0 = t.g_ptr[h], k1 = t.g_ptr[h + 1];
        if (k1 - k0 <= CHUNK) {
            for (int v : {s, k0, k1, 65535}) t.gt.push_back(u(v));
            continue;
        }
        const int p0 = part;
        for (int k = k0; k < k1; k += CHUNK)
            for (int v : {s, k, std::min(k + CHUNK, k1), part++}) t.gt.push_back(u(v));
        for (int v : {s, p0, part, 0}) t.gc.push_back(u(v));
    }
    t.rt.clear();
    t.re.clear();
    t.rc.clear();
    for (int i = 0; i < n; i++) {
        const int e0 = (int)t.re.size() / 2;
        for (int k = t.res_ptr[i]; k < t.res_ptr[i + 1]; k++) {
            t.re.push_back(u(t.res_pos[k]));
            t.re.push_back(u(t.res_xcol[k]));
        }
        const int e1 = (int)t.re.size() / 2;
        if (e1 - e0 <= CHUNK) {
            for (int v : {i, e0, e1, 65535}) t.rt.push_back(u(v));
            continue;
        }
        const int p0 = part_r;
        for (int e = e0; e < e1; e += CHUNK)
            for (int v : {i, e, std::min(e + CHUNK, e1), part_r++}) t.rt.push_back(u(v));
        for (int v : {i, p0, part_r, 0}) t.rc.push_back(u(v));
    }
    t.n_part = std::max(part, part_r);
    // factorisation chunks: L entries of each chunk as {q, pivot, k, n_updates}, updates as {src, dst}
    t.fch.clear();
    t.fl.clear();
    t.fu.clear();
    for (size_t c = 0; c < t.ch_row.size(); c++) {
        const int l0 = (int)t.fl.size() / 4, u0 = (int)t.fu.size() / 2;
        for (int q = t.ch_q0[c]; q < t.ch_q1[c]; q++) {
            const int k = t.ci[q];
            const int nu = t.upd_ptr[q + 1] - t.upd_ptr[q];
            for (int v : {q, t.diag[k], k, nu}) t.fl.push_back(u(v));
            for (int x = t.upd_ptr[q]; x < t.upd_ptr[q + 1]; x++) {
                t.fu.push_back(u(t.upd_src[x]));
                t.fu.push_back(u(t.upd_dst[x]));
            }
        }
        const int l1 = (int)t.fl.size() / 4;
        for (int v : {t.ch_row[c], l0, l1, u0}) t.fch.push_back(u(v));
    }
    // backward rows in level order with their U entries {position, column}
    t.bro.clear();
    t.bu.clear();
    for (int r = 0; r < (int)t.blvl_rows.size(); r++) {
        const int i = t.blvl_rows[r];
        const int e0 = (int)t.bu.size() / 2;
        for (int q = t.diag[i] + 1; q < t.rp[i + 1]; q++) {
            t.bu.push_back(u(q));
            t.bu.push_back(u(t.ci[q]));
        }
        const int e1 = (int)t.bu.size() / 2;
        for (int v : {i, std::max(t.diag[i], 0), e0, e1}) t.bro.push_back(u(v));
    }
}

}  // namespace s21
