/*
 * ORACLE -- test infrastructure only.
 *
 * CPU FP64 restatement of Spice21's Newton inner loop, written to reproduce the reference's
 * arithmetic operation by operation, quirks included.  It is the checker for the MI355X engine
 * and the CPU baseline of bench.py.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; the product library never links or calls it.
 *
 * Restated from (paths relative to the reference repository):
 *   spice21/src/sparse21/mod.rs  -- orthogonal linked-list matrix, Markowitz pivoting (re-run on
 *                                   every factorisation), right-looking LU, solves, residual
 *   spice21/src/analysis.rs      -- Solver::update/solve/converged, TranState, Tran, ac
 *   spice21/src/comps/mod.rs     -- Vsrc, Capacitor, Resistor, Isrc, constants
 *   spice21/src/comps/mos.rs     -- Mos1 resolve/derive/op_stamp/load_ac, Mos0
 *   spice21/src/comps/diode.rs   -- Diode derive/limit/load
 *   spice21/src/comps/bsim4/     -- BSIM4: the Component (create_matps, load, commit,
 *                                   bsim4solver.rs:28-145, 3757-3769) is the oracle's own; its device
 *                                   functions (parameter derivation, Bsim4::op, tran_op, stamp) are the
 *                                   product's host-and-device source (spice21_amd/csrc/bsim4*.h,
 *                                   bsim4_derive.cpp), compiled here by g++ for the CPU.  They are not an
 *                                   independent restatement: the reference's three BSIM4 ring-oscillator
 *                                   waveforms and its dcop known answers pin them, through this oracle.
 * Complex arithmetic follows num-complex 0.3 (spice21/Cargo.toml:23), which the reference uses
 * for AC: naive division by |z|^2, norm = hypot(re, im).
 *
 * Input is the flattened netlist of include/spice21_mi355x.h (elaboration is the product's).
 */
#include "../include/spice21_mi355x.h"
#include "../spice21_amd/csrc/bsim4_eval.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace orc {

struct SpError {
    std::string desc;
};
[[noreturn]] static void sperror(const std::string& s) { throw SpError{s}; }

/* ------------------------------------------------------------------------------------------
 * Numbers (spnum.rs:16-31; num-complex 0.3 operators)
 * ------------------------------------------------------------------------------------------ */
struct Cx {
    double re, im;
};
static inline Cx operator+(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
static inline Cx operator-(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
static inline Cx operator*(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
static inline Cx operator/(Cx a, Cx b) {
    double norm_sqr = b.re * b.re + b.im * b.im;
    double re = a.re * b.re + a.im * b.im;
    double im = a.im * b.re - a.re * b.im;
    return {re / norm_sqr, im / norm_sqr};
}
static inline Cx cx_scale_mul(Cx a, double s) { return {a.re * s, a.im * s}; }
static inline Cx cx_scale_div(Cx a, double s) { return {a.re / s, a.im / s}; }
static inline double absv(double x) { return std::fabs(x); }
static inline double absv(Cx z) { return std::hypot(z.re, z.im); }
// Diagnostics (not part of the restated algorithm): factorisations, and those whose pivot search chose
// other elements than the same matrix's previous factorisation did.
static std::atomic<long long> g_factorisations{0}, g_pivot_changes{0};
static inline bool is_zero(double x) { return x == 0.0; }
static inline bool is_zero(Cx z) { return z.re == 0.0 && z.im == 0.0; }
template <class T> static inline T zero();
template <> inline double zero<double>() { return 0.0; }
template <> inline Cx zero<Cx>() { return {0.0, 0.0}; }

/* ------------------------------------------------------------------------------------------
 * Sparse matrix: sparse21/mod.rs:12-1012
 * ------------------------------------------------------------------------------------------ */
enum Axis { ROWS = 0, COLS = 1 };
static inline Axis other(Axis a) { return a == ROWS ? COLS : ROWS; }
using Eidx = long;  // Option<Eindex>: -1 is None
static const Eidx NONE = -1;

struct Maker {
    virtual Eidx make(size_t row, size_t col) = 0;
    virtual ~Maker() {}
};

template <class T> struct Matrix : Maker {
    enum State { CREATED, FACTORING, FACTORED, RESET };
    struct Elem {
        Eidx index;
        size_t row, col;
        T val;
        bool fillin;
        Eidx next_in_row = NONE, next_in_col = NONE;
        size_t loc(Axis a) const { return a == ROWS ? row : col; }
        void set_loc(Axis a, size_t to) { (a == ROWS ? row : col) = to; }
        Eidx next(Axis a) const { return a == ROWS ? next_in_row : next_in_col; }
        void set_next(Axis a, Eidx e) { (a == ROWS ? next_in_row : next_in_col) = e; }
    };
    struct AxisData {
        std::vector<Eidx> hdrs;
        std::vector<size_t> qtys, markowitz;
        bool has_mapping = false;
        std::vector<size_t> e2i, i2e;
        void grow(size_t to) {  // :180-190
            while (hdrs.size() < to) {
                hdrs.push_back(NONE);
                qtys.push_back(0);
                markowitz.push_back(0);
            }
        }
        void setup_factoring() {  // :191-196
            markowitz = qtys;
            if (!has_mapping) {
                has_mapping = true;
                e2i.resize(hdrs.size());
                i2e.resize(hdrs.size());
                for (size_t k = 0; k < hdrs.size(); k++) e2i[k] = i2e[k] = k;
            }
        }
        void swap_int(size_t x, size_t y) {  // AxisMapping::swap_int :153-161
            size_t tmp = i2e[x];
            i2e[x] = i2e[y];
            i2e[y] = tmp;
            e2i[i2e[x]] = x;
            e2i[i2e[y]] = y;
        }
        void swap(size_t x, size_t y) {  // :197-204
            std::swap(hdrs[x], hdrs[y]);
            std::swap(qtys[x], qtys[y]);
            std::swap(markowitz[x], markowitz[y]);
            if (has_mapping) swap_int(x, y);
        }
    };

    State state = CREATED;
    std::vector<Elem> el;
    AxisData axes[2];
    std::vector<Eidx> diag;
    std::vector<Eidx> fillins;
    long factorizations = 0;

    size_t num_rows() const { return axes[ROWS].hdrs.size(); }
    size_t num_cols() const { return axes[COLS].hdrs.size(); }
    Eidx hdr(Axis a, size_t loc) const { return axes[a].hdrs[loc]; }
    void set_hdr(Axis a, size_t loc, Eidx e) { axes[a].hdrs[loc] = e; }

    Eidx make(size_t row, size_t col) override {  // :265-270
        Eidx e = get_elem(row, col);
        if (e != NONE) return e;
        return add_element_(row, col, zero<T>(), false);
    }
    void add_element(size_t row, size_t col, T val) { add_element_(row, col, val, false); }
    void reset() {  // :272-277
        for (auto& e : el) e.val = zero<T>();
        state = RESET;
    }
    void update(Eidx ei, T val) {  // :279-282
        T tmp = el[ei].val + val;
        el[ei].val = tmp;
    }
    std::vector<T> vecmul(const std::vector<T>& x) const {  // :284-297
        if (x.size() != num_cols()) sperror("Invalid Dimensions");
        std::vector<T> y(num_rows(), zero<T>());
        for (size_t row = 0; row < num_rows(); row++) {
            Eidx ep = hdr(ROWS, row);
            while (ep != NONE) {
                y[row] = y[row] + el[ep].val * x[el[ep].col];
                ep = el[ep].next_in_row;
            }
        }
        return y;
    }
    std::vector<T> res(const std::vector<T>& x, const std::vector<T>& rhs) const {  // :298-327
        if (x.size() != num_cols() || rhs.size() != num_cols() || num_rows() != num_cols())
            sperror("index out of bounds");
        std::vector<T> xi(num_cols(), zero<T>());
        if (axes[COLS].has_mapping) {
            for (size_t k = 0; k < xi.size(); k++) xi[k] = x[axes[COLS].i2e[k]];
        } else {
            for (size_t k = 0; k < xi.size(); k++) xi[k] = x[k];
        }
        std::vector<T> ri = vecmul(xi);
        std::vector<T> r(ri.size(), zero<T>());
        if (axes[ROWS].has_mapping) {
            for (size_t k = 0; k < xi.size(); k++) r[k] = rhs[k] - ri[axes[ROWS].e2i[k]];
        } else {
            for (size_t k = 0; k < xi.size(); k++) r[k] = rhs[k] - ri[k];
        }
        return r;
    }
    void insert(Elem& e) {  // :328-361
        bool expanded = false;
        if (e.row + 1 > num_rows()) {
            axes[ROWS].grow(e.row + 1);
            expanded = true;
        }
        if (e.col + 1 > num_cols()) {
            axes[COLS].grow(e.col + 1);
            expanded = true;
        }
        if (expanded) {
            size_t new_diag_len = std::min(num_rows(), num_cols());
            while (diag.size() < new_diag_len) diag.push_back(NONE);
        }
        insert_axis(COLS, e);
        insert_axis(ROWS, e);
        axes[ROWS].qtys[e.row] += 1;
        axes[COLS].qtys[e.col] += 1;
        if (state == FACTORING) {
            axes[ROWS].markowitz[e.row] += 1;
            axes[COLS].markowitz[e.col] += 1;
        }
        if (e.row == e.col) diag[e.row] = e.index;
        if (e.fillin) fillins.push_back(e.index);
    }
    void insert_axis(Axis ax, Elem& e) {  // :362-391
        Eidx head = axes[ax].hdrs[e.loc(ax)];
        if (head == NONE) {
            set_hdr(ax, e.loc(ax), e.index);
            return;
        }
        Axis off = other(ax);
        if (el[head].loc(off) > e.loc(off)) {
            e.set_next(ax, head);
            set_hdr(ax, e.loc(ax), e.index);
            return;
        }
        Eidx prev = head;
        while (el[prev].next(ax) != NONE) {
            Eidx nxt = el[prev].next(ax);
            if (el[nxt].loc(off) >= e.loc(off)) break;
            prev = nxt;
        }
        e.set_next(ax, el[prev].next(ax));
        el[prev].set_next(ax, e.index);
    }
    Eidx add_fillin(size_t row, size_t col) { return add_element_(row, col, zero<T>(), true); }  // :392
    Eidx add_element_(size_t row, size_t col, T val, bool fillin) {  // :395-402
        Elem e;
        e.index = (Eidx)el.size();
        e.row = row;
        e.col = col;
        e.val = val;
        e.fillin = fillin;
        insert(e);
        el.push_back(e);
        return e.index;
    }
    Eidx get_elem(size_t row, size_t col) const {  // :404-428
        if (row >= num_rows()) return NONE;
        if (col >= num_cols()) return NONE;
        if (row == col) return diag[row];
        Eidx ep = hdr(ROWS, row);
        while (ep != NONE) {
            const Elem& e = el[ep];
            if (e.col == col) return ep;
            if (e.col > col) return NONE;
            ep = e.next_in_row;
        }
        return NONE;
    }
    void move_element(Axis ax, Eidx idx, size_t to) {  // :436-499
        size_t loc = el[idx].loc(ax);
        if (loc == to) return;
        Axis off = other(ax);
        size_t y = el[idx].loc(off);
        if (loc < to) {
            Eidx br = before_loc(off, y, to, idx);
            if (br == NONE) sperror("ERROR");
            if (br != idx) {
                Eidx be = prev(off, idx, NONE);
                Eidx nxt = el[idx].next(off);
                if (be == NONE)
                    set_hdr(off, y, nxt);
                else
                    el[be].set_next(off, nxt);
                Eidx brn = el[br].next(off);
                el[idx].set_next(off, brn);
                el[br].set_next(off, idx);
            }
        } else {
            Eidx br = before_loc(off, y, to, NONE);
            Eidx be = prev(off, idx, NONE);
            if (br != be) {
                if (be != NONE) {
                    Eidx nxt = el[idx].next(off);
                    el[be].set_next(off, nxt);
                }
                if (br == NONE) {
                    Eidx first = hdr(off, y);
                    el[idx].set_next(off, first);
                    axes[off].hdrs[y] = idx;
                } else if (br != idx) {
                    Eidx nxt = el[br].next(off);
                    el[idx].set_next(off, nxt);
                    el[br].set_next(off, idx);
                }
            }
        }
        el[idx].set_loc(ax, to);
        if (loc == y)
            diag[loc] = NONE;
        else if (to == y)
            diag[to] = idx;
    }
    void exchange_elements(Axis ax, Eidx ix, Eidx iy) {  // :500-551
        Axis off = other(ax);
        size_t off_loc = el[ix].loc(off);
        Eidx bx = prev(off, ix, NONE);
        Eidx by = prev(off, iy, ix);
        if (by == NONE) sperror("ERROR!");
        size_t locx = el[ix].loc(ax);
        size_t locy = el[iy].loc(ax);
        el[iy].set_loc(ax, locx);
        el[ix].set_loc(ax, locy);
        if (bx == NONE)
            set_hdr(off, off_loc, iy);
        else
            el[bx].set_next(off, iy);
        if (by == ix) {
            Eidx tmp = el[iy].next(off);
            el[iy].set_next(off, ix);
            el[ix].set_next(off, tmp);
        } else {
            Eidx xnxt = el[ix].next(off);
            Eidx ynxt = el[iy].next(off);
            el[iy].set_next(off, xnxt);
            el[ix].set_next(off, ynxt);
            el[by].set_next(off, ix);
        }
        if (locx == off_loc)
            diag[off_loc] = iy;
        else if (locy == off_loc)
            diag[off_loc] = ix;
    }
    Eidx prev(Axis ax, Eidx idx, Eidx hint) const {  // :552-575
        Eidx p = (hint != NONE) ? hint : hdr(ax, el[idx].loc(ax));
        if (p == NONE) return NONE;
        if (p == idx) return NONE;
        Eidx pi = p;
        while (el[pi].next(ax) != NONE) {
            Eidx nxt = el[pi].next(ax);
            if (nxt == idx) break;
            pi = nxt;
        }
        return pi;
    }
    Eidx before_loc(Axis ax, size_t loc, size_t before, Eidx hint) const {  // :576-598
        Eidx p = (hint != NONE) ? hint : hdr(ax, loc);
        Axis off = other(ax);
        if (p == NONE) return NONE;
        if (el[p].loc(off) >= before) return NONE;
        Eidx pi = p;
        while (el[pi].next(ax) != NONE) {
            Eidx nxt = el[pi].next(ax);
            if (el[nxt].loc(off) >= before) break;
            pi = nxt;
        }
        return pi;
    }
    void swap(Axis ax, size_t a, size_t b) {  // :599-643
        if (a == b) return;
        size_t x = std::min(a, b), y = std::max(a, b);
        Eidx ix = axes[ax].hdrs[x];
        Eidx iy = axes[ax].hdrs[y];
        Axis off = other(ax);
        for (;;) {
            if (ix != NONE && iy != NONE) {
                size_t ox = el[ix].loc(off), oy = el[iy].loc(off);
                if (ox < oy) {
                    move_element(ax, ix, y);
                    ix = el[ix].next(ax);
                } else if (oy < ox) {
                    move_element(ax, iy, x);
                    iy = el[iy].next(ax);
                } else {
                    exchange_elements(ax, ix, iy);
                    ix = el[ix].next(ax);
                    iy = el[iy].next(ax);
                }
            } else if (ix == NONE && iy != NONE) {
                move_element(ax, iy, x);
                iy = el[iy].next(ax);
            } else if (ix != NONE && iy == NONE) {
                move_element(ax, ix, y);
                ix = el[ix].next(ax);
            } else {
                break;
            }
        }
        axes[ax].swap(x, y);
    }
    void lu_factorize() {  // :647-674
        if (!(diag.size() > 0)) sperror("Assert Gt Failed: 0 <= 0");
        for (size_t k = 0; k < axes[ROWS].hdrs.size(); k++)
            if (hdr(ROWS, k) == NONE) sperror("Singular Matrix");
        for (size_t k = 0; k < axes[COLS].hdrs.size(); k++)
            if (hdr(COLS, k) == NONE) sperror("Singular Matrix");
        state = FACTORING;
        axes[ROWS].setup_factoring();
        axes[COLS].setup_factoring();
        std::vector<Eidx> seq;
        for (size_t n = 0; n < diag.size() - 1; n++) {
            Eidx pivot = search_for_pivot(n);
            if (pivot == NONE) sperror("Pivot Search Fail");
            seq.push_back(pivot);
            swap(ROWS, el[pivot].row, n);
            swap(COLS, el[pivot].col, n);
            row_col_elim(pivot, n);
        }
        state = FACTORED;
        factorizations++;
        // diagnostics: did this factorisation's search pick other elements than the previous one's?
        if (factorizations > 1 && seq != last_pivots) g_pivot_changes++;
        g_factorisations++;
        last_pivots.swap(seq);
    }
    std::vector<Eidx> last_pivots;
    Eidx search_for_pivot(size_t n) const {  // :676-686
        Eidx ei = markowitz_search_diagonal(n);
        if (ei != NONE) return ei;
        ei = markowitz_search_submatrix(n);
        if (ei != NONE) return ei;
        return find_max(n);
    }
    Eidx max_after(Axis ax, Eidx after) const {  // :688-702
        Eidx best = after;
        double best_val = absv(el[after].val);
        Eidx e = el[after].next(ax);
        while (e != NONE) {
            double val = absv(el[e].val);
            if (val > best_val) {
                best = e;
                best_val = val;
            }
            e = el[e].next(ax);
        }
        return best;
    }
    Eidx max_after_loc(Axis ax, size_t in_loc, size_t after_loc) const {  // :704-724
        Eidx e = axes[ax].hdrs[in_loc];
        Axis off = other(ax);
        while (e != NONE) {
            if (el[e].loc(off) >= after_loc) break;
            e = el[e].next(ax);
        }
        if (e == NONE) return NONE;
        Eidx best = e;
        double best_val = absv(el[best].val);
        while (e != NONE) {
            double val = absv(el[e].val);
            if (val > best_val) {
                best = e;
                best_val = val;
            }
            e = el[e].next(ax);
        }
        return best;
    }
    size_t markowitz_product(Eidx ei) const {  // :726-733
        const Elem& e = el[ei];
        size_t mr = axes[ROWS].markowitz[e.row];
        size_t mc = axes[COLS].markowitz[e.col];
        if (!(mr > 0) || !(mc > 0)) sperror("markowitz count underflow");
        return (mr - 1) * (mc - 1);
    }
    Eidx markowitz_search_diagonal(size_t n) const {  // :735-783  (rel 1e-3, abs 0, ties x5)
        Eidx best_elem = NONE;
        size_t best_mark = (size_t)-1;
        double best_ratio = 0.0;
        size_t num_ties = 0;
        for (size_t k = n; k < diag.size(); k++) {
            Eidx d = diag[k];
            if (d == NONE) continue;
            Eidx max_in_col = max_after_loc(COLS, k, n);
            if (max_in_col == NONE) continue;
            double threshold = 1e-3 * absv(el[max_in_col].val) + 0.0;
            if (absv(el[d].val) < threshold) continue;
            size_t mark = markowitz_product(d);
            if (mark < best_mark) {
                num_ties = 0;
                best_elem = diag[k];
                best_mark = mark;
                best_ratio = absv(el[d].val / el[max_in_col].val);
            } else if (mark == best_mark) {
                num_ties += 1;
                double ratio = absv(el[d].val / el[max_in_col].val);
                if (ratio > best_ratio) {
                    best_elem = diag[k];
                    best_mark = mark;
                    best_ratio = ratio;
                }
                if (num_ties >= best_mark * 5) return best_elem;
            }
        }
        return best_elem;
    }
    Eidx markowitz_search_submatrix(size_t n) const {  // :785-834 (scans column n on every pass)
        Eidx best_elem = NONE;
        size_t best_mark = (size_t)-1;
        double best_ratio = 0.0;
        for (size_t k_ = n; k_ < axes[COLS].hdrs.size(); k_++) {
            Eidx e = hdr(COLS, n);
            while (e != NONE) {
                if (el[e].row >= n) break;
                e = el[e].next_in_col;
            }
            if (e == NONE) continue;
            Eidx max_in_col = max_after(COLS, e);
            while (e != NONE) {
                size_t mark = markowitz_product(e);
                if (mark < best_mark) {
                    best_elem = e;
                    best_mark = mark;
                    best_ratio = absv(el[e].val / el[max_in_col].val);
                } else if (mark == best_mark) {
                    double ratio = absv(el[e].val / el[max_in_col].val);
                    if (ratio > best_ratio) {
                        best_elem = e;
                        best_mark = mark;
                        best_ratio = ratio;
                    }
                }
                e = el[e].next_in_col;
            }
        }
        return best_elem;
    }
    Eidx find_max(size_t n) const {  // :837-863
        Eidx max_elem = NONE;
        double max_val = 0.0;
        for (size_t k = n; k < axes[COLS].hdrs.size(); k++) {
            Eidx ep = hdr(COLS, k);
            while (ep != NONE) {
                if (el[ep].row >= n) break;
                ep = el[ep].next_in_col;
            }
            while (ep != NONE) {
                double val = absv(el[ep].val);
                if (val > max_val) {
                    max_elem = ep;
                    max_val = val;
                }
                ep = el[ep].next_in_col;
            }
        }
        return max_elem;
    }
    void row_col_elim(Eidx pivot, size_t n) {  // :865-919
        Eidx de = diag[n];
        if (de == NONE) sperror("Singular Matrix");
        if (de != pivot) sperror("Assert Eq Failed: pivot is not on the diagonal");
        T pivot_val = el[pivot].val;
        if (is_zero(pivot_val)) sperror("Assert Neq Failed: zero pivot");
        Eidx plower = el[pivot].next_in_col;
        while (plower != NONE) {
            el[plower].val = el[plower].val / pivot_val;
            plower = el[plower].next_in_col;
        }
        Eidx pupper = el[pivot].next_in_row;
        while (pupper != NONE) {
            size_t pupper_col = el[pupper].col;
            plower = el[pivot].next_in_col;
            Eidx psub = el[pupper].next_in_col;
            while (plower != NONE) {
                while (psub != NONE) {
                    if (el[psub].row >= el[plower].row) break;
                    psub = el[psub].next_in_col;
                }
                Eidx pse;
                if (psub == NONE)
                    pse = add_fillin(el[plower].row, pupper_col);
                else if (el[psub].row > el[plower].row)
                    pse = add_fillin(el[plower].row, pupper_col);
                else
                    pse = psub;
                T v = el[pupper].val * el[plower].val;
                el[pse].val = el[pse].val - v;
                psub = el[pse].next_in_col;
                plower = el[plower].next_in_col;
            }
            axes[COLS].markowitz[pupper_col] -= 1;
            pupper = el[pupper].next_in_row;
        }
        axes[ROWS].markowitz[n] -= 1;
        axes[COLS].markowitz[n] -= 1;
        plower = el[pivot].next_in_col;
        while (plower != NONE) {
            size_t plower_row = el[plower].row;
            axes[ROWS].markowitz[plower_row] -= 1;
            plower = el[plower].next_in_col;
        }
    }
    std::vector<T> solve(const std::vector<T>& rhs) {  // :929-991
        if (state != FACTORED) lu_factorize();
        // The reference indexes past the end (a Rust panic) when the vector and the matrix disagree.
        if (rhs.size() != num_rows() || rhs.size() != num_cols()) sperror("index out of bounds");
        std::vector<T> c(rhs.size(), zero<T>());
        if (!axes[ROWS].has_mapping) sperror("Missing Row Mapping");
        for (size_t k = 0; k < c.size(); k++) c[k] = rhs[axes[ROWS].i2e[k]];
        for (size_t k = 0; k < diag.size(); k++) {
            if (is_zero(c[k])) continue;
            Eidx di = diag[k];
            if (di == NONE) sperror("Singular Matrix");
            Eidx e = el[di].next_in_col;
            while (e != NONE) {
                c[el[e].row] = c[el[e].row] - c[k] * el[e].val;
                e = el[e].next_in_col;
            }
        }
        for (size_t kk = diag.size(); kk-- > 0;) {
            Eidx di = diag[kk];
            if (di == NONE) sperror("Singular Matrix");
            Eidx ep = el[di].next_in_row;
            while (ep != NONE) {
                c[kk] = c[kk] - c[el[ep].col] * el[ep].val;
                ep = el[ep].next_in_row;
            }
            c[kk] = c[kk] / el[di].val;
        }
        std::vector<T> soln(c.size(), zero<T>());
        if (!axes[COLS].has_mapping) sperror("Missing Column Mapping");
        for (size_t k = 0; k < c.size(); k++) soln[k] = c[axes[COLS].e2i[k]];
        return soln;
    }
};

/* ------------------------------------------------------------------------------------------
 * Options and analysis state (analysis.rs:389-449, 641-693)
 * ------------------------------------------------------------------------------------------ */
struct Options {
    double temp = 300.15, tnom = 300.15, gmin = 1e-12, iabstol = 1e-12, reltol = 1e-3;
};
enum AnKind { AN_OP, AN_TRAN, AN_AC };
struct TranState {
    double t = 0.0, dt = 0.0;
};
struct ChargeInteg {
    double g = 0.0, i = 0.0, rhs = 0.0;
};
struct AnalysisInfo {
    AnKind kind = AN_OP;
    const TranState* tran = nullptr;
    double omega = 0.0;
};
// TranState::integrate, BE branch (analysis.rs:420-428)
// Planted mistakes (0 = none): tests set one to show that a closed-form check fails on it.
//   1  Mos1 transient: the gate-source integration kept and the gate-drain one stored apart (the
//      reference overwrites tr.gs with the gate-drain integration, mos.rs:787-802)
//   2  Mos1 AC: the (G, dr) admittance stamped once (the reference stamps it twice, mos.rs:951,954)
//   3  Diode: depletion charge without the transit-time factor (diode.rs:307-309 multiplies by tt)
//   4  Meyer gate-bulk capacitance below threshold over phi instead of 2 phi (mos.rs:727-735)
//   5  Meyer saturation gate-source capacitance cox/2 instead of cox/3 (mos.rs:740-743)
//   6  BSIM4 commit skipped: charges integrate against the device's initial point (bsim4solver.rs:3765-3767)
//   7  BSIM4 bulk-row right-hand-side stamp dropped (stamp.rs: the bNodePrime RHS entry)
static int g_plant = 0;

static inline ChargeInteg integq(const TranState& st, double dq, double dq_dv, double vguess) {
    double dt = st.dt;
    double g = dq_dv / dt;
    double i = dq / dt;
    double rhs = i - g * vguess;
    return ChargeInteg{g, i, rhs};
}

namespace consts {  // comps/mod.rs:24-37
static const double KB = 1.3806226e-23;
static const double Q = 1.6021918e-19;
static const double KB_OVER_Q = KB / Q;
static const double KELVIN_TO_C = 273.15;
static const double TEMP_REF = KELVIN_TO_C + 27.0;
static const double SIO2_PERMITTIVITY = 3.9 * 8.854214871e-12;
static const double SQRT2 = 1.4142135624;
}  // namespace consts

/* ------------------------------------------------------------------------------------------
 * Variables and stamps (analysis.rs:22-123)
 * ------------------------------------------------------------------------------------------ */
using VarIdx = long;  // Option<VarIndex>: -1 is None
template <class T> struct Stamps {
    std::vector<std::pair<Eidx, T>> g;
    std::vector<std::pair<VarIdx, T>> b;
};
template <class T> static inline T var_get(const std::vector<T>& vals, VarIdx i) {
    return i < 0 ? zero<T>() : vals[(size_t)i];
}
static Eidx make_matrix_elem(Maker& m, VarIdx row, VarIdx col) {  // comps/mod.rs:349-354
    if (row >= 0 && col >= 0) return m.make((size_t)row, (size_t)col);
    return NONE;
}

struct Component {
    virtual void create_matrix_elems(Maker& m) = 0;
    virtual void commit() {}
    virtual void update(double) {}
    virtual Stamps<double> load(const std::vector<double>& x, const AnalysisInfo& an, const Options& o) = 0;
    virtual Stamps<Cx> load_ac(const std::vector<Cx>&, const AnalysisInfo&, const Options&) {
        sperror("AC Not Implemented For This Component!");
    }
    virtual ~Component() {}
};

/* Vsrc: comps/mod.rs:95-150 */
struct Vsrc : Component {
    double v, acm;
    VarIdx p, n, ivar;
    Eidx pi = NONE, ip = NONE, ni = NONE, in_ = NONE;
    Vsrc(double vdc, double acm_, VarIdx p_, VarIdx n_, VarIdx iv) : v(vdc), acm(acm_), p(p_), n(n_), ivar(iv) {}
    void update(double val) override { v = val; }
    void create_matrix_elems(Maker& m) override {
        pi = make_matrix_elem(m, p, ivar);
        ip = make_matrix_elem(m, ivar, p);
        ni = make_matrix_elem(m, n, ivar);
        in_ = make_matrix_elem(m, ivar, n);
    }
    Stamps<double> load(const std::vector<double>&, const AnalysisInfo&, const Options&) override {
        return Stamps<double>{{{pi, 1.0}, {ip, 1.0}, {ni, -1.0}, {in_, -1.0}}, {{ivar, v}}};
    }
    Stamps<Cx> load_ac(const std::vector<Cx>&, const AnalysisInfo&, const Options&) override {
        return Stamps<Cx>{{{pi, {1.0, 0.0}}, {ip, {1.0, 0.0}}, {ni, {-1.0, 0.0}}, {in_, {-1.0, 0.0}}}, {{ivar, {acm, 0.0}}}};
    }
};

/* Capacitor: comps/mod.rs:152-239 */
struct Capacitor : Component {
    double c;
    VarIdx p, n;
    Eidx pp = NONE, nn = NONE, pn = NONE, np = NONE;
    struct Op {
        double v = 0, q = 0, i = 0;
    } op, guess;
    Capacitor(double c_, VarIdx p_, VarIdx n_) : c(c_), p(p_), n(n_) {}
    void create_matrix_elems(Maker& m) override {
        pp = make_matrix_elem(m, p, p);
        pn = make_matrix_elem(m, p, n);
        np = make_matrix_elem(m, n, p);
        nn = make_matrix_elem(m, n, n);
    }
    void commit() override { op = guess; }
    Stamps<double> load(const std::vector<double>& x, const AnalysisInfo& an, const Options&) override {
        double vd = var_get(x, p) - var_get(x, n);
        double q = c * vd;
        if (an.kind == AN_OP) {
            guess = Op{vd, q, 0.0};
            return Stamps<double>{};
        }
        if (an.kind != AN_TRAN) sperror("HOW WE GET HERE?!?");
        ChargeInteg ci = integq(*an.tran, q - op.q, c, vd);
        guess = Op{vd, q, ci.i};
        double g = ci.g, rhs = ci.rhs;
        return Stamps<double>{{{pp, g}, {nn, g}, {pn, -g}, {np, -g}}, {{p, -rhs}, {n, rhs}}};
    }
    Stamps<Cx> load_ac(const std::vector<Cx>&, const AnalysisInfo& an, const Options&) override {
        double w = an.omega;
        return Stamps<Cx>{{{pp, {0.0, w * c}}, {nn, {0.0, w * c}}, {pn, {0.0, -w * c}}, {np, {0.0, -w * c}}}, {}};
    }
};

/* Resistor: comps/mod.rs:271-323 */
struct Resistor : Component {
    double g;
    VarIdx t[2];
    Eidx m[2][2] = {{NONE, NONE}, {NONE, NONE}};
    Resistor(double g_, VarIdx p, VarIdx n) : g(g_), t{p, n} {}
    void update(double val) override { g = val; }
    void create_matrix_elems(Maker& mk) override {
        for (int l = 0; l < 2; l++)
            for (int r = 0; r < 2; r++) m[l][r] = make_matrix_elem(mk, t[l], t[r]);
    }
    Stamps<double> load(const std::vector<double>&, const AnalysisInfo&, const Options&) override {
        return Stamps<double>{{{m[0][0], g}, {m[1][1], g}, {m[0][1], -g}, {m[1][0], -g}}, {}};
    }
    Stamps<Cx> load_ac(const std::vector<Cx>&, const AnalysisInfo&, const Options&) override {
        return Stamps<Cx>{{{m[0][0], {g, 0.0}}, {m[1][1], {g, 0.0}}, {m[0][1], {-g, 0.0}}, {m[1][0], {-g, 0.0}}}, {}};
    }
};

/* Isrc: comps/mod.rs:326-346 */
struct Isrc : Component {
    double i;
    VarIdx p, n;
    Isrc(double i_, VarIdx p_, VarIdx n_) : i(i_), p(p_), n(n_) {}
    void create_matrix_elems(Maker&) override {}
    Stamps<double> load(const std::vector<double>&, const AnalysisInfo&, const Options&) override {
        return Stamps<double>{{}, {{p, i}, {n, -i}}};
    }
};

/* ------------------------------------------------------------------------------------------
 * Mos1 (comps/mos.rs)
 * ------------------------------------------------------------------------------------------ */
static inline bool has(const s21_params& p, int f) { return (p.set >> f) & 1ull; }
static inline double opt(const s21_params& p, int f, double dflt) { return has(p, f) ? p.v[f] : dflt; }

struct Mos1Model {  // mos.rs:107-138
    double p;       // polarity: NMOS +1, PMOS -1
    double vt0, kp, gamma, cox_per_area, phi, lambda, cbd, cbs, is, pb, cgso, cgdo, cgbo, cj, mj, cjsw, mjsw, js, tox, ld, fc,
        tnom, kf, af;
    bool has_rd, has_rs, has_rsh;
    double rd, rs, rsh;
};
static Mos1Model mos1_resolve(const s21_params& s) {  // Mos1Model::resolve, mos.rs:140-237
    using namespace consts;
    Mos1Model m{};
    bool pmos = has(s, 2) && s.v[2] == 1.0;
    m.p = pmos ? -1.0 : 1.0;
    double tnom = has(s, 30) ? s.v[30] + KELVIN_TO_C : TEMP_REF;
    double fact1 = tnom / TEMP_REF;
    double vtnom = tnom * KB_OVER_Q;
    double kt1 = KB * tnom;
    double egfet1 = 1.16 - (7.02e-4 * (tnom * tnom)) / (tnom + 1108.0);
    double arg1 = -egfet1 / 2.0 / kt1 + 1.1150877 / (KB * 2.0 * TEMP_REF);
    double pbfact1 = -2.0 * vtnom * (1.5 * std::log(fact1) + Q * arg1);
    (void)pbfact1;
    double cox_per_area = 0.0;
    double vt0 = opt(s, 3, 0.0);
    double kp = opt(s, 4, 2.0e-5);
    double phi = opt(s, 6, 0.6);
    double gamma = opt(s, 5, 0.0);
    if (has(s, 23)) {
        double tox = s.v[23];
        cox_per_area = SIO2_PERMITTIVITY / tox;
        if (!has(s, 4)) {
            double u0 = opt(s, 25, 600.0);
            kp = u0 * cox_per_area * 1e-4;
        }
        if (has(s, 27)) {
            double nsub = s.v[27];
            if (nsub * 1e6 <= 1.45e16) sperror("Invalid Mos1 Substrate Doping nsub < ni (1.45e16)");
            if (!has(s, 6)) {
                phi = 2.0 * vtnom * std::log(nsub * 1e6 / 1.45e16);
                phi = std::fmax(phi, 0.1);
            }
            double fermis = m.p * 0.5 * phi;
            double wkfng = 3.2;
            double gate_type = 1.0;
            if (has(s, 28)) {
                double val = s.v[28];
                if (val > 1 || val < -1) sperror("Invalid Mos1 tps");
                gate_type = val;
            }
            if (gate_type != 0.0) {
                double fermig = m.p * gate_type * 0.5 * egfet1;
                wkfng = 3.25 + 0.5 * egfet1 - fermig;
            }
            if (!has(s, 5)) gamma = std::sqrt(2.0 * 11.70 * 8.854214871e-12 * Q * nsub * 1e6) / cox_per_area;
            if (!has(s, 3)) {
                double nss = opt(s, 29, 0.0);
                double wkfngs = wkfng - (3.25 + 0.5 * egfet1 + fermis);
                double vfb = wkfngs - nss * 1e4 * Q / cox_per_area;
                vt0 = vfb + m.p * (gamma * std::sqrt(phi) + phi);
            }
        }
    }
    m.vt0 = vt0;
    m.kp = kp;
    m.cox_per_area = cox_per_area;
    m.gamma = gamma;
    m.phi = phi;
    m.tnom = tnom;
    m.lambda = opt(s, 7, 0.0);
    m.pb = opt(s, 13, 0.8);
    m.cbd = opt(s, 10, 0.0);
    m.cbs = opt(s, 11, 0.0);
    m.cgso = opt(s, 14, 0.0);
    m.cgdo = opt(s, 15, 0.0);
    m.cgbo = opt(s, 16, 0.0);
    m.cj = opt(s, 18, 0.0);
    m.cjsw = opt(s, 20, 0.0);
    m.mj = opt(s, 19, 0.5);
    m.mjsw = opt(s, 21, 0.5);
    m.is = opt(s, 12, 1.0e-14);
    m.js = opt(s, 22, 1.0e-8);
    m.tox = opt(s, 23, 1.0e-7);
    m.ld = opt(s, 24, 0.0);
    m.fc = opt(s, 26, 0.5);
    m.kf = opt(s, 31, 0.0);
    m.af = opt(s, 32, 1.0);
    m.has_rd = has(s, 8);
    m.rd = opt(s, 8, 0.0);
    m.has_rs = has(s, 9);
    m.rs = opt(s, 9, 0.0);
    m.has_rsh = has(s, 17);
    m.rsh = opt(s, 17, 0.0);
    return m;
}
struct Mos1Inst {  // Mos1InstanceParams::resolve, mos.rs:271-290
    double m, l, w, a_d, a_s, pd, ps, nrd, nrs;
    bool has_temp;
};
static Mos1Inst mos1_inst_resolve(const s21_params& s) {
    return Mos1Inst{opt(s, 2, 0.0), opt(s, 3, 1e-6), opt(s, 4, 1e-6), opt(s, 5, 1e-12), opt(s, 6, 1e-12),
                    opt(s, 7, 1e-6),  opt(s, 8, 1e-6), opt(s, 9, 1.0),  opt(s, 10, 1.0),  has(s, 11)};
}
struct MosJunction {  // mos.rs:488-501
    double area, isat, depletion_threshold, bulkpot_t, vcrit, czb, czbsw, f2, f3, f4;
    // MosJunction::qc, mos.rs:504-521
    void qc(double v, const Mos1Model& model, double& q, double& c) const {
        if (czb == 0.0 && czbsw == 0.0) {
            q = 0.0;
            c = 0.0;
            return;
        }
        if (v < depletion_threshold) {
            double arg = 1.0 - v / bulkpot_t;
            double sarg = std::exp(-model.mj * std::log(arg));
            double sargsw = std::exp(-model.mjsw * std::log(arg));
            q = bulkpot_t * (czb * (1.0 - arg * sarg) / (1.0 - model.mj) + czbsw * (1.0 - arg * sargsw) / (1.0 - model.mjsw));
            c = czb * sarg + czbsw * sargsw;
        } else {
            q = f4 + v * (f2 + v * f3 / 2.0);
            c = f2 + v * f3;
        }
    }
};
struct Mos1Intp {  // Mos1InternalParams, mos.rs:300-317
    double temp, vtherm, vt0_t, kp_t, phi_t, beta, cox, cgs_ov, cgd_ov, cgb_ov, leff;
    MosJunction drain_junc, source_junc;
    double grd, grs;
};
static Mos1Intp mos1_derive(const Mos1Model& model, const Mos1Inst& inst, const Options& opts) {  // mos.rs:320-476
    using namespace consts;
    if (inst.has_temp) sperror("Mos1 Instance Temperatures Are Not Supported");
    double temp = opts.temp;
    double fact1 = model.tnom / TEMP_REF;
    double vtnom = model.tnom * KB_OVER_Q;
    double kt1 = KB * model.tnom;
    double egfet1 = 1.16 - (7.02e-4 * (model.tnom * model.tnom)) / (model.tnom + 1108.0);
    double arg1 = -egfet1 / 2.0 / kt1 + 1.1150877 / (KB * 2.0 * TEMP_REF);
    double pbfact1 = -2.0 * vtnom * (1.5 * std::log(fact1) + Q * arg1);
    double kt = temp * KB;
    double vtherm = temp * KB_OVER_Q;
    double temp_ratio = temp / model.tnom;
    double fact2 = temp / TEMP_REF;
    double egfet = 1.16 - (7.02e-4 * (temp * temp)) / (temp + 1108.0);
    double arg = -egfet / 2.0 / kt + 1.1150877 / (KB * 2.0 * TEMP_REF);
    double pbfact = -2.0 * vtherm * (1.5 * std::log(fact2) + Q * arg);
    double leff = inst.l - 2.0 * model.ld;
    if (leff < 0.0) sperror("Mos1 Effective Length < 0");
    double phio = (model.phi - pbfact1) / fact1;
    double phi_t = fact2 * phio + pbfact;
    double vbi_t = model.vt0 - model.p * (model.gamma * std::sqrt(model.phi)) + 0.5 * (egfet1 - egfet) +
                   model.p * 0.5 * (phi_t - model.phi);
    double vt0_t = vbi_t + model.p * model.gamma * std::sqrt(phi_t);
    double isat_t = model.is * std::exp(-egfet / vtherm + egfet1 / vtnom);
    double jsat_t = model.js * std::exp(-egfet / vtherm + egfet1 / vtnom);
    double pbo = (model.pb - pbfact1) / fact1;
    double gmaold = (model.pb - pbo) / pbo;
    double capfact = 1.0 / (1.0 + model.mj * (4e-4 * (model.tnom - TEMP_REF) - gmaold));
    double cbd_t = model.cbd * capfact;
    double cbs_t = model.cbs * capfact;
    double cj_t = model.cj * capfact;
    capfact = 1.0 / (1.0 + model.mjsw * (4e-4 * (model.tnom - TEMP_REF) - gmaold));
    double cjsw_t = model.cjsw * capfact;
    double bulkpot_t = fact2 * pbo + pbfact;
    double gmanew = (bulkpot_t - pbo) / pbo;
    capfact = 1.0 / (1.0 + model.mj * (4e-4 * (temp - TEMP_REF) - gmanew));
    cbd_t *= capfact;
    cbs_t *= capfact;
    cj_t *= capfact;
    capfact = 1.0 / (1.0 + model.mjsw * (4e-4 * (temp - TEMP_REF) - gmanew));
    cjsw_t *= capfact;
    double depletion_threshold = model.fc * bulkpot_t;
    double argc = 1.0 - model.fc;
    double sarg = std::exp((-model.mj) * std::log(argc));
    double sargsw = std::exp((-model.mjsw) * std::log(argc));
    bool use_default_isat = jsat_t == 0.0 || inst.a_d == 0.0 || inst.a_s == 0.0;
    auto junc_new = [&](double area, double perim, bool is_drain) {
        MosJunction j{};
        double isat = use_default_isat ? isat_t : jsat_t * area;
        double vcrit = vtherm * std::log(vtherm / (SQRT2 * isat));
        double czb;
        if (is_drain)
            czb = (model.cbd == 0.0) ? cj_t * area : cbd_t;
        else
            czb = (model.cbs == 0.0) ? cj_t * area : cbs_t;
        double czbsw = cjsw_t * perim;
        double f2 = czb * (1.0 - model.fc * (1.0 + model.mj)) * sarg / argc +
                    czbsw * (1.0 - model.fc * (1.0 + model.mjsw)) * sargsw / argc;
        double f3 = czb * model.mj * sarg / argc / bulkpot_t + czbsw * model.mjsw * sargsw / argc / bulkpot_t;
        double f4 = czb * bulkpot_t * (1.0 - argc * sarg) / (1.0 - model.mj) +
                    czbsw * bulkpot_t * (1.0 - argc * sargsw) / (1.0 - model.mjsw) -
                    f3 / 2.0 * (depletion_threshold * depletion_threshold) - depletion_threshold * f2;
        j.area = area;
        j.isat = isat;
        j.depletion_threshold = depletion_threshold;
        j.bulkpot_t = bulkpot_t;
        j.vcrit = vcrit;
        j.czb = czb;
        j.czbsw = czbsw;
        j.f2 = f2;
        j.f3 = f3;
        j.f4 = f4;
        return j;
    };
    Mos1Intp ip{};
    ip.drain_junc = junc_new(inst.a_d, inst.pd, true);
    ip.source_junc = junc_new(inst.a_s, inst.ps, false);
    double grs = 0.0, grd = 0.0;
    if (model.has_rs)
        grs = model.rs <= 0.0 ? 0.0 : 1.0 / model.rs;
    else if (model.has_rsh)
        grs = model.rsh <= 0.0 ? 0.0 : 1.0 / model.rsh / inst.nrs;
    if (model.has_rd)
        grd = model.rd <= 0.0 ? 0.0 : 1.0 / model.rd;
    else if (model.has_rsh)
        grd = model.rsh <= 0.0 ? 0.0 : 1.0 / model.rsh / inst.nrd;
    double kp_t = model.kp / temp_ratio * std::sqrt(temp_ratio);
    ip.vt0_t = vt0_t;
    ip.kp_t = kp_t;
    ip.temp = temp;
    ip.vtherm = vtherm;
    ip.leff = leff;
    ip.cox = model.cox_per_area * leff * inst.w;
    ip.beta = kp_t * inst.w / leff;
    ip.phi_t = phi_t;
    ip.cgs_ov = inst.w * model.cgso;
    ip.cgd_ov = inst.w * model.cgdo;
    ip.cgb_ov = leff * model.cgbo;
    ip.grs = grs;
    ip.grd = grd;
    return ip;
}

struct Mos1OpPoint {  // mos.rs:524-556
    double ids = 0, vgs = 0, vds = 0, vgd = 0, vgb = 0, vdb = 0, vsb = 0, gm = 0, gds = 0, gmbs = 0, gbs = 0, gbd = 0,
           cgs = 0, cgd = 0, cgb = 0, cbs = 0, cbd = 0;
    bool reversed = false;
    ChargeInteg tr_gs, tr_gd, tr_gb, tr_bs, tr_bd;
};
enum { MD = 0, MG = 1, MS = 2, MB = 3, MDP = 4, MSP = 5 };  // Mos1Var, mos.rs:569-577

struct Mos1 : Component {
    Mos1Model model;
    Mos1Intp intp;
    VarIdx ports[6];
    Mos1OpPoint op, guess;
    Eidx matps[6][6];
    Mos1(const Mos1Model& m, const Mos1Intp& ip, const VarIdx pr[6]) : model(m), intp(ip) {
        for (int k = 0; k < 6; k++) ports[k] = pr[k];
        for (auto& r : matps)
            for (auto& c : r) c = NONE;
    }
    void create_matrix_elems(Maker& mk) override {  // mos.rs:896-903
        static const int order[6] = {MG, MD, MS, MB, MDP, MSP};
        for (int a = 0; a < 6; a++)
            for (int b = 0; b < 6; b++) matps[order[a]][order[b]] = make_matrix_elem(mk, ports[order[a]], ports[order[b]]);
    }
    void commit() override { op = guess; }
    Stamps<double> load(const std::vector<double>& x, const AnalysisInfo& an, const Options& opts) override {
        // Mos1::load + op_stamp, mos.rs:649-913
        double vd_ = var_get(x, ports[MD]), vg_ = var_get(x, ports[MG]), vs_ = var_get(x, ports[MS]),
               vb_ = var_get(x, ports[MB]);
        double gmin = opts.gmin;
        double p = model.p;
        bool reversed = p * (vd_ - vs_) < 0.0;
        double vd = reversed ? vs_ : vd_;
        double vs = reversed ? vd_ : vs_;
        double vgs = p * (vg_ - vs);
        double vgd = p * (vg_ - vd);
        double vds = p * (vd - vs);
        double vgb = p * (vg_ - vb_);
        double vsb = p * (vs - vb_);
        double vdb = p * (vd - vb_);
        double von = vsb > 0.0 ? intp.vt0_t + model.gamma * (std::sqrt(intp.phi_t + vsb) - std::sqrt(intp.phi_t)) : intp.vt0_t;
        double vov = vgs - von;
        double vdsat = std::fmax(vov, 0.0);
        double ids = 0.0, gm = 0.0, gds = 0.0, gmbs = 0.0;
        if (vov > 0.0) {
            if (vds >= vov) {
                ids = intp.beta / 2.0 * (vov * vov) * (1.0 + model.lambda * vds);
                gm = intp.beta * vov * (1.0 + model.lambda * vds);
                gds = model.lambda * intp.beta / 2.0 * (vov * vov);
            } else {
                ids = intp.beta * (vov * vds - (vds * vds) / 2.0) * (1.0 + model.lambda * vds);
                gm = intp.beta * vds * (1.0 + model.lambda * vds);
                gds = intp.beta * ((vov - vds) * (1.0 + model.lambda * vds) + model.lambda * ((vov * vds) - (vds * vds) / 2.0));
            }
            gmbs = (intp.phi_t + vsb > 0.0) ? gm * model.gamma / 2.0 / std::sqrt(intp.phi_t + vsb) : 0.0;
        }
        double vtherm = intp.vtherm;
        const MosJunction& bs_junc = !reversed ? intp.source_junc : intp.drain_junc;
        const MosJunction& bd_junc = !reversed ? intp.drain_junc : intp.source_junc;
        double ibs = bs_junc.isat * (std::exp(-vsb / vtherm) - 1.0);
        double gbs = (bs_junc.isat / vtherm) * std::exp(-vsb / vtherm) + gmin;
        double ibs_rhs = ibs + vsb * gbs;
        double ibd = bd_junc.isat * (std::exp(-vdb / vtherm) - 1.0);
        double gbd = (bd_junc.isat / vtherm) * std::exp(-vdb / vtherm) + gmin;
        double ibd_rhs = ibd + vdb * gbd;
        double cox = intp.cox;
        double cgs1, cgd1, cgb1;
        if (vov <= -intp.phi_t) {
            cgb1 = cox / 2.0;
            cgs1 = 0.0;
            cgd1 = 0.0;
        } else if (vov <= -intp.phi_t / 2.0) {
            cgb1 = -vov * cox / ((g_plant == 4 ? 1.0 : 2.0) * intp.phi_t);
            cgs1 = 0.0;
            cgd1 = 0.0;
        } else if (vov <= 0.0) {
            cgb1 = -vov * cox / ((g_plant == 4 ? 1.0 : 2.0) * intp.phi_t);
            cgs1 = vov * cox / (1.5 * intp.phi_t) + cox / 3.0;
            cgd1 = 0.0;
        } else if (vdsat <= vds) {
            cgs1 = cox / (g_plant == 5 ? 2.0 : 3.0);
            cgd1 = 0.0;
            cgb1 = 0.0;
        } else {
            double vddif = 2.0 * vdsat - vds;
            double vddif1 = vdsat - vds;
            double vddif2 = vddif * vddif;
            cgd1 = cox * (1.0 - vdsat * vdsat / vddif2) / 3.0;
            cgs1 = cox * (1.0 - vddif1 * vddif1 / vddif2) / 3.0;
            cgb1 = 0.0;
        }
        bool same = reversed == op.reversed;
        double cgs2 = (op.cgs == 0.0) ? cgs1 : (same ? op.cgs : op.cgd);
        double cgs = cgs1 + cgs2 + intp.cgs_ov;
        double cgd = cgd1 + intp.cgd_ov + (same ? op.cgd : op.cgs);
        double cgb = cgb1 + intp.cgb_ov + op.cgb;
        double qbs_, cbs, qbd_, cbd;
        bs_junc.qc(-vsb, model, qbs_, cbs);
        bd_junc.qc(-vdb, model, qbd_, cbd);
        ChargeInteg tr_gs, tr_gd, tr_gb, tr_bs, tr_bd;
        if (an.kind == AN_TRAN) {
            const TranState& st = *an.tran;
            {
                double dqgs = same ? (vgs - op.vgs) * cgs : (vgs - op.vgd) * cgs;
                tr_gs = integq(st, dqgs, cgs, vgs);
            }
            {
                double dqgd = same ? (vgd - op.vgd) * cgd : (vgd - op.vgs) * cgd;
                if (g_plant == 1)
                    tr_gd = integq(st, dqgd, cgd, vgd);
                else
                    tr_gs = integq(st, dqgd, cgd, vgd);  // overwrites tr.gs, as mos.rs:801 does
            }
            {
                double dqgb = (vgb - op.vgb) * cgb;
                tr_gb = integq(st, dqgb, cgb, vgb);
            }
            {
                double dqbs = same ? (-vsb + op.vsb) * cbs : (-vsb + op.vdb) * cbs;
                double dqbd = same ? (-vdb + op.vdb) * cbd : (-vdb + op.vsb) * cbd;
                tr_bs = integq(st, dqbs, cbs, -vsb);
                tr_bd = integq(st, dqbd, cbd, -vdb);
            }
        }
        double irhs = ids - gm * vgs - gds * vds;
        int sr = !reversed ? MSP : MDP, sx = !reversed ? MS : MD, dr = !reversed ? MDP : MSP, dx = !reversed ? MD : MS;
        double grd = intp.grd, grs = intp.grs;
        Stamps<double> st;
        st.g = {
            {matps[dr][dr], gds + grd + gbd + tr_gd.g},
            {matps[sr][sr], gm + gds + grs + gbs + gmbs + tr_gs.g},
            {matps[dr][sr], -gm - gds - gmbs},
            {matps[sr][dr], -gds},
            {matps[dr][MG], gm - tr_gd.g},
            {matps[sr][MG], -gm - tr_gs.g},
            {matps[MG][MG], (tr_gd.g + tr_gs.g + tr_gb.g)},
            {matps[MB][MB], (gbd + gbs + tr_gb.g)},
            {matps[MG][MB], -tr_gb.g},
            {matps[MG][dr], -tr_gd.g},
            {matps[MG][sr], -tr_gs.g},
            {matps[MB][MG], -tr_gb.g},
            {matps[MB][dr], -gbd},
            {matps[MB][sr], -gbs},
            {matps[dr][MB], -gbd + gmbs},
            {matps[sr][MB], -gbs - gmbs},
            {matps[dx][dr], -grd},
            {matps[dr][dx], -grd},
            {matps[dx][dx], grd},
            {matps[sx][sr], -grs},
            {matps[sr][sx], -grs},
            {matps[sx][sx], grs},
        };
        st.b = {
            {ports[dr], p * (-irhs + ibd_rhs + tr_gd.rhs)},
            {ports[sr], p * (irhs + ibs_rhs + tr_gs.rhs)},
            {ports[MG], -p * (tr_gs.rhs + tr_gb.rhs + tr_gd.rhs)},
            {ports[MB], -p * (ibd_rhs + ibs_rhs - tr_gb.rhs)},
        };
        Mos1OpPoint g;
        g.ids = ids;
        g.vgs = vgs;
        g.vds = vds;
        g.vgd = vgd;
        g.vgb = vgb;
        g.vdb = vdb;
        g.vsb = vsb;
        g.gm = gm;
        g.gds = gds;
        g.gmbs = gmbs;
        g.gbs = gbs;
        g.gbd = gbd;
        g.reversed = reversed;
        g.cgs = cgs1;
        g.cgd = cgd1;
        g.cgb = cgb1;
        g.cbs = cbs;
        g.cbd = cbd;
        g.tr_gs = tr_gs;
        g.tr_gd = tr_gd;
        g.tr_gb = tr_gb;
        g.tr_bs = tr_bs;
        g.tr_bd = tr_bd;
        guess = g;
        return st;
    }
    Stamps<Cx> load_ac(const std::vector<Cx>&, const AnalysisInfo& an, const Options&) override {  // mos.rs:914-968
        double omega = an.omega;
        double gm = op.gm, gds = op.gds, gmbs = op.gmbs, gbs = op.gbs, gbd = op.gbd;
        double gcgs = omega * op.cgs, gcgd = omega * op.cgd, gcgb = omega * op.cgb;
        int sr = !op.reversed ? MSP : MDP, sx = !op.reversed ? MS : MD, dr = !op.reversed ? MDP : MSP, dx = !op.reversed ? MD : MS;
        Stamps<Cx> st;
        st.g = {
            {matps[dr][dr], {gds + intp.grd + gbd, gcgd}},
            {matps[sr][sr], {gm + gds + intp.grs + gbs + gmbs, gcgs}},
            {matps[dr][sr], {-gm - gds - gmbs, 0.0}},
            {matps[sr][dr], {-gds, 0.0}},
            {matps[dr][MG], {gm, -gcgd}},
            {matps[sr][MG], {-gm, -gcgs}},
            {matps[MG][MG], {0.0, gcgd + gcgs + gcgb}},
            {matps[MB][MB], {gbd + gbs, gcgb}},
            {matps[MG][MB], {0.0, -gcgb}},
            {matps[MG][dr], {0.0, -gcgd}},
            {matps[MG][sr], {0.0, -gcgs}},
            {matps[MB][MG], {0.0, -gcgb}},
            {g_plant == 2 ? -1 : matps[MG][dr], {0.0, -gcgd}},  // stamped twice, as mos.rs:951,954
            {matps[MB][dr], {-gbd, 0.0}},
            {matps[MB][sr], {-gbs, 0.0}},
            {matps[dr][MB], {-gbd + gmbs, 0.0}},
            {matps[sr][MB], {-gbs - gmbs, 0.0}},
            {matps[dx][dr], {-intp.grd, 0.0}},
            {matps[dr][dx], {-intp.grd, 0.0}},
            {matps[dx][dx], {intp.grd, 0.0}},
            {matps[sx][sr], {-intp.grs, 0.0}},
            {matps[sr][sx], {-intp.grs, 0.0}},
            {matps[sx][sx], {intp.grs, 0.0}},
        };
        return st;
    }
};

/* Mos0: mos.rs:1007-1100 */
struct Mos0 : Component {
    double p, vth = 0.25, beta = 50e-3, lam = 3e-3;
    VarIdx ports[4];
    Eidx matps[4][4];
    Mos0(int mos_type, const VarIdx pr[4]) : p(mos_type == 1 ? -1.0 : 1.0) {
        for (int k = 0; k < 4; k++) ports[k] = pr[k];
        for (auto& r : matps)
            for (auto& c : r) c = NONE;
    }
    void create_matrix_elems(Maker& mk) override {
        static const int pr[6][2] = {{MD, MD}, {MS, MS}, {MD, MS}, {MS, MD}, {MD, MG}, {MS, MG}};
        for (auto& t : pr) matps[t[0]][t[1]] = make_matrix_elem(mk, ports[t[0]], ports[t[1]]);
    }
    Stamps<double> load(const std::vector<double>& x, const AnalysisInfo&, const Options& opts) override {
        double gmin = opts.gmin;
        double vg = var_get(x, ports[MG]), vd = var_get(x, ports[MD]), vs = var_get(x, ports[MS]);
        double vds1 = p * (vd - vs);
        bool reversed = vds1 < 0.0;
        double vgs = reversed ? p * (vg - vd) : p * (vg - vs);
        double vds = reversed ? -vds1 : vds1;
        double vov = vgs - vth;
        double ids = 0.0, gm = 0.0, gds = 0.0;
        if (vov > 0.0) {
            if (vds >= vov) {
                ids = beta / 2.0 * (vov * vov) * (1.0 + lam * vds);
                gm = beta * vov * (1.0 + lam * vds);
                gds = lam * beta / 2.0 * (vov * vov);
            } else {
                ids = beta * (vov * vds - (vds * vds) / 2.0) * (1.0 + lam * vds);
                gm = beta * vds * (1.0 + lam * vds);
                gds = beta * ((vov - vds) * (1.0 + lam * vds) + lam * ((vov * vds) - (vds * vds) / 2.0));
            }
        }
        int sr = !reversed ? MS : MD, dr = !reversed ? MD : MS;
        double irhs = ids - gm * vgs - gds * vds;
        Stamps<double> st;
        st.g = {{matps[dr][dr], gds + gmin},
                {matps[sr][sr], (gm + gds + gmin)},
                {matps[dr][sr], -(gm + gds + gmin)},
                {matps[sr][dr], -gds - gmin},
                {matps[dr][MG], gm},
                {matps[sr][MG], -gm}};
        st.b = {{ports[dr], -p * irhs}, {ports[sr], p * irhs}};
        return st;
    }
};

/* ------------------------------------------------------------------------------------------
 * Diode (comps/diode.rs)
 * ------------------------------------------------------------------------------------------ */
struct DiodeModel {  // diode.rs:52-70
    double tnom, is, n, tt, vj, m, eg, xti, kf, af, fc, bv, ibv, rs, cj0;
    bool has_rs() const { return rs != 0.0; }
    bool has_bv() const { return bv != 0.0; }
};
static DiodeModel diode_model_from(const s21_params& s) {
    return DiodeModel{opt(s, 2, 300.15), opt(s, 3, 1e-14), opt(s, 4, 1.0), opt(s, 5, 0.0),  opt(s, 6, 1.0),
                      opt(s, 7, 0.5),    opt(s, 8, 1.11),  opt(s, 9, 3.0), opt(s, 10, 0.0), opt(s, 11, 1.0),
                      opt(s, 12, 0.5),   opt(s, 13, 0.0),  opt(s, 14, 1e-3), opt(s, 15, 0.0), opt(s, 16, 0.0)};
}
struct DiodeIntp {  // diode.rs:127-143
    double vt, vte, vcrit, isat, gspr, cz, cz2, dep_threshold, f1, f2, f3, bv;
};
static DiodeIntp diode_derive(const DiodeModel& model, const s21_params& inst, const Options& opts) {  // diode.rs:146-212
    using namespace consts;
    double tnom = model.tnom;
    double temp = has(inst, 5) ? inst.v[5] : opts.temp;
    double area = has(inst, 4) ? inst.v[4] : 1.0;
    double gs = model.has_rs() ? 1.0 / model.rs : 0.0;
    double vt = KB_OVER_Q * temp;
    double vtnom = KB_OVER_Q * tnom;
    double fact2 = temp / TEMP_REF;
    double egfet = 1.16 - (7.02e-4 * temp * temp) / (temp + 1108.0);
    double arg = -egfet / (2.0 * KB * temp) + 1.1150877 / (2.0 * KB * TEMP_REF);
    double pbfact = -2.0 * vt * (1.5 * std::log(fact2) + Q * arg);
    double egfet1 = 1.16 - (7.02e-4 * tnom) / (tnom + 1108.0);
    double arg1 = -egfet1 / (KB * 2.0 * tnom) + 1.1150877 / (2.0 * KB * TEMP_REF);
    double fact1 = tnom / TEMP_REF;
    double pbfact1 = -2.0 * vtnom * (1.5 * std::log(fact1) + Q * arg1);
    double pbo = (model.vj - pbfact1) / fact1;
    double gmaold = (model.vj - pbo) / pbo;
    double cjunc = model.cj0 / (1.0 + model.m * (400e-6 * (tnom - TEMP_REF) - gmaold));
    double vjunc = pbfact + fact2 * pbo;
    double gmanew = (vjunc - pbo) / pbo;
    cjunc *= 1.0 + model.m * (400e-6 * (temp - TEMP_REF) - gmanew);
    (void)cjunc;
    double isat = model.is * std::exp(((temp / tnom) - 1.0) * model.eg / model.n * vt + model.xti / model.n * std::log(temp / tnom));
    double xfc = 1.0 - std::log(model.fc);
    double f1 = vjunc * (1.0 - std::exp(1.0 - model.m * xfc)) / (1.0 - model.m);
    double dep_threshold = model.fc * model.vj;
    double vte = model.n * vt;
    double vcrit = vte * (vte / std::sqrt(2.0) / isat);
    double bv = model.bv;
    if (model.has_bv()) {
        double ibv = model.ibv;
        for (int i = 0; i < 25; i++) bv = model.bv - vt * std::log(ibv / isat + 1.0 - bv / vt);
    }
    double f2 = std::exp(xfc * (1.0 + model.m));
    double f3 = 1.0 - model.fc * (1.0 + model.m);
    double gspr = gs * area;
    double cz = model.cj0 * area;
    double cz2 = cz / f2;
    return DiodeIntp{vt, vte, vcrit, isat, gspr, cz, cz2, dep_threshold, f1, f2, f3, bv};
}
struct Diode : Component {
    DiodeModel model;
    DiodeIntp intp;
    VarIdx pp_, nn_, rr_;  // ports p, n, r
    Eidx pp = NONE, pr = NONE, rp = NONE, rr = NONE, nr = NONE, rn = NONE, nn = NONE;
    struct Op {
        double vd = 0, id = 0, gd = 0, cd = 0, charge = 0, capcur = 0, p = 0;
    } op, guess;
    Diode(const DiodeModel& m, const DiodeIntp& ip, VarIdx p, VarIdx n, VarIdx r) : model(m), intp(ip), pp_(p), nn_(n), rr_(r) {}
    void create_matrix_elems(Maker& mk) override {  // diode.rs:248-256
        pp = make_matrix_elem(mk, pp_, pp_);
        pr = make_matrix_elem(mk, pp_, rr_);
        rp = make_matrix_elem(mk, rr_, pp_);
        rr = make_matrix_elem(mk, rr_, rr_);
        nr = make_matrix_elem(mk, nn_, rr_);
        rn = make_matrix_elem(mk, rr_, nn_);
        nn = make_matrix_elem(mk, nn_, nn_);
    }
    void commit() override { op = guess; }
    double limit(double vd, bool has_past, double past) const {  // diode.rs:228-245
        double vnew = vd;
        double vold = has_past ? past : guess.vd;
        if (vnew <= intp.vcrit || std::fabs(vnew - vold) <= 2.0 * intp.vte) return vnew;
        if (vold > 0.0) {
            double arg = 1.0 + (vnew - vold) / intp.vte;
            if (arg > 0.0) return vold + intp.vte * std::log(arg);
            return intp.vcrit;
        }
        return intp.vte * std::log(vnew / intp.vte);
    }
    Stamps<double> load(const std::vector<double>& x, const AnalysisInfo& an, const Options& opts) override {  // :279-354
        double gmin = opts.gmin;
        double vd = var_get(x, rr_) - var_get(x, nn_);
        if (model.has_bv() && vd < std::fmin(10.0 * intp.vte - intp.bv, 0.0)) {
            double vtemp = limit(-intp.bv, true, intp.bv - guess.vd);
            vd = vtemp - intp.bv;
        } else {
            vd = limit(vd, false, 0.0);
        }
        double id, gd;
        if (!model.has_bv() || vd >= -intp.bv) {
            double e = std::exp(vd / intp.vte);
            id = intp.isat * (e - 1.0) + gmin * vd;
            gd = intp.isat * e / intp.vte + gmin;
        } else {
            double e = std::exp((vd - intp.bv) / intp.vte);
            id = -intp.isat * e + gmin * vd;
            gd = intp.isat * e / intp.vte + gmin;
        }
        double qd, cd;
        if (vd < intp.dep_threshold) {
            double a = 1.0 - vd / model.vj;
            double s = -model.m * std::log(a);
            qd = (g_plant == 3 ? 1.0 : model.tt) * model.vj * intp.cz * (1.0 - a * s) / (1.0 - model.m);
            cd = model.tt * gd + intp.cz * s;
        } else {
            qd = model.tt * id + intp.cz * intp.f1 +
                 intp.cz2 * (intp.f3 * (vd - intp.dep_threshold) +
                             model.m / 2.0 / model.vj * (vd * vd - intp.dep_threshold * intp.dep_threshold));
            cd = model.tt + intp.cz2 * intp.f3 + model.m * vd / model.vj;
        }
        double gc = 0.0, ic = 0.0;
        if (an.kind == AN_TRAN) {
            ChargeInteg ci = integq(*an.tran, qd - op.charge, cd, vd);
            gc = ci.g;
            ic = ci.i;
        }
        id += ic;
        gd += gc;
        guess = Op{vd, id, gd, cd, qd, ic, vd * id};
        double irhs = id - vd * gd;
        Stamps<double> st;
        st.g = {{nn, gd}, {rn, -gd}, {nr, -gd}, {rr, gd + intp.gspr}, {pp, intp.gspr}, {pr, -intp.gspr}, {rp, -intp.gspr}};
        st.b = {{rr_, -irhs}, {nn_, irhs}};
        return st;
    }
};

/* ------------------------------------------------------------------------------------------
 * Solver (analysis.rs:136-346)
 * ------------------------------------------------------------------------------------------ */
/* BSIM4 at the default mode selectors: no internal nodes, so dNodePrime = dNode, gNodePrime = gNodeExt,
 * sNodePrime = sNode, bNodePrime = dbNode = sbNode = bNode and no charge node (bsim4ports.rs:31-112). */
struct Bsim4 : Component {
    s21::b4::B4Par P;
    VarIdx t[4];  // d, g, s, b
    s21::b4::B4Op op{}, guess{};
    Eidx m[4][4];
    Bsim4(const s21::b4::B4Par& par, const VarIdx nd[4]) : P(par) {
        for (int k = 0; k < 4; k++) t[k] = nd[k];
        for (auto& r : m)
            for (auto& c : r) c = NONE;
    }
    void create_matrix_elems(Maker& mk) override {  // create_matps, bsim4solver.rs:28-116 (Q entries: no node)
        enum { D = 0, G = 1, S = 2, B = 3 };
        static const int order[22][2] = {{D, B}, {G, B}, {S, B}, {B, D}, {B, G}, {B, S}, {B, B}, {D, D},
                                         {G, G}, {S, S}, {D, D}, {S, S}, {D, D}, {G, D}, {G, S}, {S, S},
                                         {D, S}, {D, D}, {D, G}, {S, G}, {S, S}, {S, D}};
        for (auto& q : order) m[q[0]][q[1]] = make_matrix_elem(mk, t[q[0]], t[q[1]]);
    }
    void commit() override {
        if (g_plant != 6) op = guess;
    }
    Stamps<double> load(const std::vector<double>& x, const AnalysisInfo& an, const Options& opts) override {
        // load_dc_tr, bsim4solver.rs:134-144: op() from the port voltages, tran_op in a transient, stamp()
        s21::b4::B4Op N;
        s21::b4::op_eval(P, guess, var_get(x, t[0]), var_get(x, t[1]), var_get(x, t[2]), var_get(x, t[3]), opts.gmin,
                         N);
        s21::b4::B4Tran T{};
        if (an.kind == AN_TRAN) s21::b4::tran_eval(P, op, N, an.tran->dt, T);
        guess = N;
        Stamps<double> st;
        s21::b4::stamp(P, N, T, [&](int k, double v) {
            const VarIdx r = t[s21::b4::b4_row(k)];
            const int c = s21::b4::b4_col(k);
            if (c < 0)
                st.b.push_back({g_plant == 7 && s21::b4::b4_row(k) == 3 ? -1 : r, v});
            else
                st.g.push_back({r < 0 || t[c] < 0 ? NONE : m[s21::b4::b4_row(k)][c], v});
        });
        return st;
    }
};

struct Circuit {
    std::vector<std::unique_ptr<Component>> comps;
    std::vector<int> kinds;
    size_t n_vars = 0;
};

static void build_components(const s21_netlist* net, const Options& opts, Circuit& ck) {
    ck.n_vars = (size_t)net->n_vars;
    ck.kinds.assign(net->var_kind, net->var_kind + net->n_vars);
    for (int32_t k = 0; k < net->n_comps; k++) {
        const s21_comp& c = net->comps[k];
        const int32_t* nd = c.node;
        switch (c.type) {
        case S21_COMP_R: ck.comps.emplace_back(new Resistor(c.value, nd[0], nd[1])); break;
        case S21_COMP_C: ck.comps.emplace_back(new Capacitor(c.value, nd[0], nd[1])); break;
        case S21_COMP_I: ck.comps.emplace_back(new Isrc(c.value, nd[0], nd[1])); break;
        case S21_COMP_V: ck.comps.emplace_back(new Vsrc(c.value, c.acm, nd[0], nd[1], nd[2])); break;
        case S21_COMP_MOS1: {
            if (c.model < 0 || c.model >= net->n_mos1_models || c.inst < 0 || c.inst >= net->n_mos1_insts)
                sperror("Model not defined");
            Mos1Model m = mos1_resolve(net->mos1_models[c.model]);
            Mos1Inst ip = mos1_inst_resolve(net->mos1_insts[c.inst]);
            Mos1Intp in = mos1_derive(m, ip, opts);
            VarIdx pr[6] = {nd[0], nd[1], nd[2], nd[3], nd[4], nd[5]};
            ck.comps.emplace_back(new Mos1(m, in, pr));
            break;
        }
        case S21_COMP_MOS0: {
            VarIdx pr[4] = {nd[0], nd[1], nd[2], nd[3]};
            ck.comps.emplace_back(new Mos0(c.model, pr));
            break;
        }
        case S21_COMP_DIODE: {
            if (c.model < 0 || c.model >= net->n_diode_models || c.inst < 0 || c.inst >= net->n_diode_insts)
                sperror("Parameters not defined");
            DiodeModel m = diode_model_from(net->diode_models[c.model]);
            DiodeIntp ip = diode_derive(m, net->diode_insts[c.inst], opts);
            ck.comps.emplace_back(new Diode(m, ip, nd[0], nd[1], nd[2]));
            break;
        }
        case S21_COMP_BSIM4: {
            if (c.model < 0 || c.model >= net->n_bsim4_models || c.inst < 0 || c.inst >= net->n_bsim4_insts)
                sperror("Parameters not defined");
            s21::b4::B4Par P;
            try {
                P = s21::b4::derive_params(net->bsim4_models[c.model], net->bsim4_insts[c.inst]);
            } catch (const std::domain_error& e) {
                sperror(e.what());
            }
            VarIdx pr[4] = {nd[0], nd[1], nd[2], nd[3]};
            ck.comps.emplace_back(new Bsim4(P, pr));
            break;
        }
        default: sperror("Invalid Component");
        }
    }
}

struct Stats {
    long long iters = 0, solves = 0, factorizations = 0;
};

struct RealSolver {
    Circuit& ck;
    Options opts;
    Matrix<double> mat;
    std::vector<double> x, rhs;
    Stats* stats;
    RealSolver(Circuit& c, const Options& o, Stats* s) : ck(c), opts(o), stats(s) {
        for (auto& comp : ck.comps) comp->create_matrix_elems(mat);
        x.assign(ck.n_vars, 0.0);
    }
    void update(const AnalysisInfo& an) {  // analysis.rs:153-168
        for (auto& comp : ck.comps) {
            Stamps<double> s = comp->load(x, an, opts);
            for (auto& u : s.g)
                if (u.first != NONE) mat.update(u.first, u.second);
            for (auto& u : s.b)
                if (u.first >= 0) rhs[(size_t)u.first] += u.second;
        }
    }
    bool converged(const std::vector<double>& dx, const std::vector<double>& res) const {  // :331-345
        for (double e : dx)
            if (absv(e) > opts.reltol) return false;
        for (double e : res)
            if (absv(e) > opts.iabstol) return false;
        return true;
    }
    void solve(const AnalysisInfo& an) {  // analysis.rs:169-210
        std::vector<double> dx(x.size(), 0.0);
        for (int k = 0; k < 100; k++) {
            mat.reset();
            rhs.assign(x.size(), 0.0);
            update(an);
            if (stats) stats->iters++;
            if (mat.num_rows() != x.size() || mat.num_cols() != x.size()) sperror("Invalid Dimensions");
            std::vector<double> res = mat.res(x, rhs);
            if (converged(dx, res)) {
                for (auto& c : ck.comps) c->commit();
                return;
            }
            long f0 = mat.factorizations;
            dx = mat.solve(res);
            if (stats) {
                stats->solves++;
                stats->factorizations += mat.factorizations - f0;
            }
            double max_step = 1000e-3;
            double max_abs = 0.0;
            for (double v : dx) max_abs = (std::fabs(v) > max_abs) ? std::fabs(v) : max_abs;
            if (max_abs > max_step)
                for (size_t r = 0; r < dx.size(); r++) dx[r] = dx[r] * max_step / max_abs;
            for (size_t r = 0; r < x.size(); r++) x[r] += dx[r];
        }
        sperror("Convergence Failed");
    }
};

struct ComplexSolver {  // analysis.rs:215-304
    Circuit& ck;
    Options opts;
    Matrix<Cx> mat;
    std::vector<Cx> x, rhs;
    Stats* stats;
    ComplexSolver(Circuit& c, const Options& o, Stats* s) : ck(c), opts(o), stats(s) {
        for (auto& comp : ck.comps) comp->create_matrix_elems(mat);
        x.assign(ck.n_vars, Cx{0.0, 0.0});
    }
    void update(const AnalysisInfo& an) {
        for (auto& comp : ck.comps) {
            Stamps<Cx> s = comp->load_ac(x, an, opts);
            for (auto& u : s.g)
                if (u.first != NONE) mat.update(u.first, u.second);
            for (auto& u : s.b)
                if (u.first >= 0) rhs[(size_t)u.first] = rhs[(size_t)u.first] + u.second;
        }
    }
    void solve(const AnalysisInfo& an) {
        std::vector<Cx> dx(x.size(), Cx{0.0, 0.0});
        for (int k = 0; k < 20; k++) {
            mat.reset();
            rhs.assign(x.size(), Cx{0.0, 0.0});
            update(an);
            if (stats) stats->iters++;
            if (mat.num_rows() != x.size() || mat.num_cols() != x.size()) sperror("Invalid Dimensions");
            std::vector<Cx> res = mat.res(x, rhs);
            bool ok = true;
            for (auto& v : dx) ok = ok && (absv(v) < 1e-3);
            bool ok2 = true;
            for (auto& v : res) ok2 = ok2 && (absv(v) < 1e-9);
            if (ok && ok2) {
                for (auto& c : ck.comps) c->commit();
                return;
            }
            long f0 = mat.factorizations;
            dx = mat.solve(res);
            if (stats) {
                stats->solves++;
                stats->factorizations += mat.factorizations - f0;
            }
            double max_step = 1.0;
            double max_abs = 0.0;
            for (auto& v : dx) max_abs = (absv(v) > max_abs) ? absv(v) : max_abs;
            if (max_abs > max_step)
                for (size_t r = 0; r < dx.size(); r++) dx[r] = cx_scale_div(cx_scale_mul(dx[r], max_step), max_abs);
            for (size_t r = 0; r < x.size(); r++) x[r] = x[r] + dx[r];
        }
        sperror("Convergence Failed");
    }
};

static Options to_opts(const s21_options* o) {
    Options r;
    if (o) {
        r.temp = o->temp;
        r.tnom = o->tnom;
        r.gmin = o->gmin;
        r.iabstol = o->iabstol;
        r.reltol = o->reltol;
    }
    return r;
}

}  // namespace orc

/* ============================================================================================
 * C API of the oracle
 * ============================================================================================ */
using namespace orc;

static void set_err(char* err, size_t len, const std::string& s) {
    if (err && len) {
        size_t n = std::min(len - 1, s.size());
        memcpy(err, s.data(), n);
        err[n] = 0;
    }
}
static int code_of(const std::string& d) {
    if (d == "Convergence Failed") return S21_CONVERGENCE;
    if (d == "Singular Matrix") return S21_SINGULAR;
    if (d == "Pivot Search Fail" || d.rfind("Assert Neq", 0) == 0) return S21_PIVOT;
    return S21_ERR;
}

extern "C" {

void oracle_plant(int k) { g_plant = k; }
void oracle_pivot_stats(long long* factorisations, long long* changes, int reset) {
    if (factorisations) *factorisations = g_factorisations.load();
    if (changes) *changes = g_pivot_changes.load();
    if (reset) {
        g_factorisations = 0;
        g_pivot_changes = 0;
    }
}

/* DC operating point (analysis.rs:383-388).  x_out[n_vars]; iters_out: Newton iterations. */
int oracle_dcop(const s21_netlist* net, const s21_options* o, double* x_out, int64_t* iters_out, int64_t* facts_out,
                char* err, size_t errlen) {
    try {
        Options opts = to_opts(o);
        Circuit ck;
        build_components(net, opts, ck);
        Stats st;
        RealSolver s(ck, opts, &st);
        AnalysisInfo an;
        an.kind = AN_OP;
        s.solve(an);
        for (size_t k = 0; k < s.x.size(); k++) x_out[k] = s.x[k];
        if (iters_out) *iters_out = st.iters;
        if (facts_out) *facts_out = st.factorizations;
        return S21_OK;
    } catch (SpError& e) {
        set_err(err, errlen, e.desc);
        return code_of(e.desc);
    }
}

/* One dcop per instance of a batch: instance i is the netlist with every override's vals[i] applied, a
 * varied Mos1/Diode getting its own copy of its model or instance parameter set -- the per-instance
 * circuit a batch group stands for.  Instances i0 .. i0+n_inst-1 run on n_threads host threads.
 * x_out[n_inst * n_vars], iters_out and status_out[n_inst]. */
int oracle_dcop_batch(const s21_netlist* net, const s21_options* o, int32_t n_over, const s21_override* over,
                      int32_t i0, int32_t n_inst, int32_t n_threads, double* x_out, int64_t* iters_out,
                      int32_t* status_out) {
    if (!net || n_inst < 0 || n_threads < 1) return S21_INVALID;
    const size_t nv = (size_t)net->n_vars;
    std::atomic<int32_t> next(0);
    auto work = [&]() {
        std::vector<s21_comp> comps;
        std::vector<s21_params> mm, mi, dm, di, bm, bi;
        for (;;) {
            const int32_t j = next.fetch_add(1);
            if (j >= n_inst) return;
            const int32_t i = i0 + j;
            comps.assign(net->comps, net->comps + net->n_comps);
            mm.assign(net->mos1_models, net->mos1_models + net->n_mos1_models);
            mi.assign(net->mos1_insts, net->mos1_insts + net->n_mos1_insts);
            dm.assign(net->diode_models, net->diode_models + net->n_diode_models);
            di.assign(net->diode_insts, net->diode_insts + net->n_diode_insts);
            bm.assign(net->bsim4_models, net->bsim4_models + net->n_bsim4_models);
            bi.assign(net->bsim4_insts, net->bsim4_insts + net->n_bsim4_insts);
            for (int32_t k = 0; k < n_over; k++) {
                const s21_override& ov = over[k];
                s21_comp& c = comps[ov.comp];
                if (ov.target == S21_OVR_VALUE) {
                    c.value = ov.vals[i];
                    continue;
                }
                const bool model = ov.target == S21_OVR_MODEL;
                std::vector<s21_params>& lst = c.type == S21_COMP_MOS1    ? (model ? mm : mi)
                                               : c.type == S21_COMP_BSIM4 ? (model ? bm : bi)
                                                                          : (model ? dm : di);
                int32_t& idx = ov.target == S21_OVR_MODEL ? c.model : c.inst;
                s21_params p = lst[idx];
                p.v[ov.field] = ov.vals[i];
                p.set |= 1ull << ov.field;
                lst.push_back(p);
                idx = (int32_t)lst.size() - 1;
            }
            s21_netlist n2 = *net;
            n2.comps = comps.data();
            n2.mos1_models = mm.data();
            n2.n_mos1_models = (int32_t)mm.size();
            n2.mos1_insts = mi.data();
            n2.n_mos1_insts = (int32_t)mi.size();
            n2.diode_models = dm.data();
            n2.n_diode_models = (int32_t)dm.size();
            n2.diode_insts = di.data();
            n2.n_diode_insts = (int32_t)di.size();
            n2.bsim4_models = bm.data();
            n2.n_bsim4_models = (int32_t)bm.size();
            n2.bsim4_insts = bi.data();
            n2.n_bsim4_insts = (int32_t)bi.size();
            int64_t it = 0;
            const int st = oracle_dcop(&n2, o, x_out + (size_t)j * nv, &it, nullptr, nullptr, 0);
            if (iters_out) iters_out[j] = it;
            if (status_out) status_out[j] = st;
        }
    };
    std::vector<std::thread> th;
    for (int32_t t = 1; t < n_threads; t++) th.emplace_back(work);
    work();
    for (std::thread& t : th) t.join();
    return S21_OK;
}

/* Transient (analysis.rs:489-574).  Writes up to max_pts points: t_out[p], x_out[p * n_total + k];
 * n_total = n_vars + 2 * n_ic.  Returns the number of points, or -status on error. */
int64_t oracle_tran(const s21_netlist* net, const s21_options* o, const s21_tran_options* tr, int64_t max_pts,
                    double* t_out, double* x_out, int32_t* n_total_out, int64_t* iters_out, char* err, size_t errlen) {
    try {
        Options opts = to_opts(o);
        Circuit ck;
        build_components(net, opts, ck);
        Stats st;
        RealSolver s(ck, opts, &st);
        std::vector<size_t> vic, ric;
        for (int32_t k = 0; k < tr->n_ic; k++) {  // Tran::ic, analysis.rs:510-525
            VarIdx fnode = (VarIdx)ck.n_vars++;
            VarIdx ivar = (VarIdx)ck.n_vars++;
            ck.kinds.push_back(S21_VAR_V);
            ck.kinds.push_back(S21_VAR_I);
            s.x.push_back(0.0);
            s.x.push_back(0.0);
            std::unique_ptr<Component> r(new Resistor(1.0, fnode, tr->ic_var[k]));
            r->create_matrix_elems(s.mat);
            ck.comps.push_back(std::move(r));
            ric.push_back(ck.comps.size() - 1);
            std::unique_ptr<Component> v(new Vsrc(tr->ic_val[k], 0.0, fnode, -1, ivar));
            v->create_matrix_elems(s.mat);
            ck.comps.push_back(std::move(v));
            vic.push_back(ck.comps.size() - 1);
        }
        size_t nt = ck.n_vars;
        if (n_total_out) *n_total_out = (int32_t)nt;
        int64_t npts = 0;
        TranState ts;
        AnalysisInfo an;
        an.kind = AN_OP;
        s.solve(an);
        auto push = [&](double t) {
            if (npts < max_pts) {
                t_out[npts] = t;
                for (size_t k = 0; k < nt; k++) x_out[npts * (int64_t)nt + (int64_t)k] = s.x[k];
            }
            npts++;
        };
        push(ts.t);
        for (size_t c : vic) ck.comps[c]->update(0.0);
        for (size_t c : ric) ck.comps[c]->update(1e-9);
        ts.t = tr->tstep;
        ts.dt = tr->tstep;
        an.kind = AN_TRAN;
        an.tran = &ts;
        size_t tpoint = 0;
        while (ts.t < tr->tstop && tpoint < (size_t)1e9) {
            s.solve(an);
            push(ts.t);
            tpoint++;
            ts.t += tr->tstep;
        }
        if (iters_out) *iters_out = st.iters;
        return npts;
    } catch (SpError& e) {
        set_err(err, errlen, e.desc);
        return -(int64_t)code_of(e.desc);
    }
}

/* AC sweep (analysis.rs:761-832).  f_out[p]; x_out[(p * n_vars + k) * 2 + {0,1}].  Returns the number
 * of frequencies, or -status on error. */
int64_t oracle_ac(const s21_netlist* net, const s21_options* o, const s21_ac_options* ac, int64_t max_pts,
                  double* f_out, double* x_out, int64_t* iters_out, char* err, size_t errlen) {
    try {
        Options opts = to_opts(o);
        Circuit ck;
        build_components(net, opts, ck);
        Stats st;
        {
            RealSolver s(ck, opts, &st);
            AnalysisInfo an;
            an.kind = AN_OP;
            s.solve(an);
        }
        ComplexSolver cs(ck, opts, &st);
        double f = (double)ac->fstart;
        double fstop = (double)ac->fstop;
        double fstep = std::pow(10.0, std::log10(fstop / f) / (double)ac->npts);
        int64_t npts = 0;
        size_t n = ck.n_vars;
        while (f <= fstop) {
            AnalysisInfo an;
            an.kind = AN_AC;
            an.omega = 2.0 * M_PI * f;
            cs.solve(an);
            if (npts < max_pts) {
                f_out[npts] = f;
                for (size_t k = 0; k < n; k++) {
                    x_out[(npts * (int64_t)n + (int64_t)k) * 2 + 0] = cs.x[k].re;
                    x_out[(npts * (int64_t)n + (int64_t)k) * 2 + 1] = cs.x[k].im;
                }
            }
            npts++;
            if (f == fstop) break;
            f = std::fmin(f * fstep, fstop);
        }
        if (iters_out) *iters_out = st.iters;
        return npts;
    } catch (SpError& e) {
        set_err(err, errlen, e.desc);
        return -(int64_t)code_of(e.desc);
    }
}

/* One OP assembly at voltages x (analysis.rs:177-182): element values in creation order with their
 * positions, plus the RHS.  Returns the element count, or -status. */
int64_t oracle_assemble_op(const s21_netlist* net, const s21_options* o, const double* x, int64_t max_elems,
                           int32_t* rows, int32_t* cols, double* vals, double* rhs_out, char* err, size_t errlen) {
    try {
        Options opts = to_opts(o);
        Circuit ck;
        build_components(net, opts, ck);
        RealSolver s(ck, opts, nullptr);
        for (size_t k = 0; k < s.x.size(); k++) s.x[k] = x[k];
        s.mat.reset();
        s.rhs.assign(s.x.size(), 0.0);
        AnalysisInfo an;
        an.kind = AN_OP;
        s.update(an);
        int64_t ne = (int64_t)s.mat.el.size();
        for (int64_t k = 0; k < ne && k < max_elems; k++) {
            rows[k] = (int32_t)s.mat.el[k].row;
            cols[k] = (int32_t)s.mat.el[k].col;
            vals[k] = s.mat.el[k].val;
        }
        for (size_t k = 0; k < s.x.size(); k++) rhs_out[k] = s.rhs[k];
        return ne;
    } catch (SpError& e) {
        set_err(err, errlen, e.desc);
        return -(int64_t)code_of(e.desc);
    }
}

/* Pivot order of the first factorisation of the OP Jacobian at voltages x: AxisMapping.i2e of rows and
 * columns after lu_factorize (sparse21/mod.rs:647-674), plus the nnz of L+U including fill. */
int oracle_first_pivots(const s21_netlist* net, const s21_options* o, const double* x, int32_t* row_i2e,
                        int32_t* col_i2e, int64_t* nnz_lu, char* err, size_t errlen) {
    try {
        Options opts = to_opts(o);
        Circuit ck;
        build_components(net, opts, ck);
        RealSolver s(ck, opts, nullptr);
        for (size_t k = 0; k < s.x.size(); k++) s.x[k] = x[k];
        s.mat.reset();
        s.rhs.assign(s.x.size(), 0.0);
        AnalysisInfo an;
        an.kind = AN_OP;
        s.update(an);
        s.mat.lu_factorize();
        for (size_t k = 0; k < s.x.size(); k++) {
            row_i2e[k] = (int32_t)s.mat.axes[ROWS].i2e[k];
            col_i2e[k] = (int32_t)s.mat.axes[COLS].i2e[k];
        }
        if (nnz_lu) *nnz_lu = (int64_t)s.mat.el.size();
        return S21_OK;
    } catch (SpError& e) {
        set_err(err, errlen, e.desc);
        return code_of(e.desc);
    }
}

/* Derived Mos1 parameters (mos.rs:140-476) in a fixed order, for parity of the host derivation:
 * [vt0_t, kp_t, phi_t, beta, cox, cgs_ov, cgd_ov, cgb_ov, leff, vtherm, grd, grs,
 *  d.isat, d.depletion_threshold, d.bulkpot_t, d.czb, d.czbsw, d.f2, d.f3, d.f4,
 *  s.isat, s.depletion_threshold, s.bulkpot_t, s.czb, s.czbsw, s.f2, s.f3, s.f4,
 *  model p, gamma, lambda, mj, mjsw]  (33 values) */
int oracle_mos1_derive(const s21_params* model, const s21_params* inst, const s21_options* o, double* out, char* err,
                       size_t errlen) {
    try {
        Options opts = to_opts(o);
        Mos1Model m = mos1_resolve(*model);
        Mos1Intp ip = mos1_derive(m, mos1_inst_resolve(*inst), opts);
        double v[] = {ip.vt0_t, ip.kp_t, ip.phi_t, ip.beta, ip.cox, ip.cgs_ov, ip.cgd_ov, ip.cgb_ov, ip.leff, ip.vtherm,
                      ip.grd, ip.grs, ip.drain_junc.isat, ip.drain_junc.depletion_threshold, ip.drain_junc.bulkpot_t,
                      ip.drain_junc.czb, ip.drain_junc.czbsw, ip.drain_junc.f2, ip.drain_junc.f3, ip.drain_junc.f4,
                      ip.source_junc.isat, ip.source_junc.depletion_threshold, ip.source_junc.bulkpot_t,
                      ip.source_junc.czb, ip.source_junc.czbsw, ip.source_junc.f2, ip.source_junc.f3, ip.source_junc.f4,
                      m.p, m.gamma, m.lambda, m.mj, m.mjsw};
        memcpy(out, v, sizeof v);
        return S21_OK;
    } catch (SpError& e) {
        set_err(err, errlen, e.desc);
        return code_of(e.desc);
    }
}

/* Derived Diode parameters (diode.rs:146-212):
 * [vt, vte, vcrit, isat, gspr, cz, cz2, dep_threshold, f1, f2, f3, bv]  (12 values) */
int oracle_diode_derive(const s21_params* model, const s21_params* inst, const s21_options* o, double* out) {
    Options opts = to_opts(o);
    DiodeModel m = diode_model_from(*model);
    DiodeIntp ip = diode_derive(m, *inst, opts);
    double v[] = {ip.vt, ip.vte, ip.vcrit, ip.isat, ip.gspr, ip.cz, ip.cz2, ip.dep_threshold, ip.f1, ip.f2, ip.f3, ip.bv};
    memcpy(out, v, sizeof v);
    return S21_OK;
}

/* ---- bare sparse-matrix access for the reference's sparse21 known-answer tests ---- */
struct OMat {
    bool cplx;
    Matrix<double> r;
    Matrix<Cx> c;
};
void* oracle_mat_new(int cplx) {
    OMat* m = new OMat();
    m->cplx = cplx != 0;
    return m;
}
void oracle_mat_free(void* m) { delete (OMat*)m; }
void oracle_mat_add(void* mp, int64_t row, int64_t col, double re, double im) {
    OMat* m = (OMat*)mp;
    if (m->cplx)
        m->c.add_element((size_t)row, (size_t)col, Cx{re, im});
    else
        m->r.add_element((size_t)row, (size_t)col, re);
}
int oracle_mat_factor(void* mp, char* err, size_t errlen) {
    OMat* m = (OMat*)mp;
    try {
        if (m->cplx)
            m->c.lu_factorize();
        else
            m->r.lu_factorize();
        return S21_OK;
    } catch (SpError& e) {
        set_err(err, errlen, e.desc);
        return code_of(e.desc);
    }
}
/* rhs/out are interleaved re/im for complex matrices. */
int oracle_mat_solve(void* mp, int64_t n, const double* rhs, double* out, char* err, size_t errlen) {
    OMat* m = (OMat*)mp;
    try {
        if (m->cplx) {
            std::vector<Cx> b(n);
            for (int64_t k = 0; k < n; k++) b[k] = Cx{rhs[2 * k], rhs[2 * k + 1]};
            std::vector<Cx> x = m->c.solve(b);
            for (int64_t k = 0; k < n; k++) {
                out[2 * k] = x[k].re;
                out[2 * k + 1] = x[k].im;
            }
        } else {
            std::vector<double> b(rhs, rhs + n);
            std::vector<double> x = m->r.solve(b);
            for (int64_t k = 0; k < n; k++) out[k] = x[k];
        }
        return S21_OK;
    } catch (SpError& e) {
        set_err(err, errlen, e.desc);
        return code_of(e.desc);
    }
}
/* Value at internal position (row, col): returns 1 and writes it when the element exists. */
int oracle_mat_get(void* mp, int64_t row, int64_t col, double* re, double* im) {
    OMat* m = (OMat*)mp;
    if (m->cplx) {
        Eidx e = m->c.get_elem((size_t)row, (size_t)col);
        if (e == NONE) return 0;
        *re = m->c.el[e].val.re;
        *im = m->c.el[e].val.im;
        return 1;
    }
    Eidx e = m->r.get_elem((size_t)row, (size_t)col);
    if (e == NONE) return 0;
    *re = m->r.el[e].val;
    *im = 0.0;
    return 1;
}
/* Row swap with mapping bookkeeping, as the reference's test_row_mappings exercises
 * (sparse21/mod.rs:1356-1375): sets state FACTORING and sets up the row mapping once. */
void oracle_mat_swap_rows(void* mp, int64_t x, int64_t y) {
    OMat* m = (OMat*)mp;
    auto& r = m->r;
    if (r.state != Matrix<double>::FACTORING) {
        r.state = Matrix<double>::FACTORING;
        r.axes[ROWS].setup_factoring();
    }
    r.swap(ROWS, (size_t)x, (size_t)y);
}
void oracle_mat_row_mapping(void* mp, int64_t n, int64_t* e2i, int64_t* i2e) {
    OMat* m = (OMat*)mp;
    for (int64_t k = 0; k < n; k++) {
        e2i[k] = (int64_t)m->r.axes[ROWS].e2i[k];
        i2e[k] = (int64_t)m->r.axes[ROWS].i2e[k];
    }
}
/* Consistency of the orthogonal lists (the reference's Matrix::checkups, sparse21/mod.rs:1088-1160):
 * every element is reachable once along its row and its column, lists are sorted, the diagonal
 * array points at diagonal elements.  Returns 0 when consistent. */
int oracle_mat_check(void* mp) {
    OMat* om = (OMat*)mp;
    auto& m = om->r;
    size_t ne = m.el.size();
    std::vector<int> seen_r(ne, 0), seen_c(ne, 0);
    for (size_t r = 0; r < m.num_rows(); r++) {
        Eidx e = m.hdr(ROWS, r);
        long last = -1;
        while (e != NONE) {
            if (m.el[e].row != r) return 1;
            if ((long)m.el[e].col <= last) return 2;
            last = (long)m.el[e].col;
            seen_r[e]++;
            e = m.el[e].next_in_row;
        }
    }
    for (size_t c = 0; c < m.num_cols(); c++) {
        Eidx e = m.hdr(COLS, c);
        long last = -1;
        while (e != NONE) {
            if (m.el[e].col != c) return 3;
            if ((long)m.el[e].row <= last) return 4;
            last = (long)m.el[e].row;
            seen_c[e]++;
            e = m.el[e].next_in_col;
        }
    }
    for (size_t k = 0; k < ne; k++)
        if (seen_r[k] != 1 || seen_c[k] != 1) return 5;
    for (size_t k = 0; k < m.diag.size(); k++) {
        Eidx d = m.diag[k];
        if (d != NONE && (m.el[d].row != k || m.el[d].col != k)) return 6;
        if (d == NONE) {
            for (Eidx e = m.hdr(ROWS, k); e != NONE; e = m.el[e].next_in_row)
                if (m.el[e].col == k) return 7;
        }
    }
    return 0;
}

}  // extern "C"
