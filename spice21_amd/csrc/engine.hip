// Batched Newton engine for MI355X (gfx950).
//
// One wavefront (a 64-lane workgroup) owns one circuit instance for the whole analysis: operating point
// and, for a transient, every time step.  All per-instance Newton state lives in LDS:
//     x[n] | rhs[n] | c[n] | contrib[n_contrib] | a[nnz] (| lu32[nnz] in mixed precision)
// Each Newton iteration (Solver::solve, analysis.rs:169-210) runs these phases, separated by barriers:
//   1. device evaluation, one lane per device, writing every live stamp into its own contribution slot
//      (no atomics; Mos1 parameters of the lane's devices stay in registers for the whole launch);
//   2. ordered gather: each matrix / RHS slot sums its contributions in component-then-stamp order,
//      which is the order Solver::update adds them (analysis.rs:153-168);
//   3. KCL residual res = rhs - A x in FP64 over the pivoted rows (sparse21/mod.rs:298-327) and the
//      convergence test on the previous update and the residual (analysis.rs:331-345);
//   4. numeric LU on the pivot order and fill pattern frozen by the host symbolic phase, level-scheduled
//      by rows (IKJ order, which applies every update to an entry in the same pivot order as the
//      reference's right-looking elimination, sparse21/mod.rs:865-919);
//   5. level-scheduled forward and backward sweeps (sparse21/mod.rs:946-979);
//   6. the max-|dx| <= 1 clamp and the update x += dx (analysis.rs:196-207).
// In mixed precision phases 4-5 run in FP32 while 1-3 and 6 stay FP64, so Newton iterates as
// iterative refinement.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "internal.h"

namespace s21 {

// ============================================================================================
// Device-side descriptors
// ============================================================================================
struct TopoGPU {
    int n, nnz, n_contrib;
    int n_m1, n_m0, n_d, n_lin;
    int n_lvl, n_blvl;
    int sym_status;
    const int *m1_nodes, *m1_base;
    const uint32_t* m1_mask;
    const int* m1_caps;
    const int *m0_nodes, *m0_base;
    const uint32_t* m0_mask;
    const double* m0_p;
    const int *d_nodes, *d_base;
    const uint32_t* d_mask;
    const int *lin_type, *lin_nodes, *lin_base;
    const uint32_t* lin_mask;
    const int *g_ptr, *g_idx;
    const int *res_ptr, *res_pos, *res_xcol;
    const int *row_i2e, *col_e2i;
    const int *rp, *ci, *diag;
    const int *upd_ptr, *upd_src, *upd_dst;
    const int *lvl_ptr, *lvl_rows, *blvl_ptr, *blvl_rows;
};

struct GroupGPU {
    TopoGPU t;
    int n_inst, blk0;
    const double* m1_core;  // [11][n_inst][n_m1]
    const double* m1_caps;  // [4][n_inst][n_m1]
    double* m1_state;       // [14][n_inst][n_m1]: committed op (7) then guess (7)
    const double* d_par;    // [14][n_inst][n_d]
    double* d_state;        // [3][n_inst][n_d]: guess vd, guess charge, committed charge
    const double* lin_val;  // [n_inst][n_lin] operating-point values
    const double* lin_val_tran;
    double* c_q;            // [n_inst][n_lin] committed capacitor charge
    double* x;              // [n_inst][n]
    int* status;            // [n_inst]
    long long* iters;       // [n_inst]
    double* wave;           // [n_inst][npts][n] (transient) or nullptr
    double* dbg;            // assembly dump [nnz + n] (assemble mode)
};

enum { MODE_OP = 0, MODE_TRAN = 1, MODE_ASSEMBLE = 2 };
struct Args {
    int mode;
    int max_iter;
    double gmin, iabstol, reltol;
    double tstep, tstop;
    long long npts;
    int dbg_inst;
};

constexpr int WAVE = 64;
constexpr int M1_CORE_F = 11, M1_CAPS_F = 4, M1_STATE_F = 7, D_PAR_F = 14, D_STATE_F = 3;

__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double w = __shfl_xor(v, o, WAVE);
        v = (w > v) ? w : v;  // the reference's fold drops NaN (analysis.rs:198)
    }
    return v;
}
__device__ inline bool wave_any(bool p) { return __ballot(p) != 0ull; }

template <int NE>
__device__ inline void write_entries(const double* v, uint32_t mask, int base, double* contrib) {
#pragma unroll
    for (int e = 0; e < NE; e++)
        if ((mask >> e) & 1u) contrib[base + __popc(mask & ((1u << e) - 1u))] = v[e];
}

// Per-lane register copy of the lane's Mos1 devices (DPL > 0).
template <int DPL>
struct Mos1Regs {
    Mos1Core P[DPL];
    int nd[DPL][4];
    uint32_t mask[DPL];
    int base[DPL];
    int caps[DPL];
};

struct Ctx {
    const GroupGPU* G;
    const TopoGPU* T;
    int inst, lane;
    double *x, *rhs, *c, *contrib, *a;
    float* lu32;
};

__device__ inline double xv(const double* x, int v) { return v < 0 ? 0.0 : x[v]; }

__device__ inline Mos1Core load_mos1_core(const GroupGPU& G, int n_m1, int inst, int d) {
    const double* p = G.m1_core;
    const size_t s = (size_t)G.n_inst * n_m1, o = (size_t)inst * n_m1 + d;
    Mos1Core P;
    P.p = p[0 * s + o];
    P.vt0_t = p[1 * s + o];
    P.beta = p[2 * s + o];
    P.phi_t = p[3 * s + o];
    P.gamma = p[4 * s + o];
    P.lambda = p[5 * s + o];
    P.vtherm = p[6 * s + o];
    P.isat_d = p[7 * s + o];
    P.isat_s = p[8 * s + o];
    P.grd = p[9 * s + o];
    P.grs = p[10 * s + o];
    return P;
}

// Mos1 with gate capacitance: parameters and state come from global memory.
template <bool TRAN>
__device__ inline void eval_mos1_caps(const Ctx& C, const Mos1Core& P, int d, double vd, double vg, double vs,
                                      double vb, double gmin, double dt, double* v) {
    const GroupGPU& G = *C.G;
    const int n_m1 = C.T->n_m1;
    const size_t s = (size_t)G.n_inst * n_m1, o = (size_t)C.inst * n_m1 + d;
    Mos1Caps K;
    K.cox = G.m1_caps[0 * s + o];
    K.cgs_ov = G.m1_caps[1 * s + o];
    K.cgd_ov = G.m1_caps[2 * s + o];
    K.cgb_ov = G.m1_caps[3 * s + o];
    Mos1State op, gs;
    double* st = G.m1_state;
    op.vgs = st[0 * s + o];
    op.vgd = st[1 * s + o];
    op.vgb = st[2 * s + o];
    op.cgs = st[3 * s + o];
    op.cgd = st[4 * s + o];
    op.cgb = st[5 * s + o];
    op.reversed = st[6 * s + o];
    mos1_load<TRAN>(P, &K, &op, vd, vg, vs, vb, gmin, dt, v, &gs);
    st[7 * s + o] = gs.vgs;
    st[8 * s + o] = gs.vgd;
    st[9 * s + o] = gs.vgb;
    st[10 * s + o] = gs.cgs;
    st[11 * s + o] = gs.cgd;
    st[12 * s + o] = gs.cgb;
    st[13 * s + o] = gs.reversed;
}

template <int DPL>
__device__ void eval_devices(const Ctx& C, const Mos1Regs<(DPL > 0 ? DPL : 1)>& R, bool tran, const Args& A) {
    const GroupGPU& G = *C.G;
    const TopoGPU& T = *C.T;
    const double* x = C.x;
    const double gmin = A.gmin, dt = A.tstep;
    // ---- Mos1 ----
    if (DPL > 0) {
#pragma unroll
        for (int j = 0; j < (DPL > 0 ? DPL : 1); j++) {
            const int d = C.lane + WAVE * j;
            if (d < T.n_m1) {
                const double vd = xv(x, R.nd[j][0]), vg = xv(x, R.nd[j][1]), vs = xv(x, R.nd[j][2]),
                             vb = xv(x, R.nd[j][3]);
                double v[MOS1_NE];
                if (R.caps[j]) {
                    if (tran)
                        eval_mos1_caps<true>(C, R.P[j], d, vd, vg, vs, vb, gmin, dt, v);
                    else
                        eval_mos1_caps<false>(C, R.P[j], d, vd, vg, vs, vb, gmin, dt, v);
                } else {
                    mos1_load<false>(R.P[j], nullptr, nullptr, vd, vg, vs, vb, gmin, dt, v, nullptr);
                }
                write_entries<MOS1_NE>(v, R.mask[j], R.base[j], C.contrib);
            }
        }
    } else {
        for (int d = C.lane; d < T.n_m1; d += WAVE) {
            const Mos1Core P = load_mos1_core(G, T.n_m1, C.inst, d);
            const int* nd = T.m1_nodes + 6 * d;
            const double vd = xv(x, nd[0]), vg = xv(x, nd[1]), vs = xv(x, nd[2]), vb = xv(x, nd[3]);
            double v[MOS1_NE];
            if (T.m1_caps[d]) {
                if (tran)
                    eval_mos1_caps<true>(C, P, d, vd, vg, vs, vb, gmin, dt, v);
                else
                    eval_mos1_caps<false>(C, P, d, vd, vg, vs, vb, gmin, dt, v);
            } else {
                mos1_load<false>(P, nullptr, nullptr, vd, vg, vs, vb, gmin, dt, v, nullptr);
            }
            write_entries<MOS1_NE>(v, T.m1_mask[d], T.m1_base[d], C.contrib);
        }
    }
    // ---- Mos0 ----
    for (int d = C.lane; d < T.n_m0; d += WAVE) {
        const int* nd = T.m0_nodes + 4 * d;
        double v[MOS0_NE];
        mos0_load(T.m0_p[d], xv(x, nd[0]), xv(x, nd[1]), xv(x, nd[2]), gmin, v);
        write_entries<MOS0_NE>(v, T.m0_mask[d], T.m0_base[d], C.contrib);
    }
    // ---- Diode ----
    for (int d = C.lane; d < T.n_d; d += WAVE) {
        const size_t s = (size_t)G.n_inst * T.n_d, o = (size_t)C.inst * T.n_d + d;
        const double* q = G.d_par;
        DiodeP P;
        P.isat = q[0 * s + o];
        P.vte = q[1 * s + o];
        P.vcrit = q[2 * s + o];
        P.gspr = q[3 * s + o];
        P.bv = q[4 * s + o];
        P.has_bv = q[5 * s + o];
        P.cz = q[6 * s + o];
        P.cz2 = q[7 * s + o];
        P.dep = q[8 * s + o];
        P.f1 = q[9 * s + o];
        P.f3 = q[10 * s + o];
        P.tt = q[11 * s + o];
        P.vj = q[12 * s + o];
        P.m = q[13 * s + o];
        double* st = G.d_state;
        const int* nd = T.d_nodes + 3 * d;
        double v[DIODE_NE];
        DiodeGuess gs;
        if (tran)
            diode_load<true>(P, xv(x, nd[2]), xv(x, nd[1]), gmin, dt, st[0 * s + o], st[2 * s + o], v, &gs);
        else
            diode_load<false>(P, xv(x, nd[2]), xv(x, nd[1]), gmin, dt, st[0 * s + o], st[2 * s + o], v, &gs);
        st[0 * s + o] = gs.vd;
        st[1 * s + o] = gs.charge;
        write_entries<DIODE_NE>(v, T.d_mask[d], T.d_base[d], C.contrib);
    }
    // ---- R, C, V, I ----
    for (int d = C.lane; d < T.n_lin; d += WAVE) {
        const int typ = T.lin_type[d];
        const int* nd = T.lin_nodes + 3 * d;
        const size_t o = (size_t)C.inst * T.n_lin + d;
        const double val = tran ? G.lin_val_tran[o] : G.lin_val[o];
        double v[6];
        if (typ == S21_COMP_R) {
            v[0] = val;
            v[1] = val;
            v[2] = -val;
            v[3] = -val;
            write_entries<4>(v, T.lin_mask[d], T.lin_base[d], C.contrib);
        } else if (typ == S21_COMP_C) {
            if (tran) {
                cap_load_tran(val, xv(x, nd[0]) - xv(x, nd[1]), G.c_q[o], dt, v);
            } else {
                for (int k = 0; k < 6; k++) v[k] = 0.0;  // no stamps in the operating point
            }
            write_entries<6>(v, T.lin_mask[d], T.lin_base[d], C.contrib);
        } else if (typ == S21_COMP_V) {
            v[0] = 1.0;
            v[1] = 1.0;
            v[2] = -1.0;
            v[3] = -1.0;
            v[4] = val;
            write_entries<5>(v, T.lin_mask[d], T.lin_base[d], C.contrib);
        } else {
            v[0] = val;
            v[1] = -val;
            write_entries<2>(v, T.lin_mask[d], T.lin_base[d], C.contrib);
        }
    }
}

// Ordered gather: slot s < nnz is LU position s, slot nnz + k is RHS entry k (external order).
__device__ inline void gather(const Ctx& C) {
    const TopoGPU& T = *C.T;
    const int ns = T.nnz + T.n;
    for (int s = C.lane; s < ns; s += WAVE) {
        double acc = 0.0;
        const int e = T.g_ptr[s + 1];
        for (int k = T.g_ptr[s]; k < e; k++) acc = acc + C.contrib[T.g_idx[k]];
        if (s < T.nnz)
            C.a[s] = acc;
        else
            C.rhs[s - T.nnz] = acc;
    }
}

// c[i] = rhs[row_i2e[i]] - (A x)[i] over internal row i; returns whether this lane saw |res| > iabstol.
__device__ inline bool residual(const Ctx& C, double iabstol) {
    const TopoGPU& T = *C.T;
    bool bad = false;
    for (int i = C.lane; i < T.n; i += WAVE) {
        double acc = 0.0;
        const int e = T.res_ptr[i + 1];
        for (int k = T.res_ptr[i]; k < e; k++) acc = acc + C.a[T.res_pos[k]] * C.x[T.res_xcol[k]];
        const double r = C.rhs[T.row_i2e[i]] - acc;
        C.c[i] = r;
        bad = bad || (fabs(r) > iabstol);
    }
    return bad;
}

// Level-scheduled IKJ factorisation.  Returns a nonzero status when a pivot other than the last is zero
// (the reference's assertion in row_col_elim, sparse21/mod.rs:872).
template <typename V>
__device__ inline int factor(const Ctx& C, V* lu) {
    const TopoGPU& T = *C.T;
    bool zero_pivot = false;
    for (int l = 0; l < T.n_lvl; l++) {
        const int e = T.lvl_ptr[l + 1];
        for (int r = T.lvl_ptr[l] + C.lane; r < e; r += WAVE) {
            const int i = T.lvl_rows[r];
            const int dg = T.diag[i];
            for (int q = T.rp[i]; q < dg; q++) {
                const V piv = lu[T.diag[T.ci[q]]];
                const V lv = lu[q] / piv;
                lu[q] = lv;
                const int ue = T.upd_ptr[q + 1];
                for (int u = T.upd_ptr[q]; u < ue; u++) {
                    const int dst = T.upd_dst[u];
                    const V prod = lu[T.upd_src[u]] * lv;
                    lu[dst] = lu[dst] - prod;
                }
            }
            if (i != T.n - 1 && lu[dg] == V(0)) zero_pivot = true;
        }
        __syncthreads();
    }
    return wave_any(zero_pivot) ? S21_PIVOT : S21_OK;
}

// Forward (unit L) and backward (U) sweeps over c in internal order.
template <typename V>
__device__ inline void sweeps(const Ctx& C, const V* lu) {
    const TopoGPU& T = *C.T;
    for (int l = 0; l < T.n_lvl; l++) {
        const int e = T.lvl_ptr[l + 1];
        for (int r = T.lvl_ptr[l] + C.lane; r < e; r += WAVE) {
            const int i = T.lvl_rows[r];
            V y = (V)C.c[i];
            const int dg = T.diag[i];
            for (int q = T.rp[i]; q < dg; q++) {
                const V prod = (V)C.c[T.ci[q]] * lu[q];
                y = y - prod;
            }
            C.c[i] = (double)y;
        }
        __syncthreads();
    }
    for (int l = 0; l < T.n_blvl; l++) {
        const int e = T.blvl_ptr[l + 1];
        for (int r = T.blvl_ptr[l] + C.lane; r < e; r += WAVE) {
            const int i = T.blvl_rows[r];
            V y = (V)C.c[i];
            const int dg = T.diag[i];
            const int re = T.rp[i + 1];
            for (int q = dg + 1; q < re; q++) {
                const V prod = (V)C.c[T.ci[q]] * lu[q];
                y = y - prod;
            }
            C.c[i] = (double)(y / lu[dg]);
        }
        __syncthreads();
    }
}

// Commit the accepted operating point (Component::commit) of every stateful device.
__device__ inline void commit(const Ctx& C) {
    const GroupGPU& G = *C.G;
    const TopoGPU& T = *C.T;
    for (int d = C.lane; d < T.n_lin; d += WAVE) {
        if (T.lin_type[d] != S21_COMP_C) continue;
        const int* nd = T.lin_nodes + 3 * d;
        const size_t o = (size_t)C.inst * T.n_lin + d;
        const double vd = xv(C.x, nd[0]) - xv(C.x, nd[1]);
        G.c_q[o] = G.lin_val[o] * vd;  // guess.q = c * vd at the accepted voltages
    }
    for (int d = C.lane; d < T.n_d; d += WAVE) {
        const size_t s = (size_t)G.n_inst * T.n_d, o = (size_t)C.inst * T.n_d + d;
        G.d_state[2 * s + o] = G.d_state[1 * s + o];
    }
    for (int d = C.lane; d < T.n_m1; d += WAVE) {
        if (!T.m1_caps[d]) continue;
        const size_t s = (size_t)G.n_inst * T.n_m1, o = (size_t)C.inst * T.n_m1 + d;
        for (int f = 0; f < M1_STATE_F; f++) G.m1_state[f * s + o] = G.m1_state[(M1_STATE_F + f) * s + o];
    }
}

// One Newton solve.  `iters` accumulates loads.  Returns an S21_* status.
template <int DPL, bool MIXED>
__device__ int newton(const Ctx& C, const Mos1Regs<(DPL > 0 ? DPL : 1)>& R, bool tran, const Args& A,
                      long long& iters) {
    const TopoGPU& T = *C.T;
    bool prev_ok = true;  // dx starts at zero
    for (int it = 0; it < A.max_iter; it++) {
        iters++;
        eval_devices<DPL>(C, R, tran, A);
        __syncthreads();
        gather(C);
        __syncthreads();
        const bool bad = wave_any(residual(C, A.iabstol));
        __syncthreads();
        if (prev_ok && !bad) {
            commit(C);
            __syncthreads();
            return S21_OK;
        }
        if (T.sym_status != S21_OK) return T.sym_status;
        int st;
        if (MIXED) {
            for (int p = C.lane; p < T.nnz; p += WAVE) C.lu32[p] = (float)C.a[p];
            __syncthreads();
            st = factor<float>(C, C.lu32);
            if (st == S21_OK) sweeps<float>(C, C.lu32);
        } else {
            st = factor<double>(C, C.a);
            if (st == S21_OK) sweeps<double>(C, C.a);
        }
        if (st != S21_OK) return st;
        // clamp and update (analysis.rs:196-207)
        double mx = 0.0;
        for (int k = C.lane; k < T.n; k += WAVE) {
            const double dx = fabs(C.c[T.col_e2i[k]]);
            mx = (dx > mx) ? dx : mx;
        }
        mx = wave_max(mx);
        bool ok = true;
        for (int k = C.lane; k < T.n; k += WAVE) {
            double dx = C.c[T.col_e2i[k]];
            if (mx > 1.0) dx = dx * 1.0 / mx;
            ok = ok && !(fabs(dx) > A.reltol);
            C.x[k] = C.x[k] + dx;
        }
        prev_ok = !wave_any(!ok);
        __syncthreads();
    }
    return S21_CONVERGENCE;
}

template <int DPL, bool MIXED>
__global__ __launch_bounds__(WAVE) void newton_kernel(const GroupGPU* __restrict__ groups, int n_groups, Args A) {
    int g = 0;
    const int b = blockIdx.x;
    while (g + 1 < n_groups && groups[g + 1].blk0 <= b) g++;
    const GroupGPU& G = groups[g];
    const TopoGPU& T = G.t;
    const int inst = (A.mode == MODE_ASSEMBLE) ? A.dbg_inst : b - G.blk0;
    if (inst < 0 || inst >= G.n_inst) return;
    extern __shared__ double smem[];
    Ctx C;
    C.G = &G;
    C.T = &T;
    C.inst = inst;
    C.lane = threadIdx.x;
    C.x = smem;
    C.rhs = C.x + T.n;
    C.c = C.rhs + T.n;
    C.contrib = C.c + T.n;
    C.a = C.contrib + T.n_contrib;
    C.lu32 = (float*)(C.a + T.nnz);

    constexpr int NR = DPL > 0 ? DPL : 1;
    Mos1Regs<NR> R;
#pragma unroll
    for (int j = 0; j < NR; j++) {
        const int d = C.lane + WAVE * j;
        if (DPL > 0 && d < T.n_m1) {
            R.P[j] = load_mos1_core(G, T.n_m1, inst, d);
            for (int k = 0; k < 4; k++) R.nd[j][k] = T.m1_nodes[6 * d + k];
            R.mask[j] = T.m1_mask[d];
            R.base[j] = T.m1_base[d];
            R.caps[j] = T.m1_caps[d];
        } else {
            R.caps[j] = 0;
        }
    }
    for (int k = C.lane; k < T.n; k += WAVE) C.x[k] = G.x[(size_t)inst * T.n + k];
    __syncthreads();

    if (A.mode == MODE_ASSEMBLE) {
        eval_devices<DPL>(C, R, false, A);
        __syncthreads();
        gather(C);
        __syncthreads();
        for (int s = C.lane; s < T.nnz; s += WAVE) G.dbg[s] = C.a[s];
        for (int k = C.lane; k < T.n; k += WAVE) G.dbg[T.nnz + k] = C.rhs[k];
        return;
    }

    long long iters = 0;
    int st = newton<DPL, MIXED>(C, R, false, A, iters);
    if (A.mode == MODE_TRAN && st == S21_OK) {
        double* w = G.wave + (size_t)inst * (size_t)A.npts * T.n;
        for (int k = C.lane; k < T.n; k += WAVE) w[k] = C.x[k];
        long long pt = 1;
        double t = A.tstep;  // analysis.rs:553-570
        while (t < A.tstop && pt < A.npts) {
            st = newton<DPL, MIXED>(C, R, true, A, iters);
            if (st != S21_OK) break;
            for (int k = C.lane; k < T.n; k += WAVE) w[(size_t)pt * T.n + k] = C.x[k];
            pt++;
            t += A.tstep;
        }
    }
    for (int k = C.lane; k < T.n; k += WAVE) G.x[(size_t)inst * T.n + k] = C.x[k];
    if (C.lane == 0) {
        G.status[inst] = st;
        G.iters[inst] = iters;
    }
}

// ============================================================================================
// Host runtime
// ============================================================================================
#define HIPCHECK(x)                                                                         \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) fail(S21_DEVICE, std::string("HIP: ") + hipGetErrorString(e_)); \
    } while (0)

struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
    DevMem() {}
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t b) {
        if (p) {
            (void)hipFree(p);
            p = nullptr;
        }
        bytes = b;
        if (b) HIPCHECK(hipMalloc(&p, b));
    }
};

struct Override {
    int comp, target, field;
    std::vector<double> vals;
};

struct GroupHost {
    // the group's netlist
    int n_vars = 0;
    std::vector<std::string> names;
    std::vector<int> kinds;
    std::vector<s21_comp> comps;
    std::vector<s21_params> mos1_models, mos1_insts, diode_models, diode_insts;
    int n_inst = 0;
    std::vector<Override> overs;
    // the last analysis
    Topology topo;
    std::unique_ptr<DevMem> d_topo, d_par, d_state, d_out;
    GroupGPU gg{};
    std::vector<double> x;
    std::vector<int> status;
    std::vector<long long> iters;
    int dpl = 0;
    size_t smem = 0;
};

}  // namespace s21

struct s21_batch {
    std::vector<s21::GroupHost> groups;
    s21_options opts;
    int precision = S21_FP64;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0, kernel_avg_ms = 0.0;
    int64_t launches = 0;
    int64_t npts = 0;
    std::vector<double> axis;
    std::unique_ptr<s21::DevMem> d_wave, d_groups;
    int last_mode = -1;
    ~s21_batch() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace s21 {

static void check_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) fail(S21_DEVICE, "no HIP device available");
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHECK(hipGetDeviceProperties(&prop, dev));
    if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0)
        fail(S21_DEVICE, std::string("engine is built for gfx950, device is ") + prop.gcnArchName);
}

// Flattened components of the group for one analysis, IC forcing appended for a transient
// (Tran::ic, analysis.rs:510-525).
static void build_comps(GroupHost& g, const s21_tran_options* tr, Topology& t) {
    t.n = g.n_vars;
    t.names = g.names;
    t.kinds = g.kinds;
    t.comps.clear();
    t.tran = tr != nullptr;
    for (const s21_comp& c : g.comps) {
        Comp k{};
        k.type = c.type;
        for (int j = 0; j < 6; j++) k.node[j] = c.node[j];
        k.model = c.model;
        k.inst = c.inst;
        k.value = c.value;
        k.value_tran = c.value;
        k.acm = c.acm;
        if (c.type == S21_COMP_MOS1) {
            if (c.model < 0 || c.model >= (int)g.mos1_models.size() || c.inst < 0 || c.inst >= (int)g.mos1_insts.size())
                fail(S21_ERR, "Parameters not defined");
        } else if (c.type == S21_COMP_DIODE) {
            if (c.model < 0 || c.model >= (int)g.diode_models.size() || c.inst < 0 ||
                c.inst >= (int)g.diode_insts.size())
                fail(S21_ERR, "Parameters not defined");
        }
        for (int j = 0; j < 6; j++)
            if (k.node[j] >= g.n_vars || k.node[j] < S21_GND) fail(S21_INVALID, "component node out of range");
        t.comps.push_back(k);
    }
    if (tr) {
        for (int k = 0; k < tr->n_ic; k++) {
            const int node = tr->ic_var[k];
            if (node < 0 || node >= g.n_vars) fail(S21_INVALID, "initial condition node out of range");
            std::string nm = tr->ic_name ? tr->ic_name[k] : g.names[node];
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

// Host half of the preparation: per-instance derived parameters and the frozen structure.
struct HostPrep {
    std::vector<double> m1_core, m1_caps, d_par, lin_val, lin_tran;
    int n_m1 = 0, n_d = 0, n_lin = 0;
};

static void host_prepare(const s21_options& o, GroupHost& g, const s21_tran_options* tr, HostPrep& hp) {
    Topology& t = g.topo;
    t = Topology();
    build_comps(g, tr, t);
    topo_elements(t);
    const int NI = g.n_inst;

    // Overrides by component.
    std::unordered_map<int, std::vector<const Override*>> ov;
    for (const Override& x : g.overs) ov[x.comp].push_back(&x);

    // Device lists in component order.
    std::vector<int> m1c, dc, linc;
    for (size_t i = 0; i < t.comps.size(); i++) {
        int ty = t.comps[i].type;
        if (ty == S21_COMP_MOS1)
            m1c.push_back((int)i);
        else if (ty == S21_COMP_DIODE)
            dc.push_back((int)i);
        else if (ty != S21_COMP_MOS0)
            linc.push_back((int)i);
    }
    const int n_m1 = (int)m1c.size(), n_d = (int)dc.size(), n_lin = (int)linc.size();
    hp.n_m1 = n_m1;
    hp.n_d = n_d;
    hp.n_lin = n_lin;
    std::vector<double>& m1_core = hp.m1_core;
    std::vector<double>& m1_caps = hp.m1_caps;
    m1_core.assign((size_t)M1_CORE_F * NI * n_m1, 0.0);
    m1_caps.assign((size_t)M1_CAPS_F * NI * n_m1, 0.0);
    std::vector<uint8_t> has_caps(n_m1, 0);
    HostInstance h0;
    h0.m1.resize(n_m1);
    for (int d = 0; d < n_m1; d++) {
        const Comp& c = t.comps[m1c[d]];
        auto it = ov.find(m1c[d]);
        auto put = [&](int inst, const Mos1Derived& D) {
            const size_t s = (size_t)NI * n_m1, off = (size_t)inst * n_m1 + d;
            const double core[M1_CORE_F] = {D.core.p,      D.core.vt0_t,  D.core.beta,   D.core.phi_t,
                                            D.core.gamma,  D.core.lambda, D.core.vtherm, D.core.isat_d,
                                            D.core.isat_s, D.core.grd,    D.core.grs};
            for (int f = 0; f < M1_CORE_F; f++) m1_core[f * s + off] = core[f];
            const double cp[M1_CAPS_F] = {D.caps.cox, D.caps.cgs_ov, D.caps.cgd_ov, D.caps.cgb_ov};
            for (int f = 0; f < M1_CAPS_F; f++) m1_caps[f * s + off] = cp[f];
            if (D.has_caps) has_caps[d] = 1;
            if (inst == 0) h0.m1[d] = D;
        };
        if (it == ov.end()) {
            Mos1Derived D = derive_mos1(g.mos1_models[c.model], g.mos1_insts[c.inst], o);
            for (int i = 0; i < NI; i++) put(i, D);
        } else {
            for (int i = 0; i < NI; i++) {
                s21_params mp = g.mos1_models[c.model], ip = g.mos1_insts[c.inst];
                for (const Override* x : it->second) {
                    s21_params& tp = x->target == S21_OVR_MODEL ? mp : ip;
                    if (x->target == S21_OVR_VALUE) fail(S21_INVALID, "value override on a Mos1");
                    tp.v[x->field] = x->vals[i];
                    tp.set |= 1ull << x->field;
                }
                put(i, derive_mos1(mp, ip, o));
            }
        }
    }
    std::vector<double>& d_par = hp.d_par;
    d_par.assign((size_t)D_PAR_F * NI * n_d, 0.0);
    h0.dio.resize(n_d);
    for (int d = 0; d < n_d; d++) {
        const Comp& c = t.comps[dc[d]];
        auto it = ov.find(dc[d]);
        for (int i = 0; i < NI; i++) {
            s21_params mp = g.diode_models[c.model], ip = g.diode_insts[c.inst];
            if (it != ov.end())
                for (const Override* x : it->second) {
                    if (x->target == S21_OVR_VALUE) fail(S21_INVALID, "value override on a Diode");
                    s21_params& tp = x->target == S21_OVR_MODEL ? mp : ip;
                    tp.v[x->field] = x->vals[i];
                    tp.set |= 1ull << x->field;
                }
            DiodeP P = derive_diode(mp, ip, o);
            if (i == 0) h0.dio[d] = P;
            const double f[D_PAR_F] = {P.isat, P.vte, P.vcrit, P.gspr, P.bv, P.has_bv, P.cz,
                                       P.cz2,  P.dep, P.f1,    P.f3,   P.tt, P.vj,     P.m};
            const size_t s = (size_t)NI * n_d, off = (size_t)i * n_d + d;
            for (int k = 0; k < D_PAR_F; k++) d_par[k * s + off] = f[k];
            if (it == ov.end() && i == 0) {
                for (int i2 = 1; i2 < NI; i2++)
                    for (int k = 0; k < D_PAR_F; k++) d_par[k * s + (size_t)i2 * n_d + d] = f[k];
                break;
            }
        }
    }
    std::vector<double>& lin_val = hp.lin_val;
    std::vector<double>& lin_tran = hp.lin_tran;
    lin_val.assign((size_t)NI * n_lin, 0.0);
    lin_tran.assign((size_t)NI * n_lin, 0.0);
    h0.lin_value.resize(n_lin);
    for (int d = 0; d < n_lin; d++) {
        const Comp& c = t.comps[linc[d]];
        auto it = ov.find(linc[d]);
        for (int i = 0; i < NI; i++) {
            double v = c.value, vt = c.value_tran;
            if (it != ov.end())
                for (const Override* x : it->second) {
                    if (x->target != S21_OVR_VALUE) fail(S21_INVALID, "parameter override on a linear element");
                    v = vt = x->vals[i];
                }
            lin_val[(size_t)i * n_lin + d] = v;
            lin_tran[(size_t)i * n_lin + d] = vt;
        }
        h0.lin_value[d] = lin_val[d];
    }

    // Symbolic phase on instance 0's first Newton iteration (x = 0, operating point).
    std::vector<double> x0(t.n, 0.0), vals, rhs;
    host_assemble_op(t, h0, o, x0, vals, rhs);
    topo_pivot(t, vals);
    topo_schedule(t, has_caps);
}

// Derive per-instance parameters, build the topology and upload everything for one analysis.
static void prepare_group(s21_batch* b, GroupHost& g, const s21_tran_options* tr) {
    HostPrep hp;
    host_prepare(b->opts, g, tr, hp);
    const Topology& t = g.topo;
    const int NI = g.n_inst;
    const int n_m1 = hp.n_m1, n_d = hp.n_d, n_lin = hp.n_lin;
    const std::vector<double>&m1_core = hp.m1_core, &m1_caps = hp.m1_caps, &d_par = hp.d_par,
                              &lin_val = hp.lin_val, &lin_tran = hp.lin_tran;

    // ---- upload topology (one int buffer) ----
    std::vector<int> ib;
    std::vector<size_t> offs;
    auto push = [&](const std::vector<int>& v) {
        offs.push_back(ib.size());
        ib.insert(ib.end(), v.begin(), v.end());
        if (v.empty()) ib.push_back(0);
    };
    auto as_int = [](const std::vector<uint32_t>& v) { return std::vector<int>(v.begin(), v.end()); };
    auto u8_int = [](const std::vector<uint8_t>& v) { return std::vector<int>(v.begin(), v.end()); };
    push(t.m1_nodes);            // 0
    push(t.m1_base);             // 1
    push(as_int(t.m1_mask));     // 2
    push(u8_int(t.m1_caps));     // 3
    push(t.m0_nodes);            // 4
    push(t.m0_base);             // 5
    push(as_int(t.m0_mask));     // 6
    push(t.d_nodes);             // 7
    push(t.d_base);              // 8
    push(as_int(t.d_mask));      // 9
    push(t.lin_type);            // 10
    push(t.lin_nodes);           // 11
    push(t.lin_base);            // 12
    push(as_int(t.lin_mask));    // 13
    push(t.g_ptr);               // 14
    push(t.g_idx);               // 15
    push(t.res_ptr);             // 16
    push(t.res_pos);             // 17
    push(t.res_xcol);            // 18
    push(t.row_i2e);             // 19
    push(t.col_e2i);             // 20
    push(t.rp);                  // 21
    push(t.ci);                  // 22
    push(t.diag);                // 23
    push(t.upd_ptr);             // 24
    push(t.upd_src);             // 25
    push(t.upd_dst);             // 26
    push(t.lvl_ptr);             // 27
    push(t.lvl_rows);            // 28
    push(t.blvl_ptr);            // 29
    push(t.blvl_rows);           // 30
    const size_t dbl_off_bytes = ((ib.size() * sizeof(int) + 255) / 256) * 256;
    const size_t n_m0 = t.m0_p.size();
    g.d_topo.reset(new DevMem());
    g.d_topo->alloc(dbl_off_bytes + std::max<size_t>(n_m0, 1) * sizeof(double));
    HIPCHECK(hipMemcpy(g.d_topo->p, ib.data(), ib.size() * sizeof(int), hipMemcpyHostToDevice));
    if (n_m0)
        HIPCHECK(hipMemcpy((char*)g.d_topo->p + dbl_off_bytes, t.m0_p.data(), n_m0 * sizeof(double),
                           hipMemcpyHostToDevice));
    const int* base = (const int*)g.d_topo->p;
    TopoGPU& T = g.gg.t;
    T.n = t.n;
    T.nnz = t.nnz;
    T.n_contrib = t.n_contrib;
    T.n_m1 = n_m1;
    T.n_m0 = (int)n_m0;
    T.n_d = n_d;
    T.n_lin = n_lin;
    T.n_lvl = (int)t.lvl_ptr.size() - 1;
    T.n_blvl = (int)t.blvl_ptr.size() - 1;
    T.sym_status = t.sym_status;
    T.m1_nodes = base + offs[0];
    T.m1_base = base + offs[1];
    T.m1_mask = (const uint32_t*)(base + offs[2]);
    T.m1_caps = base + offs[3];
    T.m0_nodes = base + offs[4];
    T.m0_base = base + offs[5];
    T.m0_mask = (const uint32_t*)(base + offs[6]);
    T.m0_p = (const double*)((const char*)g.d_topo->p + dbl_off_bytes);
    T.d_nodes = base + offs[7];
    T.d_base = base + offs[8];
    T.d_mask = (const uint32_t*)(base + offs[9]);
    T.lin_type = base + offs[10];
    T.lin_nodes = base + offs[11];
    T.lin_base = base + offs[12];
    T.lin_mask = (const uint32_t*)(base + offs[13]);
    T.g_ptr = base + offs[14];
    T.g_idx = base + offs[15];
    T.res_ptr = base + offs[16];
    T.res_pos = base + offs[17];
    T.res_xcol = base + offs[18];
    T.row_i2e = base + offs[19];
    T.col_e2i = base + offs[20];
    T.rp = base + offs[21];
    T.ci = base + offs[22];
    T.diag = base + offs[23];
    T.upd_ptr = base + offs[24];
    T.upd_src = base + offs[25];
    T.upd_dst = base + offs[26];
    T.lvl_ptr = base + offs[27];
    T.lvl_rows = base + offs[28];
    T.blvl_ptr = base + offs[29];
    T.blvl_rows = base + offs[30];

    // ---- parameters (read-only) ----
    auto pad = [](size_t n) { return ((n * sizeof(double) + 255) / 256) * 256; };
    const size_t s_core = pad(m1_core.size()), s_caps = pad(m1_caps.size()), s_dp = pad(d_par.size()),
                 s_lin = pad(lin_val.size());
    g.d_par.reset(new DevMem());
    g.d_par->alloc(s_core + s_caps + s_dp + 2 * s_lin + 256);
    char* pp = (char*)g.d_par->p;
    auto up = [&](size_t off, const std::vector<double>& v) {
        if (!v.empty()) HIPCHECK(hipMemcpy(pp + off, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
        return (const double*)(pp + off);
    };
    g.gg.m1_core = up(0, m1_core);
    g.gg.m1_caps = up(s_core, m1_caps);
    g.gg.d_par = up(s_core + s_caps, d_par);
    g.gg.lin_val = up(s_core + s_caps + s_dp, lin_val);
    g.gg.lin_val_tran = up(s_core + s_caps + s_dp + s_lin, lin_tran);

    // ---- state (zero-initialised: Default op points, comps/mos.rs:524-546) ----
    const size_t s_m1s = pad((size_t)M1_STATE_F * 2 * NI * n_m1), s_ds = pad((size_t)D_STATE_F * NI * n_d),
                 s_cq = pad((size_t)NI * n_lin);
    g.d_state.reset(new DevMem());
    g.d_state->alloc(s_m1s + s_ds + s_cq + 256);
    HIPCHECK(hipMemsetAsync(g.d_state->p, 0, g.d_state->bytes, b->stream));
    char* sp = (char*)g.d_state->p;
    g.gg.m1_state = (double*)sp;
    g.gg.d_state = (double*)(sp + s_m1s);
    g.gg.c_q = (double*)(sp + s_m1s + s_ds);

    // ---- outputs ----
    const size_t s_x = pad((size_t)NI * t.n), s_st = ((NI * sizeof(int) + 255) / 256) * 256,
                 s_it = ((NI * sizeof(long long) + 255) / 256) * 256, s_dbg = pad((size_t)t.nnz + t.n);
    g.d_out.reset(new DevMem());
    g.d_out->alloc(s_x + s_st + s_it + s_dbg + 256);
    HIPCHECK(hipMemsetAsync(g.d_out->p, 0, g.d_out->bytes, b->stream));
    char* op = (char*)g.d_out->p;
    g.gg.x = (double*)op;
    g.gg.status = (int*)(op + s_x);
    g.gg.iters = (long long*)(op + s_x + s_st);
    g.gg.dbg = (double*)(op + s_x + s_st + s_it);
    g.gg.n_inst = NI;
    g.gg.wave = nullptr;

    HIPCHECK(hipStreamSynchronize(b->stream));
    g.dpl = n_m1 == 0 ? 1 : n_m1 <= 64 ? 1 : n_m1 <= 128 ? 2 : n_m1 <= 256 ? 4 : 0;
    g.smem = (size_t)(3 * t.n + t.n_contrib + t.nnz) * sizeof(double) +
             (b->precision == S21_MIXED ? (size_t)t.nnz * sizeof(float) : 0);
}

template <int DPL, bool MIXED>
static void launch_one(s21_batch* b, const GroupGPU* dgroups, int n_groups, int n_blocks, size_t smem,
                       const Args& A) {
    auto kern = newton_kernel<DPL, MIXED>;
    if (smem > 64 * 1024) HIPCHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    hipLaunchKernelGGL(kern, dim3(n_blocks), dim3(WAVE), smem, b->stream, dgroups, n_groups, A);
    HIPCHECK(hipGetLastError());
}

template <bool MIXED>
static void launch_dpl(s21_batch* b, int dpl, const GroupGPU* dg, int ng, int nb, size_t smem, const Args& A) {
    switch (dpl) {
    case 1: launch_one<1, MIXED>(b, dg, ng, nb, smem, A); break;
    case 2: launch_one<2, MIXED>(b, dg, ng, nb, smem, A); break;
    case 4: launch_one<4, MIXED>(b, dg, ng, nb, smem, A); break;
    default: launch_one<0, MIXED>(b, dg, ng, nb, smem, A); break;
    }
}

constexpr size_t MAX_SMEM = 160 * 1024;

// Runs the kernel over every group (one launch per register-residency class).
static void run(s21_batch* b, const Args& A0, double* d_wave, int64_t npts) {
    // order groups by residency class, assign block ranges
    std::vector<int> order(b->groups.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->groups[x].dpl < b->groups[y].dpl; });
    std::vector<GroupGPU> hg;
    std::vector<int> cls;
    size_t wave_off = 0;
    for (int gi : order) {
        GroupHost& g = b->groups[gi];
        if (g.smem > MAX_SMEM)
            fail(S21_INVALID, "circuit too large for the batched kernel: " + std::to_string(g.smem) + " bytes of LDS");
        GroupGPU gg = g.gg;
        gg.wave = d_wave ? d_wave + wave_off : nullptr;
        wave_off += (size_t)g.n_inst * (size_t)npts * g.topo.n;
        hg.push_back(gg);
        cls.push_back(g.dpl);
    }
    if (!b->d_groups) b->d_groups.reset(new DevMem());
    b->d_groups->alloc(std::max<size_t>(hg.size(), 1) * sizeof(GroupGPU));
    // block ranges within each class launch
    size_t i = 0;
    b->launches = 0;
    float total = 0.0f;
    while (i < hg.size()) {
        size_t j = i;
        int blk = 0;
        size_t smem = 0;
        while (j < hg.size() && cls[j] == cls[i]) {
            hg[j].blk0 = blk;
            blk += (A0.mode == MODE_ASSEMBLE) ? 1 : hg[j].n_inst;
            smem = std::max(smem, b->groups[order[j]].smem);
            j++;
        }
        HIPCHECK(hipMemcpyAsync((GroupGPU*)b->d_groups->p + i, hg.data() + i, (j - i) * sizeof(GroupGPU),
                                hipMemcpyHostToDevice, b->stream));
        HIPCHECK(hipEventRecord(b->ev0, b->stream));
        if (b->precision == S21_MIXED)
            launch_dpl<true>(b, cls[i], (GroupGPU*)b->d_groups->p + i, (int)(j - i), blk, smem, A0);
        else
            launch_dpl<false>(b, cls[i], (GroupGPU*)b->d_groups->p + i, (int)(j - i), blk, smem, A0);
        HIPCHECK(hipEventRecord(b->ev1, b->stream));
        HIPCHECK(hipEventSynchronize(b->ev1));
        float ms = 0.0f;
        HIPCHECK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        total += ms;
        b->launches++;
        i = j;
    }
    b->last_ms = total;
    b->kernel_avg_ms = b->launches ? total / b->launches : 0.0;
}

static void fetch_results(s21_batch* b) {
    for (GroupHost& g : b->groups) {
        const int NI = g.n_inst, n = g.topo.n;
        g.x.resize((size_t)NI * n);
        g.status.resize(NI);
        g.iters.resize(NI);
        HIPCHECK(hipMemcpy(g.x.data(), g.gg.x, g.x.size() * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(g.status.data(), g.gg.status, NI * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(g.iters.data(), g.gg.iters, NI * sizeof(long long), hipMemcpyDeviceToHost));
    }
}

static Args base_args(const s21_batch* b, int mode) {
    Args A{};
    A.mode = mode;
    A.max_iter = 100;  // analysis.rs:173
    A.gmin = b->opts.gmin;
    A.iabstol = b->opts.iabstol;
    A.reltol = b->opts.reltol;
    A.tstep = 0.0;
    A.tstop = 0.0;
    A.npts = 1;
    A.dbg_inst = 0;
    return A;
}

static void group_from(GroupHost& g, const s21_group& sg) {
    if (!sg.net || sg.n_inst <= 0) fail(S21_INVALID, "s21_batch_create: empty group");
    const s21_netlist& nl = *sg.net;
    if (nl.n_vars < 0 || nl.n_comps < 0 || (nl.n_comps > 0 && !nl.comps)) fail(S21_INVALID, "malformed netlist");
    g.n_vars = nl.n_vars;
    for (int k = 0; k < nl.n_vars; k++) {
        g.names.push_back(nl.var_name && nl.var_name[k] ? nl.var_name[k] : ("v" + std::to_string(k)));
        g.kinds.push_back(nl.var_kind ? nl.var_kind[k] : S21_VAR_V);
    }
    g.comps.assign(nl.comps, nl.comps + nl.n_comps);
    if (nl.n_mos1_models) g.mos1_models.assign(nl.mos1_models, nl.mos1_models + nl.n_mos1_models);
    if (nl.n_mos1_insts) g.mos1_insts.assign(nl.mos1_insts, nl.mos1_insts + nl.n_mos1_insts);
    if (nl.n_diode_models) g.diode_models.assign(nl.diode_models, nl.diode_models + nl.n_diode_models);
    if (nl.n_diode_insts) g.diode_insts.assign(nl.diode_insts, nl.diode_insts + nl.n_diode_insts);
    g.n_inst = sg.n_inst;
    for (int k = 0; k < sg.n_over; k++) {
        const s21_override& o = sg.over[k];
        if (o.comp < 0 || o.comp >= nl.n_comps || !o.vals) fail(S21_INVALID, "override: bad component");
        if (o.target != S21_OVR_VALUE && (o.field < 2 || o.field >= S21_PARAM_SLOTS))
            fail(S21_INVALID, "override: bad field");
        Override x;
        x.comp = o.comp;
        x.target = o.target;
        x.field = o.field;
        x.vals.assign(o.vals, o.vals + sg.n_inst);
        g.overs.push_back(std::move(x));
    }
}

template <typename F>
static int guard(F&& f) {
    try {
        f();
        return S21_OK;
    } catch (Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (std::exception& e) {
        set_last_error(e.what());
        return S21_ERR;
    }
}

}  // namespace s21

using namespace s21;

extern "C" {

int s21_symbolic(const s21_netlist* net, const s21_options* opts, const s21_tran_options* tran, int32_t* row_i2e,
                 int32_t* col_i2e, int64_t* nnz_a, int64_t* nnz_lu, int64_t* levels, int64_t* flops_lu) {
    return guard([&] {
        if (!net) fail(S21_INVALID, "null netlist");
        GroupHost g;
        s21_group sg{net, 1, 0, nullptr};
        group_from(g, sg);
        s21_options o;
        if (opts)
            o = *opts;
        else
            s21_default_options(&o);
        HostPrep hp;
        host_prepare(o, g, tran, hp);
        const Topology& t = g.topo;
        if (t.sym_status != S21_OK) fail(t.sym_status, t.sym_error);
        for (int k = 0; k < t.n; k++) {
            if (row_i2e) row_i2e[k] = t.row_i2e[k];
            if (col_i2e) col_i2e[k] = t.col_i2e[k];
        }
        if (nnz_a) *nnz_a = (int64_t)t.erow.size();
        if (nnz_lu) *nnz_lu = t.nnz;
        if (levels) *levels = (int64_t)t.lvl_ptr.size() - 1 + (int64_t)t.blvl_ptr.size() - 1;
        if (flops_lu) *flops_lu = t.flops_lu;
    });
}

int s21_batch_create(int32_t n_groups, const s21_group* groups, const s21_options* opts, int32_t precision,
                     s21_batch** out) {
    return guard([&] {
        if (!out || n_groups <= 0 || !groups) fail(S21_INVALID, "s21_batch_create: bad arguments");
        *out = nullptr;
        check_device();
        std::unique_ptr<s21_batch> b(new s21_batch());
        if (opts)
            b->opts = *opts;
        else
            s21_default_options(&b->opts);
        b->precision = precision == S21_MIXED ? S21_MIXED : S21_FP64;
        b->groups.resize(n_groups);
        for (int gi = 0; gi < n_groups; gi++) group_from(b->groups[gi], groups[gi]);
        HIPCHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        HIPCHECK(hipEventCreate(&b->ev0));
        HIPCHECK(hipEventCreate(&b->ev1));
        *out = b.release();
    });
}

void s21_batch_free(s21_batch* b) { delete b; }

int s21_batch_op(s21_batch* b) {
    return guard([&] {
        if (!b) fail(S21_INVALID, "null batch");
        for (GroupHost& g : b->groups) prepare_group(b, g, nullptr);
        Args A = base_args(b, MODE_OP);
        run(b, A, nullptr, 1);
        fetch_results(b);
        b->npts = 0;
        b->axis.clear();
        b->last_mode = MODE_OP;
    });
}

int s21_batch_tran(s21_batch* b, const s21_tran_options* tr) {
    return guard([&] {
        if (!b || !tr) fail(S21_INVALID, "null argument");
        for (GroupHost& g : b->groups) prepare_group(b, g, tr);
        const int64_t npts = s21_tran_point_count(tr->tstep, tr->tstop);
        size_t total = 0;
        for (GroupHost& g : b->groups) total += (size_t)g.n_inst * (size_t)npts * g.topo.n;
        if (!b->d_wave) b->d_wave.reset(new DevMem());
        b->d_wave->alloc(std::max<size_t>(total, 1) * sizeof(double));
        Args A = base_args(b, MODE_TRAN);
        A.tstep = tr->tstep;
        A.tstop = tr->tstop;
        A.npts = npts;
        run(b, A, (double*)b->d_wave->p, npts);
        fetch_results(b);
        b->npts = npts;
        b->axis.resize(npts);
        double t = 0.0;
        b->axis[0] = 0.0;
        t = tr->tstep;
        for (int64_t p = 1; p < npts; p++) {
            b->axis[p] = t;
            t += tr->tstep;
        }
        b->last_mode = MODE_TRAN;
    });
}

int s21_batch_ac(s21_batch* b, const s21_ac_options* a) {
    (void)a;
    return guard([&] {
        if (!b) fail(S21_INVALID, "null batch");
        fail(S21_ERR, "AC analysis is not available in this build");
    });
}

int32_t s21_batch_n_groups(const s21_batch* b) { return b ? (int32_t)b->groups.size() : 0; }
int32_t s21_batch_n_inst(const s21_batch* b, int32_t g) {
    return (b && g >= 0 && g < (int)b->groups.size()) ? b->groups[g].n_inst : 0;
}
int32_t s21_batch_n_vars(const s21_batch* b, int32_t g) {
    if (!b || g < 0 || g >= (int)b->groups.size()) return 0;
    const GroupHost& gh = b->groups[g];
    return gh.topo.n ? gh.topo.n : gh.n_vars;
}
const char* s21_batch_var_name(const s21_batch* b, int32_t g, int32_t k) {
    if (!b || g < 0 || g >= (int)b->groups.size()) return nullptr;
    const GroupHost& gh = b->groups[g];
    const std::vector<std::string>& nm = gh.topo.n ? gh.topo.names : gh.names;
    if (k < 0 || k >= (int)nm.size()) return nullptr;
    return nm[k].c_str();
}
int s21_batch_get_x(const s21_batch* b, int32_t g, double* out) {
    return guard([&] {
        if (!b || g < 0 || g >= (int)b->groups.size() || !out) fail(S21_INVALID, "bad arguments");
        const GroupHost& gh = b->groups[g];
        std::copy(gh.x.begin(), gh.x.end(), out);
    });
}
int s21_batch_get_status(const s21_batch* b, int32_t g, int32_t* status, int64_t* iters) {
    return guard([&] {
        if (!b || g < 0 || g >= (int)b->groups.size()) fail(S21_INVALID, "bad arguments");
        const GroupHost& gh = b->groups[g];
        for (int i = 0; i < gh.n_inst && i < (int)gh.status.size(); i++) {
            if (status) status[i] = gh.status[i];
            if (iters) iters[i] = gh.iters[i];
        }
    });
}
int64_t s21_batch_n_points(const s21_batch* b) { return b ? b->npts : 0; }
int s21_batch_get_axis(const s21_batch* b, double* out) {
    return guard([&] {
        if (!b || !out) fail(S21_INVALID, "bad arguments");
        std::copy(b->axis.begin(), b->axis.end(), out);
    });
}
int s21_batch_get_wave(const s21_batch* b, int32_t g, int32_t inst, double* out) {
    return guard([&] {
        if (!b || g < 0 || g >= (int)b->groups.size() || !out) fail(S21_INVALID, "bad arguments");
        if (b->last_mode != MODE_TRAN || !b->d_wave) fail(S21_INVALID, "no transient waveform");
        const GroupHost& gh = b->groups[g];
        if (inst < 0 || inst >= gh.n_inst) fail(S21_INVALID, "instance out of range");
        // groups are laid out in residency-class order (see run())
        std::vector<int> order(b->groups.size());
        for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return b->groups[x].dpl < b->groups[y].dpl; });
        size_t off = 0;
        for (int gi : order) {
            if (gi == g) break;
            off += (size_t)b->groups[gi].n_inst * (size_t)b->npts * b->groups[gi].topo.n;
        }
        const size_t per = (size_t)b->npts * gh.topo.n;
        HIPCHECK(hipMemcpy(out, (double*)b->d_wave->p + off + (size_t)inst * per, per * sizeof(double),
                           hipMemcpyDeviceToHost));
    });
}
int s21_batch_timing(const s21_batch* b, double* total_ms, int64_t* launches, double* kernel_avg_ms) {
    return guard([&] {
        if (!b) fail(S21_INVALID, "null batch");
        if (total_ms) *total_ms = b->last_ms;
        if (launches) *launches = b->launches;
        if (kernel_avg_ms) *kernel_avg_ms = b->kernel_avg_ms;
    });
}
int s21_batch_pivots(const s21_batch* b, int32_t g, int32_t* row_i2e, int32_t* col_i2e) {
    return guard([&] {
        if (!b || g < 0 || g >= (int)b->groups.size()) fail(S21_INVALID, "bad arguments");
        const Topology& t = b->groups[g].topo;
        if (t.sym_status != S21_OK) fail(t.sym_status, t.sym_error);
        for (int k = 0; k < t.n; k++) {
            if (row_i2e) row_i2e[k] = t.row_i2e[k];
            if (col_i2e) col_i2e[k] = t.col_i2e[k];
        }
    });
}
int s21_batch_structure(const s21_batch* b, int32_t g, int64_t* nnz_a, int64_t* nnz_lu, int64_t* levels,
                        int64_t* n_contrib) {
    return guard([&] {
        if (!b || g < 0 || g >= (int)b->groups.size()) fail(S21_INVALID, "bad arguments");
        const Topology& t = b->groups[g].topo;
        if (nnz_a) *nnz_a = (int64_t)t.erow.size();
        if (nnz_lu) *nnz_lu = t.nnz;
        if (levels) *levels = (int64_t)t.lvl_ptr.size() - 1 + (int64_t)t.blvl_ptr.size() - 1;
        if (n_contrib) *n_contrib = t.n_contrib;
    });
}
int s21_batch_assemble(s21_batch* b, int32_t g, int32_t inst, const double* x, int32_t* row_out, int32_t* col_out,
                       double* val_out, double* rhs_out, int64_t* n_elem) {
    return guard([&] {
        if (!b || g < 0 || g >= (int)b->groups.size() || !x) fail(S21_INVALID, "bad arguments");
        GroupHost& gh = b->groups[g];
        if (inst < 0 || inst >= gh.n_inst) fail(S21_INVALID, "instance out of range");
        for (GroupHost& o : b->groups) prepare_group(b, o, nullptr);
        const Topology& t = gh.topo;
        HIPCHECK(hipMemcpy(gh.gg.x + (size_t)inst * t.n, x, t.n * sizeof(double), hipMemcpyHostToDevice));
        // run only this group
        std::vector<GroupHost> saved;
        Args A = base_args(b, MODE_ASSEMBLE);
        A.dbg_inst = inst;
        GroupGPU gg = gh.gg;
        gg.blk0 = 0;
        if (!b->d_groups) b->d_groups.reset(new DevMem());
        b->d_groups->alloc(sizeof(GroupGPU));
        HIPCHECK(hipMemcpy(b->d_groups->p, &gg, sizeof(GroupGPU), hipMemcpyHostToDevice));
        if (b->precision == S21_MIXED)
            launch_dpl<true>(b, gh.dpl, (GroupGPU*)b->d_groups->p, 1, 1, gh.smem, A);
        else
            launch_dpl<false>(b, gh.dpl, (GroupGPU*)b->d_groups->p, 1, 1, gh.smem, A);
        HIPCHECK(hipStreamSynchronize(b->stream));
        std::vector<double> dbg(t.nnz + t.n);
        HIPCHECK(hipMemcpy(dbg.data(), gh.gg.dbg, dbg.size() * sizeof(double), hipMemcpyDeviceToHost));
        const int64_t ne = (int64_t)t.erow.size();
        for (int64_t e = 0; e < ne; e++) {
            if (row_out) row_out[e] = t.erow[e];
            if (col_out) col_out[e] = t.ecol[e];
            if (val_out) val_out[e] = dbg[t.elem_pos[e]];
        }
        if (rhs_out)
            for (int k = 0; k < t.n; k++) rhs_out[k] = dbg[t.nnz + k];
        if (n_elem) *n_elem = ne;
        b->last_mode = MODE_ASSEMBLE;
    });
}

}  // extern "C"
