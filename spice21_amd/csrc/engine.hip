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
    int n, nnz, n_contrib, n_part;
    int n_m1, n_m0, n_d, n_lin;
    int n_gt, n_gc, n_rt, n_rc, n_clvl, n_blvl;
    int sym_status;
    int blob_vec;       // 16-byte vectors of the record blob copied into LDS at launch
    const uint4* blob;  // the blob in global memory
    // offsets (uint16 units, 8-byte aligned) of the record tables inside the blob
    int o_gt, o_gidx, o_gc, o_rt, o_re, o_rc, o_ri2e, o_ce2i, o_diag, o_clvl, o_fch, o_fl, o_fu, o_blvl, o_bro, o_bu;
    // per-device tables (global; Mos1 tables are read once into registers)
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
    long long* prof;        // [n_inst][8] phase cycle counters (profiling mode) or nullptr
};

enum { MODE_OP = 0, MODE_TRAN = 1, MODE_ASSEMBLE = 2 };
struct Args {
    int mode;
    int max_iter;
    double gmin, iabstol, reltol;
    double tstep, tstop;
    long long npts;
    int dbg_inst;
    int prof;
    int waves;          // instances (wavefronts) per workgroup
    int inst_bytes;     // LDS bytes per instance
};

// Address spaces: pointers read out of descriptors carry no address space, so they are re-qualified in
// device code to keep LDS accesses on ds_* and global accesses on global_* instructions.
#define LDS __attribute__((address_space(3)))
#define GLB __attribute__((address_space(1)))
template <typename T>
__device__ inline GLB T* glb(T* p) {
    return (GLB T*)p;
}

constexpr int WAVE = 64;
constexpr int MAX_WAVES = 4;  // wavefronts (instances) per workgroup; the kernel's launch bound
constexpr int M1_CORE_F = 11, M1_CAPS_F = 4, M1_STATE_F = 7, D_PAR_F = 14, D_STATE_F = 3;
constexpr int NPROF = 8;  // eval, gather, residual, factor (+forward sweep), backward sweep, update, commit, total
using u16 = unsigned short;

// Orders LDS accesses between the lanes of one wavefront (a wavefront executes its LDS operations in
// order; the fences keep the compiler from moving accesses across this point).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double w = __shfl_xor(v, o, WAVE);
        v = (w > v) ? w : v;  // the reference's fold drops NaN (analysis.rs:198)
    }
    return v;
}
__device__ inline bool wave_any(bool p) { return __ballot(p) != 0ull; }

// Writes a stamp into its contribution slot when the entry is live for this device.
struct Slot {
    LDS double* contrib;
    uint32_t mask;
    int base;
    __device__ void operator()(int e, double v) const {
        if ((mask >> e) & 1u) contrib[base + __popc(mask & ((1u << e) - 1u))] = v;
    }
};

// Per-lane register copy of the lane's Mos1 devices (DPL > 0).
template <int DPL>
struct Mos1Regs {
    Mos1Core P[DPL];
    int nd01[DPL], nd23[DPL];  // node pairs packed as 16-bit halves (+1, so ground is 0)
    uint32_t mask[DPL];
    int base[DPL];  // contribution base; bit 30 flags a device with gate capacitance
};
__device__ inline int lo16(int v) { return (v & 0xFFFF) - 1; }
__device__ inline int hi16(int v) { return ((v >> 16) & 0xFFFF) - 1; }

// LDS views of the record tables.
struct Idx {
    LDS const uint64_t *gt, *gc, *rt, *rc, *fch, *fl, *bro;  // four 16-bit fields
    LDS const uint32_t *re, *fu, *bu;                         // two 16-bit fields
    LDS const u16 *gidx, *ri2e, *ce2i, *diag, *clvl, *blvl;
};
struct R4 {
    int x, y, z, w;
};
__device__ inline R4 r4(uint64_t v) {
    return R4{(int)(v & 0xFFFF), (int)((v >> 16) & 0xFFFF), (int)((v >> 32) & 0xFFFF), (int)(v >> 48)};
}
struct R2 {
    int x, y;
};
__device__ inline R2 r2(uint32_t v) { return R2{(int)(v & 0xFFFF), (int)(v >> 16)}; }

struct Ctx {
    const GroupGPU* G;
    TopoGPU T;
    Idx I;
    int inst, lane;
    LDS double *x, *rhs, *c, *contrib, *a, *part;
    LDS float* lu32;
};

__device__ inline double xv(LDS const double* x, int v) { return v < 0 ? 0.0 : x[v]; }

__device__ inline Mos1Core load_mos1_core(const GroupGPU& G, int n_m1, int inst, int d) {
    GLB const double* p = glb(G.m1_core);
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

// Mos1 with gate capacitance: capacitances and operating-point state come from global memory.
template <bool TRAN>
__device__ inline void eval_mos1_caps(const Ctx& C, const Mos1Core& P, int d, double vd, double vg, double vs,
                                            double vb, double gmin, double dt, Slot put) {
    const GroupGPU& G = *C.G;
    const int n_m1 = C.T.n_m1;
    const size_t s = (size_t)G.n_inst * n_m1, o = (size_t)C.inst * n_m1 + d;
    Mos1Caps K;
    GLB const double* kc = glb(G.m1_caps);
    K.cox = kc[0 * s + o];
    K.cgs_ov = kc[1 * s + o];
    K.cgd_ov = kc[2 * s + o];
    K.cgb_ov = kc[3 * s + o];
    Mos1State op, gs;
    GLB double* st = glb(G.m1_state);
    op.vgs = st[0 * s + o];
    op.vgd = st[1 * s + o];
    op.vgb = st[2 * s + o];
    op.cgs = st[3 * s + o];
    op.cgd = st[4 * s + o];
    op.cgb = st[5 * s + o];
    op.reversed = st[6 * s + o];
    mos1_load<TRAN>(P, &K, &op, vd, vg, vs, vb, gmin, dt, put, &gs);
    st[7 * s + o] = gs.vgs;
    st[8 * s + o] = gs.vgd;
    st[9 * s + o] = gs.vgb;
    st[10 * s + o] = gs.cgs;
    st[11 * s + o] = gs.cgd;
    st[12 * s + o] = gs.cgb;
    st[13 * s + o] = gs.reversed;
}

__device__ inline void eval_mos1(const Ctx& C, const Mos1Core& P, int d, int n0, int n1, int n2, int n3, bool caps,
                                 Slot put, bool tran, const Args& A) {
    const double vd = xv(C.x, n0), vg = xv(C.x, n1), vs = xv(C.x, n2), vb = xv(C.x, n3);
    if (caps) {
        if (tran)
            eval_mos1_caps<true>(C, P, d, vd, vg, vs, vb, A.gmin, A.tstep, put);
        else
            eval_mos1_caps<false>(C, P, d, vd, vg, vs, vb, A.gmin, A.tstep, put);
    } else {
        mos1_load<false>(P, nullptr, nullptr, vd, vg, vs, vb, A.gmin, A.tstep, put, nullptr);
    }
}

// Diodes, level-0 MOS and linear elements (few per instance): parameters from global memory.
__device__ inline void eval_others(const Ctx& C, bool tran, const Args& A) {
    const GroupGPU& G = *C.G;
    const TopoGPU& T = C.T;
    LDS const double* x = C.x;
    const double gmin = A.gmin, dt = A.tstep;
    for (int d = C.lane; d < T.n_m0; d += WAVE) {
        GLB const int* nd = glb(T.m0_nodes) + 4 * d;
        mos0_load(glb(T.m0_p)[d], xv(x, nd[0]), xv(x, nd[1]), xv(x, nd[2]), gmin,
                  Slot{C.contrib, glb(T.m0_mask)[d], glb(T.m0_base)[d]});
    }
    for (int d = C.lane; d < T.n_d; d += WAVE) {
        const size_t s = (size_t)G.n_inst * T.n_d, o = (size_t)C.inst * T.n_d + d;
        GLB const double* q = glb(G.d_par);
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
        GLB double* st = glb(G.d_state);
        GLB const int* nd = glb(T.d_nodes) + 3 * d;
        const Slot put{C.contrib, glb(T.d_mask)[d], glb(T.d_base)[d]};
        DiodeGuess gs;
        if (tran)
            diode_load<true>(P, xv(x, nd[2]), xv(x, nd[1]), gmin, dt, st[0 * s + o], st[2 * s + o], put, &gs);
        else
            diode_load<false>(P, xv(x, nd[2]), xv(x, nd[1]), gmin, dt, st[0 * s + o], st[2 * s + o], put, &gs);
        st[0 * s + o] = gs.vd;
        st[1 * s + o] = gs.charge;
    }
    for (int d = C.lane; d < T.n_lin; d += WAVE) {
        const int typ = glb(T.lin_type)[d];
        GLB const int* nd = glb(T.lin_nodes) + 3 * d;
        const size_t o = (size_t)C.inst * T.n_lin + d;
        const double val = tran ? glb(G.lin_val_tran)[o] : glb(G.lin_val)[o];
        const Slot put{C.contrib, glb(T.lin_mask)[d], glb(T.lin_base)[d]};
        if (typ == S21_COMP_R) {
            put(0, val);
            put(1, val);
            put(2, -val);
            put(3, -val);
        } else if (typ == S21_COMP_C) {
            if (tran)
                cap_load_tran(val, xv(x, nd[0]) - xv(x, nd[1]), glb(G.c_q)[o], dt, put);
            else
                for (int k = 0; k < 6; k++) put(k, 0.0);  // no stamps in the operating point
        } else if (typ == S21_COMP_V) {
            put(0, 1.0);
            put(1, 1.0);
            put(2, -1.0);
            put(3, -1.0);
            put(4, val);
        } else {
            put(0, val);
            put(1, -val);
        }
    }
}

template <int DPL>
__device__ inline void eval_devices(const Ctx& C, const Mos1Regs<(DPL > 0 ? DPL : 1)>& R, bool tran, const Args& A) {
    const TopoGPU& T = C.T;
    if (DPL > 0) {
#pragma unroll
        for (int j = 0; j < (DPL > 0 ? DPL : 1); j++) {
            const int d = C.lane + WAVE * j;
            if (d < T.n_m1)
                eval_mos1(C, R.P[j], d, lo16(R.nd01[j]), hi16(R.nd01[j]), lo16(R.nd23[j]), hi16(R.nd23[j]),
                          (R.base[j] >> 30) & 1, Slot{C.contrib, R.mask[j], R.base[j] & 0x3FFFFFFF}, tran, A);
        }
    } else {
        for (int d = C.lane; d < T.n_m1; d += WAVE) {
            const int* nd = T.m1_nodes + 6 * d;
            eval_mos1(C, load_mos1_core(*C.G, T.n_m1, C.inst, d), d, nd[0], nd[1], nd[2], nd[3], T.m1_caps[d] != 0,
                      Slot{C.contrib, T.m1_mask[d], T.m1_base[d]}, tran, A);
        }
    }
    if (T.n_m0 + T.n_d + T.n_lin) eval_others(C, tran, A);
}

__device__ inline LDS double& slot_ref(const Ctx& C, int s) { return s < C.T.nnz ? C.a[s] : C.rhs[s - C.T.nnz]; }

// Ordered gather of the contributions into A (LU positions) and the RHS.  Each short slot is summed in
// component-then-stamp order -- the order Solver::update adds them (analysis.rs:153-168); a long slot is
// summed in chunks whose partial sums are then added in order.
__device__ inline void gather(const Ctx& C) {
    const Idx& I = C.I;
    for (int t = C.lane; t < C.T.n_gt; t += WAVE) {
        const R4 r = r4(I.gt[t]);
        double acc = 0.0;
        for (int k = r.y; k < r.z; k++) acc = acc + C.contrib[I.gidx[k]];
        if (r.w == 0xFFFF)
            slot_ref(C, r.x) = acc;
        else
            C.part[r.w] = acc;
    }
    wave_sync();
    for (int t = C.lane; t < C.T.n_gc; t += WAVE) {
        const R4 r = r4(I.gc[t]);
        double acc = C.part[r.y];
        for (int p = r.y + 1; p < r.z; p++) acc = acc + C.part[p];
        slot_ref(C, r.x) = acc;
    }
}

// c[i] = rhs[row_i2e[i]] - (A x)[i] over internal rows (sparse21/mod.rs:298-327).  Returns whether this
// lane saw |res| > iabstol.
__device__ inline bool residual(const Ctx& C, double iabstol) {
    const Idx& I = C.I;
    bool bad = false;
    for (int t = C.lane; t < C.T.n_rt; t += WAVE) {
        const R4 r = r4(I.rt[t]);
        double acc = 0.0;
        for (int e = r.y; e < r.z; e++) {
            const R2 q = r2(I.re[e]);
            acc = acc + C.a[q.x] * C.x[q.y];
        }
        if (r.w == 0xFFFF) {
            const double v = C.rhs[I.ri2e[r.x]] - acc;
            C.c[r.x] = v;
            bad = bad || (fabs(v) > iabstol);
        } else {
            C.part[r.w] = acc;
        }
    }
    wave_sync();
    for (int t = C.lane; t < C.T.n_rc; t += WAVE) {
        const R4 r = r4(I.rc[t]);
        double acc = C.part[r.y];
        for (int p = r.y + 1; p < r.z; p++) acc = acc + C.part[p];
        const double v = C.rhs[I.ri2e[r.x]] - acc;
        C.c[r.x] = v;
        bad = bad || (fabs(v) > iabstol);
    }
    return bad;
}

// Factorisation with the forward sweep fused in.  A chunk of row i processes L entries (i, k) once row k
// is final: l_ik = a_ik / u_kk, a_ij -= l_ik u_kj over row k's U entries, and c_i -= c_k l_ik -- for each
// target in increasing k, the order of the reference's right-looking elimination and column-oriented
// forward substitution (sparse21/mod.rs:865-919, 946-964).  Returns S21_PIVOT when a pivot other than
// the last is zero (the assertion at sparse21/mod.rs:872).
template <typename V>
__device__ inline int factor_forward(const Ctx& C, LDS V* lu) {
    const Idx& I = C.I;
    for (int l = 0; l < C.T.n_clvl; l++) {
        const int e = I.clvl[l + 1];
        for (int ch = I.clvl[l] + C.lane; ch < e; ch += WAVE) {
            const R4 h = r4(I.fch[ch]);
            V y = (V)C.c[h.x];
            int u = h.w;
            for (int j = h.y; j < h.z; j++) {
                const R4 L = r4(I.fl[j]);
                const V lv = lu[L.x] / lu[L.y];
                lu[L.x] = lv;
                const V cp = (V)C.c[L.z] * lv;
                y = y - cp;
                for (int m = 0; m < L.w; m++, u++) {
                    const R2 q = r2(I.fu[u]);
                    const V prod = lu[q.x] * lv;
                    lu[q.y] = lu[q.y] - prod;
                }
            }
            C.c[h.x] = (double)y;
        }
        wave_sync();
    }
    bool zero_pivot = false;
    for (int i = C.lane; i < C.T.n - 1; i += WAVE)
        if (lu[I.diag[i]] == V(0)) zero_pivot = true;
    return wave_any(zero_pivot) ? S21_PIVOT : S21_OK;
}

// Backward sweep over U by row levels (sparse21/mod.rs:966-979).
template <typename V>
__device__ inline void backward(const Ctx& C, LDS const V* lu) {
    const Idx& I = C.I;
    for (int l = 0; l < C.T.n_blvl; l++) {
        const int e = I.blvl[l + 1];
        for (int r = I.blvl[l] + C.lane; r < e; r += WAVE) {
            const R4 h = r4(I.bro[r]);
            V y = (V)C.c[h.x];
            for (int k = h.z; k < h.w; k++) {
                const R2 q = r2(I.bu[k]);
                const V prod = (V)C.c[q.y] * lu[q.x];
                y = y - prod;
            }
            C.c[h.x] = (double)(y / lu[h.y]);
        }
        wave_sync();
    }
}

// Commit the accepted operating point (Component::commit) of every stateful device.
__device__ inline void commit(const Ctx& C) {
    const GroupGPU& G = *C.G;
    const TopoGPU& T = C.T;
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

__device__ inline long long cycles() { return (long long)__builtin_readcyclecounter(); }

// One Newton solve (Solver::solve, analysis.rs:169-210).  `iters` accumulates loads.  Returns an S21_* status.
template <int DPL, bool MIXED>
__device__ int newton(const Ctx& C, const Mos1Regs<(DPL > 0 ? DPL : 1)>& R, bool tran, const Args& A,
                      long long& iters, long long* pc) {
    const TopoGPU& T = C.T;
    const bool prof = A.prof != 0;
    long long t0 = 0, t1 = 0;
    bool prev_ok = true;  // dx starts at zero
    for (int it = 0; it < A.max_iter; it++) {
        iters++;
        if (prof) t0 = cycles();
        eval_devices<DPL>(C, R, tran, A);
        wave_sync();
        if (prof) { t1 = cycles(); pc[0] += t1 - t0; t0 = t1; }
        gather(C);
        wave_sync();
        if (prof) { t1 = cycles(); pc[1] += t1 - t0; t0 = t1; }
        const bool bad = wave_any(residual(C, A.iabstol));
        wave_sync();
        if (prof) { t1 = cycles(); pc[2] += t1 - t0; t0 = t1; }
        if (prev_ok && !bad) {
            commit(C);
            wave_sync();
            if (prof) { t1 = cycles(); pc[6] += t1 - t0; }
            return S21_OK;
        }
        if (T.sym_status != S21_OK) return T.sym_status;
        int st;
        if (MIXED) {
            for (int p = C.lane; p < T.nnz; p += WAVE) C.lu32[p] = (float)C.a[p];
            wave_sync();
            st = factor_forward<float>(C, C.lu32);
            if (prof) { t1 = cycles(); pc[3] += t1 - t0; t0 = t1; }
            if (st == S21_OK) backward<float>(C, C.lu32);
        } else {
            st = factor_forward<double>(C, C.a);
            if (prof) { t1 = cycles(); pc[3] += t1 - t0; t0 = t1; }
            if (st == S21_OK) backward<double>(C, C.a);
        }
        if (prof) { t1 = cycles(); pc[4] += t1 - t0; t0 = t1; }
        if (st != S21_OK) return st;
        // clamp and update (analysis.rs:196-207)
        double mx = 0.0;
        for (int k = C.lane; k < T.n; k += WAVE) {
            const double dx = fabs(C.c[C.I.ce2i[k]]);
            mx = (dx > mx) ? dx : mx;
        }
        mx = wave_max(mx);
        bool ok = true;
        for (int k = C.lane; k < T.n; k += WAVE) {
            double dx = C.c[C.I.ce2i[k]];
            if (mx > 1.0) dx = dx * 1.0 / mx;
            ok = ok && !(fabs(dx) > A.reltol);
            C.x[k] = C.x[k] + dx;
        }
        prev_ok = !wave_any(!ok);
        wave_sync();
        if (prof) { t1 = cycles(); pc[5] += t1 - t0; }
    }
    return S21_CONVERGENCE;
}

// Workgroup = A.waves wavefronts, one circuit instance each, sharing one LDS copy of the group's records.
template <int DPL, bool MIXED>
__global__ __launch_bounds__(WAVE * MAX_WAVES) void newton_kernel(const GroupGPU* __restrict__ groups, int n_groups, Args A) {
    int g = 0;
    const int b = blockIdx.x;
    while (g + 1 < n_groups && groups[g + 1].blk0 <= b) g++;
    const GroupGPU& G = groups[g];
    Ctx C;
    C.G = &G;
    C.T = G.t;
    const TopoGPU& T = C.T;
    extern __shared__ uint4 smem_u4[];
    // records -> LDS (whole workgroup)
    {
        GLB const uint64_t* src = (GLB const uint64_t*)T.blob;
        LDS uint64_t* dst = (LDS uint64_t*)smem_u4;
        for (int k = threadIdx.x; k < 2 * T.blob_vec; k += blockDim.x) dst[k] = src[k];
    }
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int inst = (A.mode == MODE_ASSEMBLE) ? A.dbg_inst : (b - G.blk0) * A.waves + w;
    if (inst < 0 || inst >= G.n_inst || (A.mode == MODE_ASSEMBLE && w != 0)) return;
    C.inst = inst;
    C.lane = threadIdx.x % WAVE;
    LDS const u16* b16 = (LDS const u16*)smem_u4;
    C.I.gt = (LDS const uint64_t*)(b16 + T.o_gt);
    C.I.gidx = b16 + T.o_gidx;
    C.I.gc = (LDS const uint64_t*)(b16 + T.o_gc);
    C.I.rt = (LDS const uint64_t*)(b16 + T.o_rt);
    C.I.re = (LDS const uint32_t*)(b16 + T.o_re);
    C.I.rc = (LDS const uint64_t*)(b16 + T.o_rc);
    C.I.ri2e = b16 + T.o_ri2e;
    C.I.ce2i = b16 + T.o_ce2i;
    C.I.diag = b16 + T.o_diag;
    C.I.clvl = b16 + T.o_clvl;
    C.I.fch = (LDS const uint64_t*)(b16 + T.o_fch);
    C.I.fl = (LDS const uint64_t*)(b16 + T.o_fl);
    C.I.fu = (LDS const uint32_t*)(b16 + T.o_fu);
    C.I.blvl = b16 + T.o_blvl;
    C.I.bro = (LDS const uint64_t*)(b16 + T.o_bro);
    C.I.bu = (LDS const uint32_t*)(b16 + T.o_bu);
    LDS double* dbase = (LDS double*)((LDS char*)((LDS uint4*)smem_u4 + T.blob_vec) + (size_t)w * A.inst_bytes);
    C.x = dbase;
    C.rhs = C.x + T.n;
    C.c = C.rhs + T.n;
    C.contrib = C.c + T.n;
    C.a = C.contrib + T.n_contrib;
    C.part = C.a + T.nnz;
    C.lu32 = (LDS float*)(C.part + T.n_part);

    constexpr int NR = DPL > 0 ? DPL : 1;
    Mos1Regs<NR> R;
#pragma unroll
    for (int j = 0; j < NR; j++) {
        const int d = C.lane + WAVE * j;
        if (DPL > 0 && d < T.n_m1) {
            R.P[j] = load_mos1_core(G, T.n_m1, inst, d);
            const int* nd = T.m1_nodes + 6 * d;
            R.nd01[j] = (nd[0] + 1) | ((nd[1] + 1) << 16);
            R.nd23[j] = (nd[2] + 1) | ((nd[3] + 1) << 16);
            R.mask[j] = T.m1_mask[d];
            R.base[j] = T.m1_base[d] | (T.m1_caps[d] ? (1 << 30) : 0);
        }
    }
    for (int k = C.lane; k < T.n; k += WAVE) C.x[k] = G.x[(size_t)inst * T.n + k];
    wave_sync();

    if (A.mode == MODE_ASSEMBLE) {
        eval_devices<DPL>(C, R, false, A);
        wave_sync();
        gather(C);
        wave_sync();
        for (int s = C.lane; s < T.nnz; s += WAVE) G.dbg[s] = C.a[s];
        for (int k = C.lane; k < T.n; k += WAVE) G.dbg[T.nnz + k] = C.rhs[k];
        return;
    }

    long long pc[NPROF] = {0, 0, 0, 0, 0, 0, 0, 0};
    const long long tstart = A.prof ? cycles() : 0;
    long long iters = 0;
    int st = newton<DPL, MIXED>(C, R, false, A, iters, pc);
    if (A.mode == MODE_TRAN && st == S21_OK) {
        double* wv = G.wave + (size_t)inst * (size_t)A.npts * T.n;
        for (int k = C.lane; k < T.n; k += WAVE) wv[k] = C.x[k];
        long long pt = 1;
        double t = A.tstep;  // analysis.rs:553-570
        while (t < A.tstop && pt < A.npts) {
            st = newton<DPL, MIXED>(C, R, true, A, iters, pc);
            if (st != S21_OK) break;
            for (int k = C.lane; k < T.n; k += WAVE) wv[(size_t)pt * T.n + k] = C.x[k];
            pt++;
            t += A.tstep;
        }
    }
    for (int k = C.lane; k < T.n; k += WAVE) G.x[(size_t)inst * T.n + k] = C.x[k];
    if (C.lane == 0) {
        G.status[inst] = st;
        G.iters[inst] = iters;
        if (A.prof && G.prof) {
            pc[7] = cycles() - tstart;
            for (int k = 0; k < NPROF; k++) G.prof[(size_t)inst * NPROF + k] = pc[k];
        }
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
    size_t blob_bytes = 0, inst_bytes = 0;
    std::string prepared;  // analysis the uploaded structure was built for ("" = none)
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
    int profile = 0;
    int last_waves = 1;
    std::vector<double> axis;
    std::unique_ptr<s21::DevMem> d_wave, d_groups;
    int last_mode = -1;
    bool fetched = true;  // host copies of x / status / iterations are current
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

    // ---- record tables: one uint16 blob, copied into LDS by every workgroup ----
    topo_records(g.topo);
    std::vector<uint16_t> blob;
    auto sect = [&](const std::vector<uint16_t>& v) {
        const int off = (int)blob.size();
        blob.insert(blob.end(), v.begin(), v.end());
        while (blob.size() % 4) blob.push_back(0);  // 8-byte alignment for ushort4 records
        if (v.empty()) blob.insert(blob.end(), 4, 0);
        return off;
    };
    auto to16 = [](const std::vector<int>& v) {
        std::vector<uint16_t> o;
        for (int x : v) {
            if (x < 0 || x > 65535) fail(S21_INVALID, "circuit too large for the batched kernel (16-bit index)");
            o.push_back((uint16_t)x);
        }
        return o;
    };
    std::vector<int> dg(t.diag);
    for (int& v : dg) v = std::max(v, 0);  // the unpivoted fallback pattern may lack diagonals
    TopoGPU& T = g.gg.t;
    T.o_gt = sect(t.gt);
    T.o_gidx = sect(to16(t.g_idx));
    T.o_gc = sect(t.gc);
    T.o_rt = sect(t.rt);
    T.o_re = sect(t.re);
    T.o_rc = sect(t.rc);
    T.o_ri2e = sect(to16(t.row_i2e));
    T.o_ce2i = sect(to16(t.col_e2i));
    T.o_diag = sect(to16(dg));
    T.o_clvl = sect(to16(t.clvl_ptr.empty() ? std::vector<int>{0} : t.clvl_ptr));
    T.o_fch = sect(t.fch);
    T.o_fl = sect(t.fl);
    T.o_fu = sect(t.fu);
    T.o_blvl = sect(to16(t.blvl_ptr));
    T.o_bro = sect(t.bro);
    T.o_bu = sect(t.bu);
    while (blob.size() % 8) blob.push_back(0);
    T.n_gt = (int)t.gt.size() / 4;
    T.n_gc = (int)t.gc.size() / 4;
    T.n_rt = (int)t.rt.size() / 4;
    T.n_rc = (int)t.rc.size() / 4;
    T.n_part = t.n_part;

    // ---- per-device tables (global) ----
    std::vector<int> ib;
    std::vector<size_t> io;
    auto push = [&](const std::vector<int>& v) {
        io.push_back(ib.size());
        ib.insert(ib.end(), v.begin(), v.end());
        if (v.empty()) ib.push_back(0);
    };
    auto as_int = [](const std::vector<uint32_t>& v) { return std::vector<int>(v.begin(), v.end()); };
    auto u8_int = [](const std::vector<uint8_t>& v) { return std::vector<int>(v.begin(), v.end()); };
    push(t.m1_nodes);          // 0
    push(t.m1_base);           // 1
    push(as_int(t.m1_mask));   // 2
    push(u8_int(t.m1_caps));   // 3
    push(t.m0_nodes);          // 4
    push(t.m0_base);           // 5
    push(as_int(t.m0_mask));   // 6
    push(t.d_nodes);           // 7
    push(t.d_base);            // 8
    push(as_int(t.d_mask));    // 9
    push(t.lin_type);          // 10
    push(t.lin_nodes);         // 11
    push(t.lin_base);          // 12
    push(as_int(t.lin_mask));  // 13
    const size_t n_m0 = t.m0_p.size();
    while (blob.size() % 8) blob.push_back(0);  // whole 16-byte vectors
    const size_t blob_bytes = blob.size() * sizeof(uint16_t);
    const size_t ib_off = ((blob_bytes + 255) / 256) * 256;
    const size_t dbl_off = ib_off + ((ib.size() * sizeof(int) + 255) / 256) * 256;
    g.d_topo.reset(new DevMem());
    g.d_topo->alloc(dbl_off + std::max<size_t>(n_m0, 1) * sizeof(double));
    char* tb = (char*)g.d_topo->p;
    HIPCHECK(hipMemcpy(tb, blob.data(), blob_bytes, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(tb + ib_off, ib.data(), ib.size() * sizeof(int), hipMemcpyHostToDevice));
    if (n_m0) HIPCHECK(hipMemcpy(tb + dbl_off, t.m0_p.data(), n_m0 * sizeof(double), hipMemcpyHostToDevice));
    const int* base = (const int*)(tb + ib_off);
    T.n = t.n;
    T.nnz = t.nnz;
    T.n_contrib = t.n_contrib;
    T.n_m1 = n_m1;
    T.n_m0 = (int)n_m0;
    T.n_d = n_d;
    T.n_lin = n_lin;
    T.n_clvl = (int)t.clvl_ptr.size() - 1;
    if (T.n_clvl < 0) T.n_clvl = 0;
    T.n_part = t.n_part;
    T.n_blvl = (int)t.blvl_ptr.size() - 1;
    T.sym_status = t.sym_status;
    T.blob_vec = (int)(blob_bytes / 16);
    T.blob = (const uint4*)tb;
    T.m1_nodes = base + io[0];
    T.m1_base = base + io[1];
    T.m1_mask = (const uint32_t*)(base + io[2]);
    T.m1_caps = base + io[3];
    T.m0_nodes = base + io[4];
    T.m0_base = base + io[5];
    T.m0_mask = (const uint32_t*)(base + io[6]);
    T.m0_p = (const double*)(tb + dbl_off);
    T.d_nodes = base + io[7];
    T.d_base = base + io[8];
    T.d_mask = (const uint32_t*)(base + io[9]);
    T.lin_type = base + io[10];
    T.lin_nodes = base + io[11];
    T.lin_base = base + io[12];
    T.lin_mask = (const uint32_t*)(base + io[13]);

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
                 s_it = ((NI * sizeof(long long) + 255) / 256) * 256, s_dbg = pad((size_t)t.nnz + t.n),
                 s_prof = b->profile ? ((NI * NPROF * sizeof(long long) + 255) / 256) * 256 : 0;
    g.d_out.reset(new DevMem());
    g.d_out->alloc(s_x + s_st + s_it + s_dbg + s_prof + 256);
    HIPCHECK(hipMemsetAsync(g.d_out->p, 0, g.d_out->bytes, b->stream));
    char* op = (char*)g.d_out->p;
    g.gg.x = (double*)op;
    g.gg.status = (int*)(op + s_x);
    g.gg.iters = (long long*)(op + s_x + s_st);
    g.gg.dbg = (double*)(op + s_x + s_st + s_it);
    g.gg.prof = b->profile ? (long long*)(op + s_x + s_st + s_it + s_dbg) : nullptr;
    g.gg.n_inst = NI;
    g.gg.wave = nullptr;

    HIPCHECK(hipStreamSynchronize(b->stream));
    g.dpl = n_m1 == 0 ? 1 : n_m1 <= 64 ? 1 : n_m1 <= 128 ? 2 : n_m1 <= 256 ? 4 : 0;
    g.blob_bytes = blob_bytes;
    g.inst_bytes = (((size_t)(3 * t.n + t.n_contrib + t.nnz + t.n_part) * sizeof(double) +
                     (b->precision == S21_MIXED ? (size_t)t.nnz * sizeof(float) : 0)) + 15) / 16 * 16;
}

// Key of the structure an analysis needs: the operating point, or a transient with its IC set.
static std::string prep_key(const s21_tran_options* tr) {
    if (!tr) return "op";
    std::string k = "tran";
    for (int i = 0; i < tr->n_ic; i++) k += ":" + std::to_string(tr->ic_var[i]);
    return k;
}

// Prepare every group for an analysis, or only reset the Newton state when the structure is current.
static void prepare_all(s21_batch* b, const s21_tran_options* tr) {
    const std::string key = prep_key(tr);
    for (GroupHost& g : b->groups) {
        if (g.prepared == key) {
            HIPCHECK(hipMemsetAsync(g.d_state->p, 0, g.d_state->bytes, b->stream));
            HIPCHECK(hipMemsetAsync(g.d_out->p, 0, g.d_out->bytes, b->stream));
        } else {
            g.prepared.clear();
            prepare_group(b, g, tr);
            g.prepared = key;
        }
    }
    HIPCHECK(hipStreamSynchronize(b->stream));
}

template <int DPL, bool MIXED>
static void launch_one(s21_batch* b, const GroupGPU* dgroups, int n_groups, int n_blocks, size_t smem,
                       const Args& A) {
    auto kern = newton_kernel<DPL, MIXED>;
    if (smem > 64 * 1024) HIPCHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    hipLaunchKernelGGL(kern, dim3(n_blocks), dim3(WAVE * A.waves), smem, b->stream, dgroups, n_groups, A);
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

constexpr size_t LDS_PER_CU = 160 * 1024;

// Instances per workgroup: the count that keeps the most instances resident per CU (every workgroup
// carries one copy of the records plus per-instance state).
static int pick_waves(size_t blob, size_t inst, int max_inst, size_t max_block) {
    int best = 1;
    size_t best_res = 0;
    for (int k = 1; k <= MAX_WAVES && k <= std::max(max_inst, 1); k++) {
        const size_t bytes = blob + (size_t)k * inst;
        if (bytes > max_block) break;
        const size_t res = (LDS_PER_CU / bytes) * k;
        if (res > best_res) {
            best_res = res;
            best = k;
        }
    }
    return best;
}

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
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    int max_block = 0;
    HIPCHECK(hipDeviceGetAttribute(&max_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    while (i < hg.size()) {
        size_t j = i;
        size_t blob = 0, inst = 0;
        int max_inst = 0;
        while (j < hg.size() && cls[j] == cls[i]) {
            blob = std::max(blob, b->groups[order[j]].blob_bytes);
            inst = std::max(inst, b->groups[order[j]].inst_bytes);
            max_inst = std::max(max_inst, hg[j].n_inst);
            j++;
        }
        Args A = A0;
        A.waves = A0.mode == MODE_ASSEMBLE ? 1 : pick_waves(blob, inst, max_inst, (size_t)max_block);
        A.inst_bytes = (int)inst;
        const size_t smem = blob + (size_t)A.waves * inst;
        if (smem > (size_t)max_block)
            fail(S21_INVALID, "circuit too large for the batched kernel: " + std::to_string(smem) + " bytes of LDS");
        int blk = 0;
        for (size_t q = i; q < j; q++) {
            hg[q].blk0 = blk;
            blk += (A.mode == MODE_ASSEMBLE) ? 1 : (hg[q].n_inst + A.waves - 1) / A.waves;
        }
        b->last_waves = A.waves;
        HIPCHECK(hipMemcpyAsync((GroupGPU*)b->d_groups->p + i, hg.data() + i, (j - i) * sizeof(GroupGPU),
                                hipMemcpyHostToDevice, b->stream));
        HIPCHECK(hipEventRecord(b->ev0, b->stream));
        if (b->precision == S21_MIXED)
            launch_dpl<true>(b, cls[i], (GroupGPU*)b->d_groups->p + i, (int)(j - i), blk, smem, A);
        else
            launch_dpl<false>(b, cls[i], (GroupGPU*)b->d_groups->p + i, (int)(j - i), blk, smem, A);
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

// Results stay on the device after an analysis; the first query copies them to the host.
static void fetch_results(s21_batch* b) {
    if (b->fetched) return;
    b->fetched = true;
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
    A.prof = b->profile;
    A.waves = 1;
    A.inst_bytes = 0;
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
        prepare_all(b, nullptr);
        Args A = base_args(b, MODE_OP);
        run(b, A, nullptr, 1);
        b->fetched = false;
        b->npts = 0;
        b->axis.clear();
        b->last_mode = MODE_OP;
    });
}

int s21_batch_tran(s21_batch* b, const s21_tran_options* tr) {
    return guard([&] {
        if (!b || !tr) fail(S21_INVALID, "null argument");
        if (!(tr->tstep > 0.0) && tr->tstop > 0.0) fail(S21_INVALID, "tstep must be positive");
        // IC values live in the uploaded element values: rebuild when they are part of the key.
        for (GroupHost& g : b->groups) g.prepared.clear();
        prepare_all(b, tr);
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
        b->fetched = false;
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
        fetch_results(const_cast<s21_batch*>(b));
        const GroupHost& gh = b->groups[g];
        std::copy(gh.x.begin(), gh.x.end(), out);
    });
}
int s21_batch_get_status(const s21_batch* b, int32_t g, int32_t* status, int64_t* iters) {
    return guard([&] {
        if (!b || g < 0 || g >= (int)b->groups.size()) fail(S21_INVALID, "bad arguments");
        fetch_results(const_cast<s21_batch*>(b));
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
int s21_batch_set_profiling(s21_batch* b, int32_t on) {
    return guard([&] {
        if (!b) fail(S21_INVALID, "null batch");
        if (b->profile != (on ? 1 : 0))
            for (s21::GroupHost& g : b->groups) g.prepared.clear();
        b->profile = on ? 1 : 0;
    });
}
int s21_batch_profile(const s21_batch* b, int32_t g, int64_t* out) {
    return guard([&] {
        if (!b || g < 0 || g >= (int)b->groups.size() || !out) fail(S21_INVALID, "bad arguments");
        const GroupHost& gh = b->groups[g];
        if (!gh.gg.prof) fail(S21_INVALID, "profiling was off during the last analysis");
        HIPCHECK(hipMemcpy(out, gh.gg.prof, (size_t)gh.n_inst * NPROF * sizeof(long long), hipMemcpyDeviceToHost));
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
        prepare_all(b, nullptr);
        const Topology& t = gh.topo;
        HIPCHECK(hipMemcpy(gh.gg.x + (size_t)inst * t.n, x, t.n * sizeof(double), hipMemcpyHostToDevice));
        // run only this group
        std::vector<GroupHost> saved;
        Args A = base_args(b, MODE_ASSEMBLE);
        A.dbg_inst = inst;
        A.waves = 1;
        A.inst_bytes = (int)gh.inst_bytes;
        GroupGPU gg = gh.gg;
        gg.blk0 = 0;
        if (!b->d_groups) b->d_groups.reset(new DevMem());
        b->d_groups->alloc(sizeof(GroupGPU));
        HIPCHECK(hipMemcpy(b->d_groups->p, &gg, sizeof(GroupGPU), hipMemcpyHostToDevice));
        if (b->precision == S21_MIXED)
            launch_dpl<true>(b, gh.dpl, (GroupGPU*)b->d_groups->p, 1, 1, gh.blob_bytes + gh.inst_bytes, A);
        else
            launch_dpl<false>(b, gh.dpl, (GroupGPU*)b->d_groups->p, 1, 1, gh.blob_bytes + gh.inst_bytes, A);
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
