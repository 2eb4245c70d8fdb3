// Single large circuit instance on MI355X (gfx950): the whole Newton inner loop for one netlist of tens of
// thousands of unknowns, which does not fit the batched kernel's per-instance LDS region.
//
//   * device evaluation, one thread per device, every stamp to its own contribution cell in HBM
//     (models.h, the same functions the batched kernel runs);
//   * ordered gather: each matrix element and RHS entry sums its contributions in component order, then
//     stamp order -- the order Solver::update adds them (analysis.rs:153-168) -- without atomics;
//   * the KCL residual in FP64 and the reference's convergence test (analysis.rs:331-345);
//   * a multifrontal LU over the order large_sym.cpp fixes: fronts of up to SMALL_MAX entries a side
//     assembled and factored in one workgroup's LDS, many fronts per launch by elimination-tree levels;
//     larger fronts assembled in column blocks and factored by panels (a diagonal-block LU in LDS,
//     triangular solves for the panel's L rows and U columns, a tiled update of the trailing block);
//   * forward and backward sweeps front by front, level by level;
//   * the max-|dx| <= 1 clamp and update (analysis.rs:196-207), commit of device state on convergence.
// The Newton loop, the fixed-step Backward-Euler time loop and IC forcing (analysis.rs:489-574) are
// driven from the host with one 16-byte read-back per iteration.  Mixed precision factors and sweeps in
// FP32; evaluation, residual, update and the convergence test stay FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "bsim4_eval.h"
#include "internal.h"
#include "large.h"
#include "models.h"

#define LCHECK(x)                                                                           \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) s21::fail(S21_DEVICE, std::string("HIP: ") + hipGetErrorString(e_)); \
    } while (0)

namespace s21 {
namespace lg {

constexpr int TPB = 256;
constexpr int NB = 64;                   // panel width of the blocked factorisation of large fronts
constexpr int CB = 16;                   // column block of the large-front assembly
constexpr size_t SMALL_LDS = 128 * 1024; // LDS bytes a small front may take
constexpr int M1F = 11, M1C = 4, M1S = 7, DPF = 14, DSF = 4;

// Orders the LDS accesses of one wavefront's lanes (a wavefront executes its LDS operations in order; the
// fences keep the compiler from moving accesses across this point).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
constexpr int small_max() {
    return sizeof(T) == 8 ? 128 : 181;  // F * F * sizeof(T) <= SMALL_LDS
}
static_assert(small_max<double>() <= TPB && small_max<float>() <= TPB,
              "the small-front sweeps give every row of a front a thread of its own");

// ---------------------------------------------------------------------------------------------------
// Device evaluation
// ---------------------------------------------------------------------------------------------------
struct EvalArgs {
    int n_m1, n_d, n_lin, tran;
    double gmin, dt;
    const double* x;
    const int* m1_node;   // [n_m1][6] d g s b dp sp (-1 = ground)
    const double* m1_par; // [M1F + M1C][n_m1]
    const int* m1_caps;   // [n_m1] 1 when the device carries gate charge
    double* m1_state;     // [2 * M1S][n_m1]: committed op, then guess
    const int* m1_cb;     // [n_m1] first contribution cell
    const int* d_node;    // [n_d][3] p n r
    const double* d_par;  // [DPF][n_d]
    double* d_state;      // [DSF][n_d]: guess vd, guess charge, committed charge, vd the last load started from
    const int* d_cb;
    const int* l_node;    // [n_lin][3]
    const int* l_type;    // [n_lin] S21_COMP_*
    const double* l_val;  // [2][n_lin] operating-point and transient values
    const double* c_q;    // [n_lin] committed capacitor charge
    const int* l_cb;
    double* cv;           // contribution cells
    // trapezoidal integration: trap = 1 integrates by the trapezoidal rule, 2 by Backward Euler while
    // recording each charge's current (the first step, from an operating point whose capacitor currents
    // are zero), 0 by Backward Euler alone.  Per charge, the current at the last accepted point
    // (committed) and the one of the current guess.
    int trap;
    double* m1_icur;      // [4][n_m1]: committed gate-drain (tr.gs) and gate-bulk currents, then guesses
    double* d_icur;       // [2][n_d]: committed junction-charge current, then guess
    double* c_icur;       // [2][n_lin]: committed capacitor current, then guess
};

__device__ inline double xval(const double* x, int v) { return v < 0 ? 0.0 : x[v]; }

__global__ void __launch_bounds__(TPB) k_eval(EvalArgs E) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < E.n_m1) {
        const int j = t, nm = E.n_m1;
        const int* nd = E.m1_node + 6 * j;
        auto F = [&](int f) { return E.m1_par[(size_t)f * nm + j]; };
        Mos1Core P{F(0), F(1), F(2), F(3), F(4), F(5), F(6), F(7), F(8), F(9), F(10)};
        double* cv = E.cv + E.m1_cb[j];
        auto put = [&](int e, double v) { cv[e] = v; };
        const double vd = xval(E.x, nd[RD]), vg = xval(E.x, nd[RG]), vs = xval(E.x, nd[RS]), vb = xval(E.x, nd[RB]);
        if (!E.m1_caps[j]) {
            mos1_load<false>(P, nullptr, nullptr, vd, vg, vs, vb, E.gmin, E.dt, put, nullptr);
            return;
        }
        Mos1Caps K{F(M1F + 0), F(M1F + 1), F(M1F + 2), F(M1F + 3)};
        double* st = E.m1_state;
        Mos1State op{st[0 * nm + j], st[1 * nm + j], st[2 * nm + j], st[3 * nm + j],
                     st[4 * nm + j], st[5 * nm + j], st[6 * nm + j]};
        Mos1State gs{};
        if (E.tran) {
            if (E.trap) {
                const double ip[2] = {E.m1_icur[(size_t)0 * nm + j], E.m1_icur[(size_t)1 * nm + j]};
                double io[2];
                mos1_load<true>(P, &K, &op, vd, vg, vs, vb, E.gmin, E.dt, put, &gs, nullptr, E.trap == 1 ? ip : nullptr,
                                io);
                E.m1_icur[(size_t)2 * nm + j] = io[0];
                E.m1_icur[(size_t)3 * nm + j] = io[1];
            } else {
                mos1_load<true>(P, &K, &op, vd, vg, vs, vb, E.gmin, E.dt, put, &gs);
            }
        } else {
            mos1_load<false>(P, &K, &op, vd, vg, vs, vb, E.gmin, E.dt, put, &gs);
        }
        st[7 * nm + j] = gs.vgs;
        st[8 * nm + j] = gs.vgd;
        st[9 * nm + j] = gs.vgb;
        st[10 * nm + j] = gs.cgs;
        st[11 * nm + j] = gs.cgd;
        st[12 * nm + j] = gs.cgb;
        st[13 * nm + j] = gs.reversed;
        return;
    }
    if (t < E.n_m1 + E.n_d) {
        const int j = t - E.n_m1, nd_ = E.n_d;
        const int* nd = E.d_node + 3 * j;
        auto F = [&](int f) { return E.d_par[(size_t)f * nd_ + j]; };
        DiodeP P{F(0), F(1), F(2), F(3), F(4), F(5), F(6), F(7), F(8), F(9), F(10), F(11), F(12), F(13)};
        double* st = E.d_state;
        const double vd0 = st[0 * nd_ + j];
        double* cv = E.cv + E.d_cb[j];
        auto put = [&](int e, double v) { cv[e] = v; };
        DiodeGuess gs;
        const double vr = xval(E.x, nd[DR_]), vn = xval(E.x, nd[DN_]);
        if (E.tran && E.trap)
            diode_load<true>(P, vr, vn, E.gmin, E.dt, vd0, st[2 * nd_ + j], put, &gs,
                             E.trap == 1 ? E.d_icur + j : nullptr, E.d_icur + nd_ + j);
        else if (E.tran)
            diode_load<true>(P, vr, vn, E.gmin, E.dt, vd0, st[2 * nd_ + j], put, &gs);
        else
            diode_load<false>(P, vr, vn, E.gmin, E.dt, vd0, st[2 * nd_ + j], put, &gs);
        st[0 * nd_ + j] = gs.vd;
        st[1 * nd_ + j] = gs.charge;
        st[3 * nd_ + j] = vd0;
        return;
    }
    if (t < E.n_m1 + E.n_d + E.n_lin) {
        const int j = t - E.n_m1 - E.n_d;
        const int* nd = E.l_node + 3 * j;
        double* cv = E.cv + E.l_cb[j];
        const double val = E.l_val[(size_t)(E.tran ? 1 : 0) * E.n_lin + j];
        switch (E.l_type[j]) {
        case S21_COMP_R:
            cv[0] = val;
            cv[1] = val;
            cv[2] = -val;
            cv[3] = -val;
            break;
        case S21_COMP_C:
            if (!E.tran) {  // no stamps in the operating point (comps/mod.rs:205-209)
                for (int e = 0; e < C_NE; e++) cv[e] = 0.0;
            } else {
                cap_load_tran(val, xval(E.x, nd[0]) - xval(E.x, nd[1]), E.c_q[j], E.dt,
                              [&](int e, double v) { cv[e] = v; }, E.trap == 1 ? E.c_icur + j : nullptr,
                              E.trap ? E.c_icur + E.n_lin + j : nullptr);
            }
            break;
        case S21_COMP_V:
            cv[0] = 1.0;
            cv[1] = 1.0;
            cv[2] = -1.0;
            cv[3] = -1.0;
            cv[4] = val;
            break;
        default:
            cv[0] = val;
            cv[1] = -val;
        }
    }
}

// BSIM4, one thread per device (bsim4solver.rs:134-145, 3757-3769): Bsim4::op from the port voltages with
// limiting against the device's last guess, tran_op's Backward-Euler charge companions against the committed
// charges in a transient, then the 42 stamps into the device's contribution cells.  Parameters and state are
// structures of arrays in HBM (field-major, so consecutive threads read consecutive addresses): parameter
// field K of device j at par[K][j]; state [B4S_F][j]: the last load's limiting fields (vds vgs vbs vbd von)
// and charges (qg qd qb), then the committed charges.
constexpr int B4S_LIM = 0, B4S_Q = 5, B4S_QC = 8, B4S_F = 11;
struct B4Args {
    int n, tran;
    double gmin, dt;
    const double* x;
    const int* node;   // [n][4] d g s b (-1 = ground)
    const double* par; // [parameter doubles][n]
    double* st;        // [B4S_F][n]
    const int* cb;     // [n] first contribution cell
    double* cv;
};
// The parameter view of one device: field K read where it is stored.
struct B4Fields {
    const double* par;
    int n, j;
    template <int K>
    __device__ double get() const {
        return par[(size_t)K * n + j];
    }
};

__global__ void __launch_bounds__(128) k_eval_b4(B4Args A) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.n) return;
    const int* nd = A.node + 4 * j;
    const B4Fields P{A.par, A.n, j};
    const size_t n = A.n;
    b4::B4Op G;  // only the limiting fields of the last load's point are read
    G.vds = A.st[(B4S_LIM + 0) * n + j];
    G.vgs = A.st[(B4S_LIM + 1) * n + j];
    G.vbs = A.st[(B4S_LIM + 2) * n + j];
    G.vbd = A.st[(B4S_LIM + 3) * n + j];
    G.von = A.st[(B4S_LIM + 4) * n + j];
    b4::B4Op N;
    b4::op_eval(P, G, xval(A.x, nd[0]), xval(A.x, nd[1]), xval(A.x, nd[2]), xval(A.x, nd[3]), A.gmin, N);
    b4::B4Tran T{};
    if (A.tran) {
        b4::B4Op O;  // tran_op reads the committed point's charges only
        O.qg = A.st[(B4S_QC + 0) * n + j];
        O.qd = A.st[(B4S_QC + 1) * n + j];
        O.qb = A.st[(B4S_QC + 2) * n + j];
        b4::tran_eval(P, O, N, A.dt, T);
    }
    double* cv = A.cv + A.cb[j];
    b4::stamp(P, N, T, [&](int k, double v) { cv[k] = v; });
    A.st[(B4S_LIM + 0) * n + j] = N.vds;
    A.st[(B4S_LIM + 1) * n + j] = N.vgs;
    A.st[(B4S_LIM + 2) * n + j] = N.vbs;
    A.st[(B4S_LIM + 3) * n + j] = N.vbd;
    A.st[(B4S_LIM + 4) * n + j] = N.von;
    A.st[(B4S_Q + 0) * n + j] = N.qg;
    A.st[(B4S_Q + 1) * n + j] = N.qd;
    A.st[(B4S_Q + 2) * n + j] = N.qb;
}

__global__ void k_commit_b4(int n, double* __restrict__ st) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n)
        for (int f = 0; f < 3; f++) st[(size_t)(B4S_QC + f) * n + j] = st[(size_t)(B4S_Q + f) * n + j];
}

// Each matrix element (then each RHS entry) sums its contributions in component-then-stamp order.
__global__ void __launch_bounds__(TPB) k_gather(int n_slots, const int* __restrict__ ptr, const int* __restrict__ idx,
                                                const double* __restrict__ cv, double* __restrict__ out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_slots) return;
    double acc = 0.0;
    for (int q = ptr[s]; q < ptr[s + 1]; q++) acc = acc + cv[idx[q]];
    out[s] = acc;
}

// res = rhs - A x by rows of A (elements of each row in creation order); flags[0] |= 1 when some
// |res_i| > iabstol (false for NaN, as the reference's comparison); *resmax = max |res_i| (NaN skipped),
// as the bits of a non-negative double.
__global__ void __launch_bounds__(TPB) k_residual(int n, const int* __restrict__ rptr, const int* __restrict__ relem,
                                                  const int* __restrict__ ecol, const double* __restrict__ val,
                                                  const double* __restrict__ rhs, const double* __restrict__ x,
                                                  double* __restrict__ res, double iabstol, int* flags,
                                                  unsigned long long* resmax) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (r < n) {
        double acc = 0.0;
        for (int q = rptr[r]; q < rptr[r + 1]; q++) {
            const int e = relem[q];
            acc = acc + val[e] * x[ecol[e]];
        }
        const double v = rhs[r] - acc;
        res[r] = v;
        if (fabs(v) > iabstol) atomicOr(flags, 1);
        if (fabs(v) == fabs(v)) m = fabs(v);
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double u = __shfl_xor(m, o, 64);
        m = u > m ? u : m;
    }
    if ((threadIdx.x & 63) == 0 && m > 0.0) atomicMax(resmax, (unsigned long long)__double_as_longlong(m));
}

// ---------------------------------------------------------------------------------------------------
// Multifrontal factorisation
// ---------------------------------------------------------------------------------------------------
struct Fronts {
    int ns;
    const int64_t* fo;     // [ns + 1] front offsets in the pool
    const int* first;      // [ns + 1] first internal column
    const int64_t* r_ptr;  // [ns + 1]
    const int* r_idx;      // R_s
    const int* ch_ptr;     // [ns + 1]
    const int* ch;         // children, ascending
    const int64_t* x_ptr;  // [ns + 1] into x_map
    const int* x_map;      // positions of R_s in the parent's index list
    const int64_t* a_ptr;  // [ns + 1] A entries of each front
    const int* a_elem;     // element index
    const int* a_off;      // local offset (column-major) in its front
    const int64_t* w_off;  // [ns + 1] offsets of each front's sweep vector (F entries)
};

__device__ inline int fk(const Fronts& S, int s) { return S.first[s + 1] - S.first[s]; }
__device__ inline int fF(const Fronts& S, int s) { return fk(S, s) + (int)(S.r_ptr[s + 1] - S.r_ptr[s]); }

// Small fronts: assemble (zero, A entries, children's update blocks in order), partial LU, write back;
// one workgroup per front, the front in LDS.
template <typename T>
__global__ void __launch_bounds__(TPB) k_small(Fronts S, const int* __restrict__ list, const double* __restrict__ aval,
                                               T* __restrict__ pool, int* flags) {
    extern __shared__ char lds_raw[];
    T* L = (T*)lds_raw;
    const int s = list[blockIdx.x];
    const int k = fk(S, s), F = fF(S, s);
    const int tid = threadIdx.x;
    for (int i = tid; i < F * F; i += TPB) L[i] = (T)0;
    __syncthreads();
    for (int64_t q = S.a_ptr[s] + tid; q < S.a_ptr[s + 1]; q += TPB) L[S.a_off[q]] = (T)aval[S.a_elem[q]];
    __syncthreads();
    for (int c = S.ch_ptr[s]; c < S.ch_ptr[s + 1]; c++) {
        const int ch = S.ch[c];
        const int kc = fk(S, ch), Fc = fF(S, ch), mc = Fc - kc;
        const int* mp = S.x_map + S.x_ptr[ch];
        const T* U = pool + S.fo[ch];
        for (int q = tid; q < mc * mc; q += TPB) {
            const int j = q / mc, i = q - j * mc;
            L[mp[j] * F + mp[i]] += U[(size_t)(kc + j) * Fc + kc + i];
        }
        __syncthreads();
    }
    for (int j = 0; j < k; j++) {
        const T piv = L[j * F + j];
        if (piv == (T)0 && tid == 0) atomicOr(flags, 2);
        for (int i = j + 1 + tid; i < F; i += TPB) L[j * F + i] = L[j * F + i] / piv;
        __syncthreads();
        const int m = F - j - 1;
        for (int q = tid; q < m * m; q += TPB) {
            const int c = j + 1 + q / m, i = j + 1 + q % m;
            L[c * F + i] = L[c * F + i] - L[j * F + i] * L[c * F + j];
        }
        __syncthreads();
    }
    T* out = pool + S.fo[s];
    for (int i = tid; i < F * F; i += TPB) out[i] = L[i];
}

// Large fronts, assembly: block b of a launch handles column block cb of front s ({s, cb} = job[b]):
// zero its columns, place A's entries there, add the children's update-block columns that map there.
template <typename T>
__global__ void __launch_bounds__(TPB) k_assemble(Fronts S, const int2* __restrict__ job, const double* __restrict__ aval,
                                                  T* __restrict__ pool) {
    const int s = job[blockIdx.x].x, c0 = job[blockIdx.x].y * CB;
    const int F = fF(S, s), c1 = min(F, c0 + CB);
    T* A = pool + S.fo[s];
    const int tid = threadIdx.x;
    const int w = c1 - c0;
    for (int q = tid; q < w * F; q += TPB) A[(size_t)c0 * F + q] = (T)0;
    __syncthreads();
    for (int64_t q = S.a_ptr[s] + tid; q < S.a_ptr[s + 1]; q += TPB) {
        const int o = S.a_off[q], c = o / F;
        if (c >= c0 && c < c1) A[o] = (T)aval[S.a_elem[q]];
    }
    __syncthreads();
    for (int cc = S.ch_ptr[s]; cc < S.ch_ptr[s + 1]; cc++) {
        const int ch = S.ch[cc];
        const int kc = fk(S, ch), Fc = fF(S, ch), mc = Fc - kc;
        const int* mp = S.x_map + S.x_ptr[ch];
        // the child's columns landing in [c0, c1): a contiguous run, the map being increasing
        int lo = 0, hi = mc;
        while (lo < hi) {
            const int md = (lo + hi) / 2;
            if (mp[md] < c0) lo = md + 1; else hi = md;
        }
        int j1 = lo;
        while (j1 < mc && mp[j1] < c1) j1++;
        const T* U = pool + S.fo[ch];
        const int nj = j1 - lo;
        if (nj == 0) continue;  // the same for every thread of the block
        for (int q = tid; q < nj * mc; q += TPB) {
            const int j = lo + q / mc, i = q % mc;
            A[(size_t)mp[j] * F + mp[i]] += U[(size_t)(kc + j) * Fc + kc + i];
        }
        __syncthreads();
    }
}

// Value of v in lane l (l the same for every lane), through the scalar unit.
__device__ inline double lane_value(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline float lane_value(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Large fronts, panel [c, c + w) of front s = list[blockIdx.x]: LU of the w x w diagonal block by one
// wavefront, lane i holding row i in registers; step j reads the pivot row from lane j.
template <typename T>
__global__ void __launch_bounds__(64) k_diag(Fronts S, const int* __restrict__ list, int panel, T* __restrict__ pool,
                                             int* flags) {
    const int s = list[blockIdx.x];
    const int k = fk(S, s), F = fF(S, s);
    const int c = panel * NB;
    if (c >= k) return;
    const int w = min(NB, k - c);
    T* A = pool + S.fo[s];
    const int lane = threadIdx.x;
    const bool own = lane < w;
    T r[NB];
#pragma unroll
    for (int q = 0; q < NB; q++) r[q] = (q < w && own) ? A[(size_t)(c + q) * F + c + lane] : (T)0;
#pragma unroll
    for (int j = 0; j < NB; j++) {
        if (j < w) {
            const T piv = lane_value(r[j], j);
            if (piv == (T)0 && lane == 0) atomicOr(flags, 2);
            // the rows below update under the execution mask (readlane reads the pivot lane whatever the
            // mask), not through a select per entry
            if (lane > j && own) {
                r[j] = r[j] / piv;
                const T l = r[j];
#pragma unroll
                for (int q = j + 1; q < NB; q++) r[q] = r[q] - l * lane_value(r[q], j);
            }
        }
    }
    if (own) {
#pragma unroll
        for (int q = 0; q < NB; q++)
            if (q < w) A[(size_t)(c + q) * F + c + lane] = r[q];
    }
}

// Large fronts, panel triangular solves: blocks [0, nrow) take TPB rows of L21 each (x U11 = a), the
// rest TPB columns of U12 each (L11 y = a, unit lower).  grid.y indexes the front list.
constexpr int TT = 64;  // threads (rows or columns) per block of the panel triangular solves

template <typename T>
__global__ void __launch_bounds__(TT) k_trsm(Fronts S, const int* __restrict__ list, int panel, T* __restrict__ pool) {
    __shared__ T D[NB * NB];
    const int s = list[blockIdx.y];
    const int k = fk(S, s), F = fF(S, s);
    const int c = panel * NB;
    if (c >= k) return;
    const int w = min(NB, k - c);
    const int rest = F - c - w;
    const int nblk = (rest + TT - 1) / TT;
    if ((int)blockIdx.x >= 2 * nblk) return;
    T* A = pool + S.fo[s];
    const int tid = threadIdx.x;
    for (int q = tid; q < w * w; q += TT) {
        const int j = q / w, i = q % w;
        D[j * NB + i] = A[(size_t)(c + j) * F + c + i];
    }
    __syncthreads();
    T v[NB];  // one row (or column) per thread, kept in registers: every loop over it is unrolled
    if ((int)blockIdx.x < nblk) {  // L21 row i: x_j = (a_j - sum_{q<j} x_q U_qj) / U_jj
        const int i = c + w + blockIdx.x * TT + tid;
        if (i >= F) return;
#pragma unroll
        for (int j = 0; j < NB; j++) v[j] = j < w ? A[(size_t)(c + j) * F + i] : (T)0;
#pragma unroll
        for (int j = 0; j < NB; j++) {
            if (j < w) {
                T a = v[j];
#pragma unroll
                for (int q = 0; q < j; q++) a = fma(-v[q], D[j * NB + q], a);
                v[j] = a / D[j * NB + j];
            }
        }
#pragma unroll
        for (int j = 0; j < NB; j++)
            if (j < w) A[(size_t)(c + j) * F + i] = v[j];
    } else {  // U12 column cc: y_j = a_j - sum_{q<j} L_jq y_q
        const int cc = c + w + (blockIdx.x - nblk) * TT + tid;
        if (cc >= F) return;
        T* col = A + (size_t)cc * F + c;
#pragma unroll
        for (int j = 0; j < NB; j++) v[j] = j < w ? col[j] : (T)0;
#pragma unroll
        for (int j = 0; j < NB; j++) {
            if (j < w) {
                T a = v[j];
#pragma unroll
                for (int q = 0; q < j; q++) a = fma(-D[q * NB + j], v[q], a);
                v[j] = a;
            }
        }
#pragma unroll
        for (int j = 0; j < NB; j++)
            if (j < w) col[j] = v[j];
    }
}

// Large fronts, trailing update of panel [c, c + w): A22 -= L21 U12, 64 x 64 tiles, 4 x 4 per thread.
// grid.x: tiles of the largest trailing block of the launch, grid.y: the front list.
template <typename T>
__global__ void __launch_bounds__(TPB) k_gemm(Fronts S, const int* __restrict__ list, int panel, T* __restrict__ pool) {
    constexpr int TM = 64, KC = 16;
    __shared__ T As[KC][TM + 1];
    __shared__ T Bs[KC][TM + 1];
    const int s = list[blockIdx.y];
    const int k = fk(S, s), F = fF(S, s);
    const int c = panel * NB;
    if (c >= k) return;
    const int w = min(NB, k - c);
    const int r0 = c + w, rest = F - r0;
    const int tiles = (rest + TM - 1) / TM;
    if ((int)blockIdx.x >= tiles * tiles) return;
    const int ti = blockIdx.x % tiles, tj = blockIdx.x / tiles;
    const int i0 = r0 + ti * TM, j0 = r0 + tj * TM;
    T* A = pool + S.fo[s];
    const int tid = threadIdx.x;
    const int tr = (tid % 16) * 4, tc = (tid / 16) * 4;
    T acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = (T)0;
    for (int q0 = 0; q0 < w; q0 += KC) {
        for (int q = tid; q < KC * TM; q += TPB) {
            {  // L21(gi, gk): consecutive threads read consecutive rows of one column
                const int kk = q / TM, ii = q % TM;
                const int gi = i0 + ii, gk = c + q0 + kk;
                As[kk][ii] = (gi < F && q0 + kk < w) ? A[(size_t)gk * F + gi] : (T)0;
            }
            {  // U12(gk, gj): consecutive threads read consecutive rows of one column
                const int kk = q % KC, jj = q / KC;
                const int gj = j0 + jj, gk = c + q0 + kk;
                Bs[kk][jj] = (gj < F && q0 + kk < w) ? A[(size_t)gj * F + gk] : (T)0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KC; kk++) {
            T a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                a[u] = As[kk][tr + u];
                b[u] = Bs[kk][tc + u];
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) acc[u][v] = fma(a[u], b[v], acc[u][v]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int gj = j0 + tc + v;
        if (gj >= F) continue;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int gi = i0 + tr + u;
            if (gi < F) A[(size_t)gj * F + gi] -= acc[u][v];
        }
    }
}

// ---------------------------------------------------------------------------------------------------
// Sweeps: one workgroup per front, fronts of a level per launch.  y is the internal right-hand side
// (then the forward solution, then the backward one); w holds each front's sweep vector.
// ---------------------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(TPB) k_forward(Fronts S, const int* __restrict__ list, const T* __restrict__ pool,
                                                 T* __restrict__ y, T* __restrict__ wbuf) {
    const int s = list[blockIdx.x];
    const int k = fk(S, s), F = fF(S, s), f = S.first[s];
    const T* A = pool + S.fo[s];
    T* w = wbuf + S.w_off[s];
    const int tid = threadIdx.x;
    for (int i = tid; i < F; i += TPB) w[i] = i < k ? y[f + i] : (T)0;
    __syncthreads();
    for (int c = S.ch_ptr[s]; c < S.ch_ptr[s + 1]; c++) {
        const int ch = S.ch[c];
        const int kc = fk(S, ch), mc = fF(S, ch) - kc;
        const int* mp = S.x_map + S.x_ptr[ch];
        const T* wc = wbuf + S.w_off[ch] + kc;
        for (int q = tid; q < mc; q += TPB) w[mp[q]] += wc[q];
        __syncthreads();
    }
    // thread i owns row i (a small front has at most TPB rows) and loads its entries of the next eight
    // pivots' columns at once
    for (int j0 = 0; j0 < k; j0 += 8) {
        T a[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int j = j0 + u;
            a[u] = (j < k && tid > j && tid < F) ? A[(size_t)j * F + tid] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int j = j0 + u;
            if (j < k) {
                const T wj = w[j];
                if (tid > j && tid < F) w[tid] = w[tid] - a[u] * wj;
                __syncthreads();
            }
        }
    }
    for (int j = tid; j < k; j += TPB) y[f + j] = w[j];
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_backward(Fronts S, const int* __restrict__ list, const T* __restrict__ pool,
                                                  T* __restrict__ y, T* __restrict__ wbuf) {
    const int s = list[blockIdx.x];
    const int k = fk(S, s), F = fF(S, s), f = S.first[s];
    const T* A = pool + S.fo[s];
    T* w = wbuf + S.w_off[s];
    const int* R = S.r_idx + S.r_ptr[s];
    const int tid = threadIdx.x;
    // w = y - U12 x_R (the ancestors' solution is final), eight columns' loads in flight at a time,
    // subtracted in column order
    for (int j = tid; j < k; j += TPB) {
        T acc = y[f + j];
        int c = k;
        for (; c + 7 < F; c += 8) {
            int ri[8];
            T a[8], yv[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                ri[u] = R[c + u - k];
                a[u] = A[(size_t)(c + u) * F + j];
            }
#pragma unroll
            for (int u = 0; u < 8; u++) yv[u] = y[ri[u]];
#pragma unroll
            for (int u = 0; u < 8; u++) acc = acc - a[u] * yv[u];
        }
        for (; c < F; c++) acc = acc - A[(size_t)c * F + j] * y[R[c - k]];
        w[j] = acc;
    }
    __syncthreads();
    // thread i owns row i and loads its entries of the next eight pivots' columns at once (thread j
    // the pivot itself)
    for (int j0 = k - 1; j0 >= 0; j0 -= 8) {
        T a[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int j = j0 - u;
            a[u] = (j >= 0 && tid <= j) ? A[(size_t)j * F + tid] : (T)1;
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int j = j0 - u;
            if (j >= 0) {
                if (tid == j) w[j] = w[j] / a[u];
                __syncthreads();
                const T xj = w[j];
                if (tid < j) w[tid] = w[tid] - a[u] * xj;
                __syncthreads();
            }
        }
    }
    for (int j = tid; j < k; j += TPB) y[f + j] = w[j];
}

// acc - sum_q a[q F] x[q] over the panel's w columns, in q order; every load is issued before the first
// use, so a thread waits for memory once per panel.
template <typename T>
__device__ inline T panel_dot(const T* __restrict__ a, int F, int w, const T* x, T acc) {
    T av[NB];
#pragma unroll
    for (int q = 0; q < NB; q++) av[q] = q < w ? a[(size_t)q * F] : (T)0;
#pragma unroll
    for (int q = 0; q < NB; q++)
        if (q < w) acc = acc - av[q] * x[q];
    return acc;
}

// Sweeps of large fronts, the fronts of a level per launch (grid.y), panel by panel.
// Forward: panel [c, c + w) solves L11 x = w[c, c + w) (unit lower) in one wavefront, lane i holding x_i
// and row i of L11 in registers, x_j passed from lane j through the scalar unit; block 0 stores x in y,
// every block updates its rows below: w_i -= L(i, c..c+w) x.
template <typename T>
__global__ void __launch_bounds__(TPB) k_fwd_panel(Fronts S, const int* __restrict__ list, int panel,
                                                   const T* __restrict__ pool, T* __restrict__ y, T* __restrict__ wbuf) {
    __shared__ T xp[NB];
    const int s = list[blockIdx.y];
    const int k = fk(S, s), F = fF(S, s), f = S.first[s];
    const int c = panel * NB;
    if (c >= k) return;
    const int w = min(NB, k - c);
    const int r0 = c + w;
    if ((int)blockIdx.x > 0 && (int)blockIdx.x * TPB >= F - r0) return;
    const T* A = pool + S.fo[s];
    T* wv = wbuf + S.w_off[s];
    const int tid = threadIdx.x;
    if (tid < 64) {
        const bool own = tid < w;
        T lrow[NB];
#pragma unroll
        for (int j = 0; j < NB; j++) lrow[j] = (own && j < w) ? A[(size_t)(c + j) * F + c + tid] : (T)0;
        T x = own ? wv[c + tid] : (T)0;
#pragma unroll
        for (int j = 0; j < NB; j++) {
            if (j < w) {
                const T xj = lane_value(x, j);
                if (tid > j && own) x = x - lrow[j] * xj;
            }
        }
        if (own) xp[tid] = x;
    }
    __syncthreads();
    if (blockIdx.x == 0 && tid < w) y[f + c + tid] = xp[tid];
    const int i = r0 + blockIdx.x * TPB + tid;
    if (i >= F) return;
    wv[i] = panel_dot(A + (size_t)c * F + i, F, w, xp, wv[i]);
}

// Forward of large fronts: the sweep vector from the own right-hand side and the children's updates.
template <typename T>
__global__ void __launch_bounds__(TPB) k_fwd_init(Fronts S, const int* __restrict__ list, const T* __restrict__ y,
                                                  T* __restrict__ wbuf) {
    const int s = list[blockIdx.x];
    const int k = fk(S, s), F = fF(S, s), f = S.first[s];
    T* w = wbuf + S.w_off[s];
    const int tid = threadIdx.x;
    for (int i = tid; i < F; i += TPB) w[i] = i < k ? y[f + i] : (T)0;
    __syncthreads();
    for (int cc = S.ch_ptr[s]; cc < S.ch_ptr[s + 1]; cc++) {
        const int ch = S.ch[cc];
        const int kc = fk(S, ch), mc = fF(S, ch) - kc;
        const int* mp = S.x_map + S.x_ptr[ch];
        const T* wc = wbuf + S.w_off[ch] + kc;
        for (int q = tid; q < mc; q += TPB) w[mp[q]] += wc[q];
        __syncthreads();
    }
}

// Backward of large fronts: w_j = y_j - U12(j, :) x_R for the front's pivots.  A block of 1024 threads
// takes 64 pivots (grid.x): lane = pivot, its 16 wavefronts stride over the columns of R, and the 16
// partial sums are added in wavefront order.
template <typename T>
__global__ void __launch_bounds__(1024) k_bwd_init(Fronts S, const int* __restrict__ list, const T* __restrict__ pool,
                                                   const T* __restrict__ y, T* __restrict__ wbuf) {
    __shared__ T part[16][64];
    const int s = list[blockIdx.y];
    const int k = fk(S, s), F = fF(S, s), f = S.first[s];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + lane;
    if ((int)blockIdx.x * 64 >= k) return;
    const T* A = pool + S.fo[s];
    const int* R = S.r_idx + S.r_ptr[s];
    T acc = (T)0;
    if (j < k) {
        int c = k + wv;
        for (; c + 16 * 7 < F; c += 16 * 8) {  // eight columns' loads in flight, added in column order
            int ri[8];
            T a[8], yv[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                ri[u] = R[c + 16 * u - k];
                a[u] = A[(size_t)(c + 16 * u) * F + j];
            }
#pragma unroll
            for (int u = 0; u < 8; u++) yv[u] = y[ri[u]];
#pragma unroll
            for (int u = 0; u < 8; u++) acc = acc + a[u] * yv[u];
        }
        for (; c < F; c += 16) acc = acc + A[(size_t)c * F + j] * y[R[c - k]];
    }
    part[wv][lane] = acc;
    __syncthreads();
    if (wv == 0 && j < k) {
        T sum = part[0][lane];
        for (int q = 1; q < 16; q++) sum = sum + part[q][lane];
        wbuf[S.w_off[s] + j] = y[f + j] - sum;
    }
}

// Backward panel [c, c + w), panels from the last: U11 x = w[c, c + w) in one wavefront (lane i holding
// x_i and row i of U11, as forward), block 0 stores x in y, every block updates its rows above:
// w_i -= U(i, c..c+w) x.
template <typename T>
__global__ void __launch_bounds__(TPB) k_bwd_panel(Fronts S, const int* __restrict__ list, int panel,
                                                   const T* __restrict__ pool, T* __restrict__ y, T* __restrict__ wbuf) {
    __shared__ T xp[NB];
    const int s = list[blockIdx.y];
    const int k = fk(S, s), F = fF(S, s), f = S.first[s];
    const int c = panel * NB;
    if (c >= k) return;
    const int w = min(NB, k - c);
    if ((int)blockIdx.x > 0 && (int)blockIdx.x * TPB >= c) return;
    const T* A = pool + S.fo[s];
    T* wv = wbuf + S.w_off[s];
    const int tid = threadIdx.x;
    if (tid < 64) {
        const bool own = tid < w;
        T urow[NB];
#pragma unroll
        for (int j = 0; j < NB; j++) urow[j] = (own && j < w) ? A[(size_t)(c + j) * F + c + tid] : (T)0;
        T x = own ? wv[c + tid] : (T)0;
#pragma unroll
        for (int j = NB - 1; j >= 0; j--) {
            if (j < w) {
                if (tid == j) x = x / urow[j];
                const T xj = lane_value(x, j);
                if (tid < j) x = x - urow[j] * xj;
            }
        }
        if (own) xp[tid] = x;
    }
    __syncthreads();
    if (blockIdx.x == 0 && tid < w) y[f + c + tid] = xp[tid];
    const int i = blockIdx.x * TPB + tid;
    if (i >= c) return;
    wv[i] = panel_dot(A + (size_t)c * F + i, F, w, xp, wv[i]);
}

// internal right-hand side from the FP64 residual (rows in pivot order), and dx back in external order
template <typename T>
__global__ void k_permute_in(int n, const int* __restrict__ prow, const double* __restrict__ res, T* __restrict__ y) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < n) y[a] = (T)res[prow[a]];
}
template <typename T>
__global__ void k_permute_out(int n, const int* __restrict__ perm, const T* __restrict__ y, double* __restrict__ dx) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < n) dx[perm[b]] = (double)y[b];
}

// max |dx| over the non-NaN entries (the reference's fold with f64::max), as ordered bit patterns;
// *nonfinite = 1 when some entry is NaN or infinite
__global__ void k_maxabs(int n, const double* __restrict__ dx, unsigned long long* out, int* nonfinite) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const double a = fabs(dx[i]);
        if (a == a) v = a;
        if (!(a <= 1.7976931348623157e308)) atomicOr(nonfinite, 1);
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double u = __shfl_xor(v, o, 64);
        v = u > v ? u : v;
    }
    if ((threadIdx.x & 63) == 0 && v > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(v));
}
// clamp and update (analysis.rs:196-207); flags[0] |= 4 when some |dx| > reltol after the clamp.  A
// solve with reused factors whose dx is not finite is discarded whole (x unchanged, not converged): the
// iteration is repeated with fresh factors.
__global__ void k_update(int n, const double* __restrict__ dx, const unsigned long long* mxbits, double reltol,
                         double* __restrict__ x, int* flags, int reused, const int* __restrict__ nonfinite) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (reused && *nonfinite) {
        if (i == 0) atomicOr(flags, 4);
        return;
    }
    const double mx = __longlong_as_double((long long)*mxbits);
    double d = dx[i];
    if (mx > 1.0) d = d * 1.0 / mx;
    if (fabs(d) > reltol) atomicOr(flags, 4);
    x[i] = x[i] + d;
}

// commit (comps/*: guess -> op) at the accepted solution
__global__ void k_commit(EvalArgs E, double* c_q) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < E.n_m1) {
        if (!E.m1_caps[t]) return;
        for (int f = 0; f < M1S; f++) E.m1_state[(size_t)f * E.n_m1 + t] = E.m1_state[(size_t)(M1S + f) * E.n_m1 + t];
        for (int f = 0; f < 2; f++) E.m1_icur[(size_t)f * E.n_m1 + t] = E.m1_icur[(size_t)(2 + f) * E.n_m1 + t];
        return;
    }
    if (t < E.n_m1 + E.n_d) {
        const int j = t - E.n_m1;
        E.d_state[2 * E.n_d + j] = E.d_state[1 * E.n_d + j];
        E.d_icur[j] = E.d_icur[E.n_d + j];
        return;
    }
    if (t < E.n_m1 + E.n_d + E.n_lin) {
        const int j = t - E.n_m1 - E.n_d;
        if (E.l_type[j] != S21_COMP_C) return;
        const int* nd = E.l_node + 3 * j;
        c_q[j] = E.l_val[j] * (xval(E.x, nd[0]) - xval(E.x, nd[1]));  // guess.q = c * vd
        E.c_icur[j] = E.c_icur[E.n_lin + j];
    }
}

// Local truncation error of a trapezoidal step from the third divided difference DD3 of the node voltages
// over the last four accepted points (x, h1, h2, h3 at t0 > t1 > t2 > t3): the error is dt^3 / 12 times the
// third derivative, which is about 6 DD3.  *rmax = max over voltage unknowns of that error over
// trtol (reltol max(|x|, |h1|) + abstol).
__global__ void k_lte(int n, const int* __restrict__ kind, const double* __restrict__ x, const double* __restrict__ h1,
                      const double* __restrict__ h2, const double* __restrict__ h3, double t0, double t1, double t2,
                      double t3, double reltol, double abstol, double trtol, unsigned long long* rmax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double r = 0.0;
    if (i < n && kind[i] == S21_VAR_V) {
        const double d01 = (x[i] - h1[i]) / (t0 - t1), d12 = (h1[i] - h2[i]) / (t1 - t2),
                     d23 = (h2[i] - h3[i]) / (t2 - t3);
        const double d012 = (d01 - d12) / (t0 - t2), d123 = (d12 - d23) / (t1 - t3);
        const double d3 = (d012 - d123) / (t0 - t3);
        const double dt = t0 - t1;
        const double lte = dt * dt * dt / 2.0 * fabs(d3);
        const double tol = trtol * (reltol * fmax(fabs(x[i]), fabs(h1[i])) + abstol);
        r = lte / tol;
        if (!(r == r)) r = 1e300;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double u = __shfl_xor(r, o, 64);
        r = u > r ? u : r;
    }
    if ((threadIdx.x & 63) == 0 && r > 0.0) atomicMax(rmax, (unsigned long long)__double_as_longlong(r));
}

// out = a + w (b - a): an output-grid point between two accepted points
__global__ void k_interp(int n, const double* __restrict__ a, const double* __restrict__ b, double w,
                         double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] + w * (b[i] - a[i]);
}

}  // namespace lg
}  // namespace s21

// ===================================================================================================
// Host runtime
// ===================================================================================================
namespace s21 {
namespace lg {

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    Buf() {}
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t b) {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = b;
        if (b) LCHECK(hipMalloc(&p, b));
    }
    template <typename V>
    V* as() const {
        return (V*)p;
    }
};

template <typename V>
static void upload(Buf& b, const std::vector<V>& v) {
    b.alloc(std::max<size_t>(v.size(), 1) * sizeof(V));
    if (!v.empty()) LCHECK(hipMemcpy(b.p, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice));
}

// Launch lists of one elimination-tree level.
struct Level {
    int small0 = 0, n_small = 0;  // range in the small-front list
    int small_F = 0;              // largest small front of the level
    int big0 = 0, n_big = 0;      // range in the big-front list
    int job0 = 0, n_job = 0;      // range in the big-front assembly jobs
    int panels = 0;               // most panels of a big front of the level
    int trsm_blocks = 0, gemm_tiles = 0;  // per panel 0 (the largest; later panels need fewer)
    int all0 = 0, n_all = 0;      // range in the all-fronts list (sweeps)
};

}  // namespace lg
}  // namespace s21

struct s21_large {
    // the netlist
    int n_vars = 0;
    std::vector<std::string> names;
    std::vector<int> kinds;
    std::vector<s21_comp> comps;
    std::vector<s21_params> mos1_models, mos1_insts, diode_models, diode_insts, bsim4_models, bsim4_insts;
    s21_options opts;
    int precision = S21_FP64;
    int device = 0;
    double relax = 0.1;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // the structure of the last analysis
    std::string prepared;
    s21::Topology t;
    s21::LargeSym S;
    int n_m1 = 0, n_d = 0, n_lin = 0, n_b4 = 0, n_cells = 0, n_elem = 0;
    std::vector<s21::lg::Level> levels;
    std::vector<int> big_host;  // the big-front launch list (host copy)
    int reuse = -1;             // factor reuse: 1 on, 0 off (factor every iteration), -1 = on in mixed precision
    bool has_bsim4 = false;     // BSIM4 devices present: no factor reuse
    bool have_factors = false;  // the pool holds factors of this structure
    int reused = 0;             // iterations since the last factorisation
    s21::lg::Buf m1_node, m1_par, m1_caps, m1_state, m1_cb, d_node, d_par, d_state, d_cb, l_node, l_type, l_val, c_q,
        l_cb, cv, g_ptr, g_idx, slots, r_ptr, r_elem, e_col, x, res, dx, flags, mxbits;
    s21::lg::Buf f_fo, f_first, f_rptr, f_ridx, f_chptr, f_ch, f_xptr, f_xmap, f_aptr, f_aelem, f_aoff, f_woff;
    s21::lg::Buf l_small, l_big, l_jobs, l_all, perm, prow, pool, yv, wv;
    s21::lg::Buf wave;
    s21::lg::Buf b4_node, b4_par, b4_st, b4_cb;
    // trapezoidal integration: per-charge currents, unknown kinds, accepted-point history
    s21::lg::Buf m1_icur, d_icur, c_icur, kinds_dev, hist[4];
    int trap = 0;
    int64_t accepted = 0, rejected = 0;
    // results
    int status = S21_OK;
    int64_t iters = 0, npts = 0, factorisations = 0;
    std::vector<double> axis;
    double ms_total = 0, ms_factor = 0, ms_solve = 0, ms_eval = 0;
    ~s21_large() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace s21 {
namespace lg {

inline int64_t ekey(int r, int c, int n) { return (int64_t)r * (int64_t)(n + 1) + c; }

// Stamp entries of one component and the slot each reaches: a matrix element, n_elem + row for an RHS
// entry, or -1 (ground, or an entry the analysis never stamps).
static int entry_targets(const Comp& c, const std::unordered_map<int64_t, int>& elem, int n, int n_elem, int* tgt) {
    const int* nd = c.node;
    auto M = [&](int r, int col) -> int {
        if (r < 0 || col < 0) return -1;
        return elem.at(ekey(r, col, n));
    };
    auto B = [&](int r) { return r < 0 ? -1 : n_elem + r; };
    switch (c.type) {
    case S21_COMP_MOS1:
        for (int e = 0; e < MOS1_NE; e++) {
            const int r = nd[mos1_entry_row(e)];
            tgt[e] = mos1_entry_col(e) < 0 ? B(r) : M(r, nd[mos1_entry_col(e)]);
        }
        return MOS1_NE;
    case S21_COMP_DIODE:
        for (int e = 0; e < DIODE_NE; e++) {
            const int r = nd[diode_entry_row(e)];
            tgt[e] = diode_entry_col(e) < 0 ? B(r) : M(r, nd[diode_entry_col(e)]);
        }
        return DIODE_NE;
    case S21_COMP_R:
    case S21_COMP_C:
        tgt[0] = M(nd[0], nd[0]);
        tgt[1] = M(nd[1], nd[1]);
        tgt[2] = M(nd[0], nd[1]);
        tgt[3] = M(nd[1], nd[0]);
        if (c.type == S21_COMP_R) return R_NE;
        tgt[4] = B(nd[0]);
        tgt[5] = B(nd[1]);
        return C_NE;
    case S21_COMP_V:
        tgt[0] = M(nd[0], nd[2]);
        tgt[1] = M(nd[2], nd[0]);
        tgt[2] = M(nd[1], nd[2]);
        tgt[3] = M(nd[2], nd[1]);
        tgt[4] = B(nd[2]);
        return V_NE;
    case S21_COMP_I:
        tgt[0] = B(nd[0]);
        tgt[1] = B(nd[1]);
        return I_NE;
    case S21_COMP_BSIM4:
        for (int k = 0; k < b4::B4_NS; k++) {
            const int r = nd[b4::b4_row(k)];
            tgt[k] = b4::b4_col(k) < 0 ? B(r) : M(r, nd[b4::b4_col(k)]);
        }
        return b4::B4_NS;
    default: fail(S21_INVALID, "the large-instance path takes R, C, V, I, Mos1, Diode and BSIM4 components");
    }
}

template <typename T>
static void set_lds_limits() {
    LCHECK(hipFuncSetAttribute((const void*)k_small<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)SMALL_LDS));
}

static Fronts fronts_view(const s21_large* L) {
    Fronts F;
    F.ns = L->S.ns;
    F.fo = L->f_fo.as<int64_t>();
    F.first = L->f_first.as<int>();
    F.r_ptr = L->f_rptr.as<int64_t>();
    F.r_idx = L->f_ridx.as<int>();
    F.ch_ptr = L->f_chptr.as<int>();
    F.ch = L->f_ch.as<int>();
    F.x_ptr = L->f_xptr.as<int64_t>();
    F.x_map = L->f_xmap.as<int>();
    F.a_ptr = L->f_aptr.as<int64_t>();
    F.a_elem = L->f_aelem.as<int>();
    F.a_off = L->f_aoff.as<int>();
    F.w_off = L->f_woff.as<int64_t>();
    return F;
}

static EvalArgs eval_args(const s21_large* L, bool tran, double dt) {
    EvalArgs E;
    E.n_m1 = L->n_m1;
    E.n_d = L->n_d;
    E.n_lin = L->n_lin;
    E.tran = tran ? 1 : 0;
    E.gmin = L->opts.gmin;
    E.dt = dt;
    E.x = L->x.as<double>();
    E.m1_node = L->m1_node.as<int>();
    E.m1_par = L->m1_par.as<double>();
    E.m1_caps = L->m1_caps.as<int>();
    E.m1_state = L->m1_state.as<double>();
    E.m1_cb = L->m1_cb.as<int>();
    E.d_node = L->d_node.as<int>();
    E.d_par = L->d_par.as<double>();
    E.d_state = L->d_state.as<double>();
    E.d_cb = L->d_cb.as<int>();
    E.l_node = L->l_node.as<int>();
    E.l_type = L->l_type.as<int>();
    E.l_val = L->l_val.as<double>();
    E.c_q = L->c_q.as<double>();
    E.l_cb = L->l_cb.as<int>();
    E.cv = L->cv.as<double>();
    E.trap = L->trap;
    E.m1_icur = L->m1_icur.as<double>();
    E.d_icur = L->d_icur.as<double>();
    E.c_icur = L->c_icur.as<double>();
    return E;
}

// The structure of one analysis: components (IC forcing for a transient), derived parameters, the
// symbolic phase on the first iteration's matrix, gather lists and front maps, all uploaded.
static void prepare(s21_large* L, const s21_tran_options* tr) {
    Topology& t = L->t;
    t = Topology();
    build_comps(L->n_vars, L->names, L->kinds, L->comps, (int)L->mos1_models.size(), (int)L->mos1_insts.size(),
                (int)L->diode_models.size(), (int)L->diode_insts.size(), tr, t);
    topo_elements(t);
    const int n = t.n;
    const int ne = (int)t.erow.size();
    L->n_elem = ne;
    const bool tran = tr != nullptr;
    // devices by class, in component order; derived parameters
    HostInstance h0;
    std::vector<int> m1_node, d_node, l_node, l_type, m1_caps;
    std::vector<double> m1p, dpar, lval0, lval1;
    std::vector<int> m1_comp, d_comp, l_comp, b4_comp, b4_node;
    for (size_t ci = 0; ci < t.comps.size(); ci++) {
        const Comp& c = t.comps[ci];
        if (c.type == S21_COMP_BSIM4) {
            if (c.model < 0 || c.model >= (int)L->bsim4_models.size() || c.inst < 0 ||
                c.inst >= (int)L->bsim4_insts.size())
                fail(S21_ERR, "Parameters not defined");
            try {
                h0.b4.push_back(b4::derive_params(L->bsim4_models[c.model], L->bsim4_insts[c.inst]));
            } catch (const std::domain_error& e) {  // the reference panics here (bsim4inst.rs:200-223)
                fail(S21_ERR, e.what());
            }
            b4_comp.push_back((int)ci);
            for (int r = 0; r < 4; r++) b4_node.push_back(c.node[r]);
        } else if (c.type == S21_COMP_MOS1) {
            h0.m1.push_back(derive_mos1(L->mos1_models[c.model], L->mos1_insts[c.inst], L->opts));
            m1_comp.push_back((int)ci);
        } else if (c.type == S21_COMP_DIODE) {
            h0.dio.push_back(derive_diode(L->diode_models[c.model], L->diode_insts[c.inst], L->opts));
            d_comp.push_back((int)ci);
        } else if (c.type == S21_COMP_MOS0) {
            fail(S21_INVALID, "the large-instance path takes R, C, V, I, Mos1, Diode and BSIM4 components");
        } else {
            h0.lin_value.push_back(c.value);
            l_comp.push_back((int)ci);
        }
    }
    const int nm = (int)m1_comp.size(), nd = (int)d_comp.size(), nl = (int)l_comp.size(), nb4 = (int)b4_comp.size();
    L->n_b4 = nb4;
    L->n_m1 = nm;
    L->n_d = nd;
    L->n_lin = nl;
    m1p.assign((size_t)(M1F + M1C) * nm, 0.0);
    m1_caps.assign(nm, 0);
    for (int j = 0; j < nm; j++) {
        const Comp& c = t.comps[m1_comp[j]];
        for (int r = 0; r < 6; r++) m1_node.push_back(c.node[r]);
        const Mos1Derived& D = h0.m1[j];
        const double core[M1F] = {D.core.p,      D.core.vt0_t,  D.core.beta,   D.core.phi_t,
                                  D.core.gamma,  D.core.lambda, D.core.vtherm, D.core.isat_d,
                                  D.core.isat_s, D.core.grd,    D.core.grs};
        for (int f = 0; f < M1F; f++) m1p[(size_t)f * nm + j] = core[f];
        const double caps[M1C] = {D.caps.cox, D.caps.cgs_ov, D.caps.cgd_ov, D.caps.cgb_ov};
        for (int f = 0; f < M1C; f++) m1p[(size_t)(M1F + f) * nm + j] = caps[f];
        m1_caps[j] = D.has_caps && tran ? 1 : 0;
    }
    dpar.assign((size_t)DPF * nd, 0.0);
    for (int j = 0; j < nd; j++) {
        const Comp& c = t.comps[d_comp[j]];
        for (int r = 0; r < 3; r++) d_node.push_back(c.node[r]);
        const DiodeP& P = h0.dio[j];
        const double v[DPF] = {P.isat, P.vte, P.vcrit, P.gspr, P.bv, P.has_bv, P.cz,
                               P.cz2,  P.dep, P.f1,    P.f3,   P.tt, P.vj,     P.m};
        for (int f = 0; f < DPF; f++) dpar[(size_t)f * nd + j] = v[f];
    }
    lval0.resize(nl);
    lval1.resize(nl);
    for (int j = 0; j < nl; j++) {
        const Comp& c = t.comps[l_comp[j]];
        for (int r = 0; r < 3; r++) l_node.push_back(c.type == S21_COMP_V || r < 2 ? c.node[r] : -1);
        l_type.push_back(c.type);
        lval0[j] = c.value;
        lval1[j] = c.value_tran;
    }
    std::vector<double> lval(lval0);
    lval.insert(lval.end(), lval1.begin(), lval1.end());
    // contribution cells and the ordered gather lists (component order, then entry order)
    std::unordered_map<int64_t, int> elem;
    elem.reserve((size_t)ne * 2);
    for (int e = 0; e < ne; e++) elem[ekey(t.erow[e], t.ecol[e], n)] = e;
    std::vector<int> m1_cb(nm), d_cb(nd), l_cb(nl), b4_cb(nb4), comp_cb(t.comps.size(), -1);
    int cells = 0;
    for (int j = 0; j < nb4; j++) {
        b4_cb[j] = cells;
        comp_cb[b4_comp[j]] = cells;
        cells += b4::B4_NS;
    }
    for (int j = 0; j < nm; j++) {
        m1_cb[j] = cells;
        comp_cb[m1_comp[j]] = cells;
        cells += MOS1_NE;
    }
    for (int j = 0; j < nd; j++) {
        d_cb[j] = cells;
        comp_cb[d_comp[j]] = cells;
        cells += DIODE_NE;
    }
    for (int j = 0; j < nl; j++) {
        l_cb[j] = cells;
        comp_cb[l_comp[j]] = cells;
        cells += C_NE;  // the widest linear element
    }
    L->n_cells = cells;
    const int nslot = ne + n;
    std::vector<int> cnt(nslot + 1, 0);
    std::vector<std::pair<int, int>> contrib;  // (slot, cell) in component order
    int tg[b4::B4_NS > MOS1_NE ? b4::B4_NS : MOS1_NE];
    for (size_t ci = 0; ci < t.comps.size(); ci++) {
        const int k = entry_targets(t.comps[ci], elem, n, ne, tg);
        for (int e = 0; e < k; e++)
            if (tg[e] >= 0) contrib.push_back({tg[e], comp_cb[ci] + e});
    }
    for (auto& pc : contrib) cnt[pc.first + 1]++;
    for (int sl = 0; sl < nslot; sl++) cnt[sl + 1] += cnt[sl];
    std::vector<int> gidx(contrib.size()), fill(cnt.begin(), cnt.end() - 1);
    for (auto& pc : contrib) gidx[fill[pc.first]++] = pc.second;
    // residual rows: elements of each row in creation order
    std::vector<int> rptr(n + 1, 0), relem(ne);
    for (int e = 0; e < ne; e++) rptr[t.erow[e] + 1]++;
    for (int r = 0; r < n; r++) rptr[r + 1] += rptr[r];
    {
        std::vector<int> f2(rptr.begin(), rptr.end() - 1);
        for (int e = 0; e < ne; e++) relem[f2[t.erow[e]]++] = e;
    }
    // symbolic phase on the first Newton iteration's matrix (x = 0)
    std::vector<double> vals, rhs;
    host_assemble_op(t, h0, L->opts, std::vector<double>(n, 0.0), vals, rhs);
    large_symbolic(n, t.erow, t.ecol, &vals, L->S, L->relax);
    const LargeSym& S = L->S;
    // A's entries by front
    std::vector<int64_t> aptr(S.ns + 1, 0);
    std::vector<int> a_s(ne), a_loc(ne);
    for (int e = 0; e < ne; e++) {
        const int a = S.irow[t.erow[e]], b = S.icol[t.ecol[e]];
        const int lo = std::min(a, b);
        const int s = (int)(std::upper_bound(S.sn_first.begin(), S.sn_first.end(), lo) - S.sn_first.begin()) - 1;
        const int64_t o = large_entry_offset(S, a, b);
        if (o < 0) fail(S21_ERR, "matrix entry outside the symbolic structure");
        a_s[e] = s;
        a_loc[e] = (int)(o - S.front_off[s]);
        aptr[s + 1]++;
    }
    for (int s = 0; s < S.ns; s++) aptr[s + 1] += aptr[s];
    std::vector<int> aelem(ne), aoff(ne);
    {
        std::vector<int64_t> f2(aptr.begin(), aptr.end() - 1);
        for (int e = 0; e < ne; e++) {
            const int64_t q = f2[a_s[e]]++;
            aelem[q] = e;
            aoff[q] = a_loc[e];
        }
    }
    std::vector<int64_t> woff(S.ns + 1, 0);
    std::vector<int> Fs(S.ns), Ks(S.ns);
    for (int s = 0; s < S.ns; s++) {
        Ks[s] = S.sn_first[s + 1] - S.sn_first[s];
        Fs[s] = Ks[s] + (int)(S.r_ptr[s + 1] - S.r_ptr[s]);
        woff[s + 1] = woff[s] + Fs[s];
    }
    // level launch lists
    const int smax = L->precision == S21_MIXED ? small_max<float>() : small_max<double>();
    std::vector<int> lsmall, lbig, lall;
    std::vector<int2> jobs;
    L->levels.clear();
    for (size_t l = 0; l + 1 < S.lvl_ptr.size(); l++) {
        Level lv;
        lv.small0 = (int)lsmall.size();
        lv.big0 = (int)lbig.size();
        lv.job0 = (int)jobs.size();
        lv.all0 = (int)lall.size();
        for (int q = S.lvl_ptr[l]; q < S.lvl_ptr[l + 1]; q++) {
            const int s = S.lvl_sn[q];
            lall.push_back(s);
            if (Fs[s] <= smax) {
                lsmall.push_back(s);
                lv.small_F = std::max(lv.small_F, Fs[s]);
            } else {
                lbig.push_back(s);
                for (int cb = 0; cb * CB < Fs[s]; cb++) jobs.push_back(make_int2(s, cb));
                lv.panels = std::max(lv.panels, (Ks[s] + NB - 1) / NB);
            }
        }
        lv.n_small = (int)lsmall.size() - lv.small0;
        lv.n_big = (int)lbig.size() - lv.big0;
        lv.n_job = (int)jobs.size() - lv.job0;
        lv.n_all = (int)lall.size() - lv.all0;
        L->levels.push_back(lv);
    }
    // upload
    upload(L->m1_node, m1_node);
    upload(L->m1_par, m1p);
    upload(L->m1_caps, m1_caps);
    upload(L->m1_cb, m1_cb);
    upload(L->d_node, d_node);
    upload(L->d_par, dpar);
    upload(L->d_cb, d_cb);
    upload(L->l_node, l_node);
    upload(L->l_type, l_type);
    upload(L->l_val, lval);
    upload(L->l_cb, l_cb);
    upload(L->b4_node, b4_node);
    {  // parameters field-major: field K of device j at [K][j]
        constexpr int D = (int)(sizeof(b4::B4Par) / sizeof(double));
        std::vector<double> par((size_t)D * nb4);
        for (int j = 0; j < nb4; j++)
            for (int k = 0; k < D; k++) par[(size_t)k * nb4 + j] = ((const double*)&h0.b4[j])[k];
        upload(L->b4_par, par);
    }
    upload(L->b4_cb, b4_cb);
    L->b4_st.alloc(std::max<size_t>((size_t)B4S_F * nb4, 1) * sizeof(double));
    L->m1_state.alloc(std::max<size_t>((size_t)2 * M1S * nm, 1) * sizeof(double));
    L->d_state.alloc(std::max<size_t>((size_t)DSF * nd, 1) * sizeof(double));
    L->c_q.alloc(std::max<size_t>(nl, 1) * sizeof(double));
    L->m1_icur.alloc(std::max<size_t>((size_t)4 * nm, 1) * sizeof(double));
    L->d_icur.alloc(std::max<size_t>((size_t)2 * nd, 1) * sizeof(double));
    L->c_icur.alloc(std::max<size_t>((size_t)2 * nl, 1) * sizeof(double));
    upload(L->kinds_dev, t.kinds);
    L->cv.alloc(std::max<size_t>(cells, 1) * sizeof(double));
    upload(L->g_ptr, cnt);
    upload(L->g_idx, gidx);
    L->slots.alloc((size_t)std::max(nslot, 1) * sizeof(double));
    upload(L->r_ptr, rptr);
    upload(L->r_elem, relem);
    upload(L->e_col, t.ecol);
    L->x.alloc((size_t)std::max(n, 1) * sizeof(double));
    L->res.alloc((size_t)std::max(n, 1) * sizeof(double));
    L->dx.alloc((size_t)std::max(n, 1) * sizeof(double));
    L->flags.alloc(8 * sizeof(int));  // 3 flags, the non-finite-dx flag, the residual's max |.| (8 bytes)
    L->mxbits.alloc(sizeof(unsigned long long));
    upload(L->f_fo, S.front_off);
    upload(L->f_first, S.sn_first);
    upload(L->f_rptr, S.r_ptr);
    upload(L->f_ridx, S.r_idx);
    upload(L->f_chptr, S.child_ptr);
    upload(L->f_ch, S.child);
    upload(L->f_xptr, S.ext_ptr);
    upload(L->f_xmap, S.ext_map);
    upload(L->f_aptr, aptr);
    upload(L->f_aelem, aelem);
    upload(L->f_aoff, aoff);
    upload(L->f_woff, woff);
    upload(L->l_small, lsmall);
    upload(L->l_big, lbig);
    L->big_host = lbig;
    upload(L->l_jobs, jobs);
    upload(L->l_all, lall);
    upload(L->perm, S.perm);
    upload(L->prow, S.prow);
    L->have_factors = false;
    const size_t tsz = L->precision == S21_MIXED ? sizeof(float) : sizeof(double);
    L->pool.alloc((size_t)std::max<int64_t>(S.pool, 1) * tsz);
    L->yv.alloc((size_t)std::max(n, 1) * tsz);
    L->wv.alloc((size_t)std::max<int64_t>(woff[S.ns], 1) * tsz);
    if (L->precision == S21_MIXED)
        set_lds_limits<float>();
    else
        set_lds_limits<double>();
}

static inline int blocks(int64_t n) { return (int)std::max<int64_t>((n + TPB - 1) / TPB, 1); }

template <typename T>
static void factor(s21_large* L) {
    const Fronts F = fronts_view(L);
    const double* aval = L->slots.as<double>();
    T* pool = L->pool.as<T>();
    int* flags = L->flags.as<int>();
    const int* lsmall = L->l_small.as<int>();
    const int* lbig = L->l_big.as<int>();
    const int2* jobs = L->l_jobs.as<int2>();
    const LargeSym& S = L->S;
    for (const Level& lv : L->levels) {
        if (lv.n_small) {
            const size_t lds = (size_t)lv.small_F * lv.small_F * sizeof(T);
            hipLaunchKernelGGL(k_small<T>, dim3(lv.n_small), dim3(TPB), lds, L->stream, F, lsmall + lv.small0, aval,
                               pool, flags + 1);
        }
        if (!lv.n_big) continue;
        hipLaunchKernelGGL(k_assemble<T>, dim3(lv.n_job), dim3(TPB), 0, L->stream, F, jobs + lv.job0, aval, pool);
        for (int p = 0; p < lv.panels; p++) {
            int tb = 0, tiles = 0;  // grid sizes for this panel: the largest over the level's big fronts
            for (int q = 0; q < lv.n_big; q++) {
                const int s = L->big_host[lv.big0 + q];
                const int k = S.sn_first[s + 1] - S.sn_first[s];
                const int Fs = k + (int)(S.r_ptr[s + 1] - S.r_ptr[s]);
                const int c = p * NB;
                if (c >= k) continue;
                const int w = std::min(NB, k - c), rest = Fs - c - w;
                tb = std::max(tb, 2 * ((rest + TT - 1) / TT));
                const int tl = (rest + 63) / 64;
                tiles = std::max(tiles, tl * tl);
            }
            hipLaunchKernelGGL(k_diag<T>, dim3(lv.n_big), dim3(64), 0, L->stream, F, lbig + lv.big0, p, pool, flags + 1);
            if (tb > 0)
                hipLaunchKernelGGL(k_trsm<T>, dim3(tb, lv.n_big), dim3(TT), 0, L->stream, F, lbig + lv.big0, p, pool);
            if (tiles > 0)
                hipLaunchKernelGGL(k_gemm<T>, dim3(tiles, lv.n_big), dim3(TPB), 0, L->stream, F, lbig + lv.big0, p, pool);
        }
    }
    LCHECK(hipGetLastError());
    L->factorisations++;
}

template <typename T>
static void solve(s21_large* L) {
    const Fronts F = fronts_view(L);
    const int n = L->t.n;
    T* y = L->yv.as<T>();
    T* w = L->wv.as<T>();
    const T* pool = L->pool.as<T>();
    const int* lsmall = L->l_small.as<int>();
    const int* lbig = L->l_big.as<int>();
    const LargeSym& S = L->S;
    hipLaunchKernelGGL(k_permute_in<T>, dim3(blocks(n)), dim3(TPB), 0, L->stream, n, L->prow.as<int>(),
                       L->res.as<double>(), y);
    // grid sizes of a big-front panel launch: blocks of rows below (forward) or above (backward) the panel
    auto rows_below = [&](const Level& lv, int p) {
        int nb = 1;
        for (int q = 0; q < lv.n_big; q++) {
            const int s = L->big_host[lv.big0 + q];
            const int k = S.sn_first[s + 1] - S.sn_first[s], Fs = k + (int)(S.r_ptr[s + 1] - S.r_ptr[s]);
            if (p * NB < k) nb = std::max(nb, (Fs - p * NB - std::min(NB, k - p * NB) + TPB - 1) / TPB);
        }
        return nb;
    };
    for (const Level& lv : L->levels) {
        if (lv.n_small)
            hipLaunchKernelGGL(k_forward<T>, dim3(lv.n_small), dim3(TPB), 0, L->stream, F, lsmall + lv.small0, pool, y, w);
        if (!lv.n_big) continue;
        hipLaunchKernelGGL(k_fwd_init<T>, dim3(lv.n_big), dim3(TPB), 0, L->stream, F, lbig + lv.big0, y, w);
        for (int p = 0; p < lv.panels; p++)
            hipLaunchKernelGGL(k_fwd_panel<T>, dim3(rows_below(lv, p), lv.n_big), dim3(TPB), 0, L->stream, F,
                               lbig + lv.big0, p, pool, y, w);
    }
    for (auto it = L->levels.rbegin(); it != L->levels.rend(); ++it) {
        const Level& lv = *it;
        if (lv.n_small)
            hipLaunchKernelGGL(k_backward<T>, dim3(lv.n_small), dim3(TPB), 0, L->stream, F, lsmall + lv.small0, pool, y,
                               w);
        if (!lv.n_big) continue;
        int kmax = 0;
        for (int q = 0; q < lv.n_big; q++) {
            const int s = L->big_host[lv.big0 + q];
            kmax = std::max(kmax, S.sn_first[s + 1] - S.sn_first[s]);
        }
        hipLaunchKernelGGL(k_bwd_init<T>, dim3((kmax + 63) / 64, lv.n_big), dim3(1024), 0, L->stream, F,
                           lbig + lv.big0, pool, y, w);
        for (int p = lv.panels - 1; p >= 0; p--)
            hipLaunchKernelGGL(k_bwd_panel<T>, dim3(std::max(1, (p * NB + TPB - 1) / TPB), lv.n_big), dim3(TPB), 0,
                               L->stream, F, lbig + lv.big0, p, pool, y, w);
    }
    hipLaunchKernelGGL(k_permute_out<T>, dim3(blocks(n)), dim3(TPB), 0, L->stream, n, L->perm.as<int>(), y,
                       L->dx.as<double>());
    LCHECK(hipGetLastError());
}

// Commit of every device's guess (Component::commit) at the accepted solution.
static void commit_state(s21_large* L, bool tran, double dt) {
    const EvalArgs E = eval_args(L, tran, dt);
    hipLaunchKernelGGL(k_commit, dim3(blocks(L->n_m1 + L->n_d + L->n_lin)), dim3(TPB), 0, L->stream, E,
                       L->c_q.as<double>());
    if (L->n_b4)
        hipLaunchKernelGGL(k_commit_b4, dim3(blocks(L->n_b4)), dim3(TPB), 0, L->stream, L->n_b4,
                           L->b4_st.as<double>());
    LCHECK(hipGetLastError());
}

// Evaluation, gather and residual at the current x; flags[0] set when the residual fails the test.
static void assemble_residual(s21_large* L, bool tran, double dt) {
    const int n = L->t.n, nslot = L->n_elem + n;
    const EvalArgs E = eval_args(L, tran, dt);
    hipLaunchKernelGGL(k_eval, dim3(blocks(L->n_m1 + L->n_d + L->n_lin)), dim3(TPB), 0, L->stream, E);
    if (L->n_b4) {
        B4Args A{L->n_b4,
                 tran ? 1 : 0,
                 L->opts.gmin,
                 dt,
                 L->x.as<double>(),
                 L->b4_node.as<int>(),
                 L->b4_par.as<double>(),
                 L->b4_st.as<double>(),
                 L->b4_cb.as<int>(),
                 L->cv.as<double>()};
        hipLaunchKernelGGL(k_eval_b4, dim3((L->n_b4 + 127) / 128), dim3(128), 0, L->stream, A);
    }
    hipLaunchKernelGGL(k_gather, dim3(blocks(nslot)), dim3(TPB), 0, L->stream, nslot, L->g_ptr.as<int>(),
                       L->g_idx.as<int>(), L->cv.as<double>(), L->slots.as<double>());
    LCHECK(hipMemsetAsync(L->flags.p, 0, sizeof(int), L->stream));
    LCHECK(hipMemsetAsync(L->flags.as<int>() + 4, 0, sizeof(unsigned long long), L->stream));
    hipLaunchKernelGGL(k_residual, dim3(blocks(n)), dim3(TPB), 0, L->stream, n, L->r_ptr.as<int>(),
                       L->r_elem.as<int>(), L->e_col.as<int>(), L->slots.as<double>(),
                       L->slots.as<double>() + L->n_elem, L->x.as<double>(), L->res.as<double>(), L->opts.iabstol,
                       L->flags.as<int>(), (unsigned long long*)(L->flags.as<int>() + 4));
    LCHECK(hipGetLastError());
}

// One Newton solve (Solver::solve, analysis.rs:169-210) from the current x.  Every iteration assembles
// and tests in FP64.  The FP64 path factors every iteration, as the reference does; with factor reuse on
// (the mixed path's default) an iteration keeps the FP32 factors of an earlier Jacobian -- of an earlier
// iteration or time point -- while the FP64 residual still falls at least fourfold per iteration (Newton
// as iterative refinement, converging to the same solution under the same test), and factors afresh
// otherwise, or after 32 reuses.
static int newton(s21_large* L, bool tran, double dt, bool commit = true) {
    const int n = L->t.n;
    bool prev_ok = true;  // dx starts at zero
    LCHECK(hipMemsetAsync(L->flags.p, 0, 8 * sizeof(int), L->stream));
    double res_prev = 0.0;
    bool last_reused = false;
    for (int it = 0; it < 100; it++) {  // analysis.rs:173
        L->iters++;
        assemble_residual(L, tran, dt);
        int hf[8];
        LCHECK(hipMemcpyAsync(hf, L->flags.p, sizeof(hf), hipMemcpyDeviceToHost, L->stream));
        LCHECK(hipStreamSynchronize(L->stream));
        if (hf[1]) return S21_PIVOT;  // the last factorisation met a zero pivot
        if (it > 0) prev_ok = hf[2] == 0;
        if (prev_ok && hf[0] == 0) {
            if (commit) commit_state(L, tran, dt);
            return S21_OK;
        }
        double res_now;
        {
            unsigned long long bits;
            std::memcpy(&bits, hf + 4, sizeof(bits));
            res_now = __builtin_bit_cast(double, bits);
        }
        // the first iteration of a solve has no residual of its own to compare with: it keeps the factors;
        // a reused solve that came out non-finite (hf[3]) was discarded and is repeated with fresh factors
        const bool refactor = !L->reuse || !L->have_factors || L->reused >= 32 || (hf[3] && last_reused) ||
                              (it > 0 && !(res_now <= 0.25 * res_prev));
        res_prev = res_now;
        LCHECK(hipMemsetAsync(L->flags.as<int>() + 1, 0, 3 * sizeof(int), L->stream));
        if (refactor) {
            if (L->precision == S21_MIXED)
                factor<float>(L);
            else
                factor<double>(L);
            L->have_factors = true;
            L->reused = 0;
        } else {
            L->reused++;
        }
        last_reused = !refactor;
        if (L->precision == S21_MIXED)
            solve<float>(L);
        else
            solve<double>(L);
        LCHECK(hipMemsetAsync(L->mxbits.p, 0, sizeof(unsigned long long), L->stream));
        hipLaunchKernelGGL(k_maxabs, dim3(blocks(n)), dim3(TPB), 0, L->stream, n, L->dx.as<double>(),
                           L->mxbits.as<unsigned long long>(), L->flags.as<int>() + 3);
        hipLaunchKernelGGL(k_update, dim3(blocks(n)), dim3(TPB), 0, L->stream, n, L->dx.as<double>(),
                           L->mxbits.as<unsigned long long>(), L->opts.reltol, L->x.as<double>(),
                           L->flags.as<int>() + 2, last_reused ? 1 : 0, L->flags.as<int>() + 3);
        LCHECK(hipGetLastError());
    }
    return S21_CONVERGENCE;
}

static void reset_state(s21_large* L) {
    for (Buf* b : {&L->m1_state, &L->d_state, &L->c_q, &L->x, &L->m1_icur, &L->d_icur, &L->c_icur, &L->b4_st})
        LCHECK(hipMemsetAsync(b->p, 0, b->bytes, L->stream));
    L->iters = 0;
    L->factorisations = 0;
    L->have_factors = false;  // each analysis starts from a fresh factorisation
}

}  // namespace lg
}  // namespace s21

namespace s21 {
namespace lg {

template <typename F>
static int lguard(F&& f) {
    try {
        f();
        return S21_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("out of host memory");
        return S21_ERR;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return S21_ERR;
    }
}

static void check_current_device(const s21_large* L) {
    int dev = 0;
    LCHECK(hipGetDevice(&dev));
    if (dev != L->device) fail(S21_INVALID, "the current device is not the one the instance was created on");
}

}  // namespace lg
}  // namespace s21

using namespace s21;
using namespace s21::lg;

extern "C" {

int s21_large_create(const s21_netlist* net, const s21_options* opts, int32_t precision, s21_large** out) {
    return lguard([&] {
        if (!out || !net) fail(S21_INVALID, "s21_large_create: bad arguments");
        *out = nullptr;
        const s21_netlist& nl = *net;
        if (nl.n_vars < 0 || nl.n_comps < 0 || (nl.n_comps > 0 && !nl.comps)) fail(S21_INVALID, "malformed netlist");
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n == 0) fail(S21_DEVICE, "no HIP device available");
        std::unique_ptr<s21_large> L(new s21_large());
        LCHECK(hipGetDevice(&L->device));
        hipDeviceProp_t prop;
        LCHECK(hipGetDeviceProperties(&prop, L->device));
        if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0)
            fail(S21_DEVICE, std::string("engine is built for gfx950, device is ") + prop.gcnArchName);
        L->n_vars = nl.n_vars;
        for (int k = 0; k < nl.n_vars; k++) {
            L->names.push_back(nl.var_name && nl.var_name[k] ? nl.var_name[k] : ("v" + std::to_string(k)));
            L->kinds.push_back(nl.var_kind ? nl.var_kind[k] : S21_VAR_V);
        }
        L->comps.assign(nl.comps, nl.comps + nl.n_comps);
        auto take = [](int k, const s21_params* p, std::vector<s21_params>& v) {
            if (k > 0 && !p) fail(S21_INVALID, "malformed netlist");
            v.assign(p, p + std::max(k, 0));
        };
        take(nl.n_bsim4_models, nl.bsim4_models, L->bsim4_models);
        take(nl.n_bsim4_insts, nl.bsim4_insts, L->bsim4_insts);
        take(nl.n_mos1_models, nl.mos1_models, L->mos1_models);
        take(nl.n_mos1_insts, nl.mos1_insts, L->mos1_insts);
        take(nl.n_diode_models, nl.diode_models, L->diode_models);
        take(nl.n_diode_insts, nl.diode_insts, L->diode_insts);
        if (opts)
            L->opts = *opts;
        else
            s21_default_options(&L->opts);
        L->precision = precision == S21_MIXED ? S21_MIXED : S21_FP64;
        bool bsim4 = false;
        for (const s21_comp& c : L->comps) bsim4 = bsim4 || c.type == S21_COMP_BSIM4;
        L->has_bsim4 = bsim4;
        L->reuse = L->precision == S21_MIXED && !bsim4 ? 1 : 0;
        LCHECK(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
        LCHECK(hipEventCreate(&L->ev0));
        LCHECK(hipEventCreate(&L->ev1));
        *out = L.release();
    });
}

void s21_large_free(s21_large* L) { delete L; }

int s21_large_op(s21_large* L) {
    return lguard([&] {
        if (!L) fail(S21_INVALID, "null instance");
        check_current_device(L);
        if (L->prepared != "op") {
            L->prepared.clear();
            prepare(L, nullptr);
            L->prepared = "op";
        }
        reset_state(L);
        L->trap = 0;
        const auto t0 = std::chrono::steady_clock::now();
        L->status = newton(L, false, 0.0);
        LCHECK(hipStreamSynchronize(L->stream));
        L->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        L->npts = 0;
        L->axis.clear();
    });
}

int s21_large_tran(s21_large* L, const s21_tran_options* tr) {
    return lguard([&] {
        if (!L || !tr) fail(S21_INVALID, "null argument");
        if (!(tr->tstep > 0.0) && tr->tstop > 0.0) fail(S21_INVALID, "tstep must be positive");
        check_current_device(L);
        L->prepared.clear();  // the IC values are part of the structure's element values
        prepare(L, tr);
        L->prepared = "tran";
        reset_state(L);
        L->trap = 0;
        const int n = L->t.n;
        const int64_t npts = s21_tran_point_count(tr->tstep, tr->tstop);
        L->wave.alloc((size_t)npts * n * sizeof(double));
        const auto t0 = std::chrono::steady_clock::now();
        L->status = newton(L, false, 0.0);  // operating point with the ICs forced (analysis.rs:540-549)
        int64_t stored = 0;
        if (L->status == S21_OK) {
            LCHECK(hipMemcpyAsync(L->wave.as<double>(), L->x.p, n * sizeof(double), hipMemcpyDeviceToDevice, L->stream));
            stored = 1;
            for (int64_t p = 1; p < npts; p++) {  // t = tstep, 2 tstep, ... while t < tstop
                L->status = newton(L, true, tr->tstep);
                if (L->status != S21_OK) break;
                LCHECK(hipMemcpyAsync(L->wave.as<double>() + (size_t)p * n, L->x.p, n * sizeof(double),
                                      hipMemcpyDeviceToDevice, L->stream));
                stored = p + 1;
            }
        }
        LCHECK(hipStreamSynchronize(L->stream));
        L->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        L->npts = stored;
        L->axis.assign(stored, 0.0);
        double t = tr->tstep;
        for (int64_t p = 1; p < stored; p++) {
            L->axis[p] = t;
            t += tr->tstep;
        }
    });
}

int s21_large_tran_trap(s21_large* L, const s21_tran_options* tr, const s21_trap_options* o) {
    return lguard([&] {
        if (!L || !tr || !o) fail(S21_INVALID, "null argument");
        if (!(tr->tstep > 0.0) || !(tr->tstop >= 0.0)) fail(S21_INVALID, "tstep must be positive");
        check_current_device(L);
        for (const s21_comp& c : L->comps)
            if (c.type == S21_COMP_BSIM4)  // tran_op bakes ag0 = 1 / dt into its companions (tran.rs:86)
                fail(S21_INVALID, "BSIM4 devices integrate by Backward Euler only");
        L->prepared.clear();
        prepare(L, tr);
        L->prepared = "tran";
        reset_state(L);
        L->trap = 0;  // the operating point integrates nothing
        const int n = L->t.n;
        const int64_t npts = s21_tran_point_count(tr->tstep, tr->tstop);
        std::vector<double> axis(npts, 0.0);  // the reference's output grid (t += tstep)
        {
            double t = tr->tstep;
            for (int64_t p = 1; p < npts; p++) {
                axis[p] = t;
                t += tr->tstep;
            }
        }
        L->wave.alloc((size_t)npts * n * sizeof(double));
        for (Buf& h : L->hist) h.alloc((size_t)std::max(n, 1) * sizeof(double));
        L->accepted = L->rejected = 0;
        const auto t0 = std::chrono::steady_clock::now();
        L->status = newton(L, false, 0.0);
        int64_t stored = 0;
        double* W = L->wave.as<double>();
        auto copy = [&](double* dst, const void* src) {
            LCHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, L->stream));
        };
        if (L->status == S21_OK) {
            copy(W, L->x.p);
            stored = 1;
            const double dtmax = tr->tstep;
            const double dtmin = o->dt_min > 0.0 ? o->dt_min : dtmax * 1e-9;
            // hist[0] = x at the last accepted point (time th[0]), hist[1..3] the ones before
            int h[4] = {0, 1, 2, 3};
            double th[4] = {0.0, 0.0, 0.0, 0.0};
            int nh = 1;
            copy(L->hist[h[0]].as<double>(), L->x.p);
            double t = 0.0, dt = o->lte ? dtmax / 4.0 : dtmax;
            L->trap = 2;  // the first step is Backward Euler, as SPICE starts: the operating point's capacitor
                          // currents (zero) do not describe the circuit released from its IC forcing
            const double t_end = axis[npts - 1];
            int64_t k = 1;
            while (k < npts) {
                if (o->lte) dt = std::min(dt, t_end - t);
                const double tn = o->lte ? t + dt : axis[k];
                L->status = newton(L, true, dt, false);
                double r = 0.0;
                if (L->status == S21_OK && o->lte && nh >= 3) {
                    LCHECK(hipMemsetAsync(L->mxbits.p, 0, sizeof(unsigned long long), L->stream));
                    hipLaunchKernelGGL(k_lte, dim3(blocks(n)), dim3(TPB), 0, L->stream, n, L->kinds_dev.as<int>(),
                                       L->x.as<double>(), L->hist[h[0]].as<double>(), L->hist[h[1]].as<double>(),
                                       L->hist[h[2]].as<double>(), tn, th[0], th[1], th[2], o->reltol, o->abstol,
                                       o->trtol, L->mxbits.as<unsigned long long>());
                    unsigned long long bits = 0;
                    LCHECK(hipMemcpyAsync(&bits, L->mxbits.p, sizeof(bits), hipMemcpyDeviceToHost, L->stream));
                    LCHECK(hipStreamSynchronize(L->stream));
                    r = __builtin_bit_cast(double, bits);
                }
                if (L->status != S21_OK || r > 1.0) {  // reject: back to the last accepted point
                    if (!o->lte) break;
                    L->rejected++;
                    dt = L->status != S21_OK ? dt / 4.0 : dt * std::max(0.25, 0.9 * std::cbrt(1.0 / r));
                    copy(L->x.as<double>(), L->hist[h[0]].p);
                    if (dt < dtmin) {
                        L->status = S21_CONVERGENCE;
                        fail(S21_CONVERGENCE, "Timestep too small");
                    }
                    L->status = S21_OK;
                    continue;
                }
                commit_state(L, true, dt);
                L->accepted++;
                L->trap = 1;
                if (o->lte) {  // output-grid points in (t, tn], linearly between the accepted points
                    while (k < npts && axis[k] <= tn * (1.0 + 1e-12)) {
                        const double w = std::min(1.0, (axis[k] - t) / (tn - t));
                        hipLaunchKernelGGL(k_interp, dim3(blocks(n)), dim3(TPB), 0, L->stream, n,
                                           L->hist[h[0]].as<double>(), L->x.as<double>(), w, W + (size_t)k * n);
                        k++;
                    }
                } else {
                    copy(W + (size_t)k * n, L->x.p);
                    k++;
                }
                stored = k;
                // history: the new point in front
                const int last = h[3];
                for (int q = 3; q > 0; q--) {
                    h[q] = h[q - 1];
                    th[q] = th[q - 1];
                }
                h[0] = last;
                th[0] = tn;
                copy(L->hist[h[0]].as<double>(), L->x.p);
                nh = std::min(nh + 1, 4);
                t = tn;
                if (o->lte && nh >= 4) dt = std::min(dtmax, dt * (r > 0.0 ? std::min(2.0, 0.9 * std::cbrt(1.0 / r)) : 2.0));
            }
        }
        LCHECK(hipStreamSynchronize(L->stream));
        L->trap = 0;
        L->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        L->npts = stored;
        L->axis.assign(axis.begin(), axis.begin() + stored);
    });
}

int s21_large_trap_stats(const s21_large* L, int64_t* accepted, int64_t* rejected) {
    return lguard([&] {
        if (!L) fail(S21_INVALID, "null instance");
        if (accepted) *accepted = L->accepted;
        if (rejected) *rejected = L->rejected;
    });
}

int32_t s21_large_n_vars(const s21_large* L) { return L ? (L->t.n ? L->t.n : L->n_vars) : 0; }

const char* s21_large_var_name(const s21_large* L, int32_t k) {
    if (!L || k < 0) return nullptr;
    const std::vector<std::string>& nm = L->t.n ? L->t.names : L->names;
    return k < (int)nm.size() ? nm[k].c_str() : nullptr;
}

int s21_large_get_x(const s21_large* L, double* out) {
    return lguard([&] {
        if (!L || !out || !L->t.n) fail(S21_INVALID, "no analysis has run");
        LCHECK(hipMemcpy(out, L->x.p, L->t.n * sizeof(double), hipMemcpyDeviceToHost));
    });
}

int s21_large_get_status(const s21_large* L, int32_t* status, int64_t* iters) {
    return lguard([&] {
        if (!L) fail(S21_INVALID, "null instance");
        if (status) *status = L->status;
        if (iters) *iters = L->iters;
    });
}

int64_t s21_large_n_points(const s21_large* L) { return L ? L->npts : 0; }

int s21_large_get_axis(const s21_large* L, double* out) {
    return lguard([&] {
        if (!L || !out) fail(S21_INVALID, "bad arguments");
        std::copy(L->axis.begin(), L->axis.end(), out);
    });
}

int s21_large_get_wave(const s21_large* L, int32_t n_sel, const int32_t* vars, double* out) {
    return lguard([&] {
        if (!L || !out || n_sel < 0 || (n_sel > 0 && !vars)) fail(S21_INVALID, "bad arguments");
        const int n = L->t.n;
        for (int j = 0; j < n_sel; j++)
            if (vars[j] < 0 || vars[j] >= n) fail(S21_INVALID, "variable out of range");
        std::vector<double> row(n);
        for (int64_t p = 0; p < L->npts; p++) {
            LCHECK(hipMemcpy(row.data(), L->wave.as<double>() + (size_t)p * n, n * sizeof(double),
                             hipMemcpyDeviceToHost));
            for (int j = 0; j < n_sel; j++) out[(size_t)p * n_sel + j] = row[vars[j]];
        }
    });
}

int s21_large_structure(const s21_large* L, int64_t* nnz_a, int64_t* nnz_lu, int64_t* fronts, int64_t* levels,
                        int64_t* max_front, double* flops_lu) {
    return lguard([&] {
        if (!L || !L->t.n) fail(S21_INVALID, "no analysis has run");
        if (nnz_a) *nnz_a = L->S.nnz_a;
        if (nnz_lu) *nnz_lu = L->S.nnz_lu;
        if (fronts) *fronts = L->S.ns;
        if (levels) *levels = (int64_t)L->levels.size();
        if (max_front) *max_front = L->S.max_front;
        if (flops_lu) *flops_lu = L->S.flops;
    });
}

int s21_large_timing(const s21_large* L, double* total_ms, int64_t* factorisations) {
    return lguard([&] {
        if (!L) fail(S21_INVALID, "null instance");
        if (total_ms) *total_ms = L->ms_total;
        if (factorisations) *factorisations = L->factorisations;
    });
}

int s21_large_set_reuse(s21_large* L, int32_t on) {
    return lguard([&] {
        if (!L) fail(S21_INVALID, "null instance");
        if (on && L->has_bsim4) fail(S21_INVALID, "factor reuse is not available with BSIM4 devices");
        L->reuse = on ? 1 : 0;
    });
}

int s21_large_time_factor(s21_large* L, int32_t reps, double* factor_ms, double* solve_ms) {
    return lguard([&] {
        if (!L || reps <= 0 || !L->t.n) fail(S21_INVALID, "bad arguments");
        check_current_device(L);
        float ms = 0.0f;
        LCHECK(hipEventRecord(L->ev0, L->stream));
        for (int r = 0; r < reps; r++) {
            if (L->precision == S21_MIXED)
                factor<float>(L);
            else
                factor<double>(L);
        }
        LCHECK(hipEventRecord(L->ev1, L->stream));
        LCHECK(hipEventSynchronize(L->ev1));
        LCHECK(hipEventElapsedTime(&ms, L->ev0, L->ev1));
        if (factor_ms) *factor_ms = ms / reps;
        LCHECK(hipEventRecord(L->ev0, L->stream));
        for (int r = 0; r < reps; r++) {
            if (L->precision == S21_MIXED)
                solve<float>(L);
            else
                solve<double>(L);
        }
        LCHECK(hipEventRecord(L->ev1, L->stream));
        LCHECK(hipEventSynchronize(L->ev1));
        LCHECK(hipEventElapsedTime(&ms, L->ev0, L->ev1));
        if (solve_ms) *solve_ms = ms / reps;
    });
}

int s21_large_host_check(int32_t n, int64_t ne, const int32_t* rows, const int32_t* cols, const double* vals,
                         const double* b, double relax, double* x, int64_t* nnz_lu, double* flops_lu) {
    return lguard([&] {
        if (n <= 0 || ne < 0 || (ne > 0 && (!rows || !cols || !vals)) || !b || !x) fail(S21_INVALID, "bad arguments");
        std::vector<int> r(rows, rows + ne), c(cols, cols + ne);
        std::vector<double> v(vals, vals + ne);
        LargeSym S;
        large_symbolic(n, r, c, &v, S, relax);
        std::vector<double> pool, xs;
        large_factor_host(S, r, c, v, pool);
        large_solve_host(S, pool, std::vector<double>(b, b + n), xs);
        std::copy(xs.begin(), xs.end(), x);
        if (nnz_lu) *nnz_lu = S.nnz_lu;
        if (flops_lu) *flops_lu = S.flops;
    });
}

}  // extern "C"
