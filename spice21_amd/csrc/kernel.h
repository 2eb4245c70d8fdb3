// Batched Newton kernel for MI355X (gfx950): device code shared by the precompiled kernels and the
// kernels compiled at run time with a factorisation generated for one frozen pattern.
//
// One persistent launch runs every instance's whole analysis: the operating point and, for a transient,
// every time step.  A workgroup holds up to MAX_WAVES wavefronts; each wavefront holds `ipw` instances
// of one group, each owning L = 64 / ipw consecutive lanes and one LDS region:
//     x[n] | c[n] | a[nnz] | rhs[n] | hub[n_hub] | part[n_part] (| lu32[nnz] | c32[n] in mixed precision)
// next to the workgroup's shared copy of the group's records.  Each Newton iteration (Solver::solve,
// analysis.rs:169-210) runs these phases:
//   1. device evaluation by colour classes: the devices of one colour share no matrix / RHS slot other
//      than hub slots, so each lane evaluates its devices and adds their stamps straight into A and the
//      RHS without atomics; stamps into a hub slot go to their own contribution cell;
//   2. hub gather: each hub slot sums its contributions in component-then-stamp order, the order
//      Solver::update adds them (analysis.rs:153-168);
//   3. KCL residual res = rhs - A x in FP64 over the pivoted rows (sparse21/mod.rs:298-327) and the
//      convergence test on the previous update and the residual (analysis.rs:331-345);
//   4. numeric LU on the pivot order and fill pattern frozen by the host symbolic phase, level-scheduled
//      by rows (IKJ order, which applies every update to an entry in the same pivot order as the
//      reference's right-looking elimination, sparse21/mod.rs:865-919), with the forward sweep fused in;
//   5. level-scheduled backward sweep (sparse21/mod.rs:966-979);
//   6. the max-|dx| <= 1 clamp and the update x += dx (analysis.rs:196-207).
// In mixed precision phases 4-5 run in FP32 while 1-3 and 6 stay FP64, so Newton iterates as
// iterative refinement.  Instances of one wavefront that finish early leave their lanes idle.
#pragma once
#ifndef __HIPCC_RTC__  // run-time compilation provides the HIP declarations itself
#include <hip/hip_runtime.h>
#endif

#include "bsim4_eval.h"
#include "layout.h"
#include "spice21_mi355x.h"

namespace s21 {
static_assert(b4::B4_NS == BSIM4_NE, "BSIM4 code stride");

// ============================================================================================
// Device-side descriptors
// ============================================================================================
constexpr int M1_CORE_F = 11, M1_CAPS_F = 4, M1_STATE_F = 7, D_PAR_F = 14, D_STATE_F = 4;
// BSIM4: doubles of a parameter structure; state fields: the last load's limiting fields (vds, vgs, vbs,
// vbd, von) and charges (qg, qd, qb), the committed charges, and the limiting fields the last load started
// from (to resume an interrupted solve at that load)
constexpr int B4_PAR_D = (int)(sizeof(b4::B4Par) / sizeof(double));
static_assert(sizeof(b4::B4Par) % sizeof(double) == 0, "B4Par is doubles only");
enum { B4_ST_LIM = 0, B4_ST_Q = 5, B4_ST_QC = 8, B4_ST_PREV = 11, B4_ST_F = 16 };

struct TopoGPU {
    int n, nnz, n_hub, n_part;
    int n_cls[N_CLS];  // devices per class (Mos1, Mos0, Diode, linear)
    int cls_start[N_CLS + 1];
    int n_col;
    int n_gt, n_gc, n_rt, n_rc, n_clvl, n_blvl;
    int sym_status;
    int blob_vec;       // 16-byte vectors of the record blob copied into LDS at launch
    const uint4* blob;  // the blob in global memory
    // offsets (uint16 units, 8-byte aligned) of the record tables inside the blob
    int o_gt, o_gidx, o_gc, o_rt, o_re, o_rc, o_ri2e, o_ce2i, o_diag, o_clvl, o_fch, o_fl, o_fu, o_blvl, o_bro, o_bu;
    int o_col;   // colours: n_col + 1 device-position starts, then {class, class-local first index, first code}
    int o_dev;   // per device: node + 1 for six terminals, flags, 0
    int o_code;  // per device entry: slot, 0x8000 | hub index, or 0xFFFF
};

struct GroupGPU {
    TopoGPU t;
    int n_inst, blk0;
    // Parameter fields, packed: field f of device j (class-local order) of instance i is
    // par[off[f] + i * istr[f] + j]; a field no instance overrides is stored once (istr = 0).
    const double* par;
    long long m1_off[M1_CORE_F + M1_CAPS_F];  // Mos1: 11 core fields, then 4 capacitances
    int m1_istr[M1_CORE_F + M1_CAPS_F];
    long long d_off[D_PAR_F];  // diode
    int d_istr[D_PAR_F];
    long long lin_off[3];  // linear element value: operating point, transient, AC (R: g, C: c, V: acm)
    int lin_istr[3];
    double* m1_state;  // [14][n_inst][n_m1]: committed op (7) then guess (7)
    double* acop;      // [9][n_inst][n_m1] small-signal operating point (capture), or nullptr
    double* d_state;   // [4][n_inst][n_d]: guess vd, guess charge, committed charge, guess vd the last
                       // evaluation started from (to resume an interrupted solve at that evaluation)
    double* c_q;            // [n_inst][n_lin] committed capacitor charge
    double* x;              // [n_inst][n]
    int* status;            // [n_inst]
    long long* iters;       // [n_inst]
    double* wave;           // [n_inst][npts][n] (transient) or nullptr
    double* dbg;            // assembly dump [nnz + n] (assemble mode)
    long long* prof;        // [n_inst][8] phase cycle counters (profiling mode) or nullptr
    int* resume_at;         // [n_inst][2]: time point of the last solve and the iterations it had used
                            // before its current one (written at exit; read by a resuming launch)
    int* ac_rec;            // AC: [n_inst][npts] per frequency: status << 8 | Newton iterations (0 = not solved)
    double* ac_end;         // AC: [n_inst][ac_runs][n] complex: each run's last iterate (first pass)
    int* ac_state;          // AC: [n_inst][ac_runs] outcome of each run's re-solve from the previous
                            // run's end (second pass); entry 0: run to stitch from in order (third pass)
    // BSIM4 (class-local device j of instance i).  Parameters by field k of the B4Par structure: a field
    // that no instance changes is stored once per device, b4_sh[k][j]; the others (b4_map[k] = their row
    // r, -1 for a shared field) per instance, b4_var[r][i][j], so that a wavefront's lanes (the devices of
    // an instance) read consecutive addresses.  State b4_st[B4_ST_F][i][j] (see B4_ST_*).
    const int* b4_map;
    const double* b4_sh;
    const double* b4_var;
    double* b4_st;
};

enum { MODE_OP = 0, MODE_TRAN = 1, MODE_ASSEMBLE = 2, MODE_AC = 3 };
struct Args {
    int mode;
    int max_iter;
    double gmin, iabstol, reltol;
    double tstep, tstop;
    long long npts;
    int dbg_inst;
    int prof;
    int waves;       // wavefronts per workgroup
    int ipw;         // instances per wavefront (1, 2, 4, 8 or 16)
    int inst_bytes;  // LDS bytes per instance
    int capture;     // operating point: store each Mos1's small-signal operating point for AC
    int resume;      // each instance continues an interrupted solve at G.resume_at's time point: G.x holds
                     // the iterate, G.iters the iterations so far, G.resume_at those of that solve, the
                     // state arrays the state before the interrupted iteration; its first convergence
                     // test fails (the interrupted iteration had passed that test already)
    int asm_tran;    // assemble mode: the transient matrix (companions from the state arrays)
    const double* omega;  // AC: angular frequency of each point
    int ac_runs;          // AC: each instance's frequencies split into this many contiguous runs (1 outside AC)
    int ac_pass;          // AC: 0 = every run from x = 0; 1 = each run re-solved from the previous run's
                          // end until it meets the first pass; 2 = in order from G.ac_state's run
    int ac_look;          // AC, pass 2: wavefronts (a workgroup) per instance, each testing one of the next
                          // frequencies for the reference's immediate acceptance, so that a stretch of
                          // kept solutions costs one round
};

// AC launch units per instance: one per run, one per run boundary, or one.
__host__ __device__ inline int ac_units(const Args& A) {
    if (A.mode != MODE_AC) return 1;
    return A.ac_pass == 0 ? A.ac_runs : A.ac_pass == 1 ? A.ac_runs - 1 : A.ac_look;
}
enum { AC_UNMET = 0, AC_MET = 1, AC_FAILED = 2 };

// Address spaces: pointers read out of descriptors carry no address space, so they are re-qualified in
// device code to keep LDS accesses on ds_* and global accesses on global_* instructions.
#define LDS __attribute__((address_space(3)))
#define GLB __attribute__((address_space(1)))
template <typename T>
__device__ inline GLB T* glb(T* p) {
    return (GLB T*)p;
}

constexpr int WAVE = 64;
constexpr int MAX_WAVES = 8;  // wavefronts per workgroup; the kernel's launch bound
constexpr size_t LDS_PER_CU = 160 * 1024;  // gfx950
constexpr int NPROF = 8;  // eval, gather, residual, factor (+forward sweep), backward sweep, update, commit, total
constexpr int RUNNING = -1;

using u16 = unsigned short;

// Orders LDS accesses between the lanes of one wavefront (a wavefront executes its LDS operations in
// order; the fences keep the compiler from moving accesses across this point).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ inline bool wave_any(bool p) { return __ballot(p) != 0ull; }

// LDS views of the record tables.
struct Idx {
    LDS const uint64_t *gt, *gc, *rt, *rc, *fch, *fl, *bro;  // four 16-bit fields
    LDS const uint32_t *re, *fu, *bu;                         // two 16-bit fields
    LDS const u16 *gidx, *ri2e, *ce2i, *diag, *clvl, *blvl, *col, *dev, *code;
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
    int inst;       // instance of this lane's segment
    int sl, L;      // lane within the segment, lanes per segment
    int seg0;       // first lane of the segment
    LDS double *x, *c, *a, *rhs, *hub, *part;
    LDS float* lu32;
};

// Ballot restricted to the lane's segment.
__device__ inline bool seg_any(const Ctx& C, bool p) {
    const uint64_t b = __ballot(p);
    if (C.L == WAVE) return b != 0ull;
    return ((b >> C.seg0) & ((1ull << C.L) - 1ull)) != 0ull;
}
// Maximum over the lane's segment (every lane of the wavefront calls it).
__device__ inline double seg_max(const Ctx& C, double v) {
    for (int o = C.L >> 1; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, WAVE);
        v = (w > v) ? w : v;  // the reference's fold drops NaN (analysis.rs:198)
    }
    return v;
}

__device__ inline double xv(LDS const double* x, int v) { return v < 0 ? 0.0 : x[v]; }

// Sink of a device's stamps: entry e goes to its slot of A|RHS (added in order) or to its hub cell.
struct Put {
    LDS double* ar;
    LDS double* hub;
    LDS const u16* code;
    __device__ void operator()(int e, double v) const {
        const int c = code[e];
        if (c == 0xFFFF) return;
        if (c & 0x8000)
            hub[c & 0x7FFF] = v;
        else
            ar[c] = ar[c] + v;
    }
};

// A device's stamps collected in registers (entry index fixed at compile time), then applied in one batch.
// The host makes every slot that one device stamps twice a hub slot, so a device's non-hub entries reach
// distinct slots and all of its slot loads may precede all of its stores.
template <int N>
struct Stamps {
    double v[N];
    __device__ void operator()(int e, double x) { v[e] = x; }
};
struct M1Stamps {
    static constexpr bool raw_stamps = true;  // Mos1: reference stamp order, swap applied at the codes
    double v[MOS1_NE];
    bool rev;
    __device__ void stamp(int k, bool r, double x) {
        v[k] = x;
        rev = r;
    }
};
template <int N, bool SIGMA>
__device__ inline void apply(const Ctx& C, LDS const u16* code, const double (&v)[N], bool rev) {
    constexpr int B = 8;  // entries per batch
#pragma unroll
    for (int k0 = 0; k0 < N; k0 += B) {
        int c[B];
        double cur[B];
#pragma unroll
        for (int k = 0; k < B; k++)
            if (k0 + k < N) c[k] = code[(SIGMA && rev) ? mos1_sigma(k0 + k) : k0 + k];
#pragma unroll
        for (int k = 0; k < B; k++)
            if (k0 + k < N) cur[k] = c[k] < 0x8000 ? C.a[c[k]] : 0.0;
#pragma unroll
        for (int k = 0; k < B; k++) {
            if (k0 + k >= N) continue;
            if (c[k] < 0x8000)
                C.a[c[k]] = cur[k] + v[k0 + k];
            else if (c[k] != 0xFFFF)
                C.hub[c[k] & 0x7FFF] = v[k0 + k];
        }
    }
}

// Only the listed entries (stamp indices fixed at compile time): the others' codes are 0xFFFF.
template <int N, bool SIGMA>
__device__ inline void apply_some(const Ctx& C, LDS const u16* code, const double (&v)[MOS1_NE], const int (&e)[N],
                                  bool rev) {
    constexpr int B = 8;
#pragma unroll
    for (int k0 = 0; k0 < N; k0 += B) {
        int c[B];
        double cur[B];
#pragma unroll
        for (int k = 0; k < B; k++)
            if (k0 + k < N) c[k] = code[(SIGMA && rev) ? mos1_sigma(e[k0 + k]) : e[k0 + k]];
#pragma unroll
        for (int k = 0; k < B; k++)
            if (k0 + k < N) cur[k] = c[k] < 0x8000 ? C.a[c[k]] : 0.0;
#pragma unroll
        for (int k = 0; k < B; k++) {
            if (k0 + k >= N) continue;
            if (c[k] < 0x8000)
                C.a[c[k]] = cur[k] + v[e[k0 + k]];
            else if (c[k] != 0xFFFF)
                C.hub[c[k] & 0x7FFF] = v[e[k0 + k]];
        }
    }
}
// Entries a Mos1 without live charge terms can stamp (the host gives the charge-only ones no slot unless
// the device has capacitance and the analysis is a transient); of those, the ones a tied Mos1 keeps
// after its merges (it has no terminal resistances either).  Both lists are closed under the swap.
constexpr int M1_DC_ENTRIES[20] = {0, 1, 2, 3, 4, 5, 7, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 25};
constexpr int M1T_DC_ENTRIES[8] = {0, 1, 2, 3, 4, 5, 22, 23};

// Tied Mos1: stamps to entries (the source/drain swap), coinciding entries merged, then one batch.
__device__ inline void apply_tied(const Ctx& C, LDS const u16* code, const M1Stamps& sv, bool charges) {
    double v[MOS1_NE];
#pragma unroll
    for (int e = 0; e < MOS1_NE; e++) v[e] = sv.rev ? sv.v[mos1_sigma(e)] : sv.v[e];
    v[1] = ((v[1] + v[7]) + v[13]) + v[15];  // (s,s)
    v[2] = v[2] + v[14];                     // (d,s)
    v[3] = v[3] + v[12];                     // (s,d)
    v[5] = v[5] + v[11];                     // (s,g)
    v[8] = v[8] + v[10];                     // (g,s)
    v[23] = v[23] + v[25];                   // rhs s
    if (charges)
        apply<MOS1_NE, false>(C, code, v, false);
    else
        apply_some<8, false>(C, code, v, M1T_DC_ENTRIES, false);
}

__device__ inline int mos1_count(const TopoGPU& T) { return T.cls_start[CLS_MOS1T + 1] - T.cls_start[CLS_MOS1]; }

// ---- device evaluation ---------------------------------------------------------------------------
__device__ inline void eval_mos1(const Ctx& C, int j, LDS const u16* dv, LDS const u16* code, bool tran, const Args& A,
                                 bool tied) {
    const GroupGPU& G = *C.G;
    const int nm = mos1_count(C.T);
    const size_t s = (size_t)G.n_inst * nm, o = (size_t)C.inst * nm + j;
    GLB const double* par = glb(G.par);
    auto F = [&](int f) { return par[G.m1_off[f] + (long long)C.inst * G.m1_istr[f] + j]; };
    Mos1Core P;
    P.p = F(0);
    P.vt0_t = F(1);
    P.beta = F(2);
    P.phi_t = F(3);
    P.gamma = F(4);
    P.lambda = F(5);
    P.vtherm = F(6);
    P.isat_d = F(7);
    P.isat_s = F(8);
    P.grd = F(9);
    P.grs = F(10);
    const double vd = xv(C.x, (int)dv[RD] - 1), vg = xv(C.x, (int)dv[RG] - 1), vs = xv(C.x, (int)dv[RS] - 1),
                 vb = xv(C.x, (int)dv[RB] - 1);
    M1Stamps sv;
    if (!(dv[6] & DEV_CAPS)) {
        mos1_load<false>(P, nullptr, nullptr, vd, vg, vs, vb, A.gmin, A.tstep, sv, nullptr);
        if (tied)
            apply_tied(C, code, sv, false);
        else
            apply_some<20, true>(C, code, sv.v, M1_DC_ENTRIES, sv.rev);
        return;
    }
    // gate capacitance: capacitances and operating-point state from global memory
    Mos1Caps K;
    K.cox = F(M1_CORE_F + 0);
    K.cgs_ov = F(M1_CORE_F + 1);
    K.cgd_ov = F(M1_CORE_F + 2);
    K.cgb_ov = F(M1_CORE_F + 3);
    Mos1State op, gs;
    GLB double* st = glb(G.m1_state);
    op.vgs = st[0 * s + o];
    op.vgd = st[1 * s + o];
    op.vgb = st[2 * s + o];
    op.cgs = st[3 * s + o];
    op.cgd = st[4 * s + o];
    op.cgb = st[5 * s + o];
    op.reversed = st[6 * s + o];
    if (tran)
        mos1_load<true>(P, &K, &op, vd, vg, vs, vb, A.gmin, A.tstep, sv, &gs);
    else
        mos1_load<false>(P, &K, &op, vd, vg, vs, vb, A.gmin, A.tstep, sv, &gs);
    if (tied)
        apply_tied(C, code, sv, true);  // charge entries have slots (hub cells among them) in both solves
    else
        apply<MOS1_NE, true>(C, code, sv.v, sv.rev);
    st[7 * s + o] = gs.vgs;
    st[8 * s + o] = gs.vgd;
    st[9 * s + o] = gs.vgb;
    st[10 * s + o] = gs.cgs;
    st[11 * s + o] = gs.cgd;
    st[12 * s + o] = gs.cgb;
    st[13 * s + o] = gs.reversed;
}

// BSIM4 (bsim4solver.rs:134-145): Bsim4::op from the terminal voltages with limiting against the last
// load's point, tran_op's companions against the committed point in a transient, the 42 stamps.  The new
// point is written to global memory as it is computed (its 86 fields never occupy registers together).
// The parameter view of one BSIM4 device of the batched kernel: each field the evaluation reads is loaded
// where it is stored (the field's row is uniform across the wavefront, the devices of an instance are
// consecutive), no copy of the whole structure is made.
struct B4Gather {
    GLB const int* map;
    GLB const double* sh;
    GLB const double* va;
    size_t nb, row0, j;  // devices per instance; instance * nb; class-local device
    size_t n_inst;
    template <int K>
    __device__ double get() const {
        const int r = map[K];
        return r < 0 ? sh[(size_t)K * nb + j] : va[(size_t)r * n_inst * nb + row0 + j];
    }
};

__device__ inline void eval_b4(const Ctx& C, int j, LDS const u16* dv, LDS const u16* code, bool tran, const Args& A) {
    const GroupGPU& G = *C.G;
    const int nb = C.T.n_cls[CLS_BSIM4];
    const size_t s = (size_t)G.n_inst * nb, o = (size_t)C.inst * nb + j;
    const B4Gather P{glb(G.b4_map), glb(G.b4_sh), glb(G.b4_var), (size_t)nb, (size_t)C.inst * nb, (size_t)j,
                     (size_t)G.n_inst};
    GLB double* st = glb(G.b4_st);
    b4::B4Op Gl;
    Gl.vds = st[(B4_ST_LIM + 0) * s + o];
    Gl.vgs = st[(B4_ST_LIM + 1) * s + o];
    Gl.vbs = st[(B4_ST_LIM + 2) * s + o];
    Gl.vbd = st[(B4_ST_LIM + 3) * s + o];
    Gl.von = st[(B4_ST_LIM + 4) * s + o];
    st[(B4_ST_PREV + 0) * s + o] = Gl.vds;
    st[(B4_ST_PREV + 1) * s + o] = Gl.vgs;
    st[(B4_ST_PREV + 2) * s + o] = Gl.vbs;
    st[(B4_ST_PREV + 3) * s + o] = Gl.vbd;
    st[(B4_ST_PREV + 4) * s + o] = Gl.von;
    b4::B4Op N;
    b4::op_eval(P, Gl, xv(C.x, (int)dv[0] - 1), xv(C.x, (int)dv[1] - 1), xv(C.x, (int)dv[2] - 1),
                xv(C.x, (int)dv[3] - 1), A.gmin, N);
    b4::B4Tran T{};
    if (tran) {
        b4::B4Op O;  // tran_op reads the committed point's charges only
        O.qg = st[(B4_ST_QC + 0) * s + o];
        O.qd = st[(B4_ST_QC + 1) * s + o];
        O.qb = st[(B4_ST_QC + 2) * s + o];
        b4::tran_eval(P, O, N, A.tstep, T);
    }
    Stamps<BSIM4_NE> sv;
    b4::stamp(P, N, T, sv);
    apply<BSIM4_NE, false>(C, code, sv.v, false);
    st[(B4_ST_LIM + 0) * s + o] = N.vds;
    st[(B4_ST_LIM + 1) * s + o] = N.vgs;
    st[(B4_ST_LIM + 2) * s + o] = N.vbs;
    st[(B4_ST_LIM + 3) * s + o] = N.vbd;
    st[(B4_ST_LIM + 4) * s + o] = N.von;
    st[(B4_ST_Q + 0) * s + o] = N.qg;
    st[(B4_ST_Q + 1) * s + o] = N.qd;
    st[(B4_ST_Q + 2) * s + o] = N.qb;
}

__device__ inline void eval_diode(const Ctx& C, int j, LDS const u16* dv, LDS const u16* code, bool tran,
                                  const Args& A) {
    const GroupGPU& G = *C.G;
    const size_t s = (size_t)G.n_inst * C.T.n_cls[CLS_DIODE], o = (size_t)C.inst * C.T.n_cls[CLS_DIODE] + j;
    GLB const double* par = glb(G.par);
    auto F = [&](int f) { return par[G.d_off[f] + (long long)C.inst * G.d_istr[f] + j]; };
    DiodeP P;
    P.isat = F(0);
    P.vte = F(1);
    P.vcrit = F(2);
    P.gspr = F(3);
    P.bv = F(4);
    P.has_bv = F(5);
    P.cz = F(6);
    P.cz2 = F(7);
    P.dep = F(8);
    P.f1 = F(9);
    P.f3 = F(10);
    P.tt = F(11);
    P.vj = F(12);
    P.m = F(13);
    GLB double* st = glb(G.d_state);
    const double vr = xv(C.x, (int)dv[DR_] - 1), vn = xv(C.x, (int)dv[DN_] - 1);
    const double vd0 = st[0 * s + o];
    DiodeGuess gs;
    Stamps<DIODE_NE> sv;
    if (tran)
        diode_load<true>(P, vr, vn, A.gmin, A.tstep, vd0, st[2 * s + o], sv, &gs);
    else
        diode_load<false>(P, vr, vn, A.gmin, A.tstep, vd0, st[2 * s + o], sv, &gs);
    apply<DIODE_NE, false>(C, code, sv.v, false);
    st[0 * s + o] = gs.vd;
    st[1 * s + o] = gs.charge;
    st[3 * s + o] = vd0;
}

__device__ inline void eval_lin(const Ctx& C, int j, LDS const u16* dv, LDS const u16* code, bool tran, const Args& A) {
    const GroupGPU& G = *C.G;
    Stamps<6> put;
    const int typ = dv[6] >> 4;
    const size_t o = (size_t)C.inst * C.T.n_cls[CLS_LIN] + j;
    const int f = tran ? 1 : 0;
    const double val = glb(G.par)[G.lin_off[f] + (long long)C.inst * G.lin_istr[f] + j];
    if (typ == S21_COMP_R) {
        put(0, val);
        put(1, val);
        put(2, -val);
        put(3, -val);
    } else if (typ == S21_COMP_C) {
        if (!tran) {  // no stamps in the operating point (comps/mod.rs:205-209)
            // Hub cells are summed whether or not their device stamped, so this device's cells must hold
            // zero rather than whatever the LDS held before the launch.
#pragma unroll
            for (int e = 0; e < 6; e++)
                if (code[e] != 0xFFFF && (code[e] & 0x8000)) C.hub[code[e] & 0x7FFF] = 0.0;
            return;
        }
        cap_load_tran(val, xv(C.x, (int)dv[0] - 1) - xv(C.x, (int)dv[1] - 1), glb(G.c_q)[o], A.tstep, put);
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
    apply<6, false>(C, code, put.v, false);
}

__device__ inline long long cycles() { return (long long)__builtin_readcyclecounter(); }

// Assembly of A and the RHS at the lane segment's x: colours in order, then the hub gather.  With
// profiling on (prof, uniform), pg[0] gets the device evaluation's cycles and pg[1] the hub gather's; pg
// is the instance's counters in global memory, given to one lane of the instance (nullptr elsewhere).
// B4: the kernel evaluates BSIM4 devices (compiled into the kernels of batches that hold them only).
template <bool B4 = false>
__device__ inline void assemble(const Ctx& C, bool act, bool tran, const Args& A, bool prof = false,
                                long long* pg = nullptr) {
    const long long t0 = prof ? cycles() : 0;
    const TopoGPU& T = C.T;
    const Idx& I = C.I;
    const int ns = T.nnz + T.n;
    for (int k = act ? C.sl : ns; k < ns; k += C.L) C.a[k] = 0.0;
    wave_sync();
    for (int cc = 0; cc < T.n_col; cc++) {
        const int p0 = I.col[cc], p1 = I.col[cc + 1];
        LDS const u16* cd = I.col + T.n_col + 1 + 3 * cc;
        const int cls = cd[0], j0 = cd[1], code0 = cd[2];
        const int stride = code_stride(cls);
        for (int p = act ? p0 + C.sl : p1; p < p1; p += C.L) {
            LDS const u16* dv = I.dev + 8 * p;
            LDS const u16* code = I.code + code0 + (p - p0) * stride;
            const int j = j0 + (p - p0);
            if (cls_is_mos1(cls))
                eval_mos1(C, j, dv, code, tran, A, cls == CLS_MOS1T);
            else if (cls == CLS_MOS0)
                mos0_load((dv[6] & DEV_PMOS) ? -1.0 : 1.0, xv(C.x, (int)dv[RD] - 1), xv(C.x, (int)dv[RG] - 1),
                          xv(C.x, (int)dv[RS] - 1), A.gmin, Put{C.a, C.hub, code});
            else if (cls == CLS_DIODE)
                eval_diode(C, j, dv, code, tran, A);
            else if constexpr (B4) {
                if (cls == CLS_LIN)
                    eval_lin(C, j, dv, code, tran, A);
                else
                    eval_b4(C, j, dv, code, tran, A);
            } else {
                eval_lin(C, j, dv, code, tran, A);
            }
        }
        wave_sync();
    }
    const long long t1 = prof ? cycles() : 0;
    // hub slots: ordered sums of their contributions (long slots in chunks whose partials add in order)
    for (int t = act ? C.sl : T.n_gt; t < T.n_gt; t += C.L) {
        const R4 r = r4(I.gt[t]);
        double acc = 0.0;
        for (int k = r.y; k < r.z; k++) acc = acc + C.hub[I.gidx[k]];
        if (r.w == 0xFFFF)
            C.a[r.x] = acc;
        else
            C.part[r.w] = acc;
    }
    wave_sync();
    for (int t = act ? C.sl : T.n_gc; t < T.n_gc; t += C.L) {
        const R4 r = r4(I.gc[t]);
        double acc = C.part[r.y];
        for (int p = r.y + 1; p < r.z; p++) acc = acc + C.part[p];
        C.a[r.x] = acc;
    }
    if (prof) {
        wave_sync();
        const long long t2 = cycles();
        if (pg) {
            pg[0] += t1 - t0;
            pg[1] += t2 - t1;
        }
    }
}

// c[i] = rhs[row_i2e[i]] - (A x)[i] over internal rows (sparse21/mod.rs:298-327).  Returns whether this
// lane saw |res| > iabstol.
__device__ inline bool residual(const Ctx& C, bool act, double iabstol) {
    const Idx& I = C.I;
    bool bad = false;
    for (int t = act ? C.sl : C.T.n_rt; t < C.T.n_rt; t += C.L) {
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
    for (int t = act ? C.sl : C.T.n_rc; t < C.T.n_rc; t += C.L) {
        const R4 r = r4(I.rc[t]);
        double acc = C.part[r.y];
        for (int p = r.y + 1; p < r.z; p++) acc = acc + C.part[p];
        const double v = C.rhs[I.ri2e[r.x]] - acc;
        C.c[r.x] = v;
        bad = bad || (fabs(v) > iabstol);
    }
    return bad;
}

// Division inside the factorisation and sweeps: exact in FP64 (as the reference divides); in the FP32
// factorisation of mixed precision, which only preconditions the FP64 Newton update, the hardware
// reciprocal (1 ulp) times the numerator.
template <typename V>
__device__ inline V ldiv(V a, V b) {
    return a / b;
}
template <>
__device__ inline float ldiv<float>(float a, float b) {
    return a * __builtin_amdgcn_rcpf(b);
}

// Factorisation with the forward sweep fused in.  A chunk of row i processes L entries (i, k) once row k
// is final: l_ik = a_ik / u_kk, a_ij -= l_ik u_kj over row k's U entries, and c_i -= c_k l_ik -- for each
// target in increasing k, the order of the reference's right-looking elimination and column-oriented
// forward substitution (sparse21/mod.rs:865-919, 946-964).  Returns whether this lane saw a pivot the
// kept order cannot use (other than the last): a zero one (the assertion at sparse21/mod.rs:872) or one
// below the smallest normal number of the factors' precision, whose multipliers l_ik = a_ik / u_kk
// overflow.  The reference's own search exchanges a diagonal below 1e-3 of its column
// (sparse21/mod.rs:213-216, 757-762); the engine keeps such a pivot while it is normal, as its Newton
// loop tests a residual computed from the assembled matrix in FP64 and needs the factors only for the
// update.
template <typename V>
__device__ constexpr V min_normal() {
    return sizeof(V) == 8 ? V(2.2250738585072014e-308) : V(1.17549435e-38f);
}
template <typename V>
__device__ inline bool pivot_unusable(V u) {  // false for NaN, which the reference's search passes over too
    return u < min_normal<V>() && u > -min_normal<V>();
}
template <typename V>
__device__ inline bool factor_forward(const Ctx& C, bool act, LDS V* lu) {
    const Idx& I = C.I;
    for (int l = 0; l < C.T.n_clvl; l++) {
        const int e = I.clvl[l + 1];
        for (int ch = act ? I.clvl[l] + C.sl : e; ch < e; ch += C.L) {
            const R4 h = r4(I.fch[ch]);
            V y = (V)C.c[h.x];
            int u = h.w;
            for (int j = h.y; j < h.z; j++) {
                const R4 L = r4(I.fl[j]);
                const V lv = ldiv<V>(lu[L.x], lu[L.y]);
                lu[L.x] = lv;
                const V ck = (V)C.c[L.z];
                if (ck != V(0)) y = y - ck * lv;  // the sweep skips zero entries (sparse21/mod.rs:953)
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
    for (int i = act ? C.sl : C.T.n; i < C.T.n - 1; i += C.L)
        if (pivot_unusable(lu[I.diag[i]])) zero_pivot = true;
    return zero_pivot;
}

// Backward sweep over U by row levels (sparse21/mod.rs:966-979).
template <typename V>
__device__ inline void backward(const Ctx& C, bool act, LDS const V* lu) {
    const Idx& I = C.I;
    for (int l = 0; l < C.T.n_blvl; l++) {
        const int e = I.blvl[l + 1];
        for (int r = act ? I.blvl[l] + C.sl : e; r < e; r += C.L) {
            const R4 h = r4(I.bro[r]);
            V y = (V)C.c[h.x];
            for (int k = h.z; k < h.w; k++) {
                const R2 q = r2(I.bu[k]);
                const V prod = (V)C.c[q.y] * lu[q.x];
                y = y - prod;
            }
            C.c[h.x] = (double)ldiv<V>(y, lu[h.y]);
        }
        wave_sync();
    }
}

// Commit the accepted operating point (Component::commit) of every stateful device of the segment.
template <bool B4 = false>
__device__ inline void commit(const Ctx& C, bool act) {
    const GroupGPU& G = *C.G;
    const TopoGPU& T = C.T;
    const int nl = T.n_cls[CLS_LIN], nd = T.n_cls[CLS_DIODE], nm = mos1_count(T);
    for (int d = act ? C.sl : nl; d < nl; d += C.L) {
        LDS const u16* dv = C.I.dev + 8 * (T.cls_start[CLS_LIN] + d);
        if ((dv[6] >> 4) != S21_COMP_C) continue;
        const size_t o = (size_t)C.inst * nl + d;
        const double vd = xv(C.x, (int)dv[0] - 1) - xv(C.x, (int)dv[1] - 1);
        const double cval = glb(G.par)[G.lin_off[0] + (long long)C.inst * G.lin_istr[0] + d];
        glb(G.c_q)[o] = cval * vd;  // guess.q = c * vd at the accepted voltages
    }
    for (int d = act ? C.sl : nd; d < nd; d += C.L) {
        const size_t s = (size_t)G.n_inst * nd, o = (size_t)C.inst * nd + d;
        glb(G.d_state)[2 * s + o] = glb(G.d_state)[1 * s + o];
    }
    for (int d = act ? C.sl : nm; d < nm; d += C.L) {
        LDS const u16* dv = C.I.dev + 8 * (T.cls_start[CLS_MOS1] + d);
        if (!(dv[6] & DEV_CAPS)) continue;
        const size_t s = (size_t)G.n_inst * nm, o = (size_t)C.inst * nm + d;
        for (int f = 0; f < M1_STATE_F; f++) glb(G.m1_state)[f * s + o] = glb(G.m1_state)[(M1_STATE_F + f) * s + o];
    }
    if constexpr (B4) {
        const int nb = T.n_cls[CLS_BSIM4];
        const size_t s = (size_t)G.n_inst * nb;
        GLB double* st = glb(G.b4_st);
        for (int d = act ? C.sl : nb; d < nb; d += C.L) {
            const size_t o = (size_t)C.inst * nb + d;
            for (int f = 0; f < 3; f++) st[(B4_ST_QC + f) * s + o] = st[(B4_ST_Q + f) * s + o];
        }
    }
}

// One Newton solve (Solver::solve, analysis.rs:169-210) of every active segment of the wavefront.
// `iters` accumulates loads.  Returns the segment's S21_* status (S21_OK for a lane that was inactive).
// A fresh solve has it_left = max_iter and prev_ok = true (dx starts at zero); a resumed one has the
// iterations it has left and prev_ok = false.
// The LU policy supplies factor<V>(C, act, lu) -> zero pivot seen, and backward<V>(C, act, lu), run by
// each instance's own lanes; or, with lane_per_instance, factor_lane / backward_lane run by one lane
// per instance for every instance of the workgroup at once (newton_wg).
struct LuLevels {
    static constexpr bool lane_per_instance = false;
    // Analyses a policy serves: 0 every one; a generated policy 1 (operating points and transients of its
    // real patterns) or 2 (AC launches of its complex patterns).
    static constexpr int kind = 0;
    // Whether the kernel carries the phase-cycle counters (Profiled<>); without them no profiling code is
    // compiled in.
    static constexpr bool profile = false;
    // Whether the kernel evaluates BSIM4 devices (WithBsim4<>).
    static constexpr bool bsim4 = false;
    template <typename V>
    static __device__ bool factor(const Ctx& C, bool act, LDS V* lu) {
        return factor_forward<V>(C, act, lu);
    }
    template <typename V>
    static __device__ void backward(const Ctx& C, bool act, LDS const V* lu) {
        s21::backward<V>(C, act, lu);
    }
};

// A Newton solve's outcome: its status and the loads (iterations) it ran.  Returned by value, as are the
// phase cycles (added to the instance's counters in global memory), so that no counter lives in private
// memory.
struct SolveOut {
    int st, loads;
};

template <bool MIXED, class LU>
__device__ SolveOut newton(const Ctx& C, bool act, bool tran, const Args& A, long long* pg, int it_left,
                           bool prev_ok) {
    const TopoGPU& T = C.T;
    const bool prof = LU::profile && A.prof != 0;
    long long t0 = 0, t1 = 0;
    int st = act ? RUNNING : S21_OK;
    int loads = 0;
    for (int it = 0; it < A.max_iter; it++) {
        if (it >= it_left) act = false;  // the segment's share of the iteration limit is used up
        if (!wave_any(act)) break;
        if (act) loads++;
        assemble<LU::bsim4>(C, act, tran, A, prof, pg);
        wave_sync();
        if (prof) t0 = cycles();
        const bool bad = seg_any(C, residual(C, act, A.iabstol));
        wave_sync();
        if (prof) { t1 = cycles(); if (pg) pg[2] += t1 - t0; t0 = t1; }
        const bool conv = act && prev_ok && !bad;
        if (wave_any(conv)) {
            commit<LU::bsim4>(C, conv);
            wave_sync();
        }
        if (conv) {
            st = S21_OK;
            act = false;
        }
        if (act && T.sym_status != S21_OK) {
            st = T.sym_status;
            act = false;
        }
        if (prof) { t1 = cycles(); if (pg) pg[6] += t1 - t0; t0 = t1; }
        if (!wave_any(act)) break;
        bool zero;
        if (MIXED) {
            for (int p = act ? C.sl : T.nnz; p < T.nnz; p += C.L) C.lu32[p] = (float)C.a[p];
            wave_sync();
            zero = seg_any(C, LU::template factor<float>(C, act, C.lu32));
        } else {
            zero = seg_any(C, LU::template factor<double>(C, act, C.a));
        }
        if (act && zero) {
            st = S21_PIVOT;
            act = false;
        }
        if (prof) { t1 = cycles(); if (pg) pg[3] += t1 - t0; t0 = t1; }
        if (MIXED)
            LU::template backward<float>(C, act, C.lu32);
        else
            LU::template backward<double>(C, act, C.a);
        if (prof) { t1 = cycles(); if (pg) pg[4] += t1 - t0; t0 = t1; }
        // clamp and update (analysis.rs:196-207)
        double mx = 0.0;
        for (int k = act ? C.sl : T.n; k < T.n; k += C.L) {
            const double dx = fabs(C.c[C.I.ce2i[k]]);
            mx = (dx > mx) ? dx : mx;
        }
        mx = seg_max(C, mx);
        bool ok = true;
        for (int k = act ? C.sl : T.n; k < T.n; k += C.L) {
            double dx = C.c[C.I.ce2i[k]];
            if (mx > 1.0) dx = dx * 1.0 / mx;
            ok = ok && !(fabs(dx) > A.reltol);
            C.x[k] = C.x[k] + dx;
        }
        prev_ok = !seg_any(C, !ok);
        wave_sync();
        if (prof) { t1 = cycles(); if (pg) pg[5] += t1 - t0; }
    }
    return SolveOut{st == RUNNING ? S21_CONVERGENCE : st, loads};
}

// ============================================================================================
// Small-signal operating point (for AC) and the AC analysis
// ============================================================================================
struct NullSink {
    __device__ void operator()(int, double) const {}
};

// After the operating point: each Mos1's conductances and Meyer capacitances at the solution (the
// op that Mos1::commit keeps, comps/mos.rs:904-907).
__device__ inline void capture_acop(const Ctx& C, bool act, const Args& A) {
    const GroupGPU& G = *C.G;
    const TopoGPU& T = C.T;
    const int nm = T.cls_start[CLS_MOS1T + 1] - T.cls_start[CLS_MOS1];
    GLB const double* par = glb(G.par);
    for (int j = act ? C.sl : nm; j < nm; j += C.L) {
        LDS const u16* dv = C.I.dev + 8 * (T.cls_start[CLS_MOS1] + j);
        auto F = [&](int f) { return par[G.m1_off[f] + (long long)C.inst * G.m1_istr[f] + j]; };
        const Mos1Core P{F(0), F(1), F(2), F(3), F(4), F(5), F(6), F(7), F(8), F(9), F(10)};
        const Mos1Caps K{F(M1_CORE_F + 0), F(M1_CORE_F + 1), F(M1_CORE_F + 2), F(M1_CORE_F + 3)};
        Mos1AcOp op;
        mos1_load<false>(P, &K, nullptr, xv(C.x, (int)dv[RD] - 1), xv(C.x, (int)dv[RG] - 1),
                         xv(C.x, (int)dv[RS] - 1), xv(C.x, (int)dv[RB] - 1), A.gmin, A.tstep, NullSink{}, nullptr,
                         &op);
        const size_t s = (size_t)G.n_inst * nm, o = (size_t)C.inst * nm + j;
        GLB double* out = glb(G.acop);
        const double v[M1_ACOP_F] = {op.gm, op.gds, op.gmbs, op.gbs, op.gbd, op.cgs, op.cgd, op.cgb, op.reversed};
        for (int f = 0; f < M1_ACOP_F; f++) out[f * s + o] = v[f];
    }
}

// Complex values with num-complex arithmetic (the reference's Complex<f64>; division by the naive
// formula, norm = hypot).
template <typename R>
struct Z {
    R re, im;
};
using zd = Z<double>;
using zf = Z<float>;
template <typename R>
__device__ inline Z<R> operator+(Z<R> a, Z<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R>
__device__ inline Z<R> operator-(Z<R> a, Z<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R>
__device__ inline Z<R> operator*(Z<R> a, Z<R> b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline zd zdiv(zd a, zd b) {
    const double ns = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / ns, (a.im * b.re - a.re * b.im) / ns};
}
__device__ inline zf zdiv(zf a, zf b) {  // FP32 factorisation of mixed precision: reciprocal of the norm
    const float r = __builtin_amdgcn_rcpf(b.re * b.re + b.im * b.im);
    return {(a.re * b.re + a.im * b.im) * r, (a.im * b.re - a.re * b.im) * r};
}
template <typename R>
__device__ inline bool zzero(Z<R> a) { return a.re == R(0) && a.im == R(0); }
__device__ inline double znorm(zd a) { return hypot(a.re, a.im); }
template <typename V>
__device__ inline V zcvt(zd a) { return V{(decltype(V::re))a.re, (decltype(V::re))a.im}; }
template <typename V>
__device__ inline zd zwide(V a) { return zd{(double)a.re, (double)a.im}; }

// Complex arrays live in LDS as interleaved (re, im) pairs.
struct AcCtx {
    LDS double *x, *c, *a, *rhs, *hub, *part;
    LDS float* lu32;   // mixed precision: complex64 copy of A, factored, then the residual's copy
    LDS double* lu64;  // FP64: copy of A, factored (A itself stays for the residual)
    LDS double* par;  // the instance's AC parameters: M1_CORE_F fields of each Mos1, then each linear value
    int nm;           // Mos1 devices
};
template <typename R>
__device__ inline Z<R> ldz(LDS const R* p, int i) {
    return Z<R>{p[2 * i], p[2 * i + 1]};
}
template <typename R>
__device__ inline void stz(LDS R* p, int i, Z<R> v) {
    p[2 * i] = v.re;
    p[2 * i + 1] = v.im;
}

// A device's complex stamps, applied as one batch like the real ones (distinct non-hub slots).
template <int N, bool SIGMA>
__device__ inline void apply_z(const AcCtx& Z_, LDS const u16* code, const double (&vr)[N], const double (&vi)[N],
                               bool rev) {
    constexpr int B = 8;
#pragma unroll
    for (int k0 = 0; k0 < N; k0 += B) {
        int c[B];
        zd cur[B];
#pragma unroll
        for (int k = 0; k < B; k++)
            if (k0 + k < N) c[k] = code[(SIGMA && rev) ? mos1_sigma(k0 + k) : k0 + k];
#pragma unroll
        for (int k = 0; k < B; k++)
            if (k0 + k < N) cur[k] = c[k] < 0x8000 ? ldz(Z_.a, c[k]) : zd{0.0, 0.0};
#pragma unroll
        for (int k = 0; k < B; k++) {
            if (k0 + k >= N) continue;
            const zd v{vr[k0 + k], vi[k0 + k]};
            if (c[k] < 0x8000)
                stz(Z_.a, c[k], cur[k] + v);
            else if (c[k] != 0xFFFF)
                stz(Z_.hub, c[k] & 0x7FFF, v);
        }
    }
}
struct AcM1Stamps {
    static constexpr bool raw_stamps = true;
    double re[MOS1_NE], im[MOS1_NE];
    bool rev;
    __device__ void stamp(int e, bool r, double a, double b) {
        re[e] = re[e] + a;
        im[e] = im[e] + b;
        rev = r;
    }
};

__device__ inline void eval_ac_mos1(const Ctx& C, const AcCtx& Z_, int j, LDS const u16* code, double omega,
                                    bool tied) {
    auto F = [&](int f) { return Z_.par[f * Z_.nm + j]; };
    const Mos1AcOp op{F(0), F(1), F(2), F(3), F(4), F(5), F(6), F(7), F(8)};
    AcM1Stamps sv;
#pragma unroll
    for (int e = 0; e < MOS1_NE; e++) sv.re[e] = sv.im[e] = 0.0;
    sv.rev = false;
    mos1_load_ac(op, F(9), F(10), omega, sv);
    if (!tied) {
        apply_z<MOS1_NE, true>(Z_, code, sv.re, sv.im, sv.rev);
        return;
    }
    double r[MOS1_NE], i[MOS1_NE];
#pragma unroll
    for (int e = 0; e < MOS1_NE; e++) {
        r[e] = sv.rev ? sv.re[mos1_sigma(e)] : sv.re[e];
        i[e] = sv.rev ? sv.im[mos1_sigma(e)] : sv.im[e];
    }
    const int from[6][4] = {{1, 7, 13, 15}, {2, 14, -1, -1}, {3, 12, -1, -1}, {5, 11, -1, -1}, {8, 10, -1, -1},
                            {23, 25, -1, -1}};
#pragma unroll
    for (int g = 0; g < 6; g++)
#pragma unroll
        for (int k = 1; k < 4; k++)
            if (from[g][k] >= 0) {
                r[from[g][0]] = r[from[g][0]] + r[from[g][k]];
                i[from[g][0]] = i[from[g][0]] + i[from[g][k]];
            }
    apply_z<MOS1_NE, false>(Z_, code, r, i, false);
}

__device__ inline void eval_ac_lin(const Ctx& C, const AcCtx& Z_, int j, LDS const u16* dv, LDS const u16* code,
                                   double omega) {
    const int typ = dv[6] >> 4;
    const double val = Z_.par[M1_CORE_F * Z_.nm + j];
    double r[6] = {0, 0, 0, 0, 0, 0}, i[6] = {0, 0, 0, 0, 0, 0};
    if (typ == S21_COMP_R) {  // comps/mod.rs:311-322
        r[0] = val;
        r[1] = val;
        r[2] = -val;
        r[3] = -val;
    } else if (typ == S21_COMP_C) {  // comps/mod.rs:223-238
        const double wc = omega * val;
        i[0] = wc;
        i[1] = wc;
        i[2] = -wc;
        i[3] = -wc;
    } else {  // Vsrc, comps/mod.rs:139-150
        r[0] = 1.0;
        r[1] = 1.0;
        r[2] = -1.0;
        r[3] = -1.0;
        r[4] = val;
    }
    apply_z<6, false>(Z_, code, r, i, false);
}

// Complex A|RHS at angular frequency omega (Solver<Complex>::update, analysis.rs:237-251).
__device__ __attribute__((always_inline)) inline void assemble_ac(const Ctx& C, const AcCtx& Z_, bool act,
                                                                   double omega) {
    const TopoGPU& T = C.T;
    const Idx& I = C.I;
    const int ns = T.nnz + T.n;
    for (int k = act ? C.sl : ns; k < ns; k += C.L) stz(Z_.a, k, zd{0.0, 0.0});
    wave_sync();
    for (int cc = 0; cc < T.n_col; cc++) {
        const int p0 = I.col[cc], p1 = I.col[cc + 1];
        LDS const u16* cd = I.col + T.n_col + 1 + 3 * cc;
        const int cls = cd[0], j0 = cd[1], code0 = cd[2];
        const int stride = code_stride(cls);
        for (int p = act ? p0 + C.sl : p1; p < p1; p += C.L) {
            LDS const u16* code = I.code + code0 + (p - p0) * stride;
            const int j = j0 + (p - p0);
            if (cls_is_mos1(cls))
                eval_ac_mos1(C, Z_, j, code, omega, cls == CLS_MOS1T);
            else
                eval_ac_lin(C, Z_, j, I.dev + 8 * p, code, omega);
        }
        wave_sync();
    }
    for (int t = act ? C.sl : T.n_gt; t < T.n_gt; t += C.L) {
        const R4 r = r4(I.gt[t]);
        zd acc{0.0, 0.0};
        for (int k = r.y; k < r.z; k++) acc = acc + ldz(Z_.hub, I.gidx[k]);
        if (r.w == 0xFFFF)
            stz(Z_.a, r.x, acc);
        else
            stz(Z_.part, r.w, acc);
    }
    wave_sync();
    for (int t = act ? C.sl : T.n_gc; t < T.n_gc; t += C.L) {
        const R4 r = r4(I.gc[t]);
        zd acc = ldz(Z_.part, r.y);
        for (int p = r.y + 1; p < r.z; p++) acc = acc + ldz(Z_.part, p);
        stz(Z_.a, r.x, acc);
    }
}

// Complex residual; true when this lane saw |res| >= 1e-9 (the AC test is strict, analysis.rs:268-270).
__device__ inline bool residual_ac(const Ctx& C, const AcCtx& Z_, bool act) {
    const Idx& I = C.I;
    bool bad = false;
    for (int t = act ? C.sl : C.T.n_rt; t < C.T.n_rt; t += C.L) {
        const R4 r = r4(I.rt[t]);
        zd acc{0.0, 0.0};
        for (int e = r.y; e < r.z; e++) {
            const R2 q = r2(I.re[e]);
            acc = acc + ldz(Z_.a, q.x) * ldz(Z_.x, q.y);
        }
        if (r.w == 0xFFFF) {
            const zd v = ldz(Z_.rhs, I.ri2e[r.x]) - acc;
            stz(Z_.c, r.x, v);
            bad = bad || !(znorm(v) < 1e-9);
        } else {
            stz(Z_.part, r.w, acc);
        }
    }
    wave_sync();
    for (int t = act ? C.sl : C.T.n_rc; t < C.T.n_rc; t += C.L) {
        const R4 r = r4(I.rc[t]);
        zd acc = ldz(Z_.part, r.y);
        for (int p = r.y + 1; p < r.z; p++) acc = acc + ldz(Z_.part, p);
        const zd v = ldz(Z_.rhs, I.ri2e[r.x]) - acc;
        stz(Z_.c, r.x, v);
        bad = bad || !(znorm(v) < 1e-9);
    }
    return bad;
}

template <typename R>
__device__ inline bool factor_forward_ac(const Ctx& C, const AcCtx& Z_, bool act, LDS R* lu) {
    using V = Z<R>;
    const Idx& I = C.I;
    for (int l = 0; l < C.T.n_clvl; l++) {
        const int e = I.clvl[l + 1];
        for (int ch = act ? I.clvl[l] + C.sl : e; ch < e; ch += C.L) {
            const R4 h = r4(I.fch[ch]);
            V y = zcvt<V>(ldz(Z_.c, h.x));
            int u = h.w;
            for (int j = h.y; j < h.z; j++) {
                const R4 L = r4(I.fl[j]);
                const V lv = zdiv(ldz(lu, L.x), ldz(lu, L.y));
                stz(lu, L.x, lv);
                const V ck = zcvt<V>(ldz(Z_.c, L.z));
                if (!zzero(ck)) y = y - ck * lv;
                for (int m = 0; m < L.w; m++, u++) {
                    const R2 q = r2(I.fu[u]);
                    stz(lu, q.y, ldz(lu, q.y) - ldz(lu, q.x) * lv);
                }
            }
            stz(Z_.c, h.x, zwide(y));
        }
        wave_sync();
    }
    bool zero_pivot = false;
    for (int i = act ? C.sl : C.T.n; i < C.T.n - 1; i += C.L)
        if (zzero(ldz(lu, I.diag[i]))) zero_pivot = true;
    return zero_pivot;
}

// The forward sweep of factor_forward_ac alone, on factors it has left in lu: the same records and
// operations, so a matrix factored once serves every Newton iteration at one frequency (the AC matrix
// does not depend on x).
template <typename R>
__device__ inline void forward_ac(const Ctx& C, const AcCtx& Z_, bool act, LDS const R* lu) {
    using V = Z<R>;
    const Idx& I = C.I;
    for (int l = 0; l < C.T.n_clvl; l++) {
        const int e = I.clvl[l + 1];
        for (int ch = act ? I.clvl[l] + C.sl : e; ch < e; ch += C.L) {
            const R4 h = r4(I.fch[ch]);
            V y = zcvt<V>(ldz(Z_.c, h.x));
            for (int j = h.y; j < h.z; j++) {
                const R4 L = r4(I.fl[j]);
                const V ck = zcvt<V>(ldz(Z_.c, L.z));
                if (!zzero(ck)) y = y - ck * ldz(lu, L.x);
            }
            stz(Z_.c, h.x, zwide(y));
        }
        wave_sync();
    }
}

template <typename R>
__device__ inline void backward_ac(const Ctx& C, const AcCtx& Z_, bool act, LDS const R* lu) {
    using V = Z<R>;
    const Idx& I = C.I;
    for (int l = 0; l < C.T.n_blvl; l++) {
        const int e = I.blvl[l + 1];
        for (int r = act ? I.blvl[l] + C.sl : e; r < e; r += C.L) {
            const R4 h = r4(I.bro[r]);
            V y = zcvt<V>(ldz(Z_.c, h.x));
            for (int k = h.z; k < h.w; k++) {
                const R2 q = r2(I.bu[k]);
                y = y - zcvt<V>(ldz(Z_.c, q.y)) * ldz(lu, q.x);
            }
            stz(Z_.c, h.x, zwide(zdiv(y, ldz(lu, h.y))));
        }
        wave_sync();
    }
}

// Complex Newton at one frequency (Solver<Complex>::solve, analysis.rs:253-304): at most 20 iterations,
// converged when every |dx| < 1e-3 and every |res| < 1e-9, |dx| clamped to 1.  The reference assembles
// and factors at every iteration; the AC matrix and right-hand side do not depend on x, so every
// iteration would assemble the same values and factor them into the same LU: the first iteration does
// both (the factors into a copy of A, which the residual keeps reading) and later iterations reuse them.
// Inlined by force: a call would bring the function-call convention's private stack into the kernels.
template <bool MIXED>
__device__ __attribute__((always_inline)) SolveOut newton_ac(const Ctx& C, const AcCtx& Z_, bool act, double omega) {
    const TopoGPU& T = C.T;
    int st = act ? RUNNING : S21_OK;
    int loads = 0;
    bool prev_ok = true;  // dx starts at zero
    for (int it = 0; it < 20; it++) {
        if (!wave_any(act)) break;
        if (act) loads++;
        if (it == 0) assemble_ac(C, Z_, act, omega);
        wave_sync();
        const bool bad = seg_any(C, residual_ac(C, Z_, act));
        wave_sync();
        const bool conv = act && prev_ok && !bad;
        if (conv) {  // Component::commit changes nothing the next frequency reads
            st = S21_OK;
            act = false;
        }
        if (act && T.sym_status != S21_OK) {
            st = T.sym_status;
            act = false;
        }
        if (!wave_any(act)) break;
        bool zero = false;
        if (it == 0) {
            if (MIXED) {
                for (int p = act ? C.sl : T.nnz; p < T.nnz; p += C.L) stz(Z_.lu32, p, zcvt<zf>(ldz(Z_.a, p)));
                wave_sync();
                zero = seg_any(C, factor_forward_ac<float>(C, Z_, act, Z_.lu32));
            } else {
                for (int p = act ? C.sl : T.nnz; p < T.nnz; p += C.L) stz(Z_.lu64, p, ldz(Z_.a, p));
                wave_sync();
                zero = seg_any(C, factor_forward_ac<double>(C, Z_, act, Z_.lu64));
            }
        } else if (MIXED) {
            forward_ac<float>(C, Z_, act, Z_.lu32);
        } else {
            forward_ac<double>(C, Z_, act, Z_.lu64);
        }
        if (act && zero) {
            st = S21_PIVOT;
            act = false;
        }
        if (MIXED)
            backward_ac<float>(C, Z_, act, Z_.lu32);
        else
            backward_ac<double>(C, Z_, act, Z_.lu64);
        double mx = 0.0;
        for (int k = act ? C.sl : T.n; k < T.n; k += C.L) {
            const double d = znorm(ldz(Z_.c, C.I.ce2i[k]));
            mx = (d > mx) ? d : mx;
        }
        mx = seg_max(C, mx);
        bool ok = true;
        for (int k = act ? C.sl : T.n; k < T.n; k += C.L) {
            zd dx = ldz(Z_.c, C.I.ce2i[k]);
            if (mx > 1.0) dx = zd{dx.re * 1.0 / mx, dx.im * 1.0 / mx};
            ok = ok && znorm(dx) < 1e-3;
            stz(Z_.x, k, ldz(Z_.x, k) + dx);
        }
        prev_ok = !seg_any(C, !ok);
        wave_sync();
    }
    return SolveOut{st == RUNNING ? S21_CONVERGENCE : st, loads};
}

// The complex Newton of newton_ac for every instance of the workgroup in lockstep, with a factorisation
// run by one lane per instance (LU::factor_lane / backward_lane on the interleaved complex arrays; the
// mixed-precision copies of A and of the residual are complex64).
template <bool MIXED, class LU>
__device__ SolveOut newton_ac_wg(const Ctx& C, const AcCtx& Z_, bool act, double omega, LDS int* s_act, int slot,
                                 int n_slots, LDS char* region0, int stride, int gidx) {
    const TopoGPU& T = C.T;
    int st = act ? RUNNING : S21_OK;
    int loads = 0;
    bool prev_ok = true;  // dx starts at zero
    for (int it = 0; it < 20; it++) {
        if (!__syncthreads_or(act)) break;
        if (act) loads++;
        if (it == 0) assemble_ac(C, Z_, act, omega);  // the same matrix at every iteration (newton_ac)
        wave_sync();
        const bool bad = seg_any(C, residual_ac(C, Z_, act));
        wave_sync();
        const bool conv = act && prev_ok && !bad;
        if (conv) {
            st = S21_OK;
            act = false;
        }
        if (act && T.sym_status != S21_OK) {
            st = T.sym_status;
            act = false;
        }
        if (MIXED) {
            if (it == 0)
                for (int p = act ? C.sl : T.nnz; p < T.nnz; p += C.L) stz(Z_.lu32, p, zcvt<zf>(ldz(Z_.a, p)));
            for (int k = act ? C.sl : T.n; k < T.n; k += C.L) stz(Z_.lu32 + 2 * T.nnz, k, zcvt<zf>(ldz(Z_.c, k)));
        } else if (it == 0) {
            for (int p = act ? C.sl : T.nnz; p < T.nnz; p += C.L) stz(Z_.lu64, p, ldz(Z_.a, p));
        }
        if (C.sl == 0) s_act[slot] = act ? 1 : 0;
        __syncthreads();
        for (int sl = threadIdx.x; sl < n_slots; sl += blockDim.x) {
            if (!s_act[sl]) continue;
            LDS double* c = (LDS double*)(region0 + (size_t)sl * stride) + 2 * T.n;
            LDS double* a = c + 2 * T.n;
            LDS double* lu_end = a + 2 * (T.nnz + T.n + T.n_hub + T.n_part);  // the LU copy follows them
            bool z = false;
            if (MIXED) {
                LDS float* lu = (LDS float*)lu_end;
                if (it == 0)
                    z = LU::template factor_lane<float>(gidx, lu, lu + 2 * T.nnz);
                else
                    LU::template forward_lane<float>(gidx, lu, lu + 2 * T.nnz);
                if (!z) LU::template backward_lane<float>(gidx, lu, lu + 2 * T.nnz);
            } else {
                if (it == 0)
                    z = LU::template factor_lane<double>(gidx, lu_end, c);
                else
                    LU::template forward_lane<double>(gidx, lu_end, c);
                if (!z) LU::template backward_lane<double>(gidx, lu_end, c);
            }
            s_act[sl] = z ? 2 : 1;
        }
        __syncthreads();
        if (act && s_act[slot] == 2) {
            st = S21_PIVOT;
            act = false;
        }
        if (MIXED) {
            for (int k = act ? C.sl : T.n; k < T.n; k += C.L) stz(Z_.c, k, zwide(ldz(Z_.lu32 + 2 * T.nnz, k)));
            wave_sync();
        }
        double mx = 0.0;
        for (int k = act ? C.sl : T.n; k < T.n; k += C.L) {
            const double d = znorm(ldz(Z_.c, C.I.ce2i[k]));
            mx = (d > mx) ? d : mx;
        }
        mx = seg_max(C, mx);
        bool ok = true;
        for (int k = act ? C.sl : T.n; k < T.n; k += C.L) {
            zd dx = ldz(Z_.c, C.I.ce2i[k]);
            if (mx > 1.0) dx = zd{dx.re * 1.0 / mx, dx.im * 1.0 / mx};
            ok = ok && znorm(dx) < 1e-3;
            stz(Z_.x, k, ldz(Z_.x, k) + dx);
        }
        prev_ok = !seg_any(C, !ok);
        wave_sync();
    }
    return SolveOut{st == RUNNING ? S21_CONVERGENCE : st, loads};
}

// One Newton solve of every instance of the workgroup in lockstep, for an LU policy that factors one
// lane per instance: the instances' own lanes assemble, test convergence and update; between two
// barriers the first n_slots threads of the workgroup each factor and sweep one instance's matrix,
// so one instruction stream serves all of them.  s_act: per-slot flag words in LDS.
template <bool MIXED, class LU>
__device__ SolveOut newton_wg(const Ctx& C, bool act, bool tran, const Args& A, long long* pg, int it_left,
                              bool prev_ok, LDS int* s_act, int slot, int n_slots, LDS char* region0, int stride,
                              int gidx) {
    const TopoGPU& T = C.T;
    const bool prof = LU::profile && A.prof != 0;
    long long t0 = 0, t1 = 0;
    int st = act ? RUNNING : S21_OK;
    int loads = 0;
    for (int it = 0; it < A.max_iter; it++) {
        if (it >= it_left) act = false;
        if (!__syncthreads_or(act)) break;
        if (act) loads++;
        assemble<LU::bsim4>(C, act, tran, A, prof, pg);
        wave_sync();
        if (prof) t0 = cycles();
        const bool bad = seg_any(C, residual(C, act, A.iabstol));
        wave_sync();
        const bool conv = act && prev_ok && !bad;
        if (wave_any(conv)) {
            commit<LU::bsim4>(C, conv);
            wave_sync();
        }
        if (conv) {
            st = S21_OK;
            act = false;
        }
        if (act && T.sym_status != S21_OK) {
            st = T.sym_status;
            act = false;
        }
        if (MIXED) {  // FP32 copies of A and of the residual for the factorisation lanes
            for (int p = act ? C.sl : T.nnz; p < T.nnz; p += C.L) C.lu32[p] = (float)C.a[p];
            for (int k = act ? C.sl : T.n; k < T.n; k += C.L) C.lu32[T.nnz + k] = (float)C.c[k];
        }
        if (C.sl == 0) s_act[slot] = act ? 1 : 0;
        if (prof) { t1 = cycles(); if (pg) pg[2] += t1 - t0; t0 = t1; }
        __syncthreads();
        for (int sl = threadIdx.x; sl < n_slots; sl += blockDim.x) {
            if (!s_act[sl]) continue;
            LDS double* c = (LDS double*)(region0 + (size_t)sl * stride) + T.n;
            LDS double* a = c + T.n;
            bool z;
            if (MIXED) {
                LDS float* lu = (LDS float*)(a + T.nnz + T.n + T.n_hub + T.n_part);
                z = LU::template factor_lane<float>(gidx, lu, lu + T.nnz);
                if (!z) LU::template backward_lane<float>(gidx, lu, lu + T.nnz);
            } else {
                z = LU::template factor_lane<double>(gidx, a, c);
                if (!z) LU::template backward_lane<double>(gidx, a, c);
            }
            s_act[sl] = z ? 2 : 1;
        }
        __syncthreads();
        if (prof) { t1 = cycles(); if (pg) pg[3] += t1 - t0; t0 = t1; }
        if (act && s_act[slot] == 2) {  // a pivot the reference's search would not take (factor_forward)
            st = S21_PIVOT;
            act = false;
        }
        if (MIXED) {
            for (int k = act ? C.sl : T.n; k < T.n; k += C.L) C.c[k] = (double)C.lu32[T.nnz + k];
            wave_sync();
        }
        // clamp and update (analysis.rs:196-207)
        double mx = 0.0;
        for (int k = act ? C.sl : T.n; k < T.n; k += C.L) {
            const double dx = fabs(C.c[C.I.ce2i[k]]);
            mx = (dx > mx) ? dx : mx;
        }
        mx = seg_max(C, mx);
        bool ok = true;
        for (int k = act ? C.sl : T.n; k < T.n; k += C.L) {
            double dx = C.c[C.I.ce2i[k]];
            if (mx > 1.0) dx = dx * 1.0 / mx;
            ok = ok && !(fabs(dx) > A.reltol);
            C.x[k] = C.x[k] + dx;
        }
        prev_ok = !seg_any(C, !ok);
        wave_sync();
        if (prof) { t1 = cycles(); if (pg) pg[5] += t1 - t0; }
    }
    return SolveOut{st == RUNNING ? S21_CONVERGENCE : st, loads};
}

// The AC sweep of newton_body: slot's instance (C.inst) for every frequency of run u (first pass), its
// re-solve from the previous run's end (second pass, run u + 1), or its in-order continuation with u the
// wavefront's look-ahead offset (third pass).  region0: the workgroup's first instance region.
template <bool MIXED, class LU>
__device__ inline void ac_sweep(Ctx& C, const Args& A, int g, int u, int slot, bool valid, LDS char* region0) {
    const GroupGPU& G = *C.G;
    const TopoGPU& T = C.T;
    const int inst = C.inst;
    LDS double* dbase = (LDS double*)(region0 + (size_t)slot * A.inst_bytes);
    AcCtx Z_;
    Z_.x = dbase;
    Z_.c = Z_.x + 2 * T.n;
    Z_.a = Z_.c + 2 * T.n;
    Z_.rhs = Z_.a + 2 * T.nnz;
    Z_.hub = Z_.rhs + 2 * T.n;
    Z_.part = Z_.hub + 2 * T.n_hub;
    Z_.lu32 = (LDS float*)(Z_.part + 2 * T.n_part);
    Z_.lu64 = Z_.part + 2 * T.n_part;  // FP64: where the complex64 copies would be
    Z_.par = MIXED ? (LDS double*)(Z_.lu32 + 2 * (T.nnz + T.n)) : Z_.lu64 + 2 * T.nnz;
    Z_.nm = mos1_count(T);
    {  // parameters read every iteration: LDS instead of global-memory latency per colour
        const int nl = T.n_cls[CLS_LIN];
        GLB const double* par = glb(G.par);
        for (int q = valid ? C.sl : M1_CORE_F * Z_.nm; q < M1_CORE_F * Z_.nm; q += C.L) {
            const int f = q / Z_.nm, j = q - f * Z_.nm;
            Z_.par[q] = par[G.m1_off[f] + (long long)inst * G.m1_istr[f] + j];
        }
        for (int j = valid ? C.sl : nl; j < nl; j += C.L)
            Z_.par[M1_CORE_F * Z_.nm + j] = par[G.lin_off[2] + (long long)inst * G.lin_istr[2] + j];
        wave_sync();
    }
    C.x = Z_.x;
    const int R = A.ac_runs;
    const long long F = A.npts;
    auto first_pt = [&](int r) { return F * r / R; };
    GLB int* rec = glb(G.ac_rec) + (size_t)inst * F;
    GLB double* wv = glb(G.wave) + (size_t)inst * F * 2 * T.n;
    GLB double* ends = glb(G.ac_end) + (size_t)inst * R * 2 * T.n;
    auto load_x = [&](bool act, GLB const double* src) {
        for (int k = act ? C.sl : T.n; k < T.n; k += C.L) stz(Z_.x, k, zd{src[2 * k], src[2 * k + 1]});
        wave_sync();
    };
    auto store_x = [&](bool act, GLB double* dst) {
        for (int k = act ? C.sl : T.n; k < T.n; k += C.L) {
            const zd v = ldz(Z_.x, k);
            dst[2 * k] = v.re;
            dst[2 * k + 1] = v.im;
        }
    };
    // Loops run in step across the workgroup when one lane per instance factors (its barriers).
    auto any = [&](bool p) {
        if constexpr (LU::lane_per_instance)
            return __syncthreads_or(p) != 0;
        else
            return wave_any(p);
    };
    long long iters = 0;
    // Newton at frequency pt from the current x; records the point and returns its status.
    auto point = [&](bool act, long long pt, int& its) {
        const long long i0 = iters;
        const double om = act ? A.omega[pt] : 0.0;
        SolveOut r;
        if constexpr (LU::lane_per_instance) {
            __shared__ int s_act[WAVE * MAX_WAVES];
            r = newton_ac_wg<MIXED, LU>(C, Z_, act, om, (LDS int*)s_act, slot, A.waves * A.ipw, region0,
                                        A.inst_bytes, g);
        } else {
            r = newton_ac<MIXED>(C, Z_, act, om);
        }
        const int s2 = r.st;
        iters += r.loads;
        its = (int)(iters - i0);
        if (act && C.sl == 0) rec[pt] = (s2 << 8) | its;
        if (act && s2 == S21_OK) store_x(true, wv + (size_t)pt * 2 * T.n);
        return s2;
    };
    // Any frequency of [p, e) failed or unsolved (the lanes of the segment share the scan).
    auto failed_in = [&](long long p, long long e) {
        bool f = false;
        for (long long q = p + C.sl; q < e; q += C.L) f = f || rec[q] == 0 || (rec[q] >> 8) != S21_OK;
        return seg_any(C, f);
    };
    // A frequency solved (at least one full Newton step) on both paths: both hold its unique solution.
    auto meets = [&](int its, int old) {
        return its >= 2 && old != 0 && (old >> 8) == S21_OK && (old & 0xFF) >= 2;
    };
    if (A.ac_pass == 0) {  // run u from x = 0 (analysis.rs:215-233)
        bool alive = valid;
        for (int k = valid ? C.sl : T.n; k < T.n; k += C.L) stz(Z_.x, k, zd{0.0, 0.0});
        wave_sync();
        const long long p0 = first_pt(u), p1 = first_pt(u + 1);
        for (long long pt = p0;; pt++) {
            const bool here = alive && pt < p1;
            if (!any(here)) break;
            int its;
            const int s2 = point(here, pt, its);
            if (here) alive = s2 == S21_OK;
        }
        store_x(valid, ends + (size_t)u * 2 * T.n);
    } else if (A.ac_pass == 1) {  // run r from the first pass's end of run r - 1, until the paths meet
        const int r = u + 1;
        bool live = valid;
        load_x(live, ends + (size_t)(r - 1) * 2 * T.n);
        int state = AC_UNMET;
        const long long p0 = first_pt(r), p1 = first_pt(r + 1);
        for (long long pt = p0;; pt++) {
            const bool here = live && pt < p1;
            if (!any(here)) break;
            const int old = here ? rec[pt] : 0;
            int its;
            const int s2 = point(here, pt, its);
            if (here && s2 != S21_OK) {
                state = AC_FAILED;
                live = false;
            } else if (here && meets(its, old)) {
                state = AC_MET;
                live = false;
            }
        }
        if (valid && C.sl == 0) G.ac_state[(size_t)inst * R + r] = state;
    } else {
        // In order from run rs, jumping to the next run wherever the paths meet.  The A.ac_look
        // wavefronts of the workgroup serve one instance: from the current solution x, wavefront w
        // runs iteration 0 of frequency k + w (assembly and residual: the reference keeps x when
        // every |res| < 1e-9, since dx starts at zero); the first frequency that does not keep x
        // is solved by wavefront 0, and the ones before it keep x.
        __shared__ int s_flag[MAX_WAVES];
        __shared__ long long s_out[4];
        const int w = u;                       // this wavefront's offset
        const int lead = slot - w;             // the instance's first slot in the workgroup
        LDS double* x_lead = (LDS double*)(region0 + (size_t)lead * A.inst_bytes);
        const int rs = valid ? G.ac_state[(size_t)inst * R] : 0;
        bool live = valid && rs > 0;
        long long k = live ? first_pt(rs) : 1;
        load_x(live, wv + (size_t)(k - 1) * 2 * T.n);
        // The workgroup is this instance's A.ac_look wavefronts, so live is the same in all of them.
        while (__syncthreads_or(live)) {
            const long long kw = k + w;
            const bool test = live && kw < F;
            bool keep = false;
            if (wave_any(test)) {
                assemble_ac(C, Z_, test, test ? A.omega[kw] : 0.0);
                wave_sync();
                const bool bad = seg_any(C, residual_ac(C, Z_, test));
                wave_sync();
                keep = test && !bad;
            }
            if (C.sl == 0) s_flag[slot] = keep ? 1 : 0;
            __syncthreads();
            int ws = A.ac_look;  // first offset that does not keep x
            for (int j = 0; j < A.ac_look; j++)
                if (!s_flag[lead + j]) {
                    ws = j;
                    break;
                }
            if (live && w < ws) {  // kept: x stands, one iteration
                if (C.sl == 0) rec[kw] = (S21_OK << 8) | 1;
                store_x(true, wv + (size_t)kw * 2 * T.n);
            }
            const long long ks = k + ws;
            const bool solve = live && ws < A.ac_look && ks < F;
            const int old = solve && w == 0 ? rec[ks] : 0;  // before wavefront 0 records ks
            int its = 0;
            const int s2 = point(solve && w == 0, ks, its);
            if (solve && w == 0 && C.sl == 0) {
                s_out[0] = s2;
                s_out[1] = its;
                s_out[2] = old;
            }
            __syncthreads();
            if (live) {
                long long next = ks + 1;
                if (!solve) {
                    next = ks;  // every tested frequency kept x (or the sweep ended)
                } else if (s_out[0] != S21_OK) {
                    live = false;  // the sweep stops at a failed frequency (analysis.rs:801)
                } else if (meets((int)s_out[1], (int)s_out[2])) {
                    int rr = (int)((ks * R) / F);
                    while (first_pt(rr + 1) <= ks) rr++;
                    while (first_pt(rr) > ks) rr--;
                    // The rest of run rr is stored data of a path through frequency ks's solution:
                    // the sweep's (a failure there is the sweep's).
                    bool bad = failed_in(ks + 1, first_pt(rr + 1));
                    int q = rr + 1;
                    // Run rr ends as the first pass ended it unless its re-solve never met that
                    // pass; then every following run whose re-solve met is the sweep's, and so is a
                    // run whose re-solve ran to its end (its start was the sweep's) or failed (the
                    // sweep fails there).
                    if (rr == 0 || G.ac_state[(size_t)inst * R + rr] != AC_UNMET) {
                        while (!bad && q < R && G.ac_state[(size_t)inst * R + q] == AC_MET) {
                            bad = failed_in(first_pt(q), first_pt(q + 1));
                            q++;
                        }
                        if (!bad && q < R) {
                            if (G.ac_state[(size_t)inst * R + q] == AC_FAILED)
                                bad = true;
                            else
                                q++;
                        }
                    }
                    next = first_pt(q);
                    if (bad) live = false;
                }
                if (next >= F) live = false;
                if (live && solve) {  // every wavefront continues from the sweep's solution at next - 1
                    if (next - 1 == ks)
                        for (int j = C.sl; j < 2 * T.n; j += C.L) Z_.x[j] = x_lead[j];
                    else  // stored by an earlier launch
                        load_x(true, wv + (size_t)(next - 1) * 2 * T.n);
                    wave_sync();
                }
                k = next;
            }
            __syncthreads();  // wavefront 0's x is read before its next solve overwrites it
        }
    }
}

// The operating point and transient of newton_body for slot's instance (C.inst).
template <bool MIXED, class LU>
__device__ inline void op_tran(Ctx& C, const Args& A, int g, int slot, bool valid, LDS char* region0) {
    const GroupGPU& G = *C.G;
    const TopoGPU& T = C.T;
    const int inst = C.inst;
    LDS double* dbase = (LDS double*)(region0 + (size_t)slot * A.inst_bytes);
    C.x = dbase;
    C.c = C.x + T.n;
    C.a = C.c + T.n;
    C.rhs = C.a + T.nnz;
    C.hub = C.rhs + T.n;
    C.part = C.hub + T.n_hub;
    C.lu32 = (LDS float*)(C.part + T.n_part);

    for (int k = valid ? C.sl : T.n; k < T.n; k += C.L) C.x[k] = G.x[(size_t)inst * T.n + k];
    wave_sync();

    if (A.mode == MODE_ASSEMBLE) {
        assemble<LU::bsim4>(C, valid, A.asm_tran != 0, A);
        wave_sync();
        for (int s = valid ? C.sl : T.nnz + T.n; s < T.nnz + T.n; s += C.L) G.dbg[s] = C.a[s];
        return;
    }

    // phase cycle counters of this instance (profiling mode): one lane adds to them in global memory
    long long* const pg =
        (LU::profile && A.prof && G.prof && valid && C.sl == 0) ? G.prof + (size_t)inst * NPROF : nullptr;
    if (pg)
        for (int k = 0; k < NPROF; k++) pg[k] = 0;
    const long long tstart = LU::profile && A.prof ? cycles() : 0;
    long long iters = 0;
    int pt0 = 0, used0 = 0;  // a resumed instance: the solve it continues and the iterations that solve used
    if (A.resume && valid) {
        iters = G.iters[inst];
        pt0 = G.resume_at[2 * inst];
        used0 = G.resume_at[2 * inst + 1];
    }
    int last_pt = 0;
    long long last_start = 0;  // iteration count when the last solve began
    auto solve = [&](bool act, bool tran, int left, bool prev_ok) {
        SolveOut r;
        if constexpr (LU::lane_per_instance) {
            __shared__ int s_act[WAVE * MAX_WAVES];
            r = newton_wg<MIXED, LU>(C, act, tran, A, pg, left, prev_ok, (LDS int*)s_act, slot, A.waves * A.ipw,
                                     region0, A.inst_bytes, g);
        } else {
            r = newton<MIXED, LU>(C, act, tran, A, pg, left, prev_ok);
        }
        iters += r.loads;
        return r.st;
    };
    last_start = iters - used0;
    const bool op_here = pt0 == 0;  // a resumed instance may start inside the time loop
    int st = solve(valid && op_here, false, A.max_iter - (A.resume ? used0 : 0), !A.resume);
    if (A.capture && G.acop) capture_acop(C, valid && st == S21_OK, A);
    if (A.mode == MODE_TRAN) {
        bool alive = valid && st == S21_OK;
        double* wv = G.wave + (size_t)inst * (size_t)A.npts * T.n;
        for (int k = alive && op_here ? C.sl : T.n; k < T.n; k += C.L) wv[k] = C.x[k];
        // t = tstep, 2 tstep, ... while t < tstop (analysis.rs:553-570)
        for (long long pt = 1; pt < A.npts; pt++) {
            if (LU::lane_per_instance ? !__syncthreads_or(alive) : !wave_any(alive)) break;
            const bool here = alive && pt >= pt0;   // a resumed instance joins at its own point
            const bool cont = A.resume && pt == pt0;  // ... with the solve it continues
            if (here) {
                last_pt = (int)pt;
                last_start = cont ? iters - used0 : iters;
            }
            const int s2 = solve(here, true, cont ? A.max_iter - used0 : A.max_iter, !cont);
            if (here) {
                st = s2;
                alive = s2 == S21_OK;
            }
            for (int k = here && alive ? C.sl : T.n; k < T.n; k += C.L) wv[(size_t)pt * T.n + k] = C.x[k];
        }
    }
    if (!valid) return;
    for (int k = C.sl; k < T.n; k += C.L) G.x[(size_t)inst * T.n + k] = C.x[k];
    if (C.sl == 0) {
        G.status[inst] = st;
        G.iters[inst] = iters;
        if (G.resume_at) {
            G.resume_at[2 * inst] = last_pt;
            G.resume_at[2 * inst + 1] = (int)(iters - last_start);
        }
        if (pg) pg[7] = cycles() - tstart;
    }
}

// Workgroup = A.waves wavefronts x A.ipw instances, sharing one LDS copy of the group's records.
template <bool MIXED, class LU>
__device__ inline void newton_body(const GroupGPU* __restrict__ groups, int n_groups, const Args& A) {
    int g = 0;
    const int b = blockIdx.x;
    while (g + 1 < n_groups && groups[g + 1].blk0 <= b) g++;
    const GroupGPU& G = groups[g];
    Ctx C;
    C.G = &G;
    C.T = G.t;
    const TopoGPU& T = C.T;
    extern __shared__ uint4 smem_u4[];
    {  // records -> LDS (whole workgroup)
        GLB const uint64_t* src = (GLB const uint64_t*)T.blob;
        LDS uint64_t* dst = (LDS uint64_t*)smem_u4;
        for (int k = threadIdx.x; k < 2 * T.blob_vec; k += blockDim.x) dst[k] = src[k];
    }
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int lane = threadIdx.x % WAVE;
    C.L = WAVE / A.ipw;
    const int sub = lane / C.L;
    C.sl = lane % C.L;
    C.seg0 = sub * C.L;
    const int slot = w * A.ipw + sub;  // instance slot within the workgroup
    int inst;
    bool valid;
    if (A.mode == MODE_ASSEMBLE) {
        inst = A.dbg_inst;
        valid = slot == 0;
    } else {
        inst = (b - G.blk0) * A.waves * A.ipw + slot;  // AC: the run's index, instance-major
        valid = inst < G.n_inst * ac_units(A);
        if (A.mode == MODE_AC) inst /= ac_units(A);
    }
    // A workgroup-synchronous solve needs every wavefront at every barrier, instance or not.
    if (!LU::lane_per_instance && !wave_any(valid)) return;
    if (!valid) inst = 0;  // idle lanes touch no instance data
    C.inst = inst;
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
    C.I.col = b16 + T.o_col;
    C.I.dev = b16 + T.o_dev;
    C.I.code = b16 + T.o_code;
    LDS char* region0 = (LDS char*)((LDS uint4*)smem_u4 + T.blob_vec);
    if (A.mode == MODE_AC) {
        if constexpr (LU::kind != 1 && !LU::bsim4)  // BSIM4 has no AC stamps (no load_ac in the reference)
            ac_sweep<MIXED, LU>(C, A, g, valid ? (b - G.blk0) * A.waves * A.ipw + slot - inst * ac_units(A) : 0, slot,
                                valid, region0);
    } else {
        if constexpr (LU::kind != 2) op_tran<MIXED, LU>(C, A, g, slot, valid, region0);
    }
}

// An LU policy with the phase-cycle counters compiled in (profiling launches).
template <class LU>
struct Profiled : LU {
    static constexpr bool profile = true;
};
// An LU policy whose kernel evaluates BSIM4 devices (batches that hold them).
template <class LU>
struct WithBsim4 : LU {
    static constexpr bool bsim4 = true;
};

// Precompiled kernel: the level-scheduled factorisation of the pattern's records.
template <bool MIXED>
__global__ __launch_bounds__(WAVE * MAX_WAVES) void newton_kernel(const GroupGPU* __restrict__ groups, int n_groups,
                                                                  Args A) {
    newton_body<MIXED, LuLevels>(groups, n_groups, A);
}
template <bool MIXED>
__global__ __launch_bounds__(WAVE * MAX_WAVES) void newton_kernel_prof(const GroupGPU* __restrict__ groups,
                                                                       int n_groups, Args A) {
    newton_body<MIXED, Profiled<LuLevels>>(groups, n_groups, A);
}
template <bool MIXED>
__global__ __launch_bounds__(WAVE * MAX_WAVES) void newton_kernel_b4(const GroupGPU* __restrict__ groups,
                                                                     int n_groups, Args A) {
    newton_body<MIXED, WithBsim4<LuLevels>>(groups, n_groups, A);
}

}  // namespace s21
