// Batched Newton engine for MI355X (gfx950): host runtime.  The kernel is in kernel.h; jit.cpp compiles
// variants of it with a factorisation generated for one frozen pattern.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "internal.h"
#include "jit.h"
#include "kernel.h"

namespace s21 {

// ============================================================================================
// Host runtime
// ============================================================================================
#define HIPCHECK(x)                                                                         \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) fail(S21_DEVICE, std::string("HIP: ") + hipGetErrorString(e_)); \
    } while (0)

constexpr int JIT_MIN_INST = 512;      // instances that repay a run-time compilation
constexpr int64_t JIT_MIN_AC_POINTS = 100000;  // AC: instance-frequencies that repay one
constexpr int64_t JIT_MAX_OPS = 40000;         // largest generated factorisation of one pattern
constexpr int64_t JIT_MAX_TOTAL_OPS = 200000;  // ... and of all the patterns of one batch

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
    std::vector<s21_params> mos1_models, mos1_insts, diode_models, diode_insts, bsim4_models, bsim4_insts;
    int n_inst = 0;
    std::vector<Override> overs;
    // the last analysis
    Topology topo;
    std::unique_ptr<DevMem> d_topo, d_par, d_state, d_out;
    std::unique_ptr<DevMem> d_b4;  // BSIM4 parameters and states
    int b4_var_fields = 0;          // BSIM4 parameter fields stored per instance
    std::unique_ptr<DevMem> d_acop;  // Mos1 small-signal operating points captured for AC
    std::unique_ptr<DevMem> d_ac;    // AC: per-frequency records, runs' last iterates, runs' outcomes
    GroupGPU gg{};
    std::vector<double> x;
    std::vector<int> status;
    std::vector<long long> iters;
    size_t blob_bytes = 0, inst_bytes = 0;  // record blob; LDS bytes one instance needs
    std::string prepared;  // analysis the uploaded structure was built for ("" = none)
    // Pivoting: by default the search on instance 0's first matrix (x = 0); pivot_row / pivot_col give
    // the order instead; pivot_later (matrix values in element order) adds the reference's later
    // searches on those matrices.
    std::vector<int> pivot_row, pivot_col;
    std::vector<std::vector<double>> pivot_later;
};

}  // namespace s21

struct s21_batch {
    std::vector<s21::GroupHost> groups;
    s21_options opts;
    int precision = S21_FP64;
    int device = 0;  // the HIP device current at creation, which holds every buffer of the batch
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0, kernel_avg_ms = 0.0;
    int64_t launches = 0;
    int64_t resumed = 0;  // instance resumptions of the last analysis (launches of recover_pivots)
    int64_t npts = 0;
    int profile = 0;
    bool has_b4 = false;  // some group holds BSIM4 devices: the kernels that evaluate them
    int last_waves = 1, last_ipw = 1;
    int ipw = 0;  // instances per wavefront; 0 = chosen per launch
    int jit = 0;  // generated factorisation: 0 = for large single-pattern batches, 1 = always, -1 = never
    int waves = 0;  // wavefronts per workgroup, at most this many (0 = chosen per launch)
    int hub = 0;    // hub policy: 0 = chosen per batch, 1 = every shared slot, 2 = slots of many contributors
    int ac_runs = 0;  // AC: runs of frequencies per instance; 0 = chosen per sweep
    int last_ac_runs = 0;
    bool last_jit = false;
    // kernel with every group's factorisation written out (generated at the last preparation), or none
    s21::JitKernel jitk;
    bool has_jit = false;
    std::vector<double> axis;
    std::unique_ptr<s21::DevMem> d_wave, d_groups;
    std::vector<char> groups_uploaded;  // the descriptors d_groups holds
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

// Flattened components of the group for one analysis, IC forcing appended for a transient.
static void build_comps(GroupHost& g, const s21_tran_options* tr, Topology& t) {
    build_comps(g.n_vars, g.names, g.kinds, g.comps, (int)g.mos1_models.size(), (int)g.mos1_insts.size(),
                (int)g.diode_models.size(), (int)g.diode_insts.size(), tr, t);
}

// Host half of the preparation: the frozen structure, then per-instance derived parameters in the
// kernel's device order.
struct HostPrep {
    std::vector<double> m1_core, m1_caps, d_par, lin_val, lin_tran, lin_ac;
    std::vector<b4::B4Par> b4;  // [n_b4], or [n_inst][n_b4] when an instance overrides a BSIM4 parameter
    bool b4_per_inst = false;
    int n_m1 = 0, n_d = 0, n_lin = 0, n_b4 = 0;
};
// BSIM4 parameters of one device; the reference panics where derivation throws (bsim4inst.rs:200-223).
static b4::B4Par derive_b4(const s21_params& mp, const s21_params& ip) {
    try {
        return b4::derive_params(mp, ip);
    } catch (const std::domain_error& e) {
        fail(S21_ERR, e.what());
    }
}

// Inputs of an AC structure: the first angular frequency and each instance's Mos1 operating point
// captured after the DC solution, [9][n_inst][n_dc] in that solution's device order (dc_comp: the
// component of each).
struct AcInput {
    double omega0 = 0.0;
    std::vector<double> acop;
    std::vector<int> dc_comp;
};

// Parameter sets of component `ci` for instance i with the group's overrides applied.
static void inst_params(const GroupHost& g, const std::vector<const Override*>* ov, const Comp& c, int i,
                        s21_params& mp, s21_params& ip) {
    const bool mos = c.type == S21_COMP_MOS1, b4 = c.type == S21_COMP_BSIM4;
    mp = mos ? g.mos1_models[c.model] : b4 ? g.bsim4_models[c.model] : g.diode_models[c.model];
    ip = mos ? g.mos1_insts[c.inst] : b4 ? g.bsim4_insts[c.inst] : g.diode_insts[c.inst];
    if (!ov) return;
    for (const Override* x : *ov) {
        if (x->target == S21_OVR_VALUE)
            fail(S21_INVALID, mos ? "value override on a Mos1" : b4 ? "value override on a BSIM4" : "value override on a Diode");
        s21_params& tp = x->target == S21_OVR_MODEL ? mp : ip;
        tp.v[x->field] = x->vals[i];
        tp.set |= 1ull << x->field;
    }
}

static void host_prepare(const s21_options& o, GroupHost& g, const s21_tran_options* tr, HostPrep& hp,
                         const AcInput* ac = nullptr, int hub_contrib = HUB_CONTRIB) {
    Topology& t = g.topo;
    t = Topology();
    build_comps(g, tr, t);
    t.ac = ac != nullptr;
    topo_elements(t);
    const int NI = g.n_inst;
    std::unordered_map<int, std::vector<const Override*>> ov;
    for (const Override& x : g.overs) ov[x.comp].push_back(&x);
    auto ov_of = [&](int ci) -> const std::vector<const Override*>* {
        auto it = ov.find(ci);
        return it == ov.end() ? nullptr : &it->second;
    };

    // Instance 0 in component order: the symbolic phase runs on its first Newton iteration (x = 0).
    HostInstance h0;
    std::vector<uint8_t> has_caps;
    for (size_t i = 0; i < t.comps.size(); i++) {
        const Comp& c = t.comps[i];
        s21_params mp, ip;
        if (c.type == S21_COMP_MOS1) {
            inst_params(g, ov_of((int)i), c, 0, mp, ip);
            h0.m1.push_back(derive_mos1(mp, ip, o));
            has_caps.push_back(h0.m1.back().has_caps ? 1 : 0);
        } else if (c.type == S21_COMP_BSIM4) {
            if (c.model < 0 || c.model >= (int)g.bsim4_models.size() || c.inst < 0 ||
                c.inst >= (int)g.bsim4_insts.size())
                fail(S21_ERR, "Parameters not defined");
            inst_params(g, ov_of((int)i), c, 0, mp, ip);
            h0.b4.push_back(derive_b4(mp, ip));
        } else if (c.type == S21_COMP_DIODE) {
            inst_params(g, ov_of((int)i), c, 0, mp, ip);
            h0.dio.push_back(derive_diode(mp, ip, o));
        } else if (c.type != S21_COMP_MOS0) {
            double v = c.value;
            if (const auto* L = ov_of((int)i))
                for (const Override* x : *L) {
                    if (x->target != S21_OVR_VALUE) fail(S21_INVALID, "parameter override on a linear element");
                    v = x->vals[0];
                }
            h0.lin_value.push_back(v);
        }
    }
    // A Mos1 has gate-charge state when any instance has gate capacitance.
    {
        int im = 0;
        for (size_t i = 0; i < t.comps.size(); i++) {
            const Comp& c = t.comps[i];
            if (c.type != S21_COMP_MOS1) continue;
            if (!has_caps[im] && ov_of((int)i))
                for (int k = 1; k < NI && !has_caps[im]; k++) {
                    s21_params mp, ip;
                    inst_params(g, ov_of((int)i), c, k, mp, ip);
                    if (derive_mos1(mp, ip, o).has_caps) has_caps[im] = 1;
                }
            im++;
        }
    }
    // AC: the captured operating points by component (Mos1 k in component order <- DC local index).
    std::vector<int> dc_of_m1;
    if (ac) {
        std::unordered_map<int, int> dc_pos;
        for (size_t j = 0; j < ac->dc_comp.size(); j++) dc_pos[ac->dc_comp[j]] = (int)j;
        for (size_t i = 0; i < t.comps.size(); i++)
            if (t.comps[i].type == S21_COMP_MOS1) dc_of_m1.push_back(dc_pos.at((int)i));
    }
    const int n_dc = ac ? (int)ac->dc_comp.size() : 0;
    auto acop_of = [&](int k, int inst) {  // Mos1 k (component order), instance inst
        const size_t s = (size_t)NI * n_dc, off = (size_t)inst * n_dc + dc_of_m1[k];
        const double* v = ac->acop.data();
        return Mos1AcOp{v[0 * s + off], v[1 * s + off], v[2 * s + off], v[3 * s + off], v[4 * s + off],
                        v[5 * s + off], v[6 * s + off], v[7 * s + off], v[8 * s + off]};
    };
    if (ac) {  // pivot on instance 0's first AC matrix (analysis.rs:761-832: fstart, x = 0)
        std::vector<Mos1AcOp> op0;
        std::vector<double> grd0, grs0, lin0;
        int k = 0, il = 0;
        for (const Comp& c : t.comps) {
            if (c.type == S21_COMP_MOS1) {
                op0.push_back(acop_of(k, 0));
                grd0.push_back(h0.m1[k].core.grd);
                grs0.push_back(h0.m1[k].core.grs);
                k++;
            } else if (c.type == S21_COMP_R || c.type == S21_COMP_C || c.type == S21_COMP_V) {
                lin0.push_back(c.type == S21_COMP_V ? c.acm : h0.lin_value[il]);
                il++;
            } else if (c.type == S21_COMP_I) {
                il++;
            }
        }
        std::vector<double> re, im;
        host_assemble_ac(t, op0, grd0, grs0, lin0, ac->omega0, re, im);
        topo_pivot_cx(t, re, im);
    } else if (!g.pivot_row.empty()) {
        topo_pivot_forced(t, g.pivot_row, g.pivot_col);
    } else {
        std::vector<double> x0(t.n, 0.0), vals, rhs;
        host_assemble_op(t, h0, o, x0, vals, rhs);
        if (!g.pivot_later.empty())
            topo_pivot_sequence(t, vals, g.pivot_later);
        else
            topo_pivot(t, vals);
    }
    topo_schedule(t, has_caps, hub_contrib);

    // Per-instance parameters, SoA [field][instance][class-local device] in the kernel's device order.
    const int n_m1 = t.cls_start[CLS_MOS1T + 1] - t.cls_start[CLS_MOS1];  // both Mos1 classes
    const int n_d = t.cls_start[CLS_DIODE + 1] - t.cls_start[CLS_DIODE];
    const int n_lin = t.cls_start[CLS_LIN + 1] - t.cls_start[CLS_LIN];
    hp.n_m1 = n_m1;
    hp.n_d = n_d;
    hp.n_lin = n_lin;
    hp.m1_core.assign((size_t)M1_CORE_F * NI * n_m1, 0.0);
    hp.m1_caps.assign((size_t)M1_CAPS_F * NI * n_m1, 0.0);
    for (int d = 0; d < n_m1; d++) {
        const int ci = t.dev_comp[t.cls_start[CLS_MOS1] + d];
        const Comp& c = t.comps[ci];
        const auto* L = ov_of(ci);
        const size_t s = (size_t)NI * n_m1;
        Mos1Derived D{};
        int k_comp = 0;  // Mos1 index of this device in component order
        if (ac)
            for (int c2 = 0; c2 < ci; c2++) k_comp += t.comps[c2].type == S21_COMP_MOS1;
        for (int i = 0; i < NI; i++) {
            if (i == 0 || L) {
                s21_params mp, ip;
                inst_params(g, L, c, i, mp, ip);
                D = derive_mos1(mp, ip, o);
            }
            const size_t off = (size_t)i * n_m1 + d;
            double core[M1_CORE_F] = {D.core.p,      D.core.vt0_t,  D.core.beta,   D.core.phi_t,
                                      D.core.gamma,  D.core.lambda, D.core.vtherm, D.core.isat_d,
                                      D.core.isat_s, D.core.grd,    D.core.grs};
            if (ac) {  // AC fields: the operating point of load_ac, then the terminal conductances
                const Mos1AcOp op = acop_of(k_comp, i);
                const double a[M1_CORE_F] = {op.gm,  op.gds, op.gmbs,     op.gbs,     op.gbd,    op.cgs,
                                             op.cgd, op.cgb, op.reversed, D.core.grd, D.core.grs};
                for (int f = 0; f < M1_CORE_F; f++) core[f] = a[f];
            }
            for (int f = 0; f < M1_CORE_F; f++) hp.m1_core[f * s + off] = core[f];
            const double cp[M1_CAPS_F] = {D.caps.cox, D.caps.cgs_ov, D.caps.cgd_ov, D.caps.cgb_ov};
            for (int f = 0; f < M1_CAPS_F; f++) hp.m1_caps[f * s + off] = cp[f];
        }
    }
    hp.d_par.assign((size_t)D_PAR_F * NI * n_d, 0.0);
    for (int d = 0; d < n_d; d++) {
        const int ci = t.dev_comp[t.cls_start[CLS_DIODE] + d];
        const Comp& c = t.comps[ci];
        const auto* L = ov_of(ci);
        const size_t s = (size_t)NI * n_d;
        DiodeP P{};
        for (int i = 0; i < NI; i++) {
            if (i == 0 || L) {
                s21_params mp, ip;
                inst_params(g, L, c, i, mp, ip);
                P = derive_diode(mp, ip, o);
            }
            const double f[D_PAR_F] = {P.isat, P.vte, P.vcrit, P.gspr, P.bv, P.has_bv, P.cz,
                                       P.cz2,  P.dep, P.f1,    P.f3,   P.tt, P.vj,     P.m};
            const size_t off = (size_t)i * n_d + d;
            for (int k = 0; k < D_PAR_F; k++) hp.d_par[k * s + off] = f[k];
        }
    }
    // BSIM4: whole parameter structures (array of structures), per instance only when overridden
    const int n_b4 = t.cls_start[CLS_BSIM4 + 1] - t.cls_start[CLS_BSIM4];
    hp.n_b4 = n_b4;
    hp.b4_per_inst = false;
    for (int d = 0; d < n_b4; d++) hp.b4_per_inst = hp.b4_per_inst || ov_of(t.dev_comp[t.cls_start[CLS_BSIM4] + d]);
    hp.b4.resize((size_t)(hp.b4_per_inst ? NI : 1) * n_b4);
    for (int d = 0; d < n_b4; d++) {
        const int ci = t.dev_comp[t.cls_start[CLS_BSIM4] + d];
        const Comp& c = t.comps[ci];
        const auto* L = ov_of(ci);
        for (int i = 0; i < (hp.b4_per_inst ? NI : 1); i++) {
            if (i == 0 || L) {
                s21_params mp, ip;
                inst_params(g, L, c, i, mp, ip);
                hp.b4[(size_t)i * n_b4 + d] = derive_b4(mp, ip);
            } else {
                hp.b4[(size_t)i * n_b4 + d] = hp.b4[d];
            }
        }
    }
    hp.lin_val.assign((size_t)NI * n_lin, 0.0);
    hp.lin_tran.assign((size_t)NI * n_lin, 0.0);
    hp.lin_ac.assign((size_t)NI * n_lin, 0.0);
    for (int d = 0; d < n_lin; d++) {
        const int ci = t.dev_comp[t.cls_start[CLS_LIN] + d];
        const Comp& c = t.comps[ci];
        const auto* L = ov_of(ci);
        for (int i = 0; i < NI; i++) {
            double v = c.value, vt = c.value_tran;
            if (L)
                for (const Override* x : *L) v = vt = x->vals[i];
            hp.lin_val[(size_t)i * n_lin + d] = v;
            hp.lin_tran[(size_t)i * n_lin + d] = vt;
            hp.lin_ac[(size_t)i * n_lin + d] = c.type == S21_COMP_V ? c.acm : v;  // AC value: g, c or acm
        }
    }
}

// SIMDs of the current device.
static int device_simds() {
    int dev = 0, cus = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    return 4 * std::max(cus, 1);
}
// Instances per wavefront chosen for a batch: the largest power of two up to 16 that still gives every
// SIMD of the chip two wavefronts (one to run while the other waits on memory).
static int auto_ipw(int total_inst, int simds) {
    int ipw = 1;
    while (ipw < 16 && total_inst / (ipw * 2) >= 2 * simds) ipw *= 2;
    return ipw;
}

// Derive per-instance parameters, build the topology and upload everything for one analysis.
// A latency-bound batch makes every shared slot a hub, so that each device class is evaluated in one
// round instead of one per colour (more LDS cells and a longer gather, fewer dependent rounds): a batch
// of at most 64 instances (one wavefront each at most; the single 5-stage ring: 14.2 -> 9.3 us per
// iteration), and in mixed precision one of at most two instances per SIMD.  Measured: the 1024-ring
// transient batch ran 137 ms per ns with hubs against 151 without in mixed precision, but 304 against
// 278 in FP64 (whose exact-division factorisation is most of an iteration).  Not for AC, whose complex
// gather costs more than the rounds it saves (the 256-instance amplifier sweep: 5.5 ms without, 6.8 ms
// with, mixed).
constexpr int LATENCY_INST = 64;
constexpr int LATENCY_INST_PER_SIMD = 2;

static void prepare_group(s21_batch* b, GroupHost& g, const s21_tran_options* tr, const AcInput* ac = nullptr) {
    HostPrep hp;
    int64_t total = 0;
    for (const GroupHost& x : b->groups) total += x.n_inst;
    int dev = 0, cus = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const bool latency = total <= LATENCY_INST || (b->precision == S21_MIXED &&
                                                   total <= (int64_t)LATENCY_INST_PER_SIMD * 4 * std::max(cus, 1));
    // A group with BSIM4 devices, whose evaluation dominates an iteration, evaluates each class in one
    // round: with colour rounds the devices of one node (drain and gate) would take several (four in an
    // inverter chain), most lanes idle in each.
    bool b4 = false;
    for (const s21_comp& c : g.comps) b4 = b4 || c.type == S21_COMP_BSIM4;
    const bool every_shared = b->hub == 1 || (b->hub == 0 && (latency || b4));
    host_prepare(b->opts, g, tr, hp, ac, every_shared && !ac ? 1 : HUB_CONTRIB);
    const Topology& t = g.topo;
    const int NI = g.n_inst;

    // ---- record tables: one uint16 blob, copied into LDS by every workgroup ----
    topo_records(g.topo);
    std::vector<uint16_t> blob;
    auto sect = [&](const std::vector<uint16_t>& v) {
        const int off = (int)blob.size();
        blob.insert(blob.end(), v.begin(), v.end());
        while (blob.size() % 4) blob.push_back(0);  // 8-byte alignment for packed records
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
    // colours: starts, then {class, class-local first device, first code} per colour
    std::vector<int> col(t.col_ptr);
    for (size_t cc = 0; cc < t.col_cls.size(); cc++) {
        const int k = t.col_cls[cc], p0 = t.col_ptr[cc];
        col.push_back(k);
        col.push_back(p0 - t.cls_start[k == CLS_MOS1T ? CLS_MOS1 : k]);  // Mos1 parameters span both classes
        col.push_back(t.code_base[k] + (p0 - t.cls_start[k]) * code_stride(k));
    }
    TopoGPU& T = g.gg.t;
    T = TopoGPU{};
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
    T.o_col = sect(to16(col));
    T.o_dev = sect(t.dev_rec);
    T.o_code = sect(t.codes);
    while (blob.size() % 8) blob.push_back(0);  // whole 16-byte vectors
    const size_t blob_bytes = blob.size() * sizeof(uint16_t);
    T.n_gt = (int)t.gt.size() / 4;
    T.n_gc = (int)t.gc.size() / 4;
    T.n_rt = (int)t.rt.size() / 4;
    T.n_rc = (int)t.rc.size() / 4;
    T.n_part = t.n_part;
    T.n = t.n;
    T.nnz = t.nnz;
    T.n_hub = t.n_hub;
    for (int k = 0; k < N_CLS; k++) T.n_cls[k] = t.cls_start[k + 1] - t.cls_start[k];
    for (int k = 0; k <= N_CLS; k++) T.cls_start[k] = t.cls_start[k];
    T.n_col = (int)t.col_cls.size();
    T.n_clvl = std::max((int)t.clvl_ptr.size() - 1, 0);
    T.n_blvl = (int)t.blvl_ptr.size() - 1;
    T.sym_status = t.sym_status;
    T.blob_vec = (int)(blob_bytes / 16);
    g.d_topo.reset(new DevMem());
    g.d_topo->alloc(blob_bytes);
    HIPCHECK(hipMemcpy(g.d_topo->p, blob.data(), blob_bytes, hipMemcpyHostToDevice));
    T.blob = (const uint4*)g.d_topo->p;

    // ---- parameters (read-only), packed by field: a field no instance overrides is stored once ----
    auto pad = [](size_t n) { return ((n * sizeof(double) + 255) / 256) * 256; };
    std::vector<double> par;
    auto pack = [&](const std::vector<double>& full, int nf, int ndev, long long* off, int* istr) {
        const size_t s = (size_t)NI * ndev;
        for (int f = 0; f < nf; f++) {
            const double* v = full.data() + f * s;
            bool shared = true;
            for (size_t i = ndev; i < s && shared; i++) shared = v[i] == v[i % ndev];
            while (par.size() % 32) par.push_back(0.0);  // 256-byte aligned fields
            off[f] = (long long)par.size();
            istr[f] = shared ? 0 : ndev;
            par.insert(par.end(), v, v + (shared ? (size_t)ndev : s));
        }
    };
    {
        std::vector<double> m1(hp.m1_core);
        m1.insert(m1.end(), hp.m1_caps.begin(), hp.m1_caps.end());
        pack(m1, M1_CORE_F + M1_CAPS_F, hp.n_m1, g.gg.m1_off, g.gg.m1_istr);
        pack(hp.d_par, D_PAR_F, hp.n_d, g.gg.d_off, g.gg.d_istr);
        std::vector<double> lin(hp.lin_val);
        lin.insert(lin.end(), hp.lin_tran.begin(), hp.lin_tran.end());
        lin.insert(lin.end(), hp.lin_ac.begin(), hp.lin_ac.end());
        pack(lin, 3, hp.n_lin, g.gg.lin_off, g.gg.lin_istr);
    }
    g.d_par.reset(new DevMem());
    g.d_par->alloc(std::max<size_t>(par.size(), 1) * sizeof(double));
    if (!par.empty()) HIPCHECK(hipMemcpy(g.d_par->p, par.data(), par.size() * sizeof(double), hipMemcpyHostToDevice));
    g.gg.par = (const double*)g.d_par->p;

    // ---- BSIM4: the field map, shared and per-instance parameters, then the state (last) ----
    {
        const size_t nb = (size_t)hp.n_b4;
        std::vector<int> map(B4_PAR_D, -1);
        int n_var = 0;
        const size_t ni_b4 = hp.b4_per_inst ? (size_t)NI : 1;
        auto field = [&](size_t i, size_t j, int k) { return ((const double*)&hp.b4[i * nb + j])[k]; };
        for (int k = 0; k < B4_PAR_D; k++) {
            bool same = true;
            for (size_t i = 1; i < ni_b4 && same; i++)
                for (size_t j = 0; j < nb && same; j++) {
                    const double a = field(i, j, k), c = field(0, j, k);
                    same = memcmp(&a, &c, sizeof(double)) == 0;
                }
            if (!same) map[k] = n_var++;
        }
        std::vector<double> sh((size_t)B4_PAR_D * nb), var((size_t)n_var * NI * nb);
        for (int k = 0; k < B4_PAR_D; k++)
            for (size_t j = 0; j < nb; j++) {
                sh[(size_t)k * nb + j] = field(0, j, k);
                if (map[k] >= 0)
                    for (size_t i = 0; i < (size_t)NI; i++) var[((size_t)map[k] * NI + i) * nb + j] = field(i, j, k);
            }
        const size_t s_map = ((B4_PAR_D * sizeof(int) + 255) / 256) * 256, s_sh = pad(sh.size()),
                     s_var = pad(var.size()), s_st = pad((size_t)B4_ST_F * NI * nb);
        g.d_b4.reset(new DevMem());
        g.d_b4->alloc(nb ? s_map + s_sh + s_var + s_st : 0);
        char* p = (char*)g.d_b4->p;
        if (nb) {
            HIPCHECK(hipMemcpy(p, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice));
            HIPCHECK(hipMemcpy(p + s_map, sh.data(), sh.size() * sizeof(double), hipMemcpyHostToDevice));
            if (!var.empty())
                HIPCHECK(hipMemcpy(p + s_map + s_sh, var.data(), var.size() * sizeof(double), hipMemcpyHostToDevice));
            HIPCHECK(hipMemsetAsync(p + s_map + s_sh + s_var, 0, s_st, b->stream));
        }
        g.gg.b4_map = nb ? (const int*)p : nullptr;
        g.gg.b4_sh = nb ? (const double*)(p + s_map) : nullptr;
        g.gg.b4_var = nb ? (const double*)(p + s_map + s_sh) : nullptr;
        g.gg.b4_st = nb ? (double*)(p + s_map + s_sh + s_var) : nullptr;
        g.b4_var_fields = n_var;
    }

    // ---- state (zero-initialised: Default op points, comps/mos.rs:524-546) ----
    const size_t s_m1s = pad((size_t)M1_STATE_F * 2 * NI * hp.n_m1), s_ds = pad((size_t)D_STATE_F * NI * hp.n_d),
                 s_cq = pad((size_t)NI * hp.n_lin);
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
                 s_prof = b->profile ? ((NI * NPROF * sizeof(long long) + 255) / 256) * 256 : 0,
                 s_ra = ((NI * 2 * sizeof(int) + 255) / 256) * 256;
    g.d_out.reset(new DevMem());
    g.d_out->alloc(s_x + s_st + s_it + s_dbg + s_prof + s_ra + 256);
    HIPCHECK(hipMemsetAsync(g.d_out->p, 0, g.d_out->bytes, b->stream));
    char* op = (char*)g.d_out->p;
    g.gg.x = (double*)op;
    g.gg.status = (int*)(op + s_x);
    g.gg.iters = (long long*)(op + s_x + s_st);
    g.gg.dbg = (double*)(op + s_x + s_st + s_it);
    g.gg.prof = b->profile ? (long long*)(op + s_x + s_st + s_it + s_dbg) : nullptr;
    g.gg.resume_at = (int*)(op + s_x + s_st + s_it + s_dbg + s_prof);
    g.gg.n_inst = NI;
    g.gg.wave = nullptr;

    HIPCHECK(hipStreamSynchronize(b->stream));
    g.blob_bytes = blob_bytes;
    const size_t vw = ac ? 2 : 1;  // values per entry: complex for AC
    g.inst_bytes = (size_t)(3 * t.n + t.nnz + t.n_hub + t.n_part) * sizeof(double) * vw +
                   (b->precision == S21_MIXED ? (size_t)(t.nnz + t.n) * sizeof(float) * vw : 0);
    if (ac) {  // the instance's AC parameters, cached in LDS by the kernel; in FP64 the copy of A it factors
        g.inst_bytes += (size_t)(M1_CORE_F * (t.cls_start[CLS_MOS1T + 1] - t.cls_start[CLS_MOS1]) +
                                 (t.cls_start[CLS_LIN + 1] - t.cls_start[CLS_LIN])) * sizeof(double);
        if (b->precision != S21_MIXED) g.inst_bytes += (size_t)t.nnz * 2 * sizeof(double);
    }
    g.gg.acop = nullptr;
}

// Key of the structure an analysis needs: the operating point, or a transient with its IC set.
static std::string prep_key(const s21_tran_options* tr) {
    if (!tr) return "op";
    std::string k = "tran";
    for (int i = 0; i < tr->n_ic; i++) k += ":" + std::to_string(tr->ic_var[i]);
    return k;
}

// Prepare every group for an analysis, or only reset the Newton state when the structure is current.
// Generated factorisation: for a batch with enough instances to repay the compilation (or when asked),
// one kernel with every group's pattern written out -- groups of one pattern share its functions.
static void prepare_jit(s21_batch* b, bool ac = false, int64_t n_freq = 0) {
    b->has_jit = false;
    int64_t total = 0, ops = 0;
    for (const GroupHost& g : b->groups) total += g.n_inst;
    if (b->jit < 0 || (b->jit == 0 && (ac ? total * n_freq < JIT_MIN_AC_POINTS : total < JIT_MIN_INST))) return;
    if (b->has_b4) return;  // the generated kernels carry no BSIM4 evaluation: record-driven factorisation
    std::vector<const Topology*> pats;
    std::vector<int> group_pattern;
    std::unordered_map<std::string, int> seen;  // pattern -> index, keyed by what its code is written from
    for (const GroupHost& g : b->groups) {
        const Topology& t = g.topo;
        if (t.sym_status != S21_OK) return;
        std::string key = std::to_string(t.n) + ":" + std::to_string(t.nnz);
        for (const std::vector<int>* v : {&t.rp, &t.ci, &t.diag, &t.upd_ptr, &t.upd_src, &t.upd_dst})
            key.append((const char*)v->data(), v->size() * sizeof(int)).push_back('|');
        auto it = seen.find(key);
        if (it == seen.end()) {
            const int64_t k = jit_ops(t);
            if (k > JIT_MAX_OPS) return;
            ops += k;
            it = seen.emplace(key, (int)pats.size()).first;
            pats.push_back(&t);
        }
        group_pattern.push_back(it->second);
    }
    if (ops > JIT_MAX_TOTAL_OPS) return;
    b->jitk = jit_compile(jit_source(pats, group_pattern, b->precision == S21_MIXED, 2, ac, b->profile != 0));
    b->has_jit = true;
}

static void prepare_all(s21_batch* b, const s21_tran_options* tr) {
    const std::string key = prep_key(tr);
    bool changed = false;
    for (GroupHost& g : b->groups) {
        if (g.prepared == key) {
            HIPCHECK(hipMemsetAsync(g.d_state->p, 0, g.d_state->bytes, b->stream));
            HIPCHECK(hipMemsetAsync(g.d_out->p, 0, g.d_out->bytes, b->stream));
            if (g.gg.b4_st) {  // the parameters stay; the state after them restarts from zero
                char* p0 = (char*)g.gg.b4_st;
                HIPCHECK(hipMemsetAsync(p0, 0, g.d_b4->bytes - (size_t)(p0 - (char*)g.d_b4->p), b->stream));
            }
        } else {
            g.prepared.clear();
            prepare_group(b, g, tr);
            g.prepared = key;
            changed = true;
        }
    }
    if (changed) prepare_jit(b);
    // No wait here: the resets are ordered on the batch's stream, so every later write into the state or
    // output buffers (the launch, s21_batch_assemble's upload of x) must be issued on that stream too.
}

template <bool MIXED>
static void launch(s21_batch* b, const GroupGPU* dgroups, int n_groups, int n_blocks, size_t smem, const Args& A) {
    if (b->has_b4 && A.prof) fail(S21_INVALID, "the profiling kernels do not evaluate BSIM4 devices");
    auto kern = b->has_b4 ? newton_kernel_b4<MIXED> : A.prof ? newton_kernel_prof<MIXED> : newton_kernel<MIXED>;
    if (smem > 64 * 1024)
        HIPCHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    hipLaunchKernelGGL(kern, dim3(n_blocks), dim3(WAVE * A.waves), smem, b->stream, dgroups, n_groups, A);
    HIPCHECK(hipGetLastError());
}
static void launch_any(s21_batch* b, const GroupGPU* dg, int ng, int nb, size_t smem, const Args& A) {
    if (b->precision == S21_MIXED)
        launch<true>(b, dg, ng, nb, smem, A);
    else
        launch<false>(b, dg, ng, nb, smem, A);
}

// LDS bytes between the regions of consecutive instances: whole 256-byte rows plus an offset that
// starts the segments of one half-wavefront on different LDS banks.
static size_t inst_stride(size_t raw, int ipw, bool lane_per_instance = false) {
    const size_t L = WAVE / ipw;
    // The generated factorisation reads the same offset of every instance's region from one lane each:
    // regions 8 bytes apart modulo 256 put those lanes on different banks.
    return ((raw + 255) / 256) * 256 + (lane_per_instance ? 8 : (L * 8) % 256);
}
constexpr size_t JIT_STATIC_LDS = sizeof(int) * WAVE * MAX_WAVES;  // newton_wg's per-slot flags
constexpr size_t AC_STATIC_LDS = 128;  // the AC third pass's flags and outcome (64 bytes), with alignment

// Launch shape: instances per wavefront (the batch's setting, or auto_ipw) and wavefronts per workgroup:
// the most that still leave a workgroup for every CU (or as many workgroups as one wavefront per
// workgroup gives, for a smaller batch) and fit in the LDS -- a larger workgroup shares the generated
// factorisation's lanes among more instances, a smaller one spreads the batch over more CUs.
static void pick_config(const s21_batch* b, size_t blob, size_t raw, int total_inst, size_t max_block, int simds,
                        Args& A, size_t& stride) {
    int ipw = b->ipw;
    const bool jit = b->has_jit;
    blob += AC_STATIC_LDS;
    if (jit) blob += JIT_STATIC_LDS;
    if (A.mode == MODE_AC && A.ac_pass == 2) {  // a workgroup per instance, its wavefronts looking ahead
        stride = inst_stride(raw, 1, jit);
        if (blob + stride > max_block)
            fail(S21_INVALID, "circuit too large for the batched kernel: " + std::to_string(blob + stride) +
                                  " bytes of LDS per instance");
        int k = MAX_WAVES;
        while (k > 1 && blob + (size_t)k * stride > max_block) k--;
        A.ipw = 1;
        A.waves = k;
        A.ac_look = k;
        return;
    }
    if (ipw <= 0) ipw = auto_ipw(total_inst, simds);
    const long cus = std::max(simds / 4, 1);
    for (;; ipw /= 2) {
        stride = inst_stride(raw, ipw, jit);
        auto workgroups = [&](int k) {
            long n = 0;
            for (const GroupHost& g : b->groups) n += ((long)g.n_inst * ac_units(A) + k * ipw - 1) / (k * ipw);
            return n;
        };
        const long target = std::min(cus, workgroups(1));
        long largest = 0;  // launch units of the largest group: no wavefront beyond them has work
        for (const GroupHost& g : b->groups) largest = std::max(largest, (long)g.n_inst * ac_units(A));
        int best = 0;
        const int kmax = b->waves > 0 ? std::min(b->waves, MAX_WAVES) : MAX_WAVES;
        for (int k = 1; k <= kmax; k++) {
            if (blob + (size_t)k * ipw * stride > max_block) break;
            if (k > 1 && b->waves <= 0 && (workgroups(k) < target || (long)(k - 1) * ipw >= largest)) break;
            best = k;
        }
        if (best > 0) {
            A.ipw = ipw;
            A.waves = best;
            return;
        }
        if (ipw == 1)
            fail(S21_INVALID, "circuit too large for the batched kernel: " + std::to_string(blob + stride) +
                                  " bytes of LDS per instance");
    }
}

// Launch of the kernel over every group: its arguments, LDS bytes, workgroups and group descriptors.
struct LaunchPlan {
    Args A;
    size_t smem = 0;
    int blk = 0;
    std::vector<GroupGPU> hg;
};

static LaunchPlan plan_launch(const s21_batch* b, const Args& A0, double* d_wave, int64_t npts) {
    LaunchPlan P;
    size_t wave_off = 0, blob = 0, raw = 0;
    int total = 0;
    for (const GroupHost& g : b->groups) {
        GroupGPU gg = g.gg;
        gg.wave = d_wave ? d_wave + wave_off : nullptr;
        wave_off += (size_t)g.n_inst * (size_t)npts * g.topo.n;
        P.hg.push_back(gg);
        blob = std::max(blob, g.blob_bytes);
        raw = std::max(raw, g.inst_bytes);
        total += g.n_inst * ac_units(A0);
    }
    int dev = 0, max_block = 0, cus = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&max_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    P.A = A0;
    size_t stride = 0;
    pick_config(b, blob, raw, total, (size_t)std::max(max_block, 65536), 4 * std::max(cus, 1), P.A, stride);
    P.A.inst_bytes = (int)stride;
    P.smem = blob + (size_t)P.A.waves * P.A.ipw * stride;
    const int per_block = P.A.waves * P.A.ipw;
    for (GroupGPU& gg : P.hg) {
        gg.blk0 = P.blk;
        P.blk += (gg.n_inst * ac_units(P.A) + per_block - 1) / per_block;
    }
    return P;
}

// Launches a plan whose descriptors are already at `dgroups` (device memory), timed by the batch's events.
static void launch_plan(s21_batch* b, const LaunchPlan& P, const GroupGPU* dgroups) {
    const Args& A = P.A;
    const bool jit = b->has_jit && A.mode != MODE_ASSEMBLE;
    HIPCHECK(hipEventRecord(b->ev0, b->stream));
    if (jit) {
        const GroupGPU* dg = dgroups;
        int ng = (int)P.hg.size();
        Args a = A;
        void* args[] = {(void*)&dg, (void*)&ng, (void*)&a};
        HIPCHECK(hipModuleLaunchKernel(b->jitk.fn, P.blk, 1, 1, WAVE * A.waves, 1, 1, (unsigned)P.smem, b->stream,
                                       args, nullptr));
    } else {
        launch_any(b, dgroups, (int)P.hg.size(), P.blk, P.smem, A);
    }
    HIPCHECK(hipEventRecord(b->ev1, b->stream));
    HIPCHECK(hipEventSynchronize(b->ev1));
    float ms = 0.0f;
    HIPCHECK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_waves = A.waves;
    b->last_ipw = A.ipw;
    b->last_jit = jit;
    b->launches = 1;
    b->last_ms = ms;
    b->kernel_avg_ms = ms;
}

// Runs the kernel over every group in one launch.
static void run(s21_batch* b, const Args& A0, double* d_wave, int64_t npts) {
    const LaunchPlan P = plan_launch(b, A0, d_wave, npts);
    if (!b->d_groups) b->d_groups.reset(new DevMem());
    const size_t gbytes = P.hg.size() * sizeof(GroupGPU);
    if (b->d_groups->bytes < gbytes) {
        b->d_groups->alloc(gbytes);
        b->groups_uploaded.clear();
    }
    // The descriptors change only with a preparation or an analysis's buffers: a repeated call skips them.
    if (b->groups_uploaded.size() != gbytes || std::memcmp(b->groups_uploaded.data(), P.hg.data(), gbytes) != 0) {
        HIPCHECK(hipMemcpyAsync(b->d_groups->p, P.hg.data(), gbytes, hipMemcpyHostToDevice, b->stream));
        b->groups_uploaded.assign((const char*)P.hg.data(), (const char*)P.hg.data() + gbytes);
    }
    launch_plan(b, P, (const GroupGPU*)b->d_groups->p);
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
    A.ipw = 1;
    A.inst_bytes = 0;
    A.capture = 0;
    A.resume = 0;
    A.asm_tran = 0;
    A.omega = nullptr;
    A.ac_runs = 1;
    A.ac_pass = 0;
    A.ac_look = 1;
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
    if (nl.n_bsim4_models) g.bsim4_models.assign(nl.bsim4_models, nl.bsim4_models + nl.n_bsim4_models);
    if (nl.n_bsim4_insts) g.bsim4_insts.assign(nl.bsim4_insts, nl.bsim4_insts + nl.n_bsim4_insts);
    if (nl.n_mos1_models) g.mos1_models.assign(nl.mos1_models, nl.mos1_models + nl.n_mos1_models);
    if (nl.n_mos1_insts) g.mos1_insts.assign(nl.mos1_insts, nl.mos1_insts + nl.n_mos1_insts);
    if (nl.n_diode_models) g.diode_models.assign(nl.diode_models, nl.diode_models + nl.n_diode_models);
    if (nl.n_diode_insts) g.diode_insts.assign(nl.diode_insts, nl.diode_insts + nl.n_diode_insts);
    g.n_inst = sg.n_inst;
    for (int k = 0; k < sg.n_over; k++) {
        const s21_override& o = sg.over[k];
        if (o.comp < 0 || o.comp >= nl.n_comps || !o.vals) fail(S21_INVALID, "override: bad component");
        // BSIM4 parameter sets start at field 1 (Bsim4Model.mos_type, Bsim4InstParams.l)
        const int f0 = nl.comps[o.comp].type == S21_COMP_BSIM4 ? 1 : 2;
        if (o.target != S21_OVR_VALUE && (o.field < f0 || o.field >= S21_PARAM_SLOTS))
            fail(S21_INVALID, "override: bad field");
        Override x;
        x.comp = o.comp;
        x.target = o.target;
        x.field = o.field;
        x.vals.assign(o.vals, o.vals + sg.n_inst);
        g.overs.push_back(std::move(x));
    }
}

// ---- pivot recovery ----
// The reference searches its pivots again at every factorisation (Matrix::solve after each reset,
// sparse21/mod.rs:647-674, 929-932); the engine keeps the first search's order for as long as every pivot
// of it is a normal number, neither zero nor subnormal (factor_forward).  An instance whose kept order meets
// a pivot that fails this (status S21_PIVOT) resumes the interrupted iteration -- of the operating point or of a transient
// time point -- in a launch of its own: a mixed-precision instance first in FP64 on the same order, then
// on the order the reference's search picks on that iteration's matrix (assembled on the GPU from the
// instance's state), searched from the arrangement and fill its earlier factorisations left.  The
// resumed solve keeps its iteration count and the limit of 100 (analysis.rs:173); each resumption uses
// at least one iteration.
constexpr int MAX_RESUME = 8;

// An instance's device state, by component (the device order of a structure depends on its colouring).
struct InstState {
    std::vector<std::vector<double>> m1, d;  // per Mos1 / diode component: its state fields
    std::vector<double> q;                    // per linear component: committed capacitor charge
    // per BSIM4 component: the limiting fields its last load started from (vds, vgs, vbs, vbd, von), then
    // its committed point (the B4Op's doubles)
    std::vector<std::vector<double>> b4;
};
constexpr int B4_PREV_F = 5, B4_QC_F = 3;
static void dev_maps(const Topology& t, std::vector<int>& m1, std::vector<int>& d, std::vector<int>& lin) {
    m1.assign(t.dev_comp.begin() + t.cls_start[CLS_MOS1], t.dev_comp.begin() + t.cls_start[CLS_MOS1T + 1]);
    d.assign(t.dev_comp.begin() + t.cls_start[CLS_DIODE], t.dev_comp.begin() + t.cls_start[CLS_DIODE + 1]);
    lin.assign(t.dev_comp.begin() + t.cls_start[CLS_LIN], t.dev_comp.begin() + t.cls_start[CLS_LIN + 1]);
}
// State of instance i; the diode guess is the one the last evaluation started from (field 3).
static InstState read_state(const GroupHost& g, int i) {
    std::vector<int> m1, d, lin;
    dev_maps(g.topo, m1, d, lin);
    const size_t NI = g.n_inst, nm = m1.size(), nd = d.size(), nl = lin.size();
    InstState S;
    const size_t ncomp = g.topo.comps.size();
    S.m1.assign(ncomp, {});
    S.d.assign(ncomp, {});
    S.q.assign(ncomp, 0.0);
    S.b4.assign(ncomp, {});
    const Topology& t = g.topo;
    const size_t nb = (size_t)(t.cls_start[CLS_BSIM4 + 1] - t.cls_start[CLS_BSIM4]);
    if (nb) {
        std::vector<double> st((size_t)B4_ST_F * NI * nb);
        HIPCHECK(hipMemcpy(st.data(), g.gg.b4_st, st.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < nb; k++) {
            std::vector<double>& v = S.b4[t.dev_comp[t.cls_start[CLS_BSIM4] + k]];
            for (int f = 0; f < B4_PREV_F; f++) v.push_back(st[(B4_ST_PREV + f) * NI * nb + i * nb + k]);
            for (int f = 0; f < B4_QC_F; f++) v.push_back(st[(B4_ST_QC + f) * NI * nb + i * nb + k]);
        }
    }
    std::vector<double> tmp;
    auto rd = [&](const double* base, size_t nf, size_t ndev, std::vector<std::vector<double>>& out,
                  const std::vector<int>& comp) {
        tmp.resize(nf * NI * ndev);
        if (!tmp.empty())
            HIPCHECK(hipMemcpy(tmp.data(), base, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < ndev; k++) {
            std::vector<double>& v = out[comp[k]];
            v.resize(nf);
            for (size_t f = 0; f < nf; f++) v[f] = tmp[f * NI * ndev + i * ndev + k];
        }
    };
    rd(g.gg.m1_state, 2 * M1_STATE_F, nm, S.m1, m1);
    rd(g.gg.d_state, D_STATE_F, nd, S.d, d);
    for (size_t k = 0; k < nd; k++) S.d[d[k]][0] = S.d[d[k]][3];
    if (nl) {
        tmp.resize(NI * nl);
        HIPCHECK(hipMemcpy(tmp.data(), g.gg.c_q, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < nl; k++) S.q[lin[k]] = tmp[i * nl + k];
    }
    return S;
}
// State of a one-instance group.
static void write_state(GroupHost& h, const InstState& S) {
    std::vector<int> m1, d, lin;
    dev_maps(h.topo, m1, d, lin);
    const size_t nm = m1.size(), nd = d.size(), nl = lin.size();
    std::vector<double> tmp;
    auto wr = [&](double* base, size_t nf, size_t ndev, const std::vector<std::vector<double>>& in,
                  const std::vector<int>& comp) {
        tmp.assign(nf * ndev, 0.0);
        for (size_t k = 0; k < ndev; k++)
            for (size_t f = 0; f < nf && f < in[comp[k]].size(); f++) tmp[f * ndev + k] = in[comp[k]][f];
        if (!tmp.empty())
            HIPCHECK(hipMemcpy(base, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice));
    };
    wr(h.gg.m1_state, 2 * M1_STATE_F, nm, S.m1, m1);
    wr(h.gg.d_state, D_STATE_F, nd, S.d, d);
    const Topology& t = h.topo;
    const size_t nb = (size_t)(t.cls_start[CLS_BSIM4 + 1] - t.cls_start[CLS_BSIM4]);
    if (nb) {  // the load resumes from the limiting fields it started from, against the committed charges
        std::vector<double> st((size_t)B4_ST_F * nb, 0.0);
        for (size_t k = 0; k < nb; k++) {
            const std::vector<double>& v = S.b4[t.dev_comp[t.cls_start[CLS_BSIM4] + k]];
            if (v.size() != (size_t)(B4_PREV_F + B4_QC_F)) fail(S21_ERR, "internal: BSIM4 state of the wrong size");
            for (int f = 0; f < B4_PREV_F; f++) {
                st[(B4_ST_LIM + f) * nb + k] = v[f];
                st[(B4_ST_PREV + f) * nb + k] = v[f];
            }
            for (int f = 0; f < B4_QC_F; f++) st[(B4_ST_QC + f) * nb + k] = v[B4_PREV_F + f];
        }
        HIPCHECK(hipMemcpy(h.gg.b4_st, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    if (nl) {
        tmp.assign(nl, 0.0);
        for (size_t k = 0; k < nl; k++) tmp[k] = S.q[lin[k]];
        HIPCHECK(hipMemcpy(h.gg.c_q, tmp.data(), nl * sizeof(double), hipMemcpyHostToDevice));
    }
}

struct Stalled {
    int g = 0, i = 0;           // group, instance
    std::vector<double> x;      // iterate of the interrupted iteration
    InstState state;            // state before that iteration
    long long iters = 0;        // loads before the interrupted iteration
    int pt = 0, used = 0;       // its time point, and the iterations its solve had used before it
    bool fp64 = false;          // the interrupted solve ran in FP64
    std::vector<int> row, col;  // the order that met the zero pivot
    std::vector<std::vector<double>> later;  // matrices of the searches after the first
};

static Stalled stalled_from(const GroupHost& g, int gi, int i, bool fp64) {
    Stalled s;
    s.g = gi;
    s.i = i;
    s.x.resize(g.topo.n);
    HIPCHECK(hipMemcpy(s.x.data(), g.gg.x + (size_t)i * g.topo.n, g.topo.n * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(&s.iters, g.gg.iters + i, sizeof(long long), hipMemcpyDeviceToHost));
    int ra[2];
    HIPCHECK(hipMemcpy(ra, g.gg.resume_at + 2 * i, sizeof(ra), hipMemcpyDeviceToHost));
    s.iters -= 1;
    s.pt = ra[0];
    s.used = ra[1] - 1;
    s.state = read_state(g, i);
    s.fp64 = fp64;
    s.row = g.topo.row_i2e;
    s.col = g.topo.col_i2e;
    return s;
}

// Device offset of group gi's waveform in the batch's wave buffer (the layout run() gives it).
static size_t wave_offset(const s21_batch* b, int gi, int64_t npts) {
    size_t off = 0;
    for (int k = 0; k < gi; k++) off += (size_t)b->groups[k].n_inst * (size_t)npts * b->groups[k].topo.n;
    return off;
}

static void recover_pivots(s21_batch* b, const s21_tran_options* tr, int64_t npts) {
    b->resumed = 0;
    std::vector<Stalled> todo;
    for (size_t gi = 0; gi < b->groups.size(); gi++) {
        GroupHost& g = b->groups[gi];
        std::vector<int> st(g.n_inst);
        HIPCHECK(hipMemcpy(st.data(), g.gg.status, st.size() * sizeof(int), hipMemcpyDeviceToHost));
        if (g.topo.sym_status != S21_OK) continue;  // the search itself failed: nothing to resume on
        for (int i = 0; i < g.n_inst; i++)
            if (st[i] == S21_PIVOT) todo.push_back(stalled_from(g, (int)gi, i, b->precision != S21_MIXED));
    }
    if (todo.empty()) return;
    fetch_results(b);  // host copies of every result, patched below
    const int mode = tr ? MODE_TRAN : MODE_OP;
    for (int round = 0; round < MAX_RESUME && !todo.empty(); round++) {
        s21_batch sb;
        sb.opts = b->opts;
        sb.precision = S21_FP64;
        sb.has_b4 = b->has_b4;
        sb.ipw = 1;
        sb.jit = -1;
        HIPCHECK(hipStreamCreateWithFlags(&sb.stream, hipStreamNonBlocking));
        HIPCHECK(hipEventCreate(&sb.ev0));
        HIPCHECK(hipEventCreate(&sb.ev1));
        sb.groups.resize(todo.size());
        for (size_t k = 0; k < todo.size(); k++) {
            const Stalled& s = todo[k];
            const GroupHost& g = b->groups[s.g];
            GroupHost& h = sb.groups[k];
            h.n_vars = g.n_vars;
            h.names = g.names;
            h.kinds = g.kinds;
            h.comps = g.comps;
            h.mos1_models = g.mos1_models;
            h.mos1_insts = g.mos1_insts;
            h.diode_models = g.diode_models;
            h.diode_insts = g.diode_insts;
            h.bsim4_models = g.bsim4_models;
            h.bsim4_insts = g.bsim4_insts;
            h.n_inst = 1;
            for (const Override& o : g.overs) {
                Override x = o;
                x.vals.assign(1, o.vals[s.i]);
                h.overs.push_back(std::move(x));
            }
            h.pivot_row = s.row;  // the order that failed: FP64 retry, or the structure for the assembly below
            h.pivot_col = s.col;
        }
        auto load_inputs = [&](size_t k) {
            const Stalled& s = todo[k];
            GroupHost& h = sb.groups[k];
            HIPCHECK(hipMemcpy(h.gg.x, s.x.data(), s.x.size() * sizeof(double), hipMemcpyHostToDevice));
            HIPCHECK(hipMemcpy(h.gg.iters, &s.iters, sizeof(long long), hipMemcpyHostToDevice));
            const int ra[2] = {s.pt, s.used};
            HIPCHECK(hipMemcpy(h.gg.resume_at, ra, sizeof(ra), hipMemcpyHostToDevice));
            write_state(h, s.state);
        };
        prepare_all(&sb, tr);
        HIPCHECK(hipStreamSynchronize(sb.stream));  // the state resets land before the uploads below
        // Instances that already ran in FP64 need a new order: the interrupted iteration's matrix, as the
        // kernel assembles it from the iterate and the state, searched after the earlier ones.
        bool rebuilt = false;
        for (size_t k = 0; k < todo.size(); k++) {
            Stalled& s = todo[k];
            if (!s.fp64) continue;
            GroupHost& h = sb.groups[k];
            load_inputs(k);
            Args A = base_args(&sb, MODE_ASSEMBLE);
            A.asm_tran = s.pt > 0;
            A.tstep = tr ? tr->tstep : 0.0;
            A.dbg_inst = 0;
            A.waves = 1;
            A.ipw = 1;
            const size_t stride = inst_stride(h.inst_bytes, 1);
            A.inst_bytes = (int)stride;
            GroupGPU gg = h.gg;
            gg.blk0 = 0;
            if (!sb.d_groups) sb.d_groups.reset(new DevMem());
            if (sb.d_groups->bytes < sizeof(GroupGPU)) sb.d_groups->alloc(sizeof(GroupGPU));
            HIPCHECK(hipMemcpy(sb.d_groups->p, &gg, sizeof(GroupGPU), hipMemcpyHostToDevice));
            sb.groups_uploaded.clear();
            launch_any(&sb, (const GroupGPU*)sb.d_groups->p, 1, 1, h.blob_bytes + stride, A);
            HIPCHECK(hipStreamSynchronize(sb.stream));
            const Topology& t = h.topo;
            std::vector<double> dbg(t.nnz + t.n), vals(t.erow.size());
            HIPCHECK(hipMemcpy(dbg.data(), h.gg.dbg, dbg.size() * sizeof(double), hipMemcpyDeviceToHost));
            for (size_t e = 0; e < vals.size(); e++) vals[e] = dbg[t.elem_pos[e]];
            s.later.push_back(std::move(vals));
            h.pivot_row.clear();
            h.pivot_col.clear();
            h.pivot_later = s.later;
            h.prepared.clear();
            rebuilt = true;
        }
        if (rebuilt) prepare_all(&sb, tr);
        HIPCHECK(hipStreamSynchronize(sb.stream));
        for (size_t k = 0; k < todo.size(); k++) load_inputs(k);
        std::unique_ptr<DevMem> wave;
        if (tr) {
            size_t total = 0;
            for (GroupHost& h : sb.groups) total += (size_t)npts * h.topo.n;
            wave.reset(new DevMem());
            wave->alloc(std::max<size_t>(total, 1) * sizeof(double));
        }
        Args A = base_args(&sb, mode);
        A.resume = 1;
        if (tr) {
            A.tstep = tr->tstep;
            A.tstop = tr->tstop;
            A.npts = npts;
        }
        run(&sb, A, tr ? (double*)wave->p : nullptr, tr ? npts : 1);
        b->resumed += (int64_t)todo.size();
        sb.fetched = false;
        fetch_results(&sb);
        std::vector<Stalled> next;
        for (size_t k = 0; k < todo.size(); k++) {
            Stalled& s = todo[k];
            GroupHost& h = sb.groups[k];
            GroupHost& g = b->groups[s.g];
            const int n = g.topo.n;
            auto copy_wave = [&](int64_t p1) {  // points [s.pt, p1) of this round into the batch's waveform
                if (!tr || p1 <= s.pt) return;
                double* dst = (double*)b->d_wave->p + wave_offset(b, s.g, npts) + ((size_t)s.i * npts + s.pt) * n;
                const double* src = (const double*)wave->p + wave_offset(&sb, (int)k, npts) + (size_t)s.pt * n;
                HIPCHECK(hipMemcpy(dst, src, (size_t)(p1 - s.pt) * n * sizeof(double), hipMemcpyDeviceToDevice));
            };
            if (h.status[0] == S21_PIVOT && round + 1 < MAX_RESUME) {
                Stalled t2 = stalled_from(h, s.g, 0, true);
                t2.i = s.i;
                t2.later = s.later;
                copy_wave(t2.pt);
                next.push_back(std::move(t2));
                continue;
            }
            std::copy(h.x.begin(), h.x.begin() + n, g.x.begin() + (size_t)s.i * n);
            g.status[s.i] = h.status[0];
            g.iters[s.i] = h.iters[0];
            HIPCHECK(hipMemcpy(g.gg.x + (size_t)s.i * n, h.x.data(), n * sizeof(double), hipMemcpyHostToDevice));
            HIPCHECK(hipMemcpy(g.gg.status + s.i, &h.status[0], sizeof(int), hipMemcpyHostToDevice));
            HIPCHECK(hipMemcpy(g.gg.iters + s.i, &h.iters[0], sizeof(long long), hipMemcpyHostToDevice));
            copy_wave(npts);
        }
        todo.swap(next);
    }
}

// ---- allocation-free operating point ----
// s21_batch_op_noalloc keeps every instance's Newton state (committed and guess operating points of the
// devices, charges) in this fixed array instead of buffers of the batch; the group descriptors sit at its
// start.  Each device holds its own copy from module load until the process ends.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static char* g_scratch_ptr[NOALLOC_MAX_DEVICES];  // device address of `scratch`, per device

// The array's device address on the current device, resolved once per device.
static char* scratch_on_current_device() {
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) fail(S21_DEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) fail(S21_DEVICE, "scratch array has no device address");
        g_scratch_ptr[dev] = (char*)p;
    }
    return g_scratch_ptr[dev];
}

// Clears n 32-bit words with a kernel (a fill kernel rather than a memset node, so that the call can be
// captured into a HIP graph).
__global__ void fill_zero_u32(uint32_t* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}
static void zero_async(void* p, size_t bytes, hipStream_t s) {
    const size_t n = bytes / 4;
    if (!n) return;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(fill_zero_u32, dim3(blocks), dim3(256), 0, s, (uint32_t*)p, n);
    HIPCHECK(hipGetLastError());
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

int s21_jit_build(const s21_netlist* net, const s21_options* opts, const s21_tran_options* tran, int32_t precision,
                  int64_t* n_ops, int64_t* code_bytes) {
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
        if (n_ops) *n_ops = jit_ops(t);
        const std::vector<char> code = jit_code(jit_source(t, precision == S21_MIXED, 2));
        if (code_bytes) *code_bytes = (int64_t)code.size();
    });
}

int s21_jit_build_ac(const s21_netlist* net, const s21_options* opts, double fstart, int32_t precision,
                     int64_t* n_ops, int64_t* code_bytes) {
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
        AcInput in;
        in.omega0 = 2.0 * M_PI * fstart;
        for (size_t i = 0; i < g.comps.size(); i++)
            if (g.comps[i].type == S21_COMP_MOS1) in.dc_comp.push_back((int)i);
        in.acop.assign((size_t)M1_ACOP_F * in.dc_comp.size(), 0.0);
        HostPrep hp;
        host_prepare(o, g, nullptr, hp, &in);
        const Topology& t = g.topo;
        if (t.sym_status != S21_OK) fail(t.sym_status, t.sym_error);
        if (n_ops) *n_ops = jit_ops(t);
        const std::vector<char> code =
            jit_code(jit_source(std::vector<const Topology*>{&t}, std::vector<int>{0}, precision == S21_MIXED, 2, true));
        if (code_bytes) *code_bytes = (int64_t)code.size();
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
        for (const GroupHost& g : b->groups)
            for (const s21_comp& c : g.comps) b->has_b4 = b->has_b4 || c.type == S21_COMP_BSIM4;
        HIPCHECK(hipGetDevice(&b->device));
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
        recover_pivots(b, nullptr, 1);
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
        recover_pivots(b, tr, npts);
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
    return guard([&] {
        if (!b || !a) fail(S21_INVALID, "null argument");
        // Frequencies (analysis.rs:786-826): f *= fstep up to fstop, both ends included.
        std::vector<double> freqs;
        {
            double f = (double)a->fstart;
            const double fstop = (double)a->fstop;
            const double fstep = std::pow(10.0, std::log10(fstop / f) / (double)a->npts);
            while (f <= fstop) {
                freqs.push_back(f);
                if (f == fstop) break;
                f = std::fmin(f * fstep, fstop);
            }
        }
        if (freqs.empty()) fail(S21_INVALID, "empty frequency range");
        const int64_t F = (int64_t)freqs.size();
        for (GroupHost& g : b->groups)
            for (const s21_comp& c : g.comps)
                if (c.type != S21_COMP_R && c.type != S21_COMP_C && c.type != S21_COMP_V && c.type != S21_COMP_MOS1)
                    fail(S21_ERR, "AC Not Implemented For This Component!");  // comps/mod.rs:86-88

        // Operating point of every instance, each Mos1's small-signal operating point captured.
        prepare_all(b, nullptr);
        for (GroupHost& g : b->groups) {
            const int nm = g.topo.cls_start[CLS_MOS1T + 1] - g.topo.cls_start[CLS_MOS1];
            if (!g.d_acop) g.d_acop.reset(new DevMem());
            g.d_acop->alloc(std::max<size_t>((size_t)M1_ACOP_F * g.n_inst * nm, 1) * sizeof(double));
            g.gg.acop = (double*)g.d_acop->p;
        }
        Args A = base_args(b, MODE_OP);
        A.capture = 1;
        run(b, A, nullptr, 1);
        std::vector<std::vector<int>> dc_status(b->groups.size());
        std::vector<AcInput> acin(b->groups.size());
        for (size_t gi = 0; gi < b->groups.size(); gi++) {
            GroupHost& g = b->groups[gi];
            const Topology& t = g.topo;
            dc_status[gi].resize(g.n_inst);
            HIPCHECK(hipMemcpy(dc_status[gi].data(), g.gg.status, g.n_inst * sizeof(int), hipMemcpyDeviceToHost));
            AcInput& in = acin[gi];
            in.omega0 = 2.0 * M_PI * freqs[0];
            in.dc_comp.assign(t.dev_comp.begin() + t.cls_start[CLS_MOS1], t.dev_comp.begin() + t.cls_start[CLS_MOS1T + 1]);
            in.acop.resize((size_t)M1_ACOP_F * g.n_inst * in.dc_comp.size());
            if (!in.acop.empty())
                HIPCHECK(hipMemcpy(in.acop.data(), g.d_acop->p, in.acop.size() * sizeof(double), hipMemcpyDeviceToHost));
        }

        // AC structure (pivots on instance 0's matrix at fstart) and one launch over every frequency.
        for (size_t gi = 0; gi < b->groups.size(); gi++) {
            GroupHost& g = b->groups[gi];
            g.prepared.clear();
            prepare_group(b, g, nullptr, &acin[gi]);
            g.prepared = "ac";
        }
        prepare_jit(b, true, F);  // the complex factorisation written out, or the precompiled kernel

        // Runs of frequencies.  The small-signal system does not depend on x (every load_ac stamp comes
        // from the captured operating point), so a frequency whose Newton loop takes at least one full
        // step ends at that frequency's unique solution (up to rounding) whatever its start.  The warm
        // start matters only where the reference's test passes at once: a frequency whose residual at
        // the previous solution is already below 1e-9 keeps that solution.  Each instance's sweep is cut
        // into contiguous runs, solved in parallel from x = 0 (first pass); each run is then re-solved
        // from the previous run's end until both paths take a full step at the same frequency, after
        // which they agree (second pass); a run whose re-solve never meets its first pass is continued
        // in order (third pass, a workgroup per instance whose wavefronts test the next frequencies for
        // the reference's immediate acceptance together).  The result is the sequential sweep's, and so
        // are the iteration counts.
        int runs = b->ac_runs;
        if (runs <= 0) {  // about sixteen wavefronts per CU in the first pass
            int dev = 0, cus = 0;
            HIPCHECK(hipGetDevice(&dev));
            HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            int64_t inst = 0;
            for (const GroupHost& g : b->groups) inst += g.n_inst;
            runs = (int)std::min<int64_t>((64 * (int64_t)std::max(cus, 1) + inst - 1) / inst, std::max<int64_t>(F / 16, 1));
        }
        runs = (int)std::max<int64_t>(std::min<int64_t>(runs, F), 1);
        for (GroupHost& g : b->groups) {
            const size_t ni = (size_t)g.n_inst, n2 = 2 * (size_t)g.topo.n;
            if (!g.d_ac) g.d_ac.reset(new DevMem());
            g.d_ac->alloc(ni * runs * n2 * sizeof(double) + ni * (size_t)F * sizeof(int) + ni * runs * sizeof(int));
            g.gg.ac_end = (double*)g.d_ac->p;
            g.gg.ac_rec = (int*)(g.gg.ac_end + ni * runs * n2);
            g.gg.ac_state = g.gg.ac_rec + ni * (size_t)F;
            HIPCHECK(hipMemsetAsync(g.gg.ac_rec, 0, ni * (size_t)F * sizeof(int), b->stream));  // before pass 0
        }
        std::vector<double> omega(F);
        for (int64_t p = 0; p < F; p++) omega[p] = 2.0 * M_PI * freqs[p];  // AcState::omega, analysis.rs:801
        DevMem d_omega;
        d_omega.alloc(F * sizeof(double));
        HIPCHECK(hipMemcpy(d_omega.p, omega.data(), F * sizeof(double), hipMemcpyHostToDevice));
        size_t total = 0;
        for (GroupHost& g : b->groups) total += (size_t)g.n_inst * (size_t)F * g.topo.n * 2;
        if (!b->d_wave) b->d_wave.reset(new DevMem());
        b->d_wave->alloc(std::max<size_t>(total, 1) * sizeof(double));
        Args B = base_args(b, MODE_AC);
        B.npts = F;
        B.omega = (const double*)d_omega.p;
        B.ac_runs = runs;
        double ms = 0.0;
        int launches = 0;
        auto pass = [&](int k) {
            B.ac_pass = k;
            run(b, B, (double*)b->d_wave->p, 2 * F);
            ms += b->last_ms;
            launches++;
        };
        pass(0);
        if (runs > 1) {
            pass(1);
            // Third pass from the first run (in frequency order) that neither met its first pass nor failed.
            bool third = false;
            for (GroupHost& g : b->groups) {
                std::vector<int> state((size_t)g.n_inst * runs);
                HIPCHECK(hipMemcpy(state.data(), g.gg.ac_state, state.size() * sizeof(int), hipMemcpyDeviceToHost));
                for (int i = 0; i < g.n_inst; i++) {
                    int* st = state.data() + (size_t)i * runs;
                    int from = 0;
                    for (int r = 1; r < runs; r++) {
                        if (st[r] == AC_MET) continue;
                        if (st[r] == AC_UNMET && r + 1 < runs) from = r + 1;  // run r is the sweep's own
                        break;
                    }
                    st[0] = from;
                    third = third || from > 0;
                }
                HIPCHECK(hipMemcpy(g.gg.ac_state, state.data(), state.size() * sizeof(int), hipMemcpyHostToDevice));
            }
            if (third) pass(2);
        }
        b->last_ms = ms;
        b->kernel_avg_ms = ms / launches;
        b->launches = launches;
        // Each instance's status and iteration count, from its frequencies in order: the sweep stops at
        // the first that fails (analysis.rs:801).
        for (GroupHost& g : b->groups) {
            std::vector<int> rec((size_t)g.n_inst * F);
            HIPCHECK(hipMemcpy(rec.data(), g.gg.ac_rec, rec.size() * sizeof(int), hipMemcpyDeviceToHost));
            std::vector<long long> it1(g.n_inst, 0);
            std::vector<int> st1(g.n_inst, S21_OK);
            for (int i = 0; i < g.n_inst; i++)
                for (int64_t p = 0; p < F; p++) {
                    const int r = rec[(size_t)i * F + p];
                    if (r == 0) fail(S21_ERR, "internal: AC frequency left unsolved");
                    it1[i] += r & 0xFF;
                    if ((r >> 8) != S21_OK) {
                        st1[i] = r >> 8;
                        break;
                    }
                }
            HIPCHECK(hipMemcpy(g.gg.status, st1.data(), g.n_inst * sizeof(int), hipMemcpyHostToDevice));
            HIPCHECK(hipMemcpy(g.gg.iters, it1.data(), g.n_inst * sizeof(long long), hipMemcpyHostToDevice));
        }
        b->last_ac_runs = runs;
        b->fetched = false;
        b->npts = F;
        b->axis = freqs;
        b->last_mode = MODE_AC;
        // An instance whose operating point failed reports that failure.
        fetch_results(b);
        for (size_t gi = 0; gi < b->groups.size(); gi++)
            for (int i = 0; i < b->groups[gi].n_inst; i++)
                if (dc_status[gi][i] != S21_OK) b->groups[gi].status[i] = dc_status[gi][i];
    });
}

// The reference's AC result files (analysis.rs:783-827), for one instance of the last sweep: the streamed
// sequence `stream.ac.json`, one flat array [f, re0, im0, re1, im1, ...] per frequency, and the in-memory
// result `ac.json`, serialised before the reference fills its map: {"signals", "freq", "data" (per
// frequency, [re, im] per signal, as num-complex serialises a Complex), "map": {}}.  Numbers are written
// in the shortest form that reads back to the same double; non-finite ones as null, as serde_json does.
// Unlike the reference, which streams as it sweeps, the files are written once the sweep has succeeded.
static void json_num(std::string& o, double v) {
    if (!std::isfinite(v)) {
        o += "null";
        return;
    }
    char buf[40];
    for (int prec = 1; prec <= 17; prec++) {  // the shortest text that reads back to the same double
        std::snprintf(buf, sizeof buf, "%.*g", prec, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    o += buf;
}
static void json_str(std::string& o, const std::string& s) {
    o += '"';
    for (char c : s) {
        if (c == '"' || c == '\\') {
            o += '\\';
            o += c;
        } else if ((unsigned char)c < 0x20) {
            char buf[8];
            std::snprintf(buf, sizeof buf, "\\u%04x", (unsigned)(unsigned char)c);
            o += buf;
        } else {
            o += c;
        }
    }
    o += '"';
}
static void write_file(const std::string& path, const std::string& text) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) fail(S21_ERR, "cannot write " + path);
    const size_t w = std::fwrite(text.data(), 1, text.size(), f);
    const int c = std::fclose(f);
    if (w != text.size() || c != 0) fail(S21_ERR, "cannot write " + path);
}

int s21_batch_write_ac_results(const s21_batch* b, int32_t g, int32_t inst, const char* dir) {
    return guard([&] {
        if (!b || g < 0 || g >= (int)b->groups.size() || b->last_mode != MODE_AC) fail(S21_INVALID, "no AC sweep to write");
        const GroupHost& gh = b->groups[g];
        if (inst < 0 || inst >= gh.n_inst) fail(S21_INVALID, "instance out of range");
        const int n = gh.topo.n;
        const int64_t np = b->npts;
        std::vector<double> wv((size_t)np * n * 2);
        if (s21_batch_get_wave(b, g, inst, wv.data()) != S21_OK) fail(S21_ERR, "AC waveform unavailable");
        const std::string base = dir && *dir ? std::string(dir) + "/" : std::string();
        std::string st = "[";
        for (int64_t p = 0; p < np; p++) {
            if (p) st += ',';
            st += '[';
            json_num(st, b->axis[p]);
            for (int k = 0; k < n; k++) {
                st += ',';
                json_num(st, wv[((size_t)p * n + k) * 2]);
                st += ',';
                json_num(st, wv[((size_t)p * n + k) * 2 + 1]);
            }
            st += ']';
        }
        st += ']';
        std::string js = "{\"signals\":[";
        for (int k = 0; k < n; k++) {
            if (k) js += ',';
            json_str(js, gh.topo.names[k]);
        }
        js += "],\"freq\":[";
        for (int64_t p = 0; p < np; p++) {
            if (p) js += ',';
            json_num(js, b->axis[p]);
        }
        js += "],\"data\":[";
        for (int64_t p = 0; p < np; p++) {
            if (p) js += ',';
            js += '[';
            for (int k = 0; k < n; k++) {
                if (k) js += ',';
                js += '[';
                json_num(js, wv[((size_t)p * n + k) * 2]);
                js += ',';
                json_num(js, wv[((size_t)p * n + k) * 2 + 1]);
                js += ']';
            }
            js += ']';
        }
        js += "],\"map\":{}}";
        write_file(base + "stream.ac.json", st);
        write_file(base + "ac.json", js);
    });
}

int s21_batch_op_noalloc(s21_batch* b, double* x, int32_t* status, int64_t* iters) {
    return guard([&] {
        if (!b || !x || !status || !iters) fail(S21_INVALID, "null argument");
        if ((uintptr_t)x % alignof(double) || (uintptr_t)status % alignof(int32_t) ||
            (uintptr_t)iters % alignof(int64_t))
            fail(S21_INVALID, "s21_batch_op_noalloc: misaligned output buffer");
        int dev = 0;
        HIPCHECK(hipGetDevice(&dev));
        if (dev != b->device) fail(S21_INVALID, "s21_batch_op_noalloc: the current device is not the batch's");
        for (const GroupHost& g : b->groups)
            if (g.prepared != "op")
                fail(S21_INVALID, "s21_batch_op_noalloc: the batch holds no operating-point structure; "
                                  "one s21_batch_op call on it builds and uploads it");
        auto up = [](size_t v) { return (v + 255) / 256 * 256; };
        std::vector<size_t> st_off;
        size_t off = up(b->groups.size() * sizeof(GroupGPU));
        for (const GroupHost& g : b->groups) {
            st_off.push_back(off);
            off += up(g.d_state->bytes);
        }
        if (off > NOALLOC_SCRATCH_BYTES)
            fail(S21_INVALID, "s21_batch_op_noalloc: the batch's Newton state needs " + std::to_string(off) +
                                  " bytes, the scratch array holds " + std::to_string(NOALLOC_SCRATCH_BYTES));
        char* base = scratch_on_current_device();
        Args A = base_args(b, MODE_OP);
        A.prof = 0;
        LaunchPlan P = plan_launch(b, A, nullptr, 1);
        size_t xo = 0, io = 0;
        for (size_t k = 0; k < b->groups.size(); k++) {
            const GroupHost& g = b->groups[k];
            GroupGPU& gg = P.hg[k];
            char* sp = base + st_off[k];
            const char* s0 = (const char*)g.gg.m1_state;  // the batch's own state buffer, same layout
            gg.m1_state = (double*)sp;
            gg.d_state = (double*)(sp + ((const char*)g.gg.d_state - s0));
            gg.c_q = (double*)(sp + ((const char*)g.gg.c_q - s0));
            gg.x = x + xo;
            gg.status = status + io;
            gg.iters = (long long*)(iters + io);
            gg.dbg = nullptr;
            gg.prof = nullptr;
            gg.resume_at = nullptr;
            gg.acop = nullptr;
            gg.wave = nullptr;
            gg.ac_rec = nullptr;
            gg.ac_end = nullptr;
            gg.ac_state = nullptr;
            xo += (size_t)g.n_inst * g.topo.n;
            io += (size_t)g.n_inst;
        }
        // Default operating points (zeros) and the reference's start x = 0: the array keeps the last
        // call's contents, and the caller's buffers hold whatever they held.
        zero_async(base + st_off[0], off - st_off[0], b->stream);
        for (const GroupHost& g : b->groups)  // BSIM4 points stay in the batch's own buffer (no allocation)
            if (g.gg.b4_st) {
                char* p0 = (char*)g.gg.b4_st;
                zero_async(p0, g.d_b4->bytes - (size_t)(p0 - (char*)g.d_b4->p), b->stream);
            }
        zero_async(x, xo * sizeof(double), b->stream);
        zero_async(status, io * sizeof(int32_t), b->stream);
        zero_async(iters, io * sizeof(int64_t), b->stream);
        HIPCHECK(hipMemcpyAsync(base, P.hg.data(), P.hg.size() * sizeof(GroupGPU), hipMemcpyHostToDevice, b->stream));
        launch_plan(b, P, (const GroupGPU*)base);
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
        if ((b->last_mode != MODE_TRAN && b->last_mode != MODE_AC) || !b->d_wave)
            fail(S21_INVALID, "no transient or AC waveform");
        const GroupHost& gh = b->groups[g];
        if (inst < 0 || inst >= gh.n_inst) fail(S21_INVALID, "instance out of range");
        const size_t vw = b->last_mode == MODE_AC ? 2 : 1;  // AC points are complex (re, im)
        size_t off = 0;  // groups' waveforms follow each other in group order (see run())
        for (int gi = 0; gi < g; gi++)
            off += (size_t)b->groups[gi].n_inst * (size_t)b->npts * b->groups[gi].topo.n * vw;
        const size_t per = (size_t)b->npts * gh.topo.n * vw;
        HIPCHECK(hipMemcpy(out, (double*)b->d_wave->p + off + (size_t)inst * per, per * sizeof(double),
                           hipMemcpyDeviceToHost));
    });
}
int s21_batch_resumptions(const s21_batch* b, int64_t* instance_rounds) {
    return guard([&] {
        if (!b || !instance_rounds) fail(S21_INVALID, "bad arguments");
        *instance_rounds = b->resumed;
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
int s21_batch_set_ipw(s21_batch* b, int32_t ipw) {
    return guard([&] {
        if (!b) fail(S21_INVALID, "null batch");
        if (ipw != 0 && ipw != 1 && ipw != 2 && ipw != 4 && ipw != 8 && ipw != 16)
            fail(S21_INVALID, "instances per wavefront must be 0, 1, 2, 4, 8 or 16");
        b->ipw = ipw;
    });
}
int s21_batch_set_jit(s21_batch* b, int32_t mode) {
    return guard([&] {
        if (!b) fail(S21_INVALID, "null batch");
        if (mode < -1 || mode > 1) fail(S21_INVALID, "generated-factorisation mode must be -1, 0 or 1");
        if (mode != b->jit)
            for (GroupHost& g : b->groups) g.prepared.clear();
        b->jit = mode;
    });
}
int s21_batch_set_waves(s21_batch* b, int32_t waves) {
    return guard([&] {
        if (!b) fail(S21_INVALID, "null batch");
        if (waves < 0 || waves > MAX_WAVES) fail(S21_INVALID, "wavefronts per workgroup must be 0 (auto) to 8");
        b->waves = waves;
    });
}
int s21_batch_set_hub(s21_batch* b, int32_t policy) {
    return guard([&] {
        if (!b) fail(S21_INVALID, "null batch");
        if (policy < 0 || policy > 2) fail(S21_INVALID, "hub policy must be 0 (auto), 1 or 2");
        if (policy != b->hub)
            for (GroupHost& g : b->groups) g.prepared.clear();
        b->hub = policy;
    });
}
int s21_batch_set_ac_runs(s21_batch* b, int32_t runs) {
    return guard([&] {
        if (!b) fail(S21_INVALID, "null batch");
        if (runs < 0) fail(S21_INVALID, "runs of frequencies must be 0 (chosen per sweep) or positive");
        b->ac_runs = runs;
    });
}
int32_t s21_batch_last_ac_runs(const s21_batch* b) { return b ? b->last_ac_runs : 0; }
int s21_batch_config(const s21_batch* b, int32_t* ipw, int32_t* waves, int32_t* jit) {
    return guard([&] {
        if (!b) fail(S21_INVALID, "null batch");
        if (ipw) *ipw = b->last_ipw;
        if (waves) *waves = b->last_waves;
        if (jit) *jit = b->last_jit ? 1 : 0;
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
        if (n_contrib) *n_contrib = t.n_hub;
    });
}
int s21_batch_assemble(s21_batch* b, int32_t g, int32_t inst, const double* x, int32_t* row_out, int32_t* col_out,
                       double* val_out, double* rhs_out, int64_t cap, int64_t* n_elem) {
    return guard([&] {
        if (!b || g < 0 || g >= (int)b->groups.size() || !x) fail(S21_INVALID, "bad arguments");
        GroupHost& gh = b->groups[g];
        if (inst < 0 || inst >= gh.n_inst) fail(S21_INVALID, "instance out of range");
        prepare_all(b, nullptr);
        const Topology& t = gh.topo;
        const int64_t ne = (int64_t)t.erow.size();
        if (ne > cap) fail(S21_INVALID, "element buffers hold " + std::to_string(cap) + " entries, the matrix has " +
                                            std::to_string(ne));
        // Every write below is on the batch's stream, after the state resets prepare_all queued there.
        HIPCHECK(hipMemcpyAsync(gh.gg.x + (size_t)inst * t.n, x, t.n * sizeof(double), hipMemcpyHostToDevice,
                                b->stream));
        Args A = base_args(b, MODE_ASSEMBLE);
        A.dbg_inst = inst;
        A.waves = 1;
        A.ipw = 1;
        const size_t stride = inst_stride(gh.inst_bytes, 1);
        A.inst_bytes = (int)stride;
        GroupGPU gg = gh.gg;
        gg.blk0 = 0;
        if (!b->d_groups) b->d_groups.reset(new DevMem());
        if (b->d_groups->bytes < sizeof(GroupGPU)) b->d_groups->alloc(sizeof(GroupGPU));
        HIPCHECK(hipMemcpyAsync(b->d_groups->p, &gg, sizeof(GroupGPU), hipMemcpyHostToDevice, b->stream));
        b->groups_uploaded.clear();
        launch_any(b, (const GroupGPU*)b->d_groups->p, 1, 1, gh.blob_bytes + stride, A);
        std::vector<double> dbg(t.nnz + t.n);
        HIPCHECK(hipMemcpyAsync(dbg.data(), gh.gg.dbg, dbg.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
        HIPCHECK(hipStreamSynchronize(b->stream));
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
