// Kernels compiled at run time for one frozen pattern.
//
// The level-scheduled factorisation walks 16-bit records in LDS, so every dependent step pays a record
// load before its data load.  For a batch whose instances share one pattern, this file writes the
// factorisation with the fused forward sweep and the backward sweep out as straight-line code -- every
// LU position a constant offset -- in the order of the records and with the same division (ldiv)
// (IKJ by rows in pivot order:
// l_ik = a_ik / u_kk, the row-k updates, c_i -= c_k l_ik skipped when c_k is zero; then rows from the last,
// c_i = (c_i - sum u_ij c_j) / u_ii; sparse21/mod.rs:865-919, 946-979), compiles it together with the
// kernel body of kernel.h through hiprtc, and keeps the module for the life of the process.  One lane per
// instance runs the generated code, every instance of a workgroup at once (newton_wg).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstdlib>
#include <functional>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "internal.h"
#include "jit.h"

namespace s21 {

// Headers of the kernel, embedded at build time (jit_src.cpp).
extern const int kJitHeaderCount;
extern const char* const kJitHeaderNames[];
extern const char* const kJitHeaderSources[];

int64_t jit_ops(const Topology& t) {
    int64_t ops = 0;
    for (int i = 0; i < t.n; i++) {
        for (int q = t.rp[i]; t.diag[i] >= 0 && q < t.diag[i]; q++) ops += 2 + (t.upd_ptr[q + 1] - t.upd_ptr[q]);
        ops += 1 + (t.rp[i + 1] - t.diag[i] - 1);
    }
    return ops;
}

// Two shapes of the generated code, chosen by precision (both measured on the 64-stage chain batch):
//
// * FP64: straight-line statements on the LDS arrays, which the compiler schedules freely.
// * FP32 (mixed precision): values in named locals, in steps -- a row of the factorisation (long rows in
//   chunks of kChunk L entries) or of the backward sweep.  A step takes what the previous kCarry steps
//   computed from their locals and loads everything else from the LDS; those loads are issued one step
//   ahead, before the previous step's arithmetic, so their latency overlaps it.  A step stores what later
//   code reads (a finished U row, the swept RHS).  A scheduling barrier between steps keeps the compiler
//   from hoisting the loads of many steps at once, which would need more registers than a lane has.  It
//   works on FP32 copies of A and of the residual; updates are fused multiply-adds, a zero c_k is not
//   skipped, and each pivot's reciprocal (the one ldiv multiplies by) is taken once, stored in place of
//   the pivot and reused by every row and by the backward sweep: this factorisation only preconditions
//   the FP64 Newton update.
//   (With FP64 values, two registers each, this shape spills and runs slower than the straight one.)
//
// Both keep the records' operation order and division (ldiv); FP64 performs exactly their arithmetic.
constexpr int kCarry = 1;
constexpr int kChunk = 4;

namespace {
struct GenStep {
    std::string loads, body;
};
// Emits steps with their loads one step ahead, a scheduling barrier after each.
std::string emit_steps(const std::vector<GenStep>& st) {
    std::ostringstream o;
    if (!st.empty()) o << st[0].loads;
    for (size_t k = 0; k < st.size(); k++) {
        if (k + 1 < st.size()) o << st[k + 1].loads;
        o << st[k].body << "    __builtin_amdgcn_sched_barrier(0);\n";
    }
    return o.str();
}

// `fn` opens each generated function: its qualifiers and the name prefix (lu_gen / bw_gen follow it).
void gen_straight(std::ostringstream& o, const Topology& t, const std::string& fn, const std::string& sfx) {
    o << "template <typename V>\n" << fn << " bool lu_gen" << sfx
      << "(LDS V* __restrict__ lu, LDS V* __restrict__ c) {\n"
         "    V y, l, ck;\n";
    for (int i = 0; i < t.n; i++) {
        if (t.diag[i] < 0 || t.diag[i] == t.rp[i]) continue;  // no L entries
        o << "    y = (V)c[" << i << "];\n";
        for (int q = t.rp[i]; q < t.diag[i]; q++) {
            const int k = t.ci[q];
            o << "    l = ldiv<V>(lu[" << q << "], lu[" << t.diag[k] << "]); lu[" << q << "] = l; ck = (V)c[" << k
              << "]; if (ck != V(0)) y = y - ck * l;\n";
            for (int x = t.upd_ptr[q]; x < t.upd_ptr[q + 1]; x++)
                o << "    lu[" << t.upd_dst[x] << "] = lu[" << t.upd_dst[x] << "] - lu[" << t.upd_src[x] << "] * l;\n";
        }
        o << "    c[" << i << "] = (double)y;\n";
    }
    o << "    bool z = false;\n";  // a pivot zero or below the smallest normal number (factor_forward)
    for (int i = 0; i + 1 < t.n; i++) o << "    z = z || pivot_unusable(lu[" << t.diag[i] << "]);\n";
    o << "    return z;\n}\n";
    o << "template <typename V>\n" << fn << " void bw_gen" << sfx
      << "(LDS const V* __restrict__ lu, LDS V* __restrict__ c) {\n"
         "    V y;\n";
    for (int i = t.n - 1; i >= 0; i--) {
        o << "    y = (V)c[" << i << "];";
        for (int q = t.diag[i] + 1; q < t.rp[i + 1]; q++) o << " y = y - (V)c[" << t.ci[q] << "] * lu[" << q << "];";
        o << " c[" << i << "] = (double)ldiv<V>(y, lu[" << t.diag[i] << "]);\n";
    }
    o << "}\n";
}

void gen_steps(std::ostringstream& o, const Topology& t, const std::string& fn, const std::string& sfx) {
    const int NEVER = -1000000;
    // ---- factorisation + forward sweep ----
    {
        std::vector<GenStep> steps;
        std::ostringstream pre;  // pivots of rows without L entries: checked and inverted up front
        std::vector<int> def_step(t.nnz, NEVER), cdef_step(t.n, NEVER);  // step whose local holds the value
        std::vector<char> used(t.nnz, 0), cused(t.n, 0);
        int step = -1;
        for (int i = 0; i < t.n; i++) {
            const bool has_l = t.diag[i] >= 0 && t.diag[i] > t.rp[i];
            if (!has_l) {
                if (t.diag[i] >= 0) {
                    pre << "    l = lu[" << t.diag[i] << "];\n";
                    if (i + 1 < t.n) pre << "    m = fminf(m, fabsf(l));\n";
                    pre << "    lu[" << t.diag[i] << "] = __builtin_amdgcn_rcpf(l);\n";
                }
                continue;
            }
            const int row_first = step + 1;  // row i's own values, once loaded, stay in locals for the row
            for (int q0 = t.rp[i]; q0 < t.diag[i]; q0 += kChunk) {
                step++;
                GenStep g;
                std::ostringstream loads, body;
                auto fresh = [&](int defd, bool own) { return own ? defd >= row_first : step - defd <= kCarry; };
                auto use = [&](int pos) {
                    const bool own = pos >= t.rp[i] && pos < t.rp[i + 1];
                    if (!fresh(def_step[pos], own)) {
                        loads << "    v" << pos << " = lu[" << pos << "];\n";
                        def_step[pos] = step;
                    }
                    used[pos] = 1;
                    return "v" + std::to_string(pos);
                };
                auto usec = [&](int k) {
                    if (!fresh(cdef_step[k], k == i)) {
                        loads << "    y" << k << " = c[" << k << "];\n";
                        cdef_step[k] = step;
                    }
                    cused[k] = 1;
                    return "y" + std::to_string(k);
                };
                const std::string yi = usec(i);
                const int q1 = std::min(q0 + kChunk, t.diag[i]);
                for (int q = q0; q < q1; q++) {
                    const int k = t.ci[q];
                    const std::string a = use(q), rpiv = use(t.diag[k]), yk = usec(k);  // pivot reciprocal
                    body << "    l = " << a << " * " << rpiv << "; " << yi
                         << " = __builtin_fmaf(-" << yk << ", l, " << yi << ");\n";
                    for (int x = t.upd_ptr[q]; x < t.upd_ptr[q + 1]; x++) {
                        const std::string src = use(t.upd_src[x]), d = use(t.upd_dst[x]);
                        body << "    " << d << " = __builtin_fmaf(-" << src << ", l, " << d << ");\n";
                        def_step[t.upd_dst[x]] = step;
                    }
                }
                cdef_step[i] = step;
                if (q1 == t.diag[i]) {  // row finished: its U row (the pivot inverted) and swept RHS
                    const std::string d = use(t.diag[i]);
                    if (i + 1 < t.n) body << "    m = fminf(m, fabsf(" << d << "));\n";
                    body << "    " << d << " = __builtin_amdgcn_rcpf(" << d << ");\n";
                    def_step[t.diag[i]] = step;
                    for (int q = t.diag[i]; q < t.rp[i + 1]; q++)
                        if (def_step[q] >= row_first) {
                            body << "    lu[" << q << "] = v" << q << ";\n";
                            def_step[q] = step;
                        }
                    body << "    c[" << i << "] = " << yi << ";\n";
                }
                g.loads = loads.str();
                g.body = body.str();
                steps.push_back(std::move(g));
            }
        }
        o << "template <typename V>\n" << fn << " bool lu_gen" << sfx
          << "(LDS V* __restrict__ lu, LDS V* __restrict__ c) {\n"
             "    V l, m = __builtin_inff();\n";
        for (int q = 0; q < t.nnz; q++)
            if (used[q]) o << "    V v" << q << ";\n";
        for (int k = 0; k < t.n; k++)
            if (cused[k]) o << "    V y" << k << ";\n";
        o << pre.str() << emit_steps(steps) << "    return pivot_unusable(m);\n}\n";
    }
    // ---- backward sweep ----
    {
        std::vector<GenStep> steps;
        std::vector<int> xdef(t.n, NEVER);
        int step = -1;
        for (int i = t.n - 1; i >= 0; i--) {
            const int u0 = t.diag[i] + 1, u1 = t.rp[i + 1];
            int q0 = u0;
            do {
                step++;
                GenStep g;
                std::ostringstream loads, body;
                const int q1 = std::min(q0 + 2 * kChunk, u1);
                if (q0 == u0) loads << "    y" << i << " = c[" << i << "];\n";
                for (int q = q0; q < q1; q++) {
                    const int j = t.ci[q];
                    if (step - xdef[j] > kCarry) {
                        loads << "    x" << j << " = c[" << j << "];\n";
                        xdef[j] = step;
                    }
                    loads << "    u" << q << " = lu[" << q << "];\n";
                    body << "    y" << i << " = __builtin_fmaf(-x" << j << ", u" << q << ", y" << i << ");\n";
                }
                if (q1 == u1) {
                    loads << "    u" << t.diag[i] << " = lu[" << t.diag[i] << "];\n";
                    body << "    x" << i << " = y" << i << " * u" << t.diag[i] << ";\n    c[" << i << "] = x" << i
                         << ";\n";
                    xdef[i] = step;
                }
                g.loads = loads.str();
                g.body = body.str();
                steps.push_back(std::move(g));
                q0 = q1;
            } while (q0 < u1);
        }
        o << "template <typename V>\n" << fn << " void bw_gen" << sfx
          << "(LDS const V* __restrict__ lu, LDS V* __restrict__ c) {\n";
        for (int q = 0; q < t.nnz; q++) o << "    V u" << q << ";\n";
        for (int k = 0; k < t.n; k++) o << "    V x" << k << ", y" << k << ";\n";
        o << emit_steps(steps) << "}\n";
    }
}
// The complex factorisation and sweeps of an AC pattern, straight-line on the interleaved (re, im)
// arrays, with the record-driven code's operations (factor_forward_ac / backward_ac in kernel.h): the
// num-complex naive division, c_k skipped when exactly zero.  R is the component type (double, or float
// for the mixed-precision copies, whose division multiplies by the norm's reciprocal).
void gen_cx(std::ostringstream& o, const Topology& t, const std::string& fn, const std::string& sfx) {
    o << "template <typename R>\n" << fn << " bool lu_gen" << sfx
      << "(LDS R* __restrict__ lu, LDS R* __restrict__ c) {\n"
         "    using V = Z<R>;\n"
         "    V y, l, ck;\n";
    for (int i = 0; i < t.n; i++) {
        if (t.diag[i] < 0 || t.diag[i] == t.rp[i]) continue;  // no L entries
        o << "    y = ldz(c, " << i << ");\n";
        for (int q = t.rp[i]; q < t.diag[i]; q++) {
            const int k = t.ci[q];
            o << "    l = zdiv(ldz(lu, " << q << "), ldz(lu, " << t.diag[k] << ")); stz(lu, " << q << ", l); ck = ldz(c, "
              << k << "); if (!zzero(ck)) y = y - ck * l;\n";
            for (int x = t.upd_ptr[q]; x < t.upd_ptr[q + 1]; x++)
                o << "    stz(lu, " << t.upd_dst[x] << ", ldz(lu, " << t.upd_dst[x] << ") - ldz(lu, " << t.upd_src[x]
                  << ") * l);\n";
        }
        o << "    stz(c, " << i << ", y);\n";
    }
    o << "    bool z = false;\n";
    for (int i = 0; i + 1 < t.n; i++) o << "    z = z || zzero(ldz(lu, " << t.diag[i] << "));\n";
    o << "    return z;\n}\n";
    // The forward sweep alone, on the factors lu_gen left (later Newton iterations at one frequency).
    o << "template <typename R>\n" << fn << " void fw_gen" << sfx
      << "(LDS const R* __restrict__ lu, LDS R* __restrict__ c) {\n"
         "    using V = Z<R>;\n"
         "    V y, ck;\n";
    for (int i = 0; i < t.n; i++) {
        if (t.diag[i] < 0 || t.diag[i] == t.rp[i]) continue;
        o << "    y = ldz(c, " << i << ");";
        for (int q = t.rp[i]; q < t.diag[i]; q++)
            o << " ck = ldz(c, " << t.ci[q] << "); if (!zzero(ck)) y = y - ck * ldz(lu, " << q << ");";
        o << " stz(c, " << i << ", y);\n";
    }
    o << "}\n";
    o << "template <typename R>\n" << fn << " void bw_gen" << sfx
      << "(LDS const R* __restrict__ lu, LDS R* __restrict__ c) {\n"
         "    using V = Z<R>;\n"
         "    V y;\n";
    for (int i = t.n - 1; i >= 0; i--) {
        o << "    y = ldz(c, " << i << ");";
        for (int q = t.diag[i] + 1; q < t.rp[i + 1]; q++) o << " y = y - ldz(c, " << t.ci[q] << ") * ldz(lu, " << q << ");";
        o << " stz(c, " << i << ", zdiv(y, ldz(lu, " << t.diag[i] << ")));\n";
    }
    o << "}\n";
}
}  // namespace

std::string jit_source(const std::vector<const Topology*>& ts, const std::vector<int>& group_pattern, bool mixed,
                       int waves_per_simd, bool complex, bool profile) {
    std::ostringstream o;
    o << "#include \"kernel.h\"\nnamespace s21 {\n";
    // One pattern: the functions inline into the kernel.  Several: each stays a function of its own, so
    // the kernel's registers are not allocated for all of them at once.
    const std::string fn = ts.size() == 1 ? "__device__ inline" : "__device__ __noinline__";
    for (size_t p = 0; p < ts.size(); p++) {
        if (complex)
            gen_cx(o, *ts[p], fn, "_" + std::to_string(p));
        else if (mixed)
            gen_steps(o, *ts[p], fn, "_" + std::to_string(p));
        else
            gen_straight(o, *ts[p], fn, "_" + std::to_string(p));
    }
    // group of the batch -> its pattern's functions
    auto dispatch = [&](const char* ret, const char* name) {
        std::ostringstream d;
        if (ts.size() == 1) {
            d << "        " << ret << name << "_0<V>(lu, c);\n";
            return d.str();
        }
        d << "        switch (g) {\n";
        for (size_t p = 0; p < ts.size(); p++) {
            for (size_t g = 0; g < group_pattern.size(); g++)
                if (group_pattern[g] == (int)p) d << "        case " << g << ":\n";
            d << "            " << ret << name << "_" << p << "<V>(lu, c);" << (ret[0] ? "" : " break;") << "\n";
        }
        d << "        }\n";
        if (ret[0]) d << "        return true;\n";
        return d.str();
    };
    o << "struct LuGen {\n"
         "    static constexpr bool lane_per_instance = true;\n"
         "    static constexpr int kind = "
      << (complex ? 2 : 1) << ";\n"
         "    static constexpr bool profile = false;\n"
         "    static constexpr bool bsim4 = false;\n"
         "    template <typename V>\n"
         "    static __device__ bool factor_lane(int g, LDS V* lu, LDS V* c) {\n"
      << dispatch("return ", "lu_gen")
      << "    }\n"
         "    template <typename V>\n"
         "    static __device__ void backward_lane(int g, LDS const V* lu, LDS V* c) {\n"
      << dispatch("", "bw_gen")
      << "    }\n";
    if (complex)
        o << "    template <typename V>\n"
             "    static __device__ void forward_lane(int g, LDS const V* lu, LDS V* c) {\n"
          << dispatch("", "fw_gen") << "    }\n";
    o << "};\n}  // namespace s21\n";
    o << "extern \"C\" __global__ __launch_bounds__(s21::WAVE * s21::MAX_WAVES, " << waves_per_simd
      << ") void s21_newton_jit("
         "const s21::GroupGPU* __restrict__ groups, int n_groups, s21::Args A) {\n"
         "    s21::newton_body<"
      << (mixed ? "true" : "false") << (profile ? ", s21::Profiled<s21::LuGen>>" : ", s21::LuGen>")
      << "(groups, n_groups, A);\n}\n";
    return o.str();
}

std::string jit_source(const Topology& t, bool mixed, int waves_per_simd) {
    return jit_source(std::vector<const Topology*>{&t}, std::vector<int>{0}, mixed, waves_per_simd, false);
}

namespace {
struct Cache {
    std::mutex mu;
    std::unordered_map<std::string, JitKernel> by_source;
};
Cache& cache() {
    static Cache c;
    return c;
}
}  // namespace

std::vector<char> jit_code(const std::string& src) {
    if (const char* dump = std::getenv("S21_JIT_DUMP")) {  // debugging aid: the generated source
        if (FILE* f = std::fopen(dump, "w")) {
            std::fwrite(src.data(), 1, src.size(), f);
            std::fclose(f);
        }
    }
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "s21_newton_jit.hip", kJitHeaderCount, kJitHeaderSources,
                            kJitHeaderNames) != HIPRTC_SUCCESS)
        fail(S21_DEVICE, "hiprtc: cannot create program");
    const char* opts[] = {"--offload-arch=gfx950", "-O3",     "-std=c++17",      "-ffp-contract=off",
                          "-Xclang",               "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult r = hiprtcCompileProgram(prog, 8, opts);
    if (r != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        fail(S21_DEVICE, "hiprtc: compilation failed: " + log.substr(0, 2000));
    }
    size_t n = 0;
    hiprtcGetCodeSize(prog, &n);
    std::vector<char> code(n);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return code;
}

JitKernel jit_compile(const std::string& src) {
    Cache& c = cache();
    {
        std::lock_guard<std::mutex> lk(c.mu);
        auto it = c.by_source.find(src);
        if (it != c.by_source.end()) return it->second;
    }
    const std::vector<char> code = jit_code(src);
    JitKernel k;
    if (hipModuleLoadData(&k.mod, code.data()) != hipSuccess) fail(S21_DEVICE, "hipModuleLoadData failed");
    if (hipModuleGetFunction(&k.fn, k.mod, "s21_newton_jit") != hipSuccess) fail(S21_DEVICE, "hipModuleGetFunction failed");
    std::lock_guard<std::mutex> lk(c.mu);
    return c.by_source.emplace(src, k).first->second;
}

}  // namespace s21
