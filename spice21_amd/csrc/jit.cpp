// Kernels compiled at run time for one frozen pattern.
//
// The level-scheduled factorisation walks 16-bit records in LDS, so every dependent step pays a record
// load before its data load.  For a batch whose instances share one pattern, this file writes the
// factorisation with the fused forward sweep and the backward sweep out as straight-line code -- every
// LU position a constant offset -- in exactly the order of the records and with the same division
// (ldiv) (IKJ by rows in pivot order:
// l_ik = a_ik / u_kk, the row-k updates, c_i -= c_k l_ik skipped when c_k is zero; then rows from the last,
// c_i = (c_i - sum u_ij c_j) / u_ii; sparse21/mod.rs:865-919, 946-979), compiles it together with the
// kernel body of kernel.h through hiprtc, and keeps the module for the life of the process.  One lane per
// instance runs the generated code, every instance of a workgroup at once (newton_wg); the values it
// needs stay in registers.
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

std::string jit_source(const Topology& t, bool mixed, int waves_per_simd) {
    std::ostringstream o;
    o << "#include \"kernel.h\"\nnamespace s21 {\n";
    // factorisation + forward sweep
    o << "template <typename V>\n__device__ inline bool lu_gen(LDS V* __restrict__ lu, LDS double* __restrict__ c) {\n"
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
    o << "    bool z = false;\n";
    for (int i = 0; i + 1 < t.n; i++) o << "    z = z || lu[" << t.diag[i] << "] == V(0);\n";
    o << "    return z;\n}\n";
    // backward sweep
    o << "template <typename V>\n__device__ inline void bw_gen(LDS const V* __restrict__ lu, LDS double* __restrict__ c) {\n"
         "    V y;\n";
    for (int i = t.n - 1; i >= 0; i--) {
        o << "    y = (V)c[" << i << "];";
        for (int q = t.diag[i] + 1; q < t.rp[i + 1]; q++) o << " y = y - (V)c[" << t.ci[q] << "] * lu[" << q << "];";
        o << " c[" << i << "] = (double)ldiv<V>(y, lu[" << t.diag[i] << "]);\n";
    }
    o << "}\n";
    o << "struct LuGen {\n"
         "    static constexpr bool lane_per_instance = true;\n"
         "    template <typename V>\n"
         "    static __device__ bool factor_lane(LDS V* lu, LDS double* c) { return lu_gen<V>(lu, c); }\n"
         "    template <typename V>\n"
         "    static __device__ void backward_lane(LDS const V* lu, LDS double* c) { bw_gen<V>(lu, c); }\n"
         "};\n}  // namespace s21\n";
    o << "extern \"C\" __global__ __launch_bounds__(s21::WAVE * s21::MAX_WAVES, " << waves_per_simd
      << ") void s21_newton_jit("
         "const s21::GroupGPU* __restrict__ groups, int n_groups, s21::Args A) {\n"
         "    s21::newton_body<"
      << (mixed ? "true" : "false") << ", s21::LuGen>(groups, n_groups, A);\n}\n";
    return o.str();
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
