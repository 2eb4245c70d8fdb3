// Run-time compiled Newton kernels for one frozen pattern (jit.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "internal.h"

namespace s21 {

struct JitKernel {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
};
// Statements of the generated factorisation and sweeps (its size).
int64_t jit_ops(const Topology& t);
// Source of the kernel with the factorisation of t written out, for FP64 or mixed precision, compiled
// for at least `waves_per_simd` resident wavefronts per SIMD (its register budget).
std::string jit_source(const Topology& t, bool mixed, int waves_per_simd);
// The same for a batch of several patterns: ts[p] for pattern p, group_pattern[g] the pattern of group g
// (its factorisation is chosen by the group index at run time).  complex: the patterns are AC patterns
// and the kernel runs AC launches only (otherwise it runs operating points and transients).
// profile: the kernel carries the phase-cycle counters (Profiled<LuGen>).
std::string jit_source(const std::vector<const Topology*>& ts, const std::vector<int>& group_pattern, bool mixed,
                       int waves_per_simd, bool complex, bool profile = false);
// Code object of a source for gfx950 (no device needed).
std::vector<char> jit_code(const std::string& src);
// Compiled kernel for a source (compiled once per process and source).
JitKernel jit_compile(const std::string& src);

}  // namespace s21
