// The engine's entry points (batched and single-instance) that the byte ABI (circuit.cpp) calls after decoding and elaborating,
// replaced for the sanitizer build of the host code: there is no GPU here, so each reports S21_DEVICE.
#include "spice21_mi355x.h"

extern "C" {
int s21_batch_create(int32_t, const s21_group*, const s21_options*, int32_t, s21_batch** out) {
    if (out) *out = nullptr;
    return S21_DEVICE;
}
void s21_batch_free(s21_batch*) {}
int s21_batch_op(s21_batch*) { return S21_DEVICE; }
int s21_batch_tran(s21_batch*, const s21_tran_options*) { return S21_DEVICE; }
int s21_batch_ac(s21_batch*, const s21_ac_options*) { return S21_DEVICE; }
int s21_batch_get_x(const s21_batch*, int32_t, double*) { return S21_DEVICE; }
int s21_batch_get_status(const s21_batch*, int32_t, int32_t*, int64_t*) { return S21_DEVICE; }
int64_t s21_batch_n_points(const s21_batch*) { return 0; }
int s21_batch_get_axis(const s21_batch*, double*) { return S21_DEVICE; }
int s21_batch_get_wave(const s21_batch*, int32_t, int32_t, double*) { return S21_DEVICE; }
int32_t s21_batch_n_vars(const s21_batch*, int32_t) { return 0; }
const char* s21_batch_var_name(const s21_batch*, int32_t, int32_t) { return nullptr; }
int s21_batch_write_ac_results(const s21_batch*, int32_t, int32_t, const char*) { return S21_DEVICE; }
int s21_large_create(const s21_netlist*, const s21_options*, int32_t, s21_large** out) {
    if (out) *out = nullptr;
    return S21_DEVICE;
}
void s21_large_free(s21_large*) {}
int s21_large_op(s21_large*) { return S21_DEVICE; }
int s21_large_tran(s21_large*, const s21_tran_options*) { return S21_DEVICE; }
int s21_large_get_status(const s21_large*, int32_t*, int64_t*) { return S21_DEVICE; }
int s21_large_get_x(const s21_large*, double*) { return S21_DEVICE; }
int64_t s21_large_n_points(const s21_large*) { return 0; }
int32_t s21_large_n_vars(const s21_large*) { return 0; }
const char* s21_large_var_name(const s21_large*, int32_t) { return nullptr; }
int s21_large_get_axis(const s21_large*, double*) { return S21_DEVICE; }
int s21_large_get_wave(const s21_large*, int32_t, const int32_t*, double*) { return S21_DEVICE; }
}
