// Library-level C ABI entry points: error text, defaults and point counts.
#include <cmath>
#include <cstring>

#include "internal.h"

namespace s21 {
static thread_local std::string g_last_error;
void set_last_error(const std::string& s) { g_last_error = s; }
}  // namespace s21

extern "C" {

int32_t s21_abi_version(void) { return S21_ABI_VERSION; }

size_t s21_last_error(char* buf, size_t len) {
    const std::string& e = s21::g_last_error;
    if (buf && len) {
        size_t n = e.size() < len - 1 ? e.size() : len - 1;
        memcpy(buf, e.data(), n);
        buf[n] = 0;
    }
    return e.size();
}

void s21_default_options(s21_options* o) {  // analysis.rs:665-693
    if (!o) return;
    o->temp = 300.15;
    o->tnom = 300.15;
    o->gmin = 1e-12;
    o->iabstol = 1e-12;
    o->reltol = 1e-3;
}

int64_t s21_tran_point_count(double tstep, double tstop) {  // analysis.rs:551-570
    if (!(tstep > 0.0)) return 1;
    int64_t n = 1;
    double t = tstep;
    while (t < tstop && n < (int64_t)1e9) {
        n++;
        t += tstep;
    }
    return n;
}

int64_t s21_ac_point_count(const s21_ac_options* a) {  // analysis.rs:791-819
    if (!a) return 0;
    double f = (double)a->fstart;
    const double fstop = (double)a->fstop;
    const double fstep = pow(10.0, log10(fstop / f) / (double)a->npts);
    int64_t n = 0;
    while (f <= fstop) {
        n++;
        if (f == fstop || n > (int64_t)1e8) break;
        f = fmin(f * fstep, fstop);
    }
    return n;
}

}  // extern "C"
