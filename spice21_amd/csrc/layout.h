// Device classes and per-device record layout shared by the host symbolic phase and the kernels.
#pragma once
#include "models.h"

namespace s21 {

// Device classes in kernel order.  CLS_MOS1T is a Mos1 without terminal resistances whose source is tied
// to its bulk: its coinciding entries are merged in registers before they reach the matrix.
enum { CLS_MOS1 = 0, CLS_MOS1T = 1, CLS_MOS0 = 2, CLS_DIODE = 3, CLS_LIN = 4, CLS_BSIM4 = 5, N_CLS = 6 };
// BSIM4 stamp entries (bsim4_eval.h: 38 matrix stamps, then 4 right-hand-side stamps).
constexpr int BSIM4_NE = 42;
__host__ __device__ constexpr bool cls_is_mos1(int cls) { return cls == CLS_MOS1 || cls == CLS_MOS1T; }
// Entries of a tied Mos1 that merge into an earlier entry with the same slot: (s,s) <- 7, 13, 15;
// (d,s) <- 14; (s,d) <- 12; (s,g) <- 11; (g,s) <- 10; rhs s <- 25.
__host__ __device__ constexpr bool mos1t_merged(int e) {
    return e == 7 || e == 13 || e == 15 || e == 14 || e == 12 || e == 11 || e == 10 || e == 25;
}
// Device flags in the per-device record: Mos1 with gate-charge state; Mos0 PMOS; the linear element's
// component type sits in bits 4-7.
enum { DEV_CAPS = 1, DEV_PMOS = 2 };
// Slot codes per device (one per stamp entry) by class.
__host__ __device__ constexpr int code_stride(int cls) {
    return cls_is_mos1(cls)    ? MOS1_NE
           : cls == CLS_MOS0  ? MOS0_NE
           : cls == CLS_DIODE ? DIODE_NE
           : cls == CLS_BSIM4 ? BSIM4_NE
                              : 6;
}

}  // namespace s21
