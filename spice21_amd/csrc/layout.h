// Device classes and per-device record layout shared by the host symbolic phase and the kernels.
#pragma once
#include "models.h"

namespace s21 {

// Device classes in kernel order.
enum { CLS_MOS1 = 0, CLS_MOS0 = 1, CLS_DIODE = 2, CLS_LIN = 3, N_CLS = 4 };
// Device flags in the per-device record: Mos1 with gate-charge state; Mos0 PMOS; the linear element's
// component type sits in bits 4-7.
enum { DEV_CAPS = 1, DEV_PMOS = 2 };
// Slot codes per device (one per stamp entry) by class.
__host__ __device__ constexpr int code_stride(int cls) {
    return cls == CLS_MOS1 ? MOS1_NE : cls == CLS_MOS0 ? MOS0_NE : cls == CLS_DIODE ? DIODE_NE : 6;
}

}  // namespace s21
