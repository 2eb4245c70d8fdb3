"""One large circuit instance on the MI355X engine (Python view of the ``s21_large_*`` C ABI).

For a netlist of tens of thousands of unknowns -- a power grid with its loads -- whose matrix does not fit
one batched instance's LDS region: the same Newton control, device models and fixed-step transient as the
batched engine, with the matrix in HBM and a multifrontal LU over a fill-reducing order.
"""
import ctypes as C

import numpy as np

from . import _abi as A
from ._lib import check, lib
from .batch import FP64, MIXED, make_tran_options  # noqa: F401  (precision constants shared with Batch)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class Large:
    def __init__(self, cnet, opts=None, precision=FP64):
        self.cnet = cnet
        self.h = C.c_void_p()
        check(lib().s21_large_create(C.pointer(cnet.net), C.byref(opts) if opts is not None else None,
                                     int(precision), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().s21_large_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_reuse(self, on):
        """Factor reuse across Newton iterations (default on in mixed precision, off in FP64)."""
        check(lib().s21_large_set_reuse(self.h, 1 if on else 0))
        return self

    def op(self):
        check(lib().s21_large_op(self.h))
        return self

    def tran(self, tstep, tstop, ics=()):
        t = make_tran_options(tstep, tstop, ics, self.cnet.var_names)
        check(lib().s21_large_tran(self.h, C.byref(t)))
        return self

    def tran_trap(self, tstep, tstop, ics=(), lte=True, reltol=1e-3, abstol=1e-6, trtol=7.0, dt_min=0.0):
        """Trapezoidal transient: fixed steps of tstep (lte=False) or steps chosen by the local truncation
        error with tstep as the largest, interpolated onto the output grid t = k tstep."""
        t = make_tran_options(tstep, tstop, ics, self.cnet.var_names)
        o = A.s21_trap_options(1 if lte else 0, reltol, abstol, trtol, dt_min)
        check(lib().s21_large_tran_trap(self.h, C.byref(t), C.byref(o)))
        return self

    def trap_stats(self):
        """(accepted, rejected) steps of the last trapezoidal transient."""
        a, r = C.c_int64(), C.c_int64()
        check(lib().s21_large_trap_stats(self.h, C.byref(a), C.byref(r)))
        return a.value, r.value

    def n_vars(self):
        return lib().s21_large_n_vars(self.h)

    def var_names(self):
        return [lib().s21_large_var_name(self.h, k).decode() for k in range(self.n_vars())]

    def x(self):
        out = np.zeros(self.n_vars())
        check(lib().s21_large_get_x(self.h, _dp(out)))
        return out

    def status(self):
        st, it = C.c_int32(), C.c_int64()
        check(lib().s21_large_get_status(self.h, C.byref(st), C.byref(it)))
        return st.value, it.value

    def axis(self):
        n = lib().s21_large_n_points(self.h)
        out = np.zeros(n)
        check(lib().s21_large_get_axis(self.h, _dp(out)))
        return out

    def wave(self, variables=None):
        """Transient waveforms [points, len(variables)] (all unknowns when None)."""
        sel = np.arange(self.n_vars(), dtype=np.int32) if variables is None else np.asarray(variables, dtype=np.int32)
        n = lib().s21_large_n_points(self.h)
        out = np.zeros((n, len(sel)))
        check(lib().s21_large_get_wave(self.h, len(sel), _i32p(sel), _dp(out)))
        return out

    def structure(self):
        v = [C.c_int64() for _ in range(5)]
        fl = C.c_double()
        check(lib().s21_large_structure(self.h, *[C.byref(x) for x in v], C.byref(fl)))
        return dict(nnz_a=v[0].value, nnz_lu=v[1].value, fronts=v[2].value, levels=v[3].value,
                    max_front=v[4].value, flops_lu=fl.value)

    def timing(self):
        ms, nf = C.c_double(), C.c_int64()
        check(lib().s21_large_timing(self.h, C.byref(ms), C.byref(nf)))
        return ms.value, nf.value

    def time_factor(self, reps=5):
        """(factorisation ms, sweeps ms) of the last assembled matrix, averaged over reps (HIP events)."""
        f, s = C.c_double(), C.c_double()
        check(lib().s21_large_time_factor(self.h, int(reps), C.byref(f), C.byref(s)))
        return f.value, s.value


def host_check(n, rows, cols, vals, b, relax=0.1):
    """CPU reference of the large path's factorisation (no GPU): x solving A x = b, nnz(L+U), flops."""
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(n)
    nnz, fl = C.c_int64(), C.c_double()
    check(lib().s21_large_host_check(int(n), len(rows), _i32p(rows), _i32p(cols), _dp(vals), _dp(b), float(relax),
                                     _dp(x), C.byref(nnz), C.byref(fl)))
    return x, nnz.value, fl.value
