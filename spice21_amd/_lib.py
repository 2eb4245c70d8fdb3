"""Loader of the engine library (``libspice21_mi355x.so``, built by ``spice21_amd.build``).

There is no fallback: every analysis runs in the HIP library, and importing the engine without the built
library raises.
"""
import ctypes as C
import os

from . import _abi as A

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libspice21_mi355x.so")
_lib = None


class S21Error(RuntimeError):
    """An engine error; ``desc`` is the reference's SpError text, ``code`` an S21_* status."""

    def __init__(self, code, desc):
        super().__init__(desc)
        self.code = code
        self.desc = desc


def _sig(L, name, res, *args):
    f = getattr(L, name)
    f.restype = res
    f.argtypes = list(args)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP engine first (python -m spice21_amd.build or __graft_entry__.build())"
        )
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    i32, i64, d, vp = C.c_int32, C.c_int64, C.c_double, C.c_void_p
    _sig(L, "s21_abi_version", i32)
    _sig(L, "s21_last_error", C.c_size_t, C.c_char_p, C.c_size_t)
    _sig(L, "s21_default_options", None, P(A.s21_options))
    _sig(L, "s21_tran_point_count", i64, d, d)
    _sig(L, "s21_ac_point_count", i64, P(A.s21_ac_options))
    _sig(L, "s21_symbolic", C.c_int, P(A.s21_netlist), P(A.s21_options), P(A.s21_tran_options), P(i32), P(i32),
         P(i64), P(i64), P(i64), P(i64))
    _sig(L, "s21_batch_create", C.c_int, i32, P(A.s21_group), P(A.s21_options), i32, P(vp))
    _sig(L, "s21_batch_free", None, vp)
    _sig(L, "s21_batch_op", C.c_int, vp)
    _sig(L, "s21_batch_op_noalloc", C.c_int, vp, P(d), P(i32), P(i64))
    _sig(L, "s21_batch_write_ac_results", C.c_int, vp, i32, i32, C.c_char_p)
    _sig(L, "s21_batch_tran", C.c_int, vp, P(A.s21_tran_options))
    _sig(L, "s21_batch_ac", C.c_int, vp, P(A.s21_ac_options))
    _sig(L, "s21_batch_n_groups", i32, vp)
    _sig(L, "s21_batch_n_inst", i32, vp, i32)
    _sig(L, "s21_batch_n_vars", i32, vp, i32)
    _sig(L, "s21_batch_var_name", C.c_char_p, vp, i32, i32)
    _sig(L, "s21_batch_get_x", C.c_int, vp, i32, P(d))
    _sig(L, "s21_batch_get_status", C.c_int, vp, i32, P(i32), P(i64))
    _sig(L, "s21_batch_n_points", i64, vp)
    _sig(L, "s21_batch_get_axis", C.c_int, vp, P(d))
    _sig(L, "s21_batch_get_wave", C.c_int, vp, i32, i32, P(d))
    _sig(L, "s21_batch_timing", C.c_int, vp, P(d), P(i64), P(d))
    _sig(L, "s21_batch_resumptions", C.c_int, vp, P(i64))
    _sig(L, "s21_batch_set_waves", C.c_int, vp, i32)
    _sig(L, "s21_batch_set_hub", C.c_int, vp, i32)
    _sig(L, "s21_batch_pivots", C.c_int, vp, i32, P(i32), P(i32))
    _sig(L, "s21_batch_set_profiling", C.c_int, vp, i32)
    _sig(L, "s21_batch_set_ipw", C.c_int, vp, i32)
    _sig(L, "s21_batch_set_jit", C.c_int, vp, i32)
    _sig(L, "s21_jit_build", C.c_int, C.POINTER(A.s21_netlist), C.POINTER(A.s21_options),
         C.POINTER(A.s21_tran_options), i32, C.POINTER(C.c_int64), C.POINTER(C.c_int64))
    _sig(L, "s21_jit_build_ac", C.c_int, C.POINTER(A.s21_netlist), C.POINTER(A.s21_options), d, i32,
         C.POINTER(C.c_int64), C.POINTER(C.c_int64))
    _sig(L, "s21_batch_config", C.c_int, vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32))
    _sig(L, "s21_batch_set_ac_runs", C.c_int, vp, i32)
    _sig(L, "s21_batch_last_ac_runs", i32, vp)
    _sig(L, "s21_batch_profile", C.c_int, vp, i32, P(i64))
    _sig(L, "s21_batch_structure", C.c_int, vp, i32, P(i64), P(i64), P(i64), P(i64))
    _sig(L, "s21_batch_assemble", C.c_int, vp, i32, i32, P(d), P(i32), P(i32), P(d), P(d), i64, P(i64))
    # single large instance
    _sig(L, "s21_large_create", C.c_int, P(A.s21_netlist), P(A.s21_options), i32, P(vp))
    _sig(L, "s21_large_free", None, vp)
    _sig(L, "s21_large_op", C.c_int, vp)
    _sig(L, "s21_large_tran", C.c_int, vp, P(A.s21_tran_options))
    _sig(L, "s21_large_n_vars", i32, vp)
    _sig(L, "s21_large_var_name", C.c_char_p, vp, i32)
    _sig(L, "s21_large_get_x", C.c_int, vp, P(d))
    _sig(L, "s21_large_get_status", C.c_int, vp, P(i32), P(i64))
    _sig(L, "s21_large_n_points", i64, vp)
    _sig(L, "s21_large_get_axis", C.c_int, vp, P(d))
    _sig(L, "s21_large_get_wave", C.c_int, vp, i32, P(i32), P(d))
    _sig(L, "s21_large_structure", C.c_int, vp, P(i64), P(i64), P(i64), P(i64), P(i64), P(d))
    _sig(L, "s21_large_timing", C.c_int, vp, P(d), P(i64))
    _sig(L, "s21_large_time_factor", C.c_int, vp, i32, P(d), P(d))
    _sig(L, "s21_large_set_reuse", C.c_int, vp, i32)
    _sig(L, "s21_large_tran_trap", C.c_int, vp, P(A.s21_tran_options), P(A.s21_trap_options))
    _sig(L, "s21_large_trap_stats", C.c_int, vp, P(i64), P(i64))
    _sig(L, "s21_large_host_check", C.c_int, i32, i64, P(i32), P(i32), P(d), P(d), d, P(d), P(i64), P(d))
    _lib = L
    return L


def last_error():
    buf = C.create_string_buffer(4096)
    lib().s21_last_error(buf, 4096)
    return buf.value.decode(errors="replace")


def check(rc):
    if rc != 0:
        raise S21Error(rc, last_error())
    return rc
