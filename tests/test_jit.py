"""The generated factorisation compiles for gfx950 (host only: symbolic phase, code generation, hiprtc)."""
import ctypes as C

import pytest

from spice21_amd import _abi as A
from spice21_amd import workloads as W
from spice21_amd._lib import check, lib
from spice21_amd.batch import FP64, MIXED, make_tran_options


def _build(cn, precision, tran=None):
    ops, code = C.c_int64(), C.c_int64()
    check(lib().s21_jit_build(cn.ptr, None, C.byref(tran) if tran is not None else None, precision, C.byref(ops),
                              C.byref(code)))
    return ops.value, code.value


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_chain_factorisation_compiles(precision):
    nl, _ = W.chain_mc(1)
    ops, code = _build(nl.to_c(), precision)
    assert ops > 300 and code > 1000


def test_ring_transient_factorisation_compiles():
    cn = W.ring5().to_c()
    tr = make_tran_options(1e-12, 1e-11, [("1", 0.5)], cn.var_names)
    ops, code = _build(cn, FP64, tr)
    assert ops > 0 and code > 1000


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_ac_complex_factorisation_compiles(precision):
    """The AC kernel: the amplifier-and-ladder pattern's complex factorisation written out (lane per
    instance) with the AC Newton around it."""
    nl, _ = W.amp_ladder(1)
    ops, code = C.c_int64(), C.c_int64()
    check(lib().s21_jit_build_ac(nl.to_c().ptr, None, 1000.0, precision, C.byref(ops), C.byref(code)))
    assert ops.value > 26 and code.value > 1000
