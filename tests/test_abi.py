"""The engine's C ABI: the library loads on a CPU-only host and exports every declared entry point."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import oracle_py as O
from spice21_amd import _abi as A
from spice21_amd._lib import LIB_PATH, lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "spice21_mi355x.h")


def declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"S21_API[^;(]*?\b(s21_\w+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    names = declared()
    assert len(names) >= 30
    L = C.CDLL(LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_abi_version_and_defaults():
    assert lib().s21_abi_version() == 2
    o = A.s21_options()
    lib().s21_default_options(C.byref(o))
    assert (o.temp, o.tnom, o.gmin, o.iabstol, o.reltol) == (300.15, 300.15, 1e-12, 1e-12, 1e-3)


@pytest.mark.parametrize("tstep,tstop", [(1e-11, 1e-8), (1e-15, 1e-12), (1e-12, 10e-9), (1e-12, 5e-9), (0.3, 1.0)])
def test_tran_point_count(tstep, tstop):
    assert lib().s21_tran_point_count(tstep, tstop) == O.tran_point_count(tstep, tstop)


@pytest.mark.parametrize("f0,f1,n", [(1, 1_000_000, 100), (1000, 100_000_000_000, 1000), (10, 10, 5)])
def test_ac_point_count(f0, f1, n):
    from spice21_amd.netlist import Netlist

    nl = Netlist()
    nl.r(1e-3, "a", "")
    nl.v("v", 0.0, "a", "", acm=1.0)
    f, _, _ = O.ac(nl.to_c(), f0, f1, n)
    assert lib().s21_ac_point_count(C.byref(A.s21_ac_options(f0, f1, n))) == len(f)


def test_tuning_knobs_reject_bad_arguments():
    from spice21_amd._lib import lib
    assert lib().s21_batch_set_waves(None, 2) == A.S21_INVALID
    assert lib().s21_batch_set_hub(None, 1) == A.S21_INVALID
