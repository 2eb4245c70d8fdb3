"""Trapezoidal transients (a mode the reference has disabled, analysis.rs:429-435) on the large-instance
path: the fixed-step companion against the closed-form trapezoidal RC recurrence, its second order against
Backward Euler on the exact RC response, and the LTE-controlled step against a fine fixed-step run."""
import numpy as np
import pytest

from spice21_amd import _abi as A
from spice21_amd import workloads as W
from spice21_amd.large import Large
from spice21_amd.netlist import Netlist

pytestmark = pytest.mark.gpu
GIC = 1e-9  # the IC forcing conductance left after the operating point (analysis.rs:541-549)


def log(*a):
    print(*a, flush=True)


def rc(G=1e-3, C=1e-9):
    n = Netlist()
    n.v("v1", 1.0, "inp", "")
    n.r(G, "inp", "out")
    n.c(C, "out", "")
    return n


def trap_recurrence(G, C, dt, npts):
    """v' C = G (1 - v) - gic v from the operating point with the IC forcing (1 S to 0 V): a first Backward
    Euler step, C (v1 - v0)/dt = i1, then trapezoidal steps, 2C (v1 - v0)/dt - i0 = i1, with the capacitor
    current i = G - (G + gic) v."""
    v = G / (G + 1.0)
    out = [v]
    v = (C / dt * v + G) / (C / dt + G + GIC)
    i = G - (G + GIC) * v
    out.append(v)
    for _ in range(npts - 2):
        v = (2 * C / dt * v + i + G) / (2 * C / dt + G + GIC)
        i = G - (G + GIC) * v
        out.append(v)
    return np.array(out)


def test_fixed_step_trapezoidal_rc_recurrence():
    G, C, dt = 1e-3, 1e-9, 50e-9
    cn = rc(G, C).to_c()
    g = Large(cn).tran_trap(dt, 3e-6, [("out", 0.0)], lte=False)
    assert g.status()[0] == A.S21_OK
    out = g.wave()[:, cn.var_names.index("out")]
    ref = trap_recurrence(G, C, dt, len(out))
    log("trap rc max |dv|", np.abs(out - ref).max())
    assert np.abs(out - ref).max() <= 1e-9  # Newton stops at |residual| <= 1e-12 A against 2C/dt = 0.04 S
    assert out[-1] > 0.9


def test_trapezoidal_is_second_order_and_backward_euler_first():
    """Error at t = 2 us of the RC step response against v = 1 - e^{-t/tau} (tau = 1 us; the operating
    point starts at G/(G+1) instead of 0 and gic shifts the final value, both folded into the exact
    solution): halving dt divides the trapezoidal error by about 4 and Backward Euler's by about 2."""
    G, C = 1e-3, 1e-9
    cn = rc(G, C).to_c()
    k = G + GIC
    vinf, v0 = G / k, G / (G + 1.0)
    tau = C / k

    def err(dt, trap):
        g = Large(cn)
        if trap:
            g.tran_trap(dt, 2e-6, [("out", 0.0)], lte=False)
        else:
            g.tran(dt, 2e-6, [("out", 0.0)])
        assert g.status()[0] == A.S21_OK
        t = g.axis()
        out = g.wave()[:, cn.var_names.index("out")]
        exact = vinf + (v0 - vinf) * np.exp(-t / tau)
        return np.abs(out - exact)[len(t) // 2:].max()

    e_trap = [err(dt, True) for dt in (40e-9, 20e-9)]
    e_be = [err(dt, False) for dt in (40e-9, 20e-9)]
    log("trap errors", e_trap, "be errors", e_be)
    assert 3.0 < e_trap[0] / e_trap[1] < 5.0
    assert 1.6 < e_be[0] / e_be[1] < 2.4
    assert e_trap[1] < e_be[1] / 10


def test_lte_controlled_ring_against_fine_fixed_steps():
    """The 5-stage ring (0-0.5 ns, 1 ps largest step) with the step chosen by the local truncation error
    agrees with a fixed-step trapezoidal run at 0.02 ps to within the LTE tolerance's scale, and rejects
    and shrinks steps around the transitions."""
    cn = W.ring5().to_c()
    ics = [("1", 0.5)]
    ref = Large(cn).tran_trap(0.02e-12, 0.5e-9, ics, lte=False)
    assert ref.status()[0] == A.S21_OK
    tr = ref.axis()
    Xr = ref.wave()
    g = Large(cn).tran_trap(1e-12, 0.5e-9, ics, lte=True)
    st, it = g.status()
    acc, rej = g.trap_stats()
    log("lte ring: status", st, "iterations", it, "accepted", acc, "rejected", rej)
    assert st == A.S21_OK
    t = g.axis()
    X = g.wave()
    v = [k for k, kind in enumerate(list(W.ring5().kinds) + [0, 1]) if kind == 0]
    Xi = np.stack([np.interp(t, tr, Xr[:, k]) for k in v], axis=1)
    err = np.abs(X[:, v] - Xi).max()
    log("  max |v - fine reference|", err)
    assert err <= 2e-2
    assert acc >= len(t) - 1 and rej >= 0
