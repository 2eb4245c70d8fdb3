"""The oracle's charge paths against what the mathematics fixes: Mos1 Meyer gate charge in a transient,
Mos1 small-signal capacitances in AC, and diode junction charge in a transient.

No reference fixture covers these paths (the reference's golden waveforms use Mos1 without tox and no
diodes).  Each test builds a circuit whose behaviour reduces to a scalar Backward-Euler recurrence or a
small linear phasor system, solves that independently (root finding per time step, a dense complex
solve per frequency) from the model equations as the textbook gives them -- SPICE2's Meyer half
capacitances, the level-1 square law, Shockley's diode law -- plus the reference's own quirks where they
change the answer, and compares the oracle's waveform.  Each check also runs with a planted mistake in
the oracle (``oracle_py.planted``) and must fail there.
"""
import math

import numpy as np
import pytest
from scipy.optimize import brentq

from oracle import oracle_py as O
from spice21_amd.netlist import Netlist

VT = 1.3806226e-23 / 1.6021918e-19 * 300.15  # thermal voltage at the default temperature (comps/mod.rs:24-37)
EPS_OX = 3.9 * 8.854214871e-12  # SiO2 permittivity (F/m) as in Mos1Model::resolve (comps/mos.rs)
GIC = 1e-9  # IC-forcing conductance left in place for the whole transient (analysis.rs:541-549)


def root(f, lo, hi):
    return brentq(f, lo, hi, xtol=1e-18, rtol=1e-15, maxiter=500)


# ---- SPICE2 Meyer model: half capacitances (each point adds its own half and the previous point's) ----
def meyer_half(vov, vds, cox, phi):
    """(cgs, cgd, cgb) halves of the Meyer gate capacitances at overdrive vov and vds >= 0."""
    if vov <= -phi:  # accumulation: the whole oxide to the bulk
        return 0.0, 0.0, cox / 2
    if vov <= -phi / 2:  # depletion: the bulk share falls linearly to half
        return 0.0, 0.0, -vov * cox / (2 * phi)
    if vov <= 0.0:  # weak inversion: the channel builds up toward the source
        return vov * cox / (1.5 * phi) + cox / 3, 0.0, -vov * cox / (2 * phi)
    if vov <= vds:  # saturation: 2/3 of the oxide to the source
        return cox / 3, 0.0, 0.0
    d = 2 * vov - vds  # triode
    return cox / 3 * (1 - (vov - vds) ** 2 / d ** 2), cox / 3 * (1 - vov ** 2 / d ** 2), 0.0


VD = 0.3  # drain bias of the gate-charge circuit


def gate_rc(vin, g, w, l, tox, cov):
    """An NMOS with source and bulk grounded and its drain held at VD by a source, its gate charged from
    `vin` through conductance `g` from the initial condition 0 V.  The channel current flows between the
    two sources; the gate node sees only gate charge.  (With vds > 0 the Meyer capacitances are continuous
    across the regions; at vds = 0 they jump at threshold and the reference's Newton loop, whose Jacobian
    holds the capacitances as constants, does not converge there.)"""
    n = Netlist()
    n.v("vin", vin, "in", "")
    n.r(g, "in", "g")
    n.v("vd", VD, "d", "")
    nm = n.mos1_model(vt0=0.4, kp=100e-6, tox=tox, cgso=cov, cgdo=cov, cgbo=cov)
    ip = n.mos1_inst(w=w, l=l)
    n.mos1("m", "d", "g", "", "", nm, ip)
    return n


def gate_rc_step(vin, g, w, l, tox, cov, dt):
    """KCL residual of the gate node at one Backward-Euler step from u to v (amperes).  The reference
    integrates the gate-source charge and then overwrites that result with the gate-drain integration
    (comps/mos.rs:787-802), whose companion it stamps between gate and source (the tr.gs entries of
    mos.rs:838-868): g = Cgd/dt, rhs = i - g vgd.  The gate's current through that companion is
    g (vg - vs) + i - g vgd = i + g vds, so the gate sees the gate-drain and gate-bulk charge steps
    i = Cgd (v - u)/dt, Cgb (v - u)/dt plus a current Cgd vds / dt that stays after the charge settles.
    Cgd(v, u) = cgd(v) + cgd_ov + cgd(u), Cgb(v, u) = cgb(v) + cgb_ov + cgb(u) (each point's Meyer half
    plus the previous point's); the gate-source capacitance drops out."""
    cox = EPS_OX / tox * l * w
    cgd_ov, cgb_ov = w * cov, l * cov
    phi, vt0 = 0.6, 0.4

    def caps(v):
        _, cgd, cgb = meyer_half(v - vt0, VD, cox, phi)
        return cgd, cgb

    def f(v, u):
        cv, cu = caps(v), caps(u)
        cgd, cgb = cv[0] + cgd_ov + cu[0], cv[1] + cgb_ov + cu[1]
        return g * (vin - v) - GIC * v - (cgd + cgb) * (v - u) / dt - cgd * VD / dt

    return f


def gate_rc_expected(vin, g, w, l, tox, cov, dt, npts):
    """The gate waveform: the operating point (the IC forces the gate through 1 S to 0 V), then each
    step's root of gate_rc_step."""
    f = gate_rc_step(vin, g, w, l, tox, cov, dt)
    u = g * vin / (g + 1.0)
    out = [u]
    for _ in range(npts - 1):
        u = root(lambda v, u=u: f(v, u), -2.0, 2.0)
        out.append(u)
    return np.array(out)


def step_residuals(f, v):
    return np.abs([f(v[k], v[k - 1]) for k in range(1, len(v))])


@pytest.mark.parametrize("vin", [1.0, -1.0])
def test_meyer_gate_charge_transient_closed_form(vin):
    """vin = +1 V takes the gate through depletion and weak inversion to threshold, where the Cgd vds / dt
    current holds it; vin = -1 V through depletion into accumulation (cgb = cox/2), past -1 V.  Every
    step of the oracle's waveform satisfies the gate's KCL to 1e-12 A (the reference's own convergence
    bound), and the waveform equals the recurrence's to 3e-8 V."""
    args = dict(vin=vin, g=1e-4, w=1e-6, l=1e-6, tox=1e-8, cov=2e-10)
    dt, tstop = 1e-12, 4e-10
    cn = gate_rc(**args).to_c()
    t, X, names, _ = O.tran(cn, dt, tstop, [("g", 0.0)])
    vg = X[:, names.index("g")]
    ref = gate_rc_expected(**args, dt=dt, npts=len(t))
    f = gate_rc_step(**args, dt=dt)
    assert np.max(np.abs(np.diff(vg[-20:]))) < 1e-5 * abs(vin)  # the run covers the whole charge
    # every step solves the gate's KCL to the reference's current tolerance (iabstol, analysis.rs:331-345)
    assert vg[0] == pytest.approx(ref[0], abs=1e-15)
    assert np.max(step_residuals(f, vg)) <= 1e-12
    assert np.max(np.abs(vg - ref)) <= 3e-8, np.max(np.abs(vg - ref))
    # planted mistakes: the gate-source charge kept (no overwrite), the depletion bulk share doubled
    for k in (1, 4):
        with O.planted(k):
            _, Xp, _, _ = O.tran(cn, dt, tstop, [("g", 0.0)])
        assert np.max(step_residuals(f, Xp[:, names.index("g")])) > 1e-8, k
        assert np.max(np.abs(Xp[:, names.index("g")] - ref)) > 1e-4, k


# ---- AC: small-signal gate capacitances of load_ac (comps/mos.rs:914-968) ----
def test_ac_gate_rc_in_saturation_closed_form():
    """Gate biased at 0.6 V through 10 kohm, drain on a 1 V source (AC ground): saturation, where the
    reference's AC admittance holds only the committed intrinsic Meyer half cgs = cox/3 (no overlap, which
    load_ac leaves out): v_g / v_in = 1 / (1 + j w R cox/3).  The reference's AC Newton keeps the previous
    frequency's solution while its residual stays below 1e-9 A (analysis.rs:253-304), so the check is
    that bound on the gate's KCL residual, which also bounds |v_g - H| by 1e-9 / |G + j w C|."""
    w, l, tox, cov, g = 10e-6, 1e-6, 1e-8, 2e-10, 1e-4
    n = Netlist()
    n.v("vin", 0.6, "in", "", acm=1.0)
    n.r(g, "in", "g")
    n.v("vdd", 1.0, "d", "")
    nm = n.mos1_model(vt0=0.4, kp=100e-6, tox=tox, cgso=cov, cgdo=cov, cgbo=cov)
    n.mos1("m", "d", "g", "", "", nm, n.mos1_inst(w=w, l=l))
    cn = n.to_c()
    cox = EPS_OX / tox * l * w
    f, X, _ = O.ac(cn, 1000, 100_000_000_000, 40)
    Y = g + 1j * 2 * np.pi * f * (cox / 3)
    H = g / Y
    vg = X[:, n.names.index("g")]
    res = np.abs(g * (1.0 - vg) - (Y - g) * vg)
    assert np.max(res) < 1e-9
    assert np.all(np.abs(vg - H) * np.abs(Y) <= 1.000001e-9)
    assert abs(H[-1]) < 0.1  # the sweep passes the pole
    with O.planted(5):  # saturation cgs = cox/2
        _, Xp, _ = O.ac(cn, 1000, 100_000_000_000, 40)
    vp = Xp[:, n.names.index("g")]
    assert np.max(np.abs(g * (1.0 - vp) - (Y - g) * vp)) > 1e-7


def test_ac_triode_stage_closed_form():
    """Gate at 1 V through 10 kohm, drain through 10 kohm to 1 V: the stage sits in the triode region, so
    the gate-drain capacitance couples both nodes.  Small-signal equations with the textbook square law
    (gm = beta vds, gds = beta (vov - vds)), the drain junction's gbd and the Meyer triode halves, and
    the reference's (G, dr) admittance stamped twice (comps/mos.rs:951,954):
        g:  Gg (v_in - v_g) = j w (cgs + cgd + cgb) v_g - 2 j w cgd v_d
        d:  (gds + gbd + Gd + j w cgd) v_d + (gm - j w cgd) v_g = 0
    The operating point itself is checked against KCL with the square law."""
    w, l, tox, gg, gd_ = 10e-6, 1e-6, 1e-8, 1e-4, 1e-4
    n = Netlist()
    n.v("vin", 1.0, "in", "", acm=1.0)
    n.r(gg, "in", "g")
    n.v("vdd", 1.0, "vdd", "")
    n.r(gd_, "vdd", "d")
    nm = n.mos1_model(vt0=0.4, kp=100e-6, tox=tox)
    n.mos1("m", "d", "g", "", "", nm, n.mos1_inst(w=w, l=l))
    cn = n.to_c()
    x, _, _ = O.dcop(cn)
    vgs, vds = x[n.names.index("g")], x[n.names.index("d")]
    beta = 100e-6 * w / l
    vov = vgs - 0.4
    assert vgs == pytest.approx(1.0, abs=1e-12) and 0 < vds < vov  # triode
    ids = beta * (vov * vds - vds ** 2 / 2)
    isat_d = 1e-8 * 1e-12  # js * default drain area
    ibd = isat_d * (math.exp(-vds / VT) - 1) - 1e-12 * vds  # drain-bulk junction, reverse-biased, + gmin
    assert abs(gd_ * (1.0 - vds) - ids + ibd) <= 1e-12
    gm, gds = beta * vds, beta * (vov - vds)
    gbd = isat_d / VT * math.exp(-vds / VT) + 1e-12
    cox = EPS_OX / tox * l * w
    cgs, cgd, cgb = meyer_half(vov, vds, cox, 0.6)
    f, X, _ = O.ac(cn, 1000, 100_000_000_000, 40)

    def expected(double_gd):
        out = []
        for fk in f:
            jw = 1j * 2 * np.pi * fk
            M = np.array([[gg + jw * (cgs + cgd + cgb), -(2 if double_gd else 1) * jw * cgd],
                          [gm - jw * cgd, gds + gbd + gd_ + jw * cgd]])
            out.append(np.linalg.solve(M, np.array([gg * 1.0, 0.0])))
        return np.array(out)

    E = expected(True)

    def residual(V):  # KCL of both nodes (the reference's AC acceptance bound is 1e-9 A per entry)
        out = []
        for fk, v in zip(f, V):
            jw = 1j * 2 * np.pi * fk
            out.append([gg * (1.0 - v[0]) - jw * (cgs + cgd + cgb) * v[0] + 2 * jw * cgd * v[1],
                        (gds + gbd + gd_ + jw * cgd) * v[1] + (gm - jw * cgd) * v[0]])
        return np.abs(np.array(out))

    got = np.stack([X[:, n.names.index("g")], X[:, n.names.index("d")]], axis=1)
    assert np.max(residual(got)) < 1e-9
    assert np.max(np.abs(got - E) / np.abs(E)) <= 1e-4
    # the doubled stamp matters: the single-stamp reading differs at high frequency
    E1 = expected(False)
    assert np.max(np.abs(E1 - E) / np.abs(E)) > 1e-2
    with O.planted(2):
        _, Xp, _ = O.ac(cn, 1000, 100_000_000_000, 40)
    gotp = np.stack([Xp[:, n.names.index("g")], Xp[:, n.names.index("d")]], axis=1)
    assert np.max(residual(gotp)) > 1e-7


# ---- diode junction charge in a transient (comps/diode.rs:279-354) ----
def diode_rc(cj0, tt):
    """1 V charging node a through 1 kohm against 1 pF, with a diode from ground (anode) to a (cathode):
    reverse-biased as a rises."""
    n = Netlist()
    n.v("v1", 1.0, "in", "")
    n.r(1e-3, "in", "a")
    dm = n.diode_model(cj0=cj0, tt=tt)
    n.diode("d1", "", "a", dm, n.diode_inst())
    n.c(1e-12, "a", "")
    return n


def diode_rc_expected(cj0, tt, dt, npts, m=0.5, vj=1.0):
    """Node a's Backward-Euler recurrence.  KCL: G (1 - v) + id(vd) + (Q(vd) - Q(ud)) / dt - C (v - u) / dt
    - gic v = 0 with vd = -v, id = is (exp(vd / vt) - 1) + gmin vd (Shockley, is = 1e-14), and the
    reference's reverse-bias charge Q = tt vj cz (1 - a s) / (1 - m), a = 1 - vd / vj, s = -m ln a
    (comps/diode.rs:307-309): it carries the transit time as a factor, so without tt the junction stores
    no charge and its capacitance enters only the Newton Jacobian."""
    G, Cn = 1e-3, 1e-12

    def q(vd):
        a = 1 - vd / vj
        s = -m * math.log(a)
        return tt * vj * cj0 * (1 - a * s) / (1 - m)

    def idc(vd):
        return 1e-14 * (math.exp(vd / VT) - 1) + 1e-12 * vd

    # operating point with the IC forcing node a through 1 S to 0 V
    v0 = root(lambda v: G * (1 - v) + idc(-v) - 1.0 * v, -0.1, 1.1)
    out, u = [v0], v0
    for _ in range(npts - 1):
        def f(v, u=u):
            return G * (1 - v) + idc(-v) + (q(-v) - q(-u)) / dt - Cn * (v - u) / dt - GIC * v

        u = root(f, -0.1, 1.1)
        out.append(u)
    return np.array(out)


@pytest.mark.parametrize("cj0,tt", [(0.0, 0.0), (1e-12, 0.0), (1e-12, 0.3), (2e-13, 0.5)])
def test_diode_junction_charge_transient_closed_form(cj0, tt):
    """The oracle's node voltage equals the recurrence to 1e-9 V.  With tt = 0 the junction capacitance
    changes nothing in the waveform (cj0 = 1e-12 equals cj0 = 0).  The reference's charge scales with tt
    in seconds, so at a physical transit time (1e-9 s) it is below 1e-20 C; transit times of 0.3 s and
    0.5 s make it move the node by 0.2 V and 0.05 V, which pins its form.  (At several pairs in between,
    e.g. cj0 = 2e-12 with tt = 2e-10, the reference's Newton loop, whose Jacobian takes cz s for the
    charge's derivative, does not converge.)"""
    dt, tstop = 1e-11, 3e-9
    cn = diode_rc(cj0, tt).to_c()
    t, X, names, _ = O.tran(cn, dt, tstop, [("a", 0.0)])
    va = X[:, names.index("a")]
    ref = diode_rc_expected(cj0, tt, dt, len(t))
    assert va[-1] > 0.9  # the node charges through several time constants
    assert np.max(np.abs(va - ref)) <= 1e-9, np.max(np.abs(va - ref))
    if cj0 > 0:
        # planted: the depletion charge without the transit-time factor
        with O.planted(3):
            try:
                _, Xp, _, _ = O.tran(cn, dt, tstop, [("a", 0.0)])
                caught = np.max(np.abs(Xp[:, names.index("a")] - ref)) > 1e-4
            except O.OracleError:  # the mistaken charge no longer matches its own Jacobian entry
                caught = True
        assert caught
