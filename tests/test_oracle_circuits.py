"""The oracle against the reference's dcop/tran/ac known answers (spice21/src/tests.rs,
spice21py/spice21py/tests/test_spice21.py) and against closed forms."""
import math

import numpy as np
import pytest

from oracle import oracle_py as O
from spice21_amd import _abi as A
from spice21_amd.netlist import Netlist
from tests.circuits import mos1_defaults


def op(n, **kw):
    x, it, _ = O.dcop(n.to_c(), A.s21_options.default(**kw) if kw else None)
    return {nm: x[k] for k, nm in enumerate(n.names)}, x


def test_dcop1_r_only():  # tests.rs:20-27
    n = Netlist()
    n.r(0.001, "a", "")
    _, x = op(n)
    assert list(x) == [0.0]


def test_dcop2_i_r():  # tests.rs:28-44
    n = Netlist(["vdd"])
    n.i(1e-3, "vdd", "")
    n.r(1e-3, "vdd", "")
    s, _ = op(n)
    assert s["vdd"] == 1.0


def test_dcop4_divider_exact():  # tests.rs:64-83
    n = Netlist(["vdd", "div"])
    n.v("v1", 1.0, "vdd", "")
    n.r(2e-3, "vdd", "div")
    n.r(2e-3, "", "div")
    s, _ = op(n)
    assert s["vdd"] == 1.0 and s["div"] == 0.5 and s["v1"] == -1e-3


def test_dcop5_diode_round_trip():  # tests.rs:84-133
    def dio(src):
        n = Netlist(["p"])
        dm, di = n.diode_model(), n.diode_inst()
        n.diode("dd", "p", "", dm, di)
        src(n)
        return op(n)[0]

    i = abs(dio(lambda n: n.v("vin", 0.70, "p", ""))["vin"])
    assert 1e-3 < i < 100e-3
    s = dio(lambda n: n.i(i, "p", ""))
    assert abs(s["p"] - 0.70) < 1e-3
    # Shockley closed form at T = tnom: i = is (exp(v / vt) - 1) + gmin v
    vt = 1.3806226e-23 / 1.6021918e-19 * 300.15
    assert i == pytest.approx(1e-14 * (math.exp(0.7 / vt) - 1) + 1e-12 * 0.7, rel=1e-9)


def mos0_ckt(comps):
    n = Netlist()
    for c in comps:
        c(n)
    return n


@pytest.mark.parametrize(
    "pol,swap", [(0, False), (0, True), (1, False), (1, True)]
)
def test_dcop8_mos0_diode_connected(pol, swap):  # tests.rs:181-335
    i = 5e-3 if pol == 0 else -5e-3
    n = Netlist()
    n.i(i, "0", "")
    if swap:
        n.mos0("m", "", "0", "0", "", pol)
    else:
        n.mos0("m", "0", "0", "", "", pol)
    s, _ = op(n)
    target = 0.697 if pol == 0 else -0.697
    assert abs(s["0"] - target) < 1e-3


def test_dcop6_mos0_char():  # tests.rs:135-157
    n = Netlist(["g", "d"])
    n.mos0("m", "d", "g", "", "", 0)
    n.v("v1", 1.0, "g", "")
    n.v("v2", 1.0, "d", "")
    s, _ = op(n)
    assert s["g"] == 1.0 and s["d"] == 1.0 and s["v1"] == 0.0
    assert abs(s["v2"] + 14.1e-3) < 1e-4
    # closed form: saturation, beta/2 (vgs - vth)^2 (1 + lam vds) + gmin vds
    ids = 50e-3 / 2 * 0.75**2 * (1 + 3e-3) + 1e-12
    assert -s["v2"] == pytest.approx(ids, rel=1e-12)


def test_dcop12_mos0_inverter_chain():  # tests.rs:556-666
    n = Netlist()
    for k in range(5):
        n.r(1e-9, str(k), "")
    for k in range(4):
        n.mos0(f"p{k+1}", str(k + 1), str(k), "0", "0", 1)
        n.mos0(f"n{k+1}", str(k + 1), str(k), "", "", 0)
    n.v("v1", 1.0, "0", "")
    _, x = op(n)
    assert x[0] == 1.0 and abs(x[1]) < 1e-3 and abs(x[2] - 1) < 1e-3 and abs(x[3]) < 1e-3
    assert abs(x[4] - 1) < 1e-3 and abs(x[5]) < 1e-6


def test_dcop13_rc():  # tests.rs:668-692
    n = Netlist()
    n.r(1e-3, "1", "0")
    n.c(1e-9, "1", "")
    n.v("v1", 1.0, "0", "")
    assert list(op(n)[1]) == [1.0, 1.0, 0.0]
    n = Netlist()
    n.c(1e-9, "i", "o")
    n.r(1e-3, "o", "")
    n.v("v1", 1.0, "i", "")
    assert list(op(n)[1]) == [1.0, 0.0, 0.0]


def test_mos1_op_and_inverter():  # tests.rs:791-866
    n = Netlist()
    nm, pm, ip = mos1_defaults(n)
    n.mos1("m", "0", "0", "", "", nm, ip)
    n.v("v1", 1.0, "0", "")
    _, x = op(n)
    assert x[0] == 1.0 and -1e-3 < x[1] < 0.0
    n = Netlist(["inp", "out", "vdd", "vss"])
    nm, pm, ip = mos1_defaults(n)
    n.mos1("p", "out", "inp", "vdd", "vdd", pm, ip)
    n.mos1("n", "out", "inp", "vss", "vss", nm, ip)
    n.v("v1", 1.0, "vdd", "vss")
    n.v("v2", 0.0, "vss", "")
    n.v("v3", 0.0, "inp", "vss")
    s, _ = op(n)
    assert s["vdd"] == 1.0 and s["vss"] == 0.0 and s["inp"] == 0.0
    assert abs(s["out"] - 1.0) < 1e-3
    for v in ("v1", "v2", "v3"):
        assert abs(s[v]) < 1e-6


def test_mos1_square_law_at_config_params():
    """Mos1 with the configs' parameters against the level-1 closed form at T = tnom:
    Id = kp/2 W/L (Vgs - Vt)^2 in saturation, kp W/L ((Vgs-Vt)Vds - Vds^2/2) in triode."""
    for vg, vd in [(1.0, 1.0), (1.0, 0.2), (0.7, 0.9)]:
        n = Netlist(["g", "d"])
        m = n.mos1_model(vt0=0.4, kp=100e-6)
        ip = n.mos1_inst(w=1e-6, l=0.1e-6)
        n.mos1("m", "d", "g", "", "", m, ip)
        n.v("vg", vg, "g", "")
        n.v("vd", vd, "d", "")
        s, _ = op(n)
        beta = 100e-6 * 10.0
        vov = vg - 0.4
        ids = beta / 2 * vov**2 if vd >= vov else beta * (vov * vd - vd * vd / 2)
        # drain junction current js*area*(exp(-vdb/vt)-1); gmin enters only the linearised conductance
        # and cancels in the converged KCL (mos.rs:720-722)
        vt = 1.3806226e-23 / 1.6021918e-19 * 300.15
        ibd = 1e-8 * 1e-12 * (math.exp(-vd / vt) - 1)
        assert -s["vd"] == pytest.approx(ids - ibd, rel=1e-9)


def test_mos1_jacobian_matches_finite_differences():
    """The assembled OP Jacobian equals d(A(x) x - rhs(x))/dx by central differences (a sign or index
    mistake in any of the 22 stamps fails this).  gamma = 0: the reference's Norton current
    irhs = ids - gm vgs - gds vds omits gmbs vbs (mos.rs:829), so with body effect A(x)x - rhs(x) is
    ids - gmbs vsb and its derivative is not the stamped Jacobian."""
    n = Netlist(["in", "out", "vdd", "mid"])
    nm = n.mos1_model(vt0=0.4, kp=100e-6, lambda_=0.05)
    pm = n.mos1_model(mos_type=1, vt0=-0.4, kp=50e-6, lambda_=0.05)
    ip = n.mos1_inst(w=1e-6, l=0.1e-6)
    n.mos1("p", "out", "in", "vdd", "vdd", pm, ip)
    n.mos1("n", "out", "in", "mid", "", nm, ip)
    n.r(1e-4, "mid", "")
    cn = n.to_c()
    rng = np.random.default_rng(7)
    for trial in range(5):
        # bias both devices un-reversed with vsb >= 0: with the source/drain reversed the reference
        # drops the body effect from the current but still stamps gmbs (mos.rs:670-674, 696-700)
        x = np.zeros(n.n_vars)
        x[2] = rng.uniform(0.8, 1.0)  # vdd
        x[3] = rng.uniform(0.01, 0.2)  # mid
        x[0] = rng.uniform(0.2, 0.8)  # in
        x[1] = rng.uniform(x[3] + 0.05, x[2] - 0.05)  # out
        rows, cols, vals, rhs = O.assemble_op(cn, x)
        J = np.zeros((n.n_vars, n.n_vars))
        for r, c, v in zip(rows, cols, vals):
            J[r, c] += v

        def kcl(xx):  # f(x) = A(x) x - rhs(x) = device currents leaving each node
            r2, c2, v2, b2 = O.assemble_op(cn, xx)
            A2 = np.zeros_like(J)
            for r, c, v in zip(r2, c2, v2):
                A2[r, c] += v
            return A2 @ xx - b2

        h = 1e-7
        for k in range(n.n_vars):
            e = np.zeros(n.n_vars)
            e[k] = h
            fd = (kcl(x + e) - kcl(x - e)) / (2 * h)
            assert np.allclose(J[:, k], fd, rtol=1e-4, atol=1e-9), (trial, k, J[:, k], fd)


def test_tran1_rc_exact_recurrence():  # tests.rs:693-717
    n = Netlist()
    n.v("v1", 1.0, "inp", "")
    n.r(1e-3, "inp", "out")
    n.c(1e-9, "out", "")
    t, X, names, _ = O.tran(n.to_c(), 10e-9, 10e-6, [("out", 0.0)])
    out = X[:, names.index("out")]
    assert np.all(X[:, names.index("inp")] == 1.0)
    assert abs(out[0]) < 1e-3 and abs(out[-1] - 1.0) < 1e-3 and np.all(np.diff(out) > 0)
    # Backward Euler closed form: v[k+1] = (v[k] + a) / (1 + a + b), a = G dt / C, b = gic dt / C
    G, Cc, dt = 1e-3, 1e-9, 10e-9
    a, b = G * dt / Cc, 1e-9 * dt / Cc
    v = out[0]
    for k in range(1, len(out)):
        v = (v + a) / (1 + a + b)
        assert out[k] == pytest.approx(v, abs=1e-9)


def test_ac1_rc_lowpass_closed_form():  # spice21py/spice21py/tests/test_spice21.py:53-70
    n = Netlist()
    n.r(1e-3, "inp", "out")
    n.c(1e-9, "out", "")
    n.v("v", 1e-3, "inp", "", acm=1.0)
    f, X, _ = O.ac(n.to_c(), 1, 1_000_000, 100)
    inp, out = X[:, 0], X[:, 1]
    assert np.all(inp == 1.0 + 0.0j)
    assert np.all(np.abs(out[1:]) <= np.abs(out[:-1]))
    H = 1.0 / (1.0 + 1j * 2 * np.pi * f * 1e-9 / 1e-3)
    assert np.max(np.abs(out - H)) < 1e-6
    assert f[0] == 1.0 and f[-1] == 1e6 and len(f) == 101


def test_mos1_derive_closed_form_at_tnom():
    """At T = tnom the temperature scaling is the identity: vt0_t = vt0, beta = kp W/Leff,
    isat = js * area, phi_t = phi (mos.rs:320-476)."""
    m = A.s21_params.make(A.MOS1_MODEL_FIELDS, vt0=0.4, kp=100e-6, ld=0.01e-6)
    i = A.s21_params.make(A.MOS1_INST_FIELDS, w=1e-6, l=0.1e-6, a_d=2e-12)
    d = O.mos1_derive(m, i)
    assert d[0] == pytest.approx(0.4, rel=1e-12)
    assert d[3] == pytest.approx(100e-6 * 1e-6 / 0.08e-6, rel=1e-12)
    assert d[2] == pytest.approx(0.6, rel=1e-12)
    assert d[12] == pytest.approx(1e-8 * 2e-12, rel=1e-12)
    assert d[20] == pytest.approx(1e-8 * 1e-12, rel=1e-12)
    # a different temperature moves vt0 and beta
    d2 = O.mos1_derive(m, i, A.s21_options.default(temp=350.0))
    assert d2[0] != d[0] and d2[3] < d[3]


def test_mos1_reversed_body_effect_quirk():
    """Reversed NMOS (vsb < 0 after the S/D swap): the current has no body effect but gmbs is stamped."""
    n = Netlist(["g", "d", "s"])
    nm = n.mos1_model(vt0=0.4, kp=100e-6, gamma=0.5)
    ip = n.mos1_inst(w=1e-6, l=0.1e-6)
    n.mos1("m", "d", "g", "s", "", nm, ip)
    x = np.array([1.0, 0.0, 0.3])  # drain below source: reversed, vsb = 0 - 0 -> von = vt0_t
    rows, cols, vals, rhs = O.assemble_op(n.to_c(), x)
    J = np.zeros((3, 3))
    for r, c, v in zip(rows, cols, vals):
        J[r, c] += v
    # reversed: the source-side row is "d"; its bulk column is ground, so gmbs shows up only through
    # the (sr, sr) entry: gm + gds + gbs + gmbs with gmbs = gm * gamma / 2 / sqrt(phi)
    vgs = 1.0 - 0.0
    vds = 0.3
    beta = 1e-3
    vov = vgs - 0.4
    gm = beta * vds
    gds = beta * (vov - vds)
    gmbs = gm * 0.5 / 2 / np.sqrt(0.6)
    vt = 1.3806226e-23 / 1.6021918e-19 * 300.15
    gbs = 1e-20 / vt + 1e-12
    assert J[1, 1] == pytest.approx(gm + gds + gbs + gmbs, rel=1e-9)


def test_ac_common_source_gain_closed_form():
    """A level-1 NMOS common-source stage without capacitances (no tox, no junction capacitance): its AC
    response is the small-signal gain at every frequency, -gm / (gds + 1/RD + gbd), with gm and gds of the
    square law at the oracle's own operating point and gbd the drain junction's exp term plus gmin.  Pins
    the Mos1 AC stamps (gm from gate to drain, gds and gbd on the drain) and their signs."""
    for lam, rd in [(0.05, 5e3), (0.0, 2e3), (0.1, 8e3)]:
        n = Netlist()
        nm = n.mos1_model(vt0=0.4, kp=100e-6, lambda_=lam)
        ip = n.mos1_inst(w=1e-6, l=0.1e-6)
        n.v("vin", 0.6, "g", "", acm=1.0)
        n.v("vdd", 1.0, "vdd", "")
        n.r(1.0 / rd, "vdd", "d")
        n.mos1("m", "d", "g", "", "", nm, ip)
        s, _ = op(n)
        vds = s["d"]
        beta, vov = 100e-6 * 10.0, 0.2
        assert vds >= vov  # saturation
        gm = beta * vov * (1.0 + lam * vds)
        gds = lam * beta / 2.0 * vov * vov
        vt = 1.3806226e-23 / 1.6021918e-19 * 300.15
        gbd = 1e-8 * 1e-12 / vt * math.exp(-vds / vt) + 1e-12
        gain = -gm / (gds + 1.0 / rd + gbd)
        f, X, _ = O.ac(n.to_c(), 1, 1_000_000_000, 9)
        out = X[:, n.names.index("d")]
        assert np.max(np.abs(out - gain)) <= 1e-9 * abs(gain), (lam, rd, out[0], gain)
        assert np.all(X[:, n.names.index("g")] == 1.0 + 0.0j)


def test_switching_diode_repivots_and_follows_the_off_state_recurrence():
    """The switching-diode circuit with gmin = 0: the reference's pivot searches choose other elements as the
    diode turns off (pivot_stats counts factorisations whose search differs from the previous one's), and
    once the diode's conductance has underflowed to zero (a above 19.3 V) the node follows the Backward
    Euler recurrence of the RC with the reverse saturation current and the IC forcing conductance left
    after the operating point: C (a1 - a0)/dt = I - isat - (G + gic) a1."""
    from tests.circuits import no_gmin, switching_diode

    I, G, C, isat, dt, gic = 3.0, 0.1, 1e-7, 1e-3, 1e-8, 1e-9
    cn = switching_diode(I, G, C, isat).to_c()
    O.pivot_stats()
    t, X, names, _ = O.tran(cn, dt, 3e-6, [("a", 0.0)], opts=no_gmin())
    fact, changes = O.pivot_stats()
    assert fact > 300 and changes >= 1
    a = X[:, names.index("a")]
    off = np.nonzero(a > 19.5)[0]
    assert len(off) > 100
    k = off[1:]
    pred = (C / dt * a[k - 1] + I - isat) / (C / dt + G + gic)
    assert np.abs(a[k] - pred).max() <= 1e-10
    # the source holds b at 0 V and carries the diode's reverse current
    assert np.all(X[:, names.index("b")] == 0.0)
    assert np.abs(X[k, names.index("vb")] - isat).max() <= 1e-12
