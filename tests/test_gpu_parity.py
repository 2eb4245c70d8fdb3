"""MI355X engine against the oracle (and the reference's golden waveforms).  Needs the GPU.

Tolerances: device evaluation uses the GPU's exp/sqrt (within an ulp or so of glibc's), so assembled
values agree to 1e-12 relative; FP64 Newton solutions agree with the oracle to 1e-9 relative (the bar
for the plain FP64 path); mixed precision agrees with FP64 to 1e-6 V."""
import json
import os

import numpy as np
import pytest

from oracle import oracle_py as O
from spice21_amd import workloads as W
from spice21_amd.batch import FP64, MIXED, Batch, Group
from spice21_amd.elab import elaborate
from tests.circuits import GOLDEN_RUNS, materialize, py_ring3, random_ac_circuit, ro3
from tests.test_symbolic import random_circuit

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def log(*a):
    print(*a, flush=True)


def wave_within(X, R, names, kinds, what):
    """Signal by signal: node voltages within 1e-9 of their swing (of 1 uV at least: a node the reference
    holds at exactly 0 V, such as an IC source's, may read 1e-23 V here); branch currents within 1e-9 of
    their swing or 1e-12 A, the reference's own current tolerance (iabstol, analysis.rs:331-345), which
    bounds how closely two FP64 Newton paths that round exp differently in the last bit can agree on a
    current.  Logs each signal's deviation; returns whether all are within."""
    ok = True
    for k, nm in enumerate(names):
        swing = float(np.max(np.abs(R[:, k])))
        dev = float(np.max(np.abs(X[:, k] - R[:, k])))
        tol = 1e-9 * max(swing, 1e-6) if kinds[k] == 0 else max(1e-9 * swing, 1e-12)
        log(f"  {what} {nm}: max |dev| {dev:.3e} (swing {swing:.3e}, tolerance {tol:.3e})")
        ok = ok and dev <= tol
    return ok


def rel_err(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-9))) if len(a) else 0.0


@pytest.mark.parametrize("which", ["cmos", "mos0", "chain", "rand3", "rand8"])
def test_assembly_matches_oracle(which):
    if which == "chain":
        nl = W.chain(64)[0]
    elif which.startswith("rand"):
        nl = random_circuit(int(which[4:]))
    else:
        nl = ro3(which)
    cn = nl.to_c()
    b = Batch(Group(cn))
    rng = np.random.default_rng(1)
    for trial in range(3):
        x = rng.uniform(-0.2, 1.2, size=cn.n_vars)
        log("assemble", which, trial)
        r, c, v, rhs = b.assemble(x)
        orr, occ, ov, orhs = O.assemble_op(cn, x)
        assert list(r) == list(orr) and list(c) == list(occ)
        scale = np.maximum(np.abs(ov), 1e-15)
        assert np.max(np.abs(v - ov) / scale) < 1e-12, (which, trial)
        assert np.max(np.abs(rhs - orhs) / np.maximum(np.abs(orhs), 1e-15)) < 1e-12


@pytest.mark.parametrize("which", ["ring5", "chain", "cmos", "mos0", "rand1", "rand5", "rand11"])
def test_dcop_matches_oracle(which):
    if which == "ring5":
        nl = W.ring5()
    elif which == "chain":
        nl = W.chain(64)[0]
    elif which.startswith("rand"):
        nl = random_circuit(int(which[4:]))
    else:
        nl = ro3(which)
    cn = nl.to_c()
    log("dcop", which)
    b = Batch(Group(cn)).op()
    st, it = b.status()
    try:
        ox, oit, _ = O.dcop(cn)
    except O.OracleError as e:  # the reference fails on this draw: so must the engine, the same way
        assert e.desc == "Convergence Failed" and st[0] == 2, (which, e.desc, st[0])
        return
    assert st[0] == 0
    x = b.x()[0]
    assert rel_err(x, ox) < 1e-9, (which, x, ox)


@pytest.mark.parametrize("fixture", sorted(GOLDEN_RUNS))
def test_tran_matches_golden(fixture):
    kind, tstep, tstop = GOLDEN_RUNS[fixture]
    g = json.load(open(os.path.join(GOLDEN, fixture)))
    cn = ro3(kind).to_c()
    log("tran", fixture)
    b = Batch(Group(cn)).tran(tstep, tstop, [("1", 0.0)])
    st, it = b.status()
    assert st[0] == 0
    t = b.axis()
    X = b.wave()
    names = b.var_names()
    assert np.array_equal(t, np.array(g["time"]))
    # the reference's own test allows 1e-6 absolute (assert.rs:124-143); this one 1e-9 of each swing
    R = np.stack([np.array(g[nm]) for nm in names], axis=1)
    kinds = list(ro3(kind).kinds) + [0, 1]  # the IC's .1.vic (voltage) and .1.iic (current)
    assert wave_within(X, R, names, kinds, "golden")


def test_ring5_full_transient_matches_oracle_fp64_and_mixed():
    """The single 5-stage ring at its full length: dcop, then 0-10 ns at 1 ps (10,001 points) from the
    0.5 V kick on node 1.  FP64 agrees with the oracle to 1e-9 of each signal's swing with the same
    Newton iteration total; mixed precision agrees with FP64 to 1e-6 V on node voltages."""
    cn = W.ring5().to_c()
    log("ring5 full transient: oracle")
    t_o, X_o, names_o, it_o = O.tran(cn, 1e-12, 10e-9, [("1", 0.5)])
    assert len(t_o) == 10001
    log("ring5 full transient: fp64")
    b = Batch(Group(cn)).tran(1e-12, 10e-9, [("1", 0.5)])
    st, it = b.status()
    assert st[0] == 0
    X = b.wave()
    assert b.var_names() == list(names_o) and np.array_equal(b.axis(), t_o)
    log("  iterations", it[0], "oracle", it_o)
    assert wave_within(X, X_o, names_o, list(W.ring5().kinds) + [0, 1], "fp64 vs oracle")
    assert it[0] == it_o
    log("ring5 full transient: mixed")
    bm = Batch(Group(cn), precision=MIXED).tran(1e-12, 10e-9, [("1", 0.5)])
    assert bm.status()[0][0] == 0
    Xm = bm.wave()
    nv = [k for k, kind in enumerate(W.ring5().kinds) if kind == 0]
    err_v = float(np.abs(Xm[:, nv] - X[:, nv]).max())
    log("  mixed vs fp64 max node-voltage error", err_v)
    assert err_v <= 1e-6


def test_chain_mc_batch_matches_oracle_per_instance():
    n_inst = 96
    nl, over = W.chain_mc(n_inst=n_inst)
    cn = nl.to_c()
    log("chain_mc fp64", n_inst)
    b = Batch(Group(cn, n_inst, over)).op()
    st, it = b.status()
    assert np.all(st == 0)
    X = b.x()
    for i in [0, 1, 17, 63, 95]:
        ox, _, _ = O.dcop(materialize(nl, over, i).to_c())
        assert rel_err(X[i], ox) < 1e-9, i
    log("chain_mc mixed", n_inst)
    bm = Batch(Group(cn, n_inst, over), precision=MIXED).op()
    stm, itm = bm.status()
    assert np.all(stm == 0)
    Xm = bm.x()
    nv = [k for k, kind in enumerate(nl.kinds) if kind == 0]
    assert np.max(np.abs(Xm[:, nv] - X[:, nv])) <= 1e-6


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_results_bit_identical_across_runs_and_launch_shapes(precision):
    """Stamps are summed in a fixed order (slot codes, hub cells) and each instance is factored by its own
    lanes, so neither a repeated call nor another launch shape (instances per wavefront, wavefronts per
    workgroup) changes a single bit of the solution or an iteration count; nor does a repeated AC sweep."""
    nl, over = W.chain_mc(n_inst=512)
    g = Group(nl.to_c(), 512, over)
    ref = None
    for ipw in (2, 1, 4, 2):
        log("determinism", precision, "ipw", ipw)
        b = Batch(g, precision=precision).set_ipw(ipw).op()
        x, (st, it) = b.x(), b.status()
        assert np.all(st == 0)
        if ref is None:
            ref = (x, it)
        else:
            assert np.array_equal(x, ref[0]) and np.array_equal(it, ref[1]), ipw
    nl, over = W.amp_ladder(16)
    ga = Group(nl.to_c(), 16, over)
    a = Batch(ga, precision=precision).ac(1000, 100_000_000_000, 300)
    c = Batch(ga, precision=precision).ac(1000, 100_000_000_000, 300)
    for i in range(16):
        assert np.array_equal(a.wave(0, i), c.wave(0, i)), i


@pytest.mark.parametrize("jit", [-1, 1])
@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_repeated_analyses_on_one_batch_are_bit_identical(precision, jit):
    """A batch keeps its structure between calls and only resets its Newton state: op, op again, then
    an AC sweep, an assembly and op once more on the same batch give bit-identical operating points and
    iteration counts, and the assembly sees the x it was given."""
    nl, over = W.chain_mc(n_inst=64)
    cn = nl.to_c()
    b = Batch(Group(cn, 64, over), precision=precision).set_jit(jit)
    log("repeat", precision, jit, "op 1")
    x1, (st1, it1) = b.op().x(), b.status()
    assert np.all(st1 == 0)
    log("repeat", precision, jit, "op 2")
    x2, (_, it2) = b.op().x(), b.status()
    assert np.array_equal(x1, x2) and np.array_equal(it1, it2)
    log("repeat", precision, jit, "ac")
    b.ac(1000, 1_000_000_000, 20)
    assert np.all(b.status()[0] == 0)
    log("repeat", precision, jit, "assemble")
    xa = np.random.default_rng(3).uniform(-0.2, 1.2, size=cn.n_vars)
    r, c, v, rhs = b.assemble(xa)
    orr, occ, ov, orhs = O.assemble_op(materialize(nl, over, 0).to_c(), xa)
    assert list(r) == list(orr) and list(c) == list(occ)
    assert np.max(np.abs(v - ov) / np.maximum(np.abs(ov), 1e-15)) < 1e-12
    assert np.max(np.abs(rhs - orhs) / np.maximum(np.abs(orhs), 1e-15)) < 1e-12
    log("repeat", precision, jit, "op 3")
    x3, (_, it3) = b.op().x(), b.status()
    assert np.array_equal(x1, x3) and np.array_equal(it1, it3)


def test_mixed_tran_close_to_fp64():
    cn = ro3("cmos").to_c()
    a = Batch(Group(cn)).tran(1e-11, 1e-8, [("1", 0.0)]).wave()
    bm = Batch(Group(cn), precision=MIXED).tran(1e-11, 1e-8, [("1", 0.0)])
    assert bm.status()[0][0] == 0
    m = bm.wave()
    log("mixed vs fp64 tran", float(np.max(np.abs(m[:, :4] - a[:, :4]))))
    assert np.max(np.abs(m[:, :4] - a[:, :4])) <= 1e-6


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_generated_factorisation_matches_level_schedule(precision):
    """The run-time generated factorisation performs the record-driven one's operations in the same order.
    FP64: bit for bit.  Mixed precision: its FP32 updates are fused multiply-adds and it does not skip a
    zero c_k, so the FP32 preconditioner rounds differently; the Newton iterates still reach the FP64
    fixed point in as many iterations on average (within 2 %); the two mixed results differ by what
    separates either from FP64 (a few nV, within the 1e-6 V bar), and are held to 1e-8 V of each other."""
    n_inst = 600
    nl, over = W.chain_mc(n_inst=n_inst)
    g = Group(nl.to_c(), n_inst, over)
    log("levels", precision)
    a = Batch(g, precision=precision).set_jit(-1).op()
    assert a.config()[2] is False
    log("generated", precision)
    b = Batch(g, precision=precision).set_jit(1).op()
    assert b.config()[2] is True
    sa, ia = a.status()
    sb, ib = b.status()
    assert np.all(sa == 0) and np.all(sb == 0)
    if precision == FP64:
        assert np.array_equal(ia, ib)
        assert np.max(np.abs(a.x() - b.x())) <= 1e-12
    else:
        log("  iterations: level schedule", int(ia.sum()), "generated", int(ib.sum()), "max |diff|",
            int(np.max(np.abs(ia - ib))))
        assert abs(float(ib.mean()) - float(ia.mean())) <= 0.02 * float(ia.mean())
        assert np.max(np.abs(a.x() - b.x())) <= 1e-8
    if precision == FP64:
        x0, _, _ = O.dcop(materialize(nl, over, 7).to_c())
        assert rel_err(b.x()[7], x0) < 1e-9


@pytest.mark.parametrize("ipw", [1, 4])
def test_generated_factorisation_transient_matches_golden(ipw):
    kind, tstep, tstop = GOLDEN_RUNS["test_mos1_cmos_ro_tran.json"]
    g = json.load(open(os.path.join(GOLDEN, "test_mos1_cmos_ro_tran.json")))
    cn = ro3(kind).to_c()
    log("generated tran ipw", ipw)
    b = Batch(Group(cn, 3)).set_jit(1).set_ipw(ipw).tran(tstep, tstop, [("1", 0.0)])
    assert b.config()[2] is True
    st, _ = b.status()
    assert np.all(st == 0)
    names = b.var_names()
    for inst in range(3):
        X = b.wave(0, inst)
        worst = max(float(np.max(np.abs(X[:, k] - np.array(g[nm])))) for k, nm in enumerate(names))
        assert worst <= 1e-6, (inst, worst)


def test_ac_batch_matches_oracle_per_instance():
    """24 two-stage CS amplifiers with per-instance 20-section RC ladders, swept from 1 kHz to 100 GHz."""
    n_inst = 24
    nl, over = W.amp_ladder(n_inst)
    cn = nl.to_c()
    log("ac fp64", n_inst)
    b = Batch(Group(cn, n_inst, over)).ac(1000, 100_000_000_000, 1000)
    st, it = b.status()
    assert np.all(st == 0), st
    f = b.axis()
    for i in [0, 5, 23]:
        fo, Xo, _ = O.ac(materialize(nl, over, i).to_c(), 1000, 100_000_000_000, 1000, max_pts=1100)
        assert np.array_equal(f, fo)
        X = b.wave(0, i)
        err = np.max(np.abs(X - Xo) / np.maximum(np.abs(Xo), 1e-9))
        log("  inst", i, "max rel", err)
        assert err < 1e-9, (i, err)


def test_ac_runs_of_frequencies_match_one_sequential_sweep():
    """Each instance's sweep cut into runs solved in parallel and stitched: the same solutions and the
    same Newton iteration counts as one sequential sweep, which repeats the reference's warm starts (its
    counts equal the oracle's).  Without the stitching the runs' cold starts differ by up to 6e-4
    relative on this sweep: a frequency whose residual at the previous solution is below 1e-9 keeps that
    solution in the reference, so its result depends on where the sweep came from."""
    n_inst = 12
    nl, over = W.amp_ladder(n_inst)
    g = Group(nl.to_c(), n_inst, over)
    log("ac runs: sequential")
    seq = Batch(g).set_ac_runs(1).ac(1000, 100_000_000_000, 300)
    assert seq.ac_runs() == 1
    st, it = seq.status()
    assert np.all(st == 0), st
    for i in [0, 7]:
        ci = materialize(nl, over, i).to_c()
        fo, Xo, ito = O.ac(ci, 1000, 100_000_000_000, 300, max_pts=400)
        ito -= O.dcop(ci)[1]  # the oracle counts the operating point's iterations too
        assert it[i] == ito, (i, it[i], ito)
        err = np.max(np.abs(seq.wave(0, i) - Xo) / np.maximum(np.abs(Xo), 1e-9))
        assert err < 1e-9, (i, err)
    nf = len(seq.axis())
    for runs in (0, 7, nf + 5):
        log("ac runs:", runs)
        par = Batch(g).set_ac_runs(runs).ac(1000, 100_000_000_000, 300)
        assert par.ac_runs() == (min(runs, nf) if runs else par.ac_runs()) and par.ac_runs() > 1
        st2, it2 = par.status()
        assert np.all(st2 == 0), st2
        log("  iterations", int(it2.sum()), "sequential", int(it.sum()))
        assert np.array_equal(it2, it)
        worst = 0.0
        for i in range(n_inst):
            a = seq.wave(0, i)
            worst = max(worst, float(np.max(np.abs(par.wave(0, i) - a) / np.maximum(np.abs(a), 1e-9))))
        log("  max rel vs sequential", worst)
        assert worst < 1e-9, (runs, worst)


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_ac_generated_factorisation_matches_level_schedule(precision):
    """The AC kernel with the complex factorisation written out (one lane per instance) against the
    record-driven one: the same operations, so the same iteration counts, and values equal to rounding."""
    n_inst = 24
    nl, over = W.amp_ladder(n_inst)
    g = Group(nl.to_c(), n_inst, over)
    log("ac levels", precision)
    a = Batch(g, precision=precision).set_jit(-1).ac(1000, 100_000_000_000, 500)
    assert a.config()[2] is False
    log("ac generated", precision)
    b = Batch(g, precision=precision).set_jit(1).ac(1000, 100_000_000_000, 500)
    assert b.config()[2] is True
    sa, ia = a.status()
    sb, ib = b.status()
    assert np.all(sa == 0) and np.all(sb == 0)
    worst = max(rel_err(b.wave(0, i).ravel(), a.wave(0, i).ravel()) for i in range(n_inst))
    log("  iterations", int(ia.sum()), int(ib.sum()), "max rel", worst)
    assert np.array_equal(ia, ib)
    assert worst <= (1e-12 if precision == FP64 else 1e-8)
    if precision == FP64:
        fo, Xo, _ = O.ac(materialize(nl, over, 3).to_c(), 1000, 100_000_000_000, 500, max_pts=600)
        assert rel_err(b.wave(0, 3).ravel(), Xo.ravel()) < 1e-9


@pytest.mark.parametrize("jit", [-1, 1])
def test_ac_batch_of_two_patterns_matches_oracle(jit):
    """Two groups of different AC patterns in one batch (two generated complex factorisations chosen by
    group index, or the record-driven ones), each sweep cut into runs and stitched: every instance
    against the oracle, with the sequential sweep's iteration counts."""
    parts = [(W.amp_ladder(5, sections=20, seed=1), 5), (W.amp_ladder(3, sections=7, seed=2), 3)]
    groups = [Group(nl.to_c(), k, over) for (nl, over), k in parts]
    log("ac two patterns, jit", jit)
    b = Batch(groups).set_jit(jit).ac(1000, 100_000_000_000, 400)
    assert b.config()[2] is (jit == 1) and b.ac_runs() > 1
    seq = Batch(groups).set_jit(jit).set_ac_runs(1).ac(1000, 100_000_000_000, 400)
    for gi, ((nl, over), k) in enumerate(parts):
        st, it = b.status(gi)
        st1, it1 = seq.status(gi)
        assert np.all(st == 0) and np.all(st1 == 0)
        assert np.array_equal(it, it1), (gi, it, it1)
        for i in range(k):
            fo, Xo, _ = O.ac(materialize(nl, over, i).to_c(), 1000, 100_000_000_000, 400, max_pts=500)
            assert np.array_equal(b.axis(), fo)
            err = rel_err(b.wave(gi, i).ravel(), Xo.ravel())
            assert err < 1e-9, (gi, i, err)


@pytest.mark.parametrize("jit", [-1, 1])
def test_random_circuits_ac_match_oracle(jit):
    """30 random netlists of R, C, Mos1 (some with terminal resistances, some with gate capacitances) and
    AC sources, one group each in one batch, swept from 1 kHz to 10 GHz in runs of frequencies: each
    sweep's outcome and values against the oracle (3 of the 30 fail to converge there)."""
    nls = [random_ac_circuit(s) for s in range(30)]
    log("random ac, jit", jit)
    b = Batch([Group(nl.to_c()) for nl in nls]).set_jit(jit).ac(1000, 10_000_000_000, 200)
    log("  generated", b.config()[2], "runs", b.ac_runs())
    f = b.axis()
    fails = 0
    for gi, nl in enumerate(nls):
        st, _ = b.status(gi)
        try:
            fo, Xo, _ = O.ac(nl.to_c(), 1000, 10_000_000_000, 200, max_pts=300)
        except O.OracleError as e:
            assert e.desc == "Convergence Failed" and st[0] == 2, (gi, e.desc, st[0])
            fails += 1
            continue
        assert st[0] == 0, (gi, st[0])
        assert np.array_equal(f, fo)
        err = rel_err(b.wave(gi, 0).ravel(), Xo.ravel())
        assert err < 1e-9, (gi, err)
    assert fails == 3


@pytest.mark.parametrize("f0,f1,npts", [(1000, 1000, 10), (1000, 2000, 1), (1000, 100_000, 17), (7, 7_000_000, 40)])
def test_ac_short_sweeps_match_oracle(f0, f1, npts):
    """Sweeps of one, two and a few dozen frequencies (one run, or runs of a couple of frequencies)."""
    nl, over = W.amp_ladder(4)
    b = Batch(Group(nl.to_c(), 4, over)).ac(f0, f1, npts)
    f = b.axis()
    assert np.all(b.status()[0] == 0)
    for i in range(4):
        fo, Xo, _ = O.ac(materialize(nl, over, i).to_c(), f0, f1, npts, max_pts=npts + 10)
        assert np.array_equal(f, fo)
        assert rel_err(b.wave(0, i).ravel(), Xo.ravel()) < 1e-9, (i, b.ac_runs())


def test_ac_mixed_close_to_fp64():
    n_inst = 16
    nl, over = W.amp_ladder(n_inst)
    g = Group(nl.to_c(), n_inst, over)
    a = Batch(g).ac(1000, 100_000_000_000, 200)
    m = Batch(g, precision=MIXED).ac(1000, 100_000_000_000, 200)
    assert np.all(m.status()[0] == 0)
    worst = max(float(np.max(np.abs(m.wave(0, i) - a.wave(0, i)))) for i in range(n_inst))
    log("ac mixed vs fp64", worst)
    assert worst <= 1e-6


def test_ac_byte_abi_matches_oracle():
    from spice21_amd import api, protos

    ckt = protos.Circuit(name="rc")
    ckt.comps.append(protos.Instance(v=protos.Vsrc(name="vi", p="inp", n="", dc=0.0, acm=1.0)))
    ckt.comps.append(protos.Instance(r=protos.Resistor(name="r1", p="inp", n="out", g=1e-3)))
    ckt.comps.append(protos.Instance(c=protos.Capacitor(name="c1", p="out", n="", c=1e-9)))
    res = api.ac(ckt, args=protos.AcOptions(fstart=1, fstop=10_000_000, npts=70))
    nl = W.Netlist()
    nl.v("vi", 0.0, "inp", "", acm=1.0)
    nl.r(1e-3, "inp", "out")
    nl.c(1e-9, "out", "")
    f, X, _ = O.ac(nl.to_c(), 1, 10_000_000, 70)
    out = np.array(res["out"])
    assert out.shape == (len(f),)
    assert np.max(np.abs(out - X[:, 2]) / np.abs(X[:, 2])) < 1e-9


def test_uneven_ring_batch_transient_matches_oracle():
    """Uneven ring oscillators: several groups of different size in one launch, per-instance C and VDD,
    300 steps."""
    gs = W.ring_groups(48, seed=0xC3)[:4]
    groups = [Group(n.to_c(), k, over) for n, k, over, _ in gs]
    log("ring batch groups", [(g.cnet.n_vars, g.n_inst) for g in groups])
    b = Batch(groups).tran(1e-12, 3e-10, [("1", 0.0)])
    for gi, (n, k, over, _) in enumerate(gs):
        st, it = b.status(gi)
        assert np.all(st == 0), (gi, st)
        for i in sorted({0, k - 1}):
            t, X, names, _ = O.tran(materialize(n, over, i).to_c(), 1e-12, 3e-10, [("1", 0.0)])
            Xg = b.wave(gi, i)
            assert Xg.shape == X.shape
            err = float(np.max(np.abs(Xg - X)))
            log("  group", gi, "inst", i, "max |dv|", err)
            assert err <= 1e-9, (gi, i, err)


def test_mixed_ring_batch_with_hub_cells_matches_fp64_and_oracle():
    """256 uneven rings in mixed precision: at most two instances per SIMD, so every shared slot is a hub
    (one evaluation round per device class); the FP64 run of the same batch keeps the colour rounds.
    Every instance's waveform within the mixed-precision bar (1e-6 V) of the FP64 one, and a few of them
    against the oracle."""
    gs = W.ring_groups(256, seed=0xC3)
    groups = [Group(n.to_c(), k, over) for n, k, over, _ in gs]
    log("hub-cell ring batch", len(groups), "groups")
    m = Batch(groups, precision=MIXED).tran(1e-12, 2e-10, [("1", 0.0)])
    f = Batch(groups).tran(1e-12, 2e-10, [("1", 0.0)])
    assert m.structure(0)["n_contrib"] > f.structure(0)["n_contrib"]  # the hub cells
    worst = 0.0
    for gi, (n, k, over, _) in enumerate(gs):
        assert np.all(m.status(gi)[0] == 0) and np.all(f.status(gi)[0] == 0), gi
        for i in range(k):
            worst = max(worst, float(np.max(np.abs(m.wave(gi, i) - f.wave(gi, i)))))
    log("  mixed vs fp64 max |dv|", worst)
    assert worst <= 1e-6
    for gi in (0, len(gs) // 2, len(gs) - 1):
        n, k, over, _ = gs[gi]
        t, X, _, _ = O.tran(materialize(n, over, k - 1).to_c(), 1e-12, 2e-10, [("1", 0.0)])
        assert float(np.max(np.abs(m.wave(gi, k - 1) - X))) <= 1e-6
        assert float(np.max(np.abs(f.wave(gi, k - 1) - X))) <= 1e-9


S21_CONVERGENCE = 2


def _random_batch_cases(n=60):
    """Random netlists (NMOS/PMOS Mos1 with and without rd/rs, Mos0, diodes with and without rs, R, C, I,
    V; tests/test_symbolic.py) with the oracle's dcop outcome for each."""
    cases = []
    for s in range(n):
        cn = random_circuit(s).to_c()
        try:
            ox, oit, _ = O.dcop(cn)
            cases.append((s, cn, ox, oit))
        except O.OracleError as e:
            assert e.desc == "Convergence Failed", (s, e.desc)
            cases.append((s, cn, None, None))
    return cases


# Seeds whose kept pivot order meets an exact zero pivot in FP64 (seed 56: at the second iteration, with
# a diode's 1.5e15 S conductance in the matrix; the reference searches again and moves on).  The engine
# resumes such an instance on the order the reference's search picks there; the reference changes its
# order twice more later in this solve, so the last digits follow another rounding path through
# Jacobians of condition number up to 4e19, and the solution is held to 1e-9 of its norm instead.
RESUMED_FP64 = {56}


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_random_circuits_dcop_outcome_matches_oracle(precision):
    """60 random netlists as 60 groups of one launch.  Where the oracle converges the engine converges to
    the same point (FP64: 1e-9 relative, same iteration count; mixed: 1e-6 V of it); where the oracle
    reports "Convergence Failed" (analysis.rs:209) the instance's status is S21_CONVERGENCE."""
    cases = _random_batch_cases()
    log("random dcop", precision, "groups", len(cases), "oracle failures", sum(c[2] is None for c in cases))
    b = Batch([Group(cn) for _, cn, _, _ in cases], precision=precision).op()
    it_diff = []
    for gi, (s, cn, ox, oit) in enumerate(cases):
        st, it = b.status(gi)
        if ox is None:
            assert st[0] == S21_CONVERGENCE, (s, st[0])
            continue
        assert st[0] == 0, (s, st[0])
        x = b.x(gi)[0]
        if precision == FP64 and s in RESUMED_FP64:
            log("  seed", s, "resumed: iterations", int(it[0]), "oracle", oit,
                "max |dx| / max |x|", float(np.max(np.abs(x - ox)) / np.max(np.abs(ox))))
            assert np.max(np.abs(x - ox)) <= 1e-9 * np.max(np.abs(ox)), (s, x, ox)
        elif precision == FP64:
            # 1e-9 relative, with a 1 fV / 1 fA floor for unknowns the residual test (1e-12) leaves at noise
            assert np.all(np.abs(x - ox) <= 1e-9 * np.abs(ox) + 1e-15), (s, x, ox)
            if it[0] != oit:
                it_diff.append((s, int(it[0]), oit))
        else:
            v = [k for k in range(cn.n_vars) if cn.kinds[k] == 0]
            assert np.max(np.abs(x[v] - ox[v])) <= 1e-6, s
    assert not it_diff, it_diff


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_random_circuits_transient_matches_oracle(precision):
    """The 36 random netlists the oracle can bias, 100 steps of 1 ps each, one group per netlist: FP64 to
    1e-9 of the oracle (relative, 1 V / 1 A floor), mixed precision to 1e-6 V."""
    cases = [c for c in _random_batch_cases() if c[2] is not None]
    b = Batch([Group(cn) for _, cn, _, _ in cases], precision=precision).tran(1e-12, 1e-10)
    for gi, (s, cn, _, _) in enumerate(cases):
        try:
            t, X, names, _ = O.tran(cn, 1e-12, 1e-10)
        except O.OracleError as e:
            st, _ = b.status(gi)
            assert st[0] != 0, (s, e.desc)
            continue
        st, _ = b.status(gi)
        assert st[0] == 0, (s, st[0])
        Xg = b.wave(gi, 0)
        assert Xg.shape == X.shape
        if precision == FP64:
            err = float(np.max(np.abs(Xg - X) / np.maximum(np.abs(X), 1.0)))
            assert err <= 1e-9, (s, err)
        else:
            v = [k for k in range(cn.n_vars) if cn.kinds[k] == 0]
            err = float(np.max(np.abs(Xg[:, v] - X[:, v])))
            assert err <= 1e-6, (s, err)


def test_capacitor_hub_cells_read_zero_in_operating_points():
    """A capacitor stamps nothing in an operating point (comps/mod.rs:205-209).  In the three-stage ring
    every stage's capacitor ends on vss, a hub slot, so those capacitors own hub cells; in the t = 0
    operating point of a transient and in the DC point of an AC sweep the cells must read zero, not what
    an earlier launch left in the LDS.  A large batch runs first, on every CU, to leave such content."""
    nl, over = W.chain_mc(4096)
    Batch(Group(nl.to_c(), 4096, over)).op()
    e = elaborate(py_ring3())
    b = Batch(Group(e)).tran(1e-11, 5e-10, [("ring0", 0.0)])
    assert b.structure()["n_contrib"] > 0  # vss is a hub slot of the transient's structure
    assert b.status()[0][0] == 0
    t, X, names, _ = O.tran(e, 1e-11, 5e-10, [("ring0", 0.0)])
    err = float(np.max(np.abs(b.wave(0, 0) - X)))
    log("ring tran after a large batch: max |dx|", err)
    assert err <= 1e-9
    Batch(Group(nl.to_c(), 4096, over)).op()
    ea = elaborate(py_ring3(acm=1.0))
    ba = Batch(Group(ea)).ac(1, 1_000_000_000, 20)
    assert ba.status()[0][0] == 0
    f, Xa, _ = O.ac(ea, 1, 1_000_000_000, 20)
    assert np.array_equal(ba.axis(), f)
    Xg = ba.wave(0, 0)
    err = float(np.max(np.abs(Xg - Xa) / np.maximum(np.abs(Xa), 1e-12)))
    log("ring ac after a large batch: max rel", err)
    assert err <= 1e-9


def test_mos1_gate_capacitance_transient_matches_oracle():
    """Mos1 with tox and overlap capacitances (the Meyer charge path, its committed operating point and
    the reference's tr.gs quirk) in a 500-point ring transient, one instance and a batch of 32 whose
    overlap capacitance differs per instance: FP64 to 1e-9 of the oracle, mixed precision to 1e-6 V.
    Node 1 starts at 0 V: from the ring's symmetric operating point (every node at 0.5 V) the symmetry
    breaking amplifies any last-bit difference about e-fold every ten points."""
    from tests.circuits import ring_caps

    nl = ring_caps()
    cn = nl.to_c()
    t, X, names, _ = O.tran(cn, 1e-12, 5e-10, [("1", 0.0)])
    b = Batch(Group(cn)).tran(1e-12, 5e-10, [("1", 0.0)])
    assert b.status()[0][0] == 0
    assert np.array_equal(b.axis(), t)
    err = float(np.max(np.abs(b.wave(0, 0) - X)))
    log("ring with gate capacitance, FP64: max |dx|", err)
    assert err <= 1e-9
    m = Batch(Group(cn), precision=MIXED).tran(1e-12, 5e-10, [("1", 0.0)])
    assert m.status()[0][0] == 0
    assert float(np.max(np.abs(m.wave(0, 0)[:, :5] - X[:, :5]))) <= 1e-6
    # a batch: every Mos1's gate-drain overlap varies per instance (the gate-source one never reaches a
    # stamp: the reference's gd integration overwrites tr.gs, comps/mos.rs:788, 801)
    n_inst = 32
    mos = [ci for ci, c in enumerate(nl.comps) if c[0] == 4]  # S21_COMP_MOS1
    vals = 2e-10 * (1.0 + 0.5 * np.linspace(-1.0, 1.0, n_inst))
    over = [(ci, "model", "cgdo", vals) for ci in mos]
    bb = Batch(Group(cn, n_inst, over)).tran(1e-12, 5e-10, [("1", 0.0)])
    st, _ = bb.status()
    worst, failed = 0.0, []
    for i in range(n_inst):  # the reference fails to converge on some of these; so must the engine
        try:
            ti, Xi, _, _ = O.tran(materialize(nl, over, i).to_c(), 1e-12, 5e-10, [("1", 0.0)])
        except O.OracleError as e:
            assert e.desc == "Convergence Failed" and st[i] == 2, (i, e.desc, st[i])
            failed.append(i)
            continue
        assert st[i] == 0, (i, st[i])
        worst = max(worst, float(np.max(np.abs(bb.wave(0, i) - Xi))))
    log("  batch: max |dx|", worst, "instances the oracle cannot simulate", failed)
    assert worst <= 1e-9
    ok = [i for i in range(n_inst) if i not in failed]
    assert float(np.max(np.abs(bb.wave(0, ok[0]) - bb.wave(0, ok[-1])))) > 0.1  # the overrides matter


def test_mos1_gate_capacitance_ac_matches_oracle():
    """The AC sweep of an amplifier chain whose Mos1 have gate capacitance (the jwC stamps of
    Mos1::load_ac from the captured Meyer capacitances): FP64 to 1e-9 of the oracle."""
    from tests.circuits import cs_amp_caps

    cn = cs_amp_caps().to_c()
    f, X, _ = O.ac(cn, 1000, 100_000_000_000, 40)
    b = Batch(Group(cn)).ac(1000, 100_000_000_000, 40)
    assert b.status()[0][0] == 0
    assert np.array_equal(b.axis(), f)
    Xg = b.wave(0, 0)
    err = float(np.max(np.abs(Xg - X) / np.maximum(np.abs(X), 1e-12)))
    log("amplifier with gate capacitance, AC: max rel", err)
    assert err <= 1e-9


def test_diode_charge_transient_outcome_matches_oracle():
    """Diode junction charge in a transient (node a charged through 1 kohm from 0 V): reverse-biased the
    oracle converges at every point and the engine matches it to 1e-9; forward-biased the reference's
    depletion-capacitance formula (comps/diode.rs:311-318) keeps Newton from converging, and the engine
    reports the same "Convergence Failed"."""
    from tests.circuits import diode_rc

    cn = diode_rc().to_c()
    t, X, names, it = O.tran(cn, 1e-11, 2e-9, [("a", 0.0)])
    b = Batch(Group(cn)).tran(1e-11, 2e-9, [("a", 0.0)])
    st, its = b.status()
    assert st[0] == 0 and np.array_equal(b.axis(), t)
    err = float(np.max(np.abs(b.wave(0, 0) - X)))
    log("reverse-biased diode charge: max |dx|", err, "iterations", int(its[0]), "oracle", it)
    assert err <= 1e-9
    for forward, tt in ((True, 0.0), (True, 1e-10)):
        cf = diode_rc(tt=tt, forward=forward).to_c()
        with pytest.raises(O.OracleError, match="Convergence Failed"):
            O.tran(cf, 1e-11, 2e-9, [("a", 0.0)])
        assert Batch(Group(cf)).tran(1e-11, 2e-9, [("a", 0.0)]).status()[0][0] == 2


def test_generated_factorisation_several_patterns():
    """One generated kernel for a batch of several patterns (each group's factorisation chosen by its
    index): FP64 bit for bit the record-driven results on the 60 random netlists (one group each), and
    the uneven ring batch's transient against the oracle."""
    cases = _random_batch_cases()
    groups = [Group(cn) for _, cn, _, _ in cases]
    a = Batch(groups).set_jit(-1).op()
    b = Batch(groups).set_jit(1).op()
    assert a.config()[2] is False and b.config()[2] is True
    for gi in range(len(groups)):
        sa, ia = a.status(gi)
        sb, ib = b.status(gi)
        assert sa[0] == sb[0] and ia[0] == ib[0], (gi, sa[0], sb[0], ia[0], ib[0])
        if sa[0] == 0:
            assert np.array_equal(a.x(gi), b.x(gi)), gi
    gs = W.ring_groups(48, seed=0xC3)[:6]
    rg = [Group(n.to_c(), k, over) for n, k, over, _ in gs]
    for prec, tol in ((FP64, 1e-9), (MIXED, 1e-6)):
        r = Batch(rg, precision=prec).set_jit(1).tran(1e-12, 3e-10, [("1", 0.0)])
        assert r.config()[2] is True
        for gi, (n, k, over, _) in enumerate(gs):
            st, _ = r.status(gi)
            assert np.all(st == 0), (gi, st)
            t, X, names, _ = O.tran(materialize(n, over, k - 1).to_c(), 1e-12, 3e-10, [("1", 0.0)])
            err = float(np.max(np.abs(r.wave(gi, k - 1)[:, :4] - X[:, :4])))
            assert err <= tol, (prec, gi, err)


def _floating_cap():
    n = W.Netlist()
    n.v("v1", 1.0, "in", "")
    n.r(1e-3, "in", "a")
    n.c(1e-12, "a", "b")  # b has no DC path: its matrix entries exist but are zero in the operating point
    return n


def _nan_resistor():
    n = W.Netlist()
    n.v("v1", 1.0, "in", "")
    n.r(float("nan"), "in", "a")
    n.r(1e-3, "a", "")
    return n


def test_failure_outcomes_match_oracle_and_stay_per_instance():
    """Failures the reference reports the engine reports per instance, leaving the rest of the batch
    alone: a node reached only through a capacitor meets a zero pivot in the operating point (the
    oracle's "Assert Neq Failed: zero pivot", sparse21/mod.rs:872 -> S21_PIVOT); a NaN conductance makes
    every unknown NaN yet passes the convergence test, whose > comparisons are false for NaN, in the
    oracle and the engine alike (analysis.rs:331-345)."""
    fl, nn = _floating_cap().to_c(), _nan_resistor().to_c()
    with pytest.raises(O.OracleError) as e:
        O.dcop(fl)
    assert e.value.code == 4
    ox, oit, _ = O.dcop(nn)
    assert np.all(np.isnan(ox))
    chain = W.chain(64)[0].to_c()
    cx, cit, _ = O.dcop(chain)
    b = Batch([Group(fl), Group(chain), Group(nn)]).op()
    assert b.status(0)[0][0] == 4
    st, it = b.status(1)
    assert st[0] == 0 and it[0] == cit and rel_err(b.x(1)[0], cx) < 1e-9
    st, it = b.status(2)
    assert st[0] == 0 and it[0] == oit and np.all(np.isnan(b.x(2)[0]))


def test_closed_form_charge_circuits_match_oracle():
    """The circuits whose oracle waveforms are pinned by closed forms (tests/test_oracle_charges.py): the
    Meyer gate-charge transient in both directions, the diode junction charge with and without transit
    time, and the two AC gate-capacitance stages, on the engine against the oracle."""
    from tests import test_oracle_charges as P

    for vin in (1.0, -1.0):
        cn = P.gate_rc(vin=vin, g=1e-4, w=1e-6, l=1e-6, tox=1e-8, cov=2e-10).to_c()
        log("meyer gate rc", vin)
        t, X, names, _ = O.tran(cn, 1e-12, 4e-10, [("g", 0.0)])
        b = Batch(Group(cn)).tran(1e-12, 4e-10, [("g", 0.0)])
        assert b.status()[0][0] == 0 and np.array_equal(b.axis(), t)
        err = float(np.abs(b.wave() - X).max())
        log("  max |dx|", err)
        assert err <= 1e-9
    for cj0, tt in ((1e-12, 0.0), (1e-12, 0.3)):
        cn = P.diode_rc(cj0, tt).to_c()
        log("diode rc", cj0, tt)
        t, X, names, _ = O.tran(cn, 1e-11, 3e-9, [("a", 0.0)])
        b = Batch(Group(cn)).tran(1e-11, 3e-9, [("a", 0.0)])
        assert b.status()[0][0] == 0 and np.array_equal(b.axis(), t)
        err = float(np.abs(b.wave() - X).max())
        log("  max |dx|", err)
        assert err <= 1e-9
    from spice21_amd.netlist import Netlist

    n = Netlist()
    n.v("vin", 1.0, "in", "", acm=1.0)
    n.r(1e-4, "in", "g")
    n.v("vdd", 1.0, "vdd", "")
    n.r(1e-4, "vdd", "d")
    nm = n.mos1_model(vt0=0.4, kp=100e-6, tox=1e-8)
    n.mos1("m", "d", "g", "", "", nm, n.mos1_inst(w=10e-6, l=1e-6))
    cn = n.to_c()
    log("ac triode stage")
    f, X, _ = O.ac(cn, 1000, 100_000_000_000, 40)
    b = Batch(Group(cn)).ac(1000, 100_000_000_000, 40)
    assert b.status()[0][0] == 0
    W_ = b.wave()
    err = float((np.abs(W_ - X) / np.maximum(np.abs(X), 1e-12)).max())
    log("  max relative |dx|", err)
    assert err <= 1e-9


@pytest.mark.parametrize("jit", [-1, 1])
def test_profiling_kernels_count_phases_and_leave_results_unchanged(jit):
    """Profiling launches run kernels with the phase-cycle counters compiled in (the normal kernels carry
    none): the same solutions and iteration counts, bit for bit, and per instance non-negative phase
    cycles whose sum stays within the instance's total (the eighth counter)."""
    nl, over = W.chain_mc(n_inst=64)
    cn = nl.to_c()
    for precision in (FP64, MIXED):
        ref = Batch(Group(cn, 64, over), precision=precision).set_jit(jit).op()
        b = Batch(Group(cn, 64, over), precision=precision).set_jit(jit).set_profiling(True).op()
        assert np.array_equal(ref.x(), b.x()) and np.array_equal(ref.status()[1], b.status()[1])
        pr = b.profile()
        log("profile", precision, jit, pr[0])
        assert np.all(pr >= 0) and np.all(pr[:, 7] > 0)
        assert np.all(pr[:, :7].sum(axis=1) <= pr[:, 7])
        assert np.all(pr[:, 0] > 0) and np.all(pr[:, 3] > 0)  # device evaluation, factorisation


def test_tuning_knobs_keep_results_bit_identical():
    """Wavefronts per workgroup and instances per wavefront change only the time: at one hub policy and
    factorisation kind the 64-instance chain's operating points are bit-identical across them.  The hub
    policy changes the order in which a shared slot's contributions are added (policy 1 gathers every
    shared slot in component order, the reference's order; policy 2 adds those of few contributors in
    colour order) and the generated factorisation rounds like the record-driven one only to 1e-12
    (test_generated_factorisation_matches_level_schedule), so across those the solutions agree to
    rounding.  The autotuner leaves the batch at one of the configurations it timed."""
    nl, over = W.chain_mc(n_inst=64)
    cn = nl.to_c()
    v = np.asarray(nl.kinds) == 0
    for precision in (FP64, MIXED):
        refs = {}
        for ipw, waves, hub, jit in [(1, 1, 1, -1), (4, 2, 1, -1), (2, 4, 1, 1), (16, 1, 1, 1), (4, 2, 2, -1),
                                     (16, 1, 2, -1), (8, 8, 2, 1), (2, 1, 2, 1)]:
            b = Batch(Group(cn, 64, over), precision=precision)
            b.set_ipw(ipw).set_waves(waves).set_hub(hub).set_jit(jit).op()
            x, (st, it) = b.x(), b.status()
            assert np.all(st == 0)
            log("knobs", precision, (ipw, waves, hub, jit), "config", b.config())
            if (hub, jit) not in refs:
                refs[(hub, jit)] = (x, it)
            else:
                k = refs[(hub, jit)]
                assert np.array_equal(x, k[0]) and np.array_equal(it, k[1]), (ipw, waves, hub, jit)
        for k, (x, _) in refs.items():
            d = np.abs(x - refs[(1, -1)][0])[:, v].max()
            log("hub, jit", k, "differs from (1, -1) by", d)
            assert d <= (1e-12 if precision == FP64 else 1e-6)
        b = Batch(Group(cn, 64, over), precision=precision)
        times = b.autotune(lambda bb: bb.op(), repeats=1)
        log("autotune", precision, b.tuned, times)
        assert ",".join(f"{k}={b.tuned[k]}" for k in ("ipw", "waves", "jit", "hub")) in times
        assert np.abs(b.x() - refs[(1, -1)][0])[:, v].max() <= (1e-12 if precision == FP64 else 1e-6)
