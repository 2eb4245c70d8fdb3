"""The reference's Python-binding tests (spice21py/spice21py/tests/test_spice21.py), run through the MI355X
engine's byte ABI, with the oracle as the checker where the reference's own test leaves values unchecked.
Needs the GPU."""
import json

import numpy as np
import pytest

from oracle import oracle_py as O
from spice21_amd import S21Error, api, protos
from spice21_amd.api import circuit
from spice21_amd.elab import elaborate
from spice21_amd.protos import Capacitor, Diode, Isrc, Resistor, Vsrc
from tests.circuits import py_ring3

pytestmark = pytest.mark.gpu


def test_dcop1():  # test_spice21.py:21-30
    ckt = circuit([Resistor(p="1", g=1e-3), Capacitor(p="1"), Isrc(p="1", dc=1e-3)])
    res = api.dcop(protos.Op(ckt=ckt))
    assert isinstance(res, dict)
    assert res["1"] == 1.0


def test_dcop2_missing_diode_parameters():
    """test_spice21.py:33-39 is marked xfail in the reference: a Diode with no model or parameter names
    fails to resolve.  The engine fails the same way, with the reference's error text."""
    for v in (5, 6, 7, 8):
        with pytest.raises(S21Error, match="Parameters not defined"):
            api.dcop(circuit(Vsrc(name="vb", dc=0.1 * v, p="a"), Diode(p="a")))


def test_dcop2_with_parameters_matches_oracle():
    """The same diode bias points with model and parameters defined, against the oracle."""
    for v in (5, 6, 7, 8):
        ckt = circuit(protos.DiodeModel(name="dm"), protos.DiodeInstParams(name="dp"),
                      Vsrc(name="vb", dc=0.1 * v, p="a"), Diode(name="d1", p="a", model="dm", params="dp"))
        res = api.dcop(ckt)
        e = elaborate(ckt)
        x, _, _ = O.dcop(e)
        assert res["a"] == 0.1 * v
        ib = x[e.var_names.index("vb")]
        assert abs(res["vb"] - ib) <= 1e-9 * abs(ib), (v, res["vb"], ib)


def test_tran1():  # test_spice21.py:42-50
    c = circuit([Resistor(p="1", g=1e-3), Capacitor(p="1"), Isrc(p="1", dc=1e-3)])
    res = api.tran(c)
    assert isinstance(res, dict)
    assert isinstance(res["1"], list)
    for v in res["1"]:
        assert v == 1.0
    res = api.tran(c, args=protos.TranOptions(tstep=1e-9, tstop=1e-6))
    assert len(res["1"]) == O.tran_point_count(1e-9, 1e-6)
    assert all(v == 1.0 for v in res["1"])


def test_ac1(tmp_path, monkeypatch):  # test_spice21.py:53-70
    monkeypatch.chdir(tmp_path)  # the AC analysis writes its result files into the working directory
    c = circuit(Resistor(p="inp", n="out", g=1e-3), Capacitor(p="out", c=1e-9), Vsrc(p="inp", dc=1e-3, acm=1))
    for args in (None, protos.AcOptions(fstart=1, fstop=1_000_000_000, npts=90)):
        res = api.ac(c, args=args)
        assert isinstance(res, dict)
        assert isinstance(res["inp"], list)
        assert isinstance(res["out"], list)
        for v in res["inp"]:
            assert isinstance(v, complex)
            assert v == complex(1.0, 0.0)
        for k in range(1, len(res["out"])):  # low-pass: the output magnitude falls
            assert abs(res["out"][k]) <= abs(res["out"][k - 1])
    assert len(res["out"]) == 91


def test_ac_result_files(tmp_path, monkeypatch):
    """The reference's `ac` writes stream.ac.json (per frequency [f, re0, im0, ...] in variable order) and
    ac.json (signals, freq, data as [re, im] pairs, and an empty map: it is serialised before the map is
    filled) into the working directory (analysis.rs:783-827); both hold the same numbers as the result."""
    monkeypatch.chdir(tmp_path)
    c = circuit(Resistor(p="inp", n="out", g=1e-3), Capacitor(p="out", c=1e-9), Vsrc(name="v1", p="inp", acm=1))
    res = api.ac(c, args=protos.AcOptions(fstart=10, fstop=100_000_000, npts=35))
    stream = json.load(open(tmp_path / "stream.ac.json"))
    full = json.load(open(tmp_path / "ac.json"))
    assert set(full) == {"signals", "freq", "data", "map"} and full["map"] == {}
    names = full["signals"]
    assert sorted(names) == sorted(res)
    assert len(stream) == len(full["freq"]) == len(full["data"]) == len(res["out"]) == 36
    for p, (row, f, d) in enumerate(zip(stream, full["freq"], full["data"])):
        assert row[0] == f and len(row) == 1 + 2 * len(names) and len(d) == len(names)
        for k, nm in enumerate(names):
            z = res[nm][p]
            assert row[1 + 2 * k] == z.real and row[2 + 2 * k] == z.imag  # shortest round-trip text
            assert d[k] == [z.real, z.imag]
    assert full["freq"][0] == 10.0 and full["freq"][-1] == 100_000_000.0
    # the RC low-pass's closed form at every frequency
    w = 2 * np.pi * np.array(full["freq"])
    out = np.array([complex(*d[names.index("out")]) for d in full["data"]])
    assert np.abs(out - 1.0 / (1.0 + 1j * w * 1e-9 / 1e-3)).max() <= 1e-9


def test_mos1_ro_matches_oracle():
    """test_spice21.py:153-174 runs this three-stage ring and checks only the result type; here every
    waveform is compared with the oracle's, point by point."""
    ckt = py_ring3()
    args = protos.TranOptions(tstop=1e-8, tstep=1e-11)
    args.ic["ring0"] = 0.0
    res = api.tran(ckt=ckt, args=args)
    assert isinstance(res, dict)
    t, X, names, _ = O.tran(elaborate(ckt), 1e-11, 1e-8, [("ring0", 0.0)])
    assert len(res["ring1"]) == len(t)
    worst = 0.0
    for k, nm in enumerate(names):
        if nm in res:
            worst = max(worst, float(np.max(np.abs(np.array(res[nm]) - X[:, k]))))
    print("mos1 ring vs oracle: max |dx|", worst, flush=True)
    assert worst <= 1e-9


@pytest.mark.parametrize("size,why", [(24, "16-bit index"), (30, "slot index")])
def test_circuit_too_large_for_the_batch_runs_on_the_single_instance_path(size, why):
    """A power grid whose matrix does not fit the batched kernel (record offsets beyond 16 bits at
    24 x 24, more slots than its slot codes address at 30 x 30) is refused by the batch and solved by the
    byte ABI on the single-instance path: node voltages against the oracle to 1e-9 relative, branch
    currents to 1e-12 A (sums of ~1 A mesh currents)."""
    from spice21_amd import _abi as A
    from spice21_amd import workloads as W
    from spice21_amd.batch import Batch, Group
    from tests.test_elab_scale import power_grid_proto

    kw = dict(size=size, n_inv=size * size // 8, n_diode=size * size // 20, pad_every=4)
    nl = W.power_grid(**kw)
    cn = nl.to_c()
    with pytest.raises(S21Error, match="too large for the batched kernel.*" + why):
        Batch(Group(cn)).op()
    print("grid", size, "unknowns", cn.n_vars, flush=True)
    res = api.dcop(protos.Op(ckt=power_grid_proto(**kw)))
    xo, _, _ = O.dcop(cn)
    v = np.asarray(nl.kinds) == A.S21_VAR_V
    got = np.array([res[nm] for nm in nl.names])
    rel = np.abs(got - xo)[v] / np.maximum(np.abs(xo[v]), 1e-9)
    cur = np.abs(got - xo)[~v]
    print("  max relative node difference", rel.max(), "max branch current difference", cur.max(), flush=True)
    assert rel.max() < 1e-9 and cur.max() <= 1e-12


def test_transient_too_large_for_the_batch_runs_on_the_single_instance_path():
    """The transient byte call on the 24 x 24 grid (refused by the batched kernel): the single-instance
    path's waveform against the oracle's, point for point."""
    from spice21_amd import _abi as A
    from spice21_amd import workloads as W
    from tests.test_elab_scale import power_grid_proto
    from tests.test_gpu_parity import wave_within

    kw = dict(size=24, n_inv=72, n_diode=28, pad_every=4)
    nl = W.power_grid(**kw)
    args = protos.TranOptions(tstep=1e-12, tstop=4e-12)
    args.ic["g0_0"] = 0.0
    res = api.tran(protos.Tran(ckt=power_grid_proto(**kw), args=args))
    t, X, names, _ = O.tran(nl.to_c(), 1e-12, 4e-12, [("g0_0", 0.0)])
    assert np.array_equal(np.array(res["time"]), t)
    got = np.stack([np.array(res[nm]) for nm in names], axis=1)
    assert len(names) == len(nl.kinds) + 2  # the initial condition's source and its current
    assert wave_within(got, X, names, list(nl.kinds) + [A.S21_VAR_V, A.S21_VAR_I], "grid tran bytes")
