"""Circuits of the reference's own tests (spice21/src/tests.rs), as flattened netlists."""
from spice21_amd.netlist import Netlist


def mos1_defaults(n):
    """add_mos1_defaults (tests.rs:1449-1461): default NMOS, PMOS and instance parameters."""
    nmos = n.mos1_model()
    pmos = n.mos1_model(mos_type=1)
    inst = n.mos1_inst()
    return nmos, pmos, inst


def bsim4_defaults(n):
    """add_bsim4_defaults (tests.rs:1463-1474): default BSIM4 NMOS and PMOS models and instance parameters."""
    return n.bsim4_model(0), n.bsim4_model(1), n.bsim4_inst()


def ro3(kind):
    """cmos_ro3 / nmos_ro3 / pmos_ro3 (tests.rs:889-912, 966-990, 1009-1033) flattened by hand: module
    instances x1..x3 map inp/out onto "1"->"2", "2"->"3", "3"->"1"; vss is ground.  Kinds "cmos", "nmos",
    "pmos" use the Mos1 defaults, "bsim4_cmos", "bsim4_nmos", "bsim4_pmos" the BSIM4 ones."""
    n = Netlist(signals=["1", "2", "3", "vdd"])
    stages = [("1", "2"), ("2", "3"), ("3", "1")]
    if kind.startswith("bsim4_"):
        nm, pm, ip = bsim4_defaults(n)
        dev, kind = n.bsim4, kind[len("bsim4_"):]
    else:
        dev = n.mos1
    if kind == "mos0":
        n.v("v1", 1.0, "vdd", "")
        for x, (a, b) in enumerate(stages):
            n.mos0(f"x{x+1}.p", b, a, "vdd", "vdd", 1)
            n.mos0(f"x{x+1}.n", b, a, "", "", 0)
            n.c(1e-15, b, "")
        return n
    if dev == n.mos1:
        nm, pm, ip = mos1_defaults(n)
    n.v("v1", 1.0, "vdd", "")
    for x, (a, b) in enumerate(stages):
        if kind == "cmos":
            dev(f"x{x+1}.p", b, a, "vdd", "vdd", pm, ip)
            dev(f"x{x+1}.n", b, a, "", "", nm, ip)
            n.c(1e-15, b, "")
        elif kind == "nmos":
            dev(f"x{x+1}.m", b, a, "", "", nm, ip)
            n.r(1e-6, b, "vdd")
            n.c(0.5e-15, b, "vdd")
        elif kind == "pmos":
            n.c(1e-16, b, "")
            n.r(1e-6, b, "")
            dev(f"x{x+1}.m", b, a, "vdd", "vdd", pm, ip)
        else:
            raise ValueError(kind)
    return n


GOLDEN_RUNS = {
    # fixture -> (circuit kind, tstep, tstop), tests.rs:773-790, 931-947, 992-1008, 1035-1051
    "test_mos1_cmos_ro_tran.json": ("cmos", 1e-11, 1e-8),
    "test_mos1_nmos_ro_tran.json": ("nmos", 1e-11, 1e-8),
    "test_mos1_pmos_ro_tran.json": ("pmos", 1e-11, 1e-8),
    "test_mos0_cmos_ro_tran.json": ("mos0", 1e-15, 1e-12),
}
BSIM4_GOLDEN_RUNS = {
    # tests.rs:948-965, 1053-1071, 1360-1377
    "test_bsim4_cmos_ro_tran.json": ("bsim4_cmos", 1e-10, 3e-7),
    "test_bsim4_pmos_ro_tran.json": ("bsim4_pmos", 1e-11, 1e-8),
    "test_bsim4_nmos_ro_tran.json": ("bsim4_nmos", 1e-9, 1e-6),
}



def materialize(nl, overrides, i):
    """Standalone netlist of instance i of a batch (see Netlist.instance)."""
    return nl.instance(overrides, i)


def py_ring3(acm=0.0):
    """The three-stage Mos1 ring of spice21py's test_mos1_ro (spice21py/spice21py/tests/test_spice21.py:
    132-174) as a protobuf circuit: vss is a node of its own (driven by vvss), every stage's capacitor ends
    on it, and the feedback source vvfb closes the loop from ring0 to ring3 (AC magnitude ``acm``)."""
    from spice21_amd import protos as P
    from spice21_amd.api import circuit

    comps = []
    for k in range(3):
        inp, out = f"ring{k}", f"ring{k + 1}"
        p = P.Mos(name=f"stg{k}p", model="pmos", params="default")
        p.ports.g, p.ports.d, p.ports.s, p.ports.b = inp, out, "vdd", "vdd"
        n = P.Mos(name=f"stg{k}n", model="nmos", params="default")
        n.ports.g, n.ports.d, n.ports.s, n.ports.b = inp, out, "vss", "vss"
        comps += [p, n, P.Capacitor(name=f"stg{k}c", p=out, n="vss", c=1e-15)]
    return circuit(
        P.Mos1Model(name="nmos", mos_type=P.MosType.NMOS),
        P.Mos1Model(name="pmos", mos_type=P.MosType.PMOS),
        P.Mos1InstParams(name="default"),
        P.Vsrc(name="vvdd", p="vdd", n="vss", dc=1.0),
        P.Vsrc(name="vvss", p="vss", dc=0),
        P.Vsrc(name="vvfb", p="ring0", n="ring3", dc=0, acm=acm),
        *comps,
    )


def ring_caps(stages=5, tox=1e-8, overlap=2e-10, load=10e-15):
    """The 5-stage ring of the single-ring workload with gate capacitance: both Mos1 models set tox (Meyer
    gate charge, comps/mos.rs:725-806) and the gate-source/drain/bulk overlap capacitances, and each stage
    node has a 10 fF load to ground."""
    n = Netlist(signals=[str(k + 1) for k in range(stages)] + ["vdd"])
    caps = dict(tox=tox, cgso=overlap, cgdo=overlap, cgbo=overlap)
    nm = n.mos1_model(vt0=0.4, kp=100e-6, **caps)
    pm = n.mos1_model(mos_type=1, vt0=0.4, kp=50e-6, **caps)
    ni = n.mos1_inst(w=1e-6, l=0.1e-6)
    pi = n.mos1_inst(w=2e-6, l=0.1e-6)
    n.v("vdd", 1.0, "vdd", "")
    for k in range(stages):
        a, b = str(k + 1), str((k + 1) % stages + 1)
        n.mos1(f"x{k}.p", b, a, "vdd", "vdd", pm, pi)
        n.mos1(f"x{k}.n", b, a, "", "", nm, ni)
        n.c(load, b, "")
    return n


def cs_amp_caps(n_stages=2, tox=1e-8, overlap=2e-10):
    """A common-source amplifier chain with gate capacitance (Mos1 tox and overlaps), resistive loads from
    a 1 V supply and an AC-driven 0.6 V input bias: Mos1::load_ac's capacitive stamps in the AC sweep."""
    n = Netlist()
    nm = n.mos1_model(vt0=0.4, kp=100e-6, tox=tox, cgso=overlap, cgdo=overlap, cgbo=overlap)
    ni = n.mos1_inst(w=1e-6, l=0.1e-6)
    n.v("vin", 0.6, "in", "", acm=1.0)
    n.v("vdd", 1.0, "vdd", "")
    prev = "in"
    for k in range(n_stages):
        out = f"o{k + 1}"
        n.r(1.0 / 5e3, "vdd", out)
        n.mos1(f"m{k + 1}", out, prev, "", "", nm, ni)
        n.c(20e-15, out, "")
        prev = out
    return n


def diode_rc(cj0=1e-12, tt=0.0, forward=False):
    """A 1 V source charging node a through 1 kohm against 1 pF and a diode with junction capacitance cj0
    (and transit time tt), reverse-biased by default (cathode on a) or forward-biased (anode on a)."""
    n = Netlist()
    n.v("v1", 1.0, "in", "")
    n.r(1e-3, "in", "a")
    dm = n.diode_model(cj0=cj0, tt=tt)
    if forward:
        n.diode("d1", "a", "", dm, n.diode_inst())
    else:
        n.diode("d1", "", "a", dm, n.diode_inst())
    n.c(1e-12, "a", "")
    return n


def random_ac_circuit(seed):
    """A random netlist of the components AC supports (R, C, Mos1 with and without terminal resistances
    and gate capacitances, one or two sources with AC magnitudes), every node tied to ground by 1e-6 S."""
    import numpy as np

    rng = np.random.default_rng(1000 + seed)
    n = Netlist()
    nodes = [f"n{k}" for k in range(int(rng.integers(3, 12)))]
    for nd in nodes:
        n.addv(nd)
    nm = n.mos1_model(vt0=0.4, kp=100e-6, lambda_=0.02, tox=1e-8, cgso=2e-10, cgdo=2e-10)
    pm = n.mos1_model(mos_type=1, vt0=-0.4, kp=50e-6)
    rm = n.mos1_model(vt0=0.3, rd=50.0, rs=20.0, tox=2e-8)
    ip = n.mos1_inst(w=1e-6, l=0.1e-6)
    pick = lambda: nodes[int(rng.integers(len(nodes)))] if rng.random() > 0.15 else ""
    n.v("vs", float(rng.uniform(0.5, 1.5)), nodes[0], "", acm=1.0)
    if rng.random() < 0.5:
        n.v("va", float(rng.uniform(0.0, 0.5)), nodes[1], "", acm=float(rng.uniform(0.1, 2.0)))
    for k in range(int(rng.integers(4, 20))):
        c = rng.random()
        if c < 0.35:
            n.r(float(rng.uniform(1e-4, 1e-2)), pick(), pick())
        elif c < 0.6:
            n.c(float(rng.uniform(1e-15, 1e-12)), pick(), pick())
        else:
            m = [nm, pm, rm][int(rng.integers(3))]
            n.mos1(f"m{k}", pick(), pick(), pick(), pick(), m, ip)
    for nd in nodes:  # keep every node connected
        n.r(1e-6, nd, "")
    return n


def switching_diode(I=3.0, G=0.1, C=1e-7, isat=1e-3):
    """A diode from a 0 V source node b to an RC node a that a current source charges toward I/G.  At the
    first Newton iteration (x = 0) the diode's conductance isat/vte passes the reference's pivot threshold,
    so the first search pivots on it; as a charges, the diode turns off and its conductance falls below
    1e-3 of the source's unit entry in its column (a above ~0.1 V), below 1e-6 of it (a above ~0.3 V) and,
    with gmin = 0 and no junction charge, to exactly zero (exp underflow, a above ~19.3 V).  The
    reference's later searches then pivot on the source's entries instead."""
    from spice21_amd.netlist import Netlist

    n = Netlist()
    n.v("vb", 0.0, "b", "")
    n.r(G, "a", "")
    n.c(C, "a", "")
    n.i(I, "a", "")
    dm = n.diode_model(**{"is": isat})
    di = n.diode_inst()
    n.diode("d1", "b", "a", dm, di)
    return n


def no_gmin():
    from spice21_amd import _abi as A

    o = A.s21_options.default()
    o.gmin = 0.0
    return o
