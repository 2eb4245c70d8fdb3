"""Elaboration at the single large instance's size: the 200 x 200 power grid (45,040 unknowns, 131,640
components) as a protobuf Circuit through the engine's decoder and elaborator (elab.rs:35-265, with
hashed name lookup in place of the reference's linear find_or_create, analysis.rs:92-111), against the
netlist the workload builder makes directly: the same variables in the same order and the same components
with the same nodes, values and parameters."""
import time

import numpy as np

from spice21_amd import _abi as A
from spice21_amd import protos as P
from spice21_amd import workloads as W
from spice21_amd.elab import elaborate


def power_grid_proto(size=200, n_inv=5000, n_diode=2000, pad_every=20, seed=0xC4):
    """W.power_grid as a Circuit message: the same draws, names and element order."""
    c = P.Circuit(name="power_grid")
    g = lambda r, k: f"g{r}_{k}"
    nedge = 2 * size * (size - 1)
    u = W.uniform(seed, nedge + 2 * n_inv + n_diode)
    r_edge = 0.5 + 1.5 * u[:nedge]
    pick = (u[nedge:] * (size * size)).astype(np.int64)

    def add(**kw):
        c.comps.append(P.Instance(**kw))

    k = 0
    for r in range(size):
        for q in range(size - 1):
            add(r=P.Resistor(name=f"rh{r}_{q}", p=g(r, q), n=g(r, q + 1), g=1.0 / r_edge[k]))
            k += 1
    for r in range(size - 1):
        for q in range(size):
            add(r=P.Resistor(name=f"rv{r}_{q}", p=g(r, q), n=g(r + 1, q), g=1.0 / r_edge[k]))
            k += 1
    for r in range(size):
        for q in range(size):
            add(c=P.Capacitor(name=f"c{r}_{q}", p=g(r, q), n="", c=1e-12))
    walk = ([(0, q) for q in range(size)] + [(r, size - 1) for r in range(1, size)] +
            [(size - 1, q) for q in range(size - 2, -1, -1)] + [(r, 0) for r in range(size - 2, 0, -1)])
    for j, (r, q) in enumerate(walk[::pad_every]):
        add(v=P.Vsrc(name=f"vpad{j}", p=g(r, q), n="", dc=1.0))
    for name, typ, prm in (("nmos", P.MosType.NMOS, W.NMOS), ("pmos", P.MosType.PMOS, W.PMOS)):
        m = P.Mos1Model(name=name, mos_type=typ)
        m.vt0.value, m.kp.value = prm["vt0"], prm["kp"]
        c.defs.append(P.Def(mos1model=m))
    for name, prm in (("ninst", W.NINST), ("pinst", W.PINST)):
        i = P.Mos1InstParams(name=name)
        i.w.value, i.l.value = prm["w"], prm["l"]
        c.defs.append(P.Def(mos1inst=i))
    dm = P.DiodeModel(name="dmod")
    getattr(dm, "is").value = 1e-14
    dm.n.value = 1.0
    c.defs.append(P.Def(diodemodel=dm))
    c.defs.append(P.Def(diodeinst=P.DiodeInstParams(name="dinst")))
    node = lambda p: g(int(p) // size, int(p) % size)
    for j in range(n_inv):
        a, s = node(pick[2 * j]), node(pick[2 * j + 1])
        mp = P.Mos(name=f"inv{j}.p", model="pmos", params="pinst")
        mp.ports.d, mp.ports.g, mp.ports.s, mp.ports.b = f"o{j}", a, s, s
        mn = P.Mos(name=f"inv{j}.n", model="nmos", params="ninst")
        mn.ports.d, mn.ports.g, mn.ports.s, mn.ports.b = f"o{j}", a, "", ""
        add(m=mp)
        add(m=mn)
    for j in range(n_diode):
        add(d=P.Diode(name=f"d{j}", p="", n=node(pick[2 * n_inv + j]), model="dmod", params="dinst"))
    return c


def test_power_grid_elaborates_like_the_direct_netlist():
    ckt = power_grid_proto()
    t0 = time.perf_counter()
    buf = ckt.SerializeToString()
    t1 = time.perf_counter()
    e = elaborate(buf)
    t2 = time.perf_counter()
    print(f"encode {t1 - t0:.2f} s ({len(buf) / 1e6:.1f} MB), decode + elaborate {t2 - t1:.2f} s", flush=True)
    assert t2 - t1 < 30.0  # the reference's linear name lookup is quadratic here
    ref = W.power_grid()
    cn = ref.to_c()
    assert e.n_vars == cn.n_vars == 45040
    assert e.var_names == list(cn.var_names)
    assert e.var_kinds == list(ref.kinds)
    assert e.net.n_comps == cn.net.n_comps == len(ref.comps) == 131640
    for k in range(e.net.n_comps):
        a, b = e.net.comps[k], cn.net.comps[k]
        assert a.type == b.type and list(a.node) == list(b.node) and a.value == b.value, k
    # each device's model and instance parameters reach the derivation identically (values and which are
    # set; the elaborator numbers parameter sets in order of first use, the builder in order of definition)
    def same(pa, pb, type_slot):
        mask = ~(1 << type_slot) if type_slot is not None and pa.v[type_slot] == 0 else ~0
        return pa.set & mask == pb.set & mask and list(pa.v) == list(pb.v)

    seen = set()
    for k in range(e.net.n_comps):
        a, b = e.net.comps[k], cn.net.comps[k]
        if a.type not in (A.S21_COMP_MOS1, A.S21_COMP_DIODE) or (a.type, a.model, a.inst) in seen:
            continue
        seen.add((a.type, a.model, a.inst))
        if a.type == A.S21_COMP_MOS1:
            assert same(e.net.mos1_models[a.model], cn.net.mos1_models[b.model], 2), k
            assert same(e.net.mos1_insts[a.inst], cn.net.mos1_insts[b.inst], None), k
        else:
            assert same(e.net.diode_models[a.model], cn.net.diode_models[b.model], None), k
            assert same(e.net.diode_insts[a.inst], cn.net.diode_insts[b.inst], None), k
    assert len(seen) == 3
