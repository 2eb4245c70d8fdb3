"""Synthetic benchmark circuits (flattened netlists plus per-instance parameter overrides).

Random draws use splitmix64 -> 53-bit uniform -> Box-Muller, seeded per workload, so every consumer
(tests, bench, CPU baseline) sees the same instances.

  ring5()      5-stage CMOS ring oscillator, 10 fF per stage, VDD = 1 V (single instance)
  chain_mc()   N instances of a 64-stage CMOS inverter chain, per-device VTO ~ N(base, 15 mV) and
               W ~ N(base, 3 %), input tied to 0.5 V
  ring_batch() ring oscillators of uneven size: k = 3 + 2 U{0..49} stages, C ~ U[5, 50] fF, VDD ~ U[0.8, 1.2]
  amp_ladder() N instances of a two-stage common-source amplifier driving a 20-section RC ladder with
               R ~ U[1k, 10k] and C ~ U[10f, 100f] per section and instance (AC sweep workload)
  bsim4_chain_mc() N instances of a CMOS inverter chain of BSIM4 devices (default model, W 1 / 2 um,
               L 0.1 um), per-device DELVTO ~ N(0, 15 mV) and W ~ N(base, 3 %), input tied to 0.5 V
"""
import numpy as np

from .netlist import Netlist

MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed, n):
    """n successive outputs of splitmix64 started at `seed`."""
    with np.errstate(over="ignore"):
        state = np.uint64(seed) + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        z = state
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n):
    """n uniforms in [0, 1) with 53-bit resolution."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, n):
    """n standard normals by Box-Muller over pairs of uniforms."""
    m = (n + 1) // 2
    u = uniform(seed, 2 * m)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    return z[:n]


# Mos1 parameters of the benchmark circuits
NMOS = dict(vt0=0.4, kp=100e-6)
PMOS = dict(mos_type=1, vt0=-0.4, kp=50e-6)
NINST = dict(w=1e-6, l=0.1e-6)
PINST = dict(w=2e-6, l=0.1e-6)


def _models(n, pmos=None):
    nm = n.mos1_model(**NMOS)
    pm = n.mos1_model(**(pmos or PMOS))
    ni = n.mos1_inst(**NINST)
    pi = n.mos1_inst(**PINST)
    return nm, pm, ni, pi


def ring5(stages=5, cap=10e-15, vdd=1.0, pmos=None):
    """CMOS ring oscillator: stage k drives node str(k+1) from node str(k); the last stage closes the ring
    on node "1".  Returns the netlist; the transient kicks node "1" through an initial condition."""
    n = Netlist(signals=[str(k + 1) for k in range(stages)] + ["vdd"])
    nm, pm, ni, pi = _models(n, pmos)
    n.v("vdd", vdd, "vdd", "")
    for k in range(stages):
        a, b = str(k + 1), str((k + 1) % stages + 1)
        n.mos1(f"x{k}.p", b, a, "vdd", "vdd", pm, pi)
        n.mos1(f"x{k}.n", b, a, "", "", nm, ni)
        n.c(cap, b, "", name=f"x{k}.c")
    return n


def chain(stages=64, vin=0.5, vdd=1.0):
    """Inverter chain: input Vsrc on "in", stage k drives "s{k}" from "s{k-1}" ("s0" = "in")."""
    n = Netlist()
    nm, pm, ni, pi = _models(n)
    n.v("vin", vin, "in", "")
    n.v("vdd", vdd, "vdd", "")
    mos = []
    for k in range(1, stages + 1):
        a = "in" if k == 1 else f"s{k-1}"
        b = f"s{k}"
        mos.append((n.mos1(f"p{k}", b, a, "vdd", "vdd", pm, pi), "p"))
        mos.append((n.mos1(f"n{k}", b, a, "", "", nm, ni), "n"))
    return n, mos


def chain_mc(n_inst=4096, stages=64, seed=0xC2, sigma_vt=0.015, sigma_w=0.03):
    """Monte-Carlo inverter chains: returns (netlist, overrides) where overrides set every device's own
    vt0 (model field) and w (instance field) per instance."""
    n, mos = chain(stages)
    nd = len(mos)
    z = normal(seed, 2 * nd * n_inst).reshape(n_inst, nd, 2)
    over = []
    for j, (ci, kind) in enumerate(mos):
        base_vt = NMOS["vt0"] if kind == "n" else PMOS["vt0"]
        base_w = NINST["w"] if kind == "n" else PINST["w"]
        over.append((ci, "model", "vt0", base_vt + sigma_vt * z[:, j, 0]))
        over.append((ci, "inst", "w", base_w * (1.0 + sigma_w * z[:, j, 1])))
    return n, over


def ring_batch(n_inst=1024, seed=0xC3):
    """Uneven ring oscillators: per instance (stages, cap, vdd)."""
    u = uniform(seed, 3 * n_inst).reshape(n_inst, 3)
    stages = 3 + 2 * np.minimum((u[:, 0] * 50).astype(int), 49)
    cap = 5e-15 + u[:, 1] * 45e-15
    vdd = 0.8 + u[:, 2] * 0.4
    return stages, cap, vdd


def ring_groups(n_inst=1024, seed=0xC3):
    """The uneven ring-oscillator batch as batch groups, one per stage count: a list of (netlist,
    n_inst_k, overrides, instance ids) with every stage capacitor's value and the supply's value per
    instance (node "1" is unknown 0 in every group, so one initial condition applies to all)."""
    stages, cap, vdd = ring_batch(n_inst, seed)
    groups = []
    for k in sorted(set(int(v) for v in stages)):
        ids = np.nonzero(stages == k)[0]
        n = ring5(stages=k)
        caps = [ci for ci, c in enumerate(n.comps) if c[0] == 1]  # S21_COMP_C
        vsrc = [ci for ci, c in enumerate(n.comps) if c[0] == 3]  # S21_COMP_V
        over = [(ci, "value", 0, cap[ids]) for ci in caps] + [(vsrc[0], "value", 0, vdd[ids])]
        groups.append((n, len(ids), over, ids))
    return groups


def amp_ladder(n_inst=256, sections=20, seed=0xC5, vin=0.6, vdd=1.0, rd1=5e3, rd2=2e3):
    """Two-stage common-source amplifier (NMOS M1: g = in, d = n1, RD1 from vdd; M2: g = n1, d = n2, RD2
    from vdd) with an AC-driven input source (dc bias vin, acm = 1) and an RC ladder on n2.  Returns
    (netlist, overrides): every ladder resistor's conductance and capacitor's value per instance."""
    n = Netlist()
    nm = n.mos1_model(**NMOS)
    ni = n.mos1_inst(**NINST)
    n.v("vin", vin, "in", "", acm=1.0)
    n.v("vdd", vdd, "vdd", "")
    n.r(1.0 / rd1, "vdd", "n1", name="rd1")
    n.mos1("m1", "n1", "in", "", "", nm, ni)
    n.r(1.0 / rd2, "vdd", "n2", name="rd2")
    n.mos1("m2", "n2", "n1", "", "", nm, ni)
    u = uniform(seed, 2 * sections * n_inst).reshape(n_inst, sections, 2)
    over = []
    prev = "n2"
    for k in range(sections):
        node = f"l{k + 1}"
        ri = n.r(1.0 / 5.5e3, prev, node, name=f"rl{k + 1}")
        ci = n.c(55e-15, node, "", name=f"cl{k + 1}")
        over.append((ri, "value", 0, 1.0 / (1e3 + 9e3 * u[:, k, 0])))
        over.append((ci, "value", 0, 10e-15 + 90e-15 * u[:, k, 1]))
        prev = node
    return n, over


def power_grid(size=200, n_inv=5000, n_diode=2000, pad_every=20, seed=0xC4, device="mos1"):
    """One large instance: a size x size resistor mesh (edge R ~ U[0.5, 2] ohm, 1 pF from every mesh node
    to ground) with 1.0 V pads on every `pad_every`-th node of its boundary walk, `n_inv` CMOS inverters
    (input and PMOS source/bulk at random mesh nodes, NMOS to ground, each driving a new output node) and
    `n_diode` reverse-biased diodes (IS 1e-14, N 1; anode at ground, cathode at a random mesh node).
    device: "mos1" (the benchmark's Mos1 models) or "bsim4" (the default BSIM4 models, W 1 / 2 um,
    L 0.1 um)."""
    n = Netlist()
    g = lambda r, c: f"g{r}_{c}"
    nedge = 2 * size * (size - 1)
    u = uniform(seed, nedge + 2 * n_inv + n_diode)
    r_edge = 0.5 + 1.5 * u[:nedge]
    pick = (u[nedge:] * (size * size)).astype(np.int64)
    k = 0
    for r in range(size):
        for c in range(size - 1):
            n.r(1.0 / r_edge[k], g(r, c), g(r, c + 1), name=f"rh{r}_{c}")
            k += 1
    for r in range(size - 1):
        for c in range(size):
            n.r(1.0 / r_edge[k], g(r, c), g(r + 1, c), name=f"rv{r}_{c}")
            k += 1
    for r in range(size):
        for c in range(size):
            n.c(1e-12, g(r, c), "", name=f"c{r}_{c}")
    walk = ([(0, c) for c in range(size)] + [(r, size - 1) for r in range(1, size)] +
            [(size - 1, c) for c in range(size - 2, -1, -1)] + [(r, 0) for r in range(size - 2, 0, -1)])
    for j, (r, c) in enumerate(walk[::pad_every]):
        n.v(f"vpad{j}", 1.0, g(r, c), "")
    if device == "bsim4":
        nm, pm = n.bsim4_model(0), n.bsim4_model(1)
        ni, pi = n.bsim4_inst(w=1e-6, l=0.1e-6), n.bsim4_inst(w=2e-6, l=0.1e-6)
        dev = n.bsim4
    else:
        nm, pm, ni, pi = _models(n)
        dev = n.mos1
    node = lambda p: g(int(p) // size, int(p) % size)
    for j in range(n_inv):
        a, s = node(pick[2 * j]), node(pick[2 * j + 1])
        dev(f"inv{j}.p", f"o{j}", a, s, s, pm, pi)
        dev(f"inv{j}.n", f"o{j}", a, "", "", nm, ni)
    dm = n.diode_model(is_=1e-14, n=1.0)
    di = n.diode_inst()
    for j in range(n_diode):
        n.diode(f"d{j}", "", node(pick[2 * n_inv + j]), dm, di)
    return n


def bsim4_chain(stages=16, vin=0.5, vdd=1.0):
    """Inverter chain of BSIM4 devices (the default NMOS and PMOS models, W = 1 um / 2 um, L = 0.1 um)."""
    n = Netlist()
    nm, pm = n.bsim4_model(0), n.bsim4_model(1)
    ni, pi = n.bsim4_inst(w=1e-6, l=0.1e-6), n.bsim4_inst(w=2e-6, l=0.1e-6)
    n.v("vin", vin, "in", "")
    n.v("vdd", vdd, "vdd", "")
    mos = []
    for k in range(1, stages + 1):
        a = "in" if k == 1 else f"s{k-1}"
        b = f"s{k}"
        mos.append((n.bsim4(f"p{k}", b, a, "vdd", "vdd", pm, pi), "p"))
        mos.append((n.bsim4(f"n{k}", b, a, "", "", nm, ni), "n"))
    return n, mos


def bsim4_ring(stages=5, cap=1e-15, vdd=1.0):
    """Ring oscillator of BSIM4 inverters (the default NMOS and PMOS models, W = 1 um, L = 0.1 um) with a
    capacitor from every stage node to ground; the transient kicks node "1" through an initial condition
    (the ring's DC point alone is an unstable equilibrium)."""
    n = Netlist(signals=[str(k + 1) for k in range(stages)] + ["vdd"])
    nm, pm, ip = n.bsim4_model(0), n.bsim4_model(1), n.bsim4_inst(w=1e-6, l=0.1e-6)
    n.v("vdd", vdd, "vdd", "")
    for k in range(stages):
        a, b = str(k + 1), str((k + 1) % stages + 1)
        n.bsim4(f"x{k}.p", b, a, "vdd", "vdd", pm, ip)
        n.bsim4(f"x{k}.n", b, a, "", "", nm, ip)
        n.c(cap, b, "", name=f"x{k}.c")
    return n


def bsim4_ring_batch(n_inst=1024, stages=5, seed=0xC7):
    """Batch of BSIM4 rings whose stage capacitors are drawn per instance, U[1, 10] fF: (netlist,
    overrides)."""
    n = bsim4_ring(stages)
    cap = 1e-15 + 9e-15 * uniform(seed, n_inst)
    caps = [ci for ci, c in enumerate(n.comps) if c[0] == 1]  # S21_COMP_C
    return n, [(ci, "value", 0, cap) for ci in caps]


def bsim4_chain_mc(n_inst=1024, stages=16, seed=0xC6, sigma_vt=0.015, sigma_w=0.03):
    """Monte-Carlo BSIM4 inverter chains: (netlist, overrides) setting every device's own DELVTO and W
    (instance fields) per instance."""
    n, mos = bsim4_chain(stages)
    nd = len(mos)
    z = normal(seed, 2 * nd * n_inst).reshape(n_inst, nd, 2)
    over = []
    for j, (ci, kind) in enumerate(mos):
        base_w = 1e-6 if kind == "n" else 2e-6
        over.append((ci, "inst", "delvto", sigma_vt * z[:, j, 0]))
        over.append((ci, "inst", "w", base_w * (1.0 + sigma_w * z[:, j, 1])))
    return n, over
