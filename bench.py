#!/usr/bin/env python3
"""Throughput of the batched Newton inner loop on one MI355X.

Workload: 4096 instances of a 64-stage CMOS inverter chain (Mos1, NMOS VTO 0.4 V / KP 100e-6, PMOS
VTO -0.4 V / KP 50e-6, W/L 1u/0.1u and 2u/0.1u), input tied to 0.5 V, per-device VTO ~ N(base, 15 mV) and
W ~ N(base, 3 %) drawn from seed 0xC2; one DC operating point per instance (68 unknowns, 128 Mos1).

A step is one batched dcop of all 4096 instances from the zero initial guess, with the instances'
parameters already resident in HBM: one persistent kernel launch that runs every instance's Newton loop
to convergence.  The headline value is Newton instance-iterations per second of the mixed-precision
path (FP32 factor and solves, FP64 evaluation, residual, update and convergence test); the plain FP64
path is timed beside it, with the node-voltage / branch-current difference between the two.

Extra objects on the JSON line:
  roofline      algorithmic HBM bytes per launch over the kernel's HIP-event duration, against 8 TB/s
  cpu_baseline  the C++ restatement of the reference algorithm (the oracle) over the same 4096
                instances on host threads
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published
TRAFFIC_FILE = os.path.join(HERE, "profiles", "s21_newton_jit_chain_mc4096_traffic.txt")


def algorithmic_bytes(cnet, structure, factor_bytes):
    """HBM bytes one Newton iteration of one instance must move at minimum:
    sum over devices of (8 B x parameters + 4 B x terminal indices) + 8 B x Jacobian values assembled
    + 2 x factor_bytes x LU entries (written by the factorisation, read by the solves) + 4 n-vectors
    (x, rhs, residual, dx) of 8 B.  Shared topology tables are read once per launch and not counted."""
    from spice21_amd import _abi as A

    dev = 0
    for k in range(cnet.net.n_comps):
        c = cnet.comps[k]
        if c.type == A.S21_COMP_MOS1:
            terms = 4 + (c.node[4] != c.node[0]) + (c.node[5] != c.node[2])
            dev += 8 * 10 + 4 * terms
        elif c.type == A.S21_COMP_DIODE:
            dev += 8 * 10 + 4 * (2 + (c.node[2] != c.node[0]))
        elif c.type == A.S21_COMP_MOS0:
            dev += 4 * 4
        elif c.type == A.S21_COMP_V:
            dev += 8 + 4 * 3
        else:  # R, C, I
            dev += 8 + 4 * 2
    n = cnet.n_vars
    return dev + 8 * structure["nnz_a"] + 2 * factor_bytes * structure["nnz_lu"] + 4 * 8 * n


COUNTERS_FILE = os.path.join(HERE, "profiles", "s21_newton_jit_chain_mc4096_counters.txt")


def summary_fields(path):
    """`key: value` lines of a committed rocprofv3 summary (profiles/), as a dict of strings."""
    out = {}
    try:
        with open(path) as f:
            for line in f:
                if ":" in line and not line.startswith(" ") and line.split(":", 1)[0].replace("_", "").isalnum():
                    k, v = line.split(":", 1)
                    out[k.strip()] = v.strip()
    except OSError:
        pass
    return out


def measured_traffic():
    """HBM bytes per launch of the mixed kernel on this workload, copied from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known size; counters cannot be read inside
    this run), and the kernel time the summary was taken at; (None, None, None) when absent."""
    f = summary_fields(TRAFFIC_FILE)
    if "traffic_bytes_per_launch" not in f:
        return None, None, None
    at_ms = float(f["kernel_avg_ms_at_profile"]) if "kernel_avg_ms_at_profile" in f else None
    return float(f["traffic_bytes_per_launch"]), f"copied from profiles/{os.path.basename(TRAFFIC_FILE)}", at_ms


def binding_unit(kernel_avg_ms):
    """What bounds the kernel, from the committed hardware-counter summary of the same workload (rocprofv3
    --pmc on s21_newton_jit): FP64 VALU and LDS utilisation beside the HBM figure."""
    f = summary_fields(COUNTERS_FILE)
    if not f:
        return None
    out = {"source": f"copied from profiles/{os.path.basename(COUNTERS_FILE)}"}
    for k in ("bound_by", "fp64_tflops", "fp64_frac_of_peak", "valu_busy_frac", "lds_busy_frac",
              "lds_bank_conflict_frac", "hbm_gbs_measured", "hbm_frac_of_peak", "scratch_bytes_per_lane"):
        if k in f:
            v = f[k]
            try:
                v = float(v)
            except ValueError:
                pass
            out[k] = v
    return out


def copy_calibration(mib):
    """Three launches of torch's device copy kernel over mib MiB: their FETCH_SIZE / WRITE_SIZE readings
    per dispatch, against the known mib MiB, give the scale of those counters."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "copies": 3}), flush=True)


def extra_measurements(which=("tran", "ac", "ring5", "grid", "bsim4")):
    """The metric's transient rate and the other workloads, by HIP-event kernel time (each analysis call
    also rebuilds its host structures, which is not the kernel's throughput)."""
    out = {}
    if "tran" in which:
        out.update(_tran_ring_batch())
    if "ac" in which:
        out.update(_ac_amp_ladder())
    if "ring5" in which:
        out.update(_tran_single_ring5())
    if "grid" in which:
        out.update(_power_grid_single())
    if "grid-op" in which:
        out.update(_power_grid_single(tran=False))
    if "bsim4" in which:
        out.update(_bsim4_chain_mc())
        out.update(_bsim4_ring_batch())
    return out


def _bsim4_ring_batch(n_inst=1024, stages=5, tstop=0.5e-9):
    """Batched BSIM4 transient: 1024 five-stage BSIM4 ring oscillators whose stage capacitors are drawn per
    instance (U[1, 10] fF), 0-0.5 ns at 1 ps from node 1 at 0 V, both precisions (parity of such batches
    with the oracle is tests/test_gpu_bsim4.py's).  Seven of the 1024 instances end in "Convergence
    Failed", as the oracle's plain Newton iteration does on them (tests/test_gpu_bsim4.py::
    test_bsim4_ring_batch_workload_matches_the_oracle_per_instance checks two)."""
    import numpy as np

    from spice21_amd import workloads as W
    from spice21_amd.batch import FP64, MIXED, Batch, Group

    nl, over = W.bsim4_ring_batch(n_inst, stages)
    g = Group(nl.to_c(), n_inst, over)
    rec, ws = {}, {}
    for name, prec in (("fp64", FP64), ("mixed", MIXED)):
        b = Batch(g, precision=prec)
        b.tran(1e-12, tstop, [("1", 0.0)])  # untimed: structure and parameters
        b.tran(1e-12, tstop, [("1", 0.0)])
        ms = b.timing()[0]
        st, it = b.status()
        npts = len(b.axis())
        iters = int(it.sum())
        rec[name] = {"tran_pts_per_s": n_inst * npts / (ms * 1e-3), "instance_iters_per_s": iters / (ms * 1e-3),
                     "bsim4_evals_per_s": iters * 2 * stages / (ms * 1e-3), "kernel_ms": ms, "points": npts,
                     "newton_iters": iters, "failed_instances": int(np.sum(st != 0)), "launch": list(b.config())}
        ws[name] = b.wave(0, 0)
    v = np.asarray(list(nl.kinds) + [0, 1]) == 0
    rec["mixed_vs_fp64_max_node_v_instance0"] = float(np.abs(ws["mixed"] - ws["fp64"])[:, v[:ws["fp64"].shape[1]]].max())
    return {"bsim4_ring_batch": {"workload": f"{n_inst} {stages}-stage BSIM4 ring oscillators (C per instance), "
                                             f"0-{tstop * 1e9:g} ns at 1 ps", **rec}}


def _bsim4_chain_mc(n_inst=4096, stages=16):
    """Batched BSIM4: Monte-Carlo operating points of 16-stage BSIM4 inverter chains (32 devices each, DELVTO
    and W per device and instance), both precisions (parity with the oracle is tests/test_gpu_bsim4.py's)."""
    import numpy as np

    from spice21_amd import workloads as W
    from spice21_amd.batch import FP64, MIXED, Batch, Group

    nl, over = W.bsim4_chain_mc(n_inst, stages)
    g = Group(nl.to_c(), n_inst, over)
    rec, xs = {}, {}
    for name, prec in (("fp64", FP64), ("mixed", MIXED)):
        b = Batch(g, precision=prec)
        t0 = time.perf_counter()
        b.op()  # untimed: derives and uploads the structure and every instance's BSIM4 parameters
        prep_s = time.perf_counter() - t0
        b.op()
        ms = b.timing()[0]
        st, it = b.status()
        iters = int(it.sum())
        rec[name] = {"instance_iters_per_s": iters / (ms * 1e-3),
                     "bsim4_evals_per_s": iters * 2 * stages / (ms * 1e-3), "kernel_ms": ms,
                     "newton_iters": iters, "failed_instances": int(np.sum(st != 0)),
                     "first_call_s": prep_s, "launch": list(b.config())}
        xs[name] = b.x()
    rec["mixed_vs_fp64_max_node_v"] = float(np.abs(xs["mixed"] - xs["fp64"]).max())
    return {"bsim4_chain_mc": {"workload": f"{n_inst} x {stages}-stage BSIM4 inverter chain dcop "
                                           f"({2 * stages} BSIM4 per instance, DELVTO and W per device)", **rec}}


def _power_grid_single(tran=True):
    """The single large instance: 200 x 200 power-grid mesh with 5,000 inverters and 2,000 diodes (45,040
    unknowns), dcop then 0-1 ns at 1 ps from a 0.5 V kick on mesh node (100, 100), FP64 and mixed.  Rates
    are wall time of each analysis (the Newton loop is driven from the host, one read-back per iteration)."""
    import numpy as np

    from spice21_amd import workloads as W
    from spice21_amd.large import FP64, MIXED, Large

    nl = W.power_grid()
    cn = nl.to_c()
    vmask = np.asarray(list(nl.kinds) + [0, 1]) == 0
    rec = {"workload": "200x200 power grid, 5000 inverters, 2000 diodes (45040 unknowns): dcop + 0-1 ns at 1 ps, "
                       "0.5 V kick on mesh node (100, 100)"}
    xs, ws = {}, {}
    for name, prec in (("fp64", FP64), ("mixed", MIXED)):
        g = Large(cn, precision=prec).op()
        st, it = g.status()
        ms, nf = g.timing()
        fms, sms = g.time_factor(3)
        S = g.structure()
        rec[name] = {"dcop": {"status": st, "newton_iters": it, "ms": ms, "instance_iters_per_s": it / (ms * 1e-3),
                              "factorisations": nf},
                     "factor_ms": fms, "sweeps_ms": sms, "factor_tflops": S["flops_lu"] / (fms * 1e-3) / 1e12}
        xs[name] = g.x()
        rec["structure"] = S
        if not tran:
            continue
        g.tran(1e-12, 1e-9, [("g100_100", 0.5)])
        st, it = g.status()
        ms, nf = g.timing()
        npts = len(g.axis())
        rec[name]["tran"] = {"status": st, "points": npts, "newton_iters": it, "ms": ms,
                             "tran_pts_per_s": npts / (ms * 1e-3), "factorisations": nf}
        ws[name] = g.wave()
    v = np.asarray(nl.kinds) == 0
    rec["mixed_vs_fp64_dcop_max_node_v"] = float(np.abs(xs["mixed"] - xs["fp64"])[v].max())
    if not tran:
        return {"power_grid_single": rec}
    n = min(len(ws["mixed"]), len(ws["fp64"]))
    rec["mixed_vs_fp64_tran_max_node_v"] = float(np.abs(ws["mixed"][:n] - ws["fp64"][:n])[:, vmask].max())
    return {"power_grid_single": rec}


def _tran_ring_batch():
    import numpy as np

    from spice21_amd import workloads as W
    from spice21_amd.batch import FP64, MIXED, Batch, Group

    out = {}
    # batched transient of 1024 uneven ring oscillators
    gs = W.ring_groups(1024)
    groups = [Group(n.to_c(), k, over) for n, k, over, _ in gs]
    rec, waves = {}, {}
    for name, prec in (("fp64", FP64), ("mixed", MIXED)):
        b = Batch(groups, precision=prec).tran(1e-12, 5e-9, [("1", 0.0)])
        ms = b.timing()[0]
        npts = len(b.axis())
        iters = sum(int(b.status(g)[1].sum()) for g in range(len(groups)))
        failed = sum(int(np.sum(b.status(g)[0] != 0)) for g in range(len(groups)))
        rec[name] = {"tran_pts_per_s": 1024 * npts / (ms * 1e-3), "instance_iters_per_s": iters / (ms * 1e-3),
                     "kernel_ms": ms, "failed_instances": failed}
        waves[name] = [b.wave(g, 0) for g in range(0, len(groups), 7)]
    rec["mixed_vs_fp64_max_node_v"] = float(max(np.abs(a - c).max() for a, c in zip(waves["fp64"], waves["mixed"])))
    out["tran_ring_batch"] = {"workload": "1024 ring oscillators of 3..101 stages (C, VDD per instance), "
                                          "0-5 ns at 1 ps, 5000 points", **rec}
    return out


def _ac_amp_ladder():
    import numpy as np

    from spice21_amd import workloads as W
    from spice21_amd.batch import FP64, MIXED, Batch, Group

    out = {}
    # batched AC sweep
    nl, over = W.amp_ladder(256)
    g = Group(nl.to_c(), 256, over)
    rec, sw = {}, {}
    for name, prec in (("fp64", FP64), ("mixed", MIXED)):
        b = Batch(g, precision=prec).ac(1000, 100_000_000_000, 1000)  # untimed: builds the generated kernel
        b.ac(1000, 100_000_000_000, 1000)
        ms, launches, _ = b.timing()
        nf = len(b.axis())
        st, it = b.status()
        rec[name] = {"instance_freqs_per_s": 256 * nf / (ms * 1e-3), "kernel_ms": ms, "launches": launches,
                     "runs_per_instance": b.ac_runs(), "generated_factorisation": b.config()[2],
                     "newton_iters": int(it.sum()), "failed_instances": int(np.sum(st != 0))}
        sw[name] = [b.wave(0, i) for i in range(0, 256, 37)]
    rec["mixed_vs_fp64_max_abs"] = float(max(np.abs(a - c).max() for a, c in zip(sw["fp64"], sw["mixed"])))
    out["ac_amp_ladder"] = {"workload": "256 two-stage CS amplifiers with a 20-section RC ladder, 1 kHz-100 GHz, "
                                        "1002 frequencies", **rec}
    return out


def _tran_single_ring5():
    import numpy as np

    from spice21_amd import workloads as W
    from spice21_amd.batch import FP64, MIXED, Batch, Group

    out = {}
    # one 5-stage ring oscillator (latency): the default path (record-driven factorisation) and the one
    # with the factorisation generated for its pattern (a run-time compilation a single run rarely repays)
    nl = W.ring5()
    rec = {"workload": "5-stage CMOS ring, 0-10 ns at 1 ps"}
    waves = {}
    for prec, jit, key in ((FP64, 0, ""), (FP64, 1, "_generated"), (MIXED, 0, "_mixed")):
        b = Batch(Group(nl.to_c()), precision=prec).set_jit(jit).tran(1e-12, 1e-8, [("1", 0.5)])
        b.tran(1e-12, 1e-8, [("1", 0.5)])
        ms = b.timing()[0]
        it = int(b.status()[1][0])
        rec.update({"points": len(b.axis()), "newton_iters" + key: it, "kernel_ms" + key: ms,
                    "tran_pts_per_s" + key: len(b.axis()) / (ms * 1e-3), "us_per_newton_iter" + key: ms * 1e3 / it})
        waves[key] = b.wave()
    v = [k for k, kind in enumerate(nl.kinds) if kind == 0]
    i = [k for k, kind in enumerate(nl.kinds) if kind == 1]
    rec["mixed_vs_fp64_max_node_v"] = float(np.abs(waves["_mixed"][:, v] - waves[""][:, v]).max())
    rec["mixed_vs_fp64_max_branch_i"] = float(np.abs(waves["_mixed"][:, i] - waves[""][:, i]).max())
    # trapezoidal integration with local-truncation-error step control (single-instance path), 1 ps the
    # largest step, output on the same 1 ps grid; its error against a fixed-step trapezoidal run at 0.1 ps
    # over the first nanosecond
    from spice21_amd.large import Large

    g = Large(nl.to_c()).tran_trap(1e-12, 1e-8, [("1", 0.5)], lte=True)
    st, it = g.status()
    ms = g.timing()[0]
    acc, rej = g.trap_stats()
    lte = {"status": st, "points": len(g.axis()), "newton_iters": it, "ms": ms,
           "tran_pts_per_s": len(g.axis()) / (ms * 1e-3), "accepted_steps": acc, "rejected_steps": rej}
    w = g.wave()
    f = Large(nl.to_c()).tran_trap(0.1e-12, 1e-9, [("1", 0.5)], lte=False)
    tf, wf = f.axis(), f.wave()
    t = g.axis()
    m = t <= 1e-9
    ref = np.stack([np.interp(t[m], tf, wf[:, k]) for k in v], axis=1)
    lte["max_node_v_vs_fixed_0p1ps_first_ns"] = float(np.abs(w[m][:, v] - ref).max())
    rec["trap_lte"] = lte
    out["tran_single_ring5"] = rec
    return out


def time_batch(batch, steps, warmup):
    for _ in range(warmup):
        batch.op()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.op()  # returns once the launch has completed (HIP event synchronisation)
        kernel_ms.append(batch.timing()[0])
    elapsed = time.perf_counter() - t0
    st, it = batch.status()
    return elapsed, kernel_ms, st, it


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-inst", type=int, default=4096)
    ap.add_argument("--cpu-threads", type=int, default=16, help="host threads of the CPU baseline")
    ap.add_argument("--cpu-reps", type=int, default=3, help="repetitions of the CPU baseline (median reported)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="skip the instances-per-wavefront autotuning before the timed steps")
    ap.add_argument("--ipw", type=int, default=0, help="instances per wavefront (implies --no-autotune)")
    ap.add_argument("--no-extra", dest="extra", action="store_false",
                    help="skip the transient / AC / single-ring measurements")
    ap.add_argument("--extra-only", choices=["tran", "ac", "ring5", "grid", "grid-op", "bsim4"], default=None,
                    help="only run one of the other workloads and print its record (for profiling runs)")
    ap.add_argument("--precision", choices=["both", "mixed", "fp64"], default="both",
                    help="'mixed' / 'fp64' time one path only (for profiling runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (reference reading for the HBM "
                         "traffic counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.extra_only:
        print(json.dumps(extra_measurements((args.extra_only,))), flush=True)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the batched instances are independent, so a multi-GPU run would "
              "shard them by instance with no exchange during the solve; run it with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from spice21_amd import _abi as A
    from spice21_amd import workloads as W
    from spice21_amd.batch import FP64, MIXED, Batch, Group

    nl, over = W.chain_mc(args.n_inst)
    cn = nl.to_c()
    group = Group(cn, args.n_inst, over)
    kinds = np.array(nl.kinds)
    vmask, imask = kinds == A.S21_VAR_V, kinds == A.S21_VAR_I

    res = {}
    for name, prec in (("fp64", FP64), ("mixed", MIXED)):
        if args.precision not in ("both", name):
            continue
        b = Batch(group, precision=prec)
        if args.ipw:
            b.set_ipw(args.ipw)
        tune = b.autotune(lambda bb: bb.op()) if args.autotune and not args.ipw else {}
        elapsed, kms, st, it = time_batch(b, args.steps, args.warmup)
        ipw, waves, jit = b.config()
        res[name] = dict(batch=b, elapsed=elapsed, kernel_ms=kms, status=st, iters=it, x=b.x(),
                         structure=b.structure(), launch={"instances_per_wavefront": ipw, "wavefronts_per_workgroup": waves,
                                                          "generated_factorisation": jit,
                                                          "tuned": getattr(b, "tuned", None),
                                                          "autotune_ms": {str(k): round(v, 4) for k, v in tune.items()}})
        print(f"[bench] {name}: {args.steps} steps in {elapsed:.4f} s, kernel {np.mean(kms):.3f} ms/launch, "
              f"iterations/instance {it.min()}..{it.max()}, failed {int(np.sum(st != 0))}, launch {res[name]['launch']}",
              file=sys.stderr, flush=True)
    head = "mixed" if "mixed" in res else "fp64"
    h = res[head]
    inst_iters = int(h["iters"].sum())  # per step; every step repeats the same deterministic solve
    value = inst_iters * args.steps / h["elapsed"]
    kernel_avg_ms = float(np.mean(h["kernel_ms"]))
    b_iter = algorithmic_bytes(cn, h["structure"], 4 if head == "mixed" else 8)
    achieved = b_iter * inst_iters / (kernel_avg_ms * 1e-3) / 1e9
    traffic, traffic_src, traffic_at_ms = measured_traffic() if head == "mixed" else (None, None, None)

    line = {
        "metric": "batched Newton instance-iters/s & tran pts/s, mixed vs FP64 at \u22641e-6 V error",
        "value": value,
        "unit": "instance-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["elapsed"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 factor+solve / f64 eval+residual" if head == "mixed" else "f64",
        "data": "synthetic (seeded Monte Carlo over VTO and W)",
        "config": {"workload": f"chain_mc: {args.n_inst} x 64-stage CMOS inverter chain dcop", "n_inst": args.n_inst,
                   "unknowns": cn.n_vars, "mos1_per_instance": 128, "nnz_a": h["structure"]["nnz_a"],
                   "nnz_lu": h["structure"]["nnz_lu"], "newton_iters_per_step": inst_iters,
                   "failed_instances": int(np.sum(h["status"] != 0)), "launch": h["launch"]},
        # achieved: the algorithmic bytes of the Newton iterations (matrix, LU and vectors counted as if
        # each went through HBM once) over the kernel time.  They stay in LDS here: measured HBM traffic is
        # below 1 % of them, and the unit that binds the kernel is named in "binding".
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     # the copied figure describes this kernel only while its time still matches the profiled one
                     "traffic_profile_kernel_ms": traffic_at_ms,
                     "traffic_kernel_time_matches": (abs(kernel_avg_ms / traffic_at_ms - 1.0) <= 0.1
                                                     if traffic_at_ms else None),
                     "bytes_per_inst_iter": b_iter, "kernel_avg_ms": kernel_avg_ms,
                     "binding": binding_unit(kernel_avg_ms) if head == "mixed" else None},
    }
    if "fp64" in res:
        f = res["fp64"]
        f_iters = int(f["iters"].sum())
        line["fp64"] = {"value": f_iters * args.steps / f["elapsed"], "ms_per_step": f["elapsed"] / args.steps * 1e3,
                        "kernel_avg_ms": float(np.mean(f["kernel_ms"])), "newton_iters_per_step": f_iters}
    if "fp64" in res and "mixed" in res:
        d = np.abs(res["mixed"]["x"] - res["fp64"]["x"])
        line["mixed_vs_fp64"] = {"max_node_v": float(d[:, vmask].max()), "max_branch_i": float(d[:, imask].max()),
                                 "speedup": value / line["fp64"]["value"]}

    if args.extra and args.precision == "both":
        line["other_workloads"] = extra_measurements()

    if not args.no_cpu:
        from oracle import oracle_py as O  # the CPU restatement, timed as the baseline and used as the checker

        nproc = os.cpu_count() or 1
        try:
            allowed = len(os.sched_getaffinity(0))
        except AttributeError:
            allowed = nproc
        # every core this process may use, up to the host-thread share of one GPU (--cpu-threads, 16)
        threads = max(1, min(args.cpu_threads, allowed))
        walls = []
        for _ in range(args.cpu_reps):  # the median of a few repetitions: a rate steadier than one pass
            t0 = time.perf_counter()
            xo, ito, sto = O.dcop_batch(group, threads=threads)
            walls.append(time.perf_counter() - t0)
        cpu_s = float(np.median(walls))
        n1 = min(args.n_inst, 512)  # one core: a sample of the same instances
        walls1 = []
        for _ in range(args.cpu_reps):
            t0 = time.perf_counter()
            _, it1, _ = O.dcop_batch(group, n_inst=n1, threads=1)
            walls1.append(time.perf_counter() - t0)
        cpu1_s = float(np.median(walls1))
        model = ""
        try:
            with open("/proc/cpuinfo") as f:
                model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
        except OSError:
            pass
        line["cpu_baseline"] = {
            "value": float(ito.sum()) / cpu_s, "unit": "instance-iters/s", "cores": threads, "kind": "port",
            "sample": f"all {args.n_inst} instances, one dcop each ({int(ito.sum())} Newton iterations), "
                      f"{len(walls)} times on {threads} threads: median {cpu_s:.2f} s wall "
                      f"(about {sum(walls) * threads + sum(walls1):.0f} s of CPU work in all, one-core sample "
                      f"included)",
            "what": "the oracle: a C++ restatement of the reference algorithm (Markowitz re-pivoting and "
                    "linked-list LU at every Newton iteration, fixed-step BE), not the Rust reference (no "
                    "Rust toolchain on these hosts)",
            "nproc": nproc, "cores_allowed": allowed, "cpu_model": model,
            "one_core": {"value": float(it1.sum()) / cpu1_s, "unit": "instance-iters/s",
                         "sample": f"the first {n1} instances on 1 thread, median of {len(walls1)}: "
                                   f"{cpu1_s:.2f} s"}}
        if "fp64" in res:
            xf = res["fp64"]["x"]
            rel = np.abs(xf - xo) / np.maximum(np.abs(xo), 1e-9)
            line["fp64_vs_oracle"] = {"max_rel": float(rel.max()), "max_abs": float(np.abs(xf - xo).max()),
                                      "iters_equal": bool(np.array_equal(ito, res["fp64"]["iters"]))}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
