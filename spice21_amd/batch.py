"""Batched analyses on the MI355X engine (Python view of the engine C ABI).

A :class:`Batch` holds groups; a group is one flattened netlist (one topology) times ``n_inst`` instances
that differ only in parameter values given as per-instance overrides.  Each analysis runs all instances
of all groups in persistent kernel launches on the current GPU.
"""
import ctypes as C

import numpy as np

from . import _abi as A
from ._lib import check, lib

FP64, MIXED = A.S21_FP64, A.S21_MIXED


def _i32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _field(target, comp_type, name):
    if isinstance(name, int):
        return name
    if target == A.S21_OVR_VALUE:
        return 0
    table = {
        (A.S21_OVR_MODEL, A.S21_COMP_MOS1): A.MOS1_MODEL_FIELDS,
        (A.S21_OVR_INST, A.S21_COMP_MOS1): A.MOS1_INST_FIELDS,
        (A.S21_OVR_MODEL, A.S21_COMP_DIODE): A.DIODE_MODEL_FIELDS,
        (A.S21_OVR_INST, A.S21_COMP_DIODE): A.DIODE_INST_FIELDS,
        (A.S21_OVR_MODEL, A.S21_COMP_BSIM4): A.BSIM4_MODEL_FIELDS,
        (A.S21_OVR_INST, A.S21_COMP_BSIM4): A.BSIM4_INST_FIELDS,
    }[(target, comp_type)]
    return A._field(table, name)


def make_tran_options(tstep, tstop, ics=(), var_names=()):
    """ics: (node name, value) pairs resolved against ``var_names``."""
    n = len(ics)
    var = (C.c_int32 * max(n, 1))()
    val = (C.c_double * max(n, 1))()
    names_b = [str(nm).encode() for nm, _ in ics]
    names = (C.c_char_p * max(n, 1))(*names_b)
    for k, (nm, v) in enumerate(ics):
        var[k] = list(var_names).index(str(nm))
        val[k] = v
    t = A.s21_tran_options(tstep, tstop, n, var, val, names)
    t._keep = (var, val, names, names_b)
    return t


def symbolic(cnet, opts=None, tran=None):
    """Host symbolic phase only: (row_i2e, col_i2e, nnz_a, nnz_lu, levels, lu_flops)."""
    n = cnet.n_vars + (2 * tran.n_ic if tran is not None else 0)
    ri = np.zeros(max(n, 1), dtype=np.int32)
    ci = np.zeros(max(n, 1), dtype=np.int32)
    na, nl, lv, fl = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
    check(lib().s21_symbolic(cnet.ptr, C.byref(opts) if opts is not None else None,
                             C.byref(tran) if tran is not None else None, _i32p(ri), _i32p(ci), C.byref(na),
                             C.byref(nl), C.byref(lv), C.byref(fl)))
    return ri[:n], ci[:n], na.value, nl.value, lv.value, fl.value


class Group:
    def __init__(self, cnet, n_inst=1, overrides=()):
        """overrides: iterable of (comp_index, target, field, values[n_inst]) with target one of
        'value' | 'model' | 'inst' and field a parameter name (e.g. 'vt0', 'w') or proto field number."""
        self.cnet = cnet
        self.n_inst = int(n_inst)
        tmap = {"value": A.S21_OVR_VALUE, "model": A.S21_OVR_MODEL, "inst": A.S21_OVR_INST}
        self._vals = []
        self.over = (A.s21_override * max(len(overrides), 1))()
        for k, (comp, target, field, vals) in enumerate(overrides):
            tg = tmap[target] if isinstance(target, str) else int(target)
            v = np.ascontiguousarray(vals, dtype=np.float64)
            assert v.shape == (self.n_inst,)
            self._vals.append(v)
            ctype = cnet.comps[comp].type
            self.over[k] = A.s21_override(comp, tg, _field(tg, ctype, field), _dp(v))
        self.n_over = len(overrides)


class Batch:
    def __init__(self, groups, opts=None, precision=FP64):
        if isinstance(groups, Group):
            groups = [groups]
        self.groups = list(groups)
        arr = (A.s21_group * len(self.groups))()
        for k, g in enumerate(self.groups):
            arr[k] = A.s21_group(C.pointer(g.cnet.net), g.n_inst, g.n_over, g.over)
        self._arr = arr
        self.opts = opts
        self.h = C.c_void_p()
        check(lib().s21_batch_create(len(self.groups), arr, C.byref(opts) if opts is not None else None,
                                     int(precision), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().s21_batch_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- analyses ----
    def op(self):
        check(lib().s21_batch_op(self.h))
        self._ac = False
        return self

    def ac(self, fstart, fstop, npts):
        """Operating point, then the small-signal sweep from fstart to fstop Hz (npts steps, log-spaced)."""
        a = A.s21_ac_options(int(fstart), int(fstop), int(npts))
        check(lib().s21_batch_ac(self.h, C.byref(a)))
        self._ac = True
        return self

    def tran(self, tstep, tstop, ics=()):
        names = self.groups[0].cnet.var_names
        t = make_tran_options(tstep, tstop, ics, names)
        check(lib().s21_batch_tran(self.h, C.byref(t)))
        self._ac = False
        return self

    # ---- results ----
    def n_vars(self, g=0):
        return lib().s21_batch_n_vars(self.h, g)

    def var_names(self, g=0):
        return [lib().s21_batch_var_name(self.h, g, k).decode() for k in range(self.n_vars(g))]

    def x(self, g=0):
        n, ni = self.n_vars(g), self.groups[g].n_inst
        out = np.zeros((ni, n))
        check(lib().s21_batch_get_x(self.h, g, _dp(out)))
        return out

    def status(self, g=0):
        ni = self.groups[g].n_inst
        st = np.zeros(ni, dtype=np.int32)
        it = np.zeros(ni, dtype=np.int64)
        check(lib().s21_batch_get_status(self.h, g, _i32p(st), it.ctypes.data_as(C.POINTER(C.c_int64))))
        return st, it

    def axis(self):
        n = lib().s21_batch_n_points(self.h)
        out = np.zeros(n)
        check(lib().s21_batch_get_axis(self.h, _dp(out)))
        return out

    def wave(self, g=0, inst=0):
        """Transient waveform [points, vars], or after ac() the complex sweep [frequencies, vars]."""
        n = lib().s21_batch_n_points(self.h)
        if getattr(self, "_ac", False):
            out = np.zeros((n, self.n_vars(g), 2))
            check(lib().s21_batch_get_wave(self.h, g, inst, _dp(out)))
            return out[:, :, 0] + 1j * out[:, :, 1]
        out = np.zeros((n, self.n_vars(g)))
        check(lib().s21_batch_get_wave(self.h, g, inst, _dp(out)))
        return out

    def write_ac_results(self, directory, g=0, inst=0):
        """The reference's AC result files (stream.ac.json, ac.json) of one instance, into `directory`."""
        check(lib().s21_batch_write_ac_results(self.h, g, inst, str(directory).encode()))

    def timing(self):
        ms, nl, avg = C.c_double(), C.c_int64(), C.c_double()
        check(lib().s21_batch_timing(self.h, C.byref(ms), C.byref(nl), C.byref(avg)))
        return ms.value, nl.value, avg.value

    def resumptions(self):
        """Instance-launches of the last operating point / transient that resumed an instance on a new
        pivot order (its kept order met a pivot that is zero or below the smallest normal number)."""
        n = C.c_int64()
        check(lib().s21_batch_resumptions(self.h, C.byref(n)))
        return n.value

    def set_profiling(self, on=True):
        check(lib().s21_batch_set_profiling(self.h, 1 if on else 0))
        return self

    def set_ipw(self, ipw):
        """Instances per wavefront (1, 2, 4, 8, 16), or 0 to let each launch choose."""
        check(lib().s21_batch_set_ipw(self.h, int(ipw)))
        return self

    def set_jit(self, mode):
        """Generated factorisation: 1 always, -1 never, 0 for large single-pattern batches."""
        check(lib().s21_batch_set_jit(self.h, int(mode)))
        return self

    def set_ac_runs(self, runs):
        """AC: each instance's frequencies cut into this many contiguous runs solved in parallel, each
        started from x = 0 (1 = the reference's sequence of warm starts; 0 = chosen per sweep)."""
        check(lib().s21_batch_set_ac_runs(self.h, int(runs)))
        return self

    def ac_runs(self):
        """Runs of frequencies per instance of the last AC sweep."""
        return int(lib().s21_batch_last_ac_runs(self.h))

    def set_waves(self, waves):
        """Wavefronts per workgroup, at most this many (1..8), or 0 to let each launch choose."""
        check(lib().s21_batch_set_waves(self.h, int(waves)))
        return self

    def set_hub(self, policy):
        """Hub policy of the stamp scatter: 1 every shared slot, 2 only slots of many contributors, 0 auto."""
        check(lib().s21_batch_set_hub(self.h, int(policy)))
        return self

    def _time(self, run, repeats):
        run(self)  # untimed: builds the structure and, for a generated factorisation, its kernel
        ms = []
        for _ in range(repeats):
            run(self)
            ms.append(self.timing()[0])
        return min(ms)

    def autotune(self, run, candidates=(1, 2, 4, 8, 16), waves=(0, 2, 4), jit=(0, -1), hub=(0, 1, 2), repeats=2):
        """Times ``run(self)`` (an analysis call) over the launch shape and structure and keeps the fastest:
        instances per wavefront (all candidates, at the automatic waves), then, at the best of those,
        wavefronts per workgroup, the factorisation (0: generated where the batch qualifies, -1:
        record-driven) and the hub policy, each in turn at the best so far.  Results are bit-identical
        across the shapes and factorisations; the hub policy changes the order shared slots are summed in
        (agreeing to rounding).  Returns {"ipw=..,waves=..,jit=..,hub=..": kernel ms}
        for every configuration timed; ``self`` is left at the fastest."""
        best = {"ipw": 0, "waves": 0, "jit": 0, "hub": 0}
        setters = {"ipw": self.set_ipw, "waves": self.set_waves, "jit": self.set_jit, "hub": self.set_hub}
        times = {}

        def key(c):
            return ",".join(f"{k}={c[k]}" for k in ("ipw", "waves", "jit", "hub"))

        def apply(c):
            for k, v in c.items():
                setters[k](v)

        for dim, values in (("ipw", candidates), ("waves", waves), ("jit", jit), ("hub", hub)):
            for v in values:
                c = dict(best, **{dim: v})
                if key(c) in times:
                    continue
                apply(c)
                try:
                    times[key(c)] = self._time(run, repeats)
                except Exception:  # a shape the circuit does not fit (LDS) is skipped
                    continue
            best = min(({k: int(x.split("=")[1]) for k, x in zip(("ipw", "waves", "jit", "hub"), t.split(","))}
                        for t in times), key=lambda c: times[key(c)])
        apply(best)
        run(self)
        self.tuned = best
        return times

    def config(self):
        """(instances per wavefront, wavefronts per workgroup, generated factorisation used) of the last
        launch."""
        a, w, j = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().s21_batch_config(self.h, C.byref(a), C.byref(w), C.byref(j)))
        return a.value, w.value, bool(j.value)

    PHASES = ("devices", "hub", "residual", "factor", "backward", "update", "commit", "total")

    def profile(self, g=0):
        """Per-instance device cycles of the last analysis by phase: array [n_inst, 8]."""
        out = np.zeros((self.groups[g].n_inst, 8), dtype=np.int64)
        check(lib().s21_batch_profile(self.h, g, out.ctypes.data_as(C.POINTER(C.c_int64))))
        return out

    def pivots(self, g=0):
        n = self.n_vars(g)
        ri = np.zeros(n, dtype=np.int32)
        ci = np.zeros(n, dtype=np.int32)
        check(lib().s21_batch_pivots(self.h, g, _i32p(ri), _i32p(ci)))
        return ri, ci

    def structure(self, g=0):
        v = [C.c_int64() for _ in range(4)]
        check(lib().s21_batch_structure(self.h, g, *[C.byref(x) for x in v]))
        return dict(nnz_a=v[0].value, nnz_lu=v[1].value, levels=v[2].value, n_contrib=v[3].value)

    def assemble(self, x, g=0, inst=0, max_elems=1 << 20):
        """Matrix elements (rows, cols, values in creation order) and RHS of instance ``inst`` at the
        operating-point voltages ``x`` (one value per unknown of the group)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        n_vars = self.n_vars(g)
        if x.shape != (n_vars,):
            raise ValueError(f"x has shape {x.shape}, the group has {n_vars} unknowns")
        rows = np.zeros(max_elems, dtype=np.int32)
        cols = np.zeros(max_elems, dtype=np.int32)
        vals = np.zeros(max_elems)
        rhs = np.zeros(max(n_vars, 1))
        ne = C.c_int64()
        check(lib().s21_batch_assemble(self.h, g, inst, _dp(x), _i32p(rows), _i32p(cols), _dp(vals), _dp(rhs),
                                       int(max_elems), C.byref(ne)))
        n = ne.value
        return rows[:n], cols[:n], vals[:n], rhs[:n_vars]
