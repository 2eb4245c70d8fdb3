"""ctypes binding of the oracle (test infrastructure only).

The oracle is a CPU FP64 restatement of the reference's Newton loop (oracle/oracle.cpp).  Only the test
suite, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from spice21_amd import _abi as A

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-B", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        d, i32, i64 = C.c_double, C.c_int32, C.c_int64
        L.oracle_dcop.argtypes = [P(A.s21_netlist), P(A.s21_options), P(d), P(i64), P(i64), C.c_char_p, C.c_size_t]
        L.oracle_dcop_batch.argtypes = [P(A.s21_netlist), P(A.s21_options), i32, P(A.s21_override), i32, i32, i32,
                                        P(d), P(i64), P(i32)]
        L.oracle_dcop.restype = C.c_int
        L.oracle_tran.argtypes = [P(A.s21_netlist), P(A.s21_options), P(A.s21_tran_options), i64, P(d), P(d), P(i32),
                                  P(i64), C.c_char_p, C.c_size_t]
        L.oracle_tran.restype = i64
        L.oracle_ac.argtypes = [P(A.s21_netlist), P(A.s21_options), P(A.s21_ac_options), i64, P(d), P(d), P(i64),
                                C.c_char_p, C.c_size_t]
        L.oracle_ac.restype = i64
        L.oracle_assemble_op.argtypes = [P(A.s21_netlist), P(A.s21_options), P(d), i64, P(i32), P(i32), P(d), P(d),
                                         C.c_char_p, C.c_size_t]
        L.oracle_assemble_op.restype = i64
        L.oracle_first_pivots.argtypes = [P(A.s21_netlist), P(A.s21_options), P(d), P(i32), P(i32), P(i64), C.c_char_p,
                                          C.c_size_t]
        L.oracle_first_pivots.restype = C.c_int
        L.oracle_mos1_derive.argtypes = [P(A.s21_params), P(A.s21_params), P(A.s21_options), P(d), C.c_char_p, C.c_size_t]
        L.oracle_diode_derive.argtypes = [P(A.s21_params), P(A.s21_params), P(A.s21_options), P(d)]
        L.oracle_mat_new.argtypes = [C.c_int]
        L.oracle_mat_new.restype = C.c_void_p
        L.oracle_mat_free.argtypes = [C.c_void_p]
        L.oracle_mat_add.argtypes = [C.c_void_p, i64, i64, d, d]
        L.oracle_mat_factor.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t]
        L.oracle_mat_solve.argtypes = [C.c_void_p, i64, P(d), P(d), C.c_char_p, C.c_size_t]
        L.oracle_mat_get.argtypes = [C.c_void_p, i64, i64, P(d), P(d)]
        L.oracle_mat_swap_rows.argtypes = [C.c_void_p, i64, i64]
        L.oracle_mat_row_mapping.argtypes = [C.c_void_p, i64, P(i64), P(i64)]
        L.oracle_mat_check.argtypes = [C.c_void_p]
        L.oracle_plant.argtypes = [C.c_int]
        L.oracle_plant.restype = None
        L.oracle_pivot_stats.argtypes = [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.c_int]
        L.oracle_pivot_stats.restype = None
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, code, desc):
        super().__init__(desc)
        self.code = code
        self.desc = desc


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _opts(opts):
    return opts if opts is not None else A.s21_options.default()


def tran_point_count(tstep, tstop):
    """Points of the reference's loop (analysis.rs:551-570), including the OP point at t = 0."""
    t, n = tstep, 1
    while t < tstop:
        n += 1
        t += tstep
    return n


def dcop(cnet, opts=None):
    """Returns (x, newton_iterations, factorisations)."""
    x = np.zeros(max(cnet.n_vars, 1))
    it, fa = C.c_int64(), C.c_int64()
    err = C.create_string_buffer(512)
    rc = lib().oracle_dcop(cnet.ptr, C.byref(_opts(opts)), _dp(x), C.byref(it), C.byref(fa), err, 512)
    if rc != 0:
        raise OracleError(rc, err.value.decode())
    return x[: cnet.n_vars], it.value, fa.value


def dcop_batch(group, i0=0, n_inst=None, threads=1, opts=None):
    """One dcop per instance of a batch group (``spice21_amd.batch.Group``: a netlist plus per-instance
    overrides), on ``threads`` host threads.  Returns (x[n_inst, n_vars], iters[n_inst], status[n_inst])."""
    n = group.n_inst - i0 if n_inst is None else int(n_inst)
    nv = group.cnet.n_vars
    x = np.zeros((max(n, 1), max(nv, 1)))
    it = np.zeros(max(n, 1), dtype=np.int64)
    st = np.zeros(max(n, 1), dtype=np.int32)
    rc = lib().oracle_dcop_batch(group.cnet.ptr, C.byref(_opts(opts)), group.n_over, group.over, int(i0), n,
                                 int(threads), _dp(x), it.ctypes.data_as(C.POINTER(C.c_int64)),
                                 st.ctypes.data_as(C.POINTER(C.c_int32)))
    if rc != 0:
        raise OracleError(rc, "oracle_dcop_batch: bad arguments")
    return x[:n, :nv], it[:n], st[:n]


def make_tran_options(tstep, tstop, ics=(), cnet=None):
    """ics: sequence of (node name, value) resolved against cnet's variable names."""
    n = len(ics)
    var = (C.c_int32 * max(n, 1))()
    val = (C.c_double * max(n, 1))()
    names_b = [str(nm).encode() for nm, _ in ics]
    names = (C.c_char_p * max(n, 1))(*names_b)
    for k, (nm, v) in enumerate(ics):
        var[k] = cnet.var_names.index(str(nm))
        val[k] = v
    t = A.s21_tran_options(tstep, tstop, n, var, val, names)
    t._keep = (var, val, names, names_b)
    return t


def tran(cnet, tstep, tstop, ics=(), opts=None):
    """Returns (time[T], X[T, n_total], names, newton_iterations)."""
    tr = make_tran_options(tstep, tstop, ics, cnet)
    npts = tran_point_count(tstep, tstop)
    nt = cnet.n_vars + 2 * len(ics)
    t = np.zeros(npts)
    X = np.zeros((npts, nt))
    ntot, it = C.c_int32(), C.c_int64()
    err = C.create_string_buffer(512)
    r = lib().oracle_tran(cnet.ptr, C.byref(_opts(opts)), C.byref(tr), npts, _dp(t), _dp(X), C.byref(ntot),
                          C.byref(it), err, 512)
    if r < 0:
        raise OracleError(-r, err.value.decode())
    names = list(cnet.var_names)
    for nm, _ in ics:
        names += [f".{nm}.vic", f".{nm}.iic"]
    return t[:r], X[:r], names, it.value


def ac(cnet, fstart, fstop, npts, opts=None, max_pts=None):
    """Returns (freq[F], X[F, n] complex, newton_iterations)."""
    a = A.s21_ac_options(fstart, fstop, npts)
    mp = max_pts or (npts + 4)
    f = np.zeros(mp)
    X = np.zeros((mp, cnet.n_vars, 2))
    it = C.c_int64()
    err = C.create_string_buffer(512)
    r = lib().oracle_ac(cnet.ptr, C.byref(_opts(opts)), C.byref(a), mp, _dp(f), _dp(X), C.byref(it), err, 512)
    if r < 0:
        raise OracleError(-r, err.value.decode())
    r = min(r, mp)
    return f[:r], X[:r, :, 0] + 1j * X[:r, :, 1], it.value


def assemble_op(cnet, x, opts=None, max_elems=1 << 20):
    """One OP load at voltages x: (rows, cols, vals, rhs) over elements in creation order."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    rows = np.zeros(max_elems, dtype=np.int32)
    cols = np.zeros(max_elems, dtype=np.int32)
    vals = np.zeros(max_elems)
    rhs = np.zeros(max(cnet.n_vars, 1))
    err = C.create_string_buffer(512)
    ne = lib().oracle_assemble_op(cnet.ptr, C.byref(_opts(opts)), _dp(x), max_elems,
                                  rows.ctypes.data_as(C.POINTER(C.c_int32)), cols.ctypes.data_as(C.POINTER(C.c_int32)),
                                  _dp(vals), _dp(rhs), err, 512)
    if ne < 0:
        raise OracleError(-ne, err.value.decode())
    return rows[:ne], cols[:ne], vals[:ne], rhs[: cnet.n_vars]


def first_pivots(cnet, x, opts=None):
    """Row/col i2e maps after the first factorisation of the OP Jacobian at x, and nnz(L+U)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    ri = np.zeros(cnet.n_vars, dtype=np.int32)
    ci = np.zeros(cnet.n_vars, dtype=np.int32)
    nnz = C.c_int64()
    err = C.create_string_buffer(512)
    rc = lib().oracle_first_pivots(cnet.ptr, C.byref(_opts(opts)), _dp(x), ri.ctypes.data_as(C.POINTER(C.c_int32)),
                                   ci.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(nnz), err, 512)
    if rc != 0:
        raise OracleError(rc, err.value.decode())
    return ri, ci, nnz.value


def mos1_derive(model, inst, opts=None):
    out = np.zeros(33)
    err = C.create_string_buffer(512)
    rc = lib().oracle_mos1_derive(C.byref(model), C.byref(inst), C.byref(_opts(opts)), _dp(out), err, 512)
    if rc != 0:
        raise OracleError(rc, err.value.decode())
    return out


def diode_derive(model, inst, opts=None):
    out = np.zeros(12)
    lib().oracle_diode_derive(C.byref(model), C.byref(inst), C.byref(_opts(opts)), _dp(out))
    return out


class Matrix:
    """The reference's sparse21::Matrix, restated (for its known-answer tests)."""

    def __init__(self, cplx=False, entries=()):
        self.cplx = cplx
        self.h = lib().oracle_mat_new(1 if cplx else 0)
        for r, c, v in entries:
            self.add(r, c, v)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_mat_free(self.h)
            self.h = None

    def add(self, r, c, v):
        v = complex(v)
        lib().oracle_mat_add(self.h, r, c, v.real, v.imag)

    def factor(self):
        err = C.create_string_buffer(256)
        rc = lib().oracle_mat_factor(self.h, err, 256)
        if rc != 0:
            raise OracleError(rc, err.value.decode())

    def solve(self, rhs):
        n = len(rhs)
        err = C.create_string_buffer(256)
        if self.cplx:
            b = np.zeros(2 * n)
            b[0::2] = np.real(rhs)
            b[1::2] = np.imag(rhs)
            out = np.zeros(2 * n)
        else:
            b = np.array(rhs, dtype=np.float64)
            out = np.zeros(n)
        rc = lib().oracle_mat_solve(self.h, n, _dp(b), _dp(out), err, 256)
        if rc != 0:
            raise OracleError(rc, err.value.decode())
        return out[0::2] + 1j * out[1::2] if self.cplx else out

    def get(self, r, c):
        re, im = C.c_double(), C.c_double()
        ok = lib().oracle_mat_get(self.h, r, c, C.byref(re), C.byref(im))
        if not ok:
            return None
        return complex(re.value, im.value) if self.cplx else re.value

    def swap_rows(self, x, y):
        lib().oracle_mat_swap_rows(self.h, x, y)

    def row_mapping(self, n):
        e2i = np.zeros(n, dtype=np.int64)
        i2e = np.zeros(n, dtype=np.int64)
        lib().oracle_mat_row_mapping(self.h, n, e2i.ctypes.data_as(C.POINTER(C.c_int64)),
                                     i2e.ctypes.data_as(C.POINTER(C.c_int64)))
        return list(e2i), list(i2e)

    def check(self):
        return lib().oracle_mat_check(self.h)


class planted:
    """Context manager: the oracle runs with planted mistake ``k`` (see g_plant in oracle.cpp), so that a
    test can show that its check fails on that mistake."""

    def __init__(self, k):
        self.k = int(k)

    def __enter__(self):
        lib().oracle_plant(self.k)
        return self

    def __exit__(self, *exc):
        lib().oracle_plant(0)
        return False


def pivot_stats(reset=True):
    """(factorisations, factorisations whose pivot search chose other elements than the previous
    factorisation of the same matrix) since the last reset."""
    f, c = C.c_longlong(), C.c_longlong()
    lib().oracle_pivot_stats(C.byref(f), C.byref(c), int(reset))
    return f.value, c.value
