"""The single-large-instance path's host phase (no GPU): its static pivoting, fill-reducing order and
multifrontal factorisation, through the CPU reference of the numeric phase the GPU kernels mirror, against
dense LAPACK solves; and the instance ABI's argument checks."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle_py as O
from spice21_amd import _abi as A
from spice21_amd import large as LG
from spice21_amd import workloads as W
from spice21_amd._lib import lib


def dense(n, r, c, v):
    M = np.zeros((n, n))
    np.add.at(M, (np.asarray(r), np.asarray(c)), np.asarray(v))
    return M


@pytest.mark.parametrize("size,n_inv,n_diode,pad", [(6, 4, 2, 3), (12, 20, 8, 4), (20, 60, 20, 5)])
def test_power_grid_matrix_solves_like_lapack(size, n_inv, n_diode, pad):
    """The Jacobian of a small power grid (mesh, inverters, diodes, pad sources whose branch rows have no
    diagonal) at a random bias: the large path's solve matches numpy's LU solve to the matrix's
    conditioning, and its own residual is at rounding level."""
    nl = W.power_grid(size=size, n_inv=n_inv, n_diode=n_diode, pad_every=pad)
    cn = nl.to_c()
    n = cn.n_vars
    rng = np.random.default_rng(size)
    r, c, v, _ = O.assemble_op(cn, rng.uniform(0.0, 1.0, n))
    b = rng.normal(size=n)
    x, nnz_lu, flops = LG.host_check(n, r, c, v, b)
    M = dense(n, r, c, v)
    assert np.abs(M @ x - b).max() <= 1e-12 * max(1.0, np.abs(b).max()) * np.abs(M).max()
    xr = np.linalg.solve(M, b)
    assert np.abs(x - xr).max() <= 1e-8 * max(1.0, np.abs(xr).max())
    assert nnz_lu >= len(set(zip(r.tolist(), c.tolist()))) and flops > 0


@pytest.mark.parametrize("seed", range(6))
def test_random_sparse_systems(seed):
    """Random sparse systems with a dominant diagonal, some rows whose diagonal is absent (they need a
    static row exchange, as a voltage source's branch row does), and random long-range couplings."""
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(30, 300))
    rows, cols, vals = [], [], []
    for i in range(n):
        rows.append(i)
        cols.append(i)
        vals.append(5.0 + rng.uniform())
    for _ in range(4 * n):
        i, j = rng.integers(0, n, 2)
        rows.append(int(i))
        cols.append(int(j))
        vals.append(rng.uniform(-1, 1))
    # two source-like pairs: rows p_k / branch b_k coupled by +-1, the branch row with no diagonal
    for k in range(2):
        p, br = int(rng.integers(0, n)), n + k
        rows += [p, br]
        cols += [br, p]
        vals += [1.0, 1.0]
    n2 = n + 2
    b = rng.normal(size=n2)
    for relax in (0.0, 0.1, 0.3):
        x, _, _ = LG.host_check(n2, rows, cols, vals, b, relax=relax)
        M = dense(n2, rows, cols, vals)
        assert np.abs(M @ x - b).max() <= 1e-10 * np.abs(M).max() * max(1.0, np.abs(x).max()), (seed, relax)


def test_structurally_singular_matrix_is_reported():
    # column 2 empty: no transversal exists
    rows, cols, vals = [0, 1, 0], [0, 1, 1], [1.0, 2.0, 3.0]
    with pytest.raises(Exception, match="Singular Matrix"):
        LG.host_check(3, rows, cols, vals, np.ones(3))


def test_fill_reducing_order_at_full_size():
    """The 200 x 200 power grid (45,040 unknowns): the host symbolic phase orders it in well under a
    second of CPU (the reference's Markowitz search would take over an hour) and keeps L+U within 30 M
    entries; checked through a solve of a system on its pattern whose answer is known."""
    nl = W.power_grid()
    cn = nl.to_c()
    n = cn.n_vars
    r, c, v, _ = O.assemble_op(cn, np.zeros(n), max_elems=1 << 22)
    # a diagonally dominant system on the same pattern (the host factorisation of the real one takes
    # minutes of CPU): x_true = 1
    rr, cc = np.asarray(r), np.asarray(c)
    vv = np.where(rr == cc, 8.0, -1.0 / 8.0)
    # voltage-source branch rows have no diagonal: give their coupling entries unit weight
    kinds = np.asarray(nl.kinds)
    vv = np.where((kinds[rr] == A.S21_VAR_I) | (kinds[cc] == A.S21_VAR_I), 1.0, vv)
    xt = np.ones(n)
    b = np.zeros(n)
    np.add.at(b, rr, vv * xt[cc])
    import time

    t0 = time.perf_counter()
    x, nnz_lu, flops = LG.host_check(n, r, c, vv, b)
    assert time.perf_counter() - t0 < 600
    assert nnz_lu < 30_000_000
    assert np.abs(x - xt).max() < 1e-8


def test_large_abi_rejects_bad_arguments():
    assert lib().s21_large_op(None) == A.S21_INVALID
    assert lib().s21_large_get_x(None, None) == A.S21_INVALID
    assert lib().s21_large_n_points(None) == 0
    h = C.c_void_p()
    assert lib().s21_large_create(None, None, 0, C.byref(h)) == A.S21_INVALID
