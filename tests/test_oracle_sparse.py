"""The oracle's sparse matrix against the reference's sparse21 unit tests (sparse21/mod.rs:1356-1526)
and against dense LAPACK solves on random systems."""
import numpy as np
import pytest
import scipy.linalg

from oracle import oracle_py as O


def test_solve_known_answer():  # sparse21/mod.rs:1450-1471
    m = O.Matrix(entries=[(0, 0, 1.0), (0, 1, 1.0), (0, 2, 1.0), (1, 1, 2.0), (1, 2, 5.0), (2, 0, 2.0), (2, 1, 5.0),
                          (2, 2, -1.0)])
    assert m.check() == 0
    m.factor()
    assert m.check() == 0
    x = m.solve([6.0, -4.0, 27.0])
    assert np.all(np.abs(x - np.array([5.0, 3.0, -2.0])) < 1e-9)


def test_solve_identity():  # :1473-1495
    for s in range(1, 10):
        m = O.Matrix(entries=[(k, k, 1.0) for k in range(s)])
        rhs = [float(e) for e in range(s)]
        assert list(m.solve(rhs)) == rhs
    m = O.Matrix(entries=[(k, k, 1.0) for k in range(3)])
    assert list(m.solve([11.1, 30.3, 99.9])) == [11.1, 30.3, 99.9]


def test_lu_identity_and_lower():  # :1377-1419
    m = O.Matrix(entries=[(k, k, 1.0) for k in range(3)])
    m.factor()
    assert [m.get(k, k) for k in range(3)] == [1.0, 1.0, 1.0]
    ent = [(0, 0, 1.0), (1, 0, 1.0), (2, 0, 1.0), (1, 1, 1.0), (2, 1, 1.0), (2, 2, 1.0)]
    m = O.Matrix(entries=ent)
    m.factor()
    assert m.check() == 0
    for r, c, v in ent:
        assert m.get(r, c) == v


def test_row_mappings():  # :1356-1375
    m = O.Matrix(entries=[(k, k, 1.0) for k in range(4)])
    m.swap_rows(0, 3)
    assert m.check() == 0
    assert m.row_mapping(4) == ([3, 1, 2, 0], [3, 1, 2, 0])
    m.swap_rows(0, 2)
    assert m.check() == 0
    assert m.row_mapping(4) == ([3, 1, 0, 2], [2, 1, 3, 0])


def test_complex_known_answers():  # :1501-1526
    m = O.Matrix(cplx=True, entries=[(0, 0, 1.0), (1, 1, 1.0)])
    assert list(m.solve([1j, 1j])) == [1j, 1j]
    m = O.Matrix(cplx=True, entries=[(0, 0, 1.0), (1, 0, -1.0), (0, 1, -1.0), (1, 1, 1.0 + 1.0j)])
    assert list(m.solve([1.0, 0.0])) == [1.0 - 1.0j, -1.0j]


def test_singular_errors():
    m = O.Matrix(entries=[(0, 0, 1.0), (2, 2, 1.0), (0, 1, 1.0)])  # row 1 empty -> Singular (mod.rs:649-653)
    with pytest.raises(O.OracleError, match="Singular Matrix"):
        m.solve([1.0, 2.0, 3.0])
    m = O.Matrix(entries=[(0, 0, 1.0), (1, 0, 1.0)])  # 2x1: the reference panics on the column unwind
    with pytest.raises(O.OracleError, match="index out of bounds"):
        m.solve([1.0, 2.0])
    # numerically singular: the last index is never pivot-searched (mod.rs:663), so its zero pivot is
    # divided by in the backward sweep and the solution is non-finite instead of an error
    m = O.Matrix(entries=[(0, 0, 1.0), (0, 1, 1.0), (1, 0, 1.0), (1, 1, 1.0)])
    x = m.solve([1.0, 2.0])
    assert not np.all(np.isfinite(x))
    # a zero pivot before the last index is caught by row_col_elim's assertion (mod.rs:872)
    m = O.Matrix(entries=[(0, 0, 0.0), (0, 1, 0.0), (1, 0, 0.0), (1, 1, 0.0), (2, 2, 1.0)])
    with pytest.raises(O.OracleError):
        m.solve([1.0, 2.0, 3.0])


@pytest.mark.parametrize("seed", range(12))
def test_random_sparse_vs_lapack(seed):
    """Brute force: random sparse, diagonally weighted systems against a dense LAPACK solve.  A dropped
    fill-in term, a wrong permutation unwind or a transposed operand fails this."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 40))
    cplx = seed % 3 == 2
    dense = np.zeros((n, n), dtype=complex if cplx else float)
    ents = []
    for r in range(n):
        for c in range(n):
            if r == c or rng.random() < 0.15:
                v = rng.normal() + (1j * rng.normal() if cplx else 0.0)
                if r == c and seed % 2 == 0:
                    v += 4.0
                dense[r, c] = v
                ents.append((r, c, v))
    if abs(np.linalg.det(dense)) < 1e-8:
        pytest.skip("near-singular draw")
    order = rng.permutation(len(ents))
    m = O.Matrix(cplx=cplx, entries=[ents[k] for k in order])
    b = rng.normal(size=n) + (1j * rng.normal(size=n) if cplx else 0.0)
    x = m.solve(b)
    assert m.check() == 0 if not cplx else True
    xr = scipy.linalg.solve(dense, b)
    assert np.max(np.abs(x - xr)) <= 1e-8 * max(1.0, np.max(np.abs(xr)))
    assert np.max(np.abs(dense @ x - b)) <= 1e-9 * max(1.0, np.max(np.abs(b))) * n
