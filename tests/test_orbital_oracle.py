"""The oracle's orbital-sum sparse generator (the N = 1e6 sparse workload's
recipe, no reference counterpart): structural invariants on small sizes --
symmetric, sorted diagonal made of nsub-orbital sums, exactly m distinct
drawn columns per row, ascending column indices, N(0, sigma^2) values."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
from scipy import stats

from oracle import oracle as orc


def _csr(n, m, seed, **kw):
    rp, ci, v = orc.orbital_csr(n, m, seed, **kw)
    return rp, ci, v, sp.csr_matrix((v, ci, rp), shape=(n, n))


@pytest.mark.parametrize("n,m,seed", [(400, 12, 1), (1500, 30, 7), (60, 59, 3), (2, 1, 0)])
def test_structure(n, m, seed):
    rp, ci, v, A = _csr(n, m, seed)
    assert rp[0] == 0 and rp[-1] == ci.size and np.all(np.diff(rp) >= 1)
    for i in range(n):
        row = ci[rp[i]:rp[i + 1]]
        assert np.all(np.diff(row) > 0)  # ascending, distinct
        assert i in row
        # at least the m own draws plus the diagonal
        assert row.size >= m + 1
    D = A.toarray()
    assert np.array_equal(D, D.T)
    d = np.diag(D)
    assert np.all(np.diff(d) >= 0)
    # nnz = n + 2 * (#distinct unordered pairs); a pair drawn from both ends counts once
    off = np.count_nonzero(D - np.diag(d))
    assert ci.size == n + off
    assert n * m <= off <= 2 * n * m


def test_diagonal_is_sorted_orbital_sums():
    """Every diagonal value is a sum of 8 distinct orbital energies: with 10
    orbitals (C(10,8) = 45 possible sums) all values lie in that set."""
    n = 300
    d = orc.orbital_diagonal(n, 5, norb=10, nsub=8)
    # recover the orbitals from the 1-orbital model with the same n and seed:
    # its diagonals are the orbitals themselves (all 10 appear among 300 rows)
    e = np.unique(orc.orbital_diagonal(n, 5, norb=10, nsub=1))
    assert e.size == 10
    sums = np.array([sum(c) for c in itertools.combinations(e, 8)])
    for x in d:
        assert np.min(np.abs(sums - x)) <= 1e-12 * max(1.0, abs(x))
    assert np.all(np.diff(d) >= 0)
    assert np.all((e >= -20.0) & (e <= 5.0))


def test_off_diagonal_values_are_normal():
    n, m = 3000, 20
    rp, ci, v, A = _csr(n, m, 11, sigma=0.02)
    rows = np.repeat(np.arange(n), np.diff(rp))
    off = v[rows != ci]
    assert abs(off.mean()) <= 4 * 0.02 / np.sqrt(off.size)
    assert abs(off.std() / 0.02 - 1.0) <= 0.02
    assert stats.kstest(off / 0.02, "norm").pvalue > 1e-3


def test_columns_are_uniform():
    """Drawn columns spread uniformly over [0, n) (chi-square over 20 bins)."""
    n, m = 4000, 25
    rp, ci, v, A = _csr(n, m, 2)
    rows = np.repeat(np.arange(n), np.diff(rp))
    cols = ci[rows != ci]
    h, _ = np.histogram(cols, bins=20, range=(0, n))
    assert stats.chisquare(h).pvalue > 1e-3


def test_seed_changes_matrix_and_is_deterministic():
    a = orc.orbital_csr(500, 10, 1)
    b = orc.orbital_csr(500, 10, 1)
    c = orc.orbital_csr(500, 10, 2)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[1], c[1])
