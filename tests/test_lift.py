"""Degeneracy lifting of the plain CSC route (host preparation inside the
library, no GPU): the prepared matrix, permutation, rotation count and the
map back to the caller's basis equal those of the reference's own
preparation code (IPT/src/ipt_preparation.cu, built from source into
oracle/_ref)."""
import ctypes

import numpy as np
import pytest

from block_ipt_amd import abi
from oracle import oracle as orc


def degenerate_csc(n, seed, shuffle=False, symmetric=True, coupled=True):
    """Symmetric sparse matrix whose diagonal holds exactly repeated values
    (groups of 1-3 equal entries) with small couplings."""
    rng = np.random.default_rng(seed)
    levels = np.repeat(np.arange(1, n + 1, dtype=float), rng.integers(1, 4, n))[:n]
    d = levels * 0.5
    A = np.diag(d)
    # members of a degenerate group are coupled (1e-2), so the lift splits them
    for lv in (np.unique(levels) if coupled else []):
        g = np.nonzero(levels == lv)[0]
        for a in g:
            for b in g:
                if a < b:
                    A[a, b] = A[b, a] = 1e-2 * rng.uniform(0.5, 1.0)
    for _ in range(4 * n):
        i, j = rng.integers(0, n, 2)
        if i != j:
            x = 1e-3 * rng.standard_normal()
            A[i, j] = x
            A[j, i] = x if symmetric else 1e-3 * rng.standard_normal()
    if shuffle:
        p = rng.permutation(n)
        A = A[np.ix_(p, p)]
    cp = [0]
    ri, v = [], []
    for c in range(n):
        nz = np.nonzero(A[:, c])[0]
        ri += list(nz)
        v += list(A[nz, c])
        cp.append(len(ri))
    return (np.array(cp, dtype=np.int32), np.array(ri, dtype=np.int32), np.array(v)), A


def ours(csc, k, thr=0.0):
    cp, ri, v = csc
    n = len(cp) - 1
    dense = np.zeros((n, n), order="F")
    perm = np.zeros(n, dtype=np.int32)
    rot = ctypes.c_int(0)
    st = abi.lib().ipt_hip_prepare_sparse_csc(cp.ctypes.data_as(abi._ip), ri.ctypes.data_as(abi._ip),
                                              v.ctypes.data_as(abi._dp), n, k, int(cp[-1]), thr,
                                              dense.ctypes.data_as(abi._dp), perm.ctypes.data_as(abi._ip),
                                              ctypes.byref(rot))
    return st, dense, perm, rot.value


def ours_back(csc, k, X, thr=0.0):
    cp, ri, v = csc
    X = np.array(X, dtype=np.float64, order="F", copy=True)
    st = abi.lib().ipt_hip_prepared_vectors_to_original(cp.ctypes.data_as(abi._ip), ri.ctypes.data_as(abi._ip),
                                                        v.ctypes.data_as(abi._dp), len(cp) - 1, k, int(cp[-1]),
                                                        thr, X.ctypes.data_as(abi._dp), X.shape[1])
    assert st == 0
    return X


needs_ref = pytest.mark.skipif(orc.ref_prepare_sparse_csc(*degenerate_csc(8, 0)[0], 2) is None,
                               reason="oracle/_ref (the reference's preparation built from source) is absent")


@needs_ref
@pytest.mark.parametrize("n,k,seed,shuffle", [(40, 8, 1, False), (60, 20, 2, True), (120, 30, 3, True),
                                              (300, 60, 4, True), (500, 90, 6, False)])
def test_prepared_matrix_equals_reference(n, k, seed, shuffle):
    csc, A = degenerate_csc(n, seed, shuffle)
    st, D, perm, rot = ours(csc, k)
    rst, RD, rperm, rrot = orc.ref_prepare_sparse_csc(*csc, k)
    assert st == 0 and rst == 0
    assert rot == rrot and rot >= 1
    assert np.array_equal(perm, rperm)
    assert np.abs(D - RD).max() <= 1e-14 * np.abs(A).max()
    # a similarity: same spectrum as the input
    assert np.allclose(np.linalg.eigvalsh(0.5 * (D + D.T)), np.linalg.eigvalsh(A), atol=1e-12)
    X = np.random.default_rng(seed).standard_normal((n, 5))
    assert np.abs(ours_back(csc, k, X) - orc.ref_prepared_vectors_to_original(*csc, k, X)).max() <= 1e-13


def test_lift_separates_exact_degeneracies():
    """After the lift no target pair is joined any more, and mapping an
    eigenvector of the prepared matrix back gives one of the input's."""
    csc, A = degenerate_csc(50, 5)
    st, D, perm, rot = ours(csc, 10)
    assert st == 0 and rot >= 1
    w, V = np.linalg.eigh(0.5 * (D + D.T))
    X = ours_back(csc, 10, V[:, :3])
    assert np.abs(A @ X - X * w[:3]).max() <= 1e-12


@needs_ref
def test_status_agrees_with_reference():
    """Success or failure of the preparation (an uncoupled exact degeneracy
    cannot be lifted) agrees with the reference's on 40 inputs."""
    for seed in range(10):
        for n, k in ((40, 8), (90, 25)):
            for coupled in (True, False):
                csc, A = degenerate_csc(n, seed, shuffle=seed % 2 == 1, coupled=coupled)
                assert ours(csc, k)[0] == orc.ref_prepare_sparse_csc(*csc, k)[0]


@needs_ref
def test_non_symmetric_needing_lift_fails_like_reference():
    csc, _ = degenerate_csc(30, 7, symmetric=False)
    st = ours(csc, 10)[0]
    rst = orc.ref_prepare_sparse_csc(*csc, 10)[0]
    assert st == 1 and rst == 1


def test_no_degeneracy_is_a_pure_sort():
    rng = np.random.default_rng(0)
    n = 30
    d = rng.permutation(np.arange(n, dtype=float))
    A = np.diag(d)
    cp = np.arange(n + 1, dtype=np.int32)
    ri = np.arange(n, dtype=np.int32)
    st, D, perm, rot = ours((cp, ri, d.copy()), 5)
    assert st == 0 and rot == 0
    assert np.array_equal(np.diag(D), np.sort(d))
    assert np.array_equal(d[perm], np.sort(d))


@needs_ref
def test_duplicate_entries_summed_like_reference():
    """A CSC input that stores some entries twice (split values): the
    prepared matrix equals the reference's, which sums duplicates."""
    (cp, ri, v), A = degenerate_csc(60, 8, shuffle=True)
    rng = np.random.default_rng(8)
    cols = [list(zip(ri[cp[c]:cp[c + 1]], v[cp[c]:cp[c + 1]])) for c in range(60)]
    for c in range(60):
        out = []
        for r, x in cols[c]:
            if r != c and rng.uniform() < 0.3:
                f = rng.uniform(0.2, 0.8)
                out += [(r, f * x), (r, (1 - f) * x)]
            else:
                out.append((r, x))
        cols[c] = out
    cp2 = np.array([0] + list(np.cumsum([len(e) for e in cols])), dtype=np.int32)
    ri2 = np.array([r for e in cols for r, _ in e], dtype=np.int32)
    v2 = np.array([x for e in cols for _, x in e])
    st, D, perm, rot = ours((cp2, ri2, v2), 15)
    rst, RD, rperm, rrot = orc.ref_prepare_sparse_csc(cp2, ri2, v2, 15)
    assert st == rst == 0 and rot == rrot and np.array_equal(perm, rperm)
    assert np.abs(D - RD).max() <= 1e-14 * np.abs(A).max()
