"""CPU check of the host generalized symmetric eigensolver used by the
block path's Rayleigh-Ritz (block_ipt_amd/csrc/ipt_small_eig.h) against
LAPACK (scipy.linalg.eigh(P, G)), including near-degenerate clusters."""
import os
import subprocess

import numpy as np
import pytest
import scipy.linalg as sl

from conftest import ROOT


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("small_eig")
    exe = d / "small_eig"
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "cpp", "small_eig_driver.cpp"),
                    "-o", str(exe)], check=True)
    return exe, d


def _run(driver, G, P, need=None):
    exe, d = driver
    m = G.shape[0]
    inp, out = d / "in.bin", d / "out.bin"
    with open(inp, "wb") as f:
        np.array([m], dtype=np.int32).tofile(f)
        np.asfortranarray(G).ravel(order="F").tofile(f)
        np.asfortranarray(P).ravel(order="F").tofile(f)
    subprocess.run([str(exe), str(inp), str(out)] + ([] if need is None else [str(need)]), check=True)
    raw = np.fromfile(out, dtype=np.uint8)
    st = int(raw[:4].view(np.int32)[0])
    vals = raw[4:].view(np.float64)
    return st, vals[:m], vals[m:].reshape((m, m), order="F")


def _case(m, seed, cluster=False):
    rng = np.random.default_rng(seed)
    V = rng.standard_normal((5 * m, m))
    if cluster:  # nearly dependent basis and clustered spectrum, like cluster bases
        V[:, 1] = V[:, 0] + 1e-4 * V[:, 1]
    G = V.T @ V
    lam = np.sort(rng.uniform(0, 10, m))
    if cluster:
        lam[2:6] = lam[2] + 1e-9 * np.arange(4)
    Q, _ = np.linalg.qr(rng.standard_normal((m, m)))
    Asym = Q @ np.diag(lam) @ Q.T
    L = np.linalg.cholesky(G)
    P = L @ Asym @ L.T
    return G, 0.5 * (P + P.T)


@pytest.mark.parametrize("m,seed,cluster", [(1, 0, False), (2, 1, False), (7, 2, False), (70, 3, False),
                                            (70, 4, True), (173, 5, True)])
def test_matches_lapack(driver, m, seed, cluster):
    G, P = _case(m, seed, cluster)
    st, theta, C = _run(driver, G, P)
    assert st == 0
    w, W = sl.eigh(P, G)
    scale = max(1.0, np.abs(w).max())
    assert np.abs(theta - w).max() <= 1e-10 * scale * max(1.0, np.linalg.cond(G) * 1e-6)
    assert np.all(np.diff(theta) >= 0)
    # G-orthonormal eigenvectors with small residual; a Cholesky-based
    # method loses accuracy with cond(G), so the bound is LAPACK's own error
    # on the same input (times a small factor) or 1e-9, whichever is larger
    orth_lapack = np.abs(W.T @ G @ W - np.eye(m)).max()
    assert np.abs(C.T @ G @ C - np.eye(m)).max() <= max(1e-9, 20 * orth_lapack)
    res_lapack = np.abs(P @ W - G @ W * w).max()
    R = P @ C - G @ C * theta
    assert np.abs(R).max() <= max(1e-9 * scale * np.abs(G).max(), 20 * res_lapack)


def test_not_positive_definite(driver):
    G = np.diag([1.0, -1.0, 2.0])
    P = np.eye(3)
    st, _, _ = _run(driver, G, P)
    assert st == 1


def test_identity_metric_is_plain_eigh(driver):
    rng = np.random.default_rng(9)
    A = rng.standard_normal((40, 40))
    A = A + A.T
    st, theta, C = _run(driver, np.eye(40), A)
    assert st == 0
    w, V = np.linalg.eigh(A)
    assert np.abs(theta - w).max() <= 1e-12 * np.abs(w).max()
    assert np.abs(np.abs(np.sum(C * V, axis=0)) - 1.0).max() <= 1e-10


@pytest.mark.parametrize("m,dup", [(12, 1), (60, 3)])
def test_dependent_basis_falls_back_to_canonical_orthogonalisation(driver, m, dup):
    """A basis with exactly repeated columns (G singular): the rank-revealing
    variant returns the Ritz pairs of the spanned subspace (LAPACK eigh of
    the projection onto an orthonormal basis of it), with V C orthonormal."""
    rng = np.random.default_rng(m)
    n = 5 * m
    A = rng.standard_normal((n, n))
    A = A + A.T
    V = rng.standard_normal((n, m))
    for j in range(dup):
        V[:, m - 1 - j] = V[:, j]
    G = V.T @ V
    P = V.T @ A @ V
    st, _, _ = _run(driver, G, P)
    assert st == 1
    r = m - dup
    st, theta, C = _run(driver, G, P, need=r)
    assert st == 0
    Qv, _ = np.linalg.qr(V[:, :r])
    w = np.linalg.eigvalsh(Qv.T @ A @ Qv)
    assert np.abs(theta[:r] - w).max() <= 1e-9 * np.abs(w).max()
    X = V @ C[:, :r]
    assert np.abs(X.T @ X - np.eye(r)).max() <= 1e-9
    # Galerkin condition: the residuals are orthogonal to the subspace
    assert np.abs(Qv.T @ (A @ X - X * theta[:r])).max() <= 1e-9 * np.abs(w).max()
    # more pairs than the rank: an error
    st, _, _ = _run(driver, G, P, need=r + 1)
    assert st == 1
