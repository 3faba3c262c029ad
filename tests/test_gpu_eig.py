"""The device symmetric eigensolver (ipt_hip_sym_eig, block Jacobi) against
LAPACK: random, nearly diagonal (the projected problems of Davidson) and
exactly degenerate spectra, orders from 1 to 800 (the block padding and the
largest Davidson search space of the k = 256 workload)."""
import numpy as np
import pytest
import scipy.linalg as sl

pytestmark = pytest.mark.gpu

from block_ipt_amd import api  # noqa: E402


def _check(A, w, V, tol=1e-12):
    wl = sl.eigvalsh(A)
    scale = max(1.0, np.abs(wl).max())
    assert np.abs(w - wl).max() <= tol * scale, np.abs(w - wl).max()
    m = A.shape[0]
    assert np.abs(V.T @ V - np.eye(m)).max() <= 1e-12
    R = A @ V - V * w
    assert np.abs(R).max() <= 1e-12 * scale


@pytest.mark.parametrize("m", [1, 2, 5, 31, 32, 33, 64, 70, 127, 256, 427, 800])
def test_random_symmetric(m):
    rng = np.random.default_rng(m)
    A = rng.standard_normal((m, m))
    A = 0.5 * (A + A.T)
    print("sym_eig random m=%d" % m, flush=True)
    w, V = api.sym_eig(A)
    _check(A, w, V)


@pytest.mark.parametrize("m", [96, 300, 768])
def test_nearly_diagonal(m):
    """Ritz values on the diagonal, small couplings to new directions: the
    structure of Davidson's projected problems."""
    rng = np.random.default_rng(7 + m)
    A = np.diag(np.sort(rng.uniform(0.0, 1.0, m)))
    B = 1e-4 * rng.standard_normal((m, m))
    A = A + 0.5 * (B + B.T)
    w, V = api.sym_eig(A)
    _check(A, w, V)


def test_degenerate_spectrum():
    """Exactly repeated eigenvalues (multiplicities 1-5): the vectors span each
    eigenspace and stay orthonormal."""
    rng = np.random.default_rng(3)
    lam = np.repeat(np.arange(80, dtype=float), rng.integers(1, 6, 80))[:150]
    assert lam.size == 150
    Q, _ = np.linalg.qr(rng.standard_normal((150, 150)))
    A = (Q * lam) @ Q.T
    A = 0.5 * (A + A.T)
    w, V = api.sym_eig(A)
    _check(A, w, V, tol=1e-12)
