"""FP32-preconditioned Cholesky QR (ipt_hip_cholqr_mixed) checked by the
invariants that fix it: Q^T Q = I to FP64 accuracy and span(Q) = span(Z),
for bases with condition numbers up to 1e9 (beyond what a plain FP64
Cholesky QR of the Gram matrix, kappa^2 > 1/eps, can do); and the
Rayleigh-mixed entry with IPT_RAYLEIGH_QR=mpqr against its Householder
version."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

from block_ipt_amd import api  # noqa: E402


def _basis(n, k, kappa, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((n, k)))
    V, _ = np.linalg.qr(rng.standard_normal((k, k)))
    s = np.logspace(0, -np.log10(kappa), k)
    return np.asfortranarray(U @ np.diag(s) @ V.T)


@pytest.mark.parametrize("n,k,kappa", [(2000, 16, 1e2), (5000, 64, 1e6), (3000, 40, 1e9), (700, 700, 1e3)])
def test_orthonormal_and_same_span(n, k, kappa):
    hip = api.hip()
    Z = _basis(n, k, kappa, n + k)
    dQ = hip.to_device(Z.ravel(order="F"))
    try:
        api.cholqr_mixed(dQ, n, k)
        Q = hip.to_host(dQ, (n, k))
    finally:
        hip.free(dQ)
    assert np.abs(Q.T @ Q - np.eye(k)).max() <= 1e-12
    R = Z - Q @ (Q.T @ Z)
    assert np.linalg.norm(R) <= 1e-9 * np.linalg.norm(Z)
    # plain Cholesky QR in FP64 would lose orthogonality at this conditioning
    if kappa >= 1e9:
        G = Z.T @ Z
        try:
            L = np.linalg.cholesky(G)
            Qp = np.linalg.solve(L, Z.T).T
            assert np.abs(Qp.T @ Qp - np.eye(k)).max() > 1e-6
        except np.linalg.LinAlgError:
            pass


def test_zero_column_is_rejected():
    hip = api.hip()
    Z = _basis(500, 8, 10.0, 1)
    Z[:, 3] = 0.0
    dQ = hip.to_device(Z.ravel(order="F"))
    try:
        with pytest.raises(api.IptError):
            api.cholqr_mixed(dQ, 500, 8)
    finally:
        hip.free(dQ)


def test_rayleigh_mixed_with_mpqr(monkeypatch):
    for v in ("IPT_BLOCK_CLUSTER", "IPT_PRECISION"):
        monkeypatch.delenv(v, raising=False)
    hip = api.hip()
    n, k = 1024, 24
    H = orc.eps_dense(n, 0.01, 20260603 + n)
    dH = hip.to_device(np.asfortranarray(H))
    try:
        monkeypatch.delenv("IPT_RAYLEIGH_QR", raising=False)
        v1, l1, i1 = api.mixed_device_tol("rayleigh_tf32", dH, n, k, 1e-6, 100, 1e-12, 100)
        monkeypatch.setenv("IPT_RAYLEIGH_QR", "mpqr")
        v2, l2, i2 = api.mixed_device_tol("rayleigh_tf32", dH, n, k, 1e-6, 100, 1e-12, 100)
    finally:
        hip.free(dH)
    assert np.abs(np.sort(l1) - np.sort(l2)).max() <= 1e-10 * max(1.0, np.abs(l1).max())
    assert np.abs(v2.T @ v2 - np.eye(k)).max() <= 1e-10
