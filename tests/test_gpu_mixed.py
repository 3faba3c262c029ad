"""Mixed-precision path: the anchored low-precision product against an
FP64 product, the device generator against the oracle's, and mixed block
solves against the FP64 block solve and LAPACK."""
import numpy as np
import pytest
import scipy.linalg as sl

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

from block_ipt_amd import api  # noqa: E402


@pytest.mark.parametrize("mode,tol", [("fp16", 3e-3), ("bf16", 2e-2), ("bf16x3", 1e-4)])
@pytest.mark.parametrize("n,nc", [(300, 7), (517, 40), (1024, 130)])
def test_lowp_product_error_bound(mode, tol, n, nc):
    """Y = Ya + D(B-Ba) + E(B-Ba): the low-precision term's error relative
    to ||E (B-Ba)|| stays at the input precision (unit roundoff 2^-11 FP16,
    2^-8 BF16, ~2^-16 BF16x3, times a small accumulation factor)."""
    hip = api.hip()
    H = orc.rand_dense(n, 1, 0.0, 10.0, 0.1)
    rng = np.random.default_rng(n + nc)
    Ba = np.asfortranarray(rng.standard_normal((n, nc)))
    B = np.asfortranarray(Ba + 1e-3 * rng.standard_normal((n, nc)))
    Ya = np.asfortranarray(H @ Ba)
    d = np.ascontiguousarray(np.diag(H))
    bufs = [hip.to_device(np.asfortranarray(x)) for x in (H, B, Ba, Ya, d)]
    dY = hip.malloc(n * nc * 8)
    try:
        api.lowp_product(bufs[0], n, mode, bufs[1], bufs[2], bufs[3], bufs[4], nc, dY)
        Y = hip.to_host(dY, (n, nc))
    finally:
        for p in bufs + [dY]:
            hip.free(p)
    E = H - np.diag(d)
    exact = H @ B
    lowp_term = E @ (B - Ba)
    err = np.linalg.norm(Y - exact, axis=0) / np.linalg.norm(lowp_term, axis=0)
    assert err.max() <= tol, err.max()


def test_device_generator_matches_oracle():
    hip = api.hip()
    n = 700
    dH = hip.malloc(n * n * 8)
    try:
        api.generate_dense(dH, n, 3, 0.0, 50.0, 0.1)
        G = hip.to_host(dH, (n, n))
    finally:
        hip.free(dH)
    O = orc.rand_dense(n, 3, 0.0, 50.0, 0.1)
    assert np.array_equal(np.diag(G), np.diag(O))
    assert np.abs(G - O).max() <= 1e-15 * np.abs(O).max()
    assert np.array_equal(G, G.T)


@pytest.fixture
def mixed_env(monkeypatch):
    monkeypatch.setenv("IPT_BLOCK_CLUSTER", "1")
    for v in ("IPT_DAVIDSON_ENRICH", "IPT_BLOCK_CLUSTER_MAXITER", "IPT_BLOCK_CLUSTER_TOL", "IPT_PRECISION",
              "IPT_MIXED_ANCHOR_EVERY"):
        monkeypatch.delenv(v, raising=False)
    return monkeypatch


@pytest.mark.parametrize("mode", ["fp16", "bf16", "bf16x3"])
def test_mixed_block_solve_matches_fp64(mixed_env, mode):
    """C1 recipe (N=2048, k=16, d ~ U[0,10], V scale 0.05/sqrt(N))."""
    n, k = 2048, 16
    H = orc.rand_dense(n, 0, 0.0, 10.0, 0.05)
    v64, l64, i64 = api.ipt_cuda_tol(H, k, 1e-12, 400)
    mixed_env.setenv("IPT_PRECISION", mode)
    mixed_env.setenv("IPT_MIXED_ANCHOR_EVERY", "10")
    vm, lm, im = api.ipt_cuda_tol(H, k, 1e-12, 400)
    assert np.abs(lm - l64).max() <= 1e-10 * max(1.0, np.abs(l64).max())
    w = sl.eigh(H, eigvals_only=True, subset_by_index=[0, k - 1])
    assert np.abs(lm - w).max() <= 1e-10 * max(1.0, np.abs(w).max())
    res = orc.residuals(vm, lm, dense=H)
    assert res.max() <= 1e-10
    # anchored steps converge like the FP64 iteration (a broken product would not)
    assert im["iterations"] <= 2 * i64["iterations"] + 12
