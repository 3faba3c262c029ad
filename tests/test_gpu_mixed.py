"""Mixed-precision path: the anchored low-precision product against an
FP64 product, the device generator against the oracle's, and mixed block
solves against the FP64 block solve and LAPACK."""
import numpy as np
import pytest
import scipy.linalg as sl

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

from block_ipt_amd import api  # noqa: E402


@pytest.mark.parametrize("mode,tol", [("fp16", 3e-3), ("bf16", 2e-2), ("bf16x3", 1e-4), ("fp32", 1e-5),
                                      ("fp8", 0.1)])
@pytest.mark.parametrize("n,nc", [(300, 7), (517, 40), (1024, 130)])
def test_lowp_product_error_bound(mode, tol, n, nc):
    """Y = Ya + D(B-Ba) + E(B-Ba): the low-precision term's error relative
    to ||E (B-Ba)|| stays at the input precision (unit roundoff 2^-11 FP16,
    2^-8 BF16, ~2^-16 BF16x3, 2^-24 FP32, 2^-4 FP8 E4M3 on both operands,
    times a small accumulation factor).  A wrong FP8 encoding (another
    exponent bias, the FNUZ variant) is off by a factor of two and fails."""
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


def _scaled_residuals(H, X, lam):
    normH = np.abs(H).sum(axis=0).max()
    R = H @ X - X * lam
    return np.linalg.norm(R, axis=0) / (max(1.0, normH) * np.maximum(1.0, np.linalg.norm(X, axis=0)))


@pytest.fixture
def mixed_env(monkeypatch):
    monkeypatch.setenv("IPT_BLOCK_CLUSTER", "1")
    for v in ("IPT_DAVIDSON_ENRICH", "IPT_BLOCK_CLUSTER_MAXITER", "IPT_BLOCK_CLUSTER_TOL", "IPT_PRECISION",
              "IPT_MIXED_ANCHOR_EVERY", "IPT_BLOCK_LAGGED_STATUS"):
        monkeypatch.delenv(v, raising=False)
    return monkeypatch


@pytest.mark.parametrize("mode,lagged", [("fp16", "0"), ("bf16", "0"), ("bf16x3", "0"), ("fp32", "0"),
                                         ("fp16", "1")])
def test_mixed_block_solve_matches_fp64(mixed_env, mode, lagged):
    """C1 recipe (N=2048, k=16, d ~ U[0,10], V scale 0.05/sqrt(N)); the
    exact steps' status read at once or one step late."""
    n, k = 2048, 16
    H = orc.rand_dense(n, 0, 0.0, 10.0, 0.05)
    v64, l64, i64 = api.ipt_cuda_tol(H, k, 1e-12, 400)
    mixed_env.setenv("IPT_PRECISION", mode)
    mixed_env.setenv("IPT_MIXED_ANCHOR_EVERY", "10")
    mixed_env.setenv("IPT_BLOCK_LAGGED_STATUS", lagged)
    vm, lm, im = api.ipt_cuda_tol(H, k, 1e-12, 400)
    assert np.abs(lm - l64).max() <= 1e-10 * max(1.0, np.abs(l64).max())
    w = sl.eigh(H, eigvals_only=True, subset_by_index=[0, k - 1])
    assert np.abs(lm - w).max() <= 1e-10 * max(1.0, np.abs(w).max())
    res = orc.residuals(vm, lm, dense=H)
    assert res.max() <= 1e-10
    # anchored steps converge like the FP64 iteration (a broken product would not)
    assert im["iterations"] <= 2 * i64["iterations"] + 12


@pytest.mark.parametrize("mode", ["fp64", "fp16", "bf16x3"])
def test_headline_workload_lowest_pairs(mixed_env, mode):
    """The headline workload (N=16384, k=64, seed 0) with the benchmark's
    solver settings returns the 64 lowest eigenvalues: equal to LAPACK's
    (tests/golden/c2_lowest.json) to 1e-12, on the FP64 and the mixed
    paths."""
    import json
    import os
    from conftest import ROOT
    with open(os.path.join(ROOT, "tests", "golden", "c2_lowest.json")) as f:
        g = json.load(f)
    hip = api.hip()
    n, k = g["n"], 64
    dH = hip.malloc(n * n * 8)
    try:
        api.generate_dense(dH, n, g["seed"], g["lo"], g["hi"], g["vscale"])
        if mode != "fp64":
            mixed_env.setenv("IPT_PRECISION", mode)
        mixed_env.setenv("IPT_MIXED_ANCHOR_EVERY", "20")
        mixed_env.setenv("IPT_BLOCK_CLUSTER_TOL", "1e-8")
        mixed_env.setenv("IPT_DAVIDSON_ENRICH", "1")
        mixed_env.setenv("IPT_DAVIDSON_STEPS", "10")
        mixed_env.setenv("IPT_DAVIDSON_CONVERGED_TOL", "1e-10")
        print("headline workload, %s" % mode, flush=True)
        dv, dl, it, fp = api.ipt_cuda_device_tol(dH, n, k, 1e-9, 400)
        lam = hip.to_host(dl, (k,))
        X = hip.to_host(dv, (n, k))
        api.free_device_result(dv, dl)
        H = hip.to_host(dH, (n, n))
    finally:
        hip.free(dH)
    vals = np.sort(lam)
    w = np.asarray(g["values"][:k])
    assert np.abs(vals - w).max() <= 1e-12
    # every pair converged: the scaled residual of the reference's Davidson
    # (IPT/src/ipt_cuda.cu:2684-2697), computed independently on the host
    res = _scaled_residuals(H, X, lam)
    assert res.max() <= 1e-10, res.max()
    G = X.T @ X
    assert np.abs(G - np.eye(k)).max() <= 1e-10


@pytest.mark.parametrize("mode", ["fp16", "fp64"])
def test_class_launch_bitwise(mixed_env, mode):
    """Widths of one lane class updated in one launch give the iterates of
    one launch per width bit for bit (headline workload: widths 11, 12 and
    15 share a launch)."""
    hip = api.hip()
    n, k = 16384, 64
    dH = hip.malloc(n * n * 8)
    try:
        api.generate_dense(dH, n, 0, 0.0, 50.0, 0.1)
        if mode != "fp64":
            mixed_env.setenv("IPT_PRECISION", mode)
        mixed_env.setenv("IPT_BLOCK_CLUSTER_TOL", "1e-8")
        out = []
        for merge in ("0", "1"):
            print("class launch %s, %s" % (merge, mode), flush=True)
            mixed_env.setenv("IPT_BLOCK_CLASS_LAUNCH", merge)
            dv, dl, it, fp = api.ipt_cuda_device_tol(dH, n, k, 1e-9, 400)
            out.append((hip.to_host(dv, (n, k)), hip.to_host(dl, (k,)), api.last_solve_stats()))
            api.free_device_result(dv, dl)
    finally:
        hip.free(dH)
    (v0, l0, s0), (v1, l1, s1) = out
    assert s0 == s1
    assert np.array_equal(l0, l1) and np.array_equal(v0, v1)
    # the common result is a converged one, not merely the same one
    H = orc.rand_dense(n, 0, 0.0, 50.0, 0.1)
    assert _scaled_residuals(H, v1, l1).max() <= 1e-10


def test_fp8_block_solve_with_refinement(mixed_env):
    """FP8 E4M3 correction products (per-tile scaled E) anchored by FP64
    products, then the FP64 Rayleigh-Ritz + Davidson refinement: the same 16
    pairs as LAPACK to 1e-10 with residual <= 1e-10.  (Without refinement a
    slowly converging cluster stalls at the FP8 noise floor, about 1e-9.)"""
    n, k = 2048, 16
    H = orc.rand_dense(n, 0, 0.0, 10.0, 0.05)
    mixed_env.setenv("IPT_PRECISION", "fp8")
    mixed_env.setenv("IPT_MIXED_ANCHOR_EVERY", "5")
    mixed_env.setenv("IPT_DAVIDSON_ENRICH", "1")
    mixed_env.setenv("IPT_DAVIDSON_STEPS", "10")
    mixed_env.setenv("IPT_DAVIDSON_CONVERGED_TOL", "1e-11")
    vm, lm, im = api.ipt_cuda_tol(H, k, 1e-12, 400)
    w = sl.eigh(H, eigvals_only=True, subset_by_index=[0, k - 1])
    assert np.abs(lm - w).max() <= 1e-10 * max(1.0, np.abs(w).max())
    res = orc.residuals(vm, lm, dense=H)
    assert res.max() <= 1e-10
    assert api.last_solve_stats()["lowp_products"] > 0


# ------------------------------------------------------------------ two-phase entry points

def _plain_fp64(dH, n, k, tol, maxiter):
    hip = api.hip()
    dv, dl, it, fp = api.ipt_cuda_device_tol(dH, n, k, tol, maxiter)
    try:
        return hip.to_host(dv, (n, k)), hip.to_host(dl, (k,)), it
    finally:
        api.free_device_result(dv, dl)


@pytest.fixture
def plain_env(monkeypatch):
    for v in ("IPT_BLOCK_CLUSTER", "IPT_PRECISION"):
        monkeypatch.delenv(v, raising=False)
    return monkeypatch


@pytest.mark.parametrize("kind,lo_tol", [("fp16", 1e-3), ("fp32", 1e-6), ("tf32", 1e-6)])
@pytest.mark.parametrize("n,k", [(512, 512), (1000, 37)])
def test_two_phase_matches_plain_fp64(plain_env, kind, lo_tol, n, k):
    """ipt_mixed_precision{,_fp32,_tf32}_cuda_device_tol: low-precision IPT to
    lo_tol, FP64 IPT from there to 1e-12.  The result is the FP64 fixed point
    (same as the plain FP64 solve to 1e-10) and the FP64 phase needs fewer
    iterations than a cold FP64 start (the splitmix64 sweep matrix,
    eps 0.01, as in the reference's mixed benchmarks)."""
    hip = api.hip()
    H = orc.eps_dense(n, 0.01, 20260603 + n)
    dH = hip.to_device(np.asfortranarray(H))
    try:
        v64, l64, it64 = _plain_fp64(dH, n, k, 1e-12, 100)
        vm, lm, info = api.mixed_device_tol(kind, dH, n, k, lo_tol, 100, 1e-12, 100)
    finally:
        hip.free(dH)
    lo = {"fp16": "fp16", "fp32": "fp32", "tf32": "tf32"}[kind]
    assert info["n"] == n and info["k"] == k
    assert info["%s_fixed_point_residual" % lo] <= lo_tol
    assert info["fp64_fixed_point_residual"] <= 1e-12
    assert info["fp64_iterations"] < it64
    assert np.abs(lm - l64).max() <= 1e-10 * max(1.0, np.abs(l64).max())
    assert np.abs(vm - v64).max() <= 1e-9
    assert info["total_time_sec"] >= info["fp64_time_sec"] > 0.0


def test_rayleigh_mixed_tf32(plain_env):
    """ipt_rayleigh_mixed_tf32_cuda_device_tol: BF16x3 IPT -> QR -> FP64 IPT
    on Q^T H Q -> Q Z.  Ritz values of the low-precision subspace are
    accurate to the square of its error; vectors to the error itself."""
    hip = api.hip()
    n, k = 1024, 24
    H = orc.eps_dense(n, 0.01, 20260603 + n)
    dH = hip.to_device(np.asfortranarray(H))
    try:
        v64, l64, _ = _plain_fp64(dH, n, k, 1e-12, 100)
        vr, lr, info = api.mixed_device_tol("rayleigh_tf32", dH, n, k, 1e-6, 100, 1e-12, 100)
    finally:
        hip.free(dH)
    assert info["tf32_fixed_point_residual"] <= 1e-6 and info["fp64_fixed_point_residual"] <= 1e-12
    order = np.argsort(lr)
    assert np.abs(np.sort(lr) - np.sort(l64)).max() <= 1e-9 * max(1.0, np.abs(l64).max())
    # orthonormal Ritz vectors, residual at the subspace's accuracy
    G = vr.T @ vr
    assert np.abs(G - np.eye(k)).max() <= 1e-10
    R = H @ vr - vr * lr
    assert np.linalg.norm(R, axis=0).max() <= 1e-5
    assert order.size == k


def test_two_phase_invalid_arguments():
    with pytest.raises(api.IptError) as e:
        api.mixed_device_tol("fp16", 1, 4, 8, 1e-3, 10, 1e-12, 10)  # k > n
    assert e.value.status == 1


@pytest.mark.parametrize("spacing", ["5", "20"])
def test_mixed_strongly_coupled_no_rank_loss(mixed_env, spacing):
    """C2 recipe at N=4096, seed 4, cluster tolerance 1e-9: clusters whose
    block fixed point is not attracting drift when iterated past their
    convergence.  Low-precision steps stop a cluster as soon as it reaches
    the tolerance (an exact step then confirms it), and a basis that still
    comes out of rank below k is widened by unit vectors, so the solve
    returns; with the FP64 Davidson refinement its pairs are LAPACK's 64
    lowest."""
    n, k = 4096, 64
    H = orc.rand_dense(n, 4, 0.0, 50.0, 0.1)
    mixed_env.setenv("IPT_BLOCK_CLUSTER_TOL", "1e-9")
    v64, l64, i64 = api.ipt_cuda_tol(H, k, 1e-9, 400)
    mixed_env.setenv("IPT_PRECISION", "fp16")
    mixed_env.setenv("IPT_MIXED_ANCHOR_EVERY", spacing)
    print("N=4096 seed 4, spacing %s, no Davidson" % spacing, flush=True)
    vm, lm, im = api.ipt_cuda_tol(H, k, 1e-9, 400)
    assert np.all(np.isfinite(lm)) and np.all(np.isfinite(vm))
    print("values vs FP64 path: %.3e, iterations %d / %d" % (np.abs(np.sort(lm) - np.sort(l64)).max(),
                                                            im["iterations"], i64["iterations"]), flush=True)
    mixed_env.setenv("IPT_DAVIDSON_ENRICH", "1")
    mixed_env.setenv("IPT_DAVIDSON_STEPS", "60")
    mixed_env.setenv("IPT_DAVIDSON_CONVERGED_TOL", "1e-11")
    print("N=4096 seed 4, spacing %s, Davidson" % spacing, flush=True)
    vd, ld, idv = api.ipt_cuda_tol(H, k, 1e-9, 400)
    w = sl.eigh(H, eigvals_only=True, subset_by_index=[0, k - 1])
    assert np.abs(np.sort(ld) - w).max() <= 1e-10 * max(1.0, np.abs(w).max())
    assert _scaled_residuals(H, vd, ld).max() <= 1e-10


@pytest.mark.parametrize("packed", [False, True])
@pytest.mark.parametrize("n,nc", [(17, 3), (1000, 70), (4097, 129), (8192, 16)])
def test_dense_fp64_product(packed, n, nc):
    """The FP64 MFMA product of the block path, from column-major H and from
    the copy in MFMA operand order (k padded to 16, row blocks clamped),
    against numpy at ragged sizes, to a rounding bound of the summation
    (64 eps sqrt(n) |H||Z| elementwise)."""
    hip = api.hip()
    rng = np.random.default_rng(n + 7 * nc)
    H = np.asfortranarray(rng.standard_normal((n, n)))
    Z = np.asfortranarray(rng.standard_normal((n, nc)))
    dH = hip.to_device(H)
    dZ = hip.to_device(Z)
    dY = hip.malloc(n * nc * 8)
    try:
        api.dense_product(dH, n, dZ, nc, dY, packed=packed)
        Y = hip.to_host(dY, (n, nc))
    finally:
        for p in (dH, dZ, dY):
            hip.free(p)
    ref = H @ Z
    scale = np.abs(H) @ np.abs(Z)
    assert np.all(np.abs(Y - ref) <= 64 * np.finfo(float).eps * np.sqrt(n) * scale + 1e-300)
