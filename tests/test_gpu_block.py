"""GPU parity of the block-cluster -> Rayleigh-Ritz (-> Davidson) path
against the CPU oracle and LAPACK."""
import numpy as np
import pytest
import scipy.linalg as sl

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

from block_ipt_amd import api  # noqa: E402
from block_ipt_amd.abi import IptError  # noqa: E402


@pytest.fixture
def block_env(monkeypatch):
    monkeypatch.setenv("IPT_BLOCK_CLUSTER", "1")
    for v in ("IPT_DAVIDSON_ENRICH", "IPT_DAVIDSON_STEPS", "IPT_BLOCK_CLUSTER_MAXITER", "IPT_BLOCK_CLUSTER_TOL",
              "IPT_COUPLING_GAP_RATIO_TOL", "IPT_MAX_LIFT_BLOCK_SIZE", "IPT_BLOCK_LAGGED_STATUS"):
        monkeypatch.delenv(v, raising=False)
    return monkeypatch


def _dense_to_csc(H):
    n = H.shape[0]
    nz = H != 0.0
    cp = np.concatenate([[0], np.cumsum(nz.sum(0))]).astype(np.int32)
    ri = np.concatenate([np.nonzero(nz[:, c])[0] for c in range(n)]).astype(np.int32)
    v = np.concatenate([H[nz[:, c], c] for c in range(n)])
    return cp, ri, v


def _check_eigs(H, vec, val, k, rel=1e-10, angle=1e-8, res_tol=1e-10):
    """Eigenvalues to `rel`, residuals to `res_tol`, and each Ritz vector
    within `angle` of the LAPACK eigenvector -- or, when its eigenvalue has
    a near neighbour, within the Davis-Kahan bound ||r|| / gap."""
    w, V = sl.eigh(H, subset_by_index=[0, k])
    assert np.all(np.abs(val - w[:k]) <= rel * np.maximum(1, np.abs(w[:k]))), np.abs(val - w[:k]).max()
    res = orc.residuals(vec, val, dense=H)
    assert res.max() <= res_tol
    rabs = np.linalg.norm(H @ vec - vec * val, axis=0)
    for j in range(k):
        gap = np.min(np.abs(np.delete(w, j) - w[j]))
        x = vec[:, j] / np.linalg.norm(vec[:, j])
        ang = np.linalg.norm(x - V[:, j] * (V[:, j] @ x))  # sin of the angle
        assert ang <= max(angle, 10 * rabs[j] / gap), (j, ang, gap)


@pytest.mark.parametrize("n,k,seed,hi,vs", [(512, 16, 0, 10.0, 0.05), (700, 24, 2, 20.0, 0.08),
                                            (2048, 16, 0, 10.0, 0.05)])
def test_dense_block_matches_oracle_and_lapack(block_env, n, k, seed, hi, vs):
    H = orc.rand_dense(n, seed, 0.0, hi, vs)
    gv, gl, info = api.ipt_cuda_tol(H, k, 1e-12, 400)
    _check_eigs(H, gv, gl, k)
    ov, ovec, oinfo = orc.block_pipeline(k, dense=H, maxiter=400, tol=1e-12)
    assert np.abs(gl - ov).max() <= 1e-10 * max(1.0, np.abs(ov).max())
    assert info["basis_cols"] == oinfo["basis_cols"]
    assert info["iterations"] == oinfo["max_iterations"]


def _far_rows_case(n=4000, k=16, seed=5, vs=0.055):
    """k targets packed in [0, 0.01] (one cluster), the other rows dense in
    [0.05, 0.4]: about 700 rows lie within 8 ||h - c0 I||_1 of the cluster's
    centre, more than one pass of the near-row blocks (512 rows)."""
    rng = np.random.default_rng(seed)
    d = np.concatenate([np.sort(rng.uniform(0.0, 0.01, k)), np.sort(rng.uniform(0.05, 0.4, n - k))])
    V = rng.standard_normal((n, n)) * vs / np.sqrt(n)
    H = np.triu(V, 1)
    H = H + H.T
    H[np.diag_indices(n)] = d
    return H


@pytest.mark.parametrize("lagged", ["0", "1"])
def test_near_and_far_rows_of_a_wide_cluster(block_env, lagged):
    # width 16: near rows by elimination over two passes, far rows by the
    # series; the same pairs as LAPACK and the oracle's iteration count,
    # with the status read at once or one step late
    block_env.setenv("IPT_BLOCK_LAGGED_STATUS", lagged)
    H = _far_rows_case()
    rg, _ = orc.target_clusters(16, dense=H)
    assert rg == [(0, 15)]
    gv, gl, info = api.ipt_cuda_tol(H, 16, 1e-12, 400)
    _check_eigs(H, gv, gl, 16)
    ov, ovec, oinfo = orc.block_pipeline(16, dense=H, maxiter=400, tol=1e-12)
    assert np.abs(gl - ov).max() <= 1e-10 * max(1.0, np.abs(ov).max())
    assert info["iterations"] == oinfo["max_iterations"]


def test_sparse_block_matches_oracle(block_env):
    H = orc.rand_dense(600, 9, 0.0, 10.0, 0.06)
    H[np.abs(H) < 2e-3] = 0.0
    cp, ri, v = _dense_to_csc(H)
    gv, gl, info = api.ipt_cuda_sparse_csc_tol(cp, ri, v, 20, 1e-12, 100)
    _check_eigs(H, gv, gl, 20)
    ov, ovec, oinfo = orc.block_pipeline(20, csc=(cp, ri, v), maxiter=100, tol=1e-12)
    assert np.abs(gl - ov).max() <= 1e-10 * max(1.0, np.abs(ov).max())
    assert info["iterations"] == oinfo["max_iterations"]


def test_sparse_block_unsorted_input(block_env):
    H = orc.rand_dense(400, 4, 0.0, 8.0, 0.05)
    p = np.random.default_rng(2).permutation(400)
    Hp = H[np.ix_(p, p)]
    gv, gl, info = api.ipt_cuda_sparse_csc_tol(*_dense_to_csc(Hp), 12, 1e-12, 100)
    _check_eigs(Hp, gv, gl, 12)


def test_dense_block_unsorted_input(block_env):
    H = orc.rand_dense(500, 6, 0.0, 10.0, 0.05)
    p = np.random.default_rng(5).permutation(500)
    Hp = np.asfortranarray(H[np.ix_(p, p)])
    gv, gl, info = api.ipt_cuda_tol(Hp, 10, 1e-12, 100)
    _check_eigs(Hp, gv, gl, 10)


def test_davidson_refinement(block_env):
    """Few block iterations, then Davidson to the converged tolerance."""
    block_env.setenv("IPT_BLOCK_CLUSTER_MAXITER", "2")
    block_env.setenv("IPT_DAVIDSON_ENRICH", "1")
    block_env.setenv("IPT_DAVIDSON_STEPS", "40")
    block_env.setenv("IPT_DAVIDSON_CONVERGED_TOL", "1e-12")
    n, k = 800, 16
    H = orc.rand_dense(n, 3, 0.0, 10.0, 0.05)
    gv, gl, info = api.ipt_cuda_tol(H, k, 1e-12, 100)
    assert info["davidson_attempted"] == 1 and info["accepted_steps"] >= 1
    assert info["davidson_residual_after"] <= 1e-12
    _check_eigs(H, gv, gl, k, res_tol=1e-12)
    ov, ovec, _ = orc.block_pipeline(k, dense=H, maxiter=2, tol=1e-12, davidson_steps=40, conv_tol=1e-12)
    assert np.abs(gl - ov).max() <= 1e-10 * max(1.0, np.abs(ov).max())


def test_device_entry_block_path(block_env):
    hip = api.hip()
    n, k = 600, 20
    H = orc.rand_dense(n, 8, 0.0, 12.0, 0.06)
    dH = hip.to_device(np.asfortranarray(H))
    try:
        dv, dl, it, fp = api.ipt_cuda_device_tol(dH, n, k, 1e-12, 100)
        vec, val = hip.to_host(dv, (n, k)), hip.to_host(dl, (k,))
        api.free_device_result(dv, dl)
    finally:
        hip.free(dH)
    _check_eigs(H, vec, val, k)
    assert fp <= 1e-12


@pytest.mark.parametrize("n,hi,k,seed", [(1024, 8.0, 48, 1), (1024, 6.0, 40, 4)])
def test_wide_cluster_class(block_env, n, hi, k, seed):
    """A 49- / 46-column cluster (the 33..64 lane class with its own launch
    and a separate finalize): same pairs as LAPACK and the oracle, same
    iteration count as the oracle."""
    H = orc.rand_dense(n, seed, 0.0, hi, 0.1)
    gv, gl, info = api.ipt_cuda_tol(H, k, 1e-12, 400)
    _check_eigs(H, gv, gl, k)
    ov, ovec, oinfo = orc.block_pipeline(k, dense=H, maxiter=400, tol=1e-12)
    assert np.abs(gl - ov).max() <= 1e-10 * max(1.0, np.abs(ov).max())
    assert info["iterations"] == oinfo["max_iterations"]


@pytest.mark.parametrize("n,m,k", [(1024, 80, 80), (1024, 100, 96), (2048, 130, 128)])
def test_clusters_wider_than_64(block_env, n, m, k):
    """IPT_MAX_LIFT_BLOCK_SIZE raised: an 80- / 100- / 130-column cluster runs
    through the LDS-resident update; same pairs as LAPACK and the oracle,
    same iteration count as the oracle."""
    from test_oracle import wide_cluster_matrix

    block_env.setenv("IPT_MAX_LIFT_BLOCK_SIZE", "142")
    H = wide_cluster_matrix(n, m, 7)
    gv, gl, info = api.ipt_cuda_tol(H, k, 1e-12, 100)
    _check_eigs(H, gv, gl, k)
    ov, ovec, oinfo = orc.block_pipeline(k, dense=H, max_block=142, maxiter=100, tol=1e-12)
    assert np.abs(gl - ov).max() <= 1e-10 * max(1.0, np.abs(ov).max())
    assert info["iterations"] == oinfo["max_iterations"]


def test_cluster_wider_than_supported_rejected(block_env):
    from test_oracle import wide_cluster_matrix

    block_env.setenv("IPT_MAX_LIFT_BLOCK_SIZE", "200")
    H = wide_cluster_matrix(1024, 150, 7)
    with pytest.raises(IptError) as e:
        api.ipt_cuda_tol(H, 140, 1e-12, 100)
    assert e.value.status == 1


def test_too_large_cluster_rejected(block_env):
    H = orc.rand_dense(2048, 1, 0.0, 6.25, 0.1)
    with pytest.raises(IptError) as e:
        api.ipt_cuda_tol(H, 64, 1e-12, 100)
    assert e.value.status == 1


def test_nonsymmetric_rejected(block_env):
    H = orc.kat_dense(64, 0)
    with pytest.raises(IptError) as e:
        api.ipt_cuda_tol(H, 4, 1e-12, 100)
    assert e.value.status == 1
    with pytest.raises(IptError) as e:
        api.ipt_cuda_sparse_csc_tol(*_dense_to_csc(H), 4, 1e-12, 100)
    assert e.value.status == 1


@pytest.mark.parametrize("n,i,j", [(300, 250, 10), (300, 70, 64), (1000, 999, 0)])
def test_dense_scan_finds_asymmetry_off_the_diagonal_tiles(block_env, n, i, j):
    """The dense scan compares every tile pair: one entry pair that differs
    (a ragged size, the pair in tiles away from the diagonal or across a
    tile edge) is reported; the symmetric matrix passes."""
    H = orc.rand_dense(n, 2, 0.0, 10.0, 0.05)
    api.ipt_cuda_tol(H, 4, 1e-12, 200)
    H[i, j] += 1e-6
    with pytest.raises(IptError) as e:
        api.ipt_cuda_tol(H, 4, 1e-12, 200)
    assert e.value.status == 1


def _chain_csc(n, strong):
    """Tridiagonal-plus-noise CSC with diagonal 0..n-1 and weak couplings,
    plus the given strong couplings {(i, j): value}."""
    rng = np.random.default_rng(4)
    A = np.diag(np.arange(n, dtype=float))
    for i in range(n - 1):
        A[i, i + 1] = A[i + 1, i] = 1e-3 * rng.uniform(0.5, 1.0)
    for (i, j), x in strong.items():
        A[i, j] = A[j, i] = x
    return _dense_to_csc(A), A


def test_oversample_targets(block_env):
    """IPT_BLOCK_CLUSTER_OVERSAMPLE=8: eight more target clusters in the
    basis, the same k lowest pairs (IPT/src/ipt_cuda.cu:4603-4613)."""
    block_env.setenv("IPT_BLOCK_CLUSTER_OVERSAMPLE", "8")
    H = orc.rand_dense(500, 2, 0.0, 10.0, 0.05)
    gv, gl, info = api.ipt_cuda_tol(H, 12, 1e-12, 100)
    _check_eigs(H, gv, gl, 12)
    assert info["oversample"] == 8 and info["basis_cols"] >= 20
    csc, A = _chain_csc(200, {})
    gv, gl, info = api.ipt_cuda_sparse_csc_tol(*csc, 10, 1e-12, 100)
    _check_eigs(A, gv, gl, 10)
    assert info["oversample"] == 8 and info["basis_cols"] >= 18


def test_adaptive_expansion(block_env):
    """IPT_BLOCK_CLUSTER_ADAPTIVE: the cluster of target k-1 = 9 takes in row
    10 (coupling/gap 0.15: above tau 0.1, below the cluster rule's 0.2), then
    row 11 of equal diagonal... here row 11 is 2 units away with ratio 0.05
    and stays out; the fields report it and the pairs are LAPACK's."""
    block_env.setenv("IPT_BLOCK_CLUSTER_ADAPTIVE", "1")
    block_env.setenv("IPT_BLOCK_CLUSTER_OVERSAMPLE", "5")  # ignored when adaptive
    csc, A = _chain_csc(300, {(9, 10): 0.15, (10, 12): 0.1})
    gv, gl, info = api.ipt_cuda_sparse_csc_tol(*csc, 10, 1e-12, 200)
    _check_eigs(A, gv, gl, 10)
    assert info["adaptive_block_enabled"] == 1 and info["oversample"] == 0
    assert abs(info["adaptive_coupling_tau"] - 0.1) < 1e-15
    assert (info["adaptive_target_block_start"], info["adaptive_target_block_end"]) == (9, 10)
    assert info["adaptive_added_indices"] == "10" and info["adaptive_limit_hit"] == 0
    # the same through the device CSC entry (host preparation)
    hip = api.hip()
    cp, ri, v = csc
    bufs = [hip.to_device(x) for x in (cp, ri, v)]
    try:
        dv, dl, it, fp = api.ipt_cuda_sparse_csc_device_tol(*bufs, 300, 10, len(v), 1e-12, 200)
        vec, val = hip.to_host(dv, (300, 10)), hip.to_host(dl, (10,))
        api.free_device_result(dv, dl)
    finally:
        for b in bufs:
            hip.free(b)
    _check_eigs(A, vec, val, 10)


def test_adaptive_limit(block_env):
    """A coupling above tau that would widen the target block past
    IPT_MAX_LIFT_BLOCK_SIZE sets adaptive_limit_hit and is left out."""
    block_env.setenv("IPT_BLOCK_CLUSTER_ADAPTIVE", "1")
    block_env.setenv("IPT_MAX_LIFT_BLOCK_SIZE", "4")
    csc, A = _chain_csc(100, {(9, 20): 1.5})
    gv, gl, info = api.ipt_cuda_sparse_csc_tol(*csc, 10, 1e-12, 200)
    _check_eigs(A, gv, gl, 10)
    assert info["adaptive_limit_hit"] == 1 and info["adaptive_added_indices"] == "none"


def test_davidson_returns_best_state(block_env):
    """Davidson with target tolerance 0 (corrections for every pair, also
    the converged ones, whose noise directions crowd the restarted search
    space): the largest residual does not fall monotonically, so among step
    limits 16..30 there are runs whose last accepted step raised it above
    an earlier one.  For the first such limit, the returned pairs are those
    of the best state seen (IPT/src/ipt_cuda.cu:485-701, 4941-4983): bit
    for bit the pairs of the same solve stopped at that step, with the
    recorded largest residual."""
    block_env.setenv("IPT_BLOCK_CLUSTER_MAXITER", "2")
    block_env.setenv("IPT_DAVIDSON_ENRICH", "1")
    block_env.setenv("IPT_DAVIDSON_CONVERGED_TOL", "0")
    n, k = 800, 16
    H = orc.rand_dense(n, 3, 0.0, 10.0, 0.05)
    for steps in range(30, 15, -1):
        block_env.setenv("IPT_DAVIDSON_STEPS", str(steps))
        gv, gl, info = api.ipt_cuda_tol(H, k, 1e-12, 100)
        if info["final_returned_from_best_so_far"] == 1:
            break
    else:
        pytest.fail("no step limit in 16..30 ended above the best state")
    assert info["best_so_far_enabled"] == 1 and info["best_so_far_update_count"] >= 2
    best = info["best_so_far_step"]
    assert 1 <= best < info["accepted_steps"]
    assert info["best_so_far_source"] == "davidson"
    assert info["davidson_residual_after"] == info["best_so_far_max_residual"]
    assert 0 <= info["best_so_far_max_residual_index"] < k
    res = orc.residuals(gv, gl, dense=H)
    assert abs(res.max() - info["best_so_far_max_residual"]) <= 1e-6 * res.max()
    # the same solve stopped at the best step returns the same pairs
    block_env.setenv("IPT_DAVIDSON_STEPS", str(best))
    sv, sl_, sinfo = api.ipt_cuda_tol(H, k, 1e-12, 100)
    assert sinfo["final_returned_from_best_so_far"] == 0 and sinfo["best_so_far_step"] == best
    assert np.array_equal(sl_, gl) and np.array_equal(sv, gv)
    assert sinfo["davidson_residual_after"] == info["best_so_far_max_residual"]


def test_rayleigh_ritz_qr_option_and_orthogonality_stats(block_env):
    """IPT_BLOCK_CLUSTER_QR (IPT/src/ipt_cuda.cu:3071, 3138-3221): the
    projection on an explicitly orthonormalised basis gives the same pairs
    and sets rayleigh_ritz_used_qr; IPT_RR_HOST_ORTHOGONALITY_STATS
    (:3072-3073, 2489-2524) reports ||B^T B - I|| of the basis used (large for
    the raw cluster bases, rounding-level for the orthonormalised one) and
    ||X^T X - I|| of the Ritz vectors."""
    H = orc.rand_dense(900, 5, 0.0, 10.0, 0.06)
    block_env.setenv("IPT_RR_HOST_ORTHOGONALITY_STATS", "1")
    gv, gl, info = api.ipt_cuda_tol(H, 20, 1e-12, 200)
    _check_eigs(H, gv, gl, 20)
    assert info["rayleigh_ritz_used_qr"] == 0
    assert info["basis_orthogonality_max_abs_error"] > 1e-6  # cluster bases are not orthonormal
    assert info["ritz_vectors_orthogonality_max_abs_error"] <= 1e-13
    assert info["basis_has_nan_or_inf"] == 0 and info["ritz_vectors_has_nan_or_inf"] == 0
    block_env.setenv("IPT_BLOCK_CLUSTER_QR", "1")
    qv, ql, qinfo = api.ipt_cuda_tol(H, 20, 1e-12, 200)
    _check_eigs(H, qv, ql, 20)
    assert qinfo["rayleigh_ritz_used_qr"] == 1
    assert qinfo["basis_orthogonality_max_abs_error"] <= 1e-13
    assert qinfo["ritz_vectors_orthogonality_max_abs_error"] <= 1e-13
    assert np.abs(ql - gl).max() <= 1e-11 * max(1.0, np.abs(gl).max())
    block_env.delenv("IPT_RR_HOST_ORTHOGONALITY_STATS")
    _, _, ninfo = api.ipt_cuda_tol(H, 20, 1e-12, 200)
    assert np.isnan(ninfo["basis_orthogonality_frobenius_error"])
