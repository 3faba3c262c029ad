"""The CPU oracle checked against the reference's golden data, LAPACK and
closed forms (never against a re-typed copy of its own formula)."""
import numpy as np
import pytest
import scipy.linalg as sl

from conftest import load_golden
from oracle import oracle as orc


def test_splitmix64_known_answer():
    # SplitMix64 with state 0: first output 0xe220a8397b1dcdaf, second
    # 0x6e789e6aa1b965f4 (published reference sequence).
    assert orc.splitmix64(0) == 0xE220A8397B1DCDAF
    assert orc.splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4


def test_uniform01_range_and_moments():
    u = np.array([orc.uniform01(7, 100, i, j) for i in range(60) for j in range(60)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.02 and abs(u.var() - 1 / 12) < 0.01


def test_rand_dense_structure():
    n = 300
    H = orc.rand_dense(n, 5, 0.0, 50.0, 0.1)
    d = np.diag(H)
    assert np.all(np.diff(d) >= 0) and d.min() >= 0 and d.max() <= 50
    assert np.array_equal(H, H.T)
    V = H - np.diag(d)
    off = V[np.triu_indices(n, 1)] / (0.1 / np.sqrt(n))
    # standard normal moments within sampling error (44850 samples)
    assert abs(off.mean()) < 0.03 and abs(off.std() - 1.0) < 0.02
    assert abs(np.mean(off ** 4) - 3.0) < 0.15


def _kat_residual(H, vec, val):
    R = H @ vec - vec * val
    den = np.maximum(1, np.abs(val)) * np.maximum(1, np.linalg.norm(vec, axis=0))
    return (np.linalg.norm(R, axis=0) / den).max()


@pytest.mark.parametrize("case", load_golden("dense_kat.json")["cases"], ids=lambda c: c["case"])
def test_dense_kat_cases(case):
    """IPT/test_temp/runtests.cu cases: 20 iterations; eigenpairs must equal
    LAPACK's and the residual must meet the reference's 5e-9 bar."""
    g = load_golden("dense_kat.json")
    n, k, sym = case["n"], case["k"], case["symmetric"]
    H = orc.kat_dense(n, sym)
    vec, val, it, _ = orc.ipt_dense(H, k, iterations=g["iterations"])
    assert it == 20
    res = _kat_residual(H, vec, val)
    assert res <= g["residual_tol"]
    # same order as the reference's recorded residual (both at rounding level)
    assert res <= max(100 * case["max_relative_residual"], 1e-14)
    w, V = np.linalg.eig(H)
    order = np.argsort(w.real)
    w = w.real[order]
    assert np.allclose(val, w[:k], rtol=0, atol=1e-12)
    for j in range(k):
        v = V[:, order[j]].real
        v = v / v[j]          # IPT normalisation z_jj = 1
        assert np.allclose(vec[:, j], v, atol=1e-11)


def test_kat_detects_planted_mistake():
    """A sign error in G (G_rj = 1/(d_r - d_j)) or a dropped Z*diag(DZ) term
    must fail the LAPACK comparison used above."""
    n, k = 32, 5
    H = orc.kat_dense(n, 1)
    d = np.diag(H)
    w, V = np.linalg.eigh(H)
    for planted in ("sign", "drop"):
        Z = np.eye(n)[:, :k]
        for _ in range(20):
            Y = H @ Z
            c = Y[np.arange(k), np.arange(k)] - d[:k]
            with np.errstate(divide="ignore"):
                G = 1.0 / (d[None, :k] - d[:, None])
            if planted == "sign":
                G = -G
            G[np.arange(k), np.arange(k)] = 0
            cz = 0.0 if planted == "drop" else Z * c[None, :]
            Z = (Y - d[:, None] * Z - cz) * G
            Z[np.arange(k), np.arange(k)] = 1
        v = V[:, :k] / np.diag(V[:k, :k])
        assert not np.allclose(Z, v, atol=1e-11)


@pytest.mark.parametrize("row", [r for r in load_golden("dense_eps_sweep.json")["rows"] if r["n"] <= 512],
                         ids=lambda r: "n%d" % r["n"])
def test_dense_eps_sweep_iterations(row):
    """Published iteration counts of the splitmix64 sweep (k = N, tol 1e-12)
    are reproduced exactly; residuals agree to rounding."""
    g = load_golden("dense_eps_sweep.json")
    n = row["n"]
    H = orc.eps_dense(n, row["epsilon"], g["seed_base"] + n)
    vec, val, it, fp = orc.ipt_dense(H, n, tol=g["tol"], maxiter=g["maxiter"])
    assert it == row["ipt_iterations"]
    assert fp <= g["tol"]
    assert 0.2 < fp / row["ipt_fixed_point_residual"] < 5.0
    fro = np.linalg.norm(H @ vec - vec * val)
    assert 0.2 < fro / row["residual_ipt_frobenius"] < 5.0


def test_sparse_kat_h2o():
    g = load_golden("h2o_w_csc.json")
    cp, ri, v = (np.array(g["col_ptr"]), np.array(g["row_ind"]), np.array(g["values"]))
    n = g["n"]
    vec, val, it, _ = orc.ipt_csc(cp, ri, v, g["k"], iterations=g["iterations"])
    dense = np.zeros((n, n))
    for c in range(n):
        for p in range(cp[c], cp[c + 1]):
            dense[ri[p], c] += v[p]
    assert _kat_residual(dense, vec, val) <= g["residual_tol"]
    w = np.linalg.eigvals(dense).real
    assert abs(val[0] - np.sort(w)[0]) <= 1e-10


def test_sparse_kat_banded():
    cp, ri, v = orc.banded_csc(64)
    vec, val, it, _ = orc.ipt_csc(cp, ri, v, 5, iterations=80)
    dense = np.zeros((64, 64))
    for c in range(64):
        dense[ri[cp[c]:cp[c + 1]], c] = v[cp[c]:cp[c + 1]]
    assert _kat_residual(dense, vec, val) <= 1e-8
    assert np.allclose(val, np.linalg.eigvalsh(dense)[:5], atol=1e-12)


def _dense_to_csc(H):
    n = H.shape[0]
    nz = H != 0.0
    cp = np.zeros(n + 1, dtype=np.int32)
    cp[1:] = np.cumsum(nz.sum(axis=0))
    ri = np.concatenate([np.nonzero(nz[:, c])[0] for c in range(n)]).astype(np.int32)
    v = np.concatenate([H[nz[:, c], c] for c in range(n)])
    return cp, ri, v


@pytest.mark.parametrize("case", load_golden("cluster_ranges.json")["cases"], ids=lambda c: c["name"])
def test_cluster_ranges_match_reference_code(case):
    """Ranges equal those of the reference's own host code (built from source
    in oracle/_ref; recorded in tests/golden/cluster_ranges.json)."""
    n, k = case["n"], case["k"]
    if case["kind"] == "eps":
        H = orc.eps_dense(n, case["vscale"], case["seed"])
    else:
        H = orc.rand_dense(n, case["seed"], case["lo"], case["hi"], case["vscale"])
    ranges, too_large = orc.target_clusters(k, dense=H)
    multi = [r for r in ranges if r[1] > r[0]]
    assert [tuple(r) for r in case["ranges"]] == multi
    assert too_large == case["too_large"]
    # singleton completion covers every target exactly once
    cov = np.zeros(n, int)
    for a, b in ranges:
        cov[a:b + 1] += 1
    assert np.all(cov[:k] == 1)
    # CSC input gives the same answer as dense input
    r2, t2 = orc.target_clusters(k, csc=_dense_to_csc(H))
    assert r2 == ranges and t2 == too_large


def test_cluster_ranges_live_reference():
    """When oracle/_ref is built (reference checkout present), compare on
    fresh inputs too."""
    H = orc.rand_dense(700, 11, 0.0, 20.0, 0.08)
    got = orc.ref_target_clusters(*_dense_to_csc(H), 24)
    if got is None:
        pytest.skip("oracle/_ref not built (no reference checkout)")
    ranges, tl = orc.target_clusters(24, dense=H)
    assert [r for r in ranges if r[1] > r[0]] == got[0] and tl == got[1]


def test_sym_eig_matches_lapack():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((40, 40))
    A = A + A.T
    w, V = orc.sym_eig(A)
    wl = np.linalg.eigvalsh(A)
    assert np.allclose(w, wl, atol=1e-12)
    assert np.allclose(V.T @ V, np.eye(40), atol=1e-12)
    assert np.allclose(A @ V, V * w, atol=1e-11)


def test_block_cluster_solve_invariant_subspace():
    """A converged cluster basis spans the LAPACK invariant subspace of the
    cluster's eigenvalues."""
    H = orc.rand_dense(400, 2, 0.0, 10.0, 0.05)
    ranges, _ = orc.target_clusters(12, dense=H)
    a, b = max(ranges, key=lambda r: r[1] - r[0])
    m = b - a + 1
    assert m >= 2
    B, it, fp = orc.block_cluster_solve(a, b, 200, 1e-13, dense=H)
    assert fp <= 1e-13
    w, V = np.linalg.eigh(H)
    Q, _ = np.linalg.qr(B)
    # eigenvectors with largest weight in span(B) are those of the cluster
    ov = np.linalg.norm(Q.T @ V, axis=0)
    idx = np.argsort(-ov)[:m]
    ang = sl.subspace_angles(Q, V[:, idx]).max()
    assert ang < 1e-9


def wide_cluster_matrix(n, m, seed):
    """m nearly degenerate states (spacing 1e-3, couplings 2e-2) below the
    rest (U[5,10], couplings 1e-3): one cluster of width m, weakly coupled
    to the others."""
    rng = np.random.default_rng(seed)
    d = np.concatenate([1.0 + 1e-3 * np.arange(m), np.sort(rng.uniform(5.0, 10.0, n - m))])
    V = 1e-3 * rng.standard_normal((n, n))
    V[:m, :m] = 2e-2 * rng.standard_normal((m, m))
    V = np.triu(V, 1)
    return np.asfortranarray(np.diag(d) + V + V.T)


def test_wide_cluster_invariant_subspace():
    """A 100-wide cluster (beyond the 64 default, IPT_MAX_LIFT_BLOCK_SIZE
    raised): the converged basis spans the LAPACK invariant subspace of the
    100 lowest eigenvalues."""
    H = wide_cluster_matrix(600, 100, 7)
    ranges, tl = orc.target_clusters(96, dense=H, max_block=142)
    assert not tl and max(b - a + 1 for a, b in ranges) == 100
    B, it, fp = orc.block_cluster_solve(0, 99, 100, 1e-13, dense=H)
    assert fp <= 1e-13
    w, V = np.linalg.eigh(H)
    Q, _ = np.linalg.qr(B)
    assert sl.subspace_angles(Q, V[:, :100]).max() < 1e-9


def test_width_one_cluster_equals_plain_ipt_fixed_point():
    H = orc.kat_dense(48, 1)
    B, it, fp = orc.block_cluster_solve(3, 3, 200, 1e-14, dense=H)
    vec, val, _, _ = orc.ipt_dense(H, 4, tol=1e-14, maxiter=200)
    assert np.allclose(B[:, 0], vec[:, 3], atol=1e-12)


def test_rayleigh_ritz_matches_lapack_projection():
    H = orc.rand_dense(300, 4, 0.0, 10.0, 0.1)
    rng = np.random.default_rng(1)
    B = np.eye(300)[:, :30] + 0.05 * rng.standard_normal((300, 30))
    val, vec = orc.rayleigh_ritz(B, 10, dense=H)
    Q, _ = np.linalg.qr(B)
    wl = np.linalg.eigvalsh(Q.T @ H @ Q)[:10]
    assert np.allclose(val, wl, atol=1e-12)
    assert np.allclose(np.linalg.norm(vec, axis=0), 1.0)
    # Galerkin condition: residuals orthogonal to the subspace
    R = H @ vec - vec * val
    assert np.abs(Q.T @ R).max() < 1e-11


def test_block_pipeline_c1_small():
    """Clusters -> block IPT -> RR on a C1-recipe matrix (reduced to N=512 to
    keep the CPU suite short): every pair at residual <= 1e-10 and equal to
    LAPACK eigenpairs (eigenvalues 1e-10 relative, angle <= 1e-8)."""
    n, k = 512, 16
    H = orc.rand_dense(n, 0, 0.0, 10.0, 0.05)
    val, vec, info = orc.block_pipeline(k, dense=H, maxiter=100, tol=1e-12)
    res = orc.residuals(vec, val, dense=H)
    assert res.max() <= 1e-10
    w, V = sl.eigh(H, subset_by_index=[0, k - 1])
    assert np.all(np.abs(val - w) <= 1e-10 * np.maximum(1, np.abs(w)))
    assert sl.subspace_angles(vec, V).max() <= 1e-8


def test_davidson_refines_unconverged_pairs():
    """Davidson from a truncated block-IPT basis reaches the converged
    tolerance and LAPACK eigenpairs."""
    n, k = 400, 12
    H = orc.rand_dense(n, 3, 0.0, 10.0, 0.05)
    val0, vec0, _ = orc.block_pipeline(k, dense=H, maxiter=2, tol=1e-12)
    res0 = orc.residuals(vec0, val0, dense=H)
    assert res0.max() > 1e-8
    val, vec, info = orc.block_pipeline(k, dense=H, maxiter=2, tol=1e-12,
                                        davidson_steps=30, conv_tol=1e-12)
    res = orc.residuals(vec, val, dense=H)
    assert res.max() <= 1e-12
    w = sl.eigh(H, eigvals_only=True, subset_by_index=[0, k - 1])
    assert np.allclose(val, w, rtol=1e-10, atol=0)


def test_too_large_cluster_is_an_error():
    H = orc.rand_dense(2048, 1, 0.0, 6.25, 0.1)
    with pytest.raises(orc.OracleError):
        orc.block_pipeline(64, dense=H)
