"""GPU parity of the plain IPT entry points against the CPU oracle and the
reference's golden data (KAT cases, splitmix64 sweep iteration counts)."""
import numpy as np
import pytest

from conftest import load_golden
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

from block_ipt_amd import api  # noqa: E402


def _kat_res(H, vec, val):
    R = H @ vec - vec * val
    den = np.maximum(1, np.abs(val)) * np.maximum(1, np.linalg.norm(vec, axis=0))
    return (np.linalg.norm(R, axis=0) / den).max()


@pytest.mark.parametrize("case", load_golden("dense_kat.json")["cases"], ids=lambda c: c["case"])
def test_dense_kat_parity(case):
    """IPT/test_temp/runtests.cu cases: vectors/values within 5e-9 of the
    oracle (we hold 1e-12), residual <= 5e-9."""
    n, k = case["n"], case["k"]
    H = orc.kat_dense(n, case["symmetric"])
    ov, ol, _, _ = orc.ipt_dense(H, k, iterations=20)
    gv, gl, info = api.ipt_cuda(H, k, 20)
    assert info["iterations"] == 20
    assert np.abs(gv - ov).max() <= 1e-12
    assert np.abs(gl - ol).max() <= 1e-12
    assert _kat_res(H, gv, gl) <= 5e-9


@pytest.mark.parametrize("n", [32, 64, 128, 256, 512, 1024, 2048])
def test_dense_eps_sweep_iterations(n):
    """Splitmix64 sweep (k = N, tol 1e-12): the published iteration counts,
    fixed-point residuals and Frobenius residuals of the reference run."""
    g = load_golden("dense_eps_sweep.json")
    row = [r for r in g["rows"] if r["n"] == n][0]
    H = orc.eps_dense(n, row["epsilon"], g["seed_base"] + n)
    gv, gl, info = api.ipt_cuda_tol(H, n, g["tol"], g["maxiter"])
    assert info["iterations"] == row["ipt_iterations"]
    assert 0.2 < info["fixed_point_residual"] / row["ipt_fixed_point_residual"] < 5.0
    # ||HV - V Lambda||_F evaluated on the host carries its own rounding
    # noise of order eps*||H||_F, which dominates at n = 2048
    fro = np.linalg.norm(H @ gv - gv * gl)
    noise = 2.0 * np.finfo(float).eps * np.linalg.norm(H)
    assert 0.2 * row["residual_ipt_frobenius"] < fro < max(5.0 * row["residual_ipt_frobenius"], noise)
    if n <= 512:
        ov, ol, oit, ofp = orc.ipt_dense(H, n, tol=g["tol"], maxiter=g["maxiter"])
        assert oit == info["iterations"]
        assert np.abs(gl - ol).max() <= 1e-11 * n
        assert np.abs(gv - ov).max() <= 1e-12


@pytest.mark.parametrize("n,k", [(100, 1), (100, 17), (333, 48), (600, 130), (1000, 200)])
def test_dense_odd_shapes(n, k):
    """Row/column/K tails of the MFMA GEMM (n not a multiple of 16/64, k
    spanning several 128-column passes)."""
    H = orc.eps_dense(n, 0.02, 7 + n)
    ov, ol, oit, _ = orc.ipt_dense(H, k, tol=1e-12, maxiter=50)
    gv, gl, info = api.ipt_cuda_tol(H, k, 1e-12, 50)
    assert info["iterations"] == oit
    assert np.abs(gv - ov).max() <= 1e-12
    assert np.abs(gl - ol).max() <= 1e-10


def test_dense_nonsymmetric_with_initial():
    n, k = 96, 5
    H = orc.kat_dense(n, 0)
    rng = np.random.default_rng(3)
    Z0 = np.eye(n)[:, :k] + 1e-3 * rng.standard_normal((n, k))
    ov, ol, oit, ofp = orc.ipt_dense(H, k, tol=1e-13, maxiter=60, Z0=Z0)
    gv, gl, info = api.ipt_cuda_with_initial_tol(H, Z0, k, 1e-13, 60)
    assert info["iterations"] == oit
    assert np.abs(gv - ov).max() <= 1e-12 and np.abs(gl - ol).max() <= 1e-12
    gv2, gl2, _ = api.ipt_cuda_with_initial(H, Z0, k, oit)
    assert np.abs(gv2 - gv).max() == 0.0


def test_dense_zero_initial_is_invalid():
    H = orc.kat_dense(32, 1)
    with pytest.raises(Exception) as e:
        api.ipt_cuda_with_initial_tol(H, np.zeros((32, 3)), 3, 1e-12, 10)
    assert e.value.status == 1


def test_dense_device_entry_matches_host():
    hip = api.hip()
    n, k = 512, 32
    H = orc.eps_dense(n, 0.03, 11)
    dH = hip.to_device(np.asfortranarray(H))
    try:
        dv, dl, it, fp = api.ipt_cuda_device_tol(dH, n, k, 1e-12, 100)
        vec = hip.to_host(dv, (n, k))
        val = hip.to_host(dl, (k,))
        api.free_device_result(dv, dl)
        hv, hl, info = api.ipt_cuda_tol(H, k, 1e-12, 100)
        assert it == info["iterations"] and fp == info["fixed_point_residual"]
        assert np.array_equal(vec, hv) and np.array_equal(val, hl)
        # explicit handle (own stream)
        h = api.create_handle()
        dv, dl = api.ipt_cuda_device(dH, n, k, it, handle=h)
        assert np.array_equal(hip.to_host(dv, (n, k)), hv)
        api.free_device_result(dv, dl)
        api.destroy_handle(h)
    finally:
        hip.free(dH)


def _csc_dense(cp, ri, v, n):
    D = np.zeros((n, n))
    for c in range(n):
        for p in range(cp[c], cp[c + 1]):
            D[ri[p], c] += v[p]
    return D


def test_sparse_kat_h2o():
    g = load_golden("h2o_w_csc.json")
    cp, ri, v = np.array(g["col_ptr"]), np.array(g["row_ind"]), np.array(g["values"])
    n = g["n"]
    gv, gl, info = api.ipt_cuda_sparse_csc(cp, ri, v, 1, g["iterations"])
    ov, ol, _, _ = orc.ipt_csc(cp, ri, v, 1, iterations=g["iterations"])
    assert np.abs(gv - ov).max() <= g["diff_tol"] and np.abs(gl - ol).max() <= g["diff_tol"]
    assert _kat_res(_csc_dense(cp, ri, v, n), gv, gl) <= g["residual_tol"]


def test_sparse_kat_banded():
    cp, ri, v = orc.banded_csc(64)
    gv, gl, info = api.ipt_cuda_sparse_csc(cp, ri, v, 5, 80)
    ov, ol, _, _ = orc.ipt_csc(cp, ri, v, 5, iterations=80)
    assert np.abs(gv - ov).max() <= 1e-12 and np.abs(gl - ol).max() <= 1e-12
    assert _kat_res(_csc_dense(cp, ri, v, 64), gv, gl) <= 1e-8


def test_sparse_unsorted_diagonal_is_permuted_back():
    """Host CSC path sorts the diagonal and returns vectors in the caller's
    order (IPT/src/ipt_preparation.cu:1375-1420, 1275-1291)."""
    n, k = 200, 6
    H = orc.eps_dense(n, 0.02, 5)
    rng = np.random.default_rng(0)
    p = rng.permutation(n)
    Hp = H[np.ix_(p, p)]
    Hp[np.abs(Hp) < 0.004] = 0.0  # sparsify, keep symmetric
    nz = Hp != 0
    cp = np.concatenate([[0], np.cumsum(nz.sum(0))]).astype(np.int32)
    ri = np.concatenate([np.nonzero(nz[:, c])[0] for c in range(n)]).astype(np.int32)
    v = np.concatenate([Hp[nz[:, c], c] for c in range(n)])
    gv, gl, info = api.ipt_cuda_sparse_csc_tol(cp, ri, v, k, 1e-13, 100)
    w, V = np.linalg.eigh(Hp)
    assert np.allclose(gl, w[:k], atol=1e-11)
    res = np.linalg.norm(Hp @ gv - gv * gl, axis=0)
    assert res.max() <= 1e-10


def test_sparse_device_entry():
    hip = api.hip()
    cp, ri, v = orc.banded_csc(300)
    n, k = 300, 12
    dcp, dri, dv_ = hip.to_device(cp), hip.to_device(ri), hip.to_device(v)
    try:
        dv, dl, it, fp = api.ipt_cuda_sparse_csc_device_tol(dcp, dri, dv_, n, k, int(cp[-1]), 1e-13, 200)
        vec, val = hip.to_host(dv, (n, k)), hip.to_host(dl, (k,))
        api.free_device_result(dv, dl)
    finally:
        for p in (dcp, dri, dv_):
            hip.free(p)
    ov, ol, oit, _ = orc.ipt_csc(cp, ri, v, k, tol=1e-13, maxiter=200)
    assert it == oit
    assert np.abs(vec - ov).max() <= 1e-12 and np.abs(val - ol).max() <= 1e-12


@pytest.mark.parametrize("n,k,seed", [(60, 12, 1), (120, 30, 2)])
def test_sparse_plain_lifts_degeneracies(monkeypatch, n, k, seed):
    """The plain CSC route with exactly repeated diagonal entries: the
    library's preparation lifts them (the unprepared map has zero gaps), the
    solve converges to the LAPACK pairs, and the vectors equal the oracle's
    plain IPT on the reference's own prepared matrix mapped back with the
    reference's routine (oracle/_ref), when that is built."""
    from test_lift import degenerate_csc

    for v in ("IPT_BLOCK_CLUSTER", "IPT_MAX_LIFT_ROUNDS", "IPT_MAX_LIFT_BLOCK_SIZE", "IPT_DEGENERACY_THRESHOLD"):
        monkeypatch.delenv(v, raising=False)
    csc, A = degenerate_csc(n, seed, shuffle=True)
    gv, gl, info = api.ipt_cuda_sparse_csc_tol(*csc, k, 1e-12, 200)
    assert info["fixed_point_residual"] <= 1e-12
    w = np.linalg.eigvalsh(A)
    assert np.abs(np.sort(gl) - w[:k]).max() <= 1e-11
    res = np.linalg.norm(A @ gv - gv * gl, axis=0) / np.linalg.norm(gv, axis=0)
    assert res.max() <= 1e-10
    ref = orc.ref_prepare_sparse_csc(*csc, k)
    if ref is not None:
        st, D, perm, rot = ref
        assert st == 0 and rot >= 1
        ov, ol, oit, ofp = orc.ipt_dense(D, k, tol=1e-12, maxiter=200)
        ov = orc.ref_prepared_vectors_to_original(*csc, k, ov)
        assert oit == info["iterations"]
        assert np.abs(gv - ov).max() <= 1e-10
        assert np.abs(gl - ol).max() <= 1e-11
