"""ACX-accelerated plain IPT on the GPU (acx_cuda_device_tol): the
reference's published iteration / map-call counts, parity with the ACX
oracle, the maxiter and zero-denominator paths."""
import json
import os

import numpy as np
import pytest

from conftest import ROOT
from oracle import acx
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

from block_ipt_amd import api  # noqa: E402


def _golden():
    with open(os.path.join(ROOT, "tests", "golden", "acx_sweep.json")) as f:
        return json.load(f)


def _run(H, k, tol, maxiter):
    hip = api.hip()
    n = H.shape[0]
    dH = hip.to_device(np.asfortranarray(H))
    try:
        return api.acx_cuda_device_tol(dH, n, k, tol, maxiter)
    finally:
        hip.free(dH)


def _plain_iterations(H, tol, maxiter):
    hip = api.hip()
    n = H.shape[0]
    dH = hip.to_device(np.asfortranarray(H))
    try:
        dv, dl, it, fp = api.ipt_cuda_device_tol(dH, n, n, tol, maxiter)
        api.free_device_result(dv, dl)
    finally:
        hip.free(dH)
    return it, fp


_ROWS = [r for r in _golden()["rows"] if r["n"] <= 8192 or abs(r["epsilon"] - 0.12) < 1e-12]


@pytest.mark.parametrize("row", _ROWS, ids=lambda r: "n%d-eps%g" % (r["n"], r["epsilon"]))
def test_published_counts(row):
    """IPT/results/ACX/acx_vs_original_ipt_summary.csv (k = N, tol 1e-12):
    ACX iterations and map calls, and the plain IPT iterations of the same
    matrix, equal the published ones."""
    g = _golden()
    n, eps = row["n"], row["epsilon"]
    print("acx published n=%d eps=%g" % (n, eps), flush=True)
    H = orc.eps_dense(n, eps, acx.sweep_seed(n, eps, g["seed_base"]))
    vec, val, info = _run(H, n, g["tol"], g["maxiter"])
    assert (info["iterations"], info["f_calls"], info["converged"]) == (row["acx_iterations"], row["acx_f_calls"], 1)
    assert info["fixed_point_residual"] < g["tol"]
    assert 0.1 < info["fixed_point_residual"] / row["acx_fixed_point_residual"] < 10.0
    assert np.all(np.isfinite(vec)) and np.all(np.isfinite(val))
    # the values are the diagonal entries of H f1 for f1 with unit diagonal
    assert np.array_equal(np.diag(vec), np.ones(n))
    del vec
    it, fp = _plain_iterations(H, g["tol"], g["maxiter"])
    assert it == row["ipt_iterations"]
    assert 0.1 < fp / row["ipt_fixed_point_residual"] < 10.0


@pytest.mark.parametrize("n,k,eps", [(256, 256, 0.1), (700, 64, 0.12), (1000, 1000, 0.06), (333, 130, 0.08)])
def test_matches_oracle(n, k, eps):
    """Same iteration and map-call counts as the oracle, vectors and values
    to rounding (FP64 products in another summation order)."""
    H = orc.eps_dense(n, eps, acx.sweep_seed(n, eps))
    vec, val, info = _run(H, k, 1e-12, 1000)
    o = acx.acx_dense(H, k, 1e-12, 1000)
    assert (info["iterations"], info["f_calls"], info["converged"]) == (o["iterations"], o["f_calls"], o["converged"])
    assert np.abs(vec - o["vectors"]).max() <= 1e-11
    assert np.abs(val - o["values"]).max() <= 1e-12 * np.abs(o["values"]).max()
    assert abs(info["max_residual"] - o["max_residual"]) <= 1e-9 * max(1.0, o["max_residual"])
    w = np.linalg.eigvalsh(H)
    assert np.abs(np.sort(val) - w[:k]).max() <= 1e-11 * np.abs(w).max()


def test_maxiter_returns_unconverged():
    """maxiter reached: success, converged = 0, the last f1 returned
    (acx_cuda.cu:476-593); 2 iterations = 2 + 3 map calls."""
    n = 300
    H = orc.eps_dense(n, 0.1, 17)
    vec, val, info = _run(H, n, 1e-12, 2)
    assert (info["iterations"], info["f_calls"], info["converged"]) == (2, 5, 0)
    o = acx.acx_dense(H, n, 1e-12, 2)
    assert np.abs(vec - o["vectors"]).max() <= 1e-12
    assert np.abs(val - o["values"]).max() <= 1e-12 * np.abs(o["values"]).max()


def test_zero_denominator_status():
    H = orc.eps_dense(64, 0.05, 3)
    H[1, 1] = H[0, 0]
    with pytest.raises(api.IptError) as e:
        _run(H, 2, 1e-12, 10)
    assert e.value.status == 5


def test_repeatable():
    """Fixed reduction order: two runs agree bitwise."""
    n = 512
    H = orc.eps_dense(n, 0.1, 99)
    a = _run(H, 100, 1e-12, 1000)
    b = _run(H, 100, 1e-12, 1000)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert a[2]["f_calls"] == b[2]["f_calls"]


def test_device_eps_generator_matches_oracle():
    """The device sweep-matrix generator equals the oracle's bit for bit."""
    hip = api.hip()
    n, eps = 700, 0.12
    seed = acx.sweep_seed(n, eps)
    dH = hip.malloc(n * n * 8)
    try:
        api.generate_eps(dH, n, eps, seed)
        G = hip.to_host(dH, (n, n))
    finally:
        hip.free(dH)
    assert np.array_equal(G, orc.eps_dense(n, eps, seed))
