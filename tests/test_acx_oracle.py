"""The ACX oracle (oracle/acx.py) pinned: published iteration / map-call
counts of the reference's sweep, the closed form of one extrapolation step
on an affine map, the plain IPT step of the C oracle, and LAPACK."""
import json
import os

import numpy as np
import pytest

from conftest import ROOT
from oracle import acx
from oracle import oracle as orc


def _golden():
    with open(os.path.join(ROOT, "tests", "golden", "acx_sweep.json")) as f:
        return json.load(f)


def _affine(xstar, rho):
    """F(x) = x* + rho (x - x*) per column: a scalar contraction whose fixed
    point is x*."""
    def fmap(x, residual=True):
        return xstar + rho[None, :] * (x - xstar), np.zeros(x.shape[1])
    return fmap


@pytest.mark.parametrize("order", [2, 3])
def test_one_step_is_exact_on_a_scalar_contraction(order):
    """d1 = (rho-1) e, d2 = (rho-1)^2 e, d3 = (rho-1)^3 e with e = x - x*, so
    s = 1/|rho-1| and x + 2 s d1 + s^2 d2 = x + 3 s d1 + 3 s^2 d2 + s^3 d3
    = x*: one step is exact.  A wrong coefficient, sign or difference order
    leaves a multiple of e."""
    rng = np.random.default_rng(order)
    n, k = 50, 7
    xstar = rng.standard_normal((n, k))
    rho = rng.uniform(-0.9, 0.9, k)
    x = rng.standard_normal((n, k))
    f = _affine(xstar, rho)
    f1 = f(x)[0]
    f2 = f(f1)[0]
    f3 = f(f2)[0] if order == 3 else None
    got = acx.extrapolate(x, f1, f2, f3)
    assert np.abs(got - xstar).max() <= 1e-12 * np.abs(xstar).max()
    # not the plain iterate: the step is the extrapolation's own
    assert np.abs((f3 if order == 3 else f2) - xstar).max() > 1e-3


def test_loop_counts_on_an_affine_map():
    """Iteration 1 (order 2) lands on x*, iteration 2 sees F(x*) = x* and
    stops: 2 iterations, 3 map calls, the fixed point returned."""
    n, k = 40, 5
    rng = np.random.default_rng(5)
    xstar = rng.standard_normal((n, k))
    rho = rng.uniform(0.1, 0.8, k)
    out = acx.acx_dense(np.eye(n), k, 1e-12, 50, fmap=_affine(xstar, rho))
    assert (out["iterations"], out["f_calls"], out["converged"]) == (2, 3, 1)
    assert np.abs(out["vectors"] - xstar).max() <= 1e-12


def test_zero_sigma_when_steps_vanish():
    """A column already at its fixed point has d2 = 0: s = 0, x unchanged
    (acx_cuda.cu:231-235), no NaN."""
    x = np.ones((6, 2))
    got = acx.extrapolate(x, x.copy(), x.copy())
    assert np.array_equal(got, x)


@pytest.mark.parametrize("n,k", [(64, 64), (150, 20)])
def test_map_equals_plain_ipt_step(n, k):
    """The ACX map is the plain IPT map: one step from I and from another
    iterate equals one step of the C oracle's plain IPT."""
    H = orc.eps_dense(n, 0.05, 31 + n)
    d = np.diag(H).copy()
    F, _, bad = acx.ipt_map(H, d, np.eye(n)[:, :k])
    assert not bad
    ov, _, _, _ = orc.ipt_dense(H, k, iterations=1)
    assert np.abs(F - ov).max() <= 1e-14
    Z0 = np.asfortranarray(np.eye(n)[:, :k] + 0.01 * np.random.default_rng(n).standard_normal((n, k)))
    Z0[np.arange(k), np.arange(k)] = 1.0
    F2, _, _ = acx.ipt_map(H, d, Z0)
    ov2, _, _, _ = orc.ipt_dense(H, k, iterations=1, Z0=Z0)
    assert np.abs(F2 - ov2).max() <= 1e-13


def test_zero_denominator():
    H = orc.eps_dense(20, 0.05, 3)
    H[1, 1] = H[0, 0]
    with pytest.raises(acx.AcxError) as e:
        acx.acx_dense(H, 2, 1e-12, 10)
    assert e.value.status == 5


@pytest.mark.parametrize("n,k,eps", [(200, 200, 0.1), (300, 24, 0.12)])
def test_converged_pairs_match_lapack(n, k, eps):
    H = orc.eps_dense(n, eps, acx.sweep_seed(n, eps))
    out = acx.acx_dense(H, k, 1e-12, 1000)
    assert out["converged"] == 1
    V, lam = out["vectors"], out["values"]
    w = np.linalg.eigvalsh(H)
    assert np.abs(np.sort(lam) - w[:k]).max() <= 1e-11 * np.abs(w).max()
    res = np.linalg.norm(H @ V - V * lam[None, :], axis=0) / np.linalg.norm(V, axis=0)
    assert res.max() <= 1e-10


def test_sweep_seed():
    """Seeds of the sweep: base + N + splitmix64(bits(eps)) mod 1e9+7, with
    splitmix64 the C oracle's (itself checked against the reference's
    generator outputs)."""
    import struct
    for eps in (0.06, 0.08, 0.1, 0.12):
        bits = struct.unpack("<Q", struct.pack("<d", eps))[0]
        assert acx.sweep_seed(4096, eps) == 20260615 + 4096 + orc.splitmix64(bits) % 1000000007


def test_published_counts_n4096():
    """IPT/results/ACX/acx_vs_original_ipt_summary.csv, N = 4096, eps 0.06,
    k = N: ACX 5 iterations / 11 map calls, plain IPT 17 iterations; the
    fixed-point residuals agree to rounding."""
    g = _golden()
    row = [r for r in g["rows"] if r["n"] == 4096 and abs(r["epsilon"] - 0.06) < 1e-12][0]
    n = row["n"]
    H = orc.eps_dense(n, row["epsilon"], acx.sweep_seed(n, row["epsilon"], g["seed_base"]))
    out = acx.acx_dense(H, n, g["tol"], g["maxiter"])
    assert (out["iterations"], out["f_calls"], out["converged"]) == (row["acx_iterations"], row["acx_f_calls"], 1)
    assert 0.2 < out["fixed_point_residual"] / row["acx_fixed_point_residual"] < 5.0
    _, it, fp = acx.ipt_plain_dense(H, n, g["tol"], g["maxiter"])
    assert it == row["ipt_iterations"]
    assert 0.2 < fp / row["ipt_fixed_point_residual"] < 5.0
