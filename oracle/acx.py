"""CPU oracle of ACX-accelerated plain IPT: a numpy restatement of
IPT/src/ACX/acx_cuda.cu (FP64, the products by numpy's BLAS).

TEST INFRASTRUCTURE ONLY: imported by tests/ and by the CPU-baseline leg of
bench.py, always as the checker.  The product library never imports it.

Pinning (tests/test_acx_oracle.py): the published iteration and map-call
counts of IPT/results/ACX/acx_vs_original_ipt_summary.csv at N = 4096; one
extrapolation step lands exactly on the fixed point of an affine map with a
scalar contraction (a closed form that a wrong coefficient, sign or
difference order fails); the map equals one step of the C oracle's plain
IPT; converged pairs against LAPACK.
"""
import struct

import numpy as np

FIXED_POINT_RTOL = 1.0e-12      # IPT/src/ACX/acx_cuda.cu:17
ZERO_DENOMINATOR = 5            # IPT/src/ACX/acx_cuda.cu:16
SWEEP_SEED = 20260615           # IPT/test/ACX/benchmark_acx_cuda_sweep.cu:25

_M64 = 0xFFFFFFFFFFFFFFFF


class AcxError(RuntimeError):
    def __init__(self, status):
        super().__init__("acx oracle status %d" % status)
        self.status = status


def _splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & _M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _M64
    return x ^ (x >> 31)


def sweep_seed(n, epsilon, base=SWEEP_SEED):
    """Generator seed of one sweep case: base + N + splitmix64(bits of
    epsilon) mod 1000000007 (benchmark_acx_cuda_sweep.cu:273-282, 880-882).
    The matrix itself is oracle.eps_dense(n, epsilon, seed)."""
    bits = struct.unpack("<Q", struct.pack("<d", float(epsilon)))[0]
    return base + n + _splitmix64(bits) % 1000000007


def _denominators(d, k):
    den = d[None, :k] - d[:, None]
    off = np.ones(den.shape, dtype=bool)
    off[np.arange(k), np.arange(k)] = False
    return den, off & (np.abs(den) <= np.finfo(float).eps)


def ipt_map(H, d, X, den=None, residual=True):
    """One application of the plain IPT map, acx_cuda.cu:106-157, 298-358:
    Y = H X, mx_j = Y_jj, u_j = mx_j - d_j X_jj,
    F_rj = (Y_rj - d_r X_rj - X_rj u_j) / (d_j - d_r), F_jj = 1.
    Returns (F, per-column residual ||Y_j - X_j mx_j|| or None,
    zero_denominator).  `den` caches _denominators(d, k) across calls."""
    k = X.shape[1]
    j = np.arange(k)
    den, bad = den if den is not None else _denominators(d, k)
    Y = H @ X
    mx = Y[j, j].copy()
    upd = mx - d[:k] * X[j, j]
    res = np.sqrt(((Y - X * mx[None, :]) ** 2).sum(axis=0)) if residual else None
    Y -= d[:, None] * X
    Y -= X * upd[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        Y /= den
    Y[bad] = 0.0
    Y[j, j] = 1.0
    return Y, res, bool(bad.any())


def extrapolate(x, f1, f2, f3=None):
    """The ACX update, acx_cuda.cu:190-296: order 2 (f3 None)
    x + 2 s d1 + s^2 d2 with s = |<d2,d1>| / <d2,d2|; order 3
    x + 3 s d1 + 3 s^2 d2 + s^3 d3 with s = |<d3,d2>| / <d3,d3>, per column;
    s = 0 where the denominator is not positive."""
    d1 = f1 - x
    d2 = f2 - 2.0 * f1 + x
    if f3 is None:
        num = (d2 * d1).sum(axis=0)
        den = (d2 * d2).sum(axis=0)
    else:
        d3 = f3 - 3.0 * f2 + 3.0 * f1 - x
        num = (d3 * d2).sum(axis=0)
        den = (d3 * d3).sum(axis=0)
    with np.errstate(divide="ignore", invalid="ignore"):
        s = np.where(den > 0.0, np.abs(num / np.where(den > 0.0, den, 1.0)), 0.0)
    if f3 is None:
        return x + (2.0 * s * d1 + s * s * d2)
    return x + (3.0 * s * d1 + 3.0 * s * s * d2 + s * s * s * d3)


def acx_dense(H, k, tol, maxiter, fmap=None):
    """ACX loop of acx_cuda.cu:476-593 from X = I(:, 0:k).  Returns a dict
    with vectors (f1 of the last iteration), values diag(H f1), iterations,
    f_calls, converged, fixed_point_residual, max_residual.  `fmap`
    (optional) replaces ipt_map, for tests of the extrapolation itself."""
    H = np.asarray(H, dtype=np.float64)
    n = H.shape[0]
    if n <= 0 or k <= 0 or k > n or tol <= 0.0 or maxiter <= 0:
        raise AcxError(1)
    d = np.diag(H).copy()
    if fmap is None:
        den = _denominators(d, k)
        if den[1].any():
            raise AcxError(ZERO_DENOMINATOR)

        def fmap(X, residual=True):
            F, res, _ = ipt_map(H, d, X, den, residual)
            return F, res
    stop = tol if tol < FIXED_POINT_RTOL else FIXED_POINT_RTOL
    x = np.zeros((n, k))
    x[np.arange(k), np.arange(k)] = 1.0
    out = {"iterations": 0, "f_calls": 0, "converged": 0, "fixed_point_residual": np.nan,
           "max_residual": np.nan}
    f1 = None
    for it in range(1, maxiter + 1):
        order = 2 if it % 2 == 1 else 3
        out["iterations"] = it
        f1, res = fmap(x)
        out["f_calls"] += 1
        out["max_residual"] = float(res.max())
        xn = np.linalg.norm(x)
        if not (xn > 0.0) or not np.isfinite(xn):
            raise AcxError(1)
        fp = np.linalg.norm(f1 - x) / xn
        out["fixed_point_residual"] = fp
        if fp < stop:
            out["converged"] = 1
            break
        f2, _ = fmap(f1, residual=False)
        out["f_calls"] += 1
        if order == 2:
            x = extrapolate(x, f1, f2)
        else:
            f3, _ = fmap(f2, residual=False)
            out["f_calls"] += 1
            x = extrapolate(x, f1, f2, f3)
    out["vectors"] = f1
    out["values"] = np.einsum("ij,ji->i", H[:k, :], f1)
    return out


def ipt_plain_dense(H, k, tol, maxiter):
    """Plain IPT with the same map and the reference's stop test
    ||F(Z) - Z||_F / ||Z||_F <= tol (IPT/src/ipt_cuda.cu:1663-1688);
    numpy products, for sizes where the C oracle's loops are too slow.
    Returns (vectors, iterations, fixed_point_residual)."""
    H = np.asarray(H, dtype=np.float64)
    n = H.shape[0]
    d = np.diag(H).copy()
    z = np.zeros((n, k))
    z[np.arange(k), np.arange(k)] = 1.0
    den = _denominators(d, k)
    fp = np.nan
    it = 0
    for it in range(1, maxiter + 1):
        zn, _, _ = ipt_map(H, d, z, den, residual=False)
        fp = np.linalg.norm(zn - z) / np.linalg.norm(z)
        z = zn
        if fp <= tol:
            break
    return z, it, fp
