"""Sparse block path with the matrix resident on the device
(ipt_cuda_sparse_csc_device_tol, IPT_BLOCK_CLUSTER=1) on the orbital-sum
model, the standalone CSR product, and the host-preparation route for
input whose diagonal is not sorted."""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

from block_ipt_amd import api  # noqa: E402


@pytest.fixture
def block_env(monkeypatch):
    monkeypatch.setenv("IPT_BLOCK_CLUSTER", "1")
    for v in ("IPT_DAVIDSON_ENRICH", "IPT_DAVIDSON_STEPS", "IPT_BLOCK_CLUSTER_MAXITER", "IPT_BLOCK_CLUSTER_TOL",
              "IPT_COUPLING_GAP_RATIO_TOL", "IPT_MAX_LIFT_BLOCK_SIZE", "IPT_PRECISION", "IPT_DAVIDSON_PRECOND",
              "IPT_CSR_TILES", "IPT_CSR_TILES_WIDTH"):
        monkeypatch.delenv(v, raising=False)
    return monkeypatch


class _Dev:
    def __init__(self):
        self.hip = api.hip()
        self.ptrs = []

    def put(self, a):
        p = self.hip.to_device(np.ascontiguousarray(a))
        self.ptrs.append(p)
        return p

    def close(self):
        for p in self.ptrs:
            self.hip.free(p)


@pytest.fixture
def dev():
    d = _Dev()
    yield d
    d.close()


@pytest.mark.parametrize("k", [1, 7, 32, 33, 64, 100])
def test_spmm_matches_scipy(dev, k):
    n, m = 5000, 30
    rp, ci, v = orc.orbital_csr(n, m, 3)
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    X = np.random.default_rng(k).standard_normal((n, k))
    drp, dci, dv, dX = dev.put(rp), dev.put(ci), dev.put(v), dev.put(X)
    dY = dev.hip.malloc(n * k * 8)
    dev.ptrs.append(dY)
    api.csr_spmm(drp, dci, dv, n, dX, k, dY, k, k)
    Y = dev.hip.to_host(dY, (n, k), order="C")
    ref = A @ X
    assert np.abs(Y - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 8])
def test_spmm_narrow_long_rows(dev, k):
    """Widths up to 8 (the one-lane-per-pair kernel): rows of about 600
    entries (several 256-entry rounds), strided X and Y rows."""
    n, m = 3000, 300
    rp, ci, v = orc.orbital_csr(n, m, 11)
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    ldx, ldy = k + 3, k + 1
    X = np.random.default_rng(k).standard_normal((n, ldx))
    drp, dci, dv, dX = dev.put(rp), dev.put(ci), dev.put(v), dev.put(X)
    dY = dev.hip.malloc(n * ldy * 8)
    dev.ptrs.append(dY)
    api.csr_spmm(drp, dci, dv, n, dX, ldx, dY, ldy, k)
    Y = dev.hip.to_host(dY, (n, ldy), order="C")[:, :k]
    ref = A @ X[:, :k]
    assert np.abs(Y - ref).max() <= 1e-13 * np.abs(ref).max()


_ORBITAL = {}


def _orbital_cached(n, m, seed):
    key = (n, m, seed)
    if key not in _ORBITAL:
        _ORBITAL[key] = orc.orbital_csr(n, m, seed)
    return _ORBITAL[key]


@pytest.mark.parametrize("slices", [0, 1, 2, 8, 16])
@pytest.mark.parametrize("n,k", [(300000, 1), (150000, 4), (20001, 3), (40000, 9), (150000, 20)])
def test_spmm_column_slices(dev, n, k, slices):
    """Column-sliced passes of the CSR product (rows of the orbital model,
    ascending columns): every pass count, including the automatic one (X
    of 2.4, 4.8, 0.5, 2.9 and 24 MB: 4, 4, 1, 1 and 16 passes), against
    scipy; a
    slice bound off by one would drop or double the entries at a slice
    edge."""
    rp, ci, v = _orbital_cached(n, 24, 7)
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    X = np.random.default_rng(n + k).standard_normal((n, k))
    drp, dci, dv, dX = dev.put(rp), dev.put(ci), dev.put(v), dev.put(X)
    dY = dev.hip.malloc(n * k * 8)
    dev.ptrs.append(dY)
    api.csr_spmm_sliced(drp, dci, dv, n, dX, k, dY, k, k, slices)
    Y = dev.hip.to_host(dY, (n, k), order="C")
    ref = A @ X
    assert np.abs(Y - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("k", [12, 27])
def test_spmm_padded_row_stride(dev, k):
    """X at the row stride the block path uses (12 -> 16, 27 -> 32 doubles,
    ipt_hip_csr_row_stride) with the automatic pass count (X of 14.4 and
    25.9 MB: 1 and 16 passes): the padding columns never enter the sums
    (they hold NaN here), against scipy."""
    n = 150000
    ld = api.lib().ipt_hip_csr_row_stride(k, 8)
    assert ld == {12: 16, 27: 32}[k]
    rp, ci, v = _orbital_cached(n, 24, 7)
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    X = np.full((n, ld), np.nan)
    X[:, :k] = np.random.default_rng(k).standard_normal((n, k))
    drp, dci, dv, dX = dev.put(rp), dev.put(ci), dev.put(v), dev.put(X)
    dY = dev.hip.malloc(n * k * 8)
    dev.ptrs.append(dY)
    api.csr_spmm_sliced(drp, dci, dv, n, dX, ld, dY, k, k, 0)
    Y = dev.hip.to_host(dY, (n, k), order="C")
    ref = A @ X[:, :k]
    assert np.isfinite(Y).all()
    assert np.abs(Y - ref).max() <= 1e-13 * np.abs(ref).max()


def test_spmm_sliced_rejects_bad_counts(dev):
    n = 100
    rp, ci, v = orc.orbital_csr(n, 4, 1)
    drp, dci, dv = dev.put(rp), dev.put(ci), dev.put(v)
    dX = dev.hip.malloc(n * 8)
    dev.ptrs.append(dX)
    for bad in (3, 32, -1):
        with pytest.raises(api.IptError):
            api.csr_spmm_sliced(drp, dci, dv, n, dX, 1, dX, 1, 1, bad)


@pytest.mark.parametrize("n,k", [(300000, 1), (150000, 2), (20001, 3), (4095, 1), (8193, 2)])
def test_spmm_tiled(dev, n, k):
    """The LDS-tiled product (4096-row groups, 8192-column tiles): rows and
    columns past a group or tile edge, one partial group, X with ld = k
    (column c at stride k), against scipy.  The LDS sums take the products
    in the order the atomic adds land, so the bar is rounding, not bits."""
    rp, ci, v = _orbital_cached(n, 24, 7)
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    X = np.random.default_rng(n + k).standard_normal((n, k))
    drp, dci, dv, dX = dev.put(rp), dev.put(ci), dev.put(v), dev.put(X)
    dY = dev.hip.malloc(n * k * 8)
    dev.ptrs.append(dY)
    api.csr_spmm_tiled(drp, dci, dv, n, dX, k, dY, k, k)
    Y = dev.hip.to_host(dY, (n, k), order="C")
    ref = A @ X
    assert np.abs(Y - ref).max() <= 1e-13 * np.abs(ref).max()


def test_spmm_tiled_unsorted_rows_and_bad_columns(dev):
    """The bucketing takes each entry by its own column: rows whose columns
    descend (across tile edges and within one tile) give scipy's product;
    a column index outside [0, n) is refused before anything is written."""
    n = 20000
    rng = np.random.default_rng(5)
    rp, ci, v = _orbital_cached(n, 24, 7)
    ci = ci.copy()
    for r in range(0, n, 7):  # reverse every seventh row
        ci[rp[r]:rp[r + 1]] = ci[rp[r]:rp[r + 1]][::-1]
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    X = rng.standard_normal((n, 2))
    drp, dci, dv, dX = dev.put(rp), dev.put(ci), dev.put(v), dev.put(X)
    dY = dev.hip.malloc(n * 2 * 8)
    dev.ptrs.append(dY)
    api.csr_spmm_tiled(drp, dci, dv, n, dX, 2, dY, 2, 2)
    Y = dev.hip.to_host(dY, (n, 2), order="C")
    ref = A @ X
    assert np.abs(Y - ref).max() <= 1e-13 * np.abs(ref).max()
    bad = ci.copy()
    bad[5] = n
    dbad = dev.put(bad)
    with pytest.raises(api.IptError):
        api.csr_spmm_tiled(drp, dbad, dv, n, dX, 2, dY, 2, 2)


def _solve_device(dev, rp, ci, v, n, k, tol=1e-12, maxiter=300):
    drp, dci, dv = dev.put(rp), dev.put(ci), dev.put(v)
    dvec, dval, it, fp = api.ipt_cuda_sparse_csc_device_tol(drp, dci, dv, n, k, ci.size, tol, maxiter)
    try:
        vec = dev.hip.to_host(dvec, (n, k))
        val = dev.hip.to_host(dval, (k,))
    finally:
        api.free_device_result(dvec, dval)
    return vec, val, it, fp


@pytest.mark.parametrize("n,m,k,seed", [(3000, 20, 8, 1), (4000, 25, 16, 5)])
def test_device_sparse_block_orbital(block_env, dev, n, m, k, seed):
    """Eigenpairs of the orbital-sum model from the device-resident path:
    LAPACK's eigenvalues to 1e-10, residual <= 1e-10, and the oracle's
    block pipeline on the same CSC gives the same values."""
    rp, ci, v = orc.orbital_csr(n, m, seed)
    vec, val, it, fp = _solve_device(dev, rp, ci, v, n, k)
    H = sp.csr_matrix((v, ci, rp), shape=(n, n)).toarray()
    w = sl.eigh(H, eigvals_only=True, subset_by_index=[0, k - 1])
    assert np.abs(val - w).max() <= 1e-10 * max(1.0, np.abs(w).max())
    assert orc.residuals(vec, val, dense=H).max() <= 1e-10
    ov, _, oinfo = orc.block_pipeline(k, csc=(rp, ci, v), maxiter=300, tol=1e-12)
    assert np.abs(val - ov).max() <= 1e-10 * max(1.0, np.abs(ov).max())
    assert it == oinfo["max_iterations"]


def test_device_sparse_block_unsorted_diagonal(block_env, dev):
    """A permuted matrix (diagonal not sorted) takes the host preparation
    route and returns the same pairs in the caller's row order."""
    n, m, k = 1500, 12, 6
    rp, ci, v = orc.orbital_csr(n, m, 8)
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    p = np.random.default_rng(1).permutation(n)
    Ap = sp.csc_matrix(A.toarray()[np.ix_(p, p)])
    Ap.sort_indices()
    vec, val, it, fp = _solve_device(dev, Ap.indptr.astype(np.int32), Ap.indices.astype(np.int32), Ap.data, n, k)
    vec0, val0, it0, fp0 = _solve_device(dev, rp, ci, v, n, k)
    assert np.abs(val - val0).max() <= 1e-12 * max(1.0, np.abs(val0).max())
    # same vectors up to sign, rows permuted back
    for j in range(k):
        a, b = vec[:, j], vec0[p, j]
        assert min(np.abs(a - b).max(), np.abs(a + b).max()) <= 1e-9


def test_device_sparse_block_nonsymmetric_rejected(block_env, dev):
    n, m = 500, 6
    rp, ci, v = orc.orbital_csr(n, m, 2)
    v = v.copy()
    off = np.nonzero(ci != np.repeat(np.arange(n), np.diff(rp)))[0]
    v[off[3]] += 0.1
    with pytest.raises(api.IptError):
        _solve_device(dev, rp, ci, v, n, 4)


def test_device_sparse_block_tiled_narrow_products(block_env, dev):
    """The products of the sparse solve through the LDS-tiled layout (forced
    at n = 300000, below the automatic size, and for every width up to 64,
    so that each product of the solve takes it): the same pairs as with the
    row-major product to rounding, residuals <= 1e-10, and the tiled kernel
    ran."""
    n, m, k = 300000, 16, 8
    rp, ci, v = orc.orbital_csr(n, m, 3)
    block_env.setenv("IPT_CSR_TILES", "0")
    vec0, val0, _, _ = _solve_device(dev, rp, ci, v, n, k)
    block_env.setenv("IPT_CSR_TILES", "2")
    block_env.setenv("IPT_CSR_TILES_WIDTH", "64")
    api.profile_reset()
    api.profile_enable(True, only=["spmm_tiles"])
    try:
        vec, val, it, fp = _solve_device(dev, rp, ci, v, n, k)
    finally:
        api.profile_enable(False)
    _, products = api.profile_get("spmm_tiles")
    assert products > 0
    assert np.abs(val - val0).max() <= 1e-10 * max(1.0, np.abs(val0).max())
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    normH = np.abs(A).sum(axis=1).max()
    res = np.linalg.norm(A @ vec - vec * val, axis=0) / (max(1.0, normH) * np.maximum(1.0, np.linalg.norm(vec, axis=0)))
    assert res.max() <= 1e-10


@pytest.mark.parametrize("n,m,k,seed", [(3000, 20, 8, 1), (20000, 40, 32, 2), (300000, 16, 8, 3)])
def test_sparse_fp32_anchored(block_env, dev, n, m, k, seed):
    """IPT_PRECISION on CSR input: FP32 correction products anchored by FP64
    ones (the dense path's scheme); the pairs equal the FP64 block solve's
    and the solve used low-precision products.  At n = 300000 both the FP64
    and the FP32 products of at most 8 columns run in column-sliced
    passes."""
    rp, ci, v = orc.orbital_csr(n, m, seed)
    vec64, val64, _, _ = _solve_device(dev, rp, ci, v, n, k)
    st64 = api.last_solve_stats()
    block_env.setenv("IPT_PRECISION", "fp32")
    vec, val, it, fp = _solve_device(dev, rp, ci, v, n, k)
    st = api.last_solve_stats()
    assert st["lowp_products"] > 0 and st64["lowp_products"] == 0
    assert np.abs(val - val64).max() <= 1e-10 * max(1.0, np.abs(val64).max())
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    normH = np.abs(A).sum(axis=1).max()
    res = np.linalg.norm(A @ vec - vec * val, axis=0) / (max(1.0, normH) * np.maximum(1.0, np.linalg.norm(vec, axis=0)))
    assert res.max() <= 1e-10
    # host CSC entry (host preparation, same operator)
    hv, hl, info = api.ipt_cuda_sparse_csc_tol(rp, ci, v, k, 1e-12, 300)
    assert np.abs(hl - val64).max() <= 1e-10 * max(1.0, np.abs(val64).max())
