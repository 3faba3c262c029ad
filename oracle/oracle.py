"""ctypes front end of the CPU oracle (oracle/ipt_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point and
by the CPU-baseline leg of bench.py, always as the checker.  The product
library never imports this module.

Arrays are numpy float64 column-major (Fortran order) to match the C ABI of
the solver.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libipt_oracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libipt_prep_ref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

# Defaults of the degeneracy options, IPT/src/ipt_preparation.cu:16-21.
DEFAULT_DEG_OPTS = (0.0, 1000.0, 1.0e-12, 0.2, 0.02)
DEFAULT_MAX_BLOCK = 64


def build():
    """Compile the oracle (and, when the reference exists, oracle/_ref)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
    if os.path.isdir("/root/reference/IPT_C_GPU/src"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.orc_splitmix64.restype = ctypes.c_uint64
    lib.orc_splitmix64.argtypes = [ctypes.c_uint64]
    lib.orc_uniform01.restype = ctypes.c_double
    lib.orc_uniform01.argtypes = [ctypes.c_uint64, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong]
    lib.orc_gen_kat_dense.argtypes = [_dp, ctypes.c_int, ctypes.c_int]
    lib.orc_gen_eps_dense.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_uint64]
    lib.orc_gen_sorted_diagonal.argtypes = [_dp, ctypes.c_int, ctypes.c_uint64, ctypes.c_double, ctypes.c_double]
    lib.orc_gen_rand_dense.argtypes = [_dp, ctypes.c_int, ctypes.c_uint64, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.orc_gen_banded_csc.restype = ctypes.c_int
    lib.orc_gen_banded_csc.argtypes = [ctypes.c_int, _ip, _ip, _dp]
    lib.orc_inf_norm.restype = ctypes.c_double
    lib.orc_inf_norm.argtypes = [_dp, _ip, _ip, _dp, ctypes.c_int]
    ipt_args = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                _dp, _dp, _ip, _dp]
    lib.orc_ipt_dense.restype = ctypes.c_int
    lib.orc_ipt_dense.argtypes = [_dp] + ipt_args
    lib.orc_ipt_csc.restype = ctypes.c_int
    lib.orc_ipt_csc.argtypes = [_ip, _ip, _dp] + ipt_args
    lib.orc_target_clusters.restype = ctypes.c_int
    lib.orc_target_clusters.argtypes = [_dp, _ip, _ip, _dp, ctypes.c_int, ctypes.c_int, _dp,
                                        ctypes.c_int, _ip, _ip]
    lib.orc_block_cluster_solve.restype = ctypes.c_int
    lib.orc_block_cluster_solve.argtypes = [_dp, _ip, _ip, _dp, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                            ctypes.c_double, _dp, _ip, _dp]
    lib.orc_rayleigh_ritz.restype = ctypes.c_int
    lib.orc_rayleigh_ritz.argtypes = [_dp, _ip, _ip, _dp, ctypes.c_int, _dp, ctypes.c_int,
                                      ctypes.c_int, _dp, _dp]
    lib.orc_residuals.argtypes = [_dp, _ip, _ip, _dp, ctypes.c_int, _dp, _dp, ctypes.c_int,
                                  ctypes.c_double, _dp]
    lib.orc_sym_eig.argtypes = [_dp, ctypes.c_int, _dp, _dp]
    lib.orc_davidson.restype = ctypes.c_int
    lib.orc_davidson.argtypes = [_dp, _ip, _ip, _dp, ctypes.c_int, ctypes.c_int, _dp,
                                 ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double,
                                 _ip, _ip, _ip, ctypes.POINTER(ctypes.c_longlong)]
    lib.orc_block_pipeline.restype = ctypes.c_int
    lib.orc_block_pipeline.argtypes = [_dp, _ip, _ip, _dp, ctypes.c_int, ctypes.c_int, _dp,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                       _dp, _dp, _ip, _ip, ctypes.POINTER(ctypes.c_longlong)]
    lib.orc_num_threads.restype = ctypes.c_int
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(_ip)


def _f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


# ---------------------------------------------------------------- generators

def splitmix64(x):
    return lib().orc_splitmix64(int(x) & 0xFFFFFFFFFFFFFFFF)


def uniform01(seed, n, row, col):
    return lib().orc_uniform01(int(seed), int(n), int(row), int(col))


def kat_dense(n, symmetric):
    """Dense KAT matrix of IPT/test_temp/runtests.cu:59-80."""
    H = np.zeros((n, n), order="F")
    lib().orc_gen_kat_dense(_d(H), n, int(symmetric))
    return H


def eps_dense(n, eps, seed):
    """diag(i+1)+eps*U sweep matrix, benchmark_tf32_cusolver_sweep.cu:472-503."""
    H = np.zeros((n, n), order="F")
    lib().orc_gen_eps_dense(_d(H), n, float(eps), int(seed))
    return H


def sorted_diagonal(n, seed, lo, hi):
    d = np.zeros(n)
    lib().orc_gen_sorted_diagonal(_d(d), n, int(seed), float(lo), float(hi))
    return d


def rand_dense(n, seed, lo, hi, vscale):
    """diag(sorted U[lo,hi]) + symmetric N(0,1)*vscale/sqrt(n), zero diagonal."""
    H = np.zeros((n, n), order="F")
    lib().orc_gen_rand_dense(_d(H), n, int(seed), float(lo), float(hi), float(vscale))
    return H


def set_diagonal_gaps(H, positions=(10, 50, 120, 200), gap=1e-4):
    """Near-degenerate variant (in place): H[i+1, i+1] = H[i, i] + gap for
    each position i; the diagonal must stay sorted."""
    for i in positions:
        if not H[i, i] + gap < (H[i + 2, i + 2] if i + 2 < H.shape[0] else np.inf):
            raise ValueError("gap at %d would unsort the diagonal" % i)
        H[i + 1, i + 1] = H[i, i] + gap
    return H


def orbital_csr(n, m, seed, norb=40, nsub=8, elo=-20.0, ehi=5.0, sigma=0.02):
    """CSR (= CSC) of the orbital-sum sparse model: (row_ptr, col_ind, values)."""
    cap = n * (2 * m + 1)
    rp = np.zeros(n + 1, dtype=np.int32)
    ci = np.zeros(cap, dtype=np.int32)
    v = np.zeros(cap)
    f = lib().orc_gen_orbital_csr
    f.restype = ctypes.c_longlong
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                  ctypes.c_double, ctypes.c_double, _ip, _ip, _dp, ctypes.c_longlong]
    nnz = f(n, m, int(seed), norb, nsub, float(elo), float(ehi), float(sigma), _i(rp), _i(ci), _d(v), cap)
    if nnz < 0:
        raise RuntimeError("orc_gen_orbital_csr failed")
    return rp, ci[:nnz].copy(), v[:nnz].copy()


def orbital_diagonal(n, seed, norb=40, nsub=8, elo=-20.0, ehi=5.0):
    d = np.zeros(n)
    f = lib().orc_orbital_diagonal
    f.restype = None
    f.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]
    f(n, int(seed), norb, nsub, float(elo), float(ehi), _d(d))
    return d


def banded_csc(n):
    """Tridiagonal CSC of IPT/test/sparse/runtests_sparse.cu:224-268."""
    nnz = n + 2 * (n - 1)
    cp = np.zeros(n + 1, dtype=np.int32)
    ri = np.zeros(nnz, dtype=np.int32)
    v = np.zeros(nnz)
    got = lib().orc_gen_banded_csc(n, _i(cp), _i(ri), _d(v))
    assert got == nnz
    return cp, ri, v


# ---------------------------------------------------------------- solvers

class OracleError(RuntimeError):
    def __init__(self, status):
        super().__init__("oracle status %d" % status)
        self.status = status


def inf_norm(dense=None, csc=None, n=None):
    if dense is not None:
        dense = _f64(dense)
        return lib().orc_inf_norm(_d(dense), None, None, None, dense.shape[0])
    cp, ri, v = csc
    return lib().orc_inf_norm(None, _i(cp), _i(ri), _d(v), len(cp) - 1)


def ipt_dense(H, k, iterations=None, tol=None, maxiter=None, Z0=None):
    """Plain IPT; fixed `iterations`, or `tol`/`maxiter` stop.  Returns
    (vectors n x k, values k, iterations_done, fixed_point_residual)."""
    H = _f64(H)
    n = H.shape[0]
    use_tol = tol is not None
    mi = maxiter if use_tol else iterations
    vec = np.zeros((n, k), order="F")
    val = np.zeros(k)
    it = ctypes.c_int(0)
    fp = ctypes.c_double(0.0)
    z0 = None if Z0 is None else _f64(Z0)
    st = lib().orc_ipt_dense(_d(H), n, k, _d(z0), int(mi), float(tol or 0.0), int(use_tol),
                             _d(vec), _d(val), ctypes.byref(it), ctypes.byref(fp))
    if st != 0:
        raise OracleError(st)
    return vec, val, it.value, fp.value


def ipt_csc(cp, ri, v, k, iterations=None, tol=None, maxiter=None, Z0=None):
    cp = np.ascontiguousarray(cp, dtype=np.int32)
    ri = np.ascontiguousarray(ri, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.float64)
    n = len(cp) - 1
    use_tol = tol is not None
    mi = maxiter if use_tol else iterations
    vec = np.zeros((n, k), order="F")
    val = np.zeros(k)
    it = ctypes.c_int(0)
    fp = ctypes.c_double(0.0)
    z0 = None if Z0 is None else _f64(Z0)
    st = lib().orc_ipt_csc(_i(cp), _i(ri), _d(v), n, k, _d(z0), int(mi), float(tol or 0.0),
                           int(use_tol), _d(vec), _d(val), ctypes.byref(it), ctypes.byref(fp))
    if st != 0:
        raise OracleError(st)
    return vec, val, it.value, fp.value


def _mat_args(dense, csc):
    if dense is not None:
        dense = _f64(dense)
        return (dense, None, None, None, dense.shape[0])
    cp, ri, v = csc
    cp = np.ascontiguousarray(cp, dtype=np.int32)
    ri = np.ascontiguousarray(ri, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.float64)
    return (None, cp, ri, v, len(cp) - 1)


def target_clusters(k, dense=None, csc=None, opts=DEFAULT_DEG_OPTS, max_block=DEFAULT_MAX_BLOCK):
    """Cluster ranges (list of (first,last)) incl. singleton targets, too_large."""
    dn, cp, ri, v, n = _mat_args(dense, csc)
    o = np.asarray(opts, dtype=np.float64)
    out = np.zeros(4 * n + 4, dtype=np.int32)
    tl = ctypes.c_int(0)
    nr = lib().orc_target_clusters(_d(dn), _i(cp), _i(ri), _d(v), n, k, _d(o), int(max_block),
                                   _i(out), ctypes.byref(tl))
    return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(nr)], bool(tl.value)


def ref_target_clusters(cp, ri, v, k, degeneracy_threshold=0.0):
    """The reference's own ipt_target_degenerate_subspaces (oracle/_ref),
    before singleton completion.  None when oracle/_ref is not built."""
    if not os.path.exists(_REF_PATH):
        return None
    rl = ctypes.CDLL(_REF_PATH)
    rl.ref_target_clusters.restype = ctypes.c_int
    rl.ref_target_clusters.argtypes = [_ip, _ip, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                       _ip, _ip]
    cp = np.ascontiguousarray(cp, dtype=np.int32)
    ri = np.ascontiguousarray(ri, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.float64)
    n = len(cp) - 1
    out = np.zeros(4 * n + 4, dtype=np.int32)
    tl = ctypes.c_int(0)
    nr = rl.ref_target_clusters(_i(cp), _i(ri), _d(v), n, k, float(degeneracy_threshold), _i(out),
                                ctypes.byref(tl))
    return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(nr)], bool(tl.value)


def _ref_csc_args(cp, ri, v):
    return (np.ascontiguousarray(cp, dtype=np.int32), np.ascontiguousarray(ri, dtype=np.int32),
            np.ascontiguousarray(v, dtype=np.float64))


def ref_prepare_sparse_csc(cp, ri, v, k, degeneracy_threshold=0.0):
    """The reference's own plain-CSC preparation (sort + degeneracy lifting,
    IPT/src/ipt_preparation.cu:1375-1648) from oracle/_ref: returns
    (status, prepared dense n x n, perm[sorted] = original, rotations), or
    None when oracle/_ref is not built."""
    if not os.path.exists(_REF_PATH):
        return None
    rl = ctypes.CDLL(_REF_PATH)
    f = rl.ref_prepare_sparse_csc
    f.restype = ctypes.c_int
    f.argtypes = [_ip, _ip, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _ip, _ip]
    cp, ri, v = _ref_csc_args(cp, ri, v)
    n = len(cp) - 1
    dense = np.zeros((n, n), order="F")
    perm = np.zeros(n, dtype=np.int32)
    rot = ctypes.c_int(0)
    st = f(_i(cp), _i(ri), _d(v), n, int(k), int(cp[-1]), float(degeneracy_threshold), _d(dense), _i(perm),
           ctypes.byref(rot))
    return st, dense, perm, rot.value


def ref_prepared_vectors_to_original(cp, ri, v, k, X, degeneracy_threshold=0.0):
    """The reference's ipt_apply_prepared_q_to_vectors on a copy of X (n x kv)."""
    if not os.path.exists(_REF_PATH):
        return None
    rl = ctypes.CDLL(_REF_PATH)
    f = rl.ref_prepared_vectors_to_original
    f.restype = ctypes.c_int
    f.argtypes = [_ip, _ip, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int]
    cp, ri, v = _ref_csc_args(cp, ri, v)
    X = np.array(X, dtype=np.float64, order="F", copy=True)
    st = f(_i(cp), _i(ri), _d(v), len(cp) - 1, int(k), int(cp[-1]), float(degeneracy_threshold), _d(X),
           X.shape[1])
    if st != 0:
        raise OracleError(st)
    return X


def block_cluster_solve(first, last, maxiter, tol, damping=1.0, dense=None, csc=None):
    dn, cp, ri, v, n = _mat_args(dense, csc)
    m = last - first + 1
    B = np.zeros((n, m), order="F")
    it = ctypes.c_int(0)
    fp = ctypes.c_double(0.0)
    st = lib().orc_block_cluster_solve(_d(dn), _i(cp), _i(ri), _d(v), n, first, last, maxiter,
                                       float(tol), float(damping), _d(B), ctypes.byref(it),
                                       ctypes.byref(fp))
    if st != 0:
        raise OracleError(st)
    return B, it.value, fp.value


def rayleigh_ritz(B, k, dense=None, csc=None):
    dn, cp, ri, v, n = _mat_args(dense, csc)
    B = _f64(B)
    bc = B.shape[1]
    val = np.zeros(k)
    vec = np.zeros((n, k), order="F")
    st = lib().orc_rayleigh_ritz(_d(dn), _i(cp), _i(ri), _d(v), n, _d(B), bc, k, _d(val), _d(vec))
    if st != 0:
        raise OracleError(st)
    return val, vec


def residuals(X, theta, dense=None, csc=None, normH=None):
    """Scaled residuals ||Hx-theta x||/(max(1,||H||_inf) max(1,||x||))."""
    dn, cp, ri, v, n = _mat_args(dense, csc)
    X = _f64(X)
    th = np.ascontiguousarray(theta, dtype=np.float64)
    k = X.shape[1]
    if normH is None:
        normH = lib().orc_inf_norm(_d(dn), _i(cp), _i(ri), _d(v), n)
    res = np.zeros(k)
    lib().orc_residuals(_d(dn), _i(cp), _i(ri), _d(v), n, _d(X), _d(th), k, float(normH), _d(res))
    return res


def sym_eig(A):
    A = np.array(A, dtype=np.float64, order="F")
    m = A.shape[0]
    w = np.zeros(m)
    V = np.zeros((m, m), order="F")
    lib().orc_sym_eig(_d(A), m, _d(w), _d(V))
    return w, V


def block_pipeline(k, dense=None, csc=None, opts=DEFAULT_DEG_OPTS, max_block=DEFAULT_MAX_BLOCK,
                   maxiter=100, tol=1e-12, damping=1.0, davidson_steps=0, conv_tol=1e-13):
    """Clusters -> per-cluster block IPT -> Cholesky RR (-> Davidson).
    Returns (values, vectors, info dict)."""
    dn, cp, ri, v, n = _mat_args(dense, csc)
    o = np.asarray(opts, dtype=np.float64)
    val = np.zeros(k)
    vec = np.zeros((n, k), order="F")
    bc = ctypes.c_int(0)
    mi = ctypes.c_int(0)
    mv = ctypes.c_longlong(0)
    st = lib().orc_block_pipeline(_d(dn), _i(cp), _i(ri), _d(v), n, k, _d(o), int(max_block),
                                  int(maxiter), float(tol), float(damping), int(davidson_steps),
                                  float(conv_tol), _d(val), _d(vec), ctypes.byref(bc),
                                  ctypes.byref(mi), ctypes.byref(mv))
    if st != 0:
        raise OracleError(st)
    return val, vec, {"basis_cols": bc.value, "max_iterations": mi.value, "matvecs": mv.value}


def num_threads():
    return lib().orc_num_threads()


def set_num_threads(t):
    lib().orc_set_num_threads(int(t))


# ------------------------------------------------------------------ learned preconditioner

MLP_NPARAM = 4481


def gelu(x):
    f = lib().orc_gelu
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_double]
    return f(float(x))


def mlp_forward(params, feat):
    params, feat = _f64(params), _f64(feat)
    n = feat.size // 3
    y = np.zeros(n)
    f = lib().orc_mlp_forward
    f.restype = None
    f(_d(params), _d(feat), ctypes.c_long(n), _d(y))
    return y


def mlp_loss(params, feat, q, t, w):
    params, feat, q, t, w = (_f64(a) for a in (params, feat, q, t, w))
    f = lib().orc_mlp_loss
    f.restype = ctypes.c_double
    return f(_d(params), _d(feat), _d(q), _d(t), _d(w), ctypes.c_long(q.size))


def mlp_grad(params, feat, q, t, w):
    params, feat, q, t, w = (_f64(a) for a in (params, feat, q, t, w))
    g = np.zeros(MLP_NPARAM)
    f = lib().orc_mlp_grad
    f.restype = ctypes.c_double
    loss = f(_d(params), _d(feat), _d(q), _d(t), _d(w), ctypes.c_long(q.size), _d(g))
    return g, loss
