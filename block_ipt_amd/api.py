"""Python mirror of the reference's C entry points (IPT/src/ipt_cuda.cu:262-308).

Each function takes numpy arrays (column-major where the C ABI says so),
calls libipt_hip.so, and returns numpy results plus the IPTCudaResult
fields.  Device-memory entry points take raw device pointers; the small
`hip` helper below (ctypes on the HIP runtime the library already loaded)
allocates and copies device memory for callers without another framework.
"""
import ctypes
import os

import numpy as np

from . import abi
from .abi import IptError, IPTCudaResult, lib

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(_ip)


def result_to_dict(r):
    out = {}
    for name, _ in IPTCudaResult._fields_:
        if name in ("vectors", "values", "davidson_history", "davidson_selection_history",
                    "davidson_block_history"):
            continue
        v = getattr(r, name)
        out[name] = v.decode(errors="replace") if isinstance(v, bytes) else v
    return out


def _finish(status, r, where):
    if status != 0:
        raise IptError(status, where)
    n, k = r.n, r.k
    vec = np.ctypeslib.as_array(r.vectors, shape=(n * k,)).copy().reshape((n, k), order="F")
    val = np.ctypeslib.as_array(r.values, shape=(k,)).copy()
    info = result_to_dict(r)
    lib().ipt_cuda_free_result(ctypes.byref(r))
    return vec, val, info


# ------------------------------------------------------------------ dense host

def ipt_cuda(H, k, iterations):
    H = _f64(H)
    r = IPTCudaResult()
    st = lib().ipt_cuda(_d(H), H.shape[0], int(k), int(iterations), ctypes.byref(r))
    return _finish(st, r, "ipt_cuda")


def ipt_cuda_tol(H, k, tol, maxiter):
    H = _f64(H)
    r = IPTCudaResult()
    st = lib().ipt_cuda_tol(_d(H), H.shape[0], int(k), float(tol), int(maxiter), ctypes.byref(r))
    return _finish(st, r, "ipt_cuda_tol")


def ipt_cuda_with_initial(H, Z0, k, iterations):
    H, Z0 = _f64(H), _f64(Z0)
    r = IPTCudaResult()
    st = lib().ipt_cuda_with_initial(_d(H), _d(Z0), H.shape[0], int(k), int(iterations), ctypes.byref(r))
    return _finish(st, r, "ipt_cuda_with_initial")


def ipt_cuda_with_initial_tol(H, Z0, k, tol, maxiter):
    H, Z0 = _f64(H), _f64(Z0)
    r = IPTCudaResult()
    st = lib().ipt_cuda_with_initial_tol(_d(H), _d(Z0), H.shape[0], int(k), float(tol), int(maxiter),
                                         ctypes.byref(r))
    return _finish(st, r, "ipt_cuda_with_initial_tol")


# ------------------------------------------------------------------ sparse host

def ipt_cuda_sparse_csc(col_ptr, row_ind, values, k, iterations):
    cp, ri, v = _i32(col_ptr), _i32(row_ind), np.ascontiguousarray(values, dtype=np.float64)
    n = len(cp) - 1
    r = IPTCudaResult()
    st = lib().ipt_cuda_sparse_csc(_i(cp), _i(ri), _d(v), n, int(k), int(cp[-1]), int(iterations),
                                   ctypes.byref(r))
    return _finish(st, r, "ipt_cuda_sparse_csc")


def ipt_cuda_sparse_csc_tol(col_ptr, row_ind, values, k, tol, maxiter):
    cp, ri, v = _i32(col_ptr), _i32(row_ind), np.ascontiguousarray(values, dtype=np.float64)
    n = len(cp) - 1
    r = IPTCudaResult()
    st = lib().ipt_cuda_sparse_csc_tol(_i(cp), _i(ri), _d(v), n, int(k), int(cp[-1]), float(tol),
                                       int(maxiter), ctypes.byref(r))
    return _finish(st, r, "ipt_cuda_sparse_csc_tol")


def ipt_cuda_sparse_csc_with_initial_tol(col_ptr, row_ind, values, Z0, k, tol, maxiter):
    cp, ri, v = _i32(col_ptr), _i32(row_ind), np.ascontiguousarray(values, dtype=np.float64)
    Z0 = _f64(Z0)
    n = len(cp) - 1
    r = IPTCudaResult()
    st = lib().ipt_cuda_sparse_csc_with_initial_tol(_i(cp), _i(ri), _d(v), _d(Z0), n, int(k), int(cp[-1]),
                                                    float(tol), int(maxiter), ctypes.byref(r))
    return _finish(st, r, "ipt_cuda_sparse_csc_with_initial_tol")


# ------------------------------------------------------------------ HIP plumbing

class _Hip:
    """Minimal ctypes view of the HIP runtime already loaded by the library."""

    def __init__(self):
        lib()  # make sure the runtime the library uses is the one we bind
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        h.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        h.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
        h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        h.hipEventDestroy.argtypes = [ctypes.c_void_p]
        h.hipGetErrorString.restype = ctypes.c_char_p
        self.h = h

    def check(self, e, what):
        if e != 0:
            raise RuntimeError("%s: HIP error %d %s" % (what, e, self.h.hipGetErrorString(e).decode()))

    def device_count(self):
        c = ctypes.c_int(0)
        e = self.h.hipGetDeviceCount(ctypes.byref(c))
        return c.value if e == 0 else 0

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        self.check(self.h.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(int(nbytes), 1))), "hipMalloc")
        return p.value

    def free(self, p):
        if p:
            self.h.hipFree(ctypes.c_void_p(p))

    def to_device(self, arr):
        arr = np.asarray(arr)
        p = self.malloc(arr.nbytes)
        self.check(self.h.hipMemcpy(ctypes.c_void_p(p), arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, 1),
                   "hipMemcpy H2D")
        return p

    def to_host(self, p, shape, dtype=np.float64, order="F"):
        out = np.empty(shape, dtype=dtype, order=order)
        self.check(self.h.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(p), out.nbytes, 2),
                   "hipMemcpy D2H")
        return out

    def synchronize(self):
        self.check(self.h.hipDeviceSynchronize(), "hipDeviceSynchronize")


_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = _Hip()
    return _hip


# ------------------------------------------------------------------ device

def ipt_cuda_device_tol(d_H, n, k, tol, maxiter, handle=None):
    """Device entry point; returns (d_vectors, d_values, iterations, fp)
    as raw pointers (free with free_device_result)."""
    dv, dl = ctypes.c_void_p(), ctypes.c_void_p()
    it = ctypes.c_int(0)
    fp = ctypes.c_double(0.0)
    st = lib().ipt_cuda_device_tol(ctypes.c_void_p(d_H), int(n), int(k), float(tol), int(maxiter),
                                   handle, ctypes.byref(dv), ctypes.byref(dl), ctypes.byref(it),
                                   ctypes.byref(fp))
    if st != 0:
        raise IptError(st, "ipt_cuda_device_tol")
    return dv.value, dl.value, it.value, fp.value


def ipt_cuda_device(d_H, n, k, iterations, handle=None):
    dv, dl = ctypes.c_void_p(), ctypes.c_void_p()
    st = lib().ipt_cuda_device(ctypes.c_void_p(d_H), int(n), int(k), int(iterations), handle,
                               ctypes.byref(dv), ctypes.byref(dl))
    if st != 0:
        raise IptError(st, "ipt_cuda_device")
    return dv.value, dl.value


def ipt_cuda_sparse_csc_device_tol(d_cp, d_ri, d_v, n, k, nnz, tol, maxiter, handle=None):
    dv, dl = ctypes.c_void_p(), ctypes.c_void_p()
    it = ctypes.c_int(0)
    fp = ctypes.c_double(0.0)
    st = lib().ipt_cuda_sparse_csc_device_tol(ctypes.c_void_p(d_cp), ctypes.c_void_p(d_ri),
                                              ctypes.c_void_p(d_v), int(n), int(k), int(nnz), float(tol),
                                              int(maxiter), handle, ctypes.byref(dv), ctypes.byref(dl),
                                              ctypes.byref(it), ctypes.byref(fp))
    if st != 0:
        raise IptError(st, "ipt_cuda_sparse_csc_device_tol")
    return dv.value, dl.value, it.value, fp.value


def free_device_result(d_vectors, d_values):
    lib().ipt_cuda_free_device_result(ctypes.c_void_p(d_vectors), ctypes.c_void_p(d_values))


def create_handle(stream=None):
    h = ctypes.c_void_p()
    st = lib().ipt_hip_create_handle(ctypes.byref(h), ctypes.c_void_p(stream) if stream else None)
    if st != 0:
        raise IptError(st, "ipt_hip_create_handle")
    return h.value


def destroy_handle(h):
    return lib().ipt_hip_destroy_handle(ctypes.c_void_p(h))


_MIXED = {
    "fp16": ("ipt_mixed_precision_cuda_device_tol", "ipt_mixed_precision_cuda_free_device_result",
             abi.IPTMixedCudaDeviceResult),
    "fp32": ("ipt_mixed_precision_fp32_cuda_device_tol", "ipt_mixed_precision_fp32_cuda_free_device_result",
             abi.IPTMixedFP32CudaDeviceResult),
    "tf32": ("ipt_mixed_precision_tf32_cuda_device_tol", "ipt_mixed_precision_tf32_cuda_free_device_result",
             abi.IPTMixedTF32CudaDeviceResult),
    "rayleigh_tf32": ("ipt_rayleigh_mixed_tf32_cuda_device_tol", "ipt_rayleigh_mixed_cuda_free_device_result",
                      abi.IPTRayleighMixedCudaDeviceResult),
}


def mixed_device_tol(kind, d_H, n, k, lo_tol, lo_maxiter, hi_tol, hi_maxiter, handle=None):
    """Two-phase mixed-precision entry (include/ipt_hip_mixed.h); kind is
    fp16 | fp32 | tf32 | rayleigh_tf32.  Returns (vectors, values, info)
    copied to the host; the device result is freed."""
    name, free_name, R = _MIXED[kind]
    r = R()
    st = getattr(lib(), name)(ctypes.c_void_p(d_H), int(n), int(k), float(lo_tol), int(lo_maxiter),
                              float(hi_tol), int(hi_maxiter), handle, ctypes.byref(r))
    if st != 0:
        raise IptError(st, name)
    info = {f[0]: getattr(r, f[0]) for f in R._fields_ if not f[0].startswith("d_")}
    try:
        vec = hip().to_host(r.d_vectors, (n, k))
        val = hip().to_host(r.d_values, (k,))
    finally:
        getattr(lib(), free_name)(ctypes.byref(r))
    return vec, val, info


def acx_cuda_device_tol(d_H, n, k, tol, maxiter, handle=None, copy=True):
    """ACX-accelerated plain IPT (include/ipt_hip_acx.h).  Returns
    (vectors, values, info) copied to the host with the device result freed,
    or with copy=False the ACXCudaResult itself (free with acx_free_result)."""
    r = abi.ACXCudaResult()
    st = lib().acx_cuda_device_tol(ctypes.c_void_p(d_H), int(n), int(k), float(tol), int(maxiter), handle,
                                   ctypes.byref(r))
    if st != 0:
        raise IptError(st, "acx_cuda_device_tol")
    if not copy:
        return r
    info = {f[0]: getattr(r, f[0]) for f in abi.ACXCudaResult._fields_ if not f[0].startswith("d_")}
    try:
        vec = hip().to_host(r.d_vectors, (n, k))
        val = hip().to_host(r.d_values, (k,))
    finally:
        lib().acx_cuda_free_result(ctypes.byref(r))
    return vec, val, info


def acx_free_result(r):
    lib().acx_cuda_free_result(ctypes.byref(r))


_ACX_MIXED = {
    "tf32": ("acx_mixed_precision_tf32_cuda_device_tol", "acx_mixed_precision_tf32_cuda_free_result",
             abi.ACXMixedTF32CudaResult),
    "rayleigh_tf32": ("acx_rayleigh_mixed_tf32_cuda_device_tol", "acx_rayleigh_mixed_tf32_cuda_free_result",
                      abi.ACXRayleighMixedTF32CudaResult),
}


def acx_mixed_device_tol(kind, d_H, n, k, lo_tol, lo_maxiter, hi_tol, hi_maxiter, handle=None):
    """Two-phase ACX entry (include/ipt_hip_acx.h); kind is tf32 |
    rayleigh_tf32.  Returns (vectors, values, info) copied to the host."""
    name, free_name, R = _ACX_MIXED[kind]
    r = R()
    st = getattr(lib(), name)(ctypes.c_void_p(d_H), int(n), int(k), float(lo_tol), int(lo_maxiter),
                              float(hi_tol), int(hi_maxiter), handle, ctypes.byref(r))
    if st != 0:
        raise IptError(st, name)
    info = {f[0]: getattr(r, f[0]) for f in R._fields_ if not f[0].startswith("d_")}
    try:
        vec = hip().to_host(r.d_vectors, (n, k))
        val = hip().to_host(r.d_values, (k,))
    finally:
        getattr(lib(), free_name)(ctypes.byref(r))
    return vec, val, info


# ------------------------------------------------------------------ extensions

LOWP_MODES = {"fp16": 0, "bf16": 1, "bf16x3": 2, "fp32": 3, "fp8": 4}


def generate_dense(d_H, n, seed, lo, hi, vscale, handle=None):
    """Fill device buffer d_H (n*n doubles) with the synthetic dense matrix
    described in include/ipt_hip_ext.h."""
    st = lib().ipt_hip_generate_dense(ctypes.c_void_p(d_H), int(n), int(seed), float(lo), float(hi),
                                      float(vscale), handle)
    if st != 0:
        raise IptError(st, "ipt_hip_generate_dense")


def generate_orbital_csr(n, m, seed, d_row_ptr, d_col_ind, d_values, capacity, norb=40, nsub=8, elo=-20.0,
                         ehi=5.0, sigma=0.02, handle=None):
    """Device CSR of the orbital-sum sparse model; returns nnz."""
    nnz = ctypes.c_longlong(0)
    st = lib().ipt_hip_generate_orbital_csr(int(n), int(m), int(seed), int(norb), int(nsub), float(elo), float(ehi),
                                            float(sigma), _vpp(d_row_ptr), _vpp(d_col_ind), _vpp(d_values),
                                            int(capacity), ctypes.byref(nnz), handle)
    if st != 0:
        raise IptError(st, "ipt_hip_generate_orbital_csr")
    return nnz.value


# Solver options for k = 256 pairs of a strongly coupled dense matrix (the
# N=65536 FP8 workload): FP8 correction products anchored every 5 steps,
# clusters that stop contracting (or exceed the block limit) handed to
# Rayleigh-Ritz and Davidson, whose search space holds 3k columns and keeps
# the last k corrections over a restart (the reference's limits, k + 2k/3
# columns and all but one slot kept, leave one new correction per step at
# this k).
MANY_PAIRS_OPTIONS = {
    "IPT_BLOCK_CLUSTER": "1",
    "IPT_PRECISION": "fp8",
    "IPT_MIXED_ANCHOR_EVERY": "5",
    "IPT_COUPLING_GAP_RATIO_TOL": "0.5",
    "IPT_BLOCK_CLUSTER_STALL": "2",
    "IPT_BLOCK_CLUSTER_MAXITER": "40",
    "IPT_DAVIDSON_ENRICH": "1",
    "IPT_DAVIDSON_STEPS": "80",
    "IPT_DAVIDSON_CONVERGED_TOL": "1e-10",
    "IPT_DAVIDSON_MAX_BASIS": "768",
    "IPT_DAVIDSON_RETAIN": "256",
}

# the near-degenerate variant of the dense N=65536, k=256 workload
NEAR_DEGENERATE_POSITIONS = (10, 50, 120, 200)
NEAR_DEGENERATE_GAP = 1e-4


def set_diagonal_gaps(d_H, n, positions=NEAR_DEGENERATE_POSITIONS, gap=NEAR_DEGENERATE_GAP, handle=None):
    pos = np.ascontiguousarray(positions, dtype=np.int32)
    st = lib().ipt_hip_set_diagonal_gaps(ctypes.c_void_p(d_H), int(n), _i(pos), len(pos), float(gap), handle)
    if st != 0:
        raise IptError(st, "ipt_hip_set_diagonal_gaps")


def generate_eps(d_H, n, eps, seed, handle=None):
    """diag(i+1) + eps U sweep matrix on the device (include/ipt_hip_ext.h)."""
    st = lib().ipt_hip_generate_eps(ctypes.c_void_p(d_H), int(n), float(eps), int(seed), handle)
    if st != 0:
        raise IptError(st, "ipt_hip_generate_eps")


def csr_spmm(d_row_ptr, d_col_ind, d_values, n, d_X, ldx, d_Y, ldy, k, handle=None):
    st = lib().ipt_hip_csr_spmm(_vpp(d_row_ptr), _vpp(d_col_ind), _vpp(d_values), int(n), _vpp(d_X), int(ldx),
                                _vpp(d_Y), int(ldy), int(k), handle)
    if st != 0:
        raise IptError(st, "ipt_hip_csr_spmm")


def csr_spmm_sliced(d_row_ptr, d_col_ind, d_values, n, d_X, ldx, d_Y, ldy, k, slices=0, handle=None):
    """The CSR product in column-sliced passes (columns ascending per row)."""
    st = lib().ipt_hip_csr_spmm_sliced(_vpp(d_row_ptr), _vpp(d_col_ind), _vpp(d_values), int(n), _vpp(d_X),
                                       int(ldx), _vpp(d_Y), int(ldy), int(k), int(slices), handle)
    if st != 0:
        raise IptError(st, "ipt_hip_csr_spmm_sliced")


def csr_spmm_tiled(d_row_ptr, d_col_ind, d_values, n, d_X, ldx, d_Y, ldy, k, handle=None):
    """The CSR product through the LDS-tiled layout (built on each call;
    any column order within a row, n <= 2^25)."""
    st = lib().ipt_hip_csr_spmm_tiled(_vpp(d_row_ptr), _vpp(d_col_ind), _vpp(d_values), int(n), _vpp(d_X),
                                      int(ldx), _vpp(d_Y), int(ldy), int(k), handle)
    if st != 0:
        raise IptError(st, "ipt_hip_csr_spmm_tiled")


def cholqr_mixed(d_Q, n, k, handle=None):
    st = lib().ipt_hip_cholqr_mixed(_vpp(d_Q), int(n), int(k), handle)
    if st != 0:
        raise IptError(st, "ipt_hip_cholqr_mixed")


def lowp_product(d_H, n, mode, d_B, d_Ba, d_Ya, d_diag, nc, d_Y, handle=None):
    st = lib().ipt_hip_lowp_product(ctypes.c_void_p(d_H), int(n), LOWP_MODES.get(mode, mode),
                                    ctypes.c_void_p(d_B), ctypes.c_void_p(d_Ba), ctypes.c_void_p(d_Ya),
                                    ctypes.c_void_p(d_diag), int(nc), ctypes.c_void_p(d_Y), handle)
    if st != 0:
        raise IptError(st, "ipt_hip_lowp_product")


def dense_product(d_H, n, d_Z, nc, d_Y, packed=False, handle=None):
    """Y = H Z in FP64 on the block path's MFMA kernel (device pointers)."""
    st = lib().ipt_hip_dense_product(ctypes.c_void_p(d_H), int(n), ctypes.c_void_p(d_Z), int(nc),
                                     ctypes.c_void_p(d_Y), 1 if packed else 0, handle)
    if st != 0:
        raise IptError(st, "ipt_hip_dense_product")


def profile_enable(on=True, only=None):
    """Event timing of the library's kernel families; `only` (a list of
    family names) times just those."""
    if only is None:
        lib().ipt_hip_profile_enable(1 if on else 0)
        return
    st = lib().ipt_hip_profile_enable_only(",".join(only).encode() if on else b"")
    if st != 0:
        raise IptError(st, "ipt_hip_profile_enable_only")


def profile_reset():
    lib().ipt_hip_profile_reset()


def profile_get(kernel):
    ms = ctypes.c_double(0.0)
    n = ctypes.c_longlong(0)
    st = lib().ipt_hip_profile_get(kernel.encode(), ctypes.byref(ms), ctypes.byref(n))
    if st != 0:
        raise IptError(st, "ipt_hip_profile_get")
    return ms.value, n.value


def profile_columns(kernel):
    """Columns processed by the timed launches of one kernel family."""
    c = ctypes.c_longlong(0)
    st = lib().ipt_hip_profile_get_columns(kernel.encode(), ctypes.byref(c))
    if st != 0:
        raise IptError(st, "ipt_hip_profile_get_columns")
    return c.value


def last_solve_stats(handle=None):
    b, s_, dav = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    f64, lp = ctypes.c_longlong(0), ctypes.c_longlong(0)
    st = lib().ipt_hip_last_solve_stats(handle, ctypes.byref(b), ctypes.byref(s_), ctypes.byref(f64),
                                        ctypes.byref(lp), ctypes.byref(dav))
    if st != 0:
        raise IptError(st, "ipt_hip_last_solve_stats")
    return {"basis_cols": b.value, "steps": s_.value, "fp64_products": f64.value, "lowp_products": lp.value,
            "davidson_steps": dav.value}


def last_cluster_ranges(handle=None):
    """(first, last) row ranges (sorted-diagonal order, singletons
    included) of the last block-cluster solve."""
    cnt = ctypes.c_int(0)
    st = lib().ipt_hip_last_cluster_ranges(handle, None, 0, ctypes.byref(cnt))
    if st != 0:
        raise IptError(st, "ipt_hip_last_cluster_ranges")
    out = np.zeros(2 * max(cnt.value, 1), dtype=np.int32)
    st = lib().ipt_hip_last_cluster_ranges(handle, _i(out), cnt.value, ctypes.byref(cnt))
    if st != 0:
        raise IptError(st, "ipt_hip_last_cluster_ranges")
    return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(cnt.value)]


def sym_eig(A, handle=None):
    """Eigenvalues and eigenvectors of the symmetric host matrix A through
    the library's device eigensolver."""
    A = np.asfortranarray(np.asarray(A, dtype=np.float64))
    m = A.shape[0]
    h = hip()
    dA, dw, dV = h.to_device(A), h.malloc(m * 8), h.malloc(m * m * 8)
    try:
        st = lib().ipt_hip_sym_eig(ctypes.c_void_p(dA), m, m, ctypes.c_void_p(dw), ctypes.c_void_p(dV), m, handle)
        if st != 0:
            raise IptError(st, "ipt_hip_sym_eig")
        return h.to_host(dw, (m,)), h.to_host(dV, (m, m))
    finally:
        for p in (dA, dw, dV):
            h.free(p)


def last_cluster_states(handle=None):
    """Per cluster of the last block solve: 0 converged, 1 iterating at the
    limit, 2 stalled."""
    cnt = ctypes.c_int(0)
    st = lib().ipt_hip_last_cluster_states(handle, None, 0, ctypes.byref(cnt))
    if st != 0:
        raise IptError(st, "ipt_hip_last_cluster_states")
    out = np.zeros(max(cnt.value, 1), dtype=np.int32)
    st = lib().ipt_hip_last_cluster_states(handle, _i(out), cnt.value, ctypes.byref(cnt))
    if st != 0:
        raise IptError(st, "ipt_hip_last_cluster_states")
    return [int(x) for x in out[:cnt.value]]


def last_davidson_stats(handle=None):
    acc, rej, ma, mr = (ctypes.c_int(0) for _ in range(4))
    mv, res = ctypes.c_longlong(0), ctypes.c_double(0.0)
    st = lib().ipt_hip_last_davidson_stats(handle, ctypes.byref(acc), ctypes.byref(rej), ctypes.byref(ma),
                                           ctypes.byref(mr), ctypes.byref(mv), ctypes.byref(res))
    if st != 0:
        raise IptError(st, "ipt_hip_last_davidson_stats")
    return {"accepted": acc.value, "rejected": rej.value, "mlp_accepted": ma.value, "mlp_rejected": mr.value,
            "matvecs": mv.value, "residual_after": res.value}


# ------------------------------------------------------------------ learned preconditioner

MLP_NPARAM = 4481


def _vpp(p):
    return ctypes.c_void_p(p) if p else None


def mlp_infer(d_params, d_feat, d_qf, count, d_out, handle=None):
    st = lib().ipt_hip_mlp_infer(_vpp(d_params), _vpp(d_feat), _vpp(d_qf), int(count), _vpp(d_out), handle)
    if st != 0:
        raise IptError(st, "ipt_hip_mlp_infer")


def mlp_gradient(d_params, d_feat, d_q, d_t, d_w, count, d_grad, d_loss, handle=None):
    st = lib().ipt_hip_mlp_gradient(_vpp(d_params), _vpp(d_feat), _vpp(d_q), _vpp(d_t), _vpp(d_w), int(count),
                                    _vpp(d_grad), _vpp(d_loss), handle)
    if st != 0:
        raise IptError(st, "ipt_hip_mlp_gradient")


def mlp_gradient_noalloc(d_params, d_feat, d_q, d_t, d_w, count, d_grad, d_loss, handle=None):
    """ipt_hip_mlp_gradient through the entry point that allocates nothing."""
    st = lib().ipt_hip_mlp_gradient_noalloc(_vpp(d_params), _vpp(d_feat), _vpp(d_q), _vpp(d_t), _vpp(d_w),
                                            int(count), _vpp(d_grad), _vpp(d_loss), handle)
    if st != 0:
        raise IptError(st, "ipt_hip_mlp_gradient_noalloc")


def mlp_adam_step(d_params, d_m, d_v, d_grad, gscale, lr, step, handle=None):
    st = lib().ipt_hip_mlp_adam_step(_vpp(d_params), _vpp(d_m), _vpp(d_v), _vpp(d_grad), float(gscale), float(lr),
                                     int(step), handle)
    if st != 0:
        raise IptError(st, "ipt_hip_mlp_adam_step")


def mlp_features(d_R, ldr, d_diag, d_rowabs, d_theta, d_idx, n, nt, d_feat, d_qf, handle=None):
    st = lib().ipt_hip_mlp_features(_vpp(d_R), int(ldr), _vpp(d_diag), _vpp(d_rowabs), _vpp(d_theta), _vpp(d_idx),
                                    int(n), int(nt), _vpp(d_feat), _vpp(d_qf), handle)
    if st != 0:
        raise IptError(st, "ipt_hip_mlp_features")


def save_mlp_model(path, params):
    """Model file read by the library (IPT_DAVIDSON_MLP_MODEL): the 8-byte
    tag IPTMLP01, the parameter count (int32), then FP32 parameters."""
    p = np.ascontiguousarray(np.asarray(params, dtype=np.float32))
    if p.size != MLP_NPARAM:
        raise ValueError("expected %d parameters" % MLP_NPARAM)
    with open(path, "wb") as f:
        f.write(b"IPTMLP01")
        np.array([MLP_NPARAM], dtype=np.int32).tofile(f)
        p.tofile(f)


def load_mlp_model(path):
    with open(path, "rb") as f:
        if f.read(8) != b"IPTMLP01":
            raise ValueError("not an IPT MLP model file")
        cnt = int(np.fromfile(f, dtype=np.int32, count=1)[0])
        return np.fromfile(f, dtype=np.float32, count=cnt)


# ------------------------------------------------------------------ FCI inputs (host only)

def _fci_to_dict(m, with_perm=True):
    n, nnz = m.n, m.nnz
    out = {
        "n": n, "nnz": nnz,
        "col_ptr": np.ctypeslib.as_array(m.col_ptr, shape=(n + 1,)).copy(),
        "row_ind": np.ctypeslib.as_array(m.row_ind, shape=(max(nnz, 1),))[:nnz].copy(),
        "values": np.ctypeslib.as_array(m.values, shape=(max(nnz, 1),))[:nnz].copy(),
        "diagonal": np.ctypeslib.as_array(m.diagonal, shape=(n,)).copy(),
        "matrix_inf_norm": m.matrix_inf_norm, "norb": m.norb, "nelec": m.nelec, "neleca": m.neleca,
        "nelecb": m.nelecb, "na": m.na, "nb": m.nb,
    }
    out["perm"] = np.ctypeslib.as_array(m.perm, shape=(n,)).copy() if (with_perm and m.perm) else None
    return out


def fci_build_csc(norb, neleca, nelecb, h1, eri, ecore, sort_by_diagonal=True):
    """Full-CI Hamiltonian (CSC) of the active space; see ipt_hip_fci.h."""
    h1 = np.ascontiguousarray(h1, dtype=np.float64).reshape(norb * norb)
    eri = np.ascontiguousarray(eri, dtype=np.float64).reshape(norb ** 4)
    m = abi.IPTFciCsc()
    st = lib().ipt_hip_fci_build_csc(int(norb), int(neleca), int(nelecb), _d(h1), _d(eri), float(ecore),
                                     1 if sort_by_diagonal else 0, ctypes.byref(m))
    if st != 0:
        raise IptError(st, "ipt_hip_fci_build_csc")
    try:
        return _fci_to_dict(m)
    finally:
        lib().ipt_hip_fci_free_csc(ctypes.byref(m))


def fcidump_read(path):
    """(norb, nelec, ms2, h1 (norb x norb), eri (norb^4 as 4-d), ecore)."""
    norb, nelec, ms2 = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    st = lib().ipt_hip_fcidump_read_header(os.fsencode(path), ctypes.byref(norb), ctypes.byref(nelec),
                                           ctypes.byref(ms2))
    if st != 0:
        raise IptError(st, "ipt_hip_fcidump_read_header")
    no = norb.value
    h1 = np.zeros(no * no)
    eri = np.zeros(no ** 4)
    ecore = ctypes.c_double(0.0)
    st = lib().ipt_hip_fcidump_read(os.fsencode(path), no, _d(h1), _d(eri), ctypes.byref(ecore))
    if st != 0:
        raise IptError(st, "ipt_hip_fcidump_read")
    return no, nelec.value, ms2.value, h1.reshape(no, no), eri.reshape(no, no, no, no), ecore.value


def fcidump_write(path, norb, nelec, ms2, h1, eri, ecore):
    h1 = np.ascontiguousarray(h1, dtype=np.float64).reshape(norb * norb)
    eri = np.ascontiguousarray(eri, dtype=np.float64).reshape(norb ** 4)
    st = lib().ipt_hip_fcidump_write(os.fsencode(path), int(norb), int(nelec), int(ms2), _d(h1), _d(eri),
                                     float(ecore))
    if st != 0:
        raise IptError(st, "ipt_hip_fcidump_write")


def csc_cache_write(path, csc):
    """Write a CSC dict (keys of fci_build_csc) in the reference's binary
    cache format."""
    cp = np.ascontiguousarray(csc["col_ptr"], dtype=np.int32)
    ri = np.ascontiguousarray(csc["row_ind"], dtype=np.int32)
    v = np.ascontiguousarray(csc["values"], dtype=np.float64)
    dg = np.ascontiguousarray(csc["diagonal"], dtype=np.float64)
    m = abi.IPTFciCsc()
    m.n, m.nnz = len(cp) - 1, len(ri)
    m.col_ptr, m.row_ind, m.values, m.diagonal = _i(cp), _i(ri), _d(v), _d(dg)
    m.perm = None
    m.matrix_inf_norm = float(csc["matrix_inf_norm"])
    for key in ("norb", "nelec", "neleca", "nelecb", "na", "nb"):
        setattr(m, key, int(csc.get(key, 0)))
    st = lib().ipt_hip_csc_cache_write(os.fsencode(path), ctypes.byref(m))
    if st != 0:
        raise IptError(st, "ipt_hip_csc_cache_write")


def csc_cache_read(path):
    m = abi.IPTFciCsc()
    st = lib().ipt_hip_csc_cache_read(os.fsencode(path), ctypes.byref(m))
    if st != 0:
        raise IptError(st, "ipt_hip_csc_cache_read")
    try:
        return _fci_to_dict(m, with_perm=False)
    finally:
        lib().ipt_hip_fci_free_csc(ctypes.byref(m))
