"""Independent device-side checks for the GPU tests and the benchmark:
rocBLAS (the vendor library, not this repository's kernels) through ctypes,
on device pointers of the same HIP runtime the library uses.  Used where a
matrix is too large to check on the host (dense N=65536).  Not part of the
solver."""
import ctypes

import numpy as np

from . import api

_OP_N, _OP_T = 111, 112


class RocBlas:
    def __init__(self):
        api.hip()  # load the library's HIP runtime first
        self.lib = ctypes.CDLL("librocblas.so.5")
        self.h = ctypes.c_void_p()
        if self.lib.rocblas_create_handle(ctypes.byref(self.h)) != 0:
            raise RuntimeError("rocblas_create_handle failed")
        self.lib.rocblas_dgemm.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p,
                                           ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_double), ctypes.c_void_p, ctypes.c_int]
        self.lib.rocblas_dasum.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_double)]

    def close(self):
        self.lib.rocblas_destroy_handle(self.h)

    def gemm_nn(self, m, n, k, dA, lda, dB, ldb, dC, ldc):
        """C = A B, column-major, device pointers."""
        one, zero = ctypes.c_double(1.0), ctypes.c_double(0.0)
        st = self.lib.rocblas_dgemm(self.h, _OP_N, _OP_N, m, n, k, ctypes.byref(one), ctypes.c_void_p(dA), lda,
                                    ctypes.c_void_p(dB), ldb, ctypes.byref(zero), ctypes.c_void_p(dC), ldc)
        if st != 0:
            raise RuntimeError("rocblas_dgemm status %d" % st)

    def inf_norm_symmetric(self, dH, n):
        """max column sum of |H| (= the row sums for a symmetric H)."""
        out = ctypes.c_double(0.0)
        best = 0.0
        for c in range(n):
            if self.lib.rocblas_dasum(self.h, n, ctypes.c_void_p(dH + 8 * c * n), 1, ctypes.byref(out)) != 0:
                raise RuntimeError("rocblas_dasum failed")
            best = max(best, out.value)
        return best


def scaled_residuals_device(dH, n, dX, dlam, k):
    """The reference's scaled residuals ||H x - theta x|| / (max(1, ||H||_inf)
    max(1, ||x||)) of k device vectors (column-major n x k), H x by rocBLAS;
    returns (residuals, values, ||X^T X - I||_max)."""
    hip = api.hip()
    rb = RocBlas()
    dY = hip.malloc(n * k * 8)
    try:
        rb.gemm_nn(n, k, n, dH, n, dX, n, dY, n)
        hip.synchronize()
        normH = rb.inf_norm_symmetric(dH, n)
        Y = hip.to_host(dY, (n, k))
    finally:
        hip.free(dY)
        rb.close()
    X = hip.to_host(dX, (n, k))
    lam = hip.to_host(dlam, (k,))
    R = Y - X * lam
    res = np.linalg.norm(R, axis=0) / (max(1.0, normH) * np.maximum(1.0, np.linalg.norm(X, axis=0)))
    orth = np.abs(X.T @ X - np.eye(k)).max()
    return res, lam, orth
