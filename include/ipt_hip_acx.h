/*
 * ipt_hip_acx.h -- ACX-accelerated plain IPT (alternating cyclic
 * extrapolation of the IPT fixed-point map) in libipt_hip.so, drop-in for
 * the reference's IPT/src/ACX/acx_cuda.cu:11-33, 43-81, 390-620 (reference
 * checkout, IPT/ = IPT_C_GPU/).
 *
 * Iteration t (t = 1, 2, ...) starts from x (x = I(:, 0:k) at t = 1):
 *   f1 = F(x); stop when ||f1 - x||_F / ||x||_F < min(tol, 1e-12)
 *   (the returned vectors are then f1);
 *   odd t  (order 2): f2 = F(f1), d1 = f1 - x, d2 = f2 - 2 f1 + x,
 *          sigma_j = |<d2_j, d1_j>| / <d2_j, d2_j>  (per column j),
 *          x += 2 sigma d1 + sigma^2 d2;
 *   even t (order 3): f2 = F(f1), f3 = F(f2), d3 = f3 - 3 f2 + 3 f1 - x,
 *          sigma_j = |<d3_j, d2_j>| / <d3_j, d3_j>,
 *          x += 3 sigma d1 + 3 sigma^2 d2 + sigma^3 d3;
 * with F the plain IPT map.  Values are diag(H f1).  A run that reaches
 * maxiter without meeting the test returns success with converged = 0.
 *
 * The cublasHandle_t argument becomes an opaque handle (NULL: the library's
 * context on the current device; otherwise ipt_hip_create_handle of
 * ipt_hip.h).  d_vectors (n x k column-major) and d_values (k) are device
 * allocations released by acx_cuda_free_result.
 */
#ifndef IPT_HIP_ACX_H
#define IPT_HIP_ACX_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes, IPT/src/ACX/acx_cuda.cu:11-17 */
#define ACX_CUDA_SUCCESS 0
#define ACX_CUDA_INVALID_ARGUMENT 1
#define ACX_CUDA_ALLOCATION_FAILED 2
#define ACX_CUDA_CUDA_ERROR 3
#define ACX_CUDA_CUBLAS_ERROR 4
#define ACX_CUDA_ZERO_DENOMINATOR 5
#define ACX_FIXED_POINT_RTOL 1.0e-12

/* IPT/src/ACX/acx_cuda.cu:20-33 */
typedef struct {
    int n;
    int k;
    int iterations;
    int f_calls;
    int converged;
    double time_sec;
    double iteration_time_sec;
    double finalization_time_sec;
    double max_residual;        /* max_j ||(H x)_j - x_j (H x)_jj|| of the last iteration's x */
    double fixed_point_residual;
    double *d_vectors;
    double *d_values;
} ACXCudaResult;

/* IPT/src/ACX/acx_cuda.cu:43-61 */
const char *acx_cuda_status_string(int status);
/* IPT/src/ACX/acx_cuda.cu:63-71: zero fields, NaN residuals */
void acx_cuda_reset_result(ACXCudaResult *result);
/* IPT/src/ACX/acx_cuda.cu:73-81 */
void acx_cuda_free_result(ACXCudaResult *result);
/* IPT/src/ACX/acx_cuda.cu:390-620; status 5 when two diagonal entries of a
 * target column and a row differ by at most DBL_EPSILON */
int acx_cuda_device_tol(const double *d_matrix_col_major, int n, int k, double tol, int maxiter, void *handle,
                        ACXCudaResult *result);

/* Two-phase ACX (IPT/src/ACX/acx_mixed_precision_cuda.cu:19-52, 811-897):
 * ACX with the off-diagonal product in low precision ("tf32" = BF16x3 on
 * gfx950, which has no TF32; the iterate is kept in FP64) until
 * ||f1 - x|| / ||x|| <= tf32_tol, then FP64 ACX restarted from that iterate
 * (f1 when converged, else the last extrapolated x) to min(fp64_tol, 1e-12);
 * the returned vectors are the FP64 phase's f1 when it converged, else its
 * last x, and the values diag(H vectors). */
typedef struct {
    int n;
    int k;
    int tf32_iterations;
    int fp64_iterations;
    int tf32_f_calls;
    int fp64_f_calls;
    int tf32_converged;
    int fp64_converged;
    double tf32_time_sec;
    double fp64_time_sec;
    double total_time_sec;
    double finalization_time_sec;
    double tf32_fixed_point_residual;
    double fp64_fixed_point_residual;
    double fp64_max_residual;
    double *d_vectors;
    double *d_values;
} ACXMixedTF32CudaResult;

void acx_mixed_precision_tf32_cuda_reset_result(ACXMixedTF32CudaResult *result);
void acx_mixed_precision_tf32_cuda_free_result(ACXMixedTF32CudaResult *result);
int acx_mixed_precision_tf32_cuda_device_tol(const double *d_matrix_col_major, int n, int k, double tf32_tol,
                                             int tf32_maxiter, double fp64_tol, int fp64_maxiter, void *handle,
                                             ACXMixedTF32CudaResult *result);

/* Rayleigh two-phase ACX (IPT/src/ACX/acx_rayleigh_mixed_precision_cuda.cu:
 * 19-52, 289-434): low-precision ACX as above, Householder QR of its
 * iterate Q, FP64 ACX on the k x k Rayleigh quotient Q^T H Q (symmetrised),
 * vectors Q Z. */
typedef struct {
    int n;
    int k;
    int tf32_iterations;
    int fp64_iterations;
    int tf32_f_calls;
    int fp64_f_calls;
    int tf32_converged;
    int fp64_converged;
    double tf32_time_sec;
    double orthogonalize_time_sec;
    double rayleigh_transform_time_sec;
    double fp64_time_sec;
    double finalization_time_sec;
    double backtransform_time_sec;
    double total_time_sec;
    double tf32_fixed_point_residual;
    double fp64_fixed_point_residual;
    double fp64_max_residual;
    double *d_vectors;
    double *d_values;
} ACXRayleighMixedTF32CudaResult;

void acx_rayleigh_mixed_tf32_cuda_reset_result(ACXRayleighMixedTF32CudaResult *result);
void acx_rayleigh_mixed_tf32_cuda_free_result(ACXRayleighMixedTF32CudaResult *result);
int acx_rayleigh_mixed_tf32_cuda_device_tol(const double *d_matrix_col_major, int n, int k, double tf32_tol,
                                            int tf32_maxiter, double fp64_tol, int fp64_maxiter, void *handle,
                                            ACXRayleighMixedTF32CudaResult *result);

#ifdef __cplusplus
}
#endif

#endif
