/*
 * ipt_hip_mixed.h -- two-phase mixed-precision entry points of
 * libipt_hip.so, drop-in for the reference's
 * IPT/src/mixed_precision/ipt_mixed_precision_cuda.cu:19-78 and
 * IPT/src/mixed_precision/ipt_rayleigh_mixed_cuda.cu:19-41 (reference
 * checkout, IPT/ = IPT_C_GPU/).
 *
 * Each entry runs plain IPT with the off-diagonal product E Z in low
 * precision on MFMA (E = H - diag(H); FP32 accumulation) to the low
 * tolerance, then restarts plain FP64 IPT from that iterate to the FP64
 * tolerance.  Precision mapping on gfx950: "fp16" = FP16 inputs, "fp32" =
 * FP32 inputs (MFMA 16x16x4 FP32), "tf32" = BF16x3 (three BF16 products of
 * split operands; gfx950 has no TF32).  The iterate itself is kept in FP64
 * between products.
 *
 * The cublasHandle_t argument becomes an opaque handle (NULL: the
 * library's context on the current device; otherwise ipt_hip_create_handle).
 * Result structs keep the reference's field order; d_vectors (n x k,
 * column-major) and d_values (k) are device allocations released by the
 * matching *_free_device_result.  Status codes are those of ipt_hip.h.
 */
#ifndef IPT_HIP_MIXED_H
#define IPT_HIP_MIXED_H

#ifdef __cplusplus
extern "C" {
#endif

/* IPT/src/mixed_precision/ipt_mixed_precision_cuda.cu:19-31 */
typedef struct {
    int n;
    int k;
    int fp16_iterations;
    int fp64_iterations;
    double fp16_time_sec;
    double fp64_time_sec;
    double total_time_sec;
    double fp16_fixed_point_residual;
    double fp64_fixed_point_residual;
    double *d_vectors;
    double *d_values;
} IPTMixedCudaDeviceResult;

/* IPT/src/mixed_precision/ipt_mixed_precision_cuda.cu:33-45 */
typedef struct {
    int n;
    int k;
    int fp32_iterations;
    int fp64_iterations;
    double fp32_time_sec;
    double fp64_time_sec;
    double total_time_sec;
    double fp32_fixed_point_residual;
    double fp64_fixed_point_residual;
    double *d_vectors;
    double *d_values;
} IPTMixedFP32CudaDeviceResult;

/* IPT/src/mixed_precision/ipt_mixed_precision_cuda.cu:47-59 */
typedef struct {
    int n;
    int k;
    int tf32_iterations;
    int fp64_iterations;
    double tf32_time_sec;
    double fp64_time_sec;
    double total_time_sec;
    double tf32_fixed_point_residual;
    double fp64_fixed_point_residual;
    double *d_vectors;
    double *d_values;
} IPTMixedTF32CudaDeviceResult;

/* IPT/src/mixed_precision/ipt_rayleigh_mixed_cuda.cu:19-34 */
typedef struct {
    int n;
    int k;
    int tf32_iterations;
    int fp64_iterations;
    double tf32_time_sec;
    double orthogonalize_time_sec;
    double rayleigh_transform_time_sec;
    double fp64_time_sec;
    double backtransform_time_sec;
    double total_time_sec;
    double tf32_fixed_point_residual;
    double fp64_fixed_point_residual;
    double *d_vectors;
    double *d_values;
} IPTRayleighMixedCudaDeviceResult;

/* ipt_mixed_precision_cuda.cu:61-65, 1005-1080: FP16 phase + FP64 phase. */
void ipt_mixed_precision_cuda_free_device_result(IPTMixedCudaDeviceResult *result);
int ipt_mixed_precision_cuda_device_tol(const double *d_matrix_col_major, int n, int k, double fp16_tol,
                                        int fp16_maxiter, double fp64_tol, int fp64_maxiter, void *handle,
                                        IPTMixedCudaDeviceResult *result);

/* ipt_mixed_precision_cuda.cu:66-70, 1082-1155: FP32 phase + FP64 phase. */
void ipt_mixed_precision_fp32_cuda_free_device_result(IPTMixedFP32CudaDeviceResult *result);
int ipt_mixed_precision_fp32_cuda_device_tol(const double *d_matrix_col_major, int n, int k, double fp32_tol,
                                             int fp32_maxiter, double fp64_tol, int fp64_maxiter, void *handle,
                                             IPTMixedFP32CudaDeviceResult *result);

/* ipt_mixed_precision_cuda.cu:71-75, 1157-1222: "TF32" (BF16x3 here) phase
 * + FP64 phase. */
void ipt_mixed_precision_tf32_cuda_free_device_result(IPTMixedTF32CudaDeviceResult *result);
int ipt_mixed_precision_tf32_cuda_device_tol(const double *d_matrix_col_major, int n, int k, double tf32_tol,
                                             int tf32_maxiter, double fp64_tol, int fp64_maxiter, void *handle,
                                             IPTMixedTF32CudaDeviceResult *result);

/* ipt_rayleigh_mixed_cuda.cu:36-41, 274-400: "TF32" (BF16x3) IPT, Householder
 * QR of the iterate, B = Q^T H Q (symmetrised), FP64 IPT on B, vectors Q Z. */
void ipt_rayleigh_mixed_cuda_free_device_result(IPTRayleighMixedCudaDeviceResult *result);
int ipt_rayleigh_mixed_tf32_cuda_device_tol(const double *d_matrix_col_major, int n, int k, double tf32_tol,
                                            int tf32_maxiter, double fp64_tol, int fp64_maxiter, void *handle,
                                            IPTRayleighMixedCudaDeviceResult *result);

#ifdef __cplusplus
}
#endif

#endif /* IPT_HIP_MIXED_H */
