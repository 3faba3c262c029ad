/*
 * ipt_hip_ext.h -- additions to the drop-in ABI of libipt_hip.so that have
 * no counterpart in the reference: synthetic workload generators, the
 * anchored low-precision product used by the mixed-precision block path,
 * and kernel timing instrumentation.  Callers: benchmarks, tests and tools
 * that measure the library.
 */
#ifndef IPT_HIP_EXT_H
#define IPT_HIP_EXT_H

#ifdef __cplusplus
extern "C" {
#endif

/* Dense symmetric test matrix on the device, column-major n x n:
 * H = diag(d) + V with d = n draws of lo + (hi-lo) U(seed, n, i, n) sorted
 * ascending, V_ij = V_ji = vscale/sqrt(n) * z_ij (i != j), z_ij standard
 * normal by Box-Muller from the splitmix64 streams U(seed,n,min,max) and
 * U(seed,n,max,min), U as in
 * IPT/test/mixed_precision_tf32/benchmark_tf32_cusolver_sweep.cu:442-449. */
int ipt_hip_generate_dense(double *d_H, int n, unsigned long long seed, double lo, double hi, double vscale,
                           void *handle);

/* Y = Ya + diag(d)(B - Ba) + E (B - Ba), E = H - diag(H) rounded to
 * precision `mode` (0 fp16, 1 bf16, 2 bf16x3, 3 fp32, 4 fp8 E4M3 with one
 * scale per 128 x 128 tile of E), FP32-accumulated MFMA.  All device,
 * column-major n x nc (leading dimension n); H n x n. */
int ipt_hip_lowp_product(const double *d_H, int n, int mode, const double *d_B, const double *d_Ba,
                         const double *d_Ya, const double *d_diag, int nc, double *d_Y, void *handle);

/* Kernel timing with HIP events on the library stream.  enable(1) starts
 * recording; get() waits for the recorded events and returns the summed
 * duration (ms) and launch count of one kernel family: "gemm_f64",
 * "lowp_gemm", "block_update", "spmm". */
void ipt_hip_profile_enable(int on);
void ipt_hip_profile_reset(void);
int ipt_hip_profile_get(const char *kernel, double *total_ms, long long *launches);

/* Statistics of the last block-cluster solve on `handle` (NULL: library
 * context of the current device): basis columns, iteration steps, number of
 * FP64 and low-precision block products, Davidson steps accepted. */
int ipt_hip_last_solve_stats(void *handle, int *basis_cols, int *steps, long long *fp64_products,
                             long long *lowp_products, int *davidson_steps);

#ifdef __cplusplus
}
#endif

#endif /* IPT_HIP_EXT_H */
