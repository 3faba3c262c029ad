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

/* Near-degenerate variant of a dense matrix with sorted diagonal (the
 * dense N=65536, k=256 workload's variant: positions 10, 50, 120, 200, gap
 * 1e-4): H(i+1,i+1) = H(i,i) + gap for each position i.  Status 1 when a
 * position is out of range or the new entry would not stay below
 * H(i+2,i+2) (nothing is written then).  No reference counterpart. */
int ipt_hip_set_diagonal_gaps(double *d_H, int n, const int *positions, int count, double gap, void *handle);

/* The reference's epsilon-sweep matrix on the device (column-major n x n):
 * H_ii = i + 1 + eps U(seed, n, i, i), H_ij = H_ji = eps (U(seed,n,i,j) +
 * U(seed,n,j,i)) / 2, U as above; the generator of
 * IPT/test/ACX/benchmark_acx_cuda_sweep.cu:284-312 (seed there: 20260615 +
 * n + a per-epsilon offset) and benchmark_tf32_cusolver_sweep.cu:472-503. */
int ipt_hip_generate_eps(double *d_H, int n, double eps, unsigned long long seed, void *handle);

/* Sparse synthetic workload (orbital-sum model), CSR = CSC (symmetric) with
 * 32-bit indices and ascending column indices per row:
 *   norb orbital energies U[elo, ehi] (sorted); row i's diagonal is the sum
 *   of a random distinct nsub-subset, the n diagonals sorted ascending; row
 *   i keeps the first m distinct columns of its candidate stream
 *   floor(U (n-1)) (shifted past i), each with the value sigma * N(0,1), and
 *   H_ji = H_ij (mirror); when i drew j and j drew i, the larger row's draw
 *   sets both entries.  U as in ipt_hip_generate_dense, one tagged stream per
 *   purpose.  The caller provides row_ptr (n + 1 ints) and col_ind / values
 *   of `capacity` entries (n (2m + 1) always suffices); nnz_out receives the
 *   entry count.  m <= 512 and m <= n - 1. */
int ipt_hip_generate_orbital_csr(int n, int m, unsigned long long seed, int norb, int nsub, double elo, double ehi,
                                 double sigma, int *d_row_ptr, int *d_col_ind, double *d_values, long long capacity,
                                 long long *nnz_out, void *handle);

/* Y = A X for a CSR matrix A (n x n, 32-bit indices) and row-major X, Y
 * (n x k, leading dimensions ldx, ldy >= k): the sparse product of the block
 * path (the reference calls cusparseSpMM, IPT/src/ipt_cuda.cu:1535-1559). */
int ipt_hip_csr_spmm(const int *d_row_ptr, const int *d_col_ind, const double *d_values, int n, const double *d_X,
                     int ldx, double *d_Y, int ldy, int k, void *handle);

/* The same product for a CSR whose column indices ascend within each row,
 * run as column-sliced passes: pass b multiplies the entries with columns
 * in one slice of [0, n) and adds to Y, so its gathers stay within that
 * slice of X.  slices = 0: the pass count the block path uses for width k
 * (1 for k > 32 or X of at most 2 MiB; k <= 8: 4 passes up to 16 MiB of X
 * and 8 beyond; 9 <= k <= 32: 1 pass up to 16 MiB of X, beyond it 8
 * passes, 16 from k = 17); 1, 2, 4, 8, 16: that many passes.  The block
 * path also lays X out at the row stride with the fewest 128-byte lines
 * (ldx from next_pow2(k) or k rounded up to 16 where that touches fewer
 * lines). */
int ipt_hip_csr_spmm_sliced(const int *d_row_ptr, const int *d_col_ind, const double *d_values, int n,
                            const double *d_X, int ldx, double *d_Y, int ldy, int k, int slices, void *handle);

/* The row stride (in elements of elem = 8 or 4 bytes) the block path gives
 * a row-major X of w columns before a CSR product: among w, next_pow2(w)
 * and w rounded up to a 128-byte line, the smallest with the fewest
 * 128-byte lines per row on average over the row offsets.  Host only (no
 * reference counterpart; exported for the layout's tests); -1 for w <= 0
 * or another elem. */
long ipt_hip_csr_row_stride(int w, int elem);

/* The same product through the LDS-tiled layout (addition, no reference
 * counterpart; the block path takes it for its narrow products on large
 * matrices): the entries regrouped once into buckets of 4096 rows x 4096
 * columns, one workgroup per row group staging 64 KiB of X (8192 rows of
 * one column or 4096 rows of two) in LDS, so every gather is an LDS read
 * and the matrix is streamed once per two columns.  Builds the layout on
 * each call (the block path builds it once per solve); n <= 2^25, any
 * column order within a row, IPT_CUDA_INVALID_ARGUMENT for a column index
 * outside [0, n); the sums agree with the row-major product to rounding
 * (LDS atomic adds). */
int ipt_hip_csr_spmm_tiled(const int *d_row_ptr, const int *d_col_ind, const double *d_values, int n,
                           const double *d_X, int ldx, double *d_Y, int ldy, int k, void *handle);

/* In place: Q (n x k, column-major, leading dimension n, k <= n) <- the
 * orthonormal factor of its QR factorisation (positive diagonal R), by
 * shifted Cholesky QR 3 in FP64 on the device: columns normalised, one
 * Cholesky QR pass on the shifted Gram, then two plain passes.  Replaces the
 * reference's FP32-preconditioned Cholesky QR
 * (IPT/mixed/ipt_rayleigh_mixed_mpqr_cuda.cu:169-386: FP32 Householder R,
 * V = Q R^-1, one FP64 Cholesky QR), which gives the same factor.  Status 1
 * for a zero or non-finite column or a numerically dependent basis.
 * ipt_rayleigh_mixed_tf32_cuda_device_tol uses it with IPT_RAYLEIGH_QR=mpqr. */
int ipt_hip_cholqr_mixed(double *d_Q, int n, int k, void *handle);

/* Y = Ya + diag(d)(B - Ba) + E (B - Ba), E = H - diag(H) rounded to
 * precision `mode` (0 fp16, 1 bf16, 2 bf16x3, 3 fp32, 4 fp8 E4M3 with one
 * scale per 128 x 128 tile of E), FP32-accumulated MFMA.  All device,
 * column-major n x nc (leading dimension n); H n x n. */
int ipt_hip_lowp_product(const double *d_H, int n, int mode, const double *d_B, const double *d_Ba,
                         const double *d_Ya, const double *d_diag, int nc, double *d_Y, void *handle);

/* Y = H Z in FP64 on the MFMA kernel of the block path: packed = 0 reads
 * column-major H directly, packed = 1 first copies H into MFMA operand
 * order (the layout the all-FP64 dense block solve uses for n >= 4096).
 * All device, column-major: H n x n, Z and Y n x nc (leading dimension n). */
int ipt_hip_dense_product(const double *d_H, int n, const double *d_Z, int nc, double *d_Y, int packed,
                          void *handle);

/* Kernel timing with HIP events on the library stream.  enable(1) starts
 * recording; get() waits for the recorded events and returns the summed
 * duration (ms) and launch count of one kernel family: "gemm_f64",
 * "lowp_gemm", "block_update", "spmm" (FP64 CSR products), "spmm_f32"
 * (the FP32 correction products of the sparse mixed path). */
void ipt_hip_profile_enable(int on);
/* Times only the listed families (comma-separated names as above), so the
 * event records of the others stay out of the stream; "" stops timing. */
int ipt_hip_profile_enable_only(const char *kernels);
void ipt_hip_profile_reset(void);
int ipt_hip_profile_get(const char *kernel, double *total_ms, long long *launches);
/* Columns processed by the timed launches of one family, summed (the
 * algorithmic bytes of a product launch depend on its width). */
int ipt_hip_profile_get_columns(const char *kernel, long long *columns);

/* Statistics of the last block-cluster solve on `handle` (NULL: library
 * context of the current device): basis columns, iteration steps, number of
 * FP64 and low-precision block products, Davidson steps accepted. */
int ipt_hip_last_solve_stats(void *handle, int *basis_cols, int *steps, long long *fp64_products,
                             long long *lowp_products, int *davidson_steps);

/* Davidson refinement of the last block-cluster solve on `handle`: steps
 * accepted and rejected, steps whose learned (MLP) correction was kept and
 * those redone with the diagonal correction, H products of single columns,
 * and the largest scaled residual at the end (the reference reports the
 * same figures in IPTBlockClusterResult, IPT/src/ipt_cuda.cu:3318-3982). */
int ipt_hip_last_davidson_stats(void *handle, int *accepted, int *rejected, int *mlp_accepted, int *mlp_rejected,
                                long long *matvecs, double *residual_after);

/* Eigenvalues (ascending, d_w: m) and orthonormal eigenvectors (d_V, m x m,
 * column-major, leading dimension ldv) of the symmetric d_A (m x m, lda; read
 * only), on the device by two-sided block Jacobi (32 x 32 subproblems, one
 * cyclic sweep each per round in LDS, block rotations applied by a tiled
 * kernel).  The
 * projected problems of Rayleigh-Ritz and Davidson use it in place of the
 * reference's cusolverDnDsygvd / Dsyevd (IPT/src/ipt_cuda.cu:3213-3226).
 * Status 6 when it does not converge in 40 sweeps. */
int ipt_hip_sym_eig(const double *d_A, int lda, int m, double *d_w, double *d_V, int ldv, void *handle);

/* Cluster ranges of the last block-cluster solve on `handle` (NULL: the
 * library context of the current device): *count receives the number of
 * clusters, and the first min(count, capacity) of them are written as
 * (first, last) row pairs in the solver's sorted-diagonal order, singletons
 * included, in basis-column order.  The reference computes the same ranges
 * in ipt_target_degenerate_subspaces (IPT/src/ipt_preparation.cu:501-645)
 * and ipt_block_cluster_prepare (IPT/src/ipt_cuda.cu:4551-4683) but does not
 * expose them; callers: parity tests and tools that inspect a solve. */
int ipt_hip_last_cluster_ranges(void *handle, int *first_last, int capacity, int *count);
/* State of each cluster of the last block-cluster solve, in the order of
 * ipt_hip_last_cluster_ranges: 0 converged to the block tolerance, 1 still
 * iterating at the iteration limit, 2 stopped by stall detection
 * (IPT_BLOCK_CLUSTER_STALL); no reference counterpart (the reference prints
 * per-cluster iteration lines).  Callers: tests and tools. */
int ipt_hip_last_cluster_states(void *handle, int *states, int capacity, int *count);

/* Learned Davidson preconditioner (per-component MLP 3 -> 64 -> 64 -> 1,
 * GELU; 4481 FP32 parameters: W1[64][3], b1[64], W2[64][64], b2[64],
 * W3[64], b3).  Correction t_i = (r_i / s_i) * y_i with y = MLP(x_i),
 * s_i = sum_{j != i} |H_ij|, x_i = (asinh((D_i - theta)/s_i),
 * r_i / max|r|, s_i / (|D_i - theta| + s_i)).  Enabled in the block path by
 * IPT_DAVIDSON_PRECOND=mlp with IPT_DAVIDSON_MLP_MODEL=<file>; the diagonal
 * correction stays the fallback (a step whose largest residual does not
 * decrease is redone with it).  These entry points expose the kernels for
 * training and testing. */

/* Features (3 floats per (target q, component i), at (q n + i) * 3) and
 * qf = r_i / s_i for the columns idx[0..nt) of the residual block R
 * (n x ., column-major, leading dimension ldr) and Ritz values theta. */
int ipt_hip_mlp_features(const double *d_R, int ldr, const double *d_diag, const double *d_rowabs,
                         const double *d_theta, const int *d_idx, int n, int nt, float *d_feat, float *d_qf,
                         void *handle);

/* out[c] = qf[c] * MLP(feat[3c..3c+3)) (qf NULL: the bare MLP output);
 * BF16 weights of the hidden layer, FP32 accumulation. */
int ipt_hip_mlp_infer(const float *d_params, const float *d_feat, const float *d_qf, long long count, float *d_out,
                      void *handle);

/* Loss sum_c w_c (q_c MLP(x_c) - t_c)^2 and its gradient (4481 floats) in
 * FP32 over `count` samples (feat 3 per sample, q, t, w one each). */
int ipt_hip_mlp_gradient(const float *d_params, const float *d_feat, const float *d_q, const float *d_t,
                         const float *d_w, long long count, float *d_grad, float *d_loss, void *handle);

/* ipt_hip_mlp_gradient without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight training or serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results (bit for bit) as
 * ipt_hip_mlp_gradient.  Its partial gradients live in a fixed device array
 * of 1024 x 4482 floats (18.4 MB per GPU, kept until the process ends),
 * which covers every count; the entry point calls no hipMalloc/hipFree and
 * creates no library context (handle NULL: the null stream).  Every call on
 * a GPU shares that GPU's array: calls that overlap on one GPU race. */
int ipt_hip_mlp_gradient_noalloc(const float *d_params, const float *d_feat, const float *d_q, const float *d_t,
                                 const float *d_w, long long count, float *d_grad, float *d_loss, void *handle);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with gradient
 * grad * gscale; m, v are the moment buffers (4481 floats each, zero at
 * step 1). */
int ipt_hip_mlp_adam_step(float *d_params, float *d_m, float *d_v, const float *d_grad, float gscale, float lr,
                          int step, void *handle);

/* The host preparation of the plain CSC entry points (no IPT_BLOCK_CLUSTER),
 * IPT/src/ipt_preparation.cu:1375-1648: stable sort of the diagonal and
 * lifting of degenerate target subspaces (symmetric input) by local
 * eigenvector similarities.  Writes the prepared matrix densely (n x n
 * column-major, may be NULL), perm_out[sorted] = original (may be NULL) and
 * the number of rotations.  Status 1 where the reference's preparation
 * fails (a subspace wider than IPT_MAX_LIFT_BLOCK_SIZE, a non-symmetric
 * matrix that needs a lift, targets still joined after the last round).
 * Host only; no device work. */
int ipt_hip_prepare_sparse_csc(const int *col_ptr, const int *row_ind, const double *values, int n, int k, int nnz,
                               double degeneracy_threshold, double *dense_out, int *perm_out, int *rotations_out);
/* vectors (n x kv, column-major) from the prepared basis of the same call to
 * the caller's: the rotations in reverse order, then the inverse
 * permutation, IPT/src/ipt_preparation.cu:1222-1294.  Host only. */
int ipt_hip_prepared_vectors_to_original(const int *col_ptr, const int *row_ind, const double *values, int n, int k,
                                         int nnz, double degeneracy_threshold, double *vectors, int kv);

#ifdef __cplusplus
}
#endif

#endif /* IPT_HIP_EXT_H */
