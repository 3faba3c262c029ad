/*
 * ipt_hip.h -- C ABI of the MI355X block-IPT eigensolver (libipt_hip.so).
 *
 * Drop-in for the reference's CUDA entry points: every symbol, argument
 * order, status code and the IPTCudaResult layout are those of
 * IPT/src/ipt_cuda.cu:75-314 (reference checkout, IPT/ = IPT_C_GPU/).
 * The single signature change: the cublasHandle_t argument of the *_device
 * entry points is an opaque `void *handle` -- NULL selects the library's
 * own per-device context, otherwise a handle from ipt_hip_create_handle().
 *
 * Solver options are the reference's IPT_* environment variables, read at
 * call time, plus the additions documented in each declaration below.
 *
 * Threading: calls are synchronous and may run concurrently only on
 * distinct handles.
 */
#ifndef IPT_HIP_H
#define IPT_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes, IPT/src/ipt_cuda.cu:82-90.  Code 3 reports a HIP runtime
 * error, 4 a device BLAS error, 5 a sparse-kernel error, 6 a small dense
 * solver (Cholesky / symmetric eigen) failure. */
enum {
    IPT_CUDA_SUCCESS = 0,
    IPT_CUDA_INVALID_ARGUMENT = 1,
    IPT_CUDA_ALLOCATION_FAILED = 2,
    IPT_CUDA_CUDA_ERROR = 3,
    IPT_CUDA_CUBLAS_ERROR = 4,
    IPT_CUDA_CUSPARSE_ERROR = 5,
    IPT_CUDA_CUSOLVER_ERROR = 6
};

/* Per-step Davidson history records, IPT/src/ipt_cuda.cu:92-194 (layout
 * kept so result structs are binary compatible; the pair-28/29 diagnostic
 * fields are left NaN / zero by this library). */
typedef struct {
    int davidson_step;
    int active_pair_index;
    double residual_before;
    double residual_after;
    int accepted;
    int basis_cols;
    double max_relative_eigen_residual;
    int max_relative_eigen_residual_index;
    double pair_28_residual;
    double pair_29_residual;
    double orthogonality_max_abs_error;
    int restarted;
    int best_so_far_updated;
    int best_so_far_step;
    double best_so_far_max_residual;
    double pair28_best_residual;
    double pair29_best_residual;
    int accept_global_ok;
    int accept_active_ok;
    int accept_locked_safe;
    char accept_reason[64];
    int accepted_steps;
    int rejected_steps;
    int early_jump_to_continuation;
    double gap28_29;
    double relative_gap28_29;
    double cluster_residual_fro_before;
    double cluster_residual_fro_after;
    double cluster_residual_max_before;
    double cluster_residual_max_after;
    double ritz_overlap_28_28;
    double ritz_overlap_28_29;
    double ritz_overlap_29_28;
    double ritz_overlap_29_29;
    int ritz_overlap_swap_detected;
    char pair28_lock_state[16];
    char pair29_lock_state[16];
    int cluster_hard_locked;
    int cluster_soft_locked_count;
} IPTDavidsonHistoryEntry;

typedef struct {
    int davidson_step;
    int pair_index;
    double residual;
    int selected_by_residual;
    int selected_forced;
    int selected_auto_cluster;
    int skipped_converged_in_cluster;
    int skipped_linear_dependent;
    int skipped_locked_old_logic_should_not_happen;
    double correction_norm_before_ortho;
    double correction_norm_after_ortho;
} IPTDavidsonSelectionEntry;

typedef struct {
    int davidson_step;
    char active_pairs[64];
    int accepted_corrections;
    int rejected_corrections;
    char correction_norm_before_ortho[256];
    char correction_norm_after_ortho[256];
    double residual_before_global;
    double residual_after_global;
    double pair_28_before;
    double pair_28_after;
    double pair_29_before;
    double pair_29_after;
    int accepted;
    char reject_reason[64];
    int basis_cols_before;
    int basis_cols_after;
    double orthogonality_error_before;
    double orthogonality_error_after;
    int best_so_far_updated;
    int best_so_far_step;
    double best_so_far_max_residual;
    double pair28_best_residual;
    double pair29_best_residual;
    int accept_global_ok;
    int accept_active_ok;
    int accept_locked_safe;
    char accept_reason[64];
    int accepted_steps;
    int rejected_steps;
    int early_jump_to_continuation;
    double gap28_29;
    double relative_gap28_29;
    double cluster_residual_fro_before;
    double cluster_residual_fro_after;
    double cluster_residual_max_before;
    double cluster_residual_max_after;
    double ritz_overlap_28_28;
    double ritz_overlap_28_29;
    double ritz_overlap_29_28;
    double ritz_overlap_29_29;
    int ritz_overlap_swap_detected;
    char pair28_lock_state[16];
    char pair29_lock_state[16];
    int cluster_hard_locked;
    int cluster_soft_locked_count;
} IPTDavidsonBlockHistoryEntry;

/* Host result, IPT/src/ipt_cuda.cu:196-260 (identical layout). */
typedef struct {
    int n;
    int k;
    int iterations;
    double *vectors; /* column-major n x k */
    double *values;  /* length k */
    int basis_cols;
    long long matvecs;
    double preparation_time_sec;
    double transfer_setup_time_sec;
    double iteration_time_sec;
    double rayleigh_ritz_time_sec;
    double davidson_time_sec;
    double solve_time_sec;
    double fixed_point_residual;
    int oversample;
    int rayleigh_ritz_used_qr;
    double basis_orthogonality_frobenius_error;
    double basis_orthogonality_max_abs_error;
    double ritz_vectors_orthogonality_frobenius_error;
    double ritz_vectors_orthogonality_max_abs_error;
    int basis_has_nan_or_inf;
    int ritz_vectors_has_nan_or_inf;
    int davidson_attempted;
    int davidson_accepted;
    int davidson_target_index;
    double davidson_residual_before;
    double davidson_residual_after;
    double davidson_denom_clip;
    int adaptive_block_enabled;
    double adaptive_coupling_tau;
    int adaptive_limit_hit;
    int adaptive_target_block_start;
    int adaptive_target_block_end;
    char adaptive_added_indices[256];
    int davidson_history_count;
    IPTDavidsonHistoryEntry *davidson_history;
    int davidson_restart_count;
    int davidson_selection_history_count;
    IPTDavidsonSelectionEntry *davidson_selection_history;
    int davidson_block_history_count;
    IPTDavidsonBlockHistoryEntry *davidson_block_history;
    int best_so_far_enabled;
    int best_so_far_updated;
    int best_so_far_update_count;
    int best_so_far_step;
    char best_so_far_source[32];
    int best_so_far_basis_cols;
    double best_so_far_max_residual;
    int best_so_far_max_residual_index;
    int final_returned_from_best_so_far;
    double pair28_best_residual;
    double pair29_best_residual;
    int accepted_steps;
    int rejected_steps;
    int early_jump_to_continuation;
    double gap28_29;
    double relative_gap28_29;
    double cluster_residual_fro;
    double cluster_residual_max;
    char pair28_lock_state[16];
    char pair29_lock_state[16];
    int cluster_hard_locked;
    int cluster_soft_locked_count;
} IPTCudaResult;

/* ipt_cuda_status_string, IPT/src/ipt_cuda.cu:375-395. */
const char *ipt_cuda_status_string(int status);
/* ipt_cuda_free_result, IPT/src/ipt_cuda.cu:398-475. */
void ipt_cuda_free_result(IPTCudaResult *result);
/* ipt_cuda_free_device_result, IPT/src/ipt_cuda.cu:478-483 (hipFree). */
void ipt_cuda_free_device_result(double *d_vectors, double *d_values);

/* Dense plain IPT from host memory, fixed iterations / tolerance stop;
 * IPT/src/ipt_cuda.cu:1764-1777, 1975-1992 (with_initial variants).
 * With IPT_BLOCK_CLUSTER=1 and no initial vectors, dense input takes the
 * block-cluster -> Rayleigh-Ritz (-> Davidson) path of the sparse API. */
int ipt_cuda(const double *matrix_col_major, int n, int k, int iterations,
             IPTCudaResult *result);
int ipt_cuda_tol(const double *matrix_col_major, int n, int k, double tol,
                 int maxiter, IPTCudaResult *result);
int ipt_cuda_with_initial(const double *matrix_col_major,
                          const double *initial_vectors_col_major, int n,
                          int k, int iterations, IPTCudaResult *result);
int ipt_cuda_with_initial_tol(const double *matrix_col_major,
                              const double *initial_vectors_col_major, int n,
                              int k, double tol, int maxiter,
                              IPTCudaResult *result);

/* Dense plain IPT on device memory; IPT/src/ipt_cuda.cu:1741-1762.
 * Outputs are library-allocated device buffers released with
 * ipt_cuda_free_device_result. */
int ipt_cuda_device(const double *d_matrix_col_major, int n, int k,
                    int iterations, void *handle, double **d_vectors_out,
                    double **d_values_out);
int ipt_cuda_device_tol(const double *d_matrix_col_major, int n, int k,
                        double tol, int maxiter, void *handle,
                        double **d_vectors_out, double **d_values_out,
                        int *iterations_done_out,
                        double *fixed_point_residual_out);

/* Sparse CSC from host memory; IPT/src/ipt_cuda.cu:2227-2253, 5298-5316.
 * IPT_BLOCK_CLUSTER=1 (no initial vectors) selects block-cluster IPT ->
 * Rayleigh-Ritz, and IPT_DAVIDSON_ENRICH=1 adds Davidson refinement,
 * IPT/src/ipt_cuda.cu:5070-5075, 4915. */
int ipt_cuda_sparse_csc(const int *col_ptr, const int *row_ind,
                        const double *matrix_values, int n, int k, int nnz,
                        int iterations, IPTCudaResult *result);
int ipt_cuda_sparse_csc_tol(const int *col_ptr, const int *row_ind,
                            const double *matrix_values, int n, int k,
                            int nnz, double tol, int maxiter,
                            IPTCudaResult *result);
int ipt_cuda_sparse_csc_with_initial(
    const int *col_ptr, const int *row_ind, const double *matrix_values,
    const double *initial_vectors_col_major, int n, int k, int nnz,
    int iterations, IPTCudaResult *result);
int ipt_cuda_sparse_csc_with_initial_tol(
    const int *col_ptr, const int *row_ind, const double *matrix_values,
    const double *initial_vectors_col_major, int n, int k, int nnz,
    double tol, int maxiter, IPTCudaResult *result);

/* Sparse CSC on device memory; IPT/src/ipt_cuda.cu:2214-2225 (plain IPT). */
int ipt_cuda_sparse_csc_device(
    const int *d_col_ptr, const int *d_row_ind,
    const double *d_matrix_values, int n, int k, int nnz, int iterations,
    void *handle, double **d_vectors_out, double **d_values_out);
int ipt_cuda_sparse_csc_device_tol(
    const int *d_col_ptr, const int *d_row_ind,
    const double *d_matrix_values, int n, int k, int nnz, double tol,
    int maxiter, void *handle, double **d_vectors_out,
    double **d_values_out, int *iterations_done_out,
    double *fixed_point_residual_out);

/* Handle API (replaces cublasHandle_t; no counterpart in the reference).
 * A handle owns one HIP stream of default priority and a workspace.
 * stream may be NULL (handle creates its own). */
int ipt_hip_create_handle(void **handle_out, void *hip_stream);
int ipt_hip_destroy_handle(void *handle);

#ifdef __cplusplus
}
#endif

#endif /* IPT_HIP_H */
