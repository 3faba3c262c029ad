// Driver (ours) that compiles the reference's host-only preparation source
// IPT/src/ipt_preparation.cu in place and exposes its cluster detection
// through a C entry point, so tests can compare the oracle's restatement
// with the reference's own code.  Built only where the reference checkout
// exists; the output goes to oracle/_ref/.  Test infrastructure only.
#include IPT_REF_SRC

extern "C" int ref_target_clusters(const int *col_ptr, const int *row_ind,
                                   const double *values, int n, int k,
                                   double degeneracy_threshold,
                                   int *ranges_out, int *too_large)
{
    std::vector<int> cp(col_ptr, col_ptr + n + 1);
    std::vector<int> ri(row_ind, row_ind + cp[n]);
    std::vector<double> v(values, values + cp[n]);
    std::vector<double> diag;
    ipt_extract_csc_diagonal_vectors(cp, ri, v, n, &diag);
    IPTDegeneracyOptions options = ipt_degeneracy_options(degeneracy_threshold);
    std::vector<std::pair<int, int>> ranges =
        ipt_target_degenerate_subspaces(cp, ri, v, diag, k, options, too_large);
    for (size_t i = 0; i < ranges.size(); ++i) {
        ranges_out[2 * i] = ranges[i].first;
        ranges_out[2 * i + 1] = ranges[i].second;
    }
    return (int)ranges.size();
}

// The reference's plain-CSC preparation (sort + degeneracy lifting):
// prepared matrix written densely (n x n column-major), perm_out[sorted] =
// original, rotation count; returns its status (0 ok, 1 invalid, 2 memory).
extern "C" int ref_prepare_sparse_csc(const int *col_ptr, const int *row_ind, const double *values, int n, int k,
                                      int nnz, double degeneracy_threshold, double *dense_out, int *perm_out,
                                      int *rotations_out)
{
    IPTPreparedCsc prep;
    int st = ipt_prepare_sparse_csc(col_ptr, row_ind, values, n, k, nnz, 1, 1, degeneracy_threshold, &prep);
    if (st != 0) {
        return st;
    }
    for (size_t i = 0; i < (size_t)n * n; ++i) {
        dense_out[i] = 0.0;
    }
    for (int c = 0; c < n; ++c) {
        for (int p = prep.col_ptr[c]; p < prep.col_ptr[c + 1]; ++p) {
            dense_out[(size_t)prep.row_ind[p] + (size_t)c * n] += prep.values[p];
        }
    }
    for (int i = 0; i < n; ++i) {
        perm_out[i] = (prep.permuted && prep.basis_to_original != NULL) ? prep.basis_to_original[i] : i;
    }
    *rotations_out = prep.rotation_count;
    ipt_prepared_csc_free(&prep);
    return 0;
}

// Vectors from the prepared basis to the caller's (rotations, then the
// inverse permutation) with the reference's own routine.
extern "C" int ref_prepared_vectors_to_original(const int *col_ptr, const int *row_ind, const double *values,
                                                int n, int k, int nnz, double degeneracy_threshold,
                                                double *vectors, int kv)
{
    IPTPreparedCsc prep;
    int st = ipt_prepare_sparse_csc(col_ptr, row_ind, values, n, k, nnz, 1, 1, degeneracy_threshold, &prep);
    if (st != 0) {
        return st;
    }
    int ok = ipt_apply_prepared_q_to_vectors(&prep, vectors, kv);
    ipt_prepared_csc_free(&prep);
    return ok ? 0 : 3;
}
