// Host-side preparation of the sparse / block-cluster paths.
#pragma once

#include <utility>
#include <vector>

namespace ipt {

struct HostCsc {
    int n = 0;
    std::vector<int> col_ptr;
    std::vector<int> row_ind;
    std::vector<double> values;
};

// Options of the degenerate-cluster detection (IPT_* environment variables,
// IPT/src/ipt_preparation.cu:116-146).
struct DegeneracyOptions {
    double absolute_gap = 0.0;
    double eps_multiplier = 1000.0;
    double rel_gap_tol = 1.0e-12;
    double ratio_tol = 0.2;
    double eta = 0.02;
    int max_block = 64;
    int max_lift_rounds = 3;
};
DegeneracyOptions degeneracy_options_from_env(double degeneracy_threshold);

// CSC validity, IPT/src/ipt_preparation.cu:148-169.
bool csc_valid(const int *col_ptr, const int *row_ind, int n, int nnz);
// Diagonal (first stored match per column).
void csc_diagonal_host(const int *col_ptr, const int *row_ind, const double *values, int n,
                       std::vector<double> *diag);
// perm[sorted] = original; inverse[original] = sorted (stable by value, then index).
void stable_sort_permutation(const std::vector<double> &diag, std::vector<int> *perm,
                             std::vector<int> *inverse);
// B = P A P^T with duplicate entries summed and exact zeros dropped
// (IPT/src/ipt_preparation.cu:210-262).
void permute_csc(const int *col_ptr, const int *row_ind, const double *values, int n,
                 const std::vector<int> &inverse, HostCsc *out);
// CSR arrays of A from its CSC arrays (= CSC of A^T), rows sorted by column.
void csc_to_csr(const int *col_ptr, const int *row_ind, const double *values, int n,
                std::vector<int> *row_ptr, std::vector<int> *col_ind, std::vector<double> *vals);
// |a_ij - a_ji| <= 1e-12 max(1,|a_ij|,|a_ji|), IPT/src/ipt_preparation.cu:860-883.
bool csc_is_symmetric(const HostCsc &a);
// max_i sum_j |a_ij|
double csc_inf_norm(const int *col_ptr, const int *row_ind, const double *values, int n);

// Interface to the matrix for cluster detection: coupling lookups and the
// off-diagonal strengths.  Implemented for CSC and for dense column-major.
struct CouplingSource {
    virtual ~CouplingSource() = default;
    virtual int n() const = 0;
    virtual double diag(int i) const = 0;
    virtual double value(int row, int col) const = 0;  // 0 if not stored
    // max(|offdiag| in column i, |offdiag| in row i)
    virtual void offdiag_strengths(std::vector<double> *s) const = 0;
};
struct CscCoupling : CouplingSource {
    const HostCsc *a;
    explicit CscCoupling(const HostCsc *m) : a(m) {}
    int n() const override { return a->n; }
    double diag(int i) const override;
    double value(int row, int col) const override;
    void offdiag_strengths(std::vector<double> *s) const override;
};
struct DenseCoupling : CouplingSource {
    const double *h;
    long ld;
    int nn;
    DenseCoupling(const double *m, long l, int n_) : h(m), ld(l), nn(n_) {}
    int n() const override { return nn; }
    double diag(int i) const override { return h[(long)i + (long)i * ld]; }
    double value(int row, int col) const override { return h[(long)row + (long)col * ld]; }
    void offdiag_strengths(std::vector<double> *s) const override;
};

// Target clusters: ranges [first,last] of rows whose union-find component
// contains a target column < k, merged, then singleton completion for
// uncovered targets (IPT/src/ipt_preparation.cu:501-645,
// IPT/src/ipt_cuda.cu:4114-4132).  `diag` is the sorted diagonal.
std::vector<std::pair<int, int>> target_clusters(const CouplingSource &m, const std::vector<double> &diag,
                                                 int k, const DegeneracyOptions &o, bool *too_large,
                                                 bool add_targets = true);

// Number of (row, target column) pairs that the cluster rule would still
// join, IPT/src/ipt_preparation.cu:658-757 (0: every target is separated).
int target_gap_violations(const CouplingSource &m, const std::vector<double> &diag, int k,
                          const DegeneracyOptions &o);

// Sparse preparation of the plain path, IPT/src/ipt_preparation.cu:1375-1648:
// stable sort of the diagonal (symmetric permutation) and, for symmetric
// input, lifting of the degenerate target subspaces by local eigenvector
// similarities, up to max_lift_rounds rounds, until no target pair is joined.
struct LiftedCsc {
    HostCsc a;                  // prepared matrix
    std::vector<int> perm;      // perm[sorted] = original (empty: not permuted)
    std::vector<int> rot_first; // rotations in the order applied
    std::vector<int> rot_m;
    std::vector<double> rot_q;  // their m x m eigenvector blocks, column-major
};
// Status: 0 ok; 1 invalid (a subspace wider than max_block, a non-symmetric
// matrix needing a lift, or targets still joined after the last round).
int prepare_lifted_csc(const int *col_ptr, const int *row_ind, const double *values, int n, int k, int nnz,
                       double degeneracy_threshold, LiftedCsc *out);
// vectors (n x k column-major, prepared basis) -> caller's basis: the
// rotations in reverse order, then the inverse permutation,
// IPT/src/ipt_preparation.cu:1222-1294.
void lifted_to_original(const LiftedCsc &p, double *vectors, int k);
// Eigenvectors of a symmetric m x m block (column-major) by classical
// Jacobi, columns ordered by ascending Rayleigh quotient,
// IPT/src/ipt_preparation.cu:759-828, 921-977.
void sym_block_eigenvectors(const std::vector<double> &a, int m, std::vector<double> *v);

// Target count and adaptive expansion of the block-cluster path,
// IPT/src/ipt_cuda.cu:4596-4613: IPT_BLOCK_CLUSTER_OVERSAMPLE (0) more
// targets (all n when it exceeds n - k), IPT_BLOCK_CLUSTER_ADAPTIVE (off;
// forces oversample 0) with coupling threshold IPT_BLOCK_CLUSTER_COUPLING_TAU
// (0.1).
struct BlockTargets {
    int target_k = 0;
    int oversample = 0;
    bool adaptive = false;
    double tau = 0.1;
};
BlockTargets block_targets_from_env(int n, int k);
// Grow the cluster that holds target k-1 by its most strongly coupled
// outside row (|a_rc| / max(|d_c - d_r|, 1e-14) > tau), with rows of
// (numerically) equal diagonal alongside, while the widened range stays
// within max_block; overlapping ranges merge.  added: the rows brought in
// (sorted, unique).  Returns whether the width limit stopped a coupling
// above tau.  IPT/src/ipt_cuda.cu:4134-4272.
bool adaptive_expand(const HostCsc &a, const std::vector<double> &d, int k, int max_block, double tau,
                     std::vector<std::pair<int, int>> *ranges, std::vector<int> *added);

}  // namespace ipt
