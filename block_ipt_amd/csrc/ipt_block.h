// Block-cluster IPT and FP64 Rayleigh-Ritz: shared device-side plan and
// kernel launchers.
#pragma once

#include "ipt_internal.h"

namespace ipt {

struct BlockCluster {
    int first;    // first matrix row of the cluster (sorted order)
    int last;     // last matrix row
    int off;      // first basis column
    int m;        // width = last - first + 1
    int nblocks;  // row blocks of its update launch (partials per cluster)
    int pad;
};

// Update grid of a cluster of width 2..32 (ipt_block_width.h).  Widths of
// at least kSeriesWidth: near blocks first (the rows close to the
// cluster's spectrum, P = next_pow2(m) lanes per row, 256 / P rows per
// block and pass, kNearRowSlots rows per pass), then one far block per
// kFarRows rows (kFarLanes lanes per row).  Narrower widths: one block per
// 256 / P rows, every row eliminated.  Each block owns one partial slot of
// the cluster.
constexpr int kSeriesWidth = 9;
constexpr int kNearRowSlots = 512;
constexpr int kFarLanes = 4;
constexpr int kFarRows = 256 / kFarLanes;

__host__ __device__ constexpr int width_lanes(int m)
{
    return m <= 2 ? 2 : m <= 4 ? 4 : m <= 8 ? 8 : m <= 16 ? 16 : 32;
}

__host__ __device__ constexpr int near_blocks_for(int m)
{
    return kNearRowSlots / (256 / width_lanes(m));
}

__host__ __device__ constexpr int update_blocks_for(int m, int n)
{
    return m >= kSeriesWidth ? near_blocks_for(m) + (n + kFarRows - 1) / kFarRows
                             : (n + 256 / width_lanes(m) - 1) / (256 / width_lanes(m));
}

// Device-resident description of all clusters, grouped by lane class
// P = next_pow2(m) in {1,...,64}.
struct BlockPlan {
    int nclusters = 0;
    int basis_cols = 0;
    int class_count[7] = {0, 0, 0, 0, 0, 0, 0};
    int class_offset[7] = {0, 0, 0, 0, 0, 0, 0};
    std::vector<BlockCluster> host;
    DevBuf<BlockCluster> clusters;
    DevBuf<int> class_lists;  // cluster ids, class-major
    DevBuf<int> active;       // 1 while the cluster iterates
    DevBuf<int> iters;        // iterations done per cluster
    DevBuf<double> fp;        // last fixed-point residual per cluster
    DevBuf<int> tentative;    // stopped after a low-precision step, awaiting an exact update
    DevBuf<double> fp_prev;   // fixed-point residual of the previous step (stall detection)
    DevBuf<int> growth;       // consecutive steps with a growing residual
    DevBuf<int> stalled;      // stopped by stall detection
    int stall_steps = 0;      // growing steps that stall a cluster (0: no detection)
    std::vector<int> skipped; // clusters never iterated (wider than the block limit)
    DevBuf<int> flags;        // [0] singular, [1] status bits, [2] active count
    // written by each finalisation for one read-back: [0] status bits,
    // [1] active count, [2 + c] cluster c active, [2 + nclusters + c] its fp
    DevBuf<double> status;
    int partial_stride = 0;   // max nblocks over clusters
    int w1_count = 0;         // clusters of width 1 (their own kernel)
    DevBuf<int> w1_list;      // their ids
    int wide_count = 0;          // clusters of width 2..32
    int width_count[33] = {};    // clusters of each width 2..32
    int width_offset[33] = {};   // their position in width_lists
    DevBuf<int> width_lists;     // cluster ids grouped by width
    std::vector<int> wide_host;  // clusters of width 65..142 (one launch each)
    DevBuf<int> wide_list;       // their ids
    // host view of clusters frozen for good (as of the last exact step):
    // launches whose clusters are all frozen are skipped; empty: none known
    std::vector<char> frozen;
    DevBuf<double> cmax;  // operand emission: per cluster step and anchor maxima of the last update
};

struct LowpEmit;

// Build the plan for ranges [first,last] (sorted by first) on n rows.
// skip_wider > 0: clusters wider than that are left out of every update
// launch (they keep their identity blocks; block_plan_mark_skipped then
// marks them stalled, for Rayleigh-Ritz and Davidson to take over).
int block_plan_create(Context *ctx, const std::vector<std::pair<int, int>> &ranges, int n, BlockPlan *plan,
                      int skip_wider = 0);
int block_plan_mark_skipped(Context *ctx, BlockPlan *plan);
// Reset iteration state (all clusters active, counters zero).
int block_plan_reset(Context *ctx, BlockPlan *plan);
// B = identity blocks of every cluster (column-major, ld).
int block_identity(Context *ctx, const BlockPlan &plan, double *B, long ld, int n);
// One batched update B -> Bn from Y = H B; then, when count_steps > 0,
// per-cluster norms of this update, freeze converged clusters (fp <= tol),
// add count_steps to the iteration count of every cluster still active and
// set flags.  count_steps = 0 skips that finalisation: the caller then
// counts the step with the next finalised one (clusters only freeze in
// finalised steps, so the active set is the same over those steps).
// lowp_step = 1: Y came from a low-precision product; clusters reaching
// fp <= tol stop and are marked tentative (see block_reactivate_tentative).
// emit != NULL: the update also writes the operand of the next
// low-precision product (ipt_emit.h); the caller keeps the planes,
// colscale and plan.cmax; em.mpart is set here.
int block_update(Context *ctx, BlockPlan &plan, const double *B, const double *Y, double *Bn, long ld,
                 const double *d, int n, double damping, double tol, int count_steps = 1, int lowp_step = 0,
                 const LowpEmit *emit = nullptr);
// Before an exact step: tentative clusters iterate once more (with the
// exact product), after which the finalisation freezes them.
int block_reactivate_tentative(Context *ctx, BlockPlan &plan);
// B(:, columns of the clusters in list) = their identity blocks.
int block_identity_clusters(Context *ctx, const BlockPlan &plan, const std::vector<int> &list, double *B, long ld,
                            int n);

// C (m x nc, ldc) = A^T B for A (nrows x m, lda), B (nrows x nc, ldb), all
// column-major FP64; fixed-order two-stage reduction.
int gemm_tn_tall(Context *ctx, const double *A, long lda, const double *B, long ldb, double *C, long ldc,
                 long nrows, int m, int nc);
// C (nrows x nc, ldc) = A (nrows x m, lda) * W (m x nc, ldw); alpha/beta form
// C = alpha*A*W + beta*C.
int gemm_nn_small(Context *ctx, const double *A, long lda, const double *W, long ldw, double *C, long ldc,
                  long nrows, int m, int nc, double alpha, double beta);
// sq[j] = sum_r X[r + j*ld]^2
int column_sqnorms(Context *ctx, const double *X, long ld, long nrows, int nc, double *sq);
// X[:, j] *= 1/sqrt(sq[j]) (and the same for HX when non-NULL)
int scale_columns_inv_sqrt(Context *ctx, double *X, double *HX, long ld, long nrows, int nc,
                           const double *sq);
// Right-looking Cholesky of the symmetric G (t x t, ld t, overwritten by L)
// in one workgroup, column j kept iff its pivot d is positive and finite
// and, with ref != NULL, sqrt(d) > rel sqrt(ref[j]) for a finite positive
// ref[j]; keep[j] = 1/0 and *nkept on the device.  W (t x t, ld t) =
// (L^-1)^T on the kept rows and columns, zero elsewhere, so that for a
// basis T with G = T^T T the kept columns of T W are orthonormal.
int chol_reject_device(Context *ctx, double *G, int t, const double *ref, double rel, double *W, int *keep,
                       int *nkept);
// D(:, b) = S(:, idx[b]) for b < nc, both with leading dimension t.
int gather_columns(Context *ctx, const double *S, int t, const int *idx, int nc, double *D);
// Q (n x k, ld n) <- the orthonormal factor of its QR factorisation
// (positive diagonal R) by shifted Cholesky QR 3, all in FP64 on the
// device.  IPT_CUDA_INVALID_ARGUMENT for a zero, non-finite or numerically
// dependent column.
int cholqr3_device(Context *ctx, double *Q, int n, int k);
// For a symmetric positive Gram matrix G (m x m, ld m): *dev_out =
// ||D^-1/2 G D^-1/2 - I||_F with D = diag(G).  When that is at most max_dev
// (< 1), S (m x m, ld m) = D^-1/2 (D^-1/2 G D^-1/2)^-1/2, so that
// S^T G S = I, and *applied = 1; otherwise S is untouched and *applied = 0.
int gram_inverse_sqrt(Context *ctx, const double *G, int m, double max_dev, double *S, double *dev_out,
                      int *applied);
// res[j] = ||HX_j - theta_j X_j|| / (max(1,normH) max(1,||X_j||)); also
// writes R = HX - X theta when R != NULL.
int residual_norms(Context *ctx, const double *X, const double *HX, long ld, long nrows, int nc,
                   const double *theta, double normH, double *res, double *R);
// out[perm[s] + j*ld] = in[s + j*ld]
int unpermute_rows(Context *ctx, const double *in, double *out, long ld, long nrows, int nc, const int *perm);

}  // namespace ipt

namespace ipt {
// Dense preparation scan: column abs sums (colsum), column max |offdiag|
// (colmax, as doubles), symmetry flag asym (device int).
int dense_scan(Context *ctx, const double *H, long ld, int n, double *colsum, double *colmax, int *asym);
// out = P H P^T with perm[sorted] = original.
int dense_permute(Context *ctx, const double *H, long ldh, double *out, long ld, int n, const int *perm);
// vals[2i] = H(r_i, c_i), vals[2i+1] = H(c_i, r_i) for pairs rc[2i], rc[2i+1].
int gather_pairs(Context *ctx, const double *H, long ld, const int *d_rc, int count, double *d_vals);
// T(:, q) = R(:, idx[q]) / clip(d - theta[idx[q]])
int diag_correction(Context *ctx, const double *R, long ldr, const double *d, const double *theta,
                    const int *d_idx, int nt, double clip, double *T, long ldt, long nrows);
}  // namespace ipt
