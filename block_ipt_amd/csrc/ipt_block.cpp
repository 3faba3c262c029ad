// Block-cluster IPT -> FP64 Rayleigh-Ritz -> Davidson refinement.
//
// Reference pipeline: IPT/src/ipt_cuda.cu:4712-5016 (sparse CSC, selected by
// IPT_BLOCK_CLUSTER=1) with ipt_block_cluster_solve_basis_gpu (2871-3037),
// ipt_block_cluster_rayleigh_ritz_gpu (3039-3316) and
// ipt_block_davidson_enrich_gpu (3318-3982).
//
// MI355X design differences (results are the same up to rounding):
//  * all clusters iterate together: one H*B product per iteration over the
//    concatenated cluster bases (the reference solves clusters one after
//    another); a cluster that reaches its tolerance is frozen on the device
//    and its columns are carried through unchanged, so every cluster stops
//    at exactly the iteration the sequential solve would;
//  * Rayleigh-Ritz keeps H*B from the last product instead of recomputing
//    it; Davidson keeps H*V for the whole search space and multiplies only
//    the new correction block by H;
//  * Davidson corrections are orthogonalised as a block (two classical
//    Gram-Schmidt passes against the basis, then a Cholesky QR of the block
//    with rejection of numerically dependent directions) instead of one
//    vector at a time.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cstring>
#include <functional>

#include "ipt_block.h"
#include "ipt_emit.h"
#include "ipt_lowp.h"
#include "ipt_sparse_lp.h"
#include "ipt_prep.h"

namespace ipt {

// learned preconditioner (ipt_mlp.hip)
int mlp_correction(Context *ctx, const float *params, const double *R, long ldr, const double *d,
                   const double *rowabs, const double *theta, const int *idx, int n, int nt, double *T, long ldt);

namespace {

double seconds_since(const std::chrono::steady_clock::time_point &t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}


// Y(:, 0:nc) = H X(:, 0:nc), both column-major with leading dimension n.
using ApplyFn = std::function<int(const double *X, double *Y, int nc)>;
// Y = Ya + diag(d) (B - Ba) + E (B - Ba) with E (B - Ba) in low precision
using DeltaFn = std::function<int(const double *B, const double *Ba, const double *Ya, int nc, double *Y)>;

struct BlockParams {
    int maxiter = 100;
    double tol = 1e-12;
    double damping = 1.0;
    int davidson = 0;
    int davidson_steps = 1;
    int davidson_extra = 0;
    double conv_tol = 1e-13;
    int lowp_mode = -1;     // -1: every product in FP64
    int anchor_every = 0;   // low-precision steps between exact FP64 steps (0: 20 dense, 5 CSR)
    int precond_mlp = 0;    // Davidson corrections from the learned MLP
    std::string mlp_model;  // its parameter file
    int verbose = 0;
    int stall_steps = 0;    // consecutive growing steps that stop a cluster (0: never)
    int max_basis = 0;      // Davidson search-space limit (0: the reference's k + max(16, (2k+2)/3))
    int retain = -1;        // corrections kept over a restart (-1: the reference's max_basis - k - 1)
    int max_block = 64;     // IPT_MAX_LIFT_BLOCK_SIZE
    int emit = 1;           // dense FP16/BF16/BF16x3: the update writes the next product's operand
};

// Davidson search-space limit for k targets on n rows
int davidson_max_basis(const BlockParams &p, int n, int k)
{
    const int kmax = p.max_basis > 0 ? std::max(p.max_basis, k + 1) : k + std::max(16, (2 * k + 2) / 3);
    return std::min(n, kmax);
}

constexpr int kMlpParams = 4481;

// Parameter file: tag "IPTMLP01", int32 count, FP32 parameters.
int load_mlp_model(const std::string &path, DevBuf<float> *out)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (f == nullptr) {
        fprintf(stderr, "IPT: cannot open MLP model %s\n", path.c_str());
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    char tag[8];
    int cnt = 0;
    std::vector<float> h((size_t)kMlpParams);
    const bool ok = fread(tag, 1, 8, f) == 8 && memcmp(tag, "IPTMLP01", 8) == 0 && fread(&cnt, sizeof(int), 1, f) == 1 &&
                    cnt == kMlpParams && fread(h.data(), sizeof(float), h.size(), f) == h.size();
    fclose(f);
    if (!ok) {
        fprintf(stderr, "IPT: %s is not an MLP model file with %d parameters\n", path.c_str(), kMlpParams);
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    IPT_STCHK(out->alloc(h.size()));
    IPT_HIPCHK(hipMemcpy(out->p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    return IPT_CUDA_SUCCESS;
}

// Synchronous device-to-host reads of several arrays with one stream
// synchronisation, staged through pinned memory (a copy into pageable
// memory is staged by the runtime and waits on its own): the context's
// pinned scalars up to 32 KiB in all, else its pinned staging buffer (which
// must not hold data of a caller then).
struct HostRead {
    void *dst;
    const void *src;
    size_t bytes;
};

int read_host_many(Context *ctx, const std::vector<HostRead> &reads)
{
    size_t total = 0;
    for (const auto &r : reads) {
        total += (r.bytes + 7) / 8 * 8;
    }
    char *stage = reinterpret_cast<char *>(ctx->host_scalars);
    if (total > 4096 * sizeof(double)) {
        IPT_STCHK(context_host_pinned(ctx, total, (void **)&stage));
    }
    size_t o = 0;
    for (const auto &r : reads) {
        IPT_HIPCHK(hipMemcpyAsync(stage + o, r.src, r.bytes, hipMemcpyDeviceToHost, ctx->stream));
        o += (r.bytes + 7) / 8 * 8;
    }
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    o = 0;
    for (const auto &r : reads) {
        memcpy(r.dst, stage + o, r.bytes);
        o += (r.bytes + 7) / 8 * 8;
    }
    return IPT_CUDA_SUCCESS;
}

int read_host(Context *ctx, void *dst, const void *src, size_t bytes)
{
    return read_host_many(ctx, {HostRead{dst, src, bytes}});
}

// Internal status of rayleigh_ritz: the basis has numerical rank below k.
constexpr int kRankDeficient = 100;

// Options of one Rayleigh-Ritz projection (IPT/src/ipt_cuda.cu:3039-3316):
// use_qr -- IPT_BLOCK_CLUSTER_QR: the projection is formed on an explicitly
// orthonormalised basis Q and its images HQ (the reference's geqrf/orgqr,
// then Q^T H Q and syevd) instead of the Gram/projected pair; stats --
// IPT_RR_HOST_ORTHOGONALITY_STATS: ||B^T B - I|| of the basis used and
// ||X^T X - I|| of the Ritz vectors (:2489-2524, :3168-3206).
struct RROptions {
    bool use_qr = false;
    bool stats = false;
};
struct RRInfo {
    int used_qr = 0;
    double basis_fro = NAN, basis_max = NAN;
    int basis_bad = 0;
    double ritz_fro = NAN, ritz_max = NAN;
    int ritz_bad = 0;
};

RROptions rr_options_from_env()
{
    RROptions o;
    o.use_qr = env_flag("IPT_BLOCK_CLUSTER_QR") != 0;
    o.stats = env_flag("IPT_RR_HOST_ORTHOGONALITY_STATS") != 0;
    return o;
}

// ||G - I||_F and max |G - I| of a symmetric m x m device matrix (NaN-aware:
// *bad = 1 when an entry is not finite)
int identity_deviation(Context *ctx, const double *G, int m, double *fro, double *mx, int *bad)
{
    std::vector<double> h((size_t)m * m);
    IPT_HIPCHK(hipMemcpyAsync(h.data(), G, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    double f = 0.0, x = 0.0;
    int b = 0;
    for (int j = 0; j < m; ++j) {
        for (int i = 0; i < m; ++i) {
            const double v = h[(size_t)i + (size_t)j * m];
            if (!std::isfinite(v)) {
                b = 1;
                continue;
            }
            const double e = v - (i == j ? 1.0 : 0.0);
            f += e * e;
            x = std::max(x, std::fabs(e));
        }
    }
    *bad = b;
    *fro = b ? NAN : std::sqrt(f);
    *mx = b ? NAN : x;
    return IPT_CUDA_SUCCESS;
}

// Rayleigh-Ritz over V (n x cols) with HV = H V, on the device:
//   G = V^T V; S = G^-1/2 by Newton-Schulz when the column-scaled G is
//   within 0.25 of the identity, else G = W diag(g) W^T (the device
//   eigensolver) and S = the columns of W with g > 1e-13 max(g) scaled by
//   g^-1/2 (canonical orthogonalisation: V S is an orthonormal basis of
//   span(V), dependent directions dropped);
//   A = S^T (V^T H V) S = Z diag(theta) Z^T, C = S Z,
//   X = V C(:, :k), HX = HV C(:, :k), columns of X normalised (HX alike).
// The same pairs as the Cholesky-based cusolverDnDsygvd of the reference
// when G is positive definite; a nearly dependent basis (rank >= k) is
// handled instead of failing.  theta (device) receives all r values of the
// kept directions, +inf after them up to cols.  kRankDeficient when fewer
// than k directions remain.
int rayleigh_ritz(Context *ctx, const double *V, const double *HV, int n, int cols, int k, double *theta,
                  double *X, double *HX, const RROptions &opt = RROptions(), RRInfo *info = nullptr)
{
    RRInfo dummy;
    RRInfo &inf = info != nullptr ? *info : dummy;
    inf = RRInfo();
    const size_t cc = (size_t)cols * cols;
    DevBuf<double> G, P, W, gw, S, T, A, Z, C;
    IPT_STCHK(G.alloc(cc));
    IPT_STCHK(P.alloc(cc));
    IPT_STCHK(W.alloc(cc));
    IPT_STCHK(gw.alloc((size_t)cols));
    IPT_STCHK(gemm_tn_tall(ctx, V, n, V, n, G.p, cols, n, cols, cols));
    IPT_STCHK(symmetrize(ctx, G.p, cols, cols));
    if (opt.stats && !opt.use_qr) {
        IPT_STCHK(identity_deviation(ctx, G.p, cols, &inf.basis_fro, &inf.basis_max, &inf.basis_bad));
    }
    // A basis within 0.25 (Frobenius) of orthonormal after column scaling
    // (always so for Davidson's orthonormalised basis) is full rank with
    // condition number below 5/3: S = G^-1/2 by a few products, no
    // eigensolve.  Otherwise the canonical orthogonalisation below.
    int r = cols;
    {
        IPT_STCHK(S.alloc(cc));
        double dev = NAN;
        int applied = 0;
        IPT_STCHK(gram_inverse_sqrt(ctx, G.p, cols, 0.25, S.p, &dev, &applied));
        trace_mark("Gram inverse square root enqueued");
        if (!applied) {
            IPT_STCHK(sym_eig_device(ctx, G.p, cols, cols, gw.p, W.p, cols, nullptr));
            std::vector<double> hg((size_t)cols);
            IPT_STCHK(read_host(ctx, hg.data(), gw.p, hg.size() * sizeof(double)));
            const double gmax = hg.back();
            int first = cols;
            while (first > 0 && hg[(size_t)first - 1] > 1e-13 * gmax) {
                --first;
            }
            r = cols - first;
            if (!(gmax > 0.0) || !std::isfinite(gmax) || r < k) {
                return kRankDeficient;
            }
            // S (cols x r): the kept eigenvectors (the last r) over sqrt(g)
            IPT_HIPCHK(hipMemcpyAsync(S.p, W.p + (size_t)first * cols, (size_t)cols * r * sizeof(double),
                                      hipMemcpyDeviceToDevice, ctx->stream));
            IPT_STCHK(scale_columns_inv_sqrt(ctx, S.p, nullptr, cols, cols, r, gw.p + first));
        }
    }
    IPT_STCHK(A.alloc((size_t)r * r));
    DevBuf<double> Q, HQ;
    if (opt.use_qr) {
        // explicit orthonormal basis Q = V S and its images HQ = HV S, then
        // the projection Q^T (H Q)
        IPT_STCHK(Q.alloc((size_t)n * r));
        IPT_STCHK(HQ.alloc((size_t)n * r));
        IPT_STCHK(gemm_nn_small(ctx, V, n, S.p, cols, Q.p, n, n, cols, r, 1.0, 0.0));
        IPT_STCHK(gemm_nn_small(ctx, HV, n, S.p, cols, HQ.p, n, n, cols, r, 1.0, 0.0));
        IPT_STCHK(gemm_tn_tall(ctx, Q.p, n, HQ.p, n, A.p, r, n, r, r));
        if (opt.stats) {
            IPT_STCHK(T.alloc((size_t)r * r));
            IPT_STCHK(gemm_tn_tall(ctx, Q.p, n, Q.p, n, T.p, r, n, r, r));
            IPT_STCHK(identity_deviation(ctx, T.p, r, &inf.basis_fro, &inf.basis_max, &inf.basis_bad));
        }
        inf.used_qr = 1;
    } else {
        IPT_STCHK(gemm_tn_tall(ctx, V, n, HV, n, P.p, cols, n, cols, cols));
        IPT_STCHK(T.alloc((size_t)r * cols));
        IPT_STCHK(gemm_tn_tall(ctx, S.p, cols, P.p, cols, T.p, r, cols, r, cols));
        IPT_STCHK(gemm_nn_small(ctx, T.p, r, S.p, cols, A.p, r, r, cols, r, 1.0, 0.0));
    }
    IPT_STCHK(symmetrize(ctx, A.p, r, r));
    IPT_STCHK(Z.alloc((size_t)r * r));
    IPT_STCHK(sym_eig_device(ctx, A.p, r, r, theta, Z.p, r, nullptr));
    trace_mark("projected eigensolve enqueued");
    if (r < cols) {
        std::vector<double> inf_fill((size_t)(cols - r), INFINITY);
        IPT_HIPCHK(hipMemcpyAsync(theta + r, inf_fill.data(), inf_fill.size() * sizeof(double), hipMemcpyHostToDevice,
                                  ctx->stream));
        IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    if (opt.use_qr) {
        IPT_STCHK(gemm_nn_small(ctx, Q.p, n, Z.p, r, X, n, n, r, k, 1.0, 0.0));
        IPT_STCHK(gemm_nn_small(ctx, HQ.p, n, Z.p, r, HX, n, n, r, k, 1.0, 0.0));
    } else {
        IPT_STCHK(C.alloc((size_t)cols * k));
        IPT_STCHK(gemm_nn_small(ctx, S.p, cols, Z.p, r, C.p, cols, cols, r, k, 1.0, 0.0));
        IPT_STCHK(gemm_nn_small(ctx, V, n, C.p, cols, X, n, n, cols, k, 1.0, 0.0));
        IPT_STCHK(gemm_nn_small(ctx, HV, n, C.p, cols, HX, n, n, cols, k, 1.0, 0.0));
    }
    double *sq = gw.p;  // reuse (cols >= k)
    IPT_STCHK(column_sqnorms(ctx, X, n, n, k, sq));
    IPT_STCHK(scale_columns_inv_sqrt(ctx, X, HX, n, n, k, sq));
    if (opt.stats) {
        DevBuf<double> XX;
        IPT_STCHK(XX.alloc((size_t)k * k));
        IPT_STCHK(gemm_tn_tall(ctx, X, n, X, n, XX.p, k, n, k, k));
        IPT_STCHK(identity_deviation(ctx, XX.p, k, &inf.ritz_fro, &inf.ritz_max, &inf.ritz_bad));
    }
    return IPT_CUDA_SUCCESS;
}


// Project the block T (n x t) against the first `cols` columns of V (two
// classical Gram-Schmidt passes) and orthonormalise it with a Cholesky QR on
// the device; directions whose norm after projection falls below 1e-10 of
// their norm before are dropped (the pivot test of the factorisation).
// Returns the kept count in *kept; kept columns are compacted to the front
// of T.  IPT/src/ipt_cuda.cu:2711-2780 (per vector).
int project_orthonormalize(Context *ctx, const double *V, int n, int cols, double *T, int t, int *kept,
                           double *work /* n x t */)
{
    *kept = 0;
    if (t <= 0) {
        return IPT_CUDA_SUCCESS;
    }
    double *C = nullptr;
    const size_t cwords = (size_t)std::max(cols, t) * (size_t)t;
    const size_t tt = (size_t)t * t;
    // C | sq (t) | W (t x t) | Wc (t x t) | flags (t + 1 ints) | keep list (t ints)
    const size_t words = cwords + (size_t)t + 2 * tt + (size_t)t + 1;
    IPT_STCHK(context_scratch(ctx, SLOT_ORTHO, words * sizeof(double), (void **)&C));
    double *sq = C + cwords;
    double *W = sq + t;
    double *Wc = W + tt;
    int *flags = reinterpret_cast<int *>(Wc + tt);
    int *list = flags + t + 1;
    IPT_STCHK(column_sqnorms(ctx, T, n, n, t, sq));
    if (cols > 0) {
        for (int pass = 0; pass < 2; ++pass) {
            IPT_STCHK(gemm_tn_tall(ctx, V, n, T, n, C, cols, n, cols, t));
            IPT_STCHK(gemm_nn_small(ctx, V, n, C, cols, T, n, n, cols, t, -1.0, 1.0));
        }
    }
    // Gram of the projected block; Cholesky with rejection, W = L^-T
    IPT_STCHK(gemm_tn_tall(ctx, T, n, T, n, C, t, n, t, t));
    IPT_STCHK(symmetrize(ctx, C, t, t));
    IPT_STCHK(chol_reject_device(ctx, C, t, sq, 1e-10, W, flags, flags + t));
    std::vector<int> hflags((size_t)t + 1);
    IPT_STCHK(read_host(ctx, hflags.data(), flags, hflags.size() * sizeof(int)));
    std::vector<int> keep;
    for (int j = 0; j < t; ++j) {
        if (hflags[(size_t)j]) {
            keep.push_back(j);
        }
    }
    const int kc = (int)keep.size();
    *kept = kc;
    if (kc == 0) {
        return IPT_CUDA_SUCCESS;
    }
    // T <- T W(:, keep)
    IPT_HIPCHK(hipMemcpyAsync(list, keep.data(), (size_t)kc * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    IPT_STCHK(gather_columns(ctx, W, t, list, kc, Wc));
    IPT_STCHK(gemm_nn_small(ctx, T, n, Wc, t, work, n, n, t, kc, 1.0, 0.0));
    IPT_HIPCHK(hipMemcpyAsync(T, work, (size_t)n * kc * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return IPT_CUDA_SUCCESS;
}

struct DavidsonStats {
    int attempted = 0;
    int accepted_steps = 0;
    int rejected_steps = 0;
    int restarts = 0;
    double residual_before = NAN;
    double residual_after = NAN;
    int target_index = -1;
    double clip = NAN;
    long long matvecs = 0;
    int basis_cols = 0;
    int mlp_accepted = 0;  // learned correction steps kept
    int mlp_rejected = 0;  // learned steps redone with the diagonal correction
    // lowest-largest-residual state seen (IPT/src/ipt_cuda.cu:485-701): the
    // initial Ritz pairs or an accepted step; returned when a later accepted
    // step raised the largest residual
    int best_update_count = 0;
    int best_updated = 0;
    int best_step = -1;
    const char *best_source = "none";
    int best_basis_cols = 0;
    double best_max = NAN;
    int best_max_index = -1;
    int returned_best = 0;
    double pair28_best = NAN;
    double pair29_best = NAN;
};

double max_of(const std::vector<double> &v, int *idx)
{
    double m = NAN;
    int at = -1;
    for (size_t i = 0; i < v.size(); ++i) {
        if (!std::isfinite(m) || v[i] > m) {
            m = v[i];
            at = (int)i;
        }
    }
    if (idx != nullptr) {
        *idx = at;
    }
    return m;
}

// Residual-only Davidson refinement, IPT/src/ipt_cuda.cu:3318-3982 (HEAD):
// corrections for every pair above conv_tol, thick restart to the k Ritz
// vectors plus the last correction block, trial accepted when finite.
// With `mlp` (device parameters; rowabs = off-diagonal row sums of |H|)
// the corrections come from the learned preconditioner; a trial step that
// does not lower the largest residual is redone with the diagonal one, which
// also serves the next two steps.
// V/HV: n x cap buffers holding the search space (first *cols columns).
int davidson(Context *ctx, const ApplyFn &apply, const double *d, const double *rowabs, const float *mlp, int n, int k,
             double normH, double *V, double *HV, int *cols_io, int cap, double *theta, double *X, double *HX,
             const BlockParams &p, DavidsonStats *st, const RROptions &rro, RRInfo *rr_info)
{
    RRInfo best_rr = *rr_info, trial_rr;
    const int steps = std::max(1, std::max(0, p.davidson_steps) + std::max(0, p.davidson_extra));
    int kmax = davidson_max_basis(p, n, k);
    if (kmax > cap) {
        kmax = cap;
    }
    const int retain = p.retain >= 0 ? std::min(p.retain, std::max(0, kmax - k - 1)) : std::max(0, kmax - k - 1);
    int cols = *cols_io;
    st->attempted = 1;

    DevBuf<double> R, T, HT, work, tX, tHX, ttheta, res, recent, hrecent;
    DevBuf<int> idx;
    IPT_STCHK(R.alloc((size_t)n * k));
    IPT_STCHK(T.alloc((size_t)n * k));
    IPT_STCHK(HT.alloc((size_t)n * k));
    IPT_STCHK(work.alloc((size_t)n * std::max(k, retain + 1)));
    IPT_STCHK(tX.alloc((size_t)n * k));
    IPT_STCHK(tHX.alloc((size_t)n * k));
    IPT_STCHK(ttheta.alloc((size_t)cap));
    IPT_STCHK(res.alloc((size_t)k));
    IPT_STCHK(recent.alloc((size_t)n * k));
    IPT_STCHK(hrecent.alloc((size_t)n * k));
    IPT_STCHK(idx.alloc((size_t)k));
    int recent_count = 0;
    int diag_hold = 0;  // steps left on the diagonal correction after a rejected learned step

    std::vector<double> htheta((size_t)k), hres((size_t)k);
    IPT_STCHK(residual_norms(ctx, X, HX, n, n, k, theta, normH, res.p, R.p));
    IPT_STCHK(read_host_many(ctx, {HostRead{htheta.data(), theta, (size_t)k * sizeof(double)},
                                   HostRead{hres.data(), res.p, (size_t)k * sizeof(double)}}));
    {
        double scale = std::max(1.0, std::fabs(normH));
        for (double th : htheta) {
            if (std::isfinite(th)) {
                scale = std::max(scale, std::fabs(th));
            }
        }
        st->clip = std::pow(DBL_EPSILON, 2.0 / 3.0) * scale;
        if (!std::isfinite(st->clip) || st->clip <= 0.0) {
            st->clip = std::pow(DBL_EPSILON, 2.0 / 3.0);
        }
    }
    st->residual_before = max_of(hres, &st->target_index);

    // best-so-far bookkeeping: the best state is either the current one
    // (best_current) or a copy kept in bX / bHX / btheta
    DevBuf<double> bX, bHX, btheta;
    bool best_current = false;
    auto note_pairs = [&](const std::vector<double> &r) {
        // IPT/src/ipt_cuda.cu:538-555 (the two pairs the reference tracks)
        if (r.size() > 28 && std::isfinite(r[28]) && !(r[28] >= st->pair28_best)) {
            st->pair28_best = r[28];
        }
        if (r.size() > 29 && std::isfinite(r[29]) && !(r[29] >= st->pair29_best)) {
            st->pair29_best = r[29];
        }
    };
    // returns true when the state with residuals r becomes the best
    auto note_best = [&](const std::vector<double> &r, const char *source, int step, int basis) -> bool {
        note_pairs(r);
        int at = -1;
        const double m = max_of(r, &at);
        if (!std::isfinite(m)) {
            return false;
        }
        const bool has = st->best_update_count > 0;
        if (has && !(m < st->best_max)) {
            return false;
        }
        if (has) {
            st->best_updated = 1;
        }
        ++st->best_update_count;
        st->best_step = step;
        st->best_source = source;
        st->best_basis_cols = basis;
        st->best_max = m;
        st->best_max_index = at;
        return true;
    };
    best_current = note_best(hres, "davidson_initial", 0, cols);

    auto restart = [&]() -> int {
        // basis = k Ritz vectors + up to `retain` previous corrections
        IPT_HIPCHK(hipMemcpyAsync(V, X, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        IPT_HIPCHK(hipMemcpyAsync(HV, HX, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        int nc = k;
        const int lim = std::min(retain, recent_count);
        if (lim > 0) {
            // project recent against X, orthonormalise; H*recent follows the
            // same linear combination: recompute H on the kept block.
            IPT_HIPCHK(hipMemcpyAsync(T.p, recent.p, (size_t)n * lim * sizeof(double), hipMemcpyDeviceToDevice,
                                      ctx->stream));
            int kept = 0;
            IPT_STCHK(project_orthonormalize(ctx, V, n, nc, T.p, lim, &kept, work.p));
            if (kept > 0) {
                IPT_STCHK(apply(T.p, HT.p, kept));
                st->matvecs += kept;
                IPT_HIPCHK(hipMemcpyAsync(V + (size_t)nc * n, T.p, (size_t)n * kept * sizeof(double),
                                          hipMemcpyDeviceToDevice, ctx->stream));
                IPT_HIPCHK(hipMemcpyAsync(HV + (size_t)nc * n, HT.p, (size_t)n * kept * sizeof(double),
                                          hipMemcpyDeviceToDevice, ctx->stream));
                nc += kept;
            }
        }
        cols = nc;
        ++st->restarts;
        return IPT_CUDA_SUCCESS;
    };

    for (int step = 1; step <= steps; ++step) {
        std::vector<int> cand;
        for (int j = 0; j < k; ++j) {
            if (!std::isfinite(hres[(size_t)j]) || hres[(size_t)j] > p.conv_tol) {
                cand.push_back(j);
            }
        }
        if (cand.empty()) {
            break;
        }
        if (cols > k && cols + (int)cand.size() > kmax) {
            IPT_STCHK(restart());
        }
        const int round_limit = std::min(k, std::max(0, kmax - cols));
        if ((int)cand.size() > round_limit) {
            cand.resize((size_t)round_limit);
        }
        if (cand.empty()) {
            ++st->rejected_steps;
            break;
        }
        const int nt = (int)cand.size();
        IPT_HIPCHK(hipMemcpyAsync(idx.p, cand.data(), (size_t)nt * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        // one trial step: corrections (learned or diagonal), projection,
        // products, Rayleigh-Ritz on the enlarged basis, trial residuals
        std::vector<double> tres((size_t)k);
        int kept = 0;
        auto trial = [&](bool learned) -> int {
            if (learned) {
                IPT_STCHK(mlp_correction(ctx, mlp, R.p, n, d, rowabs, theta, idx.p, n, nt, T.p, n));
            } else {
                IPT_STCHK(diag_correction(ctx, R.p, n, d, theta, idx.p, nt, st->clip, T.p, n, n));
            }
            kept = 0;
            IPT_STCHK(project_orthonormalize(ctx, V, n, cols, T.p, nt, &kept, work.p));
            if (kept == 0) {
                return IPT_CUDA_SUCCESS;
            }
            IPT_STCHK(apply(T.p, HT.p, kept));
            st->matvecs += kept;
            IPT_HIPCHK(hipMemcpyAsync(V + (size_t)cols * n, T.p, (size_t)n * kept * sizeof(double),
                                      hipMemcpyDeviceToDevice, ctx->stream));
            IPT_HIPCHK(hipMemcpyAsync(HV + (size_t)cols * n, HT.p, (size_t)n * kept * sizeof(double),
                                      hipMemcpyDeviceToDevice, ctx->stream));
            int rs = rayleigh_ritz(ctx, V, HV, n, cols + kept, k, ttheta.p, tX.p, tHX.p, rro, &trial_rr);
            if (rs == kRankDeficient) {
                fprintf(stderr, "IPT Davidson: search space of rank below k\n");
                rs = IPT_CUDA_CUSOLVER_ERROR;
            }
            IPT_STCHK(rs);
            // the current residual block R is still needed if the trial is
            // redone, so the trial residual vectors go to `work` first
            IPT_STCHK(residual_norms(ctx, tX.p, tHX.p, n, n, k, ttheta.p, normH, res.p, work.p));
            IPT_STCHK(read_host(ctx, tres.data(), res.p, (size_t)k * sizeof(double)));
            return IPT_CUDA_SUCCESS;
        };
        const double before = max_of(hres, nullptr);
        bool learned = mlp != nullptr && diag_hold == 0;
        if (diag_hold > 0) {
            --diag_hold;
        }
        IPT_STCHK(trial(learned));
        if (learned) {
            // the diagonal preconditioner is the judge: a learned step that
            // does not lower the largest residual is redone with it, and the
            // next two steps use the diagonal correction directly
            const double after = kept > 0 ? max_of(tres, nullptr) : INFINITY;
            if (!(after < before)) {
                ++st->mlp_rejected;
                learned = false;
                diag_hold = 2;
                IPT_STCHK(trial(false));
            } else {
                ++st->mlp_accepted;
            }
        }
        if (kept == 0) {
            ++st->rejected_steps;
            break;
        }
        const int tcols = cols + kept;
        IPT_HIPCHK(hipMemcpyAsync(R.p, work.p, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice,
                                  ctx->stream));
        bool finite = true;
        for (double r : tres) {
            if (!std::isfinite(r)) {
                finite = false;
            }
        }
        if (!finite) {
            // rejected trial: residual vectors R must be those of the
            // current pairs again
            IPT_STCHK(residual_norms(ctx, X, HX, n, n, k, theta, normH, res.p, R.p));
            ++st->rejected_steps;
            break;
        }
        cols = tcols;
        {
            // an accepted step that does not lower the largest residual keeps
            // the best state aside before it is overwritten
            int tat = -1;
            const double tmax = max_of(tres, &tat);
            const bool improves = std::isfinite(tmax) && (st->best_update_count == 0 || tmax < st->best_max);
            if (!improves && best_current) {
                if (bX.p == nullptr) {
                    IPT_STCHK(bX.alloc((size_t)n * k));
                    IPT_STCHK(bHX.alloc((size_t)n * k));
                    IPT_STCHK(btheta.alloc((size_t)k));
                }
                IPT_HIPCHK(hipMemcpyAsync(bX.p, X, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice,
                                          ctx->stream));
                IPT_HIPCHK(hipMemcpyAsync(bHX.p, HX, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice,
                                          ctx->stream));
                IPT_HIPCHK(hipMemcpyAsync(btheta.p, theta, (size_t)k * sizeof(double), hipMemcpyDeviceToDevice,
                                          ctx->stream));
                best_current = false;
            }
            if (note_best(tres, "davidson", step, tcols)) {
                best_current = true;
                best_rr = trial_rr;
            }
            *rr_info = trial_rr;
        }
        IPT_HIPCHK(hipMemcpyAsync(X, tX.p, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        IPT_HIPCHK(hipMemcpyAsync(HX, tHX.p, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        IPT_HIPCHK(hipMemcpyAsync(theta, ttheta.p, (size_t)tcols * sizeof(double), hipMemcpyDeviceToDevice,
                                  ctx->stream));
        IPT_HIPCHK(hipMemcpyAsync(recent.p, T.p, (size_t)n * kept * sizeof(double), hipMemcpyDeviceToDevice,
                                  ctx->stream));
        recent_count = kept;
        hres = tres;
        ++st->accepted_steps;
        if (p.verbose) {
            int conv = 0;
            for (double r : hres) {
                conv += r <= p.conv_tol;
            }
            fprintf(stderr, "IPT Davidson: step %d corrections %d kept %d basis %d, %d/%d pairs <= %.1e, largest %.3e\n",
                    step, nt, kept, cols, conv, k, p.conv_tol, max_of(hres, nullptr));
        }
        if (cols >= kmax) {
            IPT_STCHK(restart());
        }
    }
    int mi = -1;
    st->residual_after = max_of(hres, &mi);
    st->basis_cols = cols;
    if (!best_current && st->best_update_count > 0) {
        // IPT/src/ipt_cuda.cu:4941-4983: the best state is returned
        IPT_HIPCHK(hipMemcpyAsync(X, bX.p, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        IPT_HIPCHK(hipMemcpyAsync(HX, bHX.p, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice,
                                  ctx->stream));
        IPT_HIPCHK(hipMemcpyAsync(theta, btheta.p, (size_t)k * sizeof(double), hipMemcpyDeviceToDevice,
                                  ctx->stream));
        IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
        st->returned_best = 1;
        st->residual_after = st->best_max;
        st->basis_cols = st->best_basis_cols;
        *rr_info = best_rr;
    }
    *cols_io = cols;
    return IPT_CUDA_SUCCESS;
}

}  // namespace

struct BlockResult {
    DevBuf<double> theta;  // >= k
    DevBuf<double> X;      // n x k (sorted basis)
    DevBuf<double> HX;     // n x k
    int iterations = 0;
    double fp_max = NAN;
    int basis_cols = 0;
    long long matvecs = 0;
    double iteration_time = 0.0;
    double rr_time = 0.0;
    double davidson_time = 0.0;
    long long fp64_products = 0;  // H * block products in FP64
    long long lowp_products = 0;  // low-precision correction products
    int widened = 0;              // final basis widened by unit vectors (rank below k)
    int stalled = 0;              // clusters stopped by stall detection
    RRInfo rr;                    // the Rayleigh-Ritz of the returned pairs
    DavidsonStats dav;
};

void record_ranges(Context *ctx, const std::vector<std::pair<int, int>> &ranges)
{
    ctx->last_ranges.clear();
    ctx->last_cluster_state.clear();
    for (const auto &r : ranges) {
        ctx->last_ranges.push_back(r.first);
        ctx->last_ranges.push_back(r.second);
    }
}

void record_stats(Context *ctx, const BlockResult &br)
{
    ctx->last_basis_cols = br.basis_cols;
    ctx->last_steps = br.iterations;
    ctx->last_fp64_products = br.fp64_products;
    ctx->last_lowp_products = br.lowp_products;
    ctx->last_davidson_steps = br.dav.accepted_steps;
    ctx->last_davidson_rejected = br.dav.rejected_steps;
    ctx->last_mlp_accepted = br.dav.mlp_accepted;
    ctx->last_mlp_rejected = br.dav.mlp_rejected;
    ctx->last_davidson_matvecs = br.dav.matvecs;
    ctx->last_davidson_residual = br.dav.residual_after;
}

// Block-cluster IPT on a prepared (sorted-diagonal) operator.  With a
// low-precision operator the iteration alternates `anchor_every` steps whose
// product is the anchored correction (see ipt_gemm_lp.hip) with one exact
// FP64 step; clusters are frozen only on exact steps, so the stopping rule
// is the FP64 one.  `identity_product` (optional) fills H * identity blocks
// without a product (dense: gathered columns of H).
int block_pipeline(Context *ctx, const ApplyFn &apply, const ApplyFn &identity_product, const DeltaFn *lowp,
                   const double *d, const double *rowabs, int n, int k,
                   const std::vector<std::pair<int, int>> &ranges, double normH, const BlockParams &p,
                   BlockResult *out, bool sparse_products = false, const LowpOperator *dense_op = nullptr,
                   const std::function<int()> &before_davidson = std::function<int()>())
{
    BlockPlan plan;
    // with stall detection on, clusters wider than the block limit are not
    // an error: they are never iterated and Davidson refines their pairs
    trace_mark("pipeline start");
    IPT_STCHK(block_plan_create(ctx, ranges, n, &plan, p.stall_steps > 0 ? p.max_block : 0));
    trace_mark("plan created");
    const int bc = plan.basis_cols;
    if (bc < k) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    // room for k identity columns should the final basis be rank deficient
    int cap = std::min(bc + k, std::max(bc, n));
    plan.stall_steps = p.stall_steps;
    if (p.davidson) {
        const int kmax = davidson_max_basis(p, n, k);
        cap = std::max(bc, kmax) + k + 1;
        cap = std::min(std::max(cap, bc + k), std::max(bc, n));
    }
    const bool mixed = lowp != nullptr;
    DevBuf<double> Ba, Bb, Y, Banc, Yanc;
    IPT_STCHK(Ba.alloc((size_t)n * cap));
    IPT_STCHK(Bb.alloc((size_t)n * bc));
    IPT_STCHK(Y.alloc((size_t)n * cap));
    IPT_STCHK(block_identity(ctx, plan, Ba.p, n, n));
    if (mixed) {
        IPT_STCHK(Banc.alloc((size_t)n * bc));
        IPT_STCHK(Yanc.alloc((size_t)n * bc));
        IPT_HIPCHK(hipMemcpyAsync(Banc.p, Ba.p, (size_t)n * bc * sizeof(double), hipMemcpyDeviceToDevice,
                                  ctx->stream));
        if (identity_product) {
            IPT_STCHK(identity_product(Banc.p, Yanc.p, bc));
        } else {
            IPT_STCHK(apply(Banc.p, Yanc.p, bc));
            out->fp64_products += 1;
        }
    }
    // Operand emission (dense FP16 / BF16 / BF16x3 correction products):
    // every update writes the operand of the next low-precision product
    // (ipt_emit.h), so a low-precision step is the product, its epilogue
    // and the update.  The wide-cluster kernel does not emit.
    const bool fused = mixed && p.emit && dense_op != nullptr && plan.wide_host.empty() &&
                       (dense_op->mode == LOWP_FP16 || dense_op->mode == LOWP_BF16 ||
                        dense_op->mode == LOWP_BF16X3);
    const int nb_all = (bc + 15) / 16;
    DevBuf<unsigned short> zhi, zlo;
    DevBuf<double> colscale;
    if (fused) {
        const size_t zel = (size_t)dense_op->nkb * nb_all * 512;  // 16-bit elements per plane
        IPT_STCHK(zhi.alloc(zel));
        IPT_HIPCHK(hipMemsetAsync(zhi.p, 0, zel * sizeof(unsigned short), ctx->stream));
        if (dense_op->mode == LOWP_BF16X3) {
            IPT_STCHK(zlo.alloc(zel));
            IPT_HIPCHK(hipMemsetAsync(zlo.p, 0, zel * sizeof(unsigned short), ctx->stream));
        }
        IPT_STCHK(colscale.alloc((size_t)nb_all * 16));
        IPT_STCHK(fill_device(ctx, colscale.p, (long)nb_all * 16, 1.0));
        IPT_STCHK(plan.cmax.alloc((size_t)2 * plan.nclusters));
        IPT_HIPCHK(hipMemsetAsync(plan.cmax.p, 0, (size_t)2 * plan.nclusters * sizeof(double), ctx->stream));
    }
    plan.frozen.assign((size_t)plan.nclusters, 0);
    trace_mark("iteration set up");
    double *cur = Ba.p, *nxt = Bb.p;
    auto t0 = std::chrono::steady_clock::now();
    // Exact steps are scheduled: at most anchor_every low-precision steps
    // apart, earlier when the per-cluster convergence rate measured between
    // the last two exact steps predicts that every active cluster reaches
    // tol sooner (only exact steps may freeze a cluster, so steps after
    // convergence would otherwise be wasted).  The last step is always
    // exact, so the returned basis comes out of an exact step.
    // CSR products cost in proportion to their width and an exact one
    // about twice a low-precision one: short spacing, so clusters freeze
    // (and drop out of the products) soon after they converge
    const int anchor_every = p.anchor_every > 0 ? p.anchor_every : (sparse_products ? 5 : 20);
    int next_exact = anchor_every;
    int t_prev = -1;
    std::vector<double> fp_prev((size_t)plan.nclusters, NAN), fp_now((size_t)plan.nclusters);
    std::vector<int> act((size_t)plan.nclusters, 1);
    // Products cover the columns of the clusters still active (as of the
    // last exact step): a frozen cluster's update copies its columns and
    // reads nothing of H B, so its columns of Y (and of the anchor Y_a) may
    // go stale.  A dense low-precision product streams all of E whatever
    // its width, so it takes the hull of the active columns in one call.
    // The other products take one call per run of adjacent active clusters
    // when that is cheaper, by a cost in columns: a dense FP64 product is
    // bound by streaming H below about 24 columns (cost max(24, w)); a CSR
    // product streams the matrix once per call, worth about 12 columns of
    // gathers (cost 12 + w).
    const int kFloorCols = sparse_products ? 0 : 24;
    const int kFixedCols = sparse_products ? 12 : 0;
    const bool lowp_hull = !sparse_products;
    std::vector<std::pair<int, int>> runs;
    int run_cols = 0, hull_lo = 0, hull_w = bc;
    bool per_run = false;
    auto rebuild_runs = [&]() {
        runs.clear();
        run_cols = 0;
        long run_cost = 0;
        for (int c = 0; c < plan.nclusters; ++c) {
            if (!act[(size_t)c]) {
                continue;
            }
            const int o = plan.host[(size_t)c].off, m = plan.host[(size_t)c].m;
            if (!runs.empty() && runs.back().first + runs.back().second == o) {
                runs.back().second += m;
            } else {
                runs.emplace_back(o, m);
            }
            run_cols += m;
        }
        if (runs.empty()) {
            hull_w = 0;
            return;
        }
        for (const auto &r : runs) {
            run_cost += kFixedCols + std::max(kFloorCols, r.second);
        }
        hull_lo = runs.front().first;
        hull_w = runs.back().first + runs.back().second - hull_lo;
        per_run = run_cost < kFixedCols + std::max(kFloorCols, hull_w);
    };
    if (!plan.skipped.empty()) {
        // skipped clusters are in no update launch: both iterate buffers
        // hold their identity blocks
        IPT_HIPCHK(hipMemcpyAsync(Bb.p, Ba.p, (size_t)n * bc * sizeof(double), hipMemcpyDeviceToDevice,
                                  ctx->stream));
        IPT_STCHK(block_plan_mark_skipped(ctx, &plan));
        for (int c : plan.skipped) {
            act[(size_t)c] = 0;
        }
        if (p.verbose) {
            fprintf(stderr, "IPT block-cluster: %zu clusters wider than %d left to Rayleigh-Ritz and Davidson\n",
                    plan.skipped.size(), p.max_block);
        }
    }
    rebuild_runs();
    // Status of the exact steps: the finalisation's status block (status
    // bits, active flags, fixed-point residuals).  Lagged (IPT_BLOCK_LAGGED_
    // STATUS=1; the default when every step is exact, the mixed path's
    // exact steps being few): it is queued to pinned memory and read after
    // the next step has been enqueued, so the GPU does not wait for the
    // host's round trip.  The host's view (products, frozen launches, the
    // next exact step, the stop) is then one step behind: a cluster found
    // inactive has its columns in both iterate buffers after that next step
    // (its update copies them), and the loop may run one step past the last
    // cluster's convergence, which changes no iterate.
    const size_t nstat = (size_t)2 + 2 * plan.nclusters;
    const bool lagged = nstat <= 4096 && ctx->host_status != nullptr &&
                        env_int("IPT_BLOCK_LAGGED_STATUS", mixed ? 0 : 1) != 0;
    struct EventHolder {
        hipEvent_t e = nullptr;
        ~EventHolder()
        {
            if (e != nullptr) {
                (void)hipEventDestroy(e);
            }
        }
    } stat_ev;
    if (lagged) {
        IPT_HIPCHK(hipEventCreateWithFlags(&stat_ev.e, hipEventDisableTiming));
    }
    int pending = -1;  // exact step whose status is queued and not yet read
    bool done = false;
    // Host side of an exact step's status (sit: its step): frozen clusters,
    // the products' runs, the next exact step, errors and the stop.
    // copy_frozen: copy newly frozen clusters to the other iterate buffer
    // (not needed when the following step has already been enqueued).
    auto process_status = [&](int sit, const double *stv, bool copy_frozen) -> int {
        int flags[3];
        flags[0] = 0;
        flags[1] = (int)stv[0];
        flags[2] = (int)stv[1];
        for (int c = 0; c < plan.nclusters; ++c) {
            act[(size_t)c] = stv[(size_t)2 + c] != 0.0 ? 1 : 0;
            fp_now[(size_t)c] = stv[(size_t)2 + plan.nclusters + c];
        }
        // After an exact step an inactive cluster is frozen for good: its
        // columns are in both iterate buffers, and launches whose clusters
        // are all frozen are skipped from here on.
        for (int c = 0; c < plan.nclusters; ++c) {
            if (!act[(size_t)c] && !plan.frozen[(size_t)c]) {
                if (copy_frozen) {
                    const size_t o = (size_t)plan.host[(size_t)c].off * n;
                    IPT_HIPCHK(hipMemcpyAsync(nxt + o, cur + o, (size_t)n * plan.host[(size_t)c].m * sizeof(double),
                                              hipMemcpyDeviceToDevice, ctx->stream));
                }
                plan.frozen[(size_t)c] = 1;
            }
        }
        rebuild_runs();
        if (mixed) {
            int steps = anchor_every + 1;
            if (t_prev >= 0) {
                // Dense input: the low-precision product costs the same
                // whatever the active width, so the next exact step waits
                // for the slowest cluster.  CSR input: products cost in
                // proportion to the active width, so the next exact step
                // comes when the first cluster is predicted to converge,
                // which freezes it and narrows the following products.
                double need = 0.0, need_first = INFINITY;
                bool known = true;
                for (int c = 0; c < plan.nclusters; ++c) {
                    if (!act[(size_t)c]) {
                        continue;
                    }
                    const double f1 = fp_now[(size_t)c], f0 = fp_prev[(size_t)c];
                    if (!(f1 > 0.0) || !(f0 > f1) || !std::isfinite(f0)) {
                        known = false;
                        continue;
                    }
                    const double rate = std::log(f1 / f0) / (double)(sit - t_prev);  // < 0
                    const double nc = std::log(p.tol / f1) / rate;
                    need = std::max(need, nc);
                    need_first = std::min(need_first, nc);
                }
                // CSR: 10% margin, a late exact step costing less than an
                // extra one.  Dense: none; by the time one cluster is left
                // the exact product covers only its columns and costs about
                // two low-precision steps, so an early exact step costs no
                // more than a late one.
                if (sparse_products && std::isfinite(need_first)) {
                    steps = std::min(steps, std::max(2, (int)std::ceil(1.1 * need_first) + 1));
                } else if (!sparse_products && known) {
                    steps = std::min(steps, std::max(2, (int)std::ceil(need)));
                }
            }
            fp_prev = fp_now;
            t_prev = sit;
            next_exact = sit + steps;
            if (p.verbose) {
                double worst = 0.0;
                int worst_c = -1;
                for (int c = 0; c < plan.nclusters; ++c) {
                    if (act[(size_t)c] && !(fp_now[(size_t)c] <= worst)) {
                        worst = fp_now[(size_t)c];
                        worst_c = c;
                    }
                }
                fprintf(stderr, "IPT block-cluster mixed: exact step %d, %d clusters active, largest fp %.3e "
                        "(cluster %d), next exact step %d\n", sit, flags[2], worst, worst_c, next_exact);
            }
        }
        if (flags[1] != 0) {
            fprintf(stderr,
                    "IPT block-cluster: block solve failed (%s) at iteration %d; identity fallback is disabled\n",
                    (flags[1] & 2) ? "singular" : "non-finite norm", sit + 1);
            return IPT_CUDA_INVALID_ARGUMENT;
        }
        done = flags[2] == 0;
        return IPT_CUDA_SUCCESS;
    };
    for (int it = 0; it < p.maxiter && !runs.empty(); ++it) {
        const bool exact = !mixed || it == next_exact || it == p.maxiter - 1;
        const size_t ho = (size_t)hull_lo * n;
        if (mixed && exact) {
            // clusters stopped by a low-precision step since the last exact
            // one take this exact step (they are inside its product: they
            // were active at the last exact step) and then freeze
            IPT_STCHK(block_reactivate_tentative(ctx, plan));
        }
        if (exact && per_run) {
            for (const auto &r : runs) {
                const size_t o = (size_t)r.first * n;
                IPT_STCHK(apply(cur + o, Y.p + o, r.second));
            }
        } else if (exact) {
            IPT_STCHK(apply(cur + ho, Y.p + ho, hull_w));
        } else if (fused) {
            IPT_STCHK(lowp_apply_emitted(ctx, *dense_op, zhi.p, zlo.p, nb_all, hull_lo, hull_w, colscale.p, cur,
                                         Banc.p, Yanc.p, n, d, bc, Y.p));
        } else if (lowp_hull || !per_run) {
            IPT_STCHK((*lowp)(cur + ho, Banc.p + ho, Yanc.p + ho, hull_w, Y.p + ho));
        } else {
            for (const auto &r : runs) {
                const size_t o = (size_t)r.first * n;
                IPT_STCHK((*lowp)(cur + o, Banc.p + o, Yanc.p + o, r.second, Y.p + o));
            }
        }
        if (exact) {
            out->fp64_products += 1;
            out->matvecs += per_run ? run_cols : hull_w;
        } else {
            out->lowp_products += 1;
        }
        // a low-precision step stops clusters that reach tol there; the next
        // exact step confirms them (block_update, k_block_finalize)
        LowpEmit em;
        if (fused) {
            em.zhi = zhi.p;
            em.zlo = zlo.p;
            em.mode = dense_op->mode;
            em.nb_all = nb_all;
            // an exact step's pair becomes the anchor of the next product
            em.anc = exact ? nullptr : Banc.p;
            em.cmax = plan.cmax.p;
            em.fp = plan.fp.p;
            em.colscale = colscale.p;
        }
        IPT_STCHK(block_update(ctx, plan, cur, Y.p, nxt, n, d, n, p.damping, p.tol, 1, exact ? 0 : 1,
                               fused ? &em : nullptr));
        if (mixed && exact) {
            // the exact pair (B, H B) becomes the anchor of the next steps
            IPT_HIPCHK(hipMemcpyAsync(Banc.p, cur, (size_t)n * bc * sizeof(double), hipMemcpyDeviceToDevice,
                                      ctx->stream));
            IPT_HIPCHK(hipMemcpyAsync(Yanc.p, Y.p, (size_t)n * bc * sizeof(double), hipMemcpyDeviceToDevice,
                                      ctx->stream));
        }
        std::swap(cur, nxt);
        if (pending >= 0) {
            // the status of the previous exact step, read while this step
            // runs on the GPU
            IPT_HIPCHK(hipEventSynchronize(stat_ev.e));
            trace_mark("exact step status on the host");
            const int sit = pending;
            pending = -1;
            IPT_STCHK(process_status(sit, ctx->host_status, false));
            if (done) {
                break;
            }
        }
        if (!exact) {
            continue;
        }
        if (lagged) {
            IPT_HIPCHK(hipMemcpyAsync(ctx->host_status, plan.status.p, nstat * sizeof(double), hipMemcpyDeviceToHost,
                                      ctx->stream));
            IPT_HIPCHK(hipEventRecord(stat_ev.e, ctx->stream));
            pending = it;
            continue;
        }
        {
            // one read-back per exact step (the finalisation's status block)
            std::vector<double> stv(nstat);
            IPT_STCHK(read_host(ctx, stv.data(), plan.status.p, stv.size() * sizeof(double)));
            trace_mark("exact step status on the host");
            IPT_STCHK(process_status(it, stv.data(), true));
        }
        if (done) {
            break;
        }
    }
    if (pending >= 0) {
        // the last step's status (the loop ended at the step limit)
        IPT_HIPCHK(hipEventSynchronize(stat_ev.e));
        IPT_STCHK(process_status(pending, ctx->host_status, false));
    }
    // per-cluster iterations and residuals; stalled clusters start the
    // Rayleigh-Ritz basis again from their identity blocks
    {
        std::vector<int> its((size_t)plan.nclusters), stl((size_t)plan.nclusters), live((size_t)plan.nclusters);
        std::vector<double> fps((size_t)plan.nclusters);
        IPT_STCHK(read_host_many(ctx, {HostRead{stl.data(), plan.stalled.p, stl.size() * sizeof(int)},
                                       HostRead{live.data(), plan.active.p, live.size() * sizeof(int)},
                                       HostRead{its.data(), plan.iters.p, its.size() * sizeof(int)},
                                       HostRead{fps.data(), plan.fp.p, fps.size() * sizeof(double)}}));
        std::vector<int> reset;
        ctx->last_cluster_state.assign((size_t)plan.nclusters, 0);
        for (int c = 0; c < plan.nclusters; ++c) {
            out->iterations = std::max(out->iterations, its[(size_t)c]);
            if (stl[(size_t)c]) {
                reset.push_back(c);
                ctx->last_cluster_state[(size_t)c] = 2;
                continue;
            }
            if (live[(size_t)c]) {
                ctx->last_cluster_state[(size_t)c] = 1;
            }
            if (std::isfinite(fps[(size_t)c]) && (!std::isfinite(out->fp_max) || fps[(size_t)c] > out->fp_max)) {
                out->fp_max = fps[(size_t)c];
            }
        }
        out->stalled = (int)reset.size();
        if (p.verbose && !reset.empty()) {
            fprintf(stderr, "IPT block-cluster: %zu clusters stalled; their bases restart from identity blocks\n",
                    reset.size());
        }
        IPT_STCHK(block_identity_clusters(ctx, plan, reset, cur, n, n));
    }
    trace_mark("iteration done");
    out->iteration_time = seconds_since(t0);
    // the basis lives in `cur`; Davidson needs it in the big buffer Ba
    if (cur != Ba.p) {
        IPT_HIPCHK(hipMemcpyAsync(Ba.p, cur, (size_t)n * bc * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    }
    Bb.release();
    auto t1 = std::chrono::steady_clock::now();
    IPT_STCHK(apply(Ba.p, Y.p, bc));
    trace_mark("final product enqueued");
    out->matvecs += bc;
    out->fp64_products += 1;
    IPT_STCHK(out->theta.alloc((size_t)cap));
    IPT_STCHK(out->X.alloc((size_t)n * k));
    IPT_STCHK(out->HX.alloc((size_t)n * k));
    int rr_cols = bc;
    const RROptions rro = rr_options_from_env();
    int rs = rayleigh_ritz(ctx, Ba.p, Y.p, n, bc, k, out->theta.p, out->X.p, out->HX.p, rro, &out->rr);
    if (rs == kRankDeficient && bc + k <= cap) {
        // Clusters that drifted onto overlapping subspaces leave a basis of
        // rank below k: the unit vectors of the k target rows widen it, so
        // that Rayleigh-Ritz (and Davidson after it) can still find k pairs.
        if (p.verbose) {
            fprintf(stderr, "IPT block-cluster: basis of rank below k; widened by %d unit vectors\n", k);
        }
        IPT_STCHK(set_identity(ctx, Ba.p + (size_t)bc * n, n, n, k));
        IPT_STCHK(apply(Ba.p + (size_t)bc * n, Y.p + (size_t)bc * n, k));
        out->matvecs += k;
        out->fp64_products += 1;
        out->widened = 1;
        rr_cols = bc + k;
        rs = rayleigh_ritz(ctx, Ba.p, Y.p, n, rr_cols, k, out->theta.p, out->X.p, out->HX.p, rro, &out->rr);
    }
    if (rs == kRankDeficient) {
        fprintf(stderr, "IPT Rayleigh-Ritz failed: basis of rank below k\n");
        rs = IPT_CUDA_CUSOLVER_ERROR;
    }
    IPT_STCHK(rs);
    trace_mark("Rayleigh-Ritz enqueued");
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    trace_mark("Rayleigh-Ritz done");
    out->rr_time = seconds_since(t1);
    out->basis_cols = rr_cols;
    if (p.davidson) {
        auto t2 = std::chrono::steady_clock::now();
        int cols = rr_cols;
        if (out->widened) {
            // the widened basis is far from orthonormal: search from the k
            // Ritz vectors instead (the projections assume orthonormal columns)
            IPT_HIPCHK(hipMemcpyAsync(Ba.p, out->X.p, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice,
                                      ctx->stream));
            IPT_HIPCHK(hipMemcpyAsync(Y.p, out->HX.p, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice,
                                      ctx->stream));
            cols = k;
        }
        DevBuf<float> mlp;
        if (p.precond_mlp && rowabs != nullptr) {
            IPT_STCHK(load_mlp_model(p.mlp_model, &mlp));
        }
        // before_davidson runs once, at Davidson's first product (a solve
        // whose Rayleigh-Ritz pairs have converged makes none)
        bool dav_started = false;
        const ApplyFn dav_apply = [&](const double *X, double *Y, int nc) -> int {
            if (!dav_started && before_davidson) {
                dav_started = true;
                IPT_STCHK(before_davidson());
            }
            return apply(X, Y, nc);
        };
        IPT_STCHK(davidson(ctx, dav_apply, d, rowabs, mlp.p, n, k, normH, Ba.p, Y.p, &cols, cap, out->theta.p, out->X.p,
                           out->HX.p, p, &out->dav, rro, &out->rr));
        out->matvecs += out->dav.matvecs;
        out->basis_cols = out->dav.basis_cols;
        IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
        trace_mark("Davidson done");
        out->davidson_time = seconds_since(t2);
    }
    return IPT_CUDA_SUCCESS;
}

namespace {

BlockParams params_from_env(int maxiter, double tol, int use_tol)
{
    // IPT/src/ipt_cuda.cu:4672-4680, 3330-3340, 4915
    BlockParams p;
    p.maxiter = env_int("IPT_BLOCK_CLUSTER_MAXITER", maxiter > 0 ? maxiter : 100);
    if (p.maxiter <= 0) {
        p.maxiter = 100;
    }
    p.tol = env_double("IPT_BLOCK_CLUSTER_TOL", use_tol ? tol : 1.0e-12);
    p.damping = env_double("IPT_BLOCK_CLUSTER_DAMPING", 1.0);
    p.davidson = env_flag("IPT_DAVIDSON_ENRICH");
    p.davidson_steps = env_int("IPT_DAVIDSON_STEPS", 1);
    p.davidson_extra = env_int("IPT_DAVIDSON_EXTRA_STEPS", 0);
    p.conv_tol = env_double("IPT_DAVIDSON_CONVERGED_TOL", 1.0e-13);
    // additions: precision of the correction products and anchor interval
    const char *prec = getenv("IPT_PRECISION");
    if (prec != nullptr) {
        if (!strcmp(prec, "fp16")) {
            p.lowp_mode = LOWP_FP16;
        } else if (!strcmp(prec, "bf16")) {
            p.lowp_mode = LOWP_BF16;
        } else if (!strcmp(prec, "bf16x3") || !strcmp(prec, "tf32")) {
            p.lowp_mode = LOWP_BF16X3;
        } else if (!strcmp(prec, "fp32")) {
            p.lowp_mode = LOWP_FP32;
        } else if (!strcmp(prec, "fp8")) {
            p.lowp_mode = LOWP_FP8;
        }
    }
    p.anchor_every = std::max(0, env_int("IPT_MIXED_ANCHOR_EVERY", 0));
    p.emit = env_int("IPT_MIXED_EMIT", 1) != 0;
    // additions: stall detection and the Davidson search-space limits
    p.stall_steps = std::max(0, env_int("IPT_BLOCK_CLUSTER_STALL", 0));
    p.max_basis = std::max(0, env_int("IPT_DAVIDSON_MAX_BASIS", 0));
    p.retain = env_int("IPT_DAVIDSON_RETAIN", -1);
    p.max_block = degeneracy_options_from_env(0.0).max_block;
    p.verbose = env_flag("IPT_HIP_VERBOSE");
    // addition: IPT_DAVIDSON_PRECOND=mlp (model file IPT_DAVIDSON_MLP_MODEL)
    const char *pc = getenv("IPT_DAVIDSON_PRECOND");
    if (pc != nullptr && !strcmp(pc, "mlp")) {
        const char *mf = getenv("IPT_DAVIDSON_MLP_MODEL");
        if (mf != nullptr && mf[0] != '\0') {
            p.precond_mlp = 1;
            p.mlp_model = mf;
        } else {
            fprintf(stderr, "IPT: IPT_DAVIDSON_PRECOND=mlp without IPT_DAVIDSON_MLP_MODEL; diagonal corrections used\n");
        }
    }
    return p;
}

void fill_result(IPTCudaResult *r, const BlockResult &b, int n, int k)
{
    r->n = n;
    r->k = k;
    r->iterations = b.iterations;
    r->basis_cols = b.basis_cols;
    r->matvecs = b.matvecs;
    r->fixed_point_residual = b.fp_max;
    r->iteration_time_sec = b.iteration_time;
    r->rayleigh_ritz_time_sec = b.rr_time;
    r->davidson_time_sec = b.davidson_time;
    r->davidson_attempted = b.dav.attempted;
    r->davidson_accepted = b.dav.accepted_steps > 0;
    r->davidson_target_index = b.dav.target_index;
    r->davidson_residual_before = b.dav.residual_before;
    r->davidson_residual_after = b.dav.residual_after;
    r->davidson_denom_clip = b.dav.clip;
    r->davidson_restart_count = b.dav.restarts;
    r->accepted_steps = b.dav.accepted_steps;
    r->rejected_steps = b.dav.rejected_steps;
    // IPT/src/ipt_cuda.cu:3168-3206 (NaN unless IPT_RR_HOST_ORTHOGONALITY_STATS)
    r->rayleigh_ritz_used_qr = b.rr.used_qr;
    r->basis_orthogonality_frobenius_error = b.rr.basis_fro;
    r->basis_orthogonality_max_abs_error = b.rr.basis_max;
    r->basis_has_nan_or_inf = b.rr.basis_bad;
    r->ritz_vectors_orthogonality_frobenius_error = b.rr.ritz_fro;
    r->ritz_vectors_orthogonality_max_abs_error = b.rr.ritz_max;
    r->ritz_vectors_has_nan_or_inf = b.rr.ritz_bad;
    // IPT/src/ipt_cuda.cu:678-701
    const DavidsonStats &d = b.dav;
    const bool has = d.best_update_count > 0;
    r->best_so_far_enabled = 1;
    r->best_so_far_updated = d.best_updated;
    r->best_so_far_update_count = d.best_update_count;
    r->best_so_far_step = has ? d.best_step : -1;
    snprintf(r->best_so_far_source, sizeof(r->best_so_far_source), "%s", has ? d.best_source : "none");
    r->best_so_far_basis_cols = has ? d.best_basis_cols : 0;
    r->best_so_far_max_residual = has ? d.best_max : NAN;
    r->best_so_far_max_residual_index = has ? d.best_max_index : -1;
    r->final_returned_from_best_so_far = d.returned_best;
    r->pair28_best_residual = d.pair28_best;
    r->pair29_best_residual = d.pair29_best;
}

void result_init_block(IPTCudaResult *r)
{
    // IPT/src/ipt_cuda.cu:4745-4771
    memset(r, 0, sizeof(*r));
    r->fixed_point_residual = NAN;
    r->basis_orthogonality_frobenius_error = NAN;
    r->basis_orthogonality_max_abs_error = NAN;
    r->ritz_vectors_orthogonality_frobenius_error = NAN;
    r->ritz_vectors_orthogonality_max_abs_error = NAN;
    r->davidson_target_index = -1;
    r->davidson_residual_before = NAN;
    r->davidson_residual_after = NAN;
    r->davidson_denom_clip = NAN;
    r->adaptive_coupling_tau = NAN;
    r->adaptive_target_block_start = -1;
    r->adaptive_target_block_end = -1;
    r->best_so_far_enabled = 1;
    r->best_so_far_step = -1;
    snprintf(r->best_so_far_source, sizeof(r->best_so_far_source), "none");
    r->best_so_far_max_residual = NAN;
    r->best_so_far_max_residual_index = -1;
    r->pair28_best_residual = NAN;
    r->pair29_best_residual = NAN;
    r->gap28_29 = NAN;
    r->relative_gap28_29 = NAN;
    r->cluster_residual_fro = NAN;
    r->cluster_residual_max = NAN;
}

// Target-count and adaptive-expansion fields, IPT/src/ipt_cuda.cu:4613,
// 4636-4667.
void fill_target_fields(IPTCudaResult *r, const BlockTargets &bt, const std::vector<std::pair<int, int>> &ranges,
                        int k, const std::vector<int> &added, bool limit_hit)
{
    if (r == nullptr) {
        return;
    }
    r->oversample = bt.oversample;
    r->adaptive_block_enabled = bt.adaptive ? 1 : 0;
    r->adaptive_coupling_tau = bt.tau;
    r->adaptive_limit_hit = limit_hit ? 1 : 0;
    std::string text;
    for (int i : added) {
        if (!text.empty()) {
            text += ";";
        }
        text += std::to_string(i);
    }
    snprintf(r->adaptive_added_indices, sizeof(r->adaptive_added_indices), "%s",
             text.empty() ? "none" : text.c_str());
    for (const auto &rg : ranges) {
        if (rg.first <= k - 1 && k - 1 <= rg.second) {
            r->adaptive_target_block_start = rg.first;
            r->adaptive_target_block_end = rg.second;
        }
    }
}

// Dense matrix on the device: GPU-side preparation (diagonal, symmetry,
// off-diagonal strengths, infinity norm), cluster detection on the host
// from the few coupling values it needs, permutation to sorted order when
// the diagonal is not ascending.
struct DensePrep {
    std::vector<int> perm;              // perm[sorted] = original
    bool permuted = false;
    DevBuf<double> Hs;                  // permuted copy when needed
    const double *H = nullptr;          // matrix used by the solver
    DevBuf<double> d;                   // sorted diagonal (device)
    DevBuf<double> rowabs;              // sorted off-diagonal row sums of |H|
    DevBuf<int> dperm;
    std::vector<double> hd;             // sorted diagonal (host)
    double normH = 0.0;
    double max_offdiag = 0.0;
    std::vector<std::pair<int, int>> ranges;
};

struct WindowCoupling : CouplingSource {
    int nn;
    const std::vector<double> *dg;
    const std::vector<double> *off;
    std::vector<std::pair<long, double>> lut;  // key row*n+col -> value
    int n() const override { return nn; }
    double diag(int i) const override { return (*dg)[(size_t)i]; }
    double value(int row, int col) const override
    {
        const long key = (long)row * nn + col;
        auto it = std::lower_bound(lut.begin(), lut.end(), key,
                                   [](const std::pair<long, double> &e, long kk) { return e.first < kk; });
        if (it != lut.end() && it->first == key) {
            return it->second;
        }
        return 0.0;
    }
    void offdiag_strengths(std::vector<double> *s) const override { *s = *off; }
};

// Fetches H(r, c) and H(c, r) for the pairs rc[2t], rc[2t+1] (device
// arrays) into out[2t], out[2t+1].
using GatherFn = std::function<int(const int *rc, int npairs, double *out)>;

// Cluster detection on the sorted diagonal `sd` and the per-row largest
// off-diagonal magnitudes `soff`: the (row, target) pairs inside each
// target's detection window are fetched with `gather`, then union-find
// (IPT/src/ipt_preparation.cu:501-645).
int detect_clusters(Context *ctx, const std::vector<double> &sd, const std::vector<double> &soff, int n, int k,
                    const GatherFn &gather, std::vector<std::pair<int, int>> *ranges,
                    const std::function<int()> &overlap = std::function<int()>())
{
    // candidate (row, target) pairs of the detection windows
    const DegeneracyOptions o = degeneracy_options_from_env(env_double("IPT_DEGENERACY_THRESHOLD", 0.0));
    const double eps = DBL_EPSILON;
    double maxabs = 0.0;
    for (double x : sd) {
        maxabs = std::max(maxabs, std::fabs(x));
    }
    const double tau_num = o.eps_multiplier * eps * std::max(1.0, maxabs);
    const double ratio_tol = std::max(o.ratio_tol, eps);
    std::vector<int> rc;
    for (int col = 0; col < k; ++col) {
        const double dc = sd[(size_t)col];
        const double scale = std::max(1.0, std::fabs(dc));
        const double fl = std::max(std::max(o.absolute_gap, tau_num), o.rel_gap_tol * scale);
        const double tau_c = std::max(fl, soff[(size_t)col] / ratio_tol);
        const int lo = (int)(std::lower_bound(sd.begin(), sd.end(), dc - tau_c) - sd.begin());
        const int hi = (int)(std::upper_bound(sd.begin(), sd.end(), dc + tau_c) - sd.begin());
        for (int r = lo; r < hi; ++r) {
            if (r != col) {
                rc.push_back(r);
                rc.push_back(col);
            }
        }
    }
    trace_mark("detection pairs listed");
    WindowCoupling wc;
    wc.nn = n;
    wc.dg = &sd;
    wc.off = &soff;
    const int npairs = (int)(rc.size() / 2);
    // the entries are fetched and copied back first; work enqueued by
    // `overlap` (dense: packing the low-precision operator) runs behind
    // the copy while the host waits only for the copy
    DevBuf<int> drc;
    DevBuf<double> dv;
    double *hv = nullptr;
    hipEvent_t ready = nullptr;
    if (npairs > 0) {
        IPT_STCHK(drc.alloc(rc.size()));
        IPT_STCHK(dv.alloc(rc.size()));
        IPT_STCHK(context_host_pinned(ctx, rc.size() * sizeof(double), (void **)&hv));
        IPT_HIPCHK(hipMemcpyAsync(drc.p, rc.data(), rc.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        IPT_STCHK(gather(drc.p, npairs, dv.p));
        IPT_HIPCHK(hipMemcpyAsync(hv, dv.p, rc.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        IPT_HIPCHK(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
        const hipError_t er = hipEventRecord(ready, ctx->stream);
        if (er != hipSuccess) {
            (void)hipEventDestroy(ready);
            return IPT_CUDA_CUDA_ERROR;
        }
    }
    trace_mark("pair gather enqueued");
    if (overlap) {
        const int st = overlap();
        if (st != IPT_CUDA_SUCCESS) {
            if (ready != nullptr) {
                (void)hipEventSynchronize(ready);
                (void)hipEventDestroy(ready);
            }
            return st;
        }
    }
    trace_mark("overlapped work enqueued");
    if (npairs > 0) {
        const hipError_t es = hipEventSynchronize(ready);
        (void)hipEventDestroy(ready);
        IPT_HIPCHK(es);
        for (int i = 0; i < npairs; ++i) {
            const int r = rc[(size_t)2 * i], c = rc[(size_t)2 * i + 1];
            wc.lut.emplace_back((long)r * n + c, hv[(size_t)2 * i]);
            wc.lut.emplace_back((long)c * n + r, hv[(size_t)2 * i + 1]);
        }
        std::sort(wc.lut.begin(), wc.lut.end());
    }
    trace_mark("pair values on the host");
    bool too_large = false;
    *ranges = target_clusters(wc, sd, k, o, &too_large);
    record_ranges(ctx, *ranges);
    trace_mark("clusters detected");
    if (too_large && env_int("IPT_BLOCK_CLUSTER_STALL", 0) <= 0) {
        int worst = 0;
        for (const auto &r : *ranges) {
            worst = std::max(worst, r.second - r.first + 1);
        }
        fprintf(stderr, "IPT block-cluster invalid argument at block detection: block_size=%d max_block_size=%d\n",
                worst, o.max_block);
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    return IPT_CUDA_SUCCESS;
}

// `early` (may be empty) is called as soon as the working matrix and its
// largest off-diagonal magnitude are known, before the host-side cluster
// detection, so device work it enqueues overlaps that host work.
int dense_prepare(Context *ctx, const double *dH, int n, int k, DensePrep *pr,
                  const std::function<int(const DensePrep &)> &early)
{
    IPT_STCHK(pr->d.alloc((size_t)n));
    DevBuf<double> colsum, colmax;
    DevBuf<int> asym;
    IPT_STCHK(colsum.alloc((size_t)n));
    IPT_STCHK(colmax.alloc((size_t)n));
    IPT_STCHK(asym.alloc(1));
    trace_mark("preparation buffers");
    IPT_STCHK(dense_diagonal(ctx, dH, n, n, pr->d.p));
    trace_mark("diagonal enqueued");
    IPT_STCHK(dense_scan(ctx, dH, n, n, colsum.p, colmax.p, asym.p));
    std::vector<double> hd((size_t)n), hsum((size_t)n), hmax((size_t)n);
    int hasym = 0;
    trace_mark("scan enqueued");
    IPT_STCHK(read_host_many(ctx, {HostRead{hd.data(), pr->d.p, (size_t)n * sizeof(double)},
                                   HostRead{hsum.data(), colsum.p, (size_t)n * sizeof(double)},
                                   HostRead{hmax.data(), colmax.p, (size_t)n * sizeof(double)},
                                   HostRead{&hasym, asym.p, sizeof(int)}}));
    trace_mark("scan results on the host");
    if (hasym) {
        fprintf(stderr, "IPT block-cluster path currently requires a real symmetric matrix\n");
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    // symmetric: row sums == column sums
    pr->normH = 0.0;
    for (double s : hsum) {
        pr->normH = std::max(pr->normH, s);
    }
    std::vector<int> inv;
    stable_sort_permutation(hd, &pr->perm, &inv);
    pr->permuted = false;
    for (int i = 0; i < n; ++i) {
        if (pr->perm[(size_t)i] != i) {
            pr->permuted = true;
            break;
        }
    }
    if (pr->permuted) {
        IPT_STCHK(pr->dperm.alloc((size_t)n));
        IPT_HIPCHK(hipMemcpyAsync(pr->dperm.p, pr->perm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice,
                                  ctx->stream));
    }
    std::vector<double> sd((size_t)n), soff((size_t)n);
    for (int s = 0; s < n; ++s) {
        sd[(size_t)s] = hd[(size_t)pr->perm[(size_t)s]];
        soff[(size_t)s] = hmax[(size_t)pr->perm[(size_t)s]];
    }
    // column sums of |H| (= row sums, symmetric) minus the diagonal, in
    // sorted order, formed on the device from the scan's column sums and the
    // unsorted diagonal (pr->d is overwritten by the sorted one after it)
    IPT_STCHK(pr->rowabs.alloc((size_t)n));
    IPT_STCHK(offdiag_row_sums(ctx, colsum.p, pr->d.p, pr->permuted ? pr->dperm.p : nullptr, n, pr->rowabs.p));
    if (pr->permuted) {
        IPT_STCHK(pr->Hs.alloc((size_t)n * n));
        IPT_STCHK(dense_permute(ctx, dH, n, pr->Hs.p, n, n, pr->dperm.p));
        IPT_HIPCHK(hipMemcpyAsync(pr->d.p, sd.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice,
                                  ctx->stream));
        pr->H = pr->Hs.p;
    } else {
        pr->H = dH;
    }
    trace_mark("diagonal order and row sums sent");
    pr->hd = sd;
    pr->max_offdiag = 0.0;
    for (double x : soff) {
        pr->max_offdiag = std::max(pr->max_offdiag, x);
    }
    GatherFn gather = [&](const int *drc, int npairs, double *dv) {
        return gather_pairs(ctx, pr->H, n, drc, npairs, dv);
    };
    std::function<int()> overlap;
    if (early) {
        overlap = [&]() { return early(*pr); };
    }
    // IPT_BLOCK_CLUSTER_OVERSAMPLE widens the targets; adaptive expansion
    // reads CSC columns and applies to sparse input only
    return detect_clusters(ctx, sd, soff, n, block_targets_from_env(n, k).target_k, gather, &pr->ranges, overlap);
}

}  // namespace

// Dense block-cluster solve on a device matrix; outputs in the caller's
// row order.  Used by the dense host entry points and the device entry
// points when IPT_BLOCK_CLUSTER=1.
int block_dense_device(Context *ctx, const double *dH, int n, int k, int maxiter, double tol, int use_tol,
                       double **d_vec_out, double **d_val_out, IPTCudaResult *info)
{
    auto t0 = std::chrono::steady_clock::now();
    trace_begin();
    const BlockParams p = params_from_env(maxiter, tol, use_tol);
    trace_mark("options read");
    DensePrep pr;
    LowpOperator lowp;
    const LowpOperator *lp = nullptr;
    // the low-precision copy of E is packed while the host detects clusters
    auto pack = [&](const DensePrep &q) -> int {
        if (p.lowp_mode >= 0) {
            IPT_STCHK(lowp_operator_create(ctx, q.H, n, n, p.lowp_mode, q.max_offdiag, &lowp));
            lp = &lowp;
        }
        return IPT_CUDA_SUCCESS;
    };
    IPT_STCHK(dense_prepare(ctx, dH, n, k, &pr, pack));
    trace_mark("dense preparation returned");
    const double prep = seconds_since(t0);
    const double *H = pr.H;
    // When every product is FP64, a copy of H in MFMA operand order makes
    // each of them stream contiguously; one pass over H, repaid within a
    // few products.
    F64Packed packed;
    const bool pack_ok = n >= 4096 && env_int("IPT_F64_PACKED", 1) != 0;
    if (p.lowp_mode < 0 && pack_ok) {
        IPT_STCHK(f64_packed_create(ctx, H, n, n, &packed));
    }
    // the mixed path's FP64 products are few until Davidson, whose
    // corrections are all FP64 products: the copy is made when it starts
    const std::function<int()> before_davidson = [&]() -> int {
        if (packed.p.p == nullptr && pack_ok) {
            return f64_packed_create(ctx, H, n, n, &packed);
        }
        return IPT_CUDA_SUCCESS;
    };
    ApplyFn apply = [&](const double *X, double *Y, int nc) {
        if (packed.p.p != nullptr) {
            return gemm_f64_packed(ctx, packed, X, n, Y, n, nc);
        }
        return gemm_f64(ctx, H, n, X, n, Y, n, n, nc);
    };
    // H * (identity block of cluster [first,last]) = columns first..last of H
    const std::vector<std::pair<int, int>> &rg = pr.ranges;
    ApplyFn identity_product = [&](const double *, double *Y, int) -> int {
        int off = 0;
        for (const auto &r : rg) {
            const int m = r.second - r.first + 1;
            IPT_HIPCHK(hipMemcpyAsync(Y + (size_t)off * n, H + (size_t)r.first * n, (size_t)n * m * sizeof(double),
                                      hipMemcpyDeviceToDevice, ctx->stream));
            off += m;
        }
        return IPT_CUDA_SUCCESS;
    };
    const DeltaFn delta = [&](const double *B, const double *Ba, const double *Ya, int nc, double *Y) {
        return lowp_apply_delta(ctx, lowp, B, Ba, Ya, n, pr.d.p, nc, Y);
    };
    BlockResult br;
    IPT_STCHK(block_pipeline(ctx, apply, identity_product, lp != nullptr ? &delta : nullptr, pr.d.p, pr.rowabs.p, n,
                             k, pr.ranges, pr.normH, p, &br, false, lp, before_davidson));
    record_stats(ctx, br);
    trace_mark("block pipeline returned");
    if (p.verbose) {
        fprintf(stderr,
                "IPT block-cluster dense: n=%d k=%d clusters=%zu basis_cols=%d precision=%d steps=%d "
                "fp64_products=%lld lowp_products=%lld davidson_steps=%d max_fp=%.3e\n",
                n, k, pr.ranges.size(), br.basis_cols, p.lowp_mode, br.iterations, br.fp64_products,
                br.lowp_products, br.dav.accepted_steps, br.fp_max);
    }
    DevBuf<double> vec, val;
    IPT_STCHK(vec.alloc((size_t)n * k));
    IPT_STCHK(val.alloc((size_t)k));
    if (pr.permuted) {
        IPT_STCHK(unpermute_rows(ctx, br.X.p, vec.p, n, n, k, pr.dperm.p));
    } else {
        IPT_HIPCHK(hipMemcpyAsync(vec.p, br.X.p, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    }
    IPT_HIPCHK(hipMemcpyAsync(val.p, br.theta.p, (size_t)k * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    if (info != nullptr) {
        fill_result(info, br, n, k);
        fill_target_fields(info, block_targets_from_env(n, k), pr.ranges, k, std::vector<int>(), false);
        info->preparation_time_sec = prep;
        info->solve_time_sec = seconds_since(t0);
    }
    *d_vec_out = vec.take();
    *d_val_out = val.take();
    return IPT_CUDA_SUCCESS;
}

// The LDS-tiled layout for the narrow products of a large CSR
// (ipt_csr_tiles.hip): IPT_CSR_TILES = 0 never, 1 (default) when one column
// of X (8n bytes) exceeds an XCD's 4 MiB L2, 2 always.  tl->groups stays 0
// when not taken, also for n > 2^25 or a column index outside [0, n) (the
// layout's own checks); *maxw = widest product taken through it (IPT_CSR_TILES_WIDTH).
static int narrow_tiles(Context *ctx, const int *rp, const int *ci, const double *v, int n, CsrTiles *tl, int *maxw)
{
    const int mode = env_int("IPT_CSR_TILES", 1);
    *maxw = std::max(1, env_int("IPT_CSR_TILES_WIDTH", 4));
    if (mode == 0 || (mode == 1 && n < (1 << 19))) {
        return IPT_CUDA_SUCCESS;
    }
    int nnz = 0;
    IPT_HIPCHK(hipMemcpyAsync(&nnz, rp + n, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    const int st = csr_tiles_create(ctx, rp, ci, v, n, nnz, tl);
    if (st == IPT_CUDA_INVALID_ARGUMENT) {
        tl->groups = 0;
        tl->keys.release();
        tl->vals.release();
        tl->start.release();
        tl->ltile.release();
        tl->lstart.release();
        tl->glen.release();
        return IPT_CUDA_SUCCESS;
    }
    trace_mark("narrow-product tiles built");
    return st;
}

// Sparse block-cluster solve with the matrix resident on the device (CSC of
// a symmetric matrix = its CSR).  Runs when every row's indices ascend and
// the diagonal is already sorted (no permutation needed); otherwise sets
// *host_path and returns, and the caller takes the host preparation route
// (IPT/src/ipt_cuda.cu:4551-4683 permutes on the host as well).
int block_sparse_device(Context *ctx, const int *d_rp, const int *d_ci, const double *d_v, int n, int k,
                        int maxiter, double tol, int use_tol, double **d_vec_out, double **d_val_out,
                        int *iters_out, double *fp_out, bool *host_path)
{
    *host_path = false;
    auto t0 = std::chrono::steady_clock::now();
    const BlockParams p = params_from_env(maxiter, tol, use_tol);
    DevBuf<double> d, asum, omax, rowabs;
    DevBuf<int> flags;
    IPT_STCHK(d.alloc((size_t)n));
    IPT_STCHK(asum.alloc((size_t)n));
    IPT_STCHK(omax.alloc((size_t)n));
    IPT_STCHK(rowabs.alloc((size_t)n));
    IPT_STCHK(flags.alloc(1));
    IPT_HIPCHK(hipMemsetAsync(flags.p, 0, sizeof(int), ctx->stream));
    IPT_STCHK(csr_scan(ctx, d_rp, d_ci, d_v, n, d.p, asum.p, omax.p, flags.p));
    int hflags = 0;
    IPT_STCHK(read_host(ctx, &hflags, flags.p, sizeof(int)));
    if (hflags & 1) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    if (hflags != 0) {
        *host_path = true;
        return IPT_CUDA_SUCCESS;
    }
    // exact symmetry by fingerprint (one streaming pass); only a mismatch
    // runs the entry-by-entry check with the reference's tolerance
    DevBuf<unsigned long long> fps;
    IPT_STCHK(fps.alloc(2));
    IPT_HIPCHK(hipMemsetAsync(fps.p, 0, 2 * sizeof(unsigned long long), ctx->stream));
    IPT_STCHK(csr_symhash(ctx, d_rp, d_ci, d_v, n, fps.p));
    unsigned long long hfp[2] = {0, 0};
    IPT_STCHK(read_host(ctx, hfp, fps.p, sizeof(hfp)));
    if (hfp[0] != hfp[1]) {
        IPT_STCHK(csr_symcheck(ctx, d_rp, d_ci, d_v, n, flags.p));
    }
    std::vector<double> hd((size_t)n), hsum((size_t)n), hmax((size_t)n);
    IPT_HIPCHK(hipMemcpyAsync(hd.data(), d.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    IPT_HIPCHK(hipMemcpyAsync(hsum.data(), asum.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    IPT_HIPCHK(hipMemcpyAsync(hmax.data(), omax.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    IPT_STCHK(read_host(ctx, &hflags, flags.p, sizeof(int)));
    if (hflags & 4) {
        fprintf(stderr, "IPT block-cluster path currently requires a real Hermitian CSC matrix\n");
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    for (int i = 0; i + 1 < n; ++i) {
        if (hd[(size_t)i + 1] < hd[(size_t)i]) {
            *host_path = true;
            return IPT_CUDA_SUCCESS;
        }
    }
    double normH = 0.0;
    std::vector<double> rabs((size_t)n);
    for (int i = 0; i < n; ++i) {
        normH = std::max(normH, hsum[(size_t)i]);
        rabs[(size_t)i] = std::max(0.0, hsum[(size_t)i] - std::fabs(hd[(size_t)i]));
    }
    IPT_HIPCHK(hipMemcpyAsync(rowabs.p, rabs.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    std::vector<std::pair<int, int>> ranges;
    GatherFn gather = [&](const int *drc, int npairs, double *dv) {
        return csr_gather_pairs(ctx, d_rp, d_ci, d_v, drc, npairs, dv);
    };
    IPT_STCHK(detect_clusters(ctx, hd, hmax, n, block_targets_from_env(n, k).target_k, gather, &ranges));
    const double prep = seconds_since(t0);
    int maxcols = 0;
    for (const auto &r : ranges) {
        maxcols += r.second - r.first + 1;
    }
    maxcols = std::max(maxcols, std::min(n, k + std::max(16, (2 * k + 2) / 3)) + k + 1);
    const int chunk = std::min(maxcols, 128);
    DevBuf<double> xr, yr;
    // rows of X at csr_row_stride(w) <= round_up(w, 16) doubles
    IPT_STCHK(xr.alloc((size_t)n * ((chunk + 15) / 16 * 16)));
    IPT_STCHK(yr.alloc((size_t)n * chunk));
    // column slices for the narrow products (rows ascend on this path)
    CsrSlices slices;
    if (env_int("IPT_CSR_SLICES", 1) != 0) {
        IPT_STCHK(csr_slices_create(ctx, d_rp, d_ci, n, &slices));
    }
    CsrTiles tiles;
    int tiles_w = 0;
    IPT_STCHK(narrow_tiles(ctx, d_rp, d_ci, d_v, n, &tiles, &tiles_w));
    ApplyFn apply = [&](const double *X, double *Y, int nc) -> int {
        if (tiles.groups > 0 && nc <= tiles_w) {
            // column-major in and out: no transposition
            return spmm_csr_tiles(ctx, tiles, X, 1, n, Y, 1, n, nc);
        }
        for (int c0 = 0; c0 < nc; c0 += chunk) {
            const int w = std::min(chunk, nc - c0);
            const long ld = csr_row_stride(w);
            IPT_STCHK(colmajor_to_rowmajor(ctx, X + (size_t)c0 * n, n, xr.p, ld, n, w));
            IPT_STCHK(spmm_csr_rowmajor(ctx, d_rp, d_ci, d_v, n, xr.p, ld, yr.p, w, w,
                                        slices.nslices > 0 ? &slices : nullptr));
            IPT_STCHK(rowmajor_to_colmajor(ctx, yr.p, w, Y + (size_t)c0 * n, n, n, w));
        }
        return IPT_CUDA_SUCCESS;
    };
    // IPT_PRECISION on CSR input: the FP32 anchored product (any
    // requested low precision maps to FP32 for sparse input)
    SparseLowp slp;
    DeltaFn delta;
    if (p.lowp_mode >= 0) {
        int nnz = 0;
        IPT_HIPCHK(hipMemcpy(&nnz, d_rp + n, sizeof(int), hipMemcpyDeviceToHost));
        IPT_STCHK(sparse_lowp_create(ctx, d_rp, d_ci, d_v, n, nnz, &slp));
        slp.slices = slices.nslices > 0 ? &slices : nullptr;
        delta = [&](const double *B, const double *Ba, const double *Ya, int nc, double *Y) {
            return sparse_lowp_apply_delta(ctx, slp, B, Ba, Ya, n, d.p, nc, Y);
        };
    }
    BlockResult br;
    IPT_STCHK(block_pipeline(ctx, apply, ApplyFn(), p.lowp_mode >= 0 ? &delta : nullptr, d.p, rowabs.p, n, k, ranges,
                             normH, p, &br, true));
    record_stats(ctx, br);
    if (p.verbose) {
        fprintf(stderr,
                "IPT block-cluster sparse: n=%d k=%d clusters=%zu basis_cols=%d steps=%d fp64_products=%lld "
                "davidson_steps=%d max_fp=%.3e prep=%.3fs\n",
                n, k, ranges.size(), br.basis_cols, br.iterations, br.fp64_products, br.dav.accepted_steps, br.fp_max,
                prep);
    }
    DevBuf<double> vec, val;
    IPT_STCHK(vec.alloc((size_t)n * k));
    IPT_STCHK(val.alloc((size_t)k));
    IPT_HIPCHK(hipMemcpyAsync(vec.p, br.X.p, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    IPT_HIPCHK(hipMemcpyAsync(val.p, br.theta.p, (size_t)k * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    if (iters_out) *iters_out = br.iterations;
    if (fp_out) *fp_out = br.fp_max;
    *d_vec_out = vec.take();
    *d_val_out = val.take();
    return IPT_CUDA_SUCCESS;
}

int block_dense_host_entry(const double *H, int n, int k, int maxiter, double tol, int use_tol,
                           IPTCudaResult *result)
{
    result_init_block(result);
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(nullptr, &st);
    if (ctx == nullptr) {
        return st;
    }
    auto run = [&]() -> int {
        const size_t nn = (size_t)n * n, nk = (size_t)n * k;
        auto t0 = std::chrono::steady_clock::now();
        DevBuf<double> dH;
        IPT_STCHK(dH.alloc(nn));
        IPT_HIPCHK(hipMemcpy(dH.p, H, nn * sizeof(double), hipMemcpyHostToDevice));
        result->transfer_setup_time_sec = seconds_since(t0);
        double *dv = nullptr, *dl = nullptr;
        IPT_STCHK(block_dense_device(ctx, dH.p, n, k, maxiter, tol, use_tol, &dv, &dl, result));
        result->vectors = (double *)calloc(nk, sizeof(double));
        result->values = (double *)calloc((size_t)k, sizeof(double));
        hipError_t e1 = hipSuccess, e2 = hipSuccess;
        if (result->vectors != nullptr && result->values != nullptr) {
            e1 = hipMemcpy(result->vectors, dv, nk * sizeof(double), hipMemcpyDeviceToHost);
            e2 = hipMemcpy(result->values, dl, (size_t)k * sizeof(double), hipMemcpyDeviceToHost);
        }
        ipt_cuda_free_device_result(dv, dl);
        if (result->vectors == nullptr || result->values == nullptr) {
            return IPT_CUDA_ALLOCATION_FAILED;
        }
        IPT_HIPCHK(e1);
        IPT_HIPCHK(e2);
        return IPT_CUDA_SUCCESS;
    };
    st = run();
    if (st != IPT_CUDA_SUCCESS) {
        ipt_cuda_free_result(result);
    }
    return st;
}

int block_sparse_host_entry(const int *col_ptr, const int *row_ind, const double *values, int n, int k, int nnz,
                            int maxiter, double tol, int use_tol, IPTCudaResult *result)
{
    result_init_block(result);
    if (!csc_valid(col_ptr, row_ind, n, nnz)) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(nullptr, &st);
    if (ctx == nullptr) {
        return st;
    }
    auto run = [&]() -> int {
        auto t0 = std::chrono::steady_clock::now();
        // IPT/src/ipt_cuda.cu:4551-4683
        std::vector<double> d0;
        csc_diagonal_host(col_ptr, row_ind, values, n, &d0);
        std::vector<int> perm, inv;
        stable_sort_permutation(d0, &perm, &inv);
        HostCsc a;
        permute_csc(col_ptr, row_ind, values, n, inv, &a);
        std::vector<double> d;
        csc_diagonal_host(a.col_ptr.data(), a.row_ind.data(), a.values.data(), n, &d);
        if (!csc_is_symmetric(a)) {
            fprintf(stderr, "IPT block-cluster path currently requires a real Hermitian CSC matrix\n");
            return IPT_CUDA_INVALID_ARGUMENT;
        }
        const DegeneracyOptions o = degeneracy_options_from_env(env_double("IPT_DEGENERACY_THRESHOLD", 0.0));
        bool too_large = false;
        CscCoupling cc(&a);
        const BlockTargets bt = block_targets_from_env(n, k);
        std::vector<std::pair<int, int>> ranges = target_clusters(cc, d, bt.target_k, o, &too_large);
        record_ranges(ctx, ranges);
        if (too_large && env_int("IPT_BLOCK_CLUSTER_STALL", 0) <= 0) {
            int worst = 0;
            for (const auto &r : ranges) {
                worst = std::max(worst, r.second - r.first + 1);
            }
            fprintf(stderr,
                    "IPT block-cluster invalid argument at block detection: block_size=%d max_block_size=%d\n",
                    worst, o.max_block);
            return IPT_CUDA_INVALID_ARGUMENT;
        }
        std::vector<int> added;
        const bool limit_hit = bt.adaptive && adaptive_expand(a, d, k, o.max_block, bt.tau, &ranges, &added);
        fill_target_fields(result, bt, ranges, k, added, limit_hit);
        record_ranges(ctx, ranges);
        const double normH = csc_inf_norm(col_ptr, row_ind, values, n);
        result->preparation_time_sec = seconds_since(t0);
        auto t1 = std::chrono::steady_clock::now();
        // symmetric: CSC arrays are the CSR arrays of the same matrix
        DevBuf<int> rp, ci, dperm;
        DevBuf<double> vv, dd, xr, yr;
        const int pn = a.col_ptr[(size_t)n];
        IPT_STCHK(rp.alloc((size_t)n + 1));
        IPT_STCHK(ci.alloc((size_t)std::max(pn, 1)));
        IPT_STCHK(vv.alloc((size_t)std::max(pn, 1)));
        IPT_STCHK(dd.alloc((size_t)n));
        IPT_STCHK(dperm.alloc((size_t)n));
        // off-diagonal absolute row sums (symmetric: column sums), for the
        // learned Davidson preconditioner
        std::vector<double> rabs((size_t)n, 0.0);
        for (int c = 0; c < n; ++c) {
            for (int e = a.col_ptr[(size_t)c]; e < a.col_ptr[(size_t)c + 1]; ++e) {
                if (a.row_ind[(size_t)e] != c) {
                    rabs[(size_t)c] += std::fabs(a.values[(size_t)e]);
                }
            }
        }
        DevBuf<double> drabs;
        IPT_STCHK(drabs.alloc((size_t)n));
        // uploads on the library stream (device buffers may come from the
        // allocation cache of that stream)
        IPT_HIPCHK(hipMemcpyAsync(rp.p, a.col_ptr.data(), (size_t)(n + 1) * sizeof(int), hipMemcpyHostToDevice,
                                  ctx->stream));
        if (pn > 0) {
            IPT_HIPCHK(hipMemcpyAsync(ci.p, a.row_ind.data(), (size_t)pn * sizeof(int), hipMemcpyHostToDevice,
                                      ctx->stream));
            IPT_HIPCHK(hipMemcpyAsync(vv.p, a.values.data(), (size_t)pn * sizeof(double), hipMemcpyHostToDevice,
                                      ctx->stream));
        }
        IPT_HIPCHK(hipMemcpyAsync(dd.p, d.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        IPT_HIPCHK(hipMemcpyAsync(dperm.p, perm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        IPT_HIPCHK(hipMemcpyAsync(drabs.p, rabs.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice,
                                  ctx->stream));
        IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
        result->transfer_setup_time_sec = seconds_since(t1);
        auto t2 = std::chrono::steady_clock::now();
        const BlockParams p = params_from_env(maxiter, tol, use_tol);
        int maxcols = 0;
        for (const auto &r : ranges) {
            maxcols += r.second - r.first + 1;
        }
        maxcols = std::max(maxcols, std::min(n, k + std::max(16, (2 * k + 2) / 3)) + k + 1);
        IPT_STCHK(xr.alloc((size_t)n * ((std::min(maxcols, 128) + 15) / 16 * 16)));
        IPT_STCHK(yr.alloc((size_t)n * std::min(maxcols, 128)));
        CsrSlices slices;
        if (env_int("IPT_CSR_SLICES", 1) != 0) {
            IPT_STCHK(csr_slices_create(ctx, rp.p, ci.p, n, &slices));
        }
        CsrTiles tiles;
        int tiles_w = 0;
        IPT_STCHK(narrow_tiles(ctx, rp.p, ci.p, vv.p, n, &tiles, &tiles_w));
        ApplyFn apply = [&](const double *X, double *Y, int nc) -> int {
            if (tiles.groups > 0 && nc <= tiles_w) {
                return spmm_csr_tiles(ctx, tiles, X, 1, n, Y, 1, n, nc);
            }
            for (int c0 = 0; c0 < nc; c0 += 128) {
                const int w = std::min(128, nc - c0);
                const long ld = csr_row_stride(w);
                IPT_STCHK(colmajor_to_rowmajor(ctx, X + (size_t)c0 * n, n, xr.p, ld, n, w));
                IPT_STCHK(spmm_csr_rowmajor(ctx, rp.p, ci.p, vv.p, n, xr.p, ld, yr.p, w, w,
                                            slices.nslices > 0 ? &slices : nullptr));
                IPT_STCHK(rowmajor_to_colmajor(ctx, yr.p, w, Y + (size_t)c0 * n, n, n, w));
            }
            return IPT_CUDA_SUCCESS;
        };
        SparseLowp slp;
        DeltaFn delta;
        if (p.lowp_mode >= 0) {
            IPT_STCHK(sparse_lowp_create(ctx, rp.p, ci.p, vv.p, n, pn, &slp));
            slp.slices = slices.nslices > 0 ? &slices : nullptr;
            delta = [&](const double *B, const double *Ba, const double *Ya, int nc, double *Y) {
                return sparse_lowp_apply_delta(ctx, slp, B, Ba, Ya, n, dd.p, nc, Y);
            };
        }
        BlockResult br;
        IPT_STCHK(block_pipeline(ctx, apply, ApplyFn(), p.lowp_mode >= 0 ? &delta : nullptr, dd.p, drabs.p, n, k,
                                 ranges, normH, p, &br, true));
        record_stats(ctx, br);
        DevBuf<double> vec;
        IPT_STCHK(vec.alloc((size_t)n * k));
        IPT_STCHK(unpermute_rows(ctx, br.X.p, vec.p, n, n, k, dperm.p));
        result->vectors = (double *)calloc((size_t)n * k, sizeof(double));
        result->values = (double *)calloc((size_t)k, sizeof(double));
        if (result->vectors == nullptr || result->values == nullptr) {
            return IPT_CUDA_ALLOCATION_FAILED;
        }
        IPT_HIPCHK(hipMemcpy(result->vectors, vec.p, (size_t)n * k * sizeof(double), hipMemcpyDeviceToHost));
        IPT_HIPCHK(hipMemcpy(result->values, br.theta.p, (size_t)k * sizeof(double), hipMemcpyDeviceToHost));
        fill_result(result, br, n, k);
        result->solve_time_sec = seconds_since(t2);
        return IPT_CUDA_SUCCESS;
    };
    st = run();
    if (st != IPT_CUDA_SUCCESS) {
        ipt_cuda_free_result(result);
    }
    return st;
}

}  // namespace ipt

using namespace ipt;

extern "C" int ipt_hip_last_solve_stats(void *handle, int *basis_cols, int *steps, long long *fp64_products,
                                        long long *lowp_products, int *davidson_steps)
{
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    if (basis_cols) *basis_cols = ctx->last_basis_cols;
    if (steps) *steps = ctx->last_steps;
    if (fp64_products) *fp64_products = ctx->last_fp64_products;
    if (lowp_products) *lowp_products = ctx->last_lowp_products;
    if (davidson_steps) *davidson_steps = ctx->last_davidson_steps;
    return IPT_CUDA_SUCCESS;
}

extern "C" int ipt_hip_last_davidson_stats(void *handle, int *accepted, int *rejected, int *mlp_accepted,
                                           int *mlp_rejected, long long *matvecs, double *residual_after)
{
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    if (accepted) *accepted = ctx->last_davidson_steps;
    if (rejected) *rejected = ctx->last_davidson_rejected;
    if (mlp_accepted) *mlp_accepted = ctx->last_mlp_accepted;
    if (mlp_rejected) *mlp_rejected = ctx->last_mlp_rejected;
    if (matvecs) *matvecs = ctx->last_davidson_matvecs;
    if (residual_after) *residual_after = ctx->last_davidson_residual;
    return IPT_CUDA_SUCCESS;
}

// Cluster ranges of the last block-cluster solve on `handle`: pairs
// (first, last) of rows in the solver's sorted order, singletons included,
// in basis-column order.  Writes min(count, capacity) pairs.
extern "C" int ipt_hip_last_cluster_ranges(void *handle, int *first_last, int capacity, int *count)
{
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    if (capacity < 0 || (capacity > 0 && first_last == nullptr)) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    const int total = (int)(ctx->last_ranges.size() / 2);
    if (count) *count = total;
    for (int i = 0; i < std::min(total, capacity); ++i) {
        first_last[2 * i] = ctx->last_ranges[(size_t)2 * i];
        first_last[2 * i + 1] = ctx->last_ranges[(size_t)2 * i + 1];
    }
    return IPT_CUDA_SUCCESS;
}

// State of each cluster of the last block-cluster solve, in the order of
// ipt_hip_last_cluster_ranges: 0 converged, 1 still iterating when the
// iteration limit was reached, 2 stopped by stall detection.
extern "C" int ipt_hip_last_cluster_states(void *handle, int *states, int capacity, int *count)
{
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    if (capacity < 0 || (capacity > 0 && states == nullptr)) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    const int total = (int)ctx->last_cluster_state.size();
    if (count) *count = total;
    for (int i = 0; i < std::min(total, capacity); ++i) {
        states[i] = ctx->last_cluster_state[(size_t)i];
    }
    return IPT_CUDA_SUCCESS;
}
