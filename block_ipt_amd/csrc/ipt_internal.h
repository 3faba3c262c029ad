// Internal declarations shared by the translation units of libipt_hip.so.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ipt_hip.h"

#define IPT_HIPCHK(call)                                                      \
    do {                                                                      \
        hipError_t ipt_e_ = (call);                                           \
        if (ipt_e_ != hipSuccess) {                                           \
            fprintf(stderr, "HIP error at %s:%d: %s\n", __FILE__, __LINE__,   \
                    hipGetErrorString(ipt_e_));                               \
            return IPT_CUDA_CUDA_ERROR;                                       \
        }                                                                     \
    } while (0)

#define IPT_STCHK(call)                                                       \
    do {                                                                      \
        int ipt_s_ = (call);                                                  \
        if (ipt_s_ != IPT_CUDA_SUCCESS) {                                     \
            return ipt_s_;                                                    \
        }                                                                     \
    } while (0)

namespace ipt {

// ------------------------------------------------------------------ context

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    double *host_scalars = nullptr;  // pinned, 4096 doubles
    int *host_flags = nullptr;       // pinned, 4096 ints
    double *host_status = nullptr;   // pinned, 4096 doubles: the block iteration's queued status read
    void *host_pinned = nullptr;     // pinned staging for device-to-host reads, grows
    size_t host_pinned_bytes = 0;
    void *slots[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    // statistics of the last block-cluster solve
    int last_basis_cols = 0;
    int last_steps = 0;
    long long last_fp64_products = 0;
    long long last_lowp_products = 0;
    int last_davidson_steps = 0;
    int last_davidson_rejected = 0;
    int last_mlp_accepted = 0;
    int last_mlp_rejected = 0;
    long long last_davidson_matvecs = 0;
    double last_davidson_residual = 0.0;
    // cluster ranges [first, last] (sorted row order) of the last solve
    std::vector<int> last_ranges;
    // per cluster: 0 converged, 1 iterating at the limit, 2 stalled
    std::vector<int> last_cluster_state;
    size_t slot_bytes[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

// Reusable device scratch owned by the context (grown on demand, freed with
// the context).  Slots are private to one library routine each:
enum {
    SLOT_GEMM_SPLITK = 0,
    SLOT_REDUCE = 1,
    SLOT_SMALL = 2,
    SLOT_SOLVER = 3,
    SLOT_RR = 4,
    SLOT_ORTHO = 5,
    SLOT_LOWP = 6,
    SLOT_MISC = 7
};
int context_scratch(Context *ctx, int slot, size_t bytes, void **out);
// Pinned host staging buffer of at least `bytes` (grows; freed with the
// context), so that device-to-host reads can run asynchronously.
int context_host_pinned(Context *ctx, size_t bytes, void **out);

// NULL handle -> library context of the current device.
Context *context_from_handle(void *handle, int *status);

// Device buffer owned by the caller's scope.
// Device memory of library routines comes from a cache keyed by the
// library stream it is used on (set by context_from_handle for the calling
// thread): a block released by one call is handed to a later call on the
// same stream, whose work is ordered after every use of the block, so
// repeated solves make no hipMalloc/hipFree calls (hipFree synchronises the
// device).  Blocks handed to callers are detached and owned by them.
int pool_alloc(void **p, size_t bytes);
void pool_free(void *p);
void pool_detach(void *p);
void pool_set_stream(hipStream_t s);
void pool_release_stream(hipStream_t s);

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t count = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    int alloc(size_t n)
    {
        release();
        if (n == 0) {
            return IPT_CUDA_SUCCESS;
        }
        void *q = nullptr;
        if (pool_alloc(&q, n * sizeof(T)) != IPT_CUDA_SUCCESS) {
            fprintf(stderr, "IPT: device allocation of %zu bytes failed\n", n * sizeof(T));
            return IPT_CUDA_ALLOCATION_FAILED;
        }
        p = static_cast<T *>(q);
        count = n;
        return IPT_CUDA_SUCCESS;
    }
    void release()
    {
        if (p != nullptr) {
            pool_free(p);
        }
        p = nullptr;
        count = 0;
    }
    // hand the allocation to the caller (it is then released with hipFree)
    T *take()
    {
        T *q = p;
        pool_detach(q);
        p = nullptr;
        count = 0;
        return q;
    }
};

// ------------------------------------------------------------------ profiling

enum {
    PROF_GEMM_F64 = 0,
    PROF_LOWP_GEMM = 1,
    PROF_BLOCK_UPDATE = 2,
    PROF_SPMM = 3,
    PROF_SPMM_F32 = 4,
    PROF_SPMM_TILES = 5,
    PROF_COUNT = 6
};
struct ProfileToken {
    int id = -1;
    hipEvent_t start = nullptr;
    hipEvent_t stop = nullptr;
    bool markers = true;  // events recorded as stream markers (else by the launches)
};
bool profile_on();
// A span of launches timed by events recorded on the stream before and after
// it (marker packets: each costs a few microseconds between the kernels).
ProfileToken profile_begin(Context *ctx, int id);
// A span whose first and last launches go through launch_timed: the events
// take those dispatches' own start and end timestamps, so timing adds no
// packets to the stream and no gaps between the kernels.
ProfileToken profile_begin_launches(int id);
// cols: the columns the timed launch processed (summed per family)
void profile_end(Context *ctx, const ProfileToken &t, long long cols = 0);

// One launch of a span opened by profile_begin_launches: the first launch
// carries the start event, the last the stop event (hipExtLaunchKernelGGL);
// untimed launches and spans go out as ordinary launches.
template <typename... Args, typename F = void (*)(Args...)>
inline void launch_timed(const ProfileToken &t, bool first, bool last, F kernel, dim3 grid, dim3 block,
                         unsigned shmem, hipStream_t st, Args... args)
{
    if (t.id >= 0 && !t.markers && (first || last)) {
        hipExtLaunchKernelGGL(kernel, grid, block, shmem, st, first ? t.start : nullptr, last ? t.stop : nullptr, 0,
                              args...);
    } else {
        hipLaunchKernelGGL(kernel, grid, block, shmem, st, args...);
    }
}

// IPT_HIP_HOST_TRACE=1: the host's time at the stages of a solve
// (microseconds since trace_begin, on stderr), to place the host work that
// leaves the GPU idle; device allocations and releases are marked too.
void trace_begin();
void trace_mark(const char *what);
bool trace_on();

// ------------------------------------------------------------------ options

double env_double(const char *name, double default_value);
int env_int(const char *name, int default_value);
int env_flag(const char *name);
int env_flag_default(const char *name, int default_value);

// ------------------------------------------------------------------ kernels
// All launchers enqueue on ctx->stream and return an IPT status.

// Y = H * Z; H column-major n x n (ldh), Z column-major n x ncols (ldz),
// Y column-major (ldy).  FP64 MFMA.
int gemm_f64(Context *ctx, const double *H, long ldh, const double *Z, long ldz, double *Y,
             long ldy, int n, int ncols);
// FP64 H packed in MFMA A-operand order (ipt_gemm_f64.hip): one extra copy
// of H, streamed contiguously by gemm_f64_packed; worth its one pass over H
// when a solve runs many FP64 products.
struct F64Packed {
    int n = 0;
    int nkp = 0;  // k pairs of 8
    DevBuf<double2> p;
};
int f64_packed_create(Context *ctx, const double *H, long ld, int n, F64Packed *op);
// Y (n x ncols, ldy) = H Z (Z n x ncols, ldz) from the packed copy
int gemm_f64_packed(Context *ctx, const F64Packed &op, const double *Z, long ldz, double *Y, long ldy, int ncols);

// d[i] = H[i + i*ldh]
int dense_diagonal(Context *ctx, const double *H, long ldh, int n, double *d);
// x[0..count) = v on the device (no host buffer, no synchronisation)
int fill_device(Context *ctx, double *x, long count, double v);
// out[s] = max(0, colsum[p] - |d[p]|), p = perm[s] (perm NULL: identity):
// off-diagonal row sums of |H| (symmetric H, column sums of |H| in colsum)
int offdiag_row_sums(Context *ctx, const double *colsum, const double *d, const int *perm, int n, double *out);
// CSC diagonal (first match per column), IPT/src/ipt_cuda.cu:1469-1489.
int csc_diagonal(Context *ctx, const int *col_ptr, const int *row_ind, const double *values, int n,
                 double *d);

// Z = I(:, 0:k) in column-major layout (ld).
int set_identity(Context *ctx, double *Z, long ld, int n, int k);

// Plain IPT update (one iteration after Y = H Z) with the squared Frobenius
// norms of F(Z)-Z and Z written to d_norms[0], d_norms[1] (device).
// Layout: element (r,j) at r*rs + j*cs.  Znew must not alias Y or Z.
int plain_update(Context *ctx, const double *Y, const double *Z, double *Znew, const double *d,
                 int n, int k, long rs, long cs, double *d_norms);

// values[j] = Y[j*rs + j*cs]
int gather_diagonal(Context *ctx, const double *Y, int k, long rs, long cs, double *values);

// Symmetric sparse preparation (ipt_sparse.hip), CSR with ascending columns.
// flags |= 1 index out of range, 2 columns not strictly ascending (scan),
// 4 not symmetric (symcheck; call only when the scan set no flag).
int csr_scan(Context *ctx, const int *rp, const int *ci, const double *v, int n, double *diag, double *abssum,
             double *offmax, int *flags);
int csr_symcheck(Context *ctx, const int *rp, const int *ci, const double *v, int n, int *flags);
// sums[0] / sums[1]: order-independent 64-bit fingerprints of the entries
// (r, c, v) and of their transposes (c, r, v); equal for an exactly
// symmetric matrix (sums must be zero on entry).
int csr_symhash(Context *ctx, const int *rp, const int *ci, const double *v, int n, unsigned long long *sums);
int csr_gather_pairs(Context *ctx, const int *rp, const int *ci, const double *v, const int *rc, int npairs,
                     double *out);

// CSR SpMM: Y(n x k, row-major ld ldy) = A(CSR) * X(n x k, row-major ld ldx).
// Column slices of a CSR whose columns ascend within each row: for slice
// b of kCsrMaxSlices (columns [b w, (b+1) w), w = ceil(n / kCsrMaxSlices)),
// row r's entries are [split[b n + r], split[(b+1) n + r]).  Bounds only;
// the entries stay where they are.
constexpr int kCsrMaxSlices = 16;
struct CsrSlices {
    int n = 0;
    int nslices = 0;
    DevBuf<int> split;  // (nslices + 1) x n
};
int csr_slices_create(Context *ctx, const int *row_ptr, const int *col_ind, int n, CsrSlices *sl);
// passes the product of width k takes (1: the whole matrix at once)
int csr_slices_for(const CsrSlices &sl, int k);
// Leading dimension for a row-major X of w columns of elem bytes whose rows
// the CSR product gathers: the smallest of w, next_pow2(w) and w rounded up
// to a 128-byte line with the fewest lines per row (at most that rounding)
long csr_row_stride(int w, int elem = 8);
// The entries of a CSR (any column order within a row) regrouped into
// buckets of (4096-row group, 4096-column tile) for the narrow products
// with x in LDS (ipt_csr_tiles.hip): per entry a 32-bit key (row in the
// group << 12 | column in the tile) and the value; start has
// groups * tiles + 1 offsets.
struct CsrTiles {
    int n = 0;
    int groups = 0;
    int tiles = 0;
    DevBuf<long> start;
    DevBuf<unsigned> keys;
    DevBuf<double> vals;
    // per group (stride tiles, tiles + 1): its nonempty buckets' tiles and
    // starts (then the group's end), and their count
    DevBuf<int> ltile;
    DevBuf<long> lstart;
    DevBuf<int> glen;
};
int csr_tiles_create(Context *ctx, const int *row_ptr, const int *col_ind, const double *vals, int n, long long nnz,
                     CsrTiles *out);
// Y = A X through the tiles, two columns per launch: element (r, c) of X at
// X[r * xrs + c * xcs], of Y at Y[r * yrs + c * ycs] (row-major: xrs = ld,
// xcs = 1; column-major: xrs = 1, xcs = ld)
int spmm_csr_tiles(Context *ctx, const CsrTiles &tl, const double *X, long xrs, long xcs, double *Y, long yrs,
                   long ycs, int k);
// Y = A X, X and Y row-major (ld >= k); with slices, narrow products run
// as one pass per column slice (gathers confined to a slice of X).
int spmm_csr_rowmajor(Context *ctx, const int *row_ptr, const int *col_ind, const double *vals,
                      int n, const double *X, long ldx, double *Y, long ldy, int k,
                      const CsrSlices *slices = nullptr);
// the same product in exactly ns passes (a power of two dividing
// slices.nslices; 1: the whole matrix in one pass)
int spmm_csr_sliced_passes(Context *ctx, const int *row_ptr, const int *col_ind, const double *vals, int n,
                           const double *X, long ldx, double *Y, long ldy, int k, const CsrSlices &slices, int ns);

// A <- (A + A^T)/2 for a square column-major n x n (ld) matrix
// (IPT/src/mixed_precision/ipt_rayleigh_mixed_cuda.cu:84-106).
int symmetrize(Context *ctx, double *A, long ld, int n);

// Symmetric eigendecomposition on the device (ipt_eig.hip): eigenvalues
// ascending in w (m), orthonormal eigenvectors in V (m x m, ldv); A (m x m,
// lda) is read only.  *sweeps_out (may be NULL): block-Jacobi sweeps used.
int sym_eig_device(Context *ctx, const double *A, long lda, int m, double *w, double *V, long ldv, int *sweeps_out);

// Layout copies between column-major (ld) and row-major (ld) n x k.
int colmajor_to_rowmajor(Context *ctx, const double *A, long lda, double *B, long ldb, int n, int k);
int rowmajor_to_colmajor(Context *ctx, const double *A, long lda, double *B, long ldb, int n, int k);

}  // namespace ipt
