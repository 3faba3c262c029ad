// CSR product of narrow widths with the gathered vector in LDS.
//
// The row-major CSR product gathers X rows from caches one 128-byte line per
// entry: for one or two columns (8 or 16 bytes used of each line) the
// product is bound by the line traffic from L2, not by the matrix stream.
// Here the entries are regrouped once per matrix into buckets (row group of
// kGroupRows rows, column tile of kTileCols columns); a workgroup owns a row
// group, stages 64 KiB of x into LDS (two column tiles of one column, or one
// tile of two columns) and streams the buckets' entries (a 32-bit key: row
// within the group, column within the tile; and the value), adding
// v * x[col] into its rows' sums in LDS.  The matrix is read once per one or
// two columns, coalesced, x once per workgroup from the caches, and every
// gather is an LDS read.  The LDS sums take the products in the order the
// hardware applies the atomic adds, so results agree with the row-major
// product to rounding, not bit for bit.
#include <algorithm>

#include "ipt_internal.h"

namespace ipt {

namespace {

constexpr int kGroupRowsLog = 12;
constexpr int kTileColsLog = 12;
constexpr int kGroupRows = 1 << kGroupRowsLog;  // 4096 rows: 32 KiB of sums
constexpr int kTileCols = 1 << kTileColsLog;    // 4096 columns: 32 KiB per column of x
constexpr int kThreads = 1024;
constexpr int kMaxTiles = 8192;  // column tiles (n <= 2^25): 64 KiB of LDS counters

// Exclusive prefix over the block of one value per thread (kThreads
// threads, row order = thread order); *total gets the block's sum.
__device__ __forceinline__ long block_exclusive_scan(long v, long *wsum, long *total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    long inc = v;
    for (int o = 1; o < 64; o <<= 1) {
        const long up = __shfl_up(inc, o, 64);
        if (lane >= o) {
            inc += up;
        }
    }
    if (lane == 63) {
        wsum[w] = inc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long s = 0;
        for (int i = 0; i < kThreads / 64; ++i) {
            const long t = wsum[i];
            wsum[i] = s;
            s += t;
        }
        *total = s;
    }
    __syncthreads();
    const long r = wsum[w] + inc - v;
    __syncthreads();
    return r;
}

// Bucket counts (pass 0) or the scatter into the buckets (pass 1) of row
// group blockIdx.x: a wavefront per row reads the row's entries 64 at a
// time (coalesced), the bucket of an entry is its column tile, and an LDS
// counter per tile (pass 0: the count; pass 1: the next free position,
// starting at the bucket's offset) takes one atomic add per entry.  Within
// a bucket the entries land in the order of those adds; the product adds
// them in LDS in arbitrary order as well.  Pass 0 sets *bad for a column
// outside [0, n).  ntiles <= kMaxTiles (LDS).
template <int PASS>
__global__ __launch_bounds__(kThreads) void k_tiles_build(const int *__restrict__ rp, const int *__restrict__ ci,
                                                          const double *__restrict__ v, int n, int ntiles,
                                                          long *__restrict__ count, const long *__restrict__ start,
                                                          unsigned *__restrict__ keys, double *__restrict__ vals,
                                                          int *__restrict__ bad)
{
    __shared__ unsigned long long ctr[kMaxTiles];
    const int g = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int t = tid; t < ntiles; t += kThreads) {
        ctr[t] = PASS == 0 ? 0ull : (unsigned long long)start[(long)g * ntiles + t];
    }
    __syncthreads();
    bool wrong = false;
    const long r_end = min((long)n, ((long)g + 1) * kGroupRows);
    for (long r = (long)g * kGroupRows + wave; r < r_end; r += kThreads / 64) {
        const unsigned rl = (unsigned)(r - (long)g * kGroupRows) << kTileColsLog;
        const int pe = rp[r + 1];
        for (int e = rp[r] + lane; e < pe; e += 64) {
            const int c = ci[e];
            if (c < 0 || c >= n) {
                wrong = true;
                continue;
            }
            if (PASS == 0) {
                atomicAdd(&ctr[c >> kTileColsLog], 1ull);
            } else {
                const unsigned long long dst = atomicAdd(&ctr[c >> kTileColsLog], 1ull);
                keys[dst] = rl | ((unsigned)c & (kTileCols - 1));
                vals[dst] = v[e];
            }
        }
    }
    if (PASS == 0) {
        if (wrong) {
            atomicOr(bad, 1);
        }
        __syncthreads();
        for (int t = tid; t < ntiles; t += kThreads) {
            count[(long)g * ntiles + t] = (long)ctr[t];
        }
    }
}

// start[b] = sum of count[0..b), start[nb] = the total: one block, each
// thread a contiguous chunk of the bucket counts.
__global__ __launch_bounds__(kThreads) void k_tiles_offsets(const long *__restrict__ count, long nb,
                                                            long *__restrict__ start)
{
    __shared__ long wsum[kThreads / 64];
    __shared__ long total;
    const long per = (nb + kThreads - 1) / kThreads;
    const long b0 = min(nb, (long)threadIdx.x * per), b1 = min(nb, b0 + per);
    long mine = 0;
    for (long b = b0; b < b1; ++b) {
        mine += count[b];
    }
    long s = block_exclusive_scan(mine, wsum, &total);
    for (long b = b0; b < b1; ++b) {
        start[b] = s;
        s += count[b];
    }
    if (threadIdx.x == 0) {
        start[nb] = total;
    }
}

// Per group (one thread each): its nonempty buckets in tile order, tile
// index ltile[i] and start lstart[i] (lstart[m] = the group's end; the
// buckets between two listed ones are empty, so bucket i ends where bucket
// i + 1 starts), and their count glen.
__global__ __launch_bounds__(256) void k_tiles_lists(const long *__restrict__ start, int groups, int ntiles,
                                                     int *__restrict__ ltile, long *__restrict__ lstart,
                                                     int *__restrict__ glen)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= groups) {
        return;
    }
    const long *bs = start + (long)g * ntiles;
    int *lt = ltile + (long)g * ntiles;
    long *ls = lstart + (long)g * (ntiles + 1);
    int m = 0;
    for (int t = 0; t < ntiles; ++t) {
        if (bs[t + 1] > bs[t]) {
            lt[m] = t;
            ls[m] = bs[t];
            ++m;
        }
    }
    ls[m] = bs[ntiles];
    glen[g] = m;
}

// Y (rows of group blockIdx.x, KC columns) = A X: element (r, c) of X at
// x[r * xrs + c * xcs], of Y at y[r * yrs + c * ycs].  A stage stages 64 KiB
// of x: KC = 1 two nonempty tiles (their buckets are adjacent, the
// second's columns offset by kTileCols in LDS), KC = 2 one tile of both
// columns.  Software-pipelined: while a stage's entries are added, each
// thread holds the next stage's x values and first batch of entries in
// registers, so the stage boundary waits on no memory.
template <int KC>
__global__ __launch_bounds__(kThreads) void k_spmm_tiles(const int *__restrict__ ltile,
                                                         const long *__restrict__ lstart,
                                                         const int *__restrict__ glen,
                                                         const unsigned *__restrict__ keys,
                                                         const double *__restrict__ vals, int n, int ntiles,
                                                         const double *__restrict__ x, long xrs, long xcs,
                                                         double *__restrict__ y, long yrs, long ycs)
{
    constexpr int TS = 2 / KC;               // tiles per stage
    constexpr int SV = TS * kTileCols * KC;  // staged values per stage
    constexpr int XPT = SV / kThreads;       // of them per thread
    constexpr int U = 8;                     // entries per thread per batch
    constexpr long BATCH = (long)U * kThreads;
    __shared__ double xs[SV];
    __shared__ double ys[kGroupRows * KC];
    const int g = blockIdx.x;
    const int tid = threadIdx.x;
    const int m = glen[g];
    const int *lt = ltile + (long)g * ntiles;
    const long *ls = lstart + (long)g * (ntiles + 1);
    const int nst = (m + TS - 1) / TS;
    for (int i = tid; i < kGroupRows * KC; i += kThreads) {
        ys[i] = 0.0;
    }
    double xr[XPT];
    unsigned kn[U];
    double vn[U];
    long ne0 = 0, nem = 0, ne1 = 0;
    // x and the first batch of entries of stage j into registers
    auto prefetch = [&](int j) {
        const int i0 = TS * j;
        ne0 = ls[i0];
        nem = ls[i0 + 1];
        ne1 = ls[min(i0 + TS, m)];
#pragma unroll
        for (int q = 0; q < XPT; ++q) {
            constexpr int per_tile = kTileCols * KC;
            const int half = q * kThreads / per_tile;
            const int within = tid + q * kThreads - half * per_tile;
            const int tile = i0 + half < m ? lt[i0 + half] : -1;
            const long col = (long)tile * kTileCols + within / KC;
            xr[q] = tile >= 0 && col < n ? x[col * xrs + (within % KC) * xcs] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long e = ne0 + tid + (long)u * kThreads;
            kn[u] = e < ne1 ? __builtin_nontemporal_load(keys + e) : 0u;
            vn[u] = e < ne1 ? __builtin_nontemporal_load(vals + e) : 0.0;
        }
    };
    // ys[row] += v * x[col] for the batch of entries eb + tid + u * kThreads
    auto add = [&](const unsigned *kb, const double *vb, long eb, long em, long e1) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long e = eb + tid + (long)u * kThreads;
            if (e < e1) {
                const int row = (int)(kb[u] >> kTileColsLog);
                int col = (int)(kb[u] & (kTileCols - 1));
                if (TS == 2 && e >= em) {
                    col += kTileCols;
                }
#pragma unroll
                for (int c = 0; c < KC; ++c) {
                    atomicAdd(&ys[row * KC + c], vb[u] * xs[col * KC + c]);
                }
            }
        }
    };
    if (nst > 0) {
        prefetch(0);
    }
    for (int j = 0; j < nst; ++j) {
        const long e0 = ne0, em = nem, e1 = ne1;
        unsigned kc[U];
        double vc[U];
#pragma unroll
        for (int q = 0; q < XPT; ++q) {
            xs[tid + q * kThreads] = xr[q];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            kc[u] = kn[u];
            vc[u] = vn[u];
        }
        __syncthreads();
        if (j + 1 < nst) {
            prefetch(j + 1);
        }
        add(kc, vc, e0, em, e1);
        for (long eb = e0 + BATCH; eb < e1; eb += BATCH) {
            unsigned kb[U];
            double vb[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const long e = eb + tid + (long)u * kThreads;
                kb[u] = e < e1 ? __builtin_nontemporal_load(keys + e) : 0u;
                vb[u] = e < e1 ? __builtin_nontemporal_load(vals + e) : 0.0;
            }
            add(kb, vb, eb, em, e1);
        }
        __syncthreads();
    }
    for (int i = tid; i < kGroupRows; i += kThreads) {
        const long r = (long)g * kGroupRows + i;
        if (r < n) {
#pragma unroll
            for (int c = 0; c < KC; ++c) {
                y[r * yrs + c * ycs] = ys[i * KC + c];
            }
        }
    }
}

}  // namespace

int csr_tiles_create(Context *ctx, const int *row_ptr, const int *col_ind, const double *vals, int n, long long nnz,
                     CsrTiles *out)
{
    if (n <= 0 || nnz < 0 || (n + kTileCols - 1) / kTileCols > kMaxTiles) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    out->n = n;
    out->groups = (n + kGroupRows - 1) / kGroupRows;
    out->tiles = (n + kTileCols - 1) / kTileCols;
    const long nb = (long)out->groups * out->tiles;
    DevBuf<long> count;
    DevBuf<int> bad;
    IPT_STCHK(count.alloc((size_t)nb));
    IPT_STCHK(bad.alloc(1));
    IPT_STCHK(out->start.alloc((size_t)nb + 1));
    IPT_STCHK(out->keys.alloc((size_t)std::max(1LL, nnz)));
    IPT_STCHK(out->vals.alloc((size_t)std::max(1LL, nnz)));
    IPT_HIPCHK(hipMemsetAsync(bad.p, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL(k_tiles_build<0>, dim3(out->groups), dim3(kThreads), 0, ctx->stream, row_ptr, col_ind, vals,
                       n, out->tiles, count.p, (const long *)nullptr, (unsigned *)nullptr, (double *)nullptr, bad.p);
    hipLaunchKernelGGL(k_tiles_offsets, dim3(1), dim3(kThreads), 0, ctx->stream, count.p, nb, out->start.p);
    IPT_HIPCHK(hipGetLastError());
    // the scatter trusts the counts: columns in range, and the counts add
    // up to the entries (row pointers from 0), checked before pass 1
    int hb[1] = {0};
    long total = 0;
    IPT_HIPCHK(hipMemcpyAsync(hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    IPT_HIPCHK(hipMemcpyAsync(&total, out->start.p + nb, sizeof(long), hipMemcpyDeviceToHost, ctx->stream));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    if (hb[0] != 0 || total != nnz) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    hipLaunchKernelGGL(k_tiles_build<1>, dim3(out->groups), dim3(kThreads), 0, ctx->stream, row_ptr, col_ind, vals,
                       n, out->tiles, count.p, out->start.p, out->keys.p, out->vals.p, (int *)nullptr);
    IPT_STCHK(out->ltile.alloc((size_t)nb));
    IPT_STCHK(out->lstart.alloc((size_t)nb + out->groups));
    IPT_STCHK(out->glen.alloc((size_t)out->groups));
    hipLaunchKernelGGL(k_tiles_lists, dim3((out->groups + 255) / 256), dim3(256), 0, ctx->stream,
                       (const long *)out->start.p, out->groups, out->tiles, out->ltile.p, out->lstart.p, out->glen.p);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

int spmm_csr_tiles(Context *ctx, const CsrTiles &tl, const double *X, long xrs, long xcs, double *Y, long yrs,
                   long ycs, int k)
{
    const ProfileToken prof = profile_begin_launches(PROF_SPMM_TILES);
    for (int c = 0; c < k; c += 2) {
        const bool first = c == 0, last = c + 2 >= k;
        if (k - c >= 2) {
            launch_timed(prof, first, last, k_spmm_tiles<2>, dim3(tl.groups), dim3(kThreads), 0, ctx->stream,
                         (const int *)tl.ltile.p, (const long *)tl.lstart.p, (const int *)tl.glen.p,
                         (const unsigned *)tl.keys.p, (const double *)tl.vals.p, tl.n, tl.tiles, X + c * xcs, xrs,
                         xcs, Y + c * ycs, yrs, ycs);
        } else {
            launch_timed(prof, first, last, k_spmm_tiles<1>, dim3(tl.groups), dim3(kThreads), 0, ctx->stream,
                         (const int *)tl.ltile.p, (const long *)tl.lstart.p, (const int *)tl.glen.p,
                         (const unsigned *)tl.keys.p, (const double *)tl.vals.p, tl.n, tl.tiles, X + c * xcs, xrs,
                         xcs, Y + c * ycs, yrs, ycs);
        }
    }
    profile_end(ctx, prof, k);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

}  // namespace ipt

using namespace ipt;

extern "C" int ipt_hip_csr_spmm_tiled(const int *d_row_ptr, const int *d_col_ind, const double *d_values, int n,
                                      const double *d_X, int ldx, double *d_Y, int ldy, int k, void *handle)
{
    if (d_row_ptr == nullptr || d_col_ind == nullptr || d_values == nullptr || d_X == nullptr || d_Y == nullptr ||
        n <= 0 || k <= 0 || ldx < k || ldy < k) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    int nnz = 0;
    IPT_HIPCHK(hipMemcpy(&nnz, d_row_ptr + n, sizeof(int), hipMemcpyDeviceToHost));
    CsrTiles tl;
    IPT_STCHK(csr_tiles_create(ctx, d_row_ptr, d_col_ind, d_values, n, nnz, &tl));
    IPT_STCHK(spmm_csr_tiles(ctx, tl, d_X, ldx, 1, d_Y, ldy, 1, k));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return IPT_CUDA_SUCCESS;
}
