// Elementwise / reduction kernels of the plain IPT path and the CSR SpMM.
#include <algorithm>

#include "ipt_internal.h"

namespace ipt {

static inline int blocks_for(long total, int threads, int cap)
{
    long b = (total + threads - 1) / threads;
    if (b < 1) {
        b = 1;
    }
    return (int)(b > cap ? cap : b);
}

// ------------------------------------------------------------------ basics

__global__ void k_dense_diag(const double *__restrict__ H, long ldh, int n, double *__restrict__ d)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        d[i] = H[i + i * ldh];
    }
}

int dense_diagonal(Context *ctx, const double *H, long ldh, int n, double *d)
{
    hipLaunchKernelGGL(k_dense_diag, dim3(blocks_for(n, 256, 4096)), dim3(256), 0, ctx->stream, H,
                       ldh, n, d);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

__global__ void k_fill(double *__restrict__ x, long count, double v)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x) {
        x[i] = v;
    }
}

int fill_device(Context *ctx, double *x, long count, double v)
{
    if (count <= 0) {
        return IPT_CUDA_SUCCESS;
    }
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(count, 256, 4096)), dim3(256), 0, ctx->stream, x, count, v);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

// out[s] = max(0, colsum[p] - |d[p]|), p = perm[s] (perm NULL: p = s): the
// off-diagonal row sums of |H| of a symmetric matrix in sorted row order
__global__ void k_offdiag_row_sums(const double *__restrict__ colsum, const double *__restrict__ d,
                                   const int *__restrict__ perm, int n, double *__restrict__ out)
{
    for (long s = (long)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (long)gridDim.x * blockDim.x) {
        const long p = perm != nullptr ? perm[s] : s;
        out[s] = fmax(0.0, colsum[p] - fabs(d[p]));
    }
}

int offdiag_row_sums(Context *ctx, const double *colsum, const double *d, const int *perm, int n, double *out)
{
    hipLaunchKernelGGL(k_offdiag_row_sums, dim3(blocks_for(n, 256, 4096)), dim3(256), 0, ctx->stream, colsum, d,
                       perm, n, out);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

// first stored entry with row == col, IPT/src/ipt_cuda.cu:1469-1489
__global__ void k_csc_diag(const int *__restrict__ cp, const int *__restrict__ ri,
                           const double *__restrict__ v, int n, double *__restrict__ d)
{
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x) {
        double x = 0.0;
        for (int p = cp[c]; p < cp[c + 1]; ++p) {
            if (ri[p] == c) {
                x = v[p];
                break;
            }
        }
        d[c] = x;
    }
}

int csc_diagonal(Context *ctx, const int *col_ptr, const int *row_ind, const double *values, int n,
                 double *d)
{
    hipLaunchKernelGGL(k_csc_diag, dim3(blocks_for(n, 256, 4096)), dim3(256), 0, ctx->stream,
                       col_ptr, row_ind, values, n, d);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

__global__ void k_identity(double *__restrict__ Z, long ld, int n, int k)
{
    const long total = (long)n * k;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)gridDim.x * blockDim.x) {
        const long r = idx % n;
        const long c = idx / n;
        Z[r + c * ld] = (r == c) ? 1.0 : 0.0;
    }
}

int set_identity(Context *ctx, double *Z, long ld, int n, int k)
{
    hipLaunchKernelGGL(k_identity, dim3(blocks_for((long)n * k, 256, 16384)), dim3(256), 0,
                       ctx->stream, Z, ld, n, k);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

// Upper and lower triangle replaced by their average (one thread per pair).
__global__ void k_symmetrize(double *A, long ld, int n)
{
    const long total = (long)n * n;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const long r = t % n, c = t / n;
        if (r < c) {
            const double avg = 0.5 * (A[r + c * ld] + A[c + r * ld]);
            A[r + c * ld] = avg;
            A[c + r * ld] = avg;
        }
    }
}

int symmetrize(Context *ctx, double *A, long ld, int n)
{
    hipLaunchKernelGGL(k_symmetrize, dim3(blocks_for((long)n * n, 256, 4096)), dim3(256), 0, ctx->stream, A, ld, n);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

// ------------------------------------------------------------------ plain update

// Block-wide sum of two doubles; result valid in thread 0.
__device__ inline void block_sum2(double &a, double &b)
{
    __shared__ double sa[16], sb[16];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sa[w] = a;
        sb[w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        a = 0.0;
        b = 0.0;
        for (int i = 0; i < nw; ++i) {
            a += sa[i];
            b += sb[i];
        }
    }
}

// One plain IPT step, IPT/src/ipt_cuda.cu:1403-1443 fused with 1457-1467:
//   c_j   = (HZ)_jj - d_j Z_jj
//   Z'_rj = (Y_rj - d_r Z_rj - Z_rj c_j) * G_rj,  G_rj = 1/(d_j - d_r)
//           (NaN when the gap is zero or not finite, 1369-1396 with min_gap 0)
//   Z'_jj = 1
// and per-block partial sums of |Z'-Z|^2, |Z|^2.
__global__ __launch_bounds__(256) void k_plain_update(const double *Y, const double *__restrict__ Z,
                                                      double *Znew, const double *__restrict__ d,
                                                      int n, int k, long rs, long cs, int rowmajor,
                                                      double *__restrict__ partials)
{
    const long total = (long)n * k;
    double sdn = 0.0, szn = 0.0;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)gridDim.x * blockDim.x) {
        long r, j;
        if (rowmajor) {
            r = idx / k;
            j = idx % k;
        } else {
            r = idx % n;
            j = idx / n;
        }
        const long off = r * rs + j * cs;
        const long dj_off = j * rs + j * cs;
        const double z = Z[off];
        const double dj = d[j];
        double nz;
        if (r == j) {
            nz = 1.0;
        } else {
            const double cj = Y[dj_off] - dj * Z[dj_off];
            const double gap = dj - d[r];
            const double g = (isfinite(gap) && gap != 0.0) ? 1.0 / gap : (double)NAN;
            nz = (Y[off] - d[r] * z - z * cj) * g;
        }
        const double dl = nz - z;
        sdn += dl * dl;
        szn += z * z;
        // Y(j,j) is read by every thread of column j: write after the grid
        // finished reading would need a barrier, so the caller never aliases
        // Znew with Y.
        Znew[off] = nz;
    }
    block_sum2(sdn, szn);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = sdn;
        partials[2 * blockIdx.x + 1] = szn;
    }
}

__global__ void k_sum_pairs(const double *__restrict__ partials, int count, double *__restrict__ out)
{
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
        a += partials[2 * i];
        b += partials[2 * i + 1];
    }
    block_sum2(a, b);
    if (threadIdx.x == 0) {
        out[0] = a;
        out[1] = b;
    }
}

int plain_update(Context *ctx, const double *Y, const double *Z, double *Znew, const double *d,
                 int n, int k, long rs, long cs, double *d_norms)
{
    const int blocks = blocks_for((long)n * k, 256, 2048);
    double *partials = nullptr;
    IPT_STCHK(context_scratch(ctx, SLOT_REDUCE, (size_t)blocks * 2 * sizeof(double), (void **)&partials));
    const int rowmajor = (cs == 1 && k > 1) ? 1 : 0;
    hipLaunchKernelGGL(k_plain_update, dim3(blocks), dim3(256), 0, ctx->stream, Y, Z, Znew, d, n, k,
                       rs, cs, rowmajor, partials);
    IPT_HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(256), 0, ctx->stream, partials, blocks, d_norms);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

__global__ void k_gather_diag(const double *__restrict__ Y, int k, long rs, long cs, double *__restrict__ v)
{
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < k; j += gridDim.x * blockDim.x) {
        v[j] = Y[(long)j * rs + (long)j * cs];
    }
}

int gather_diagonal(Context *ctx, const double *Y, int k, long rs, long cs, double *values)
{
    hipLaunchKernelGGL(k_gather_diag, dim3(blocks_for(k, 256, 64)), dim3(256), 0, ctx->stream, Y, k,
                       rs, cs, values);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

// ------------------------------------------------------------------ SpMM

// Wavefront-per-row CSR SpMM with row-major X / Y.  The 64 lanes are split
// into G = 64/S groups of S lanes (S = next power of two >= k, capped at
// 64); lane s of a group accumulates column s (and s+S, ...).  The row's
// (column, value) pairs are read 64 at a time, one per lane (coalesced),
// and handed to the groups by cross-lane shuffles; each step gathers G
// contiguous rows of X, and the unrolled steps keep several gathers in
// flight per lane -- the gathers, not the compulsory bytes, bound this
// kernel when X (n x k FP64) is larger than the L2.
//
// Row r's entries are [rs[r], re[r]): rs = row_ptr, re = row_ptr + 1 for
// the whole matrix, or the bounds of one column slice (CsrSlices); with
// accumulate the products are added to Y.
template <int S>
__global__ __launch_bounds__(256) void k_spmm_rowmajor(const int *__restrict__ rs, const int *__restrict__ re,
                                                       const int *__restrict__ ci,
                                                       const double *__restrict__ vals, int n,
                                                       const double *__restrict__ X, long ldx,
                                                       double *__restrict__ Y, long ldy, int k, int accumulate)
{
    constexpr int G = 64 / S;
    constexpr int STEPS = 64 / G;
    const int lane = threadIdx.x & 63;
    const int g = lane / S;
    const int s = lane % S;
    const long wave_global = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long waves = ((long)gridDim.x * blockDim.x) >> 6;
    for (long row = wave_global; row < n; row += waves) {
        const int pb = rs[row], pe = re[row];
        for (int c0 = 0; c0 < k; c0 += S) {
            const int col = c0 + s;
            const bool active = col < k;
            const double *xc = X + (active ? col : 0);
            double acc = 0.0;
            for (int base = pb; base < pe; base += 64) {
                const int cnt = min(64, pe - base);
                // padding entries: column 0 (a valid row of X), value 0
                const int myc = lane < cnt ? ci[base + lane] : 0;
                const double myv = lane < cnt ? vals[base + lane] : 0.0;
                // U gathers issued back to back, then their FMAs; the step
                // count is padded to a multiple of U (U * G divides 64)
                constexpr int U = STEPS < 8 ? STEPS : 8;
                const int steps = ((cnt + G * U - 1) / (G * U)) * U;
                for (int j0 = 0; j0 < steps; j0 += U) {
                    int c[U];
                    double v[U], x[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        c[u] = __shfl(myc, (j0 + u) * G + g, 64);
                        v[u] = __shfl(myv, (j0 + u) * G + g, 64);
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        x[u] = xc[(long)c[u] * ldx];
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        acc += v[u] * x[u];
                    }
                }
            }
#pragma unroll
            for (int off = S; off < 64; off <<= 1) {
                acc += __shfl_xor(acc, off, 64);
            }
            if (g == 0 && active) {
                if (accumulate) {
                    Y[row * ldy + col] += acc;
                } else {
                    Y[row * ldy + col] = acc;
                }
            }
        }
    }
}

// one pass of the product; first / last: the pass opens / closes the timed span
static void launch_spmm(const ProfileToken &prof, bool first, bool last, hipStream_t st, const int *rs,
                        const int *re, const int *col_ind, const double *vals, int n, const double *X, long ldx,
                        double *Y, long ldy, int k, int accumulate)
{
    const int blocks = blocks_for((long)n * 64, 256, 65536);
    dim3 g(blocks), b(256);
    if (k <= 1) {
        launch_timed(prof, first, last, k_spmm_rowmajor<1>, g, b, 0, st, rs, re, col_ind, vals, n, X, ldx, Y, ldy,
                     k, accumulate);
    } else if (k <= 2) {
        launch_timed(prof, first, last, k_spmm_rowmajor<2>, g, b, 0, st, rs, re, col_ind, vals, n, X, ldx, Y, ldy,
                     k, accumulate);
    } else if (k <= 4) {
        launch_timed(prof, first, last, k_spmm_rowmajor<4>, g, b, 0, st, rs, re, col_ind, vals, n, X, ldx, Y, ldy,
                     k, accumulate);
    } else if (k <= 8) {
        launch_timed(prof, first, last, k_spmm_rowmajor<8>, g, b, 0, st, rs, re, col_ind, vals, n, X, ldx, Y, ldy,
                     k, accumulate);
    } else if (k <= 16) {
        launch_timed(prof, first, last, k_spmm_rowmajor<16>, g, b, 0, st, rs, re, col_ind, vals, n, X, ldx, Y, ldy,
                     k, accumulate);
    } else if (k <= 32) {
        launch_timed(prof, first, last, k_spmm_rowmajor<32>, g, b, 0, st, rs, re, col_ind, vals, n, X, ldx, Y, ldy,
                     k, accumulate);
    } else {
        launch_timed(prof, first, last, k_spmm_rowmajor<64>, g, b, 0, st, rs, re, col_ind, vals, n, X, ldx, Y, ldy,
                     k, accumulate);
    }
}

int spmm_csr_sliced_passes(Context *ctx, const int *row_ptr, const int *col_ind, const double *vals, int n,
                           const double *X, long ldx, double *Y, long ldy, int k, const CsrSlices &slices, int ns)
{
    const ProfileToken prof = profile_begin_launches(PROF_SPMM);
    if (ns > 1) {
        // column slices of the same CSR, one pass each: a pass gathers
        // rows of X from one slice (an L2-sized part of X), the later
        // passes add to Y
        const int *sp = slices.split.p;
        const long stride = (long)slices.n;
        const int step = slices.nslices / ns;
        for (int b = 0; b < ns; ++b) {
            launch_spmm(prof, b == 0, b == ns - 1, ctx->stream, sp + (long)(b * step) * stride,
                        sp + (long)((b + 1) * step) * stride, col_ind, vals, n, X, ldx, Y, ldy, k, b > 0);
        }
    } else {
        launch_spmm(prof, true, true, ctx->stream, row_ptr, row_ptr + 1, col_ind, vals, n, X, ldx, Y, ldy, k, 0);
    }
    profile_end(ctx, prof, k);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

int spmm_csr_rowmajor(Context *ctx, const int *row_ptr, const int *col_ind, const double *vals,
                      int n, const double *X, long ldx, double *Y, long ldy, int k, const CsrSlices *slices)
{
    if (slices == nullptr) {
        const CsrSlices none;
        return spmm_csr_sliced_passes(ctx, row_ptr, col_ind, vals, n, X, ldx, Y, ldy, k, none, 1);
    }
    return spmm_csr_sliced_passes(ctx, row_ptr, col_ind, vals, n, X, ldx, Y, ldy, k, *slices,
                                  csr_slices_for(*slices, k));
}

// split[b n + r] = first entry of row r whose column is >= b * width
// (width = ceil(n / nslices)), split[0] = row_ptr, split[nslices] =
// row_ptr + 1; columns ascend within each row.  One thread per row.
__global__ __launch_bounds__(256) void k_csr_split(const int *__restrict__ rp, const int *__restrict__ ci, int n,
                                                   int nslices, int width, int *__restrict__ split)
{
    for (long r = (long)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x) {
        const int pb = rp[r], pe = rp[r + 1];
        split[r] = pb;
        split[(long)nslices * n + r] = pe;
        int lo = pb;
        for (int b = 1; b < nslices; ++b) {
            const int c = b * width;
            int hi = pe;
            while (lo < hi) {  // first index in [lo, pe) with ci >= c
                const int mid = (lo + hi) >> 1;
                if (ci[mid] < c) {
                    lo = mid + 1;
                } else {
                    hi = mid;
                }
            }
            split[(long)b * n + r] = lo;
        }
    }
}

int csr_slices_create(Context *ctx, const int *row_ptr, const int *col_ind, int n, CsrSlices *sl)
{
    sl->n = n;
    sl->nslices = kCsrMaxSlices;
    IPT_STCHK(sl->split.alloc((size_t)(kCsrMaxSlices + 1) * n));
    const int width = (n + kCsrMaxSlices - 1) / kCsrMaxSlices;
    hipLaunchKernelGGL(k_csr_split, dim3(blocks_for(n, 256, 8192)), dim3(256), 0, ctx->stream, row_ptr, col_ind, n,
                       kCsrMaxSlices, width, sl->split.p);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

int csr_slices_for(const CsrSlices &sl, int k)
{
    // k <= 8 and X beyond 2 MiB: 4 or 8 passes, at most about 4 MiB of X
    // per pass (an XCD's L2) where 8 passes allow it; 9 <= k <= 32 and X
    // beyond 16 MiB: 8 passes, 16 from 17 columns.  Measured on the N = 1e6
    // orbital model (200 draws per row, X rows at csr_row_stride): the best
    // counts were 4 at k = 1 and 2, 8 at k = 4 to 16 (k = 16: 6.49 ms
    // against 7.58 in one pass and 7.53 in 16), 16 at k = 24 to 32 (k = 27:
    // 12.86 ms against 13.17 in 8; k = 32: 13.02 against 13.21); more passes
    // re-read Y and the row bounds more often
    const long xbytes = (long)sl.n * k * 8;
    if (k > 32 || xbytes <= (2L << 20) || sl.nslices < 8) {
        return 1;
    }
    if (k > 8) {
        return xbytes <= (16L << 20) ? 1 : k > 16 && sl.nslices >= 16 ? 16 : 8;
    }
    return xbytes > (16L << 20) ? 8 : 4;
}

long csr_row_stride(int w, int elem)
{
    // 128-byte lines a row of w elements of elem bytes touches, averaged
    // over the row offsets that stride ld produces (they repeat every 128 /
    // gcd(elem ld, 128) rows); the smallest ld with the fewest lines.  w = 12: ld 16,
    // one line instead of 1.5 (10.5 -> 7.6 ms at N = 1e6); w = 27: ld 32,
    // 2 lines instead of 2.6 (18.2 -> 14.0 ms)
    if (w <= 1) {
        return 1;
    }
    auto lines = [w, elem](long ld) {
        const long stride = ld * elem, bytes = (long)w * elem;
        long g = 128, a = stride % 128;
        while (a != 0) {
            const long t = g % a;
            g = a;
            a = t;
        }
        const long period = 128 / g;
        long total = 0;
        for (long r = 0; r < period; ++r) {
            const long off = (r * stride) % 128;
            total += (off + bytes - 1) / 128 + 1;
        }
        return (double)total / (double)period;
    };
    long p2 = 1;
    while (p2 < w) {
        p2 <<= 1;
    }
    const long per_line = 128 / elem;
    const long cands[3] = {(long)w, p2, ((long)w + per_line - 1) / per_line * per_line};
    long best = w;
    double best_lines = lines(w);
    for (long c : cands) {
        const double l = lines(c);
        if (l < best_lines - 1e-9 || (l < best_lines + 1e-9 && c < best)) {
            best = c;
            best_lines = l;
        }
    }
    return best;
}

// ------------------------------------------------------------------ layouts

// Layout changes between column-major (ld >= n) and row-major (ld >= k)
// n x k blocks, k <= 128 per launch column chunk: 32-row tiles staged in
// LDS so both the column-major side (32 consecutive rows of a column) and
// the row-major side (whole rows) are accessed contiguously.
constexpr int kTileRows = 32;
constexpr int kTileCols = 128;

__global__ __launch_bounds__(256) void k_col_to_row(const double *__restrict__ A, long lda,
                                                    double *__restrict__ B, long ldb, int n, int k)
{
    __shared__ double t[kTileRows][kTileCols + 1];
    const int c0 = blockIdx.y * kTileCols;
    const int kc = min(kTileCols, k - c0);
    for (long r0 = (long)blockIdx.x * kTileRows; r0 < n; r0 += (long)gridDim.x * kTileRows) {
        const int rows = (int)min((long)kTileRows, n - r0);
        // read: lane -> row within the tile, wave-strided over columns
        for (int e = threadIdx.x; e < kTileRows * kc; e += 256) {
            const int rr = e % kTileRows, cc = e / kTileRows;
            if (rr < rows) {
                t[rr][cc] = A[r0 + rr + (long)(c0 + cc) * lda];
            }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < rows * kc; e += 256) {
            const int rr = e / kc, cc = e % kc;
            B[(r0 + rr) * ldb + c0 + cc] = t[rr][cc];
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_row_to_col(const double *__restrict__ A, long lda,
                                                    double *__restrict__ B, long ldb, int n, int k)
{
    __shared__ double t[kTileRows][kTileCols + 1];
    const int c0 = blockIdx.y * kTileCols;
    const int kc = min(kTileCols, k - c0);
    for (long r0 = (long)blockIdx.x * kTileRows; r0 < n; r0 += (long)gridDim.x * kTileRows) {
        const int rows = (int)min((long)kTileRows, n - r0);
        for (int e = threadIdx.x; e < rows * kc; e += 256) {
            const int rr = e / kc, cc = e % kc;
            t[rr][cc] = A[(r0 + rr) * lda + c0 + cc];
        }
        __syncthreads();
        for (int e = threadIdx.x; e < kTileRows * kc; e += 256) {
            const int rr = e % kTileRows, cc = e / kTileRows;
            if (rr < rows) {
                B[r0 + rr + (long)(c0 + cc) * ldb] = t[rr][cc];
            }
        }
        __syncthreads();
    }
}

static dim3 tile_grid(int n, int k)
{
    const long tiles = ((long)n + kTileRows - 1) / kTileRows;
    return dim3((unsigned)std::min<long>(tiles, 16384), (unsigned)((k + kTileCols - 1) / kTileCols));
}

int colmajor_to_rowmajor(Context *ctx, const double *A, long lda, double *B, long ldb, int n, int k)
{
    if (n <= 0 || k <= 0) {
        return IPT_CUDA_SUCCESS;
    }
    hipLaunchKernelGGL(k_col_to_row, tile_grid(n, k), dim3(256), 0, ctx->stream, A, lda, B, ldb, n, k);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

int rowmajor_to_colmajor(Context *ctx, const double *A, long lda, double *B, long ldb, int n, int k)
{
    if (n <= 0 || k <= 0) {
        return IPT_CUDA_SUCCESS;
    }
    hipLaunchKernelGGL(k_row_to_col, tile_grid(n, k), dim3(256), 0, ctx->stream, A, lda, B, ldb, n, k);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

}  // namespace ipt

using namespace ipt;

extern "C" long ipt_hip_csr_row_stride(int w, int elem)
{
    if (w <= 0 || (elem != 4 && elem != 8)) {
        return -1;
    }
    return csr_row_stride(w, elem);
}
