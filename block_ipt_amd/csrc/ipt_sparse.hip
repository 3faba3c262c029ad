// Device-side preparation of a symmetric sparse matrix for the block path
// (the CSC of a symmetric matrix is also its CSR): diagonal, absolute row
// sums and largest off-diagonal magnitude per row, structure checks, the
// symmetry check, and lookups of single entries for cluster detection.
// Host counterparts: IPT/src/ipt_cuda.cu:4551-4683 (block preparation),
// IPT/src/ipt_preparation.cu:148-169 (CSC validation), 860-883 (symmetry).
#include <algorithm>

#include "ipt_internal.h"
#include "../../include/ipt_hip_ext.h"

namespace ipt {

namespace {

// first position in the ascending row segment [lo, hi) whose column is >= c
// (hi if none)
__device__ inline int find_in_row(const int *__restrict__ ci, int lo, int hi, int c)
{
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int x = ci[mid];
        if (x < c) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    return lo;
}

// One wavefront per row.
__global__ __launch_bounds__(256) void k_csr_scan(const int *__restrict__ rp, const int *__restrict__ ci,
                                                  const double *__restrict__ v, int n, double *__restrict__ diag,
                                                  double *__restrict__ abssum, double *__restrict__ offmax,
                                                  int *__restrict__ flags)
{
    const int lane = threadIdx.x & 63;
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long waves = ((long)gridDim.x * blockDim.x) >> 6;
    int bad = 0;
    for (long row = wave; row < n; row += waves) {
        const int pb = rp[row], pe = rp[row + 1];
        double dg = 0.0, s = 0.0, om = 0.0;
        for (int p = pb + lane; p < pe; p += 64) {
            const int c = ci[p];
            const double x = v[p];
            if (c < 0 || c >= n) {
                bad |= 1;
                continue;
            }
            if (p > pb && ci[p - 1] >= c) {
                bad |= 2;  // not strictly ascending
            }
            s += fabs(x);
            if (c == row) {
                dg += x;
            } else {
                om = fmax(om, fabs(x));
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            dg += __shfl_xor(dg, o, 64);
            s += __shfl_xor(s, o, 64);
            om = fmax(om, __shfl_xor(om, o, 64));
        }
        if (lane == 0) {
            diag[row] = dg;
            abssum[row] = s;
            offmax[row] = om;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        bad |= __shfl_xor(bad, o, 64);
    }
    if (lane == 0 && bad) {
        atomicOr(flags, bad);
    }
}

__device__ inline unsigned long long mix64(unsigned long long x)
{
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

__device__ inline unsigned long long entry_hash(unsigned r, unsigned c, double v)
{
    return mix64(mix64(((unsigned long long)r << 32) | c) ^ (unsigned long long)__double_as_longlong(v));
}

// Exact-symmetry fingerprint in one streaming pass: sums[0] = sum of
// hash(r, c, v) over the stored entries above the diagonal (r < c),
// sums[1] = sum of hash(c, r, v) over those below it (wrapping 64-bit sums,
// order independent, one hash per entry).  A symmetric matrix has equal
// sums (each (r, c, v) above has its (c, r, v) below); unequal sums send
// the caller to the tolerance check below.
__global__ __launch_bounds__(256) void k_csr_symhash(const int *__restrict__ rp, const int *__restrict__ ci,
                                                     const double *__restrict__ v, int n,
                                                     unsigned long long *__restrict__ sums)
{
    const int lane = threadIdx.x & 63;
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long waves = ((long)gridDim.x * blockDim.x) >> 6;
    unsigned long long a = 0, b = 0;
    for (long row = wave; row < n; row += waves) {
        const int pb = rp[row], pe = rp[row + 1];
        for (int p = pb + lane; p < pe; p += 64) {
            const unsigned c = (unsigned)ci[p];
            const unsigned r = (unsigned)row;
            if (r != c) {
                const unsigned long long h = entry_hash(r < c ? r : c, r < c ? c : r, v[p]);
                if (r < c) {
                    a += h;
                } else {
                    b += h;
                }
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    // one pair of atomics per workgroup: the two sums are single addresses,
    // so every atomic on them serialises
    __shared__ unsigned long long wa[4], wb[4];
    const int w = threadIdx.x >> 6;
    if (lane == 0) {
        wa[w] = a;
        wb[w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&sums[0], wa[0] + wa[1] + wa[2] + wa[3]);
        atomicAdd(&sums[1], wb[0] + wb[1] + wb[2] + wb[3]);
    }
}

// |H_rc - H_cr| <= 1e-12 max(1, |H_rc|, |H_cr|) for every stored entry; the
// mirror is found by binary search in the (ascending) row c.
__global__ __launch_bounds__(256) void k_csr_symcheck(const int *__restrict__ rp, const int *__restrict__ ci,
                                                      const double *__restrict__ v, int n, int *__restrict__ flags)
{
    const int lane = threadIdx.x & 63;
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long waves = ((long)gridDim.x * blockDim.x) >> 6;
    int bad = 0;
    for (long row = wave; row < n; row += waves) {
        const int pb = rp[row], pe = rp[row + 1];
        for (int p = pb + lane; p < pe; p += 64) {
            const int c = ci[p];
            if (c == row) {
                continue;
            }
            const int lo = rp[c], hi = rp[c + 1];
            const int q = find_in_row(ci, lo, hi, (int)row);
            const double m = (q < hi && ci[q] == row) ? v[q] : 0.0;
            const double x = v[p];
            if (!(fabs(x - m) <= 1e-12 * fmax(1.0, fmax(fabs(x), fabs(m))))) {
                bad |= 4;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        bad |= __shfl_xor(bad, o, 64);
    }
    if (lane == 0 && bad) {
        atomicOr(flags, bad);
    }
}

// out[2t] = H(r, c), out[2t+1] = H(c, r) for pairs rc[2t], rc[2t+1]
__global__ void k_csr_gather_pairs(const int *__restrict__ rp, const int *__restrict__ ci,
                                   const double *__restrict__ v, const int *__restrict__ rc, int npairs,
                                   double *__restrict__ out)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < 2L * npairs; t += (long)gridDim.x * blockDim.x) {
        const long pr = t >> 1;
        const int r = rc[2 * pr + (t & 1)], c = rc[2 * pr + 1 - (t & 1)];
        const int lo = rp[r], hi = rp[r + 1];
        const int q = find_in_row(ci, lo, hi, c);
        out[t] = (q < hi && ci[q] == c) ? v[q] : 0.0;
    }
}

int grid_waves(long rows)
{
    long b = (rows + 3) / 4;
    return (int)std::max(1L, std::min(b, 65536L));
}

}  // namespace

int csr_scan(Context *ctx, const int *rp, const int *ci, const double *v, int n, double *diag, double *abssum,
             double *offmax, int *flags)
{
    hipLaunchKernelGGL(k_csr_scan, dim3(grid_waves(n)), dim3(256), 0, ctx->stream, rp, ci, v, n, diag, abssum, offmax,
                       flags);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

int csr_symhash(Context *ctx, const int *rp, const int *ci, const double *v, int n, unsigned long long *sums)
{
    // a grid of 16 workgroups per CU walking the rows: 2 x 4096 atomics
    // instead of two per row (at N = 1e6, nnz 4e8: 6.4 ms with two per row,
    // 2.0 ms with 1024 workgroups, 1.3 ms with 4096)
    hipLaunchKernelGGL(k_csr_symhash, dim3((unsigned)std::min(grid_waves(n), 4096)), dim3(256), 0, ctx->stream, rp, ci,
                       v, n, sums);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

int csr_symcheck(Context *ctx, const int *rp, const int *ci, const double *v, int n, int *flags)
{
    hipLaunchKernelGGL(k_csr_symcheck, dim3(grid_waves(n)), dim3(256), 0, ctx->stream, rp, ci, v, n, flags);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

int csr_gather_pairs(Context *ctx, const int *rp, const int *ci, const double *v, const int *rc, int npairs,
                     double *out)
{
    if (npairs <= 0) {
        return IPT_CUDA_SUCCESS;
    }
    const long total = 2L * npairs;
    hipLaunchKernelGGL(k_csr_gather_pairs, dim3((unsigned)std::min<long>((total + 255) / 256, 16384)), dim3(256), 0,
                       ctx->stream, rp, ci, v, rc, npairs, out);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

}  // namespace ipt

using namespace ipt;

extern "C" int ipt_hip_csr_spmm(const int *d_row_ptr, const int *d_col_ind, const double *d_values, int n,
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
    IPT_STCHK(spmm_csr_rowmajor(ctx, d_row_ptr, d_col_ind, d_values, n, d_X, ldx, d_Y, ldy, k));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return IPT_CUDA_SUCCESS;
}

extern "C" int ipt_hip_csr_spmm_sliced(const int *d_row_ptr, const int *d_col_ind, const double *d_values, int n,
                                       const double *d_X, int ldx, double *d_Y, int ldy, int k, int slices,
                                       void *handle)
{
    if (d_row_ptr == nullptr || d_col_ind == nullptr || d_values == nullptr || d_X == nullptr || d_Y == nullptr ||
        n <= 0 || k <= 0 || ldx < k || ldy < k || slices < 0 || slices > kCsrMaxSlices ||
        (slices & (slices - 1)) != 0) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    CsrSlices sl;
    IPT_STCHK(csr_slices_create(ctx, d_row_ptr, d_col_ind, n, &sl));
    if (slices > 0) {
        // a fixed pass count: the slice table coarsened to `slices` passes
        IPT_STCHK(spmm_csr_sliced_passes(ctx, d_row_ptr, d_col_ind, d_values, n, d_X, ldx, d_Y, ldy, k, sl, slices));
    } else {
        IPT_STCHK(spmm_csr_rowmajor(ctx, d_row_ptr, d_col_ind, d_values, n, d_X, ldx, d_Y, ldy, k, &sl));
    }
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return IPT_CUDA_SUCCESS;
}
