// Synthetic sparse workload generator (device side): the orbital-sum
// configuration-interaction model used for the N = 1e6 sparse benchmark.
//
//  * `norb` orbital energies e_o ~ U[elo, ehi], sorted ascending.
//  * Row i's diagonal is the sum of a random distinct `nsub`-subset of the
//    orbitals (partial Fisher-Yates); the n diagonals are sorted ascending,
//    so row i holds the i-th smallest.
//  * Row i draws the candidate columns c_t = floor(u_t (n-1)), shifted past
//    i (t = 0, 1, ...), and keeps the first `m` distinct ones -- rejection
//    sampling of m distinct off-diagonal columns.  The draw at sequence
//    index t carries the value sigma * N(0,1) (Box-Muller on two uniforms).
//  * Symmetric by mirroring: the draw (i -> j) sets H_ij = H_ji.  When i drew
//    j and j drew i, the draw of the larger row wins for both entries.
//
// Every uniform is the splitmix64 counter stream of the dense generator
// (IPT/test/mixed_precision_tf32/benchmark_tf32_cusolver_sweep.cu:434-449),
// keyed by a per-purpose tag, so the matrix is a pure function of the
// parameters.  Output: CSR (= CSC, symmetric) with 32-bit indices, column
// indices ascending within each row.
//
// Construction, all on the device except the n diagonal values:
//   draw (wave per row) -> incoming counts (atomics) -> segment offsets
//   (scan) -> scatter own + mirrored entries -> per-row bitonic sort in LDS
//   by (column, source row), keep the last of each column run -> scan of the
//   kept counts -> compact into the caller's arrays.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <vector>

#include "ipt_internal.h"
#include "../../include/ipt_hip_ext.h"

namespace ipt {

namespace {

constexpr unsigned long long kTagOrbital = 0x6f72626974616c73ULL;
constexpr unsigned long long kTagSubset = 0x0073756273657473ULL;
constexpr unsigned long long kTagColumn = 0x00636f6c756d6e73ULL;
constexpr unsigned long long kTagValue = 0x000076616c756573ULL;
constexpr int kMaxDraws = 512;     // m
constexpr int kMaxSegment = 2048;  // entries of one row before deduplication (LDS sort)

__host__ __device__ inline unsigned long long mix64(unsigned long long x)
{
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

__host__ __device__ inline double unif(unsigned long long seed, long long n, long long row, long long col)
{
    unsigned long long x = seed + (unsigned long long)n * 0xd1b54a32d192ed03ULL;
    x ^= (unsigned long long)(row + 1) * 0x9e3779b97f4a7c15ULL;
    x ^= (unsigned long long)(col + 1) * 0xbf58476d1ce4e5b9ULL;
    return (double)(mix64(x) >> 11) * (1.0 / 9007199254740992.0);
}

__device__ inline int candidate_column(unsigned long long seed, int n, int i, long long t)
{
    int c = (int)(unif(seed ^ kTagColumn, n, i, t) * (double)(n - 1));
    c = c > n - 2 ? n - 2 : c;
    return c >= i ? c + 1 : c;
}

__device__ inline double draw_value(unsigned long long seed, int n, int i, long long t, double sigma)
{
    const double u1 = unif(seed ^ kTagValue, n, i, 2 * t);
    const double u2 = unif(seed ^ kTagValue, n, i, 2 * t + 1);
    return sigma * sqrt(-2.0 * log1p(-u1)) * cos(6.283185307179586 * u2);
}

// One wavefront per row: 64 candidates per round, each checked against the
// accepted list and the earlier lanes of its round; the first-occurrence
// lanes are appended in sequence order until m columns are accepted.
__global__ __launch_bounds__(256) void k_orb_draw(int n, int m, unsigned long long seed, double sigma,
                                                  int *__restrict__ own_col, double *__restrict__ own_val,
                                                  int *__restrict__ incoming)
{
    __shared__ int accepted[4][kMaxDraws];
    __shared__ int round_cols[4][64];
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long waves = ((long)gridDim.x * blockDim.x) >> 6;
    for (long row = wave; row < n; row += waves) {
        const int i = (int)row;
        int cnt = 0;
        long long t0 = 0;
        while (cnt < m) {
            const long long t = t0 + lane;
            const int c = candidate_column(seed, n, i, t);
            round_cols[w][lane] = c;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            bool first = true;
            for (int q = 0; q < cnt; ++q) {
                first = first && accepted[w][q] != c;
            }
            for (int l = 0; l < lane; ++l) {
                first = first && round_cols[w][l] != c;
            }
            const unsigned long long firsts = __ballot(first);
            const int pos = cnt + __popcll(firsts & ((1ULL << lane) - 1ULL));
            if (first && pos < m) {
                accepted[w][pos] = c;
                own_col[row * m + pos] = c;
                own_val[row * m + pos] = draw_value(seed, n, i, t, sigma);
                atomicAdd(&incoming[c], 1);
            }
            cnt = min(m, cnt + __popcll(firsts));
            t0 += 64;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
}

__global__ void k_orb_seglen(int n, int m, const int *__restrict__ incoming, long long *__restrict__ len,
                             int *__restrict__ maxlen)
{
    int local = 0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int l = 1 + m + incoming[i];
        len[i] = l;
        local = max(local, l);
    }
    for (int o = 32; o > 0; o >>= 1) {
        local = max(local, __shfl_down(local, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(maxlen, local);
    }
}

// ---- exclusive scan of n int64 values into out[0..n] (out[n] = total)

constexpr int kScanBlock = 1024;  // elements per 256-thread block

__device__ inline long long block_inclusive_scan(long long v, long long *sh)
{
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        const long long u = __shfl_up(v, o, 64);
        if (lane >= o) {
            v += u;
        }
    }
    if (lane == 63) {
        sh[w] = v;
    }
    __syncthreads();
    long long add = 0;
    for (int q = 0; q < w; ++q) {
        add += sh[q];
    }
    __syncthreads();
    return v + add;
}

__global__ __launch_bounds__(256) void k_scan_local(const long long *__restrict__ in, long n,
                                                    long long *__restrict__ out, long long *__restrict__ sums)
{
    __shared__ long long sh[4];
    const long base = (long)blockIdx.x * kScanBlock + threadIdx.x * 4;
    long long v[4], s = 0;
    for (int j = 0; j < 4; ++j) {
        v[j] = base + j < n ? in[base + j] : 0;
        s += v[j];
    }
    const long long incl = block_inclusive_scan(s, sh);
    long long run = incl - s;
    for (int j = 0; j < 4; ++j) {
        if (base + j < n) {
            out[base + j] = run;
        }
        run += v[j];
    }
    if (threadIdx.x == 255) {
        sums[blockIdx.x] = incl;
    }
}

__global__ __launch_bounds__(256) void k_scan_sums(long long *__restrict__ sums, int nb, long long *__restrict__ total)
{
    __shared__ long long sh[4];
    long long carry = 0;
    for (int b0 = 0; b0 < nb; b0 += 256) {
        const int b = b0 + threadIdx.x;
        const long long v = b < nb ? sums[b] : 0;
        const long long incl = block_inclusive_scan(v, sh);
        if (b < nb) {
            sums[b] = carry + incl - v;
        }
        __shared__ long long last;
        if (threadIdx.x == 255) {
            last = incl;
        }
        __syncthreads();
        carry += last;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *total = carry;
    }
}

__global__ void k_scan_add(long long *__restrict__ out, long n, const long long *__restrict__ sums,
                           const long long *__restrict__ total)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += (long)gridDim.x * blockDim.x) {
        if (i == n) {
            out[n] = *total;
        } else {
            out[i] += sums[i / kScanBlock];
        }
    }
}

int exclusive_scan(Context *ctx, const long long *in, long n, long long *out)
{
    const int nb = (int)((n + kScanBlock - 1) / kScanBlock);
    DevBuf<long long> sums, total;
    IPT_STCHK(sums.alloc((size_t)std::max(nb, 1)));
    IPT_STCHK(total.alloc(1));
    hipLaunchKernelGGL(k_scan_local, dim3(std::max(nb, 1)), dim3(256), 0, ctx->stream, in, n, out, sums.p);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, ctx->stream, sums.p, nb, total.p);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)std::min<long>((n + 256) / 256 + 1, 8192)), dim3(256), 0,
                       ctx->stream, out, n, sums.p, total.p);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

// ---- scatter

__host__ __device__ inline unsigned long long seg_key(int col, int src)
{
    return ((unsigned long long)(unsigned)col << 32) | (unsigned)src;
}

__global__ void k_orb_scatter(int n, int m, const int *__restrict__ own_col, const double *__restrict__ own_val,
                              const double *__restrict__ diag, const long long *__restrict__ seg,
                              int *__restrict__ cursor, unsigned long long *__restrict__ key,
                              double *__restrict__ val)
{
    const long total = (long)n * m;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const int i = (int)(e / m), q = (int)(e % m);
        const int c = own_col[e];
        const double v = own_val[e];
        const long long a = seg[i] + 1 + q;
        key[a] = seg_key(c, i);
        val[a] = v;
        const long long b = seg[c] + 1 + m + atomicAdd(&cursor[c], 1);
        key[b] = seg_key(i, i);
        val[b] = v;
        if (q == 0) {
            key[seg[i]] = seg_key(i, i);
            val[seg[i]] = diag[i];
        }
    }
}

// ---- per-row sort + deduplication (one 256-thread workgroup per row)

template <int P>
__global__ __launch_bounds__(256) void k_orb_sort(int n, const long long *__restrict__ seg,
                                                  unsigned long long *__restrict__ key, double *__restrict__ val,
                                                  long long *__restrict__ kept)
{
    __shared__ unsigned long long sk[P];
    __shared__ double sv[P];
    __shared__ long long wsum[4];
    constexpr int PER = P / 256;
    for (long row = blockIdx.x; row < n; row += gridDim.x) {
        const long long base = seg[row];
        const int len = (int)(seg[row + 1] - base);
        for (int e = threadIdx.x; e < P; e += 256) {
            sk[e] = e < len ? key[base + e] : ~0ULL;
            sv[e] = e < len ? val[base + e] : 0.0;
        }
        __syncthreads();
        // bitonic sort, ascending
        for (int size = 2; size <= P; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int t = threadIdx.x; t < P / 2; t += 256) {
                    const int lo = 2 * t - (t & (stride - 1));
                    const int hi = lo + stride;
                    const bool up = (lo & size) == 0;
                    const unsigned long long a = sk[lo], b = sk[hi];
                    if ((a > b) == up) {
                        sk[lo] = b;
                        sk[hi] = a;
                        const double x = sv[lo];
                        sv[lo] = sv[hi];
                        sv[hi] = x;
                    }
                }
                __syncthreads();
            }
        }
        // keep the last entry of each column run (largest source row)
        int flag[PER];
        long long cnt = 0;
        for (int j = 0; j < PER; ++j) {
            const int e = threadIdx.x * PER + j;
            flag[j] = e < len && (e + 1 == len || (sk[e + 1] >> 32) != (sk[e] >> 32));
            cnt += flag[j];
        }
        const long long incl = block_inclusive_scan(cnt, wsum);
        long long pos = incl - cnt;
        unsigned long long ok[PER];
        double ov[PER];
        for (int j = 0; j < PER; ++j) {
            ok[j] = sk[threadIdx.x * PER + j];
            ov[j] = sv[threadIdx.x * PER + j];
        }
        __syncthreads();
        for (int j = 0; j < PER; ++j) {
            if (flag[j]) {
                key[base + pos] = ok[j] >> 32;
                val[base + pos] = ov[j];
                ++pos;
            }
        }
        if (threadIdx.x == 255) {
            kept[row] = incl;
        }
        __syncthreads();
    }
}

__global__ void k_orb_compact(int n, const long long *__restrict__ seg, const unsigned long long *__restrict__ key,
                              const double *__restrict__ val, const long long *__restrict__ rowp,
                              int *__restrict__ row_ptr, int *__restrict__ col_ind, double *__restrict__ values)
{
    const int lane = threadIdx.x & 63;
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long waves = ((long)gridDim.x * blockDim.x) >> 6;
    for (long row = wave; row <= n; row += waves) {
        if (lane == 0) {
            row_ptr[row] = (int)rowp[row];
        }
        if (row == n) {
            continue;
        }
        const long long src = seg[row], dst = rowp[row];
        const int cnt = (int)(rowp[row + 1] - dst);
        for (int e = lane; e < cnt; e += 64) {
            col_ind[dst + e] = (int)key[src + e];
            values[dst + e] = val[src + e];
        }
    }
}

template <int P>
void launch_sort(Context *ctx, int n, const long long *seg, unsigned long long *key, double *val, long long *kept)
{
    hipLaunchKernelGGL(k_orb_sort<P>, dim3((unsigned)std::min(n, 65536)), dim3(256), 0, ctx->stream, n, seg, key,
                       val, kept);
}

int sync_copy(Context *ctx, void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    IPT_HIPCHK(hipMemcpyAsync(dst, src, bytes, kind, ctx->stream));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return IPT_CUDA_SUCCESS;
}

}  // namespace

// Host part: the n sorted diagonal values.
void orbital_diagonal(int n, unsigned long long seed, int norb, int nsub, double elo, double ehi,
                      std::vector<double> *out)
{
    std::vector<double> e((size_t)norb);
    for (int o = 0; o < norb; ++o) {
        e[(size_t)o] = elo + (ehi - elo) * unif(seed ^ kTagOrbital, n, o, 0);
    }
    std::sort(e.begin(), e.end());
    out->assign((size_t)n, 0.0);
    std::vector<int> perm((size_t)norb);
    for (int i = 0; i < n; ++i) {
        std::iota(perm.begin(), perm.end(), 0);
        double s = 0.0;
        for (int t = 0; t < nsub; ++t) {
            int r = t + (int)(unif(seed ^ kTagSubset, n, i, t) * (double)(norb - t));
            r = std::min(r, norb - 1);
            std::swap(perm[(size_t)t], perm[(size_t)r]);
            s += e[(size_t)perm[(size_t)t]];
        }
        (*out)[(size_t)i] = s;
    }
    std::sort(out->begin(), out->end());
}

}  // namespace ipt

using namespace ipt;

extern "C" int ipt_hip_generate_orbital_csr(int n, int m, unsigned long long seed, int norb, int nsub, double elo,
                                            double ehi, double sigma, int *d_row_ptr, int *d_col_ind,
                                            double *d_values, long long capacity, long long *nnz_out,
                                            void *handle)
{
    if (nnz_out) *nnz_out = 0;
    if (n < 2 || m < 0 || m > kMaxDraws || m > n - 1 || norb < 1 || nsub < 1 || nsub > norb ||
        d_row_ptr == nullptr || d_col_ind == nullptr || d_values == nullptr || nnz_out == nullptr ||
        !(ehi >= elo) || !(sigma >= 0.0)) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    std::vector<double> hd;
    orbital_diagonal(n, seed, norb, nsub, elo, ehi, &hd);
    const long own = (long)n * m;
    DevBuf<int> own_col, incoming, cursor, maxlen;
    DevBuf<double> own_val, diag;
    IPT_STCHK(own_col.alloc((size_t)std::max(own, 1L)));
    IPT_STCHK(own_val.alloc((size_t)std::max(own, 1L)));
    IPT_STCHK(incoming.alloc((size_t)n));
    IPT_STCHK(cursor.alloc((size_t)n));
    IPT_STCHK(maxlen.alloc(1));
    IPT_STCHK(diag.alloc((size_t)n));
    IPT_HIPCHK(hipMemsetAsync(incoming.p, 0, (size_t)n * sizeof(int), ctx->stream));
    IPT_HIPCHK(hipMemsetAsync(cursor.p, 0, (size_t)n * sizeof(int), ctx->stream));
    IPT_HIPCHK(hipMemsetAsync(maxlen.p, 0, sizeof(int), ctx->stream));
    IPT_HIPCHK(hipMemcpyAsync(diag.p, hd.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    const unsigned draw_blocks = (unsigned)std::min<long>(((long)n + 3) / 4, 65536);
    if (m > 0) {
        hipLaunchKernelGGL(k_orb_draw, dim3(draw_blocks), dim3(256), 0, ctx->stream, n, m, seed, sigma, own_col.p,
                           own_val.p, incoming.p);
        IPT_HIPCHK(hipGetLastError());
    }
    DevBuf<long long> len, seg;
    IPT_STCHK(len.alloc((size_t)n));
    IPT_STCHK(seg.alloc((size_t)n + 1));
    hipLaunchKernelGGL(k_orb_seglen, dim3((unsigned)std::min<long>(((long)n + 255) / 256, 4096)), dim3(256), 0,
                       ctx->stream, n, m, incoming.p, len.p, maxlen.p);
    IPT_HIPCHK(hipGetLastError());
    IPT_STCHK(exclusive_scan(ctx, len.p, n, seg.p));
    long long total = 0;
    int hmax = 0;
    IPT_STCHK(sync_copy(ctx, &total, seg.p + n, sizeof(long long), hipMemcpyDeviceToHost));
    IPT_STCHK(sync_copy(ctx, &hmax, maxlen.p, sizeof(int), hipMemcpyDeviceToHost));
    if (hmax > kMaxSegment) {
        fprintf(stderr, "IPT orbital generator: a row collects %d entries (limit %d)\n", hmax, kMaxSegment);
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    DevBuf<unsigned long long> key;
    DevBuf<double> val;
    IPT_STCHK(key.alloc((size_t)total));
    IPT_STCHK(val.alloc((size_t)total));
    hipLaunchKernelGGL(k_orb_scatter, dim3((unsigned)std::min<long>((std::max(own, (long)n) + 255) / 256, 65536)),
                       dim3(256), 0, ctx->stream, n, m, own_col.p, own_val.p, diag.p, seg.p, cursor.p, key.p, val.p);
    IPT_HIPCHK(hipGetLastError());
    if (m == 0) {
        // diagonal only: the scatter loop above has no entries to visit
        std::vector<unsigned long long> hk((size_t)n);
        for (int i = 0; i < n; ++i) {
            hk[(size_t)i] = seg_key(i, i);
        }
        IPT_HIPCHK(hipMemcpyAsync(key.p, hk.data(), (size_t)n * sizeof(unsigned long long), hipMemcpyHostToDevice,
                                  ctx->stream));
        IPT_HIPCHK(hipMemcpyAsync(val.p, diag.p, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    own_col.release();
    own_val.release();
    DevBuf<long long> kept, rowp;
    IPT_STCHK(kept.alloc((size_t)n));
    IPT_STCHK(rowp.alloc((size_t)n + 1));
    if (hmax <= 256) {
        launch_sort<256>(ctx, n, seg.p, key.p, val.p, kept.p);
    } else if (hmax <= 512) {
        launch_sort<512>(ctx, n, seg.p, key.p, val.p, kept.p);
    } else if (hmax <= 1024) {
        launch_sort<1024>(ctx, n, seg.p, key.p, val.p, kept.p);
    } else {
        launch_sort<2048>(ctx, n, seg.p, key.p, val.p, kept.p);
    }
    IPT_HIPCHK(hipGetLastError());
    IPT_STCHK(exclusive_scan(ctx, kept.p, n, rowp.p));
    long long nnz = 0;
    IPT_STCHK(sync_copy(ctx, &nnz, rowp.p + n, sizeof(long long), hipMemcpyDeviceToHost));
    if (nnz > INT_MAX || nnz > capacity) {
        fprintf(stderr, "IPT orbital generator: nnz %lld exceeds %s\n", nnz,
                nnz > INT_MAX ? "32-bit indexing" : "the output capacity");
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    hipLaunchKernelGGL(k_orb_compact, dim3((unsigned)std::min<long>(((long)n + 4) / 4, 65536)), dim3(256), 0,
                       ctx->stream, n, seg.p, key.p, val.p, rowp.p, d_row_ptr, d_col_ind, d_values);
    IPT_HIPCHK(hipGetLastError());
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    *nnz_out = nnz;
    return IPT_CUDA_SUCCESS;
}
