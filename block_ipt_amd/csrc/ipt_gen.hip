// Synthetic dense workload generator (device side).  The recipe is stated
// in include/ipt_hip_ext.h; the splitmix64 counter stream follows
// IPT/test/mixed_precision_tf32/benchmark_tf32_cusolver_sweep.cu:434-449.
#include <algorithm>

#include "ipt_internal.h"
#include "../../include/ipt_hip_ext.h"

namespace ipt {

__host__ __device__ inline unsigned long long splitmix64(unsigned long long x)
{
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

__host__ __device__ inline double uniform01(unsigned long long seed, long long n, long long row, long long col)
{
    unsigned long long x = seed + (unsigned long long)n * 0xd1b54a32d192ed03ULL;
    x ^= (unsigned long long)(row + 1) * 0x9e3779b97f4a7c15ULL;
    x ^= (unsigned long long)(col + 1) * 0xbf58476d1ce4e5b9ULL;
    return (double)(splitmix64(x) >> 11) * (1.0 / 9007199254740992.0);
}

__global__ void k_gen_dense(double *__restrict__ H, int n, unsigned long long seed, double s,
                            const double *__restrict__ d)
{
    const long total = (long)n * n;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const long r = t % n, c = t / n;
        double v;
        if (r == c) {
            v = d[r];
        } else {
            const long a = r < c ? r : c, b = r < c ? c : r;
            const double u1 = uniform01(seed, n, a, b);
            const double u2 = uniform01(seed, n, b, a);
            v = s * sqrt(-2.0 * log1p(-u1)) * cos(6.283185307179586 * u2);
        }
        H[t] = v;
    }
}

// diag(i + 1) + eps U: H_ii = i + 1 + eps u(i, i), H_ij = H_ji =
// eps (u(i, j) + u(j, i)) / 2 (IPT/test/ACX/benchmark_acx_cuda_sweep.cu:284-312,
// the same recipe as benchmark_tf32_cusolver_sweep.cu:472-503).
__global__ void k_gen_eps(double *__restrict__ H, int n, unsigned long long seed, double eps)
{
    const long total = (long)n * n;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const long r = t % n, c = t / n;
        double v;
        if (r == c) {
            // separately rounded product and sum, as the host generators
            // compute it (no contraction into an FMA)
            v = __dadd_rn((double)(r + 1), __dmul_rn(eps, uniform01(seed, n, r, c)));
        } else {
            // the reference writes the pair from (row <= col): the same value
            // both ways
            const long a = r < c ? r : c, b = r < c ? c : r;
            v = 0.5 * eps * (uniform01(seed, n, a, b) + uniform01(seed, n, b, a));
        }
        H[t] = v;
    }
}

}  // namespace ipt

using namespace ipt;

extern "C" int ipt_hip_generate_eps(double *d_H, int n, double eps, unsigned long long seed, void *handle)
{
    if (d_H == nullptr || n <= 0) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    const long total = (long)n * n;
    const long blocks = std::min((total + 255) / 256, 65536L);
    hipLaunchKernelGGL(k_gen_eps, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, d_H, n, seed, eps);
    IPT_HIPCHK(hipGetLastError());
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return IPT_CUDA_SUCCESS;
}

extern "C" int ipt_hip_generate_dense(double *d_H, int n, unsigned long long seed, double lo, double hi,
                                      double vscale, void *handle)
{
    if (d_H == nullptr || n <= 0) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    std::vector<double> d((size_t)n);
    for (int i = 0; i < n; ++i) {
        d[(size_t)i] = lo + (hi - lo) * uniform01(seed, n, i, n);
    }
    std::sort(d.begin(), d.end());
    DevBuf<double> dd;
    IPT_STCHK(dd.alloc((size_t)n));
    IPT_HIPCHK(hipMemcpyAsync(dd.p, d.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    const long total = (long)n * n;
    long blocks = (total + 255) / 256;
    if (blocks > 65536) {
        blocks = 65536;
    }
    hipLaunchKernelGGL(k_gen_dense, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, d_H, n, seed,
                       vscale / std::sqrt((double)n), dd.p);
    IPT_HIPCHK(hipGetLastError());
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return IPT_CUDA_SUCCESS;
}

// Near-degenerate variant of a generated matrix: for each position i the
// diagonal entry i+1 becomes H(i,i) + gap.  The diagonal must stay sorted:
// a position whose next diagonal entry would not stay above the new value
// (H(i+2,i+2) <= H(i,i) + gap), or out of range, is rejected before any
// entry is written.
extern "C" int ipt_hip_set_diagonal_gaps(double *d_H, int n, const int *positions, int count, double gap,
                                         void *handle)
{
    if (d_H == nullptr || n <= 1 || count < 0 || (count > 0 && positions == nullptr) || !(gap > 0.0)) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    auto at = [&](int i) { return d_H + (size_t)i * n + i; };
    std::vector<double> nv((size_t)count);
    for (int q = 0; q < count; ++q) {
        const int i = positions[q];
        if (i < 0 || i + 1 >= n) {
            fprintf(stderr, "IPT: diagonal gap position %d out of range for n=%d\n", i, n);
            return IPT_CUDA_INVALID_ARGUMENT;
        }
        double di = 0.0, dnext = INFINITY;
        IPT_HIPCHK(hipMemcpy(&di, at(i), sizeof(double), hipMemcpyDeviceToHost));
        if (i + 2 < n) {
            IPT_HIPCHK(hipMemcpy(&dnext, at(i + 2), sizeof(double), hipMemcpyDeviceToHost));
        }
        nv[(size_t)q] = di + gap;
        if (!(nv[(size_t)q] < dnext)) {
            fprintf(stderr, "IPT: a gap of %g after diagonal entry %d would unsort the diagonal\n", gap, i);
            return IPT_CUDA_INVALID_ARGUMENT;
        }
    }
    for (int q = 0; q < count; ++q) {
        IPT_HIPCHK(hipMemcpy(at(positions[q] + 1), &nv[(size_t)q], sizeof(double), hipMemcpyHostToDevice));
    }
    return IPT_CUDA_SUCCESS;
}
