// ACX-accelerated plain IPT (IPT/src/ACX/acx_cuda.cu): alternating order-2 /
// order-3 extrapolation of the plain IPT fixed-point map, FP64 throughout.
//
// One application of the map F is the FP64 MFMA product Y = H X followed by
// one fused elementwise kernel (k_acx_map) that forms F(X), the per-column
// residual ||Y_j - X_j Y_jj|| and, depending on where the call sits in the
// iteration, the sums that the stop test or the extrapolation needs: the
// reference's separate metadata / finalize / delta / nrm2 / sigma passes
// (acx_cuda.cu:106-157, 178-296, 298-358, 496-520) become one pass over the
// n x k block per map.  Per-(column, row chunk) partial sums are reduced in
// a fixed order, so a run is bitwise reproducible.  The values diag(H f1)
// come from a tiled diagonal-of-product kernel that reads k rows of H once
// instead of a full product (acx_cuda.cu:360-388).
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cstring>
#include <functional>

#include "ipt_acx.h"
#include "../../include/ipt_hip_acx.h"

namespace ipt {
namespace {

constexpr int kThreads = 256;
constexpr int kRowsPerChunk = 1024;

// What k_acx_map accumulates besides the residual:
//   MAP_PLAIN        nothing
//   MAP_FIXED_POINT  sum (F(X) - X)^2 and sum X^2 (stop test, X = x)
//   MAP_ORDER2       sum d2 d1 and sum d2^2 with X = f1, F(X) = f2
//   MAP_ORDER3       sum d3 d2 and sum d3^2 with X = f2, F(X) = f3
enum { MAP_PLAIN = 0, MAP_FIXED_POINT = 1, MAP_ORDER2 = 2, MAP_ORDER3 = 3 };

__device__ inline double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) {
        v += __shfl_xor(v, o, 64);
    }
    return v;
}

// Sum of three values over a 256-thread block, valid in thread 0.
__device__ inline void block_sum3(double &a, double &b, double &c)
{
    __shared__ double s[3][kThreads / 64];
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        s[0][w] = a;
        s[1][w] = b;
        s[2][w] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = s[0][0] + s[0][1] + s[0][2] + s[0][3];
        b = s[1][0] + s[1][1] + s[1][2] + s[1][3];
        c = s[2][0] + s[2][1] + s[2][2] + s[2][3];
    }
}

// F(X) of IPT/src/ACX/acx_cuda.cu:106-157 for one (column c, row chunk):
//   mx = Y_cc, u = mx - d_c X_cc,
//   F_rc = (Y_rc - d_r X_rc - X_rc u) / (d_c - d_r),  F_cc = 1,
// a zero denominator (|d_c - d_r| <= DBL_EPSILON) sets *flag and writes 0.
// part[3 b + 0..2] (b = c * nchunks + chunk): residual sum, then the two
// MODE sums.  Y must not alias F (every chunk of column c reads Y_cc).
template <int MODE>
__global__ __launch_bounds__(kThreads) void k_acx_map(const double *__restrict__ Y, const double *__restrict__ X,
                                                      const double *__restrict__ d, int n, int nchunks,
                                                      double *__restrict__ F, const double *__restrict__ x0,
                                                      const double *__restrict__ f1, double *__restrict__ part,
                                                      int *__restrict__ flag)
{
    const long b = blockIdx.x;
    const int c = (int)(b / nchunks);
    const int chunk = (int)(b - (long)c * nchunks);
    const long cn = (long)c * n;
    const double mx = Y[cn + c];
    const double dc = d[c];
    const double upd = mx - dc * X[cn + c];
    const int r0 = chunk * kRowsPerChunk;
    const int r1 = min(n, r0 + kRowsPerChunk);
    double s_res = 0.0, s_a = 0.0, s_b = 0.0;
    int bad = 0;
    for (int r = r0 + threadIdx.x; r < r1; r += kThreads) {
        const long idx = cn + r;
        const double y = Y[idx];
        const double x = X[idx];
        const double res = y - x * mx;
        s_res += res * res;
        double f;
        if (r == c) {
            f = 1.0;
        } else {
            const double dr = d[r];
            const double den = dc - dr;
            if (fabs(den) <= DBL_EPSILON) {
                bad = 1;
                f = 0.0;
            } else {
                f = (y - dr * x - x * upd) / den;
            }
        }
        F[idx] = f;
        if (MODE == MAP_FIXED_POINT) {
            const double dl = f - x;
            s_a += dl * dl;
            s_b += x * x;
        } else if (MODE == MAP_ORDER2) {
            // X = f1, f = f2
            const double xv = x0[idx];
            const double d1 = x - xv;
            const double d2 = f - 2.0 * x + xv;
            s_a += d2 * d1;
            s_b += d2 * d2;
        } else if (MODE == MAP_ORDER3) {
            // X = f2, f = f3
            const double xv = x0[idx];
            const double f1v = f1[idx];
            const double d2 = x - 2.0 * f1v + xv;
            const double d3 = f - 3.0 * x + 3.0 * f1v - xv;
            s_a += d3 * d2;
            s_b += d3 * d3;
        }
    }
    if (bad) {
        atomicOr(flag, 1);
    }
    block_sum3(s_res, s_a, s_b);
    if (threadIdx.x == 0) {
        part[3 * b + 0] = s_res;
        part[3 * b + 1] = s_a;
        part[3 * b + 2] = s_b;
    }
}

// out[0] = max_c sqrt(sum of column c's residual partials),
// out[1], out[2] = totals of the two MODE sums.  One block of 1024 threads.
__global__ __launch_bounds__(1024) void k_acx_reduce(const double *__restrict__ part, int k, int nchunks,
                                                     double *__restrict__ out)
{
    __shared__ double sm[16], sa[16], sb[16];
    double m = 0.0, a = 0.0, bsum = 0.0;
    for (int c = threadIdx.x; c < k; c += blockDim.x) {
        double r = 0.0;
        for (int ch = 0; ch < nchunks; ++ch) {
            const long b = (long)c * nchunks + ch;
            r += part[3 * b];
            a += part[3 * b + 1];
            bsum += part[3 * b + 2];
        }
        m = fmax(m, sqrt(r));
    }
    for (int o = 32; o > 0; o >>= 1) {
        m = fmax(m, __shfl_xor(m, o, 64));
        a += __shfl_xor(a, o, 64);
        bsum += __shfl_xor(bsum, o, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sm[w] = m;
        sa[w] = a;
        sb[w] = bsum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        for (int i = 1; i < nw; ++i) {
            m = fmax(m, sm[i]);
            a += sa[i];
            bsum += sb[i];
        }
        out[0] = m;
        out[1] = a;
        out[2] = bsum;
    }
}

// x += 2 s d1 + s^2 d2 (ORDER 2) or 3 s d1 + 3 s^2 d2 + s^3 d3 (ORDER 3),
// s = |num / den| from column c's partials (0 unless den > 0),
// IPT/src/ACX/acx_cuda.cu:190-296.
template <int ORDER>
__global__ __launch_bounds__(kThreads) void k_acx_extrapolate(double *__restrict__ x, const double *__restrict__ f1,
                                                              const double *__restrict__ f2,
                                                              const double *__restrict__ f3, int n, int nchunks,
                                                              const double *__restrict__ part)
{
    const long b = blockIdx.x;
    const int c = (int)(b / nchunks);
    const int chunk = (int)(b - (long)c * nchunks);
    double num = 0.0, den = 0.0;
    for (int ch = 0; ch < nchunks; ++ch) {
        const long q = (long)c * nchunks + ch;
        num += part[3 * q + 1];
        den += part[3 * q + 2];
    }
    const double s = den > 0.0 ? fabs(num / den) : 0.0;
    const double s2 = s * s;
    const long cn = (long)c * n;
    const int r1 = min(n, (chunk + 1) * kRowsPerChunk);
    for (int r = chunk * kRowsPerChunk + threadIdx.x; r < r1; r += kThreads) {
        const long idx = cn + r;
        const double xv = x[idx];
        const double a1 = f1[idx];
        const double a2 = f2[idx];
        const double d1 = a1 - xv;
        const double d2 = a2 - 2.0 * a1 + xv;
        if (ORDER == 2) {
            x[idx] = xv + (2.0 * s * d1 + s2 * d2);
        } else {
            const double d3 = f3[idx] - 3.0 * a2 + 3.0 * a1 - xv;
            x[idx] = xv + (3.0 * s * d1 + 3.0 * s2 * d2 + s2 * s * d3);
        }
    }
}

// part[split * k + c] = sum over j in the split's row range of H(c, j) V(j, c).
// Block: 64 consecutive targets c, V staged through LDS 64 rows at a time so
// that both operands are read along their contiguous dimension.
__global__ __launch_bounds__(kThreads) void k_acx_diag_product(const double *__restrict__ H, long ldh,
                                                               const double *__restrict__ V, int n, int k,
                                                               int split_len, double *__restrict__ part)
{
    __shared__ double T[64][65];
    __shared__ double red[4][64];
    const int c0 = blockIdx.x * 64;
    const int js = blockIdx.y * split_len;
    const int je = min(n, js + split_len);
    const int cl = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int c = c0 + cl;
    double acc = 0.0;
    for (int j0 = js; j0 < je; j0 += 64) {
        for (int e = threadIdx.x; e < 64 * 64; e += kThreads) {
            const int jl = e & 63, tc = e >> 6;
            const int j = j0 + jl, cc = c0 + tc;
            T[tc][jl] = (j < je && cc < k) ? V[(long)cc * n + j] : 0.0;
        }
        __syncthreads();
        if (c < k) {
            for (int jl = q; jl < 64; jl += 4) {
                const int j = j0 + jl;
                if (j < je) {
                    acc += H[(long)j * ldh + c] * T[cl][jl];
                }
            }
        }
        __syncthreads();
    }
    red[q][cl] = acc;
    __syncthreads();
    if (q == 0 && c < k) {
        part[(long)blockIdx.y * k + c] = ((red[0][cl] + red[1][cl]) + red[2][cl]) + red[3][cl];
    }
}

__global__ void k_acx_diag_reduce(const double *__restrict__ part, int splits, int k, double *__restrict__ values)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < k; c += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int i = 0; i < splits; ++i) {
            s += part[(long)i * k + c];
        }
        values[c] = s;
    }
}

struct AcxWork {
    const std::function<int(const double *, double *)> *product;  // Y = H X, n x k
    int n, k, nchunks;
    const double *d;
    double *Y;
    double *part;
    int *flag;
};

// F(X) -> Fout with the MODE sums (see k_acx_map); one product.
int acx_map(Context *ctx, const AcxWork &w, int mode, const double *X, double *Fout, const double *x0,
            const double *f1)
{
    IPT_STCHK((*w.product)(X, w.Y));
    const dim3 grid((unsigned)((long)w.k * w.nchunks));
    switch (mode) {
    case MAP_FIXED_POINT:
        hipLaunchKernelGGL(k_acx_map<MAP_FIXED_POINT>, grid, dim3(kThreads), 0, ctx->stream, w.Y, X, w.d, w.n,
                           w.nchunks, Fout, x0, f1, w.part, w.flag);
        break;
    case MAP_ORDER2:
        hipLaunchKernelGGL(k_acx_map<MAP_ORDER2>, grid, dim3(kThreads), 0, ctx->stream, w.Y, X, w.d, w.n, w.nchunks,
                           Fout, x0, f1, w.part, w.flag);
        break;
    case MAP_ORDER3:
        hipLaunchKernelGGL(k_acx_map<MAP_ORDER3>, grid, dim3(kThreads), 0, ctx->stream, w.Y, X, w.d, w.n, w.nchunks,
                           Fout, x0, f1, w.part, w.flag);
        break;
    default:
        hipLaunchKernelGGL(k_acx_map<MAP_PLAIN>, grid, dim3(kThreads), 0, ctx->stream, w.Y, X, w.d, w.n, w.nchunks,
                           Fout, x0, f1, w.part, w.flag);
        break;
    }
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

}  // namespace

// values[c] = (H V)(c, c) for c < k.
int acx_diag_product(Context *ctx, const double *H, int n, const double *V, int k, double *values)
{
    const int ctiles = (k + 63) / 64;
    int splits = 1024 / ctiles;
    splits = std::max(1, std::min(splits, (n + 63) / 64));
    int split_len = (n + splits - 1) / splits;
    split_len = (split_len + 63) / 64 * 64;
    splits = (n + split_len - 1) / split_len;
    double *part = nullptr;
    IPT_STCHK(context_scratch(ctx, SLOT_REDUCE, (size_t)splits * k * sizeof(double), (void **)&part));
    hipLaunchKernelGGL(k_acx_diag_product, dim3(ctiles, splits), dim3(kThreads), 0, ctx->stream, H, (long)n, V, n, k,
                       split_len, part);
    IPT_HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_acx_diag_reduce, dim3((k + 255) / 256), dim3(256), 0, ctx->stream, part, splits, k, values);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

static double now_seconds()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int acx_loop(Context *ctx, const double *d_diag, int n, int k, AcxRun *run)
{
    const size_t nk = (size_t)n * (size_t)k;
    const int nchunks = (n + kRowsPerChunk - 1) / kRowsPerChunk;
    if ((long)k * nchunks > 0x7fffffffL) {
        return ACX_CUDA_INVALID_ARGUMENT;
    }
    DevBuf<double> x, f1, f2, f3, Y, part, sums;
    DevBuf<int> flag;
    IPT_STCHK(x.alloc(nk));
    IPT_STCHK(f1.alloc(nk));
    IPT_STCHK(f2.alloc(nk));
    IPT_STCHK(f3.alloc(nk));
    IPT_STCHK(Y.alloc(nk));
    IPT_STCHK(part.alloc(3 * (size_t)k * nchunks));
    IPT_STCHK(sums.alloc(3));
    IPT_STCHK(flag.alloc(1));
    IPT_HIPCHK(hipMemsetAsync(flag.p, 0, sizeof(int), ctx->stream));
    if (run->x0 != nullptr) {
        IPT_HIPCHK(hipMemcpyAsync(x.p, run->x0, nk * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    } else {
        IPT_STCHK(set_identity(ctx, x.p, n, n, k));
    }
    const AcxWork w{run->product, n, k, nchunks, d_diag, Y.p, part.p, flag.p};
    double *hs = ctx->host_scalars;
    int *hf = ctx->host_flags;
    bool final_f1 = true;
    for (int iter = 1; iter <= run->maxiter; ++iter) {
        const int order = (iter % 2 == 1) ? 2 : 3;
        run->iterations = iter;
        IPT_STCHK(acx_map(ctx, w, MAP_FIXED_POINT, x.p, f1.p, nullptr, nullptr));
        run->f_calls += 1;
        hipLaunchKernelGGL(k_acx_reduce, dim3(1), dim3(1024), 0, ctx->stream, part.p, k, nchunks, sums.p);
        IPT_HIPCHK(hipGetLastError());
        IPT_HIPCHK(hipMemcpyAsync(hs, sums.p, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        IPT_HIPCHK(hipMemcpyAsync(hf, flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
        if (hf[0] != 0) {
            return ACX_CUDA_ZERO_DENOMINATOR;
        }
        run->max_residual = hs[0];
        const double delta_norm = std::sqrt(hs[1]);
        const double current_norm = std::sqrt(hs[2]);
        if (!(current_norm > 0.0) || !std::isfinite(current_norm)) {
            return ACX_CUDA_INVALID_ARGUMENT;
        }
        const double fp = delta_norm / current_norm;
        run->fixed_point_residual = fp;
        if (run->stop_inclusive ? fp <= run->stop_tol : fp < run->stop_tol) {
            run->converged = 1;
            final_f1 = true;
            break;
        }
        if (order == 2) {
            IPT_STCHK(acx_map(ctx, w, MAP_ORDER2, f1.p, f2.p, x.p, nullptr));
            run->f_calls += 1;
            hipLaunchKernelGGL(k_acx_extrapolate<2>, dim3((unsigned)((long)k * nchunks)), dim3(kThreads), 0,
                               ctx->stream, x.p, f1.p, f2.p, nullptr, n, nchunks, part.p);
        } else {
            IPT_STCHK(acx_map(ctx, w, MAP_PLAIN, f1.p, f2.p, nullptr, nullptr));
            IPT_STCHK(acx_map(ctx, w, MAP_ORDER3, f2.p, f3.p, x.p, f1.p));
            run->f_calls += 2;
            hipLaunchKernelGGL(k_acx_extrapolate<3>, dim3((unsigned)((long)k * nchunks)), dim3(kThreads), 0,
                               ctx->stream, x.p, f1.p, f2.p, f3.p, n, nchunks, part.p);
        }
        IPT_HIPCHK(hipGetLastError());
        final_f1 = !run->final_x_unconverged;
    }
    // maps after the last stop test are checked here
    IPT_HIPCHK(hipMemcpyAsync(hf, flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    if (hf[0] != 0) {
        return ACX_CUDA_ZERO_DENOMINATOR;
    }
    run->vectors.release();
    if (final_f1) {
        run->vectors.p = f1.take();
    } else {
        run->vectors.p = x.take();
    }
    run->vectors.count = nk;
    return ACX_CUDA_SUCCESS;
}

namespace {

int acx_run(Context *ctx, const double *dH, int n, int k, double tol, int maxiter, ACXCudaResult *result)
{
    const double t_start = now_seconds();
    DevBuf<double> d, vals;
    IPT_STCHK(d.alloc((size_t)n));
    IPT_STCHK(vals.alloc((size_t)k));
    IPT_STCHK(dense_diagonal(ctx, dH, n, n, d.p));
    const std::function<int(const double *, double *)> product = [&](const double *X, double *Y) {
        return gemm_f64(ctx, dH, n, X, n, Y, n, n, k);
    };
    AcxRun run;
    run.product = &product;
    run.stop_tol = tol < ACX_FIXED_POINT_RTOL ? tol : ACX_FIXED_POINT_RTOL;
    run.maxiter = maxiter;
    const double t_iter = now_seconds();
    const int st = acx_loop(ctx, d.p, n, k, &run);
    result->iterations = run.iterations;
    result->f_calls = run.f_calls;
    result->converged = run.converged;
    result->max_residual = run.max_residual;
    result->fixed_point_residual = run.fixed_point_residual;
    if (st != ACX_CUDA_SUCCESS) {
        return st;
    }
    result->iteration_time_sec = now_seconds() - t_iter;
    const double t_fin = now_seconds();
    IPT_STCHK(acx_diag_product(ctx, dH, n, run.vectors.p, k, vals.p));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    result->finalization_time_sec = now_seconds() - t_fin;
    result->n = n;
    result->k = k;
    result->d_vectors = run.vectors.take();
    result->d_values = vals.take();
    result->time_sec = now_seconds() - t_start;
    return ACX_CUDA_SUCCESS;
}

}  // namespace
}  // namespace ipt

using namespace ipt;

extern "C" const char *acx_cuda_status_string(int status)
{
    switch (status) {
    case ACX_CUDA_SUCCESS: return "success";
    case ACX_CUDA_INVALID_ARGUMENT: return "invalid argument";
    case ACX_CUDA_ALLOCATION_FAILED: return "allocation failed";
    case ACX_CUDA_CUDA_ERROR: return "HIP runtime error";
    case ACX_CUDA_CUBLAS_ERROR: return "device BLAS error";
    case ACX_CUDA_ZERO_DENOMINATOR: return "zero diagonal denominator";
    default: return "unknown error";
    }
}

extern "C" void acx_cuda_reset_result(ACXCudaResult *result)
{
    if (result == nullptr) {
        return;
    }
    memset(result, 0, sizeof(*result));
    result->max_residual = NAN;
    result->fixed_point_residual = NAN;
}

extern "C" void acx_cuda_free_result(ACXCudaResult *result)
{
    if (result == nullptr) {
        return;
    }
    if (result->d_vectors != nullptr) {
        (void)hipFree(result->d_vectors);
    }
    if (result->d_values != nullptr) {
        (void)hipFree(result->d_values);
    }
    acx_cuda_reset_result(result);
}

extern "C" int acx_cuda_device_tol(const double *d_matrix_col_major, int n, int k, double tol, int maxiter,
                                   void *handle, ACXCudaResult *result)
{
    if (result == nullptr) {
        return ACX_CUDA_INVALID_ARGUMENT;
    }
    acx_cuda_reset_result(result);
    if (d_matrix_col_major == nullptr || n <= 0 || k <= 0 || k > n || tol <= 0.0 || maxiter <= 0) {
        return ACX_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    st = acx_run(ctx, d_matrix_col_major, n, k, tol, maxiter, result);
    if (st != ACX_CUDA_SUCCESS) {
        acx_cuda_free_result(result);
    }
    return st;
}
