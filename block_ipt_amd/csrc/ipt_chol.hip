// Cholesky factorisation with column rejection and the inverse of the
// factor, on the device, for the small Gram matrices of Cholesky QR and of
// Davidson's block orthonormalisation (orders up to a few hundred).
//
// One workgroup runs the right-looking factorisation of G (t x t) column by
// column.  At step j the pivot d = G_jj (after the updates of the columns
// kept before j) decides whether column j is kept; a kept column becomes
// L(:, j) = G(j:, j) / sqrt(d) and updates the trailing block, a rejected
// one is left out (a zero column), so that the kept rows and columns carry
// the Cholesky factor of the kept principal submatrix.  The same row
// eliminations applied to an identity matrix produce E = L^-1 on the kept
// rows and columns (rejected rows of E are unit rows that no kept row uses),
// and the result is handed out transposed: W = E^T, so that V W is the
// orthonormalised V.  Orders up to kLdsMax work in LDS, larger ones in
// global memory (the caller's G, overwritten, and E in the output).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <vector>

#include "ipt_block.h"
#include "ipt_internal.h"

namespace ipt {

namespace {

constexpr int CT = 1024;     // threads of the factorisation workgroup
constexpr int kLdsMax = 96;  // 2 t^2 doubles + 2 t in LDS: 145.5 KiB at t = 96

// G (t x t, column-major, both triangles) -> keep[j], W = (L^-1)^T (t x t),
// G overwritten.  Column j is rejected when d <= 0 or not finite, or, with
// ref != NULL, when ref[j] is not a finite positive norm or
// sqrt(d) <= rel sqrt(ref[j]).  *nkept = number of kept columns.
template <bool kLds>
__global__ __launch_bounds__(CT) void k_chol_reject(double *__restrict__ G, int t, const double *__restrict__ ref,
                                                    double rel, int *__restrict__ keep, double *__restrict__ W,
                                                    int *__restrict__ nkept)
{
    extern __shared__ double sh[];
    __shared__ double piv;
    __shared__ int kp, count;
    const int tid = threadIdx.x;
    const long tt = (long)t * t;
    double *A = kLds ? sh : G;
    double *E = kLds ? sh + tt : W;
    double *c = kLds ? sh + 2 * tt : sh;  // column j of L
    double *er = c + t;                   // row j of E, scaled
    for (long e = tid; e < tt; e += CT) {
        if (kLds) {
            A[e] = G[e];
        }
        E[e] = (e % t) == (e / t) ? 1.0 : 0.0;
    }
    if (tid == 0) {
        count = 0;
    }
    __syncthreads();
    for (int j = 0; j < t; ++j) {
        if (tid == 0) {
            const double d = A[j + (long)j * t];
            int k = d > 0.0 && isfinite(d);
            if (k && ref != nullptr) {
                const double r = ref[j];
                k = isfinite(r) && r > DBL_MIN && sqrt(d) > rel * sqrt(r);
            }
            kp = k;
            piv = k ? sqrt(d) : 0.0;
            keep[j] = k;
            count += k;
        }
        __syncthreads();
        if (!kp) {
            // uniform; the barrier keeps thread 0 from rewriting kp before
            // every wavefront has read it
            __syncthreads();
            continue;
        }
        const double p = piv;
        const double pinv = 1.0 / p;
        for (int i = tid; i < t; i += CT) {
            // column j of L: zero above the diagonal
            const double v = i < j ? 0.0 : (i == j ? p : A[i + (long)j * t] * pinv);
            c[i] = v;
            A[i + (long)j * t] = v;
            // row j of E (nonzero in columns <= j) divided by the pivot
            const double w = i <= j ? E[j + (long)i * t] * pinv : 0.0;
            er[i] = w;
            E[j + (long)i * t] = w;
        }
        __syncthreads();
        // trailing update of G (both triangles) and of the rows of E below j
        const int w = t - j - 1;
        for (long e = tid; e < (long)w * w; e += CT) {
            const int i = j + 1 + (int)(e % w), l = j + 1 + (int)(e / w);
            A[i + (long)l * t] -= c[i] * c[l];
        }
        for (long e = tid; e < (long)w * (j + 1); e += CT) {
            const int r = j + 1 + (int)(e % w), col = (int)(e / w);
            E[r + (long)col * t] -= c[r] * er[col];
        }
        __syncthreads();
    }
    // W = E^T over the kept rows and columns; L (lower) back into G
    for (long e = tid; e < tt; e += CT) {
        const int i = (int)(e % t), l = (int)(e / t);
        if (kLds) {
            G[e] = (i >= l && keep[l]) ? A[e] : 0.0;
            W[e] = (keep[i] && keep[l]) ? E[l + (long)i * t] : 0.0;
        } else if (i < l || !keep[l]) {
            G[e] = 0.0;
        }
    }
    if (tid == 0) {
        *nkept = count;
    }
    if (!kLds) {
        // transpose E in place (W holds it); pairs (i, l), i < l, swapped once
        __syncthreads();
        for (long e = tid; e < tt; e += CT) {
            const int i = (int)(e % t), l = (int)(e / t);
            if (i < l) {
                const double a = (keep[i] && keep[l]) ? W[i + (long)l * t] : 0.0;
                const double b = (keep[i] && keep[l]) ? W[l + (long)i * t] : 0.0;
                W[i + (long)l * t] = b;
                W[l + (long)i * t] = a;
            } else if (i == l && !keep[i]) {
                W[e] = 0.0;
            }
        }
    }
}

// D(:, b) = S(:, idx[b]) for b < nc (column gather; ld t for both)
__global__ void k_gather_columns(const double *__restrict__ S, int t, const int *__restrict__ idx, int nc,
                                 double *__restrict__ D)
{
    const long total = (long)t * nc;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const int i = (int)(e % t), b = (int)(e / t);
        D[e] = S[i + (long)idx[b] * t];
    }
}

// G(i, i) += s
__global__ void k_add_diagonal(double *__restrict__ G, int t, double s)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < t) {
        G[i + (long)i * t] += s;
    }
}

}  // namespace

int chol_reject_device(Context *ctx, double *G, int t, const double *ref, double rel, double *W, int *keep,
                       int *nkept)
{
    if (t <= 0) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    if (t <= kLdsMax) {
        const size_t lds = (2 * (size_t)t * t + 2 * (size_t)t) * sizeof(double);
        static bool attr_set = false;
        if (!attr_set) {
            IPT_HIPCHK(hipFuncSetAttribute((const void *)k_chol_reject<true>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)((2 * kLdsMax * kLdsMax + 2 * kLdsMax) * sizeof(double))));
            attr_set = true;
        }
        hipLaunchKernelGGL(k_chol_reject<true>, dim3(1), dim3(CT), lds, ctx->stream, G, t, ref, rel, keep, W, nkept);
    } else {
        const size_t lds = 2 * (size_t)t * sizeof(double);
        hipLaunchKernelGGL(k_chol_reject<false>, dim3(1), dim3(CT), lds, ctx->stream, G, t, ref, rel, keep, W,
                           nkept);
    }
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

int gather_columns(Context *ctx, const double *S, int t, const int *idx, int nc, double *D)
{
    if (nc <= 0) {
        return IPT_CUDA_SUCCESS;
    }
    const long total = (long)t * nc;
    hipLaunchKernelGGL(k_gather_columns, dim3((unsigned)std::min<long>((total + 255) / 256, 4096)), dim3(256), 0,
                       ctx->stream, S, t, idx, nc, D);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

// One Cholesky QR pass on Q (n x k, ld n): G = Q^T Q + shift I = L L^T,
// Q <- Q L^-T.  IPT_CUDA_INVALID_ARGUMENT when G is not numerically positive
// definite.
static int cholqr_pass(Context *ctx, double *Q, int n, int k, double shift, double *G, double *W, double *T,
                       int *flags)
{
    IPT_STCHK(gemm_tn_tall(ctx, Q, n, Q, n, G, k, n, k, k));
    IPT_STCHK(symmetrize(ctx, G, k, k));
    if (shift > 0.0) {
        hipLaunchKernelGGL(k_add_diagonal, dim3((k + 255) / 256), dim3(256), 0, ctx->stream, G, k, shift);
        IPT_HIPCHK(hipGetLastError());
    }
    IPT_STCHK(chol_reject_device(ctx, G, k, nullptr, 0.0, W, flags, flags + k));
    int kept = 0;
    IPT_HIPCHK(hipMemcpyAsync(&kept, flags + k, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    if (kept != k) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    IPT_STCHK(gemm_nn_small(ctx, Q, n, W, k, T, n, n, k, k, 1.0, 0.0));
    IPT_HIPCHK(hipMemcpyAsync(Q, T, (size_t)n * k * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    return IPT_CUDA_SUCCESS;
}

// Shifted Cholesky QR 3 (Fukaya, Kannan, Nakatsukasa, Yamamoto, Yanagisawa,
// SIAM J. Sci. Comput. 42(1), 2020): a first pass on G + s I with
// s = 11 (n k + k (k + 1)) u ||Q||_F^2 brings the condition number to
// O(u^-1/2) for any numerically full-rank Q, then two plain passes (Cholesky
// QR 2) make Q orthonormal to O(u).  Columns are normalised first, so
// ||Q||_F^2 = k.  The orthonormal factor is the Q of Q = Q_new R with R
// upper triangular and positive diagonal.
int cholqr3_device(Context *ctx, double *Q, int n, int k)
{
    if (Q == nullptr || n <= 0 || k <= 0 || k > n) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    DevBuf<double> sq, G, W, T;
    DevBuf<int> flags;
    IPT_STCHK(sq.alloc((size_t)k));
    IPT_STCHK(column_sqnorms(ctx, Q, n, n, k, sq.p));
    std::vector<double> hsq((size_t)k);
    IPT_HIPCHK(hipMemcpyAsync(hsq.data(), sq.p, hsq.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    for (double x : hsq) {
        if (!(x > 0.0) || !std::isfinite(x)) {
            return IPT_CUDA_INVALID_ARGUMENT;  // zero or non-finite column
        }
    }
    IPT_STCHK(scale_columns_inv_sqrt(ctx, Q, nullptr, n, n, k, sq.p));
    IPT_STCHK(G.alloc((size_t)k * k));
    IPT_STCHK(W.alloc((size_t)k * k));
    IPT_STCHK(T.alloc((size_t)n * k));
    IPT_STCHK(flags.alloc((size_t)k + 1));
    const double u = DBL_EPSILON / 2;
    const double shift = 11.0 * ((double)n * k + (double)k * (k + 1)) * u * (double)k;
    IPT_STCHK(cholqr_pass(ctx, Q, n, k, shift, G.p, W.p, T.p, flags.p));
    IPT_STCHK(cholqr_pass(ctx, Q, n, k, 0.0, G.p, W.p, T.p, flags.p));
    IPT_STCHK(cholqr_pass(ctx, Q, n, k, 0.0, G.p, W.p, T.p, flags.p));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return IPT_CUDA_SUCCESS;
}

}  // namespace ipt
