// Symmetric eigendecomposition on the device for the projected problems of
// Rayleigh-Ritz (orders up to a few thousand), by two-sided block Jacobi.
//
// The matrix is padded to M = 32 NBK rows (NBK even; padding rows and
// columns zero, padding diagonal far above the spectrum, so padding never
// rotates with real indices and its eigenvalues sort last).  A sweep visits
// every pair (p, q) of 32-row blocks once, in round-robin order: the NBK/2
// pairs of a round are disjoint.  Per round
//   k_jac_pair : one workgroup per pair diagonalises its 64 x 64 block
//                [[A_pp, A_pq], [A_qp, A_qq]] by cyclic Jacobi in LDS
//                (Golub & Van Loan, Matrix Computations, 4th ed., 8.5,
//                parallel ordering of section 8.5.5) and writes its 64 x 64
//                orthogonal factor Q;
//   k_jac_rows : [A_p; A_q] <- Q^T [A_p; A_q]   (64-column tiles)
//   k_jac_cols : [A_.p A_.q] <- [A_.p A_.q] Q, the same for V (64-row tiles)
// The off-diagonal Frobenius norm is checked after each sweep (one read of
// two scalars); converged when it falls below 1e-15 of the matrix norm, or
// below 1e-12 of it and no longer shrinking (the rounding floor).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "ipt_internal.h"

namespace ipt {

namespace {

constexpr int JB = 32;       // block rows
constexpr int JS = 2 * JB;   // subproblem order
constexpr int JT = 256;      // threads per workgroup

// pair i of round r among nb (even) players by the circle method: player 0
// fixed, the other nb - 1 on a circle rotated by one place per round; pair
// i > 0 joins the places r + i and r - i.  Over rounds 0..nb-2 every pair
// meets once, and the pairs of a round are disjoint.
__host__ __device__ inline void round_pair(int nb, int r, int i, int *p, int *q)
{
    const int m = nb - 1;
    int a, b;
    if (i == 0) {
        a = 0;
        b = 1 + r % m;
    } else {
        a = 1 + (r + i) % m;
        b = 1 + (r + m - i) % m;
    }
    *p = a < b ? a : b;
    *q = a < b ? b : a;
}

// pair i of inner round r among JS indices: the same tournament
__device__ inline void inner_pair(int r, int i, int *p, int *q)
{
    round_pair(JS, r, i, p, q);
}

__global__ void k_jac_init(const double *__restrict__ A, long lda, int m, double *__restrict__ Ap, double *__restrict__ V,
                           int M, double big)
{
    const long total = (long)M * M;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % M), j = (int)(t / M);
        double a = 0.0;
        if (i < m && j < m) {
            a = A[i + (long)j * lda];
        } else if (i == j) {
            a = big + (double)(i - m);  // distinct padding values, above the spectrum
        }
        Ap[t] = a;
        V[t] = i == j ? 1.0 : 0.0;
    }
}

// Diagonalise the 64 x 64 block of pair (round, blockIdx.x) in LDS; Q out.
__global__ __launch_bounds__(JT) void k_jac_pair(const double *__restrict__ Ap, int M, int nb, int round,
                                                 double *__restrict__ Qout, int inner_sweeps)
{
    __shared__ double S[JS][JS + 1];
    __shared__ double Q[JS][JS + 1];
    __shared__ double rc[JS / 2], rs[JS / 2];
    __shared__ double red[JT / 64];
    __shared__ int done, rotated;
    static_assert(JT % 64 == 0, "whole wavefronts");
    int bp, bq;
    round_pair(nb, round, blockIdx.x, &bp, &bq);
    const int t = threadIdx.x;
    // local index l < 32 -> global row 32 bp + l, l >= 32 -> 32 bq + l - 32
    for (int e = t; e < JS * JS; e += JT) {
        const int i = e % JS, j = e / JS;
        const int gi = i < JB ? JB * bp + i : JB * bq + i - JB;
        const int gj = j < JB ? JB * bp + j : JB * bq + j - JB;
        S[i][j] = Ap[gi + (long)gj * M];
        Q[i][j] = i == j ? 1.0 : 0.0;
    }
    __syncthreads();
    // the Frobenius norm is invariant under the rotations
    if (t == 0) {
        rotated = 0;
    }
    double tot = 0.0;
    for (int e = t; e < JS * JS; e += JT) {
        tot += S[e % JS][e / JS] * S[e % JS][e / JS];
    }
    for (int o = 32; o > 0; o >>= 1) {
        tot += __shfl_down(tot, o, 64);
    }
    if ((t & 63) == 0) {
        red[t >> 6] = tot;
    }
    __syncthreads();
    tot = 0.0;
    for (int w = 0; w < JT / 64; ++w) {
        tot += red[w];
    }
    __syncthreads();
    for (int sweep = 0; sweep < inner_sweeps; ++sweep) {
        double off = 0.0;
        for (int e = t; e < JS * JS; e += JT) {
            const int i = e % JS, j = e / JS;
            if (i != j) {
                off += S[i][j] * S[i][j];
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            off += __shfl_down(off, o, 64);
        }
        if ((t & 63) == 0) {
            red[t >> 6] = off;
        }
        __syncthreads();
        if (t == 0) {
            double so = 0.0;
            for (int w = 0; w < JT / 64; ++w) {
                so += red[w];
            }
            // off-diagonal below 1e-14 of the norm, or the last sweep found
            // nothing above the rotation threshold
            done = !(so > 1e-28 * tot) || (sweep > 0 && !rotated);
            rotated = 0;
        }
        __syncthreads();
        if (done) {
            break;
        }
        for (int r = 0; r < JS - 1; ++r) {
            if (t < JS / 2) {
                int p, q;
                inner_pair(r, t, &p, &q);
                const double a = S[p][p], b = S[q][q], c = S[p][q];
                double cs = 1.0, sn = 0.0;
                if (c != 0.0 && fabs(c) > 1e-17 * sqrt(fabs(a) * fabs(b))) {
                    rotated = 1;
                    const double tau = (b - a) / (2.0 * c);
                    const double tt = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                    cs = 1.0 / sqrt(1.0 + tt * tt);
                    sn = tt * cs;
                }
                rc[t] = cs;
                rs[t] = sn;
            }
            __syncthreads();
            // rows: S <- J^T S
            for (int e = t; e < (JS / 2) * JS; e += JT) {
                const int i = e % (JS / 2), j = e / (JS / 2);
                int p, q;
                inner_pair(r, i, &p, &q);
                const double cs = rc[i], sn = rs[i];
                const double sp = S[p][j], sq = S[q][j];
                S[p][j] = cs * sp - sn * sq;
                S[q][j] = sn * sp + cs * sq;
            }
            __syncthreads();
            // columns: S <- S J, Q <- Q J
            for (int e = t; e < (JS / 2) * JS; e += JT) {
                const int i = e % (JS / 2), j = e / (JS / 2);
                int p, q;
                inner_pair(r, i, &p, &q);
                const double cs = rc[i], sn = rs[i];
                const double sp = S[j][p], sq = S[j][q];
                S[j][p] = cs * sp - sn * sq;
                S[j][q] = sn * sp + cs * sq;
                const double qp = Q[j][p], qq = Q[j][q];
                Q[j][p] = cs * qp - sn * qq;
                Q[j][q] = sn * qp + cs * qq;
            }
            __syncthreads();
        }
    }
    // Q is a product of up to a few hundred rotations per column; one
    // Newton-Schulz step Q (3I - Q^T Q) / 2 (quadratic in the departure
    // from orthogonality) keeps the outer similarity orthogonal to rounding,
    // so the spectrum does not drift over the outer rounds.  S is free now.
    for (int e = t; e < JS * JS; e += JT) {
        const int i = e % JS, j = e / JS;
        double acc = 0.0;
        for (int l = 0; l < JS; ++l) {
            acc = fma(Q[l][i], Q[l][j], acc);
        }
        S[i][j] = acc;
    }
    __syncthreads();
    double *Qo = Qout + (long)blockIdx.x * JS * JS;
    for (int e = t; e < JS * JS; e += JT) {
        const int i = e % JS, j = e / JS;
        double acc = 0.0;
        for (int l = 0; l < JS; ++l) {
            acc = fma(Q[i][l], S[l][j], acc);
        }
        Qo[e] = 1.5 * Q[i][j] - 0.5 * acc;
    }
}

// [A_p; A_q](:, tile) <- Q^T [A_p; A_q](:, tile): one workgroup per (pair,
// 64-column tile)
__global__ __launch_bounds__(JT) void k_jac_rows(double *__restrict__ Ap, int M, int nb, int round,
                                                 const double *__restrict__ Qall)
{
    __shared__ double Qs[JS][JS + 1];
    __shared__ double X[JS][JS + 1];
    int bp, bq;
    round_pair(nb, round, blockIdx.x, &bp, &bq);
    const double *Q = Qall + (long)blockIdx.x * JS * JS;
    const int c0 = blockIdx.y * JS;
    const int t = threadIdx.x;
    for (int e = t; e < JS * JS; e += JT) {
        const int i = e % JS, j = e / JS;
        Qs[i][j] = Q[e];
        const int gi = i < JB ? JB * bp + i : JB * bq + i - JB;
        const int gj = c0 + j;
        X[i][j] = gj < M ? Ap[gi + (long)gj * M] : 0.0;
    }
    __syncthreads();
    for (int e = t; e < JS * JS; e += JT) {
        const int i = e % JS, j = e / JS;
        const int gj = c0 + j;
        double acc = 0.0;
#pragma unroll 8
        for (int l = 0; l < JS; ++l) {
            acc = fma(Qs[l][i], X[l][j], acc);
        }
        const int gi = i < JB ? JB * bp + i : JB * bq + i - JB;
        if (gj < M) {
            Ap[gi + (long)gj * M] = acc;
        }
    }
}

// [A_.p A_.q](tile, :) <- [A_.p A_.q](tile, :) Q and the same for V: one
// workgroup per (pair, 64-row tile)
__global__ __launch_bounds__(JT) void k_jac_cols(double *__restrict__ Ap, double *__restrict__ V, int M, int nb,
                                                 int round, const double *__restrict__ Qall)
{
    __shared__ double Qs[JS][JS + 1];
    __shared__ double X[JS][JS + 1];
    int bp, bq;
    round_pair(nb, round, blockIdx.x, &bp, &bq);
    const double *Q = Qall + (long)blockIdx.x * JS * JS;
    const int r0 = blockIdx.y * JS;
    const int t = threadIdx.x;
    for (int e = t; e < JS * JS; e += JT) {
        Qs[e % JS][e / JS] = Q[e];
    }
    for (int which = 0; which < 2; ++which) {
        double *Mx = which == 0 ? Ap : V;
        __syncthreads();
        for (int e = t; e < JS * JS; e += JT) {
            const int i = e % JS, j = e / JS;
            const int gi = r0 + i;
            const int gj = j < JB ? JB * bp + j : JB * bq + j - JB;
            X[i][j] = gi < M ? Mx[gi + (long)gj * M] : 0.0;
        }
        __syncthreads();
        for (int e = t; e < JS * JS; e += JT) {
            const int i = e % JS, j = e / JS;
            const int gi = r0 + i;
            double acc = 0.0;
#pragma unroll 8
            for (int l = 0; l < JS; ++l) {
                acc = fma(X[i][l], Qs[l][j], acc);
            }
            const int gj = j < JB ? JB * bp + j : JB * bq + j - JB;
            if (gi < M) {
                Mx[gi + (long)gj * M] = acc;
            }
        }
    }
}

// out[0] = sum of squared off-diagonal entries, out[1] = sum of squares
// (fixed-order two-level reduction: per-block partials, then one block)
__global__ __launch_bounds__(256) void k_jac_norms(const double *__restrict__ Ap, int M, double *__restrict__ part)
{
    double off = 0.0, all = 0.0;
    const long total = (long)M * M;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const double v = Ap[e] * Ap[e];
        all += v;
        if (e % M != e / M) {
            off += v;
        }
    }
    __shared__ double so[4], sa[4];
    for (int o = 32; o > 0; o >>= 1) {
        off += __shfl_down(off, o, 64);
        all += __shfl_down(all, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        so[threadIdx.x >> 6] = off;
        sa[threadIdx.x >> 6] = all;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = so[0] + so[1] + so[2] + so[3];
        part[2 * blockIdx.x + 1] = sa[0] + sa[1] + sa[2] + sa[3];
    }
}

__global__ void k_jac_norms_final(const double *__restrict__ part, int nparts, double *__restrict__ out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double off = 0.0, all = 0.0;
        for (int b = 0; b < nparts; ++b) {
            off += part[2 * b];
            all += part[2 * b + 1];
        }
        out[0] = off;
        out[1] = all;
    }
}

// w[i] = Ap[perm[i], perm[i]], Vout(:, i) = V(0:m, perm[i]) for i < m
__global__ void k_jac_gather(const double *__restrict__ Ap, const double *__restrict__ V, int M, int m,
                             const int *__restrict__ perm, double *__restrict__ w, double *__restrict__ Vout, long ldv)
{
    const long total = (long)m * m;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % m), j = (int)(t / m);
        const int pj = perm[j];
        Vout[i + (long)j * ldv] = V[i + (long)pj * M];
        if (i == 0) {
            w[j] = Ap[pj + (long)pj * M];
        }
    }
}

__global__ void k_jac_diag(const double *__restrict__ Ap, int M, double *__restrict__ d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) {
        d[i] = Ap[i + (long)i * M];
    }
}

}  // namespace

// Eigenvalues (ascending) w[0..m) and eigenvectors V (m x m, ldv) of the
// symmetric A (m x m, lda; read only).  *sweeps_out (may be NULL): outer
// sweeps used.  Returns IPT_CUDA_CUSOLVER_ERROR when 40 sweeps do not
// converge.
int sym_eig_device(Context *ctx, const double *A, long lda, int m, double *w, double *V, long ldv, int *sweeps_out)
{
    if (m <= 0) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int nb = (m + JB - 1) / JB;
    nb += nb & 1;
    if (nb < 2) {
        nb = 2;
    }
    const int M = nb * JB;
    const int npairs = nb / 2;
    const int nparts = 256;
    DevBuf<double> Ap, Vp, Q, part, red, dg;
    DevBuf<int> dperm;
    IPT_STCHK(Ap.alloc((size_t)M * M));
    IPT_STCHK(Vp.alloc((size_t)M * M));
    IPT_STCHK(Q.alloc((size_t)npairs * JS * JS));
    IPT_STCHK(part.alloc(2 * (size_t)nparts));
    IPT_STCHK(red.alloc(2));
    IPT_STCHK(dg.alloc((size_t)M));
    IPT_STCHK(dperm.alloc((size_t)m));
    // padding above the spectrum: |lambda| <= ||A||_F; the norms pass gives it
    hipLaunchKernelGGL(k_jac_init, dim3(std::min(4096, (M * M + 255) / 256)), dim3(256), 0, ctx->stream, A, lda, m,
                       Ap.p, Vp.p, M, 0.0);
    IPT_HIPCHK(hipGetLastError());
    double h[2] = {0.0, 0.0};
    auto norms = [&]() -> int {
        hipLaunchKernelGGL(k_jac_norms, dim3(nparts), dim3(256), 0, ctx->stream, Ap.p, M, part.p);
        hipLaunchKernelGGL(k_jac_norms_final, dim3(1), dim3(64), 0, ctx->stream, part.p, nparts, red.p);
        IPT_HIPCHK(hipGetLastError());
        IPT_HIPCHK(hipMemcpyAsync(h, red.p, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
        return IPT_CUDA_SUCCESS;
    };
    IPT_STCHK(norms());
    if (!std::isfinite(h[1])) {
        return IPT_CUDA_CUSOLVER_ERROR;
    }
    if (M > m) {
        const double big = 2.0 * std::sqrt(h[1]) + 1.0;
        hipLaunchKernelGGL(k_jac_init, dim3(std::min(4096, (M * M + 255) / 256)), dim3(256), 0, ctx->stream, A, lda,
                           m, Ap.p, Vp.p, M, big);
        IPT_HIPCHK(hipGetLastError());
        IPT_STCHK(norms());
    }
    // converged when the off-diagonal Frobenius norm is below 1e-15 of the
    // matrix's, or below 1e-12 of it and no longer halving (the rounding
    // floor of an order-M matrix is about eps M of its norm)
    const double total = h[1];
    int sweep = 0;
    bool converged = !(h[0] > 1e-30 * total);
    double prev = h[0];
    const dim3 tiles(npairs, M / JS);
    for (sweep = 0; sweep < 40 && !converged; ++sweep) {
        for (int r = 0; r < nb - 1; ++r) {
            hipLaunchKernelGGL(k_jac_pair, dim3(npairs), dim3(JT), 0, ctx->stream, Ap.p, M, nb, r, Q.p, 12);
            hipLaunchKernelGGL(k_jac_rows, tiles, dim3(JT), 0, ctx->stream, Ap.p, M, nb, r, Q.p);
            hipLaunchKernelGGL(k_jac_cols, tiles, dim3(JT), 0, ctx->stream, Ap.p, Vp.p, M, nb, r, Q.p);
            IPT_HIPCHK(hipGetLastError());
        }
        IPT_STCHK(norms());
        converged = !(h[0] > 1e-30 * total) || (h[0] <= 1e-24 * total && h[0] >= 0.25 * prev);
        prev = h[0];
    }
    if (sweeps_out) {
        *sweeps_out = sweep;
    }
    if (!converged || !std::isfinite(h[0])) {
        fprintf(stderr, "IPT: block Jacobi eigensolver did not converge (off-diagonal %.3e of %.3e)\n",
                std::sqrt(h[0]), std::sqrt(total));
        return IPT_CUDA_CUSOLVER_ERROR;
    }
    // ascending order (the padding values are above every real one)
    std::vector<double> hd((size_t)M);
    hipLaunchKernelGGL(k_jac_diag, dim3((M + 255) / 256), dim3(256), 0, ctx->stream, Ap.p, M, dg.p);
    IPT_HIPCHK(hipGetLastError());
    IPT_HIPCHK(hipMemcpyAsync(hd.data(), dg.p, (size_t)M * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    std::vector<int> all((size_t)M);
    std::iota(all.begin(), all.end(), 0);
    std::stable_sort(all.begin(), all.end(), [&](int a, int b) { return hd[(size_t)a] < hd[(size_t)b]; });
    const std::vector<int> perm(all.begin(), all.begin() + m);
    IPT_HIPCHK(hipMemcpyAsync(dperm.p, perm.data(), (size_t)m * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_jac_gather, dim3(std::min(4096, (m * m + 255) / 256)), dim3(256), 0, ctx->stream, Ap.p, Vp.p,
                       M, m, dperm.p, w, V, ldv);
    IPT_HIPCHK(hipGetLastError());
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return IPT_CUDA_SUCCESS;
}

}  // namespace ipt

using namespace ipt;

extern "C" int ipt_hip_sym_eig(const double *d_A, int lda, int m, double *d_w, double *d_V, int ldv, void *handle)
{
    if (d_A == nullptr || d_w == nullptr || d_V == nullptr || m <= 0 || lda < m || ldv < m) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    return sym_eig_device(ctx, d_A, lda, m, d_w, d_V, ldv, nullptr);
}
