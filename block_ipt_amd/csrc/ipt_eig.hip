// Symmetric eigendecomposition on the device for the projected problems of
// Rayleigh-Ritz (orders up to a few thousand), by two-sided block Jacobi.
//
// The matrix is padded to M = JB NBK rows (NBK even; padding rows and
// columns zero, padding diagonal far above the spectrum, so padding never
// rotates with real indices and its eigenvalues sort last).  A sweep visits
// every pair (p, q) of JB-row blocks once, in round-robin order: the NBK/2
// pairs of a round are disjoint and cover every block.  Per round
//   k_jac_pair  : one workgroup per pair runs cyclic Jacobi sweeps (Golub &
//                 Van Loan, Matrix Computations, 4th ed., 8.5, parallel
//                 ordering of section 8.5.5) on its 2JB x 2JB block
//                 [[A_pp, A_pq], [A_qp, A_qq]] in LDS and writes the
//                 accumulated orthogonal factor Q (or a flag that it did not
//                 rotate).  Each inner round updates the block by its JB
//                 disjoint rotations at once: every 2 x 2 sub-block (pair i,
//                 pair j) is rotated from both sides by one thread.  A
//                 bounded number of inner sweeps (one by default) suffices:
//                 the outer iteration converges in as many sweeps as with
//                 fully diagonalised subproblems, at a fraction of the
//                 serial depth;
//   k_jac_apply : every tile (pair P, pair R) of A becomes Q_P^T A_PR Q_R
//                 (the upper tiles, mirrored), and every row tile of V
//                 becomes V_TR Q_R; pairs that did not rotate are skipped.
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

constexpr int JB = 16;       // block rows
constexpr int JS = 2 * JB;   // subproblem order
constexpr int NR = JB;       // rotations per inner round
constexpr int PT = 256;      // threads of the subproblem workgroup
constexpr int AT = 256;      // threads of the apply workgroup
constexpr int NPARTS = 256;  // partial sums of the norm pass

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

// global row of local index l of the pair (bp, bq)
__device__ inline int pair_row(int l, int bp, int bq)
{
    return l < JB ? JB * bp + l : JB * bq + l - JB;
}

// (i, j), i <= j < n, of the b-th entry of an upper triangle stored by
// columns: j = floor((sqrt(8b + 1) - 1) / 2), corrected for rounding
__device__ inline void tri_index(int b, int *i, int *j)
{
    int c = (int)((sqrtf(8.0f * (float)b + 1.0f) - 1.0f) * 0.5f);
    while (c * (c + 1) / 2 > b) {
        --c;
    }
    while ((c + 1) * (c + 2) / 2 <= b) {
        ++c;
    }
    *j = c;
    *i = b - c * (c + 1) / 2;
}

// 1/x and 1/sqrt(x) from the hardware estimates (relative error near 2^-24)
// and two Newton steps each (2^-48, then rounding level)
__device__ inline double rcp_refined(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

__device__ inline double rsq_refined(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-(x * y), y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-(x * y), y, 1.0);
    return fma(0.5 * y, e, y);
}

// The rotation J = [[c, s], [-s, c]] with J^T [[a, x], [x, b]] J diagonal,
// |angle| <= pi/4 (Golub & Van Loan, Algorithm 8.5.1): t = tan(angle) =
// sign(d) 2x / (|d| + sqrt(d^2 + 4x^2)), d = b - a, evaluated on operands
// scaled by a power of two (exact) so that the squares neither overflow nor
// underflow; c = (1 + t^2)^-1/2, s = t c.  Requires x != 0.
__device__ inline void jacobi_rotation(double a, double b, double x, double *cs, double *sn, double *tn)
{
    const double d = b - a;
    int ex;
    (void)frexp(fmax(fabs(d), 2.0 * fabs(x)), &ex);
    const double dn = ldexp(d, -ex), xn = ldexp(x, -ex);
    const double h2 = fma(dn, dn, 4.0 * xn * xn);  // in [1/4, 2)
    const double h = h2 * rsq_refined(h2);
    const double t = (d >= 0.0 ? 2.0 : -2.0) * xn * rcp_refined(fabs(dn) + h);
    const double c = rsq_refined(fma(t, t, 1.0));
    *cs = c;
    *sn = t * c;
    *tn = t;
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

// Cyclic Jacobi on the 2JB x 2JB block of pair (round, blockIdx.x), at most
// `inner_sweeps` sweeps, stopping early when its off-diagonal sum of squares
// is at most abs_thr2 or a sweep rotated nothing.  Writes qflag (1 when any
// rotation was applied) and, if so, the factor Q (column-major JS x JS).
// Thread t < JB(JB+1)/2 owns the 2 x 2 sub-block (pair i, pair j), i <= j,
// of its triangle index in every inner round, and two (row, pair) entries
// of Q; the pair indices of a round are recomputed in registers.
__global__ __launch_bounds__(PT) void k_jac_pair(const double *__restrict__ Ap, int M, int nb, int round,
                                                 double *__restrict__ Qout, int *__restrict__ qflag,
                                                 int inner_sweeps, double abs_thr2)
{
    __shared__ double S[JS][JS + 1];
    __shared__ double Q[JS][JS + 1];
    __shared__ double rc[NR], rs[NR], rt[NR];
    __shared__ double red[PT / 64];
    __shared__ int sweep_rot;
    static_assert(PT % 64 == 0 && PT >= NR * (NR + 1) / 2 && (JS * NR) % PT == 0, "thread mapping");
    constexpr int QPT = JS * NR / PT;  // Q entries per thread
    int bp, bq;
    round_pair(nb, round, blockIdx.x, &bp, &bq);
    const int t = threadIdx.x;
    for (int e = t; e < JS * JS; e += PT) {
        const int i = e % JS, j = e / JS;
        S[i][j] = Ap[pair_row(i, bp, bq) + (long)pair_row(j, bp, bq) * M];
        Q[i][j] = i == j ? 1.0 : 0.0;
    }
    const bool owns_block = t < NR * (NR + 1) / 2;
    int bi = 0, bj = 0;
    if (owns_block) {
        tri_index(t, &bi, &bj);
    }
    int any_rot = 0;
    __syncthreads();
    for (int sweep = 0; sweep < inner_sweeps; ++sweep) {
        double off = 0.0;
        for (int e = t; e < JS * JS; e += PT) {
            const int i = e % JS, j = e / JS;
            if (i != j) {
                off += S[i][j] * S[i][j];
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            off += __shfl_xor(off, o, 64);
        }
        if ((t & 63) == 0) {
            red[t >> 6] = off;
        }
        if (t == 0) {
            sweep_rot = 0;
        }
        __syncthreads();
        off = 0.0;
        for (int w = 0; w < PT / 64; ++w) {
            off += red[w];
        }
        if (!(off > abs_thr2)) {
            break;
        }
        for (int r = 0; r < JS - 1; ++r) {
            if (t < NR) {
                int p, q;
                round_pair(JS, r, t, &p, &q);
                const double a = S[p][p], b = S[q][q], c = S[p][q];
                double cs = 1.0, sn = 0.0, tn = 0.0;
                // rotate unless |c| <= 1e-17 sqrt(|a b|)
                if (c * c > (1e-17 * fabs(a)) * (1e-17 * fabs(b))) {
                    jacobi_rotation(a, b, c, &cs, &sn, &tn);
                    sweep_rot = 1;
                }
                rc[t] = cs;
                rs[t] = sn;
                rt[t] = tn;
            }
            __syncthreads();
            // S <- J^T S J by 2 x 2 sub-blocks, the lower ones mirrored;
            // J = [[c, s], [-s, c]] per pair
            if (owns_block) {
                int pi, qi;
                round_pair(JS, r, bi, &pi, &qi);
                if (bi == bj) {
                    const double tn = rt[bi];
                    if (tn != 0.0) {
                        const double x = S[pi][qi];
                        S[pi][pi] -= tn * x;
                        S[qi][qi] += tn * x;
                        S[pi][qi] = 0.0;
                        S[qi][pi] = 0.0;
                    }
                } else {
                    int pj, qj;
                    round_pair(JS, r, bj, &pj, &qj);
                    const double ci = rc[bi], si = rs[bi], cj = rc[bj], sj = rs[bj];
                    const double x00 = S[pi][pj], x01 = S[pi][qj], x10 = S[qi][pj], x11 = S[qi][qj];
                    const double y00 = ci * x00 - si * x10, y01 = ci * x01 - si * x11;
                    const double y10 = si * x00 + ci * x10, y11 = si * x01 + ci * x11;
                    const double z00 = cj * y00 - sj * y01, z01 = sj * y00 + cj * y01;
                    const double z10 = cj * y10 - sj * y11, z11 = sj * y10 + cj * y11;
                    S[pi][pj] = z00;
                    S[pi][qj] = z01;
                    S[qi][pj] = z10;
                    S[qi][qj] = z11;
                    S[pj][pi] = z00;
                    S[qj][pi] = z01;
                    S[pj][qi] = z10;
                    S[qj][qi] = z11;
                }
            }
            // Q <- Q J
#pragma unroll
            for (int u = 0; u < QPT; ++u) {
                const int e = t + u * PT;
                const int row = e % JS, j = e / JS;
                int pj, qj;
                round_pair(JS, r, j, &pj, &qj);
                const double cj = rc[j], sj = rs[j];
                const double a = Q[row][pj], b = Q[row][qj];
                Q[row][pj] = cj * a - sj * b;
                Q[row][qj] = sj * a + cj * b;
            }
            __syncthreads();
        }
        const int rotated = sweep_rot;
        __syncthreads();  // every wavefront has read the flag before it is reset
        any_rot |= rotated;
        if (!rotated) {
            break;
        }
    }
    if (t == 0) {
        qflag[blockIdx.x] = any_rot;
    }
    if (!any_rot) {
        return;
    }
    // Q is a product of a few hundred rotations per column; one Newton-Schulz
    // step Q (3I - Q^T Q) / 2 (quadratic in the departure from orthogonality)
    // keeps the outer similarity orthogonal to rounding, so the spectrum does
    // not drift over the outer rounds.  S is free now.
    __syncthreads();
    for (int e = t; e < JS * JS; e += PT) {
        const int i = e % JS, j = e / JS;
        double acc = 0.0;
        for (int l = 0; l < JS; ++l) {
            acc = fma(Q[l][i], Q[l][j], acc);
        }
        S[i][j] = acc;
    }
    __syncthreads();
    double *Qo = Qout + (long)blockIdx.x * JS * JS;
    for (int e = t; e < JS * JS; e += PT) {
        const int i = e % JS, j = e / JS;
        double acc = 0.0;
        for (int l = 0; l < JS; ++l) {
            acc = fma(Q[i][l], S[l][j], acc);
        }
        Qo[e] = 1.5 * Q[i][j] - 0.5 * acc;
    }
}

// Y = op(L) X for JS x JS LDS tiles, 256 threads with 2 x 2 outputs each
// (rows ti, ti + 16; columns tj, tj + 16); transpose_left: L^T.
template <bool transpose_left>
__device__ inline void tile_mul(const double (*L)[JS + 1], const double (*X)[JS + 1], double acc[2][2])
{
    const int ti = threadIdx.x & 15, tj = threadIdx.x >> 4;
    acc[0][0] = acc[0][1] = acc[1][0] = acc[1][1] = 0.0;
#pragma unroll 8
    for (int l = 0; l < JS; ++l) {
        const double a0 = transpose_left ? L[l][ti] : L[ti][l];
        const double a1 = transpose_left ? L[l][ti + 16] : L[ti + 16][l];
        const double b0 = X[l][tj], b1 = X[l][tj + 16];
        acc[0][0] = fma(a0, b0, acc[0][0]);
        acc[0][1] = fma(a0, b1, acc[0][1]);
        acc[1][0] = fma(a1, b0, acc[1][0]);
        acc[1][1] = fma(a1, b1, acc[1][1]);
    }
}

// Workgroups [0, ntile_a): the upper tiles (P <= R) of A by pairs,
// A_PR <- Q_P^T A_PR Q_R and A_RP <- its transpose (a diagonal tile is
// symmetrised); the rest: row tile T (JS rows) of V and pair R,
// V_TR <- V_TR Q_R.  Identity factors (qflag 0) are skipped.
__global__ __launch_bounds__(AT) void k_jac_apply(double *__restrict__ Ap, double *__restrict__ V, int M, int nb,
                                                  int round, const double *__restrict__ Qall,
                                                  const int *__restrict__ qflag, int ntile_a, int v_rows)
{
    static_assert(AT == 256 && JS == 32, "2 x 2 outputs per thread cover a 32 x 32 tile");
    __shared__ double X[JS][JS + 1];
    __shared__ double Y[JS][JS + 1];
    __shared__ double QL[JS][JS + 1];
    __shared__ double QR[JS][JS + 1];
    const int np = nb / 2;
    const int t = threadIdx.x;
    const int ti = t & 15, tj = t >> 4;
    int id = blockIdx.x;
    if (id < ntile_a) {
        int P, R;
        tri_index(id, &P, &R);
        const int fP = qflag[P], fR = qflag[R];
        if (!fP && !fR) {
            return;
        }
        int pp, pq, rp_, rq_;
        round_pair(nb, round, P, &pp, &pq);
        round_pair(nb, round, R, &rp_, &rq_);
        const double *Qp = Qall + (long)P * JS * JS;
        const double *Qr = Qall + (long)R * JS * JS;
        for (int e = t; e < JS * JS; e += AT) {
            const int i = e % JS, j = e / JS;
            X[i][j] = Ap[pair_row(i, pp, pq) + (long)pair_row(j, rp_, rq_) * M];
            if (fP) {
                QL[i][j] = Qp[e];
            }
            if (fR) {
                QR[i][j] = Qr[e];
            }
        }
        __syncthreads();
        double acc[2][2];
        if (fP) {
            tile_mul<true>(QL, X, acc);
#pragma unroll
            for (int a = 0; a < 2; ++a) {
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    Y[ti + 16 * a][tj + 16 * b] = acc[a][b];
                }
            }
        } else {
            for (int e = t; e < JS * JS; e += AT) {
                Y[e % JS][e / JS] = X[e % JS][e / JS];
            }
        }
        __syncthreads();
        // X <- Y Q_R: (Q_R^T Y^T)^T, with Y read transposed
        if (fR) {
            const int ri = t & 15, rj = t >> 4;
            double o[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
#pragma unroll 8
            for (int l = 0; l < JS; ++l) {
                const double a0 = Y[ri][l], a1 = Y[ri + 16][l];
                const double b0 = QR[l][rj], b1 = QR[l][rj + 16];
                o[0][0] = fma(a0, b0, o[0][0]);
                o[0][1] = fma(a0, b1, o[0][1]);
                o[1][0] = fma(a1, b0, o[1][0]);
                o[1][1] = fma(a1, b1, o[1][1]);
            }
            __syncthreads();
#pragma unroll
            for (int a = 0; a < 2; ++a) {
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    X[ri + 16 * a][rj + 16 * b] = o[a][b];
                }
            }
        } else {
            for (int e = t; e < JS * JS; e += AT) {
                X[e % JS][e / JS] = Y[e % JS][e / JS];
            }
        }
        __syncthreads();
        for (int e = t; e < JS * JS; e += AT) {
            const int i = e % JS, j = e / JS;
            double v = X[i][j];
            if (P == R) {
                v = 0.5 * (v + X[j][i]);
            }
            Ap[pair_row(i, pp, pq) + (long)pair_row(j, rp_, rq_) * M] = v;
            if (P != R) {
                Ap[pair_row(i, rp_, rq_) + (long)pair_row(j, pp, pq) * M] = X[j][i];
            }
        }
        return;
    }
    id -= ntile_a;
    const int R = id % np, T = id / np;
    if (!qflag[R]) {
        return;
    }
    int rp_, rq_;
    round_pair(nb, round, R, &rp_, &rq_);
    const double *Qr = Qall + (long)R * JS * JS;
    const int r0 = T * JS;
    for (int e = t; e < JS * JS; e += AT) {
        const int i = e % JS, j = e / JS;
        const int gi = r0 + i;
        X[i][j] = gi < v_rows ? V[gi + (long)pair_row(j, rp_, rq_) * M] : 0.0;
        QR[i][j] = Qr[e];
    }
    __syncthreads();
    double o[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
#pragma unroll 8
    for (int l = 0; l < JS; ++l) {
        const double a0 = X[ti][l], a1 = X[ti + 16][l];
        const double b0 = QR[l][tj], b1 = QR[l][tj + 16];
        o[0][0] = fma(a0, b0, o[0][0]);
        o[0][1] = fma(a0, b1, o[0][1]);
        o[1][0] = fma(a1, b0, o[1][0]);
        o[1][1] = fma(a1, b1, o[1][1]);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a) {
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int gi = r0 + ti + 16 * a;
            if (gi < v_rows) {
                V[gi + (long)pair_row(tj + 16 * b, rp_, rq_) * M] = o[a][b];
            }
        }
    }
}

// part[2b] = sum of squared off-diagonal entries, part[2b + 1] = sum of
// squares, over the entries of grid-stride block b
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

// out = sums of the NPARTS partial pairs (fixed order: one wavefront, a
// strided sum per lane, then a butterfly)
__global__ __launch_bounds__(64) void k_jac_norms_final(const double *__restrict__ part, double *__restrict__ out)
{
    double off = 0.0, all = 0.0;
    for (int b = threadIdx.x; b < NPARTS; b += 64) {
        off += part[2 * b];
        all += part[2 * b + 1];
    }
    for (int o = 32; o > 0; o >>= 1) {
        off += __shfl_xor(off, o, 64);
        all += __shfl_xor(all, o, 64);
    }
    if (threadIdx.x == 0) {
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

// Padding diagonal entries i = m..M-1 set to big + (i - m), big = 2 ||Ap||_F + 1
// from red[1] (the sum of squares after k_jac_init with big = 0): distinct
// values above the spectrum, computed without a host round trip.
__global__ void k_jac_pad(double *__restrict__ Ap, int M, int m, const double *__restrict__ red)
{
    const int i = m + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) {
        Ap[i + (long)i * M] = (2.0 * sqrt(red[1]) + 1.0) + (double)(i - m);
    }
}

// perm[r] = the index of the diagonal entry of rank r (ascending, ties by
// index: a stable sort) for r < m, by counting; d has M entries.
__global__ void k_jac_order(const double *__restrict__ d, int M, int m, int *__restrict__ perm)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) {
        return;
    }
    const double di = d[i];
    int r = 0;
    for (int j = 0; j < M; ++j) {
        const double dj = d[j];
        r += (dj < di || (dj == di && j < i)) ? 1 : 0;
    }
    if (r < m) {
        perm[r] = i;
    }
}

}  // namespace

// Eigenvalues (ascending) w[0..m) and eigenvectors V (m x m, ldv) of the
// symmetric A (m x m, lda; read only).  *sweeps_out (may be NULL): outer
// sweeps used.  Returns IPT_CUDA_CUSOLVER_ERROR when 40 sweeps do not
// converge.  The results are complete in stream order on ctx->stream (the
// host does not wait for them).
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
    const int ntile_a = npairs * (npairs + 1) / 2;
    // row tiles of V that hold real rows (padding rows of V stay zero in
    // the real columns and never mix with them)
    const int vtiles = (m + JS - 1) / JS;
    const int inner = std::max(1, env_int("IPT_EIG_INNER_SWEEPS", 1));
    DevBuf<double> Ap, Vp, Q, part, red, dg;
    DevBuf<int> dperm, qflag;
    IPT_STCHK(Ap.alloc((size_t)M * M));
    IPT_STCHK(Vp.alloc((size_t)M * M));
    IPT_STCHK(Q.alloc((size_t)npairs * JS * JS));
    IPT_STCHK(qflag.alloc((size_t)npairs));
    IPT_STCHK(part.alloc(2 * (size_t)NPARTS));
    IPT_STCHK(red.alloc(2));
    IPT_STCHK(dg.alloc((size_t)M));
    IPT_STCHK(dperm.alloc((size_t)m));
    const int init_blocks = std::min(4096, (M * M + 255) / 256);
    // padding above the spectrum: |lambda| <= ||A||_F; the norms pass gives it
    hipLaunchKernelGGL(k_jac_init, dim3(init_blocks), dim3(256), 0, ctx->stream, A, lda, m, Ap.p, Vp.p, M, 0.0);
    IPT_HIPCHK(hipGetLastError());
    double h[2] = {0.0, 0.0};
    auto norms = [&]() -> int {
        hipLaunchKernelGGL(k_jac_norms, dim3(NPARTS), dim3(256), 0, ctx->stream, Ap.p, M, part.p);
        hipLaunchKernelGGL(k_jac_norms_final, dim3(1), dim3(64), 0, ctx->stream, part.p, red.p);
        IPT_HIPCHK(hipGetLastError());
        IPT_HIPCHK(hipMemcpyAsync(h, red.p, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
        return IPT_CUDA_SUCCESS;
    };
    if (M > m) {
        // the padding values follow from the norm of the matrix on the device
        hipLaunchKernelGGL(k_jac_norms, dim3(NPARTS), dim3(256), 0, ctx->stream, Ap.p, M, part.p);
        hipLaunchKernelGGL(k_jac_norms_final, dim3(1), dim3(64), 0, ctx->stream, part.p, red.p);
        hipLaunchKernelGGL(k_jac_pad, dim3((M - m + 255) / 256), dim3(256), 0, ctx->stream, Ap.p, M, m, red.p);
        IPT_HIPCHK(hipGetLastError());
    }
    IPT_STCHK(norms());
    if (!std::isfinite(h[1])) {
        return IPT_CUDA_CUSOLVER_ERROR;
    }
    // converged when the off-diagonal Frobenius norm is below 1e-15 of the
    // matrix's, or below 1e-12 of it and no longer halving (the rounding
    // floor of an order-M matrix is about eps M of its norm)
    const double total = h[1];
    const double abs_thr2 = 1e-30 * total;
    int sweep = 0;
    bool converged = !(h[0] > abs_thr2);
    double prev = h[0];
    // one pair: its subproblem is the whole matrix, solved in one round
    const int sweeps_per_round = nb == 2 ? 30 : inner;
    const int apply_blocks = ntile_a + vtiles * npairs;
    for (sweep = 0; sweep < 40 && !converged; ++sweep) {
        for (int r = 0; r < nb - 1; ++r) {
            hipLaunchKernelGGL(k_jac_pair, dim3(npairs), dim3(PT), 0, ctx->stream, Ap.p, M, nb, r, Q.p, qflag.p,
                               sweeps_per_round, abs_thr2);
            hipLaunchKernelGGL(k_jac_apply, dim3(apply_blocks), dim3(AT), 0, ctx->stream, Ap.p, Vp.p, M, nb, r, Q.p,
                               qflag.p, ntile_a, m);
            IPT_HIPCHK(hipGetLastError());
        }
        IPT_STCHK(norms());
        converged = !(h[0] > abs_thr2) || (h[0] <= 1e-24 * total && h[0] >= 0.25 * prev);
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
    // ascending order (the padding values are above every real one), sorted
    // on the device; the results are ordered on the stream behind it
    hipLaunchKernelGGL(k_jac_diag, dim3((M + 255) / 256), dim3(256), 0, ctx->stream, Ap.p, M, dg.p);
    hipLaunchKernelGGL(k_jac_order, dim3((M + 255) / 256), dim3(256), 0, ctx->stream, dg.p, M, m, dperm.p);
    hipLaunchKernelGGL(k_jac_gather, dim3(std::min(4096, (m * m + 255) / 256)), dim3(256), 0, ctx->stream, Ap.p, Vp.p,
                       M, m, dperm.p, w, V, ldv);
    IPT_HIPCHK(hipGetLastError());
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
    IPT_STCHK(sym_eig_device(ctx, d_A, lda, m, d_w, d_V, ldv, nullptr));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return IPT_CUDA_SUCCESS;
}
