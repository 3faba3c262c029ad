// Block-cluster update with the cluster width M (2..32) fixed at compile
// time, one launch per width present in a solve.
//
// The elimination is that of block_update_rows<P> in ipt_block_kernels.hip
// (row r owned by P = next_pow2(M) lanes, lane t holding equation t of
// (d_r I - h^T) x = d_r B_r - (HB)_r, implicit partial pivoting, the
// column-dominance shortcut); the loops run to M instead of P, the pivot
// rows are not scaled (each unknown is divided by its pivot once at the
// end, one FMA per eliminated entry), the pivot search is a DPP butterfly
// instead of a shuffle reduction, and when every column of a row's system
// is dominant the pivot broadcasts are DPP moves instead of LDS shuffles:
// the same pivots, results within a few units of rounding.  A width's
// kernel then holds only the pivot steps and
// columns it uses: one kernel for every width up to 32 with the loops run
// to P needs ~120 KB of instructions (twice the 64 KB instruction cache a
// CU pair shares) and the registers of its widest path in every block.
#pragma once

#include <cfloat>

#include "ipt_block.h"
#include "ipt_emit.h"

namespace ipt {

struct BlockUpdateArgs {
    const double *B;
    const double *Y;
    double *Bn;
    long ld;
    const double *d;
    int n;
    const BlockCluster *clusters;
    const int *active;
    double damping;
    double *partials;
    int pstride;
    int *flags;
    LowpEmit em;  // operand of the next low-precision product (em.zhi NULL: none)
};

// Broadcast of lane L of each P-lane group by DPP: row_newbcast (lane L of
// every 16-lane row) for P = 16 and, with a select on lane bit 3, for P = 8;
// quad_perm for P = 4 and 2; for P = 32, row_newbcast followed by gfx950's
// permlane16_swap across the two rows of the group.  No LDS traffic.
template <int P, int L>
__device__ __forceinline__ double group_bcast(double v)
{
    if constexpr (P == 16) {
        return __builtin_amdgcn_update_dpp(v, v, 0x150 + L, 0xf, 0xf, false);
    } else if constexpr (P == 8) {
        const double lo = __builtin_amdgcn_update_dpp(v, v, 0x150 + L, 0xf, 0xf, false);
        const double hi = __builtin_amdgcn_update_dpp(v, v, 0x150 + 8 + L, 0xf, 0xf, false);
        return (threadIdx.x & 8) ? hi : lo;
    } else if constexpr (P == 4) {
        return __builtin_amdgcn_update_dpp(v, v, L | (L << 2) | (L << 4) | (L << 6), 0xf, 0xf, false);
    } else if constexpr (P == 2) {
        return __builtin_amdgcn_update_dpp(v, v, L | (L << 2) | ((2 + L) << 4) | ((2 + L) << 6), 0xf, 0xf, false);
    } else {
        // P = 32: row_newbcast gives each 16-lane row its own lane L % 16;
        // permlane16_swap then hands the source row's value to the other
        // row of the pair (the swapped-in operand for an even source row,
        // the swapped-out one for an odd source row), per 32-bit half.
        const double b = __builtin_amdgcn_update_dpp(v, v, 0x150 + (L & 15), 0xf, 0xf, false);
        const unsigned long long u = (unsigned long long)__double_as_longlong(b);
        const unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
        const auto rl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        const auto rh = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        const unsigned olo = L < 16 ? rl[0] : rl[1];
        const unsigned ohi = L < 16 ? rh[0] : rh[1];
        return __longlong_as_double((long long)(((unsigned long long)ohi << 32) | olo));
    }
}

// The value of the lane selected by the DPP control CTRL (quad_perm,
// row_half_mirror, row_mirror), per 32-bit half for doubles.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xf, 0xf, false);
}

template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v)
{
    return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xf, 0xf, false);
}

// The value of lane i ^ 16 (the other 16-lane row of a 32-lane group), by
// gfx950's permlane16_swap: of the swap of v with itself, a lane of an even
// row takes the second result, a lane of an odd row the first.
__device__ __forceinline__ unsigned other_row_u32(unsigned v)
{
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (threadIdx.x & 16) ? r[0] : r[1];
}

__device__ __forceinline__ double other_row_f64(double v)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = other_row_u32((unsigned)u), hi = other_row_u32((unsigned)(u >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Arg-max over the P lanes of a group (largest cand, ties to the lowest
// index), every lane of the group left with the result: a butterfly of DPP
// exchanges (lane ^ 1, lane ^ 2 by quad_perm; the mirrored half row and the
// mirrored row pair up the quads and the halves; permlane16_swap the two
// rows of a 32-lane group).  The order is total, so the result is that of
// any pairing: the same pivot as a shuffle reduction, without LDS traffic.
template <int P>
__device__ __forceinline__ void group_argmax(double &cand, int &cidx)
{
    auto take = [&](double oc, int oi) {
        if (oc > cand || (oc == cand && oi < cidx)) {
            cand = oc;
            cidx = oi;
        }
    };
    take(dpp_f64<0xB1>(cand), dpp_i32<0xB1>(cidx));  // quad_perm [1,0,3,2]
    if constexpr (P >= 4) {
        take(dpp_f64<0x4E>(cand), dpp_i32<0x4E>(cidx));  // quad_perm [2,3,0,1]
    }
    if constexpr (P >= 8) {
        take(dpp_f64<0x141>(cand), dpp_i32<0x141>(cidx));  // row_half_mirror
    }
    if constexpr (P >= 16) {
        take(dpp_f64<0x140>(cand), dpp_i32<0x140>(cidx));  // row_mirror
    }
    if constexpr (P >= 32) {
        const double oc = other_row_f64(cand);
        const int oi = (int)other_row_u32((unsigned)cidx);
        take(oc, oi);
    }
}

// Pivot steps PC..M-1 of a system whose columns are diagonally dominant
// (pivot of step pc on lane pc, known at compile time).  Gauss-Jordan
// without row scaling: the pivot lane keeps its row, every other lane
// subtracts (a_pc / piv) times it; the unknown of the lane that pivoted at
// step pc is its right-hand side over its pivot at the end.
template <int M, int P, int PC>
__device__ __forceinline__ void dominant_steps(double (&a)[M + 1], int t, bool &bad, int &mypc, double &mypiv)
{
    if constexpr (PC < M) {
        const double piv = group_bcast<P, PC>(a[PC]);
        if (!(fabs(piv) > 64.0 * DBL_EPSILON) || !isfinite(piv)) {
            bad = true;
            return;
        }
        const bool me = t == PC;
        const double alpha = me ? 0.0 : -(a[PC] / piv);
#pragma unroll
        for (int c = PC + 1; c <= M; ++c) {
            a[c] = fma(alpha, group_bcast<P, PC>(a[c]), a[c]);
        }
        if (me) {
            mypc = PC;
            mypiv = piv;
        }
        dominant_steps<M, P, PC + 1>(a, t, bad, mypc, mypiv);
    }
}

// Elimination of one row (the P lanes of group `grp` of the calling wave;
// lane t holds equation t of (d_r I - h^T) x = d_r B_r - (HB)_r).  Adds the
// row's norm and operand-maximum contributions to the lane's accumulators.
template <int M>
__device__ __forceinline__ void eliminate_row(long row, int t, const BlockCluster &cl, const BlockUpdateArgs &g,
                                              const double (*hs)[M + 1], const double *hsum, double escale,
                                              double &sdn, double &sbn, double &mstep, double &manc)
{
    constexpr int P = width_lanes(M);
    const long ld = g.ld;
    const bool emit = g.em.zhi != nullptr;
    const double *Bc = g.B + (long)cl.off * ld;
    const double *Yc = g.Y + (long)cl.off * ld;
    double *Bnc = g.Bn + (long)cl.off * ld;
    if (row >= cl.first && row <= cl.last) {
        if (t < M) {
            const double nv = (row == cl.first + t) ? 1.0 : 0.0;
            const double ov = Bc[row + (long)t * ld];
            Bnc[row + (long)t * ld] = nv;
            sdn += (nv - ov) * (nv - ov);
            sbn += ov * ov;
            if (emit) {
                const double av = g.em.anc != nullptr ? g.em.anc[row + (long)(cl.off + t) * ld] : ov;
                emit_store(g.em, g.em.mode, row, cl.off + t, nv, av, escale);
                mstep = fmax(mstep, fabs(nv - ov));
                manc = fmax(manc, fabs(nv - av));
            }
        }
        return;
    }
    const double dr = g.d[row];
    // Lanes t >= M (padding up to P) load the equation of lane M-1: they
    // are never the pivot lane, nothing reads their values, and unpredicated
    // loads keep the register count of a power-of-two width.
    const int te = t < M ? t : M - 1;
    double a[M + 1];
#pragma unroll
    for (int u = 0; u < M; ++u) {
        // h(u, t) = Y[first+u, off+t]
        a[u] = (u == t ? dr : 0.0) - hs[te][u];
    }
    a[M] = dr * Bc[row + (long)te * ld] - Yc[row + (long)te * ld];
    int mypc = -1;
    double mypiv = 1.0;
    bool bad = false;
    const bool mine = t >= M || fabs(dr - hs[te][te]) > 2.0 * hsum[t];
    const unsigned long long ball = __ballot(mine);
    const int base = (threadIdx.x & 63) & ~(P - 1);
    const unsigned long long gmask = ((1ULL << P) - 1ULL) << base;
    const bool dominant = (ball & gmask) == gmask;
    if (dominant) {
        dominant_steps<M, P, 0>(a, t, bad, mypc, mypiv);
    } else {
        bool used = false;
#pragma unroll
        for (int pc = 0; pc < M; ++pc) {
            if (!bad) {
                double cand = (t < M && !used) ? fabs(a[pc]) : -1.0;
                int cidx = t;
                group_argmax<P>(cand, cidx);
                if (!(cand > 64.0 * DBL_EPSILON) || !isfinite(cand)) {
                    bad = true;
                } else {
                    const double piv = __shfl(a[pc], cidx, P);
                    const bool me = t == cidx;
                    const double alpha = me ? 0.0 : -(a[pc] / piv);
#pragma unroll
                    for (int c = pc + 1; c <= M; ++c) {
                        a[c] = fma(alpha, __shfl(a[c], cidx, P), a[c]);
                    }
                    if (me) {
                        used = true;
                        mypc = pc;
                        mypiv = piv;
                    }
                }
            }
        }
    }
    if (bad) {
        if (t == 0) {
            atomicOr(&g.flags[0], 1);
        }
    } else if (mypc >= 0) {
        double v = a[M] / mypiv;
        const double ov = Bc[row + (long)mypc * ld];
        if (g.damping < 1.0) {
            v = (1.0 - g.damping) * ov + g.damping * v;
        }
        if (!isfinite(v)) {
            atomicOr(&g.flags[0], 1);
            v = 0.0;
        }
        Bnc[row + (long)mypc * ld] = v;
        sdn += (v - ov) * (v - ov);
        sbn += ov * ov;
        if (emit) {
            const double av = g.em.anc != nullptr ? g.em.anc[row + (long)(cl.off + mypc) * ld] : ov;
            emit_store(g.em, g.em.mode, row, cl.off + mypc, v, av, escale);
            mstep = fmax(mstep, fabs(v - ov));
            manc = fmax(manc, fabs(v - av));
        }
    }
}

// Far row (kFarLanes = 4 lanes): with c0 the mean of diag(h) and
// nrm = ||h - c0 I||_1 (largest column sum), a row whose diagonal lies
// farther than 8 nrm from c0 has s = 1 / (d_r - c0) and rho = |s| nrm <= 1/8,
// so
//     x^T = b^T (d_r I - h)^-1 = s b^T sum_k (s (h - c0 I))^k,
// a series whose k-th term is at most rho^k of the first.  Horner's rule
// z <- b^T + s z (h - c0 I), K - 1 times, then x^T = s z, with K the wave's
// term count (rho^K <= 2^-54 for its largest rho: the truncation is below
// the rounding of the elimination).  Lane q of the row's four owns columns
// [qS, qS + S), S = ceil(M / 4); each step takes the other lanes' z by DPP
// quad broadcasts and reads h(i, j) = ht[i][j] for its columns from LDS.
// acc[t] += sum over the columns i of lane SRC of z_i h(i, j0 + t), z_i
// taken from lane SRC of the quad by DPP (quad_perm broadcast)
template <int M, int SRC>
__device__ __forceinline__ void series_terms(const double (&z)[(M + 3) / 4], const double (*ht)[4 * ((M + 3) / 4)],
                                             int j0, double (&acc)[(M + 3) / 4])
{
    constexpr int S = (M + 3) / 4;
    constexpr int U = M - SRC * S < S ? M - SRC * S : S;  // columns i of lane SRC
    if constexpr (U > 0) {
        // all of this lane's h values first (one wait for the LDS), then the
        // broadcasts and FMAs; columns past M hold zeros, so no lane-
        // dependent guard is needed
        double hv[U][S];
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int t = 0; t < S; ++t) {
                hv[u][t] = ht[SRC * S + u][j0 + t];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double zi = __builtin_amdgcn_update_dpp(z[u], z[u], SRC | (SRC << 2) | (SRC << 4) | (SRC << 6),
                                                          0xf, 0xf, false);
#pragma unroll
            for (int t = 0; t < S; ++t) {
                acc[t] = fma(zi, hv[u][t], acc[t]);
            }
        }
    }
}

template <int M>
__device__ __forceinline__ void series_row(long row, int q, const BlockCluster &cl, const BlockUpdateArgs &g,
                                           const double (*ht)[4 * ((M + 3) / 4)], double c0, double s, int K,
                                           double escale,
                                           double &sdn, double &sbn, double &mstep, double &manc)
{
    constexpr int S = (M + 3) / 4;
    const long ld = g.ld;
    const bool emit = g.em.zhi != nullptr;
    const double *Bc = g.B + (long)cl.off * ld;
    const double *Yc = g.Y + (long)cl.off * ld;
    double *Bnc = g.Bn + (long)cl.off * ld;
    const double dr = g.d[row];
    const int j0 = q * S;
    double b[S], z[S];
#pragma unroll
    for (int t = 0; t < S; ++t) {
        const int j = j0 + t;
        b[t] = j < M ? dr * Bc[row + (long)j * ld] - Yc[row + (long)j * ld] : 0.0;
        z[t] = b[t];
    }
#pragma unroll 1
    for (int k = 1; k < K; ++k) {
        // h is read from LDS on every pass: hoisted out of the loop its
        // values would not fit in registers
        __asm__ volatile("" ::: "memory");
        double acc[S];
#pragma unroll
        for (int t = 0; t < S; ++t) {
            acc[t] = -c0 * z[t];
        }
        series_terms<M, 0>(z, ht, j0, acc);
        series_terms<M, 1>(z, ht, j0, acc);
        series_terms<M, 2>(z, ht, j0, acc);
        series_terms<M, 3>(z, ht, j0, acc);
#pragma unroll
        for (int t = 0; t < S; ++t) {
            z[t] = fma(s, acc[t], b[t]);
        }
    }
    bool bad = false;
#pragma unroll
    for (int t = 0; t < S; ++t) {
        const int j = j0 + t;
        if (j < M) {
            const double ov = Bc[row + (long)j * ld];
            double v = s * z[t];
            if (g.damping < 1.0) {
                v = (1.0 - g.damping) * ov + g.damping * v;
            }
            if (!isfinite(v)) {
                bad = true;
                v = 0.0;
            }
            Bnc[row + (long)j * ld] = v;
            sdn += (v - ov) * (v - ov);
            sbn += ov * ov;
            if (emit) {
                const double av = g.em.anc != nullptr ? g.em.anc[row + (long)(cl.off + j) * ld] : ov;
                emit_store(g.em, g.em.mode, row, cl.off + j, v, av, escale);
                mstep = fmax(mstep, fabs(v - ov));
                manc = fmax(manc, fabs(v - av));
            }
        }
    }
    if (bad) {
        atomicOr(&g.flags[0], 1);
    }
}

// One block of a width-M cluster's update: bx < near_blocks_for(M) is a
// near block, the rest are far blocks of kFarRows rows.  A row r is near
// when |d_r - c0| <= 8 nrm or it is a row of the cluster (not a number:
// far); near and far blocks apply the same test to the same values, so
// every row is updated exactly once.
template <int M>
__device__ __forceinline__ void block_update_rows_m(int cid, int bx, const BlockUpdateArgs &g)
{
    constexpr int P = width_lanes(M);
    constexpr int GROUPS = 256 / P;
    constexpr int NB = near_blocks_for(M);
    const BlockCluster cl = g.clusters[cid];
    const long ld = g.ld;
    const int t = threadIdx.x % P;
    const int grp = threadIdx.x / P;
    const bool live = g.active[cid] != 0;
    double sdn = 0.0, sbn = 0.0;
    const bool emit = g.em.zhi != nullptr;
    const double escale = emit ? emit_scale(g.em, cid) : 1.0;
    double mstep = 0.0, manc = 0.0;
    if (emit && live && bx == 0 && threadIdx.x < M) {
        g.em.colscale[cl.off + threadIdx.x] = escale;
    }
    // h staged in LDS once per block (coalesced: column u of h is M
    // consecutive rows of Y), hs[t][u] = h(u, t)
    __shared__ double hs[M][M + 1];
    __shared__ double ht[M][4 * ((M + 3) / 4)];  // ht[u][t] = h(u, t) (zero past M), for the far rows
    __shared__ double hsum[P];
    __shared__ double cinfo[2];
    if (live) {
        const double *Yc = g.Y + (long)cl.off * ld;
        for (int e = threadIdx.x; e < M * M; e += 256) {
            const int u = e % M, tc = e / M;
            const double v = Yc[cl.first + u + (long)tc * ld];
            hs[tc][u] = v;
            ht[u][tc] = v;
        }
        for (int e = threadIdx.x; e < M * (4 * ((M + 3) / 4) - M); e += 256) {
            const int w = 4 * ((M + 3) / 4) - M;
            ht[e / w][M + e % w] = 0.0;
        }
    }
    __syncthreads();
    if (live && threadIdx.x < 64) {
        // off-diagonal absolute row sums of h (the elimination's dominance
        // test), the centre c0 and the bound 8 ||h - c0 I||_1 of the near
        // rows; lane sums in a fixed order, so every block gets the same
        const int u0 = threadIdx.x;
        double acc = 0.0, dg = 0.0;
        if (u0 < M) {
            dg = hs[u0][u0];
#pragma unroll
            for (int u = 0; u < M; ++u) {
                if (u != u0) {
                    acc += fabs(hs[u][u0]);
                }
            }
        }
        if (u0 < P) {
            hsum[u0] = acc;
        }
        double sdg = dg;
        for (int o = 1; o < 64; o <<= 1) {
            sdg += __shfl_xor(sdg, o, 64);
        }
        const double c0 = sdg / (double)M;
        double col = 0.0;
        if (u0 < M) {
#pragma unroll
            for (int i = 0; i < M; ++i) {
                col += fabs(hs[u0][i] - (i == u0 ? c0 : 0.0));
            }
        }
        for (int o = 1; o < 64; o <<= 1) {
            col = fmax(col, __shfl_xor(col, o, 64));
        }
        if (u0 == 0) {
            cinfo[0] = c0;
            cinfo[1] = 8.0 * col;
        }
    }
    __syncthreads();
    if constexpr (M < kSeriesWidth) {
        // narrow: every row eliminated, one row per lane group
        const long row = (long)bx * GROUPS + grp;
        if (row < g.n) {
            if (live) {
                eliminate_row<M>(row, t, cl, g, hs, hsum, escale, sdn, sbn, mstep, manc);
            } else if (t < M) {
                g.Bn[row + (long)(cl.off + t) * ld] = g.B[row + (long)(cl.off + t) * ld];
            }
        }
    }
    const double c0 = live ? cinfo[0] : 0.0, thr = live ? cinfo[1] : 0.0;
    // not a number (c0 or thr): no row but the cluster's is near, and the
    // far path flags the non-finite results
    auto is_near = [&](long r) { return (r >= cl.first && r <= cl.last) || fabs(g.d[r] - c0) <= thr; };
    if constexpr (M < kSeriesWidth) {
    } else if (bx < NB) {
        if (live) {
            // hull [lo, hi) of the near rows: the cluster's rows [first,
            // last] widened by the rows next to them with d in [c0 - thr,
            // c0 + thr], counted 256 at a time on each side (d ascending, so
            // those rows are a run ending at first - 1 and a run starting at
            // last + 1); no number: no rows counted
            long lo = cl.first, hi = cl.last + 1;
            for (;;) {
                const long r = lo - 1 - threadIdx.x;
                const int c = __syncthreads_count(r >= 0 && g.d[r] >= c0 - thr);
                lo -= c;
                if (c < 256) {
                    break;
                }
            }
            for (;;) {
                const long r = hi + threadIdx.x;
                const int c = __syncthreads_count(r < g.n && g.d[r] <= c0 + thr);
                hi += c;
                if (c < 256) {
                    break;
                }
            }
#pragma unroll 1
            for (long base = lo + (long)bx * GROUPS; base < hi; base += (long)NB * GROUPS) {
                const long row = base + grp;
                if (row < hi && row < g.n && is_near(row)) {
                    eliminate_row<M>(row, t, cl, g, hs, hsum, escale, sdn, sbn, mstep, manc);
                }
            }
        }
    } else {
        const long row = (long)(bx - NB) * kFarRows + threadIdx.x / kFarLanes;
        const int q = threadIdx.x % kFarLanes;
        const bool in = row < g.n;
        if (!live) {
            if (in) {
                const double *Bc = g.B + (long)cl.off * ld;
                double *Bnc = g.Bn + (long)cl.off * ld;
                for (int j = q; j < M; j += kFarLanes) {
                    Bnc[row + (long)j * ld] = Bc[row + (long)j * ld];
                }
            }
        } else {
            const bool far = in && !is_near(row);
            const double s = far ? 1.0 / (g.d[row] - c0) : 0.0;
            // the wave's term count from its largest ratio rho = |s| nrm
            double rho = far ? fabs(s) * 0.125 * thr : 0.0;
            for (int o = 1; o < 64; o <<= 1) {
                rho = fmax(rho, __shfl_xor(rho, o, 64));
            }
            int K = 1;
            if (rho > 0.0) {
                K = (int)ceil(54.0 / -log2(rho));
                K = K < 1 ? 1 : K > 20 ? 20 : K;
            }
            // the quad's lanes exchange values by DPP: all four take the
            // row's path together (far is the same for the four)
            if (far) {
                series_row<M>(row, q, cl, g, ht, c0, s, K, escale, sdn, sbn, mstep, manc);
            }
        }
    }
    if (emit) {
        emit_block_max(g.em, (long)cid * g.pstride + bx, mstep, manc);
    }
    __shared__ double sa[4], sb[4];
    for (int o = 32; o > 0; o >>= 1) {
        sdn += __shfl_down(sdn, o, 64);
        sbn += __shfl_down(sbn, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sa[threadIdx.x >> 6] = sdn;
        sb[threadIdx.x >> 6] = sbn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const long slot = (long)cid * g.pstride + bx;
        g.partials[2 * slot] = sa[0] + sa[1] + sa[2] + sa[3];
        g.partials[2 * slot + 1] = sb[0] + sb[1] + sb[2] + sb[3];
    }
}

// blockIdx.y = cluster list[blockIdx.y] (all of width M), blockIdx.x = near
// or far block.
template <int M>
__global__ __launch_bounds__(256) void k_block_update_m(BlockUpdateArgs g, const int *__restrict__ list)
{
    block_update_rows_m<M>(list[blockIdx.y], blockIdx.x, g);
}

// Launch the width-M update of `count` clusters listed at device `list`.
template <int M>
int launch_block_width(const int *list, int count, const BlockUpdateArgs &g, hipStream_t stream)
{
    const dim3 grid((unsigned)update_blocks_for(M, g.n), (unsigned)count);
    hipLaunchKernelGGL(k_block_update_m<M>, grid, dim3(256), 0, stream, g, list);
    return hipGetLastError() == hipSuccess ? IPT_CUDA_SUCCESS : IPT_CUDA_CUDA_ERROR;
}

// Widths 65..kMaxWideCluster (ipt_block_wide.hip): one workgroup per row,
// the augmented system in LDS; block_wide_rows() rows per workgroup.
constexpr int kMaxWideCluster = 142;
int block_wide_rows();
int launch_block_wide(int m, const int *list, int count, const BlockUpdateArgs &g, hipStream_t stream);

// Width 2..32 dispatch (ipt_block_kernels.hip); the instantiations live in
// ipt_block_w*.hip so that they compile in parallel.
int block_update_width(int m, const int *list, int count, const BlockUpdateArgs &g, hipStream_t stream);
// One launch for the clusters of widths lo..2lo-2 (lo = 3, 5, 9: lane
// classes 4, 8, 16), list ordered as the per-width lists (ipt_block_class.hip).
int block_update_class(int lo, const int *list, int count, const BlockUpdateArgs &g, hipStream_t stream);

#define IPT_BLOCK_WIDTH_INSTANTIATE(W) \
    template int launch_block_width<W>(const int *, int, const BlockUpdateArgs &, hipStream_t);
#define IPT_BLOCK_WIDTH_EXTERN(W) \
    extern template int launch_block_width<W>(const int *, int, const BlockUpdateArgs &, hipStream_t);

}  // namespace ipt
