// Block-cluster update with the cluster width M (2..32) fixed at compile
// time, one launch per width present in a solve.
//
// The elimination is that of block_update_rows<P> in ipt_block_kernels.hip
// (row r owned by P = next_pow2(M) lanes, lane t holding equation t of
// (d_r I - h^T) x = d_r B_r - (HB)_r, implicit partial pivoting by a
// shuffle arg-max, the column-dominance shortcut); the loops run to M
// instead of P, the pivot rows are not scaled (each unknown is divided by
// its pivot once at the end, one FMA per eliminated entry), and when every
// column of a row's system is dominant the pivot broadcasts are DPP moves
// instead of LDS shuffles: the same pivots, results within a few units of
// rounding.  A width's kernel then holds only the pivot steps and
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

template <int M>
__device__ __forceinline__ void block_update_rows_m(int cid, int lblk, const BlockUpdateArgs &g)
{
    constexpr int P = M <= 2 ? 2 : M <= 4 ? 4 : M <= 8 ? 8 : M <= 16 ? 16 : 32;
    constexpr int GROUPS = 256 / P;
    const BlockCluster cl = g.clusters[cid];
    const long ld = g.ld;
    const int t = threadIdx.x % P;
    const int grp = threadIdx.x / P;
    const long row = (long)lblk * GROUPS + grp;
    const bool live = g.active[cid] != 0;
    double sdn = 0.0, sbn = 0.0;
    const bool emit = g.em.zhi != nullptr;
    const double escale = emit ? emit_scale(g.em, cid) : 1.0;
    double mstep = 0.0, manc = 0.0;
    if (emit && live && lblk == 0 && grp == 0 && t < M) {
        g.em.colscale[cl.off + t] = escale;
    }
    // h staged in LDS once per block (coalesced: column u of h is M
    // consecutive rows of Y), hs[t][u] = h(u, t), so that the equations of
    // every row of the block read it from there
    __shared__ double hs[M][M + 1];
    __shared__ double hsum[P];
    if (live) {
        const double *Yc = g.Y + (long)cl.off * ld;
        for (int e = threadIdx.x; e < M * M; e += 256) {
            const int u = e % M, tc = e / M;
            hs[tc][u] = Yc[cl.first + u + (long)tc * ld];
        }
    }
    __syncthreads();
    // off-diagonal absolute row sums of h (see block_update_rows)
    if (threadIdx.x < P) {
        double acc = 0.0;
        if (threadIdx.x < M && live) {
#pragma unroll
            for (int u = 0; u < M; ++u) {
                if (u != (int)threadIdx.x) {
                    acc += fabs(hs[u][threadIdx.x]);
                }
            }
        }
        hsum[threadIdx.x] = acc;
    }
    __syncthreads();

    if (row < g.n) {
        const double *Bc = g.B + (long)cl.off * ld;
        const double *Yc = g.Y + (long)cl.off * ld;
        double *Bnc = g.Bn + (long)cl.off * ld;
        if (!live) {
            if (t < M) {
                Bnc[row + (long)t * ld] = Bc[row + (long)t * ld];
            }
        } else if (row >= cl.first && row <= cl.last) {
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
        } else {
            const double dr = g.d[row];
            // Lanes t >= M (padding up to P) load the equation of lane M-1:
            // they are never the pivot lane, nothing reads their values, and
            // unpredicated loads keep the register count of a power-of-two
            // width.
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
#pragma unroll
                        for (int o = 1; o < P; o <<= 1) {
                            const double oc = __shfl_xor(cand, o, P);
                            const int oi = __shfl_xor(cidx, o, P);
                            if (oc > cand || (oc == cand && oi < cidx)) {
                                cand = oc;
                                cidx = oi;
                            }
                        }
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
    }
    if (emit) {
        emit_block_max(g.em, (long)cid * g.pstride + lblk, mstep, manc);
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
        const long slot = (long)cid * g.pstride + lblk;
        g.partials[2 * slot] = sa[0] + sa[1] + sa[2] + sa[3];
        g.partials[2 * slot + 1] = sb[0] + sb[1] + sb[2] + sb[3];
    }
}

// blockIdx.y = cluster list[blockIdx.y] (all of width M), blockIdx.x = row block.
template <int M>
__global__ __launch_bounds__(256) void k_block_update_m(BlockUpdateArgs g, const int *__restrict__ list)
{
    block_update_rows_m<M>(list[blockIdx.y], blockIdx.x, g);
}

// Launch the width-M update of `count` clusters listed at device `list`.
template <int M>
int launch_block_width(const int *list, int count, const BlockUpdateArgs &g, hipStream_t stream)
{
    constexpr int P = M <= 2 ? 2 : M <= 4 ? 4 : M <= 8 ? 8 : M <= 16 ? 16 : 32;
    constexpr int rows = 256 / P;
    const dim3 grid((unsigned)((g.n + rows - 1) / rows), (unsigned)count);
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
