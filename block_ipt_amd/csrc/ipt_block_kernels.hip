// Batched block-cluster (Riccati) IPT update and the FP64 Rayleigh-Ritz
// building blocks (tall-skinny A^T B, tall x small products, column norms,
// residual norms, unpermute).
//
// Block update (IPT/src/ipt_cuda.cu:2280-2384): for cluster rows
// [first,last] of width m with basis columns [off, off+m) and h =
// (HB)[first:last, off:off+m], every row r outside the cluster solves
//     (d_r I - h^T) x = d_r B_r - (HB)_r,   B'_r = x  (optional damping)
// and cluster rows are reset to the identity.  The reference gives each
// row a thread block with the pivot search done by one thread; here a row
// is owned by a group of P = next_pow2(m) lanes of one wavefront, lane t
// holding equation t in registers; partial pivoting is implicit (the pivot
// lane is picked by a shuffle arg-max among lanes not yet used) so no row
// swap moves data.  A pivot with |a| <= 64*DBL_EPSILON or a non-finite
// result flags the solve as singular, as in the reference.
#include <cfloat>

#include "ipt_block.h"
#include "ipt_block_width.h"
#include "ipt_emit.h"

namespace ipt {

static inline int blocks_for(long total, int threads, int cap)
{
    long b = (total + threads - 1) / threads;
    if (b < 1) {
        b = 1;
    }
    return (int)(b > cap ? cap : b);
}

static inline int lane_class(int m)
{
    int c = 0;
    while ((1 << c) < m) {
        ++c;
    }
    return c;
}

// ------------------------------------------------------------------ block update

// Rows [lblk * 256/P, (lblk + 1) * 256/P) of cluster cid.
template <int P>
__device__ __forceinline__ void block_update_rows(
    int cid, int lblk, const double *__restrict__ B, const double *__restrict__ Y, double *__restrict__ Bn,
    long ld, const double *__restrict__ d, int n, const BlockCluster *__restrict__ clusters,
    const int *__restrict__ active, double damping, double *__restrict__ partials, int pstride,
    int *__restrict__ flags, const LowpEmit &em)
{
    constexpr int GROUPS = 256 / P;
    const BlockCluster cl = clusters[cid];
    const int m = cl.m;
    const int t = threadIdx.x % P;
    const int grp = threadIdx.x / P;
    const long row = (long)lblk * GROUPS + grp;
    const bool live = active[cid] != 0;
    double sdn = 0.0, sbn = 0.0;
    const bool emit = em.zhi != nullptr;
    const double escale = emit ? emit_scale(em, cid) : 1.0;
    double mstep = 0.0, manc = 0.0;
    if (emit && live && lblk == 0 && grp == 0 && t < m) {
        em.colscale[cl.off + t] = escale;
    }
    // Off-diagonal absolute row sums of the cluster matrix h, shared by
    // every row of the block: a row whose system matrix (d_r I - h^T) is
    // column diagonally dominant by a factor of 2 keeps its pivots on the
    // diagonal through the whole elimination (dominance survives Gaussian
    // elimination, and the factor 2 absorbs rounding), so partial pivoting
    // would pick exactly those lanes: the arg-max search is skipped and the
    // arithmetic is the same.
    __shared__ double hsum[P];
    if (P > 1) {
        if (threadIdx.x < P) {
            double acc = 0.0;
            if (threadIdx.x < m && live) {
                const double *Yc = Y + (long)cl.off * ld;
                for (int u = 0; u < m; ++u) {
                    if (u != (int)threadIdx.x) {
                        acc += fabs(Yc[cl.first + threadIdx.x + (long)u * ld]);
                    }
                }
            }
            hsum[threadIdx.x] = acc;
        }
        __syncthreads();
    }

    if (row < n) {
        const double *Bc = B + (long)cl.off * ld;
        const double *Yc = Y + (long)cl.off * ld;
        double *Bnc = Bn + (long)cl.off * ld;
        if (!live) {
            if (t < m) {
                Bnc[row + (long)t * ld] = Bc[row + (long)t * ld];
            }
        } else if (row >= cl.first && row <= cl.last) {
            if (t < m) {
                const double nv = (row == cl.first + t) ? 1.0 : 0.0;
                const double ov = Bc[row + (long)t * ld];
                Bnc[row + (long)t * ld] = nv;
                sdn += (nv - ov) * (nv - ov);
                sbn += ov * ov;
                if (emit) {
                    const double av = em.anc != nullptr ? em.anc[row + (long)(cl.off + t) * ld] : ov;
                    emit_store(em, em.mode, row, cl.off + t, nv, av, escale);
                    mstep = fmax(mstep, fabs(nv - ov));
                    manc = fmax(manc, fabs(nv - av));
                }
            }
        } else {
            const double dr = d[row];
            double a[P + 1];
#pragma unroll
            for (int u = 0; u < P; ++u) {
                a[u] = 0.0;
                if (t < m && u < m) {
                    // h(u, t) = Y[first+u, off+t]
                    a[u] = (u == t ? dr : 0.0) - Yc[cl.first + u + (long)t * ld];
                }
            }
            a[P] = (t < m) ? dr * Bc[row + (long)t * ld] - Yc[row + (long)t * ld] : 0.0;
            bool used = false;
            int mypc = -1;
            bool bad = false;
            // lane t checks column t: |d_r - h_tt| > 2 sum_{u != t} |h_tu|
            bool dominant = true;
            if (P > 1) {
                const bool mine = t >= m || fabs(dr - Yc[cl.first + t + (long)t * ld]) > 2.0 * hsum[t];
                const unsigned long long ball = __ballot(mine);
                const int base = (threadIdx.x & 63) & ~(P - 1);
                const unsigned long long gmask = (P == 64) ? ~0ULL : (((1ULL << P) - 1ULL) << base);
                dominant = (ball & gmask) == gmask;
            }
#pragma unroll
            for (int pc = 0; pc < P; ++pc) {
                if (pc < m && !bad) {
                    double cand;
                    int cidx;
                    if (dominant) {
                        cidx = pc;
                        cand = fabs(__shfl(a[pc], pc, P));
                    } else {
                        cand = (t < m && !used) ? fabs(a[pc]) : -1.0;
                        cidx = t;
#pragma unroll
                        for (int o = 1; o < P; o <<= 1) {
                            const double oc = __shfl_xor(cand, o, P);
                            const int oi = __shfl_xor(cidx, o, P);
                            if (oc > cand || (oc == cand && oi < cidx)) {
                                cand = oc;
                                cidx = oi;
                            }
                        }
                    }
                    if (!(cand > 64.0 * DBL_EPSILON) || !isfinite(cand)) {
                        bad = true;
                    } else {
                        const double inv = 1.0 / __shfl(a[pc], cidx, P);
                        const double f = a[pc];
#pragma unroll
                        for (int c = pc; c <= P; ++c) {
                            // columns m..P-1 are zero padding (m is uniform per block)
                            if (c < m || c == P) {
                                const double pr = __shfl(a[c], cidx, P) * inv;
                                if (t == cidx) {
                                    a[c] = pr;
                                } else {
                                    a[c] -= f * pr;
                                }
                            }
                        }
                        if (t == cidx) {
                            used = true;
                            mypc = pc;
                        }
                    }
                }
            }
            if (bad) {
                if (t == 0) {
                    atomicOr(&flags[0], 1);
                }
            } else if (mypc >= 0) {
                double v = a[P];
                const double ov = Bc[row + (long)mypc * ld];
                if (damping < 1.0) {
                    v = (1.0 - damping) * ov + damping * v;
                }
                if (!isfinite(v)) {
                    atomicOr(&flags[0], 1);
                    v = 0.0;
                }
                Bnc[row + (long)mypc * ld] = v;
                sdn += (v - ov) * (v - ov);
                sbn += ov * ov;
                if (emit) {
                    const double av = em.anc != nullptr ? em.anc[row + (long)(cl.off + mypc) * ld] : ov;
                    emit_store(em, em.mode, row, cl.off + mypc, v, av, escale);
                    mstep = fmax(mstep, fabs(v - ov));
                    manc = fmax(manc, fabs(v - av));
                }
            }
        }
    }
    if (emit) {
        emit_block_max(em, (long)cid * pstride + lblk, mstep, manc);
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
        const long slot = (long)cid * pstride + lblk;
        partials[2 * slot] = sa[0] + sa[1] + sa[2] + sa[3];
        partials[2 * slot + 1] = sb[0] + sb[1] + sb[2] + sb[3];
    }
}

// One class: blockIdx.y = cluster of the class list, blockIdx.x = row block.
template <int P>
__global__ __launch_bounds__(256) void k_block_update(
    const double *__restrict__ B, const double *__restrict__ Y, double *__restrict__ Bn, long ld,
    const double *__restrict__ d, int n, const BlockCluster *__restrict__ clusters,
    const int *__restrict__ class_list, const int *__restrict__ active, double damping,
    double *__restrict__ partials, int pstride, int *__restrict__ flags, LowpEmit em)
{
    block_update_rows<P>(class_list[blockIdx.y], blockIdx.x, B, Y, Bn, ld, d, n, clusters, active, damping,
                         partials, pstride, flags, em);
}

// Widths 2..32: one kernel per width (ipt_block_width.h).
IPT_BLOCK_WIDTH_EXTERN(2)
IPT_BLOCK_WIDTH_EXTERN(3)
IPT_BLOCK_WIDTH_EXTERN(4)
IPT_BLOCK_WIDTH_EXTERN(5)
IPT_BLOCK_WIDTH_EXTERN(6)
IPT_BLOCK_WIDTH_EXTERN(7)
IPT_BLOCK_WIDTH_EXTERN(8)
IPT_BLOCK_WIDTH_EXTERN(9)
IPT_BLOCK_WIDTH_EXTERN(10)
IPT_BLOCK_WIDTH_EXTERN(11)
IPT_BLOCK_WIDTH_EXTERN(12)
IPT_BLOCK_WIDTH_EXTERN(13)
IPT_BLOCK_WIDTH_EXTERN(14)
IPT_BLOCK_WIDTH_EXTERN(15)
IPT_BLOCK_WIDTH_EXTERN(16)
IPT_BLOCK_WIDTH_EXTERN(17)
IPT_BLOCK_WIDTH_EXTERN(18)
IPT_BLOCK_WIDTH_EXTERN(19)
IPT_BLOCK_WIDTH_EXTERN(20)
IPT_BLOCK_WIDTH_EXTERN(21)
IPT_BLOCK_WIDTH_EXTERN(22)
IPT_BLOCK_WIDTH_EXTERN(23)
IPT_BLOCK_WIDTH_EXTERN(24)
IPT_BLOCK_WIDTH_EXTERN(25)
IPT_BLOCK_WIDTH_EXTERN(26)
IPT_BLOCK_WIDTH_EXTERN(27)
IPT_BLOCK_WIDTH_EXTERN(28)
IPT_BLOCK_WIDTH_EXTERN(29)
IPT_BLOCK_WIDTH_EXTERN(30)
IPT_BLOCK_WIDTH_EXTERN(31)
IPT_BLOCK_WIDTH_EXTERN(32)

int block_update_width(int m, const int *list, int count, const BlockUpdateArgs &g, hipStream_t stream)
{
    switch (m) {
    case 2: return launch_block_width<2>(list, count, g, stream);
    case 3: return launch_block_width<3>(list, count, g, stream);
    case 4: return launch_block_width<4>(list, count, g, stream);
    case 5: return launch_block_width<5>(list, count, g, stream);
    case 6: return launch_block_width<6>(list, count, g, stream);
    case 7: return launch_block_width<7>(list, count, g, stream);
    case 8: return launch_block_width<8>(list, count, g, stream);
    case 9: return launch_block_width<9>(list, count, g, stream);
    case 10: return launch_block_width<10>(list, count, g, stream);
    case 11: return launch_block_width<11>(list, count, g, stream);
    case 12: return launch_block_width<12>(list, count, g, stream);
    case 13: return launch_block_width<13>(list, count, g, stream);
    case 14: return launch_block_width<14>(list, count, g, stream);
    case 15: return launch_block_width<15>(list, count, g, stream);
    case 16: return launch_block_width<16>(list, count, g, stream);
    case 17: return launch_block_width<17>(list, count, g, stream);
    case 18: return launch_block_width<18>(list, count, g, stream);
    case 19: return launch_block_width<19>(list, count, g, stream);
    case 20: return launch_block_width<20>(list, count, g, stream);
    case 21: return launch_block_width<21>(list, count, g, stream);
    case 22: return launch_block_width<22>(list, count, g, stream);
    case 23: return launch_block_width<23>(list, count, g, stream);
    case 24: return launch_block_width<24>(list, count, g, stream);
    case 25: return launch_block_width<25>(list, count, g, stream);
    case 26: return launch_block_width<26>(list, count, g, stream);
    case 27: return launch_block_width<27>(list, count, g, stream);
    case 28: return launch_block_width<28>(list, count, g, stream);
    case 29: return launch_block_width<29>(list, count, g, stream);
    case 30: return launch_block_width<30>(list, count, g, stream);
    case 31: return launch_block_width<31>(list, count, g, stream);
    case 32: return launch_block_width<32>(list, count, g, stream);
    default: return IPT_CUDA_INVALID_ARGUMENT;
    }
}

// Width-1 clusters (most of the targets): the update is the scalar
//     x = (d_r B_r - (HB)_r) / (d_r - h),  h = (HB)[first, off],
// one thread per row and kClustersPerBlock clusters per block, so the
// registers (and occupancy) are those of a scalar update and a row's d_r is
// loaded once.  Same arithmetic as the P = 1 path of block_update_rows
// (pivot test, product with the reciprocal, damping, finite check).  The
// norm partials are per 64-row wavefront (nblocks = ceil(n / 64)).
constexpr int kClustersPerBlock = 4;

__global__ __launch_bounds__(256) void k_block_update_w1(
    const double *__restrict__ B, const double *__restrict__ Y, double *__restrict__ Bn, long ld,
    const double *__restrict__ d, int n, const BlockCluster *__restrict__ clusters,
    const int *__restrict__ w1_list, int nw1, const int *__restrict__ active, double damping,
    double *__restrict__ partials, int pstride, int *__restrict__ flags, LowpEmit em)
{
    const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool emit = em.zhi != nullptr;
    const int chunk = (int)(row >> 6);
    const bool in = row < n;
    const double dr = in ? d[row] : 0.0;
    int bad = 0;
    for (int j = 0; j < kClustersPerBlock; ++j) {
        const int q = blockIdx.y * kClustersPerBlock + j;
        if (q >= nw1) {
            break;
        }
        const int cid = w1_list[q];
        const BlockCluster cl = clusters[cid];
        const bool live = active[cid] != 0;
        const long col = (long)cl.off * ld;
        double sdn = 0.0, sbn = 0.0;
        double mstep = 0.0, manc = 0.0;
        const double escale = emit ? emit_scale(em, cid) : 1.0;
        if (emit && live && row == 0) {
            em.colscale[cl.off] = escale;
        }
        if (in) {
            const double ov = B[row + col];
            double nv = ov;
            bool wrote = false;
            if (!live) {
                Bn[row + col] = ov;
            } else if (row == cl.first) {
                Bn[row + col] = 1.0;
                sdn = (1.0 - ov) * (1.0 - ov);
                sbn = ov * ov;
                nv = 1.0;
                wrote = true;
            } else {
                const double a = dr - Y[cl.first + col];
                const double rhs = dr * ov - Y[row + col];
                if (!(fabs(a) > 64.0 * DBL_EPSILON) || !isfinite(a)) {
                    bad = 1;
                } else {
                    double v = rhs * (1.0 / a);
                    if (damping < 1.0) {
                        v = (1.0 - damping) * ov + damping * v;
                    }
                    if (!isfinite(v)) {
                        bad = 1;
                        v = 0.0;
                    }
                    Bn[row + col] = v;
                    sdn = (v - ov) * (v - ov);
                    sbn = ov * ov;
                    nv = v;
                    wrote = true;
                }
            }
            if (emit && wrote) {
                const double av = em.anc != nullptr ? em.anc[row + col] : ov;
                emit_store(em, em.mode, row, cl.off, nv, av, escale);
                mstep = fabs(nv - ov);
                manc = fabs(nv - av);
            }
        }
        // a wavefront entirely past the last row has no partial slot
        if (live && (long)chunk * 64 < n) {
            for (int o = 32; o > 0; o >>= 1) {
                sdn += __shfl_down(sdn, o, 64);
                sbn += __shfl_down(sbn, o, 64);
                if (emit) {
                    mstep = fmax(mstep, __shfl_down(mstep, o, 64));
                    manc = fmax(manc, __shfl_down(manc, o, 64));
                }
            }
            if ((threadIdx.x & 63) == 0) {
                const long slot = (long)cid * pstride + chunk;
                partials[2 * slot] = sdn;
                partials[2 * slot + 1] = sbn;
                if (emit) {
                    em.mpart[2 * slot] = mstep;
                    em.mpart[2 * slot + 1] = manc;
                }
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        bad |= __shfl_down(bad, o, 64);
    }
    if ((threadIdx.x & 63) == 0 && bad) {
        atomicOr(&flags[0], 1);
    }
}

// Per-cluster fixed-point residual fp = ||B'-B||_F / max(1, ||B||_F)
// (IPT/src/ipt_cuda.cu:2984-3003); clusters reaching fp <= tol are frozen.
// After a low-precision step (lowp_step = 1) a cluster reaching fp <= tol
// stops iterating at once but is marked tentative: the next exact step
// updates it once more with an exact product and then freezes it, so no
// cluster iterates on past its convergence waiting for an exact step.
// One block per cluster, fixed-order reduction of the update's partials.
// flags[1] |= 1 on a non-finite norm.  The last block to finish (ticket in
// flags[3]) also sets flags[1] |= 2 when a solve was singular (flags[0],
// then cleared) and flags[2] = clusters still active, and resets the
// ticket.
__global__ __launch_bounds__(256) void k_block_finalize(const double *__restrict__ partials, int pstride,
                                                        const BlockCluster *__restrict__ clusters,
                                                        int *__restrict__ active, double *__restrict__ fp_out,
                                                        int *__restrict__ iters, double tol, int count_steps,
                                                        int nclusters, int *__restrict__ flags,
                                                        int *__restrict__ tentative, int lowp_step,
                                                        double *__restrict__ fp_prev, int *__restrict__ growth,
                                                        int *__restrict__ stalled, int stall_steps,
                                                        const double *__restrict__ mpart, double *__restrict__ cmax,
                                                        double *__restrict__ status)
{
    const int c = blockIdx.x;
    __shared__ double sa[4], sb[4], sm[4], sn[4];
    __shared__ int wl[4];
    __shared__ int last;
    if (active[c]) {
        const int nb = clusters[c].nblocks;
        double dn = 0.0, bn = 0.0, ms = 0.0, ma = 0.0;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) {
            dn += partials[2 * ((long)c * pstride + b)];
            bn += partials[2 * ((long)c * pstride + b) + 1];
            if (mpart != nullptr) {
                ms = fmax(ms, mpart[2 * ((long)c * pstride + b)]);
                ma = fmax(ma, mpart[2 * ((long)c * pstride + b) + 1]);
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            dn += __shfl_down(dn, o, 64);
            bn += __shfl_down(bn, o, 64);
            ms = fmax(ms, __shfl_down(ms, o, 64));
            ma = fmax(ma, __shfl_down(ma, o, 64));
        }
        if ((threadIdx.x & 63) == 0) {
            sa[threadIdx.x >> 6] = dn;
            sb[threadIdx.x >> 6] = bn;
            sm[threadIdx.x >> 6] = ms;
            sn[threadIdx.x >> 6] = ma;
        }
        __syncthreads();
        if (threadIdx.x == 0 && mpart != nullptr) {
            // the update's operand maxima of this step scale the next one
            cmax[2 * c] = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
            cmax[2 * c + 1] = fmax(fmax(sn[0], sn[1]), fmax(sn[2], sn[3]));
        }
        if (threadIdx.x == 0) {
            dn = sqrt(sa[0] + sa[1] + sa[2] + sa[3]);
            bn = sqrt(sb[0] + sb[1] + sb[2] + sb[3]);
            const double fp = dn / fmax(1.0, bn);
            fp_out[c] = fp;
            iters[c] += count_steps;
            bool stall = false;
            if (stall_steps > 0) {
                // stall detection: a cluster whose fixed-point residual grew
                // over stall_steps consecutive steps (or is not finite) stops
                // iterating; the caller resets its basis
                const double prev = fp_prev[c];
                fp_prev[c] = fp;
                const bool finite = isfinite(dn) && isfinite(bn) && bn > 0.0;
                growth[c] = (finite && isfinite(prev) && fp > prev) ? growth[c] + 1 : 0;
                stall = !finite || growth[c] >= stall_steps;
            } else if (!isfinite(dn) || !isfinite(bn) || bn <= 0.0) {
                atomicOr(&flags[1], 1);
            }
            if (stall) {
                active[c] = 0;
                tentative[c] = 0;
                stalled[c] = 1;
            } else if (lowp_step) {
                if (fp <= tol) {
                    active[c] = 0;
                    tentative[c] = 1;
                }
            } else if (tentative[c] || fp <= tol) {
                active[c] = 0;
                tentative[c] = 0;
            }
        }
    }
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(&flags[3], 1) == nclusters - 1;
    }
    __syncthreads();
    if (last) {
        __threadfence();
        int live = 0;
        for (int q = threadIdx.x; q < nclusters; q += blockDim.x) {
            live += __atomic_load_n(&active[q], __ATOMIC_RELAXED) != 0;
        }
        for (int o = 32; o > 0; o >>= 1) {
            live += __shfl_down(live, o, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            wl[threadIdx.x >> 6] = live;
        }
        __syncthreads();
        // one status block for the host: status bits, active count, then
        // per cluster active flag and fixed-point residual
        for (int q = threadIdx.x; q < nclusters; q += blockDim.x) {
            status[2 + q] = __atomic_load_n(&active[q], __ATOMIC_RELAXED) != 0 ? 1.0 : 0.0;
            const unsigned long long fb =
                __atomic_load_n(reinterpret_cast<unsigned long long *>(&fp_out[q]), __ATOMIC_RELAXED);
            status[2 + nclusters + q] = __longlong_as_double((long long)fb);
        }
        if (threadIdx.x == 0) {
            if (__atomic_load_n(&flags[0], __ATOMIC_RELAXED)) {
                atomicOr(&flags[1], 2);
            }
            flags[0] = 0;
            flags[2] = wl[0] + wl[1] + wl[2] + wl[3];
            flags[3] = 0;
            status[0] = (double)__atomic_load_n(&flags[1], __ATOMIC_RELAXED);
            status[1] = (double)flags[2];
        }
    }
}

__global__ void k_block_reset(int *__restrict__ active, int *__restrict__ iters, double *__restrict__ fp,
                              int nclusters, int *__restrict__ flags, int *__restrict__ tentative,
                              double *__restrict__ fp_prev, int *__restrict__ growth, int *__restrict__ stalled)
{
    for (int c = threadIdx.x; c < nclusters; c += blockDim.x) {
        active[c] = 1;
        iters[c] = 0;
        fp[c] = NAN;
        tentative[c] = 0;
        fp_prev[c] = NAN;
        growth[c] = 0;
        stalled[c] = 0;
    }
    if (threadIdx.x == 0) {
        flags[0] = 0;
        flags[1] = 0;
        flags[2] = nclusters;
        flags[3] = 0;
    }
}

__global__ void k_block_identity(double *__restrict__ B, long ld, int n, const BlockCluster *__restrict__ clusters,
                                 int nclusters, int basis_cols)
{
    const long total = (long)n * basis_cols;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)gridDim.x * blockDim.x) {
        const long r = idx % n;
        const int col = (int)(idx / n);
        // cluster owning column col (clusters are few; linear scan)
        int c = 0;
        while (c + 1 < nclusters && clusters[c + 1].off <= col) {
            ++c;
        }
        const BlockCluster cl = clusters[c];
        B[r + (long)col * ld] = (r == cl.first + (col - cl.off)) ? 1.0 : 0.0;
    }
}

// B(:, columns of cluster c) = its identity block, for the clusters in list
__global__ void k_block_identity_some(double *__restrict__ B, long ld, int n,
                                      const BlockCluster *__restrict__ clusters, const int *__restrict__ list)
{
    const BlockCluster cl = clusters[list[blockIdx.y]];
    const long total = (long)n * cl.m;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        const long r = idx % n;
        const int j = (int)(idx / n);
        B[r + (long)(cl.off + j) * ld] = (r == cl.first + j) ? 1.0 : 0.0;
    }
}

int block_identity_clusters(Context *ctx, const BlockPlan &plan, const std::vector<int> &list, double *B, long ld,
                            int n)
{
    if (list.empty()) {
        return IPT_CUDA_SUCCESS;
    }
    DevBuf<int> dl;
    IPT_STCHK(dl.alloc(list.size()));
    IPT_HIPCHK(hipMemcpyAsync(dl.p, list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_block_identity_some, dim3(256, (unsigned)list.size()), dim3(256), 0, ctx->stream, B, ld, n,
                       plan.clusters.p, dl.p);
    IPT_HIPCHK(hipGetLastError());
    // the list must outlive the launch
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return IPT_CUDA_SUCCESS;
}

int block_plan_create(Context *ctx, const std::vector<std::pair<int, int>> &ranges, int n, BlockPlan *plan,
                      int skip_wider)
{
    plan->nclusters = (int)ranges.size();
    plan->host.clear();
    plan->wide_host.clear();
    plan->skipped.clear();
    std::vector<std::vector<int>> lists(7);
    int off = 0;
    plan->partial_stride = 1;
    for (int c = 0; c < plan->nclusters; ++c) {
        BlockCluster cl;
        cl.first = ranges[(size_t)c].first;
        cl.last = ranges[(size_t)c].second;
        cl.m = cl.last - cl.first + 1;
        cl.off = off;
        const int cls = lane_class(cl.m);
        if (skip_wider > 0 && cl.m > skip_wider) {
            // never iterated: its columns keep their identity block
            cl.nblocks = 1;
            cl.pad = 0;
            off += cl.m;
            plan->skipped.push_back(c);
            plan->host.push_back(cl);
            continue;
        }
        if (cl.m > kMaxWideCluster) {
            fprintf(stderr, "IPT block-cluster: a cluster of width %d exceeds the widest supported update (%d)\n",
                    cl.m, kMaxWideCluster);
            return IPT_CUDA_INVALID_ARGUMENT;
        }
        // width 1: partials per 64-row wavefront; 2..32: per near and far
        // block; wider: per row block
        const int rows = cl.m == 1 ? 64 : cl.m > 64 ? block_wide_rows() : 256 >> cls;
        cl.nblocks = cl.m >= 2 && cl.m <= 32 ? update_blocks_for(cl.m, n) : (n + rows - 1) / rows;
        cl.pad = 0;
        plan->partial_stride = std::max(plan->partial_stride, cl.nblocks);
        off += cl.m;
        if (cl.m > 64) {
            plan->wide_host.push_back(c);
        } else {
            lists[(size_t)cls].push_back(c);
        }
        plan->host.push_back(cl);
    }
    plan->basis_cols = off;
    std::vector<int> flat;
    for (int cls = 0; cls < 7; ++cls) {
        plan->class_offset[cls] = (int)flat.size();
        plan->class_count[cls] = (int)lists[(size_t)cls].size();
        flat.insert(flat.end(), lists[(size_t)cls].begin(), lists[(size_t)cls].end());
    }
    const int nc = std::max(plan->nclusters, 1);
    IPT_STCHK(plan->clusters.alloc((size_t)nc));
    IPT_STCHK(plan->class_lists.alloc(std::max<size_t>(flat.size(), 1)));
    IPT_STCHK(plan->active.alloc((size_t)nc));
    IPT_STCHK(plan->iters.alloc((size_t)nc));
    IPT_STCHK(plan->fp.alloc((size_t)nc));
    IPT_STCHK(plan->tentative.alloc((size_t)nc));
    IPT_STCHK(plan->fp_prev.alloc((size_t)nc));
    IPT_STCHK(plan->growth.alloc((size_t)nc));
    IPT_STCHK(plan->stalled.alloc((size_t)nc));
    IPT_STCHK(plan->flags.alloc(4));
    IPT_STCHK(plan->status.alloc((size_t)2 + 2 * (size_t)nc));
    // width-1 clusters: their own kernel
    plan->w1_count = (int)lists[0].size();
    IPT_STCHK(plan->w1_list.alloc(std::max<size_t>(lists[0].size(), 1)));
    if (!lists[0].empty()) {
        IPT_HIPCHK(hipMemcpyAsync(plan->w1_list.p, lists[0].data(), lists[0].size() * sizeof(int),
                                  hipMemcpyHostToDevice, ctx->stream));
    }
    IPT_STCHK(plan->wide_list.alloc(std::max<size_t>(plan->wide_host.size(), 1)));
    if (!plan->wide_host.empty()) {
        IPT_HIPCHK(hipMemcpyAsync(plan->wide_list.p, plan->wide_host.data(), plan->wide_host.size() * sizeof(int),
                                  hipMemcpyHostToDevice, ctx->stream));
    }
    // clusters of width 2..32 grouped by width (one launch per width)
    {
        std::vector<int> wl;
        plan->wide_count = 0;
        for (int m = 0; m <= 32; ++m) {
            plan->width_offset[m] = (int)wl.size();
            plan->width_count[m] = 0;
            if (m < 2) {
                continue;
            }
            for (int c = 0; c < plan->nclusters; ++c) {
                if (plan->host[(size_t)c].m == m) {
                    wl.push_back(c);
                    plan->width_count[m] += 1;
                }
            }
            plan->wide_count += plan->width_count[m];
        }
        IPT_STCHK(plan->width_lists.alloc(std::max<size_t>(wl.size(), 1)));
        if (!wl.empty()) {
            IPT_HIPCHK(hipMemcpyAsync(plan->width_lists.p, wl.data(), wl.size() * sizeof(int), hipMemcpyHostToDevice,
                                      ctx->stream));
        }
    }
    if (plan->nclusters > 0) {
        IPT_HIPCHK(hipMemcpyAsync(plan->clusters.p, plan->host.data(), plan->host.size() * sizeof(BlockCluster),
                                  hipMemcpyHostToDevice, ctx->stream));
        IPT_HIPCHK(hipMemcpyAsync(plan->class_lists.p, flat.data(), flat.size() * sizeof(int),
                                  hipMemcpyHostToDevice, ctx->stream));
    }
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return block_plan_reset(ctx, plan);
}

__global__ void k_block_mark_stalled(int *__restrict__ active, int *__restrict__ stalled, const int *__restrict__ list,
                                     int count)
{
    for (int q = threadIdx.x; q < count; q += blockDim.x) {
        active[list[q]] = 0;
        stalled[list[q]] = 1;
    }
}

int block_plan_mark_skipped(Context *ctx, BlockPlan *plan)
{
    if (plan->skipped.empty()) {
        return IPT_CUDA_SUCCESS;
    }
    DevBuf<int> dl;
    IPT_STCHK(dl.alloc(plan->skipped.size()));
    IPT_HIPCHK(hipMemcpyAsync(dl.p, plan->skipped.data(), plan->skipped.size() * sizeof(int), hipMemcpyHostToDevice,
                              ctx->stream));
    hipLaunchKernelGGL(k_block_mark_stalled, dim3(1), dim3(256), 0, ctx->stream, plan->active.p, plan->stalled.p,
                       dl.p, (int)plan->skipped.size());
    IPT_HIPCHK(hipGetLastError());
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return IPT_CUDA_SUCCESS;
}

int block_plan_reset(Context *ctx, BlockPlan *plan)
{
    hipLaunchKernelGGL(k_block_reset, dim3(1), dim3(256), 0, ctx->stream, plan->active.p, plan->iters.p,
                       plan->fp.p, plan->nclusters, plan->flags.p, plan->tentative.p, plan->fp_prev.p,
                       plan->growth.p, plan->stalled.p);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

__global__ void k_block_reactivate(int *__restrict__ active, const int *__restrict__ tentative, int nclusters)
{
    for (int c = threadIdx.x; c < nclusters; c += blockDim.x) {
        if (tentative[c]) {
            active[c] = 1;
        }
    }
}

int block_reactivate_tentative(Context *ctx, BlockPlan &plan)
{
    hipLaunchKernelGGL(k_block_reactivate, dim3(1), dim3(256), 0, ctx->stream, plan.active.p, plan.tentative.p,
                       plan.nclusters);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

int block_identity(Context *ctx, const BlockPlan &plan, double *B, long ld, int n)
{
    hipLaunchKernelGGL(k_block_identity, dim3(blocks_for((long)n * plan.basis_cols, 256, 16384)), dim3(256), 0,
                       ctx->stream, B, ld, n, plan.clusters.p, plan.nclusters, plan.basis_cols);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

int block_update(Context *ctx, BlockPlan &plan, const double *B, const double *Y, double *Bn, long ld,
                 const double *d, int n, double damping, double tol, int count_steps, int lowp_step,
                 const LowpEmit *emit)
{
    // norm partials (2 doubles per slot), then the operand maxima partials
    const size_t slots = (size_t)std::max(plan.nclusters, 1) * plan.partial_stride;
    double *partials = nullptr;
    IPT_STCHK(context_scratch(ctx, SLOT_REDUCE, slots * 4 * sizeof(double), (void **)&partials));
    LowpEmit em;
    if (emit != nullptr) {
        em = *emit;
        em.mpart = partials + 2 * slots;
    }
    // A launch whose clusters are all frozen (host view of the last exact
    // step) is skipped: their columns are the same in both iterate buffers.
    auto all_frozen = [&](int lo, int hi) {
        if (plan.frozen.empty()) {
            return false;
        }
        for (int c = 0; c < plan.nclusters; ++c) {
            const int m = plan.host[(size_t)c].m;
            if (m >= lo && m <= hi && !plan.frozen[(size_t)c]) {
                return false;
            }
        }
        return true;
    };
    const ProfileToken prof = profile_begin(ctx, PROF_BLOCK_UPDATE);
    if (plan.w1_count > 0 && !all_frozen(1, 1)) {
        dim3 grid((n + 255) / 256, (plan.w1_count + kClustersPerBlock - 1) / kClustersPerBlock);
        hipLaunchKernelGGL(k_block_update_w1, grid, dim3(256), 0, ctx->stream, B, Y, Bn, ld, d, n, plan.clusters.p,
                           plan.w1_list.p, plan.w1_count, plan.active.p, damping, partials, plan.partial_stride,
                           plan.flags.p, em);
        IPT_HIPCHK(hipGetLastError());
    }
    if (plan.wide_count > 0) {
        BlockUpdateArgs g{B, Y, Bn, ld, d, n, plan.clusters.p, plan.active.p, damping, partials,
                          plan.partial_stride, plan.flags.p};
        g.em = em;
        // widest first: their blocks take longest.  Widths 3..16 that share
        // a lane class go out as one launch when more than one is present.
        const bool merge = env_int("IPT_BLOCK_CLASS_LAUNCH", 1) != 0;
        for (int m = 32; m >= 2; --m) {
            const int lo = m >= 9 && m <= 16 ? 9 : m >= 5 && m <= 8 ? 5 : m >= 3 && m <= 4 ? 3 : 0;
            if (merge && lo != 0 && m == 2 * lo - 2) {
                int present = 0, count = 0;
                for (int q = lo; q <= m; ++q) {
                    present += plan.width_count[q] > 0;
                    count += plan.width_count[q];
                }
                if (present > 1) {
                    if (!all_frozen(lo, m)) {
                        IPT_STCHK(block_update_class(lo, plan.width_lists.p + plan.width_offset[lo], count, g,
                                                     ctx->stream));
                    }
                    m = lo;  // the loop's decrement moves on to lo - 1
                    continue;
                }
            }
            if (plan.width_count[m] > 0 && !all_frozen(m, m)) {
                IPT_STCHK(block_update_width(m, plan.width_lists.p + plan.width_offset[m], plan.width_count[m], g,
                                             ctx->stream));
            }
        }
    }
    if (plan.class_count[6] > 0 && !all_frozen(33, 64)) {
        // widths 33..64 keep their own launch (their registers would lower
        // the occupancy of every fused block)
        const int rows = 256 >> 6;
        dim3 grid((n + rows - 1) / rows, plan.class_count[6]);
        hipLaunchKernelGGL(k_block_update<64>, grid, dim3(256), 0, ctx->stream, B, Y, Bn, ld, d, n,
                           plan.clusters.p, plan.class_lists.p + plan.class_offset[6], plan.active.p, damping,
                           partials, plan.partial_stride, plan.flags.p, em);
        IPT_HIPCHK(hipGetLastError());
    }
    if (!plan.wide_host.empty()) {
        // the wide kernel writes no operand (the caller keeps emission off
        // when wide clusters are present)
        const BlockUpdateArgs g{B, Y, Bn, ld, d, n, plan.clusters.p, plan.active.p, damping, partials,
                                plan.partial_stride, plan.flags.p};
        for (size_t i = 0; i < plan.wide_host.size(); ++i) {
            const int c = plan.wide_host[i];
            if (!plan.frozen.empty() && plan.frozen[(size_t)c]) {
                continue;
            }
            IPT_STCHK(launch_block_wide(plan.host[(size_t)c].m, plan.wide_list.p + i, 1, g, ctx->stream));
        }
    }
    profile_end(ctx, prof);
    if (count_steps <= 0) {
        return IPT_CUDA_SUCCESS;
    }
    hipLaunchKernelGGL(k_block_finalize, dim3(plan.nclusters), dim3(256), 0, ctx->stream, partials,
                       plan.partial_stride, plan.clusters.p, plan.active.p, plan.fp.p, plan.iters.p, tol,
                       count_steps, plan.nclusters, plan.flags.p, plan.tentative.p, lowp_step, plan.fp_prev.p,
                       plan.growth.p, plan.stalled.p, plan.stall_steps, emit != nullptr ? em.mpart : nullptr,
                       emit != nullptr ? plan.cmax.p : nullptr, plan.status.p);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

// ------------------------------------------------------------------ A^T B (tall)

// C_partial[z] (m x nc) = A[rows of chunk z]^T B[rows of chunk z].  Tile of
// 32 x 32 outputs per block (16x16 threads, 2x2 each), rows staged through
// LDS 64 at a time.
__global__ __launch_bounds__(256) void k_gemm_tn_partial(const double *__restrict__ A, long lda,
                                                         const double *__restrict__ B, long ldb,
                                                         double *__restrict__ P, long nrows, int m, int nc,
                                                         long chunk)
{
    __shared__ double as[64][33];
    __shared__ double bs[64][33];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
    const long r0 = (long)blockIdx.z * chunk;
    long r1 = r0 + chunk;
    if (r1 > nrows) {
        r1 = nrows;
    }
    double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (long rb = r0; rb < r1; rb += 64) {
        for (int e = threadIdx.x; e < 64 * 32; e += 256) {
            const int rr = e & 63, cc = e >> 6;
            const long r = rb + rr;
            as[rr][cc] = (r < r1 && i0 + cc < m) ? A[r + (long)(i0 + cc) * lda] : 0.0;
            bs[rr][cc] = (r < r1 && j0 + cc < nc) ? B[r + (long)(j0 + cc) * ldb] : 0.0;
        }
        __syncthreads();
#pragma unroll 8
        for (int rr = 0; rr < 64; ++rr) {
            const double a0 = as[rr][ty], a1 = as[rr][ty + 16];
            const double b0 = bs[rr][tx], b1 = bs[rr][tx + 16];
            acc[0][0] += a0 * b0;
            acc[0][1] += a0 * b1;
            acc[1][0] += a1 * b0;
            acc[1][1] += a1 * b1;
        }
        __syncthreads();
    }
    double *Pz = P + (long)blockIdx.z * m * nc;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int i = i0 + ty + 16 * a, j = j0 + tx + 16 * b;
            if (i < m && j < nc) {
                Pz[i + (long)j * m] = acc[a][b];
            }
        }
    }
}

__global__ void k_sum_partials(const double *__restrict__ P, int splits, int m, int nc, double *__restrict__ C,
                               long ldc)
{
    const long total = (long)m * nc;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int z = 0; z < splits; ++z) {
            s += P[(long)z * total + idx];
        }
        const long i = idx % m, j = idx / m;
        C[i + j * ldc] = s;
    }
}

int gemm_tn_tall(Context *ctx, const double *A, long lda, const double *B, long ldb, double *C, long ldc,
                 long nrows, int m, int nc)
{
    if (m <= 0 || nc <= 0) {
        return IPT_CUDA_SUCCESS;
    }
    const int gx = (m + 31) / 32, gy = (nc + 31) / 32;
    long splits = 2048 / (gx * gy);
    if (splits < 1) {
        splits = 1;
    }
    long chunk = (nrows + splits - 1) / splits;
    if (chunk < 256) {
        chunk = 256;
    }
    chunk = (chunk + 63) / 64 * 64;
    splits = (nrows + chunk - 1) / chunk;
    if (splits < 1) {
        splits = 1;
    }
    double *P = nullptr;
    IPT_STCHK(context_scratch(ctx, SLOT_SMALL, (size_t)splits * m * nc * sizeof(double), (void **)&P));
    hipLaunchKernelGGL(k_gemm_tn_partial, dim3(gx, gy, (unsigned)splits), dim3(256), 0, ctx->stream, A, lda, B,
                       ldb, P, nrows, m, nc, chunk);
    IPT_HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_sum_partials, dim3(blocks_for((long)m * nc, 256, 1024)), dim3(256), 0, ctx->stream, P,
                       (int)splits, m, nc, C, ldc);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

// ------------------------------------------------------------------ A W (tall x small)

// 256 threads = 64 rows x 4 column groups; W staged in LDS 64 rows at a
// time; each thread accumulates up to 16 output columns of one row.
__global__ __launch_bounds__(256) void k_gemm_nn_small(const double *__restrict__ A, long lda,
                                                       const double *__restrict__ W, long ldw,
                                                       double *__restrict__ C, long ldc, long nrows, int m,
                                                       int nc, double alpha, double beta)
{
    __shared__ double ws[64][65];
    const int lr = threadIdx.x & 63;
    const int cg = threadIdx.x >> 6;  // 0..3
    const long r = (long)blockIdx.x * 64 + lr;
    const int c0 = blockIdx.y * 64;
    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        acc[q] = 0.0;
    }
    for (int i0 = 0; i0 < m; i0 += 64) {
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * 64; e += 256) {
            const int ii = e & 63, jj = e >> 6;
            ws[ii][jj] = (i0 + ii < m && c0 + jj < nc) ? W[(i0 + ii) + (long)(c0 + jj) * ldw] : 0.0;
        }
        __syncthreads();
        const int ilim = (m - i0) < 64 ? (m - i0) : 64;
        if (r < nrows) {
            for (int ii = 0; ii < ilim; ++ii) {
                const double a = A[r + (long)(i0 + ii) * lda];
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    acc[q] += a * ws[ii][cg + 4 * q];
                }
            }
        }
    }
    if (r < nrows) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int c = c0 + cg + 4 * q;
            if (c < nc) {
                double *dst = C + r + (long)c * ldc;
                *dst = (beta == 0.0) ? alpha * acc[q] : alpha * acc[q] + beta * (*dst);
            }
        }
    }
}

int gemm_nn_small(Context *ctx, const double *A, long lda, const double *W, long ldw, double *C, long ldc,
                  long nrows, int m, int nc, double alpha, double beta)
{
    if (nrows <= 0 || nc <= 0) {
        return IPT_CUDA_SUCCESS;
    }
    dim3 grid((unsigned)((nrows + 63) / 64), (unsigned)((nc + 63) / 64));
    hipLaunchKernelGGL(k_gemm_nn_small, grid, dim3(256), 0, ctx->stream, A, lda, W, ldw, C, ldc, nrows, m, nc,
                       alpha, beta);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

// ------------------------------------------------------------------ norms

// one block per column, fixed-order reduction
__global__ void k_col_sqnorm(const double *__restrict__ X, long ld, long nrows, double *__restrict__ sq)
{
    const int j = blockIdx.x;
    double s = 0.0;
    for (long r = threadIdx.x; r < nrows; r += blockDim.x) {
        const double x = X[r + (long)j * ld];
        s += x * x;
    }
    __shared__ double sh[4];
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_down(s, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[threadIdx.x >> 6] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        sq[j] = sh[0] + sh[1] + sh[2] + sh[3];
    }
}

int column_sqnorms(Context *ctx, const double *X, long ld, long nrows, int nc, double *sq)
{
    if (nc <= 0) {
        return IPT_CUDA_SUCCESS;
    }
    hipLaunchKernelGGL(k_col_sqnorm, dim3(nc), dim3(256), 0, ctx->stream, X, ld, nrows, sq);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

__global__ void k_scale_cols(double *__restrict__ X, double *__restrict__ HX, long ld, long nrows, int nc,
                             const double *__restrict__ sq)
{
    const long total = nrows * nc;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)gridDim.x * blockDim.x) {
        const long r = idx % nrows, j = idx / nrows;
        const double s = 1.0 / sqrt(sq[j]);
        X[r + j * ld] *= s;
        if (HX != nullptr) {
            HX[r + j * ld] *= s;
        }
    }
}

int scale_columns_inv_sqrt(Context *ctx, double *X, double *HX, long ld, long nrows, int nc, const double *sq)
{
    hipLaunchKernelGGL(k_scale_cols, dim3(blocks_for(nrows * nc, 256, 16384)), dim3(256), 0, ctx->stream, X, HX,
                       ld, nrows, nc, sq);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

// ------------------------------------------------------------------ Gram inverse square root

// dinv[i] = G_ii^-1/2; E = D^-1/2 G D^-1/2 - I (m x m, ld m); part[b] = sum
// of E^2 over the grid-stride share of block b (256 blocks).
__global__ __launch_bounds__(256) void k_gram_deviation(const double *__restrict__ G, int m,
                                                        double *__restrict__ E, double *__restrict__ part)
{
    double acc = 0.0;
    const long total = (long)m * m;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)gridDim.x * blockDim.x) {
        const int i = (int)(idx % m), j = (int)(idx / m);
        const double gi = G[i + (long)i * m], gj = G[j + (long)j * m];
        const double e = G[idx] / (sqrt(gi) * sqrt(gj)) - (i == j ? 1.0 : 0.0);
        E[idx] = e;
        acc += e * e;
    }
    __shared__ double sh[4];
    for (int o = 32; o > 0; o >>= 1) {
        acc += __shfl_down(acc, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[threadIdx.x >> 6] = acc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
    }
}

// out[0] = sum of the 256 partials (fixed order); out[1..m] = G_ii^-1/2
__global__ __launch_bounds__(64) void k_gram_deviation_final(const double *__restrict__ part,
                                                             const double *__restrict__ G, int m,
                                                             double *__restrict__ out)
{
    double s = 0.0;
    for (int b = threadIdx.x; b < 256; b += 64) {
        s += part[b];
    }
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, 64);
    }
    if (threadIdx.x == 0) {
        out[0] = s;
    }
    for (int i = threadIdx.x; i < m; i += 64) {
        out[1 + i] = 1.0 / sqrt(G[i + (long)i * m]);
    }
}

// T = a I + b W (m x m)
__global__ void k_shifted_scale(const double *__restrict__ W, int m, double a, double b, double *__restrict__ T)
{
    const long total = (long)m * m;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)gridDim.x * blockDim.x) {
        const int i = (int)(idx % m), j = (int)(idx / m);
        T[idx] = b * W[idx] + (i == j ? a : 0.0);
    }
}

// S[i + j m] = dinv[i] Z[i + j m]
__global__ void k_scale_rows(const double *__restrict__ Z, int m, const double *__restrict__ dinv,
                             double *__restrict__ S)
{
    const long total = (long)m * m;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)gridDim.x * blockDim.x) {
        S[idx] = dinv[idx % m] * Z[idx];
    }
}

int gram_inverse_sqrt(Context *ctx, const double *G, int m, double max_dev, double *S, double *dev_out,
                      int *applied)
{
    *applied = 0;
    *dev_out = NAN;
    if (m <= 0) {
        return IPT_CUDA_SUCCESS;
    }
    const size_t mm = (size_t)m * m;
    DevBuf<double> E, Y, Z, W, T, aux;
    IPT_STCHK(E.alloc(mm));
    IPT_STCHK(aux.alloc(256 + 1 + (size_t)m));
    double *part = aux.p, *out = aux.p + 256;
    const int blocks = blocks_for((long)mm, 256, 4096);
    hipLaunchKernelGGL(k_gram_deviation, dim3(256), dim3(256), 0, ctx->stream, G, m, E.p, part);
    hipLaunchKernelGGL(k_gram_deviation_final, dim3(1), dim3(64), 0, ctx->stream, part, G, m, out);
    IPT_HIPCHK(hipGetLastError());
    double fro2 = 0.0;
    IPT_HIPCHK(hipMemcpyAsync(&fro2, out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    const double dev = std::sqrt(fro2);
    *dev_out = dev;
    if (!(dev <= max_dev) || !(max_dev < 1.0)) {
        return IPT_CUDA_SUCCESS;
    }
    const double *dinv = out + 1;
    IPT_STCHK(Z.alloc(mm));
    // ||E||_2 <= dev.  Below 1e-8 the series (I + E)^-1/2 = I - E/2 + 3E^2/8
    // - ... is exact to rounding after its first-order term.  Otherwise the
    // coupled Newton-Schulz iteration Y <- Y T, Z <- T Z, T = (3I - Z Y)/2
    // from Y = I + E, Z = I (Higham, Functions of Matrices, 2008, 6.3)
    // converges to Z = (I + E)^-1/2; an eigenvalue 1 - delta of Z Y maps to
    // 1 - (3/4) delta^2 - (1/4) delta^3, so the step count follows from dev.
    if (dev <= 1e-8) {
        hipLaunchKernelGGL(k_shifted_scale, dim3(blocks), dim3(256), 0, ctx->stream, E.p, m, 1.0, -0.5, Z.p);
        IPT_HIPCHK(hipGetLastError());
    } else {
        IPT_STCHK(Y.alloc(mm));
        IPT_STCHK(W.alloc(mm));
        IPT_STCHK(T.alloc(mm));
        hipLaunchKernelGGL(k_shifted_scale, dim3(blocks), dim3(256), 0, ctx->stream, E.p, m, 1.0, 1.0, Y.p);
        hipLaunchKernelGGL(k_shifted_scale, dim3(blocks), dim3(256), 0, ctx->stream, E.p, m, 1.0, 0.0, Z.p);
        IPT_HIPCHK(hipGetLastError());
        double delta = dev;
        while (delta > 1e-17) {
            IPT_STCHK(gemm_f64(ctx, Z.p, m, Y.p, m, W.p, m, m, m));
            hipLaunchKernelGGL(k_shifted_scale, dim3(blocks), dim3(256), 0, ctx->stream, W.p, m, 1.5, -0.5, T.p);
            IPT_HIPCHK(hipGetLastError());
            IPT_STCHK(gemm_f64(ctx, Y.p, m, T.p, m, W.p, m, m, m));
            std::swap(Y.p, W.p);
            IPT_STCHK(gemm_f64(ctx, T.p, m, Z.p, m, W.p, m, m, m));
            std::swap(Z.p, W.p);
            delta = 0.75 * delta * delta + 0.25 * delta * delta * delta;
        }
    }
    hipLaunchKernelGGL(k_scale_rows, dim3(blocks), dim3(256), 0, ctx->stream, Z.p, m, dinv, S);
    IPT_HIPCHK(hipGetLastError());
    *applied = 1;
    return IPT_CUDA_SUCCESS;
}

// IPT/src/ipt_cuda.cu:2399-2447, 2684-2697: R = HX - X theta and the scaled
// residual norm per column.  One block per column.
__global__ void k_residual_norms(const double *__restrict__ X, const double *__restrict__ HX, long ld, long nrows,
                                 const double *__restrict__ theta, double scale, double *__restrict__ res,
                                 double *__restrict__ R)
{
    const int j = blockIdx.x;
    const double th = theta[j];
    double rs = 0.0, xs = 0.0;
    for (long r = threadIdx.x; r < nrows; r += blockDim.x) {
        const double x = X[r + (long)j * ld];
        const double rr = HX[r + (long)j * ld] - th * x;
        if (R != nullptr) {
            R[r + (long)j * ld] = rr;
        }
        rs += rr * rr;
        xs += x * x;
    }
    __shared__ double sa[4], sb[4];
    for (int o = 32; o > 0; o >>= 1) {
        rs += __shfl_down(rs, o, 64);
        xs += __shfl_down(xs, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sa[threadIdx.x >> 6] = rs;
        sb[threadIdx.x >> 6] = xs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double rn = sqrt(sa[0] + sa[1] + sa[2] + sa[3]);
        const double xn = sqrt(sb[0] + sb[1] + sb[2] + sb[3]);
        res[j] = rn / (scale * fmax(1.0, xn));
    }
}

int residual_norms(Context *ctx, const double *X, const double *HX, long ld, long nrows, int nc,
                   const double *theta, double normH, double *res, double *R)
{
    if (nc <= 0) {
        return IPT_CUDA_SUCCESS;
    }
    hipLaunchKernelGGL(k_residual_norms, dim3(nc), dim3(256), 0, ctx->stream, X, HX, ld, nrows, theta,
                       std::max(1.0, normH), res, R);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

__global__ void k_unpermute(const double *__restrict__ in, double *__restrict__ out, long ld, long nrows, int nc,
                            const int *__restrict__ perm)
{
    const long total = nrows * nc;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)gridDim.x * blockDim.x) {
        const long s = idx % nrows, j = idx / nrows;
        out[(long)perm[s] + j * ld] = in[s + j * ld];
    }
}

int unpermute_rows(Context *ctx, const double *in, double *out, long ld, long nrows, int nc, const int *perm)
{
    hipLaunchKernelGGL(k_unpermute, dim3(blocks_for(nrows * nc, 256, 16384)), dim3(256), 0, ctx->stream, in, out,
                       ld, nrows, nc, perm);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

}  // namespace ipt

namespace ipt {

// ------------------------------------------------------------------ dense prep

// One pass over tile pairs (bi <= bj) of a dense column-major H:
//  * symmetry: |H_rc - H_cr| <= 1e-12 max(1,|H_rc|,|H_cr|) (IPT/src/ipt_preparation.cu:860-883)
//  * per-column partial abs sums over each 64-row block (for ||H||_inf of
//    a symmetric matrix) written to colsum_part[rowblock * n + col]
//  * per-column partial max |offdiag| over each 64-row block, colmax_part
__global__ __launch_bounds__(256) void k_dense_scan(const double *__restrict__ H, long ld, int n, int nb,
                                                    double *__restrict__ colsum_part,
                                                    double *__restrict__ colmax_part, int *__restrict__ asym)
{
    // ta = T = H[bi rows, bj cols], tb = T' = H[bj rows, bi cols], both
    // [row][col] with a padded stride (conflict-free row and column walks)
    __shared__ double ta[64][65];
    __shared__ double tb[64][65];
    __shared__ double rs[2][4][64], rm[2][4][64];
    // blockIdx.x -> (bi, bj), bi <= bj, row-major over the upper triangle of
    // tile pairs: row bi starts at bi*nb - bi*(bi-1)/2
    const long idx = blockIdx.x;
    const double b2 = 2.0 * nb + 1.0;
    int bi = (int)((b2 - sqrt(b2 * b2 - 8.0 * (double)idx)) * 0.5);
    bi = max(0, min(bi, nb - 1));
    while (bi > 0 && (long)bi * nb - (long)bi * (bi - 1) / 2 > idx) {
        --bi;
    }
    while ((long)(bi + 1) * nb - (long)(bi + 1) * bi / 2 <= idx) {
        ++bi;
    }
    const int bj = bi + (int)(idx - ((long)bi * nb - (long)bi * (bi - 1) / 2));
    const bool diag = bi == bj;
    const long R0 = (long)bi * 64, C0 = (long)bj * 64;
    // loads: lane = row of the tile, wavefront w = columns w, w+4, ..., so
    // every load instruction reads 512 contiguous bytes of one column; both
    // tiles' 32 loads per thread are issued before the first is used
    const int lr = threadIdx.x & 63, w = threadIdx.x >> 6;
    double tv[16], uv[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const long gr = R0 + lr, gc = C0 + w + 4 * j;
        tv[j] = (gr < n && gc < n) ? H[gr + gc * ld] : 0.0;
    }
    if (!diag) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const long gr = C0 + lr, gc = R0 + w + 4 * j;
            uv[j] = (gr < n && gc < n) ? H[gr + gc * ld] : 0.0;
        }
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        ta[lr][w + 4 * j] = tv[j];
        if (!diag) {
            tb[lr][w + 4 * j] = uv[j];
        }
    }
    __syncthreads();
    // thread (c = lane, rows 16w .. 16w+15): column c of T (global column
    // C0 + c) against row c of T' (T' transposed), and column c of T'
    // (global column R0 + c)
    const int c = lr;
    int bad = 0;
    double sa = 0.0, ma = 0.0, sb = 0.0, mb = 0.0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int r = 16 * w + i;
        const double v = ta[r][c];
        const double u = diag ? ta[c][r] : tb[c][r];  // T(r, c) pairs with H(C0 + c, R0 + r)
        const double scale = fmax(1.0, fmax(fabs(v), fabs(u)));
        if (fabs(v - u) > 1e-12 * scale) {
            bad = 1;
        }
        sa += fabs(v);
        if (!(diag && r == c)) {
            ma = fmax(ma, fabs(v));
        }
        if (!diag) {
            // rows of block bj never meet the diagonal of columns R0 + c
            const double x = tb[r][c];
            sb += fabs(x);
            mb = fmax(mb, fabs(x));
        }
    }
    rs[0][w][c] = sa;
    rm[0][w][c] = ma;
    rs[1][w][c] = sb;
    rm[1][w][c] = mb;
    __syncthreads();
    if (threadIdx.x < 128) {
        const int h = threadIdx.x >> 6;  // 0: columns of T (C0 + c), 1: of T' (R0 + c)
        if (h == 0 || !diag) {
            const double s = ((rs[h][0][c] + rs[h][1][c]) + rs[h][2][c]) + rs[h][3][c];
            const double m = fmax(fmax(rm[h][0][c], rm[h][1][c]), fmax(rm[h][2][c], rm[h][3][c]));
            const long gc = (h == 0 ? C0 : R0) + c;
            const long slot = h == 0 ? bi : bj;
            if (gc < n) {
                colsum_part[slot * n + gc] = s;
                colmax_part[slot * n + gc] = m;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        bad |= __shfl_xor(bad, o, 64);
    }
    if ((threadIdx.x & 63) == 0 && bad) {
        atomicOr(asym, 1);
    }
}

// colsum[c] / colmax[c]: fixed-order reductions over the nb row-block
// partials: 32 columns per block, 8 threads per column each summing every
// 8th partial, the 8 sums then added in a fixed order (enough blocks to
// keep HBM busy: the partials are 2 x nb x n doubles)
__global__ __launch_bounds__(256) void k_colsum_reduce(const double *__restrict__ part,
                                                       const double *__restrict__ mpart, int nb, int n,
                                                       double *__restrict__ colsum, double *__restrict__ colmax)
{
    __shared__ double ss[8][33], sm[8][33];
    const int lc = threadIdx.x & 31, j = threadIdx.x >> 5;
    const long c = (long)blockIdx.x * 32 + lc;
    double s = 0.0, m = 0.0;
    if (c < n) {
        for (int b = j; b < nb; b += 8) {
            s += part[(long)b * n + c];
            m = fmax(m, mpart[(long)b * n + c]);
        }
    }
    ss[j][lc] = s;
    sm[j][lc] = m;
    __syncthreads();
    if (j == 0 && c < n) {
        double t = ss[0][lc], u = sm[0][lc];
        for (int q = 1; q < 8; ++q) {
            t += ss[q][lc];
            u = fmax(u, sm[q][lc]);
        }
        colsum[c] = t;
        colmax[c] = u;
    }
}

int dense_scan(Context *ctx, const double *H, long ld, int n, double *colsum, double *colmax, int *asym)
{
    const int nb = (n + 63) / 64;
    // 2 nb n doubles (n^2 / 4 bytes: 1 GiB at n = 65536), needed only here:
    // a pooled buffer, returned at exit (and trimmed under memory pressure)
    // instead of a scratch slot held for the context's lifetime
    DevBuf<double> partbuf;
    IPT_STCHK(partbuf.alloc((size_t)2 * nb * n));
    double *part = partbuf.p;
    IPT_HIPCHK(hipMemsetAsync(asym, 0, sizeof(int), ctx->stream));
    const long pairs = (long)nb * (nb + 1) / 2;
    hipLaunchKernelGGL(k_dense_scan, dim3((unsigned)pairs), dim3(256), 0, ctx->stream, H, ld, n, nb, part,
                       part + (size_t)nb * n, asym);
    IPT_HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_colsum_reduce, dim3((unsigned)((n + 31) / 32)), dim3(256), 0, ctx->stream, part,
                       part + (size_t)nb * n, nb, n, colsum, colmax);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

// out[r + c*ld] = H[perm[r] + perm[c]*ldh]  (P H P^T)
__global__ void k_dense_permute(const double *__restrict__ H, long ldh, double *__restrict__ out, long ld, int n,
                                const int *__restrict__ perm)
{
    const long total = (long)n * n;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)gridDim.x * blockDim.x) {
        const long r = idx % n, c = idx / n;
        out[r + c * ld] = H[(long)perm[r] + (long)perm[c] * ldh];
    }
}

int dense_permute(Context *ctx, const double *H, long ldh, double *out, long ld, int n, const int *perm)
{
    hipLaunchKernelGGL(k_dense_permute, dim3(blocks_for((long)n * n, 256, 65536)), dim3(256), 0, ctx->stream, H,
                       ldh, out, ld, n, perm);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

// vals[2i] = H[rows[i] + cols[i]*ld], vals[2i+1] = H[cols[i] + rows[i]*ld]
__global__ void k_gather_pairs(const double *__restrict__ H, long ld, const int *__restrict__ rc, int count,
                               double *__restrict__ vals)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const long r = rc[2 * i], c = rc[2 * i + 1];
        vals[2 * i] = H[r + c * ld];
        vals[2 * i + 1] = H[c + r * ld];
    }
}

int gather_pairs(Context *ctx, const double *H, long ld, const int *d_rc, int count, double *d_vals)
{
    if (count <= 0) {
        return IPT_CUDA_SUCCESS;
    }
    hipLaunchKernelGGL(k_gather_pairs, dim3(blocks_for(count, 256, 1024)), dim3(256), 0, ctx->stream, H, ld, d_rc,
                       count, d_vals);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

// Davidson correction t_i = r_i / clip(D_i - theta), IPT/src/ipt_cuda.cu:2449-2465,
// for several target pairs at once: T(:, q) from R(:, idx[q]).
__global__ void k_diag_correction(const double *__restrict__ R, long ldr, const double *__restrict__ d,
                                  const double *__restrict__ theta, const int *__restrict__ idx, int nt,
                                  double clip, double *__restrict__ T, long ldt, long nrows)
{
    const long total = nrows * nt;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const long r = e % nrows;
        const int q = (int)(e / nrows);
        const int j = idx[q];
        double den = d[r] - theta[j];
        if (fabs(den) < clip) {
            den = den < 0.0 ? -clip : clip;
        }
        T[r + (long)q * ldt] = R[r + (long)j * ldr] / den;
    }
}

int diag_correction(Context *ctx, const double *R, long ldr, const double *d, const double *theta,
                    const int *d_idx, int nt, double clip, double *T, long ldt, long nrows)
{
    if (nt <= 0) {
        return IPT_CUDA_SUCCESS;
    }
    hipLaunchKernelGGL(k_diag_correction, dim3(blocks_for(nrows * nt, 256, 16384)), dim3(256), 0, ctx->stream, R,
                       ldr, d, theta, d_idx, nt, clip, T, ldt, nrows);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

}  // namespace ipt
