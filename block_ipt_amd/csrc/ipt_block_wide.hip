// Block-cluster update for cluster widths 65..kMaxWideCluster: one
// workgroup per row with the augmented system (d_r I - h^T | d_r B_r -
// (HB)_r) in LDS and Gauss-Jordan elimination with partial pivoting by row
// exchange, as the reference's kernel does it for every width
// (IPT/src/ipt_cuda.cu:2280-2384; pivot rule and singular threshold
// 64 DBL_EPSILON as there, and as the oracle's orc_gauss_jordan).  Reached
// only when IPT_MAX_LIFT_BLOCK_SIZE is raised above 64.  A 142-wide system
// fills the 160 KiB of LDS one workgroup can hold on gfx950.
#include <cfloat>

#include "ipt_block_width.h"

namespace ipt {

namespace {

constexpr int kWideRows = 4;  // rows per workgroup, one after the other

__global__ __launch_bounds__(256) void k_block_update_wide(BlockUpdateArgs g, const int *__restrict__ list)
{
    extern __shared__ double aug[];  // m x (m + 1), equation e at aug[e * (m + 1)]
    __shared__ double rv[4];
    __shared__ int ri[4];
    __shared__ int s_piv;
    __shared__ double sa[4], sb[4];
    const int cid = list[blockIdx.y];
    const BlockCluster cl = g.clusters[cid];
    const int m = cl.m, w = m + 1;
    const long ld = g.ld;
    const bool live = g.active[cid] != 0;
    const double *Bc = g.B + (long)cl.off * ld;
    const double *Yc = g.Y + (long)cl.off * ld;
    double *Bnc = g.Bn + (long)cl.off * ld;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    static_assert(kMaxWideCluster + 1 <= 3 * 64, "three pivot-row entries per lane");
    double sdn = 0.0, sbn = 0.0;
    for (int q = 0; q < kWideRows; ++q) {
        const long row = (long)blockIdx.x * kWideRows + q;
        if (row >= g.n) {
            break;
        }
        if (!live) {
            for (int e = tid; e < m; e += 256) {
                Bnc[row + (long)e * ld] = Bc[row + (long)e * ld];
            }
            continue;
        }
        if (row >= cl.first && row <= cl.last) {
            for (int e = tid; e < m; e += 256) {
                const double nv = (row == cl.first + e) ? 1.0 : 0.0;
                const double ov = Bc[row + (long)e * ld];
                Bnc[row + (long)e * ld] = nv;
                sdn += (nv - ov) * (nv - ov);
                sbn += ov * ov;
            }
            continue;
        }
        const double dr = g.d[row];
        // column u of h is m consecutive rows of Y: the 64 lanes of a
        // wavefront read consecutive u of one equation e (wavefront ty takes
        // every fourth equation)
        const int tx = tid & 63, ty = tid >> 6;
        for (int e = ty; e < m; e += 4) {
            const double *Ye = Yc + (long)e * ld;
            for (int u = tx; u < m; u += 64) {
                aug[e * w + u] = (e == u ? dr : 0.0) - Ye[cl.first + u];
            }
            if (tx == 0) {
                aug[e * w + m] = dr * Bc[row + (long)e * ld] - Ye[row];
            }
        }
        __syncthreads();
        bool bad = false;
        for (int pc = 0; pc < m; ++pc) {
            // arg-max of |aug[e][pc]|, e >= pc, lowest index on ties
            double best = -1.0;
            int bi = m;
            for (int e = pc + tid; e < m; e += 256) {
                const double a = fabs(aug[e * w + pc]);
                if (a > best) {
                    best = a;
                    bi = e;
                }
            }
            for (int o = 32; o > 0; o >>= 1) {
                const double ob = __shfl_xor(best, o, 64);
                const int oi = __shfl_xor(bi, o, 64);
                if (ob > best || (ob == best && oi < bi)) {
                    best = ob;
                    bi = oi;
                }
            }
            if (lane == 0) {
                rv[wave] = best;
                ri[wave] = bi;
            }
            __syncthreads();
            if (tid == 0) {
                double b = rv[0];
                int i = ri[0];
                for (int k = 1; k < 4; ++k) {
                    if (rv[k] > b || (rv[k] == b && ri[k] < i)) {
                        b = rv[k];
                        i = ri[k];
                    }
                }
                s_piv = (!(b > 64.0 * DBL_EPSILON) || !isfinite(b)) ? -1 : i;
            }
            __syncthreads();
            const int pr = s_piv;
            if (pr < 0) {
                bad = true;
                break;
            }
            // exchange rows pc and pr (columns < pc are zero in both)
            if (pr != pc) {
                for (int c = pc + tid; c < w; c += 256) {
                    const double t = aug[pc * w + c];
                    aug[pc * w + c] = aug[pr * w + c];
                    aug[pr * w + c] = t;
                }
                __syncthreads();
            }
            const double piv = aug[pc * w + pc];
            __syncthreads();
            for (int c = pc + 1 + tid; c < w; c += 256) {
                aug[pc * w + c] /= piv;
            }
            __syncthreads();
            // eliminate column pc from every other equation (columns > pc;
            // column pc itself is not read again): lane tx owns columns
            // pc + 1 + tx + 64 j, whose pivot-row entries it keeps in
            // registers, wavefront ty every fourth equation
            double prow[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int c = pc + 1 + tx + 64 * j;
                prow[j] = c < w ? aug[pc * w + c] : 0.0;
            }
            for (int e = ty; e < m; e += 4) {
                if (e == pc) {
                    continue;
                }
                const double f = aug[e * w + pc];
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const int c = pc + 1 + tx + 64 * j;
                    if (c < w) {
                        aug[e * w + c] -= f * prow[j];
                    }
                }
            }
            __syncthreads();
        }
        if (bad) {
            if (tid == 0) {
                atomicOr(&g.flags[0], 1);
            }
        } else {
            for (int e = tid; e < m; e += 256) {
                double v = aug[e * w + m];
                const double ov = Bc[row + (long)e * ld];
                if (g.damping < 1.0) {
                    v = (1.0 - g.damping) * ov + g.damping * v;
                }
                if (!isfinite(v)) {
                    atomicOr(&g.flags[0], 1);
                    v = 0.0;
                }
                Bnc[row + (long)e * ld] = v;
                sdn += (v - ov) * (v - ov);
                sbn += ov * ov;
            }
        }
        __syncthreads();  // aug is rebuilt for the next row
    }
    for (int o = 32; o > 0; o >>= 1) {
        sdn += __shfl_down(sdn, o, 64);
        sbn += __shfl_down(sbn, o, 64);
    }
    if (lane == 0) {
        sa[wave] = sdn;
        sb[wave] = sbn;
    }
    __syncthreads();
    if (tid == 0) {
        const long slot = (long)cid * g.pstride + blockIdx.x;
        g.partials[2 * slot] = sa[0] + sa[1] + sa[2] + sa[3];
        g.partials[2 * slot + 1] = sb[0] + sb[1] + sb[2] + sb[3];
    }
}

}  // namespace

int block_wide_rows()
{
    return kWideRows;
}

int launch_block_wide(int m, const int *list, int count, const BlockUpdateArgs &g, hipStream_t stream)
{
    if (m <= 64 || m > kMaxWideCluster) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    const size_t lds = (size_t)m * (m + 1) * sizeof(double);
    IPT_HIPCHK(hipFuncSetAttribute((const void *)k_block_update_wide, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
    const dim3 grid((unsigned)((g.n + kWideRows - 1) / kWideRows), (unsigned)count);
    hipLaunchKernelGGL(k_block_update_wide, grid, dim3(256), lds, stream, g, list);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

}  // namespace ipt
