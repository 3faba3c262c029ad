// Operand emission of the block update for the next low-precision product.
//
// In the mixed-precision block IPT the next correction product multiplies
// E by (B_next - A), A the anchor of that product: the iterate B of this
// step when this step is exact (the exact pair becomes the anchor), the
// current anchor otherwise.  Instead of a separate pass that finds the
// column maxima and converts (B_next - A) into the MFMA operand, every
// update kernel writes each new element straight into the operand, in the
// GEMM's B-fragment order (16-bit planes: FP16, BF16, or BF16 hi + lo for
// BF16x3), with one power-of-two scale per cluster taken from the previous
// step (lagged):
//
//   s = 2^(e - 8) for the largest |B - A| of the cluster's columns one step
//   earlier (= f 2^e, f in [0.5, 1)); this step's largest scaled value is
//   then about [128, 256) times its growth over one step.
//
// FP16 keeps full relative precision on [2^-14, 65504]: the lagged value
// may grow 256-fold or shrink two-million-fold before the operand loses
// range, far beyond what one contracting step does; values are saturated
// at the FP16 limit, never turned into infinities.  The per-cluster maxima
// of this step (|B_next - B| and |B_next - A|) go to per-block partials
// that k_block_finalize reduces for the next step.
#pragma once

#include <hip/hip_runtime.h>

#include "ipt_lowp.h"

namespace ipt {

struct LowpEmit {
    unsigned short *zhi = nullptr;  // first plane (NULL: no emission)
    unsigned short *zlo = nullptr;  // second plane (BF16x3)
    int mode = -1;                  // LOWP_FP16 / LOWP_BF16 / LOWP_BF16X3
    int nb_all = 0;                 // column blocks of 16 over all basis columns
    const double *anc = nullptr;    // anchor A of the next product; NULL: A = B (exact step)
    const double *cmax = nullptr;   // per cluster: [2c] step maximum, [2c+1] anchor maximum (previous step)
    const double *fp = nullptr;     // per cluster fixed-point residual (fallback scale)
    double *colscale = nullptr;     // per basis column: the scale written with the operand
    double *mpart = nullptr;        // per (cluster, block) maxima partials, 2 doubles
};

// Scale of cluster cid for this step's operand (see above).
__device__ __forceinline__ double emit_scale(const LowpEmit &em, int cid)
{
    double mx = em.anc == nullptr ? em.cmax[2 * cid] : em.cmax[2 * cid + 1];
    if (!(mx > 0.0) || !isfinite(mx)) {
        mx = em.fp[cid];
    }
    if (!(mx > 0.0) || !isfinite(mx)) {
        return 1.0;
    }
    int e = 0;
    (void)frexp(mx, &e);
    return ldexp(1.0, e - 8);
}

__device__ __forceinline__ unsigned short emit_half_bits(float x)
{
    x = fminf(fmaxf(x, -65504.0f), 65504.0f);  // NaN passes through
    const _Float16 h = (_Float16)x;
    return __builtin_bit_cast(unsigned short, h);
}

__device__ __forceinline__ unsigned short emit_bf16_bits(float x)
{
    unsigned int u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) {
        return (unsigned short)(u >> 16);
    }
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// Write (v - a) / s for basis column cabs of matrix row `row`.  Element
// (row, column) of the operand: k block row/32, column block cabs/16,
// lane (cabs % 16) + 16 ((row % 32) / 8), element row % 8 of the lane's 8.
__device__ __forceinline__ void emit_store(const LowpEmit &em, int mode, long row, int cabs, double v, double a,
                                           double s)
{
    const float x = (float)((v - a) / s);
    const long u4 = ((row >> 5) * em.nb_all + (cabs >> 4)) * 64 + (cabs & 15) + 16 * ((row & 31) >> 3);
    const long e = u4 * 8 + (row & 7);
    if (mode == LOWP_FP16) {
        em.zhi[e] = emit_half_bits(x);
    } else {
        const unsigned short h = emit_bf16_bits(x);
        em.zhi[e] = h;
        if (mode == LOWP_BF16X3) {  // remainder plane
            em.zlo[e] = emit_bf16_bits((float)((double)x - (double)__uint_as_float(((unsigned int)h) << 16)));
        }
    }
}

// Block-wide maxima of (mstep, manc) into mpart[2 slot], [2 slot + 1]; every
// thread of the block must call it (one barrier).
__device__ __forceinline__ void emit_block_max(const LowpEmit &em, long slot, double mstep, double manc)
{
    __shared__ double ms[4], ma[4];
    for (int o = 32; o > 0; o >>= 1) {
        mstep = fmax(mstep, __shfl_down(mstep, o, 64));
        manc = fmax(manc, __shfl_down(manc, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        ms[threadIdx.x >> 6] = mstep;
        ma[threadIdx.x >> 6] = manc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        double a = ms[0], b = ma[0];
        for (int w = 1; w < nw; ++w) {
            a = fmax(a, ms[w]);
            b = fmax(b, ma[w]);
        }
        em.mpart[2 * slot] = a;
        em.mpart[2 * slot + 1] = b;
    }
}

}  // namespace ipt
