// Low-precision MFMA GEMM kernel of the mixed-precision block IPT: the
// correction product E (B - B_a) on gfx950 (FP16 / BF16 / BF16x3 inputs,
// FP32 accumulation).  Operand layouts are produced by k_pack_E and
// k_delta_convert in ipt_gemm_lp.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "ipt_lowp.h"

namespace ipt {

typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef __bf16 b8_t __attribute__((ext_vector_type(8)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u4_t __attribute__((ext_vector_type(4)));

// streamed once: non-temporal 16-byte load
__device__ inline uint4 load_nt(const uint4 *p)
{
    const u4_t v = __builtin_nontemporal_load(reinterpret_cast<const u4_t *>(p));
    uint4 r;
    r.x = v.x;
    r.y = v.y;
    r.z = v.z;
    r.w = v.w;
    return r;
}

// ------------------------------------------------------------------ GEMM

// Partial products of split z (k blocks [8 KW z, 8 KW (z + 1)), KW = 8 or
// 4 k blocks per wavefront) for every row block.
//
// The B fragments of a split are small (KW k blocks x NB column blocks x
// 1 KiB per wavefront) and are held in registers for the whole kernel:
// wavefront w of the 8 in a workgroup owns k blocks [kbw, kbw + KW), loads
// their B fragments once, then walks the workgroup's row blocks.  Per row
// block it issues the KW E-tile loads of the next row block (one 16-byte
// non-temporal load per lane each; the 8 wavefronts together read 64 KiB
// of consecutive HBM), runs KW x NB MFMAs on the current tiles, and parks
// its 16 x 16NB partial in LDS.  After one barrier the workgroup sums the 8
// partials in a fixed order and writes one 16 x 16NB FP32 tile of P.  So
// the inner loop has no B traffic at all, the E stream keeps 8 KiB per
// wavefront in flight across the MFMAs and the barrier, and the result is
// deterministic.
//
// tscale: FP8 only (else unused).
// The operand holds zstride column blocks per k block; the product takes
// the NB blocks from zcb0 on.
// P layout: tile (z, rb, cb) of 1 KiB at ((z * nrb + rb) * NB + cb) * 256
// floats, lane l, element v = MFMA D layout (row 4(l/16) + v, column l%16).
// Row blocks past nrb are handled by clamped loads whose results are never
// stored; nkb is a multiple of 64, so every split is full.
//
// FP8: a 16-byte lane entry holds two k blocks (bytes 0-7 the even one), so
// a wavefront's KW k blocks are KW/2 loads; the partial over each 128-wide k
// tile (4 k blocks) is scaled by that E tile's scale before it is added.
template <int NB, int MODE, int KW>
__global__ __launch_bounds__(512) void k_lowp_gemm(const uint4 *__restrict__ Ehi, const uint4 *__restrict__ Elo,
                                                   const uint4 *__restrict__ Zhi, const uint4 *__restrict__ Zlo,
                                                   const float *__restrict__ tscale, float *__restrict__ P, int nkb,
                                                   int nrb, int rb_per_wg, int zstride, int zcb0)
{
    constexpr int W = 8;  // wavefronts per workgroup
    constexpr bool X3 = MODE == LOWP_BF16X3;
    constexpr bool F8 = MODE == LOWP_FP8;
    constexpr bool TWO = lowp_two_planes(MODE);
    constexpr int NZ = TWO ? 2 : 1;
    constexpr int LK = F8 ? KW / 2 : KW;  // 16-byte lane entries per row block and wavefront
    static_assert(!F8 || KW % 4 == 0, "FP8 wavefronts own whole 128-wide k tiles");
    __shared__ f4_t red[2][W][NB][64];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int z = blockIdx.y;
    const int kbw = (z * W + wave) * KW;
    const int rbA = blockIdx.x * rb_per_wg;
    const int rbB = rbA + rb_per_wg < nrb ? rbA + rb_per_wg : nrb;
    const int nke = F8 ? nkb / 2 : nkb;  // lane entries per row block
    const int ke0 = F8 ? kbw / 2 : kbw;  // first entry of this wavefront

    uint4 bz[NZ][LK][NB];
#pragma unroll
    for (int kk = 0; kk < LK; ++kk) {
#pragma unroll
        for (int cb = 0; cb < NB; ++cb) {
            const long i = ((long)(ke0 + kk) * zstride + zcb0 + cb) * 64 + lane;
            bz[0][kk][cb] = Zhi[i];
            if (TWO) {
                bz[NZ - 1][kk][cb] = Zlo[i];
            }
        }
    }
    // entry (rb, ke0 + kk) of this lane: rowbase(rb) + kk * 64
    const long kofs = (long)ke0 * 64 + lane;
    uint4 eh[2][LK], el[2][LK];
    auto loadE = [&](int rb, int slot) {
        const long o = (long)(rb < rbB ? rb : rbB - 1) * nke * 64 + kofs;
        const uint4 *ph = Ehi + o;
        const uint4 *pl = Elo + o;
#pragma unroll
        for (int kk = 0; kk < LK; ++kk) {
            eh[slot][kk] = load_nt(ph + kk * 64);
            if (TWO) {
                el[slot][kk] = load_nt(pl + kk * 64);
            }
        }
    };
    auto step = [&](int rb, int slot) {
        loadE(rb + 1, slot ^ 1);
        // keep the prefetch ahead of the MFMAs that do not need it
        __builtin_amdgcn_sched_barrier(0);
        f4_t acc[NB];
#pragma unroll
        for (int cb = 0; cb < NB; ++cb) {
            acc[cb] = (f4_t){0.f, 0.f, 0.f, 0.f};
        }
        if constexpr (F8) {
            const int rbs = rb < rbB ? rb : rbB - 1;
            const float *ts = tscale + (long)(rbs >> 3) * (nkb >> 2) + (kbw >> 2);
#pragma unroll
            for (int g = 0; g < KW / 4; ++g) {
                f4_t part[NB];
#pragma unroll
                for (int cb = 0; cb < NB; ++cb) {
                    part[cb] = (f4_t){0.f, 0.f, 0.f, 0.f};
                }
#pragma unroll
                for (int kk = 2 * g; kk < 2 * g + 2; ++kk) {
                    const uint4 a = eh[slot][kk];
                    const long a0 = (long)(((unsigned long)a.y << 32) | a.x);
                    const long a1 = (long)(((unsigned long)a.w << 32) | a.z);
#pragma unroll
                    for (int cb = 0; cb < NB; ++cb) {
                        const uint4 b = bz[0][kk][cb];
                        const long b0 = (long)(((unsigned long)b.y << 32) | b.x);
                        const long b1 = (long)(((unsigned long)b.w << 32) | b.z);
                        part[cb] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a0, b0, part[cb], 0, 0, 0);
                        part[cb] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a1, b1, part[cb], 0, 0, 0);
                    }
                }
                const float sc = ts[g];
#pragma unroll
                for (int cb = 0; cb < NB; ++cb) {
                    acc[cb] += sc * part[cb];
                }
            }
        }
#pragma unroll
        for (int kk = 0; kk < (F8 ? 0 : KW); ++kk) {
#pragma unroll
            for (int cb = 0; cb < NB; ++cb) {
                if (MODE == LOWP_FP32) {
                    // lane l holds k = 4j + l/16 of the k block, j = 0..7
                    const uint4 a0 = eh[slot][kk], a1 = el[slot][kk];
                    const uint4 b0 = bz[0][kk][cb], b1 = bz[NZ - 1][kk][cb];
                    const unsigned int av[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
                    const unsigned int bv[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        acc[cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(av[j]), __uint_as_float(bv[j]),
                                                                       acc[cb], 0, 0, 0);
                    }
                } else if (MODE == LOWP_FP16) {
                    acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8_t, eh[slot][kk]),
                                                                     __builtin_bit_cast(h8_t, bz[0][kk][cb]), acc[cb],
                                                                     0, 0, 0);
                } else {
                    const b8_t bv = __builtin_bit_cast(b8_t, bz[0][kk][cb]);
                    acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8_t, eh[slot][kk]), bv,
                                                                      acc[cb], 0, 0, 0);
                    if (X3) {
                        acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(b8_t, eh[slot][kk]), __builtin_bit_cast(b8_t, bz[NZ - 1][kk][cb]),
                            acc[cb], 0, 0, 0);
                        acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8_t, el[slot][kk]), bv,
                                                                          acc[cb], 0, 0, 0);
                    }
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        // red[slot] was last read two row blocks ago, before the barrier
        // of the previous row block
#pragma unroll
        for (int cb = 0; cb < NB; ++cb) {
            red[slot][wave][cb][lane] = acc[cb];
        }
        __syncthreads();
        const int t = threadIdx.x;
        if (t < NB * 64 && rb < rbB) {
            const int cb = t >> 6, l = t & 63;
            f4_t s = red[slot][0][cb][l];
#pragma unroll
            for (int w = 1; w < W; ++w) {
                s += red[slot][w][cb][l];
            }
            reinterpret_cast<f4_t *>(P)[(((long)z * nrb + rb) * NB + cb) * 64 + l] = s;
        }
    };

    loadE(rbA, 0);
    // an even number of steps keeps the ring slot a compile-time constant;
    // a step past rbB only repeats clamped loads and stores nothing
    for (int rb = rbA; rb < rbB; rb += 2) {
        step(rb, 0);
        step(rb + 1, 1);
    }
}

}  // namespace ipt
