// Low-precision correction product for mixed-precision block IPT on gfx950.
//
// The reference's mixed paths (IPT/src/mixed_precision/ipt_mixed_precision_cuda.cu:
// 189-216, 520-523, 713-716) multiply an FP16/TF32 copy of the off-diagonal
// part E = H - diag(H) by the whole iterate.  Here the low-precision
// operand is the *change* of the iterate since the last exact FP64 product
// (the anchor):
//     H B = H B_a + diag(d)(B - B_a) + E (B - B_a),
// with H B_a known exactly, diag(d)(B - B_a) formed in FP64 and only
// E (B - B_a) computed with MFMA (FP16 or BF16 inputs, FP32 accumulation;
// BF16x3 = three BF16 products of split operands replaces TF32).  Its
// rounding error scales with ||B - B_a||, not with ||B||, which is what lets
// the iteration reach FP64 accuracy with rare FP64 products.
//
// Layouts.  E is packed once into MFMA A-operand order: 16 x 32 tiles
// (row block rb, k block kb), tile-major with kb fastest, lane l's eight
// elements (row 16rb + l%16, k 32kb + 8(l/16) + j) contiguous, so a
// wavefront fetches one 1 KiB tile with one 16-byte load per lane and the
// row stream of a wavefront is sequential in HBM.  The change B - B_a is
// converted per column (power-of-two column scale) into half/bfloat16 in
// MFMA B-operand order (one KiB per k block and 16-column block), which
// workgroups stage through LDS.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "ipt_block.h"
#include "ipt_lowp.h"
#include "ipt_gemm_lp_kernel.h"

#include "../../include/ipt_hip_ext.h"

namespace ipt {

static inline int blocks_for(long total, int threads, int cap)
{
    long b = (total + threads - 1) / threads;
    if (b < 1) {
        b = 1;
    }
    return (int)(b > cap ? cap : b);
}

__device__ inline unsigned short to_half_bits(float x)
{
    _Float16 h = (_Float16)x;
    return __builtin_bit_cast(unsigned short, h);
}

// round-to-nearest-even float -> bfloat16 bits
__device__ inline unsigned short to_bf16_bits(float x)
{
    unsigned int u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) {
        return (unsigned short)(u >> 16);
    }
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ inline float bf16_bits_to_float(unsigned short b) { return __uint_as_float(((unsigned int)b) << 16); }

// float -> OCP FP8 E4M3 bits (1-4-3, bias 7, max 448, no infinity), round to
// nearest even, saturating at 448; x finite.
__device__ inline unsigned int to_e4m3(float x)
{
    const unsigned int sign = (__float_as_uint(x) >> 24) & 0x80u;
    const float a = fminf(fabsf(x), 448.0f);
    unsigned int code;
    if (a < 0.015625f) {
        // below 2^-6: subnormal grid of 2^-9 (a rounded multiple of 8 is the
        // smallest normal, whose code is also 8)
        code = (unsigned int)__builtin_rintf(a * 512.0f);
    } else {
        const unsigned int ua = __float_as_uint(a);
        unsigned int e = ((ua >> 23) & 0xFFu) - 127u + 7u;
        unsigned int r = (ua >> 20) & 7u;
        const unsigned int rem = ua & 0xFFFFFu;
        if (rem > 0x80000u || (rem == 0x80000u && (r & 1u))) {
            ++r;
        }
        if (r == 8u) {
            r = 0u;
            ++e;
        }
        code = (e << 3) | r;
        if (code > 0x7Eu) {
            code = 0x7Eu;
        }
    }
    return sign | code;
}

// eight floats -> 8 E4M3 bytes (element j in byte j)
__device__ inline void pack_e4m3(const float *v, unsigned int &lo, unsigned int &hi)
{
    lo = to_e4m3(v[0]) | (to_e4m3(v[1]) << 8) | (to_e4m3(v[2]) << 16) | (to_e4m3(v[3]) << 24);
    hi = to_e4m3(v[4]) | (to_e4m3(v[5]) << 8) | (to_e4m3(v[6]) << 16) | (to_e4m3(v[7]) << 24);
}

// ------------------------------------------------------------------ pack E

// One thread per (tile, lane): 8 elements.  FP16: fp16 of scale*E;
// BF16: bf16 of E; BF16X3: bf16 hi + bf16 lo of E; FP32: E in FP32 with
// k = 4j + l/16 (plane 0: j < 4, plane 1: j >= 4); the 16-bit modes hold
// k = 8(l/16) + j.
__global__ __launch_bounds__(256) void k_pack_E(const double *__restrict__ H, long ld, int n, int nkb, int nrb,
                                                int mode, float scale, uint4 *__restrict__ hi, uint4 *__restrict__ lo)
{
    const long total = (long)nrb * nkb * 64;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const long tile = t >> 6;
        const int l = (int)(t & 63);
        const long rb = tile / nkb, kb = tile % nkb;
        const long r = rb * 16 + (l & 15);
        double e[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const long k = mode == LOWP_FP32 ? kb * 32 + 4 * j + (l >> 4) : kb * 32 + 8 * (l >> 4) + j;
            e[j] = (r < n && k < n && r != k) ? H[r + k * ld] : 0.0;
        }
        if (mode == LOWP_FP32) {
            uint4 p0, p1;
            p0.x = __float_as_uint((float)e[0]);
            p0.y = __float_as_uint((float)e[1]);
            p0.z = __float_as_uint((float)e[2]);
            p0.w = __float_as_uint((float)e[3]);
            p1.x = __float_as_uint((float)e[4]);
            p1.y = __float_as_uint((float)e[5]);
            p1.z = __float_as_uint((float)e[6]);
            p1.w = __float_as_uint((float)e[7]);
            hi[t] = p0;
            lo[t] = p1;
            continue;
        }
        unsigned short vh[8], vl[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (mode == LOWP_FP16) {
                vh[j] = to_half_bits((float)(e[j] * (double)scale));
                vl[j] = 0;
            } else {
                vh[j] = to_bf16_bits((float)e[j]);
                vl[j] = to_bf16_bits((float)(e[j] - (double)bf16_bits_to_float(vh[j])));
            }
        }
        uint4 ph, pl;
        ph.x = vh[0] | ((unsigned)vh[1] << 16);
        ph.y = vh[2] | ((unsigned)vh[3] << 16);
        ph.z = vh[4] | ((unsigned)vh[5] << 16);
        ph.w = vh[6] | ((unsigned)vh[7] << 16);
        hi[t] = ph;
        if (mode == LOWP_BF16X3) {
            pl.x = vl[0] | ((unsigned)vl[1] << 16);
            pl.y = vl[2] | ((unsigned)vl[3] << 16);
            pl.z = vl[4] | ((unsigned)vl[5] << 16);
            pl.w = vl[6] | ((unsigned)vl[7] << 16);
            lo[t] = pl;
        }
    }
}

// FP8 tile scales: one block per 128 x 128 tile (blockIdx.x = k tile,
// blockIdx.y = row tile) of E = offdiag(H); scale = max|E_tile| / 448, or 1
// for an all-zero tile.
__global__ __launch_bounds__(256) void k_tile_scale(const double *__restrict__ H, long ld, int n, int nkt,
                                                    float *__restrict__ tscale)
{
    const long r0 = (long)blockIdx.y * 128, c0 = (long)blockIdx.x * 128;
    double m = 0.0;
    for (int e = threadIdx.x; e < 128 * 128; e += 256) {
        const long r = r0 + (e & 127), c = c0 + (e >> 7);
        if (r < n && c < n && r != c) {
            m = fmax(m, fabs(H[r + c * ld]));
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        m = fmax(m, __shfl_down(m, o, 64));
    }
    __shared__ double sm[4];
    if ((threadIdx.x & 63) == 0) {
        sm[threadIdx.x >> 6] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
        tscale[(long)blockIdx.y * nkt + blockIdx.x] = m > 0.0 ? (float)(m / 448.0) : 1.0f;
    }
}

// FP8 E: one thread per (row block, k-block pair, lane); bytes 0-7 the
// even k block, 8-15 the odd one, each k = 32 kb + 8(l/16) + j.
__global__ __launch_bounds__(256) void k_pack_E_fp8(const double *__restrict__ H, long ld, int n, int nkb, int nrb,
                                                    const float *__restrict__ tscale, uint4 *__restrict__ out)
{
    const int nkp = nkb / 2;
    const long total = (long)nrb * nkp * 64;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int l = (int)(t & 63);
        const long rb = (t >> 6) / nkp, kp = (t >> 6) % nkp;
        const long r = rb * 16 + (l & 15);
        unsigned int w[4];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const long kb = 2 * kp + h;
            const float inv = 1.0f / tscale[(rb >> 3) * (nkb >> 2) + (kb >> 2)];
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const long k = kb * 32 + 8 * (l >> 4) + j;
                v[j] = (r < n && k < n && r != k) ? (float)H[r + k * ld] * inv : 0.0f;
            }
            pack_e4m3(v, w[2 * h], w[2 * h + 1]);
        }
        uint4 q;
        q.x = w[0];
        q.y = w[1];
        q.z = w[2];
        q.w = w[3];
        out[t] = q;
    }
}

int lowp_operator_create(Context *ctx, const double *H, long ld, int n, int mode, double max_abs_offdiag,
                         LowpOperator *op)
{
    op->n = n;
    op->mode = mode;
    // k blocks padded to a multiple of 64 (zero tiles) so that the GEMM's
    // splits of 64 (or 32) k blocks are all full
    op->nkb = (n + 2047) / 2048 * 64;
    op->nrb = (n + 15) / 16;
    op->np = op->nkb * 32;
    op->scale = 1.0f;
    if (mode == LOWP_FP16) {
        // power-of-two scale putting max|E| just below 2^15 (FP16 max 65504)
        double m = max_abs_offdiag > 0.0 ? max_abs_offdiag : 1.0;
        int e = (int)std::floor(std::log2(32768.0 / m));
        if (e > 100) {
            e = 100;
        }
        if (e < -100) {
            e = -100;
        }
        op->scale = (float)std::ldexp(1.0, e);
    }
    if (mode == LOWP_FP8) {
        const int nkt = op->nkb / 4, nrt = (op->nrb + 7) / 8;
        IPT_STCHK(op->tscale.alloc((size_t)nrt * nkt));
        IPT_STCHK(op->hi.alloc((size_t)op->nrb * op->nkb * 256));  // one byte per element
        hipLaunchKernelGGL(k_tile_scale, dim3(nkt, nrt), dim3(256), 0, ctx->stream, H, ld, n, nkt, op->tscale.p);
        IPT_HIPCHK(hipGetLastError());
        const long total = (long)op->nrb * (op->nkb / 2) * 64;
        hipLaunchKernelGGL(k_pack_E_fp8, dim3(blocks_for(total, 256, 65536)), dim3(256), 0, ctx->stream, H, ld, n,
                           op->nkb, op->nrb, op->tscale.p, reinterpret_cast<uint4 *>(op->hi.p));
        IPT_HIPCHK(hipGetLastError());
        return IPT_CUDA_SUCCESS;
    }
    const size_t elems = (size_t)op->nrb * op->nkb * 512;
    IPT_STCHK(op->hi.alloc(elems));
    if (lowp_two_planes(mode)) {
        IPT_STCHK(op->lo.alloc(elems));
    }
    const long total = (long)op->nrb * op->nkb * 64;
    hipLaunchKernelGGL(k_pack_E, dim3(blocks_for(total, 256, 65536)), dim3(256), 0, ctx->stream, H, ld, n, op->nkb,
                       op->nrb, mode, op->scale, reinterpret_cast<uint4 *>(op->hi.p),
                       reinterpret_cast<uint4 *>(op->lo.p));
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

// ------------------------------------------------------------------ delta

// colmax_part[b * nc + c] = max over row slice b of |B - Ba| (Ba NULL:
// zero): blockIdx.y = column, blockIdx.x = row slice b.  The converter
// takes the maximum over the slices.
__global__ __launch_bounds__(256) void k_delta_colmax(const double *__restrict__ B, const double *__restrict__ Ba,
                                                      long ld, int n, int nc, double *__restrict__ colmax_part)
{
    const int c = blockIdx.y;
    double m = 0.0;
    for (long r = (long)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x) {
        m = fmax(m, fabs(B[r + (long)c * ld] - (Ba != nullptr ? Ba[r + (long)c * ld] : 0.0)));
    }
    for (int o = 32; o > 0; o >>= 1) {
        m = fmax(m, __shfl_down(m, o, 64));
    }
    __shared__ double sm[4];
    if ((threadIdx.x & 63) == 0) {
        sm[threadIdx.x >> 6] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // one partial per block: no zero-initialised accumulator needed
        colmax_part[(long)blockIdx.x * nc + c] = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
    }
}

// s_c = 2^ceil(log2 max) (1 when the column did not move); Z = (B-Ba)/s_c
// written fragment-major: for k block kb and column block cb the 64 lanes'
// MFMA B fragments (column 16cb + l%16; k as in k_pack_E) are one
// contiguous KiB per plane, so a wavefront loads its fragments with plain
// coalesced 16-byte loads.  Columns >= nc and rows >= n are zero.  Ba may be
// NULL (zero anchor).
__global__ __launch_bounds__(256) void k_delta_convert(const double *__restrict__ B, const double *__restrict__ Ba,
                                                       long ld, int n, int nc, int nkb, int nb, int mode,
                                                       const double *__restrict__ colmax_part, int nparts,
                                                       double *__restrict__ colscale, uint4 *__restrict__ zhi,
                                                       uint4 *__restrict__ zlo)
{
    const long total = (long)(mode == LOWP_FP8 ? nkb / 2 : nkb) * nb * 64;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int l = (int)(t & 63);
        const int cb = (int)((t >> 6) % nb);
        const long kb = (t >> 6) / nb;  // FP8: k-block pair
        const int c = cb * 16 + (l & 15);
        double s = 1.0;
        if (c < nc) {
            double mx = 0.0;
            for (int q = 0; q < nparts; ++q) {
                mx = fmax(mx, colmax_part[(long)q * nc + c]);
            }
            if (mx > 0.0) {
                int e = 0;
                (void)frexp(mx, &e);  // mx = f * 2^e, f in [0.5,1)
                // FP8: scaled column maximum in [128, 256) (E4M3 max 448)
                s = ldexp(1.0, mode == LOWP_FP8 ? e - 8 : e);
            }
            if (t < 64 * nb && (l >> 4) == 0) {
                colscale[c] = s;  // k block 0, first lane group: once per column
            }
        }
        if (mode == LOWP_FP8) {
            unsigned int w[4];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                float v[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const long k = (2 * kb + h) * 32 + 8 * (l >> 4) + j;
                    double x = 0.0;
                    if (c < nc && k < n) {
                        const long o = k + (long)c * ld;
                        x = (B[o] - (Ba != nullptr ? Ba[o] : 0.0)) / s;
                    }
                    v[j] = (float)x;
                }
                pack_e4m3(v, w[2 * h], w[2 * h + 1]);
            }
            uint4 q;
            q.x = w[0];
            q.y = w[1];
            q.z = w[2];
            q.w = w[3];
            zhi[t] = q;
            continue;
        }
        double x[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const long k = mode == LOWP_FP32 ? kb * 32 + 4 * j + (l >> 4) : kb * 32 + 8 * (l >> 4) + j;
            x[j] = 0.0;
            if (c < nc && k < n) {
                const long o = k + (long)c * ld;
                x[j] = (B[o] - (Ba != nullptr ? Ba[o] : 0.0)) / s;
            }
        }
        if (mode == LOWP_FP32) {
            uint4 p0, p1;
            p0.x = __float_as_uint((float)x[0]);
            p0.y = __float_as_uint((float)x[1]);
            p0.z = __float_as_uint((float)x[2]);
            p0.w = __float_as_uint((float)x[3]);
            p1.x = __float_as_uint((float)x[4]);
            p1.y = __float_as_uint((float)x[5]);
            p1.z = __float_as_uint((float)x[6]);
            p1.w = __float_as_uint((float)x[7]);
            zhi[t] = p0;
            zlo[t] = p1;
            continue;
        }
        unsigned short vh[8], vl[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (mode == LOWP_FP16) {
                vh[j] = to_half_bits((float)x[j]);
                vl[j] = 0;
            } else {
                vh[j] = to_bf16_bits((float)x[j]);
                vl[j] = to_bf16_bits((float)(x[j] - (double)bf16_bits_to_float(vh[j])));
            }
        }
        uint4 ph;
        ph.x = vh[0] | ((unsigned)vh[1] << 16);
        ph.y = vh[2] | ((unsigned)vh[3] << 16);
        ph.z = vh[4] | ((unsigned)vh[5] << 16);
        ph.w = vh[6] | ((unsigned)vh[7] << 16);
        zhi[t] = ph;
        if (mode == LOWP_BF16X3) {
            uint4 pl;
            pl.x = vl[0] | ((unsigned)vl[1] << 16);
            pl.y = vl[2] | ((unsigned)vl[3] << 16);
            pl.z = vl[4] | ((unsigned)vl[5] << 16);
            pl.w = vl[6] | ((unsigned)vl[7] << 16);
            zlo[t] = pl;
        }
    }
}

// Y = Ya + d (B - Ba) + colscale/scale * sum_z P[z]; one thread per P
// lane entry (4 rows of one column), splits summed in order.  Ya and Ba
// may be NULL (zero).
__global__ __launch_bounds__(256) void k_lowp_epilogue(const f4_t *__restrict__ P, int splits, int n, int nc, int nrb,
                                                       int nb, const double *__restrict__ Ya,
                                                       const double *__restrict__ B, const double *__restrict__ Ba,
                                                       long ld, const double *__restrict__ d,
                                                       const double *__restrict__ colscale, double inv_scale,
                                                       double *__restrict__ Y)
{
    const long per_split = (long)nrb * nb * 64;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < per_split; t += (long)gridDim.x * blockDim.x) {
        const int l = (int)(t & 63);
        const int cb = (int)((t >> 6) % nb);
        const long rb = (t >> 6) / nb;
        const int c = cb * 16 + (l & 15);
        if (c >= nc) {
            continue;
        }
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        for (int z = 0; z < splits; ++z) {
            const f4_t v = P[(long)z * per_split + t];
            s0 += (double)v[0];
            s1 += (double)v[1];
            s2 += (double)v[2];
            s3 += (double)v[3];
        }
        const double f = colscale[c] * inv_scale;
        const double sv[4] = {s0, s1, s2, s3};
        const long r0 = rb * 16 + 4 * (l >> 4);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const long r = r0 + v;
            if (r < n) {
                const long o = r + (long)c * ld;
                const double ba = Ba != nullptr ? Ba[o] : 0.0;
                Y[o] = (Ya != nullptr ? Ya[o] : 0.0) + d[r] * (B[o] - ba) + sv[v] * f;
            }
        }
    }
}

// k blocks per wavefront: the B fragments it holds in registers take
// KW x NB x 4 VGPRs per plane (FP8: half, two k blocks per 16-byte entry);
// chosen so no variant spills (256 VGPRs at two wavefronts per SIMD; FP8 at
// NB = 8 and KW = 8: 230).  Fewer k blocks per wavefront mean more splits,
// so more FP32 partials per product: FP8 at 128 columns with KW = 4 wrote
// as many partial bytes as it read of E
constexpr int lowp_kw(int mode, int nb)
{
    return mode == LOWP_FP8 ? 8
                            : lowp_two_planes(mode) ? (nb == 1 ? 8 : nb <= 5 ? 4 : 2) : (nb <= 5 ? 8 : 4);
}

template <int MODE, int NB>
static void launch_lowp_nb(const ProfileToken &prof, dim3 grid, hipStream_t st, const uint4 *Ehi, const uint4 *Elo,
                           const uint4 *Zhi, const uint4 *Zlo, const float *ts, float *P, int nkb, int nrb, int rbw,
                           int zstride, int zcb0)
{
    launch_timed(prof, true, true, k_lowp_gemm<NB, MODE, lowp_kw(MODE, NB)>, grid, dim3(512), 0, st, Ehi, Elo, Zhi,
                 Zlo, ts, P, nkb, nrb, rbw, zstride, zcb0);
}

template <int MODE>
static void launch_lowp(const ProfileToken &prof, int NB, dim3 grid, hipStream_t st, const uint4 *Ehi,
                        const uint4 *Elo, const uint4 *Zhi, const uint4 *Zlo, const float *ts, float *P, int nkb,
                        int nrb, int rbw, int zstride, int zcb0)
{
    switch (NB) {
    case 1: launch_lowp_nb<MODE, 1>(prof, grid, st, Ehi, Elo, Zhi, Zlo, ts, P, nkb, nrb, rbw, zstride, zcb0); break;
    case 2: launch_lowp_nb<MODE, 2>(prof, grid, st, Ehi, Elo, Zhi, Zlo, ts, P, nkb, nrb, rbw, zstride, zcb0); break;
    case 3: launch_lowp_nb<MODE, 3>(prof, grid, st, Ehi, Elo, Zhi, Zlo, ts, P, nkb, nrb, rbw, zstride, zcb0); break;
    case 4: launch_lowp_nb<MODE, 4>(prof, grid, st, Ehi, Elo, Zhi, Zlo, ts, P, nkb, nrb, rbw, zstride, zcb0); break;
    case 5: launch_lowp_nb<MODE, 5>(prof, grid, st, Ehi, Elo, Zhi, Zlo, ts, P, nkb, nrb, rbw, zstride, zcb0); break;
    case 6: launch_lowp_nb<MODE, 6>(prof, grid, st, Ehi, Elo, Zhi, Zlo, ts, P, nkb, nrb, rbw, zstride, zcb0); break;
    case 7: launch_lowp_nb<MODE, 7>(prof, grid, st, Ehi, Elo, Zhi, Zlo, ts, P, nkb, nrb, rbw, zstride, zcb0); break;
    default: launch_lowp_nb<MODE, 8>(prof, grid, st, Ehi, Elo, Zhi, Zlo, ts, P, nkb, nrb, rbw, zstride, zcb0); break;
    }
}

// Grid of one product: splits of 8 wavefronts' k blocks, row groups so that
// about one workgroup per CU runs.
struct LowpGrid {
    int splits, rbw, gx;
};

static LowpGrid lowp_grid(const LowpOperator &op, int NB)
{
    const int kw = lowp_kw(op.mode, NB);
    const int kspan = 8 * kw;
    LowpGrid g;
    g.splits = op.nkb / kspan;  // nkb is a multiple of 64 >= kspan
    const int groups = std::max(1, std::min(op.nrb, (256 + g.splits - 1) / g.splits));
    g.rbw = (op.nrb + groups - 1) / groups;
    g.gx = (op.nrb + g.rbw - 1) / g.rbw;
    return g;
}

static void launch_lowp_mode(const ProfileToken &prof, const LowpOperator &op, int NB, dim3 grid, hipStream_t st,
                             const uint4 *zhi, const uint4 *zlo, float *P, int rbw, int zstride, int zcb0)
{
    const uint4 *Ehi = reinterpret_cast<const uint4 *>(op.hi.p);
    const uint4 *Elo = reinterpret_cast<const uint4 *>(op.lo.p);
    const float *ts = op.tscale.p;
    if (op.mode == LOWP_BF16X3) {
        launch_lowp<LOWP_BF16X3>(prof, NB, grid, st, Ehi, Elo, zhi, zlo, ts, P, op.nkb, op.nrb, rbw, zstride, zcb0);
    } else if (op.mode == LOWP_FP32) {
        launch_lowp<LOWP_FP32>(prof, NB, grid, st, Ehi, Elo, zhi, zlo, ts, P, op.nkb, op.nrb, rbw, zstride, zcb0);
    } else if (op.mode == LOWP_FP8) {
        launch_lowp<LOWP_FP8>(prof, NB, grid, st, Ehi, Elo, zhi, zlo, ts, P, op.nkb, op.nrb, rbw, zstride, zcb0);
    } else if (op.mode == LOWP_FP16) {
        launch_lowp<LOWP_FP16>(prof, NB, grid, st, Ehi, Elo, zhi, zlo, ts, P, op.nkb, op.nrb, rbw, zstride, zcb0);
    } else {
        launch_lowp<LOWP_BF16>(prof, NB, grid, st, Ehi, Elo, zhi, zlo, ts, P, op.nkb, op.nrb, rbw, zstride, zcb0);
    }
}

int lowp_apply_delta(Context *ctx, const LowpOperator &op, const double *B, const double *Ba, const double *Ya,
                     long ld, const double *d, int nc, double *Y)
{
    const int n = op.n;
    for (int c0 = 0; c0 < nc; c0 += 128) {
        const int w = std::min(128, nc - c0);
        const int NB = (w + 15) / 16;
        // one split = the k blocks of 8 wavefronts; about one workgroup of
        // 8 wavefronts per CU over 256 CUs
        const LowpGrid lg = lowp_grid(op, NB);
        const int splits = lg.splits, rbw = lg.rbw, gx = lg.gx;
        // scratch: colmax partials (nparts x w) | colscale (w) | Z hi/lo
        // (fragment-major) | partial sums
        const int nparts = std::min(8, (n + 255) / 256);
        const size_t zvec = (size_t)op.nkb * NB * 64;  // uint4 per plane
        const size_t zoff = (size_t)(nparts + 1) * w * 8;
        const size_t zoff_al = (zoff + 255) / 256 * 256;
        const size_t poff = zoff_al + 2 * zvec * 16;
        const size_t bytes = poff + (size_t)splits * op.nrb * NB * 1024 + 256;
        char *base = nullptr;
        IPT_STCHK(context_scratch(ctx, SLOT_LOWP, bytes, (void **)&base));
        double *colmax = reinterpret_cast<double *>(base);
        double *colscale = reinterpret_cast<double *>(base + (size_t)nparts * w * 8);
        uint4 *zhi = reinterpret_cast<uint4 *>(base + zoff_al);
        uint4 *zlo = zhi + zvec;
        float *P = reinterpret_cast<float *>(base + poff);
        const double *Bc = B + (size_t)c0 * ld;
        // a NULL anchor stays NULL in every column chunk (zero anchor)
        const double *Bac = Ba != nullptr ? Ba + (size_t)c0 * ld : nullptr;
        const double *Yac = Ya != nullptr ? Ya + (size_t)c0 * ld : nullptr;
        hipLaunchKernelGGL(k_delta_colmax, dim3(nparts, w), dim3(256), 0, ctx->stream, Bc, Bac, ld, n, w, colmax);
        IPT_HIPCHK(hipGetLastError());
        const size_t zconv = op.mode == LOWP_FP8 ? zvec / 2 : zvec;  // FP8: one entry per k-block pair
        hipLaunchKernelGGL(k_delta_convert, dim3(blocks_for((long)zconv, 256, 16384)), dim3(256), 0, ctx->stream, Bc,
                           Bac, ld, n, w, op.nkb, NB, op.mode, colmax, nparts, colscale, zhi, zlo);
        IPT_HIPCHK(hipGetLastError());
        dim3 grid(gx, splits);
        const ProfileToken prof = profile_begin_launches(PROF_LOWP_GEMM);
        launch_lowp_mode(prof, op, NB, grid, ctx->stream, zhi, zlo, P, rbw, NB, 0);
        profile_end(ctx, prof, w);
        IPT_HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_lowp_epilogue, dim3(blocks_for((long)op.nrb * NB * 64, 256, 16384)), dim3(256), 0,
                           ctx->stream, reinterpret_cast<const f4_t *>(P), splits, n, w, op.nrb, NB,
                           Yac, Bc, Bac, ld, d, colscale, 1.0 / (double)op.scale,
                           Y + (size_t)c0 * ld);
        IPT_HIPCHK(hipGetLastError());
    }
    return IPT_CUDA_SUCCESS;
}

int lowp_apply_emitted(Context *ctx, const LowpOperator &op, const unsigned short *zhi, const unsigned short *zlo,
                       int nb_all, int c_lo, int c_w, const double *colscale, const double *B, const double *Ba,
                       const double *Ya, long ld, const double *d, int bc, double *Y)
{
    const int n = op.n;
    const int cb_end = (c_lo + c_w + 15) / 16;
    for (int cb = c_lo / 16; cb < cb_end; cb += 8) {
        const int NB = std::min(8, cb_end - cb);
        const LowpGrid lg = lowp_grid(op, NB);
        float *P = nullptr;
        IPT_STCHK(context_scratch(ctx, SLOT_LOWP, (size_t)lg.splits * op.nrb * NB * 1024 + 256, (void **)&P));
        const ProfileToken prof = profile_begin_launches(PROF_LOWP_GEMM);
        launch_lowp_mode(prof, op, NB, dim3(lg.gx, lg.splits), ctx->stream, reinterpret_cast<const uint4 *>(zhi),
                         reinterpret_cast<const uint4 *>(zlo), P, lg.rbw, nb_all, cb);
        profile_end(ctx, prof, std::min(16 * NB, bc - 16 * cb));
        IPT_HIPCHK(hipGetLastError());
        // columns past the basis (the last block's padding) are not written
        const int c0 = 16 * cb;
        const int w = std::min(16 * NB, bc - c0);
        hipLaunchKernelGGL(k_lowp_epilogue, dim3(blocks_for((long)op.nrb * NB * 64, 256, 16384)), dim3(256), 0,
                           ctx->stream, reinterpret_cast<const f4_t *>(P), lg.splits, n, w, op.nrb, NB,
                           Ya + (size_t)c0 * ld, B + (size_t)c0 * ld, Ba + (size_t)c0 * ld, ld, d, colscale + c0,
                           1.0 / (double)op.scale, Y + (size_t)c0 * ld);
        IPT_HIPCHK(hipGetLastError());
    }
    return IPT_CUDA_SUCCESS;
}

}  // namespace ipt

using namespace ipt;

extern "C" int ipt_hip_lowp_product(const double *d_H, int n, int mode, const double *d_B, const double *d_Ba,
                                    const double *d_Ya, const double *d_diag, int nc, double *d_Y, void *handle)
{
    if (d_H == nullptr || d_B == nullptr || d_Ba == nullptr || d_Ya == nullptr || d_diag == nullptr ||
        d_Y == nullptr || n <= 0 || nc <= 0 || mode < LOWP_FP16 || mode > LOWP_FP8) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    // max |offdiag| sizes the FP16 scale
    double mx = 0.0;
    {
        DevBuf<double> colsum, colmax;
        DevBuf<int> asym;
        IPT_STCHK(colsum.alloc((size_t)n));
        IPT_STCHK(colmax.alloc((size_t)n));
        IPT_STCHK(asym.alloc(1));
        IPT_STCHK(dense_scan(ctx, d_H, n, n, colsum.p, colmax.p, asym.p));
        std::vector<double> h((size_t)n);
        IPT_HIPCHK(hipMemcpyAsync(h.data(), colmax.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
        for (double x : h) {
            mx = std::max(mx, x);
        }
    }
    LowpOperator op;
    IPT_STCHK(lowp_operator_create(ctx, d_H, n, n, mode, mx, &op));
    IPT_STCHK(lowp_apply_delta(ctx, op, d_B, d_Ba, d_Ya, n, d_diag, nc, d_Y));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return IPT_CUDA_SUCCESS;
}
