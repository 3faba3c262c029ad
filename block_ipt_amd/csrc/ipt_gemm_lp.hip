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
// converted per column (power-of-two column scale) into a column-major
// half/bfloat16 array with k contiguous, so a lane's B fragment is one
// 16-byte load served from L2.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "ipt_block.h"
#include "ipt_lowp.h"

#include "../../include/ipt_hip_ext.h"

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

// ------------------------------------------------------------------ pack E

// One thread per (tile, lane): 8 elements.  mode 0: fp16 of scale*E;
// mode 1: bf16 of E; mode 2: bf16 hi + bf16 lo of E.
__global__ __launch_bounds__(256) void k_pack_E(const double *__restrict__ H, long ld, int n, int nkb, int nrb,
                                                int mode, float scale, unsigned short *__restrict__ hi,
                                                unsigned short *__restrict__ lo)
{
    const long total = (long)nrb * nkb * 64;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const long tile = t >> 6;
        const int l = (int)(t & 63);
        const long rb = tile / nkb, kb = tile % nkb;
        const long r = rb * 16 + (l & 15);
        const long k0 = kb * 32 + 8 * (l >> 4);
        unsigned short vh[8], vl[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const long k = k0 + j;
            double e = 0.0;
            if (r < n && k < n && r != k) {
                e = H[r + k * ld];
            }
            if (mode == 0) {
                vh[j] = to_half_bits((float)(e * (double)scale));
                vl[j] = 0;
            } else {
                const float f = (float)e;
                vh[j] = to_bf16_bits(f);
                vl[j] = to_bf16_bits((float)(e - (double)bf16_bits_to_float(vh[j])));
            }
        }
        uint4 ph, pl;
        ph.x = vh[0] | ((unsigned)vh[1] << 16);
        ph.y = vh[2] | ((unsigned)vh[3] << 16);
        ph.z = vh[4] | ((unsigned)vh[5] << 16);
        ph.w = vh[6] | ((unsigned)vh[7] << 16);
        reinterpret_cast<uint4 *>(hi)[t] = ph;
        if (mode == 2) {
            pl.x = vl[0] | ((unsigned)vl[1] << 16);
            pl.y = vl[2] | ((unsigned)vl[3] << 16);
            pl.z = vl[4] | ((unsigned)vl[5] << 16);
            pl.w = vl[6] | ((unsigned)vl[7] << 16);
            reinterpret_cast<uint4 *>(lo)[t] = pl;
        }
    }
}

int lowp_operator_create(Context *ctx, const double *H, long ld, int n, int mode, double max_abs_offdiag,
                         LowpOperator *op)
{
    op->n = n;
    op->mode = mode;
    op->nkb = (n + 31) / 32;
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
    const size_t elems = (size_t)op->nrb * op->nkb * 512;
    IPT_STCHK(op->hi.alloc(elems));
    if (mode == LOWP_BF16X3) {
        IPT_STCHK(op->lo.alloc(elems));
    }
    const long total = (long)op->nrb * op->nkb * 64;
    hipLaunchKernelGGL(k_pack_E, dim3(blocks_for(total, 256, 65536)), dim3(256), 0, ctx->stream, H, ld, n, op->nkb,
                       op->nrb, mode, op->scale, op->hi.p, op->lo.p);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

// ------------------------------------------------------------------ delta

// colmax[c] = max_r |B - Ba|: blockIdx.y = column, blockIdx.x = row
// quarter; one atomicMax on the IEEE bits per block (order free).
__global__ __launch_bounds__(256) void k_delta_colmax(const double *__restrict__ B, const double *__restrict__ Ba,
                                                      long ld, int n, int nc, unsigned long long *__restrict__ colmax)
{
    const int c = blockIdx.y;
    double m = 0.0;
    for (long r = (long)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (long)gridDim.x * blockDim.x) {
        m = fmax(m, fabs(B[r + (long)c * ld] - Ba[r + (long)c * ld]));
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
        if (m > 0.0) {
            atomicMax(&colmax[c], (unsigned long long)__double_as_longlong(m));
        }
    }
}

// s_c = 2^ceil(log2 max) (1 when the column did not move); Z = (B-Ba)/s_c
__global__ void k_delta_convert(const double *__restrict__ B, const double *__restrict__ Ba, long ld, int n, int nc,
                                int np, int mode, const unsigned long long *__restrict__ colmax,
                                double *__restrict__ colscale, unsigned short *__restrict__ zhi,
                                unsigned short *__restrict__ zlo)
{
    const long total = (long)np * nc;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const long k = t % np;
        const int c = (int)(t / np);
        const double mx = __longlong_as_double((long long)colmax[c]);
        int e = 0;
        if (mx > 0.0) {
            (void)frexp(mx, &e);  // mx = f * 2^e, f in [0.5,1)
        }
        const double s = mx > 0.0 ? ldexp(1.0, e) : 1.0;
        if (k == 0) {
            colscale[c] = s;
        }
        double x = 0.0;
        if (k < n) {
            x = (B[k + (long)c * ld] - Ba[k + (long)c * ld]) / s;
        }
        if (mode == LOWP_FP16) {
            zhi[t] = to_half_bits((float)x);
        } else {
            const unsigned short h = to_bf16_bits((float)x);
            zhi[t] = h;
            if (mode == LOWP_BF16X3) {
                zlo[t] = to_bf16_bits((float)(x - (double)bf16_bits_to_float(h)));
            }
        }
    }
}

// ------------------------------------------------------------------ GEMM

// Partial products P[z] (FP32, column-major n x nc) over the k blocks of
// split z.  A workgroup (4 wavefronts) owns 8 row blocks (128 rows); all
// waves walk the same k blocks so each B fragment is fetched from L2 once
// per workgroup and re-served from L1 to the other waves, and each
// fragment feeds two MFMAs per wave.  E tiles are streamed with
// non-temporal loads (no reuse, keep L2 for B) two k blocks ahead of use.
template <int NB, int MODE, int DEPTH>
__global__ __launch_bounds__(256) void k_lowp_gemm(const uint4 *__restrict__ Ehi, const uint4 *__restrict__ Elo,
                                                   const uint4 *__restrict__ Zhi, const uint4 *__restrict__ Zlo,
                                                   float *__restrict__ P, int n, int nc, int nkb, int nrb, int np,
                                                   int kb_per_split)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int rb0 = (blockIdx.x * 4 + wave) * 2;
    const int kb0 = blockIdx.z * kb_per_split;
    int kb1 = kb0 + kb_per_split;
    if (kb1 > nkb) {
        kb1 = nkb;
    }
    const bool ok0 = rb0 < nrb, ok1 = rb0 + 1 < nrb;
    const int li = lane & 15, lk = lane >> 4;
    const uint4 zero = {0u, 0u, 0u, 0u};
    f4_t acc[2][NB];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            acc[a][c] = (f4_t){0.f, 0.f, 0.f, 0.f};
        }
    }
    const long arow0 = (long)rb0 * nkb * 64 + lane;
    const long arow1 = arow0 + (long)nkb * 64;
    auto loadA = [&](const uint4 *E, int kb, uint4 &x0, uint4 &x1) {
        x0 = zero;
        x1 = zero;
        if (kb < kb1) {
            if (ok0) {
                x0 = load_nt(E + arow0 + (long)kb * 64);
            }
            if (ok1) {
                x1 = load_nt(E + arow1 + (long)kb * 64);
            }
        }
    };
    auto loadB = [&](const uint4 *Z, int kb, uint4 *b) {
#pragma unroll
        for (int cb = 0; cb < NB; ++cb) {
            const int col = cb * 16 + li;
            b[cb] = (col < nc && kb < kb1) ? Z[((long)col * np + (long)kb * 32 + 8 * lk) >> 3] : zero;
        }
    };
    // E tiles and B fragments of k block kb+DEPTH are issued together while
    // kb is consumed: vector loads complete in issue order, so everything a
    // step needs was issued DEPTH steps earlier than anything it does not.
    uint4 a0[DEPTH], a1[DEPTH], a0l[DEPTH], a1l[DEPTH], bh[DEPTH][NB], bl[DEPTH][NB];
#pragma unroll
    for (int u = 0; u < DEPTH; ++u) {
        loadA(Ehi, kb0 + u, a0[u], a1[u]);
        loadB(Zhi, kb0 + u, bh[u]);
        if (MODE == LOWP_BF16X3) {
            loadA(Elo, kb0 + u, a0l[u], a1l[u]);
            loadB(Zlo, kb0 + u, bl[u]);
        }
    }
    for (int kb = kb0; kb < kb1; kb += DEPTH) {
#pragma unroll
        for (int u = 0; u < DEPTH; ++u) {
            const int kc = kb + u;
            if (kc < kb1) {
                // consume slot u, then refill it with k block kc + DEPTH
#pragma unroll
                for (int cb = 0; cb < NB; ++cb) {
                    if (MODE == LOWP_FP16) {
                        const h8_t bv = __builtin_bit_cast(h8_t, bh[u][cb]);
                        acc[0][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8_t, a0[u]), bv, acc[0][cb], 0, 0, 0);
                        acc[1][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8_t, a1[u]), bv, acc[1][cb], 0, 0, 0);
                    } else {
                        const b8_t bv = __builtin_bit_cast(b8_t, bh[u][cb]);
                        acc[0][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8_t, a0[u]), bv, acc[0][cb], 0, 0, 0);
                        acc[1][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8_t, a1[u]), bv, acc[1][cb], 0, 0, 0);
                        if (MODE == LOWP_BF16X3) {
                            const b8_t blv = __builtin_bit_cast(b8_t, bl[u][cb]);
                            acc[0][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8_t, a0[u]), blv, acc[0][cb], 0, 0, 0);
                            acc[1][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8_t, a1[u]), blv, acc[1][cb], 0, 0, 0);
                            acc[0][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8_t, a0l[u]), bv, acc[0][cb], 0, 0, 0);
                            acc[1][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8_t, a1l[u]), bv, acc[1][cb], 0, 0, 0);
                        }
                    }
                }
                loadA(Ehi, kc + DEPTH, a0[u], a1[u]);
                loadB(Zhi, kc + DEPTH, bh[u]);
                if (MODE == LOWP_BF16X3) {
                    loadA(Elo, kc + DEPTH, a0l[u], a1l[u]);
                    loadB(Zlo, kc + DEPTH, bl[u]);
                }
            }
        }
    }
    // D layout (f32 16x16x32): lane l, element v -> row 4*(l/16)+v, col l%16
    float *Pz = P + (long)blockIdx.z * n * nc;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int rb = rb0 + a;
#pragma unroll
        for (int cb = 0; cb < NB; ++cb) {
            const int c = cb * 16 + li;
            if (c >= nc) {
                continue;
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long r = (long)rb * 16 + 4 * lk + v;
                if (r < n) {
                    Pz[r + (long)c * n] = acc[a][cb][v];
                }
            }
        }
    }
}

// Y = Ya + d (B - Ba) + colscale/scale * sum_z P[z]
__global__ void k_lowp_epilogue(const float *__restrict__ P, int splits, int n, int nc,
                                const double *__restrict__ Ya, const double *__restrict__ B,
                                const double *__restrict__ Ba, long ld, const double *__restrict__ d,
                                const double *__restrict__ colscale, double inv_scale, double *__restrict__ Y)
{
    const long total = (long)n * nc;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const long r = t % n, c = t / n;
        double s = 0.0;
        for (int z = 0; z < splits; ++z) {
            s += (double)P[(long)z * total + t];
        }
        const long o = r + c * ld;
        Y[o] = Ya[o] + d[r] * (B[o] - Ba[o]) + s * colscale[c] * inv_scale;
    }
}

template <int MODE, int DEP>
static void launch_lowp(int NB, dim3 grid, hipStream_t st, const uint4 *Ehi, const uint4 *Elo, const uint4 *Zhi,
                        const uint4 *Zlo, float *P, int n, int nc, int nkb, int nrb, int np, int kbs)
{
#define IPT_LP(NBV)                                                                                              \
    hipLaunchKernelGGL((k_lowp_gemm<NBV, MODE, DEP>), grid, dim3(256), 0, st, Ehi, Elo, Zhi, Zlo, P, n, nc, nkb, nrb, np, \
                       kbs)
    switch (NB) {
    case 1: IPT_LP(1); break;
    case 2: IPT_LP(2); break;
    case 3: IPT_LP(3); break;
    case 4: IPT_LP(4); break;
    case 5: IPT_LP(5); break;
    case 6: IPT_LP(6); break;
    case 7: IPT_LP(7); break;
    default: IPT_LP(8); break;
    }
#undef IPT_LP
}

int lowp_apply_delta(Context *ctx, const LowpOperator &op, const double *B, const double *Ba, const double *Ya,
                     long ld, const double *d, int nc, double *Y)
{
    const int n = op.n;
    for (int c0 = 0; c0 < nc; c0 += 128) {
        const int w = std::min(128, nc - c0);
        // scratch: colmax (w u64) | colscale (w) | Z hi/lo (np*w halves each) | partials
        const int wgx = (op.nrb + 7) / 8;
        int splits = std::max(1, 1024 / std::max(wgx, 1));
        splits = std::min(splits, op.nkb);
        const int kbs = (op.nkb + splits - 1) / splits;
        splits = (op.nkb + kbs - 1) / kbs;
        const size_t zwords = (size_t)op.np * w;
        const size_t bytes = 2 * (size_t)w * 8 + 2 * zwords * 2 + (size_t)splits * n * w * 4 + 1024;
        char *base = nullptr;
        IPT_STCHK(context_scratch(ctx, SLOT_LOWP, bytes, (void **)&base));
        unsigned long long *colmax = reinterpret_cast<unsigned long long *>(base);
        double *colscale = reinterpret_cast<double *>(base + (size_t)w * 8);
        unsigned short *zhi = reinterpret_cast<unsigned short *>(base + 2 * (size_t)w * 8);
        unsigned short *zlo = zhi + zwords;
        float *P = reinterpret_cast<float *>(base + ((2 * (size_t)w * 8 + 2 * zwords * 2 + 255) / 256) * 256);
        const double *Bc = B + (size_t)c0 * ld;
        const double *Bac = Ba + (size_t)c0 * ld;
        IPT_HIPCHK(hipMemsetAsync(colmax, 0, (size_t)w * 8, ctx->stream));
        hipLaunchKernelGGL(k_delta_colmax, dim3(std::min(8, (n + 255) / 256), w), dim3(256), 0, ctx->stream, Bc, Bac,
                           ld, n, w, colmax);
        IPT_HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_delta_convert, dim3(blocks_for((long)op.np * w, 256, 16384)), dim3(256), 0, ctx->stream,
                           Bc, Bac, ld, n, w, op.np, op.mode, colmax, colscale, zhi, zlo);
        IPT_HIPCHK(hipGetLastError());
        const int NB = (w + 15) / 16;
        dim3 grid(wgx, 1, splits);
        const uint4 *Ehi = reinterpret_cast<const uint4 *>(op.hi.p);
        const uint4 *Elo = reinterpret_cast<const uint4 *>(op.lo.p);
        const uint4 *Zh = reinterpret_cast<const uint4 *>(zhi);
        const uint4 *Zl = reinterpret_cast<const uint4 *>(zlo);
        const ProfileToken prof = profile_begin(ctx, PROF_LOWP_GEMM);
        // prefetch depth (k blocks in flight per wavefront); tuning knob
        static const int depth = env_int("IPT_LOWP_DEPTH", 1);
        if (op.mode == LOWP_BF16X3) {
            launch_lowp<LOWP_BF16X3, 1>(NB, grid, ctx->stream, Ehi, Elo, Zh, Zl, P, n, w, op.nkb, op.nrb, op.np, kbs);
        } else if (op.mode == LOWP_FP16) {
            if (depth >= 4) {
                launch_lowp<LOWP_FP16, 4>(NB, grid, ctx->stream, Ehi, Elo, Zh, Zl, P, n, w, op.nkb, op.nrb, op.np, kbs);
            } else if (depth >= 2) {
                launch_lowp<LOWP_FP16, 2>(NB, grid, ctx->stream, Ehi, Elo, Zh, Zl, P, n, w, op.nkb, op.nrb, op.np, kbs);
            } else {
                launch_lowp<LOWP_FP16, 1>(NB, grid, ctx->stream, Ehi, Elo, Zh, Zl, P, n, w, op.nkb, op.nrb, op.np, kbs);
            }
        } else {
            if (depth >= 2) {
                launch_lowp<LOWP_BF16, 2>(NB, grid, ctx->stream, Ehi, Elo, Zh, Zl, P, n, w, op.nkb, op.nrb, op.np, kbs);
            } else {
                launch_lowp<LOWP_BF16, 1>(NB, grid, ctx->stream, Ehi, Elo, Zh, Zl, P, n, w, op.nkb, op.nrb, op.np, kbs);
            }
        }
        profile_end(ctx, prof);
        IPT_HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_lowp_epilogue, dim3(blocks_for((long)n * w, 256, 16384)), dim3(256), 0, ctx->stream, P,
                           splits, n, w, Ya + (size_t)c0 * ld, Bc, Bac, ld, d, colscale, 1.0 / (double)op.scale,
                           Y + (size_t)c0 * ld);
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
        d_Y == nullptr || n <= 0 || nc <= 0 || mode < 0 || mode > 2) {
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
