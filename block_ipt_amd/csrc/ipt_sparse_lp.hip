// Low-precision (FP32) correction product of the sparse block path:
//     Y = Y_a + diag(d) (B - B_a) + E (B - B_a),   E = A - diag(A),
// with E (B - B_a) from FP32 values and an FP32 row-major copy of the delta
// (FP32 accumulation), the diagonal and anchor terms in FP64 -- the anchored
// scheme of the dense path (ipt_gemm_lp.hip) for CSR input.  The product's
// time is set by the gathered rows of the iterate (ipt_kernels.hip,
// k_spmm_rowmajor): FP32 rows halve the gathered bytes and the compulsory
// value bytes.
#include <algorithm>

#include "ipt_internal.h"
#include "ipt_sparse_lp.h"

namespace ipt {

namespace {

__global__ void k_to_f32(const double *__restrict__ x, float *__restrict__ y, long count)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x) {
        y[i] = (float)x[i];
    }
}

constexpr int kTr = 32;   // rows per tile
constexpr int kTc = 128;  // columns per tile

// D (row-major FP32, ld ldd >= w) = B - B_a (column-major FP64, ld ldb) for
// columns [0, w), through an LDS tile so both sides are read / written
// contiguously.
__global__ __launch_bounds__(256) void k_delta_rows_f32(const double *__restrict__ B, const double *__restrict__ Ba,
                                                        long ldb, int n, int w, float *__restrict__ D, long ldd)
{
    __shared__ float t[kTr][kTc + 1];
    const int r0 = blockIdx.x * kTr;
    for (int e = threadIdx.x; e < kTr * kTc; e += 256) {
        const int r = e % kTr, c = e / kTr;
        float v = 0.0f;
        if (r0 + r < n && c < w) {
            const long i = (long)c * ldb + r0 + r;
            v = (float)(B[i] - Ba[i]);
        }
        t[r][c] = v;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kTr * kTc; e += 256) {
        const int c = e % kTc, r = e / kTc;
        if (r0 + r < n && c < w) {
            D[(long)(r0 + r) * ldd + c] = t[r][c];
        }
    }
}

// Z (row-major FP32, ld w) = E D (D row-major, ld ldx): a wavefront per
// row, S lanes per column
// group, the row's (column, value) pairs read 64 at a time and handed out
// by shuffles (as k_spmm_rowmajor); the diagonal entry is skipped.  Row r's
// entries are [rs[r], re[r]) (the whole row, or one column slice; with
// accumulate the slice's products are added to Z).
template <int S>
__global__ __launch_bounds__(256) void k_spmm_offdiag_f32(const int *__restrict__ rs, const int *__restrict__ re,
                                                          const int *__restrict__ ci,
                                                          const float *__restrict__ vals, int n,
                                                          const float *__restrict__ X, long ldx, int w,
                                                          float *__restrict__ Z, int accumulate)
{
    constexpr int G = 64 / S;
    constexpr int STEPS = 64 / G;
    constexpr int U = STEPS < 8 ? STEPS : 8;
    const int lane = threadIdx.x & 63;
    const int g = lane / S;
    const int s = lane % S;
    const long wave_global = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long waves = ((long)gridDim.x * blockDim.x) >> 6;
    for (long row = wave_global; row < n; row += waves) {
        const int pb = rs[row], pe = re[row];
        for (int c0 = 0; c0 < w; c0 += S) {
            const int col = c0 + s;
            const bool active = col < w;
            const float *xc = X + (active ? col : 0);
            float acc = 0.0f;
            for (int base = pb; base < pe; base += 64) {
                const int cnt = min(64, pe - base);
                // padding and the diagonal: column 0 (a valid row), value 0
                int myc = 0;
                float myv = 0.0f;
                if (lane < cnt) {
                    myc = ci[base + lane];
                    myv = myc == row ? 0.0f : vals[base + lane];
                }
                const int steps = ((cnt + G * U - 1) / (G * U)) * U;
                for (int j0 = 0; j0 < steps; j0 += U) {
                    int c[U];
                    float v[U], x[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        c[u] = __shfl(myc, (j0 + u) * G + g, 64);
                        v[u] = __shfl(myv, (j0 + u) * G + g, 64);
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        x[u] = xc[(long)c[u] * ldx];
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        acc = fmaf(v[u], x[u], acc);
                    }
                }
            }
#pragma unroll
            for (int off = S; off < 64; off <<= 1) {
                acc += __shfl_xor(acc, off, 64);
            }
            if (g == 0 && active) {
                if (accumulate) {
                    Z[row * w + col] += acc;
                } else {
                    Z[row * w + col] = acc;
                }
            }
        }
    }
}

// Y (column-major FP64, ld ldy) = Y_a + d (B - B_a) + Z, Z row-major FP32
// (ld w), through an LDS tile.
__global__ __launch_bounds__(256) void k_delta_epilogue(const float *__restrict__ Z, int w, int n,
                                                        const double *__restrict__ Ya, const double *__restrict__ B,
                                                        const double *__restrict__ Ba, long ld,
                                                        const double *__restrict__ d, double *__restrict__ Y)
{
    __shared__ float t[kTr][kTc + 1];
    const int r0 = blockIdx.x * kTr;
    for (int e = threadIdx.x; e < kTr * kTc; e += 256) {
        const int c = e % kTc, r = e / kTc;
        t[r][c] = (r0 + r < n && c < w) ? Z[(long)(r0 + r) * w + c] : 0.0f;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kTr * kTc; e += 256) {
        const int r = e % kTr, c = e / kTr;
        if (r0 + r < n && c < w) {
            const long i = (long)c * ld + r0 + r;
            Y[i] = Ya[i] + d[r0 + r] * (B[i] - Ba[i]) + (double)t[r][c];
        }
    }
}

template <int S>
void launch_spmm_f32(const ProfileToken &prof, bool first, bool last, dim3 g, hipStream_t st, const int *rs,
                     const int *re, const int *ci, const float *v, int n, const float *X, long ldx, int w,
                     float *Z, int accumulate)
{
    launch_timed(prof, first, last, k_spmm_offdiag_f32<S>, g, dim3(256), 0, st, rs, re, ci, v, n, X, ldx, w, Z,
                 accumulate);
}

// one pass of the product; first / last: the pass opens / closes the timed span
void spmm_f32_pass(const ProfileToken &prof, bool first, bool last, dim3 g, hipStream_t st, const int *rs,
                   const int *re, const int *ci, const float *v, int n, const float *X, long ldx, int w,
                   float *Z, int accumulate)
{
    if (w <= 1) {
        launch_spmm_f32<1>(prof, first, last, g, st, rs, re, ci, v, n, X, ldx, w, Z, accumulate);
    } else if (w <= 2) {
        launch_spmm_f32<2>(prof, first, last, g, st, rs, re, ci, v, n, X, ldx, w, Z, accumulate);
    } else if (w <= 4) {
        launch_spmm_f32<4>(prof, first, last, g, st, rs, re, ci, v, n, X, ldx, w, Z, accumulate);
    } else if (w <= 8) {
        launch_spmm_f32<8>(prof, first, last, g, st, rs, re, ci, v, n, X, ldx, w, Z, accumulate);
    } else if (w <= 16) {
        launch_spmm_f32<16>(prof, first, last, g, st, rs, re, ci, v, n, X, ldx, w, Z, accumulate);
    } else if (w <= 32) {
        launch_spmm_f32<32>(prof, first, last, g, st, rs, re, ci, v, n, X, ldx, w, Z, accumulate);
    } else {
        launch_spmm_f32<64>(prof, first, last, g, st, rs, re, ci, v, n, X, ldx, w, Z, accumulate);
    }
}

}  // namespace

int sparse_lowp_create(Context *ctx, const int *rp, const int *ci, const double *vals, int n, long long nnz,
                       SparseLowp *op)
{
    op->rp = rp;
    op->ci = ci;
    op->n = n;
    IPT_STCHK(op->vals.alloc((size_t)std::max(nnz, 1LL)));
    const long blocks = std::min((nnz + 255) / 256, 65536LL);
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)std::max(1L, blocks)), dim3(256), 0, ctx->stream, vals, op->vals.p,
                       (long)nnz);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

int sparse_lowp_apply_delta(Context *ctx, const SparseLowp &op, const double *B, const double *Ba, const double *Ya,
                            long ld, const double *d, int nc, double *Y)
{
    const int n = op.n;
    const int chunk = std::min(nc, kTc);
    // D rows at csr_row_stride(w, 4) <= round_up(w, 32) floats
    const size_t dcols = (size_t)(chunk + 31) / 32 * 32;
    float *D = nullptr;
    IPT_STCHK(context_scratch(ctx, SLOT_LOWP, (dcols + chunk) * n * sizeof(float), (void **)&D));
    float *Z = D + (size_t)n * dcols;
    const dim3 tiles((unsigned)((n + kTr - 1) / kTr));
    const long wblocks = std::min(((long)n * 64 + 255) / 256, 65536L);
    const dim3 g((unsigned)wblocks);
    for (int c0 = 0; c0 < nc; c0 += chunk) {
        const int w = std::min(chunk, nc - c0);
        const size_t o = (size_t)c0 * ld;
        const long ldd = csr_row_stride(w, 4);
        hipLaunchKernelGGL(k_delta_rows_f32, tiles, dim3(256), 0, ctx->stream, B + o, Ba + o, ld, n, w, D, ldd);
        IPT_HIPCHK(hipGetLastError());
        const ProfileToken prof = profile_begin_launches(PROF_SPMM_F32);
        // FP32 rows of X: the pass count of an FP64 product half as wide
        const int ns = op.slices != nullptr ? csr_slices_for(*op.slices, (w + 1) / 2) : 1;
        if (ns > 1) {
            const int *sp = op.slices->split.p;
            const long stride = (long)op.slices->n;
            const int step = op.slices->nslices / ns;
            for (int b = 0; b < ns; ++b) {
                spmm_f32_pass(prof, b == 0, b == ns - 1, g, ctx->stream, sp + (long)(b * step) * stride,
                              sp + (long)((b + 1) * step) * stride, op.ci, op.vals.p, n, D, ldd, w, Z, b > 0);
            }
        } else {
            spmm_f32_pass(prof, true, true, g, ctx->stream, op.rp, op.rp + 1, op.ci, op.vals.p, n, D, ldd, w, Z, 0);
        }
        profile_end(ctx, prof, w);
        IPT_HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_delta_epilogue, tiles, dim3(256), 0, ctx->stream, Z, w, n, Ya + o, B + o, Ba + o, ld, d,
                           Y + o);
        IPT_HIPCHK(hipGetLastError());
    }
    return IPT_CUDA_SUCCESS;
}

}  // namespace ipt
