// Dense FP64 product Y = H * Z for the IPT iteration on gfx950.
//
// Replaces the cublasDgemm(N,N,n,k,n) call sites of IPT/src/ipt_cuda.cu:1649,
// 1888 with a skinny-N MFMA kernel: the k (<= 128 per pass) columns of Z
// stay resident in LDS per K step while each wavefront streams 16 rows of
// H straight from HBM into the A operand of v_mfma_f64_16x16x4f64.
//
// Work split: a workgroup (4 waves) owns 64 rows x 16*NB columns and one
// K-chunk; K is split so the grid has >= ~1024 workgroups (8 XCDs x 32
// CUs x 4); split-K partials are summed in a fixed order by a second
// kernel, so results are deterministic.
#include "ipt_internal.h"

namespace ipt {

typedef double v4d __attribute__((ext_vector_type(4)));

template <int NB>
__global__ __launch_bounds__(256) void k_gemm_f64(const double *__restrict__ H, long ldh,
                                                  const double *__restrict__ Z, long ldz,
                                                  double *__restrict__ Y, long ldy,
                                                  long split_stride, int n, int ncols,
                                                  long kchunk)
{
    constexpr int BN = 16 * NB;
    constexpr int BK = 16;
    __shared__ double zs[2][BK][BN + 1];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int kq = lane >> 4;          // k offset inside a 4-wide MFMA step
    const int li = lane & 15;
    const long rowbase = (long)blockIdx.x * 64 + wave * 16;
    const long row = rowbase + li;
    const int col0 = blockIdx.y * BN;
    const long kbeg = (long)blockIdx.z * kchunk;
    long kend = kbeg + kchunk;
    if (kend > n) {
        kend = n;
    }
    const bool rowok = row < n;

    v4d acc[NB];
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) {
        acc[cb] = (v4d){0.0, 0.0, 0.0, 0.0};
    }

    double zr[NB];
    double ar[4];
    const int nsteps = kend > kbeg ? (int)((kend - kbeg + BK - 1) / BK) : 0;

    auto load_z = [&](long k0) {
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int e = tid + 256 * i;
            const int kk = e & (BK - 1);
            const int cc = e >> 4;
            const long kg = k0 + kk;
            const int cg = col0 + cc;
            zr[i] = (kg < kend && cg < ncols) ? Z[kg + (long)cg * ldz] : 0.0;
        }
    };
    auto store_z = [&](int buf) {
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int e = tid + 256 * i;
            zs[buf][e & (BK - 1)][e >> 4] = zr[i];
        }
    };
    auto load_a = [&](long k0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const long kg = k0 + 4 * q + kq;
            ar[q] = (rowok && kg < kend) ? H[row + kg * ldh] : 0.0;
        }
    };

    if (nsteps > 0) {
        load_z(kbeg);
        store_z(0);
        load_a(kbeg);
    }
    __syncthreads();
    for (int s = 0; s < nsteps; ++s) {
        const int buf = s & 1;
        double a_cur[4] = {ar[0], ar[1], ar[2], ar[3]};
        const bool more = s + 1 < nsteps;
        if (more) {
            const long kn = kbeg + (long)(s + 1) * BK;
            load_z(kn);
            load_a(kn);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
            for (int cb = 0; cb < NB; ++cb) {
                const double b = zs[buf][4 * q + kq][16 * cb + li];
                acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_cur[q], b, acc[cb], 0, 0, 0);
            }
        }
        if (more) {
            store_z(buf ^ 1);
        }
        __syncthreads();
    }

    // D layout of v_mfma_f64_16x16x4f64: lane l, element i -> row 4*i + l/16, col l%16.
    double *Yp = Y + (long)blockIdx.z * split_stride;
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) {
        const int c = col0 + 16 * cb + li;
        if (c < ncols) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const long r = rowbase + 4 * i + kq;
                if (r < n) {
                    Yp[r + (long)c * ldy] = acc[cb][i];
                }
            }
        }
    }
}

// Y[r + c*ldy] = sum_s P[s*stride + r + c*n]  (fixed order)
__global__ void k_splitk_reduce(const double *__restrict__ P, long stride, int splits, int n,
                                int ncols, double *__restrict__ Y, long ldy)
{
    const long total = (long)n * ncols;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long)gridDim.x * blockDim.x) {
        const long r = idx % n;
        const long c = idx / n;
        double s = 0.0;
        for (int z = 0; z < splits; ++z) {
            s += P[z * stride + idx];
        }
        Y[r + c * ldy] = s;
    }
}

template <int NB>
static void launch_gemm(dim3 grid, hipStream_t st, const double *H, long ldh, const double *Z,
                        long ldz, double *Y, long ldy, long split_stride, int n, int ncols,
                        long kchunk)
{
    hipLaunchKernelGGL(k_gemm_f64<NB>, grid, dim3(256), 0, st, H, ldh, Z, ldz, Y, ldy,
                       split_stride, n, ncols, kchunk);
}

static int gemm_f64_pass(Context *ctx, const double *H, long ldh, const double *Z, long ldz,
                         double *Y, long ldy, int n, int ncols)
{
    const int NB = (ncols + 15) / 16;  // 1..8
    const int gx = (n + 63) / 64;
    int splits = (1024 + gx - 1) / gx;
    const int max_splits = (n + 255) / 256;
    if (splits > max_splits) {
        splits = max_splits;
    }
    if (splits < 1) {
        splits = 1;
    }
    long kchunk = ((long)n + splits - 1) / splits;
    kchunk = (kchunk + 15) / 16 * 16;
    splits = (int)(((long)n + kchunk - 1) / kchunk);
    dim3 grid(gx, 1, splits);

    double *out = Y;
    long out_ld = ldy;
    long stride = 0;
    double *part = nullptr;
    if (splits > 1) {
        IPT_STCHK(context_scratch(ctx, SLOT_GEMM_SPLITK, (size_t)splits * n * ncols * sizeof(double),
                                  (void **)&part));
        out = part;
        out_ld = n;
        stride = (long)n * ncols;
    }
    const ProfileToken prof = profile_begin(ctx, PROF_GEMM_F64);
    switch (NB) {
    case 1: launch_gemm<1>(grid, ctx->stream, H, ldh, Z, ldz, out, out_ld, stride, n, ncols, kchunk); break;
    case 2: launch_gemm<2>(grid, ctx->stream, H, ldh, Z, ldz, out, out_ld, stride, n, ncols, kchunk); break;
    case 3: launch_gemm<3>(grid, ctx->stream, H, ldh, Z, ldz, out, out_ld, stride, n, ncols, kchunk); break;
    case 4: launch_gemm<4>(grid, ctx->stream, H, ldh, Z, ldz, out, out_ld, stride, n, ncols, kchunk); break;
    case 5: launch_gemm<5>(grid, ctx->stream, H, ldh, Z, ldz, out, out_ld, stride, n, ncols, kchunk); break;
    case 6: launch_gemm<6>(grid, ctx->stream, H, ldh, Z, ldz, out, out_ld, stride, n, ncols, kchunk); break;
    case 7: launch_gemm<7>(grid, ctx->stream, H, ldh, Z, ldz, out, out_ld, stride, n, ncols, kchunk); break;
    default: launch_gemm<8>(grid, ctx->stream, H, ldh, Z, ldz, out, out_ld, stride, n, ncols, kchunk); break;
    }
    profile_end(ctx, prof, ncols);
    IPT_HIPCHK(hipGetLastError());
    if (splits > 1) {
        const long total = (long)n * ncols;
        int blocks = (int)((total + 255) / 256);
        if (blocks > 8192) {
            blocks = 8192;
        }
        hipLaunchKernelGGL(k_splitk_reduce, dim3(blocks), dim3(256), 0, ctx->stream, part,
                           stride, splits, n, ncols, Y, ldy);
        IPT_HIPCHK(hipGetLastError());
    }
    return IPT_CUDA_SUCCESS;
}

int gemm_f64(Context *ctx, const double *H, long ldh, const double *Z, long ldz, double *Y,
             long ldy, int n, int ncols)
{
    if (n <= 0 || ncols <= 0) {
        return IPT_CUDA_SUCCESS;
    }
    for (int c0 = 0; c0 < ncols; c0 += 128) {
        const int w = ncols - c0 < 128 ? ncols - c0 : 128;
        IPT_STCHK(gemm_f64_pass(ctx, H, ldh, Z + (long)c0 * ldz, ldz, Y + (long)c0 * ldy, ldy, n, w));
    }
    return IPT_CUDA_SUCCESS;
}

}  // namespace ipt
