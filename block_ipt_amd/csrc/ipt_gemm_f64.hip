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
#include <algorithm>
#include <map>
#include <mutex>

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
static void launch_gemm(const ProfileToken &prof, dim3 grid, hipStream_t st, const double *H, long ldh,
                        const double *Z, long ldz, double *Y, long ldy, long split_stride, int n, int ncols,
                        long kchunk)
{
    launch_timed(prof, true, true, k_gemm_f64<NB>, grid, dim3(256), 0, st, H, ldh, Z, ldz, Y, ldy, split_stride, n,
                 ncols, kchunk);
}

// Workgroups of `kernel` (256 threads, static LDS only) resident at once on
// the current device: compute units x workgroups per unit at the kernel's
// register and LDS footprint (cached per kernel).
static int resident_workgroups(const void *kernel)
{
    static std::mutex mu;
    static std::map<const void *, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    const auto it = cache.find(kernel);
    if (it != cache.end()) {
        return it->second;
    }
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 256, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 1024;
    }
    const int v = std::max(1, cus * per);
    cache[kernel] = v;
    return v;
}

// k splits of a product with gx row tiles: about 1024 workgroups, rounded to
// whole rounds of the resident ones (a last round that fills a third of
// the machine would run at a third of its rate), each k chunk at least 256.
static int product_splits(int gx, int resident, int n)
{
    const int rounds = std::max(1, (1024 + resident / 2) / resident);
    int s = (rounds * resident + gx / 2) / gx;
    s = std::min(s, std::max(1, (n + 255) / 256));
    return std::max(1, s);
}

template <int NB>
__global__ __launch_bounds__(256) void k_gemm_f64_packed(const double2 *__restrict__ PA, int nkp,
                                                         const double *__restrict__ Z, long ldz,
                                                         double *__restrict__ Y, long ldy, long split_stride, int n,
                                                         int ncols, long kchunk);

template <int NB>
static const void *gemm_kernel(bool packed)
{
    return packed ? reinterpret_cast<const void *>(k_gemm_f64_packed<NB>)
                  : reinterpret_cast<const void *>(k_gemm_f64<NB>);
}

static const void *gemm_kernel_nb(int NB, bool packed)
{
    switch (NB) {
    case 1: return gemm_kernel<1>(packed);
    case 2: return gemm_kernel<2>(packed);
    case 3: return gemm_kernel<3>(packed);
    case 4: return gemm_kernel<4>(packed);
    case 5: return gemm_kernel<5>(packed);
    case 6: return gemm_kernel<6>(packed);
    case 7: return gemm_kernel<7>(packed);
    default: return gemm_kernel<8>(packed);
    }
}

static int gemm_f64_pass(Context *ctx, const double *H, long ldh, const double *Z, long ldz,
                         double *Y, long ldy, int n, int ncols)
{
    const int NB = (ncols + 15) / 16;  // 1..8
    const int gx = (n + 63) / 64;
    int splits = product_splits(gx, resident_workgroups(gemm_kernel_nb(NB, false)), n);
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
    const ProfileToken prof = profile_begin_launches(PROF_GEMM_F64);
    switch (NB) {
    case 1: launch_gemm<1>(prof, grid, ctx->stream, H, ldh, Z, ldz, out, out_ld, stride, n, ncols, kchunk); break;
    case 2: launch_gemm<2>(prof, grid, ctx->stream, H, ldh, Z, ldz, out, out_ld, stride, n, ncols, kchunk); break;
    case 3: launch_gemm<3>(prof, grid, ctx->stream, H, ldh, Z, ldz, out, out_ld, stride, n, ncols, kchunk); break;
    case 4: launch_gemm<4>(prof, grid, ctx->stream, H, ldh, Z, ldz, out, out_ld, stride, n, ncols, kchunk); break;
    case 5: launch_gemm<5>(prof, grid, ctx->stream, H, ldh, Z, ldz, out, out_ld, stride, n, ncols, kchunk); break;
    case 6: launch_gemm<6>(prof, grid, ctx->stream, H, ldh, Z, ldz, out, out_ld, stride, n, ncols, kchunk); break;
    case 7: launch_gemm<7>(prof, grid, ctx->stream, H, ldh, Z, ldz, out, out_ld, stride, n, ncols, kchunk); break;
    default: launch_gemm<8>(prof, grid, ctx->stream, H, ldh, Z, ldz, out, out_ld, stride, n, ncols, kchunk); break;
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

// ------------------------------------------------------------------ packed H

// H in MFMA A-operand order: entry (row block rb, k pair kp) is 64 lanes x
// 2 doubles, lane l holding H(16 rb + l%16, 8 kp + l/16) and
// H(16 rb + l%16, 8 kp + 4 + l/16), the A operands of two consecutive
// v_mfma_f64_16x16x4f64 steps; entries k pair fastest, so a wavefront's
// stream of its row block is one contiguous run of 1 KiB loads (column-major
// H gives 128-byte pieces of columns 8 n bytes apart instead).
__global__ __launch_bounds__(256) void k_pack_f64(const double *__restrict__ H, long ld, int n, int nkp,
                                                  double2 *__restrict__ P)
{
    const long total = (long)((n + 15) / 16) * nkp * 64;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int l = (int)(t & 63);
        const long e = t >> 6;
        const long rb = e / nkp, kp = e % nkp;
        const long r = 16 * rb + (l & 15);
        const long k0 = 8 * kp + (l >> 4), k1 = k0 + 4;
        double2 v;
        v.x = (r < n && k0 < n) ? H[r + k0 * ld] : 0.0;
        v.y = (r < n && k1 < n) ? H[r + k1 * ld] : 0.0;
        P[t] = v;
    }
}

// Y = H Z from packed H: a workgroup (4 wavefronts) owns 128 rows (two
// 16-row blocks per wavefront, so each Z value read from LDS feeds two
// MFMAs) x 16NB columns and one k chunk; Z staged through LDS 16 k at a
// time, double-buffered; split-K partials as in k_gemm_f64.
template <int NB>
__global__ __launch_bounds__(256) void k_gemm_f64_packed(const double2 *__restrict__ PA, int nkp,
                                                         const double *__restrict__ Z, long ldz,
                                                         double *__restrict__ Y, long ldy, long split_stride, int n,
                                                         int ncols, long kchunk)
{
    constexpr int BN = 16 * NB;
    constexpr int BK = 16;
    constexpr int RPW = 2;
    constexpr int NP = BK / 8;  // k pairs per stage
    __shared__ double zs[2][BK][BN + 1];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int kq = lane >> 4;
    const int li = lane & 15;
    const long rb0 = (long)blockIdx.x * 4 * RPW + wave * RPW;  // first row block of this wavefront
    const long kbeg = (long)blockIdx.z * kchunk;               // kchunk is a multiple of BK
    long kend = kbeg + kchunk;
    if (kend > n) {
        kend = n;
    }
    const long nrb = (n + 15) / 16;
    v4d acc[RPW][NB];
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
#pragma unroll
        for (int cb = 0; cb < NB; ++cb) {
            acc[r][cb] = (v4d){0.0, 0.0, 0.0, 0.0};
        }
    }
    double zr[NB];
    double2 ar[RPW][NP];
    const int nsteps = kend > kbeg ? (int)((kend - kbeg + BK - 1) / BK) : 0;
    auto load_z = [&](long k0) {
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int e = tid + 256 * i;
            const long kg = k0 + (e & (BK - 1));
            const int cg = e >> 4;
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
    // k beyond kend reads the next chunk's (or the zero padding's) values,
    // which meet zero rows of the Z tile
    auto load_a = [&](long k0) {
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const long rb = rb0 + r < nrb ? rb0 + r : nrb - 1;
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                ar[r][p] = PA[(rb * nkp + k0 / 8 + p) * 64 + lane];
            }
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
        double2 a_cur[RPW][NP];
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                a_cur[r][p] = ar[r][p];
            }
        }
        const bool more = s + 1 < nsteps;
        if (more) {
            const long kn = kbeg + (long)(s + 1) * BK;
            load_z(kn);
            load_a(kn);
        }
#pragma unroll
        for (int q = 0; q < BK / 4; ++q) {
#pragma unroll
            for (int cb = 0; cb < NB; ++cb) {
                const double b = zs[buf][4 * q + kq][16 * cb + li];
#pragma unroll
                for (int r = 0; r < RPW; ++r) {
                    const double av = (q & 1) ? a_cur[r][q >> 1].y : a_cur[r][q >> 1].x;
                    acc[r][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b, acc[r][cb], 0, 0, 0);
                }
            }
        }
        if (more) {
            store_z(buf ^ 1);
        }
        __syncthreads();
    }
    double *Yp = Y + (long)blockIdx.z * split_stride;
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
#pragma unroll
        for (int cb = 0; cb < NB; ++cb) {
            const int c = 16 * cb + li;
            if (c < ncols) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const long row = (rb0 + r) * 16 + 4 * i + kq;
                    if (row < n) {
                        Yp[row + (long)c * ldy] = acc[r][cb][i];
                    }
                }
            }
        }
    }
}

int f64_packed_create(Context *ctx, const double *H, long ld, int n, F64Packed *op)
{
    op->n = n;
    // k extent padded to a multiple of 16 (zeros): a 16-wide stage reads
    // two whole k pairs
    op->nkp = (n + 15) / 16 * 2;
    const long entries = (long)((n + 15) / 16) * op->nkp * 64;
    IPT_STCHK(op->p.alloc((size_t)entries));
    int blocks = (int)std::min<long>((entries + 255) / 256, 65536);
    hipLaunchKernelGGL(k_pack_f64, dim3(blocks), dim3(256), 0, ctx->stream, H, ld, n, op->nkp, op->p.p);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

template <int NB>
static void launch_packed(const ProfileToken &prof, dim3 grid, hipStream_t st, const F64Packed &op, const double *Z,
                          long ldz, double *Y, long ldy, long split_stride, int ncols, long kchunk)
{
    launch_timed(prof, true, true, k_gemm_f64_packed<NB>, grid, dim3(256), 0, st, (const double2 *)op.p.p, op.nkp, Z,
                 ldz, Y, ldy, split_stride, op.n, ncols, kchunk);
}

static int gemm_packed_pass(Context *ctx, const F64Packed &op, const double *Z, long ldz, double *Y, long ldy,
                            int ncols)
{
    const int n = op.n;
    const int NB = (ncols + 15) / 16;
    const int gx = (n + 127) / 128;
    int splits = product_splits(gx, resident_workgroups(gemm_kernel_nb(NB, true)), n);
    long kchunk = ((long)n + splits - 1) / splits;
    kchunk = (kchunk + 15) / 16 * 16;
    splits = (int)(((long)n + kchunk - 1) / kchunk);
    dim3 grid(gx, 1, splits);
    double *out = Y;
    long out_ld = ldy, stride = 0;
    double *part = nullptr;
    if (splits > 1) {
        IPT_STCHK(context_scratch(ctx, SLOT_GEMM_SPLITK, (size_t)splits * n * ncols * sizeof(double),
                                  (void **)&part));
        out = part;
        out_ld = n;
        stride = (long)n * ncols;
    }
    const ProfileToken prof = profile_begin_launches(PROF_GEMM_F64);
    switch (NB) {
    case 1: launch_packed<1>(prof, grid, ctx->stream, op, Z, ldz, out, out_ld, stride, ncols, kchunk); break;
    case 2: launch_packed<2>(prof, grid, ctx->stream, op, Z, ldz, out, out_ld, stride, ncols, kchunk); break;
    case 3: launch_packed<3>(prof, grid, ctx->stream, op, Z, ldz, out, out_ld, stride, ncols, kchunk); break;
    case 4: launch_packed<4>(prof, grid, ctx->stream, op, Z, ldz, out, out_ld, stride, ncols, kchunk); break;
    case 5: launch_packed<5>(prof, grid, ctx->stream, op, Z, ldz, out, out_ld, stride, ncols, kchunk); break;
    case 6: launch_packed<6>(prof, grid, ctx->stream, op, Z, ldz, out, out_ld, stride, ncols, kchunk); break;
    case 7: launch_packed<7>(prof, grid, ctx->stream, op, Z, ldz, out, out_ld, stride, ncols, kchunk); break;
    default: launch_packed<8>(prof, grid, ctx->stream, op, Z, ldz, out, out_ld, stride, ncols, kchunk); break;
    }
    profile_end(ctx, prof, ncols);
    IPT_HIPCHK(hipGetLastError());
    if (splits > 1) {
        const long total = (long)n * ncols;
        const int blocks = (int)std::min<long>((total + 255) / 256, 8192);
        hipLaunchKernelGGL(k_splitk_reduce, dim3(blocks), dim3(256), 0, ctx->stream, part, stride, splits, n, ncols,
                           Y, ldy);
        IPT_HIPCHK(hipGetLastError());
    }
    return IPT_CUDA_SUCCESS;
}

int gemm_f64_packed(Context *ctx, const F64Packed &op, const double *Z, long ldz, double *Y, long ldy, int ncols)
{
    for (int c0 = 0; c0 < ncols; c0 += 128) {
        const int w = std::min(128, ncols - c0);
        IPT_STCHK(gemm_packed_pass(ctx, op, Z + (long)c0 * ldz, ldz, Y + (long)c0 * ldy, ldy, w));
    }
    return IPT_CUDA_SUCCESS;
}

}  // namespace ipt

extern "C" int ipt_hip_dense_product(const double *d_H, int n, const double *d_Z, int nc, double *d_Y, int packed,
                                     void *handle)
{
    using namespace ipt;
    if (d_H == nullptr || d_Z == nullptr || d_Y == nullptr || n <= 0 || nc <= 0) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    if (packed) {
        F64Packed op;
        IPT_STCHK(f64_packed_create(ctx, d_H, n, n, &op));
        IPT_STCHK(gemm_f64_packed(ctx, op, d_Z, n, d_Y, n, nc));
    } else {
        IPT_STCHK(gemm_f64(ctx, d_H, n, d_Z, n, d_Y, n, n, nc));
    }
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return IPT_CUDA_SUCCESS;
}
