// Learned Davidson preconditioner: a per-component MLP 3 -> 64 -> 64 -> 1
// (GELU) that replaces the diagonal (Jacobi) correction
//     t_i = r_i / clip(D_i - theta)          (IPT/src/ipt_cuda.cu:2449-2465)
// by
//     t_i = (r_i / s_i) * y_i,   y_i = MLP(x_i),   s_i = rowabs_i (> 0),
//     x_i = (asinh((D_i - theta) / s_i),  r_i / max_j |r_j|,
//            s_i / (|D_i - theta| + s_i)),
// rowabs_i = sum_{j != i} |H_ij|.  The Jacobi correction is y = s / (D - theta);
// the network learns a better per-component map from residual to exact
// correction (trained on pairs from plain solves).
//
// Parameters, FP32, row-major: W1[64][3] at 0, b1[64] at 192, W2[64][64] at
// 256 (out, in), b2[64] at 4352, W3[64] at 4416, b3 at 4480; 4481 floats.
// Inference uses BF16 copies of W2 in MFMA operand order (FP32 accumulate);
// training and the gradient run in FP32.
#include <cmath>
#include <mutex>

#include "ipt_internal.h"

#include "../../include/ipt_hip_ext.h"

namespace ipt {

constexpr int MLP_W1 = 0, MLP_B1 = 192, MLP_W2 = 256, MLP_B2 = 4352, MLP_W3 = 4416, MLP_B3 = 4480;
constexpr int MLP_NPARAM = 4481;

typedef __bf16 mb8_t __attribute__((ext_vector_type(8)));
typedef float mf4_t __attribute__((ext_vector_type(4)));

// GELU, tanh form (Hendrycks & Gimpel 2016), and its derivative
__device__ inline float gelu(float x)
{
    const float c = 0.7978845608028654f;  // sqrt(2/pi)
    return 0.5f * x * (1.0f + tanhf(c * (x + 0.044715f * x * x * x)));
}

__device__ inline float gelu_grad(float x)
{
    const float c = 0.7978845608028654f;
    const float u = c * (x + 0.044715f * x * x * x);
    const float th = tanhf(u);
    return 0.5f * (1.0f + th) + 0.5f * x * (1.0f - th * th) * c * (1.0f + 3.0f * 0.044715f * x * x);
}

__device__ inline unsigned short f2bf(float x)
{
    unsigned int u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) {
        return (unsigned short)(u >> 16);
    }
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ inline float bf2f(unsigned short b) { return __uint_as_float(((unsigned int)b) << 16); }

// ------------------------------------------------------------------ features

// Per (component i, target q): x (3 floats), qf = r_i / s_i, for targets
// idx[q] of the Ritz values theta and residual columns R (n x ., ldr);
// rmax[q] = max_i |R_i,idx[q]| (precomputed).  Layout: feat[(q * n + i) * 3].
__global__ void k_mlp_features(const double *__restrict__ R, long ldr, const double *__restrict__ d,
                               const double *__restrict__ rowabs, const double *__restrict__ theta,
                               const int *__restrict__ idx, const double *__restrict__ rmax, int n, int nt,
                               float *__restrict__ feat, float *__restrict__ qf)
{
    const long total = (long)n * nt;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % n);
        const int q = (int)(t / n);
        const int j = idx[q];
        const double a = d[i] - theta[j];
        const double s = fmax(rowabs[i], 1e-300);
        const double r = R[i + (long)j * ldr];
        const double rm = rmax[q] > 0.0 ? rmax[q] : 1.0;
        feat[3 * t + 0] = (float)asinh(a / s);
        feat[3 * t + 1] = (float)(r / rm);
        feat[3 * t + 2] = (float)(s / (fabs(a) + s));
        qf[t] = (float)(r / s);
    }
}

// rmax[q] = max_i |R[i, idx[q]]|; one block per target
__global__ __launch_bounds__(256) void k_col_absmax(const double *__restrict__ R, long ldr,
                                                    const int *__restrict__ idx, int n, double *__restrict__ rmax)
{
    const int j = idx[blockIdx.x];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        m = fmax(m, fabs(R[i + (long)j * ldr]));
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
        rmax[blockIdx.x] = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
    }
}

// T[:, q] (ldt) = out[q * n + i] widened to FP64
__global__ void k_mlp_store(const float *__restrict__ out, int n, int nt, double *__restrict__ T, long ldt)
{
    const long total = (long)n * nt;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        T[(t % n) + (t / n) * ldt] = (double)out[t];
    }
}

// ------------------------------------------------------------------ inference

// One wavefront per 16 components.  Layer 1 is evaluated straight into the
// MFMA A-operand layout (lane l: component l%16, hidden units
// 32kb + 8(l/16) + j), layer 2 is 8 MFMAs 16x16x32 BF16 (FP32 accumulate)
// against W2 held as B fragments (column = output unit), layer 3 a 16-lane
// shuffle reduction.  out[c] = qf[c] * y_c (or y_c when qf is NULL).
__global__ __launch_bounds__(256) void k_mlp_infer(const float *__restrict__ params, const float *__restrict__ feat,
                                                   const float *__restrict__ qf, long count, float *__restrict__ out)
{
    __shared__ float w1[192], b1[64], b2[64], w3[64];
    __shared__ unsigned short w2b[64 * 64];
    for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) {
        w2b[e] = f2bf(params[MLP_W2 + e]);
    }
    for (int e = threadIdx.x; e < 192; e += blockDim.x) {
        w1[e] = params[MLP_W1 + e];
    }
    if (threadIdx.x < 64) {
        b1[threadIdx.x] = params[MLP_B1 + threadIdx.x];
        b2[threadIdx.x] = params[MLP_B2 + threadIdx.x];
        w3[threadIdx.x] = params[MLP_W3 + threadIdx.x];
    }
    __syncthreads();
    const float b3 = params[MLP_B3];
    const int lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    // B fragments: B[k][col] = W2[col][k], col = 16cb + li, k = 32kb + 8lk + j
    mb8_t bw[2][4];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            unsigned short v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                v[j] = w2b[(16 * cb + li) * 64 + 32 * kb + 8 * lk + j];
            }
            bw[kb][cb] = __builtin_bit_cast(mb8_t, v);
        }
    }
    const long waves = (long)gridDim.x * (blockDim.x >> 6);
    for (long tile = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); tile * 16 < count; tile += waves) {
        const long c0 = tile * 16;
        const long ci = c0 + li;
        float x0 = 0.f, x1 = 0.f, x2 = 0.f;
        if (ci < count) {
            x0 = feat[3 * ci];
            x1 = feat[3 * ci + 1];
            x2 = feat[3 * ci + 2];
        }
        mb8_t a[2];
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            unsigned short v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int h = 32 * kb + 8 * lk + j;
                const float z = w1[3 * h] * x0 + w1[3 * h + 1] * x1 + w1[3 * h + 2] * x2 + b1[h];
                v[j] = f2bf(gelu(z));
            }
            a[kb] = __builtin_bit_cast(mb8_t, v);
        }
        float part[4] = {0.f, 0.f, 0.f, 0.f};  // rows 4lk + v
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
            mf4_t acc = {0.f, 0.f, 0.f, 0.f};
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], bw[0][cb], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], bw[1][cb], acc, 0, 0, 0);
            const int col = 16 * cb + li;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                part[v] += w3[col] * gelu(acc[v] + b2[col]);
            }
        }
        // sum over the 16 lanes of a row group (same lk, all li)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                part[v] += __shfl_xor(part[v], o, 64);
            }
        }
        if (li < 4) {
            // lane (lk, li=v) writes component 4lk + v
            const float y = (li == 0 ? part[0] : li == 1 ? part[1] : li == 2 ? part[2] : part[3]) + b3;
            const long c = c0 + 4 * lk + li;
            if (c < count) {
                out[c] = qf != nullptr ? qf[c] * y : y;
            }
        }
    }
}

// ------------------------------------------------------------------ training

// Loss L = sum_s w_s (q_s y_s - t_s)^2 over samples s (x_s: feat[3s],
// q_s, t_s, w_s); gradient of L with respect to the parameters.  A block
// takes 64 samples at a time: the forward and backward activations of the
// tile live in LDS, every thread owns a strided set of parameters and adds
// its gradient over the tile's samples in a fixed order.  Per-block sums
// go to gpart[block][4481] (reduced in a fixed order by k_mlp_reduce).
constexpr int MLP_TILE = 64;

__global__ __launch_bounds__(256) void k_mlp_grad(const float *__restrict__ params, const float *__restrict__ feat,
                                                  const float *__restrict__ q, const float *__restrict__ tgt,
                                                  const float *__restrict__ wt, long count,
                                                  float *__restrict__ gpart, float *__restrict__ lpart)
{
    __shared__ float w2[64 * 64];
    __shared__ float gacc[MLP_NPARAM];  // entry p is owned by thread p % 256
    __shared__ float x[MLP_TILE][3];
    __shared__ float z1[MLP_TILE][65], h1[MLP_TILE][65], z2[MLP_TILE][65], d2[MLP_TILE][65], d1[MLP_TILE][65];
    __shared__ float dy[MLP_TILE], lsum[MLP_TILE];
    const int tid = threadIdx.x;
    for (int e = tid; e < 4096; e += 256) {
        w2[e] = params[MLP_W2 + e];
    }
    for (int p = tid; p < MLP_NPARAM; p += 256) {
        gacc[p] = 0.f;
    }
    float lacc = 0.f;
    __syncthreads();
    const long tiles = (count + MLP_TILE - 1) / MLP_TILE;
    for (long tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const long s0 = tile * MLP_TILE;
        if (tid < MLP_TILE) {
            const long s = s0 + tid;
            const bool ok = s < count;
            x[tid][0] = ok ? feat[3 * s] : 0.f;
            x[tid][1] = ok ? feat[3 * s + 1] : 0.f;
            x[tid][2] = ok ? feat[3 * s + 2] : 0.f;
        }
        __syncthreads();
        // layer 1
        for (int e = tid; e < MLP_TILE * 64; e += 256) {
            const int s = e >> 6, h = e & 63;
            const float z = params[MLP_W1 + 3 * h] * x[s][0] + params[MLP_W1 + 3 * h + 1] * x[s][1] +
                            params[MLP_W1 + 3 * h + 2] * x[s][2] + params[MLP_B1 + h];
            z1[s][h] = z;
            h1[s][h] = gelu(z);
        }
        __syncthreads();
        // layer 2
        for (int e = tid; e < MLP_TILE * 64; e += 256) {
            const int s = e >> 6, c = e & 63;
            float z = params[MLP_B2 + c];
            for (int k = 0; k < 64; ++k) {
                z += w2[c * 64 + k] * h1[s][k];
            }
            z2[s][c] = z;
        }
        __syncthreads();
        // output, loss and dL/dy (one thread per sample)
        if (tid < MLP_TILE) {
            const long s = s0 + tid;
            float y = params[MLP_B3];
            for (int c = 0; c < 64; ++c) {
                y += params[MLP_W3 + c] * gelu(z2[tid][c]);
            }
            float gy = 0.f, l = 0.f;
            if (s < count) {
                const float e = q[s] * y - tgt[s];
                l = wt[s] * e * e;
                gy = 2.f * wt[s] * e * q[s];
            }
            dy[tid] = gy;
            lsum[tid] = l;
        }
        __syncthreads();
        // delta2 = dy W3 gelu'(z2)
        for (int e = tid; e < MLP_TILE * 64; e += 256) {
            const int s = e >> 6, c = e & 63;
            d2[s][c] = dy[s] * params[MLP_W3 + c] * gelu_grad(z2[s][c]);
        }
        __syncthreads();
        // delta1 = (W2^T delta2) gelu'(z1)
        for (int e = tid; e < MLP_TILE * 64; e += 256) {
            const int s = e >> 6, k = e & 63;
            float acc = 0.f;
            for (int c = 0; c < 64; ++c) {
                acc += w2[c * 64 + k] * d2[s][c];
            }
            d1[s][k] = acc * gelu_grad(z1[s][k]);
        }
        __syncthreads();
        // parameter gradients over the tile (fixed sample order), each
        // entry updated only by its owner thread
        for (int p = tid; p < MLP_NPARAM; p += 256) {
            float acc = 0.f;
            if (p < MLP_B1) {  // W1[h][k]
                const int h = p / 3, k = p % 3;
                for (int s = 0; s < MLP_TILE; ++s) {
                    acc += d1[s][h] * x[s][k];
                }
            } else if (p < MLP_W2) {  // b1[h]
                const int h = p - MLP_B1;
                for (int s = 0; s < MLP_TILE; ++s) {
                    acc += d1[s][h];
                }
            } else if (p < MLP_B2) {  // W2[c][k]
                const int c = (p - MLP_W2) >> 6, k = (p - MLP_W2) & 63;
                for (int s = 0; s < MLP_TILE; ++s) {
                    acc += d2[s][c] * h1[s][k];
                }
            } else if (p < MLP_W3) {  // b2[c]
                const int c = p - MLP_B2;
                for (int s = 0; s < MLP_TILE; ++s) {
                    acc += d2[s][c];
                }
            } else if (p < MLP_B3) {  // W3[c]
                const int c = p - MLP_W3;
                for (int s = 0; s < MLP_TILE; ++s) {
                    acc += dy[s] * gelu(z2[s][c]);
                }
            } else {  // b3
                for (int s = 0; s < MLP_TILE; ++s) {
                    acc += dy[s];
                }
            }
            gacc[p] += acc;
        }
        if (tid == 0) {
            for (int s = 0; s < MLP_TILE; ++s) {
                lacc += lsum[s];
            }
        }
        __syncthreads();
    }
    for (int p = tid; p < MLP_NPARAM; p += 256) {
        gpart[(long)blockIdx.x * MLP_NPARAM + p] = gacc[p];
    }
    if (tid == 0) {
        lpart[blockIdx.x] = lacc;
    }
}

// grad[p] = sum_b gpart[b][p] (fixed order); loss = sum_b lpart[b]
__global__ void k_mlp_reduce(const float *__restrict__ gpart, const float *__restrict__ lpart, int nblocks,
                             float *__restrict__ grad, float *__restrict__ loss)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < MLP_NPARAM) {
        float s = 0.f;
        for (int b = 0; b < nblocks; ++b) {
            s += gpart[(long)b * MLP_NPARAM + p];
        }
        grad[p] = s;
    }
    if (p == 0) {
        float l = 0.f;
        for (int b = 0; b < nblocks; ++b) {
            l += lpart[b];
        }
        *loss = l;
    }
}

// Adam (Kingma & Ba 2015) on the parameters with gradient grad * gscale.
__global__ void k_mlp_adam(float *__restrict__ params, float *__restrict__ m, float *__restrict__ v,
                           const float *__restrict__ grad, float gscale, float lr, float b1, float b2, float eps,
                           int step)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) {
        return;
    }
    const float g = grad[p] * gscale;
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / (1.f - powf(b1, (float)step));
    const float vh = vv / (1.f - powf(b2, (float)step));
    params[p] -= lr * mh / (sqrtf(vh) + eps);
}

// ------------------------------------------------------------------ host side

int mlp_features(Context *ctx, const double *R, long ldr, const double *d, const double *rowabs,
                 const double *theta, const int *idx, const double *rmax, int n, int nt, float *feat, float *qf)
{
    const long total = (long)n * nt;
    const int blocks = (int)std::min<long>((total + 255) / 256, 16384);
    hipLaunchKernelGGL(k_mlp_features, dim3(std::max(blocks, 1)), dim3(256), 0, ctx->stream, R, ldr, d, rowabs, theta,
                       idx, rmax, n, nt, feat, qf);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

int mlp_infer(Context *ctx, const float *params, const float *feat, const float *qf, long count, float *out)
{
    if (count <= 0) {
        return IPT_CUDA_SUCCESS;
    }
    const long tiles = (count + 15) / 16;
    const int blocks = (int)std::min<long>((tiles + 3) / 4, 4096);
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, ctx->stream, params, feat, qf, count, out);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

// Davidson corrections T (n x nt, ldt) for the targets idx from the
// residual block R: features, inference, widening.
int mlp_correction(Context *ctx, const float *params, const double *R, long ldr, const double *d,
                   const double *rowabs, const double *theta, const int *idx, int n, int nt, double *T, long ldt)
{
    const size_t cnt = (size_t)n * nt;
    DevBuf<float> feat, qf, out;
    DevBuf<double> rmax;
    IPT_STCHK(feat.alloc(3 * cnt));
    IPT_STCHK(qf.alloc(cnt));
    IPT_STCHK(out.alloc(cnt));
    IPT_STCHK(rmax.alloc((size_t)nt));
    hipLaunchKernelGGL(k_col_absmax, dim3(nt), dim3(256), 0, ctx->stream, R, ldr, idx, n, rmax.p);
    IPT_HIPCHK(hipGetLastError());
    IPT_STCHK(mlp_features(ctx, R, ldr, d, rowabs, theta, idx, rmax.p, n, nt, feat.p, qf.p));
    IPT_STCHK(mlp_infer(ctx, params, feat.p, qf.p, (long)cnt, out.p));
    const int blocks = (int)std::min<size_t>((cnt + 255) / 256, 16384);
    hipLaunchKernelGGL(k_mlp_store, dim3(std::max(blocks, 1)), dim3(256), 0, ctx->stream, out.p, n, nt, T, ldt);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

// Partial gradients: one slice of MLP_NPARAM floats per block (and one loss
// float), at most kGradSlices blocks whatever the sample count (the blocks
// stride over the 64-sample tiles).
constexpr int kGradSlices = 1024;

static int mlp_gradient_blocks(long count)
{
    const long tiles = (count + MLP_TILE - 1) / MLP_TILE;
    return (int)std::max<long>(1, std::min<long>(tiles, kGradSlices));
}

// The two launches of the gradient; part holds nb * (MLP_NPARAM + 1) floats,
// every one of them written by k_mlp_grad before k_mlp_reduce reads it.
static int mlp_gradient_launch(hipStream_t stream, const float *params, const float *feat, const float *q,
                               const float *t, const float *w, long count, float *grad, float *loss, int nb,
                               float *part)
{
    hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(256), 0, stream, params, feat, q, t, w, count, part,
                       part + (size_t)nb * MLP_NPARAM);
    IPT_HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, stream, part,
                       part + (size_t)nb * MLP_NPARAM, nb, grad, loss);
    IPT_HIPCHK(hipGetLastError());
    return IPT_CUDA_SUCCESS;
}

static int mlp_gradient(Context *ctx, const float *params, const float *feat, const float *q, const float *t,
                        const float *w, long count, float *grad, float *loss)
{
    const int nb = mlp_gradient_blocks(count);
    float *part = nullptr;
    IPT_STCHK(context_scratch(ctx, SLOT_MISC, ((size_t)nb * MLP_NPARAM + nb) * sizeof(float), (void **)&part));
    return mlp_gradient_launch(ctx->stream, params, feat, q, t, w, count, grad, loss, nb, part);
}

// ------------------------------------------------------------------ gradient without allocation

// Partial gradients of ipt_hip_mlp_gradient_noalloc: a fixed array of
// kGradSlices slices (18.4 MB), which covers every sample count.  Each GPU
// holds its own copy until the process ends.  In host code `scratch` is a
// host-side stand-in for the symbol: kernels get the device address that
// hipGetSymbolAddress returns for the current device (scratch_address).
static __device__ __align__(256) float scratch[(size_t)kGradSlices * (MLP_NPARAM + 1)];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mutex;
static float *g_scratch_ptr[kMaxScratchDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
static int scratch_address(float **out)
{
    int dev = -1;
    IPT_HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices) {
        fprintf(stderr, "IPT: device %d beyond the %d devices the gradient scratch supports\n", dev,
                kMaxScratchDevices);
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            fprintf(stderr, "IPT: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
                    hipGetErrorString(e));
            return IPT_CUDA_CUDA_ERROR;
        }
        g_scratch_ptr[dev] = static_cast<float *>(p);
    }
    *out = g_scratch_ptr[dev];
    return IPT_CUDA_SUCCESS;
}

}  // namespace ipt

using namespace ipt;

extern "C" int ipt_hip_mlp_infer(const float *d_params, const float *d_feat, const float *d_qf, long long count,
                                 float *d_out, void *handle)
{
    if (d_params == nullptr || d_feat == nullptr || d_out == nullptr || count < 0) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    IPT_STCHK(mlp_infer(ctx, d_params, d_feat, d_qf, (long)count, d_out));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return IPT_CUDA_SUCCESS;
}

extern "C" int ipt_hip_mlp_gradient(const float *d_params, const float *d_feat, const float *d_q, const float *d_t,
                                    const float *d_w, long long count, float *d_grad, float *d_loss, void *handle)
{
    if (d_params == nullptr || d_feat == nullptr || d_q == nullptr || d_t == nullptr || d_w == nullptr ||
        d_grad == nullptr || d_loss == nullptr || count <= 0) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    IPT_STCHK(mlp_gradient(ctx, d_params, d_feat, d_q, d_t, d_w, (long)count, d_grad, d_loss));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return IPT_CUDA_SUCCESS;
}

extern "C" int ipt_hip_mlp_gradient_noalloc(const float *d_params, const float *d_feat, const float *d_q,
                                            const float *d_t, const float *d_w, long long count, float *d_grad,
                                            float *d_loss, void *handle)
{
    if (d_params == nullptr || d_feat == nullptr || d_q == nullptr || d_t == nullptr || d_w == nullptr ||
        d_grad == nullptr || d_loss == nullptr || count <= 0) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    const int nb = mlp_gradient_blocks((long)count);
    if (nb > kGradSlices) {
        fprintf(stderr, "IPT: %d partial gradients exceed the %d the fixed scratch holds\n", nb, kGradSlices);
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    float *part = nullptr;
    IPT_STCHK(scratch_address(&part));
    // a handle's stream, else the null stream: no library context is created
    hipStream_t stream = handle != nullptr ? static_cast<Context *>(handle)->stream : nullptr;
    IPT_STCHK(mlp_gradient_launch(stream, d_params, d_feat, d_q, d_t, d_w, (long)count, d_grad, d_loss, nb, part));
    IPT_HIPCHK(hipStreamSynchronize(stream));
    return IPT_CUDA_SUCCESS;
}

extern "C" int ipt_hip_mlp_adam_step(float *d_params, float *d_m, float *d_v, const float *d_grad, float gscale,
                                     float lr, int step, void *handle)
{
    if (d_params == nullptr || d_m == nullptr || d_v == nullptr || d_grad == nullptr || step <= 0) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, ctx->stream, d_params, d_m, d_v,
                       d_grad, gscale, lr, 0.9f, 0.999f, 1e-8f, step);
    IPT_HIPCHK(hipGetLastError());
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return IPT_CUDA_SUCCESS;
}

extern "C" int ipt_hip_mlp_features(const double *d_R, int ldr, const double *d_diag, const double *d_rowabs,
                                    const double *d_theta, const int *d_idx, int n, int nt, float *d_feat,
                                    float *d_qf, void *handle)
{
    if (d_R == nullptr || d_diag == nullptr || d_rowabs == nullptr || d_theta == nullptr || d_idx == nullptr ||
        d_feat == nullptr || d_qf == nullptr || n <= 0 || nt <= 0 || ldr < n) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    DevBuf<double> rmax;
    IPT_STCHK(rmax.alloc((size_t)nt));
    hipLaunchKernelGGL(k_col_absmax, dim3(nt), dim3(256), 0, ctx->stream, d_R, (long)ldr, d_idx, n, rmax.p);
    IPT_HIPCHK(hipGetLastError());
    IPT_STCHK(mlp_features(ctx, d_R, ldr, d_diag, d_rowabs, d_theta, d_idx, rmax.p, n, nt, d_feat, d_qf));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return IPT_CUDA_SUCCESS;
}
