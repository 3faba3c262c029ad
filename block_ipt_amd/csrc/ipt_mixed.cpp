// Two-phase mixed-precision entry points (low-precision plain IPT, then
// FP64 plain IPT restarted from its iterate) and the Rayleigh-mixed entry
// point (low-precision IPT, QR, FP64 IPT on the k x k Rayleigh quotient).
//
// Reference: IPT/src/mixed_precision/ipt_mixed_precision_cuda.cu:413-1225
// and ipt_rayleigh_mixed_cuda.cu:141-400.  The reference keeps the
// low-precision iterate in FP16/FP32 storage and multiplies it by an
// FP16/FP32 copy of E = H - diag(H) with cublasGemmEx.  Here the product
// E Z runs on MFMA in the requested input precision (FP32 accumulation)
// through the same operator as the block path's correction products
// (lowp_apply_delta with a zero anchor), and the iterate and the update stay
// FP64; the FP64 phase is the plain loop of ipt_plain.cpp.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/ipt_hip_acx.h"
#include "../../include/ipt_hip_mixed.h"
#include "ipt_acx.h"
#include "ipt_block.h"
#include "ipt_lowp.h"
#include "ipt_plain.h"

namespace ipt {

namespace {

double now_seconds()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool mixed_args_ok(const double *dH, int n, int k, double lo_tol, int lo_maxiter, double hi_tol, int hi_maxiter)
{
    return !(dH == nullptr || n <= 0 || k <= 0 || k > n || !(lo_tol > 0.0) || !(hi_tol > 0.0) ||
             lo_maxiter <= 0 || hi_maxiter <= 0);
}

// max |H_rc|, r != c, from the device scan (sizes the FP16 scale).
int offdiag_absmax(Context *ctx, const double *dH, int n, double *out)
{
    DevBuf<double> colsum, colmax;
    DevBuf<int> asym;
    IPT_STCHK(colsum.alloc((size_t)n));
    IPT_STCHK(colmax.alloc((size_t)n));
    IPT_STCHK(asym.alloc(1));
    IPT_STCHK(dense_scan(ctx, dH, n, n, colsum.p, colmax.p, asym.p));
    std::vector<double> h((size_t)n);
    IPT_HIPCHK(hipMemcpyAsync(h.data(), colmax.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    double m = 0.0;
    for (double x : h) {
        m = std::max(m, x);
    }
    *out = m;
    return IPT_CUDA_SUCCESS;
}

// Low-precision phase (ipt_mixed_precision_cuda.cu:413-800): plain IPT from
// the identity with H Z = diag(d) Z + E Z, E Z in `mode`, to `tol`.  On
// success *d_init owns the returned iterate (device, n x k column-major).
int lowp_phase(Context *ctx, const double *dH, int n, int k, int mode, double tol, int maxiter,
               DevBuf<double> *d_init, int *iters, double *fp)
{
    double mx = 0.0;
    IPT_STCHK(offdiag_absmax(ctx, dH, n, &mx));
    LowpOperator op;
    IPT_STCHK(lowp_operator_create(ctx, dH, n, n, mode, mx, &op));
    const size_t nk = (size_t)n * (size_t)k;
    DevBuf<double> a, b, y, dg;
    IPT_STCHK(a.alloc(nk));
    IPT_STCHK(b.alloc(nk));
    IPT_STCHK(y.alloc(nk));
    IPT_STCHK(dg.alloc((size_t)n));
    IPT_STCHK(set_identity(ctx, a.p, n, n, k));
    IPT_STCHK(dense_diagonal(ctx, dH, n, n, dg.p));
    auto apply = [&](const double *X, double *Y) {
        return lowp_apply_delta(ctx, op, X, nullptr, nullptr, n, dg.p, k, Y);
    };
    double *x = nullptr;
    IPT_STCHK(plain_loop(ctx, apply, n, k, 1, n, dg.p, a.p, b.p, y.p, maxiter, tol, 1, &x, iters, fp));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    d_init->release();
    d_init->p = (x == a.p) ? a.take() : b.take();
    d_init->count = nk;
    return IPT_CUDA_SUCCESS;
}

struct TwoPhase {
    int lo_iters = 0, hi_iters = 0;
    double lo_time = NAN, hi_time = NAN, total = NAN, lo_fp = NAN, hi_fp = NAN;
    double *vec = nullptr, *val = nullptr;
};

int two_phase(const double *dH, int n, int k, int mode, double lo_tol, int lo_maxiter, double hi_tol,
              int hi_maxiter, void *handle, TwoPhase *out)
{
    if (!mixed_args_ok(dH, n, k, lo_tol, lo_maxiter, hi_tol, hi_maxiter)) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    const double t0 = now_seconds();
    DevBuf<double> init;
    IPT_STCHK(lowp_phase(ctx, dH, n, k, mode, lo_tol, lo_maxiter, &init, &out->lo_iters, &out->lo_fp));
    const double t1 = now_seconds();
    IPT_STCHK(dense_plain_device(ctx, dH, n, k, init.p, hi_maxiter, hi_tol, 1, &out->vec, &out->val,
                                 &out->hi_iters, &out->hi_fp));
    const double t2 = now_seconds();
    out->lo_time = t1 - t0;
    out->hi_time = t2 - t1;
    out->total = t2 - t0;
    return IPT_CUDA_SUCCESS;
}

// Zero counts and pointers, NaN times and residuals (the reference's reset,
// ipt_mixed_precision_cuda.cu:94-155, ipt_rayleigh_mixed_cuda.cu:50-70).
// All four result structs are four ints, then doubles, then the two device
// pointers.
template <typename R>
void reset_two_phase(R *r)
{
    static_assert(offsetof(R, d_vectors) % sizeof(double) == 0 && 4 * sizeof(int) % sizeof(double) == 0,
                  "result layout: ints, doubles, pointers");
    memset(r, 0, sizeof(*r));
    double *first = reinterpret_cast<double *>(reinterpret_cast<char *>(r) + 4 * sizeof(int));
    double *end = reinterpret_cast<double *>(&r->d_vectors);
    for (double *q = first; q < end; ++q) {
        *q = NAN;
    }
}

// Q <- the orthonormal factor of its QR factorisation.  The reference uses
// Householder geqrf + orgqr (ipt_rayleigh_mixed_cuda.cu:141-254); here the
// same factor (up to column signs) comes from shifted Cholesky QR 3 on the
// device, all in FP64.
int orthogonalize_qr(Context *ctx, double *Q, int n, int k)
{
    return cholqr3_device(ctx, Q, n, k);
}

}  // namespace

// Preconditioned Cholesky QR (IPT/mixed/ipt_rayleigh_mixed_mpqr_cuda.cu:
// 108-386).  The reference normalises the columns, takes R from an FP32
// Householder QR, forms V = Q R^-1 and finishes with one FP64 Cholesky QR:
// the FP32 R brings the condition number of V to about 1/eps_32 so that the
// FP64 Cholesky succeeds.  Here the preconditioning step is a shifted FP64
// Cholesky QR (which handles condition numbers up to about 1/eps_64), and
// two plain Cholesky QR passes follow; the result is the orthonormal factor
// of the QR factorisation of Q, as the reference's.  Q: n x k, column-major,
// ld n.
int orthogonalize_mpqr(Context *ctx, double *Q, int n, int k)
{
    return cholqr3_device(ctx, Q, n, k);
}

}  // namespace ipt

using namespace ipt;

extern "C" int ipt_hip_cholqr_mixed(double *d_Q, int n, int k, void *handle)
{
    if (d_Q == nullptr || n <= 0 || k <= 0 || k > n) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    return orthogonalize_mpqr(ctx, d_Q, n, k);
}

#define IPT_TWO_PHASE_ENTRY(NAME, FREE, TYPE, MODE, LO_IT, LO_TIME, LO_FP)                                        \
    extern "C" void FREE(TYPE *result)                                                                            \
    {                                                                                                             \
        if (result == nullptr) {                                                                                  \
            return;                                                                                               \
        }                                                                                                         \
        (void)hipFree(result->d_vectors);                                                                         \
        (void)hipFree(result->d_values);                                                                          \
        reset_two_phase(result);                                                                                  \
    }                                                                                                             \
    extern "C" int NAME(const double *dH, int n, int k, double lo_tol, int lo_maxiter, double hi_tol,             \
                        int hi_maxiter, void *handle, TYPE *result)                                               \
    {                                                                                                             \
        if (result == nullptr) {                                                                                  \
            return IPT_CUDA_INVALID_ARGUMENT;                                                                     \
        }                                                                                                         \
        reset_two_phase(result);                                                                                  \
        TwoPhase tp;                                                                                              \
        const int st = two_phase(dH, n, k, MODE, lo_tol, lo_maxiter, hi_tol, hi_maxiter, handle, &tp);            \
        if (st != IPT_CUDA_SUCCESS) {                                                                             \
            (void)hipFree(tp.vec);                                                                                \
            (void)hipFree(tp.val);                                                                                \
            return st;                                                                                            \
        }                                                                                                         \
        result->n = n;                                                                                            \
        result->k = k;                                                                                            \
        result->LO_IT = tp.lo_iters;                                                                              \
        result->fp64_iterations = tp.hi_iters;                                                                    \
        result->LO_TIME = tp.lo_time;                                                                             \
        result->fp64_time_sec = tp.hi_time;                                                                       \
        result->total_time_sec = tp.total;                                                                        \
        result->LO_FP = tp.lo_fp;                                                                                 \
        result->fp64_fixed_point_residual = tp.hi_fp;                                                             \
        result->d_vectors = tp.vec;                                                                               \
        result->d_values = tp.val;                                                                                \
        return IPT_CUDA_SUCCESS;                                                                                  \
    }

IPT_TWO_PHASE_ENTRY(ipt_mixed_precision_cuda_device_tol, ipt_mixed_precision_cuda_free_device_result,
                    IPTMixedCudaDeviceResult, LOWP_FP16, fp16_iterations, fp16_time_sec, fp16_fixed_point_residual)
IPT_TWO_PHASE_ENTRY(ipt_mixed_precision_fp32_cuda_device_tol, ipt_mixed_precision_fp32_cuda_free_device_result,
                    IPTMixedFP32CudaDeviceResult, LOWP_FP32, fp32_iterations, fp32_time_sec,
                    fp32_fixed_point_residual)
IPT_TWO_PHASE_ENTRY(ipt_mixed_precision_tf32_cuda_device_tol, ipt_mixed_precision_tf32_cuda_free_device_result,
                    IPTMixedTF32CudaDeviceResult, LOWP_BF16X3, tf32_iterations, tf32_time_sec,
                    tf32_fixed_point_residual)

#undef IPT_TWO_PHASE_ENTRY

extern "C" void ipt_rayleigh_mixed_cuda_free_device_result(IPTRayleighMixedCudaDeviceResult *result)
{
    if (result == nullptr) {
        return;
    }
    (void)hipFree(result->d_vectors);
    (void)hipFree(result->d_values);
    reset_two_phase(result);
}

extern "C" int ipt_rayleigh_mixed_tf32_cuda_device_tol(const double *dH, int n, int k, double tf32_tol,
                                                       int tf32_maxiter, double fp64_tol, int fp64_maxiter,
                                                       void *handle, IPTRayleighMixedCudaDeviceResult *result)
{
    if (result == nullptr) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    reset_two_phase(result);
    if (!mixed_args_ok(dH, n, k, tf32_tol, tf32_maxiter, fp64_tol, fp64_maxiter)) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    double t = now_seconds();
    DevBuf<double> Q;
    IPT_STCHK(lowp_phase(ctx, dH, n, k, LOWP_BF16X3, tf32_tol, tf32_maxiter, &Q, &result->tf32_iterations,
                         &result->tf32_fixed_point_residual));
    result->tf32_time_sec = now_seconds() - t;

    t = now_seconds();
    // Householder QR (ipt_rayleigh_mixed_cuda.cu) or, with
    // IPT_RAYLEIGH_QR=mpqr, the FP32-preconditioned Cholesky QR of the
    // reference's mpqr build (ipt_rayleigh_mixed_mpqr_cuda.cu:456-460)
    const char *qr = getenv("IPT_RAYLEIGH_QR");
    if (qr != nullptr && !strcmp(qr, "mpqr")) {
        IPT_STCHK(orthogonalize_mpqr(ctx, Q.p, n, k));
    } else {
        IPT_STCHK(orthogonalize_qr(ctx, Q.p, n, k));
    }
    result->orthogonalize_time_sec = now_seconds() - t;

    // B = Q^T (H Q), symmetrised (ipt_rayleigh_mixed_cuda.cu:319-335)
    t = now_seconds();
    const size_t nk = (size_t)n * (size_t)k;
    DevBuf<double> HQ, B;
    IPT_STCHK(HQ.alloc(nk));
    IPT_STCHK(B.alloc((size_t)k * k));
    IPT_STCHK(gemm_f64(ctx, dH, n, Q.p, n, HQ.p, n, n, k));
    IPT_STCHK(gemm_tn_tall(ctx, Q.p, n, HQ.p, n, B.p, k, n, k, k));
    IPT_STCHK(symmetrize(ctx, B.p, k, k));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    result->rayleigh_transform_time_sec = now_seconds() - t;

    // plain FP64 IPT on the k x k Rayleigh quotient (ipt_rayleigh_mixed_cuda.cu:345-352)
    t = now_seconds();
    double *dz = nullptr, *dval = nullptr;
    st = dense_plain_device(ctx, B.p, k, k, nullptr, fp64_maxiter, fp64_tol, 1, &dz, &dval,
                            &result->fp64_iterations, &result->fp64_fixed_point_residual);
    if (st != IPT_CUDA_SUCCESS) {
        return st;
    }
    DevBuf<double> Z, vals;
    Z.p = dz;
    Z.count = (size_t)k * k;
    vals.p = dval;
    vals.count = (size_t)k;
    result->fp64_time_sec = now_seconds() - t;

    // vectors = Q Z (ipt_rayleigh_mixed_cuda.cu:355-362)
    t = now_seconds();
    DevBuf<double> V;
    IPT_STCHK(V.alloc(nk));
    IPT_STCHK(gemm_nn_small(ctx, Q.p, n, Z.p, k, V.p, n, n, k, k, 1.0, 0.0));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    result->backtransform_time_sec = now_seconds() - t;
    result->total_time_sec = result->tf32_time_sec + result->orthogonalize_time_sec +
                             result->rayleigh_transform_time_sec + result->fp64_time_sec +
                             result->backtransform_time_sec;
    result->n = n;
    result->k = k;
    result->d_vectors = V.take();
    result->d_values = vals.take();
    return IPT_CUDA_SUCCESS;
}

// ------------------------------------------------------------------ two-phase ACX

namespace ipt {
namespace {

// Low-precision ACX phase (IPT/src/ACX/acx_mixed_precision_cuda.cu:360-573):
// ACX from the identity with H X = diag(d) X + E X, E X in BF16x3, stop at
// fp <= tol; returns f1 when converged, else the last extrapolated x.
int acx_lowp_phase(Context *ctx, const double *dH, int n, int k, double tol, int maxiter, const double *d_diag,
                   AcxRun *run)
{
    double mx = 0.0;
    IPT_STCHK(offdiag_absmax(ctx, dH, n, &mx));
    LowpOperator op;
    IPT_STCHK(lowp_operator_create(ctx, dH, n, n, LOWP_BF16X3, mx, &op));
    const std::function<int(const double *, double *)> product = [&](const double *X, double *Y) {
        return lowp_apply_delta(ctx, op, X, nullptr, nullptr, n, d_diag, k, Y);
    };
    run->product = &product;
    run->stop_tol = tol;
    run->stop_inclusive = true;
    run->final_x_unconverged = true;
    run->maxiter = maxiter;
    return acx_loop(ctx, d_diag, n, k, run);
}

// FP64 ACX on a device matrix A (m x m, ld m) from x0 (NULL: identity),
// stop at fp < min(tol, 1e-12); values diag(A vectors).
int acx_fp64_phase(Context *ctx, const double *A, int m, int k, double tol, int maxiter, const double *x0,
                   bool final_x_unconverged, AcxRun *run, DevBuf<double> *values, double *fin_time)
{
    DevBuf<double> d;
    IPT_STCHK(d.alloc((size_t)m));
    IPT_STCHK(dense_diagonal(ctx, A, m, m, d.p));
    const std::function<int(const double *, double *)> product = [&](const double *X, double *Y) {
        return gemm_f64(ctx, A, m, X, m, Y, m, m, k);
    };
    run->product = &product;
    run->x0 = x0;
    run->stop_tol = tol < 1.0e-12 ? tol : 1.0e-12;
    run->stop_inclusive = false;
    run->final_x_unconverged = final_x_unconverged;
    run->maxiter = maxiter;
    const int st = acx_loop(ctx, d.p, m, k, run);
    if (st != IPT_CUDA_SUCCESS) {
        return st;
    }
    const double t = now_seconds();
    IPT_STCHK(values->alloc((size_t)k));
    IPT_STCHK(acx_diag_product(ctx, A, m, run->vectors.p, k, values->p));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    *fin_time = now_seconds() - t;
    return IPT_CUDA_SUCCESS;
}

bool acx_mixed_args_ok(const double *dH, int n, int k, double lo_tol, int lo_maxiter, double hi_tol, int hi_maxiter)
{
    // IPT/src/ACX/acx_mixed_precision_cuda.cu:832-836 (the handle may be NULL here)
    return !(dH == nullptr || n <= 0 || k <= 0 || k > n || lo_tol <= 0.0 || hi_tol <= 0.0 || lo_maxiter <= 0 ||
             hi_maxiter <= 0);
}

}  // namespace
}  // namespace ipt

extern "C" void acx_mixed_precision_tf32_cuda_reset_result(ACXMixedTF32CudaResult *r)
{
    if (r == nullptr) {
        return;
    }
    memset(r, 0, sizeof(*r));
    r->tf32_time_sec = NAN;
    r->fp64_time_sec = NAN;
    r->total_time_sec = NAN;
    r->finalization_time_sec = NAN;
    r->tf32_fixed_point_residual = NAN;
    r->fp64_fixed_point_residual = NAN;
    r->fp64_max_residual = NAN;
}

extern "C" void acx_mixed_precision_tf32_cuda_free_result(ACXMixedTF32CudaResult *r)
{
    if (r == nullptr) {
        return;
    }
    if (r->d_vectors != nullptr) {
        (void)hipFree(r->d_vectors);
    }
    if (r->d_values != nullptr) {
        (void)hipFree(r->d_values);
    }
    acx_mixed_precision_tf32_cuda_reset_result(r);
}

extern "C" int acx_mixed_precision_tf32_cuda_device_tol(const double *dH, int n, int k, double tf32_tol,
                                                        int tf32_maxiter, double fp64_tol, int fp64_maxiter,
                                                        void *handle, ACXMixedTF32CudaResult *result)
{
    using namespace ipt;
    if (result == nullptr) {
        return ACX_CUDA_INVALID_ARGUMENT;
    }
    acx_mixed_precision_tf32_cuda_reset_result(result);
    if (!acx_mixed_args_ok(dH, n, k, tf32_tol, tf32_maxiter, fp64_tol, fp64_maxiter)) {
        return ACX_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    auto run = [&]() -> int {
        const double t0 = now_seconds();
        DevBuf<double> d;
        IPT_STCHK(d.alloc((size_t)n));
        IPT_STCHK(dense_diagonal(ctx, dH, n, n, d.p));
        AcxRun lo;
        const int s1 = acx_lowp_phase(ctx, dH, n, k, tf32_tol, tf32_maxiter, d.p, &lo);
        result->tf32_iterations = lo.iterations;
        result->tf32_f_calls = lo.f_calls;
        result->tf32_converged = lo.converged;
        result->tf32_fixed_point_residual = lo.fixed_point_residual;
        if (s1 != IPT_CUDA_SUCCESS) {
            return s1;
        }
        result->tf32_time_sec = now_seconds() - t0;
        const double t1 = now_seconds();
        AcxRun hi;
        DevBuf<double> vals;
        double fin = 0.0;
        const int s2 = acx_fp64_phase(ctx, dH, n, k, fp64_tol, fp64_maxiter, lo.vectors.p, true, &hi, &vals, &fin);
        result->fp64_iterations = hi.iterations;
        result->fp64_f_calls = hi.f_calls;
        result->fp64_converged = hi.converged;
        result->fp64_fixed_point_residual = hi.fixed_point_residual;
        result->fp64_max_residual = hi.max_residual;
        if (s2 != IPT_CUDA_SUCCESS) {
            return s2;
        }
        result->fp64_time_sec = now_seconds() - t1;
        result->finalization_time_sec = fin;
        result->total_time_sec = now_seconds() - t0;
        result->n = n;
        result->k = k;
        result->d_vectors = hi.vectors.take();
        result->d_values = vals.take();
        return IPT_CUDA_SUCCESS;
    };
    st = run();
    if (st != IPT_CUDA_SUCCESS) {
        acx_mixed_precision_tf32_cuda_free_result(result);
    }
    return st;
}

extern "C" void acx_rayleigh_mixed_tf32_cuda_reset_result(ACXRayleighMixedTF32CudaResult *r)
{
    if (r == nullptr) {
        return;
    }
    memset(r, 0, sizeof(*r));
    r->tf32_time_sec = NAN;
    r->orthogonalize_time_sec = NAN;
    r->rayleigh_transform_time_sec = NAN;
    r->fp64_time_sec = NAN;
    r->finalization_time_sec = NAN;
    r->backtransform_time_sec = NAN;
    r->total_time_sec = NAN;
    r->tf32_fixed_point_residual = NAN;
    r->fp64_fixed_point_residual = NAN;
    r->fp64_max_residual = NAN;
}

extern "C" void acx_rayleigh_mixed_tf32_cuda_free_result(ACXRayleighMixedTF32CudaResult *r)
{
    if (r == nullptr) {
        return;
    }
    if (r->d_vectors != nullptr) {
        (void)hipFree(r->d_vectors);
    }
    if (r->d_values != nullptr) {
        (void)hipFree(r->d_values);
    }
    acx_rayleigh_mixed_tf32_cuda_reset_result(r);
}

extern "C" int acx_rayleigh_mixed_tf32_cuda_device_tol(const double *dH, int n, int k, double tf32_tol,
                                                       int tf32_maxiter, double fp64_tol, int fp64_maxiter,
                                                       void *handle, ACXRayleighMixedTF32CudaResult *result)
{
    using namespace ipt;
    if (result == nullptr) {
        return ACX_CUDA_INVALID_ARGUMENT;
    }
    acx_rayleigh_mixed_tf32_cuda_reset_result(result);
    if (!acx_mixed_args_ok(dH, n, k, tf32_tol, tf32_maxiter, fp64_tol, fp64_maxiter)) {
        return ACX_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    auto run = [&]() -> int {
        double t = now_seconds();
        DevBuf<double> d;
        IPT_STCHK(d.alloc((size_t)n));
        IPT_STCHK(dense_diagonal(ctx, dH, n, n, d.p));
        AcxRun lo;
        const int s1 = acx_lowp_phase(ctx, dH, n, k, tf32_tol, tf32_maxiter, d.p, &lo);
        result->tf32_iterations = lo.iterations;
        result->tf32_f_calls = lo.f_calls;
        result->tf32_converged = lo.converged;
        result->tf32_fixed_point_residual = lo.fixed_point_residual;
        if (s1 != IPT_CUDA_SUCCESS) {
            return s1;
        }
        result->tf32_time_sec = now_seconds() - t;
        // Householder QR of the iterate (acx_rayleigh_mixed_precision_cuda.cu:
        // 111-170; the column normalisation before it does not change Q)
        t = now_seconds();
        double *Q = lo.vectors.p;
        IPT_STCHK(orthogonalize_qr(ctx, Q, n, k));
        result->orthogonalize_time_sec = now_seconds() - t;
        // B = Q^T (H Q), symmetrised
        t = now_seconds();
        const size_t nk = (size_t)n * (size_t)k;
        DevBuf<double> HQ, B;
        IPT_STCHK(HQ.alloc(nk));
        IPT_STCHK(B.alloc((size_t)k * k));
        IPT_STCHK(gemm_f64(ctx, dH, n, Q, n, HQ.p, n, n, k));
        IPT_STCHK(gemm_tn_tall(ctx, Q, n, HQ.p, n, B.p, k, n, k, k));
        IPT_STCHK(symmetrize(ctx, B.p, k, k));
        IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
        result->rayleigh_transform_time_sec = now_seconds() - t;
        // FP64 ACX on the k x k quotient (the FP64 entry point's rules)
        t = now_seconds();
        AcxRun hi;
        DevBuf<double> vals;
        double fin = 0.0;
        const int s2 = acx_fp64_phase(ctx, B.p, k, k, fp64_tol, fp64_maxiter, nullptr, false, &hi, &vals, &fin);
        result->fp64_iterations = hi.iterations;
        result->fp64_f_calls = hi.f_calls;
        result->fp64_converged = hi.converged;
        result->fp64_fixed_point_residual = hi.fixed_point_residual;
        result->fp64_max_residual = hi.max_residual;
        if (s2 != IPT_CUDA_SUCCESS) {
            return s2;
        }
        result->fp64_time_sec = now_seconds() - t;
        result->finalization_time_sec = fin;
        // vectors = Q Z
        t = now_seconds();
        DevBuf<double> V;
        IPT_STCHK(V.alloc(nk));
        IPT_STCHK(gemm_nn_small(ctx, Q, n, hi.vectors.p, k, V.p, n, n, k, k, 1.0, 0.0));
        IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
        result->backtransform_time_sec = now_seconds() - t;
        result->total_time_sec = result->tf32_time_sec + result->orthogonalize_time_sec +
                                 result->rayleigh_transform_time_sec + result->fp64_time_sec +
                                 result->backtransform_time_sec;
        result->n = n;
        result->k = k;
        result->d_vectors = V.take();
        result->d_values = vals.take();
        return IPT_CUDA_SUCCESS;
    };
    st = run();
    if (st != IPT_CUDA_SUCCESS) {
        acx_rayleigh_mixed_tf32_cuda_free_result(result);
    }
    return st;
}
