// Plain (non-block) IPT entry points: dense and sparse, host and device.
//
// Semantics follow IPT/src/ipt_cuda.cu:1562-2253 and 5018-5316: argument
// validation, identity start (or caller's initial vectors), fixed iteration
// count or the tolerance stop ||F(Z)-Z||_F / ||Z||_F <= tol with the
// converged F(Z) returned, eigenvalue estimates lambda_j = (H Z)_jj.
#include <chrono>
#include <climits>
#include <cstring>
#include <functional>

#include "ipt_internal.h"
#include "ipt_plain.h"
#include "ipt_prep.h"

namespace ipt {

int block_dense_host_entry(const double *H, int n, int k, int maxiter, double tol, int use_tol,
                           IPTCudaResult *result);
int block_sparse_host_entry(const int *col_ptr, const int *row_ind, const double *values, int n,
                            int k, int nnz, int maxiter, double tol, int use_tol,
                            IPTCudaResult *result);
int block_dense_device(Context *ctx, const double *dH, int n, int k, int maxiter, double tol, int use_tol,
                       double **d_vec_out, double **d_val_out, IPTCudaResult *info);
int block_sparse_device(Context *ctx, const int *d_rp, const int *d_ci, const double *d_v, int n, int k,
                        int maxiter, double tol, int use_tol, double **d_vec_out, double **d_val_out,
                        int *iters_out, double *fp_out, bool *host_path);

namespace {

double seconds_since(const std::chrono::steady_clock::time_point &t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

// Generic fixed-point loop.  apply(X, Y) computes Y = H X in the layout
// (rs, cs); X/Y/Xn are n x k buffers in that layout.  On return *x_final
// points at the buffer that holds the returned iterate.
int plain_loop(Context *ctx, const std::function<int(const double *, double *)> &apply, int n, int k,
               long rs, long cs, const double *d, double *bufA, double *bufB, double *Y, int maxiter,
               double tol, int use_tol, double **x_final, int *iters_out, double *fp_out)
{
    double *d_norms = nullptr;
    IPT_STCHK(context_scratch(ctx, SLOT_SMALL, 2 * sizeof(double), (void **)&d_norms));
    double *x = bufA, *xn = bufB;
    int done = 0;
    double fp = NAN;
    for (int it = 0; it < maxiter; ++it) {
        IPT_STCHK(apply(x, Y));
        IPT_STCHK(plain_update(ctx, Y, x, xn, d, n, k, rs, cs, d_norms));
        done = it + 1;
        if (use_tol) {
            IPT_HIPCHK(hipMemcpyAsync(ctx->host_scalars, d_norms, 2 * sizeof(double),
                                      hipMemcpyDeviceToHost, ctx->stream));
            IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
            const double dn = std::sqrt(ctx->host_scalars[0]);
            const double zn = std::sqrt(ctx->host_scalars[1]);
            if (zn <= 0.0 || !std::isfinite(zn)) {
                return IPT_CUDA_INVALID_ARGUMENT;
            }
            fp = dn / zn;
            if (fp <= tol) {
                std::swap(x, xn);
                break;
            }
        }
        std::swap(x, xn);
    }
    *x_final = x;
    if (iters_out) {
        *iters_out = done;
    }
    if (fp_out) {
        *fp_out = fp;
    }
    return IPT_CUDA_SUCCESS;
}

bool plain_args_ok(int n, int k, int maxiter, double tol, int use_tol)
{
    return !(n <= 0 || k <= 0 || k > n || maxiter < 0 || (use_tol && (tol <= 0.0 || maxiter <= 0)));
}

// Dense plain IPT on device memory.  Z0 (device, column-major n x k) may be
// NULL.  Outputs are fresh device allocations.
int dense_plain_device(Context *ctx, const double *dH, int n, int k, const double *dZ0, int maxiter,
                       double tol, int use_tol, double **d_vec_out, double **d_val_out,
                       int *iters_out, double *fp_out)
{
    const size_t nk = (size_t)n * (size_t)k;
    DevBuf<double> a, b, y, dg, vals;
    IPT_STCHK(a.alloc(nk));
    IPT_STCHK(b.alloc(nk));
    IPT_STCHK(y.alloc(nk));
    IPT_STCHK(dg.alloc((size_t)n));
    IPT_STCHK(vals.alloc((size_t)k));
    if (dZ0 != nullptr) {
        IPT_HIPCHK(hipMemcpyAsync(a.p, dZ0, nk * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    } else {
        IPT_STCHK(set_identity(ctx, a.p, n, n, k));
    }
    IPT_STCHK(dense_diagonal(ctx, dH, n, n, dg.p));
    auto apply = [&](const double *X, double *Y) { return gemm_f64(ctx, dH, n, X, n, Y, n, n, k); };
    double *x = nullptr;
    IPT_STCHK(plain_loop(ctx, apply, n, k, 1, n, dg.p, a.p, b.p, y.p, maxiter, tol, use_tol, &x,
                         iters_out, fp_out));
    IPT_STCHK(apply(x, y.p));
    IPT_STCHK(gather_diagonal(ctx, y.p, k, 1, n, vals.p));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    *d_vec_out = (x == a.p) ? a.take() : b.take();
    *d_val_out = vals.take();
    return IPT_CUDA_SUCCESS;
}

namespace {

// Sparse plain IPT with a device CSR copy of H; iterate kept row-major.
int sparse_plain_device_csr(Context *ctx, const int *d_rp, const int *d_ci, const double *d_v,
                            const double *d_diag, int n, int k, const double *dZ0_colmajor,
                            int maxiter, double tol, int use_tol, double **d_vec_out,
                            double **d_val_out, int *iters_out, double *fp_out)
{
    const size_t nk = (size_t)n * (size_t)k;
    DevBuf<double> a, b, y, vec, vals;
    IPT_STCHK(a.alloc(nk));
    IPT_STCHK(b.alloc(nk));
    IPT_STCHK(y.alloc(nk));
    IPT_STCHK(vec.alloc(nk));
    IPT_STCHK(vals.alloc((size_t)k));
    if (dZ0_colmajor != nullptr) {
        IPT_STCHK(colmajor_to_rowmajor(ctx, dZ0_colmajor, n, a.p, k, n, k));
    } else {
        IPT_STCHK(set_identity(ctx, vec.p, n, n, k));
        IPT_STCHK(colmajor_to_rowmajor(ctx, vec.p, n, a.p, k, n, k));
    }
    auto apply = [&](const double *X, double *Y) {
        return spmm_csr_rowmajor(ctx, d_rp, d_ci, d_v, n, X, k, Y, k, k);
    };
    double *x = nullptr;
    IPT_STCHK(plain_loop(ctx, apply, n, k, k, 1, d_diag, a.p, b.p, y.p, maxiter, tol, use_tol, &x,
                         iters_out, fp_out));
    IPT_STCHK(apply(x, y.p));
    IPT_STCHK(gather_diagonal(ctx, y.p, k, k, 1, vals.p));
    IPT_STCHK(rowmajor_to_colmajor(ctx, x, k, vec.p, n, n, k));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    *d_vec_out = vec.take();
    *d_val_out = vals.take();
    return IPT_CUDA_SUCCESS;
}

void result_reset(IPTCudaResult *r)
{
    if (r != nullptr) {
        r->n = 0;
        r->k = 0;
        r->iterations = 0;
        r->vectors = nullptr;
        r->values = nullptr;
        r->basis_cols = 0;
        r->fixed_point_residual = NAN;
    }
}

int dense_with_initial_impl(const double *H, const double *Z0, int n, int k, int maxiter, double tol,
                            int use_tol, IPTCudaResult *result)
{
    result_reset(result);
    if (H == nullptr || result == nullptr || !plain_args_ok(n, k, maxiter, tol, use_tol)) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    if (Z0 == nullptr && env_flag("IPT_BLOCK_CLUSTER")) {
        return block_dense_host_entry(H, n, k, maxiter, tol, use_tol, result);
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(nullptr, &st);
    if (ctx == nullptr) {
        return st;
    }
    const size_t nn = (size_t)n * (size_t)n, nk = (size_t)n * (size_t)k;
    auto run = [&]() -> int {
        result->vectors = (double *)calloc(nk, sizeof(double));
        result->values = (double *)calloc((size_t)k, sizeof(double));
        if (result->vectors == nullptr || result->values == nullptr) {
            return IPT_CUDA_ALLOCATION_FAILED;
        }
        auto t0 = std::chrono::steady_clock::now();
        DevBuf<double> dH, dZ0;
        IPT_STCHK(dH.alloc(nn));
        IPT_HIPCHK(hipMemcpyAsync(dH.p, H, nn * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        if (Z0 != nullptr) {
            IPT_STCHK(dZ0.alloc(nk));
            IPT_HIPCHK(hipMemcpyAsync(dZ0.p, Z0, nk * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        }
        IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
        result->transfer_setup_time_sec = seconds_since(t0);
        auto t1 = std::chrono::steady_clock::now();
        double *dv = nullptr, *dl = nullptr;
        int iters = 0;
        double fp = NAN;
        IPT_STCHK(dense_plain_device(ctx, dH.p, n, k, dZ0.p, maxiter, tol, use_tol, &dv, &dl, &iters, &fp));
        result->iteration_time_sec = seconds_since(t1);
        result->solve_time_sec = result->iteration_time_sec;
        hipError_t e1 = hipMemcpy(result->vectors, dv, nk * sizeof(double), hipMemcpyDeviceToHost);
        hipError_t e2 = hipMemcpy(result->values, dl, (size_t)k * sizeof(double), hipMemcpyDeviceToHost);
        ipt_cuda_free_device_result(dv, dl);
        IPT_HIPCHK(e1);
        IPT_HIPCHK(e2);
        result->n = n;
        result->k = k;
        result->iterations = iters;
        result->basis_cols = k;
        result->matvecs = (long long)(iters + 1) * k;
        if (use_tol) {
            result->fixed_point_residual = fp;
        }
        return IPT_CUDA_SUCCESS;
    };
    st = run();
    if (st != IPT_CUDA_SUCCESS) {
        ipt_cuda_free_result(result);
    }
    return st;
}

int dense_device_impl(const double *dH, int n, int k, int maxiter, double tol, int use_tol, void *handle,
                      double **d_vec_out, double **d_val_out, int *iters_out, double *fp_out)
{
    if (d_vec_out) *d_vec_out = nullptr;
    if (d_val_out) *d_val_out = nullptr;
    if (iters_out) *iters_out = 0;
    if (fp_out) *fp_out = NAN;
    if (dH == nullptr || d_vec_out == nullptr || d_val_out == nullptr ||
        !plain_args_ok(n, k, maxiter, tol, use_tol)) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    int iters = 0;
    double fp = NAN;
    if (env_flag("IPT_BLOCK_CLUSTER")) {
        // extension: the reference routes only CSC input to the block path
        IPTCudaResult info;
        memset(&info, 0, sizeof(info));
        st = block_dense_device(ctx, dH, n, k, maxiter, tol, use_tol, d_vec_out, d_val_out, &info);
        iters = info.iterations;
        fp = info.fixed_point_residual;
    } else {
        st = dense_plain_device(ctx, dH, n, k, nullptr, maxiter, tol, use_tol, d_vec_out, d_val_out, &iters, &fp);
    }
    if (st != IPT_CUDA_SUCCESS) {
        *d_vec_out = nullptr;
        *d_val_out = nullptr;
        return st;
    }
    if (iters_out) *iters_out = iters;
    if (fp_out) *fp_out = fp;
    return IPT_CUDA_SUCCESS;
}

// Upload a host CSC as device CSR of the same matrix (+ diagonal).
struct DeviceCsr {
    DevBuf<int> rp, ci;
    DevBuf<double> v, diag;
};

int upload_csr(Context *ctx, const int *cp, const int *ri, const double *vals, int n, DeviceCsr *out)
{
    std::vector<int> rp, ci;
    std::vector<double> v;
    csc_to_csr(cp, ri, vals, n, &rp, &ci, &v);
    std::vector<double> dg;
    csc_diagonal_host(cp, ri, vals, n, &dg);
    IPT_STCHK(out->rp.alloc((size_t)n + 1));
    IPT_STCHK(out->ci.alloc(ci.size() > 0 ? ci.size() : 1));
    IPT_STCHK(out->v.alloc(v.size() > 0 ? v.size() : 1));
    IPT_STCHK(out->diag.alloc((size_t)n));
    IPT_HIPCHK(hipMemcpyAsync(out->rp.p, rp.data(), rp.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    if (!ci.empty()) {
        IPT_HIPCHK(hipMemcpyAsync(out->ci.p, ci.data(), ci.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        IPT_HIPCHK(hipMemcpyAsync(out->v.p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    }
    IPT_HIPCHK(hipMemcpyAsync(out->diag.p, dg.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
    return IPT_CUDA_SUCCESS;
}

int sparse_with_initial_impl(const int *cp, const int *ri, const double *vals, const double *Z0, int n,
                             int k, int nnz, int maxiter, double tol, int use_tol, IPTCudaResult *result)
{
    result_reset(result);
    if (cp == nullptr || ri == nullptr || vals == nullptr || result == nullptr || nnz < 0 ||
        !plain_args_ok(n, k, maxiter, tol, use_tol)) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    if (Z0 == nullptr && env_flag("IPT_BLOCK_CLUSTER")) {
        return block_sparse_host_entry(cp, ri, vals, n, k, nnz, maxiter, tol, use_tol, result);
    }
    if (!csc_valid(cp, ri, n, nnz)) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(nullptr, &st);
    if (ctx == nullptr) {
        return st;
    }
    const size_t nk = (size_t)n * (size_t)k;
    auto run = [&]() -> int {
        auto t0 = std::chrono::steady_clock::now();
        // Preparation, IPT/src/ipt_preparation.cu:1375-1648 (the reference's
        // plain CSC route, IPT/src/ipt_cuda.cu:5076-5092): stable sort of the
        // diagonal and lifting of degenerate target subspaces.
        LiftedCsc prep;
        IPT_STCHK(prepare_lifted_csc(cp, ri, vals, n, k, nnz, env_double("IPT_DEGENERACY_THRESHOLD", 0.0), &prep));
        const int *ucp = prep.a.col_ptr.data(), *uri = prep.a.row_ind.data();
        const double *uv = prep.a.values.data();
        result->preparation_time_sec = seconds_since(t0);
        result->vectors = (double *)calloc(nk, sizeof(double));
        result->values = (double *)calloc((size_t)k, sizeof(double));
        if (result->vectors == nullptr || result->values == nullptr) {
            return IPT_CUDA_ALLOCATION_FAILED;
        }
        auto t1 = std::chrono::steady_clock::now();
        DeviceCsr csr;
        IPT_STCHK(upload_csr(ctx, ucp, uri, uv, n, &csr));
        DevBuf<double> dZ0;
        if (Z0 != nullptr) {
            IPT_STCHK(dZ0.alloc(nk));
            IPT_HIPCHK(hipMemcpy(dZ0.p, Z0, nk * sizeof(double), hipMemcpyHostToDevice));
        }
        result->transfer_setup_time_sec = seconds_since(t1);
        auto t2 = std::chrono::steady_clock::now();
        double *dv = nullptr, *dl = nullptr;
        int iters = 0;
        double fp = NAN;
        IPT_STCHK(sparse_plain_device_csr(ctx, csr.rp.p, csr.ci.p, csr.v.p, csr.diag.p, n, k, dZ0.p,
                                          maxiter, tol, use_tol, &dv, &dl, &iters, &fp));
        result->iteration_time_sec = seconds_since(t2);
        result->solve_time_sec = result->iteration_time_sec;
        hipError_t e1 = hipMemcpy(result->vectors, dv, nk * sizeof(double), hipMemcpyDeviceToHost);
        hipError_t e2 = hipMemcpy(result->values, dl, (size_t)k * sizeof(double), hipMemcpyDeviceToHost);
        ipt_cuda_free_device_result(dv, dl);
        IPT_HIPCHK(e1);
        IPT_HIPCHK(e2);
        // back to the caller's basis, IPT/src/ipt_preparation.cu:1222-1294
        lifted_to_original(prep, result->vectors, k);
        result->n = n;
        result->k = k;
        result->iterations = iters;
        result->basis_cols = k;
        result->matvecs = (long long)(iters + 1) * k;
        if (use_tol) {
            result->fixed_point_residual = fp;
        }
        return IPT_CUDA_SUCCESS;
    };
    st = run();
    if (st != IPT_CUDA_SUCCESS) {
        ipt_cuda_free_result(result);
    }
    return st;
}

// Block path for a device CSC whose rows are not ascending or whose
// diagonal is not sorted: host preparation (permutation), results copied
// back to library-allocated device buffers.
int block_sparse_via_host(Context *ctx, const int *d_cp, const int *d_ri, const double *d_v, int n, int k, int nnz,
                          int maxiter, double tol, int use_tol, double **d_vec_out, double **d_val_out,
                          int *iters_out, double *fp_out)
{
    std::vector<int> cp((size_t)n + 1), ri((size_t)std::max(nnz, 1));
    std::vector<double> v((size_t)std::max(nnz, 1));
    IPT_HIPCHK(hipMemcpy(cp.data(), d_cp, cp.size() * sizeof(int), hipMemcpyDeviceToHost));
    if (nnz > 0) {
        IPT_HIPCHK(hipMemcpy(ri.data(), d_ri, (size_t)nnz * sizeof(int), hipMemcpyDeviceToHost));
        IPT_HIPCHK(hipMemcpy(v.data(), d_v, (size_t)nnz * sizeof(double), hipMemcpyDeviceToHost));
    }
    IPTCudaResult r;
    memset(&r, 0, sizeof(r));
    IPT_STCHK(block_sparse_host_entry(cp.data(), ri.data(), v.data(), n, k, nnz, maxiter, tol, use_tol, &r));
    DevBuf<double> vec, val;
    int st = vec.alloc((size_t)n * k);
    if (st == IPT_CUDA_SUCCESS) st = val.alloc((size_t)k);
    if (st == IPT_CUDA_SUCCESS &&
        (hipMemcpy(vec.p, r.vectors, (size_t)n * k * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(val.p, r.values, (size_t)k * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)) {
        st = IPT_CUDA_CUDA_ERROR;
    }
    if (iters_out) *iters_out = r.iterations;
    if (fp_out) *fp_out = r.fixed_point_residual;
    ipt_cuda_free_result(&r);
    IPT_STCHK(st);
    (void)ctx;
    *d_vec_out = vec.take();
    *d_val_out = val.take();
    return IPT_CUDA_SUCCESS;
}

int sparse_device_impl(const int *d_cp, const int *d_ri, const double *d_v, int n, int k, int nnz,
                       int maxiter, double tol, int use_tol, void *handle, double **d_vec_out,
                       double **d_val_out, int *iters_out, double *fp_out)
{
    if (d_vec_out) *d_vec_out = nullptr;
    if (d_val_out) *d_val_out = nullptr;
    if (iters_out) *iters_out = 0;
    if (fp_out) *fp_out = NAN;
    if (d_cp == nullptr || d_ri == nullptr || d_v == nullptr || d_vec_out == nullptr ||
        d_val_out == nullptr || nnz < 0 || !plain_args_ok(n, k, maxiter, tol, use_tol)) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int st = IPT_CUDA_SUCCESS;
    Context *ctx = context_from_handle(handle, &st);
    if (ctx == nullptr) {
        return st;
    }
    if (env_flag("IPT_BLOCK_CLUSTER")) {
        // symmetric input: the CSC is the CSR, used in place when its rows
        // ascend and its diagonal is sorted
        std::vector<int> hcp((size_t)n + 1);
        IPT_HIPCHK(hipMemcpy(hcp.data(), d_cp, hcp.size() * sizeof(int), hipMemcpyDeviceToHost));
        bool ok = hcp[0] == 0 && hcp[(size_t)n] == nnz;
        for (int i = 0; ok && i < n; ++i) {
            ok = hcp[(size_t)i] <= hcp[(size_t)i + 1];
        }
        if (!ok) {
            return IPT_CUDA_INVALID_ARGUMENT;
        }
        // adaptive expansion reads the CSC columns on the host
        bool host_path = block_targets_from_env(n, k).adaptive;
        if (!host_path) {
            IPT_STCHK(block_sparse_device(ctx, d_cp, d_ri, d_v, n, k, maxiter, tol, use_tol, d_vec_out, d_val_out,
                                          iters_out, fp_out, &host_path));
        }
        if (!host_path) {
            return IPT_CUDA_SUCCESS;
        }
        return block_sparse_via_host(ctx, d_cp, d_ri, d_v, n, k, nnz, maxiter, tol, use_tol, d_vec_out, d_val_out,
                                     iters_out, fp_out);
    }
    // The wavefront-per-row SpMM needs rows: build the CSR of H from the
    // caller's CSC (= CSR of H^T) on the host.
    std::vector<int> cp((size_t)n + 1), ri((size_t)nnz);
    std::vector<double> v((size_t)nnz);
    IPT_HIPCHK(hipMemcpy(cp.data(), d_cp, cp.size() * sizeof(int), hipMemcpyDeviceToHost));
    if (nnz > 0) {
        IPT_HIPCHK(hipMemcpy(ri.data(), d_ri, ri.size() * sizeof(int), hipMemcpyDeviceToHost));
        IPT_HIPCHK(hipMemcpy(v.data(), d_v, v.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    if (!csc_valid(cp.data(), ri.data(), n, nnz)) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    DeviceCsr csr;
    IPT_STCHK(upload_csr(ctx, cp.data(), ri.data(), v.data(), n, &csr));
    int iters = 0;
    double fp = NAN;
    st = sparse_plain_device_csr(ctx, csr.rp.p, csr.ci.p, csr.v.p, csr.diag.p, n, k, nullptr, maxiter, tol,
                                 use_tol, d_vec_out, d_val_out, &iters, &fp);
    if (st != IPT_CUDA_SUCCESS) {
        *d_vec_out = nullptr;
        *d_val_out = nullptr;
        return st;
    }
    if (iters_out) *iters_out = iters;
    if (fp_out) *fp_out = fp;
    return IPT_CUDA_SUCCESS;
}

}  // namespace
}  // namespace ipt

using namespace ipt;

extern "C" int ipt_cuda(const double *H, int n, int k, int iterations, IPTCudaResult *result)
{
    return dense_with_initial_impl(H, nullptr, n, k, iterations, 0.0, 0, result);
}

extern "C" int ipt_cuda_tol(const double *H, int n, int k, double tol, int maxiter, IPTCudaResult *result)
{
    return dense_with_initial_impl(H, nullptr, n, k, maxiter, tol, 1, result);
}

extern "C" int ipt_cuda_with_initial(const double *H, const double *Z0, int n, int k, int iterations,
                                     IPTCudaResult *result)
{
    return dense_with_initial_impl(H, Z0, n, k, iterations, 0.0, 0, result);
}

extern "C" int ipt_cuda_with_initial_tol(const double *H, const double *Z0, int n, int k, double tol,
                                         int maxiter, IPTCudaResult *result)
{
    return dense_with_initial_impl(H, Z0, n, k, maxiter, tol, 1, result);
}

extern "C" int ipt_cuda_device(const double *dH, int n, int k, int iterations, void *handle,
                               double **d_vectors_out, double **d_values_out)
{
    return dense_device_impl(dH, n, k, iterations, 0.0, 0, handle, d_vectors_out, d_values_out, nullptr,
                             nullptr);
}

extern "C" int ipt_cuda_device_tol(const double *dH, int n, int k, double tol, int maxiter, void *handle,
                                   double **d_vectors_out, double **d_values_out, int *iterations_done_out,
                                   double *fixed_point_residual_out)
{
    return dense_device_impl(dH, n, k, maxiter, tol, 1, handle, d_vectors_out, d_values_out,
                             iterations_done_out, fixed_point_residual_out);
}

extern "C" int ipt_cuda_sparse_csc(const int *cp, const int *ri, const double *v, int n, int k, int nnz,
                                   int iterations, IPTCudaResult *result)
{
    return sparse_with_initial_impl(cp, ri, v, nullptr, n, k, nnz, iterations, 0.0, 0, result);
}

extern "C" int ipt_cuda_sparse_csc_tol(const int *cp, const int *ri, const double *v, int n, int k, int nnz,
                                       double tol, int maxiter, IPTCudaResult *result)
{
    return sparse_with_initial_impl(cp, ri, v, nullptr, n, k, nnz, maxiter, tol, 1, result);
}

extern "C" int ipt_cuda_sparse_csc_with_initial(const int *cp, const int *ri, const double *v,
                                                const double *Z0, int n, int k, int nnz, int iterations,
                                                IPTCudaResult *result)
{
    return sparse_with_initial_impl(cp, ri, v, Z0, n, k, nnz, iterations, 0.0, 0, result);
}

extern "C" int ipt_cuda_sparse_csc_with_initial_tol(const int *cp, const int *ri, const double *v,
                                                    const double *Z0, int n, int k, int nnz, double tol,
                                                    int maxiter, IPTCudaResult *result)
{
    return sparse_with_initial_impl(cp, ri, v, Z0, n, k, nnz, maxiter, tol, 1, result);
}

extern "C" int ipt_cuda_sparse_csc_device(const int *d_cp, const int *d_ri, const double *d_v, int n, int k,
                                          int nnz, int iterations, void *handle, double **d_vectors_out,
                                          double **d_values_out)
{
    return sparse_device_impl(d_cp, d_ri, d_v, n, k, nnz, iterations, 0.0, 0, handle, d_vectors_out,
                              d_values_out, nullptr, nullptr);
}

extern "C" int ipt_cuda_sparse_csc_device_tol(const int *d_cp, const int *d_ri, const double *d_v, int n,
                                              int k, int nnz, double tol, int maxiter, void *handle,
                                              double **d_vectors_out, double **d_values_out,
                                              int *iterations_done_out, double *fixed_point_residual_out)
{
    return sparse_device_impl(d_cp, d_ri, d_v, n, k, nnz, maxiter, tol, 1, handle, d_vectors_out,
                              d_values_out, iterations_done_out, fixed_point_residual_out);
}

extern "C" int ipt_hip_prepare_sparse_csc(const int *col_ptr, const int *row_ind, const double *values, int n, int k,
                                          int nnz, double degeneracy_threshold, double *dense_out, int *perm_out,
                                          int *rotations_out)
{
    if (col_ptr == nullptr || row_ind == nullptr || values == nullptr || n <= 0 || k <= 0 || k > n || nnz < 0 ||
        !csc_valid(col_ptr, row_ind, n, nnz)) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    LiftedCsc prep;
    IPT_STCHK(prepare_lifted_csc(col_ptr, row_ind, values, n, k, nnz, degeneracy_threshold, &prep));
    if (dense_out != nullptr) {
        memset(dense_out, 0, (size_t)n * n * sizeof(double));
        for (int c = 0; c < n; ++c) {
            for (int p = prep.a.col_ptr[(size_t)c]; p < prep.a.col_ptr[(size_t)c + 1]; ++p) {
                dense_out[(size_t)prep.a.row_ind[(size_t)p] + (size_t)c * n] += prep.a.values[(size_t)p];
            }
        }
    }
    if (perm_out != nullptr) {
        for (int i = 0; i < n; ++i) {
            perm_out[i] = prep.perm.empty() ? i : prep.perm[(size_t)i];
        }
    }
    if (rotations_out != nullptr) {
        *rotations_out = (int)prep.rot_first.size();
    }
    return IPT_CUDA_SUCCESS;
}

extern "C" int ipt_hip_prepared_vectors_to_original(const int *col_ptr, const int *row_ind, const double *values,
                                                    int n, int k, int nnz, double degeneracy_threshold,
                                                    double *vectors, int kv)
{
    if (col_ptr == nullptr || row_ind == nullptr || values == nullptr || vectors == nullptr || n <= 0 || k <= 0 ||
        k > n || nnz < 0 || kv <= 0 || !csc_valid(col_ptr, row_ind, n, nnz)) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    LiftedCsc prep;
    IPT_STCHK(prepare_lifted_csc(col_ptr, row_ind, values, n, k, nnz, degeneracy_threshold, &prep));
    lifted_to_original(prep, vectors, kv);
    return IPT_CUDA_SUCCESS;
}
