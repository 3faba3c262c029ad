// ACX iteration shared by the FP64 entry point (ipt_acx.hip) and the
// two-phase entry points (ipt_mixed.cpp).
#pragma once

#include <functional>

#include "ipt_internal.h"

namespace ipt {

// One ACX run (IPT/src/ACX/acx_cuda.cu:476-593 and its two-phase twins,
// acx_mixed_precision_cuda.cu:454-536, 680-760).
struct AcxRun {
    // inputs
    const std::function<int(const double *, double *)> *product;
    const double *x0 = nullptr;  // start (device, n x k); NULL: I(:, 0:k)
    double stop_tol = 0.0;
    bool stop_inclusive = false;   // fp <= tol (low-precision phase) or fp < tol
    bool final_x_unconverged = false;  // unconverged: return x (two-phase) or f1
    int maxiter = 0;
    // outputs
    int iterations = 0, f_calls = 0, converged = 0;
    double fixed_point_residual = NAN, max_residual = NAN;
    DevBuf<double> vectors;  // the returned iterate
};

// Runs the ACX loop from run->x0 (or the identity) with run->product as
// H X; diag is H's diagonal on the device.  ACX_CUDA_* status.
int acx_loop(Context *ctx, const double *d_diag, int n, int k, AcxRun *run);
// values[c] = (H V)(c, c), c < k, from k rows of H.
int acx_diag_product(Context *ctx, const double *H, int n, const double *V, int k, double *values);

}  // namespace ipt
