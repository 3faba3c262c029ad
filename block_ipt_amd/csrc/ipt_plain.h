// Plain IPT building blocks shared by the entry points of ipt_plain.cpp and
// the two-phase mixed-precision entry points of ipt_mixed.cpp.
#pragma once

#include <functional>

#include "ipt_internal.h"

namespace ipt {

// Fixed-point loop Z <- F(Z) (IPT/src/ipt_cuda.cu:1644-1693).  apply(X, Y)
// computes Y = H X in the layout (rs, cs); bufA holds the start iterate,
// bufB and Y are scratch of the same size.  With use_tol the loop stops
// when ||F(Z)-Z||_F / ||Z||_F <= tol and returns F(Z).  *x_final points at
// the buffer holding the returned iterate.
int plain_loop(Context *ctx, const std::function<int(const double *, double *)> &apply, int n, int k,
               long rs, long cs, const double *d, double *bufA, double *bufB, double *Y, int maxiter,
               double tol, int use_tol, double **x_final, int *iters_out, double *fp_out);

// Dense plain FP64 IPT on device memory from Z0 (device, n x k; NULL:
// identity).  Outputs are fresh device allocations.
int dense_plain_device(Context *ctx, const double *dH, int n, int k, const double *dZ0, int maxiter,
                       double tol, int use_tol, double **d_vec_out, double **d_val_out, int *iters_out,
                       double *fp_out);

}  // namespace ipt
