// Low-precision (FP32) anchored correction product for CSR input
// (ipt_sparse_lp.hip).
#pragma once

#include "ipt_internal.h"

namespace ipt {

struct SparseLowp {
    const int *rp = nullptr;  // CSR row pointers / column indices (caller's)
    const int *ci = nullptr;
    int n = 0;
    DevBuf<float> vals;       // FP32 copy of the values
    const CsrSlices *slices = nullptr;  // column slices of the same CSR (narrow products), or none
};

int sparse_lowp_create(Context *ctx, const int *rp, const int *ci, const double *vals, int n, long long nnz,
                       SparseLowp *op);
// Y = Ya + diag(d) (B - Ba) + E (B - Ba), all n x nc column-major (ld),
// E (B - Ba) from FP32 values and an FP32 delta.
int sparse_lowp_apply_delta(Context *ctx, const SparseLowp &op, const double *B, const double *Ba, const double *Ya,
                            long ld, const double *d, int nc, double *Y);

}  // namespace ipt
