// Low-precision operator for the mixed-precision block IPT.
#pragma once

#include "ipt_internal.h"

namespace ipt {

enum { LOWP_FP16 = 0, LOWP_BF16 = 1, LOWP_BF16X3 = 2 };

// E = H - diag(H) packed in MFMA A-operand order (16 x 32 tiles).
struct LowpOperator {
    int n = 0;
    int mode = LOWP_FP16;
    int nkb = 0;      // k blocks of 32
    int nrb = 0;      // row blocks of 16
    int np = 0;       // padded k extent (nkb * 32)
    float scale = 1.f;  // E was multiplied by `scale` before rounding (FP16)
    DevBuf<unsigned short> hi;
    DevBuf<unsigned short> lo;  // BF16x3 only
};

// Pack E from a dense column-major FP64 H; max_abs_offdiag sizes the FP16 scale.
int lowp_operator_create(Context *ctx, const double *H, long ld, int n, int mode, double max_abs_offdiag,
                         LowpOperator *op);

// Y(:, 0:nc) = Ya + diag(d)(B - Ba) + E (B - Ba), the low-precision part in
// FP32-accumulated MFMA; all column-major with leading dimension ld.
int lowp_apply_delta(Context *ctx, const LowpOperator &op, const double *B, const double *Ba, const double *Ya,
                     long ld, const double *d, int nc, double *Y);

}  // namespace ipt
