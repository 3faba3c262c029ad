// Low-precision operator for the mixed-precision block IPT.
#pragma once

#include "ipt_internal.h"

namespace ipt {

// FP16 / BF16: one 16-bit plane of E and of the operand.  BF16X3: hi and lo
// BF16 planes (three products).  FP32: E and the operand in FP32, the two
// planes holding the first and second four of a lane's eight k values
// (MFMA 16x16x4 FP32).  FP8: OCP E4M3 with one scale per 128 x 128 tile of E
// (max|E_tile| / 448) and a power-of-two scale per operand column; a lane's
// 16 bytes hold two consecutive k blocks (MFMA 16x16x32 FP8).
enum { LOWP_FP16 = 0, LOWP_BF16 = 1, LOWP_BF16X3 = 2, LOWP_FP32 = 3, LOWP_FP8 = 4 };

// Modes whose E and operand occupy two 16-byte planes per lane and tile.
inline constexpr bool lowp_two_planes(int mode) { return mode == LOWP_BF16X3 || mode == LOWP_FP32; }

// E = H - diag(H) packed in MFMA A-operand order (16 x 32 tiles).
struct LowpOperator {
    int n = 0;
    int mode = LOWP_FP16;
    int nkb = 0;      // k blocks of 32
    int nrb = 0;      // row blocks of 16
    int np = 0;       // padded k extent (nkb * 32)
    float scale = 1.f;  // E was multiplied by `scale` before rounding (FP16)
    DevBuf<unsigned short> hi;
    DevBuf<unsigned short> lo;  // second plane (BF16x3, FP32)
    DevBuf<float> tscale;       // FP8: scale of tile (row tile rt, k tile kt) at rt * (nkb/4) + kt
};

// Pack E from a dense column-major FP64 H; max_abs_offdiag sizes the FP16 scale.
int lowp_operator_create(Context *ctx, const double *H, long ld, int n, int mode, double max_abs_offdiag,
                         LowpOperator *op);

// Y(:, 0:nc) = Ya + diag(d)(B - Ba) + E (B - Ba), the low-precision part in
// FP32-accumulated MFMA; all column-major with leading dimension ld.  Ba and
// Ya may be NULL (zero anchor: Y = diag(d) B + E B).
int lowp_apply_delta(Context *ctx, const LowpOperator &op, const double *B, const double *Ba, const double *Ya,
                     long ld, const double *d, int nc, double *Y);

// The same product from an operand the block update already wrote
// (ipt_emit.h): planes over nb_all column blocks of 16, scale colscale[c]
// per basis column; the columns [c_lo, c_lo + c_w) are covered, rounded out
// to whole column blocks (at most 16 * nb_all = the basis padded, bc the
// basis width: no column past it is written).  Ya, B, Ba, Y: the full basis
// (column 0 first).
int lowp_apply_emitted(Context *ctx, const LowpOperator &op, const unsigned short *zhi, const unsigned short *zlo,
                       int nb_all, int c_lo, int c_w, const double *colscale, const double *B, const double *Ba,
                       const double *Ya, long ld, const double *d, int bc, double *Y);

}  // namespace ipt
