// Device-side scalars and launchers of the interior-point iteration (ipm.hip, ipm.cpp).
#pragma once
#include "internal.h"

namespace tsvm {

// Slots of the device scalar array (double sc[IPM_NSLOT]).  Reductions write the raw
// sums; the consumers normalise.
enum IpmSlot {
  SL_Y = 0,      // equality multiplier y
  SL_MU,         // sum x s + (C - x) xi
  SL_NDX2,       // |r_x|^2
  SL_AX,         // a^T x
  SL_XQX,        // x^T Z Z^T x
  SL_SX,         // sum x
  SL_SXI,        // sum xi
  SL_L1,         // sum |x|
  SL_DMAX,       // max D
  SL_DMIN,       // min D
  SL_AR7,        // a^T r7 (predictor)
  SL_AB,         // a^T b, b = (D + Z Z^T)^-1 a
  SL_ALPHA_AFF,  // predictor step bound
  SL_MUAFF,      // sum of complementarity at the predictor point
  SL_AR7C,       // a^T r7 (corrector)
  SL_ALPHA,      // step bound of the combined direction
  SL_DY,         // predictor dy
  SL_DYC,        // corrector dy
  SL_CHOLFAIL,   // 1 when the capacitance Cholesky failed (the step was not taken)
  SL_QQ,         // init: |Z Z^T 1|^2
  SL_QE,         // init: 1^T Z Z^T 1
  SL_R2B,        // bordered refinement: a^T dx of the step being refined
  SL_AR7B,       // a^T S(r5')
  SL_DYB,        // bordered refinement correction of dy
  SL_BRES,       // |rho_1|^2 of the bordered residual
  SL_BRHS,       // |r5|^2
  SL_HSB,        // hybrid capacitance: error bound sqrt(sum t_i^2) of the BF16 rows
  SL_HSN,        // hybrid capacitance: number of FP64 rows (must follow SL_HSB)
  SL_HS8B,       // hybrid capacitance: error bound of the FP8 rows (must follow SL_HSN)
  SL_HS8N,       // hybrid capacitance: number of FP8 rows (must follow SL_HS8B)
  SL_HS2B,       // hybrid capacitance: bound of the compensated rows (must follow SL_HS8N)
  SL_HS2N,       // hybrid capacitance: number of compensated rows (must follow SL_HS2B)
  IPM_NSLOT      // number of slots (32: one 256-byte copy to the host per iteration)
};

static_assert(IPM_NSLOT == 32, "the scalar array is one 256-byte copy per iteration");
enum RedOp { RED_SUM = 0, RED_MAX = 1, RED_MIN = 2 };

// T (p x kp) = Z^T V with V p vectors of length n (row q at V + q*ldv); Z n x kp FP32.
tsvm_status ipm_gemv_t(Ctx* c, const float* Z, int64_t n, int kp, const double* V, int64_t ldv,
                       int p, double* T);
// E (p x n, row q at E + q*lde) = Z T, T p x kp.
tsvm_status ipm_gemv_n(Ctx* c, const float* Z, int64_t n, int kp, const double* T, int p,
                       double* E, int64_t lde);

}  // namespace tsvm
