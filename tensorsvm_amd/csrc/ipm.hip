// Primal-dual interior-point solver on the GPU (replaces mpc, tensorsvm-train.cu:1507-1785,
// NewtonStep 1794-1822 and SMWSolve 1824-1840).
//
// State (x, s, xi, the step, D) lives in HBM as FP64 n-vectors; the factor Z stays
// resident as FP32 (n x kp, kp = k rounded up to 32 with zero columns).  Per iteration:
//   residual pass     Z^T x, Z (Z^T x), then one fused O(n) pass producing r_x, r_s, r_xi,
//                     D, 1/D and every reduction of the stop test (mu, |r_x|, a^T x,
//                     x^T Q x, sums, max/min D) -- one 256-byte device->host copy
//   capacitance       M = I + Z^T D^-1 Z: FP64 MFMA from FP32 Z (plain form) or BF16
//                     MFMA with FP32 accumulation (BF16 form); FP64 Cholesky on the GPU
//   predictor         SMW solves for [r5, a] batched into one pass pair over Z; the
//                     reference's third solve is a linear combination (S^-1 is linear:
//                     S^-1 (a dy - r5) = dy S^-1 a - S^-1 r5); step bound and mu_aff as
//                     device reductions
//   corrector         the combined step starts from the affine one, and a sweep of
//                     bordered refinement against the combined right-hand sides (the
//                     residual of the bordered Newton system against the exact operator
//                     D + Z Z^T, FP32 factor, FP64 accumulation; one SMW solve; the
//                     correction again combined with S^-1 a) adds the corrector's step and
//                     corrects the affine one in one solve; update.  (refine > 0 instead
//                     solves the corrector's system on its own and refines each step.)
// The last pass of each SMW solve also accumulates Z^T sol, so Z^T of the combined step is
// a combination of k-vectors, and the bordered residual pass also accumulates Z^T of the
// next solve's right-hand side; Z^T x is carried from iteration to iteration.  Per
// iteration that is 4 reads of Z for 256 <= kp <= 768 (the residual pass, which also forms
// the predictor's right-hand sides and their Z^T; the predictor's finish; the sweep's
// residual and finish), 8 with refine = 1.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "internal.h"
#include "ipm_kernels.h"

namespace tsvm {

constexpr int EW_BLOCKS = 1024;  // blocks of the fused O(n) passes
constexpr int MAXR = 10;         // reduction slots per pass
// Error budget of the hybrid capacitance: the BF16 rows satisfy sqrt(sum t_i^2) <= 16
// (t_i = w_i |z_i|^2), which keeps |M~ - M|_2, and with lambda_min(M) >= 1 the
// refinement's contraction factor, near 0.1 (three bf16 roundings per term, zero-mean
// and independent across rows: ~c 2^-9 16).  Measured at n = 1.6M: budgets 4, 16 and 64
// take the same iteration count (163, 163, 162) at the same held-out accuracy, with 68%,
// 88% and 94% of the rows in BF16.
constexpr double IPM_SYRK_BUDGET = 16.0;

// ------------------------------------------------------------------------------------
// block reduction helper: each block writes its R partials to red[blockIdx.x * R + r]
// ------------------------------------------------------------------------------------
template <int R>
__device__ inline void block_reduce_store(double (&v)[R], const int (&op)[R], double* red) {
  __shared__ double sh[4][R];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < R; r++) {
    double x = v[r];
    if (op[r] == RED_SUM) x = wave_sum(x);
    else if (op[r] == RED_MAX) x = wave_max(x);
    else x = wave_min(x);
    if (lane == 0) sh[wave][r] = x;
  }
  __syncthreads();
  if (threadIdx.x < R) {
    int r = threadIdx.x;
    double x = sh[0][r];
    for (int w = 1; w < 4; w++) {
      double y = sh[w][r];
      x = op[r] == RED_SUM ? x + y : (op[r] == RED_MAX ? fmax(x, y) : fmin(x, y));
    }
    red[blockIdx.x * R + r] = x;
  }
}

// which reductions a pass left in its per-block partials, and where their results go
// (passed by value: no host-to-device copy per reduction)
struct RedSpec {
  int n;
  int op[MAXR];
  int slot[MAXR];
};

// sc[slot[r]] = reduction over nblk blocks of red[b * R + r]  (one block, fixed order)
__global__ __launch_bounds__(256) void reduce_slots(const double* __restrict__ red, int nblk,
                                                    const RedSpec rs, double* __restrict__ sc) {
  __shared__ double sh[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int R = rs.n;
  for (int r = 0; r < R; r++) {
    const int op = rs.op[r];
    double x = op == RED_SUM ? 0.0 : (op == RED_MAX ? -INFINITY : INFINITY);
    for (int b = threadIdx.x; b < nblk; b += 256) {
      double y = red[b * R + r];
      x = op == RED_SUM ? x + y : (op == RED_MAX ? fmax(x, y) : fmin(x, y));
    }
    x = op == RED_SUM ? wave_sum(x) : (op == RED_MAX ? wave_max(x) : wave_min(x));
    if (lane == 0) sh[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      double z = sh[0];
      for (int w = 1; w < 4; w++)
        z = op == RED_SUM ? z + sh[w] : (op == RED_MAX ? fmax(z, sh[w]) : fmin(z, sh[w]));
      sc[rs.slot[r]] = z;
    }
    __syncthreads();
  }
}

// Fixed-order sum of per-block partials: T[e] = sum_b part[b * m + e] in two stages
// (slices of 32 blocks, then the slices), so that no wave walks all nblk partials.
__global__ __launch_bounds__(256) void reduce_parts_kernel(const double* __restrict__ part, int nblk,
                                                           int m, double* __restrict__ dst) {
  __shared__ double sh[4][64];
  const int o = blockIdx.x * 64 + (threadIdx.x & 63), g = threadIdx.x >> 6;
  const int b0 = blockIdx.y * 32, b1 = b0 + 32 < nblk ? b0 + 32 : nblk;
  double s = 0.0;
  if (o < m)
#pragma unroll 8
    for (int b = b0 + g; b < b1; b += 4) s += part[(int64_t)b * m + o];
  sh[g][threadIdx.x & 63] = s;
  __syncthreads();
  if (g == 0 && o < m)
    dst[(int64_t)blockIdx.y * m + o] = sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
}

static tsvm_status reduce_parts(Ctx* c, const double* part, int nblk, int m, double* T) {
  if (nblk < 1 || nblk > 32 * 32) return fail(c, TSVM_ERR_ARG, "reduce_parts: %d partials", nblk);
  const int S = (int)ceil_div(nblk, 32);
  const dim3 b(256);
  if (S == 1) {
    hipLaunchKernelGGL(reduce_parts_kernel, dim3((unsigned)ceil_div(m, 64), 1), b, 0, c->stream,
                       part, nblk, m, T);
  } else {
    void* p2;
    TSVM_HIP(c, c->ws.get("reduce_parts2", (size_t)S * m * 8, &p2));
    hipLaunchKernelGGL(reduce_parts_kernel, dim3((unsigned)ceil_div(m, 64), (unsigned)S), b, 0,
                       c->stream, part, nblk, m, (double*)p2);
    hipLaunchKernelGGL(reduce_parts_kernel, dim3((unsigned)ceil_div(m, 64), 1), b, 0, c->stream,
                       (const double*)p2, S, m, T);
  }
  TSVM_HIP(c, hipGetLastError());
  return TSVM_OK;
}


// Fused passes over Z (defined below): modes and launcher.
enum { ZP_SMW = 0, ZP_BORDER = 1, ZP_T = 2, ZP_N = 3, ZP_RES = 4 };
// operands and outputs of the fused residual pass (ZP_RES)
struct ResArgs {
  const double *x, *s, *xi;
  double *dx, *ds, *dxi, *D, *winv, *rhs, *V;
  double C;
};
static tsvm_status ipm_zpass(Ctx* c, int mode, int p, int q, const float* Z, int64_t n, int kp,
                             const double* T, const double* V, const double* winv,
                             const double* a, const double* D, const double* r5, const double* dx,
                             const double* sc, int corr, double* out, double* red, double* Tout,
                             const ResArgs* ra = nullptr, int* nred = nullptr);

// ------------------------------------------------------------------------------------
// Z^T V: block b owns a row range; thread t handles float4 column group t % (kp/4) of
// rows t / (kp/4) + step * R.  Partials per block, then a fixed-order reduction.
// ------------------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(256) void gemv_t_kernel(const float* __restrict__ Z, int64_t n, int kp,
                                                     const double* __restrict__ V, int64_t ldv,
                                                     int64_t rows_per_blk, double* __restrict__ part) {
  extern __shared__ double shm[];
  const int L = kp / 4;
  const int R = 256 / L;
  const int slot = threadIdx.x / L, c4 = threadIdx.x % L;
  const bool active = slot < R;
  const int64_t r0 = blockIdx.x * rows_per_blk;
  const int64_t r1 = r0 + rows_per_blk < n ? r0 + rows_per_blk : n;
  double acc[P][4];
#pragma unroll
  for (int q = 0; q < P; q++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[q][j] = 0.0;
  if (active) {
    int64_t i = r0 + slot;
    for (; i + 3 * R < r1; i += 4 * R) {
      f32x4_t z[4];
      double v[4][P];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        z[u] = *(const f32x4_t*)(Z + (i + u * R) * kp + 4 * c4);
#pragma unroll
        for (int q = 0; q < P; q++) v[u][q] = V[q * ldv + i + u * R];
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int q = 0; q < P; q++)
#pragma unroll
          for (int j = 0; j < 4; j++) acc[q][j] += v[u][q] * (double)z[u][j];
    }
    for (; i < r1; i += R) {
      f32x4_t z = *(const f32x4_t*)(Z + i * kp + 4 * c4);
#pragma unroll
      for (int q = 0; q < P; q++) {
        double vq = V[q * ldv + i];
#pragma unroll
        for (int j = 0; j < 4; j++) acc[q][j] += vq * (double)z[j];
      }
    }
  }
  // combine the R row slots: shm[slot][q][c]
  if (active)
#pragma unroll
    for (int q = 0; q < P; q++)
#pragma unroll
      for (int j = 0; j < 4; j++) shm[(slot * P + q) * kp + 4 * c4 + j] = acc[q][j];
  __syncthreads();
  for (int e = threadIdx.x; e < P * kp; e += 256) {
    double s = 0.0;
    for (int sl = 0; sl < R; sl++) s += shm[sl * P * kp + e];
    part[blockIdx.x * (int64_t)(P * kp) + e] = s;
  }
}

static int gemv_blocks(Ctx* c, int64_t n) {
  int64_t b = (int64_t)c->cu_count * 2;
  int64_t mx = ceil_div(n, 64);
  if (b > mx) b = mx;
  return (int)(b < 1 ? 1 : b);
}

tsvm_status ipm_gemv_t(Ctx* c, const float* Z, int64_t n, int kp, const double* V, int64_t ldv,
                       int p, double* T) {
  if (kp % 32 != 0 || kp > 1024 || p < 1 || p > 4)
    return fail(c, TSVM_ERR_ARG, "gemv_t: kp=%d p=%d unsupported", kp, p);
  // two vectors: the wide fused kernel (the form below reads Z at half speed for p = 2)
  if (kp >= 256 && p == 2 && ldv == n)
    return ipm_zpass(c, ZP_T, p, p, Z, n, kp, nullptr, V, nullptr, nullptr, nullptr, nullptr,
                     nullptr, nullptr, 0, nullptr, nullptr, T);
  const int nblk = gemv_blocks(c, n);
  const int64_t rpb = ceil_div(n, nblk);
  void* part;
  TSVM_HIP(c, c->ws.get("gemvt_part", (size_t)nblk * p * kp * 8, &part));
  const int R = 256 / (kp / 4);
  size_t sm = (size_t)R * p * kp * 8;
  dim3 g(nblk), b(256);
  switch (p) {
    case 1: hipLaunchKernelGGL(gemv_t_kernel<1>, g, b, sm, c->stream, Z, n, kp, V, ldv, rpb, (double*)part); break;
    case 2: hipLaunchKernelGGL(gemv_t_kernel<2>, g, b, sm, c->stream, Z, n, kp, V, ldv, rpb, (double*)part); break;
    case 3: hipLaunchKernelGGL(gemv_t_kernel<3>, g, b, sm, c->stream, Z, n, kp, V, ldv, rpb, (double*)part); break;
    default: hipLaunchKernelGGL(gemv_t_kernel<4>, g, b, sm, c->stream, Z, n, kp, V, ldv, rpb, (double*)part); break;
  }
  TSVM_HIP(c, hipGetLastError());
  return reduce_parts(c, (const double*)part, nblk, p * kp, T);
}

// ------------------------------------------------------------------------------------
// Z T: G = kp/4 float4 groups per row, L = the largest power of two <= min(64, G) lanes
// per row (the lane segment is reduced with xor shuffles), each lane F = ceil(G/L) groups
// (groups past G contribute zero).  T columns of a lane stay in registers.
// ------------------------------------------------------------------------------------
template <int P, int F>
__global__ __launch_bounds__(256) void gemv_n_kernel(const float* __restrict__ Z, int64_t n, int kp,
                                                     const double* __restrict__ T,
                                                     double* __restrict__ E, int64_t lde) {
  const int G = kp / 4;
  int L = 64;
  while (L > G) L >>= 1;
  const int RW = 64 / L;  // rows per wave per step
  const int lane = threadIdx.x & 63;
  const int seg = lane / L, li = lane % L;
  double t[P][F][4];
  int grp[F];  // float4 group of each f, clamped to a valid one (its t is zero)
#pragma unroll
  for (int f = 0; f < F; f++) {
    const int g = li + f * L;
    grp[f] = g < G ? g : 0;
#pragma unroll
    for (int q = 0; q < P; q++)
#pragma unroll
      for (int j = 0; j < 4; j++) t[q][f][j] = g < G ? T[q * kp + 4 * g + j] : 0.0;
  }
  const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / 64;
  const int64_t nwaves = (int64_t)gridDim.x * blockDim.x / 64;
  for (int64_t base = wave * RW; base < n; base += nwaves * RW * 2) {
    f32x4_t z[2][F];
    int64_t rows[2];
#pragma unroll
    for (int u = 0; u < 2; u++) {
      rows[u] = base + u * nwaves * RW + seg;
      int64_t rr = rows[u] < n ? rows[u] : n - 1;
#pragma unroll
      for (int f = 0; f < F; f++) z[u][f] = *(const f32x4_t*)(Z + rr * kp + 4 * grp[f]);
    }
#pragma unroll
    for (int u = 0; u < 2; u++) {
      double acc[P];
#pragma unroll
      for (int q = 0; q < P; q++) {
        double s = 0.0;
#pragma unroll
        for (int f = 0; f < F; f++)
#pragma unroll
          for (int j = 0; j < 4; j++) s += (double)z[u][f][j] * t[q][f][j];
        acc[q] = s;
      }
#pragma unroll
      for (int q = 0; q < P; q++)
        for (int o = L / 2; o > 0; o >>= 1) acc[q] += __shfl_xor(acc[q], o, 64);
      if (li == 0 && rows[u] < n)
#pragma unroll
        for (int q = 0; q < P; q++) E[q * lde + rows[u]] = acc[q];
    }
  }
}

tsvm_status ipm_gemv_n(Ctx* c, const float* Z, int64_t n, int kp, const double* T, int p,
                       double* E, int64_t lde) {
  if (kp % 32 != 0 || kp > 1024 || p < 1 || p > 2)
    return fail(c, TSVM_ERR_ARG, "gemv_n: kp=%d p=%d unsupported", kp, p);
  if (kp >= 256 && p == 1 && lde == n)
    return ipm_zpass(c, ZP_N, 1, 0, Z, n, kp, T, nullptr, nullptr, nullptr, nullptr, nullptr,
                     nullptr, nullptr, 0, E, nullptr, nullptr);
  int L = 64;
  while (L > kp / 4) L >>= 1;
  const int F = (int)ceil_div(kp / 4, L);  // <= 4 for kp <= 1024
  int64_t blocks = (int64_t)c->cu_count * 8;
  int64_t need = ceil_div(n, 4 * (64 / L) * 2);
  if (blocks > need) blocks = need < 1 ? 1 : need;
  dim3 g((unsigned)blocks), b(256);
#define GN(PP, FF) hipLaunchKernelGGL((gemv_n_kernel<PP, FF>), g, b, 0, c->stream, Z, n, kp, T, E, lde)
  if (p == 1) {
    if (F == 1) GN(1, 1); else if (F == 2) GN(1, 2); else if (F == 3) GN(1, 3); else GN(1, 4);
  } else {
    if (F == 1) GN(2, 1); else if (F == 2) GN(2, 2); else if (F == 3) GN(2, 3); else GN(2, 4);
  }
#undef GN
  TSVM_HIP(c, hipGetLastError());
  return TSVM_OK;
}

// ------------------------------------------------------------------------------------
// Fused passes over Z.  One read of each row z_i serves three things: the products
// z_i . T_q (the layout of gemv_n_kernel), the row-wise update that consumes them, and the
// accumulation of Z^T u for the vector u that the update produces (per-block partials,
// reduced by reduce_parts in a fixed order).
//   ZP_SMW     last step of the SMW solve (1828-1838): sol_q = V_q - (z_i . T_q) / D_i for
//              q < P with V_q = D^-1 rhs_q and T_q = M^-1 Z^T V_q; reduces a^T sol_q;
//              Q = P also accumulates Z^T sol_q, Q = 0 skips it
//   ZP_BORDER  residual of the bordered Newton system for the step dx, with T = Z^T dx:
//              r5'_i = D_i dx_i + z_i . T - a_i dy + r5_i (= -rho_1); writes
//              V = r5' / D (the next SMW solve's scaled right-hand side), accumulates
//              Z^T V and reduces a^T dx, |r5'|^2, |r5|^2
// zpass_kernel serves kp < 256: L < 64 lanes share a row as in gemv_n_kernel, each lane
// loads the per-row operands of its row.  ZP_U rows are in flight per lane.
// ------------------------------------------------------------------------------------
constexpr int ZP_U = 4;

template <int MODE, int P, int Q, int F>
__global__ __launch_bounds__(256) void zpass_kernel(
    const float* __restrict__ Z, int64_t n, int kp, const double* __restrict__ T,
    const double* __restrict__ V, const double* __restrict__ winv, const double* __restrict__ a,
    const double* __restrict__ D, const double* __restrict__ r5, const double* __restrict__ dxv,
    const double* __restrict__ sc, int corr, double* __restrict__ out, double* __restrict__ red,
    double* __restrict__ part, const ResArgs ra) {
  extern __shared__ double shz[];  // 2 x Q x kp: waves 0, 1 store, waves 2, 3 add
  constexpr int R = MODE == ZP_SMW ? P : 3;
  constexpr int QA = Q > 0 ? Q : 1;
  constexpr int NO = MODE == ZP_SMW ? P + 2 : 5;  // per-row operands
  const int G = kp / 4;
  int L = 64;
  while (L > G) L >>= 1;
  const int RW = 64 / L;  // rows per wave per step
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int seg = lane / L, li = lane % L;
  double t[P][F][4], acc[QA][F][4];
  int grp[F];  // float4 group of each f, clamped to a valid one (its t is zero)
#pragma unroll
  for (int f = 0; f < F; f++) {
    const int g = li + f * L;
    grp[f] = g < G ? g : 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
#pragma unroll
      for (int q = 0; q < P; q++) t[q][f][j] = g < G ? T[q * kp + 4 * g + j] : 0.0;
#pragma unroll
      for (int q = 0; q < QA; q++) acc[q][f][j] = 0.0;
    }
  }
  double rv[R];
  int op[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    rv[r] = 0.0;
    op[r] = RED_SUM;
  }
  const double dy = MODE == ZP_BORDER ? sc[corr ? SL_DYC : SL_DY] : 0.0;
  const int64_t wave = (int64_t)blockIdx.x * 4 + wid;
  const int64_t step = (int64_t)gridDim.x * 4 * RW;
  for (int64_t base = wave * RW + seg; base < n + seg; base += step * ZP_U) {
    f32x4_t z[ZP_U][F];
    double o[ZP_U][NO];
    int64_t row[ZP_U];
#pragma unroll
    for (int u = 0; u < ZP_U; u++) {
      row[u] = base + u * step;
      const int64_t rr = row[u] < n ? row[u] : n - 1;
#pragma unroll
      for (int f = 0; f < F; f++) z[u][f] = *(const f32x4_t*)(Z + rr * kp + 4 * grp[f]);
      if constexpr (MODE == ZP_SMW) {
#pragma unroll
        for (int q = 0; q < P; q++) o[u][q] = V[q * n + rr];
        o[u][P] = winv[rr];
        o[u][P + 1] = a[rr];
      } else {
        o[u][0] = D[rr];
        o[u][1] = dxv[rr];
        o[u][2] = winv[rr];
        o[u][3] = a[rr];
        o[u][4] = r5[rr];
      }
    }
#pragma unroll
    for (int u = 0; u < ZP_U; u++) {
      double d[P];
#pragma unroll
      for (int q = 0; q < P; q++) {
        double s = 0.0;
#pragma unroll
        for (int f = 0; f < F; f++)
#pragma unroll
          for (int j = 0; j < 4; j++) s += (double)z[u][f][j] * t[q][f][j];
        d[q] = s;
      }
#pragma unroll
      for (int q = 0; q < P; q++)
        for (int m = L / 2; m > 0; m >>= 1) d[q] += __shfl_xor(d[q], m, 64);
      const bool ok = row[u] < n;
      const bool wr = ok && li == 0;
      double uq[QA];
      if constexpr (MODE == ZP_SMW) {
#pragma unroll
        for (int q = 0; q < P; q++) {
          const double sq = o[u][q] - d[q] * o[u][P];
          if (wr) {
            out[q * n + row[u]] = sq;
            rv[q] += o[u][P + 1] * sq;
          }
          if (q < QA) uq[q] = ok ? sq : 0.0;
        }
      } else {
        const double r = o[u][0] * o[u][1] + d[0] - o[u][3] * dy + o[u][4];
        const double vv = r * o[u][2];
        if (wr) {
          out[row[u]] = vv;
          rv[0] += o[u][3] * o[u][1];
          rv[1] += r * r;
          rv[2] += o[u][4] * o[u][4];
        }
        uq[0] = ok ? vv : 0.0;
      }
      if constexpr (Q > 0) {
#pragma unroll
        for (int q = 0; q < Q; q++)
#pragma unroll
          for (int f = 0; f < F; f++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[q][f][j] += uq[q] * (double)z[u][f][j];
      }
    }
  }
  if constexpr (Q > 0) {
    // rows of the RW segments of a wave, then the four waves (two LDS rounds)
#pragma unroll
    for (int q = 0; q < Q; q++)
#pragma unroll
      for (int f = 0; f < F; f++)
#pragma unroll
        for (int j = 0; j < 4; j++)
          for (int m = L; m < 64; m <<= 1) acc[q][f][j] += __shfl_xor(acc[q][f][j], m, 64);
    const int half = wid & 1;
    for (int rnd = 0; rnd < 2; rnd++) {
      if ((wid >> 1) == rnd && seg == 0) {
#pragma unroll
        for (int q = 0; q < Q; q++)
#pragma unroll
          for (int f = 0; f < F; f++) {
            const int g = li + f * L;
            if (g < G)
#pragma unroll
              for (int j = 0; j < 4; j++) {
                double* dst = &shz[(half * Q + q) * kp + 4 * g + j];
                *dst = rnd == 0 ? acc[q][f][j] : *dst + acc[q][f][j];
              }
          }
      }
      __syncthreads();
    }
    for (int e = threadIdx.x; e < Q * kp; e += 256)
      part[(int64_t)blockIdx.x * Q * kp + e] = shz[e] + shz[Q * kp + e];
  }
  block_reduce_store<R>(rv, op, red);
}

__device__ inline double f64_from(int lo, int hi) {
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline double readlane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  return f64_from(__builtin_amdgcn_readlane((int)b, l), __builtin_amdgcn_readlane((int)(b >> 32), l));
}
// value of a DPP-selected lane (all rows and banks enabled, every source lane exists)
template <int CTRL>
__device__ inline double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  return f64_from(__builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false),
                  __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false));
}
constexpr int DPP_XOR1 = 0xB1;         // quad_perm [1, 0, 3, 2]
constexpr int DPP_XOR2 = 0x4E;         // quad_perm [2, 3, 0, 1]
constexpr int DPP_HALF_MIRROR = 0x141; // lane i <- 7 - i within 8
constexpr int DPP_MIRROR = 0x140;      // lane i <- 15 - i within 16
constexpr int DPP_ROR8 = 0x128;        // lane i <- (i + 8) mod 16 within 16
// v_permlane32_swap: lanes 0-31 return a(l) + a(l + 32), lanes 32-63 b(l - 32) + b(l)
__device__ inline double swap_sum32(double a, double b) {
  const long long A = __double_as_longlong(a), B = __double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)A, (unsigned)B, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(A >> 32), (unsigned)(B >> 32), false, false);
  return f64_from(lo[0], hi[0]) + f64_from(lo[1], hi[1]);
}
// v_permlane16_swap: rows (16 lanes) 0 and 2 return a(l) + a(l + 16), rows 1 and 3
// b(l - 16) + b(l)
__device__ inline double swap_sum16(double a, double b) {
  const long long A = __double_as_longlong(a), B = __double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)A, (unsigned)B, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(A >> 32), (unsigned)(B >> 32), false, false);
  return f64_from(lo[0], hi[0]) + f64_from(lo[1], hi[1]);
}

// The passes for kp >= 256, where a wave owns whole rows.  Each wave takes ZP_U consecutive
// rows per step, so that the per-row operands and results of a step share cache lines.
// Each lane holds F float4 groups of the step's rows.  The P x ZP_U partial products are
// summed over the wave by a reduce-scatter (lane bit 5 keeps half of the set, bit 4 half
// of the rest, ... by permlane swaps and DPP moves, no LDS), after which lane l holds product
// m = l >> (6 - log2(P ZP_U)) (q = m / ZP_U, row u = m % ZP_U) and has loaded that row's
// operands itself, so the row-wise update runs for all rows at once.  The updated values
// return to every lane by readlane for the Z^T accumulation.  Besides ZP_SMW and
// ZP_BORDER:
//   ZP_T       Z^T V_q for q < P (Q = P), no products
//   ZP_N       out = Z T (P = 1, Q = 0)
//   ZP_RES     the residual pass with T = Z^T x (1606-1643) fused with the predictor's
//              right-hand sides (1805-1808), as k_resid + k_rhs_pred + k_scale: e = Z T,
//              r = e - a y - s + xi - 1, ds = -s x, dxi = -xi (C - x), D, 1/D and the
//              nine stop-test reductions, r5 = r - ds/x + dxi/(C - x), then V = [r5, a]/D
//              and Z^T V (P = 1, Q = 2)
template <int MODE, int P, int Q, int F>
__global__ __launch_bounds__(256) void zpass_wide_kernel(
    const float* __restrict__ Z, int64_t n, int kp, const double* __restrict__ T,
    const double* __restrict__ V, const double* __restrict__ winv, const double* __restrict__ a,
    const double* __restrict__ D, const double* __restrict__ r5, const double* __restrict__ dxv,
    const double* __restrict__ sc, int corr, double* __restrict__ out, double* __restrict__ red,
    double* __restrict__ part, const ResArgs ra) {
  extern __shared__ double shz[];  // 2 x Q x kp: waves 0, 1 store, waves 2, 3 add
  constexpr bool DOT = MODE != ZP_T;
  constexpr bool RED = MODE == ZP_SMW || MODE == ZP_BORDER || MODE == ZP_RES;
  constexpr int R = MODE == ZP_BORDER ? 3 : (MODE == ZP_RES ? 9 : P);
  constexpr int QA = Q > 0 ? Q : 1;
  constexpr int M = P * ZP_U;  // 4 or 8 products per step
  constexpr int LOGM = M == 4 ? 2 : 3;
  constexpr int GL = 64 >> LOGM;  // lanes per product after the reduction
  constexpr int NO = MODE == ZP_BORDER ? 5 : (MODE == ZP_SMW ? 3 : (MODE == ZP_RES ? 4 : 1));
  static_assert(M == 4 || M == 8, "P * ZP_U must be 4 or 8");
  static_assert(MODE != ZP_RES || (P == 1 && Q == 2), "the residual pass is P = 1, Q = 2");
  const int G = kp / 4;
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ml = lane >> (6 - LOGM);
  const int ql = ml / ZP_U, ul = ml % ZP_U;
  const bool lead = (lane & (GL - 1)) == 0;
  double t[DOT ? P : 1][F][4], acc[QA][F][4];
  int grp[F];
#pragma unroll
  for (int f = 0; f < F; f++) {
    const int g = lane + f * 64;
    grp[f] = g < G ? g : 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if constexpr (DOT)
#pragma unroll
        for (int q = 0; q < P; q++) t[q][f][j] = g < G ? T[q * kp + 4 * g + j] : 0.0;
#pragma unroll
      for (int q = 0; q < QA; q++) acc[q][f][j] = 0.0;
    }
  }
  double rv[R];
  int op[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    rv[r] = 0.0;
    op[r] = RED_SUM;
  }
  if constexpr (MODE == ZP_RES) {
    rv[7] = -INFINITY;
    op[7] = RED_MAX;
    rv[8] = INFINITY;
    op[8] = RED_MIN;
  }
  const double dy = MODE == ZP_BORDER ? sc[corr ? SL_DYC : SL_DY] : 0.0;
  const double yeq = MODE == ZP_RES ? sc[SL_Y] : 0.0;
  // The grid sweeps Z as one front: a block step is 4 ZP_U consecutive rows (the block's
  // waves take ZP_U each), block b takes block steps b, b + grid, ...  (measured at
  // n = 1.6M, kp = 512 against contiguous row ranges per wave: 1-3% faster)
  constexpr int64_t SB = 4 * ZP_U;
  const int64_t r0 = (int64_t)blockIdx.x * SB + wid * ZP_U;
  const int64_t r1 = n;
  const int64_t stride = (int64_t)gridDim.x * SB;
  // Software pipeline: the rows and per-row operands of the next step are loaded before
  // the current step's products, reductions and updates run (two register sets, the loop
  // unrolled by two so that they alternate without copies)
  auto load_step = [&](int64_t base, f32x4_t(&z)[ZP_U][F], double(&o)[NO]) {
#pragma unroll
    for (int u = 0; u < ZP_U; u++) {
      const int64_t rr = base + u < n ? base + u : n - 1;
#pragma unroll
      for (int f = 0; f < F; f++) z[u][f] = *(const f32x4_t*)(Z + rr * kp + 4 * grp[f]);
    }
    const int64_t rl = base + ul < n ? base + ul : n - 1;
    if constexpr (MODE == ZP_BORDER) {
      o[0] = D[rl];
      o[1] = dxv[rl];
      o[2] = winv[rl];
      o[3] = a[rl];
      o[4] = r5[rl];
    } else if constexpr (MODE == ZP_SMW) {
      o[0] = V[ql * n + rl];
      o[1] = winv[rl];
      o[2] = a[rl];
    } else if constexpr (MODE == ZP_T) {
      o[0] = V[ql * n + rl];
    } else if constexpr (MODE == ZP_RES) {
      o[0] = ra.x[rl];
      o[1] = ra.s[rl];
      o[2] = ra.xi[rl];
      o[3] = a[rl];
    }
  };
  // rows base .. base + ZP_U - 1 (rows past r1 were clamped for the loads and are masked)
  auto step = [&](int64_t base, const f32x4_t(&z)[ZP_U][F], const double(&o)[NO]) {
    const int64_t myrow = base + ul;
    const bool ok = myrow < r1;
    double w, w1 = 0.0;  // this lane's product(s) after the row-wise update
    if constexpr (DOT) {
      double v[M];
#pragma unroll
      for (int q = 0; q < P; q++)
#pragma unroll
        for (int u = 0; u < ZP_U; u++) {
          double s = 0.0;
#pragma unroll
          for (int f = 0; f < F; f++)
#pragma unroll
            for (int j = 0; j < 4; j++) s += (double)z[u][f][j] * t[q][f][j];
          v[q * ZP_U + u] = s;
        }
      // reduce-scatter over lane bits 5, 4 (and 3 for M = 8), then sums within GL lanes
#pragma unroll
      for (int i = 0; i < M / 2; i++) v[i] = swap_sum32(v[i], v[i + M / 2]);
#pragma unroll
      for (int i = 0; i < M / 4; i++) v[i] = swap_sum16(v[i], v[i + M / 4]);
      if constexpr (M == 8) {
        const bool hi = (lane & 8) != 0;
        v[0] = (hi ? v[1] : v[0]) + dpp_f64<DPP_ROR8>(hi ? v[0] : v[1]);
      }
      v[0] += dpp_f64<DPP_XOR1>(v[0]);
      v[0] += dpp_f64<DPP_XOR2>(v[0]);
      v[0] += dpp_f64<DPP_HALF_MIRROR>(v[0]);
      if constexpr (M == 4) v[0] += dpp_f64<DPP_MIRROR>(v[0]);
      if constexpr (MODE == ZP_SMW) {
        w = o[0] - v[0] * o[1];
        if (ok && lead) {
          out[ql * n + myrow] = w;
#pragma unroll
          for (int q = 0; q < P; q++)
            if (q == ql) rv[q] += o[2] * w;
        }
      } else if constexpr (MODE == ZP_BORDER) {
        const double r = o[0] * o[1] + v[0] - o[3] * dy + o[4];
        w = r * o[2];
        if (ok && lead) {
          out[myrow] = w;
          rv[0] += o[3] * o[1];
          rv[1] += r * r;
          rv[2] += o[4] * o[4];
        }
      } else if constexpr (MODE == ZP_RES) {
        const double xv = o[0], sv = o[1], xiv = o[2], av = o[3], ev = v[0];
        const double cx = ra.C - xv;
        const double r = ev - av * yeq - sv + xiv - 1.0;
        const double dsv = -sv * xv, dxiv = -xiv * cx;
        const double Dv = sv / xv + xiv / cx;
        const double wv = 1.0 / Dv;
        const double r5v = r - dsv / xv + dxiv / cx;
        w = r5v * wv;
        w1 = av * wv;
        if (ok && lead) {
          ra.dx[myrow] = r;
          ra.ds[myrow] = dsv;
          ra.dxi[myrow] = dxiv;
          ra.D[myrow] = Dv;
          ra.winv[myrow] = wv;
          ra.rhs[myrow] = r5v;
          ra.V[myrow] = w;
          ra.V[n + myrow] = w1;
          rv[0] += xv * sv + cx * xiv;
          rv[1] += r * r;
          rv[2] += av * xv;
          rv[3] += xv * ev;
          rv[4] += xv;
          rv[5] += xiv;
          rv[6] += fabs(xv);
          rv[7] = fmax(rv[7], Dv);
          rv[8] = fmin(rv[8], Dv);
        }
      } else {
        w = v[0];
        if (ok && lead) out[myrow] = w;
      }
    } else {
      w = o[0];
    }
    if constexpr (Q > 0) {
#pragma unroll
      for (int u = 0; u < ZP_U; u++) {
        double uq[Q];
#pragma unroll
        for (int q = 0; q < Q; q++)
          // ZP_RES: both values of row u sit in its lane group
          uq[q] = base + u < r1 ? (MODE == ZP_RES ? readlane_f64(q == 0 ? w : w1, u * GL)
                                                  : readlane_f64(w, (q * ZP_U + u) * GL))
                                : 0.0;
#pragma unroll
        for (int f = 0; f < F; f++)
#pragma unroll
          for (int j = 0; j < 4; j++) {
            // converted again rather than kept from the dot products: the FP64 copies of
            // the step's rows would otherwise stay live across the reduction (64 VGPRs)
            float zf = z[u][f][j];
            asm("" : "+v"(zf));
            const double zd = (double)zf;
#pragma unroll
            for (int q = 0; q < Q; q++) acc[q][f][j] += uq[q] * zd;
          }
      }
    }
  };
  f32x4_t za[ZP_U][F], zb[ZP_U][F];
  double oa[NO], ob[NO];
  if (r0 < r1) load_step(r0, za, oa);
  for (int64_t base = r0; base < r1; base += 2 * stride) {
    if (base + stride < r1) load_step(base + stride, zb, ob);
    step(base, za, oa);
    if (base + stride >= r1) break;
    if (base + 2 * stride < r1) load_step(base + 2 * stride, za, oa);
    step(base + stride, zb, ob);
  }
  if constexpr (Q > 0) {
    const int half = wid & 1;
    for (int rnd = 0; rnd < 2; rnd++) {
      if ((wid >> 1) == rnd) {
#pragma unroll
        for (int q = 0; q < Q; q++)
#pragma unroll
          for (int f = 0; f < F; f++) {
            const int g = lane + f * 64;
            if (g < G)
#pragma unroll
              for (int j = 0; j < 4; j++) {
                double* dst = &shz[(half * Q + q) * kp + 4 * g + j];
                *dst = rnd == 0 ? acc[q][f][j] : *dst + acc[q][f][j];
              }
          }
      }
      __syncthreads();
    }
    for (int e = threadIdx.x; e < Q * kp; e += 256)
      part[(int64_t)blockIdx.x * Q * kp + e] = shz[e] + shz[Q * kp + e];
  }
  if constexpr (RED) block_reduce_store<R>(rv, op, red);
}

// Blocks of a pass over Z: as many as the CUs hold at once (at most four per CU, from the
// kernel's register and LDS demand), so that every wave's chunk of rows is in flight from
// the start and no second round of blocks runs behind the first.
static int zpass_blocks(Ctx* c, const void* fn, size_t sm) {
  static std::mutex mu;
  static std::map<std::pair<const void*, size_t>, int> occ_of;
  int occ = 0;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = occ_of.find({fn, sm});
    if (it != occ_of.end()) occ = it->second;
  }
  if (occ == 0) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, 256, sm) != hipSuccess || occ < 1)
      occ = 1;
    if (occ > 4) occ = 4;
    std::lock_guard<std::mutex> lk(mu);
    occ_of[{fn, sm}] = occ;
  }
  const int nb = c->cu_count * occ;
  return nb < EW_BLOCKS ? nb : EW_BLOCKS;
}

static tsvm_status ipm_zpass(Ctx* c, int mode, int p, int q, const float* Z, int64_t n, int kp,
                             const double* T, const double* V, const double* winv,
                             const double* a, const double* D, const double* r5, const double* dx,
                             const double* sc, int corr, double* out, double* red, double* Tout,
                             const ResArgs* ra, int* nred) {
  const int G = kp / 4;
  const bool wide = G >= 64;
  bool ok_shape = false;
  if (mode == ZP_SMW) ok_shape = (p == 2 && q == 2) || (p == 1 && (q == 0 || q == 1));
  else if (mode == ZP_BORDER) ok_shape = p == 1 && q == 1;
  else if (mode == ZP_T) ok_shape = wide && (p == 1 || p == 2) && q == p;
  else if (mode == ZP_N) ok_shape = wide && p == 1 && q == 0;
  else if (mode == ZP_RES) ok_shape = wide && p == 1 && q == 2 && ra;
  if (kp % 32 != 0 || kp > 1024 || !ok_shape)
    return fail(c, TSVM_ERR_ARG, "zpass: kp=%d mode=%d p=%d q=%d unsupported", kp, mode, p, q);
  int L = 64;
  while (L > G) L >>= 1;
  const int F = (int)ceil_div(G, L);  // wide: <= 4, else <= 2
  const void* fn = nullptr;
#define ZK(KN, MO, PP, QQ, FF) fn = (const void*)&KN<MO, PP, QQ, FF>
#define ZW(MO, PP, QQ)                                         \
  if (F == 1) ZK(zpass_wide_kernel, MO, PP, QQ, 1);            \
  else if (F == 2) ZK(zpass_wide_kernel, MO, PP, QQ, 2);       \
  else if (F == 3) ZK(zpass_wide_kernel, MO, PP, QQ, 3);       \
  else ZK(zpass_wide_kernel, MO, PP, QQ, 4);
#define ZF(MO, PP, QQ)                                         \
  if (wide) {                                                  \
    ZW(MO, PP, QQ)                                             \
  } else {                                                     \
    if (F == 1) ZK(zpass_kernel, MO, PP, QQ, 1);               \
    else ZK(zpass_kernel, MO, PP, QQ, 2);                      \
  }
  if (mode == ZP_BORDER) {
    ZF(ZP_BORDER, 1, 1);
  } else if (mode == ZP_T) {
    if (p == 2) {
      ZW(ZP_T, 2, 2);
    } else {
      ZW(ZP_T, 1, 1);
    }
  } else if (mode == ZP_N) {
    ZW(ZP_N, 1, 0);
  } else if (mode == ZP_RES) {
    ZW(ZP_RES, 1, 2);
  } else if (p == 2) {
    ZF(ZP_SMW, 2, 2);
  } else if (q == 1) {
    ZF(ZP_SMW, 1, 1);
  } else {
    ZF(ZP_SMW, 1, 0);
  }
#undef ZF
#undef ZW
#undef ZK
  const size_t sm = (size_t)2 * q * kp * 8;
  const int nblk = zpass_blocks(c, fn, sm);
  if (nred) *nred = nblk;
  double* part = nullptr;
  if (q > 0) {
    // sized for the largest grid and two vectors, so that it is allocated once
    const size_t mx = (size_t)(c->cu_count * 4 < EW_BLOCKS ? c->cu_count * 4 : EW_BLOCKS);
    void* pv;
    TSVM_HIP(c, c->ws.get("zpass_part", (nblk > (int)mx ? (size_t)nblk : mx) * 2 * kp * 8, &pv));
    part = (double*)pv;
  }
  ResArgs rargs = ra ? *ra : ResArgs{};
  void* args[] = {(void*)&Z,  (void*)&n,   (void*)&kp,  (void*)&T,  (void*)&V,     (void*)&winv,
                  (void*)&a,  (void*)&D,   (void*)&r5,  (void*)&dx, (void*)&sc,    (void*)&corr,
                  (void*)&out, (void*)&red, (void*)&part, (void*)&rargs};
  TSVM_HIP(c, hipLaunchKernel(fn, dim3(nblk), dim3(256), args, sm, c->stream));
  if (q > 0) TSVM_CHECK(reduce_parts(c, (const double*)part, nblk, q * kp, Tout));
  return TSVM_OK;
}

// ------------------------------------------------------------------------------------
// fused O(n) passes
// ------------------------------------------------------------------------------------
#define GRID_LOOP(i, n) \
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

// residual, D, stop-test reductions (1606-1643)
__global__ __launch_bounds__(256) void k_resid(int64_t n, double C, const double* __restrict__ x,
                                               const double* __restrict__ s, const double* __restrict__ xi,
                                               const double* __restrict__ a, const double* __restrict__ e,
                                               const double* __restrict__ sc, double* __restrict__ dx,
                                               double* __restrict__ ds, double* __restrict__ dxi,
                                               double* __restrict__ D, double* __restrict__ winv,
                                               double* __restrict__ red) {
  const double y = sc[SL_Y];
  double v[9] = {0, 0, 0, 0, 0, 0, 0, -INFINITY, INFINITY};
  const int op[9] = {RED_SUM, RED_SUM, RED_SUM, RED_SUM, RED_SUM, RED_SUM, RED_SUM, RED_MAX, RED_MIN};
  GRID_LOOP(i, n) {
    double xv = x[i], sv = s[i], xiv = xi[i], av = a[i], ev = e[i];
    double cx = C - xv;
    double r = ev - av * y - sv + xiv - 1.0;
    dx[i] = r;
    ds[i] = -sv * xv;
    dxi[i] = -xiv * cx;
    double Dv = sv / xv + xiv / cx;
    D[i] = Dv;
    winv[i] = 1.0 / Dv;
    v[0] += xv * sv + cx * xiv;
    v[1] += r * r;
    v[2] += av * xv;
    v[3] += xv * ev;
    v[4] += xv;
    v[5] += xiv;
    v[6] += fabs(xv);
    v[7] = fmax(v[7], Dv);
    v[8] = fmin(v[8], Dv);
  }
  block_reduce_store<9>(v, op, red);
}

// predictor right-hand sides: rhs0 = r5 = r1 - r3/x + r4/(C-x), rhs1 = a (1805-1808)
__global__ void k_rhs_pred(int64_t n, double C, const double* __restrict__ x,
                           const double* __restrict__ dx, const double* __restrict__ ds,
                           const double* __restrict__ dxi, const double* __restrict__ a,
                           double* __restrict__ rhs, int64_t ldr) {
  GRID_LOOP(i, n) {
    double xv = x[i];
    rhs[i] = dx[i] - ds[i] / xv + dxi[i] / (C - xv);
    rhs[ldr + i] = a[i];
  }
}

// V = rhs / D (p vectors)
__global__ void k_scale(int64_t n, int p, const double* __restrict__ rhs, int64_t ldr,
                        const double* __restrict__ winv, double* __restrict__ V, int64_t ldv) {
  GRID_LOOP(i, n) {
    double w = winv[i];
    for (int q = 0; q < p; q++) V[q * ldv + i] = rhs[q * ldr + i] * w;
  }
}

// third solve of the Newton step (1815-1816): dx = S^-1 (a dy - r5) with
// dy = (r2 + a^T r7) / (a^T b), r2 = -a^T x (predictor) or 0 (corrector).  The SMW
// operator S^-1 = (D + Z Z^T)^-1 (with the capacitance as factored) is linear and
// r7 = S^-1 r5, b = S^-1 a are already at hand, so dx = dy b - r7 without another solve.
// With Tr7 = Z^T r7 and Tb = Z^T b given (kp entries), block 0 also forms
// T3 = Z^T dx = dy Tb - Tr7 for the bordered refinement.
__global__ void k_sol3_lin(int64_t n, int corr, const double* __restrict__ b,
                           const double* __restrict__ r7, double* __restrict__ sc,
                           double* __restrict__ sol3, int kp, const double* __restrict__ Tr7,
                           const double* __restrict__ Tb, double* __restrict__ T3) {
  const double r2 = corr ? 0.0 : -sc[SL_AX];
  const double dy = (r2 + sc[corr ? SL_AR7C : SL_AR7]) / sc[SL_AB];
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[corr ? SL_DYC : SL_DY] = dy;
  if (Tr7 && blockIdx.x == 0)
    for (int e = threadIdx.x; e < kp; e += blockDim.x) T3[e] = dy * Tb[e] - Tr7[e];
  GRID_LOOP(i, n) sol3[i] = dy * b[i] - r7[i];
}

// bordered correction: (ddx, ddy) solve the same bordered system for the residual
// (r5', rho2): ddy = (rho2 + a^T S^-1 r5') / a^T b, ddx = S^-1 (a ddy - r5') = ddy b - s5
// with s5 = S^-1 r5' and b = S^-1 a (S^-1 linear); dx += ddx, dy += ddy.  With Ts5 =
// Z^T s5 given, block 0 also updates T3 = Z^T dx by ddy Tb - Ts5 (for a further sweep).
// corr: 0 the affine step, 1 the corrector's, 2 the combined step (whose r2 is the affine
// one's, the corrector's being 0).
__global__ void k_border_lin(int64_t n, int corr, const double* __restrict__ b,
                             const double* __restrict__ s5, double* __restrict__ sc,
                             double* __restrict__ sol3, int kp, const double* __restrict__ Ts5,
                             const double* __restrict__ Tb, double* __restrict__ T3) {
  const double r2 = corr == 1 ? 0.0 : -sc[SL_AX];
  const double rho2 = r2 - sc[SL_R2B];
  const double ddy = (rho2 + sc[SL_AR7B]) / sc[SL_AB];
  GRID_LOOP(i, n) sol3[i] += ddy * b[i] - s5[i];
  if (Ts5 && blockIdx.x == 0)
    for (int e = threadIdx.x; e < kp; e += blockDim.x) T3[e] += ddy * Tb[e] - Ts5[e];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[SL_DYB] = ddy;
    sc[corr ? SL_DYC : SL_DY] += ddy;
  }
}

// Newton step completion (1817-1821) and step bound (1733-1740 / 1762-1768):
//   ds = (r3 - s dx)/x;  dxi = (r4 + xi dx)/(C-x), dx from k_sol3_lin
// corr = 0: predictor, results to (dx, ds, dxi) in place.
// corr = 1: corrector; the total step aff + corr goes to (dx, ds, dxi), the affine
//           step is read from (adx, ads, adxi).
__global__ __launch_bounds__(256) void k_step(int64_t n, double C, int corr, const double* __restrict__ x,
                                              const double* __restrict__ s, const double* __restrict__ xi,
                                              const double* __restrict__ dxs, double* __restrict__ dx,
                                              double* __restrict__ ds, double* __restrict__ dxi,
                                              const double* __restrict__ adx, const double* __restrict__ ads,
                                              const double* __restrict__ adxi, double* __restrict__ red) {
  double v[1] = {1.0};
  const int op[1] = {RED_MIN};
  GRID_LOOP(i, n) {
    double xv = x[i], cx = C - xv;
    double ddx = dxs[i];
    double dds = (ds[i] - s[i] * ddx) / xv;
    double ddxi = (dxi[i] + xi[i] * ddx) / cx;
    if (corr) {
      ddx += adx[i];
      dds += ads[i];
      ddxi += adxi[i];
    }
    dx[i] = ddx;
    ds[i] = dds;
    dxi[i] = ddxi;
    double al = v[0];
    if (ddx < 0) al = fmin(al, -xv / ddx);
    else if (ddx > 0) al = fmin(al, cx / ddx);
    if (dds < 0) al = fmin(al, -s[i] / dds);
    if (ddxi < 0) al = fmin(al, -xi[i] / ddxi);
    v[0] = al;
  }
  block_reduce_store<1>(v, op, red);
}

// complementarity at the predictor point (1742-1745)
__global__ __launch_bounds__(256) void k_muaff(int64_t n, double C, const double* __restrict__ x,
                                               const double* __restrict__ s, const double* __restrict__ xi,
                                               const double* __restrict__ dx, const double* __restrict__ ds,
                                               const double* __restrict__ dxi, const double* __restrict__ sc,
                                               double* __restrict__ red) {
  const double al = sc[SL_ALPHA_AFF];
  double v[1] = {0.0};
  const int op[1] = {RED_SUM};
  GRID_LOOP(i, n) {
    double xa = x[i] + al * dx[i];
    v[0] += xa * (s[i] + al * ds[i]) + (C - xa) * (xi[i] + al * dxi[i]);
  }
  block_reduce_store<1>(v, op, red);
}

// corrector right-hand sides (1750-1757): keep the affine step, r3 = sigma mu - dx ds,
// r4 = sigma mu + dx dxi, r1 = r2 = 0, rhs0 = r5 = -r3/x + r4/(C-x).
// comb: r3, r4, rhs0 are left as the sums of the affine right-hand sides (r3 = -s x,
// r4 = -xi (C - x), r5 in rhs0) and the corrector's, the right-hand sides of the combined
// step (and V = r5 / D of the corrector's part when V is given).
__global__ void k_rhs_corr(int64_t n, double C, const double* __restrict__ x,
                           double* __restrict__ dx, double* __restrict__ ds, double* __restrict__ dxi,
                           double* __restrict__ adx, double* __restrict__ ads, double* __restrict__ adxi,
                           const double* __restrict__ sc, double* __restrict__ rhs, int comb,
                           const double* __restrict__ s, const double* __restrict__ xi,
                           const double* __restrict__ winv, double* __restrict__ V) {
  const double n2 = 2.0 * (double)n;
  const double mu = sc[SL_MU] / n2, mua = sc[SL_MUAFF] / n2;
  double sig = mua / mu;
  sig = sig * sig * sig;
  const double sm = sig * mu;
  GRID_LOOP(i, n) {
    double a1 = dx[i], a3 = ds[i], a4 = dxi[i];
    adx[i] = a1;
    ads[i] = a3;
    adxi[i] = a4;
    double r3 = sm - a1 * a3, r4 = sm + a1 * a4;
    const double xv = x[i], cx = C - xv;
    const double r5 = -r3 / xv + r4 / cx;
    if (comb) {
      if (V) V[i] = r5 * winv[i];
      r3 -= s[i] * xv;
      r4 -= xi[i] * cx;
      rhs[i] += r5;
    } else {
      rhs[i] = r5;
    }
    ds[i] = r3;
    dxi[i] = r4;
  }
}

// x, s, xi += 0.99 alpha step; y += 0.99 alpha (dy + dyc)   (1769-1777); comb: the
// corrector's dy slot holds the combined step's.  Skipped when the capacitance Cholesky
// failed (the reference returns the current iterate, 1711-1714).
__global__ void k_update(int64_t n, double* __restrict__ x, double* __restrict__ s, double* __restrict__ xi,
                         const double* __restrict__ dx, const double* __restrict__ ds,
                         const double* __restrict__ dxi, double* __restrict__ sc, const int* __restrict__ info,
                         int comb) {
  if (*info != 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[SL_CHOLFAIL] = 1.0;
    return;
  }
  const double al = 0.99 * sc[SL_ALPHA];
  GRID_LOOP(i, n) {
    x[i] += al * dx[i];
    s[i] += al * ds[i];
    xi[i] += al * dxi[i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[SL_Y] += al * (comb ? sc[SL_DYC] : sc[SL_DY] + sc[SL_DYC]);
}

// Z^T x after k_update: x += 0.99 alpha (dx_aff + dx_corr), and T3a, T3 hold Z^T of the two
// steps (T3a null: T3 holds Z^T of the combined step); skipped, as k_update, after a failed
// factorisation
__global__ void k_carry_tx(int kp, double* __restrict__ Tx, const double* __restrict__ T3a,
                           const double* __restrict__ T3c, const double* __restrict__ sc,
                           const int* __restrict__ info) {
  if (*info != 0) return;
  const double al = 0.99 * sc[SL_ALPHA];
  for (int e = threadIdx.x; e < kp; e += blockDim.x) Tx[e] += al * ((T3a ? T3a[e] : 0.0) + T3c[e]);
}

// the combined step starting from the affine one: sol3 = dx_aff, T3 = Z^T dx_aff and the
// corrector's dy slot = dy_aff; the refinement sweeps against the combined right-hand sides
// then add the corrector's part
__global__ void k_start_affine(int64_t n, const double* __restrict__ adx, double* __restrict__ sol3,
                               int kp, const double* __restrict__ T3a, double* __restrict__ T3,
                               double* __restrict__ sc) {
  GRID_LOOP(i, n) sol3[i] = adx[i];
  if (blockIdx.x == 0) {
    for (int e = threadIdx.x; e < kp; e += blockDim.x) T3[e] = T3a[e];
    if (threadIdx.x == 0) sc[SL_DYC] = sc[SL_DY];
  }
}

// M += I (kp x kp) and copy lower to upper is not needed (Cholesky reads the lower part)
__global__ void k_add_identity(double* __restrict__ M, int64_t kp) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < kp) M[i * kp + i] += 1.0;
}

// initial point (1535-1579): x = clamp(1^T Q 1 / |Q 1|^2) (RBF) or clamp(1) (linear), s = xi = 1
__global__ __launch_bounds__(256) void k_init_red(int64_t n, const double* __restrict__ e,
                                                  double* __restrict__ red) {
  double v[2] = {0, 0};
  const int op[2] = {RED_SUM, RED_SUM};
  GRID_LOOP(i, n) {
    v[0] += e[i] * e[i];
    v[1] += e[i];
  }
  block_reduce_store<2>(v, op, red);
}
__global__ void k_fill(int64_t n, double* __restrict__ p, double v) {
  GRID_LOOP(i, n) p[i] = v;
}
__global__ void k_init_xs(int64_t n, double C, int linear, const double* __restrict__ sc,
                          double* __restrict__ x, double* __restrict__ s, double* __restrict__ xi) {
  double ox = linear ? 1.0 : sc[SL_QE] / sc[SL_QQ];
  if (ox >= 0.99 * C) ox = 0.99 * C;
  else if (ox <= 0.01 * C) ox = 0.01 * C;
  GRID_LOOP(i, n) {
    x[i] = ox;
    s[i] = 1.0;
    xi[i] = 1.0;
  }
}
__global__ void k_init_linear_adjust(int64_t n, const double* __restrict__ e, const double* __restrict__ a,
                                     double* __restrict__ s, double* __restrict__ xi) {
  GRID_LOOP(i, n) {
    double t = e[i] - 1.0;  // y = 0 at the start
    (void)a;
    if (t > 0) s[i] += t;
    else xi[i] -= t;
  }
}
__global__ void k_labels_to_f64(int64_t n, const float* __restrict__ af, double* __restrict__ a) {
  GRID_LOOP(i, n) a[i] = (double)af[i];
}
__global__ void k_pad_z(int64_t n, int64_t k, int kp, const float* __restrict__ Z, float* __restrict__ Zp) {
  GRID_LOOP(e, n * kp) {
    int64_t i = e / kp, c = e - i * kp;
    Zp[e] = c < k ? Z[i * k + c] : 0.0f;
  }
}

// ------------------------------------------------------------------------------------
// host driver
// ------------------------------------------------------------------------------------
namespace {
double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

tsvm_status ipm_dev(Ctx* c, const float* Zin, const float* af, int64_t n, int64_t k,
                    const tsvm_ipm_params* prm, double* x, double* xi, double* s_out,
                    tsvm_ipm_info* info) {
  const double t_start = now();
  if (n <= 0 || k <= 0 || k > TSVM_MAX_RANK)
    return fail(c, TSVM_ERR_ARG, "ipm: need n > 0, 0 < k <= %d", TSVM_MAX_RANK);
  if (!(prm->C > 0)) return fail(c, TSVM_ERR_ARG, "ipm: C must be positive");
  hipStream_t st = c->stream;
  const int kp = (int)round_up(k, 32);
  const double C = prm->C;
  // Low-precision form: the capacitance's bulk rows on BF16 MFMA, the high-weight rows
  // on FP64 (split per iteration so that the BF16 error bound stays within
  // IPM_SYRK_BUDGET; lambda_min(M) >= 1 makes that bound the refinement's contraction
  // factor up to a constant).  The plain form builds it in FP64 throughout.
  const bool fp8 = prm->precision == TSVM_PREC_FP8;
  const bool bf16 = prm->precision == TSVM_PREC_BF16 || prm->precision == TSVM_PREC_FP16X2 || fp8;
  void* p;
  auto ws = [&](const char* name, int64_t bytes, void** out) -> tsvm_status {
    TSVM_HIP(c, c->ws.get(name, (size_t)bytes, out));
    return TSVM_OK;
  };
  double *a, *e, *dx, *ds, *dxi, *adx, *ads, *adxi, *D, *winv, *rhs, *V, *solp, *solc, *T, *M,
      *sc, *red, *s, *sol3;
  int* dinfo;
#define WS(name, bytes, var)              \
  TSVM_CHECK(ws(name, (bytes), &p));      \
  var = (decltype(var))p;
  WS("ipm_a", n * 8, a);
  WS("ipm_e", n * 8, e);
  WS("ipm_dx", n * 8, dx);
  WS("ipm_ds", n * 8, ds);
  WS("ipm_dxi", n * 8, dxi);
  WS("ipm_adx", n * 8, adx);
  WS("ipm_ads", n * 8, ads);
  WS("ipm_adxi", n * 8, adxi);
  WS("ipm_D", n * 8, D);
  WS("ipm_winv", n * 8, winv);
  WS("ipm_rhs", 2 * n * 8, rhs);
  WS("ipm_V", 2 * n * 8, V);
  WS("ipm_solp", 2 * n * 8, solp);
  WS("ipm_solc", n * 8, solc);
  WS("ipm_sol3", n * 8, sol3);
  WS("ipm_T", 8 * kp * 8, T);
  WS("ipm_M", (int64_t)kp * kp * 8, M);
  WS("ipm_sc", IPM_NSLOT * 8, sc);
  WS("ipm_red", EW_BLOCKS * MAXR * 8, red);
  WS("ipm_info", 16, dinfo);
  if (s_out) {
    s = s_out;
  } else {
    WS("ipm_s", n * 8, s);
  }
  const float* Z = Zin;
  if (kp != k) {
    float* Zp;
    WS("ipm_Zp", n * kp * 4, Zp);
    hipLaunchKernelGGL(k_pad_z, dim3(4096), dim3(256), 0, st, n, k, kp, Zin, Zp);
    Z = Zp;
  }
  HybridSyrk hs;
  if (bf16) TSVM_CHECK(hybrid_syrk_setup(c, Z, n, kp, &hs, fp8 ? hybrid_fp8_eps() : 0.0));
#undef WS
  const dim3 eg(EW_BLOCKS), eb(256);
  hipLaunchKernelGGL(k_labels_to_f64, eg, eb, 0, st, n, af, a);
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, (int64_t)IPM_NSLOT, sc, 0.0);

  auto reduce = [&](std::vector<int> ops, std::vector<int> slots, int nblk = EW_BLOCKS) -> tsvm_status {
    RedSpec rs{};
    rs.n = (int)ops.size();
    for (int r = 0; r < rs.n; r++) {
      rs.op[r] = ops[r];
      rs.slot[r] = slots[r];
    }
    hipLaunchKernelGGL(reduce_slots, dim3(1), dim3(256), 0, st, (const double*)red, nblk, rs, sc);
    return TSVM_OK;
  };
  // the fused residual pass runs on the wide-row kernel; above kp = 768 (four float4 groups
  // per lane) its register demand spills, and the separate passes are kept
  const bool fuse_res = kp >= 256 && kp <= 768;
  // k-vectors: the capacitance solves (two rows), Z^T r7 and Z^T b of the predictor,
  // Z^T of the latest single solve, Z^T dx of the step being refined
  double* Tsol = T;
  double* Tr7 = T + 2 * kp;
  double* Tb = T + 3 * kp;
  double* Tc = T + 4 * kp;
  double* T3 = T + 5 * kp;   // Z^T of the corrector's step
  double* Tx = T + 6 * kp;   // Z^T x
  double* T3a = T + 7 * kp;  // Z^T of the predictor's (affine) step
  // SMW solve: sol = (D + Z Z^T)^{-1} rhs with the capacitance as factored, p right-hand
  // sides (its error is corrected by the bordered refinement below, whose residual uses
  // the exact operator).  V = D^-1 rhs must be in place; Z^T V, M^-1 Z^T V, then one
  // fused pass forms sol, its a^T sol_q (to slots) and, with Tz, Z^T sol_q.
  // have_tv: Tsol already holds Z^T V (the fused residual pass formed it)
  auto smw = [&](int pp, double* sol, double* Tz, std::vector<int> slots,
                 bool have_tv = false) -> tsvm_status {
    if (!have_tv) TSVM_CHECK(ipm_gemv_t(c, Z, n, kp, V, n, pp, Tsol));
    TSVM_CHECK(chol_solve_f64(c, M, kp, Tsol, pp, dinfo));
    int nr = 0;
    TSVM_CHECK(ipm_zpass(c, ZP_SMW, pp, Tz ? pp : 0, Z, n, kp, Tsol, V, winv, a, nullptr, nullptr,
                         nullptr, sc, 0, sol, red, Tz, nullptr, &nr));
    return reduce(std::vector<int>(slots.size(), RED_SUM), slots, nr);
  };

  // Refinement of the bordered Newton system for the step in sol3 (r5 in rhs row 0):
  //   rho = [-r5 - (A dx - a dy); r2 - a^T dx];  (ddx, ddy) from the same Schur
  //   elimination; dx += ddx, dy += ddy.  T3 = Z^T sol3 comes in from the linear
  //   combination that formed sol3, so the residual pass reads Z once.
  // refine = m > 0: m sweeps for the affine step and m for the corrector's, each solved
  // first.  refine = -m: the combined step (the one taken) starts from the affine step as
  // solved and m sweeps against the combined right-hand sides follow (m + 1 near the stop):
  // the first one's residual is the corrector's right-hand side plus the affine step's own
  // residual, so its correction is the corrector's step and the affine step's refinement
  // together (the corrector needs no solve of its own); further sweeps refine the sum.
  const bool comb = prm->refine < 0;
  const int border = comb ? -prm->refine : prm->refine;
  // With refinement, Z^T of each step is known from the passes that form it (T3 of the
  // predictor and of the corrector, the last sweep's Z^T s5 included), so Z^T x is carried
  // from iteration to iteration as Z^T x += alpha (T3a + T3) instead of read from Z again.
  const bool carry_tx = border >= 1;
  auto border_refine = [&](int corr, double* T3v, int nsw) -> tsvm_status {
    for (int sw = 0; sw < nsw; sw++) {
      const bool zt = sw + 1 < nsw || carry_tx;  // Z^T s5 wanted: T3v stays Z^T sol3
      // residual, V = r5' / D and Z^T V in one pass
      int nr = 0;
      TSVM_CHECK(ipm_zpass(c, ZP_BORDER, 1, 1, Z, n, kp, T3v, nullptr, winv, a, D, rhs, sol3, sc,
                           corr, V, red, Tsol, nullptr, &nr));
      TSVM_CHECK(reduce({RED_SUM, RED_SUM, RED_SUM}, {SL_R2B, SL_BRES, SL_BRHS}, nr));
      // s5 = S^-1 r5'; the correction is a combination of s5 and b = S^-1 a
      TSVM_CHECK(chol_solve_f64(c, M, kp, Tsol, 1, dinfo));
      TSVM_CHECK(ipm_zpass(c, ZP_SMW, 1, zt ? 1 : 0, Z, n, kp, Tsol, V, winv, a, nullptr,
                           nullptr, nullptr, sc, 0, solc, red, zt ? Tc : nullptr, nullptr, &nr));
      TSVM_CHECK(reduce({RED_SUM}, {SL_AR7B}, nr));
      hipLaunchKernelGGL(k_border_lin, eg, eb, 0, st, n, corr, (const double*)(solp + n),
                         (const double*)solc, sc, sol3, kp, (const double*)(zt ? Tc : nullptr),
                         (const double*)Tb, T3v);
    }
    return TSVM_OK;
  };

  // ---- initial point (1535-1579)
  const bool linear = prm->kernel_type == 0;
  if (!linear) {
    hipLaunchKernelGGL(k_fill, eg, eb, 0, st, n, V, 1.0);
    TSVM_CHECK(ipm_gemv_t(c, Z, n, kp, V, n, 1, T));
    TSVM_CHECK(ipm_gemv_n(c, Z, n, kp, T, 1, e, n));
    hipLaunchKernelGGL(k_init_red, eg, eb, 0, st, n, (const double*)e, red);
    TSVM_CHECK(reduce({RED_SUM, RED_SUM}, {SL_QQ, SL_QE}));
  }
  hipLaunchKernelGGL(k_init_xs, eg, eb, 0, st, n, C, linear ? 1 : 0, (const double*)sc, x, s, xi);
  if (linear || prm->init == 1) {
    TSVM_CHECK(ipm_gemv_t(c, Z, n, kp, x, n, 1, T));
    TSVM_CHECK(ipm_gemv_n(c, Z, n, kp, T, 1, e, n));
    hipLaunchKernelGGL(k_init_linear_adjust, eg, eb, 0, st, n, (const double*)e, (const double*)a, s, xi);
  }
  TSVM_HIP(c, hipGetLastError());

  Event ev0, ev1, evb0, evb1, evb2;
  TSVM_HIP(c, ev0.create());
  TSVM_HIP(c, ev1.create());
  TSVM_HIP(c, evb0.create());
  TSVM_HIP(c, evb1.create());
  TSVM_HIP(c, evb2.create());
  hipEvent_t evb[3] = {evb0, evb1, evb2};
  double t_syrk = 0.0, t_bf16 = 0.0, row_frac_sum = 0.0, fp8_frac_sum = 0.0;
  double x2_frac_sum = 0.0, flops_low = 0.0, pending_flops = 0.0;
  double t_dense = 0.0, flops_dense = 0.0, pending_dense = 0.0;
  int n_syrk = 0, n_bf16 = 0;
  bool pending_syrk = false, pending_bf16 = false;
  double hsc[IPM_NSLOT];
  tsvm_ipm_info inf;
  memset(&inf, 0, sizeof(inf));
  int reason = TSVM_MAXIT;
  int iter = 0;
  const int maxit = prm->maxit > 0 ? prm->maxit : 100;
  for (iter = 0; iter < maxit; iter++) {
    // ---- residual pass (1606-1643); with wide rows it also forms the predictor's scaled
    // right-hand sides [r5, a]/D and their Z^T (into Tsol) from the same read of Z
    if (!carry_tx || iter == 0) TSVM_CHECK(ipm_gemv_t(c, Z, n, kp, x, n, 1, Tx));
    int nres = EW_BLOCKS;
    if (fuse_res) {
      const ResArgs ra{x, s, xi, dx, ds, dxi, D, winv, rhs, V, C};
      TSVM_CHECK(ipm_zpass(c, ZP_RES, 1, 2, Z, n, kp, Tx, nullptr, nullptr, a, nullptr, nullptr,
                           nullptr, sc, 0, nullptr, red, Tsol, &ra, &nres));
    } else {
      TSVM_CHECK(ipm_gemv_n(c, Z, n, kp, Tx, 1, e, n));
      hipLaunchKernelGGL(k_resid, eg, eb, 0, st, n, C, (const double*)x, (const double*)s,
                         (const double*)xi, (const double*)a, (const double*)e, (const double*)sc,
                         dx, ds, dxi, D, winv, red);
    }
    TSVM_CHECK(reduce({RED_SUM, RED_SUM, RED_SUM, RED_SUM, RED_SUM, RED_SUM, RED_SUM, RED_MAX, RED_MIN},
                      {SL_MU, SL_NDX2, SL_AX, SL_XQX, SL_SX, SL_SXI, SL_L1, SL_DMAX, SL_DMIN},
                      nres));
    // hybrid capacitance: split the rows for this iteration's weights (the FP64 row
    // count comes back with the scalars)
    if (bf16) TSVM_CHECK(hybrid_syrk_split(c, &hs, winv, IPM_SYRK_BUDGET, sc + SL_HSB));
    TSVM_HIP(c, hipMemcpyAsync(hsc, sc, sizeof(hsc), hipMemcpyDeviceToHost, st));
    TSVM_HIP(c, hipStreamSynchronize(st));
    if (pending_syrk) {
      float ms = 0.f;
      TSVM_HIP(c, hipEventElapsedTime(&ms, ev0, ev1));
      t_syrk += ms * 1e-3;
      n_syrk++;
      if (pending_bf16) {
        TSVM_HIP(c, hipEventElapsedTime(&ms, evb[0], evb[2]));
        t_bf16 += ms * 1e-3;
        flops_low += pending_flops;
        n_bf16++;
        if (pending_dense > 0) {
          TSVM_HIP(c, hipEventElapsedTime(&ms, evb[0], evb[1]));
          t_dense += ms * 1e-3;
          flops_dense += pending_dense;
        }
      }
      pending_syrk = false;
      pending_bf16 = false;
    }
    if (hsc[SL_CHOLFAIL] != 0.0) {
      reason = TSVM_CHOL_FAIL;
      if (prm->verbose)
        printf("Cholesky fact info error; solver becomes unstable. Terminate.\n");
      break;
    }
    const double mu = hsc[SL_MU] / (2.0 * n);
    const double normdx = sqrt(hsc[SL_NDX2]);
    const double normdy = fabs(hsc[SL_AX]);
    const double tmp = 0.5 * hsc[SL_XQX];
    const double pobj = tmp - hsc[SL_SX];
    const double dobj = -tmp - C * hsc[SL_SXI];
    const double gap = (pobj - dobj) / (1.0 + fabs(pobj));
    const double cond = hsc[SL_DMAX] / hsc[SL_DMIN];
    inf.mu = mu;
    inf.normdx = normdx;
    inf.normdy = normdy;
    inf.pobj = pobj;
    inf.dobj = dobj;
    inf.gap_rel = gap;
    inf.condD = cond;
    inf.y = hsc[SL_Y];
    if (hsc[SL_BRHS] > 0) inf.refine_res = sqrt(hsc[SL_BRES] / hsc[SL_BRHS]);
    if (prm->verbose && (iter % 5 == 0 || prm->verbose > 1))
      printf("iter %d, mu=%.3e, normdx=%.3e, normdy=%.3e max/min(D)=%.3e pobj=%.9e dobj=%.9e%s\n", iter,
             mu, normdx, normdy, cond, pobj, dobj,
             prm->verbose > 1 ? (std::string(" newton-res=") + std::to_string(inf.refine_res) + " ay-res=" + std::to_string(hsc[SL_R2B])).c_str() : "");
    if (c->progress && iter % 10 == 0) {
      fprintf(stderr, "ipm: iteration %d, relative gap %.3e (%.1f s)\n", iter, gap, now() - t_start);
      fflush(stderr);
    }
    bool conv;
    const bool late = prm->stop_mode == TSVM_STOP_REF ? mu < 1e-5 : gap <= 100.0 * prm->tol;
    if (prm->stop_mode == TSVM_STOP_REF) {
      conv = mu < 1e-7 && normdx < 1e-7 && normdy < 1e-7;
    } else {
      conv = gap <= prm->tol && normdx / (1.0 + sqrt((double)n)) <= prm->tol &&
             normdy <= prm->tol * (1.0 + hsc[SL_L1] / C);
    }
    if (conv) {
      reason = TSVM_CONVERGED;
      if (prm->verbose) {
        printf("Converged!\n");
        printf("iter %d, mu=%.3e, normdx=%.3e, normdy=%.3e max/min(D)=%.3e pobj=%.9e dobj=%.9e\n",
               iter, mu, normdx, normdy, cond, pobj, dobj);
      }
      break;
    }
    if (cond > 1e18) {
      reason = TSVM_ILL_COND;
      if (prm->verbose) {
        printf("D is too ill-conditioned %.3e! Terminating.\n", cond);
        printf("iter %d, mu=%.3e, normdx=%.3e, normdy=%.3e max/min(D)=%.3e pobj=%.9e dobj=%.9e\n",
               iter, mu, normdx, normdy, cond, pobj, dobj);
      }
      break;
    }
    // ---- capacitance M = I + Z^T D^-1 Z and its Cholesky (1660-1714)
    TSVM_HIP(c, hipEventRecord(ev0, st));
    if (bf16) {
      static_assert(SL_HS2N == SL_HSB + HYBRID_NSLOT - 1, "hybrid slots are contiguous");
      const HybridCounts hc = hybrid_counts(&hs, hsc + SL_HSB);
      TSVM_CHECK(hybrid_syrk_form(c, &hs, Z, winv, hc, M, evb));
      pending_bf16 = hc.nh < n;
      row_frac_sum += 1.0 - (double)hc.nh / (double)n;
      fp8_frac_sum += (double)hc.n8 / (double)n;
      x2_frac_sum += (double)hc.n2 / (double)n;
      pending_flops = hybrid_syrk_flops(&hs, hc);
      pending_dense = hybrid_dense_mfma_flops(&hs, hc);
    } else {
      TSVM_CHECK(gemm_tn_f64(c, Z, kp, kp, Z, kp, kp, n, winv, true, M));
      hipLaunchKernelGGL(k_add_identity, dim3((unsigned)ceil_div(kp, 256)), dim3(256), 0, st, M, (int64_t)kp);
    }
    TSVM_HIP(c, hipEventRecord(ev1, st));
    pending_syrk = true;
    TSVM_CHECK(chol_f64(c, M, kp, dinfo));
    // ---- predictor (1727-1748)
    if (!fuse_res) {
      hipLaunchKernelGGL(k_rhs_pred, eg, eb, 0, st, n, C, (const double*)x, (const double*)dx,
                         (const double*)ds, (const double*)dxi, (const double*)a, rhs, n);
      hipLaunchKernelGGL(k_scale, eg, eb, 0, st, n, 2, (const double*)rhs, n, (const double*)winv, V, n);
    }
    TSVM_CHECK(smw(2, solp, Tr7, {SL_AR7, SL_AB}, fuse_res));  // r7 = S^-1 r5, b = S^-1 a
    hipLaunchKernelGGL(k_sol3_lin, eg, eb, 0, st, n, 0, (const double*)(solp + n), (const double*)solp,
                       sc, sol3, kp, (const double*)(border ? Tr7 : nullptr), (const double*)Tb, T3a);
    if (!comb) TSVM_CHECK(border_refine(0, T3a, border));
    hipLaunchKernelGGL(k_step, eg, eb, 0, st, n, C, 0, (const double*)x, (const double*)s,
                       (const double*)xi, (const double*)sol3, dx, ds, dxi, (const double*)adx,
                       (const double*)ads, (const double*)adxi, red);
    TSVM_CHECK(reduce({RED_MIN}, {SL_ALPHA_AFF}));
    hipLaunchKernelGGL(k_muaff, eg, eb, 0, st, n, C, (const double*)x, (const double*)s,
                       (const double*)xi, (const double*)dx, (const double*)ds, (const double*)dxi,
                       (const double*)sc, red);
    TSVM_CHECK(reduce({RED_SUM}, {SL_MUAFF}));
    // ---- corrector (1750-1777)
    hipLaunchKernelGGL(k_rhs_corr, eg, eb, 0, st, n, C, (const double*)x, dx, ds, dxi, adx, ads, adxi,
                       (const double*)sc, rhs, comb ? 1 : 0, (const double*)s, (const double*)xi,
                       (const double*)winv, (double*)nullptr);
    if (comb) {  // the sweeps below solve for the corrector's part
      hipLaunchKernelGGL(k_start_affine, eg, eb, 0, st, n, (const double*)adx, sol3, kp,
                         (const double*)T3a, T3, sc);
    } else {
      hipLaunchKernelGGL(k_scale, eg, eb, 0, st, n, 1, (const double*)rhs, n, (const double*)winv, V, n);
      TSVM_CHECK(smw(1, solc, border ? Tc : nullptr, {SL_AR7C}));
      hipLaunchKernelGGL(k_sol3_lin, eg, eb, 0, st, n, 1, (const double*)(solp + n),
                         (const double*)solc, sc, sol3, kp, (const double*)(border ? Tc : nullptr),
                         (const double*)Tb, T3);
    }
    // refine = -m: one sweep more once the stop is near (relative gap within 100 times the
    // tolerance, or mu below 1e-5 under the reference's rule): the first sweep leaves the
    // corrector's step with the error of one solve on the factored capacitance, which the
    // last iterations, with D at its most ill-conditioned, need refined
    const int nsw = comb ? border + (late ? 1 : 0) : border;
    TSVM_CHECK(border_refine(comb ? 2 : 1, T3, nsw));
    hipLaunchKernelGGL(k_step, eg, eb, 0, st, n, C, comb ? 0 : 1, (const double*)x, (const double*)s,
                       (const double*)xi, (const double*)sol3, dx, ds, dxi, (const double*)adx,
                       (const double*)ads, (const double*)adxi, red);
    TSVM_CHECK(reduce({RED_MIN}, {SL_ALPHA}));
    hipLaunchKernelGGL(k_update, eg, eb, 0, st, n, x, s, xi, (const double*)dx, (const double*)ds,
                       (const double*)dxi, sc, (const int*)dinfo, comb ? 1 : 0);
    if (carry_tx)
      hipLaunchKernelGGL(k_carry_tx, dim3(1), dim3(256), 0, st, kp, Tx,
                         (const double*)(comb ? nullptr : T3a), (const double*)T3, (const double*)sc,
                         (const int*)dinfo);
    TSVM_HIP(c, hipGetLastError());
  }
  TSVM_HIP(c, hipStreamSynchronize(st));
  // a failed factorisation is seen at the top of the next iteration, whose update was
  // skipped: the iterate returned is the one of the failing iteration
  inf.iters = reason == TSVM_CHOL_FAIL && iter > 0 ? iter - 1 : iter;
  inf.reason = reason;
  inf.t_total = now() - t_start;
  inf.t_syrk = t_syrk;
  if (t_syrk > 0) inf.syrk_tflops = (double)n * k * (k + 1) * n_syrk / t_syrk * 1e-12;
  inf.t_syrk_bf16 = t_bf16;
  inf.syrk_bf16_iters = n_bf16;
  if (t_bf16 > 0) inf.syrk_bf16_tflops = flops_low / t_bf16 * 1e-12;
  inf.t_syrk_dense = t_dense;
  if (t_dense > 0) inf.syrk_dense_mfma_tflops = flops_dense / t_dense * 1e-12;
  if (n_syrk > 0) inf.bf16_row_frac = row_frac_sum / n_syrk;
  if (n_syrk > 0) inf.fp8_row_frac = fp8_frac_sum / n_syrk;
  if (n_syrk > 0) inf.x2_row_frac = x2_frac_sum / n_syrk;
  if (info) *info = inf;
  return TSVM_OK;
}

}  // namespace tsvm

using tsvm::Ctx;

extern "C" {

void tsvm_ipm_default_params(tsvm_ipm_params* p) {
  if (!p) return;
  p->C = 1.0;
  p->kernel_type = 2;
  p->maxit = 100;
  p->stop_mode = TSVM_STOP_REF;
  p->tol = 1e-5;
  p->precision = TSVM_PREC_FP32;
  p->refine = -1;
  p->verbose = 0;
  p->init = 0;
}

tsvm_status tsvm_ipm_dev(tsvm_ctx* ctx, const float* Z, const float* a, int64_t n, int64_t k,
                         const tsvm_ipm_params* prm, double* alpha, double* xi, double* s,
                         tsvm_ipm_info* info) {
  if (!ctx || !Z || !a || !prm || !alpha || !xi) return TSVM_ERR_ARG;
  (void)hipSetDevice(ctx->c.device);
  return tsvm::ipm_dev(&ctx->c, Z, a, n, k, prm, alpha, xi, s, info);
}

tsvm_status tsvm_ipm(tsvm_ctx* ctx, const float* Z, const float* a, int64_t n, int64_t k,
                     const tsvm_ipm_params* prm, double* alpha, double* xi, double* s,
                     tsvm_ipm_info* info) {
  if (!ctx || !Z || !a || !prm || !alpha || !xi || n <= 0 || k <= 0) return TSVM_ERR_ARG;
  Ctx* c = &ctx->c;
  (void)hipSetDevice(c->device);
  hipStream_t st = c->stream;
  void *dZ, *da, *dal, *dxi, *ds;
  TSVM_HIP(c, c->ws.get("h_ipmZ", n * k * 4, &dZ));
  TSVM_HIP(c, c->ws.get("h_ipma", n * 4, &da));
  TSVM_HIP(c, c->ws.get("h_ipmx", n * 8, &dal));
  TSVM_HIP(c, c->ws.get("h_ipmxi", n * 8, &dxi));
  TSVM_HIP(c, c->ws.get("h_ipms", n * 8, &ds));
  TSVM_HIP(c, hipMemcpyAsync(dZ, Z, n * k * 4, hipMemcpyHostToDevice, st));
  TSVM_HIP(c, hipMemcpyAsync(da, a, n * 4, hipMemcpyHostToDevice, st));
  TSVM_CHECK(tsvm::ipm_dev(c, (float*)dZ, (float*)da, n, k, prm, (double*)dal, (double*)dxi,
                           (double*)ds, info));
  TSVM_HIP(c, hipMemcpyAsync(alpha, dal, n * 8, hipMemcpyDeviceToHost, st));
  TSVM_HIP(c, hipMemcpyAsync(xi, dxi, n * 8, hipMemcpyDeviceToHost, st));
  if (s) TSVM_HIP(c, hipMemcpyAsync(s, ds, n * 8, hipMemcpyDeviceToHost, st));
  TSVM_HIP(c, hipStreamSynchronize(st));
  return TSVM_OK;
}

tsvm_status tsvm_mpc(tsvm_ctx* ctx, const double* Z, int64_t ldz, const double* a, double C,
                     double* X, double* Xi, int64_t n, int64_t k, const tsvm_ipm_params* prm,
                     tsvm_ipm_info* info) {
  if (!ctx || !Z || !a || !X || !Xi || n <= 0 || k <= 0 || ldz < k) return TSVM_ERR_ARG;
  tsvm_ipm_params p;
  if (prm)
    p = *prm;
  else
    tsvm_ipm_default_params(&p);
  p.C = C;
  std::vector<float> Zf(n * k), af(n);
  for (int64_t i = 0; i < n; i++) {
    for (int64_t j = 0; j < k; j++) Zf[i * k + j] = (float)Z[i * ldz + j];
    af[i] = (float)a[i];
  }
  return tsvm_ipm(ctx, Zf.data(), af.data(), n, k, &p, X, Xi, nullptr, info);
}

}  // extern "C"
