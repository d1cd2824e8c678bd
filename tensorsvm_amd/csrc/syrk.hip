// Weighted Gram of the interior-point capacitance, M = I + Z^T diag(w) Z
// (replaces the Zs = D^{-1/2} Z scaling + cblas_dsyrk / chunked cublasDsyrk of
// tensorsvm-train.cu:1660-1706).
//
// Plain form: one FP64-MFMA GEMM over all n rows (gemm_tn_f64).
//
// Low-precision (hybrid) form: the rows are split by t_i = w_i |z_i|^2.  The bulk B
// (small t) goes through a BF16 MFMA SYRK with FP32 accumulation; the rows that carry
// the large weights (the free support vectors late in the solve) go through the FP64
// GEMM, read through an index list.  The split keeps the BF16 part's error below a
// budget: each bulk term w_i z_i z_i^T is computed with three bf16 roundings (z once
// per solve on both sides, w z once), a relative error of about 3 * 2^-9 that is
// zero-mean and independent across rows, so |E|_2 ~ c 2^-9 sqrt(sum_B t_i^2).  Because
// M = I + ... has lambda_min(M) >= 1, |M^-1 E| <= |E|, which is the contraction factor
// of the refinement sweeps that restore the Newton step; sqrt(sum_B t_i^2) <= budget
// keeps it well below one.  sum_B t_i^2 is bounded from integer counts of t_i per binary
// exponent (t in [2^e, 2^(e+1)) counts as 4^(e+1)), so the split is deterministic.
//
// Data: ZbT (n_pad x kpt bf16) = Z rounded once per solve (Z does not change during the
// interior-point iterations), stored in chunks of SY_BK rows, each chunk column-major:
// element (i, c) sits at ((i / SY_BK) kpt + c) SY_BK + i % SY_BK, so that one stage of a
// SYRK tile (SY_BK rows x 128 or 256 columns) is one contiguous block of memory and not
// hundreds of rows a full column length apart (a TLB miss per load); nz = |z_i|^2
// (FP64); per iteration wb (FP32 weights of the bulk rows, 0 for the FP64 rows) and rows
// (indices of the FP64 rows, in order).
#include <cstring>

#include "common.h"
#include "internal.h"

namespace tsvm {

constexpr int HS_BLOCKS = 1024;  // blocks of the split passes (each a contiguous row range)
// error budget of the FP8 tier (TSVM_PREC_FP8): the BF16 tier's 2^-9 x its budget of 16
constexpr double HYBRID_FP8_EPS = 1.0 / 32.0;
// per-row costs of the FP8 tier's choice, in units of one BF16-pass row (measured on
// MI355X at n = 1.6M / 4M, kpt = 512 / 1024): the FP8 SYRK runs at half the BF16 pass's
// time, the e4m3 packing (3 bytes per entry of HBM traffic) adds a share that falls as
// 1 / kpt against the k^2 work (0.61 at 512, 0.35 at 1024); one FP64 row through the
// index list costs about 20 BF16-pass rows
__host__ inline double hybrid_fp8_cost(int kpt) { return 0.52 + 0.35 * 1024.0 / kpt; }
constexpr double HYBRID_F64_COST = 20.0;
// one compensated row: three BF16 SYRK rows plus its gather and hi/lo packing (FP32 row
// read, two bf16 rows written: about 1.5 BF16-pass rows at kpt = 512, falling as 1 / kpt)
__host__ inline double hybrid_x2_cost(int kpt) { return 3.0 + 0.75 * 1024.0 / kpt; }
constexpr int HS_BINS = 128;     // bins of t_i by binary exponent: bin = floor(log2 t) + 64
constexpr int SY_BK = 64;        // rows (the GEMM's K) per SYRK stage
constexpr int SY_LS = SY_BK + 8;  // LDS row stride (bf16) of a staged panel

__device__ inline int hs_bin(double t) {
  if (t == 0.0) return 0;
  if (!(t > 0.0 && t < 1e300)) return HS_BINS - 1;  // negative, NaN, huge: the FP64 part
  int e = ilogb(t) + 64;
  return e < 0 ? 0 : (e > HS_BINS - 1 ? HS_BINS - 1 : e);
}

// ZbT(i, c) = bf16(Z[i][c]) for c < kp, 0 for kp <= c < kpt or i >= n (32 x 32 tiles)
__global__ void make_zbt_kernel(const float* __restrict__ Z, int64_t n, int kp, int kpt,
                                int64_t n_pad, __bf16* __restrict__ ZbT) {
  __shared__ float tile[32][33];
  const int64_t i0 = (int64_t)blockIdx.x * 32;
  const int c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    int64_t i = i0 + r;
    int c = c0 + tx;
    tile[r][tx] = (i < n && c < kp) ? Z[i * kp + c] : 0.0f;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    int c = c0 + r;
    if (c < kpt) ZbT[((i0 / SY_BK) * kpt + c) * SY_BK + i0 % SY_BK + tx] = (__bf16)tile[tx][r];
  }
}

// nz[i] = |z_i|^2 in FP64; one wave per row, lanes stride the columns
__global__ __launch_bounds__(256) void rownorm2_kernel(const float* __restrict__ Z, int64_t n,
                                                       int kp, double* __restrict__ nz) {
  const int lane = threadIdx.x & 63;
  const int64_t w0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t i = w0; i < n; i += nw) {
    double s = 0.0;
    for (int c = lane; c < kp; c += 64) {
      double v = Z[i * kp + c];
      s += v * v;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) nz[i] = s;
  }
}

// per-block counts of rows per bin of t_i = w_i nz_i (contiguous row ranges)
__global__ __launch_bounds__(256) void hs_hist_kernel(const double* __restrict__ w,
                                                      const double* __restrict__ nz, int64_t n,
                                                      int64_t rpb, int* __restrict__ part) {
  __shared__ int h[HS_BINS];
  for (int b = threadIdx.x; b < HS_BINS; b += blockDim.x) h[b] = 0;
  __syncthreads();
  const int64_t r0 = blockIdx.x * rpb, r1 = r0 + rpb < n ? r0 + rpb : n;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) atomicAdd(&h[hs_bin(w[i] * nz[i])], 1);
  __syncthreads();
  for (int b = threadIdx.x; b < HS_BINS; b += blockDim.x) part[blockIdx.x * HS_BINS + b] = h[b];
}

// Classes of a row by its bin of t_i (cuts c8 <= c1 <= c2, written by hs_cutoff_kernel
// into cut[HC_C8], cut[HC_C1], cut[HC_C2]): bin < c8 the FP8 pass, c8 <= bin < c1 the
// BF16 pass, c1 <= bin < c2 the compensated (hi + lo bf16) rows, bin >= c2 FP64.
enum { HC_C2 = 0, HC_NH = 1, HC_C8 = 2, HC_W8MAX = 3, HC_C1 = 4, HC_N2 = 5, HC_INTS = 8 };

// one block of 1024 threads: total counts per bin (32 groups of 32 threads, each thread
// summing four bins of every 32nd block's partials with 16-byte loads; integer sums, so
// the order does not matter), then the cuts.  Each tier keeps a bound on its error:
// * BF16 pass: sqrt(sum t^2) over its rows <= budget (error about 2^-9 of that);
// * compensated rows: sqrt(sum t^2) <= budget 2^7 (error about 2^-16 of that: the same
//   absolute error as the BF16 pass);
// * FP8 pass (eps8 > 0): 2^-3 sqrt(sum t^2) (three e4m3 roundings per term, 2^-4 each)
//   plus the flush bound n8 t_max 2^-17 (an entry below the subnormal step of its
//   column's scale, at most 2^-16 of the largest, becomes 0) <= eps8.
// Three layouts are priced with the per-row costs (units of one BF16-pass row: the dense
// passes stream every row, r2 per compensated row, r64 per FP64 row) and the cheapest
// taken: (A) the BF16 pass over the smallest-t rows, then compensated, then FP64;
// (B, eps8 > 0) the FP8 pass, then compensated, then FP64; (C) no dense pass: compensated
// from the smallest t up, then FP64.  x2 = false: no compensated rows.  slots: [0] bound
// of the BF16 rows, [2] error bound of the FP8 rows, [3] their number, [4] bound of the
// compensated rows ([1], [5]: the FP64 and compensated counts, from hs_scan_kernel).
constexpr int HS_CUT_GROUPS = 32;
constexpr int HS_CUT_THREADS = HS_CUT_GROUPS * HS_BINS / 4;
static_assert(HS_BINS % 4 == 0, "16-byte loads of the bin partials");
__global__ __launch_bounds__(HS_CUT_THREADS) void hs_cutoff_kernel(
    const int* __restrict__ part, int nblocks, double budget, double eps8, bool x2, double r8,
    double r2, double r64, int* __restrict__ cut_out, double* __restrict__ slots) {
  __shared__ long long grp[HS_CUT_GROUPS][HS_BINS];
  __shared__ long long cnt[HS_BINS];
  const int qd = threadIdx.x % (HS_BINS / 4), g = threadIdx.x / (HS_BINS / 4);
  long long s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll 8
  for (int k = g; k < nblocks; k += HS_CUT_GROUPS) {
    const int4 v = *(const int4*)&part[k * HS_BINS + 4 * qd];
    s0 += v.x;
    s1 += v.y;
    s2 += v.z;
    s3 += v.w;
  }
  grp[g][4 * qd] = s0;
  grp[g][4 * qd + 1] = s1;
  grp[g][4 * qd + 2] = s2;
  grp[g][4 * qd + 3] = s3;
  __syncthreads();
  if (threadIdx.x < HS_BINS) {
    long long t = 0;
    for (int q = 0; q < HS_CUT_GROUPS; q++) t += grp[q][threadIdx.x];
    cnt[threadIdx.x] = t;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  auto sq = [](int e) { return ldexp(1.0, 2 * (e - 63)); };  // bound of t^2 in bin e
  // largest cut >= e0 with sqrt(sum over bins [e0, cut) of cnt 4^(e-63)) <= bud (compared
  // squared: no square root per bin)
  auto cut_from = [&](int e0, double bud, double* cum_out) {
    double cum = 0.0;
    int cut = HS_BINS;
    const double bsq = bud * bud;
    for (int e = e0; e < HS_BINS; e++) {
      if (cnt[e] == 0) continue;
      const double add = (double)cnt[e] * sq(e);
      if (cum + add > bsq) {
        cut = e;
        break;
      }
      cum += add;
    }
    *cum_out = cum;
    return cut;
  };
  auto rows_in = [&](int e0, int e1) {
    double r = 0.0;
    for (int e = e0; e < e1; e++) r += (double)cnt[e];
    return r;
  };
  const double ntot = rows_in(0, HS_BINS);
  const double bud2 = x2 ? budget * 128.0 : 0.0;
  // (A)
  double cumA, cum2A;
  const int c1A = cut_from(0, budget, &cumA);
  const int c2A = x2 ? cut_from(c1A, bud2, &cum2A) : c1A;
  if (!x2) cum2A = 0.0;
  const double costA = ntot + r2 * rows_in(c1A, c2A) + r64 * rows_in(c2A, HS_BINS);
  int c8 = 0, c1 = c1A, c2 = c2A;
  double cum1 = cumA, cum2 = cum2A, cost = costA;
  // (C)
  if (x2) {
    double cum2C;
    const int c2C = cut_from(0, bud2, &cum2C);
    const double costC = r2 * rows_in(0, c2C) + r64 * rows_in(c2C, HS_BINS);
    if (costC < cost) {
      c8 = 0, c1 = 0, c2 = c2C, cum1 = 0.0, cum2 = cum2C, cost = costC;
    }
  }
  // (B)
  double cum8 = 0.0, n8 = 0.0;
  if (eps8 > 0.0) {
    int cut8 = 0;
    for (int e = 0; e < c1A; e++) {
      if (cnt[e] == 0) continue;  // cut8 stays one above the last non-empty bin taken
      const double add = (double)cnt[e] * sq(e);
      const double err = ldexp(sqrt(cum8 + add), -3) + (n8 + cnt[e]) * ldexp(1.0, e - 63 - 17);
      if (err > eps8) break;
      cum8 += add;
      n8 += cnt[e];
      cut8 = e + 1;
    }
    if (n8 > 0.0) {
      double cum2B = 0.0;
      const int c2B = x2 ? cut_from(cut8, bud2, &cum2B) : cut8;
      const double costB = r8 * ntot + r2 * rows_in(cut8, c2B) + r64 * rows_in(c2B, HS_BINS);
      if (costB < cost) {
        c8 = cut8, c1 = cut8, c2 = c2B, cum1 = 0.0, cum2 = cum2B, cost = costB;
      } else {
        n8 = 0.0;
      }
    }
  }
  if (c8 == 0) {
    n8 = 0.0;
    cum8 = 0.0;
  }
  cut_out[HC_C2] = c2;
  cut_out[HC_C8] = c8;
  cut_out[HC_C1] = c1;
  slots[0] = sqrt(cum1);
  slots[2] = n8 > 0.0 ? ldexp(sqrt(cum8), -3) + n8 * ldexp(1.0, c8 - 1 - 63 - 17) : 0.0;
  slots[3] = n8;
  slots[4] = sqrt(cum2);
}

// per-block counts of the FP64 rows (bin >= c2) and of the compensated rows
// (c1 <= bin < c2)
__global__ __launch_bounds__(256) void hs_count_kernel(const double* __restrict__ w,
                                                       const double* __restrict__ nz, int64_t n,
                                                       int64_t rpb, const int* __restrict__ cut,
                                                       int* __restrict__ cnt,
                                                       int* __restrict__ cnt2) {
  __shared__ int tot[2];
  if (threadIdx.x < 2) tot[threadIdx.x] = 0;
  __syncthreads();
  const int c1 = cut[HC_C1], c2 = cut[HC_C2];
  const int64_t r0 = blockIdx.x * rpb, r1 = r0 + rpb < n ? r0 + rpb : n;
  int ch = 0, cm = 0;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
    const int bin = hs_bin(w[i] * nz[i]);
    ch += bin >= c2;
    cm += bin >= c1 && bin < c2;
  }
  atomicAdd(&tot[0], ch);
  atomicAdd(&tot[1], cm);
  __syncthreads();
  if (threadIdx.x == 0) {
    cnt[blockIdx.x] = tot[0];
    cnt2[blockIdx.x] = tot[1];
  }
}

// one block of HS_BLOCKS threads: exclusive scan of the block counts (Hillis-Steele in
// LDS); *slot = *total = the sum
__global__ __launch_bounds__(HS_BLOCKS) void hs_scan_kernel(const int* __restrict__ cnt,
                                                            int nblocks, int* __restrict__ off,
                                                            double* __restrict__ slot,
                                                            int* __restrict__ total) {
  __shared__ int v[2][HS_BLOCKS];
  const int t = threadIdx.x;
  v[0][t] = t < nblocks ? cnt[t] : 0;
  __syncthreads();
  int cur = 0;
  for (int o = 1; o < HS_BLOCKS; o <<= 1) {
    v[cur ^ 1][t] = v[cur][t] + (t >= o ? v[cur][t - o] : 0);
    cur ^= 1;
    __syncthreads();
  }
  // v[cur] holds inclusive sums
  if (t < nblocks) off[t] = v[cur][t] - (t < nblocks ? cnt[t] : 0);
  if (t == HS_BLOCKS - 1) {
    *slot = (double)v[cur][t];
    *total = v[cur][t];
  }
}

// rows of the FP64 part and of the compensated part, each in increasing order; wb = FP32
// weight of the BF16-pass rows, 0 elsewhere and on the padding; w8 (when given) =
// sqrt(w) (FP32) on the FP8-pass rows, 0 elsewhere, and w8max (bits of a non-negative
// float) their maximum
__global__ __launch_bounds__(256) void hs_scatter_kernel(const double* __restrict__ w,
                                                         const double* __restrict__ nz, int64_t n,
                                                         int64_t n_pad, int64_t rpb,
                                                         const int* __restrict__ cut,
                                                         const int* __restrict__ off,
                                                         const int* __restrict__ off2,
                                                         int* __restrict__ rows,
                                                         int* __restrict__ rows2,
                                                         float* __restrict__ wb,
                                                         float* __restrict__ w8,
                                                         unsigned int* __restrict__ w8max) {
  __shared__ int wsum[2][4];
  const int c2 = cut[HC_C2], c1 = cut[HC_C1];
  const int c8 = w8 ? cut[HC_C8] : 0;
  float mx = 0.0f;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t r0 = blockIdx.x * rpb;
  int64_t r1 = r0 + rpb < n_pad ? r0 + rpb : n_pad;
  int base = off[blockIdx.x], base2 = off2[blockIdx.x];
  for (int64_t c0 = r0; c0 < r1; c0 += 256) {
    const int64_t i = c0 + threadIdx.x;
    bool hi = false, mid = false;
    float wf = 0.0f, w8f = 0.0f;
    if (i < r1 && i < n) {
      const double wi = w[i];
      const int bin = hs_bin(wi * nz[i]);
      hi = bin >= c2;
      mid = bin >= c1 && bin < c2;
      if (bin < c8) {  // wi >= 0 here: negative and NaN weights are in the last bin
        w8f = (float)sqrt(wi);
        mx = fmaxf(mx, w8f);
      } else if (bin < c1) {
        wf = (float)wi;
      }
    }
    if (i < r1) {
      wb[i] = wf;
      if (w8) w8[i] = w8f;
    }
    const unsigned long long m = __ballot(hi), m2 = __ballot(mid);
    const unsigned long long below_mask = (1ull << lane) - 1ull;
    if (lane == 0) {
      wsum[0][wave] = __popcll(m);
      wsum[1][wave] = __popcll(m2);
    }
    __syncthreads();
    int pre = 0, pre2 = 0;
    for (int k = 0; k < wave; k++) {
      pre += wsum[0][k];
      pre2 += wsum[1][k];
    }
    if (hi) rows[base + pre + __popcll(m & below_mask)] = (int)i;
    if (mid) rows2[base2 + pre2 + __popcll(m2 & below_mask)] = (int)i;
    base += wsum[0][0] + wsum[0][1] + wsum[0][2] + wsum[0][3];
    base2 += wsum[1][0] + wsum[1][1] + wsum[1][2] + wsum[1][3];
    __syncthreads();
  }
  if (w8) {
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if (lane == 0) atomicMax(w8max, __float_as_uint(mx));
  }
}

// per-column power-of-two scales of the FP8 tier: |y_ic| = sqrt(w_i) |z_ic| is below
// min(sqrt(t_max), w8max zmax_c) (t_max = 2^(cut8 - 64) bounds t_i = w_i |z_i|^2 in the
// tier); sc8[c] = 2^e with that bound times sc8[c] in [128, 256) (e4m3 holds 448)
__global__ void fp8_scales_kernel(const float* __restrict__ zmax, int kpt,
                                  const int* __restrict__ cut, const unsigned int* __restrict__ w8max,
                                  float* __restrict__ sc8) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= kpt) return;
  const double tb = sqrt(ldexp(1.0, cut[2] - 64));
  const double cb = (double)__uint_as_float(*w8max) * (double)zmax[col];
  const double bnd = cb < tb ? cb : tb;
  int e = 0;
  if (bnd > 0.0) {
    int ex;
    (void)frexp(bnd, &ex);  // bnd = f 2^ex, f in [0.5, 1)
    e = 8 - ex;
    e = e > 100 ? 100 : (e < -100 ? -100 : e);
  }
  sc8[col] = ldexpf(1.0f, e);
}

// zmax[c] = max_i |z_ic| (the FP8 tier's static column bounds)
__global__ void colmax_kernel(const float* __restrict__ Z, int64_t n, int kp, int kpt,
                              unsigned int* __restrict__ zmax_bits) {
  const int col = blockIdx.y * 64 + (threadIdx.x & 63);
  if (col >= kpt) return;
  float mx = 0.0f;
  if (col < kp)
    for (int64_t i = blockIdx.x * 4 + (threadIdx.x >> 6); i < n; i += (int64_t)gridDim.x * 4)
      mx = fmaxf(mx, fabsf(Z[i * kp + col]));
  atomicMax(zmax_bits + col, __float_as_uint(mx));
}

// scales 8 bf16 by 8 FP32 factors, one rounding each
__device__ inline bf16x8_t scale8(u32x4_t v, const f32x4_t* w) {
  bf16x8_t a8 = __builtin_bit_cast(bf16x8_t, v);
#pragma unroll
  for (int j = 0; j < 8; j++) a8[j] = (__bf16)((float)a8[j] * w[j >> 2][j & 3]);
  return a8;
}

// BF16 SYRK for kpt = 128: part[split] (128 x 128) = sum over the split's rows of
// wb_i z_i z_i^T.  Block = 4 waves (wave 64 x 64 = 2 x 2 MFMA 32x32x16 tiles); SY_BK
// rows per stage, double-buffered LDS with the next stage's global loads in flight; the
// weights scale the A panel while it is staged.  The tiles of one split
// sit 8 blocks apart (on one XCD) so that its rows are read once into L2.
__global__ __launch_bounds__(256) void syrk_bf16_kernel(const __bf16* __restrict__ ZbT,
                                                        int64_t n_pad, int kpt,
                                                        const float* __restrict__ wb, int ntile,
                                                        int nsplit, int64_t stages_per_split,
                                                        float* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) __bf16 sA[2][128 * SY_LS];
  __shared__ __attribute__((aligned(16))) __bf16 sB[2][128 * SY_LS];
  const int nlow = ntile * (ntile + 1) / 2;
  const int bid = blockIdx.x;
  const int xcd = bid % 8, grp = bid / 8;
  const int tile = grp % nlow;
  const int split = (grp / nlow) * 8 + xcd;
  if (split >= nsplit) return;
  int ta = 0;
  while ((ta + 1) * (ta + 2) / 2 <= tile) ta++;
  const int tb = tile - ta * (ta + 1) / 2;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l32 = lane & 31, h = lane >> 5;
  const int wa = (wave >> 1) * 64, wbo = (wave & 1) * 64;
  const int64_t s0 = (int64_t)split * stages_per_split;
  const int64_t nst_all = n_pad / SY_BK;
  const int64_t s1 = s0 + stages_per_split < nst_all ? s0 + stages_per_split : nst_all;
  const int64_t nst = s1 > s0 ? s1 - s0 : 0;
  f32x16_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int q = 0; q < 16; q++) acc[i][j][q] = 0.0f;
  const __bf16* Apan = ZbT + (int64_t)ta * 128 * SY_BK;
  const __bf16* Bpan = ZbT + (int64_t)tb * 128 * SY_BK;
  // staging map: thread -> (panel row = idx >> 3, 8 consecutive K at (idx & 7) * 8) for
  // idx = tid + 256 v; the K offset depends on tid only, so 8 weights cover all four
  const int scol = (tid & 7) * 8;
  u32x4_t va[4], vb[4];
  f32x4_t wv[2];
  auto gload = [&](int64_t st) {
    const int64_t n0 = (s0 + st) * SY_BK;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int row = (tid + v * 256) >> 3;
      va[v] = *(const u32x4_t*)(Apan + ((s0 + st) * kpt + row) * SY_BK + scol);
      vb[v] = *(const u32x4_t*)(Bpan + ((s0 + st) * kpt + row) * SY_BK + scol);
    }
    wv[0] = *(const f32x4_t*)(wb + n0 + scol);
    wv[1] = *(const f32x4_t*)(wb + n0 + scol + 4);
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int row = (tid + v * 256) >> 3;
      *(bf16x8_t*)&sA[buf][row * SY_LS + scol] = scale8(va[v], wv);
      *(u32x4_t*)&sB[buf][row * SY_LS + scol] = vb[v];
    }
  };
  if (nst > 0) {
    gload(0);
    sstore(0);
    __syncthreads();
  }
  for (int64_t st = 0; st < nst; st++) {
    const int buf = (int)(st & 1);
    gload(st + 1 < nst ? st + 1 : st);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int ks = 0; ks < SY_BK / 16; ks++) {
      bf16x8_t af[2], bfr[2];
#pragma unroll
      for (int i = 0; i < 2; i++) {
        af[i] = *(const bf16x8_t*)&sA[buf][(wa + 32 * i + l32) * SY_LS + ks * 16 + 8 * h];
        bfr[i] = *(const bf16x8_t*)&sB[buf][(wbo + 32 * i + l32) * SY_LS + ks * 16 + 8 * h];
      }
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("" ::: "memory");
    sstore(buf ^ 1);
    __syncthreads();
  }
  // C layout of the 32x32 MFMA: register q holds row 8(q/4) + 4h + q%4, column l32
  float* P = part + (int64_t)split * kpt * kpt;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int a = ta * 128 + wa + 32 * i + 8 * (q / 4) + 4 * h + (q % 4);
        const int bcol = tb * 128 + wbo + 32 * j + l32;
        P[(int64_t)a * kpt + bcol] = acc[i][j][q];
      }
}

// BF16 SYRK for kpt a multiple of 256: part[split] (kpt x kpt, lower 256 x 256 tiles) =
// sum over the split's rows of wb_i z_i z_i^T.  One block per CU; 4 waves, each
// a 128 x 128 quarter of the tile (4 x 4 MFMA 32x32x16 tiles, 256 accumulators), so that
// each 16-byte fragment read from LDS feeds 4 MFMAs.  Software pipeline per stage of SB_K
// rows (4 steps of 16): the fragments of step s + 1 are read while the MFMAs of step s
// run (one wave per SIMD leaves nothing else to hide LDS latency); the next stage is
// staged into the other LDS buffer (the A panel scaled by the weights) during the last
// two steps, half a stage per step, and each half's registers are reloaded at once with
// the stage after it (a whole stage for the loads to arrive); one barrier per stage.
// The result of the wave above the diagonal of a diagonal tile is not stored.
constexpr int SB_K = SY_BK;      // the chunk of the ZbT layout
constexpr int SB_LS = SB_K + 8;  // LDS row stride (bf16): conflict-free 16-byte reads

// APAN(st) / BPAN(st): the A / B panel (256 columns) of the operand chunk of stage st;
// SCALE: the A panel is scaled by the row weights wb (stage st covers rows
// (s0 + st) SB_K ..), else both panels are staged as loaded
template <bool SCALE, class APAN, class BPAN>
__device__ __forceinline__ void syrk256_stages(APAN apan, BPAN bpan,
                                               const float* __restrict__ wb, int64_t s0,
                                               int64_t nst, __bf16 (*sm)[256 * SB_LS],
                                               f32x16_t (&acc)[4][4], int tid, int wa, int wbo,
                                               int l32, int h) {
  // Both panels are loaded on diagonal tiles too (the same rows twice, the second from
  // L2): the three tiles of a split then stream at one pace and share each row in their
  // XCD's L2.  With one panel the diagonal tiles ran ahead and the pass fetched 3.0 GB
  // from HBM at 1.6M x 512 instead of 1.65 GB (in the same time: the pass is not
  // HBM-bound, but the traffic is waste).
  constexpr int NP = 2;  // panels loaded
  const int scol = (tid & 7) * 8;   // staging map: panel rows (tid >> 3) + 32 q, 8 K at scol
  const int srow = tid >> 3;
  u32x4_t v[NP][8];
  f32x4_t wv[2];
  // half hf of stage st's loads: panel rows (tid >> 3) + 32 q for q in [4 hf, 4 hf + 4);
  // the weights come with the second half (both halves' stores scale by them)
  auto gload = [&](int64_t st, int hf) {
    const int64_t base = (int64_t)srow * SB_K + scol;
#pragma unroll
    for (int p = 0; p < NP; p++) {
      const __bf16* pan = p ? bpan(st) : apan(st);
#pragma unroll
      for (int q = 4 * hf; q < 4 * hf + 4; q++)
        v[p][q] = *(const u32x4_t*)(pan + base + (int64_t)32 * q * SB_K);
    }
    if (SCALE && hf == 1) {
      wv[0] = *(const f32x4_t*)(wb + (s0 + st) * SB_K + scol);
      wv[1] = *(const f32x4_t*)(wb + (s0 + st) * SB_K + scol + 4);
    }
  };
  // half hf of the staging: panel rows (tid >> 3) + 32 q for q in [4 hf, 4 hf + 4);
  // A scaled (SCALE), B as loaded
  auto sstore = [&](int buf, int hf) {
#pragma unroll
    for (int q = 4 * hf; q < 4 * hf + 4; q++) {
      if (SCALE)
        *(bf16x8_t*)&sm[2 * buf][(srow + 32 * q) * SB_LS + scol] = scale8(v[0][q], wv);
      else
        *(u32x4_t*)&sm[2 * buf][(srow + 32 * q) * SB_LS + scol] = v[0][q];
      *(u32x4_t*)&sm[2 * buf + 1][(srow + 32 * q) * SB_LS + scol] = v[NP - 1][q];
    }
  };
  auto fread = [&](bf16x8_t(&fa)[4], bf16x8_t(&fb)[4], int buf, int ks) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      fa[i] = *(const bf16x8_t*)&sm[2 * buf][(wa + 32 * i + l32) * SB_LS + ks * 16 + 8 * h];
      fb[i] = *(const bf16x8_t*)&sm[2 * buf + 1][(wbo + 32 * i + l32) * SB_LS + ks * 16 + 8 * h];
    }
  };
  auto mm = [&](const bf16x8_t(&fa)[4], const bf16x8_t(&fb)[4]) {
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
  };
  if (nst <= 0) return;
  gload(0, 0);
  gload(0, 1);
  sstore(0, 0);
  sstore(0, 1);
  gload(nst > 1 ? 1 : 0, 0);
  gload(nst > 1 ? 1 : 0, 1);
  __syncthreads();
  bf16x8_t fa0[4], fb0[4], fa1[4], fb1[4];
  // Phases are fenced with sched_barrier.  Inside one: the next step's 8 fragment reads
  // first, then the 16 MFMAs; in the last two, the staging VALU work is spread between
  // the MFMAs (5 per MFMA), its LDS writes come next and then the loads of the same half of
  // stage st + 2 into the registers just written out, so that each load has a whole stage
  // to arrive (the last stages reload the last one: no branch in the body).
  for (int64_t st = 0; st < nst; st++) {
    const int buf = (int)(st & 1);
    const int64_t snext = st + 2 < nst ? st + 2 : nst - 1;
    fread(fa0, fb0, buf, 0);
    __builtin_amdgcn_sched_barrier(0);
    fread(fa1, fb1, buf, 1);
    mm(fa0, fb0);
    __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 16, 0);
    __builtin_amdgcn_sched_barrier(0);
    fread(fa0, fb0, buf, 2);
    mm(fa1, fb1);
    __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 16, 0);
    __builtin_amdgcn_sched_barrier(0);
    fread(fa1, fb1, buf, 3);
    sstore(buf ^ 1, 0);  // buf ^ 1 was last read before the previous barrier
    gload(snext, 0);
    mm(fa0, fb0);
    __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
#pragma unroll
    for (int m = 0; m < 16; m++) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x002, 5, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x200, 8, 0);
    __builtin_amdgcn_sched_group_barrier(0x020, 4 * NP, 0);
    __builtin_amdgcn_sched_barrier(0);
    sstore(buf ^ 1, 1);
    gload(snext, 1);
    mm(fa1, fb1);
#pragma unroll
    for (int m = 0; m < 16; m++) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x002, 5, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x200, 8, 0);
    __builtin_amdgcn_sched_group_barrier(0x020, 4 * NP + (SCALE ? 2 : 0), 0);
    __builtin_amdgcn_sched_barrier(0);
    __syncthreads();
  }
}

__global__ __launch_bounds__(256, 1) void syrk_bf16_256_kernel(const __bf16* __restrict__ ZbT,
                                                               int64_t n_pad, int kpt,
                                                               const float* __restrict__ wb,
                                                               int ntile, int nsplit,
                                                               int64_t stages_per_split,
                                                               float* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) __bf16 sm[4][256 * SB_LS];  // [buf * 2 + panel]
  const int nlow = ntile * (ntile + 1) / 2;
  const int bid = blockIdx.x;
  const int xcd = bid % 8, grp = bid / 8;
  const int tile = grp % nlow;
  const int split = (grp / nlow) * 8 + xcd;
  if (split >= nsplit) return;
  int ta = 0;
  while ((ta + 1) * (ta + 2) / 2 <= tile) ta++;
  const int tb = tile - ta * (ta + 1) / 2;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l32 = lane & 31, h = lane >> 5;
  const int wa = (wave >> 1) * 128, wbo = (wave & 1) * 128;
  const int64_t s0 = (int64_t)split * stages_per_split;
  const int64_t nst_all = n_pad / SB_K;
  const int64_t s1 = s0 + stages_per_split < nst_all ? s0 + stages_per_split : nst_all;
  const int64_t nst = s1 > s0 ? s1 - s0 : 0;
  f32x16_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int q = 0; q < 16; q++) acc[i][j][q] = 0.0f;
  const __bf16* Apan = ZbT + (int64_t)ta * 256 * SB_K;
  const __bf16* Bpan = ZbT + (int64_t)tb * 256 * SB_K;
  const int64_t cstride = (int64_t)kpt * SB_K;  // elements per chunk
  auto apan = [&](int64_t st) { return Apan + (s0 + st) * cstride; };
  auto bpan = [&](int64_t st) { return Bpan + (s0 + st) * cstride; };
  syrk256_stages<true>(apan, bpan, wb, s0, nst, sm, acc, tid, wa, wbo, l32, h);
  if (ta == tb && wbo > wa) return;  // strictly above the diagonal
  // C layout of the 32x32 MFMA: register q holds row 8(q/4) + 4h + q%4, column l32
  float* P = part + (int64_t)split * kpt * kpt;
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int a = ta * 256 + wa + 32 * i + 8 * (q / 4) + 4 * h + (q % 4);
        const int bcol = tb * 256 + wbo + 32 * j + l32;
        P[(int64_t)a * kpt + bcol] = acc[i][j][q];
      }
}

// M (kp x kp FP64) += sum over splits of the BF16 partials, in split order.  Threads
// read lower entries (a >= b, coalesced along b) and add each sum to both mirrored
// entries, so M stays exactly symmetric.
// sc (FP8 tier, may be null): the per-column scales of the operands; the sum is divided
// by sc_a sc_b
__global__ void syrk_accum_kernel(const float* __restrict__ part, int nsplit, int kpt, int kp,
                                  double* __restrict__ M, const float* __restrict__ sc = nullptr) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= (int64_t)kp * kp) return;
  const int a = (int)(idx / kp), b = (int)(idx % kp);
  if (b > a) return;
  double s = 0.0;
  for (int t = 0; t < nsplit; t++) s += (double)part[(int64_t)t * kpt * kpt + (int64_t)a * kpt + b];
  if (sc) s /= (double)sc[a] * (double)sc[b];
  M[idx] += s;
  if (a != b) M[(int64_t)b * kp + a] += s;
}

// FP8 pass (kpt a multiple of 256): Z8 (chunked like ZbT, one byte per entry) holds
// y_ic = e4m3(sc8_c sqrt(w_i) z_ic), w8 = sqrt(w_i) on the FP8 rows and 0 elsewhere, so
// sum_i y_i y_i^T / (sc8_a sc8_b) is the FP8 rows' part of Z^T diag(w) Z.  sc8_c is a power
// of two: the products are exact, one e4m3 rounding (round to nearest even) per entry.
// Block: one 64-row chunk x 256 columns; a thread converts 8 rows of a column.
__device__ inline unsigned int pack_fp8x4(float a, float b, float c, float d) {
  int v = 0;
  v = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, v, false);
  v = __builtin_amdgcn_cvt_pk_fp8_f32(c, d, v, true);
  return (unsigned int)v;
}
typedef __attribute__((ext_vector_type(8))) int i32x8_t;
typedef __attribute__((ext_vector_type(2))) unsigned int u32x2_t;

__global__ __launch_bounds__(256) void pack_fp8_kernel(const __bf16* __restrict__ ZbT, int kpt,
                                                       const float* __restrict__ w8,
                                                       const float* __restrict__ sc8,
                                                       unsigned char* __restrict__ Z8) {
  const int64_t ch = blockIdx.x;
  const int rg = threadIdx.x & 7, cl = threadIdx.x >> 3;
  const f32x4_t w0 = *(const f32x4_t*)(w8 + ch * SY_BK + rg * 8);
  const f32x4_t w1 = *(const f32x4_t*)(w8 + ch * SY_BK + rg * 8 + 4);
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int col = blockIdx.y * 256 + cl + 32 * q;
    const int64_t off = (ch * kpt + col) * SY_BK + rg * 8;
    const bf16x8_t zb = __builtin_bit_cast(bf16x8_t, *(const u32x4_t*)(ZbT + off));
    const float sc = sc8[col];
    float y[8];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      y[t] = (float)zb[t] * w0[t] * sc;
      y[t + 4] = (float)zb[t + 4] * w1[t] * sc;
    }
    u32x2_t pk;
    pk[0] = pack_fp8x4(y[0], y[1], y[2], y[3]);
    pk[1] = pack_fp8x4(y[4], y[5], y[6], y[7]);
    *(u32x2_t*)(Z8 + off) = pk;
  }
}

// FP8 SYRK (kpt a multiple of 256): part[split] (lower 256 x 256 tiles) = sum over the
// split's rows of y_i y_i^T from Z8.  Same tiling as syrk_bf16_256_kernel (4 waves, each a
// 128 x 128 quarter, 256 accumulators); a 64-row chunk is one K step of
// v_mfma_scale_f32_32x32x64_f8f6f4 with unit block scales (the per-column scales are
// divided out when the partials are summed).  LDS rows of 80 bytes per column (64 K bytes
// + padding: 16 lanes' 16-byte reads at a stride of 20 banks cover all 64 banks); a lane's
// fragment is the 32 K bytes 32 (lane / 32) .. + 32 of its column.  Double-buffered LDS
// with the next chunk's loads in registers, one barrier per chunk; on a diagonal tile one
// panel is staged.
constexpr int S8_LS = 80;  // LDS bytes per staged column

__global__ __launch_bounds__(256, 1) void syrk_fp8_256_kernel(const unsigned char* __restrict__ Z8,
                                                              int64_t n_pad, int kpt, int ntile,
                                                              int nsplit, int64_t stages_per_split,
                                                              float* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) unsigned char sm[2][2][256 * S8_LS];
  const int nlow = ntile * (ntile + 1) / 2;
  const int bid = blockIdx.x;
  const int xcd = bid % 8, grp = bid / 8;
  const int tile = grp % nlow;
  const int split = (grp / nlow) * 8 + xcd;
  if (split >= nsplit) return;
  int ta = 0;
  while ((ta + 1) * (ta + 2) / 2 <= tile) ta++;
  const int tb = tile - ta * (ta + 1) / 2;
  const bool diag = ta == tb;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l32 = lane & 31, h = lane >> 5;
  const int wa = (wave >> 1) * 128, wbo = (wave & 1) * 128;
  const int64_t s0 = (int64_t)split * stages_per_split;
  const int64_t nst_all = n_pad / SY_BK;
  const int64_t s1 = s0 + stages_per_split < nst_all ? s0 + stages_per_split : nst_all;
  const int64_t nst = s1 > s0 ? s1 - s0 : 0;
  f32x16_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int q = 0; q < 16; q++) acc[i][j][q] = 0.0f;
  // staging map: 16-byte piece idx = tid + 256 v of a panel is column idx / 4, K bytes
  // 16 (idx % 4) .. + 16
  u32x4_t v[2][4];
  auto gload = [&](int64_t st) {
#pragma unroll
    for (int p = 0; p < 2; p++) {
      const unsigned char* pan = Z8 + ((s0 + st) * kpt + (p && !diag ? tb : ta) * 256) * SY_BK;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int idx = tid + 256 * q;
        if (p == 0 || !diag) v[p][q] = *(const u32x4_t*)(pan + (idx >> 2) * SY_BK + (idx & 3) * 16);
      }
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int idx = tid + 256 * q;
        if (p == 0 || !diag) *(u32x4_t*)&sm[buf][p][(idx >> 2) * S8_LS + (idx & 3) * 16] = v[p][q];
      }
  };
  if (nst > 0) {
    gload(0);
    sstore(0);
    __syncthreads();
  }
  const int pb_sel = diag ? 0 : 1;
  for (int64_t st = 0; st < nst; st++) {
    const int buf = (int)(st & 1);
    gload(st + 1 < nst ? st + 1 : st);
    i32x8_t fa[4], fb[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const unsigned char* pa = &sm[buf][0][(wa + 32 * i + l32) * S8_LS + 32 * h];
      const unsigned char* pb = &sm[buf][pb_sel][(wbo + 32 * i + l32) * S8_LS + 32 * h];
      const u32x4_t a0 = *(const u32x4_t*)pa, a1 = *(const u32x4_t*)(pa + 16);
      const u32x4_t b0 = *(const u32x4_t*)pb, b1 = *(const u32x4_t*)(pb + 16);
      fa[i] = (i32x8_t){(int)a0[0], (int)a0[1], (int)a0[2], (int)a0[3],
                        (int)a1[0], (int)a1[1], (int)a1[2], (int)a1[3]};
      fb[i] = (i32x8_t){(int)b0[0], (int)b0[1], (int)b0[2], (int)b0[3],
                        (int)b1[0], (int)b1[1], (int)b1[2], (int)b1[3]};
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++)
        acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fa[i], fb[j], acc[i][j], 0, 0,
                                                                   0, 127, 0, 127);
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("" ::: "memory");
    sstore(buf ^ 1);
    __syncthreads();
  }
  if (diag && wbo > wa) return;  // strictly above the diagonal
  // C layout of the 32x32 MFMA: register q holds row 8(q/4) + 4h + q%4, column l32
  float* P = part + (int64_t)split * kpt * kpt;
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int a = ta * 256 + wa + 32 * i + 8 * (q / 4) + 4 * h + (q % 4);
        const int bcol = tb * 256 + wbo + 32 * j + l32;
        P[(int64_t)a * kpt + bcol] = acc[i][j][q];
      }
}

// Compensated rows: y_i = sqrt(w_i) z_i in FP64, hi = bf16(y), lo = bf16(y - hi), so that
// y_a y_b = hi_a hi_b + hi_a lo_b + lo_a hi_b up to about 3 2^-18 |y_a| |y_b|.  H and L
// hold hi and lo of the listed rows rows2[0..m2) (compacted, zero padded to m2p rows),
// chunked like ZbT.  Block: 64 compacted rows x 64 columns; rows read as float4 (16 lanes
// per 256-byte row segment), hi / lo transposed through LDS, chunk columns written as
// 16-byte pieces (8 rows each).
constexpr int PX_LS = 72;  // LDS row stride (bf16) of the transposed tiles
__global__ __launch_bounds__(256) void pack_x2_kernel(const float* __restrict__ Z, int kp,
                                                      int kpt, const double* __restrict__ w,
                                                      const int* __restrict__ rows2, int64_t m2,
                                                      __bf16* __restrict__ H,
                                                      __bf16* __restrict__ L) {
  __shared__ __attribute__((aligned(16))) __bf16 th[64 * PX_LS];
  __shared__ __attribute__((aligned(16))) __bf16 tl[64 * PX_LS];
  __shared__ double sw[64];
  __shared__ int src[64];
  const int64_t q = blockIdx.x;
  const int c0 = blockIdx.y * 64;
  const int t = threadIdx.x;
  if (t < 64) {
    const int64_t r = q * 64 + t;
    const int i = r < m2 ? rows2[r] : -1;
    src[t] = i;
    sw[t] = i >= 0 ? sqrt(w[i]) : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < 4; p++) {
    const int idx = p * 256 + t, r = idx >> 4, c4 = (idx & 15) * 4;
    const int i = src[r];
    float z[4] = {0.f, 0.f, 0.f, 0.f};
    if (i >= 0) {
      const float* zr = Z + (int64_t)i * kp + c0 + c4;
      if (c0 + c4 + 4 <= kp && (kp & 3) == 0) {
        const f32x4_t v = *(const f32x4_t*)zr;
        z[0] = v[0], z[1] = v[1], z[2] = v[2], z[3] = v[3];
      } else {
        for (int u = 0; u < 4; u++)
          if (c0 + c4 + u < kp) z[u] = zr[u];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const double y = sw[r] * (double)z[u];
      const __bf16 hi = (__bf16)(float)y;
      th[(c4 + u) * PX_LS + r] = hi;
      tl[(c4 + u) * PX_LS + r] = (__bf16)(float)(y - (double)(float)hi);
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < 2; p++) {
    const int idx = p * 256 + t, cc = idx >> 3, r8 = (idx & 7) * 8;
    const int64_t o = (q * kpt + c0 + cc) * SY_BK + r8;
    *(u32x4_t*)(H + o) = *(const u32x4_t*)&th[cc * PX_LS + r8];
    *(u32x4_t*)(L + o) = *(const u32x4_t*)&tl[cc * PX_LS + r8];
  }
}

// Compensated SYRK (kpt a multiple of 256): part[split] (lower 256 x 256 tiles) = the
// split's share of sum over the compacted rows of hi hi^T + hi lo^T + lo hi^T, as one
// product over the K-concatenation [hi | hi | lo] [hi | lo | hi]^T: stage g of 3 S
// (S = m2p / 64 chunks) takes its A panel from chunk g % S of H (g < 2 S) or L, its B
// panel from H (g < S or g >= 2 S) or L.  The sum is symmetric in (a, b), so only lower
// tiles are formed; the tiling is syrk_bf16_256_kernel's, both panels always staged.
__global__ __launch_bounds__(256, 1) void syrk_x2_256_kernel(const __bf16* __restrict__ H,
                                                             const __bf16* __restrict__ L,
                                                             int64_t S, int kpt, int ntile,
                                                             int nsplit, int64_t stages_per_split,
                                                             float* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) __bf16 sm[4][256 * SB_LS];  // [buf * 2 + panel]
  const int nlow = ntile * (ntile + 1) / 2;
  const int bid = blockIdx.x;
  const int xcd = bid % 8, grp = bid / 8;
  const int tile = grp % nlow;
  const int split = (grp / nlow) * 8 + xcd;
  if (split >= nsplit) return;
  int ta = 0;
  while ((ta + 1) * (ta + 2) / 2 <= tile) ta++;
  const int tb = tile - ta * (ta + 1) / 2;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l32 = lane & 31, h = lane >> 5;
  const int wa = (wave >> 1) * 128, wbo = (wave & 1) * 128;
  const int64_t s0 = (int64_t)split * stages_per_split;
  const int64_t nst_all = 3 * S;
  const int64_t s1 = s0 + stages_per_split < nst_all ? s0 + stages_per_split : nst_all;
  const int64_t nst = s1 > s0 ? s1 - s0 : 0;
  f32x16_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int q = 0; q < 16; q++) acc[i][j][q] = 0.0f;
  const int64_t cstride = (int64_t)kpt * SB_K;
  const int64_t aoff = (int64_t)ta * 256 * SB_K, boff = (int64_t)tb * 256 * SB_K;
  auto apan = [&](int64_t st) {
    const int64_t g = s0 + st, c = g % S;
    return (g < 2 * S ? H : L) + c * cstride + aoff;
  };
  auto bpan = [&](int64_t st) {
    const int64_t g = s0 + st, c = g % S;
    return (g >= S && g < 2 * S ? L : H) + c * cstride + boff;
  };
  syrk256_stages<false>(apan, bpan, nullptr, s0, nst, sm, acc, tid, wa, wbo, l32, h);
  if (ta == tb && wbo > wa) return;  // strictly above the diagonal
  // C layout of the 32x32 MFMA: register q holds row 8(q/4) + 4h + q%4, column l32
  float* P = part + (int64_t)split * kpt * kpt;
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int a = ta * 256 + wa + 32 * i + 8 * (q / 4) + 4 * h + (q % 4);
        const int bcol = tb * 256 + wbo + 32 * j + l32;
        P[(int64_t)a * kpt + bcol] = acc[i][j][q];
      }
}

__global__ void add_identity_kernel(double* __restrict__ M, int64_t k) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < k) M[i * k + i] += 1.0;
}

// ---- host side ----------------------------------------------------------------------

__global__ void zero_u32s_kernel(unsigned int* __restrict__ p, int cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) p[i] = 0u;
}

tsvm_status hybrid_syrk_setup(Ctx* c, const float* Z, int64_t n, int kp, HybridSyrk* hs,
                              double fp8_eps) {
  hs->n = n;
  hs->kp = kp;
  hs->kpt = kp <= 128 ? 128 : (int)round_up(kp, 256);
  hs->fp8_eps = hs->kpt >= 256 ? fp8_eps : 0.0;
  hs->x2 = hs->kpt >= 256;
  hs->n_pad = round_up(n > 0 ? n : 1, SY_BK);
  hs->r8 = hybrid_fp8_cost(hs->kpt);
  hs->r2 = hybrid_x2_cost(hs->kpt);
  hs->r64 = HYBRID_F64_COST;
  void* p;
  TSVM_HIP(c, c->ws.get("hs_ZbT", (size_t)hs->kpt * hs->n_pad * 2, &p));
  hs->ZbT = p;
  TSVM_HIP(c, c->ws.get("hs_nz", (size_t)(n > 0 ? n : 1) * 8, &p));
  hs->nz = (double*)p;
  TSVM_HIP(c, c->ws.get("hs_wb", (size_t)hs->n_pad * 4, &p));
  hs->wb = (float*)p;
  TSVM_HIP(c, c->ws.get("hs_rows", (size_t)(n > 0 ? n : 1) * 4, &p));
  hs->rows = (int*)p;
  TSVM_HIP(c, c->ws.get("hs_rows2", (size_t)(n > 0 ? n : 1) * 4, &p));
  hs->rows2 = (int*)p;
  TSVM_HIP(c, c->ws.get("hs_ints", (size_t)(HS_BLOCKS * (HS_BINS + 4) + HC_INTS) * 4, &p));
  hs->ints = (int*)p;
  hipStream_t s = c->stream;
  if (hs->fp8_eps > 0.0) {
    TSVM_HIP(c, c->ws.get("hs_w8", (size_t)hs->n_pad * 4, &p));
    hs->w8 = (float*)p;
    TSVM_HIP(c, c->ws.get("hs_zmax", (size_t)hs->kpt * 4, &p));
    hs->zmax = (float*)p;
    TSVM_HIP(c, c->ws.get("hs_sc8", (size_t)hs->kpt * 4, &p));
    hs->sc8 = (float*)p;
    TSVM_HIP(c, c->ws.get("hs_Z8", (size_t)hs->kpt * hs->n_pad, &p));
    hs->Z8 = p;
    hipLaunchKernelGGL(zero_u32s_kernel, dim3((unsigned)ceil_div(hs->kpt, 256)), dim3(256), 0, s,
                       (unsigned int*)hs->zmax, hs->kpt);
    hipLaunchKernelGGL(colmax_kernel, dim3(1024, (unsigned)(hs->kpt / 64)), dim3(256), 0, s, Z, n,
                       kp, hs->kpt, (unsigned int*)hs->zmax);
  }
  dim3 g((unsigned)(hs->n_pad / 32), (unsigned)(hs->kpt / 32));
  hipLaunchKernelGGL(make_zbt_kernel, g, dim3(256), 0, s, Z, n, kp, hs->kpt, hs->n_pad,
                     (__bf16*)hs->ZbT);
  hipLaunchKernelGGL(rownorm2_kernel, dim3(4096), dim3(256), 0, s, Z, n, kp, hs->nz);
  TSVM_HIP(c, hipGetLastError());
  return TSVM_OK;
}

tsvm_status hybrid_syrk_split(Ctx* c, HybridSyrk* hs, const double* w, double budget,
                              double* slots) {
  hipStream_t s = c->stream;
  const int64_t n = hs->n;
  int* part = hs->ints;                       // HS_BLOCKS x HS_BINS
  int* cnt = part + HS_BLOCKS * HS_BINS;      // HS_BLOCKS
  int* off = cnt + HS_BLOCKS;                 // HS_BLOCKS
  int* cnt2 = off + HS_BLOCKS;                // HS_BLOCKS
  int* off2 = cnt2 + HS_BLOCKS;               // HS_BLOCKS
  int* cut = off2 + HS_BLOCKS;                // HC_INTS
  const int64_t rpb = round_up(ceil_div(hs->n_pad, HS_BLOCKS), 256);
  const int nb = (int)ceil_div(hs->n_pad, rpb);
  hipLaunchKernelGGL(hs_hist_kernel, dim3(nb), dim3(256), 0, s, w, hs->nz, n, rpb, part);
  hipLaunchKernelGGL(hs_cutoff_kernel, dim3(1), dim3(HS_CUT_THREADS), 0, s,
                     (const int*)part, nb, budget, hs->fp8_eps, hs->x2, hs->r8, hs->r2, hs->r64,
                     cut, slots);
  hipLaunchKernelGGL(hs_count_kernel, dim3(nb), dim3(256), 0, s, w, hs->nz, n, rpb,
                     (const int*)cut, cnt, cnt2);
  hipLaunchKernelGGL(hs_scan_kernel, dim3(1), dim3(HS_BLOCKS), 0, s, (const int*)cnt, nb, off,
                     slots + 1, cut + HC_NH);
  hipLaunchKernelGGL(hs_scan_kernel, dim3(1), dim3(HS_BLOCKS), 0, s, (const int*)cnt2, nb, off2,
                     slots + 5, cut + HC_N2);
  unsigned int* w8max = (unsigned int*)(cut + HC_W8MAX);
  if (hs->w8) hipLaunchKernelGGL(zero_u32s_kernel, dim3(1), dim3(64), 0, s, w8max, 1);
  hipLaunchKernelGGL(hs_scatter_kernel, dim3(nb), dim3(256), 0, s, w, hs->nz, n, hs->n_pad, rpb,
                     (const int*)cut, (const int*)off, (const int*)off2, hs->rows, hs->rows2,
                     hs->wb, hs->w8, w8max);
  if (hs->w8)
    hipLaunchKernelGGL(fp8_scales_kernel, dim3((unsigned)ceil_div(hs->kpt, 256)), dim3(256), 0, s,
                       (const float*)hs->zmax, hs->kpt, (const int*)cut,
                       (const unsigned int*)w8max, hs->sc8);
  TSVM_HIP(c, hipGetLastError());
  return TSVM_OK;
}

HybridCounts hybrid_counts(const HybridSyrk* hs, const double* slots) {
  HybridCounts hc;
  hc.nh = (int64_t)slots[1];
  hc.n8 = hs->fp8_eps > 0.0 ? (int64_t)slots[3] : 0;
  hc.n2 = hs->x2 ? (int64_t)slots[5] : 0;
  hc.nb = hs->n - hc.nh - hc.n8 - hc.n2;
  return hc;
}

// splits of a dense or compensated pass over nst stages: as many per XCD as fit its CUs
// with every tile of a split resident at once (the 128 kernel runs two blocks per CU, the
// 256 kernels one)
static void pass_splits(const Ctx* c, int kpt, int64_t nst, int* ntile, int* nlow, int* nsplit,
                        int64_t* sps) {
  const bool big = kpt >= 256;
  *ntile = kpt / (big ? 256 : 128);
  *nlow = *ntile * (*ntile + 1) / 2;
  const int slots_xcd = (c->cu_count / 8) * (big ? 1 : 2);
  int ns = 8 * (slots_xcd / *nlow > 1 ? slots_xcd / *nlow : 1);
  if (ns > round_up(nst, 8)) ns = (int)round_up(nst, 8);
  *nsplit = ns;
  *sps = ceil_div(nst, ns);
}

tsvm_status hybrid_syrk_form(Ctx* c, HybridSyrk* hs, const float* Z, const double* w,
                             const HybridCounts& hc, double* M, hipEvent_t* ev_low) {
  hipStream_t s = c->stream;
  const int64_t n = hs->n;
  const int kp = hs->kp, kpt = hs->kpt;
  // FP64 rows (all rows when every row is FP64: no index list); this GEMM also writes
  // M's first value (zero for no rows), which the other parts add to
  if (hc.nh >= n)
    TSVM_CHECK(gemm_tn_f64(c, Z, kp, kp, Z, kp, kp, n, w, true, M));
  else
    TSVM_CHECK(gemm_tn_f64(c, Z, kp, kp, Z, kp, kp, hc.nh, w, true, M, hs->rows));
  hipLaunchKernelGGL(add_identity_kernel, dim3((unsigned)ceil_div(kp, 256)), dim3(256), 0, s, M,
                     (int64_t)kp);
  if (ev_low) TSVM_HIP(c, hipEventRecord(ev_low[0], s));
  const dim3 accg((unsigned)ceil_div((int64_t)kp * kp, 256));
  int ntile, nlow, nsplit;
  int64_t sps;
  if (hc.n8 > 0 || hc.nb > 0) {
    // the dense pass over every row: FP8 (FP8 rows, w8 = 0 elsewhere) or BF16 (wb = 0
    // outside the BF16 rows)
    const bool fp8 = hc.n8 > 0 && hs->Z8 && kpt >= 256;
    const int64_t nst = hs->n_pad / SY_BK;
    pass_splits(c, kpt, nst, &ntile, &nlow, &nsplit, &sps);
    void* part;
    TSVM_HIP(c, c->ws.get("hs_part", (size_t)nsplit * kpt * kpt * 4, &part));
    if (fp8) {
      hipLaunchKernelGGL(pack_fp8_kernel, dim3((unsigned)nst, (unsigned)(kpt / 256)), dim3(256), 0, s,
                         (const __bf16*)hs->ZbT, kpt, (const float*)hs->w8,
                         (const float*)hs->sc8, (unsigned char*)hs->Z8);
      hipLaunchKernelGGL(syrk_fp8_256_kernel, dim3((unsigned)(nlow * nsplit)), dim3(256), 0, s,
                         (const unsigned char*)hs->Z8, hs->n_pad, kpt, ntile, nsplit, sps,
                         (float*)part);
      hipLaunchKernelGGL(syrk_accum_kernel, accg, dim3(256), 0, s, (const float*)part, nsplit,
                         kpt, kp, M, (const float*)hs->sc8);
    } else {
      if (kpt >= 256)
        hipLaunchKernelGGL(syrk_bf16_256_kernel, dim3((unsigned)(nlow * nsplit)), dim3(256), 0, s,
                           (const __bf16*)hs->ZbT, hs->n_pad, kpt, (const float*)hs->wb, ntile,
                           nsplit, sps, (float*)part);
      else
        hipLaunchKernelGGL(syrk_bf16_kernel, dim3((unsigned)(nlow * nsplit)), dim3(256), 0, s,
                           (const __bf16*)hs->ZbT, hs->n_pad, kpt, (const float*)hs->wb, ntile,
                           nsplit, sps, (float*)part);
      hipLaunchKernelGGL(syrk_accum_kernel, accg, dim3(256), 0, s, (const float*)part, nsplit,
                         kpt, kp, M, (const float*)nullptr);
    }
  }
  if (ev_low) TSVM_HIP(c, hipEventRecord(ev_low[1], s));
  if (hc.n2 > 0 && hs->x2) {
    // the compensated rows, compacted: capacity grown in steps of an eighth of the rows
    const int64_t m2p = round_up(hc.n2, SY_BK);
    const int64_t cap = round_up(m2p, round_up(ceil_div(hs->n_pad, 8), SY_BK));
    void *ph, *pl;
    TSVM_HIP(c, c->ws.get("hs_x2H", (size_t)cap * kpt * 2, &ph));
    TSVM_HIP(c, c->ws.get("hs_x2L", (size_t)cap * kpt * 2, &pl));
    const int64_t S = m2p / SY_BK;
    hipLaunchKernelGGL(pack_x2_kernel, dim3((unsigned)S, (unsigned)(kpt / 64)), dim3(256), 0, s, Z,
                       kp, kpt, w, (const int*)hs->rows2, hc.n2, (__bf16*)ph, (__bf16*)pl);
    pass_splits(c, kpt, 3 * S, &ntile, &nlow, &nsplit, &sps);
    void* part;
    TSVM_HIP(c, c->ws.get("hs_part2", (size_t)nsplit * kpt * kpt * 4, &part));
    hipLaunchKernelGGL(syrk_x2_256_kernel, dim3((unsigned)(nlow * nsplit)), dim3(256), 0, s,
                       (const __bf16*)ph, (const __bf16*)pl, S, kpt, ntile, nsplit, sps,
                       (float*)part);
    hipLaunchKernelGGL(syrk_accum_kernel, accg, dim3(256), 0, s, (const float*)part, nsplit, kpt,
                       kp, M, (const float*)nullptr);
  }
  if (ev_low) TSVM_HIP(c, hipEventRecord(ev_low[2], s));
  TSVM_HIP(c, hipGetLastError());
  return TSVM_OK;
}

double hybrid_fp8_eps() { return HYBRID_FP8_EPS; }

double hybrid_syrk_flops(const HybridSyrk* hs, const HybridCounts& hc) {
  // rows formed in low precision (each once) times the SYRK count k (k + 1)
  return (double)(hs->n - hc.nh) * hs->kp * (hs->kp + 1.0);
}

double hybrid_dense_mfma_flops(const HybridSyrk* hs, const HybridCounts& hc) {
  // the dense pass streams every row through the lower tiles (256 or 128 wide, diagonal
  // tiles whole): the MFMA work it issues
  if (hc.n8 <= 0 && hc.nb <= 0) return 0.0;
  const int bt = hs->kpt >= 256 ? 256 : 128, nt = hs->kpt / bt;
  return 2.0 * bt * bt * (nt * (nt + 1) / 2) * (double)hs->n_pad;
}

tsvm_status weighted_syrk_dev(Ctx* c, const float* Z, int64_t n, int64_t k, const double* w,
                              int precision, double budget, double* M, tsvm_syrk_info* info) {
  if (n < 0 || k <= 0 || k > 4096 || !Z || !w || !M)
    return fail(c, TSVM_ERR_ARG, "weighted_syrk: bad arguments (n=%ld k=%ld)", (long)n, (long)k);
  if (precision != TSVM_PREC_FP32 && precision != TSVM_PREC_BF16 && precision != TSVM_PREC_FP8)
    return fail(c, TSVM_ERR_ARG, "weighted_syrk: unknown precision %d", precision);
  hipStream_t s = c->stream;
  tsvm_syrk_info inf;
  memset(&inf, 0, sizeof(inf));
  inf.n_fp64_rows = n;
  if (precision == TSVM_PREC_FP32 || n == 0) {
    TSVM_CHECK(gemm_tn_f64(c, Z, k, k, Z, k, k, n, w, true, M));
    hipLaunchKernelGGL(add_identity_kernel, dim3((unsigned)ceil_div(k, 256)), dim3(256), 0, s, M, k);
    TSVM_HIP(c, hipGetLastError());
  } else {
    HybridSyrk hs;
    TSVM_CHECK(hybrid_syrk_setup(c, Z, n, (int)k, &hs,
                                 precision == TSVM_PREC_FP8 ? HYBRID_FP8_EPS : 0.0));
    void* p;
    TSVM_HIP(c, c->ws.get("hs_slots", HYBRID_NSLOT * 8, &p));
    double* slots = (double*)p;
    TSVM_CHECK(hybrid_syrk_split(c, &hs, w, budget > 0 ? budget : 4.0, slots));
    double hsl[HYBRID_NSLOT] = {0};
    TSVM_HIP(c, hipMemcpyAsync(hsl, slots, sizeof(hsl), hipMemcpyDeviceToHost, s));
    TSVM_HIP(c, hipStreamSynchronize(s));
    Event ev0, ev1, ev2;
    TSVM_HIP(c, ev0.create());
    TSVM_HIP(c, ev1.create());
    TSVM_HIP(c, ev2.create());
    hipEvent_t ev[3] = {ev0, ev1, ev2};
    const HybridCounts hc = hybrid_counts(&hs, hsl);
    TSVM_CHECK(hybrid_syrk_form(c, &hs, Z, w, hc, M, ev));
    TSVM_HIP(c, hipStreamSynchronize(s));
    float ms = 0.f, msd = 0.f;
    TSVM_HIP(c, hipEventElapsedTime(&ms, ev[0], ev[2]));
    TSVM_HIP(c, hipEventElapsedTime(&msd, ev[0], ev[1]));
    inf.n_fp64_rows = hc.nh;
    inf.bf16_bound = hsl[0];
    inf.n_fp8_rows = hc.n8;
    inf.fp8_bound = hsl[2];
    inf.n_x2_rows = hc.n2;
    inf.x2_bound = hsl[4];
    if (hc.nh < n) {
      inf.t_bf16 = ms * 1e-3;
      if (ms > 0) inf.bf16_tflops = hybrid_syrk_flops(&hs, hc) / (ms * 1e-3) * 1e-12;
    }
    const double fd = hybrid_dense_mfma_flops(&hs, hc);
    if (fd > 0) {
      inf.t_dense = msd * 1e-3;
      if (msd > 0) inf.dense_mfma_tflops = fd / (msd * 1e-3) * 1e-12;
    }
  }
  TSVM_HIP(c, hipStreamSynchronize(s));
  if (info) *info = inf;
  return TSVM_OK;
}

}  // namespace tsvm
