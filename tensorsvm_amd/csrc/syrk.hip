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

// one block of 8 HS_BINS threads: total counts per bin (8 groups of partials, summed in
// a fixed order), then the cutoff bin: the bulk is every bin below `cut`, the largest
// prefix whose bound sqrt(sum cnt_b 4^(b-63)) stays within the budget.  out: cut[0] =
// cut, cut[2] = cut8 (0: no FP8 pass), slots[0] = bound of the bulk, slots[2] = bound of
// the FP8 rows, slots[3] = their number.
// FP8 tier (eps8 > 0): the smallest-t bins while 2^-3 sqrt(sum t^2) (three e4m3 roundings
// per term, 2^-4 each) plus the flush bound n8 t_max 2^-17 (an entry below the subnormal
// step of its column's scale, at most 2^-16 of the largest, becomes 0) stays within eps8.
// The iteration takes the FP8 pass (over bins < cut8, the rest in FP64) when the per-row
// cost model says it is cheaper than the BF16 pass: r8 n + r64 nb < n, nb the bulk rows
// that would move to FP64 (costs in units of one BF16-pass row).  Otherwise the split is
// exactly the BF16 form's.
constexpr int HS_CUT_GROUPS = 8;
__global__ __launch_bounds__(HS_CUT_GROUPS * HS_BINS) void hs_cutoff_kernel(
    const int* __restrict__ part, int nblocks, double budget, double eps8, double r8, double r64,
    int* __restrict__ cut_out, double* __restrict__ slots) {
  __shared__ long long grp[HS_CUT_GROUPS][HS_BINS];
  __shared__ long long cnt[HS_BINS];
  const int b = threadIdx.x % HS_BINS, g = threadIdx.x / HS_BINS;
  long long s = 0;
#pragma unroll 8
  for (int k = g; k < nblocks; k += HS_CUT_GROUPS) s += part[k * HS_BINS + b];
  grp[g][b] = s;
  __syncthreads();
  if (g == 0) {
    long long t = 0;
    for (int q = 0; q < HS_CUT_GROUPS; q++) t += grp[q][b];
    cnt[b] = t;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double cum = 0.0;
  int cut = HS_BINS;
  for (int e = 0; e < HS_BINS; e++) {
    if (cnt[e] == 0) continue;
    const double add = (double)cnt[e] * ldexp(1.0, 2 * (e - 63));
    if (sqrt(cum + add) > budget) {
      cut = e;
      break;
    }
    cum += add;
  }
  int cut8 = 0;
  double cum8 = 0.0, n8 = 0.0;
  if (eps8 > 0.0) {
    for (int e = 0; e < cut; e++) {
      if (cnt[e] == 0) continue;  // cut8 stays one above the last non-empty bin taken
      const double add = (double)cnt[e] * ldexp(1.0, 2 * (e - 63));
      const double err = ldexp(sqrt(cum8 + add), -3) + (n8 + cnt[e]) * ldexp(1.0, e - 63 - 17);
      if (err > eps8) break;
      cum8 += add;
      n8 += cnt[e];
      cut8 = e + 1;
    }
    double ntot = 0.0, nb = 0.0;
    for (int e = 0; e < HS_BINS; e++) ntot += (double)cnt[e];
    for (int e = cut8; e < cut; e++) nb += (double)cnt[e];
    if (!(n8 > 0.0 && r8 * ntot + r64 * nb < ntot)) {
      cut8 = 0;
      cum8 = 0.0;
      n8 = 0.0;
    }
  }
  if (n8 > 0.0) {  // FP8 pass: no BF16 rows, FP64 from cut8 up
    cut = cut8;
    cum = 0.0;
  }
  cut_out[0] = cut;
  cut_out[2] = cut8;
  slots[0] = sqrt(cum);
  slots[2] = n8 > 0.0 ? ldexp(sqrt(cum8), -3) + n8 * ldexp(1.0, cut8 - 1 - 63 - 17) : 0.0;
  slots[3] = n8;
}

// per-block counts of the FP64 rows (bin >= cut)
__global__ __launch_bounds__(256) void hs_count_kernel(const double* __restrict__ w,
                                                       const double* __restrict__ nz, int64_t n,
                                                       int64_t rpb, const int* __restrict__ cut,
                                                       int* __restrict__ cnt) {
  __shared__ int tot;
  if (threadIdx.x == 0) tot = 0;
  __syncthreads();
  const int cb = *cut;
  const int64_t r0 = blockIdx.x * rpb, r1 = r0 + rpb < n ? r0 + rpb : n;
  int c = 0;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) c += hs_bin(w[i] * nz[i]) >= cb;
  atomicAdd(&tot, c);
  __syncthreads();
  if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
}

// one block of HS_BLOCKS threads: exclusive scan of the block counts (Hillis-Steele in
// LDS); slots[1] = number of FP64 rows
__global__ __launch_bounds__(HS_BLOCKS) void hs_scan_kernel(const int* __restrict__ cnt,
                                                            int nblocks, int* __restrict__ off,
                                                            double* __restrict__ slots,
                                                            int* __restrict__ nh) {
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
    slots[1] = (double)v[cur][t];
    *nh = v[cur][t];
  }
}

// rows of the FP64 part in increasing order; wb = FP32 weight of the bulk rows, 0 for
// the FP64 rows and the padding
// FP8 rows (bin < cut[2], when w8 is given): wb = 0, w8 = sqrt(w) (FP32) and
// w8max (bits of a non-negative float) = their maximum
__global__ __launch_bounds__(256) void hs_scatter_kernel(const double* __restrict__ w,
                                                         const double* __restrict__ nz, int64_t n,
                                                         int64_t n_pad, int64_t rpb,
                                                         const int* __restrict__ cut,
                                                         const int* __restrict__ off,
                                                         int* __restrict__ rows,
                                                         float* __restrict__ wb,
                                                         float* __restrict__ w8,
                                                         unsigned int* __restrict__ w8max) {
  __shared__ int wsum[4];
  const int cb = cut[0];
  const int c8 = w8 ? cut[2] : 0;
  float mx = 0.0f;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t r0 = blockIdx.x * rpb;
  int64_t r1 = r0 + rpb < n_pad ? r0 + rpb : n_pad;
  int base = off[blockIdx.x];
  for (int64_t c0 = r0; c0 < r1; c0 += 256) {
    const int64_t i = c0 + threadIdx.x;
    bool hi = false;
    float wf = 0.0f, w8f = 0.0f;
    if (i < r1 && i < n) {
      const double wi = w[i];
      const int bin = hs_bin(wi * nz[i]);
      hi = bin >= cb;
      if (bin < c8) {  // wi >= 0 here: negative and NaN weights are in the last bin
        w8f = (float)sqrt(wi);
        mx = fmaxf(mx, w8f);
      } else if (!hi) {
        wf = (float)wi;
      }
    }
    if (i < r1) {
      wb[i] = wf;
      if (w8) w8[i] = w8f;
    }
    const unsigned long long m = __ballot(hi);
    const int below = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wave] = __popcll(m);
    __syncthreads();
    int pre = 0;
    for (int k = 0; k < wave; k++) pre += wsum[k];
    const int tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    if (hi) rows[base + pre + below] = (int)i;
    base += tot;
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
// run (one wave per SIMD leaves nothing else to hide LDS latency), the next stage's
// global loads are issued at the top of the stage and staged into the other LDS buffer
// (the A panel scaled by the weights) during the last two steps, one barrier per stage.
// On a diagonal tile one panel is loaded (staged scaled and as is) and the result of the
// wave above the diagonal is not stored.
constexpr int SB_K = SY_BK;      // the chunk of the ZbT layout
constexpr int SB_LS = SB_K + 8;  // LDS row stride (bf16): conflict-free 16-byte reads

template <bool DIAG>
__device__ __forceinline__ void syrk256_stages(const __bf16* __restrict__ Apan,
                                               const __bf16* __restrict__ Bpan,
                                               const float* __restrict__ wb, int kpt, int64_t s0,
                                               int64_t nst, __bf16 (*sm)[256 * SB_LS],
                                               f32x16_t (&acc)[4][4], int tid, int wa, int wbo,
                                               int l32, int h) {
  constexpr int NP = DIAG ? 1 : 2;  // panels loaded
  const int scol = (tid & 7) * 8;   // staging map: panel rows (tid >> 3) + 32 q, 8 K at scol
  const int srow = tid >> 3;
  u32x4_t v[NP][8];
  f32x4_t wv[2];
  auto gload = [&](int64_t st) {
    const int64_t base = ((s0 + st) * kpt + srow) * SB_K + scol;
#pragma unroll
    for (int p = 0; p < NP; p++)
#pragma unroll
      for (int q = 0; q < 8; q++)
        v[p][q] = *(const u32x4_t*)((p ? Bpan : Apan) + base + (int64_t)32 * q * SB_K);
    wv[0] = *(const f32x4_t*)(wb + (s0 + st) * SB_K + scol);
    wv[1] = *(const f32x4_t*)(wb + (s0 + st) * SB_K + scol + 4);
  };
  // half hf of the staging: panel rows (tid >> 3) + 32 q for q in [4 hf, 4 hf + 4);
  // A scaled, B as loaded
  auto sstore = [&](int buf, int hf) {
#pragma unroll
    for (int q = 4 * hf; q < 4 * hf + 4; q++) {
      *(bf16x8_t*)&sm[2 * buf][(srow + 32 * q) * SB_LS + scol] = scale8(v[0][q], wv);
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
  gload(0);
  sstore(0, 0);
  sstore(0, 1);
  __syncthreads();
  bf16x8_t fa0[4], fb0[4], fa1[4], fb1[4];
  // Phases are fenced with sched_barrier.  Inside one: the next step's 8 fragment reads
  // first, then the 16 MFMAs; in the last two, the staging VALU work is spread between
  // the MFMAs (5 per MFMA) and its LDS writes come last.
  for (int64_t st = 0; st < nst; st++) {
    const int buf = (int)(st & 1);
    fread(fa0, fb0, buf, 0);
    // the last stage reloads itself into the unused buffer: no branch in the body
    gload(st + 1 < nst ? st + 1 : st);
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
    mm(fa0, fb0);
    __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
#pragma unroll
    for (int m = 0; m < 16; m++) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x002, 5, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x200, 8, 0);
    __builtin_amdgcn_sched_barrier(0);
    sstore(buf ^ 1, 1);
    mm(fa1, fb1);
#pragma unroll
    for (int m = 0; m < 16; m++) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x002, 5, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x200, 8, 0);
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
  if (ta == tb)
    syrk256_stages<true>(Apan, Bpan, wb, kpt, s0, nst, sm, acc, tid, wa, wbo, l32, h);
  else
    syrk256_stages<false>(Apan, Bpan, wb, kpt, s0, nst, sm, acc, tid, wa, wbo, l32, h);
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
  hs->n_pad = round_up(n > 0 ? n : 1, SY_BK);
  void* p;
  TSVM_HIP(c, c->ws.get("hs_ZbT", (size_t)hs->kpt * hs->n_pad * 2, &p));
  hs->ZbT = p;
  TSVM_HIP(c, c->ws.get("hs_nz", (size_t)(n > 0 ? n : 1) * 8, &p));
  hs->nz = (double*)p;
  TSVM_HIP(c, c->ws.get("hs_wb", (size_t)hs->n_pad * 4, &p));
  hs->wb = (float*)p;
  TSVM_HIP(c, c->ws.get("hs_rows", (size_t)(n > 0 ? n : 1) * 4, &p));
  hs->rows = (int*)p;
  TSVM_HIP(c, c->ws.get("hs_ints", (size_t)(HS_BLOCKS * (HS_BINS + 2) + 8) * 4, &p));
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
    hs->r8 = hybrid_fp8_cost(hs->kpt);
    hs->r64 = HYBRID_F64_COST;
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
  int* cut = off + HS_BLOCKS;                 // 1
  int* nh = cut + 1;                          // 1
  const int64_t rpb = round_up(ceil_div(hs->n_pad, HS_BLOCKS), 256);
  const int nb = (int)ceil_div(hs->n_pad, rpb);
  hipLaunchKernelGGL(hs_hist_kernel, dim3(nb), dim3(256), 0, s, w, hs->nz, n, rpb, part);
  hipLaunchKernelGGL(hs_cutoff_kernel, dim3(1), dim3(HS_CUT_GROUPS * HS_BINS), 0, s,
                     (const int*)part, nb, budget, hs->fp8_eps, hs->r8, hs->r64, cut, slots);
  hipLaunchKernelGGL(hs_count_kernel, dim3(nb), dim3(256), 0, s, w, hs->nz, n, rpb,
                     (const int*)cut, cnt);
  hipLaunchKernelGGL(hs_scan_kernel, dim3(1), dim3(HS_BLOCKS), 0, s, (const int*)cnt, nb, off, slots, nh);
  unsigned int* w8max = (unsigned int*)(cut + 3);  // cut[0] cut, cut[1] FP64 rows, cut[2] cut8
  if (hs->w8) hipLaunchKernelGGL(zero_u32s_kernel, dim3(1), dim3(64), 0, s, w8max, 1);
  hipLaunchKernelGGL(hs_scatter_kernel, dim3(nb), dim3(256), 0, s, w, hs->nz, n, hs->n_pad, rpb,
                     (const int*)cut, (const int*)off, hs->rows, hs->wb, hs->w8, w8max);
  if (hs->w8)
    hipLaunchKernelGGL(fp8_scales_kernel, dim3((unsigned)ceil_div(hs->kpt, 256)), dim3(256), 0, s,
                       (const float*)hs->zmax, hs->kpt, (const int*)cut,
                       (const unsigned int*)w8max, hs->sc8);
  TSVM_HIP(c, hipGetLastError());
  return TSVM_OK;
}

tsvm_status hybrid_syrk_form(Ctx* c, HybridSyrk* hs, const float* Z, const double* w,
                             int64_t nh, double* M, hipEvent_t* ev_bf16, int64_t n8) {
  hipStream_t s = c->stream;
  const int64_t n = hs->n;
  const int kp = hs->kp, kpt = hs->kpt;
  // FP64 rows (all rows when every row is FP64: no index list)
  if (nh >= n)
    TSVM_CHECK(gemm_tn_f64(c, Z, kp, kp, Z, kp, kp, n, w, true, M));
  else
    TSVM_CHECK(gemm_tn_f64(c, Z, kp, kp, Z, kp, kp, nh, w, true, M, hs->rows));
  hipLaunchKernelGGL(add_identity_kernel, dim3((unsigned)ceil_div(kp, 256)), dim3(256), 0, s, M,
                     (int64_t)kp);
  if (ev_bf16) TSVM_HIP(c, hipEventRecord(ev_bf16[0], s));
  const bool fp8 = n8 > 0 && hs->Z8 && kpt >= 256;
  if (nh < n) {
    // splits: as many per XCD as fit its CUs with every tile of a split resident at once
    // (the 128 kernel runs two blocks per CU, the 256 kernels one)
    const bool big = kpt >= 256;
    const int ntile = kpt / (big ? 256 : 128);
    const int nlow = ntile * (ntile + 1) / 2;
    const int64_t nst = hs->n_pad / SY_BK;
    const int slots_xcd = (c->cu_count / 8) * (big ? 1 : 2);
    int nsplit = 8 * (slots_xcd / nlow > 1 ? slots_xcd / nlow : 1);
    if (nsplit > round_up(nst, 8)) nsplit = (int)round_up(nst, 8);
    const int64_t sps = ceil_div(nst, nsplit);
    void* part;
    TSVM_HIP(c, c->ws.get("hs_part", (size_t)nsplit * kpt * kpt * 4, &part));
    if (fp8) {
      // the FP8 pass over the FP8 rows (the others have w8 = 0) in place of the BF16 pass
      hipLaunchKernelGGL(pack_fp8_kernel, dim3((unsigned)nst, (unsigned)(kpt / 256)), dim3(256), 0, s,
                         (const __bf16*)hs->ZbT, kpt, (const float*)hs->w8,
                         (const float*)hs->sc8, (unsigned char*)hs->Z8);
      hipLaunchKernelGGL(syrk_fp8_256_kernel, dim3((unsigned)(nlow * nsplit)), dim3(256), 0, s,
                         (const unsigned char*)hs->Z8, hs->n_pad, kpt, ntile, nsplit, sps,
                         (float*)part);
      hipLaunchKernelGGL(syrk_accum_kernel, dim3((unsigned)ceil_div((int64_t)kp * kp, 256)),
                         dim3(256), 0, s, (const float*)part, nsplit, kpt, kp, M,
                         (const float*)hs->sc8);
    } else {
      if (big)
        hipLaunchKernelGGL(syrk_bf16_256_kernel, dim3((unsigned)(nlow * nsplit)), dim3(256), 0, s,
                           (const __bf16*)hs->ZbT, hs->n_pad, kpt, (const float*)hs->wb, ntile,
                           nsplit, sps, (float*)part);
      else
        hipLaunchKernelGGL(syrk_bf16_kernel, dim3((unsigned)(nlow * nsplit)), dim3(256), 0, s,
                           (const __bf16*)hs->ZbT, hs->n_pad, kpt, (const float*)hs->wb, ntile,
                           nsplit, sps, (float*)part);
      hipLaunchKernelGGL(syrk_accum_kernel, dim3((unsigned)ceil_div((int64_t)kp * kp, 256)),
                         dim3(256), 0, s, (const float*)part, nsplit, kpt, kp, M,
                         (const float*)nullptr);
    }
  }
  if (ev_bf16) TSVM_HIP(c, hipEventRecord(ev_bf16[1], s));
  TSVM_HIP(c, hipGetLastError());
  return TSVM_OK;
}

double hybrid_fp8_eps() { return HYBRID_FP8_EPS; }

double hybrid_syrk_flops(const HybridSyrk* hs) {
  // rows streamed by the BF16 kernel times the SYRK count k (k + 1)
  return (double)hs->n_pad * hs->kp * (hs->kp + 1.0);
}

tsvm_status weighted_syrk_dev(Ctx* c, const float* Z, int64_t n, int64_t k, const double* w,
                              int precision, double budget, double* M, tsvm_syrk_info* info) {
  if (n < 0 || k <= 0 || k > 4096 || !Z || !w || !M)
    return fail(c, TSVM_ERR_ARG, "weighted_syrk: bad arguments (n=%ld k=%ld)", (long)n, (long)k);
  if (precision != TSVM_PREC_FP32 && precision != TSVM_PREC_BF16 && precision != TSVM_PREC_FP8)
    return fail(c, TSVM_ERR_ARG, "weighted_syrk: unknown precision %d", precision);
  hipStream_t s = c->stream;
  tsvm_syrk_info inf;
  inf.n_fp64_rows = n;
  inf.bf16_bound = 0.0;
  inf.t_bf16 = 0.0;
  inf.bf16_tflops = 0.0;
  inf.n_fp8_rows = 0;
  inf.fp8_bound = 0.0;
  if (precision == TSVM_PREC_FP32 || n == 0) {
    TSVM_CHECK(gemm_tn_f64(c, Z, k, k, Z, k, k, n, w, true, M));
    hipLaunchKernelGGL(add_identity_kernel, dim3((unsigned)ceil_div(k, 256)), dim3(256), 0, s, M, k);
    TSVM_HIP(c, hipGetLastError());
  } else {
    HybridSyrk hs;
    TSVM_CHECK(hybrid_syrk_setup(c, Z, n, (int)k, &hs,
                                 precision == TSVM_PREC_FP8 ? HYBRID_FP8_EPS : 0.0));
    void* p;
    TSVM_HIP(c, c->ws.get("hs_slots", 32, &p));
    double* slots = (double*)p;
    TSVM_CHECK(hybrid_syrk_split(c, &hs, w, budget > 0 ? budget : 4.0, slots));
    double hsl[4] = {0, 0, 0, 0};
    TSVM_HIP(c, hipMemcpyAsync(hsl, slots, 32, hipMemcpyDeviceToHost, s));
    TSVM_HIP(c, hipStreamSynchronize(s));
    Event ev0, ev1;
    TSVM_HIP(c, ev0.create());
    TSVM_HIP(c, ev1.create());
    hipEvent_t ev[2] = {ev0, ev1};
    const int64_t nh = (int64_t)hsl[1];
    const int64_t n8 = hs.fp8_eps > 0.0 ? (int64_t)hsl[3] : 0;
    tsvm_status st = hybrid_syrk_form(c, &hs, Z, w, nh, M, ev, n8);
    if (st == TSVM_OK) {
      TSVM_HIP(c, hipStreamSynchronize(s));
      float ms = 0.f;
      TSVM_HIP(c, hipEventElapsedTime(&ms, ev[0], ev[1]));
      inf.n_fp64_rows = nh;
      inf.bf16_bound = hsl[0];
      inf.n_fp8_rows = n8;
      inf.fp8_bound = hsl[2];
      if (nh < n) {
        inf.t_bf16 = ms * 1e-3;
        if (ms > 0) inf.bf16_tflops = hybrid_syrk_flops(&hs) / (ms * 1e-3) * 1e-12;
      }
    }
    if (st != TSVM_OK) return st;
  }
  TSVM_HIP(c, hipStreamSynchronize(s));
  if (info) *info = inf;
  return TSVM_OK;
}

}  // namespace tsvm
