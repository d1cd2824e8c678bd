// Fused labelled-RBF kernel product on gfx950 MFMA.
//
//   W[i, c] = y1_i * sum_j exp(-gamma |x1_i - x2_j|^2) * (y2_j A[j, c])
//
// Replaces the reference's per-8192^2-tile pipeline (tensorsvm-train.cu:2476-2567):
// vecnorm (2569-2581) + XIJ GEMM (2506-2518) + rbf_kergen (2598-2612, which writes the
// tile to HBM) + the K*A GEMM (2537-2555).  Here one kernel keeps the exponentiated
// tile in registers, flash-attention style:
//
//   S^T (32 j x 32 i) = Qop_j . Pop_i^T      MFMA (the distance, in log2 units)
//   E = exp2(S^T)                             VALU v_exp_f32, packed to the MFMA A operand
//   O (i x c) += E . A'_j                     MFMA, O stays in registers over all j
//
// The distance is folded into one GEMM with augmented operands (norms ride along as
// extra K columns):  Pop_i . Qop_j = 2 g' x_i.x_j - g'|x_i|^2 - g'|x_j|^2 with
// g' = gamma*log2(e), so exp2 gives the kernel directly.  In the BF16 form each FP32
// value is split into bf16 hi + lo parts and the GEMM carries hi*hi + hi*lo + lo*hi
// (about 16 mantissa bits for the distance) while the big product E.A' is plain BF16
// with FP32 accumulation.  The fp16 forms keep the same distance GEMM and run E.A' on
// the fp16 MFMA (same rate as bf16, 3 more mantissa bits; E lies in (0, 1] and A' is
// brought into range by one power-of-two scale): FP16 plain, FP16X2 with E split into
// two fp16 terms, FP16R2 with A' split into two fp16 terms (two MFMA per tile for the
// split forms).  The FP32 form uses v_mfma_f32_32x32x2f32 throughout.
//
// Layout produced by the prep kernels (padding is zero-filled):
//   BF16: Pop  [m_pad][KD] bf16, Qop [n_pad][KD] bf16, KD = roundup(3d+4, 16)
//         AT   [n_pad/32][r_pad][32] bf16 = y2_j A[j,c], j permuted inside each chunk
//              so one 16-byte LDS read is a whole MFMA B operand (FP16X2: fp16 of
//              sc * y2_j A[j,c], sc = 2^e with sc * max|A| in [2^13, 2^14))
//   FP32: Pop  [m_pad][KF] f32, Qop [n_pad][KF] f32, KF = roundup(d+2, 2)
//         Af   [n_pad][r_pad] f32 = y2_j A[j,c]
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "common.h"
#include "internal.h"

namespace tsvm {

// 32-j chunks per pipeline stage of the BF16 kernel (the j padding is a multiple of it)
#define KBF16_CH 2
// Widest operand (KD) taken by the 512-thread fp16 form.  At KD = 176 and 256 its resident
// P rows push it past 256 VGPRs (339 and 192 spilled) while the 256-thread form spills
// none, so those widths keep 256-thread blocks (one per CU there: 84-105 KB of LDS each);
// at KD = 96-128 both forms spill and the 512-thread form less (4-94 against 92-129).
#define KBF16_WIDE_MAX_KD 128

// position of j (0..31) inside a permuted 32-j chunk row (see kprod_bf16 B operand)
__device__ __host__ inline int perm32(int j) {
  int kap = j >> 4, jj = j & 15;
  int half = (jj & 7) >> 2;
  int t = ((jj >> 3) << 2) | (jj & 3);
  return ((2 * kap + half) << 3) | t;
}

__device__ inline void split_bf16(float v, __bf16& hi, __bf16& lo) {
  hi = (__bf16)v;
  lo = (__bf16)(v - (float)hi);
}

// rows side: [xh, xh, xl, 1, 1, bh, bl, 0...]
__global__ void prep_rows_bf16(const float* __restrict__ X, int64_t m, int64_t m_pad, int d,
                               int KD, double gl2, __bf16* __restrict__ P) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= m_pad) return;
  __bf16* p = P + i * KD;
  const __bf16 z = (__bf16)0.0f;
  if (i >= m) {
    for (int t = 0; t < KD; t++) p[t] = z;
    return;
  }
  const float* x = X + i * d;
  double nrm = 0.0;
  for (int t = 0; t < d; t++) {
    float v = x[t];
    nrm += (double)v * v;
    __bf16 hi, lo;
    split_bf16(v, hi, lo);
    p[t] = hi;
    p[d + t] = hi;
    p[2 * d + t] = lo;
  }
  float b = (float)(-gl2 * nrm);
  __bf16 bh, bl;
  split_bf16(b, bh, bl);
  p[3 * d] = (__bf16)1.0f;
  p[3 * d + 1] = (__bf16)1.0f;
  p[3 * d + 2] = bh;
  p[3 * d + 3] = bl;
  for (int t = 3 * d + 4; t < KD; t++) p[t] = z;
}

// columns side: [uh, ul, uh, bh, bl, 1, 1, 0...] with u = 2 g' x
__global__ void prep_cols_bf16(const float* __restrict__ X, int64_t n, int64_t n_pad, int d,
                               int KD, double gl2, __bf16* __restrict__ Q) {
  int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= n_pad) return;
  __bf16* q = Q + j * KD;
  const __bf16 z = (__bf16)0.0f;
  if (j >= n) {
    for (int t = 0; t < KD; t++) q[t] = z;
    return;
  }
  const float* x = X + j * d;
  double nrm = 0.0;
  for (int t = 0; t < d; t++) {
    float v = x[t];
    nrm += (double)v * v;
    float u = (float)(2.0 * gl2 * (double)v);
    __bf16 hi, lo;
    split_bf16(u, hi, lo);
    q[t] = hi;
    q[d + t] = lo;
    q[2 * d + t] = hi;
  }
  float b = (float)(-gl2 * nrm);
  __bf16 bh, bl;
  split_bf16(b, bh, bl);
  q[3 * d] = bh;
  q[3 * d + 1] = bl;
  q[3 * d + 2] = (__bf16)1.0f;
  q[3 * d + 3] = (__bf16)1.0f;
  for (int t = 3 * d + 4; t < KD; t++) q[t] = z;
}

// AT[chunk][c][perm32(jj)] = bf16(y_j A[j, c]); a 32(j) x 32(c) tile per block (256 thr)
__global__ void prep_A_bf16(const float* __restrict__ A, int64_t lda, const float* __restrict__ y,
                            int64_t n, int64_t r, int64_t r_pad, __bf16* __restrict__ AT) {
  __shared__ float tile[32][33];
  const int64_t chunk = blockIdx.x;      // j chunk
  const int64_t cb = blockIdx.y * 32;    // column block
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int jj = ty; jj < 32; jj += 8) {
    int64_t j = chunk * 32 + jj, c = cb + tx;
    float v = 0.0f;
    if (j < n && c < r) v = y[j] * A[j * lda + c];
    tile[jj][tx] = v;
  }
  __syncthreads();
  for (int cc = ty; cc < 32; cc += 8) {
    int64_t c = cb + cc;
    if (c >= r_pad) continue;
    AT[(chunk * r_pad + c) * 32 + perm32(tx)] = (__bf16)tile[tx][cc];
  }
}

// max |y_j A[j, c]| over the n x r block, as the bits of a non-negative float
// (atomicMax on the int view orders non-negative floats correctly)
__global__ void amax_kernel(const float* __restrict__ A, int64_t lda, int64_t n, int64_t r,
                            unsigned int* __restrict__ out) {
  float mx = 0.0f;
  const int64_t total = n * r;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t j = idx / r, c = idx - j * r;
    mx = fmaxf(mx, fabsf(A[j * lda + c]));
  }
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  __shared__ float part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) mx = fmaxf(mx, part[w]);
    atomicMax(out, __float_as_uint(mx));
  }
}

// clears the amax slot (a kernel rather than a memset node, so that the call can be
// captured into a graph and replayed)
__global__ void zero_u32_kernel(unsigned int* __restrict__ p) { *p = 0u; }

// power-of-two scale sc with sc * amax in [2^13, 2^14) (1 for an all-zero A): slot 0
// holds sc, slot 1 holds 1/sc (both exact)
__global__ void fp16_scale_kernel(const unsigned int* __restrict__ amax_bits,
                                  float* __restrict__ sc) {
  float a = __uint_as_float(*amax_bits);
  int e = 0;
  if (a > 0.0f && a < INFINITY) {
    int ex;
    (void)frexpf(a, &ex);  // a = f 2^ex, f in [0.5, 1)
    e = 14 - ex;
    e = e > 100 ? 100 : (e < -100 ? -100 : e);
  }
  sc[0] = ldexpf(1.0f, e);
  sc[1] = ldexpf(1.0f, -e);
}

// fp16 right-hand side: AT[chunk][c][perm32(jj)] = fp16(sc * y_j A[j, c]); with R2 the
// row carries 64 values, hi at [0, 32) and lo = fp16(sc * y_j A[j, c] - hi) at [32, 64)
template <bool R2>
__global__ void prep_A_f16(const float* __restrict__ A, int64_t lda, const float* __restrict__ y,
                           int64_t n, int64_t r, int64_t r_pad, const float* __restrict__ sc,
                           _Float16* __restrict__ AT) {
  constexpr int AW = R2 ? 64 : 32;
  __shared__ float tile[32][33];
  const int64_t chunk = blockIdx.x;
  const int64_t cb = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const float scale = sc[0];
  for (int jj = ty; jj < 32; jj += 8) {
    int64_t j = chunk * 32 + jj, c = cb + tx;
    float v = 0.0f;
    if (j < n && c < r) v = y[j] * A[j * lda + c];
    tile[jj][tx] = v * scale;
  }
  __syncthreads();
  for (int cc = ty; cc < 32; cc += 8) {
    int64_t c = cb + cc;
    if (c >= r_pad) continue;
    const float v = tile[tx][cc];
    const _Float16 hi = (_Float16)v;
    _Float16* row = AT + (chunk * r_pad + c) * AW;
    row[perm32(tx)] = hi;
    if (R2) row[32 + perm32(tx)] = (_Float16)(v - (float)hi);
  }
}

// ---- operands in MFMA fragment order (the wide-tile kernel) ------------------------
// Qf[chunk][ks][lane][8]: the A operand of the distance MFMA for 32 columns j of X2 and
// K-step ks: lane l holds Qop[j = 32 chunk + l % 32][16 ks + 8 (l / 32) + 0..7].  A chunk
// is one contiguous run of 64 KD bytes.
__device__ inline int64_t qfrag_index(int64_t j, int t) {  // offset within j's chunk
  const int jr = (int)(j & 31), ks = t >> 4, hh = (t >> 3) & 1, e = t & 7;
  return ((int64_t)ks * 64 + (jr + 32 * hh)) * 8 + e;
}
__global__ void prep_cols_bf16_frag(const float* __restrict__ X, int64_t n, int64_t n_pad, int d,
                                    int KD, double gl2, __bf16* __restrict__ Qf) {
  int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= n_pad) return;
  __bf16* q = Qf + (j >> 5) * (int64_t)(32 * KD);
  auto put = [&](int t, __bf16 v) { q[qfrag_index(j, t)] = v; };
  const __bf16 z = (__bf16)0.0f;
  if (j >= n) {
    for (int t = 0; t < KD; t++) put(t, z);
    return;
  }
  const float* x = X + j * d;
  double nrm = 0.0;
  for (int t = 0; t < d; t++) {
    float v = x[t];
    nrm += (double)v * v;
    float u = (float)(2.0 * gl2 * (double)v);
    __bf16 hi, lo;
    split_bf16(u, hi, lo);
    put(t, hi);
    put(d + t, lo);
    put(2 * d + t, hi);
  }
  float bb = (float)(-gl2 * nrm);
  __bf16 bh, bl;
  split_bf16(bb, bh, bl);
  put(3 * d, bh);
  put(3 * d + 1, bl);
  put(3 * d + 2, (__bf16)1.0f);
  put(3 * d + 3, (__bf16)1.0f);
  for (int t = 3 * d + 4; t < KD; t++) put(t, z);
}

// Af[chunk][ct][kap][lane][tt]: the B operand of the product MFMA for K-step kap of a
// 32-j chunk and column tile ct: lane l holds A'[j][c] for c = 32 ct + l % 32 and
// j = 32 chunk + 16 kap + 8 (tt / 4) + 4 (l / 32) + tt % 4 -- the order in which the
// distance MFMA's C layout leaves the kernel values in the A operand.  One chunk's column
// tiles are consecutive 2 KB runs.  EM_BF16: bf16 of y_j A[j,c]; otherwise fp16 of
// sc y_j A[j,c] (sc from fp16_scale_kernel).
template <typename ET>
__global__ void prep_A_frag(const float* __restrict__ A, int64_t lda, const float* __restrict__ y,
                            int64_t n, int64_t r, int64_t r_tiles, const float* __restrict__ sc,
                            ET* __restrict__ Af) {
  __shared__ float tile[32][33];
  const int64_t chunk = blockIdx.x;
  const int64_t ct = blockIdx.y;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const float scale = sc ? sc[0] : 1.0f;
  for (int jj = ty; jj < 32; jj += 8) {
    int64_t j = chunk * 32 + jj, c = ct * 32 + tx;
    float v = 0.0f;
    if (j < n && c < r) v = y[j] * A[j * lda + c];
    tile[jj][tx] = v * scale;
  }
  __syncthreads();
  ET* out = Af + (chunk * r_tiles + ct) * 1024;
  // thread -> (lane, tt) pairs: 64 lanes x 16 values (2 kap x 8 tt) = 1024 per tile
  for (int e = threadIdx.x; e < 1024; e += 256) {
    const int kap = e >> 9, lane = (e >> 3) & 63, tt = e & 7;
    const int hh = lane >> 5, cc = lane & 31;
    const int jr = 16 * kap + 8 * (tt >> 2) + 4 * hh + (tt & 3);
    out[e] = (ET)tile[jr][cc];
  }
}

__global__ void prep_rows_f32(const float* __restrict__ X, int64_t m, int64_t m_pad, int d, int KF,
                              double gl2, float* __restrict__ P) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= m_pad) return;
  float* p = P + i * KF;
  if (i >= m) {
    for (int t = 0; t < KF; t++) p[t] = 0.0f;
    return;
  }
  const float* x = X + i * d;
  double nrm = 0.0;
  for (int t = 0; t < d; t++) {
    nrm += (double)x[t] * x[t];
    p[t] = x[t];
  }
  p[d] = 1.0f;
  p[d + 1] = (float)(-gl2 * nrm);
  for (int t = d + 2; t < KF; t++) p[t] = 0.0f;
}

__global__ void prep_cols_f32(const float* __restrict__ X, int64_t n, int64_t n_pad, int d, int KF,
                              double gl2, float* __restrict__ Q) {
  int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= n_pad) return;
  float* q = Q + j * KF;
  if (j >= n) {
    for (int t = 0; t < KF; t++) q[t] = 0.0f;
    return;
  }
  const float* x = X + j * d;
  double nrm = 0.0;
  for (int t = 0; t < d; t++) {
    nrm += (double)x[t] * x[t];
    q[t] = (float)(2.0 * gl2 * (double)x[t]);
  }
  q[d] = (float)(-gl2 * nrm);
  q[d + 1] = 1.0f;
  for (int t = d + 2; t < KF; t++) q[t] = 0.0f;
}

__global__ void prep_A_f32(const float* __restrict__ A, int64_t lda, const float* __restrict__ y,
                           int64_t n, int64_t n_pad, int64_t r, int64_t r_pad, float* __restrict__ Af) {
  int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= n_pad * r_pad) return;
  int64_t j = idx / r_pad, c = idx - j * r_pad;
  float v = 0.0f;
  if (j < n && c < r) v = y[j] * A[j * lda + c];
  Af[idx] = v;
}

// -------------------------------------------------------------------------------------
// Low-precision main kernel.  THR threads = THR/64 waves (256, or 512 for the fp16
// 128-column form with KD <= 128: one block per CU whose eight waves share each staged
// stage); wave w
// owns rows [64w, 64w+64) of the block's THR rows (two 32-row MFMA blocks) and all
// BN = 32*NT columns of the block.
// Per 32-j chunk: S^T via KD/16 MFMA 32x32x16 bf16 per row block, exp2 on 16 values per
// lane per row block, then the E.A' MFMA (2*NT per row block and term) for O += E.A'.
// LDS (double-buffered): Qop chunk 32 x KD, AT chunk BN x AW (row stride padded).
//
// EM selects the E.A' arithmetic:
//   EM_BF16   E bf16, A' bf16                      1 MFMA per tile, E to 2^-9
//   EM_F16    E fp16, A' fp16 (scaled)             1 MFMA per tile, E and A' to 2^-11
//   EM_F16X2  E = e_hi + e_lo fp16, A' fp16        2 MFMA per tile, E to ~2^-22
//   EM_F16R2  E fp16, A' = a_hi + a_lo fp16        2 MFMA per tile, A' to ~2^-22
// -------------------------------------------------------------------------------------
enum { EM_BF16 = 0, EM_F16 = 1, EM_F16X2 = 2, EM_F16R2 = 3 };

template <int NT, int KD, int CH, int EM, int OCC, int THR = 256>
__global__ __launch_bounds__(THR, OCC) void kprod_bf16(const __bf16* __restrict__ P,
                                                  const __bf16* __restrict__ Q,
                                                  const void* __restrict__ ATv,
                                                  const float* __restrict__ y1,
                                                  float* __restrict__ W, int64_t ldw, int64_t m,
                                                  int64_t r, int64_t n_stages, int64_t r_pad,
                                                  const float* __restrict__ sc, int64_t nxb,
                                                  int64_t ntiles) {
  using ET = typename std::conditional<EM == EM_BF16, __bf16, _Float16>::type;
  using ET8 = typename std::conditional<EM == EM_BF16, bf16x8_t, f16x8_t>::type;
  const ET* __restrict__ AT = (const ET*)ATv;
  constexpr int AW = EM == EM_F16R2 ? 64 : 32;  // AT elements per (chunk, column): hi | lo
  constexpr int BN = NT * 32;
  constexpr int KS = KD / 16;
  constexpr int QS = KD + 8;        // LDS row stride (bf16) of the Qop stage
  constexpr int AS = AW * CH + 8;   // LDS row stride (elements) of the AT stage
  constexpr int QV = 32 * CH * KD / 8;   // 16-byte vectors per Qop stage (contiguous)
  constexpr int AV = BN * AW * CH / 8;   // 16-byte vectors per AT stage
  constexpr int NQ = (QV + THR - 1) / THR;
  constexpr int NA = (AV + THR - 1) / THR;
  __shared__ __attribute__((aligned(16))) __bf16 sQ[2][32 * CH * QS];
  __shared__ __attribute__((aligned(16))) ET sA[2][BN * AS];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int l32 = lane & 31, h = lane >> 5;
  // Persistent grid: the block walks output tiles blockIdx.x, +gridDim.x, ...; a round of
  // resident blocks covers consecutive row blocks of one column block, so they stream
  // the same Qop and AT chunks at nearly the same time and share them through L2.
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
  const int64_t rowbase = (tile % nxb) * THR + wave * 64;
  const int64_t c0 = (tile / nxb) * BN;

  // S^T B operand: this wave's P rows, resident in registers for the whole j loop.
  bf16x8_t pb[2][KS];
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int ks = 0; ks < KS; ks++)
      pb[rb][ks] = *(const bf16x8_t*)(P + (rowbase + 32 * rb + l32) * KD + ks * 16 + 8 * h);

  f32x16_t acc[2][NT];
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int ct = 0; ct < NT; ct++)
#pragma unroll
      for (int q = 0; q < 16; q++) acc[rb][ct][q] = 0.0f;

  // Staging: the Qop stage is one contiguous run; the AT stage is CH chunks, each a
  // contiguous [r_pad][AW] block of which this column block reads BN rows.  Consecutive
  // lanes take one whole LDS row of the AT stage (the CH chunks' pieces side by side), so
  // a group of lanes stores one padded row instead of half-rows of two.  Loads are
  // branch-free (out-of-range lanes re-read a valid vector and skip the store) so that
  // they stay in flight across the compute of the current stage.
  u32x4_t sq[NQ], sa[NA];
  auto gload = [&](int64_t t) {
    const u32x4_t* gq = (const u32x4_t*)(Q + t * (32 * CH) * KD);
#pragma unroll
    for (int v = 0; v < NQ; v++) {
      int idx = tid + v * THR;
      sq[v] = gq[idx < QV ? idx : QV - 1];
    }
#pragma unroll
    for (int v = 0; v < NA; v++) {
      int idx = tid + v * THR;
      idx = idx < AV ? idx : AV - 1;
      const int row = idx / (CH * AW / 8), j = idx - row * (CH * AW / 8);
      const int ch = j / (AW / 8), q = j - ch * (AW / 8);
      const u32x4_t* ga = (const u32x4_t*)(AT + ((t * CH + ch) * r_pad + c0) * AW);
      sa[v] = ga[row * (AW / 8) + q];
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int v = 0; v < NQ; v++) {
      int idx = tid + v * THR;
      if (QV % THR == 0 || idx < QV) {
        int row = idx / (KD / 8), col = (idx % (KD / 8)) * 8;
        *(u32x4_t*)&sQ[buf][row * QS + col] = sq[v];
      }
    }
#pragma unroll
    for (int v = 0; v < NA; v++) {
      int idx = tid + v * THR;
      if (AV % THR == 0 || idx < AV) {
        const int row = idx / (CH * AW / 8), j = idx - row * (CH * AW / 8);
        *(u32x4_t*)&sA[buf][row * AS + j * 8] = sa[v];
      }
    }
  };

  gload(0);
  sstore(0);
  __syncthreads();
  for (int64_t t = 0; t < n_stages; t++) {
    const int buf = (int)(t & 1);
    // prefetch the next stage into registers (the last iteration re-reads its own
    // stage: unconditional, so the loads stay in flight across the MFMA work below)
    gload(t + 1 < n_stages ? t + 1 : t);
    __builtin_amdgcn_sched_barrier(0);  // issue the prefetch before the MFMA work
#pragma unroll
    for (int ch = 0; ch < CH; ch++) {
      // distance tile S^T (32 j x 32 i) for both row blocks
      f32x16_t s[2];
#pragma unroll
      for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int q = 0; q < 16; q++) s[rb][q] = 0.0f;
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        bf16x8_t qa = *(const bf16x8_t*)&sQ[buf][(ch * 32 + l32) * QS + ks * 16 + 8 * h];
        s[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(qa, pb[0][ks], s[0], 0, 0, 0);
        s[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(qa, pb[1][ks], s[1], 0, 0, 0);
      }
      // E = exp2(S^T): lane holds i = l32 and j = 8(q/4) + 4h + q%4; registers 8k..8k+7
      // form the A operand of K-half k (j in 16k + {4h..4h+3, 8+4h..8+4h+3}).
      ET8 e[2][2], el[2][2];
#pragma unroll
      for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int kap = 0; kap < 2; kap++)
#pragma unroll
          for (int tt = 0; tt < 8; tt++) {
            const float v = __builtin_amdgcn_exp2f(s[rb][8 * kap + tt]);
            const ET hi = (ET)v;
            e[rb][kap][tt] = hi;
            if constexpr (EM == EM_F16X2) el[rb][kap][tt] = (ET)(v - (float)hi);
          }
      // O += E . A'_j (per term of the split forms).  In the 512-thread form the wave
      // raises its issue priority over the product MFMAs, so that of the two waves on a
      // SIMD the one with MFMAs ready keeps the matrix pipe fed while the other runs its
      // exp2 work (3.14 -> 3.06 s per product at 1.6M x 1.6M x 512).
      if constexpr (THR == 512) __builtin_amdgcn_s_setprio(1);
#pragma unroll
      for (int ct = 0; ct < NT; ct++) {
#pragma unroll
        for (int kap = 0; kap < 2; kap++) {
          const int base = (ct * 32 + l32) * AS + ch * AW + (2 * kap + h) * 8;
          const ET8 bv = *(const ET8*)&sA[buf][base];
          if constexpr (EM == EM_BF16) {
            acc[0][ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(e[0][kap], bv, acc[0][ct], 0, 0, 0);
            acc[1][ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(e[1][kap], bv, acc[1][ct], 0, 0, 0);
          } else {
            acc[0][ct] = __builtin_amdgcn_mfma_f32_32x32x16_f16(e[0][kap], bv, acc[0][ct], 0, 0, 0);
            acc[1][ct] = __builtin_amdgcn_mfma_f32_32x32x16_f16(e[1][kap], bv, acc[1][ct], 0, 0, 0);
          }
          if constexpr (EM == EM_F16X2) {
            acc[0][ct] = __builtin_amdgcn_mfma_f32_32x32x16_f16(el[0][kap], bv, acc[0][ct], 0, 0, 0);
            acc[1][ct] = __builtin_amdgcn_mfma_f32_32x32x16_f16(el[1][kap], bv, acc[1][ct], 0, 0, 0);
          }
          if constexpr (EM == EM_F16R2) {
            const ET8 bl = *(const ET8*)&sA[buf][base + 32];
            acc[0][ct] = __builtin_amdgcn_mfma_f32_32x32x16_f16(e[0][kap], bl, acc[0][ct], 0, 0, 0);
            acc[1][ct] = __builtin_amdgcn_mfma_f32_32x32x16_f16(e[1][kap], bl, acc[1][ct], 0, 0, 0);
          }
        }
      }
      if constexpr (THR == 512) __builtin_amdgcn_s_setprio(0);
    }
    // keep the LDS stores of the next stage after this stage's MFMA work
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("" ::: "memory");
    sstore(buf ^ 1);
    __syncthreads();
  }
  // epilogue: W[i, c] = y1_i * O[i, c]; acc register q holds row 8(q/4) + 4h + q%4.  The
  // fp16 forms undo the right-hand side's power-of-two scale (exact).
  const float unscale = EM == EM_BF16 ? 1.0f : sc[1];
#pragma unroll
  for (int rb = 0; rb < 2; rb++) {
#pragma unroll
    for (int q = 0; q < 16; q++) {
      int64_t i = rowbase + 32 * rb + 8 * (q / 4) + 4 * h + (q % 4);
      if (i >= m) continue;
      float yi = y1[i] * unscale;
#pragma unroll
      for (int ct = 0; ct < NT; ct++) {
        int64_t c = c0 + ct * 32 + l32;
        if (c < r) W[i * ldw + c] = yi * acc[rb][ct][q];
      }
    }
  }
  }  // tile
}

// The distance MFMA of the wide-tile kernel with its result in architectural VGPRs.  The
// kernel's 256 product accumulators fill the accumulation registers; the compiler would
// give this MFMA an accumulation register too and move a product tile out and back around
// it (96 register moves per 32-j chunk).  Its inputs come from LDS reads and from registers
// loaded once per tile (never from a VALU write just before), its result is read by v_exp
// only after several later MFMAs (the loop body's groups are fenced with sched_barrier), and
// back-to-back accumulation into the same registers needs no wait states, as for the
// compiler's own MFMA chains.
template <bool LAST = false>
__device__ __forceinline__ f32x16_t mfma_bf16_vgpr_first(bf16x8_t a, bf16x8_t b) {
  f32x16_t d;
  if constexpr (LAST)
    asm volatile("v_mfma_f32_32x32x16_bf16 %0, %1, %2, 0\n\ts_nop 7\n\ts_nop 1"
                 : "=&v"(d) : "v"(a), "v"(b));
  else
    asm volatile("v_mfma_f32_32x32x16_bf16 %0, %1, %2, 0" : "=&v"(d) : "v"(a), "v"(b));
  return d;
}
__device__ __forceinline__ void mfma_bf16_vgpr_acc(f32x16_t& d, bf16x8_t a, bf16x8_t b) {
  asm volatile("v_mfma_f32_32x32x16_bf16 %0, %1, %2, %0" : "+v"(d) : "v"(a), "v"(b));
}
// the last MFMA of a chain whose result VALU instructions read next: ten wait states
// after it (an 8-pass XDL result read by VALU; the MFMA pipe is busy meanwhile anyway)
__device__ __forceinline__ void mfma_bf16_vgpr_last(f32x16_t& d, bf16x8_t a, bf16x8_t b) {
  asm volatile("v_mfma_f32_32x32x16_bf16 %0, %1, %2, %0\n\ts_nop 7\n\ts_nop 1"
               : "+v"(d) : "v"(a), "v"(b));
}

// -------------------------------------------------------------------------------------
// Wide-tile low-precision kernel (right-hand sides wider than 128 columns).  WAVES waves
// per block, one block per CU; wave w owns RB 32-row blocks and NT 32-column tiles of the
// output (RB NT 16 accumulator registers).  Used as NT = 8, WAVES = 4, RB = 2: one wave per
// SIMD holding 64 rows x 256 columns in all 256 accumulation registers, so that every B
// fragment read from LDS feeds two MFMAs (with 32 rows per wave each product MFMA needed
// its own 1 KB fragment), and a distance tile and its exp2 are formed once per 256
// columns (8 distance MFMAs per 32 product MFMAs).  The distance MFMAs write
// architectural registers (mfma_bf16_vgpr_*): with the product tiles filling the
// accumulation registers the compiler would otherwise shuttle a product tile out and back
// around each of them.  Operands come in MFMA fragment order (Qf, Af above): a chunk of
// 32 j is one contiguous run in global memory and in LDS, staging is a linear copy and
// every fragment read is a 16-byte read at lane * 16 (no padding, no bank conflicts).
// LDS ring of three chunk buffers: iteration c reads A'_c and Q_{c+1} and writes chunk
// c + 2; one barrier per chunk.
// The loop body is laid out by hand in groups fenced with sched_barrier, the two row
// blocks' product MFMAs for one B fragment per group:
//  * the groups take K step 0 of the eight column tiles, then K step 1, so the two updates
//    of an accumulator within a chunk are eight groups apart (back to back, each waited
//    for the other);
//  * B fragments (A'_c) come through a ring of four registers, read four groups ahead; the
//    first four of chunk c + 1 are read before the barrier that ends chunk c (its buffer
//    was written two chunks earlier), so no product MFMA waits on a fresh LDS read;
//  * the Q fragments of chunk c + 1 are read after the barrier, its distance MFMAs sit in
//    the first groups and the exp2 / conversion of the next E tile (v_exp_f32 issues at a
//    quarter of the VALU rate) is spread over the rest.
// Measured at 400k x 400k x 512 on MI355X: the MFMA pipe is busy 58% of the cycles at
// 1.91 GHz.  The chip's power limit bounds what a better schedule can win: with the global
// loads and LDS stores taken out (a timing copy, wrong results) the pipe is busy 75% but
// the clock falls to 1.72 GHz (17% more MFMA work per second), and a register-only loop of
// fp16 MFMAs on random operands runs at 1.64 PF/s, 66% of the 2.5 PF/s dense peak.
// The blocks of an XCD stream the same chunks of A' and Q and must start each tile
// together to share them in L2 (in a persistent launch where every block walked its own
// tiles they drifted apart: at 1.6M x 1.6M x 512 the HBM fetch grew 30-fold and the clock
// fell from 1.9 to 1.6 GHz).  With sched, one cooperative launch runs the XCD-round
// schedule (kp_next_unit): the blocks of an XCD take a unit of consecutive tiles together
// and the next unit once all have finished, from a counter that every XCD draws on, so a
// faster XCD takes more units.  Without it the block handles tiles tile0 + blockIdx.x,
// + gridDim.x, ... below ntiles and the host launches one round of grid tiles at a time.
// The outputs are bit-identical to the 128-column kernel's (same operands, same
// accumulation order) on either schedule.
// -------------------------------------------------------------------------------------
// XCD-round schedule of the persistent wide launch.  Workgroups go to the eight XCDs
// round-robin, so block b runs on XCD b % 8; the gs = gridDim.x / 8 blocks of an XCD take
// the gs consecutive tiles of one unit together (they stream the same column chunks
// through their XCD's L2), and the XCD's block 0 takes the next unit from a device-wide
// counter once all gs have finished the last one: an XCD that runs faster takes more
// units, where one launch per round of tiles waits for the slowest XCD every round.
// sched (zeroed before the launch): [0] next unit, [1 + x] (round << 32) | unit published
// for XCD x, [1 + 8 + x] tiles finished on XCD x, [17] error flag.  The waits are bounded
// (the launch is cooperative, so every block is resident and they end; the bound, in polls
// of the counter, is set from the tile's length with a wide margin); a wait that runs out
// sets the flag and the blocks leave the loop.
constexpr int KP_XCDS = 8;
constexpr int KP_SCHED_WORDS = 2 + 2 * KP_XCDS;
__device__ __forceinline__ unsigned long long kp_ld(unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// called by one thread per block; returns the unit of this round, or -1 after a failed wait
__device__ long long kp_next_unit(unsigned long long* sched, int xcd, int slot, int gs,
                                  long long round, long long spin_max) {
  unsigned long long* pub = sched + 1 + xcd;
  unsigned long long* done = sched + 1 + KP_XCDS + xcd;
  unsigned long long* err = sched + 1 + 2 * KP_XCDS;
  long long spins = 0;
  if (slot == 0) {
    while (kp_ld(done) < (unsigned long long)gs * (unsigned long long)round) {
      if (++spins > spin_max) {
        atomicOr(err, 1ull);
        return -1;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    const unsigned long long u = atomicAdd(sched, 1ull);
    __hip_atomic_store(pub, ((unsigned long long)(round + 1) << 32) | u, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    return (long long)u;
  }
  for (;;) {
    const unsigned long long v = kp_ld(pub);
    if ((v >> 32) >= (unsigned long long)(round + 1)) return (long long)(v & 0xffffffffull);
    if (++spins > spin_max) {
      atomicOr(err, 1ull);
      return -1;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

template <int NT, int KD, int EM, int WAVES, int RB>
__global__ __launch_bounds__(64 * WAVES, 1) void kprod_wide(const __bf16* __restrict__ P,
                                                      const __bf16* __restrict__ Qf,
                                                      const void* __restrict__ Afv,
                                                      const float* __restrict__ y1,
                                                      float* __restrict__ W, int64_t ldw,
                                                      int64_t m, int64_t r, int64_t n_chunks,
                                                      int64_t r_tiles, const float* __restrict__ sc,
                                                      int64_t nxb, int64_t ntiles, int64_t tile0,
                                                      unsigned long long* __restrict__ sched) {
  using ET = typename std::conditional<EM == EM_BF16, __bf16, _Float16>::type;
  using ET8 = typename std::conditional<EM == EM_BF16, bf16x8_t, f16x8_t>::type;
  constexpr int KS = KD / 16;
  constexpr int NG = 2 * NT;            // B fragments (groups) per chunk
  constexpr int NE = RB * 16;           // exp2 per lane per chunk
  constexpr int QB = 64 * KD;
  constexpr int AB = 2048 * NT;
  constexpr int CB = QB + AB;
  constexpr int THR = 64 * WAVES;
  constexpr int NV = (CB / 16 + THR - 1) / THR;
  constexpr int CBP = NV * THR * 16;
  constexpr int BR = 32 * RB * WAVES;
  // Groups D0 .. D0 + KS - 1 carry the distance MFMAs of the next chunk (one K step of
  // every row block per group), the groups from X0 on its exp2 work (v_exp_f32 issues at a
  // quarter of the VALU rate: three per group of two product MFMAs keep the MFMA pipe fed)
  constexpr int D0 = 1;
  constexpr int X0 = D0 + KS;
  constexpr int XPG = (NE + (NG - X0) - 1) / (NG - X0);  // exp2 per group
  static_assert(KS >= 1 && X0 < NG, "distance MFMAs must fit before the exp2 groups");
  __shared__ __attribute__((aligned(16))) unsigned char lds[3 * CBP];
  const unsigned char* __restrict__ Af = (const unsigned char*)Afv;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int l32 = lane & 31, h = lane >> 5;

  // sched: the XCD-round schedule (one persistent launch); null: tiles tile0 + blockIdx.x,
  // + gridDim.x, ... (one launch per round)
  const bool dyn = sched != nullptr;
  const int gs = (int)(gridDim.x / KP_XCDS);
  const int xcd = (int)(blockIdx.x % KP_XCDS), slot = (int)(blockIdx.x / KP_XCDS);
  const int64_t nunits = dyn ? (ntiles + gs - 1) / gs : 0;
  __shared__ long long s_unit;
  long long round = 0;
  for (int64_t tile = tile0 + blockIdx.x;; tile += gridDim.x) {
    if (dyn) {
      // the wait for a unit lasts about a tile (n_chunks chunks of 32 columns, some 2000
      // cycles each): 2^20 polls plus 2^12 per chunk
      if (tid == 0) s_unit = kp_next_unit(sched, xcd, slot, gs, round, (1ll << 20) + (n_chunks << 12));
      __syncthreads();
      const long long u = s_unit;
      __syncthreads();  // s_unit is rewritten next round
      if (u < 0 || u >= nunits) break;
      round++;
      tile = u * gs + slot;
      if (tile >= ntiles) {  // the last unit is partial: this block's share is empty
        if (tid == 0) atomicAdd(sched + 1 + KP_XCDS + xcd, 1ull);
        continue;
      }
    } else if (tile >= ntiles) {
      break;
    }
    const int64_t rowbase = (tile % nxb) * BR + wave * 32 * RB;
    const int64_t ct0 = (tile / nxb) * NT;
    bf16x8_t pb[RB][KS];
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
      for (int ks = 0; ks < KS; ks++)
        pb[rb][ks] = *(const bf16x8_t*)(P + (rowbase + 32 * rb + l32) * KD + ks * 16 + 8 * h);
    f32x16_t acc[RB][NT];
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
      for (int ct = 0; ct < NT; ct++)
#pragma unroll
        for (int q = 0; q < 16; q++) acc[rb][ct][q] = 0.0f;
    u32x4_t stg[NV];
    auto gload = [&](int64_t t) {
      const u32x4_t* gq = (const u32x4_t*)((const unsigned char*)Qf + t * QB);
      const u32x4_t* ga = (const u32x4_t*)(Af + (t * r_tiles + ct0) * 2048);
#pragma unroll
      for (int v = 0; v < NV; v++) {
        const int idx = tid + v * THR;
        const int ia = idx - QB / 16;
        const u32x4_t* pq = gq + (idx < QB / 16 ? idx : 0);
        const u32x4_t* pa = ga + (ia < 0 ? 0 : (ia < AB / 16 ? ia : AB / 16 - 1));
        stg[v] = *(idx < QB / 16 ? pq : pa);
      }
    };
    auto sstore = [&](int buf) {
#pragma unroll
      for (int v = 0; v < NV; v++) *(u32x4_t*)(lds + buf * CBP + (tid + v * THR) * 16) = stg[v];
    };
    auto bread = [&](int buf, int g) { return *(const ET8*)(lds + buf * CBP + QB + g * 1024 + lane * 16); };
    auto qread = [&](int buf, int ks) { return *(const bf16x8_t*)(lds + buf * CBP + ks * 1024 + lane * 16); };

    const int64_t nc = n_chunks;
    gload(0);
    sstore(0);
    gload(nc > 1 ? 1 : 0);
    sstore(1);
    __syncthreads();
    ET8 e[RB][2];
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
      f32x16_t s0;
#pragma unroll
      for (int q = 0; q < 16; q++) s0[q] = 0.0f;
#pragma unroll
      for (int ks = 0; ks < KS; ks++)
        s0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(qread(0, ks), pb[rb][ks], s0, 0, 0, 0);
#pragma unroll
      for (int kap = 0; kap < 2; kap++)
#pragma unroll
        for (int tt = 0; tt < 8; tt++) e[rb][kap][tt] = (ET)__builtin_amdgcn_exp2f(s0[8 * kap + tt]);
    }
    // B fragment of group g: column tile g % NT, K step g / NT (fragment order: ct, kap)
    auto bfrag = [](int g) { return (g % NT) * 2 + g / NT; };
    ET8 b[4];
#pragma unroll
    for (int i = 0; i < 4; i++) b[i] = bread(0, bfrag(i));
    int bc = 0;
    for (int64_t c = 0; c < nc; c++) {
      const int bn = bc == 2 ? 0 : bc + 1;
      const int bw = bn == 2 ? 0 : bn + 1;
      gload(c + 2 < nc ? c + 2 : nc - 1);
      bf16x8_t qf[KS];
#pragma unroll
      for (int ks = 0; ks < KS; ks++) qf[ks] = qread(bn, ks);
      f32x16_t sd[RB];
      ET8 en[RB][2];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int g = 0; g < NG; g++) {
        // K step kap of every column tile, then the other: the two updates of one
        // accumulator are NT groups apart (back to back they wait on each other)
        const int ct = g % NT, kap = g / NT;
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
          if constexpr (EM == EM_BF16)
            acc[rb][ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(e[rb][kap], b[g & 3], acc[rb][ct], 0, 0, 0);
          else
            acc[rb][ct] = __builtin_amdgcn_mfma_f32_32x32x16_f16(e[rb][kap], b[g & 3], acc[rb][ct], 0, 0, 0);
        }
        if (g + 4 < NG) b[g & 3] = bread(bc, bfrag(g + 4));
        if (g >= D0 && g < X0) {
#pragma unroll
          for (int rb = 0; rb < RB; rb++) {
            if (g == D0)
              sd[rb] = mfma_bf16_vgpr_first<KS == 1>(qf[0], pb[rb][0]);
            else if (g < X0 - 1)
              mfma_bf16_vgpr_acc(sd[rb], qf[g - D0], pb[rb][g - D0]);
            else
              mfma_bf16_vgpr_last(sd[rb], qf[g - D0], pb[rb][g - D0]);
          }
        }
        if (g >= X0) {
#pragma unroll
          for (int u = 0; u < XPG; u++) {
            const int q = (g - X0) * XPG + u;
            if (q < NE) en[q >> 4][(q >> 3) & 1][q & 7] = (ET)__builtin_amdgcn_exp2f(sd[q >> 4][q & 15]);
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int rb = 0; rb < RB; rb++) {
        e[rb][0] = en[rb][0];
        e[rb][1] = en[rb][1];
      }
      // the first B fragments of chunk c + 1 (buffer bn, complete since the last barrier)
#pragma unroll
      for (int i = 0; i < 4; i++) b[i] = bread(bn, bfrag(i));
      sstore(bw);
      __syncthreads();
      bc = bn;
    }
    const float unscale = EM == EM_BF16 ? 1.0f : sc[1];
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int64_t i = rowbase + 32 * rb + 8 * (q / 4) + 4 * h + (q % 4);
        if (i >= m) continue;
        const float yi = y1[i] * unscale;
#pragma unroll
        for (int ct = 0; ct < NT; ct++) {
          const int64_t cc = (ct0 + ct) * 32 + l32;
          if (cc < r) W[i * ldw + cc] = yi * acc[rb][ct][q];
        }
      }
    __syncthreads();
    if (dyn && tid == 0) atomicAdd(sched + 1 + KP_XCDS + xcd, 1ull);
  }
}

// -------------------------------------------------------------------------------------
// FP32 main kernel.  256 threads = 4 waves; wave w owns rows [32w, 32w+32) of the block's
// 128 rows and all BN = 32*NT columns.  MFMA 32x32x2 f32 throughout.  LDS (single
// buffer, register prefetch): Qop chunk 32 x KF, A' chunk 32 x BN (row stride padded).
// -------------------------------------------------------------------------------------
template <int NT, int KF>
__global__ __launch_bounds__(256) void kprod_f32(const float* __restrict__ P,
                                                 const float* __restrict__ Q,
                                                 const float* __restrict__ Af,
                                                 const float* __restrict__ y1,
                                                 float* __restrict__ W, int64_t ldw, int64_t m,
                                                 int64_t r, int64_t n_chunks, int64_t r_pad,
                                                 int64_t seg_chunks, int64_t seg_stride) {
  constexpr int BN = NT * 32;
  constexpr int KSF = KF / 2;
  constexpr int AS = BN + 8;          // LDS row stride (floats) of the A' chunk
  constexpr int QV = 32 * KF / 4;     // float4 vectors per Qop chunk
  constexpr int AV = 32 * BN / 4;     // float4 vectors per A' chunk
  constexpr int NV = (QV + AV + 255) / 256;
  __shared__ __attribute__((aligned(16))) float sQ[32 * KF];
  __shared__ __attribute__((aligned(16))) float sA[32 * AS];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int l32 = lane & 31, h = lane >> 5;
  const int64_t rowbase = (int64_t)blockIdx.x * 128 + wave * 32;
  const int64_t c0 = (int64_t)blockIdx.y * BN;

  float pb[KSF];
#pragma unroll
  for (int ks = 0; ks < KSF; ks++) pb[ks] = P[(rowbase + l32) * KF + 2 * ks + h];

  f32x16_t acc[NT];
#pragma unroll
  for (int ct = 0; ct < NT; ct++)
#pragma unroll
    for (int q = 0; q < 16; q++) acc[ct][q] = 0.0f;

  f32x4_t stage[NV];
  auto gload = [&](int64_t t) {
    const f32x4_t* gq = (const f32x4_t*)(Q + t * 32 * KF);
#pragma unroll
    for (int v = 0; v < NV; v++) {
      int idx = tid + v * 256;
      // branch-free: both addresses are formed, one is selected (in-range index clamped)
      int a = idx - QV;
      a = a < 0 ? 0 : (a < AV ? a : AV - 1);
      int row = a / (BN / 4), col = (a % (BN / 4)) * 4;
      const f32x4_t* pa = (const f32x4_t*)(Af + (t * 32 + row) * r_pad + c0 + col);
      const f32x4_t* pq = gq + (idx < QV ? idx : 0);
      stage[v] = *(idx < QV ? pq : pa);
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int v = 0; v < NV; v++) {
      int idx = tid + v * 256;
      if (idx < QV) {
        *(f32x4_t*)&sQ[idx * 4] = stage[v];
      } else if (idx < QV + AV) {
        int a = idx - QV;
        int row = a / (BN / 4), col = (a % (BN / 4)) * 4;
        *(f32x4_t*)&sA[row * AS + col] = stage[v];
      }
    }
  };

  // j segment of this block (blockIdx.z): its FP32 sum goes to its own output slice
  const int64_t t0 = (int64_t)blockIdx.z * seg_chunks;
  const int64_t t1 = t0 + seg_chunks < n_chunks ? t0 + seg_chunks : n_chunks;
  W += (int64_t)blockIdx.z * seg_stride;
  gload(t0);
  for (int64_t t = t0; t < t1; t++) {
    sstore();
    __syncthreads();
    gload(t + 1 < t1 ? t + 1 : t);
    __builtin_amdgcn_sched_barrier(0);
    f32x16_t s;
#pragma unroll
    for (int q = 0; q < 16; q++) s[q] = 0.0f;
#pragma unroll
    for (int ks = 0; ks < KSF; ks++)
      s = __builtin_amdgcn_mfma_f32_32x32x2f32(sQ[l32 * KF + 2 * ks + h], pb[ks], s, 0, 0, 0);
    float e[16];
#pragma unroll
    for (int q = 0; q < 16; q++) e[q] = __builtin_amdgcn_exp2f(s[q]);
#pragma unroll
    for (int ct = 0; ct < NT; ct++) {
#pragma unroll
      for (int q = 0; q < 16; q++) {
        int j = 8 * (q / 4) + 4 * h + (q % 4);
        acc[ct] = __builtin_amdgcn_mfma_f32_32x32x2f32(e[q], sA[j * AS + ct * 32 + l32], acc[ct],
                                                        0, 0, 0);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("" ::: "memory");
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 16; q++) {
    int64_t i = rowbase + 8 * (q / 4) + 4 * h + (q % 4);
    if (i >= m) continue;
    float yi = y1[i];
#pragma unroll
    for (int ct = 0; ct < NT; ct++) {
      int64_t c = c0 + ct * 32 + l32;
      if (c < r) W[i * ldw + c] = yi * acc[ct][q];
    }
  }
}

// -------------------------------------------------------------------------------------
// FP64 kernel (TSVM_PREC_FP64): the distance, exp and product in double on
// v_mfma_f64_16x16x4f64, the FP64 reference of the low-precision and FP32 forms.
// Operands: P64 [m_pad][KF] = [x, 1, -g'|x|^2], Q64 [n_pad][KF] = [2 g' x, -g'|x|^2, 1]
// (KF = roundup(d + 2, 4)), A64 [n_pad][r_pad] = y_j A[j, c].  256 threads; wave w owns
// 32 rows (two 16-row tiles) and the block's 128 columns (8 tiles of 16): 64 FP64
// accumulators per tile pair.  Per 16-column chunk j: S^T (16 j x 16 i) = Q64_j P64_i^T
// (KF / 4 MFMAs per row tile; item q of lane l holds j = l / 16 + 4 q, i = l % 16), E =
// exp2(S^T) in FP64, and O (i x c) += E A64_j with K-step q taking j = 4 q + l / 16, the
// j that item q of the lane holds.  One LDS stage (register prefetch of the next).
// -------------------------------------------------------------------------------------
__global__ void prep_rows_f64(const float* __restrict__ X, int64_t m, int64_t m_pad, int d, int KF,
                              double gl2, double* __restrict__ P) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= m_pad) return;
  double* p = P + i * KF;
  if (i >= m) {
    for (int t = 0; t < KF; t++) p[t] = 0.0;
    return;
  }
  const float* x = X + i * d;
  double nrm = 0.0;
  for (int t = 0; t < d; t++) {
    nrm += (double)x[t] * x[t];
    p[t] = x[t];
  }
  p[d] = 1.0;
  p[d + 1] = -gl2 * nrm;
  for (int t = d + 2; t < KF; t++) p[t] = 0.0;
}
__global__ void prep_cols_f64(const float* __restrict__ X, int64_t n, int64_t n_pad, int d, int KF,
                              double gl2, double* __restrict__ Q) {
  int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= n_pad) return;
  double* q = Q + j * KF;
  if (j >= n) {
    for (int t = 0; t < KF; t++) q[t] = 0.0;
    return;
  }
  const float* x = X + j * d;
  double nrm = 0.0;
  for (int t = 0; t < d; t++) {
    nrm += (double)x[t] * x[t];
    q[t] = 2.0 * gl2 * (double)x[t];
  }
  q[d] = -gl2 * nrm;
  q[d + 1] = 1.0;
  for (int t = d + 2; t < KF; t++) q[t] = 0.0;
}
__global__ void prep_A_f64(const float* __restrict__ A, int64_t lda, const float* __restrict__ y,
                           int64_t n, int64_t n_pad, int64_t r, int64_t r_pad, double* __restrict__ Ad) {
  int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= n_pad * r_pad) return;
  int64_t j = idx / r_pad, c = idx - j * r_pad;
  Ad[idx] = (j < n && c < r) ? (double)y[j] * (double)A[j * lda + c] : 0.0;
}

template <int KF>
__global__ __launch_bounds__(256) void kprod_f64(const double* __restrict__ P,
                                                 const double* __restrict__ Q,
                                                 const double* __restrict__ Ad,
                                                 const float* __restrict__ y1,
                                                 float* __restrict__ W, int64_t ldw, int64_t m,
                                                 int64_t r, int64_t n_chunks, int64_t r_pad) {
  constexpr int NT = 8, BN = 16 * NT, KS = KF / 4;
  constexpr int QV = 16 * KF / 2;      // 16-byte vectors (two doubles) of a Q chunk
  constexpr int AV = 16 * BN / 2;      // of an A64 chunk slice
  constexpr int NV = (QV + AV + 255) / 256;
  __shared__ __attribute__((aligned(16))) double sQ[16 * KF];
  __shared__ __attribute__((aligned(16))) double sA[16 * BN];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const int64_t rowbase = (int64_t)blockIdx.x * 128 + wave * 32;
  const int64_t c0 = (int64_t)blockIdx.y * BN;
  // B operand of the distance MFMA: row i = l16 of each 16-row tile, K index 4 s + kq
  double pb[2][KS];
#pragma unroll
  for (int rt = 0; rt < 2; rt++)
#pragma unroll
    for (int ks = 0; ks < KS; ks++) pb[rt][ks] = P[(rowbase + 16 * rt + l16) * KF + 4 * ks + kq];
  f64x4_t acc[2][NT];
#pragma unroll
  for (int rt = 0; rt < 2; rt++)
#pragma unroll
    for (int ct = 0; ct < NT; ct++) acc[rt][ct] = (f64x4_t){0, 0, 0, 0};
  f64x2_t stage[NV];
  auto gload = [&](int64_t t) {
    const f64x2_t* gq = (const f64x2_t*)(Q + t * 16 * KF);
#pragma unroll
    for (int v = 0; v < NV; v++) {
      const int idx = tid + v * 256;
      int a = idx - QV;
      a = a < 0 ? 0 : (a < AV ? a : AV - 1);
      const int row = a / (BN / 2), col = (a % (BN / 2)) * 2;
      const f64x2_t* pa = (const f64x2_t*)(Ad + (t * 16 + row) * r_pad + c0 + col);
      const f64x2_t* pq = gq + (idx < QV ? idx : 0);
      stage[v] = *(idx < QV ? pq : pa);
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int v = 0; v < NV; v++) {
      const int idx = tid + v * 256;
      if (idx < QV) {
        *(f64x2_t*)&sQ[idx * 2] = stage[v];
      } else if (idx < QV + AV) {
        const int a = idx - QV;
        *(f64x2_t*)&sA[(a / (BN / 2)) * BN + (a % (BN / 2)) * 2] = stage[v];
      }
    }
  };
  gload(0);
  for (int64_t t = 0; t < n_chunks; t++) {
    sstore();
    __syncthreads();
    gload(t + 1 < n_chunks ? t + 1 : t);
    f64x4_t sd[2];
#pragma unroll
    for (int rt = 0; rt < 2; rt++) sd[rt] = (f64x4_t){0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      const double qa = sQ[l16 * KF + 4 * ks + kq];
#pragma unroll
      for (int rt = 0; rt < 2; rt++)
        sd[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(qa, pb[rt][ks], sd[rt], 0, 0, 0);
    }
    double e[2][4];
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
      for (int q = 0; q < 4; q++) e[rt][q] = exp2(sd[rt][q]);
#pragma unroll
    for (int ct = 0; ct < NT; ct++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const double bv = sA[(4 * q + kq) * BN + ct * 16 + l16];
#pragma unroll
        for (int rt = 0; rt < 2; rt++)
          acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(e[rt][q], bv, acc[rt][ct], 0, 0, 0);
      }
    __syncthreads();
  }
  // item q of lane l holds row l / 16 + 4 q, column l % 16 of each 16 x 16 tile
#pragma unroll
  for (int rt = 0; rt < 2; rt++)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int64_t i = rowbase + 16 * rt + kq + 4 * q;
      if (i >= m) continue;
      const double yi = y1[i];
#pragma unroll
      for (int ct = 0; ct < NT; ct++) {
        const int64_t c = c0 + ct * 16 + l16;
        if (c < r) W[i * ldw + c] = (float)(yi * acc[rt][ct][q]);
      }
    }
}

// -------------------------------------------------------------------------------------
// host launchers
// -------------------------------------------------------------------------------------
// Tiling per arithmetic (measured on the 1.6M x 1.6M x 512 sketch shape): the
// one-term forms run 128-column tiles at two waves per SIMD, the two-term forms
// 256-column tiles (one distance tile per 256 columns) with one chunk per stage.
template <int NT, int KD>
static void launch_bf16(int em, const __bf16* P, const __bf16* Q, const void* AT,
                        const float* y1, float* W, int64_t ldw, int64_t m, int64_t r,
                        int64_t n_chunks, int64_t r_pad, int64_t m_pad, const float* sc,
                        int cu_count, hipStream_t s) {
  constexpr int CH = KBF16_CH;
  constexpr int OCC = NT <= 4 ? 2 : 1;
  const int64_t nxb = m_pad / 256, ntiles = nxb * (r_pad / (NT * 32));
  const int64_t resident = (int64_t)cu_count * OCC;
  dim3 grid((unsigned)(ntiles < resident ? ntiles : resident));
  switch (em) {
    case EM_BF16:
      hipLaunchKernelGGL((kprod_bf16<NT, KD, CH, EM_BF16, OCC>), grid, dim3(256), 0, s, P, Q, AT,
                         y1, W, ldw, m, r, n_chunks / CH, r_pad, sc, nxb, ntiles);
      break;
    case EM_F16:
      if constexpr (NT == 4 && KD <= KBF16_WIDE_MAX_KD) {
        // one 512-thread block per CU: eight waves share each staged stage
        const int64_t nxb2 = m_pad / 512, ntiles2 = nxb2 * (r_pad / (NT * 32));
        dim3 grid2((unsigned)(ntiles2 < cu_count ? ntiles2 : cu_count));
        hipLaunchKernelGGL((kprod_bf16<NT, KD, CH, EM_F16, 1, 512>), grid2, dim3(512), 0, s, P,
                           Q, AT, y1, W, ldw, m, r, n_chunks / CH, r_pad, sc, nxb2, ntiles2);
      } else {
        hipLaunchKernelGGL((kprod_bf16<NT, KD, CH, EM_F16, OCC>), grid, dim3(256), 0, s, P, Q,
                           AT, y1, W, ldw, m, r, n_chunks / CH, r_pad, sc, nxb, ntiles);
      }
      break;
    case EM_F16X2:
      hipLaunchKernelGGL((kprod_bf16<NT, KD, 1, EM_F16X2, 1>), grid, dim3(256), 0, s, P, Q, AT,
                         y1, W, ldw, m, r, n_chunks, r_pad, sc, nxb, ntiles);
      break;
    default:  // EM_F16R2 (its AT rows are twice as wide)
      hipLaunchKernelGGL((kprod_bf16<NT, KD, 1, EM_F16R2, 1>), grid, dim3(256), 0, s, P, Q, AT,
                         y1, W, ldw, m, r, n_chunks, r_pad, sc, nxb, ntiles);
      break;
  }
}
template <int NT, int KF>
static void launch_f32(const float* P, const float* Q, const float* Af, const float* y1,
                       float* W, int64_t ldw, int64_t m, int64_t r, int64_t n_chunks,
                       int64_t r_pad, int64_t m_pad, int64_t nseg, int64_t seg_chunks,
                       int64_t seg_stride, hipStream_t s) {
  dim3 grid((unsigned)(m_pad / 128), (unsigned)(r_pad / (NT * 32)), (unsigned)nseg);
  hipLaunchKernelGGL((kprod_f32<NT, KF>), grid, dim3(256), 0, s, P, Q, Af, y1, W, ldw, m, r,
                     n_chunks, r_pad, seg_chunks, seg_stride);
}

// W[i, c] = FP32(sum over segments of part[s][i][c], in FP64 and segment order)
__global__ void seg_reduce_kernel(const float* __restrict__ part, int64_t nseg, int64_t m,
                                  int64_t r, float* __restrict__ W, int64_t ldw) {
  const int64_t tot = m * r;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int64_t t = 0; t < nseg; t++) acc += (double)part[t * tot + e];
    const int64_t i = e / r, c = e - i * r;
    W[i * ldw + c] = (float)acc;
  }
}

// Where the operand buffers of one product come from: the context's workspace (grown on
// demand), or a fixed device region carved from its start (the allocation-free entry
// point).  Every buffer of a call is taken before its first launch.
struct KpBuffers {
  unsigned char* base = nullptr;  // fixed region, or nullptr for the workspace
  size_t cap = 0, used = 0;
  tsvm_status get(Ctx* c, const char* name, size_t bytes, void** out) {
    if (!base) {
      TSVM_HIP(c, c->ws.get(name, bytes, out));
      return TSVM_OK;
    }
    const size_t off = (used + 255) / 256 * 256;
    if (off + bytes > cap)
      return fail(c, TSVM_ERR_ARG,
                  "kernel_product_noalloc: operands need more than the %zu MiB of scratch "
                  "(%zu bytes requested so far)", cap >> 20, off + bytes);
    *out = base + off;
    used = off + bytes;
    return TSVM_OK;
  }
};

// Widest operand (KD) of the wide-tile kernel
#define KPW_MAX_KD 128

__global__ void zero_u64_kernel(unsigned long long* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

// One cooperative launch of cu_count blocks on the XCD-round schedule (sched: device words,
// KP_SCHED_WORDS of them) when there are at least two rounds of tiles, cu_count splits into
// the eight XCDs and the stream is not being captured; otherwise (or if the cooperative
// launch is refused) one launch per round of cu_count tiles (one block per CU).  *dyn_used
// tells which ran (the caller then waits for the product and checks the schedule's flag).
template <int NT, int KD, int WAVES, int RB>
static hipError_t launch_wide(int em, const __bf16* P, const __bf16* Qf, const void* Af,
                              const float* y1, float* W, int64_t ldw, int64_t m, int64_t r,
                              int64_t n_chunks, int64_t r_tiles, int64_t nxb, const float* sc,
                              int cu_count, unsigned long long* sched, hipStream_t s,
                              bool* dyn_used) {
  const int64_t ntiles = nxb * (r_tiles / NT);
  *dyn_used = false;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cap) != hipSuccess) cap = hipStreamCaptureStatusActive;
  if (sched && cap == hipStreamCaptureStatusNone && cu_count % KP_XCDS == 0 &&
      ntiles >= 2 * (int64_t)cu_count) {
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, sched, KP_SCHED_WORDS);
    int64_t t0 = 0;
    unsigned long long* sp = sched;
    void* args[] = {(void*)&P, (void*)&Qf, (void*)&Af, (void*)&y1, (void*)&W, (void*)&ldw,
                    (void*)&m, (void*)&r, (void*)&n_chunks, (void*)&r_tiles, (void*)&sc,
                    (void*)&nxb, (void*)&ntiles, (void*)&t0, (void*)&sp};
    const void* fn = em == EM_BF16 ? (const void*)&kprod_wide<NT, KD, EM_BF16, WAVES, RB>
                                   : (const void*)&kprod_wide<NT, KD, EM_F16, WAVES, RB>;
    const hipError_t e = hipLaunchCooperativeKernel(fn, dim3((unsigned)cu_count),
                                                    dim3(64 * WAVES), args, 0, s);
    if (e == hipSuccess) {
      *dyn_used = true;
      return hipSuccess;
    }
    (void)hipGetLastError();  // refused: fall back to one launch per round
  }
  unsigned long long* none = nullptr;
  for (int64_t t0 = 0; t0 < ntiles; t0 += cu_count) {
    const int64_t t1 = t0 + cu_count < ntiles ? t0 + cu_count : ntiles;
    const dim3 grid((unsigned)(t1 - t0));
    if (em == EM_BF16)
      hipLaunchKernelGGL((kprod_wide<NT, KD, EM_BF16, WAVES, RB>), grid, dim3(64 * WAVES), 0, s, P, Qf,
                         Af, y1, W, ldw, m, r, n_chunks, r_tiles, sc, nxb, t1, t0, none);
    else
      hipLaunchKernelGGL((kprod_wide<NT, KD, EM_F16, WAVES, RB>), grid, dim3(64 * WAVES), 0, s, P, Qf,
                         Af, y1, W, ldw, m, r, n_chunks, r_tiles, sc, nxb, t1, t0, none);
  }
  return hipGetLastError();
}

// The one-term low-precision forms with more than 128 right-hand-side columns.
static tsvm_status product_wide(Ctx* c, KpBuffers& bufs, int em, int KD, const float* X1,
                                const float* y1, int64_t m, const float* X2, const float* y2,
                                int64_t n, int64_t d, const float* A, int64_t lda, int64_t r,
                                double gl2, float* W, int64_t ldw) {
  hipStream_t s = c->stream;
  constexpr int NT = 8, RB = 2, WAVES = 4, BR = 32 * RB * WAVES;
  const int64_t n_pad = round_up(n > 0 ? n : 1, 32), n_chunks = n_pad / 32;
  const int64_t m_pad = round_up(m, BR), r_pad = round_up(r, 32 * NT);
  const int64_t r_tiles = r_pad / 32, nxb = m_pad / BR;
  void *P, *Qf, *Af, *scv = nullptr, *schedv = nullptr;
  TSVM_CHECK(bufs.get(c, "kp_sched", KP_SCHED_WORDS * 8, &schedv));
  TSVM_CHECK(bufs.get(c, "kp_P", m_pad * KD * 2, &P));
  TSVM_CHECK(bufs.get(c, "kp_Q", n_pad * KD * 2, &Qf));
  TSVM_CHECK(bufs.get(c, "kp_A", n_pad * r_pad * 2, &Af));
  if (em != EM_BF16) TSVM_CHECK(bufs.get(c, "kp_sc", 16, &scv));
  hipLaunchKernelGGL(prep_rows_bf16, dim3((unsigned)ceil_div(m_pad, 256)), dim3(256), 0, s, X1, m,
                     m_pad, (int)d, KD, gl2, (__bf16*)P);
  hipLaunchKernelGGL(prep_cols_bf16_frag, dim3((unsigned)ceil_div(n_pad, 256)), dim3(256), 0, s,
                     X2, n, n_pad, (int)d, KD, gl2, (__bf16*)Qf);
  float* sc = nullptr;
  const dim3 ga((unsigned)n_chunks, (unsigned)r_tiles);
  if (em != EM_BF16) {
    // one power-of-two scale brings max |y_j A[j,c]| to [2^13, 2^14) in fp16
    sc = (float*)scv;
    unsigned int* amax = (unsigned int*)(sc + 2);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, amax);
    const int64_t tot = (n > 0 ? n : 1) * r;
    const unsigned nb = (unsigned)std::min<int64_t>(ceil_div(tot, 256), 4096);
    hipLaunchKernelGGL(amax_kernel, dim3(nb), dim3(256), 0, s, A, lda, n, r, amax);
    hipLaunchKernelGGL(fp16_scale_kernel, dim3(1), dim3(1), 0, s, amax, sc);
    hipLaunchKernelGGL(prep_A_frag<_Float16>, ga, dim3(256), 0, s, A, lda, y2, n, r, r_tiles,
                       (const float*)sc, (_Float16*)Af);
  } else {
    hipLaunchKernelGGL(prep_A_frag<__bf16>, ga, dim3(256), 0, s, A, lda, y2, n, r, r_tiles,
                       (const float*)nullptr, (__bf16*)Af);
  }
  TSVM_HIP(c, hipGetLastError());
  const __bf16 *Pb = (const __bf16*)P, *Qb = (const __bf16*)Qf;
  unsigned long long* sched = (unsigned long long*)schedv;
  bool dyn = false;
  hipError_t le = hipSuccess;
#define KPW_CASE(kd)                                                                            \
  case kd:                                                                                      \
    le = launch_wide<NT, kd, WAVES, RB>(em, Pb, Qb, Af, y1, W, ldw, m, r, n_chunks, r_tiles,    \
                                        nxb, sc, c->cu_count, sched, s, &dyn);                  \
    break;
  switch (KD) {
    KPW_CASE(16)
    KPW_CASE(32)
    KPW_CASE(48)
    KPW_CASE(64)
    KPW_CASE(80)
    KPW_CASE(96)
    KPW_CASE(112)
    KPW_CASE(128)
    default:
      return fail(c, TSVM_ERR_ARG, "kernel_product(wide): KD=%d not compiled", KD);
  }
#undef KPW_CASE
  TSVM_HIP(c, le);
  if (dyn) {
    // the schedule's error flag (a bounded wait that ran out)
    unsigned long long flag = 0;
    TSVM_HIP(c, hipMemcpyAsync(&flag, sched + KP_SCHED_WORDS - 1, 8, hipMemcpyDeviceToHost, s));
    TSVM_HIP(c, hipStreamSynchronize(s));
    if (flag != 0) return fail(c, TSVM_ERR_HIP, "kernel_product(wide): tile schedule wait timed out");
  }
  return TSVM_OK;
}

static tsvm_status kernel_product_impl(Ctx* c, KpBuffers& bufs, const float* X1, const float* y1,
                                       int64_t m, const float* X2, const float* y2, int64_t n,
                                       int64_t d, const float* A, int64_t lda, int64_t r,
                                       double gamma, int precision, float* W, int64_t ldw) {
  if (m <= 0 || r <= 0) return TSVM_OK;
  if (n < 0 || d <= 0 || lda < r || ldw < r || !(gamma >= 0.0))
    return fail(c, TSVM_ERR_ARG, "kernel_product: bad shape m=%ld n=%ld d=%ld r=%ld", (long)m,
                (long)n, (long)d, (long)r);
  hipStream_t s = c->stream;
  const double gl2 = gamma * TSVM_LOG2E;
  if (precision == TSVM_PREC_BF16 || precision == TSVM_PREC_FP16 ||
      precision == TSVM_PREC_FP16X2 || precision == TSVM_PREC_FP16R2) {
    const int em = precision == TSVM_PREC_BF16     ? EM_BF16
                   : precision == TSVM_PREC_FP16   ? EM_F16
                   : precision == TSVM_PREC_FP16X2 ? EM_F16X2
                                                   : EM_F16R2;
    const int64_t aw = em == EM_F16R2 ? 64 : 32;
    const int64_t n_pad = round_up(n > 0 ? n : 1, 32 * KBF16_CH), n_chunks = n_pad / 32;
    // operand width: the smallest compiled size holding [hi, hi, lo, norm terms]
    static const int kds[] = {16, 32, 48, 64, 80, 96, 112, 128, 176, 256};
    int KD = 0;
    for (int v : kds)
      if (v >= 3 * d + 4) { KD = v; break; }
    if (KD == 0)
      return fail(c, TSVM_ERR_ARG, "kernel_product(bf16): d=%ld exceeds 84 features", (long)d);
    const bool two_term = em == EM_F16X2 || em == EM_F16R2;
    if (!two_term && r > 128 && KD <= KPW_MAX_KD)
      return product_wide(c, bufs, em, KD, X1, y1, m, X2, y2, n, d, A, lda, r, gl2, W, ldw);
    const int NT = (r <= 32) ? 1 : (r <= 64 ? 2 : (r <= 128 || !two_term ? 4 : 8));
    const int64_t BN = NT * 32, m_pad = round_up(
                      m, em == EM_F16 && NT == 4 && KD <= KBF16_WIDE_MAX_KD ? 512 : 256),
                  r_pad = round_up(r, BN);
    void *P, *Q, *AT, *scv = nullptr;
    TSVM_CHECK(bufs.get(c, "kp_P", m_pad * KD * 2, &P));
    TSVM_CHECK(bufs.get(c, "kp_Q", n_pad * KD * 2, &Q));
    TSVM_CHECK(bufs.get(c, "kp_A", n_pad * r_pad * aw * 2 / 32, &AT));
    if (em != EM_BF16) TSVM_CHECK(bufs.get(c, "kp_sc", 16, &scv));
    hipLaunchKernelGGL(prep_rows_bf16, dim3((unsigned)ceil_div(m_pad, 256)), dim3(256), 0, s, X1, m,
                       m_pad, (int)d, KD, gl2, (__bf16*)P);
    hipLaunchKernelGGL(prep_cols_bf16, dim3((unsigned)ceil_div(n_pad, 256)), dim3(256), 0, s, X2, n,
                       n_pad, (int)d, KD, gl2, (__bf16*)Q);
    float* sc = nullptr;
    if (em != EM_BF16) {
      // one power-of-two scale brings max |y_j A[j,c]| to [2^13, 2^14) in fp16
      sc = (float*)scv;
      unsigned int* amax = (unsigned int*)(sc + 2);
      hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, amax);
      const int64_t tot = (n > 0 ? n : 1) * r;
      const unsigned nb = (unsigned)std::min<int64_t>(ceil_div(tot, 256), 4096);
      hipLaunchKernelGGL(amax_kernel, dim3(nb), dim3(256), 0, s, A, lda, n, r, amax);
      hipLaunchKernelGGL(fp16_scale_kernel, dim3(1), dim3(1), 0, s, amax, sc);
      if (em == EM_F16R2)
        hipLaunchKernelGGL(prep_A_f16<true>, dim3((unsigned)n_chunks, (unsigned)(r_pad / 32)),
                           dim3(256), 0, s, A, lda, y2, n, r, r_pad, sc, (_Float16*)AT);
      else
        hipLaunchKernelGGL(prep_A_f16<false>, dim3((unsigned)n_chunks, (unsigned)(r_pad / 32)),
                           dim3(256), 0, s, A, lda, y2, n, r, r_pad, sc, (_Float16*)AT);
    } else {
      hipLaunchKernelGGL(prep_A_bf16, dim3((unsigned)n_chunks, (unsigned)(r_pad / 32)), dim3(256),
                         0, s, A, lda, y2, n, r, r_pad, (__bf16*)AT);
    }
    TSVM_HIP(c, hipGetLastError());
    const __bf16 *Pb = (const __bf16*)P, *Qb = (const __bf16*)Q;
    const void* Ab = AT;
#define KP_BF16_CASE(kd)                                                                        \
  case kd:                                                                                      \
    if (NT == 1)                                                                                \
      launch_bf16<1, kd>(em, Pb, Qb, Ab, y1, W, ldw, m, r, n_chunks, r_pad, m_pad, sc, c->cu_count, s);   \
    else if (NT == 2)                                                                           \
      launch_bf16<2, kd>(em, Pb, Qb, Ab, y1, W, ldw, m, r, n_chunks, r_pad, m_pad, sc, c->cu_count, s);   \
    else if (NT == 4)                                                                           \
      launch_bf16<4, kd>(em, Pb, Qb, Ab, y1, W, ldw, m, r, n_chunks, r_pad, m_pad, sc, c->cu_count, s);   \
    else                                                                                        \
      launch_bf16<8, kd>(em, Pb, Qb, Ab, y1, W, ldw, m, r, n_chunks, r_pad, m_pad, sc, c->cu_count, s);   \
    break;
    switch (KD) {
      KP_BF16_CASE(16)
      KP_BF16_CASE(32)
      KP_BF16_CASE(48)
      KP_BF16_CASE(64)
      KP_BF16_CASE(80)
      KP_BF16_CASE(96)
      KP_BF16_CASE(112)
      KP_BF16_CASE(128)
      KP_BF16_CASE(176)
      KP_BF16_CASE(256)
      default:
        return fail(c, TSVM_ERR_ARG, "kernel_product(bf16): KD=%d not compiled", KD);
    }
#undef KP_BF16_CASE
  } else if (precision == TSVM_PREC_FP64) {
    const int64_t n_pad = round_up(n > 0 ? n : 1, 16), n_chunks = n_pad / 16;
    // operand width: the smallest compiled size holding [x, norm terms] (zero padded)
    static const int kfs[] = {4, 8, 12, 16, 20, 24, 28, 32, 36, 40, 48, 56, 64, 80, 96, 112, 132};
    int KF = 0;
    for (int v : kfs)
      if (v >= d + 2) { KF = v; break; }
    if (KF == 0)
      return fail(c, TSVM_ERR_ARG, "kernel_product(fp64): d=%ld exceeds 130 features", (long)d);
    const int64_t m_pad = round_up(m, 128), r_pad = round_up(r, 128);
    void *P, *Q, *Ad;
    TSVM_CHECK(bufs.get(c, "kp_P", m_pad * KF * 8, &P));
    TSVM_CHECK(bufs.get(c, "kp_Q", n_pad * KF * 8, &Q));
    TSVM_CHECK(bufs.get(c, "kp_A", n_pad * r_pad * 8, &Ad));
    hipLaunchKernelGGL(prep_rows_f64, dim3((unsigned)ceil_div(m_pad, 256)), dim3(256), 0, s, X1, m,
                       m_pad, (int)d, KF, gl2, (double*)P);
    hipLaunchKernelGGL(prep_cols_f64, dim3((unsigned)ceil_div(n_pad, 256)), dim3(256), 0, s, X2, n,
                       n_pad, (int)d, KF, gl2, (double*)Q);
    hipLaunchKernelGGL(prep_A_f64, dim3((unsigned)ceil_div(n_pad * r_pad, 256)), dim3(256), 0, s,
                       A, lda, y2, n, n_pad, r, r_pad, (double*)Ad);
    TSVM_HIP(c, hipGetLastError());
    const dim3 grid((unsigned)(m_pad / 128), (unsigned)(r_pad / 128));
#define KP_F64_CASE(kf)                                                                          \
  case kf:                                                                                       \
    hipLaunchKernelGGL((kprod_f64<kf>), grid, dim3(256), 0, s, (const double*)P, (const double*)Q, \
                       (const double*)Ad, y1, W, ldw, m, r, n_chunks, r_pad);                     \
    break;
    switch (KF) {
      KP_F64_CASE(4) KP_F64_CASE(8) KP_F64_CASE(12) KP_F64_CASE(16) KP_F64_CASE(20)
      KP_F64_CASE(24) KP_F64_CASE(28) KP_F64_CASE(32) KP_F64_CASE(36) KP_F64_CASE(40)
      KP_F64_CASE(48) KP_F64_CASE(56) KP_F64_CASE(64) KP_F64_CASE(80) KP_F64_CASE(96)
      KP_F64_CASE(112) KP_F64_CASE(132)
      default:
        return fail(c, TSVM_ERR_ARG, "kernel_product(fp64): KF=%d not compiled", KF);
    }
#undef KP_F64_CASE
  } else if (precision == TSVM_PREC_FP32) {
    const int64_t n_pad = round_up(n > 0 ? n : 1, 32), n_chunks = n_pad / 32;
    static const int kfs[] = {4, 8, 12, 16, 20, 24, 34, 56, 64, 96, 130};
    int KF = 0;
    for (int v : kfs)
      if (v >= d + 2) { KF = v; break; }
    if (KF == 0)
      return fail(c, TSVM_ERR_ARG, "kernel_product(fp32): d=%ld exceeds 128 features", (long)d);
    const int NT = (r <= 32) ? 1 : (r <= 64 ? 2 : (r <= 128 ? 4 : 8));
    const int64_t BN = NT * 32, m_pad = round_up(m, 128), r_pad = round_up(r, BN);
    // FP32 sums over segments of 8192 columns (the reference's tile, whose products it
    // accumulates with beta = 1, 2537-2555), added in FP64 in segment order, bounded by
    // 8 GiB of partial sums: at most max(m, n) r nseg floats.  Where that bound binds
    // (max(m, n) r > 2^28: 1.6M x 512 takes two segments of 800k columns) each FP32 sum
    // spans more columns; the reference accumulates all n columns of a product in FP32.
    // The bound takes the larger of m and n so that a product computed in row slabs (m < n)
    // is segmented, and rounded, exactly as the whole one.
    const int64_t seg_len = 256;  // chunks of 32 columns
    int64_t nseg = ceil_div(n_chunks, seg_len);
    const double seg_rows = (double)(m > n ? m : n);
    const int64_t max_seg = (int64_t)(8.0 * (1ll << 30) / (seg_rows * r * 4.0));
    if (nseg > max_seg) nseg = max_seg > 1 ? max_seg : 1;
    const int64_t seg_chunks = ceil_div(n_chunks, nseg);
    nseg = ceil_div(n_chunks, seg_chunks);
    void *P, *Q, *Af, *pp = nullptr;
    TSVM_CHECK(bufs.get(c, "kp_P", m_pad * KF * 4, &P));
    TSVM_CHECK(bufs.get(c, "kp_Q", n_pad * KF * 4, &Q));
    TSVM_CHECK(bufs.get(c, "kp_A", n_pad * r_pad * 4, &Af));
    if (nseg > 1) TSVM_CHECK(bufs.get(c, "kp_seg", (size_t)(nseg * m * r * 4), &pp));
    hipLaunchKernelGGL(prep_rows_f32, dim3((unsigned)ceil_div(m_pad, 256)), dim3(256), 0, s, X1, m,
                       m_pad, (int)d, KF, gl2, (float*)P);
    hipLaunchKernelGGL(prep_cols_f32, dim3((unsigned)ceil_div(n_pad, 256)), dim3(256), 0, s, X2, n,
                       n_pad, (int)d, KF, gl2, (float*)Q);
    hipLaunchKernelGGL(prep_A_f32, dim3((unsigned)ceil_div(n_pad * r_pad, 256)), dim3(256), 0, s,
                       A, lda, y2, n, n_pad, r, r_pad, (float*)Af);
    TSVM_HIP(c, hipGetLastError());
    float* Wout = W;
    int64_t ldo = ldw, seg_stride = 0;
    if (nseg > 1) {
      Wout = (float*)pp;
      ldo = r;
      seg_stride = m * r;
    }
    const float *Pf = (const float*)P, *Qf = (const float*)Q, *Aff = (const float*)Af;
#define KP_F32_ARGS Pf, Qf, Aff, y1, Wout, ldo, m, r, n_chunks, r_pad, m_pad, nseg, seg_chunks, seg_stride, s
#define KP_F32_CASE(kf)                                   \
  case kf:                                                \
    if (NT == 1) launch_f32<1, kf>(KP_F32_ARGS);          \
    else if (NT == 2) launch_f32<2, kf>(KP_F32_ARGS);     \
    else if (NT == 4) launch_f32<4, kf>(KP_F32_ARGS);     \
    else launch_f32<8, kf>(KP_F32_ARGS);                  \
    break;
    switch (KF) {
      KP_F32_CASE(4)
      KP_F32_CASE(8)
      KP_F32_CASE(12)
      KP_F32_CASE(16)
      KP_F32_CASE(20)
      KP_F32_CASE(24)
      KP_F32_CASE(34)
      KP_F32_CASE(56)
      KP_F32_CASE(64)
      KP_F32_CASE(96)
      KP_F32_CASE(130)
      default:
        return fail(c, TSVM_ERR_ARG, "kernel_product(fp32): KF=%d not compiled", KF);
    }
#undef KP_F32_CASE
#undef KP_F32_ARGS
    if (nseg > 1)
      hipLaunchKernelGGL(seg_reduce_kernel, dim3(4096), dim3(256), 0, s, (const float*)Wout, nseg, m,
                         r, W, ldw);
  } else {
    return fail(c, TSVM_ERR_ARG, "kernel_product: unknown precision %d", precision);
  }
  TSVM_HIP(c, hipGetLastError());
  return TSVM_OK;
}

tsvm_status kernel_product_dev(Ctx* c, const float* X1, const float* y1, int64_t m,
                               const float* X2, const float* y2, int64_t n, int64_t d,
                               const float* A, int64_t lda, int64_t r, double gamma,
                               int precision, float* W, int64_t ldw) {
  KpBuffers ws;
  return kernel_product_impl(c, ws, X1, y1, m, X2, y2, n, d, A, lda, r, gamma, precision, W,
                             ldw);
}

// ---- allocation-free kernel product ------------------------------------------------
// The operand buffers of tsvm_kernel_product_noalloc_dev live in this file-scope device
// array: one copy per GPU, held until the process ends.  Host code never passes its name
// to a kernel (on the host it is a stand-in, not a device address); the address of each
// device's copy is looked up once, with that device current, and kept in g_scratch_dev.
static __device__ __align__(256) unsigned char scratch[TSVM_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};
}  // namespace

tsvm_status noalloc_scratch_setup(Ctx* c) {
  if (c->device < 0 || c->device >= kMaxDevices)
    return fail(c, TSVM_ERR_ARG, "noalloc scratch: device %d beyond %d", c->device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[c->device]) {
    void* ptr = nullptr;
    TSVM_HIP(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(c, TSVM_ERR_HIP, "noalloc scratch: no device address");
    g_scratch_dev[c->device] = ptr;
  }
  c->scratch = g_scratch_dev[c->device];
  return TSVM_OK;
}

tsvm_status kernel_product_noalloc_dev(Ctx* c, const float* X1, const float* y1, int64_t m,
                                       const float* X2, const float* y2, int64_t n, int64_t d,
                                       const float* A, int64_t r, double gamma, int precision,
                                       float* W) {
  if (!c->scratch) return fail(c, TSVM_ERR_HIP, "kernel_product_noalloc: scratch not set up");
  KpBuffers fixed;
  fixed.base = (unsigned char*)c->scratch;
  fixed.cap = TSVM_NOALLOC_SCRATCH_BYTES;
  return kernel_product_impl(c, fixed, X1, y1, m, X2, y2, n, d, A, r, r, gamma, precision, W, r);
}

}  // namespace tsvm
