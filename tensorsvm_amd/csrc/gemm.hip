// Tall-skinny dense GEMMs of the factor construction, on FP64 MFMA
// (v_mfma_f64_16x16x4f64, FP64 accumulation of FP32 operands):
//
//   gemm_tn_f64:  C (ka x kb) = A^T diag(w) B over n rows   (Gram W^T W, C = W^T Q,
//                 the plain-form capacitance Z^T D^-1 Z)   -- replaces the chunked
//                 cublasDgemm of tensorsvm-train.cu:2122-2149 and cblas_dsyrk/cublasDsyrk
//                 of 1674-1697
//   gemm_nn_tall: Y (n x kb, FP32) = A (n x ka, FP32) * B (ka x kb, FP64)   (W R^-1,
//                 U = W X) -- replaces cublasSgemm of 799-806 and 2231-2233
//
// gemm_tn_f64 splits the n rows over a third grid dimension; each split writes its
// partial tile and a second kernel sums the splits in a fixed order (deterministic).
#include "common.h"
#include "internal.h"

namespace tsvm {

// 256 threads; the block owns a 64 x 64 output tile, wave w a 32 x 32 quarter
// (2 x 2 MFMA tiles of 16 x 16).  Rows are staged 32 at a time through LDS as FP64
// (the A panel already multiplied by the row weight), double-buffered: the next stage's
// global loads are in flight while the MFMAs consume the current one.
constexpr int TN_R = 32;  // rows per stage

__global__ __launch_bounds__(256) void gemm_tn_f64_kernel(
    const float* __restrict__ A, int64_t lda, int64_t ka, const float* __restrict__ B,
    int64_t ldb, int64_t kb, int64_t n, const double* __restrict__ w, int lower_only,
    int64_t rows_per_split, const int* __restrict__ rows, double* __restrict__ part) {
  const int ta = blockIdx.x, tb = blockIdx.y;
  if (lower_only && tb > ta) return;
  __shared__ double sA[2][TN_R][64];
  __shared__ double sB[2][TN_R][64];
  const int64_t split = blockIdx.z;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const int wa = (wave >> 1) * 32, wb = (wave & 1) * 32;
  const int64_t a0 = (int64_t)ta * 64, b0 = (int64_t)tb * 64;
  const int64_t r0 = split * rows_per_split;
  const int64_t r1 = r0 + rows_per_split < n ? r0 + rows_per_split : n;
  const int64_t nst = (r1 - r0 + TN_R - 1) / TN_R;
  // staging map: thread -> (row, 4 columns) for two passes of 16 rows
  const int sr = tid >> 4, sc4 = (tid & 15) * 4;
  const bool avec = (lda % 4 == 0) && (a0 + 64 <= ka);
  const bool bvec = (ldb % 4 == 0) && (b0 + 64 <= kb);
  f32x4_t ra[2], rb[2];
  double rw[2];
  auto gload = [&](int64_t st) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      int64_t row = r0 + st * TN_R + sr + 16 * h;
      bool ok = row < r1;
      int64_t rr = ok ? row : r0;
      if (rows) rr = rows[rr];
      rw[h] = ok ? (w ? w[rr] : 1.0) : 0.0;
      const float* pa = A + rr * lda + a0 + sc4;
      const float* pb = B + rr * ldb + b0 + sc4;
      if (avec) {
        ra[h] = *(const f32x4_t*)pa;
      } else {
#pragma unroll
        for (int j = 0; j < 4; j++) ra[h][j] = (a0 + sc4 + j < ka) ? pa[j] : 0.0f;
      }
      if (bvec) {
        rb[h] = *(const f32x4_t*)pb;
      } else {
#pragma unroll
        for (int j = 0; j < 4; j++) rb[h][j] = (b0 + sc4 + j < kb) ? pb[j] : 0.0f;
      }
      if (!ok) rb[h] = (f32x4_t){0.f, 0.f, 0.f, 0.f};
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        sA[buf][sr + 16 * h][sc4 + j] = rw[h] * (double)ra[h][j];
        sB[buf][sr + 16 * h][sc4 + j] = (double)rb[h][j];
      }
  };
  f64x4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = (f64x4_t){0, 0, 0, 0};
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
    for (int s4 = 0; s4 < TN_R / 4; s4++) {
      const int row = 4 * s4 + kq;
      double av0 = sA[buf][row][wa + l16], av1 = sA[buf][row][wa + 16 + l16];
      double bv0 = sB[buf][row][wb + l16], bv1 = sB[buf][row][wb + 16 + l16];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(av0, bv0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(av0, bv1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(av1, bv0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(av1, bv1, acc[1][1], 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("" ::: "memory");
    sstore(buf ^ 1);
    __syncthreads();
  }
  // D layout of v_mfma_f64_16x16x4f64: item q of a lane holds row lane/16 + 4q,
  // column lane%16
  double* P = part + split * ka * kb;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        int64_t a = a0 + wa + 16 * i + kq + 4 * q, b = b0 + wb + 16 * j + l16;
        if (a < ka && b < kb) P[a * kb + b] = acc[i][j][q];
      }
}

__global__ void reduce_splits_kernel(const double* __restrict__ part, int64_t nsplit,
                                     int64_t ka, int64_t kb, int lower_only,
                                     double* __restrict__ C) {
  int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= ka * kb) return;
  int64_t a = idx / kb, b = idx - a * kb;
  // lower_only: the lower triangle is summed and mirrored entry by entry, so C is
  // exactly symmetric (inside a diagonal tile (w z_a) z_b and (w z_b) z_a round apart)
  if (lower_only && b > a) return;
  double s = 0.0;
  for (int64_t t = 0; t < nsplit; t++) s += part[t * ka * kb + idx];
  C[idx] = s;
  if (lower_only && b < a) C[b * kb + a] = s;
}

tsvm_status gemm_tn_f64(Ctx* c, const float* A, int64_t lda, int64_t ka, const float* B,
                        int64_t ldb, int64_t kb, int64_t n, const double* w, bool lower_only,
                        double* C, const int* rows) {
  if (ka <= 0 || kb <= 0) return TSVM_OK;
  if (lower_only && ka != kb) return fail(c, TSVM_ERR_ARG, "gemm_tn_f64: lower_only needs ka==kb");
  const int64_t ta = ceil_div(ka, 64), tb = ceil_div(kb, 64);
  const int64_t tiles = lower_only ? ta * (ta + 1) / 2 : ta * tb;
  // enough blocks to fill the chip several times; at least 256 rows per split
  int64_t nsplit = ceil_div((int64_t)c->cu_count * 8, tiles);
  int64_t max_split = ceil_div(n > 0 ? n : 1, 256);
  if (nsplit > max_split) nsplit = max_split;
  if (nsplit < 1) nsplit = 1;
  int64_t rps = round_up(ceil_div(n > 0 ? n : 1, nsplit), TN_R);
  nsplit = ceil_div(n > 0 ? n : 1, rps);
  void* part;
  TSVM_HIP(c, c->ws.get("gemm_part", (size_t)(nsplit * ka * kb * 8), &part));
  dim3 grid((unsigned)ta, (unsigned)tb, (unsigned)nsplit);
  hipLaunchKernelGGL(gemm_tn_f64_kernel, grid, dim3(256), 0, c->stream, A, lda, ka, B, ldb, kb,
                     n, w, lower_only ? 1 : 0, rps, rows, (double*)part);
  hipLaunchKernelGGL(reduce_splits_kernel, dim3((unsigned)ceil_div(ka * kb, 256)), dim3(256), 0,
                     c->stream, (const double*)part, nsplit, ka, kb, lower_only ? 1 : 0, C);
  TSVM_HIP(c, hipGetLastError());
  return TSVM_OK;
}

// 256 threads; block owns 64 rows x 64 columns of Y, wave w rows [16w, 16w+16) x 64 columns
// (4 MFMA tiles).  K (= ka) is walked 4 at a time.
__global__ __launch_bounds__(256) void gemm_nn_tall_kernel(const float* __restrict__ A,
                                                           int64_t lda, int64_t ka,
                                                           const double* __restrict__ B,
                                                           int64_t kb, int64_t n,
                                                           float* __restrict__ Y, int64_t ldy) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const int64_t row0 = (int64_t)blockIdx.x * 64 + wave * 16;
  const int64_t col0 = (int64_t)blockIdx.y * 64;
  const int64_t arow = row0 + l16;
  const bool rok = arow < n;
  const float* pa = A + (rok ? arow : 0) * lda;
  f64x4_t acc[4];
#pragma unroll
  for (int j = 0; j < 4; j++) acc[j] = (f64x4_t){0, 0, 0, 0};
  for (int64_t k0 = 0; k0 < ka; k0 += 16) {
    double av[4], bv[4][4];
#pragma unroll
    for (int s = 0; s < 4; s++) {
      int64_t kk = k0 + 4 * s + kq;
      bool kok = kk < ka;
      av[s] = (rok && kok) ? (double)pa[kk] : 0.0;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        int64_t col = col0 + 16 * j + l16;
        bv[s][j] = (kok && col < kb) ? B[kk * kb + col] : 0.0;
      }
    }
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int j = 0; j < 4; j++)
        acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s], bv[s][j], acc[j], 0, 0, 0);
  }
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int64_t r = row0 + kq + 4 * q, col = col0 + 16 * j + l16;  // D: row lane/16 + 4q
      if (r < n && col < kb) Y[r * ldy + col] = (float)acc[j][q];
    }
}

tsvm_status gemm_nn_tall(Ctx* c, const float* A, int64_t lda, int64_t ka, const double* B,
                         int64_t kb, int64_t n, float* Y, int64_t ldy) {
  if (n <= 0 || kb <= 0) return TSVM_OK;
  dim3 grid((unsigned)ceil_div(n, 64), (unsigned)ceil_div(kb, 64));
  hipLaunchKernelGGL(gemm_nn_tall_kernel, grid, dim3(256), 0, c->stream, A, lda, ka, B, kb, n, Y,
                     ldy);
  TSVM_HIP(c, hipGetLastError());
  return TSVM_OK;
}

}  // namespace tsvm
