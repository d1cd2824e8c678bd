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
// (2 x 2 MFMA tiles of 16 x 16).  Rows of one split are walked 4 at a time (MFMA K = 4).
__global__ __launch_bounds__(256) void gemm_tn_f64_kernel(
    const float* __restrict__ A, int64_t lda, int64_t ka, const float* __restrict__ B,
    int64_t ldb, int64_t kb, int64_t n, const double* __restrict__ w, int lower_only,
    int64_t rows_per_split, double* __restrict__ part) {
  const int ta = blockIdx.x, tb = blockIdx.y;
  if (lower_only && tb > ta) return;
  const int64_t split = blockIdx.z;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const int64_t a0 = (int64_t)ta * 64 + (wave >> 1) * 32;
  const int64_t b0 = (int64_t)tb * 64 + (wave & 1) * 32;
  const int64_t r0 = split * rows_per_split;
  const int64_t r1 = r0 + rows_per_split < n ? r0 + rows_per_split : n;
  f64x4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = (f64x4_t){0, 0, 0, 0};
  const bool va0 = a0 + l16 < ka, va1 = a0 + 16 + l16 < ka;
  const bool vb0 = b0 + l16 < kb, vb1 = b0 + 16 + l16 < kb;
  for (int64_t rr = r0; rr < r1; rr += 16) {
    double av[4][2], bv[4][2];
#pragma unroll
    for (int s = 0; s < 4; s++) {
      int64_t row = rr + 4 * s + kq;
      bool ok = row < r1;
      double ws = ok ? (w ? w[row] : 1.0) : 0.0;
      const float* pa = A + (ok ? row : r0) * lda + a0 + l16;
      const float* pb = B + (ok ? row : r0) * ldb + b0 + l16;
      av[s][0] = va0 ? ws * (double)pa[0] : 0.0;
      av[s][1] = va1 ? ws * (double)pa[16] : 0.0;
      bv[s][0] = (vb0 && ok) ? (double)pb[0] : 0.0;
      bv[s][1] = (vb1 && ok) ? (double)pb[16] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s][i], bv[s][j], acc[i][j], 0, 0, 0);
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
        int64_t a = a0 + 16 * i + kq + 4 * q, b = b0 + 16 * j + l16;
        if (a < ka && b < kb) P[a * kb + b] = acc[i][j][q];
      }
}

__global__ void reduce_splits_kernel(const double* __restrict__ part, int64_t nsplit,
                                     int64_t ka, int64_t kb, int lower_only,
                                     double* __restrict__ C) {
  int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= ka * kb) return;
  int64_t a = idx / kb, b = idx - a * kb;
  if (lower_only && (b / 64) > (a / 64)) return;
  double s = 0.0;
  for (int64_t t = 0; t < nsplit; t++) s += part[t * ka * kb + idx];
  C[idx] = s;
  if (lower_only && (b / 64) < (a / 64)) C[b * kb + a] = s;
}

tsvm_status gemm_tn_f64(Ctx* c, const float* A, int64_t lda, int64_t ka, const float* B,
                        int64_t ldb, int64_t kb, int64_t n, const double* w, bool lower_only,
                        double* C) {
  if (ka <= 0 || kb <= 0) return TSVM_OK;
  if (lower_only && ka != kb) return fail(c, TSVM_ERR_ARG, "gemm_tn_f64: lower_only needs ka==kb");
  const int64_t ta = ceil_div(ka, 64), tb = ceil_div(kb, 64);
  const int64_t tiles = lower_only ? ta * (ta + 1) / 2 : ta * tb;
  // enough blocks to fill the chip several times; at least 256 rows per split
  int64_t nsplit = ceil_div((int64_t)c->cu_count * 8, tiles);
  int64_t max_split = ceil_div(n > 0 ? n : 1, 256);
  if (nsplit > max_split) nsplit = max_split;
  if (nsplit < 1) nsplit = 1;
  int64_t rps = round_up(ceil_div(n > 0 ? n : 1, nsplit), 16);
  nsplit = ceil_div(n > 0 ? n : 1, rps);
  void* part;
  TSVM_HIP(c, c->ws.get("gemm_part", (size_t)(nsplit * ka * kb * 8), &part));
  dim3 grid((unsigned)ta, (unsigned)tb, (unsigned)nsplit);
  hipLaunchKernelGGL(gemm_tn_f64_kernel, grid, dim3(256), 0, c->stream, A, lda, ka, B, ldb, kb,
                     n, w, lower_only ? 1 : 0, rps, (double*)part);
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
