// -analysis: |K_Y|_F and |K_Y - U U^T|_F (replaces rbf_fnorm_res, tensorsvm-train.cu:
// 2388-2467, which materialises 8192^2 FP64 tiles of K and U U^T and reduces them with
// thrust).  One persistent kernel walks the 128 x 128 tiles of the lower triangle: U_I U_J^T
// on FP32 MFMA (32x32x2, the factor is FP32) staged through LDS in 32-column slices, the
// kernel values from the rows of the tile (FP32 distances, FP64 exp), and both squared
// sums accumulated per block in FP64 (off-diagonal tiles count twice).  Each block walks a
// fixed sequence of tiles and the host adds the per-block sums in block order, so the
// result is deterministic.  Work: n^2 (2k + 3d) / 2 flops; nothing of size n^2 is stored.
#include <math.h>

#include <vector>

#include "common.h"
#include "internal.h"

namespace tsvm {

constexpr int FN_T = 128;  // tile edge
constexpr int FN_K = 32;   // factor columns per LDS slice
constexpr int FN_LS = FN_K + 1;
constexpr int FN_DMAX = 64;  // features held in LDS for the tile's rows

__global__ __launch_bounds__(256) void fnorm_res_kernel(const float* __restrict__ X,
                                                        const float* __restrict__ y, int64_t n,
                                                        int d, double gamma,
                                                        const float* __restrict__ U, int64_t ldu,
                                                        int k, int64_t nb, int64_t ntiles,
                                                        double* __restrict__ part) {
  __shared__ float sA[FN_T * FN_LS];
  __shared__ float sB[FN_T * FN_LS];
  __shared__ double red[2][256];
  __shared__ float sx[2 * FN_T * FN_DMAX];  // rows of the tile: [2][FN_T][d]
  float* xI = sx;
  float* xJ = sx + FN_T * d;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l32 = lane & 31, h = lane >> 5;
  const int wa = (wave >> 1) * 64, wb = (wave & 1) * 64;
  double s_k = 0.0, s_r = 0.0;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    // lower-triangle tile t -> (bi >= bj)
    int64_t bi = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (bi * (bi + 1) / 2 > t) bi--;
    while ((bi + 1) * (bi + 2) / 2 <= t) bi++;
    const int64_t bj = t - bi * (bi + 1) / 2;
    const int64_t i0 = bi * FN_T, j0 = bj * FN_T;
    const double wgt = bi == bj ? 1.0 : 2.0;
    __syncthreads();  // previous tile's LDS reads are done
    for (int e = tid; e < FN_T * d; e += 256) {
      const int r = e / d, c = e - r * d;
      xI[e] = i0 + r < n ? X[(i0 + r) * d + c] : 0.0f;
      xJ[e] = j0 + r < n ? X[(j0 + r) * d + c] : 0.0f;
    }
    f32x16_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int q = 0; q < 16; q++) acc[a][b][q] = 0.0f;
    for (int c0 = 0; c0 < k; c0 += FN_K) {
      __syncthreads();
      for (int e = tid; e < FN_T * FN_K; e += 256) {
        const int r = e / FN_K, c = e - r * FN_K;
        const bool cok = c0 + c < k;
        sA[r * FN_LS + c] = (i0 + r < n && cok) ? U[(i0 + r) * ldu + c0 + c] : 0.0f;
        sB[r * FN_LS + c] = (j0 + r < n && cok) ? U[(j0 + r) * ldu + c0 + c] : 0.0f;
      }
      __syncthreads();
#pragma unroll 4
      for (int ks = 0; ks < FN_K / 2; ks++) {
        const float a0 = sA[(wa + l32) * FN_LS + 2 * ks + h];
        const float a1 = sA[(wa + 32 + l32) * FN_LS + 2 * ks + h];
        const float b0 = sB[(wb + l32) * FN_LS + 2 * ks + h];
        const float b1 = sB[(wb + 32 + l32) * FN_LS + 2 * ks + h];
        acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, acc[1][1], 0, 0, 0);
      }
    }
    // C layout: register q of block (a, b) holds row wa + 32a + 8(q/4) + 4h + q%4 and
    // column wb + 32b + l32 of the tile
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int q = 0; q < 16; q++) {
          const int r = wa + 32 * a + 8 * (q / 4) + 4 * h + (q % 4);
          const int cc = wb + 32 * b + l32;
          const int64_t i = i0 + r, j = j0 + cc;
          if (i >= n || j >= n) continue;
          float dist = 0.0f;
          for (int f = 0; f < d; f++) {
            const float df = xI[r * d + f] - xJ[cc * d + f];
            dist = fmaf(df, df, dist);
          }
          const double kv = (double)y[i] * (double)y[j] * exp(-gamma * (double)dist);
          const double rv = kv - (double)acc[a][b][q];
          s_k += wgt * kv * kv;
          s_r += wgt * rv * rv;
        }
  }
  red[0][tid] = s_k;
  red[1][tid] = s_r;
  __syncthreads();
  if (tid == 0) {
    double a = 0.0, b = 0.0;
    for (int t = 0; t < 256; t++) {
      a += red[0][t];
      b += red[1][t];
    }
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

tsvm_status fnorm_res_dev(Ctx* c, const float* X, const float* y, int64_t n, int64_t d,
                          double gamma, const float* U, int64_t ldu, int64_t k, double* fnorm,
                          double* fnorm_res) {
  if (n <= 0 || d <= 0 || d > FN_DMAX || k <= 0 || ldu < k)
    return fail(c, TSVM_ERR_ARG, "fnorm_res: bad shape n=%ld d=%ld k=%ld", (long)n, (long)d,
                (long)k);
  const int64_t nb = ceil_div(n, FN_T);
  const int64_t ntiles = nb * (nb + 1) / 2;
  const int64_t grid = ntiles < (int64_t)c->cu_count * 4 ? ntiles : (int64_t)c->cu_count * 4;
  void* p;
  TSVM_HIP(c, c->ws.get("fn_part", (size_t)grid * 16, &p));
  double* part = (double*)p;
  hipLaunchKernelGGL(fnorm_res_kernel, dim3((unsigned)grid), dim3(256), 0, c->stream, X, y, n,
                     (int)d, gamma, U, ldu, (int)k, nb, ntiles, part);
  TSVM_HIP(c, hipGetLastError());
  std::vector<double> h((size_t)grid * 2);
  TSVM_HIP(c, hipMemcpyAsync(h.data(), part, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
  TSVM_HIP(c, hipStreamSynchronize(c->stream));
  double a = 0.0, b = 0.0;
  for (int64_t g = 0; g < grid; g++) {
    a += h[2 * g];
    b += h[2 * g + 1];
  }
  *fnorm = sqrt(a);
  *fnorm_res = sqrt(b);
  return TSVM_OK;
}

}  // namespace tsvm
