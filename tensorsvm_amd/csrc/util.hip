// Small device helpers of the trainer (bias column sums).
#include "common.h"
#include "internal.h"

#include <vector>

namespace tsvm {

// part[b][c] = sum over the block's rows of m_i U[i, c], m_i = alpha_i if alpha_i > thr
__global__ __launch_bounds__(256) void masked_colsum_kernel(const float* __restrict__ U, int64_t n,
                                                            int64_t k, int64_t ldu,
                                                            const double* __restrict__ alpha,
                                                            double thr, int64_t rpb,
                                                            double* __restrict__ part) {
  const int64_t r0 = blockIdx.x * rpb;
  const int64_t r1 = r0 + rpb < n ? r0 + rpb : n;
  for (int64_t c = threadIdx.x; c < k; c += blockDim.x) {
    double s = 0.0;
    for (int64_t i = r0; i < r1; i++) {
      double a = alpha[i];
      if (a > thr) s += a * (double)U[i * ldu + c];
    }
    part[blockIdx.x * k + c] = s;
  }
}

tsvm_status masked_colsum(Ctx* c, const float* U, int64_t n, int64_t k, int64_t ldu,
                          const double* alpha, double thr, double* v_host) {
  int64_t nblk = (int64_t)c->cu_count * 4;
  if (nblk > n) nblk = n > 0 ? n : 1;
  const int64_t rpb = ceil_div(n, nblk);
  void* part;
  TSVM_HIP(c, c->ws.get("colsum_part", (size_t)(nblk * k * 8), &part));
  hipLaunchKernelGGL(masked_colsum_kernel, dim3((unsigned)nblk), dim3(256), 0, c->stream, U, n, k, ldu,
                     alpha, thr, rpb, (double*)part);
  TSVM_HIP(c, hipGetLastError());
  std::vector<double> h(nblk * k);
  TSVM_HIP(c, hipMemcpyAsync(h.data(), part, nblk * k * 8, hipMemcpyDeviceToHost, c->stream));
  TSVM_HIP(c, hipStreamSynchronize(c->stream));
  for (int64_t j = 0; j < k; j++) {
    double s = 0.0;
    for (int64_t b = 0; b < nblk; b++) s += h[b * k + j];
    v_host[j] = s;
  }
  return TSVM_OK;
}

}  // namespace tsvm
