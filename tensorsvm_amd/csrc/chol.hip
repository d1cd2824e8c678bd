// FP64 Cholesky of the k x k capacitance matrix M = I + Z^T D^-1 Z and its solves, on
// the GPU (replaces LAPACKE_dpotrf, tensorsvm-train.cu:1709, and LAPACKE_dpotrs, 1834).
//
// Blocked right-looking factorisation with 32-column panels; per panel three launches:
//   chol_diag    one block factors the 32 x 32 diagonal block in LDS and inverts it
//   chol_panel   rows below: L_i = A_i L_jj^{-T} (one thread per row)
//   chol_update  trailing lower triangle A -= L_p L_p^T (one block per 32 x 32 tile)
// A failure (non-positive pivot, as dpotrf's info > 0) is recorded in a device int and
// every later launch (and chol_solve_f64) becomes a no-op, so the host reads one flag.
// The inverted diagonal blocks are kept and reused by the triangular solves.
#include "common.h"
#include "internal.h"

namespace tsvm {

constexpr int CB = 32;

__global__ __launch_bounds__(256) void chol_diag(double* __restrict__ M, int64_t k, int64_t j0,
                                                 double* __restrict__ Dinv, int* __restrict__ info) {
  if (*info != 0) return;
  __shared__ double A[CB][CB + 1];
  __shared__ double Li[CB][CB + 1];
  __shared__ int bad;
  const int tid = threadIdx.x;
  const int jb = (int)(k - j0 < CB ? k - j0 : CB);
  if (tid == 0) bad = 0;
  for (int e = tid; e < CB * CB; e += 256) {
    int i = e / CB, j = e % CB;
    A[i][j] = (i < jb && j < jb) ? M[(j0 + i) * k + j0 + j] : (i == j ? 1.0 : 0.0);
    Li[i][j] = 0.0;
  }
  __syncthreads();
  for (int j = 0; j < jb; j++) {
    if (tid == 0) {
      double s = A[j][j];
      if (!(s > 0.0)) {
        bad = j + 1;
        s = 1.0;
      }
      A[j][j] = sqrt(s);
    }
    __syncthreads();
    if (bad) break;
    if (tid > j && tid < jb) A[tid][j] /= A[j][j];
    __syncthreads();
    for (int e = tid; e < CB * CB; e += 256) {
      int i = e / CB, l = e % CB;
      if (i > j && l > j && l <= i && i < jb) A[i][l] -= A[i][j] * A[l][j];
    }
    __syncthreads();
  }
  if (bad) {
    if (tid == 0) *info = (int)(j0 + bad);
    return;
  }
  // inverse of the lower-triangular block: thread c solves column c
  if (tid < CB) {
    const int c = tid;
    for (int i = 0; i < CB; i++) {
      if (i < c) continue;
      double v = (i == c) ? 1.0 : 0.0;
      for (int t = c; t < i; t++) v -= A[i][t] * Li[t][c];
      Li[i][c] = v / A[i][i];
    }
  }
  __syncthreads();
  for (int e = tid; e < CB * CB; e += 256) {
    int i = e / CB, j = e % CB;
    if (i < jb && j < jb) M[(j0 + i) * k + j0 + j] = (j <= i) ? A[i][j] : 0.0;
    Dinv[(j0 / CB) * CB * CB + i * CB + j] = (j <= i) ? Li[i][j] : 0.0;
  }
}

__global__ __launch_bounds__(256) void chol_panel(double* __restrict__ M, int64_t k, int64_t j0,
                                                  const double* __restrict__ Dinv,
                                                  const int* __restrict__ info) {
  if (*info != 0) return;
  __shared__ double Li[CB][CB + 1];
  const double* D = Dinv + (j0 / CB) * CB * CB;
  for (int e = threadIdx.x; e < CB * CB; e += 256) Li[e / CB][e % CB] = D[e];
  __syncthreads();
  const int64_t i = j0 + CB + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= k) return;
  double a[CB];
  double* row = M + i * k + j0;
#pragma unroll
  for (int t = 0; t < CB; t++) a[t] = row[t];
  // L_i[c] = sum_{t <= c} A_i[t] Linv[c][t]
#pragma unroll
  for (int c = 0; c < CB; c++) {
    double v = 0.0;
#pragma unroll
    for (int t = 0; t <= c; t++) v += a[t] * Li[c][t];
    row[c] = v;
  }
}

__global__ __launch_bounds__(256) void chol_update(double* __restrict__ M, int64_t k, int64_t j0,
                                                   const int* __restrict__ info) {
  if (*info != 0) return;
  const int ti = blockIdx.x, tl = blockIdx.y;
  if (tl > ti) return;
  __shared__ double Pi[CB][CB + 1];
  __shared__ double Pl[CB][CB + 1];
  const int64_t r0 = j0 + CB;
  const int64_t ib = r0 + (int64_t)ti * CB, lb = r0 + (int64_t)tl * CB;
  for (int e = threadIdx.x; e < CB * CB; e += 256) {
    int r = e / CB, t = e % CB;
    Pi[r][t] = (ib + r < k) ? M[(ib + r) * k + j0 + t] : 0.0;
    Pl[r][t] = (lb + r < k) ? M[(lb + r) * k + j0 + t] : 0.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < CB * CB; e += 256) {
    int r = e / CB, q = e % CB;
    int64_t i = ib + r, l = lb + q;
    if (i >= k || l >= k || l > i) continue;
    double v = 0.0;
#pragma unroll
    for (int t = 0; t < CB; t++) v += Pi[r][t] * Pl[q][t];
    M[i * k + l] -= v;
  }
}

__global__ void zero_upper(double* __restrict__ M, int64_t k, const int* __restrict__ info) {
  if (*info != 0) return;
  int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= k * k) return;
  int64_t i = idx / k, j = idx - i * k;
  if (j > i) M[idx] = 0.0;
}

__global__ void set_int(int* p, int v) { *p = v; }

tsvm_status chol_f64(Ctx* c, double* M, int64_t k, int* info) {
  void* dinv;
  TSVM_HIP(c, c->ws.get("chol_dinv", (size_t)(ceil_div(k, CB) * CB * CB * 8), &dinv));
  hipStream_t s = c->stream;
  hipLaunchKernelGGL(set_int, dim3(1), dim3(1), 0, s, info, 0);
  for (int64_t j0 = 0; j0 < k; j0 += CB) {
    hipLaunchKernelGGL(chol_diag, dim3(1), dim3(256), 0, s, M, k, j0, (double*)dinv, info);
    const int64_t rest = k - j0 - CB;
    if (rest <= 0) break;
    hipLaunchKernelGGL(chol_panel, dim3((unsigned)ceil_div(rest, 256)), dim3(256), 0, s, M, k, j0,
                       (const double*)dinv, (const int*)info);
    const unsigned T = (unsigned)ceil_div(rest, CB);
    hipLaunchKernelGGL(chol_update, dim3(T, T), dim3(256), 0, s, M, k, j0, (const int*)info);
  }
  hipLaunchKernelGGL(zero_upper, dim3((unsigned)ceil_div(k * k, 256)), dim3(256), 0, s, M, k,
                     (const int*)info);
  TSVM_HIP(c, hipGetLastError());
  return TSVM_OK;
}

// (L L^T) X = B for p <= 4 right-hand sides (rows of B), one block of 256 threads.
// Forward: y_b = Linv_bb (b_b - L_{b,<b} y_{<b}); backward: x_b = Linv_bb^T (y_b - L_{>b,b}^T x_{>b}).
__global__ __launch_bounds__(256) void chol_solve_kernel(const double* __restrict__ L, int64_t k,
                                                         const double* __restrict__ Dinv,
                                                         double* __restrict__ B, int p,
                                                         const int* __restrict__ info) {
  if (*info != 0) return;
  extern __shared__ double smem[];
  double* X = smem;                 // p x k
  double* part = smem + 4 * k;      // [8][CB][4] partial sums
  double* tmp = part + 8 * CB * 4;  // [CB][4]
  const int tid = threadIdx.x;
  const int r = tid % CB, lane8 = tid / CB;  // 32 rows x 8 partial lanes
  for (int64_t e = tid; e < (int64_t)p * k; e += 256) X[e] = B[e];
  __syncthreads();
  const int64_t nb = ceil_div(k, CB);
  // forward
  for (int64_t b = 0; b < nb; b++) {
    const int64_t j0 = b * CB;
    const int jb = (int)(k - j0 < CB ? k - j0 : CB);
    // forward: 8 consecutive threads share a row of L (coalesced along t)
    const int fr = tid >> 3, fg = tid & 7;
    double acc[4] = {0, 0, 0, 0};
    if (fr < jb) {
      const double* Lr = L + (j0 + fr) * k;
      int64_t t = fg;
      for (; t + 24 < j0; t += 32) {
        double l0 = Lr[t], l1 = Lr[t + 8], l2 = Lr[t + 16], l3 = Lr[t + 24];
        for (int q = 0; q < p; q++)
          acc[q] += l0 * X[q * k + t] + l1 * X[q * k + t + 8] + l2 * X[q * k + t + 16] +
                    l3 * X[q * k + t + 24];
      }
      for (; t < j0; t += 8) {
        double l = Lr[t];
        for (int q = 0; q < p; q++) acc[q] += l * X[q * k + t];
      }
    }
    for (int q = 0; q < 4; q++) part[(fg * CB + fr) * 4 + q] = acc[q];
    __syncthreads();
    if (tid < CB * 4) {
      int rr = tid / 4, q = tid % 4;
      if (rr < jb && q < p) {
        double s = 0.0;
        for (int u = 0; u < 8; u++) s += part[(u * CB + rr) * 4 + q];
        tmp[rr * 4 + q] = X[q * k + j0 + rr] - s;
      }
    }
    __syncthreads();
    if (tid < CB * 4) {
      int rr = tid / 4, q = tid % 4;
      if (rr < jb && q < p) {
        const double* Di = Dinv + b * CB * CB + rr * CB;
        double s = 0.0;
        for (int t = 0; t <= rr; t++) s += Di[t] * tmp[t * 4 + q];
        X[q * k + j0 + rr] = s;
      }
    }
    __syncthreads();
  }
  // backward
  for (int64_t b = nb - 1; b >= 0; b--) {
    const int64_t j0 = b * CB;
    const int jb = (int)(k - j0 < CB ? k - j0 : CB);
    double acc[4] = {0, 0, 0, 0};
    if (r < jb) {
      int64_t t = j0 + CB + lane8;
      for (; t + 24 < k; t += 32) {
        double l0 = L[t * k + j0 + r], l1 = L[(t + 8) * k + j0 + r];
        double l2 = L[(t + 16) * k + j0 + r], l3 = L[(t + 24) * k + j0 + r];
        for (int q = 0; q < p; q++)
          acc[q] += l0 * X[q * k + t] + l1 * X[q * k + t + 8] + l2 * X[q * k + t + 16] +
                    l3 * X[q * k + t + 24];
      }
      for (; t < k; t += 8) {
        double l = L[t * k + j0 + r];
        for (int q = 0; q < p; q++) acc[q] += l * X[q * k + t];
      }
    }
    for (int q = 0; q < 4; q++) part[(lane8 * CB + r) * 4 + q] = acc[q];
    __syncthreads();
    if (tid < CB * 4) {
      int rr = tid / 4, q = tid % 4;
      if (rr < jb && q < p) {
        double s = 0.0;
        for (int u = 0; u < 8; u++) s += part[(u * CB + rr) * 4 + q];
        tmp[rr * 4 + q] = X[q * k + j0 + rr] - s;
      }
    }
    __syncthreads();
    if (tid < CB * 4) {
      int rr = tid / 4, q = tid % 4;
      if (rr < jb && q < p) {
        // x_rr = sum_{t >= rr} Linv[t][rr] tmp[t]
        const double* Db = Dinv + b * CB * CB;
        double s = 0.0;
        for (int t = rr; t < jb; t++) s += Db[t * CB + rr] * tmp[t * 4 + q];
        X[q * k + j0 + rr] = s;
      }
    }
    __syncthreads();
  }
  for (int64_t e = tid; e < (int64_t)p * k; e += 256) B[e] = X[e];
}

tsvm_status chol_solve_f64(Ctx* c, const double* L, int64_t k, double* B, int64_t p,
                           const int* info) {
  if (p < 1 || p > 4) return fail(c, TSVM_ERR_ARG, "chol_solve: 1 <= p <= 4");
  void* dinv;
  TSVM_HIP(c, c->ws.get("chol_dinv", (size_t)(ceil_div(k, CB) * CB * CB * 8), &dinv));
  size_t sm = (size_t)(4 * k + 8 * CB * 4 + CB * 4) * 8;
  if (sm > 160 * 1024) return fail(c, TSVM_ERR_ARG, "chol_solve: k=%ld too large", (long)k);
  hipLaunchKernelGGL(chol_solve_kernel, dim3(1), dim3(256), sm, c->stream, L, k,
                     (const double*)dinv, B, (int)p, info);
  TSVM_HIP(c, hipGetLastError());
  return TSVM_OK;
}

}  // namespace tsvm
