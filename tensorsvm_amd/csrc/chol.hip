// FP64 Cholesky of the k x k capacitance matrix M = I + Z^T D^-1 Z and its solves, on
// the GPU (replaces LAPACKE_dpotrf, tensorsvm-train.cu:1709, and LAPACKE_dpotrs, 1834).
//
// Blocked right-looking factorisation with 32-column panels; per panel three launches:
//   chol_diag    one wave factors the 32 x 32 diagonal block in registers and inverts it
//   chol_panel   rows below: L_i = A_i L_jj^{-T} (one thread per entry)
//   chol_update  trailing lower triangle A -= L_p L_p^T (one block per 32 x 32 tile)
// A failure (non-positive pivot, as dpotrf's info > 0) is recorded in a device int and
// every later launch (and chol_solve_f64) becomes a no-op, so the host reads one flag.
// The inverted diagonal blocks and a transposed copy of the factor are kept for the
// triangular solves (chol_solve_f64 follows chol_f64 on the same context).
#include "common.h"
#include "internal.h"

namespace tsvm {

constexpr int CB = 32;

__device__ inline double readlane_f64(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// One wave factors the 32 x 32 diagonal block in registers (lane i < 32 holds row i) and
// inverts the factor (lane c < 32 then holds column c of the inverse); the entries
// another lane needs come by readlane, so there is no barrier.  Rows past the matrix
// end are identity.
__global__ __launch_bounds__(64) void chol_diag(double* __restrict__ M, int64_t k, int64_t j0,
                                                double* __restrict__ Dinv, int* __restrict__ info) {
  if (*info != 0) return;
  const int lane = threadIdx.x;
  const int jb = (int)(k - j0 < CB ? k - j0 : CB);
  const bool row = lane < jb;
  double a[CB];
#pragma unroll
  for (int t = 0; t < CB; t++)
    a[t] = row && t < jb ? M[(j0 + lane) * k + j0 + t] : (lane == t ? 1.0 : 0.0);
  int bad = 0;
#pragma unroll
  for (int j = 0; j < CB; j++) {
    const double d = readlane_f64(a[j], j);
    if (!(d > 0.0)) {
      bad = j + 1;
      break;
    }
    // 1/sqrt(d): the hardware estimate and two Newton steps y <- y (3 - d y^2) / 2 (each
    // doubles the correct bits), so the pivot step needs no division on its critical path
    double rs = __builtin_amdgcn_rsq(d);
    rs = rs * (1.5 - 0.5 * d * rs * rs);
    rs = rs * (1.5 - 0.5 * d * rs * rs);
    const double s = d * rs;
    a[j] = lane == j ? s : (lane > j ? a[j] * rs : a[j]);
#pragma unroll
    for (int l = j + 1; l < CB; l++) {
      const double llj = readlane_f64(a[j], l);
      if (lane >= l) a[l] -= a[j] * llj;
    }
  }
  if (bad) {
    if (lane == 0) *info = (int)(j0 + bad);
    return;
  }
  if (row)
#pragma unroll
    for (int t = 0; t < CB; t++)
      if (t < jb) M[(j0 + lane) * k + j0 + t] = t <= lane ? a[t] : 0.0;
  // column c of the inverse: Li[i][c] = (e_c[i] - sum_{t < i} L[i][t] Li[t][c]) / L[i][i],
  // with Li[t][c] = 0 above the diagonal (t < c).  Lane c keeps its column in registers
  // (unrolled loops, no LDS round trip between the dependent steps); the factor and its
  // reciprocal diagonal are read from LDS, every lane at the same address.
  __shared__ double Ls[CB][CB + 1];
  __shared__ double Rd[CB];
  if (lane < CB) {
#pragma unroll
    for (int t = 0; t < CB; t++) Ls[lane][t] = a[t];
    double dl = 0.0;  // L[lane][lane]
#pragma unroll
    for (int t = 0; t < CB; t++)
      if (t == lane) dl = a[t];
    Rd[lane] = 1.0 / dl;
  }
  __syncthreads();
  if (lane < CB) {
    const int c = lane;
    double x[CB];
#pragma unroll
    for (int i = 0; i < CB; i++) {
      double v = i == c ? 1.0 : 0.0;
#pragma unroll
      for (int t = 0; t < i; t++) v -= Ls[i][t] * x[t];
      x[i] = i >= c ? v * Rd[i] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < CB; i++) Dinv[(j0 / CB) * CB * CB + i * CB + c] = x[i];
  }
}

// rows below the diagonal block: L_i[c] = sum_{t <= c} A_i[t] Linv[c][t]; a block takes 8
// rows, one thread per (row, column)
__global__ __launch_bounds__(256) void chol_panel(double* __restrict__ M, int64_t k, int64_t j0,
                                                  const double* __restrict__ Dinv,
                                                  const int* __restrict__ info) {
  if (*info != 0) return;
  __shared__ double Li[CB][CB + 1];
  __shared__ double Ar[8][CB + 1];
  const double* D = Dinv + (j0 / CB) * CB * CB;
  for (int e = threadIdx.x; e < CB * CB; e += 256) Li[e / CB][e % CB] = D[e];
  const int r = threadIdx.x / CB, c = threadIdx.x % CB;
  const int64_t i = j0 + CB + blockIdx.x * 8 + r;
  Ar[r][c] = i < k ? M[i * k + j0 + c] : 0.0;
  __syncthreads();
  if (i >= k) return;
  double v = 0.0;
#pragma unroll
  for (int t = 0; t < CB; t++)
    if (t <= c) v += Ar[r][t] * Li[c][t];
  M[i * k + j0 + c] = v;
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

// zero the strict upper triangle and write Lt = L^T (32 x 32 tiles through LDS)
__global__ __launch_bounds__(256) void chol_finish(double* __restrict__ M, int64_t k,
                                                   double* __restrict__ Lt,
                                                   const int* __restrict__ info) {
  if (*info != 0) return;
  __shared__ double tile[CB][CB + 1];
  const int64_t i0 = (int64_t)blockIdx.y * CB, j0 = (int64_t)blockIdx.x * CB;
  for (int e = threadIdx.x; e < CB * CB; e += 256) {
    const int r = e / CB, c = e % CB;
    const int64_t i = i0 + r, j = j0 + c;
    double v = 0.0;
    if (i < k && j < k) {
      v = j <= i ? M[i * k + j] : 0.0;
      if (j > i) M[i * k + j] = 0.0;
    }
    tile[r][c] = v;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < CB * CB; e += 256) {
    const int r = e / CB, c = e % CB;  // Lt[j0 + r][i0 + c] = L[i0 + c][j0 + r]
    if (j0 + r < k && i0 + c < k) Lt[(j0 + r) * k + i0 + c] = tile[c][r];
  }
}

__global__ void set_int(int* p, int v) { *p = v; }

tsvm_status chol_f64(Ctx* c, double* M, int64_t k, int* info) {
  void* dinv;
  TSVM_HIP(c, c->ws.get("chol_dinv", (size_t)(ceil_div(k, CB) * CB * CB * 8), &dinv));
  hipStream_t s = c->stream;
  hipLaunchKernelGGL(set_int, dim3(1), dim3(1), 0, s, info, 0);
  for (int64_t j0 = 0; j0 < k; j0 += CB) {
    hipLaunchKernelGGL(chol_diag, dim3(1), dim3(64), 0, s, M, k, j0, (double*)dinv, info);
    const int64_t rest = k - j0 - CB;
    if (rest <= 0) break;
    hipLaunchKernelGGL(chol_panel, dim3((unsigned)ceil_div(rest, 8)), dim3(256), 0, s, M, k, j0,
                       (const double*)dinv, (const int*)info);
    const unsigned T = (unsigned)ceil_div(rest, CB);
    hipLaunchKernelGGL(chol_update, dim3(T, T), dim3(256), 0, s, M, k, j0, (const int*)info);
  }
  void* lt;
  TSVM_HIP(c, c->ws.get("chol_lt", (size_t)(k * k * 8), &lt));
  const unsigned T = (unsigned)ceil_div(k, CB);
  hipLaunchKernelGGL(chol_finish, dim3(T, T), dim3(256), 0, s, M, k, (double*)lt, (const int*)info);
  TSVM_HIP(c, hipGetLastError());
  return TSVM_OK;
}

// (L L^T) X = B for p <= 4 right-hand sides (rows of B) in one block of 1024 threads,
// right-looking:
//   forward, block b ascending:   y_b = Linv_bb x_b, then x_r -= L_rb y_b for every row below
//   backward, block b descending: x_b = Linv_bb^T y_b, then x_c -= L_bc^T x_b for every
//                                 column before
// At the start of a step every thread issues its loads at once (one entry of Linv_bb
// into LDS, the 32 entries of L its row or column needs, read from Lt = L^T in the
// forward sweep so that a wave's loads are coalesced), so a step costs one round of load
// latency.  Needs k - 32 <= 1024 (one row or column per thread).
constexpr int SOLVE_THREADS = 1024;
template <int p>
__global__ __launch_bounds__(SOLVE_THREADS) void chol_solve_kernel(
    const double* __restrict__ L, const double* __restrict__ Lt, int64_t k,
    const double* __restrict__ Dinv, double* __restrict__ B, const int* __restrict__ info) {
  if (*info != 0) return;
  extern __shared__ double smem[];
  double* X = smem;              // p x k
  double* Y = smem + 4 * k;      // [4][CB] the current block
  double* Db = Y + 4 * CB;       // [CB][CB] Linv_bb
  const int tid = threadIdx.x;
  const int q = tid / CB, rr = tid % CB;  // the block solve: q < p, rr < jb
  for (int64_t e = tid; e < (int64_t)p * k; e += SOLVE_THREADS) X[e] = B[e];
  const int64_t nb = ceil_div(k, CB);
  double l[CB];
  for (int64_t b = 0; b < nb; b++) {
    const int64_t j0 = b * CB;
    const int jb = (int)(k - j0 < CB ? k - j0 : CB);
    const int64_t r = j0 + jb + tid;  // this thread's row below the block
    Db[tid] = Dinv[b * CB * CB + tid];
    if (r < k)
#pragma unroll
      for (int t = 0; t < CB; t++) l[t] = t < jb ? Lt[(j0 + t) * k + r] : 0.0;
    __syncthreads();
    if (q < p && rr < jb) {
      double s = 0.0;
#pragma unroll
      for (int t = 0; t < CB; t++)
        if (t <= rr) s += Db[rr * CB + t] * X[q * k + j0 + t];
      Y[q * CB + rr] = s;
    }
    __syncthreads();
    if (q < p && rr < jb) X[q * k + j0 + rr] = Y[q * CB + rr];
    if (r < k) {
      double acc[p] = {};
#pragma unroll
      for (int t = 0; t < CB; t++)
#pragma unroll
        for (int qq = 0; qq < p; qq++) acc[qq] += l[t] * Y[qq * CB + t];
#pragma unroll
      for (int qq = 0; qq < p; qq++) X[qq * k + r] -= acc[qq];
    }
    __syncthreads();
  }
  for (int64_t b = nb - 1; b >= 0; b--) {
    const int64_t j0 = b * CB;
    const int jb = (int)(k - j0 < CB ? k - j0 : CB);
    const int64_t cc = tid;  // this thread's column before the block
    Db[tid] = Dinv[b * CB * CB + tid];
    if (cc < j0)
#pragma unroll
      for (int t = 0; t < CB; t++) l[t] = t < jb ? L[(j0 + t) * k + cc] : 0.0;
    __syncthreads();
    if (q < p && rr < jb) {
      double s = 0.0;
#pragma unroll
      for (int t = 0; t < CB; t++)
        if (t >= rr && t < jb) s += Db[t * CB + rr] * X[q * k + j0 + t];
      Y[q * CB + rr] = s;
    }
    __syncthreads();
    if (q < p && rr < jb) X[q * k + j0 + rr] = Y[q * CB + rr];
    if (cc < j0) {
      double acc[p] = {};
#pragma unroll
      for (int t = 0; t < CB; t++)
#pragma unroll
        for (int qq = 0; qq < p; qq++) acc[qq] += l[t] * Y[qq * CB + t];
#pragma unroll
      for (int qq = 0; qq < p; qq++) X[qq * k + cc] -= acc[qq];
    }
    __syncthreads();
  }
  for (int64_t e = tid; e < (int64_t)p * k; e += SOLVE_THREADS) B[e] = X[e];
}

tsvm_status chol_solve_f64(Ctx* c, const double* L, int64_t k, double* B, int64_t p,
                           const int* info) {
  if (p < 1 || p > 4) return fail(c, TSVM_ERR_ARG, "chol_solve: 1 <= p <= 4");
  void* dinv;
  TSVM_HIP(c, c->ws.get("chol_dinv", (size_t)(ceil_div(k, CB) * CB * CB * 8), &dinv));
  if (k - CB > SOLVE_THREADS) return fail(c, TSVM_ERR_ARG, "chol_solve: k=%ld too large", (long)k);
  const size_t sm = (size_t)(4 * k + 4 * CB + CB * CB) * 8;
  void* lt;
  TSVM_HIP(c, c->ws.get("chol_lt", (size_t)(k * k * 8), &lt));
  const double* di = (const double*)dinv;
  const double* Lt = (const double*)lt;
  const dim3 g(1), b(SOLVE_THREADS);
  switch (p) {
    case 1: hipLaunchKernelGGL(chol_solve_kernel<1>, g, b, sm, c->stream, L, Lt, k, di, B, info); break;
    case 2: hipLaunchKernelGGL(chol_solve_kernel<2>, g, b, sm, c->stream, L, Lt, k, di, B, info); break;
    case 3: hipLaunchKernelGGL(chol_solve_kernel<3>, g, b, sm, c->stream, L, Lt, k, di, B, info); break;
    default: hipLaunchKernelGGL(chol_solve_kernel<4>, g, b, sm, c->stream, L, Lt, k, di, B, info); break;
  }
  TSVM_HIP(c, hipGetLastError());
  return TSVM_OK;
}

}  // namespace tsvm
