// FP64 Cholesky of the k x k capacitance matrix M = I + Z^T D^-1 Z and its solves, on
// the GPU (replaces LAPACKE_dpotrf, tensorsvm-train.cu:1709, and LAPACKE_dpotrs, 1834).
//
// Blocked right-looking factorisation with 32-column panels, one launch per panel:
//   chol_diag    (first panel only) one wave factors the 32 x 32 diagonal block in
//                registers and inverts it
//   chol_step    block (ti, tl) of the trailing lower triangle forms the two panel tiles it
//                needs, L_i = A_i L_jj^{-T}, subtracts L_ti L_tl^T from its tile, and the
//                diagonal blocks write their panel tile out; block (0, 0) then factors the
//                next diagonal block from its updated tile, so the next launch has it
// The factor is built in a workspace copy (the panel columns of M are read by every block
// of a launch while the diagonal blocks write theirs); chol_finish copies it into M.
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

// Factors the diagonal block at (j0, j0): A points at its first entry (row stride lda, in
// global memory or LDS), the factor goes to Lw (k x k, row-major) and its inverse to Dinv.
// Called by every thread of a block of at least 64 threads (it has a block barrier); the
// first wave computes.  Lane i < 32 holds row i and the entries another lane needs come by
// readlane, so the factorisation needs no barrier.  The inverse is formed in the same loop:
// lane c keeps column c of L^-1 as partial sums s_i (starting from e_c); once column j of
// L is final, x_j = s_j / L_jj and s_i -= L_ij x_j for i > j, with the L_ij the
// factorisation has just read.  Rows past the matrix end are identity.
__device__ __forceinline__ void chol_diag_dev(const double* A, int64_t lda, double* __restrict__ Lw,
                                              int64_t k, int64_t j0, double* __restrict__ Dinv,
                                              int* __restrict__ info) {
  __shared__ int sbad;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const bool w0 = tid < 64;
  const int jb = (int)(k - j0 < CB ? k - j0 : CB);
  const bool row = w0 && lane < jb;
  if (w0) {
    double a[CB], sx[CB];
#pragma unroll
    for (int t = 0; t < CB; t++) {
      a[t] = row && t < jb ? A[lane * lda + t] : (lane == t ? 1.0 : 0.0);
      sx[t] = lane == t ? 1.0 : 0.0;
    }
    int bad = 0;
#pragma unroll
    for (int j = 0; j < CB; j++) {
      const double d = readlane_f64(a[j], j);
      if (!(d > 0.0)) {
        bad = j + 1;
        break;
      }
      // 1/sqrt(d): the hardware estimate and two Newton steps y <- y (3 - d y^2) / 2 (each
      // doubles the correct bits), so the pivot step needs no division on its critical path;
      // rs is also 1 / L_jj for the inverse
      double rs = __builtin_amdgcn_rsq(d);
      rs = rs * (1.5 - 0.5 * d * rs * rs);
      rs = rs * (1.5 - 0.5 * d * rs * rs);
      // Every lane updates every entry: those above a lane's diagonal (t > lane) are never
      // read (the loop reads a[j] only on lanes l >= j, and only t <= lane is written out),
      // so the updates need no lane masks; on lane j, a[j] rs = d rs is L_jj.
      a[j] *= rs;
      const double xj = sx[j] * rs;
      sx[j] = xj;
#pragma unroll
      for (int l = j + 1; l < CB; l++) {
        const double llj = readlane_f64(a[j], l);
        a[l] -= a[j] * llj;
        sx[l] -= llj * xj;
      }
    }
    if (lane == 0) sbad = bad;  // wave-uniform
    if (!bad && row)
#pragma unroll
      for (int t = 0; t < CB; t++)
        if (t < jb) Lw[(j0 + lane) * k + j0 + t] = t <= lane ? a[t] : 0.0;
    if (!bad && lane < CB)
#pragma unroll
      for (int i = 0; i < CB; i++) Dinv[(j0 / CB) * CB * CB + i * CB + lane] = sx[i];
  }
  __syncthreads();
  if (sbad && tid == 0) *info = (int)(j0 + sbad);
}

// the first diagonal block
__global__ __launch_bounds__(64) void chol_diag(const double* __restrict__ M, double* __restrict__ Lw,
                                                int64_t k, double* __restrict__ Dinv,
                                                int* __restrict__ info) {
  if (*info != 0) return;
  chol_diag_dev(M, k, Lw, k, 0, Dinv, info);
}

// Panel j0 (its diagonal block factored: inverse in Dinv) and the trailing update.  Block
// x covers tile (ti, tl), tl <= ti, x = ti (ti + 1) / 2 + tl, of the trailing lower
// triangle (rows and columns from j0 + 32).  It forms the panel tiles of its row and
// column blocks, P[r][c] = sum_{t <= c} A[r][t] Linv[c][t] (A: M's columns j0 .. j0 + 31),
// and subtracts P_ti P_tl^T from its tile of M; the diagonal blocks write their panel tile
// to Lw.  Block 0 keeps its updated tile, the next diagonal block, in LDS and factors it.
__global__ __launch_bounds__(256) void chol_step(double* __restrict__ M, double* __restrict__ Lw,
                                                 int64_t k, int64_t j0, double* __restrict__ Dinv,
                                                 int* __restrict__ info) {
  if (*info != 0) return;
  const int x = blockIdx.x;
  int ti = (int)((sqrt(8.0 * x + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= x) ti++;
  while (ti * (ti + 1) / 2 > x) ti--;
  const int tl = x - ti * (ti + 1) / 2;
  const bool dg = ti == tl;
  __shared__ double Li[CB][CB + 1];
  __shared__ double Pi[CB][CB + 1];
  __shared__ double Pl[CB][CB + 1];
  __shared__ double Tn[CB][CB + 1];  // block 0: the updated tile
  const double* D = Dinv + (j0 / CB) * CB * CB;
  const int64_t r0 = j0 + CB;
  const int64_t ib = r0 + (int64_t)ti * CB, lb = r0 + (int64_t)tl * CB;
  const int tid = threadIdx.x;
  for (int e = tid; e < CB * CB; e += 256) {
    const int r = e / CB, t = e % CB;
    Li[r][t] = D[e];
    Pi[r][t] = ib + r < k ? M[(ib + r) * k + j0 + t] : 0.0;
    if (!dg) Pl[r][t] = lb + r < k ? M[(lb + r) * k + j0 + t] : 0.0;
  }
  __syncthreads();
  constexpr int PER = CB * CB / 256;
  double pv[PER], qv[PER];
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int e = tid + 256 * u, r = e / CB, c = e % CB;
    double v = 0.0, w = 0.0;
#pragma unroll
    for (int t = 0; t < CB; t++)
      if (t <= c) {
        v += Pi[r][t] * Li[c][t];
        if (!dg) w += Pl[r][t] * Li[c][t];
      }
    pv[u] = v;
    qv[u] = w;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int e = tid + 256 * u, r = e / CB, c = e % CB;
    Pi[r][c] = pv[u];
    if (!dg) Pl[r][c] = qv[u];
    if (dg && ib + r < k) Lw[(ib + r) * k + j0 + c] = pv[u];
  }
  __syncthreads();
  const double(*Q)[CB + 1] = dg ? Pi : Pl;
  for (int e = tid; e < CB * CB; e += 256) {
    const int r = e / CB, q = e % CB;
    const int64_t i = ib + r, l = lb + q;
    if (i >= k || l >= k || l > i) {
      if (x == 0) Tn[r][q] = 0.0;
      continue;
    }
    double v = 0.0;
#pragma unroll
    for (int t = 0; t < CB; t++) v += Pi[r][t] * Q[q][t];
    const double nv = M[i * k + l] - v;
    M[i * k + l] = nv;
    if (x == 0) Tn[r][q] = nv;
  }
  if (x == 0) {
    __syncthreads();
    chol_diag_dev(&Tn[0][0], CB + 1, Lw, k, r0, Dinv, info);
  }
}

// M = the factor (strict upper triangle zero) and Lt = L^T, from the factor's workspace
// copy (32 x 32 tiles through LDS)
__global__ __launch_bounds__(256) void chol_finish(double* __restrict__ M, const double* __restrict__ Lw,
                                                   int64_t k, double* __restrict__ Lt,
                                                   const int* __restrict__ info) {
  if (*info != 0) return;
  __shared__ double tile[CB][CB + 1];
  const int64_t i0 = (int64_t)blockIdx.y * CB, j0 = (int64_t)blockIdx.x * CB;
  for (int e = threadIdx.x; e < CB * CB; e += 256) {
    const int r = e / CB, c = e % CB;
    const int64_t i = i0 + r, j = j0 + c;
    double v = 0.0;
    if (i < k && j < k) {
      v = j <= i ? Lw[i * k + j] : 0.0;
      M[i * k + j] = v;
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
  void *dinv, *lw, *lt;
  TSVM_HIP(c, c->ws.get("chol_dinv", (size_t)(ceil_div(k, CB) * CB * CB * 8), &dinv));
  TSVM_HIP(c, c->ws.get("chol_lw", (size_t)(k * k * 8), &lw));
  TSVM_HIP(c, c->ws.get("chol_lt", (size_t)(k * k * 8), &lt));
  hipStream_t s = c->stream;
  hipLaunchKernelGGL(set_int, dim3(1), dim3(1), 0, s, info, 0);
  hipLaunchKernelGGL(chol_diag, dim3(1), dim3(64), 0, s, (const double*)M, (double*)lw, k,
                     (double*)dinv, info);
  for (int64_t j0 = 0; j0 + CB < k; j0 += CB) {
    const int64_t T = ceil_div(k - j0 - CB, CB);
    hipLaunchKernelGGL(chol_step, dim3((unsigned)(T * (T + 1) / 2)), dim3(256), 0, s, M,
                       (double*)lw, k, j0, (double*)dinv, info);
  }
  const unsigned T = (unsigned)ceil_div(k, CB);
  hipLaunchKernelGGL(chol_finish, dim3(T, T), dim3(256), 0, s, M, (const double*)lw, k,
                     (double*)lt, (const int*)info);
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
// latency.  Needs k - 32 <= 1024 (one row or column per thread).  The block products run
// over all 32 entries without branches: Linv_bb is stored with exact zeros above its
// diagonal and in the identity rows and columns of a partial last block.
constexpr int SOLVE_THREADS = 1024;
template <int p>
__global__ __launch_bounds__(SOLVE_THREADS) void chol_solve_kernel(
    const double* __restrict__ L, const double* __restrict__ Lt, int64_t k,
    const double* __restrict__ Dinv, double* __restrict__ B, const int* __restrict__ info) {
  if (*info != 0) return;
  extern __shared__ double smem[];
  double* X = smem;              // p x k
  double* Y = smem + 4 * k;      // [4][CB] the current block
  double* Db = Y + 4 * CB;       // [CB][CB + 1] Linv_bb (padded rows: the forward block
                                 // product reads a row per thread)
  const int tid = threadIdx.x;
  const int q = tid / CB, rr = tid % CB;  // the block solve: q < p, rr < jb
  for (int64_t e = tid; e < (int64_t)p * k; e += SOLVE_THREADS) X[e] = B[e];
  const int64_t nb = ceil_div(k, CB);
  double l[CB];
  for (int64_t b = 0; b < nb; b++) {
    const int64_t j0 = b * CB;
    const int jb = (int)(k - j0 < CB ? k - j0 : CB);
    const int64_t r = j0 + jb + tid;  // this thread's row below the block
    Db[(tid / CB) * (CB + 1) + tid % CB] = Dinv[b * CB * CB + tid];
    if (r < k)
#pragma unroll
      for (int t = 0; t < CB; t++) l[t] = t < jb ? Lt[(j0 + t) * k + r] : 0.0;
    __syncthreads();
    if (q < p && rr < jb) {
      double s = 0.0;
#pragma unroll
      for (int t = 0; t < CB; t++)
        s += Db[rr * (CB + 1) + t] * X[q * k + j0 + (t < jb ? t : 0)];
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
    Db[(tid / CB) * (CB + 1) + tid % CB] = Dinv[b * CB * CB + tid];
    if (cc < j0)
#pragma unroll
      for (int t = 0; t < CB; t++) l[t] = t < jb ? L[(j0 + t) * k + cc] : 0.0;
    __syncthreads();
    if (q < p && rr < jb) {
      double s = 0.0;
#pragma unroll
      for (int t = 0; t < CB; t++)
        s += Db[t * (CB + 1) + rr] * X[q * k + j0 + (t < jb ? t : 0)];
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
  const size_t sm = (size_t)(4 * k + 4 * CB + CB * (CB + 1)) * 8;
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
