// Gradients of the trained model with respect to theta = (log gamma, s_1..s_d), where
// the kernel is exp(-gamma sum_f e^{2 s_f} (x_if - x_jf)^2) (per-feature input scales),
// alpha and b held fixed (envelope theorem at the solver's optimum).  No counterpart in
// the reference; the dual objective and the factor follow tensorsvm-train.cu:1636-1643
// (pobj/dobj) and LRA 1941-2249.
//
//   D(theta) = sum alpha - 1/2 alpha^T P K_Y P alpha,  P = projector onto range(K_Y Omega)
//              (the trained dual through the factor: U U^T = P K_Y P when W^T K_Y W is
//              positive definite, i.e. no eigenvalue is clipped)
//   L(theta) = mean_t h(y_t f(x_t)),  f = sum_i alpha_i y_i K(x_t, x_i) + b, h the
//              smoothed hinge (1/2 - z | (1 - z)^2 / 2 | 0)
//
// With Y = K_Y Omega = W R (W orthonormal, R = W^T Y), g = K_Y P alpha:
//   dD = -1/2 (P a)^T dK_Y (P a) - ((I-P) a)^T dK_Y (Omega R^-1 W^T g)
//        - ((I-P) g)^T dK_Y (Omega R^-1 W^T a)            (a = alpha)
// (from dP = (I-P) dY Y^+ + (Y^+)^T dY^T (I-P), dY = dK_Y Omega).  Every term is a bilinear
// form u^T dK_Y v = (y u)^T dK (y v), and per feature
//   S_f(p, q) = sum_ij p_i q_j K_ij (x_if - x_jf)^2
//             = sum_i p_i [x_if^2 (K q)_i - 2 x_if (K (q x_f))_i + (K (q x_f^2))_i],
// with dK/ds_f = -2 gamma (x_if - x_jf)^2 K and dK/dlog gamma = -gamma |x_i - x_j|^2 K.
// So the backward pass is the fused RBF product applied to the right-hand sides
// [q, q x_f, q x_f^2] (2d + 1 columns per pair; the kernel tiles are recomputed on MFMA,
// nothing of size n^2 is stored) followed by a reduction over rows.  The features are
// centred first (the differences, hence K, are unchanged; the expansion loses less).
#include <math.h>
#include <string.h>

#include <chrono>
#include <vector>

#include <stdio.h>

#include "common.h"
#include "internal.h"

namespace tsvm {

namespace {
double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

// Xs = X * es - mu (n x d)
__global__ void grad_scale_kernel(const float* __restrict__ X, int64_t n, int d,
                                  const float* __restrict__ es_mu, float* __restrict__ Xs) {
  const int64_t tot = n * d;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int f = (int)(e % d);
    Xs[e] = X[e] * es_mu[f] - es_mu[d + f];
  }
}

// column sums of X (n x d FP32) per block: part[block][f] (fixed row ranges)
__global__ __launch_bounds__(256) void grad_colsum_kernel(const float* __restrict__ X, int64_t n,
                                                          int d, int64_t rpb,
                                                          double* __restrict__ part) {
  const int64_t r0 = blockIdx.x * rpb, r1 = r0 + rpb < n ? r0 + rpb : n;
  for (int f0 = 0; f0 < d; f0 += 256) {
    const int nf = d - f0 < 256 ? d - f0 : 256;
    const int lanes = 256 / nf;
    const int t = threadIdx.x, f = f0 + t % nf, rl = t / nf;
    double s = 0.0;
    if (rl < lanes)
      for (int64_t i = r0 + rl; i < r1; i += lanes) s += X[i * d + f];
    __shared__ double red[256];
    red[t] = rl < lanes ? s : 0.0;
    __syncthreads();
    if (t < nf) {
      double a = 0.0;
      for (int l = 0; l < lanes; l++) a += red[l * nf + t];
      part[blockIdx.x * (int64_t)d + f0 + t] = a;
    }
    __syncthreads();
  }
}

// out_i = base_i + sgn * sum_c A[i, c] v[c]   (A n x k FP32, v FP64; one wave per row)
__global__ __launch_bounds__(256) void grad_rowdot_kernel(const float* __restrict__ A, int64_t n,
                                                          int k, const double* __restrict__ v,
                                                          const double* __restrict__ base,
                                                          double sgn, double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t w0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t i = w0; i < n; i += nw) {
    double s = 0.0;
    for (int c = lane; c < k; c += 64) s += (double)A[i * k + c] * v[c];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) out[i] = (base ? base[i] : 0.0) + sgn * s;
  }
}

// right-hand side of one bilinear pair: R[i, c0 + (0 | 1+f | 1+d+f)] = q_i (1 | x_if | x_if^2),
// q_i = ysel_i v_i (ysel = labels or null for 1)
__global__ void grad_rhs_kernel(const float* __restrict__ Xs, int64_t n, int d,
                                const double* __restrict__ v, const float* __restrict__ ysel,
                                int64_t ldr, int64_t c0, float* __restrict__ R) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double q = v[i] * (ysel ? (double)ysel[i] : 1.0);
    float* r = R + i * ldr + c0;
    r[0] = (float)q;
    for (int f = 0; f < d; f++) {
      const double x = Xs[i * d + f];
      r[1 + f] = (float)(q * x);
      r[1 + d + f] = (float)(q * x * x);
    }
  }
}

// per-block partial S_f = sum_i p_i [x_if^2 W[i,c0] - 2 x_if W[i,c0+1+f] + W[i,c0+1+d+f]],
// p_i = psel_i u_i (fixed row ranges, fixed reduction order)
__global__ __launch_bounds__(256) void grad_reduce_kernel(const float* __restrict__ Xs, int64_t m,
                                                          int d, const double* __restrict__ u,
                                                          const float* __restrict__ psel,
                                                          const float* __restrict__ Wd,
                                                          int64_t ldw, int64_t c0, int64_t rpb,
                                                          double* __restrict__ part) {
  const int64_t r0 = blockIdx.x * rpb, r1 = r0 + rpb < m ? r0 + rpb : m;
  __shared__ double red[256];
  for (int f0 = 0; f0 < d; f0 += 256) {
    const int nf = d - f0 < 256 ? d - f0 : 256;
    const int lanes = 256 / nf;
    const int t = threadIdx.x, f = f0 + t % nf, rl = t / nf;
    double s = 0.0;
    if (rl < lanes)
      for (int64_t i = r0 + rl; i < r1; i += lanes) {
        const double p = u[i] * (psel ? (double)psel[i] : 1.0);
        const double x = Xs[i * d + f];
        const float* w = Wd + i * ldw + c0;
        s += p * (x * x * (double)w[0] - 2.0 * x * (double)w[1 + f] + (double)w[1 + d + f]);
      }
    red[t] = rl < lanes ? s : 0.0;
    __syncthreads();
    if (t < nf) {
      double a = 0.0;
      for (int l = 0; l < lanes; l++) a += red[l * nf + t];
      part[blockIdx.x * (int64_t)d + f0 + t] = a;
    }
    __syncthreads();
  }
}

namespace {

constexpr int GR_BLOCKS = 512;

// host sums of per-block partials (fixed order)
tsvm_status sum_parts(Ctx* c, const double* part, int nb, int d, std::vector<double>& out) {
  std::vector<double> h((size_t)nb * d);
  TSVM_HIP(c, hipMemcpyAsync(h.data(), part, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
  TSVM_HIP(c, hipStreamSynchronize(c->stream));
  out.assign(d, 0.0);
  for (int b = 0; b < nb; b++)
    for (int f = 0; f < d; f++) out[f] += h[(size_t)b * d + f];
  return TSVM_OK;
}

// x <- A^{-1} x (k x k row-major, Gaussian elimination with partial pivoting)
bool solve_dense(std::vector<double> A, int64_t k, double* x) {
  for (int64_t col = 0; col < k; col++) {
    int64_t piv = col;
    for (int64_t r = col + 1; r < k; r++)
      if (fabs(A[r * k + col]) > fabs(A[piv * k + col])) piv = r;
    if (A[piv * k + col] == 0.0) return false;
    if (piv != col) {
      for (int64_t j = 0; j < k; j++) std::swap(A[col * k + j], A[piv * k + j]);
      std::swap(x[col], x[piv]);
    }
    for (int64_t r = col + 1; r < k; r++) {
      const double l = A[r * k + col] / A[col * k + col];
      if (l == 0.0) continue;
      for (int64_t j = col; j < k; j++) A[r * k + j] -= l * A[col * k + j];
      x[r] -= l * x[col];
    }
  }
  for (int64_t col = k - 1; col >= 0; col--) {
    double s = x[col];
    for (int64_t j = col + 1; j < k; j++) s -= A[col * k + j] * x[j];
    x[col] = s / A[col * k + col];
  }
  return true;
}

}  // namespace

tsvm_status grad_dev(Ctx* c, const float* X, const float* y, int64_t n, int64_t d,
                     const double* alpha, double b, int64_t k, double log_gamma,
                     const double* log_scales, const float* Om, const float* Xt,
                     const float* yt, int64_t nt, int precision, double* vals, double* gD,
                     double* gL, tsvm_grad_info* info) {
  if (n <= 0 || d <= 0 || d > 128 || k <= 0 || k > n || nt < 0 || !X || !y || !alpha || !Om ||
      !log_scales || !vals || (nt > 0 && (!Xt || !yt)) ||
      (precision != TSVM_PREC_FP32 && precision != TSVM_PREC_BF16))
    return fail(c, TSVM_ERR_ARG, "grad: bad arguments (n=%ld d=%ld k=%ld nt=%ld)", (long)n,
                (long)d, (long)k, (long)nt);
  const double t0 = now();
  hipStream_t s = c->stream;
  const double gamma = exp(log_gamma);
  tsvm_grad_info inf;
  memset(&inf, 0, sizeof(inf));
  const int64_t rtr = 3 * (2 * d + 1), rte = 2 * d + 1;
  const int64_t nmax = n > nt ? n : nt;
  void* p;
#define GWS(name, bytes, T, var)                             \
  TSVM_HIP(c, c->ws.get(name, (size_t)(bytes), &p));         \
  T* var = (T*)p;
  GWS("gr_Xs", n * d * 4, float, Xs);
  GWS("gr_Xts", (nt > 0 ? nt : 1) * d * 4, float, Xts);
  GWS("gr_esmu", 2 * d * 4, float, esmu);
  GWS("gr_part", 3 * GR_BLOCKS * d * 8, double, part);
  GWS("gr_Y", n * k * 4, float, Yb);
  GWS("gr_W", n * k * 4, float, Wb);
  GWS("gr_Q", n * k * 4, float, Qb);
  GWS("gr_kk", k * k * 8, double, dkk);
  GWS("gr_kv", 2 * k * 8, double, dkv);
  GWS("gr_vec", (7 * n + nt) * 8, double, vec);
  GWS("gr_ones", nmax * 4, float, ones);
#undef GWS
  // ---- scaled, centred features
  std::vector<double> mu;
  const int64_t rpb = ceil_div(n, GR_BLOCKS);
  const int nbr = (int)ceil_div(n, rpb);
  hipLaunchKernelGGL(grad_colsum_kernel, dim3(nbr), dim3(256), 0, s, X, n, (int)d, rpb, part);
  TSVM_CHECK(sum_parts(c, part, nbr, (int)d, mu));
  std::vector<float> hes(2 * d);
  for (int64_t f = 0; f < d; f++) {
    const double e = exp(log_scales[f]);
    hes[f] = (float)e;
    hes[d + f] = (float)(mu[f] / (double)n * e);
  }
  TSVM_HIP(c, hipMemcpyAsync(esmu, hes.data(), 2 * d * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(grad_scale_kernel, dim3(2048), dim3(256), 0, s, X, n, (int)d,
                     (const float*)esmu, Xs);
  if (nt > 0)
    hipLaunchKernelGGL(grad_scale_kernel, dim3(2048), dim3(256), 0, s, Xt, nt, (int)d,
                       (const float*)esmu, Xts);
  {
    std::vector<float> h1(nmax, 1.0f);
    TSVM_HIP(c, hipMemcpyAsync(ones, h1.data(), nmax * 4, hipMemcpyHostToDevice, s));
    TSVM_HIP(c, hipStreamSynchronize(s));
  }
  // ---- factor: Y = K_Y Omega, W = orth(Y), R = W^T Y, Q = K_Y W, C = W^T Q (the two
  // products in FP32, or in the fp16 product)
  const int fprec = precision == TSVM_PREC_FP32 ? TSVM_PREC_FP32 : TSVM_PREC_FP16;
  auto progress = [&](const char* what) -> tsvm_status {
    if (c->progress) {
      TSVM_HIP(c, hipStreamSynchronize(s));
      fprintf(stderr, "grad: %s (%.1f s)\n", what, now() - t0);
      fflush(stderr);
    }
    return TSVM_OK;
  };
  TSVM_CHECK(kernel_product_dev(c, Xs, y, n, Xs, y, n, d, Om, k, k, gamma, fprec, Yb, k));
  TSVM_CHECK(progress("factor product K_Y Omega"));
  TSVM_HIP(c, hipMemcpyAsync(Wb, Yb, n * k * 4, hipMemcpyDeviceToDevice, s));
  std::vector<double> G;
  int passes = 0;
  TSVM_CHECK(cholqr_orth(c, Wb, Qb, n, k, G, &passes));
  std::vector<double> R(k * k), C(k * k);
  TSVM_CHECK(gemm_tn_f64(c, Wb, k, k, Yb, k, k, n, nullptr, false, dkk));
  TSVM_HIP(c, hipMemcpyAsync(R.data(), dkk, k * k * 8, hipMemcpyDeviceToHost, s));
  TSVM_CHECK(kernel_product_dev(c, Xs, y, n, Xs, y, n, d, Wb, k, k, gamma, fprec, Qb, k));
  TSVM_CHECK(progress("factor product K_Y W"));
  TSVM_CHECK(gemm_tn_f64(c, Wb, k, k, Qb, k, k, n, nullptr, false, dkk));
  TSVM_HIP(c, hipMemcpyAsync(C.data(), dkk, k * k * 8, hipMemcpyDeviceToHost, s));
  TSVM_HIP(c, hipStreamSynchronize(s));
  for (int64_t a = 0; a < k; a++)
    for (int64_t bb = 0; bb < a; bb++) {
      const double v = 0.5 * (C[a * k + bb] + C[bb * k + a]);
      C[a * k + bb] = v;
      C[bb * k + a] = v;
    }
  {
    const int chol = dev_chol_inv_t(c, C.data(), k, 0.0, dkk);  // positive definiteness
    if (chol < 0) return fail(c, TSVM_ERR_HIP, "grad: rocSOLVER Cholesky failed");
    if (chol > 0)
      return fail(c, TSVM_ERR_NUMERIC,
                  "grad: W^T K_Y W is not positive definite (the factor clips eigenvalues)");
  }
  std::vector<double> beta(k);
  TSVM_CHECK(masked_colsum(c, Wb, n, k, k, alpha, -HUGE_VAL, beta.data()));
  std::vector<double> ha(n);
  TSVM_HIP(c, hipMemcpyAsync(ha.data(), alpha, n * 8, hipMemcpyDeviceToHost, s));
  TSVM_HIP(c, hipStreamSynchronize(s));
  double sa = 0.0, bcb = 0.0;
  for (int64_t i = 0; i < n; i++) sa += ha[i];
  for (int64_t a = 0; a < k; a++)
    for (int64_t bb = 0; bb < k; bb++) bcb += beta[a] * C[a * k + bb] * beta[bb];
  vals[0] = sa - 0.5 * bcb;
  vals[1] = 0.0;
  inf.t_forward = now() - t0;
  const dim3 rg(2048), rb(256);
  double* Pa = vec;
  double* gv = vec + n;
  double* u2 = vec + 2 * n;
  double* u3 = vec + 3 * n;
  double* v2 = vec + 4 * n;
  double* v3 = vec + 5 * n;
  double* qv = vec + 6 * n;
  Event e0, e1;
  TSVM_HIP(c, e0.create());
  TSVM_HIP(c, e1.create());
  double flops = 0.0;
  float ms_bwd = 0.f, ms_dual = 0.f;
  if (gD) {
    // ---- the three bilinear pairs of dD
    TSVM_HIP(c, hipMemcpyAsync(dkv, beta.data(), k * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(grad_rowdot_kernel, rg, rb, 0, s, (const float*)Wb, n, (int)k,
                       (const double*)dkv, (const double*)nullptr, 1.0, Pa);  // P a = W beta
    hipLaunchKernelGGL(grad_rowdot_kernel, rg, rb, 0, s, (const float*)Qb, n, (int)k,
                       (const double*)dkv, (const double*)nullptr, 1.0, gv);  // g = Q beta
    hipLaunchKernelGGL(grad_rowdot_kernel, rg, rb, 0, s, (const float*)Wb, n, (int)k,
                       (const double*)dkv, alpha, -1.0, u2);  // (I - P) a
    std::vector<double> wg(k);
    TSVM_CHECK(masked_colsum(c, Wb, n, k, k, gv, -HUGE_VAL, wg.data()));
    std::vector<double> c1 = wg, c2 = beta;
    if (!solve_dense(R, k, c1.data()) || !solve_dense(R, k, c2.data()))
      return fail(c, TSVM_ERR_NUMERIC, "grad: the sketch K_Y Omega is rank deficient");
    TSVM_HIP(c, hipMemcpyAsync(dkv, wg.data(), k * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(grad_rowdot_kernel, rg, rb, 0, s, (const float*)Wb, n, (int)k,
                       (const double*)dkv, (const double*)gv, -1.0, u3);  // (I - P) g
    TSVM_HIP(c, hipMemcpyAsync(dkv, c1.data(), k * 8, hipMemcpyHostToDevice, s));
    TSVM_HIP(c, hipMemcpyAsync(dkv + k, c2.data(), k * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(grad_rowdot_kernel, rg, rb, 0, s, Om, n, (int)k, (const double*)dkv,
                       (const double*)nullptr, 1.0, v2);  // Omega Y^+ g
    hipLaunchKernelGGL(grad_rowdot_kernel, rg, rb, 0, s, Om, n, (int)k, (const double*)(dkv + k),
                       (const double*)nullptr, 1.0, v3);  // Omega Y^+ a
    // right-hand sides [q, q x, q x^2] of the three pairs side by side (n x rtr)
    void* pr;
    TSVM_HIP(c, c->ws.get("gr_rhs", (size_t)(n * rtr * 4), &pr));
    float* Rh = (float*)pr;
    void* pw;
    TSVM_HIP(c, c->ws.get("gr_wd", (size_t)(nmax * rtr * 4), &pw));
    float* Wd = (float*)pw;
    const double* qs[3] = {Pa, v2, v3};
    const double* us[3] = {Pa, u2, u3};
    for (int m = 0; m < 3; m++)
      hipLaunchKernelGGL(grad_rhs_kernel, rg, rb, 0, s, (const float*)Xs, n, (int)d, qs[m], y, rtr,
                         (int64_t)m * (2 * d + 1), Rh);
    TSVM_HIP(c, hipEventRecord(e0, s));
    TSVM_CHECK(kernel_product_dev(c, Xs, ones, n, Xs, ones, n, d, Rh, rtr, rtr, gamma,
                                  TSVM_PREC_FP32, Wd, rtr));
    TSVM_CHECK(progress("backward product of the dual"));
    for (int m = 0; m < 3; m++)
      hipLaunchKernelGGL(grad_reduce_kernel, dim3(nbr), dim3(256), 0, s, (const float*)Xs, n,
                         (int)d, us[m], y, (const float*)Wd, rtr, (int64_t)m * (2 * d + 1), rpb,
                         part + (int64_t)m * GR_BLOCKS * d);
    TSVM_HIP(c, hipEventRecord(e1, s));
    std::vector<double> T(d, 0.0), S;
    for (int m = 0; m < 3; m++) {
      TSVM_CHECK(sum_parts(c, part + (int64_t)m * GR_BLOCKS * d, nbr, (int)d, S));
      for (int64_t f = 0; f < d; f++) T[f] += (m == 0 ? 1.0 : 2.0) * S[f];
    }
    float ms = 0.f;
    TSVM_HIP(c, hipEventElapsedTime(&ms, e0, e1));
    ms_bwd += ms;
    ms_dual = ms;
    flops += 2.0 * n * n * (d + rtr);
    double tot = 0.0;
    for (int64_t f = 0; f < d; f++) {
      gD[1 + f] = gamma * T[f];
      tot += T[f];
    }
    gD[0] = 0.5 * gamma * tot;
  }
  // ---- held-out smoothed hinge: f_t = (K q)_t + b with q = alpha y; one product gives f
  // and the derivative columns
  if (nt > 0) {
    void* pr;
    TSVM_HIP(c, c->ws.get("gr_rhs", (size_t)(n * rtr * 4), &pr));
    float* Rh = (float*)pr;
    void* pw;
    TSVM_HIP(c, c->ws.get("gr_wd", (size_t)(nmax * rtr * 4), &pw));
    float* Wd = (float*)pw;
    hipLaunchKernelGGL(grad_rhs_kernel, rg, rb, 0, s, (const float*)Xs, n, (int)d, alpha, y, rte,
                       (int64_t)0, Rh);
    // On the FP64 product: the per-feature sums below combine x_t^2 (Kq)_t,
    // -2 x_t (K(q x))_t and (K(q x^2))_t, which cancel to a small remainder, so FP32
    // accumulation (~1e-6 relative per column) left ~1e-3 relative error in dL/dlog gamma
    // at 200k x 50k rows, d = 32; FP64 accumulation leaves the final FP32 rounding.
    // Column 0 (q = alpha y) is the decision values behind L, with the same accuracy.
    TSVM_HIP(c, hipEventRecord(e0, s));
    TSVM_CHECK(kernel_product_dev(c, Xts, ones, nt, Xs, ones, n, d, Rh, rte, rte, gamma,
                                  TSVM_PREC_FP64, Wd, rte));
    TSVM_CHECK(progress("backward product of the held-out loss"));
    TSVM_HIP(c, hipEventRecord(e1, s));
    std::vector<float> hf(nt);
    TSVM_HIP(c, hipMemcpy2DAsync(hf.data(), 4, Wd, (size_t)rte * 4, 4, (size_t)nt,
                                 hipMemcpyDeviceToHost, s));
    std::vector<float> hy(nt);
    TSVM_HIP(c, hipMemcpyAsync(hy.data(), yt, nt * 4, hipMemcpyDeviceToHost, s));
    TSVM_HIP(c, hipStreamSynchronize(s));
    float ms = 0.f;
    TSVM_HIP(c, hipEventElapsedTime(&ms, e0, e1));
    std::vector<double> pt(nt);
    double loss = 0.0;
    for (int64_t t = 0; t < nt; t++) {
      const double z = (double)hy[t] * ((double)hf[t] + b);
      double h, dh;
      if (z <= 0.0) {
        h = 0.5 - z;
        dh = -1.0;
      } else if (z < 1.0) {
        h = 0.5 * (1.0 - z) * (1.0 - z);
        dh = -(1.0 - z);
      } else {
        h = 0.0;
        dh = 0.0;
      }
      loss += h;
      pt[t] = dh * (double)hy[t] / (double)nt;
    }
    vals[1] = loss / (double)nt;
    if (gL) {
      TSVM_HIP(c, hipMemcpyAsync(qv, pt.data(), nt * 8, hipMemcpyHostToDevice, s));
      Event e2, e3;
      TSVM_HIP(c, e2.create());
      TSVM_HIP(c, e3.create());
      TSVM_HIP(c, hipEventRecord(e2, s));
      const int64_t rpt = ceil_div(nt, GR_BLOCKS);
      const int nbt = (int)ceil_div(nt, rpt);
      hipLaunchKernelGGL(grad_reduce_kernel, dim3(nbt), dim3(256), 0, s, (const float*)Xts, nt,
                         (int)d, (const double*)qv, (const float*)nullptr, (const float*)Wd, rte,
                         (int64_t)0, rpt, part);
      TSVM_HIP(c, hipEventRecord(e3, s));
      std::vector<double> S;
      TSVM_CHECK(sum_parts(c, part, nbt, (int)d, S));
      float ms2 = 0.f;
      TSVM_HIP(c, hipEventElapsedTime(&ms2, e2, e3));
      double tot = 0.0;
      for (int64_t f = 0; f < d; f++) {
        gL[1 + f] = -2.0 * gamma * S[f];
        tot += S[f];
      }
      gL[0] = -gamma * tot;
      ms_bwd += ms + ms2;
      inf.t_backward_heldout = (ms + ms2) * 1e-3;
    }
  } else if (gL) {
    for (int64_t f = 0; f <= d; f++) gL[f] = 0.0;
  }
  inf.realk = k;
  inf.orth_passes = passes;
  inf.t_backward = ms_bwd * 1e-3;
  inf.backward_flops = flops;
  if (ms_dual > 0) inf.backward_tflops = flops / (ms_dual * 1e-3) * 1e-12;
  inf.t_total = now() - t0;
  if (info) *info = inf;
  return TSVM_OK;
}

}  // namespace tsvm
