// Low-rank factor construction (replaces LRA, tensorsvm-train.cu:1941-2249).
//
// Everything of size n stays resident in HBM; only k x k matrices cross to the host.
//   1. W = K_Y Omega                      fused kernel-product kernel (FP32 MFMA; the BF16
//                                         form runs this sketch in the fp16 product, the
//                                         FP16X2 form in the two-term product: rounding
//                                         noise in the sketch adds non-kernel directions to
//                                         range(W), which the two-term form keeps ~2^-11
//                                         smaller)
//   2. W <- orth(W)                       shifted Cholesky-QR (Fukaya et al., SIAM J. Sci.
//                                         Comput. 2020): G = W^T W on FP64 MFMA, Cholesky
//                                         of G + s I and L^{-T} (rocSOLVER), W <- W L^{-T};
//                                         repeated until G = I to 1e-6 (at most 3 passes)
//   3. Gp = W^T W                         Gram of the FP32 basis as stored
//   4. Q = K_Y W                          second kernel product (BF16 form: the fp16 product;
//                                         the basis itself stays FP32 -- rounding it would
//                                         rotate span(W) by ~u and move the projected
//                                         kernel by ~u sigma_1, while the product's rounding
//                                         only perturbs C inside span(W), where the
//                                         trailing block sees ~u sigma_tail)
//   5. Cb = W^T Q (FP64), symmetrised     (the reference's C, 2122-2156)
//   6. k x k on the GPU (rocSOLVER / rocBLAS): Gp = L L^T, C = L^-1 Cb L^-T = V diag(w) V^T
//      (dsyevd), X = L^-T V sqrt(w+)
//      -- with Wo = W L^-T orthonormal this is exactly the reference's C = Wo^T K_Y Wo,
//      so U U^T = Wo C+ Wo^T is the projected kernel even where W is only nearly
//      orthonormal
//   7. U = W X                            FP64 MFMA, FP32 output (the reference's U = W X)
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <vector>

#include "internal.h"

namespace tsvm {

namespace {
double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

// Shifted Cholesky-QR of the n x k FP32 block W (in place; tmp is an n x k scratch):
// G = W^T W on FP64 MFMA, Cholesky of G + s I and L^{-T} on the GPU, W <- W L^{-T},
// repeated until G = I to 1e-6 (at most 3 passes).  On return G holds the Gram of the
// basis as stored.
tsvm_status cholqr_orth(Ctx* c, float* W, float* tmp, int64_t n, int64_t k,
                        std::vector<double>& G, int* passes) {
  hipStream_t s = c->stream;
  void *pG, *pB;
  TSVM_HIP(c, c->ws.get("qr_G", (size_t)(k * k * 8), &pG));
  TSVM_HIP(c, c->ws.get("qr_B", (size_t)(k * k * 8), &pB));
  double* dG = (double*)pG;
  double* dB = (double*)pB;
  G.assign(k * k, 0.0);
  float* cur = W;
  float* oth = tmp;
  bool have_gram = false;
  *passes = 0;
  for (int pass = 0; pass < 3; pass++) {
    TSVM_CHECK(gemm_tn_f64(c, cur, k, k, cur, k, k, n, nullptr, true, dG));
    TSVM_HIP(c, hipMemcpyAsync(G.data(), dG, k * k * 8, hipMemcpyDeviceToHost, s));
    TSVM_HIP(c, hipStreamSynchronize(s));
    double dev = 0.0, tr = 0.0;
    for (int64_t i = 0; i < k; i++) {
      tr += G[i * k + i];
      for (int64_t j = 0; j < k; j++) dev = fmax(dev, fabs(G[i * k + j] - (i == j ? 1.0 : 0.0)));
    }
    if (!(tr > 0.0) || !std::isfinite(tr))
      return fail(c, TSVM_ERR_NUMERIC, "cholqr: block has norm %g", tr);
    have_gram = true;
    if (dev < 1e-6) break;  // orthonormal to FP32 storage accuracy
    *passes = pass + 1;
    // shift (pass 0 only) keeps the Cholesky of an ill-conditioned Gram defined
    double shift = pass == 0 ? 11.0 * ((double)n * k + (double)k * (k + 1)) * 1.11e-16 * tr : 0.0;
    int fails = 0;
    for (;;) {  // dB = L^{-T}
      const int r = dev_chol_inv_t(c, G.data(), k, shift, dB);
      if (r < 0) return fail(c, TSVM_ERR_HIP, "cholqr: rocSOLVER Cholesky / inverse failed");
      if (r == 0) break;
      shift = shift > 0 ? shift * 10.0 : 1e-14 * tr;
      if (++fails > 30) return fail(c, TSVM_ERR_NUMERIC, "cholqr: Cholesky failed");
    }
    TSVM_CHECK(gemm_nn_tall(c, cur, k, k, dB, k, n, oth, k));
    std::swap(cur, oth);
    have_gram = false;
  }
  if (cur != W) TSVM_HIP(c, hipMemcpyAsync(W, cur, n * k * 4, hipMemcpyDeviceToDevice, s));
  if (!have_gram) {
    TSVM_CHECK(gemm_tn_f64(c, W, k, k, W, k, k, n, nullptr, true, dG));
    TSVM_HIP(c, hipMemcpyAsync(G.data(), dG, k * k * 8, hipMemcpyDeviceToHost, s));
  }
  TSVM_HIP(c, hipStreamSynchronize(s));
  return TSVM_OK;
}

// W = K_Y A over all n rows.  With progress lines on, the rows go in slabs of about
// 2e15 flops (multiples of 256 rows), one stderr line per slab; each row's result does
// not depend on the slabbing (the FP32 product's column segments follow n, not the slab).
static tsvm_status product_rows(Ctx* c, const char* name, const float* X, const float* y,
                                int64_t n, int64_t d, const float* A, int64_t k, double gamma,
                                int prec, float* W) {
  int64_t slabs = 1;
  if (c->progress) {
    slabs = (int64_t)ceil(2.0 * (double)n * (double)n * (double)(d + k) / 2e15);
    if (slabs < 1) slabs = 1;
  }
  const int64_t rows = ((n + slabs - 1) / slabs + 255) / 256 * 256;
  const double t0 = now();
  for (int64_t r0 = 0; r0 < n; r0 += rows) {
    const int64_t m = rows < n - r0 ? rows : n - r0;
    TSVM_CHECK(kernel_product_dev(c, X + r0 * d, y + r0, m, X, y, n, d, A, k, k, gamma, prec,
                                  W + r0 * k, k));
    if (c->progress) {
      TSVM_HIP(c, hipStreamSynchronize(c->stream));
      fprintf(stderr, "%s: rows %ld / %ld (%.1f s)\n", name, (long)(r0 + m), (long)n, now() - t0);
      fflush(stderr);
    }
  }
  return TSVM_OK;
}

tsvm_status lra_dev(Ctx* c, const float* X, const float* y, int64_t n, int64_t d, int64_t k,
                    double gamma, int precision, float* Om, float* U, tsvm_lra_info* info) {
  if (n <= 0 || d <= 0 || k <= 0 || k > n)
    return fail(c, TSVM_ERR_ARG, "lra: need 0 < k <= n (n=%ld k=%ld)", (long)n, (long)k);
  if (precision != TSVM_PREC_FP32 && precision != TSVM_PREC_BF16 &&
      precision != TSVM_PREC_FP16X2 && precision != TSVM_PREC_FP64 && precision != TSVM_PREC_FP8)
    return fail(c, TSVM_ERR_ARG, "lra: unknown precision %d", precision);
  if (precision == TSVM_PREC_FP8) precision = TSVM_PREC_BF16;  // its factor is the BF16 form's
  hipStream_t s = c->stream;
  tsvm_lra_info inf;
  memset(&inf, 0, sizeof(inf));
  const double t0 = now();
  void *pW, *pG, *pB;
  TSVM_HIP(c, c->ws.get("lra_W", (size_t)(n * k * 4), &pW));
  TSVM_HIP(c, c->ws.get("lra_G", (size_t)(k * k * 8), &pG));
  TSVM_HIP(c, c->ws.get("lra_B", (size_t)(k * k * 8), &pB));
  float* W = (float*)pW;
  double* dG = (double*)pG;
  double* dB = (double*)pB;
  std::vector<double> G(k * k), T(k * k);

  // 1. W = K_Y Omega (timed with events on the stream it runs on)
  Event e0, e1;
  TSVM_HIP(c, e0.create());
  TSVM_HIP(c, e1.create());
  TSVM_HIP(c, hipEventRecord(e0, s));
  // low-precision forms: the sketch in fp16 (BF16 form) or in the two-term product
  const int sketch_prec = precision == TSVM_PREC_BF16 ? TSVM_PREC_FP16 : precision;
  TSVM_CHECK(product_rows(c, "factor: K_Y Omega", X, y, n, d, Om, k, gamma, sketch_prec, W));
  TSVM_HIP(c, hipEventRecord(e1, s));
  TSVM_HIP(c, hipStreamSynchronize(s));
  float ms = 0.f;
  TSVM_HIP(c, hipEventElapsedTime(&ms, e0, e1));
  double t1 = now();
  inf.t_product1 = ms * 1e-3;

  // 2-3. shifted Cholesky-QR (U is the scratch basis); G = Gram of the basis as stored
  TSVM_CHECK(cholqr_orth(c, W, U, n, k, G, &inf.orth_passes));
  TSVM_HIP(c, hipStreamSynchronize(s));
  double t2 = now();
  inf.t_orth = t2 - t1;

  // 4. Q = K_Y W (into the Omega buffer)
  float* Q = Om;
  TSVM_HIP(c, hipEventRecord(e0, s));
  const int q_prec = precision == TSVM_PREC_FP32 || precision == TSVM_PREC_FP64 ? precision
                                                                               : TSVM_PREC_FP16;
  TSVM_CHECK(product_rows(c, "factor: K_Y W", X, y, n, d, W, k, gamma, q_prec, Q));
  TSVM_HIP(c, hipEventRecord(e1, s));
  TSVM_HIP(c, hipStreamSynchronize(s));
  TSVM_HIP(c, hipEventElapsedTime(&ms, e0, e1));
  double t3 = now();
  inf.t_product2 = ms * 1e-3;

  // 5. Cb = W^T Q
  std::vector<double> Cb(k * k), Cs(k * k), w(k), V(k * k);
  TSVM_CHECK(gemm_tn_f64(c, W, k, k, Q, k, k, n, nullptr, false, dG));
  TSVM_HIP(c, hipMemcpyAsync(Cb.data(), dG, k * k * 8, hipMemcpyDeviceToHost, s));
  TSVM_HIP(c, hipStreamSynchronize(s));
  const double tc0 = now();
  for (int64_t i = 0; i < k; i++)
    for (int64_t j = 0; j < k; j++) Cs[i * k + j] = 0.5 * (Cb[i * k + j] + Cb[j * k + i]);
  // 6. C = L^-1 Cs L^-T, eig, X = L^-T V sqrt(w+) (the k x k products on the GPU: dLt =
  // L^-T, dA and dT scratch, X into dB)
  void *pLt, *pA, *pT;
  TSVM_HIP(c, c->ws.get("lra_Lt", (size_t)(k * k * 8), &pLt));
  TSVM_HIP(c, c->ws.get("lra_A", (size_t)(k * k * 8), &pA));
  TSVM_HIP(c, c->ws.get("lra_T", (size_t)(k * k * 8), &pT));
  double* dLt = (double*)pLt;
  double* dA = (double*)pA;
  double* dT = (double*)pT;
  double tc1, tc2;
  int64_t realk = k;
  const int chol = dev_chol_inv_t(c, G.data(), k, 0.0, dLt);
  if (chol < 0) return fail(c, TSVM_ERR_HIP, "lra: rocSOLVER Cholesky / inverse failed");
  if (chol == 0) {
    TSVM_HIP(c, hipMemcpyAsync(dA, Cs.data(), k * k * 8, hipMemcpyHostToDevice, s));
    TSVM_CHECK(dev_gemm_rm(c, true, false, dLt, dA, k, dT));   // T = L^-1 Cs
    TSVM_CHECK(dev_gemm_rm(c, false, false, dT, dLt, k, dA));  // L^-1 Cs L^-T
    TSVM_HIP(c, hipMemcpyAsync(Cb.data(), dA, k * k * 8, hipMemcpyDeviceToHost, s));
    TSVM_HIP(c, hipStreamSynchronize(s));
    for (int64_t i = 0; i < k; i++)
      for (int64_t j = 0; j < i; j++) {
        double v = 0.5 * (Cb[i * k + j] + Cb[j * k + i]);
        Cb[i * k + j] = v;
        Cb[j * k + i] = v;
      }
    tc1 = now();
    TSVM_CHECK(host_sym_eig(c, Cb.data(), k, w.data(), V.data()));
    tc2 = now();
    for (int64_t j = 0; j < k; j++) {
      double sc = 0.0;
      if (w[j] > 0) sc = sqrt(w[j]);
      else realk--;
      for (int64_t i = 0; i < k; i++) V[i * k + j] *= sc;
    }
    TSVM_HIP(c, hipMemcpyAsync(dA, V.data(), k * k * 8, hipMemcpyHostToDevice, s));
    TSVM_CHECK(dev_gemm_rm(c, false, false, dLt, dA, k, dB));  // X = L^-T V sqrt(w+)
  } else {
    // Rank-deficient sketch (gamma = 0 gives K = 1 1^T; duplicated rows): range(K_Y Omega)
    // has dimension r < k and the basis Gram is singular.  With G = Vg diag(g) Vg^T and the
    // r directions g_i > 1e-8 max g kept, B = Vg_r diag(g_r^-1/2) makes W B an orthonormal
    // basis of range(W); C_r = B^T Cs B (r x r) = Vc diag(w) Vc^T and X = B Vc sqrt(w+),
    // padded with k - r zero columns (zero eigenvalues, clipped like the reference's
    // non-positive ones, 2199-2210).  U U^T is again the projected kernel P K_Y P.
    std::vector<double> gw(k), gv(k * k);
    TSVM_CHECK(host_sym_eig(c, G.data(), k, gw.data(), gv.data()));
    const double gmax = gw[k - 1];
    if (!(gmax > 0)) return fail(c, TSVM_ERR_NUMERIC, "lra: the sketch K_Y Omega is zero");
    int64_t r0 = 0;  // first kept direction (ascending eigenvalues)
    while (r0 < k && !(gw[r0] > 1e-8 * gmax)) r0++;
    const int64_t r = k - r0;
    std::vector<double> B(k * r), Cr(r * r), Tr(k * r), wr(r), Vr(r * r);
    for (int64_t i = 0; i < k; i++)
      for (int64_t j = 0; j < r; j++) B[i * r + j] = gv[i * k + r0 + j] / sqrt(gw[r0 + j]);
    for (int64_t i = 0; i < k; i++)  // Tr = Cs B (k x r)
      for (int64_t j = 0; j < r; j++) {
        double acc = 0.0;
        for (int64_t t = 0; t < k; t++) acc += Cs[i * k + t] * B[t * r + j];
        Tr[i * r + j] = acc;
      }
    for (int64_t i = 0; i < r; i++)  // Cr = B^T Tr, symmetrised
      for (int64_t j = 0; j < r; j++) {
        double acc = 0.0;
        for (int64_t t = 0; t < k; t++) acc += B[t * r + i] * Tr[t * r + j];
        Cr[i * r + j] = acc;
      }
    for (int64_t i = 0; i < r; i++)
      for (int64_t j = 0; j < i; j++) {
        double v = 0.5 * (Cr[i * r + j] + Cr[j * r + i]);
        Cr[i * r + j] = v;
        Cr[j * r + i] = v;
      }
    tc1 = now();
    TSVM_CHECK(host_sym_eig(c, Cr.data(), r, wr.data(), Vr.data()));
    tc2 = now();
    realk = 0;
    for (int64_t j = 0; j < k; j++) w[j] = j < r0 ? 0.0 : wr[j - r0];
    for (int64_t j = 0; j < r; j++) {
      const double sc = wr[j] > 0 ? sqrt(wr[j]) : 0.0;
      if (wr[j] > 0) realk++;
      for (int64_t i = 0; i < r; i++) Vr[i * r + j] *= sc;
    }
    for (int64_t i = 0; i < k; i++)  // X = B Vr sqrt(w+), columns r0.. (ascending order)
      for (int64_t j = 0; j < k; j++) {
        double acc = 0.0;
        if (j >= r0)
          for (int64_t t = 0; t < r; t++) acc += B[i * r + t] * Vr[t * r + (j - r0)];
        T[i * k + j] = acc;
      }
    TSVM_HIP(c, hipMemcpyAsync(dB, T.data(), k * k * 8, hipMemcpyHostToDevice, s));
  }
  inf.realk = realk;
  inf.eig_max = w[k - 1];
  inf.eig_min = w[0];
  double t4 = now();
  inf.t_core = t4 - t3;
  if (c->progress) {
    fprintf(stderr,
            "factor: core %.3f s (C = W^T Q %.3f s, projection %.3f s, eig %.3f s, right "
            "factor %.3f s)\n",
            t4 - t3, tc0 - t3, tc1 - tc0, tc2 - tc1, t4 - tc2);
    fflush(stderr);
  }

  // 7. U = W X
  TSVM_CHECK(gemm_nn_tall(c, W, k, k, dB, k, n, U, k));
  TSVM_HIP(c, hipStreamSynchronize(s));
  double t5 = now();
  inf.t_factor = t5 - t4;
  inf.t_total = t5 - t0;
  if (info) *info = inf;
  return TSVM_OK;
}

}  // namespace tsvm

using tsvm::Ctx;

extern "C" {

tsvm_status tsvm_lra_dev(tsvm_ctx* ctx, const float* X, const float* y, int64_t n, int64_t d,
                         int64_t k, double gamma, int precision, float* omega, float* U,
                         tsvm_lra_info* info) {
  if (!ctx || !X || !y || !omega || !U) return TSVM_ERR_ARG;
  Ctx* c = &ctx->c;
  hipSetDevice(c->device);
  return tsvm::lra_dev(c, X, y, n, d, k, gamma, precision, omega, U, info);
}

tsvm_status tsvm_lra(tsvm_ctx* ctx, const float* X, const float* y, int64_t n, int64_t d,
                     int64_t k, double gamma, int precision, const float* omega, float* U,
                     tsvm_lra_info* info) {
  if (!ctx || !X || !y || !U || n <= 0 || d <= 0 || k <= 0) return TSVM_ERR_ARG;
  Ctx* c = &ctx->c;
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  void *dX, *dy, *dOm, *dU;
  TSVM_HIP(c, c->ws.get("h_lraX", n * d * 4, &dX));
  TSVM_HIP(c, c->ws.get("h_lray", n * 4, &dy));
  TSVM_HIP(c, c->ws.get("h_lraOm", n * k * 4, &dOm));
  TSVM_HIP(c, c->ws.get("h_lraU", n * k * 4, &dU));
  TSVM_HIP(c, hipMemcpyAsync(dX, X, n * d * 4, hipMemcpyHostToDevice, s));
  TSVM_HIP(c, hipMemcpyAsync(dy, y, n * 4, hipMemcpyHostToDevice, s));
  if (omega) {
    TSVM_HIP(c, hipMemcpyAsync(dOm, omega, n * k * 4, hipMemcpyHostToDevice, s));
    TSVM_HIP(c, hipStreamSynchronize(s));
  } else {
    std::vector<float> om(n * k);
    TSVM_CHECK(tsvm_gaussrand_fill(om.data(), n * k, 0));
    TSVM_HIP(c, hipMemcpyAsync(dOm, om.data(), n * k * 4, hipMemcpyHostToDevice, s));
    TSVM_HIP(c, hipStreamSynchronize(s));
  }
  TSVM_CHECK(tsvm::lra_dev(c, (float*)dX, (float*)dy, n, d, k, gamma, precision, (float*)dOm,
                           (float*)dU, info));
  TSVM_HIP(c, hipMemcpyAsync(U, dU, n * k * 4, hipMemcpyDeviceToHost, s));
  TSVM_HIP(c, hipStreamSynchronize(s));
  return TSVM_OK;
}

tsvm_status tsvm_lra_f64(tsvm_ctx* ctx, const float* X, int64_t ldx, const float* y, int64_t n,
                         int64_t d, int64_t k, double gamma, int precision, const float* omega,
                         double* U, int64_t ldu, double* l1, tsvm_lra_info* info) {
  if (!ctx || !X || !y || !U || n <= 0 || d <= 0 || k <= 0 || ldx < d || ldu < k)
    return TSVM_ERR_ARG;
  std::vector<float> packed;
  const float* Xp = X;
  if (ldx != d) {
    packed.resize(n * d);
    for (int64_t i = 0; i < n; i++)
      std::copy(X + i * ldx, X + i * ldx + d, packed.begin() + i * d);
    Xp = packed.data();
  }
  std::vector<float> Uf(n * k);
  tsvm_lra_info inf;
  TSVM_CHECK(tsvm_lra(ctx, Xp, y, n, d, k, gamma, precision, omega, Uf.data(), &inf));
  for (int64_t i = 0; i < n; i++)
    for (int64_t j = 0; j < k; j++) U[i * ldu + j] = (double)Uf[i * k + j];
  if (l1) *l1 = inf.eig_max;
  if (info) *info = inf;
  return TSVM_OK;
}

}  // extern "C"
