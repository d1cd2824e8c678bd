// Whole training runs (main, tensorsvm-train.cu:858-1059) and the decision function.
//
// RBF (T=2):  U = LRA(X) on the device, alpha = IPM(U, y) on the device, then the bias
//             of writemodel (353-376): b = mean over the first <= 50 free SVs j of
//             y_j - y_j U_j . v with v = sum_{alpha_i > 1e-6} alpha_i U_i (one pass over U)
// Linear (T=0): Z = diag(y) X, IPM with the linear initial point, w = sum alpha_i y_i x_i,
//             b = mean over the first <= 5000 free SVs of y_j - w . x_j (301-331)
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <chrono>
#include <vector>

#include "internal.h"

namespace tsvm {

namespace {
double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void to_ipm_params(const tsvm_train_params* p, tsvm_ipm_params* q) {
  tsvm_ipm_default_params(q);
  q->C = p->C;
  q->kernel_type = p->kernel_type == 0 ? 0 : 2;
  q->maxit = p->maxit > 0 ? p->maxit : 100;
  q->stop_mode = p->stop_mode;
  q->tol = p->tol;
  q->precision = p->precision;
  q->refine = p->refine;
  q->verbose = p->verbose;
  q->init = p->init;
}
}  // namespace

// RBF branch with every n-sized array on the device.  U (n x k) is left in the workspace
// buffer "tr_U".
tsvm_status train_rbf_dev(Ctx* c, const float* X, const float* y, int64_t n, int64_t d,
                          const tsvm_train_params* prm, float* Om, double* alpha,
                          tsvm_train_info* info) {
  const double t0 = now();
  const int64_t k = prm->k;
  if (k <= 0 || k > n || k > TSVM_MAX_RANK)
    return fail(c, TSVM_ERR_ARG, "train: need 0 < k <= min(n, %d) (k=%ld, n=%ld)", TSVM_MAX_RANK,
                (long)k, (long)n);
  struct ProgressScope {  // progress lines for the duration of this call
    Ctx* c;
    int saved;
    ProgressScope(Ctx* cc, int p) : c(cc), saved(cc->progress) { c->progress = p || saved; }
    ~ProgressScope() { c->progress = saved; }
  } progress_scope(c, prm->progress);
  tsvm_train_info inf;
  memset(&inf, 0, sizeof(inf));
  void *pU, *pxi, *ps;
  TSVM_HIP(c, c->ws.get("tr_U", (size_t)(n * k * 4), &pU));
  TSVM_HIP(c, c->ws.get("tr_xi", (size_t)(n * 8), &pxi));
  TSVM_HIP(c, c->ws.get("tr_s", (size_t)(n * 8), &ps));
  float* U = (float*)pU;
  if (prm->verbose) printf("\x1b[34mRBF kernel: gamma=%.3e, C=%.3e Approximation Rank K=%ld\x1b[39m\n",
                           prm->gamma, prm->C, (long)k);
  TSVM_CHECK(lra_dev(c, X, y, n, d, k, prm->gamma, prm->precision, Om, U, &inf.lra));
  if (prm->verbose) {
    printf("[LRA]: C: largest ew=%.3e, smallest ew=%.3e\n", inf.lra.eig_max, inf.lra.eig_min);
    printf("real rank is %ld\n", (long)inf.lra.realk);
    printf("\x1b[95mLRA elapsed time = %.3f seconds\x1b[39m\n", inf.lra.t_total);
  }
  if (prm->analysis) {  // 2328-2360
    const double ta = now();
    TSVM_CHECK(fnorm_res_dev(c, X, y, n, d, prm->gamma, U, k, k, &inf.fnorm, &inf.fnorm_res));
    inf.t_analysis = now() - ta;
    if (prm->verbose) {
      printf("fnorm(K)=%.3e fnorm(K-U*U=%.3e')\n", inf.fnorm, inf.fnorm_res);
      printf("rbf_fnorm_res took %.3f seconds\n", inf.t_analysis);
    }
  }
  tsvm_ipm_params ip;
  to_ipm_params(prm, &ip);
  TSVM_CHECK(ipm_dev(c, U, y, n, k, &ip, alpha, (double*)pxi, (double*)ps, &inf.ipm));
  if (prm->verbose) printf("\x1b[95mMPC elapsed time = %.3f seconds\x1b[39m\n", inf.ipm.t_total);
  // bias (writemodel 353-376).  The samples are the first <= 50 free SVs
  // (1e-6 < alpha < C - 1e-6), the reference's set, under the reference stop rule.  Under
  // the relative-gap stop the multipliers of non-SVs have not yet vanished (~mu/s, above
  // 1e-6 at c2), so there a free SV must also have IPM slacks below its distances to the
  // bounds (s < alpha, xi < C - alpha); when no multiplier passes that test the
  // reference's set is taken, so the bias is NaN only without free SVs, as in the
  // reference (372).
  const double tb = now();
  std::vector<double> ha(n), hs(n), hxi(n);
  TSVM_HIP(c, hipMemcpyAsync(ha.data(), alpha, n * 8, hipMemcpyDeviceToHost, c->stream));
  TSVM_HIP(c, hipMemcpyAsync(hs.data(), ps, n * 8, hipMemcpyDeviceToHost, c->stream));
  TSVM_HIP(c, hipMemcpyAsync(hxi.data(), pxi, n * 8, hipMemcpyDeviceToHost, c->stream));
  TSVM_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<int64_t> freeidx, freeref;
  const bool slack_rule = prm->stop_mode != TSVM_STOP_REF;
  for (int64_t i = 0; i < n; i++)
    if (ha[i] > 1e-6) {
      inf.nsv++;
      if (ha[i] < prm->C - 1e-6) {
        inf.nfree++;
        if (freeref.size() < 50) freeref.push_back(i);
        if (slack_rule && freeidx.size() < 50 && hs[i] < ha[i] && hxi[i] < prm->C - ha[i])
          freeidx.push_back(i);
      }
    }
  inf.b_slack_rule = slack_rule && !freeidx.empty() ? 1 : 0;
  if (!inf.b_slack_rule) freeidx.swap(freeref);
  std::vector<double> v(k);
  TSVM_CHECK(masked_colsum(c, U, n, k, k, alpha, 1e-6, v.data()));
  std::vector<float> urow(k);
  double acc = 0.0;
  std::vector<double> bs;
  for (int64_t j : freeidx) {
    float yj;
    TSVM_HIP(c, hipMemcpy(urow.data(), U + j * k, k * 4, hipMemcpyDeviceToHost));
    TSVM_HIP(c, hipMemcpy(&yj, y + j, 4, hipMemcpyDeviceToHost));
    double dot = 0.0;
    for (int64_t t = 0; t < k; t++) dot += (double)urow[t] * v[t];
    double bj = (double)yj - (double)yj * dot;
    acc += bj;
    bs.push_back(bj);
  }
  const double nb = (double)bs.size();
  inf.b = acc / nb;  // NaN without free SVs, as in the reference (372)
  double sq = 0.0;
  for (double bj : bs) sq += (bj - inf.b) * (bj - inf.b);
  inf.b_std = sqrt(sq / nb);
  inf.b_samples = (int)bs.size();
  inf.t_bias = now() - tb;
  inf.t_total = now() - t0 - inf.t_analysis;
  if (prm->verbose) {
    printf("#BSV %ld, #SV %ld\n", (long)inf.nfree, (long)inf.nsv);
    printf("approx mean b=%.6e std b=%.6e, #samples=%d\n ", inf.b, inf.b_std, inf.b_samples);
  }
  if (info) *info = inf;
  return TSVM_OK;
}

}  // namespace tsvm

using tsvm::Ctx;

extern "C" {

void tsvm_train_default_params(tsvm_train_params* p) {
  if (!p) return;
  p->k = 50;
  p->C = 1.0;
  p->gamma = 0.0;
  p->kernel_type = 2;
  p->precision = TSVM_PREC_FP32;
  p->stop_mode = TSVM_STOP_REF;
  p->tol = 1e-5;
  p->maxit = 100;
  p->refine = -1;
  p->verbose = 0;
  p->init = 0;
  p->analysis = 0;
  p->progress = 0;
}

tsvm_status tsvm_train_dev(tsvm_ctx* ctx, const float* X, const float* y, int64_t n, int64_t d,
                           const tsvm_train_params* prm, float* omega, double* alpha,
                           tsvm_train_info* info) {
  if (!ctx || !X || !y || !prm || !omega || !alpha || n <= 0 || d <= 0) return TSVM_ERR_ARG;
  if (prm->kernel_type != 2)
    return tsvm::fail(&ctx->c, TSVM_ERR_ARG, "tsvm_train_dev: RBF (kernel_type 2) only");
  (void)hipSetDevice(ctx->c.device);
  return tsvm::train_rbf_dev(&ctx->c, X, y, n, d, prm, omega, alpha, info);
}

tsvm_status tsvm_train(tsvm_ctx* ctx, const float* X, const float* y, int64_t n, int64_t d,
                       const tsvm_train_params* prm, const float* omega, double* alpha,
                       double* w_out, tsvm_train_info* info) {
  if (!ctx || !X || !y || !prm || !alpha || n <= 0 || d <= 0) return TSVM_ERR_ARG;
  Ctx* c = &ctx->c;
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  const double t0 = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  void *dX, *dy, *dal;
  TSVM_HIP(c, c->ws.get("h_trX", n * d * 4, &dX));
  TSVM_HIP(c, c->ws.get("h_try", n * 4, &dy));
  TSVM_HIP(c, c->ws.get("h_tral", n * 8, &dal));
  TSVM_HIP(c, hipMemcpyAsync(dy, y, n * 4, hipMemcpyHostToDevice, s));
  if (prm->kernel_type == 0) {
    // linear branch (907-953): Z = diag(y) X, k = d
    if (d > TSVM_MAX_RANK)
      return tsvm::fail(c, TSVM_ERR_ARG, "train: the linear branch needs d <= %d (d=%ld)",
                        TSVM_MAX_RANK, (long)d);
    std::vector<float> Z(n * d);
    for (int64_t i = 0; i < n; i++)
      for (int64_t t = 0; t < d; t++) Z[i * d + t] = y[i] * X[i * d + t];
    TSVM_HIP(c, hipMemcpyAsync(dX, Z.data(), n * d * 4, hipMemcpyHostToDevice, s));
    tsvm_ipm_params ip;
    tsvm_ipm_default_params(&ip);
    ip.C = prm->C;
    ip.kernel_type = 0;
    ip.maxit = prm->maxit > 0 ? prm->maxit : 100;
    ip.stop_mode = prm->stop_mode;
    ip.tol = prm->tol;
    ip.precision = prm->precision;
    ip.refine = prm->refine;
    ip.verbose = prm->verbose;
    ip.init = prm->init;
    void* pxi;
    TSVM_HIP(c, c->ws.get("tr_xi", (size_t)(n * 8), &pxi));
    tsvm_train_info inf;
    memset(&inf, 0, sizeof(inf));
    if (prm->verbose) printf("Linear SVM\n");
    TSVM_CHECK(tsvm::ipm_dev(c, (float*)dX, (float*)dy, n, d, &ip, (double*)dal, (double*)pxi, nullptr,
                             &inf.ipm));
    TSVM_HIP(c, hipMemcpyAsync(alpha, dal, n * 8, hipMemcpyDeviceToHost, s));
    TSVM_HIP(c, hipStreamSynchronize(s));
    // w and b (301-331)
    std::vector<double> w(d, 0.0);
    for (int64_t i = 0; i < n; i++)
      if (alpha[i] > 1e-6) {
        inf.nsv++;
        if (alpha[i] < prm->C - 1e-6) inf.nfree++;
        for (int64_t t = 0; t < d; t++) w[t] += alpha[i] * y[i] * X[i * d + t];
      }
    double sum = 0.0;
    std::vector<double> bs;
    for (int64_t i = 0; i < n && bs.size() < 5000; i++)
      if (alpha[i] > 1e-6 && alpha[i] < prm->C - 1e-6) {
        double tmp = y[i];
        for (int64_t t = 0; t < d; t++) tmp -= w[t] * X[i * d + t];
        bs.push_back(tmp);
        sum += tmp;
      }
    if (!bs.empty()) {
      inf.b = sum / (double)bs.size();
      double sq = 0.0;
      for (double b : bs) sq += (b - inf.b) * (b - inf.b);
      inf.b_std = sqrt(sq / (double)bs.size());
    } else {
      inf.b = 0.0;  // "Empty boundary SV! Give up." (329-331)
    }
    inf.b_samples = (int)bs.size();
    if (w_out) memcpy(w_out, w.data(), sizeof(double) * d);
    inf.t_total = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count() - t0;
    if (info) *info = inf;
    return TSVM_OK;
  }
  if (prm->kernel_type != 2) return tsvm::fail(c, TSVM_ERR_ARG, "unimplemented -t %d", prm->kernel_type);
  const int64_t k = prm->k;
  if (k <= 0 || k > n || k > TSVM_MAX_RANK)
    return tsvm::fail(c, TSVM_ERR_ARG, "train: need 0 < k <= min(n, %d) (k=%ld, n=%ld)",
                      TSVM_MAX_RANK, (long)k, (long)n);
  void* dOm;
  TSVM_HIP(c, c->ws.get("h_trOm", n * k * 4, &dOm));
  TSVM_HIP(c, hipMemcpyAsync(dX, X, n * d * 4, hipMemcpyHostToDevice, s));
  if (omega) {
    TSVM_HIP(c, hipMemcpyAsync(dOm, omega, n * k * 4, hipMemcpyHostToDevice, s));
    TSVM_HIP(c, hipStreamSynchronize(s));
  } else {
    std::vector<float> om(n * k);
    TSVM_CHECK(tsvm_gaussrand_fill(om.data(), n * k, 0));
    TSVM_HIP(c, hipMemcpyAsync(dOm, om.data(), n * k * 4, hipMemcpyHostToDevice, s));
    TSVM_HIP(c, hipStreamSynchronize(s));
  }
  TSVM_CHECK(tsvm::train_rbf_dev(c, (float*)dX, (float*)dy, n, d, prm, (float*)dOm, (double*)dal, info));
  TSVM_HIP(c, hipMemcpyAsync(alpha, dal, n * 8, hipMemcpyDeviceToHost, s));
  TSVM_HIP(c, hipStreamSynchronize(s));
  return TSVM_OK;
}

tsvm_status tsvm_decision_function(tsvm_ctx* ctx, const float* X, const float* y,
                                   const double* alpha, int64_t n, int64_t d, double gamma,
                                   double b, const float* T, int64_t m, double* f) {
  if (!ctx || !X || !y || !alpha || !T || !f || n < 0 || d <= 0 || m < 0) return TSVM_ERR_ARG;
  if (m == 0) return TSVM_OK;
  Ctx* c = &ctx->c;
  (void)hipSetDevice(c->device);
  hipStream_t s = c->stream;
  // support vectors only; coefficient alpha_i (y_i enters through the labelled product)
  std::vector<float> xs, ys, as;
  for (int64_t i = 0; i < n; i++)
    if (alpha[i] > 1e-6) {
      xs.insert(xs.end(), X + i * d, X + (i + 1) * d);
      ys.push_back(y[i]);
      as.push_back((float)alpha[i]);
    }
  const int64_t ns = (int64_t)ys.size();
  std::vector<float> ones(m, 1.0f), out(m, 0.0f);
  void *dS, *dys, *das, *dT, *d1, *dW;
  TSVM_HIP(c, c->ws.get("df_S", (ns ? ns : 1) * d * 4, &dS));
  TSVM_HIP(c, c->ws.get("df_y", (ns ? ns : 1) * 4, &dys));
  TSVM_HIP(c, c->ws.get("df_a", (ns ? ns : 1) * 4, &das));
  TSVM_HIP(c, c->ws.get("df_T", m * d * 4, &dT));
  TSVM_HIP(c, c->ws.get("df_1", m * 4, &d1));
  TSVM_HIP(c, c->ws.get("df_W", m * 4, &dW));
  if (ns > 0) {
    TSVM_HIP(c, hipMemcpyAsync(dS, xs.data(), ns * d * 4, hipMemcpyHostToDevice, s));
    TSVM_HIP(c, hipMemcpyAsync(dys, ys.data(), ns * 4, hipMemcpyHostToDevice, s));
    TSVM_HIP(c, hipMemcpyAsync(das, as.data(), ns * 4, hipMemcpyHostToDevice, s));
  }
  TSVM_HIP(c, hipMemcpyAsync(dT, T, m * d * 4, hipMemcpyHostToDevice, s));
  TSVM_HIP(c, hipMemcpyAsync(d1, ones.data(), m * 4, hipMemcpyHostToDevice, s));
  // FP32 form: the decision values feed a sign and an accuracy, no low precision here
  TSVM_CHECK(tsvm::kernel_product_dev(c, (float*)dT, (float*)d1, m, (float*)dS, (float*)dys, ns, d,
                                      (float*)das, 1, 1, gamma, TSVM_PREC_FP32, (float*)dW, 1));
  TSVM_HIP(c, hipMemcpyAsync(out.data(), dW, m * 4, hipMemcpyDeviceToHost, s));
  TSVM_HIP(c, hipStreamSynchronize(s));
  for (int64_t t = 0; t < m; t++) f[t] = (double)out[t] + b;
  return TSVM_OK;
}

}  // extern "C"
