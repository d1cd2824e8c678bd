// Model reader and GPU predictor for the model files the trainer writes: the LIBSVM text
// model of the RBF branch (writemodel, tensorsvm-train.cu:353-406) and the LIBLINEAR
// model of the linear branch (340-350).  The reference delegates prediction to an
// external ThunderSVM/LIBSVM predictor (README.md:27-37); its own predict() (166-270) is
// not reachable from main.  Decision values follow those predictors:
//   RBF:    f(x) = sum_j coef_j exp(-gamma |x - sv_j|^2) - rho,  coef_j = y_j alpha_j
//   linear: f(x) = w . x + w_bias * bias
// and the predicted label is label[0] when f > 0, else label[1].
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "internal.h"

struct tsvm_model {
  int kernel_type = 2;  // 2 rbf, 0 linear
  double gamma = 0.0, rho = 0.0, bias = -1.0;
  int label[2] = {1, -1};
  int64_t d = 0;              // feature count (max index seen)
  std::vector<double> coef;   // RBF: per SV
  std::vector<float> sv;      // RBF: nsv x d dense
  std::vector<double> w;      // linear: d weights (+ bias weight last when bias >= 0)
};

namespace {

bool read_all(const char* path, std::string& out) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  char buf[1 << 16];
  size_t r;
  while ((r = fread(buf, 1, sizeof(buf), f)) > 0) out.append(buf, r);
  fclose(f);
  return true;
}

std::vector<std::string> split_lines(const std::string& s) {
  std::vector<std::string> lines;
  size_t p = 0;
  while (p < s.size()) {
    size_t e = s.find('\n', p);
    if (e == std::string::npos) e = s.size();
    std::string ln = s.substr(p, e - p);
    if (!ln.empty() && ln.back() == '\r') ln.pop_back();
    lines.push_back(ln);
    p = e + 1;
  }
  return lines;
}

}  // namespace

extern "C" {

tsvm_status tsvm_read_model(const char* path, tsvm_model** out) {
  if (!path || !out) return TSVM_ERR_ARG;
  *out = nullptr;
  std::string text;
  if (!read_all(path, text)) return TSVM_ERR_IO;
  std::vector<std::string> lines = split_lines(text);
  tsvm_model* m = new tsvm_model();
  bool linear = false, in_sv = false, in_w = false, have_kernel = false;
  std::vector<std::vector<std::pair<long, float>>> rows;
  for (const std::string& ln : lines) {
    if (in_sv) {
      if (ln.find_first_not_of(" \t") == std::string::npos) continue;
      const char* p = ln.c_str();
      char* e;
      double cf = strtod(p, &e);
      if (e == p) goto bad;
      m->coef.push_back(cf);
      rows.emplace_back();
      p = e;
      for (;;) {
        while (*p == ' ' || *p == '\t') p++;
        if (!*p) break;
        long idx = strtol(p, &e, 10);
        if (e == p || *e != ':' || idx < 1) goto bad;
        p = e + 1;
        double v = strtod(p, &e);
        if (e == p) goto bad;
        p = e;
        rows.back().emplace_back(idx, (float)v);
        if (idx > m->d) m->d = idx;
      }
      continue;
    }
    if (in_w) {
      if (ln.find_first_not_of(" \t") == std::string::npos) continue;
      char* e;
      double v = strtod(ln.c_str(), &e);
      if (e == ln.c_str()) goto bad;
      m->w.push_back(v);
      continue;
    }
    char key[64] = {0};
    if (sscanf(ln.c_str(), "%63s", key) != 1) continue;
    const char* rest = ln.c_str() + strlen(key);
    if (!strcmp(key, "svm_type")) {
      if (!strstr(rest, "c_svc")) goto bad;
    } else if (!strcmp(key, "kernel_type")) {
      have_kernel = true;
      if (strstr(rest, "rbf")) m->kernel_type = 2;
      else if (strstr(rest, "linear")) m->kernel_type = 0;
      else goto bad;
    } else if (!strcmp(key, "gamma")) {
      m->gamma = atof(rest);
    } else if (!strcmp(key, "rho")) {
      m->rho = atof(rest);
    } else if (!strcmp(key, "label")) {
      if (sscanf(rest, "%d %d", &m->label[0], &m->label[1]) != 2) goto bad;
    } else if (!strcmp(key, "nr_class")) {
      if (atoi(rest) != 2) goto bad;
    } else if (!strcmp(key, "solver_type")) {
      linear = true;
      m->kernel_type = 0;
    } else if (!strcmp(key, "nr_feature")) {
      m->d = atol(rest);
    } else if (!strcmp(key, "bias")) {
      m->bias = atof(rest);
    } else if (!strcmp(key, "SV")) {
      in_sv = true;
    } else if (!strcmp(key, "w")) {
      in_w = true;
    }
    // total_sv, nr_sv and unknown keys carry nothing the decision needs
  }
  if (linear) {
    const size_t need = (size_t)m->d + (m->bias >= 0 ? 1 : 0);
    if (m->d <= 0 || m->w.size() != need) goto bad;
  } else {
    if (!have_kernel || m->kernel_type != 2 || !in_sv) goto bad;
    if (m->d == 0) m->d = 1;
    m->sv.assign(m->coef.size() * (size_t)m->d, 0.0f);
    for (size_t j = 0; j < rows.size(); j++)
      for (auto& iv : rows[j]) m->sv[j * m->d + (iv.first - 1)] = iv.second;
  }
  *out = m;
  return TSVM_OK;
bad:
  delete m;
  return TSVM_ERR_FORMAT;
}

void tsvm_model_free(tsvm_model* m) { delete m; }

tsvm_status tsvm_model_get(const tsvm_model* m, int* kernel_type, double* gamma, double* rho,
                           int* labels, int64_t* nsv, int64_t* d) {
  if (!m) return TSVM_ERR_ARG;
  if (kernel_type) *kernel_type = m->kernel_type;
  if (gamma) *gamma = m->gamma;
  if (rho) *rho = m->kernel_type == 2 ? m->rho : -(m->bias >= 0 ? m->w.back() * m->bias : 0.0);
  if (labels) {
    labels[0] = m->label[0];
    labels[1] = m->label[1];
  }
  if (nsv) *nsv = (int64_t)m->coef.size();
  if (d) *d = m->d;
  return TSVM_OK;
}

tsvm_status tsvm_predict(tsvm_ctx* ctx, const tsvm_model* m, const float* T, int64_t nt,
                         int64_t d, double* f, float* pred) {
  if (!ctx || !m || (nt > 0 && (!T || d <= 0))) return TSVM_ERR_ARG;
  tsvm::Ctx* c = &ctx->c;
  if (nt == 0) return TSVM_OK;
  std::vector<double> dec(nt);
  if (m->kernel_type == 0) {
    for (int64_t t = 0; t < nt; t++) {
      double s = m->bias >= 0 ? m->w.back() * m->bias : 0.0;
      for (int64_t j = 0; j < d && j < m->d; j++) s += m->w[j] * (double)T[t * d + j];
      dec[t] = s;
    }
  } else {
    // features beyond either side's width are zero on that side
    const int64_t dd = d > m->d ? d : m->d;
    const int64_t ns = (int64_t)m->coef.size();
    std::vector<float> tt, ss;
    const float* Tp = T;
    const float* Sp = m->sv.data();
    if (dd != d) {
      tt.assign(nt * dd, 0.0f);
      for (int64_t t = 0; t < nt; t++) memcpy(&tt[t * dd], T + t * d, d * 4);
      Tp = tt.data();
    }
    if (dd != m->d) {
      ss.assign(ns * dd, 0.0f);
      for (int64_t j = 0; j < ns; j++) memcpy(&ss[j * dd], m->sv.data() + j * m->d, m->d * 4);
      Sp = ss.data();
    }
    std::vector<float> cf(ns), ones(nt > ns ? nt : ns, 1.0f);
    for (int64_t j = 0; j < ns; j++) cf[j] = (float)m->coef[j];
    (void)hipSetDevice(c->device);
    hipStream_t s = c->stream;
    void *dS, *dc, *dT, *d1, *dW;
    TSVM_HIP(c, c->ws.get("pr_S", (ns ? ns : 1) * dd * 4, &dS));
    TSVM_HIP(c, c->ws.get("pr_c", (ns ? ns : 1) * 4, &dc));
    TSVM_HIP(c, c->ws.get("pr_T", nt * dd * 4, &dT));
    TSVM_HIP(c, c->ws.get("pr_1", ones.size() * 4, &d1));
    TSVM_HIP(c, c->ws.get("pr_W", nt * 4, &dW));
    if (ns > 0) {
      TSVM_HIP(c, hipMemcpyAsync(dS, Sp, ns * dd * 4, hipMemcpyHostToDevice, s));
      TSVM_HIP(c, hipMemcpyAsync(dc, cf.data(), ns * 4, hipMemcpyHostToDevice, s));
    }
    TSVM_HIP(c, hipMemcpyAsync(dT, Tp, nt * dd * 4, hipMemcpyHostToDevice, s));
    TSVM_HIP(c, hipMemcpyAsync(d1, ones.data(), ones.size() * 4, hipMemcpyHostToDevice, s));
    std::vector<float> out(nt, 0.0f);
    // coefficient j already carries y_j: unlabelled product (ones on both sides)
    TSVM_CHECK(tsvm::kernel_product_dev(c, (float*)dT, (float*)d1, nt, (float*)dS, (float*)d1, ns,
                                        dd, (float*)dc, 1, 1, m->gamma, TSVM_PREC_FP32,
                                        (float*)dW, 1));
    TSVM_HIP(c, hipMemcpyAsync(out.data(), dW, nt * 4, hipMemcpyDeviceToHost, s));
    TSVM_HIP(c, hipStreamSynchronize(s));
    for (int64_t t = 0; t < nt; t++) dec[t] = (double)out[t] - m->rho;
  }
  for (int64_t t = 0; t < nt; t++) {
    if (f) f[t] = dec[t];
    if (pred) pred[t] = (float)(dec[t] > 0 ? m->label[0] : m->label[1]);
  }
  return TSVM_OK;
}

}  // extern "C"
