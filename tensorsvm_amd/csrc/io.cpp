// Host I/O of the trainer: LIBSVM reader, label mapping, model writers.
//
// Reader semantics follow libsvmread (tensorsvm-train.cu:1162-1278): a line is
// "label idx:val idx:val ...", fields separated by spaces or tabs, indices strictly
// increasing within a line, the label must parse completely, a value may be followed by
// white space; d = max_index - min_index + 1 with min_index = min(1, every index) (equal
// to the reference's d for 1-based files; the reference overruns rows for 0-based ones).
// The file is read into memory once and parsed in two passes over the buffer.
#include <ctype.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../include/tsvm.h"

namespace {

inline bool is_sep(char ch) { return ch == ' ' || ch == '\t'; }

// Parse one line [p, e).  mode 0: validate and report max/min index; mode 1: fill row.
// Returns false on a format error.
bool parse_line(const char* p, const char* e, int mode, long* maxi, long* mini, float* label,
                float* row, long min_index, int scale256) {
  std::vector<char> buf(p, e);
  buf.push_back('\0');
  char* s = buf.data();
  char* q = s;
  while (*q && is_sep(*q)) q++;
  char* lab = q;
  while (*q && !is_sep(*q) && *q != '\n' && *q != '\r') q++;
  if (q == lab) return false;  // empty line
  char save = *q;
  *q = '\0';
  char* endp;
  double lv = strtod(lab, &endp);
  if (endp == lab || *endp != '\0') return false;
  if (label) *label = (float)lv;
  *q = save;
  long last = -1;
  for (;;) {
    while (*q && (is_sep(*q) || *q == '\n' || *q == '\r')) q++;
    if (!*q) break;
    char* idx = q;
    while (*q && *q != ':' && !is_sep(*q)) q++;
    if (*q != ':') return false;
    *q = '\0';
    errno = 0;
    long index = strtol(idx, &endp, 10);
    if (endp == idx || errno != 0 || *endp != '\0' || index <= last) return false;
    last = index;
    char* val = q + 1;
    errno = 0;
    double v = strtod(val, &endp);
    if (endp == val || errno != 0 || (*endp != '\0' && !isspace((unsigned char)*endp))) return false;
    q = endp;
    if (mode == 0) {
      if (index < *mini) *mini = index;
    } else {
      float fv = (float)v;
      if (scale256) fv /= 256;
      row[index - min_index] = fv;
    }
  }
  if (mode == 0 && last > *maxi) *maxi = last;
  return true;
}

}  // namespace

extern "C" {

void tsvm_free(void* p) { free(p); }

tsvm_status tsvm_read_libsvm(const char* path, int scale256, float** labels, float** inst,
                             int64_t* n_out, int64_t* d_out) {
  if (!path || !labels || !inst || !n_out || !d_out) return TSVM_ERR_ARG;
  FILE* f = fopen(path, "rb");
  if (!f) return TSVM_ERR_IO;
  std::vector<char> data;
  {
    char chunk[1 << 16];
    size_t r;
    while ((r = fread(chunk, 1, sizeof(chunk), f)) > 0) data.insert(data.end(), chunk, chunk + r);
  }
  fclose(f);
  std::vector<std::pair<size_t, size_t>> lines;
  size_t st = 0;
  for (size_t i = 0; i < data.size(); i++)
    if (data[i] == '\n') {
      lines.push_back({st, i});
      st = i + 1;
    }
  if (st < data.size()) lines.push_back({st, data.size()});
  long maxi = 0, mini = 1;
  for (auto& ln : lines)
    if (!parse_line(data.data() + ln.first, data.data() + ln.second, 0, &maxi, &mini, nullptr,
                    nullptr, 0, 0))
      return TSVM_ERR_FORMAT;
  const int64_t n = (int64_t)lines.size();
  int64_t d = maxi - mini + 1;
  if (d < 0) d = 0;
  float* lab = (float*)malloc(sizeof(float) * (n > 0 ? n : 1));
  float* x = (float*)calloc((size_t)(n * d > 0 ? n * d : 1), sizeof(float));
  if (!lab || !x) {
    free(lab);
    free(x);
    return TSVM_ERR_NOMEM;
  }
  for (int64_t i = 0; i < n; i++) {
    long a = 0, b = 0;
    if (!parse_line(data.data() + lines[i].first, data.data() + lines[i].second, 1, &a, &b,
                    &lab[i], x + i * d, mini, scale256)) {
      free(lab);
      free(x);
      return TSVM_ERR_FORMAT;
    }
  }
  *labels = lab;
  *inst = x;
  *n_out = n;
  *d_out = d;
  return TSVM_OK;
}

int64_t tsvm_map_labels(float* labels, int64_t n, int* pos_out, int* neg_out) {
  int pos = -42, neg = -42;
  for (int64_t i = 0; i < n; i++) {
    if (pos == -42) {
      pos = (int)labels[i];
    } else if (neg == -42 && pos != (int)labels[i]) {
      neg = (int)labels[i];
      break;
    }
  }
  int64_t bad = 0;
  for (int64_t i = 0; i < n; i++) {
    if (labels[i] == (float)pos) labels[i] = 1;
    else if (labels[i] == (float)neg) labels[i] = -1;
    else bad++;
  }
  if (pos_out) *pos_out = pos;
  if (neg_out) *neg_out = neg;
  return bad;
}

tsvm_status tsvm_write_model_rbf(const char* path, const double* alpha, const float* labels,
                                 const float* inst, int64_t n, int64_t d, double C,
                                 double gamma, double b, int pos, int neg) {
  int nSV = 0, nfree = 0;
  for (int64_t i = 0; i < n; i++)
    if (alpha[i] > 1e-6) {
      nSV++;
      if (alpha[i] < C - 1e-6) nfree++;
    }
  FILE* f = fopen(path, "w");
  if (!f) return TSVM_ERR_IO;
  fprintf(f, "svm_type c_svc\n");
  fprintf(f, "kernel_type rbf\n");
  fprintf(f, "gamma %.7f\n", gamma);
  fprintf(f, "nr_class 2\n");
  fprintf(f, "total_sv %d\n", nSV);
  fprintf(f, "rho %f\n", -b);
  fprintf(f, "label %d %d\n", pos, neg);
  fprintf(f, "nr_sv %d %d\n", nfree, nSV - nfree);
  fprintf(f, "SV\n");
  for (int64_t j = 0; j < n; j++) {
    if (!(alpha[j] > 1e-6)) continue;
    fprintf(f, "%7f ", labels[j] * alpha[j]);
    const float* r = inst + j * d;
    for (int64_t t = 0; t < d; t++)
      if (r[t] > 0 || r[t] < 0) fprintf(f, "%d:%7f ", (int)(t + 1), r[t]);
    fprintf(f, "\n");
  }
  if (fclose(f) != 0) return TSVM_ERR_IO;
  return TSVM_OK;
}

tsvm_status tsvm_write_model_linear(const char* path, const double* w, int64_t d, double b,
                                    int pos, int neg) {
  FILE* f = fopen(path, "w");
  if (!f) return TSVM_ERR_IO;
  fprintf(f, "solver_type L2R_L1LOSS_SVC_DUAL\n");
  fprintf(f, "nr_class 2\n");
  fprintf(f, "label %d %d\n", pos, neg);
  fprintf(f, "nr_feature %d\n", (int)d);
  fprintf(f, "bias %d\n", 1);
  fprintf(f, "w\n");
  for (int64_t i = 0; i < d; i++) fprintf(f, "%.16f \n", w[i]);
  fprintf(f, "%.16f\n", b);
  if (fclose(f) != 0) return TSVM_ERR_IO;
  return TSVM_OK;
}

}  // extern "C"
