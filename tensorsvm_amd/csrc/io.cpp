// Host I/O of the trainer: LIBSVM reader and writer, label mapping, model writers.
//
// Reader semantics follow libsvmread (tensorsvm-train.cu:1162-1278): a line is
// "label idx:val idx:val ...", fields separated by spaces or tabs, indices strictly
// increasing within a line, the label must parse completely, a value may be followed by
// white space; d = max_index - min_index + 1 with min_index = min(1, every index) (equal
// to the reference's d for 1-based files; the reference overruns rows for 0-based ones).
// Values are converted like the reference's atof: strtod, then rounded to float.
//
// The reference reads with fgets/strtok on one thread.  Here the file is read into one
// buffer (with a NUL after the last byte), cut into per-thread chunks at line
// boundaries, and parsed in place, each chunk by one thread: the parse validates every
// line and keeps its label and (index, value) pairs; once the index range and the line
// counts are known, each chunk's pairs are scattered into its rows.  Every scan is
// bounded by the line's end, so no conversion reads into the next line.
#include <ctype.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tsvm.h"

namespace {

inline bool is_sep(char ch) { return ch == ' ' || ch == '\t'; }

// Parses the line [p, e) (*e is '\n' or the buffer's NUL): the label, then every
// (index, value) pair through put(index, float value).  Returns false on a format error;
// *maxi receives the line's last (largest) index.
template <class Put>
bool parse_line(const char* p, const char* e, float* label, long* maxi, Put&& put) {
  const char* q = p;
  while (q < e && is_sep(*q)) q++;
  const char* lab = q;
  while (q < e && !is_sep(*q) && *q != '\r') q++;
  if (q == lab) return false;  // empty line
  char* endp;
  double lv = strtod(lab, &endp);
  if (endp == lab || endp != q) return false;  // the label token must parse completely
  *label = (float)lv;
  long last = -1;
  for (;;) {
    while (q < e && (is_sep(*q) || *q == '\r')) q++;
    if (q >= e) break;
    const char* idx = q;
    while (q < e && *q != ':' && !is_sep(*q)) q++;
    if (q >= e || *q != ':') return false;
    errno = 0;
    long index = strtol(idx, &endp, 10);  // stops at the ':' at the latest
    if (endp == idx || errno != 0 || endp != q || index <= last) return false;
    last = index;
    const char* val = q + 1;
    // strtod skips leading white space; do that here so that it stays within the line
    const char* v0 = val;
    while (v0 < e && isspace((unsigned char)*v0)) v0++;
    if (v0 >= e) return false;
    errno = 0;
    double v = strtod(v0, &endp);
    if (endp == v0 || errno != 0 || endp > e || (endp < e && !isspace((unsigned char)*endp)))
      return false;
    q = endp;
    put(index, (float)v);
  }
  if (last > *maxi) *maxi = last;
  return true;
}

int io_threads(int64_t bytes) {
  unsigned hc = std::thread::hardware_concurrency();
  int t = hc == 0 ? 1 : (int)std::min(hc, 32u);
  const int64_t per = 1 << 20;  // at least 1 MiB per thread
  if ((int64_t)t * per > bytes) t = (int)std::max<int64_t>(1, bytes / per);
  return t;
}

struct Chunk {
  size_t b = 0, e = 0;  // byte range (whole lines)
  long maxi = 0, mini = 1;
  bool ok = true;
  std::vector<float> lab;        // per line
  std::vector<int64_t> cnt;      // pairs per line
  std::vector<long> idx;         // pairs
  std::vector<float> val;
};

}  // namespace

extern "C" {

void tsvm_free(void* p) { free(p); }

tsvm_status tsvm_read_libsvm(const char* path, int scale256, float** labels, float** inst,
                             int64_t* n_out, int64_t* d_out) {
  if (!path || !labels || !inst || !n_out || !d_out) return TSVM_ERR_ARG;
  FILE* f = fopen(path, "rb");
  if (!f) return TSVM_ERR_IO;
  if (fseek(f, 0, SEEK_END) != 0) {
    fclose(f);
    return TSVM_ERR_IO;
  }
  const long fsz = ftell(f);
  if (fsz < 0 || fseek(f, 0, SEEK_SET) != 0) {
    fclose(f);
    return TSVM_ERR_IO;
  }
  const size_t size = (size_t)fsz;
  std::vector<char> data(size + 1);
  const size_t got = size ? fread(data.data(), 1, size, f) : 0;
  fclose(f);
  if (got != size) return TSVM_ERR_IO;
  data[size] = '\0';
  const char* buf = data.data();
  // chunks end just after a '\n' (or at the end of the file)
  const int nt = io_threads((int64_t)size);
  std::vector<Chunk> ch(nt);
  size_t pos = 0;
  for (int t = 0; t < nt; t++) {
    ch[t].b = pos;
    size_t target = t == nt - 1 ? size : std::max(pos, size * (size_t)(t + 1) / nt);
    if (target < size) {
      const void* nl = memchr(buf + target, '\n', size - target);
      target = nl ? (size_t)((const char*)nl - buf) + 1 : size;
    }
    ch[t].e = target;
    pos = target;
  }
  auto each_line = [&](const Chunk& c, auto&& fn) -> bool {
    size_t st = c.b;
    while (st < c.e) {
      const void* nl = memchr(buf + st, '\n', c.e - st);
      const size_t en = nl ? (size_t)((const char*)nl - buf) : c.e;
      if (!fn(buf + st, buf + en)) return false;
      st = en + 1;
    }
    return true;
  };
  {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
      th.emplace_back([&, t]() {
        Chunk& c = ch[t];
        c.ok = each_line(c, [&](const char* p, const char* e) {
          float lv = 0.0f;
          const size_t before = c.idx.size();
          const bool r = parse_line(p, e, &lv, &c.maxi, [&](long index, float v) {
            if (index < c.mini) c.mini = index;
            c.idx.push_back(index);
            c.val.push_back(v);
          });
          c.lab.push_back(lv);
          c.cnt.push_back((int64_t)(c.idx.size() - before));
          return r;
        });
      });
    for (auto& x : th) x.join();
  }
  long maxi = 0, mini = 1;
  int64_t n = 0;
  std::vector<int64_t> off(nt);
  for (int t = 0; t < nt; t++) {
    if (!ch[t].ok) return TSVM_ERR_FORMAT;
    maxi = std::max(maxi, ch[t].maxi);
    mini = std::min(mini, ch[t].mini);
    off[t] = n;
    n += (int64_t)ch[t].lab.size();
  }
  int64_t d = maxi - mini + 1;
  if (d < 0) d = 0;
  float* lab = (float*)malloc(sizeof(float) * (n > 0 ? n : 1));
  float* x = (float*)calloc((size_t)(n * d > 0 ? n * d : 1), sizeof(float));
  if (!lab || !x) {
    free(lab);
    free(x);
    return TSVM_ERR_NOMEM;
  }
  {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
      th.emplace_back([&, t]() {
        const Chunk& c = ch[t];
        size_t q = 0;
        for (size_t l = 0; l < c.lab.size(); l++) {
          const int64_t i = off[t] + (int64_t)l;
          lab[i] = c.lab[l];
          float* row = x + i * d;
          for (int64_t u = 0; u < c.cnt[l]; u++, q++) {
            float fv = c.val[q];
            if (scale256) fv /= 256;
            row[c.idx[q] - mini] = fv;
          }
        }
      });
    for (auto& y : th) y.join();
  }
  *labels = lab;
  *inst = x;
  *n_out = n;
  *d_out = d;
  return TSVM_OK;
}

tsvm_status tsvm_write_libsvm(const char* path, const float* labels, const float* inst,
                              int64_t n, int64_t d) {
  if (!path || n < 0 || d < 0 || (n > 0 && (!labels || (d > 0 && !inst)))) return TSVM_ERR_ARG;
  FILE* f = fopen(path, "wb");
  if (!f) return TSVM_ERR_IO;
  // rows formatted in parallel blocks, written in order; "%.9g" round-trips every float
  const int64_t BR = 65536;
  const int64_t nb = (n + BR - 1) / BR;
  unsigned hc = std::thread::hardware_concurrency();
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(hc == 0 ? 1 : std::min(hc, 32u), nb));
  bool io_ok = true;
  for (int64_t b0 = 0; b0 < nb && io_ok; b0 += nt) {
    std::vector<std::string> out(nt);
    std::vector<std::thread> th;
    for (int t = 0; t < nt && b0 + t < nb; t++)
      th.emplace_back([&, t]() {
        std::string& o = out[t];
        char tmp[64];
        const int64_t r0 = (b0 + t) * BR, r1 = std::min(n, r0 + BR);
        for (int64_t i = r0; i < r1; i++) {
          o.append(tmp, (size_t)snprintf(tmp, sizeof(tmp), "%.9g", (double)labels[i]));
          for (int64_t j = 0; j < d; j++) {
            const float v = inst[i * d + j];
            if (v == 0.0f) continue;
            o.append(tmp, (size_t)snprintf(tmp, sizeof(tmp), " %ld:%.9g", (long)(j + 1), (double)v));
          }
          o.push_back('\n');
        }
      });
    for (auto& y : th) y.join();
    for (auto& o : out)
      if (!o.empty() && fwrite(o.data(), 1, o.size(), f) != o.size()) io_ok = false;
  }
  if (fclose(f) != 0) io_ok = false;
  return io_ok ? TSVM_OK : TSVM_ERR_IO;
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
