/*
 * TEST INFRASTRUCTURE ONLY -- the CPU oracle (checker) for the TensorSVM training
 * hot path.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it.  It restates the reference trainer's algorithm (tensorsvm-train.cu) in
 * plain C with FP64 arithmetic; every function cites the reference lines it follows.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#define _GNU_SOURCE
#include "tsvm_oracle.h"

#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void orc_set_num_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}
void orc_free(void *p) { free(p); }

/* ------------------------------------------------------------------------- */
/* Omega: gaussrand (2621-2643), filled row-major (1966-1974).                 */
/* The reference is C++, so log/sqrt on float arguments resolve to the float   */
/* overloads: every step below is float arithmetic, no contraction.            */
/* ------------------------------------------------------------------------- */
void orc_glibc_rand(int32_t *out, int64_t count) {
  srand(1); /* glibc's default state equals srand(1); the reference never seeds */
  for (int64_t i = 0; i < count; i++) out[i] = rand();
}

void orc_gaussrand_fill(float *out, int64_t count) {
  float V1 = 0.f, V2 = 0.f, S = 0.f;
  int phase = 0;
  srand(1);
  for (int64_t t = 0; t < count; t++) {
    float X;
    if (phase == 0) {
      do {
        float U1 = (float)rand() / (float)RAND_MAX;
        float U2 = (float)rand() / (float)RAND_MAX;
        V1 = 2.0f * U1 - 1.0f;
        V2 = 2.0f * U2 - 1.0f;
        float a = V1 * V1;
        float b = V2 * V2;
        S = a + b;
      } while (S >= 1.0f || S == 0.0f);
      float l = logf(S);
      float q = (-2.0f * l) / S;
      X = V1 * sqrtf(q);
    } else {
      float l = logf(S);
      float q = (-2.0f * l) / S;
      X = V2 * sqrtf(q);
    }
    phase = 1 - phase;
    out[t] = X;
  }
}

/* ------------------------------------------------------------------------- */
/* LIBSVM reader (1162-1278) and label map (879-896).                          */
/* ------------------------------------------------------------------------- */
#define ORC_LINE 1000000
int orc_libsvm_read(const char *path, int scale256, float **labels, float **inst,
                    int64_t *n, int64_t *d) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  char *line = (char *)malloc(ORC_LINE);
  char *endptr;
  long max_index = 0, min_index = 1;
  int64_t l = 0;
  /* pass 1: rows, max index, strictly increasing indices (1182-1212) */
  while (fgets(line, ORC_LINE, f)) {
    long inst_max_index = -1;
    strtok(line, " \t");
    while (1) {
      char *idx = strtok(NULL, ":");
      char *val = strtok(NULL, " \t");
      if (val == NULL) break;
      errno = 0;
      long index = strtol(idx, &endptr, 10);
      if (endptr == idx || errno != 0 || *endptr != '\0' || index <= inst_max_index) {
        free(line);
        fclose(f);
        return -2;
      }
      inst_max_index = index;
      if (index < min_index) min_index = index;
    }
    if (inst_max_index > max_index) max_index = inst_max_index;
    l++;
  }
  int64_t nf = max_index - min_index + 1;
  if (nf < 0) nf = 0;
  float *lab = (float *)malloc(sizeof(float) * (l > 0 ? l : 1));
  float *x = (float *)calloc((size_t)((l * nf) > 0 ? l * nf : 1), sizeof(float));
  rewind(f);
  /* pass 2: labels and dense row-major values (1222-1266) */
  for (int64_t i = 0; i < l; i++) {
    if (!fgets(line, ORC_LINE, f)) break;
    char *label = strtok(line, " \t\n");
    if (label == NULL) {
      free(line); free(lab); free(x); fclose(f);
      return -2;
    }
    lab[i] = (float)strtod(label, &endptr);
    if (endptr == label || *endptr != '\0') {
      free(line); free(lab); free(x); fclose(f);
      return -2;
    }
    while (1) {
      char *idx = strtok(NULL, ":");
      char *val = strtok(NULL, " \t");
      if (val == NULL) break;
      long index = strtol(idx, &endptr, 10) - min_index;
      errno = 0;
      float v = (float)strtod(val, &endptr);
      if (scale256) v /= 256;
      if (endptr == val || errno != 0 || (*endptr != '\0' && !isspace((unsigned char)*endptr))) {
        free(line); free(lab); free(x); fclose(f);
        return -2;
      }
      x[i * nf + index] = v;
    }
  }
  free(line);
  fclose(f);
  *labels = lab;
  *inst = x;
  *n = l;
  *d = nf;
  return 0;
}

int64_t orc_label_map(float *labels, int64_t n, int *pos_out, int *neg_out) {
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
  *pos_out = pos;
  *neg_out = neg;
  return bad;
}

/* ------------------------------------------------------------------------- */
/* Kernel product (2476-2612): W = diag(y1) K(X1, X2) diag(y2) A.              */
/* ------------------------------------------------------------------------- */
void orc_kernel_product(const float *X1, const float *y1, int64_t m, const float *X2,
                        const float *y2, int64_t n, int64_t d, const double *A, int64_t r,
                        double gamma, double *W) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < m; i++) {
    double *w = W + i * r;
    for (int64_t c = 0; c < r; c++) w[c] = 0.0;
    const float *xi = X1 + i * d;
    for (int64_t j = 0; j < n; j++) {
      const float *xj = X2 + j * d;
      double dist = 0.0;
      for (int64_t t = 0; t < d; t++) {
        double df = (double)xi[t] - (double)xj[t];
        dist += df * df;
      }
      double kv = (double)y2[j] * exp(-gamma * dist);
      const double *aj = A + j * r;
      for (int64_t c = 0; c < r; c++) w[c] += kv * aj[c];
    }
    for (int64_t c = 0; c < r; c++) w[c] *= (double)y1[i];
  }
}

/* The same product in the reference's FP32 arithmetic: float norms (vecnorm 2569-2581),
 * float X_i . X_j (the XIJ sgemm, 2506-2518), the kernel value
 * y_i y_j expf(-gamma (XI + XJ - 2 XIJ)) with the distance in float (rbf_kergen 2598-2612)
 * and float accumulation of W_i += buf A_j over all n columns in index order (the K A
 * sgemm with beta = 1, 2537-2555, and its chunk loop 2017-2058).  cuBLAS's blocking order
 * and __expf's error are not reproduced (libm's expf instead). */
void orc_kernel_product_f32(const float *X1, const float *y1, int64_t m, const float *X2,
                            const float *y2, int64_t n, int64_t d, const float *A, int64_t r,
                            double gamma, float *W) {
  float *xn2 = (float *)malloc(sizeof(float) * (n > 0 ? n : 1));
  for (int64_t j = 0; j < n; j++) {
    float sum = 0.0f;
    for (int64_t t = 0; t < d; t++) sum += X2[j * d + t] * X2[j * d + t];
    xn2[j] = sum;
  }
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < m; i++) {
    float *w = W + i * r;
    for (int64_t c = 0; c < r; c++) w[c] = 0.0f;
    const float *xi = X1 + i * d;
    float xn1 = 0.0f;
    for (int64_t t = 0; t < d; t++) xn1 += xi[t] * xi[t];
    for (int64_t j = 0; j < n; j++) {
      const float *xj = X2 + j * d;
      float dot = 0.0f;
      for (int64_t t = 0; t < d; t++) dot += xi[t] * xj[t];
      const float dist = xn1 + xn2[j] - 2.0f * dot;
      const float kv = y1[i] * y2[j] * expf((float)(-gamma * (double)dist));
      const float *aj = A + j * r;
      for (int64_t c = 0; c < r; c++) w[c] += kv * aj[c];
    }
  }
  free(xn2);
}

/* ------------------------------------------------------------------------- */
/* Orthonormalisation (replaces ortho/CAQR/QR 617-856): thin Householder Q.    */
/* ------------------------------------------------------------------------- */
void orc_householder_q(double *W, int64_t n, int64_t k) {
  double *V = (double *)calloc((size_t)(n * k), sizeof(double)); /* reflectors, row-major */
  double *beta = (double *)calloc((size_t)k, sizeof(double));
  double *tmp = (double *)malloc(sizeof(double) * k);
  for (int64_t j = 0; j < k && j < n; j++) {
    double alpha2 = 0.0;
    for (int64_t i = j; i < n; i++) alpha2 += W[i * k + j] * W[i * k + j];
    double nrm = sqrt(alpha2);
    double x0 = W[j * k + j];
    double s = (x0 >= 0) ? -nrm : nrm;
    /* v = x - s e1, beta = 2 / (v^T v) */
    for (int64_t i = j; i < n; i++) V[i * k + j] = W[i * k + j];
    V[j * k + j] = x0 - s;
    double vtv = alpha2 - x0 * x0 + (x0 - s) * (x0 - s);
    beta[j] = (vtv > 0) ? 2.0 / vtv : 0.0;
    /* apply H = I - beta v v^T to columns j..k-1 */
    for (int64_t c = j; c < k; c++) tmp[c] = 0.0;
    for (int64_t i = j; i < n; i++) {
      double vi = V[i * k + j];
      for (int64_t c = j; c < k; c++) tmp[c] += vi * W[i * k + c];
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = j; i < n; i++) {
      double vi = V[i * k + j] * beta[j];
      for (int64_t c = j; c < k; c++) W[i * k + c] -= vi * tmp[c];
    }
  }
  /* form Q = H_0 H_1 ... H_{k-1} [I_k; 0] */
  for (int64_t i = 0; i < n * k; i++) W[i] = 0.0;
  for (int64_t i = 0; i < k && i < n; i++) W[i * k + i] = 1.0;
  for (int64_t j = (k < n ? k : n) - 1; j >= 0; j--) {
    for (int64_t c = 0; c < k; c++) tmp[c] = 0.0;
    for (int64_t i = j; i < n; i++) {
      double vi = V[i * k + j];
      for (int64_t c = 0; c < k; c++) tmp[c] += vi * W[i * k + c];
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = j; i < n; i++) {
      double vi = V[i * k + j] * beta[j];
      for (int64_t c = 0; c < k; c++) W[i * k + c] -= vi * tmp[c];
    }
  }
  free(V);
  free(beta);
  free(tmp);
}

/* ------------------------------------------------------------------------- */
/* Symmetric eigen-decomposition by cyclic Jacobi (replaces dsyevd, 2191).     */
/* ------------------------------------------------------------------------- */
void orc_sym_eig(const double *A0, int64_t k, double *w, double *V) {
  double *A = (double *)malloc(sizeof(double) * k * k);
  memcpy(A, A0, sizeof(double) * k * k);
  for (int64_t i = 0; i < k * k; i++) V[i] = 0.0;
  for (int64_t i = 0; i < k; i++) V[i * k + i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int64_t p = 0; p < k; p++)
      for (int64_t q = 0; q < k; q++) {
        double v = A[p * k + q] * A[p * k + q];
        tot += v;
        if (p != q) off += v;
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int64_t p = 0; p < k - 1; p++) {
      for (int64_t q = p + 1; q < k; q++) {
        double apq = A[p * k + q];
        if (fabs(apq) < 1e-300) continue;
        double app = A[p * k + p], aqq = A[q * k + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int64_t r = 0; r < k; r++) { /* columns p, q */
          double arp = A[r * k + p], arq = A[r * k + q];
          A[r * k + p] = c * arp - s * arq;
          A[r * k + q] = s * arp + c * arq;
        }
        for (int64_t r = 0; r < k; r++) { /* rows p, q */
          double apr = A[p * k + r], aqr = A[q * k + r];
          A[p * k + r] = c * apr - s * aqr;
          A[q * k + r] = s * apr + c * aqr;
        }
        for (int64_t r = 0; r < k; r++) {
          double vrp = V[r * k + p], vrq = V[r * k + q];
          V[r * k + p] = c * vrp - s * vrq;
          V[r * k + q] = s * vrp + c * vrq;
        }
      }
    }
  }
  /* sort ascending */
  int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * k);
  for (int64_t i = 0; i < k; i++) idx[i] = i;
  for (int64_t i = 1; i < k; i++) {
    int64_t t = idx[i];
    int64_t j = i - 1;
    while (j >= 0 && A[idx[j] * k + idx[j]] > A[t * k + t]) {
      idx[j + 1] = idx[j];
      j--;
    }
    idx[j + 1] = t;
  }
  double *Vs = (double *)malloc(sizeof(double) * k * k);
  for (int64_t c = 0; c < k; c++) {
    w[c] = A[idx[c] * k + idx[c]];
    for (int64_t r = 0; r < k; r++) Vs[r * k + c] = V[r * k + idx[c]];
  }
  memcpy(V, Vs, sizeof(double) * k * k);
  free(Vs);
  free(idx);
  free(A);
}

/* ------------------------------------------------------------------------- */
/* LRA (1941-2249).                                                            */
/* ------------------------------------------------------------------------- */
/* round every entry to FP32 (the reference keeps W, Q, C, X and U in float storage) */
static void round_f32(double *a, int64_t cnt) {
  for (int64_t i = 0; i < cnt; i++) a[i] = (double)(float)a[i];
}

int64_t orc_lra_ex(const float *X, const float *y, int64_t n, int64_t d, int64_t k,
                   double gamma, const float *omega, int mode, double *U, double *eig_out) {
  const int f32 = mode == ORC_LRA_REF_FP32;
  if (n <= 0 || k <= 0) return 0;
  double *Om = (double *)malloc(sizeof(double) * n * k);
  double *W = (double *)malloc(sizeof(double) * n * k);
  double *Q = (double *)malloc(sizeof(double) * n * k);
  for (int64_t i = 0; i < n * k; i++) Om[i] = (double)omega[i];
  float *Wf = f32 ? (float *)malloc(sizeof(float) * n * k) : NULL;
  float *Af = f32 ? (float *)malloc(sizeof(float) * n * k) : NULL;
  if (f32) { /* W = K_Y Omega in FP32 arithmetic (2067) */
    orc_kernel_product_f32(X, y, n, X, y, n, d, omega, k, gamma, Wf);
    for (int64_t i = 0; i < n * k; i++) W[i] = (double)Wf[i];
  } else {
    orc_kernel_product(X, y, n, X, y, n, d, Om, k, gamma, W); /* W = K_Y Omega (2067) */
  }
  orc_householder_q(W, n, k); /* ortho (2075) */
  if (f32) {                  /* float W out of ortho(float *W, ...); Q = K_Y W (2085) */
    for (int64_t i = 0; i < n * k; i++) Af[i] = (float)W[i];
    round_f32(W, n * k);
    orc_kernel_product_f32(X, y, n, X, y, n, d, Af, k, gamma, Wf);
    for (int64_t i = 0; i < n * k; i++) Q[i] = (double)Wf[i];
    free(Wf);
    free(Af);
  } else {
    orc_kernel_product(X, y, n, X, y, n, d, W, k, gamma, Q); /* Q = K_Y W (2085) */
  }
  double *C = (double *)calloc((size_t)(k * k), sizeof(double));
  for (int64_t i = 0; i < n; i++) /* C = W^T Q, FP64 accumulation (2122-2149) */
    for (int64_t a = 0; a < k; a++) {
      double wa = W[i * k + a];
      for (int64_t b = 0; b < k; b++) C[a * k + b] += wa * Q[i * k + b];
    }
  if (f32) round_f32(C, k * k); /* the FP64 sum lands in float C (2150) */
  for (int64_t a = 0; a < k; a++) /* symmetrise from the lower triangle (2152-2156) */
    for (int64_t b = 0; b < a; b++) {
      double v = 0.5 * (C[a * k + b] + C[b * k + a]);
      if (f32) v = (double)(float)v;
      C[a * k + b] = v;
      C[b * k + a] = v;
    }
  double *w = (double *)malloc(sizeof(double) * k);
  double *V = (double *)malloc(sizeof(double) * k * k);
  orc_sym_eig(C, k, w, V); /* dsyevd (2191) */
  int64_t realk = k;
  for (int64_t c = 0; c < k; c++) { /* X = V sqrt(max(w,0)) (2197-2210) */
    double s = 0.0;
    if (w[c] > 0) s = sqrt(w[c]);
    else realk--;
    for (int64_t r = 0; r < k; r++) V[r * k + c] *= s;
  }
  if (f32) round_f32(V, k * k); /* float X (2187, 2208) */
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) /* U = W X (2213-2241): sgemm in the FP32 form */
    for (int64_t c = 0; c < k; c++) {
      if (f32) {
        float acc = 0.0f;
        for (int64_t t = 0; t < k; t++) acc += (float)W[i * k + t] * (float)V[t * k + c];
        U[i * k + c] = (double)acc;
      } else {
        double acc = 0.0;
        for (int64_t t = 0; t < k; t++) acc += W[i * k + t] * V[t * k + c];
        U[i * k + c] = acc;
      }
    }
  if (eig_out) memcpy(eig_out, w, sizeof(double) * k);
  free(Om); free(W); free(Q); free(C); free(w); free(V);
  return realk;
}

int64_t orc_lra(const float *X, const float *y, int64_t n, int64_t d, int64_t k,
                double gamma, const float *omega, double *U, double *eig_out) {
  return orc_lra_ex(X, y, n, d, k, gamma, omega, ORC_LRA_FP64, U, eig_out);
}

/* ------------------------------------------------------------------------- */
/* Gradients of the trained dual and of the held-out smoothed hinge loss      */
/* (no counterpart in the reference; the dual and the factor follow LRA      */
/* 1941-2249 and mpc 1636-1643).                                              */
/* ------------------------------------------------------------------------- */
static double kval(const double *xs, int64_t i, const double *xt, int64_t t, int64_t d,
                   double gamma) {
  double s = 0.0;
  for (int64_t f = 0; f < d; f++) {
    double df = xs[i * d + f] - xt[t * d + f];
    s += df * df;
  }
  return exp(-gamma * s);
}

/* out (n x r) = K_Y A over the scaled rows (labels on both sides) */
static void kprod64(const double *xs, const float *y, int64_t n, int64_t d, double gamma,
                    const double *A, int64_t r, double *out) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < n; i++) {
    double *o = out + i * r;
    for (int64_t c = 0; c < r; c++) o[c] = 0.0;
    for (int64_t j = 0; j < n; j++) {
      double kij = (double)y[i] * (double)y[j] * kval(xs, i, xs, j, d, gamma);
      for (int64_t c = 0; c < r; c++) o[c] += kij * A[j * r + c];
    }
  }
}

/* S_f = sum_ij p_i q_j K(a_i, c_j) (a_if - c_jf)^2 for rows a (m) and c (n) */
static void bilinear_dk(const double *xa, int64_t m, const double *xc, int64_t n, int64_t d,
                        double gamma, const double *p, const double *q, double *S) {
  for (int64_t f = 0; f < d; f++) S[f] = 0.0;
#pragma omp parallel
  {
    double *loc = (double *)calloc((size_t)d, sizeof(double));
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < m; i++) {
      if (p[i] == 0.0) continue;
      for (int64_t j = 0; j < n; j++) {
        double w = p[i] * q[j] * kval(xa, i, xc, j, d, gamma);
        for (int64_t f = 0; f < d; f++) {
          double df = xa[i * d + f] - xc[j * d + f];
          loc[f] += w * df * df;
        }
      }
    }
#pragma omp critical
    for (int64_t f = 0; f < d; f++) S[f] += loc[f];
    free(loc);
  }
}

/* x <- A^{-1} x for a k x k matrix (Gaussian elimination, partial pivoting) */
static int solve_dense(const double *A, int64_t k, double *x) {
  double *M = (double *)malloc(sizeof(double) * k * k);
  memcpy(M, A, sizeof(double) * k * k);
  for (int64_t c = 0; c < k; c++) {
    int64_t piv = c;
    for (int64_t r = c + 1; r < k; r++)
      if (fabs(M[r * k + c]) > fabs(M[piv * k + c])) piv = r;
    if (M[piv * k + c] == 0.0) { free(M); return -1; }
    if (piv != c) {
      for (int64_t j = 0; j < k; j++) { double t = M[c * k + j]; M[c * k + j] = M[piv * k + j]; M[piv * k + j] = t; }
      double t = x[c]; x[c] = x[piv]; x[piv] = t;
    }
    for (int64_t r = c + 1; r < k; r++) {
      double l = M[r * k + c] / M[c * k + c];
      if (l == 0.0) continue;
      for (int64_t j = c; j < k; j++) M[r * k + j] -= l * M[c * k + j];
      x[r] -= l * x[c];
    }
  }
  for (int64_t c = k - 1; c >= 0; c--) {
    double s = x[c];
    for (int64_t j = c + 1; j < k; j++) s -= M[c * k + j] * x[j];
    x[c] = s / M[c * k + c];
  }
  free(M);
  return 0;
}

int orc_grad(const float *X, const float *y, int64_t n, int64_t d, int64_t k,
             const float *omega, const double *alpha, double b, double log_gamma,
             const double *log_scales, const float *Xt, const float *yt, int64_t nt,
             double *vals, double *gD, double *gL) {
  const double gamma = exp(log_gamma);
  double *xs = (double *)malloc(sizeof(double) * n * d);
  double *xt = (double *)malloc(sizeof(double) * (nt > 0 ? nt : 1) * d);
  for (int64_t i = 0; i < n; i++)
    for (int64_t f = 0; f < d; f++) xs[i * d + f] = (double)X[i * d + f] * exp(log_scales[f]);
  for (int64_t t = 0; t < nt; t++)
    for (int64_t f = 0; f < d; f++) xt[t * d + f] = (double)Xt[t * d + f] * exp(log_scales[f]);
  double *Om = (double *)malloc(sizeof(double) * n * k);
  double *Y = (double *)malloc(sizeof(double) * n * k);
  double *W = (double *)malloc(sizeof(double) * n * k);
  double *Q = (double *)malloc(sizeof(double) * n * k);
  for (int64_t i = 0; i < n * k; i++) Om[i] = (double)omega[i];
  kprod64(xs, y, n, d, gamma, Om, k, Y); /* Y = K_Y Omega */
  memcpy(W, Y, sizeof(double) * n * k);
  orc_householder_q(W, n, k);            /* range(W) = range(Y) */
  kprod64(xs, y, n, d, gamma, W, k, Q);  /* Q = K_Y W */
  double *C = (double *)calloc((size_t)(k * k), sizeof(double));
  double *R = (double *)calloc((size_t)(k * k), sizeof(double));
  double *beta = (double *)calloc((size_t)k, sizeof(double));
  for (int64_t i = 0; i < n; i++)
    for (int64_t a = 0; a < k; a++) {
      double wa = W[i * k + a];
      beta[a] += wa * alpha[i];
      for (int64_t c = 0; c < k; c++) {
        C[a * k + c] += wa * Q[i * k + c];
        R[a * k + c] += wa * Y[i * k + c];
      }
    }
  for (int64_t a = 0; a < k; a++)
    for (int64_t c = 0; c < a; c++) {
      double v = 0.5 * (C[a * k + c] + C[c * k + a]);
      C[a * k + c] = v;
      C[c * k + a] = v;
    }
  int ret = 0;
  { /* positive definiteness of C (no eigenvalue clipping in the factor) */
    double *L = (double *)malloc(sizeof(double) * k * k);
    memcpy(L, C, sizeof(double) * k * k);
    for (int64_t j = 0; j < k && ret == 0; j++) {
      double s = L[j * k + j];
      for (int64_t t = 0; t < j; t++) s -= L[j * k + t] * L[j * k + t];
      if (!(s > 0.0)) { ret = -1; break; }
      L[j * k + j] = sqrt(s);
      for (int64_t i = j + 1; i < k; i++) {
        double v = L[i * k + j];
        for (int64_t t = 0; t < j; t++) v -= L[i * k + t] * L[j * k + t];
        L[i * k + j] = v / L[j * k + j];
      }
    }
    free(L);
  }
  double sa = 0.0, bcb = 0.0;
  for (int64_t i = 0; i < n; i++) sa += alpha[i];
  for (int64_t a = 0; a < k; a++)
    for (int64_t c = 0; c < k; c++) bcb += beta[a] * C[a * k + c] * beta[c];
  vals[0] = sa - 0.5 * bcb;
  if (gD && ret == 0) {
    double *Pa = (double *)malloc(sizeof(double) * n), *g = (double *)malloc(sizeof(double) * n);
    double *wg = (double *)calloc((size_t)k, sizeof(double));
    for (int64_t i = 0; i < n; i++) {
      double s1 = 0.0, s2 = 0.0;
      for (int64_t a = 0; a < k; a++) {
        s1 += W[i * k + a] * beta[a];
        s2 += Q[i * k + a] * beta[a];
      }
      Pa[i] = s1; /* P alpha */
      g[i] = s2;  /* K_Y P alpha */
    }
    for (int64_t i = 0; i < n; i++)
      for (int64_t a = 0; a < k; a++) wg[a] += W[i * k + a] * g[i];
    double *c1 = (double *)malloc(sizeof(double) * k), *c2 = (double *)malloc(sizeof(double) * k);
    memcpy(c1, wg, sizeof(double) * k);
    memcpy(c2, beta, sizeof(double) * k);
    solve_dense(R, k, c1); /* Y^+ g = R^{-1} W^T g */
    solve_dense(R, k, c2); /* Y^+ alpha */
    double *pp = (double *)malloc(sizeof(double) * n * 3), *qq = (double *)malloc(sizeof(double) * n * 3);
    for (int64_t i = 0; i < n; i++) {
      double wwg = 0.0, o1 = 0.0, o2 = 0.0;
      for (int64_t a = 0; a < k; a++) {
        wwg += W[i * k + a] * wg[a];
        o1 += Om[i * k + a] * c1[a];
        o2 += Om[i * k + a] * c2[a];
      }
      double yi = y[i];
      pp[i] = yi * Pa[i];              qq[i] = yi * Pa[i];       /* (P a, P a) */
      pp[n + i] = yi * (alpha[i] - Pa[i]); qq[n + i] = yi * o1;  /* ((I-P) a, Omega Y^+ g) */
      pp[2 * n + i] = yi * (g[i] - wwg);   qq[2 * n + i] = yi * o2; /* ((I-P) g, Omega Y^+ a) */
    }
    double *S = (double *)malloc(sizeof(double) * d), *T = (double *)calloc((size_t)d, sizeof(double));
    for (int m = 0; m < 3; m++) {
      bilinear_dk(xs, n, xs, n, d, gamma, pp + m * n, qq + m * n, S);
      for (int64_t f = 0; f < d; f++) T[f] += (m == 0 ? 1.0 : 2.0) * S[f];
    }
    double tot = 0.0;
    for (int64_t f = 0; f < d; f++) {
      gD[1 + f] = gamma * T[f];
      tot += T[f];
    }
    gD[0] = 0.5 * gamma * tot;
    free(Pa); free(g); free(wg); free(c1); free(c2); free(pp); free(qq); free(S); free(T);
  }
  /* held-out smoothed hinge */
  vals[1] = 0.0;
  if (nt > 0) {
    double *q = (double *)malloc(sizeof(double) * n), *p = (double *)malloc(sizeof(double) * nt);
    for (int64_t i = 0; i < n; i++) q[i] = alpha[i] * (double)y[i];
    double loss = 0.0;
    for (int64_t t = 0; t < nt; t++) {
      double f = b;
      for (int64_t i = 0; i < n; i++) f += q[i] * kval(xt, t, xs, i, d, gamma);
      double z = (double)yt[t] * f, h, dh;
      if (z <= 0.0) { h = 0.5 - z; dh = -1.0; }
      else if (z < 1.0) { h = 0.5 * (1.0 - z) * (1.0 - z); dh = -(1.0 - z); }
      else { h = 0.0; dh = 0.0; }
      loss += h;
      p[t] = dh * (double)yt[t] / (double)nt;
    }
    vals[1] = loss / (double)nt;
    if (gL) {
      double *S = (double *)malloc(sizeof(double) * d);
      bilinear_dk(xt, nt, xs, n, d, gamma, p, q, S);
      double tot = 0.0;
      for (int64_t f = 0; f < d; f++) {
        gL[1 + f] = -2.0 * gamma * S[f];
        tot += S[f];
      }
      gL[0] = -gamma * tot;
      free(S);
    }
    free(q); free(p);
  } else if (gL) {
    for (int64_t f = 0; f <= d; f++) gL[f] = 0.0;
  }
  free(xs); free(xt); free(Om); free(Y); free(W); free(Q); free(C); free(R); free(beta);
  return ret;
}

/* ------------------------------------------------------------------------- */
/* IPM (mpc 1507-1785, NewtonStep 1794-1822, SMWSolve 1824-1840).              */
/* ------------------------------------------------------------------------- */
static void gemv_t(const double *Z, int64_t n, int64_t k, const double *v, double *out) {
  /* out = Z^T v */
  int nt = orc_num_threads();
  double *part = (double *)calloc((size_t)(nt * k), sizeof(double));
#pragma omp parallel
  {
#ifdef _OPENMP
    int t = omp_get_thread_num();
#else
    int t = 0;
#endif
    double *p = part + (int64_t)t * k;
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      double vi = v[i];
      const double *z = Z + i * k;
      for (int64_t c = 0; c < k; c++) p[c] += vi * z[c];
    }
  }
  for (int64_t c = 0; c < k; c++) {
    double s = 0.0;
    for (int t = 0; t < nt; t++) s += part[(int64_t)t * k + c];
    out[c] = s;
  }
  free(part);
}
static void gemv_n(const double *Z, int64_t n, int64_t k, const double *q, double *out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    double s = 0.0;
    const double *z = Z + i * k;
    for (int64_t c = 0; c < k; c++) s += z[c] * q[c];
    out[i] = s;
  }
}
/* lower Cholesky in place (dpotrf 'L'); returns 0 or the failing column + 1 */
static int chol_lower(double *M, int64_t k) {
  for (int64_t j = 0; j < k; j++) {
    double s = M[j * k + j];
    for (int64_t t = 0; t < j; t++) s -= M[j * k + t] * M[j * k + t];
    if (!(s > 0.0)) return (int)(j + 1);
    double ljj = sqrt(s);
    M[j * k + j] = ljj;
#pragma omp parallel for schedule(static) if (k - j > 128)
    for (int64_t i = j + 1; i < k; i++) {
      double v = M[i * k + j];
      for (int64_t t = 0; t < j; t++) v -= M[i * k + t] * M[j * k + t];
      M[i * k + j] = v / ljj;
    }
  }
  return 0;
}
/* dpotrs 'L', one right-hand side */
static void chol_solve(const double *L, int64_t k, double *b) {
  for (int64_t i = 0; i < k; i++) {
    double v = b[i];
    for (int64_t t = 0; t < i; t++) v -= L[i * k + t] * b[t];
    b[i] = v / L[i * k + i];
  }
  for (int64_t i = k - 1; i >= 0; i--) {
    double v = b[i];
    for (int64_t t = i + 1; t < k; t++) v -= L[t * k + i] * b[t];
    b[i] = v / L[i * k + i];
  }
}
/* M = I + Z^T D^{-1} Z, factored (1660-1714) */
static int build_capacitance(const double *Z, const double *D, int64_t n, int64_t k, double *M) {
  int nt = orc_num_threads();
  double *part = (double *)calloc((size_t)nt * k * k, sizeof(double));
#pragma omp parallel
  {
#ifdef _OPENMP
    int t = omp_get_thread_num();
#else
    int t = 0;
#endif
    double *p = part + (int64_t)t * k * k;
    double *zs = (double *)malloc(sizeof(double) * k);
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      double sc = 1.0 / sqrt(D[i]); /* rows of Z scaled by D^{-1/2} (1669) */
      for (int64_t c = 0; c < k; c++) zs[c] = Z[i * k + c] * sc;
      for (int64_t a = 0; a < k; a++) {
        double za = zs[a];
        double *pa = p + a * k;
        for (int64_t b = 0; b <= a; b++) pa[b] += za * zs[b];
      }
    }
    free(zs);
  }
  for (int64_t a = 0; a < k; a++)
    for (int64_t b = 0; b <= a; b++) {
      double s = 0.0;
      for (int t = 0; t < nt; t++) s += part[(int64_t)t * k * k + a * k + b];
      M[a * k + b] = s;
      M[b * k + a] = s;
    }
  free(part);
  for (int64_t a = 0; a < k; a++) M[a * k + a] += 1.0; /* 1706 */
  return chol_lower(M, k);                             /* 1709 */
}
static void smw_solve_fact(const double *Z, const double *D, const double *L, int64_t n,
                           int64_t k, double *b) {
  double *c = (double *)malloc(sizeof(double) * n);
  double *bb = (double *)malloc(sizeof(double) * k);
  for (int64_t i = 0; i < n; i++) { b[i] /= D[i]; c[i] = b[i]; }
  gemv_t(Z, n, k, b, bb);
  chol_solve(L, k, bb);
  gemv_n(Z, n, k, bb, b);
  for (int64_t i = 0; i < n; i++) b[i] = c[i] - b[i] / D[i];
  free(c);
  free(bb);
}
void orc_smw_solve(const double *Z, const double *D, int64_t n, int64_t k, double *b) {
  double *M = (double *)malloc(sizeof(double) * k * k);
  build_capacitance(Z, D, n, k, M);
  smw_solve_fact(Z, D, M, n, k, b);
  free(M);
}
/* NewtonStep (1794-1822): r = [r1 (n); r2 (1); r3 (n); r4 (n)] overwritten by the step */
static void newton_step(const double *Z, const double *D, const double *L, double C,
                        const double *a, const double *X, const double *S, const double *Xi,
                        double *r, int64_t n, int64_t k) {
  double *r1 = r, *r2 = r + n, *r3 = r + n + 1, *r4 = r + 2 * n + 1;
  double *r5 = (double *)malloc(sizeof(double) * n);
  double *r7 = (double *)malloc(sizeof(double) * n);
  double *b = (double *)malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; i++) {
    r5[i] = r1[i] - r3[i] / X[i] + r4[i] / (C - X[i]);
    r7[i] = r5[i];
  }
  smw_solve_fact(Z, D, L, n, k, r7);
  double r6 = r2[0];
  double dot1 = 0.0, dot2 = 0.0;
  for (int64_t i = 0; i < n; i++) dot1 += a[i] * r7[i];
  r6 += dot1;
  for (int64_t i = 0; i < n; i++) b[i] = a[i];
  smw_solve_fact(Z, D, L, n, k, b);
  for (int64_t i = 0; i < n; i++) dot2 += a[i] * b[i];
  r2[0] = r6 / dot2;
  for (int64_t i = 0; i < n; i++) r1[i] = a[i] * r2[0] - r5[i];
  smw_solve_fact(Z, D, L, n, k, r1);
  for (int64_t i = 0; i < n; i++) {
    r3[i] = (r3[i] - S[i] * r1[i]) / X[i];
    r4[i] = (r4[i] + Xi[i] * r1[i]) / (C - X[i]);
  }
  free(r5); free(r7); free(b);
}
void orc_newton_direction(const double *Z, const double *a, int64_t n, int64_t k, double C,
                          const double *X, const double *S, const double *Xi, double *r) {
  double *D = (double *)malloc(sizeof(double) * n);
  double *M = (double *)malloc(sizeof(double) * k * k);
  for (int64_t i = 0; i < n; i++) D[i] = S[i] / X[i] + Xi[i] / (C - X[i]);
  build_capacitance(Z, D, n, k, M);
  newton_step(Z, D, M, C, a, X, S, Xi, r, n, k);
  free(D);
  free(M);
}

static double max_step(const double *r, const double *X, const double *S, const double *Xi,
                       double C, int64_t n) {
  const double *dx = r, *ds = r + n + 1, *dxi = r + 2 * n + 1;
  double alpha = 1.0;
  for (int64_t i = 0; i < n; i++) { /* 1733-1740, 1762-1768 */
    if (dx[i] < 0) alpha = fmin(alpha, -X[i] / dx[i]);
    else if (dx[i] > 0) alpha = fmin(alpha, (C - X[i]) / dx[i]);
    if (ds[i] < 0) alpha = fmin(alpha, -S[i] / ds[i]);
    if (dxi[i] < 0) alpha = fmin(alpha, -Xi[i] / dxi[i]);
  }
  return alpha;
}

int orc_mpc(const double *Z, const double *a, int64_t n, int64_t k, double C,
            int kernel_type, int maxit, int stop_mode, double tol, double *X, double *Xi,
            double *S, orc_ipm_info *info) {
  double y = 0.0;
  double *r = (double *)calloc((size_t)(3 * n + 1), sizeof(double));
  double *r_aff = (double *)calloc((size_t)(3 * n + 1), sizeof(double));
  double *q = (double *)malloc(sizeof(double) * k);
  double *e = (double *)malloc(sizeof(double) * n);
  double *D = (double *)malloc(sizeof(double) * n);
  double *M = (double *)malloc(sizeof(double) * k * k);
  for (int64_t i = 0; i < n; i++) e[i] = 1.0;
  if (kernel_type == 2) { /* 1535-1552 */
    gemv_t(Z, n, k, e, q);
    gemv_n(Z, n, k, q, e);
  }
  double qq = 0.0, qe = 0.0;
  for (int64_t i = 0; i < n; i++) { qq += e[i] * e[i]; qe += e[i]; }
  double ox = qe / qq;
  if (ox >= 0.99 * C) ox = 0.99 * C;
  else if (ox <= 0.01 * C) ox = 0.01 * C;
  for (int64_t i = 0; i < n; i++) { X[i] = ox; S[i] = 1.0; Xi[i] = 1.0; }
  if (kernel_type == 0) { /* 1569-1578: zero the initial dual residual */
    gemv_t(Z, n, k, X, q);
    gemv_n(Z, n, k, q, e);
    for (int64_t i = 0; i < n; i++) {
      double t = e[i] - a[i] * y - 1.0;
      if (t > 0) S[i] += t;
      else Xi[i] -= t;
    }
  }
  int iter;
  int reason = ORC_MAXIT;
  double mu = 0, normdx = 0, normdy = 0, pobj = 0, dobj = 0, gap = 0, cond = 0;
  for (iter = 0; iter < maxit; iter++) {
    mu = 0.0;
    for (int64_t i = 0; i < n; i++) mu += X[i] * S[i] + (C - X[i]) * Xi[i];
    mu /= (2.0 * n);
    double *dx = r, *dy = r + n, *ds = r + n + 1, *dxi = r + 2 * n + 1;
    gemv_t(Z, n, k, X, q); /* 1616-1619 */
    gemv_n(Z, n, k, q, e);
    double xqx = 0.0;
    for (int64_t i = 0; i < n; i++) xqx += X[i] * e[i];
    for (int64_t i = 0; i < n; i++) dx[i] = e[i] - a[i] * y - S[i] + Xi[i] - 1.0;
    double ax = 0.0;
    for (int64_t i = 0; i < n; i++) ax += a[i] * X[i];
    dy[0] = -ax;
    for (int64_t i = 0; i < n; i++) ds[i] = -S[i] * X[i];
    for (int64_t i = 0; i < n; i++) dxi[i] = -Xi[i] * (C - X[i]);
    double dmax = -1, dmin = 1e308;
    for (int64_t i = 0; i < n; i++) { /* 1624-1630 */
      D[i] = S[i] / X[i] + Xi[i] / (C - X[i]);
      if (D[i] > dmax) dmax = D[i];
      if (D[i] < dmin) dmin = D[i];
    }
    cond = dmax / dmin;
    normdx = 0.0;
    for (int64_t i = 0; i < n; i++) normdx += dx[i] * dx[i];
    normdx = sqrt(normdx);
    normdy = fabs(dy[0]);
    double tmp = 0.5 * xqx; /* 1636-1643 */
    double sx = 0.0, sxi = 0.0, l1 = 0.0;
    for (int64_t i = 0; i < n; i++) { sx += X[i]; sxi += Xi[i]; l1 += fabs(X[i]); }
    pobj = tmp - sx;
    dobj = -tmp - C * sxi;
    gap = (pobj - dobj) / (1.0 + fabs(pobj));
    if (stop_mode == ORC_STOP_REF) { /* 1647 */
      if (mu < 1e-7 && normdx < 1e-7 && normdy < 1e-7) { reason = ORC_CONVERGED; break; }
    } else {
      if (gap <= tol && normdx / (1.0 + sqrt((double)n)) <= tol &&
          normdy <= tol * (1.0 + l1 / C)) { reason = ORC_CONVERGED; break; }
    }
    if (cond > 1e18) { reason = ORC_ILL_COND; break; } /* 1653 */
    if (build_capacitance(Z, D, n, k, M) != 0) { reason = ORC_CHOL_FAIL; break; }
    newton_step(Z, D, M, C, a, X, S, Xi, r, n, k); /* predictor (1727) */
    double alpha = max_step(r, X, S, Xi, C, n);
    double mu_aff = 0.0;
    for (int64_t i = 0; i < n; i++)
      mu_aff += (X[i] + alpha * dx[i]) * (S[i] + alpha * ds[i]) +
                (C - X[i] - alpha * dx[i]) * (Xi[i] + alpha * dxi[i]);
    mu_aff /= (2.0 * n);
    double sigma = mu_aff / mu;
    sigma = sigma * sigma * sigma;
    memcpy(r_aff, r, sizeof(double) * (3 * n + 1));
    for (int64_t i = 0; i < n; i++) { /* corrector RHS (1750-1757) */
      ds[i] = sigma * mu - r_aff[i] * r_aff[i + n + 1];
      dxi[i] = sigma * mu + r_aff[i] * r_aff[i + 2 * n + 1];
      dx[i] = 0.0;
    }
    dy[0] = 0.0;
    newton_step(Z, D, M, C, a, X, S, Xi, r, n, k); /* corrector (1758) */
    for (int64_t i = 0; i < 3 * n + 1; i++) r[i] += r_aff[i];
    alpha = 0.99 * max_step(r, X, S, Xi, C, n);
    for (int64_t i = 0; i < n; i++) { /* 1771-1777 */
      X[i] += alpha * dx[i];
      S[i] += alpha * ds[i];
      Xi[i] += alpha * dxi[i];
    }
    y += alpha * dy[0];
  }
  if (info) {
    info->iters = iter;
    info->reason = reason;
    info->mu = mu; info->normdx = normdx; info->normdy = normdy;
    info->pobj = pobj; info->dobj = dobj; info->gap_rel = gap; info->condD = cond;
    info->y = y;
  }
  free(r); free(r_aff); free(q); free(e); free(D); free(M);
  return reason;
}

/* ------------------------------------------------------------------------- */
/* Bias and model files (writemodel 273-417).                                  */
/* ------------------------------------------------------------------------- */
double orc_bias_rbf(const double *alpha, const float *labels, const double *U, int64_t n,
                    int64_t k, double C, const double *S, const double *Xi, int *nfree_used,
                    double *std_out) {
  int64_t nSV = 0, nfree = 0;
  for (int64_t i = 0; i < n; i++)
    if (alpha[i] > 1e-6) { nSV++; if (alpha[i] < C - 1e-6) nfree++; }
  int64_t *iSV = (int64_t *)malloc(sizeof(int64_t) * (nSV + 1));
  int64_t *iF = (int64_t *)malloc(sizeof(int64_t) * (nfree + 1));
  int64_t s = 0, f = 0;
  for (int64_t i = 0; i < n; i++)
    if (alpha[i] > 1e-6) {
      iSV[s++] = i;
      /* with slacks given, a free SV must also have s < alpha and xi < C - alpha */
      if (alpha[i] < C - 1e-6 && (!S || (S[i] < alpha[i] && Xi[i] < C - alpha[i]))) iF[f++] = i;
    }
  int64_t ns = f < 50 ? f : 50;
  double acc = 0.0;
  double *bs = (double *)malloc(sizeof(double) * (ns + 1));
  for (int64_t j = 0; j < ns; j++) { /* 355-372 */
    int64_t jj = iF[j];
    double yj = labels[jj];
    for (int64_t i = 0; i < nSV; i++) {
      int64_t ii = iSV[i];
      double sum = 0.0;
      for (int64_t c = 0; c < k; c++) sum += U[ii * k + c] * U[jj * k + c];
      yj -= alpha[ii] * labels[jj] * sum;
    }
    acc += yj;
    bs[j] = yj;
  }
  double b = acc / (double)ns; /* NaN when there is no free SV, as in the reference */
  double sq = 0.0;
  for (int64_t j = 0; j < ns; j++) sq += (bs[j] - b) * (bs[j] - b);
  if (std_out) *std_out = sqrt(sq / (double)ns);
  if (nfree_used) *nfree_used = (int)ns;
  free(iSV); free(iF); free(bs);
  return b;
}

double orc_bias_linear(const double *alpha, const float *labels, const float *inst,
                       int64_t n, int64_t d, double C, double *w) {
  for (int64_t t = 0; t < d; t++) w[t] = 0.0;
  for (int64_t i = 0; i < n; i++) /* 302-308 */
    if (alpha[i] > 1e-6)
      for (int64_t t = 0; t < d; t++) w[t] += alpha[i] * labels[i] * inst[i * d + t];
  int64_t cnt = 0;
  double sum = 0.0;
  for (int64_t i = 0; i < n && cnt < 5000; i++) { /* 310-331 */
    if (alpha[i] > 1e-6 && alpha[i] < C - 1e-6) {
      double tmp = labels[i];
      for (int64_t t = 0; t < d; t++) tmp -= w[t] * inst[i * d + t];
      sum += tmp;
      cnt++;
    }
  }
  return cnt > 0 ? sum / (double)cnt : 0.0;
}

int orc_write_model_rbf(const char *path, const double *alpha, const float *labels,
                        const float *inst, int64_t n, int64_t d, double C, double gamma,
                        double b, int pos, int neg) {
  int nSV = 0, nfree = 0;
  for (int64_t i = 0; i < n; i++)
    if (alpha[i] > 1e-6) { nSV++; if (alpha[i] < C - 1e-6) nfree++; }
  FILE *f = fopen(path, "w");
  if (!f) return -1;
  fprintf(f, "svm_type c_svc\n"); /* 384-396 */
  fprintf(f, "kernel_type rbf\n");
  fprintf(f, "gamma %.7f\n", gamma);
  fprintf(f, "nr_class 2\n");
  fprintf(f, "total_sv %d\n", nSV);
  fprintf(f, "rho %f\n", -b);
  fprintf(f, "label %d %d\n", pos, neg);
  fprintf(f, "nr_sv %d %d\n", nfree, nSV - nfree);
  fprintf(f, "SV\n");
  for (int64_t j = 0; j < n; j++) { /* 397-406 */
    if (!(alpha[j] > 1e-6)) continue;
    fprintf(f, "%7f ", labels[j] * alpha[j]);
    for (int64_t t = 0; t < d; t++) {
      float v = inst[j * d + t];
      if (v > 0 || v < 0) fprintf(f, "%d:%7f ", (int)(t + 1), v);
    }
    fprintf(f, "\n");
  }
  fclose(f);
  return 0;
}

int orc_write_model_linear(const char *path, const double *w, int64_t d, double b, int pos,
                           int neg) {
  FILE *f = fopen(path, "w");
  if (!f) return -1;
  fprintf(f, "solver_type L2R_L1LOSS_SVC_DUAL\n"); /* 340-350 */
  fprintf(f, "nr_class 2\n");
  fprintf(f, "label %d %d\n", pos, neg);
  fprintf(f, "nr_feature %d\n", (int)d);
  fprintf(f, "bias %d\n", 1);
  fprintf(f, "w\n");
  for (int64_t i = 0; i < d; i++) fprintf(f, "%.16f \n", w[i]);
  fprintf(f, "%.16f\n", b);
  fclose(f);
  return 0;
}

void orc_decision(const double *alpha, const float *labels, const float *inst, int64_t n,
                  int64_t d, double gamma, double b, const float *test, int64_t nt,
                  double *f) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t t = 0; t < nt; t++) {
    double acc = b;
    for (int64_t i = 0; i < n; i++) {
      if (!(alpha[i] > 1e-6)) continue;
      double dist = 0.0;
      for (int64_t c = 0; c < d; c++) {
        double df = (double)inst[i * d + c] - (double)test[t * d + c];
        dist += df * df;
      }
      acc += alpha[i] * labels[i] * exp(-gamma * dist);
    }
    f[t] = acc;
  }
}

void orc_fnorm_res(const float *X, const float *y, int64_t n, int64_t d, double gamma,
                   const double *U, int64_t k, double *fnorm, double *fnorm_res) {
  double a1 = 0.0, a2 = 0.0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : a1, a2)
  for (int64_t i = 0; i < n; i++) {
    for (int64_t j = 0; j < n; j++) {
      double dist = 0.0;
      for (int64_t c = 0; c < d; c++) {
        double df = (double)X[i * d + c] - (double)X[j * d + c];
        dist += df * df;
      }
      double kv = (double)y[i] * (double)y[j] * exp(-gamma * dist);
      double uu = 0.0;
      for (int64_t c = 0; c < k; c++) uu += U[i * k + c] * U[j * k + c];
      a1 += kv * kv;
      a2 += (kv - uu) * (kv - uu);
    }
  }
  *fnorm = sqrt(a1);
  *fnorm_res = sqrt(a2);
}
