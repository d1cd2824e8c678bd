// Small dense FP64 linear algebra on the host for the factor construction (k x k,
// k <= a few thousand): Cholesky, triangular inverse, and the symmetric eigen-solver.
//
// The reference computes the core eigen-decomposition on the host with LAPACKE_dsyevd
// (tensorsvm-train.cu:2191).  Here it is rocSOLVER's dsyevd on the GPU (k x k, once per
// factor); a cyclic Jacobi solver of our own is the fallback.
#include <math.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "internal.h"

namespace tsvm {

namespace {

// parallel-for over [0, n) on up to nt threads (coarse chunks)
template <class F>
void pfor(int64_t n, F f) {
  int nt = (int)std::thread::hardware_concurrency();
  if (nt > 16) nt = 16;
  if (n < 256 || nt <= 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; t++) {
    int64_t a = t * chunk, b = a + chunk < n ? a + chunk : n;
    if (a >= b) break;
    th.emplace_back([=]() { f(a, b); });
  }
  for (auto& x : th) x.join();
}

void jacobi_eig(const double* A0, int64_t k, double* w, double* V) {
  std::vector<double> A(A0, A0 + k * k);
  for (int64_t i = 0; i < k * k; i++) V[i] = 0.0;
  for (int64_t i = 0; i < k; i++) V[i * k + i] = 1.0;
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int64_t p = 0; p < k; p++)
      for (int64_t q = 0; q < k; q++) {
        double v = A[p * k + q] * A[p * k + q];
        tot += v;
        if (p != q) off += v;
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int64_t p = 0; p < k - 1; p++)
      for (int64_t q = p + 1; q < k; q++) {
        double apq = A[p * k + q];
        if (fabs(apq) < 1e-300) continue;
        double theta = (A[q * k + q] - A[p * k + p]) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int64_t r = 0; r < k; r++) {
          double arp = A[r * k + p], arq = A[r * k + q];
          A[r * k + p] = c * arp - s * arq;
          A[r * k + q] = s * arp + c * arq;
        }
        for (int64_t r = 0; r < k; r++) {
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
  std::vector<int64_t> idx(k);
  for (int64_t i = 0; i < k; i++) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return A[a * k + a] < A[b * k + b]; });
  std::vector<double> Vs(k * k);
  for (int64_t c = 0; c < k; c++) {
    w[c] = A[idx[c] * k + idx[c]];
    for (int64_t r = 0; r < k; r++) Vs[r * k + c] = V[r * k + idx[c]];
  }
  memcpy(V, Vs.data(), sizeof(double) * k * k);
}

}  // namespace

const char* host_lapack_name() { return "rocsolver_dsyevd"; }

// Symmetric eigen-decomposition of the k x k core matrix on the GPU (rocSOLVER dsyevd on
// the context's stream); false if the library call fails.
static bool gpu_sym_eig(Ctx* c, const double* A, int64_t k, double* w, double* V) {
  if (!c->rocblas) {
    rocblas_handle h;
    if (rocblas_create_handle(&h) != rocblas_status_success) return false;
    c->rocblas = (void*)h;
    rocblas_set_stream(h, c->stream);
  }
  void *dA, *dD, *dE, *dI;
  if (c->ws.get("eig_A", (size_t)(k * k * 8), &dA) != hipSuccess ||
      c->ws.get("eig_D", (size_t)(k * 8), &dD) != hipSuccess ||
      c->ws.get("eig_E", (size_t)(k * 8), &dE) != hipSuccess ||
      c->ws.get("eig_I", 16, &dI) != hipSuccess)
    return false;
  // a symmetric row-major matrix is its own column-major form
  int info = -1;
  bool ok = hipMemcpyAsync(dA, A, k * k * 8, hipMemcpyHostToDevice, c->stream) == hipSuccess &&
            rocsolver_dsyevd((rocblas_handle)c->rocblas, rocblas_evect_original,
                             rocblas_fill_lower, (rocblas_int)k, (double*)dA, (rocblas_int)k,
                             (double*)dD, (double*)dE, (rocblas_int*)dI) == rocblas_status_success &&
            hipMemcpyAsync(w, dD, k * 8, hipMemcpyDeviceToHost, c->stream) == hipSuccess &&
            hipMemcpyAsync(&info, dI, 4, hipMemcpyDeviceToHost, c->stream) == hipSuccess;
  std::vector<double> a(k * k);
  ok = ok && hipMemcpyAsync(a.data(), dA, k * k * 8, hipMemcpyDeviceToHost, c->stream) == hipSuccess &&
       hipStreamSynchronize(c->stream) == hipSuccess;
  if (!ok || info != 0) return false;
  // column j of the column-major result = eigenvector j; V[:, j] row-major
  for (int64_t r = 0; r < k; r++)
    for (int64_t j = 0; j < k; j++) V[r * k + j] = a[r + j * k];
  return true;
}

// the host Jacobi solver is the fallback if the library call fails
tsvm_status host_sym_eig(Ctx* c, const double* A, int64_t k, double* w, double* V) {
  if (!gpu_sym_eig(c, A, k, w, V)) jacobi_eig(A, k, w, V);
  return TSVM_OK;
}

// One eigen-decomposition of a 512 x 512 tridiagonal matrix, so that the rocBLAS handle
// and rocSOLVER's kernels (about 0.17 s the first time in a process) are set up when the
// context is created rather than inside the first factor construction.
void warm_sym_eig(Ctx* c) {
  const int64_t k = 512;
  std::vector<double> A(k * k, 0.0), w(k), V(k * k);
  for (int64_t i = 0; i < k; i++) {
    A[i * k + i] = (double)(i + 1);
    if (i + 1 < k) A[i * k + i + 1] = A[(i + 1) * k + i] = 0.5;
  }
  (void)gpu_sym_eig(c, A.data(), k, w.data(), V.data());
}

int host_chol(double* M, int64_t k) {
  const int64_t B = 64;
  for (int64_t j0 = 0; j0 < k; j0 += B) {
    const int64_t jb = j0 + B < k ? B : k - j0;
    for (int64_t j = j0; j < j0 + jb; j++) {  // diagonal block, unblocked
      double s = M[j * k + j];
      for (int64_t t = j0; t < j; t++) s -= M[j * k + t] * M[j * k + t];
      if (!(s > 0.0)) return (int)(j + 1);
      double l = sqrt(s);
      M[j * k + j] = l;
      for (int64_t i = j + 1; i < j0 + jb; i++) {
        double v = M[i * k + j];
        for (int64_t t = j0; t < j; t++) v -= M[i * k + t] * M[j * k + t];
        M[i * k + j] = v / l;
      }
    }
    const int64_t r0 = j0 + jb;
    if (r0 >= k) break;
    pfor(k - r0, [&](int64_t a, int64_t b) {  // panel: L_i = A_i L_jj^{-T}
      for (int64_t i = r0 + a; i < r0 + b; i++) {
        double* Mi = M + i * k;
        for (int64_t j = j0; j < j0 + jb; j++) {
          double v = Mi[j];
          const double* Mj = M + j * k;
          for (int64_t t = j0; t < j; t++) v -= Mi[t] * Mj[t];
          Mi[j] = v / Mj[j];
        }
      }
    });
    pfor(k - r0, [&](int64_t a, int64_t b) {  // trailing update (lower part)
      for (int64_t i = r0 + a; i < r0 + b; i++) {
        double* Mi = M + i * k;
        for (int64_t l = r0; l <= i; l++) {
          const double* Ml = M + l * k;
          double v = 0.0;
          for (int64_t t = j0; t < j0 + jb; t++) v += Mi[t] * Ml[t];
          Mi[l] -= v;
        }
      }
    });
  }
  for (int64_t i = 0; i < k; i++)
    for (int64_t j = i + 1; j < k; j++) M[i * k + j] = 0.0;
  return 0;
}

// Linv = L^{-1} (both lower, row-major): column c solves L x = e_c
void host_tri_inv_lower(const double* L, int64_t k, double* Linv) {
  for (int64_t i = 0; i < k * k; i++) Linv[i] = 0.0;
  pfor(k, [&](int64_t a, int64_t b) {
    std::vector<double> x(k);
    for (int64_t c = a; c < b; c++) {
      for (int64_t i = 0; i < c; i++) x[i] = 0.0;
      for (int64_t i = c; i < k; i++) {
        double v = (i == c) ? 1.0 : 0.0;
        const double* Li = L + i * k;
        for (int64_t t = c; t < i; t++) v -= Li[t] * x[t];
        x[i] = v / Li[i];
      }
      for (int64_t i = c; i < k; i++) Linv[i * k + c] = x[i];
    }
  });
}

// C = A * B (k x k, row-major), parallel over rows
void host_gemm(const double* A, const double* B, int64_t k, double* C) {
  pfor(k, [&](int64_t a, int64_t b) {
    for (int64_t i = a; i < b; i++) {
      double* Ci = C + i * k;
      for (int64_t j = 0; j < k; j++) Ci[j] = 0.0;
      for (int64_t t = 0; t < k; t++) {
        double ait = A[i * k + t];
        if (ait == 0.0) continue;
        const double* Bt = B + t * k;
        for (int64_t j = 0; j < k; j++) Ci[j] += ait * Bt[j];
      }
    }
  });
}

}  // namespace tsvm
