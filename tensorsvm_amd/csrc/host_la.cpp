// Small dense FP64 linear algebra of the factor construction (k x k, k <= a few
// thousand): Cholesky with the triangular inverse, products, and the symmetric
// eigen-solver, on the GPU through rocSOLVER / rocBLAS on the context's stream.
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
#include <vector>

#include "internal.h"

namespace tsvm {

namespace {

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
// the context's rocBLAS handle (created on first use, bound to the context's stream)
static rocblas_handle blas_handle(Ctx* c) {
  if (!c->rocblas) {
    rocblas_handle h;
    if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
    c->rocblas = (void*)h;
    rocblas_set_stream(h, c->stream);
  }
  return (rocblas_handle)c->rocblas;
}

static bool gpu_sym_eig(Ctx* c, const double* A, int64_t k, double* w, double* V) {
  if (!blas_handle(c)) return false;
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

// A row-major k x k matrix read column-major is its transpose, so a symmetric matrix with
// its row-major strict lower part zeroed is, for rocSOLVER, the lower triangle of the same
// matrix with zeros above.  dpotrf (lower) then dtrtri (lower) leave L^{-1} in the column-
// major lower triangle, which read row-major is L^{-T}, upper triangular, the zeros below
// untouched (both routines reference only the lower triangle).
int dev_chol_inv_t(Ctx* c, const double* A, int64_t k, double shift, double* dB) {
  rocblas_handle h = blas_handle(c);
  void* dI;
  if (!h || c->ws.get("chol_inv_info", 16, &dI) != hipSuccess) return -1;
  std::vector<double> M(k * k);
  for (int64_t i = 0; i < k; i++)
    for (int64_t j = 0; j < k; j++) M[i * k + j] = j < i ? 0.0 : A[i * k + j] + (i == j ? shift : 0.0);
  rocblas_int info[2] = {-1, -1};
  rocblas_int* di = (rocblas_int*)dI;
  // every path waits for the stream before returning: M and info are read by queued copies
  bool ok = hipMemcpyAsync(dB, M.data(), k * k * 8, hipMemcpyHostToDevice, c->stream) == hipSuccess &&
            rocsolver_dpotrf(h, rocblas_fill_lower, (rocblas_int)k, dB, (rocblas_int)k, di) ==
                rocblas_status_success &&
            hipMemcpyAsync(&info[0], di, sizeof(rocblas_int), hipMemcpyDeviceToHost, c->stream) ==
                hipSuccess;
  if (hipStreamSynchronize(c->stream) != hipSuccess || !ok) return -1;
  if (info[0] != 0) return (int)info[0];  // leading minor of order info[0] not positive definite
  ok = rocsolver_dtrtri(h, rocblas_fill_lower, rocblas_diagonal_non_unit, (rocblas_int)k, dB,
                        (rocblas_int)k, di + 1) == rocblas_status_success &&
       hipMemcpyAsync(&info[1], di + 1, sizeof(rocblas_int), hipMemcpyDeviceToHost, c->stream) ==
           hipSuccess;
  if (hipStreamSynchronize(c->stream) != hipSuccess || !ok || info[1] != 0) return -1;
  return 0;
}

// Row-major C = op(A) op(B) is column-major C^T = op(B)^T op(A)^T: the same operands in the
// other order, each with its own transpose flag.
tsvm_status dev_gemm_rm(Ctx* c, bool ta, bool tb, const double* A, const double* B, int64_t k,
                        double* C) {
  rocblas_handle h = blas_handle(c);
  if (!h) return fail(c, TSVM_ERR_HIP, "rocBLAS handle could not be created");
  const double one = 1.0, zero = 0.0;
  const rocblas_int kk = (rocblas_int)k;
  if (rocblas_dgemm(h, tb ? rocblas_operation_transpose : rocblas_operation_none,
                    ta ? rocblas_operation_transpose : rocblas_operation_none, kk, kk, kk, &one, B,
                    kk, A, kk, &zero, C, kk) != rocblas_status_success)
    return fail(c, TSVM_ERR_HIP, "rocblas_dgemm (k = %ld) failed", (long)k);
  return TSVM_OK;
}

// The k x k library calls of the factor construction once at k = 512 (an eigen-
// decomposition of a tridiagonal matrix, a Cholesky with its triangular inverse, and the
// two product forms), so that the rocBLAS handle and the libraries' kernels (about 0.4 s
// the first time in a process) are set up when the context is created rather than inside
// the first factor construction.  The buffers are the factor construction's own.
void warm_dense_la(Ctx* c) {
  const int64_t k = 512;
  std::vector<double> A(k * k, 0.0), w(k), V(k * k);
  for (int64_t i = 0; i < k; i++) {
    A[i * k + i] = (double)(i + 1);
    if (i + 1 < k) A[i * k + i + 1] = A[(i + 1) * k + i] = 0.5;
  }
  (void)gpu_sym_eig(c, A.data(), k, w.data(), V.data());
  void *pLt, *pA, *pT;
  if (c->ws.get("lra_Lt", (size_t)(k * k * 8), &pLt) != hipSuccess ||
      c->ws.get("lra_A", (size_t)(k * k * 8), &pA) != hipSuccess ||
      c->ws.get("lra_T", (size_t)(k * k * 8), &pT) != hipSuccess)
    return;
  if (dev_chol_inv_t(c, A.data(), k, 0.0, (double*)pLt) != 0) return;
  if (hipMemcpyAsync(pA, A.data(), k * k * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess) return;
  const tsvm_status st1 = dev_gemm_rm(c, true, false, (double*)pLt, (double*)pA, k, (double*)pT);
  const tsvm_status st2 = dev_gemm_rm(c, false, false, (double*)pT, (double*)pLt, k, (double*)pA);
  (void)hipStreamSynchronize(c->stream);
  if (st1 != TSVM_OK || st2 != TSVM_OK) c->last_error.clear();
}

}  // namespace tsvm
