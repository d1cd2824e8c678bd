// Internal host-side interfaces of libtsvm (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>

#include "../../include/tsvm.h"

namespace tsvm {

// Device workspace: named buffers that grow on demand and are reused across calls.
struct Workspace {
  std::map<std::string, std::pair<void*, size_t>> bufs;
  hipError_t get(const char* name, size_t bytes, void** out);
  void release_all();
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  Workspace ws;
  std::string last_error;
  int cu_count = 256;
  void* rocblas = nullptr;  // rocblas_handle for the k x k eigen-solver, created on first use
  // phase timers of the last tsvm_lra / tsvm_ipm call (seconds)
  double t_phase[16] = {0};
};

// Record a failure on the context and return its status.
tsvm_status fail(Ctx* c, tsvm_status st, const char* fmt, ...);
#define TSVM_HIP(ctx, expr)                                                          \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess)                                                            \
      return ::tsvm::fail((ctx), TSVM_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__,  \
                          #expr, hipGetErrorString(_e));                             \
  } while (0)
#define TSVM_CHECK(expr)                     \
  do {                                       \
    tsvm_status _s = (expr);                 \
    if (_s != TSVM_OK) return _s;            \
  } while (0)

// ---- kernel product (kprod.hip) -------------------------------------------------
// W[i, c] = y1_i * sum_j y2_j exp(-gamma |x1_i - x2_j|^2) A[j, c], all device pointers.
// X1 m x d, X2 n x d row-major; A n x r with leading dimension lda; W m x r, ldw.
tsvm_status kernel_product_dev(Ctx* c, const float* X1, const float* y1, int64_t m,
                               const float* X2, const float* y2, int64_t n, int64_t d,
                               const float* A, int64_t lda, int64_t r, double gamma,
                               int precision, float* W, int64_t ldw);

// ---- dense tall-skinny GEMMs (gemm.hip) -----------------------------------------
// C (ka x kb, FP64, row-major) = A^T diag(w) B over n rows; A n x ka (lda), B n x kb
// (ldb) FP32 row-major; w FP64 or nullptr (= ones).  lower_only: only blocks with
// block-row >= block-col are computed (for SYRK; the caller mirrors).
tsvm_status gemm_tn_f64(Ctx* c, const float* A, int64_t lda, int64_t ka, const float* B,
                        int64_t ldb, int64_t kb, int64_t n, const double* w, bool lower_only,
                        double* C);
// Y (n x kb, FP32, ldy) = A (n x ka FP32, lda) * B (ka x kb FP64 row-major).
tsvm_status gemm_nn_tall(Ctx* c, const float* A, int64_t lda, int64_t ka, const double* B,
                         int64_t kb, int64_t n, float* Y, int64_t ldy);

// ---- small dense FP64 linear algebra on the device (chol.hip) --------------------
// In-place lower Cholesky of the k x k row-major matrix M (device); the strict upper
// triangle is zeroed.  info_dev (device int) receives 0 or the 1-based failing column.
tsvm_status chol_f64(Ctx* c, double* M, int64_t k, int* info_dev);
// Solve (L L^T) X = B for p right-hand sides; B is p x k row-major (one RHS per row),
// overwritten by X.  Skipped (B unchanged) when *info_dev != 0.
tsvm_status chol_solve_f64(Ctx* c, const double* L, int64_t k, double* B, int64_t p,
                           const int* info_dev);

// ---- elementwise helpers (util.hip) -----------------------------------------------
// v (host, k) = sum_i m_i U[i, :], m_i = alpha_i if alpha_i > thr else 0 (device U, alpha)
tsvm_status masked_colsum(Ctx* c, const float* U, int64_t n, int64_t k, int64_t ldu,
                          const double* alpha, double thr, double* v_host);

// ---- host small dense FP64 linear algebra (host_la.cpp) ---------------------------
// Symmetric eigen-decomposition (ascending), V column i = eigenvector i, row-major.
tsvm_status host_sym_eig(Ctx* c, const double* A, int64_t k, double* w, double* V);
const char* host_lapack_name();
int host_chol(double* M, int64_t k);  // lower, in place; 0 or failing column + 1
void host_tri_inv_lower(const double* L, int64_t k, double* Linv);
void host_gemm(const double* A, const double* B, int64_t k, double* C);  // k x k

// ---- factor construction (lra.cpp) ------------------------------------------------
// U (n x k, FP32, device) with U U^T ~ K_Y; Om (n x k, device) is consumed (reused as
// workspace).  Fills info (may be null).
tsvm_status lra_dev(Ctx* c, const float* X, const float* y, int64_t n, int64_t d, int64_t k,
                    double gamma, int precision, float* Om, float* U, tsvm_lra_info* info);

// ---- interior-point solver (ipm.hip) ---------------------------------------------
tsvm_status ipm_dev(Ctx* c, const float* Z, const float* a, int64_t n, int64_t k,
                    const tsvm_ipm_params* prm, double* x, double* xi, double* s,
                    tsvm_ipm_info* info);

}  // namespace tsvm

// the opaque C-ABI handle
struct tsvm_ctx {
  tsvm::Ctx c;
};
