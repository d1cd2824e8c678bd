// Internal host-side interfaces of libtsvm (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

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
  void* rocblas = nullptr;  // rocblas_handle of the k x k eigen-solver (set up with the context)
  // phase timers of the last tsvm_lra / tsvm_ipm call (seconds)
  double t_phase[16] = {0};
  int progress = 0;  // progress lines on stderr during a training call
  void* scratch = nullptr;  // this device's copy of the allocation-free product's scratch
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

// A HIP event destroyed with its scope, so that early returns (TSVM_HIP / TSVM_CHECK
// inside a timed region) do not leak it.
struct Event {
  hipEvent_t e = nullptr;
  Event() = default;
  Event(const Event&) = delete;
  Event& operator=(const Event&) = delete;
  ~Event() {
    if (e) (void)hipEventDestroy(e);
  }
  hipError_t create() { return hipEventCreate(&e); }
  operator hipEvent_t() const { return e; }
};

// ---- kernel product (kprod.hip) -------------------------------------------------
// W[i, c] = y1_i * sum_j y2_j exp(-gamma |x1_i - x2_j|^2) A[j, c], all device pointers.
// X1 m x d, X2 n x d row-major; A n x r with leading dimension lda; W m x r, ldw.
tsvm_status kernel_product_dev(Ctx* c, const float* X1, const float* y1, int64_t m,
                               const float* X2, const float* y2, int64_t n, int64_t d,
                               const float* A, int64_t lda, int64_t r, double gamma,
                               int precision, float* W, int64_t ldw);

// The same product without allocating: every operand buffer is carved from the device's
// fixed scratch array (TSVM_NOALLOC_SCRATCH_BYTES); lda = r, ldw = r.
tsvm_status kernel_product_noalloc_dev(Ctx* c, const float* X1, const float* y1, int64_t m,
                                       const float* X2, const float* y2, int64_t n, int64_t d,
                                       const float* A, int64_t r, double gamma, int precision,
                                       float* W);
// Looks up (once per device, with the device current) the address of that device's copy
// of the scratch array and stores it in c->scratch.
tsvm_status noalloc_scratch_setup(Ctx* c);

// ---- dense tall-skinny GEMMs (gemm.hip) -----------------------------------------
// C (ka x kb, FP64, row-major) = A^T diag(w) B over n rows; A n x ka (lda), B n x kb
// (ldb) FP32 row-major; w FP64 or nullptr (= ones).  lower_only: only blocks with
// block-row >= block-col are computed (for SYRK; the caller mirrors).  rows (device,
// optional): the n rows taken are rows[0..n) of A, B and w (read in place) instead of 0..n.
tsvm_status gemm_tn_f64(Ctx* c, const float* A, int64_t lda, int64_t ka, const float* B,
                        int64_t ldb, int64_t kb, int64_t n, const double* w, bool lower_only,
                        double* C, const int* rows = nullptr);

// ---- factor construction helpers (lra.cpp) ---------------------------------------
// Shifted Cholesky-QR of the n x k FP32 block W in place (tmp: n x k scratch); G receives
// the Gram of the basis as stored (k x k, host), passes the number of QR passes taken.
tsvm_status cholqr_orth(Ctx* c, float* W, float* tmp, int64_t n, int64_t k,
                        std::vector<double>& G, int* passes);

// ---- weighted Gram M = I + Z^T diag(w) Z of the capacitance (syrk.hip) -----------
// Hybrid form: rows split by t_i = w_i |z_i|^2 into tiers, each under an error budget:
// a dense pass over every row (BF16, or FP8 with the FP8 tier, for the smallest t),
// compensated hi + lo bf16 rows (kpt >= 256, gathered through an index list), and FP64
// rows (the largest t, through an index list).  Which dense pass runs, if any, is chosen
// per iteration by a per-row cost model.
struct HybridSyrk {
  int64_t n = 0, n_pad = 0;
  int kp = 0, kpt = 0;
  void* ZbT = nullptr;     // kpt x n_pad bf16, Z^T rounded once
  double* nz = nullptr;    // |z_i|^2
  float* wb = nullptr;     // FP32 weights of the bulk rows (0 for FP64 rows), n_pad
  int* rows = nullptr;     // indices of the FP64 rows
  int* rows2 = nullptr;    // indices of the compensated rows
  bool x2 = false;         // compensated tier available (kpt >= 256)
  int* ints = nullptr;     // split scratch (histograms, counts, offsets, cut, count)
  // FP8 tier (fp8_eps > 0): the smallest-t rows on the FP8 MFMA with per-column scales
  double fp8_eps = 0.0;    // error budget of the tier (0: no FP8 tier)
  float* w8 = nullptr;     // sqrt(w_i) of the FP8 rows, 0 elsewhere (n_pad)
  float* zmax = nullptr;   // max_i |z_ic| per column (kpt, set up once per solve)
  float* sc8 = nullptr;    // per-column power-of-two scales of the iteration (kpt)
  void* Z8 = nullptr;      // e4m3 of sc8_c sqrt(w_i) z_ic (n_pad x kpt, chunked as ZbT)
  double r8 = 0.0;         // cost of the FP8 pass (packing included) / the BF16 pass
  double r2 = 0.0;         // cost of one compensated row / one BF16-pass row
  double r64 = 0.0;        // cost of one FP64 row / one BF16-pass row
};
// rows per tier of one split: FP64, FP8 pass, compensated, BF16 pass
struct HybridCounts {
  int64_t nh = 0, n8 = 0, n2 = 0, nb = 0;
};
constexpr int HYBRID_NSLOT = 6;  // device doubles written by hybrid_syrk_split
// Workspace and the once-per-solve operands (ZbT, nz) for Z (n x kp FP32, device).
// fp8_eps > 0 (kp > 128 only) adds the FP8 tier with that error budget.
tsvm_status hybrid_syrk_setup(Ctx* c, const float* Z, int64_t n, int kp, HybridSyrk* hs,
                              double fp8_eps = 0.0);
// error budget of the FP8 tier of TSVM_PREC_FP8
double hybrid_fp8_eps();
// Splits the rows for weights w (device) with the BF16 tier's budget (the compensated
// tier's is 2^7 times it): HYBRID_NSLOT device doubles, slots[0] = bound on sqrt(sum t_i^2)
// over the BF16-pass rows, [1] = number of FP64 rows, [2] = error bound of the FP8-pass
// rows, [3] = their number (0 unless the iteration takes the FP8 pass), [4] = bound on
// sqrt(sum t_i^2) over the compensated rows, [5] = their number; fills rows, rows2, wb
// (and w8, sc8).
tsvm_status hybrid_syrk_split(Ctx* c, HybridSyrk* hs, const double* w, double budget,
                              double* slots);
// the tier sizes from the slots (host copy)
HybridCounts hybrid_counts(const HybridSyrk* hs, const double* slots);
// M = I + FP64 part over the listed rows + the dense pass (FP8 when hc.n8 > 0, BF16 when
// hc.nb > 0) + the compensated part.  ev (optional, three events) marks the start of the
// low-precision parts, the end of the dense pass and the end of the compensated part.
tsvm_status hybrid_syrk_form(Ctx* c, HybridSyrk* hs, const float* Z, const double* w,
                             const HybridCounts& hc, double* M, hipEvent_t* ev);
// flops of the low-precision parts, SYRK convention (rows formed x k (k + 1))
double hybrid_syrk_flops(const HybridSyrk* hs, const HybridCounts& hc);
// MFMA flops the dense pass issues (every row, lower tiles with whole diagonal tiles)
double hybrid_dense_mfma_flops(const HybridSyrk* hs, const HybridCounts& hc);
// tsvm_fnorm_res_dev (fnorm.hip): |K_Y|_F, |K_Y - U U^T|_F; U has row stride ldu
tsvm_status fnorm_res_dev(Ctx* c, const float* X, const float* y, int64_t n, int64_t d,
                          double gamma, const float* U, int64_t ldu, int64_t k, double* fnorm,
                          double* fnorm_res);
// tsvm_grad_dev (grad.hip)
tsvm_status grad_dev(Ctx* c, const float* X, const float* y, int64_t n, int64_t d,
                     const double* alpha, double b, int64_t k, double log_gamma,
                     const double* log_scales, const float* Om, const float* Xt,
                     const float* yt, int64_t nt, int precision, double* vals, double* gD,
                     double* gL, tsvm_grad_info* info);
// tsvm_weighted_syrk_dev (synchronises before returning)
tsvm_status weighted_syrk_dev(Ctx* c, const float* Z, int64_t n, int64_t k, const double* w,
                              int precision, double budget, double* M, tsvm_syrk_info* info);
// Y (n x kb, FP32, ldy) = A (n x ka FP32, lda) * B (ka x kb FP64 row-major).
tsvm_status gemm_nn_tall(Ctx* c, const float* A, int64_t lda, int64_t ka, const double* B,
                         int64_t kb, int64_t n, float* Y, int64_t ldy);

// ---- small dense FP64 linear algebra on the device (chol.hip) --------------------
// In-place lower Cholesky of the k x k row-major matrix M (device); the strict upper
// triangle is zeroed.  info_dev (device int) receives 0 or the 1-based failing column.
tsvm_status chol_f64(Ctx* c, double* M, int64_t k, int* info_dev);
// Solve (L L^T) X = B for p right-hand sides; B is p x k row-major (one RHS per row),
// overwritten by X.  Skipped (B unchanged) when *info_dev != 0.  Uses the inverted
// diagonal blocks and the transposed factor that the last chol_f64 on the context left
// in its workspace.
tsvm_status chol_solve_f64(Ctx* c, const double* L, int64_t k, double* B, int64_t p,
                           const int* info_dev);

// ---- elementwise helpers (util.hip) -----------------------------------------------
// v (host, k) = sum_i m_i U[i, :], m_i = alpha_i if alpha_i > thr else 0 (device U, alpha)
tsvm_status masked_colsum(Ctx* c, const float* U, int64_t n, int64_t k, int64_t ldu,
                          const double* alpha, double thr, double* v_host);

// ---- small dense FP64 linear algebra, k x k (host_la.cpp) ---------------------------
// Symmetric eigen-decomposition (ascending), V column i = eigenvector i, row-major; A, w
// and V on the host.
tsvm_status host_sym_eig(Ctx* c, const double* A, int64_t k, double* w, double* V);
void warm_dense_la(Ctx* c);  // first-use set-up of the k x k library calls (context creation)
const char* host_lapack_name();
// dB (device, row-major) = L^{-T}, upper triangular, for A + shift I = L L^T (A symmetric,
// row-major, on the host).  0; j > 0 if the leading minor of order j is not positive
// definite; -1 if a library call failed.  Synchronises the stream.
int dev_chol_inv_t(Ctx* c, const double* A, int64_t k, double shift, double* dB);
// C = op(A) op(B), row-major k x k device matrices (op = transpose where the flag is set)
tsvm_status dev_gemm_rm(Ctx* c, bool ta, bool tb, const double* A, const double* B, int64_t k,
                        double* C);

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
