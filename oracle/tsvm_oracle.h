/*
 * TEST INFRASTRUCTURE ONLY.  CPU oracle for the TensorSVM training hot path.
 *
 * A plain-C restatement of the reference trainer's algorithm (tensorsvm-train.cu),
 * used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  Nothing in the product (tensorsvm_amd/) links, loads or calls it.
 *
 * Arithmetic: FP64 throughout (the reference's plain method keeps the factor in
 * FP32 and the SMW/IPM in FP64; the oracle is at least that accurate).
 * Layout: every matrix is row-major, 64-bit sizes.
 */
#ifndef TSVM_ORACLE_H
#define TSVM_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Omega fill: the reference's gaussrand() (tensorsvm-train.cu:2621-2643) called
 * count times in the row-major order of the fill loop (1966-1974), starting from
 * glibc's never-seeded rand() state.  Uses glibc rand() itself (srand(1) first). */
void orc_gaussrand_fill(float *out, int64_t count);
/* First count raw glibc rand() values from the default (seed 1) state. */
void orc_glibc_rand(int32_t *out, int64_t count);

/* LIBSVM reader (1162-1278).  Returns 0 on success, -1 open failure, -2 format error
 * (the reference prints and returns silently; here the error is a status).
 * d = max_index - min_index + 1 (equals the reference's d for 1-based files). */
int orc_libsvm_read(const char *path, int scale256, float **labels, float **inst,
                    int64_t *n, int64_t *d);
void orc_free(void *p);
/* Label mapping (879-896): first label -> +1 (POS), first different -> -1 (NEG).
 * Returns the number of labels outside {POS, NEG} (left unchanged). */
int64_t orc_label_map(float *labels, int64_t n, int *pos, int *neg);

/* Labelled RBF kernel product (rbf_kermatmul 2476-2567 with vecnorm 2569-2581 and
 * rbf_kergen 2598-2612):  W[i,c] = y1_i * sum_j y2_j exp(-gamma |x1_i - x2_j|^2) A[j,c]
 * X1 m x d, X2 n x d, A n x r, W m x r (overwritten).  FP64 accumulation. */
void orc_kernel_product(const float *X1, const float *y1, int64_t m,
                        const float *X2, const float *y2, int64_t n, int64_t d,
                        const double *A, int64_t r, double gamma, double *W);

/* The same product in the reference's FP32 arithmetic (float norms, float XIJ, float
 * distance XI + XJ - 2 XIJ, expf, float accumulation over all n columns in index order).
 * A n x r and W m x r are float. */
void orc_kernel_product_f32(const float *X1, const float *y1, int64_t m, const float *X2,
                            const float *y2, int64_t n, int64_t d, const float *A, int64_t r,
                            double gamma, float *W);

/* Thin Q of a Householder QR (replaces TSQR ortho/CAQR/QR 617-856): W (n x k) <- Q. */
void orc_householder_q(double *W, int64_t n, int64_t k);
/* Symmetric eigen-decomposition by cyclic Jacobi (replaces LAPACKE_dsyevd 2191):
 * A (k x k, symmetric) -> eigenvalues w ascending, V row-major with V[:,i] the
 * eigenvector of w[i]. */
void orc_sym_eig(const double *A, int64_t k, double *w, double *V);

/* Low-rank factor (LRA 1941-2249, default C_EIG path):
 *   W = K_Y Omega; W = orth(W); Q = K_Y W; C = W^T Q (symmetrised from the lower
 *   triangle, 2152-2156); C = V diag(w) V^T; U = W V diag(sqrt(max(w,0))).
 * U n x k.  eig_out (may be NULL) receives the k eigenvalues; returns realk. */
int64_t orc_lra(const float *X, const float *y, int64_t n, int64_t d, int64_t k,
                double gamma, const float *omega, double *U, double *eig_out);
/* The same with a choice of storage precision:
 *   ORC_LRA_FP64      every intermediate in FP64 (orc_lra);
 *   ORC_LRA_REF_FP32  the reference's plain-method arithmetic: both kernel products in
 *                     FP32 (orc_kernel_product_f32), the orthonormal basis rounded to
 *                     FP32 (float W), C = W^T Q accumulated in FP64 and stored as float
 *                     (2143-2157), X = V sqrt(w) as float (2187, 2208) and U = W X as an
 *                     FP32 product (2213-2241).  The orthonormaliser is the FP64
 *                     Householder QR (the reference's FP32 TSQR rounding is not
 *                     restated). */
enum { ORC_LRA_FP64 = 0, ORC_LRA_REF_FP32 = 1 };
int64_t orc_lra_ex(const float *X, const float *y, int64_t n, int64_t d, int64_t k,
                   double gamma, const float *omega, int mode, double *U, double *eig_out);

/* Interior-point solver (mpc 1507-1785, NewtonStep 1794-1822, SMWSolve 1824-1840). */
enum { ORC_STOP_REF = 0, ORC_STOP_GAP = 1 };
enum { ORC_CONVERGED = 0, ORC_ILL_COND = 1, ORC_MAXIT = 2, ORC_CHOL_FAIL = 3 };
typedef struct {
  int iters;
  int reason;      /* ORC_CONVERGED / ORC_ILL_COND / ORC_MAXIT / ORC_CHOL_FAIL */
  double mu, normdx, normdy, pobj, dobj, gap_rel, condD;
  double y;        /* the equality multiplier (= rho at a free SV) */
} orc_ipm_info;
/* Z n x k row-major (labelled factor), a labels (+-1).  kernel_type 2 uses the RBF
 * initial point (1535-1552), 0 the linear one (1553-1579).  stop_mode ORC_STOP_REF
 * is the reference rule (mu, |r_x|, |r_y| < 1e-7, 1647); ORC_STOP_GAP stops when
 * (pobj-dobj)/(1+|pobj|) <= tol, |r_x|/(1+sqrt n) <= tol and |r_y| <= tol(1+|x|_1/C). */
int orc_mpc(const double *Z, const double *a, int64_t n, int64_t k, double C,
            int kernel_type, int maxit, int stop_mode, double tol, double *X,
            double *Xi, double *S, orc_ipm_info *info);
/* One SMW solve (1824-1840): b <- (D + Z Z^T)^{-1} b, M = chol(I + Z^T D^-1 Z). */
void orc_smw_solve(const double *Z, const double *D, int64_t n, int64_t k, double *b);

/* One NewtonStep (1794-1822) at (X, S, Xi): r = [r_x (n); r_y (1); r_s (n); r_xi (n)]
 * is overwritten with the step (dx; dy; ds; dxi). */
void orc_newton_direction(const double *Z, const double *a, int64_t n, int64_t k, double C,
                          const double *X, const double *S, const double *Xi, double *r);
/* Bias (writemodel 353-372): mean over the first <= 50 free SVs of
 * y_j - sum_i alpha_i y_j (U_i . U_j).  S and Xi NULL: the reference's free set
 * (1e-6 < alpha < C - 1e-6); given the IPM's final slacks, a free SV must also have
 * S < alpha and Xi < C - alpha (the trainer's rule).  nfree_used receives the sample
 * count. */
double orc_bias_rbf(const double *alpha, const float *labels, const double *U,
                    int64_t n, int64_t k, double C, const double *S, const double *Xi,
                    int *nfree_used, double *std_out);
/* Bias of the linear branch (301-331). */
double orc_bias_linear(const double *alpha, const float *labels, const float *inst,
                       int64_t n, int64_t d, double C, double *w_out);
/* Model writers (384-406 RBF, 340-350 linear). Return 0 or -1 on open failure. */
int orc_write_model_rbf(const char *path, const double *alpha, const float *labels,
                        const float *inst, int64_t n, int64_t d, double C, double gamma,
                        double b, int pos, int neg);
int orc_write_model_linear(const char *path, const double *w, int64_t d, double b,
                           int pos, int neg);

/* Decision values f(x_t) = sum_i alpha_i y_i exp(-gamma|x_i-x_t|^2) + b (FP64). */
void orc_decision(const double *alpha, const float *labels, const float *inst, int64_t n,
                  int64_t d, double gamma, double b, const float *test, int64_t nt,
                  double *f);
/* -analysis (2328-2360, rbf_fnorm_res 2388-2467): |K_Y|_F and |K_Y - U U^T|_F. */
void orc_fnorm_res(const float *X, const float *y, int64_t n, int64_t d, double gamma,
                   const double *U, int64_t k, double *fnorm, double *fnorm_res);

/* Gradients of the trained model with respect to theta = (log gamma, s_1..s_d), the
 * kernel being exp(-gamma sum_f e^{2 s_f} (x_if - x_jf)^2) (per-feature input scales),
 * with alpha and b held fixed (envelope theorem at the solver's optimum):
 *   D(theta) = sum alpha - 1/2 alpha^T P K_Y P alpha, P the orthogonal projector onto
 *              range(K_Y Omega) -- the trained dual objective through the factor
 *              construction (U U^T = P K_Y P when W^T K_Y W is positive definite);
 *   L(theta) = mean over held-out rows of the smoothed hinge h(y_t f(x_t)), f the exact-
 *              kernel decision function with bias b; h(z) = 1/2 - z (z <= 0),
 *              (1 - z)^2 / 2 (0 < z < 1), 0 (z >= 1).
 * FP64 throughout, brute force over pairs (no expansion of (x_i - x_j)^2).  vals[0] = D,
 * vals[1] = L; gD, gL (d + 1 each, component 0 = d/dlog gamma) may be NULL (values only).
 * Returns 0, or -1 when W^T K_Y W is not positive definite. */
int orc_grad(const float *X, const float *y, int64_t n, int64_t d, int64_t k,
             const float *omega, const double *alpha, double b, double log_gamma,
             const double *log_scales, const float *Xt, const float *yt, int64_t nt,
             double *vals, double *gD, double *gL);

/* Threads the oracle's OpenMP regions use (the CPU baseline reports it). */
int orc_num_threads(void);
void orc_set_num_threads(int t);

#ifdef __cplusplus
}
#endif
#endif
