/*
 * tsvm.h -- C-ABI of libtsvm, the MI355X (gfx950) TensorSVM training hot path.
 *
 * Plain C types only: host pointers are caller-owned, row-major, 64-bit sizes; device
 * memory is owned by the library (a tsvm_ctx holds the stream and a reusable device
 * workspace).  Every entry point returns a tsvm_status instead of exiting; the text of
 * the last failure is available from tsvm_last_error().  One context per host thread;
 * calls on one context are serialised by the caller.
 *
 * Each entry point names the reference interface it replaces
 * (robbwu/tensorsvm, tensorsvm-train.cu:line).
 */
#ifndef TSVM_H
#define TSVM_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef int tsvm_status;
#define TSVM_OK 0
#define TSVM_ERR_ARG 1         /* invalid argument / shape */
#define TSVM_ERR_HIP 2         /* HIP runtime or launch failure */
#define TSVM_ERR_NOMEM 3       /* device or host allocation failure */
#define TSVM_ERR_IO 4          /* file open / write failure */
#define TSVM_ERR_FORMAT 5      /* malformed LIBSVM input */
#define TSVM_ERR_NUMERIC 6     /* eigen-solver failure, non-finite factor */
#define TSVM_ERR_NODEVICE 7    /* no HIP device */

/* Arithmetic of the kernel-product and Z^T D^-1 Z GEMMs. */
#define TSVM_PREC_FP32 0       /* plain form: FP32 MFMA products, FP64 capacitance */
#define TSVM_PREC_BF16 1       /* low-precision MFMA form: kernel products with the distance
                                  in split bf16 and E.A on the fp16 MFMA (both products of
                                  the factor in TSVM_PREC_FP16), FP32 accumulate; hybrid
                                  BF16/FP64 capacitance, FP64 Cholesky + refinement in the
                                  IPM */
/* Kernel-product forms (tsvm_kernel_product[_dev]); the distance GEMM is the BF16 form's,
 * the product E.A runs on the fp16 MFMA with A under one power-of-two scale.
 * TSVM_PREC_FP16X2 is also a training / factor form: the BF16 form with the factor's
 * sketch in the two-term product (decision values closer to the FP32 form's, one more
 * MFMA per tile in the sketch). */
#define TSVM_PREC_FP16X2 2     /* E as two fp16 terms (~2^-22), A fp16 (2^-11) */
#define TSVM_PREC_FP16 3       /* E and A fp16 (2^-11 each), one MFMA per tile */
#define TSVM_PREC_FP16R2 4     /* E fp16 (2^-11), A as two fp16 terms (~2^-22) */
/* FP64 throughout (distance, exp, product on the FP64 MFMA; FP32 output): the reference
 * against which the other forms' error is stated.  As a training / factor form: both
 * products of the factor in FP64, the rest as TSVM_PREC_FP32 (factor stored in FP32,
 * FP64 capacitance). */
#define TSVM_PREC_FP64 5
/* Training / interior-point form: TSVM_PREC_BF16 with a third capacitance tier -- the rows
 * with the smallest t_i = w_i |z_i|^2 (the non-support vectors late in a solve) on the
 * FP8 e4m3 MFMA with per-column power-of-two scales and FP32 accumulation, under an error
 * budget of 2^-5 (tsvm_weighted_syrk_dev: the same split). */
#define TSVM_PREC_FP8 6

/* Largest rank k of the factor (and width of Z in the interior-point solver): the
 * capacitance Cholesky and its solves keep a k x k FP64 matrix per workgroup.  The
 * reference's own limit is the k x k host eigen-solve (LRA, tensorsvm-train.cu:2180-2212);
 * its usage lines take k = 32 and 400 (README.md). */
#define TSVM_MAX_RANK 1024

typedef struct tsvm_ctx tsvm_ctx;

/* Context: device ordinal, one HIP stream, reusable device workspace.
 * Replaces the per-call cublasCreate/cusolverDnCreate/cudaStreamCreate handles of
 * LRA (tensorsvm-train.cu:1946-1961) and mpc (1509-1513).  Creation also sets up the
 * k x k library calls of the factor construction (rocBLAS handle; rocSOLVER dsyevd,
 * dpotrf, dtrtri and rocBLAS dgemm run once so that their kernels are loaded; ~0.4 s once
 * per process). */
tsvm_status tsvm_create(int device, tsvm_ctx **out);
void tsvm_destroy(tsvm_ctx *ctx);
const char *tsvm_last_error(const tsvm_ctx *ctx);
/* Progress lines on stderr during long calls on this context (kernel-product slabs of
 * the factor, IPM iterations, gradient phases).  No counterpart in the reference: for
 * runs of many minutes under a supervisor that watches for output.  Off by default;
 * tsvm_train_params.progress turns it on for one training call. */
void tsvm_set_progress(tsvm_ctx *ctx, int on);
/* Synchronise the context's stream. */
tsvm_status tsvm_sync(tsvm_ctx *ctx);

/* Random test matrix Omega (count = n*k values, row-major): the reference's gaussrand()
 * (tensorsvm-train.cu:2621-2643) called in the order of the fill loop (1966-1974) from
 * glibc's never-seeded rand() state; bit-identical, computed on nthreads host threads
 * (<= 0: all cores) with a jump-ahead of the additive generator. */
tsvm_status tsvm_gaussrand_fill(float *omega, int64_t count, int nthreads);

/* Labelled RBF kernel product
 *   W[i,c] = y1_i * sum_j y2_j exp(-gamma |x1_i - x2_j|^2) A[j,c]
 * X1 m x d, X2 n x d, A n x r, W m x r (host pointers, row-major).
 * Replaces rbf_kermatmul (tensorsvm-train.cu:2476-2567, with vecnorm 2569-2581 and
 * rbf_kergen 2598-2612) and its out-of-core driver RBF_KERMATMUL (1999-2063): one fused
 * MFMA kernel, the kernel matrix is never materialised.  precision: TSVM_PREC_*. */
tsvm_status tsvm_kernel_product(tsvm_ctx *ctx, const float *X1, const float *y1, int64_t m,
                                const float *X2, const float *y2, int64_t n, int64_t d,
                                const float *A, int64_t r, double gamma, int precision,
                                float *W);
/* Same with device pointers (enqueued on the context's stream, not synchronised, except
 * that a product of at least two rounds of 256-row x 256-column tiles per CU, which runs
 * as one cooperative launch whose blocks take tiles per XCD from a device counter, waits
 * for its end to check that schedule; under stream capture it runs as one launch per
 * round instead). */
tsvm_status tsvm_kernel_product_dev(tsvm_ctx *ctx, const float *X1, const float *y1,
                                    int64_t m, const float *X2, const float *y2, int64_t n,
                                    int64_t d, const float *A, int64_t r, double gamma,
                                    int precision, float *W);

/* Allocation-free form of tsvm_kernel_product_dev.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops such as scoring test rows against a model's support vectors, harnesses that
 * forbid hipMalloc).  Same arguments, arithmetic and results as tsvm_kernel_product_dev
 * (the same kernels); the operand copies it prepares live in a fixed device array of
 * TSVM_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends, instead of the
 * context's workspace.  TSVM_ERR_ARG (before any work is queued) when the operands need
 * more: about 2(m + n)(3d + 4) + 2 n r bytes in the BF16/fp16 forms and
 * 4(m + n)(d + 2) + 4 n r (+ 4 m r per 8192-column segment when n > 8192) in FP32.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU (two
 * contexts or streams) race and must be serialised by the caller.  Enqueued on the
 * context's stream and synchronised as tsvm_kernel_product_dev. */
#define TSVM_NOALLOC_SCRATCH_BYTES (64ull << 20)
tsvm_status tsvm_kernel_product_noalloc_dev(tsvm_ctx *ctx, const float *X1, const float *y1,
                                            int64_t m, const float *X2, const float *y2,
                                            int64_t n, int64_t d, const float *A, int64_t r,
                                            double gamma, int precision, float *W);

/* Timings and diagnostics of the low-rank factor construction. */
typedef struct {
  double t_product1;   /* W = K_Y Omega (s, HIP events on the context's stream) */
  double t_orth;       /* orthonormalisation of W (s) */
  double t_product2;   /* Q = K_Y W (s, HIP events) */
  double t_core;       /* W^T W, C = W^T Q, host eigen-decomposition (s) */
  double t_factor;     /* U = W X (s) */
  double t_total;      /* whole call (s) */
  int64_t realk;       /* eigenvalues > 0 kept (the reference's "real rank", 2198-2211) */
  double eig_max, eig_min;
  int orth_passes;     /* Cholesky-QR passes taken */
} tsvm_lra_info;

/* Rank-k factor U (n x k) of the labelled RBF kernel, U U^T ~ diag(y) K diag(y):
 *   W = K_Y Omega; W = orth(W); Q = K_Y W; C = W^T Q; C = V diag(w) V^T;
 *   U = W V diag(sqrt(max(w, 0))).
 * Replaces LRA (tensorsvm-train.cu:1941-2249; default C_EIG path with one range-finder
 * pass).  The orthonormaliser is a shifted Cholesky-QR on FP64 MFMA instead of the
 * reference's TSQR (617-856): U U^T depends only on range(W).  omega: n x k row-major
 * test matrix, or NULL to draw the reference's gaussrand() stream (tsvm_gaussrand_fill).
 * U: n x k row-major FP32 output (host). */
tsvm_status tsvm_lra(tsvm_ctx *ctx, const float *X, const float *y, int64_t n, int64_t d,
                     int64_t k, double gamma, int precision, const float *omega, float *U,
                     tsvm_lra_info *info);
/* The reference's own types: double LRA(float *Z, int ldz, double *U, int ldu, long n,
 * long k) (tensorsvm-train.cu:111, 1941-2249, called at 960) writes U as double with
 * leading dimension ldu and returns the largest eigenvalue kept (w[k-1]); gamma and the
 * labels, globals there, are arguments here.  X: n rows of d features, leading dimension
 * ldx >= d; U: n x k FP64 with ldu >= k (host).  U holds the FP32 factor widened (the
 * reference's U = W X is an FP32 GEMM, 2231-2233, widened the same way).  *l1 (may be
 * NULL) = the largest eigenvalue kept. */
tsvm_status tsvm_lra_f64(tsvm_ctx *ctx, const float *X, int64_t ldx, const float *y, int64_t n,
                         int64_t d, int64_t k, double gamma, int precision, const float *omega,
                         double *U, int64_t ldu, double *l1, tsvm_lra_info *info);
/* Device-pointer form: X, y, omega (n x k, consumed: used as workspace) and U are
 * device pointers; synchronised on return. */
tsvm_status tsvm_lra_dev(tsvm_ctx *ctx, const float *X, const float *y, int64_t n, int64_t d,
                         int64_t k, double gamma, int precision, float *omega, float *U,
                         tsvm_lra_info *info);

/* Weighted Gram of the interior-point capacitance, M = I + Z^T diag(w) Z (replaces the
 * D^{-1/2} scaling + cblas_dsyrk / chunked cublasDsyrk of mpc, tensorsvm-train.cu:
 * 1660-1706).  Z: n x k FP32 row-major, w: n FP64, M: k x k FP64 row-major (all
 * device pointers; synchronised on return).  TSVM_PREC_FP32: FP64 MFMA over all rows.
 * TSVM_PREC_BF16 (hybrid): rows with small t_i = w_i |z_i|^2 run on BF16 MFMA (FP32
 * accumulate) while a bound on sqrt(sum t_i^2) over them stays within budget (<= 0: 4,
 * the interior-point default), the others on FP64 MFMA.  The BF16 part's error is
 * about 2^-9 times that bound in norm, against lambda_min(M) >= 1.  TSVM_PREC_FP8: the
 * same with the smallest-t rows first taken into an FP8 e4m3 tier (per-column scales)
 * while its error bound stays within 2^-5 (k > 128). */
typedef struct {
  int64_t n_fp64_rows; /* rows computed in FP64 */
  double bf16_bound;   /* bound on sqrt(sum t_i^2) over the BF16 rows */
  double t_bf16;       /* s in the BF16 (and FP8) part */
  double bf16_tflops;  /* n k (k+1) / t_bf16 */
  int64_t n_fp8_rows;  /* TSVM_PREC_FP8: rows in the FP8 tier */
  double fp8_bound;    /* TSVM_PREC_FP8: error bound of the FP8 tier */
  int64_t n_x2_rows;   /* rows in the compensated (hi + lo bf16) tier */
  double x2_bound;     /* bound on sqrt(sum t_i^2) over them */
  double t_dense;      /* s in the dense pass (every row; lower tiles, diagonal tiles whole) */
  double dense_mfma_tflops; /* MFMA flops the dense pass issues / t_dense */
} tsvm_syrk_info;
tsvm_status tsvm_weighted_syrk_dev(tsvm_ctx *ctx, const float *Z, int64_t n, int64_t k,
                                   const double *w, int precision, double budget, double *M,
                                   tsvm_syrk_info *info);

/* Gradients of the trained model with respect to theta = (log gamma, s_1..s_d), the
 * kernel being exp(-gamma sum_f e^{2 s_f} (x_if - x_jf)^2) (per-feature input log-scales
 * s), with alpha and b held fixed (envelope theorem at the solver's optimum).  No
 * counterpart in the reference (its dual objective: tensorsvm-train.cu:1636-1643).
 *   vals[0] = D = sum alpha - 1/2 alpha^T P K_Y P alpha, P the projector onto
 *             range(K_Y Omega): the trained dual through the factor construction (the
 *             same Omega as training; U U^T = P K_Y P when no eigenvalue is clipped --
 *             TSVM_ERR_NUMERIC otherwise)
 *   vals[1] = L = mean over held-out rows of the smoothed hinge h(y_t f(x_t)), f the
 *             exact-kernel decision function with bias b; h(z) = 1/2 - z (z <= 0),
 *             (1 - z)^2 / 2 (0 < z < 1), 0 (z >= 1)
 * g_dual, g_loss (d + 1 each, host; component 0 = d/dlog gamma) may be NULL.  The
 * backward pass recomputes kernel tiles in the fused FP32 product over derivative
 * right-hand sides; nothing of size n^2 is stored.  The held-out part (the decision values
 * behind L and the columns of its derivative) runs on the FP64 product: its per-feature
 * sums cancel to a small remainder, and differences of L must resolve its derivative.  precision: the forward pass's two
 * products of the factor (K_Y Omega, K_Y W) in TSVM_PREC_FP32, or in the fp16 product
 * with TSVM_PREC_BF16 (about 8x faster at large n; the backward stays FP32).  Device: X
 * (n x d), y, alpha (FP64), omega (n x k, not modified), Xt (nt x d), yt; host:
 * log_scales (d), vals (2). */
typedef struct {
  double t_forward;       /* s: scaled features, factor and dual value */
  double t_backward;      /* s in the backward products and reductions (HIP events) */
  double backward_flops;  /* the dual's FP32 backward product, 2 n^2 (d + 3 (2d+1)) */
  double backward_tflops; /* backward_flops / the time of that product and its reductions */
  double t_total;
  int64_t realk;
  int orth_passes;
  double t_backward_heldout; /* s of t_backward in the held-out part (FP64 product, 2 nt n
                                (d + 2d+1) flops, and its reduction) */
} tsvm_grad_info;
tsvm_status tsvm_grad_dev(tsvm_ctx *ctx, const float *X, const float *y, int64_t n, int64_t d,
                          const double *alpha, double b, int64_t k, double log_gamma,
                          const double *log_scales, const float *omega, const float *Xt,
                          const float *yt, int64_t nt, int precision, double *vals,
                          double *g_dual, double *g_loss, tsvm_grad_info *info);

/* ---- model files and prediction ------------------------------------------------------
 * The model files the trainer writes (RBF: LIBSVM text, writemodel 353-406; linear:
 * LIBLINEAR text, 340-350), read back and scored on the GPU -- what the reference leaves
 * to an external ThunderSVM/LIBSVM predictor (README.md:27-37).  Decision values:
 * RBF f = sum_j coef_j exp(-gamma |x - sv_j|^2) - rho; linear f = w . x + w_bias * bias;
 * the prediction is label[0] if f > 0, else label[1]. */
typedef struct tsvm_model tsvm_model;
/* TSVM_ERR_IO when the file cannot be read, TSVM_ERR_FORMAT on a malformed model. */
tsvm_status tsvm_read_model(const char *path, tsvm_model **model);
void tsvm_model_free(tsvm_model *model);
/* Any output pointer may be NULL; rho of a linear model is -w_bias * bias. */
tsvm_status tsvm_model_get(const tsvm_model *model, int *kernel_type, double *gamma,
                           double *rho, int *labels, int64_t *nsv, int64_t *d);
/* T: nt x d host rows (features past either side's width count as 0); f (decision
 * values) and pred (predicted labels) are host arrays of nt, either may be NULL. */
tsvm_status tsvm_predict(tsvm_ctx *ctx, const tsvm_model *model, const float *T, int64_t nt,
                         int64_t d, double *f, float *pred);

/* -analysis (tensorsvm-train.cu:2328-2360, rbf_fnorm_res 2388-2467): |K_Y|_F and
 * |K_Y - U U^T|_F over all n^2 pairs, on the GPU without storing any n x n block.
 * Device X (n x d, d <= 64), y, U (n x k FP32 row-major); results on the host. */
tsvm_status tsvm_fnorm_res_dev(tsvm_ctx *ctx, const float *X, const float *y, int64_t n,
                               int64_t d, double gamma, const float *U, int64_t k,
                               double *fnorm, double *fnorm_res);

/* Interior-point solver parameters (the reference's constants are the defaults). */
#define TSVM_STOP_REF 0 /* mu, |r_x|, |r_y| < 1e-7 (tensorsvm-train.cu:1647) */
#define TSVM_STOP_GAP 1 /* (pobj-dobj)/(1+|pobj|) <= tol, |r_x|/(1+sqrt n) <= tol,
                           |r_y| <= tol (1 + |x|_1 / C) */
#define TSVM_CONVERGED 0
#define TSVM_ILL_COND 1 /* max D / min D > 1e18 (1653) */
#define TSVM_MAXIT 2
#define TSVM_CHOL_FAIL 3 /* capacitance Cholesky failed: current iterate returned (1711) */
typedef struct {
  double C;          /* box bound */
  int kernel_type;   /* 2: RBF initial point (1535-1552); 0: linear one (1553-1579) */
  int maxit;         /* MAX_MPC_ITER = 100 (62) */
  int stop_mode;     /* TSVM_STOP_REF or TSVM_STOP_GAP */
  double tol;        /* for TSVM_STOP_GAP */
  int precision;     /* TSVM_PREC_FP32: FP64 capacitance; TSVM_PREC_BF16 (or FP16X2): hybrid
                        capacitance (tsvm_weighted_syrk_dev, bulk rows on BF16 MFMA) +
                        refinement of the Newton step against the FP32 factor */
  int refine;        /* refinement sweeps of the bordered Newton system: m > 0 sweeps for each of
                        the affine and the corrector's steps (each solved first); -m (default
                        -1): the combined step starts from the affine one and m sweeps against
                        the combined right-hand sides follow, the first adding the corrector's
                        step (no solve of its own) */
  int verbose;       /* print the reference's iteration line every 5 iterations */
  int init;          /* 0: the reference's initial point for kernel_type; 1: the same x with
                        s, xi shifted so the initial dual residual is zero (what the linear
                        branch does, 1569-1578) -- far fewer iterations when Q is large */
} tsvm_ipm_params;
typedef struct {
  int iters, reason;
  double mu, normdx, normdy, pobj, dobj, gap_rel, condD;
  double y;          /* equality multiplier (= rho = -b at a free SV) */
  double t_total;    /* s */
  double t_syrk;     /* s spent building Z^T D^-1 Z (all iterations) */
  double syrk_tflops;/* algorithmic n k (k+1) flops / t_syrk */
  double refine_res; /* last relative refinement residual of the bordered Newton system (with
                        refine = -1, the residual of the affine step against the combined
                        right-hand sides: the corrector's relative size) */
  /* low-precision form: the capacitance's bulk rows on BF16 MFMA (hybrid split) */
  double t_syrk_bf16;      /* s spent in the BF16 part (all iterations) */
  double syrk_bf16_tflops; /* n k (k+1) flops per BF16 part / its average time */
  double bf16_row_frac;    /* rows in the BF16 part, averaged over iterations */
  int syrk_bf16_iters;     /* iterations with a BF16 part */
  double fp8_row_frac;     /* TSVM_PREC_FP8: rows in the FP8 tier, averaged over iterations */
  double x2_row_frac;      /* rows in the compensated tier, averaged over iterations */
  double t_syrk_dense;     /* s spent in the dense (BF16 or FP8) pass over all rows */
  double syrk_dense_mfma_tflops; /* MFMA flops the dense pass issues / t_syrk_dense */
} tsvm_ipm_info;
void tsvm_ipm_default_params(tsvm_ipm_params *p);
/* Mehrotra predictor-corrector for the SVM dual with Q = Z Z^T, every O(n) and O(nk)
 * step on the GPU.  Replaces mpc (tensorsvm-train.cu:1507-1785), NewtonStep (1794-1822)
 * and SMWSolve (1824-1840): the Newton system (D + Z Z^T) is solved by
 * Sherman-Morrison-Woodbury with the capacitance I + Z^T D^-1 Z built on MFMA and
 * factored by an FP64 Cholesky on the GPU.  Z: n x k labelled factor (FP32, row-major),
 * a: labels (+-1).  alpha (= X), xi, s: n FP64 outputs (s may be NULL). */
tsvm_status tsvm_ipm(tsvm_ctx *ctx, const float *Z, const float *a, int64_t n, int64_t k,
                     const tsvm_ipm_params *prm, double *alpha, double *xi, double *s,
                     tsvm_ipm_info *info);
/* The reference's own types: void mpc(double *Z, double *a, double C, double *X,
 * double *Xi, int n, int k) (tensorsvm-train.cu:100, 1507-1785).  Z: n x k FP64 with
 * leading dimension ldz >= k, a: n FP64 labels, X (= alpha) and Xi: n FP64 outputs (host).
 * Z is narrowed to the solver's FP32 factor, which is exact for a factor from
 * tsvm_lra_f64 or the reference's LRA (FP32 values widened) and a 2^-24 relative rounding
 * otherwise.  prm: NULL for the defaults (tsvm_ipm_default_params); C overrides prm->C. */
tsvm_status tsvm_mpc(tsvm_ctx *ctx, const double *Z, int64_t ldz, const double *a, double C,
                     double *X, double *Xi, int64_t n, int64_t k, const tsvm_ipm_params *prm,
                     tsvm_ipm_info *info);
/* Device-pointer form (Z, a, alpha, xi, s on the device; s may be NULL). */
tsvm_status tsvm_ipm_dev(tsvm_ctx *ctx, const float *Z, const float *a, int64_t n, int64_t k,
                         const tsvm_ipm_params *prm, double *alpha, double *xi, double *s,
                         tsvm_ipm_info *info);

/* Whole training run: factor (T=2) or labelled data (T=0), IPM, bias. */
typedef struct {
  int64_t k;          /* rank (-k, default 50) */
  double C;           /* -c (default 1) */
  double gamma;       /* -g (default 0) */
  int kernel_type;    /* -t: 2 RBF (LRA + IPM), 0 linear (IPM on diag(y) X) */
  int precision;      /* TSVM_PREC_FP32 / TSVM_PREC_BF16 (-tensorcore) / TSVM_PREC_FP16X2 */
  int stop_mode;      /* TSVM_STOP_REF (default) or TSVM_STOP_GAP */
  double tol;         /* gap tolerance for TSVM_STOP_GAP */
  int maxit;          /* IPM iterations (100) */
  int refine;         /* refinement sweeps, as tsvm_ipm_params.refine (default -1) */
  int verbose;        /* reference-style progress lines on stdout */
  int init;           /* IPM initial point, as tsvm_ipm_params.init */
  int analysis;       /* -analysis: |K_Y|_F and |K_Y - U U^T|_F after the factor (RBF) */
  int progress;       /* progress lines on stderr: the kernel products run in row slabs,
                         one line per slab, and one line per 10 IPM iterations (results
                         are the same; meant for runs of many minutes) */
} tsvm_train_params;
typedef struct {
  tsvm_lra_info lra;
  tsvm_ipm_info ipm;
  double b;           /* bias; rho = -b */
  double b_std;       /* spread of the per-free-SV bias samples (reference "std b") */
  int b_samples;      /* free SVs averaged (<= 50 RBF, <= 5000 linear) */
  int b_slack_rule;   /* RBF: 1 when the samples are the slack-confirmed free SVs (relative-
                         gap stop), 0 for the reference's set (1e-6 < alpha < C - 1e-6) */
  int64_t nsv, nfree; /* alpha > 1e-6; and alpha < C - 1e-6 */
  double t_bias, t_total;  /* s */
  double fnorm, fnorm_res; /* -analysis results (0 unless requested) */
  double t_analysis;       /* s (not part of t_total) */
} tsvm_train_info;
void tsvm_train_default_params(tsvm_train_params *p);
/* Train on X (n x d row-major FP32) with labels y in {+1,-1}: the RBF branch of main
 * (tensorsvm-train.cu:957-987: LRA then mpc) or the linear branch (907-918), then the bias
 * of writemodel (301-331 linear, 353-376 RBF).  omega: n x k test matrix or NULL (the
 * reference's gaussrand() stream).  alpha: n FP64 output.  w_out (linear only, may be
 * NULL): d weights. */
tsvm_status tsvm_train(tsvm_ctx *ctx, const float *X, const float *y, int64_t n, int64_t d,
                       const tsvm_train_params *prm, const float *omega, double *alpha,
                       double *w_out, tsvm_train_info *info);
/* Device-pointer form of the RBF branch: X, y, omega (n x k, consumed) and alpha on the
 * device; the factor stays on the device (no host copies of size n k). */
tsvm_status tsvm_train_dev(tsvm_ctx *ctx, const float *X, const float *y, int64_t n, int64_t d,
                           const tsvm_train_params *prm, float *omega, double *alpha,
                           tsvm_train_info *info);

/* Decision values f(t) = sum_i alpha_i y_i exp(-gamma |x_i - t|^2) + b for m test rows
 * (host pointers; only alpha_i > 1e-6 contribute).  The predictor the reference delegates
 * to ThunderSVM (README.md:27-37) -- its own predict() (166-270) is dead code -- as a
 * kernel product with one right-hand side. */
tsvm_status tsvm_decision_function(tsvm_ctx *ctx, const float *X, const float *y,
                                   const double *alpha, int64_t n, int64_t d, double gamma,
                                   double b, const float *T, int64_t m, double *f);

/* LIBSVM text input (libsvmread, tensorsvm-train.cu:1162-1278): dense row-major FP32
 * features (*inst, n x d) and labels (*labels), malloc'ed; free with tsvm_free.
 * scale256: divide values by 256 (-scale256).  TSVM_ERR_IO / TSVM_ERR_FORMAT on failure
 * (the reference prints and returns with partial arrays). */
tsvm_status tsvm_read_libsvm(const char *path, int scale256, float **labels, float **inst,
                             int64_t *n, int64_t *d);
void tsvm_free(void *p);
/* LIBSVM text output of a dense data set (no counterpart in the reference: the writer of
 * generated data sets for the command-line path).  One line per row: the label, then
 * "j:v" for every non-zero feature (1-based), values as "%.9g", which tsvm_read_libsvm
 * reads back to the same floats.  Rows are formatted on all host threads. */
tsvm_status tsvm_write_libsvm(const char *path, const float *labels, const float *inst,
                              int64_t n, int64_t d);
/* Label mapping of main (879-896): first label -> +1 (POS), first different -> -1 (NEG).
 * Returns the count of labels outside {POS, NEG} (left unchanged). */
int64_t tsvm_map_labels(float *labels, int64_t n, int *pos, int *neg);
/* Model files of writemodel: RBF LIBSVM text (384-406) and linear LIBLINEAR text
 * (340-350), formats byte-identical to the reference's fprintf calls. */
tsvm_status tsvm_write_model_rbf(const char *path, const double *alpha, const float *labels,
                                 const float *inst, int64_t n, int64_t d, double C,
                                 double gamma, double b, int pos, int neg);
tsvm_status tsvm_write_model_linear(const char *path, const double *w, int64_t d, double b,
                                    int pos, int neg);

#ifdef __cplusplus
}
#endif
#endif /* TSVM_H */
