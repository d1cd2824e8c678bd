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
#define TSVM_PREC_BF16 1       /* BF16 MFMA (FP32 accumulate) + FP64 Cholesky + refinement */

typedef struct tsvm_ctx tsvm_ctx;

/* Context: device ordinal, one HIP stream, reusable device workspace.
 * Replaces the per-call cublasCreate/cusolverDnCreate/cudaStreamCreate handles of
 * LRA (tensorsvm-train.cu:1946-1961) and mpc (1509-1513). */
tsvm_status tsvm_create(int device, tsvm_ctx **out);
void tsvm_destroy(tsvm_ctx *ctx);
const char *tsvm_last_error(const tsvm_ctx *ctx);
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
/* Same with device pointers (enqueued on the context's stream, not synchronised). */
tsvm_status tsvm_kernel_product_dev(tsvm_ctx *ctx, const float *X1, const float *y1,
                                    int64_t m, const float *X2, const float *y2, int64_t n,
                                    int64_t d, const float *A, int64_t r, double gamma,
                                    int precision, float *W);

#ifdef __cplusplus
}
#endif
#endif /* TSVM_H */
