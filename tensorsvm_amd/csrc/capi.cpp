// C-ABI entry points of libtsvm: context, workspace, host-pointer wrappers.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include <rocblas/rocblas.h>

#include "internal.h"

namespace tsvm {

tsvm_status fail(Ctx* c, tsvm_status st, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->last_error = buf;
  return st;
}

hipError_t Workspace::get(const char* name, size_t bytes, void** out) {
  if (bytes == 0) bytes = 256;
  auto it = bufs.find(name);
  if (it != bufs.end()) {
    if (it->second.second >= bytes) {
      *out = it->second.first;
      return hipSuccess;
    }
    // the buffer may still be read by queued work: drain the device before freeing it
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return e;
    e = hipFree(it->second.first);
    if (e != hipSuccess) return e;
    bufs.erase(it);
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) return e;
  bufs[name] = {p, bytes};
  *out = p;
  return hipSuccess;
}

void Workspace::release_all() {
  hipDeviceSynchronize();
  for (auto& kv : bufs) hipFree(kv.second.first);
  bufs.clear();
}

}  // namespace tsvm

using tsvm::Ctx;

extern "C" {

tsvm_status tsvm_create(int device, tsvm_ctx** out) {
  if (!out) return TSVM_ERR_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return TSVM_ERR_NODEVICE;
  if (device < 0 || device >= count) return TSVM_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return TSVM_ERR_HIP;
  tsvm_ctx* x = new tsvm_ctx();
  x->c.device = device;
  if (hipStreamCreate(&x->c.stream) != hipSuccess) {
    delete x;
    return TSVM_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) x->c.cu_count = prop.multiProcessorCount;
  tsvm::warm_dense_la(&x->c);
  if (tsvm::noalloc_scratch_setup(&x->c) != TSVM_OK) {
    hipStreamDestroy(x->c.stream);
    delete x;
    return TSVM_ERR_HIP;
  }
  *out = x;
  return TSVM_OK;
}

void tsvm_destroy(tsvm_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->c.device);
  hipStreamSynchronize(ctx->c.stream);
  ctx->c.ws.release_all();
  if (ctx->c.rocblas) rocblas_destroy_handle((rocblas_handle)ctx->c.rocblas);
  hipStreamDestroy(ctx->c.stream);
  delete ctx;
}

const char* tsvm_last_error(const tsvm_ctx* ctx) {
  return ctx ? ctx->c.last_error.c_str() : "no context";
}

void tsvm_set_progress(tsvm_ctx* ctx, int on) {
  if (ctx) ctx->c.progress = on ? 1 : 0;
}

tsvm_status tsvm_sync(tsvm_ctx* ctx) {
  if (!ctx) return TSVM_ERR_ARG;
  TSVM_HIP(&ctx->c, hipStreamSynchronize(ctx->c.stream));
  return TSVM_OK;
}

tsvm_status tsvm_kernel_product_dev(tsvm_ctx* ctx, const float* X1, const float* y1, int64_t m,
                                    const float* X2, const float* y2, int64_t n, int64_t d,
                                    const float* A, int64_t r, double gamma, int precision,
                                    float* W) {
  if (!ctx) return TSVM_ERR_ARG;
  hipSetDevice(ctx->c.device);
  return tsvm::kernel_product_dev(&ctx->c, X1, y1, m, X2, y2, n, d, A, r, r, gamma, precision,
                                  W, r);
}

tsvm_status tsvm_kernel_product_noalloc_dev(tsvm_ctx* ctx, const float* X1, const float* y1,
                                            int64_t m, const float* X2, const float* y2,
                                            int64_t n, int64_t d, const float* A, int64_t r,
                                            double gamma, int precision, float* W) {
  if (!ctx) return TSVM_ERR_ARG;
  hipSetDevice(ctx->c.device);
  return tsvm::kernel_product_noalloc_dev(&ctx->c, X1, y1, m, X2, y2, n, d, A, r, gamma,
                                          precision, W);
}

tsvm_status tsvm_weighted_syrk_dev(tsvm_ctx* ctx, const float* Z, int64_t n, int64_t k,
                                   const double* w, int precision, double budget, double* M,
                                   tsvm_syrk_info* info) {
  if (!ctx) return TSVM_ERR_ARG;
  hipSetDevice(ctx->c.device);
  return tsvm::weighted_syrk_dev(&ctx->c, Z, n, k, w, precision, budget, M, info);
}

tsvm_status tsvm_grad_dev(tsvm_ctx* ctx, const float* X, const float* y, int64_t n, int64_t d,
                          const double* alpha, double b, int64_t k, double log_gamma,
                          const double* log_scales, const float* omega, const float* Xt,
                          const float* yt, int64_t nt, int precision, double* vals,
                          double* g_dual, double* g_loss, tsvm_grad_info* info) {
  if (!ctx) return TSVM_ERR_ARG;
  hipSetDevice(ctx->c.device);
  return tsvm::grad_dev(&ctx->c, X, y, n, d, alpha, b, k, log_gamma, log_scales, omega, Xt, yt,
                        nt, precision, vals, g_dual, g_loss, info);
}

tsvm_status tsvm_fnorm_res_dev(tsvm_ctx* ctx, const float* X, const float* y, int64_t n,
                               int64_t d, double gamma, const float* U, int64_t k,
                               double* fnorm, double* fnorm_res) {
  if (!ctx || !X || !y || !U || !fnorm || !fnorm_res) return TSVM_ERR_ARG;
  hipSetDevice(ctx->c.device);
  return tsvm::fnorm_res_dev(&ctx->c, X, y, n, d, gamma, U, k, k, fnorm, fnorm_res);
}

tsvm_status tsvm_kernel_product(tsvm_ctx* ctx, const float* X1, const float* y1, int64_t m,
                                const float* X2, const float* y2, int64_t n, int64_t d,
                                const float* A, int64_t r, double gamma, int precision,
                                float* W) {
  if (!ctx) return TSVM_ERR_ARG;
  Ctx* c = &ctx->c;
  if (m < 0 || n < 0 || d <= 0 || r < 0 || !X1 || !y1 || !W || (n > 0 && (!X2 || !y2 || !A)))
    return tsvm::fail(c, TSVM_ERR_ARG, "tsvm_kernel_product: bad arguments");
  if (m == 0 || r == 0) return TSVM_OK;
  hipSetDevice(c->device);
  hipStream_t s = c->stream;
  void *dX1, *dy1, *dX2, *dy2, *dA, *dW;
  TSVM_HIP(c, c->ws.get("h_X1", m * d * 4, &dX1));
  TSVM_HIP(c, c->ws.get("h_y1", m * 4, &dy1));
  TSVM_HIP(c, c->ws.get("h_X2", (n ? n : 1) * d * 4, &dX2));
  TSVM_HIP(c, c->ws.get("h_y2", (n ? n : 1) * 4, &dy2));
  TSVM_HIP(c, c->ws.get("h_A", (n ? n : 1) * r * 4, &dA));
  TSVM_HIP(c, c->ws.get("h_W", m * r * 4, &dW));
  TSVM_HIP(c, hipMemcpyAsync(dX1, X1, m * d * 4, hipMemcpyHostToDevice, s));
  TSVM_HIP(c, hipMemcpyAsync(dy1, y1, m * 4, hipMemcpyHostToDevice, s));
  if (n > 0) {
    TSVM_HIP(c, hipMemcpyAsync(dX2, X2, n * d * 4, hipMemcpyHostToDevice, s));
    TSVM_HIP(c, hipMemcpyAsync(dy2, y2, n * 4, hipMemcpyHostToDevice, s));
    TSVM_HIP(c, hipMemcpyAsync(dA, A, n * r * 4, hipMemcpyHostToDevice, s));
  }
  TSVM_CHECK(tsvm::kernel_product_dev(c, (float*)dX1, (float*)dy1, m, (float*)dX2, (float*)dy2, n,
                                      d, (float*)dA, r, r, gamma, precision, (float*)dW, r));
  TSVM_HIP(c, hipMemcpyAsync(W, dW, m * r * 4, hipMemcpyDeviceToHost, s));
  TSVM_HIP(c, hipStreamSynchronize(s));
  return TSVM_OK;
}

}  // extern "C"
