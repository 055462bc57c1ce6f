// Context, error handling, workspace and device-memory helpers.
#include "iso_internal.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace iso {

__global__ void fill_u32_kernel(uint32_t *p, uint32_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// Plain 16-byte-per-lane copy (also the reference kernel for HBM counter
// calibration: it reads and writes exactly n*16 bytes).
__global__ void copy_u128_kernel(const uint4 *src, uint4 *dst, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

// Clears memory with a kernel (graph-capture safe, unlike a memset node).
int fill_u32(iso_ctx *ctx, void *p, uint32_t v, size_t count) {
  if (count == 0) return ISO_OK;
  size_t blocks = (count + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                     (uint32_t *)p, v, count);
  ISO_HIP_CHECK(ctx, hipGetLastError());
  return ISO_OK;
}

int fail(iso_ctx *ctx, int code, const char *fmt, ...) {
  if (ctx) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    ctx->last_error = buf;
  }
  return code;
}

int hip_fail(iso_ctx *ctx, hipError_t e, const char *what, const char *file, int line) {
  int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? ISO_ERR_OUT_OF_MEMORY
                                                                         : ISO_ERR_HIP;
  return fail(ctx, code, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

int enter(iso_ctx *ctx) {
  if (!ctx) return ISO_ERR_INVALID_ARG;
  ctx->last_error.clear();
  ISO_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  return ISO_OK;
}

int ws_get(iso_ctx *ctx, iso_ws_slot slot, size_t bytes, bool zero_at_rest, void **out) {
  iso_ws_buf &b = ctx->ws[slot];
  if (bytes == 0) bytes = 256;
  if (b.bytes < bytes) {
    if (b.ptr) {
      ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      ISO_HIP_CHECK(ctx, hipFree(b.ptr));
      b.ptr = nullptr;
      b.bytes = 0;
    }
    size_t grow = bytes + bytes / 4;  // amortise growth
    grow = (grow + 4095) & ~size_t(4095);
    ISO_HIP_CHECK(ctx, hipMalloc(&b.ptr, grow));
    b.bytes = grow;
    b.zero_at_rest = zero_at_rest;
    if (zero_at_rest) ISO_TRY(fill_u32(ctx, b.ptr, 0u, grow / 4));
  }
  *out = b.ptr;
  return ISO_OK;
}

int pinned_get(iso_ctx *ctx, size_t bytes, void **out, int area) {
  void *&p = ctx->pinned[area];
  size_t &have = ctx->pinned_bytes[area];
  if (have < bytes) {
    if (p) {
      ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      ISO_HIP_CHECK(ctx, hipHostFree(p));
      p = nullptr;
      have = 0;
    }
    size_t n = bytes < 4096 ? 4096 : bytes;
    ISO_HIP_CHECK(ctx, hipHostMalloc(&p, n, hipHostMallocDefault));
    have = n;
  }
  *out = p;
  return ISO_OK;
}

}  // namespace iso

extern "C" {

int iso_ctx_create(iso_ctx **out, int device) {
  if (!out) return ISO_ERR_INVALID_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ISO_ERR_HIP;
  if (device < 0 || device >= ndev) return ISO_ERR_INVALID_ARG;
  iso_ctx *ctx = new iso_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess) { delete ctx; return ISO_ERR_HIP; }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->cu_count = prop.multiProcessorCount;
  // default-priority stream (no priority games)
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return ISO_ERR_HIP;
  }
  iso::noalloc_setup(ctx);  // the device's address of the no-alloc scratch array
  *out = ctx;
  return ISO_OK;
}

int iso_ctx_destroy(iso_ctx *ctx) {
  if (!ctx) return ISO_ERR_INVALID_ARG;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &b : ctx->ws)
    if (b.ptr) (void)hipFree(b.ptr);
  for (auto &e : ctx->walk_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : ctx->graphs) {
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    if (e.buf) (void)hipFree(e.buf);
  }
  for (void *p : ctx->pinned)
    if (p) (void)hipHostFree(p);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return ISO_OK;
}

int iso_ctx_synchronize(iso_ctx *ctx) {
  ISO_TRY(iso::enter(ctx));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return ISO_OK;
}

void *iso_ctx_stream(iso_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

const char *iso_ctx_last_error(const iso_ctx *ctx) { return ctx ? ctx->last_error.c_str() : ""; }

const char *iso_ctx_tuning(const iso_ctx *ctx) { return ctx ? ctx->tune_log.c_str() : ""; }

const char *iso_status_string(int s) {
  switch (s) {
    case ISO_OK: return "ok";
    case ISO_ERR_INVALID_ARG: return "invalid argument";
    case ISO_ERR_HIP: return "HIP runtime error";
    case ISO_ERR_OUT_OF_MEMORY: return "out of device memory";
    case ISO_ERR_CAPACITY: return "output capacity too small";
    case ISO_ERR_UNSUPPORTED: return "unsupported parameters";
    case ISO_ERR_EMPTY: return "empty point set";
    default: return "unknown status";
  }
}

int iso_device_alloc(iso_ctx *ctx, void **d_ptr, size_t bytes) {
  ISO_TRY(iso::enter(ctx));
  if (!d_ptr) return ISO_ERR_INVALID_ARG;
  ISO_HIP_CHECK(ctx, hipMalloc(d_ptr, bytes ? bytes : 4));
  return ISO_OK;
}

int iso_device_free(iso_ctx *ctx, void *d_ptr) {
  ISO_TRY(iso::enter(ctx));
  if (d_ptr) {
    ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    ISO_HIP_CHECK(ctx, hipFree(d_ptr));
  }
  return ISO_OK;
}

int iso_copy_to_device(iso_ctx *ctx, void *d_dst, const void *h_src, size_t bytes) {
  ISO_TRY(iso::enter(ctx));
  if (bytes == 0) return ISO_OK;
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return ISO_OK;
}

int iso_copy_to_host(iso_ctx *ctx, void *h_dst, const void *d_src, size_t bytes) {
  ISO_TRY(iso::enter(ctx));
  if (bytes == 0) return ISO_OK;
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return ISO_OK;
}

int iso_device_copy(iso_ctx *ctx, void *d_dst, const void *d_src, size_t bytes) {
  ISO_TRY(iso::enter(ctx));
  if (bytes == 0) return ISO_OK;
  if (bytes % 16 == 0 && ((uintptr_t)d_dst & 15) == 0 && ((uintptr_t)d_src & 15) == 0) {
    const size_t n = bytes / 16;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(iso::copy_u128_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                       (const uint4 *)d_src, (uint4 *)d_dst, n);
    ISO_HIP_CHECK(ctx, hipGetLastError());
  } else {
    ISO_HIP_CHECK(ctx, hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  }
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return ISO_OK;
}

int iso_device_memset(iso_ctx *ctx, void *d_dst, int value, size_t bytes) {
  ISO_TRY(iso::enter(ctx));
  if (bytes == 0) return ISO_OK;
  if (bytes % 4 == 0 && ((uintptr_t)d_dst & 3) == 0) {
    uint32_t b = (uint32_t)(value & 0xff);
    return iso::fill_u32(ctx, d_dst, b | b << 8 | b << 16 | b << 24, bytes / 4);
  }
  ISO_HIP_CHECK(ctx, hipMemsetAsync(d_dst, value, bytes, ctx->stream));
  return ISO_OK;
}

}  // extern "C"
