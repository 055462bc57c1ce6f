// Volume ingest and isovalue selection, the front of the `single` module.
//
// Ingest (filehandler/src/BinaryFloatReader.cpp:29-49, types/src/Isosurface.cpp:
// 32-76, 185-200): a raw float32 file stored x-fastest becomes the z-fastest
// image img[(x*Y + y)*Z + z] = file[x + y*X + z*X*Y]; realMin/realMax are the
// extremes of the values inside [cfg.min, cfg.max], starting from 100000 and
// -100000.  Here: an LDS-tiled 32x32 transpose per y plane (reads coalesced
// along x, writes along z) with the restricted min/max folded in.
//
// Isovalues (runtime/src/Single.cpp:286-392, uniform = true): a 65536-bin
// histogram of the image over [realMin, realMax] on the device, then the
// reference's equalisation walk over the bins on the host (65536 steps).
// Bin index = int(65536 * (v - gmin) * oneovergrange) with oneovergrange the
// float of 1.0/(double)(gmax - gmin), clamped to [0, 65535]; the reference's
// loop bound is a float (dataSize), so voxel i counts iff float(i) < float(n).
#include <cfloat>
#include <vector>

#include "iso_internal.hpp"

namespace {

constexpr int kT = 32;

__global__ __launch_bounds__(256) void ingest_kernel(const float *file, int X, int Y, int Z,
                                                     float cmin, float cmax, float *img,
                                                     float2 *part) {
  __shared__ float tile[kT][kT + 1];
  __shared__ float s_lo[4], s_hi[4];
  const int x0 = blockIdx.x * kT, z0 = blockIdx.y * kT, y = blockIdx.z;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  float lo = 100000.0f, hi = -100000.0f;
  for (int r = ty; r < kT; r += 8) {
    const int x = x0 + tx, z = z0 + r;
    if (x < X && z < Z) {
      const float v = file[(int64_t)x + (int64_t)y * X + (int64_t)z * X * Y];
      tile[r][tx] = v;
      if (v < lo && v >= cmin) lo = v;
      if (v > hi && v <= cmax) hi = v;
    }
  }
  __syncthreads();
  for (int r = ty; r < kT; r += 8) {
    const int x = x0 + r, z = z0 + tx;
    if (x < X && z < Z) img[((int64_t)x * Y + y) * Z + z] = tile[tx][r];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    float l2 = __shfl_xor(lo, o, 64), h2 = __shfl_xor(hi, o, 64);
    lo = l2 < lo ? l2 : lo;
    hi = h2 > hi ? h2 : hi;
  }
  if ((threadIdx.x & 63) == 0) { s_lo[threadIdx.x >> 6] = lo; s_hi[threadIdx.x >> 6] = hi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      lo = s_lo[w] < lo ? s_lo[w] : lo;
      hi = s_hi[w] > hi ? s_hi[w] : hi;
    }
    const int64_t b = ((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    part[b] = make_float2(lo, hi);
  }
}

__global__ void minmax_final_kernel(const float2 *part, int64_t n, float2 *out) {
  __shared__ float s_lo[16], s_hi[16];
  float lo = 100000.0f, hi = -100000.0f;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    float2 v = part[i];
    lo = v.x < lo ? v.x : lo;
    hi = v.y > hi ? v.y : hi;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    float l2 = __shfl_xor(lo, o, 64), h2 = __shfl_xor(hi, o, 64);
    lo = l2 < lo ? l2 : lo;
    hi = h2 > hi ? h2 : hi;
  }
  if ((threadIdx.x & 63) == 0) { s_lo[threadIdx.x >> 6] = lo; s_hi[threadIdx.x >> 6] = hi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      lo = s_lo[w] < lo ? s_lo[w] : lo;
      hi = s_hi[w] > hi ? s_hi[w] : hi;
    }
    *out = make_float2(lo, hi);
  }
}

constexpr int kBins = 65536;

__global__ __launch_bounds__(256) void uniform_hist_kernel(const float *img, int64_t n, float gmin,
                                                           float oneover, uint32_t *hist) {
  const float fn = (float)n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (!((float)i < fn)) continue;  // the reference's float loop bound
    int idx = (int)(((float)kBins * (img[i] - gmin)) * oneover);
    idx = idx > kBins - 1 ? kBins - 1 : idx;
    idx = idx < 0 ? 0 : idx;
    atomicAdd(&hist[idx], 1u);
  }
}

}  // namespace

extern "C" {

int iso_volume_ingest(iso_ctx *ctx, const float *d_file, int X, int Y, int Z, float cfg_min,
                      float cfg_max, float *d_img, float *h_real_min, float *h_real_max) {
  ISO_TRY(iso::enter(ctx));
  if (!d_file || !d_img || X <= 0 || Y <= 0 || Z <= 0 || d_file == d_img)
    return iso::fail(ctx, ISO_ERR_INVALID_ARG, "iso_volume_ingest: bad arguments");
  dim3 grid((X + kT - 1) / kT, (Z + kT - 1) / kT, Y);
  const int64_t nb = (int64_t)grid.x * grid.y * grid.z;
  void *part = nullptr, *res = nullptr;
  ISO_TRY(iso::ws_get(ctx, WS_INTERVAL_PART, sizeof(float2) * (size_t)nb, false, &part));
  ISO_TRY(iso::ws_get(ctx, WS_SMALL, 64, false, &res));
  hipLaunchKernelGGL(ingest_kernel, grid, dim3(256), 0, ctx->stream, d_file, X, Y, Z, cfg_min,
                     cfg_max, d_img, (float2 *)part);
  hipLaunchKernelGGL(minmax_final_kernel, dim3(1), dim3(1024), 0, ctx->stream,
                     (const float2 *)part, nb, (float2 *)res);
  ISO_HIP_CHECK(ctx, hipGetLastError());
  float2 mm;
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(&mm, res, sizeof mm, hipMemcpyDeviceToHost, ctx->stream));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (h_real_min) *h_real_min = mm.x;
  if (h_real_max) *h_real_max = mm.y;
  return ISO_OK;
}

int iso_equalized_isovalues(iso_ctx *ctx, const float *d_img, int64_t n, float real_min,
                            float real_max, int k, float *h_isos) {
  ISO_TRY(iso::enter(ctx));
  if (!d_img || n <= 0 || k <= 0 || !h_isos)
    return iso::fail(ctx, ISO_ERR_INVALID_ARG, "iso_equalized_isovalues: bad arguments");
  void *dh = nullptr;
  ISO_TRY(iso::ws_get(ctx, WS_MC_KEYS_B, sizeof(uint32_t) * kBins, false, &dh));
  ISO_TRY(iso::fill_u32(ctx, dh, 0u, kBins));
  const float oneover = (float)(1.0 / (double)(real_max - real_min));
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(uniform_hist_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, d_img,
                     n, real_min, oneover, (uint32_t *)dh);
  ISO_HIP_CHECK(ctx, hipGetLastError());
  std::vector<int> hist(kBins);
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(hist.data(), dh, sizeof(int) * kBins, hipMemcpyDeviceToHost,
                                    ctx->stream));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  // equalisation walk (Single.cpp:362-389), bins relabelled with the isovalue index
  double sum = 0.0, acc = 0.0, threshold;
  for (int i = 0; i < kBins; ++i) sum += hist[i];
  int index = 0;
  threshold = sum / double(k);
  for (int i = 0; i < k; ++i) {
    while ((acc == 0.0 || acc + hist[index] < threshold) && index < kBins - 1) {
      acc += hist[index];
      hist[index] = i;
      index++;
    }
    sum -= acc;
    acc = 0.0;
    if (i < k - 1) threshold = sum / double(k - (i + 1));
  }
  // bins -> isovalues (Single.cpp:304-319)
  int idx = 0, last = 0;
  const float step = (real_max - real_min) / (float)kBins;
  for (int i = 0; i < k; ++i) {
    while (idx < kBins && hist[idx] <= i) ++idx;
    const int range = idx - last;
    const int mid = last + range / 2;
    const float off = step * (float)mid;
    h_isos[i] = real_min + off;
    last = idx;
  }
  return ISO_OK;
}

}  // extern "C"
