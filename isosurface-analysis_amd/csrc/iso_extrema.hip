// Local extrema of a volume: the values of the voxels that no neighbour both
// undercuts and exceeds, in voxel order.
//
// Reference: findExtrema_h / find_extrema_g / checkNeighborhood
// (isosurfaces_cuda/src/IsosurfaceExtrema.cu:22-138), compaction with
// is_not_dummy_value (hdr/thrustWrapper.h:29-41):
//   - voxel idx = (x, y, z), z fastest (findPointForIndex, hdr/helpers.h:68-77);
//   - its value and its neighbours' are read at x + y*dimY + z*dimY*dimZ, the
//     reference's own index (for a cube it is the x-fastest index);
//   - neighbours are the voxels x-1..x+3, y-1..y+3, z-1..z+3 inside the grid,
//     the voxel itself included; the voxel keeps its value unless more than 5
//     of them equal it or some are lower and some higher, which give -1e10;
//   - values above -1e10 are kept in voxel order (a stable compaction).
// One lane per voxel with the 125 reads served by L1/L2, then a flag scan and
// a scatter of the kept values.
#include "iso_internal.hpp"

namespace iso {
int exclusive_scan_u32(iso_ctx *ctx, const uint32_t *d_in, uint32_t *d_out, int64_t n,
                       uint32_t *d_total, uint32_t *d_tmp);
size_t scan_tmp_words(int64_t n);
}  // namespace iso

namespace {

constexpr float kDummy = -1e10f;  // DUMMY_VALUE (thrustWrapper.h:29), exact in float

__global__ void extrema_kernel(const float *__restrict__ data, int X, int Y, int Z, int64_t N,
                               float *__restrict__ val, uint32_t *__restrict__ keep) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N) return;
  const int z = (int)(idx % Z), y = (int)((idx / Z) % Y), x = (int)(idx / ((int64_t)Y * Z));
  const int64_t sy = Y, sz = (int64_t)Y * Z;
  const float c = data[x + y * sy + z * sz];
  int higher = 0, lower = 0, same = 0;
  for (int i = -1; i < 4; ++i) {
    const int nx = x + i;
    if (nx < 0 || nx >= X) continue;
    for (int j = -1; j < 4; ++j) {
      const int ny = y + j;
      if (ny < 0 || ny >= Y) continue;
      const int64_t row = nx + ny * sy;
#pragma unroll
      for (int k = -1; k < 4; ++k) {
        const int nz = z + k;
        if (nz < 0 || nz >= Z) continue;
        const float v = data[row + nz * sz];
        higher += v > c;
        lower += v < c;
        same += v == c;
      }
    }
  }
  float out = c;
  if (same > 5) out = kDummy;
  if (lower > 0 && higher > 0) out = kDummy;
  val[idx] = out;
  keep[idx] = out > kDummy ? 1u : 0u;
}

__global__ void extrema_scatter_kernel(const float *val, const uint32_t *keep, const uint32_t *pos,
                                       int64_t N, float *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N && keep[i]) out[pos[i]] = val[i];
}

}  // namespace

extern "C" {

int iso_find_extrema(iso_ctx *ctx, const float *d_data, int X, int Y, int Z, float *d_out,
                     int64_t cap, int64_t *h_count) {
  ISO_TRY(iso::enter(ctx));
  if (!d_data || !h_count || X <= 0 || Y <= 0 || Z <= 0 || cap < 0 || (cap > 0 && !d_out))
    return iso::fail(ctx, ISO_ERR_INVALID_ARG, "iso_find_extrema: bad arguments");
  const int64_t N = (int64_t)X * Y * Z;
  // the reference's index x + y*Y + z*Y*Z must stay inside the buffer
  const int64_t last = (int64_t)(X - 1) + (int64_t)(Y - 1) * Y + (int64_t)(Z - 1) * Y * Z;
  if (last >= N)
    return iso::fail(ctx, ISO_ERR_UNSUPPORTED,
                     "iso_find_extrema: %dx%dx%d: the reference's index x + y*%d + z*%d*%d "
                     "reaches %lld, past the %lld values of the volume",
                     X, Y, Z, Y, Y, Z, (long long)last, (long long)N);
  if (N >= ((int64_t)1 << 32))
    return iso::fail(ctx, ISO_ERR_UNSUPPORTED, "iso_find_extrema: 2^32 or more voxels");
  void *val = nullptr, *keep = nullptr, *pos = nullptr, *tmp = nullptr, *hb = nullptr;
  ISO_TRY(iso::ws_get(ctx, WS_PIPE_E, sizeof(float) * (size_t)N, false, &val));
  ISO_TRY(iso::ws_get(ctx, WS_PIPE_F, sizeof(uint32_t) * (size_t)N, false, &keep));
  ISO_TRY(iso::ws_get(ctx, WS_PIPE_G, sizeof(uint32_t) * (size_t)(N + 1), false, &pos));
  ISO_TRY(iso::ws_get(ctx, WS_SCAN_TMP, sizeof(uint32_t) * iso::scan_tmp_words(N), false, &tmp));
  const unsigned blocks = (unsigned)((N + 255) / 256);
  hipLaunchKernelGGL(extrema_kernel, dim3(blocks), dim3(256), 0, ctx->stream, d_data, X, Y, Z, N,
                     (float *)val, (uint32_t *)keep);
  ISO_HIP_CHECK(ctx, hipGetLastError());
  uint32_t *d_total = (uint32_t *)pos + N;
  ISO_TRY(iso::exclusive_scan_u32(ctx, (const uint32_t *)keep, (uint32_t *)pos, N, d_total,
                                  (uint32_t *)tmp));
  ISO_TRY(iso::pinned_get(ctx, 16, &hb));
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(hb, d_total, 4, hipMemcpyDeviceToHost, ctx->stream));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  const int64_t n = *(const uint32_t *)hb;
  *h_count = n;
  if (n > cap) return iso::fail(ctx, ISO_ERR_CAPACITY, "iso_find_extrema: %lld values", (long long)n);
  if (n > 0)
    hipLaunchKernelGGL(extrema_scatter_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                       (const float *)val, (const uint32_t *)keep, (const uint32_t *)pos, N, d_out);
  ISO_HIP_CHECK(ctx, hipGetLastError());
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return ISO_OK;
}

}  // extern "C"
