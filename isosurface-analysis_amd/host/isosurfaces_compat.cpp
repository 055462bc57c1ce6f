// Host-side mirror of the reference's device-library interface on libiso_amd.so.
// See isosurfaces_compat.hpp for the contract.
#include "isosurfaces_compat.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

namespace {

std::mutex g_mu;
iso_ctx *g_ctx = nullptr;

// grow-only device buffers reused across calls
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  void *get(size_t need) {
    if (need > bytes) {
      if (p) iso_device_free(g_ctx, p);
      p = nullptr;
      bytes = 0;
      size_t n = need + need / 4 + 4096;
      check(iso_device_alloc(g_ctx, &p, n), "device allocation");
      bytes = n;
    }
    return p;
  }
  static void check(int st, const char *what);
};
DevBuf g_in, g_out, g_aux, g_aux2;

void die(int st, const char *what) {
  fprintf(stderr, "isosurfaces (libiso_amd) error in %s: %s %s\n", what, iso_status_string(st),
          g_ctx ? iso_ctx_last_error(g_ctx) : "");
  exit(EXIT_FAILURE);
}

void DevBuf::check(int st, const char *what) {
  if (st != ISO_OK) die(st, what);
}

void ok(int st, const char *what) {
  if (st != ISO_OK) die(st, what);
}

void ensure_ctx() {
  if (!g_ctx) ok(iso_ctx_create(&g_ctx, 0), "iso_ctx_create");
}

void df_host_points(const float *h_xyz, long n, bool normalised, int X, int Y, int Z, int ds,
                    float *h_distances) {
  const size_t F = (size_t)(X / ds) * (Y / ds) * (Z / ds);
  float *d_pts = (float *)g_in.get(sizeof(float) * 3 * (size_t)std::max(n, 1L));
  float *d_df = (float *)g_out.get(sizeof(float) * F);
  ok(iso_copy_to_device(g_ctx, d_pts, h_xyz, sizeof(float) * 3 * (size_t)n), "upload points");
  int st = normalised ? iso_df_points_normalized(g_ctx, d_pts, n, X, Y, Z, ds, ISO_DF_LBVH, d_df)
                      : iso_df_points_f3(g_ctx, d_pts, n, X, Y, Z, ds, ISO_DF_LBVH, d_df);
  ok(st, "distance field");
  ok(iso_copy_to_host(g_ctx, h_distances, d_df, sizeof(float) * F), "download field");
}

}  // namespace

iso_ctx *isosurfaces_compat_context() {
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_ctx();
  return g_ctx;
}

extern "C" long approximate_isosurfaces_h(float *h_data, float isovalue, const int dimX,
                                          const int dimY, const int dimZ, const int scale,
                                          int3 **h_points) {
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_ctx();
  if (scale != 1) die(ISO_ERR_UNSUPPORTED, "approximate_isosurfaces_h (scale != 1)");
  const size_t N = (size_t)dimX * dimY * dimZ;
  float *d_img = (float *)g_in.get(sizeof(float) * N);
  ok(iso_copy_to_device(g_ctx, d_img, h_data, sizeof(float) * N), "upload volume");
  int64_t n = 0;
  size_t cap = g_out.bytes / sizeof(int3);
  int st = iso_approx_extract(g_ctx, d_img, dimX, dimY, dimZ, isovalue,
                              (int32_t *)g_out.get(sizeof(int3) * std::max<size_t>(cap, 1)),
                              (int64_t)std::max<size_t>(cap, 1), &n);
  if (st == ISO_ERR_CAPACITY) {
    int32_t *d_out = (int32_t *)g_out.get(sizeof(int3) * (size_t)n);
    st = iso_approx_extract(g_ctx, d_img, dimX, dimY, dimZ, isovalue, d_out, n, &n);
  }
  ok(st, "approximate_isosurfaces_h");
  *h_points = (int3 *)malloc(sizeof(int3) * (size_t)std::max<int64_t>(n, 1));
  if (!*h_points) die(ISO_ERR_OUT_OF_MEMORY, "approximate_isosurfaces_h (host malloc)");
  ok(iso_copy_to_host(g_ctx, *h_points, g_out.p, sizeof(int3) * (size_t)n), "download cells");
  return (long)n;
}

long approximate_isosurfaces_W(float *data, float isovalue, const int dimX, const int dimY,
                               const int dimZ, const int scale, int3 **h_points) {
  return approximate_isosurfaces_h(data, isovalue, dimX, dimY, dimZ, scale, h_points);
}

extern "C" void calculate_histogram_h(float **distanceFields, float *minValues, float *maxValues,
                                      const int fieldSize, const unsigned long numFields,
                                      const unsigned long histogramSize, unsigned int *h_jointHist,
                                      unsigned int *h_colSums, unsigned int *h_rowSums,
                                      float *h_simMap, const int simMapSize, const bool multi) {
  (void)h_jointHist;  // the reference never copies joint histograms back
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_ctx();
  const size_t n = numFields, nf = multi ? 2 * n : n, F = (size_t)fieldSize, H = histogramSize;
  float *d_fields = (float *)g_in.get(sizeof(float) * F * nf);
  for (size_t i = 0; i < nf; ++i)
    ok(iso_copy_to_device(g_ctx, d_fields + i * F, distanceFields[i], sizeof(float) * F),
       "upload field");
  float *d_mm = (float *)g_aux.get(sizeof(float) * 2 * nf);
  ok(iso_copy_to_device(g_ctx, d_mm, minValues, sizeof(float) * nf), "upload mins");
  ok(iso_copy_to_device(g_ctx, d_mm + nf, maxValues, sizeof(float) * nf), "upload maxs");
  const size_t sums = n * n * H;
  char *d_out = (char *)g_out.get(sizeof(float) * n * n + 2 * sizeof(uint32_t) * sums);
  float *d_map = (float *)d_out;
  uint32_t *d_rows = (uint32_t *)(d_out + sizeof(float) * n * n);
  uint32_t *d_cols = d_rows + sums;
  ok(iso_hist_nmi(g_ctx, d_fields, (int)n, (int64_t)F, (int)H, multi ? 1 : 0, ISO_COMPAT_ALL, d_mm,
                  d_mm + nf, d_map, nullptr, d_rows, d_cols),
     "calculate_histogram_h");
  std::vector<float> map(n * n);
  ok(iso_copy_to_host(g_ctx, map.data(), d_map, sizeof(float) * n * n), "download map");
  if (h_rowSums) ok(iso_copy_to_host(g_ctx, h_rowSums, d_rows, sizeof(uint32_t) * sums), "rows");
  if (h_colSums) ok(iso_copy_to_host(g_ctx, h_colSums, d_cols, sizeof(uint32_t) * sums), "cols");
  // simMap[id1 * simMapSize + id2] (JointHistogram.cu:157-158)
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) h_simMap[i * (size_t)simMapSize + j] = map[i * n + j];
}

extern "C" void generateHierarchy_h(std::vector<float3> points, const int dimX, const int dimY,
                                    const int dimZ, float *h_distances, const int numSamples,
                                    const int dfDownscale, const bool sampleGiven,
                                    const unsigned long sampleSize,
                                    std::vector<float3> samplePoints) {
  (void)numSamples;
  (void)sampleSize;
  (void)samplePoints;
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_ctx();
  if (sampleGiven) die(ISO_ERR_UNSUPPORTED, "generateHierarchy_h (given sample points)");
  static_assert(sizeof(float3) == 12, "float3 is three packed floats");
  df_host_points((const float *)points.data(), (long)points.size(), true, dimX, dimY, dimZ,
                 dfDownscale, h_distances);
}

void generateHierarchy_W(std::vector<float3> points, const int dimX, const int dimY, const int dimZ,
                         float *distances, int numSamples, const int dfDownscale) {
  (void)numSamples;
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_ctx();
  df_host_points((const float *)points.data(), (long)points.size(), false, dimX, dimY, dimZ,
                 dfDownscale, distances);
}

void generateHierarchyi3_W(int3 *points, const int dimX, const int dimY, const int dimZ,
                           float *distances, const long numPoints, const int numSamples,
                           const int dfDownscale) {
  (void)numSamples;
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_ctx();
  const size_t F = (size_t)(dimX / dfDownscale) * (dimY / dfDownscale) * (dimZ / dfDownscale);
  int32_t *d_pts = (int32_t *)g_in.get(sizeof(int3) * (size_t)std::max(numPoints, 1L));
  float *d_df = (float *)g_out.get(sizeof(float) * F);
  ok(iso_copy_to_device(g_ctx, d_pts, points, sizeof(int3) * (size_t)numPoints), "upload cells");
  ok(iso_df_points_i3(g_ctx, d_pts, numPoints, dimX, dimY, dimZ, dfDownscale, ISO_DF_LBVH, d_df),
     "distance field");
  ok(iso_copy_to_host(g_ctx, distances, d_df, sizeof(float) * F), "download field");
}

void cudaMarchingCube(const float &iso_val, const vec3i &dim, float *data, const int &maxVerts,
                      std::vector<vec4f> &ret_points) {
  (void)maxVerts;
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_ctx();
  const size_t N = (size_t)dim.x * dim.y * dim.z;
  float *d_vol = (float *)g_aux2.get(sizeof(float) * N);
  ok(iso_copy_to_device(g_ctx, d_vol, data, sizeof(float) * N), "upload volume");
  int64_t n = 0;
  size_t cap = std::max<size_t>(g_out.bytes / 12, 1);
  const uint32_t flags = ISO_COMPAT_MC | ISO_MC_SORTED;
  int st = iso_mc_extract(g_ctx, d_vol, dim.x, dim.y, dim.z, iso_val, flags,
                          (float *)g_out.get(12 * cap), (int64_t)cap, &n, nullptr);
  if (st == ISO_ERR_CAPACITY)
    st = iso_mc_extract(g_ctx, d_vol, dim.x, dim.y, dim.z, iso_val, flags,
                        (float *)g_out.get(12 * (size_t)n), n, &n, nullptr);
  ok(st, "cudaMarchingCube");
  std::vector<float> xyz(3 * (size_t)n);
  ok(iso_copy_to_host(g_ctx, xyz.data(), g_out.p, 12 * (size_t)n), "download vertices");
  ret_points.resize((size_t)n);
  for (int64_t i = 0; i < n; ++i)
    ret_points[i] = vec4f{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 1.0f};
}

extern "C" void calculate_distances_h(std::vector<float3> points, int dimX, int dimY, int dimZ,
                                      int dfDownscale, float *h_distances) {
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_ctx();
  const long n = (long)points.size();
  const size_t F = (size_t)dimX * dimY * dimZ;  // SIZE (DistanceField.cu:71)
  float *d_pts = (float *)g_in.get(sizeof(float) * 3 * (size_t)std::max(n, 1L));
  float *d_df = (float *)g_out.get(sizeof(float) * std::max<size_t>(F, 1));
  if (n > 0)
    ok(iso_copy_to_device(g_ctx, d_pts, points.data(), sizeof(float) * 3 * (size_t)n),
       "upload points");
  ok(iso_df_voxel_units(g_ctx, d_pts, n, dimX, dimY, dimZ, dfDownscale, d_df),
     "calculate_distances_h");
  ok(iso_copy_to_host(g_ctx, h_distances, d_df, sizeof(float) * F), "download field");
}

extern "C" long findExtrema_h(float *h_data, int3 dimensions, float **h_points) {
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_ctx();
  const int X = dimensions.x, Y = dimensions.y, Z = dimensions.z;
  const size_t N = (size_t)X * Y * Z;
  float *d_vol = (float *)g_aux2.get(sizeof(float) * std::max<size_t>(N, 1));
  ok(iso_copy_to_device(g_ctx, d_vol, h_data, sizeof(float) * N), "upload volume");
  int64_t n = 0;
  size_t cap = std::max<size_t>(g_out.bytes / 4, 1);
  int st = iso_find_extrema(g_ctx, d_vol, X, Y, Z, (float *)g_out.get(4 * cap), (int64_t)cap, &n);
  if (st == ISO_ERR_CAPACITY)
    st = iso_find_extrema(g_ctx, d_vol, X, Y, Z, (float *)g_out.get(4 * (size_t)std::max<int64_t>(n, 1)),
                          n, &n);
  ok(st, "findExtrema_h");
  // the callee mallocs the values, the caller frees them (IsosurfaceExtrema.cu:130)
  *h_points = (float *)malloc(sizeof(float) * (size_t)std::max<int64_t>(n, 1));
  if (!*h_points) die(ISO_ERR_OUT_OF_MEMORY, "findExtrema_h (host values)");
  if (n > 0) ok(iso_copy_to_host(g_ctx, *h_points, g_out.p, sizeof(float) * (size_t)n), "values");
  return (long)n;
}
