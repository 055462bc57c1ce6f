// The batched ISM pipeline: one volume, k isovalues, one call.
//
// The reference runs, per isovalue, an upload of the whole volume, an
// extraction, a device->host->device round trip of the points, an LBVH
// build + traversal and a download of the field; then one histogram launch
// per 1M-voxel chunk (runtime/src/Single.cpp:193-231, 445-456).
//
// Here everything stays in HBM and the k isovalues travel together:
//   1. one extraction pass counts every isovalue's vertices (one read of the
//      volume for all k), the host sizes the segments (the only sync);
//   2. a second pass writes the vertices of all isosurfaces end to end;
//   3. all k distance fields: one normalise, one radix sort by (isosurface,
//      Morton code), one Karras build, one bottom-up pass, one traversal
//      launch over (bricks x isosurfaces);
//   4. field intervals for all fields, then every unordered pair in the
//      fused histogram + NMI kernel.
// Empty isosurfaces are skipped and the map is over the non-empty ones
// (Single.cpp:212-222, 240); their ids come back in h_ids.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "iso_df.hpp"

namespace iso {
int mc_multi(iso_ctx *ctx, const float *d_vol, int gx, int gy, int gz, uint32_t flags,
             const float *d_isos_sorted, const int *d_slot, int k, uint32_t *d_counts,
             unsigned long long *d_stats, float *d_out, const int64_t *d_seg_off,
             const int64_t *d_seg_cap);
int approx_multi(iso_ctx *ctx, const float *d_img, int X, int Y, int Z, const float *d_isos_sorted,
                 const int *d_slot, int k, uint32_t *d_counts, int32_t *d_out,
                 const int64_t *d_seg_off, const int64_t *d_seg_cap);
int field_interval(iso_ctx *ctx, const float *d_fields, int nfields, int64_t F, float *d_mins,
                   float *d_maxs);
int hist_nmi_tasks(iso_ctx *ctx, const float *d_fields, int n, int64_t F, int H, int multi,
                   uint32_t flags, const float *d_mins, const float *d_maxs, float *d_map,
                   const int2 *d_tasks, int ntasks, uint32_t *d_joint, uint32_t *d_rows,
                   uint32_t *d_cols);
void build_tasks(int n, int multi, std::vector<int2> &t);
}  // namespace iso

namespace {

// Layout of the per-call argument block (device copy of a pinned host block).
struct IsmBlock {
  static size_t bytes(int k) {
    return align(sizeof(float) * k) + align(sizeof(int) * k) + align(sizeof(uint32_t) * k) +
           align(sizeof(unsigned long long) * 2 * k) + 3 * align(sizeof(int64_t) * (k + 1));
  }
  static size_t align(size_t b) { return (b + 255) & ~size_t(255); }
  float *isos;
  int *slot;
  uint32_t *counts;
  unsigned long long *stats;
  int64_t *seg_off, *seg_cap, *ne_off;
  void carve(void *base, int k) {
    char *p = (char *)base;
    isos = (float *)p; p += align(sizeof(float) * k);
    slot = (int *)p; p += align(sizeof(int) * k);
    counts = (uint32_t *)p; p += align(sizeof(uint32_t) * k);
    stats = (unsigned long long *)p; p += align(sizeof(unsigned long long) * 2 * k);
    seg_off = (int64_t *)p; p += align(sizeof(int64_t) * (k + 1));
    seg_cap = (int64_t *)p; p += align(sizeof(int64_t) * (k + 1));
    ne_off = (int64_t *)p;
  }
};

struct StageTimer {
  iso_ctx *ctx;
  bool on;
  hipEvent_t ev[6];
  int n = 0;
  StageTimer(iso_ctx *c, bool enable) : ctx(c), on(enable) {
    if (on)
      for (auto &e : ev) (void)hipEventCreate(&e);
  }
  ~StageTimer() {
    if (on)
      for (auto &e : ev) (void)hipEventDestroy(e);
  }
  void mark() {
    if (on && n < 6) (void)hipEventRecord(ev[n++], ctx->stream);
  }
  double ms(int a, int b) {
    float t = 0;
    if (on && b < n) (void)hipEventElapsedTime(&t, ev[a], ev[b]);
    return t;
  }
};

}  // namespace

extern "C" {

int iso_ism_volume(iso_ctx *ctx, const float *d_vol, int X, int Y, int Z, const float *h_isos,
                   int k, int extractor, int ds, int H, int df_method, uint32_t flags,
                   float *d_map, float *d_fields, int *h_ids, iso_ism_stats *h_stats) {
  ISO_TRY(iso::enter(ctx));
  if (!d_vol || !h_isos || !d_map || !h_ids || k <= 0 || X < 2 || Y < 2 || Z < 2 || ds <= 0 ||
      X < ds || Y < ds || Z < ds)
    return iso::fail(ctx, ISO_ERR_INVALID_ARG, "iso_ism_volume: bad arguments");
  if (extractor != ISO_EXTRACT_MC && extractor != ISO_EXTRACT_APPROX)
    return iso::fail(ctx, ISO_ERR_INVALID_ARG, "iso_ism_volume: unknown extractor %d", extractor);
  if (df_method != ISO_DF_LBVH && df_method != ISO_DF_BRUTE)
    return iso::fail(ctx, ISO_ERR_INVALID_ARG, "iso_ism_volume: unknown DF method %d", df_method);
  if (H < 1 || H > 256) return iso::fail(ctx, ISO_ERR_UNSUPPORTED, "iso_ism_volume: H=%d", H);
  StageTimer tm(ctx, h_stats != nullptr);

  // ---- isovalues, ascending, with their caller slots
  std::vector<int> order(k);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return h_isos[a] < h_isos[b]; });
  const size_t bb = IsmBlock::bytes(k);
  void *dblk = nullptr, *hblk = nullptr;
  ISO_TRY(iso::ws_get(ctx, WS_PIPE_A, bb, false, &dblk));
  ISO_TRY(iso::pinned_get(ctx, bb, &hblk));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  IsmBlock H_, D_;
  H_.carve(hblk, k);
  D_.carve(dblk, k);
  for (int j = 0; j < k; ++j) {
    H_.isos[j] = h_isos[order[j]];
    H_.slot[j] = order[j];
    H_.seg_off[j] = 0;
    H_.seg_cap[j] = 0;
  }
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(dblk, hblk, bb, hipMemcpyHostToDevice, ctx->stream));
  tm.mark();

  // ---- 1. count
  const bool mc = extractor == ISO_EXTRACT_MC;
  if (mc)
    ISO_TRY(iso::mc_multi(ctx, d_vol, X, Y, Z, flags, D_.isos, D_.slot, k, D_.counts, D_.stats,
                          nullptr, D_.seg_off, D_.seg_cap));
  else
    ISO_TRY(iso::approx_multi(ctx, d_vol, X, Y, Z, D_.isos, D_.slot, k, D_.counts, nullptr,
                              D_.seg_off, D_.seg_cap));
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(H_.counts, D_.counts, sizeof(uint32_t) * k,
                                    hipMemcpyDeviceToHost, ctx->stream));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));

  // ---- 2. segments of the non-empty isosurfaces, in caller order
  int m = 0;
  int64_t total = 0;
  for (int j = 0; j < k; ++j) {
    H_.seg_off[j] = total;
    H_.seg_cap[j] = H_.counts[j];
    if (H_.counts[j] > 0) {
      h_ids[m] = j;
      H_.ne_off[m] = total;
      ++m;
      total += H_.counts[j];
    }
  }
  H_.ne_off[m] = total;
  if (h_stats) {
    memset(h_stats, 0, sizeof(*h_stats));
    h_stats->n_nonempty = m;
    h_stats->total_points = total;
  }
  if (m == 0) return ISO_OK;
  const size_t pbytes = (size_t)total * 3 * 4;
  void *pts = nullptr;
  ISO_TRY(iso::ws_get(ctx, WS_PIPE_B, pbytes, false, &pts));
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(D_.seg_off, H_.seg_off, 3 * IsmBlock::align(sizeof(int64_t) * (k + 1)),
                                    hipMemcpyHostToDevice, ctx->stream));
  if (mc)
    ISO_TRY(iso::mc_multi(ctx, d_vol, X, Y, Z, flags, D_.isos, D_.slot, k, D_.counts, D_.stats,
                          (float *)pts, D_.seg_off, D_.seg_cap));
  else
    ISO_TRY(iso::approx_multi(ctx, d_vol, X, Y, Z, D_.isos, D_.slot, k, D_.counts, (int32_t *)pts,
                              D_.seg_off, D_.seg_cap));
  tm.mark();

  // ---- 3. distance fields of all non-empty isosurfaces
  iso::SampleGrid g{X, Y, Z, ds, X / ds, Y / ds, Z / ds};
  const int64_t F = (int64_t)g.nx * g.ny * g.nz;
  float *fields = d_fields;
  if (!fields) {
    void *f = nullptr;
    ISO_TRY(iso::ws_get(ctx, WS_PIPE_C, sizeof(float) * (size_t)F * m, false, &f));
    fields = (float *)f;
  }
  ISO_TRY(iso::df_segments(ctx, pts, !mc, total, D_.ne_off, m, g, df_method, fields));
  tm.mark();

  // ---- 4. intervals, then all unordered pairs: histogram + NMI
  void *mm = nullptr, *dt = nullptr;
  ISO_TRY(iso::ws_get(ctx, WS_PIPE_D, sizeof(float) * 2 * (size_t)m, false, &mm));
  ISO_TRY(iso::field_interval(ctx, fields, m, F, (float *)mm, (float *)mm + m));
  std::vector<int2> tasks;
  iso::build_tasks(m, 0, tasks);
  const size_t tb = sizeof(int2) * tasks.size();
  ISO_TRY(iso::ws_get(ctx, WS_HIST_TASKS, tb, false, &dt));
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(dt, tasks.data(), tb, hipMemcpyHostToDevice, ctx->stream));
  tm.mark();
  ISO_TRY(iso::hist_nmi_tasks(ctx, fields, m, F, H, 0, flags, (const float *)mm,
                              (const float *)mm + m, d_map, (const int2 *)dt, (int)tasks.size(),
                              nullptr, nullptr, nullptr));
  tm.mark();
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (h_stats) {
    h_stats->ms_extract = tm.ms(0, 1);
    h_stats->ms_df = tm.ms(1, 2);
    h_stats->ms_hist = tm.ms(3, 4);  // the histogram + NMI kernel launches only
  }
  return ISO_OK;
}

}  // extern "C"
