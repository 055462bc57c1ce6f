// The batched ISM pipeline: one volume, k isovalues, one call.
//
// The reference runs, per isovalue, an upload of the whole volume, an
// extraction, a device->host->device round trip of the points, an LBVH
// build + traversal and a download of the field; then one histogram launch
// per 1M-voxel chunk (runtime/src/Single.cpp:193-231, 445-456).
//
// Here everything stays in HBM and the k isovalues travel together:
//   1. one extraction pass writes the vertices of all k isosurfaces into one
//      pool, each tagged with its isovalue slot (one read of the volume);
//   2. all k distance fields: one normalise, one radix sort by (slot, Morton
//      code) that regroups the pool, one Karras build, one bottom-up sweep,
//      one traversal launch over (bricks x isosurfaces);
//   3. field intervals for all fields, then every unordered pair in the
//      fused histogram + NMI kernel.
// The host synchronises once after extraction (set sizes) and once at the
// end.  Empty isosurfaces are skipped and the map is over the non-empty ones
// (Single.cpp:212-222, 240); their ids come back in h_ids.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "iso_df.hpp"

namespace iso {
int mc_multi(iso_ctx *ctx, const float *d_vol, int gx, int gy, int gz, uint32_t flags,
             const float *d_isos_sorted, const int *d_slot, int k, uint32_t *d_counts,
             unsigned long long *d_stats, float *d_pool, uint32_t *d_pool_slot, int64_t cap,
             int64_t *h_total, uint32_t *h_counts);
int approx_multi(iso_ctx *ctx, const float *d_img, int X, int Y, int Z, const float *d_isos_sorted,
                 const int *d_slot, int k, uint32_t *d_counts, int32_t *d_pool,
                 uint32_t *d_pool_slot, int64_t cap, int64_t *h_total);
int field_interval(iso_ctx *ctx, const float *d_fields, int nfields, int64_t F, float *d_mins,
                   float *d_maxs);
int hist_nmi_tasks(iso_ctx *ctx, const float *d_fields, int n, int64_t F, int H, int multi,
                   uint32_t flags, const float *d_mins, const float *d_maxs, float *d_map,
                   const int2 *d_tasks, int ntasks, uint32_t *d_joint, uint32_t *d_rows,
                   uint32_t *d_cols);
void build_tasks(int n, int multi, std::vector<int2> &t);
}  // namespace iso

namespace {

// Per-call argument block: ascending isovalues, their caller slots, per-slot
// vertex counts and cell statistics (device copy of a host block).
struct IsmBlock {
  static size_t align(size_t b) { return (b + 255) & ~size_t(255); }
  static size_t bytes(int k) {
    return align(sizeof(float) * k) + align(sizeof(int) * k) + align(sizeof(uint32_t) * k) +
           align(sizeof(unsigned long long) * 2 * k);
  }
  float *isos;
  int *slot;
  uint32_t *counts;
  unsigned long long *stats;
  void carve(void *base, int k) {
    char *p = (char *)base;
    isos = (float *)p; p += align(sizeof(float) * k);
    slot = (int *)p; p += align(sizeof(int) * k);
    counts = (uint32_t *)p; p += align(sizeof(uint32_t) * k);
    stats = (unsigned long long *)p;
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
  if (h_stats) memset(h_stats, 0, sizeof(*h_stats));
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
  IsmBlock Hb, Db;
  Hb.carve(hblk, k);
  Db.carve(dblk, k);
  for (int j = 0; j < k; ++j) {
    Hb.isos[j] = h_isos[order[j]];
    Hb.slot[j] = order[j];
  }
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(dblk, hblk, bb, hipMemcpyHostToDevice, ctx->stream));
  tm.mark();

  // ---- 1. extraction into the pool (grown and redone if it overflows)
  const bool mc = extractor == ISO_EXTRACT_MC;
  int64_t total = 0;
  std::vector<uint32_t> counts(k);
  void *pool = nullptr, *pslot = nullptr;
  for (int attempt = 0; attempt < 2; ++attempt) {
    const int64_t cap = (int64_t)(ctx->ws[WS_PIPE_B].bytes / 12);
    ISO_TRY(iso::ws_get(ctx, WS_PIPE_B, 12 * (size_t)std::max<int64_t>(cap, 1), false, &pool));
    ISO_TRY(iso::ws_get(ctx, WS_PIPE_SLOTS, 4 * (size_t)std::max<int64_t>(cap, 1), false, &pslot));
    if (mc)
      ISO_TRY(iso::mc_multi(ctx, d_vol, X, Y, Z, flags, Db.isos, Db.slot, k, Db.counts, nullptr,
                            (float *)pool, (uint32_t *)pslot, cap, &total, counts.data()));
    else
      ISO_TRY(iso::approx_multi(ctx, d_vol, X, Y, Z, Db.isos, Db.slot, k, Db.counts,
                                (int32_t *)pool, (uint32_t *)pslot, cap, &total));
    if (total <= cap) break;
    ISO_TRY(iso::ws_get(ctx, WS_PIPE_B, 12 * (size_t)total, false, &pool));
    ISO_TRY(iso::ws_get(ctx, WS_PIPE_SLOTS, 4 * (size_t)total, false, &pslot));
  }
  if (!mc) {  // marching cubes hands its counts back with its total
    ISO_HIP_CHECK(ctx, hipMemcpyAsync(Hb.counts, Db.counts, sizeof(uint32_t) * k,
                                      hipMemcpyDeviceToHost, ctx->stream));
    ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    counts.assign(Hb.counts, Hb.counts + k);
  }
  tm.mark();

  // ---- 2. distance fields of all non-empty isosurfaces
  int m = 0;
  for (int j = 0; j < k; ++j)
    if (counts[j] > 0) h_ids[m++] = j;
  if (h_stats) {
    h_stats->n_nonempty = m;
    h_stats->total_points = total;
  }
  if (m == 0) return ISO_OK;
  iso::SampleGrid g{X, Y, Z, ds, X / ds, Y / ds, Z / ds};
  const int64_t F = (int64_t)g.nx * g.ny * g.nz;
  float *fields = d_fields;
  if (!fields) {
    void *f = nullptr;
    ISO_TRY(iso::ws_get(ctx, WS_PIPE_C, sizeof(float) * (size_t)F * m, false, &f));
    fields = (float *)f;
  }
  // the pair list goes up now, while the extraction still runs, so nothing is
  // left for the host between the fields and the histogram launch; it is
  // staged in the pinned block (free: the extraction's read-back synchronised
  // the stream, and the fields' set-up synchronises before reusing it)
  void *dt = nullptr, *tpin = nullptr;
  std::vector<int2> tasks;
  iso::build_tasks(m, 0, tasks);
  const size_t tb = sizeof(int2) * tasks.size();
  ISO_TRY(iso::ws_get(ctx, WS_HIST_TASKS, tb, false, &dt));
  ISO_TRY(iso::pinned_get(ctx, tb, &tpin));
  memcpy(tpin, tasks.data(), tb);
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(dt, tpin, tb, hipMemcpyHostToDevice, ctx->stream));
  // stage times requested: the walk kernel is timed too (events made once)
  if (h_stats && !ctx->walk_ev[0]) {
    ISO_HIP_CHECK(ctx, hipEventCreate(&ctx->walk_ev[0]));
    ISO_HIP_CHECK(ctx, hipEventCreate(&ctx->walk_ev[1]));
  }
  ISO_TRY(iso::df_pool(ctx, pool, !mc, total, (const uint32_t *)pslot, k, counts.data(), g,
                       df_method, fields, false, /*coherent=*/true));
  tm.mark();

  // ---- 3. intervals, then all unordered pairs: histogram + NMI
  void *mm = nullptr;
  ISO_TRY(iso::ws_get(ctx, WS_PIPE_D, sizeof(float) * 2 * (size_t)m, false, &mm));
  ISO_TRY(iso::field_interval(ctx, fields, m, F, (float *)mm, (float *)mm + m));
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
    if (ctx->walk_ev[0] && df_method == ISO_DF_LBVH) {
      float w = 0.0f;
      if (hipEventElapsedTime(&w, ctx->walk_ev[0], ctx->walk_ev[1]) == hipSuccess)
        h_stats->ms_walk = w;
    }
  }
  return ISO_OK;
}

}  // extern "C"
