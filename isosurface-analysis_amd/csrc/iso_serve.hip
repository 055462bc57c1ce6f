// Serving form: a request loop over HBM-resident volumes that answers ISM,
// TSM and VSM requests as they arrive.
//
// The reference runs each map as a separate process invocation: `single` per
// (variable, time step) and `multi` for the 16x16 cross maps between the
// representative fields of two volumes, re-reading every field from disk
// (runtime/src/MultiVar.cpp:76-249, JointHistogram.cu:309-318).
//
// Design:
//   - volumes are uploaded once and stay in HBM;
//   - an ISM computes all fields of a volume with the batched pipeline and
//     keeps them, the map and the volume's representatives in an HBM
//     residency cache keyed by (variable, time step); TSM/VSM requests reuse
//     cached fields and only run the cross-map histogram kernel; a request
//     that needs a volume another request is computing waits for it instead
//     of recomputing;
//   - `slots` worker threads, each with its own context (stream + workspace),
//     take requests in arrival order, so work of requests that arrive while
//     others run overlaps on the device (continuous batching);
//   - the fields of a volume live in one block of a pool allocated when the
//     server starts (mem_budget bytes, at least two blocks, at most one per
//     volume): requests take and return blocks under the server's lock and
//     never allocate or free device memory;
//   - admission is by memory footprint: a request starts when a slot is free
//     and its new fields fit the budget beside the cache and in-flight work,
//     evicting least-recently-used unpinned volumes if needed; the resident
//     volumes a request uses are pinned from its admission to its end, so its
//     estimate stays true; an idle server always admits, so no request waits
//     forever;
//   - every result carries arrival, start and completion times (steady clock).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <iterator>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "iso_internal.hpp"

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

int64_t vkey(int var, int step) { return (int64_t)var * 1000000 + step; }

struct Volume {
  float *d_img = nullptr;
  int X = 0, Y = 0, Z = 0;
  std::vector<float> isos;
};

struct Entry {
  enum State { EMPTY, COMPUTING, READY, FAILED } state = EMPTY;
  float *d_fields = nullptr;
  int m = 0;
  std::vector<int> ids;    // non-empty isovalue indices
  std::vector<float> map;  // m x m
  std::vector<int> reps;   // positions in 0..m-1 of the representatives, ascending
  size_t bytes = 0;
  int pins = 0;
  uint64_t last_use = 0;
  int status = ISO_OK;
};

struct Job {
  uint64_t ticket;
  iso_request req;
  double t_submit;
  size_t estimate;
  // resident entries the job was admitted against, pinned from admission to
  // completion so that no other admission evicts them in between
  Entry *held[2] = {nullptr, nullptr};
};

struct Done {
  int status = ISO_OK;
  int m = 0;
  int slot = -1;
  int computed = 0;             // ISMs this request computed
  int64_t computed_pairs = 0;   // pairs of their maps
  double t_submit = 0, t_start = 0, t_done = 0;
  std::vector<float> map;
  std::vector<int> ids;
};

struct Slot {
  iso_ctx *ctx = nullptr;
  // the context's autotuner counters already added to the server's stats
  int64_t tune_measured_seen = 0, tune_nearest_seen = 0;
  float *d_map = nullptr;    // k*k
  float *d_cross = nullptr;  // 2*reps*F
  float *d_mm = nullptr;     // 4*reps
};

}  // namespace

struct iso_server {
  iso_server_config cfg;
  int64_t F = 0;
  std::vector<Slot> slots;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_queue, cv_done, cv_cache;
  std::deque<Job> queue;
  std::unordered_map<uint64_t, Done> done;
  std::map<int64_t, Volume> volumes;
  std::map<int64_t, Entry> cache;
  // field blocks of k * F floats, allocated by iso_server_start
  std::vector<float *> blocks, free_blocks;
  size_t cache_bytes = 0, inflight = 0;
  bool stop = false;
  uint64_t next_ticket = 1, tick = 0;
  int busy = 0;
  iso_server_stats stats{};
};

namespace {

// Fields bytes a request would add (volumes not resident yet).
size_t estimate_locked(iso_server *s, const iso_request &r) {
  size_t per = (size_t)s->cfg.k * (size_t)s->F * sizeof(float);
  size_t b = 0;
  auto need = [&](int v, int t) {
    auto it = s->cache.find(vkey(v, t));
    if (it == s->cache.end() || it->second.state != Entry::READY) b += per;
  };
  need(r.var1, r.step1);
  if (r.kind != ISO_REQ_ISM) need(r.var2, r.step2);
  return b;
}

void evict_locked(iso_server *s, size_t want) {
  while (s->cache_bytes + s->inflight + want > s->cfg.mem_budget) {
    auto victim = s->cache.end();
    for (auto it = s->cache.begin(); it != s->cache.end(); ++it)
      if (it->second.state == Entry::READY && it->second.pins == 0 &&
          (victim == s->cache.end() || it->second.last_use < victim->second.last_use))
        victim = it;
    if (victim == s->cache.end()) return;
    s->free_blocks.push_back(victim->second.d_fields);  // back to the pool
    s->cache_bytes -= victim->second.bytes;
    s->cache.erase(victim);
    s->stats.evictions++;
  }
}

// Pins the READY entries a request will use (they are then outside the
// estimate and must stay resident until the job ends).
void hold_locked(iso_server *s, Job &j) {
  int n = 0;
  auto hold = [&](int v, int t) {
    auto it = s->cache.find(vkey(v, t));
    if (it != s->cache.end() && it->second.state == Entry::READY) {
      it->second.pins++;
      j.held[n++] = &it->second;
    }
  };
  hold(j.req.var1, j.req.step1);
  if (j.req.kind != ISO_REQ_ISM && vkey(j.req.var2, j.req.step2) != vkey(j.req.var1, j.req.step1))
    hold(j.req.var2, j.req.step2);
}

bool admissible_locked(iso_server *s, Job &j) {
  j.estimate = estimate_locked(s, j.req);
  // the entries this request would use are not eviction candidates for it
  Job probe = j;
  hold_locked(s, probe);
  evict_locked(s, j.estimate);
  for (Entry *e : probe.held)
    if (e) e->pins--;
  if (s->busy == 0) return true;  // an idle server always admits (after evicting)
  return s->cache_bytes + s->inflight + j.estimate <= s->cfg.mem_budget;
}

// A free field block, evicting the least recently used unpinned volume when
// the pool is empty, else waiting for one (a block comes back when a volume
// is evicted or a computation fails).  Called with the lock held.
float *take_block_locked(iso_server *s, std::unique_lock<std::mutex> &lk) {
  while (s->free_blocks.empty()) {
    auto victim = s->cache.end();
    for (auto it = s->cache.begin(); it != s->cache.end(); ++it)
      if (it->second.state == Entry::READY && it->second.pins == 0 &&
          (victim == s->cache.end() || it->second.last_use < victim->second.last_use))
        victim = it;
    if (victim != s->cache.end()) {
      s->free_blocks.push_back(victim->second.d_fields);
      s->cache_bytes -= victim->second.bytes;
      s->cache.erase(victim);
      s->stats.evictions++;
      break;
    }
    s->cv_cache.wait(lk);
  }
  float *b = s->free_blocks.back();
  s->free_blocks.pop_back();
  return b;
}

// Returns the pinned READY entry of a volume, computing it if needed (then
// counted in d.computed and d.computed_pairs).
int ensure_entry(iso_server *s, Slot &sl, int var, int step, Entry **out, Done &d) {
  const int64_t key = vkey(var, step);
  std::unique_lock<std::mutex> lk(s->mu);
  auto vit = s->volumes.find(key);
  if (vit == s->volumes.end()) return ISO_ERR_INVALID_ARG;
  while (true) {
    Entry &e = s->cache[key];
    if (e.state == Entry::READY) {
      e.pins++;
      e.last_use = ++s->tick;
      s->stats.cache_hits++;
      *out = &e;
      return ISO_OK;
    }
    if (e.state == Entry::COMPUTING) {
      s->cv_cache.wait(lk);
      continue;
    }
    e.state = Entry::COMPUTING;
    break;
  }
  Volume &vol = vit->second;
  float *d_fields = take_block_locked(s, lk);
  lk.unlock();

  const int k = s->cfg.k;
  int st = ISO_OK;
  std::vector<int> ids(k);
  iso_ism_stats is{};
  std::vector<float> map;
  std::vector<int> reps;
  if (st == ISO_OK)
    st = iso_ism_volume(sl.ctx, vol.d_img, vol.X, vol.Y, vol.Z, vol.isos.data(), k,
                        s->cfg.extractor, s->cfg.ds, s->cfg.H_ism, s->cfg.df_method, s->cfg.flags,
                        sl.d_map, d_fields, ids.data(), &is);
  if (st == ISO_OK) {
    map.resize((size_t)is.n_nonempty * is.n_nonempty);
    st = iso_copy_to_host(sl.ctx, map.data(), sl.d_map, sizeof(float) * map.size());
  }
  if (st == ISO_OK && is.n_nonempty > 0) {
    std::vector<int> pos(is.n_nonempty);
    for (int i = 0; i < is.n_nonempty; ++i) pos[i] = i;
    std::vector<int> sel(std::min(s->cfg.reps, is.n_nonempty));
    int got = 0;
    st = iso_select_representatives(map.data(), is.n_nonempty, pos.data(), (int)sel.size(),
                                    sel.data(), nullptr, &got);
    reps.assign(sel.begin(), sel.begin() + got);
  }

  lk.lock();
  Entry &e = s->cache[key];
  if (st == ISO_OK) {
    e.state = Entry::READY;
    e.d_fields = d_fields;
    e.m = is.n_nonempty;
    ids.resize(e.m);
    e.ids = ids;
    e.map = std::move(map);
    e.reps = std::move(reps);
    e.bytes = sizeof(float) * (size_t)k * s->F;
    s->cache_bytes += e.bytes;
    e.pins = 1;
    e.last_use = ++s->tick;
    s->stats.ism_computed++;
    d.computed++;
    d.computed_pairs += (int64_t)e.m * (e.m - 1) / 2;
    *out = &e;
  } else {
    s->free_blocks.push_back(d_fields);
    s->cache.erase(key);
  }
  s->cv_cache.notify_all();
  return st;
}

void unpin(iso_server *s, Entry *e) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    e->pins--;
  }
  s->cv_cache.notify_all();  // an unpinned volume may free a block
}

int run_job(iso_server *s, Slot &sl, const Job &j, Done &d) {
  const iso_request &r = j.req;
  Entry *a = nullptr, *b = nullptr;
  ISO_TRY(ensure_entry(s, sl, r.var1, r.step1, &a, d));
  if (r.kind == ISO_REQ_ISM) {
    d.m = a->m;
    d.map = a->map;
    d.ids = a->ids;
    unpin(s, a);
    return ISO_OK;
  }
  int st = ensure_entry(s, sl, r.var2, r.step2, &b, d);
  if (st != ISO_OK) {
    unpin(s, a);
    return st;
  }
  const int n = (int)std::min(a->reps.size(), b->reps.size());
  const size_t fb = sizeof(float) * (size_t)s->F;
  hipStream_t strm = (hipStream_t)iso_ctx_stream(sl.ctx);
  for (int i = 0; i < n && st == ISO_OK; ++i) {
    if (hipMemcpyAsync(sl.d_cross + (size_t)i * s->F, a->d_fields + (size_t)a->reps[i] * s->F, fb,
                       hipMemcpyDeviceToDevice, strm) != hipSuccess ||
        hipMemcpyAsync(sl.d_cross + (size_t)(n + i) * s->F, b->d_fields + (size_t)b->reps[i] * s->F,
                       fb, hipMemcpyDeviceToDevice, strm) != hipSuccess)
      st = ISO_ERR_HIP;
  }
  if (st == ISO_OK && n > 0) st = iso_field_interval(sl.ctx, sl.d_cross, 2 * n, s->F, sl.d_mm, sl.d_mm + 2 * n);
  if (st == ISO_OK && n > 0)
    st = iso_hist_nmi(sl.ctx, sl.d_cross, n, s->F, s->cfg.H_multi, 1, s->cfg.flags, sl.d_mm,
                      sl.d_mm + 2 * n, sl.d_map, nullptr, nullptr, nullptr);
  if (st == ISO_OK) {
    d.m = n;
    d.map.resize((size_t)n * n);
    st = iso_copy_to_host(sl.ctx, d.map.data(), sl.d_map, sizeof(float) * d.map.size());
    for (int i = 0; i < n; ++i) d.ids.push_back(a->ids[a->reps[i]]);
    for (int i = 0; i < n; ++i) d.ids.push_back(b->ids[b->reps[i]]);
  }
  unpin(s, a);
  unpin(s, b);
  return st;
}

void worker(iso_server *s, int slot) {
  Slot &sl = s->slots[slot];
  while (true) {
    Job j;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_queue.wait(lk, [&] { return s->stop || (!s->queue.empty() && admissible_locked(s, s->queue.front())); });
      if (s->stop && s->queue.empty()) return;
      j = s->queue.front();
      s->queue.pop_front();
      hold_locked(s, j);
      s->inflight += j.estimate;
      s->busy++;
    }
    Done d;
    d.t_submit = j.t_submit;
    d.t_start = now_s();
    d.slot = slot;
    d.status = run_job(s, sl, j, d);
    d.t_done = now_s();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      for (Entry *e : j.held)
        if (e) e->pins--;
      s->inflight -= j.estimate;
      s->busy--;
      s->stats.completed++;
      s->stats.tunings_in_requests += sl.ctx->tune_measured - sl.tune_measured_seen;
      s->stats.untuned_shapes += sl.ctx->tune_nearest - sl.tune_nearest_seen;
      sl.tune_measured_seen = sl.ctx->tune_measured;
      sl.tune_nearest_seen = sl.ctx->tune_nearest;
      s->done[j.ticket] = std::move(d);
    }
    s->cv_done.notify_all();
    s->cv_queue.notify_all();
    s->cv_cache.notify_all();
  }
}

}  // namespace

extern "C" {

int iso_server_create(iso_server **out, const iso_server_config *cfg) {
  if (!out || !cfg || cfg->slots <= 0 || cfg->k <= 0 || cfg->reps <= 0 || cfg->ds <= 0)
    return ISO_ERR_INVALID_ARG;
  *out = nullptr;
  iso_server *s = new iso_server();
  s->cfg = *cfg;
  s->slots.resize(cfg->slots);
  for (auto &sl : s->slots)
    if (iso_ctx_create(&sl.ctx, cfg->device) != ISO_OK) {
      iso_server_destroy(s);
      return ISO_ERR_HIP;
    }
  *out = s;
  return ISO_OK;
}

int iso_server_add_volume(iso_server *s, int var, int step, const float *h_img, int X, int Y,
                          int Z) {
  if (!s || !h_img || X < 2 || Y < 2 || Z < 2) return ISO_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  if (!s->workers.empty()) return ISO_ERR_INVALID_ARG;  // volumes are added before serving starts
  const int64_t F = (int64_t)(X / s->cfg.ds) * (Y / s->cfg.ds) * (Z / s->cfg.ds);
  if (s->F && F != s->F) return ISO_ERR_UNSUPPORTED;  // one grid size per server
  s->F = F;
  Volume v;
  v.X = X;
  v.Y = Y;
  v.Z = Z;
  const size_t n = (size_t)X * Y * Z;
  iso_ctx *ctx = s->slots[0].ctx;
  ISO_TRY(iso_device_alloc(ctx, (void **)&v.d_img, sizeof(float) * n));
  ISO_TRY(iso_copy_to_device(ctx, v.d_img, h_img, sizeof(float) * n));
  // k isovalues evenly spaced strictly inside the volume's range
  float lo = h_img[0], hi = h_img[0];
  for (size_t i = 1; i < n; ++i) { lo = std::min(lo, h_img[i]); hi = std::max(hi, h_img[i]); }
  for (int i = 0; i < s->cfg.k; ++i)
    v.isos.push_back((float)(lo + (double)(i + 1) * ((double)hi - lo) / (s->cfg.k + 1)));
  s->volumes[vkey(var, step)] = v;
  return ISO_OK;
}

int iso_server_start(iso_server *s) {
  if (!s || s->F == 0) return ISO_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  if (!s->workers.empty()) return ISO_OK;
  const int k = s->cfg.k, R = s->cfg.reps;
  for (auto &sl : s->slots) {
    ISO_TRY(iso_device_alloc(sl.ctx, (void **)&sl.d_map, sizeof(float) * (size_t)k * k));
    ISO_TRY(iso_device_alloc(sl.ctx, (void **)&sl.d_cross, sizeof(float) * 2 * (size_t)R * s->F));
    ISO_TRY(iso_device_alloc(sl.ctx, (void **)&sl.d_mm, sizeof(float) * 4 * (size_t)R));
  }
  // the field-block pool: the budget's worth of blocks, at least the two a
  // cross map needs, at most one per volume
  {
    const size_t per = sizeof(float) * (size_t)k * s->F;
    size_t nb = std::max<size_t>(s->cfg.mem_budget / per, 2);
    nb = std::min(nb, std::max<size_t>(s->volumes.size(), 2));
    for (size_t i = 0; i < nb; ++i) {
      float *b = nullptr;
      ISO_TRY(iso_device_alloc(s->slots[0].ctx, (void **)&b, per));
      s->blocks.push_back(b);
    }
    s->free_blocks = s->blocks;
  }
  // Warm-up, results dropped (the cache stays cold): every slot runs an ISM
  // of one volume and a cross map of its fields, slot i on the i-th volume,
  // so every slot's workspace is sized before the first request (a workspace
  // that grows frees its old buffer, which waits for the whole device).  The
  // autotuner measures its launch shapes on slot 0's run only; every slot
  // then takes those choices and measures nothing more: a request whose
  // sizes round to keys not seen here takes the nearest measured choice, so
  // no request pays for measurements.
  auto warm = [&](Slot &sl, const Volume &vol) -> int {
    float *d_fields = s->blocks[0];  // free again when requests start
    std::vector<int> ids(k);
    iso_ism_stats is{};
    int st = iso_ism_volume(sl.ctx, vol.d_img, vol.X, vol.Y, vol.Z, vol.isos.data(), k,
                            s->cfg.extractor, s->cfg.ds, s->cfg.H_ism, s->cfg.df_method,
                            s->cfg.flags, sl.d_map, d_fields, ids.data(), &is);
    const int n = std::min(R, is.n_nonempty);
    hipStream_t strm = (hipStream_t)iso_ctx_stream(sl.ctx);
    const size_t fb = sizeof(float) * (size_t)s->F;
    for (int i = 0; i < n && st == ISO_OK; ++i)
      for (int h = 0; h < 2 && st == ISO_OK; ++h)
        if (hipMemcpyAsync(sl.d_cross + (size_t)(h * n + i) * s->F, d_fields + (size_t)i * s->F,
                           fb, hipMemcpyDeviceToDevice, strm) != hipSuccess)
          st = ISO_ERR_HIP;
    if (st == ISO_OK && n > 0)
      st = iso_field_interval(sl.ctx, sl.d_cross, 2 * n, s->F, sl.d_mm, sl.d_mm + 2 * n);
    if (st == ISO_OK && n > 0)
      st = iso_hist_nmi(sl.ctx, sl.d_cross, n, s->F, s->cfg.H_multi, 1, s->cfg.flags, sl.d_mm,
                        sl.d_mm + 2 * n, sl.d_map, nullptr, nullptr, nullptr);
    if (st == ISO_OK) st = iso_ctx_synchronize(sl.ctx);
    return st;
  };
  auto vit = s->volumes.begin();
  ISO_TRY(warm(s->slots[0], vit->second));
  for (size_t i = 1; i < s->slots.size(); ++i) {
    s->slots[i].ctx->tune = s->slots[0].ctx->tune;
    s->slots[i].ctx->tune_log = s->slots[0].ctx->tune_log;
  }
  for (auto &slot : s->slots) slot.ctx->tune_frozen = true;
  for (size_t i = 1; i < s->slots.size(); ++i) {
    if (std::next(vit) != s->volumes.end()) ++vit;
    ISO_TRY(warm(s->slots[i], vit->second));
  }
  // the warm-up is not a request: what it measured or looked up is not counted
  for (auto &slot : s->slots) {
    slot.tune_measured_seen = slot.ctx->tune_measured;
    slot.tune_nearest_seen = slot.ctx->tune_nearest;
  }
  for (int i = 0; i < (int)s->slots.size(); ++i) s->workers.emplace_back(worker, s, i);
  return ISO_OK;
}

int iso_server_submit(iso_server *s, const iso_request *req, uint64_t *ticket) {
  if (!s || !req || !ticket) return ISO_ERR_INVALID_ARG;
  if (req->kind != ISO_REQ_ISM && req->kind != ISO_REQ_TSM && req->kind != ISO_REQ_VSM)
    return ISO_ERR_INVALID_ARG;
  if (req->kind == ISO_REQ_TSM && req->var1 != req->var2) return ISO_ERR_INVALID_ARG;
  if (req->kind == ISO_REQ_VSM && req->step1 != req->step2) return ISO_ERR_INVALID_ARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->workers.empty()) return ISO_ERR_INVALID_ARG;  // iso_server_start first
    if (!s->volumes.count(vkey(req->var1, req->step1)) ||
        (req->kind != ISO_REQ_ISM && !s->volumes.count(vkey(req->var2, req->step2))))
      return ISO_ERR_INVALID_ARG;
    Job j{s->next_ticket++, *req, now_s(), 0};
    *ticket = j.ticket;
    s->queue.push_back(j);
    s->stats.submitted++;
  }
  s->cv_queue.notify_all();
  return ISO_OK;
}

int iso_server_wait(iso_server *s, uint64_t ticket, float *h_map, int map_cap, int *h_ids,
                    int ids_cap, iso_request_info *info) {
  if (!s) return ISO_ERR_INVALID_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->done.count(ticket) > 0; });
  Done d = std::move(s->done[ticket]);
  s->done.erase(ticket);
  lk.unlock();
  if (info) {
    info->status = d.status;
    info->m = d.m;
    info->slot = d.slot;
    info->t_submit = d.t_submit;
    info->t_start = d.t_start;
    info->t_done = d.t_done;
    info->computed = d.computed;
    info->computed_pairs = d.computed_pairs;
  }
  if (d.status != ISO_OK) return d.status;
  if (h_map) {
    if ((int64_t)map_cap < (int64_t)d.map.size()) return ISO_ERR_CAPACITY;
    memcpy(h_map, d.map.data(), sizeof(float) * d.map.size());
  }
  if (h_ids) {
    if (ids_cap < (int)d.ids.size()) return ISO_ERR_CAPACITY;
    memcpy(h_ids, d.ids.data(), sizeof(int) * d.ids.size());
  }
  return ISO_OK;
}

int iso_server_get_stats(iso_server *s, iso_server_stats *out) {
  if (!s || !out) return ISO_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  out->cache_bytes = s->cache_bytes;
  out->resident_volumes = (int)s->cache.size();
  return ISO_OK;
}

int iso_server_destroy(iso_server *s) {
  if (!s) return ISO_ERR_INVALID_ARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_queue.notify_all();
  for (auto &t : s->workers) t.join();
  for (float *b : s->blocks) (void)hipFree(b);
  for (auto &kv : s->volumes)
    if (kv.second.d_img) (void)hipFree(kv.second.d_img);
  for (auto &sl : s->slots) {
    if (sl.d_map) (void)hipFree(sl.d_map);
    if (sl.d_cross) (void)hipFree(sl.d_cross);
    if (sl.d_mm) (void)hipFree(sl.d_mm);
    if (sl.ctx) iso_ctx_destroy(sl.ctx);
  }
  delete s;
  return ISO_OK;
}

}  // extern "C"
