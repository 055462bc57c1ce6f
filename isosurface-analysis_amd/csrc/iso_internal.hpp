// Internal declarations shared by the translation units of libiso_amd.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>
#include <string>

#include "iso_amd.h"

// Workspace slots: device buffers owned by a context, grown on demand and
// reused across calls (the reference cudaMalloc's and frees per call).
enum iso_ws_slot {
  WS_HIST_GLOBAL = 0,  // per-task 32-bit joint histograms, zero at rest
  WS_HIST_ARRIVE,      // per-task slice arrival counters, zero at rest
  WS_HIST_TASKS,       // int2 task list
  WS_INTERVAL_PART,    // partial min/max per block
  WS_MC_COUNTERS,      // extraction counters
  WS_MC_CORNER_TABLE,  // hash set of grid-point vertices, EMPTY at rest
  WS_MC_BLOCKS,        // marching cubes: per-block vertex totals and their scan
  WS_MC_PRIV,          // marching cubes: privatised per-slot counters
  WS_MC_WAVES,         // marching cubes: vertices per (block, round, isovalue, wave)
  WS_MC_KEYS_A,        // sort buffers
  WS_MC_KEYS_B,
  WS_MC_TMP,
  WS_SMALL,            // per-call argument/result block
  WS_DF_CODES_A,
  WS_DF_CODES_B,
  WS_DF_IDX_A,
  WS_DF_IDX_B,
  WS_DF_NODES,
  WS_DF_PENDING,        // LBVH build: (left, right, first, last) halves per split
  WS_DF_ROOTS,          // LBVH root per point set
  WS_DF_LEAVES,
  WS_DF_FLAGS,
  WS_DF_SEG_A,
  WS_DF_SEG_B,
  WS_DF_SEG_TMP,
  WS_DF_SORTED_CODES,
  WS_DF_LEVELS,
  WS_DF_LEVEL_COUNTS,
  WS_DF_GROUPS,          // LBVH leaves: boxes, sets, deltas
  WS_DF_LEAF_FIRST,      // LBVH leaves: first point of each (+ end)
  WS_DF_BOUNDS,          // LBVH leaf boundaries: per-block counts and their scan
  WS_DF_POOL_FIRST,      // LBVH leaves taken from pool runs: first point of each, pool order
  WS_DF_OFFSETS,
  WS_PIPE_SLOTS,
  WS_SORT_HIST,
  WS_SCAN_TMP,
  WS_PIPE_A,
  WS_PIPE_B,
  WS_PIPE_C,
  WS_PIPE_D,
  WS_PIPE_E,
  WS_PIPE_F,
  WS_PIPE_G,
  WS_PIPE_H,
  WS_NMI_STAGE,        // float-order NMI: per-slot staging of the chain lists
  WS_NMI_SLOTS,        // float-order NMI: free-slot bitmap, zero at rest
  WS_COUNT
};

struct iso_ws_buf {
  void *ptr = nullptr;
  size_t bytes = 0;
  bool zero_at_rest = false;
};

struct iso_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int cu_count = 256;
  int xcd_count = 8;
  std::string last_error;
  iso_ws_buf ws[WS_COUNT];
  // small pinned host staging areas: [0] shared, [1] the distance fields'
  // set-up only (so it needs no wait for the shared one's copies)
  void *pinned[2] = {nullptr, nullptr};
  size_t pinned_bytes[2] = {0, 0};
  // autotuner: (kind, shape key) -> launch variant, and a log of its choices
  std::map<uint64_t, int> tune;
  std::string tune_log;
  // no measuring: a key not yet seen takes the choice of the nearest tuned
  // key of the same kind (a server sets it after its warm-up, so no request
  // pays for measurements)
  bool tune_frozen = false;
  // measurement sessions run, and frozen lookups that took a nearest key
  int64_t tune_measured = 0, tune_nearest = 0;
  // when set, the LBVH walk launch is bracketed by these events (the last
  // launch of a call leaves its time there)
  hipEvent_t walk_ev[2] = {nullptr, nullptr};
  // captured pair stages of iso_ism_batch (ISO_BATCH_GRAPH), one per layout;
  // each owns the device buffers its kernels read
  struct graph_entry {
    uint64_t key = 0;
    hipGraphExec_t exec = nullptr;
    void *buf = nullptr;  // tasks | maps | queue head | drain area
    size_t bytes = 0;
    uint64_t replays = 0;
  };
  std::vector<graph_entry> graphs;
};

namespace iso {

int fail(iso_ctx *ctx, int code, const char *fmt, ...);
int hip_fail(iso_ctx *ctx, hipError_t e, const char *what, const char *file, int line);

// Returns a device buffer of at least `bytes` for `slot`; zero_at_rest buffers
// are cleared when (re)allocated and every user restores zeros after use.
int ws_get(iso_ctx *ctx, iso_ws_slot slot, size_t bytes, bool zero_at_rest, void **out);

// Small pinned host buffer for count read-backs (area 0, or area 1 for the
// distance fields' set-up).  Growing an area waits for the stream.
int pinned_get(iso_ctx *ctx, size_t bytes, void **out, int area = 0);

int enter(iso_ctx *ctx);  // hipSetDevice(ctx->device) + clears last_error

// Fills count 32-bit words with v using a kernel on the context's stream.
int fill_u32(iso_ctx *ctx, void *p, uint32_t v, size_t count);

// Device buffers one launch of the histogram + NMI kernel works in: merged
// 32-bit histograms and slice arrival counters per task (zero at rest) and,
// for the reference-order NMI, staging slots (slot_words bitmap words per
// XCD, a free bit = 0) with nmi32_stage_bytes(H) per slot.  Only the low
// slots_per_word bits of a word have staging behind them (the others are
// kept set): slot w*32 + b stages at index w*slots_per_word + b.
struct HistBufs {
  uint32_t *ghist = nullptr;
  uint32_t *arrive = nullptr;
  float *stage = nullptr;
  uint32_t *slot_bits = nullptr;
  int slot_words = 0;
  int slots_per_word = 32;
};

// Looks up this context's device address of the no-alloc entry point's
// scratch array (once per device; see iso_noalloc.hip).
void noalloc_setup(iso_ctx *ctx);

// One launch over tasks d_tasks[0, nt) with the caller's buffers; variant
// 0 or 1 is the kernel shape (the autotuner's two candidates).
int hist_launch(iso_ctx *ctx, const float *d_fields, int n, int64_t F, int H, int multi,
                uint32_t flags, const float *d_mins, const float *d_maxs, float *d_map,
                const int2 *d_tasks, int nt, uint32_t *d_joint, uint32_t *d_rows,
                uint32_t *d_cols, const HistBufs &b, int variant);

inline int ceil_log2(uint64_t v) {
  int b = 0;
  while (((uint64_t)1 << b) < v && b < 63) ++b;
  return b;
}

enum iso_tune_kind { kTuneHist = 1, kTuneTraverse = 2 };

// Autotuner.  run(v) enqueues variant v (0 <= v < ncand) of one launch on the
// context's stream; every variant computes the same result.  The first time
// a (kind, key) shape class is seen, each variant runs once on the real data,
// timed with HIP events, the two fastest run once more, and the one of those
// two with the smaller total is kept for the context's lifetime;
// later calls run the kept variant only.  ISO_AUTOTUNE=0 in the environment
// keeps the default variant instead of measuring.
template <class Run>
int tune_pick(iso_ctx *ctx, int kind, uint64_t key, int ncand, int dflt, const char *what,
              Run &&run) {
  const uint64_t k = ((uint64_t)kind << 56) ^ key;
  auto it = ctx->tune.find(k);
  if (it != ctx->tune.end()) return run(it->second);
  if (ctx->tune_frozen) {
    // the same kind's tuned key nearest to this one (keys pack log2 sizes,
    // most significant first), else the default
    int v = dflt;
    uint64_t gap = ~0ull;
    for (const auto &e : ctx->tune) {
      if ((e.first >> 56) != (uint64_t)kind) continue;
      const uint64_t d = e.first > k ? e.first - k : k - e.first;
      if (d < gap) {
        gap = d;
        v = e.second;
      }
    }
    ctx->tune[k] = v;
    ctx->tune_nearest++;
    return run(v);
  }
  static const bool on = [] {
    const char *e = std::getenv("ISO_AUTOTUNE");
    return !(e && e[0] == '0');
  }();
  if (!on || ncand <= 1) {
    ctx->tune[k] = dflt;
    return run(dflt);
  }
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return run(dflt);
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return run(dflt);
  }
  int best = dflt, st = ISO_OK;
  ctx->tune_measured++;
  char buf[512];
  int pos = snprintf(buf, sizeof buf, "%s [key %llx]:", what, (unsigned long long)key);
  auto timed = [&](int v, float &ms) -> bool {
    (void)hipEventRecord(e0, ctx->stream);
    st = run(v);
    (void)hipEventRecord(e1, ctx->stream);
    if (st != ISO_OK || hipEventSynchronize(e1) != hipSuccess) return false;
    (void)hipEventElapsedTime(&ms, e0, e1);
    return true;
  };
  // one pass over every variant, then the two fastest once more (the first
  // pass also warms caches and allocations); the kept variant is the one with
  // the smaller sum of its two times
  std::vector<float> t(ncand, 3.0e38f);
  bool ok = true;
  for (int v = 0; v < ncand && ok; ++v) {
    ok = timed(v, t[v]);
    if (ok && pos < (int)sizeof buf) pos += snprintf(buf + pos, sizeof buf - pos, " %.3f", t[v]);
  }
  if (ok) {
    int a = 0, b = -1;
    for (int v = 1; v < ncand; ++v) {
      if (t[v] < t[a]) { b = a; a = v; }
      else if (b < 0 || t[v] < t[b]) b = v;
    }
    float ra = 0.0f, rb = 0.0f;
    ok = timed(a, ra) && timed(b, rb);
    if (ok) {
      best = (t[a] + ra <= t[b] + rb) ? a : b;
      if (pos < (int)sizeof buf)
        pos += snprintf(buf + pos, sizeof buf - pos, " | again %d: %.3f, %d: %.3f", a, ra, b, rb);
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != ISO_OK) return st;
  ctx->tune[k] = best;
  ctx->tune_log += std::string(buf) + " ms -> variant " + std::to_string(best) + "\n";
  return ISO_OK;
}

}  // namespace iso

#define ISO_HIP_CHECK(ctx, call)                                           \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess) return iso::hip_fail(ctx, e_, #call, __FILE__, __LINE__); \
  } while (0)

#define ISO_TRY(expr)               \
  do {                              \
    int s_ = (expr);                \
    if (s_ != ISO_OK) return s_;    \
  } while (0)

// Maps hardware workgroup id -> logical id so that consecutive logical ids run
// on the same XCD (workgroups are dispatched round-robin over the 8 XCDs;
// each XCD has its own L2).  Bijective for any grid size.
__device__ __forceinline__ uint32_t iso_xcd_remap(uint32_t b, uint32_t G, uint32_t nxcd) {
  uint32_t q = G / nxcd, r = G % nxcd;
  uint32_t x = b % nxcd, k = b / nxcd;
  uint32_t base = x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q;
  return base + k;
}
