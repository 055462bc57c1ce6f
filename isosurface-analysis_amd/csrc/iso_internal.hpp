// Internal declarations shared by the translation units of libiso_amd.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
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
  WS_MC_KEYS_A,        // sort buffers
  WS_MC_KEYS_B,
  WS_MC_TMP,
  WS_SMALL,            // per-call argument/result block
  WS_DF_POINTS,        // normalised points
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
  void *pinned = nullptr;  // small pinned host staging area
  size_t pinned_bytes = 0;
};

namespace iso {

int fail(iso_ctx *ctx, int code, const char *fmt, ...);
int hip_fail(iso_ctx *ctx, hipError_t e, const char *what, const char *file, int line);

// Returns a device buffer of at least `bytes` for `slot`; zero_at_rest buffers
// are cleared when (re)allocated and every user restores zeros after use.
int ws_get(iso_ctx *ctx, iso_ws_slot slot, size_t bytes, bool zero_at_rest, void **out);

// Small pinned host buffer for count read-backs.
int pinned_get(iso_ctx *ctx, size_t bytes, void **out);

int enter(iso_ctx *ctx);  // hipSetDevice(ctx->device) + clears last_error

// Fills count 32-bit words with v using a kernel on the context's stream.
int fill_u32(iso_ctx *ctx, void *p, uint32_t v, size_t count);

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
