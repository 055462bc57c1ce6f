// Batched form over many volumes of uneven size: the ISMs of a whole batch.
//
// Per volume, extraction and distance fields run as in iso_ism_volume; the
// fields of every volume land in one buffer.  Then all unordered pairs of
// all ISMs go through ONE persistent histogram+NMI launch: a few workgroups
// per CU pull pair tasks (largest field first) from a device queue until it
// is empty, so volumes from 48^3 to 256^3 share the machine without tail
// effects and without a launch per ISM.
//
// Per-pair semantics are those of hist_nmi_kernel (JointHistogram.cu:24-162):
// exact integer joint counts privatised in LDS; NMI in the reference's float
// order with ISO_COMPAT_NMI32 (iso_nmi.hpp), else in double.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <vector>

#include "iso_bins.hpp"
#include "iso_df.hpp"
#include "iso_nmi.hpp"

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
}  // namespace iso

namespace {

struct BatchTask {
  int64_t fi, fj;  // element offsets of the two fields in the field buffer
  int64_t F;       // voxels per field
  int64_t processed;
  int mm_i, mm_j;  // the two fields' entries in the (mins, maxs) arrays
  int pad0, pad1;
  int map_i, map_j;  // entries (i,j) and (j,i) of the job's map
  int job, n;        // job index and its map size
};

struct BatchArgs {
  const float *fields;
  const float *mins, *maxs;  // per field of the batch (DistanceField::getInterval)
  const BatchTask *tasks;
  int ntasks;
  uint32_t *next;          // queue head, cleared before each launch
  float *const *maps;      // per job: map (row-major n x n)
  uint32_t *ghist;         // per workgroup drain area (16-bit path), zero at rest
  int H;
  int nmi32;     // reference float-order NMI (ISO_COMPAT_NMI32)
  int list_cap;  // float2 entries per chain list in the dynamic LDS (nmi32)
  float *stage;  // nmi32: a staging area of nmi32_stage_bytes(H) per workgroup
};

constexpr int kBBlock = 1024;
constexpr int kBWaves = kBBlock / 64;
constexpr int kBRoundVpt = 8;
constexpr int kBRound = kBBlock * kBRoundVpt;
constexpr int kBDepth = 1;  // rounds per drain decision (2 spills registers here)
constexpr int kBSuper = kBRound * kBDepth;
constexpr uint32_t kBLimit16 = 65535u - kBSuper;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <bool PACK16>
__global__ __launch_bounds__(kBBlock) void hist_batch_kernel(BatchArgs a) {
  extern __shared__ uint32_t s_hist[];
  __shared__ uint32_t s_rows[256], s_cols[256];
  __shared__ double s_red[3][kBWaves];
  __shared__ int s_task, s_flag[3], s_flushed;
  __shared__ iso::Nmi32Lds s_nmi;
  const int tid = threadIdx.x, H = a.H, HH = H * H;
  const int words = PACK16 ? (HH + 1) >> 1 : HH;
  uint32_t *gh = a.ghist + (int64_t)blockIdx.x * HH;
  for (int w = tid; w < words; w += kBBlock) s_hist[w] = 0u;
  while (true) {
    if (tid == 0) s_task = (int)atomicAdd(a.next, 1u);
    if (tid < 256) { s_rows[tid] = 0u; s_cols[tid] = 0u; }
    if (tid == 0) { s_flag[0] = s_flag[1] = s_flag[2] = 0; s_flushed = 0; }
    __syncthreads();
    // wave-uniform (the LDS read is not known to be): the task's pointers
    // then live in scalar registers and the buffer resources built from them
    // need no per-lane waterfall
    const int ti = __builtin_amdgcn_readfirstlane(s_task);
    if (ti >= a.ntasks) break;
    const BatchTask t = a.tasks[ti];
    const float *f1 = a.fields + t.fi, *f2 = a.fields + t.fj;
    // the pair's bins (JointHistogram.cu:127-129): union range, H bins
    iso::Binner bin;
    bin.init(fminf(a.mins[t.mm_i], a.mins[t.mm_j]), fmaxf(a.maxs[t.mm_i], a.maxs[t.mm_j]), H);
    auto drain = [&]() {
      for (int w = tid; w < words; w += kBBlock) {
        uint32_t v = s_hist[w];
        if (v) {
          if (PACK16) {
            if (v & 0xffffu) gh[2 * w] += v & 0xffffu;
            if (v >> 16) gh[2 * w + 1] += v >> 16;
          } else {
            gh[w] += v;
          }
          s_hist[w] = 0u;
        }
      }
    };
    if constexpr (PACK16) {
      // The round structure of hist_nmi_kernel's 16-bit path: flag rotation
      // over three slots, kBDepth rounds per drain decision, buffer loads
      // bounded by the pair's end, and each round's registers refilled with
      // the next super-round's values as soon as they are binned.
      const int64_t end = t.processed;
      uint32_t voff[kBRoundVpt];
#pragma unroll
      for (int u = 0; u < kBRoundVpt; ++u) voff[u] = 4u * (uint32_t)(u * kBBlock + tid);
      auto rsrc = [&](const float *f, int64_t b) {
        const int64_t left = end - b;
        const int n = left <= 0 ? 0 : (left >= kBRound ? kBRound : (int)left);
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(left > 0 ? f + b : f), 0,
                                                 4 * n, iso::kBufferWord3);
      };
      float x[kBDepth][kBRoundVpt], y[kBDepth][kBRoundVpt];
      auto load_round = [&](float (&xr)[kBRoundVpt], float (&yr)[kBRoundVpt], int64_t b) {
        const auto r1 = rsrc(f1, b), r2 = rsrc(f2, b);
#pragma unroll
        for (int u = 0; u < kBRoundVpt; ++u) {
          xr[u] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r1, voff[u], 0, 0));
          yr[u] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r2, voff[u], 0, 0));
        }
      };
#pragma unroll
      for (int r = 0; r < kBDepth; ++r) load_round(x[r], y[r], (int64_t)r * kBRound);
      auto super_round = [&](auto full_tag, int64_t base) {
        constexpr bool FULL = decltype(full_tag)::value;
        const int rem = FULL ? kBSuper : (int)(end - base);
        bool hot = false;
#pragma unroll
        for (int r = 0; r < kBDepth; ++r) {
          uint32_t e[kBRoundVpt], old[kBRoundVpt], sh[kBRoundVpt];
          iso::joint_bins(bin, x[r], y[r], e);
          load_round(x[r], y[r], base + kBSuper + (int64_t)r * kBRound);
#pragma unroll
          for (int u = 0; u < kBRoundVpt; ++u) {  // atomics back to back, returns examined after
            const int o = (r * kBRoundVpt + u) * kBBlock;
            old[u] = 0u;
            sh[u] = 0u;
            if (FULL || o + tid < rem) {
              sh[u] = (e[u] & 1u) << 4;
              old[u] = atomicAdd(&s_hist[e[u] >> 1], 1u << sh[u]);
            }
          }
#pragma unroll
          for (int u = 0; u < kBRoundVpt; ++u) hot |= ((old[u] >> sh[u]) & 0xffffu) >= kBLimit16;
        }
        return hot;
      };
      int fb = 0;
      for (int64_t base = 0; base < end; base += kBSuper) {
        const bool hot = base + kBSuper <= end
                             ? super_round(std::integral_constant<bool, true>(), base)
                             : super_round(std::integral_constant<bool, false>(), base);
        const int fn = fb == 2 ? 0 : fb + 1;
        if (hot) s_flag[fb] = 1;
        if (tid == 0) s_flag[fn] = 0;
        iso::lds_barrier();
        if (s_flag[fb]) {
          drain();  // this workgroup owns its drain area: plain adds
          if (tid == 0) s_flushed = 1;
          __syncthreads();
        }
        fb = fn;
      }
    } else {
      // 8 values per lane and field per step through bounded buffer loads,
      // the next step's loaded as soon as this step's are binned (as in
      // hist_nmi_kernel's 32-bit path)
      constexpr int kVpt = 8, kStep = kVpt * kBBlock;
      const int64_t end = t.processed;
      uint32_t voff[kVpt];
#pragma unroll
      for (int u = 0; u < kVpt; ++u) voff[u] = 4u * (uint32_t)(u * kBBlock + tid);
      float x[kVpt], y[kVpt];
      auto load = [&](int64_t b) {
        const int64_t left = end - b;
        const int n = left <= 0 ? 0 : (left >= kStep ? kStep : (int)left);
        const auto r1 = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float *>(left > 0 ? f1 + b : f1), 0, 4 * n, iso::kBufferWord3);
        const auto r2 = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float *>(left > 0 ? f2 + b : f2), 0, 4 * n, iso::kBufferWord3);
#pragma unroll
        for (int u = 0; u < kVpt; ++u) {
          x[u] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r1, voff[u], 0, 0));
          y[u] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r2, voff[u], 0, 0));
        }
      };
      load(0);
      for (int64_t b = 0; b < end; b += kStep) {
        uint32_t e[kVpt];
        iso::joint_bins(bin, x, y, e);
        load(b + kStep);
        if (b + kStep <= end) {
#pragma unroll
          for (int u = 0; u < kVpt; ++u) atomicAdd(&s_hist[e[u]], 1u);
        } else {
          const int rem = (int)(end - b);
#pragma unroll
          for (int u = 0; u < kVpt; ++u)
            if (u * kBBlock + tid < rem) atomicAdd(&s_hist[e[u]], 1u);
        }
      }
    }
    __syncthreads();
    const bool use_global = s_flushed;
    if (use_global) {
      drain();
      __syncthreads();
    }
    if (a.nmi32) {
      // the reference's float order (iso_nmi.hpp) with this workgroup's own
      // staging area; its windows overwrite the histogram's LDS, which is
      // cleared again afterwards
      // H made opaque here, so the tail's bin addresses are not hoisted out of
      // the task loop (they would stay live, and spill, across the histogram)
      int Hv = H;
      asm volatile("" : "+v"(Hv));
      auto count_at = [&](int e) -> uint32_t {
        if (use_global) return gh[e];
        return PACK16 ? (s_hist[e >> 1] >> ((e & 1) << 4)) & 0xffffu : s_hist[e];
      };
      auto take_at = [&](int e) -> uint32_t {
        if (!use_global) return PACK16 ? (s_hist[e >> 1] >> ((e & 1) << 4)) & 0xffffu : s_hist[e];
        const uint32_t c = gh[e];
        gh[e] = 0u;  // zero bins already are
        return c;
      };
      float *map = a.maps[t.job];
      // this workgroup's own staging area (the grid is persistent)
      auto claim = [&]() -> float * {
        return (float *)((char *)a.stage + (size_t)blockIdx.x * iso::nmi32_stage_bytes(H));
      };
      auto release = [&]() {};
      float *out_ij = map + (int64_t)t.map_i * t.n + t.map_j;
      float *out_ji = map + (int64_t)t.map_j * t.n + t.map_i;
      const bool mirror = t.map_i != t.map_j;
      // the counts stay in registers (16 bits each with the packed
      // histogram), unless the pair drained to the global one or an odd H
      // puts a lane's bins across three packed words
      if (PACK16 && (use_global || (H & 1)))
        iso::nmi32_pair<16, 4, false>(count_at, s_hist, take_at, true, Hv, t.F, mirror, s_rows,
                                      s_cols, claim, release, (float2 *)s_hist, a.list_cap,
                                      s_nmi, out_ij, out_ji);
      else
        iso::nmi32_pair<PACK16 ? 16 : 8, PACK16 ? 4 : 2, true>(
            count_at, s_hist, take_at, use_global, Hv, t.F, mirror, s_rows, s_cols, claim,
            release, (float2 *)s_hist, a.list_cap, s_nmi, out_ij, out_ji);
      for (int w = tid; w < words; w += kBBlock) s_hist[w] = 0u;
      __syncthreads();
      continue;
    }
    double sxy = 0.0;
    for (int e = tid; e < HH; e += kBBlock) {
      uint32_t c;
      if (use_global) {
        c = gh[e];
        gh[e] = 0u;
      } else {
        c = PACK16 ? (s_hist[e >> 1] >> ((e & 1) << 4)) & 0xffffu : s_hist[e];
      }
      if (c) {
        const double dc = (double)c;
        sxy += dc * log(dc);
        atomicAdd(&s_rows[e / H], c);
        atomicAdd(&s_cols[e % H], c);
      }
    }
    __syncthreads();
    // LDS back to zero for the next task
    for (int w = tid; w < words; w += kBBlock) s_hist[w] = 0u;
    double sx = 0.0, sy = 0.0;
    if (tid < H) {
      const uint32_t r = s_rows[tid], c = s_cols[tid];
      if (r) sx = (double)r * log((double)r);
      if (c) sy = (double)c * log((double)c);
    }
    sxy = wave_sum(sxy);
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    if ((tid & 63) == 0) { s_red[0][tid >> 6] = sxy; s_red[1][tid >> 6] = sx; s_red[2][tid >> 6] = sy; }
    __syncthreads();
    if (tid == 0) {
      double txy = 0, tx = 0, ty = 0;
      for (int q = 0; q < kBWaves; ++q) { txy += s_red[0][q]; tx += s_red[1][q]; ty += s_red[2][q]; }
      const double dN = (double)t.F, lnN = log(dN);
      const double hx = lnN - tx / dN, hy = lnN - ty / dN, hxy = lnN - txy / dN;
      const double val = 2.0 * (hx + hy - hxy) / (hx + hy);
      const float out = (val != val) ? 0.0f : (float)val;
      float *map = a.maps[t.job];
      map[(int64_t)t.map_i * t.n + t.map_j] = out;
      map[(int64_t)t.map_j * t.n + t.map_i] = out;
    }
    __syncthreads();
  }
}

}  // namespace

extern "C" {

int iso_ism_batch(iso_ctx *ctx, iso_ism_job *jobs, int njobs, int extractor, int ds, int H,
                  int df_method, uint32_t flags, iso_ism_stats *h_stats) {
  ISO_TRY(iso::enter(ctx));
  if (!jobs || njobs <= 0 || ds <= 0 || H < 1 || H > 256)
    return iso::fail(ctx, ISO_ERR_INVALID_ARG, "iso_ism_batch: bad arguments");
  if (h_stats) memset(h_stats, 0, sizeof(*h_stats));
  // the stage events, destroyed on every return (early ones included)
  struct StageEvents {
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    ~StageEvents() {
      for (auto &x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } stage_events;
  hipEvent_t *ev = stage_events.e;
  for (int q = 0; q < 3; ++q) ISO_HIP_CHECK(ctx, hipEventCreate(&ev[q]));
  ISO_HIP_CHECK(ctx, hipEventRecord(ev[0], ctx->stream));
  // ---- per volume: extraction + distance fields into one field buffer
  std::vector<int64_t> fbase(njobs), F(njobs);
  int64_t ftotal = 0;
  for (int b = 0; b < njobs; ++b) {
    iso_ism_job &jb = jobs[b];
    if (!jb.d_vol || !jb.h_isos || !jb.d_map || !jb.h_ids || jb.k <= 0 || jb.X < ds ||
        jb.Y < ds || jb.Z < ds || jb.X < 2 || jb.Y < 2 || jb.Z < 2)
      return iso::fail(ctx, ISO_ERR_INVALID_ARG, "iso_ism_batch: bad job %d", b);
    F[b] = (int64_t)(jb.X / ds) * (jb.Y / ds) * (jb.Z / ds);
    fbase[b] = ftotal;
    ftotal += F[b] * jb.k;
  }
  void *fields = nullptr, *dmm = nullptr;
  ISO_TRY(iso::ws_get(ctx, WS_PIPE_E, sizeof(float) * (size_t)ftotal, false, &fields));
  int ktotal = 0;
  for (int b = 0; b < njobs; ++b) ktotal += jobs[b].k;
  // (mins, maxs) of every field of the batch, filled on the device
  ISO_TRY(iso::ws_get(ctx, WS_PIPE_D, sizeof(float) * 2 * (size_t)ktotal, false, &dmm));
  float *d_mins = (float *)dmm, *d_maxs = d_mins + ktotal;
  const bool mc = extractor == ISO_EXTRACT_MC;
  std::vector<BatchTask> tasks;
  std::vector<float *> maps(njobs);
  int64_t points = 0;
  int mmbase = 0;  // fields of the earlier jobs
  for (int b = 0; b < njobs; ++b) {
    iso_ism_job &jb = jobs[b];
    const int k = jb.k;
    std::vector<int> order(k);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return jb.h_isos[x] < jb.h_isos[y]; });
    const size_t blk = 4 * (size_t)k * 2 + 4 * (size_t)k + 16 * (size_t)k;
    void *dblk = nullptr, *hblk = nullptr;
    ISO_TRY(iso::ws_get(ctx, WS_PIPE_A, blk + 1024, false, &dblk));
    ISO_TRY(iso::pinned_get(ctx, blk + 1024, &hblk));
    ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    float *h_iso = (float *)hblk;
    int *h_slot = (int *)(h_iso + k);
    for (int j = 0; j < k; ++j) {
      h_iso[j] = jb.h_isos[order[j]];
      h_slot[j] = order[j];
    }
    float *d_iso = (float *)dblk;
    int *d_slot = (int *)(d_iso + k);
    uint32_t *d_counts = (uint32_t *)(d_slot + k);
    ISO_HIP_CHECK(ctx, hipMemcpyAsync(dblk, hblk, 8 * (size_t)k, hipMemcpyHostToDevice, ctx->stream));
    int64_t total = 0;
    void *pool = nullptr, *pslot = nullptr;
    for (int attempt = 0; attempt < 2; ++attempt) {
      const int64_t cap = (int64_t)(ctx->ws[WS_PIPE_B].bytes / 12);
      ISO_TRY(iso::ws_get(ctx, WS_PIPE_B, 12 * (size_t)std::max<int64_t>(cap, 1), false, &pool));
      ISO_TRY(iso::ws_get(ctx, WS_PIPE_SLOTS, 4 * (size_t)std::max<int64_t>(cap, 1), false, &pslot));
      if (mc)
        ISO_TRY(iso::mc_multi(ctx, jb.d_vol, jb.X, jb.Y, jb.Z, flags, d_iso, d_slot, k, d_counts,
                              nullptr, (float *)pool, (uint32_t *)pslot, cap, &total, nullptr));
      else
        ISO_TRY(iso::approx_multi(ctx, jb.d_vol, jb.X, jb.Y, jb.Z, d_iso, d_slot, k, d_counts,
                                  (int32_t *)pool, (uint32_t *)pslot, cap, &total));
      if (total <= cap) break;
      ISO_TRY(iso::ws_get(ctx, WS_PIPE_B, 12 * (size_t)total, false, &pool));
      ISO_TRY(iso::ws_get(ctx, WS_PIPE_SLOTS, 4 * (size_t)total, false, &pslot));
    }
    std::vector<uint32_t> counts(k);
    ISO_HIP_CHECK(ctx, hipMemcpyAsync(counts.data(), d_counts, 4 * (size_t)k, hipMemcpyDeviceToHost,
                                      ctx->stream));
    ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    int m = 0;
    for (int j = 0; j < k; ++j)
      if (counts[j] > 0) jb.h_ids[m++] = j;
    jb.n_nonempty = m;
    maps[b] = jb.d_map;
    points += total;
    if (m == 0) continue;
    iso::SampleGrid g{jb.X, jb.Y, jb.Z, ds, jb.X / ds, jb.Y / ds, jb.Z / ds};
    float *fb = (float *)fields + fbase[b];
    ISO_TRY(iso::df_pool(ctx, pool, !mc, total, (const uint32_t *)pslot, k, counts.data(), g,
                         df_method, fb, false, /*coherent=*/true));
    // bin parameters come from the field intervals; the pair kernel reads
    // them on the device, so nothing is read back here
    ISO_TRY(iso::field_interval(ctx, fb, m, F[b], d_mins + mmbase, d_maxs + mmbase));
    int64_t processed = F[b];
    if ((flags & ISO_COMPAT_HIST_CHUNKS) && F[b] > 0 && F[b] % 1000000 == 0) processed -= 1000000;
    for (int i = 0; i < m; ++i)
      for (int j = i; j < m; ++j) {
        BatchTask t;
        t.fi = fbase[b] + (int64_t)i * F[b];
        t.fj = fbase[b] + (int64_t)j * F[b];
        t.F = F[b];
        t.processed = processed;
        t.mm_i = mmbase + i;
        t.mm_j = mmbase + j;
        t.pad0 = t.pad1 = 0;
        t.map_i = i;
        t.map_j = j;
        t.job = b;
        t.n = m;
        tasks.push_back(t);
      }
    mmbase += m;
  }
  ISO_HIP_CHECK(ctx, hipEventRecord(ev[1], ctx->stream));
  // ---- all pairs of all ISMs: one persistent launch, largest fields first
  std::stable_sort(tasks.begin(), tasks.end(),
                   [](const BatchTask &x, const BatchTask &y) { return x.F > y.F; });
  const int nt = (int)tasks.size();
  if (nt > 0) {
    const bool pack16 = H > 128;
    const int per_cu = 1;  // one 1024-lane workgroup per CU (registers)
    const int groups = std::min(nt, per_cu * ctx->cu_count);
    const bool nmi32 = (flags & ISO_COMPAT_NMI32) != 0;
    // the float-order NMI lays its chain lists out in the histogram's LDS
    const size_t hist_lds = 4 * (pack16 ? ((size_t)H * H + 1) / 2 : (size_t)H * H);
    const size_t lds = nmi32 ? std::max<size_t>(hist_lds, 16384) : hist_lds;
    if (pack16)
      ISO_HIP_CHECK(ctx, hipFuncSetAttribute((const void *)hist_batch_kernel<true>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, 131072));
    else
      ISO_HIP_CHECK(ctx, hipFuncSetAttribute((const void *)hist_batch_kernel<false>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, 65536));
    BatchArgs ba;
    ba.fields = (const float *)fields;
    ba.mins = d_mins;
    ba.maxs = d_maxs;
    ba.ntasks = nt;
    ba.H = H;
    ba.nmi32 = nmi32 ? 1 : 0;
    ba.list_cap = (int)(lds / 16);
    auto launch = [&]() -> int {
      ISO_TRY(iso::fill_u32(ctx, ba.next, 0u, 1));
      if (pack16)
        hipLaunchKernelGGL(hist_batch_kernel<true>, dim3(groups), dim3(kBBlock), lds, ctx->stream, ba);
      else
        hipLaunchKernelGGL(hist_batch_kernel<false>, dim3(groups), dim3(kBBlock), lds, ctx->stream, ba);
      ISO_HIP_CHECK(ctx, hipGetLastError());
      return ISO_OK;
    };
    const size_t tbytes = sizeof(BatchTask) * (size_t)nt, mbytes = sizeof(float *) * (size_t)njobs;
    const size_t hbytes = sizeof(uint32_t) * (size_t)groups * H * H;
    const size_t sbytes = nmi32 ? (size_t)groups * iso::nmi32_stage_bytes(H) : 0;
    ba.stage = nullptr;
    if (flags & ISO_BATCH_GRAPH) {
      // the graph owns one buffer: tasks | maps | queue head | drain area |
      // NMI staging; its key is the layout: the tasks, the launch shape and
      // the inputs' addresses
      uint64_t key = 1469598103934665603ull;
      auto mix = [&](const void *p, size_t n) {
        const unsigned char *c = (const unsigned char *)p;
        for (size_t i = 0; i < n; ++i) key = (key ^ c[i]) * 1099511628211ull;
      };
      mix(tasks.data(), tbytes);
      // every kernel argument the capture bakes in: d_maxs = d_mins + ktotal
      // moves with ktotal even where d_mins and the tasks stay the same
      const uint64_t shape[9] = {(uint64_t)H, (uint64_t)groups, (uint64_t)njobs,
                                 (uint64_t)(uintptr_t)fields, (uint64_t)(uintptr_t)d_mins,
                                 (uint64_t)(uintptr_t)d_maxs, (uint64_t)nt,
                                 (uint64_t)ba.nmi32, (uint64_t)lds};
      mix(shape, sizeof shape);
      iso_ctx::graph_entry *ge = nullptr;
      for (auto &e : ctx->graphs)
        if (e.key == key) ge = &e;
      const size_t o_maps = (tbytes + 255) & ~(size_t)255, o_head = o_maps + ((mbytes + 255) & ~(size_t)255);
      const size_t o_hist = o_head + 256, o_stage = o_hist + ((hbytes + 255) & ~(size_t)255);
      const size_t need = o_stage + sbytes;
      if (!ge) {
        if (ctx->graphs.size() >= 8) {  // keep a few layouts: drop the oldest
          ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
          auto &old = ctx->graphs.front();
          if (old.exec) (void)hipGraphExecDestroy(old.exec);
          if (old.buf) (void)hipFree(old.buf);
          ctx->graphs.erase(ctx->graphs.begin());
        }
        iso_ctx::graph_entry e;
        e.key = key;
        e.bytes = need;
        ISO_HIP_CHECK(ctx, hipMalloc(&e.buf, need));
        char *gb = (char *)e.buf;
        ISO_HIP_CHECK(ctx, hipMemcpyAsync(gb, tasks.data(), tbytes, hipMemcpyHostToDevice, ctx->stream));
        ISO_TRY(iso::fill_u32(ctx, gb + o_hist, 0u, hbytes / 4));  // drain area: zero at rest
        ba.tasks = (const BatchTask *)gb;
        ba.maps = (float *const *)(gb + o_maps);
        ba.next = (uint32_t *)(gb + o_head);
        ba.ghist = (uint32_t *)(gb + o_hist);
        if (nmi32) ba.stage = (float *)(gb + o_stage);
        hipGraph_t graph = nullptr;
        ISO_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const int rc = launch();
        const hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
        if (rc != ISO_OK) {
          if (graph) (void)hipGraphDestroy(graph);
          (void)hipFree(e.buf);
          return rc;
        }
        ISO_HIP_CHECK(ctx, ce);
        const hipError_t ie = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) {
          (void)hipFree(e.buf);
          ISO_HIP_CHECK(ctx, ie);
        }
        ctx->graphs.push_back(e);
        ge = &ctx->graphs.back();
      } else {
        ++ge->replays;
      }
      // the maps may be new buffers each call: refresh their addresses
      ISO_HIP_CHECK(ctx, hipMemcpyAsync((char *)ge->buf + o_maps, maps.data(), mbytes,
                                        hipMemcpyHostToDevice, ctx->stream));
      ISO_HIP_CHECK(ctx, hipGraphLaunch(ge->exec, ctx->stream));
      if (h_stats) h_stats->graph_replays = ge->replays;
    } else {
      void *dt = nullptr, *dq = nullptr, *dmaps = nullptr, *gh = nullptr;
      ISO_TRY(iso::ws_get(ctx, WS_HIST_TASKS, tbytes, false, &dt));
      ISO_TRY(iso::ws_get(ctx, WS_PIPE_G, 256, false, &dq));
      ISO_TRY(iso::ws_get(ctx, WS_PIPE_F, mbytes, false, &dmaps));
      ISO_TRY(iso::ws_get(ctx, WS_HIST_GLOBAL, hbytes, true, &gh));
      if (nmi32) {
        void *st = nullptr;
        ISO_TRY(iso::ws_get(ctx, WS_NMI_STAGE, sbytes, false, &st));
        ba.stage = (float *)st;
      }
      ISO_HIP_CHECK(ctx, hipMemcpyAsync(dt, tasks.data(), tbytes, hipMemcpyHostToDevice,
                                        ctx->stream));
      ISO_HIP_CHECK(ctx, hipMemcpyAsync(dmaps, maps.data(), mbytes, hipMemcpyHostToDevice,
                                        ctx->stream));
      ba.tasks = (const BatchTask *)dt;
      ba.next = (uint32_t *)dq;
      ba.maps = (float *const *)dmaps;
      ba.ghist = (uint32_t *)gh;
      ISO_TRY(launch());
    }
  }
  ISO_HIP_CHECK(ctx, hipEventRecord(ev[2], ctx->stream));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (h_stats) {
    float t01 = 0, t12 = 0;
    (void)hipEventElapsedTime(&t01, ev[0], ev[1]);
    (void)hipEventElapsedTime(&t12, ev[1], ev[2]);
    int ne = 0;
    for (int b = 0; b < njobs; ++b) ne += jobs[b].n_nonempty;
    h_stats->n_nonempty = ne;
    h_stats->total_points = points;
    h_stats->ms_df = t01;  // extraction + fields of all volumes
    h_stats->ms_hist = t12;
  }
  return ISO_OK;
}

}  // extern "C"
