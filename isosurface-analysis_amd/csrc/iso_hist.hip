// Joint histograms + normalised mutual information over pairs of distance
// fields, and per-field intervals.
//
// Reference semantics (isosurfaces_cuda/src/JointHistogram.cu):
//   pair (i,j): min = min(min_i,min_j), max = max(max_i,max_j),
//   step = (max-min)/H, bin(v) = (int)((v-min)/step) clamped to H-1, step==0 -> 0
//   (findBucket_d, :63-75; :127-129); hist[row*H+col], row from field i.
//   NMI = 2 (Hx+Hy-Hxy)/(Hx+Hy), NaN -> 0 (calculateMutualInformation, :24-60).
//
// MI355X design:
//   - one workgroup (1024 lanes, 16 waves) owns one unordered pair (or one
//     voxel slice of it when there are too few pairs to fill 256 CUs); the
//     pair's joint histogram is privatised in LDS (32-bit counters for
//     H <= 128: 64 KB; 16-bit counters packed two per dword for H <= 256:
//     128 KB of the 160 KB LDS);
//   - bins are computed exactly: q = (v-min) * (1/step) is within H*2^-23 of
//     the true quotient, so whenever frac(q) is at least H*2^-19 away from an
//     integer, trunc(q) equals the reference's trunc((v-min)/step); the rare
//     remaining values take the IEEE division;
//   - 16-bit counters never wrap: after every 8192-voxel round the workgroup
//     checks a flag raised by any counter above 65535-8192 and, if set,
//     drains LDS into a 32-bit global histogram;
//   - slices of one pair (and drained pairs) merge in a global 32-bit
//     histogram; the last slice to arrive (acq/rel counter) evaluates NMI
//     and restores the zeros, so the buffer is zero at rest;
//   - NMI is fused into the same kernel; only the float result leaves the
//     chip.  With ISO_COMPAT_NMI32 it is the reference's float evaluation
//     in the reference's order (iso_nmi.hpp); without it, the same quantity
//     from the exact integer counts in double (sum c ln c);
//   - workgroup ids are remapped so consecutive pairs (sharing field i) run
//     on the same XCD and hit its L2.
#include "iso_internal.hpp"
#include "iso_bins.hpp"
#include "iso_nmi.hpp"

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

namespace {

constexpr int kBlock = 1024;
constexpr int kWaves = kBlock / 64;
constexpr int kRoundVpt = 8;                     // voxels per lane per 16-bit round
constexpr int kRound = kBlock * kRoundVpt;       // voxels per round
// DEPTH (a template parameter, picked by the autotuner): with 16-bit
// counters, the rounds between drain decisions (a counter grows by at most
// DEPTH * kRound between them, which the drain threshold leaves room for);
// with 32-bit counters, the values per lane and field loaded per step.
constexpr int kMaxTasksPerLaunch = 16384;  // one launch per 128-isovalue ISM (no tail wave)

struct HistArgs {
  const float *fields;
  int64_t F;          // field size (N in the entropy normalisation)
  int64_t processed;  // voxels visited per pair (F, or fewer under Q16)
  const int2 *tasks;
  int ntasks;
  int slices;
  int64_t chunk;      // voxels per slice
  const float *mins;
  const float *maxs;
  uint32_t *ghist;    // ntasks * H*H, zero at rest
  uint32_t *arrive;   // ntasks, zero at rest
  float *map;
  int n;
  int multi;
  int H;
  uint32_t *joint;
  uint32_t *rows;
  uint32_t *cols;
  int nxcd;
  int nmi32;     // reference float-order NMI (ISO_COMPAT_NMI32)
  int list_cap;  // float2 entries per chain list in the dynamic LDS (nmi32)
  float *stage;         // nmi32: nslots staging areas of nmi32_stage_bytes(H)
  uint32_t *slot_bits;  // nmi32: free-slot bitmap, zero at rest
  int slot_words;       // nmi32: bitmap words per XCD
  int slots_per_word;   // nmi32: staging slots behind each bitmap word
};

using iso::Binner;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

using iso::lds_barrier;
using iso::kBufferWord3;

template <bool PACK16>
__device__ __forceinline__ uint32_t lds_count(const uint32_t *s, int e) {
  if (PACK16) return (s[e >> 1] >> ((e & 1) << 4)) & 0xffffu;
  return s[e];
}

template <bool PACK16, int DEPTH>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(PACK16 ? 4 : 8))) void hist_nmi_kernel(HistArgs a) {
  extern __shared__ uint32_t s_hist[];
  __shared__ uint32_t s_rows[256];
  __shared__ uint32_t s_cols[256];
  __shared__ double s_red[3][kWaves];
  __shared__ int s_flag[3], s_flushed, s_last;
  __shared__ uint32_t s_slot;
  __shared__ iso::Nmi32Lds s_nmi;

  const int tid = threadIdx.x;
  const int H = a.H;
  const int HH = H * H;
  const int words = PACK16 ? (HH + 1) >> 1 : HH;

  const uint32_t G = gridDim.x;
  const uint32_t L = iso_xcd_remap(blockIdx.x, G, (uint32_t)a.nxcd);
  const int task = (int)(L / (uint32_t)a.slices);
  const int slice = (int)(L % (uint32_t)a.slices);
  const int2 tk = a.tasks[task];
  const int fi = tk.x, fj = tk.y;

  for (int w = tid; w < words; w += kBlock) s_hist[w] = 0u;
  if (tid < 256) { s_rows[tid] = 0u; s_cols[tid] = 0u; }
  if (tid == 0) { s_flag[0] = s_flag[1] = s_flag[2] = 0; s_flushed = 0; s_last = 0; }

  Binner bin;
  {
    float mi = a.mins[fi], mj = a.mins[fj];
    float xi = a.maxs[fi], xj = a.maxs[fj];
    bin.init(mi < mj ? mi : mj, xi > xj ? xi : xj, H);
  }
  const float *f1 = a.fields + (int64_t)fi * a.F;
  const float *f2 = a.fields + (int64_t)fj * a.F;
  const int64_t begin = (int64_t)slice * a.chunk;
  int64_t end = begin + a.chunk;
  if (end > a.processed) end = a.processed;
  uint32_t *gh = a.ghist + (int64_t)task * HH;
  __syncthreads();

  auto drain = [&]() {
    // LDS -> global 32-bit histogram of this task, leaves LDS zero
    for (int w = tid; w < words; w += kBlock) {
      uint32_t v = s_hist[w];
      if (v) {
        if (PACK16) {
          uint32_t lo = v & 0xffffu, hi = v >> 16;
          if (lo) atomicAdd(&gh[2 * w], lo);
          if (hi) atomicAdd(&gh[2 * w + 1], hi);
        } else {
          atomicAdd(&gh[w], v);
        }
        s_hist[w] = 0u;
      }
    }
  };

  if constexpr (PACK16) {
    // Round r raises flag r%3; it is read after the round's barrier and
    // cleared by thread 0 in round r+2 before that round's barrier, so no
    // wave of a later round can touch it while a slow wave still reads it.
    // A super-round is DEPTH rounds of 8 values per lane, loaded during the
    // previous one: the rounds' atomics (8 back to back before any return
    // value is examined), then one LDS-only barrier for the drain decision.
    constexpr int kSuperRound = kRound * DEPTH;
    constexpr uint32_t kLimit16 = 65535u - kSuperRound;  // drain threshold
    int fb = 0;
    // Loads go through buffer resources based at one round of the slice
    // whose record count stops at the slice's end: a value past it reads as
    // 0 without a memory access, so no load needs a bound.  A round's values
    // are binned, then the same registers take the next super-round's values
    // for that round, so those loads are in flight over this super-round's
    // remaining atomics and its barrier.
    uint32_t voff[kRoundVpt];
#pragma unroll
    for (int u = 0; u < kRoundVpt; ++u) voff[u] = 4u * (uint32_t)(u * kBlock + tid);
    auto rsrc = [&](const float *f, int64_t b) {
      const int64_t left = end - b;
      const int n = left <= 0 ? 0 : (left >= kRound ? kRound : (int)left);
      return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(left > 0 ? f + b : f), 0,
                                               4 * n, kBufferWord3);
    };
    float x[DEPTH][kRoundVpt], y[DEPTH][kRoundVpt];
    auto load_round = [&](float (&xr)[kRoundVpt], float (&yr)[kRoundVpt], int64_t b) {
      const auto r1 = rsrc(f1, b), r2 = rsrc(f2, b);
#pragma unroll
      for (int u = 0; u < kRoundVpt; ++u) {
        xr[u] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r1, voff[u], 0, 0));
        yr[u] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r2, voff[u], 0, 0));
      }
    };
#pragma unroll
    for (int r = 0; r < DEPTH; ++r) load_round(x[r], y[r], begin + (int64_t)r * kRound);
    // FULL: a whole super-round inside the slice, so no per-voxel bound
    auto super_round = [&](auto full_tag, int64_t base) {
      constexpr bool FULL = decltype(full_tag)::value;
      const int rem = FULL ? kSuperRound : (int)(end - base);
      bool hot = false;
#pragma unroll
      for (int r = 0; r < DEPTH; ++r) {
        uint32_t e[kRoundVpt], old[kRoundVpt], sh[kRoundVpt];
        iso::joint_bins(bin, x[r], y[r], e);
        load_round(x[r], y[r], base + kSuperRound + (int64_t)r * kRound);
#pragma unroll
        for (int u = 0; u < kRoundVpt; ++u) {
          const int o = (r * kRoundVpt + u) * kBlock;
          old[u] = 0u;
          sh[u] = 0u;
          if (FULL || o + tid < rem) {
            sh[u] = (e[u] & 1u) << 4;
            old[u] = atomicAdd(&s_hist[e[u] >> 1], 1u << sh[u]);
          }
        }
#pragma unroll
        for (int u = 0; u < kRoundVpt; ++u) hot |= ((old[u] >> sh[u]) & 0xffffu) >= kLimit16;
      }
      return hot;
    };
    for (int64_t base = begin; base < end; base += kSuperRound) {
      const bool hot = base + kSuperRound <= end
                           ? super_round(std::integral_constant<bool, true>(), base)
                           : super_round(std::integral_constant<bool, false>(), base);
      const int fn = fb == 2 ? 0 : fb + 1;
      if (hot) s_flag[fb] = 1;
      if (tid == 0) s_flag[fn] = 0;
      lds_barrier();
      if (s_flag[fb]) {
        drain();
        if (tid == 0) s_flushed = 1;
        lds_barrier();
      }
      fb = fn;
    }
  } else {
    // DEPTH values per lane and field per step through bounded buffer loads
    // (as in the 16-bit path); the next step's values load as soon as this
    // step's are binned.  32-bit counters cannot wrap within a slice.
    constexpr int kStep = DEPTH * kBlock;
    uint32_t voff[DEPTH];
#pragma unroll
    for (int u = 0; u < DEPTH; ++u) voff[u] = 4u * (uint32_t)(u * kBlock + tid);
    float x[DEPTH], y[DEPTH];
    auto load = [&](int64_t b) {
      const int64_t left = end - b;
      const int n = left <= 0 ? 0 : (left >= kStep ? kStep : (int)left);
      const auto r1 = __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(left > 0 ? f1 + b : f1),
                                                        0, 4 * n, kBufferWord3);
      const auto r2 = __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(left > 0 ? f2 + b : f2),
                                                        0, 4 * n, kBufferWord3);
#pragma unroll
      for (int u = 0; u < DEPTH; ++u) {
        x[u] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r1, voff[u], 0, 0));
        y[u] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r2, voff[u], 0, 0));
      }
    };
    load(begin);
    for (int64_t b = begin; b < end; b += kStep) {
      uint32_t e[DEPTH];
      iso::joint_bins(bin, x, y, e);
      load(b + kStep);
      if (b + kStep <= end) {
#pragma unroll
        for (int u = 0; u < DEPTH; ++u) atomicAdd(&s_hist[e[u]], 1u);
      } else {
        const int rem = (int)(end - b);
#pragma unroll
        for (int u = 0; u < DEPTH; ++u)
          if (u * kBlock + tid < rem) atomicAdd(&s_hist[e[u]], 1u);
      }
    }
  }
  __syncthreads();

  const bool use_global = a.slices > 1 || s_flushed;
  if (use_global) {
    drain();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (a.slices > 1) {
      if (tid == 0) {
        uint32_t old = __hip_atomic_fetch_add(&a.arrive[task], 1u, __ATOMIC_ACQ_REL,
                                              __HIP_MEMORY_SCOPE_AGENT);
        s_last = (old == (uint32_t)a.slices - 1);
      }
      __syncthreads();
      if (!s_last) return;
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
  }

  const bool mirror = !a.multi && fi != fj;
  const int mj = a.multi ? fj - a.n : fj;
  uint32_t *jout = a.joint ? a.joint + ((int64_t)fi * a.n + mj) * HH : nullptr;
  uint32_t *jtr = (a.joint && mirror) ? a.joint + ((int64_t)fj * a.n + fi) * HH : nullptr;
  auto write_marginals = [&]() {
    if (tid < H && a.rows) {
      const uint32_t r = s_rows[tid], c = s_cols[tid];
      int64_t o = ((int64_t)fi * a.n + mj) * H + tid;
      a.rows[o] = r;
      a.cols[o] = c;
      if (mirror) {
        int64_t o2 = ((int64_t)fj * a.n + fi) * H + tid;
        a.rows[o2] = c;
        a.cols[o2] = r;
      }
    }
  };

  // ---- fused NMI in the reference's float order ----
  if (a.nmi32) {
    if (jout) {
      // the joint histogram out (and its transpose for the mirrored entry),
      // before the NMI tail reuses the LDS
      for (int e = tid; e < HH; e += kBlock) {
        const uint32_t c = use_global ? gh[e] : lds_count<PACK16>(s_hist, e);
        jout[e] = c;
        if (jtr) jtr[(e % H) * H + e / H] = c;
      }
    }
    auto count_at = [&](int e) -> uint32_t {
      return use_global ? gh[e] : lds_count<PACK16>(s_hist, e);
    };
    auto take_at = [&](int e) -> uint32_t {
      if (!use_global) return lds_count<PACK16>(s_hist, e);
      const uint32_t c = gh[e];
      gh[e] = 0u;  // restore zero at rest (zero bins already are)
      return c;
    };
    // a staging slot of this XCD, claimed only when the chain lists do not
    // fit in LDS
    auto claim = [&]() -> float * {
      if (tid == 0) s_slot = iso::nmi32_claim(a.slot_bits, a.slot_words, a.nxcd);
      __syncthreads();
      const uint32_t slot = s_slot;
      // only the low slots_per_word bits of each word have staging behind them
      const size_t stage_idx = (size_t)(slot >> 5) * (size_t)a.slots_per_word + (slot & 31u);
      return (float *)((char *)a.stage + stage_idx * iso::nmi32_stage_bytes(H));
    };
    auto release = [&]() {
      if (tid == 0) iso::nmi32_release(a.slot_bits, s_slot);
    };
    float *out_ij = a.map + (int64_t)fi * a.n + mj, *out_ji = a.map + (int64_t)fj * a.n + fi;
    // the counts stay in registers (16 bits each with the packed histogram),
    // unless the pair merged in the global one (where they may need 32) or
    // an odd H puts a lane's bins across three packed words
    if (PACK16 && (use_global || (H & 1)))
      iso::nmi32_pair<16, 4, false>(count_at, s_hist, take_at, true, H, a.F, mirror, s_rows,
                                    s_cols, claim, release, (float2 *)s_hist, a.list_cap, s_nmi,
                                    out_ij, out_ji);
    else if (PACK16)
      iso::nmi32_pair<16, 4, true>(count_at, s_hist, take_at, false, H, a.F, mirror, s_rows,
                                   s_cols, claim, release, (float2 *)s_hist, a.list_cap, s_nmi,
                                   out_ij, out_ji);
    else
      iso::nmi32_pair<8, 2, true>(count_at, s_hist, take_at, use_global, H, a.F, mirror, s_rows,
                                  s_cols, claim, release, (float2 *)s_hist, a.list_cap, s_nmi,
                                  out_ij, out_ji);
    write_marginals();
    if (tid == 0 && use_global && a.slices > 1) a.arrive[task] = 0u;
    return;
  }

  // ---- fused NMI from exact counts ----
  double sxy = 0.0;
  for (int e = tid; e < HH; e += kBlock) {
    uint32_t c;
    if (use_global) {
      c = gh[e];
      gh[e] = 0u;  // restore zero at rest
    } else {
      c = lds_count<PACK16>(s_hist, e);
    }
    if (c) {
      double dc = (double)c;
      sxy += dc * log(dc);
      atomicAdd(&s_rows[e / H], c);
      atomicAdd(&s_cols[e % H], c);
    }
    if (jout) jout[e] = c;
    if (jtr) jtr[(e % H) * H + e / H] = c;
  }
  __syncthreads();
  double sx = 0.0, sy = 0.0;
  if (tid < H) {
    uint32_t r = s_rows[tid], c = s_cols[tid];
    if (r) sx = (double)r * log((double)r);
    if (c) sy = (double)c * log((double)c);
  }
  write_marginals();
  sxy = wave_sum(sxy);
  sx = wave_sum(sx);
  sy = wave_sum(sy);
  const int wv = tid >> 6;
  if ((tid & 63) == 0) { s_red[0][wv] = sxy; s_red[1][wv] = sx; s_red[2][wv] = sy; }
  __syncthreads();
  if (tid == 0) {
    double txy = 0, tx = 0, ty = 0;
    for (int k = 0; k < kWaves; ++k) { txy += s_red[0][k]; tx += s_red[1][k]; ty += s_red[2][k]; }
    double dN = (double)a.F, lnN = log(dN);
    double hx = lnN - tx / dN, hy = lnN - ty / dN, hxy = lnN - txy / dN;
    double v = 2.0 * (hx + hy - hxy) / (hx + hy);
    float out = (v != v) ? 0.0f : (float)v;
    a.map[(int64_t)fi * a.n + mj] = out;
    if (mirror) a.map[(int64_t)fj * a.n + fi] = out;
    if (use_global && a.slices > 1) a.arrive[task] = 0u;
  }
}

// ---- per-field (min, max), DistanceField::getInterval semantics ----
constexpr int kIntBlock = 256;

__global__ __launch_bounds__(kIntBlock) void interval_partial_kernel(const float *fields, int64_t F,
                                                                     int parts, float2 *part) {
  const int field = blockIdx.y;
  const int p = blockIdx.x;
  const float *f = fields + (int64_t)field * F;
  int64_t chunk = (F + parts - 1) / parts;
  int64_t b = (int64_t)p * chunk, e = b + chunk;
  if (e > F) e = F;
  // min and max by these comparisons ignore NaN and, over values that are
  // never -0 (distances), do not depend on the order they are taken in
  float lo = FLT_MAX, hi = FLT_MIN;
  int64_t i = b + threadIdx.x;
  for (; i + 3 * kIntBlock < e; i += 4 * kIntBlock) {  // four loads in flight
    const float v[4] = {f[i], f[i + kIntBlock], f[i + 2 * kIntBlock], f[i + 3 * kIntBlock]};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      lo = v[u] < lo ? v[u] : lo;
      hi = v[u] > hi ? v[u] : hi;
    }
  }
  for (; i < e; i += kIntBlock) {
    const float v = f[i];
    lo = v < lo ? v : lo;
    hi = v > hi ? v : hi;
  }
  __shared__ float s_lo[kIntBlock / 64], s_hi[kIntBlock / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    float l2 = __shfl_xor(lo, o, 64), h2 = __shfl_xor(hi, o, 64);
    lo = l2 < lo ? l2 : lo;
    hi = h2 > hi ? h2 : hi;
  }
  if ((threadIdx.x & 63) == 0) { s_lo[threadIdx.x >> 6] = lo; s_hi[threadIdx.x >> 6] = hi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kIntBlock / 64; ++k) {
      lo = s_lo[k] < lo ? s_lo[k] : lo;
      hi = s_hi[k] > hi ? s_hi[k] : hi;
    }
    part[(int64_t)field * parts + p] = make_float2(lo, hi);
  }
}

__global__ void interval_final_kernel(const float2 *part, int parts, int nfields, float *mins,
                                      float *maxs) {
  int field = blockIdx.x * blockDim.x + threadIdx.x;
  if (field >= nfields) return;
  float lo = FLT_MAX, hi = FLT_MIN;
  for (int p = 0; p < parts; ++p) {
    float2 v = part[(int64_t)field * parts + p];
    lo = v.x < lo ? v.x : lo;
    hi = v.y > hi ? v.y : hi;
  }
  mins[field] = lo;
  maxs[field] = hi;
}

}  // namespace

namespace iso {

int field_interval(iso_ctx *ctx, const float *d_fields, int nfields, int64_t F, float *d_mins,
                   float *d_maxs) {
  // about 16K values per workgroup: enough workgroups to fill the chip
  int parts = (int)std::min<int64_t>(256, (F + 16383) / 16384);
  if (parts < 1) parts = 1;
  void *part = nullptr;
  ISO_TRY(ws_get(ctx, WS_INTERVAL_PART, sizeof(float2) * (size_t)parts * nfields, false, &part));
  hipLaunchKernelGGL(interval_partial_kernel, dim3(parts, nfields), dim3(kIntBlock), 0,
                     ctx->stream, d_fields, F, parts, (float2 *)part);
  ISO_HIP_CHECK(ctx, hipGetLastError());
  hipLaunchKernelGGL(interval_final_kernel, dim3((nfields + 127) / 128), dim3(128), 0, ctx->stream,
                     (const float2 *)part, parts, nfields, d_mins, d_maxs);
  ISO_HIP_CHECK(ctx, hipGetLastError());
  return ISO_OK;
}

// Launches the histogram+NMI kernel over a device task list.
int hist_nmi_tasks(iso_ctx *ctx, const float *d_fields, int n, int64_t F, int H, int multi,
                   uint32_t flags, const float *d_mins, const float *d_maxs, float *d_map,
                   const int2 *d_tasks, int ntasks, uint32_t *d_joint, uint32_t *d_rows,
                   uint32_t *d_cols) {
  if (ntasks <= 0) return ISO_OK;
  const bool pack16 = H > 128;
  const bool nmi32 = (flags & ISO_COMPAT_NMI32) != 0;
  // float-order NMI: a staging slot per workgroup that fits on the device
  // at once (one per CU with 16-bit counters, two with 32-bit ones), in
  // whole bitmap words per XCD (slots are claimed on the claimant's XCD)
  HistBufs b;
  const int nxcd = std::max(ctx->xcd_count, 1);
  const int per_xcd = ((pack16 ? 1 : 2) * ctx->cu_count + nxcd - 1) / nxcd;
  b.slot_words = (per_xcd + 31) / 32;
  const size_t nslots = (size_t)b.slot_words * 32 * nxcd;
  if (nmi32) {
    void *stage = nullptr, *bits = nullptr;
    ISO_TRY(ws_get(ctx, WS_NMI_STAGE, nslots * nmi32_stage_bytes(H), false, &stage));
    ISO_TRY(ws_get(ctx, WS_NMI_SLOTS, 4 * (size_t)b.slot_words * nxcd, true, &bits));
    b.stage = (float *)stage;
    b.slot_bits = (uint32_t *)bits;
  }
  int64_t processed = F;
  if (flags & ISO_COMPAT_HIST_CHUNKS) {
    const int64_t C = 1000000;
    if (F > 0 && F % C == 0) processed = F - C;
  }
  auto run = [&](int variant) -> int {
    for (int t0 = 0; t0 < ntasks; t0 += kMaxTasksPerLaunch) {
      int nt = ntasks - t0 < kMaxTasksPerLaunch ? ntasks - t0 : kMaxTasksPerLaunch;
      void *gh = nullptr, *arr = nullptr;
      ISO_TRY(ws_get(ctx, WS_HIST_GLOBAL, sizeof(uint32_t) * (size_t)nt * H * H, true, &gh));
      ISO_TRY(ws_get(ctx, WS_HIST_ARRIVE, sizeof(uint32_t) * (size_t)nt, true, &arr));
      b.ghist = (uint32_t *)gh;
      b.arrive = (uint32_t *)arr;
      ISO_TRY(hist_launch(ctx, d_fields, n, F, H, multi, flags, d_mins, d_maxs, d_map,
                          d_tasks + t0, nt, d_joint, d_rows, d_cols, b, variant));
    }
    return ISO_OK;
  };
  const uint64_t key = ((uint64_t)H << 32) | ((uint64_t)ceil_log2((uint64_t)processed) << 16) |
                       (uint64_t)ceil_log2((uint64_t)ntasks);
  return tune_pick(ctx, kTuneHist, key, 2, 1, pack16 ? "hist16 rounds/sync {1,2}" : "hist32 loads/step {4,8}", run);
}

// One launch of the histogram + NMI kernel over nt tasks with the caller's
// buffers (ghist: nt * H * H words and arrive: nt words, both zero at rest;
// for ISO_COMPAT_NMI32 the staging slots and their bitmap).  Launch variant
// 0 or 1 (the autotuner's choices).
int hist_launch(iso_ctx *ctx, const float *d_fields, int n, int64_t F, int H, int multi,
                uint32_t flags, const float *d_mins, const float *d_maxs, float *d_map,
                const int2 *d_tasks, int nt, uint32_t *d_joint, uint32_t *d_rows,
                uint32_t *d_cols, const HistBufs &b, int variant) {
  if (nt <= 0) return ISO_OK;
  const bool pack16 = H > 128;
  const bool nmi32 = (flags & ISO_COMPAT_NMI32) != 0;
  const size_t words = pack16 ? ((size_t)H * H + 1) / 2 : (size_t)H * H;
  // the float-order NMI reads its chain lists back through the histogram's
  // LDS: at least 16 KB, windows of 1024 entries per list
  const size_t lds = nmi32 ? std::max<size_t>(words * 4, 16384) : words * 4;
  int64_t processed = F;
  if (flags & ISO_COMPAT_HIST_CHUNKS) {
    const int64_t C = 1000000;
    if (F > 0 && F % C == 0) processed = F - C;
  }
  // variants: 16-bit counters with 1 or 2 rounds per drain decision,
  // 32-bit counters with 4 or 8 values per step (the autotuner picks; with
  // the prefetch, 16 would take 96 registers and one workgroup per CU).
  // The LDS limit is raised once per process, whatever thread gets here.
  static std::once_flag attr_once;
  static hipError_t attr_err = hipSuccess;
  std::call_once(attr_once, [] {
    const void *kern[4] = {(const void *)hist_nmi_kernel<false, 4>,
                           (const void *)hist_nmi_kernel<false, 8>,
                           (const void *)hist_nmi_kernel<true, 1>,
                           (const void *)hist_nmi_kernel<true, 2>};
    for (int v = 0; v < 4 && attr_err == hipSuccess; ++v)
      attr_err = hipFuncSetAttribute(kern[v], hipFuncAttributeMaxDynamicSharedMemorySize,
                                     v < 2 ? 65536 : 131072);
  });
  ISO_HIP_CHECK(ctx, attr_err);
  // enough workgroups to fill the chip, but at least 16K voxels per slice
  int want = (2 * ctx->cu_count + nt - 1) / nt;
  int64_t max_slices = processed / 16384;
  if (max_slices < 1) max_slices = 1;
  int slices = want < max_slices ? want : (int)max_slices;
  if (slices < 1) slices = 1;
  const int64_t chunk = (processed + slices - 1) / slices;
  HistArgs a;
  a.fields = d_fields;
  a.F = F;
  a.processed = processed;
  a.tasks = d_tasks;
  a.ntasks = nt;
  a.slices = slices;
  a.chunk = chunk;
  a.mins = d_mins;
  a.maxs = d_maxs;
  a.ghist = b.ghist;
  a.arrive = b.arrive;
  a.map = d_map;
  a.n = n;
  a.multi = multi;
  a.H = H;
  a.joint = d_joint;
  a.rows = d_rows;
  a.cols = d_cols;
  a.nxcd = ctx->xcd_count;
  a.nmi32 = nmi32 ? 1 : 0;
  a.list_cap = (int)(lds / 16);
  a.stage = b.stage;
  a.slot_bits = b.slot_bits;
  a.slot_words = b.slot_words;
  a.slots_per_word = b.slots_per_word;
  dim3 grid((unsigned)(nt * slices));
  if (pack16) {
    if (variant == 0)
      hipLaunchKernelGGL((hist_nmi_kernel<true, 1>), grid, dim3(kBlock), lds, ctx->stream, a);
    else
      hipLaunchKernelGGL((hist_nmi_kernel<true, 2>), grid, dim3(kBlock), lds, ctx->stream, a);
  } else {
    if (variant == 0)
      hipLaunchKernelGGL((hist_nmi_kernel<false, 4>), grid, dim3(kBlock), lds, ctx->stream, a);
    else
      hipLaunchKernelGGL((hist_nmi_kernel<false, 8>), grid, dim3(kBlock), lds, ctx->stream, a);
  }
  ISO_HIP_CHECK(ctx, hipGetLastError());
  return ISO_OK;
}

// Unordered pairs (i <= j) row by row, or all n*n (i, n+j) for multi.
void build_tasks(int n, int multi, std::vector<int2> &t) {
  t.clear();
  if (multi) {
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) t.push_back(make_int2(i, n + j));
  } else {
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j) t.push_back(make_int2(i, j));
  }
}

}  // namespace iso

extern "C" {

int iso_field_interval(iso_ctx *ctx, const float *d_fields, int nfields, int64_t F,
                       float *d_mins, float *d_maxs) {
  ISO_TRY(iso::enter(ctx));
  if (!d_fields || !d_mins || !d_maxs || nfields <= 0 || F <= 0)
    return iso::fail(ctx, ISO_ERR_INVALID_ARG, "iso_field_interval: bad arguments");
  return iso::field_interval(ctx, d_fields, nfields, F, d_mins, d_maxs);
}

int iso_hist_nmi(iso_ctx *ctx, const float *d_fields, int n, int64_t F, int H, int multi,
                 uint32_t flags, const float *d_mins, const float *d_maxs, float *d_map,
                 uint32_t *d_joint, uint32_t *d_rows, uint32_t *d_cols) {
  ISO_TRY(iso::enter(ctx));
  if (!d_fields || !d_mins || !d_maxs || !d_map || n <= 0 || F <= 0)
    return iso::fail(ctx, ISO_ERR_INVALID_ARG, "iso_hist_nmi: bad arguments");
  if (H < 1 || H > 256)
    return iso::fail(ctx, ISO_ERR_UNSUPPORTED, "iso_hist_nmi: H=%d outside [1,256]", H);
  if ((d_rows == nullptr) != (d_cols == nullptr))
    return iso::fail(ctx, ISO_ERR_INVALID_ARG, "iso_hist_nmi: rows and cols go together");
  std::vector<int2> tasks;
  iso::build_tasks(n, multi, tasks);
  void *dt = nullptr, *pin = nullptr;
  size_t tb = sizeof(int2) * tasks.size();
  ISO_TRY(iso::ws_get(ctx, WS_HIST_TASKS, tb, false, &dt));
  ISO_TRY(iso::pinned_get(ctx, tb, &pin));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));  // pinned buffer reuse
  memcpy(pin, tasks.data(), tb);
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(dt, pin, tb, hipMemcpyHostToDevice, ctx->stream));
  ISO_TRY(iso::hist_nmi_tasks(ctx, d_fields, n, F, H, multi, flags, d_mins, d_maxs, d_map,
                              (const int2 *)dt, (int)tasks.size(), d_joint, d_rows, d_cols));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return ISO_OK;
}

}  // extern "C"
