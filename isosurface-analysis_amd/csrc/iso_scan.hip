// Device-wide exclusive scan of 32-bit counts and a stable LSD radix sort of
// (32-bit key, 32-bit value) pairs, used by compaction, the lexicographic
// vertex sort and the Morton sort of the LBVH.
//
// Radix sort, one 8-bit digit per pass:
//   1. count:   each 1024-lane workgroup ranks its 4096-key tile; per-digit
//               tile counts go to hist[digit][tile] (digit-major);
//   2. scan:    exclusive scan of hist gives every (digit, tile) its base;
//   3. scatter: the tile is re-ranked stably in LDS (per-wave 8-ballot digit
//               match, wave offsets scanned per digit) and written to
//               base + rank.  Keys are moved once per pass.
#include "iso_internal.hpp"

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

// block exclusive scan of one value per lane; returns the exclusive prefix and
// writes the block total to *total
template <int BLOCK>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *s_tmp, uint32_t *total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc = wave_incl_scan(v);
  if (lane == 63) s_tmp[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int k = 0; k < BLOCK / 64; ++k) {
      uint32_t t = s_tmp[k];
      s_tmp[k] = run;
      run += t;
    }
    s_tmp[BLOCK / 64] = run;
  }
  __syncthreads();
  uint32_t ex = inc - v + s_tmp[w];
  *total = s_tmp[BLOCK / 64];
  __syncthreads();
  return ex;
}

__global__ __launch_bounds__(kScanBlock) void scan_reduce_kernel(const uint32_t *in, int64_t n,
                                                                 uint32_t *sums) {
  __shared__ uint32_t s_tmp[kScanBlock / 64 + 1];
  int64_t base = (int64_t)blockIdx.x * kScanTile;
  uint32_t v = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + (int64_t)k * kScanBlock + threadIdx.x;
    if (i < n) v += in[i];
  }
  uint32_t tot;
  block_excl_scan<kScanBlock>(v, s_tmp, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// single workgroup: exclusive scan of m block sums in place, total to *total
__global__ __launch_bounds__(1024) void scan_sums_kernel(uint32_t *sums, int64_t m,
                                                         uint32_t *total) {
  __shared__ uint32_t s_tmp[1024 / 64 + 1];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < m; base += 1024) {
    int64_t i = base + threadIdx.x;
    uint32_t v = i < m ? sums[i] : 0u;
    uint32_t tot;
    uint32_t ex = block_excl_scan<1024>(v, s_tmp, &tot);
    if (i < m) sums[i] = ex + s_carry;
    __syncthreads();
    if (threadIdx.x == 0) s_carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0 && total) *total = s_carry;
}

__global__ __launch_bounds__(kScanBlock) void scan_down_kernel(const uint32_t *in, int64_t n,
                                                               const uint32_t *sums,
                                                               uint32_t *out) {
  __shared__ uint32_t s_tmp[kScanBlock / 64 + 1];
  int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    v[k] = i < n ? in[i] : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t ex = block_excl_scan<kScanBlock>(s, s_tmp, &tot) + sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    if (i < n) out[i] = ex;
    ex += v[k];
  }
}

// ---------------------------------------------------------------- radix sort
constexpr int kSortBlock = 1024;
constexpr int kSortWaves = kSortBlock / 64;

// For lane's digit d (8 bits), mask of lanes in the wave holding the same digit.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
  uint64_t m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    bool bit = (d >> b) & 1u;
    uint64_t bb = __ballot(bit && valid);
    m &= bit ? bb : ~bb;
  }
  return m;
}

// Tile element order: item k of the tile is at tile_base + k*kSortBlock + tid,
// processed round by round (k) so ranks are stable in global index order.
template <int ITEMS>
__global__ __launch_bounds__(kSortBlock) void radix_count_kernel(const uint32_t *keys, int64_t n,
                                                                 int shift, uint32_t *hist,
                                                                 int ntiles) {
  __shared__ uint32_t s_cnt[256];
  if (threadIdx.x < 256) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * (kSortBlock * ITEMS);
  // every key requested before the first count (a load inside the counting
  // loop would wait for its value before the next one is issued)
  uint32_t keyv[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const int64_t i = base + (int64_t)k * kSortBlock + threadIdx.x;
    keyv[k] = i < n ? keys[i] : 0u;
  }
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const int64_t i = base + (int64_t)k * kSortBlock + threadIdx.x;
    if (i < n) atomicAdd(&s_cnt[(keyv[k] >> shift) & 255u], 1u);
  }
  __syncthreads();
  if (threadIdx.x < 256) hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = s_cnt[threadIdx.x];
}

template <int ITEMS>
__global__ __launch_bounds__(kSortBlock) void radix_scatter_kernel(
    const uint32_t *keys, const uint32_t *vals, int64_t n, int shift, const uint32_t *hist,
    int ntiles, uint32_t *keys_out, uint32_t *vals_out) {
  __shared__ uint32_t s_wcnt[kSortWaves][256];  // per-wave digit counts of this round
  __shared__ uint32_t s_run[256];               // running per-digit count in the tile
  __shared__ uint32_t s_base[256];              // global base of (digit, tile)
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid < 256) {
    s_run[tid] = 0;
    s_base[tid] = hist[(int64_t)tid * ntiles + blockIdx.x];
  }
  int64_t base = (int64_t)blockIdx.x * (kSortBlock * ITEMS);
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // the whole tile's keys and values are requested up front, so one memory
  // latency is exposed per tile instead of one per round
  uint32_t keyv[ITEMS], valv[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const int64_t i = base + (int64_t)k * kSortBlock + tid;
    keyv[k] = i < n ? keys[i] : 0u;
    valv[k] = i < n ? vals[i] : 0u;
  }
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    for (int d = lane; d < 256; d += 64) s_wcnt[w][d] = 0;
    __syncthreads();
    int64_t i = base + (int64_t)k * kSortBlock + tid;
    bool valid = i < n;
    uint32_t key = keyv[k];
    uint32_t val = valv[k];
    uint32_t d = (key >> shift) & 255u;
    uint64_t m = match_digit(d, valid);
    uint32_t rank_in_wave = __popcll(m & lt);
    if (valid && rank_in_wave == 0) s_wcnt[w][d] = (uint32_t)__popcll(m);
    __syncthreads();
    // exclusive scan over waves per digit (+ running tile count)
    if (tid < 256) {
      uint32_t run = s_run[tid];
      for (int q = 0; q < kSortWaves; ++q) {
        uint32_t c = s_wcnt[q][tid];
        s_wcnt[q][tid] = run;
        run += c;
      }
      s_run[tid] = run;
    }
    __syncthreads();
    if (valid) {
      uint32_t pos = s_base[d] + s_wcnt[w][d] + rank_in_wave;
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
  }
}

__global__ void iota_kernel(uint32_t *p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}

}  // namespace

namespace iso {

// out[i] = sum(in[0..i)); *d_total (device, may be null) = sum of all.
// tmp: workspace of at least scan_tmp_words(n) words.
size_t scan_tmp_words(int64_t n) { return (size_t)((n + kScanTile - 1) / kScanTile) + 1; }

int exclusive_scan_u32(iso_ctx *ctx, const uint32_t *d_in, uint32_t *d_out, int64_t n,
                       uint32_t *d_total, uint32_t *d_tmp) {
  if (n <= 0) return ISO_OK;
  int64_t nb = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)nb), dim3(kScanBlock), 0, ctx->stream,
                     d_in, n, d_tmp);
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(1024), 0, ctx->stream, d_tmp, nb, d_total);
  hipLaunchKernelGGL(scan_down_kernel, dim3((unsigned)nb), dim3(kScanBlock), 0, ctx->stream, d_in,
                     n, (const uint32_t *)d_tmp, d_out);
  ISO_HIP_CHECK(ctx, hipGetLastError());
  return ISO_OK;
}

int iota_u32(iso_ctx *ctx, uint32_t *d, int64_t n) {
  if (n <= 0) return ISO_OK;
  hipLaunchKernelGGL(iota_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, d,
                     n);
  ISO_HIP_CHECK(ctx, hipGetLastError());
  return ISO_OK;
}

// Stable sort of (key, val) by the low `bits` bits of key.  Buffers a/b are
// ping-ponged; the sorted result ends in (*keys, *vals) which point to one of
// the two buffer pairs on return.
int radix_sort_pairs(iso_ctx *ctx, uint32_t **keys, uint32_t **vals, uint32_t *keys_alt,
                     uint32_t *vals_alt, int64_t n, int bits) {
  if (n <= 1) return ISO_OK;
  // keys per tile = 1024 lanes x items.  16 items: a quarter of the tile histograms
  // to scan and one exposed load latency per 16K keys (175M keys, 4 passes:
  // count 0.40 -> 0.34 ms and scatter 0.83 -> 0.79 ms per pass against 4
  // items).  ISO_SORT_ITEMS=4 or 8 in the environment selects the smaller tiles.
  int items = 16;
  if (const char *e = std::getenv("ISO_SORT_ITEMS")) {
    const int v = std::atoi(e);
    if (v == 4 || v == 8) items = v;
  }
  const int64_t tile = (int64_t)kSortBlock * items;
  int ntiles = (int)((n + tile - 1) / tile);
  size_t hwords = (size_t)256 * ntiles;
  void *hist = nullptr, *tmp = nullptr;
  ISO_TRY(ws_get(ctx, WS_SORT_HIST, sizeof(uint32_t) * hwords, false, &hist));
  ISO_TRY(ws_get(ctx, WS_SCAN_TMP, sizeof(uint32_t) * scan_tmp_words((int64_t)hwords), false, &tmp));
  uint32_t *ka = *keys, *va = *vals, *kb = keys_alt, *vb = vals_alt;
  for (int shift = 0; shift < bits; shift += 8) {
    auto launch = [&](auto count, auto scatter) {
      hipLaunchKernelGGL(count, dim3(ntiles), dim3(kSortBlock), 0, ctx->stream, ka, n, shift,
                         (uint32_t *)hist, ntiles);
      ISO_TRY(exclusive_scan_u32(ctx, (uint32_t *)hist, (uint32_t *)hist, (int64_t)hwords, nullptr,
                                 (uint32_t *)tmp));
      hipLaunchKernelGGL(scatter, dim3(ntiles), dim3(kSortBlock), 0, ctx->stream, ka, va, n, shift,
                         (const uint32_t *)hist, ntiles, kb, vb);
      return ISO_OK;
    };
    if (items == 16) ISO_TRY(launch(radix_count_kernel<16>, radix_scatter_kernel<16>));
    else if (items == 8) ISO_TRY(launch(radix_count_kernel<8>, radix_scatter_kernel<8>));
    else ISO_TRY(launch(radix_count_kernel<4>, radix_scatter_kernel<4>));
    ISO_HIP_CHECK(ctx, hipGetLastError());
    uint32_t *t = ka; ka = kb; kb = t;
    t = va; va = vb; vb = t;
  }
  *keys = ka;
  *vals = va;
  return ISO_OK;
}

}  // namespace iso

// ---------------------------------------------------------------- vertex sort
namespace {

__global__ void key_from_xyz_kernel(const float *xyz, const uint32_t *idx, int64_t n, int comp,
                                    uint32_t *key) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) key[i] = __float_as_uint(xyz[3 * (int64_t)idx[i] + comp]);
}

__global__ void gather_xyz_kernel(const float *xyz, const uint32_t *idx, int64_t n, float *out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    int64_t s = 3 * (int64_t)idx[i];
    out[3 * i] = xyz[s];
    out[3 * i + 1] = xyz[s + 1];
    out[3 * i + 2] = xyz[s + 2];
  }
}

__global__ void copy_f32_kernel(const float *src, float *dst, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace

namespace iso {

// Sorts n xyz triples lexicographically (x, then y, then z), in place, as the
// reference's thrust::sort over vec4f does (utils/typeOperation.h:1268-1279).
// Coordinates are non-negative, so float order equals the order of their bits;
// three stable LSD passes (z, y, x) of a 32-bit radix sort give the order.
int sort_vertices_lex(iso_ctx *ctx, float *d_xyz, int64_t n) {
  if (n <= 1) return ISO_OK;
  void *ka, *kb, *va, *vb, *tmp;
  ISO_TRY(ws_get(ctx, WS_MC_KEYS_A, 4 * (size_t)n, false, &ka));
  ISO_TRY(ws_get(ctx, WS_MC_KEYS_B, 4 * (size_t)n, false, &kb));
  ISO_TRY(ws_get(ctx, WS_DF_IDX_A, 4 * (size_t)n, false, &va));
  ISO_TRY(ws_get(ctx, WS_DF_IDX_B, 4 * (size_t)n, false, &vb));
  ISO_TRY(ws_get(ctx, WS_MC_TMP, 12 * (size_t)n, false, &tmp));
  const unsigned g = (unsigned)((n + 255) / 256);
  uint32_t *idx = (uint32_t *)va, *idx_alt = (uint32_t *)vb;
  ISO_TRY(iota_u32(ctx, idx, n));
  for (int comp = 2; comp >= 0; --comp) {
    uint32_t *keys = (uint32_t *)ka;
    hipLaunchKernelGGL(key_from_xyz_kernel, dim3(g), dim3(256), 0, ctx->stream, d_xyz, idx, n, comp,
                       keys);
    uint32_t *k0 = keys, *v0 = idx;
    uint32_t *kalt = (uint32_t *)kb, *valt = idx_alt;
    ISO_TRY(radix_sort_pairs(ctx, &k0, &v0, kalt, valt, n, 32));
    // 4 passes: the result is back in the original buffers
    idx = v0;
    idx_alt = (v0 == (uint32_t *)va) ? (uint32_t *)vb : (uint32_t *)va;
  }
  hipLaunchKernelGGL(gather_xyz_kernel, dim3(g), dim3(256), 0, ctx->stream, d_xyz, idx, n,
                     (float *)tmp);
  hipLaunchKernelGGL(copy_f32_kernel, dim3((unsigned)((3 * n + 255) / 256)), dim3(256), 0,
                     ctx->stream, (const float *)tmp, d_xyz, 3 * n);
  ISO_HIP_CHECK(ctx, hipGetLastError());
  return ISO_OK;
}

}  // namespace iso
