// Distance fields: for every sample s of the (X/ds, Y/ds, Z/ds) grid,
// sqrtf(min over points p of d2(s, p)) (d2 itself for a single point).
//
// Reference semantics (isosurfaces_cuda/funcs.cpp:209-233, src/BVH.cu:393-613):
//   points are normalised p/max(X,Y,Z); sample (x,y,z) is
//   ((x*ds+ds/2)/X, (y*ds+ds/2)/Y, (z*ds+ds/2)/Z); the LBVH traversal prunes
//   with a point-box lower bound, so it returns the exact minimum.
//
// Squared distance, identical in every path of this library:
//   d2 = fmaf(dz, dz, fmaf(dy, dy, dx*dx)),  dx = s.x - p.x.
// The point-box bound uses the same expression on per-axis gaps
//   s - clamp(s, bmin, bmax), whose magnitudes are <= |s - p| for every p in
//   the box (rounding is monotone), so bound <= d2(p) and pruning with
//   `bound < best` never discards a closer point: LBVH == brute force, bit for
//   bit.
//
// Every kernel works on many point sets at once (the isosurfaces of one ISM,
// one pool of points tagged with their set), so an ISM's k trees are built
// and traversed in a handful of launches.
//
// MI355X design:
//   - brute force: a lane holds 8 samples in registers and the points are
//     staged through LDS 256 at a time and broadcast to the wave, 7 VALU ops
//     per (sample, point); point ranges split over workgroups merge with an
//     integer atomicMin on the float bits (valid for non-negative floats);
//   - LBVH: 30-bit Morton codes (the reference's quantisation), a stable LDS
//     radix sort by code then by segment (one sort of (segment, code) keys
//     for up to 256 segments); the leaves are groups of 16 or 32 consecutive
//     sorted points with their boxes, and per segment the binary radix tree
//     over the groups' keys (first code, group index) (height <= 64, each
//     node covers a contiguous point range) is built bottom-up together with
//     its boxes inside chunks of 1024 leaves in LDS, then the few
//     chunk-crossing nodes by one workgroup, one tree level per barrier;
//   - traversal is wave-cooperative: a wave walks the tree once per 4x4x4
//     brick of samples with a wave-uniform stack (64 entries in LDS; the
//     height bound means it cannot overflow); a child is entered if any
//     lane's bound beats that lane's best; a leaf group, or a subtree of at
//     most `bucket` points, is scanned from one coalesced load and broadcast
//     to the wave through LDS.
#include "iso_df.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace iso {
int radix_sort_pairs(iso_ctx *ctx, uint32_t **keys, uint32_t **vals, uint32_t *keys_alt,
                     uint32_t *vals_alt, int64_t n, int bits);
int exclusive_scan_u32(iso_ctx *ctx, const uint32_t *d_in, uint32_t *d_out, int64_t n,
                       uint32_t *d_total, uint32_t *d_tmp);
size_t scan_tmp_words(int64_t n);
}  // namespace iso

namespace {


__device__ __forceinline__ float d2_point(float sx, float sy, float sz, float px, float py,
                                          float pz) {
  float dx = sx - px, dy = sy - py, dz = sz - pz;
  return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// Squared distance to a box, in two steps: the x and y part, shared by the
// samples of a lane (same x and y, different z), then the z part.  The gap
// per axis is s - clamp(s, lo, hi) (one v_med3_f32 and one subtraction); its
// square equals that of max(lo - s, s - hi, 0), because RN(s - lo) = -RN(lo - s)
__device__ __forceinline__ float d2_box_xy(float sx, float sy, const float *lo, const float *hi) {
  const float dx = sx - __builtin_amdgcn_fmed3f(sx, lo[0], hi[0]);
  const float dy = sy - __builtin_amdgcn_fmed3f(sy, lo[1], hi[1]);
  return __builtin_fmaf(dy, dy, dx * dx);
}
__device__ __forceinline__ float d2_box_z(float sz, const float *lo, const float *hi, float xy) {
  const float dz = sz - __builtin_amdgcn_fmed3f(sz, lo[2], hi[2]);
  return __builtin_fmaf(dz, dz, xy);
}

__device__ __forceinline__ void sample_coords(const iso::SampleGrid &g, int x, int y, int z,
                                              float &sx, float &sy, float &sz) {
  sx = (float)(x * g.ds + g.ds / 2) / (float)g.X;
  sy = (float)(y * g.ds + g.ds / 2) / (float)g.Y;
  sz = (float)(z * g.ds + g.ds / 2) / (float)g.Z;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// number of set bits of m in lanes below this one
__device__ __forceinline__ int lane_rank(uint64_t m) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// ---- point preparation ------------------------------------------------------
__device__ __forceinline__ uint32_t spread10(float c) {
  c = fminf(fmaxf(c * 1024.0f, 0.0f), 1023.0f);
  uint32_t v = (uint32_t)c;
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

// A point as the prepared arrays hold it: 12 bytes, no padding (a leaf's
// points are one contiguous run of 12-byte records; a lane's point is one
// three-dword load).
struct Pt {
  float x, y, z;
};
static_assert(sizeof(Pt) == 12, "points are 12-byte records");

// Point i normalised: p / dm per coordinate (funcs.cpp:151-175; int cells
// go through float first).  Recomputed where needed instead of stored: a
// division is cheaper than a 12-byte round trip through HBM.
template <typename T>
__device__ __forceinline__ Pt normalized(const T *xyz, int64_t i, float dm) {
  return Pt{(float)xyz[3 * i] / dm, (float)xyz[3 * i + 1] / dm, (float)xyz[3 * i + 2] / dm};
}

// Sort keys of the normalised points.  keyed == false: the 30-bit Morton
// code (morton3D_d, BVH.cu:95-106).  keyed == true: one 32-bit key, the set
// in the top 8 bits above the top 24 bits of the code, so a single sort both
// groups the sets and orders each set along the curve (sets < 256).
template <typename T>
__global__ void key_kernel(const T *xyz, const uint32_t *seg, bool keyed, int64_t n, float dm,
                           uint32_t *codes, uint32_t *idx) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Pt p = normalized(xyz, i, dm);
  const uint32_t c = spread10(p.x) * 4 + spread10(p.y) * 2 + spread10(p.z);
  codes[i] = keyed ? ((seg ? seg[i] << 24 : 0u) | (c >> 6)) : c;
  idx[i] = (uint32_t)i;
}

// normalised points in sorted order; the set of each from its key's top 8 bits
template <typename T>
__global__ void gather_keyed_kernel(const T *xyz, float dm, const uint32_t *idx,
                                    const uint32_t *keys, int64_t n, Pt *pts_out,
                                    uint32_t *seg_out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  pts_out[i] = normalized(xyz, idx[i], dm);
  seg_out[i] = keys[i] >> 24;
}

// normalised points in final order and their Morton codes
template <typename T>
__global__ void gather_kernel(const T *xyz, float dm, const uint32_t *idx, int64_t n,
                              Pt *pts_out, uint32_t *codes_out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Pt p = normalized(xyz, idx[i], dm);
  pts_out[i] = p;
  codes_out[i] = spread10(p.x) * 4 + spread10(p.y) * 2 + spread10(p.z);
}

__global__ void gather_u32_kernel(const uint32_t *src, const uint32_t *idx, int64_t n,
                                  uint32_t *out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = src[idx[i]];
}

// ---- brute force ----------------------------------------------------------
constexpr int kBfBlock = 256;
constexpr int kBfSamples = 8;

// Points of the workgroup's chunk are staged through LDS 256 at a time
// (one coalesced load per lane) and read back at a uniform address, which
// broadcasts each point to the wave: one LDS read serves 8 samples x 7 VALU.
__global__ __launch_bounds__(kBfBlock) void brute_kernel(iso::SampleGrid g, int64_t F,
                                                         const Pt *__restrict__ pts,
                                                         const int64_t *__restrict__ off,
                                                         int nchunks, uint32_t *best_bits) {
  __shared__ float4 s_pts[kBfBlock];
  const int seg = blockIdx.z;
  const int64_t a = off[seg], b = off[seg + 1];
  const int64_t n = b - a;
  const int64_t pchunk = (n + nchunks - 1) / nchunks;
  const int64_t p0 = a + (int64_t)blockIdx.y * pchunk;
  int64_t p1 = p0 + pchunk;
  if (p1 > b) p1 = b;
  if (p0 >= p1) return;  // the whole workgroup leaves together
  const int64_t s0 = ((int64_t)blockIdx.x * kBfBlock + threadIdx.x) * kBfSamples;
  float sx[kBfSamples], sy[kBfSamples], sz[kBfSamples], best[kBfSamples];
#pragma unroll
  for (int q = 0; q < kBfSamples; ++q) {
    int64_t idx = s0 + q < F ? s0 + q : F - 1;
    int z = (int)(idx % g.nz), y = (int)((idx / g.nz) % g.ny), x = (int)(idx / ((int64_t)g.nz * g.ny));
    sample_coords(g, x, y, z, sx[q], sy[q], sz[q]);
    best[q] = INFINITY;
  }
  for (int64_t c = p0; c < p1; c += kBfBlock) {
    const int cnt = (int)(p1 - c < kBfBlock ? p1 - c : kBfBlock);
    __syncthreads();  // the previous chunk's reads are done
    if ((int)threadIdx.x < cnt) {
      const Pt p = pts[c + threadIdx.x];
      s_pts[threadIdx.x] = make_float4(p.x, p.y, p.z, 0.0f);
    }
    __syncthreads();
    int i = 0;
    for (; i + 3 < cnt; i += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float4 P = s_pts[i + u];
#pragma unroll
        for (int q = 0; q < kBfSamples; ++q)  // d2 is never NaN: fminf == the smaller one
          best[q] = fminf(best[q], d2_point(sx[q], sy[q], sz[q], P.x, P.y, P.z));
      }
    }
    for (; i < cnt; ++i) {
      const float4 P = s_pts[i];
#pragma unroll
      for (int q = 0; q < kBfSamples; ++q)
        best[q] = fminf(best[q], d2_point(sx[q], sy[q], sz[q], P.x, P.y, P.z));
    }
  }
  uint32_t *out = best_bits + (int64_t)seg * F;
#pragma unroll
  for (int q = 0; q < kBfSamples; ++q)
    if (s0 + q < F) atomicMin(&out[s0 + q], __float_as_uint(best[q]));
}

__global__ void finish_kernel(uint32_t *bits, int64_t F, const int64_t *off, int S, float *df) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= F * S) return;
  int seg = (int)(i / F);
  float d = __uint_as_float(bits[i]);
  df[i] = (off[seg + 1] - off[seg] == 1) ? d : sqrtf(d);
}


// the next run of points starts from the square of this run's root
// (DistanceField.cu:54-56: min = distances[idx]; min = min * min)
__global__ void square_into_bits_kernel(const float *df, int64_t F, uint32_t *bits) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < F) bits[i] = __float_as_uint(df[i] * df[i]);
}

__global__ void root_of_bits_kernel(const uint32_t *bits, int64_t F, float *df) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < F) df[i] = sqrtf(__uint_as_float(bits[i]));
}

// ---- LBVH -----------------------------------------------------------------
// The leaves of the tree are the maximal subtrees of at most B points (B in
// 16..64) of the binary radix tree over the sorted point keys (Morton code,
// index): a leaf is a Morton cell's points, so it never straddles a jump of
// the curve.  Only the tree above them is built: B-fold fewer nodes than a
// tree over points, so one set's tree (tens of thousands of 64-byte records)
// stays in L2 while the samples of that set walk it.
//
// Split x (between sorted points x and x+1) is the root of the point-tree
// node [l, r] with l-1 and r the nearest splits on either side with a
// smaller delta; the node has r - l + 1 = r - (l-1) points.  x separates two
// leaves iff that is more than B (a set boundary has delta -1: always).
// Leaves are the runs between separating splits.  The tree over leaves takes
// the point-level delta of the split after each leaf, so it is exactly the
// point tree with every leaf subtree collapsed.
constexpr int kMaxLeaf = 64;
constexpr int kBndBlock = 1024;
constexpr int kBndHalo = kMaxLeaf + 1;
constexpr int kBndSub = 16;  // 1024-point sub-blocks per flat_bound / leaf_emit block

struct Node {
  float lmin[3], lmax[3];
  float rmin[3], rmax[3];
  // >= 0: internal node.  < 0: leaf; -(x + 1) where x is the first point of
  // the right child (for a left leaf: the end of its points, which start at
  // `first`; for a right leaf: its first point, its points end at `last`)
  int left, right;
  int first, last;  // sorted point range covered (global)
};
static_assert(sizeof(Node) == 64, "node record is one 64-byte line");

// Split strength between sorted points x and x+1: the common-prefix length of
// the 64-bit keys (Morton code, index); -1 across a set boundary or past the
// end.  Keys are unique, so the binary radix tree over them is unique and at
// most 64 levels high.
__device__ __forceinline__ int split_delta(const uint32_t *codes, const uint32_t *seg, int64_t ntot,
                                           int64_t x) {
  if (x < 0 || x + 1 >= ntot) return -1;
  if (seg[x] != seg[x + 1]) return -1;
  const uint32_t a = codes[x], b = codes[x + 1];
  if (a != b) return __clz(a ^ b);
  return 32 + __clz((uint32_t)x ^ (uint32_t)(x + 1));
}

// The point-tree node of split x holds exactly the points whose keys share
// the first delta(x) bits with key x, a contiguous run around x.  Its ends are
// found by two branch-free binary searches over a window of B points on each
// side; the node has more than B points iff the run reaches past either
// search window or its length exceeds B.
//
// Each wave writes its 64 separator flags as one bit mask and each block its
// count; after a scan of the counts, leaf_emit_kernel turns the masks into
// leaf_first.  Split ntot-1 (past the last point) always separates.
__global__ __launch_bounds__(kBndBlock) void leaf_bound_kernel(const uint32_t *codes,
                                                               const uint32_t *seg, int64_t ntot,
                                                               int B, uint32_t *block_cnt,
                                                               uint64_t *sepmask) {
  constexpr int kW = kBndBlock + 2 * kBndHalo;
  __shared__ uint32_t s_code[kW], s_seg[kW];
  __shared__ uint32_t s_wave[kBndBlock / 64];
  const int64_t x0 = (int64_t)blockIdx.x * kBndBlock;
  for (int i = threadIdx.x; i < kW; i += kBndBlock) {
    const int64_t y = x0 - kBndHalo + i;
    const bool in = y >= 0 && y < ntot;
    s_code[i] = in ? codes[y] : 0u;
    s_seg[i] = in ? seg[y] : 0xffffffffu;  // outside: a set of its own
  }
  __syncthreads();
  const int64_t x = x0 + threadIdx.x;
  bool sep = false;
  if (x < ntot) {
    const int c = kBndHalo + threadIdx.x;
    const uint32_t sx = s_seg[c];
    const uint64_t kx = ((uint64_t)s_code[c] << 32) | (uint32_t)x;
    if (x + 1 >= ntot || s_seg[c + 1] != sx) {
      sep = true;  // set boundary or past the end
    } else {
      const uint64_t k1 = ((uint64_t)s_code[c + 1] << 32) | (uint32_t)(x + 1);
      const int d = __clzll((long long)(kx ^ k1));  // 0..63
      // in(y): point y shares the first d key bits with x (d == 0: same set)
      auto in = [&](int i) {
        const uint64_t ky = ((uint64_t)s_code[i] << 32) | (uint32_t)(x - c + i);
        return s_seg[i] == sx && (d == 0 || ((ky ^ kx) >> (64 - d)) == 0);
      };
      if (in(c - B) || in(c + B)) {
        sep = true;
      } else {
        int lo = c - B, hi = c;  // in(lo) false, in(hi) true
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (in(mid)) hi = mid;
          else lo = mid;
        }
        const int l = hi;
        lo = c + 1;  // in(c + 1) true (it shares d bits), in(c + B) false
        hi = c + B;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (in(mid)) lo = mid;
          else hi = mid;
        }
        sep = lo - l + 1 > B;
      }
    }
  }
  const uint64_t m = __ballot(sep);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    s_wave[w] = (uint32_t)__popcll(m);
    sepmask[(x0 >> 6) + w] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int i = 0; i < kBndBlock / 64; ++i) t += s_wave[i];
    block_cnt[blockIdx.x] = t;
  }
}

// Separating split x writes x + 1 (the first point of the next leaf) at its
// rank: leaf_first[1 + rank].  The last entry written is ntot, and the number
// of leaves is the total count.  A block covers kBndSub 1024-point
// sub-blocks (block_base holds one entry per sub-block).
__global__ __launch_bounds__(kBndBlock) void leaf_emit_kernel(int64_t ntot, int64_t nbb,
                                                              const uint64_t *sepmask,
                                                              const uint32_t *block_base,
                                                              int *leaf_first) {
  constexpr int kWaves = kBndBlock / 64;
  static_assert(kBndSub * kWaves <= kBndBlock, "one thread per (sub-block, wave)");
  __shared__ uint32_t s_cnt[kBndSub][kWaves];
  __shared__ uint32_t s_before[kBndSub][kWaves];  // leaves before each wave's first point
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint64_t m[kBndSub];
  // every mask load issued before the first count is used
#pragma unroll
  for (int k = 0; k < kBndSub; ++k) {
    const int64_t sb = (int64_t)blockIdx.x * kBndSub + k;
    m[k] = sb < nbb ? sepmask[sb * kWaves + w] : 0ull;
  }
#pragma unroll
  for (int k = 0; k < kBndSub; ++k)
    if (lane == 0) s_cnt[k][w] = (uint32_t)__popcll(m[k]);
  __syncthreads();
  if (threadIdx.x < kBndSub * kWaves) {
    const int k = threadIdx.x / kWaves, v = threadIdx.x % kWaves;
    const int64_t sb = (int64_t)blockIdx.x * kBndSub + k;
    uint32_t before = sb < nbb ? block_base[sb] : 0u;
    for (int i = 0; i < v; ++i) before += s_cnt[k][i];
    s_before[k][v] = before;
  }
  __syncthreads();
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int k = 0; k < kBndSub; ++k) {
    const int64_t sb = (int64_t)blockIdx.x * kBndSub + k;
    if (sb >= nbb) break;
    const int64_t x = sb * kBndBlock + threadIdx.x;
    const uint32_t before = s_before[k][w];
    if (x < ntot && ((m[k] >> lane) & 1ull))
      leaf_first[1 + before + __popcll(m[k] & lt)] = (int)(x + 1);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) leaf_first[0] = 0;
}

// One lane per leaf: its box (min/max over its at most 64 contiguous points),
// its set, and the delta of the split after it (-1 after a set's last leaf).
constexpr int kBoxLanes = 8, kBoxPer = kMaxLeaf / kBoxLanes;

// kBoxLanes lanes per leaf (at most kMaxLeaf points): lane `sub` takes points
// sub, sub + kBoxLanes, ..., all requested before the first is used (a slot
// past the leaf reloads its first point), so a load instruction reads a few
// contiguous runs instead of one point of each of 64 leaves.  min and max are
// exact, so the box does not depend on the order.
__global__ __launch_bounds__(256) void leaf_box_kernel(const Pt *__restrict__ pts,
                                                       const uint32_t *__restrict__ codes,
                                                       const uint32_t *__restrict__ seg,
                                                       int64_t ntot,
                                                       const int *__restrict__ leaf_first,
                                                       int64_t G, float4 *gbox, uint32_t *gseg,
                                                       int *gdelta) {
  const int64_t j = ((int64_t)blockIdx.x * 256 + threadIdx.x) / kBoxLanes;
  const int sub = threadIdx.x % kBoxLanes;
  const bool live = j < G;
  int first = 0, end = 0;  // point 0 exists: the set holds at least one point
  if (live) {
    first = leaf_first[j];
    end = leaf_first[j + 1];
  }
  Pt q[kBoxPer];
#pragma unroll
  for (int t = 0; t < kBoxPer; ++t) {
    const int i = first + sub + kBoxLanes * t;
    q[t] = pts[i < end ? i : first];
  }
  float lx = INFINITY, ly = INFINITY, lz = INFINITY, hx = -INFINITY, hy = -INFINITY,
        hz = -INFINITY;
#pragma unroll
  for (int t = 0; t < kBoxPer; ++t)
    if (first + sub + kBoxLanes * t < end) {
      lx = fminf(lx, q[t].x); ly = fminf(ly, q[t].y); lz = fminf(lz, q[t].z);
      hx = fmaxf(hx, q[t].x); hy = fmaxf(hy, q[t].y); hz = fmaxf(hz, q[t].z);
    }
#pragma unroll
  for (int o = 1; o < kBoxLanes; o <<= 1) {  // stays within the leaf's lanes
    lx = fminf(lx, __shfl_xor(lx, o, 64)); ly = fminf(ly, __shfl_xor(ly, o, 64));
    lz = fminf(lz, __shfl_xor(lz, o, 64)); hx = fmaxf(hx, __shfl_xor(hx, o, 64));
    hy = fmaxf(hy, __shfl_xor(hy, o, 64)); hz = fmaxf(hz, __shfl_xor(hz, o, 64));
  }
  if (live && sub == 0) {
    gbox[2 * j] = make_float4(lx, ly, lz, 0.0f);
    gbox[2 * j + 1] = make_float4(hx, hy, hz, 0.0f);
    gseg[j] = seg[first];
    gdelta[j] = split_delta(codes, seg, ntot, (int64_t)end - 1);
  }
}

// ---- leaves taken from the runs of an extraction pool ----------------------
// Marching cubes and the approximation emit a set's points in runs that follow
// the volume (one block's brick of cells, isovalue by isovalue), so B
// consecutive points of one set lie close together.  Here the leaves are
// those runs, cut where the set changes and at every multiple of B; only the
// leaves are sorted (set above the top 24 bits of the Morton code of a leaf's
// middle point), and their points are copied, normalised, leaf by leaf into
// that order, which also gives each leaf's box.  The tree and the walk are those of the point-sorted path; every
// box holds exactly its leaf's points, so the fields are the same.

// separator after x: x is the last point, x + 1 is a multiple of B, or the
// set changes (the separator format of leaf_bound_kernel: per 1024-point
// sub-block, 16 mask words and a count).  A block covers kBndSub sub-blocks
// with all its loads issued first.

__global__ __launch_bounds__(kBndBlock) void flat_bound_kernel(const uint32_t *seg, int64_t ntot,
                                                               int B, int64_t nbb,
                                                               uint32_t *block_cnt,
                                                               uint64_t *sepmask) {
  __shared__ uint32_t s_wave[kBndSub][kBndBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // all the set ids first, then the tests: a load inside the short-circuit
  // test waits for its value before the next sub-block's load is issued
  uint32_t s0[kBndSub], s1[kBndSub];
#pragma unroll
  for (int k = 0; k < kBndSub; ++k) {
    const int64_t x = ((int64_t)blockIdx.x * kBndSub + k) * kBndBlock + threadIdx.x;
    s0[k] = 0u;
    s1[k] = 0u;
    if (seg && x + 1 < ntot) {
      s0[k] = seg[x];
      s1[k] = seg[x + 1];
    }
  }
  bool sep[kBndSub];
#pragma unroll
  for (int k = 0; k < kBndSub; ++k) {
    const int64_t x = ((int64_t)blockIdx.x * kBndSub + k) * kBndBlock + threadIdx.x;
    // x + 1 <= ntot < 2^31
    sep[k] = x < ntot && (x + 1 == ntot || (uint32_t)(x + 1) % (uint32_t)B == 0 || s0[k] != s1[k]);
  }
#pragma unroll
  for (int k = 0; k < kBndSub; ++k) {
    const int64_t sb = (int64_t)blockIdx.x * kBndSub + k;
    const uint64_t m = __ballot(sep[k]);
    if (sb < nbb && lane == 0) {
      s_wave[k][w] = (uint32_t)__popcll(m);
      sepmask[sb * (kBndBlock / 64) + w] = m;
    }
  }
  __syncthreads();
  if (threadIdx.x < kBndSub) {
    const int64_t sb = (int64_t)blockIdx.x * kBndSub + threadIdx.x;
    if (sb < nbb) {
      uint32_t t = 0;
      for (int i = 0; i < kBndBlock / 64; ++i) t += s_wave[threadIdx.x][i];
      block_cnt[sb] = t;
    }
  }
}

// One lane per leaf in pool order: its sort key, from its middle point.
template <typename T>
__global__ void pool_leaf_kernel(const T *__restrict__ xyz, const uint32_t *__restrict__ seg,
                                 float dm, const int *__restrict__ pfirst, int64_t G,
                                 uint32_t *key, uint32_t *idx) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= G) return;
  const int first = pfirst[j], mid = first + ((pfirst[j + 1] - first) >> 1);
  const Pt p = normalized(xyz, mid, dm);
  const uint32_t c = spread10(p.x) * 4 + spread10(p.y) * 2 + spread10(p.z);
  key[j] = (seg ? seg[first] << 24 : 0u) | (c >> 6);
  idx[j] = (uint32_t)j;
}

// points of the leaf at sorted position j (pool leaf idx[j])
__global__ void leaf_count_kernel(const uint32_t *idx, const int *pfirst, int64_t G,
                                  uint32_t *cnt) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= G) return;
  const uint32_t L = idx[j];
  cnt[j] = (uint32_t)(pfirst[L + 1] - pfirst[L]);
}

// kCopyLanes lanes per sorted leaf j: each lane copies up to 64 / kCopyLanes
// of its at most 64 points, normalised, to gfirst[j]... (all its loads issued
// before any is used); the group's min and max are the leaf's box; its first
// lane writes the box, the set and the delta of the split after it (-1 after
// a set's last leaf; equal keys split by position).
constexpr int kCopyLanes = 16, kCopyPer = 64 / kCopyLanes;

template <typename T>
__global__ __launch_bounds__(256) void leaf_copy_kernel(
    const T *__restrict__ xyz, float dm, const uint32_t *__restrict__ keys,
    const uint32_t *__restrict__ idx, const int *__restrict__ pfirst,
    const int *__restrict__ gfirst, int64_t G, Pt *pts, float4 *gbox, uint32_t *gseg,
    int *gdelta) {
  const int64_t j = ((int64_t)blockIdx.x * 256 + threadIdx.x) / kCopyLanes;
  const int sub = threadIdx.x % kCopyLanes;
  const bool live = j < G;
  int src = 0, c = 0, dst = 0;
  if (live) {
    const uint32_t L = idx[j];
    src = pfirst[L];
    c = pfirst[L + 1] - src;
    dst = gfirst[j];
  }
  // every point requested before the first division (with the division in
  // the same guarded loop each load waited for its value before the next
  // one was issued); slots past the leaf load its first point again (point
  // src exists: the pool holds at least one point), so no load is guarded
  T raw[kCopyPer][3];
#pragma unroll
  for (int t = 0; t < kCopyPer; ++t) {
    const int e = sub + kCopyLanes * t < c ? src + sub + kCopyLanes * t : src;
    const T *q = xyz + 3 * (int64_t)e;
    raw[t][0] = q[0];
    raw[t][1] = q[1];
    raw[t][2] = q[2];
  }
  Pt p[kCopyPer];
#pragma unroll
  for (int t = 0; t < kCopyPer; ++t)
    if (sub + kCopyLanes * t < c)  // the division of normalized()
      p[t] = Pt{(float)raw[t][0] / dm, (float)raw[t][1] / dm, (float)raw[t][2] / dm};
  float lx = INFINITY, ly = INFINITY, lz = INFINITY, hx = -INFINITY, hy = -INFINITY,
        hz = -INFINITY;
#pragma unroll
  for (int t = 0; t < kCopyPer; ++t)
    if (sub + kCopyLanes * t < c) {
      pts[dst + sub + kCopyLanes * t] = p[t];
      lx = fminf(lx, p[t].x); ly = fminf(ly, p[t].y); lz = fminf(lz, p[t].z);
      hx = fmaxf(hx, p[t].x); hy = fmaxf(hy, p[t].y); hz = fmaxf(hz, p[t].z);
    }
#pragma unroll
  for (int o = 1; o < kCopyLanes; o <<= 1) {  // stays within the group
    lx = fminf(lx, __shfl_xor(lx, o, 64)); ly = fminf(ly, __shfl_xor(ly, o, 64));
    lz = fminf(lz, __shfl_xor(lz, o, 64)); hx = fmaxf(hx, __shfl_xor(hx, o, 64));
    hy = fmaxf(hy, __shfl_xor(hy, o, 64)); hz = fmaxf(hz, __shfl_xor(hz, o, 64));
  }
  if (live && sub == 0) {
    gbox[2 * j] = make_float4(lx, ly, lz, 0.0f);
    gbox[2 * j + 1] = make_float4(hx, hy, hz, 0.0f);
    const uint32_t k = keys[j];
    gseg[j] = k >> 24;
    int d = -1;
    if (j + 1 < G) {
      const uint32_t k2 = keys[j + 1];
      if ((k2 >> 24) == (k >> 24))
        d = k != k2 ? __clz(k ^ k2) : 32 + __clz((uint32_t)j ^ (uint32_t)(j + 1));
    }
    gdelta[j] = d;
  }
}

// delta between leaves x and x+1 (-1 past either end)
__device__ __forceinline__ int leaf_delta(const int *gdelta, int64_t G, int64_t x) {
  return (x < 0 || x + 1 >= G) ? -1 : gdelta[x];
}

// child reference as stored in a record: internal nodes keep their index; a
// leaf ~j becomes -(x + 1), x = the first point of the right child (the end
// of a left leaf j, the start of a right leaf j)
__device__ __forceinline__ int child_ref(int c, bool as_left, const int *leaf_first) {
  return c >= 0 ? c : -(leaf_first[as_left ? ~c + 1 : ~c] + 1);
}

struct Box3 {
  float lx, ly, lz, hx, hy, hz;
};

__device__ __forceinline__ Box3 group_box(const float4 *gbox, int j) {
  const float4 a = gbox[2 * j], b = gbox[2 * j + 1];
  return Box3{a.x, a.y, a.z, b.x, b.y, b.z};
}

__device__ __forceinline__ void child_box(const Node *nodes, const float4 *gbox, int c, float *lo,
                                          float *hi) {
  if (c < 0) {
    const Box3 q = group_box(gbox, ~c);
    lo[0] = q.lx; lo[1] = q.ly; lo[2] = q.lz;
    hi[0] = q.hx; hi[1] = q.hy; hi[2] = q.hz;
  } else {
    const Node &m = nodes[c];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[k] = fminf(m.lmin[k], m.rmin[k]);
      hi[k] = fmaxf(m.lmax[k], m.rmax[k]);
    }
  }
}

// The radix tree is built bottom-up with boxes in one pass (Apetrei,
// "Fast and Simple Agglomerative LBVH Construction", 2014): internal node p is
// the split between sorted leaves p and p+1; a node covering leaves [l, r]
// hangs under split r (as left child) if delta(r) > delta(l-1), else under
// split l-1 (as right child).  Each arriving child records itself and its
// outer bound in the parent's pending slot; the second arrival has both
// children and the parent's range.
//
// Phase 1, one workgroup per chunk of kChunk leaves, all in LDS: splits with
// both leaves in the chunk keep pending slots and arrival counters in LDS
// (workgroup-scope atomics).  A walk that reaches a split on a chunk boundary
// records its half in global memory and counts one arrival there (relaxed);
// the second arrival queues the split.  At the end, finished in-chunk nodes
// are written as whole 64-byte records, and half-finished ones hand their
// pending half and arrival count to global memory.
// Phase 2, one workgroup: the queued splits get their records and climb (see
// crossing_build_kernel).
constexpr int kXBlock = 1024;
constexpr int kXGrid = 256;  // at most this many workgroups share out the sets
constexpr int kChunk = 1024;
constexpr int kChunkBlock = 512;  // 2 leaves per lane, 3 workgroups per CU
constexpr int kUnset = INT_MIN;   // no child recorded (leaf ids are ~j > INT_MIN)
constexpr int kNoRoot = INT_MIN;  // root_of of a set without points

__device__ __forceinline__ Box3 leaf_or_lds_box(int c, int64_t c0, const float4 *gbox,
                                                const float (*s_lo)[kChunk],
                                                const float (*s_hi)[kChunk]) {
  if (c < 0) return group_box(gbox, ~c);
  const int i = (int)(c - c0);
  return Box3{s_lo[0][i], s_lo[1][i], s_lo[2][i], s_hi[0][i], s_hi[1][i], s_hi[2][i]};
}

// pend[p] = (left child, right child, first leaf, last leaf) of split p;
// the leaves here are the collapsed subtrees (ntot of them)
__global__ __launch_bounds__(kChunkBlock) void chunk_build_kernel(
    int64_t ntot, const int *gdelta, const uint32_t *seg, const float4 *gbox, const int *gfirst,
    Node *nodes, int4 *pend, uint32_t *arrive, int *list, uint32_t *count, int *root_of) {
  __shared__ float s_lo[3][kChunk], s_hi[3][kChunk];
  __shared__ int s_delta[kChunk + 1];  // split_delta(c0 - 1 + i)
  __shared__ int2 s_child[kChunk], s_range[kChunk];
  __shared__ uint32_t s_cnt[kChunk];
  const int64_t c0 = (int64_t)blockIdx.x * kChunk;
  const int len = (int)(c0 + kChunk < ntot ? kChunk : ntot - c0);
  const int64_t c1 = c0 + len;
  for (int i = threadIdx.x; i <= len; i += kChunkBlock)
    s_delta[i] = leaf_delta(gdelta, ntot, c0 - 1 + i);
  for (int t = threadIdx.x; t < len; t += kChunkBlock) {
    s_cnt[t] = 0u;
    s_child[t] = make_int2(kUnset, kUnset);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < len; t += kChunkBlock) {
    int64_t l = c0 + t, r = l;
    int cur = ~(int)l;
    while (true) {
      const int dl = s_delta[l - c0], dr = s_delta[r - c0 + 1];  // delta(l-1), delta(r)
      if (dl < 0 && dr < 0) {  // the whole set: cur is its root
        root_of[seg[l]] = cur >= 0 ? cur : -1;  // -1: the set is one leaf
        break;
      }
      const bool as_left = dr > dl;
      const int64_t p = as_left ? r : l - 1;
      if (p < c0 || p > c1 - 2) {  // split on a chunk boundary
        int *h = reinterpret_cast<int *>(pend + p);
        if (as_left) { h[0] = cur; h[2] = (int)l; }
        else { h[1] = cur; h[3] = (int)r; }
        if (__hip_atomic_fetch_add(&arrive[p], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 1u)
          list[atomicAdd(count, 1u)] = (int)p;
        break;
      }
      const int slot = (int)(p - c0);
      if (as_left) { s_child[slot].x = cur; s_range[slot].x = (int)l; }
      else { s_child[slot].y = cur; s_range[slot].y = (int)r; }
      if (__hip_atomic_fetch_add(&s_cnt[slot], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) ==
          0u)
        break;  // the sibling's walk finishes this node
      const int2 ch = s_child[slot], rg = s_range[slot];
      const Box3 a = leaf_or_lds_box(ch.x, c0, gbox, s_lo, s_hi);
      const Box3 b = leaf_or_lds_box(ch.y, c0, gbox, s_lo, s_hi);
      s_lo[0][slot] = fminf(a.lx, b.lx);
      s_lo[1][slot] = fminf(a.ly, b.ly);
      s_lo[2][slot] = fminf(a.lz, b.lz);
      s_hi[0][slot] = fmaxf(a.hx, b.hx);
      s_hi[1][slot] = fmaxf(a.hy, b.hy);
      s_hi[2][slot] = fmaxf(a.hz, b.hz);
      l = rg.x;
      r = rg.y;
      cur = (int)p;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < len - 1; t += kChunkBlock) {
    const int64_t p = c0 + t;
    const uint32_t n = s_cnt[t];
    const int2 ch = s_child[t], rg = s_range[t];
    if (n == 2u) {
      const Box3 a = leaf_or_lds_box(ch.x, c0, gbox, s_lo, s_hi);
      const Box3 b = leaf_or_lds_box(ch.y, c0, gbox, s_lo, s_hi);
      float4 *rec = reinterpret_cast<float4 *>(nodes + p);  // Node field order
      rec[0] = make_float4(a.lx, a.ly, a.lz, a.hx);
      rec[1] = make_float4(a.hy, a.hz, b.lx, b.ly);
      rec[2] = make_float4(b.lz, b.hx, b.hy, b.hz);
      rec[3] = make_float4(__int_as_float(child_ref(ch.x, true, gfirst)),
                           __int_as_float(child_ref(ch.y, false, gfirst)),
                           __int_as_float(gfirst[rg.x]), __int_as_float(gfirst[rg.y + 1] - 1));
    } else if (n == 1u) {  // the other child finishes later, in phase 2
      int *h = reinterpret_cast<int *>(pend + p);
      if (ch.x != kUnset) { h[0] = ch.x; h[2] = rg.x; }
      else { h[1] = ch.y; h[3] = rg.y; }
      arrive[p] = 1u;
    }
  }
}

// Phase 2: the chunk-crossing nodes.  Trees of different sets never meet,
// so workgroup k finishes the sets s with s % gridDim.x == k.  Each lane takes
// splits queued by phase 1 and climbs: it writes the node's record and its
// half of the parent's pending slot, then counts its arrival at the parent
// with a workgroup-scope acquire-release atomic; the second arrival goes on
// with the parent, whose other half and child records it now sees.  A set is
// finished on one CU, so workgroup scope orders it, and the kernel boundary
// orders what phase 1 handed over.  No level barriers: every lane climbs at
// its own pace and stops at a first arrival or a set's root.
__global__ __launch_bounds__(kXBlock) void crossing_build_kernel(
    int64_t ntot, const int *gdelta, const uint32_t *seg, const float4 *gbox, const int *gfirst,
    Node *nodes, int4 *pend, uint32_t *arrive, const int *list, const uint32_t *count0,
    int *root_of) {
  const uint32_t n = *count0;
  for (uint32_t t = threadIdx.x; t < n; t += kXBlock) {
    if (seg[list[t]] % gridDim.x != blockIdx.x) continue;
    int p = list[t];
    while (true) {
      const int4 pd = pend[p];
      Node r;
      child_box(nodes, gbox, pd.x, r.lmin, r.lmax);
      child_box(nodes, gbox, pd.y, r.rmin, r.rmax);
      r.left = child_ref(pd.x, true, gfirst);
      r.right = child_ref(pd.y, false, gfirst);
      r.first = gfirst[pd.z];
      r.last = gfirst[pd.w + 1] - 1;
      nodes[p] = r;
      const int dl = leaf_delta(gdelta, ntot, (int64_t)pd.z - 1);
      const int dr = leaf_delta(gdelta, ntot, pd.w);
      if (dl < 0 && dr < 0) {
        root_of[seg[pd.z]] = p;
        break;
      }
      const bool as_left = dr > dl;
      const int parent = as_left ? pd.w : pd.z - 1;
      int *h = reinterpret_cast<int *>(pend + parent);
      if (as_left) { h[0] = p; h[2] = pd.z; }
      else { h[1] = p; h[3] = pd.w; }
      if (__hip_atomic_fetch_add(&arrive[parent], 1u, __ATOMIC_ACQ_REL,
                                 __HIP_MEMORY_SCOPE_WORKGROUP) == 0u)
        break;  // the sibling is not finished: its lane goes on
      p = parent;
    }
  }
}

constexpr int kTrBlock = 256;
constexpr int kTrWaves = kTrBlock / 64;

// Persistent: a fixed grid of waves (enough to fill every SIMD) works
// through the S * bricks sample bricks, set by set in the order `order`.
// With chunk > 0 a wave takes tickets for runs of consecutive bricks (whole
// columns, then half and quarter columns towards the end) from counters:
// bricks cost very different amounts (samples near a surface are cheap), so
// a wave that finishes early takes more work instead of idling in its
// workgroup.  With slots == 8 the whole columns are dealt to the 8 XCDs
// first: workgroups go to the XCDs in turn (workgroup b to XCD b % 8), and
// the waves of XCD x take, from a counter of their own, the runs of `group`
// columns assigned to x, about an eighth of a set each, so the waves sharing
// an L2 walk one compact part of each set's tree and its points.  Run j goes
// to XCD (j + j / 8) % 8, which moves each XCD to another part of the volume
// from one set to the next.  (Were the workgroups placed otherwise, every
// column would still be walked once; only the L2 sharing would be lost.)  The
// half and quarter columns then come from one shared counter.  With chunk == 0
// (few bricks) wave i takes bricks i, i + waves, ... and touches no counter.
// A brick is 4 x 4 x 4*SPL samples: lane (x, y, z) holds samples
// (x, y, z + 4s), s < SPL.  A wave walks the tree once per brick.  A child is
// entered if any sample's box bound beats that sample's best.  A leaf child
// (at most BUCKET points) is scanned on the spot, before the internal sibling
// is tested, so its points tighten that test.
// A leaf is scanned only for the samples whose box bound beats their best (no
// other sample can improve: the box bound is never above a point's d2, in
// floating point too, since rounding is monotonic).  When at most
// kCompact * SPL samples want it, which is most scans (near half of them want
// 8 or fewer), those samples are packed into the wave (a full scan evaluates
// SPL samples per lane): P = the next power of two >= their
// count, lane l evaluates sample l % P against points l / P, l / P + 64 / P,
// ..., and a min over the lanes of each sample hands the result back to its
// owner.  min is exact, so the result is the one of the full scan.
constexpr int kCompact = 32;

// Occupancy: 8 waves per SIMD (64 VGPRs) for one or two samples per lane; 6
// (80 VGPRs, no spills) for four, which the compiler would otherwise give 87
// VGPRs and 5 waves (a 600x248x248 front, 100 isovalues: walk 5.79 -> 5.12 ms;
// 7 and 8 waves spill and run at 5.25 and 5.22 ms)
template <int BUCKET, int SPL>
__global__ __launch_bounds__(kTrBlock) __attribute__((amdgpu_waves_per_eu(SPL >= 4 ? 6 : 8, 8))) void traverse_kernel(iso::SampleGrid g,
                                                            const Node *__restrict__ nodes,
                                                            const Pt *__restrict__ pts,
                                                            const int64_t *__restrict__ off,
                                                            const int *__restrict__ root_of,
                                                            const int *__restrict__ ne_set,
                                                            const int *__restrict__ order, int S,
                                                            int bx, int by, int bz, int64_t F,
                                                            uint32_t chunk, uint32_t nbig,
                                                            uint32_t nhalf, uint32_t small,
                                                            uint32_t slots, uint32_t group,
                                                            uint32_t *work,
                                                            float *df) {
  static_assert(BUCKET <= 64, "one point per lane");
  __shared__ int s_stack[kTrWaves][64];  // one deferred node per level at most
  // the points being scanned by each wave; slots from BUCKET on stay at +inf
  // (a packed scan reads up to 63 slots past the leaf)
  __shared__ float4 s_pts[kTrWaves][BUCKET + 64];
  // packed samples of a compacted scan: at most one per lane
  constexpr int kPack = kCompact * SPL < 64 ? kCompact * SPL : 64;
  __shared__ float4 s_cand[kTrWaves][kPack];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  s_pts[w][BUCKET + lane] = make_float4(INFINITY, 0.f, 0.f, 0.f);
  const uint32_t nbricks = (uint32_t)bx * by * bz;
  const uint32_t total = nbricks * (uint32_t)S;
  // Tickets: the first nbig are chunks of `chunk` bricks, the next nhalf
  // chunks of chunk / 2, the rest chunks of `small` bricks, so the end of the
  // queue is dealt out ever more finely and the waves run out of work
  // together.  work[0] counts the shared tickets, work[2 + x] those of XCD x.
  auto grab = [&](uint32_t *ctr) -> uint32_t {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(ctr, 1u);
    return (uint32_t)uniform(__shfl((int)t, 0, 64));
  };
  const uint32_t half = chunk / 2;
  auto first_of = [&](uint32_t k) -> uint32_t {
    if (k < nbig) return k * chunk;
    if (k < nbig + nhalf) return nbig * chunk + (k - nbig) * half;
    return nbig * chunk + nhalf * half + (k - nbig - nhalf) * small;
  };
  // the previous brick this wave walked (consecutive bricks of a chunk follow
  // each other along z): its samples' results bound the next brick's
  uint32_t prev_t = 0xffffffffu;
  float prev_best[SPL];
#pragma unroll
  for (int q = 0; q < SPL; ++q) prev_best[q] = -1.0f;
  auto walk = [&](uint32_t t) {
    const uint32_t pos = t / nbricks;  // the sets are walked in the order `order`
    const int seg = uniform(order[pos]);
    const uint32_t brick = t - pos * nbricks;
    const int64_t base = off[seg];
    const int n = (int)(off[seg + 1] - base);
    const int bzi = (int)(brick % bz), byi = (int)((brick / bz) % by);
    const int bxi = (int)(brick / ((uint32_t)bz * by));
    const int xs = bxi * 4 + (lane >> 4), ys = byi * 4 + ((lane >> 2) & 3);
    int zs[SPL];
    bool valid[SPL];
    // a lane's samples share x and y, so the x and y parts of their distances
    // are computed once; an invalid sample's coordinates are finite and never
    // used (its best of -1 asks for no node and its result is not written)
    float sx, sy, sz[SPL], best[SPL];
#pragma unroll
    for (int q = 0; q < SPL; ++q) {
      zs[q] = bzi * 4 * SPL + 4 * q + (lane & 3);
      valid[q] = xs < g.nx && ys < g.ny && zs[q] < g.nz;
      sample_coords(g, xs, ys, zs[q], sx, sy, sz[q]);
      best[q] = valid[q] ? INFINITY : -1.0f;  // invalid samples never ask for a node
    }
    // A start below +inf from the brick before (same sample column, the z
    // block just below): its top sample (x, y, z') has squared distance p to
    // its nearest point P, so this sample's distance to P, hence its minimum,
    // is at most sqrt(p) + (z - z') in sample units; squared and raised by
    // 2^-16 relative (far above the few ulps of rounding in p, the sum and
    // the float d2 of P) it is an upper bound of the float minimum the walk
    // computes, so the walk still finds that minimum (a box or point is
    // skipped only when it cannot go below the bound) and only prunes sooner.
    if (t == prev_t + 1u && bzi != 0) {
      const float p = __shfl(prev_best[SPL - 1], (lane & ~3) | 3, 64);
      const int zp = bzi * 4 * SPL - 1;
      const float szp = (float)(zp * g.ds + g.ds / 2) / (float)g.Z;
#pragma unroll
      for (int q = 0; q < SPL; ++q) {
        if (valid[q] && p >= 0.0f && p < INFINITY) {
          const float b = sqrtf(p) + (sz[q] - szp);
          best[q] = fminf(best[q], b * b * (1.0f + 0x1p-16f));
        }
      }
    }

    // points [first, first + cnt), cnt <= BUCKET: one coalesced load, through
    // this wave's LDS slots, read back at a uniform address (a broadcast), four
    // per step.  Slots past cnt hold a point at +inf, whose d2 is +inf, so the
    // last step needs no mask.  d2 is never NaN, so fminf keeps the smaller.
    // d: each sample's bound for the leaf's box.
    auto scan = [&](int first, int cnt, const float (&d)[SPL]) {
      const Pt pv = pts[first + (lane < cnt ? lane : 0)];
      if (lane < BUCKET)
        s_pts[w][lane] = lane < cnt ? make_float4(pv.x, pv.y, pv.z, 0.f)
                                    : make_float4(INFINITY, 0.f, 0.f, 0.f);
      uint64_t mq[SPL];
      int pc = 0;
#pragma unroll
      for (int q = 0; q < SPL; ++q) {
        mq[q] = __ballot(d[q] < best[q]);
        pc += (int)__popcll(mq[q]);
      }
      if (pc <= kPack) {
        if (pc == 0) return;
        int rank0 = 0;
#pragma unroll
        for (int q = 0; q < SPL; ++q) {
          if ((mq[q] >> lane) & 1u)
            s_cand[w][rank0 + lane_rank(mq[q])] = make_float4(sx, sy, sz[q], 0.f);
          rank0 += (int)__popcll(mq[q]);
        }
        const int lg = pc <= 1 ? 0 : 32 - __builtin_clz((unsigned)(pc - 1));  // P = 1 << lg
        const int j = lane & ((1 << lg) - 1), step = 64 >> lg;
        // The partial minima are kept as bit patterns: a d2 is never NaN or
        // -0 (a sum of squares, +inf at the padding), so its bits order as
        // its values and an integer min needs no NaN canonicalisation; the
        // exchanges are plain ds_bpermute (lanes wrap mod 64, no bound test).
        uint32_t part = 0x7f800000u;  // +inf
        if (j < pc) {
          const float4 c = s_cand[w][j];
          const float4 *row = &s_pts[w][lane >> lg];
          for (int i0 = 0; i0 < cnt; i0 += step) {  // a uniform trip count; +inf past cnt
            const float4 p = row[i0];
            part = min(part, __float_as_uint(d2_point(c.x, c.y, c.z, p.x, p.y, p.z)));
          }
        }
        const int lane4 = lane << 2;
        for (int o = 1 << lg; o < 64; o <<= 1)
          part = min(part, (uint32_t)__builtin_amdgcn_ds_bpermute(lane4 ^ (o << 2), (int)part));
        rank0 = 0;
#pragma unroll
        for (int q = 0; q < SPL; ++q) {
          const float v = __uint_as_float(
              (uint32_t)__builtin_amdgcn_ds_bpermute((rank0 + lane_rank(mq[q])) << 2, (int)part));
          if ((mq[q] >> lane) & 1u) best[q] = fminf(best[q], v);
          rank0 += (int)__popcll(mq[q]);
        }
        return;
      }
      for (int i0 = 0; i0 < cnt; i0 += 4) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float4 p = s_pts[w][(i0 + i) & (BUCKET - 1)];
          // d2_point with the x and y part shared by the lane's samples
          const float dx = sx - p.x, dy = sy - p.y;
          const float pxy = __builtin_fmaf(dy, dy, dx * dx);
#pragma unroll
          for (int q = 0; q < SPL; ++q) {
            const float dz = sz[q] - p.z;
            best[q] = fminf(best[q], __builtin_fmaf(dz, dz, pxy));
          }
        }
      }
    };
    // does any sample of any lane find bound d below its best?
    // (one ballot per comparison: a ballot of an OR-ed bool costs two more
    // vector instructions to rebuild the mask)
    auto wanted = [&](const float (&d)[SPL]) {
      uint64_t m = 0;
#pragma unroll
      for (int q = 0; q < SPL; ++q) m |= __ballot(d[q] < best[q]);
      return m != 0;
    };
    int all = 0;  // valid samples of the brick
#pragma unroll
    for (int q = 0; q < SPL; ++q) all += __popcll(__ballot(valid[q]));

    int node = uniform(root_of[ne_set[seg]]);
    if (node < 0) {
      if (node != kNoRoot) {  // the set is one leaf
        float z[SPL];
#pragma unroll
        for (int q = 0; q < SPL; ++q) z[q] = 0.0f;
        scan((int)base, n, z);
      }
    } else {
      int sp = 0;
      while (true) {
        const Node &r = nodes[uniform(node)];
        const int first = uniform(r.first), last = uniform(r.last);
        const int left = uniform(r.left), right = uniform(r.right);
        float dl[SPL], dr[SPL];
        const float lxy = d2_box_xy(sx, sy, r.lmin, r.lmax), rxy = d2_box_xy(sx, sy, r.rmin, r.rmax);
#pragma unroll
        for (int q = 0; q < SPL; ++q) {
          dl[q] = d2_box_z(sz[q], r.lmin, r.lmax, lxy);
          dr[q] = d2_box_z(sz[q], r.rmin, r.rmax, rxy);
        }
        bool nl = wanted(dl), nr = wanted(dr);
        if (left < 0 || right < 0) {
          bool scanned = false;
          if (left < 0 && nl) { scan(first, -left - 1 - first, dl); scanned = true; }
          if (right < 0 && nr) { scan(-right - 1, last + 2 + right, dr); scanned = true; }
          if (left < 0) nl = false;
          if (right < 0) nr = false;
          if (scanned) {  // the leaf's points may already rule the sibling out
            if (nl) nl = wanted(dl);
            if (nr) nr = wanted(dr);
          }
        }
        int next;
        if (nl && nr) {
          // nearer child first: the one most samples find closer
          int closer = 0;
#pragma unroll
          for (int q = 0; q < SPL; ++q) closer += __popcll(__ballot(valid[q] && dl[q] <= dr[q]));
          const bool left_first = closer * 2 >= all;
          s_stack[w][sp++] = left_first ? right : left;
          next = left_first ? left : right;
        } else if (nl) {
          next = left;
        } else if (nr) {
          next = right;
        } else {
          if (sp == 0) break;
          next = uniform(s_stack[w][--sp]);
        }
        node = next;
      }
    }
#pragma unroll
    for (int q = 0; q < SPL; ++q) {
      if (valid[q] && n > 0) {
        const int64_t o = ((int64_t)xs * g.ny + ys) * g.nz + zs[q];
        df[(int64_t)seg * F + o] = n == 1 ? best[q] : sqrtf(best[q]);
      }
      prev_best[q] = (valid[q] && n > 0) ? best[q] : -1.0f;
    }
    prev_t = t;
  };
  if (chunk == 0) {
    const uint32_t nw = gridDim.x * kTrWaves;
    for (uint32_t t = blockIdx.x * kTrWaves + w; t < total; t += nw) walk(t);
  } else {
    // a ticket is taken only when the wave is free: one taken ahead would
    // wait behind the wave's current chunk, and chunks of the costliest sets
    // run for a millisecond or more.  A wave's first ticket is its own index
    // (the counter hands out the tickets after those), so the waves do not
    // all meet at a counter when they start.
    // XCD x's i-th ticket is column i % group of its run m = i / group, which
    // is run 8m + (x - m) mod 8 of the whole queue: increasing in i, so the first
    // ticket past nbig ends the XCD's part.  With slots == 1 every ticket is
    // shared.  The walk is written once (one copy of its registers).
    const uint32_t nw = gridDim.x * kTrWaves;
    const uint32_t x = blockIdx.x % slots, nwx = (gridDim.x / slots) * kTrWaves;
    uint32_t k = blockIdx.x * kTrWaves + w;    // shared ticket
    uint32_t i = (blockIdx.x / slots) * kTrWaves + w;  // ticket of XCD x
    bool own = slots > 1;
    while (true) {
      uint32_t t0, len;
      if (own) {
        const uint32_t m = i / group;
        const uint32_t kk = (slots * m + ((x - m) & (slots - 1))) * group + i % group;
        if (kk >= nbig) {  // the XCD's columns are all taken
          own = false;
          k = nbig + grab(work);
          continue;
        }
        t0 = kk * chunk;
        len = chunk;
      } else {
        t0 = first_of(k);
        if (t0 >= total) break;
        len = k < nbig ? chunk : k < nbig + nhalf ? half : small;
      }
      const uint32_t end = total - t0 < len ? total : t0 + len;
      for (uint32_t t = t0; t < end; ++t) walk(t);
      if (own) i = nwx + grab(work + 2 + x);
      else k = (slots > 1 ? nbig : nw) + grab(work);
    }
  }
}

}  // namespace

namespace iso {

int df_pool(iso_ctx *ctx, const void *d_xyz, bool is_int, int64_t ntot, const uint32_t *d_seg,
            int nseg, const uint32_t *h_counts, const SampleGrid &g, int method, float *d_df,
            bool normalised, bool coherent) {
  if (ntot <= 0 || nseg <= 0) return ISO_OK;
  if (ntot >= ((int64_t)1 << 31)) return fail(ctx, ISO_ERR_UNSUPPORTED, "more than 2^31 points");
  const int64_t F = (int64_t)g.nx * g.ny * g.nz;
  // x / 1.0f is exact, so pre-normalised points pass through unchanged
  const float dm = normalised ? 1.0f : (float)std::max(g.X, std::max(g.Y, g.Z));
  const bool lbvh = method == ISO_DF_LBVH;

  // set offsets after grouping (all sets, in set order) and of the non-empty
  // ones, then the set id of each non-empty set and the order the LBVH walk
  // takes the non-empty sets in; a word past the block receives the LBVH
  // leaf count
  const size_t ob = sizeof(int64_t) * (size_t)(2 * nseg + 2) + sizeof(int) * (size_t)(2 * nseg);
  const size_t ob_read = (ob + 15) & ~(size_t)15;
  void *d_offs = nullptr, *h_offs = nullptr;
  ISO_TRY(ws_get(ctx, WS_DF_OFFSETS, ob, false, &d_offs));
  // its own pinned area: every copy through it ended before the previous call
  // returned (the leaf count read-back synchronises the stream after them),
  // so the host prepares the offsets while the stream is still busy
  ISO_TRY(pinned_get(ctx, ob_read + 16, &h_offs, 1));
  int64_t *off_set = (int64_t *)h_offs, *off_ne = off_set + nseg + 1;
  int *ne_set = (int *)(off_ne + nseg + 1);
  int *walk_order = ne_set + nseg;
  uint32_t *h_leaves = (uint32_t *)((char *)h_offs + ob_read);
  int S = 0;
  int64_t run = 0;
  for (int s = 0; s < nseg; ++s) {
    off_set[s] = run;
    if (h_counts[s] > 0) {
      ne_set[S] = s;
      off_ne[S++] = run;
    }
    run += h_counts[s];
  }
  off_set[nseg] = run;
  off_ne[S] = run;
  if (run != ntot) return fail(ctx, ISO_ERR_INVALID_ARG, "df_pool: set sizes do not add up");
  // The walk takes the sets with the most points first: on a 600x248x248
  // front with 100 isovalues (samples every 4 voxels) a brick of a set costs
  // about in proportion to the set's points (correlation 0.93 over the 100
  // sets; the extreme isovalues wrap the noise
  // of the whole volume, 4-17M points and up to 3x the cost per brick of the
  // 1M-point middle ones), and walked last such sets left most waves idle
  // while a few finished them.  Cheap sets at the end of the queue let the
  // waves run out of work together (walk 6.2 -> 5.2 ms).
  for (int i = 0; i < S; ++i) walk_order[i] = i;
  std::stable_sort(walk_order, walk_order + S, [&](int a, int b) {
    return off_ne[a + 1] - off_ne[a] > off_ne[b + 1] - off_ne[b];
  });
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(d_offs, h_offs, ob, hipMemcpyHostToDevice, ctx->stream));
  const int64_t *d_off_set = (const int64_t *)d_offs, *d_off_ne = d_off_set + nseg + 1;
  const int *d_ne_set = (const int *)(d_off_ne + nseg + 1);
  const int *d_walk_order = d_ne_set + nseg;

  void *ca = nullptr, *cb = nullptr, *ia = nullptr, *ib = nullptr;
  ISO_TRY(ws_get(ctx, WS_DF_CODES_A, 4 * (size_t)ntot, false, &ca));
  ISO_TRY(ws_get(ctx, WS_DF_CODES_B, 4 * (size_t)ntot, false, &cb));
  ISO_TRY(ws_get(ctx, WS_DF_IDX_A, 4 * (size_t)ntot, false, &ia));
  ISO_TRY(ws_get(ctx, WS_DF_IDX_B, 4 * (size_t)ntot, false, &ib));
  const unsigned gn = (unsigned)((ntot + 255) / 256);
  // LBVH with < 256 sets: one sort of (set, 24-bit code) keys; otherwise the
  // 30-bit code sort (LBVH only) then a stable sort by set
  const bool keyed = lbvh && nseg <= 256;
  const uint32_t *kseg = (nseg > 1) ? d_seg : nullptr;
  // leaves from the pool's runs (ISO_DF_POOL_LEAVES=0 keeps the point sort)
  // Leaves from the pool's runs or from the sorted points: which one is
  // faster depends on the workload (a 600x248x248 front at ds=4: runs, 12.6 against
  // 19.9 ms; small blob volumes at ds=2: sorted points), so the autotuner
  // times both.  ISO_DF_POOL_LEAVES=0 or 1 (read per call) fixes the choice.
  bool can_pool = keyed && coherent, can_sort = true;
  if (const char *e = std::getenv("ISO_DF_POOL_LEAVES")) {
    if (e[0] == '0') can_pool = false;
    else if (e[0] == '1' && can_pool) can_sort = false;
  }
  uint32_t *keys = (uint32_t *)ca, *vals = (uint32_t *)ia;
  uint32_t *sk = nullptr;
  void *sorted = nullptr, *scodes = nullptr;
  ISO_TRY(ws_get(ctx, WS_DF_LEAVES, sizeof(Pt) * (size_t)ntot, false, &sorted));
  // sorted-point path: keys, sort, normalised points in key order
  auto prep_sorted = [&]() -> int {
    keys = (uint32_t *)ca;
    vals = (uint32_t *)ia;
    if (is_int)
      hipLaunchKernelGGL(key_kernel<int32_t>, dim3(gn), dim3(256), 0, ctx->stream,
                         (const int32_t *)d_xyz, kseg, keyed, ntot, dm, (uint32_t *)ca,
                         (uint32_t *)ia);
    else
      hipLaunchKernelGGL(key_kernel<float>, dim3(gn), dim3(256), 0, ctx->stream,
                         (const float *)d_xyz, kseg, keyed, ntot, dm, (uint32_t *)ca,
                         (uint32_t *)ia);
    ISO_HIP_CHECK(ctx, hipGetLastError());
    void *segk = nullptr, *segk2 = nullptr;
    ISO_TRY(ws_get(ctx, WS_DF_SEG_A, 4 * (size_t)ntot, false, &segk));
    ISO_TRY(ws_get(ctx, WS_DF_SEG_B, 4 * (size_t)ntot, false, &segk2));
    sk = (uint32_t *)segk;
    if (keyed) {
      ISO_TRY(radix_sort_pairs(ctx, &keys, &vals, (uint32_t *)cb, (uint32_t *)ib, ntot,
                               kseg ? 32 : 24));
      if (is_int)
        hipLaunchKernelGGL(gather_keyed_kernel<int32_t>, dim3(gn), dim3(256), 0, ctx->stream,
                           (const int32_t *)d_xyz, dm, (const uint32_t *)vals,
                           (const uint32_t *)keys, ntot, (Pt *)sorted, sk);
      else
        hipLaunchKernelGGL(gather_keyed_kernel<float>, dim3(gn), dim3(256), 0, ctx->stream,
                           (const float *)d_xyz, dm, (const uint32_t *)vals,
                           (const uint32_t *)keys, ntot, (Pt *)sorted, sk);
      scodes = keys;  // sorted keys: same set => same top 8 bits
    } else {
      if (lbvh) ISO_TRY(radix_sort_pairs(ctx, &keys, &vals, (uint32_t *)cb, (uint32_t *)ib, ntot, 30));
      if (nseg > 1 && d_seg) {
        hipLaunchKernelGGL(gather_u32_kernel, dim3(gn), dim3(256), 0, ctx->stream, d_seg,
                           (const uint32_t *)vals, ntot, sk);
        uint32_t *sv = vals;
        uint32_t *sv_alt = (vals == (uint32_t *)ia) ? (uint32_t *)ib : (uint32_t *)ia;
        const int sbits = nseg <= 256 ? 8 : (nseg <= 65536 ? 16 : 24);
        ISO_TRY(radix_sort_pairs(ctx, &sk, &sv, (uint32_t *)segk2, sv_alt, ntot, sbits));
        vals = sv;
      } else {
        ISO_TRY(fill_u32(ctx, sk, 0u, (size_t)ntot));
      }
      // sk[i] = set of sorted position i; vals[i] = its point
      ISO_TRY(ws_get(ctx, WS_DF_SORTED_CODES, 4 * (size_t)ntot, false, &scodes));
      if (is_int)
        hipLaunchKernelGGL(gather_kernel<int32_t>, dim3(gn), dim3(256), 0, ctx->stream,
                           (const int32_t *)d_xyz, dm, (const uint32_t *)vals, ntot,
                           (Pt *)sorted, (uint32_t *)scodes);
      else
        hipLaunchKernelGGL(gather_kernel<float>, dim3(gn), dim3(256), 0, ctx->stream,
                           (const float *)d_xyz, dm, (const uint32_t *)vals, ntot,
                           (Pt *)sorted, (uint32_t *)scodes);
    }
    return ISO_OK;
  };
  // one path only: the sort runs once, outside the timed variants
  if (!can_pool) ISO_TRY(prep_sorted());

  if (!lbvh) {
    void *bits = nullptr;
    ISO_TRY(ws_get(ctx, WS_DF_FLAGS, 4 * (size_t)F * S, false, &bits));
    ISO_TRY(fill_u32(ctx, bits, 0x7f800000u, (size_t)F * S));
    const int64_t sblocks = (F + kBfBlock * kBfSamples - 1) / (kBfBlock * kBfSamples);
    const int64_t avg = (ntot + S - 1) / S;
    int64_t chunks = (8 * (int64_t)ctx->cu_count + sblocks * S - 1) / (sblocks * S);
    chunks = std::max<int64_t>(1, std::min<int64_t>(chunks, (avg + 63) / 64));
    hipLaunchKernelGGL(brute_kernel, dim3((unsigned)sblocks, (unsigned)chunks, (unsigned)S),
                       dim3(kBfBlock), 0, ctx->stream, g, F, (const Pt *)sorted, d_off_ne,
                       (int)chunks, (uint32_t *)bits);
    hipLaunchKernelGGL(finish_kernel, dim3((unsigned)((F * S + 255) / 256)), dim3(256), 0,
                       ctx->stream, (uint32_t *)bits, F, d_off_ne, S, d_df);
    ISO_HIP_CHECK(ctx, hipGetLastError());
    return ISO_OK;
  }

  // One variant = B, the most points per leaf: the leaves, the tree over
  // them and the traversal.  All variants return the same bits.
  if ((uint64_t)((g.nx + 3) / 4) * ((g.ny + 3) / 4) * ((g.nz + 3) / 4) * (uint64_t)S >=
      ((uint64_t)1 << 31))
    return fail(ctx, ISO_ERR_UNSUPPORTED, "distance fields: 2^31 or more sample bricks");
  const int64_t nbb = (ntot + kBndBlock - 1) / kBndBlock;
  void *bnd = nullptr, *lfirst = nullptr, *roots = nullptr, *counts = nullptr;
  // separator masks (16 words per block), then block counts, their scan,
  // the total and the scan's workspace
  ISO_TRY(ws_get(ctx, WS_DF_BOUNDS,
                 8 * 16 * (size_t)nbb + 4 * (2 * (size_t)nbb + scan_tmp_words(nbb) + 4), false,
                 &bnd));
  uint64_t *smask = (uint64_t *)bnd;
  uint32_t *bcnt = (uint32_t *)(smask + 16 * nbb), *bbase = bcnt + nbb, *btot = bbase + nbb,
           *btmp = btot + 4;
  ISO_TRY(ws_get(ctx, WS_DF_LEAF_FIRST, 4 * ((size_t)ntot + 1), false, &lfirst));
  int *gfirst = (int *)lfirst;
  ISO_TRY(ws_get(ctx, WS_DF_ROOTS, sizeof(int) * (size_t)nseg, false, &roots));
  ISO_TRY(ws_get(ctx, WS_DF_LEVEL_COUNTS, sizeof(uint32_t) * 16, false, &counts));
  uint32_t *cnt = (uint32_t *)counts;

  int *pfirst = nullptr;  // pool leaves: first point of each, pool order
  if (can_pool) {
    void *pf = nullptr;
    ISO_TRY(ws_get(ctx, WS_DF_POOL_FIRST, 4 * ((size_t)ntot + 1), false, &pf));
    pfirst = (int *)pf;
  }

  auto run_variant = [&](int B, int spl, bool pool_leaves) -> int {
    if (!pool_leaves && can_pool) ISO_TRY(prep_sorted());  // both paths tuned: part of the cost
    if (pool_leaves)
      hipLaunchKernelGGL(flat_bound_kernel, dim3((unsigned)((nbb + kBndSub - 1) / kBndSub)),
                         dim3(kBndBlock), 0, ctx->stream, kseg, ntot, B, nbb, bcnt, smask);
    else
      hipLaunchKernelGGL(leaf_bound_kernel, dim3((unsigned)nbb), dim3(kBndBlock), 0, ctx->stream,
                         (const uint32_t *)scodes, (const uint32_t *)sk, ntot, B, bcnt, smask);
    ISO_TRY(exclusive_scan_u32(ctx, bcnt, bbase, nbb, btot, btmp));
    hipLaunchKernelGGL(leaf_emit_kernel, dim3((unsigned)((nbb + kBndSub - 1) / kBndSub)),
                       dim3(kBndBlock), 0, ctx->stream, ntot, nbb, (const uint64_t *)smask,
                       (const uint32_t *)bbase, pool_leaves ? pfirst : gfirst);
    ISO_HIP_CHECK(ctx, hipGetLastError());
    ISO_HIP_CHECK(ctx, hipMemcpyAsync(h_leaves, btot, 4, hipMemcpyDeviceToHost, ctx->stream));
    ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const int64_t G = *h_leaves;  // leaves; gfirst[0..G] is written
    if (G < S || G > ntot) return fail(ctx, ISO_ERR_HIP, "lbvh: bad leaf count %lld", (long long)G);
    void *nodes = nullptr, *pend = nullptr, *arr = nullptr, *grp = nullptr, *lists = nullptr;
    ISO_TRY(ws_get(ctx, WS_DF_NODES, sizeof(Node) * (size_t)G, false, &nodes));
    ISO_TRY(ws_get(ctx, WS_DF_PENDING, sizeof(int4) * (size_t)G, false, &pend));
    ISO_TRY(ws_get(ctx, WS_DF_FLAGS, sizeof(uint32_t) * (size_t)G, false, &arr));
    // per leaf: box (2 float4), set, delta to the next leaf
    ISO_TRY(ws_get(ctx, WS_DF_GROUPS, (sizeof(float4) * 2 + 8) * (size_t)G, false, &grp));
    ISO_TRY(ws_get(ctx, WS_DF_LEVELS, sizeof(int) * (size_t)G, false, &lists));
    float4 *gbox = (float4 *)grp;
    uint32_t *gseg = (uint32_t *)(gbox + 2 * G);
    int *gdelta = (int *)(gseg + G);
    int *la = (int *)lists;  // splits queued by phase 1
    ISO_TRY(fill_u32(ctx, arr, 0u, (size_t)G));
    ISO_TRY(fill_u32(ctx, roots, (uint32_t)kNoRoot, (size_t)nseg));
    ISO_TRY(fill_u32(ctx, cnt, 0u, 16));
    if (pool_leaves) {
      // keys in pool order, sort of the leaves, their new first points (a
      // scan of their sizes in sorted order), then the copy and the boxes
      void *stmp = nullptr;
      const unsigned lg = (unsigned)((G + 255) / 256);
      if (is_int)
        hipLaunchKernelGGL(pool_leaf_kernel<int32_t>, dim3(lg), dim3(256), 0, ctx->stream,
                           (const int32_t *)d_xyz, kseg, dm, (const int *)pfirst, G,
                           (uint32_t *)ca, (uint32_t *)ia);
      else
        hipLaunchKernelGGL(pool_leaf_kernel<float>, dim3(lg), dim3(256), 0, ctx->stream,
                           (const float *)d_xyz, kseg, dm, (const int *)pfirst, G,
                           (uint32_t *)ca, (uint32_t *)ia);
      ISO_HIP_CHECK(ctx, hipGetLastError());
      uint32_t *lk = (uint32_t *)ca, *lv = (uint32_t *)ia;
      ISO_TRY(radix_sort_pairs(ctx, &lk, &lv, (uint32_t *)cb, (uint32_t *)ib, G, kseg ? 32 : 24));
      hipLaunchKernelGGL(leaf_count_kernel, dim3(lg), dim3(256), 0, ctx->stream,
                         (const uint32_t *)lv, (const int *)pfirst, G, (uint32_t *)la);
      ISO_TRY(ws_get(ctx, WS_SCAN_TMP, 4 * (size_t)scan_tmp_words(G), false, &stmp));
      ISO_TRY(exclusive_scan_u32(ctx, (uint32_t *)la, (uint32_t *)gfirst, G,
                                 (uint32_t *)gfirst + G, (uint32_t *)stmp));
      const unsigned cg = (unsigned)((G * kCopyLanes + 255) / 256);
      if (is_int)
        hipLaunchKernelGGL(leaf_copy_kernel<int32_t>, dim3(cg), dim3(256), 0, ctx->stream,
                           (const int32_t *)d_xyz, dm, (const uint32_t *)lk, (const uint32_t *)lv,
                           (const int *)pfirst, (const int *)gfirst, G, (Pt *)sorted, gbox,
                           gseg, gdelta);
      else
        hipLaunchKernelGGL(leaf_copy_kernel<float>, dim3(cg), dim3(256), 0, ctx->stream,
                           (const float *)d_xyz, dm, (const uint32_t *)lk, (const uint32_t *)lv,
                           (const int *)pfirst, (const int *)gfirst, G, (Pt *)sorted, gbox,
                           gseg, gdelta);
    } else {
      hipLaunchKernelGGL(leaf_box_kernel, dim3((unsigned)((G * kBoxLanes + 255) / 256)), dim3(256), 0,
                         ctx->stream, (const Pt *)sorted, (const uint32_t *)scodes,
                         (const uint32_t *)sk, ntot, (const int *)gfirst, G, gbox, gseg, gdelta);
    }
    hipLaunchKernelGGL(chunk_build_kernel, dim3((unsigned)((G + kChunk - 1) / kChunk)),
                       dim3(kChunkBlock), 0, ctx->stream, G, (const int *)gdelta,
                       (const uint32_t *)gseg, (const float4 *)gbox, (const int *)gfirst,
                       (Node *)nodes, (int4 *)pend, (uint32_t *)arr, la, cnt, (int *)roots);
    hipLaunchKernelGGL(crossing_build_kernel, dim3((unsigned)std::min(nseg, kXGrid)),
                       dim3(kXBlock), 0, ctx->stream, G,
                       (const int *)gdelta, (const uint32_t *)gseg, (const float4 *)gbox,
                       (const int *)gfirst, (Node *)nodes, (int4 *)pend, (uint32_t *)arr,
                       (const int *)la, (const uint32_t *)cnt, (int *)roots);
    // bricks of 4 x 4 x 4*spl samples
    const int bx = (g.nx + 3) / 4, by = (g.ny + 3) / 4, bz = (g.nz + 4 * spl - 1) / (4 * spl);
    // persistent: up to 8 waves per SIMD (4 SIMDs per CU); with at least 4
    // bricks per wave they pull chunks of about 1/16 of their share from
    // cnt[2], rounded up to whole columns of bz bricks (a chunk then starts a
    // column, and every brick but a column's first starts from the bound of
    // the brick below it), else each takes a fixed stride
    const uint64_t total = (uint64_t)bx * by * bz * (uint64_t)S;
    const uint64_t max_waves = (uint64_t)ctx->cu_count * 32;
    const uint64_t waves = std::min<uint64_t>(max_waves, (total + 1) / 2);
    uint32_t chunk = 0u, nbig = 0u, nhalf = 0u, small = 1u;
    if (total >= 4 * waves) {
      const uint64_t c = std::min<uint64_t>(256, std::max<uint64_t>(1, total / (waves * 16)));
      chunk = (uint32_t)(((c + bz - 1) / bz) * bz);
      // the last waves x chunk / 2 bricks go out `small` (a quarter column)
      // at a time: enough to even out when the waves finish their last whole
      // columns, and every other brick of the stretch still starts from the
      // brick below it
      const uint64_t tail = std::min<uint64_t>(total, waves * chunk / 2);
      small = std::max<uint32_t>(1u, chunk / 4);
      // before it, another waves x chunk / 2 bricks in half columns
      const uint64_t mid = std::min<uint64_t>(total - tail, waves * chunk / 2);
      nbig = (uint32_t)((total - tail - mid) / chunk);
      nhalf = chunk >= 2 ? (uint32_t)((total - tail - (uint64_t)nbig * chunk) / (chunk / 2)) : 0u;
    }
    const dim3 tg((unsigned)((waves + kTrWaves - 1) / kTrWaves));
    // whole columns per XCD: runs of an eighth of a set's columns (a
    // 600x248x248 front at ds=4, 100 isovalues: walk 4.64 -> 4.59 ms and
    // 6.2 -> 4.7 GB of L2 misses; stealing another XCD's runs once the own
    // ones are taken made it 4.8 ms); 8 XCDs need a grid of whole rounds of 8
    uint32_t slots = 1, group = 1;
    if (chunk > 0 && tg.x % 8 == 0) {
      slots = 8;
      group = (uint32_t)std::max<uint64_t>(1, (uint64_t)bx * by * bz / chunk / 8);
    }
    auto launch = [&](auto kernel) {
      hipLaunchKernelGGL(kernel, tg, dim3(kTrBlock), 0, ctx->stream, g, (const Node *)nodes,
                         (const Pt *)sorted, d_off_ne, (const int *)roots, d_ne_set,
                         d_walk_order, S, bx,
                         by, bz, F, chunk, nbig, nhalf, small, slots, group, cnt + 2, d_df);
    };
    // the walk alone, timed on the context's stream when a caller asks
    if (ctx->walk_ev[0]) ISO_HIP_CHECK(ctx, hipEventRecord(ctx->walk_ev[0], ctx->stream));
    if (B <= 32) {
      if (spl == 4) launch(traverse_kernel<32, 4>);
      else if (spl == 2) launch(traverse_kernel<32, 2>);
      else launch(traverse_kernel<32, 1>);
    } else {
      if (spl == 4) launch(traverse_kernel<64, 4>);
      else if (spl == 2) launch(traverse_kernel<64, 2>);
      else launch(traverse_kernel<64, 1>);
    }
    ISO_HIP_CHECK(ctx, hipGetLastError());
    if (ctx->walk_ev[1]) ISO_HIP_CHECK(ctx, hipEventRecord(ctx->walk_ev[1], ctx->stream));
    return ISO_OK;
  };
  // (B, samples per lane) per variant, picked by the autotuner; ISO_DF_LEAF=B
  // (8..64) forces B, with ISO_DF_SPL=1, 2 or 4 (default 1)
  static const int kLeaf[4] = {64, 48, 64, 64};
  static const int kSpl[4] = {4, 2, 2, 1};
  if (const char *e = std::getenv("ISO_DF_LEAF")) {  // read per call
    const int v = std::atoi(e);
    const char *q = std::getenv("ISO_DF_SPL");
    const int spl = q && (q[0] == '2' || q[0] == '4') ? q[0] - '0' : 1;
    if (v >= 8 && v <= kMaxLeaf) return run_variant(v, spl, can_pool);
  }
  // variants 0-3: leaves from the pool's runs; 4-7 (or 0-3 when the pool path
  // is out): leaves from the sorted points
  const bool both = can_pool && can_sort;
  const uint64_t key = ((uint64_t)(can_pool ? 1 + both : 0) << 32) |
                       ((uint64_t)ceil_log2((uint64_t)(ntot / S + 1)) << 16) |
                       (uint64_t)ceil_log2((uint64_t)F);
  const char *what = both ? "lbvh (leaves, leaf points, samples per lane) {runs 64,4|runs 48,2|"
                            "runs 64,2|runs 64,1|sorted 64,4|sorted 48,2|sorted 64,2|sorted 64,1}"
                     : can_pool ? "lbvh (leaves from runs: leaf points, samples per lane) "
                                  "{64,4|48,2|64,2|64,1}"
                                : "lbvh (leaves from sorted points: leaf points, samples per lane) "
                                  "{64,4|48,2|64,2|64,1}";
  return tune_pick(ctx, kTuneTraverse, key, both ? 8 : 4, 2, what, [&](int v) {
    return run_variant(kLeaf[v & 3], kSpl[v & 3], can_pool && v < 4);
  });
}

}  // namespace iso

namespace {

int df_single(iso_ctx *ctx, const void *d_xyz, bool is_int, int64_t n, int X, int Y, int Z, int ds,
              int method, float *d_df, bool normalised = false) {
  if (method != ISO_DF_LBVH && method != ISO_DF_BRUTE)
    return iso::fail(ctx, ISO_ERR_INVALID_ARG, "iso_df_points: unknown method %d", method);
  if (n <= 0) return iso::fail(ctx, ISO_ERR_EMPTY, "distance field of an empty point set");
  if (n >= ((int64_t)1 << 31)) return iso::fail(ctx, ISO_ERR_UNSUPPORTED, "more than 2^31 points");
  if (!d_xyz || !d_df || ds <= 0 || X < ds || Y < ds || Z < ds)
    return iso::fail(ctx, ISO_ERR_INVALID_ARG, "iso_df_points: bad arguments");
  iso::SampleGrid g{X, Y, Z, ds, X / ds, Y / ds, Z / ds};
  const uint32_t count = (uint32_t)n;
  ISO_TRY(iso::df_pool(ctx, d_xyz, is_int, n, nullptr, 1, &count, g, method, d_df, normalised));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return ISO_OK;
}

}  // namespace

extern "C" {

int iso_df_points_f3(iso_ctx *ctx, const float *d_xyz, int64_t n, int X, int Y, int Z, int ds,
                     int method, float *d_df) {
  ISO_TRY(iso::enter(ctx));
  return df_single(ctx, d_xyz, false, n, X, Y, Z, ds, method, d_df);
}

int iso_df_points_i3(iso_ctx *ctx, const int32_t *d_xyz, int64_t n, int X, int Y, int Z, int ds,
                     int method, float *d_df) {
  ISO_TRY(iso::enter(ctx));
  return df_single(ctx, d_xyz, true, n, X, Y, Z, ds, method, d_df);
}

int iso_df_points_normalized(iso_ctx *ctx, const float *d_xyz, int64_t n, int X, int Y, int Z,
                             int ds, int method, float *d_df) {
  ISO_TRY(iso::enter(ctx));
  return df_single(ctx, d_xyz, false, n, X, Y, Z, ds, method, d_df, true);
}

// calculate_distances_h (isosurfaces_cuda/src/DistanceField.cu:15-98): the
// brute-force kernel over a grid of nx*ny*nz samples at (x*ds + ds/2, ...)
// (sample coordinates divided by 1, points by 1: both exact), in runs of
// 2,000,000 points as the reference launches them.
int iso_df_voxel_units(iso_ctx *ctx, const float *d_xyz, int64_t n, int nx, int ny, int nz,
                       int ds, float *d_df) {
  ISO_TRY(iso::enter(ctx));
  if (!d_df || (n > 0 && !d_xyz) || n < 0 || nx <= 0 || ny <= 0 || nz <= 0 || ds <= 0)
    return iso::fail(ctx, ISO_ERR_INVALID_ARG, "iso_df_voxel_units: bad arguments");
  if (n >= ((int64_t)1 << 31))
    return iso::fail(ctx, ISO_ERR_UNSUPPORTED, "iso_df_voxel_units: more than 2^31 points");
  constexpr int64_t kRun = 2000000;  // MAX_PER_RUN (DistanceField.cu:14)
  const int64_t F = (int64_t)nx * ny * nz;
  const iso::SampleGrid g{1, 1, 1, ds, nx, ny, nz};
  const int64_t runs = n > 0 ? (n + kRun - 1) / kRun : 0;
  void *bits = nullptr, *doff = nullptr, *hoff = nullptr;
  ISO_TRY(iso::ws_get(ctx, WS_DF_FLAGS, 4 * (size_t)F, false, &bits));
  const size_t ob = sizeof(int64_t) * (size_t)(runs + 1);
  ISO_TRY(iso::ws_get(ctx, WS_DF_OFFSETS, ob, false, &doff));
  ISO_TRY(iso::pinned_get(ctx, ob, &hoff));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));  // pinned block reuse
  int64_t *ho = (int64_t *)hoff;
  for (int64_t r = 0; r <= runs; ++r) ho[r] = std::min(r * kRun, n);
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(doff, hoff, ob, hipMemcpyHostToDevice, ctx->stream));
  const unsigned fb = (unsigned)((F + 255) / 256);
  // the first run's minimum starts at 1e38 (DistanceField.cu:52)
  ISO_TRY(iso::fill_u32(ctx, bits, 0x7e967699u, (size_t)F));
  const int64_t sblocks = (F + kBfBlock * kBfSamples - 1) / (kBfBlock * kBfSamples);
  for (int64_t r = 0; r < runs; ++r) {
    if (r > 0)
      hipLaunchKernelGGL(square_into_bits_kernel, dim3(fb), dim3(256), 0, ctx->stream, d_df, F,
                         (uint32_t *)bits);
    const int64_t len = ho[r + 1] - ho[r];
    int64_t chunks = (8 * (int64_t)ctx->cu_count + sblocks - 1) / sblocks;
    chunks = std::max<int64_t>(1, std::min<int64_t>(chunks, (len + 63) / 64));
    hipLaunchKernelGGL(brute_kernel, dim3((unsigned)sblocks, (unsigned)chunks, 1u), dim3(kBfBlock),
                       0, ctx->stream, g, F, (const Pt *)d_xyz, (const int64_t *)doff + r,
                       (int)chunks, (uint32_t *)bits);
    hipLaunchKernelGGL(root_of_bits_kernel, dim3(fb), dim3(256), 0, ctx->stream,
                       (const uint32_t *)bits, F, d_df);
  }
  if (runs == 0)
    hipLaunchKernelGGL(root_of_bits_kernel, dim3(fb), dim3(256), 0, ctx->stream,
                       (const uint32_t *)bits, F, d_df);
  ISO_HIP_CHECK(ctx, hipGetLastError());
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return ISO_OK;
}

}  // extern "C"
