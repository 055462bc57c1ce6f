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
//   max(bmin - s, s - bmax, 0), which are <= |s - p| for every p in the box
//   (rounding is monotone), so bound <= d2(p) and pruning with `bound < best`
//   never discards a closer point: LBVH == brute force, bit for bit.
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
//     radix sort by code then by segment, and per segment the binary radix
//     tree over the keys (code, index) (height <= 64, each node covers a
//     contiguous leaf range), built bottom-up together with its boxes inside
//     chunks of 1024 leaves in LDS, then the few chunk-crossing nodes one
//     tree level per launch;
//   - traversal is wave-cooperative: a wave walks the tree once per 4x4x4
//     brick of samples with a wave-uniform stack (64 entries in LDS; the
//     height bound means it cannot overflow); a child is entered if any
//     lane's bound beats that lane's best; subtrees of <= 16 points are
//     scanned from one coalesced load, broadcast to the wave through LDS.
#include "iso_df.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace iso {
int radix_sort_pairs(iso_ctx *ctx, uint32_t **keys, uint32_t **vals, uint32_t *keys_alt,
                     uint32_t *vals_alt, int64_t n, int bits);
}  // namespace iso

namespace {


__device__ __forceinline__ float d2_point(float sx, float sy, float sz, float px, float py,
                                          float pz) {
  float dx = sx - px, dy = sy - py, dz = sz - pz;
  return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

__device__ __forceinline__ float d2_box(float sx, float sy, float sz, const float *lo,
                                        const float *hi) {
  float dx = fmaxf(fmaxf(lo[0] - sx, sx - hi[0]), 0.0f);
  float dy = fmaxf(fmaxf(lo[1] - sy, sy - hi[1]), 0.0f);
  float dz = fmaxf(fmaxf(lo[2] - sz, sz - hi[2]), 0.0f);
  return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

__device__ __forceinline__ void sample_coords(const iso::SampleGrid &g, int x, int y, int z,
                                              float &sx, float &sy, float &sz) {
  sx = (float)(x * g.ds + g.ds / 2) / (float)g.X;
  sy = (float)(y * g.ds + g.ds / 2) / (float)g.Y;
  sz = (float)(z * g.ds + g.ds / 2) / (float)g.Z;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

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

// Normalised points and their sort keys.  keyed == false: the 30-bit Morton
// code (morton3D_d, BVH.cu:95-106).  keyed == true: one 32-bit key, the set
// in the top 8 bits above the top 24 bits of the code, so a single sort both
// groups the sets and orders each set along the curve (sets < 256).
template <typename T>
__global__ void normalize_kernel(const T *xyz, const uint32_t *seg, bool keyed, int64_t n,
                                 float dm, float4 *pts, uint32_t *codes, uint32_t *idx) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x = (float)xyz[3 * i] / dm, y = (float)xyz[3 * i + 1] / dm, z = (float)xyz[3 * i + 2] / dm;
  pts[i] = make_float4(x, y, z, 0.0f);
  if (codes) {
    const uint32_t c = spread10(x) * 4 + spread10(y) * 2 + spread10(z);
    codes[i] = keyed ? ((seg ? seg[i] << 24 : 0u) | (c >> 6)) : c;
    idx[i] = (uint32_t)i;
  }
}

// points in sorted order; the set of each from its key's top 8 bits
__global__ void gather_keyed_kernel(const float4 *pts, const uint32_t *idx, const uint32_t *keys,
                                    int64_t n, float4 *pts_out, uint32_t *seg_out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  pts_out[i] = pts[idx[i]];
  seg_out[i] = keys[i] >> 24;
}

// points in final order and their Morton codes (recomputed from the point)
__global__ void gather_kernel(const float4 *pts, const uint32_t *idx, int64_t n, float4 *pts_out,
                              uint32_t *codes_out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 p = pts[idx[i]];
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
                                                         const float4 *__restrict__ pts,
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
    if ((int)threadIdx.x < cnt) s_pts[threadIdx.x] = pts[c + threadIdx.x];
    __syncthreads();
    int i = 0;
    for (; i + 3 < cnt; i += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float4 P = s_pts[i + u];
#pragma unroll
        for (int q = 0; q < kBfSamples; ++q) {
          const float d = d2_point(sx[q], sy[q], sz[q], P.x, P.y, P.z);
          best[q] = d < best[q] ? d : best[q];
        }
      }
    }
    for (; i < cnt; ++i) {
      const float4 P = s_pts[i];
#pragma unroll
      for (int q = 0; q < kBfSamples; ++q) {
        const float d = d2_point(sx[q], sy[q], sz[q], P.x, P.y, P.z);
        best[q] = d < best[q] ? d : best[q];
      }
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

// ---- LBVH -----------------------------------------------------------------
struct Node {
  float lmin[3], lmax[3];
  float rmin[3], rmax[3];
  int left, right;  // >= 0 internal node, < 0 leaf ~index (global indices)
  int first, last;  // sorted leaf range covered (global)
};
static_assert(sizeof(Node) == 64, "node record is one 64-byte line");

// Split strength between sorted leaves x and x+1 (global indices): the
// common-prefix length of the 64-bit keys (Morton code, index); -1 across a
// set boundary or past the end.  Keys are unique, so the binary radix tree
// over them is unique and at most 64 levels high.
__device__ __forceinline__ int split_delta(const uint32_t *codes, const uint32_t *seg, int64_t ntot,
                                           int64_t x) {
  if (x < 0 || x + 1 >= ntot) return -1;
  if (seg[x] != seg[x + 1]) return -1;
  const uint32_t a = codes[x], b = codes[x + 1];
  if (a != b) return __clz(a ^ b);
  return 32 + __clz((uint32_t)x ^ (uint32_t)(x + 1));
}

__device__ __forceinline__ void child_box(const Node *nodes, const float4 *pts, int c, float *lo,
                                          float *hi) {
  if (c < 0) {
    float4 p = pts[~c];
    lo[0] = hi[0] = p.x;
    lo[1] = hi[1] = p.y;
    lo[2] = hi[2] = p.z;
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
// Phase 2, one launch per tree level: queued splits get their records, then
// arrive at their parents; the second arrival queues the parent for the next
// launch.  Kernel boundaries order all data handed between workgroups, so no
// thread pays for an agent-scope release/acquire (an L2 writeback/invalidate
// on each of the 8 XCDs).
constexpr int kMaxLevels = 64;  // tree height bound
constexpr int kLvlBlock = 256;
constexpr int kLvlGrid = 2048;
constexpr int kChunk = 1024;
constexpr int kChunkBlock = 512;  // 2 leaves per lane, 3 workgroups per CU
constexpr int kUnset = INT_MIN;   // no child recorded (leaf ids are ~j > INT_MIN)

__device__ __forceinline__ void push_ready(int parent, bool want, uint32_t *arrive, int *list,
                                           uint32_t *count) {
  bool ready = false;
  if (want && parent >= 0)
    ready = __hip_atomic_fetch_add(&arrive[parent], 1u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT) == 1u;
  const uint64_t m = __ballot(ready);
  if (m == 0) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
  base = __shfl(base, leader, 64);
  if (ready) {
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    list[base + __popcll(m & lt)] = parent;
  }
}

struct Box3 {
  float lx, ly, lz, hx, hy, hz;
};

__device__ __forceinline__ Box3 leaf_or_lds_box(int c, int64_t c0, const float4 *pts,
                                                const float (*s_lo)[kChunk],
                                                const float (*s_hi)[kChunk]) {
  if (c < 0) {
    const float4 q = pts[~c];
    return Box3{q.x, q.y, q.z, q.x, q.y, q.z};
  }
  const int i = (int)(c - c0);
  return Box3{s_lo[0][i], s_lo[1][i], s_lo[2][i], s_hi[0][i], s_hi[1][i], s_hi[2][i]};
}

// pend[p] = (left child, right child, first leaf, last leaf) of split p
__global__ __launch_bounds__(kChunkBlock) void chunk_build_kernel(
    int64_t ntot, const uint32_t *codes, const uint32_t *seg, const float4 *pts, Node *nodes,
    int4 *pend, uint32_t *arrive, int *list, uint32_t *count, int *root_of) {
  __shared__ float s_lo[3][kChunk], s_hi[3][kChunk];
  __shared__ int s_delta[kChunk + 1];  // split_delta(c0 - 1 + i)
  __shared__ int2 s_child[kChunk], s_range[kChunk];
  __shared__ uint32_t s_cnt[kChunk];
  const int64_t c0 = (int64_t)blockIdx.x * kChunk;
  const int len = (int)(c0 + kChunk < ntot ? kChunk : ntot - c0);
  const int64_t c1 = c0 + len;
  for (int i = threadIdx.x; i <= len; i += kChunkBlock)
    s_delta[i] = split_delta(codes, seg, ntot, c0 - 1 + i);
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
        root_of[seg[l]] = cur;
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
      const Box3 a = leaf_or_lds_box(ch.x, c0, pts, s_lo, s_hi);
      const Box3 b = leaf_or_lds_box(ch.y, c0, pts, s_lo, s_hi);
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
      const Box3 a = leaf_or_lds_box(ch.x, c0, pts, s_lo, s_hi);
      const Box3 b = leaf_or_lds_box(ch.y, c0, pts, s_lo, s_hi);
      float4 *rec = reinterpret_cast<float4 *>(nodes + p);  // Node field order
      rec[0] = make_float4(a.lx, a.ly, a.lz, a.hx);
      rec[1] = make_float4(a.hy, a.hz, b.lx, b.ly);
      rec[2] = make_float4(b.lz, b.hx, b.hy, b.hz);
      rec[3] = make_float4(__int_as_float(ch.x), __int_as_float(ch.y), __int_as_float(rg.x),
                           __int_as_float(rg.y));
    } else if (n == 1u) {  // the other child finishes later, level by level
      int *h = reinterpret_cast<int *>(pend + p);
      if (ch.x != kUnset) { h[0] = ch.x; h[2] = rg.x; }
      else { h[1] = ch.y; h[3] = rg.y; }
      arrive[p] = 1u;
    }
  }
}

__global__ __launch_bounds__(kLvlBlock) void level_build_kernel(
    int64_t ntot, const uint32_t *codes, const uint32_t *seg, const float4 *pts, Node *nodes,
    int4 *pend, uint32_t *arrive, const int *list_in, const uint32_t *count_in, int *list_out,
    uint32_t *count_out, int *root_of) {
  const uint32_t n = *count_in;
  for (uint32_t base = blockIdx.x * kLvlBlock; base < n; base += gridDim.x * kLvlBlock) {
    const uint32_t t = base + threadIdx.x;
    const bool live = t < n;
    int parent = -1;
    if (live) {
      const int p = list_in[t];
      const int4 pd = pend[p];
      Node r;
      child_box(nodes, pts, pd.x, r.lmin, r.lmax);
      child_box(nodes, pts, pd.y, r.rmin, r.rmax);
      r.left = pd.x;
      r.right = pd.y;
      r.first = pd.z;
      r.last = pd.w;
      nodes[p] = r;
      const int dl = split_delta(codes, seg, ntot, (int64_t)pd.z - 1);
      const int dr = split_delta(codes, seg, ntot, pd.w);
      if (dl < 0 && dr < 0) {
        root_of[seg[pd.z]] = p;
      } else {
        const bool as_left = dr > dl;
        parent = as_left ? pd.w : pd.z - 1;
        int *h = reinterpret_cast<int *>(pend + parent);
        if (as_left) { h[0] = p; h[2] = pd.z; }
        else { h[1] = p; h[3] = pd.w; }
      }
    }
    push_ready(parent, live, arrive, list_out, count_out);
  }
}

constexpr int kTrBlock = 256;
constexpr int kTrWaves = kTrBlock / 64;

// grid: (bricks / (waves * IL), segments).  A wave owns IL bricks of 4x4x4
// samples (one sample per lane per brick) and walks the tree once per brick;
// the IL walks advance in lock step, so their node loads are in flight
// together.  A child is entered if any lane's box bound beats its best.
template <int IL, int BUCKET>
__global__ __launch_bounds__(kTrBlock) void traverse_kernel(iso::SampleGrid g,
                                                            const Node *__restrict__ nodes,
                                                            const float4 *__restrict__ pts,
                                                            const int64_t *__restrict__ off,
                                                            const int *__restrict__ root_of,
                                                            const int *__restrict__ ne_set,
                                                            int bx, int by, int bz, int64_t F,
                                                            float *df) {
  __shared__ int s_stack[kTrWaves][IL][64];
  __shared__ float4 s_pts[kTrWaves][BUCKET];  // the current bucket of each wave
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int seg = blockIdx.y;
  const int64_t base = off[seg];
  const int n = (int)(off[seg + 1] - base);
  const int64_t nbricks = (int64_t)bx * by * bz;
  const int64_t brick0 = ((int64_t)blockIdx.x * kTrWaves + w) * IL;
  if (brick0 >= nbricks || n <= 0) return;  // whole wave exits together
  float sx[IL], sy[IL], sz[IL], best[IL];
  bool valid[IL], active[IL];
  int node[IL], sp[IL], xs[IL], ys[IL], zs[IL];
#pragma unroll
  for (int t = 0; t < IL; ++t) {
    const int64_t brick = brick0 + t;
    active[t] = brick < nbricks;  // wave-uniform
    const int64_t b = active[t] ? brick : brick0;
    const int bzi = (int)(b % bz), byi = (int)((b / bz) % by);
    const int bxi = (int)(b / ((int64_t)bz * by));
    xs[t] = bxi * 4 + (lane >> 4);
    ys[t] = byi * 4 + ((lane >> 2) & 3);
    zs[t] = bzi * 4 + (lane & 3);
    valid[t] = active[t] && xs[t] < g.nx && ys[t] < g.ny && zs[t] < g.nz;
    sample_coords(g, valid[t] ? xs[t] : 0, valid[t] ? ys[t] : 0, valid[t] ? zs[t] : 0, sx[t],
                  sy[t], sz[t]);
    best[t] = valid[t] ? INFINITY : -1.0f;  // invalid lanes never ask for a node
    sp[t] = 0;
  }
  if (n == 1) {
    const float4 p = pts[base];
#pragma unroll
    for (int t = 0; t < IL; ++t) {
      best[t] = d2_point(sx[t], sy[t], sz[t], p.x, p.y, p.z);
      active[t] = false;
    }
  } else {
    const int root = uniform(root_of[ne_set[seg]]);
#pragma unroll
    for (int t = 0; t < IL; ++t) {
      node[t] = root;
      active[t] = active[t] && root >= 0;  // a set of >= 2 points always has an internal root
    }
  }
  while (true) {
    bool any = false;
#pragma unroll
    for (int t = 0; t < IL; ++t) any |= active[t];
    if (!any) break;
    // all IL node records are loaded and tested before any walk branches, so
    // their loads are in flight together (a finished walk re-reads its last node)
    int firsts[IL], lasts[IL], lefts[IL], rights[IL];
    float dls[IL], drs[IL];
#pragma unroll
    for (int t = 0; t < IL; ++t) {
      const Node &r = nodes[uniform(node[t])];
      firsts[t] = uniform(r.first);
      lasts[t] = uniform(r.last);
      lefts[t] = uniform(r.left);
      rights[t] = uniform(r.right);
      dls[t] = d2_box(sx[t], sy[t], sz[t], r.lmin, r.lmax);
      drs[t] = d2_box(sx[t], sy[t], sz[t], r.rmin, r.rmax);
    }
#pragma unroll
    for (int t = 0; t < IL; ++t) {
      if (!active[t]) continue;
      const int first = firsts[t], last = lasts[t], left = lefts[t], right = rights[t];
      const float dl = dls[t], dr = drs[t];
      int next = -1;
      if (__ballot(dl < best[t] || dr < best[t]) != 0) {
        const int cnt = last - first + 1;
        if (cnt <= BUCKET) {
          // the bucket's points come in one coalesced load and go through
          // this wave's LDS slots; reads at a uniform address broadcast them,
          // 4 at a time and branch-free (past the bucket: +inf)
          const float4 pv = pts[first + (lane < cnt ? lane : 0)];
          if (lane < cnt) s_pts[w][lane] = pv;
          for (int i0 = 0; i0 < cnt; i0 += 4) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const float4 q = s_pts[w][(i0 + i) & (BUCKET - 1)];
              const float d = d2_point(sx[t], sy[t], sz[t], q.x, q.y, q.z);
              const float dv = i0 + i < cnt ? d : INFINITY;
              best[t] = dv < best[t] ? dv : best[t];
            }
          }
        } else {
          bool nl, nr;
          if (left < 0) { best[t] = dl < best[t] ? dl : best[t]; nl = false; }  // leaf: bound == d2
          else nl = __ballot(dl < best[t]) != 0;
          if (right < 0) { best[t] = dr < best[t] ? dr : best[t]; nr = false; }
          else nr = __ballot(dr < best[t]) != 0;
          if (nl && nr) {
            // nearer child first: the one most lanes find closer
            const bool left_first = __popcll(__ballot(valid[t] && dl <= dr)) * 2 >=
                                    __popcll(__ballot(valid[t]));
            if (left_first) { s_stack[w][t][sp[t]++] = right; next = left; }
            else { s_stack[w][t][sp[t]++] = left; next = right; }
          } else if (nl) {
            next = left;
          } else if (nr) {
            next = right;
          }
        }
      }
      if (next < 0) {
        if (sp[t] == 0) active[t] = false;
        else next = uniform(s_stack[w][t][--sp[t]]);
      }
      if (next >= 0) node[t] = next;
    }
  }
#pragma unroll
  for (int t = 0; t < IL; ++t)
    if (valid[t]) {
      const int64_t o = ((int64_t)xs[t] * g.ny + ys[t]) * g.nz + zs[t];
      df[(int64_t)seg * F + o] = n == 1 ? best[t] : sqrtf(best[t]);
    }
}

}  // namespace

namespace iso {

int df_pool(iso_ctx *ctx, const void *d_xyz, bool is_int, int64_t ntot, const uint32_t *d_seg,
            int nseg, const uint32_t *h_counts, const SampleGrid &g, int method, float *d_df,
            bool normalised) {
  if (ntot <= 0 || nseg <= 0) return ISO_OK;
  if (ntot >= ((int64_t)1 << 31)) return fail(ctx, ISO_ERR_UNSUPPORTED, "more than 2^31 points");
  const int64_t F = (int64_t)g.nx * g.ny * g.nz;
  // x / 1.0f is exact, so pre-normalised points pass through unchanged
  const float dm = normalised ? 1.0f : (float)std::max(g.X, std::max(g.Y, g.Z));
  const bool lbvh = method == ISO_DF_LBVH;

  // set offsets after grouping (all sets, in set order) and of the non-empty ones
  // ... followed by the set id of each non-empty set
  const size_t ob = sizeof(int64_t) * (size_t)(2 * nseg + 2) + sizeof(int) * (size_t)nseg;
  void *d_offs = nullptr, *h_offs = nullptr;
  ISO_TRY(ws_get(ctx, WS_DF_OFFSETS, ob, false, &d_offs));
  ISO_TRY(pinned_get(ctx, ob, &h_offs));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));  // pinned block reuse
  int64_t *off_set = (int64_t *)h_offs, *off_ne = off_set + nseg + 1;
  int *ne_set = (int *)(off_ne + nseg + 1);
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
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(d_offs, h_offs, ob, hipMemcpyHostToDevice, ctx->stream));
  const int64_t *d_off_set = (const int64_t *)d_offs, *d_off_ne = d_off_set + nseg + 1;
  const int *d_ne_set = (const int *)(d_off_ne + nseg + 1);

  void *pts = nullptr, *ca = nullptr, *cb = nullptr, *ia = nullptr, *ib = nullptr;
  ISO_TRY(ws_get(ctx, WS_DF_POINTS, sizeof(float4) * (size_t)ntot, false, &pts));
  ISO_TRY(ws_get(ctx, WS_DF_CODES_A, 4 * (size_t)ntot, false, &ca));
  ISO_TRY(ws_get(ctx, WS_DF_CODES_B, 4 * (size_t)ntot, false, &cb));
  ISO_TRY(ws_get(ctx, WS_DF_IDX_A, 4 * (size_t)ntot, false, &ia));
  ISO_TRY(ws_get(ctx, WS_DF_IDX_B, 4 * (size_t)ntot, false, &ib));
  const unsigned gn = (unsigned)((ntot + 255) / 256);
  // LBVH with < 256 sets: one sort of (set, 24-bit code) keys; otherwise the
  // 30-bit code sort (LBVH only) then a stable sort by set
  const bool keyed = lbvh && nseg <= 256;
  const uint32_t *kseg = (nseg > 1) ? d_seg : nullptr;
  if (is_int)
    hipLaunchKernelGGL(normalize_kernel<int32_t>, dim3(gn), dim3(256), 0, ctx->stream,
                       (const int32_t *)d_xyz, kseg, keyed, ntot, dm, (float4 *)pts,
                       (uint32_t *)ca, (uint32_t *)ia);
  else
    hipLaunchKernelGGL(normalize_kernel<float>, dim3(gn), dim3(256), 0, ctx->stream,
                       (const float *)d_xyz, kseg, keyed, ntot, dm, (float4 *)pts,
                       (uint32_t *)ca, (uint32_t *)ia);
  ISO_HIP_CHECK(ctx, hipGetLastError());

  uint32_t *keys = (uint32_t *)ca, *vals = (uint32_t *)ia;
  void *segk = nullptr, *segk2 = nullptr;
  ISO_TRY(ws_get(ctx, WS_DF_SEG_A, 4 * (size_t)ntot, false, &segk));
  ISO_TRY(ws_get(ctx, WS_DF_SEG_B, 4 * (size_t)ntot, false, &segk2));
  uint32_t *sk = (uint32_t *)segk;
  void *sorted = nullptr, *scodes = nullptr;
  ISO_TRY(ws_get(ctx, WS_DF_LEAVES, sizeof(float4) * (size_t)ntot, false, &sorted));
  if (keyed) {
    ISO_TRY(radix_sort_pairs(ctx, &keys, &vals, (uint32_t *)cb, (uint32_t *)ib, ntot,
                             kseg ? 32 : 24));
    hipLaunchKernelGGL(gather_keyed_kernel, dim3(gn), dim3(256), 0, ctx->stream,
                       (const float4 *)pts, (const uint32_t *)vals, (const uint32_t *)keys, ntot,
                       (float4 *)sorted, sk);
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
    hipLaunchKernelGGL(gather_kernel, dim3(gn), dim3(256), 0, ctx->stream, (const float4 *)pts,
                       (const uint32_t *)vals, ntot, (float4 *)sorted, (uint32_t *)scodes);
  }

  if (!lbvh) {
    void *bits = nullptr;
    ISO_TRY(ws_get(ctx, WS_DF_FLAGS, 4 * (size_t)F * S, false, &bits));
    ISO_TRY(fill_u32(ctx, bits, 0x7f800000u, (size_t)F * S));
    const int64_t sblocks = (F + kBfBlock * kBfSamples - 1) / (kBfBlock * kBfSamples);
    const int64_t avg = (ntot + S - 1) / S;
    int64_t chunks = (8 * (int64_t)ctx->cu_count + sblocks * S - 1) / (sblocks * S);
    chunks = std::max<int64_t>(1, std::min<int64_t>(chunks, (avg + 63) / 64));
    hipLaunchKernelGGL(brute_kernel, dim3((unsigned)sblocks, (unsigned)chunks, (unsigned)S),
                       dim3(kBfBlock), 0, ctx->stream, g, F, (const float4 *)sorted, d_off_ne,
                       (int)chunks, (uint32_t *)bits);
    hipLaunchKernelGGL(finish_kernel, dim3((unsigned)((F * S + 255) / 256)), dim3(256), 0,
                       ctx->stream, (uint32_t *)bits, F, d_off_ne, S, d_df);
    ISO_HIP_CHECK(ctx, hipGetLastError());
    return ISO_OK;
  }

  void *nodes = nullptr, *pend = nullptr, *arr = nullptr, *roots = nullptr;
  ISO_TRY(ws_get(ctx, WS_DF_NODES, sizeof(Node) * (size_t)ntot, false, &nodes));
  ISO_TRY(ws_get(ctx, WS_DF_PENDING, sizeof(int4) * (size_t)ntot, false, &pend));
  ISO_TRY(ws_get(ctx, WS_DF_FLAGS, sizeof(uint32_t) * (size_t)ntot, false, &arr));
  ISO_TRY(ws_get(ctx, WS_DF_ROOTS, sizeof(int) * (size_t)nseg, false, &roots));
  ISO_TRY(fill_u32(ctx, arr, 0u, (size_t)ntot));
  ISO_TRY(fill_u32(ctx, roots, 0xffffffffu, (size_t)nseg));  // -1: no tree (yet)
  // level worklists: two ping-pong lists + one counter per level
  void *lists = nullptr, *counts = nullptr;
  ISO_TRY(ws_get(ctx, WS_DF_LEVELS, sizeof(int) * 2 * (size_t)ntot, false, &lists));
  ISO_TRY(ws_get(ctx, WS_DF_LEVEL_COUNTS, sizeof(uint32_t) * (kMaxLevels + 2), false, &counts));
  ISO_TRY(fill_u32(ctx, counts, 0u, kMaxLevels + 2));
  int *la = (int *)lists, *lb = (int *)lists + ntot;
  uint32_t *cnt = (uint32_t *)counts;
  hipLaunchKernelGGL(chunk_build_kernel, dim3((unsigned)((ntot + kChunk - 1) / kChunk)),
                     dim3(kChunkBlock), 0, ctx->stream, ntot, (const uint32_t *)scodes,
                     (const uint32_t *)sk, (const float4 *)sorted, (Node *)nodes, (int4 *)pend,
                     (uint32_t *)arr, la, cnt, (int *)roots);
  // nodes crossing chunk boundaries: about (leaves / kChunk) * log2(leaves) of them
  const unsigned lg = (unsigned)std::min<int64_t>(
      kLvlGrid, std::max<int64_t>(1, (ntot / kChunk * 32 + kLvlBlock - 1) / kLvlBlock));
  for (int lvl = 0; lvl < kMaxLevels; ++lvl) {
    hipLaunchKernelGGL(level_build_kernel, dim3(lg), dim3(kLvlBlock), 0, ctx->stream, ntot,
                       (const uint32_t *)scodes, (const uint32_t *)sk, (const float4 *)sorted,
                       (Node *)nodes, (int4 *)pend, (uint32_t *)arr, (const int *)la,
                       (const uint32_t *)(cnt + lvl), lb, cnt + lvl + 1, (int *)roots);
    std::swap(la, lb);
  }
  // Traversal variants (walks per wave, bucket size), picked by the
  // autotuner; ISO_DF_TRAVERSE="walks,bucket" forces one (1 or 2; 8..64).
  static const int2 kVar[4] = {make_int2(1, 16), make_int2(1, 8), make_int2(1, 32),
                               make_int2(2, 16)};
  static const int2 forced = [] {
    int2 c = make_int2(0, 0);
    if (const char *e = std::getenv("ISO_DF_TRAVERSE")) {
      int a = 0, b = 0;
      if (std::sscanf(e, "%d,%d", &a, &b) == 2 && (a == 1 || a == 2) &&
          (b == 8 || b == 16 || b == 32 || b == 64))
        c = make_int2(a, b);
    }
    return c;
  }();
  auto launch_cfg = [&](int2 c) -> int {
    const int bx = (g.nx + 3) / 4, by = (g.ny + 3) / 4, bz = (g.nz + 3) / 4;
    const int64_t bricks = (int64_t)bx * by * bz;
    const int64_t per_block = (int64_t)kTrWaves * c.x;
    const dim3 tg((unsigned)((bricks + per_block - 1) / per_block), (unsigned)S);
    auto launch = [&](auto kernel) {
      hipLaunchKernelGGL(kernel, tg, dim3(kTrBlock), 0, ctx->stream, g, (const Node *)nodes,
                         (const float4 *)sorted, d_off_ne, (const int *)roots, d_ne_set, bx, by,
                         bz, F, d_df);
    };
    switch (c.x * 100 + c.y) {
      case 108: launch(traverse_kernel<1, 8>); break;
      case 132: launch(traverse_kernel<1, 32>); break;
      case 164: launch(traverse_kernel<1, 64>); break;
      case 208: launch(traverse_kernel<2, 8>); break;
      case 216: launch(traverse_kernel<2, 16>); break;
      case 232: launch(traverse_kernel<2, 32>); break;
      case 264: launch(traverse_kernel<2, 64>); break;
      default: launch(traverse_kernel<1, 16>); break;
    }
    ISO_HIP_CHECK(ctx, hipGetLastError());
    return ISO_OK;
  };
  if (forced.x) return launch_cfg(forced);
  const uint64_t key = ((uint64_t)ceil_log2((uint64_t)(ntot / S + 1)) << 16) |
                       (uint64_t)ceil_log2((uint64_t)F);
  return tune_pick(ctx, kTuneTraverse, key, 4, 0, "traverse (walks,bucket) {1,16|1,8|1,32|2,16}",
                   [&](int v) { return launch_cfg(kVar[v]); });
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

}  // extern "C"
