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
//   - brute force: a lane holds 8 samples in registers and the points stream
//     through scalar loads (one point broadcast to the wave), 7 VALU ops per
//     (sample, point); point ranges split over workgroups merge with an
//     integer atomicMin on the float bits (valid for non-negative floats);
//   - LBVH: 30-bit Morton codes (the reference's quantisation), a stable LDS
//     radix sort by code then by segment, Karras' radix tree per segment with
//     the index tie-break for equal codes (height <= 64, each node covers a
//     contiguous leaf range); bottom-up boxes inside chunks of 2048 leaves in
//     LDS, then the few chunk-crossing nodes one tree level per launch;
//   - traversal is wave-cooperative: a wave owns a (4*SPL)x4x4 brick of
//     samples, SPL per lane, and walks the tree once with a wave-uniform stack
//     (64 entries, 256 B of LDS; the height bound means it cannot overflow); a
//     child is entered if any sample's bound beats that sample's best;
//     subtrees of <= 8 points are scanned with scalar-loaded points.
#include "iso_df.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace iso {
int radix_sort_pairs(iso_ctx *ctx, uint32_t **keys, uint32_t **vals, uint32_t *keys_alt,
                     uint32_t *vals_alt, int64_t n, int bits);
}  // namespace iso

namespace {

constexpr int kBucket = 8;

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

template <typename T>
__global__ void normalize_kernel(const T *xyz, int64_t n, float dm, float4 *pts, uint32_t *codes,
                                 uint32_t *idx) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x = (float)xyz[3 * i] / dm, y = (float)xyz[3 * i + 1] / dm, z = (float)xyz[3 * i + 2] / dm;
  pts[i] = make_float4(x, y, z, 0.0f);
  if (codes) {
    // morton3D_d (BVH.cu:95-106)
    codes[i] = spread10(x) * 4 + spread10(y) * 2 + spread10(z);
    idx[i] = (uint32_t)i;
  }
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

__global__ __launch_bounds__(kBfBlock) void brute_kernel(iso::SampleGrid g, int64_t F,
                                                         const float4 *__restrict__ pts,
                                                         const int64_t *__restrict__ off,
                                                         int nchunks, uint32_t *best_bits) {
  const int seg = blockIdx.z;
  const int64_t a = off[seg], b = off[seg + 1];
  const int64_t n = b - a;
  const int64_t pchunk = (n + nchunks - 1) / nchunks;
  const int64_t p0 = a + (int64_t)blockIdx.y * pchunk;
  int64_t p1 = p0 + pchunk;
  if (p1 > b) p1 = b;
  if (p0 >= p1) return;
  const int64_t s0 = ((int64_t)blockIdx.x * kBfBlock + threadIdx.x) * kBfSamples;
  float sx[kBfSamples], sy[kBfSamples], sz[kBfSamples], best[kBfSamples];
#pragma unroll
  for (int q = 0; q < kBfSamples; ++q) {
    int64_t idx = s0 + q < F ? s0 + q : F - 1;
    int z = (int)(idx % g.nz), y = (int)((idx / g.nz) % g.ny), x = (int)(idx / ((int64_t)g.nz * g.ny));
    sample_coords(g, x, y, z, sx[q], sy[q], sz[q]);
    best[q] = INFINITY;
  }
  for (int64_t p = p0; p < p1; ++p) {
    const float4 P = pts[p];
#pragma unroll
    for (int q = 0; q < kBfSamples; ++q) {
      float d = d2_point(sx[q], sy[q], sz[q], P.x, P.y, P.z);
      best[q] = d < best[q] ? d : best[q];
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

// delta over local indices i, j of a segment of n leaves starting at `base`
__device__ __forceinline__ int delta(const uint32_t *codes, int64_t base, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  uint32_t a = codes[base + i], b = codes[base + j];
  if (a != b) return __clz(a ^ b);
  return 32 + __clz((uint32_t)(i ^ j));
}

// Karras, "Maximizing Parallelism in the Construction of BVHs, Octrees, and
// k-d Trees" (HPG 2012), keys (code, index) so equal codes split evenly.
// Internal node of local index i lives at global index base + i; the root of
// a segment is its first index.
__global__ void karras_kernel(const uint32_t *codes, const uint32_t *segid, const int64_t *off,
                              int64_t ntot, Node *nodes, int *parent_node, int *parent_leaf) {
  int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= ntot) return;
  const int seg = (int)segid[gi];
  const int64_t base = off[seg];
  const int n = (int)(off[seg + 1] - base);
  const int i = (int)(gi - base);
  if (i >= n - 1) return;
  int d = delta(codes, base, n, i, i + 1) - delta(codes, base, n, i, i - 1) > 0 ? 1 : -1;
  int dmin = delta(codes, base, n, i, i - d);
  int lmax = 2;
  while (delta(codes, base, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(codes, base, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = delta(codes, base, n, i, j);
  int s = 0;
  for (int div = 2;; div <<= 1) {
    int t = (l + div - 1) / div;
    if (delta(codes, base, n, i, i + (s + t) * d) > dnode) s += t;
    if (t <= 1) break;
  }
  int gamma = i + s * d + (d < 0 ? d : 0);
  int lo = i < j ? i : j, hi = i < j ? j : i;
  Node &nd = nodes[gi];
  nd.first = (int)(base + lo);
  nd.last = (int)(base + hi);
  const int G = (int)(base + gamma);
  if (lo == gamma) { nd.left = ~G; parent_leaf[G] = (int)gi; }
  else { nd.left = G; parent_node[G] = (int)gi; }
  if (hi == gamma + 1) { nd.right = ~(G + 1); parent_leaf[G + 1] = (int)gi; }
  else { nd.right = G + 1; parent_node[G + 1] = (int)gi; }
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

// Bottom-up boxes, one launch per tree level.  A node is ready once both
// children are done; the second child to finish (relaxed atomic counter, no
// fence) appends the parent to the next level's worklist.  Box data written
// in one launch is read in a later one, so kernel boundaries order it and no
// thread pays for an agent-wide cache writeback/invalidate (8 XCDs, 8 L2s).
constexpr int kMaxLevels = 64;  // tree height bound (delta strictly grows to 63)
constexpr int kLvlBlock = 256;
constexpr int kLvlGrid = 2048;

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

// Boxes of every node whose leaf range lies inside one chunk of kChunk sorted
// leaves, one workgroup per chunk.  Each leaf's thread walks up: at an
// in-chunk parent it bumps an LDS arrival counter (workgroup scope); the
// second arriver reads both child boxes (points, or LDS boxes of in-chunk
// children), writes the parent's record and its box, and goes on upwards.
// A walk that reaches a node crossing the chunk boundary counts one arrival
// there in global memory (relaxed); the second arrival queues that node for
// the level-by-level pass, which runs after this kernel has ended.
constexpr int kChunk = 2048;
constexpr int kChunkBlock = 1024;  // 2 leaves per lane: 16 waves to overlap the walks

__global__ __launch_bounds__(kChunkBlock) void chunk_boxes_kernel(
    int64_t ntot, Node *nodes, const float4 *pts, const int *parent_node, const int *parent_leaf,
    uint32_t *arrive, int *list, uint32_t *count) {
  __shared__ float s_lo[3][kChunk], s_hi[3][kChunk];
  __shared__ uint32_t s_cnt[kChunk];
  const int64_t c0 = (int64_t)blockIdx.x * kChunk;
  const int64_t c1 = c0 + kChunk < ntot ? c0 + kChunk : ntot;
  for (int t = threadIdx.x; t < kChunk; t += kChunkBlock) s_cnt[t] = 0u;
  __syncthreads();
  for (int64_t j = c0 + threadIdx.x; j < c1; j += kChunkBlock) {
    int p = parent_leaf[j];
    while (p >= 0) {
      const int first = nodes[p].first, last = nodes[p].last;
      if (first < c0 || last >= c1) {
        // crossing node: finished later, level by level
        if (__hip_atomic_fetch_add(&arrive[p], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 1u)
          list[atomicAdd(count, 1u)] = p;
        break;
      }
      const int slot = p - (int)c0;
      if (__hip_atomic_fetch_add(&s_cnt[slot], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) ==
          0u)
        break;  // the sibling's walk finishes this node
      const int left = nodes[p].left, right = nodes[p].right;
      float lo[3], hi[3], rlo[3], rhi[3];
      if (left < 0) {
        const float4 q = pts[~left];
        lo[0] = hi[0] = q.x; lo[1] = hi[1] = q.y; lo[2] = hi[2] = q.z;
      } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) { lo[k] = s_lo[k][left - c0]; hi[k] = s_hi[k][left - c0]; }
      }
      if (right < 0) {
        const float4 q = pts[~right];
        rlo[0] = rhi[0] = q.x; rlo[1] = rhi[1] = q.y; rlo[2] = rhi[2] = q.z;
      } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) { rlo[k] = s_lo[k][right - c0]; rhi[k] = s_hi[k][right - c0]; }
      }
      Node &nd = nodes[p];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        nd.lmin[k] = lo[k]; nd.lmax[k] = hi[k];
        nd.rmin[k] = rlo[k]; nd.rmax[k] = rhi[k];
        s_lo[k][slot] = fminf(lo[k], rlo[k]);
        s_hi[k][slot] = fmaxf(hi[k], rhi[k]);
      }
      p = parent_node[p];
    }
  }
}

__global__ __launch_bounds__(kLvlBlock) void level_boxes_kernel(Node *nodes, const float4 *pts,
                                                                const int *parent_node,
                                                                uint32_t *arrive, const int *list_in,
                                                                const uint32_t *count_in,
                                                                int *list_out, uint32_t *count_out) {
  const uint32_t n = *count_in;
  for (uint32_t base = blockIdx.x * kLvlBlock; base < n; base += gridDim.x * kLvlBlock) {
    const uint32_t t = base + threadIdx.x;
    const bool live = t < n;
    int parent = -1;
    if (live) {
      const int node = list_in[t];
      Node &nd = nodes[node];
      float lo[3], hi[3];
      child_box(nodes, pts, nd.left, lo, hi);
#pragma unroll
      for (int k = 0; k < 3; ++k) { nd.lmin[k] = lo[k]; nd.lmax[k] = hi[k]; }
      child_box(nodes, pts, nd.right, lo, hi);
#pragma unroll
      for (int k = 0; k < 3; ++k) { nd.rmin[k] = lo[k]; nd.rmax[k] = hi[k]; }
      parent = parent_node[node];
    }
    push_ready(parent, live, arrive, list_out, count_out);
  }
}

constexpr int kTrBlock = 256;
constexpr int kTrWaves = kTrBlock / 64;

// grid: (bricks / waves, segments).  A wave owns a brick of (4*SPL) x 4 x 4
// samples, SPL per lane, and walks the tree once for all of them: a child is
// entered if any sample's box bound beats that sample's best.
template <int SPL>
__global__ __launch_bounds__(kTrBlock) void traverse_kernel(iso::SampleGrid g,
                                                            const Node *__restrict__ nodes,
                                                            const float4 *__restrict__ pts,
                                                            const int64_t *__restrict__ off,
                                                            int bx, int by, int bz, int64_t F,
                                                            float *df) {
  __shared__ int s_stack[kTrWaves][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int seg = blockIdx.y;
  const int64_t base = off[seg];
  const int n = (int)(off[seg + 1] - base);
  const int64_t brick = (int64_t)blockIdx.x * kTrWaves + w;
  const int64_t nbricks = (int64_t)bx * by * bz;
  if (brick >= nbricks || n <= 0) return;  // whole wave exits together
  const int bzi = (int)(brick % bz), byi = (int)((brick / bz) % by);
  const int bxi = (int)(brick / ((int64_t)bz * by));
  const int y = byi * 4 + ((lane >> 2) & 3), z = bzi * 4 + (lane & 3);
  float sx[SPL], sy[SPL], sz[SPL], best[SPL];
  bool valid[SPL];
#pragma unroll
  for (int q = 0; q < SPL; ++q) {
    const int x = (bxi * SPL + q) * 4 + (lane >> 4);
    valid[q] = x < g.nx && y < g.ny && z < g.nz;
    sample_coords(g, valid[q] ? x : 0, valid[q] ? y : 0, valid[q] ? z : 0, sx[q], sy[q], sz[q]);
    best[q] = valid[q] ? INFINITY : -1.0f;  // invalid samples never ask for a node
  }

  if (n == 1) {
    const float4 p = pts[base];
#pragma unroll
    for (int q = 0; q < SPL; ++q) best[q] = d2_point(sx[q], sy[q], sz[q], p.x, p.y, p.z);
  } else {
    int sp = 0;
    int node = uniform((int)base);
    while (true) {
      const Node &r = nodes[node];
      const int first = uniform(r.first), last = uniform(r.last);
      const int left = uniform(r.left), right = uniform(r.right);
      float dl[SPL], dr[SPL];
      bool wl = false, wr = false;
      float ml = INFINITY, mr = INFINITY;
#pragma unroll
      for (int q = 0; q < SPL; ++q) {
        dl[q] = d2_box(sx[q], sy[q], sz[q], r.lmin, r.lmax);
        dr[q] = d2_box(sx[q], sy[q], sz[q], r.rmin, r.rmax);
        wl |= dl[q] < best[q];
        wr |= dr[q] < best[q];
        if (valid[q]) { ml = fminf(ml, dl[q]); mr = fminf(mr, dr[q]); }
      }
      int next = -1;
      if (__ballot(wl || wr) != 0) {
        if (last - first + 1 <= kBucket) {
          for (int p = first; p <= last; ++p) {
            const float4 P = pts[uniform(p)];
#pragma unroll
            for (int q = 0; q < SPL; ++q) {
              const float d = d2_point(sx[q], sy[q], sz[q], P.x, P.y, P.z);
              best[q] = d < best[q] ? d : best[q];
            }
          }
        } else {
          bool nl, nr;
          if (left < 0) {  // leaf: its bound is its d2
#pragma unroll
            for (int q = 0; q < SPL; ++q) best[q] = dl[q] < best[q] ? dl[q] : best[q];
            nl = false;
          } else {
            nl = __ballot(wl) != 0;
          }
          if (right < 0) {
#pragma unroll
            for (int q = 0; q < SPL; ++q) best[q] = dr[q] < best[q] ? dr[q] : best[q];
            nr = false;
          } else {
            nr = __ballot(wr) != 0;
          }
          if (nl && nr) {
            // nearer child first: the one most lanes find closer
            bool any_valid = false;
#pragma unroll
            for (int q = 0; q < SPL; ++q) any_valid |= valid[q];
            const bool left_first =
                __popcll(__ballot(any_valid && ml <= mr)) * 2 >= __popcll(__ballot(any_valid));
            if (left_first) { s_stack[w][sp++] = right; next = left; }
            else { s_stack[w][sp++] = left; next = right; }
          } else if (nl) {
            next = left;
          } else if (nr) {
            next = right;
          }
        }
      }
      if (next < 0) {
        if (sp == 0) break;
        next = s_stack[w][--sp];
      }
      node = uniform(next);
    }
  }
#pragma unroll
  for (int q = 0; q < SPL; ++q)
    if (valid[q]) {
      const int x = (bxi * SPL + q) * 4 + (lane >> 4);
      const int64_t o = ((int64_t)x * g.ny + y) * g.nz + z;
      df[(int64_t)seg * F + o] = n == 1 ? best[q] : sqrtf(best[q]);
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
  const size_t ob = sizeof(int64_t) * (size_t)(2 * nseg + 2);
  void *d_offs = nullptr, *h_offs = nullptr;
  ISO_TRY(ws_get(ctx, WS_DF_OFFSETS, ob, false, &d_offs));
  ISO_TRY(pinned_get(ctx, ob, &h_offs));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));  // pinned block reuse
  int64_t *off_set = (int64_t *)h_offs, *off_ne = off_set + nseg + 1;
  int S = 0;
  int64_t run = 0;
  for (int s = 0; s < nseg; ++s) {
    off_set[s] = run;
    if (h_counts[s] > 0) off_ne[S++] = run;
    run += h_counts[s];
  }
  off_set[nseg] = run;
  off_ne[S] = run;
  if (run != ntot) return fail(ctx, ISO_ERR_INVALID_ARG, "df_pool: set sizes do not add up");
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(d_offs, h_offs, ob, hipMemcpyHostToDevice, ctx->stream));
  const int64_t *d_off_set = (const int64_t *)d_offs, *d_off_ne = d_off_set + nseg + 1;

  void *pts = nullptr, *ca = nullptr, *cb = nullptr, *ia = nullptr, *ib = nullptr;
  ISO_TRY(ws_get(ctx, WS_DF_POINTS, sizeof(float4) * (size_t)ntot, false, &pts));
  ISO_TRY(ws_get(ctx, WS_DF_CODES_A, 4 * (size_t)ntot, false, &ca));
  ISO_TRY(ws_get(ctx, WS_DF_CODES_B, 4 * (size_t)ntot, false, &cb));
  ISO_TRY(ws_get(ctx, WS_DF_IDX_A, 4 * (size_t)ntot, false, &ia));
  ISO_TRY(ws_get(ctx, WS_DF_IDX_B, 4 * (size_t)ntot, false, &ib));
  const unsigned gn = (unsigned)((ntot + 255) / 256);
  if (is_int)
    hipLaunchKernelGGL(normalize_kernel<int32_t>, dim3(gn), dim3(256), 0, ctx->stream,
                       (const int32_t *)d_xyz, ntot, dm, (float4 *)pts, (uint32_t *)ca,
                       (uint32_t *)ia);
  else
    hipLaunchKernelGGL(normalize_kernel<float>, dim3(gn), dim3(256), 0, ctx->stream,
                       (const float *)d_xyz, ntot, dm, (float4 *)pts, (uint32_t *)ca,
                       (uint32_t *)ia);
  ISO_HIP_CHECK(ctx, hipGetLastError());

  // order: by Morton code (LBVH only), then stably by set
  uint32_t *keys = (uint32_t *)ca, *vals = (uint32_t *)ia;
  if (lbvh) ISO_TRY(radix_sort_pairs(ctx, &keys, &vals, (uint32_t *)cb, (uint32_t *)ib, ntot, 30));
  void *segk = nullptr, *segk2 = nullptr;
  ISO_TRY(ws_get(ctx, WS_DF_SEG_A, 4 * (size_t)ntot, false, &segk));
  ISO_TRY(ws_get(ctx, WS_DF_SEG_B, 4 * (size_t)ntot, false, &segk2));
  uint32_t *sk = (uint32_t *)segk;
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
  void *sorted = nullptr, *scodes = nullptr;
  ISO_TRY(ws_get(ctx, WS_DF_LEAVES, sizeof(float4) * (size_t)ntot, false, &sorted));
  ISO_TRY(ws_get(ctx, WS_DF_SORTED_CODES, 4 * (size_t)ntot, false, &scodes));
  hipLaunchKernelGGL(gather_kernel, dim3(gn), dim3(256), 0, ctx->stream, (const float4 *)pts,
                     (const uint32_t *)vals, ntot, (float4 *)sorted, (uint32_t *)scodes);

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

  void *nodes = nullptr, *par = nullptr, *arr = nullptr;
  ISO_TRY(ws_get(ctx, WS_DF_NODES, sizeof(Node) * (size_t)ntot, false, &nodes));
  ISO_TRY(ws_get(ctx, WS_DF_PARENTS, sizeof(int) * (size_t)(2 * ntot), false, &par));
  ISO_TRY(ws_get(ctx, WS_DF_FLAGS, sizeof(uint32_t) * (size_t)ntot, false, &arr));
  int *parent_node = (int *)par, *parent_leaf = (int *)par + ntot;
  ISO_TRY(fill_u32(ctx, par, 0xffffffffu, (size_t)(2 * ntot)));
  ISO_TRY(fill_u32(ctx, arr, 0u, (size_t)ntot));
  hipLaunchKernelGGL(karras_kernel, dim3(gn), dim3(256), 0, ctx->stream, (const uint32_t *)scodes,
                     (const uint32_t *)sk, d_off_set, ntot, (Node *)nodes, parent_node, parent_leaf);
  // level worklists: two ping-pong lists + one counter per level
  void *lists = nullptr, *counts = nullptr;
  ISO_TRY(ws_get(ctx, WS_DF_LEVELS, sizeof(int) * 2 * (size_t)ntot, false, &lists));
  ISO_TRY(ws_get(ctx, WS_DF_LEVEL_COUNTS, sizeof(uint32_t) * (kMaxLevels + 2), false, &counts));
  ISO_TRY(fill_u32(ctx, counts, 0u, kMaxLevels + 2));
  int *la = (int *)lists, *lb = (int *)lists + ntot;
  uint32_t *cnt = (uint32_t *)counts;
  hipLaunchKernelGGL(chunk_boxes_kernel, dim3((unsigned)((ntot + kChunk - 1) / kChunk)),
                     dim3(kChunkBlock), 0, ctx->stream, ntot, (Node *)nodes, (const float4 *)sorted,
                     (const int *)parent_node, (const int *)parent_leaf, (uint32_t *)arr, la, cnt);
  // nodes crossing chunk boundaries: about (leaves / kChunk) * log2(leaves) of them
  const unsigned lg = (unsigned)std::min<int64_t>(
      kLvlGrid, std::max<int64_t>(1, (ntot / kChunk * 32 + kLvlBlock - 1) / kLvlBlock));
  for (int lvl = 0; lvl < kMaxLevels; ++lvl) {
    hipLaunchKernelGGL(level_boxes_kernel, dim3(lg), dim3(kLvlBlock), 0, ctx->stream,
                       (Node *)nodes, (const float4 *)sorted, (const int *)parent_node,
                       (uint32_t *)arr, (const int *)la, (const uint32_t *)(cnt + lvl), lb,
                       cnt + lvl + 1);
    std::swap(la, lb);
  }
  // samples per lane: 1 by default; ISO_DF_SPL=1|2|4 overrides (tuning)
  static const int spl = [] {
    const char *e = std::getenv("ISO_DF_SPL");
    const int v = e ? std::atoi(e) : 1;
    return (v == 1 || v == 2 || v == 4) ? v : 1;
  }();
  const int bx = (g.nx + 4 * spl - 1) / (4 * spl), by = (g.ny + 3) / 4, bz = (g.nz + 3) / 4;
  const int64_t bricks = (int64_t)bx * by * bz;
  const dim3 tg((unsigned)((bricks + kTrWaves - 1) / kTrWaves), (unsigned)S);
  if (spl == 1)
    hipLaunchKernelGGL(traverse_kernel<1>, tg, dim3(kTrBlock), 0, ctx->stream, g,
                       (const Node *)nodes, (const float4 *)sorted, d_off_ne, bx, by, bz, F, d_df);
  else if (spl == 2)
    hipLaunchKernelGGL(traverse_kernel<2>, tg, dim3(kTrBlock), 0, ctx->stream, g,
                       (const Node *)nodes, (const float4 *)sorted, d_off_ne, bx, by, bz, F, d_df);
  else
    hipLaunchKernelGGL(traverse_kernel<4>, tg, dim3(kTrBlock), 0, ctx->stream, g,
                       (const Node *)nodes, (const float4 *)sorted, d_off_ne, bx, by, bz, F, d_df);
  ISO_HIP_CHECK(ctx, hipGetLastError());
  return ISO_OK;
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
