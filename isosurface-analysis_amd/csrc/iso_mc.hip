// Surface extraction: marching cubes (unique vertices) and the reference's
// voxel approximation, for one or many isovalues per pass over the volume.
//
// Reference semantics, marching cubes (isosurfaces_cuda/src/MarchingCube.cu,
// MarchingCube.cpp:588-710, thrustWrapper.cu:44-56):
//   every classified cell (cube index bit k = corner k < iso, corners clamped
//   to the grid) emits, for every crossed edge, lerp(v_a, v_b, t) with
//   t = (iso - f_a)/(f_b - f_a) in the cell's own edge orientation; the
//   vertex list is sorted and exact duplicates are removed.
//
// MI355X design -- one lane per grid point P, owning the three edges
// P->P+e_x, P->P+e_y, P->P+e_z:
//   - a vertex whose along-edge coordinate is not an integer lies strictly
//     inside its grid edge, so it can only coincide with vertices of the same
//     edge; those come in two orientations (+ from cells on one side, - from
//     the other), so deduplication is a compare of at most two values, done
//     in registers by the edge's owner;
//   - a vertex that lands exactly on a grid point (t rounds to 0 or 1, or a
//     sample equals iso) can coincide across edges; it is inserted into a
//     hash set keyed by (isovalue, grid point) and emitted by the winner;
//   - which orientations exist depends only on which cells contain the edge
//     (grid extent and the reference's classify limit), computed once;
//   - the workgroup loops over only the isovalues inside the [min,max] of its
//     samples (block culling), so one read of the volume serves all k
//     isovalues of an ISM;
//   - output slots come from one atomic per wave per isovalue (wave prefix
//     sum in registers); the vertex set is exact, its order unspecified.
#include "iso_internal.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

namespace iso {
int exclusive_scan_u32(iso_ctx *ctx, const uint32_t *d_in, uint32_t *d_out, int64_t n,
                       uint32_t *d_total, uint32_t *d_tmp);
size_t scan_tmp_words(int64_t n);
int radix_sort_pairs(iso_ctx *ctx, uint32_t **keys, uint32_t **vals, uint32_t *keys_alt,
                     uint32_t *vals_alt, int64_t n, int bits);
int iota_u32(iso_ctx *ctx, uint32_t *d, int64_t n);
}  // namespace iso

namespace {

constexpr int kMcBlock = 256;
constexpr int kMcWaves = kMcBlock / 64;
constexpr int kMcRound = 32;  // isovalues per pool reservation (bit masks of 32)
constexpr unsigned long long kEmpty = ~0ull;

// ---- case tables, generated from first principles -------------------------
// Corners 0..7 of a cell sit at (0,0,0) (1,0,0) (1,1,0) (0,1,0) (0,0,1)
// (1,0,1) (1,1,1) (0,1,1); edge e joins kEdge[e][0] -> kEdge[e][1] (the
// reference's numbering, MarchingCube.cu:208-265).  A case's vertex count is
// 3*(#crossed edges - 2*#polygons); polygons join crossed edges face by face,
// cutting off the corners below iso on faces with four crossed edges.
const int kEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                          {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
const int kFace[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4},
                         {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}};

int find_root(int *p, int x) {
  while (p[x] != x) x = p[x] = p[p[x]];
  return x;
}

struct CaseTables {
  uint8_t nverts[256];
  CaseTables() {
    for (int c = 0; c < 256; ++c) {
      int crossed = 0, ne = 0, parent[12];
      for (int e = 0; e < 12; ++e) {
        parent[e] = e;
        if (((c >> kEdge[e][0]) & 1) != ((c >> kEdge[e][1]) & 1)) { crossed |= 1 << e; ++ne; }
      }
      for (int f = 0; f < 6; ++f) {
        int fe[4], ce[4], nc = 0;
        for (int i = 0; i < 4; ++i) {
          int a = kFace[f][i], b = kFace[f][(i + 1) & 3];
          for (int e = 0; e < 12; ++e)
            if ((kEdge[e][0] == a && kEdge[e][1] == b) || (kEdge[e][0] == b && kEdge[e][1] == a))
              fe[i] = e;
          if (crossed >> fe[i] & 1) ce[nc++] = fe[i];
        }
        if (nc == 2) parent[find_root(parent, ce[0])] = find_root(parent, ce[1]);
        if (nc == 4)
          for (int i = 0; i < 4; ++i)
            if ((c >> kFace[f][i]) & 1)
              parent[find_root(parent, fe[(i + 3) & 3])] = find_root(parent, fe[i]);
      }
      int polys = 0;
      for (int e = 0; e < 12; ++e)
        if ((crossed >> e & 1) && find_root(parent, e) == e) ++polys;
      nverts[c] = (uint8_t)(3 * (ne - 2 * polys));
    }
  }
};

const CaseTables &case_tables() {
  static CaseTables t;
  return t;
}

struct McArgs {
  const float *vol;
  int gx, gy, gz;          // sample extent per logical axis (corner clamp)
  int64_t sx, sy, sz;      // element strides of logical x, y, z
  int ncx, ncy, ncz;       // cells per axis
  int npx, npy, npz;       // grid points per axis (ncells + 1)
  int64_t npoints;
  int64_t ncls;            // cells with linear index x + gx*(y + gy*z) below this are classified
  int zfast;               // lane order: z fastest (1) or x fastest (0)
  const float *isos;       // ascending
  const int *slot;         // sorted index -> caller's isovalue index
  int k;
  uint32_t *counts;        // per slot: vertices found
  unsigned long long *stats;  // per slot: active cells, total vertices
  float *out;              // vertex pool, xyz triples (all isovalues, any order)
  uint32_t *out_slot;      // per pool vertex: isovalue slot (may be null)
  int64_t cap;             // pool capacity in vertices (0 = count only)
  unsigned long long *corner_tab;
  uint32_t *corner_owner;  // per table slot: id of the candidate that inserted the key
  uint32_t corner_mask;
  uint32_t *overflow;
  uint32_t *block_total;       // count pass: vertices per block
  const uint32_t *block_base;  // emit pass: exclusive scan of block_total
  // vertices per (isovalue of a round, wave) of each block (at most 64 x 5,
  // 16 bits), written by the count pass and read by the emit pass:
  // max_rounds x kMcRound x kMcWaves entries per block
  uint16_t *wave_counts;
  int max_rounds;
  uint32_t *counts_priv;       // [kMcPriv][k] vertices per slot, privatised
  unsigned long long *stats_priv;  // [kMcPriv][2k] active cells, case vertices
  uint8_t nverts[256];
};

constexpr int kMcPriv = 64;   // copies of the per-slot counters (blocks spread over them)
constexpr int kMcAcc = 256;   // per-block LDS accumulators for the first isovalues of its range

__device__ __forceinline__ uint64_t hash64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Count pass: true if this candidate inserted the key (first claim of a
// grid-point vertex); the inserter records its id beside the key.
__device__ __forceinline__ bool corner_claim(const McArgs &a, uint64_t key, uint32_t id) {
  uint32_t h = (uint32_t)hash64(key) & a.corner_mask;
  for (int probe = 0; probe < 128; ++probe) {
    unsigned long long old = atomicCAS(&a.corner_tab[h], kEmpty, (unsigned long long)key);
    if (old == kEmpty) {
      a.corner_owner[h] = id;
      return true;
    }
    if (old == key) return false;
    h = (h + 1) & a.corner_mask;
  }
  atomicAdd(a.overflow, 1u);  // table too full: the host retries with a larger one
  return false;
}

// Emit pass (a later launch): true if this candidate inserted the key.
__device__ __forceinline__ bool corner_owned(const McArgs &a, uint64_t key, uint32_t id) {
  uint32_t h = (uint32_t)hash64(key) & a.corner_mask;
  for (int probe = 0; probe < 128; ++probe) {
    const unsigned long long k = a.corner_tab[h];
    if (k == key) return a.corner_owner[h] == id;
    if (k == kEmpty) return false;
    h = (h + 1) & a.corner_mask;
  }
  return false;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

// number of set bits of m in lanes below this one
__device__ __forceinline__ uint32_t lane_prefix(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ bool cell_ok(const McArgs &a, int cx, int cy, int cz) {
  if (cx < 0 || cy < 0 || cz < 0 || cx >= a.ncx || cy >= a.ncy || cz >= a.ncz) return false;
  int64_t lin = (int64_t)cx + (int64_t)a.gx * ((int64_t)cy + (int64_t)a.gy * cz);
  return lin < a.ncls;
}

// EMIT = false: count pass (claims grid-point vertices, per-block totals,
// per-slot counts and, when STATS, cell statistics).  EMIT = true: the same
// decisions again, grid-point vertices taken by their recorded owner,
// vertices written from the block's scanned base.  No pool counter is shared
// between blocks.
template <bool EMIT, bool STATS>
__global__ __launch_bounds__(kMcBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void mc_kernel(McArgs a) {
  __shared__ uint8_t s_nv[256];
  __shared__ float s_lo[kMcBlock / 64], s_hi[kMcBlock / 64];
  __shared__ int s_range[2];
  // per (isovalue of the round, wave): vertices, active cells, case vertices, offsets
  __shared__ uint32_t s_wt[kMcRound][kMcWaves], s_wa[kMcRound][kMcWaves];
  __shared__ uint32_t s_wv[kMcRound][kMcWaves], s_off[kMcRound][kMcWaves];
  // per isovalue of the block's range (first kMcAcc): vertices, active cells, case vertices
  __shared__ uint32_t s_accv[kMcAcc], s_acca[kMcAcc], s_acct[kMcAcc];
  __shared__ uint32_t s_run;  // vertices of the block so far
  const int tid = threadIdx.x, w = tid >> 6;
  s_nv[tid] = a.nverts[tid];
  if (!EMIT) {
    for (int i = tid; i < kMcAcc; i += kMcBlock) { s_accv[i] = 0u; s_acca[i] = 0u; s_acct[i] = 0u; }
  }
  if (tid == 0) s_run = 0u;

  // A block owns a brick of 16 x 4 x 4 grid points, 16 along the memory-fast
  // axis: a small brick spans a narrower value range than a 256-point line,
  // so fewer isovalues reach it.
  int px, py, pz;
  {
    const int f = tid & 15, m1 = (tid >> 4) & 3, m2 = tid >> 6;
    const int64_t b = blockIdx.x;
    if (a.zfast) {  // z fastest: brick 4 (x) x 4 (y) x 16 (z)
      const int nbz = (a.npz + 15) >> 4, nby = (a.npy + 3) >> 2;
      pz = (int)(b % nbz) * 16 + f;
      py = (int)((b / nbz) % nby) * 4 + m1;
      px = (int)(b / ((int64_t)nbz * nby)) * 4 + m2;
    } else {  // x fastest: brick 16 (x) x 4 (y) x 4 (z)
      const int nbx = (a.npx + 15) >> 4, nby = (a.npy + 3) >> 2;
      px = (int)(b % nbx) * 16 + f;
      py = (int)((b / nbx) % nby) * 4 + m1;
      pz = (int)(b / ((int64_t)nbx * nby)) * 4 + m2;
    }
  }
  const bool live = px < a.npx && py < a.npy && pz < a.npz;
  if (!live) { px = 0; py = 0; pz = 0; }
  // the eight corner samples of the cell at P, reference corner order
  float s[8];
  {
    const int x0 = min(px, a.gx - 1), x1 = min(px + 1, a.gx - 1);
    const int y0 = min(py, a.gy - 1), y1 = min(py + 1, a.gy - 1);
    const int z0 = min(pz, a.gz - 1), z1 = min(pz + 1, a.gz - 1);
    const float *v = a.vol;
    s[0] = v[x0 * a.sx + y0 * a.sy + z0 * a.sz];
    s[1] = v[x1 * a.sx + y0 * a.sy + z0 * a.sz];
    s[2] = v[x1 * a.sx + y1 * a.sy + z0 * a.sz];
    s[3] = v[x0 * a.sx + y1 * a.sy + z0 * a.sz];
    s[4] = v[x0 * a.sx + y0 * a.sy + z1 * a.sz];
    s[5] = v[x1 * a.sx + y0 * a.sy + z1 * a.sz];
    s[6] = v[x1 * a.sx + y1 * a.sy + z1 * a.sz];
    s[7] = v[x0 * a.sx + y1 * a.sy + z1 * a.sz];
  }
  // which cells around P exist and are classified (isovalue independent)
  const bool c000 = live && cell_ok(a, px, py, pz);
  const bool c001 = live && cell_ok(a, px, py, pz - 1);
  const bool c010 = live && cell_ok(a, px, py - 1, pz);
  const bool c011 = live && cell_ok(a, px, py - 1, pz - 1);
  const bool c100 = live && cell_ok(a, px - 1, py, pz);
  const bool c101 = live && cell_ok(a, px - 1, py, pz - 1);
  const bool c110 = live && cell_ok(a, px - 1, py - 1, pz);
  const bool xp = c000 || c001, xm = c010 || c011;   // x edge: + / - orientation present
  const bool yp = c100 || c101, ym = c000 || c001;   // y edge
  const bool zp = c000 || c100 || c010 || c110;      // z edge (always +)
  const bool any = xp || xm || yp || ym || zp;

  float lo = 3.4e38f, hi = -3.4e38f;
  if (any) {
#pragma unroll
    for (int q = 0; q < 8; ++q) { lo = fminf(lo, s[q]); hi = fmaxf(hi, s[q]); }
  }
  // an isovalue can touch this lane only if lane_lo < iso <= lane_hi (an edge
  // crosses iff min(a,b) < iso <= max(a,b); outside, the cell case is 0 or 255)
  const float lane_lo = lo, lane_hi = hi;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo = fminf(lo, __shfl_xor(lo, o, 64));
    hi = fmaxf(hi, __shfl_xor(hi, o, 64));
  }
  if ((tid & 63) == 0) { s_lo[tid >> 6] = lo; s_hi[tid >> 6] = hi; }
  __syncthreads();
  if (tid == 0) {
    float L = s_lo[0], Hh = s_hi[0];
    for (int w = 1; w < kMcBlock / 64; ++w) { L = fminf(L, s_lo[w]); Hh = fmaxf(Hh, s_hi[w]); }
    // isovalues with L < iso <= Hh: [upper_bound(L), upper_bound(Hh))
    int b0 = 0, e0 = a.k;
    while (b0 < e0) { int m = (b0 + e0) >> 1; if (a.isos[m] > L) e0 = m; else b0 = m + 1; }
    int b1 = b0, e1 = a.k;
    while (b1 < e1) { int m = (b1 + e1) >> 1; if (a.isos[m] > Hh) e1 = m; else b1 = m + 1; }
    s_range[0] = b0;
    s_range[1] = L <= Hh ? b1 : b0;
  }
  __syncthreads();
  const int j0 = s_range[0], j1 = s_range[1];
  const float fx = (float)px, fy = (float)py, fz = (float)pz;
  const int lane = tid & 63;
  const uint32_t pid = (uint32_t)(((int64_t)pz * a.npy + py) * a.npx + px);  // candidate ids
  const int priv = (int)(blockIdx.x % kMcPriv);

  const int64_t cap = a.cap;
  float *o = a.out;
  uint32_t *os = a.out_slot;
  const int64_t bbase = EMIT ? (int64_t)a.block_base[blockIdx.x] : 0;
  for (int jc = j0; jc < j1; jc += kMcRound) {
    const int jn = min(kMcRound, j1 - jc);
    uint16_t *wc = a.wave_counts + ((int64_t)blockIdx.x * a.max_rounds + (jc - j0) / kMcRound) *
                                       (kMcRound * kMcWaves);
    if (EMIT) {
      // this round's vertices per (isovalue, wave) from the count pass, then
      // the offsets of (isovalue, wave) within the block
      if (tid < jn * kMcWaves) s_wt[tid / kMcWaves][tid % kMcWaves] = wc[tid];
      __syncthreads();
      if (tid == 0) {
        uint32_t run = s_run;
        for (int u = 0; u < jn; ++u)
          for (int q = 0; q < kMcWaves; ++q) {
            s_off[u][q] = run;
            run += s_wt[u][q];
          }
        s_run = run;
      }
      __syncthreads();
    }
    for (int u = 0; u < jn; ++u) {
      const float iso = a.isos[jc + u];
      if (__ballot(lane_lo < iso && iso <= lane_hi) == 0) {  // nothing for this wave
        if (!EMIT && lane == 0) { s_wt[u][w] = 0u; s_wa[u][w] = 0u; s_wv[u][w] = 0u; }
        continue;
      }
      const int slot = a.slot[jc + u];
      const bool in0 = s[0] < iso;
      // candidate vertices: x+, x-, y+, y-, z+ (value along the edge axis)
      float vxp = 0, vxm = 0, vyp = 0, vym = 0, vzp = 0;
      bool exp_ = false, exm = false, eyp = false, eym = false, ezp = false;
      if ((xp || xm) && in0 != (s[1] < iso)) {
        if (xp) { vxp = fx + (iso - s[0]) / (s[1] - s[0]); exp_ = true; }
        if (xm) { vxm = (fx + 1.0f) - (iso - s[1]) / (s[0] - s[1]); exm = true; }
        if (exp_ && exm && vxp == vxm) exm = false;
      }
      if ((yp || ym) && in0 != (s[3] < iso)) {
        if (yp) { vyp = fy + (iso - s[0]) / (s[3] - s[0]); eyp = true; }
        if (ym) { vym = (fy + 1.0f) - (iso - s[3]) / (s[0] - s[3]); eym = true; }
        if (eyp && eym && vyp == vym) eym = false;
      }
      if (zp && in0 != (s[4] < iso)) { vzp = fz + (iso - s[0]) / (s[4] - s[0]); ezp = true; }
      // grid-point vertices: the first claim wins (count pass), its owner emits
      const uint64_t kslot = (uint64_t)slot << 40;
      auto gp = [&](int gxp, int gyp, int gzp) {
        return kslot | (uint64_t)((int64_t)gxp + (int64_t)a.npx * ((int64_t)gyp + (int64_t)a.npy * gzp));
      };
      auto take = [&](uint64_t key, uint32_t kind) {
        const uint32_t id = pid * 5u + kind;
        return EMIT ? corner_owned(a, key, id) : corner_claim(a, key, id);
      };
      if (exp_ && (vxp == fx || vxp == fx + 1.0f)) exp_ = take(gp((int)vxp, py, pz), 0);
      if (exm && (vxm == fx || vxm == fx + 1.0f)) exm = take(gp((int)vxm, py, pz), 1);
      if (eyp && (vyp == fy || vyp == fy + 1.0f)) eyp = take(gp(px, (int)vyp, pz), 2);
      if (eym && (vym == fy || vym == fy + 1.0f)) eym = take(gp(px, (int)vym, pz), 3);
      if (ezp && (vzp == fz || vzp == fz + 1.0f)) ezp = take(gp(px, py, (int)vzp), 4);
      const uint32_t cnt = (uint32_t)exp_ + exm + eyp + eym + ezp;
      // cnt <= 5 (3 bits) over bit-sliced ballots: the wave's sum (count
      // pass) or this lane's place in the wave's run (emit pass)
      const uint64_t b0 = __ballot(cnt & 1u), b1 = __ballot(cnt & 2u), b2 = __ballot(cnt & 4u);
      if (EMIT) {
        if (cnt) {
          int64_t pos = bbase + s_off[u][w] + lane_prefix(b0) + 2 * lane_prefix(b1) +
                        4 * lane_prefix(b2);
          auto put = [&](float x, float y, float z) {
            if (pos < cap) {
              o[3 * pos] = x;
              o[3 * pos + 1] = y;
              o[3 * pos + 2] = z;
              if (os) os[pos] = (uint32_t)slot;
            }
            ++pos;
          };
          if (exp_) put(vxp, fy, fz);
          if (exm) put(vxm, fy, fz);
          if (eyp) put(fx, vyp, fz);
          if (eym) put(fx, vym, fz);
          if (ezp) put(fx, fy, vzp);
        }
        continue;
      }
      const uint32_t cnt_w = __popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2);
      if (lane == 0) s_wt[u][w] = cnt_w;
      if (!EMIT && STATS) {
        // cell statistics of the cell at P
        uint32_t act = 0, tot = 0;
        if (c000) {
          uint32_t ci = 0;
#pragma unroll
          for (int q = 0; q < 8; ++q) ci |= (uint32_t)(s[q] < iso) << q;
          tot = s_nv[ci];
          act = tot > 0;
        }
        const uint32_t act_w = __popcll(__ballot(act != 0));
        const uint32_t tot_w = __popcll(__ballot(tot & 1u)) + 2 * __popcll(__ballot(tot & 2u)) +
                               4 * __popcll(__ballot(tot & 4u)) + 8 * __popcll(__ballot(tot & 8u));
        if (lane == 0) { s_wa[u][w] = act_w; s_wv[u][w] = tot_w; }
      }
    }
    __syncthreads();
    if (!EMIT) {
      if (tid < jn * kMcWaves) wc[tid] = (uint16_t)s_wt[tid / kMcWaves][tid % kMcWaves];
      // one lane per isovalue of the round folds the wave sums into the block's
      if (tid < jn) {
        uint32_t vj = 0, aj = 0, tj = 0;
#pragma unroll
        for (int q = 0; q < kMcWaves; ++q) {
          vj += s_wt[tid][q];
          if (STATS) { aj += s_wa[tid][q]; tj += s_wv[tid][q]; }
        }
        const int r = jc + tid - j0;
        if (r < kMcAcc) {
          s_accv[r] += vj;
          if (STATS) { s_acca[r] += aj; s_acct[r] += tj; }
        } else if (vj || aj) {  // past the LDS accumulators: straight to the counters
          const int slot = a.slot[jc + tid];
          atomicAdd(&a.counts_priv[(int64_t)priv * a.k + slot], vj);
          if (STATS) {
            atomicAdd(&a.stats_priv[((int64_t)priv * a.k + slot) * 2], (unsigned long long)aj);
            atomicAdd(&a.stats_priv[((int64_t)priv * a.k + slot) * 2 + 1], (unsigned long long)tj);
          }
        }
        if (vj) atomicAdd(&s_run, vj);
      }
      __syncthreads();  // s_wt is rewritten by the next round
    }
    // (emit pass: the barrier above ends the round's reads of s_off and s_wt)
  }
  if (!EMIT) {
    __syncthreads();
    if (tid == 0) a.block_total[blockIdx.x] = s_run;
    const int nacc = min(j1 - j0, kMcAcc);
    for (int r = tid; r < nacc; r += kMcBlock) {
      const uint32_t vj = s_accv[r], aj = s_acca[r], tj = s_acct[r];
      if (vj || aj) {
        const int slot = a.slot[j0 + r];
        atomicAdd(&a.counts_priv[(int64_t)priv * a.k + slot], vj);
        if (STATS) {
          atomicAdd(&a.stats_priv[((int64_t)priv * a.k + slot) * 2], (unsigned long long)aj);
          atomicAdd(&a.stats_priv[((int64_t)priv * a.k + slot) * 2 + 1], (unsigned long long)tj);
        }
      }
    }
  }
}

// Per-isovalue totals from the privatised counters (counts[s] = the sum of
// the kMcPriv copies; stats likewise unless stats is null), and their 64-bit
// sum into *total64 (which the 32-bit vertex offsets are checked against).
__global__ void mc_reduce_kernel(const uint32_t *counts_priv, const unsigned long long *stats_priv,
                                 int k, uint32_t *counts, unsigned long long *stats,
                                 unsigned long long *total64) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= k) return;
  uint32_t c = 0;
  unsigned long long a0 = 0, a1 = 0;
  for (int p = 0; p < kMcPriv; ++p) {
    c += counts_priv[(int64_t)p * k + s];
    if (stats) {
      a0 += stats_priv[((int64_t)p * k + s) * 2];
      a1 += stats_priv[((int64_t)p * k + s) * 2 + 1];
    }
  }
  counts[s] = c;
  atomicAdd(total64, (unsigned long long)c);
  if (stats) {
    stats[2 * s] = a0;
    stats[2 * s + 1] = a1;
  }
}

// ---- voxel approximation ----------------------------------------------------
// Cell (x,y,z), x<X-1, y<Y-1, z<Z-1 of the z-fastest image is emitted iff one
// corner is > iso and one is < iso (IsosurfaceApproximation.cu:67-135);
// output order is the cell order x-major / z-fastest (stable copy_if).
constexpr int kApBlock = 256;
constexpr int kApItems = 4;
constexpr int kApTile = kApBlock * kApItems;

struct ApArgs {
  const float *img;
  int X, Y, Z;
  int64_t ncells;
  float iso;
  uint32_t *tile_counts;
  const uint32_t *tile_offsets;
  int32_t *out;
  int64_t cap;
};

__device__ __forceinline__ bool approx_flag(const ApArgs &a, int64_t o) {
  const int64_t Zc = a.Z - 1, Yc = a.Y - 1;
  const int z = (int)(o % Zc);
  const int y = (int)((o / Zc) % Yc);
  const int x = (int)(o / (Zc * Yc));
  const int64_t YZ = (int64_t)a.Y * a.Z;
  const int64_t lu = x * YZ + (int64_t)y * a.Z + z;
  const float *p = a.img + lu;
  const float c[8] = {p[0], p[1], p[a.Z], p[a.Z + 1], p[YZ], p[YZ + 1], p[YZ + a.Z], p[YZ + a.Z + 1]};
  bool above = false, below = false;
#pragma unroll
  for (int q = 0; q < 8; ++q) { above |= c[q] > a.iso; below |= c[q] < a.iso; }
  return above && below;
}

// tile t covers cells [t*kApTile, (t+1)*kApTile), item q of lane l is cell
// t*kApTile + l*kApItems + q (contiguous per lane, so tile order = cell order)
__global__ __launch_bounds__(kApBlock) void approx_count_kernel(ApArgs a) {
  __shared__ uint32_t s_w[kApBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kApTile + (int64_t)threadIdx.x * kApItems;
  uint32_t c = 0;
#pragma unroll
  for (int q = 0; q < kApItems; ++q)
    if (base + q < a.ncells) c += approx_flag(a, base + q);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < kApBlock / 64; ++w) t += s_w[w];
    a.tile_counts[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kApBlock) void approx_emit_kernel(ApArgs a) {
  __shared__ uint32_t s_w[kApBlock / 64];
  const int64_t base = (int64_t)blockIdx.x * kApTile + (int64_t)threadIdx.x * kApItems;
  bool f[kApItems];
  uint32_t c = 0;
#pragma unroll
  for (int q = 0; q < kApItems; ++q) {
    f[q] = base + q < a.ncells && approx_flag(a, base + q);
    c += f[q];
  }
  const uint32_t inc = wave_incl_scan(c);
  if ((threadIdx.x & 63) == 63) s_w[threadIdx.x >> 6] = inc;
  __syncthreads();
  uint32_t wpre = 0;
  for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) wpre += s_w[w];
  int64_t pos = (int64_t)a.tile_offsets[blockIdx.x] + wpre + inc - c;
  const int64_t Zc = a.Z - 1, Yc = a.Y - 1;
#pragma unroll
  for (int q = 0; q < kApItems; ++q) {
    if (f[q]) {
      if (pos < a.cap) {
        const int64_t o = base + q;
        a.out[3 * pos] = (int32_t)(o / (Zc * Yc));
        a.out[3 * pos + 1] = (int32_t)((o / Zc) % Yc);
        a.out[3 * pos + 2] = (int32_t)(o % Zc);
      }
      ++pos;
    }
  }
}

// Many isovalues per pass (unordered pool output, for the pipeline): lane =
// cell of a 4 x 4 x 16 brick (z fastest), block-culled loop over the
// isovalues strictly inside the block's [min, max] in rounds of 32: pass 1
// counts per wave, one pool reservation per block and round, pass 2 writes.
struct ApMultiArgs {
  const float *img;
  int X, Y, Z;
  const float *isos;  // ascending
  const int *slot;
  int k;
  uint32_t *counts;
  int32_t *out;        // pool of xyz cells
  uint32_t *out_slot;  // per pool entry: isovalue slot
  int64_t cap;
  unsigned long long *pool_count;
};

constexpr int kApWaves = kApBlock / 64;
constexpr int kApRound = 32;

__global__ __launch_bounds__(kApBlock) void approx_multi_kernel(ApMultiArgs a) {
  __shared__ float s_lo[kApWaves], s_hi[kApWaves];
  __shared__ int s_range[2];
  __shared__ uint32_t s_wt[kApRound][kApWaves], s_off[kApRound][kApWaves];
  __shared__ unsigned long long s_base;
  __shared__ int s_any;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int Xc = a.X - 1, Yc = a.Y - 1, Zc = a.Z - 1;
  int x, y, z;
  {
    const int nbz = (Zc + 15) >> 4, nby = (Yc + 3) >> 2;
    const int64_t b = blockIdx.x;
    z = (int)(b % nbz) * 16 + (tid & 15);
    y = (int)((b / nbz) % nby) * 4 + ((tid >> 4) & 3);
    x = (int)(b / ((int64_t)nbz * nby)) * 4 + (tid >> 6);
  }
  const bool live = x < Xc && y < Yc && z < Zc;
  if (!live) { x = 0; y = 0; z = 0; }
  const int64_t YZ = (int64_t)a.Y * a.Z;
  const float *p = a.img + (x * YZ + (int64_t)y * a.Z + z);
  float lo = 3.4e38f, hi = -3.4e38f;
  if (live) {
    const float c[8] = {p[0], p[1], p[a.Z], p[a.Z + 1], p[YZ], p[YZ + 1], p[YZ + a.Z], p[YZ + a.Z + 1]};
#pragma unroll
    for (int q = 0; q < 8; ++q) { lo = fminf(lo, c[q]); hi = fmaxf(hi, c[q]); }
  }
  float blo = lo, bhi = hi;
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) {
    blo = fminf(blo, __shfl_xor(blo, s, 64));
    bhi = fmaxf(bhi, __shfl_xor(bhi, s, 64));
  }
  if (lane == 0) { s_lo[w] = blo; s_hi[w] = bhi; }
  __syncthreads();
  if (tid == 0) {
    float L = s_lo[0], Hh = s_hi[0];
    for (int q = 1; q < kApWaves; ++q) { L = fminf(L, s_lo[q]); Hh = fmaxf(Hh, s_hi[q]); }
    int b0 = 0, e0 = a.k;  // first iso > L
    while (b0 < e0) { int m = (b0 + e0) >> 1; if (a.isos[m] > L) e0 = m; else b0 = m + 1; }
    int b1 = b0, e1 = a.k;  // first iso >= Hh
    while (b1 < e1) { int m = (b1 + e1) >> 1; if (a.isos[m] >= Hh) e1 = m; else b1 = m + 1; }
    s_range[0] = b0;
    s_range[1] = L < Hh ? b1 : b0;
  }
  __syncthreads();
  const int j0 = s_range[0], j1 = s_range[1];
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (int jc = j0; jc < j1; jc += kApRound) {
    const int jn = min(kApRound, j1 - jc);
    for (int u = 0; u < jn; ++u) {
      const float iso = a.isos[jc + u];
      const uint64_t m = __ballot(lo < iso && iso < hi);  // a corner below and one above
      if (lane == 0) s_wt[u][w] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t run = 0;
      for (int u = 0; u < jn; ++u) {
        uint32_t vj = 0;
        for (int q = 0; q < kApWaves; ++q) {
          s_off[u][q] = run + vj;
          vj += s_wt[u][q];
        }
        run += vj;
        if (vj) atomicAdd(&a.counts[a.slot[jc + u]], vj);
      }
      s_base = run ? atomicAdd(a.pool_count, (unsigned long long)run) : 0ull;
      s_any = run != 0;
    }
    __syncthreads();
    if (!s_any) continue;  // block-uniform
    for (int u = 0; u < jn; ++u) {
      const float iso = a.isos[jc + u];
      const bool f = lo < iso && iso < hi;
      const uint64_t m = __ballot(f);
      if (f) {
        const int64_t pos = (int64_t)s_base + s_off[u][w] + __popcll(m & lt);
        if (pos < a.cap) {
          int32_t *q = a.out + 3 * pos;
          q[0] = x;
          q[1] = y;
          q[2] = z;
          if (a.out_slot) a.out_slot[pos] = (uint32_t)a.slot[jc + u];
        }
      }
    }
  }
}

}  // namespace

namespace iso {

int sort_vertices_lex(iso_ctx *ctx, float *d_xyz, int64_t n);

// Voxel approximation at k isovalues in one pass, pool output like mc_multi.
int approx_multi(iso_ctx *ctx, const float *d_img, int X, int Y, int Z, const float *d_isos_sorted,
                 const int *d_slot, int k, uint32_t *d_counts, int32_t *d_pool,
                 uint32_t *d_pool_slot, int64_t cap, int64_t *h_total) {
  *h_total = 0;
  ISO_TRY(fill_u32(ctx, d_counts, 0u, (size_t)k));
  if (X < 2 || Y < 2 || Z < 2 || k <= 0) return ISO_OK;
  if ((int64_t)(X - 1) * (Y - 1) * (Z - 1) > (int64_t)UINT32_MAX)
    return fail(ctx, ISO_ERR_UNSUPPORTED, "voxel approximation: more than 2^32-1 cells");
  void *ctr = nullptr, *pin = nullptr;
  ISO_TRY(ws_get(ctx, WS_MC_COUNTERS, 64, false, &ctr));
  ISO_TRY(fill_u32(ctx, ctr, 0u, 4));
  ApMultiArgs a;
  a.img = d_img;
  a.X = X;
  a.Y = Y;
  a.Z = Z;
  a.isos = d_isos_sorted;
  a.slot = d_slot;
  a.k = k;
  a.counts = d_counts;
  a.out = d_pool;
  a.out_slot = d_pool_slot;
  a.cap = d_pool ? cap : 0;
  a.pool_count = (unsigned long long *)ctr + 1;
  const int64_t blocks =
      (int64_t)((X - 1 + 3) / 4) * ((Y - 1 + 3) / 4) * ((Z - 1 + 15) / 16);
  hipLaunchKernelGGL(approx_multi_kernel, dim3((unsigned)blocks), dim3(kApBlock), 0, ctx->stream,
                     a);
  ISO_HIP_CHECK(ctx, hipGetLastError());
  ISO_TRY(pinned_get(ctx, 16, &pin));
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(pin, ctr, 16, hipMemcpyDeviceToHost, ctx->stream));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *h_total = (int64_t)((unsigned long long *)pin)[1];
  return ISO_OK;
}

// Marching cubes at k isovalues in one pass over the volume.  Every vertex
// goes to one pool (xyz, and its isovalue slot if d_pool_slot != null);
// counts/stats are device arrays indexed by slot (stats may be null: no cell
// statistics are then gathered); d_isos_sorted/d_slot give
// the ascending isovalues and their slots.  *h_total = vertices found, which
// may exceed cap (nothing past cap is written); h_counts (k words, may be
// null) receives the per-slot counts.  The vertex-writing pass is left
// running on the stream when this returns.  Synchronises the stream and
// retries when the grid-point hash set fills up.
int mc_multi(iso_ctx *ctx, const float *d_vol, int gx, int gy, int gz, uint32_t flags,
             const float *d_isos_sorted, const int *d_slot, int k, uint32_t *d_counts,
             unsigned long long *d_stats, float *d_pool, uint32_t *d_pool_slot, int64_t cap,
             int64_t *h_total, uint32_t *h_counts) {
  *h_total = 0;
  McArgs a;
  a.vol = d_vol;
  a.gx = gx;
  a.gy = gy;
  a.gz = gz;
  const bool compat = flags & ISO_COMPAT_MC;
  if (compat) {
    a.sx = 1;
    a.sy = gx;
    a.sz = (int64_t)gx * gy;
    a.ncx = gx;
    a.ncy = gy;
    a.ncz = gz;
    int64_t n = (int64_t)gx * gy * gz;
    int64_t bx = n / 128, by = 1;  // MarchingCube.cpp:599-608
    if (bx > 65535) { by = bx / 32768; bx = 32768; }
    a.ncls = std::min<int64_t>(128 * bx * by, n);
    a.zfast = 0;
  } else {
    a.sx = (int64_t)gy * gz;
    a.sy = gz;
    a.sz = 1;
    a.ncx = gx - 1;
    a.ncy = gy - 1;
    a.ncz = gz - 1;
    a.ncls = INT64_MAX;
    a.zfast = 1;
  }
  a.npx = a.ncx + 1;
  a.npy = a.ncy + 1;
  a.npz = a.ncz + 1;
  a.npoints = (int64_t)a.npx * a.npy * a.npz;
  // candidate vertex ids are 32-bit: point id * 5 + edge kind
  if (a.npoints > (int64_t)(UINT32_MAX / 5u))
    return fail(ctx, ISO_ERR_UNSUPPORTED,
                "marching cubes: %lld grid points, at most %u (32-bit candidate vertex ids)",
                (long long)a.npoints, UINT32_MAX / 5u);
  a.isos = d_isos_sorted;
  a.slot = d_slot;
  a.k = k;
  a.counts = d_counts;
  a.stats = d_stats;
  a.out = d_pool;
  a.out_slot = d_pool_slot;
  a.cap = d_pool ? cap : 0;
  memcpy(a.nverts, case_tables().nverts, 256);
  if (a.ncx <= 0 || a.ncy <= 0 || a.ncz <= 0 || k <= 0) {
    ISO_TRY(fill_u32(ctx, d_counts, 0u, (size_t)k));
    if (d_stats) ISO_TRY(fill_u32(ctx, d_stats, 0u, (size_t)4 * k));
    return ISO_OK;
  }
  const int64_t fast = a.zfast ? a.npz : a.npx, slow = a.zfast ? a.npx : a.npz;
  const int64_t blocks = ((fast + 15) / 16) * ((a.npy + 3) / 4) * ((slow + 3) / 4);
  void *blk = nullptr, *priv = nullptr, *tmp = nullptr;
  ISO_TRY(ws_get(ctx, WS_MC_BLOCKS, sizeof(uint32_t) * 2 * (size_t)(blocks + 1), false, &blk));
  // the count pass's vertices per (round, isovalue, wave) of each block, which
  // the emit pass starts each round from (every entry it reads is written)
  a.max_rounds = (k + kMcRound - 1) / kMcRound;
  void *wcb = nullptr;
  ISO_TRY(ws_get(ctx, WS_MC_WAVES,
                 sizeof(uint16_t) * (size_t)blocks * a.max_rounds * kMcRound * kMcWaves, false,
                 &wcb));
  a.wave_counts = (uint16_t *)wcb;
  ISO_TRY(ws_get(ctx, WS_MC_PRIV, (sizeof(uint32_t) + 2 * sizeof(unsigned long long)) * kMcPriv * (size_t)k,
                 false, &priv));
  ISO_TRY(ws_get(ctx, WS_SCAN_TMP, sizeof(uint32_t) * scan_tmp_words(blocks), false, &tmp));
  a.block_total = (uint32_t *)blk;
  a.block_base = (const uint32_t *)blk + (blocks + 1);
  a.stats_priv = (unsigned long long *)priv;
  a.counts_priv = (uint32_t *)(a.stats_priv + 2 * (size_t)kMcPriv * k);
  uint32_t log2cap = 16;
  for (int attempt = 0; attempt < 8; ++attempt, log2cap += 2) {
    const size_t slots = (size_t)1 << log2cap;
    void *tab = nullptr, *ctr = nullptr, *pin = nullptr;
    ISO_TRY(ws_get(ctx, WS_MC_CORNER_TABLE, slots * 12, false, &tab));  // keys, then owners
    ISO_TRY(ws_get(ctx, WS_MC_COUNTERS, 64, false, &ctr));
    ISO_TRY(fill_u32(ctx, tab, 0xffffffffu, slots * 2));
    ISO_TRY(fill_u32(ctx, ctr, 0u, 4));  // [overflow, vertex total, 64-bit vertex total]
    ISO_TRY(fill_u32(ctx, priv, 0u, (size_t)kMcPriv * k * 5));
    a.corner_tab = (unsigned long long *)tab;
    a.corner_owner = (uint32_t *)(a.corner_tab + slots);
    a.corner_mask = (uint32_t)(slots - 1);
    a.overflow = (uint32_t *)ctr;
    // cell statistics only when the caller takes them
    if (d_stats)
      hipLaunchKernelGGL((mc_kernel<false, true>), dim3((unsigned)blocks), dim3(kMcBlock), 0,
                         ctx->stream, a);
    else
      hipLaunchKernelGGL((mc_kernel<false, false>), dim3((unsigned)blocks), dim3(kMcBlock), 0,
                         ctx->stream, a);
    ISO_HIP_CHECK(ctx, hipGetLastError());
    ISO_TRY(exclusive_scan_u32(ctx, a.block_total, (uint32_t *)a.block_base, blocks,
                               (uint32_t *)ctr + 1, (uint32_t *)tmp));
    hipLaunchKernelGGL(mc_reduce_kernel, dim3((unsigned)((k + 255) / 256)), dim3(256), 0,
                       ctx->stream, (const uint32_t *)a.counts_priv,
                       (const unsigned long long *)a.stats_priv, k, d_counts, d_stats,
                       (unsigned long long *)ctr + 1);
    // the counters, and the per-slot counts when the caller takes them (they
    // are final after the count pass), in one read-back
    ISO_TRY(pinned_get(ctx, 16 + (h_counts ? 4 * (size_t)k : 0), &pin));
    ISO_HIP_CHECK(ctx, hipMemcpyAsync(pin, ctr, 16, hipMemcpyDeviceToHost, ctx->stream));
    if (h_counts)
      ISO_HIP_CHECK(ctx, hipMemcpyAsync((char *)pin + 16, d_counts, 4 * (size_t)k,
                                        hipMemcpyDeviceToHost, ctx->stream));
    ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (((uint32_t *)pin)[0] != 0u) continue;  // grid-point table full: again, larger
    if (h_counts) memcpy(h_counts, (const char *)pin + 16, 4 * (size_t)k);
    *h_total = (int64_t)((uint32_t *)pin)[1];
    // the vertex offsets are a 32-bit scan: its total must be the exact one
    if ((unsigned long long)*h_total != ((unsigned long long *)pin)[1])
      return fail(ctx, ISO_ERR_UNSUPPORTED,
                  "marching cubes: %llu vertices over %d isovalues, at most %u per call",
                  ((unsigned long long *)pin)[1], k, UINT32_MAX);
    if (a.cap > 0 && *h_total <= a.cap && *h_total > 0)
      hipLaunchKernelGGL((mc_kernel<true, false>), dim3((unsigned)blocks), dim3(kMcBlock), 0,
                         ctx->stream,
                         a);
    ISO_HIP_CHECK(ctx, hipGetLastError());
    return ISO_OK;
  }
  return fail(ctx, ISO_ERR_OUT_OF_MEMORY, "marching cubes: grid-point vertex table overflow");
}

int approx_one(iso_ctx *ctx, const float *d_img, int X, int Y, int Z, float iso, int32_t *d_out,
               int64_t cap, int64_t *count) {
  *count = 0;
  if (X < 2 || Y < 2 || Z < 2) return ISO_OK;
  if ((int64_t)(X - 1) * (Y - 1) * (Z - 1) > (int64_t)UINT32_MAX)
    return fail(ctx, ISO_ERR_UNSUPPORTED, "voxel approximation: more than 2^32-1 cells");
  ApArgs a;
  a.img = d_img;
  a.X = X;
  a.Y = Y;
  a.Z = Z;
  a.ncells = (int64_t)(X - 1) * (Y - 1) * (Z - 1);
  a.iso = iso;
  a.out = d_out;
  a.cap = cap;
  const int64_t tiles = (a.ncells + kApTile - 1) / kApTile;
  void *cnt = nullptr, *tmp = nullptr;
  ISO_TRY(ws_get(ctx, WS_MC_KEYS_A, sizeof(uint32_t) * (tiles + 1), false, &cnt));
  ISO_TRY(ws_get(ctx, WS_SCAN_TMP, sizeof(uint32_t) * scan_tmp_words(tiles), false, &tmp));
  a.tile_counts = (uint32_t *)cnt;
  a.tile_offsets = (uint32_t *)cnt;
  hipLaunchKernelGGL(approx_count_kernel, dim3((unsigned)tiles), dim3(kApBlock), 0, ctx->stream, a);
  ISO_TRY(exclusive_scan_u32(ctx, (uint32_t *)cnt, (uint32_t *)cnt, tiles, (uint32_t *)cnt + tiles,
                             (uint32_t *)tmp));
  hipLaunchKernelGGL(approx_emit_kernel, dim3((unsigned)tiles), dim3(kApBlock), 0, ctx->stream, a);
  ISO_HIP_CHECK(ctx, hipGetLastError());
  uint32_t total = 0;
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(&total, (uint32_t *)cnt + tiles, 4, hipMemcpyDeviceToHost,
                                    ctx->stream));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *count = total;
  return ISO_OK;
}

}  // namespace iso

extern "C" {

int iso_mc_extract(iso_ctx *ctx, const float *d_vol, int gx, int gy, int gz, float iso,
                   uint32_t flags, float *d_xyz, int64_t cap, int64_t *h_count,
                   iso_mc_stats *h_stats) {
  ISO_TRY(iso::enter(ctx));
  if (!d_vol || gx <= 0 || gy <= 0 || gz <= 0 || !h_count || cap < 0 || (cap > 0 && !d_xyz))
    return iso::fail(ctx, ISO_ERR_INVALID_ARG, "iso_mc_extract: bad arguments");
  // one isovalue: small device block with [iso, slot, count, stats]
  struct Small {
    float iso;
    int slot;
    uint32_t count;
    uint32_t pad;
    unsigned long long stats[2];
  };
  void *dsm = nullptr, *pin = nullptr;
  ISO_TRY(iso::ws_get(ctx, WS_SMALL, sizeof(Small), false, &dsm));
  ISO_TRY(iso::pinned_get(ctx, 64, &pin));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  Small hs{};
  hs.iso = iso;
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(dsm, &hs, sizeof(Small), hipMemcpyHostToDevice, ctx->stream));
  Small *ds = (Small *)dsm;
  int64_t total = 0;
  ISO_TRY(iso::mc_multi(ctx, d_vol, gx, gy, gz, flags, &ds->iso, &ds->slot, 1, &ds->count,
                        ds->stats, d_xyz, nullptr, d_xyz ? cap : 0, &total, nullptr));
  ISO_HIP_CHECK(ctx, hipMemcpyAsync(&hs, dsm, sizeof(Small), hipMemcpyDeviceToHost, ctx->stream));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  *h_count = hs.count;
  if (h_stats) {
    h_stats->active_cells = (int64_t)hs.stats[0];
    h_stats->total_verts = (int64_t)hs.stats[1];
    h_stats->triangles = (int64_t)hs.stats[1] / 3;
    h_stats->unique_verts = hs.count;
  }
  if ((int64_t)hs.count > cap)
    return iso::fail(ctx, ISO_ERR_CAPACITY, "iso_mc_extract: %lld vertices, capacity %lld",
                     (long long)hs.count, (long long)cap);
  if (flags & ISO_MC_SORTED) ISO_TRY(iso::sort_vertices_lex(ctx, d_xyz, hs.count));
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return ISO_OK;
}

int iso_approx_extract(iso_ctx *ctx, const float *d_img, int X, int Y, int Z, float iso,
                       int32_t *d_xyz, int64_t cap, int64_t *h_count) {
  ISO_TRY(iso::enter(ctx));
  if (!d_img || X <= 0 || Y <= 0 || Z <= 0 || !h_count || cap < 0 || (cap > 0 && !d_xyz))
    return iso::fail(ctx, ISO_ERR_INVALID_ARG, "iso_approx_extract: bad arguments");
  int64_t n = 0;
  ISO_TRY(iso::approx_one(ctx, d_img, X, Y, Z, iso, d_xyz, cap, &n));
  *h_count = n;
  if (n > cap) return iso::fail(ctx, ISO_ERR_CAPACITY, "iso_approx_extract: %lld cells, capacity %lld", (long long)n, (long long)cap);
  return ISO_OK;
}

}  // extern "C"
