// Normalised mutual information in the reference's float order, evaluated by
// a histogram workgroup once its pair's joint counts are complete.
//
// Reference: calculateMutualInformation (isosurfaces_cuda/src/JointHistogram.cu:
// 24-60), run per ordered pair (i, j) on the device:
//   hXY -= c * logf(c) over the nonzero bins, rows outer, columns inner;
//   hX  -= c * logf(c) over colSums[0..H), hY likewise over rowSums;
//   h = h / N + logf(N);  NMI = 2 (hX + hY - hXY) / (hX + hY), NaN -> 0.
// Each `-=` is one fused multiply-add (nvcc contracts a - b*c), so every
// accumulator is a chain of fmaf roundings in a fixed order; the result
// depends on that order and on nothing else.  The map entry (j, i) of the
// same unordered pair comes from the transposed histogram: its hXY chain
// visits the bins column by column, and its hX, hY are this pair's hY, hX
// (their sum is the same float).
//
// On the GPU the chains are serial (one lane each) but short: only the
// nonzero bins take part (about 1,500 of 65,536 for the 600x248x248 front at
// 256 bins).  The whole workgroup first writes the nonzero counts in chain
// order to a staging area in global memory; positions come from counts of
// nonzero bins: a row-major position is the row's base (exclusive prefix of
// the nonzero bins per row) plus the nonzero bins before it in its row (a
// ballot prefix across the wave); a column-major position is the column's
// base plus the nonzero bins above it in its column (per-wave column counts
// and their prefix over the waves before).  Then the lists come back into
// LDS window by window (coalesced), each lane taking the logarithms of the
// counts it brings in, and one lane per chain folds the terms (c, logf(c)).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace iso {

constexpr int kNmiBlock = 1024;  // lanes of the histogram workgroups

// logf, correctly rounded for the counts it is applied to: the double
// logarithm of the value, rounded once to float.  For every integer from 1
// to 16,200,281 the exact logarithm lies at least 2^-48 (relative) away from
// a float rounding boundary (checked exhaustively on the CPU), far more than
// the error of a double logarithm, so any double log rounds to the one
// correct float.  Counts and field sizes stay below that bound unless a
// field has more than 16.2 million samples.
__device__ __forceinline__ float logf_cr(float x) { return (float)log((double)x); }

// number of set bits of m in the lanes below this one
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// LDS the float-order NMI uses beside the histogram kernels' own.
struct Nmi32Lds {
  uint8_t colpre[16][256];  // per wave: nonzero bins of column j in its 16 rows,
                            // then those in the rows of the waves before it
  uint32_t rowbase[256];    // nonzero bins per row, then their exclusive prefix
  uint32_t colbase[256];    // exclusive prefix of the nonzero bins per column
  // the marginals' chain terms (c, logf(c)), (0, 0) where c == 0
  alignas(16) float2 mrow[256];
  alignas(16) float2 mcol[256];
  float2 mslack[16];  // read past mcol by the chain fold, never folded
  uint32_t wtot[8];
  float acc[4];
};

// Staging of one workgroup: 2 * H * H float counts (row-major list, then
// column-major list).
__host__ __device__ constexpr size_t nmi32_stage_bytes(int H) {
  return (size_t)2 * H * H * sizeof(float);
}

// Terms per block of the chain fold: the lists in LDS are padded with (0, 0)
// to a multiple of it, and fmaf(-0, 0, acc) is acc exactly.
constexpr int kChainBlock = 16;

// acc = fmaf(-c, t, acc) over the terms (c, t) = v[0], ..., v[n-1] in order.
// v is 16-byte aligned; from n up to the next multiple of kChainBlock it
// holds (0, 0), and the kChainBlock entries after that are readable (read,
// never folded).  One lane runs it: blocks of 2 Q terms in two register
// buffers of Q float4, each block's 16-byte LDS reads issued one block
// ahead of its fold, no predication: per term one fma and half a read.
template <int Q>
__device__ __forceinline__ float fma_chain(const float2 *v, int n, float acc) {
  static_assert(kChainBlock % (2 * Q) == 0, "block");
  const float4 *p = reinterpret_cast<const float4 *>(v);
  const int nq = ((n + kChainBlock - 1) / kChainBlock) * (kChainBlock / 2);
  if (nq == 0) return acc;
  float4 a[Q], b[Q];
  auto load = [&](float4(&r)[Q], int q0) {
#pragma unroll
    for (int u = 0; u < Q; ++u) r[u] = p[q0 + u];
  };
  auto fold = [&](const float4(&r)[Q]) {
#pragma unroll
    for (int u = 0; u < Q; ++u) {
      acc = fmaf(-r[u].x, r[u].y, acc);
      acc = fmaf(-r[u].z, r[u].w, acc);
    }
  };
  // reads stay below nq + Q float4: nq + 2 Q terms
  load(a, 0);
  load(b, Q);
  for (int q = 0;; q += 2 * Q) {
    fold(a);
    if (q + Q >= nq) break;
    load(a, q + 2 * Q);
    fold(b);
    if (q + 2 * Q >= nq) break;
    load(b, q + 3 * Q);
  }
  return acc;
}

// Bins a lane covers: lane (w, l) of the 16 waves holds bins (RW w + r,
// CL l + c), r < RW, c < CL, so 16 RW rows and 64 CL columns must cover H
// (RW 16, CL 4 up to 256 bins; RW 8, CL 2 up to 128, a quarter of the
// registers).
template <int RW, int CL>
struct Nmi32Tile {
  static_assert(RW * CL <= 64 && 16 * RW <= 256 && 64 * CL <= 256, "tile");
  static constexpr int kBits = RW * CL;
  // bits c, c + CL, c + 2 CL, ... of a lane's mask: column c of its tile
  static constexpr uint64_t col_mask(int c) {
    uint64_t m = 0;
    for (int r = 0; r < RW; ++r) m |= 1ull << (CL * r + c);
    return m;
  }
  static constexpr uint64_t kRowBits = (1ull << CL) - 1ull;
};

// The pair's joint counts are count_at(e), e = row * H + col (row from the
// first field), H <= 16 RW and H <= 64 CL; with KEEP and the 64-bin tiles
// they are read instead from `packed`, the LDS histogram of 16-bit counts two
// to a word (H even).  KEEP: the counts stay in the lanes' registers from the
// structure pass to the placement.  take_at(e) returns the count of a nonzero
// bin; it is called once per nonzero bin, after every read of the counts,
// where they are not kept or need_take is set (it may restore a zero behind
// it).  Called by every lane of a 1024-lane workgroup, with uniform flags.
// s_rows / s_cols: 256 zeroed LDS words, left holding the marginals (row
// sums, column sums).  win: LDS room for 2 * cap float2, 16-byte aligned,
// cap >= 2 kChainBlock, that no lane reads once every lane has passed the
// first barrier in here.  With KEEP, when both chain lists fit there, they
// are built in it directly; otherwise claim_stage() (called by every lane)
// returns a global staging area of nmi32_stage_bytes(H) for them, and
// release_stage() (every lane) hands it back after its last read.  Lane 0
// writes the map entry (i, j) to *out_ij and, with mirror, the entry (j, i)
// to *out_ji.  N is the field size.
template <int RW, int CL, bool KEEP, class CountAt, class TakeAt, class Claim, class Release>
__device__ __attribute__((always_inline)) void nmi32_pair(
    CountAt count_at, const uint32_t *packed, TakeAt take_at, bool need_take, int H, int64_t N,
    bool mirror, uint32_t *s_rows, uint32_t *s_cols, Claim claim_stage, Release release_stage,
    float2 *win, int cap, Nmi32Lds &s, float *out_ij, float *out_ji) {
  using T = Nmi32Tile<RW, CL>;
  static_assert(RW == 8 || RW == 16, "rows per wave");
  constexpr int kLogRW = RW == 16 ? 4 : 3;
  // chain buffers: 2 x 8 float4 with the 256-bin tiles, 2 x 4 float4 with
  // the smaller ones (whose kernels fit more workgroups per CU)
  constexpr int CQ = RW >= 16 ? 8 : 4;
  // KEEP: a lane's counts are kept from the structure pass to the placement,
  // two 16-bit counts per register for the 64-bin tiles, one per register
  // for the smaller ones
  constexpr bool kPack = RW * CL > 16;
  constexpr int kRegs = KEEP ? (kPack ? RW * CL / 2 : RW * CL) : 1;
  // the lane id made opaque: inside a persistent task loop, addresses derived
  // from it must not be hoisted out of the loop (they would stay live across
  // the histogram phase and spill)
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int w = tid >> 6, l = tid & 63;
  // bit CL r + c of `nzm` says whether bin (RW w + r, CL l + c) is nonzero
  uint64_t nzm = 0ull;
  // per row r: nonzero bins of the lanes below in the row, a byte each (at
  // most CL * 63)
  uint32_t rowpre[(RW + 3) / 4];
  uint32_t reg[kRegs];
#pragma unroll
  for (int q = 0; q < (RW + 3) / 4; ++q) rowpre[q] = 0u;
  {
    // without KEEP, the lane's row sums are kept instead
    uint32_t csum[CL], rsum[KEEP ? 1 : RW];
#pragma unroll
    for (int c = 0; c < CL; ++c) csum[c] = 0u;
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const int i = RW * w + r;
      uint32_t n = 0u, rs = 0u;
      if constexpr (KEEP && kPack) {
        // the lane's 4 bins of row i are 16-bit halves of two words of the
        // packed histogram (H even); halves past the row's end read as 0
        uint32_t w0 = 0u, w1 = 0u;
        if (i < H && CL * l < H) {
          const int b = (i * H + CL * l) >> 1;
          w0 = packed[b];
          if (CL * l + 2 < H) w1 = packed[b + 1];
        }
        reg[2 * r] = w0;
        reg[2 * r + 1] = w1;
      }
#pragma unroll
      for (int c = 0; c < CL; ++c) {
        const int j = CL * l + c;
        const int k = r * CL + c;
        uint32_t v;
        if constexpr (KEEP && kPack) {
          v = (reg[k >> 1] >> (16 * (k & 1))) & 0xffffu;
        } else {
          v = (i < H && j < H) ? count_at(i * H + j) : 0u;
          if (KEEP) reg[KEEP && !kPack ? k : 0] = v;
        }
        csum[c] += v;
        rs += v;
        if (v) {
          nzm |= 1ull << k;
          ++n;
        }
      }
      if (!KEEP) rsum[KEEP ? 0 : r] = rs;
      // the lane's count n <= CL < 8, bit-sliced over three ballots
      const uint64_t b0 = __ballot(n & 1u), b1 = __ballot(n & 2u), b2 = __ballot(n & 4u);
      if (l == 0)
        s.rowbase[i] =
            (uint32_t)__popcll(b0) + 2u * (uint32_t)__popcll(b1) + 4u * (uint32_t)__popcll(b2);
      const uint32_t pre = lanes_below(b0) + 2u * lanes_below(b1) + 4u * lanes_below(b2);
      rowpre[r >> 2] |= pre << (8 * (r & 3));
    }
    // row sums over the wave's 64 lanes by halving exchanges: across lane
    // bit 32, 16, ... a lane keeps the half of its rows that the bit selects
    // and adds its partner's copy of that half (RW - 1 shuffles), then the
    // low lane bits are summed; lane l ends with row l >> (6 - log2 RW).  The
    // lane's own row sums come from its counts as the first exchange needs
    // them.
    auto own_row = [&](int r) -> uint32_t {
      if (!KEEP) return rsum[r];
      uint32_t t = 0u;
#pragma unroll
      for (int c = 0; c < CL; ++c) {
        const int k = r * CL + c;
        t += kPack ? (reg[KEEP ? k >> 1 : 0] >> (16 * (k & 1))) & 0xffffu : reg[KEEP ? k : 0];
      }
      return t;
    };
    uint32_t rsv[RW / 2];
#pragma unroll
    for (int t = 0; t < RW / 2; ++t) {
      const bool up = (l & 32) != 0;
      const uint32_t lo = own_row(t), hi = own_row(RW / 2 + t);
      rsv[t] = (up ? hi : lo) + __shfl_xor(up ? lo : hi, 32, 64);
    }
#pragma unroll
    for (int st = 1; st < kLogRW; ++st) {
      const int half = RW >> (st + 1), sbit = 32 >> st;
      const bool up = (l & sbit) != 0;
#pragma unroll
      for (int t = 0; t < half; ++t) {
        const uint32_t give = up ? rsv[t] : rsv[half + t];
        const uint32_t keep = up ? rsv[half + t] : rsv[t];
        rsv[t] = keep + __shfl_xor(give, sbit, 64);
      }
    }
    uint32_t tot = rsv[0];
#pragma unroll
    for (int o = 32 >> kLogRW; o > 0; o >>= 1) tot += __shfl_xor(tot, o, 64);
    {
      const int i = RW * w + (l >> (6 - kLogRW));
      if ((l & ((64 >> kLogRW) - 1)) == 0 && i < H) s_rows[i] = tot;
    }
#pragma unroll
    for (int c = 0; c < CL; ++c) {
      if (csum[c]) atomicAdd(&s_cols[CL * l + c], csum[c]);
      s.colpre[w][CL * l + c] = (uint8_t)__popcll(nzm & T::col_mask(c));
    }
    // rows and columns beyond the tiles count no bins
#pragma unroll
    for (int c = CL; c < 4; ++c) s.colpre[w][64 * CL + (4 - CL) * l + (c - CL)] = 0;
    if (16 * RW < 256 && tid >= 16 * RW && tid < 256) s.rowbase[tid] = 0u;
  }
  __syncthreads();
  // column bases (waves 0-3), row bases (waves 4-7), marginal logs (8-15)
  uint32_t mine = 0u, incl = 0u;
  if (tid < 512) {
    if (tid < 256) {
      uint32_t run = 0u;
      for (int q = 0; q < 16; ++q) {
        const uint32_t v = s.colpre[q][tid];
        s.colpre[q][tid] = (uint8_t)run;  // at most 15 RW
        run += v;
      }
      mine = run;
    } else {
      mine = s.rowbase[tid - 256];
    }
    incl = mine;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(incl, o, 64);
      if (l >= o) incl += t;
    }
    if (l == 63) s.wtot[w] = incl;
  } else if (tid < 768) {
    const int q = tid - 512;
    const uint32_t c = q < H ? s_cols[q] : 0u;
    s.mcol[q] = c ? make_float2((float)c, logf_cr((float)c)) : make_float2(0.0f, 0.0f);
  } else {
    const int q = tid - 768;
    const uint32_t c = q < H ? s_rows[q] : 0u;
    s.mrow[q] = c ? make_float2((float)c, logf_cr((float)c)) : make_float2(0.0f, 0.0f);
  }
  __syncthreads();
  if (tid < 512) {
    uint32_t off = 0u;
    for (int q = tid < 256 ? 0 : 4; q < w; ++q) off += s.wtot[q];
    if (tid < 256)
      s.colbase[tid] = incl - mine + off;
    else
      s.rowbase[tid - 256] = incl - mine + off;
  }
  const uint32_t nz = s.wtot[4] + s.wtot[5] + s.wtot[6] + s.wtot[7];
  __syncthreads();

  // each list's room in win: capw terms, a multiple of kChainBlock (the
  // mirror list starts 16-byte aligned at win + capw); a list is folded from
  // at most capw - kChainBlock terms at a time, so its padding and the fold's
  // read-ahead stay inside its room
  const int capw = cap / kChainBlock * kChainBlock;
  const int step = capw - kChainBlock;
  // both lists straight into win when they fit and the registers hold the
  // counts; else through the global staging area, window by window
  const bool direct = KEEP && (int)nz <= step;
  float *stage = direct ? nullptr : claim_stage();
  const uint32_t HH = (uint32_t)(H * H);
  // positions in chain order, row by row of the lane's tile: a row's nonzero
  // bins are consecutive in the row-major list from the row's base plus the
  // lanes before; in the column-major list column j continues from its base
  // plus the bins above in earlier waves, counted up row by row here
  {
    uint32_t colrun[CL];
#pragma unroll
    for (int c = 0; c < CL; ++c) colrun[c] = s.colbase[CL * l + c] + s.colpre[w][CL * l + c];
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const uint32_t bits = (uint32_t)(nzm >> (CL * r)) & (uint32_t)T::kRowBits;
      if (bits) {
        const int i = RW * w + r;
        uint32_t pa = s.rowbase[i] + ((rowpre[r >> 2] >> (8 * (r & 3))) & 0xffu);
#pragma unroll
        for (int c = 0; c < CL; ++c) {
          if (bits & (1u << c)) {
            const int e = i * H + CL * l + c;
            const int k = r * CL + c;
            uint32_t v;
            if (KEEP && direct) {
              v = kPack ? (reg[KEEP ? k >> 1 : 0] >> (16 * (k & 1))) & 0xffffu : reg[KEEP ? k : 0];
              if (need_take) (void)take_at(e);
            } else {
              v = take_at(e);
            }
            const float p = (float)v;
            // positions are a permutation of [0, nz); the bounds only keep a
            // wrong count from writing past the lists
            if (direct) {
              if (pa < nz) win[pa] = make_float2(p, 0.0f);
              if (mirror && colrun[c] < nz) win[capw + colrun[c]] = make_float2(p, 0.0f);
            } else {
              if (pa < HH) stage[pa] = p;
              if (mirror && colrun[c] < HH) stage[HH + colrun[c]] = p;
            }
            ++pa;
            ++colrun[c];
          }
        }
      }
    }
  }
  __syncthreads();

  // one lane per chain: (i, j) on wave 0, (j, i) on wave 1, the column and
  // row marginals on waves 2 and 3
  float acc_ij = 0.0f, acc_ji = 0.0f, hx = 0.0f, hy = 0.0f;
  auto fold = [&](int len, bool first) {
    if (tid == 0) {
      acc_ij = fma_chain<CQ>(win, len, acc_ij);
    } else if (tid == 64) {
      if (mirror) acc_ji = fma_chain<CQ>(win + capw, len, acc_ji);
    } else if (tid == 128) {
      // zero column sums, and the entries from H to 256, are (0, 0) no-ops
      if (first) hx = fma_chain<CQ>(s.mcol, H, hx);
    } else if (tid == 192) {
      if (first) hy = fma_chain<CQ>(s.mrow, H, hy);
    }
  };
  if (direct) {
    // the logarithms in place; (0, 0) from nz to the next kChainBlock
    const int nzp = ((int)nz + kChainBlock - 1) / kChainBlock * kChainBlock;
    for (int q = tid; q < nzp; q += kNmiBlock) {
      if (q < (int)nz) {
        win[q].y = logf_cr(win[q].x);
        if (mirror) win[capw + q].y = logf_cr(win[capw + q].x);
      } else {
        win[q] = make_float2(0.0f, 0.0f);
        if (mirror) win[capw + q] = make_float2(0.0f, 0.0f);
      }
    }
    __syncthreads();
    fold((int)nz, true);
  } else {
    // window by window through LDS, each entry's logarithm taken by the lane
    // that brings it in
    for (uint32_t w0 = 0;; w0 += (uint32_t)step) {
      const uint32_t w1 = w0 + (uint32_t)step < nz ? w0 + (uint32_t)step : nz;
      const int len = w0 < nz ? (int)(w1 - w0) : 0;
      const int lenp = (len + kChainBlock - 1) / kChainBlock * kChainBlock;
      for (int q = tid; q < lenp; q += kNmiBlock) {
        float2 t = make_float2(0.0f, 0.0f), u = t;
        if (q < len) {
          const float c = stage[w0 + q];
          t = make_float2(c, logf_cr(c));
          if (mirror) {
            const float d = stage[HH + w0 + q];
            u = make_float2(d, logf_cr(d));
          }
        }
        win[q] = t;
        if (mirror) win[capw + q] = u;
      }
      __syncthreads();
      fold(len, w0 == 0);
      __syncthreads();
      if (w1 >= nz) break;
    }
    release_stage();
  }
  if (tid == 64) s.acc[1] = acc_ji;
  if (tid == 128) s.acc[2] = hx;
  if (tid == 192) s.acc[3] = hy;
  __syncthreads();
  if (tid == 0) {
    const float fN = (float)N;
    const float lN = logf_cr(fN);
    const float fx = s.acc[2] / fN + lN, fy = s.acc[3] / fN + lN;
    const float sxy = fx + fy;
    float fxy = acc_ij / fN + lN;
    float v = 2.0f * (sxy - fxy) / sxy;
    *out_ij = v != v ? 0.0f : v;
    if (mirror) {
      fxy = s.acc[1] / fN + lN;
      v = 2.0f * (sxy - fxy) / sxy;
      *out_ji = v != v ? 0.0f : v;
    }
  }
}

// Staging slots of a launch whose workgroups are not persistent: a bitmap of
// free slots (zero at rest), `words` 32-bit words per XCD.  A workgroup only
// claims slots of the XCD it runs on (the hardware XCC id): each XCD has its
// own write-back L2, not coherent with the others within a kernel, so a slot
// written on one XCD and then reused on another could have the first XCD's
// dirty lines written back over the second's data.  Within one XCD the next
// holder's stores land in the same L2 lines.  Each XCD has at least as many
// slots as workgroups of the kernel that fit on its CUs at once, so a claim
// finds a free bit on its first sweep; a workgroup that holds a slot never
// waits, so a claim that found none still ends once one is released.
__device__ __forceinline__ uint32_t nmi32_claim(uint32_t *bits, int words, int nxcd) {
  // s_getreg of HW_REG_XCC_ID (register 20), bits [3:0]
  const uint32_t xcd = (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) %
                       (uint32_t)nxcd;
  for (;;) {
    for (int k = 0; k < words; ++k) {
      const int w = (int)xcd * words + k;
      uint32_t v = __hip_atomic_load(&bits[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      while (~v) {
        const uint32_t bit = 1u << (__ffs(~v) - 1);
        const uint32_t old = atomicOr(&bits[w], bit);
        if (!(old & bit)) return (uint32_t)w * 32u + (uint32_t)(__ffs(bit) - 1);
        v = old | bit;
      }
    }
    __builtin_amdgcn_s_sleep(8);
  }
}

// Called by one lane after the workgroup's last read of the slot has returned
// (the window copies end in LDS stores of the loaded values).  The next
// holder writes before it reads, and its writes depend on the claimed index,
// so relaxed atomics order the hand-over.
__device__ __forceinline__ void nmi32_release(uint32_t *bits, uint32_t slot) {
  atomicAnd(&bits[slot >> 5], ~(1u << (slot & 31u)));
}

}  // namespace iso
