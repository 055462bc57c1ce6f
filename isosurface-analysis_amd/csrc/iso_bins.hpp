// Exact histogram binning shared by the histogram kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace iso {

// Bin of v for a pair with range [mn, mx] and H bins, as the reference
// computes it (JointHistogram.cu:63-75, :127-129): step = (mx - mn) / H,
// bin = (int)((v - mn) / step) clamped to H - 1, step == 0 -> bin 0.
// With a = v - mn (the reference's own rounded difference) and t = a / step
// exactly, q = a * RN(1 / step) is t (1 + d1)(1 + d2), |d1|, |d2| <= 2^-24,
// and t <= H (1 + 2^-23), so |q - t| <= H 2^-23 (1 + 2^-22); the reference's
// RN(a / step) is within H 2^-24 of t.  When frac(q) is at least
// margin = H 2^-21 (more than twice their distance) away from an integer,
// no integer lies between q and RN(a / step) and trunc(q) is the reference's
// bin; the rare rest take the IEEE division.
struct Binner {
  float mn, step, rstep, lo, hi, half_in;
  int H;
  int mode;  // 0: step == 0 (everything in bin 0); 1: fast + exact fix-up; 2: exact only

  __device__ __forceinline__ void init(float mn_, float mx_, int H_) {
    H = H_;
    mn = mn_;
    step = (mx_ - mn_) / (float)H_;
    if (step == 0.0f) { mode = 0; return; }
    rstep = 1.0f / step;
    float margin = (float)H_ * 0x1p-21f;
    lo = margin;
    hi = 1.0f - margin;
    // |frac - 1/2| <= 1/2 - margin is lo <= frac <= hi: exact for frac >=
    // 1/4; below that frac - 1/2 rounds by at most 2^-26, far inside the
    // margin's slack
    half_in = 0.5f - margin;
    mode = (rstep < 1e30f && rstep > 0.0f && margin < 0.25f) ? 1 : 2;
  }

  __device__ __forceinline__ int operator()(float v) const {
    if (mode == 0) return 0;
    float a = v - mn;
    int k;
    if (mode == 1) {
      const float q = a * rstep;
      k = (int)q;
      // q - trunc(q) in one v_fract_f32 (q >= 0, so floor == trunc; exact)
      const float fr = __builtin_amdgcn_fractf(q);
      if (!(fr >= lo && fr <= hi)) k = (int)(a / step);
    } else {
      k = (int)(a / step);
    }
    // v >= mn for every value of the pair (mn is their minimum), so a >= 0
    // and k >= 0 (a NaN converts to 0)
    return min(k, H - 1);
  }
};

// Joint bins e[u] = bin(x[u]) * H + bin(y[u]) of N value pairs, as
// Binner::operator() gives them.  In the fast mode every value takes the
// quotient; a pair whose value lies within the margin of an integer sends its
// wave (a uniform branch, rarely taken) through the exact binner for that
// pair.  No clamp is needed on the fast path: q <= H (1 + 2^-24)(1 + 2^-23)
// (a <= RN(mx - mn), step >= RN(mx - mn) / H (1 - 2^-24)), so a q at or past
// H has frac(q) below H 2^-22, inside the margin, and takes the exact path.
// H <= 256, so the index is a 24-bit multiply-add.
template <int N>
__device__ __forceinline__ void joint_bins(const Binner &b, const float (&x)[N],
                                           const float (&y)[N], uint32_t (&e)[N]) {
  if (b.mode == 1) {
#pragma unroll
    for (int u = 0; u < N; ++u) {
      const float qx = (x[u] - b.mn) * b.rstep, qy = (y[u] - b.mn) * b.rstep;
      const bool bad = !(fabsf(__builtin_amdgcn_fractf(qx) - 0.5f) <= b.half_in) ||
                       !(fabsf(__builtin_amdgcn_fractf(qy) - 0.5f) <= b.half_in);
      const uint32_t kx = (uint32_t)qx, ky = (uint32_t)qy;
      e[u] = __umul24(kx, (uint32_t)b.H) + ky;
      if (__builtin_expect(__ballot(bad) != 0ull, 0))
        e[u] = (uint32_t)(b(x[u]) * b.H + b(y[u]));
    }
  } else {
#pragma unroll
    for (int u = 0; u < N; ++u) e[u] = (uint32_t)(b(x[u]) * b.H + b(y[u]));
  }
}


// Workgroup barrier that orders LDS only: it does not wait for global loads
// or stores in flight (a round loop keeps its next loads outstanding).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// word 3 of a raw 32-bit buffer resource on the gfx9 family
constexpr int kBufferWord3 = 0x00020000;

}  // namespace iso
