// Exact histogram binning shared by the histogram kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace iso {

// Bin of v for a pair with range [mn, mx] and H bins, as the reference
// computes it (JointHistogram.cu:63-75, :127-129): step = (mx - mn) / H,
// bin = (int)((v - mn) / step) clamped to H - 1, step == 0 -> bin 0.
// q = (v - mn) * (1 / step) is within H * 2^-23 of the true quotient, so when
// frac(q) is at least H * 2^-19 away from an integer, trunc(q) is the
// reference's trunc((v - mn) / step); the rare rest take the IEEE division.
struct Binner {
  float mn, step, rstep, lo, hi;
  int H;
  int mode;  // 0: step == 0 (everything in bin 0); 1: fast + exact fix-up; 2: exact only

  __device__ __forceinline__ void init(float mn_, float mx_, int H_) {
    H = H_;
    mn = mn_;
    step = (mx_ - mn_) / (float)H_;
    if (step == 0.0f) { mode = 0; return; }
    rstep = 1.0f / step;
    float margin = (float)H_ * 0x1p-19f;
    lo = margin;
    hi = 1.0f - margin;
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
