// px_device.h -- per-pixel projections shared by the fused passes
// (kern_pass.inc) and the any-length passes (kern_any.inc).
#pragma once
#include "fft_device.h"

namespace nadia {

// BaseCDI::support_constraint per pixel (NADIA/src/BaseCDI.c++:136-154):
// 0 -> exact zero, (0,1) -> scaled, >= 1 -> unchanged.
template <typename T>
__device__ __forceinline__ c2_t<T> support_px(c2_t<T> v, T s) {
  if (s == (T)0) return c2_t<T>{(T)0, (T)0};
  if (s < (T)1) return c2_t<T>{v.x * s, v.y * s};
  return v;
}

// BaseCDI::scale_intensity per pixel (NADIA/src/BaseCDI.c++:163-189) with
// Complex_2D::set_mag (NADIA/include/Complex_2D.h:205-215).  q = sqrt(I); a
// set sign bit marks a beam-stop pixel, which is left unchanged and excluded
// from the error sums.
template <typename T>
__device__ __forceinline__ c2_t<T> modulus_px(c2_t<T> y, T q, double& diff, double& norm) {
  if (signbit(q)) return y;
  const T mag = sqrt(y.x * y.x + y.y * y.y);
  const double dm = (double)mag - (double)q;
  diff += dm * dm;
  norm += (double)q * (double)q;
  if (mag == (T)0) return c2_t<T>{q, (T)0};
  const T r = q / mag;
  return c2_t<T>{y.x * r, y.y * r};
}

// Single-precision bulk path: |y| and q/|y| from one hardware reciprocal
// square root (about 1 ulp) instead of an IEEE square root and division; the
// error sums stay in double.  Tiny magnitudes take the IEEE path.
template <>
__device__ __forceinline__ float2 modulus_px<float>(float2 y, float q, double& diff, double& norm) {
  if (signbit(q)) return y;
  const float s2 = __builtin_fmaf(y.x, y.x, y.y * y.y);
  float mag, r;
  if (s2 > 1e-30f) {
    const float rs = __builtin_amdgcn_rsqf(s2);
    mag = s2 * rs;
    r = q * rs;
  } else {
    mag = sqrtf(s2);
    r = mag == 0.f ? 0.f : q / mag;
  }
  const double dm = (double)mag - (double)q;
  diff += dm * dm;
  norm += (double)q * (double)q;
  if (mag == 0.f) return float2{q, 0.f};
  return float2{y.x * r, y.y * r};
}

}  // namespace nadia
