// tc_device.h -- per-pixel pieces of TransmissionConstraint and of the
// FresnelCDI transmission function <-> exit-surface wave conversion, shared by
// the element-wise kernels and the row pass that applies the constraint while
// it rebuilds the row transform.
#pragma once

#include <cmath>

#include "fft_device.h"
#include "kernels.h"

namespace nadia {

// FresnelCDI::set_transmission_function (NADIA/src/FresnelCDI.c++:253-283):
// ESW = T L - L for one pixel
template <typename T>
__device__ __forceinline__ c2_t<T> esw_of(c2_t<T> t, c2_t<T> l) {
  const double lr = l.x, li = l.y, tr_ = t.x, ti = t.y;
  return c2_t<T>{(T)(lr * tr_ - li * ti - lr), (T)(lr * ti + li * tr_ - li)};
}

// FresnelCDI::get_transmission_function (NADIA/src/FresnelCDI.c++:298-355):
// T = (ESW + L) / L for one pixel, 1 where L = 0
template <typename T>
__device__ __forceinline__ c2_t<T> trans_of(c2_t<T> e, c2_t<T> l) {
  const double lr = l.x, li = l.y;
  const double den = lr * lr + li * li;
  if (den == 0) return c2_t<T>{(T)1, (T)0};
  const double er = e.x, ei = e.y;
  const double rn = lr * (er + lr) + li * (ei + li);
  const double in = lr * (ei + li) - li * (er + lr);
  return c2_t<T>{(T)(rn / den), (T)(in / den)};
}

// Complex_2D accessors (NADIA/include/Complex_2D.h:205-290): get_mag in T,
// set_mag / set_phase through double temporaries.
template <typename T>
__device__ __forceinline__ T tc_mag(c2_t<T> v) { return sqrt(v.x * v.x + v.y * v.y); }

template <typename T>
__device__ __forceinline__ double tc_phase(c2_t<T> v) { return (double)(T)atan2((double)v.y, (double)v.x); }

template <typename T>
__device__ __forceinline__ void tc_set_mag(c2_t<T>& v, double value) {
  const double mag = (double)tc_mag<T>(v);
  if (mag == 0) {
    v.x = (T)value;
    v.y = (T)0;
  } else {
    v.x = (T)((double)v.x * (value / mag));
    v.y = (T)((double)v.y * (value / mag));
  }
}

template <typename T>
__device__ __forceinline__ void tc_set_phase(c2_t<T>& v, double value) {
  const double mag = (double)tc_mag<T>(v);
  v.x = (T)(mag * cos(value));
  v.y = (T)(mag * sin(value));
}

// the field the constraint acts on at pixel p: the buffer's value, or the
// transmission function of the exit-surface wave it holds
template <typename T>
__device__ __forceinline__ c2_t<T> tc_load(const c2_t<T>* t, const TcDev& d, long long p) {
  return d.ill ? trans_of<T>(t[p], ((const c2_t<T>*)d.ill)[p]) : t[p];
}

// TransmissionConstraint::apply_constraint for one pixel of region m (0 =
// none): ComplexConstraint::get_new_mag / get_new_phase from the region's
// mean, then charge flipping and the unity cap
// (NADIA/src/TransmissionConstraint.c++:81-172)
template <typename T>
__device__ __forceinline__ c2_t<T> tc_px(c2_t<T> v, const TcDev& d, int m) {
  const double phase_old = tc_phase<T>(v);
  double mag_old = (double)tc_mag<T>(v);
  if (m > 0) {
    const int k = m - 1;
    const double c = d.c_mean[k], a1 = d.alpha1[k], a2 = d.alpha2[k];
    const double mag_new = exp((1 - a1) * log(mag_old) + a1 * c * phase_old);
    const double phase_new = (1 - a2) * phase_old + a2 * log(mag_old) / c;
    tc_set_mag<T>(v, mag_new);
    tc_set_phase<T>(v, phase_new);
    mag_old = mag_new;
  }
  if (d.charge_flip && (T)d.flip_sign * v.y > (T)0) v.y = -v.y;
  if (d.unity && mag_old > 1) tc_set_mag<T>(v, 1.0);
  return v;
}

}  // namespace nadia
