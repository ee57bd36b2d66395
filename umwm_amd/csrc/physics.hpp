// Per-bin and per-point physics shared by the operator kernels and the
// fused kernels.  Every function keeps the reference's expression order so
// the FP64 build reproduces a serial FP64 run of the reference up to the
// rounding of the device math library.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace umwm {

constexpr double kTwoPiD = 2 * 3.141592653589793;  // src/umwm_module.F90:87-89

// x**y (Fortran real power) for the spectral densities, y > 0.  FP64
// evaluates exp(y log x) with the device library's FP64 exp and log
// (relative error ~|y log x| ulp, below 1e-14 for the spectra here); FP32
// uses the hardware log2/exp2 pair, inside the FP32 path's error budget
// (Hs within 0.5 % of FP64).  A non-positive base gives 0: round-off can
// leave a bin at -1e-30 where refraction or propagation empties it exactly
// (Courant number 1), and the reference's x**y would turn that into NaN.
__device__ inline double pw(double x, double y) { return x > 0.0 ? exp(y * log(x)) : 0.0; }
__device__ inline float pw(float x, float y) {
  return x > 0.f ? __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)) : 0.f;
}

__device__ inline double ex(double x) { return exp(x); }
__device__ inline float ex(float x) { return __expf(x); }

// src/umwm_sheltering.F90:49-79 sheltering_coare35
template <typename T>
__device__ inline T sheltering_coare35(T wspd) {
  const T x1 = 15, x2 = 33, x3 = 60;
  const T y1 = T(0.10), y2 = T(0.09), y3 = T(0.06);
  const T intercept = T(0.04), curvature = T(0.65), decay = T(1.6);
  const T slope1 = (y1 - intercept) / x1;
  const T slope2 = (y3 - y2) / (x3 - x2);
  const T c = curvature * (slope2 - slope1) / (x2 - x1);
  const T b = slope1 - 2 * c * x1;
  const T a = y1 - b * x1 - c * (x1 * x1);
  const T s1 = a + b * x2 + c * (x2 * x2);
  if (wspd <= x1) return intercept + slope1 * wspd;
  if (wspd <= x2) return a + b * wspd + c * (wspd * wspd);
  return s1 * ex(-(wspd - x2) / (decay * wspd));
}

// Per-point inputs of sin_d12 (src/umwm_source_functions.F90:16-82).
template <typename T>
struct SinPoint {
  T a;        // 2.5 * ustar
  T wspd;
  T pre;      // (1 - fice) * twopi * rhorat
  T shelt;
  T uc, vc;
  int oc;     // 0-based last prognostic bin
};

// ssin(o,p,i) for one bin; cwt = cos(wdir - th(p)).
template <typename T>
__device__ inline T ssin_bin(const SinPoint<T>& s, T logl2, T cwt, T cp0, T cth, T sth,
                             T fkovg, T sin_fac, T fs1, T fs2, int o) {
  T x = (s.wspd + s.a * logl2) * cwt - cp0 - s.uc * cth - s.vc * sth;
  x = sin_fac * fabs(x) * x;
  if (x > T(0)) x = x * s.shelt / sin_fac;
  if (x < T(0)) x = x * fs1;
  if (x < T(0) && cwt > T(0)) x = x * fs2;
  x = s.pre * x * fkovg;
  if (o > s.oc && x < T(0)) x = T(0);  // diagnostic tail: max(ssin, 0)
  return x;
}

// ssin_bin with ba = wspd + 2.5 ustar logl2overz(o) evaluated once per
// frequency.
template <typename T>
__device__ inline T ssin_bin2(const SinPoint<T>& s, T ba, T cwt, T cp0, T cth, T sth, T fkovg,
                              T sin_fac, T fs1, T fs2, int o) {
  T x = ba * cwt - cp0 - s.uc * cth - s.vc * sth;
  x = sin_fac * fabs(x) * x;
  if (x > T(0)) x = x * s.shelt / sin_fac;
  if (x < T(0)) x = x * fs1;
  if (x < T(0) && cwt > T(0)) x = x * fs2;
  x = s.pre * x * fkovg;
  if (o > s.oc && x < T(0)) x = T(0);
  return x;
}

// ssin_bin with the wind projection expanded by the angle-sum identity:
// (wspd + 2.5 ustar logl2overz) cos(wdir - th) - uc cth - vc sth
//   = A cth + B sth,  A = ba cos(wdir) - uc,  B = ba sin(wdir) - vc,
// A and B evaluated once per frequency; cwt = cos(wdir - th) is still
// needed for the swell-overrunning-wind test.
template <typename T>
__device__ inline T ssin_ab(const SinPoint<T>& s, T A, T B, T cwt, T cp0, T cth, T sth, T fkovg,
                            T sin_fac, T fs1, T fs2, int o) {
  T x = A * cth + B * sth - cp0;
  x = sin_fac * fabs(x) * x;
  if (x > T(0)) x = x * s.shelt / sin_fac;
  if (x < T(0)) x = x * fs1;
  if (x < T(0) && cwt > T(0)) x = x * fs2;
  x = s.pre * x * fkovg;
  if (o > s.oc && x < T(0)) x = T(0);
  return x;
}

// Cut-off bin search, src/umwm_source_functions.F90:24-36.  Returns the
// 0-based index of the highest bin <= om-3 whose frequency is below
// fcutoff (bin 0 if none).
template <typename T>
__device__ inline int cutoff_bin(T fcutoff, const T* f, int om) {
  int oc = 0;
  for (int o = om - 3; o >= 0; --o) {
    oc = o;
    if (fcutoff > f[o]) break;
  }
  return oc;
}

// s_ice attenuation for a point with significant wave height ht
// (src/umwm_source_functions.F90:108-153); multiplied by 2 cg0 per bin.
__device__ inline double sice_coef(double ht) {
  const double H_th = 3.0, C1 = -5.35e-6, C2 = C1 * H_th;
  return ht < H_th ? C1 * ht : C2;
}

// ---------------------------------------------------------------------
// stress('atm') per-point chain, src/umwm_stress.F90:216-264 and the
// helpers at :268-344, always in FP64.
// ---------------------------------------------------------------------
__device__ inline double stress_tail(double wspd, double kmax) {
  const double kcap = 1e3, a = 0.000112, b = -0.01451, c = -1.0186;
  double tail = a * (wspd * wspd) + b * wspd + c;
  double kpt = pow(kmax, tail);
  return (pow(kcap, tail + 1) - kmax * kpt) / (kpt * (tail + 1));
}

__device__ inline double friction_velocity_skin(double wspd, double z, double nu,
                                                double kappa) {
  double z0 = 1e-3, ust = 0;
  for (int n = 0; n < 6; ++n) {  // exactly six fixed-point passes
    ust = kappa * wspd / log(z / z0);
    z0 = 0.132 * nu / ust;
  }
  return ust;
}

struct StressIn {
  double sx, sy;      // sum_o taux_util(o) * kdk(o)
  double dx, dy;      // same over o >= oc
  double tomx, tomy;  // taux_util(om), tauy_util(om)
  double wspd, wdir, uc, vc, us1, vs1, rhoa, rhow, kom;
};

struct StressOut {
  double taux, tauy, taux_form, tauy_form, taux_skin, tauy_skin;
  double taux_diag, tauy_diag, tailatmx, tailatmy, cd, ustar;
};

__device__ inline StressOut stress_atm_point(const StressIn& s, double dthg, double z,
                                             double nu_air, double kappa) {
  StressOut r;
  const double tail = stress_tail(s.wspd, s.kom);
  r.tailatmx = s.tomx * s.kom * tail * s.rhow * dthg;
  r.tailatmy = s.tomy * s.kom * tail * s.rhow * dthg;
  r.taux_form = s.sx * s.rhow * dthg + r.tailatmx;
  r.tauy_form = s.sy * s.rhow * dthg + r.tailatmy;
  r.taux_diag = s.dx * s.rhow * dthg;
  r.tauy_diag = s.dy * s.rhow * dthg;
  // wind_relative with the surface Stokes drift (:240-243, :333-344)
  const double urel = s.wspd * cos(s.wdir) - (s.uc + s.us1);
  const double vrel = s.wspd * sin(s.wdir) - (s.vc + s.vs1);
  const double wspdrel = sqrt(urel * urel + vrel * vrel);
  const double wdirrel = atan2(vrel, urel);
  const double w2 = s.wspd * s.wspd;
  const double cd_form =
      sqrt(r.taux_form * r.taux_form + r.tauy_form * r.tauy_form) / (s.rhoa * w2);
  const double u = friction_velocity_skin(wspdrel, z, nu_air, kappa);
  double cd_skin = (u * u) / w2;
  cd_skin = cd_skin * (1 + 2 * cd_skin / (cd_skin + cd_form + 2.2250738585072014e-308)) / 3;
  if (cd_skin > 1e-2) cd_skin = 1e-2;
  r.taux_skin = s.rhoa * cd_skin * (wspdrel * wspdrel) * cos(wdirrel);
  r.tauy_skin = s.rhoa * cd_skin * (wspdrel * wspdrel) * sin(wdirrel);
  r.taux = r.taux_form + r.taux_skin;
  r.tauy = r.tauy_form + r.tauy_skin;
  const double tn = sqrt(r.taux * r.taux + r.tauy * r.tauy);
  r.cd = tn / (s.rhoa * w2);
  r.ustar = sqrt(tn / s.rhoa);
  return r;
}

template <typename T>
__device__ inline void store_stress(const DomainDev<T>& d, int j, const StressOut& r) {
  d.taux[j] = r.taux;
  d.tauy[j] = r.tauy;
  d.taux_form[j] = r.taux_form;
  d.tauy_form[j] = r.tauy_form;
  d.taux_skin[j] = r.taux_skin;
  d.tauy_skin[j] = r.tauy_skin;
  d.taux_diag[j] = r.taux_diag;
  d.tauy_diag[j] = r.tauy_diag;
  d.tailatmx[j] = r.tailatmx;
  d.tailatmy[j] = r.tailatmy;
  d.cd[j] = r.cd;
  d.ustar[j] = r.ustar;
}

// ---------------------------------------------------------------------
// wavefront reductions
// ---------------------------------------------------------------------
__device__ inline double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Order-preserving map of doubles (NaN aside) onto unsigned 64-bit keys, so
// that an integer atomic max combines doubles; ~okey(x) does the same for a
// min.  No double maps to key 0, which therefore stands for "no value".
__device__ inline unsigned long long okey(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ inline double okey_inv(unsigned long long k) {
  return __longlong_as_double((long long)((k >> 63) ? (k & 0x7fffffffffffffffull) : ~k));
}

}  // namespace umwm
