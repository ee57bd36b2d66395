// FP64 exp, log and reciprocal for the sweeps of the FP64 path.
//
// exp: Tang's table-driven method (P. T. P. Tang, "Table-driven
//   implementation of the exponential function in IEEE floating-point
//   arithmetic", ACM TOMS 15(2), 1989) with 64 table points:
//   x = (64 m + j) ln2/64 + r, |r| <= ln2/128, and
//   exp x = 2^m 2^(j/64) (1 + p(r)), p the degree-5 Taylor polynomial of
//   e^r - 1 (truncation below 3.5e-17 relative).  2^(j/64) is held as a
//   double-double, so the result carries one rounding plus the polynomial's.
// log: Tang's table-driven method for the logarithm (ACM TOMS 16(4), 1990)
//   with 129 breakpoints: x = 2^e y, y in [1/2, 1), F the breakpoint
//   1/2 + j/256 nearest y, f = y - F (exact), u = f / F, |u| <= 1/256, and
//   log x = e ln2 + log F + log1p(u), log1p(u) - u from its degree-7 Taylor
//   polynomial (truncation below 2e-18 relative).  e ln2_hi + (log F)_hi is
//   exact, and u's rounding is recovered as (f - u F) / F, so the result
//   carries one rounding plus the small terms'.
//
// Both exp and log stay within 1 ulp of the correctly rounded result
// (checked against 80-bit long double on the CPU, where the same source
// compiles).  Every multiply-add is an explicit fma, so the device build and
// the host build give the same bits.  The table pointers may address global
// memory or a copy in LDS; ExpTab is 64 {hi, lo} pairs, LogTab 129 {hi, lo}
// pairs and InvTab 129 values (fmath64_tables.inc).
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define UMWM_FM64_HD __host__ __device__
#else
#define UMWM_FM64_HD
#endif

namespace umwm {
namespace fm64 {

#include "fmath64_tables.inc"

constexpr int kExpTabLen = 2 * 64;
constexpr int kLogTabLen = 2 * 129;
constexpr int kInvTabLen = 129;

UMWM_FM64_HD inline uint64_t bits_of(double x) { return __builtin_bit_cast(uint64_t, x); }
UMWM_FM64_HD inline double double_of(uint64_t b) { return __builtin_bit_cast(double, b); }

// e^x.  Arguments below -746 give 0 and above 710 give +inf; NaN stays NaN
// (with kNaN false NaN gives 0 or +inf: for callers whose NaN reaches the
// result by another operand).
template <bool kNaN = true>
UMWM_FM64_HD inline double exp_t(double x, const double* __restrict__ tab) {
  const double xc = __builtin_fmin(__builtin_fmax(x, -746.0), 710.0);
  const double t = __builtin_rint(xc * kExpInvL);
  const int n = int(t);
  const double r = __builtin_fma(t, -kExpL2, __builtin_fma(t, -kExpL1, xc));
  double a = __builtin_fma(r, 1.0 / 120, 1.0 / 24);
  a = __builtin_fma(r, a, 1.0 / 6);
  a = __builtin_fma(r, a, 0.5);
  const double p = __builtin_fma(r * r, a, r);
  const int j = n & 63;
  const double hi = tab[2 * j], lo = tab[2 * j + 1];
  const double v = __builtin_ldexp(hi + __builtin_fma(hi, p, lo), n >> 6);
  if constexpr (kNaN) return x == x ? v : x;
  return v;
}

// log x.  0 gives -inf, negative arguments NaN, +inf stays +inf (with
// kSpecial false only 0 is treated apart: for finite x >= 0).
template <bool kSpecial = true>
UMWM_FM64_HD inline double log_t(double x, const double* __restrict__ ltab,
                                 const double* __restrict__ itab) {
  int e;
  const double y = __builtin_frexp(x, &e);  // x = y 2^e, y in [1/2, 1)
  const uint32_t top = uint32_t(bits_of(y) >> 32);
  const uint32_t j = ((top & 0xFFFFFu) + 0x1000u) >> 13;  // nearest of 1/2 + j/256
  const double F = double_of(uint64_t(0x3FE00000u + (j << 13)) << 32);
  const double f = y - F, iF = itab[j];
  // u as an fma with a +0 addend: a plain product would be contracted into
  // the u + (...) below, counting u's rounding twice with rho
  const double u = __builtin_fma(f, iF, 0.0);
  const double rho = __builtin_fma(-u, F, f);  // f - u F: u's rounding, recovered
  double a = __builtin_fma(u, 1.0 / 7, -1.0 / 6);
  a = __builtin_fma(u, a, 1.0 / 5);
  a = __builtin_fma(u, a, -0.25);
  a = __builtin_fma(u, a, 1.0 / 3);
  a = __builtin_fma(u, a, -0.5);
  const double ed = double(e);
  const double hi = __builtin_fma(ed, kLn2Hi, ltab[2 * j]);  // exact
  const double lo = __builtin_fma(rho, iF, __builtin_fma(ed, kLn2Lo, ltab[2 * j + 1]));
  const double v = hi + (u + __builtin_fma(u * u, a, lo));
  // finite positive arguments (normal or not) take v; a select, not a
  // branch, so that the compiler does not sink the evaluation into a
  // conditional block
  if constexpr (!kSpecial) return x == 0.0 ? -__builtin_inf() : v;
  const double sp = x == 0.0 ? -__builtin_inf() : (x < 0.0 ? __builtin_nan("") : x);
  return (x > 0.0 && x < __builtin_inf()) ? v : sp;
}

}  // namespace fm64
}  // namespace umwm
