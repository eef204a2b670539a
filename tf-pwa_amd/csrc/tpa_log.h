// Natural logarithm and reciprocal of a positive normal double, for clip_log's x > eps branch
// (tf_pwa/model/model.py:283-292) in the evaluation passes, where the library log and IEEE
// division take ~90 of the ~540 vector instructions per event group.
//
// ln x: x = m 2^e with m in [sqrt(1/2), sqrt(2)) (frexp, one doubling), then
// ln m = 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.1716, from the series
// atanh(s) = s (1 + z/3 + z^2/5 + ...), z = s^2 <= 0.0295, summed to z^10 (truncation below
// 2^-57 of the result), and ln 2 split in a high part with trailing zero bits (e ln2_hi exact
// for |e| < 2^11) and a low part (Cody and Waite's argument reduction; J.-M. Muller,
// Elementary Functions, ch. 11). Error at most 2 ulp (mean 0.08) against the C library's log
// on 4M arguments over 1e-6 .. 1e6, from a reciprocal seed good to 2^-22 (the CPU tests).
//
// The reciprocal starts from an approximation (the hardware's v_rcp_f64 on the device) and
// takes two Newton steps, r <- r + r (1 - d r): the relative error squares per step, so an
// approximation good to 2^-20 ends within an ulp.
#pragma once
#include <math.h>

#ifndef TPA_LOG_FN
#define TPA_LOG_FN __device__ __forceinline__
// A series coefficient as an opaque scalar-register value at its point of use: each FMA of
// the series then reads it from a scalar register instead of the compiler keeping it in a
// vector register for the whole evaluation loop of a pass.
__device__ __forceinline__ double tpa_log_coef(double v) {
  asm volatile("" : "+s"(v));
  return v;
}
#define TPA_LOG_COEF(x) tpa_log_coef(x)
#endif
#ifndef TPA_LOG_COEF
#define TPA_LOG_COEF(x) (x)
#endif

template <class Approx>
TPA_LOG_FN double tpa_recip_pos(double d, Approx approx) {
  double r = approx(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}

template <class Approx>
TPA_LOG_FN double tpa_log_pos(double x, Approx approx) {
  const double ln2_hi = 6.93147180369123816490e-01;  // 0x3FE62E42FEE00000
  const double ln2_lo = 1.90821492927058770002e-10;  // ln 2 - ln2_hi
  int e;
  double m = frexp(x, &e);  // x = m 2^e, m in [1/2, 1)
  if (m < 0.70710678118654752440) {
    m += m;
    e -= 1;
  }
  const double num = m - 1.0, den = m + 1.0;  // both exact
  const double r = tpa_recip_pos(den, approx);
  double s = num * r;
  s = fma(fma(-s, den, num), r, s);  // s = num / den to an ulp
  const double z = s * s;
  // Horner's scheme over the ten coefficients 1/3 .. 1/21: each step takes one constant, so
  // on the device every coefficient can sit in a scalar register (an FMA with two non-inline
  // constants keeps one of them in a vector register for the whole pass)
  double p = fma(z, TPA_LOG_COEF(1.0 / 21.0), TPA_LOG_COEF(1.0 / 19.0));
  p = fma(z, p, TPA_LOG_COEF(1.0 / 17.0));
  p = fma(z, p, TPA_LOG_COEF(1.0 / 15.0));
  p = fma(z, p, TPA_LOG_COEF(1.0 / 13.0));
  p = fma(z, p, TPA_LOG_COEF(1.0 / 11.0));
  p = fma(z, p, TPA_LOG_COEF(1.0 / 9.0));
  p = fma(z, p, TPA_LOG_COEF(1.0 / 7.0));
  p = fma(z, p, TPA_LOG_COEF(1.0 / 5.0));
  p = fma(z, p, TPA_LOG_COEF(1.0 / 3.0));
  const double s2 = s + s;
  const double fe = (double)e;
  // e ln2_hi + 2s carries the leading bits; 2s z p + e ln2_lo is the small tail
  return fma(fe, ln2_hi, s2) + fma(s2 * z, p, fe * ln2_lo);
}
