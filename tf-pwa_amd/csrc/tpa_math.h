// Device math for the likelihood kernels: complex numbers, Lorentz boosts, SU(2) frames,
// Wigner d-functions, barrier factors and Breit-Wigner lineshapes, all in FP64.
//
// Each function states the tf-pwa function whose arithmetic it reproduces; the operation
// order follows the reference so results agree to rounding.
#pragma once
#ifndef __HIPCC_RTC__  // hipRTC (the fused pass's runtime compile) provides these itself
#include <hip/hip_runtime.h>
#include <math.h>
#endif

#include "tpa_model.h"

#define TPA_DEV __device__ __forceinline__

struct cd {
  double r, i;
};

TPA_DEV cd c_make(double r, double i) { return cd{r, i}; }
TPA_DEV cd c_add(cd a, cd b) { return cd{a.r + b.r, a.i + b.i}; }
TPA_DEV cd c_sub(cd a, cd b) { return cd{a.r - b.r, a.i - b.i}; }
TPA_DEV cd c_mul(cd a, cd b) { return cd{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
TPA_DEV cd c_scale(cd a, double s) { return cd{a.r * s, a.i * s}; }
TPA_DEV cd c_conj(cd a) { return cd{a.r, -a.i}; }
TPA_DEV cd c_neg(cd a) { return cd{-a.r, -a.i}; }
TPA_DEV double c_abs2(cd a) { return a.r * a.r + a.i * a.i; }
// 1/a (used for the SU(2) diagonal elements; |a| is O(1) there)
TPA_DEV cd c_inv(cd a) {
  double d = a.r * a.r + a.i * a.i;
  return cd{a.r / d, -a.i / d};
}
TPA_DEV cd c_expi(double x) {
  double s, c;
  sincos(x, &s, &c);
  return cd{c, s};
}

// ---------------------------------------------------------------- 3-vectors (angle.py:13-95)
struct v3 {
  double x, y, z;
};
TPA_DEV v3 v_make(double x, double y, double z) { return v3{x, y, z}; }
TPA_DEV double v_dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
TPA_DEV double v_norm(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
TPA_DEV v3 v_scale(v3 a, double s) { return v3{a.x * s, a.y * s, a.z * s}; }
TPA_DEV v3 v_div(v3 a, double s) { return v3{a.x / s, a.y / s, a.z / s}; }
TPA_DEV v3 v_unit(v3 a) { return v_div(a, v_norm(a)); }
TPA_DEV v3 v_cross(v3 a, v3 b) {
  return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// Vector3.cross_unit (angle.py:57-67): fall back to a x (1 + b) for parallel inputs.
TPA_DEV v3 v_cross_unit(v3 a, v3 b) {
  v3 c = v_cross(a, b);
  if (v_norm(c) < 1.0e-14) c = v_cross(a, v3{1.0 + b.x, 1.0 + b.y, 1.0 + b.z});
  return v_unit(c);
}
TPA_DEV double v_angle_from(v3 v, v3 x, v3 y) { return atan2(v_dot(v, y), v_dot(v, x)); }

// --------------------------------------------------------- 4-vectors (angle.py:98-242)
struct v4 {
  double t, x, y, z;
};
TPA_DEV v3 lv_vect(v4 p) { return v3{p.x, p.y, p.z}; }
TPA_DEV v4 lv_add(v4 a, v4 b) { return v4{a.t + b.t, a.x + b.x, a.y + b.y, a.z + b.z}; }
TPA_DEV double lv_M(v4 p) {
  double m2 = p.t * p.t - p.x * p.x - p.y * p.y - p.z * p.z;
  return sqrt(fabs(m2));
}
// LorentzVector.boost (angle.py:148-163)
TPA_DEV v4 lv_boost(v4 p, v3 pb) {
  double beta2 = v_dot(pb, pb);
  double gamma = 1.0 / sqrt(1.0 - beta2);
  double bp = v_dot(pb, lv_vect(p));
  double gamma2 = beta2 > 1.0e-14 ? (gamma - 1.0) / beta2 : 0.0;
  double a = gamma2 * bp;
  double b = gamma * p.t;
  v4 r;
  r.x = (p.x + a * pb.x) + b * pb.x;
  r.y = (p.y + a * pb.y) + b * pb.y;
  r.z = (p.z + a * pb.z) + b * pb.z;
  r.t = gamma * (p.t + bp);
  return r;
}
// LorentzVector.rest_vector (angle.py:140-146): `other` in the rest frame of `frame`.
TPA_DEV v4 lv_rest(v4 frame, v4 other) {
  v3 pb = v3{-(frame.x / frame.t), -(frame.y / frame.t), -(frame.z / frame.t)};
  return lv_boost(other, pb);
}
// LorentzVector.omega (angle.py:198-213)
TPA_DEV double lv_omega(v4 p) {
  v3 pb = v3{p.x / p.t, p.y / p.t, p.z / p.t};
  double beta2 = v_dot(pb, pb);
  beta2 = beta2 < 1.0 ? beta2 : 0.0;
  double gamma = 1.0 / sqrt(1.0 - beta2);
  return acosh(gamma);
}

// EulerAngle.angle_zx_z_getx (angle.py:285-308)
TPA_DEV void angle_zx_z_getx(v3 z1, v3 x1, v3 z2, double* alpha, double* beta, v3* x2) {
  v3 u_z1 = v_unit(z1);
  v3 u_z2 = v_unit(z2);
  v3 u_y1 = v_cross_unit(z1, x1);
  v3 u_x1 = v_cross_unit(u_y1, z1);
  v3 u_yr = v_cross_unit(z1, z2);
  v3 u_xr = v_cross_unit(u_yr, z1);
  *alpha = v_angle_from(u_xr, u_x1, u_y1);
  *beta = v_angle_from(u_z2, u_z1, u_xr);
  *x2 = v_cross_unit(u_yr, u_z2);
}

// Python float modulo (floor mod) as used by the alpha folding of cal_angle.py:291-303
TPA_DEV double fmod_floor(double a, double b) {
  double r = fmod(a, b);
  if (r != 0.0 && ((r < 0.0) != (b < 0.0))) r += b;
  return r;
}

// ------------------------------------------------------------ SU(2) (angle.py:329-392)
struct su2 {
  cd a, b, c, d;  // [[a, b], [c, d]]
};
TPA_DEV su2 su2_mul(su2 x, su2 y) {
  su2 r;
  r.a = c_add(c_mul(x.a, y.a), c_mul(x.b, y.c));
  r.b = c_add(c_mul(x.a, y.b), c_mul(x.b, y.d));
  r.c = c_add(c_mul(x.c, y.a), c_mul(x.d, y.c));
  r.d = c_add(c_mul(x.c, y.b), c_mul(x.d, y.d));
  return r;
}
TPA_DEV su2 su2_inv(su2 x) { return su2{x.d, c_neg(x.b), c_neg(x.c), x.a}; }
TPA_DEV su2 su2_boost_z(double omega) {
  cd a = cd{exp(omega / 2.0), 0.0};
  return su2{c_inv(a), cd{0, 0}, cd{0, 0}, a};
}
TPA_DEV su2 su2_rot_z(double alpha) {
  cd a = c_expi(alpha / 2.0);
  return su2{c_inv(a), cd{0, 0}, cd{0, 0}, a};
}
TPA_DEV su2 su2_rot_y(double beta) {
  double s = sin(beta / 2.0), c = cos(beta / 2.0);
  return su2{cd{c, 0}, cd{-s, 0}, cd{s, 0}, cd{c, 0}};
}
// SU2M.get_euler_angle (angle.py:374-392)
TPA_DEV void su2_euler(su2 x, double* alpha, double* beta, double* gamma) {
  double cb = c_add(c_mul(x.a, x.d), c_mul(x.b, x.c)).r;
  cb = fmin(fmax(cb, -1.0), 1.0);
  *beta = acos(cb);
  double apg = atan2(x.d.i, x.d.r);
  double amg = -atan2(x.c.i, x.c.r);
  *alpha = apg + amg;
  *gamma = apg - amg;
}

// ------------------------------------------------------ Wigner d (dfun.py:118-218)
// d^{j/2}_{m/2, n/2}(beta) = sum_l w[l][m][n] sin^l(beta/2) cos^{j-l}(beta/2) with the weight
// table of small_d_weight; `w` points at the (j+1)^3 table of this j.
TPA_DEV double wigner_d(const double* __restrict__ w, int j, int m2, int n2, const double* sp,
                        const double* cp) {
  int im = (m2 + j) / 2, in = (n2 + j) / 2;
  double acc = 0.0;
  for (int l = 0; l <= j; ++l) acc += w[(l * (j + 1) + im) * (j + 1) + in] * (sp[l] * cp[j - l]);
  return acc;
}
TPA_DEV void half_angle_powers(double beta, int j, double* sp, double* cp) {
  double s = sin(0.5 * beta), c = cos(0.5 * beta);
  sp[0] = 1.0;
  cp[0] = 1.0;
  for (int l = 1; l <= j; ++l) {
    sp[l] = sp[l - 1] * s;
    cp[l] = cp[l - 1] * c;
  }
}
TPA_DEV int dweight_offset(int j) {
  int off = 0;
  for (int k = 0; k < j; ++k) off += (k + 1) * (k + 1) * (k + 1);
  return off;
}

// ------------------------------------- barrier factors and lineshapes (breit_wigner.py)
// Bprime_polynomial (breit_wigner.py:332-355): |theta_l(i w)|^2 coefficients, highest first.
TPA_DEV double bprime_poly(int l, double z) {
  switch (l) {
    case 0: return 1.0;
    case 1: return z + 1.0;
    case 2: return (z + 3.0) * z + 9.0;
    case 3: return ((z + 6.0) * z + 45.0) * z + 225.0;
    case 4: return (((z + 10.0) * z + 135.0) * z + 1575.0) * z + 11025.0;
    default: return ((((z + 15.0) * z + 315.0) * z + 6300.0) * z + 99225.0) * z + 893025.0;
  }
}
TPA_DEV double bprime_poly_deriv(int l, double z) {
  switch (l) {
    case 0: return 0.0;
    case 1: return 1.0;
    case 2: return 2.0 * z + 3.0;
    case 3: return (3.0 * z + 12.0) * z + 45.0;
    case 4: return ((4.0 * z + 30.0) * z + 270.0) * z + 1575.0;
    default: return (((5.0 * z + 60.0) * z + 945.0) * z + 12600.0) * z + 99225.0;
  }
}
// get_relative_p2 (amp/core.py:324-334)
TPA_DEV double rel_p2(double m0, double m1, double m2) {
  double s = m1 + m2, d = m1 - m2;
  double p = (m0 - s) * (m0 + s) * (m0 - d) * (m0 + d);
  double t = 2.0 * m0;
  return p / (t * t);
}
// get_relative_p (amp/core.py:311-321), m_eff = max(m0, m1 + m2)
TPA_DEV double rel_p(double m0, double m1, double m2) {
  double s = m1 + m2, d = m1 - m2;
  double me = m0 > s ? m0 : s;
  double p = (me - s) * (me + s) * (me - d) * (me + d);
  return sqrt(p) / (2.0 * me);
}

// BWR (breit_wigner.py:74-91) with Gamma (231-245) and Bprime (290-306); BW (56-71).
// Lineshape codes: bits 0-1 the propagator (0 none: 1; 1 BW, constant width; 2 BWR, running
// width), bit 2 (width_norm) multiplies the propagator by the width g0 (Particle.get_amp,
// tf_pwa/amp/core.py:399-418).
#define TPA_SHAPE_KIND(s) ((s) & 3)
#define TPA_SHAPE_NORM 4

TPA_DEV cd lineshape(int shape, double m, double m0, double g0, double q, double q0, int L,
                     double d) {
  const int kind = TPA_SHAPE_KIND(shape);
  if (kind == 0) return cd{1.0, 0.0};
  double gamma = g0;
  if (kind == 2) {
    double qq0 = q0 > 1.0e-15 ? pow(q / q0, (double)(2 * L + 1)) : 1.0;
    double mm0 = m0 / m;
    double qd0 = q0 * d, qd = q * d;
    double bp = sqrt(bprime_poly(L, qd0 * qd0)) / sqrt(bprime_poly(L, qd * qd));
    gamma = g0 * qq0 * mm0 * (bp * bp);
  }
  double x = m0 * m0 - m * m;
  double y = m0 * gamma;
  double s = x * x + y * y;
  const double nf = (shape & TPA_SHAPE_NORM) ? g0 : 1.0;
  return cd{nf * x / s, nf * y / s};
}

// Second-order forward derivatives in two variables (a: the pole mass m0, b: the width g0):
// value, gradient and Hessian of a real function, propagated exactly through each operation.
struct jet2 {
  double v, a, b, aa, ab, bb;
};
TPA_DEV jet2 j2_const(double x) { return jet2{x, 0.0, 0.0, 0.0, 0.0, 0.0}; }
TPA_DEV jet2 j2_add(jet2 u, jet2 w) {
  return jet2{u.v + w.v, u.a + w.a, u.b + w.b, u.aa + w.aa, u.ab + w.ab, u.bb + w.bb};
}
TPA_DEV jet2 j2_sub(jet2 u, jet2 w) {
  return jet2{u.v - w.v, u.a - w.a, u.b - w.b, u.aa - w.aa, u.ab - w.ab, u.bb - w.bb};
}
TPA_DEV jet2 j2_addc(jet2 u, double c) {
  u.v += c;
  return u;
}
TPA_DEV jet2 j2_scale(jet2 u, double c) {
  return jet2{u.v * c, u.a * c, u.b * c, u.aa * c, u.ab * c, u.bb * c};
}
TPA_DEV jet2 j2_mul(jet2 u, jet2 w) {
  return jet2{u.v * w.v,
              u.a * w.v + u.v * w.a,
              u.b * w.v + u.v * w.b,
              u.aa * w.v + 2.0 * u.a * w.a + u.v * w.aa,
              u.ab * w.v + u.a * w.b + u.b * w.a + u.v * w.ab,
              u.bb * w.v + 2.0 * u.b * w.b + u.v * w.bb};
}
// g(u) from g(u.v), g'(u.v), g''(u.v) (the chain rule to second order)
TPA_DEV jet2 j2_apply(jet2 u, double g, double g1, double g2) {
  return jet2{g, g1 * u.a, g1 * u.b, g2 * u.a * u.a + g1 * u.aa, g2 * u.a * u.b + g1 * u.ab,
              g2 * u.b * u.b + g1 * u.bb};
}
TPA_DEV jet2 j2_recip(jet2 u) {
  const double r = 1.0 / u.v;
  return j2_apply(u, r, -r * r, 2.0 * r * r * r);
}
TPA_DEV jet2 j2_sqrt(jet2 u) {
  const double r = sqrt(u.v);
  return j2_apply(u, r, 0.5 / r, -0.25 / (r * u.v));
}
TPA_DEV jet2 j2_powi(jet2 u, int n) {
  const double x = u.v;
  double p2 = 1.0;  // x^(n-2)
  for (int i = 0; i < n - 2; ++i) p2 *= x;
  if (n == 0) return j2_const(1.0);
  if (n == 1) return u;
  return j2_apply(u, p2 * x * x, n * p2 * x, n * (n - 1.0) * p2);
}
TPA_DEV jet2 j2_bprime_poly(int l, jet2 z) {
  static constexpr double cof[6][6] = {{1.0},
                                       {1.0, 1.0},
                                       {1.0, 3.0, 9.0},
                                       {1.0, 6.0, 45.0, 225.0},
                                       {1.0, 10.0, 135.0, 1575.0, 11025.0},
                                       {1.0, 15.0, 315.0, 6300.0, 99225.0, 893025.0}};
  const int L = l < 5 ? l : 5;
  jet2 r = j2_const(cof[L][0]);
  for (int i = 1; i <= L; ++i) r = j2_addc(j2_mul(r, z), cof[L][i]);
  return r;
}

// The lineshape as lineshape_grad computes it, with its first and second derivatives in
// (m0, g0): out[0] = B, out[1] = dB/dm0, out[2] = dB/dg0, out[3] = d2B/dm0^2,
// out[4] = d2B/dm0 dg0, out[5] = d2B/dg0^2 (the q0 clamp m_eff = max(m0, m1 + m2) of
// get_relative_p has zero derivatives on its clamped branch, as in the reference's tape).
TPA_DEV void lineshape_hess(int shape, double m, double m0v, double g0v, double q, double my,
                            double mz, int L, double d, cd* out) {
  const int kind = TPA_SHAPE_KIND(shape);
  if (kind == 0) {
    out[0] = cd{1.0, 0.0};
    for (int i = 1; i < 6; ++i) out[i] = cd{0.0, 0.0};
    return;
  }
  const jet2 m0 = jet2{m0v, 1.0, 0.0, 0.0, 0.0, 0.0};
  const jet2 g0 = jet2{g0v, 0.0, 1.0, 0.0, 0.0, 0.0};
  jet2 gamma = g0;
  if (kind == 2) {
    const double s = my + mz, df = my - mz;
    const jet2 me = m0v > s ? m0 : j2_const(s);
    const jet2 p = j2_mul(j2_mul(j2_addc(me, -s), j2_addc(me, s)),
                          j2_mul(j2_addc(me, -df), j2_addc(me, df)));
    const jet2 q0 = p.v > 0.0 ? j2_mul(j2_sqrt(p), j2_recip(j2_scale(me, 2.0)))
                              : j2_const(0.0);
    jet2 qq0 = j2_const(1.0);
    if (q0.v > 1.0e-15) qq0 = j2_powi(j2_scale(j2_recip(q0), q), 2 * L + 1);
    const jet2 qd0 = j2_scale(q0, d);
    const jet2 p0 = j2_bprime_poly(L, j2_mul(qd0, qd0));
    const double pz = bprime_poly(L, q * d * q * d);
    const jet2 bp2 = j2_scale(p0, 1.0 / pz);
    gamma = j2_mul(j2_mul(g0, qq0), j2_mul(j2_scale(m0, 1.0 / m), bp2));
  }
  const jet2 x = j2_addc(j2_mul(m0, m0), -m * m);
  const jet2 y = j2_mul(m0, gamma);
  const jet2 rs = j2_recip(j2_add(j2_mul(x, x), j2_mul(y, y)));
  jet2 br = j2_mul(x, rs), bi = j2_mul(y, rs);
  if (shape & TPA_SHAPE_NORM) {
    br = j2_mul(br, g0);
    bi = j2_mul(bi, g0);
  }
  out[0] = cd{br.v, bi.v};
  out[1] = cd{br.a, bi.a};
  out[2] = cd{br.b, bi.b};
  out[3] = cd{br.aa, bi.aa};
  out[4] = cd{br.ab, bi.ab};
  out[5] = cd{br.bb, bi.bb};
}

// Lineshape with its derivatives in the pole mass m0 and width g0 (the reverse-mode
// gradient of BWR / BW, tf_pwa/breit_wigner.py:56-91, 231-245, 290-306, with q0 from
// get_relative_p's m_eff = max(m0, m1+m2) clamp, whose clamped branch has zero derivative).
TPA_DEV void lineshape_grad(int shape, double m, double m0, double g0, double q, double my,
                            double mz, int L, double d, cd* B, cd* dBm, cd* dBg) {
  const int kind = TPA_SHAPE_KIND(shape);
  if (kind == 0) {
    *B = cd{1.0, 0.0};
    *dBm = *dBg = cd{0.0, 0.0};
    return;
  }
  double gamma = g0, dgm = 0.0, dgg = 1.0;
  if (kind == 2) {
    const double s = my + mz, df = my - mz;
    const double me = m0 > s ? m0 : s;
    const double p = (me - s) * (me + s) * (me - df) * (me + df);
    const double q0 = sqrt(p) / (2.0 * me);
    double dq0 = 0.0;
    if (m0 > s && p > 0.0) {
      const double dp = 2.0 * me * ((me * me - df * df) + (me * me - s * s));
      dq0 = dp / (4.0 * me * sqrt(p)) - q0 / me;
    }
    double qq0 = 1.0, dqq0 = 0.0;
    if (q0 > 1.0e-15) {
      qq0 = pow(q / q0, (double)(2 * L + 1));
      dqq0 = -(2.0 * L + 1.0) * qq0 / q0 * dq0;
    }
    const double qd0 = q0 * d, qd = q * d;
    const double p0 = bprime_poly(L, qd0 * qd0), pz = bprime_poly(L, qd * qd);
    const double bp = sqrt(p0) / sqrt(pz);
    const double bp2 = bp * bp;
    const double dbp2 = bprime_poly_deriv(L, qd0 * qd0) * 2.0 * q0 * d * d * dq0 / pz;
    const double mm0 = m0 / m;
    gamma = g0 * qq0 * mm0 * bp2;
    dgm = g0 * (dqq0 * mm0 * bp2 + qq0 * (1.0 / m) * bp2 + qq0 * mm0 * dbp2);
    dgg = qq0 * mm0 * bp2;
  } else {
    dgm = 0.0;
    dgg = 1.0;
  }
  const double x = m0 * m0 - m * m;
  const double y = m0 * gamma;
  const double ss = x * x + y * y;
  *B = cd{x / ss, y / ss};
  // B = 1 / D with D = x - i y; dB = -B^2 dD
  const cd B2 = c_mul(*B, *B);
  const cd dDm = cd{2.0 * m0, -(gamma + m0 * dgm)};
  const cd dDg = cd{0.0, -(m0 * dgg)};
  *dBm = c_neg(c_mul(B2, dDm));
  *dBg = c_neg(c_mul(B2, dDg));
  if (shape & TPA_SHAPE_NORM) {  // g0 B: d/dm0 = g0 dB/dm0, d/dg0 = B + g0 dB/dg0
    *dBg = c_add(*B, c_scale(*dBg, g0));
    *dBm = c_scale(*dBm, g0);
    *B = c_scale(*B, g0);
  }
}

// Neumaier-compensated summation: s + c carries the sum of the added terms to about twice the
// working precision.
TPA_DEV void neumaier_add(double& s, double& c, double y) {
  const double t = s + y;
  if (fabs(s) >= fabs(y))
    c += (s - t) + y;
  else
    c += (y - t) + s;
  s = t;
}
TPA_DEV void neumaier_merge(double& s, double& c, double s2, double c2) {
  const double t = s + s2;
  double err;
  if (fabs(s) >= fabs(s2))
    err = (s - t) + s2;
  else
    err = (s2 - t) + s;
  c = (c + c2) + err;
  s = t;
}
