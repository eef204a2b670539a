// Quasi-Newton minimiser shared by the serving runtime (one per fit) and the single-fit
// session: BFGS with a strong-Wolfe line search in the unbounded variables x of the "asin"
// bound transforms (tf_pwa/variable.py:1029-1135), the scheme of scipy.optimize.minimize
// (method BFGS) that tf-pwa's fit_scipy drives (tf_pwa/fit.py:189-300).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "tpa_coupling.h"

namespace tpa_cpl {

inline double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

inline double norm_inf(const std::vector<double>& a) {
  double m = 0.0;
  for (double v : a) m = std::max(m, std::fabs(v));
  return m;
}

struct Trial {
  double a, f, d;
  std::vector<double> g;
};

// Strong-Wolfe line search (Nocedal & Wright, Numerical Optimization, Algorithms 3.5 and 3.6,
// cubic interpolation of eq. 3.59 with bisection as safeguard). obj(x, g) -> f. Returns false
// on failure.
template <class Obj>
bool line_search(Obj& obj, const std::vector<double>& x, const std::vector<double>& p, double f0,
                 double d0, double a1, Trial& out) {
  const double c1 = 1e-4, c2 = 0.9, amax = 1e10;
  const int n = (int)x.size();
  std::vector<double> xt(n);
  auto eval = [&](double a, Trial& t) {
    for (int i = 0; i < n; ++i) xt[i] = x[i] + a * p[i];
    t.a = a;
    t.f = obj(xt, t.g);
    t.d = dot(t.g, p);
  };
  auto armijo_fails = [&](const Trial& t) {
    return !std::isfinite(t.f) || t.f > f0 + c1 * t.a * d0;
  };
  // below this change of f the comparisons are rounding noise (|f| ~ 1e4..1e6 for NLLs)
  const double noise = 1e-14 * (1.0 + std::fabs(f0));
  auto zoom = [&](Trial lo, Trial hi) -> bool {
    for (int it = 0; it < 30; ++it) {
      if (std::fabs(hi.a - lo.a) * std::fabs(d0) <= noise) break;
      double a = 0.5 * (lo.a + hi.a);
      if (std::isfinite(hi.f) && std::isfinite(hi.d)) {
        const double d1 = lo.d + hi.d - 3.0 * (lo.f - hi.f) / (lo.a - hi.a);
        const double disc = d1 * d1 - lo.d * hi.d;
        if (disc >= 0.0) {
          const double d2 = std::copysign(std::sqrt(disc), hi.a - lo.a);
          const double den = hi.d - lo.d + 2.0 * d2;
          if (den != 0.0) {
            const double ac = hi.a - (hi.a - lo.a) * (hi.d + d2 - d1) / den;
            const double lo_b = std::min(lo.a, hi.a), hi_b = std::max(lo.a, hi.a);
            const double margin = 0.1 * (hi_b - lo_b);
            if (std::isfinite(ac) && ac > lo_b + margin && ac < hi_b - margin) a = ac;
          }
        }
      }
      if (std::fabs(hi.a - lo.a) <= 1e-16 * std::max(1.0, std::fabs(lo.a))) break;
      Trial t;
      eval(a, t);
      if (armijo_fails(t) || t.f >= lo.f) {
        hi = std::move(t);
      } else {
        if (std::fabs(t.d) <= -c2 * d0) {
          out = std::move(t);
          return true;
        }
        if (t.d * (hi.a - lo.a) >= 0.0) hi = lo;
        lo = std::move(t);
      }
    }
    // no strong-Wolfe point; accept a point of sufficient decrease if we have one
    if (lo.a > 0.0 && lo.f < f0) {
      out = std::move(lo);
      return true;
    }
    return false;
  };
  Trial prev;
  prev.a = 0.0;
  prev.f = f0;
  prev.d = d0;
  double a = a1;
  for (int i = 0; i < 40; ++i) {
    Trial t;
    eval(a, t);
    if (armijo_fails(t) || (i > 0 && t.f >= prev.f)) return zoom(std::move(prev), std::move(t));
    if (std::fabs(t.d) <= -c2 * d0) {
      out = std::move(t);
      return true;
    }
    if (t.d >= 0.0) return zoom(std::move(t), std::move(prev));
    prev = std::move(t);
    a = std::min(2.0 * a, amax);
  }
  return false;
}

struct BfgsResult {
  std::vector<double> y;  // physical values at the end
  double f = NAN;
  int status = 1;  // 0 converged (|g|_inf <= gtol in x), 1 iteration limit, 2 line search failed
  int64_t niter = 0;
};

// Minimise fy(y, gy) -> f from y0 (physical values). H0: the inverse Hessian in x to start
// from (n x n, row-major), or the identity when null; with H0 the first step is the full
// Newton step. maxiter <= 0: 200 n iterations.
template <class FY>
BfgsResult bfgs_minimize(const CouplingMap* S, const double* y0, FY&& fy, double gtol,
                         int maxiter, const std::vector<double>* H0) {
  const int n = S->nv;
  std::vector<double> yb(n), dydx(n), gy(n);
  auto obj = [&](const std::vector<double>& xv, std::vector<double>& gx) {
    to_y(S, xv.data(), yb.data(), dydx.data());
    const double f = fy(yb.data(), gy.data());
    gx.resize(n);
    for (int j = 0; j < n; ++j) gx[j] = gy[j] * dydx[j];
    return f;
  };
  std::vector<double> x(n), g, p(n), s(n), yv(n), Hy(n);
  for (int j = 0; j < n; ++j) x[j] = to_x(S, j, y0[j]);
  std::vector<double> H((size_t)n * n, 0.0);
  const bool newton_start = H0 != nullptr;
  if (H0)
    H = *H0;
  else
    for (int j = 0; j < n; ++j) H[(size_t)j * n + j] = 1.0;
  double f = obj(x, g);
  double f_old_old = f + std::sqrt(dot(g, g)) / 2.0;
  int status = 1;
  int64_t k = 0;
  const int64_t kmax = maxiter > 0 ? maxiter : 200 * (int64_t)n;
  if (norm_inf(g) <= gtol) status = 0;
  while (status == 1 && k < kmax) {
    for (int i = 0; i < n; ++i) {
      double acc = 0.0;
      for (int j = 0; j < n; ++j) acc -= H[(size_t)i * n + j] * g[j];
      p[i] = acc;
    }
    double d0 = dot(g, p);
    if (!(d0 < 0.0)) {  // not a descent direction: restart from steepest descent
      std::fill(H.begin(), H.end(), 0.0);
      for (int j = 0; j < n; ++j) H[(size_t)j * n + j] = 1.0;
      for (int j = 0; j < n; ++j) p[j] = -g[j];
      d0 = -dot(g, g);
    }
    // first trial step from the previous decrease (Nocedal & Wright eq. 3.60)
    double a1 = 1.0;
    const double guess = 1.01 * 2.0 * (f - f_old_old) / d0;
    if (!(newton_start && k == 0) && std::isfinite(guess) && guess > 0.0) a1 = std::min(1.0, guess);
    Trial t;
    if (!line_search(obj, x, p, f, d0, a1, t)) {
      status = 2;
      break;
    }
    for (int j = 0; j < n; ++j) {
      s[j] = t.a * p[j];
      x[j] += s[j];
      yv[j] = t.g[j] - g[j];
    }
    f_old_old = f;
    f = t.f;
    g = std::move(t.g);
    ++k;
    if (norm_inf(g) <= gtol) {
      status = 0;
      break;
    }
    const double ys = dot(yv, s);
    if (!(ys > 1e-12 * std::sqrt(dot(yv, yv) * dot(s, s)))) continue;  // keep H (curvature)
    const double rho = 1.0 / ys;
    for (int i = 0; i < n; ++i) {
      double acc = 0.0;
      for (int j = 0; j < n; ++j) acc += H[(size_t)i * n + j] * yv[j];
      Hy[i] = acc;
    }
    const double yHy = dot(yv, Hy);
    const double cs = rho * rho * yHy + rho;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        H[(size_t)i * n + j] += -rho * (s[i] * Hy[j] + Hy[i] * s[j]) + cs * s[i] * s[j];
  }
  BfgsResult r;
  r.niter = k;
  r.status = status;
  r.f = f;
  r.y.resize(n);
  to_y(S, x.data(), r.y.data(), dydx.data());
  return r;
}

}  // namespace tpa_cpl
