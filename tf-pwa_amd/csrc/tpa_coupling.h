// Coupling map of the partial-wave coefficients and the host-side chain rules shared by the
// serving runtime (tpa_serve.cpp) and the single-fit likelihood session (tpa_fcn.cpp):
// c_k = wave_const_k * v_f0 * ... * v_f(NS-1) with polar (r e^{i phi}) or cartesian factors whose
// real parts are trainable variables or fixed values, the "asin" bound transforms of
// tf_pwa/variable.py:1029-1135, and the Wirtinger-calculus chain rules from derivatives in c to
// derivatives in the variables.
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <string>
#include <vector>

#include "tfpwa_amd.h"
#include "tpa_internal.h"

namespace tpa_cpl {

using cplx = std::complex<double>;
constexpr int NS = TPA_CPL_SLOTS;  // factors per partial wave (total + one per vertex)

struct CouplingMap {
  int P = 0, nf = 0, nv = 0;
  std::vector<int> wave_factor;  // [P][NS]
  std::vector<double> wave_const;
  std::vector<int> factor_polar, factor_var;  // [nf], [nf][2]
  std::vector<double> factor_fixed;           // [nf][2]
  std::vector<int> bound_kind;                // [nv]
  std::vector<double> bound_ab;               // [nv][2]
  std::vector<cplx> gram;                     // [P][P] MC normalisation (may be empty)

  // Validates and copies a C-ABI map (and the Gram matrix when given).
  int load(const TpaCouplingMap* map, const double* g) {
    if (!map) return tpa_set_error(TPA_ERR_ARG, "null pointer");
    if (map->P < 1 || map->n_factor < 1 || map->n_var < 1)
      return tpa_set_error(TPA_ERR_ARG, "bad coupling map sizes");
    if (!map->wave_factor || !map->wave_const || !map->factor_polar || !map->factor_var ||
        !map->factor_fixed || !map->bound_kind || !map->bound_ab)
      return tpa_set_error(TPA_ERR_ARG, "incomplete coupling map");
    P = map->P;
    nf = map->n_factor;
    nv = map->n_var;
    for (int i = 0; i < NS * P; ++i)
      if (map->wave_factor[i] < 0 || map->wave_factor[i] >= nf)
        return tpa_set_error(TPA_ERR_ARG, "wave factor index out of range");
    for (int i = 0; i < 2 * nf; ++i)
      if (map->factor_var[i] < -1 || map->factor_var[i] >= nv)
        return tpa_set_error(TPA_ERR_ARG, "factor variable index out of range");
    for (int j = 0; j < nv; ++j)
      if (map->bound_kind[j] < 0 || map->bound_kind[j] > 3)
        return tpa_set_error(TPA_ERR_ARG, "bad bound kind");
    wave_factor.assign(map->wave_factor, map->wave_factor + NS * P);
    wave_const.assign(map->wave_const, map->wave_const + P);
    factor_polar.assign(map->factor_polar, map->factor_polar + nf);
    factor_var.assign(map->factor_var, map->factor_var + 2 * nf);
    factor_fixed.assign(map->factor_fixed, map->factor_fixed + 2 * nf);
    bound_kind.assign(map->bound_kind, map->bound_kind + nv);
    bound_ab.assign(map->bound_ab, map->bound_ab + 2 * nv);
    gram.clear();
    if (g) {
      gram.resize((size_t)P * P);
      for (size_t i = 0; i < (size_t)P * P; ++i) gram[i] = cplx(g[2 * i], g[2 * i + 1]);
    }
    return TPA_OK;
  }
};

// ------------------------------------------------------------------ parameters -> coefficients
inline void factor_values(const CouplingMap* S, const double* y, cplx* v, cplx* dr, cplx* di) {
  for (int f = 0; f < S->nf; ++f) {
    const int vr = S->factor_var[2 * f], vi = S->factor_var[2 * f + 1];
    const double r = vr >= 0 ? y[vr] : S->factor_fixed[2 * f];
    const double i = vi >= 0 ? y[vi] : S->factor_fixed[2 * f + 1];
    if (S->factor_polar[f]) {
      const cplx e(std::cos(i), std::sin(i));
      v[f] = r * e;
      dr[f] = e;
      di[f] = cplx(0.0, 1.0) * r * e;
    } else {
      v[f] = cplx(r, i);
      dr[f] = 1.0;
      di[f] = cplx(0.0, 1.0);
    }
  }
}

// Bound transform y(x) and dy/dx (the "asin" bounds of tf_pwa/variable.py:1029-1135).
inline void to_y(const CouplingMap* S, const double* x, double* y, double* dydx) {
  for (int j = 0; j < S->nv; ++j) {
    const double a = S->bound_ab[2 * j], b = S->bound_ab[2 * j + 1], t = x[j];
    switch (S->bound_kind[j]) {
      case 1:
        y[j] = (b - a) * (std::sin(t) + 1.0) / 2.0 + a;
        dydx[j] = (b - a) * std::cos(t) / 2.0;
        break;
      case 2:
        y[j] = a + std::sqrt(t * t + 1.0) - 1.0;
        dydx[j] = t / std::sqrt(t * t + 1.0);
        break;
      case 3:
        y[j] = b + 1.0 - std::sqrt(t * t + 1.0);
        dydx[j] = -t / std::sqrt(t * t + 1.0);
        break;
      default:
        y[j] = t;
        dydx[j] = 1.0;
    }
  }
}

// Inverse of to_y. y is clamped into the bound first (tf_pwa/variable.py Bound.get_y2x), so a
// start value outside a one-sided range maps to its edge instead of NaN.
inline double to_x(const CouplingMap* S, int j, double y) {
  const double a = S->bound_ab[2 * j], b = S->bound_ab[2 * j + 1];
  switch (S->bound_kind[j]) {
    case 1: return std::asin(std::min(1.0, std::max(-1.0, 2.0 * (y - a) / (b - a) - 1.0)));
    case 2: {
      const double u = std::max(y, a) - a + 1.0;
      return std::sqrt(u * u - 1.0);
    }
    case 3: {
      const double u = b - std::min(y, b) + 1.0;
      return std::sqrt(u * u - 1.0);
    }
    default: return y;
  }
}

inline double d2y_dx2(const CouplingMap* S, int j, double t) {
  const double a = S->bound_ab[2 * j], b = S->bound_ab[2 * j + 1];
  switch (S->bound_kind[j]) {
    case 1: return -(b - a) * std::sin(t) / 2.0;
    case 2: return 1.0 / std::pow(t * t + 1.0, 1.5);
    case 3: return -1.0 / std::pow(t * t + 1.0, 1.5);
    default: return 0.0;
  }
}

// Inverse of A + tau I with the smallest tau >= 0 (doubling from 1e-3 max|A_ii|) for which the
// Cholesky factorisation succeeds (Nocedal & Wright, Algorithm 3.3). A is n x n, row-major.
inline void shifted_cholesky_inverse(const std::vector<double>& A, int n, std::vector<double>& inv) {
  double maxd = 0.0, mind = INFINITY;
  for (int i = 0; i < n; ++i) {
    maxd = std::max(maxd, std::fabs(A[(size_t)i * n + i]));
    mind = std::min(mind, A[(size_t)i * n + i]);
  }
  const double beta = 1e-3 * std::max(maxd, 1e-300);
  double tau = mind > 0.0 ? 0.0 : -mind + beta;
  std::vector<double> L((size_t)n * n);
  for (int it = 0; it < 200; ++it) {
    bool ok = true;
    std::fill(L.begin(), L.end(), 0.0);
    for (int j = 0; j < n && ok; ++j) {
      double d = A[(size_t)j * n + j] + tau;
      for (int k = 0; k < j; ++k) d -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
      if (!(d > 0.0) || !std::isfinite(d)) {
        ok = false;
        break;
      }
      const double ljj = std::sqrt(d);
      L[(size_t)j * n + j] = ljj;
      for (int i = j + 1; i < n; ++i) {
        double s = A[(size_t)i * n + j];
        for (int k = 0; k < j; ++k) s -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
        L[(size_t)i * n + j] = s / ljj;
      }
    }
    if (ok) break;
    tau = std::max(2.0 * tau, beta);
  }
  inv.assign((size_t)n * n, 0.0);
  std::vector<double> z(n);
  for (int col = 0; col < n; ++col) {  // solve L L^T x = e_col
    for (int i = 0; i < n; ++i) {
      double s = (i == col) ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * z[k];
      z[i] = s / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = z[i];
      for (int k = i + 1; k < n; ++k) s -= L[(size_t)k * n + i] * inv[(size_t)k * n + col];
      inv[(size_t)i * n + col] = s / L[(size_t)i * n + i];
    }
  }
}

// Gradient in y of a function whose Wirtinger derivative in the coefficients is F1
// (dF = 2 Re sum_k F1_k dc_k): the chain rule through the three factors of each c_k.
inline void grad_from_F1(const CouplingMap* S, const cplx* v, const cplx* dr, const cplx* di,
                  const cplx* F1, double* gy) {
  std::fill(gy, gy + S->nv, 0.0);
  for (int k = 0; k < S->P; ++k) {
    const cplx comb = S->wave_const[k] * F1[k];
    const int* w = &S->wave_factor[NS * k];
    for (int a = 0; a < NS; ++a) {
      cplx rest = comb;
      for (int b = 0; b < NS; ++b)
        if (b != a) rest *= v[w[b]];
      const int f = w[a];
      const int vr = S->factor_var[2 * f], vi = S->factor_var[2 * f + 1];
      if (vr >= 0) gy[vr] += 2.0 * (dr[f] * rest).real();
      if (vi >= 0) gy[vi] += 2.0 * (di[f] * rest).real();
    }
  }
}

inline void coefficients(const CouplingMap* S, const double* y, std::vector<cplx>& v,
                  std::vector<cplx>& dr, std::vector<cplx>& di, std::vector<cplx>& c) {
  v.resize(S->nf);
  dr.resize(S->nf);
  di.resize(S->nf);
  c.resize(S->P);
  factor_values(S, y, v.data(), dr.data(), di.data());
  for (int k = 0; k < S->P; ++k) {
    const int* w = &S->wave_factor[NS * k];
    cplx ck = S->wave_const[k];
    for (int a = 0; a < NS; ++a) ck *= v[w[a]];
    c[k] = ck;
  }
}

// MC normalisation through the Gram matrix: S = c^dagger I c and dS/dc_k = conj((I c)_k).
inline double mc_term(const CouplingMap* S, const std::vector<cplx>& c, std::vector<cplx>& Sk) {
  const int P = S->P;
  Sk.resize(P);
  double smc = 0.0;
  for (int k = 0; k < P; ++k) {
    cplx acc = 0.0;
    const cplx* row = &S->gram[(size_t)k * P];
    for (int l = 0; l < P; ++l) acc += row[l] * c[l];
    Sk[k] = std::conj(acc);
    smc += (std::conj(c[k]) * acc).real();
  }
  return smc;
}

// Hessian in y of a real function F of the coefficients from its Wirtinger derivatives
// F1_k = dF/dc_k, Fcc_kl = d2F/dc_k dc_l, Fcb_kl = d2F/dc_k dconj(c_l):
// H = 2 Re[J Fcc J^T + J Fcb J^H + sum_k F1_k d2c_k/dy dy] with J = dc/dy.
inline void hess_chain(const CouplingMap* S, const cplx* v, const cplx* dr, const cplx* di,
                       const cplx* F1, const cplx* Fcc, const cplx* Fcb, double* H) {
  const int P = S->P, nv = S->nv;
  // J[a][k] = d c_k / d y_a
  std::vector<cplx> J((size_t)nv * P, 0.0);
  for (int k = 0; k < P; ++k) {
    const int* w = &S->wave_factor[NS * k];
    for (int a = 0; a < NS; ++a) {
      cplx rest = S->wave_const[k];
      for (int b = 0; b < NS; ++b)
        if (b != a) rest *= v[w[b]];
      const int f = w[a];
      const int vr = S->factor_var[2 * f], vi = S->factor_var[2 * f + 1];
      if (vr >= 0) J[(size_t)vr * P + k] += dr[f] * rest;
      if (vi >= 0) J[(size_t)vi * P + k] += di[f] * rest;
    }
  }
  std::vector<cplx> X((size_t)nv * P), Y((size_t)nv * P);
  for (int a = 0; a < nv; ++a)
    for (int l = 0; l < P; ++l) {
      cplx xa = 0.0, ya = 0.0;
      for (int k = 0; k < P; ++k) {
        xa += J[(size_t)a * P + k] * Fcc[(size_t)k * P + l];
        ya += J[(size_t)a * P + k] * Fcb[(size_t)k * P + l];
      }
      X[(size_t)a * P + l] = xa;
      Y[(size_t)a * P + l] = ya;
    }
  std::vector<cplx> M((size_t)nv * nv, 0.0);
  for (int a = 0; a < nv; ++a)
    for (int b = 0; b < nv; ++b) {
      cplx acc = 0.0;
      for (int l = 0; l < P; ++l)
        acc += X[(size_t)a * P + l] * J[(size_t)b * P + l] +
               Y[(size_t)a * P + l] * std::conj(J[(size_t)b * P + l]);
      M[(size_t)a * nv + b] = acc;
    }
  // sum_k F1_k d2c_k/dy dy: one polar factor twice, or two different factors once each
  for (int k = 0; k < P; ++k) {
    const cplx CF = S->wave_const[k] * F1[k];
    const int* w = &S->wave_factor[NS * k];
    for (int a = 0; a < NS; ++a) {
      const int f = w[a];
      const int fr = S->factor_var[2 * f], fi = S->factor_var[2 * f + 1];
      cplx rest = CF;
      for (int b = 0; b < NS; ++b)
        if (b != a) rest *= v[w[b]];
      if (S->factor_polar[f]) {
        const cplx e = dr[f];  // e^{i phi}
        if (fr >= 0 && fi >= 0) {
          M[(size_t)fr * nv + fi] += cplx(0.0, 1.0) * e * rest;
          M[(size_t)fi * nv + fr] += cplx(0.0, 1.0) * e * rest;
        }
        if (fi >= 0) M[(size_t)fi * nv + fi] += -v[f] * rest;
      }
      for (int b = 0; b < NS; ++b) {
        if (b == a) continue;
        const int g = w[b];
        cplx rest2 = CF;
        for (int t = 0; t < NS; ++t)
          if (t != a && t != b) rest2 *= v[w[t]];
        const int va[2] = {S->factor_var[2 * f], S->factor_var[2 * f + 1]};
        const int vb[2] = {S->factor_var[2 * g], S->factor_var[2 * g + 1]};
        const cplx da[2] = {dr[f], di[f]}, db[2] = {dr[g], di[g]};
        for (int p = 0; p < 2; ++p)
          for (int q = 0; q < 2; ++q)
            if (va[p] >= 0 && vb[q] >= 0) M[(size_t)va[p] * nv + vb[q]] += da[p] * db[q] * rest2;
      }
    }
  }
  for (int a = 0; a < nv; ++a)
    for (int b = 0; b < nv; ++b)
      H[(size_t)a * nv + b] = (M[(size_t)a * nv + b] + M[(size_t)b * nv + a]).real();
}

}  // namespace tpa_cpl
