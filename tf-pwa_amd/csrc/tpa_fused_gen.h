// Source generator of the fused evaluation pass: one HIP kernel specialised to a three-body
// model description (TpaModel) that computes, per event, every partial wave from the angle
// cache (Wigner D of both vertices, CG x barrier factors, the lineshape, the contraction over
// the resonance helicity), the amplitude, f, the weighted (clip_log) sum and the adjoint sums
// G_k, without a partial-wave cache in memory. The spins, helicity lists, LS lists, CG values
// and d-function weights become literals of the generated code, so every intermediate stays in
// registers; coefficients, lineshape parameters and barrier clamp flags are kernel arguments.
//
// Arithmetic follows the streaming build of the same waves (k_build_pw): the product of the
// vertex factors summed over the resonance helicity, times the lineshape, in the frame of the
// wave's topology. Aligned D factors of spin > 0 finals are applied to the per-topology sums
// instead of to each wave: A = sum_t R_t (sum_{k in t} c_k h'_k) with R_t the aligned rotation
// of topology t, and G_k = sum_e u sum_slot h'_k,slot B_t,slot with B_t = R_t^T conj(A).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <sstream>
#include <string>
#include <vector>

#include "tpa_model.h"

namespace tpa_fused {

// small_d_weight (tf_pwa/dfun.py:118-143): d^{j/2}_{m/2, n/2}(b) = sum_l w_l s^l c^(j-l) with
// s, c the sine and cosine of b/2.
inline std::vector<double> dweights(int j, int m, int n) {
  auto f = [](int x) {
    double r = 1.0;
    for (int i = 2; i <= (x >> 1); ++i) r *= i;
    return r;
  };
  std::vector<double> w(j + 1, 0.0);
  for (int k = std::max(0, n - m); k <= std::min(j - m, j + n); k += 2) {
    const int l = (2 * k + (m - n)) / 2;
    const double sign = (((k + m - n) / 2) % 2 == 0) ? 1.0 : -1.0;
    double tmp = sign * std::sqrt(1.0 * f(j + m) * f(j - m) * f(j + n) * f(j - n));
    tmp /= f(j - m - k) * f(j + n - k) * f(k + m - n) * f(k);
    w[l] = tmp;
  }
  return w;
}

inline std::string lit(double x) {
  char b[64];
  std::snprintf(b, sizeof b, "%.17g", x);
  std::string s(b);
  if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
  return s;
}

// Kernel arguments (by value): pointers and sizes, then the coefficients c_k (re, im), the
// lineshape parameters m0, g0, q0 per chain, the mode and gradient switches and the barrier
// clamp flags per chain (bit s: act1[s], bit 16 + s: act2[s]). The host lays the same bytes
// out (fused_args_layout).
struct ArgsLayout {
  int P, NC;
  size_t off_c, off_m0, off_g0, off_q0, off_mode, off_grad, off_act, size;
};
inline ArgsLayout fused_args_layout(int P, int NC) {
  ArgsLayout L;
  L.P = P;
  L.NC = NC;
  L.off_c = 40;
  L.off_m0 = L.off_c + 16 * (size_t)P;
  L.off_g0 = L.off_m0 + 8 * (size_t)NC;
  L.off_q0 = L.off_g0 + 8 * (size_t)NC;
  L.off_mode = L.off_q0 + 8 * (size_t)NC;
  L.off_grad = L.off_mode + 4;
  L.off_act = L.off_grad + 4;
  L.size = (L.off_act + 4 * (size_t)NC + 7) / 8 * 8;
  return L;
}

class Gen {
 public:
  Gen(const TpaModel& M, const double* cg, int n_cg) : M_(M), cg_(cg, cg + n_cg) {}

  std::string source() {
    prelude();
    event_function();
    kernel();
    return o_.str();
  }

 private:
  const TpaModel& M_;
  std::vector<double> cg_;
  std::ostringstream o_;

  int nfh(int f) const { return M_.n_fh[f]; }
  int n_slot() const { return M_.n_top * nfh(0) * nfh(1) * nfh(2); }
  // slot index of (iA, j0, j1, j2), the order of k_build_pw's lambda index
  int slot(int iA, const int j[3]) const {
    return ((iA * nfh(0) + j[0]) * nfh(1) + j[1]) * nfh(2) + j[2];
  }
  void unslot(int s, int* iA, int j[3]) const {
    j[2] = s % nfh(2);
    s /= nfh(2);
    j[1] = s % nfh(1);
    s /= nfh(1);
    j[0] = s % nfh(0);
    *iA = s / nfh(0);
  }
  static std::string cvar(const std::string& p, int t) { return p + std::to_string(t); }

  void prelude() {
    o_ << "// generated by tpa_fused_gen.h for one model description\n"
          "#include \"tpa_math.h\"\n"
          "#ifndef TPA_FUSED_KEEP\n"
          "#define TPA_FUSED_KEEP(x) asm volatile(\"\" : \"+v\"(x))\n"
          "#endif\n";
    o_ << "#define TPA_P " << M_.P << "\n#define TPA_NC " << M_.n_chain << "\n#define TPA_NSLOT "
       << n_slot() << "\n#define TPA_LAM " << n_slot() << "\n";
    o_ << "struct FusedArgs {\n  const double* ang;\n  const double* w;\n  double* partials;\n"
          "  long long N, S;\n  double c[2 * TPA_P];\n  double m0[TPA_NC], g0[TPA_NC], "
          "q0[TPA_NC];\n  int mode, grad;\n  int act[TPA_NC];\n};\n";
  }

  // e^{i m a / 2} for the m in `ms` from the base E = e^{i a/2} (powers by multiplication).
  void phases(const std::string& name, const std::string& angle, int mmax, bool keep) {
    o_ << "    cd " << name << "_1 = c_expi(0.5 * " << angle << ");\n";
    if (keep) o_ << "    TPA_FUSED_KEEP(" << name << "_1.r); TPA_FUSED_KEEP(" << name << "_1.i);\n";
    for (int m = 2; m <= mmax; ++m)
      o_ << "    const cd " << name << "_" << m << " = c_mul(" << name << "_" << (m - 1) << ", "
         << name << "_1);\n";
  }
  static std::string phase(const std::string& name, int m) {
    if (m == 0) return "c_make(1.0, 0.0)";
    if (m > 0) return name + "_" + std::to_string(m);
    return "c_conj(" + name + "_" + std::to_string(-m) + ")";
  }
  void half_powers(const std::string& name, const std::string& angle, int jmax, bool keep) {
    o_ << "    double " << name << "s_1, " << name << "c_1;\n    sincos(0.5 * " << angle
       << ", &" << name << "s_1, &" << name << "c_1);\n";
    if (keep) o_ << "    TPA_FUSED_KEEP(" << name << "s_1); TPA_FUSED_KEEP(" << name << "c_1);\n";
    for (int l = 2; l <= jmax; ++l) {
      o_ << "    const double " << name << "s_" << l << " = " << name << "s_" << (l - 1) << " * "
         << name << "s_1;\n";
      o_ << "    const double " << name << "c_" << l << " = " << name << "c_" << (l - 1) << " * "
         << name << "c_1;\n";
    }
  }
  // d^{j/2}_{m/2,n/2} as a literal polynomial in the half-angle powers of `name`
  static std::string dpoly(const std::string& name, int j, int m, int n) {
    const std::vector<double> w = dweights(j, m, n);
    std::string e;
    for (int l = 0; l <= j; ++l) {
      if (w[l] == 0.0) continue;
      std::string term = lit(w[l]);
      if (l > 0) term += " * " + name + "s_" + std::to_string(l);
      if (j - l > 0) term += " * " + name + "c_" + std::to_string(j - l);
      e += (e.empty() ? "" : " + ") + term;
    }
    return e.empty() ? "0.0" : "(" + e + ")";
  }
  // sqrt(q2)^l as products
  static std::string qpow(const std::string& q2, int l) {
    if (l == 0) return "1.0";
    std::string e = (l % 2) ? "sqrt(" + q2 + ")" : "";
    for (int i = 0; i < l / 2; ++i) e += (e.empty() ? "" : " * ") + q2;
    return "(" + e + ")";
  }

  // Per event: the per-topology sums A'_t (forward) or the adjoint accumulation (backward).
  void chain_waves(int c, bool backward) {
    const TpaChain& C = M_.chain[c];
    const int t = C.topo;
    const TpaTopo& T = M_.topo[t];
    const int ja = M_.top_j2, jr = C.jr2;
    const int fx = T.x, fy = C.r_y_first ? T.y : T.z, fz = C.r_y_first ? T.z : T.y;
    const std::string tv = "t" + std::to_string(t) + "_";
    const std::string cv = "c" + std::to_string(c) + "_";
    const std::string v1 = tv + (C.top_r_first ? "R" : "X");
    const std::string v2 = tv + ((fy == T.y) ? "Y" : "Z");
    o_ << "  {  // chain " << c << "\n";
    // barrier factors and lineshape
    o_ << "    const double q1sq = " << (C.top_r_first ? "rel_p2(mtop, " + tv + "mR, mf" +
                                                             std::to_string(fx) + ")"
                                                       : "rel_p2(mtop, mf" + std::to_string(fx) +
                                                             ", " + tv + "mR)")
       << ";\n";
    o_ << "    const double q2sq = rel_p2(" << tv << "mR, mf" << fy << ", mf" << fz << ");\n";
    o_ << "    const int act = a.act[" << c << "];\n";
    const double dd = C.d * C.d;
    for (int s = 0; s < C.n_ls1; ++s)
      o_ << "    const double bf1_" << s << " = ((act >> " << s << ") & 1) ? "
         << qpow("q1sq", C.l1[s]) << " / sqrt(bprime_poly(" << C.l1[s] << ", q1sq * "
         << lit(dd) << ")) : " << qpow("q1sq", C.l1[s]) << ";\n";
    for (int s = 0; s < C.n_ls2; ++s)
      o_ << "    const double bf2_" << s << " = ((act >> " << (16 + s) << ") & 1) ? "
         << qpow("q2sq", C.l2[s]) << " / sqrt(bprime_poly(" << C.l2[s] << ", q2sq * "
         << lit(dd) << ")) : " << qpow("q2sq", C.l2[s]) << ";\n";
    if (C.shape)
      o_ << "    const cd bw = lineshape(" << C.shape << ", " << tv << "mR, a.m0[" << c
         << "], a.g0[" << c << "], rel_p(" << tv << "mR, mf" << fy << ", mf" << fz << "), a.q0["
         << c << "], " << C.bw_l << ", " << lit(C.d) << ");\n";
    else
      o_ << "    const cd bw = c_make(1.0, 0.0);\n";
    // vertex factor products P(iA, iR, slot) and the partial sums over the resonance helicity
    const int nX = nfh(fx);
    const int nb1 = C.top_r_first ? C.n_r : nX;
    const int nc1 = C.top_r_first ? nX : C.n_r;
    const int nb2 = nfh(fy), nc2 = nfh(fz);
    std::vector<std::string> D1seen, D2seen;
    auto d1name = [&](int iA, int d1) {
      const std::string n = cv + "D1_" + std::to_string(iA) + "_" + std::to_string(d1 + ja);
      if (std::find(D1seen.begin(), D1seen.end(), n) == D1seen.end()) {
        D1seen.push_back(n);
        o_ << "    const cd " << n << " = c_scale(" << phase(v1 + "e", M_.hel_top[iA]) << ", "
           << dpoly(v1, ja, M_.hel_top[iA], d1) << ");\n";
      }
      return n;
    };
    auto d2name = [&](int iR, int d2) {
      const std::string n = cv + "D2_" + std::to_string(iR) + "_" + std::to_string(d2 + jr);
      if (std::find(D2seen.begin(), D2seen.end(), n) == D2seen.end()) {
        D2seen.push_back(n);
        o_ << "    const cd " << n << " = c_scale(" << phase(v2 + "e", C.hel_r[iR]) << ", "
           << dpoly(v2, jr, C.hel_r[iR], d2) << ");\n";
      }
      return n;
    };
    const int ns = n_slot();
    for (int sl = 0; sl < ns; ++sl) {
      int iA, j[3];
      unslot(sl, &iA, j);
      const int jX = j[fx];
      const int lX = M_.hel_f[fx][jX];
      const int ib2 = j[fy], ic2 = j[fz];
      const int d2 = M_.hel_f[fy][ib2] - M_.hel_f[fz][ic2];
      if (d2 > jr || d2 < -jr) continue;
      // T_s1 = sum_iR cg1(s1, iR, jX) D1[iA][d1] D2[iR][d2]
      std::vector<std::string> terms(C.n_ls1);
      for (int iR = 0; iR < C.n_r; ++iR) {
        const int lR = C.hel_r[iR];
        const int d1 = C.top_r_first ? lR - lX : lX - lR;
        if (d1 > ja || d1 < -ja) continue;
        const int ib1 = C.top_r_first ? iR : jX;
        const int ic1 = C.top_r_first ? jX : iR;
        bool any = false;
        for (int s1 = 0; s1 < C.n_ls1; ++s1)
          if (cg_[C.cg1_off + (s1 * nb1 + ib1) * nc1 + ic1] != 0.0) any = true;
        if (!any) continue;
        const std::string pn = cv + "P_" + std::to_string(sl) + "_" + std::to_string(iR);
        const std::string n1 = d1name(iA, d1);  // declares the factors before the product
        const std::string n2 = d2name(iR, d2);
        o_ << "    const cd " << pn << " = c_mul(" << n1 << ", " << n2 << ");\n";
        for (int s1 = 0; s1 < C.n_ls1; ++s1) {
          const double g = cg_[C.cg1_off + (s1 * nb1 + ib1) * nc1 + ic1];
          if (g == 0.0) continue;
          terms[s1] += (terms[s1].empty() ? "" : ", ") + std::string("{") + pn + ", " + lit(g) +
                       "}";
        }
      }
      for (int s1 = 0; s1 < C.n_ls1; ++s1) {
        // U = bw * bf1 * sum_iR cg1 P
        std::string sum = "c_make(0.0, 0.0)";
        {
          // expand the {P, g} list into explicit multiply-adds
          std::string rest = terms[s1];
          std::string acc;
          size_t pos = 0;
          while (pos < rest.size()) {
            const size_t a0 = rest.find('{', pos);
            if (a0 == std::string::npos) break;
            const size_t a1 = rest.find('}', a0);
            const std::string item = rest.substr(a0 + 1, a1 - a0 - 1);
            const size_t comma = item.find(", ");
            const std::string pv = item.substr(0, comma), gv = item.substr(comma + 2);
            acc = acc.empty() ? "c_scale(" + pv + ", " + gv + ")"
                              : "c_add(" + acc + ", c_scale(" + pv + ", " + gv + "))";
            pos = a1 + 1;
          }
          if (!acc.empty()) sum = acc;
        }
        if (sum == "c_make(0.0, 0.0)") continue;
        const std::string un = cv + "U_" + std::to_string(s1) + "_" + std::to_string(sl);
        o_ << "    const cd " << un << " = c_scale(c_mul(bw, " << sum << "), bf1_" << s1
           << ");\n";
        for (int s2 = 0; s2 < C.n_ls2; ++s2) {
          const double g2 = cg_[C.cg2_off + (s2 * nb2 + ib2) * nc2 + ic2];
          if (g2 == 0.0) continue;
          const int k = C.k0 + s1 * C.n_ls2 + s2;
          const std::string hs = "bf2_" + std::to_string(s2) + " * " + lit(g2);
          if (!backward) {
            // A'_t[slot] += c_k h'_k,slot
            o_ << "    " << tv << "Ap[" << sl << "] = c_add(" << tv << "Ap[" << sl
               << "], c_scale(c_mul(c_make(a.c[" << 2 * k << "], a.c[" << 2 * k + 1 << "]), "
               << un << "), " << hs << "));\n";
          } else {
            // G_k += u h'_k,slot B_t[slot]
            o_ << "    { const cd x = c_scale(c_mul(" << un << ", " << tv << "B[" << sl
               << "]), u * " << hs << "); gr[" << k << "] += x.r; gi[" << k << "] += x.i; }\n";
          }
        }
      }
    }
    o_ << "  }\n";
  }

  struct TopoNeeds {
    bool R = false, X = false, Y = false, Z = false;
    int jR = 0, jX = 0, jY = 0, jZ = 0;  // highest 2J of the D functions on each angle pair
  };
  TopoNeeds needs(int t) const {
    TopoNeeds n;
    for (int c = 0; c < M_.n_chain; ++c) {
      const TpaChain& C = M_.chain[c];
      if (C.topo != t) continue;
      const TpaTopo& T = M_.topo[t];
      const int fy = C.r_y_first ? T.y : T.z;
      if (C.top_r_first) {
        n.R = true;
        n.jR = std::max(n.jR, M_.top_j2);
      } else {
        n.X = true;
        n.jX = std::max(n.jX, M_.top_j2);
      }
      if (fy == T.y) {
        n.Y = true;
        n.jY = std::max(n.jY, C.jr2);
      } else {
        n.Z = true;
        n.jZ = std::max(n.jZ, C.jr2);
      }
    }
    return n;
  }
  bool aligned(int t, int f) const { return M_.topo[t].aligned[f] && M_.final_j2[f] != 0; }

  // Loads and trig of every topology; keep = true marks them opaque so that the backward
  // sweep recomputes the wave factors from them instead of keeping those factors live.
  void load_topologies(bool keep) {
    const long gb = (long)TPA_NF_TOPO * M_.n_topo;
    o_ << "  double mtop = ang[" << gb << "LL * S + e];\n";
    for (int f = 0; f < 3; ++f)
      o_ << "  double mf" << f << " = ang[" << (gb + 1 + f) << "LL * S + e];\n";
    if (keep)
      o_ << "  TPA_FUSED_KEEP(mtop); TPA_FUSED_KEEP(mf0); TPA_FUSED_KEEP(mf1); "
            "TPA_FUSED_KEEP(mf2);\n";
    for (int t = 0; t < M_.n_topo; ++t) {
      const TopoNeeds n = needs(t);
      const std::string tv = "t" + std::to_string(t) + "_";
      const long b = (long)TPA_NF_TOPO * t;
      o_ << "  double " << tv << "mR = ang[" << b + TPA_F_MR << "LL * S + e];\n";
      if (keep) o_ << "  TPA_FUSED_KEEP(" << tv << "mR);\n";
      auto pair = [&](bool need, const char* nm, int fld, int j) {
        if (!need) return;
        o_ << "  const double " << tv << nm << "a = ang[" << b + fld << "LL * S + e], " << tv
           << nm << "b = ang[" << b + fld + 1 << "LL * S + e];\n";
        (void)j;
      };
      pair(n.R, "R", TPA_F_AR, n.jR);
      pair(n.X, "X", TPA_F_AX, n.jX);
      pair(n.Y, "Y", TPA_F_AY, n.jY);
      pair(n.Z, "Z", TPA_F_AZ, n.jZ);
      for (int f = 0; f < 3; ++f) {
        if (!aligned(t, f)) continue;
        for (int q = 0; q < 3; ++q)
          o_ << "  const double " << tv << "al" << f << "_" << q << " = ang["
             << b + TPA_F_AL + 3 * f + q << "LL * S + e];\n";
      }
    }
  }
  // trig bases (phases and half-angle powers) of the angle pairs a topology's chains use
  void trig(int t, bool keep) {
    const TopoNeeds n = needs(t);
    const std::string tv = "t" + std::to_string(t) + "_";
    auto pair = [&](bool need, const char* nm, int j) {
      if (!need) return;
      const std::string base = tv + nm;
      phases(base + "e", base + "a", j, keep);
      half_powers(base, base + "b", j, keep);
    };
    pair(n.R, "R", n.jR);
    pair(n.X, "X", n.jX);
    pair(n.Y, "Y", n.jY);
    pair(n.Z, "Z", n.jZ);
  }
  // the aligned rotation R_t[slot][lambda] of topology t as a list of nonzero products
  void rotation_terms(int t, std::vector<std::pair<std::pair<int, int>, std::string>>* out,
                      bool keep) {
    const std::string tv = "t" + std::to_string(t) + "_";
    const int ns = n_slot();
    for (int f = 0; f < 3; ++f) {
      if (!aligned(t, f)) continue;
      const int jf = M_.final_j2[f];
      const std::string al = tv + "al" + std::to_string(f) + "_";
      phases(al + "ea", al + "0", jf, keep);
      phases(al + "eg", al + "2", jf, keep);
      half_powers(al + "b", al + "1", jf, keep);
      for (int i = 0; i < nfh(f); ++i)
        for (int i2 = 0; i2 < nfh(f); ++i2) {
          const int m1 = M_.hel_f[f][i], m2 = M_.hel_f[f][i2];
          o_ << "    const cd " << al << "Da_" << i << "_" << i2 << " = c_scale(c_mul("
             << phase(al + "ea", m1) << ", " << phase(al + "eg", m2) << "), "
             << dpoly(al + "b", jf, m1, m2) << ");\n";
        }
    }
    for (int sl = 0; sl < ns; ++sl) {
      int iA, j[3];
      unslot(sl, &iA, j);
      // every output lambda reachable from this slot
      int lo[3], hi[3];
      for (int f = 0; f < 3; ++f) {
        lo[f] = aligned(t, f) ? 0 : j[f];
        hi[f] = aligned(t, f) ? nfh(f) : j[f] + 1;
      }
      int o[3];
      for (o[0] = lo[0]; o[0] < hi[0]; ++o[0])
        for (o[1] = lo[1]; o[1] < hi[1]; ++o[1])
          for (o[2] = lo[2]; o[2] < hi[2]; ++o[2]) {
            std::string fac;
            for (int f = 0; f < 3; ++f) {
              if (!aligned(t, f)) continue;
              const std::string d = tv + "al" + std::to_string(f) + "_Da_" +
                                    std::to_string(j[f]) + "_" + std::to_string(o[f]);
              fac = fac.empty() ? d : "c_mul(" + fac + ", " + d + ")";
            }
            out->push_back({{sl, slot(iA, o)}, fac});
          }
    }
  }

  void event_function() {
    const int ns = n_slot();
    // forward: A[lambda]; returns nothing else
    o_ << "__device__ __forceinline__ void tpa_fused_amp(const FusedArgs& a, const double* "
          "__restrict__ ang, long long S, long long e, cd* __restrict__ A) {\n";
    load_topologies(false);
    for (int l = 0; l < ns; ++l) o_ << "  A[" << l << "] = c_make(0.0, 0.0);\n";
    for (int t = 0; t < M_.n_topo; ++t) {
      const std::string tv = "t" + std::to_string(t) + "_";
      o_ << "  {  // topology " << t << "\n";
      trig(t, false);
      o_ << "    cd " << tv << "Ap[" << ns << "];\n    for (int s = 0; s < " << ns
         << "; ++s) " << tv << "Ap[s] = c_make(0.0, 0.0);\n";
      for (int c = 0; c < M_.n_chain; ++c)
        if (M_.chain[c].topo == t) chain_waves(c, false);
      std::vector<std::pair<std::pair<int, int>, std::string>> rt;
      rotation_terms(t, &rt, false);
      for (auto& x : rt) {
        const int sl = x.first.first, lam = x.first.second;
        o_ << "    A[" << lam << "] = c_add(A[" << lam << "], "
           << (x.second.empty() ? tv + "Ap[" + std::to_string(sl) + "]"
                                : "c_mul(" + tv + "Ap[" + std::to_string(sl) + "], " + x.second +
                                      ")")
           << ");\n";
      }
      o_ << "  }\n";
    }
    o_ << "}\n";
    // backward: G_k += u sum_slot h'_k,slot B_t[slot], B_t = R_t^T conj(A)
    o_ << "__device__ __forceinline__ void tpa_fused_adjoint(const FusedArgs& a, const double* "
          "__restrict__ ang, long long S, long long e, const cd* __restrict__ A, double u, "
          "double* __restrict__ gr, double* __restrict__ gi) {\n";
    load_topologies(true);
    for (int t = 0; t < M_.n_topo; ++t) {
      const std::string tv = "t" + std::to_string(t) + "_";
      o_ << "  {  // topology " << t << "\n";
      trig(t, true);
      std::vector<std::pair<std::pair<int, int>, std::string>> rt;
      rotation_terms(t, &rt, true);
      o_ << "    cd " << tv << "B[" << ns << "];\n    for (int s = 0; s < " << ns << "; ++s) "
         << tv << "B[s] = c_make(0.0, 0.0);\n";
      for (auto& x : rt) {
        const int sl = x.first.first, lam = x.first.second;
        o_ << "    " << tv << "B[" << sl << "] = c_add(" << tv << "B[" << sl << "], "
           << (x.second.empty() ? "c_conj(A[" + std::to_string(lam) + "])"
                                : "c_mul(c_conj(A[" + std::to_string(lam) + "]), " + x.second +
                                      ")")
           << ");\n";
      }
      for (int c = 0; c < M_.n_chain; ++c)
        if (M_.chain[c].topo == t) chain_waves(c, true);
      o_ << "  }\n";
    }
    o_ << "}\n";
  }

  void kernel() {
    const int P = M_.P, ns = n_slot();
    o_ << R"(
// One pass over events [0, N) of the angle cache ang ([field][S]) with weights w: the sum of
// the mode (0: w clip_log f, 1: w f) and, with a.grad, the adjoint sums G_k; per-block
// partials slot-major (slot j of block b at partials[j * gridDim.x + b]) for k_reduce.
extern "C" __global__ __launch_bounds__(256) void tpa_fused_pass(FusedArgs a) {
  const double eps = 1.0e-6;
  double vs = 0.0, vc = 0.0;
  double gr[TPA_P], gi[TPA_P];
#pragma unroll
  for (int k = 0; k < TPA_P; ++k) gr[k] = gi[k] = 0.0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < a.N; e += stride) {
    cd A[TPA_LAM];
    tpa_fused_amp(a, a.ang, a.S, e, A);
    double f = 0.0;
#pragma unroll
    for (int l = 0; l < TPA_LAM; ++l) f += A[l].r * A[l].r + A[l].i * A[l].i;
    const double w = a.w[e];
    double y, u;
    if (a.mode == 0) {
      double cl, dcl;
      if (f > eps) {
        cl = log(f);
        dcl = 1.0 / f;
      } else {
        const double t = (f - eps) / eps;
        cl = (log(eps) + t) - (t * t) / 2.0;
        dcl = 1.0 / eps - t / eps;
      }
      y = w * cl;
      u = w * dcl;
    } else {
      y = w * f;
      u = w;
    }
    neumaier_add(vs, vc, y);
    if (a.grad) tpa_fused_adjoint(a, a.ang, a.S, e, A, u, gr, gi);
  }
  // block reduction: shuffles within each wave, then the four waves in a fixed order
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const double s2 = __shfl_xor(vs, m, 64);
    const double c2 = __shfl_xor(vc, m, 64);
    neumaier_merge(vs, vc, s2, c2);
  }
  __shared__ double sh[4][2 + 2 * TPA_P];
  if (a.grad) {
#pragma unroll
    for (int k = 0; k < TPA_P; ++k) {
      double xr = gr[k], xi = gi[k];
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) {
        xr += __shfl_xor(xr, m, 64);
        xi += __shfl_xor(xi, m, 64);
      }
      if (lane == 0) {
        sh[wave][2 + 2 * k] = xr;
        sh[wave][3 + 2 * k] = xi;
      }
    }
  }
  if (lane == 0) {
    sh[wave][0] = vs;
    sh[wave][1] = vc;
  }
  __syncthreads();
  double* out = a.partials + blockIdx.x;
  if (threadIdx.x == 0) {
    double s = sh[0][0], c = sh[0][1];
    for (int wv = 1; wv < 4; ++wv) neumaier_merge(s, c, sh[wv][0], sh[wv][1]);
    out[0] = s;
    out[gridDim.x] = c;
  }
  if (a.grad) {
    for (int j = 2 + threadIdx.x; j < 2 + 2 * TPA_P; j += blockDim.x)
      out[(long long)j * gridDim.x] = ((sh[0][j] + sh[1][j]) + sh[2][j]) + sh[3][j];
  }
}
)";
    (void)P;
    (void)ns;
  }
};

inline std::string fused_source(const TpaModel& M, const double* cg, int n_cg) {
  Gen g(M, cg, n_cg);
  return g.source();
}

}  // namespace tpa_fused
