// HIP kernels for four-body final states of spin-0 particles (trees of three two-body vertices).
//
//   k_cal_angle4   p4 (SoA) -> per topology and vertex: the core's invariant mass and the
//                  helicity angles (alpha, beta) of both outs in the core's helicity frame
//                  (cal_angle.py:186-343 chain boosts and helicity angles, angle.py:285-308).
//   k_build_pw4    angles -> partial waves h[k][event][lambda_top] with every parameter-free
//                  factor of DecayChain.get_amp (amp/core.py:1392-1445): per vertex the Wigner
//                  D of its core, the CG x barrier factor of each LS coupling, summed over the
//                  helicities of both resonances, times their lineshapes when folded.
#pragma once
#include "tpa_kernels.h"

// ----------------------------------------------------------------------------- cal_angle
__global__ __launch_bounds__(256) void k_cal_angle4(const TpaModel4* __restrict__ M,
                                                    const double* __restrict__ p4, int64_t n,
                                                    int64_t off, int64_t S,
                                                    double* __restrict__ ang) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int64_t col = off + e;
  const int nt = M->n_topo;
  const int64_t gbase = (int64_t)TPA4_NF_TOPO * nt;
  v4 p[4];
  for (int f = 0; f < 4; ++f) {
    p[f].t = p4[(int64_t)(f * 4 + 0) * n + e];
    p[f].x = p4[(int64_t)(f * 4 + 1) * n + e];
    p[f].y = p4[(int64_t)(f * 4 + 2) * n + e];
    p[f].z = p4[(int64_t)(f * 4 + 3) * n + e];
  }
  const v4 ptop = lv_add(lv_add(p[0], p[1]), lv_add(p[2], p[3]));
  ang[(gbase + 0) * S + col] = lv_M(ptop);
  for (int f = 0; f < 4; ++f) ang[(gbase + 1 + f) * S + col] = lv_M(p[f]);
  v3 base_z = v_make(0.0, 0.0, 1.0);
  if (M->random_z) {
    const v3 p3 = lv_vect(ptop);
    if (!(v_norm(p3) < 1.0e-5)) base_z = p3;
  }
  const v3 base_x = v_make(1.0, 0.0, 0.0);
  for (int t = 0; t < nt; ++t) {
    const Tpa4Topo& T = M->topo[t];
    double* out = ang + (int64_t)(TPA4_NF_TOPO * t) * S + col;
    // per vertex: every final in the core's rest frame, the frame's z and x axes, and the
    // x axis that each out carries to its own decay
    v4 fr[3][4];
    v3 zax[3], xax[3], xout[3][2];
    for (int v = 0; v < 3; ++v) {
      if (v == 0) {
        for (int f = 0; f < 4; ++f) fr[0][f] = lv_rest(ptop, p[f]);
        zax[0] = base_z;
        xax[0] = base_x;
      } else {
        const int pv = T.parent[v], po = T.parent_out[v];
        v4 pc = v4{0.0, 0.0, 0.0, 0.0};
        for (int f = 0; f < 4; ++f)
          if (T.core[v] >> f & 1) pc = lv_add(pc, fr[pv][f]);
        for (int f = 0; f < 4; ++f) fr[v][f] = lv_rest(pc, fr[pv][f]);
        zax[v] = lv_vect(pc);
        xax[v] = xout[pv][po];
      }
      v4 pcore = v4{0.0, 0.0, 0.0, 0.0};
      for (int f = 0; f < 4; ++f)
        if (T.core[v] >> f & 1) pcore = lv_add(pcore, p[f]);
      out[(5 * v + 0) * S] = lv_M(pcore);
      double bias = -M_PI;
      for (int o = 0; o < 2; ++o) {
        v4 po4 = v4{0.0, 0.0, 0.0, 0.0};
        for (int f = 0; f < 4; ++f)
          if (T.out[v][o] >> f & 1) po4 = lv_add(po4, fr[v][f]);
        double a, b;
        angle_zx_z_getx(zax[v], xax[v], lv_vect(po4), &a, &b, &xout[v][o]);
        a = fmod_floor(a - bias, 2.0 * M_PI) + bias;
        bias -= M_PI;
        out[(5 * v + 1 + 2 * o) * S] = a;
        out[(5 * v + 2 + 2 * o) * S] = b;
      }
    }
  }
}

// ------------------------------------------------------------------ partial-wave build
__global__ __launch_bounds__(128) void k_build_pw4(const TpaModel4* __restrict__ M,
                                                   const double* __restrict__ ang, int64_t n,
                                                   int64_t off, int64_t S,
                                                   const double* __restrict__ cg,
                                                   const double* __restrict__ dw, int fold,
                                                   double2* __restrict__ h, int64_t hS,
                                                   int64_t hoff) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int64_t col = off + e;
  const int nt = M->n_topo;
  const int lam_pad = M->lam_pad;
  const int64_t gbase = (int64_t)TPA4_NF_TOPO * nt;
  const double mtop = ang[(gbase + 0) * S + col];
  double mf[4];
  for (int f = 0; f < 4; ++f) mf[f] = ang[(gbase + 1 + f) * S + col];
  const int nA = M->n_top;
  double sp[TPA_MAX_J2 + 1], cp[TPA_MAX_J2 + 1];
  for (int c = 0; c < M->n_chain; ++c) {
    const Tpa4Chain& C = M->chain[c];
    const double* A = ang + (int64_t)(TPA4_NF_TOPO * C.topo) * S + col;
    // masses of the node slots: top, the two resonances, the finals
    double mslot[7];
    mslot[0] = mtop;
    mslot[1] = A[5 * S];
    mslot[2] = A[10 * S];
    for (int f = 0; f < 4; ++f) mslot[3 + f] = mf[f];
    // per vertex: barrier factors per LS and D^{J_core*}_{lambda_core, delta}(alpha, beta, 0)
    const double dd = C.d * C.d;
    double bf[3][TPA_MAX_LS];
    cd D[3][TPA_MAX_HEL][TPA_MAX_J2 + 1];
    int jcore[3], ncore[3];
    for (int v = 0; v < 3; ++v) {
      const double q2 = rel_p2(mslot[v], mslot[C.out_slot[v][0]], mslot[C.out_slot[v][1]]);
      for (int s = 0; s < C.n_ls[v]; ++s) {
        const int l = C.l[v][s];
        const double ql = pow(q2, 0.5 * l);
        bf[v][s] = C.act[v][s] ? ql / sqrt(bprime_poly(l, q2 * dd)) : ql;
      }
      const double a = A[(5 * v + (C.swap[v] ? 3 : 1)) * S];
      const double b = A[(5 * v + (C.swap[v] ? 4 : 2)) * S];
      const int j = v == 0 ? M->top_j2 : C.r_j2[v - 1];
      const int nh = v == 0 ? nA : C.r_n[v - 1];
      jcore[v] = j;
      ncore[v] = nh;
      half_angle_powers(b, j, sp, cp);
      const double* wj = dw + dweight_offset(j);
      for (int i = 0; i < nh; ++i) {
        const int m = v == 0 ? M->hel_top[i] : C.r_hel[v - 1][i];
        const cd ph = c_expi((0.5 * m) * a);
        for (int dl = -j; dl <= j; dl += 2) D[v][i][(dl + j) / 2] = c_scale(ph, wigner_d(wj, j, m, dl, sp, cp));
      }
    }
    cd bw = c_make(1.0, 0.0);
    if (fold) {
      for (int r = 0; r < 2; ++r) {
        if (!C.r_shape[r]) continue;
        const int v = r + 1;  // the resonance of slot r + 1 decays at vertex r + 1
        const double mR = mslot[r + 1];
        const double q = rel_p(mR, mslot[C.out_slot[v][0]], mslot[C.out_slot[v][1]]);
        const double q0 = rel_p(C.r_m0[r], C.r_q0m[r][0], C.r_q0m[r][1]);
        bw = c_mul(bw, lineshape(C.r_shape[r], mR, C.r_m0[r], C.r_g0[r], q, q0, C.r_bw_l[r], C.d));
      }
    }
    // helicity index -> 2*lambda of a slot (finals: spin 0)
    auto hel = [&](int slot, int iA, int i1, int i2) {
      return slot == 0 ? M->hel_top[iA] : slot == 1 ? C.r_hel[0][i1] : slot == 2 ? C.r_hel[1][i2] : 0;
    };
    auto hidx = [&](int slot, int iA, int i1, int i2) {
      return slot == 0 ? iA : slot == 1 ? i1 : slot == 2 ? i2 : 0;
    };
    auto nhel = [&](int slot) { return slot == 0 ? nA : slot == 1 ? C.r_n[0] : slot == 2 ? C.r_n[1] : 1; };
    for (int s0 = 0; s0 < C.n_ls[0]; ++s0)
      for (int s1 = 0; s1 < C.n_ls[1]; ++s1)
        for (int s2 = 0; s2 < C.n_ls[2]; ++s2) {
          const int64_t k = C.k0 + (s0 * C.n_ls[1] + s1) * C.n_ls[2] + s2;
          const int sv[3] = {s0, s1, s2};
          double2* hk = h + ((int64_t)k * hS + hoff + e) * lam_pad;
          for (int iA = 0; iA < nA; ++iA) {
            cd sum = c_make(0.0, 0.0);
            for (int i1 = 0; i1 < C.r_n[0]; ++i1)
              for (int i2 = 0; i2 < C.r_n[1]; ++i2) {
                cd term = c_make(1.0, 0.0);
                bool ok = true;
                for (int v = 0; v < 3 && ok; ++v) {
                  const int o0 = C.out_slot[v][0], o1 = C.out_slot[v][1];
                  const int dl = hel(o0, iA, i1, i2) - hel(o1, iA, i1, i2);
                  if (dl > jcore[v] || dl < -jcore[v]) {
                    ok = false;
                    break;
                  }
                  const int ic = hidx(v, iA, i1, i2);
                  const int ib0 = hidx(o0, iA, i1, i2), ib1 = hidx(o1, iA, i1, i2);
                  const double cgv = cg[C.cg_off[v] + (sv[v] * nhel(o0) + ib0) * nhel(o1) + ib1];
                  term = c_mul(term, c_scale(D[v][ic][(dl + jcore[v]) / 2], cgv * bf[v][sv[v]]));
                }
                if (ok) sum = c_add(sum, term);
              }
            sum = c_mul(sum, bw);
            hk[iA] = make_double2(sum.r, sum.i);
          }
          for (int l = nA; l < lam_pad; ++l) hk[l] = make_double2(0.0, 0.0);
        }
  }
}
