// HIP kernels of the unbinned-likelihood engine (gfx950, FP64 path).
//
//   k_cal_angle3   p4 (SoA) -> per-topology helicity + aligned angles and masses
//                  (tf_pwa/cal_angle.py:262-490, angle.py:98-392)
//   k_build_pw     angles -> partial-wave amplitudes h[k][event][lambda] that carry every
//                  parameter-independent factor: Wigner D of both vertices, CG x barrier
//                  factor per LS coupling, the contraction over the resonance helicity and
//                  the aligned D of spin>0 finals (amp/core.py:931-1164, 1392-1445,
//                  dfun.py:118-266); optionally the Breit-Wigner is folded in.
//   k_eval_pipe    per evaluation: A_lambda = sum_k c_k h_k,lambda, f = sum |A|^2, the
//                  weighted log-sum (clip_log) or linear MC sum with a compensated
//                  accumulator, and the adjoint sums G_k = sum_e u_e sum_lambda conj(A) h_k
//                  (model/model.py:41-94, 283-292, 407-449); software-pipelined loads.
//                  k_eval_fold is the same pass without the pipeline (kept for comparison).
//   k_reduce       deterministic fixed-order reduction of the per-block partial sums.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "tpa_log.h"
#include "tpa_math.h"

// ----------------------------------------------------------------------------- cal_angle
__global__ __launch_bounds__(256) void k_cal_angle3(const TpaModel* __restrict__ M,
                                                    const double* __restrict__ p4, int64_t n,
                                                    int64_t off, int64_t S,
                                                    double* __restrict__ ang) {
  // events e < n of the input p4 [f*4+c][n] go to column off + e of the cache (stride S)
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int64_t col = off + e;
  const int nt = M->n_topo;
  const int64_t gbase = (int64_t)TPA_NF_TOPO * nt;
  v4 p[TPA_MAX_FINAL];
  for (int f = 0; f < 3; ++f) {
    p[f].t = p4[(int64_t)(f * 4 + 0) * n + e];
    p[f].x = p4[(int64_t)(f * 4 + 1) * n + e];
    p[f].y = p4[(int64_t)(f * 4 + 2) * n + e];
    p[f].z = p4[(int64_t)(f * 4 + 3) * n + e];
  }
  v4 ptop = lv_add(lv_add(p[0], p[1]), p[2]);
  ang[(gbase + 0) * S + col] = lv_M(ptop);
  for (int f = 0; f < 3; ++f) ang[(gbase + 1 + f) * S + col] = lv_M(p[f]);

  v3 base_z = v_make(0.0, 0.0, 1.0);
  if (M->random_z) {
    v3 p3 = lv_vect(ptop);
    if (!(v_norm(p3) < 1.0e-5)) base_z = p3;
  }
  const v3 base_x = v_make(1.0, 0.0, 0.0);

  su2 rm[TPA_MAX_TOPO][TPA_MAX_FINAL];
  su2 bm[TPA_MAX_TOPO][TPA_MAX_FINAL];
  for (int t = 0; t < nt; ++t) {
    const TpaTopo T = M->topo[t];
    double* out = ang + (int64_t)(TPA_NF_TOPO * t) * S + col;
    v4 pR = lv_add(p[T.y < T.z ? T.y : T.z], p[T.y < T.z ? T.z : T.y]);
    out[TPA_F_MR * S] = lv_M(pR);
    // cal_chain_boost: everything into the top rest frame, daughters of R into R's frame
    v4 pRt = lv_rest(ptop, pR);
    v4 pXt = lv_rest(ptop, p[T.x]);
    v4 pYt = lv_rest(ptop, p[T.y]);
    v4 pZt = lv_rest(ptop, p[T.z]);
    // top decay in the representative order (alpha folded to [-pi,pi) then [-2pi,0))
    double bias = -M_PI;
    v3 xR = v_make(1.0, 0.0, 0.0);
    su2 rR = su2_rot_y(0.0), bR = su2_rot_y(0.0);
    for (int pos = 0; pos < 2; ++pos) {
      const bool isR = (pos == 0) == (T.r_first != 0);
      const v4 pj = isR ? pRt : pXt;
      double a, b;
      v3 x;
      angle_zx_z_getx(base_z, base_x, lv_vect(pj), &a, &b, &x);
      a = fmod_floor(a - bias, 2.0 * M_PI) + bias;
      bias -= M_PI;
      su2 bmat = su2_boost_z(lv_omega(pj));
      su2 rmat = su2_mul(su2_rot_y(b), su2_rot_z(a));
      if (isR) {
        out[(TPA_F_AR + 0) * S] = a;
        out[(TPA_F_AR + 1) * S] = b;
        xR = x;
        rR = rmat;
        bR = bmat;
      } else {
        out[(TPA_F_AX + 0) * S] = a;
        out[(TPA_F_AX + 1) * S] = b;
        rm[t][T.x] = rmat;
        bm[t][T.x] = bmat;
      }
    }
    // resonance decay, daughters y then z
    bias = -M_PI;
    const v3 zR = lv_vect(pRt);
    for (int pos = 0; pos < 2; ++pos) {
      const int f = pos == 0 ? T.y : T.z;
      const v4 pjr = lv_rest(pRt, pos == 0 ? pYt : pZt);
      double a, b;
      v3 x;
      angle_zx_z_getx(zR, xR, lv_vect(pjr), &a, &b, &x);
      a = fmod_floor(a - bias, 2.0 * M_PI) + bias;
      bias -= M_PI;
      const int fld = pos == 0 ? TPA_F_AY : TPA_F_AZ;
      out[(fld + 0) * S] = a;
      out[(fld + 1) * S] = b;
      bm[t][f] = su2_boost_z(lv_omega(pjr));
      rm[t][f] = su2_mul(su2_mul(su2_mul(su2_rot_y(b), su2_rot_z(a)), bR), rR);
    }
  }
  // aligned angles (cal_angle.py:444-474, rule 1 reference frames, r_boost + final_rest)
  for (int t = 0; t < nt; ++t) {
    double* out = ang + (int64_t)(TPA_NF_TOPO * t) * S + col;
    for (int f = 0; f < 3; ++f) {
      double al = 0.0, be = 0.0, ga = 0.0;
      if (M->topo[t].aligned[f]) {
        const int r0 = M->ref_topo[f];
        su2 R = su2_mul(rm[r0][f], su2_inv(rm[t][f]));
        R = su2_mul(su2_mul(bm[r0][f], R), su2_inv(bm[t][f]));
        su2_euler(R, &al, &be, &ga);
      }
      out[(TPA_F_AL + 3 * f + 0) * S] = al;
      out[(TPA_F_AL + 3 * f + 1) * S] = be;
      out[(TPA_F_AL + 3 * f + 2) * S] = ga;
    }
  }
}

// ------------------------------------------------------------------ partial-wave build
__global__ __launch_bounds__(128) void k_build_pw(const TpaModel* __restrict__ M,
                                                  const double* __restrict__ ang, int64_t n,
                                                  int64_t off, int64_t S,
                                                  const double* __restrict__ cg,
                                                  const double* __restrict__ dw, int fold,
                                                  double2* __restrict__ h, int64_t hS,
                                                  int64_t hoff) {
  // columns off .. off+n-1 of the angle cache (stride S) into columns hoff .. hoff+n-1 of h
  // (row stride hS: the cache itself with hS = S, hoff = off, or a chunk buffer)
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int64_t col = off + e;
  const int nt = M->n_topo;
  const int lam_pad = M->lam_pad;
  const int64_t gbase = (int64_t)TPA_NF_TOPO * nt;
  const double mtop = ang[(gbase + 0) * S + col];
  double mf[3];
  for (int f = 0; f < 3; ++f) mf[f] = ang[(gbase + 1 + f) * S + col];
  const int ja = M->top_j2;
  const int nA = M->n_top;
  int nfh[3];
  for (int f = 0; f < 3; ++f) nfh[f] = M->n_fh[f];

  for (int c = 0; c < M->n_chain; ++c) {
    const TpaChain& C = M->chain[c];
    const TpaTopo T = M->topo[C.topo];
    const double* A = ang + (int64_t)(TPA_NF_TOPO * C.topo) * S + col;
    const double mR = A[TPA_F_MR * S];
    const int fx = T.x;
    const int fy = C.r_y_first ? T.y : T.z;  // this chain's first daughter of R
    const int fz = C.r_y_first ? T.z : T.y;
    // vertex 1: top -> (R, X) in this chain's order; D uses the angle of outs[0]
    double a1, b1;
    if (C.top_r_first) {
      a1 = A[(TPA_F_AR + 0) * S];
      b1 = A[(TPA_F_AR + 1) * S];
    } else {
      a1 = A[(TPA_F_AX + 0) * S];
      b1 = A[(TPA_F_AX + 1) * S];
    }
    const double q1sq = C.top_r_first ? rel_p2(mtop, mR, mf[fx]) : rel_p2(mtop, mf[fx], mR);
    // vertex 2: R -> (fy, fz)
    const int fld2 = (fy == T.y) ? TPA_F_AY : TPA_F_AZ;
    const double a2 = A[(fld2 + 0) * S];
    const double b2 = A[(fld2 + 1) * S];
    const double q2sq = rel_p2(mR, mf[fy], mf[fz]);
    const double dd = C.d * C.d;
    double bf1[TPA_MAX_LS], bf2[TPA_MAX_LS];
    for (int s = 0; s < C.n_ls1; ++s) {
      const int l = C.l1[s];
      const double ql = pow(q1sq, 0.5 * l);
      bf1[s] = C.act1[s] ? ql / sqrt(bprime_poly(l, q1sq * dd)) : ql;
    }
    for (int s = 0; s < C.n_ls2; ++s) {
      const int l = C.l2[s];
      const double ql = pow(q2sq, 0.5 * l);
      bf2[s] = C.act2[s] ? ql / sqrt(bprime_poly(l, q2sq * dd)) : ql;
    }
    // D^{J_top*}_{lambda_A, delta}(a1, b1, 0) for all delta in [-J, J]
    double sp[TPA_MAX_J2 + 1], cp[TPA_MAX_J2 + 1];
    cd D1[TPA_MAX_HEL][TPA_MAX_J2 + 1];
    half_angle_powers(b1, ja, sp, cp);
    const double* wA = dw + dweight_offset(ja);
    for (int iA = 0; iA < nA; ++iA) {
      const int m = M->hel_top[iA];
      const cd ph = c_expi((0.5 * m) * a1);
      for (int dl = -ja; dl <= ja; dl += 2)
        D1[iA][(dl + ja) / 2] = c_scale(ph, wigner_d(wA, ja, m, dl, sp, cp));
    }
    const int jr = C.jr2;
    cd D2[TPA_MAX_HEL][TPA_MAX_J2 + 1];
    half_angle_powers(b2, jr, sp, cp);
    const double* wR = dw + dweight_offset(jr);
    for (int iR = 0; iR < C.n_r; ++iR) {
      const int m = C.hel_r[iR];
      const cd ph = c_expi((0.5 * m) * a2);
      for (int dl = -jr; dl <= jr; dl += 2)
        D2[iR][(dl + jr) / 2] = c_scale(ph, wigner_d(wR, jr, m, dl, sp, cp));
    }
    // aligned D^{J_f*}_{lambda, lambda'}(alpha, beta, gamma) of spin>0 finals
    cd Da[3][3][3];
    int has_al[3];
    for (int f = 0; f < 3; ++f) {
      const int jf = M->final_j2[f];
      has_al[f] = (T.aligned[f] && jf != 0) ? 1 : 0;
      if (!has_al[f]) continue;
      const double al = A[(TPA_F_AL + 3 * f + 0) * S];
      const double be = A[(TPA_F_AL + 3 * f + 1) * S];
      const double ga = A[(TPA_F_AL + 3 * f + 2) * S];
      half_angle_powers(be, jf, sp, cp);
      const double* wf = dw + dweight_offset(jf);
      for (int i = 0; i < nfh[f]; ++i)
        for (int i2 = 0; i2 < nfh[f]; ++i2) {
          const int m1 = M->hel_f[f][i], m2 = M->hel_f[f][i2];
          const cd eae = c_mul(c_expi((0.5 * m1) * al), c_expi((0.5 * m2) * ga));
          Da[f][i][i2] = c_scale(eae, wigner_d(wf, jf, m1, m2, sp, cp));
        }
    }
    cd bw = c_make(1.0, 0.0);
    if (fold && C.shape) {
      const double q = rel_p(mR, mf[fy], mf[fz]);
      const double q0 = rel_p(C.m0, C.m_y, C.m_z);
      bw = lineshape(C.shape, mR, C.m0, C.g0, q, q0, C.bw_l, C.d);
    }
    // helicity bookkeeping of vertex 1 (b1, c1) and vertex 2 (b2, c2)
    const int nX = nfh[fx];
    const int nb1 = C.top_r_first ? C.n_r : nX;
    const int nc1 = C.top_r_first ? nX : C.n_r;
    const int nb2 = nfh[fy], nc2 = nfh[fz];
    for (int s1 = 0; s1 < C.n_ls1; ++s1) {
      for (int s2 = 0; s2 < C.n_ls2; ++s2) {
        const int64_t k = C.k0 + s1 * C.n_ls2 + s2;
        double2* hk = h + ((int64_t)k * hS + hoff + e) * lam_pad;
        int lf = 0;
        for (int iA = 0; iA < nA; ++iA)
          for (int o0 = 0; o0 < nfh[0]; ++o0)
            for (int o1 = 0; o1 < nfh[1]; ++o1)
              for (int o2 = 0; o2 < nfh[2]; ++o2, ++lf) {
                const int oo[3] = {o0, o1, o2};
                cd sum = c_make(0.0, 0.0);
                int lo[3], hi[3];
                for (int f = 0; f < 3; ++f) {
                  lo[f] = has_al[f] ? 0 : oo[f];
                  hi[f] = has_al[f] ? nfh[f] : oo[f] + 1;
                }
                int j[3];
                for (j[0] = lo[0]; j[0] < hi[0]; ++j[0])
                  for (j[1] = lo[1]; j[1] < hi[1]; ++j[1])
                    for (j[2] = lo[2]; j[2] < hi[2]; ++j[2]) {
                      cd dal = c_make(1.0, 0.0);
                      for (int f = 0; f < 3; ++f)
                        if (has_al[f]) dal = c_mul(dal, Da[f][j[f]][oo[f]]);
                      const int jX = j[fx];
                      const int lX = M->hel_f[fx][jX];
                      const int ib2 = j[fy], ic2 = j[fz];
                      const int d2 = M->hel_f[fy][ib2] - M->hel_f[fz][ic2];
                      if (d2 > jr || d2 < -jr) continue;
                      const double cg2v = cg[C.cg2_off + (s2 * nb2 + ib2) * nc2 + ic2];
                      cd acc = c_make(0.0, 0.0);
                      for (int iR = 0; iR < C.n_r; ++iR) {
                        const int lR = C.hel_r[iR];
                        const int d1 = C.top_r_first ? lR - lX : lX - lR;
                        if (d1 > ja || d1 < -ja) continue;
                        const int ib1 = C.top_r_first ? iR : jX;
                        const int ic1 = C.top_r_first ? jX : iR;
                        const double cg1v = cg[C.cg1_off + (s1 * nb1 + ib1) * nc1 + ic1];
                        const cd v1 = c_scale(D1[iA][(d1 + ja) / 2], cg1v * bf1[s1]);
                        const cd v2 = c_scale(D2[iR][(d2 + jr) / 2], cg2v * bf2[s2]);
                        acc = c_add(acc, c_mul(v1, v2));
                      }
                      sum = c_add(sum, c_mul(acc, dal));
                    }
                sum = c_mul(sum, bw);
                hk[lf] = make_double2(sum.r, sum.i);
              }
        for (; lf < lam_pad; ++lf) hk[lf] = make_double2(0.0, 0.0);
      }
    }
  }
}

// ------------------------------------------------------------------------- evaluation
#define TPA_EVAL_CIN 32  // coefficients passed in the kernel arguments (P <= 32)

struct EvalArgs {
  const double2* __restrict__ h;     // [P][N][LG]
  const float2* __restrict__ h32;    // [P][N][LG], FP32 copy of h (passes that read FP32)
  const double* __restrict__ w;      // [N]
  const double2* __restrict__ coef;  // [P] in device memory, or null: cin holds them
  double* __restrict__ partials;     // [grid][nacc]
  double* __restrict__ f_out;        // [N] or null
  int64_t N;
  int P;
  int mode;  // 0: data, sum w * clip_log(f); 1: MC, sum w * f
  int grad;
  int nacc;  // 2 + 2P
  double2 cin[TPA_EVAL_CIN];  // coefficients by value (k_eval_pipe / k_eval_pipe32, coef null)
  // the helicity-halved passes (k_eval_half): chain phases [chain][N] of the FP32 copy (rho)
  // or the FP64 copy (rho64) and the chain of each wave (h32, resp. h, then holds the stored
  // half, [P][N][NH])
  const float2* __restrict__ rho;
  const double2* __restrict__ rho64;
  signed char kch[TPA_EVAL_CIN];
  // k_eval_half_chain: the wave of lane kg's slot j at plan[kg * PML + j] (-1: pad) and the
  // chain of its two runs at pch[kg * 2 + run] (-1: none)
  signed char plan[TPA_EVAL_CIN];
  signed char pch[8];
};

// Coefficient k of an evaluation pass (from the kernel arguments when a.coef is null).
TPA_DEV double2 eval_coef(const EvalArgs& a, int k) { return a.coef ? a.coef[k] : a.cin[k]; }


#define TPA_NACC_MAX (2 + 2 * TPA_MAX_P)

// Lane layout: lane = ev * (LG * KG) + kg * LG + lam. The LG lanes of a k-group hold the
// helicity components of one event, and the KG groups split the partial waves
// k = kg + KG * j (j < PML). A_lambda is completed by a reduction over the k-groups and f by a
// reduction over lambda, so each lane keeps only PML amplitudes and PML gradient sums.
// NT: stream the cache with non-temporal loads (it is read once per evaluation).
template <bool NT>
TPA_DEV double2 load_pw(const double2* p) {
  if constexpr (NT) {
    typedef double d2v __attribute__((ext_vector_type(2)));
    const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v*>(p));
    return make_double2(v.x, v.y);
  } else {
    return *p;
  }
}

TPA_DEV float2 load_pw32(const float2* p) {
  typedef float f2v __attribute__((ext_vector_type(2)));
  const f2v v = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p));
  return make_float2(v.x, v.y);
}

// The cache an evaluation pass reads: FP64 (h) or its FP32 copy (h32).
template <typename HT>
TPA_DEV const HT* pw_cache(const EvalArgs& a) {
  if constexpr (sizeof(HT) == sizeof(float2))
    return a.h32;
  else
    return a.h;
}
template <typename HT>
TPA_DEV HT load_stream(const HT* p) {
  if constexpr (sizeof(HT) == sizeof(float2))
    return load_pw32(p);
  else
    return load_pw<true>(p);
}

template <int LG, int KG, int PML, int U, bool NT = false>
__global__ __launch_bounds__(256) void k_eval_fold(EvalArgs a) {
  constexpr int LPE = LG * KG;  // lanes per event
  constexpr int EPW = 64 / LPE;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int lam = lane & (LG - 1);
  const int kg = (lane / LG) & (KG - 1);
  const int el = lane / LPE;
  const double eps = 1.0e-6;
  const double logeps = log(1.0e-6);
  double vs = 0.0, vc = 0.0;
  double gr[PML], gi[PML], cr[PML], ci[PML];
#pragma unroll
  for (int j = 0; j < PML; ++j) {
    const int k = kg + KG * j;
    gr[j] = gi[j] = 0.0;
    const double2 c = k < a.P ? a.coef[k] : make_double2(0.0, 0.0);
    cr[j] = c.x;
    ci[j] = c.y;
  }
  // each wave takes U * EPW consecutive events per step and issues all their loads first
  const int64_t stride = (int64_t)gridDim.x * 4 * EPW * U;
  for (int64_t base = ((int64_t)blockIdx.x * 4 + wave) * EPW * U; base < a.N; base += stride) {
    double2 hv[U][PML];
    int64_t ev[U];
    bool ok[U];
#pragma unroll
    for (int uu = 0; uu < U; ++uu) {
      ev[uu] = base + uu * EPW + el;
      ok[uu] = ev[uu] < a.N;
#pragma unroll
      for (int j = 0; j < PML; ++j) {
        const int k = kg + KG * j;
        hv[uu][j] = (ok[uu] && k < a.P) ? load_pw<NT>(&a.h[((int64_t)k * a.N + ev[uu]) * LG + lam])
                                        : make_double2(0.0, 0.0);
      }
    }
#pragma unroll
    for (int uu = 0; uu < U; ++uu) {
      const int64_t e = ev[uu];
      const bool valid = ok[uu];
      double ar = 0.0, ai = 0.0;
#pragma unroll
      for (int j = 0; j < PML; ++j) {
        ar += cr[j] * hv[uu][j].x - ci[j] * hv[uu][j].y;
        ai += cr[j] * hv[uu][j].y + ci[j] * hv[uu][j].x;
      }
#pragma unroll
      for (int m = LG; m < LPE; m <<= 1) {
        ar += __shfl_xor(ar, m, 64);
        ai += __shfl_xor(ai, m, 64);
      }
      double f = ar * ar + ai * ai;
#pragma unroll
      for (int m = 1; m < LG; m <<= 1) f += __shfl_xor(f, m, 64);
      double y = 0.0, u = 0.0;
      if (valid) {
        const double w = a.w[e];
        if (a.mode == 0) {
          const double dom = (w == 0.0) ? 1.0 : w;
          const double x = (f * w) / dom;
          double cl, dcl;
          if (x > eps) {
            cl = log(x);
            dcl = 1.0 / x;
          } else {
            const double dx = x - eps;
            const double t = dx / eps;
            cl = (logeps + t) - (t * t) / 2.0;
            dcl = 1.0 / eps - t / eps;
          }
          y = w * cl;
          u = (w * dcl) * (w / dom);
        } else {
          y = w * f;
          u = w;
        }
        if (a.f_out != nullptr && lam == 0 && kg == 0) a.f_out[e] = f;
      }
      if (lam == 0 && kg == 0) neumaier_add(vs, vc, y);
      if (a.grad) {
        const double tr = u * ar, ti = -(u * ai);
#pragma unroll
        for (int j = 0; j < PML; ++j) {
          gr[j] += tr * hv[uu][j].x - ti * hv[uu][j].y;
          gi[j] += tr * hv[uu][j].y + ti * hv[uu][j].x;
        }
      }
    }
  }
  // value: full wave reduction of the compensated pair
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const double s2 = __shfl_xor(vs, m, 64);
    const double c2 = __shfl_xor(vc, m, 64);
    neumaier_merge(vs, vc, s2, c2);
  }
  __shared__ double sh[4][TPA_NACC_MAX];
  if (a.grad) {
    // gradient slots: reduce over lambda and events, keep the k-group bits
#pragma unroll
    for (int j = 0; j < PML; ++j) {
      double xr = gr[j], xi = gi[j];
#pragma unroll
      for (int m = 1; m < LG; m <<= 1) {
        xr += __shfl_xor(xr, m, 64);
        xi += __shfl_xor(xi, m, 64);
      }
#pragma unroll
      for (int m = LPE; m < 64; m <<= 1) {
        xr += __shfl_xor(xr, m, 64);
        xi += __shfl_xor(xi, m, 64);
      }
      const int k = kg + KG * j;
      if (el == 0 && lam == 0 && k < a.P) {
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
  // slot-major partials: slot j of this block at out[j * gridDim.x] (k_reduce reads each
  // slot's blocks contiguously)
  double* out = a.partials + blockIdx.x;
  if (threadIdx.x == 0) {
    double s = sh[0][0], c = sh[0][1];
    for (int wv = 1; wv < 4; ++wv) neumaier_merge(s, c, sh[wv][0], sh[wv][1]);
    out[0] = s;
    out[gridDim.x] = c;
  }
  if (a.grad) {
    for (int j = 2 + threadIdx.x; j < a.nacc; j += blockDim.x)
      out[(int64_t)j * gridDim.x] = ((sh[0][j] + sh[1][j]) + sh[2][j]) + sh[3][j];
  }
}

// Cross-lane exchange of a double with lane ^ M: DPP quad permutations for M = 1, 2 (no LDS
// traffic), ds_swizzle's bit mode (xor within 32 lanes, no address operand) for M = 4 .. 16,
// ds_bpermute beyond.
template <int M>
TPA_DEV double xor_lanes(double v) {
  static_assert(M >= 1 && M < 64 && (M & (M - 1)) == 0, "M must be a power of two below 64");
  int lo = __double2loint(v), hi = __double2hiint(v);
  if constexpr (M == 1 || M == 2) {
    constexpr int ctrl = M == 1 ? 0xB1 : 0x4E;  // quad_perm [1,0,3,2] / [2,3,0,1]
    lo = __builtin_amdgcn_update_dpp(0, lo, ctrl, 0xF, 0xF, true);
    hi = __builtin_amdgcn_update_dpp(0, hi, ctrl, 0xF, 0xF, true);
  } else if constexpr (M < 32) {
    constexpr int pattern = 0x1F | (M << 10);  // and_mask 31, or_mask 0, xor_mask M
    lo = __builtin_amdgcn_ds_swizzle(lo, pattern);
    hi = __builtin_amdgcn_ds_swizzle(hi, pattern);
  } else {
    return __shfl_xor(v, M, 64);
  }
  return __hiloint2double(hi, lo);
}

// Sum of v over the lanes lane ^ m for m = FROM, 2 FROM, .. < TO (a butterfly).
template <int FROM, int TO>
TPA_DEV double butterfly(double v) {
  if constexpr (FROM < TO) {
    v += xor_lanes<FROM>(v);
    return butterfly<2 * FROM, TO>(v);
  } else {
    return v;
  }
}

// The hardware's reciprocal approximation (v_rcp_f64), the seed of tpa_recip_pos.
struct HwRcp {
  TPA_DEV double operator()(double d) const { return __builtin_amdgcn_rcp(d); }
};
constexpr HwRcp hw_rcp{};

// One event group of the evaluation pass: A_lambda from the lane's PML partial waves, f, the
// mode's value and derivative, the value into the compensated pair (vs, vc) of the event's
// leader lane and the adjoint into (gr, gi). Lanes past the last event (valid = false) carry
// finite stand-in waves and add exactly zero. Every lane of the wave must be active. HT is
// the cache's element type (double2, or float2 for an FP32 copy: its values are widened to
// FP64 exactly and every operation after the load is FP64).
template <int LG, int KG, int PML, typename HT>
TPA_DEV void eval_group(int mode, bool grad, double* __restrict__ f_out,
                        const HT (&hs)[PML], double w, int64_t e, bool valid, bool leader,
                        const double (&cr)[PML], const double (&ci)[PML], double& vs, double& vc,
                        double (&gr)[PML], double (&gi)[PML]) {
  constexpr int LPE = LG * KG;
  const double eps = 1.0e-6;
  double2 hv[PML];
#pragma unroll
  for (int j = 0; j < PML; ++j) hv[j] = make_double2(hs[j].x, hs[j].y);
  // complex multiply-accumulates as chained FMAs into the accumulators (two per real part
  // instead of a multiply, an FMA and an add)
  double ar = 0.0, ai = 0.0;
#pragma unroll
  for (int j = 0; j < PML; ++j) {
    ar = fma(cr[j], hv[j].x, ar);
    ar = fma(-ci[j], hv[j].y, ar);
    ai = fma(cr[j], hv[j].y, ai);
    ai = fma(ci[j], hv[j].x, ai);
  }
  ar = butterfly<LG, LPE>(ar);
  ai = butterfly<LG, LPE>(ai);
  const double f = butterfly<1, LG>(ar * ar + ai * ai);
  // lanes past the last event take weight zero (their f is finite: clamped waves); the weight
  // is read outside any branch, so no path leaves its load pending
  const double wv = valid ? w : 0.0;
  double y, u;
  if (mode == 0) {
    // clip_log(f) (tf_pwa/model/model.py:283-292) weighted by w
    double cl, dcl;
    if (f > eps) {
      cl = tpa_log_pos(f, hw_rcp);
      dcl = tpa_recip_pos(f, hw_rcp);
    } else {
      const double t = (f - eps) / eps;
      cl = (log(eps) + t) - (t * t) / 2.0;
      dcl = 1.0 / eps - t / eps;
    }
    y = wv * cl;
    u = wv * dcl;
  } else {
    y = wv * f;
    u = wv;
  }
  if (f_out != nullptr && valid && leader) f_out[e] = f;
  neumaier_add(vs, vc, leader ? y : 0.0);
  if (grad) {
    const double tr = u * ar, ti = -(u * ai);
#pragma unroll
    for (int j = 0; j < PML; ++j) {
      gr[j] = fma(tr, hv[j].x, gr[j]);
      gr[j] = fma(-ti, hv[j].y, gr[j]);
      gi[j] = fma(tr, hv[j].y, gi[j]);
      gi[j] = fma(ti, hv[j].x, gi[j]);
    }
  }
}

// The evaluation pass, software-pipelined: each wave holds two register stages of one event
// group each (its partial waves and weights) and issues the loads of the next group before
// it processes the current one, so its loads stay in flight while it computes. Loads are
// unconditional (no branch around them, so the wait before a stage's first use counts only
// the older loads): past the last event or partial wave they read a clamped, valid address,
// whose values meet a zero coefficient or a zero weight. The loop bounds are wave-uniform
// (scalar), so every lane stays active for the cross-lane exchanges. Requires N >= 1.
// HT: the cache read (double2: a.h; float2: the FP32 copy a.h32). WPE: the waves per SIMD
// the register budget is fitted to (without spills: FP32 stages at PML 6 need 162 VGPRs,
// at PML 4 128, at PML 7-8 179-196).
template <int LG, int KG, int PML, typename HT = double2,
          int WPE = (PML <= 4 && sizeof(HT) == 8) ? 4 : PML <= 6 ? 3 : (sizeof(HT) == 8 ? 2 : 1)>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void k_eval_pipe(EvalArgs a) {
  constexpr int LPE = LG * KG;
  constexpr int EPW = 64 / LPE;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int lam = lane & (LG - 1);
  const int kg = (lane / LG) & (KG - 1);
  const int el = lane / LPE;
  const bool leader = lam == 0 && kg == 0;
  double vs = 0.0, vc = 0.0;
  double gr[PML], gi[PML], cr[PML], ci[PML];
  const HT* base[PML];
  const HT* cache = pw_cache<HT>(a);
#pragma unroll
  for (int j = 0; j < PML; ++j) {
    const int k = kg + KG * j;
    gr[j] = gi[j] = 0.0;
    const double2 c = k < a.P ? eval_coef(a, k) : make_double2(0.0, 0.0);
    cr[j] = c.x;
    ci[j] = c.y;
    const int kc = k < a.P ? k : a.P - 1;
    base[j] = cache + (int64_t)kc * a.N * LG + lam;
  }
  const int64_t last = a.N - 1;
  auto load = [&](HT (&hv)[PML], double& wv, int64_t e) {
    const int64_t ec = e < a.N ? e : last;
#pragma unroll
    for (int j = 0; j < PML; ++j) hv[j] = load_stream<HT>(base[j] + ec * LG);
    wv = a.w[ec];
  };
  const int64_t stride = (int64_t)gridDim.x * 4 * EPW;
  HT h0[PML], h1[PML];
  double w0, w1;
  const int wave_u = __builtin_amdgcn_readfirstlane(wave);
  const int64_t first = ((int64_t)blockIdx.x * 4 + wave_u) * EPW;
  int64_t e = first + el;
  load(h0, w0, e);
  for (int64_t b = first; b < a.N; b += 2 * stride, e += 2 * stride) {
    load(h1, w1, e + stride);
    // keep the next stage's loads ahead of this stage's first use (no scheduling across)
    __builtin_amdgcn_sched_barrier(0);
    eval_group<LG, KG, PML>(a.mode, a.grad, a.f_out, h0, w0, e, e < a.N, leader, cr, ci, vs, vc,
                            gr, gi);
    if (b + stride >= a.N) break;
    load(h0, w0, e + 2 * stride);
    __builtin_amdgcn_sched_barrier(0);
    eval_group<LG, KG, PML>(a.mode, a.grad, a.f_out, h1, w1, e + stride, e + stride < a.N, leader,
                            cr, ci, vs, vc, gr, gi);
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const double s2 = __shfl_xor(vs, m, 64);
    const double c2 = __shfl_xor(vc, m, 64);
    neumaier_merge(vs, vc, s2, c2);
  }
  __shared__ double sh[4][TPA_NACC_MAX];
  if (a.grad) {
#pragma unroll
    for (int j = 0; j < PML; ++j) {
      double xr = butterfly<1, LG>(gr[j]), xi = butterfly<1, LG>(gi[j]);
      xr = butterfly<LPE, 64>(xr);
      xi = butterfly<LPE, 64>(xi);
      const int k = kg + KG * j;
      if (el == 0 && lam == 0 && k < a.P) {
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
  // slot-major partials: slot j of this block at out[j * gridDim.x] (k_reduce reads each
  // slot's blocks contiguously)
  double* out = a.partials + blockIdx.x;
  if (threadIdx.x == 0) {
    double s = sh[0][0], c = sh[0][1];
    for (int wv = 1; wv < 4; ++wv) neumaier_merge(s, c, sh[wv][0], sh[wv][1]);
    out[0] = s;
    out[gridDim.x] = c;
  }
  if (a.grad) {
    for (int j = 2 + threadIdx.x; j < a.nacc; j += blockDim.x)
      out[(int64_t)j * gridDim.x] = ((sh[0][j] + sh[1][j]) + sh[2][j]) + sh[3][j];
  }
}

// One event group of the FP32-cache pass with two helicity components per lane (one 16-byte
// load of two float2 per partial wave): as eval_group, with the lane's two A_lambda summed
// into f before the reduction over the LG/2 lanes of a k-group.
// FL: clip_log's log and reciprocal from tpa_log.h (about 45 fewer FP64 instructions per event
// group) or from the math library. The FP32 passes at PML 6 sit at the 168-VGPR limit of three
// waves per SIMD; tpa_log.h's form fits there only with two helicities per lane and 2 to 8
// k-groups (the P = 17 .. 24 three-body shapes), so the other PML-6 shapes keep the library's.
template <int LGL, int KG, int PML>
constexpr bool fast_log2 = PML != 6 || (LGL == 2 && KG >= 2 && KG <= 8);

template <int LGL, int KG, int PML>
TPA_DEV void eval_group2(int mode, bool grad, double* __restrict__ f_out,
                         const float4 (&hs)[PML], double w, int64_t e, bool valid, bool leader,
                         const double (&cr)[PML], const double (&ci)[PML], double& vs,
                         double& vc, double (&gr)[PML], double (&gi)[PML]) {
  constexpr int LPE = LGL * KG;
  const double eps = 1.0e-6;
  double ar0 = 0.0, ai0 = 0.0, ar1 = 0.0, ai1 = 0.0;
#pragma unroll
  for (int j = 0; j < PML; ++j) {
    const double x0 = hs[j].x, y0 = hs[j].y, x1 = hs[j].z, y1 = hs[j].w;
    ar0 = fma(cr[j], x0, ar0);
    ar0 = fma(-ci[j], y0, ar0);
    ai0 = fma(cr[j], y0, ai0);
    ai0 = fma(ci[j], x0, ai0);
    ar1 = fma(cr[j], x1, ar1);
    ar1 = fma(-ci[j], y1, ar1);
    ai1 = fma(cr[j], y1, ai1);
    ai1 = fma(ci[j], x1, ai1);
  }
  ar0 = butterfly<LGL, LPE>(ar0);
  ai0 = butterfly<LGL, LPE>(ai0);
  ar1 = butterfly<LGL, LPE>(ar1);
  ai1 = butterfly<LGL, LPE>(ai1);
  const double f = butterfly<1, LGL>((ar0 * ar0 + ai0 * ai0) + (ar1 * ar1 + ai1 * ai1));
  const double wv = valid ? w : 0.0;
  double y, u;
  if (mode == 0) {
    double cl, dcl;
    if (f > eps) {
      if constexpr (fast_log2<LGL, KG, PML>) {
        dcl = tpa_recip_pos(f, hw_rcp);
        cl = tpa_log_pos(f, hw_rcp);
      } else {
        cl = log(f);
        dcl = 1.0 / f;
      }
    } else {
      const double t = (f - eps) / eps;
      cl = (log(eps) + t) - (t * t) / 2.0;
      dcl = 1.0 / eps - t / eps;
    }
    y = wv * cl;
    u = wv * dcl;
  } else {
    y = wv * f;
    u = wv;
  }
  if (f_out != nullptr && valid && leader) f_out[e] = f;
  neumaier_add(vs, vc, leader ? y : 0.0);
  if (grad) {
    const double tr0 = u * ar0, ti0 = -(u * ai0), tr1 = u * ar1, ti1 = -(u * ai1);
#pragma unroll
    for (int j = 0; j < PML; ++j) {
      // widen again rather than keep the first loop's FP64 values live (an empty asm makes
      // the FP32 registers opaque to common-subexpression elimination; no instruction)
      float4 q = hs[j];
      asm("" : "+v"(q.x), "+v"(q.y), "+v"(q.z), "+v"(q.w));
      const double x0 = q.x, y0 = q.y, x1 = q.z, y1 = q.w;
      gr[j] = fma(tr0, x0, gr[j]);
      gr[j] = fma(-ti0, y0, gr[j]);
      gr[j] = fma(tr1, x1, gr[j]);
      gr[j] = fma(-ti1, y1, gr[j]);
      gi[j] = fma(tr0, y0, gi[j]);
      gi[j] = fma(ti0, x0, gi[j]);
      gi[j] = fma(tr1, y1, gi[j]);
      gi[j] = fma(ti1, x1, gi[j]);
    }
  }
}

TPA_DEV float4 load_pw32x2(const float4* p) {
  typedef float f4v __attribute__((ext_vector_type(4)));
  const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}

// The evaluation pass over the FP32 copy of the cache (a.h32), for LG >= 2: the pipelined
// scheme of k_eval_pipe with each lane holding two helicity components of a partial wave
// (lane = ev * (LG/2 * KG) + kg * LG/2 + lambda/2), so a lane's loads stay 16 bytes wide and
// a wave covers twice the events of the FP64 pass. The FP32 values are widened to FP64 on
// load; every sum is FP64. Requires N >= 1.
template <int LG, int KG, int PML>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PML <= 6 ? 3 : 1))) void k_eval_pipe32(EvalArgs a) {
  static_assert(LG >= 2, "two helicity components per lane");
  constexpr int LGL = LG / 2;
  constexpr int LPE = LGL * KG;
  constexpr int EPW = 64 / LPE;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int lamg = lane & (LGL - 1);
  const int kg = (lane / LGL) & (KG - 1);
  const int el = lane / LPE;
  const bool leader = lamg == 0 && kg == 0;
  double vs = 0.0, vc = 0.0;
  double gr[PML], gi[PML], cr[PML], ci[PML];
  const float4* base[PML];
#pragma unroll
  for (int j = 0; j < PML; ++j) {
    const int k = kg + KG * j;
    gr[j] = gi[j] = 0.0;
    const double2 c = k < a.P ? eval_coef(a, k) : make_double2(0.0, 0.0);
    cr[j] = c.x;
    ci[j] = c.y;
    const int kc = k < a.P ? k : a.P - 1;
    base[j] = reinterpret_cast<const float4*>(a.h32 + (int64_t)kc * a.N * LG) + lamg;
  }
  const int64_t last = a.N - 1;
  auto load = [&](float4 (&hv)[PML], double& wv, int64_t e) {
    const int64_t ec = e < a.N ? e : last;
#pragma unroll
    for (int j = 0; j < PML; ++j) hv[j] = load_pw32x2(base[j] + ec * LGL);
    wv = a.w[ec];
  };
  const int64_t stride = (int64_t)gridDim.x * 4 * EPW;
  float4 h0[PML], h1[PML];
  double w0, w1;
  const int wave_u = __builtin_amdgcn_readfirstlane(wave);
  const int64_t first = ((int64_t)blockIdx.x * 4 + wave_u) * EPW;
  int64_t e = first + el;
  load(h0, w0, e);
  for (int64_t b = first; b < a.N; b += 2 * stride, e += 2 * stride) {
    load(h1, w1, e + stride);
    __builtin_amdgcn_sched_barrier(0);
    eval_group2<LGL, KG, PML>(a.mode, a.grad, a.f_out, h0, w0, e, e < a.N, leader, cr, ci, vs,
                              vc, gr, gi);
    if (b + stride >= a.N) break;
    load(h0, w0, e + 2 * stride);
    __builtin_amdgcn_sched_barrier(0);
    eval_group2<LGL, KG, PML>(a.mode, a.grad, a.f_out, h1, w1, e + stride, e + stride < a.N,
                              leader, cr, ci, vs, vc, gr, gi);
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const double s2 = __shfl_xor(vs, m, 64);
    const double c2 = __shfl_xor(vc, m, 64);
    neumaier_merge(vs, vc, s2, c2);
  }
  __shared__ double sh[4][TPA_NACC_MAX];
  if (a.grad) {
#pragma unroll
    for (int j = 0; j < PML; ++j) {
      double xr = butterfly<1, LGL>(gr[j]), xi = butterfly<1, LGL>(gi[j]);
      xr = butterfly<LPE, 64>(xr);
      xi = butterfly<LPE, 64>(xi);
      const int k = kg + KG * j;
      if (el == 0 && lamg == 0 && k < a.P) {
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
  // slot-major partials: slot j of this block at out[j * gridDim.x] (k_reduce reads each
  // slot's blocks contiguously)
  double* out = a.partials + blockIdx.x;
  if (threadIdx.x == 0) {
    double s = sh[0][0], c = sh[0][1];
    for (int wv = 1; wv < 4; ++wv) neumaier_merge(s, c, sh[wv][0], sh[wv][1]);
    out[0] = s;
    out[gridDim.x] = c;
  }
  if (a.grad) {
    for (int j = 2 + threadIdx.x; j < a.nacc; j += blockDim.x)
      out[(int64_t)j * gridDim.x] = ((sh[0][j] + sh[1][j]) + sh[2][j]) + sh[3][j];
  }
}

// Fixed-order reduction over the per-block partial sums: one block per output slot. Slot 0/1
// are the (sum, compensation) pair of the value and fold into out[0]; every other slot is a
// compensated sum. Each thread reduces a fixed strided subset, a fixed butterfly merges the
// threads of a wave and thread 0 merges the four waves in order, so the result does not
// depend on scheduling.
// host_out: pinned host memory for the results. With cnt null every block writes its slot
// there (and fences at system scope), or to out (device memory) when host_out is null; with
// cnt non-null every block writes its slot to host_out and fences at system scope before it
// is counted on cnt, and the block counted last (which resets cnt to zero) writes seq to
// *flag, so the host can poll for the result instead of waiting on the stream.
__global__ __launch_bounds__(256) void k_reduce(const double* __restrict__ partials, int nblocks,
                                                int nacc, int slot_major,
                                                double* __restrict__ out,
                                                double* __restrict__ host_out,
                                                unsigned* __restrict__ cnt,
                                                volatile uint64_t* __restrict__ flag,
                                                uint64_t seq) {
  const int j = blockIdx.x;
  const int t = threadIdx.x;
  __shared__ double ss[4], cc[4];
  double s = 0.0, c = 0.0;
  // partial of block b, slot q: partials[q * nblocks + b] (slot-major, contiguous per slot) or
  // partials[b * nacc + q]
  const int64_t bs = slot_major ? 1 : nacc, qs = slot_major ? nblocks : 1;
  if (j != 1) {
    for (int b = t; b < nblocks; b += 256) neumaier_add(s, c, partials[b * bs + j * qs]);
    if (j == 0)
      for (int b = t; b < nblocks; b += 256) neumaier_add(s, c, partials[b * bs + qs]);
  }
  // compensated pairs merged over the wave (a fixed butterfly), then the four waves in order
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const double s2 = __shfl_xor(s, m, 64);
    const double c2 = __shfl_xor(c, m, 64);
    neumaier_merge(s, c, s2, c2);
  }
  if ((t & 63) == 0) {
    ss[t >> 6] = s;
    cc[t >> 6] = c;
  }
  __syncthreads();
  if (t != 0) return;
  s = ss[0];
  c = cc[0];
  for (int w = 1; w < 4; ++w) neumaier_merge(s, c, ss[w], cc[w]);
  const double v = (j == 1) ? 0.0 : s + c;
  if (cnt == nullptr) {
    if (host_out != nullptr) {
      host_out[j] = v;
      __threadfence_system();
    } else {
      out[j] = v;
    }
    return;
  }
  // each block's sum straight into the pinned output, made visible to the host before the
  // block is counted; the last block counted raises the flag the host polls
  host_out[j] = v;
  __threadfence_system();
  if (atomicAdd(cnt, 1u) != gridDim.x - 1) return;
  atomicExch(cnt, 0u);
  *flag = seq;
  __threadfence_system();
}

// Plain streaming copy (16-byte loads and stores), the calibration kernel for HBM byte
// counters: it moves exactly n * 16 bytes in and n * 16 bytes out.
__global__ __launch_bounds__(256) void k_copy(const double2* __restrict__ src,
                                              double2* __restrict__ dst, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

// FP32 copy of a partial-wave cache (round to nearest), for the passes that read FP32.
__global__ __launch_bounds__(256) void k_pw_to_f32(const double2* __restrict__ src,
                                                   float2* __restrict__ dst, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double2 v = src[i];
    dst[i] = make_float2((float)v.x, (float)v.y);
  }
}

// ------------------------------------------------------------- MFMA Gram (MC normalisation)
// G = sum_rows w_row Z_row Z_row^T with Z_row = [Re h_k, Im h_k] over rows = (event, lambda),
// from which I_kl = sum w conj(h_k) h_l = (Gxx + Gyy)_kl + i (Gxy_kl - Gxy_lk)
// (the cached_int matrix of tf_pwa/experimental/opt_int.py:30-52). Rows are staged through
// LDS in blocks of 256 (each k a coalesced 4 KB read), converted to FP32, and each wave
// accumulates 64 rows with v_mfma_f32_32x32x2f32 into three 32x32 tiles (xx, xy, yy) in FP32,
// folding them into FP64 accumulators every 16 rows. Requires P <= 32.
typedef float tpa_floatx16 __attribute__((ext_vector_type(16)));

#define TPA_GRAM_ROWS 256
#define TPA_GRAM_LDS_STRIDE (TPA_GRAM_ROWS + 1)
#define TPA_GRAM_SLOTS (3 * 16 * 64)

template <int RF>
__global__ __launch_bounds__(256) void k_gram_f32(const double2* __restrict__ h,
                                                  const double* __restrict__ w, double wscale,
                                                  int64_t N, int lam_pad, int lam, int P,
                                                  double* __restrict__ partials) {
  __shared__ float Xs[32 * TPA_GRAM_LDS_STRIDE];
  __shared__ float Ys[32 * TPA_GRAM_LDS_STRIDE];
  __shared__ float Ws[TPA_GRAM_ROWS];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int64_t R = N * lam_pad;
  double acc[3][16];
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[q][v] = 0.0;
  for (int64_t rb = (int64_t)blockIdx.x * TPA_GRAM_ROWS; rb < R;
       rb += (int64_t)gridDim.x * TPA_GRAM_ROWS) {
    const int64_t row = rb + t;
    float wr = 0.0f;
    if (row < R) {
      const int64_t e = row / lam_pad;
      const int l = (int)(row - e * lam_pad);
      // weights enter relative to their mean (wscale = 1/mean), so a uniform MC weight is
      // exactly 1.0f and its FP32 rounding cannot scale the whole matrix
      wr = l < lam ? (float)(w[e] * wscale) : 0.0f;
    }
    Ws[t] = wr;
    double2 hv[32];
#pragma unroll
    for (int k = 0; k < 32; ++k)
      hv[k] = (k < P && row < R) ? load_pw<true>(&h[(int64_t)k * R + row]) : make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      Xs[k * TPA_GRAM_LDS_STRIDE + t] = (float)hv[k].x;
      Ys[k * TPA_GRAM_LDS_STRIDE + t] = (float)hv[k].y;
    }
    __syncthreads();
    const int i = lane & 31;
    // FP32 accumulation runs over 16 rows (8 MFMA steps) before each FP64 refold, which keeps
    // the FP32 rounding of the partial sums at the level of single products
    for (int s0 = 0; s0 < 32; s0 += RF) {
      tpa_floatx16 cxx, cxy, cyy;
#pragma unroll
      for (int v = 0; v < 16; ++v) cxx[v] = cxy[v] = cyy[v] = 0.0f;
#pragma unroll
      for (int s = s0; s < s0 + RF; ++s) {
        const int rl = wave * 64 + 2 * s + (lane >> 5);
        const float x = Xs[i * TPA_GRAM_LDS_STRIDE + rl];
        const float y = Ys[i * TPA_GRAM_LDS_STRIDE + rl];
        const float ww = Ws[rl];
        cxx = __builtin_amdgcn_mfma_f32_32x32x2f32(x, ww * x, cxx, 0, 0, 0);
        cxy = __builtin_amdgcn_mfma_f32_32x32x2f32(x, ww * y, cxy, 0, 0, 0);
        cyy = __builtin_amdgcn_mfma_f32_32x32x2f32(y, ww * y, cyy, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        acc[0][v] += (double)cxx[v];
        acc[1][v] += (double)cxy[v];
        acc[2][v] += (double)cyy[v];
      }
    }
    __syncthreads();
  }
  // block sum of the four waves in a fixed order through LDS (reusing the staging area)
  double* red = reinterpret_cast<double*>(Xs);  // 3*16*64 doubles = 24 KB < sizeof(Xs)
  for (int wv = 0; wv < 4; ++wv) {
    if (wave == wv) {
#pragma unroll
      for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int slot = (q * 16 + v) * 64 + lane;
          red[slot] = (wv == 0) ? acc[q][v] : red[slot] + acc[q][v];
        }
    }
    __syncthreads();
  }
  double* out = partials + (int64_t)blockIdx.x * TPA_GRAM_SLOTS;
  for (int j = t; j < TPA_GRAM_SLOTS; j += blockDim.x) out[j] = red[j];
}

// Compensated fixed-order sum of every slot over the blocks (one block per slot).
__global__ __launch_bounds__(256) void k_reduce_slots(const double* __restrict__ partials,
                                                      int nblocks, int nslots,
                                                      double* __restrict__ out) {
  const int j = blockIdx.x;
  const int t = threadIdx.x;
  __shared__ double ss[256], cc[256];
  double s = 0.0, c = 0.0;
  for (int b = t; b < nblocks; b += 256) neumaier_add(s, c, partials[(int64_t)b * nslots + j]);
  ss[t] = s;
  cc[t] = c;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      double s1 = ss[t], c1 = cc[t];
      neumaier_merge(s1, c1, ss[t + w], cc[t + w]);
      ss[t] = s1;
      cc[t] = c1;
    }
    __syncthreads();
  }
  if (t == 0) out[j] = ss[0] + cc[0];
}

// ------------------------------------------------- lineshapes for floating masses / widths
// Per chain and event: B, dB/dm0, dB/dg0 at the chain's current (m0, g0), stored as
// bw[(q * n_chain + c) * N + e] for q = 0, 1, 2. One thread per (chain, event), the events of
// one chain in consecutive threads: the chains' lineshapes (divisions, square roots and the
// barrier polynomials, long dependent chains) run side by side instead of one after another
// in the event's thread, which at the 10^5-event sets of a fit left the pass latency-bound.
__global__ __launch_bounds__(256) void k_lineshape(const TpaModel* __restrict__ M,
                                                   const double* __restrict__ ang, int64_t N,
                                                   double2* __restrict__ bw) {
  const int nc = M->n_chain;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N * nc) return;
  const int c = (int)(i / N);
  const int64_t e = i - (int64_t)c * N;
  const int64_t gbase = (int64_t)TPA_NF_TOPO * M->n_topo;
  const TpaChain& C = M->chain[c];
  const TpaTopo T = M->topo[C.topo];
  const int fy = C.r_y_first ? T.y : T.z;
  const int fz = C.r_y_first ? T.z : T.y;
  const double mR = ang[(int64_t)(TPA_NF_TOPO * C.topo + TPA_F_MR) * N + e];
  const double q = rel_p(mR, ang[(gbase + 1 + fy) * N + e], ang[(gbase + 1 + fz) * N + e]);
  cd B, dBm, dBg;
  lineshape_grad(C.shape, mR, C.m0, C.g0, q, C.m_y, C.m_z, C.bw_l, C.d, &B, &dBm, &dBg);
  bw[(int64_t)(0 * nc + c) * N + e] = make_double2(B.r, B.i);
  bw[(int64_t)(1 * nc + c) * N + e] = make_double2(dBm.r, dBm.i);
  bw[(int64_t)(2 * nc + c) * N + e] = make_double2(dBg.r, dBg.i);
}

// Evaluation with per-event lineshapes (partial-wave cache built without them):
// A = sum_k c_k B_c(k) h_k. Besides G_k = sum u B H_k it returns Dm_k = sum u dB/dm0 H_k and
// Dg_k = sum u dB/dg0 H_k with H_k = sum_lambda conj(A_lambda) h_k,lambda, from which the host
// forms the mass and width gradients.
struct EvalShapeArgs {
  EvalArgs a;
  const double2* __restrict__ bw;  // [3][n_chain][N]
  const int* __restrict__ kchain;  // [P] chain of each partial wave
  int n_chain;
};

#define TPA_NACC_SHAPE_MAX (2 + 6 * TPA_MAX_P)

template <int LG, int KG, int PML>
__global__ __launch_bounds__(256) void k_eval_shape(EvalShapeArgs s) {
  // The evaluation pass with the lineshapes per event (floating masses and widths): the cache
  // holds the waves without their Breit-Wigner; bw holds B_c, dB_c/dm0 and dB_c/dg0 per chain
  // and event (k_lineshape). Loads are unconditional (clamped to a valid event and wave) and
  // issued together; lane exchanges as in k_eval_pipe.
  const EvalArgs& a = s.a;
  constexpr int LPE = LG * KG;
  constexpr int EPW = 64 / LPE;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int lam = lane & (LG - 1);
  const int kg = (lane / LG) & (KG - 1);
  const int el = lane / LPE;
  const bool leader = lam == 0 && kg == 0;
  const double eps = 1.0e-6;
  double vs = 0.0, vc = 0.0;
  double gr[PML], gi[PML], mr[PML], mi[PML], wr[PML], wi[PML], cr[PML], ci[PML];
  const double2* hb[PML];
  const double2* bb[PML];
#pragma unroll
  for (int j = 0; j < PML; ++j) {
    const int k = kg + KG * j;
    gr[j] = gi[j] = mr[j] = mi[j] = wr[j] = wi[j] = 0.0;
    const double2 c = k < a.P ? a.coef[k] : make_double2(0.0, 0.0);
    cr[j] = c.x;
    ci[j] = c.y;
    const int kc = k < a.P ? k : a.P - 1;
    hb[j] = a.h + (int64_t)kc * a.N * LG + lam;
    bb[j] = s.bw + (int64_t)s.kchain[kc] * a.N;
  }
  const int64_t nc = (int64_t)s.n_chain * a.N;  // B, dB/dm0, dB/dg0 rows are nc apart
  const int64_t last = a.N - 1;
  const int64_t stride = (int64_t)gridDim.x * 4 * EPW;
  const int wave_u = __builtin_amdgcn_readfirstlane(wave);
  // the waves, lineshapes and weight of the next event group are loaded before the current
  // one is evaluated (one group ahead; the lineshape derivatives at the start of their group)
  double2 hn[PML], bn[PML];
  double wn;
  auto load_next = [&](int64_t bnext) {
    const int64_t en = bnext + el;
    const int64_t ecn = en < a.N ? en : last;
#pragma unroll
    for (int j = 0; j < PML; ++j) {
      hn[j] = load_pw<true>(hb[j] + ecn * LG);
      bn[j] = bb[j][ecn];
    }
    wn = a.w[ecn];
  };
  load_next(((int64_t)blockIdx.x * 4 + wave_u) * EPW);
  for (int64_t b = ((int64_t)blockIdx.x * 4 + wave_u) * EPW; b < a.N; b += stride) {
    const int64_t e = b + el;
    const bool valid = e < a.N;
    const int64_t ec = valid ? e : last;
    double2 hv[PML], bv[PML], dm[PML], dg[PML];
#pragma unroll
    for (int j = 0; j < PML; ++j) {
      hv[j] = hn[j];
      bv[j] = bn[j];
    }
    const double w = valid ? wn : 0.0;
    // the derivatives of this group before the next group's loads: the waits for them then
    // leave the next group's loads in flight
    if (a.grad) {
#pragma unroll
      for (int j = 0; j < PML; ++j) {
        dm[j] = bb[j][nc + ec];
        dg[j] = bb[j][2 * nc + ec];
      }
    }
    load_next(b + stride);
    double ar = 0.0, ai = 0.0;
#pragma unroll
    for (int j = 0; j < PML; ++j) {
      const double cbr = cr[j] * bv[j].x - ci[j] * bv[j].y;
      const double cbi = cr[j] * bv[j].y + ci[j] * bv[j].x;
      ar += cbr * hv[j].x - cbi * hv[j].y;
      ai += cbr * hv[j].y + cbi * hv[j].x;
    }
    ar = butterfly<LG, LPE>(ar);
    ai = butterfly<LG, LPE>(ai);
    const double f = butterfly<1, LG>(ar * ar + ai * ai);
    double y, u;
    if (a.mode == 0) {
      // clip_log(f) (tf_pwa/model/model.py:283-292) weighted by w; the log and reciprocal of
      // tpa_log.h (about half the library's FP64 instructions)
      double cl, dcl;
      if (f > eps) {
        dcl = tpa_recip_pos(f, hw_rcp);
        cl = tpa_log_pos(f, hw_rcp);
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
    if (a.f_out != nullptr && valid && leader) a.f_out[e] = f;
    neumaier_add(vs, vc, leader ? y : 0.0);
    if (a.grad) {
#pragma unroll
      for (int j = 0; j < PML; ++j) {
        // u * conj(A) h_k, then times B, dB/dm0, dB/dg0 of the wave's chain
        const double hr = u * (ar * hv[j].x + ai * hv[j].y);
        const double hi = u * (ar * hv[j].y - ai * hv[j].x);
        gr[j] += bv[j].x * hr - bv[j].y * hi;
        gi[j] += bv[j].x * hi + bv[j].y * hr;
        mr[j] += dm[j].x * hr - dm[j].y * hi;
        mi[j] += dm[j].x * hi + dm[j].y * hr;
        wr[j] += dg[j].x * hr - dg[j].y * hi;
        wi[j] += dg[j].x * hi + dg[j].y * hr;
      }
    }
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const double s2 = __shfl_xor(vs, m, 64);
    const double c2 = __shfl_xor(vc, m, 64);
    neumaier_merge(vs, vc, s2, c2);
  }
  __shared__ double sh[4][TPA_NACC_SHAPE_MAX];
  if (a.grad) {
#pragma unroll
    for (int j = 0; j < PML; ++j) {
      double x[6] = {gr[j], gi[j], mr[j], mi[j], wr[j], wi[j]};
#pragma unroll
      for (int q = 0; q < 6; ++q) x[q] = butterfly<LPE, 64>(butterfly<1, LG>(x[q]));
      const int k = kg + KG * j;
      if (el == 0 && lam == 0 && k < a.P) {
        sh[wave][2 + 2 * k] = x[0];
        sh[wave][3 + 2 * k] = x[1];
        sh[wave][2 + 2 * a.P + 2 * k] = x[2];
        sh[wave][3 + 2 * a.P + 2 * k] = x[3];
        sh[wave][2 + 4 * a.P + 2 * k] = x[4];
        sh[wave][3 + 4 * a.P + 2 * k] = x[5];
      }
    }
  }
  if (lane == 0) {
    sh[wave][0] = vs;
    sh[wave][1] = vc;
  }
  __syncthreads();
  // slot-major partials: slot j of this block at out[j * gridDim.x] (k_reduce reads each
  // slot's blocks contiguously)
  double* out = a.partials + blockIdx.x;
  if (threadIdx.x == 0) {
    double s0 = sh[0][0], c0 = sh[0][1];
    for (int wv = 1; wv < 4; ++wv) neumaier_merge(s0, c0, sh[wv][0], sh[wv][1]);
    out[0] = s0;
    out[gridDim.x] = c0;
  }
  if (a.grad) {
    for (int j = 2 + threadIdx.x; j < a.nacc; j += blockDim.x)
      out[(int64_t)j * gridDim.x] = ((sh[0][j] + sh[1][j]) + sh[2][j]) + sh[3][j];
  }
}

// ------------------------------------------------------------ batched multi-fit evaluation
// Serving: many fits keep their data in one device pool (stride S). A batch is a list of
// chunks (<= TPA_CHUNK events of one request each); block b evaluates chunk b with that
// request's coefficients and writes its partial sums, and k_reduce_items sums each request's
// chunks in order. Counts live in device memory so one captured graph serves any batch up to
// its grid size.
#define TPA_CHUNK 2048

// A batch's work list, one blob (one copy to the device per batch), in int64 words:
// [0] chunks n, [1] requests m, then chunk_start[n], chunk_end[n], chunk_item[n],
// item_chunk[m + 1] (each request's first chunk, and the end), then at the next even word the
// coefficients [m][P] (double2).
struct BatchView {
  int64_t n, m;
  const int64_t* start;
  const int64_t* end;
  const int64_t* item;
  const int64_t* item_chunk;
  const double2* coef;
};
TPA_DEV BatchView batch_view(const int64_t* __restrict__ blob) {
  BatchView v;
  v.n = blob[0];
  v.m = blob[1];
  v.start = blob + 2;
  v.end = v.start + v.n;
  v.item = v.end + v.n;
  v.item_chunk = v.item + v.n;
  v.coef = reinterpret_cast<const double2*>(blob + ((2 + 3 * v.n + v.m + 1 + 1) & ~int64_t(1)));
  return v;
}

struct MultiArgs {
  const double2* __restrict__ h;     // pool cache [P][S][LG]
  const float2* __restrict__ h32;    // its FP32 copy [P][S][LG] (k_eval_multi32), or null
  const double* __restrict__ w;      // pool weights [S]
  const int64_t* __restrict__ blob;  // the batch (BatchView)
  double* __restrict__ partials;     // [chunk][nacc]
  int64_t S;
  int P;
  int mode;
  int nacc;
};

template <int LG, int KG, int PML>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PML <= 6 ? 3 : 1))) void k_eval_multi(MultiArgs a) {
  const int chunk = blockIdx.x;
  const BatchView bv = batch_view(a.blob);
  if (chunk >= bv.n) return;  // uniform per block, before any barrier
  constexpr int LPE = LG * KG;
  constexpr int EPW = 64 / LPE;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int lam = lane & (LG - 1);
  const int kg = (lane / LG) & (KG - 1);
  const int el = lane / LPE;
  const bool leader = lam == 0 && kg == 0;
  const int64_t item = bv.item[chunk];
  const int64_t end = bv.end[chunk];  // > start: chunks are never empty
  const double2* __restrict__ coef = bv.coef + item * a.P;
  double vs = 0.0, vc = 0.0;
  double gr[PML], gi[PML], cr[PML], ci[PML];
  const double2* base[PML];
#pragma unroll
  for (int j = 0; j < PML; ++j) {
    const int k = kg + KG * j;
    gr[j] = gi[j] = 0.0;
    const double2 c = k < a.P ? coef[k] : make_double2(0.0, 0.0);
    cr[j] = c.x;
    ci[j] = c.y;
    const int kc = k < a.P ? k : a.P - 1;
    base[j] = a.h + (int64_t)kc * a.S * LG + lam;
  }
  // the chunk's events, software-pipelined as in k_eval_pipe (two register stages, loads of
  // the next group issued before the current one is processed, clamped unconditional loads)
  const int64_t last = end - 1;
  auto load = [&](double2 (&hv)[PML], double& wv, int64_t e) {
    const int64_t ec = e < end ? e : last;
#pragma unroll
    for (int j = 0; j < PML; ++j) hv[j] = load_pw<true>(base[j] + ec * LG);
    wv = a.w[ec];
  };
  constexpr int64_t stride = 4 * EPW;
  double2 h0[PML], h1[PML];
  double w0, w1;
  const int wave_u = __builtin_amdgcn_readfirstlane(wave);
  const int64_t first = bv.start[chunk] + wave_u * EPW;
  int64_t e = first + el;
  load(h0, w0, e);
  for (int64_t b = first; b < end; b += 2 * stride, e += 2 * stride) {
    load(h1, w1, e + stride);
    __builtin_amdgcn_sched_barrier(0);
    eval_group<LG, KG, PML>(a.mode, true, nullptr, h0, w0, e, e < end, leader, cr, ci, vs, vc, gr,
                            gi);
    if (b + stride >= end) break;
    load(h0, w0, e + 2 * stride);
    __builtin_amdgcn_sched_barrier(0);
    eval_group<LG, KG, PML>(a.mode, true, nullptr, h1, w1, e + stride, e + stride < end, leader,
                            cr, ci, vs, vc, gr, gi);
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const double s2 = __shfl_xor(vs, m, 64);
    const double c2 = __shfl_xor(vc, m, 64);
    neumaier_merge(vs, vc, s2, c2);
  }
  __shared__ double sh[4][TPA_NACC_MAX];
#pragma unroll
  for (int j = 0; j < PML; ++j) {
    double xr = butterfly<1, LG>(gr[j]), xi = butterfly<1, LG>(gi[j]);
    xr = butterfly<LPE, 64>(xr);
    xi = butterfly<LPE, 64>(xi);
    const int k = kg + KG * j;
    if (el == 0 && lam == 0 && k < a.P) {
      sh[wave][2 + 2 * k] = xr;
      sh[wave][3 + 2 * k] = xi;
    }
  }
  if (lane == 0) {
    sh[wave][0] = vs;
    sh[wave][1] = vc;
  }
  __syncthreads();
  double* out = a.partials + (int64_t)chunk * a.nacc;
  if (threadIdx.x == 0) {
    double s0 = sh[0][0], c0 = sh[0][1];
    for (int wv = 1; wv < 4; ++wv) neumaier_merge(s0, c0, sh[wv][0], sh[wv][1]);
    out[0] = s0;
    out[1] = c0;
  }
  for (int j = 2 + threadIdx.x; j < a.nacc; j += blockDim.x)
    out[j] = ((sh[0][j] + sh[1][j]) + sh[2][j]) + sh[3][j];
}

// k_eval_multi over the pool's FP32 copy (a.h32), two helicity components per lane as in
// k_eval_pipe32 (LG >= 2); FP64 arithmetic after the load.
template <int LG, int KG, int PML>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PML <= 6 ? 3 : 1))) void k_eval_multi32(MultiArgs a) {
  static_assert(LG >= 2, "two helicity components per lane");
  const int chunk = blockIdx.x;
  const BatchView bv = batch_view(a.blob);
  if (chunk >= bv.n) return;  // uniform per block, before any barrier
  constexpr int LGL = LG / 2;
  constexpr int LPE = LGL * KG;
  constexpr int EPW = 64 / LPE;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int lamg = lane & (LGL - 1);
  const int kg = (lane / LGL) & (KG - 1);
  const int el = lane / LPE;
  const bool leader = lamg == 0 && kg == 0;
  const int64_t item = bv.item[chunk];
  const int64_t end = bv.end[chunk];  // > start: chunks are never empty
  const double2* __restrict__ coef = bv.coef + item * a.P;
  double vs = 0.0, vc = 0.0;
  double gr[PML], gi[PML], cr[PML], ci[PML];
  const float4* base[PML];
#pragma unroll
  for (int j = 0; j < PML; ++j) {
    const int k = kg + KG * j;
    gr[j] = gi[j] = 0.0;
    const double2 c = k < a.P ? coef[k] : make_double2(0.0, 0.0);
    cr[j] = c.x;
    ci[j] = c.y;
    const int kc = k < a.P ? k : a.P - 1;
    base[j] = reinterpret_cast<const float4*>(a.h32 + (int64_t)kc * a.S * LG) + lamg;
  }
  const int64_t last = end - 1;
  auto load = [&](float4 (&hv)[PML], double& wv, int64_t e) {
    const int64_t ec = e < end ? e : last;
#pragma unroll
    for (int j = 0; j < PML; ++j) hv[j] = load_pw32x2(base[j] + ec * LGL);
    wv = a.w[ec];
  };
  constexpr int64_t stride = 4 * EPW;
  float4 h0[PML], h1[PML];
  double w0, w1;
  const int wave_u = __builtin_amdgcn_readfirstlane(wave);
  const int64_t first = bv.start[chunk] + wave_u * EPW;
  int64_t e = first + el;
  load(h0, w0, e);
  for (int64_t b = first; b < end; b += 2 * stride, e += 2 * stride) {
    load(h1, w1, e + stride);
    __builtin_amdgcn_sched_barrier(0);
    eval_group2<LGL, KG, PML>(a.mode, true, nullptr, h0, w0, e, e < end, leader, cr, ci, vs, vc,
                              gr, gi);
    if (b + stride >= end) break;
    load(h0, w0, e + 2 * stride);
    __builtin_amdgcn_sched_barrier(0);
    eval_group2<LGL, KG, PML>(a.mode, true, nullptr, h1, w1, e + stride, e + stride < end,
                              leader, cr, ci, vs, vc, gr, gi);
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const double s2 = __shfl_xor(vs, m, 64);
    const double c2 = __shfl_xor(vc, m, 64);
    neumaier_merge(vs, vc, s2, c2);
  }
  __shared__ double sh[4][TPA_NACC_MAX];
#pragma unroll
  for (int j = 0; j < PML; ++j) {
    double xr = butterfly<1, LGL>(gr[j]), xi = butterfly<1, LGL>(gi[j]);
    xr = butterfly<LPE, 64>(xr);
    xi = butterfly<LPE, 64>(xi);
    const int k = kg + KG * j;
    if (el == 0 && lamg == 0 && k < a.P) {
      sh[wave][2 + 2 * k] = xr;
      sh[wave][3 + 2 * k] = xi;
    }
  }
  if (lane == 0) {
    sh[wave][0] = vs;
    sh[wave][1] = vc;
  }
  __syncthreads();
  double* out = a.partials + (int64_t)chunk * a.nacc;
  if (threadIdx.x == 0) {
    double s0 = sh[0][0], c0 = sh[0][1];
    for (int wv = 1; wv < 4; ++wv) neumaier_merge(s0, c0, sh[wv][0], sh[wv][1]);
    out[0] = s0;
    out[1] = c0;
  }
  for (int j = 2 + threadIdx.x; j < a.nacc; j += blockDim.x)
    out[j] = ((sh[0][j] + sh[1][j]) + sh[2][j]) + sh[3][j];
}

// FP32 copy of rows [start, start + count) of every partial-wave row k < P of a pool cache
// (row stride `stride` complex values), rounded to nearest.
__global__ __launch_bounds__(256) void k_pw_rows_to_f32(const double2* __restrict__ src,
                                                        float2* __restrict__ dst, int P,
                                                        int64_t stride, int64_t start,
                                                        int64_t count) {
  const int64_t total = (int64_t)P * count;
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += step) {
    const int64_t k = i / count, r = i - k * count;
    const int64_t x = k * stride + start + r;
    const double2 v = src[x];
    dst[x] = make_float2((float)v.x, (float)v.y);
  }
}

// Per request: the compensated, fixed-order sum of its chunks (item_chunk[i] .. item_chunk[i+1]).
// R = blockDim.x / nacc threads share a slot: thread (slot j, replica r) sums chunks r, r + R, ..
// of the request in order (the loads of one thread are independent, so several stay in flight),
// and the R partial sums are merged in r order. R depends only on nacc, so a request's result
// does not depend on the batch around it.
__global__ __launch_bounds__(256) void k_reduce_items(const double* __restrict__ partials,
                                                      const int64_t* __restrict__ blob,
                                                      int nacc, double* __restrict__ out) {
  const int item = blockIdx.x;
  const BatchView bv = batch_view(blob);
  if (item >= bv.m) return;  // uniform per block, before any barrier
  const int64_t c0 = bv.item_chunk[item], c1 = bv.item_chunk[item + 1];
  const int R = nacc <= (int)blockDim.x ? (int)blockDim.x / nacc : 1;
  __shared__ double ss[TPA_NACC_MAX > 256 ? TPA_NACC_MAX : 256];
  __shared__ double sc[TPA_NACC_MAX > 256 ? TPA_NACC_MAX : 256];
  for (int x = threadIdx.x; x < R * nacc; x += blockDim.x) {
    const int j = x % nacc, r = x / nacc;
    double s = 0.0, c = 0.0;
    if (j != 1) {
#pragma unroll 4
      for (int64_t b = c0 + r; b < c1; b += R) neumaier_add(s, c, partials[b * nacc + j]);
      if (j == 0) {
#pragma unroll 4
        for (int64_t b = c0 + r; b < c1; b += R) neumaier_add(s, c, partials[b * nacc + 1]);
      }
    }
    ss[x] = s;
    sc[x] = c;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < nacc; j += blockDim.x) {
    double s = ss[j], c = sc[j];
    for (int r = 1; r < R; ++r) neumaier_merge(s, c, ss[r * nacc + j], sc[r * nacc + j]);
    out[(int64_t)item * nacc + j] = (j == 1) ? 0.0 : s + c;
  }
}

// ------------------------------------------------------------------ second derivatives
// One pass over events [off, off+n) of a fold-1 cache with coefficients c that accumulates,
// besides the sum and its adjoint G_k, the two real second-moment matrices from which the
// exact Hessian of the sum in (c, conj c) follows (Q = 2P):
//   O1 = sum_e v_e z_e z_e^T,            z_e = (Re q_e, Im q_e), q_ek = sum_l conj(A_el) h_ekl
//   O2 = sum_e u_e sum_l y_el y_el^T,    y_el = (Re h_e.l, Im h_e.l)
// with u_e = w dL/df and v_e = w d2L/df2 (data: L = clip_log of f w / dom; MC: L = f, v = 0,
// and O2 is the FP64 Gram matrix). Both are stored as packed upper triangles (row-major, i <= j)
// after the value and G (interleaved re / im, the layout of k_eval_fold's partials).
// A block stages TPA_HESS_PAIRS (event, helicity) pairs of the cache in LDS per tile and each
// thread owns one or two 4x4 register blocks of O1 / O2.
#define TPA_HESS_PAIRS 128
#define TPA_HESS_PMAX 32

struct HessArgs {
  const double2* __restrict__ h;     // [P][S][LG]
  const double* __restrict__ w;      // [S]
  const double2* __restrict__ coef;  // [P]
  double* __restrict__ partials;     // [grid][nacc]
  int64_t off, n, S;
  int P;
  int mode;
  int nacc;  // 2 + 2P + Q(Q+1)
};

TPA_DEV int tri_index(int i, int j, int Q) { return i * Q - (i * (i - 1)) / 2 + (j - i); }

template <int LG>
__global__ __launch_bounds__(256) void k_hess(HessArgs a) {
  constexpr int E = TPA_HESS_PAIRS / LG;  // events per tile
  const int P = a.P, Q = 2 * P;
  const int nb = (Q + 3) / 4;             // 4-wide row/column blocks
  const int NB = nb * (nb + 1) / 2;        // upper-triangle blocks per matrix
  __shared__ double2 hs[TPA_HESS_PMAX][TPA_HESS_PAIRS];
  __shared__ double2 As[TPA_HESS_PAIRS];
  __shared__ double zs[E][2 * TPA_HESS_PMAX + 4];
  __shared__ double us[E], vs[E];
  __shared__ double red_s[256], red_c[256];
  const int t = threadIdx.x;
  const double eps = 1.0e-6;
  const double logeps = log(1.0e-6);
  // the (up to two) register blocks of this thread: ids < NB are O1 blocks, then O2
  int blk[2], bi[2], bj[2];
  for (int s = 0; s < 2; ++s) {
    blk[s] = t + 256 * s;
    bi[s] = bj[s] = 0;
    if (blk[s] >= 2 * NB) continue;  // no block for this slot
    int id = blk[s] < NB ? blk[s] : blk[s] - NB;  // id < NB = nb (nb + 1) / 2
    int r = 0;
    while (id >= nb - r) {
      id -= nb - r;
      ++r;
    }
    bi[s] = r;
    bj[s] = r + id;
  }
  double acc[2][4][4];
  for (int s = 0; s < 2; ++s)
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) acc[s][i][j] = 0.0;
  double gacc = 0.0;  // thread t < Q: component t of (Re G, Im G)
  double vsum = 0.0, vcomp = 0.0;
  const int64_t tiles = (a.n + E - 1) / E;
  for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const int64_t e0 = a.off + tile * E;
    const int64_t e_end = a.off + a.n;
    // stage the tile's cache: for each wave, E * LG contiguous pairs
    for (int idx = t; idx < P * TPA_HESS_PAIRS; idx += 256) {
      const int k = idx / TPA_HESS_PAIRS, p = idx % TPA_HESS_PAIRS;
      const int64_t e = e0 + p / LG;
      hs[k][p] = (e < e_end) ? load_pw<true>(&a.h[((int64_t)k * a.S + e) * LG + (p % LG)])
                             : make_double2(0.0, 0.0);
    }
    __syncthreads();
    if (t < TPA_HESS_PAIRS) {
      double ar = 0.0, ai = 0.0;
      for (int k = 0; k < P; ++k) {
        const double2 c = a.coef[k], hv = hs[k][t];
        ar += c.x * hv.x - c.y * hv.y;
        ai += c.x * hv.y + c.y * hv.x;
      }
      As[t] = make_double2(ar, ai);
    }
    __syncthreads();
    if (t < E) {
      const int64_t e = e0 + t;
      double u = 0.0, v = 0.0;
      if (e < e_end) {
        double f = 0.0;
        for (int l = 0; l < LG; ++l) f += As[t * LG + l].x * As[t * LG + l].x + As[t * LG + l].y * As[t * LG + l].y;
        const double w = a.w[e];
        double y;
        if (a.mode == 0) {
          const double dom = (w == 0.0) ? 1.0 : w;
          const double r = w / dom;
          const double x = f * r;
          double cl, d1, d2;
          if (x > eps) {
            cl = log(x);
            d1 = 1.0 / x;
            d2 = -1.0 / (x * x);
          } else {
            const double tt = (x - eps) / eps;
            cl = (logeps + tt) - (tt * tt) / 2.0;
            d1 = 1.0 / eps - tt / eps;
            d2 = -1.0 / (eps * eps);
          }
          y = w * cl;
          u = (w * d1) * r;
          v = (w * d2) * (r * r);
        } else {
          y = w * f;
          u = w;
        }
        neumaier_add(vsum, vcomp, y);
      }
      us[t] = u;
      vs[t] = v;
    }
    for (int idx = t; idx < E * P; idx += 256) {
      const int e = idx / P, k = idx % P;
      double qr = 0.0, qi = 0.0;
      for (int l = 0; l < LG; ++l) {
        const double2 A = As[e * LG + l], hv = hs[k][e * LG + l];
        qr += A.x * hv.x + A.y * hv.y;  // conj(A) h
        qi += A.x * hv.y - A.y * hv.x;
      }
      zs[e][k] = qr;
      zs[e][P + k] = qi;
    }
    __syncthreads();
    if (t < Q)
      for (int e = 0; e < E; ++e) gacc += us[e] * zs[e][t];
    for (int s = 0; s < 2; ++s) {
      if (blk[s] >= 2 * NB) continue;
      const int r0 = 4 * bi[s], c0 = 4 * bj[s];
      if (blk[s] < NB) {
        for (int e = 0; e < E; ++e) {
          double x[4], y[4];
          for (int i = 0; i < 4; ++i) {
            x[i] = (r0 + i < Q) ? vs[e] * zs[e][r0 + i] : 0.0;
            y[i] = (c0 + i < Q) ? zs[e][c0 + i] : 0.0;
          }
          for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) acc[s][i][j] += x[i] * y[j];
        }
      } else {
        for (int p = 0; p < TPA_HESS_PAIRS; ++p) {
          const double u = us[p / LG];
          double x[4], y[4];
          for (int i = 0; i < 4; ++i) {
            const int ri = r0 + i, ci = c0 + i;
            const double hr = ri < P ? hs[ri][p].x : (ri < Q ? hs[ri - P][p].y : 0.0);
            const double hc = ci < P ? hs[ci][p].x : (ci < Q ? hs[ci - P][p].y : 0.0);
            x[i] = u * hr;
            y[i] = hc;
          }
          for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) acc[s][i][j] += x[i] * y[j];
        }
      }
    }
    __syncthreads();
  }
  // write this block's partials: [0] value, [1] compensation, [2..2+Q) G, O1, O2 triangles
  double* out = a.partials + (int64_t)blockIdx.x * a.nacc;
  red_s[t] = vsum;
  red_c[t] = vcomp;
  __syncthreads();
  for (int m = 128; m > 0; m >>= 1) {
    if (t < m) {
      double s1 = red_s[t], c1 = red_c[t];
      neumaier_merge(s1, c1, red_s[t + m], red_c[t + m]);
      red_s[t] = s1;
      red_c[t] = c1;
    }
    __syncthreads();
  }
  if (t == 0) {
    out[0] = red_s[0];
    out[1] = red_c[0];
  }
  if (t < Q) out[t < P ? 2 + 2 * t : 3 + 2 * (t - P)] = gacc;  // interleaved re / im
  const int T = Q * (Q + 1) / 2;
  for (int s = 0; s < 2; ++s) {
    if (blk[s] >= 2 * NB) continue;
    double* o = out + 2 + Q + (blk[s] < NB ? 0 : T);
    const int r0 = 4 * bi[s], c0 = 4 * bj[s];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        const int ri = r0 + i, cj = c0 + j;
        if (ri < Q && cj < Q && ri <= cj) o[tri_index(ri, cj, Q)] = acc[s][i][j];
      }
  }
}

// Four doubles: the accumulator of one v_mfma_f64_16x16x4f64 tile per lane.
typedef double tpa_doublex4 __attribute__((ext_vector_type(4)));

// Second moments on the matrix cores. The same sums as k_hess, reorganised as two weighted
// Gram products per tile of R = 128 (event, helicity) rows staged in LDS:
//   O2 = Z^T diag(u) Z with Z the rows' [Re h, Im h]      (every row, weight u of its event)
//   O1 = Y^T diag(v) Y with Y the events' [Re q, Im q]    (every event)
// each on v_mfma_f64_16x16x4f64 over the upper 16 x 16 tiles (QT = ceil(2P / 16) tiles a side;
// wave w takes rows 32w .. 32w + 31 of the tile for O2 and every fourth group of 4 events for
// O1). A_lambda (one thread per row), the mode's u, v and value (one thread per event), q (one
// thread per event and wave) and G = sum u q are FP64 VALU work on the staged tile.
// Partials per block: value, compensation, G (interleaved re / im), then the O1 and O2 tiles in
// the accumulator layout (tile q, register r, lane l at (4 q + r) * 64 + l).
template <int LG, int QT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_hess_mfma(HessArgs a) {
  constexpr int R = 128;
  constexpr int E = R / LG;
  constexpr int C = 16 * QT;
  constexpr int NT = QT * (QT + 1) / 2;
  constexpr int ZS = R + 1, YS = E + 1;
  __shared__ double Zs[C * ZS];
  __shared__ double Ys[C * YS];
  __shared__ double2 As[R];
  __shared__ double ur[R];
  __shared__ double ve[E];
  const int P = a.P, Q = 2 * P;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int il = lane & 15, kl = lane >> 4;
  const double eps = 1.0e-6;
  // padding columns stay zero
  for (int idx = t; idx < (C - Q) * ZS; idx += 256) Zs[Q * ZS + idx] = 0.0;
  for (int idx = t; idx < (C - Q) * YS; idx += 256) Ys[Q * YS + idx] = 0.0;
  tpa_doublex4 o1[NT], o2[NT];
#pragma unroll
  for (int q = 0; q < NT; ++q) o1[q] = o2[q] = tpa_doublex4{0.0, 0.0, 0.0, 0.0};
  double gacc = 0.0;  // thread t < Q: component t of (Re G, Im G)
  double vsum = 0.0, vcomp = 0.0;
  const int64_t tiles = (a.n + E - 1) / E;
  const int64_t e_end = a.off + a.n;
  // staging map: thread t holds (event, helicity) row t % R of waves k = t / R + 2 j; a tile's
  // loads are issued into registers during the previous tile's work and written to LDS at the
  // start of its own (without this overlap 68% of the wave cycles went to waiting on them)
  static_assert(R == 128, "the staging map assumes 128 rows per tile");
  constexpr int NPF = 4 * QT;  // R x P <= 8 QT waves over 256 threads
  const int r_st = t & (R - 1), k_st = t >> 7;
  double2 pf[NPF];
  auto fetch = [&](int64_t tile) {
    const int64_t e = a.off + tile * E + r_st / LG;
#pragma unroll
    for (int j = 0; j < NPF; ++j) {
      const int k = k_st + 2 * j;
      pf[j] = (k < P && e < e_end)
                  ? load_pw<true>(&a.h[((int64_t)k * a.S + e) * LG + (r_st % LG)])
                  : make_double2(0.0, 0.0);
    }
  };
  if ((int64_t)blockIdx.x < tiles) fetch(blockIdx.x);
  for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const int64_t e0 = a.off + tile * E;
    // stage: for each wave k, R contiguous (event, helicity) pairs of the cache
#pragma unroll
    for (int j = 0; j < NPF; ++j) {
      const int k = k_st + 2 * j;
      if (k < P) {
        Zs[k * ZS + r_st] = pf[j].x;
        Zs[(P + k) * ZS + r_st] = pf[j].y;
      }
    }
    __syncthreads();
    if (tile + gridDim.x < tiles) fetch(tile + gridDim.x);
    if (t < R) {
      double ar = 0.0, ai = 0.0;
#pragma unroll 2
      for (int k = 0; k < P; ++k) {
        const double2 c = a.coef[k];
        const double hr = Zs[k * ZS + t], hi = Zs[(P + k) * ZS + t];
        ar += c.x * hr - c.y * hi;
        ai += c.x * hi + c.y * hr;
      }
      As[t] = make_double2(ar, ai);
    }
    __syncthreads();
    if (t < E) {
      const int64_t e = e0 + t;
      double u = 0.0, v = 0.0;
      if (e < e_end) {
        double f = 0.0;
        for (int l = 0; l < LG; ++l) f += As[t * LG + l].x * As[t * LG + l].x + As[t * LG + l].y * As[t * LG + l].y;
        const double w = a.w[e];
        double y;
        if (a.mode == 0) {
          // clip_log(f) and its first two derivatives (tf_pwa/model/model.py:283-292)
          double cl, d1, d2;
          if (f > eps) {
            cl = log(f);
            d1 = 1.0 / f;
            d2 = -1.0 / (f * f);
          } else {
            const double tt = (f - eps) / eps;
            cl = (log(eps) + tt) - (tt * tt) / 2.0;
            d1 = 1.0 / eps - tt / eps;
            d2 = -1.0 / (eps * eps);
          }
          y = w * cl;
          u = w * d1;
          v = w * d2;
        } else {
          y = w * f;
          u = w;
        }
        neumaier_add(vsum, vcomp, y);
      }
      for (int l = 0; l < LG; ++l) ur[t * LG + l] = u;
      ve[t] = v;
    }
    // q_ek = sum_l conj(A_el) h_ekl into Y (every event row, zero past the end)
#pragma unroll 1
    for (int idx = t; idx < E * P; idx += 256) {
      const int e = idx / P, k = idx - e * P;
      double qr = 0.0, qi = 0.0;
      for (int l = 0; l < LG; ++l) {
        const double2 A = As[e * LG + l];
        const double hr = Zs[k * ZS + e * LG + l], hi = Zs[(P + k) * ZS + e * LG + l];
        qr += A.x * hr + A.y * hi;
        qi += A.x * hi - A.y * hr;
      }
      Ys[k * YS + e] = qr;
      Ys[(P + k) * YS + e] = qi;
    }
    __syncthreads();
    if (t < Q)
      for (int e = 0; e < E; ++e) gacc += ur[e * LG] * Ys[t * YS + e];
    // O2 over the wave's 32 rows, 4 per MFMA step
#pragma unroll
    for (int st = 0; st < 8; ++st) {
      const int rl = wave * 32 + 4 * st + kl;
      const double uw = ur[rl];
      double z[QT];
#pragma unroll
      for (int ti = 0; ti < QT; ++ti) z[ti] = Zs[(16 * ti + il) * ZS + rl];
      int q = 0;
#pragma unroll
      for (int ti = 0; ti < QT; ++ti)
#pragma unroll
        for (int tj = ti; tj < QT; ++tj, ++q)
          o2[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(z[ti], uw * z[tj], o2[q], 0, 0, 0);
    }
    if (a.mode == 0) {
      // O1 over groups of 4 events, group g on wave g % 4
      for (int g = wave; g < E / 4; g += 4) {
        const int ev = 4 * g + kl;
        const double vw = ve[ev];
        double y[QT];
#pragma unroll
        for (int ti = 0; ti < QT; ++ti) y[ti] = Ys[(16 * ti + il) * YS + ev];
        int q = 0;
#pragma unroll
        for (int ti = 0; ti < QT; ++ti)
#pragma unroll
          for (int tj = ti; tj < QT; ++tj, ++q)
            o1[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(y[ti], vw * y[tj], o1[q], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // block partials: value pair, G, then the four waves' tiles summed in a fixed order
  double* out = a.partials + (int64_t)blockIdx.x * a.nacc;
  double* red = Zs;  // 2 * NT * 256 doubles fit in the staging area
  for (int wv = 0; wv < 4; ++wv) {
    if (wave == wv) {
#pragma unroll
      for (int q = 0; q < NT; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int s1 = (4 * q + r) * 64 + lane, s2 = NT * 256 + s1;
          red[s1] = wv == 0 ? o1[q][r] : red[s1] + o1[q][r];
          red[s2] = wv == 0 ? o2[q][r] : red[s2] + o2[q][r];
        }
    }
    __syncthreads();
  }
  for (int j = t; j < 2 * NT * 256; j += 256) out[2 + Q + j] = red[j];
  __syncthreads();
  double* rs = Zs;
  double* rc = Zs + 256;
  rs[t] = vsum;
  rc[t] = vcomp;
  __syncthreads();
  for (int m = 128; m > 0; m >>= 1) {
    if (t < m) {
      double s1 = rs[t], c1 = rc[t];
      neumaier_merge(s1, c1, rs[t + m], rc[t + m]);
      rs[t] = s1;
      rc[t] = c1;
    }
    __syncthreads();
  }
  if (t == 0) {
    out[0] = rs[0];
    out[1] = rc[0];
  }
  if (t < Q) out[t < P ? 2 + 2 * t : 3 + 2 * (t - P)] = gacc;  // interleaved re / im
}

// ------------------------------------------ second derivatives with floating lineshapes
// Per chain and event: the lineshape and its first and second derivatives in (m0, g0),
// bw[(q * n_chain + c) * N + e] for q = 0 .. 5 (B, dB/dm0, dB/dg0, d2B/dm0^2, d2B/dm0dg0,
// d2B/dg0^2; lineshape_hess).
__global__ __launch_bounds__(256) void k_lineshape_hess(const TpaModel* __restrict__ M,
                                                        const double* __restrict__ ang,
                                                        int64_t N, double2* __restrict__ bw) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N) return;
  const int nt = M->n_topo;
  const int nc = M->n_chain;
  const int64_t gbase = (int64_t)TPA_NF_TOPO * nt;
  double mf[3];
  for (int f = 0; f < 3; ++f) mf[f] = ang[(gbase + 1 + f) * N + e];
  for (int c = 0; c < nc; ++c) {
    const TpaChain& C = M->chain[c];
    const TpaTopo T = M->topo[C.topo];
    const double mR = ang[(int64_t)(TPA_NF_TOPO * C.topo + TPA_F_MR) * N + e];
    const int fy = C.r_y_first ? T.y : T.z;
    const int fz = C.r_y_first ? T.z : T.y;
    const double q = rel_p(mR, mf[fy], mf[fz]);
    cd out[6];
    lineshape_hess(C.shape, mR, C.m0, C.g0, q, C.m_y, C.m_z, C.bw_l, C.d, out);
    for (int k = 0; k < 6; ++k) bw[(int64_t)(k * nc + c) * N + e] = make_double2(out[k].r, out[k].i);
  }
}

// One pass over a fold-0 cache (waves without their lineshapes) that gives every moment the
// exact Hessian in couplings, masses and widths needs. Per event and helicity it forms the
// extended waves X_(t,k) = b^t h_k, t = 0, 1, 2 (b^0 = B, b^1 = dB/dm0, b^2 = dB/dg0 of wave
// k's chain), the amplitude A = sum_k c_k X_(0,k), and with u, v the first and second
// derivative weights of the mode (as k_hess_mfma) accumulates
//   F_j       = sum_e u z_ej,           z_ej = sum_l conj(A_el) X_j,el  (j < 3P; F_(0,k) = G_k)
//   F2_(t,k)  = sum_e u b^(3+t) q_ek,   q_ek = sum_l conj(A_el) h_ekl   (t: mm, mg, gg)
//   O1        = sum_e v y_e y_e^T,      y_e  = (Re z_e, Im z_e)            (6P columns)
//   O2        = sum_e u sum_l x x^T,    x_el = (Re X_.,el, Im X_.,el)
// as FP64 MFMA Gram products over tiles of R (event, helicity) rows staged in LDS. The
// QT(QT+1)/2 16x16 tiles of each matrix are owned by the block's four waves in turn (tile q
// by wave q % 4), which run every row of the staged tile through their own tiles. Partials per
// block: value pair, F (interleaved), F2, then O1 and O2 as raw MFMA accumulator tiles.
struct HessShapeArgs {
  const double2* __restrict__ h;     // [P][N][LG], fold-0 cache
  const double* __restrict__ w;      // [N]
  const double2* __restrict__ coef;  // [P]
  const double2* __restrict__ bw;    // [6][n_chain][N] (k_lineshape_hess)
  const int* __restrict__ kchain;    // [P]
  double* __restrict__ partials;     // [grid][nacc]
  int64_t N;
  int P, n_chain, mode, nacc;
};

#define TPA_HESS_SHAPE_PMAX 24

template <int LG, int QT>
__global__ __launch_bounds__(256) void k_hess_shape(HessShapeArgs a) {
  constexpr int R = 32;
  constexpr int E = R / LG;
  constexpr int C = 16 * QT;
  constexpr int NT = QT * (QT + 1) / 2;
  constexpr int NTW = (NT + 3) / 4;
  constexpr int ZS = R + 1, YS = E + 1;
  __shared__ double Zs[C * ZS];
  __shared__ double Ys[C * YS];
  __shared__ double2 As[R];
  __shared__ double ur[R];
  __shared__ double ve[E];
  const int P = a.P, P3 = 3 * P, Q = 6 * P;
  const int nc = a.n_chain;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int il = lane & 15, kl = lane >> 4;
  const double eps = 1.0e-6;
  for (int idx = t; idx < (C - Q) * ZS; idx += 256) Zs[Q * ZS + idx] = 0.0;
  for (int idx = t; idx < (C - Q) * YS; idx += 256) Ys[Q * YS + idx] = 0.0;
  tpa_doublex4 o1[NTW], o2[NTW];
#pragma unroll
  for (int q = 0; q < NTW; ++q) o1[q] = o2[q] = tpa_doublex4{0.0, 0.0, 0.0, 0.0};
  double gacc = 0.0;                        // thread t < 6P: component t of (Re F, Im F)
  double f2[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // thread t < P: F2_(mm|mg|gg, t) re, im
  double vsum = 0.0, vcomp = 0.0;
  const int64_t tiles = (a.N + E - 1) / E;
  for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const int64_t e0 = tile * E;
    // stage the extended waves: columns t P + k (real) and 3P + t P + k (imaginary)
    for (int idx = t; idx < P * R; idx += 256) {
      const int k = idx / R, r = idx - k * R;
      const int64_t e = e0 + r / LG;
      const int c = a.kchain[k];
      double2 hv = make_double2(0.0, 0.0);
      double2 b[3] = {hv, hv, hv};
      if (e < a.N) {
        hv = a.h[((int64_t)k * a.N + e) * LG + (r % LG)];
#pragma unroll
        for (int tt = 0; tt < 3; ++tt) b[tt] = a.bw[(int64_t)(tt * nc + c) * a.N + e];
      }
#pragma unroll
      for (int tt = 0; tt < 3; ++tt) {
        Zs[(tt * P + k) * ZS + r] = b[tt].x * hv.x - b[tt].y * hv.y;
        Zs[(P3 + tt * P + k) * ZS + r] = b[tt].x * hv.y + b[tt].y * hv.x;
      }
    }
    __syncthreads();
    if (t < R) {
      double ar = 0.0, ai = 0.0;
      for (int k = 0; k < P; ++k) {
        const double2 c = a.coef[k];
        const double xr = Zs[k * ZS + t], xi = Zs[(P3 + k) * ZS + t];
        ar += c.x * xr - c.y * xi;
        ai += c.x * xi + c.y * xr;
      }
      As[t] = make_double2(ar, ai);
    }
    __syncthreads();
    if (t < E) {
      const int64_t e = e0 + t;
      double u = 0.0, v = 0.0;
      if (e < a.N) {
        double f = 0.0;
        for (int l = 0; l < LG; ++l)
          f += As[t * LG + l].x * As[t * LG + l].x + As[t * LG + l].y * As[t * LG + l].y;
        const double w = a.w[e];
        double y;
        if (a.mode == 0) {
          double cl, d1, d2;
          if (f > eps) {
            cl = log(f);
            d1 = 1.0 / f;
            d2 = -1.0 / (f * f);
          } else {
            const double tt = (f - eps) / eps;
            cl = (log(eps) + tt) - (tt * tt) / 2.0;
            d1 = 1.0 / eps - tt / eps;
            d2 = -1.0 / (eps * eps);
          }
          y = w * cl;
          u = w * d1;
          v = w * d2;
        } else {
          y = w * f;
          u = w;
        }
        neumaier_add(vsum, vcomp, y);
      }
      for (int l = 0; l < LG; ++l) ur[t * LG + l] = u;
      ve[t] = v;
    }
    // z_ej = sum_l conj(A_el) X_j,el for the 3P extended waves (zero past the end)
    for (int idx = t; idx < E * P3; idx += 256) {
      const int e = idx / P3, j = idx - e * P3;
      double zr = 0.0, zi = 0.0;
      for (int l = 0; l < LG; ++l) {
        const double2 A = As[e * LG + l];
        const double xr = Zs[j * ZS + e * LG + l], xi = Zs[(P3 + j) * ZS + e * LG + l];
        zr += A.x * xr + A.y * xi;
        zi += A.x * xi - A.y * xr;
      }
      Ys[j * YS + e] = zr;
      Ys[(P3 + j) * YS + e] = zi;
    }
    __syncthreads();
    if (t < Q)
      for (int e = 0; e < E; ++e) gacc += ur[e * LG] * Ys[t * YS + e];
    if (t < P) {
      // F2 from q_ek = z_(0,k) / B of the wave's chain
      const int c = a.kchain[t];
      for (int e = 0; e < E; ++e) {
        const int64_t ev = e0 + e;
        if (ev >= a.N) break;
        const double u = ur[e * LG];
        const double2 B = a.bw[(int64_t)c * a.N + ev];
        const double zr = Ys[t * YS + e], zi = Ys[(P3 + t) * YS + e];
        const double den = B.x * B.x + B.y * B.y;
        const double qr = (zr * B.x + zi * B.y) / den, qi = (zi * B.x - zr * B.y) / den;
#pragma unroll
        for (int tt = 0; tt < 3; ++tt) {
          const double2 b2 = a.bw[(int64_t)((3 + tt) * nc + c) * a.N + ev];
          f2[2 * tt] += u * (b2.x * qr - b2.y * qi);
          f2[2 * tt + 1] += u * (b2.x * qi + b2.y * qr);
        }
      }
    }
    // O2 over the tile's R rows, 4 per MFMA step; tile q of the upper triangle on wave q % 4
#pragma unroll
    for (int st = 0; st < R / 4; ++st) {
      const int rl = 4 * st + kl;
      const double uw = ur[rl];
      double z[QT];
#pragma unroll
      for (int ti = 0; ti < QT; ++ti) z[ti] = Zs[(16 * ti + il) * ZS + rl];
      int q = 0;
#pragma unroll
      for (int ti = 0; ti < QT; ++ti)
#pragma unroll
        for (int tj = ti; tj < QT; ++tj, ++q)
          if (q % 4 == wave)
            o2[q / 4] = __builtin_amdgcn_mfma_f64_16x16x4f64(z[ti], uw * z[tj], o2[q / 4], 0, 0, 0);
    }
    if (a.mode == 0) {
#pragma unroll
      for (int g = 0; g < E / 4; ++g) {
        const int ev = 4 * g + kl;
        const double vw = ve[ev];
        double y[QT];
#pragma unroll
        for (int ti = 0; ti < QT; ++ti) y[ti] = Ys[(16 * ti + il) * YS + ev];
        int q = 0;
#pragma unroll
        for (int ti = 0; ti < QT; ++ti)
#pragma unroll
          for (int tj = ti; tj < QT; ++tj, ++q)
            if (q % 4 == wave)
              o1[q / 4] = __builtin_amdgcn_mfma_f64_16x16x4f64(y[ti], vw * y[tj], o1[q / 4], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  double* out = a.partials + (int64_t)blockIdx.x * a.nacc;
  const int64_t base1 = 2 + 4 * P3, base2 = base1 + (int64_t)NT * 256;
#pragma unroll
  for (int q = 0; q < NT; ++q)
    if (q % 4 == wave) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        out[base1 + (4 * q + r) * 64 + lane] = o1[q / 4][r];
        out[base2 + (4 * q + r) * 64 + lane] = o2[q / 4][r];
      }
    }
  double* rs = Zs;
  double* rc = Zs + 256;
  rs[t] = vsum;
  rc[t] = vcomp;
  __syncthreads();
  for (int m = 128; m > 0; m >>= 1) {
    if (t < m) {
      double s1 = rs[t], c1 = rc[t];
      neumaier_merge(s1, c1, rs[t + m], rc[t + m]);
      rs[t] = s1;
      rc[t] = c1;
    }
    __syncthreads();
  }
  if (t == 0) {
    out[0] = rs[0];
    out[1] = rc[0];
  }
  if (t < Q) out[t < P3 ? 2 + 2 * t : 3 + 2 * (t - P3)] = gacc;
  if (t < P) {
#pragma unroll
    for (int tt = 0; tt < 3; ++tt) {
      out[2 + 2 * P3 + 2 * (tt * P + t)] = f2[2 * tt];
      out[3 + 2 * P3 + 2 * (tt * P + t)] = f2[2 * tt + 1];
    }
  }
}

// ------------------------------------------------------------------ phase space on the device
// Counter-based random numbers: Philox4x32 with 10 rounds (J. K. Salmon et al., "Parallel
// random numbers: as easy as 1, 2, 3", SC11). Every event draws from its own counter, so a
// sample is reproducible from (seed, event index) whatever the launch shape.
struct philox_ctr {
  uint32_t x[4];
};

TPA_DEV philox_ctr philox4x32(philox_ctr c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c.x[0], p1 = (uint64_t)M1 * c.x[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    philox_ctr n;
    n.x[0] = hi1 ^ c.x[1] ^ k0;
    n.x[1] = lo1;
    n.x[2] = hi0 ^ c.x[3] ^ k1;
    n.x[3] = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// Uniform doubles in [0, 1) with 53 random bits, drawn two per Philox block.
struct uniform_stream {
  uint32_t k0, k1;
  uint64_t event;
  uint32_t block;
  double cache[2];
  int left;
  TPA_DEV double next() {
    if (left == 0) {
      philox_ctr c;
      c.x[0] = (uint32_t)event;
      c.x[1] = (uint32_t)(event >> 32);
      c.x[2] = block++;
      c.x[3] = 0x7f4a7c15u;
      c = philox4x32(c, k0, k1);
      const uint64_t a = ((uint64_t)c.x[0] << 32) | c.x[1];
      const uint64_t b = ((uint64_t)c.x[2] << 32) | c.x[3];
      cache[0] = (double)(a >> 11) * 0x1.0p-53;
      cache[1] = (double)(b >> 11) * 0x1.0p-53;
      left = 2;
    }
    return cache[--left];
  }
};

TPA_DEV double pdk(double M, double m1, double m2) {
  const double s = m1 + m2, d = m1 - m2;
  const double p2 = (M * M - s * s) * (M * M - d * d);
  return p2 > 0.0 ? sqrt(p2) / (2.0 * M) : 0.0;
}

// Back-to-back pair in the rest frame of M with an isotropic direction.
TPA_DEV void iso_pair(uniform_stream& u, double M, double ma, double mb, v4* pa, v4* pb) {
  const double c = 2.0 * u.next() - 1.0;
  const double s = sqrt(fmax(0.0, 1.0 - c * c));
  const double phi = 2.0 * M_PI * u.next();
  const double q = pdk(M, ma, mb);
  const double px = q * s * cos(phi), py = q * s * sin(phi), pz = q * c;
  *pa = v4{sqrt(q * q + ma * ma), px, py, pz};
  *pb = v4{sqrt(q * q + mb * mb), -px, -py, -pz};
}

// Boost p by the velocity of the four-vector frame (frame.t, frame.x, ...) given in the same
// system (the momentum of a rest frame seen from outside).
TPA_DEV v4 boost_by(v4 p, v4 frame) {
  const v3 b = v_make(frame.x / frame.t, frame.y / frame.t, frame.z / frame.t);
  return lv_boost(p, b);
}

// Flat phase space of m0 -> masses[0..nf) by GENBOD (F. James, CERN 68-15): sorted uniform
// intermediate masses accepted with probability prod_i p_i / wtmax, then a chain of
// back-to-back pairs boosted along. Writes p4 [f*4 + c][n]; one thread per event, each with
// its own Philox counter (event index + first), retrying until accepted.
__global__ __launch_bounds__(256) void k_phsp(int nf, double m0, const double* __restrict__ masses,
                                              double wtmax, int64_t n, int64_t first, uint32_t k0,
                                              uint32_t k1, double* __restrict__ p4) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double m[4];
  double msum = 0.0;
  for (int i = 0; i < nf; ++i) {
    m[i] = masses[i];
    msum += m[i];
  }
  const double tecm = m0 - msum;
  uniform_stream u{k0, k1, (uint64_t)(first + e), 0u, {0.0, 0.0}, 0};
  double inv[4];
  for (int attempt = 0; attempt < 4096; ++attempt) {
    double r[2] = {u.next(), u.next()};
    if (nf == 4 && r[0] > r[1]) {
      const double t = r[0];
      r[0] = r[1];
      r[1] = t;
    }
    inv[0] = m[0];
    double cs = m[0];
    for (int i = 1; i < nf - 1; ++i) {
      cs += m[i];
      inv[i] = r[i - 1] * tecm + cs;
    }
    inv[nf - 1] = m0;
    double w = 1.0;
    for (int i = 1; i < nf; ++i) w *= pdk(inv[i], inv[i - 1], m[i]);
    if (w > u.next() * wtmax) break;
  }
  v4 p[4];
  iso_pair(u, inv[1], m[0], m[1], &p[0], &p[1]);
  for (int i = 2; i < nf; ++i) {
    v4 sys, pi;
    iso_pair(u, inv[i], inv[i - 1], m[i], &sys, &pi);
    for (int j = 0; j < i; ++j) p[j] = boost_by(p[j], sys);
    p[i] = pi;
  }
  for (int f = 0; f < nf; ++f) {
    p4[(int64_t)(4 * f + 0) * n + e] = p[f].t;
    p4[(int64_t)(4 * f + 1) * n + e] = p[f].x;
    p4[(int64_t)(4 * f + 2) * n + e] = p[f].y;
    p4[(int64_t)(4 * f + 3) * n + e] = p[f].z;
  }
}

// Gather columns idx[0..n) of an angle cache (nf rows, stride S_src) into a new cache.
__global__ __launch_bounds__(256) void k_gather_cols(const double* __restrict__ src, int nf,
                                                     int64_t S_src, const int64_t* __restrict__ idx,
                                                     int64_t n, double* __restrict__ dst) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int64_t j = idx[e];
  for (int r = 0; r < nf; ++r) dst[(int64_t)r * n + e] = src[(int64_t)r * S_src + j];
}

// ------------------------------------------------------------------ FP64 MFMA Gram matrix
// G = sum_rows w_row Z_row Z_row^T in FP64 on the matrix cores (v_mfma_f64_16x16x4f64), with
// Z_row = [Re h_0 .. Re h_{P-1}, Im h_0 .. Im h_{P-1}] (Q = 2P <= 64 columns, QT = ceil(Q/16)
// tiles). A block stages 128 rows in LDS; each wave multiplies its 32 rows into the upper
// tiles (ti <= tj) of G. No refolds are needed: products and sums are FP64 throughout.
// The loads of a block's next 128 rows are issued into registers before the MFMAs of the
// current ones and written to LDS after them, so HBM latency overlaps the matrix work (without
// it 64% of the wave cycles went to waiting on the loads).

#define TPA_GRAM64_ROWS 128
#define TPA_GRAM64_STRIDE (TPA_GRAM64_ROWS + 1)
#define TPA_GRAM64_SLOTS (10 * 256)

template <int QT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(QT <= 3 ? 3 : 2))) void k_gram_f64(const double2* __restrict__ h,
                                                  const double* __restrict__ w, int64_t N,
                                                  int lam_pad, int lam, int P,
                                                  double* __restrict__ partials) {
  constexpr int NT = QT * (QT + 1) / 2;
  // staged values per thread: 128 rows x P <= 8 QT waves over 256 threads; thread t holds row
  // t % 128 of waves k = t / 128 + 2 j
  constexpr int NPF = 4 * QT;
  static_assert(TPA_GRAM64_ROWS == 128, "the staging map assumes 128 rows per step");
  __shared__ double Zs[16 * QT * TPA_GRAM64_STRIDE];
  __shared__ double Ws[TPA_GRAM64_ROWS];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int Q = 2 * P;
  const int64_t R = N * lam_pad;
  const int r_st = t & (TPA_GRAM64_ROWS - 1), k_st = t >> 7;
  tpa_doublex4 acc[NT];
#pragma unroll
  for (int q = 0; q < NT; ++q) acc[q] = tpa_doublex4{0.0, 0.0, 0.0, 0.0};
  // the padding columns Q .. 16 QT - 1 are never written by the staging: zero them once
  for (int idx = t; idx < (16 * QT - Q) * TPA_GRAM64_ROWS; idx += 256)
    Zs[(Q + idx / TPA_GRAM64_ROWS) * TPA_GRAM64_STRIDE + idx % TPA_GRAM64_ROWS] = 0.0;
  double2 pf[NPF];
  double wpf = 0.0;
  auto fetch = [&](int64_t rb) {
    const int64_t row = rb + r_st;
#pragma unroll
    for (int j = 0; j < NPF; ++j) {
      const int k = k_st + 2 * j;
      pf[j] = (k < P && row < R) ? load_pw<true>(&h[(int64_t)k * R + row])
                                 : make_double2(0.0, 0.0);
    }
    wpf = 0.0;
    if (t < TPA_GRAM64_ROWS && row < R) {
      const int64_t e = row / lam_pad;
      const int l = (int)(row - e * lam_pad);
      wpf = l < lam ? w[e] : 0.0;
    }
  };
  const int64_t step = (int64_t)gridDim.x * TPA_GRAM64_ROWS;
  int64_t rb = (int64_t)blockIdx.x * TPA_GRAM64_ROWS;
  if (rb < R) fetch(rb);
  for (; rb < R; rb += step) {
    // stage the rows fetched during the previous step (128 rows x Q columns)
#pragma unroll
    for (int j = 0; j < NPF; ++j) {
      const int k = k_st + 2 * j;
      if (k < P) {
        Zs[k * TPA_GRAM64_STRIDE + r_st] = pf[j].x;
        Zs[(P + k) * TPA_GRAM64_STRIDE + r_st] = pf[j].y;
      }
    }
    if (t < TPA_GRAM64_ROWS) Ws[t] = wpf;
    __syncthreads();
    if (rb + step < R) fetch(rb + step);  // in flight during this step's MFMAs
    const int i = lane & 15, kk = lane >> 4;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int rl = wave * 32 + 4 * s + kk;
      const double ww = Ws[rl];
      double a[QT];
#pragma unroll
      for (int ti = 0; ti < QT; ++ti) a[ti] = Zs[(16 * ti + i) * TPA_GRAM64_STRIDE + rl];
      int q = 0;
#pragma unroll
      for (int ti = 0; ti < QT; ++ti)
#pragma unroll
        for (int tj = ti; tj < QT; ++tj, ++q)
          acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ti], ww * a[tj], acc[q], 0, 0, 0);
    }
    __syncthreads();
  }
  // block sum of the four waves in a fixed order (reusing the staging area)
  double* red = Zs;  // NT * 256 doubles <= 2560 < 16 * QT * TPA_GRAM64_STRIDE for QT >= 2
  for (int wv = 0; wv < 4; ++wv) {
    if (wave == wv) {
#pragma unroll
      for (int q = 0; q < NT; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int slot = (q * 4 + v) * 64 + lane;
          red[slot] = (wv == 0) ? acc[q][v] : red[slot] + acc[q][v];
        }
    }
    __syncthreads();
  }
  double* out = partials + (int64_t)blockIdx.x * TPA_GRAM64_SLOTS;
  for (int j = t; j < NT * 256; j += blockDim.x) out[j] = red[j];
}
