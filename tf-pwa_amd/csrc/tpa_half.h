// Helicity-halved copies (FP32 or FP64) of a folded partial-wave cache, and the evaluation
// pass over them.
//
// Flipping the sign of every helicity in a partial wave conjugates its angular part: the
// Wigner functions obey D^j_{-m,-n} = (-1)^(m-n) conj(D^j_{m,n}) (d^j is real), the LS
// couplings obey the CG symmetry <j1 -m1 j2 -m2|J -M> = (-1)^(j1+j2-J) <j1 m1 j2 m2|J M>, and
// the barrier factors are real. With the lineshape product B of the wave's chain folded in,
// the folded wave h_k,o (o an outer helicity slot, -o the slot with every helicity negated)
// satisfies, per event,
//
//     h_k,-o = eta_k s_o rho conj(h_k,o),   rho = B / conj(B),  eta_k, s_o in {+1, -1},
//
// where s_o depends only on the slot and eta_k only on the wave. So half of the slots carry
// the whole cache. The copy stores, for each wave and event, the NH = Lambda/2 slots of H (one
// slot of each mirror pair) in FP32, each wave multiplied by conj(phi_k) with phi_k in {1, -i}
// chosen so that conj(phi_k)^2 eta_k is the same for all waves of a chain, and per chain and
// event the phase rho' = g_-o / (s_o conj(g_o)) (FP32), with g = conj(phi) h. Then
//
//     A_o = sum_k c~_k g_k,o,   A_-o = s_o rho' sum_k c~_k conj(g_k,o),   c~_k = phi_k c_k,
//
// and f = sum_o |A_o|^2 + |A_-o|^2. The copy kernel measures, on every event, how far the
// stored half and rho' are from reproducing the mirrored slots of the FP64 cache; the host
// refuses the copy when that residual exceeds FP64 rounding (a model whose helicity lists
// are not symmetric, say).
//
// Bytes per event at six resonances (P = 22, Lambda = 4): 22 x 2 x 8 B of waves + 6 x 8 B of
// phases + the weight = 408 B, against 712 B for the full FP32 copy. The same layout in FP64
// (S = double: 808 B against 1416 B for the FP64 cache) reproduces the FP64 cache to rounding.
#pragma once
#include "tpa_kernels.h"

#define TPA_HALF_MAX_NH 4

struct HalfCopyArgs {
  const double2* __restrict__ h;        // [P][S][lam_pad] folded FP64 cache
  void* __restrict__ g;                 // [P][n][nh] stored half, times conj(phi_k) (S pairs)
  void* __restrict__ rho;               // [n_chain][n] chain phases rho' (S pairs)
  unsigned long long* __restrict__ resid;  // largest relative residual (bits of a double >= 0)
  int64_t n, S;
  int P, lam_pad, nh, n_chain;
  int slot[TPA_HALF_MAX_NH], mirror[TPA_HALF_MAX_NH], sgn[TPA_HALF_MAX_NH];
  int ch_k0[TPA_MAX_CHAIN], ch_n[TPA_MAX_CHAIN];
  unsigned char rot[TPA_MAX_P];         // 1: phi_k = -i (the wave is stored times i)
};

// conj(phi) v: v for phi = 1, i v for phi = -i
TPA_DEV cd half_rot(double2 v, int r) {
  return r ? c_make(-v.y, v.x) : c_make(v.x, v.y);
}

// A complex value stored as a pair of S (float: rounded to nearest; double: exact).
template <typename S>
struct HalfPair;
template <>
struct HalfPair<float> {
  typedef float2 T;
  TPA_DEV static T make(const cd& v) { return make_float2((float)v.r, (float)v.i); }
};
template <>
struct HalfPair<double> {
  typedef double2 T;
  TPA_DEV static T make(const cd& v) { return make_double2(v.r, v.i); }
};

// One thread per event: every chain's reference phase from its largest stored wave, the
// stored half and phases in S, and the event's residual max |g_-o - s_o rho' conj(g_o)|
// relative to the event's largest wave (FP64, before rounding to S); one atomic max per wave
// of threads.
template <typename S>
__global__ __launch_bounds__(256) void k_pw_to_half(HalfCopyArgs a) {
  typedef HalfPair<S> HP;
  typename HP::T* const gout = static_cast<typename HP::T*>(a.g);
  typename HP::T* const rout = static_cast<typename HP::T*>(a.rho);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double worst = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < a.n; e += stride) {
    double res = 0.0, scale = 0.0;
    for (int c = 0; c < a.n_chain; ++c) {
      const int k0 = a.ch_k0[c], k1 = a.ch_k0[c] + a.ch_n[c];
      double best = 0.0;
      cd gb = c_make(0.0, 0.0), gm = c_make(0.0, 0.0);
      int sb = 1;
      for (int k = k0; k < k1; ++k) {
        const double2* hk = a.h + ((int64_t)k * a.S + e) * a.lam_pad;
        for (int o = 0; o < a.nh; ++o) {
          const double2 v = hk[a.slot[o]];
          const double m = v.x * v.x + v.y * v.y;
          if (m > best) {
            best = m;
            gb = half_rot(v, a.rot[k]);
            gm = half_rot(hk[a.mirror[o]], a.rot[k]);
            sb = a.sgn[o];
          }
        }
      }
      // rho' = g_-o / (s_o conj(g_o)) = g_-o g_o / (s_o |g_o|^2); any phase where the chain
      // vanishes (it multiplies zero)
      const cd rho = best > 0.0 ? c_scale(c_mul(gm, gb), (double)sb / best) : c_make(1.0, 0.0);
      rout[(int64_t)c * a.n + e] = HP::make(rho);
      for (int k = k0; k < k1; ++k) {
        const double2* hk = a.h + ((int64_t)k * a.S + e) * a.lam_pad;
        typename HP::T* gk = gout + ((int64_t)k * a.n + e) * a.nh;
        for (int o = 0; o < a.nh; ++o) {
          const cd x = half_rot(hk[a.slot[o]], a.rot[k]);
          const cd xm = half_rot(hk[a.mirror[o]], a.rot[k]);
          gk[o] = HP::make(x);
          const cd pred = c_scale(c_mul(rho, c_conj(x)), (double)a.sgn[o]);
          const double dr = xm.r - pred.r, di = xm.i - pred.i;
          res = fmax(res, dr * dr + di * di);
          scale = fmax(scale, fmax(x.r * x.r + x.i * x.i, xm.r * xm.r + xm.i * xm.i));
        }
      }
    }
    // a NaN anywhere makes the residual NaN, whose bits order above every finite value
    const double rel = scale > 0.0 ? sqrt(res / scale) : (res == 0.0 ? 0.0 : res);
    worst = (rel > worst || rel != rel) ? rel : worst;
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const double o = __shfl_xor(worst, m, 64);
    worst = (o > worst || o != o) ? o : worst;
  }
  if ((threadIdx.x & 63) == 0)
    atomicMax(a.resid, (unsigned long long)__double_as_longlong(fabs(worst)) |
                           (worst != worst ? 0x7ff8000000000000ull : 0ull));
}

// The NH stored slots of one wave and event: FP32 pairs in one 8- or 16-byte load, FP64 pairs
// in one or two 16-byte loads; get() widens slot o to FP64.
template <int NH, typename S>
struct HalfVec;
template <>
struct HalfVec<1, float> {
  typedef float2 T;
  TPA_DEV static T load(const T* p) {
    typedef float f2v __attribute__((ext_vector_type(2)));
    const f2v v = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p));
    return make_float2(v.x, v.y);
  }
  TPA_DEV static void get(const T& v, int o, double& x, double& y) {
    (void)o;
    x = v.x;
    y = v.y;
  }
};
template <>
struct HalfVec<2, float> {
  typedef float4 T;
  TPA_DEV static T load(const T* p) {
    typedef float f4v __attribute__((ext_vector_type(4)));
    const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
  }
  TPA_DEV static void get(const T& v, int o, double& x, double& y) {
    x = o == 0 ? v.x : v.z;
    y = o == 0 ? v.y : v.w;
  }
};
typedef double tpa_d2v __attribute__((ext_vector_type(2)));
template <>
struct HalfVec<1, double> {
  typedef double2 T;
  TPA_DEV static T load(const T* p) {
    const tpa_d2v v = __builtin_nontemporal_load(reinterpret_cast<const tpa_d2v*>(p));
    return make_double2(v.x, v.y);
  }
  TPA_DEV static void get(const T& v, int o, double& x, double& y) {
    (void)o;
    x = v.x;
    y = v.y;
  }
};
struct HalfD2x2 {
  double2 a, b;
};
template <>
struct HalfVec<2, double> {
  typedef HalfD2x2 T;
  TPA_DEV static T load(const T* p) {
    const tpa_d2v* q = reinterpret_cast<const tpa_d2v*>(p);
    const tpa_d2v v0 = __builtin_nontemporal_load(q), v1 = __builtin_nontemporal_load(q + 1);
    return HalfD2x2{make_double2(v0.x, v0.y), make_double2(v1.x, v1.y)};
  }
  TPA_DEV static void get(const T& v, int o, double& x, double& y) {
    x = o == 0 ? v.a.x : v.b.x;
    y = o == 0 ? v.a.y : v.b.y;
  }
};

// Empty asm that makes the FP32 registers opaque to common-subexpression elimination, so the
// adjoint widens them again rather than keeping the first loop's FP64 values live (FP64
// storage needs no widening: nothing to do).
TPA_DEV void half_opaque(float2& v) { asm("" : "+v"(v.x), "+v"(v.y)); }
TPA_DEV void half_opaque(float4& v) { asm("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w)); }
TPA_DEV void half_opaque(double2&) {}
TPA_DEV void half_opaque(HalfD2x2&) {}

// One event group of the halved pass. A lane holds PML waves of one event (k = kg + KG j) and
// the chain phase of each; A_o and B_o = A_-o / s_o are summed over the KG lanes of the event.
// The adjoint per wave: G~_k += u sum_o [conj(A_o) g_k,o + conj(B_o) rho' conj(g_k,o)]
//                            = sum_o conj(u A_o) g_k,o + rho' conj(sum_o u B_o g_k,o).
template <int NH, int KG, int PML, typename S>
TPA_DEV void eval_half_group(int mode, bool grad, double* __restrict__ f_out,
                             const typename HalfVec<NH, S>::T (&hs)[PML],
                             const typename HalfPair<S>::T (&rs)[PML],
                             double w, int64_t e, bool valid, bool leader,
                             const double (&cr)[PML], const double (&ci)[PML], double& vs,
                             double& vc, double (&gr)[PML], double (&gi)[PML]) {
  typedef HalfVec<NH, S> V;
  const double eps = 1.0e-6;
  double ar[NH], ai[NH], br[NH], bi[NH];
#pragma unroll
  for (int o = 0; o < NH; ++o) ar[o] = ai[o] = br[o] = bi[o] = 0.0;
#pragma unroll
  for (int j = 0; j < PML; ++j) {
    const double rr = rs[j].x, ri = rs[j].y;
    // t = c~ rho'
    const double tr = fma(cr[j], rr, -(ci[j] * ri));
    const double ti = fma(cr[j], ri, ci[j] * rr);
#pragma unroll
    for (int o = 0; o < NH; ++o) {
      double x, y;
      V::get(hs[j], o, x, y);
      ar[o] = fma(cr[j], x, ar[o]);
      ar[o] = fma(-ci[j], y, ar[o]);
      ai[o] = fma(cr[j], y, ai[o]);
      ai[o] = fma(ci[j], x, ai[o]);
      // B_o += t conj(g)
      br[o] = fma(tr, x, br[o]);
      br[o] = fma(ti, y, br[o]);
      bi[o] = fma(ti, x, bi[o]);
      bi[o] = fma(-tr, y, bi[o]);
    }
  }
  double f2 = 0.0;
#pragma unroll
  for (int o = 0; o < NH; ++o) {
    ar[o] = butterfly<1, KG>(ar[o]);
    ai[o] = butterfly<1, KG>(ai[o]);
    br[o] = butterfly<1, KG>(br[o]);
    bi[o] = butterfly<1, KG>(bi[o]);
    f2 += (ar[o] * ar[o] + ai[o] * ai[o]) + (br[o] * br[o] + bi[o] * bi[o]);
  }
  const double f = f2;
  const double wv = valid ? w : 0.0;
  double y, u;
  if (mode == 0) {
    // clip_log(f) (tf_pwa/model/model.py:283-292) weighted by w
    double cl, dcl;
    if (f > eps) {
      dcl = tpa_recip_pos(f, hw_rcp);
      cl = tpa_log_pos(f, hw_rcp);
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
    double uar[NH], uai[NH], ubr[NH], ubi[NH];
#pragma unroll
    for (int o = 0; o < NH; ++o) {
      uar[o] = u * ar[o];
      uai[o] = u * ai[o];
      ubr[o] = u * br[o];
      ubi[o] = u * bi[o];
    }
#pragma unroll
    for (int j = 0; j < PML; ++j) {
      typename V::T q = hs[j];
      half_opaque(q);
      typename HalfPair<S>::T rq = rs[j];
      half_opaque(rq);
      const double rr = rq.x, ri = rq.y;
      double qr = 0.0, qi = 0.0;
#pragma unroll
      for (int o = 0; o < NH; ++o) {
        double x, yy;
        V::get(q, o, x, yy);
        // conj(u A_o) g
        gr[j] = fma(uar[o], x, gr[j]);
        gr[j] = fma(uai[o], yy, gr[j]);
        gi[j] = fma(uar[o], yy, gi[j]);
        gi[j] = fma(-uai[o], x, gi[j]);
        // Q += u B_o g
        qr = fma(ubr[o], x, qr);
        qr = fma(-ubi[o], yy, qr);
        qi = fma(ubr[o], yy, qi);
        qi = fma(ubi[o], x, qi);
      }
      // rho' conj(Q)
      gr[j] = fma(rr, qr, gr[j]);
      gr[j] = fma(ri, qi, gr[j]);
      gi[j] = fma(ri, qr, gi[j]);
      gi[j] = fma(-rr, qi, gi[j]);
    }
  }
}

// The evaluation pass over a helicity-halved copy: FP32 (S = float; a.h32: [P][N][NH] float2,
// a.rho: [chain][N] float2) or FP64 (S = double; a.h: [P][N][NH] double2, a.rho64); a.kch[k]:
// the chain of wave k; coefficients c~ in a.cin. Lane layout: lane = ev * KG + kg, the wave
// k = kg + KG j in slot j < PML; the NH stored slots of a wave come in one 8- or 16-byte load
// (FP32) or one or two 16-byte loads (FP64). Software-pipelined as k_eval_pipe32:
// two register stages, loads of the next event group issued before the current one is used,
// unconditional loads at clamped addresses past the last event or wave. Per-block partials
// slot-major as the other passes (k_reduce). Requires N >= 1.
template <int NH, int KG, int PML, typename S>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_eval_half(EvalArgs a) {
  typedef HalfVec<NH, S> V;
  typedef typename V::T HV;
  typedef typename HalfPair<S>::T RV;
  const HV* const hdata = sizeof(S) == 4 ? reinterpret_cast<const HV*>(a.h32)
                                         : reinterpret_cast<const HV*>(a.h);
  const RV* const rdata = sizeof(S) == 4 ? reinterpret_cast<const RV*>(a.rho)
                                         : reinterpret_cast<const RV*>(a.rho64);
  constexpr int EPW = 64 / KG;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int kg = lane & (KG - 1);
  const int el = lane / KG;
  const bool leader = kg == 0;
  double vs = 0.0, vc = 0.0;
  double gr[PML], gi[PML], cr[PML], ci[PML];
  const HV* base[PML];
  const RV* rbase[PML];
#pragma unroll
  for (int j = 0; j < PML; ++j) {
    const int k = kg + KG * j;
    gr[j] = gi[j] = 0.0;
    const double2 c = k < a.P ? eval_coef(a, k) : make_double2(0.0, 0.0);
    cr[j] = c.x;
    ci[j] = c.y;
    const int kc = k < a.P ? k : a.P - 1;
    base[j] = hdata + (int64_t)kc * a.N;
    rbase[j] = rdata + (int64_t)a.kch[kc] * a.N;
  }
  const int64_t last = a.N - 1;
  auto load = [&](HV (&hv)[PML], RV (&rv)[PML], double& wv, int64_t e) {
    const int64_t ec = e < a.N ? e : last;
#pragma unroll
    for (int j = 0; j < PML; ++j) {
      hv[j] = V::load(base[j] + ec);
      rv[j] = rbase[j][ec];
    }
    wv = a.w[ec];
  };
  const int64_t stride = (int64_t)gridDim.x * 4 * EPW;
  HV h0[PML], h1[PML];
  RV r0[PML], r1[PML];
  double w0, w1;
  const int wave_u = __builtin_amdgcn_readfirstlane(wave);
  const int64_t first = ((int64_t)blockIdx.x * 4 + wave_u) * EPW;
  int64_t e = first + el;
  load(h0, r0, w0, e);
  for (int64_t b = first; b < a.N; b += 2 * stride, e += 2 * stride) {
    load(h1, r1, w1, e + stride);
    __builtin_amdgcn_sched_barrier(0);
    eval_half_group<NH, KG, PML, S>(a.mode, a.grad, a.f_out, h0, r0, w0, e, e < a.N, leader, cr,
                                 ci, vs, vc, gr, gi);
    if (b + stride >= a.N) break;
    load(h0, r0, w0, e + 2 * stride);
    __builtin_amdgcn_sched_barrier(0);
    eval_half_group<NH, KG, PML, S>(a.mode, a.grad, a.f_out, h1, r1, w1, e + stride,
                                 e + stride < a.N, leader, cr, ci, vs, vc, gr, gi);
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
      const double xr = butterfly<KG, 64>(gr[j]), xi = butterfly<KG, 64>(gi[j]);
      const int k = kg + KG * j;
      if (el == 0 && k < a.P) {
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
    for (int j = 2 + threadIdx.x; j < a.nacc; j += blockDim.x)
      out[(int64_t)j * gridDim.x] = ((sh[0][j] + sh[1][j]) + sh[2][j]) + sh[3][j];
  }
}

// Raw buffer resource over the first `bytes` bytes at p (gfx9 descriptor word 3): a load at
// an offset at or beyond `bytes` returns zero instead of touching memory.
TPA_DEV __amdgpu_buffer_rsrc_t half_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes,
                                           0x00020000);
}

// One event group of k_eval_half_chain: the lane's PML waves come in two runs of one chain
// each (slots j < J0 of chain-slot 0, the rest of chain-slot 1), so the chain phase rho' of
// a wave is that of its run: B_o = rho'_0 S_0,o + rho'_1 S_1,o with S_s,o = sum over the run
// of c~_k conj(g_k,o), and the adjoint's phase term folds into W_s,o = conj(rho'_s) u B_o:
// G~_k += sum_o conj(u A_o) g_k,o + conj(W_s,o g_k,o). The same sums as eval_half_group
// (A, B and f in FP64 from the widened FP32 waves, the adjoint in FP64), with the phase
// multiplied per run instead of per wave; the adjoint reuses the waves widened for the
// forward sums (206 VGPRs, two waves per SIMD: 2-3% faster than widening them again at 166
// VGPRs and three waves per SIMD). Returns f: the mode's value is summed by the caller, four
// groups at a time.
template <int KG, int PML, int J0>
TPA_DEV double eval_half_chain_group(int mode, bool grad, double* __restrict__ f_out,
                                     const float4 (&hs)[PML], const float2 (&rs)[2], double w,
                                     int64_t e, bool valid, bool leader, const double (&cr)[PML],
                                     const double (&ci)[PML], double (&gr)[PML],
                                     double (&gi)[PML]) {
  typedef HalfVec<2, float> V;
  const double eps = 1.0e-6;
  double ar[2], ai[2], sr[2][2], si[2][2];
#pragma unroll
  for (int o = 0; o < 2; ++o) {
    ar[o] = ai[o] = 0.0;
    sr[0][o] = si[0][o] = sr[1][o] = si[1][o] = 0.0;
  }
#pragma unroll
  for (int j = 0; j < PML; ++j) {
    const int s = j < J0 ? 0 : 1;
#pragma unroll
    for (int o = 0; o < 2; ++o) {
      double x, y;
      V::get(hs[j], o, x, y);
      ar[o] = fma(cr[j], x, ar[o]);
      ar[o] = fma(-ci[j], y, ar[o]);
      ai[o] = fma(cr[j], y, ai[o]);
      ai[o] = fma(ci[j], x, ai[o]);
      // S += c~ conj(g)
      sr[s][o] = fma(cr[j], x, sr[s][o]);
      sr[s][o] = fma(ci[j], y, sr[s][o]);
      si[s][o] = fma(ci[j], x, si[s][o]);
      si[s][o] = fma(-cr[j], y, si[s][o]);
    }
  }
  const double p0r = rs[0].x, p0i = rs[0].y, p1r = rs[1].x, p1i = rs[1].y;
  double br[2], bi[2];
  double f = 0.0;
#pragma unroll
  for (int o = 0; o < 2; ++o) {
    // B_o = rho'_0 S_0 + rho'_1 S_1 (this lane's part)
    br[o] = fma(p0r, sr[0][o], fma(-p0i, si[0][o], fma(p1r, sr[1][o], -(p1i * si[1][o]))));
    bi[o] = fma(p0r, si[0][o], fma(p0i, sr[0][o], fma(p1r, si[1][o], p1i * sr[1][o])));
    ar[o] = butterfly<1, KG>(ar[o]);
    ai[o] = butterfly<1, KG>(ai[o]);
    br[o] = butterfly<1, KG>(br[o]);
    bi[o] = butterfly<1, KG>(bi[o]);
    f += (ar[o] * ar[o] + ai[o] * ai[o]) + (br[o] * br[o] + bi[o] * bi[o]);
  }
  if (f_out != nullptr && valid && leader) f_out[e] = f;
  if (grad) {
    const double wv = valid ? w : 0.0;
    double u;
    if (mode == 0) {
      // clip_log'(f) (tf_pwa/model/model.py:283-292)
      if (f > eps) {
        u = wv * tpa_recip_pos(f, hw_rcp);
      } else {
        const double t = (f - eps) / eps;
        u = wv * (1.0 / eps - t / eps);
      }
    } else {
      u = wv;
    }
    // W_s,o = conj(rho'_s) u B_o
    double uar[2], uai[2], wr[2][2], wi[2][2];
#pragma unroll
    for (int o = 0; o < 2; ++o) {
      uar[o] = u * ar[o];
      uai[o] = u * ai[o];
      const double ubr = u * br[o], ubi = u * bi[o];
      wr[0][o] = fma(p0r, ubr, p0i * ubi);
      wi[0][o] = fma(p0r, ubi, -(p0i * ubr));
      wr[1][o] = fma(p1r, ubr, p1i * ubi);
      wi[1][o] = fma(p1r, ubi, -(p1i * ubr));
    }
#pragma unroll
    for (int j = 0; j < PML; ++j) {
      const int s = j < J0 ? 0 : 1;
      float4 q = hs[j];
#pragma unroll
      for (int o = 0; o < 2; ++o) {
        double x, yy;
        V::get(q, o, x, yy);
        // conj(u A_o) g
        gr[j] = fma(uar[o], x, gr[j]);
        gr[j] = fma(uai[o], yy, gr[j]);
        gi[j] = fma(uar[o], yy, gi[j]);
        gi[j] = fma(-uai[o], x, gi[j]);
        // conj(W g)
        gr[j] = fma(wr[s][o], x, gr[j]);
        gr[j] = fma(-wi[s][o], yy, gr[j]);
        gi[j] = fma(-wr[s][o], yy, gi[j]);
        gi[j] = fma(-wi[s][o], x, gi[j]);
      }
    }
  }
  return f;
}

// The helicity-halved FP32 pass (NH = 2 stored slots, KG = 4 lanes per event, PML waves per
// lane) with every lane's waves in two runs of one chain each (a lane plan the host makes with
// the copy, tpa_events_keep_half): the chain phase rho' is loaded and multiplied once per run
// instead of once per wave (two phase loads per lane and event group instead of PML; per wave
// the t = c~ rho' product and the adjoint's rho' conj(Q) are gone). The FP64 sums of
// k_eval_half<2, 4, PML, float>, with fewer vector instructions per event group:
// - the mode's value (clip_log with its logarithm, or w f) and its compensated sum run once
//   per four event groups: the four lanes of an event each take the stored f and w of one of
//   the four groups, instead of all of them evaluating the same logarithm every group (the
//   derivative, a reciprocal, stays per group);
// - every load is a buffer load whose lane offset is fixed for the pass (the lane's row of the
//   copy per slot, of the chain phases per run) and whose event-group offset sits in a scalar
//   register: no address arithmetic per load; a load at or past the end of the copy, the chain
//   phases or the weights returns zero, so pad slots need no clamp and events past the last
//   one read weight zero (their waves come from the next row, finite, and add exactly zero);
// - the coefficients sit in LDS and are read at each event group, so they hold no registers
//   between groups.
// The host checks that every buffer offset stays below 2^31 bytes. Partial sums per block,
// slot-major, as the other passes (k_reduce).
template <int KG, int PML, int J0>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_eval_half_chain(
    EvalArgs a, uint32_t h_bytes, uint32_t rho_bytes) {
  typedef float4 HV;
  typedef float2 RV;
  static_assert(J0 > 0 && J0 < PML, "two runs per lane");
  static_assert(KG == 4, "the value is shared out over four lanes per event");
  constexpr int EPW = 64 / KG;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int kg = lane & (KG - 1);
  const int el = lane / KG;
  const bool leader = kg == 0;
  __shared__ double2 csh[KG * PML];
  __shared__ double sh[4][TPA_NACC_MAX];
  // lane kg's slot j holds wave plan[kg * PML + j] (negative: a pad slot)
  if (threadIdx.x < KG * PML) {
    const int k = a.plan[threadIdx.x];
    csh[threadIdx.x] = k >= 0 ? eval_coef(a, k) : make_double2(0.0, 0.0);
  }
  __syncthreads();
  const __amdgpu_buffer_rsrc_t hr = half_rsrc(a.h32, h_bytes);
  const __amdgpu_buffer_rsrc_t rr = half_rsrc(a.rho, rho_bytes);
  const uint32_t N = (uint32_t)a.N;
  const __amdgpu_buffer_rsrc_t wr = half_rsrc(a.w, N * 8u);
  // per slot the lane's row of the copy (a pad slot reads past its end: zeros), per run the
  // lane's row of chain phases
  uint32_t vh[PML];
#pragma unroll
  for (int j = 0; j < PML; ++j) {
    const int k = a.plan[kg * PML + j];
    vh[j] = k >= 0 ? ((uint32_t)k * N + (uint32_t)el) * 16u : h_bytes;
  }
  uint32_t vr[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int c = a.pch[kg * 2 + q];
    vr[q] = c >= 0 ? ((uint32_t)c * N + (uint32_t)el) * 8u : rho_bytes;
  }
  const uint32_t vw = (uint32_t)el * 8u;
  // the argument fields the loop needs, as locals (a lambda that reaches into the argument
  // struct makes the compiler copy all of it into scratch memory)
  const int mode = a.mode;
  const bool grad = a.grad != 0;
  double* const f_out = a.f_out;
  const int64_t n_ev = a.N;
  double vs = 0.0, vc = 0.0;
  double gr[PML], gi[PML];
#pragma unroll
  for (int j = 0; j < PML; ++j) gr[j] = gi[j] = 0.0;
  typedef unsigned int u4v __attribute__((ext_vector_type(4)));
  typedef unsigned int u2v __attribute__((ext_vector_type(2)));
  // a group that starts past the last event (the loads issued ahead of the loop's end) reads
  // at offsets past every buffer: zeros, and no memory traffic
  const uint32_t w_bytes = N * 8u;
  auto load = [&](HV (&hv)[PML], RV (&rv)[2], double& wv, uint32_t gb)
                  __attribute__((always_inline)) {
    const bool past = gb >= N;
    const int so_h = past ? (int)h_bytes : (int)(gb * 16u);
    const int so_r = past ? (int)rho_bytes : (int)(gb * 8u);
    const int so_w = past ? (int)w_bytes : (int)(gb * 8u);
#pragma unroll
    for (int j = 0; j < PML; ++j) {
      const u4v x = __builtin_amdgcn_raw_buffer_load_b128(hr, vh[j], so_h, 2);
      hv[j] = make_float4(__uint_as_float(x.x), __uint_as_float(x.y), __uint_as_float(x.z),
                          __uint_as_float(x.w));
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const u2v r = __builtin_amdgcn_raw_buffer_load_b64(rr, vr[q], so_r, 0);
      rv[q] = make_float2(__uint_as_float(r.x), __uint_as_float(r.y));
    }
    const u2v x = __builtin_amdgcn_raw_buffer_load_b64(wr, vw, so_w, 0);
    wv = __hiloint2double((int)x.y, (int)x.x);
  };
  auto group = [&](const HV (&hv)[PML], const RV (&rv)[2], double wv, uint32_t gb)
                   __attribute__((always_inline)) {
    // the coefficients from LDS at every group (an opaque index keeps them out of registers
    // between groups)
    int cb = kg * PML;
    asm volatile("" : "+v"(cb));
    double cr[PML], ci[PML];
#pragma unroll
    for (int j = 0; j < PML; ++j) {
      const double2 c = csh[cb + j];
      cr[j] = c.x;
      ci[j] = c.y;
    }
    const int64_t e = (int64_t)gb + el;
    return eval_half_chain_group<KG, PML, J0>(mode, grad, f_out, hv, rv, wv, e, e < n_ev,
                                              leader, cr, ci, gr, gi);
  };
  // the mode's value of up to four stored groups (f_q, w_q), q < nq: lane kg takes group kg
  // (its weight is zero past the last event); then the FP32 adjoint sums go into the FP64 ones
  auto values = [&](double f0, double w0_, double f1, double w1_, double f2, double w2_, double f3,
                    double w3_, int nq) __attribute__((always_inline)) {
    const bool k1 = kg & 1, k2 = kg & 2;
    const double f = k2 ? (k1 ? f3 : f2) : (k1 ? f1 : f0);
    const double w = k2 ? (k1 ? w3_ : w2_) : (k1 ? w1_ : w0_);
    double y;
    if (mode == 0) {
      // clip_log(f) (tf_pwa/model/model.py:283-292) weighted by w
      const double eps = 1.0e-6;
      double cl;
      if (f > eps) {
        cl = tpa_log_pos(f, hw_rcp);
      } else {
        const double t = (f - eps) / eps;
        cl = (log(eps) + t) - (t * t) / 2.0;
      }
      y = w * cl;
    } else {
      y = w * f;
    }
    neumaier_add(vs, vc, kg < nq ? y : 0.0);
  };
  const uint32_t stride = gridDim.x * 4u * EPW;
  HV h0[PML], h1[PML];
  RV r0[2], r1[2];
  double w0, w1;
  const uint32_t wave_u = __builtin_amdgcn_readfirstlane(wave);
  const uint32_t first = (blockIdx.x * 4u + wave_u) * EPW;
  // two event groups per iteration; the first pair of a four-group cycle waits in (fp0, wp0),
  // (fp1, wp1) until the second pair is done
  double fp0 = 0.0, wp0 = 0.0, fp1 = 0.0, wp1 = 0.0;
  int cyc = 0;  // uniform: 1 while a first pair waits
  load(h0, r0, w0, first);
  uint32_t b = first;
  for (; b < N; b += 2 * stride) {
    load(h1, r1, w1, b + stride);
    __builtin_amdgcn_sched_barrier(0);
    const double fa = group(h0, r0, w0, b);
    const double wa = w0;
    if (b + stride >= N) {
      if (cyc)
        values(fp0, wp0, fp1, wp1, fa, wa, 0.0, 0.0, 3);
      else
        values(fa, wa, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1);
      cyc = 0;
      break;
    }
    load(h0, r0, w0, b + 2 * stride);
    __builtin_amdgcn_sched_barrier(0);
    const double fb = group(h1, r1, w1, b + stride);
    const double wb = w1;
    if (cyc) {
      values(fp0, wp0, fp1, wp1, fa, wa, fb, wb, 4);
    } else {
      fp0 = fa;
      wp0 = wa;
      fp1 = fb;
      wp1 = wb;
    }
    cyc ^= 1;
  }
  if (cyc) values(fp0, wp0, fp1, wp1, 0.0, 0.0, 0.0, 0.0, 2);
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const double s2 = __shfl_xor(vs, m, 64);
    const double c2 = __shfl_xor(vc, m, 64);
    neumaier_merge(vs, vc, s2, c2);
  }
  if (a.grad) {
#pragma unroll
    for (int j = 0; j < PML; ++j) {
      const double xr = butterfly<KG, 64>(gr[j]), xi = butterfly<KG, 64>(gi[j]);
      const int k = a.plan[kg * PML + j];
      if (el == 0 && k >= 0) {
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
    for (int j = 2 + threadIdx.x; j < a.nacc; j += blockDim.x)
      out[(int64_t)j * gridDim.x] = ((sh[0][j] + sh[1][j]) + sh[2][j]) + sh[3][j];
  }
}

// One event group of k_eval_pipe32d (the full FP32 copy, two helicity components per lane,
// KG lanes per event): A_lambda of both components and f in FP64 as eval_group2, the mode's
// derivative u, and the adjoint on the FP64 waves widened for the forward sums. Returns f:
// the mode's value is summed by the caller, once per KG groups.
template <int KG, int PML>
TPA_DEV double eval_pipe32d_group(int mode, bool grad, double* __restrict__ f_out,
                                  const float4 (&hs)[PML], double w, int64_t e, bool valid,
                                  bool leader, const double (&cr)[PML], const double (&ci)[PML],
                                  double (&gr)[PML], double (&gi)[PML]) {
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
  ar0 = butterfly<1, KG>(ar0);
  ai0 = butterfly<1, KG>(ai0);
  ar1 = butterfly<1, KG>(ar1);
  ai1 = butterfly<1, KG>(ai1);
  const double f = (ar0 * ar0 + ai0 * ai0) + (ar1 * ar1 + ai1 * ai1);
  if (f_out != nullptr && valid && leader) f_out[e] = f;
  if (grad) {
    const double wv = valid ? w : 0.0;
    double u;
    if (mode == 0) {
      // clip_log'(f) (tf_pwa/model/model.py:283-292)
      if (f > eps) {
        u = wv * tpa_recip_pos(f, hw_rcp);
      } else {
        const double t = (f - eps) / eps;
        u = wv * (1.0 / eps - t / eps);
      }
    } else {
      u = wv;
    }
    const double tr0 = u * ar0, ti0 = -(u * ai0), tr1 = u * ar1, ti1 = -(u * ai1);
#pragma unroll
    for (int j = 0; j < PML; ++j) {
      const double x0 = hs[j].x, y0 = hs[j].y, x1 = hs[j].z, y1 = hs[j].w;
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
  return f;
}

// The evaluation pass over the full FP32 copy (a.h32, two helicity slots: the four-body
// models' fastest data pass) with the scheme of k_eval_half_chain: KG = 4 lanes per event,
// wave k = kg + KG j in slot j < PML; the mode's value and its compensated sum once per four
// event groups (lane kg takes group kg); buffer loads with a fixed lane offset and scalar
// group and wave offsets, zeros at and past the end of the copy and the weights (pad waves,
// groups past the last event); coefficients from LDS at every group; the adjoint on the
// forward's widened waves. The same FP64 sums as k_eval_pipe32<2, KG, PML>. The host checks
// that every buffer offset stays below 2^31 bytes.
template <int KG, int PML>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_eval_pipe32d(
    EvalArgs a, uint32_t h_bytes) {
  typedef float4 HV;
  static_assert(KG == 4, "the value is shared out over four lanes per event");
  constexpr int EPW = 64 / KG;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int kg = lane & (KG - 1);
  const int el = lane / KG;
  const bool leader = kg == 0;
  __shared__ double2 csh[KG * PML];
  __shared__ double sh[4][TPA_NACC_MAX];
  if (threadIdx.x < KG * PML)
    csh[threadIdx.x] = (int)threadIdx.x < a.P ? eval_coef(a, threadIdx.x) : make_double2(0.0, 0.0);
  __syncthreads();
  const uint32_t N = (uint32_t)a.N;
  const __amdgpu_buffer_rsrc_t hr = half_rsrc(a.h32, h_bytes);
  const __amdgpu_buffer_rsrc_t wr = half_rsrc(a.w, N * 8u);
  // row k of the copy: [k][event][2 slots] float2, 16 bytes per event
  const uint32_t vh = ((uint32_t)kg * N + (uint32_t)el) * 16u;
  const uint32_t vw = (uint32_t)el * 8u;
  const int mode = a.mode;
  const bool grad = a.grad != 0;
  double* const f_out = a.f_out;
  const int64_t n_ev = a.N;
  const uint32_t w_bytes = N * 8u;
  double vs = 0.0, vc = 0.0;
  double gr[PML], gi[PML];
#pragma unroll
  for (int j = 0; j < PML; ++j) gr[j] = gi[j] = 0.0;
  typedef unsigned int u4v __attribute__((ext_vector_type(4)));
  typedef unsigned int u2v __attribute__((ext_vector_type(2)));
  auto load = [&](HV (&hv)[PML], double& wv, uint32_t gb) __attribute__((always_inline)) {
    const bool past = gb >= N;
#pragma unroll
    for (int j = 0; j < PML; ++j) {
      const int so = past ? (int)h_bytes : (int)(((uint32_t)(KG * j) * N + gb) * 16u);
      const u4v x = __builtin_amdgcn_raw_buffer_load_b128(hr, vh, so, 2);
      hv[j] = make_float4(__uint_as_float(x.x), __uint_as_float(x.y), __uint_as_float(x.z),
                          __uint_as_float(x.w));
    }
    const u2v x = __builtin_amdgcn_raw_buffer_load_b64(wr, vw, past ? (int)w_bytes : (int)(gb * 8u), 0);
    wv = __hiloint2double((int)x.y, (int)x.x);
  };
  auto group = [&](const HV (&hv)[PML], double wv, uint32_t gb) __attribute__((always_inline)) {
    int cb = kg;
    asm volatile("" : "+v"(cb));
    double cr[PML], ci[PML];
#pragma unroll
    for (int j = 0; j < PML; ++j) {
      const double2 c = csh[cb + KG * j];
      cr[j] = c.x;
      ci[j] = c.y;
    }
    const int64_t e = (int64_t)gb + el;
    return eval_pipe32d_group<KG, PML>(mode, grad, f_out, hv, wv, e, e < n_ev, leader, cr, ci,
                                       gr, gi);
  };
  auto values = [&](double f0, double w0_, double f1, double w1_, double f2, double w2_, double f3,
                    double w3_, int nq) __attribute__((always_inline)) {
    const bool k1 = kg & 1, k2 = kg & 2;
    const double f = k2 ? (k1 ? f3 : f2) : (k1 ? f1 : f0);
    const double w = k2 ? (k1 ? w3_ : w2_) : (k1 ? w1_ : w0_);
    double y;
    if (mode == 0) {
      // clip_log(f) (tf_pwa/model/model.py:283-292) weighted by w
      const double eps = 1.0e-6;
      double cl;
      if (f > eps) {
        cl = tpa_log_pos(f, hw_rcp);
      } else {
        const double t = (f - eps) / eps;
        cl = (log(eps) + t) - (t * t) / 2.0;
      }
      y = w * cl;
    } else {
      y = w * f;
    }
    neumaier_add(vs, vc, kg < nq ? y : 0.0);
  };
  const uint32_t stride = gridDim.x * 4u * EPW;
  HV h0[PML], h1[PML];
  double w0, w1;
  const uint32_t wave_u = __builtin_amdgcn_readfirstlane(wave);
  const uint32_t first = (blockIdx.x * 4u + wave_u) * EPW;
  double fp0 = 0.0, wp0 = 0.0, fp1 = 0.0, wp1 = 0.0;
  int cyc = 0;  // uniform: 1 while a first pair of groups waits for its value
  load(h0, w0, first);
  for (uint32_t b = first; b < N; b += 2 * stride) {
    load(h1, w1, b + stride);
    __builtin_amdgcn_sched_barrier(0);
    const double fa = group(h0, w0, b);
    const double wa = w0;
    if (b + stride >= N) {
      if (cyc)
        values(fp0, wp0, fp1, wp1, fa, wa, 0.0, 0.0, 3);
      else
        values(fa, wa, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1);
      cyc = 0;
      break;
    }
    load(h0, w0, b + 2 * stride);
    __builtin_amdgcn_sched_barrier(0);
    const double fb = group(h1, w1, b + stride);
    const double wb = w1;
    if (cyc) {
      values(fp0, wp0, fp1, wp1, fa, wa, fb, wb, 4);
    } else {
      fp0 = fa;
      wp0 = wa;
      fp1 = fb;
      wp1 = wb;
    }
    cyc ^= 1;
  }
  if (cyc) values(fp0, wp0, fp1, wp1, 0.0, 0.0, 0.0, 0.0, 2);
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const double s2 = __shfl_xor(vs, m, 64);
    const double c2 = __shfl_xor(vc, m, 64);
    neumaier_merge(vs, vc, s2, c2);
  }
  if (grad) {
#pragma unroll
    for (int j = 0; j < PML; ++j) {
      const double xr = butterfly<KG, 64>(gr[j]), xi = butterfly<KG, 64>(gi[j]);
      const int k = kg + KG * j;
      if (el == 0 && k < a.P) {
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
  if (grad) {
    for (int j = 2 + threadIdx.x; j < a.nacc; j += blockDim.x)
      out[(int64_t)j * gridDim.x] = ((sh[0][j] + sh[1][j]) + sh[2][j]) + sh[3][j];
  }
}
