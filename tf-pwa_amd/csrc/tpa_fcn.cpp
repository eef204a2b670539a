// Single-fit likelihood session: FCN.nll_grad (tf_pwa/model/model.py:1207-1240) for fixed
// lineshapes as one native call. The trainable values y go through the coupling map to the
// partial-wave coefficients, each data set and (FP64 path) each MC set is one tpa_eval pass, or
// (mixed path) the MC normalisation is the Gram form c^dagger I c, and the gradient in y follows
// by the coupling chain rule. The same arithmetic as Model._nll_grad of the Python host, without
// its per-call interpreter work.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "tfpwa_amd.h"
#include "tpa_bfgs.h"
#include "tpa_coupling.h"
#include "tpa_internal.h"
#include "tpa_model.h"

using namespace tpa_cpl;

struct tpa_fcn_s : CouplingMap {
  std::vector<tpa_events_t> data, mc;
  double sw = 0.0;
  std::vector<double> coef, out;
  // floating lineshapes (tpa_fcn_set_shape_map)
  bool shape = false;
  tpa_model_t model = nullptr;
  TpaShapeMap sm{};
  int64_t rebuilds = 0;  // cache rebuilds after barrier clamp flags changed
  int data_flags = 0;    // extra tpa_eval flags of the data passes (TPA_FLAG_F32, ..F32H, ..F64H)
  int mc_flags = 0;      // extra tpa_eval flags of the MC passes (TPA_FLAG_F64H)
  // shape passes: the MC sets run on their own stream, overlapping the data sets' passes
  hipStream_t mc_stream = nullptr;
  hipEvent_t shapes_set = nullptr;  // the lineshape parameters are on the device
  ~tpa_fcn_s() {
    if (mc_stream) (void)hipStreamDestroy(mc_stream);
    if (shapes_set) (void)hipEventDestroy(shapes_set);
  }
};

namespace {

// Host forms of the barrier polynomial |theta_l(i w)|^2 (tf_pwa/breit_wigner.py:332-355), its
// derivative, and get_relative_p2 (tf_pwa/amp/core.py:324-334) with its mass derivatives.
double bp_poly(int l, double z) {
  static const double C[6][6] = {{1}, {1, 1}, {1, 3, 9}, {1, 6, 45, 225}, {1, 10, 135, 1575, 11025},
                                 {1, 15, 315, 6300, 99225, 893025}};
  const int L = std::min(l, 5);
  double r = 0.0;
  for (int i = 0; i <= L; ++i) r = r * z + C[L][i];
  return r;
}
double bp_poly_deriv(int l, double z) {
  static const double C[6][6] = {{1}, {1, 1}, {1, 3, 9}, {1, 6, 45, 225}, {1, 10, 135, 1575, 11025},
                                 {1, 15, 315, 6300, 99225, 893025}};
  const int L = std::min(l, 5);
  double r = 0.0;
  for (int i = 0; i < L; ++i) r = r * z + C[L][i] * (L - i);
  return r;
}
double q2_of(double M, double m1, double m2) {
  const double s = m1 + m2, d = m1 - m2;
  return (M - s) * (M + s) * (M - d) * (M + d) / (4.0 * M * M);
}
double dq2_dparent(double M, double m1, double m2) {
  const double s = m1 + m2, d = m1 - m2;
  const double F = (M * M - s * s) * (M * M - d * d);
  const double dF = 2.0 * M * (M * M - d * d) + 2.0 * M * (M * M - s * s);
  return dF / (4.0 * M * M) - F / (2.0 * M * M * M);
}
double dq2_ddaughter(double M, double m1, double m2) {  // d / d m1
  const double s = m1 + m2, d = m1 - m2;
  const double dF = -2.0 * s * (M * M - d * d) - 2.0 * d * (M * M - s * s);
  return dF / (4.0 * M * M);
}

// The clamp flags (P_l(q0^2 d^2) > 0) of both vertices of chain c at mass m0.
void barrier_flags(const TpaShapeMap& S, int c, double m0, int32_t* a1, int32_t* a2) {
  const double d2 = S.d[c] * S.d[c];
  const double q01 = q2_of(S.m_top, m0, S.m_bach[c]);
  const double q02 = q2_of(m0, S.m_y[c], S.m_z[c]);
  for (int s = 0; s < S.n_l1[c]; ++s) a1[s] = bp_poly(S.l1[c][s], q01 * d2) > 0 ? 1 : 0;
  for (int s = 0; s < S.n_l2[c]; ++s) a2[s] = bp_poly(S.l2[c][s], q02 * d2) > 0 ? 1 : 0;
}

// Per chain: sqrt(P_l(q0^2 d^2)) of every LS of both vertices (1 when clamped, P_l <= 0) and
// d log / d m0 of it; false when a clamp flag differs from the cache's.
bool barrier_constants(const TpaShapeMap& S, int c, double m0, double* c1, double* c2,
                       double* g1, double* g2) {
  const double d2 = S.d[c] * S.d[c];
  const double q01 = q2_of(S.m_top, m0, S.m_bach[c]);
  const double dq01 = dq2_ddaughter(S.m_top, m0, S.m_bach[c]);
  const double q02 = q2_of(m0, S.m_y[c], S.m_z[c]);
  const double dq02 = dq2_dparent(m0, S.m_y[c], S.m_z[c]);
  for (int s = 0; s < S.n_l1[c]; ++s) {
    const int l = S.l1[c][s];
    const double p0 = bp_poly(l, q01 * d2);
    if ((p0 > 0) != (S.act1[c][s] != 0)) return false;
    c1[s] = p0 > 0 ? std::sqrt(p0) : 1.0;
    g1[s] = p0 > 0 ? bp_poly_deriv(l, q01 * d2) * d2 * dq01 / (2.0 * p0) : 0.0;
  }
  for (int s = 0; s < S.n_l2[c]; ++s) {
    const int l = S.l2[c][s];
    const double p0 = bp_poly(l, q02 * d2);
    if ((p0 > 0) != (S.act2[c][s] != 0)) return false;
    c2[s] = p0 > 0 ? std::sqrt(p0) : 1.0;
    g2[s] = p0 > 0 ? bp_poly_deriv(l, q02 * d2) * d2 * dq02 / (2.0 * p0) : 0.0;
  }
  return true;
}

// NLL and gradient with floating lineshapes (one data and one MC pass per event set, each
// with TPA_FLAG_SHAPE).
int shape_nll_grad(tpa_fcn_s* F, const double* y, double* nll, double* grad) {
  const TpaShapeMap& S = F->sm;
  const int P = F->P, nc = S.n_chain;
  std::vector<double> m0(nc), g0(nc), dlog(P);
  std::vector<double> c1((size_t)nc * TPA_SHAPE_MAX_LS), c2(c1.size()), g1(c1.size()), g2(c1.size());
  bool stale = false;
  for (int c = 0; c < nc; ++c) {
    m0[c] = S.mass_var[c] >= 0 ? y[S.mass_var[c]] : S.mass_fixed[c];
    g0[c] = S.width_var[c] >= 0 ? y[S.width_var[c]] : S.width_fixed[c];
    if (!barrier_constants(S, c, m0[c], &c1[c * TPA_SHAPE_MAX_LS], &c2[c * TPA_SHAPE_MAX_LS],
                           &g1[c * TPA_SHAPE_MAX_LS], &g2[c * TPA_SHAPE_MAX_LS]))
      stale = true;
  }
  if (stale) {
    // a resonance mass crossed the point where P_l(q0^2 d^2) changes sign: new flags on the
    // device, every cache rebuilt with them (the barrier factor per event depends on the flag)
    std::vector<int32_t> a1((size_t)TPA_MAX_CHAIN * TPA_MAX_LS, 0), a2(a1.size(), 0);
    for (int c = 0; c < nc; ++c) {
      barrier_flags(S, c, m0[c], F->sm.act1[c], F->sm.act2[c]);
      for (int s = 0; s < S.n_l1[c]; ++s) a1[c * TPA_MAX_LS + s] = F->sm.act1[c][s];
      for (int s = 0; s < S.n_l2[c]; ++s) a2[c * TPA_MAX_LS + s] = F->sm.act2[c][s];
    }
    int rc = tpa_model_set_barrier_flags(F->model, nc, a1.data(), a2.data());
    for (int pass = 0; pass < 2 && rc == TPA_OK; ++pass)
      for (tpa_events_t ev : pass == 0 ? F->data : F->mc)
        if (rc == TPA_OK) rc = tpa_events_build(ev, 0);
    if (rc != TPA_OK) return rc;
    F->rebuilds += 1;
    for (int c = 0; c < nc; ++c)
      barrier_constants(S, c, m0[c], &c1[c * TPA_SHAPE_MAX_LS], &c2[c * TPA_SHAPE_MAX_LS],
                        &g1[c * TPA_SHAPE_MAX_LS], &g2[c * TPA_SHAPE_MAX_LS]);
  }
  for (int k = 0; k < P; ++k) {
    const int c = S.wave_chain[k], s1 = S.wave_ls1[k], s2 = S.wave_ls2[k];
    F->wave_const[k] = c1[c * TPA_SHAPE_MAX_LS + s1] * c2[c * TPA_SHAPE_MAX_LS + s2];
    dlog[k] = g1[c * TPA_SHAPE_MAX_LS + s1] + g2[c * TPA_SHAPE_MAX_LS + s2];
  }
  int rc = tpa_model_set_lineshapes(F->model, nc, m0.data(), g0.data());
  if (rc != TPA_OK) return rc;
  std::vector<cplx> v, dr, di, c;
  coefficients(F, y, v, dr, di, c);
  for (int k = 0; k < P; ++k) {
    F->coef[2 * k] = c[k].real();
    F->coef[2 * k + 1] = c[k].imag();
  }
  const int flags = TPA_FLAG_SHAPE | (grad ? TPA_FLAG_GRAD : 0);
  // data passes on the model's stream, MC passes on the session's second stream once the
  // lineshape parameters are on the device: the two overlap; then every pass is collected in
  // the order of the sums below
  void* ms = nullptr;
  if ((rc = tpa_model_stream(F->model, &ms)) != TPA_OK) return rc;
  if (!F->mc_stream) {
    if (hipStreamCreateWithFlags(&F->mc_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&F->shapes_set, hipEventDisableTiming) != hipSuccess)
      return tpa_set_error(TPA_ERR_HIP, "session stream");
  }
  if (hipEventRecord(F->shapes_set, (hipStream_t)ms) != hipSuccess ||
      hipStreamWaitEvent(F->mc_stream, F->shapes_set, 0) != hipSuccess)
    return tpa_set_error(TPA_ERR_HIP, "session stream order");
  int queued = TPA_OK;
  for (int pass = 0; pass < 2 && queued == TPA_OK; ++pass)
    for (tpa_events_t ev : pass == 0 ? F->data : F->mc)
      if (queued == TPA_OK)
        queued = tpa_eval_enqueue(ev, F->coef.data(), pass == 0 ? TPA_MODE_DATA : TPA_MODE_MC,
                                  flags, pass == 0 ? ms : (void*)F->mc_stream);
  std::vector<cplx> Gd(3 * P, 0.0), Gm(3 * P, 0.0);
  double sd = 0.0, smc = 0.0;
  for (int pass = 0; pass < 2; ++pass) {
    auto& sets = pass == 0 ? F->data : F->mc;
    auto& G = pass == 0 ? Gd : Gm;
    double& acc = pass == 0 ? sd : smc;
    for (tpa_events_t ev : sets) {
      // every queued pass is collected (waited for) even after a failure
      const int rcv = tpa_eval_collect(ev, F->out.data());
      if (queued != TPA_OK || rcv != TPA_OK) {
        if (rc == TPA_OK) rc = queued != TPA_OK ? queued : rcv;
        continue;
      }
      acc += F->out[0];
      if (grad)
        for (int j = 0; j < 3 * P; ++j) G[j] += cplx(F->out[2 + 2 * j], F->out[3 + 2 * j]);
    }
  }
  if (rc != TPA_OK) {
    (void)hipStreamSynchronize((hipStream_t)ms);
    (void)hipStreamSynchronize(F->mc_stream);
    return rc;
  }
  *nll = -sd + F->sw * std::log(smc);
  if (!grad) return TPA_OK;
  const double r = F->sw / smc;
  std::vector<cplx> F1(P), Fm(P), Fg(P);
  for (int k = 0; k < P; ++k) {
    F1[k] = -Gd[k] + r * Gm[k];
    Fm[k] = -Gd[P + k] + r * Gm[P + k];
    Fg[k] = -Gd[2 * P + k] + r * Gm[2 * P + k];
  }
  grad_from_F1(F, v.data(), dr.data(), di.data(), F1.data(), grad);
  for (int k = 0; k < P; ++k) {
    const int ch = S.wave_chain[k];
    if (S.mass_var[ch] >= 0)
      grad[S.mass_var[ch]] += 2.0 * (c[k] * (dlog[k] * F1[k] + Fm[k])).real();
    if (S.width_var[ch] >= 0) grad[S.width_var[ch]] += 2.0 * (c[k] * Fg[k]).real();
  }
  return TPA_OK;
}

}  // namespace

extern "C" {

int tpa_fcn_create(const TpaCouplingMap* map, const tpa_events_t* data, int n_data,
                   const tpa_events_t* mc, int n_mc, const double* gram, double sw,
                   tpa_fcn_t* out) {
  if (!map || !out || (n_data > 0 && !data) || (n_mc > 0 && !mc))
    return tpa_set_error(TPA_ERR_ARG, "null pointer");
  if (n_data < 1) return tpa_set_error(TPA_ERR_ARG, "no data event set");
  if (!gram && n_mc < 1) return tpa_set_error(TPA_ERR_ARG, "no MC event set and no Gram matrix");
  tpa_fcn_s* F = new tpa_fcn_s();
  const int rc = F->load(map, gram);
  if (rc != TPA_OK) {
    delete F;
    return rc;
  }
  F->data.assign(data, data + n_data);
  if (!gram) F->mc.assign(mc, mc + n_mc);
  F->sw = sw;
  F->coef.resize(2 * F->P);
  F->out.resize(2 + 6 * F->P);  // room for the shape pass (G, Dm, Dg)
  *out = F;
  return TPA_OK;
}

int tpa_fcn_destroy(tpa_fcn_t F) {
  delete F;
  return TPA_OK;
}

int tpa_fcn_set_shape_map(tpa_fcn_t F, tpa_model_t model, const TpaShapeMap* map) {
  if (!F || !model || !map) return tpa_set_error(TPA_ERR_ARG, "null pointer");
  if (!F->gram.empty() || F->mc.empty())
    return tpa_set_error(TPA_ERR_ARG, "floating lineshapes need the MC event sets, not a Gram matrix");
  if (F->data_flags)
    return tpa_set_error(TPA_ERR_UNSUPPORTED, "floating lineshapes read the FP64 data caches");
  if (map->P != F->P || map->n_chain < 1 || map->n_chain > TPA_SHAPE_MAX_CHAIN)
    return tpa_set_error(TPA_ERR_ARG, "shape map sizes do not match the session");
  for (int c = 0; c < map->n_chain; ++c) {
    if (map->mass_var[c] >= F->nv || map->width_var[c] >= F->nv)
      return tpa_set_error(TPA_ERR_ARG, "shape variable index out of range");
    if (map->n_l1[c] < 1 || map->n_l1[c] > TPA_SHAPE_MAX_LS || map->n_l2[c] < 1 ||
        map->n_l2[c] > TPA_SHAPE_MAX_LS)
      return tpa_set_error(TPA_ERR_ARG, "bad LS counts in the shape map");
  }
  for (int k = 0; k < F->P; ++k) {
    const int c = map->wave_chain[k];
    if (c < 0 || c >= map->n_chain || map->wave_ls1[k] < 0 || map->wave_ls1[k] >= map->n_l1[c] ||
        map->wave_ls2[k] < 0 || map->wave_ls2[k] >= map->n_l2[c])
      return tpa_set_error(TPA_ERR_ARG, "bad wave entry in the shape map");
  }
  F->sm = *map;
  F->model = model;
  F->shape = true;
  return TPA_OK;
}

int tpa_fcn_set_data_storage(tpa_fcn_t F, int data_f32) {
  if (!F) return tpa_set_error(TPA_ERR_ARG, "null pointer");
  if (data_f32 && F->shape)
    return tpa_set_error(TPA_ERR_UNSUPPORTED, "FP32 data caches need fixed lineshapes");
  for (tpa_events_t ev : F->data) {
    const int rc = tpa_events_keep_f32(ev, data_f32 ? 1 : 0);
    if (rc != TPA_OK) return rc;
  }
  F->data_flags = data_f32 ? TPA_FLAG_F32 : 0;
  return TPA_OK;
}

int tpa_fcn_set_data_half(tpa_fcn_t F, int kind) {
  if (!F) return tpa_set_error(TPA_ERR_ARG, "null pointer");
  if (kind < 0 || kind > 2) return tpa_set_error(TPA_ERR_ARG, "kind: 0, 1 (FP32) or 2 (FP64)");
  if (kind && F->shape)
    return tpa_set_error(TPA_ERR_UNSUPPORTED, "halved data caches need fixed lineshapes");
  for (tpa_events_t ev : F->data) {
    const int rc = tpa_events_keep_half(ev, kind);
    if (rc != TPA_OK) return rc;
  }
  F->data_flags = kind == 1 ? TPA_FLAG_F32H : kind == 2 ? TPA_FLAG_F64H : 0;
  return TPA_OK;
}

int tpa_fcn_set_mc_half(tpa_fcn_t F, int on) {
  if (!F) return tpa_set_error(TPA_ERR_ARG, "null pointer");
  if (on && F->shape)
    return tpa_set_error(TPA_ERR_UNSUPPORTED, "halved MC caches need fixed lineshapes");
  for (tpa_events_t ev : F->mc) {
    const int rc = tpa_events_keep_half(ev, on ? 2 : 0);
    if (rc != TPA_OK) return rc;
  }
  F->mc_flags = on ? TPA_FLAG_F64H : 0;
  return TPA_OK;
}

int tpa_fcn_cache_rebuilds(tpa_fcn_t F, int64_t* n) {
  if (!F || !n) return tpa_set_error(TPA_ERR_ARG, "null pointer");
  *n = F->rebuilds;
  return TPA_OK;
}

int tpa_fcn_nll_grad(tpa_fcn_t F, const double* y, double* nll, double* grad) {
  if (!F || !y || !nll) return tpa_set_error(TPA_ERR_ARG, "null pointer");
  if (F->shape) return shape_nll_grad(F, y, nll, grad);
  const int P = F->P;
  std::vector<cplx> v, dr, di, c, Sk, F1(P, 0.0);
  coefficients(F, y, v, dr, di, c);
  for (int k = 0; k < P; ++k) {
    F->coef[2 * k] = c[k].real();
    F->coef[2 * k + 1] = c[k].imag();
  }
  const int flags = grad ? TPA_FLAG_GRAD : 0;
  double sd = 0.0;
  for (tpa_events_t ev : F->data) {
    const int rc = tpa_eval(ev, F->coef.data(), TPA_MODE_DATA, flags | F->data_flags,
                            F->out.data(), nullptr);
    if (rc != TPA_OK) return rc;
    sd += F->out[0];
    if (grad)
      for (int k = 0; k < P; ++k) F1[k] -= cplx(F->out[2 + 2 * k], F->out[3 + 2 * k]);
  }
  double smc = 0.0;
  if (!F->gram.empty()) {
    smc = mc_term(F, c, Sk);
  } else {
    Sk.assign(P, 0.0);
    for (tpa_events_t ev : F->mc) {
      const int rc = tpa_eval(ev, F->coef.data(), TPA_MODE_MC, flags | F->mc_flags,
                              F->out.data(), nullptr);
      if (rc != TPA_OK) return rc;
      smc += F->out[0];
      if (grad)
        for (int k = 0; k < P; ++k) Sk[k] += cplx(F->out[2 + 2 * k], F->out[3 + 2 * k]);
    }
  }
  *nll = -sd + F->sw * std::log(smc);
  if (grad) {
    const double r = F->sw / smc;
    for (int k = 0; k < P; ++k) F1[k] += r * Sk[k];
    grad_from_F1(F, v.data(), dr.data(), di.data(), F1.data(), grad);
  }
  return TPA_OK;
}

int tpa_fcn_minimize(tpa_fcn_t F, const double* y0, double gtol, int maxiter, double* y_out,
                     double* nll_out, int* status, int64_t* nfev, int64_t* niter) {
  if (!F || !y0 || !y_out || !nll_out || !status) return tpa_set_error(TPA_ERR_ARG, "null pointer");
  if (!(gtol > 0.0)) return tpa_set_error(TPA_ERR_ARG, "gtol must be positive");
  struct Failed {
    int rc;
  };
  int64_t calls = 0;
  auto fy = [&](const double* y, double* gy) {
    double f = 0.0;
    const int rc = tpa_fcn_nll_grad(F, y, &f, gy);
    if (rc != TPA_OK) throw Failed{rc};
    ++calls;
    return f;
  };
  BfgsResult r;
  try {
    r = bfgs_minimize(F, y0, fy, gtol, maxiter, nullptr);
  } catch (const Failed& e) {
    if (nfev) *nfev = calls;
    return e.rc;  // the message is already set
  }
  std::copy(r.y.begin(), r.y.end(), y_out);
  *nll_out = r.f;
  *status = r.status;
  if (nfev) *nfev = calls;
  if (niter) *niter = r.niter;
  return TPA_OK;
}

}  // extern "C"
