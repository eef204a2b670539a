// Single-fit likelihood session: FCN.nll_grad (tf_pwa/model/model.py:1207-1240) for fixed
// lineshapes as one native call. The trainable values y go through the coupling map to the
// partial-wave coefficients, each data set and (FP64 path) each MC set is one tpa_eval pass, or
// (mixed path) the MC normalisation is the Gram form c^dagger I c, and the gradient in y follows
// by the coupling chain rule. The same arithmetic as Model._nll_grad of the Python host, without
// its per-call interpreter work.
#include <cmath>
#include <cstring>
#include <vector>

#include "tfpwa_amd.h"
#include "tpa_coupling.h"
#include "tpa_internal.h"

using namespace tpa_cpl;

struct tpa_fcn_s : CouplingMap {
  std::vector<tpa_events_t> data, mc;
  double sw = 0.0;
  std::vector<double> coef, out;
};

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
  F->out.resize(2 + 2 * F->P);
  *out = F;
  return TPA_OK;
}

int tpa_fcn_destroy(tpa_fcn_t F) {
  delete F;
  return TPA_OK;
}

int tpa_fcn_nll_grad(tpa_fcn_t F, const double* y, double* nll, double* grad) {
  if (!F || !y || !nll) return tpa_set_error(TPA_ERR_ARG, "null pointer");
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
    const int rc = tpa_eval(ev, F->coef.data(), TPA_MODE_DATA, flags, F->out.data(), nullptr);
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
      const int rc = tpa_eval(ev, F->coef.data(), TPA_MODE_MC, flags, F->out.data(), nullptr);
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

}  // extern "C"
