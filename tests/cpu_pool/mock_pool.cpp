// CPU stand-in for the device event pool, for host tests of the serving runtime
// (tpa_serve.cpp) without a GPU: the same tpa_pool_eval contract (per request: the weighted
// clip_log data sum and its adjoint G_k) computed in plain loops from a host partial-wave cache
// h [P][capacity][lam_pad] that the test supplies.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "tfpwa_amd.h"
#include "tpa_internal.h"

struct tpa_pool_s {
  int P = 0, lg = 0;
  int64_t cap = 0, used = 0;
  std::vector<double> h;  // interleaved complex
  std::vector<double> w;
  int64_t evals = 0;
  // batch slots of the pipelined evaluation: results are computed at submit time
  std::vector<double> slot_out[2];
  bool slot_busy[2] = {false, false};
};

static thread_local std::string g_err;

int tpa_set_error(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

extern "C" {

const char* tpa_last_error(void) { return g_err.c_str(); }

int mock_pool_create(int P, int lam_pad, int64_t capacity, tpa_pool_t* out) {
  tpa_pool_s* p = new tpa_pool_s();
  p->P = P;
  p->lg = lam_pad;
  p->cap = capacity;
  p->h.assign((size_t)2 * P * capacity * lam_pad, 0.0);
  p->w.assign(capacity, 0.0);
  *out = p;
  return 0;
}

int mock_pool_destroy(tpa_pool_t p) {
  delete p;
  return 0;
}

// h: [P][n][lam_pad] complex (interleaved) of n events, appended after the last sample.
int mock_pool_add(tpa_pool_t p, const double* h, int64_t n, const double* w, int64_t* offset) {
  if (p->used + n > p->cap) return tpa_set_error(-1, "pool full");
  const int64_t off = p->used;
  for (int k = 0; k < p->P; ++k)
    std::memcpy(&p->h[2 * (((size_t)k * p->cap + off) * p->lg)], &h[2 * ((size_t)k * n * p->lg)],
                sizeof(double) * 2 * n * p->lg);
  std::memcpy(&p->w[off], w, sizeof(double) * n);
  p->used += n;
  *offset = off;
  return 0;
}

int64_t mock_pool_evals(tpa_pool_t p) { return p->evals; }

int tpa_pool_eval(tpa_pool_t p, int n_items, const int64_t* ranges, const double* coefs, int mode,
                  double* out, double* ms) {
  if (mode != TPA_MODE_DATA) return tpa_set_error(-3, "data sums only");
  const int P = p->P, lg = p->lg, nacc = 2 + 2 * P;
  const double eps = 1e-6, logeps = std::log(1e-6);
  for (int i = 0; i < n_items; ++i) {
    const int64_t off = ranges[2 * i], n = ranges[2 * i + 1];
    if (off < 0 || n < 1 || off + n > p->used) return tpa_set_error(-1, "range outside the pool");
    const double* c = coefs + (size_t)2 * P * i;
    double* o = out + (size_t)nacc * i;
    std::memset(o, 0, sizeof(double) * nacc);
    std::vector<double> ar(lg), ai(lg);
    for (int64_t e = off; e < off + n; ++e) {
      double f = 0.0;
      for (int l = 0; l < lg; ++l) {
        double sr = 0.0, si = 0.0;
        for (int k = 0; k < P; ++k) {
          const double* hv = &p->h[2 * (((size_t)k * p->cap + e) * lg + l)];
          sr += c[2 * k] * hv[0] - c[2 * k + 1] * hv[1];
          si += c[2 * k] * hv[1] + c[2 * k + 1] * hv[0];
        }
        ar[l] = sr;
        ai[l] = si;
        f += sr * sr + si * si;
      }
      const double w = p->w[e];
      double cl, dcl;
      if (f > eps) {
        cl = std::log(f);
        dcl = 1.0 / f;
      } else {
        const double t = (f - eps) / eps;
        cl = (logeps + t) - (t * t) / 2.0;
        dcl = 1.0 / eps - t / eps;
      }
      o[0] += w * cl;
      const double u = w * dcl;
      for (int k = 0; k < P; ++k)
        for (int l = 0; l < lg; ++l) {
          const double* hv = &p->h[2 * (((size_t)k * p->cap + e) * lg + l)];
          // u conj(A) h
          o[2 + 2 * k] += u * (ar[l] * hv[0] + ai[l] * hv[1]);
          o[3 + 2 * k] += u * (ar[l] * hv[1] - ai[l] * hv[0]);
        }
    }
    p->evals += 1;
  }
  if (ms) *ms = 0.0;
  return 0;
}

// Second moments of one range in the packed layout of tpa_pool_hess (data mode).
int tpa_pool_hess(tpa_pool_t p, int64_t off, int64_t n, const double* c, double* out) {
  const int P = p->P, lg = p->lg, Q = 2 * P;
  const int T = Q * (Q + 1) / 2;
  if (off < 0 || n < 1 || off + n > p->used) return tpa_set_error(-1, "range outside the pool");
  const double eps = 1e-6, logeps = std::log(1e-6);
  std::fill(out, out + 2 + 2 * P + 2 * T, 0.0);
  double* O1 = out + 2 + 2 * P;
  double* O2 = O1 + T;
  std::vector<double> ar(lg), ai(lg), z(Q), y((size_t)lg * Q);
  for (int64_t e = off; e < off + n; ++e) {
    double f = 0.0;
    for (int l = 0; l < lg; ++l) {
      double sr = 0.0, si = 0.0;
      for (int k = 0; k < P; ++k) {
        const double* hv = &p->h[2 * (((size_t)k * p->cap + e) * lg + l)];
        sr += c[2 * k] * hv[0] - c[2 * k + 1] * hv[1];
        si += c[2 * k] * hv[1] + c[2 * k + 1] * hv[0];
        y[(size_t)l * Q + k] = hv[0];
        y[(size_t)l * Q + P + k] = hv[1];
      }
      ar[l] = sr;
      ai[l] = si;
      f += sr * sr + si * si;
    }
    const double w = p->w[e];
    const double dom = (w == 0.0) ? 1.0 : w;
    const double r = w / dom, x = f * r;
    double cl, d1, d2;
    if (x > eps) {
      cl = std::log(x);
      d1 = 1.0 / x;
      d2 = -1.0 / (x * x);
    } else {
      const double t = (x - eps) / eps;
      cl = (logeps + t) - (t * t) / 2.0;
      d1 = 1.0 / eps - t / eps;
      d2 = -1.0 / (eps * eps);
    }
    const double u = (w * d1) * r, v = (w * d2) * (r * r);
    out[0] += w * cl;
    for (int k = 0; k < P; ++k) {
      double qr = 0.0, qi = 0.0;
      for (int l = 0; l < lg; ++l) {
        const double hr = y[(size_t)l * Q + k], hi = y[(size_t)l * Q + P + k];
        qr += ar[l] * hr + ai[l] * hi;
        qi += ar[l] * hi - ai[l] * hr;
      }
      z[k] = qr;
      z[P + k] = qi;
      out[2 + 2 * k] += u * qr;
      out[3 + 2 * k] += u * qi;
    }
    int idx = 0;
    for (int i = 0; i < Q; ++i)
      for (int j = i; j < Q; ++j, ++idx) {
        O1[idx] += v * z[i] * z[j];
        double s2 = 0.0;
        for (int l = 0; l < lg; ++l) s2 += y[(size_t)l * Q + i] * y[(size_t)l * Q + j];
        O2[idx] += u * s2;
      }
  }
  return 0;
}

}  // extern "C"

// The pipelined evaluation of the serving batcher (declared in tpa_internal.h): the sums are
// computed at submit time into the slot and handed out by finish.
int tpa_pool_eval_submit(tpa_pool_t p, int* slot_io, int n_items, const int64_t* ranges,
                         const double* coefs, int mode) {
  int si = *slot_io;
  if (si < 0)
    for (int i = 0; i < 2 && si < 0; ++i)
      if (!p->slot_busy[i]) si = i;
  if (si < 0 || si > 1 || p->slot_busy[si]) return tpa_set_error(-1, "every batch slot is in flight");
  p->slot_out[si].assign((size_t)n_items * (2 + 2 * p->P), 0.0);
  const int rc = tpa_pool_eval(p, n_items, ranges, coefs, mode, p->slot_out[si].data(), nullptr);
  if (rc != 0) return rc;
  p->slot_busy[si] = true;
  *slot_io = si;
  return 0;
}

int tpa_pool_eval_ready(tpa_pool_t p, int slot) { return p->slot_busy[slot] ? 1 : -1; }

int tpa_pool_eval_finish(tpa_pool_t p, int slot, double* out, double* ms) {
  if (!p->slot_busy[slot]) return tpa_set_error(-1, "no batch in this slot");
  std::memcpy(out, p->slot_out[slot].data(), sizeof(double) * p->slot_out[slot].size());
  p->slot_busy[slot] = false;
  if (ms) *ms = 0.0;
  return 0;
}
