// Many concurrent fits on one GPU (host runtime of the serving path).
//
// Every fit runs its own BFGS minimisation (the role of fit_scipy's scipy.optimize.minimize
// loop, tf_pwa/fit.py:189-300) on a host thread. An NLL + gradient evaluation of a fit
// (FCN.nll_grad, tf_pwa/model/model.py:1216-1240) computes the partial-wave coefficients
// c_k = const_k * v_total * v_g1 * v_g2 from the fit's parameters, hands them to one batcher
// thread and waits. The batcher gathers the pending requests of all fits and evaluates them in
// ONE batched pool launch (data sum and its adjoint per request), with the next batch queued
// behind the running one; the MC normalisation is the
// shared partial-wave Gram matrix I, so int_mc = c^dagger I c and its adjoint conj(I c) cost
// O(P^2) on the host. The gradient in the fit's parameters follows from the adjoints by the
// chain rule through the three complex factors of each c_k.
//
// A fit's result does not depend on which other fits share its batches: tpa_pool_eval's
// per-request sums have a fixed order, and each fit's BFGS is deterministic.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "tfpwa_amd.h"
#include "tpa_bfgs.h"
#include "tpa_coupling.h"
#include "tpa_internal.h"

namespace {

using namespace tpa_cpl;
using Clock = std::chrono::steady_clock;
// The batcher's one timed wait (the gathering window) runs on the monotonic clock. GCC 11's
// ThreadSanitizer runtime does not intercept pthread_cond_clockwait, which libstdc++ uses for
// steady_clock waits, so it misses the mutex being released during them and reports every
// such wait as a double lock; the ThreadSanitizer build of the host tests defines
// TPA_SERVE_WAIT_SYSTEM_CLOCK to wait on the system clock there (pthread_cond_timedwait).
#ifdef TPA_SERVE_WAIT_SYSTEM_CLOCK
using WaitClock = std::chrono::system_clock;
#else
using WaitClock = std::chrono::steady_clock;
#endif

struct Cancelled {};
struct EngineFailure {};

struct Fit {
  int id = -1;
  int64_t off = 0, n = 0;
  double sw = 0.0;
  std::vector<double> x0;
  // result (y space: physical values of the trainable parameters)
  std::vector<double> y;
  double nll = NAN;
  int status = -1;
  int64_t nfev = 0, niter = 0;
  // the request slot shared with the batcher: coef and t_submit guarded by the server mutex;
  // the result (out, failed, ready) by the fit's own mutex, so a fit woken with its result does
  // not queue on the server mutex (with many fits per batch that convoy idled the GPU)
  std::vector<double> coef, out;
  bool ready = false, failed = false;
  Clock::time_point t_submit;
  std::mutex rmu;
  std::condition_variable cv;
  std::thread th;
};

// Wake a fit waiting for its result: the fit's mutex is taken so a waiter between its
// predicate check and its wait cannot miss the notification.
void wake(Fit* F) {
  { std::lock_guard<std::mutex> fl(F->rmu); }
  F->cv.notify_all();
}

}  // namespace

struct tpa_server_s : tpa_cpl::CouplingMap {
  tpa_pool_t pool = nullptr;
  double gtol = 1e-5;
  int maxiter = 0;
  int max_batch = 1;
  double window_us = 200.0;
  int hess_start = 0;  // BFGS starts from the exact inverse Hessian at the first point

  std::mutex mu;
  std::condition_variable cv;  // batcher wake-up: new request, a client leaving, stop
  std::vector<Fit*> fits;      // by id
  std::deque<Fit*> pending;
  int clients = 0;             // fit threads alive + direct evaluations in progress
  // copies of pending.size() and clients, written under mu, read without it by the batcher
  // while it waits for a running batch
  std::atomic<int> n_pending{0}, n_clients{0};
  std::vector<int> finished;   // ids not yet polled
  std::atomic<bool> stop{false};  // written under mu; read also by fits waiting on their own mutex
  std::string batch_error;
  std::thread batcher;
  std::vector<double> lat_us;
  int64_t n_requests = 0, n_batches = 0, n_events = 0;
  double gpu_ms = 0.0;
};

namespace {

// One request through the batcher: coefficients in, data sum and adjoint out.
void request(tpa_server_s* S, Fit* F, const std::vector<cplx>& c) {
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) throw Cancelled();
    F->coef.resize(2 * S->P);
    for (int k = 0; k < S->P; ++k) {
      F->coef[2 * k] = c[k].real();
      F->coef[2 * k + 1] = c[k].imag();
    }
    {
      std::lock_guard<std::mutex> fl(F->rmu);
      F->ready = false;
      F->failed = false;
    }
    F->t_submit = Clock::now();
    S->pending.push_back(F);
    S->n_pending.store((int)S->pending.size());
  }
  S->cv.notify_all();
  std::unique_lock<std::mutex> fl(F->rmu);
  F->cv.wait(fl, [&] { return F->ready || S->stop; });
  if (!F->ready) throw Cancelled();
  if (F->failed) throw EngineFailure();
}

// NLL and its gradient in y: NLL = -sum w clip_log f + sw log(c^dagger I c) (the Model.nll of
// tf_pwa/model/model.py:407-449 with the MC sum replaced by its Gram form).
double nll_grad_y(tpa_server_s* S, Fit* F, const double* y, double* gy) {
  const int P = S->P;
  std::vector<cplx> v, dr, di, c, Sk;
  coefficients(S, y, v, dr, di, c);
  request(S, F, c);
  const double sd = F->out[0];
  const double smc = mc_term(S, c, Sk);
  const double nll = -sd + F->sw * std::log(smc);
  if (gy) {
    std::vector<cplx> F1(P);
    const double r = F->sw / smc;
    for (int k = 0; k < P; ++k) F1[k] = -cplx(F->out[2 + 2 * k], F->out[3 + 2 * k]) + r * Sk[k];
    grad_from_F1(S, v.data(), dr.data(), di.data(), F1.data(), gy);
  }
  F->nfev += 1;
  return nll;
}

// NLL, gradient and exact Hessian in y from one second-moment pass over the fit's data
// (tpa_pool_hess) and the Gram matrix: with F1, Fcc = d2/dc dc, Fcb = d2/dc dconj(c),
// H = 2 Re[J Fcc J^T + J Fcb J^H + sum_k F1_k d2c_k/dy dy] (BaseModel.nll_grad_hessian,
// tf_pwa/model/model.py:513-560, for the couplings).
double nll_grad_hess_y(tpa_server_s* S, Fit* F, const double* y, double* gy, double* H) {
  const int P = S->P, Q = 2 * P, nv = S->nv;
  const int T = Q * (Q + 1) / 2;
  std::vector<cplx> v, dr, di, c, Sk;
  coefficients(S, y, v, dr, di, c);
  std::vector<double> coef(2 * P), out(2 + 2 * P + 2 * T);
  for (int k = 0; k < P; ++k) {
    coef[2 * k] = c[k].real();
    coef[2 * k + 1] = c[k].imag();
  }
  if (tpa_pool_hess(S->pool, F->off, F->n, coef.data(), out.data()) != TPA_OK) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->batch_error = tpa_last_error();
    throw EngineFailure();
  }
  const double* O1 = out.data() + 2 + 2 * P;
  const double* O2 = O1 + T;
  auto O = [Q](const double* m, int i, int j) {
    if (i > j) std::swap(i, j);
    return m[i * Q - (i * (i - 1)) / 2 + (j - i)];
  };
  const double smc = mc_term(S, c, Sk);
  const double sw = F->sw;
  std::vector<cplx> F1(P), Fcc((size_t)P * P), Fcb((size_t)P * P);
  for (int k = 0; k < P; ++k) F1[k] = -cplx(out[2 + 2 * k], out[3 + 2 * k]) + (sw / smc) * Sk[k];
  for (int k = 0; k < P; ++k)
    for (int l = 0; l < P; ++l) {
      const cplx dcc(O(O1, k, l) - O(O1, P + k, P + l), O(O1, k, P + l) + O(O1, P + k, l));
      const cplx dcb(O(O1, k, l) + O(O1, P + k, P + l) + O(O2, k, l) + O(O2, P + k, P + l),
                     O(O1, P + k, l) - O(O1, k, P + l) + O(O2, P + k, l) - O(O2, k, P + l));
      const cplx scb = S->gram[(size_t)l * P + k];  // d2S/dc_k dconj(c_l) = I_lk
      Fcc[(size_t)k * P + l] = -dcc - (sw / (smc * smc)) * Sk[k] * Sk[l];
      Fcb[(size_t)k * P + l] = -dcb + sw * (scb / smc - Sk[k] * std::conj(Sk[l]) / (smc * smc));
    }
  grad_from_F1(S, v.data(), dr.data(), di.data(), F1.data(), gy);
  hess_chain(S, v.data(), dr.data(), di.data(), F1.data(), Fcc.data(), Fcb.data(), H);
  F->nfev += 1;
  return -out[0] + sw * std::log(smc);
}

// ------------------------------------------------------------------ BFGS
// One fit's minimisation through the shared quasi-Newton driver (tpa_bfgs.h); every objective
// evaluation is a request to the batcher.
void run_bfgs(tpa_server_s* S, Fit* F) {
  const int n = S->nv;
  std::vector<double> Hx0;
  if (S->hess_start) {
    // inverse of the exact Hessian in x: dy/dx Hy dy/dx + diag(g_y d2y/dx2)
    std::vector<double> x(n), y(n), dydx(n), gy(n), Hyy((size_t)n * n), Hx((size_t)n * n);
    for (int j = 0; j < n; ++j) x[j] = to_x(S, j, F->x0[j]);
    to_y(S, x.data(), y.data(), dydx.data());
    nll_grad_hess_y(S, F, y.data(), gy.data(), Hyy.data());
    for (int a = 0; a < n; ++a) {
      for (int b = 0; b < n; ++b) Hx[(size_t)a * n + b] = dydx[a] * Hyy[(size_t)a * n + b] * dydx[b];
      Hx[(size_t)a * n + a] += gy[a] * d2y_dx2(S, a, x[a]);
    }
    shifted_cholesky_inverse(Hx, n, Hx0);
  }
  auto fy = [&](const double* y, double* gy) { return nll_grad_y(S, F, y, gy); };
  BfgsResult r = bfgs_minimize(S, F->x0.data(), fy, S->gtol, S->maxiter,
                               S->hess_start ? &Hx0 : nullptr);
  F->niter = r.niter;
  F->status = r.status;
  F->nll = r.f;
  F->y = std::move(r.y);
}

void fit_main(tpa_server_s* S, Fit* F) {
  try {
    run_bfgs(S, F);
  } catch (const Cancelled&) {
    F->status = 3;
  } catch (const EngineFailure&) {
    F->status = 4;
  }
  std::lock_guard<std::mutex> lk(S->mu);
  S->clients -= 1;
  S->n_clients.store(S->clients);
  S->finished.push_back(F->id);
  S->cv.notify_all();
}

// ------------------------------------------------------------------ batcher
// One batch on its way through the pool: the requests, their ranges and coefficients, and the
// pool batch slot it occupies while it runs.
struct Batch {
  std::vector<Fit*> fits;
  std::vector<int64_t> ranges;
  std::vector<double> coefs, out;
  int slot = -1;
  bool failed = false;
};

// Hand a finished batch's sums to its fits. The counters and latencies are recorded under the
// server mutex (held on entry and on return); each fit then gets its result under its own mutex
// with the server mutex released, so the fits woken by the batch and the fits submitting new
// requests do not queue behind the batcher while it wakes them one by one. (The fit's mutex is
// held across the notification: a fit of a direct evaluation lives on its caller's stack and
// may end as soon as it sees its result.)
void deliver(tpa_server_s* S, Batch& B, double ms, const std::string& err,
             std::unique_lock<std::mutex>& lk) {
  const int nacc = 2 + 2 * S->P;
  const int m = (int)B.fits.size();
  const auto now = Clock::now();
  S->n_batches += 1;
  S->n_requests += m;
  for (int i = 0; i < m; ++i) S->n_events += B.ranges[2 * i + 1];
  S->gpu_ms += ms;
  if (B.failed) S->batch_error = err;
  for (int i = 0; i < m; ++i)
    S->lat_us.push_back(std::chrono::duration<double, std::micro>(now - B.fits[i]->t_submit).count());
  lk.unlock();
  for (int i = 0; i < m; ++i) {
    Fit* F = B.fits[i];
    std::lock_guard<std::mutex> fl(F->rmu);
    if (!B.failed)
      F->out.assign(B.out.begin() + (size_t)nacc * i, B.out.begin() + (size_t)nacc * (i + 1));
    F->failed = B.failed;
    F->ready = true;
    F->cv.notify_one();
  }
  lk.lock();
  B.fits.clear();
  B.slot = -1;
}

// Wait for a running batch and deliver it (called and returning with the mutex held; the wait
// itself runs unlocked).
void complete(tpa_server_s* S, Batch& B, std::unique_lock<std::mutex>& lk) {
  lk.unlock();
  double ms = 0.0;
  B.out.resize((size_t)(2 + 2 * S->P) * B.fits.size());
  const int rc = tpa_pool_eval_finish(S->pool, B.slot, B.out.data(), &ms);
  const std::string err = rc == TPA_OK ? std::string() : std::string(tpa_last_error());
  lk.lock();
  B.failed = rc != TPA_OK;
  deliver(S, B, ms, err, lk);
}

// The batcher keeps up to two batches on the GPU: while one runs, the requests of the other
// fits are gathered and the next batch is queued behind it, so the GPU does not idle between
// batches and a request does not wait for the batch ahead to return before it is launched.
// A batch launches when every fit outside the running batch is waiting, when max_batch
// requests are pending, when the running batch returns (the requests gathered meanwhile go
// at once), or, with none running, window_us after the first request.
// While a batch runs the batcher checks it between yields rather than sleeping on a timed
// wait: a timed wait oversleeps by the kernel's timer slack (~50 us by default), which would
// land on every request's latency. The spinning costs one host core while the GPU works; it
// runs without the server mutex, so the fits queue requests meanwhile.
void batcher_main(tpa_server_s* S) {
  const int P = S->P;
  Batch bt[2];
  int run = -1;  // index in bt of the batch on the GPU (-1: none)
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    // wait for requests; meanwhile deliver the running batch as soon as it is done (that wait
    // runs without the server mutex, watching the pending count)
    while (!S->stop && S->pending.empty()) {
      if (run >= 0) {
        const int slot = bt[run].slot;
        lk.unlock();
        while (!S->stop && S->n_pending.load() == 0 && tpa_pool_eval_ready(S->pool, slot) == 0)
          std::this_thread::yield();
        lk.lock();
        if (S->pending.empty() && tpa_pool_eval_ready(S->pool, slot) != 0) {
          complete(S, bt[run], lk);
          run = -1;
        }
      } else {
        S->cv.wait(lk);
      }
    }
    if (S->stop) break;
    // gather
    const Clock::time_point deadline =
        Clock::now() + std::chrono::duration_cast<Clock::duration>(
                           std::chrono::duration<double, std::micro>(S->window_us));
    for (;;) {
      const int busy = run >= 0 ? (int)bt[run].fits.size() : 0;
      if (S->stop || (int)S->pending.size() >= S->clients - busy ||
          (int)S->pending.size() >= S->max_batch)
        break;
      if (run >= 0) {
        // a batch is on the GPU: gather until it returns, then launch at once (the running
        // batch is delivered right after the launch), so the GPU only waits for the launch.
        // The wait runs without the server mutex, watching the counters the fits update, so
        // fits queue their requests meanwhile without waiting for the batcher.
        const int slot = bt[run].slot;
        lk.unlock();
        while (!S->stop && tpa_pool_eval_ready(S->pool, slot) == 0 &&
               S->n_pending.load() < S->n_clients.load() - busy &&
               S->n_pending.load() < S->max_batch)
          std::this_thread::yield();
        lk.lock();
        break;
      }
      const Clock::time_point now = Clock::now();
      if (now >= deadline) break;
      S->cv.wait_until(lk, WaitClock::now() + (deadline - now));
    }
    if (S->stop) break;
    const int nb = run == 0 ? 1 : 0;
    Batch& B = bt[nb];
    B.fits.clear();
    while (!S->pending.empty() && (int)B.fits.size() < S->max_batch) {
      B.fits.push_back(S->pending.front());
      S->pending.pop_front();
    }
    S->n_pending.store((int)S->pending.size());
    const int m = (int)B.fits.size();
    B.ranges.resize(2 * m);
    B.coefs.resize((size_t)2 * P * m);
    for (int i = 0; i < m; ++i) {
      B.ranges[2 * i] = B.fits[i]->off;
      B.ranges[2 * i + 1] = B.fits[i]->n;
      std::memcpy(&B.coefs[(size_t)2 * P * i], B.fits[i]->coef.data(), sizeof(double) * 2 * P);
    }
    lk.unlock();
    int slot = -1;
    const int rc = tpa_pool_eval_submit(S->pool, &slot, m, B.ranges.data(), B.coefs.data(),
                                        TPA_MODE_DATA);
    const std::string err = rc == TPA_OK ? std::string() : std::string(tpa_last_error());
    lk.lock();
    if (rc != TPA_OK) {
      B.failed = true;
      deliver(S, B, 0.0, err, lk);
      continue;
    }
    B.slot = slot;
    B.failed = false;
    // the new batch is queued behind the running one: deliver that one when it returns
    if (run >= 0) complete(S, bt[run], lk);
    run = nb;
  }
  if (run >= 0) complete(S, bt[run], lk);
  // wake every waiting client so it sees the stop
  for (Fit* F : S->fits) wake(F);
  for (Fit* F : S->pending) wake(F);
}

}  // namespace

extern "C" {

int tpa_server_create(tpa_pool_t pool, const TpaCouplingMap* map, const double* gram,
                      double gtol, int maxiter, int max_batch, double window_us, int hess_start,
                      tpa_server_t* out) {
  if (!pool || !map || !gram || !out) return tpa_set_error(TPA_ERR_ARG, "null pointer");
  if (max_batch < 1 || !(gtol > 0.0) || window_us < 0.0)
    return tpa_set_error(TPA_ERR_ARG, "bad server options");
  tpa_server_s* S = new tpa_server_s();
  const int rc = S->load(map, gram);
  if (rc != TPA_OK) {
    delete S;
    return rc;
  }
  S->pool = pool;
  S->gtol = gtol;
  S->maxiter = maxiter;
  S->max_batch = max_batch;
  S->window_us = window_us;
  S->hess_start = hess_start ? 1 : 0;
  S->batcher = std::thread(batcher_main, S);
  *out = S;
  return TPA_OK;
}

int tpa_server_destroy(tpa_server_t S) {
  if (!S) return TPA_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
    S->cv.notify_all();
    for (Fit* F : S->fits) wake(F);
  }
  if (S->batcher.joinable()) S->batcher.join();
  for (Fit* F : S->fits)
    if (F->th.joinable()) F->th.join();
  for (Fit* F : S->fits) delete F;
  delete S;
  return TPA_OK;
}

int tpa_server_nll_grad(tpa_server_t S, int64_t offset, int64_t n, double sw, const double* y,
                        double* nll, double* grad) {
  if (!S || !y || !nll) return tpa_set_error(TPA_ERR_ARG, "null pointer");
  Fit F;
  F.off = offset;
  F.n = n;
  F.sw = sw;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return tpa_set_error(TPA_ERR_ARG, "server stopped");
    S->clients += 1;
    S->n_clients.store(S->clients);
  }
  int rc = TPA_OK;
  try {
    *nll = nll_grad_y(S, &F, y, grad);
  } catch (const Cancelled&) {
    rc = tpa_set_error(TPA_ERR_ARG, "server stopped");
  } catch (const EngineFailure&) {
    std::lock_guard<std::mutex> lk(S->mu);
    rc = tpa_set_error(TPA_ERR_HIP, "batched evaluation failed: " + S->batch_error);
  }
  std::lock_guard<std::mutex> lk(S->mu);
  S->clients -= 1;
  S->n_clients.store(S->clients);
  S->cv.notify_all();
  return rc;
}

int tpa_server_nll_grad_hess(tpa_server_t S, int64_t offset, int64_t n, double sw,
                             const double* y, double* nll, double* grad, double* hess) {
  if (!S || !y || !nll || !grad || !hess) return tpa_set_error(TPA_ERR_ARG, "null pointer");
  Fit F;
  F.off = offset;
  F.n = n;
  F.sw = sw;
  try {
    *nll = nll_grad_hess_y(S, &F, y, grad, hess);
  } catch (const EngineFailure&) {
    std::lock_guard<std::mutex> lk(S->mu);
    return tpa_set_error(TPA_ERR_HIP, "second-moment pass failed: " + S->batch_error);
  }
  return TPA_OK;
}

int tpa_server_submit(tpa_server_t S, int64_t offset, int64_t n, double sw, const double* y0,
                      int* fit_id) {
  if (!S || !y0 || !fit_id) return tpa_set_error(TPA_ERR_ARG, "null pointer");
  if (n < 1) return tpa_set_error(TPA_ERR_ARG, "event count must be positive");
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->stop) return tpa_set_error(TPA_ERR_ARG, "server stopped");
  Fit* F = new Fit();
  F->id = (int)S->fits.size();
  F->off = offset;
  F->n = n;
  F->sw = sw;
  F->x0.assign(y0, y0 + S->nv);
  S->fits.push_back(F);
  S->clients += 1;
  S->n_clients.store(S->clients);
  F->th = std::thread(fit_main, S, F);
  *fit_id = F->id;
  return TPA_OK;
}

int tpa_server_poll(tpa_server_t S, int* ids, int max_ids, int* n_done) {
  if (!S || !n_done || (max_ids > 0 && !ids)) return tpa_set_error(TPA_ERR_ARG, "null pointer");
  std::vector<Fit*> done;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    const int m = std::min<int>(max_ids, (int)S->finished.size());
    for (int i = 0; i < m; ++i) {
      ids[i] = S->finished[i];
      done.push_back(S->fits[S->finished[i]]);
    }
    S->finished.erase(S->finished.begin(), S->finished.begin() + m);
    *n_done = m;
  }
  for (Fit* F : done)
    if (F->th.joinable()) F->th.join();
  return TPA_OK;
}

int tpa_server_result(tpa_server_t S, int fit_id, double* y, double* nll, int* status,
                      int64_t* nfev, int64_t* niter) {
  if (!S) return tpa_set_error(TPA_ERR_ARG, "null pointer");
  std::lock_guard<std::mutex> lk(S->mu);
  if (fit_id < 0 || fit_id >= (int)S->fits.size()) return tpa_set_error(TPA_ERR_ARG, "bad fit id");
  Fit* F = S->fits[fit_id];
  if (F->status < 0) return tpa_set_error(TPA_ERR_ARG, "fit still running");
  if (y && !F->y.empty()) std::memcpy(y, F->y.data(), sizeof(double) * S->nv);
  if (nll) *nll = F->nll;
  if (status) *status = F->status;
  if (nfev) *nfev = F->nfev;
  if (niter) *niter = F->niter;
  return TPA_OK;
}

int tpa_server_latencies(tpa_server_t S, double* us, int64_t max_n, int64_t* n) {
  if (!S || !n) return tpa_set_error(TPA_ERR_ARG, "null pointer");
  std::lock_guard<std::mutex> lk(S->mu);
  const int64_t m = std::min<int64_t>(max_n, (int64_t)S->lat_us.size());
  if (m > 0) {
    if (!us) return tpa_set_error(TPA_ERR_ARG, "null pointer");
    std::memcpy(us, S->lat_us.data(), sizeof(double) * m);
  }
  *n = (int64_t)S->lat_us.size();
  return TPA_OK;
}

int tpa_server_counters(tpa_server_t S, int64_t* requests, int64_t* batches, int64_t* events,
                        double* gpu_ms) {
  if (!S) return tpa_set_error(TPA_ERR_ARG, "null pointer");
  std::lock_guard<std::mutex> lk(S->mu);
  if (requests) *requests = S->n_requests;
  if (batches) *batches = S->n_batches;
  if (events) *events = S->n_events;
  if (gpu_ms) *gpu_ms = S->gpu_ms;
  return TPA_OK;
}

}  // extern "C"
