// The host scenarios of the serving runtime (tpa_serve.cpp) in one C++ program, linked against
// the CPU stand-in of the device pool (mock_pool.cpp), so that the runtime's threads, the
// batcher and the BFGS can run under AddressSanitizer + UndefinedBehaviorSanitizer and under
// ThreadSanitizer (tests/test_serving_sanitizers.py builds and runs each variant). The
// scenarios follow tests/test_serving_host.py: NLL and gradient against a direct restatement,
// fits alone and in a mix (bit-identical), the Hessian start, destroy with fits running, a
// rejected map, and starts outside one-sided bounds. Exit status 0 when every scenario passed.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <random>
#include <thread>
#include <vector>

#include "tfpwa_amd.h"

extern "C" {
int mock_pool_create(int P, int lam_pad, int64_t capacity, tpa_pool_t* out);
int mock_pool_destroy(tpa_pool_t p);
int mock_pool_add(tpa_pool_t p, const double* h, int64_t n, const double* w, int64_t* offset);
}

namespace {

using cplx = std::complex<double>;
constexpr int P = 6, LG = 2, NF = 11;

int g_failed = 0;

#define CHECK(cond)                                                            \
  do {                                                                         \
    if (!(cond)) {                                                             \
      std::printf("FAIL %s:%d: %s (%s)\n", __FILE__, __LINE__, #cond,          \
                  tpa_last_error());                                           \
      ++g_failed;                                                              \
      return;                                                                  \
    }                                                                          \
  } while (0)

// Synthetic partial waves: 3 chains x 2 waves, c_k = total_chain * g_k * const_k with polar
// factors; chain 0's total and g_0 are fixed. Variable 0 (a magnitude) is bounded to [0.1, 3].
struct Problem {
  std::vector<int64_t> sizes;
  int64_t n_all = 0;
  std::vector<cplx> h;     // [P][n_all][LG]
  std::vector<cplx> gram;  // [P][P]
  std::vector<int32_t> wave_factor, factor_polar, factor_var, bound_kind;
  std::vector<double> wave_const, factor_fixed, bound_ab;
  int nv = 0;
  std::vector<double> x_true;
  TpaCouplingMap map{};
  tpa_pool_t pool = nullptr;
  std::vector<std::pair<int64_t, int64_t>> slots;

  Problem(std::vector<int64_t> sz, uint64_t seed) : sizes(std::move(sz)) {
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> N(0.0, 1.0);
    for (int64_t n : sizes) n_all += n;
    h.resize((size_t)P * n_all * LG);
    for (auto& v : h) v = cplx(N(rng), N(rng));
    for (int64_t i = 0; i < n_all * LG; ++i) h[(size_t)n_all * LG + i] *= 0.4;
    const int64_t nm = 20000;
    std::vector<cplx> hm((size_t)P * nm * LG);
    for (auto& v : hm) v = cplx(N(rng), N(rng));
    gram.assign((size_t)P * P, 0.0);
    for (int k = 0; k < P; ++k)
      for (int m = 0; m < P; ++m) {
        cplx s = 0.0;
        for (int64_t r = 0; r < nm * LG; ++r)
          s += std::conj(hm[(size_t)k * nm * LG + r]) * hm[(size_t)m * nm * LG + r];
        gram[(size_t)k * P + m] = s / double(nm);
      }
    // factors: 0..2 chain totals, 3..8 wave couplings, 9 a fixed cartesian factor, 10 unit
    for (int k = 0; k < P; ++k) {
      wave_factor.push_back(k / 2);
      wave_factor.push_back(3 + k);
      wave_factor.push_back(9);
      wave_factor.push_back(10);
      wave_const.push_back(0.8 + 0.5 * k / (P - 1));
    }
    factor_polar.assign(NF, 1);
    factor_polar[9] = factor_polar[10] = 0;
    factor_var.assign(2 * NF, -1);
    for (int f : {1, 2, 4, 5, 6, 7, 8}) {
      factor_var[2 * f] = nv;
      factor_var[2 * f + 1] = nv + 1;
      nv += 2;
    }
    factor_fixed.assign(2 * NF, 0.0);
    for (int f = 0; f < NF; ++f) factor_fixed[2 * f] = 1.0;
    factor_fixed[18] = 0.9;
    factor_fixed[19] = 0.2;
    bound_kind.assign(nv, 0);
    bound_ab.assign(2 * nv, 0.0);
    bound_kind[0] = 1;
    bound_ab[0] = 0.1;
    bound_ab[1] = 3.0;
    x_true = {1.2, 0.4, 0.8, -1.0, 0.7, 0.3, 1.1, -0.5, 0.6, 2.0, 0.9, 1.4, 1.3, -2.2};
    refresh_map();
    if (mock_pool_create(P, LG, n_all, &pool) != 0) return;
    int64_t a = 0;
    for (int64_t n : sizes) {
      std::vector<double> hs((size_t)2 * P * n * LG);
      for (int k = 0; k < P; ++k)
        for (int64_t r = 0; r < n * LG; ++r) {
          const cplx v = h[(size_t)k * n_all * LG + a * LG + r];
          hs[2 * ((size_t)k * n * LG + r)] = v.real();
          hs[2 * ((size_t)k * n * LG + r) + 1] = v.imag();
        }
      std::vector<double> w(n, 1.0);
      int64_t off = -1;
      mock_pool_add(pool, hs.data(), n, w.data(), &off);
      slots.push_back({off, n});
      a += n;
    }
  }
  ~Problem() {
    if (pool) mock_pool_destroy(pool);
  }
  void refresh_map() {
    map.P = P;
    map.n_factor = NF;
    map.n_var = nv;
    map.wave_factor = wave_factor.data();
    map.wave_const = wave_const.data();
    map.factor_polar = factor_polar.data();
    map.factor_var = factor_var.data();
    map.factor_fixed = factor_fixed.data();
    map.bound_kind = bound_kind.data();
    map.bound_ab = bound_ab.data();
  }
  tpa_server_t server(double gtol = 1e-5, int maxiter = 0, int hess_start = 0) {
    tpa_server_t s = nullptr;
    if (tpa_server_create(pool, &map, reinterpret_cast<const double*>(gram.data()), gtol,
                          maxiter, 64, 200.0, hess_start, &s) != 0)
      return nullptr;
    return s;
  }
  std::vector<cplx> coefs(const std::vector<double>& y) const {
    std::vector<cplx> v(NF);
    for (int f = 0; f < NF; ++f) {
      const double r = factor_var[2 * f] >= 0 ? y[factor_var[2 * f]] : factor_fixed[2 * f];
      const double i =
          factor_var[2 * f + 1] >= 0 ? y[factor_var[2 * f + 1]] : factor_fixed[2 * f + 1];
      v[f] = factor_polar[f] ? std::polar(1.0, i) * r : cplx(r, i);
    }
    std::vector<cplx> c(P);
    for (int k = 0; k < P; ++k) {
      cplx x = wave_const[k];
      for (int a = 0; a < 4; ++a) x *= v[wave_factor[4 * k + a]];
      c[k] = x;
    }
    return c;
  }
  // -sum log f + n log(c^dagger I c) (unit weights; f >> 1e-6 here)
  double nll(std::pair<int64_t, int64_t> slot, const std::vector<double>& y) const {
    const std::vector<cplx> c = coefs(y);
    double s = 0.0;
    for (int64_t e = slot.first; e < slot.first + slot.second; ++e) {
      double f = 0.0;
      for (int l = 0; l < LG; ++l) {
        cplx A = 0.0;
        for (int k = 0; k < P; ++k) A += c[k] * h[((size_t)k * n_all + e) * LG + l];
        f += std::norm(A);
      }
      s -= std::log(f);
    }
    cplx q = 0.0;
    for (int k = 0; k < P; ++k)
      for (int m = 0; m < P; ++m) q += std::conj(c[k]) * gram[(size_t)k * P + m] * c[m];
    return s + double(slot.second) * std::log(q.real());
  }
};

bool nll_grad(tpa_server_t s, std::pair<int64_t, int64_t> slot, const std::vector<double>& y,
              double* f, std::vector<double>* g) {
  g->assign(y.size(), 0.0);
  return tpa_server_nll_grad(s, slot.first, slot.second, double(slot.second), y.data(), f,
                             g->data()) == 0;
}

std::vector<int> wait_for(tpa_server_t s, size_t n) {
  std::vector<int> got;
  int ids[64], nd = 0;
  const auto t0 = std::chrono::steady_clock::now();
  while (got.size() < n) {
    if (tpa_server_poll(s, ids, 64, &nd) != 0) break;
    got.insert(got.end(), ids, ids + nd);
    if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(120)) break;
    std::this_thread::sleep_for(std::chrono::milliseconds(1));
  }
  return got;
}

struct Result {
  std::vector<double> y;
  double f = NAN;
  int status = -1;
  int64_t nfev = 0, niter = 0;
};

Result result(tpa_server_t s, int fid, int nv) {
  Result r;
  r.y.assign(nv, 0.0);
  tpa_server_result(s, fid, r.y.data(), &r.f, &r.status, &r.nfev, &r.niter);
  return r;
}

void scenario_nll_and_gradient() {
  Problem pb({1500, 900}, 1);
  tpa_server_t s = pb.server();
  CHECK(s != nullptr);
  std::mt19937_64 rng(3);
  std::normal_distribution<double> N(0.0, 0.2);
  for (int rep = 0; rep < 2; ++rep) {
    std::vector<double> y = pb.x_true;
    for (auto& v : y) v += N(rng);
    for (auto slot : pb.slots) {
      double f = 0.0;
      std::vector<double> g;
      CHECK(nll_grad(s, slot, y, &f, &g));
      const double ref = pb.nll(slot, y);
      CHECK(std::fabs(f - ref) <= 1e-10 * std::fabs(ref));
      double gmax = 0.0;
      for (double v : g) gmax = std::max(gmax, std::fabs(v));
      for (int j = 0; j < pb.nv; ++j) {
        std::vector<double> yp = y, ym = y;
        yp[j] += 1e-6;
        ym[j] -= 1e-6;
        const double fd = (pb.nll(slot, yp) - pb.nll(slot, ym)) / 2e-6;
        CHECK(std::fabs(fd - g[j]) <= 1e-5 * std::max(1.0, gmax));
      }
    }
  }
  CHECK(tpa_server_destroy(s) == 0);
  std::printf("PASS nll_and_gradient\n");
}

void scenario_fits_alone_and_in_the_mix(int hess_start) {
  Problem pb({2000, 1200, 1700, 800}, 2);
  std::mt19937_64 rng(5);
  std::normal_distribution<double> N(0.0, 0.1);
  std::vector<std::vector<double>> starts;
  for (size_t i = 0; i < pb.slots.size(); ++i) {
    std::vector<double> y = pb.x_true;
    for (auto& v : y) v += N(rng);
    starts.push_back(y);
  }
  tpa_server_t s = pb.server(1e-5, 0, hess_start);
  CHECK(s != nullptr);
  int fid = -1;
  auto slot0 = pb.slots[0];
  CHECK(tpa_server_submit(s, slot0.first, slot0.second, double(slot0.second), starts[0].data(),
                          &fid) == 0);
  CHECK(wait_for(s, 1).size() == 1);
  const Result alone = result(s, fid, pb.nv);
  std::vector<int> ids;
  for (size_t i = 0; i < pb.slots.size(); ++i) {
    auto sl = pb.slots[i];
    CHECK(tpa_server_submit(s, sl.first, sl.second, double(sl.second), starts[i].data(), &fid) ==
          0);
    ids.push_back(fid);
  }
  CHECK(wait_for(s, ids.size()).size() == ids.size());
  const Result mix = result(s, ids[0], pb.nv);
  CHECK(alone.y == mix.y && alone.f == mix.f && alone.status == mix.status &&
        alone.nfev == mix.nfev && alone.niter == mix.niter);
  for (size_t i = 0; i < ids.size(); ++i) {
    const Result r = result(s, ids[i], pb.nv);
    CHECK(r.status == 0 || r.status == 2);
    CHECK(r.f < pb.nll(pb.slots[i], starts[i]));
    CHECK(std::fabs(r.f - pb.nll(pb.slots[i], r.y)) <= 1e-9 * std::fabs(r.f));
    double f = 0.0;
    std::vector<double> g;
    CHECK(nll_grad(s, pb.slots[i], r.y, &f, &g));
    for (double v : g) CHECK(std::fabs(v) < 1e-3);
    CHECK(r.y[0] >= 0.1 && r.y[0] <= 3.0);
  }
  int64_t req = 0, bat = 0, ev = 0;
  CHECK(tpa_server_counters(s, &req, &bat, &ev, nullptr) == 0);
  CHECK(bat < req && ev > req);
  CHECK(tpa_server_destroy(s) == 0);
  std::printf("PASS fits_alone_and_in_the_mix hess_start=%d\n", hess_start);
}

void scenario_destroy_cancels_running_fits() {
  Problem pb({3000}, 3);
  tpa_server_t s = pb.server(1e-300, 100000, 0);  // would run for a long time
  CHECK(s != nullptr);
  std::vector<double> y0 = pb.x_true;
  for (auto& v : y0) v += 0.3;
  int fid = -1;
  for (int i = 0; i < 4; ++i) CHECK(tpa_server_submit(s, 0, 3000, 3000.0, y0.data(), &fid) == 0);
  std::this_thread::sleep_for(std::chrono::milliseconds(50));
  const auto t0 = std::chrono::steady_clock::now();
  CHECK(tpa_server_destroy(s) == 0);
  CHECK(std::chrono::steady_clock::now() - t0 < std::chrono::seconds(30));
  std::printf("PASS destroy_cancels_running_fits\n");
}

void scenario_rejects_bad_maps() {
  Problem pb({100}, 4);
  pb.wave_factor[0] = 99;
  pb.refresh_map();
  CHECK(pb.server() == nullptr);
  std::printf("PASS rejects_bad_maps\n");
}

void scenario_starts_outside_one_sided_bounds() {
  // the magnitude of chain 2's total (variable 2) bounded below at 0.5 and that of g_1
  // (variable 4) above at 0.5; the fit starts below and above them: the start maps to the
  // bound's edge and the fit runs on finite values inside the bounds
  Problem pb({2500}, 6);
  pb.bound_kind[2] = 2;
  pb.bound_ab[4] = 0.5;
  pb.bound_kind[4] = 3;
  pb.bound_ab[9] = 0.5;
  pb.refresh_map();
  tpa_server_t s = pb.server();
  CHECK(s != nullptr);
  std::vector<double> y0 = pb.x_true;
  y0[2] = 0.2;
  y0[4] = 0.9;
  int fid = -1;
  CHECK(tpa_server_submit(s, 0, 2500, 2500.0, y0.data(), &fid) == 0);
  CHECK(wait_for(s, 1).size() == 1);
  const Result r = result(s, fid, pb.nv);
  CHECK(std::isfinite(r.f));
  for (double v : r.y) CHECK(std::isfinite(v));
  CHECK(r.y[2] >= 0.5 - 1e-12 && r.y[4] <= 0.5 + 1e-12);
  CHECK(r.status == 0 || r.status == 2);
  CHECK(tpa_server_destroy(s) == 0);
  std::printf("PASS starts_outside_one_sided_bounds\n");
}

}  // namespace

int main() {
  scenario_nll_and_gradient();
  scenario_fits_alone_and_in_the_mix(0);
  scenario_fits_alone_and_in_the_mix(1);
  scenario_destroy_cancels_running_fits();
  scenario_rejects_bad_maps();
  scenario_starts_outside_one_sided_bounds();
  std::printf("%s (%d failed)\n", g_failed ? "FAILED" : "ALL PASSED", g_failed);
  return g_failed ? 1 : 0;
}
