// C ABI of the engine (declared in include/tfpwa_amd.h): device buffers, launches and the
// kernel dispatcher. No C++ exception crosses the boundary; errors are status codes plus a
// per-thread message.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include <hip/hiprtc.h>

#include "tfpwa_amd.h"
#include "tpa_fused_gen.h"
#include "tpa_internal.h"
#include "tpa_kernels.h"
#include "tpa_kernels4.h"
#include "tpa_half.h"
#include "tpa_rtc_headers.inc"  // kRtcHeaders / kRtcHeaderNames (written by build.py)

// Pass timing: armed by a timed evaluation, the next pass launched through launch_pass records
// its start and end events from its own dispatch (hipExtLaunchKernelGGL), so the measured span
// is the kernel's execution, as a kernel trace reports it, without the queue's dispatch latency.
struct PassTimer {
  hipEvent_t t0 = nullptr, t1 = nullptr;
};
static thread_local PassTimer* tl_pass_timer = nullptr;

template <typename F, typename... Args>
static void launch_pass(F kernel, dim3 grid, dim3 block, hipStream_t s, Args... args) {
  if (PassTimer* t = tl_pass_timer) {
    tl_pass_timer = nullptr;
    hipExtLaunchKernelGGL(kernel, grid, block, 0, s, t->t0, t->t1, 0, args...);
  } else {
    hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
  }
}

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                 \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess)                                                         \
      return fail(TPA_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));   \
  } while (0)

std::mutex g_timing_mu;
bool g_timing = false;
double g_timing_ms = 0.0;
int64_t g_timing_n = 0;
int64_t g_timing_every = 1;  // time one launch in this many (tpa_timing_enable's argument)
int64_t g_timing_tick = 0;

// Whether the launch about to be made is timed: timing on, and this launch the first of its
// sampling period. The caller holds g_timing_mu.
bool timing_take_locked() {
  if (!g_timing) return false;
  return (g_timing_tick++ % g_timing_every) == 0;
}

constexpr int kEvalGridMax = 8192;     // partial-sum slots per event set
constexpr int kEvalGridDefault = 1024;

// small_d_weight (tf_pwa/dfun.py:118-143) for 2J = 0..TPA_MAX_J2, tables concatenated.
std::vector<double> wigner_weights() {
  std::vector<double> out;
  auto f = [](int x) {
    double r = 1.0;
    for (int i = 2; i <= (x >> 1); ++i) r *= i;
    return r;
  };
  for (int j = 0; j <= TPA_MAX_J2; ++j) {
    std::vector<double> w((j + 1) * (j + 1) * (j + 1), 0.0);
    for (int m = -j; m <= j; m += 2)
      for (int n = -j; n <= j; n += 2)
        for (int k = std::max(0, n - m); k <= std::min(j - m, j + n); k += 2) {
          int l = (2 * k + (m - n)) / 2;
          double sign = (((k + m - n) / 2) % 2 == 0) ? 1.0 : -1.0;
          double tmp = sign * std::sqrt(1.0 * f(j + m) * f(j - m) * f(j + n) * f(j - n));
          tmp /= f(j - m - k) * f(j + n - k) * f(k + m - n) * f(k);
          w[(l * (j + 1) + (m + j) / 2) * (j + 1) + (n + j) / 2] = tmp;
        }
    out.insert(out.end(), w.begin(), w.end());
  }
  return out;
}

bool valid_hel(int j2, int h) { return h >= -j2 && h <= j2 && ((h + j2) % 2 == 0); }

std::string check_model(const TpaModel& M, int n_cg) {
  if (M.n_final != 3) return "only three-body final states are supported";
  if (M.n_topo < 1 || M.n_topo > TPA_MAX_TOPO) return "n_topo out of range";
  if (M.n_chain < 1 || M.n_chain > TPA_MAX_CHAIN) return "n_chain out of range";
  if (M.P < 1 || M.P > TPA_MAX_P) return "P out of range";
  if (M.top_j2 < 0 || M.top_j2 > TPA_MAX_J2) return "top spin out of range";
  if (M.n_top < 1 || M.n_top > TPA_MAX_HEL) return "top helicity count out of range";
  for (int i = 0; i < M.n_top; ++i)
    if (!valid_hel(M.top_j2, M.hel_top[i])) return "bad top helicity";
  int lam = M.n_top;
  for (int f = 0; f < 3; ++f) {
    if (M.final_j2[f] < 0 || M.final_j2[f] > 2) return "final spin above 1 unsupported";
    if (M.n_fh[f] < 1 || M.n_fh[f] > 3) return "final helicity count out of range";
    for (int i = 0; i < M.n_fh[f]; ++i)
      if (!valid_hel(M.final_j2[f], M.hel_f[f][i])) return "bad final helicity";
    if (M.ref_topo[f] < 0 || M.ref_topo[f] >= M.n_topo) return "bad reference topology";
    lam *= M.n_fh[f];
  }
  if (lam != M.lam) return "lam does not match the helicity counts";
  if (M.lam_pad < M.lam || M.lam_pad > TPA_MAX_LAM || (M.lam_pad & (M.lam_pad - 1)))
    return "lam_pad must be a power of two >= lam";
  for (int t = 0; t < M.n_topo; ++t) {
    const TpaTopo& T = M.topo[t];
    int s[3] = {T.x, T.y, T.z};
    for (int i = 0; i < 3; ++i)
      if (s[i] < 0 || s[i] > 2) return "bad topology final index";
    if (T.x == T.y || T.x == T.z || T.y == T.z) return "topology finals must differ";
  }
  int next_k = 0;
  for (int c = 0; c < M.n_chain; ++c) {
    const TpaChain& C = M.chain[c];
    if (C.topo < 0 || C.topo >= M.n_topo) return "chain topology out of range";
    if (C.jr2 < 0 || C.jr2 > TPA_MAX_J2 || C.n_r < 1 || C.n_r > C.jr2 + 1)
      return "bad resonance spin";
    for (int i = 0; i < C.n_r; ++i)
      if (!valid_hel(C.jr2, C.hel_r[i])) return "bad resonance helicity";
    if (C.n_ls1 < 1 || C.n_ls1 > TPA_MAX_LS || C.n_ls2 < 1 || C.n_ls2 > TPA_MAX_LS)
      return "LS count out of range";
    for (int s2 = 0; s2 < C.n_ls1; ++s2)
      if (C.l1[s2] < 0 || C.l1[s2] > 5) return "l out of range";
    for (int s2 = 0; s2 < C.n_ls2; ++s2)
      if (C.l2[s2] < 0 || C.l2[s2] > 5) return "l out of range";
    if (C.k0 != next_k) return "chains must tile the partial-wave index range";
    next_k += C.n_ls1 * C.n_ls2;
    if (C.shape < 0 || C.shape > 7 || (C.shape & 3) == 3 || C.bw_l < 0 || C.bw_l > 5)
      return "bad lineshape";
    const TpaTopo& T = M.topo[C.topo];
    int nX = M.n_fh[T.x];
    int n1 = C.n_ls1 * C.n_r * nX;
    int fy = C.r_y_first ? T.y : T.z, fz = C.r_y_first ? T.z : T.y;
    int n2 = C.n_ls2 * M.n_fh[fy] * M.n_fh[fz];
    if (C.cg1_off < 0 || C.cg1_off + n1 > n_cg || C.cg2_off < 0 || C.cg2_off + n2 > n_cg)
      return "CG table offsets out of range";
  }
  if (next_k != M.P) return "sum of chain partial waves differs from P";
  if (n_cg < 0 || n_cg > TPA_MAX_CG || M.n_cg != n_cg) return "bad CG table size";
  return "";
}

std::string check_model4(const TpaModel4& M, int n_cg) {
  if (M.n_final != 4) return "four-body description needs n_final = 4";
  if (M.n_topo < 1 || M.n_topo > TPA4_MAX_TOPO) return "n_topo out of range";
  if (M.n_chain < 1 || M.n_chain > TPA4_MAX_CHAIN) return "n_chain out of range";
  if (M.P < 1 || M.P > TPA_MAX_P) return "P out of range";
  if (M.top_j2 < 0 || M.top_j2 > TPA_MAX_J2) return "top spin out of range";
  if (M.n_top < 1 || M.n_top > TPA_MAX_HEL) return "top helicities out of range";
  if (M.lam != M.n_top) return "four-body finals must be spin 0 (lam = top helicities)";
  if (M.lam_pad < M.lam || M.lam_pad > 16 || (M.lam_pad & (M.lam_pad - 1))) return "bad lam_pad";
  if (M.n_cg != n_cg) return "n_cg mismatch";
  for (int i = 0; i < M.n_top; ++i)
    if (std::abs(M.hel_top[i]) > M.top_j2 || ((M.hel_top[i] - M.top_j2) & 1))
      return "bad top helicity";
  for (int t = 0; t < M.n_topo; ++t) {
    const Tpa4Topo& T = M.topo[t];
    if (T.core[0] != 15 || T.parent[0] != -1) return "vertex 0 must be the top decay";
    for (int v = 0; v < 3; ++v) {
      if ((T.out[v][0] | T.out[v][1]) != T.core[v] || (T.out[v][0] & T.out[v][1]) ||
          !T.out[v][0] || !T.out[v][1])
        return "vertex outs must split the core's finals";
      if (v > 0) {
        if (T.parent[v] < 0 || T.parent[v] >= v || T.parent_out[v] < 0 || T.parent_out[v] > 1)
          return "bad vertex parent";
        if (T.out[T.parent[v]][T.parent_out[v]] != T.core[v]) return "parent out is not the core";
      }
    }
  }
  int covered = 0;
  for (int c = 0; c < M.n_chain; ++c) {
    const Tpa4Chain& C = M.chain[c];
    if (C.topo < 0 || C.topo >= M.n_topo) return "chain topology out of range";
    int np = 1;
    for (int v = 0; v < 3; ++v) {
      if (C.n_ls[v] < 1 || C.n_ls[v] > TPA_MAX_LS) return "n_ls out of range";
      for (int s = 0; s < C.n_ls[v]; ++s)
        if (C.l[v][s] < 0 || C.l[v][s] > 5) return "l out of range";
      np *= C.n_ls[v];
      for (int o = 0; o < 2; ++o)
        if (C.out_slot[v][o] < 1 || C.out_slot[v][o] > 6 || C.out_slot[v][o] == v)
          return "bad out slot";
    }
    for (int r = 0; r < 2; ++r) {
      if (C.r_j2[r] < 0 || C.r_j2[r] > TPA_MAX_J2) return "resonance spin out of range";
      if (C.r_n[r] < 1 || C.r_n[r] > TPA_MAX_HEL) return "resonance helicities out of range";
      for (int i = 0; i < C.r_n[r]; ++i)
        if (std::abs(C.r_hel[r][i]) > C.r_j2[r] || ((C.r_hel[r][i] - C.r_j2[r]) & 1))
          return "bad resonance helicity";
      if (C.r_shape[r] < 0 || C.r_shape[r] > 7 || (C.r_shape[r] & 3) == 3 || C.r_bw_l[r] < 0 ||
          C.r_bw_l[r] > 5)
        return "bad lineshape";
    }
    auto nh = [&](int slot) { return slot == 0 ? M.n_top : slot <= 2 ? C.r_n[slot - 1] : 1; };
    for (int v = 0; v < 3; ++v) {
      const int sz = C.n_ls[v] * nh(C.out_slot[v][0]) * nh(C.out_slot[v][1]);
      if (C.cg_off[v] < 0 || C.cg_off[v] + sz > n_cg) return "CG offset out of range";
    }
    if (C.k0 != covered) return "chains must cover the partial waves in order";
    covered += np;
  }
  if (covered != M.P) return "partial-wave count mismatch";
  return "";
}

}  // namespace

struct tpa_model_s {
  int device = 0;
  int kind = 3;        // 3: three-body chains (TpaModel); 4: four-body trees (TpaModel4)
  TpaModel h{};        // for kind 4 only the header fields (sizes, helicities) are set
  TpaModel* d = nullptr;
  TpaModel4 h4{};
  TpaModel4* d4 = nullptr;
  double* d_cg = nullptr;
  double* d_dw = nullptr;
  int* d_kchain = nullptr;
  hipStream_t stream = nullptr;
  TpaModel* h_pin = nullptr;  // pinned staging of tpa_model_set_lineshapes
  std::vector<double> cg_host;  // the CG table (the fused pass's generator bakes it in)
  // the fused pass (tpa_eval_fused): compiled on first use, regenerated when the description
  // changes (kept if its generated source is the same)
  std::mutex fused_mu;
  std::string fused_src;
  hipModule_t fused_mod = nullptr;
  hipFunction_t fused_fn = nullptr;
  bool fused_stale = false;
  double fused_compile_s = 0.0;
};

struct tpa_events_s {
  tpa_model_s* m = nullptr;
  int64_t n = 0;
  int nf = 0;
  double* d_ang = nullptr;
  double2* d_h = nullptr;
  double* d_w = nullptr;
  double* d_part = nullptr;
  double* d_out = nullptr;
  double2* d_coef = nullptr;
  double* d_f = nullptr;
  double* d_gpart = nullptr;
  double* d_gout = nullptr;
  double2* d_bw = nullptr;
  double* d_hpart = nullptr;  // second-moment pass: partials and result
  double* d_hout = nullptr;
  double2* d_hcoef = nullptr;
  double2* d_bw2 = nullptr;    // lineshapes with second derivatives (tpa_hess_shape)
  double* d_hspart = nullptr;  // its pass's partials and result
  double* d_hsout = nullptr;
  int built_fold = -1;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  double* h_io = nullptr;  // pinned staging of the coefficients and the sums of tpa_eval
  float2* d_h32 = nullptr;  // FP32 copy of a folded cache (tpa_events_keep_f32)
  bool keep32 = false;      // refresh d_h32 after every folded build
  bool h32_valid = false;
  // helicity-halved FP32 copy (tpa_events_keep_half): stored half [P][n][nh], chain phases
  // [n_chain][n], the phi_k of each wave (1: -i) and the chain of each wave
  float2* d_g32h = nullptr;
  float2* d_rho = nullptr;
  double2* d_g64h = nullptr;  // the same layout in FP64 (TPA_FLAG_F64H)
  double2* d_rho64 = nullptr;
  unsigned long long* d_resid = nullptr;
  // per-call rebuild (tpa_eval_recompute): one chunk of waves and the chunks' sums
  double2* d_rchunk = nullptr;
  int64_t rchunk_cap = 0;
  double* d_rout = nullptr;
  int64_t rout_cap = 0;
  bool keep32h = false;
  bool h32h_valid = false;
  bool keep64h = false;
  bool h64h_valid = false;
  int half_nh = 0;
  int half_slot[TPA_HALF_MAX_NH] = {}, half_mirror[TPA_HALF_MAX_NH] = {},
      half_sgn[TPA_HALF_MAX_NH] = {};
  double half_resid = -1.0;
  std::vector<unsigned char> half_rot;
  std::vector<signed char> half_kch;
  // lane plan of the halved FP32 pass (k_eval_half_chain; half_plan_pml 0: none fits)
  int half_plan_pml = 0;
  signed char half_plan[TPA_EVAL_CIN];
  signed char half_pch[8];
  unsigned* d_cnt = nullptr;  // k_reduce's finishing-block counter (zero between calls)
  uint64_t* h_flag = nullptr; // pinned: the sequence number of the last completed reduction
  uint64_t seq = 0;
  int pending_nout = 0;                 // a queued evaluation (tpa_eval_enqueue) not collected
  hipStream_t pending_stream = nullptr;
};

// Pinned staging of tpa_eval: 2 * TPA_MAX_P doubles of coefficients, then the sums (at most
// 2 + 6 * TPA_MAX_P doubles). Copies from and to pageable caller memory would go through the
// runtime's own staging on every call.
#define TPA_IO_DOUBLES (2 * TPA_MAX_P + 2 + 6 * TPA_MAX_P)

// --------------------------------------------------------------------------- dispatcher
// Partial waves are split over KG lane groups so that each lane holds at most 8 of them.
// The default evaluation kernel is k_eval_pipe with PML = ceil(P / KG) rounded up to 4, 6, 7
// or 8 (the register stages and accumulators of unused slots cost occupancy; up to 6 it runs
// three waves per SIMD): on the MI355X it streams 6.5-6.7 TB/s on 1M and 10M events, against
// 5.0-5.3 TB/s for k_eval_fold<.., U = 2>.
// TFPWA_EVAL_UNROLL selects k_eval_fold instead, for comparison: 1 (U = 1), 2 (U = 2),
// 3 (U = 2 with ordinary loads).
static int eval_unroll() {
  static int u = [] {
    const char* v = getenv("TFPWA_EVAL_UNROLL");
    int x = v ? atoi(v) : 0;
    return (x >= 1 && x <= 3) ? x : 0;
  }();
  return u;
}

template <int LG, int KG, typename HT>
static void launch_pipe(const EvalArgs& a, int grid, hipStream_t s) {
  const int need = (a.P + KG - 1) / KG;
  if (need <= 4)
    launch_pass((k_eval_pipe<LG, KG, 4, HT>), dim3(grid), dim3(256), s, a);
  else if (need <= 6)
    launch_pass((k_eval_pipe<LG, KG, 6, HT>), dim3(grid), dim3(256), s, a);
  else if (need <= 7)
    launch_pass((k_eval_pipe<LG, KG, 7, HT>), dim3(grid), dim3(256), s, a);
  else
    launch_pass((k_eval_pipe<LG, KG, 8, HT>), dim3(grid), dim3(256), s, a);
}

template <int LG, int KG>
static void launch_pipe32(const EvalArgs& a, int grid, hipStream_t s) {
  if constexpr (LG >= 2) {
    const int need = (a.P + KG - 1) / KG;
    if (need <= 4)
      launch_pass((k_eval_pipe32<LG, KG, 4>), dim3(grid), dim3(256), s, a);
    else if (need <= 6)
      launch_pass((k_eval_pipe32<LG, KG, 6>), dim3(grid), dim3(256), s, a);
    else if (need <= 7)
      launch_pass((k_eval_pipe32<LG, KG, 7>), dim3(grid), dim3(256), s, a);
    else
      launch_pass((k_eval_pipe32<LG, KG, 8>), dim3(grid), dim3(256), s, a);
  } else {
    launch_pipe<LG, KG, float2>(a, grid, s);
  }
}

// The deferred-value pass over the full FP32 copy (k_eval_pipe32d): two helicity slots, four
// k-groups (17 to 32 waves: the four-body models), every buffer offset below 2^31 bytes; its
// grid from the kernel's occupancy. Elsewhere k_eval_pipe32.
static bool pipe32d_fits(const EvalArgs& a, int lg, int kg) {
  if (lg != 2 || kg != 4 || a.P <= 16 || a.P > 32) return false;
  const uint64_t past = 2ull * 64 * kEvalGridMax;
  return ((uint64_t)a.P * (uint64_t)a.N + past) * 16 < (1ull << 31);
}

template <int PML>
static int pipe32d_grid_t(int64_t n) {
  static const int gmax = [] {
    const char* e = std::getenv("TFPWA_EVAL_GRID");
    if (e) {
      const int v = std::atoi(e);
      if (v >= 64 && v <= kEvalGridMax) return v;
    }
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(&k_eval_pipe32d<4, PML>), 256, 0) !=
            hipSuccess ||
        cus < 1 || per_cu < 1)
      return kEvalGridDefault;
    return std::min(kEvalGridMax, std::max(64, per_cu * cus));
  }();
  int64_t g = (n + 63) / 64;
  if (g < 1) g = 1;
  return (int)std::min<int64_t>(g, gmax);
}

static int pipe32d_pml(int P) { return P <= 24 ? 6 : P <= 28 ? 7 : 8; }

static int pipe32d_grid(int64_t n, int P) {
  const int pml = pipe32d_pml(P);
  return pml == 6 ? pipe32d_grid_t<6>(n) : pml == 7 ? pipe32d_grid_t<7>(n) : pipe32d_grid_t<8>(n);
}

static void launch_pipe32d(const EvalArgs& a, int grid, hipStream_t s) {
  const uint32_t hb = (uint32_t)((uint64_t)a.P * a.N * 16);
  const int pml = pipe32d_pml(a.P);
  if (pml == 6)
    launch_pass((k_eval_pipe32d<4, 6>), dim3(grid), dim3(256), s, a, hb);
  else if (pml == 7)
    launch_pass((k_eval_pipe32d<4, 7>), dim3(grid), dim3(256), s, a, hb);
  else
    launch_pass((k_eval_pipe32d<4, 8>), dim3(grid), dim3(256), s, a, hb);
}

// Lanes per event of the FP32-cache pass: two helicity components per lane when LG >= 2.
static int f32_lanes_per_event(int lg, int kg) { return (lg >= 2 ? lg / 2 : 1) * kg; }

// a.h32 set: the pass reads the FP32 copy (k_eval_pipe32; k_eval_pipe over float2 for LG 1).
template <int LG, int KG>
static void launch_eval(const EvalArgs& a, int grid, hipStream_t s) {
  if (a.h32 != nullptr) {
    launch_pipe32<LG, KG>(a, grid, s);
    return;
  }
  switch (eval_unroll()) {
    case 0:
      launch_pipe<LG, KG, double2>(a, grid, s);
      break;
    case 1:
      launch_pass((k_eval_fold<LG, KG, 8, 1, true>), dim3(grid), dim3(256), s, a);
      break;
    case 3:  // U = 2 with ordinary (cached) loads, for comparison
      launch_pass((k_eval_fold<LG, KG, 8, 2, false>), dim3(grid), dim3(256), s, a);
      break;
    default:
      launch_pass((k_eval_fold<LG, KG, 8, 2, true>), dim3(grid), dim3(256), s, a);
      break;
  }
}

// k-groups of the evaluation passes: the fewest that leave at most 8 waves per lane, with an
// event's lanes in one wave. (More k-groups for fewer waves per lane raise the occupancy but
// cut the loads each wave keeps in flight: at P = 25, two helicities, 8 groups of 4 waves ran
// 13% slower than 4 groups of 7.)
static int eval_kgroups(int lg, int P) {
  int kg = 1;
  while (kg * 8 < P) kg <<= 1;
  if (lg * kg > 64) return -1;
  return kg;
}

template <int LG>
static int dispatch_kg(int kg, const EvalArgs& a, int grid, hipStream_t s) {
  switch (kg) {
    case 1: launch_eval<LG, 1>(a, grid, s); return TPA_OK;
    case 2: launch_eval<LG, 2>(a, grid, s); return TPA_OK;
    case 4: launch_eval<LG, 4>(a, grid, s); return TPA_OK;
    case 8: if (LG * 8 <= 64) { launch_eval<LG, (LG * 8 <= 64 ? 8 : 1)>(a, grid, s); return TPA_OK; }
            return TPA_ERR_UNSUPPORTED;
    case 16: if (LG * 16 <= 64) { launch_eval<LG, (LG * 16 <= 64 ? 16 : 1)>(a, grid, s); return TPA_OK; }
             return TPA_ERR_UNSUPPORTED;
    default: return TPA_ERR_UNSUPPORTED;
  }
}

// The helicity-halved FP32 pass (k_eval_half<.., float>): KG lanes per event, PML =
// ceil(P / KG) rounded up to 4, 6 or 8 (P <= 32: the coefficients travel in the kernel
// arguments).
static int half_kgroups(int P) { return P <= 8 ? 1 : P <= 16 ? 2 : 4; }

template <int NH, int KG>
static void launch_half(const EvalArgs& a, int grid, hipStream_t s) {
  const int need = (a.P + KG - 1) / KG;
  if (need <= 4)
    launch_pass((k_eval_half<NH, KG, 4, float>), dim3(grid), dim3(256), s, a);
  else if (need <= 6)
    launch_pass((k_eval_half<NH, KG, 6, float>), dim3(grid), dim3(256), s, a);
  else
    launch_pass((k_eval_half<NH, KG, 8, float>), dim3(grid), dim3(256), s, a);
}

// The helicity-halved FP64 pass (k_eval_half<.., double>): a lane's register stage holds twice
// the bytes of the FP32 pass per wave, so the waves split over twice the k-groups (at most 4
// per lane; 3 at P = 17 .. 24).
static int half64_kgroups(int P) { return P <= 8 ? 2 : P <= 16 ? 4 : 8; }

template <int NH>
static void launch_half64(const EvalArgs& a, int grid, hipStream_t s) {
  if (a.P <= 8)
    launch_pass((k_eval_half<NH, 2, 4, double>), dim3(grid), dim3(256), s, a);
  else if (a.P <= 16)
    launch_pass((k_eval_half<NH, 4, 4, double>), dim3(grid), dim3(256), s, a);
  else if (a.P <= 24)
    launch_pass((k_eval_half<NH, 8, 3, double>), dim3(grid), dim3(256), s, a);
  else
    launch_pass((k_eval_half<NH, 8, 4, double>), dim3(grid), dim3(256), s, a);
}

static int dispatch_half64(int nh, const EvalArgs& a, int grid, hipStream_t s) {
  if (a.P > 32) return TPA_ERR_UNSUPPORTED;
  if (nh == 1) {
    launch_half64<1>(a, grid, s);
    return TPA_OK;
  }
  if (nh == 2) {
    launch_half64<2>(a, grid, s);
    return TPA_OK;
  }
  return TPA_ERR_UNSUPPORTED;
}

// The chain-grouped FP32 halved pass (k_eval_half_chain): two stored slots, four k-groups,
// every buffer offset below 2^31 bytes (the waves' rows up to KG * PML and the chain phases,
// each up to two grid strides past the last event), and a lane plan (made with the copy).
// Elsewhere k_eval_half.
static bool half_chain_fits(const EvalArgs& a, int nh, int kg, int n_chain) {
  if (nh != 2 || kg != 4 || a.P > 32) return false;
  // the loads of the group after the last one reach up to two grid strides past N
  const uint64_t N = (uint64_t)a.N, lim = 1ull << 31, past = 2ull * 64 * kEvalGridMax;
  return ((uint64_t)(4 * 8 + 1) * N + past) * 16 < lim && ((uint64_t)n_chain * N + past) * 8 < lim;
}

// Lane plan of k_eval_half_chain<4, PML, 4>: each of the four lanes of an event holds two
// runs of slots, [0, 4) and [4, PML), each run the waves of one chain (or pad). A chain may
// span runs of several lanes. Greedy over the chains from the largest: a free run of exactly
// the chain's remaining size, else the largest free run it fills, else the smallest free run
// that takes the rest. False when the chains do not fit.
static bool plan_half_lanes(int P, const signed char* kch, int pml, signed char* plan,
                            signed char* pch) {
  constexpr int KG = 4, J0 = 4;
  if (pml <= J0 || KG * pml > TPA_EVAL_CIN) return false;
  std::vector<std::vector<int>> waves;
  for (int k = 0; k < P; ++k) {
    const int c = kch[k];
    if (c < 0) return false;
    if ((int)waves.size() <= c) waves.resize(c + 1);
    waves[c].push_back(k);
  }
  std::vector<int> order;
  for (int c = 0; c < (int)waves.size(); ++c)
    if (!waves[c].empty()) order.push_back(c);
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return waves[x].size() > waves[y].size(); });
  // runs r = 2 kg + q: start q ? J0 : 0, size q ? pml - J0 : J0
  bool used[2 * KG] = {false};
  for (int i = 0; i < KG * pml; ++i) plan[i] = -1;
  for (int i = 0; i < 2 * KG; ++i) pch[i] = -1;
  for (int c : order) {
    size_t done = 0;
    while (done < waves[c].size()) {
      const int rem = (int)(waves[c].size() - done);
      int pick = -1, best = -1;
      for (int r = 0; r < 2 * KG; ++r)  // exact fit
        if (!used[r] && (r & 1 ? pml - J0 : J0) == rem) { pick = r; break; }
      if (pick < 0)
        for (int r = 0; r < 2 * KG; ++r) {  // largest run the chain fills
          const int sz = r & 1 ? pml - J0 : J0;
          if (!used[r] && sz <= rem && sz > best) { pick = r; best = sz; }
        }
      if (pick < 0) {
        best = 1 << 30;
        for (int r = 0; r < 2 * KG; ++r) {  // smallest run that takes the rest
          const int sz = r & 1 ? pml - J0 : J0;
          if (!used[r] && sz >= rem && sz < best) { pick = r; best = sz; }
        }
      }
      if (pick < 0) return false;
      used[pick] = true;
      const int kg = pick >> 1, q = pick & 1;
      const int start = q ? J0 : 0, sz = q ? pml - J0 : J0;
      pch[kg * 2 + q] = (signed char)c;
      for (int j = 0; j < sz && done < waves[c].size(); ++j, ++done)
        plan[kg * pml + start + j] = (signed char)waves[c][done];
    }
  }
  return true;
}

// Blocks of the chain-grouped pass: as many as are resident at once on the device (the
// kernel's occupancy, from the runtime) up to one per 64 events; TFPWA_EVAL_GRID overrides.
template <int PML>
static int half_chain_grid_t(int64_t n) {
  static const int gmax = [] {
    const char* e = std::getenv("TFPWA_EVAL_GRID");
    if (e) {
      const int v = std::atoi(e);
      if (v >= 64 && v <= kEvalGridMax) return v;
    }
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(&k_eval_half_chain<4, PML, 4>), 256, 0) !=
            hipSuccess ||
        cus < 1 || per_cu < 1)
      return kEvalGridDefault;
    return std::min(kEvalGridMax, std::max(64, per_cu * cus));
  }();
  int64_t g = (n + 63) / 64;
  if (g < 1) g = 1;
  return (int)std::min<int64_t>(g, gmax);
}

static int half_chain_grid(int64_t n, int pml) {
  return pml == 6 ? half_chain_grid_t<6>(n) : half_chain_grid_t<8>(n);
}

static void launch_half_chain(const EvalArgs& a, int pml, int n_chain, int grid,
                              hipStream_t s) {
  const uint32_t hb = (uint32_t)((uint64_t)a.P * a.N * 16);
  const uint32_t rb = (uint32_t)((uint64_t)n_chain * a.N * 8);
  if (pml == 6)
    launch_pass((k_eval_half_chain<4, 6, 4>), dim3(grid), dim3(256), s, a, hb, rb);
  else
    launch_pass((k_eval_half_chain<4, 8, 4>), dim3(grid), dim3(256), s, a, hb, rb);
}

static int dispatch_half(int nh, int kg, const EvalArgs& a, int grid, hipStream_t s) {
  if (a.P > 4 * 8) return TPA_ERR_UNSUPPORTED;
  if (nh == 1) {
    if (kg == 1) launch_half<1, 1>(a, grid, s);
    else if (kg == 2) launch_half<1, 2>(a, grid, s);
    else launch_half<1, 4>(a, grid, s);
    return TPA_OK;
  }
  if (nh == 2) {
    if (kg == 1) launch_half<2, 1>(a, grid, s);
    else if (kg == 2) launch_half<2, 2>(a, grid, s);
    else launch_half<2, 4>(a, grid, s);
    return TPA_OK;
  }
  return TPA_ERR_UNSUPPORTED;
}

static int dispatch_eval(int lg, int kg, const EvalArgs& a, int grid, hipStream_t s) {
  switch (lg) {
    case 1: return dispatch_kg<1>(kg, a, grid, s);
    case 2: return dispatch_kg<2>(kg, a, grid, s);
    case 4: return dispatch_kg<4>(kg, a, grid, s);
    case 8: return dispatch_kg<8>(kg, a, grid, s);
    case 16: return dispatch_kg<16>(kg, a, grid, s);
    default: return TPA_ERR_UNSUPPORTED;
  }
}

template <int LG, int KG>
static void launch_shape(const EvalShapeArgs& a, int grid, hipStream_t s) {
  const int need = (a.a.P + KG - 1) / KG;  // PML fitted to P as for the folded pass
  if (need <= 4)
    launch_pass((k_eval_shape<LG, KG, 4>), dim3(grid), dim3(256), s, a);
  else if (need <= 6)
    launch_pass((k_eval_shape<LG, KG, 6>), dim3(grid), dim3(256), s, a);
  else
    launch_pass((k_eval_shape<LG, KG, 8>), dim3(grid), dim3(256), s, a);
}

template <int LG>
static int dispatch_shape_kg(int kg, const EvalShapeArgs& a, int grid, hipStream_t s) {
  switch (kg) {
    case 1: launch_shape<LG, 1>(a, grid, s); return TPA_OK;
    case 2: launch_shape<LG, 2>(a, grid, s); return TPA_OK;
    case 4: launch_shape<LG, 4>(a, grid, s); return TPA_OK;
    case 8: if (LG * 8 <= 64) { launch_shape<LG, (LG * 8 <= 64 ? 8 : 1)>(a, grid, s); return TPA_OK; }
            return TPA_ERR_UNSUPPORTED;
    case 16: if (LG * 16 <= 64) { launch_shape<LG, (LG * 16 <= 64 ? 16 : 1)>(a, grid, s); return TPA_OK; }
             return TPA_ERR_UNSUPPORTED;
    default: return TPA_ERR_UNSUPPORTED;
  }
}

static int dispatch_shape(int lg, int kg, const EvalShapeArgs& a, int grid, hipStream_t s) {
  switch (lg) {
    case 1: return dispatch_shape_kg<1>(kg, a, grid, s);
    case 2: return dispatch_shape_kg<2>(kg, a, grid, s);
    case 4: return dispatch_shape_kg<4>(kg, a, grid, s);
    case 8: return dispatch_shape_kg<8>(kg, a, grid, s);
    case 16: return dispatch_shape_kg<16>(kg, a, grid, s);
    default: return TPA_ERR_UNSUPPORTED;
  }
}

// Blocks of the evaluation grid: at most kEvalGridDefault (TFPWA_EVAL_GRID: 64 .. kEvalGridMax).
static int eval_grid(int64_t n, int lanes_per_event) {
  static const int gmax = [] {
    const char* e = std::getenv("TFPWA_EVAL_GRID");
    const int v = e ? std::atoi(e) : kEvalGridDefault;
    return (v >= 64 && v <= kEvalGridMax) ? v : kEvalGridDefault;
  }();
  const int64_t epb = 4 * (64 / lanes_per_event);
  int64_t g = (n + epb - 1) / epb;
  if (g < 1) g = 1;
  return (int)std::min<int64_t>(g, gmax);
}

// Blocks of the helicity-halved passes, which run two waves per SIMD: one resident block of
// four waves per two SIMDs of every CU (2 x the CU count: 512 on the MI355X), so each wave
// slot holds one wave for the whole pass and no second, partial round of blocks trails it
// (1M-event data pass at config 2: 85.6 us at 512 blocks, 88.0 at 768, 88.9-90.0 at 1024,
// one process each on one box). TFPWA_EVAL_GRID overrides it as for the other passes.
static int half_grid(int64_t n, int lanes_per_event) {
  static const int gmax = [] {
    const char* e = std::getenv("TFPWA_EVAL_GRID");
    if (e) {
      const int v = std::atoi(e);
      if (v >= 64 && v <= kEvalGridMax) return v;
    }
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus < 1)
      return kEvalGridDefault;
    return std::min(kEvalGridMax, std::max(64, 2 * cus));
  }();
  const int64_t epb = 4 * (64 / lanes_per_event);
  int64_t g = (n + epb - 1) / epb;
  if (g < 1) g = 1;
  return (int)std::min<int64_t>(g, gmax);
}

#define TPA_POOL_ALIGN 64  // events: the granule of pool ranges
#define TPA_POOL_SLOTS 2   // batches that can be in flight on a pool at once

struct tpa_pool_s {
  std::mutex mu;  // one call at a time: the range maps, staging buffers and graphs are shared
  tpa_model_s* m = nullptr;
  int64_t cap = 0;
  double* d_ang = nullptr;
  double2* d_h = nullptr;
  double* d_w = nullptr;
  std::map<int64_t, int64_t> free_ranges;  // offset -> length
  std::map<int64_t, int64_t> used_ranges;
  int max_items = 0, max_chunks = 0;
  // Batched evaluations run through TPA_POOL_SLOTS independent sets of work lists, partials
  // and results, so one batch can be queued on the stream while the previous one runs.
  struct Slot {
    int64_t* d_blob = nullptr;  // the batch's work list and coefficients (BatchView)
    double* d_part = nullptr;
    double* d_out = nullptr;
    int64_t* h_blob = nullptr;  // its pinned mirror, copied in one piece per batch
    double* h_out = nullptr;
    std::map<int, hipGraphExec_t> graphs;  // grid bucket -> instantiated graph
    hipEvent_t t0 = nullptr, t1 = nullptr, done = nullptr;
    int n_items = 0;
    bool busy = false;  // submitted and not yet finished
  } slot[TPA_POOL_SLOTS];
  hipStream_t stream = nullptr;  // the pool's own stream (graph capture never meets other work)
  // adds: serialised among themselves, overlapping the batched evaluations on their own stream
  std::mutex add_mu;
  hipStream_t add_stream = nullptr;
  double* d_stage = nullptr;  // p4 [12][n] then weight [n]
  double* h_stage = nullptr;  // pinned mirror
  int64_t stage_n = 0;
  double* d_hpart = nullptr;  // second-moment pass (tpa_pool_hess)
  double* d_hout = nullptr;
  double2* d_hcoef = nullptr;
  float2* d_h32 = nullptr;  // FP32 copy of the cache (tpa_pool_keep_f32): batched passes read it
};

static void release_range(tpa_pool_s* p, int64_t offset);

// ---- second-moment pass (k_hess) shared by event sets and the pool
constexpr int kHessGridMax = 512;

static int hess_nacc(int P) { return 2 + 2 * P + 2 * P * (2 * P + 1); }

template <int LG>
static void launch_hess(const HessArgs& a, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_hess<LG>, dim3(grid), dim3(256), 0, s, a);
}

// Partials of the MFMA second-moment kernel: value pair, G, then the O1 and O2 tiles.
static int hess_mfma_nacc(int P) {
  const int QT = (2 * P + 15) / 16;
  return 2 + 2 * P + 2 * (QT * (QT + 1) / 2) * 256;
}

template <int LG>
static int launch_hess_mfma(const HessArgs& a, int QT, int grid, hipStream_t s) {
  switch (QT) {
    case 1: hipLaunchKernelGGL((k_hess_mfma<LG, 1>), dim3(grid), dim3(256), 0, s, a); return TPA_OK;
    case 2: hipLaunchKernelGGL((k_hess_mfma<LG, 2>), dim3(grid), dim3(256), 0, s, a); return TPA_OK;
    case 3: hipLaunchKernelGGL((k_hess_mfma<LG, 3>), dim3(grid), dim3(256), 0, s, a); return TPA_OK;
    case 4: hipLaunchKernelGGL((k_hess_mfma<LG, 4>), dim3(grid), dim3(256), 0, s, a); return TPA_OK;
    default: return TPA_ERR_UNSUPPORTED;
  }
}

// The second-moment pass: k_hess_mfma (default) or, with TFPWA_HESS_KERNEL=vec, the FP64 VALU
// kernel k_hess; both deliver the packed layout of tpa_hess.
static int run_hess(const TpaModel& M, const double2* h, const double* w, int64_t off, int64_t n,
                    int64_t S, const double* coef, int mode, hipStream_t s, double** d_part,
                    double** d_out, double2** d_coef, double* out) {
  if (M.P > TPA_HESS_PMAX) return fail(TPA_ERR_UNSUPPORTED, "second moments need P <= 32");
  static const bool vec = [] {
    const char* v = getenv("TFPWA_HESS_KERNEL");
    return v && std::string(v) == "vec";
  }();
  const int P = M.P, Q = 2 * P;
  const int nacc = vec ? hess_nacc(P) : hess_mfma_nacc(P);
  const int cap_nacc = std::max(hess_nacc(TPA_HESS_PMAX), hess_mfma_nacc(TPA_HESS_PMAX));
  if (!*d_part) HIPCHK(hipMalloc(d_part, sizeof(double) * (size_t)kHessGridMax * cap_nacc));
  if (!*d_out) HIPCHK(hipMalloc(d_out, sizeof(double) * cap_nacc));
  if (!*d_coef) HIPCHK(hipMalloc(d_coef, sizeof(double2) * TPA_HESS_PMAX));
  HIPCHK(hipMemcpyAsync(*d_coef, coef, sizeof(double2) * P, hipMemcpyHostToDevice, s));
  HessArgs a;
  a.h = h;
  a.w = w;
  a.coef = *d_coef;
  a.partials = *d_part;
  a.off = off;
  a.n = n;
  a.S = S;
  a.P = P;
  a.mode = mode;
  a.nacc = nacc;
  const int E = (vec ? TPA_HESS_PAIRS : 128) / M.lam_pad;
  const int64_t tiles = (n + E - 1) / E;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, kHessGridMax));
  const int QT = (Q + 15) / 16;
  int rc = TPA_OK;
  if (vec) {
    switch (M.lam_pad) {
      case 1: launch_hess<1>(a, grid, s); break;
      case 2: launch_hess<2>(a, grid, s); break;
      case 4: launch_hess<4>(a, grid, s); break;
      case 8: launch_hess<8>(a, grid, s); break;
      case 16: launch_hess<16>(a, grid, s); break;
      default: rc = TPA_ERR_UNSUPPORTED;
    }
  } else {
    switch (M.lam_pad) {
      case 1: rc = launch_hess_mfma<1>(a, QT, grid, s); break;
      case 2: rc = launch_hess_mfma<2>(a, QT, grid, s); break;
      case 4: rc = launch_hess_mfma<4>(a, QT, grid, s); break;
      case 8: rc = launch_hess_mfma<8>(a, QT, grid, s); break;
      case 16: rc = launch_hess_mfma<16>(a, QT, grid, s); break;
      default: rc = TPA_ERR_UNSUPPORTED;
    }
  }
  if (rc != TPA_OK) return fail(rc, "no second-moment kernel for this helicity count");
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_reduce, dim3(nacc), dim3(256), 0, s, *d_part, grid, nacc, 0, *d_out,
                     nullptr, nullptr, nullptr, (uint64_t)0);
  HIPCHK(hipGetLastError());
  if (vec) {
    HIPCHK(hipMemcpyAsync(out, *d_out, sizeof(double) * nacc, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return TPA_OK;
  }
  std::vector<double> raw(nacc);
  HIPCHK(hipMemcpyAsync(raw.data(), *d_out, sizeof(double) * nacc, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  // value, G, then each matrix's upper triangle (row-major, i <= j) from the 16x16 FP64 MFMA
  // accumulator layout: element (i, j) of tile q sits in lane j + 16 (i % 4), register i / 4
  for (int j = 0; j < 2 + Q; ++j) out[j] = raw[j];
  std::vector<int> tile(QT * QT, -1);
  for (int ti = 0, q = 0; ti < QT; ++ti)
    for (int tj = ti; tj < QT; ++tj, ++q) tile[ti * QT + tj] = q;
  const int NT = QT * (QT + 1) / 2;
  const int T = Q * (Q + 1) / 2;
  for (int m = 0; m < 2; ++m) {
    const double* src = raw.data() + 2 + Q + m * NT * 256;
    double* dst = out + 2 + Q + m * T;
    int idx = 0;
    for (int r = 0; r < Q; ++r)
      for (int c = r; c < Q; ++c, ++idx) {
        const int q = tile[(r / 16) * QT + c / 16];
        const int i = r % 16, j = c % 16;
        dst[idx] = src[(q * 4 + i / 4) * 64 + j + 16 * (i % 4)];
      }
  }
  return TPA_OK;
}

int tpa_set_error(int code, const std::string& msg) { return fail(code, msg); }

// Per-event angle-cache fields and four-momentum rows of a model's final state.
static int angle_fields(const tpa_model_s* m) {
  return m->kind == 4 ? TPA4_NF_TOPO * m->h.n_topo + 5 : TPA_NF_TOPO * m->h.n_topo + 4;
}
static int p4_rows(const tpa_model_s* m) { return 4 * (m->kind == 4 ? 4 : 3); }

// Angles of n events (p4 [rows][n] on the device) into columns off.. of a cache of stride S.
static void launch_cal_angle(const tpa_model_s* m, const double* d_p4, int64_t n, int64_t off,
                             int64_t S, double* ang, hipStream_t s) {
  const dim3 grid((unsigned)((n + 255) / 256));
  if (m->kind == 4)
    hipLaunchKernelGGL(k_cal_angle4, grid, dim3(256), 0, s, m->d4, d_p4, n, off, S, ang);
  else
    hipLaunchKernelGGL(k_cal_angle3, grid, dim3(256), 0, s, m->d, d_p4, n, off, S, ang);
}

// Build columns off .. off+n-1 of the angle cache (stride S) into h: the event set's or pool's
// cache (hS = S, hoff = off: hS < 0 selects these) or a chunk buffer (row stride hS, columns
// from hoff).
static void launch_build_pw(const tpa_model_s* m, const double* ang, int64_t n, int64_t off,
                            int64_t S, int fold, double2* h, hipStream_t s, int64_t hS = -1,
                            int64_t hoff = 0) {
  if (hS < 0) {
    hS = S;
    hoff = off;
  }
  const dim3 grid((unsigned)((n + 127) / 128));
  if (m->kind == 4)
    hipLaunchKernelGGL(k_build_pw4, grid, dim3(128), 0, s, m->d4, ang, n, off, S, m->d_cg,
                       m->d_dw, fold, h, hS, hoff);
  else
    hipLaunchKernelGGL(k_build_pw, grid, dim3(128), 0, s, m->d, ang, n, off, S, m->d_cg,
                       m->d_dw, fold, h, hS, hoff);
}

template <int LG, int KG>
static void launch_multi(const MultiArgs& a, int grid, hipStream_t s) {
  const int need = (a.P + KG - 1) / KG;  // PML as for the single-set pass
  if constexpr (LG >= 2) {
    if (a.h32 != nullptr) {  // the pool's FP32 copy
      if (need <= 4)
        launch_pass((k_eval_multi32<LG, KG, 4>), dim3(grid), dim3(256), s, a);
      else if (need <= 6)
        launch_pass((k_eval_multi32<LG, KG, 6>), dim3(grid), dim3(256), s, a);
      else if (need <= 7)
        launch_pass((k_eval_multi32<LG, KG, 7>), dim3(grid), dim3(256), s, a);
      else
        launch_pass((k_eval_multi32<LG, KG, 8>), dim3(grid), dim3(256), s, a);
      return;
    }
  }
  if (need <= 4)
    launch_pass((k_eval_multi<LG, KG, 4>), dim3(grid), dim3(256), s, a);
  else if (need <= 6)
    launch_pass((k_eval_multi<LG, KG, 6>), dim3(grid), dim3(256), s, a);
  else if (need <= 7)
    launch_pass((k_eval_multi<LG, KG, 7>), dim3(grid), dim3(256), s, a);
  else
    launch_pass((k_eval_multi<LG, KG, 8>), dim3(grid), dim3(256), s, a);
}

template <int LG>
static int dispatch_multi_kg(int kg, const MultiArgs& a, int grid, hipStream_t s) {
  switch (kg) {
    case 1: launch_multi<LG, 1>(a, grid, s); return TPA_OK;
    case 2: launch_multi<LG, 2>(a, grid, s); return TPA_OK;
    case 4: launch_multi<LG, 4>(a, grid, s); return TPA_OK;
    case 8: if (LG * 8 <= 64) { launch_multi<LG, (LG * 8 <= 64 ? 8 : 1)>(a, grid, s); return TPA_OK; }
            return TPA_ERR_UNSUPPORTED;
    case 16: if (LG * 16 <= 64) { launch_multi<LG, (LG * 16 <= 64 ? 16 : 1)>(a, grid, s); return TPA_OK; }
             return TPA_ERR_UNSUPPORTED;
    default: return TPA_ERR_UNSUPPORTED;
  }
}

static int dispatch_multi(int lg, int kg, const MultiArgs& a, int grid, hipStream_t s) {
  switch (lg) {
    case 1: return dispatch_multi_kg<1>(kg, a, grid, s);
    case 2: return dispatch_multi_kg<2>(kg, a, grid, s);
    case 4: return dispatch_multi_kg<4>(kg, a, grid, s);
    case 8: return dispatch_multi_kg<8>(kg, a, grid, s);
    case 16: return dispatch_multi_kg<16>(kg, a, grid, s);
    default: return TPA_ERR_UNSUPPORTED;
  }
}

extern "C" {

int tpa_abi_version(void) { return 1; }

const char* tpa_last_error(void) { return g_err.c_str(); }

int tpa_device_count(int* n) {
  if (!n) return fail(TPA_ERR_ARG, "null pointer");
  HIPCHK(hipGetDeviceCount(n));
  return TPA_OK;
}

int tpa_model_create(const void* desc, int64_t desc_bytes, const double* cg, int n_cg,
                     tpa_model_t* out) {
  if (!desc || !out || (n_cg > 0 && !cg)) return fail(TPA_ERR_ARG, "null pointer");
  if (desc_bytes != (int64_t)sizeof(TpaModel))
    return fail(TPA_ERR_ARG, "model description size mismatch: got " +
                                 std::to_string(desc_bytes) + " expected " +
                                 std::to_string(sizeof(TpaModel)));
  tpa_model_s* m = new tpa_model_s();
  std::memcpy(&m->h, desc, sizeof(TpaModel));
  if (n_cg > 0) m->cg_host.assign(cg, cg + n_cg);
  std::string why = check_model(m->h, n_cg);
  if (!why.empty()) {
    delete m;
    return fail(TPA_ERR_ARG, "invalid model: " + why);
  }
  hipError_t e;
  std::vector<double> dw = wigner_weights();
  if ((e = hipGetDevice(&m->device)) != hipSuccess ||
      (e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipMalloc(&m->d, sizeof(TpaModel))) != hipSuccess ||
      (e = hipMalloc(&m->d_cg, sizeof(double) * std::max(1, n_cg))) != hipSuccess ||
      (e = hipMalloc(&m->d_dw, sizeof(double) * dw.size())) != hipSuccess ||
      (e = hipMemcpy(m->d, &m->h, sizeof(TpaModel), hipMemcpyHostToDevice)) != hipSuccess ||
      (n_cg > 0 &&
       (e = hipMemcpy(m->d_cg, cg, sizeof(double) * n_cg, hipMemcpyHostToDevice)) != hipSuccess) ||
      (e = hipMemcpy(m->d_dw, dw.data(), sizeof(double) * dw.size(), hipMemcpyHostToDevice)) !=
          hipSuccess) {
    tpa_model_destroy(m);
    return fail(TPA_ERR_HIP, std::string("model upload: ") + hipGetErrorString(e));
  }
  std::vector<int> kchain(m->h.P);
  for (int c = 0; c < m->h.n_chain; ++c)
    for (int k = 0; k < m->h.chain[c].n_ls1 * m->h.chain[c].n_ls2; ++k)
      kchain[m->h.chain[c].k0 + k] = c;
  if ((e = hipMalloc(&m->d_kchain, sizeof(int) * m->h.P)) != hipSuccess ||
      (e = hipMemcpy(m->d_kchain, kchain.data(), sizeof(int) * m->h.P, hipMemcpyHostToDevice)) !=
          hipSuccess) {
    tpa_model_destroy(m);
    return fail(TPA_ERR_HIP, std::string("model upload: ") + hipGetErrorString(e));
  }
  *out = m;
  return TPA_OK;
}

int tpa_model4_create(const void* desc, int64_t desc_bytes, const double* cg, int n_cg,
                      tpa_model_t* out) {
  if (!desc || !out || (n_cg > 0 && !cg)) return fail(TPA_ERR_ARG, "null pointer");
  if (desc_bytes != (int64_t)sizeof(TpaModel4))
    return fail(TPA_ERR_ARG, "model description size mismatch: got " +
                                 std::to_string(desc_bytes) + " expected " +
                                 std::to_string(sizeof(TpaModel4)));
  tpa_model_s* m = new tpa_model_s();
  m->kind = 4;
  std::memcpy(&m->h4, desc, sizeof(TpaModel4));
  std::string why = check_model4(m->h4, n_cg);
  if (!why.empty()) {
    delete m;
    return fail(TPA_ERR_ARG, "invalid model: " + why);
  }
  // the header the generic passes (evaluation, Gram, second moments, pool) read
  m->h.n_final = 4;
  m->h.n_topo = m->h4.n_topo;
  m->h.n_chain = m->h4.n_chain;
  m->h.P = m->h4.P;
  m->h.lam = m->h4.lam;
  m->h.lam_pad = m->h4.lam_pad;
  m->h.top_j2 = m->h4.top_j2;
  m->h.n_top = m->h4.n_top;
  for (int i = 0; i < TPA_MAX_HEL; ++i) m->h.hel_top[i] = m->h4.hel_top[i];
  m->h.random_z = m->h4.random_z;
  m->h.n_cg = n_cg;
  hipError_t e;
  std::vector<double> dw = wigner_weights();
  if ((e = hipGetDevice(&m->device)) != hipSuccess ||
      (e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipMalloc(&m->d4, sizeof(TpaModel4))) != hipSuccess ||
      (e = hipMalloc(&m->d_cg, sizeof(double) * std::max(1, n_cg))) != hipSuccess ||
      (e = hipMalloc(&m->d_dw, sizeof(double) * dw.size())) != hipSuccess ||
      (e = hipMemcpy(m->d4, &m->h4, sizeof(TpaModel4), hipMemcpyHostToDevice)) != hipSuccess ||
      (n_cg > 0 &&
       (e = hipMemcpy(m->d_cg, cg, sizeof(double) * n_cg, hipMemcpyHostToDevice)) != hipSuccess) ||
      (e = hipMemcpy(m->d_dw, dw.data(), sizeof(double) * dw.size(), hipMemcpyHostToDevice)) !=
          hipSuccess) {
    tpa_model_destroy(m);
    return fail(TPA_ERR_HIP, std::string("model upload: ") + hipGetErrorString(e));
  }
  *out = m;
  return TPA_OK;
}

int tpa_model4_update(tpa_model_t m, const void* desc, int64_t desc_bytes) {
  if (!m || !desc) return fail(TPA_ERR_ARG, "null pointer");
  if (m->kind != 4) return fail(TPA_ERR_ARG, "not a four-body model");
  if (desc_bytes != (int64_t)sizeof(TpaModel4))
    return fail(TPA_ERR_ARG, "model description size mismatch");
  TpaModel4 next;
  std::memcpy(&next, desc, sizeof(TpaModel4));
  if (next.P != m->h4.P || next.n_cg != m->h4.n_cg || next.lam_pad != m->h4.lam_pad)
    return fail(TPA_ERR_ARG, "partial-wave layout cannot change");
  std::string why = check_model4(next, next.n_cg);
  if (!why.empty()) return fail(TPA_ERR_ARG, "invalid model: " + why);
  HIPCHK(hipStreamSynchronize(m->stream));
  m->h4 = next;
  HIPCHK(hipMemcpy(m->d4, &m->h4, sizeof(TpaModel4), hipMemcpyHostToDevice));
  return TPA_OK;
}

int tpa_model_destroy(tpa_model_t m) {
  if (!m) return TPA_OK;
  if (m->d) (void)hipFree(m->d);
  if (m->d4) (void)hipFree(m->d4);
  if (m->d_cg) (void)hipFree(m->d_cg);
  if (m->d_dw) (void)hipFree(m->d_dw);
  if (m->d_kchain) (void)hipFree(m->d_kchain);
  if (m->h_pin) (void)hipHostFree(m->h_pin);
  if (m->fused_mod) (void)hipModuleUnload(m->fused_mod);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return TPA_OK;
}

int tpa_model_update(tpa_model_t m, const void* desc, int64_t desc_bytes) {
  if (!m || !desc) return fail(TPA_ERR_ARG, "null pointer");
  if (m->kind != 3) return fail(TPA_ERR_ARG, "use tpa_model4_update for a four-body model");
  if (desc_bytes != (int64_t)sizeof(TpaModel))
    return fail(TPA_ERR_ARG, "model description size mismatch");
  TpaModel next;
  std::memcpy(&next, desc, sizeof(TpaModel));
  if (next.n_cg != m->h.n_cg) return fail(TPA_ERR_ARG, "CG table size cannot change");
  std::string why = check_model(next, next.n_cg);
  if (!why.empty()) return fail(TPA_ERR_ARG, "invalid model: " + why);
  HIPCHK(hipStreamSynchronize(m->stream));
  m->h = next;
  {
    std::lock_guard<std::mutex> lk(m->fused_mu);
    m->fused_stale = true;
  }
  HIPCHK(hipMemcpy(m->d, &m->h, sizeof(TpaModel), hipMemcpyHostToDevice));
  return TPA_OK;
}

int tpa_model_set_lineshapes(tpa_model_t m, int n_chain, const double* m0, const double* g0) {
  if (!m || !m0 || !g0) return fail(TPA_ERR_ARG, "null pointer");
  if (m->kind != 3) return fail(TPA_ERR_UNSUPPORTED, "lineshape parameters of a three-body model");
  if (n_chain != m->h.n_chain) return fail(TPA_ERR_ARG, "chain count mismatch");
  if (!m->h_pin) HIPCHK(hipHostMalloc((void**)&m->h_pin, sizeof(TpaModel)));
  // the pinned copy is read by the asynchronous copy: wait for the previous one to finish
  HIPCHK(hipStreamSynchronize(m->stream));
  for (int c = 0; c < n_chain; ++c) {
    m->h.chain[c].m0 = m0[c];
    m->h.chain[c].g0 = g0[c];
  }
  std::memcpy(m->h_pin, &m->h, sizeof(TpaModel));
  HIPCHK(hipMemcpyAsync(m->d, m->h_pin, sizeof(TpaModel), hipMemcpyHostToDevice, m->stream));
  return TPA_OK;
}

int tpa_model_set_barrier_flags(tpa_model_t m, int n_chain, const int32_t* act1,
                                const int32_t* act2) {
  if (!m || !act1 || !act2) return fail(TPA_ERR_ARG, "null pointer");
  if (m->kind != 3) return fail(TPA_ERR_UNSUPPORTED, "barrier flags of a three-body model");
  if (n_chain != m->h.n_chain) return fail(TPA_ERR_ARG, "chain count mismatch");
  if (!m->h_pin) HIPCHK(hipHostMalloc((void**)&m->h_pin, sizeof(TpaModel)));
  HIPCHK(hipStreamSynchronize(m->stream));
  for (int c = 0; c < n_chain; ++c) {
    for (int s = 0; s < m->h.chain[c].n_ls1; ++s) m->h.chain[c].act1[s] = act1[c * TPA_MAX_LS + s];
    for (int s = 0; s < m->h.chain[c].n_ls2; ++s) m->h.chain[c].act2[s] = act2[c * TPA_MAX_LS + s];
  }
  std::memcpy(m->h_pin, &m->h, sizeof(TpaModel));
  HIPCHK(hipMemcpyAsync(m->d, m->h_pin, sizeof(TpaModel), hipMemcpyHostToDevice, m->stream));
  return TPA_OK;
}

int tpa_events_destroy(tpa_events_t ev) {
  if (!ev) return TPA_OK;
  if (ev->d_ang) (void)hipFree(ev->d_ang);
  if (ev->d_h) (void)hipFree(ev->d_h);
  if (ev->d_h32) (void)hipFree(ev->d_h32);
  if (ev->d_g32h) (void)hipFree(ev->d_g32h);
  if (ev->d_rho) (void)hipFree(ev->d_rho);
  if (ev->d_g64h) (void)hipFree(ev->d_g64h);
  if (ev->d_rho64) (void)hipFree(ev->d_rho64);
  if (ev->d_rchunk) (void)hipFree(ev->d_rchunk);
  if (ev->d_rout) (void)hipFree(ev->d_rout);
  if (ev->d_resid) (void)hipFree(ev->d_resid);
  if (ev->d_cnt) (void)hipFree(ev->d_cnt);
  if (ev->h_flag) (void)hipHostFree(ev->h_flag);
  if (ev->d_w) (void)hipFree(ev->d_w);
  if (ev->d_part) (void)hipFree(ev->d_part);
  if (ev->d_out) (void)hipFree(ev->d_out);
  if (ev->d_coef) (void)hipFree(ev->d_coef);
  if (ev->d_f) (void)hipFree(ev->d_f);
  if (ev->d_gpart) (void)hipFree(ev->d_gpart);
  if (ev->d_gout) (void)hipFree(ev->d_gout);
  if (ev->d_bw) (void)hipFree(ev->d_bw);
  if (ev->d_hpart) (void)hipFree(ev->d_hpart);
  if (ev->d_hout) (void)hipFree(ev->d_hout);
  if (ev->d_hcoef) (void)hipFree(ev->d_hcoef);
  if (ev->d_bw2) (void)hipFree(ev->d_bw2);
  if (ev->d_hspart) (void)hipFree(ev->d_hspart);
  if (ev->d_hsout) (void)hipFree(ev->d_hsout);
  if (ev->t0) (void)hipEventDestroy(ev->t0);
  if (ev->t1) (void)hipEventDestroy(ev->t1);
  if (ev->h_io) (void)hipHostFree(ev->h_io);
  delete ev;
  return TPA_OK;
}

// An event set of n events: its buffers, weights (host, or 1 when null) and angle cache.
// The angles come from device four-momenta (p4_dev, [rows][n]) when given; otherwise the
// caller fills ev->d_ang.
static int events_make(tpa_model_t m, int64_t n, const double* weight, const double* p4_dev,
                       tpa_events_s** out) {
  tpa_events_s* ev = new tpa_events_s();
  ev->m = m;
  ev->n = n;
  ev->nf = angle_fields(m);
  const int P = m->h.P;
  hipError_t e;
  auto bail = [&](hipError_t er, const char* what) {
    tpa_events_destroy(ev);
    return fail(TPA_ERR_HIP, std::string(what) + ": " + hipGetErrorString(er));
  };
  if ((e = hipMalloc(&ev->d_ang, sizeof(double) * ev->nf * n)) != hipSuccess)
    return bail(e, "alloc angles");
  if ((e = hipMalloc(&ev->d_w, sizeof(double) * n)) != hipSuccess) return bail(e, "alloc w");
  if ((e = hipMalloc(&ev->d_part, sizeof(double) * kEvalGridMax * (2 + 6 * std::max(P, 1)))) !=
      hipSuccess)
    return bail(e, "alloc partials");
  if ((e = hipMalloc(&ev->d_out, sizeof(double) * TPA_NACC_SHAPE_MAX)) != hipSuccess)
    return bail(e, "alloc out");
  if ((e = hipMalloc(&ev->d_coef, sizeof(double2) * std::max(P, 1))) != hipSuccess)
    return bail(e, "alloc coef");
  if ((e = hipEventCreate(&ev->t0)) != hipSuccess || (e = hipEventCreate(&ev->t1)) != hipSuccess)
    return bail(e, "event create");
  if (weight) {
    if ((e = hipMemcpyAsync(ev->d_w, weight, sizeof(double) * n, hipMemcpyHostToDevice,
                            m->stream)) != hipSuccess)
      return bail(e, "upload weights");
  } else {
    std::vector<double> ones(n, 1.0);
    if ((e = hipMemcpyAsync(ev->d_w, ones.data(), sizeof(double) * n, hipMemcpyHostToDevice,
                            m->stream)) != hipSuccess)
      return bail(e, "upload weights");
  }
  if (p4_dev) {
    launch_cal_angle(m, p4_dev, n, 0, n, ev->d_ang, m->stream);
    if ((e = hipGetLastError()) != hipSuccess) return bail(e, "launch cal_angle");
  }
  if ((e = hipStreamSynchronize(m->stream)) != hipSuccess) return bail(e, "cal_angle");
  *out = ev;
  return TPA_OK;
}

int tpa_events_create(tpa_model_t m, const double* p4, int64_t n, const double* weight,
                      tpa_events_t* out) {
  if (!m || !p4 || !out) return fail(TPA_ERR_ARG, "null pointer");
  if (n < 1) return fail(TPA_ERR_ARG, "event count must be positive");
  const int rows = p4_rows(m);
  double* d_p4 = nullptr;
  HIPCHK(hipMalloc(&d_p4, sizeof(double) * rows * n));
  hipError_t e = hipMemcpyAsync(d_p4, p4, sizeof(double) * rows * n, hipMemcpyHostToDevice,
                                m->stream);
  if (e != hipSuccess) {
    (void)hipFree(d_p4);
    return fail(TPA_ERR_HIP, std::string("upload p4: ") + hipGetErrorString(e));
  }
  const int rc = events_make(m, n, weight, d_p4, out);
  (void)hipFree(d_p4);
  return rc;
}

int tpa_events_generate(tpa_model_t m, double m_top, const double* masses, int64_t n,
                        uint64_t seed, uint64_t first, const double* weight, double* p4_out,
                        tpa_events_t* out) {
  if (!m || !masses || !out) return fail(TPA_ERR_ARG, "null pointer");
  if (n < 1) return fail(TPA_ERR_ARG, "event count must be positive");
  const int nf = m->kind == 4 ? 4 : 3;
  double msum = 0.0;
  for (int i = 0; i < nf; ++i) {
    if (!(masses[i] >= 0.0)) return fail(TPA_ERR_ARG, "masses must be non-negative");
    msum += masses[i];
  }
  if (!(m_top > msum)) return fail(TPA_ERR_ARG, "the top mass is below the final-state masses");
  // GENBOD's weight maximum (emmin, emmax running bounds of the intermediate masses)
  double emmax = m_top - msum + masses[0], emmin = 0.0, wtmax = 1.0;
  for (int i = 1; i < nf; ++i) {
    emmin += masses[i - 1];
    emmax += masses[i];
    const double s = emmin + masses[i], d = emmin - masses[i];
    const double p2 = (emmax * emmax - s * s) * (emmax * emmax - d * d);
    wtmax *= p2 > 0.0 ? std::sqrt(p2) / (2.0 * emmax) : 0.0;
  }
  const int rows = 4 * nf;
  double* d_p4 = nullptr;
  double* d_m = nullptr;
  HIPCHK(hipMalloc(&d_p4, sizeof(double) * rows * n));
  hipError_t e = hipMalloc(&d_m, sizeof(double) * 4);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_m, masses, sizeof(double) * nf, hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_phsp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->stream, nf,
                       m_top, d_m, wtmax, n, (int64_t)first, (uint32_t)seed,
                       (uint32_t)(seed >> 32), d_p4);
    e = hipGetLastError();
  }
  if (e == hipSuccess && p4_out)
    e = hipMemcpyAsync(p4_out, d_p4, sizeof(double) * rows * n, hipMemcpyDeviceToHost,
                       m->stream);
  int rc = TPA_OK;
  if (e != hipSuccess)
    rc = fail(TPA_ERR_HIP, std::string("phase space: ") + hipGetErrorString(e));
  else
    rc = events_make(m, n, weight, d_p4, out);
  (void)hipFree(d_p4);
  if (d_m) (void)hipFree(d_m);
  return rc;
}

int tpa_events_select(tpa_events_t src, const int64_t* idx, int64_t n, const double* weight,
                      tpa_events_t* out) {
  if (!src || !idx || !out) return fail(TPA_ERR_ARG, "null pointer");
  if (n < 1) return fail(TPA_ERR_ARG, "event count must be positive");
  for (int64_t i = 0; i < n; ++i)
    if (idx[i] < 0 || idx[i] >= src->n) return fail(TPA_ERR_ARG, "index out of range");
  tpa_model_s* m = src->m;
  tpa_events_s* ev = nullptr;
  int rc = events_make(m, n, weight, nullptr, &ev);
  if (rc != TPA_OK) return rc;
  int64_t* d_idx = nullptr;
  hipError_t e = hipMalloc(&d_idx, sizeof(int64_t) * n);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_idx, idx, sizeof(int64_t) * n, hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_gather_cols, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->stream,
                       src->d_ang, src->nf, src->n, d_idx, n, ev->d_ang);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (d_idx) (void)hipFree(d_idx);
  if (e != hipSuccess) {
    tpa_events_destroy(ev);
    return fail(TPA_ERR_HIP, std::string("select: ") + hipGetErrorString(e));
  }
  *out = ev;
  return TPA_OK;
}

int tpa_events_concat(const tpa_events_t* srcs, int n_src, tpa_events_t* out) {
  if (!srcs || !out) return fail(TPA_ERR_ARG, "null pointer");
  if (n_src < 1) return fail(TPA_ERR_ARG, "no source event sets");
  int64_t total = 0;
  for (int i = 0; i < n_src; ++i) {
    if (!srcs[i]) return fail(TPA_ERR_ARG, "null source event set");
    if (srcs[i]->m != srcs[0]->m) return fail(TPA_ERR_ARG, "sources of different models");
    total += srcs[i]->n;
  }
  tpa_model_s* m = srcs[0]->m;
  tpa_events_s* ev = nullptr;
  int rc = events_make(m, total, nullptr, nullptr, &ev);
  if (rc != TPA_OK) return rc;
  hipError_t e = hipSuccess;
  int64_t off = 0;
  for (int i = 0; i < n_src && e == hipSuccess; ++i) {
    const tpa_events_s* src = srcs[i];
    // angle cache [field][n]: a (nf x n_i) block at column off of the (nf x total) cache
    e = hipMemcpy2DAsync(ev->d_ang + off, sizeof(double) * total, src->d_ang,
                         sizeof(double) * src->n, sizeof(double) * src->n, ev->nf,
                         hipMemcpyDeviceToDevice, m->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(ev->d_w + off, src->d_w, sizeof(double) * src->n,
                         hipMemcpyDeviceToDevice, m->stream);
    off += src->n;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (e != hipSuccess) {
    tpa_events_destroy(ev);
    return fail(TPA_ERR_HIP, std::string("concat: ") + hipGetErrorString(e));
  }
  *out = ev;
  return TPA_OK;
}

int64_t tpa_events_size(tpa_events_t ev) { return ev ? ev->n : -1; }

int tpa_events_set_weight(tpa_events_t ev, const double* weight) {
  if (!ev || !weight) return fail(TPA_ERR_ARG, "null pointer");
  HIPCHK(hipMemcpyAsync(ev->d_w, weight, sizeof(double) * ev->n, hipMemcpyHostToDevice,
                        ev->m->stream));
  HIPCHK(hipStreamSynchronize(ev->m->stream));
  return TPA_OK;
}

int tpa_events_get_angles(tpa_events_t ev, double* out, int64_t out_len) {
  if (!ev || !out) return fail(TPA_ERR_ARG, "null pointer");
  if (out_len < (int64_t)ev->nf * ev->n) return fail(TPA_ERR_ARG, "output buffer too small");
  HIPCHK(hipMemcpyAsync(out, ev->d_ang, sizeof(double) * ev->nf * ev->n, hipMemcpyDeviceToHost,
                        ev->m->stream));
  HIPCHK(hipStreamSynchronize(ev->m->stream));
  return TPA_OK;
}

// Refresh the FP32 copy of ev's folded cache (queued on the model's stream).
static int convert_f32(tpa_events_s* ev) {
  const TpaModel& M = ev->m->h;
  const int64_t n = (int64_t)M.P * ev->n * M.lam_pad;
  if (!ev->d_h32) HIPCHK(hipMalloc(&ev->d_h32, sizeof(float2) * (size_t)n));
  const int64_t grid = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_pw_to_f32, dim3((unsigned)grid), dim3(256), 0, ev->m->stream, ev->d_h,
                     ev->d_h32, n);
  HIPCHK(hipGetLastError());
  ev->h32_valid = true;
  return TPA_OK;
}

// The helicity-halved copy of ev's folded cache (see tpa_half.h). The host finds each slot's
// mirror and sign s_o from the helicity lists, and each wave's sign relative to its chain's
// reference wave from a sample of events (the sign of sum Re[h_k,-o conj(h_r,-o) h_k,o
// conj(h_r,o)] = eta_k eta_r sum |h_k,o|^2 |h_r,o|^2); k_pw_to_half writes the copy and
// measures the residual of the relation on every event, which the host checks.
#define TPA_HALF_SAMPLE 512
#define TPA_HALF_TOL 1e-9
static int convert_half(tpa_events_s* ev, bool fp64) {
  const tpa_model_s* m = ev->m;
  const TpaModel& M = m->h;
  const int L = M.lam;
  const int Lp = M.lam_pad;
  if (M.P > TPA_EVAL_CIN)
    return fail(TPA_ERR_UNSUPPORTED, "the helicity-halved copy serves at most 32 partial waves");
  // helicity (2 lambda) of every outer slot: the top's, then the finals' (three-body); the
  // top's alone (four-body finals are spin 0)
  std::vector<std::vector<int>> hel(L);
  if (m->kind == 4) {
    if (L != M.n_top) return fail(TPA_ERR_UNSUPPORTED, "unexpected four-body slot count");
    for (int i = 0; i < L; ++i) hel[i] = {M.hel_top[i]};
  } else {
    int o = 0;
    for (int iA = 0; iA < M.n_top; ++iA)
      for (int j0 = 0; j0 < M.n_fh[0]; ++j0)
        for (int j1 = 0; j1 < M.n_fh[1]; ++j1)
          for (int j2 = 0; j2 < M.n_fh[2]; ++j2, ++o) {
            if (o >= L) return fail(TPA_ERR_UNSUPPORTED, "unexpected slot count");
            hel[o] = {M.hel_top[iA], M.hel_f[0][j0], M.hel_f[1][j1], M.hel_f[2][j2]};
          }
    if (o != L) return fail(TPA_ERR_UNSUPPORTED, "unexpected slot count");
  }
  HalfCopyArgs a{};
  int nh = 0;
  for (int o = 0; o < L; ++o) {
    std::vector<int> neg(hel[o].size());
    int sum = 0;
    for (size_t i = 0; i < neg.size(); ++i) {
      neg[i] = -hel[o][i];
      sum += hel[o][i];
    }
    int mo = -1;
    for (int q = 0; q < L; ++q)
      if (hel[q] == neg) mo = q;
    if (mo < 0 || mo == o || (sum % 2) != 0)
      return fail(TPA_ERR_UNSUPPORTED, "helicity lists without a mirror for every slot");
    if (o < mo) {
      if (nh >= TPA_HALF_MAX_NH) return fail(TPA_ERR_UNSUPPORTED, "too many helicity slots");
      a.slot[nh] = o;
      a.mirror[nh] = mo;
      a.sgn[nh] = ((sum / 2) % 2 == 0) ? 1 : -1;  // (-1)^(sum of helicities)
      ++nh;
    }
  }
  if (nh != 1 && nh != 2)
    return fail(TPA_ERR_UNSUPPORTED, "the helicity-halved pass serves 2 or 4 helicity slots");
  const int n_chain = m->kind == 4 ? m->h4.n_chain : M.n_chain;
  std::vector<int> k0(n_chain), kn(n_chain);
  for (int c = 0; c < n_chain; ++c) {
    if (m->kind == 4) {
      const Tpa4Chain& C = m->h4.chain[c];
      k0[c] = C.k0;
      kn[c] = C.n_ls[0] * C.n_ls[1] * C.n_ls[2];
    } else {
      const TpaChain& C = M.chain[c];
      k0[c] = C.k0;
      kn[c] = C.n_ls1 * C.n_ls2;
    }
  }
  // wave signs relative to each chain's reference wave, from a sample of events
  const int64_t ns = std::min<int64_t>(ev->n, TPA_HALF_SAMPLE);
  std::vector<double2> smp((size_t)M.P * ns * Lp);
  HIPCHK(hipMemcpy2DAsync(smp.data(), sizeof(double2) * ns * Lp, ev->d_h,
                          sizeof(double2) * ev->n * Lp, sizeof(double2) * ns * Lp, M.P,
                          hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  auto H = [&](int k, int64_t e, int o) { return smp[((size_t)k * ns + e) * Lp + o]; };
  ev->half_rot.assign(M.P, 0);
  ev->half_kch.assign(TPA_EVAL_CIN, 0);
  for (int c = 0; c < n_chain; ++c) {
    int r = k0[c];
    double rn = -1.0;
    for (int k = k0[c]; k < k0[c] + kn[c]; ++k) {
      double nk = 0.0;
      for (int64_t e = 0; e < ns; ++e)
        for (int q = 0; q < nh; ++q) {
          const double2 v = H(k, e, a.slot[q]);
          nk += v.x * v.x + v.y * v.y;
        }
      if (nk > rn) {
        rn = nk;
        r = k;
      }
    }
    for (int k = k0[c]; k < k0[c] + kn[c]; ++k) {
      double sk = 0.0;
      for (int64_t e = 0; e < ns; ++e)
        for (int q = 0; q < nh; ++q) {
          const double2 hm = H(k, e, a.mirror[q]), rm = H(r, e, a.mirror[q]);
          const double2 hk = H(k, e, a.slot[q]), rk = H(r, e, a.slot[q]);
          // x = h_k,-o conj(h_r,-o), y = h_k,o conj(h_r,o); Re(x y)
          const double xr = hm.x * rm.x + hm.y * rm.y, xi = hm.y * rm.x - hm.x * rm.y;
          const double yr = hk.x * rk.x + hk.y * rk.y, yi = hk.y * rk.x - hk.x * rk.y;
          sk += xr * yr - xi * yi;
        }
      ev->half_rot[k] = sk < 0.0 ? 1 : 0;
      ev->half_kch[k] = (signed char)c;
    }
  }
  // lane plan of the chain-grouped FP32 pass (two stored slots, 17 to 32 waves)
  ev->half_plan_pml = 0;
  if (nh == 2 && M.P > 16 && M.P <= 32) {
    const int pml = M.P <= 24 ? 6 : 8;
    if (plan_half_lanes(M.P, ev->half_kch.data(), pml, ev->half_plan, ev->half_pch))
      ev->half_plan_pml = pml;
  }
  const size_t ng = (size_t)M.P * ev->n * nh, nr = (size_t)n_chain * ev->n;
  if (ev->half_nh != nh) {
    if (ev->d_g32h) HIPCHK(hipFree(ev->d_g32h));
    if (ev->d_g64h) HIPCHK(hipFree(ev->d_g64h));
    ev->d_g32h = nullptr;
    ev->d_g64h = nullptr;
  }
  if (fp64) {
    if (!ev->d_g64h) HIPCHK(hipMalloc(&ev->d_g64h, sizeof(double2) * ng));
    if (!ev->d_rho64) HIPCHK(hipMalloc(&ev->d_rho64, sizeof(double2) * nr));
  } else {
    if (!ev->d_g32h) HIPCHK(hipMalloc(&ev->d_g32h, sizeof(float2) * ng));
    if (!ev->d_rho) HIPCHK(hipMalloc(&ev->d_rho, sizeof(float2) * nr));
  }
  if (!ev->d_resid) HIPCHK(hipMalloc(&ev->d_resid, sizeof(unsigned long long)));
  ev->half_nh = nh;
  for (int q = 0; q < nh; ++q) {
    ev->half_slot[q] = a.slot[q];
    ev->half_mirror[q] = a.mirror[q];
    ev->half_sgn[q] = a.sgn[q];
  }
  a.h = ev->d_h;
  a.g = fp64 ? (void*)ev->d_g64h : (void*)ev->d_g32h;
  a.rho = fp64 ? (void*)ev->d_rho64 : (void*)ev->d_rho;
  a.resid = ev->d_resid;
  a.n = ev->n;
  a.S = ev->n;
  a.P = M.P;
  a.lam_pad = Lp;
  a.nh = nh;
  a.n_chain = n_chain;
  for (int c = 0; c < n_chain; ++c) {
    a.ch_k0[c] = k0[c];
    a.ch_n[c] = kn[c];
  }
  for (int k = 0; k < M.P; ++k) a.rot[k] = ev->half_rot[k];
  unsigned long long zero = 0, bits = 0;
  HIPCHK(hipMemcpyAsync(ev->d_resid, &zero, sizeof(zero), hipMemcpyHostToDevice, m->stream));
  const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((ev->n + 255) / 256, 4096));
  if (fp64)
    hipLaunchKernelGGL(k_pw_to_half<double>, dim3((unsigned)grid), dim3(256), 0, m->stream, a);
  else
    hipLaunchKernelGGL(k_pw_to_half<float>, dim3((unsigned)grid), dim3(256), 0, m->stream, a);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(&bits, ev->d_resid, sizeof(bits), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  double r;
  std::memcpy(&r, &bits, sizeof(r));
  ev->half_resid = r;
  bool& valid = fp64 ? ev->h64h_valid : ev->h32h_valid;
  if (!(r <= TPA_HALF_TOL)) {
    valid = false;
    return fail(TPA_ERR_UNSUPPORTED,
                "the helicity-flip relation does not hold for this model's cache (relative "
                "residual " + std::to_string(r) + ")");
  }
  valid = true;
  return TPA_OK;
}

int tpa_events_build(tpa_events_t ev, int fold) {
  if (!ev) return fail(TPA_ERR_ARG, "null pointer");
  const TpaModel& M = ev->m->h;
  if (!ev->d_h) HIPCHK(hipMalloc(&ev->d_h, sizeof(double2) * (size_t)M.P * ev->n * M.lam_pad));
  if (ev->m->kind == 4 && !fold)
    return fail(TPA_ERR_UNSUPPORTED, "four-body models fold fixed lineshapes into the cache");
  launch_build_pw(ev->m, ev->d_ang, ev->n, 0, ev->n, fold, ev->d_h, ev->m->stream);
  HIPCHK(hipGetLastError());
  ev->built_fold = fold ? 1 : 0;
  ev->h32_valid = false;
  ev->h32h_valid = ev->h64h_valid = false;
  if (ev->keep32 && fold) {
    const int rc = convert_f32(ev);
    if (rc != TPA_OK) return rc;
  }
  if (ev->keep32h && fold) {
    const int rc = convert_half(ev, false);
    if (rc != TPA_OK) return rc;
  }
  if (ev->keep64h && fold) {
    const int rc = convert_half(ev, true);
    if (rc != TPA_OK) return rc;
  }
  HIPCHK(hipStreamSynchronize(ev->m->stream));
  return TPA_OK;
}

int tpa_events_keep_f32(tpa_events_t ev, int on) {
  if (!ev) return fail(TPA_ERR_ARG, "null pointer");
  if (!on) {
    HIPCHK(hipStreamSynchronize(ev->m->stream));
    if (ev->d_h32) HIPCHK(hipFree(ev->d_h32));
    ev->d_h32 = nullptr;
    ev->keep32 = ev->h32_valid = false;
    return TPA_OK;
  }
  ev->keep32 = true;
  if (ev->d_h && ev->built_fold == 1 && !ev->h32_valid) {
    const int rc = convert_f32(ev);
    if (rc != TPA_OK) return rc;
    HIPCHK(hipStreamSynchronize(ev->m->stream));
  }
  return TPA_OK;
}

int tpa_events_keep_half(tpa_events_t ev, int kinds) {
  if (!ev) return fail(TPA_ERR_ARG, "null pointer");
  if (kinds & ~3) return fail(TPA_ERR_ARG, "kinds: bit 0 the FP32 copy, bit 1 the FP64 copy");
  HIPCHK(hipStreamSynchronize(ev->m->stream));
  if (!(kinds & 1)) {
    if (ev->d_g32h) HIPCHK(hipFree(ev->d_g32h));
    if (ev->d_rho) HIPCHK(hipFree(ev->d_rho));
    ev->d_g32h = nullptr;
    ev->d_rho = nullptr;
    ev->keep32h = ev->h32h_valid = false;
  }
  if (!(kinds & 2)) {
    if (ev->d_g64h) HIPCHK(hipFree(ev->d_g64h));
    if (ev->d_rho64) HIPCHK(hipFree(ev->d_rho64));
    ev->d_g64h = nullptr;
    ev->d_rho64 = nullptr;
    ev->keep64h = ev->h64h_valid = false;
  }
  for (int fp64 = 0; fp64 < 2; ++fp64) {
    if (!(kinds & (1 << fp64))) continue;
    bool& keep = fp64 ? ev->keep64h : ev->keep32h;
    const bool valid = fp64 ? ev->h64h_valid : ev->h32h_valid;
    keep = true;
    if (ev->d_h && ev->built_fold == 1 && !valid) {
      const int rc = convert_half(ev, fp64 != 0);
      if (rc != TPA_OK) {
        keep = false;
        return rc;
      }
    }
  }
  return TPA_OK;
}

int tpa_events_get_half(tpa_events_t ev, float* g, int64_t g_len, float* rho, int64_t rho_len,
                        int* info, int info_len) {
  if (!ev || !g || !rho || !info) return fail(TPA_ERR_ARG, "null pointer");
  if (!ev->h32h_valid) return fail(TPA_ERR_ARG, "no helicity-halved copy");
  const tpa_model_s* m = ev->m;
  const int P = m->h.P;
  const int n_chain = m->kind == 4 ? m->h4.n_chain : m->h.n_chain;
  const int64_t ng = 2 * (int64_t)P * ev->n * ev->half_nh, nr = 2 * (int64_t)n_chain * ev->n;
  if (g_len < ng || rho_len < nr || info_len < 13 + 2 * P)
    return fail(TPA_ERR_ARG, "output buffer too small");
  HIPCHK(hipMemcpyAsync(g, ev->d_g32h, sizeof(float) * ng, hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipMemcpyAsync(rho, ev->d_rho, sizeof(float) * nr, hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  std::fill(info, info + 13 + 2 * P, 0);
  info[0] = ev->half_nh;
  for (int q = 0; q < ev->half_nh; ++q) {
    info[1 + q] = ev->half_slot[q];
    info[5 + q] = ev->half_mirror[q];
    info[9 + q] = ev->half_sgn[q];
  }
  for (int k = 0; k < P; ++k) {
    info[13 + k] = ev->half_rot[k];
    info[13 + P + k] = ev->half_kch[k];
  }
  // the chain-grouped pass's lane plan, when the caller's buffer takes it
  if (info_len >= 13 + 2 * P + 1 + TPA_EVAL_CIN) {
    info[13 + 2 * P] = ev->half_plan_pml;
    for (int i = 0; i < TPA_EVAL_CIN; ++i)
      info[14 + 2 * P + i] = ev->half_plan_pml > 0 ? ev->half_plan[i] : -1;
  }
  return TPA_OK;
}

int tpa_events_half_residual(tpa_events_t ev, double* out) {
  if (!ev || !out) return fail(TPA_ERR_ARG, "null pointer");
  *out = ev->half_resid;
  return TPA_OK;
}

int tpa_events_get_pw(tpa_events_t ev, double* out, int64_t out_len) {
  if (!ev || !out) return fail(TPA_ERR_ARG, "null pointer");
  if (!ev->d_h) return fail(TPA_ERR_ARG, "partial-wave cache not built");
  const TpaModel& M = ev->m->h;
  const int64_t len = 2 * (int64_t)M.P * ev->n * M.lam_pad;
  if (out_len < len) return fail(TPA_ERR_ARG, "output buffer too small");
  HIPCHK(hipMemcpyAsync(out, ev->d_h, sizeof(double) * len, hipMemcpyDeviceToHost,
                        ev->m->stream));
  HIPCHK(hipStreamSynchronize(ev->m->stream));
  return TPA_OK;
}

// One evaluation pass on stream s. async: queue it and return (shape passes without f_out
// only); eval_collect then waits for the stream and copies the sums.
static int eval_impl(tpa_events_s* ev, const double* coef, int mode, int flags, double* out,
                     double* f_out, hipStream_t s, bool async) {
  if (!ev || !coef || (!out && !async)) return fail(TPA_ERR_ARG, "null pointer");
  if (!ev->d_h) return fail(TPA_ERR_ARG, "partial-wave cache not built");
  if (mode != TPA_MODE_DATA && mode != TPA_MODE_MC) return fail(TPA_ERR_ARG, "bad mode");
  const TpaModel& M = ev->m->h;
  const bool want_f = (flags & TPA_FLAG_FOUT) != 0;
  if (async && (want_f || !(flags & TPA_FLAG_SHAPE)))
    return fail(TPA_ERR_UNSUPPORTED, "queued evaluation serves shape passes without f_out");
  if (want_f && !f_out) return fail(TPA_ERR_ARG, "f_out required");
  if (want_f && !ev->d_f) HIPCHK(hipMalloc(&ev->d_f, sizeof(double) * ev->n));
  if (!ev->h_io) HIPCHK(hipHostMalloc((void**)&ev->h_io, sizeof(double) * TPA_IO_DOUBLES));
  double* h_coef = ev->h_io;
  double* h_out = ev->h_io + 2 * TPA_MAX_P;
  const bool shape = (flags & TPA_FLAG_SHAPE) != 0;
  // the folded passes take P <= 32 coefficients in their kernel arguments, and k_reduce writes
  // the sums into the pinned staging and raises a flag the host polls: no copy before or after
  // the pass and no stream synchronisation. The shape pass keeps the copies and the stream
  // wait (with the flag it ran 10-15% slower per call, measured; cause not isolated).
  static const bool io_env = [] {
    const char* v = std::getenv("TFPWA_EVAL_IO");
    return !(v && std::atoi(v) == 0);
  }();
  // (a pass that also returns f waits on the stream for that copy anyway: its sums take the
  // device-memory route, one copy, rather than a system-scope fence in every k_reduce block)
  const bool direct_io = io_env && !shape && !want_f;
  const bool args_coef = io_env && !shape && M.P <= TPA_EVAL_CIN && eval_unroll() == 0;
  if (!args_coef) {
    std::memcpy(h_coef, coef, sizeof(double2) * M.P);
    HIPCHK(hipMemcpyAsync(ev->d_coef, h_coef, sizeof(double2) * M.P, hipMemcpyHostToDevice, s));
  }
  const bool f32 = (flags & TPA_FLAG_F32) != 0;
  if (f32 && shape) return fail(TPA_ERR_UNSUPPORTED, "FP32 cache reads need a folded cache");
  if (f32 && !ev->h32_valid)
    return fail(TPA_ERR_ARG, "no FP32 copy of the cache (tpa_events_keep_f32)");
  const bool half32 = (flags & TPA_FLAG_F32H) != 0, half64 = (flags & TPA_FLAG_F64H) != 0;
  const bool half = half32 || half64;
  if (half && (shape || f32 || (half32 && half64)))
    return fail(TPA_ERR_UNSUPPORTED,
                "TPA_FLAG_F32H / TPA_FLAG_F64H exclude TPA_FLAG_SHAPE, TPA_FLAG_F32 and each other");
  if ((half32 && !ev->h32h_valid) || (half64 && !ev->h64h_valid))
    return fail(TPA_ERR_ARG, "no helicity-halved copy of the cache (tpa_events_keep_half)");
  if (half && !args_coef)
    return fail(TPA_ERR_UNSUPPORTED, "the helicity-halved passes take at most 32 coefficients");
  EvalArgs a;
  a.h = half64 ? ev->d_g64h : ev->d_h;
  a.h32 = f32 ? ev->d_h32 : half32 ? ev->d_g32h : nullptr;
  a.rho = half32 ? ev->d_rho : nullptr;
  a.rho64 = half64 ? ev->d_rho64 : nullptr;
  a.w = ev->d_w;
  a.coef = args_coef ? nullptr : ev->d_coef;
  if (args_coef) std::memcpy(a.cin, coef, sizeof(double2) * M.P);
  if (half) {
    // c~_k = phi_k c_k: phi = -i multiplies (re, im) into (im, -re)
    for (int k = 0; k < M.P; ++k) {
      if (ev->half_rot[k]) a.cin[k] = make_double2(coef[2 * k + 1], -coef[2 * k]);
      a.kch[k] = ev->half_kch[k];
    }
  }
  a.partials = ev->d_part;
  a.f_out = want_f ? ev->d_f : nullptr;
  a.N = ev->n;
  a.P = M.P;
  a.mode = mode;
  a.grad = (flags & TPA_FLAG_GRAD) ? 1 : 0;
  if (shape && ev->m->kind == 4)
    return fail(TPA_ERR_UNSUPPORTED, "floating lineshapes need a three-body model");
  if (shape && ev->built_fold != 0)
    return fail(TPA_ERR_ARG, "shape evaluation needs a cache built without lineshapes (fold=0)");
  if (!shape && ev->built_fold != 1)
    return fail(TPA_ERR_ARG, "evaluation without TPA_FLAG_SHAPE needs a folded cache (fold=1)");
  a.nacc = shape ? 2 + 6 * M.P : 2 + 2 * M.P;
  const int kg = eval_kgroups(M.lam_pad, M.P);
  if (kg < 0) return fail(TPA_ERR_UNSUPPORTED, "too many partial waves for this helicity count");
  const int kgh = half64 ? half64_kgroups(M.P) : half_kgroups(M.P);
  const int half_chains = ev->m->kind == 4 ? ev->m->h4.n_chain : M.n_chain;  // rows of rho
  const bool half_chain =
      half32 && ev->half_plan_pml > 0 && half_chain_fits(a, ev->half_nh, kgh, half_chains);
  const bool pipe32d = f32 && !shape && pipe32d_fits(a, M.lam_pad, kg);
  int grid = half_chain ? half_chain_grid(ev->n, ev->half_plan_pml)
             : pipe32d  ? pipe32d_grid(ev->n, M.P)
             : half     ? half_grid(ev->n, kgh)
                        : eval_grid(ev->n, f32 ? f32_lanes_per_event(M.lam_pad, kg) : M.lam_pad * kg);
  bool timing = false;
  if (!async) {
    std::lock_guard<std::mutex> lk(g_timing_mu);
    timing = timing_take_locked();
  }
  if (shape) {
    if (!ev->d_bw)
      HIPCHK(hipMalloc(&ev->d_bw, sizeof(double2) * 3 * (size_t)M.n_chain * ev->n));
    const int64_t lgrid = (ev->n * M.n_chain + 255) / 256;  // one thread per chain and event
    hipLaunchKernelGGL(k_lineshape, dim3((unsigned)lgrid), dim3(256), 0, s, ev->m->d, ev->d_ang,
                       ev->n, ev->d_bw);
    HIPCHK(hipGetLastError());
  }
  PassTimer pass_timer{ev->t0, ev->t1};
  if (timing) tl_pass_timer = &pass_timer;
  int rc;
  if (shape) {
    EvalShapeArgs sa;
    sa.a = a;
    sa.bw = ev->d_bw;
    sa.kchain = ev->m->d_kchain;
    sa.n_chain = M.n_chain;
    // the shape pass keeps four accumulators more per wave: at most 4 waves per lane
    // (PML = 4) while the lanes of an event fit in the wave; the reduction below sums the
    // partials of this launch's grid
    int kgs = kg;
    while (kgs * 4 < M.P && M.lam_pad * kgs * 2 <= 64) kgs <<= 1;
    grid = eval_grid(ev->n, M.lam_pad * kgs);
    rc = dispatch_shape(M.lam_pad, kgs, sa, grid, s);
  } else {
    if (half_chain) {
      std::memcpy(a.plan, ev->half_plan, sizeof(a.plan));
      std::memcpy(a.pch, ev->half_pch, sizeof(a.pch));
      launch_half_chain(a, ev->half_plan_pml, half_chains, grid, s);
      rc = TPA_OK;
    } else if (pipe32d) {
      launch_pipe32d(a, grid, s);
      rc = TPA_OK;
    } else {
      rc = half64 ? dispatch_half64(ev->half_nh, a, grid, s)
           : half ? dispatch_half(ev->half_nh, kgh, a, grid, s)
                  : dispatch_eval(M.lam_pad, kg, a, grid, s);
    }
  }
  // a pass that did not take the timer (none launched) is not counted
  if (timing && tl_pass_timer != nullptr) timing = false;
  tl_pass_timer = nullptr;
  if (rc != TPA_OK) return fail(rc, "no evaluation kernel for this helicity/partial-wave count");
  HIPCHK(hipGetLastError());
  const int nout = a.grad ? a.nacc : 2;
  // the sums land in pinned memory; without f_out the host polls the reduction's flag rather
  // than waiting on the stream (a bounded spin: after 2 s it waits on the stream, which also
  // reports a failed launch)
  const bool poll = direct_io && !want_f;
  if (poll && !ev->d_cnt) {
    HIPCHK(hipMalloc(&ev->d_cnt, sizeof(unsigned)));
    HIPCHK(hipMemsetAsync(ev->d_cnt, 0, sizeof(unsigned), s));
    HIPCHK(hipHostMalloc((void**)&ev->h_flag, 64, 0));
    *(volatile uint64_t*)ev->h_flag = 0;
  }
  const uint64_t seq = poll ? ++ev->seq : 0;
  hipLaunchKernelGGL(k_reduce, dim3(nout), dim3(256), 0, s, ev->d_part, grid, a.nacc, 1,
                     ev->d_out, direct_io ? h_out : nullptr, poll ? ev->d_cnt : nullptr,
                     poll ? (volatile uint64_t*)ev->h_flag : nullptr, seq);
  HIPCHK(hipGetLastError());
  if (!direct_io)
    HIPCHK(hipMemcpyAsync(h_out, ev->d_out, sizeof(double) * nout, hipMemcpyDeviceToHost, s));
  if (want_f)
    HIPCHK(hipMemcpyAsync(f_out, ev->d_f, sizeof(double) * ev->n, hipMemcpyDeviceToHost, s));
  if (async) {
    ev->pending_nout = nout;
    ev->pending_stream = s;
    return TPA_OK;
  }
  bool done = false;
  if (poll) {
    const volatile uint64_t* flag = ev->h_flag;
    const auto t_start = std::chrono::steady_clock::now();
    for (unsigned it = 0; !done; ++it) {
      if (*flag == seq) {
        done = true;
      } else if ((it & 1023u) == 1023u &&
                 std::chrono::steady_clock::now() - t_start > std::chrono::seconds(2)) {
        break;
      }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
  }
  if (!done) HIPCHK(hipStreamSynchronize(s));
  std::memcpy(out, h_out, sizeof(double) * nout);
  if (half && a.grad) {
    // G_k = phi_k G~_k: phi = -i maps (re, im) to (im, -re)
    for (int k = 0; k < M.P; ++k)
      if (ev->half_rot[k]) {
        const double re = out[2 + 2 * k], im = out[3 + 2 * k];
        out[2 + 2 * k] = im;
        out[3 + 2 * k] = -re;
      }
  }
  if (timing) {
    HIPCHK(hipEventSynchronize(ev->t1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, ev->t0, ev->t1));
    std::lock_guard<std::mutex> lk(g_timing_mu);
    g_timing_ms += ms;
    g_timing_n += 1;
  }
  return TPA_OK;
}

int tpa_eval(tpa_events_t ev, const double* coef, int mode, int flags, double* out,
             double* f_out) {
  if (!ev) return fail(TPA_ERR_ARG, "null pointer");
  return eval_impl(ev, coef, mode, flags, out, f_out, ev->m->stream, false);
}

// Per-call rebuild: the partial waves of `chunk` events at a time built from the angle cache
// into one chunk buffer (k_build_pw / k_build_pw4, lineshapes folded at the model's current
// parameters), each chunk streamed by the evaluation pass and reduced into its own slot, the
// chunks' sums added on the host in chunk order (compensated for the value). Needs only the
// angle cache and the weights resident.
#define TPA_RECOMPUTE_CHUNK 262144
int tpa_eval_recompute(tpa_events_t ev, const double* coef, int mode, int flags, int64_t chunk,
                       double* out) {
  if (!ev || !coef || !out) return fail(TPA_ERR_ARG, "null pointer");
  if (mode != TPA_MODE_DATA && mode != TPA_MODE_MC) return fail(TPA_ERR_ARG, "bad mode");
  if (flags & ~TPA_FLAG_GRAD)
    return fail(TPA_ERR_UNSUPPORTED, "tpa_eval_recompute takes TPA_FLAG_GRAD only");
  tpa_model_s* m = ev->m;
  const TpaModel& M = m->h;
  if (M.P > TPA_EVAL_CIN)
    return fail(TPA_ERR_UNSUPPORTED, "tpa_eval_recompute serves at most 32 partial waves");
  if (ev->n < 1) return fail(TPA_ERR_ARG, "empty event set");
  const int kg = eval_kgroups(M.lam_pad, M.P);
  if (kg < 0) return fail(TPA_ERR_UNSUPPORTED, "too many partial waves for this helicity count");
  if (chunk <= 0) chunk = TPA_RECOMPUTE_CHUNK;
  chunk = std::min<int64_t>(chunk, ev->n);
  const int64_t nch = (ev->n + chunk - 1) / chunk;
  const int nacc = 2 + 2 * M.P;
  const int grad = (flags & TPA_FLAG_GRAD) ? 1 : 0;
  const int nout = grad ? nacc : 2;
  hipStream_t s = m->stream;
  if (ev->rchunk_cap < chunk) {
    HIPCHK(hipStreamSynchronize(s));
    if (ev->d_rchunk) HIPCHK(hipFree(ev->d_rchunk));
    ev->d_rchunk = nullptr;
    ev->rchunk_cap = 0;
    HIPCHK(hipMalloc(&ev->d_rchunk, sizeof(double2) * (size_t)M.P * chunk * M.lam_pad));
    ev->rchunk_cap = chunk;
  }
  if (ev->rout_cap < nch * nacc) {
    HIPCHK(hipStreamSynchronize(s));
    if (ev->d_rout) HIPCHK(hipFree(ev->d_rout));
    ev->d_rout = nullptr;
    ev->rout_cap = 0;
    HIPCHK(hipMalloc(&ev->d_rout, sizeof(double) * (size_t)nch * nacc));
    ev->rout_cap = nch * nacc;
  }
  bool timing;
  {
    std::lock_guard<std::mutex> lk(g_timing_mu);
    timing = timing_take_locked();
  }
  if (timing) HIPCHK(hipEventRecord(ev->t0, s));
  for (int64_t c = 0; c < nch; ++c) {
    const int64_t off = c * chunk, cn = std::min<int64_t>(chunk, ev->n - off);
    // the chunk's rows have stride cn, the event count of the pass that reads them
    launch_build_pw(m, ev->d_ang, cn, off, ev->n, 1, ev->d_rchunk, s, cn, 0);
    HIPCHK(hipGetLastError());
    EvalArgs a{};
    a.h = ev->d_rchunk;
    a.h32 = nullptr;
    a.w = ev->d_w + off;
    a.coef = nullptr;
    std::memcpy(a.cin, coef, sizeof(double2) * M.P);
    a.partials = ev->d_part;
    a.f_out = nullptr;
    a.N = cn;
    a.P = M.P;
    a.mode = mode;
    a.grad = grad;
    a.nacc = nacc;
    a.rho = nullptr;
    a.rho64 = nullptr;
    const int grid = eval_grid(cn, M.lam_pad * kg);
    const int rc = dispatch_eval(M.lam_pad, kg, a, grid, s);
    if (rc != TPA_OK) return fail(rc, "no evaluation kernel for this helicity/partial-wave count");
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce, dim3(nout), dim3(256), 0, s, ev->d_part, grid, nacc, 1,
                       ev->d_rout + c * nacc, nullptr, nullptr, nullptr, (uint64_t)0);
    HIPCHK(hipGetLastError());
  }
  if (timing) HIPCHK(hipEventRecord(ev->t1, s));
  std::vector<double> r((size_t)nch * nacc);
  HIPCHK(hipMemcpyAsync(r.data(), ev->d_rout, sizeof(double) * r.size(), hipMemcpyDeviceToHost,
                        s));
  HIPCHK(hipStreamSynchronize(s));
  double vs = 0.0, vc = 0.0;
  for (int64_t c = 0; c < nch; ++c) {
    const double y = r[(size_t)c * nacc];
    const double t = vs + y;
    vc += std::fabs(vs) >= std::fabs(y) ? (vs - t) + y : (y - t) + vs;
    vs = t;
  }
  out[0] = vs + vc;
  out[1] = 0.0;
  for (int j = 2; j < nout; ++j) {
    double g = 0.0;
    for (int64_t c = 0; c < nch; ++c) g += r[(size_t)c * nacc + j];
    out[j] = g;
  }
  if (timing) {
    HIPCHK(hipEventSynchronize(ev->t1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, ev->t0, ev->t1));
    std::lock_guard<std::mutex> lk(g_timing_mu);
    g_timing_ms += ms;
    g_timing_n += 1;
  }
  return TPA_OK;
}

int tpa_eval_enqueue(tpa_events_t ev, const double* coef, int mode, int flags, void* stream) {
  if (!ev) return fail(TPA_ERR_ARG, "null pointer");
  return eval_impl(ev, coef, mode, flags, nullptr, nullptr,
                   stream ? (hipStream_t)stream : ev->m->stream, true);
}

int tpa_eval_collect(tpa_events_t ev, double* out) {
  if (!ev || !out) return fail(TPA_ERR_ARG, "null pointer");
  if (ev->pending_nout <= 0) return fail(TPA_ERR_ARG, "no queued evaluation");
  const int nout = ev->pending_nout;
  ev->pending_nout = 0;
  HIPCHK(hipStreamSynchronize(ev->pending_stream));
  std::memcpy(out, ev->h_io + 2 * TPA_MAX_P, sizeof(double) * nout);
  return TPA_OK;
}

}  // extern "C"

// ---- the fused pass (tpa_eval_fused): waves recomputed from the angles on every evaluation
static double host_rel_p(double m0, double m1, double m2) {
  const double s = m1 + m2, d = m1 - m2;
  const double me = m0 > s ? m0 : s;
  const double p = (me - s) * (me + s) * (me - d) * (me + d);
  return std::sqrt(p) / (2.0 * me);
}

// Compile (hipRTC, gfx950) and load the model's fused kernel if it is missing or its
// description changed; m->fused_mu held.
static int fused_ensure(tpa_model_s* m) {
  if (m->fused_fn && !m->fused_stale) return TPA_OK;
  std::string src = tpa_fused::fused_source(m->h, m->cg_host.data(), (int)m->cg_host.size());
  m->fused_stale = false;
  if (m->fused_fn && src == m->fused_src) return TPA_OK;
  const auto t0 = std::chrono::steady_clock::now();
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "tpa_fused.hip", kRtcHeaderCount, kRtcHeaders,
                          kRtcHeaderNames) != HIPRTC_SUCCESS)
    return fail(TPA_ERR_HIP, "hiprtcCreateProgram failed");
  // packed FP32 instructions stay off here too (the pass is FP64, so none are expected)
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang",
                        "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, 7, opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    return fail(TPA_ERR_HIP, "fused pass compile failed: " + log.substr(0, 2000));
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  std::vector<char> code(cs);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  HIPCHK(hipModuleLoadData(&mod, code.data()));
  hipError_t e = hipModuleGetFunction(&fn, mod, "tpa_fused_pass");
  if (e != hipSuccess) {
    (void)hipModuleUnload(mod);
    return fail(TPA_ERR_HIP, std::string("hipModuleGetFunction: ") + hipGetErrorString(e));
  }
  if (m->fused_mod) (void)hipModuleUnload(m->fused_mod);
  m->fused_mod = mod;
  m->fused_fn = fn;
  m->fused_src = std::move(src);
  m->fused_compile_s =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return TPA_OK;
}

static int eval_fused_impl(tpa_events_s* ev, const double* coef, int mode, int flags,
                           double* out) {
  tpa_model_s* m = ev->m;
  if (m->kind != 3) return fail(TPA_ERR_UNSUPPORTED, "the fused pass covers three-body models");
  if (mode != TPA_MODE_DATA && mode != TPA_MODE_MC) return fail(TPA_ERR_ARG, "bad mode");
  if (flags & ~TPA_FLAG_GRAD)
    return fail(TPA_ERR_UNSUPPORTED, "tpa_eval_fused takes TPA_FLAG_GRAD only");
  const TpaModel& M = m->h;
  // The generated kernel of the 22-wave Lambda_c model took a memory-aperture fault on gfx950;
  // its addresses, kernel descriptor and launch arguments check out and the cause is not
  // identified, so models above the size that has run clean stay refused.
  if (M.P > TPA_FUSED_MAX_P)
    return fail(TPA_ERR_UNSUPPORTED,
                "the fused pass is disabled above " + std::to_string(TPA_FUSED_MAX_P) +
                    " partial waves (a 22-wave model's kernel faulted, cause not identified)");
  {
    std::lock_guard<std::mutex> lk(m->fused_mu);
    const int rc = fused_ensure(m);
    if (rc != TPA_OK) return rc;
  }
  const int P = M.P, NC = M.n_chain;
  const tpa_fused::ArgsLayout L = tpa_fused::fused_args_layout(P, NC);
  std::vector<unsigned char> buf(L.size, 0);
  auto put = [&](size_t off, const void* src, size_t n) { std::memcpy(buf.data() + off, src, n); };
  const double* ang = ev->d_ang;
  const double* w = ev->d_w;
  double* part = ev->d_part;
  const long long N = ev->n, S = ev->n;
  put(0, &ang, 8);
  put(8, &w, 8);
  put(16, &part, 8);
  put(24, &N, 8);
  put(32, &S, 8);
  put(L.off_c, coef, sizeof(double) * 2 * P);
  for (int c = 0; c < NC; ++c) {
    const TpaChain& C = M.chain[c];
    const double q0 = host_rel_p(C.m0, C.m_y, C.m_z);
    int act = 0;
    for (int k = 0; k < C.n_ls1; ++k) act |= (C.act1[k] ? 1 : 0) << k;
    for (int k = 0; k < C.n_ls2; ++k) act |= (C.act2[k] ? 1 : 0) << (16 + k);
    put(L.off_m0 + 8 * c, &C.m0, 8);
    put(L.off_g0 + 8 * c, &C.g0, 8);
    put(L.off_q0 + 8 * c, &q0, 8);
    put(L.off_act + 4 * c, &act, 4);
  }
  const int grad = (flags & TPA_FLAG_GRAD) ? 1 : 0;
  put(L.off_mode, &mode, 4);
  put(L.off_grad, &grad, 4);
  const int nacc = 2 + 2 * P;
  const int64_t blocks = (ev->n + 255) / 256;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(blocks, kEvalGridDefault));
  hipStream_t s = m->stream;
  bool timing;
  {
    std::lock_guard<std::mutex> lk(g_timing_mu);
    timing = timing_take_locked();
  }
  if (timing) HIPCHK(hipEventRecord(ev->t0, s));
  void* params[] = {buf.data()};
  HIPCHK(hipModuleLaunchKernel(m->fused_fn, grid, 1, 1, 256, 1, 1, 0, s, params, nullptr));
  if (timing) HIPCHK(hipEventRecord(ev->t1, s));
  const int nout = grad ? nacc : 2;
  hipLaunchKernelGGL(k_reduce, dim3(nout), dim3(256), 0, s, part, grid, nacc, 1, ev->d_out,
                     nullptr, nullptr, nullptr, (uint64_t)0);
  HIPCHK(hipGetLastError());
  if (!ev->h_io) HIPCHK(hipHostMalloc((void**)&ev->h_io, sizeof(double) * TPA_IO_DOUBLES));
  double* h_out = ev->h_io + 2 * TPA_MAX_P;
  HIPCHK(hipMemcpyAsync(h_out, ev->d_out, sizeof(double) * nout, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  std::memcpy(out, h_out, sizeof(double) * nout);
  if (timing) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, ev->t0, ev->t1));
    std::lock_guard<std::mutex> lk(g_timing_mu);
    g_timing_ms += ms;
    g_timing_n += 1;
  }
  return TPA_OK;
}

extern "C" {

int tpa_fused_source(const void* desc, int64_t desc_bytes, const double* cg, int n_cg,
                     char* buf, int64_t buf_len, int64_t* len) {
  if (!desc || !len || (n_cg > 0 && !cg)) return fail(TPA_ERR_ARG, "null pointer");
  if (desc_bytes != (int64_t)sizeof(TpaModel))
    return fail(TPA_ERR_ARG, "model description size mismatch");
  TpaModel M;
  std::memcpy(&M, desc, sizeof(TpaModel));
  const std::string why = check_model(M, n_cg);
  if (!why.empty()) return fail(TPA_ERR_ARG, "invalid model: " + why);
  const std::string src = tpa_fused::fused_source(M, cg, n_cg);
  *len = (int64_t)src.size();
  if (buf && buf_len > 0) {
    const size_t n = std::min<size_t>(src.size(), (size_t)buf_len - 1);
    std::memcpy(buf, src.data(), n);
    buf[n] = '\0';
  }
  return TPA_OK;
}

int tpa_eval_fused(tpa_events_t ev, const double* coef, int mode, int flags, double* out) {
  if (!ev || !coef || !out) return fail(TPA_ERR_ARG, "null pointer");
  return eval_fused_impl(ev, coef, mode, flags, out);
}

int tpa_fused_compile_seconds(tpa_model_t m, double* s) {
  if (!m || !s) return fail(TPA_ERR_ARG, "null pointer");
  std::lock_guard<std::mutex> lk(m->fused_mu);
  *s = m->fused_compile_s;
  return TPA_OK;
}

}  // extern "C"

// ---- the allocation-free pass (tpa_eval_noalloc)
// Per-block partial sums of the pass live in this file-scope device array instead of the event
// set's buffer: kNoAllocBlocks blocks of at most kNoAllocSlots slots (2 + 2 * 32 doubles), slot-
// major as k_reduce reads them, 528 KiB on every GPU the library runs on, held until the process
// ends. Every block of a pass writes each slot that k_reduce then reads, so nothing needs to
// start at zero. In host code `scratch` names a host-side stand-in: kernels get the device
// address that hipGetSymbolAddress returns for the current device (noalloc_scratch).
constexpr int kNoAllocBlocks = kEvalGridDefault;
constexpr int kNoAllocSlots = 2 + 2 * TPA_EVAL_CIN;
static __device__ __align__(256) double scratch[kNoAllocBlocks * kNoAllocSlots];

constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kNoAllocMaxDevices] = {};

// Device address of `scratch` on device `dev` (current when called), resolved once per device.
static int noalloc_scratch(int dev, double** ptr) {
  if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(TPA_ERR_UNSUPPORTED, "device index");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(TPA_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                   hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *ptr = g_scratch_ptr[dev];
  return TPA_OK;
}

static int eval_noalloc_impl(tpa_events_s* ev, const double* coef, int mode, int flags,
                             double* d_out, hipStream_t s) {
  const TpaModel& M = ev->m->h;
  if (mode != TPA_MODE_DATA && mode != TPA_MODE_MC) return fail(TPA_ERR_ARG, "bad mode");
  if (flags & ~(TPA_FLAG_GRAD | TPA_FLAG_F32))
    return fail(TPA_ERR_UNSUPPORTED, "tpa_eval_noalloc takes TPA_FLAG_GRAD and TPA_FLAG_F32 only");
  if (M.P > TPA_EVAL_CIN)
    return fail(TPA_ERR_UNSUPPORTED, "tpa_eval_noalloc: P = " + std::to_string(M.P) +
                                         " partial waves exceed its maximum of " +
                                         std::to_string(TPA_EVAL_CIN));
  if (eval_unroll() != 0)
    return fail(TPA_ERR_UNSUPPORTED, "tpa_eval_noalloc runs the default pass only "
                                     "(TFPWA_EVAL_UNROLL is set)");
  if (!ev->d_h || ev->built_fold != 1)
    return fail(TPA_ERR_ARG, "tpa_eval_noalloc needs a folded cache (tpa_events_build fold=1)");
  const bool f32 = (flags & TPA_FLAG_F32) != 0;
  if (f32 && !ev->h32_valid)
    return fail(TPA_ERR_ARG, "no FP32 copy of the cache (tpa_events_keep_f32)");
  const int kg = eval_kgroups(M.lam_pad, M.P);
  if (kg < 0) return fail(TPA_ERR_UNSUPPORTED, "too many partial waves for this helicity count");
  // the same grid as tpa_eval, so both passes sum in the same order; it is at most
  // kEvalGridDefault blocks unless TFPWA_EVAL_GRID raises it past the scratch
  const int grid = eval_grid(ev->n, f32 ? f32_lanes_per_event(M.lam_pad, kg) : M.lam_pad * kg);
  if (grid > kNoAllocBlocks)
    return fail(TPA_ERR_UNSUPPORTED, "tpa_eval_noalloc: a grid of " + std::to_string(grid) +
                                         " blocks exceeds its scratch of " +
                                         std::to_string(kNoAllocBlocks));
  double* part = nullptr;
  int rc = noalloc_scratch(ev->m->device, &part);
  if (rc != TPA_OK) return rc;
  EvalArgs a;
  a.h = ev->d_h;
  a.h32 = f32 ? ev->d_h32 : nullptr;
  a.w = ev->d_w;
  a.coef = nullptr;  // coefficients travel in the kernel arguments
  std::memcpy(a.cin, coef, sizeof(double2) * M.P);
  a.partials = part;
  a.f_out = nullptr;
  a.N = ev->n;
  a.P = M.P;
  a.mode = mode;
  a.grad = (flags & TPA_FLAG_GRAD) ? 1 : 0;
  a.nacc = 2 + 2 * M.P;
  rc = dispatch_eval(M.lam_pad, kg, a, grid, s);
  if (rc != TPA_OK) return fail(rc, "no evaluation kernel for this helicity/partial-wave count");
  HIPCHK(hipGetLastError());
  const int nout = a.grad ? a.nacc : 2;
  hipLaunchKernelGGL(k_reduce, dim3(nout), dim3(256), 0, s, part, grid, a.nacc, 1, d_out,
                     nullptr, nullptr, nullptr, (uint64_t)0);
  HIPCHK(hipGetLastError());
  return TPA_OK;
}

extern "C" {

int tpa_eval_noalloc(tpa_events_t ev, const double* coef, int mode, int flags, double* d_out,
                     void* stream) {
  if (!ev || !coef || !d_out) return fail(TPA_ERR_ARG, "null pointer");
  // the scratch array of the event set's device, with that device current for the launches
  int cur = 0;
  HIPCHK(hipGetDevice(&cur));
  const int dev = ev->m->device;
  if (cur != dev) HIPCHK(hipSetDevice(dev));
  const int rc = eval_noalloc_impl(ev, coef, mode, flags, d_out,
                                   stream ? (hipStream_t)stream : ev->m->stream);
  if (cur != dev) {
    const hipError_t e = hipSetDevice(cur);
    if (rc == TPA_OK && e != hipSuccess)
      return fail(TPA_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  }
  return rc;
}

int tpa_model_stream(tpa_model_t m, void** stream) {
  if (!m || !stream) return fail(TPA_ERR_ARG, "null pointer");
  *stream = (void*)m->stream;
  return TPA_OK;
}

int tpa_gram(tpa_events_t ev, double w_mean, double* out, double* ms) {
  if (!ev || !out) return fail(TPA_ERR_ARG, "null pointer");
  if (!(w_mean > 0.0) && !(w_mean < 0.0)) return fail(TPA_ERR_ARG, "w_mean must be nonzero");
  const double wscale = 1.0 / w_mean;
  if (!ev->d_h) return fail(TPA_ERR_ARG, "partial-wave cache not built");
  const TpaModel& M = ev->m->h;
  if (M.P > 32) return fail(TPA_ERR_UNSUPPORTED, "the MFMA Gram supports up to 32 partial waves");
  hipStream_t s = ev->m->stream;
  const int64_t rows = ev->n * (int64_t)M.lam_pad;
  const int grid = (int)std::min<int64_t>((rows + TPA_GRAM_ROWS - 1) / TPA_GRAM_ROWS, 1024);
  if (!ev->d_gpart) HIPCHK(hipMalloc(&ev->d_gpart, sizeof(double) * 1024 * TPA_GRAM_SLOTS));
  if (!ev->d_gout) HIPCHK(hipMalloc(&ev->d_gout, sizeof(double) * TPA_GRAM_SLOTS));
  HIPCHK(hipEventRecord(ev->t0, s));
  static const int refold = [] {
    const char* v = std::getenv("TFPWA_GRAM_REFOLD");
    const int x = v ? std::atoi(v) : 8;
    return (x == 1 || x == 2 || x == 4 || x == 8) ? x : 8;
  }();
  switch (refold) {
    case 1:
      hipLaunchKernelGGL(k_gram_f32<1>, dim3(grid), dim3(256), 0, s, ev->d_h, ev->d_w, wscale,
                         ev->n, M.lam_pad, M.lam, M.P, ev->d_gpart);
      break;
    case 2:
      hipLaunchKernelGGL(k_gram_f32<2>, dim3(grid), dim3(256), 0, s, ev->d_h, ev->d_w, wscale,
                         ev->n, M.lam_pad, M.lam, M.P, ev->d_gpart);
      break;
    case 4:
      hipLaunchKernelGGL(k_gram_f32<4>, dim3(grid), dim3(256), 0, s, ev->d_h, ev->d_w, wscale,
                         ev->n, M.lam_pad, M.lam, M.P, ev->d_gpart);
      break;
    default:
      hipLaunchKernelGGL(k_gram_f32<8>, dim3(grid), dim3(256), 0, s, ev->d_h, ev->d_w, wscale,
                         ev->n, M.lam_pad, M.lam, M.P, ev->d_gpart);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ev->t1, s));
  hipLaunchKernelGGL(k_reduce_slots, dim3(TPA_GRAM_SLOTS), dim3(256), 0, s, ev->d_gpart, grid,
                     TPA_GRAM_SLOTS, ev->d_gout);
  HIPCHK(hipGetLastError());
  std::vector<double> raw(TPA_GRAM_SLOTS);
  HIPCHK(hipMemcpyAsync(raw.data(), ev->d_gout, sizeof(double) * TPA_GRAM_SLOTS,
                        hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (ms) {
    float f = 0.0f;
    HIPCHK(hipEventElapsedTime(&f, ev->t0, ev->t1));
    *ms = f;
  }
  // decode the 32x32 MFMA accumulator layout: element (i, j) of tile q sits in lane
  // j + 32 * ((i / 4) % 2), accumulator register (i % 4) + 4 * (i / 8)
  auto G = [&](int q, int i, int j) {
    const int lane = j + 32 * ((i / 4) % 2);
    const int v = (i % 4) + 4 * (i / 8);
    return raw[(q * 16 + v) * 64 + lane];
  };
  const int P = M.P;
  for (int k = 0; k < P; ++k)
    for (int l = 0; l < P; ++l) {
      // the two accumulator positions of a symmetric entry round differently in FP32;
      // their mean makes I exactly Hermitian
      out[2 * (k * P + l)] =
          w_mean * (0.5 * ((G(0, k, l) + G(2, k, l)) + (G(0, l, k) + G(2, l, k))));
      out[2 * (k * P + l) + 1] = w_mean * (G(1, k, l) - G(1, l, k));
    }
  return TPA_OK;
}

int tpa_gram_f64(tpa_events_t ev, double* out, double* ms) {
  if (!ev || !out) return fail(TPA_ERR_ARG, "null pointer");
  if (!ev->d_h) return fail(TPA_ERR_ARG, "partial-wave cache not built");
  const TpaModel& M = ev->m->h;
  const int P = M.P, lam_pad = M.lam_pad, lam = M.lam;
  if (P > 32) return fail(TPA_ERR_UNSUPPORTED, "the FP64 MFMA Gram supports up to 32 partial waves");
  const int QT = (2 * P + 15) / 16;
  hipStream_t s = ev->m->stream;
  const int64_t rows = ev->n * (int64_t)lam_pad;
  const int grid = (int)std::min<int64_t>((rows + TPA_GRAM64_ROWS - 1) / TPA_GRAM64_ROWS, 1024);
  if (!ev->d_gpart) HIPCHK(hipMalloc(&ev->d_gpart, sizeof(double) * 1024 * TPA_GRAM_SLOTS));
  if (!ev->d_gout) HIPCHK(hipMalloc(&ev->d_gout, sizeof(double) * TPA_GRAM_SLOTS));
  HIPCHK(hipEventRecord(ev->t0, s));
  switch (QT) {
    case 1:
      hipLaunchKernelGGL(k_gram_f64<1>, dim3(grid), dim3(256), 0, s, ev->d_h, ev->d_w, ev->n,
                         lam_pad, lam, P, ev->d_gpart);
      break;
    case 2:
      hipLaunchKernelGGL(k_gram_f64<2>, dim3(grid), dim3(256), 0, s, ev->d_h, ev->d_w, ev->n,
                         lam_pad, lam, P, ev->d_gpart);
      break;
    case 3:
      hipLaunchKernelGGL(k_gram_f64<3>, dim3(grid), dim3(256), 0, s, ev->d_h, ev->d_w, ev->n,
                         lam_pad, lam, P, ev->d_gpart);
      break;
    default:
      hipLaunchKernelGGL(k_gram_f64<4>, dim3(grid), dim3(256), 0, s, ev->d_h, ev->d_w, ev->n,
                         lam_pad, lam, P, ev->d_gpart);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ev->t1, s));
  // only the NT = QT (QT + 1) / 2 tiles' 256 slots each are written
  hipLaunchKernelGGL(k_reduce_slots, dim3(QT * (QT + 1) / 2 * 256), dim3(256), 0, s, ev->d_gpart,
                     grid, TPA_GRAM64_SLOTS, ev->d_gout);
  HIPCHK(hipGetLastError());
  std::vector<double> raw(TPA_GRAM64_SLOTS);
  HIPCHK(hipMemcpyAsync(raw.data(), ev->d_gout, sizeof(double) * TPA_GRAM64_SLOTS,
                        hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (ms) {
    float f = 0.0f;
    HIPCHK(hipEventElapsedTime(&f, ev->t0, ev->t1));
    *ms = f;
  }
  // decode the 16x16 FP64 MFMA accumulator layout: element (i, j) of a tile sits in lane
  // j + 16 * (i % 4), accumulator register i / 4; tiles q enumerate (ti <= tj) row-major
  std::vector<int> tile(QT * QT, -1);
  for (int ti = 0, q = 0; ti < QT; ++ti)
    for (int tj = ti; tj < QT; ++tj, ++q) tile[ti * QT + tj] = q;
  auto G = [&](int a, int b) {
    if (a / 16 > b / 16) std::swap(a, b);
    const int q = tile[(a / 16) * QT + b / 16];
    const int i = a % 16, j = b % 16;
    return raw[(q * 4 + i / 4) * 64 + j + 16 * (i % 4)];
  };
  for (int k = 0; k < P; ++k)
    for (int l = 0; l < P; ++l) {
      // the two positions of a symmetric entry round differently; their mean makes I
      // exactly Hermitian
      out[2 * (k * P + l)] =
          0.5 * ((G(k, l) + G(P + k, P + l)) + (G(l, k) + G(P + l, P + k)));
      out[2 * (k * P + l) + 1] = 0.5 * ((G(k, P + l) - G(P + k, l)) - (G(l, P + k) - G(P + l, k)));
    }
  return TPA_OK;
}

// ------------------------------------------------------------------------ serving pool
// FP32 copy of the pool events [off, off + n) of every partial-wave row.
static void launch_rows_to_f32(tpa_pool_s* p, int64_t off, int64_t n, hipStream_t s) {
  const TpaModel& M = p->m->h;
  const int64_t total = (int64_t)M.P * n * M.lam_pad;
  const int64_t grid = std::min<int64_t>((total + 255) / 256, 8192);
  hipLaunchKernelGGL(k_pw_rows_to_f32, dim3((unsigned)grid), dim3(256), 0, s, p->d_h, p->d_h32,
                     M.P, p->cap * M.lam_pad, off * M.lam_pad, n * M.lam_pad);
}

int tpa_pool_destroy(tpa_pool_t p) {
  if (!p) return TPA_OK;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  for (auto& q : p->slot) {
    for (auto& g : q.graphs) (void)hipGraphExecDestroy(g.second);
    void* dev[] = {q.d_blob, q.d_part, q.d_out};
    for (void* x : dev)
      if (x) (void)hipFree(x);
    void* host[] = {q.h_blob, q.h_out};
    for (void* x : host)
      if (x) (void)hipHostFree(x);
    hipEvent_t evs[] = {q.t0, q.t1, q.done};
    for (hipEvent_t x : evs)
      if (x) (void)hipEventDestroy(x);
  }
  void* dev[] = {p->d_ang, p->d_h, p->d_w, p->d_h32};
  for (void* x : dev)
    if (x) (void)hipFree(x);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  if (p->add_stream) (void)hipStreamDestroy(p->add_stream);
  if (p->d_stage) (void)hipFree(p->d_stage);
  if (p->h_stage) (void)hipHostFree(p->h_stage);
  if (p->d_hpart) (void)hipFree(p->d_hpart);
  if (p->d_hout) (void)hipFree(p->d_hout);
  if (p->d_hcoef) (void)hipFree(p->d_hcoef);
  delete p;
  return TPA_OK;
}

int tpa_pool_create(tpa_model_t m, int64_t capacity, int max_items, tpa_pool_t* out) {
  if (!m || !out) return fail(TPA_ERR_ARG, "null pointer");
  if (capacity < 1 || max_items < 1) return fail(TPA_ERR_ARG, "bad pool size");
  // The capacity is also the row stride of every per-event array of the pool ([row][cap]):
  // rounded up to the range granule it puts each row on a 512-byte boundary. With an odd stride
  // the rows of the FP32 copy (32 B per event) start off the 128-byte line grid, each wave's
  // 256-byte group spans three lines and a full batch reads up to 1.24x its bytes.
  capacity = (capacity + TPA_POOL_ALIGN - 1) / TPA_POOL_ALIGN * TPA_POOL_ALIGN;
  const TpaModel& M = m->h;
  tpa_pool_s* p = new tpa_pool_s();
  p->m = m;
  p->cap = capacity;
  p->max_items = max_items;
  p->max_chunks = (int)std::min<int64_t>((capacity + TPA_CHUNK - 1) / TPA_CHUNK + max_items,
                                         1 << 22);
  const int nf = angle_fields(m);
  const int nacc = 2 + 2 * M.P;
  hipError_t e = hipSuccess;
  auto D = [&](void** ptr, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(ptr, bytes);
  };
  auto H = [&](void** ptr, size_t bytes) {
    if (e == hipSuccess) e = hipHostMalloc(ptr, bytes, 0);
  };
  D((void**)&p->d_ang, sizeof(double) * nf * capacity);
  D((void**)&p->d_h, sizeof(double2) * (size_t)M.P * capacity * M.lam_pad);
  D((void**)&p->d_w, sizeof(double) * capacity);
  // blob words: header 2, three chunk arrays, item_chunk m + 1, one alignment word, 2 P per request
  const size_t blob_words = 2 + 3 * (size_t)p->max_chunks + (size_t)max_items + 2 +
                            2 * (size_t)max_items * M.P;
  for (auto& q : p->slot) {
    D((void**)&q.d_blob, sizeof(int64_t) * blob_words);
    D((void**)&q.d_part, sizeof(double) * (size_t)p->max_chunks * nacc);
    D((void**)&q.d_out, sizeof(double) * (size_t)max_items * nacc);
    H((void**)&q.h_blob, sizeof(int64_t) * blob_words);
    H((void**)&q.h_out, sizeof(double) * (size_t)max_items * nacc);
    if (e == hipSuccess) e = hipEventCreate(&q.t0);
    if (e == hipSuccess) e = hipEventCreate(&q.t1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&q.done, hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->add_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    tpa_pool_destroy(p);
    return fail(TPA_ERR_HIP, std::string("pool allocation: ") + hipGetErrorString(e));
  }
  p->free_ranges[0] = capacity;
  *out = p;
  return TPA_OK;
}

int64_t tpa_pool_free_events(tpa_pool_t p) {
  if (!p) return -1;
  std::lock_guard<std::mutex> lk(p->mu);
  int64_t n = 0;
  for (auto& r : p->free_ranges) n += r.second;
  return n;
}

}  // extern "C"

// First free range of at least n events, taken from the free list (p->mu held): its offset,
// or -1 when none is long enough. Ranges are whole multiples of TPA_POOL_ALIGN events, like
// the capacity (the row stride), so every range and every 2048-event chunk starts on a
// cache-line boundary of each partial-wave row: the streaming loads of two chunks never share
// a line.
static int64_t pool_take_range(tpa_pool_s* p, int64_t n) {
  auto it = p->free_ranges.begin();
  for (; it != p->free_ranges.end(); ++it)
    if (it->second >= n) break;
  if (it == p->free_ranges.end()) return -1;
  const int64_t off = it->first;
  const int64_t len = it->second;
  const int64_t take =
      std::min<int64_t>(len, (n + TPA_POOL_ALIGN - 1) / TPA_POOL_ALIGN * TPA_POOL_ALIGN);
  p->free_ranges.erase(it);
  if (len > take) p->free_ranges[off + take] = len - take;
  p->used_ranges[off] = take;
  return off;
}

extern "C" {

int tpa_pool_add(tpa_pool_t p, const double* p4, int64_t n, const double* weight,
                 int64_t* offset) {
  if (!p || !p4 || !weight || !offset) return fail(TPA_ERR_ARG, "null pointer");
  if (n < 1) return fail(TPA_ERR_ARG, "event count must be positive");
  std::lock_guard<std::mutex> add_lk(p->add_mu);
  int64_t off = 0;
  {
    std::lock_guard<std::mutex> lk(p->mu);
    off = pool_take_range(p, n);
    if (off < 0) return fail(TPA_ERR_ARG, "pool full");
  }
  auto undo = [&](hipError_t e, const char* what) {
    std::lock_guard<std::mutex> lk(p->mu);
    release_range(p, off);
    return fail(TPA_ERR_HIP, std::string("pool add: ") + what + ": " + hipGetErrorString(e));
  };
  hipError_t e = hipSuccess;
  if (p->stage_n < n) {  // grow the staging buffers (rare: sizes only increase)
    const int64_t want = std::max<int64_t>(n, p->stage_n + p->stage_n / 2);
    if (p->d_stage) (void)hipFree(p->d_stage);
    if (p->h_stage) (void)hipHostFree(p->h_stage);
    p->d_stage = nullptr;
    p->h_stage = nullptr;
    p->stage_n = 0;
    if ((e = hipMalloc(&p->d_stage, sizeof(double) * 17 * want)) != hipSuccess)
      return undo(e, "staging");
    if ((e = hipHostMalloc(&p->h_stage, sizeof(double) * 17 * want, 0)) != hipSuccess)
      return undo(e, "pinned staging");
    p->stage_n = want;
  }
  tpa_model_s* m = p->m;
  const int rows = p4_rows(m);
  std::memcpy(p->h_stage, p4, sizeof(double) * rows * n);
  std::memcpy(p->h_stage + rows * n, weight, sizeof(double) * n);
  hipStream_t s = p->add_stream;
  if ((e = hipMemcpyAsync(p->d_stage, p->h_stage, sizeof(double) * (rows + 1) * n,
                          hipMemcpyHostToDevice, s)) != hipSuccess)
    return undo(e, "upload");
  if ((e = hipMemcpyAsync(p->d_w + off, p->d_stage + rows * n, sizeof(double) * n,
                          hipMemcpyDeviceToDevice, s)) != hipSuccess)
    return undo(e, "weights");
  launch_cal_angle(m, p->d_stage, n, off, p->cap, p->d_ang, s);
  launch_build_pw(m, p->d_ang, n, off, p->cap, 1, p->d_h, s);
  if (p->d_h32) launch_rows_to_f32(p, off, n, s);
  if ((e = hipGetLastError()) != hipSuccess) return undo(e, "launch");
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return undo(e, "build");
  *offset = off;
  return TPA_OK;
}

int tpa_pool_add_events(tpa_pool_t p, tpa_events_t ev, const double* weight,
                        int64_t* offset) {
  if (!p || !ev || !weight || !offset) return fail(TPA_ERR_ARG, "null pointer");
  if (ev->m != p->m) return fail(TPA_ERR_ARG, "the event set belongs to another model");
  const int64_t n = ev->n;
  std::lock_guard<std::mutex> add_lk(p->add_mu);
  int64_t off = 0;
  {
    std::lock_guard<std::mutex> lk(p->mu);
    off = pool_take_range(p, n);
    if (off < 0) return fail(TPA_ERR_ARG, "pool full");
  }
  auto undo = [&](hipError_t e, const char* what) {
    std::lock_guard<std::mutex> lk(p->mu);
    release_range(p, off);
    return fail(TPA_ERR_HIP, std::string("pool add: ") + what + ": " + hipGetErrorString(e));
  };
  tpa_model_s* m = p->m;
  hipStream_t s = p->add_stream;
  hipError_t e = hipStreamSynchronize(m->stream);  // the set's angles are complete
  if (e != hipSuccess) return undo(e, "event set");
  // angles [field][n] of the set into rows [field][cap] of the pool at off
  if ((e = hipMemcpy2DAsync(p->d_ang + off, sizeof(double) * p->cap, ev->d_ang,
                            sizeof(double) * n, sizeof(double) * n, ev->nf,
                            hipMemcpyDeviceToDevice, s)) != hipSuccess)
    return undo(e, "angles");
  if ((e = hipMemcpyAsync(p->d_w + off, weight, sizeof(double) * n, hipMemcpyHostToDevice,
                          s)) != hipSuccess)
    return undo(e, "weights");
  launch_build_pw(m, p->d_ang, n, off, p->cap, 1, p->d_h, s);
  if (p->d_h32) launch_rows_to_f32(p, off, n, s);
  if ((e = hipGetLastError()) != hipSuccess) return undo(e, "launch");
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return undo(e, "build");
  *offset = off;
  return TPA_OK;
}

int tpa_pool_keep_f32(tpa_pool_t p, int on) {
  if (!p) return fail(TPA_ERR_ARG, "null pointer");
  std::lock_guard<std::mutex> add_lk(p->add_mu);
  std::lock_guard<std::mutex> lk(p->mu);
  for (auto& q : p->slot)
    if (q.busy) return fail(TPA_ERR_ARG, "a batch is in flight");
  const TpaModel& M = p->m->h;
  if (on && M.lam_pad < 2) return fail(TPA_ERR_UNSUPPORTED, "FP32 pool caches need >= 2 helicity slots");
  if ((p->d_h32 != nullptr) == (on != 0)) return TPA_OK;
  HIPCHK(hipStreamSynchronize(p->stream));
  // the batched graphs captured the cache pointer: capture again on the next batch
  for (auto& q : p->slot) {
    for (auto& g : q.graphs) (void)hipGraphExecDestroy(g.second);
    q.graphs.clear();
  }
  if (!on) {
    HIPCHK(hipFree(p->d_h32));
    p->d_h32 = nullptr;
    return TPA_OK;
  }
  HIPCHK(hipMalloc(&p->d_h32, sizeof(float2) * (size_t)M.P * p->cap * M.lam_pad));
  for (auto& r : p->used_ranges) launch_rows_to_f32(p, r.first, r.second, p->add_stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(p->add_stream));
  return TPA_OK;
}

int tpa_pool_release(tpa_pool_t p, int64_t offset) {
  if (!p) return fail(TPA_ERR_ARG, "null pointer");
  std::lock_guard<std::mutex> lk(p->mu);
  if (!p->used_ranges.count(offset)) return fail(TPA_ERR_ARG, "no allocation at this offset");
  release_range(p, offset);
  return TPA_OK;
}

static void release_range(tpa_pool_s* p, int64_t offset) {
  auto it = p->used_ranges.find(offset);
  int64_t off = offset, len = it->second;
  p->used_ranges.erase(it);
  auto nx = p->free_ranges.lower_bound(off);
  if (nx != p->free_ranges.end() && nx->first == off + len) {
    len += nx->second;
    p->free_ranges.erase(nx);
  }
  auto pv = p->free_ranges.lower_bound(off);
  if (pv != p->free_ranges.begin()) {
    --pv;
    if (pv->first + pv->second == off) {
      off = pv->first;
      len += pv->second;
      p->free_ranges.erase(pv);
    }
  }
  p->free_ranges[off] = len;
}

}  // extern "C"

// Queue one batched evaluation on a batch slot of the pool's stream and return without
// waiting. *slot_io: the slot to use, or -1 for any free one (the chosen slot is returned).
int tpa_pool_eval_submit(tpa_pool_t p, int* slot_io, int n_items, const int64_t* ranges,
                         const double* coefs, int mode) {
  if (!p || !slot_io || !ranges || !coefs) return fail(TPA_ERR_ARG, "null pointer");
  if (*slot_io >= TPA_POOL_SLOTS) return fail(TPA_ERR_ARG, "bad batch slot");
  if (n_items < 1 || n_items > p->max_items) return fail(TPA_ERR_ARG, "bad request count");
  if (mode != TPA_MODE_DATA) return fail(TPA_ERR_UNSUPPORTED, "batched evaluation serves data sums");
  std::lock_guard<std::mutex> lk(p->mu);
  int si = *slot_io;
  if (si < 0)
    for (int i = 0; i < TPA_POOL_SLOTS && si < 0; ++i)
      if (!p->slot[i].busy) si = i;
  if (si < 0) return fail(TPA_ERR_ARG, "every batch slot is in flight");
  auto& q = p->slot[si];
  if (q.busy) return fail(TPA_ERR_ARG, "batch slot still in flight");
  const TpaModel& M = p->m->h;
  const int nacc = 2 + 2 * M.P;
  int64_t nch = 0;
  for (int i = 0; i < n_items; ++i) {
    const int64_t off = ranges[2 * i], n = ranges[2 * i + 1];
    auto it = p->used_ranges.upper_bound(off);
    if (n < 1 || it == p->used_ranges.begin()) return fail(TPA_ERR_ARG, "range outside the pool");
    --it;
    if (off < it->first || off + n > it->first + it->second)
      return fail(TPA_ERR_ARG, "range outside an allocation");
    nch += (n + TPA_CHUNK - 1) / TPA_CHUNK;
  }
  if (nch > p->max_chunks) return fail(TPA_ERR_ARG, "too many chunks in one batch");
  // the blob (BatchView's layout): header, chunk starts, ends and requests, each request's
  // first chunk, then the coefficients at an even word
  int64_t* hb = q.h_blob;
  int64_t* cs = hb + 2;
  int64_t* ce = cs + nch;
  int64_t* ci = ce + nch;
  int64_t* ic = ci + nch;
  hb[0] = nch;
  hb[1] = n_items;
  int64_t c = 0;
  for (int i = 0; i < n_items; ++i) {
    const int64_t off = ranges[2 * i], n = ranges[2 * i + 1];
    ic[i] = c;
    for (int64_t st = off; st < off + n; st += TPA_CHUNK, ++c) {
      cs[c] = st;
      ce[c] = std::min<int64_t>(st + TPA_CHUNK, off + n);
      ci[c] = i;
    }
  }
  ic[n_items] = c;
  const int64_t coef_word = (2 + 3 * nch + n_items + 1 + 1) & ~int64_t(1);
  std::memcpy(hb + coef_word, coefs, sizeof(double2) * (size_t)n_items * M.P);
  const size_t blob_bytes = sizeof(int64_t) * (size_t)(coef_word + 2 * (int64_t)n_items * M.P);
  hipStream_t s = p->stream;
  HIPCHK(hipMemcpyAsync(q.d_blob, q.h_blob, blob_bytes, hipMemcpyHostToDevice, s));
  int bucket = 64;
  while (bucket < nch) bucket <<= 1;
  auto g = q.graphs.find(bucket);
  if (g == q.graphs.end()) {
    MultiArgs a;
    a.h = p->d_h;
    a.h32 = p->d_h32;
    a.w = p->d_w;
    a.blob = q.d_blob;
    a.partials = q.d_part;
    a.S = p->cap;
    a.P = M.P;
    a.mode = mode;
    a.nacc = nacc;
    const int kg = eval_kgroups(M.lam_pad, M.P);
    if (kg < 0) return fail(TPA_ERR_UNSUPPORTED, "too many partial waves for this helicity count");
    hipGraph_t graph;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = dispatch_multi(M.lam_pad, kg, a, bucket, s);
    hipLaunchKernelGGL(k_reduce_items, dim3(p->max_items), dim3(256), 0, s, q.d_part, q.d_blob,
                       nacc, q.d_out);
    hipError_t ce = hipStreamEndCapture(s, &graph);
    if (rc != TPA_OK) return fail(rc, "no batched kernel for this helicity count");
    if (ce != hipSuccess) return fail(TPA_ERR_HIP, std::string("capture: ") + hipGetErrorString(ce));
    hipGraphExec_t exec;
    hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) return fail(TPA_ERR_HIP, std::string("instantiate: ") + hipGetErrorString(ie));
    g = q.graphs.emplace(bucket, exec).first;
  }
  HIPCHK(hipEventRecord(q.t0, s));
  HIPCHK(hipGraphLaunch(g->second, s));
  HIPCHK(hipEventRecord(q.t1, s));
  HIPCHK(hipMemcpyAsync(q.h_out, q.d_out, sizeof(double) * (size_t)n_items * nacc, hipMemcpyDeviceToHost, s));
  HIPCHK(hipEventRecord(q.done, s));
  q.n_items = n_items;
  q.busy = true;
  *slot_io = si;
  return TPA_OK;
}

// 1 when slot `si`'s batch has completed (its results are on the host), 0 while it runs.
int tpa_pool_eval_ready(tpa_pool_t p, int si) {
  if (!p || si < 0 || si >= TPA_POOL_SLOTS) return fail(TPA_ERR_ARG, "bad batch slot");
  auto& q = p->slot[si];
  if (!q.busy) return fail(TPA_ERR_ARG, "no batch in this slot");
  const hipError_t e = hipEventQuery(q.done);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return fail(TPA_ERR_HIP, std::string("batch: ") + hipGetErrorString(e));
}

// Wait for slot `si`'s batch; its results (n_items x (2 + 2P)) go to out, the graph's GPU time
// to ms. Frees the slot.
int tpa_pool_eval_finish(tpa_pool_t p, int si, double* out, double* ms) {
  if (!p || !out || si < 0 || si >= TPA_POOL_SLOTS) return fail(TPA_ERR_ARG, "bad batch slot");
  auto& q = p->slot[si];
  if (!q.busy) return fail(TPA_ERR_ARG, "no batch in this slot");
  const hipError_t e = hipEventSynchronize(q.done);
  float f = 0.0f;
  const hipError_t et = e == hipSuccess ? hipEventElapsedTime(&f, q.t0, q.t1) : e;
  const int nacc = 2 + 2 * p->m->h.P;
  if (e == hipSuccess) std::memcpy(out, q.h_out, sizeof(double) * (size_t)q.n_items * nacc);
  {
    std::lock_guard<std::mutex> lk(p->mu);
    q.busy = false;
  }
  if (e != hipSuccess) return fail(TPA_ERR_HIP, std::string("batch: ") + hipGetErrorString(e));
  if (ms) *ms = et == hipSuccess ? f : 0.0;
  return TPA_OK;
}

extern "C" {

int tpa_pool_eval(tpa_pool_t p, int n_items, const int64_t* ranges, const double* coefs,
                  int mode, double* out, double* ms) {
  if (!p || !ranges || !coefs || !out) return fail(TPA_ERR_ARG, "null pointer");
  int si = -1;  // any free slot (a serving batcher may hold the others)
  const int rc = tpa_pool_eval_submit(p, &si, n_items, ranges, coefs, mode);
  if (rc != TPA_OK) return rc;
  return tpa_pool_eval_finish(p, si, out, ms);
}

int tpa_hess(tpa_events_t ev, const double* coef, int mode, double* out) {
  if (!ev || !coef || !out) return fail(TPA_ERR_ARG, "null pointer");
  if (mode != TPA_MODE_DATA && mode != TPA_MODE_MC) return fail(TPA_ERR_ARG, "bad mode");
  if (ev->built_fold != 1) return fail(TPA_ERR_ARG, "second moments need a fold = 1 cache");
  return run_hess(ev->m->h, ev->d_h, ev->d_w, 0, ev->n, ev->n, coef, mode, ev->m->stream,
                  &ev->d_hpart, &ev->d_hout, &ev->d_hcoef, out);
}

int tpa_hess_shape(tpa_events_t ev, const double* coef, int mode, double* out) {
  if (!ev || !coef || !out) return fail(TPA_ERR_ARG, "null pointer");
  if (mode != TPA_MODE_DATA && mode != TPA_MODE_MC) return fail(TPA_ERR_ARG, "bad mode");
  if (ev->m->kind != 3)
    return fail(TPA_ERR_UNSUPPORTED, "floating lineshapes need a three-body model");
  if (ev->built_fold != 0)
    return fail(TPA_ERR_ARG, "lineshape second moments need a cache built with fold = 0");
  const TpaModel& M = ev->m->h;
  const int P = M.P, Q = 6 * P;
  if (P > TPA_HESS_SHAPE_PMAX)
    return fail(TPA_ERR_UNSUPPORTED, "lineshape second moments need P <= " +
                                         std::to_string(TPA_HESS_SHAPE_PMAX));
  const int QT = Q <= 48 ? 3 : Q <= 96 ? 6 : 9;
  const int nacc = 2 + 12 * P + 2 * (QT * (QT + 1) / 2) * 256;
  const int LG = M.lam_pad;
  if (LG != 1 && LG != 2 && LG != 4)
    return fail(TPA_ERR_UNSUPPORTED, "lineshape second moments need <= 4 helicity slots");
  const int E = 32 / LG;
  const int64_t tiles = (ev->n + E - 1) / E;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, kHessGridMax));
  hipStream_t s = ev->m->stream;
  if (!ev->d_bw2)
    HIPCHK(hipMalloc(&ev->d_bw2, sizeof(double2) * 6 * (size_t)M.n_chain * ev->n));
  if (!ev->d_hspart) {
    const int cap = 2 + 12 * TPA_HESS_SHAPE_PMAX + 2 * 45 * 256;
    HIPCHK(hipMalloc(&ev->d_hspart, sizeof(double) * (size_t)kHessGridMax * cap));
    HIPCHK(hipMalloc(&ev->d_hsout, sizeof(double) * cap));
  }
  if (!ev->d_hcoef) HIPCHK(hipMalloc(&ev->d_hcoef, sizeof(double2) * TPA_HESS_PMAX));
  HIPCHK(hipMemcpyAsync(ev->d_hcoef, coef, sizeof(double2) * P, hipMemcpyHostToDevice, s));
  const int64_t lgrid = (ev->n + 255) / 256;
  hipLaunchKernelGGL(k_lineshape_hess, dim3((unsigned)lgrid), dim3(256), 0, s, ev->m->d,
                     ev->d_ang, ev->n, ev->d_bw2);
  HIPCHK(hipGetLastError());
  HessShapeArgs a;
  a.h = ev->d_h;
  a.w = ev->d_w;
  a.coef = ev->d_hcoef;
  a.bw = ev->d_bw2;
  a.kchain = ev->m->d_kchain;
  a.partials = ev->d_hspart;
  a.N = ev->n;
  a.P = P;
  a.n_chain = M.n_chain;
  a.mode = mode;
  a.nacc = nacc;
#define TPA_HS_LAUNCH(lg, qt) \
  hipLaunchKernelGGL((k_hess_shape<lg, qt>), dim3(grid), dim3(256), 0, s, a)
  if (LG == 1) {
    if (QT == 3) TPA_HS_LAUNCH(1, 3); else if (QT == 6) TPA_HS_LAUNCH(1, 6); else TPA_HS_LAUNCH(1, 9);
  } else if (LG == 2) {
    if (QT == 3) TPA_HS_LAUNCH(2, 3); else if (QT == 6) TPA_HS_LAUNCH(2, 6); else TPA_HS_LAUNCH(2, 9);
  } else {
    if (QT == 3) TPA_HS_LAUNCH(4, 3); else if (QT == 6) TPA_HS_LAUNCH(4, 6); else TPA_HS_LAUNCH(4, 9);
  }
#undef TPA_HS_LAUNCH
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_reduce, dim3(nacc), dim3(256), 0, s, ev->d_hspart, grid, nacc, 0,
                     ev->d_hsout, nullptr, nullptr, nullptr, (uint64_t)0);
  HIPCHK(hipGetLastError());
  std::vector<double> raw(nacc);
  HIPCHK(hipMemcpyAsync(raw.data(), ev->d_hsout, sizeof(double) * nacc, hipMemcpyDeviceToHost,
                        s));
  HIPCHK(hipStreamSynchronize(s));
  // value, F, F2, then each matrix's upper triangle (row-major, i <= j) from the 16x16 FP64
  // MFMA accumulator layout: element (i, j) of tile q sits in lane j + 16 (i % 4), register i / 4
  const int head = 2 + 12 * P;
  for (int j = 0; j < head; ++j) out[j] = raw[j];
  std::vector<int> tile(QT * QT, -1);
  for (int ti = 0, q = 0; ti < QT; ++ti)
    for (int tj = ti; tj < QT; ++tj, ++q) tile[ti * QT + tj] = q;
  const int NT = QT * (QT + 1) / 2;
  const int T = Q * (Q + 1) / 2;
  for (int mtx = 0; mtx < 2; ++mtx) {
    const double* src = raw.data() + head + (size_t)mtx * NT * 256;
    double* dst = out + head + (size_t)mtx * T;
    int idx = 0;
    for (int r = 0; r < Q; ++r)
      for (int c = r; c < Q; ++c, ++idx) {
        const int q = tile[(r / 16) * QT + c / 16];
        const int i = r % 16, j = c % 16;
        dst[idx] = src[(q * 4 + i / 4) * 64 + j + 16 * (i % 4)];
      }
  }
  return TPA_OK;
}

int tpa_pool_hess(tpa_pool_t p, int64_t offset, int64_t n, const double* coef, double* out) {
  if (!p || !coef || !out) return fail(TPA_ERR_ARG, "null pointer");
  std::lock_guard<std::mutex> lk(p->mu);
  auto it = p->used_ranges.upper_bound(offset);
  if (n < 1 || it == p->used_ranges.begin()) return fail(TPA_ERR_ARG, "range outside the pool");
  --it;
  if (offset < it->first || offset + n > it->first + it->second)
    return fail(TPA_ERR_ARG, "range outside an allocation");
  return run_hess(p->m->h, p->d_h, p->d_w, offset, n, p->cap, coef, TPA_MODE_DATA, p->stream,
                  &p->d_hpart, &p->d_hout, &p->d_hcoef, out);
}

int tpa_device_synchronize(void) {
  HIPCHK(hipDeviceSynchronize());
  return TPA_OK;
}

int tpa_calib_copy(int64_t bytes, int reps, double* ms) {
  if (bytes < 16 || reps < 1) return fail(TPA_ERR_ARG, "bad copy size");
  const int64_t n = bytes / 16;
  double2 *a = nullptr, *b = nullptr;
  hipEvent_t t0, t1;
  HIPCHK(hipMalloc(&a, n * 16));
  if (hipMalloc(&b, n * 16) != hipSuccess) {
    (void)hipFree(a);
    return fail(TPA_ERR_HIP, "copy buffer allocation failed");
  }
  HIPCHK(hipMemset(a, 0, n * 16));
  HIPCHK(hipEventCreate(&t0));
  HIPCHK(hipEventCreate(&t1));
  HIPCHK(hipEventRecord(t0, 0));
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, a, b, n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(t1, 0));
  HIPCHK(hipEventSynchronize(t1));
  float f = 0.0f;
  HIPCHK(hipEventElapsedTime(&f, t0, t1));
  if (ms) *ms = f / reps;
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  (void)hipFree(a);
  (void)hipFree(b);
  return TPA_OK;
}

int tpa_timing_enable(int on) {
  std::lock_guard<std::mutex> lk(g_timing_mu);
  g_timing = on > 0;
  g_timing_every = on > 0 ? on : 1;
  g_timing_tick = 0;
  return TPA_OK;
}

int tpa_timing_read(double* ms, int64_t* launches) {
  std::lock_guard<std::mutex> lk(g_timing_mu);
  if (ms) *ms = g_timing_ms;
  if (launches) *launches = g_timing_n;
  return TPA_OK;
}

int tpa_timing_reset(void) {
  std::lock_guard<std::mutex> lk(g_timing_mu);
  g_timing_ms = 0.0;
  g_timing_n = 0;
  return TPA_OK;
}

}  // extern "C"
