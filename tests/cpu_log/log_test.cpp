// Host build of the evaluation passes' ln / reciprocal (tpa_log.h) for the CPU tests: the
// device's v_rcp_f64 approximation is stood in for by the exact reciprocal perturbed by a
// relative error of up to 2^-22.
#define TPA_LOG_FN inline
#include <stdint.h>

#include <cstring>

#include "tpa_log.h"

namespace {
// 1/d with its low mantissa bits replaced by a pattern: a relative error of up to ~2^-22
double rough_recip(double d) {
  double r = 1.0 / d;
  uint64_t b;
  std::memcpy(&b, &r, 8);
  b = (b & ~((uint64_t(1) << 30) - 1)) | (uint64_t(0x2AAAAAAA) & ((uint64_t(1) << 30) - 1));
  std::memcpy(&r, &b, 8);
  return r;
}
}  // namespace

extern "C" void log_pos_batch(const double* x, long n, double* out_log, double* out_rcp) {
  for (long i = 0; i < n; ++i) {
    out_log[i] = tpa_log_pos(x[i], rough_recip);
    out_rcp[i] = tpa_recip_pos(x[i], rough_recip);
  }
}
