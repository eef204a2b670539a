/* CPU restatement (C, OpenMP, FP64) of one likelihood pass over a partial-wave cache: the
 * CPU baseline of the evaluation benchmark and a checker of the GPU pass. Test and measurement
 * infrastructure only; the engine never calls it.
 *
 * For events e with weights w_e and partial waves h_k,l(e) (k < P, helicity slot l < LG):
 *   A_l = sum_k c_k h_k,l,   f = sum_l |A_l|^2,
 *   mode 0 (data): value = sum_e w clip_log(f), u = w clip_log'(f)
 *                  (clip_log: tf_pwa/model/model.py:283-292, log above 1e-6, its second-order
 *                  Taylor expansion at 1e-6 below),
 *   mode 1 (MC):   value = sum_e w f,           u = w,
 *   G_k = sum_e u sum_l conj(A_l) h_k,l  (the adjoint: dvalue/dRe c_k = 2 Re G_k,
 *                                         dvalue/dIm c_k = -2 Im G_k),
 * the sums of tf_pwa/model/model.py:41-94 (sum_gradient) with the amplitude in its
 * partial-wave form. Events are processed in blocks of CPU_BLOCK so that a block's waves stay
 * in cache for the adjoint pass; threads own contiguous ranges of blocks and their partial
 * sums are combined in thread order (deterministic for a given thread count). */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define CPU_BLOCK 256
#define CPU_MAX_P 128
#define CPU_MAX_LG 16

int cpu_eval_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* h: [P][N][LG] complex as (re, im) pairs; coef: [P] complex; out: 2 + 2P doubles. */
int cpu_eval(const double* h, const double* w, int64_t N, int P, int LG, const double* coef,
             int mode, double* out) {
  if (P < 1 || P > CPU_MAX_P || LG < 1 || LG > CPU_MAX_LG || N < 1) return -1;
  const int nthr = cpu_eval_threads();
  double* part = (double*)calloc((size_t)nthr * (2 + 2 * P), sizeof(double));
  if (!part) return -2;
  const int64_t nblk = (N + CPU_BLOCK - 1) / CPU_BLOCK;
#pragma omp parallel num_threads(nthr)
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
    const int nt = omp_get_num_threads();
#else
    const int tid = 0, nt = 1;
#endif
    double* acc = part + (size_t)tid * (2 + 2 * P);
    double vs = 0.0, vc = 0.0;  /* compensated value */
    double ar[CPU_BLOCK * CPU_MAX_LG], ai[CPU_BLOCK * CPU_MAX_LG], u[CPU_BLOCK];
    const int64_t b0 = nblk * tid / nt, b1 = nblk * (tid + 1) / nt;
    for (int64_t b = b0; b < b1; ++b) {
      const int64_t e0 = b * CPU_BLOCK;
      const int ne = (int)((e0 + CPU_BLOCK <= N) ? CPU_BLOCK : N - e0);
      const int nr = ne * LG;
      memset(ar, 0, sizeof(double) * nr);
      memset(ai, 0, sizeof(double) * nr);
      for (int k = 0; k < P; ++k) {
        const double cr = coef[2 * k], ci = coef[2 * k + 1];
        const double* hk = h + 2 * ((size_t)k * N * LG + (size_t)e0 * LG);
        for (int r = 0; r < nr; ++r) {
          ar[r] += cr * hk[2 * r] - ci * hk[2 * r + 1];
          ai[r] += cr * hk[2 * r + 1] + ci * hk[2 * r];
        }
      }
      for (int i = 0; i < ne; ++i) {
        double f = 0.0;
        for (int l = 0; l < LG; ++l) f += ar[i * LG + l] * ar[i * LG + l] + ai[i * LG + l] * ai[i * LG + l];
        const double we = w[e0 + i];
        double y;
        if (mode == 0) {
          const double eps = 1.0e-6;
          double cl, dcl;
          if (f > eps) {
            cl = log(f);
            dcl = 1.0 / f;
          } else {
            const double t = (f - eps) / eps;
            cl = (log(eps) + t) - (t * t) / 2.0;
            dcl = 1.0 / eps - t / eps;
          }
          y = we * cl;
          u[i] = we * dcl;
        } else {
          y = we * f;
          u[i] = we;
        }
        /* Neumaier summation */
        const double t = vs + y;
        if (fabs(vs) >= fabs(y))
          vc += (vs - t) + y;
        else
          vc += (y - t) + vs;
        vs = t;
      }
      for (int k = 0; k < P; ++k) {
        const double* hk = h + 2 * ((size_t)k * N * LG + (size_t)e0 * LG);
        double gr = 0.0, gi = 0.0;
        for (int i = 0; i < ne; ++i) {
          double qr = 0.0, qi = 0.0;
          for (int l = 0; l < LG; ++l) {
            const int r = i * LG + l;
            const double hr = hk[2 * r], hi = hk[2 * r + 1];
            qr += ar[r] * hr + ai[r] * hi; /* conj(A) h */
            qi += ar[r] * hi - ai[r] * hr;
          }
          gr += u[i] * qr;
          gi += u[i] * qi;
        }
        acc[2 + 2 * k] += gr;
        acc[3 + 2 * k] += gi;
      }
    }
    acc[0] = vs;
    acc[1] = vc;
  }
  memset(out, 0, sizeof(double) * (2 + 2 * P));
  double vs = 0.0;
  for (int t = 0; t < nthr; ++t) {
    const double* acc = part + (size_t)t * (2 + 2 * P);
    vs += acc[0] + acc[1];
    for (int j = 2; j < 2 + 2 * P; ++j) out[j] += acc[j];
  }
  out[0] = vs;
  free(part);
  return 0;
}
