/* C ABI of the MI355X unbinned-likelihood engine (libtfpwa_amd.so).
 *
 * The engine replaces the data-parallel part of tf-pwa's likelihood path: the per-event
 * amplitude |A|^2, the weighted log-sum over data, the weighted MC normalisation sum and
 * their reverse-mode derivatives. Everything crosses this boundary as plain pointers and
 * sizes; no exceptions or C++ types cross it. Every function returns 0 on success and a
 * negative status on failure; tpa_last_error() then describes the failure (per thread).
 *
 * Ownership: a model handle owns the device copy of the model description; an event-set
 * handle owns the device buffers of one data set (angles, partial-wave cache, weights).
 * Results are written into caller-provided host buffers. Calls on one event set are
 * synchronous and must not overlap; different event sets may be used from different
 * threads.
 */
#ifndef TFPWA_AMD_H
#define TFPWA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tpa_model_s* tpa_model_t;
typedef struct tpa_events_s* tpa_events_t;

#define TPA_OK 0
#define TPA_ERR_ARG -1
#define TPA_ERR_HIP -2
#define TPA_ERR_UNSUPPORTED -3

/* Evaluation modes (tpa_eval). DATA: sum_e w_e * clip_log(f_e) (tf_pwa/model/model.py:41-54
 * with trans=clip_log, 283-292). MC: sum_e w_e * f_e (the same function with identity). */
#define TPA_MODE_DATA 0
#define TPA_MODE_MC 1
/* Flags (tpa_eval). */
#define TPA_FLAG_GRAD 1 /* also return the adjoint sums G_k */
#define TPA_FLAG_FOUT 2 /* also return the per-event f_e */
/* Lineshapes evaluated per event at the model's current (m0, g0) instead of being folded
 * into the cache (cache built with fold = 0); with TPA_FLAG_GRAD the output then also holds
 * Dm_k = sum_e u_e dB/dm0 H_k at out[2+2P+2k..] and Dg_k (dB/dg0) at out[2+4P+2k..],
 * H_k = sum_lambda conj(A_lambda) h_k,lambda, so out needs 2+6P doubles. */
#define TPA_FLAG_SHAPE 4

/* Version of this ABI. */
int tpa_abi_version(void);

/* Text of the last error on the calling thread ("" if none). */
const char* tpa_last_error(void);

/* Number of visible GPUs. */
int tpa_device_count(int* n);

/* Lower a three-body decay group into a device-resident model.
 * desc: a TpaModel (tf-pwa_amd/csrc/tpa_model.h) built by the host from config.yml;
 * cg: the CG x sqrt((2l+1)/(2J+1)) tables referenced by the chains (n_cg doubles).
 * Replaces the amplitude construction of tf_pwa/config_loader/config_loader.py:210-237
 * (get_amplitude -> DecayGroup.init_params, tf_pwa/amp/core.py:1650-1671) and the constant
 * tables of HelicityDecay._get_cg_matrix (tf_pwa/amp/core.py:851-900). */
int tpa_model_create(const void* desc, int64_t desc_bytes, const double* cg, int n_cg,
                     tpa_model_t* out);
int tpa_model_destroy(tpa_model_t model);

/* Upload n events (p4: host array [n_final*4][n], finals in sorted-name order, components
 * E, px, py, pz) and compute their helicity and aligned angles on the device.
 * Replaces cal_angle_from_momentum (tf_pwa/cal_angle.py:762-865) for the data and MC sets
 * of a fit (FCN.__init__, tf_pwa/model/model.py:1126-1164). weight: host [n] or NULL (1). */
int tpa_events_create(tpa_model_t model, const double* p4, int64_t n, const double* weight,
                      tpa_events_t* out);
int tpa_events_destroy(tpa_events_t ev);
int64_t tpa_events_size(tpa_events_t ev);

/* Replace the per-event weights (host [n]). The host applies tf-pwa's weight conventions
 * (data: alpha = sum w / sum w^2 rescale, get_weight_data tf_pwa/model/model.py:603-641;
 * MC: normalised to unit sum, FCN.__init__ 1155-1162). */
int tpa_events_set_weight(tpa_events_t ev, const double* weight);

/* Copy the angle cache to the host: out [n_fields][n] with n_fields = 18*n_topo + 4. */
int tpa_events_get_angles(tpa_events_t ev, double* out, int64_t out_len);

/* Build the partial-wave cache h[k][e][lambda] from the angle cache. fold != 0 multiplies
 * each chain's lineshape (with the m0/g0 stored in the model description) into the cache.
 * Replaces the D-matrix cache of get_D_matrix_for_angle (tf_pwa/dfun.py:221-238) and the
 * parameter-free factors of HelicityDecay.get_amp / DecayChain.get_amp
 * (tf_pwa/amp/core.py:931-1164, 1392-1445). */
int tpa_events_build(tpa_events_t ev, int fold);

/* Replace the model description (same chains and CG table; new lineshape parameters or
 * barrier clamp flags). Event sets of the model must be rebuilt (tpa_events_build) before
 * their next evaluation. */
int tpa_model_update(tpa_model_t model, const void* desc, int64_t desc_bytes);

/* Copy the partial-wave cache to the host: out [P][n][lam_pad] complex (interleaved). */
int tpa_events_get_pw(tpa_events_t ev, double* out, int64_t out_len);

/* One likelihood pass over an event set with partial-wave coefficients c_k
 * (coef: host [2P] interleaved re/im): out[0] = the weighted sum of the mode; with
 * TPA_FLAG_GRAD out[2+2k], out[3+2k] = Re/Im of G_k = sum_e u_e sum_lambda conj(A) h_k
 * where u_e is the derivative weight of the mode, so that d(sum)/d theta =
 * 2 Re sum_k (d c_k / d theta) G_k. out needs 2+2P doubles (2+6P with TPA_FLAG_SHAPE).
 * The cache must be built with fold = 1 (fold = 0 for TPA_FLAG_SHAPE). With TPA_FLAG_FOUT, f_out
 * (host [n]) receives f_e = sum_lambda |A_lambda|^2.
 * Replaces sum_gradient / _batch_sum (tf_pwa/model/model.py:41-94) as called from
 * BaseModel.nll_grad_batch (407-449) and BaseModel.nll (322-339). */
int tpa_eval(tpa_events_t ev, const double* coef, int mode, int flags, double* out,
             double* f_out);

/* Kernel timing with HIP events on the engine's stream: when enabled, every likelihood
 * kernel launch is bracketed by events; read returns the summed milliseconds and the
 * number of launches since the last reset. */
int tpa_timing_enable(int on);
int tpa_timing_read(double* ms, int64_t* launches);
int tpa_timing_reset(void);

/* Partial-wave Gram matrix of an event set on FP32 MFMA with FP64 block refolds:
 * I_kl = sum_e w_e sum_lambda conj(h_k) h_l (out: host [P][P] complex, interleaved). With it
 * the weighted MC sum at any coefficients is c^dagger I c, and its adjoint G = conj(I c).
 * Replaces build_int_matrix of tf-pwa's cached-integral model
 * (tf_pwa/experimental/opt_int.py:30-52, tf_pwa/model/opt_int.py:128-296). P <= 32.
 * w_mean: the mean event weight (the kernel works with w / w_mean in FP32).
 * ms (optional): the MFMA kernel's duration. */
int tpa_gram(tpa_events_t ev, double w_mean, double* out, double* ms);

/* Wait for all device work of the process (the timing barrier of benchmarks). */
int tpa_device_synchronize(void);

/* Measurement helper: copy `bytes` between two fresh device buffers `reps` times with a
 * plain 16-byte streaming kernel (k_copy) and return the average milliseconds per copy.
 * It is the calibration kernel for the FETCH_SIZE / WRITE_SIZE counters. */
int tpa_calib_copy(int64_t bytes, int reps, double* ms);

#ifdef __cplusplus
}
#endif
#endif
