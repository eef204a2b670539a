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
typedef struct tpa_pool_s* tpa_pool_t;
typedef struct tpa_server_s* tpa_server_t;
typedef struct tpa_fcn_s* tpa_fcn_t;

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
/* Read the FP32 copy of a folded cache (tpa_events_keep_f32) instead of the FP64 cache: half
 * the bytes per event; the values are widened to FP64 on load and every operation after it is
 * FP64. Not with TPA_FLAG_SHAPE. */
#define TPA_FLAG_F32 8
/* Read the helicity-halved FP32 copy (tpa_events_keep_half) instead of the FP64 cache: the
 * slots of one helicity of each mirror pair in FP32 plus one phase per chain and event, from
 * which the mirrored slots follow (h_k,-o = eta_k s_o rho conj(h_k,o)); every operation after
 * the load is FP64. At most 32 partial waves; not with TPA_FLAG_SHAPE or TPA_FLAG_F32. */
#define TPA_FLAG_F32H 16
/* The same from the helicity-halved FP64 copy (tpa_events_keep_half with bit 1): the FP64
 * cache's values to rounding, 808 instead of 1416 B per event for the six-resonance model. */
#define TPA_FLAG_F64H 32
/* Largest model (partial waves) tpa_eval_fused accepts. */
#define TPA_FUSED_MAX_P 8

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
/* Flat phase-space events of m_top -> masses[0..n_final) (the model's finals, in the engine's
 * final order) generated on the device: GENBOD with a per-event Philox4x32 counter
 * (seed, first + event index), so samples are reproducible and can be produced in pieces.
 * The four-momenta stay on the device (copied to p4_out, host [n_final*4][n], when not NULL);
 * the set's angles are computed from them. weight: host [n] or NULL (1).
 * Replaces PhaseSpaceGenerator.generate (tf_pwa/phasespace.py:93-202) for MC samples. */
int tpa_events_generate(tpa_model_t model, double m_top, const double* masses, int64_t n,
                        uint64_t seed, uint64_t first, const double* weight, double* p4_out,
                        tpa_events_t* out);

/* A new event set made of events idx[0..n) of src (their angles; the partial-wave cache is
 * built on demand), e.g. the accepted events of an accept-reject pass over a generated sample
 * (tf_pwa/applications.py:430-453). weight: host [n] or NULL (1). */
int tpa_events_select(tpa_events_t src, const int64_t* idx, int64_t n, const double* weight,
                      tpa_events_t* out);

/* A new event set made of all events of srcs[0..n_src) in order (their angles and weights;
 * the partial-wave cache is built on demand), so that one pass covers them. Every source must
 * belong to the same model. Replaces data_merge (tf_pwa/data.py:527-540) of event dicts. */
int tpa_events_concat(const tpa_events_t* srcs, int n_src, tpa_events_t* out);

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

/* on != 0: keep an FP32 copy (round to nearest) of the folded partial-wave cache beside the
 * FP64 one, made now if the cache is built and refreshed by every later folded build, for
 * tpa_eval with TPA_FLAG_F32; on = 0 frees it. No reference counterpart: the storage option
 * of the data pass of the mixed-precision likelihood (the reference evaluates
 * sum_gradient, tf_pwa/model/model.py:41-94, in FP64 from the angle cache). */
int tpa_events_keep_f32(tpa_events_t ev, int on);

/* kinds bit 0: keep the helicity-halved FP32 copy of the folded cache for tpa_eval with
 * TPA_FLAG_F32H; bit 1: the FP64 copy for TPA_FLAG_F64H; each made now if the cache is built
 * and refreshed by every later folded build; a copy whose bit is clear is freed. Flipping every
 * helicity of a partial wave conjugates its angular part, so for each mirror pair of outer
 * helicity slots (o, -o) only o is stored, with the phase B/conj(B) of each chain's lineshape
 * product per event: 408 instead of 712 B per event (FP32), 808 instead of 1416 (FP64), for
 * the six-resonance model. Making a copy checks on every event that the stored half
 * reproduces the mirrored slots of the FP64 cache; a relative residual above 1e-9 (a model
 * whose helicity lists are not symmetric) refuses the copy with TPA_ERR_UNSUPPORTED.
 * No reference counterpart: a storage option of the data pass of the mixed-precision
 * likelihood (the reference evaluates sum_gradient, tf_pwa/model/model.py:41-94, in FP64 from
 * the angle cache; the helicity amplitudes follow tf_pwa/amp/core.py:931-1164). */
int tpa_events_keep_half(tpa_events_t ev, int kinds);

/* The largest relative residual the last helicity-halved copy of ev measured
 * (max |h_k,-o - eta_k s_o rho conj(h_k,o)| / max |h| per event, in FP64 before rounding);
 * -1 if no copy was made. No reference counterpart (a diagnostic of tpa_events_keep_half). */
int tpa_events_half_residual(tpa_events_t ev, double* out);

/* Copy the helicity-halved copy to the host: g [P][n][nh] and rho [n_chain][n] as interleaved
 * FP32 complex, and info: info[0] = nh, info[1+q] the stored slot q, info[5+q] its mirror,
 * info[9+q] its sign s_o (q < nh), info[13+k] = 1 where wave k is stored times i (phi_k = -i),
 * info[13+P+k] the chain of wave k (info_len >= 13 + 2P). So h_k,o = phi_k g_k,o and
 * h_k,-o = phi_k s_o rho_c conj(g_k,o). With info_len >= 46 + 2P also the lane plan of the
 * chain-grouped FP32 pass: info[13+2P] its waves per lane (0: no plan fits, the pass is
 * k_eval_half), info[14+2P+i] the wave of lane slot i (-1: pad). No reference counterpart (a
 * diagnostic). */
int tpa_events_get_half(tpa_events_t ev, float* g, int64_t g_len, float* rho, int64_t rho_len,
                        int* info, int info_len);

/* Replace the model description (same chains and CG table; new lineshape parameters or
 * barrier clamp flags). Event sets of the model must be rebuilt (tpa_events_build) before
 * their next evaluation. */
int tpa_model_update(tpa_model_t model, const void* desc, int64_t desc_bytes);

/* Four-body final states of spin-0 particles (desc: TpaModel4, tpa_model.h): trees of three
 * two-body vertices, quasi-two-body (A -> R1 R2) or sequential (A -> R1 f, R1 -> R2 f). The
 * handle works with every event-set, evaluation, Gram, second-moment and pool call; the
 * partial-wave cache always folds the lineshapes (masses and widths fixed). Replaces
 * DecayGroup/DecayChain.get_amp with cal_angle_from_momentum for such decay groups
 * (tf_pwa/amp/core.py:1392-1445, 1696-1728; tf_pwa/cal_angle.py:186-343, 762-865). */
int tpa_model4_create(const void* desc, int64_t desc_bytes, const double* cg, int n_cg,
                      tpa_model_t* out);
int tpa_model4_update(tpa_model_t model, const void* desc, int64_t desc_bytes);

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

/* tpa_eval's sums (modes DATA and MC; flags: TPA_FLAG_GRAD) with the partial waves rebuilt
 * from the angle cache on every call, `chunk` events at a time (chunk <= 0: 262144): each
 * chunk's waves are built into one chunk buffer (lineshapes folded at the model's current
 * parameters, as tpa_events_build(ev, 1)), streamed by the evaluation pass and reduced, and the
 * chunks' sums added in chunk order. Only the angle cache and the weights need to stay
 * resident (the event set's own partial-wave cache is neither built nor read), so samples too
 * large for a resident cache can be evaluated; each call pays the build. At most 32 partial
 * waves. Replaces sum_gradient (tf_pwa/model/model.py:41-94) over the reference's default
 * path, which recomputes the amplitude from the cached angles on every evaluation
 * (tf_pwa/amp/core.py:1392-1445, tf_pwa/dfun.py:221-238). */
int tpa_eval_recompute(tpa_events_t ev, const double* coef, int mode, int flags, int64_t chunk,
                       double* out);

/* tpa_eval without allocation, for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). No counterpart in the
 * reference. The pass of tpa_eval over the event set's resident cache (flags: TPA_FLAG_GRAD,
 * TPA_FLAG_F32; the cache built with fold = 1), with its per-block partial sums in a static
 * device array of the library (1024 blocks x 66 doubles, 528 KiB on each GPU the library runs
 * on, held until the process ends) instead of the event set's buffer. Writes tpa_eval's sums
 * (2 doubles, or 2 + 2P with TPA_FLAG_GRAD) to d_out, which must be device memory, on `stream`
 * (a hipStream_t; NULL: the model's stream), and returns once the work is queued: no
 * allocation, copy or synchronisation. Results equal tpa_eval's bit for bit. Rejects P > 32
 * and grids beyond the scratch (TFPWA_EVAL_GRID above 1024). All calls on one GPU share its
 * scratch array: calls that overlap on one GPU race, so queue them on one stream. */
int tpa_eval_noalloc(tpa_events_t ev, const double* coef, int mode, int flags, double* d_out,
                     void* stream);

/* The fused pass: tpa_eval's sums (modes DATA and MC; TPA_FLAG_GRAD for G_k) computed from the
 * angle cache alone, every partial wave recomputed per event (Wigner D of both vertices, CG x
 * barrier factors, the lineshape at the model's current m0/g0, the resonance-helicity
 * contraction and the aligned rotation of spin > 0 finals), so no partial-wave cache is
 * needed. The kernel is generated for the model's description and compiled for gfx950 with
 * hipRTC on first use (regenerated after tpa_model_update only if its source changes).
 * Three-body models with at most TPA_FUSED_MAX_P partial waves (TPA_ERR_UNSUPPORTED above: the
 * generated kernel of a 22-wave model took a memory-aperture fault on gfx950 whose cause is
 * not identified, so larger models are refused). Replaces the per-call amplitude of tf-pwa's default model, which
 * recomputes the chains from the cached angle and D-matrix data on every evaluation
 * (DecayGroup.get_amp / DecayChain.get_amp, tf_pwa/amp/core.py:1392-1445, 1696-1728; the
 * D-matrix cache of tf_pwa/dfun.py:221-238), inside sum_gradient (tf_pwa/model/model.py:41-94). */
int tpa_eval_fused(tpa_events_t ev, const double* coef, int mode, int flags, double* out);

/* The HIP source of a three-body description's fused pass (host only, no GPU needed): *len
 * receives its length; with buf (buf_len bytes) the text is copied, NUL-terminated. The same
 * interface as tpa_model_create (tf_pwa/config_loader/config_loader.py:210-237). */
int tpa_fused_source(const void* desc, int64_t desc_bytes, const double* cg, int n_cg,
                     char* buf, int64_t buf_len, int64_t* len);

/* Seconds the model's last fused-pass compile took (0 before the first). Bookkeeping only, no
 * reference counterpart. */
int tpa_fused_compile_seconds(tpa_model_t model, double* s);

/* Kernel timing with HIP events on the engine's stream. enable(k), k >= 1: one likelihood
 * pass in k is timed, counting from the call (k = 1: every pass; k <= 0: off). A timed pass
 * of tpa_eval records its events from its own dispatch (the kernel's execution, as a kernel
 * trace reports it); the rebuild and fused entries bracket their whole sequence. read
 * returns the summed milliseconds and the number of timed launches since the last reset. */
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

/* The same Gram matrix on FP64 MFMA (v_mfma_f64_16x16x4f64): products and sums in FP64, for
 * amplitudes whose dynamic range makes the FP32 partial sums of tpa_gram drift (narrow peaks
 * over a broad MC sample). Same output layout and the same reference interface as tpa_gram
 * (tf_pwa/experimental/opt_int.py:30-52). P <= 32. ms (optional): the MFMA kernel's duration. */
int tpa_gram_f64(tpa_events_t ev, double* out, double* ms);

/* Second moments for the exact Hessian (fold = 1 cache, P <= 32). out (host, 2 + 2P + Q(Q+1)
 * doubles, Q = 2P): out[0] the sum of the mode, out[2+2k], out[3+2k] Re/Im G_k as in tpa_eval,
 * then the packed upper triangles (row-major, i <= j) of
 *   O1 = sum_e v_e z_e z_e^T with z_e = (Re q_e, Im q_e), q_ek = sum_lambda conj(A) h_k,
 *   O2 = sum_e u_e sum_lambda y y^T with y = (Re h_.lambda, Im h_.lambda),
 * u_e, v_e the first and second derivative weights of the mode (MC: u = w, v = 0, so O2 is the
 * FP64 Gram matrix). The second derivatives of the sum in (c, conj c) are
 *   d2/dc_k dc_l = sum v q_k q_l,  d2/dc_k dconj(c_l) = sum v q_k conj(q_l) + sum u h_k conj(h_l).
 * Replaces the nested tf.GradientTape Hessians (sum_hessian, tf_pwa/model/model.py:97-150) of
 * BaseModel.nll_grad_hessian (513-560) for the coupling parameters. */
int tpa_hess(tpa_events_t ev, const double* coef, int mode, double* out);

/* Moments of the exact Hessian when masses and widths float (three-body, fold = 0 cache, the
 * model's current lineshape parameters, P <= 24). With X_(t,k) = b^t h_k the extended waves
 * (t = 0, 1, 2: the wave's lineshape B, dB/dm0 and dB/dg0), A = sum_k c_k X_(0,k), and u, v
 * the mode's first and second derivative weights, out (host, 2 + 12P + Q(Q+1) doubles,
 * Q = 6P) receives: out[0] the sum of the mode; then F_j = sum_e u z_j, z_j = sum_lambda
 * conj(A) X_j (j < 3P, interleaved re/im; F_(0,k) is tpa_eval's G_k); then
 * F2_(t,k) = sum_e u d2B_t q_k (t: d2/dm0^2, d2/dm0dg0, d2/dg0^2; q_k = sum_lambda conj(A) h_k;
 * 3P complex); then the packed upper triangles of O1 = sum_e v y y^T, y = (Re z, Im z), and
 * O2 = sum_e u sum_lambda x x^T, x = (Re X, Im X). Replaces the nested tf.GradientTape
 * Hessians (sum_hessian, tf_pwa/model/model.py:97-150) of BaseModel.nll_grad_hessian
 * (513-560) for models with `float: mg` resonances (the mass and width Variables of
 * tf_pwa/amp/core.py:380-390). */
int tpa_hess_shape(tpa_events_t ev, const double* coef, int mode, double* out);

/* The same for the data range [offset, offset+n) of a pool allocation (TPA_MODE_DATA). */
int tpa_pool_hess(tpa_pool_t pool, int64_t offset, int64_t n, const double* coef, double* out);

/* Wait for all device work of the process (the timing barrier of benchmarks). */
int tpa_device_synchronize(void);

/* Measurement helper: copy `bytes` between two fresh device buffers `reps` times with a
 * plain 16-byte streaming kernel (k_copy) and return the average milliseconds per copy.
 * It is the calibration kernel for the FETCH_SIZE / WRITE_SIZE counters. */
int tpa_calib_copy(int64_t bytes, int reps, double* ms);

/* ---- Many fits on one GPU (the serving path).
 * A pool is one device event store (angles, fold-1 partial-wave cache, weights) that holds the
 * data samples of many concurrent fits. Each fit owns a contiguous range. One call of
 * tpa_pool_eval evaluates the data sum and its adjoint for every pending request of every fit in
 * a single batched launch (a hipGraph per grid size class, replayed), replacing one
 * BaseModel.nll_grad_batch call per fit (tf_pwa/model/model.py:407-449, driven by
 * fit_scipy's minimize loop, tf_pwa/fit.py:189-300) when many fits run at once.
 * The MC normalisation of these fits is shared and comes from tpa_gram on an event set. */

/* capacity: events the pool can hold, rounded up to a multiple of 64 (the row stride of the
 * pool's per-event arrays, kept on 512-byte boundaries; tpa_pool_free_events of a new pool
 * gives the rounded value); max_items: the most requests one tpa_pool_eval serves. */
int tpa_pool_create(tpa_model_t model, int64_t capacity, int max_items, tpa_pool_t* out);
int tpa_pool_destroy(tpa_pool_t pool);

/* Upload n events (p4 host [n_final*4][n], weight host [n]) into a free range, compute their
 * angles and partial-wave cache there, and return the range's first event in *offset.
 * Fails with "pool full" when no free range is long enough (the caller's admission control). */
int tpa_pool_add(tpa_pool_t pool, const double* p4, int64_t n, const double* weight,
                 int64_t* offset);

/* tpa_pool_add for the events of a device event set of the same model (their angles, copied on
 * the device; weight host [n]): a sample generated on the GPU (tpa_events_generate +
 * tpa_events_select) enters the pool without a host round trip. Replaces the hand-off of each
 * toy sample of tf-pwa's toy studies (gen_data's accept-reject sample, tf_pwa/applications.py:
 * 394-487, written out and read back as the fit's data). */
int tpa_pool_add_events(tpa_pool_t pool, tpa_events_t ev, const double* weight,
                        int64_t* offset);

/* on != 0: keep an FP32 copy (round to nearest) of the pool's partial-wave cache, made now for
 * every allocated range and by every later tpa_pool_add; the batched passes (tpa_pool_eval and
 * the serving runtime) then read it (half the bytes per event, FP64 arithmetic after the load).
 * Needs >= 2 helicity slots and no batch in flight. tpa_pool_hess keeps reading the FP64 cache.
 * No reference counterpart: the storage option of the mixed-precision data pass. */
int tpa_pool_keep_f32(tpa_pool_t pool, int on);

/* Return the range that starts at offset to the free list (merged with free neighbours). */
int tpa_pool_release(tpa_pool_t pool, int64_t offset);

/* Free events left in the pool. */
int64_t tpa_pool_free_events(tpa_pool_t pool);

/* Batched data sums: request i covers events [ranges[2i], ranges[2i] + ranges[2i+1]) of one
 * allocation, with coefficients coefs[i*2P .. i*2P+2P) (interleaved re/im). out[i*(2+2P) + ...]
 * has tpa_eval's layout for TPA_MODE_DATA with TPA_FLAG_GRAD: out[0] = sum w clip_log(f),
 * out[2+2k], out[3+2k] = Re/Im G_k. Each request's result does not depend on the other
 * requests of the batch (fixed per-request chunking and reduction order). mode must be
 * TPA_MODE_DATA. ms (optional): the batched graph's duration. */
int tpa_pool_eval(tpa_pool_t pool, int n_items, const int64_t* ranges, const double* coefs,
                  int mode, double* out, double* ms);

/* ---- The serving runtime: many BFGS fits at once over one pool.
 * The coupling map says how the partial-wave coefficients follow from a fit's trainable
 * parameters y (tf-pwa's trainable_vars order): c_k = wave_const[k] * prod_a v[f_a] over the
 * TPA_CPL_SLOTS factors f = wave_factor[4k..4k+3] (the chain total and the g_ls of each vertex,
 * tf_pwa/amp/core.py:1000-1080; a fixed factor 1 fills unused slots) and
 * v[f] = r e^{i phi} (factor_polar) or r + i phi, where r and
 * phi are y[factor_var[2f]] and y[factor_var[2f+1]], or factor_fixed[2f], [2f+1] when the index
 * is -1. bound_kind[j] (0 none, 1 [a, b], 2 lower a, 3 upper b, with a, b = bound_ab[2j],
 * [2j+1]) is the bound transform of y[j] (tf_pwa/variable.py:1029-1135); BFGS runs in the
 * unbounded variables. */
#define TPA_CPL_SLOTS 4

typedef struct {
  int32_t P;
  int32_t n_factor;
  int32_t n_var;
  int32_t reserved;
  const int32_t* wave_factor;   /* [P][TPA_CPL_SLOTS] */
  const double* wave_const;     /* [P] */
  const int32_t* factor_polar;  /* [n_factor] */
  const int32_t* factor_var;    /* [n_factor][2] */
  const double* factor_fixed;   /* [n_factor][2] */
  const int32_t* bound_kind;    /* [n_var] */
  const double* bound_ab;       /* [n_var][2] */
} TpaCouplingMap;

/* A server evaluates NLL = -sum_e w_e clip_log f_e + sw log(c^dagger I c) (Model.nll,
 * tf_pwa/model/model.py:407-449) for every fit: the data sums of all pending requests go to the
 * GPU in one tpa_pool_eval batch, the MC normalisation is the shared Gram matrix `gram`
 * (host [P][P] complex, from tpa_gram on MC with unit-sum weights). Each fit minimises with
 * BFGS on a host thread (strong-Wolfe line search, stop at max |g| <= gtol; maxiter <= 0 means
 * 200 * n_var), like fit_scipy (tf_pwa/fit.py:189-300). The batcher launches when every fit is
 * waiting, when max_batch requests are pending, or window_us after the first one arrived.
 * hess_start != 0: each fit first takes one exact Hessian (tpa_pool_hess) at its start point and
 * starts BFGS from its inverse (shifted to positive definite) with a unit first step, instead
 * of from the identity. */
int tpa_server_create(tpa_pool_t pool, const TpaCouplingMap* map, const double* gram,
                      double gtol, int maxiter, int max_batch, double window_us, int hess_start,
                      tpa_server_t* out);

/* Stops the batcher, cancels running fits (status 3) and joins every thread. */
int tpa_server_destroy(tpa_server_t server);

/* One NLL and gradient (in y) evaluation of the data range [offset, offset+n) with data weight
 * sum sw, through the batcher (FCN.nll_grad, tf_pwa/model/model.py:1216-1240). */
int tpa_server_nll_grad(tpa_server_t server, int64_t offset, int64_t n, double sw,
                        const double* y, double* nll, double* grad);

/* NLL, gradient and exact Hessian (n_var x n_var, row-major) in y from one second-moment pass
 * (tpa_pool_hess) and the Gram matrix (FCN.nll_grad_hessian, tf_pwa/model/model.py:1258-1265). */
int tpa_server_nll_grad_hess(tpa_server_t server, int64_t offset, int64_t n, double sw,
                             const double* y, double* nll, double* grad, double* hess);

/* Start a fit on the data range [offset, offset+n) from parameters y0; returns its id. */
int tpa_server_submit(tpa_server_t server, int64_t offset, int64_t n, double sw,
                      const double* y0, int* fit_id);

/* Ids of fits that ended since the last poll (at most max_ids). */
int tpa_server_poll(tpa_server_t server, int* ids, int max_ids, int* n_done);

/* Result of an ended fit: y (n_var), NLL at y, status (0 converged, 1 iteration limit, 2 line
 * search failed, 3 cancelled, 4 engine error), evaluations and iterations. */
int tpa_server_result(tpa_server_t server, int fit_id, double* y, double* nll, int* status,
                      int64_t* nfev, int64_t* niter);

/* Per-request latency (submit of the request to its result) of every request so far, in
 * microseconds; *n receives the total count, at most max_n values are copied. */
int tpa_server_latencies(tpa_server_t server, double* us, int64_t max_n, int64_t* n);

/* Requests served, batches launched, events summed over all requests and summed
 * batched-graph milliseconds (any pointer may be NULL). */
int tpa_server_counters(tpa_server_t server, int64_t* requests, int64_t* batches,
                        int64_t* events, double* gpu_ms);

/* ---- Single-fit likelihood session (fixed lineshapes).
 * NLL = -sum_data w clip_log f + sw log S_mc and its gradient in the trainable values y, with
 * the coefficients from the coupling map (see TpaCouplingMap): one tpa_eval pass per data set
 * and, when gram is NULL, per MC set (FP64 path); with a Gram matrix (host [P][P] complex of the
 * unit-sum MC weights) S_mc = c^dagger I c (mixed path). The event sets' caches (fold = 1) and
 * weights must be in place; the session keeps the handles, not copies. Replaces one
 * FCN.nll_grad call (tf_pwa/model/model.py:1207-1240) for couplings-only fits. */
int tpa_fcn_create(const TpaCouplingMap* map, const tpa_events_t* data, int n_data,
                   const tpa_events_t* mc, int n_mc, const double* gram, double sw,
                   tpa_fcn_t* out);
int tpa_fcn_destroy(tpa_fcn_t fcn);

/* grad may be NULL (value only). */
int tpa_fcn_nll_grad(tpa_fcn_t fcn, const double* y, double* nll, double* grad);

/* Floating masses and widths of three-body chains (resonance R of chain c decaying at the top
 * vertex top -> R + bachelor and at R -> y + z). Per evaluation the session takes m0, g0 of
 * each chain from y (or the fixed values), recomputes the barrier constants sqrt(P_l(z0)) of
 * both vertices (tf_pwa/amp/core.py:1072-1097 with q0 from the masses, core.py:311-334), sets
 * the device lineshape parameters, evaluates with per-event lineshapes and their derivatives
 * (TPA_FLAG_SHAPE) and adds the mass and width gradient:
 *   dNLL/dm0_c = 2 Re sum_{k in c} c_k (dlog_k F_k + Dm_k),  dNLL/dg0_c = 2 Re sum_{k in c} c_k Dg_k
 * with dlog_k the m0 derivative of log(wave constant). The caches must be built without
 * lineshapes (fold 0) and with the clamp flags act1/act2; when a mass crosses the point where
 * a flag changes, the session sets the new flags on the device and rebuilds the caches of its
 * event sets (tpa_fcn_cache_rebuilds counts this, so that a host keeping its own view of the
 * caches can refresh it). The MC normalisation is the full MC
 * pass (a Gram matrix cannot carry per-event lineshapes). Replaces FCN.nll_grad of models with
 * `float: mg` resonances (tf_pwa/model/model.py:1207-1240, the mass/width Variables of
 * tf_pwa/amp/core.py:380-390). */
#define TPA_SHAPE_MAX_CHAIN 32
#define TPA_SHAPE_MAX_LS 8
#define TPA_SHAPE_MAX_P 128
typedef struct TpaShapeMap {
  int32_t n_chain, P;
  int32_t mass_var[TPA_SHAPE_MAX_CHAIN], width_var[TPA_SHAPE_MAX_CHAIN]; /* y index or -1 */
  double mass_fixed[TPA_SHAPE_MAX_CHAIN], width_fixed[TPA_SHAPE_MAX_CHAIN];
  double m_top;
  double m_bach[TPA_SHAPE_MAX_CHAIN], m_y[TPA_SHAPE_MAX_CHAIN], m_z[TPA_SHAPE_MAX_CHAIN];
  double d[TPA_SHAPE_MAX_CHAIN];                                     /* barrier radius */
  int32_t n_l1[TPA_SHAPE_MAX_CHAIN], n_l2[TPA_SHAPE_MAX_CHAIN];
  int32_t l1[TPA_SHAPE_MAX_CHAIN][TPA_SHAPE_MAX_LS], l2[TPA_SHAPE_MAX_CHAIN][TPA_SHAPE_MAX_LS];
  int32_t act1[TPA_SHAPE_MAX_CHAIN][TPA_SHAPE_MAX_LS], act2[TPA_SHAPE_MAX_CHAIN][TPA_SHAPE_MAX_LS];
  int32_t wave_chain[TPA_SHAPE_MAX_P], wave_ls1[TPA_SHAPE_MAX_P], wave_ls2[TPA_SHAPE_MAX_P];
} TpaShapeMap;

/* Attach a shape map to a session built without a Gram matrix; model: the event sets' model
 * (its device lineshape parameters are set per evaluation). */
int tpa_fcn_set_shape_map(tpa_fcn_t fcn, tpa_model_t model, const TpaShapeMap* map);

/* data_f32 != 0: the session's data passes read the FP32 copies of the data caches
 * (TPA_FLAG_F32; the copies are kept from here on, tpa_events_keep_f32). Fixed lineshapes
 * only. The MC side is unchanged (FP64 passes or the Gram matrix). No reference counterpart:
 * the storage option of the data pass behind FCN.nll_grad (tf_pwa/model/model.py:1207-1229). */
int tpa_fcn_set_data_storage(tpa_fcn_t fcn, int data_f32);

/* The session's data passes read the helicity-halved copies of the data caches
 * (tpa_events_keep_half): FP32 (TPA_FLAG_F32H) when kind = 1, FP64 (TPA_FLAG_F64H) when
 * kind = 2, the FP64 caches when kind = 0. Fixed lineshapes only. No reference counterpart: a
 * storage option of the data pass behind FCN.nll_grad (tf_pwa/model/model.py:1207-1229). */
int tpa_fcn_set_data_half(tpa_fcn_t fcn, int kind);

/* on != 0: the session's MC passes (FP64 path) read the helicity-halved FP64 copies of the MC
 * caches (TPA_FLAG_F64H); on = 0: the FP64 caches. Fixed lineshapes only. No reference
 * counterpart: a storage option of the MC pass behind FCN.nll_grad
 * (tf_pwa/model/model.py:1207-1229). */
int tpa_fcn_set_mc_half(tpa_fcn_t fcn, int on);

/* Cache rebuilds the session made after barrier clamp flags changed (no reference
 * counterpart: a bookkeeping counter for hosts that track the caches' state). */
int tpa_fcn_cache_rebuilds(tpa_fcn_t fcn, int64_t* n);

/* A whole fit on the session: BFGS from y0 (physical values) in the unbounded variables of the
 * bound transforms, strong-Wolfe line search, until |grad|_inf <= gtol in those variables or
 * maxiter iterations (<= 0: 200 n), every step one tpa_fcn_nll_grad call. The loop of
 * fit_scipy / scipy.optimize.minimize(method="BFGS") (tf_pwa/fit.py:189-300) without a
 * round trip to the host interpreter per evaluation. status: 0 converged, 1 iteration limit,
 * 2 line search failed (precision loss). Returns the error of a failed evaluation with nfev
 * the evaluations done. */
int tpa_fcn_minimize(tpa_fcn_t fcn, const double* y0, double gtol, int maxiter, double* y_out,
                     double* nll_out, int* status, int64_t* nfev, int64_t* niter);

/* Set the lineshape parameters (m0, g0) of every chain of a three-body model on the device,
 * ordered with later work on the model's stream (the fields of TpaChain that
 * tpa_model_update would replace; no other field changes). */
int tpa_model_set_lineshapes(tpa_model_t model, int n_chain, const double* m0, const double* g0);

#ifdef __cplusplus
}
#endif
#endif
