// Internal glue between the library's translation units (not part of the C ABI).
#pragma once
#include <stdint.h>

#include <string>

#include "tfpwa_amd.h"

// Sets the calling thread's tpa_last_error() message and returns code.
int tpa_set_error(int code, const std::string& msg);

// Pipelined batched evaluation on an event pool (tpa_pool_eval split in two): submit queues
// a batch on one of the pool's batch slots (*slot_io = -1: any free slot; the slot used is
// returned) without waiting; ready polls it (1 done, 0 running); finish waits, copies the
// n_items x (2 + 2P) sums to out and frees the slot. A slot belongs to its submitter until
// finished.
int tpa_pool_eval_submit(tpa_pool_t p, int* slot_io, int n_items, const int64_t* ranges,
                         const double* coefs, int mode);
int tpa_pool_eval_ready(tpa_pool_t p, int slot);
int tpa_pool_eval_finish(tpa_pool_t p, int slot, double* out, double* ms);

// Replace the barrier clamp flags (act1 / act2, [n_chain][TPA_MAX_LS]) of a three-body model's
// chains on the device; the event sets' partial-wave caches must be rebuilt afterwards.
extern "C" int tpa_model_set_barrier_flags(tpa_model_t model, int n_chain,
                                           const int32_t* act1, const int32_t* act2);

// Queue one shape-mode evaluation pass (TPA_FLAG_SHAPE, no f_out) on `stream` (a hipStream_t;
// null: the model's stream) without waiting; tpa_eval_collect waits for that stream and copies
// the sums (tpa_eval's layout). Passes of different event sets may be queued on different
// streams and overlap.
extern "C" int tpa_eval_enqueue(tpa_events_t ev, const double* coef, int mode, int flags,
                                void* stream);
extern "C" int tpa_eval_collect(tpa_events_t ev, double* out);
// The model's stream (a hipStream_t): the lineshape parameters are set on it.
extern "C" int tpa_model_stream(tpa_model_t model, void** stream);
