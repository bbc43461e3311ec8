/* fxcorr_serve.h — the request loop over libfxcorr_gpu: subintegrations of
 * several concurrent correlation configurations on one GPU.
 *
 * Replaces FxManager's dispatch of subintegrations to Core slots
 * (mpifxcorr/src/fxmanager.cpp:306-534: round-robin send, resend on return,
 * RECEIVE_RING_LENGTH slots per Core, core.cpp:166) with a continuous-batching
 * loop: each configuration gets "slots" (one fxg_job each: its own stream and
 * its own device-resident spectra, accumulators and tables), created on demand
 * while their HBM footprint (fxg_plan_info.device_bytes) fits the budget; a
 * request that arrives while every slot of its configuration is busy waits
 * only for a slot of that configuration, so work of different configurations
 * and slots overlaps on the device.  Latency = completion - submission.
 *
 * Thread-safe; one server per process and device.  All functions return 0 or
 * a negative FXG_E* code; fxg_last_error() has the text for this thread. */
#ifndef FXCORR_SERVE_H
#define FXCORR_SERVE_H

#include <stdint.h>
#include "fxcorr_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fxs_server fxs_server;

typedef struct {
  int64_t budget_bytes;     /* HBM the server may hold                      */
  int64_t used_bytes;       /* footprint of the slots created so far        */
  int32_t nconfigs;
  int32_t nslots;           /* slots (job instances) over all configs       */
  int64_t submitted;
  int64_t completed;
  int64_t queued;           /* requests waiting for a slot                  */
  int64_t refused;          /* slot creations refused by the budget         */
} fxs_stats_t;

/* budget_bytes <= 0: 90% of the device's free HBM at creation. */
int fxs_create(int device, int64_t budget_bytes, fxs_server** out);
/* Waits for every submitted request, then frees all slots. */
int fxs_destroy(fxs_server* s);
/* A configuration whose requests may use up to max_slots concurrent slots.
 * Fails with FXG_ENOMEM when not even one slot fits the budget. */
int fxs_add_config(fxs_server* s, const fxg_job_desc* desc, int max_slots, int* config_id);
/* One subintegration: stations[ndatastreams] with device data pointers
 * (valid until the request completes), results a device buffer of
 * result_length cf32 (fxg_plan).  Returns at once with a request id. */
int fxs_submit(fxs_server* s, int config_id, const fxg_subint_desc* sub,
               const fxg_station_input* stations, float* results_device, int64_t* request);
/* Blocks until the request completes; its latency in ns (may be NULL). */
int fxs_wait(fxs_server* s, int64_t request, int64_t* latency_ns);
/* At most n requests on the device at once (0: no limit, the default);
 * requests start in submission order among those a slot has taken, so a
 * bound trades overlap for first-come service of large requests.  No
 * reference counterpart (FxManager hands every free Core a subintegration,
 * fxmanager.cpp:306-342). */
int fxs_set_max_inflight(fxs_server* s, int n);
int fxs_stats(fxs_server* s, fxs_stats_t* out);

#ifdef __cplusplus
}
#endif
#endif
