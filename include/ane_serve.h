/* ane_serve.h -- persistent probe-update request loop over many resident scenes.
 *
 * No counterpart in the reference: AneRend updates the probes of its one scene
 * from the render loop (IrradianceProbeRayTracingPass.cpp:95-165,
 * IrradianceProbeConvolvePass.cpp:109-143).  This is the serving form of the
 * same update (ane_ddgi_update) for many scenes at once:
 *
 *  - a session = a resident scene (ane_scene, its LBVH) + a probe grid
 *    (atlas and work buffers); sessions are admitted against an HBM budget
 *    (context bytes, including the traversal spill and sun grid the updates
 *    use, allocated at admission, + each distinct scene's bytes once) or
 *    refused with ANE_ERR_ADMISSION_REFUSED;
 *  - requests (one update of one session with its frame constants) are queued
 *    from any host thread; a dispatcher thread launches every queued request
 *    whose session has no update on the device while fewer than max_inflight
 *    are in flight, on the least-loaded of num_streams HIP streams, so work of
 *    different sessions overlaps on the device and a request that arrives while
 *    others run starts as soon as a slot frees;
 *  - with fuse_stages the requests launched together run as one batch: their
 *    shading stages are one launch and their convolutions one launch (the
 *    per-probe stages of all of them, addressed through a job table), after
 *    each request's own trace (the traces spread over the streams);
 *  - a session's updates run in submission order and its results depend only
 *    on its own frames: a request gives the same atlas alone and in a mix.
 *
 * Threading: every entry point may be called from any host thread.  Results
 * are kept until polled. */
#ifndef ANE_SERVE_H
#define ANE_SERVE_H

#include <stdint.h>

#include "ane_ddgi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ane_serve_config {
  int64_t hbm_budget_bytes; /* admission budget for resident sessions */
  int32_t max_inflight;     /* requests on the device at once, 1..256 */
  int32_t num_streams;      /* HIP streams the requests are spread over, 1..16 */
  int32_t capture_graphs;   /* fuse_stages 0 only. 1: each update is captured into the session's
                               HIP graph (updated in place, one launch per request); 0: direct launches */
  int32_t fuse_stages;      /* 1: the requests launched together form a batch: each request's frame
                               setup and trace, on the streams in turn (so the traces overlap), then
                               on the batch's stream ONE launch shading the probes of every request
                               in the batch and ONE launch convolving them (per precision present),
                               through a request-descriptor table; 0: each request is launched on
                               its own (graph or direct) */
} ane_serve_config;

typedef struct ane_serve_result {
  int32_t status;     /* ANE_OK or the update's status */
  int32_t session;
  double queue_ms;    /* submission -> launch */
  double latency_ms;  /* submission -> completion seen by the dispatcher */
  int64_t rays;       /* primary rays of the request */
} ane_serve_result;

typedef struct ane_serve_stats {
  int32_t sessions, queued, inflight, peak_inflight;
  uint64_t completed, failed;
  int64_t rays_done;
  int64_t used_bytes, budget_bytes;
  int64_t held_bytes; /* device bytes the sessions and their scenes hold now (= used_bytes:
                         admission charges each session's spill and sun grid up front) */
} ane_serve_stats;

typedef struct ane_serve ane_serve;

int ane_serve_open(const ane_serve_config* config, ane_serve** out);
/* Creates the session's probe grid for `scene` (the caller keeps the scene
 * alive until the session is removed).  ANE_ERR_ADMISSION_REFUSED when the
 * budget would be exceeded (nothing is kept). */
int ane_serve_add_session(ane_serve* srv, const ane_scene* scene, const ane_ddgi_config* config,
                          int32_t* session);
/* Waits for the session's queued requests, then frees its grid. */
int ane_serve_remove_session(ane_serve* srv, int32_t session);
/* Queues one update (layers of `frame`) of a session; never blocks on the device. */
int ane_serve_submit(ane_serve* srv, int32_t session, const ane_ddgi_frame* frame,
                     uint64_t* ticket);
/* ANE_OK with the result (the ticket is then released), or ANE_ERR_NOT_READY
 * after `timeout_ms` (0 = do not wait, < 0 = wait for completion). */
int ane_serve_poll(ane_serve* srv, uint64_t ticket, int32_t timeout_ms, ane_serve_result* out);
/* Atlas of a session after all its submitted requests (FP32 RGBA, reference layout). */
int ane_serve_session_atlas(ane_serve* srv, int32_t session, float* host_rgba);
int ane_serve_stats_get(ane_serve* srv, ane_serve_stats* out);
/* Drains queued and running requests, frees every session and the server. */
int ane_serve_close(ane_serve* srv);

#ifdef __cplusplus
}
#endif
#endif /* ANE_SERVE_H */
