/*
 * prt_server.h — device-resident progressive-render serving loop (continuous batching).
 *
 * Replaces the per-request pipeline of the reference web server: handleRender builds a scene and a
 * ProgressiveRaytracer per request and streams RenderProgressive's passes and tiles as SSE events
 * (web/server/render.go:45-81, 183-273, 285-368). Here scenes (BVH + materials + lights) stay resident on
 * the GPU and are shared by requests; each admitted request gets a film slot on its own HIP stream, and
 * requests that arrive while others render are admitted as slots and HBM budget free up, so their passes
 * overlap on the device. Events mirror the SSE stream: pass complete (passComplete, render.go:285-317),
 * done (complete), error, cancelled (context cancellation, progressive.go:307-314). Tiles are read with
 * prt_server_read_rgba8 (TileUpdate, render.go:22-30, 332-368).
 *
 * Results do not depend on the mix: each request renders with its own counter-RNG key and its own
 * chunking, so its image is bit-identical alone and among other requests.
 */
#ifndef PRT_SERVER_H
#define PRT_SERVER_H

#include <stdint.h>

#include "prt.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PRT_EVENT_ADMITTED 0      /* film slot allocated, first pass enqueued */
#define PRT_EVENT_PASS_COMPLETE 1 /* pass `pass_number` is readable */
#define PRT_EVENT_DONE 2          /* last pass complete */
#define PRT_EVENT_ERROR 3
#define PRT_EVENT_CANCELLED 4

typedef struct prt_server_event {
    int64_t request_id;
    int32_t type;
    int32_t pass_number;
    int32_t target_samples;
    int32_t error_code;
    double t_ms;        /* milliseconds since the request was submitted */
    double device_ms;   /* device time of this pass (PASS_COMPLETE) */
} prt_server_event;

typedef struct prt_server_stats {
    int32_t queued;
    int32_t active;
    int64_t completed;
    int64_t budget_bytes;
    int64_t used_bytes;  /* film slots of admitted, unreleased requests */
    int64_t idle_bytes;  /* device buffers kept from released films for reuse (within the budget) */
} prt_server_stats;

typedef struct prt_server prt_server;

/* hbm_budget_bytes bounds the memory of all active film slots together with the device buffers the
 * server keeps from released films for reuse (films of later requests take them instead of allocating;
 * a reused buffer may exceed what a film asked for by up to a quarter); max_active bounds concurrent
 * requests. */
int prt_server_create(int32_t device, int64_t hbm_budget_bytes, int32_t max_active, prt_server** out);
/* Queues a progressive render of `scene` (params: integrator, width/height override, depth, RR, pass
 * schedule, seed). The scene must outlive the request: prt_scene_destroy fails while a request that
 * has not been released (or a film) still uses it. */
int prt_server_submit(prt_server* server, const prt_scene* scene, const prt_render_params* params,
                      int64_t* request_id);
/* Flags of prt_server_submit_ex. */
#define PRT_SUBMIT_SNAPSHOTS 1u /* keep a copy of the film at the end of each pass (the newest three) */
/* prt_server_submit with flags. A PRT_SUBMIT_SNAPSHOTS request also holds three film copies in the HBM
 * budget, so that its passes can be read exactly as they ended while later passes render
 * (prt_server_read_pass; RenderPass hands each pass's image to its caller, progressive.go:126-223). */
int prt_server_submit_ex(prt_server* server, const prt_scene* scene, const prt_render_params* params,
                         uint32_t flags, int64_t* request_id);
/* Copies up to max_events pending events; waits up to timeout_ms for the first one (0: no wait). */
int prt_server_poll(prt_server* server, prt_server_event* events, int32_t max_events, int32_t timeout_ms,
                    int32_t* count);
int prt_server_cancel(prt_server* server, int64_t request_id);
/* Film snapshot after the last enqueued pass of the request (mean RGB, per-pixel sample counts). */
int prt_server_read(prt_server* server, int64_t request_id, float* mean_rgb, uint32_t* counts);
int prt_server_read_rgba8(prt_server* server, int64_t request_id, int32_t x0, int32_t y0, int32_t x1,
                          int32_t y1, uint8_t* rgba);
/* The film as pass `pass_number` left it (mean RGB, counts), for PRT_SUBMIT_SNAPSHOTS requests; fails
 * for a pass that has not completed or is older than the three newest. */
int prt_server_read_pass(prt_server* server, int64_t request_id, int32_t pass_number, float* mean_rgb,
                         uint32_t* counts);
/* Frees the film slot of a finished or cancelled request. */
int prt_server_release(prt_server* server, int64_t request_id);
int prt_server_stats_get(prt_server* server, prt_server_stats* stats);
int prt_server_destroy(prt_server* server);

#ifdef __cplusplus
}
#endif
#endif /* PRT_SERVER_H */
