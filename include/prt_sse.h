/*
 * prt_sse.h — the progressive-render event stream of the reference web server, as text.
 *
 * Replaces the SSE side of handleRender (web/server/render.go:45-81, 183-413): for every finished pass
 * the reference sends each 64x64 tile twice (once as the tiles complete, once after the pass's splats
 * are applied, RenderPass progressive.go:161-222) as `event: tile` with TileUpdate JSON holding a
 * base64 PNG of the tile, then `event: passComplete` with the pass statistics, and at the end
 * `event: complete`. Failures are `event: error`. This library produces that byte stream from a film
 * snapshot (mean RGB and per-pixel sample counts, as prt_film_read or prt_server_read_pass return
 * them); the HTTP side (headers, flushing, client disconnects) stays with the caller.
 *
 * Differences from the reference stream: both tile rounds carry the pass's final image (the GPU pass
 * applies its splats as it renders, so there is no pre-splat state to show) and tiles come in grid
 * order rather than worker completion order. PNG bytes differ from Go's encoder (zlib level and
 * filters); the decoded pixels are the same RGBA8 (gamma 2, clamp, truncation, vec3ToColor
 * progressive.go:482-495), written as 8-bit RGB because the image is opaque, as Go's png.Encode does.
 *
 * Returned buffers are owned by the caller and freed with prt_free. All functions here are host code
 * and need no GPU.
 */
#ifndef PRT_SSE_H
#define PRT_SSE_H

#include <stdint.h>

#include "prt.h"
#include "prt_server.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct prt_sse_pass_info {
    int32_t pass_number;          /* PassResult.PassNumber */
    int32_t total_passes;         /* ProgressiveConfig.MaxPasses */
    int32_t target_samples;       /* getSamplesForPass(pass) = RenderStats.MaxSamples */
    int32_t max_samples_per_pixel;/* ProgressiveConfig.MaxSamplesPerPixel (MinSamples starts there) */
    int32_t tile_size;            /* DefaultTileSize = 64 */
    int32_t pad;
    int64_t primitive_count;      /* Scene.GetPrimitiveCount (prt_scene_info.reference_primitives) */
    int64_t elapsed_ms;           /* since the render started */
} prt_sse_pass_info;

/* PNG (8-bit RGB, opaque) of a width x height RGBA8 image (imageToBase64PNG before base64,
 * render.go:407-413). */
int prt_png_encode_rgba8(const uint8_t* rgba, int32_t width, int32_t height, uint8_t** png, int64_t* length);

/* The tile events (twice per tile) and the passComplete event of one pass, from the film's mean RGB
 * (width*height*3) and sample counts (width*height). */
int prt_sse_format_pass(int32_t width, int32_t height, const float* mean_rgb, const uint32_t* counts,
                        const prt_sse_pass_info* info, char** text, int64_t* length);

/* `event: complete` (render.go:258-266) or `event: error` with the message (handleError :415-422). */
int prt_sse_format_complete(char** text, int64_t* length);
int prt_sse_format_error(const char* message, char** text, int64_t* length);

/* The SSE text for one event of a request that was submitted with PRT_SUBMIT_SNAPSHOTS: a
 * PASS_COMPLETE event reads that pass's snapshot (prt_server_read_pass) and formats it, DONE gives
 * `complete`, ERROR gives `error: Rendering failed: ...`; ADMITTED and CANCELLED give no text
 * (*length = 0; the reference sends nothing when the client goes away). `params` are the request's
 * parameters, `tile_size` 64 as in the reference. */
int prt_sse_server_event(prt_server* server, const prt_scene* scene, const prt_render_params* params,
                         int32_t tile_size, const prt_server_event* event, char** text, int64_t* length);

void prt_free(void* p);

#ifdef __cplusplus
}
#endif
#endif /* PRT_SSE_H */
