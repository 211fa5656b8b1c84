// prt_internal.h — host-side objects behind the opaque C handles.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <atomic>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/prt.h"
#include "prt_common.h"

namespace prt {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define PRT_HIP(call)                                      \
    do {                                                   \
        hipError_t _e = (call);                            \
        if (_e != hipSuccess) return hip_fail(_e, #call);  \
    } while (0)

// Device allocation owned by a scene or film.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// A per-scene autotuning job: timing launches enqueued with events on the stream of the film that
// started it. No host thread waits for them; a later chunk of any film of the scene reads the events once
// the last has completed and records the winner.
struct TuneJob {
    std::mutex mu;
    bool pending = false;
    int runs = 0;              // timed launches per round; events ev[0 .. 2 * runs]
    hipEvent_t ev[24] = {};
};

// Survival profile of a scene's path-tracing chunks: the fraction of a chunk's paths still traced at each
// bounce, copied back from the device without a host wait (the copy is enqueued after a chunk and read by
// a later chunk of any film of the scene once its event has completed). Chunks pick the bounce from which
// one launch finishes the remaining paths by it.
struct PathProfile {
    std::mutex mu;
    std::vector<double> frac;    // frac[b]: paths at bounce b / paths at bounce 0; negative: unknown
    uint32_t* host = nullptr;    // pinned copy target
    int host_len = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
    int pending_upto = -1;       // bounces 0 .. pending_upto of the pending copy were recorded
};

// Device buffers and streams that a serving loop reuses across its films. Admitting a request takes
// them from here and releasing it gives them back, so that once the pool holds buffers of the sizes the
// load needs, neither calls hipMalloc or hipFree nor creates or destroys a stream: those calls wait on
// the device and cost 0.4-6 ms per film, time in which the scheduler could not dispatch chunks.
// Buffers are handed out uninitialised (a film initialises what it reads). Thread-safe.
struct FilmPool {
    std::mutex mu;
    int device = 0;
    std::multimap<size_t, void*> bufs;  // idle buffers by size
    size_t cached = 0;                  // bytes of the idle buffers
    std::vector<hipStream_t> streams;   // idle streams (non-blocking, default priority)
    // The smallest idle buffer of bytes .. 1.25 * bytes, else a new one (after freeing idle buffers if the
    // device is out of memory); *got is its size.
    int take(size_t bytes, void** p, size_t* got);
    void give(void* p, size_t bytes);
    // Frees idle buffers, largest first, until at most `keep` bytes stay idle; returns the bytes freed.
    size_t trim(size_t keep);
    int take_stream(hipStream_t* st);
    void give_stream(hipStream_t st);
    // Frees every idle buffer and stream (their work must have finished).
    void clear();
    size_t idle_bytes();
};

}  // namespace prt

struct prt_scene {
    int device = 0;
    std::atomic<int> users{0};  // films and queued serving requests that read this scene
    prt::DevScene dev{};
    prt_scene_info info{};
    std::vector<prt::DevBuf> bufs;
    uint2* gstack = nullptr;     // traversal stack overflow for the intersection kernels
    uint32_t gstack_slots = 0;   // thread slots covered by gstack
    uint32_t* park = nullptr;    // park queue of the intersection kernels (count, then ray indices)
    int32_t park_steps = -1;     // their step budget before a ray is parked: -1 per batch size, 0 none
    int32_t max_stack = 0;
    int32_t num_cus = 0;
    int trace_variant[2] = {-1, -1};  // autotuned kernel variant (closest, any)
    int pt_trace_variant = 0;               // occupancy variant of the path tracer's trace kernel (PRT_PT_TRACE_VARIANT)
    std::atomic<int> pt_trace_gate{-1};     // autotuned leaf gate of the path tracer's trace kernel (eighths)
    std::atomic<int> bdpt_trace_gate{-1};   // autotuned leaf gate of the BDPT kernels' traversals (eighths)
    prt::TuneJob pt_tune, bdpt_tune;
    prt::PathProfile pt_profile;
    int forced_variant = -1;          // PRT_TRACE_VARIANT override
    int variant_blocks[16] = {0};
    std::vector<uint32_t> orig_id;  // host copy: leaf order -> original primitive id
    // float64 camera/scene facts needed by the host-side renderer
    int32_t image_width = 0, image_height = 0;
    double aspect = 1.0;
    prt_camera cam_desc{};
};

namespace prt {
// Allocates device memory owned by `s` and copies `bytes` from host (if src != nullptr).
int scene_alloc(prt_scene* s, size_t bytes, const void* src, void** out);
// Grid size used by persistent kernels (work-group count).
int persistent_blocks(const prt_scene* s, int block, int per_cu);
// NewCamera (camera.go:70-170) in float64, rounded to float32; force_height > 0 overrides
// int(width / aspect).
DevCamera make_camera(const prt_camera& c, int& height, int force_height);

// Film internals shared with the serving loop.
// With a pool, the film's buffers and stream come from it and go back to it when the film is destroyed.
int film_create_ex(const prt_scene* scene, const prt_render_params* p, int64_t capacity_cap, prt_film** out,
                   FilmPool* pool = nullptr);
size_t film_device_bytes(const prt_scene* scene, const prt_render_params* p, int64_t capacity_cap);
int film_enqueue(prt_film* f, int32_t spp);  // enqueue spp more samples per pixel, no wait
int film_plan(prt_film* f, int32_t spp);     // split spp more samples into chunks (not enqueued)
int film_plan_pass(prt_film* f, int32_t target);  // plan the pass to `target` spp (uniform or adaptive)
bool film_round_exhausted(const prt_film* f);     // every chunk planned so far has been enqueued
bool film_plan_done(const prt_film* f);      // the pass needs no more chunks
int film_advance(prt_film* f);               // adaptive films, when idle: plan the next round or end the pass
int film_enqueue_chunk(prt_film* f);         // enqueue the next planned chunk, no wait
double film_chunk_ms(prt_film* f);           // accumulate the finished chunk's device time
bool film_idle(prt_film* f);                 // all enqueued work finished
int film_track_chunks(prt_film* f);          // bracket every later chunk with ring events
int film_chunks_pending(prt_film* f);        // tracked chunks not finished yet (adds finished ones' time)
double film_plan_ms(const prt_film* f);      // device time of the plan's finished tracked chunks
int64_t film_capacity_paths(const prt_film* f);
void film_set_chunk_paths(prt_film* f, int64_t n);  // paths per chunk of later plans (at most capacity)
int64_t bdpt_grid_paths(const prt_scene* scene);    // BDPT paths of one full grid (resident blocks twice)
int film_wait(prt_film* f, float* ms);       // wait for enqueued work; device time of the last enqueue
int32_t film_spp_done(const prt_film* f);
size_t film_snapshot_bytes(const prt_film* f);
int film_snapshot(prt_film* f, void* dst, hipEvent_t done);  // enqueue copies of the accumulators, record done
int film_read_snapshot(const prt_film* f, const void* src, hipEvent_t done, hipStream_t stream, float* mean_rgb,
                       uint32_t* counts);
}  // namespace prt
