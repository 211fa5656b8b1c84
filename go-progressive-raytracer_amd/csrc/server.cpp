// server.cpp — continuous-batching serving loop for progressive renders (include/prt_server.h).
//
// One scheduler thread owns admission and pass sequencing. Each admitted request has a film slot with
// its own HIP stream; the scheduler enqueues one pass at a time per request and polls the pass's
// completion event, so passes of different requests overlap on the device while each request's passes
// stay ordered (RenderProgressive's pass loop, pkg/renderer/progressive.go:286-368).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/prt_server.h"
#include "prt_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

enum class St { Queued, Running, Done, Cancelled, Failed };

struct Req {
    int64_t id = 0;
    const prt_scene* scene = nullptr;
    prt_render_params params{};
    prt_film* film = nullptr;
    St state = St::Queued;
    int32_t pass = 0;            // last enqueued pass
    int32_t completed_pass = 0;  // last pass reported complete
    int32_t target = 0;          // samples per pixel of the enqueued pass
    size_t bytes = 0;
    Clock::time_point t_submit;
    bool cancel = false;
    bool release = false;
    int in_flight = 0;       // chunks of this request on the device
    int readers = 0;
    // PRT_SUBMIT_SNAPSHOTS: copies of the film taken when passes completed, the newest kSnapSlots kept
    bool snapshots = false;
    struct Snap {
        int32_t pass = 0;  // 0: empty
        int readers = 0;
        void* buf = nullptr;
        size_t bytes = 0;  // of buf, as taken from the pool
        hipEvent_t done = nullptr;
    };
    static constexpr int kSnapSlots = 3;
    Snap snap[kSnapSlots];
};

// A pass is enqueued chunk by chunk, so the scheduler can give first passes the device between
// chunks of longer renders. A first pass (one sample per pixel) runs as one chunk over the whole
// image where the film's path slots allow it (path tracing: up to 2^21 paths; BDPT: the subpath
// stores' 2 GiB bound, at least one full grid of the BDPT kernels): each chunk ends in a latency-bound
// tail of long paths, paid once instead of per chunk. Later passes run in chunks large enough to
// amortise that tail; a BDPT sample costs several path-tracing samples, so its later chunks hold
// fewer paths. Few such chunks are in flight at once (kBackgroundChunks), so a new request's first
// pass still waits little behind them. Measured at 20 requests/s (Poisson, mixed scenes, sizes and
// integrators): later chunks of 2^19 / 2^17 paths with 8 in flight gave 284 M samples/s at
// time-to-first-pass p50 8.4 ms, p99 26.8 ms; 2^20 / 2^18 with 6 in flight 380 M at 8.1 ms / 23-26 ms;
// 2^21 / 2^18 with 4 in flight 396 M at 7.3 ms / 20.5 ms (with 8 in flight: 406 M at 11.8 / 35 ms).
constexpr int64_t kFirstChunkMax = (int64_t)1 << 21;
constexpr int64_t kLaterChunk = (int64_t)1 << 21;
constexpr int64_t kLaterChunkBdpt = (int64_t)1 << 18;
// Chunks of later passes in flight on the device at once (over all requests). A new request's first
// pass waits for at most these, whichever hardware queue its stream lands on.
constexpr int kBackgroundChunks = 4;
// Chunks of a first pass in flight: one running, one queued behind it.
constexpr int kFirstPassChunks = 2;

int64_t serve_capacity(const prt_scene* s, const prt_render_params* p) {
    const int64_t w = p->width > 0 ? p->width : s->image_width;
    const int64_t h = p->height > 0 ? p->height : s->image_height;
    if (p->integrator == PRT_INTEGRATOR_BDPT) return std::max(w * h, prt::bdpt_grid_paths(s));
    return std::max(kLaterChunk, std::min(w * h, kFirstChunkMax));
}

}  // namespace

struct prt_server {
    int device = 0;
    int64_t budget = 0;
    int max_active = 1;
    int64_t used = 0;
    int64_t completed = 0;
    int64_t next_id = 1;
    std::mutex mu;
    std::condition_variable cv_work, cv_events;
    std::deque<Req*> queue;
    std::map<int64_t, Req*> reqs;
    std::vector<Req*> active;
    std::deque<prt_server_event> events;
    std::thread worker;
    bool stop = false;
    size_t rr_next = 0;  // round-robin position of background dispatch
    hipStream_t read_stream = nullptr;  // snapshot reads (non-blocking, default priority)
    // films' and snapshots' device buffers and streams, reused across requests; its idle buffers count
    // against the HBM budget together with the active films
    prt::FilmPool pool;

    double since(const Req* r) const {
        return std::chrono::duration<double, std::milli>(Clock::now() - r->t_submit).count();
    }
    void emit(const Req* r, int type, int32_t pass, int32_t target, int err, double dev_ms) {
        prt_server_event e{};
        e.request_id = r->id;
        e.type = type;
        e.pass_number = pass;
        e.target_samples = target;
        e.error_code = err;
        e.t_ms = since(r);
        e.device_ms = dev_ms;
        events.push_back(e);
        cv_events.notify_all();
    }
    void free_snapshots(Req* r) {
        for (auto& sn : r->snap) {
            if (sn.done) (void)hipEventSynchronize(sn.done);
            if (sn.buf) pool.give(sn.buf, sn.bytes);
            if (sn.done) (void)hipEventDestroy(sn.done);
            sn = Req::Snap{};
        }
    }
    int alloc_snapshots(Req* r) {
        const size_t b = std::max<size_t>(prt::film_snapshot_bytes(r->film), 16);
        for (auto& sn : r->snap) {
            int rc = pool.take(b, &sn.buf, &sn.bytes);
            if (rc != PRT_OK) return rc;
            if (hipEventCreateWithFlags(&sn.done, hipEventDisableTiming) != hipSuccess) return PRT_ERR_DEVICE;
        }
        return PRT_OK;
    }
    // Copy of the film as the finished pass left it, enqueued before the next pass's chunks. The slot
    // is the oldest one nobody reads; if all are being read the pass is not kept.
    int take_snapshot(Req* r) {
        Req::Snap* best = nullptr;
        for (auto& sn : r->snap)
            if (sn.readers == 0 && (!best || sn.pass < best->pass)) best = &sn;
        if (!best) return PRT_OK;
        best->pass = 0;
        int rc = prt::film_snapshot(r->film, best->buf, best->done);
        if (rc == PRT_OK) best->pass = r->pass;
        return rc;
    }
    void free_film(Req* r) {
        free_snapshots(r);
        if (r->film) {
            prt_film_destroy(r->film);
            r->film = nullptr;
            used -= (int64_t)r->bytes;
        }
    }
    // plan the next pass of an active request; its chunks are enqueued by the scheduler.
    // false on error (event emitted)
    bool next_pass(Req* r) {
        int32_t target = 0;
        int rc = prt_samples_for_pass(&r->params, r->pass + 1, &target);
        if (rc == PRT_OK) {
            r->pass++;
            r->target = target;
            const bool bdpt = r->params.integrator == PRT_INTEGRATOR_BDPT;
            prt::film_set_chunk_paths(r->film, r->pass == 1 ? prt::film_capacity_paths(r->film)
                                                            : (bdpt ? kLaterChunkBdpt : kLaterChunk));
            rc = prt::film_plan_pass(r->film, target);
        }
        if (rc != PRT_OK) {
            r->state = St::Failed;
            emit(r, PRT_EVENT_ERROR, r->pass, 0, rc, 0);
            return false;
        }
        return true;
    }
    void run() {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        while (!stop) {
            bool progressed = false;
            // ---- cancelled while queued: leave the queue now, whatever the slots and budget ----
            for (auto it = queue.begin(); it != queue.end();) {
                Req* r = *it;
                if (!r->cancel) {
                    ++it;
                    continue;
                }
                it = queue.erase(it);
                r->state = St::Cancelled;
                emit(r, PRT_EVENT_CANCELLED, 0, 0, 0, 0);
                progressed = true;
            }
            // ---- deferred releases (first: their buffers and budget go to the requests admitted next) ----
            for (auto it = reqs.begin(); it != reqs.end();) {
                Req* r = it->second;
                bool finished = r->state == St::Done || r->state == St::Cancelled || r->state == St::Failed;
                if (r->release && finished && r->readers == 0) {
                    // a request that failed at dispatch is still listed as active until the next sweep
                    active.erase(std::remove(active.begin(), active.end(), r), active.end());
                    if (r->film) (void)prt::film_wait(r->film, nullptr);
                    free_film(r);
                    const_cast<prt_scene*>(r->scene)->users--;
                    delete r;
                    it = reqs.erase(it);
                    progressed = true;
                } else {
                    ++it;
                }
            }
            // ---- admission (FIFO, HBM budget, slot count) ----
            while (!queue.empty() && (int)active.size() < max_active) {
                Req* r = queue.front();
                if ((int64_t)r->bytes > budget) {
                    queue.pop_front();
                    r->state = St::Failed;
                    emit(r, PRT_EVENT_ERROR, 0, 0, PRT_ERR_OUT_OF_MEMORY, 0);
                    continue;
                }
                if (used + (int64_t)r->bytes > budget) break;  // wait for memory to free up
                // idle pooled buffers give way to the new film when the budget would not hold both
                if (used + (int64_t)r->bytes + (int64_t)pool.idle_bytes() > budget)
                    pool.trim((size_t)(budget - used - (int64_t)r->bytes));
                queue.pop_front();
                int rc = prt::film_create_ex(r->scene, &r->params, serve_capacity(r->scene, &r->params), &r->film,
                                             &pool);
                if (rc == PRT_OK) {
                    used += (int64_t)r->bytes;  // free_film returns it with the film
                    rc = prt::film_track_chunks(r->film);
                    if (rc == PRT_OK && r->snapshots) rc = alloc_snapshots(r);
                }
                if (rc != PRT_OK) {
                    free_film(r);
                    r->state = St::Failed;
                    emit(r, PRT_EVENT_ERROR, 0, 0, rc, 0);
                    continue;
                }
                r->state = St::Running;
                active.push_back(r);
                if (next_pass(r)) emit(r, PRT_EVENT_ADMITTED, r->pass, r->target, 0, 0);
                progressed = true;
            }
            // ---- chunk and pass completion ----
            for (size_t k = 0; k < active.size();) {
                Req* r = active[k];
                bool remove = false;
                const int was_in_flight = r->in_flight;
                if (r->state != St::Running) {
                    remove = true;
                } else if (r->cancel) {
                    // no more chunks are dispatched; the request ends once its chunks on the device finish
                    if (r->in_flight > 0) r->in_flight = prt::film_chunks_pending(r->film);
                    if (r->in_flight == 0) {
                        r->state = St::Cancelled;
                        emit(r, PRT_EVENT_CANCELLED, r->completed_pass, r->target, 0, 0);
                        remove = true;
                    }
                } else if (r->in_flight == 0 || (r->in_flight = prt::film_chunks_pending(r->film)) == 0) {
                    const double ms = prt::film_plan_ms(r->film);
                    // adaptive passes: the finished round decides the next one (film idle, quick read)
                    if (prt::film_round_exhausted(r->film) && !prt::film_plan_done(r->film)) {
                        if (prt::film_advance(r->film) != PRT_OK) {
                            r->state = St::Failed;
                            emit(r, PRT_EVENT_ERROR, r->pass, r->target, PRT_ERR_DEVICE, 0);
                            active.erase(active.begin() + k);
                            continue;
                        }
                        progressed = true;
                    }
                    if (prt::film_plan_done(r->film) && r->pass > r->completed_pass) {
                        r->completed_pass = r->pass;
                        if (r->snapshots && take_snapshot(r) != PRT_OK) {
                            r->state = St::Failed;
                            emit(r, PRT_EVENT_ERROR, r->pass, r->target, PRT_ERR_DEVICE, 0);
                            active.erase(active.begin() + k);
                            continue;
                        }
                        emit(r, PRT_EVENT_PASS_COMPLETE, r->pass, r->target, 0, ms);
                        progressed = true;
                        if (r->pass >= r->params.max_passes ||
                                   prt::film_spp_done(r->film) >= r->params.max_samples_per_pixel) {
                            r->state = St::Done;
                            completed++;
                            emit(r, PRT_EVENT_DONE, r->pass, r->target, 0, 0);
                            remove = true;
                        } else if (!next_pass(r)) {
                            remove = true;
                        }
                    }
                }
                if (r->in_flight != was_in_flight) progressed = true;
                if (remove) {
                    active.erase(active.begin() + k);
                } else {
                    k++;
                }
            }
            // ---- chunk dispatch ----
            // Requests that still owe their first pass get the device first: each keeps
            // kFirstPassChunks chunks queued, and while any is owed the other requests receive no new chunks (theirs in
            // flight drain). Otherwise requests take turns (round robin) to keep up to two chunks each
            // queued, within kBackgroundChunks over all requests, so the device never waits for the
            // scheduler and a newly admitted request never waits behind a long backlog.
            bool first_owed = false;
            int background = 0;
            for (Req* r : active) {
                if (r->state != St::Running) continue;
                if (r->completed_pass == 0 && !r->cancel) first_owed = true;
                else background += r->in_flight;
            }
            auto enqueue = [&](Req* r) {
                if (prt::film_enqueue_chunk(r->film) != PRT_OK) {
                    r->state = St::Failed;
                    emit(r, PRT_EVENT_ERROR, r->pass, r->target, PRT_ERR_DEVICE, 0);
                    return false;
                }
                r->in_flight++;
                progressed = true;
                return true;
            };
            for (Req* r : active) {
                if (r->state != St::Running || r->cancel || r->completed_pass != 0) continue;
                while (r->in_flight < kFirstPassChunks && !prt::film_round_exhausted(r->film) && enqueue(r)) {
                }
            }
            if (!first_owed && !active.empty()) {
                const size_t n = active.size();
                const size_t start = rr_next % n;
                for (int round = 0; round < 2 && background < kBackgroundChunks; round++) {
                    for (size_t j = 0; j < n && background < kBackgroundChunks; j++) {
                        const size_t idx = (start + j) % n;
                        Req* r = active[idx];
                        if (r->state != St::Running || r->cancel || r->in_flight > round ||
                            prt::film_round_exhausted(r->film))
                            continue;
                        if (!enqueue(r)) continue;
                        background++;
                        rr_next = idx + 1;
                    }
                }
            }
            if (!progressed) cv_work.wait_for(lk, std::chrono::microseconds(100));
        }
    }
};

using namespace prt;

extern "C" {

int prt_server_create(int32_t device, int64_t budget, int32_t max_active, prt_server** out) {
    if (!out || budget <= 0 || max_active < 1) return fail(PRT_ERR_INVALID_ARGUMENT, "bad server arguments");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
        return fail(PRT_ERR_NO_DEVICE, "no such HIP device");
    prt_server* s = new prt_server();
    s->device = device;
    s->pool.device = device;
    s->budget = budget;
    s->max_active = max_active;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&s->read_stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return fail(PRT_ERR_DEVICE, "cannot create the server's read stream");
    }
    s->worker = std::thread([s]() { s->run(); });
    *out = s;
    return PRT_OK;
}

int prt_server_submit(prt_server* s, const prt_scene* scene, const prt_render_params* p, int64_t* id) {
    return prt_server_submit_ex(s, scene, p, 0u, id);
}

int prt_server_submit_ex(prt_server* s, const prt_scene* scene, const prt_render_params* p, uint32_t flags,
                         int64_t* id) {
    if (!s || !scene || !p || !id) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL argument");
    if (flags & ~PRT_SUBMIT_SNAPSHOTS) return fail(PRT_ERR_INVALID_ARGUMENT, "unknown submit flags");
    if (scene->device != s->device) return fail(PRT_ERR_INVALID_ARGUMENT, "scene lives on another device");
    int32_t t = 0;
    int rc = prt_samples_for_pass(p, 1, &t);
    if (rc != PRT_OK) return rc;
    if (p->integrator != PRT_INTEGRATOR_PATH && p->integrator != PRT_INTEGRATOR_BDPT)
        return fail(PRT_ERR_INVALID_ARGUMENT, "unknown integrator");
    Req* r = new Req();
    r->scene = scene;
    const_cast<prt_scene*>(scene)->users++;  // until the request is released
    r->params = *p;
    r->bytes = film_device_bytes(scene, p, serve_capacity(scene, p));
    if (flags & PRT_SUBMIT_SNAPSHOTS) {  // sums, counts and (BDPT) splats per pixel, per slot
        const int64_t w = p->width > 0 ? p->width : scene->image_width;
        const int64_t h = p->height > 0 ? p->height : scene->image_height;
        const size_t px = 16 + 4 + (p->integrator == PRT_INTEGRATOR_BDPT ? 24 : 0);
        r->snapshots = true;
        r->bytes += (size_t)Req::kSnapSlots * px * (size_t)(w * h);
    }
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        s->reqs[r->id] = r;
        s->queue.push_back(r);
    }
    s->cv_work.notify_all();
    *id = r->id;
    return PRT_OK;
}

int prt_server_poll(prt_server* s, prt_server_event* ev, int32_t max_events, int32_t timeout_ms, int32_t* count) {
    if (!s || !count || (max_events > 0 && !ev)) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->events.empty() && timeout_ms > 0)
        s->cv_events.wait_for(lk, std::chrono::milliseconds(timeout_ms), [s]() { return !s->events.empty(); });
    int32_t n = 0;
    while (n < max_events && !s->events.empty()) {
        ev[n++] = s->events.front();
        s->events.pop_front();
    }
    *count = n;
    return PRT_OK;
}

int prt_server_cancel(prt_server* s, int64_t id) {
    if (!s) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(PRT_ERR_INVALID_ARGUMENT, "unknown request");
    it->second->cancel = true;
    s->cv_work.notify_all();
    return PRT_OK;
}

static int with_film(prt_server* s, int64_t id, const std::function<int(prt_film*)>& fn) {
    Req* r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) return fail(PRT_ERR_INVALID_ARGUMENT, "unknown request");
        r = it->second;
        if (!r->film) return fail(PRT_ERR_INVALID_ARGUMENT, "request has no film (queued, failed or released)");
        r->readers++;
    }
    int rc = fn(r->film);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->readers--;
    }
    s->cv_work.notify_all();
    return rc;
}

int prt_server_read(prt_server* s, int64_t id, float* mean_rgb, uint32_t* counts) {
    if (!s) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL server");
    return with_film(s, id, [&](prt_film* f) { return prt_film_read(f, mean_rgb, nullptr, nullptr, counts); });
}

int prt_server_read_pass(prt_server* s, int64_t id, int32_t pass, float* mean_rgb, uint32_t* counts) {
    if (!s) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL server");
    Req* r;
    Req::Snap* sn = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) return fail(PRT_ERR_INVALID_ARGUMENT, "unknown request");
        r = it->second;
        if (!r->snapshots) return fail(PRT_ERR_INVALID_ARGUMENT, "request was not submitted with PRT_SUBMIT_SNAPSHOTS");
        if (!r->film) return fail(PRT_ERR_INVALID_ARGUMENT, "request has no film (queued, failed or released)");
        for (auto& x : r->snap)
            if (pass > 0 && x.pass == pass) sn = &x;
        if (!sn) return fail(PRT_ERR_INVALID_ARGUMENT, "no snapshot of that pass (not complete or no longer kept)");
        sn->readers++;
        r->readers++;
    }
    int rc = prt::film_read_snapshot(r->film, sn->buf, sn->done, s->read_stream, mean_rgb, counts);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        sn->readers--;
        r->readers--;
    }
    s->cv_work.notify_all();
    return rc;
}

int prt_server_read_rgba8(prt_server* s, int64_t id, int32_t x0, int32_t y0, int32_t x1, int32_t y1, uint8_t* rgba) {
    if (!s) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL server");
    return with_film(s, id, [&](prt_film* f) { return prt_film_read_rgba8(f, x0, y0, x1, y1, rgba); });
}

int prt_server_release(prt_server* s, int64_t id) {
    if (!s) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(PRT_ERR_INVALID_ARGUMENT, "unknown request");
    it->second->release = true;
    if (it->second->state == St::Queued || it->second->state == St::Running) it->second->cancel = true;
    s->cv_work.notify_all();
    return PRT_OK;
}

int prt_server_stats_get(prt_server* s, prt_server_stats* st) {
    if (!s || !st) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    st->queued = (int32_t)s->queue.size();
    st->active = (int32_t)s->active.size();
    st->completed = s->completed;
    st->budget_bytes = s->budget;
    st->used_bytes = s->used;
    st->idle_bytes = (int64_t)s->pool.idle_bytes();
    return PRT_OK;
}

int prt_server_destroy(prt_server* s) {
    if (!s) return PRT_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (auto& kv : s->reqs) {
        Req* r = kv.second;
        if (r->film) (void)prt::film_wait(r->film, nullptr);
        s->free_film(r);
        const_cast<prt_scene*>(r->scene)->users--;
        delete r;
    }
    s->pool.clear();
    if (s->read_stream) (void)hipStreamDestroy(s->read_stream);
    delete s;
    return PRT_OK;
}

}  // extern "C"
