// pt.hip — wavefront path tracer and device-resident progressive film.
//
// Estimator: PathTracingIntegrator.rayColorRecursive (pkg/integrator/path_tracing.go:35-234)
// unrolled into bounce stages over compacted path queues:
//   camera -> [extend (closest hit) -> shade (emission, BSDF sample, NEE, MIS, RR) -> shadow (any hit)]*
// The recursion's return value is linear in the child radiance, so each level's factor
// (Russian-roulette compensation, specular attenuation or att*cos*w/pdf including the MIS
// weight applied to the whole recursive radiance, :172-184) is folded into a path weight
// `beta`; the Russian-roulette throughput `thr` is tracked separately exactly as the
// recursion passes it (:178, :91). Film accumulation follows PixelStats (pkg/renderer/stats.go)
// with Kahan-compensated float32 sums; the per-pixel order of additions is fixed, so a
// render is bit-reproducible independent of scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "prt_internal.h"
#include "bdpt_device.h"
#include "shading.h"
#include "traverse.h"

namespace prt {

constexpr int kBlock = 256;
constexpr int kLds = 6;
constexpr int kNodeCache = 64;  // top BVH nodes kept in LDS by the trace kernels

// The Russian-roulette throughput rides in the free fourth lanes of o, d and beta; a shadow ray starts
// at the path's current origin o (written for paths that continue or cast a shadow ray).
struct PathBuffers {
    float4* o;      // origin, throughput x
    float4* d;      // direction (unnormalised as the reference passes scatter.Scattered), throughput y
    float4* beta;   // path weight, throughput z
    float4* L;      // radiance gathered so far
    float4* hit;    // t, prim (bits), u, v
    float4* sd;     // shadow ray direction, tmax (tmin 0.001)
    float4* sc;     // shadow contribution
    uint32_t* key;
    uint32_t* queue[2];
    uint32_t* squeue[2];  // shadow rays of even / odd bounces
    uint32_t* counters;   // queue lengths, one per 256-B line: path queues 0 and 1, shadow queues 2 and 3,
                          // long-ray queue 4
    uint32_t* lq;         // long rays parked by a trace launch (path slot; kLongAny set for shadow rays)
    uint32_t lcap;        // capacity of lq
    uint32_t* bounce_paths;  // paths traced at each bounce of the last chunk (written by k_trace)
};

// Long rays: a trace launch parks a ray that is still traversing after its step budget in the long-ray
// queue (up to its capacity; a ray the full queue refuses keeps traversing), and k_trace_long then
// finishes each parked ray with a whole wave (coop_trace).
constexpr uint32_t kLongAny = 0x80000000u;
constexpr uint32_t kLongQueue = 4;
constexpr uint32_t kNumCounters = 5;

struct PassParams {
    int32_t width, height;
    int32_t px0, npix;      // pixel range of this chunk (linear pixel ids; active-list range if act)
    int32_t spp;            // samples per pixel in this chunk
    int32_t sample_base;    // first sample index (uniform films)
    int32_t max_depth, rr_min;
    uint64_t seed;
    int32_t trace_gate;     // leaf gate of the traversals inside the BDPT kernels (eighths of a wave, 0 = none;
                            // kGateFlat: flat traversal)
    int32_t tune_stride;    // the subpath kernel generates every tune_stride-th slot only (autotuning runs), else 1
    // adaptive rounds: pixels come from an active list and each pixel's next sample index is its count
    const uint32_t* act;
    const uint32_t* count;
};

// Pixel and sample index of path slot i of a chunk (slot = sample-major over the chunk's pixels).
__device__ __forceinline__ void slot_sample(const PassParams& pp, uint32_t i, uint32_t& pix, uint32_t& sample) {
    const uint32_t s_local = i / (uint32_t)pp.npix, a = i % (uint32_t)pp.npix;
    if (pp.act) {
        pix = pp.act[pp.px0 + a];
        sample = pp.count[pix] + s_local;
    } else {
        pix = (uint32_t)pp.px0 + a;
        sample = (uint32_t)pp.sample_base + s_local;
    }
}

// Queue length q lives at counters[q * kCounterStride] (separate L2 lines, so the queues' atomics
// do not serialize on one line).
constexpr uint32_t kCounterStride = 64;
__device__ __forceinline__ uint32_t* qlen(const PathBuffers& pb, uint32_t q) { return pb.counters + q * kCounterStride; }

__device__ __forceinline__ bool park_ray(const PathBuffers& pb, uint32_t v) {
    const uint32_t k = atomicAdd(qlen(pb, kLongQueue), 1u);
    if (k >= pb.lcap) return false;
    pb.lq[k] = v;
    return true;
}

// Block-aggregated appends to two queues: the block's waves reserve their ranges with LDS atomics and
// one thread per block reserves the block's range of each queue with one global atomic. Every thread
// of the block must call it (it synchronizes the block); s_lds holds 4 words.
__device__ __forceinline__ void block_push2(uint32_t* s_lds, uint32_t* q0, uint32_t* c0, bool a0, uint32_t* q1,
                                            uint32_t* c1, bool a1, uint32_t value) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t lt = (1ull << lane) - 1ull;
    if (threadIdx.x < 2) s_lds[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t m0 = __ballot(a0), m1 = __ballot(a1);
    uint32_t w0 = 0, w1 = 0;
    if (lane == 0) {
        if (m0) w0 = atomicAdd(&s_lds[0], (uint32_t)__popcll(m0));
        if (m1) w1 = atomicAdd(&s_lds[1], (uint32_t)__popcll(m1));
    }
    w0 = __shfl(w0, 0, 64);
    w1 = __shfl(w1, 0, 64);
    __syncthreads();
    if (threadIdx.x == 0) {
        s_lds[2] = s_lds[0] ? atomicAdd(c0, s_lds[0]) : 0u;
        s_lds[3] = s_lds[1] ? atomicAdd(c1, s_lds[1]) : 0u;
    }
    __syncthreads();
    if (a0) q0[s_lds[2] + w0 + (uint32_t)__popcll(m0 & lt)] = value;
    if (a1) q1[s_lds[3] + w1 + (uint32_t)__popcll(m1 & lt)] = value;
}

// wave-aggregated append to a queue
__device__ __forceinline__ void queue_push(uint32_t* q, uint32_t* counter, uint32_t value, bool active) {
    uint64_t mask = __ballot(active);
    if (mask == 0) return;
    uint32_t lane = threadIdx.x & 63u;
    uint32_t leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
    if (active) {
        uint32_t rank = __popcll(mask & ((1ull << lane) - 1ull));
        q[base + rank] = value;
    }
}

// Russian roulette at the start of bounce level `b` (ApplyRussianRoulette, :193-220).
// Returns false if the path terminates; multiplies beta by the compensation otherwise.
__device__ __forceinline__ bool russian_roulette(const PassParams& pp, uint32_t key, int b, V3 thr, V3& beta) {
    float u = draw(key, (uint32_t)(b + 1) * 16u);
    if (b < pp.rr_min) return true;
    float q = fminf(0.95f, fmaxf(0.5f, lum(thr)));
    if (u > q) return false;
    beta = beta * (1.0f / q);
    return true;
}

__global__ __launch_bounds__(kBlock) void k_camera(DevScene sc, PassParams pp, PathBuffers pb) {
    const uint32_t n = (uint32_t)pp.npix * (uint32_t)pp.spp;
    uint32_t slot = blockIdx.x * kBlock + threadIdx.x;
    bool active = false;
    if (slot < n) {
        uint32_t pix, sample;
        slot_sample(pp, slot, pix, sample);
        int i = (int)(pix % (uint32_t)pp.width), j = (int)(pix / (uint32_t)pp.width);
        uint32_t key = path_key(pp.seed, pix, sample);
        // GetRay(i, j, Get2D(), Get2D()) (tile_renderer.go:56)
        float l0 = draw(key, 0), l1 = draw(key, 1), j0 = draw(key, 2), j1 = draw(key, 3);
        V3 o, d;
        camera_ray(sc.cam, i, j, l0, l1, j0, j1, o, d);
        V3 beta = v3(1, 1, 1);
        if (pp.max_depth > 0) active = russian_roulette(pp, key, 0, v3(1, 1, 1), beta);
        pb.o[slot] = make_float4(o.x, o.y, o.z, 1.f);
        pb.d[slot] = make_float4(d.x, d.y, d.z, 1.f);
        pb.beta[slot] = make_float4(beta.x, beta.y, beta.z, 1.f);
        pb.L[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
        pb.key[slot] = key;
    }
    if (pp.rr_min > 0 && pp.max_depth > 0) {
        // Russian roulette cannot end a path at level 0: every slot is active, the queue is the identity
        if (slot < n) pb.queue[0][slot] = slot;
        if (slot == 0) *qlen(pb, 0) = n;
    } else {
        queue_push(pb.queue[0], qlen(pb, 0), slot, active);
    }
}

// One trace launch per bounce: the shadow rays the previous bounce cast (any hit, shadow queue sq, or
// none when sq < 0) and, when `closest`, this bounce's extension rays (closest hit, path queue bounce & 1).
// Neither depends on the other; both read the path origin. A wave takes 64-ray batches of one kind
// (the extension rays' range is padded to whole batches), so its traversal mode is uniform.
// Traversal statistics of the counting trace kernel (prt_film_probe): extension rays, their node
// visits, box tests and primitive tests, then the same four for shadow rays.
constexpr int kProbeCounts = 8;

// Occupancy variants of the trace kernel: the kernel compiled for at least 7 (the default) or 8 waves per
// SIMD, and the default with its resident work-groups per CU capped at 6 or 5 by unused dynamic LDS. The
// same traversal, so every variant gives the same hits. A scene uses the default unless
// PRT_PT_TRACE_VARIANT pins another (timing them on a sample of one bounce's rays picked variants that were
// slower over whole renders).
constexpr int kTraceVariants = 4;
// Leaf gates (eighths of a wave's running lanes, see leaf_gate) the autotuners choose from; 0 first.
// Scenes of at most kFlatMaxPrims primitives also time the flat traversal (kGateFlat).
constexpr int kGateChoices = 4;
constexpr int kGateChoice[kGateChoices] = {0, 4, 8, kGateFlat};
inline int gate_choices(const DevScene& sc) { return sc.num_prims <= kFlatMaxPrims ? kGateChoices : kGateChoices - 1; }
constexpr int kTraceKernelWaves[kTraceVariants] = {7, 8, 7, 7};
constexpr int kTraceGroupCap[kTraceVariants] = {0, 0, 6, 5};  // 0: as resources allow

template <bool MIXED, bool COUNT = false, int WAVES = 7>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_trace(DevScene sc, PathBuffers pb,
                                                                                         int bounce, int closest, int sq,
                                                  uint2* gstack, uint32_t gstride,
                                                  unsigned long long* probe = nullptr, int gate = 0,
                                                  uint32_t tune_stride = 1, uint32_t budget = 0,
                                                  uint32_t budget_small = 0, uint32_t small_rays_per_lane = 1) {
    __shared__ uint32_t s_node[kLds * kBlock];
    __shared__ float s_t[kLds * kBlock];
    __shared__ float4 s_cache[kNodeCache * 8];
    const uint32_t cur = (uint32_t)bounce & 1u;
    // autotuning runs (tune_stride > 1) trace every tune_stride-th extension ray of the queue only
    const uint32_t nc = closest ? (*qlen(pb, cur) + tune_stride - 1u) / tune_stride : 0u;
    const uint32_t ns = sq >= 0 ? *qlen(pb, 2u + (uint32_t)sq) : 0u;
    const uint32_t* qc = pb.queue[cur];
    const uint32_t* qs = pb.squeue[sq > 0 ? 1 : 0];
    if (closest && blockIdx.x == 0 && threadIdx.x == 0) {
        *qlen(pb, cur ^ 1u) = 0;  // next bounce's queue (its last reader finished earlier)
        *qlen(pb, 2u + cur) = 0;  // the shadow queue this bounce's shading fills (the last launch read it)
        if (tune_stride == 1) pb.bounce_paths[bounce] = *qlen(pb, cur);
    }
    const uint32_t nc64 = (nc + 63u) & ~63u;
    const uint32_t stride = gridDim.x * kBlock;
    // a launch with few rays per lane lasts as long as its longest rays (each wave's batches are fixed by
    // the grid stride): park early
    if (nc + ns <= small_rays_per_lane * stride) budget = budget_small;
    // work-groups without rays leave before loading the node cache (the deep bounces of a pass trace few
    // rays on the full persistent grid)
    if (blockIdx.x * kBlock >= nc64 + ns) return;
    TStack<kLds, kBlock> st;
    st.s_node = s_node;
    st.s_t = s_t;
    st.g = gstack;
    st.lane = threadIdx.x;
    st.gslot = blockIdx.x * kBlock + threadIdx.x;
    st.gstride = gstride;
    load_node_cache<kNodeCache, kBlock>(sc, s_cache);
    st.ncache = s_cache;
    TCounts ce{0, 0, 0}, cs{0, 0, 0};
    uint32_t re = 0, rs = 0;
    // uniform trip count per wave so every lane reaches the loop exit together
    for (uint32_t base = blockIdx.x * kBlock + (threadIdx.x & ~63u); base < nc64 + ns; base += stride) {
        const bool any = base >= nc64;  // wave-uniform
        const uint32_t i = (any ? base - nc64 : base) + (threadIdx.x & 63u);
        if (i >= (any ? ns : nc)) continue;
        if (COUNT) {
            if (any) rs++;
            else re++;
        }
        const uint32_t slot = any ? qs[i] : qc[i * tune_stride];
        const float4 o = pb.o[slot];
        const float4 d = any ? pb.sd[slot] : pb.d[slot];
        TRay r;
        r.ox = o.x;
        r.oy = o.y;
        r.oz = o.z;
        r.dx = d.x;
        r.dy = d.y;
        r.dz = d.z;
        r.tmin = 0.001f;
        THit h;
        // rays still traversing after `budget` steps are parked for k_trace_long (closest hit: with the
        // best hit so far in the hit record)
        if (any) {
            r.tmax = d.w;
            auto park = [&]() { return park_ray(pb, slot | kLongAny); };
            if (trace_budget<true, MIXED, COUNT, kLds, kBlock, kNodeCache>(sc, r, h, st, &cs, gate, budget, park) ==
                kTraceMiss) {
                float4 c = pb.sc[slot];
                float4 l = pb.L[slot];
                pb.L[slot] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, 0.f);
            }
        } else {
            r.tmax = __int_as_float(0x7f800000);
            auto park = [&]() { return park_ray(pb, slot); };
            trace_budget<false, MIXED, COUNT, kLds, kBlock, kNodeCache>(sc, r, h, st, &ce, gate, budget, park);
            pb.hit[slot] = make_float4(h.t, __uint_as_float(h.prim), h.u, h.v);
        }
    }
    if (COUNT) {  // wave sums, one 64-bit atomic per count and wave
        uint32_t v[kProbeCounts] = {re, ce.nodes, ce.boxes, ce.prims, rs, cs.nodes, cs.boxes, cs.prims};
        for (int k = 0; k < kProbeCounts; k++) {
            unsigned long long x = v[k];
            for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
            if ((threadIdx.x & 63u) == 0 && x) atomicAdd(&probe[k], x);
        }
    }
}

// The rays the last trace launch parked, 64 / G per one-wave work-group (coop_trace with groups of G
// lanes): closest hits complete their hit records, unoccluded shadow rays add their contribution. Their
// results equal the one-lane traversal's.
template <bool MIXED, int G>
__global__ __launch_bounds__(64) void k_trace_long(DevScene sc, PathBuffers pb, uint2* gstack, uint32_t gstride) {
    constexpr uint32_t R = 64 / G;  // rays per work-group and round
    __shared__ uint32_t s_node[kLds * 64];
    __shared__ float s_t[kLds * 64];
    __shared__ uint32_t f_node[kCoopEntries];
    __shared__ float f_t[kCoopEntries];
    const uint32_t n = min(*qlen(pb, kLongQueue), pb.lcap);
    if (blockIdx.x * R >= n) return;
    TStack<kLds, 64> st;
    st.s_node = s_node;
    st.s_t = s_t;
    st.g = gstack;
    st.lane = threadIdx.x;
    st.gslot = blockIdx.x * 64 + threadIdx.x;
    st.gstride = gstride;
    const float kInfF = __int_as_float(0x7f800000);
    for (uint32_t base = blockIdx.x * R; base < n; base += gridDim.x * R) {  // uniform over the work-group
        const uint32_t i = base + threadIdx.x / (uint32_t)G;
        const bool has = i < n;
        const uint32_t v = has ? pb.lq[i] : 0u;
        const bool any = (v & kLongAny) != 0;
        const uint32_t slot = v & ~kLongAny;
        float4 o = make_float4(0.f, 0.f, 0.f, 0.f), d = o;
        THit h;
        h.t = kInfF;
        h.prim = 0xFFFFFFFFu;
        h.u = h.v = 0.0f;
        if (has) {
            o = pb.o[slot];
            d = any ? pb.sd[slot] : pb.d[slot];
            if (!any) {
                const float4 hp = pb.hit[slot];
                h.t = hp.x;
                h.prim = __float_as_uint(hp.y);
                h.u = hp.z;
                h.v = hp.w;
            }
        }
        const TRay r{o.x, o.y, o.z, 0.001f, d.x, d.y, d.z, any ? d.w : kInfF};
        const bool found = coop_trace<MIXED, kLds, G>(sc, r, h, st, f_node, f_t, has, any);
        if (has && (threadIdx.x & (uint32_t)(G - 1)) == 0) {
            if (any) {
                if (!found) {
                    const float4 c = pb.sc[slot];
                    const float4 l = pb.L[slot];
                    pb.L[slot] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, 0.f);
                }
            } else {
                pb.hit[slot] = make_float4(h.t, __uint_as_float(h.prim), h.u, h.v);
            }
        }
        __syncthreads();  // the frontier buffers are reused by the next rays
    }
}

// One path's bounce b (PathTracingIntegrator.rayColorRecursive, path_tracing.go:35-189) from its hit
// record: adds the emission (or the infinite lights on a miss) to L, samples the BSDF and a light, and
// writes the continuation (cont: origin, direction, weights) and the shadow ray (shadow: direction, tmax,
// contribution) of path slot `slot`.
__device__ __forceinline__ void shade_path(const DevScene& sc, const PassParams& pp, const PathBuffers& pb,
                                           uint32_t slot, int b, bool& cont, bool& shadow) {
    cont = false;
    shadow = false;
    float4 o4 = pb.o[slot], d4 = pb.d[slot], h4 = pb.hit[slot];
    float4 b4 = pb.beta[slot];
    uint32_t key = pb.key[slot];
    V3 o = v3(o4.x, o4.y, o4.z), d = v3(d4.x, d4.y, d4.z);
    V3 beta = v3(b4.x, b4.y, b4.z), thr = v3(o4.w, d4.w, b4.w);
    // L is read and written only when this bounce adds to it (a zero term leaves it unchanged;
    // a non-finite one is added as the reference adds it)
    V3 add = v3(0, 0, 0);
    THit h;
    h.t = h4.x;
    h.prim = __float_as_uint(h4.y);
    h.u = h4.z;
    h.v = h4.w;
    const uint32_t pos = (uint32_t)(b + 1) * 16u;
    if (h.prim == 0xFFFFFFFFu) {
        // miss: sum of infinite-light emission (:49-60)
        V3 e = v3(0, 0, 0);
        for (int li = 0; li < sc.num_lights; li++)
            if (light_is_infinite(sc.lights[li].type)) e = e + light_emit(sc, sc.lights[li], d);
        add = mulv(beta, e);
    } else {
        Surf s = make_surf(sc, o, d, h);
        add = mulv(beta, emit_hit(sc, s.mat, d, s.front));  // getEmittedLight :100-105
        ScatterOut so;
        if (scatter(sc, s.mat, d, s, key, pos + 1, so)) {
            V3 nd = so.dir;
            if (so.pdf <= 0.0f) {  // specular (:88-97)
                beta = mulv(beta, so.att);
                thr = mulv(thr, so.att);
                cont = true;
            } else {
                // direct lighting (:116-157), light draws at positions 8..10
                if (sc.num_lights > 0) {
                    float prob;
                    int li = select_light(sc, draw(key, pos + 8), prob);
                    LSample ls = light_sample(sc, sc.lights[li], s.p, s.n, draw(key, pos + 9), draw(key, pos + 10));
                    ls.pdf *= prob;
                    float cosine = dot(ls.dir, s.n);
                    if (lum(ls.emission) > 0.0f && ls.pdf > 0.0f && cosine > 0.0f) {
                        bool delta;
                        float mp = mat_pdf(sc, s.mat, d, ls.dir, s.n, delta);
                        if (!delta) {
                            float w = power_heuristic(ls.pdf, mp);
                            V3 f = brdf(sc, s.mat, d, ls.dir, s, 0);
                            V3 c = mulv(beta, mulv(f, ls.emission) * (cosine * w / ls.pdf));
                            if (c.x != 0.0f || c.y != 0.0f || c.z != 0.0f) {
                                float tmax = ls.dist - 0.001f;
                                pb.sd[slot] = make_float4(ls.dir.x, ls.dir.y, ls.dir.z, tmax);
                                pb.sc[slot] = make_float4(c.x, c.y, c.z, 0.f);
                                shadow = true;
                            }
                        }
                    }
                }
                // indirect lighting (:160-189)
                V3 sd = normalize(nd);
                float cosine = dot(sd, s.n);
                if (cosine > 0.0f) {
                    float lp = total_light_pdf(sc, s.p, s.n, sd);
                    float w = power_heuristic(so.pdf, lp);
                    thr = mulv(thr, so.att) * (cosine / so.pdf);
                    beta = mulv(beta, so.att) * (cosine * w / so.pdf);
                    cont = true;
                }
            }
            if (cont) d = nd;
            if (cont || shadow) o = s.p;
        }
    }
    // next level: depth limit (:37-39) and Russian roulette (:42-45)
    if (cont) {
        if (b + 1 >= pp.max_depth) cont = false;
        else cont = russian_roulette(pp, key, b + 1, thr, beta);
    }
    if (add.x != 0.0f || add.y != 0.0f || add.z != 0.0f) {
        const float4 l4 = pb.L[slot];
        pb.L[slot] = make_float4(l4.x + add.x, l4.y + add.y, l4.z + add.z, 0.f);
    }
    if (cont || shadow) pb.o[slot] = make_float4(o.x, o.y, o.z, thr.x);
    if (cont) {
        pb.d[slot] = make_float4(d.x, d.y, d.z, thr.y);
        pb.beta[slot] = make_float4(beta.x, beta.y, beta.z, thr.z);
    }
}

__global__ __launch_bounds__(kBlock) void k_shade(DevScene sc, PassParams pp, PathBuffers pb, int b) {
    __shared__ uint32_t s_push[4];
    // the long-ray queue of the next trace launch starts empty (this bounce's k_trace_long has finished)
    if (blockIdx.x == 0 && threadIdx.x == 0) *qlen(pb, kLongQueue) = 0;
    const uint32_t cur = (uint32_t)b & 1u;
    const uint32_t count = *qlen(pb, cur);
    const uint32_t* q = pb.queue[cur];
    const uint32_t stride = gridDim.x * kBlock;
    // block-uniform trip count (the queue pushes synchronize the block)
    for (uint32_t base = blockIdx.x * kBlock; base < count; base += stride) {
        uint32_t i = base + threadIdx.x;
        bool cont = false, shadow = false;
        uint32_t slot = 0;
        if (i < count) {
            slot = q[i];
            shade_path(sc, pp, pb, slot, b, cont, shadow);
        }
        block_push2(s_push, pb.queue[cur ^ 1u], qlen(pb, cur ^ 1u), cont, pb.squeue[cur], qlen(pb, 2u + cur), shadow,
                    slot);
    }
}

// The last bounces of a chunk once few paths are left: a lane takes a path of bounce b0 (its hit traced)
// and runs it to the end, shading, tracing its shadow ray and its next extension ray bounce after bounce,
// in one launch instead of two per bounce. A path's radiance terms are added in the wavefront order
// (emission of bounce b, shadow ray of bounce b, emission of bounce b + 1, ...), so films are identical.
template <bool MIXED>
__global__ __launch_bounds__(kBlock) void k_tail(DevScene sc, PassParams pp, PathBuffers pb, int b0, uint2* gstack,
                                                 uint32_t gstride, int gate) {
    __shared__ uint32_t s_node[kLds * kBlock];
    __shared__ float s_t[kLds * kBlock];
    __shared__ float4 s_cache[kNodeCache * 8];
    const uint32_t cur = (uint32_t)b0 & 1u;
    const uint32_t count = *qlen(pb, cur);
    if (blockIdx.x * kBlock >= count) return;
    TStack<kLds, kBlock> st;
    st.s_node = s_node;
    st.s_t = s_t;
    st.g = gstack;
    st.lane = threadIdx.x;
    st.gslot = blockIdx.x * kBlock + threadIdx.x;
    st.gstride = gstride;
    load_node_cache<kNodeCache, kBlock>(sc, s_cache);
    st.ncache = s_cache;
    const uint32_t* q = pb.queue[cur];
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < count; i += gridDim.x * kBlock) {
        const uint32_t slot = q[i];
        for (int b = b0;; b++) {
            bool cont, shadow;
            shade_path(sc, pp, pb, slot, b, cont, shadow);
            if (!cont && !shadow) break;
            const float4 o = pb.o[slot];
            THit h;
            if (shadow) {
                const float4 sd = pb.sd[slot];
                const TRay r{o.x, o.y, o.z, 0.001f, sd.x, sd.y, sd.z, sd.w};
                if (!trace<true, MIXED, false, kLds, kBlock, kNodeCache>(sc, r, h, st, nullptr, flat_gate(sc, gate) ? gate : 0)) {
                    const float4 c = pb.sc[slot];
                    const float4 l = pb.L[slot];
                    pb.L[slot] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, 0.f);
                }
            }
            if (!cont) break;
            const float4 d = pb.d[slot];
            const TRay r{o.x, o.y, o.z, 0.001f, d.x, d.y, d.z, __int_as_float(0x7f800000)};
            trace<false, MIXED, false, kLds, kBlock, kNodeCache>(sc, r, h, st, nullptr, flat_gate(sc, gate) ? gate : 0);
            pb.hit[slot] = make_float4(h.t, __uint_as_float(h.prim), h.u, h.v);
        }
    }
}

// ---- bidirectional path tracing (BDPTIntegrator.RayColor, pkg/integrator/bdpt.go:73-102) ----------

// Distance along d from the float64 point o_hi + o_lo to the surface of primitive `prim` near the
// float32 hit distance t32 (plane distance for triangles, quads and discs; the nearer sphere root).
__device__ double prim_t_f64(const DevScene& sc, uint32_t prim, V3 o_hi, const float* o_lo, V3 d, float t32) {
    const float4* pp = sc.prims + 3 * (size_t)prim;
    const float4 a = pp[0], b = pp[1];
    const uint32_t kind = prim_kind(__float_as_uint(a.w));
    const double ox = (double)o_hi.x + o_lo[0], oy = (double)o_hi.y + o_lo[1], oz = (double)o_hi.z + o_lo[2];
    const double dx = d.x, dy = d.y, dz = d.z;
    double nx, ny, nz, px, py, pz;  // plane normal and a point on it
    if (kind == kPrimTriangle) {
        const float4 c = pp[2];
        double e1x = (double)b.x - a.x, e1y = (double)b.y - a.y, e1z = (double)b.z - a.z;
        double e2x = (double)c.x - a.x, e2y = (double)c.y - a.y, e2z = (double)c.z - a.z;
        nx = e1y * e2z - e1z * e2y;
        ny = e1z * e2x - e1x * e2z;
        nz = e1x * e2y - e1y * e2x;
        px = a.x, py = a.y, pz = a.z;
    } else if (kind == kPrimQuad) {
        const QuadExt& q = sc.quads[prim_side(__float_as_uint(a.w))];
        nx = q.nx, ny = q.ny, nz = q.nz;
        px = a.x, py = a.y, pz = a.z;
    } else if (kind == kPrimDisc) {
        nx = b.x, ny = b.y, nz = b.z;
        px = a.x, py = a.y, pz = a.z;
    } else if (kind != kPrimSphere) {  // cylinders and cones keep their float32 distance
        return (double)t32;
    } else {  // sphere: the root of |o + t d - c|^2 = r^2 nearest t32
        double fx = ox - a.x, fy = oy - a.y, fz = oz - a.z;
        double A = dx * dx + dy * dy + dz * dz, hb = fx * dx + fy * dy + fz * dz;
        double C = fx * fx + fy * fy + fz * fz - (double)b.x * b.x;
        double disc = hb * hb - A * C;
        if (disc < 0) return (double)t32;
        double sq = sqrt(disc), t0 = (-hb - sq) / A, t1 = (-hb + sq) / A;
        return fabs(t0 - t32) <= fabs(t1 - t32) ? t0 : t1;
    }
    double den = nx * dx + ny * dy + nz * dz;
    if (den == 0.0) return (double)t32;
    return (nx * (px - ox) + ny * (py - oy) + nz * (pz - oz)) / den;
}

// extendPath, bdpt.go:206-291. Appends vertices from index n; returns the new length. `point_lo`
// is the low part of the origin when the first segment leaves a point light, else null: the light's
// float32 position can lie off its float64 position by half an ulp, which moves float32 hit
// distances of surfaces within ~tmin of the light across tmin = 0.001 (the reference's Cornell
// point light hangs exactly 0.001 below the ceiling). Hits of that segment nearer than 0.01 are
// re-decided in float64 from the light's low part; a rejected hit is skipped and the search resumes
// beyond it (after four rejections, beyond 0.01). One traversal call site keeps the kernel small.
template <bool MIXED>
__device__ int bdpt_extend(const DevScene& sc, const VertexStore& VS, uint32_t i, int n, V3 o, V3 d, V3 beta,
                           float pdf_fwd, int max_bounces, bool camera_path, uint32_t key, uint32_t slot0,
                           TStack<kLds, kBlock>& st, const float* point_lo, int gate) {
    for (int b = 0; b < max_bounces; b++) {
        BV prev = VS.get(n - 1, i);
        THit h;
        bool hit;
        const bool redecide = b == 0 && point_lo;
        float tmin = redecide ? 0.0f : 0.001f;
        for (int k = 0;; k++) {
            TRay r{o.x, o.y, o.z, tmin, d.x, d.y, d.z, __int_as_float(0x7f800000)};
            hit = trace<false, MIXED, false, kLds, kBlock, kNodeCache>(sc, r, h, st, nullptr, gate);
            if (!redecide || !hit || h.t >= 0.01f || k == 4) break;
            if (prim_t_f64(sc, h.prim, o, point_lo, d, h.t) >= 0.001) break;
            const float next = __uint_as_float(__float_as_uint(h.t) + 1u);
            tmin = k == 3 ? fmaxf(next, 0.01f) : next;
        }
        if (!hit) {
            if (camera_path) {  // createBackgroundVertex, bdpt.go:585-601
                V3 e = v3(0, 0, 0);
                for (int li = 0; li < sc.num_lights; li++)
                    if (light_is_infinite(sc.lights[li].type)) e = e + light_emit(sc, sc.lights[li], d);
                BV v;
                v.p = o + d * 1000.0f;
                v.n = d * -1.0f;
                v.wi = d * -1.0f;
                v.beta = beta;
                v.Le = e;
                v.pdf_fwd = pdf_fwd;
                v.pdf_rev = 0.0f;
                v.mat = -1;
                v.light = -1;
                v.flags = VF_INF | ((e.x != 0.0f || e.y != 0.0f || e.z != 0.0f) ? VF_LIGHT : 0u);
                VS.put(n++, i, v);
            }
            break;
        }
        Surf s = make_surf(sc, o, d, h);
        BV v;
        v.p = s.p;
        v.n = s.n;
        v.wi = d * -1.0f;
        v.beta = beta;
        v.mat = s.mat;
        v.light = -1;
        v.Le = emit_hit(sc, s.mat, d, s.front);
        v.flags = (s.front ? VF_FRONT : 0u) | ((v.Le.x != 0.0f || v.Le.y != 0.0f || v.Le.z != 0.0f) ? VF_LIGHT : 0u);
        v.u = s.u;
        v.v = s.v;
        v.pdf_fwd = convert_density(sc, prev, v, pdf_fwd);
        v.pdf_rev = 0.0f;
        ScatterOut so;
        if (!scatter(sc, s.mat, d, s, key, (slot0 + (uint32_t)b) * 16u + 1u, so)) {
            VS.put(n++, i, v);
            break;
        }
        pdf_fwd = so.pdf;
        bool spec = so.pdf <= 0.0f;
        float ct = fabsf(dot(so.dir, s.n));
        beta = spec ? mulv(beta, so.att) : mulv(beta, so.att) * (ct / so.pdf);
        bool rdelta;
        float pdf_rev = mat_pdf(sc, s.mat, so.dir, d * -1.0f, s.n, rdelta);
        if (rdelta) {
            spec = true;
            pdf_rev = 0.0f;
            pdf_fwd = 0.0f;
        }
        if (spec) v.flags |= VF_SPECULAR;
        VS.set_pdf_rev(n - 1, i, convert_density(sc, v, prev, pdf_rev));
        VS.put(n++, i, v);
        o = s.p;
        d = so.dir;
    }
    return n;
}

// Subpath generation: one thread per sample writes its camera and light subpaths and their lengths.
#ifndef PRT_PATHS_WAVES
#define PRT_PATHS_WAVES 3  // waves per SIMD of k_bdpt_paths
#endif
// SIDES: 3 both subpaths, 1 the camera subpath only, 2 the light subpath only (a kernel per side is smaller; each
// writes its half of the packed lengths).
template <bool MIXED, int SIDES = 3>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(PRT_PATHS_WAVES))) void k_bdpt_paths(DevScene sc, PassParams pp, VertexStore CP, VertexStore LP,
                                                       uint32_t* lens, uint2* gstack, uint32_t gstride) {
    __shared__ uint32_t s_node[kLds * kBlock];
    __shared__ float s_t[kLds * kBlock];
    __shared__ float4 s_cache[kNodeCache * 8];
    TStack<kLds, kBlock> st;
    st.s_node = s_node;
    st.s_t = s_t;
    st.g = gstack;
    st.lane = threadIdx.x;
    st.gslot = blockIdx.x * kBlock + threadIdx.x;
    st.gstride = gstride;
    load_node_cache<kNodeCache, kBlock>(sc, s_cache);
    st.ncache = s_cache;
    const DevCamera& cam = sc.cam;
    const uint32_t ts = (uint32_t)pp.tune_stride;
    const uint32_t n = ((uint32_t)pp.npix * (uint32_t)pp.spp + ts - 1u) / ts;
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < n; j += stride) {
        const uint32_t i = j * ts;
        uint32_t pix, sample;
        slot_sample(pp, i, pix, sample);
        int px = (int)(pix % (uint32_t)pp.width), py = (int)(pix / (uint32_t)pp.width);
        uint32_t key = path_key(pp.seed, pix, sample);
        uint32_t nc = 0, nl = 0;
        if constexpr ((SIDES & 1) != 0) {
            V3 o, d;
            camera_ray(cam, px, py, draw(key, 0), draw(key, 1), draw(key, 2), draw(key, 3), o, d);
            // camera subpath (generateCameraPath, bdpt.go:106-135)
            BV cv;
            cv.p = o;
            cv.n = d * -1.0f;
            cv.wi = v3(0, 0, 0);
            cv.beta = v3(1, 1, 1);
            cv.Le = v3(0, 0, 0);
            cv.pdf_fwd = cv.pdf_rev = 0.0f;
            cv.mat = -1;
            cv.light = -1;
            cv.flags = VF_CAMERA;
            CP.put(0, i, cv);
            nc = (uint32_t)bdpt_extend<MIXED>(sc, CP, i, 1, o, d, v3(1, 1, 1), cam_dir_pdf(cam, o, d),
                                                             pp.max_depth, true, key, 1, st, nullptr, pp.trace_gate);
        }
        // light subpath (generateLightPath, bdpt.go:139-203)
        if constexpr ((SIDES & 2) != 0) {
            if (sc.num_lights > 0) {
                const uint32_t lp0 = kLightSlot * 16u;
                float sel;
                int li = select_light(sc, draw(key, lp0), sel);
                const DevLight& L = sc.lights[li];
                ESample es = light_sample_emission(sc, L, draw(key, lp0 + 1), draw(key, lp0 + 2), draw(key, lp0 + 3),
                                                   draw(key, lp0 + 4));
                float ct = fabsf(dot(es.dir, es.n));
                BV lv;
                lv.p = es.p;
                lv.n = es.n;
                lv.wi = v3(0, 0, 0);
                lv.beta = es.emission;
                lv.Le = es.emission;
                lv.pdf_fwd = es.area_pdf * sel;
                lv.pdf_rev = 0.0f;
                lv.mat = -1;
                lv.light = li;
                bool inf = light_is_infinite(L.type);
                lv.flags = light_vertex_flags(sc, li);
                LP.put(0, i, lv);
                V3 beta = es.emission * (ct / (sel * es.area_pdf * es.dir_pdf));
                nl = (uint32_t)bdpt_extend<MIXED>(sc, LP, i, 1, es.p, es.dir, beta, es.dir_pdf, pp.max_depth - 1, false,
                                                  key, kLightSlot + 1, st,
                                                  L.type == PRT_LIGHT_POINT_SPOT ? L.from_lo : nullptr, pp.trace_gate);
                if (inf) {  // infinite-light density fix-ups (bdpt.go:182-199)
                    if (nl > 1) {
                        BV fb = LP.get(1, i);
                        float pf = es.area_pdf * sel;
                        if (fb.mat >= 0) pf *= fabsf(dot(es.dir, fb.n));
                        LP.set_pdf_fwd(1, i, pf);
                    }
                    LP.set_pdf_fwd(0, i, infinite_density(sc, es.p, es.n, es.dir));
                }
            }
        }
        if constexpr (SIDES == 3)
            lens[i] = nc | (nl << 16);
        else if constexpr (SIDES == 1)
            reinterpret_cast<uint16_t*>(lens)[2 * i] = (uint16_t)nc;
        else
            reinterpret_cast<uint16_t*>(lens)[2 * i + 1] = (uint16_t)nl;
    }
}

// Maximum of a small non-negative value (below 128) over the active lanes of the wave.
__device__ __forceinline__ int wave_max_small(int v) {
    int m = 0;
    for (int bit = 6; bit >= 0; bit--) {
        const int c = m | (1 << bit);
        if (__ballot(v >= c) != 0) m = c;
    }
    return m;
}

// The MIS walk sums of both subpaths of every sample (store_walk_sums), between the subpath and the
// strategy kernels.
__global__ __launch_bounds__(kBlock) void k_bdpt_walks(PassParams pp, VertexStore CP, VertexStore LP,
                                                       const uint32_t* lens) {
    const uint32_t n = (uint32_t)pp.npix * (uint32_t)pp.spp;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t len = lens[i];
    store_walk_sums(CP, i, (int)(len & 0xFFFFu), false);
    store_walk_sums(LP, i, (int)(len >> 16), true);
}

// Strategy (s, t) of sample i (key: its counter-RNG key) for s = 0 (t = nc), s = 1 or t = 1, or a general
// connection: evaluatePathTracingStrategy (bdpt.go:316-330), evaluateDirectLightingStrategy (:333-391),
// evaluateLightTracingStrategy (:394-471), evaluateConnectionStrategy (:489-561), each with its visibility ray
// and calculateMISWeight. Light-tracing strategies (t = 1) splat their contribution themselves and return
// false; the others return true with the weighted camera-side radiance in `out` when it is added.
// Strategy lists of the BDPT work lists (k_bdpt_pairs): general pairs, direct lighting (1, t), light tracing
// (s, 1), the camera path's own emission (0, nc). A kernel compiled for one kind evaluates only its branch.
constexpr int kKindAny = -1, kKindGeneral = 0, kKindDirect = 1, kKindLightTrace = 2, kKindEmission = 3;
constexpr int kNumLists = 4;

template <bool MIXED, int KIND = kKindAny>
__device__ __forceinline__ bool eval_strategy(const DevScene& sc, const PassParams& pp, const VertexStore& CP,
                                              const VertexStore& LP, uint32_t i, uint32_t key, int s, int t,
                                              long long* splat, TStack<kLds, kBlock>& st, V3& out) {
    if constexpr (KIND == kKindDirect) s = 1;
    if constexpr (KIND == kKindLightTrace) t = 1;
    if constexpr (KIND == kKindEmission) s = 0;
    const DevCamera& cam = sc.cam;
    V3 Lc = v3(0, 0, 0);
    bool splatted = false;
    int sx = 0, sy = 0;
    BV sampled, ct1, ls1;
    TRay vr;  // visibility ray of the strategy (one trace call site keeps the kernel call-free)
    bool vis = false;
    ct1 = CP.get(t - 1, i);
    if (s > 0) ls1 = LP.get(s - 1, i);
    if (s == 0) {  // evaluatePathTracingStrategy :316-330
        Lc = mulv(ct1.Le, ct1.beta);
    } else if (t == 1) {  // evaluateLightTracingStrategy :394-471
        const BV& lv = ls1;
        if (lv.flags & VF_SPECULAR) return false;
        const uint32_t sp0 = (kSplatSlot + (uint32_t)s) * 16u;
        V3 lc = sample_point_in_unit_disk(draw(key, sp0), draw(key, sp0 + 1)) * cam.lens_radius;
        V3 lens_p = v3(cam.center) + v3(cam.u) * lc.x + v3(cam.v) * lc.y;
        V3 dir = normalize(lv.p - lens_p);
        float dist = length(lv.p - lens_p);
        float cf = fabsf(dot(v3(cam.forward), dir));
        float cpdf = (dist * dist) / (cf * cam.lens_area);
        float ctw = dot(dir, v3(cam.forward));
        float we = ctw < cam.cos_total_width
                       ? 0.0f
                       : 1.0f / (cam.plane_area * cam.lens_area * ctw * ctw * ctw * ctw);
        if (we == 0.0f) return false;
        V3 back = dir * -1.0f;
        V3 f = vertex_brdf(sc, lv, back, 1);
        V3 cbeta = v3(we, we, we) * (1.0f / cpdf);
        float cosine = dot(back, lv.n);
        if (cosine <= 0.0f) return false;
        V3 contrib = mulv(mulv(f, cbeta), lv.beta);
        if (on_surface(sc, lv)) contrib = contrib * cosine;
        if (contrib.x == 0.0f && contrib.y == 0.0f && contrib.z == 0.0f) return false;
        // the splat lands outside the image: SplatQueue drops it, nothing else uses it
        if (!cam_map(cam, lens_p, dir, sx, sy)) return false;
        vr = TRay{lv.p.x, lv.p.y, lv.p.z, 0.001f, back.x, back.y, back.z, dist - 0.001f};
        vis = true;
        sampled.p = lens_p;
        sampled.n = back;
        sampled.wi = v3(0, 0, 0);
        sampled.beta = cbeta;
        sampled.Le = v3(0, 0, 0);
        sampled.pdf_fwd = cpdf;
        sampled.pdf_rev = total_light_pdf(sc, lens_p, back, dir);
        sampled.mat = -1;
        sampled.light = -1;
        sampled.flags = VF_CAMERA;
        Lc = contrib;
        splatted = true;
    } else if (s == 1) {  // evaluateDirectLightingStrategy :333-391
        const BV& cvt = ct1;
        if ((cvt.flags & VF_SPECULAR) || cvt.mat < 0) return false;
        const uint32_t dp0 = (kDirectSlot + (uint32_t)t) * 16u;
        float sel;
        int li = select_light(sc, draw(key, dp0), sel);
        LSample ls = light_sample(sc, sc.lights[li], cvt.p, cvt.n, draw(key, dp0 + 1), draw(key, dp0 + 2));
        ls.pdf *= sel;
        if ((ls.emission.x == 0.0f && ls.emission.y == 0.0f && ls.emission.z == 0.0f) || ls.pdf <= 0.0f)
            return false;
        float ct = fabsf(dot(ls.dir, cvt.n));
        if (ct <= 0.0f) return false;
        V3 f = brdf(sc, cvt.mat, cvt.wi, ls.dir, surf_of(cvt), 0);
        V3 lbeta = ls.emission * (1.0f / ls.pdf);
        V3 contrib = mulv(mulv(f, cvt.beta), lbeta) * ct;
        if (contrib.x == 0.0f && contrib.y == 0.0f && contrib.z == 0.0f) return false;
        vr = TRay{cvt.p.x, cvt.p.y, cvt.p.z, 0.001f, ls.dir.x, ls.dir.y, ls.dir.z, ls.dist - 0.001f};
        vis = true;
        sampled.p = ls.p;
        sampled.n = ls.n;
        sampled.wi = v3(0, 0, 0);
        sampled.beta = lbeta;
        sampled.Le = ls.emission;
        sampled.pdf_fwd = ls.pdf;
        sampled.pdf_rev = 0.0f;
        sampled.mat = -1;
        sampled.light = li;
        sampled.flags = light_vertex_flags(sc, li);
        Lc = contrib;
    } else {  // evaluateConnectionStrategy :489-561
        const BV &lv = ls1, &cvt = ct1;
        if ((lv.flags & VF_SPECULAR) || (cvt.flags & VF_SPECULAR)) return false;
        V3 dd = lv.p - cvt.p;
        float dist = length(dd);
        if (dist < 0.001f) return false;
        dd = dd * (1.0f / dist);
        float cc = dot(dd, cvt.n), cl = dot(dd * -1.0f, lv.n);
        if (cc <= 0.0f || cl <= 0.0f) return false;
        float G = (cc * cl) / (dist * dist);
        V3 fc = vertex_brdf(sc, cvt, dd, 0);
        V3 fl = vertex_brdf(sc, lv, dd * -1.0f, 1);
        V3 contrib = mulv(mulv(mulv(lv.beta, fl), fc), cvt.beta) * G;
        if (contrib.x == 0.0f && contrib.y == 0.0f && contrib.z == 0.0f) return false;
        vr = TRay{cvt.p.x, cvt.p.y, cvt.p.z, 0.001f, dd.x, dd.y, dd.z, dist - 0.001f};
        vis = true;
        Lc = contrib;
    }
    if (vis) {
        THit vh;
        if (trace<true, MIXED, false, kLds, kBlock, kNodeCache>(sc, vr, vh, st, nullptr, pp.trace_gate)) return false;
    }
    if (Lc.x == 0.0f && Lc.y == 0.0f && Lc.z == 0.0f && !splatted) return false;
    float w = mis_weight(sc, cam, CP, LP, i, sampled, ct1, ls1, s, t);
    if (splatted) {
        V3 c = Lc * w;
        // fixed point holds |c| < 2^36 per splat (non-finite or larger splats are dropped)
        if (!(fabsf(c.x) < 6.8e10f && fabsf(c.y) < 6.8e10f && fabsf(c.z) < 6.8e10f)) return false;
        long long* o3 = splat + 3 * ((size_t)sy * pp.width + sx);
        atomicAdd((unsigned long long*)(o3 + 0), (unsigned long long)llrintf(c.x * kSplatScale));
        atomicAdd((unsigned long long*)(o3 + 1), (unsigned long long)llrintf(c.y * kSplatScale));
        atomicAdd((unsigned long long*)(o3 + 2), (unsigned long long)llrintf(c.z * kSplatScale));
        return false;
    }
    out = Lc * w;
    return true;
}

// Strategies with a sampled or single vertex: one thread per sample evaluates (0, nc), (1, t) and (s, 1) of
// RayColor (bdpt.go:83-99 with evaluateBDPTStrategy :294-313), camera-side radiance to pb.L, light-tracing
// splats to `splat`. k_bdpt_connect_general adds the general connections to pb.L afterwards.
template <bool MIXED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_bdpt_connect(DevScene sc, PassParams pp, PathBuffers pb, VertexStore CP,
                                                         VertexStore LP, const uint32_t* lens, long long* splat,
                                                         uint2* gstack, uint32_t gstride) {
    __shared__ uint32_t s_node[kLds * kBlock];
    __shared__ float s_t[kLds * kBlock];
    __shared__ float4 s_cache[kNodeCache * 8];
    TStack<kLds, kBlock> st;
    st.s_node = s_node;
    st.s_t = s_t;
    st.g = gstack;
    st.lane = threadIdx.x;
    st.gslot = blockIdx.x * kBlock + threadIdx.x;
    st.gstride = gstride;
    load_node_cache<kNodeCache, kBlock>(sc, s_cache);
    st.ncache = s_cache;
    const DevCamera& cam = sc.cam;
    const uint32_t n = (uint32_t)pp.npix * (uint32_t)pp.spp;
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        uint32_t pix, sample;
        slot_sample(pp, i, pix, sample);
        uint32_t key = path_key(pp.seed, pix, sample);
        const uint32_t len = lens[i];
        const int nc = (int)(len & 0xFFFFu), nl = (int)(len >> 16);
        V3 total = v3(0, 0, 0);
        // The reference's pairs (s = 0 with t = nc, every other s <= nl and t <= nc but (1, 1)) in four
        // phases: (0, nc); (1, t) for t = 2..nc; (s, 1) for s = 2..nl; (s, t >= 2) for s = 2..nl in s-major
        // order. The first three run in steps of the same length across the wave, so that a wave evaluates
        // one kind of strategy at a time; the camera-side sum keeps the reference's (s outer, t inner)
        // order (phase 3 only splats, with integer atomics).
        const int lb = nl >= 1 ? nc - 1 : 0, lc = nl >= 2 ? nl - 1 : 0;
        const int wb = wave_max_small(lb), wc = wave_max_small(lc);
        // the general connections (s >= 2, t >= 2) run in k_bdpt_connect_general after this kernel
        const int npairs = 1 + wb + wc;
        int ds = 2, dt = 1;  // phase 4 position
        for (int j = 0; j < npairs; j++) {
            int s, t;
            if (j == 0) {
                s = 0;
                t = nc;
            } else if (j <= wb) {
                if (j > lb) continue;
                s = 1;
                t = j + 1;
            } else if (j <= wb + wc) {
                const int k = j - wb;
                if (k > lc) continue;
                s = k + 1;
                t = 1;
            } else {
                if (++dt > nc) {
                    dt = 2;
                    ds++;
                }
                s = ds;
                t = dt;
            }
            V3 c;
            if (eval_strategy<MIXED>(sc, pp, CP, LP, i, key, s, t, splat, st, c)) total = total + c;
        }
        pb.L[i] = make_float4(total.x, total.y, total.z, 0.f);
    }
}

// calculateMISWeight for a general connection (s >= 2, t >= 2): mis_weight's last case, with the light side's
// walk past ls2 (l_tail, l_fwd2) and the segment ls1 -> ls2 (c) computed once per light vertex.
__device__ __forceinline__ float mis_general(const DevScene& sc, const DevCamera& cam, const BV& ct1, const BV& ct2,
                                            const BV& ls1, const BV& ls2, const Seg& c, float l_tail, float l_fwd2,
                                            int t) {
    const bool c_first = !(ct2.flags & VF_SPECULAR);
    const float c_tail = ((ct2.flags & VF_CONN) ? 1.0f : 0.0f) + ct2.q;
    const float c_fwd2 = remap0(ct2.pdf_fwd);
    const Seg a = seg_of(ls1.p - ct1.p), b = seg_of(ct2.p - ct1.p);
    const Seg ar = seg_rev(a);
    const float c_rev1 = vtx_pdf_seg(sc, cam, ls1, c, ct1, ar);
    const float c_rev2 = vtx_pdf_seg(sc, cam, ct1, a, ct2, b);
    const float l_rev1 = vtx_pdf_seg(sc, cam, ct1, b, ls1, a);
    const float l_rev2 = vtx_pdf_seg(sc, cam, ls1, ar, ls2, c);
    float sum = 0.0f;
    float ri = remap0(c_rev1) / remap0(ct1.pdf_fwd);
    if (c_first) sum += ri;
    if (t > 2) sum += ri * (remap0(c_rev2) / c_fwd2) * c_tail;
    ri = remap0(l_rev1) / remap0(ls1.pdf_fwd);
    sum += ri + ri * (remap0(l_rev2) / l_fwd2) * l_tail;
    return 1.0f / (1.0f + sum);
}

// General connections (evaluateConnectionStrategy, bdpt.go:489-561): every (s, t) with s, t >= 2 in the
// reference's order (s outer, t inner), added to the camera-side radiance pb.L. Each light vertex is read
// once per sample (vertex s - 2 of a row is vertex s - 1 of the previous row) and each camera vertex once
// per row (vertex t - 2 is the previous step's t - 1), instead of four vertex reads per strategy.
template <bool MIXED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_bdpt_connect_general(
    DevScene sc, PassParams pp, PathBuffers pb, VertexStore CP, VertexStore LP, const uint32_t* lens, uint2* gstack,
    uint32_t gstride) {
    __shared__ uint32_t s_node[kLds * kBlock];
    __shared__ float s_t[kLds * kBlock];
    __shared__ float4 s_cache[kNodeCache * 8];
    TStack<kLds, kBlock> st;
    st.s_node = s_node;
    st.s_t = s_t;
    st.g = gstack;
    st.lane = threadIdx.x;
    st.gslot = blockIdx.x * kBlock + threadIdx.x;
    st.gstride = gstride;
    load_node_cache<kNodeCache, kBlock>(sc, s_cache);
    st.ncache = s_cache;
    const DevCamera& cam = sc.cam;
    const uint32_t n = (uint32_t)pp.npix * (uint32_t)pp.spp;
    const uint32_t stride = gridDim.x * kBlock;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        const uint32_t len = lens[i];
        const int nc = (int)(len & 0xFFFFu), nl = (int)(len >> 16);
        if (nl < 2 || nc < 2) continue;
        const float4 l4 = pb.L[i];
        V3 total = v3(l4.x, l4.y, l4.z);
        BV ls2 = LP.get(0, i);
        for (int s = 2; s <= nl; s++) {
            const BV ls1 = LP.get(s - 1, i);
            if (!(ls1.flags & VF_SPECULAR)) {
                const float l_tail = ((ls2.flags & VF_CONN) ? 1.0f : 0.0f) + ls2.q, l_fwd2 = remap0(ls2.pdf_fwd);
                const Seg c = seg_of(ls2.p - ls1.p);
                BV ct2 = CP.get(0, i);
                for (int t = 2; t <= nc; t++) {
                    const BV ct1 = CP.get(t - 1, i);
                    if (!(ct1.flags & VF_SPECULAR)) {
                        V3 dd = ls1.p - ct1.p;
                        const float dist = length(dd);
                        bool ok = dist >= 0.001f;
                        V3 contrib = v3(0, 0, 0);
                        TRay vr;
                        if (ok) {
                            dd = dd * (1.0f / dist);
                            const float cc = dot(dd, ct1.n), cl = dot(dd * -1.0f, ls1.n);
                            ok = cc > 0.0f && cl > 0.0f;
                            if (ok) {
                                const float G = (cc * cl) / (dist * dist);
                                const V3 fc = vertex_brdf(sc, ct1, dd, 0);
                                const V3 fl = vertex_brdf(sc, ls1, dd * -1.0f, 1);
                                contrib = mulv(mulv(mulv(ls1.beta, fl), fc), ct1.beta) * G;
                                ok = contrib.x != 0.0f || contrib.y != 0.0f || contrib.z != 0.0f;
                                vr = TRay{ct1.p.x, ct1.p.y, ct1.p.z, 0.001f, dd.x, dd.y, dd.z, dist - 0.001f};
                            }
                        }
                        if (ok) {
                            // the weight before the visibility ray: only the weighted contribution stays live
                            // across the traversal
                            const V3 lw = contrib * mis_general(sc, cam, ct1, ct2, ls1, ls2, c, l_tail, l_fwd2, t);
                            THit vh;
                            if (!trace<true, MIXED, false, kLds, kBlock, kNodeCache>(sc, vr, vh, st, nullptr,
                                                                                     pp.trace_gate))
                                total = total + lw;
                        }
                    }
                    ct2 = ct1;
                }
            }
            ls2 = ls1;
        }
        pb.L[i] = make_float4(total.x, total.y, total.z, 0.f);
    }
}

// BDPT strategies through per-strategy work lists (films whose strategies fit kMaxQueues, max_depth <= 16;
// deeper films run k_bdpt_connect and k_bdpt_connect_general, one thread per sample looping over its
// strategies). Strategy q of RayColor (bdpt.go:83-99), nt = max_depth:
//   q in [0, np):            general connection s = q / nt + 2, t = q % nt + 2 (np = (nt - 1) * nt);
//   q in [np, np + nt):      direct lighting (1, t), t = q - np + 2;
//   q in [np + nt, nq - 1):  light tracing (s, 1), s = q - np - nt + 2;
//   q = nq - 1:              the camera path's own emission (0, nc).
// Result planes (the weighted camera-side radiance of a queued sample, w = 1 when it is added): one per general
// pair and per direct-lighting strategy, then one for (0, nc); light tracing splats directly.
//   k_bdpt_pairs:         stores the MIS walk sums of both subpaths (as k_bdpt_walks), then each block of
//                         samples lists (sample, q) for every strategy the per-sample loops
//                         evaluate (end vertices not specular, a material for direct lighting, an emitting last
//                         camera vertex for (0, nc)) in one contiguous segment per list, ordered by q and then
//                         sample, and records which strategies in `mask`;
//   k_bdpt_edge_connect:  one lane per listed direct-lighting, light-tracing or (0, nc) strategy;
//   k_bdpt_pair_connect:  one lane per listed general pair (the contribution, MIS weight and visibility ray of
//                         k_bdpt_connect_general);
//   k_bdpt_pair_sum:      one thread per sample adds its results in the per-sample loops' order ((0, nc), then
//                         (1, t) by t, then the general pairs s-major), so the radiance is that of the per-sample
//                         kernels up to the contraction of multiply-adds.
// Lanes of a wave evaluate one strategy for consecutive samples (no lane waits for another lane's longer
// subpaths; coalesced vertex reads), and the lists are consumed in ranges that are contiguous per XCD, so the
// strategies of a block of samples run together on one XCD while its subpath vertices are in that XCD's L2.
constexpr uint32_t kMaxQueues = 512;
constexpr uint32_t kItemSlotBits = 23;  // list item: slot | q << 23 (slots per chunk < 2^23)
constexpr uint32_t kNumXcd = 8;

struct PairBuffers {
    uint32_t* items[kNumLists];  // per strategy kind: slot | (strategy index within the kind) << kItemSlotBits
    float4* res;       // [plane][slot]: weighted contribution, w = 1 if it is added
    uint2* mask;       // per slot: x = light vertices s - 1 not specular (bit s - 2) | camera vertices t - 1 not
                       // specular (bit 16 + t - 2); y = direct-lighting strategies (bit t - 2) | (0, nc) (bit 16)
    uint32_t* count;   // length of list k at count[k * kCounterStride]
    uint32_t nt;       // max_depth
    uint32_t np;       // general pairs
    uint32_t nq;       // strategies: np + 2 * nt
    uint32_t stride;   // slots per result plane
    __device__ __forceinline__ float4* plane(uint32_t k) const { return res + (size_t)k * stride; }
};

// Exclusive prefix sums of v[0 .. kMaxQueues) in place (two entries per thread); returns the total. Every
// thread of the block calls it.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t* v, uint32_t* s_tmp) {
    static_assert(kMaxQueues == 2 * kBlock, "two entries per thread");
    const uint32_t k = 2 * threadIdx.x;
    const uint32_t a = v[k], b = v[k + 1];
    s_tmp[threadIdx.x] = a + b;
    __syncthreads();
    for (uint32_t o = 1; o < kBlock; o <<= 1) {
        const uint32_t x = threadIdx.x >= o ? s_tmp[threadIdx.x - o] : 0u;
        __syncthreads();
        s_tmp[threadIdx.x] += x;
        __syncthreads();
    }
    const uint32_t before = s_tmp[threadIdx.x] - (a + b);
    const uint32_t total = s_tmp[kBlock - 1];
    v[k] = before;
    v[k + 1] = before + a;
    __syncthreads();
    return total;
}

__global__ __launch_bounds__(kBlock) void k_bdpt_pairs(PassParams pp, VertexStore CP, VertexStore LP,
                                                       const uint32_t* lens, PairBuffers pq) {
    __shared__ uint32_t s_cnt[kBlock / 64][kMaxQueues];  // per wave and strategy: entries, then offsets
    __shared__ uint32_t s_off[kMaxQueues];               // the block's entries per strategy, then offsets
    __shared__ uint32_t s_tmp[kBlock];
    __shared__ uint32_t s_base[kNumLists];
    const uint32_t n = (uint32_t)pp.npix * (uint32_t)pp.spp;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t lm = 0, cm = 0, dm = 0, zm = 0;
    if (i < n) {
        const uint32_t len = lens[i];
        const int nc = (int)(len & 0xFFFFu), nl = (int)(len >> 16);
        // the MIS walk sums of both subpaths first (k_bdpt_walks of the per-sample kernels)
        store_walk_sums(CP, i, nc, false);
        store_walk_sums(LP, i, nl, true);
        for (int s = 2; s <= nl; s++)
            if (!rec_specular(LP.rec(s - 1, i))) lm |= 1u << (s - 2);
        for (int t = 2; t <= nc; t++) {
            const uint32_t fm = __float_as_uint(CP.rec(t - 1, i).z);
            if (fm & VF_SPECULAR) continue;
            cm |= 1u << (t - 2);
            if (nl >= 1 && (fm >> 8) != 0u) dm |= 1u << (t - 2);  // a material (mat >= 0)
        }
        if (nc >= 1 && (__float_as_uint(CP.rec(nc - 1, i).z) & VF_LIGHT)) zm = 1;
        pq.mask[i] = make_uint2(lm | (cm << 16), dm | (zm << 16));
    }
    for (uint32_t q = threadIdx.x; q < kMaxQueues; q += kBlock) {
        s_off[q] = 0;
        for (uint32_t v = 0; v < kBlock / 64; v++) s_cnt[v][q] = 0;
    }
    __syncthreads();
    // wave-uniform bounds of the strategies any lane has
    uint32_t ns = 0, nc2 = 0, nd = 0;
    while (__ballot((lm >> ns) != 0u) != 0) ns++;
    while (__ballot((cm >> nc2) != 0u) != 0) nc2++;
    while (__ballot((dm >> nd) != 0u) != 0) nd++;
    const uint32_t qd = pq.np, ql = pq.np + pq.nt, qz = pq.nq - 1;
    auto count = [&](uint32_t q, bool a) {
        const uint64_t m = __ballot(a);
        if (lane == 0) s_cnt[w][q] = (uint32_t)__popcll(m);
    };
    for (uint32_t s = 0; s < ns; s++)
        for (uint32_t t = 0; t < nc2; t++) count(s * pq.nt + t, ((lm >> s) & (cm >> t) & 1u) != 0);
    for (uint32_t t = 0; t < nd; t++) count(qd + t, ((dm >> t) & 1u) != 0);
    for (uint32_t s = 0; s < ns; s++) count(ql + s, ((lm >> s) & 1u) != 0);
    count(qz, zm != 0);
    __syncthreads();
    // per strategy: the waves' offsets within the block's entries, and the block's count
    for (uint32_t q = threadIdx.x; q < pq.nq; q += kBlock) {
        uint32_t a = 0;
        for (uint32_t v = 0; v < kBlock / 64; v++) {
            const uint32_t c = s_cnt[v][q];
            s_cnt[v][q] = a;
            a += c;
        }
        s_off[q] = a;
    }
    __syncthreads();
    // the block's segment of each list: strategies in order (general pairs, direct lighting, light tracing,
    // emission), one reservation per list
    const uint32_t all = block_exclusive_scan(s_off, s_tmp);
    if (threadIdx.x < (uint32_t)kNumLists) {
        const uint32_t q0[kNumLists + 1] = {0u, pq.np, pq.np + pq.nt, pq.nq - 1, pq.nq};
        const uint32_t k = threadIdx.x;
        const uint32_t b0 = s_off[q0[k]], b1 = k + 1 < (uint32_t)kNumLists ? s_off[q0[k + 1]] : all;
        s_base[k] = (b1 > b0 ? atomicAdd(pq.count + k * kCounterStride, b1 - b0) : 0u) - b0;
    }
    __syncthreads();
    auto push = [&](uint32_t q, bool a) {
        const uint64_t m = __ballot(a);
        if (!a) return;
        const uint32_t pos = s_off[q] + s_cnt[w][q] + (uint32_t)__popcll(m & lt);
        const uint32_t k = q < qd ? 0u : q < ql ? 1u : q < qz ? 2u : 3u;
        const uint32_t e = q - (k == 0 ? 0u : k == 1 ? qd : k == 2 ? ql : qz);
        pq.items[k][s_base[k] + pos] = i | (e << kItemSlotBits);
    };
    for (uint32_t s = 0; s < ns; s++)
        for (uint32_t t = 0; t < nc2; t++) push(s * pq.nt + t, ((lm >> s) & (cm >> t) & 1u) != 0);
    for (uint32_t t = 0; t < nd; t++) push(qd + t, ((dm >> t) & 1u) != 0);
    for (uint32_t s = 0; s < ns; s++) push(ql + s, ((lm >> s) & 1u) != 0);
    push(qz, zm != 0);
}

// The items of a list a block takes: the list is cut into kNumXcd contiguous ranges, one per XCD (work-groups
// are dealt to the XCDs in turn, so block b runs on XCD b % kNumXcd), and the blocks of an XCD stride through
// its range. The grid is a multiple of kNumXcd.
struct XcdRange {
    uint32_t j, end, step;
    __device__ __forceinline__ XcdRange(uint32_t total) {
        const uint32_t x = blockIdx.x % kNumXcd, y = blockIdx.x / kNumXcd, nb = gridDim.x / kNumXcd;
        const uint32_t per = (total + kNumXcd - 1) / kNumXcd;
        j = x * per + y * kBlock + threadIdx.x;
        end = min(total, (x + 1) * per);
        step = nb * kBlock;
    }
};

#ifndef PRT_PAIR_WAVES
#define PRT_PAIR_WAVES 4  // waves per SIMD of k_bdpt_pair_connect (4: 127 VGPRs, no spills)
#endif
template <bool MIXED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(PRT_PAIR_WAVES))) void k_bdpt_pair_connect(
    DevScene sc, PassParams pp, VertexStore CP, VertexStore LP, PairBuffers pq, uint2* gstack, uint32_t gstride) {
    __shared__ uint32_t s_node[kLds * kBlock];
    __shared__ float s_t[kLds * kBlock];
    __shared__ float4 s_cache[kNodeCache * 8];
    TStack<kLds, kBlock> st;
    st.s_node = s_node;
    st.s_t = s_t;
    st.g = gstack;
    st.lane = threadIdx.x;
    st.gslot = blockIdx.x * kBlock + threadIdx.x;
    st.gstride = gstride;
    load_node_cache<kNodeCache, kBlock>(sc, s_cache);
    st.ncache = s_cache;
    const DevCamera& cam = sc.cam;
    for (XcdRange r(*(volatile uint32_t*)pq.count); r.j < r.end; r.j += r.step) {
        const uint32_t item = pq.items[kKindGeneral][r.j];
        const uint32_t i = item & ((1u << kItemSlotBits) - 1u), p = item >> kItemSlotBits;
        const int s = (int)(p / pq.nt) + 2, t = (int)(p % pq.nt) + 2;
        const BV ls1 = LP.get(s - 1, i), ct1 = CP.get(t - 1, i);
        float4 res = make_float4(0.f, 0.f, 0.f, 0.f);
        V3 dd = ls1.p - ct1.p;
        const float dist = length(dd);
        if (dist >= 0.001f) {
            dd = dd * (1.0f / dist);
            const float cc = dot(dd, ct1.n), cl = dot(dd * -1.0f, ls1.n);
            if (cc > 0.0f && cl > 0.0f) {
                const float G = (cc * cl) / (dist * dist);
                const V3 fc = vertex_brdf(sc, ct1, dd, 0);
                const V3 fl = vertex_brdf(sc, ls1, dd * -1.0f, 1);
                const V3 contrib = mulv(mulv(mulv(ls1.beta, fl), fc), ct1.beta) * G;
                if (contrib.x != 0.0f || contrib.y != 0.0f || contrib.z != 0.0f) {
                    const BV ls2 = LP.get(s - 2, i), ct2 = CP.get(t - 2, i);
                    const float l_tail = ((ls2.flags & VF_CONN) ? 1.0f : 0.0f) + ls2.q, l_fwd2 = remap0(ls2.pdf_fwd);
                    const Seg c = seg_of(ls2.p - ls1.p);
                    const V3 lw = contrib * mis_general(sc, cam, ct1, ct2, ls1, ls2, c, l_tail, l_fwd2, t);
                    const TRay vr{ct1.p.x, ct1.p.y, ct1.p.z, 0.001f, dd.x, dd.y, dd.z, dist - 0.001f};
                    THit vh;
                    if (!trace<true, MIXED, false, kLds, kBlock, kNodeCache>(sc, vr, vh, st, nullptr, pp.trace_gate))
                        res = make_float4(lw.x, lw.y, lw.z, 1.0f);
                }
            }
        }
        pq.plane(p)[i] = res;
    }
}

// Direct-lighting (KIND = kKindDirect), light-tracing (kKindLightTrace) or (0, nc) (kKindEmission) strategies,
// one lane per listed sample; a kernel per kind, each compiled for its branch of the strategy.
#ifndef PRT_EDGE_WAVES
#define PRT_EDGE_WAVES 4  // waves per SIMD of k_bdpt_edge_connect
#endif
template <bool MIXED, int KIND>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(PRT_EDGE_WAVES))) void k_bdpt_edge_connect(
    DevScene sc, PassParams pp, VertexStore CP, VertexStore LP, const uint32_t* lens, long long* splat,
    PairBuffers pq, uint2* gstack, uint32_t gstride) {
    __shared__ uint32_t s_node[kLds * kBlock];
    __shared__ float s_t[kLds * kBlock];
    __shared__ float4 s_cache[kNodeCache * 8];
    TStack<kLds, kBlock> st;
    st.s_node = s_node;
    st.s_t = s_t;
    st.g = gstack;
    st.lane = threadIdx.x;
    st.gslot = blockIdx.x * kBlock + threadIdx.x;
    st.gstride = gstride;
    load_node_cache<kNodeCache, kBlock>(sc, s_cache);
    st.ncache = s_cache;
    const uint32_t* items = pq.items[KIND];
    for (XcdRange r(*(volatile uint32_t*)(pq.count + KIND * kCounterStride)); r.j < r.end; r.j += r.step) {
        const uint32_t item = items[r.j];
        const uint32_t i = item & ((1u << kItemSlotBits) - 1u), e = item >> kItemSlotBits;
        int s = 0, t = 0;
        uint32_t plane = 0;
        if constexpr (KIND == kKindDirect) {
            s = 1;
            t = (int)e + 2;
            plane = pq.np + e;
        } else if constexpr (KIND == kKindLightTrace) {
            s = (int)e + 2;
            t = 1;
        } else {
            s = 0;
            t = (int)(lens[i] & 0xFFFFu);
            plane = pq.np + pq.nt;
        }
        uint32_t pix, sample;
        slot_sample(pp, i, pix, sample);
        const uint32_t key = path_key(pp.seed, pix, sample);
        V3 c;
        const bool add = eval_strategy<MIXED, KIND>(sc, pp, CP, LP, i, key, s, t, splat, st, c);
        if constexpr (KIND != kKindLightTrace)
            pq.plane(plane)[i] = add ? make_float4(c.x, c.y, c.z, 1.0f) : make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

__global__ __launch_bounds__(kBlock) void k_bdpt_pair_sum(PassParams pp, PathBuffers pb, PairBuffers pq) {
    const uint32_t n = (uint32_t)pp.npix * (uint32_t)pp.spp;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint2 m = pq.mask[i];
    const uint32_t lm = m.x & 0xFFFFu, cm = m.x >> 16, dm = m.y & 0xFFFFu;
    V3 total = v3(0, 0, 0);
    if (m.y >> 16) {
        const float4 r = pq.plane(pq.np + pq.nt)[i];
        if (r.w != 0.0f) total = total + v3(r.x, r.y, r.z);
    }
    for (uint32_t t = 0; (dm >> t) != 0u; t++) {
        if (!((dm >> t) & 1u)) continue;
        const float4 r = pq.plane(pq.np + t)[i];
        if (r.w != 0.0f) total = total + v3(r.x, r.y, r.z);
    }
    if (cm)
        for (uint32_t s = 0; (lm >> s) != 0u; s++) {
            if (!((lm >> s) & 1u)) continue;
            for (uint32_t t = 0; (cm >> t) != 0u; t++) {
                if (!((cm >> t) & 1u)) continue;
                const float4 r = pq.plane(s * pq.nt + t)[i];
                if (r.w != 0.0f) total = total + v3(r.x, r.y, r.z);
            }
        }
    pb.L[i] = make_float4(total.x, total.y, total.z, 0.f);
}

// Film (PixelStats.AddSample, stats.go:41-47): per pixel, each sample is added to the running sums with
// Kahan compensation, in sample order.
struct FilmBuffers {
    float4* sum;   // rgb sum, luminance sum
    float4* comp;  // Kahan compensation of sum
    float2* sq;    // luminance^2 sum, compensation
    uint32_t* count;
    long long* splat;  // BDPT light-tracing splats, 3 per pixel, fixed point (kSplatScale units)
};

__device__ __forceinline__ void kahan(float& s, float& c, float v) {
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

// vec3ToColor of GetColor (progressive.go:482-495, stats.go:36-55) for the pixels of a rectangle, in the
// operations and order of the host mean (film_mean): (sum + splats) * (1 / count), then uint8(255 *
// min(1, sqrt(mean))) in double with sqrt correctly rounded; NaN and negative means give 0.
__global__ __launch_bounds__(kBlock) void k_film_rgba8(int32_t width, int32_t x0, int32_t y0, int32_t tw, int64_t n,
                                                       const float4* sum, const uint32_t* count,
                                                       const long long* splat, uchar4* out) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const int64_t i = (int64_t)(y0 + (int32_t)(k / tw)) * width + (x0 + (int32_t)(k % tw));
    const uint32_t c = count[i];
    const float inv = c ? 1.0f / (float)c : 0.0f;
    const float4 S = sum[i];
    uint8_t o[3];
    for (int ch = 0; ch < 3; ch++) {
        float v = ch == 0 ? S.x : (ch == 1 ? S.y : S.z);
        if (splat) v += (float)((double)splat[3 * i + ch] * (1.0 / (double)kSplatScale));
        const float mean = c ? v * inv : 0.0f;
        double g = sqrt((double)mean);
        g = g > 0.0 ? fmin(1.0, g) : 0.0;
        o[ch] = (uint8_t)(255.0 * g);
    }
    out[k] = make_uchar4(o[0], o[1], o[2], 255);
}

__global__ __launch_bounds__(kBlock) void k_film(PassParams pp, PathBuffers pb, FilmBuffers fb) {
    int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= pp.npix) return;
    // every sample is added to the running sums with its own Kahan step, in sample order: the film depends
    // only on the pixel's sequence of samples, not on how a pass's samples are grouped into chunks
    const int px = pp.px0 + p;
    float4 S = fb.sum[px], C = fb.comp[px];
    float2 Q = fb.sq[px];
    for (int s = 0; s < pp.spp; s++) {
        const float4 v = pb.L[(size_t)s * pp.npix + p];
        const float lv = 0.2126f * v.x + 0.7152f * v.y + 0.0722f * v.z;
        kahan(S.x, C.x, v.x);
        kahan(S.y, C.y, v.y);
        kahan(S.z, C.z, v.z);
        kahan(S.w, C.w, lv);
        kahan(Q.x, Q.y, lv * lv);
    }
    fb.sum[px] = S;
    fb.comp[px] = C;
    fb.sq[px] = Q;
    fb.count[px] += (uint32_t)pp.spp;
}

// ---- adaptive sampling (adaptiveSamplePixelWithSplats / shouldStopSampling, tile_renderer.go:50-101) ----
// A pass runs in rounds over a list of active pixels. Each round gives every active pixel k samples with
// sample indices count, count+1, ...; the film kernel replays the reference's per-sample loop: before each
// sample it applies the stop test and discards the pixel's remaining samples of the round once it stops.
// A sample's value depends only on (pixel, sample index), so the result equals the sequential loop.
struct AdaptiveParams {
    int32_t target;       // samples per pixel of the pass
    int32_t min_samples;  // max(1, int(target * AdaptiveMinSamples))
    double threshold;     // AdaptiveThreshold
    uint32_t* next;       // active list of the next round
    uint32_t* state;      // [0]: next list length, [1]: min count, [2]: max count of the next list
};

// shouldStopSampling (tile_renderer.go:77-101) on the Kahan-compensated float32 moments, in float64.
__device__ __forceinline__ bool adaptive_stop(const AdaptiveParams& ap, uint32_t cnt, double ls, double lq) {
    if ((int32_t)cnt < ap.min_samples) return false;
    const double mean = ls / (double)cnt, mean_sq = lq / (double)cnt;
    const double var = fmax(0.0, mean_sq - mean * mean);
    if (mean <= 1e-8) return var < 1e-6;
    return sqrt(var) / mean < ap.threshold;
}

__device__ __forceinline__ void adaptive_keep(const AdaptiveParams& ap, uint32_t px, uint32_t cnt, bool keep) {
    uint64_t mask = __ballot(keep);
    if (mask == 0) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&ap.state[0], (uint32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
    if (keep) {
        ap.next[base + __popcll(mask & ((1ull << lane) - 1ull))] = px;
        atomicMin(&ap.state[1], cnt);
        atomicMax(&ap.state[2], cnt);
    }
}

// Start of a pass: every pixel that has not reached the target and is not stopped is active.
__global__ __launch_bounds__(kBlock) void k_adaptive_select(int32_t npix, FilmBuffers fb, AdaptiveParams ap) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    bool keep = false;
    uint32_t cnt = 0;
    if (p < npix) {
        cnt = fb.count[p];
        const float4 S = fb.sum[p], C = fb.comp[p];
        const float2 Q = fb.sq[p];
        keep = (int32_t)cnt < ap.target &&
               !adaptive_stop(ap, cnt, (double)S.w - (double)C.w, (double)Q.x - (double)Q.y);
    }
    adaptive_keep(ap, (uint32_t)p, cnt, keep);
}

// Per active pixel of the chunk: the reference loop over the round's k samples, each accepted sample
// added to the film with Kahan compensation; pixels still active join the next round's list.
__global__ __launch_bounds__(kBlock) void k_film_adaptive(PassParams pp, PathBuffers pb, FilmBuffers fb,
                                                          AdaptiveParams ap) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    bool keep = false;
    uint32_t px = 0, cnt = 0;
    if (p < pp.npix) {
        px = pp.act[pp.px0 + p];
        cnt = fb.count[px];
        float4 S = fb.sum[px], C = fb.comp[px];
        float2 Q = fb.sq[px];
        bool stopped = false;
        for (int s = 0; s < pp.spp; s++) {
            if ((int32_t)cnt >= ap.target ||
                adaptive_stop(ap, cnt, (double)S.w - (double)C.w, (double)Q.x - (double)Q.y)) {
                stopped = true;
                break;
            }
            const float4 v = pb.L[(size_t)s * pp.npix + p];
            const float lv = 0.2126f * v.x + 0.7152f * v.y + 0.0722f * v.z;
            kahan(S.x, C.x, v.x);
            kahan(S.y, C.y, v.y);
            kahan(S.z, C.z, v.z);
            kahan(S.w, C.w, lv);
            kahan(Q.x, Q.y, lv * lv);
            cnt++;
        }
        if (!stopped)
            stopped = (int32_t)cnt >= ap.target ||
                      adaptive_stop(ap, cnt, (double)S.w - (double)C.w, (double)Q.x - (double)Q.y);
        fb.sum[px] = S;
        fb.comp[px] = C;
        fb.sq[px] = Q;
        fb.count[px] = cnt;
        keep = !stopped;
    }
    adaptive_keep(ap, px, cnt, keep);
}

}  // namespace prt

using namespace prt;

struct prt_film {
    const prt_scene* scene = nullptr;
    DevScene sc{};  // the scene view with this film's camera
    prt_render_params params{};
    int32_t width = 0, height = 0;
    int64_t npix = 0;
    int64_t capacity = 0;  // path slots
    int32_t spp_done = 0;  // samples per pixel enqueued so far
    std::vector<DevBuf> bufs;  // sizes as allocated (a pooled buffer may be larger than asked for)
    size_t bytes = 0;          // bytes asked for
    FilmPool* pool = nullptr;  // where the buffers and stream came from and go back to (nullptr: own)
    PathBuffers pb{};
    FilmBuffers fb{};
    uint2* gstack = nullptr;
    uint32_t gslots = 0;
    int blocks = 0;
    int trace_grid[4] = {0, 0, 0, 0};  // persistent grid of each trace-kernel variant
    size_t trace_pad[4] = {0, 0, 0, 0};  // its dynamic LDS
    uint32_t park_steps = 0;  // step budget of the path tracer's traversals before a ray is parked (0: none)
    uint32_t park_steps_small = 0;  // the budget of launches with few rays per lane of the trace grid
    uint32_t park_small_rays_per_lane = 1;  // launches of at most this many rays per lane are "few"
    int64_t tail_paths = 0;  // the tail kernel finishes a chunk from the first bounce expected to hold this few paths
    int tail_blocks = 0;     // its grid
    int long_grid = 0;        // work-groups (one wave each) of k_trace_long
    int coop_lanes = 64;      // its lanes per parked ray
    // BDPT: camera and light subpath stores and subpath lengths, per path slot
    bool bdpt = false;
    int bdpt_blocks = 0;
    int bdpt_general_blocks = 0;  // grid of k_bdpt_connect_general (its own occupancy)
    bool bdpt_pairs = false;      // strategies through per-strategy queues
    int bdpt_pair_blocks = 0;     // grid of k_bdpt_pair_connect
    int bdpt_edge_blocks[kNumLists] = {0, 0, 0, 0};  // grids of k_bdpt_edge_connect per strategy kind
    PairBuffers pairs{};
    VertexStore cpath{}, lpath{};
    uint32_t* lens = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // serving films (film_track_chunks): each chunk bracketed by a pair of ring events, so that the
    // scheduler sees chunks finish one by one; at most kRing chunks may be in flight
    static constexpr int kRing = 8;
    bool track = false;
    hipEvent_t rev0[kRing] = {}, rev1[kRing] = {};
    int64_t enq = 0, seen = 0;
    int64_t chunk_paths = 0;  // paths per planned chunk (at most capacity)
    // RGBA8 reads (prt_film_read_rgba8): device image allocated with the film, shared by the film's readers
    // under rgba_mu (a served film may be read from several threads); host staging from a process-wide pool
    std::mutex rgba_mu;
    uchar4* rgba_dev = nullptr;
    // chunk plan of the samples being rendered (see film_plan)
    struct Chunk {
        int32_t px0, npix, spp, sample_base;
    };
    std::vector<Chunk> plan;
    size_t plan_pos = 0;
    double plan_ms = 0;
    // adaptive passes (AdaptiveMinSamples < 1): rounds over an active-pixel list
    bool adaptive = false;
    bool ad_running = false;   // a pass is in progress
    bool ad_selected = false;  // its first active list has been built
    int32_t ad_target = 0, ad_min = 0;
    int ad_cur = 0;            // index of the current list
    uint32_t* act[2] = {nullptr, nullptr};
    uint32_t* ad_state = nullptr;  // per list: length, min count, max count
    uint32_t ad_n = 0;             // length of the current list
    // measurement probe (prt_film_probe): counting trace kernels and per-kernel-kind event timing
    uint32_t probe_flags = 0;
    unsigned long long* probe_dev = nullptr;  // traversal counts of the counting trace kernel
    uint64_t probe_launches = 0;              // trace launches enqueued while probing
    struct Mark {
        int kind;  // PRT_PROBE_KIND_*
        hipEvent_t a, b;
    };
    std::vector<Mark> marks;
};

namespace prt {

int FilmPool::take(size_t bytes, void** p, size_t* got) {
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = bufs.lower_bound(bytes);
        if (it != bufs.end() && it->first <= bytes + bytes / 4) {
            *p = it->second;
            *got = it->first;
            cached -= it->first;
            bufs.erase(it);
            return PRT_OK;
        }
    }
    PRT_HIP(hipSetDevice(device));
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        trim(0);
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    *got = bytes;
    return PRT_OK;
}

void FilmPool::give(void* p, size_t bytes) {
    std::lock_guard<std::mutex> lk(mu);
    bufs.emplace(bytes, p);
    cached += bytes;
}

size_t FilmPool::trim(size_t keep) {
    std::lock_guard<std::mutex> lk(mu);
    size_t freed = 0;
    (void)hipSetDevice(device);
    while (cached > keep && !bufs.empty()) {
        auto it = std::prev(bufs.end());
        (void)hipFree(it->second);
        cached -= it->first;
        freed += it->first;
        bufs.erase(it);
    }
    return freed;
}

int FilmPool::take_stream(hipStream_t* st) {
    {
        std::lock_guard<std::mutex> lk(mu);
        if (!streams.empty()) {
            *st = streams.back();
            streams.pop_back();
            return PRT_OK;
        }
    }
    PRT_HIP(hipSetDevice(device));
    PRT_HIP(hipStreamCreateWithFlags(st, hipStreamNonBlocking));
    return PRT_OK;
}

void FilmPool::give_stream(hipStream_t st) {
    std::lock_guard<std::mutex> lk(mu);
    streams.push_back(st);
}

size_t FilmPool::idle_bytes() {
    std::lock_guard<std::mutex> lk(mu);
    return cached;
}

void FilmPool::clear() {
    trim(0);
    std::lock_guard<std::mutex> lk(mu);
    for (hipStream_t st : streams) (void)hipStreamDestroy(st);
    streams.clear();
}

// GetColor: (samples + splats) / SampleCount, zero when no samples (stats.go:36-55)
void film_mean(int64_t npix, const float4* sum, const uint32_t* cnt, const long long* spl, float* mean_rgb) {
    for (int64_t i = 0; i < npix; i++) {
        float inv = cnt[i] ? 1.0f / (float)cnt[i] : 0.0f;
        for (int c = 0; c < 3; c++) {
            float v = c == 0 ? sum[i].x : (c == 1 ? sum[i].y : sum[i].z);
            if (spl) v += (float)((double)spl[3 * i + c] * (1.0 / (double)kSplatScale));
            mean_rgb[3 * i + c] = cnt[i] ? v * inv : 0.f;
        }
    }
}

// Film snapshots: device copies of the accumulators (sums, counts, splats) taken on the film's stream,
// so that they hold the film as it was when the work enqueued before them finished.
size_t film_snapshot_bytes(const prt_film* f) {
    return (size_t)f->npix * (sizeof(float4) + sizeof(uint32_t) + (f->fb.splat ? 3 * sizeof(long long) : 0));
}
int film_snapshot(prt_film* f, void* dst, hipEvent_t done) {
    char* p = (char*)dst;
    PRT_HIP(hipSetDevice(f->scene->device));
    PRT_HIP(hipMemcpyAsync(p, f->fb.sum, sizeof(float4) * f->npix, hipMemcpyDeviceToDevice, f->stream));
    p += sizeof(float4) * f->npix;
    PRT_HIP(hipMemcpyAsync(p, f->fb.count, sizeof(uint32_t) * f->npix, hipMemcpyDeviceToDevice, f->stream));
    p += sizeof(uint32_t) * f->npix;
    if (f->fb.splat)
        PRT_HIP(hipMemcpyAsync(p, f->fb.splat, 3 * sizeof(long long) * f->npix, hipMemcpyDeviceToDevice, f->stream));
    PRT_HIP(hipEventRecord(done, f->stream));
    return PRT_OK;
}
// Reads a snapshot (after `done`) through `stream`: mean RGB as prt_film_read gives it, and counts.
int film_read_snapshot(const prt_film* f, const void* src, hipEvent_t done, hipStream_t stream, float* mean_rgb,
                       uint32_t* counts) {
    PRT_HIP(hipSetDevice(f->scene->device));
    std::vector<float4> sum(f->npix);
    std::vector<uint32_t> cnt(f->npix);
    std::vector<long long> spl(f->fb.splat ? 3 * f->npix : 0);
    const char* p = (const char*)src;
    PRT_HIP(hipStreamWaitEvent(stream, done, 0));
    PRT_HIP(hipMemcpyAsync(sum.data(), p, sizeof(float4) * f->npix, hipMemcpyDeviceToHost, stream));
    p += sizeof(float4) * f->npix;
    PRT_HIP(hipMemcpyAsync(cnt.data(), p, sizeof(uint32_t) * f->npix, hipMemcpyDeviceToHost, stream));
    p += sizeof(uint32_t) * f->npix;
    if (f->fb.splat)
        PRT_HIP(hipMemcpyAsync(spl.data(), p, 3 * sizeof(long long) * f->npix, hipMemcpyDeviceToHost, stream));
    PRT_HIP(hipStreamSynchronize(stream));
    if (mean_rgb) film_mean(f->npix, sum.data(), cnt.data(), f->fb.splat ? spl.data() : nullptr, mean_rgb);
    if (counts) std::memcpy(counts, cnt.data(), sizeof(uint32_t) * f->npix);
    return PRT_OK;
}


namespace {

// Pinned host staging of RGBA8 reads, shared by all films of the process: pinning pages costs
// milliseconds, so buffers are kept for reuse; each read holds one buffer while it copies.
std::mutex g_pinned_mu;
std::vector<std::pair<size_t, void*>> g_pinned_free;  // (capacity, buffer) not in use
std::map<void*, size_t> g_pinned_cap;                  // capacity of every buffer

int pinned_take(size_t bytes, void** out) {
    {
        std::lock_guard<std::mutex> lock(g_pinned_mu);
        size_t best = g_pinned_free.size();
        for (size_t i = 0; i < g_pinned_free.size(); i++)
            if (g_pinned_free[i].first >= bytes && (best == g_pinned_free.size() || g_pinned_free[i].first <
                                                                                          g_pinned_free[best].first))
                best = i;
        if (best < g_pinned_free.size()) {
            *out = g_pinned_free[best].second;
            g_pinned_free.erase(g_pinned_free.begin() + (ptrdiff_t)best);
            return PRT_OK;
        }
    }
    // round up to 1 MiB so that reads of nearby sizes share buffers
    const size_t cap = (bytes + ((size_t)1 << 20) - 1) & ~(((size_t)1 << 20) - 1);
    hipError_t e = hipHostMalloc(out, cap, hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(e, "hipHostMalloc(RGBA8 staging)");
    std::lock_guard<std::mutex> lock(g_pinned_mu);
    g_pinned_cap[*out] = cap;
    return PRT_OK;
}

void pinned_give(void* p, size_t) {
    std::lock_guard<std::mutex> lock(g_pinned_mu);
    g_pinned_free.push_back({g_pinned_cap[p], p});
}

int film_alloc(prt_film* f, size_t bytes, void** out) {
    void* p = nullptr;
    bytes = std::max<size_t>(bytes, 16);
    size_t got = bytes;
    if (f->pool) {
        int rc = f->pool->take(bytes, &p, &got);
        if (rc != PRT_OK) return rc;
    } else {
        PRT_HIP(hipMalloc(&p, bytes));
    }
    f->bufs.push_back({p, got});
    f->bytes += bytes;
    *out = p;
    return PRT_OK;
}


constexpr size_t kPathBytes = 7 * sizeof(float4) + 5 * sizeof(uint32_t);
constexpr size_t kPixelBytes = 2 * sizeof(float4) + sizeof(float2) + sizeof(uint32_t);

int64_t film_capacity(int64_t npix, int64_t cap) {
    return std::max<int64_t>(std::min<int64_t>(npix * 64, cap), std::min<int64_t>(npix, cap));
}

template <bool MIXED, bool COUNT>
const void* trace_fn_of(int v) {
    if (kTraceKernelWaves[v] == 8) return (const void*)k_trace<MIXED, COUNT, 8>;
    return (const void*)k_trace<MIXED, COUNT, 7>;
}
const void* trace_fn(bool mixed, bool count, int v) {
    if (mixed) return count ? trace_fn_of<true, true>(v) : trace_fn_of<true, false>(v);
    return count ? trace_fn_of<false, true>(v) : trace_fn_of<false, false>(v);
}

// Dynamic LDS (unused by the kernel) that caps a variant at kTraceGroupCap work-groups per CU.
size_t trace_lds_pad(int device, int v) {
    if (!kTraceGroupCap[v]) return 0;
    int cu_lds = 0;
    if (hipDeviceGetAttribute(&cu_lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) != hipSuccess ||
        cu_lds <= 0)
        cu_lds = 160 * 1024;
    hipFuncAttributes fa{};
    if (hipFuncGetAttributes(&fa, trace_fn(false, false, v)) != hipSuccess) return 0;
    const size_t per_group = ((size_t)cu_lds / (size_t)kTraceGroupCap[v]) & ~(size_t)511;
    return per_group > fa.sharedSizeBytes ? per_group - fa.sharedSizeBytes : 0;
}

// Persistent grid of a trace variant: the work-groups it keeps resident (at most 8 per CU).
int trace_grid(const prt_scene* scene, int v) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, trace_fn(scene->dev.mixed != 0, false, v), kBlock,
                                                     trace_lds_pad(scene->device, v)) != hipSuccess ||
        per_cu <= 0)
        per_cu = 4;
    return std::max(1, scene->num_cus) * std::min(per_cu, 8);
}

int film_blocks(const prt_scene* scene) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_trace<true>, kBlock, 0) !=
            hipSuccess ||
        per_cu <= 0)
        per_cu = 4;
    per_cu = std::min(per_cu, 8);
    return std::max(1, scene->num_cus) * per_cu;
}

// Long-ray parking of the path tracer: scenes of at least 2^16 primitives (whose deep bounces end in
// grazing rays along a dense mesh) park rays after 128 steps in launches that fill the grid, after 64 in
// smaller ones; PRT_PT_PARK_STEPS and PRT_PT_PARK_STEPS_SMALL override the budgets (0: no parking).
constexpr int64_t kLongCapacity = (int64_t)1 << 20;
// Lanes per parked ray: a whole wave (C3 first pass 4.1-4.4 ms against 4.5-4.7 ms with 16 lanes and 4 rays
// per wave: a launch parks too few rays for the groups' throughput to outweigh their longer latency).
constexpr int kDefaultCoopLanes = 64;
uint32_t park_steps(const prt_scene* scene, bool small) {
    if (const char* e = std::getenv(small ? "PRT_PT_PARK_STEPS_SMALL" : "PRT_PT_PARK_STEPS")) {
        const long v = std::atol(e);
        if (v >= 0 && v < (1l << 30)) return (uint32_t)v;
    }
    if (scene->info.num_primitives < (1 << 16)) return 0u;
    return small ? 64u : 128u;
}

uint32_t park_small_rays_per_lane() {
    if (const char* e = std::getenv("PRT_PT_PARK_SMALL_RAYS_PER_LANE")) {
        const long v = std::atol(e);
        if (v >= 1 && v < (1l << 20)) return (uint32_t)v;
    }
    return 4u;
}

// Resident work-groups of k_tail (at most the trace grid, whose spill-stack slots it uses).
int tail_grid(const prt_scene* scene, int trace_blocks) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_tail<true>, kBlock, 0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 4;
    return std::max(1, std::min(trace_blocks, std::max(1, scene->num_cus) * per_cu));
}

// Paths below which the last bounces of a chunk run as one launch (k_tail): the threads k_tail keeps
// resident (C0's first pass 1.4 -> 0.9 ms and 64-spp render 1.52 -> 1.79 G samples/s at 2^18 on 256 CUs;
// 2^16 gave 0.9 ms and 1.75 G, 2^20 0.85 ms and 1.32 G; C3 with parking 5.3 -> 4.1-4.4 ms and
// 783 -> 789 M samples/s). PRT_PT_TAIL_PATHS overrides (0: never).
int64_t tail_paths(int tail_blocks) {
    if (const char* e = std::getenv("PRT_PT_TAIL_PATHS")) {
        const long long v = std::atoll(e);
        if (v >= 0) return (int64_t)v;
    }
    return (int64_t)tail_blocks * kBlock;
}

// k_trace_long with groups of `lanes` lanes per parked ray.
const void* long_fn(bool mixed, int lanes) {
    if (lanes == 16) return mixed ? (const void*)k_trace_long<true, 16> : (const void*)k_trace_long<false, 16>;
    if (lanes == 32) return mixed ? (const void*)k_trace_long<true, 32> : (const void*)k_trace_long<false, 32>;
    return mixed ? (const void*)k_trace_long<true, 64> : (const void*)k_trace_long<false, 64>;
}

// Lanes per parked ray of k_trace_long (16, 32 or 64); PRT_COOP_LANES overrides.
int coop_lanes() {
    if (const char* e = std::getenv("PRT_COOP_LANES")) {
        const int v = std::atoi(e);
        if (v == 16 || v == 32 || v == 64) return v;
    }
    return kDefaultCoopLanes;
}

// Resident one-wave work-groups of k_trace_long.
int long_grid(const prt_scene* scene, int lanes) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, long_fn(true, lanes), 64, 0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 16;
    return std::max(1, scene->num_cus) * std::min(per_cu, 32);
}

size_t film_gstack_entries(const prt_scene* scene) {
    return scene->max_stack > kLds ? (size_t)(scene->max_stack - kLds + 1) : 1;
}

// Grid of the BDPT kernels: resident blocks only (each thread loops over samples), at most the
// trace grid so the traversal spill stack covers it.
int bdpt_blocks(const prt_scene* scene, int trace_blocks) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_bdpt_connect<true>, kBlock, 0) !=
            hipSuccess ||
        per_cu <= 0)
        per_cu = 2;
    return std::max(1, std::min(trace_blocks, std::max(1, scene->num_cus) * per_cu * 2));
}

// Grid of the general-connection kernel: its resident blocks twice, at most the trace grid (stack slots).
int bdpt_general_blocks(const prt_scene* scene, int trace_blocks) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_bdpt_connect_general<true>, kBlock,
                                                     0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 2;
    return std::max(1, std::min(trace_blocks, std::max(1, scene->num_cus) * per_cu * 2));
}

// Bytes of one subpath store: (max_depth + 1) vertices of 5 float4 (6 with UVs) per path slot.
size_t bdpt_store_bytes(int max_depth, int64_t slots, bool uv) {
    return (size_t)(max_depth + 1) * (uv ? 6 : 5) * sizeof(float4) * (size_t)slots;
}

// Whether a BDPT film of this depth evaluates its strategies through per-strategy queues (k_bdpt_pairs,
// k_bdpt_edge_connect, k_bdpt_pair_connect, k_bdpt_pair_sum): when its (max_depth + 2) * max_depth queues fit
// kMaxQueues and its vertex masks 16 bits (max_depth <= 16); deeper films loop over the strategies per sample
// (k_bdpt_connect, k_bdpt_connect_general). PRT_BDPT_PAIR_QUEUES=0 selects the per-sample kernels at every
// depth.
bool bdpt_pairs_used(int max_depth) {
    if (const char* e = std::getenv("PRT_BDPT_PAIR_QUEUES"))
        if (std::atoi(e) == 0) return false;
    return max_depth >= 2 && max_depth <= 16;
}

// Bytes per path slot of the strategy lists, result planes and masks.
size_t bdpt_pair_slot_bytes(int max_depth) {
    if (!bdpt_pairs_used(max_depth)) return 0;
    const size_t nt = (size_t)max_depth, np = (nt - 1) * nt;
    return (np + 2 * nt) * sizeof(uint32_t) + (np + nt + 1) * sizeof(float4) + sizeof(uint2);
}

// Bytes of the strategy lists, result planes and masks of `slots` path slots, and the list lengths.
size_t bdpt_pair_bytes(int max_depth, int64_t slots) {
    if (!bdpt_pairs_used(max_depth)) return 0;
    return bdpt_pair_slot_bytes(max_depth) * (size_t)slots + kNumLists * kCounterStride * sizeof(uint32_t);
}

// Path slots of a BDPT film: the two subpath stores stay within 2 GiB, and with the pair queues within 4 GiB.
int64_t bdpt_capacity(int max_depth, int64_t cap, bool uv) {
    const int64_t per_slot = (int64_t)(2 * bdpt_store_bytes(max_depth, 1, uv));
    int64_t c = ((int64_t)1 << 31) / per_slot;
    if (bdpt_pairs_used(max_depth)) {
        c = std::min<int64_t>(c, ((int64_t)1 << 32) / (per_slot + (int64_t)bdpt_pair_slot_bytes(max_depth)));
        c = std::min<int64_t>(c, (int64_t)1 << kItemSlotBits);  // slot ids of the strategy lists
    }
    if (const char* e = std::getenv("PRT_BDPT_SLOTS")) {  // pins the path slots (chunk size) of BDPT films
        const long long v = std::atoll(e);
        if (v > 0) c = std::min<int64_t>(c, (int64_t)v);
    }
    c = std::max<int64_t>(c, (int64_t)1 << 16) & ~(int64_t)255;
    return std::min(cap, c);
}

// Grids of k_bdpt_pair_connect and k_bdpt_edge_connect: their resident blocks twice, at most the trace grid
// (stack slots), a multiple of kNumXcd.
int bdpt_queue_blocks(const prt_scene* scene, int trace_blocks, const void* fn) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, 0) != hipSuccess || per_cu <= 0) per_cu = 2;
    const int g = std::min(trace_blocks, std::max(1, scene->num_cus) * per_cu * 2);
    return std::max((int)kNumXcd, g / (int)kNumXcd * (int)kNumXcd);
}

}  // namespace

// Splits `spp` more samples per pixel into chunks of at most `capacity` paths. The split depends
// only on the film's size and sample count, so a render is reproducible however its chunks are
// scheduled.
int film_plan(prt_film* f, int32_t spp) {
    f->plan.clear();
    f->plan_pos = 0;
    f->plan_ms = 0;
    int64_t done = 0;
    while (done < spp) {
        int64_t per = f->chunk_paths / f->npix;
        int32_t chunk_spp = (int32_t)std::max<int64_t>(1, std::min<int64_t>(per, spp - done));
        int64_t pix_step = per >= 1 ? f->npix : f->chunk_paths;
        for (int64_t p0 = 0; p0 < f->npix; p0 += pix_step) {
            prt_film::Chunk c;
            c.px0 = (int32_t)p0;
            c.npix = (int32_t)std::min<int64_t>(pix_step, f->npix - p0);
            c.spp = chunk_spp;
            c.sample_base = f->spp_done + (int32_t)done;
            f->plan.push_back(c);
        }
        done += chunk_spp;
    }
    f->spp_done += spp;
    return PRT_OK;
}

bool film_round_exhausted(const prt_film* f) { return f->plan_pos >= f->plan.size(); }

bool film_plan_done(const prt_film* f) { return f->adaptive ? !f->ad_running : film_round_exhausted(f); }

namespace {
__global__ void k_adaptive_reset(uint32_t* st) {
    st[0] = 0;
    st[1] = 0xFFFFFFFFu;
    st[2] = 0;
}
}  // namespace

// Plans the pass that brings every pixel to `target` samples per pixel (RenderPass). Uniform films
// split the missing samples into chunks; adaptive films start a sequence of rounds (film_advance).
int film_plan_pass(prt_film* f, int32_t target) {
    if (!f->adaptive) {
        int32_t need = target - f->spp_done;
        return film_plan(f, need > 0 ? need : 0);
    }
    f->plan.clear();
    f->plan_pos = 0;
    f->plan_ms = 0;
    f->ad_running = true;
    f->ad_selected = false;
    f->ad_target = target;
    f->ad_min = std::max<int32_t>(1, (int32_t)((double)target * f->params.adaptive_min_samples));
    f->spp_done = std::max(f->spp_done, target);
    return PRT_OK;
}

// Adaptive films: when the enqueued rounds have finished (the film is idle), reads the active list the
// last round (or the pass start) left and plans the next round's chunks, or ends the pass. Each round
// gives every active pixel k samples: the mandatory samples up to the minimum count plus, for path
// tracing, an eighth of the pass's remaining budget that the film kernel may discard; BDPT takes no
// speculative samples (a discarded sample's splats would already be on the film).
int film_advance(prt_film* f) {
    if (!f->adaptive || !f->ad_running || !film_round_exhausted(f)) return PRT_OK;
    PRT_HIP(hipSetDevice(f->scene->device));
    hipStream_t st = f->stream;
    FilmBuffers fb = f->fb;
    AdaptiveParams ap;
    ap.target = f->ad_target;
    ap.min_samples = f->ad_min;
    ap.threshold = f->params.adaptive_threshold;
    if (!f->ad_selected) {
        f->ad_cur = 0;
        ap.next = f->act[0];
        ap.state = f->ad_state;
        hipLaunchKernelGGL(k_adaptive_reset, dim3(1), dim3(1), 0, st, ap.state);
        hipLaunchKernelGGL(k_adaptive_select, dim3((int)((f->npix + kBlock - 1) / kBlock)), dim3(kBlock), 0, st,
                           (int32_t)f->npix, fb, ap);
        PRT_HIP(hipGetLastError());
        f->ad_selected = true;
    } else {
        f->ad_cur ^= 1;  // the list the last round wrote
    }
    uint32_t h[3] = {0, 0, 0};
    PRT_HIP(hipMemcpyAsync(h, f->ad_state + 3 * f->ad_cur, sizeof(h), hipMemcpyDeviceToHost, st));
    PRT_HIP(hipStreamSynchronize(st));
    f->ad_n = h[0];
    f->plan.clear();
    f->plan_pos = 0;
    if (f->ad_n == 0) {
        f->ad_running = false;
        return PRT_OK;
    }
    const int32_t mincnt = (int32_t)h[1], maxcnt = (int32_t)h[2];
    int64_t k;
    if (f->bdpt) {
        k = maxcnt < f->ad_min ? f->ad_min - maxcnt : 1;
    } else {
        const int64_t spec = std::max<int64_t>(1, (f->ad_target - f->ad_min + 7) / 8);
        k = (mincnt < f->ad_min ? f->ad_min - mincnt : 0) + spec;
    }
    k = std::max<int64_t>(1, std::min<int64_t>({k, (int64_t)f->ad_target - mincnt, f->chunk_paths}));
    const int64_t per_chunk = std::max<int64_t>(1, f->chunk_paths / k);
    for (int64_t a0 = 0; a0 < (int64_t)f->ad_n; a0 += per_chunk) {
        prt_film::Chunk c;
        c.px0 = (int32_t)a0;
        c.npix = (int32_t)std::min<int64_t>(per_chunk, (int64_t)f->ad_n - a0);
        c.spp = (int32_t)k;
        c.sample_base = 0;
        f->plan.push_back(c);
    }
    // the next list starts empty
    hipLaunchKernelGGL(k_adaptive_reset, dim3(1), dim3(1), 0, st, f->ad_state + 3 * (f->ad_cur ^ 1));
    PRT_HIP(hipGetLastError());
    return PRT_OK;
}

// Per-kernel timing of prt_film_probe: an event pair around each launch of a kernel kind.
hipError_t probe_mark(prt_film* f, int kind, bool begin) {
    if (!(f->probe_flags & PRT_PROBE_TIME)) return hipSuccess;
    hipEvent_t e = nullptr;
    hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) return r;
    if (begin) {
        f->marks.push_back({kind, e, nullptr});
    } else {
        f->marks.back().b = e;
    }
    return hipEventRecord(e, f->stream);
}

// Starts a tuning job of `runs` timed launches on stream st: ev[0] is recorded here; the caller then
// enqueues one warm-up launch, records ev[1], and records ev[2 + k] after timed launch k.
static hipError_t tune_begin(TuneJob& j, int runs, hipStream_t st) {
    if (runs + 2 > (int)(sizeof(j.ev) / sizeof(j.ev[0]))) return hipErrorInvalidValue;
    for (int k = 0; k < runs + 2; k++)
        if (!j.ev[k]) {
            hipError_t e = hipEventCreate(&j.ev[k]);
            if (e != hipSuccess) return e;
        }
    j.runs = runs;
    return hipEventRecord(j.ev[0], st);
}

// Device time of each timed launch of a pending job, once its last event has completed; false while it
// runs or when no job is pending (a job whose events report an error is dropped).
static bool tune_collect(TuneJob& j, float* t) {
    if (!j.pending) return false;
    hipError_t q = hipEventQuery(j.ev[j.runs + 1]);
    if (q == hipErrorNotReady) return false;
    j.pending = false;
    if (q != hipSuccess) {
        (void)hipGetLastError();  // handled here: not an error of the caller's next launch
        return false;
    }
    for (int k = 0; k < j.runs; k++) {
        t[k] = 0.0f;
        if (hipEventElapsedTime(&t[k], j.ev[1 + k], j.ev[2 + k]) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
    }
    return true;
}

// The leaf gate of a finished tuning job whose timed launches ran every gate twice, interleaved
// (t[k] = gate kGateChoice[k % nch]): the fastest by the shorter of its two times; gate 0 unless
// another is more than 3 % faster.
static int pick_gate(const float* t, int nch) {
    float best = std::min(t[0], t[nch]);
    int bestg = 0;
    for (int g = 1; g < nch; g++) {
        const float tg = std::min(t[g], t[nch + g]) * 1.03f;
        if (tg < best) {
            best = tg;
            bestg = kGateChoice[g];
        }
    }
    return bestg;
}

// Reads a completed profile copy into the scene's survival fractions (caller holds the profile lock).
static void profile_collect(PathProfile& pf) {
    if (!pf.pending) return;
    const hipError_t q = hipEventQuery(pf.ev);
    if (q == hipErrorNotReady) return;
    pf.pending = false;
    if (q != hipSuccess) {
        (void)hipGetLastError();  // a failed copy is dropped; the next chunk records another
        return;
    }
    const double n0 = (double)pf.host[0];
    if (n0 <= 0.0) return;
    if ((int)pf.frac.size() < pf.pending_upto + 1) pf.frac.resize(pf.pending_upto + 1, -1.0);
    for (int b = 0; b <= pf.pending_upto; b++) pf.frac[b] = (double)pf.host[b] / n0;
}

// A finished leaf-gate tuning job of the scene sets its gate (PT: under the job's lock; BDPT: the caller
// holds it).
static void collect_pt_tuning(prt_scene* ms) {
    std::lock_guard<std::mutex> lk(ms->pt_tune.mu);
    float t[2 * kGateChoices];
    if (tune_collect(ms->pt_tune, t)) {
        const int bestg = pick_gate(t, gate_choices(ms->dev));
        int expect = -1;
        ms->pt_trace_gate.compare_exchange_strong(expect, bestg);
        ms->info.pt_trace_leaf_gate = ms->pt_trace_gate.load();
    }
}
static void collect_bdpt_tuning_locked(prt_scene* ms) {
    float t[2 * kGateChoices];
    if (tune_collect(ms->bdpt_tune, t)) {
        const int bestg = pick_gate(t, gate_choices(ms->dev));
        int expect = -1;
        ms->bdpt_trace_gate.compare_exchange_strong(expect, bestg);
        ms->info.bdpt_trace_leaf_gate = ms->bdpt_trace_gate.load();
    }
}

// The scene's tuning jobs and survival-profile copy are recorded on the stream of the film that started
// them, but belong to the scene, which outlives the film. Once a film's stream has drained they are
// complete: collect them before the stream can be destroyed, since querying an event whose stream no
// longer exists is an error.
static void collect_scene_events(prt_scene* ms) {
    collect_pt_tuning(ms);
    {
        std::lock_guard<std::mutex> lk(ms->bdpt_tune.mu);
        collect_bdpt_tuning_locked(ms);
    }
    std::lock_guard<std::mutex> lk(ms->pt_profile.mu);
    profile_collect(ms->pt_profile);
}

// First bounce b >= 1 at which a chunk of n paths is expected to trace at most `limit` paths (by the
// scene's survival profile), else max_depth (no tail).
static int tail_bounce(prt_scene* s, int64_t n, int64_t limit, int max_depth) {
    if (limit <= 0 || max_depth < 2) return max_depth;
    std::lock_guard<std::mutex> lk(s->pt_profile.mu);
    profile_collect(s->pt_profile);
    const std::vector<double>& fr = s->pt_profile.frac;
    for (int b = 1; b < max_depth && b < (int)fr.size(); b++)
        if (fr[b] >= 0.0 && (double)n * fr[b] <= (double)limit) return b;
    return max_depth;
}

// After a chunk: copies its paths per bounce (bounces 0 .. upto) to the scene's pinned buffer, unless a
// copy is pending.
static hipError_t profile_record(prt_scene* s, prt_film* f, int upto, hipStream_t st) {
    PathProfile& pf = s->pt_profile;
    std::lock_guard<std::mutex> lk(pf.mu);
    profile_collect(pf);
    if (pf.pending) return hipSuccess;
    if (pf.host_len < upto + 1) {
        if (pf.host) {
            hipError_t e = hipHostFree(pf.host);
            pf.host = nullptr;
            pf.host_len = 0;
            if (e != hipSuccess) return e;
        }
        const int len = std::max(upto + 1, 64);
        hipError_t e = hipHostMalloc((void**)&pf.host, sizeof(uint32_t) * (size_t)len, hipHostMallocDefault);
        if (e != hipSuccess) return e;
        pf.host_len = len;
    }
    if (!pf.ev) {
        hipError_t e = hipEventCreateWithFlags(&pf.ev, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    hipError_t e = hipMemcpyAsync(pf.host, f->pb.bounce_paths, sizeof(uint32_t) * (size_t)(upto + 1),
                                  hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipEventRecord(pf.ev, st);
    if (e == hipSuccess) {
        pf.pending = true;
        pf.pending_upto = upto;
    }
    return e;
}

// Enqueues the next planned chunk on the film's stream (no wait); ev0/ev1 bracket it.
int film_enqueue_chunk(prt_film* f) {
    if (film_round_exhausted(f)) return PRT_OK;
    if (f->track && f->enq - f->seen >= prt_film::kRing)
        return fail(PRT_ERR_INVALID_ARGUMENT, "too many tracked chunks in flight");
    const prt_scene* s = f->scene;
    PRT_HIP(hipSetDevice(s->device));
    const DevScene& sc = f->sc;
    hipStream_t st = f->stream;
    const prt_film::Chunk& c = f->plan[f->plan_pos++];
    PassParams pp;
    pp.width = f->width;
    pp.height = f->height;
    pp.px0 = c.px0;
    pp.npix = c.npix;
    pp.spp = c.spp;
    pp.sample_base = c.sample_base;
    pp.max_depth = f->params.max_depth;
    pp.rr_min = f->params.rr_min_bounces;
    pp.seed = f->params.seed;
    pp.trace_gate = 0;
    pp.tune_stride = 1;
    pp.act = f->adaptive ? f->act[f->ad_cur] : nullptr;
    pp.count = f->fb.count;
    AdaptiveParams ap{};
    if (f->adaptive) {
        ap.target = f->ad_target;
        ap.min_samples = f->ad_min;
        ap.threshold = f->params.adaptive_threshold;
        ap.next = f->act[f->ad_cur ^ 1];
        ap.state = f->ad_state + 3 * (f->ad_cur ^ 1);
    }
    int64_t n = (int64_t)pp.npix * pp.spp;
    const int ring = (int)(f->enq % prt_film::kRing);
    if (f->track) PRT_HIP(hipEventRecord(f->rev0[ring], st));
    PRT_HIP(hipEventRecord(f->ev0, st));
    if (f->bdpt) {
        int grid = (int)std::min<int64_t>(f->bdpt_blocks, (n + kBlock - 1) / kBlock);
        int grid_g = (int)std::min<int64_t>(f->bdpt_general_blocks, (n + kBlock - 1) / kBlock);
        const int walk_blocks = (int)((n + kBlock - 1) / kBlock);
        // tuning runs time both subpaths in one kernel; regular launches run a kernel per side
        auto launch_paths = [&](const PassParams& q, bool split = false) {
#define PRT_PATHS_LAUNCH(M, S)                                                                                  \
    hipLaunchKernelGGL((k_bdpt_paths<M, S>), dim3(grid), dim3(kBlock), 0, st, sc, q, f->cpath, f->lpath, f->lens, \
                       f->gstack, f->gslots)
            if (split && sc.mixed) {
                PRT_PATHS_LAUNCH(true, 1);
                PRT_PATHS_LAUNCH(true, 2);
            } else if (split) {
                PRT_PATHS_LAUNCH(false, 1);
                PRT_PATHS_LAUNCH(false, 2);
            } else if (sc.mixed) {
                PRT_PATHS_LAUNCH(true, 3);
            } else {
                PRT_PATHS_LAUNCH(false, 3);
            }
#undef PRT_PATHS_LAUNCH
            return hipGetLastError();
        };
        // per-scene leaf gate of the BDPT traversals, on a chunk of at least 2^18 paths that is not a one-sample
        // first chunk (a latency-bound first pass): after a warm-up launch, the subpath kernel is
        // timed twice with each gate on about 2^18 of the chunk's samples (the regular launch writes their
        // subpaths again). Nothing waits for the timings: a later chunk of a film of the scene reads them
        // and keeps the fastest gate (gate 0 unless another is more than 3 % faster) for every BDPT kernel.
        prt_scene* ms = const_cast<prt_scene*>(s);
        {
            std::lock_guard<std::mutex> lk(ms->bdpt_tune.mu);
            const int nch = gate_choices(sc);
            collect_bdpt_tuning_locked(ms);
            if (ms->bdpt_trace_gate.load() < 0 && !ms->bdpt_tune.pending && n >= ((int64_t)1 << 18) &&
                !f->probe_flags && !f->adaptive && (f->enq > 0 || pp.spp > 1)) {
                PassParams q = pp;
                q.tune_stride = (int32_t)std::max<int64_t>(1, n >> 18);
                hipError_t e = tune_begin(ms->bdpt_tune, 2 * nch, st);
                if (e == hipSuccess) e = launch_paths(q);
                if (e == hipSuccess) e = hipEventRecord(ms->bdpt_tune.ev[1], st);
                for (int k = 0; k < 2 * nch && e == hipSuccess; k++) {
                    q.trace_gate = kGateChoice[k % nch];
                    e = launch_paths(q);
                    if (e == hipSuccess) e = hipEventRecord(ms->bdpt_tune.ev[2 + k], st);
                }
                PRT_HIP(e);
                ms->bdpt_tune.pending = true;
            }
        }
        pp.trace_gate = std::max(0, ms->bdpt_trace_gate.load());
        PRT_HIP(probe_mark(f, PRT_PROBE_KIND_BDPT_PATHS, true));
        PRT_HIP(launch_paths(pp, true));
        PRT_HIP(probe_mark(f, PRT_PROBE_KIND_BDPT_PATHS, false));
        PRT_HIP(probe_mark(f, PRT_PROBE_KIND_BDPT_WALKS, true));
        if (!f->bdpt_pairs)  // the strategy lists store the walk sums in k_bdpt_pairs
            hipLaunchKernelGGL(k_bdpt_walks, dim3(walk_blocks), dim3(kBlock), 0, st, pp, f->cpath, f->lpath, f->lens);
        PRT_HIP(probe_mark(f, PRT_PROBE_KIND_BDPT_WALKS, false));
        PRT_HIP(probe_mark(f, PRT_PROBE_KIND_BDPT_CONNECT, true));
        if (f->bdpt_pairs) {
            // strategy queues: filled after the subpaths, drained by the strategy kernels, results added per sample
            // the strategy kernels' grids: a multiple of kNumXcd, at most about one item per lane and sample
            const int64_t cap_blocks = std::max<int64_t>(kNumXcd, ((n + kBlock - 1) / kBlock) / kNumXcd * kNumXcd);
            const int grid_p = (int)std::min<int64_t>(f->bdpt_pair_blocks, cap_blocks);
            auto grid_e = [&](int k) { return (int)std::min<int64_t>(f->bdpt_edge_blocks[k], cap_blocks); };
            PRT_HIP(hipMemsetAsync(f->pairs.count, 0, kNumLists * kCounterStride * sizeof(uint32_t), st));
            hipLaunchKernelGGL(k_bdpt_pairs, dim3(walk_blocks), dim3(kBlock), 0, st, pp, f->cpath, f->lpath, f->lens,
                               f->pairs);
#define PRT_EDGE_LAUNCH(M, K)                                                                                    \
    hipLaunchKernelGGL((k_bdpt_edge_connect<M, K>), dim3(grid_e(K)), dim3(kBlock), 0, st, sc, pp, f->cpath, f->lpath, \
                       f->lens, f->fb.splat, f->pairs, f->gstack, f->gslots)
            if (sc.mixed) {
                PRT_EDGE_LAUNCH(true, kKindEmission);
                PRT_EDGE_LAUNCH(true, kKindDirect);
                PRT_EDGE_LAUNCH(true, kKindLightTrace);
                hipLaunchKernelGGL((k_bdpt_pair_connect<true>), dim3(grid_p), dim3(kBlock), 0, st, sc, pp, f->cpath,
                                   f->lpath, f->pairs, f->gstack, f->gslots);
            } else {
                PRT_EDGE_LAUNCH(false, kKindEmission);
                PRT_EDGE_LAUNCH(false, kKindDirect);
                PRT_EDGE_LAUNCH(false, kKindLightTrace);
                hipLaunchKernelGGL((k_bdpt_pair_connect<false>), dim3(grid_p), dim3(kBlock), 0, st, sc, pp, f->cpath,
                                   f->lpath, f->pairs, f->gstack, f->gslots);
            }
#undef PRT_EDGE_LAUNCH
            hipLaunchKernelGGL(k_bdpt_pair_sum, dim3(walk_blocks), dim3(kBlock), 0, st, pp, f->pb, f->pairs);
        } else if (sc.mixed) {
            hipLaunchKernelGGL((k_bdpt_connect<true>), dim3(grid), dim3(kBlock), 0, st, sc, pp, f->pb, f->cpath,
                               f->lpath, f->lens, f->fb.splat, f->gstack, f->gslots);
            hipLaunchKernelGGL((k_bdpt_connect_general<true>), dim3(grid_g), dim3(kBlock), 0, st, sc,
                               pp, f->pb, f->cpath, f->lpath, f->lens, f->gstack, f->gslots);
        } else {
            hipLaunchKernelGGL((k_bdpt_connect<false>), dim3(grid), dim3(kBlock), 0, st, sc, pp, f->pb, f->cpath,
                               f->lpath, f->lens, f->fb.splat, f->gstack, f->gslots);
            hipLaunchKernelGGL((k_bdpt_connect_general<false>), dim3(grid_g), dim3(kBlock), 0, st,
                               sc, pp, f->pb, f->cpath, f->lpath, f->lens, f->gstack, f->gslots);
        }
        PRT_HIP(probe_mark(f, PRT_PROBE_KIND_BDPT_CONNECT, false));
        PRT_HIP(probe_mark(f, PRT_PROBE_KIND_FILM, true));
        if (f->adaptive)
            hipLaunchKernelGGL(k_film_adaptive, dim3((pp.npix + kBlock - 1) / kBlock), dim3(kBlock), 0, st, pp,
                               f->pb, f->fb, ap);
        else
            hipLaunchKernelGGL(k_film, dim3((pp.npix + kBlock - 1) / kBlock), dim3(kBlock), 0, st, pp, f->pb,
                               f->fb);
        PRT_HIP(probe_mark(f, PRT_PROBE_KIND_FILM, false));
        PRT_HIP(hipGetLastError());
        PRT_HIP(hipEventRecord(f->ev1, st));
        if (f->track) PRT_HIP(hipEventRecord(f->rev1[ring], st));
        f->enq++;
        return PRT_OK;
    }
    PRT_HIP(hipMemsetAsync(f->pb.counters, 0, kNumCounters * kCounterStride * sizeof(uint32_t), st));
    int cam_blocks = (int)((n + kBlock - 1) / kBlock);
    PRT_HIP(probe_mark(f, PRT_PROBE_KIND_CAMERA, true));
    hipLaunchKernelGGL(k_camera, dim3(cam_blocks), dim3(kBlock), 0, st, sc, pp, f->pb);
    PRT_HIP(probe_mark(f, PRT_PROBE_KIND_CAMERA, false));
    // bounce b: one trace launch (the shadow rays of bounce b - 1 and the extension rays of bounce b),
    // then shading; a last launch traces the shadow rays of the last bounce
    const bool counting = (f->probe_flags & PRT_PROBE_COUNT) != 0;
    // the regular trace launches park long rays (not the counting or tuning launches)
    const uint32_t park = counting ? 0u : f->park_steps, park_small = counting ? 0u : f->park_steps_small;
    auto launch_trace = [&](int v, int gate, int b, int closest, int sq, uint32_t tstride = 1, uint32_t budget = 0,
                            uint32_t budget_small = 0) {
        PathBuffers pbv = f->pb;
        uint2* gs = f->gstack;
        uint32_t gsl = f->gslots;
        unsigned long long* pr = f->probe_dev;
        int g = gate;
        uint32_t small_per_lane = f->park_small_rays_per_lane;
        void* args[] = {(void*)&sc, (void*)&pbv, (void*)&b,  (void*)&closest, (void*)&sq,
                        (void*)&gs, (void*)&gsl, (void*)&pr, (void*)&g,       (void*)&tstride,
                        (void*)&budget, (void*)&budget_small, (void*)&small_per_lane};
        hipError_t e = hipLaunchKernel(trace_fn(sc.mixed != 0, counting, v), dim3(f->trace_grid[v]), dim3(kBlock),
                                       args, f->trace_pad[v], st);
        if (e != hipSuccess || (budget == 0 && budget_small == 0)) return e;
        PathBuffers pbl = f->pb;
        void* largs[] = {(void*)&sc, (void*)&pbl, (void*)&gs, (void*)&gsl};
        e = hipLaunchKernel(long_fn(sc.mixed != 0, f->coop_lanes), dim3(f->long_grid), dim3(64), largs, 0, st);
        if (e != hipSuccess) return e;
        return hipGetLastError();
    };
    // per-scene leaf gate of the trace kernel, on a chunk of at least 2^20 paths that is not a one-sample first
    // chunk (a latency-bound first pass): between the first bounce's shading and the second bounce's
    // trace, after a warm-up launch, the second bounce's closest-hit trace of about 2^20 of its rays (it
    // only writes their hit records, which the regular trace writes again) is timed twice with each gate.
    // Nothing waits for the timings: a later chunk of a film of the scene reads them and keeps the
    // fastest gate (gate 0 unless another is more than 3 % faster).
    prt_scene* ms = const_cast<prt_scene*>(s);
    const int tv = ms->pt_trace_variant;
    ms->info.pt_trace_waves = f->trace_grid[tv] / std::max(1, s->num_cus);  // work-groups of 4 waves per CU
    collect_pt_tuning(ms);
    const bool tune_here = n >= ((int64_t)1 << 20) && !f->probe_flags && pp.max_depth > 1 && !f->adaptive &&
                           (f->enq > 0 || pp.spp > 1) && ms->pt_trace_gate.load() < 0;
    auto tune_bounce1 = [&]() -> hipError_t {
        std::lock_guard<std::mutex> lk(ms->pt_tune.mu);
        if (ms->pt_trace_gate.load() >= 0 || ms->pt_tune.pending) return hipSuccess;
        const uint32_t tstride = (uint32_t)std::max<int64_t>(1, n >> 20);
        const int nch = gate_choices(sc);
        hipError_t e = tune_begin(ms->pt_tune, 2 * nch, st);
        if (e == hipSuccess) e = launch_trace(tv, 0, 1, 1, -1, tstride);
        if (e == hipSuccess) e = hipEventRecord(ms->pt_tune.ev[1], st);
        for (int k = 0; k < 2 * nch && e == hipSuccess; k++) {
            e = launch_trace(tv, kGateChoice[k % nch], 1, 1, -1, tstride);
            if (e == hipSuccess) e = hipEventRecord(ms->pt_tune.ev[2 + k], st);
        }
        if (e == hipSuccess) ms->pt_tune.pending = true;
        return e;
    };
    const int tgate = std::max(0, ms->pt_trace_gate.load());
    auto trace = [&](int b, int closest, int sq) {
        hipError_t e = probe_mark(f, PRT_PROBE_KIND_TRACE, true);
        if (e == hipSuccess) e = launch_trace(tv, tgate, b, closest, sq, 1, park, park_small);
        if (e == hipSuccess) e = probe_mark(f, PRT_PROBE_KIND_TRACE, false);
        if (f->probe_flags) f->probe_launches++;
        return e;
    };
    // the bounce from which k_tail finishes the chunk: the first the scene's survival profile expects to
    // hold at most tail_paths paths (probed chunks run every bounce as wavefront launches)
    const int b0 = f->probe_flags ? pp.max_depth : tail_bounce(ms, n, f->tail_paths, pp.max_depth);
    for (int b = 0; b < b0; b++) {
        if (b == 1 && tune_here) PRT_HIP(tune_bounce1());
        PRT_HIP(trace(b, 1, b > 0 ? ((b - 1) & 1) : -1));
        PRT_HIP(probe_mark(f, PRT_PROBE_KIND_SHADE, true));
        hipLaunchKernelGGL(k_shade, dim3(f->blocks), dim3(kBlock), 0, st, sc, pp, f->pb, b);
        PRT_HIP(probe_mark(f, PRT_PROBE_KIND_SHADE, false));
    }
    if (b0 < pp.max_depth) {
        // bounce b0's trace (with the shadow rays of b0 - 1), then every later bounce in one launch
        PRT_HIP(trace(b0, 1, b0 > 0 ? ((b0 - 1) & 1) : -1));
        if (sc.mixed)
            hipLaunchKernelGGL(k_tail<true>, dim3(f->tail_blocks), dim3(kBlock), 0, st, sc, pp, f->pb, b0, f->gstack,
                               f->gslots, tgate);
        else
            hipLaunchKernelGGL(k_tail<false>, dim3(f->tail_blocks), dim3(kBlock), 0, st, sc, pp, f->pb, b0, f->gstack,
                               f->gslots, tgate);
    } else if (pp.max_depth > 0) {
        PRT_HIP(trace(pp.max_depth, 0, (pp.max_depth - 1) & 1));
    }
    // paths per bounce of this chunk (bounces 0 .. b0 ran as wavefront trace launches) for the profile
    if (!f->probe_flags && pp.max_depth > 0 && !f->adaptive) PRT_HIP(profile_record(ms, f, std::min(b0, pp.max_depth - 1), st));
    int film_blocks_n = (pp.npix + kBlock - 1) / kBlock;
    PRT_HIP(probe_mark(f, PRT_PROBE_KIND_FILM, true));
    if (f->adaptive)
        hipLaunchKernelGGL(k_film_adaptive, dim3(film_blocks_n), dim3(kBlock), 0, st, pp, f->pb, f->fb, ap);
    else
        hipLaunchKernelGGL(k_film, dim3(film_blocks_n), dim3(kBlock), 0, st, pp, f->pb, f->fb);
    PRT_HIP(probe_mark(f, PRT_PROBE_KIND_FILM, false));
    PRT_HIP(hipGetLastError());
    PRT_HIP(hipEventRecord(f->ev1, st));
    if (f->track) PRT_HIP(hipEventRecord(f->rev1[ring], st));
    f->enq++;
    return PRT_OK;
}

// Device time of the last enqueued chunk (call when idle), accumulated over the plan.
double film_chunk_ms(prt_film* f) {
    float t = 0;
    if (hipEventElapsedTime(&t, f->ev0, f->ev1) == hipSuccess) f->plan_ms += t;
    return f->plan_ms;
}

int film_track_chunks(prt_film* f) {
    if (f->track) return PRT_OK;
    PRT_HIP(hipSetDevice(f->scene->device));
    for (int i = 0; i < prt_film::kRing; i++) {
        PRT_HIP(hipEventCreate(&f->rev0[i]));
        PRT_HIP(hipEventCreate(&f->rev1[i]));
    }
    f->track = true;
    f->seen = f->enq;
    return PRT_OK;
}

// Chunks still on the device; each chunk seen finished adds its device time to the plan's.
int film_chunks_pending(prt_film* f) {
    while (f->seen < f->enq) {
        const int i = (int)(f->seen % prt_film::kRing);
        if (hipEventQuery(f->rev1[i]) == hipErrorNotReady) break;
        float t = 0;
        if (hipEventElapsedTime(&t, f->rev0[i], f->rev1[i]) == hipSuccess) f->plan_ms += t;
        f->seen++;
    }
    return (int)(f->enq - f->seen);
}

double film_plan_ms(const prt_film* f) { return f->plan_ms; }

int64_t film_capacity_paths(const prt_film* f) { return f->capacity; }

void film_set_chunk_paths(prt_film* f, int64_t n) { f->chunk_paths = std::max<int64_t>(1, std::min(n, f->capacity)); }

int64_t bdpt_grid_paths(const prt_scene* scene) { return (int64_t)bdpt_blocks(scene, film_blocks(scene)) * kBlock; }

// Enqueues `spp` samples per pixel for every pixel on the film's stream (no wait).
int film_enqueue(prt_film* f, int32_t spp) {
    int rc = film_plan(f, spp);
    if (rc != PRT_OK) return rc;
    hipEvent_t first = nullptr;
    PRT_HIP(hipEventCreate(&first));
    (void)hipEventRecord(first, f->stream);
    while (!film_round_exhausted(f)) {
        rc = film_enqueue_chunk(f);
        if (rc != PRT_OK) {
            (void)hipEventDestroy(first);
            return rc;
        }
    }
    // ev0 brackets the whole request: replace the last chunk's start event by the first one
    std::swap(first, f->ev0);
    (void)hipEventDestroy(first);
    return PRT_OK;
}

bool film_idle(prt_film* f) { return hipEventQuery(f->ev1) != hipErrorNotReady; }

int film_wait(prt_film* f, float* ms) {
    PRT_HIP(hipSetDevice(f->scene->device));
    PRT_HIP(hipEventSynchronize(f->ev1));
    float t = 0;
    PRT_HIP(hipEventElapsedTime(&t, f->ev0, f->ev1));
    if (ms) *ms = t;
    return PRT_OK;
}

int32_t film_spp_done(const prt_film* f) { return f->spp_done; }

size_t film_device_bytes(const prt_scene* scene, const prt_render_params* p, int64_t cap) {
    int32_t w = p->width > 0 ? p->width : scene->image_width;
    int32_t h = p->height > 0 ? p->height : scene->image_height;
    int64_t npix = (int64_t)w * h;
    const bool bd = p->integrator == PRT_INTEGRATOR_BDPT;
    const bool uv = scene->dev.textured != 0;
    int64_t C = film_capacity(npix, bd ? bdpt_capacity(p->max_depth, cap, uv) : cap);
    int blocks = film_blocks(scene);
    for (int v = 0; v < kTraceVariants; v++) blocks = std::max(blocks, trace_grid(scene, v));
    size_t g = film_gstack_entries(scene) * (size_t)blocks * kBlock * sizeof(uint2);
    size_t b = 0;
    if (bd)
        b = 2 * bdpt_store_bytes(p->max_depth, C, uv) + bdpt_pair_bytes(p->max_depth, C) + sizeof(uint32_t) * C +
            3 * sizeof(long long) * npix;
    if (p->adaptive_min_samples < 1.0) b += 2 * sizeof(uint32_t) * npix + 6 * sizeof(uint32_t);
    b += sizeof(uchar4) * (size_t)npix;  // the device image of RGBA8 reads, allocated at the first read
    b += sizeof(uint32_t) * (size_t)std::min<int64_t>(2 * C, kLongCapacity);  // long-ray queue
    b += sizeof(uint32_t) * (size_t)(p->max_depth + 1);                          // paths per bounce
    return (size_t)C * kPathBytes + (size_t)npix * kPixelBytes + g + b + 3 * sizeof(uint32_t) * (size_t)C;
}

int film_create_ex(const prt_scene* scene, const prt_render_params* p, int64_t cap, prt_film** out, FilmPool* pool) {
    if (!scene || !p || !out) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (p->integrator != PRT_INTEGRATOR_PATH && p->integrator != PRT_INTEGRATOR_BDPT)
        return fail(PRT_ERR_INVALID_ARGUMENT, "unknown integrator");
    if (p->max_depth < 0 || p->max_depth > 1024) return fail(PRT_ERR_INVALID_ARGUMENT, "max_depth out of range");
    if (p->integrator == PRT_INTEGRATOR_BDPT && p->max_depth > kBdptMaxDepth)
        return fail(PRT_ERR_UNSUPPORTED, "BDPT max_depth above 64");
    if (p->max_passes < 1 || p->max_samples_per_pixel < 1 || p->initial_samples < 0)
        return fail(PRT_ERR_INVALID_ARGUMENT, "bad pass schedule");
    if (!std::isfinite(p->adaptive_min_samples) || !std::isfinite(p->adaptive_threshold))
        return fail(PRT_ERR_INVALID_ARGUMENT, "adaptive sampling parameters must be finite");
    int32_t w = p->width > 0 ? p->width : scene->image_width;
    int32_t h = p->height > 0 ? p->height : scene->image_height;
    if (w <= 0 || h <= 0 || (int64_t)w * h > (1ll << 28)) return fail(PRT_ERR_INVALID_ARGUMENT, "bad image size");
    PRT_HIP(hipSetDevice(scene->device));
    prt_film* f = new prt_film();
    f->scene = scene;
    f->pool = pool;
    const_cast<prt_scene*>(scene)->users++;
    f->sc = scene->dev;
    if (w != scene->image_width || h != scene->image_height) {
        // image-size override: camera width and aspect from the requested size (as the web server's
        // width/height parameters override the scene camera, web/server/render.go:371-404)
        prt_camera c = scene->cam_desc;
        c.width = w;
        c.aspect_ratio = (double)w / (double)h;
        int H = 0;
        f->sc.cam = make_camera(c, H, h);
    }
    f->params = *p;
    f->width = w;
    f->height = h;
    f->npix = (int64_t)w * h;
    f->capacity = film_capacity(
        f->npix, p->integrator == PRT_INTEGRATOR_BDPT ? bdpt_capacity(p->max_depth, cap, scene->dev.textured != 0) : cap);
    f->chunk_paths = f->capacity;
    int rc;
    auto bail = [&](int code) {
        prt_film_destroy(f);
        return code;
    };
    const int64_t C = f->capacity;
    void* ptr;
#define FALLOC(field, type, count)                                                               \
    if ((rc = film_alloc(f, sizeof(type) * (size_t)(count), &ptr)) != PRT_OK) return bail(rc); \
    field = (type*)ptr;
    FALLOC(f->pb.o, float4, C);
    FALLOC(f->pb.d, float4, C);
    FALLOC(f->pb.beta, float4, C);
    FALLOC(f->pb.L, float4, C);
    FALLOC(f->pb.hit, float4, C);
    FALLOC(f->pb.sd, float4, C);
    FALLOC(f->pb.sc, float4, C);
    FALLOC(f->pb.key, uint32_t, C);
    FALLOC(f->pb.queue[0], uint32_t, C);
    FALLOC(f->pb.queue[1], uint32_t, C);
    FALLOC(f->pb.squeue[0], uint32_t, C);
    FALLOC(f->pb.squeue[1], uint32_t, C);
    FALLOC(f->pb.counters, uint32_t, kNumCounters * kCounterStride);
    f->pb.lcap = (uint32_t)std::min<int64_t>(2 * C, kLongCapacity);
    FALLOC(f->pb.lq, uint32_t, f->pb.lcap);
    FALLOC(f->pb.bounce_paths, uint32_t, p->max_depth + 1);
    FALLOC(f->fb.sum, float4, f->npix);
    FALLOC(f->fb.comp, float4, f->npix);
    FALLOC(f->fb.sq, float2, f->npix);
    FALLOC(f->fb.count, uint32_t, f->npix);
#undef FALLOC
    f->blocks = film_blocks(scene);
    int slot_blocks = f->blocks;
    for (int v = 0; v < kTraceVariants; v++) {
        f->trace_grid[v] = trace_grid(scene, v);
        f->trace_pad[v] = trace_lds_pad(scene->device, v);
        slot_blocks = std::max(slot_blocks, f->trace_grid[v]);
    }
    f->gslots = (uint32_t)slot_blocks * kBlock;
    f->park_steps = park_steps(scene, false);
    f->park_steps_small = park_steps(scene, true);
    f->park_small_rays_per_lane = park_small_rays_per_lane();
    f->tail_blocks = tail_grid(scene, f->trace_grid[scene->pt_trace_variant]);
    f->tail_paths = tail_paths(f->tail_blocks);
    f->coop_lanes = coop_lanes();
    f->long_grid = (int)std::min<uint32_t>(f->gslots / 64u, (uint32_t)long_grid(scene, f->coop_lanes));
    if ((rc = film_alloc(f, film_gstack_entries(scene) * f->gslots * sizeof(uint2), &ptr)) != PRT_OK) return bail(rc);
    f->gstack = (uint2*)ptr;
    if (p->integrator == PRT_INTEGRATOR_BDPT) {
        f->bdpt = true;
        f->bdpt_blocks = bdpt_blocks(scene, f->blocks);
        f->bdpt_general_blocks = bdpt_general_blocks(scene, f->blocks);
        const uint32_t uv = scene->dev.textured ? 1u : 0u;
        const size_t vb = bdpt_store_bytes(p->max_depth, C, uv != 0);
        if ((rc = film_alloc(f, vb, &ptr)) != PRT_OK) return bail(rc);
        f->cpath = VertexStore{(float4*)ptr, (uint32_t)C, (uint32_t)(p->max_depth + 1), uv};
        if ((rc = film_alloc(f, vb, &ptr)) != PRT_OK) return bail(rc);
        f->lpath = VertexStore{(float4*)ptr, (uint32_t)C, (uint32_t)(p->max_depth + 1), uv};
        if ((rc = film_alloc(f, sizeof(uint32_t) * (size_t)C, &ptr)) != PRT_OK) return bail(rc);
        f->lens = (uint32_t*)ptr;
        if ((rc = film_alloc(f, 3 * sizeof(long long) * (size_t)f->npix, &ptr)) != PRT_OK) return bail(rc);
        f->fb.splat = (long long*)ptr;
        if (bdpt_pairs_used(p->max_depth)) {
            f->bdpt_pairs = true;
            f->bdpt_pair_blocks = bdpt_queue_blocks(scene, f->blocks, (const void*)k_bdpt_pair_connect<true>);
            f->bdpt_edge_blocks[kKindDirect] =
                bdpt_queue_blocks(scene, f->blocks, (const void*)k_bdpt_edge_connect<true, kKindDirect>);
            f->bdpt_edge_blocks[kKindLightTrace] =
                bdpt_queue_blocks(scene, f->blocks, (const void*)k_bdpt_edge_connect<true, kKindLightTrace>);
            f->bdpt_edge_blocks[kKindEmission] =
                bdpt_queue_blocks(scene, f->blocks, (const void*)k_bdpt_edge_connect<true, kKindEmission>);
            PairBuffers& pq = f->pairs;
            pq.nt = (uint32_t)p->max_depth;
            pq.np = (uint32_t)(p->max_depth - 1) * pq.nt;
            pq.nq = pq.np + 2 * pq.nt;
            pq.stride = (uint32_t)C;
            const uint32_t list_len[kNumLists] = {pq.np, pq.nt, pq.nt - 1, 1};  // strategies per kind
            for (int k = 0; k < kNumLists; k++) {
                if ((rc = film_alloc(f, sizeof(uint32_t) * (size_t)list_len[k] * (size_t)C, &ptr)) != PRT_OK)
                    return bail(rc);
                pq.items[k] = (uint32_t*)ptr;
            }
            if ((rc = film_alloc(f, sizeof(float4) * (size_t)(pq.np + pq.nt + 1) * (size_t)C, &ptr)) != PRT_OK)
                return bail(rc);
            pq.res = (float4*)ptr;
            if ((rc = film_alloc(f, sizeof(uint2) * (size_t)C, &ptr)) != PRT_OK) return bail(rc);
            pq.mask = (uint2*)ptr;
            if ((rc = film_alloc(f, kNumLists * kCounterStride * sizeof(uint32_t), &ptr)) != PRT_OK) return bail(rc);
            pq.count = (uint32_t*)ptr;
        }
    }
    if (p->adaptive_min_samples < 1.0) {
        f->adaptive = true;
        for (int l = 0; l < 2; l++) {
            if ((rc = film_alloc(f, sizeof(uint32_t) * (size_t)f->npix, &ptr)) != PRT_OK) return bail(rc);
            f->act[l] = (uint32_t*)ptr;
        }
        if ((rc = film_alloc(f, 6 * sizeof(uint32_t), &ptr)) != PRT_OK) return bail(rc);
        f->ad_state = (uint32_t*)ptr;
    }
    // the device image of RGBA8 reads (allocated here, so a pass's first read does not allocate)
    if ((rc = film_alloc(f, sizeof(uchar4) * (size_t)f->npix, &ptr)) != PRT_OK) return bail(rc);
    f->rgba_dev = (uchar4*)ptr;
    hipError_t e = hipSuccess;
    if (pool) {
        if ((rc = pool->take_stream(&f->stream)) != PRT_OK) return bail(rc);
    } else {
        e = hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking);
        if (e != hipSuccess) return bail(hip_fail(e, "hipStreamCreate"));
    }
    e = hipEventCreate(&f->ev0);
    if (e == hipSuccess) e = hipEventCreate(&f->ev1);
    if (e != hipSuccess) return bail(hip_fail(e, "hipEventCreate"));
    if ((rc = prt_film_reset(f)) != PRT_OK) return bail(rc);
    *out = f;
    return PRT_OK;
}

}  // namespace prt

namespace {
// Path slots of a standalone film: chunks of up to 2^26 paths (~9 GB of path state), at most an eighth of
// the device's memory. A chunk's deep bounces are launches whose duration is set by its longest paths
// (glass paths along the inside of a mesh take ~0.35 ms per bounce with few paths left), whatever the
// number of paths, so larger chunks spread them over more samples: the 1.8M-triangle glass scene at
// 1024x1024 renders 565 M samples/s in chunks of 2^23 paths, 643 M at 2^24, 713-752 M at 2^25, 793 M at
// 2^26 and 756 M at 2^27; the Cornell boxes gain 1 %.
int64_t default_capacity() {
    int64_t cap = (int64_t)1 << 26;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b > 0)
        cap = std::min<int64_t>(cap, (int64_t)(total_b / 8 / kPathBytes));
    return std::max<int64_t>(cap, (int64_t)1 << 20);
}

int render_samples(prt_film* f, int32_t spp, float* ms) {
    int rc = film_enqueue(f, spp);
    if (rc != PRT_OK) return rc;
    return film_wait(f, ms);
}

// One RenderPass to `target` samples per pixel; adaptive films run their rounds to the end.
int render_to(prt_film* f, int32_t target, float* ms) {
    if (!f->adaptive) {
        int32_t need = target - f->spp_done;
        if (need <= 0) return PRT_OK;
        return render_samples(f, need, ms);
    }
    PRT_HIP(hipSetDevice(f->scene->device));
    hipEvent_t start = nullptr;
    PRT_HIP(hipEventCreate(&start));
    int rc = PRT_OK;
    if (hipEventRecord(start, f->stream) != hipSuccess) rc = fail(PRT_ERR_DEVICE, "hipEventRecord");
    if (rc == PRT_OK) rc = film_plan_pass(f, target);
    while (rc == PRT_OK && !film_plan_done(f)) {
        rc = film_advance(f);  // waits for the enqueued round
        while (rc == PRT_OK && !film_round_exhausted(f)) rc = film_enqueue_chunk(f);
    }
    if (rc == PRT_OK && hipEventRecord(f->ev1, f->stream) != hipSuccess) rc = fail(PRT_ERR_DEVICE, "hipEventRecord");
    if (rc == PRT_OK && hipEventSynchronize(f->ev1) != hipSuccess) rc = fail(PRT_ERR_DEVICE, "hipEventSynchronize");
    if (rc == PRT_OK && ms) {
        float t = 0;
        (void)hipEventElapsedTime(&t, start, f->ev1);
        *ms = t;
    }
    (void)hipEventDestroy(start);
    return rc;
}

// RenderStats from the per-pixel counts (assembleCurrentImage, progressive.go:395-431).
int count_stats(const prt_film* f, int64_t& total, int32_t& mn, int32_t& mx) {
    std::vector<uint32_t> cnt(f->npix);
    PRT_HIP(hipMemcpyAsync(cnt.data(), f->fb.count, sizeof(uint32_t) * f->npix, hipMemcpyDeviceToHost, f->stream));
    PRT_HIP(hipStreamSynchronize(f->stream));
    total = 0;
    mn = f->params.max_samples_per_pixel;
    mx = 0;
    for (uint32_t c : cnt) {
        total += c;
        mn = std::min<int32_t>(mn, (int32_t)c);
        mx = std::max<int32_t>(mx, (int32_t)c);
    }
    return PRT_OK;
}
}  // namespace

extern "C" {

int prt_film_create(const prt_scene* scene, const prt_render_params* p, prt_film** out) {
    if (!scene) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL argument");
    PRT_HIP(hipSetDevice(scene->device));
    return film_create_ex(scene, p, default_capacity(), out);
}

int prt_film_reset(prt_film* f) {
    if (!f) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL film");
    PRT_HIP(hipSetDevice(f->scene->device));
    PRT_HIP(hipMemsetAsync(f->fb.sum, 0, sizeof(float4) * f->npix, f->stream));
    PRT_HIP(hipMemsetAsync(f->fb.comp, 0, sizeof(float4) * f->npix, f->stream));
    PRT_HIP(hipMemsetAsync(f->fb.sq, 0, sizeof(float2) * f->npix, f->stream));
    PRT_HIP(hipMemsetAsync(f->fb.count, 0, sizeof(uint32_t) * f->npix, f->stream));
    if (f->fb.splat) PRT_HIP(hipMemsetAsync(f->fb.splat, 0, 3 * sizeof(long long) * f->npix, f->stream));
    PRT_HIP(hipStreamSynchronize(f->stream));
    f->spp_done = 0;
    f->plan.clear();
    f->plan_pos = 0;
    f->ad_running = false;
    f->ad_selected = false;
    return PRT_OK;
}

int prt_film_size(const prt_film* f, int32_t* width, int32_t* height) {
    if (!f || !width || !height) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL argument");
    *width = f->width;
    *height = f->height;
    return PRT_OK;
}

int prt_film_destroy(prt_film* f) {
    if (!f) return PRT_OK;
    (void)hipSetDevice(f->scene->device);
    if (f->stream) (void)hipStreamSynchronize(f->stream);
    collect_scene_events(const_cast<prt_scene*>(f->scene));
    for (auto& m : f->marks) {
        if (m.a) (void)hipEventDestroy(m.a);
        if (m.b) (void)hipEventDestroy(m.b);
    }
    for (const DevBuf& b : f->bufs) {
        if (f->pool)
            f->pool->give(b.p, b.bytes);
        else
            (void)hipFree(b.p);
    }
    if (f->ev0) (void)hipEventDestroy(f->ev0);
    if (f->ev1) (void)hipEventDestroy(f->ev1);
    for (int i = 0; i < prt_film::kRing; i++) {
        if (f->rev0[i]) (void)hipEventDestroy(f->rev0[i]);
        if (f->rev1[i]) (void)hipEventDestroy(f->rev1[i]);
    }
    if (f->stream) {
        if (f->pool)
            f->pool->give_stream(f->stream);
        else
            (void)hipStreamDestroy(f->stream);
    }
    const_cast<prt_scene*>(f->scene)->users--;
    delete f;
    return PRT_OK;
}

static void fill_stats(const prt_film* f, int32_t target, float ms, prt_pass_stats* st) {
    if (!st) return;
    st->total_pixels = f->npix;
    st->total_samples = f->npix * (int64_t)f->spp_done;
    st->average_samples = (double)f->spp_done;
    st->max_samples = target;
    st->min_samples = f->spp_done;
    st->max_samples_used = f->spp_done;
    st->target_samples = target;
    st->device_ms = ms;
}

int prt_render_samples(prt_film* f, int32_t spp, prt_pass_stats* st) {
    if (!f) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL film");
    if (spp < 0) return fail(PRT_ERR_INVALID_ARGUMENT, "negative spp");
    if (f->adaptive)
        return fail(PRT_ERR_UNSUPPORTED, "adaptive films (AdaptiveMinSamples < 1) render by pass (prt_render_pass)");
    float ms = 0;
    if (spp > 0) {
        int rc = render_samples(f, spp, &ms);
        if (rc != PRT_OK) return rc;
    }
    fill_stats(f, f->spp_done, ms, st);
    return PRT_OK;
}

int prt_render_pass(prt_film* f, int32_t pass, prt_pass_stats* st) {
    if (!f) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL film");
    int32_t target = 0;
    int rc = prt_samples_for_pass(&f->params, pass, &target);
    if (rc != PRT_OK) return rc;
    float ms = 0;
    rc = render_to(f, target, &ms);
    if (rc != PRT_OK) return rc;
    fill_stats(f, target, ms, st);
    if (st && f->adaptive) {
        int64_t total = 0;
        int32_t mn = 0, mx = 0;
        if ((rc = count_stats(f, total, mn, mx)) != PRT_OK) return rc;
        st->total_samples = total;
        st->average_samples = (double)total / (double)f->npix;
        st->min_samples = mn;
        st->max_samples_used = mx;
    }
    return PRT_OK;
}

int prt_film_read(const prt_film* f, float* mean_rgb, float* lum_sum, float* lum_sq_sum, uint32_t* counts) {
    if (!f) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL film");
    PRT_HIP(hipSetDevice(f->scene->device));
    std::vector<float4> sum(f->npix);
    std::vector<float2> sq(f->npix);
    std::vector<uint32_t> cnt(f->npix);
    std::vector<long long> spl(f->fb.splat ? 3 * f->npix : 0);
    // ordered after the work already enqueued on the film's stream
    PRT_HIP(hipMemcpyAsync(sum.data(), f->fb.sum, sizeof(float4) * f->npix, hipMemcpyDeviceToHost, f->stream));
    PRT_HIP(hipMemcpyAsync(sq.data(), f->fb.sq, sizeof(float2) * f->npix, hipMemcpyDeviceToHost, f->stream));
    PRT_HIP(hipMemcpyAsync(cnt.data(), f->fb.count, sizeof(uint32_t) * f->npix, hipMemcpyDeviceToHost, f->stream));
    if (f->fb.splat)
        PRT_HIP(hipMemcpyAsync(spl.data(), f->fb.splat, sizeof(long long) * 3 * f->npix, hipMemcpyDeviceToHost,
                               f->stream));
    PRT_HIP(hipStreamSynchronize(f->stream));
    if (mean_rgb) prt::film_mean(f->npix, sum.data(), cnt.data(), f->fb.splat ? spl.data() : nullptr, mean_rgb);
    for (int64_t i = 0; i < f->npix; i++) {
        if (lum_sum) lum_sum[i] = sum[i].w;
        if (lum_sq_sum) lum_sq_sum[i] = sq[i].x;
        if (counts) counts[i] = cnt[i];
    }
    return PRT_OK;
}

int prt_film_read_splats(const prt_film* f, float* splat_rgb) {
    if (!f || !splat_rgb) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!f->fb.splat) {
        std::memset(splat_rgb, 0, sizeof(float) * 3 * (size_t)f->npix);
        return PRT_OK;
    }
    PRT_HIP(hipSetDevice(f->scene->device));
    std::vector<long long> spl(3 * f->npix);
    PRT_HIP(hipMemcpyAsync(spl.data(), f->fb.splat, sizeof(long long) * 3 * f->npix, hipMemcpyDeviceToHost,
                           f->stream));
    PRT_HIP(hipStreamSynchronize(f->stream));
    for (int64_t k = 0; k < 3 * f->npix; k++) splat_rgb[k] = (float)((double)spl[k] * (1.0 / (double)kSplatScale));
    return PRT_OK;
}

int prt_film_probe(prt_film* f, uint32_t flags, prt_film_probe_stats* out) {
    if (!f) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL film");
    if (flags & ~(PRT_PROBE_COUNT | PRT_PROBE_TIME)) return fail(PRT_ERR_INVALID_ARGUMENT, "unknown probe flags");
    PRT_HIP(hipSetDevice(f->scene->device));
    PRT_HIP(hipStreamSynchronize(f->stream));
    prt_film_probe_stats st{};
    if (f->probe_dev) {
        unsigned long long c[kProbeCounts] = {};
        PRT_HIP(hipMemcpy(c, f->probe_dev, sizeof(c), hipMemcpyDeviceToHost));
        uint64_t* o = &st.extension_rays;
        for (int k = 0; k < kProbeCounts; k++) o[k] = c[k];
    }
    st.trace_launches = f->probe_launches;
    double* ms[] = {&st.camera_ms, &st.trace_ms, &st.shade_ms, &st.film_ms, &st.bdpt_paths_ms, &st.bdpt_walks_ms,
                    &st.bdpt_connect_ms};
    for (auto& m : f->marks) {
        float t = 0;
        if (m.a && m.b && hipEventElapsedTime(&t, m.a, m.b) == hipSuccess && m.kind >= 0 && m.kind < 7)
            *ms[m.kind] += t;
        if (m.a) (void)hipEventDestroy(m.a);
        if (m.b) (void)hipEventDestroy(m.b);
    }
    f->marks.clear();
    f->probe_launches = 0;
    if ((flags & PRT_PROBE_COUNT) && !f->probe_dev) {
        void* p = nullptr;
        int rc = film_alloc(f, sizeof(unsigned long long) * kProbeCounts, &p);
        if (rc != PRT_OK) return rc;
        f->probe_dev = (unsigned long long*)p;
    }
    if (f->probe_dev) PRT_HIP(hipMemset(f->probe_dev, 0, sizeof(unsigned long long) * kProbeCounts));
    f->probe_flags = flags;
    if (out) *out = st;
    return PRT_OK;
}

int prt_film_read_rgba8(const prt_film* f, int32_t x0, int32_t y0, int32_t x1, int32_t y1, uint8_t* rgba) {
    if (!f || !rgba) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL argument");
    if (x0 < 0 || y0 < 0 || x1 > f->width || y1 > f->height || x0 > x1 || y0 > y1)
        return fail(PRT_ERR_INVALID_ARGUMENT, "rectangle outside the image");
    const int32_t tw = x1 - x0, th = y1 - y0;
    const int64_t n = (int64_t)tw * th;
    if (n == 0) return PRT_OK;
    prt_film* m = const_cast<prt_film*>(f);
    std::lock_guard<std::mutex> lock(m->rgba_mu);
    PRT_HIP(hipSetDevice(f->scene->device));
    const size_t bytes = sizeof(uchar4) * (size_t)n;
    void* stage = nullptr;
    int rc = pinned_take(bytes, &stage);
    if (rc != PRT_OK) return rc;
    // ordered after the work already enqueued on the film's stream
    hipLaunchKernelGGL(k_film_rgba8, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, f->stream,
                       f->width, x0, y0, tw, n, f->fb.sum, f->fb.count, f->fb.splat, m->rgba_dev);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(stage, m->rgba_dev, bytes, hipMemcpyDeviceToHost, f->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(f->stream);
    if (e == hipSuccess) std::memcpy(rgba, stage, bytes);
    pinned_give(stage, bytes);
    if (e != hipSuccess) return hip_fail(e, "RGBA8 read");
    return PRT_OK;
}

}  // extern "C"
