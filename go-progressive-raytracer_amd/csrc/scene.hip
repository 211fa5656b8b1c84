// scene.hip — scene flattening/upload (replaces Scene.Preprocess, pkg/scene/scene.go:45-78) and
// the batched closest-hit / any-hit entry points (replace BVH.Hit, pkg/geometry/bvh.go:183-239).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "bvh_build.h"
#include "prt_internal.h"
#include "traverse.h"

namespace prt {

namespace {

struct D3 {
    double x, y, z;
};
D3 d3(const double* p) { return {p[0], p[1], p[2]}; }
D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double len(D3 a) { return std::sqrt(dot(a, a)); }
D3 normalize(D3 a) {
    double l = len(a);
    if (l == 0) return {0, 0, 0};
    return {a.x / l, a.y / l, a.z / l};
}
double comp(D3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// Vec3.Rotate X->Y->Z (pkg/core/vec3.go:162-198)
D3 rotate(D3 v, D3 r) {
    if (r.x == 0 && r.y == 0 && r.z == 0) return v;
    if (r.x != 0) {
        double c = std::cos(r.x), s = std::sin(r.x);
        v = {v.x, v.y * c - v.z * s, v.y * s + v.z * c};
    }
    if (r.y != 0) {
        double c = std::cos(r.y), s = std::sin(r.y);
        v = {v.x * c + v.z * s, v.y, -v.x * s + v.z * c};
    }
    if (r.z != 0) {
        double c = std::cos(r.z), s = std::sin(r.z);
        v = {v.x * c - v.y * s, v.x * s + v.y * c, v.z};
    }
    return v;
}

struct QuadF64 {
    D3 corner, u, v;
    int mat;
};

// Quad.BoundingBox with the +-0.001 slab for axis-aligned quads (quad.go:44-77,183-210).
// A disc in float64 with the reference's in-plane frame (NewDisc, disc.go:20-43).
struct DiscF64 {
    D3 c, n, right, up;
    double r;
    int mat;
};
DiscF64 make_disc(D3 c, D3 normal, double r, int mat) {
    DiscF64 dc{c, normalize(normal), {}, {}, r, mat};
    D3 a = std::fabs(dc.n.x) > 0.1 ? D3{0, 1, 0} : D3{1, 0, 0};
    dc.right = normalize(cross(a, dc.n));
    dc.up = normalize(cross(dc.n, dc.right));
    return dc;
}
// Disc.BoundingBox: the box of the four corners c +- r*right +- r*up (disc.go:91-118).
PrimBounds disc_bounds(const DiscF64& dc) {
    D3 re = dc.right * dc.r, ue = dc.up * dc.r;
    D3 cs[4] = {dc.c + re + ue, dc.c + re - ue, dc.c - re + ue, dc.c - re - ue};
    PrimBounds b;
    for (int a = 0; a < 3; a++) {
        b.lo[a] = b.hi[a] = comp(cs[0], a);
        for (int k = 1; k < 4; k++) {
            b.lo[a] = std::min(b.lo[a], comp(cs[k], a));
            b.hi[a] = std::max(b.hi[a], comp(cs[k], a));
        }
    }
    return b;
}

PrimBounds quad_bounds(const QuadF64& q) {
    D3 n = normalize(cross(q.u, q.v));
    D3 cs[4] = {q.corner, q.corner + q.u, q.corner + q.v, q.corner + q.u + q.v};
    PrimBounds b;
    for (int a = 0; a < 3; a++) {
        b.lo[a] = b.hi[a] = comp(cs[0], a);
        for (int k = 1; k < 4; k++) {
            b.lo[a] = std::min(b.lo[a], comp(cs[k], a));
            b.hi[a] = std::max(b.hi[a], comp(cs[k], a));
        }
    }
    const double thr = 0.9999, tol = 0.0001;
    int al = -1;
    if (std::fabs(n.x) > thr && std::fabs(n.y) < tol && std::fabs(n.z) < tol) al = 0;
    else if (std::fabs(n.y) > thr && std::fabs(n.x) < tol && std::fabs(n.z) < tol) al = 1;
    else if (std::fabs(n.z) > thr && std::fabs(n.x) < tol && std::fabs(n.y) < tol) al = 2;
    if (al >= 0) {
        double f = comp(cs[0], al);
        b.lo[al] = f - 0.001;
        b.hi[al] = f + 0.001;
    }
    return b;
}

void grow(PrimBounds& w, const PrimBounds& b) {
    for (int a = 0; a < 3; a++) {
        w.lo[a] = std::min(w.lo[a], b.lo[a]);
        w.hi[a] = std::max(w.hi[a], b.hi[a]);
    }
}

float4 f4(double x, double y, double z, float w) { return make_float4((float)x, (float)y, (float)z, w); }
float bits(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

}  // namespace

// NewCamera (pkg/geometry/camera.go:70-170), float64 then rounded to float32. force_height > 0
// replaces int(width / aspect) (used when a film overrides the image size).
DevCamera make_camera(const prt_camera& c, int& height, int force_height) {
    const double pi = 3.14159265358979323846;
    D3 C = d3(c.center), L = d3(c.look_at), U = d3(c.up);
    D3 w = normalize(C - L);
    D3 u = normalize(cross(w, U));
    D3 v = cross(w, u);
    double focus = c.focus_distance;
    if (focus <= 0) focus = len(C - L);
    double lr = c.aperture / 2;
    double la = pi * lr * lr;
    if (c.aperture == 0) la = 1.0;
    int W = c.width;
    int H = force_height > 0 ? force_height : (int)((double)W / c.aspect_ratio);
    double theta = c.vfov * pi / 180.0;
    double hh = std::tan(theta / 2.0);
    double vh = 2.0 * hh * focus, vw = vh * c.aspect_ratio;
    double vh1 = 2.0 * std::tan(theta / 2.0);
    double plane = (vh1 * c.aspect_ratio) * vh1;
    D3 VU = u * vw, VV = v * vh;
    D3 du = VU * (1.0 / (double)W), dv = VV * (1.0 / (double)H);
    D3 ul = C - w * focus - VU * 0.5 - VV * 0.5;
    D3 p00 = ul + (du + dv) * 0.5;
    D3 disk_u = u * (c.aperture / 2), disk_v = v * (c.aperture / 2);
    D3 fwd = normalize(L - C);
    D3 corner = normalize(D3{-vw / 2, -vh / 2, -focus});
    double ctw = dot(corner, D3{0, 0, -1});
    DevCamera k{};
    auto put = [](float* o, D3 a) {
        o[0] = (float)a.x;
        o[1] = (float)a.y;
        o[2] = (float)a.z;
    };
    put(k.center, C);
    put(k.pixel00, p00);
    put(k.du, du);
    put(k.dv, dv);
    put(k.disk_u, disk_u);
    put(k.disk_v, disk_v);
    put(k.u, u);
    put(k.v, v);
    put(k.w, w);
    put(k.forward, fwd);
    k.lens_radius = (float)lr;
    k.lens_area = (float)la;
    k.vw = (float)vw;
    k.vh = (float)vh;
    k.plane_area = (float)plane;
    k.cos_total_width = (float)ctw;
    k.focus = (float)focus;
    k.width = W;
    k.height = H;
    k.has_lens = len(disk_u) > 0 ? 1 : 0;
    height = H;
    return k;
}

namespace {
}  // namespace

int scene_alloc(prt_scene* s, size_t bytes, const void* src, void** out) {
    void* p = nullptr;
    if (bytes == 0) bytes = 16;
    PRT_HIP(hipMalloc(&p, bytes));
    s->bufs.push_back({p, bytes});
    s->info.device_bytes += (int64_t)bytes;
    if (src) PRT_HIP(hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
    *out = p;
    return PRT_OK;
}

int persistent_blocks(const prt_scene* s, int block, int per_cu) {
    (void)block;
    return std::max(1, s->num_cus) * per_cu;
}

// ---- intersection kernels -------------------------------------------------------------
constexpr int kTraceBlock = 256;
constexpr int kTraceLds = 16;

struct ScratchState {
    std::mutex mu;
    hipEvent_t last = nullptr;
};

template <bool ANY, bool MIXED, bool COUNT>
__global__ __launch_bounds__(kTraceBlock) void k_intersect(DevScene sc, const prt_ray* __restrict__ rays,
                                                           int64_t n, prt_hit* __restrict__ hits,
                                                           uint8_t* __restrict__ occluded, uint32_t* cnt_nodes,
                                                           uint32_t* cnt_boxes, uint32_t* cnt_prims, uint2* gstack,
                                                           uint32_t gstride, unsigned long long* work) {
    __shared__ uint32_t s_node[kTraceLds * kTraceBlock];
    __shared__ float s_t[kTraceLds * kTraceBlock];
    TStack<kTraceLds, kTraceBlock> st;
    st.s_node = s_node;
    st.s_t = s_t;
    st.g = gstack;
    st.lane = threadIdx.x;
    st.gslot = blockIdx.x * kTraceBlock + threadIdx.x;
    st.gstride = gstride;
    const uint32_t lane = threadIdx.x & 63u;
    for (;;) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(work, 64ull);
        base = __shfl(base, 0, 64);
        if (base >= (unsigned long long)n) break;
        int64_t i = (int64_t)base + lane;
        if (i < n) {
            const float4* rp = reinterpret_cast<const float4*>(rays + i);
            float4 a = rp[0], b = rp[1];
            TRay r{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
            THit h;
            TCounts c{0, 0, 0};
            bool hit = trace<ANY, MIXED, COUNT, kTraceLds, kTraceBlock>(sc, r, h, st, &c);
            if (COUNT) {
                cnt_nodes[i] = c.nodes;
                cnt_boxes[i] = c.boxes;
                cnt_prims[i] = c.prims;
            } else if (ANY) {
                occluded[i] = hit ? 1 : 0;
            } else {
                prt_hit o;
                o.t = h.t;
                o.prim_id = hit ? (int32_t)__float_as_uint(sc.prims[3 * (size_t)h.prim + 2].w) : -1;
                o.u = hit ? h.u : 0.0f;
                o.v = hit ? h.v : 0.0f;
                reinterpret_cast<float4*>(hits)[i] =
                    make_float4(o.t, __int_as_float(o.prim_id), o.u, o.v);
            }
        }
    }
}

// Persistent closest/any-hit kernel with lane refill: a lane whose ray is finished takes a new
// ray from the global work counter as soon as REFILL lanes of its wave are idle (REFILL = 64
// degenerates to fetching whole 64-ray batches), so long rays do not idle the rest of the wave.
template <bool ANY, bool MIXED, int LDS, int REFILL, int NCACHE, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_intersect_refill(DevScene sc, const prt_ray* __restrict__ rays,
                                                            int64_t n, prt_hit* __restrict__ hits,
                                                            uint8_t* __restrict__ occluded, uint2* gstack,
                                                            uint32_t gstride, unsigned long long* work) {
    __shared__ uint32_t s_node[LDS * BLOCK];
    __shared__ float s_t[LDS * BLOCK];
    __shared__ float4 s_cache[NCACHE > 0 ? NCACHE * 8 : 1];
    TStack<LDS, BLOCK> st;
    st.s_node = s_node;
    st.s_t = s_t;
    st.g = gstack;
    st.lane = threadIdx.x;
    st.gslot = blockIdx.x * BLOCK + threadIdx.x;
    st.gstride = gstride;
    if (NCACHE > 0) {
        load_node_cache<NCACHE, BLOCK>(sc, s_cache);
        st.ncache = s_cache;
    }
    const uint32_t lane = threadIdx.x & 63u;
    TState s;
    bool active = false;
    bool exhausted = false;  // wave-uniform
    int64_t ridx = 0;
    for (;;) {
        uint64_t idle = __ballot(!active);
        int nidle = __popcll(idle);
        if (!exhausted && (nidle >= REFILL || nidle == 64)) {
            uint32_t leader = __ffsll((unsigned long long)idle) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(work, (unsigned long long)nidle);
            base = __shfl(base, leader, 64);
            if (base + (unsigned long long)nidle >= (unsigned long long)n) exhausted = true;
            if (!active) {
                uint32_t rank = __popcll(idle & ((1ull << lane) - 1ull));
                int64_t i = (int64_t)base + rank;
                if (i < n) {
                    const float4* rp = reinterpret_cast<const float4*>(rays + i);
                    float4 a = rp[0], b = rp[1];
                    tstate_init(s, TRay{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w});
                    ridx = i;
                    active = true;
                }
            }
        }
        if (__ballot(active) == 0) {
            if (exhausted) break;
            continue;
        }
        if (active && tstate_step<ANY, MIXED, false, LDS, BLOCK, NCACHE>(sc, s, st, nullptr)) {
            active = false;
            bool hit = s.best != 0xFFFFFFFFu;
            if (ANY) {
                occluded[ridx] = hit ? 1 : 0;
            } else {
                int32_t pid = hit ? (int32_t)__float_as_uint(sc.prims[3 * (size_t)s.best + 2].w) : -1;
                reinterpret_cast<float4*>(hits)[ridx] =
                    make_float4(hit ? s.tbest : __int_as_float(0x7f800000), __int_as_float(pid),
                                hit ? s.bu : 0.0f, hit ? s.bv : 0.0f);
            }
        }
    }
}

// Persistent closest/any-hit kernel with prefetched lane refill. Ray lengths are heavy-tailed
// (most rays of a uniform distribution leave the root box at once, a few visit dozens of nodes), so
// whole-wave batches idle most lanes. Here every lane holds the NEXT ray in registers, loaded while
// the current one traverses: a lane that finishes starts its next ray in the same step with no wait,
// and reloads its next-ray registers from the wave's ray chunk. Each wave's first chunk is static; later
// ones are reserved one ahead with an atomic on its part's work counter whose result is read only when the
// current chunk runs out, so neither the ray loads nor the work counters add latency to the steps.
#ifndef PRT_PF_CHUNK
#define PRT_PF_CHUNK 256
#endif
constexpr uint32_t kPfChunk = PRT_PF_CHUNK;

// Work counters. Every chunk reservation is an atomic add on a counter, and one address sustains about
// one such add per 14-20 ns device-wide: with one counter for the whole batch, 256-ray chunks of a
// 16.8M-ray batch (65,536 reservations per ~1.3 ms launch) already kept it near saturation, and smaller
// chunks made the launch slower (64 rays: 3.5 ms). The batch is split into kWorkParts contiguous parts,
// each with its own counter on its own 256-B line (kWorkStride 64-bit words apart); work-group b starts
// on part b % kWorkParts (work-groups are dispatched to the 8 XCDs in turn) and, once that part is used
// up, reads every part's counter in one round trip and takes chunks of the next part with rays left.
constexpr uint32_t kWorkParts = 8;
constexpr uint32_t kWorkStride = 32;
constexpr size_t kWorkBytes = 2 * (size_t)kWorkParts * kWorkStride * sizeof(unsigned long long);

// Rays per chunk: `rays` / `parts` in multiples of 64, from one ray per lane up to kPfChunk.
__device__ __forceinline__ unsigned long long pf_chunk(unsigned long long rays, unsigned long long parts) {
    // 32-bit division: a quotient at or above kPfChunk is clamped anyway
    const uint32_t r = (uint32_t)min(rays, 0xFFFFFFFFull);
    const uint32_t d = (uint32_t)max(1ull, min(parts, 0xFFFFFFFFull));
    const uint32_t c = ((r / d) + 63u) & ~63u;
    return c < 64u ? 64ull : (c > kPfChunk ? (unsigned long long)kPfChunk : (unsigned long long)c);
}

// Part p of a batch of N rays: rays [lo, hi); its waves' first chunks (static, c0 rays each, at most half
// the part over its waves) come first, then the chunks of its counter, from `dyn`.
struct PfPart {
    unsigned long long lo, hi, dyn, waves, c0;
};
__device__ __forceinline__ PfPart pf_part(unsigned long long N, uint32_t p, uint32_t grid, uint32_t waves_per_block) {
    PfPart q;
    q.lo = N * p / kWorkParts;
    q.hi = N * (p + 1) / kWorkParts;
    const unsigned long long blocks = grid / kWorkParts + (p < grid % kWorkParts ? 1u : 0u);
    q.waves = blocks * waves_per_block;
    q.c0 = pf_chunk(q.hi - q.lo, 2ull * q.waves);
    q.dyn = min(q.lo + q.waves * q.c0, q.hi);
    return q;
}
typedef float f32x4 __attribute__((ext_vector_type(4)));  // a hit record, for vector non-temporal stores

// Lane 0's value, as a wave-uniform (scalar) value.
__device__ __forceinline__ unsigned long long lane0_u64(unsigned long long x) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, 0);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), 0);
    return ((unsigned long long)hi << 32) | lo;
}

// The part table follows the counters, one 256-B line per part (words 1-5: dyn, hi, lo, c0, waves), so that
// reading it does not touch the lines the counters' atomics work on.
__device__ __forceinline__ PfPart pf_part_read(const unsigned long long* work, uint32_t p) {
    const unsigned long long* w = work + (kWorkParts + p) * kWorkStride;
    PfPart q;
    q.dyn = lane0_u64(w[1]);
    q.hi = lane0_u64(w[2]);
    q.lo = lane0_u64(w[3]);
    q.c0 = lane0_u64(w[4]);
    q.waves = lane0_u64(w[5]);
    return q;
}

// Size of the next chunk taken from a part whose counter is past `pos` (guided self-scheduling): a quarter
// of a wave's share of the part's rays left, kPfChunk while much is left, down to one ray per lane, so
// that the part's waves run out of rays at about the same time. (A wave runs through its chunk before it
// looks for more; with fixed 256-ray chunks the waves ran out up to four rays per lane apart.)
__device__ __forceinline__ unsigned long long pf_next_chunk(const PfPart& q, unsigned long long pos) {
    return pf_chunk(q.hi > pos ? q.hi - pos : 0ull, 4ull * q.waves);
}

// Triangle-only scenes: 6 waves per SIMD (80 VGPRs, the LDS bound of the 8-entry stack and 64-node cache)
// hide more node-fetch latency than the register count the compiler picks on its own (84, 5 waves).
// Long rays: a lane whose ray is still traversing after `budget` steps (0: never) parks it in the park
// queue (up to its capacity; a refused ray keeps traversing) with its best hit so far in the hit record
// (leaf-ordered primitive index), and takes its next ray; k_intersect_long then finishes every parked ray
// with a whole wave. A ray's result does not depend on which kernel finishes it.
struct ParkQueue {
    uint32_t* count;  // on its own 256-B line
    uint32_t* q;      // ray indices
    uint32_t cap;
    uint32_t budget;
};
constexpr uint32_t kParkCapacity = 1u << 18;

// Step budget of a batch of n rays on a grid of `lanes` threads: a wave per parked ray pays off where
// lanes would otherwise idle, so batches of a few rays per lane (latency-bound, their time set by the
// longest rays) park early and large batches park only their longest rays. PRT_TRACE_PARK_STEPS pins
// one budget for every batch (0: no parking). Rays are indexed with 32 bits in the park queue: batches
// beyond 2^32 rays do not park.
uint32_t park_budget(const prt_scene* s, int64_t n, int64_t lanes) {
    if (n > (int64_t)0xFFFFFFFFll) return 0u;
    if (s->park_steps >= 0) return (uint32_t)s->park_steps;
    const int64_t per_lane = n / std::max<int64_t>(1, lanes);
    return per_lane <= 1 ? 32u : (per_lane <= 16 ? 48u : 96u);
}
constexpr int kLongLds = 8;  // LDS stack entries per lane of k_intersect_long (deeper entries spill)
constexpr int kLongLanes = 64;  // lanes per parked ray in k_intersect_long (large batches park few rays)

template <bool ANY, bool MIXED, int LDS, int NCACHE, int BLOCK, int GATE = 0>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(MIXED ? 1 : 6))) void k_intersect_pf(DevScene sc, const prt_ray* __restrict__ rays, int64_t n,
                                                        prt_hit* __restrict__ hits, uint8_t* __restrict__ occluded,
                                                        uint2* gstack, uint32_t gstride, unsigned long long* work,
                                                        ParkQueue park) {
    __shared__ uint32_t s_node[LDS * BLOCK];
    __shared__ float s_t[LDS * BLOCK];
    __shared__ float4 s_cache[NCACHE > 0 ? NCACHE * 8 : 1];
    TStack<LDS, BLOCK> st;
    st.s_node = s_node;
    st.s_t = s_t;
    st.g = gstack;
    st.lane = threadIdx.x;
    st.gslot = blockIdx.x * BLOCK + threadIdx.x;
    st.gstride = gstride;
    if (NCACHE > 0) {
        load_node_cache<NCACHE, BLOCK>(sc, s_cache);
        st.ncache = s_cache;
    }
    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long N = (unsigned long long)n;
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    // wave-uniform chunk state: current chunk [cpos, cend) of part `part`; lane 0's pending atomic result
    // (psize rays of the part's counter) locates the next one
    // a wave's first chunk is static (its index among its part's waves); later chunks come from the
    // part's counter, reserved one ahead (no wave waits on an atomic to start), then from the next parts
    // a wave's first chunk is static (its index among its part's waves); later chunks come from the
    // part's counter, reserved one ahead (no wave waits on an atomic to start), then from other parts
    uint32_t part = blockIdx.x % kWorkParts;
    PfPart q = pf_part_read(work, part);
    // the wave's index within its work-group, as a scalar (the chunk state is wave-uniform)
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const unsigned long long wid = (unsigned long long)(blockIdx.x / kWorkParts) * (BLOCK / 64) + wave;
    unsigned long long cpos = min(q.lo + wid * q.c0, q.hi), cend = min(cpos + q.c0, q.hi), nxt = 0;
    unsigned long long pend = 0, psize = pf_next_chunk(q, q.dyn);
    bool exhausted = false;
    if (lane == 0) pend = atomicAdd(work + part * kWorkStride, psize);
    TState s;
    bool active = false;
    int64_t ridx = 0, nidx = -1;
    uint32_t steps = 0;
    float4 na = make_float4(0.f, 0.f, 0.f, 0.f), nb = na;
    for (;;) {
        // start the prefetched ray on an idle lane
        if (!active && nidx >= 0) {
            tstate_init(s, TRay{na.x, na.y, na.z, na.w, nb.x, nb.y, nb.z, nb.w});
            ridx = nidx;
            nidx = -1;
            active = true;
            steps = 0;
        }
        // hand out the next rays of the chunk to lanes without a prefetched ray
        uint64_t need = __ballot(nidx < 0);
        while (need != 0 && !exhausted) {
            if (cpos >= cend) {
                nxt = q.dyn + lane0_u64(pend);
                if (nxt >= q.hi) {
                    // this part is used up: lane k reads part k's counter (all in one round trip) and
                    // the wave takes one-ray-per-lane chunks of the next part that has rays left
                    // (the empty asm keeps the lane-indexed addressing here instead of hoisted into
                    // registers held across the traversal loop)
                    uint32_t k = lane;
                    asm volatile("" : "+v"(k));
                    bool left = false;
                    if (k < kWorkParts) {
                        const unsigned long long* wk = work + k * kWorkStride;
                        const unsigned long long taken =
                            __hip_atomic_load(wk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        left = wk[kWorkParts * kWorkStride + 1] + taken < wk[kWorkParts * kWorkStride + 2];
                    }
                    const uint32_t open = (uint32_t)__ballot(left);
                    if (open == 0u) {
                        exhausted = true;
                        break;
                    }
                    // the first open part after this one, cyclically
                    const uint32_t rot = (open >> (part + 1)) | (open << (kWorkParts - part - 1));
                    part = (part + 1 + (uint32_t)__builtin_ctz(rot & ((1u << kWorkParts) - 1u))) % kWorkParts;
                    q = pf_part_read(work, part);
                    psize = 64;
                    if (lane == 0) pend = atomicAdd(work + part * kWorkStride, psize);
                    continue;
                }
                cpos = nxt;
                cend = min(nxt + psize, q.hi);
                psize = pf_next_chunk(q, cend);
                if (lane == 0) pend = atomicAdd(work + part * kWorkStride, psize);
                continue;
            }
            const uint32_t avail = (uint32_t)(cend - cpos);
            const uint32_t rank = __popcll(need & lt_mask);
            const bool mine = ((need >> lane) & 1ull) != 0 && rank < avail;
            if (mine) {
                nidx = (int64_t)(cpos + rank);
                const float4* rp = reinterpret_cast<const float4*>(rays + nidx);
                na = rp[0];
                nb = rp[1];
            }
            const uint32_t took = min((uint32_t)__popcll(need), avail);
            cpos += took;
            need = __ballot(nidx < 0);
        }
        if (__ballot(active || nidx >= 0) == 0) break;  // exhausted and drained
        const bool leaf_ok = GATE > 0 ? leaf_gate(active, s.cur, GATE) : true;
        if (active) {
            // a lane held at a leaf by the gate makes no progress this step and does not count it
            const bool progress = leaf_ok || !child_is_leaf(s.cur);
            if (tstate_step<ANY, MIXED, false, LDS, BLOCK, NCACHE>(sc, s, st, nullptr, leaf_ok)) {
                active = false;
                bool hit = s.best != 0xFFFFFFFFu;
                // results are written with non-temporal stores: nothing here reads them again, and
                // streaming them past the L2 leaves its lines to the nodes (16.8M-ray launch: 0.53 ->
                // 0.47 GB written, 1.195 -> 1.144 ms)
                if (ANY) {
                    __builtin_nontemporal_store((uint8_t)(hit ? 1 : 0), occluded + ridx);
                } else {
                    int32_t pid = hit ? (int32_t)__float_as_uint(sc.prims[3 * (size_t)s.best + 2].w) : -1;
                    const f32x4 hv = {hit ? s.tbest : __int_as_float(0x7f800000), __int_as_float(pid),
                                      hit ? s.bu : 0.0f, hit ? s.bv : 0.0f};
                    __builtin_nontemporal_store(hv, reinterpret_cast<f32x4*>(hits) + ridx);
                }
            } else if (progress && ++steps == park.budget) {
                const uint32_t k = atomicAdd(park.count, 1u);
                if (k < park.cap) {
                    park.q[k] = (uint32_t)ridx;
                    if (!ANY)
                        reinterpret_cast<float4*>(hits)[ridx] =
                            make_float4(s.tbest, __uint_as_float(s.best), s.bu, s.bv);
                    active = false;
                }
            }
        }
    }
}

// Finishes the rays k_intersect_pf parked, 64 / kLongLanes per one-wave work-group (coop_trace with groups
// of kLongLanes lanes), and writes their results as k_intersect_pf would have.
template <bool ANY, bool MIXED>
__global__ __launch_bounds__(64) void k_intersect_long(DevScene sc, const prt_ray* __restrict__ rays,
                                                       prt_hit* __restrict__ hits, uint8_t* __restrict__ occluded,
                                                       uint2* gstack, uint32_t gstride, ParkQueue park) {
    constexpr uint32_t R = 64 / kLongLanes;  // rays per work-group and round
    __shared__ uint32_t s_node[kLongLds * 64];
    __shared__ float s_t[kLongLds * 64];
    __shared__ uint32_t f_node[kCoopEntries];
    __shared__ float f_t[kCoopEntries];
    const uint32_t n = min(*park.count, park.cap);
    if (blockIdx.x * R >= n) return;
    TStack<kLongLds, 64> st;
    st.s_node = s_node;
    st.s_t = s_t;
    st.g = gstack;
    st.lane = threadIdx.x;
    st.gslot = blockIdx.x * 64 + threadIdx.x;
    st.gstride = gstride;
    for (uint32_t base = blockIdx.x * R; base < n; base += gridDim.x * R) {  // uniform over the work-group
        const uint32_t i = base + threadIdx.x / (uint32_t)kLongLanes;
        const bool has = i < n;
        const uint32_t ridx = has ? park.q[i] : 0u;
        float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a;
        THit h;
        h.t = __int_as_float(0x7f800000);
        h.prim = 0xFFFFFFFFu;
        h.u = h.v = 0.0f;
        if (has) {
            const float4* rp = reinterpret_cast<const float4*>(rays + ridx);
            a = rp[0];
            b = rp[1];
            if (!ANY) {
                const float4 hp = reinterpret_cast<const float4*>(hits)[ridx];
                h.t = hp.x;
                h.prim = __float_as_uint(hp.y);
                h.u = hp.z;
                h.v = hp.w;
            }
        }
        const TRay r{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        const bool hit = coop_trace<MIXED, kLongLds, kLongLanes>(sc, r, h, st, f_node, f_t, has, ANY);
        if (has && (threadIdx.x & (uint32_t)(kLongLanes - 1)) == 0) {
            if (ANY) {
                occluded[ridx] = hit ? 1 : 0;
            } else {
                const int32_t pid = hit ? (int32_t)__float_as_uint(sc.prims[3 * (size_t)h.prim + 2].w) : -1;
                reinterpret_cast<float4*>(hits)[ridx] = make_float4(hit ? h.t : __int_as_float(0x7f800000),
                                                                    __int_as_float(pid), hit ? h.u : 0.0f,
                                                                    hit ? h.v : 0.0f);
            }
        }
        __syncthreads();  // the frontier buffers are reused by the next rays
    }
}

// Zeroes the work counters and the park count before an intersection launch
// and, for k_intersect_pf, writes the part table of a batch of n rays on `grid` work-groups of
// `waves_per_block` waves (one thread per part).
__global__ void k_clear_counters(unsigned long long* work, uint32_t* park_count, unsigned long long n, uint32_t grid,
                                 uint32_t waves_per_block) {
    const uint32_t p = threadIdx.x;
    if (p < kWorkParts) {
        const PfPart q = pf_part(n, p, grid, waves_per_block);
        work[p * kWorkStride] = 0ull;
        unsigned long long* w = work + (kWorkParts + p) * kWorkStride;
        w[1] = q.dyn;
        w[2] = q.hi;
        w[3] = q.lo;
        w[4] = q.c0;
        w[5] = q.waves;
    }
    if (p == 0) *park_count = 0u;
}

// Kernel variants the per-scene autotuner chooses from (identical per-ray traversal): kernel
// (0 = batch / threshold refill, 1 = prefetched refill), LDS stack entries per lane, refill
// threshold (kernel 0; 64 = whole-wave batches) or leaf gate (kernel 1; a leaf step waits for gate/8 of
// the running lanes, 0 = no gate), LDS node-cache size, block.
#define PRT_TRACE_VARIANTS(X)      \
    X(0, 0, 8, 64, 0, 256)         \
    X(1, 1, 8, 0, 0, 256)          \
    X(2, 1, 8, 0, 64, 256)         \
    X(3, 1, 12, 0, 0, 256)         \
    X(4, 1, 6, 0, 64, 512)         \
    X(5, 1, 8, 2, 64, 256)         \
    X(6, 1, 8, 4, 64, 256)         \
    X(7, 1, 8, 8, 64, 256)
constexpr int kNumVariants = 8;
constexpr int kVariantLds[kNumVariants] = {
#define X(id, kind, lds, refill, nc, blk) lds,
    PRT_TRACE_VARIANTS(X)
#undef X
};
constexpr int kMinLds = 6;
static_assert(kMinLds <= kLongLds, "the scene's spill stack must cover k_intersect_long");
constexpr int kDefaultVariant = 5;
constexpr int kVariantBlock[kNumVariants] = {
#define X(id, kind, lds, refill, nc, blk) blk,
    PRT_TRACE_VARIANTS(X)
#undef X
};

template <bool ANY, bool MIXED, int KIND, int LDS, int REFILL, int NCACHE, int BLOCK>
const void* kernel_of() {
    if (KIND == 1) return (const void*)k_intersect_pf<ANY, MIXED, LDS, NCACHE, BLOCK, REFILL>;
    return (const void*)k_intersect_refill<ANY, MIXED, LDS, REFILL, NCACHE, BLOCK>;
}

template <bool ANY, bool MIXED>
const void* variant_fn(int v) {
    switch (v) {
#define X(id, kind, lds, refill, nc, blk) \
    case id:                              \
        return kernel_of<ANY, MIXED, kind, lds, refill, nc, blk>();
        PRT_TRACE_VARIANTS(X)
#undef X
    }
    return nullptr;
}

// Launches variant v; the prefetched-refill variants park long rays (park.budget > 0) and are followed
// by k_intersect_long on `long_blocks` one-wave work-groups (at most gstride / 64).
// Returns the launch error of the variant's kernel, then of k_intersect_long.
template <bool ANY, bool MIXED>
hipError_t launch_variant(int v, int blocks, hipStream_t st, const DevScene& sc, const prt_ray* rays, int64_t n,
                          prt_hit* hits, uint8_t* occ, uint2* gs, uint32_t gstride, unsigned long long* work,
                          ParkQueue park, int long_blocks) {
    bool pf = false;
    switch (v) {
#define X(id, kind, lds, refill, nc, blk)                                                                     \
    case id:                                                                                                \
        if (kind == 1) {                                                                                    \
            hipLaunchKernelGGL((k_intersect_pf<ANY, MIXED, lds, nc, blk, refill>), dim3(blocks), dim3(blk), 0, st, sc, \
                               rays, n, hits, occ, gs, gstride, work, park);                                \
            pf = true;                                                                                      \
        } else {                                                                                            \
            hipLaunchKernelGGL((k_intersect_refill<ANY, MIXED, lds, refill, nc, blk>), dim3(blocks),          \
                               dim3(blk), 0, st, sc, rays, n, hits, occ, gs, gstride, work);                \
        }                                                                                                   \
        break;
        PRT_TRACE_VARIANTS(X)
#undef X
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (pf && park.budget > 0) {
        hipLaunchKernelGGL((k_intersect_long<ANY, MIXED>), dim3(long_blocks), dim3(64), 0, st, sc, rays, hits, occ, gs,
                           gstride, park);
        e = hipGetLastError();
    }
    return e;
}

// Resident one-wave work-groups of k_intersect_long, at most `slots` / 64 (its stack slots).
int long_blocks(const prt_scene* s, uint32_t slots) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_intersect_long<false, true>, 64, 0) !=
            hipSuccess ||
        per_cu <= 0) {
        (void)hipGetLastError();  // the query's failure is handled here, not by the next launch check
        per_cu = 16;
    }
    return (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)std::max(1, s->num_cus) * std::min(per_cu, 32),
                                                       slots / 64));
}

}  // namespace prt

using namespace prt;

namespace {
std::mutex g_scratch_mu;
}

static int launch_intersect(const prt_scene* scene, int mode, const prt_ray* rays, int64_t n, void* out0,
                            uint32_t* c1, uint32_t* c2, uint32_t flags, void* stream) {
    if (!scene) return fail(PRT_ERR_INVALID_ARGUMENT, "scene is NULL");
    if (n < 0) return fail(PRT_ERR_INVALID_ARGUMENT, "negative ray count");
    if (n == 0) return PRT_OK;
    if (!rays || !out0) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (flags > PRT_BUFFERS_DEVICE) return fail(PRT_ERR_INVALID_ARGUMENT, "unknown flags");
    PRT_HIP(hipSetDevice(scene->device));
    hipStream_t st = (hipStream_t)stream;
    const bool dev = flags == PRT_BUFFERS_DEVICE;
    const size_t out_elem = mode == 0 ? sizeof(prt_hit) : (mode == 1 ? 1 : 4);
    void* d_rays = (void*)rays;
    void* d_out0 = out0;
    uint32_t* d_c1 = c1;
    uint32_t* d_c2 = c2;
    std::vector<void*> temps;
    auto cleanup = [&]() {
        for (void* p : temps) (void)hipFree(p);
    };
    if (!dev) {
        hipError_t e;
        e = hipMalloc(&d_rays, (size_t)n * sizeof(prt_ray));
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(rays)");
        temps.push_back(d_rays);
        e = hipMalloc(&d_out0, (size_t)n * out_elem);
        if (e != hipSuccess) { cleanup(); return hip_fail(e, "hipMalloc(out)"); }
        temps.push_back(d_out0);
        if (mode == 2) {
            e = hipMalloc((void**)&d_c1, (size_t)n * 4);
            if (e != hipSuccess) { cleanup(); return hip_fail(e, "hipMalloc(counts)"); }
            temps.push_back(d_c1);
            e = hipMalloc((void**)&d_c2, (size_t)n * 4);
            if (e != hipSuccess) { cleanup(); return hip_fail(e, "hipMalloc(counts)"); }
            temps.push_back(d_c2);
        }
        e = hipMemcpyAsync(d_rays, rays, (size_t)n * sizeof(prt_ray), hipMemcpyHostToDevice, st);
        if (e != hipSuccess) { cleanup(); return hip_fail(e, "hipMemcpyAsync(rays)"); }
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    // scene scratch (work counter + overflow stack) is shared: order launches on it
    unsigned long long* work = (unsigned long long*)scene->bufs[0].p;
    hipError_t e;
    const DevScene& sc = scene->dev;
    uint2* gs = scene->gstack;
    uint32_t gstride = scene->gstack_slots;
    prt_scene* ms = const_cast<prt_scene*>(scene);
    // park queue: count on the first 256-B line, then the ray indices
    ParkQueue park{scene->park, scene->park + 64, kParkCapacity, 0u};
    const int lb = long_blocks(scene, gstride);
    // an error left by an earlier call on this thread (its caller saw it or chose to ignore it) is not
    // this launch's: clear it, so that the checks below report only these launches
    const hipError_t stale = hipGetLastError();
    const char* failed = "intersect kernel launch";
    auto run = [&](int v, int64_t count) -> hipError_t {
        int blocks = ms->variant_blocks[v];
        hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(kWorkParts), 0, st, work, park.count,
                           (unsigned long long)count, (uint32_t)blocks, (uint32_t)(kVariantBlock[v] / 64));
        hipError_t e0 = hipGetLastError();
        if (e0 != hipSuccess) {
            failed = "counter clear launch";
            return e0;
        }
        park.budget = park_budget(scene, count, (int64_t)blocks * kVariantBlock[v]);
        if (mode == 0) {
            if (sc.mixed) e0 = launch_variant<false, true>(v, blocks, st, sc, (const prt_ray*)d_rays, count, (prt_hit*)d_out0, nullptr, gs, gstride, work, park, lb);
            else e0 = launch_variant<false, false>(v, blocks, st, sc, (const prt_ray*)d_rays, count, (prt_hit*)d_out0, nullptr, gs, gstride, work, park, lb);
        } else {
            if (sc.mixed) e0 = launch_variant<true, true>(v, blocks, st, sc, (const prt_ray*)d_rays, count, nullptr, (uint8_t*)d_out0, gs, gstride, work, park, lb);
            else e0 = launch_variant<true, false>(v, blocks, st, sc, (const prt_ray*)d_rays, count, nullptr, (uint8_t*)d_out0, gs, gstride, work, park, lb);
        }
        return e0;
    };
    if (mode == 2) {
        e = hipMemsetAsync(work, 0, sizeof(unsigned long long), st);
        if (e != hipSuccess) { cleanup(); return hip_fail(e, "hipMemsetAsync(work)"); }
        int blocks = (int)(scene->gstack_slots / kTraceBlock);
        blocks = std::min(blocks, std::max(1, scene->num_cus) * 4);
        if (sc.mixed)
            hipLaunchKernelGGL((k_intersect<false, true, true>), dim3(blocks), dim3(kTraceBlock), 0, st, sc, (const prt_ray*)d_rays, n,
                               nullptr, nullptr, (uint32_t*)d_out0, d_c1, d_c2, gs, gstride, work);
        else
            hipLaunchKernelGGL((k_intersect<false, false, true>), dim3(blocks), dim3(kTraceBlock), 0, st, sc, (const prt_ray*)d_rays, n,
                               nullptr, nullptr, (uint32_t*)d_out0, d_c1, d_c2, gs, gstride, work);
    } else {
        int& tuned = ms->trace_variant[mode];
        if (ms->forced_variant < 0 && tuned < 0 && n >= (1 << 20)) {
            // autotune: time every variant on the first (up to) 8M rays of this batch, keep the fastest
            hipEvent_t a, b;
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            float best = 1e30f;
            int bestv = kDefaultVariant;
            const int64_t m = std::min<int64_t>(n, (int64_t)1 << 23);
            for (int v = 0; v < kNumVariants; v++) {
                float tot = 0;
                for (int rep = 0; rep < 4; rep++) {
                    (void)hipEventRecord(a, st);
                    e = run(v, m);
                    (void)hipEventRecord(b, st);
                    (void)hipEventSynchronize(b);
                    float ms_v = 0;
                    (void)hipEventElapsedTime(&ms_v, a, b);
                    if (rep > 0) tot += ms_v;
                }
                if (e != hipSuccess) break;
                // another variant replaces the default only if it is clearly (3 %) faster
                float score = v == kDefaultVariant ? tot : tot * 1.03f;
                if (score < best) {
                    best = score;
                    bestv = v;
                }
            }
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            if (e != hipSuccess) { cleanup(); return hip_fail(e, "intersect autotune"); }
            tuned = bestv;
            if (mode == 0) ms->info.trace_variant = bestv;
        }
        int v = ms->forced_variant >= 0 ? ms->forced_variant : (tuned >= 0 ? tuned : kDefaultVariant);
        e = run(v, n);
        if (e != hipSuccess) {
            cleanup();
            return fail(PRT_ERR_DEVICE, std::string(failed) + " (variant " + std::to_string(v) + "): " +
                                            hipGetErrorString(e) + (stale != hipSuccess ? std::string(" (earlier error on this thread: ") + hipGetErrorString(stale) + ")" : std::string()));
        }
    }
    e = hipGetLastError();
    if (e != hipSuccess) { cleanup(); return hip_fail(e, "intersect kernel launch"); }
    if (!dev) {
        e = hipMemcpyAsync(out0, d_out0, (size_t)n * out_elem, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && mode == 2) e = hipMemcpyAsync(c1, d_c1, (size_t)n * 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && mode == 2) e = hipMemcpyAsync(c2, d_c2, (size_t)n * 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        cleanup();
        if (e != hipSuccess) return hip_fail(e, "intersect copy-back");
    } else {
        // the shared scratch is reused by the next call: keep calls ordered on the device
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "intersect synchronize");
    }
    return PRT_OK;
}

// ---- allocation-free closest hit (prt_intersect_closest_static) ------------------------------------
// The work counter, the park queue and the traversal overflow stack of the intersection kernels live in
// one static device array instead of scene memory: kWorkBytes for the work counters, the park count
// (256 B) and kStaticParkCapacity ray indices, then the stack, entry e of thread slot g at [e * slots + g] as in
// scene memory. Each device holds its own copy (resolved once per device
// with hipGetSymbolAddress); calls on one device share it, so calls that overlap on one device race.
constexpr size_t kStaticScratchBytes = (size_t)PRT_STATIC_SCRATCH_BYTES;
constexpr uint32_t kStaticParkCapacity = 1u << 16;
constexpr size_t kStaticParkOffset = kWorkBytes;
constexpr size_t kStaticStackOffset = kStaticParkOffset + 256 + sizeof(uint32_t) * (size_t)kStaticParkCapacity;
__device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_static_mu;
void* g_static_ptr[kMaxDevices] = {};


// Device address of this device's `scratch` (the device must be current).
int static_scratch(int device, void** out) {
    if (device < 0 || device >= kMaxDevices) return fail(PRT_ERR_INVALID_ARGUMENT, "device index out of range");
    std::lock_guard<std::mutex> lock(g_static_mu);
    if (!g_static_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress");
        g_static_ptr[device] = p;
    }
    *out = g_static_ptr[device];
    return PRT_OK;
}
}  // namespace

extern "C" {

int prt_scene_create(int32_t device, const prt_scene_desc* d, prt_scene** out) {
    if (!d || !out) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(PRT_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return fail(PRT_ERR_INVALID_ARGUMENT, "device index out of range");
    if (d->num_materials < 0 || d->num_spheres < 0 || d->num_quads < 0 || d->num_boxes < 0 ||
        d->num_meshes < 0 || d->num_lights < 0 || d->num_discs < 0 || d->num_cylinders < 0 || d->num_cones < 0 ||
        d->num_textures < 0)
        return fail(PRT_ERR_INVALID_ARGUMENT, "negative count");
    for (int i = 0; i < d->num_textures; i++)
        if (d->textures[i].width <= 0 || d->textures[i].height <= 0 || !d->textures[i].rgb)
            return fail(PRT_ERR_INVALID_ARGUMENT, "texture needs pixels");
    if (d->camera.width <= 0 || !(d->camera.aspect_ratio > 0))
        return fail(PRT_ERR_INVALID_ARGUMENT, "camera width/aspect must be positive");
    const int nm = d->num_materials;
    auto bad_mat = [&](int m) { return m < 0 || m >= nm; };
    // composite materials (Layered, Mix) take two basic components (Lambertian, Metal, Dielectric,
    // Emissive): the device evaluates one level of composition
    auto basic = [&](int m) { return !bad_mat(m) && d->materials[m].type <= PRT_MAT_EMISSIVE; };
    bool textured = false;
    for (int i = 0; i < nm; i++) {
        const prt_material& m = d->materials[i];
        if (m.type < PRT_MAT_LAMBERTIAN || (m.type > PRT_MAT_LAYERED && m.type != PRT_MAT_MIX))
            return fail(PRT_ERR_INVALID_ARGUMENT, "unknown material type");
        if (m.type == PRT_MAT_LAYERED && !(basic(m.outer) && basic(m.inner)))
            return fail(PRT_ERR_INVALID_ARGUMENT, "layered material needs two basic components");
        if (m.type == PRT_MAT_MIX && !(basic(m.outer) && basic(m.inner)))
            return fail(PRT_ERR_INVALID_ARGUMENT, "mix material needs two basic components");
        if (m.type == PRT_MAT_LAMBERTIAN || m.type == PRT_MAT_METAL) {
            if (m.texture < 0 || m.texture > d->num_textures)
                return fail(PRT_ERR_INVALID_ARGUMENT, "material texture out of range");
            textured |= m.texture > 0;
        }
    }
    PRT_HIP(hipSetDevice(device));
    prt_scene* s = new prt_scene();
    s->device = device;
    int rc = PRT_OK;
    auto bail = [&](int code) {
        prt_scene_destroy(s);
        return code;
    };

    // ---- gather primitives in global id order ----
    std::vector<QuadF64> quads;
    for (int i = 0; i < d->num_quads; i++) {
        const prt_quad& q = d->quads[i];
        if (bad_mat(q.material)) return bail(fail(PRT_ERR_INVALID_ARGUMENT, "quad material out of range"));
        quads.push_back({d3(q.corner), d3(q.u), d3(q.v), q.material});
    }
    std::vector<PrimBounds> world_boxes;  // reference shape boxes (for world centre/radius)
    for (int i = 0; i < d->num_boxes; i++) {  // NewBox / generateFaces box.go:34-89
        const prt_box& b = d->boxes[i];
        if (bad_mat(b.material)) return bail(fail(PRT_ERR_INVALID_ARGUMENT, "box material out of range"));
        D3 c[8] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1}};
        PrimBounds wb;
        for (int k = 0; k < 8; k++) {
            c[k] = D3{c[k].x * b.size[0], c[k].y * b.size[1], c[k].z * b.size[2]};
            c[k] = rotate(c[k], d3(b.rotation));
            c[k] = c[k] + d3(b.center);
            for (int a = 0; a < 3; a++) {
                double v = comp(c[k], a);
                if (k == 0) wb.lo[a] = wb.hi[a] = v;
                wb.lo[a] = std::min(wb.lo[a], v);
                wb.hi[a] = std::max(wb.hi[a], v);
            }
        }
        int m = b.material;
        quads.push_back({c[4], c[5] - c[4], c[7] - c[4], m});
        quads.push_back({c[1], c[0] - c[1], c[2] - c[1], m});
        quads.push_back({c[5], c[1] - c[5], c[6] - c[5], m});
        quads.push_back({c[0], c[4] - c[0], c[3] - c[0], m});
        quads.push_back({c[3], c[7] - c[3], c[2] - c[3], m});
        quads.push_back({c[4], c[0] - c[4], c[5] - c[4], m});
        world_boxes.push_back(wb);
    }
    // discs: the scene's own, then the disc of each disc spot light with an internal emitting
    // material appended after the caller's (scene.go:117-122 adds those discs to the shapes)
    std::vector<DiscF64> discs;
    for (int i = 0; i < d->num_discs; i++) {
        const prt_disc& dc = d->discs[i];
        if (bad_mat(dc.material)) return bail(fail(PRT_ERR_INVALID_ARGUMENT, "disc material out of range"));
        if (!(dc.radius > 0)) return bail(fail(PRT_ERR_INVALID_ARGUMENT, "disc radius must be positive"));
        D3 n = d3(dc.normal);
        if (!(len(n) > 0)) return bail(fail(PRT_ERR_INVALID_ARGUMENT, "disc normal must be non-zero"));
        discs.push_back(make_disc(d3(dc.center), n, dc.radius, dc.material));
    }
    // cylinders and cones (NewCylinder cylinder.go:24-39, NewCone cone.go:27-80 with its checks)
    struct AxialF64 {
        D3 base, top, axis;
        double rb, rt, height;
        bool capped, cone;
        int mat;
    };
    std::vector<AxialF64> axials;
    for (int i = 0; i < d->num_cylinders; i++) {
        const prt_cylinder& c = d->cylinders[i];
        if (bad_mat(c.material)) return bail(fail(PRT_ERR_INVALID_ARGUMENT, "cylinder material out of range"));
        D3 b = d3(c.base_center), t = d3(c.top_center);
        double h = len(t - b);
        if (!(c.radius > 0) || !(h > 0))
            return bail(fail(PRT_ERR_INVALID_ARGUMENT, "cylinder needs a positive radius and height"));
        axials.push_back({b, t, (t - b) * (1.0 / h), c.radius, c.radius, h, c.capped != 0, false, c.material});
    }
    for (int i = 0; i < d->num_cones; i++) {
        const prt_cone& c = d->cones[i];
        if (bad_mat(c.material)) return bail(fail(PRT_ERR_INVALID_ARGUMENT, "cone material out of range"));
        D3 b = d3(c.base_center), t = d3(c.top_center);
        double h = len(t - b);
        if (!(c.base_radius > 0)) return bail(fail(PRT_ERR_INVALID_ARGUMENT, "cone base radius must be positive"));
        if (!(c.top_radius >= 0)) return bail(fail(PRT_ERR_INVALID_ARGUMENT, "cone top radius must be non-negative"));
        if (!(c.base_radius > c.top_radius))
            return bail(fail(PRT_ERR_INVALID_ARGUMENT, "cone base radius must exceed its top radius"));
        if (!(h > 0)) return bail(fail(PRT_ERR_INVALID_ARGUMENT, "cone height must be positive"));
        axials.push_back({b, t, (t - b) * (1.0 / h), c.base_radius, c.top_radius, h, c.capped != 0, true, c.material});
    }
    // lights that bring their own material, appended after the caller's in light order: the disc
    // spot light's emitter (its disc joins the shapes) and the emissive material of a sphere light
    // whose sphere is not a shape (NewSphereLight without AddSphereLight)
    std::vector<int> spot_index(std::max(d->num_lights, 1), -1);  // light -> internal material nm + k
    int nspot = 0;
    for (int i = 0; i < d->num_lights; i++) {
        const prt_light& L = d->lights[i];
        if (L.type == PRT_LIGHT_SPHERE && L.shape == -1) {
            if (!(L.radius != 0)) return bail(fail(PRT_ERR_INVALID_ARGUMENT, "detached sphere light needs a radius"));
            spot_index[i] = nspot++;
            continue;
        }
        if (L.type != PRT_LIGHT_DISC_SPOT && L.type != PRT_LIGHT_POINT_SPOT) continue;
        if (!(len(d3(L.to) - d3(L.from)) > 0))
            return bail(fail(PRT_ERR_INVALID_ARGUMENT, "spot light needs from != to"));
        if (L.type != PRT_LIGHT_DISC_SPOT) continue;
        if (!(L.radius > 0)) return bail(fail(PRT_ERR_INVALID_ARGUMENT, "disc spot light radius must be positive"));
        spot_index[i] = nspot++;
        discs.push_back(make_disc(d3(L.from), d3(L.to) - d3(L.from), L.radius, nm + spot_index[i]));
    }
    int64_t ntri = 0;
    for (int i = 0; i < d->num_meshes; i++) {
        const prt_mesh& m = d->meshes[i];
        if (m.num_faces < 0 || m.num_vertices < 0 || (m.num_faces > 0 && (!m.faces || !m.vertices)))
            return bail(fail(PRT_ERR_INVALID_ARGUMENT, "bad mesh arrays"));
        if (!m.face_materials && m.num_faces > 0 && bad_mat(m.material))
            return bail(fail(PRT_ERR_INVALID_ARGUMENT, "mesh material out of range"));
        ntri += m.num_faces;
    }
    const int64_t nsph = d->num_spheres;
    const int64_t nq = (int64_t)quads.size();
    const int64_t ndisc = (int64_t)discs.size();
    const int64_t nuser_disc = d->num_discs;
    const int64_t naxial = (int64_t)axials.size();
    const int64_t nprim = ntri + nsph + nq + ndisc + naxial;
    if (nprim >= (1ll << 28)) return bail(fail(PRT_ERR_UNSUPPORTED, "more than 2^28 primitives"));

    std::vector<PrimBounds> bounds(nprim);
    std::vector<float4> rec(3 * nprim);
    PrimBounds world;
    for (int a = 0; a < 3; a++) {
        world.lo[a] = INFINITY;
        world.hi[a] = -INFINITY;
    }
    bool world_any = false;
    int64_t gid = 0;
    std::vector<float4> trin(1, make_float4(0.f, 0.f, 0.f, 0.f));  // custom triangle normals, 1-based
    std::vector<float2> uvg;  // vertex UVs of mesh triangles by global id (x = NaN: none), if any mesh has them
    std::vector<double> rotated;
    for (int i = 0; i < d->num_meshes; i++) {
        const prt_mesh& m = d->meshes[i];
        const double* verts = m.vertices;
        if (m.rotation && m.num_vertices > 0) {  // NewTriangleMesh :57-70 with rotateVertex :170-199
            rotated.assign(m.vertices, m.vertices + 3 * m.num_vertices);
            const double* rt = m.rotation;
            for (int64_t k = 0; k < m.num_vertices; k++) {
                double* v = &rotated[3 * k];
                if (m.center)
                    for (int a = 0; a < 3; a++) v[a] -= m.center[a];
                if (rt[0] != 0) {
                    double c = std::cos(rt[0]), sn = std::sin(rt[0]);
                    double y = v[1] * c - v[2] * sn, z = v[1] * sn + v[2] * c;
                    v[1] = y;
                    v[2] = z;
                }
                if (rt[1] != 0) {
                    double c = std::cos(rt[1]), sn = std::sin(rt[1]);
                    double x = v[0] * c + v[2] * sn, z = -v[0] * sn + v[2] * c;
                    v[0] = x;
                    v[2] = z;
                }
                if (rt[2] != 0) {
                    double c = std::cos(rt[2]), sn = std::sin(rt[2]);
                    double x = v[0] * c - v[1] * sn, y = v[0] * sn + v[1] * c;
                    v[0] = x;
                    v[1] = y;
                }
                if (m.center)
                    for (int a = 0; a < 3; a++) v[a] += m.center[a];
            }
            verts = rotated.data();
        }
        PrimBounds mb;
        for (int a = 0; a < 3; a++) {
            mb.lo[a] = INFINITY;
            mb.hi[a] = -INFINITY;
        }
        for (int64_t f = 0; f < m.num_faces; f++, gid++) {
            const int32_t* fi = m.faces + 3 * f;
            if (fi[0] < 0 || fi[1] < 0 || fi[2] < 0 || fi[0] >= m.num_vertices || fi[1] >= m.num_vertices ||
                fi[2] >= m.num_vertices)
                return bail(fail(PRT_ERR_INVALID_ARGUMENT, "face index out of bounds"));
            int mat = m.face_materials ? m.face_materials[f] : m.material;
            if (bad_mat(mat)) return bail(fail(PRT_ERR_INVALID_ARGUMENT, "face material out of range"));
            const double* p0 = verts + 3 * fi[0];
            const double* p1 = verts + 3 * fi[1];
            const double* p2 = verts + 3 * fi[2];
            PrimBounds& b = bounds[gid];
            for (int a = 0; a < 3; a++) {
                b.lo[a] = std::min(p0[a], std::min(p1[a], p2[a]));
                b.hi[a] = std::max(p0[a], std::max(p1[a], p2[a]));
            }
            grow(mb, b);
            uint32_t side = 0;
            if (m.normals) {  // NewTriangleWithNormal (triangle.go:35-49): normalized custom normal
                const double* cn = m.normals + 3 * f;
                double l = std::sqrt(cn[0] * cn[0] + cn[1] * cn[1] + cn[2] * cn[2]);
                double inv = l > 0 ? 1.0 / l : 0.0;
                side = (uint32_t)trin.size();
                if (side > kSideMask) return bail(fail(PRT_ERR_UNSUPPORTED, "too many custom normals"));
                trin.push_back(make_float4((float)(cn[0] * inv), (float)(cn[1] * inv), (float)(cn[2] * inv), 0.f));
            }
            if (m.uvs) {  // TriangleMeshOptions.VertexUVs (triangle_mesh.go:97-112)
                if (uvg.empty()) uvg.assign(3 * (size_t)nprim, make_float2(NAN, NAN));
                for (int k = 0; k < 3; k++)
                    uvg[3 * gid + k] = make_float2((float)m.uvs[2 * fi[k]], (float)m.uvs[2 * fi[k] + 1]);
            }
            rec[3 * gid + 0] = f4(p0[0], p0[1], p0[2], bits(prim_tag(kPrimTriangle, side)));
            rec[3 * gid + 1] = f4(p1[0], p1[1], p1[2], bits((uint32_t)mat));
            rec[3 * gid + 2] = f4(p2[0], p2[1], p2[2], bits((uint32_t)gid));
        }
        if (m.num_faces > 0) {
            grow(world, mb);
            world_any = true;
        }
    }
    for (int64_t i = 0; i < nsph; i++, gid++) {
        const prt_sphere& sp = d->spheres[i];
        if (bad_mat(sp.material)) return bail(fail(PRT_ERR_INVALID_ARGUMENT, "sphere material out of range"));
        double r = sp.radius, ar = std::fabs(r);
        PrimBounds& b = bounds[gid];
        PrimBounds ref;  // reference box c-r .. c+r (inverted for r < 0, sphere.go:83-90)
        for (int a = 0; a < 3; a++) {
            b.lo[a] = sp.center[a] - ar;
            b.hi[a] = sp.center[a] + ar;
            ref.lo[a] = sp.center[a] - r;
            ref.hi[a] = sp.center[a] + r;
        }
        grow(world, ref);
        world_any = true;
        rec[3 * gid + 0] = f4(sp.center[0], sp.center[1], sp.center[2], bits(prim_tag(kPrimSphere, (uint32_t)i)));
        rec[3 * gid + 1] = make_float4((float)r, 0.f, 0.f, bits((uint32_t)sp.material));
        rec[3 * gid + 2] = make_float4(0.f, 0.f, 0.f, bits((uint32_t)gid));
    }
    std::vector<QuadExt> qext(std::max<int64_t>(nq, 1));
    for (int64_t i = 0; i < nq; i++, gid++) {
        const QuadF64& q = quads[i];
        bounds[gid] = quad_bounds(q);
        if (i < d->num_quads) {
            grow(world, bounds[gid]);
            world_any = true;
        }
        D3 n = normalize(cross(q.u, q.v));
        double D = dot(n, q.corner);
        D3 w = n * (1.0 / dot(n, cross(q.u, q.v)));
        // alpha = W . (hv x v) = hv . (v x W), beta = W . (u x hv) = hv . (W x u) (quad.go:133-180)
        D3 A = cross(q.v, w), B = cross(w, q.u);
        qext[i] = QuadExt{(float)n.x, (float)n.y, (float)n.z, (float)D,   (float)A.x,  (float)A.y,
                          (float)A.z, 0.f,        (float)B.x, (float)B.y, (float)B.z, 0.f, {0.f, 0.f, 0.f, 0.f}};
        rec[3 * gid + 0] = f4(q.corner.x, q.corner.y, q.corner.z, bits(prim_tag(kPrimQuad, (uint32_t)i)));
        rec[3 * gid + 1] = f4(q.u.x, q.u.y, q.u.z, bits((uint32_t)q.mat));
        rec[3 * gid + 2] = f4(q.v.x, q.v.y, q.v.z, bits((uint32_t)gid));
    }
    for (auto& wb : world_boxes) {
        grow(world, wb);
        world_any = true;
    }
    // user discs, cylinders, cones, then the discs of disc spot lights
    auto put_disc = [&](const DiscF64& dc) {
        PrimBounds ref = disc_bounds(dc);
        grow(world, ref);
        world_any = true;
        // the BVH box is padded: the corner box touches the disc rim, which a float32 plane hit
        // can overshoot by a few ulps
        PrimBounds& b = bounds[gid];
        for (int a = 0; a < 3; a++) {
            double pad = 1e-5 * dc.r + 1e-6 * std::fabs(comp(dc.c, a));
            b.lo[a] = ref.lo[a] - pad;
            b.hi[a] = ref.hi[a] + pad;
        }
        rec[3 * gid + 0] = f4(dc.c.x, dc.c.y, dc.c.z, bits(prim_tag(kPrimDisc, 0)));
        rec[3 * gid + 1] = f4(dc.n.x, dc.n.y, dc.n.z, bits((uint32_t)dc.mat));
        rec[3 * gid + 2] = make_float4((float)dc.r, (float)(dc.r * dc.r), 0.f, bits((uint32_t)gid));
        gid++;
    };
    for (int64_t i = 0; i < nuser_disc; i++) put_disc(discs[i]);
    for (const AxialF64& c : axials) {
        // world box: Cylinder/Cone.BoundingBox, the base-top segment box widened by the (base)
        // radius on every axis the axis is not parallel to (cylinder.go:41-88, cone.go:83-137)
        PrimBounds ref, &b = bounds[gid];
        for (int a = 0; a < 3; a++) {
            double lo = std::min(comp(c.base, a), comp(c.top, a)), hi = std::max(comp(c.base, a), comp(c.top, a));
            double ext = std::fabs(comp(c.axis, a)) > 0.9999 ? 0.0 : c.rb;
            ref.lo[a] = lo - ext;
            ref.hi[a] = hi + ext;
            // traversal box: the exact box of the two end discs (a disc of radius r about unit
            // axis v spans r * sqrt(1 - v_a^2) along axis a), padded
            double s = std::sqrt(std::max(0.0, 1.0 - comp(c.axis, a) * comp(c.axis, a)));
            double eb = c.rb * s, et = c.rt * s;
            double pad = 1e-5 * c.rb + 1e-6 * std::max(std::fabs(comp(c.base, a)), std::fabs(comp(c.top, a)));
            b.lo[a] = std::min(comp(c.base, a) - eb, comp(c.top, a) - et) - pad;
            b.hi[a] = std::max(comp(c.base, a) + eb, comp(c.top, a) + et) + pad;
        }
        grow(world, ref);
        world_any = true;
        const uint32_t kind = c.cone ? kPrimCone : kPrimCylinder;
        rec[3 * gid + 0] = f4(c.base.x, c.base.y, c.base.z, bits(prim_tag(kind, c.capped ? 1u : 0u)));
        rec[3 * gid + 1] = f4(c.axis.x, c.axis.y, c.axis.z, bits((uint32_t)c.mat));
        rec[3 * gid + 2] = c.cone ? make_float4((float)c.rb, (float)c.rt, (float)c.height, bits((uint32_t)gid))
                                  : make_float4((float)c.rb, (float)c.height, 0.f, bits((uint32_t)gid));
        gid++;
    }
    for (int64_t i = nuser_disc; i < ndisc; i++) put_disc(discs[i]);
    // world centre / radius of the reference BVH root box (bvh.go:38-48)
    if (world_any) {
        for (int a = 0; a < 3; a++) s->info.world_center[a] = 0.5 * (world.lo[a] + world.hi[a]);
        double dx = world.hi[0] - s->info.world_center[0], dy = world.hi[1] - s->info.world_center[1],
               dz = world.hi[2] - s->info.world_center[2];
        s->info.world_radius = std::sqrt(dx * dx + dy * dy + dz * dz);
    }

    // ---- BVH ----
    BVHBuildResult bvh = build_bvh4(bounds, 4);
    std::vector<float4> ordered(3 * std::max<int64_t>(nprim, 1));
    s->orig_id.resize(nprim);
    for (int64_t k = 0; k < nprim; k++) {
        uint32_t i = bvh.order[k];
        ordered[3 * k + 0] = rec[3 * i + 0];
        ordered[3 * k + 1] = rec[3 * i + 1];
        ordered[3 * k + 2] = rec[3 * i + 2];
        s->orig_id[k] = i;
    }
    s->info.num_primitives = nprim;
    s->info.reference_primitives = ntri + nsph + d->num_quads + d->num_boxes + ndisc + naxial;
    s->info.num_triangles = ntri;
    s->info.bvh_nodes = (int64_t)bvh.nodes.size();
    s->info.bvh_depth = bvh.depth;
    s->info.max_stack = bvh.max_stack;
    s->max_stack = bvh.max_stack;

    // ---- device upload ----
    hipDeviceProp_t prop;
    PRT_HIP(hipGetDeviceProperties(&prop, device));
    s->num_cus = prop.multiProcessorCount;
    void* p;
    // bufs[0]: work counters (kept first: the intersection launcher uses them)
    if ((rc = scene_alloc(s, kWorkBytes, nullptr, &p)) != PRT_OK) return bail(rc);
    if ((rc = scene_alloc(s, bvh.nodes.size() * sizeof(Node4), bvh.nodes.data(), &p)) != PRT_OK) return bail(rc);
    s->dev.nodes = (const Node4*)p;
    if ((rc = scene_alloc(s, ordered.size() * sizeof(float4), ordered.data(), &p)) != PRT_OK) return bail(rc);
    s->dev.prims = (const float4*)p;
    if ((rc = scene_alloc(s, qext.size() * sizeof(QuadExt), qext.data(), &p)) != PRT_OK) return bail(rc);
    s->dev.quads = (const QuadExt*)p;
    if ((rc = scene_alloc(s, trin.size() * sizeof(float4), trin.data(), &p)) != PRT_OK) return bail(rc);
    s->dev.tri_normals = (const float4*)p;
    if (!uvg.empty()) {  // vertex UVs in leaf order
        std::vector<float2> uvl(3 * (size_t)nprim);
        for (int64_t k = 0; k < nprim; k++)
            for (int j = 0; j < 3; j++) uvl[3 * k + j] = uvg[3 * (size_t)bvh.order[k] + j];
        if ((rc = scene_alloc(s, uvl.size() * sizeof(float2), uvl.data(), &p)) != PRT_OK) return bail(rc);
        s->dev.tri_uv = (const float2*)p;
    }
    s->dev.textured = textured ? 1 : 0;
    if (d->num_textures > 0) {
        std::vector<DevTexture> tex(d->num_textures);
        for (int i = 0; i < d->num_textures; i++) {
            const prt_texture& t = d->textures[i];
            if ((rc = scene_alloc(s, (size_t)t.width * t.height * 3 * sizeof(float), t.rgb, &p)) != PRT_OK)
                return bail(rc);
            tex[i] = DevTexture{t.width, t.height, (const float*)p};
        }
        if ((rc = scene_alloc(s, tex.size() * sizeof(DevTexture), tex.data(), &p)) != PRT_OK) return bail(rc);
        s->dev.textures = (const DevTexture*)p;
    }
    s->dev.mixed = (nsph + nq + ndisc + naxial) > 0 ? 1 : 0;
    s->dev.max_stack = bvh.max_stack;
    s->dev.num_nodes = (int32_t)bvh.nodes.size();
    s->dev.num_prims = (int32_t)nprim;

    // materials (float32 copies)
    std::vector<DevMaterial> mats(std::max(nm + nspot, 1));
    for (int i = 0; i < nm; i++) {
        const prt_material& m = d->materials[i];
        DevMaterial& o = mats[i];
        o.type = m.type;
        o.outer = m.outer;
        o.inner = m.inner;
        for (int a = 0; a < 3; a++) {
            o.albedo[a] = (float)m.albedo[a];
            o.emission[a] = (float)m.emission[a];
        }
        o.fuzz = (float)std::min(1.0, std::max(0.0, m.fuzz));  // NewMetal clamps (metal.go:19-29)
        o.ior = (float)m.ior;
        o.tex = (m.type == PRT_MAT_LAMBERTIAN || m.type == PRT_MAT_METAL) ? m.texture : 0;
        o.ratio = (float)std::min(1.0, std::max(0.0, m.ratio));  // NewMix clamps (mix.go:17-27)
    }
    for (int i = 0; i < d->num_lights; i++) {  // discSpotLightMaterial (disc_spot_light.go:12-46)
        if (spot_index[i] < 0) continue;
        const prt_light& L = d->lights[i];
        DevMaterial& o = mats[nm + spot_index[i]];
        std::memset(&o, 0, sizeof(o));
        if (L.type == PRT_LIGHT_SPHERE) {  // NewEmissive(emission)
            o.type = PRT_MAT_EMISSIVE;
            for (int a = 0; a < 3; a++) o.emission[a] = (float)L.emission[a];
            continue;
        }
        o.type = kMatSpotEmitter;
        D3 dir = normalize(d3(L.to) - d3(L.from));
        for (int a = 0; a < 3; a++) {
            o.emission[a] = (float)L.emission[a];
            o.spot_dir[a] = (float)comp(dir, a);
        }
    }
    if ((rc = scene_alloc(s, mats.size() * sizeof(DevMaterial), mats.data(), &p)) != PRT_OK) return bail(rc);
    s->dev.mats = (const DevMaterial*)p;

    // lights
    const int nl = d->num_lights;
    std::vector<DevLight> lights(std::max(nl, 1));
    std::vector<float> weights(std::max(nl, 1), 0.f);
    int ninf = 0, nenv = 0;
    for (int i = 0; i < nl; i++) {
        const prt_light& L = d->lights[i];
        DevLight& o = lights[i];
        std::memset(&o, 0, sizeof(o));
        o.type = L.type;
        for (int a = 0; a < 3; a++) {
            o.emission[a] = (float)L.emission[a];
            o.top[a] = (float)L.top[a];
            o.bottom[a] = (float)L.bottom[a];
        }
        if (L.type == PRT_LIGHT_QUAD) {
            if (L.shape < 0 || L.shape >= d->num_quads)
                return bail(fail(PRT_ERR_INVALID_ARGUMENT, "quad light shape out of range"));
            const prt_quad& q = d->quads[L.shape];
            D3 u = d3(q.u), v = d3(q.v), c = d3(q.corner);
            D3 n = normalize(cross(u, v));
            D3 w = n * (1.0 / dot(n, cross(u, v)));
            for (int a = 0; a < 3; a++) {
                o.corner[a] = (float)comp(c, a);
                o.u[a] = (float)comp(u, a);
                o.v[a] = (float)comp(v, a);
                o.n[a] = (float)comp(n, a);
                o.W[a] = (float)comp(w, a);
            }
            o.D = (float)dot(n, c);
            o.area = (float)len(cross(u, v));  // NewQuadLight quad_light.go:19-29
            o.mat = q.material;
        } else if (L.type == PRT_LIGHT_SPHERE && L.shape == -1) {  // sphere not among the shapes
            for (int a = 0; a < 3; a++) o.center[a] = (float)L.from[a];
            o.radius = (float)L.radius;
            o.mat = nm + spot_index[i];
        } else if (L.type == PRT_LIGHT_SPHERE) {
            if (L.shape < 0 || L.shape >= d->num_spheres)
                return bail(fail(PRT_ERR_INVALID_ARGUMENT, "sphere light shape out of range"));
            const prt_sphere& sp = d->spheres[L.shape];
            for (int a = 0; a < 3; a++) o.center[a] = (float)sp.center[a];
            o.radius = (float)sp.radius;
            o.mat = sp.material;
        } else if (L.type == PRT_LIGHT_DISC || L.type == PRT_LIGHT_DISC_SPOT || L.type == PRT_LIGHT_POINT_SPOT) {
            if (L.type == PRT_LIGHT_DISC && (L.shape < 0 || L.shape >= d->num_discs))
                return bail(fail(PRT_ERR_INVALID_ARGUMENT, "disc light shape out of range"));
            if (L.type != PRT_LIGHT_DISC) {  // NewDiscSpotLight / NewPointSpotLight
                const double deg = 3.14159265358979323846 / 180.0;
                D3 dir = normalize(d3(L.to) - d3(L.from));
                for (int a = 0; a < 3; a++) {
                    o.from[a] = (float)L.from[a];
                    o.from_lo[a] = (float)(L.from[a] - (double)o.from[a]);
                    o.spot_dir[a] = (float)comp(dir, a);
                }
                o.cos_total = (float)std::cos(L.cone_angle * deg);
                o.cos_falloff = (float)std::cos((L.cone_angle - L.cone_delta) * deg);
            }
            if (L.type != PRT_LIGHT_POINT_SPOT) {
                const DiscF64& dc =
                    L.type == PRT_LIGHT_DISC ? discs[L.shape] : discs[d->num_discs + spot_index[i]];
                for (int a = 0; a < 3; a++) {
                    o.center[a] = (float)comp(dc.c, a);
                    o.n[a] = (float)comp(dc.n, a);
                    o.right[a] = (float)comp(dc.right, a);
                    o.up[a] = (float)comp(dc.up, a);
                }
                o.radius = (float)dc.r;
                o.mat = dc.mat;
            }
        } else if (L.type == PRT_LIGHT_UNIFORM_INFINITE || L.type == PRT_LIGHT_GRADIENT_INFINITE) {
            ninf++;
        } else if (L.type == PRT_LIGHT_ENVMAP) {
            ninf++;
            nenv++;
            if (nenv > 1) return bail(fail(PRT_ERR_UNSUPPORTED, "at most one environment map light"));
            if (!L.env_rgb || L.env_width <= 0 || L.env_height <= 0)
                return bail(fail(PRT_ERR_INVALID_ARGUMENT, "environment map needs pixels"));
            // piecewise-constant 2D distribution over luminance * sin(theta), built in float64
            const int W = L.env_width, H = L.env_height;
            const double pi = 3.14159265358979323846;
            std::vector<float> func((size_t)W * H), col((size_t)H * (W + 1)), row(H + 1);
            std::vector<double> rs(H), cc(W + 1);
            for (int y = 0; y < H; y++) {
                double st = std::sin((y + 0.5) / H * pi);
                cc[0] = 0;
                for (int x = 0; x < W; x++) {
                    const float* px = L.env_rgb + ((size_t)y * W + x) * 3;
                    double l = 0.2126 * (double)px[0] + 0.7152 * (double)px[1] + 0.0722 * (double)px[2];
                    double f = std::max(l, 0.0) * st;
                    func[(size_t)y * W + x] = (float)f;
                    cc[x + 1] = cc[x] + f;
                }
                rs[y] = cc[W];
                for (int x = 0; x <= W; x++)
                    col[(size_t)y * (W + 1) + x] = (float)(rs[y] > 0 ? cc[x] / rs[y] : (double)x / W);
                col[(size_t)y * (W + 1) + W] = 1.0f;
            }
            double acc = 0;
            std::vector<double> rc64(H + 1);
            rc64[0] = 0;
            for (int y = 0; y < H; y++) rc64[y + 1] = rc64[y] + rs[y];
            acc = rc64[H];
            for (int y = 0; y <= H; y++) row[y] = (float)(acc > 0 ? rc64[y] / acc : (double)y / H);
            row[H] = 1.0f;
            s->dev.env.width = W;
            s->dev.env.height = H;
            s->dev.env.integral = (float)(acc / ((double)W * H));
            if ((rc = scene_alloc(s, (size_t)W * H * 3 * 4, L.env_rgb, &p)) != PRT_OK) return bail(rc);
            s->dev.env.rgb = (const float*)p;
            if ((rc = scene_alloc(s, func.size() * 4, func.data(), &p)) != PRT_OK) return bail(rc);
            s->dev.env.func = (const float*)p;
            if ((rc = scene_alloc(s, row.size() * 4, row.data(), &p)) != PRT_OK) return bail(rc);
            s->dev.env.row_cdf = (const float*)p;
            if ((rc = scene_alloc(s, col.size() * 4, col.data(), &p)) != PRT_OK) return bail(rc);
            s->dev.env.col_cdf = (const float*)p;
        } else {
            return bail(fail(PRT_ERR_UNSUPPORTED, "unknown light type"));
        }
        if ((L.type == PRT_LIGHT_QUAD || (L.type == PRT_LIGHT_SPHERE && L.shape != -1)) && bad_mat(o.mat))
            return bail(fail(PRT_ERR_INVALID_ARGUMENT, "light material out of range"));
    }
    // light selection probabilities (NewUniformLightSampler / NewWeightedLightSampler)
    if (nl > 0) {
        double tot = 0;
        if (d->light_weights)
            for (int i = 0; i < nl; i++) {
                if (d->light_weights[i] < 0) return bail(fail(PRT_ERR_INVALID_ARGUMENT, "negative light weight"));
                tot += d->light_weights[i];
            }
        for (int i = 0; i < nl; i++)
            weights[i] = (float)((d->light_weights && tot > 0) ? d->light_weights[i] / tot : 1.0 / nl);
    }
    if ((rc = scene_alloc(s, lights.size() * sizeof(DevLight), lights.data(), &p)) != PRT_OK) return bail(rc);
    s->dev.lights = (const DevLight*)p;
    if ((rc = scene_alloc(s, weights.size() * 4, weights.data(), &p)) != PRT_OK) return bail(rc);
    s->dev.light_weights = (const float*)p;
    s->dev.num_lights = nl;
    s->dev.num_infinite = ninf;
    for (int a = 0; a < 3; a++) s->dev.world_center[a] = (float)s->info.world_center[a];
    s->dev.world_radius = (float)s->info.world_radius;
    int H = 0;
    s->dev.cam = make_camera(d->camera, H, 0);
    s->cam_desc = d->camera;
    s->image_width = d->camera.width;
    s->image_height = H;
    s->aspect = d->camera.aspect_ratio;

    // per-variant grid (occupancy-limited, at most 12 work-groups per CU) and one overflow-stack
    // slot per resident thread of the largest grid
    int max_blocks = 1;
    for (int v = 0; v < kNumVariants; v++) {
        int per_cu = 0;
        const void* fn = s->dev.mixed ? variant_fn<false, true>(v) : variant_fn<false, false>(v);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kVariantBlock[v], 0) != hipSuccess ||
            per_cu <= 0)
            per_cu = 2;
        per_cu = std::min(per_cu, 12 * 256 / kVariantBlock[v]);
        s->variant_blocks[v] = s->num_cus * per_cu;
        max_blocks = std::max(max_blocks, s->variant_blocks[v] * kVariantBlock[v] / kTraceBlock);
    }
    s->trace_variant[0] = s->trace_variant[1] = -1;
    s->forced_variant = -1;
    if (const char* fv = std::getenv("PRT_TRACE_VARIANT")) {
        int v = std::atoi(fv);
        if (v >= 0 && v < kNumVariants) s->forced_variant = v;
    }
    s->info.trace_variant = s->forced_variant;
    s->info.pt_trace_leaf_gate = -1;
    s->info.bdpt_trace_leaf_gate = -1;
    if (const char* fp = std::getenv("PRT_PT_TRACE_VARIANT")) {  // pins the path tracer's trace occupancy variant
        int v = std::atoi(fp);
        if (v >= 0 && v < 4) s->pt_trace_variant = v;
    }
    if (const char* fg = std::getenv("PRT_TRACE_LEAF_GATE")) {  // pins the render kernels' leaf gate (0-8, 9 flat)
        int g = std::atoi(fg);
        if ((g >= 0 && g <= 8) || (g == kGateFlat && nprim <= kFlatMaxPrims)) {
            s->pt_trace_gate = s->bdpt_trace_gate = g;
            s->info.pt_trace_leaf_gate = s->info.bdpt_trace_leaf_gate = g;
        }
    }
    uint32_t slots = (uint32_t)(max_blocks * kTraceBlock);
    size_t depth = (size_t)std::max(1, s->max_stack - kMinLds + 1);
    if ((rc = scene_alloc(s, depth * slots * sizeof(uint2), nullptr, &p)) != PRT_OK) return bail(rc);
    s->gstack = (uint2*)p;
    s->gstack_slots = slots;
    // park queue of the intersection kernels: count (256 B), then kParkCapacity ray indices
    if ((rc = scene_alloc(s, 256 + sizeof(uint32_t) * (size_t)kParkCapacity, nullptr, &p)) != PRT_OK) return bail(rc);
    s->park = (uint32_t*)p;
    s->park_steps = -1;  // chosen per batch (park_budget)
    if (const char* fp = std::getenv("PRT_TRACE_PARK_STEPS")) {  // pins the batch kernels' park budget (0: none)
        long b = std::atol(fp);
        if (b >= 0 && b < (1l << 30)) s->park_steps = (int32_t)b;
    }
    *out = s;
    return PRT_OK;
}

int prt_scene_info_get(const prt_scene* scene, prt_scene_info* info) {
    if (!scene || !info) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL argument");
    *info = scene->info;
    return PRT_OK;
}

int prt_scene_destroy(prt_scene* scene) {
    if (!scene) return PRT_OK;
    if (int n = scene->users.load()) {
        return fail(PRT_ERR_INVALID_ARGUMENT,
                    "scene is still used by " + std::to_string(n) + " film(s) or serving request(s)");
    }
    (void)hipSetDevice(scene->device);
    for (prt::TuneJob* j : {&scene->pt_tune, &scene->bdpt_tune})
        for (hipEvent_t& e : j->ev)
            if (e) {
                (void)hipEventSynchronize(e);
                (void)hipEventDestroy(e);
            }
    if (scene->pt_profile.ev) {
        (void)hipEventSynchronize(scene->pt_profile.ev);
        (void)hipEventDestroy(scene->pt_profile.ev);
    }
    if (scene->pt_profile.host) (void)hipHostFree(scene->pt_profile.host);
    for (auto& b : scene->bufs) (void)hipFree(b.p);
    delete scene;
    return PRT_OK;
}

int prt_intersect_closest(const prt_scene* scene, const prt_ray* rays, int64_t n, prt_hit* hits, uint32_t flags,
                          void* stream) {
    return launch_intersect(scene, 0, rays, n, hits, nullptr, nullptr, flags, stream);
}

int prt_intersect_any(const prt_scene* scene, const prt_ray* rays, int64_t n, uint8_t* occluded, uint32_t flags,
                      void* stream) {
    return launch_intersect(scene, 1, rays, n, occluded, nullptr, nullptr, flags, stream);
}

int prt_intersect_closest_static(const prt_scene* scene, const prt_ray* rays, int64_t n, prt_hit* hits,
                                 void* stream) {
    if (!scene) return fail(PRT_ERR_INVALID_ARGUMENT, "scene is NULL");
    if (n < 0) return fail(PRT_ERR_INVALID_ARGUMENT, "negative ray count");
    if (n == 0) return PRT_OK;
    if (!rays || !hits) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL buffer");
    PRT_HIP(hipSetDevice(scene->device));
    void* base = nullptr;
    int rc = static_scratch(scene->device, &base);
    if (rc != PRT_OK) return rc;
    // the variant the autotuner picked for this scene (no tuning here: it times launches), else the default
    const int v = scene->forced_variant >= 0 ? scene->forced_variant
                                             : (scene->trace_variant[0] >= 0 ? scene->trace_variant[0] : kDefaultVariant);
    const int block = kVariantBlock[v];
    // stack entries per thread slot: the variant's spill, or k_intersect_long's if that spills sooner
    const size_t entries = (size_t)std::max(1, scene->max_stack - std::min(kVariantLds[v], kLongLds) + 1);
    const size_t per_block = entries * (size_t)block * sizeof(uint2);
    const int64_t fit = (int64_t)((kStaticScratchBytes - kStaticStackOffset) / per_block);
    if (fit < 1)
        return fail(PRT_ERR_UNSUPPORTED, "the scene's traversal stack bound (" + std::to_string(scene->max_stack) +
                                             " entries) needs more than the static scratch of " +
                                             std::to_string(kStaticScratchBytes) + " bytes");
    // the full occupancy grid where the stack fits, else as many work-groups as it holds
    const int blocks = (int)std::min<int64_t>(scene->variant_blocks[v], fit);
    unsigned long long* work = (unsigned long long*)base;
    uint32_t* pk = (uint32_t*)((char*)base + kStaticParkOffset);
    const ParkQueue park{pk, pk + 64, kStaticParkCapacity, park_budget(scene, n, (int64_t)blocks * block)};
    uint2* gs = (uint2*)((char*)base + kStaticStackOffset);
    const uint32_t slots = (uint32_t)blocks * (uint32_t)block;
    hipStream_t st = (hipStream_t)stream;
    const int lb = long_blocks(scene, slots);
    (void)hipGetLastError();  // an earlier call's error on this thread is not this launch's
    hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(kWorkParts), 0, st, work, park.count, (unsigned long long)n,
                       (uint32_t)blocks, (uint32_t)(block / 64));
    PRT_HIP(hipGetLastError());
    if (scene->dev.mixed)
        PRT_HIP((launch_variant<false, true>(v, blocks, st, scene->dev, rays, n, hits, nullptr, gs, slots, work, park, lb)));
    else
        PRT_HIP((launch_variant<false, false>(v, blocks, st, scene->dev, rays, n, hits, nullptr, gs, slots, work, park, lb)));
    return PRT_OK;
}

int prt_traversal_counts(const prt_scene* scene, const prt_ray* rays, int64_t n, uint32_t* node_visits,
                         uint32_t* box_tests, uint32_t* prim_tests, uint32_t flags, void* stream) {
    if (!box_tests || !prim_tests) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL buffer");
    return launch_intersect(scene, 2, rays, n, node_visits, box_tests, prim_tests, flags, stream);
}

}  // extern "C"
