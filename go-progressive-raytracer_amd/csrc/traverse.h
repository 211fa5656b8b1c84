// traverse.h — device-side ray traversal of the 4-wide BVH (closest hit / any hit).
//
// Semantics follow BVH.Hit (pkg/geometry/bvh.go:183-239): a hit is any primitive hit with
// t in [tmin, tmax] (inclusive), the closest one is reported. Triangles are two-sided
// (triangle.go:107-175) and use the watertight test of Woop, Benthin and Wald (JCGT 2013)
// in float32 with a float64 fallback on exact edge hits; boxes use the robust slab test
// with the far distance widened by 1 + 2*gamma(4) (Ize 2013, with one more rounding for the
// hardware reciprocal of the direction) so that float32 boxes never reject a ray that the exact box
// would accept.
#pragma once
#include "prt_common.h"

namespace prt {

// Primitive tests round every product and sum on its own (no fused multiply-add): the compiler fuses
// differently in each kernel it inlines a test into, and a hit must come out bit-identical whichever
// traversal kernel finds it. The watertight triangle test also needs it: an edge shared by two
// triangles must give exactly negated edge functions in both.
#define PRT_EXACT_ROUNDING _Pragma("clang fp contract(off)")

struct TRay {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
};

struct THit {
    float t;
    uint32_t prim;  // index into the leaf-ordered primitive array, 0xFFFFFFFF on a miss
    float u, v;
};

struct TCounts {
    uint32_t nodes, boxes, prims;
};

__device__ __forceinline__ float sel3(float x, float y, float z, int k) {
    return k == 0 ? x : (k == 1 ? y : z);
}

__device__ __forceinline__ uint32_t pick4(uint4 c, uint32_t i) {
    return i == 0 ? c.x : (i == 1 ? c.y : (i == 2 ? c.z : c.w));
}

// Per-ray constants of the watertight triangle test.
struct WRay {
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

__device__ __forceinline__ WRay make_wray(float dx, float dy, float dz) {
    WRay w;
    float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dkz = sel3(dx, dy, dz, kz);
    if (dkz < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    w.kx = kx;
    w.ky = ky;
    w.kz = kz;
    w.Sz = 1.0f / dkz;
    w.Sx = sel3(dx, dy, dz, kx) * w.Sz;
    w.Sy = sel3(dx, dy, dz, ky) * w.Sz;
    return w;
}

// Watertight two-sided ray/triangle test. On a hit with t in [tmin, tbest] writes t, u, v
// (u, v = barycentric weights of v1, v2) and returns true.
__device__ __forceinline__ bool tri_test(const TRay& r, const WRay& w, float4 a, float4 b, float4 c,
                                         float tbest, float& t_out, float& u_out, float& v_out) {
    PRT_EXACT_ROUNDING
    float Ax = a.x - r.ox, Ay = a.y - r.oy, Az = a.z - r.oz;
    float Bx = b.x - r.ox, By = b.y - r.oy, Bz = b.z - r.oz;
    float Cx = c.x - r.ox, Cy = c.y - r.oy, Cz = c.z - r.oz;
    float akx = sel3(Ax, Ay, Az, w.kx), aky = sel3(Ax, Ay, Az, w.ky), akz = sel3(Ax, Ay, Az, w.kz);
    float bkx = sel3(Bx, By, Bz, w.kx), bky = sel3(Bx, By, Bz, w.ky), bkz = sel3(Bx, By, Bz, w.kz);
    float ckx = sel3(Cx, Cy, Cz, w.kx), cky = sel3(Cx, Cy, Cz, w.ky), ckz = sel3(Cx, Cy, Cz, w.kz);
    float ax = akx - w.Sx * akz, ay = aky - w.Sy * akz;
    float bx = bkx - w.Sx * bkz, by = bky - w.Sy * bkz;
    float cx = ckx - w.Sx * ckz, cy = cky - w.Sy * ckz;
    float U = cx * by - cy * bx;
    float V = ax * cy - ay * cx;
    float W = bx * ay - by * ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        double U64 = (double)cx * (double)by - (double)cy * (double)bx;
        double V64 = (double)ax * (double)cy - (double)ay * (double)cx;
        double W64 = (double)bx * (double)ay - (double)by * (double)ax;
        U = (float)U64;
        V = (float)V64;
        W = (float)W64;
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float az2 = w.Sz * akz, bz2 = w.Sz * bkz, cz2 = w.Sz * ckz;
    float T = U * az2 + V * bz2 + W * cz2;
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= r.tmin && t <= tbest)) return false;
    t_out = t;
    u_out = V * rcp;
    v_out = W * rcp;
    return true;
}

// Sphere test (sphere.go:27-81): nearer root, else farther root, in [tmin, tbest].
// Ray/sphere roots (sphere.go:27-81 solves the same quadratic in float64). In float32 the textbook
// discriminant hb^2 - A*C cancels catastrophically once |o - c| >> r (rays arriving from an
// infinite light's sampling disk thousands of units away), so the discriminant is formed from the
// perpendicular offset l = f - (f.d / d.d) d (disc = A * (r^2 - l.l)) and the roots by the
// cancellation-free pair q / A, C / q (Haines et al., "Precision Improvements for Ray/Sphere
// Intersection", Ray Tracing Gems, 2019). Returns false if the ray's line misses the sphere;
// t0 <= t1 otherwise. Quotients and the square root use the hardware reciprocal and square root
// (deterministic, 1 ulp each) instead of the correctly rounded sequences.
__device__ __forceinline__ bool sphere_roots(float fx, float fy, float fz, float dx, float dy, float dz, float rad,
                                             float& t0, float& t1) {
    PRT_EXACT_ROUNDING
    float A = dx * dx + dy * dy + dz * dz;
    float hb = fx * dx + fy * dy + fz * dz;
    const float inv_a = __builtin_amdgcn_rcpf(A);
    float k = hb * inv_a;
    float lx = fx - k * dx, ly = fy - k * dy, lz = fz - k * dz;
    float disc = A * (rad * rad - (lx * lx + ly * ly + lz * lz));
    if (disc < 0.0f) return false;
    float C = fx * fx + fy * fy + fz * fz - rad * rad;
    float q = -(hb + copysignf(__builtin_amdgcn_sqrtf(disc), hb));
    float a = q * inv_a, b = q != 0.0f ? C * __builtin_amdgcn_rcpf(q) : a;
    t0 = fminf(a, b);
    t1 = fmaxf(a, b);
    return true;
}

// Sphere test: the nearer root in [tmin, tbest], else the farther one (sphere.go:46-57).
__device__ __forceinline__ bool sphere_test(const TRay& r, float4 a, float4 b, float tbest, float& t_out) {
    float t0, t1;
    if (!sphere_roots(r.ox - a.x, r.oy - a.y, r.oz - a.z, r.dx, r.dy, r.dz, b.x, t0, t1)) return false;
    float root = t0;
    if (root < r.tmin || root > tbest) {
        root = t1;
        if (root < r.tmin || root > tbest) return false;
    }
    t_out = root;
    return true;
}

// Quad test (quad.go:133-180): plane hit with |n.d| >= 1e-8 and t in [tmin, tbest], then the barycentric
// coordinates alpha = W . (hv x v) and beta = W . (u x hv) of hv = p - corner, in [0, 1], as the dot
// products hv . (v x W) and hv . (W x u) with both vectors precomputed per quad. t uses the hardware
// reciprocal (deterministic, within 2 ulp of the quotient).
__device__ __forceinline__ bool quad_test(const TRay& r, float4 a, const QuadExt& q, float tbest, float& t_out,
                                          float& u_out, float& v_out) {
    PRT_EXACT_ROUNDING
    float den = r.dx * q.nx + r.dy * q.ny + r.dz * q.nz;
    if (fabsf(den) < 1e-8f) return false;
    float t = (q.D - (r.ox * q.nx + r.oy * q.ny + r.oz * q.nz)) * __builtin_amdgcn_rcpf(den);
    if (t < r.tmin || t > tbest) return false;
    float px = r.ox + t * r.dx - a.x, py = r.oy + t * r.dy - a.y, pz = r.oz + t * r.dz - a.z;
    float alpha = px * q.ax + py * q.ay + pz * q.az;
    float beta = px * q.bx + py * q.by + pz * q.bz;
    if (alpha < 0.0f || alpha > 1.0f || beta < 0.0f || beta > 1.0f) return false;
    t_out = t;
    u_out = alpha;
    v_out = beta;
    return true;
}

// Disc test (disc.go:46-88): plane hit with |n.d| >= 1e-6, t in [tmin, tbest], |p - c|^2 <= r^2.
__device__ __forceinline__ bool disc_test(const TRay& r, float4 a, float4 b, float4 c, float tbest, float& t_out) {
    PRT_EXACT_ROUNDING
    float den = b.x * r.dx + b.y * r.dy + b.z * r.dz;
    if (fabsf(den) < 1e-6f) return false;
    float t = (b.x * (a.x - r.ox) + b.y * (a.y - r.oy) + b.z * (a.z - r.oz)) / den;
    if (t < r.tmin || t > tbest) return false;
    float px = r.ox + t * r.dx - a.x, py = r.oy + t * r.dy - a.y, pz = r.oz + t * r.dz - a.z;
    if (px * px + py * py + pz * pz > c.y) return false;
    t_out = t;
    return true;
}

// End-cap disc of a cylinder or cone (hitCap, cylinder.go:213-262, cone.go:299-349): plane through
// (cx,cy,cz) with normal +-(vx,vy,vz), |d.n| >= 1e-8, t in [tmin, tmax], |p - c| <= rad.
__device__ __forceinline__ bool cap_test(const TRay& r, float cx, float cy, float cz, float vx, float vy, float vz,
                                         float rad, float tmax, float& t_out) {
    PRT_EXACT_ROUNDING
    float den = vx * r.dx + vy * r.dy + vz * r.dz;
    if (fabsf(den) < 1e-8f) return false;
    float t = (vx * (cx - r.ox) + vy * (cy - r.oy) + vz * (cz - r.oz)) / den;
    if (t < r.tmin || t > tmax) return false;
    float px = r.ox + t * r.dx - cx, py = r.oy + t * r.dy - cy, pz = r.oz + t * r.dz - cz;
    if (px * px + py * py + pz * pz > rad * rad) return false;
    t_out = t;
    return true;
}

// Cylinder (Cylinder.Hit, cylinder.go:91-210): the lateral surface's first root (near, then far)
// with t in [tmin, tmax] and height in [0, H], then the base and top caps when capped, each
// accepted at t <= the closest so far. The quadratic is solved with the perpendicular offset l of
// the ray from the axis (discriminant a (r^2 - |l|^2)) and the cancellation-free root pair, so
// distant origins keep their accuracy. part: 0 body, 1 base cap, 2 top cap.
__device__ __forceinline__ bool cylinder_test(const TRay& r, float4 a, float4 b, float4 c, float tbest, float& t_out,
                                              float& part) {
    PRT_EXACT_ROUNDING
    const float vx = b.x, vy = b.y, vz = b.z, rad = c.x, H = c.y;
    float ex = r.ox - a.x, ey = r.oy - a.y, ez = r.oz - a.z;
    float dv = r.dx * vx + r.dy * vy + r.dz * vz, ev = ex * vx + ey * vy + ez * vz;
    float px = r.dx - dv * vx, py = r.dy - dv * vy, pz = r.dz - dv * vz;  // d perpendicular
    float qx = ex - ev * vx, qy = ey - ev * vy, qz = ez - ev * vz;        // (o - base) perpendicular
    float A = px * px + py * py + pz * pz;
    bool hit = false;
    float best = tbest;
    if (fabsf(A) >= 1e-8f) {
        float hb = qx * px + qy * py + qz * pz;
        float s = hb / A;
        float lx = qx - s * px, ly = qy - s * py, lz = qz - s * pz;
        float disc = A * (rad * rad - (lx * lx + ly * ly + lz * lz));
        if (disc >= 0.0f) {
            float qq = -(hb + copysignf(sqrtf(disc), hb));
            float C = qx * qx + qy * qy + qz * qz - rad * rad;
            float r0 = qq / A, r1 = qq != 0.0f ? C / qq : r0;
            float tn = fminf(r0, r1), tf = fmaxf(r0, r1);
#pragma unroll
            for (int k = 0; k < 2; k++) {
                float t = k == 0 ? tn : tf;
                if (t < r.tmin || t > best) continue;
                float h = ev + t * dv;
                if (h < 0.0f || h > H) continue;
                best = t;
                part = 0.0f;
                hit = true;
                break;
            }
        }
    }
    if (prim_side(__float_as_uint(a.w))) {
        float t;
        if (cap_test(r, a.x, a.y, a.z, -vx, -vy, -vz, rad, best, t)) {
            best = t;
            part = 1.0f;
            hit = true;
        }
        if (cap_test(r, a.x + vx * H, a.y + vy * H, a.z + vz * H, vx, vy, vz, rad, best, t)) {
            best = t;
            part = 2.0f;
            hit = true;
        }
    }
    t_out = best;
    return hit;
}

// Cone or frustum (Cone.Hit, cone.go:140-296). The lateral quadratic about the apex is solved from
// an origin moved along the ray to the point nearest the axis midpoint (coefficients of the
// shape's size, not the origin's distance); like the reference, the root (-b - sqrt D) / 2a is
// tried first, then (-b + sqrt D) / 2a, each valid for t in [tmin, tmax], height within
// [-1e-8, H + 1e-8] and on the apex's lower nappe. Caps: the base, and the top of a frustum.
__device__ __forceinline__ bool cone_test(const TRay& r, float4 a, float4 b, float4 c, float tbest, float& t_out,
                                          float& part) {
    PRT_EXACT_ROUNDING
    const float vx = b.x, vy = b.y, vz = b.z, rb = c.x, rt = c.y, H = c.z;
    const float tn = (rb - rt) / H, k1 = 1.0f + tn * tn;
    const float ha = rb / tn;  // apex height above the base
    float dd = r.dx * r.dx + r.dy * r.dy + r.dz * r.dz;
    // shift: o' = o + ts d, the ray point nearest the axis midpoint
    float mx = a.x + vx * (0.5f * H) - r.ox, my = a.y + vy * (0.5f * H) - r.oy, mz = a.z + vz * (0.5f * H) - r.oz;
    float ts = (mx * r.dx + my * r.dy + mz * r.dz) / dd;
    float ox = r.ox + ts * r.dx, oy = r.oy + ts * r.dy, oz = r.oz + ts * r.dz;
    float cx = ox - (a.x + vx * ha), cy = oy - (a.y + vy * ha), cz = oz - (a.z + vz * ha);  // o' - apex
    float dv = r.dx * vx + r.dy * vy + r.dz * vz, cv = cx * vx + cy * vy + cz * vz;
    float A = dd - k1 * dv * dv;
    bool hit = false;
    float best = tbest;
    if (fabsf(A) >= 1e-8f) {
        float hb = (r.dx * cx + r.dy * cy + r.dz * cz) - k1 * dv * cv;
        float C = (cx * cx + cy * cy + cz * cz) - k1 * cv * cv;
        float disc = hb * hb - A * C;
        if (disc >= 0.0f) {
            float sq = sqrtf(disc);
            float qq = -(hb + copysignf(sq, hb));
            float ra = qq / A, rc = qq != 0.0f ? C / qq : ra;
            // (-b - sqrt D) / 2a is qq / A when b >= 0, else C / qq
            float first = hb >= 0.0f ? ra : rc, second = hb >= 0.0f ? rc : ra;
            float hbase = cv + ha;  // height of o' above the base
#pragma unroll
            for (int k = 0; k < 2; k++) {
                float tp = k == 0 ? first : second;
                float t = ts + tp;
                if (t < r.tmin || t > best) continue;
                float h = hbase + tp * dv;
                if (h < -1e-8f || h > H + 1e-8f) continue;
                if (cv + tp * dv > 1e-8f) continue;  // upper nappe
                best = t;
                part = 0.0f;
                hit = true;
                break;
            }
        }
    }
    if (prim_side(__float_as_uint(a.w))) {
        float t;
        if (cap_test(r, a.x, a.y, a.z, -vx, -vy, -vz, rb, best, t)) {
            best = t;
            part = 1.0f;
            hit = true;
        }
        if (rt > 0.0f && cap_test(r, a.x + vx * H, a.y + vy * H, a.z + vz * H, vx, vy, vz, rt, best, t)) {
            best = t;
            part = 2.0f;
            hit = true;
        }
    }
    t_out = best;
    return hit;
}

// Far-distance widening 1 + 2*gamma(4), gamma(n) = n*u / (1 - n*u) with u = 2^-24: a slab distance
// (b - o) * (1/d) carries three roundings, the subtraction and the product (u each) and the hardware
// reciprocal of the direction (1 ulp, at most 2u relative), so its relative error is below gamma(4).
constexpr float kBoxWiden = 1.0f + 2.0f * (4.0f * 5.9604645e-08f) / (1.0f - 4.0f * 5.9604645e-08f);

// Traversal stack: the first LDS_DEPTH entries per lane live in LDS (lane-major so each lane
// owns one bank column), deeper entries spill to a global array indexed by the thread slot.
template <int LDS_DEPTH, int BLOCK>
struct TStack {
    uint32_t* s_node;  // [LDS_DEPTH][BLOCK]
    float* s_t;        // [LDS_DEPTH][BLOCK]
    uint2* g;          // [(max_stack - LDS_DEPTH)][gstride]
    const float4* ncache = nullptr;  // top BVH nodes copied to LDS (breadth-first prefix)
    uint32_t lane;     // threadIdx.x
    uint32_t gslot;    // global thread slot
    uint32_t gstride;
    __device__ __forceinline__ void push(int sp, uint32_t node, float t) {
        if (sp < LDS_DEPTH) {
            s_node[sp * BLOCK + lane] = node;
            s_t[sp * BLOCK + lane] = t;
        } else {
            g[(size_t)(sp - LDS_DEPTH) * gstride + gslot] = make_uint2(node, __float_as_uint(t));
        }
    }
    __device__ __forceinline__ void pop(int sp, uint32_t& node, float& t) {
        if (sp < LDS_DEPTH) {
            node = s_node[sp * BLOCK + lane];
            t = s_t[sp * BLOCK + lane];
        } else {
            uint2 e = g[(size_t)(sp - LDS_DEPTH) * gstride + gslot];
            node = e.x;
            t = __uint_as_float(e.y);
        }
    }
};

// Per-lane traversal state; the traversal advances one node or one leaf per step so that a
// lane that finishes can take a new ray while the rest of its wave keeps traversing.
struct TState {
    TRay r;
    float idx, idy, idz;
    WRay w;
    float tbest, bu, bv;
    uint32_t best;
    uint32_t cur;
    int sp;
};

__device__ __forceinline__ void tstate_init(TState& s, const TRay& r_in) {
    s.r = r_in;
    if (!(s.r.tmin >= 0.0f)) s.r.tmin = 0.0f;
    // reciprocal direction; zero components become huge so that 0 * inf never occurs
    s.idx = __builtin_amdgcn_rcpf(fabsf(s.r.dx) > 1e-30f ? s.r.dx : copysignf(1e-30f, s.r.dx));
    s.idy = __builtin_amdgcn_rcpf(fabsf(s.r.dy) > 1e-30f ? s.r.dy : copysignf(1e-30f, s.r.dy));
    s.idz = __builtin_amdgcn_rcpf(fabsf(s.r.dz) > 1e-30f ? s.r.dz : copysignf(1e-30f, s.r.dz));
    s.w = make_wray(s.r.dx, s.r.dy, s.r.dz);
    s.tbest = s.r.tmax;
    s.best = 0xFFFFFFFFu;
    s.bu = s.bv = 0.0f;
    s.sp = 0;
    s.cur = 0;  // root node
}

__device__ __forceinline__ void tstate_result(const TState& s, THit& hit) {
    hit.t = s.best == 0xFFFFFFFFu ? __int_as_float(0x7f800000) : s.tbest;
    hit.prim = s.best;
    hit.u = s.bu;
    hit.v = s.bv;
}

// Wave-uniform leaf gate. A wave executes the leaf branch of a step whenever any of its lanes stands at a
// leaf, and a leaf step costs about as much as a node step; when only a few lanes stand at leaves the rest
// of the wave pays for them in every step. With the gate, lanes at a leaf wait (keep their state) until at
// least `eighths`/8 of the wave's running lanes stand at one or no lane has an inner node to test, so the
// leaf branch runs for more lanes at a time. Each ray still visits the same nodes and leaves in the same
// order.
__device__ __forceinline__ bool leaf_gate(bool running, uint32_t cur, int eighths) {
    const bool at_leaf = running && child_is_leaf(cur);
    const uint64_t inner = __ballot(running && !at_leaf);
    const uint64_t leaf = __ballot(at_leaf);
    return inner == 0 || __popcll(leaf) * 8 >= __popcll(leaf | inner) * eighths;
}

// One traversal step. Returns true when the ray is finished (stack empty, or any hit found). A lane at
// a leaf leaves the step without progress when `leaf_ok` is false (see leaf_gate).
template <bool ANY, bool MIXED, bool COUNT, int LDS_DEPTH, int BLOCK, int NCACHE = 0>
__device__ __forceinline__ bool tstate_step(const DevScene& sc, TState& s, TStack<LDS_DEPTH, BLOCK>& st,
                                            TCounts* cnt, bool leaf_ok = true) {
    const TRay& r = s.r;
    if (!child_is_leaf(s.cur)) {
        // ---- inner node: test the 4 child boxes (top levels from LDS when cached) ----
        // A node holds the lower and upper x planes of its four children in its first two 16-B slots,
        // then y, then z. The near plane of an axis is the lower one when the ray's direction is
        // positive along it: the slot is chosen by the load address (16 B further when the sign bit
        // of the inverse direction is set) instead of selecting among loaded values.
        const char* nb = (NCACHE > 0 && s.cur < (uint32_t)NCACHE)
                             ? reinterpret_cast<const char*>(st.ncache + 8 * s.cur)
                             : reinterpret_cast<const char*>(sc.nodes + s.cur);
        const uint32_t sx = (__float_as_uint(s.idx) >> 27) & 16u;
        const uint32_t sy = (__float_as_uint(s.idy) >> 27) & 16u;
        const uint32_t sz = (__float_as_uint(s.idz) >> 27) & 16u;
        const float4 nX = *reinterpret_cast<const float4*>(nb + sx);
        const float4 fX = *reinterpret_cast<const float4*>(nb + (16u - sx));
        const float4 nY = *reinterpret_cast<const float4*>(nb + 32u + sy);
        const float4 fY = *reinterpret_cast<const float4*>(nb + (48u - sy));
        const float4 nZ = *reinterpret_cast<const float4*>(nb + 64u + sz);
        const float4 fZ = *reinterpret_cast<const float4*>(nb + (80u - sz));
        const uint4 ch = *reinterpret_cast<const uint4*>(nb + 96);
        if (COUNT) {
            cnt->nodes++;
            cnt->boxes += 4;
        }
        const float idx = s.idx, idy = s.idy, idz = s.idz;
        float nx0 = (nX.x - r.ox) * idx, fx0 = (fX.x - r.ox) * idx;
        float nx1 = (nX.y - r.ox) * idx, fx1 = (fX.y - r.ox) * idx;
        float nx2 = (nX.z - r.ox) * idx, fx2 = (fX.z - r.ox) * idx;
        float nx3 = (nX.w - r.ox) * idx, fx3 = (fX.w - r.ox) * idx;
        float ny0 = (nY.x - r.oy) * idy, fy0 = (fY.x - r.oy) * idy;
        float ny1 = (nY.y - r.oy) * idy, fy1 = (fY.y - r.oy) * idy;
        float ny2 = (nY.z - r.oy) * idy, fy2 = (fY.z - r.oy) * idy;
        float ny3 = (nY.w - r.oy) * idy, fy3 = (fY.w - r.oy) * idy;
        float nz0 = (nZ.x - r.oz) * idz, fz0 = (fZ.x - r.oz) * idz;
        float nz1 = (nZ.y - r.oz) * idz, fz1 = (fZ.y - r.oz) * idz;
        float nz2 = (nZ.z - r.oz) * idz, fz2 = (fZ.z - r.oz) * idz;
        float nz3 = (nZ.w - r.oz) * idz, fz3 = (fZ.w - r.oz) * idz;
        float tn0 = fmaxf(fmaxf(nx0, ny0), fmaxf(nz0, r.tmin));
        float tn1 = fmaxf(fmaxf(nx1, ny1), fmaxf(nz1, r.tmin));
        float tn2 = fmaxf(fmaxf(nx2, ny2), fmaxf(nz2, r.tmin));
        float tn3 = fmaxf(fmaxf(nx3, ny3), fmaxf(nz3, r.tmin));
        float tf0 = fminf(fminf(fminf(fx0, fy0), fz0) * kBoxWiden, s.tbest);
        float tf1 = fminf(fminf(fminf(fx1, fy1), fz1) * kBoxWiden, s.tbest);
        float tf2 = fminf(fminf(fminf(fx2, fy2), fz2) * kBoxWiden, s.tbest);
        float tf3 = fminf(fminf(fminf(fx3, fy3), fz3) * kBoxWiden, s.tbest);
        // sort keys: non-negative float bits with the slot in the low 2 bits (rounds the
        // entry distance down, which only weakens culling)
        uint32_t k0 = (tn0 <= tf0 && ch.x != kEmptyChild) ? ((__float_as_uint(tn0) & ~3u) | 0u) : 0xFFFFFFFFu;
        uint32_t k1 = (tn1 <= tf1 && ch.y != kEmptyChild) ? ((__float_as_uint(tn1) & ~3u) | 1u) : 0xFFFFFFFFu;
        uint32_t k2 = (tn2 <= tf2 && ch.z != kEmptyChild) ? ((__float_as_uint(tn2) & ~3u) | 2u) : 0xFFFFFFFFu;
        uint32_t k3 = (tn3 <= tf3 && ch.w != kEmptyChild) ? ((__float_as_uint(tn3) & ~3u) | 3u) : 0xFFFFFFFFu;
        // 4-element sorting network (ascending)
        uint32_t t0 = min(k0, k1), t1 = max(k0, k1), t2 = min(k2, k3), t3 = max(k2, k3);
        uint32_t s0 = min(t0, t2), s2 = max(t0, t2), s1 = min(t1, t3), s3 = max(t1, t3);
        uint32_t m1 = min(s1, s2), m2 = max(s1, s2);
        s1 = m1;
        s2 = m2;
        if (s0 != 0xFFFFFFFFu) {
            if (s3 != 0xFFFFFFFFu) st.push(s.sp++, pick4(ch, s3 & 3u), __uint_as_float(s3 & ~3u));
            if (s2 != 0xFFFFFFFFu) st.push(s.sp++, pick4(ch, s2 & 3u), __uint_as_float(s2 & ~3u));
            if (s1 != 0xFFFFFFFFu) st.push(s.sp++, pick4(ch, s1 & 3u), __uint_as_float(s1 & ~3u));
            s.cur = pick4(ch, s0 & 3u);
            return false;
        }
    } else {
        // ---- leaf: test its primitives ----
        if (!leaf_ok) return false;
        uint32_t first = leaf_first(s.cur), count = leaf_count(s.cur);
        for (uint32_t k = 0; k < count; k++) {
            const float4* pp = sc.prims + 3 * (size_t)(first + k);
            float4 a = pp[0];
            float t, u = 0.0f, v = 0.0f;
            bool h;
            if (COUNT) cnt->prims++;
            if (!MIXED) {
                float4 b = pp[1], c = pp[2];
                h = tri_test(r, s.w, a, b, c, s.tbest, t, u, v);
            } else {
                uint32_t tag = __float_as_uint(a.w);
                uint32_t kind = prim_kind(tag);
                float4 b = pp[1];
                if (kind == kPrimTriangle) {
                    float4 c = pp[2];
                    h = tri_test(r, s.w, a, b, c, s.tbest, t, u, v);
                } else if (kind == kPrimSphere) {
                    h = sphere_test(r, a, b, s.tbest, t);
                } else if (kind == kPrimQuad) {
                    h = quad_test(r, a, sc.quads[prim_side(tag)], s.tbest, t, u, v);
                } else if (kind == kPrimDisc) {
                    h = disc_test(r, a, b, pp[2], s.tbest, t);
                } else if (kind == kPrimCylinder) {
                    h = cylinder_test(r, a, b, pp[2], s.tbest, t, u);
                } else {
                    h = cone_test(r, a, b, pp[2], s.tbest, t, u);
                }
            }
            // closest hit: the least (t, primitive index) pair, so that the result does not depend on the
            // order in which leaves are visited (an equal t is kept only from a lower index)
            if (h && (ANY || t < s.tbest || first + k < s.best)) {
                s.tbest = t;
                s.best = first + k;
                s.bu = u;
                s.bv = v;
                if (ANY) return true;
            }
        }
    }
    // pop the nearest pending entry that can still hold a closer hit
    for (;;) {
        if (s.sp == 0) return true;
        float tn;
        st.pop(--s.sp, s.cur, tn);
        if (tn <= s.tbest) return false;
    }
}

// Copies the first NCACHE nodes (the top levels in breadth-first order) into LDS; all threads
// of the block must call it before traversing.
template <int NCACHE, int BLOCK>
__device__ __forceinline__ void load_node_cache(const DevScene& sc, float4* cache) {
    const int n = (NCACHE < sc.num_nodes ? NCACHE : sc.num_nodes) * 8;
    const float4* src = reinterpret_cast<const float4*>(sc.nodes);
    for (int k = threadIdx.x; k < n; k += BLOCK) cache[k] = src[k];
    __syncthreads();
}

// Flat traversal for scenes of at most kFlatMaxPrims primitives: every primitive is tested in a loop that
// is uniform across the wave (no node steps, no stack, no divergence), keeping the least (t, primitive
// index) pair as the BVH traversal does, so hits are identical. The render kernels' traversals use it
// when their leaf-gate argument is kGateFlat (a per-scene autotuner choice).
constexpr int kGateFlat = 9;
constexpr int32_t kFlatMaxPrims = 64;
template <bool ANY, bool MIXED, bool COUNT>
__device__ __forceinline__ void trace_flat(const DevScene& sc, TState& s, TCounts* cnt) {
    const TRay& r = s.r;
    for (uint32_t k = 0; k < (uint32_t)sc.num_prims; k++) {
        const float4* pp = sc.prims + 3 * (size_t)k;
        const float4 a = pp[0];
        float t, u = 0.0f, v = 0.0f;
        bool h;
        if (COUNT) cnt->prims++;
        if (!MIXED) {
            const float4 b = pp[1], c = pp[2];
            h = tri_test(r, s.w, a, b, c, s.tbest, t, u, v);
        } else {
            const uint32_t tag = __float_as_uint(a.w);
            const uint32_t kind = prim_kind(tag);
            const float4 b = pp[1];
            if (kind == kPrimTriangle) {
                h = tri_test(r, s.w, a, b, pp[2], s.tbest, t, u, v);
            } else if (kind == kPrimSphere) {
                h = sphere_test(r, a, b, s.tbest, t);
            } else if (kind == kPrimQuad) {
                h = quad_test(r, a, sc.quads[prim_side(tag)], s.tbest, t, u, v);
            } else if (kind == kPrimDisc) {
                h = disc_test(r, a, b, pp[2], s.tbest, t);
            } else if (kind == kPrimCylinder) {
                h = cylinder_test(r, a, b, pp[2], s.tbest, t, u);
            } else {
                h = cone_test(r, a, b, pp[2], s.tbest, t, u);
            }
        }
        if (h && (ANY || t < s.tbest || k < s.best)) {
            s.tbest = t;
            s.best = k;
            s.bu = u;
            s.bv = v;
            if (ANY) break;
        }
    }
}

__device__ __forceinline__ bool flat_gate(const DevScene& sc, int gate) {
    return gate == kGateFlat && sc.num_prims <= kFlatMaxPrims;
}

// Closest hit (ANY = false) or any hit (ANY = true) of one ray. Returns true on a hit.
// Always inlined: its callers have one call site each, and a call would pass the ray, hit and stack
// through scratch memory (AMDGPU's inliner gives up on large kernels by basic-block count).
// gate > 0 (wave-uniform) gates leaves (leaf_gate) at gate/8 of the running lanes.
template <bool ANY, bool MIXED, bool COUNT, int LDS_DEPTH, int BLOCK, int NCACHE = 0>
__device__ __forceinline__ bool trace(const DevScene& sc, const TRay& r_in, THit& hit, TStack<LDS_DEPTH, BLOCK>& st,
                                      TCounts* cnt, int gate = 0) {
    TState s;
    tstate_init(s, r_in);
    if (flat_gate(sc, gate)) {
        trace_flat<ANY, MIXED, COUNT>(sc, s, cnt);
    } else {
        for (;;) {
            const bool ok = gate > 0 ? leaf_gate(true, s.cur, gate) : true;
            if (tstate_step<ANY, MIXED, COUNT, LDS_DEPTH, BLOCK, NCACHE>(sc, s, st, cnt, ok)) break;
        }
    }
    tstate_result(s, hit);
    return s.best != 0xFFFFFFFFu;
}

// trace() with a step budget. After `budget` steps that made progress (0: no budget) an unfinished ray
// offers itself to `park` (a callable returning true when it took the ray) and, when taken, stops with
// its best hit so far in `hit`; a ray the callable refuses keeps traversing. Returns kTraceMiss, kTraceHit or
// kTraceParked. Whether a ray is parked does not change its result: the closest hit is the least
// (t, primitive index) pair whichever traversal finds it.
constexpr int kTraceMiss = 0, kTraceHit = 1, kTraceParked = 2;
template <bool ANY, bool MIXED, bool COUNT, int LDS_DEPTH, int BLOCK, int NCACHE, typename Park>
__device__ __forceinline__ int trace_budget(const DevScene& sc, const TRay& r_in, THit& hit,
                                            TStack<LDS_DEPTH, BLOCK>& st, TCounts* cnt, int gate, uint32_t budget,
                                            Park park) {
    TState s;
    tstate_init(s, r_in);
    uint32_t steps = 0;
    bool parked = false;
    if (flat_gate(sc, gate)) {  // a flat traversal has no long rays to park
        trace_flat<ANY, MIXED, COUNT>(sc, s, cnt);
        tstate_result(s, hit);
        return s.best != 0xFFFFFFFFu ? kTraceHit : kTraceMiss;
    }
    for (;;) {
        const bool ok = gate > 0 ? leaf_gate(true, s.cur, gate) : true;
        const bool progress = ok || !child_is_leaf(s.cur);  // a lane held at a leaf by the gate does not count
        if (tstate_step<ANY, MIXED, COUNT, LDS_DEPTH, BLOCK, NCACHE>(sc, s, st, cnt, ok)) break;
        if (progress && ++steps == budget && park()) {
            parked = true;
            break;
        }
    }
    tstate_result(s, hit);
    return parked ? kTraceParked : (s.best != 0xFFFFFFFFu ? kTraceHit : kTraceMiss);
}

// Minimum of an unsigned value over each group of G lanes (G = 16, 32 or 64, groups aligned to G); every
// lane of the wave must be active. DPP steps reduce each row of 16 lanes (swap neighbours, swap pairs,
// mirror halves of 8, mirror the row); rows are then combined by lane swaps (G = 32) or read as scalars
// (G = 64). Every lane receives its group's minimum.
template <int G>
__device__ __forceinline__ uint32_t group_min_u32(uint32_t v) {
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x140, 0xF, 0xF, false));  // row_mirror
    if constexpr (G == 32) {
        v = min(v, (uint32_t)__shfl_xor((int)v, 16, 64));
    } else if constexpr (G == 64) {
        const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)v, 0), b = (uint32_t)__builtin_amdgcn_readlane((int)v, 16);
        const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)v, 32), d = (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
        v = min(min(a, b), min(c, d));
    }
    return v;
}

// Minimum of a non-negative float (or +inf) over each group of G lanes: such floats order like their bits.
template <int G>
__device__ __forceinline__ float group_min_nonneg(float x) {
    return __uint_as_float(group_min_u32<G>(__float_as_uint(x)));
}

// Exclusive prefix sum of a count in [0, 7] over the lanes of the group `gmask`, and the group's total.
__device__ __forceinline__ uint32_t group_scan7(uint32_t c, uint64_t gmask, uint32_t& total) {
    const uint64_t lt = ((1ull << (threadIdx.x & 63u)) - 1ull) & gmask;
    const uint64_t b0 = __ballot(c & 1u) & gmask, b1 = __ballot(c & 2u) & gmask, b2 = __ballot(c & 4u) & gmask;
    total = (uint32_t)(__popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2));
    return (uint32_t)(__popcll(b0 & lt) + 2 * __popcll(b1 & lt) + 4 * __popcll(b2 & lt));
}

// Cooperative traversal of one ray per group of G lanes (G = 16, 32 or 64; a one-wave work-group holds
// 64 / G groups, every lane of a group passes its group's ray). For rays that visit hundreds of nodes
// (grazing rays along a dense mesh), whose one-lane traversal is a long chain of dependent node fetches:
//  1. the group expands a frontier breadth-first, a node per lane per round, until it holds at least G
//     entries or only leaves (at most 4 (G - 1): fewer than G nodes expand into at most 4 each);
//  2. each lane traverses frontier entries depth-first with the one-lane step (its own stack), taking
//     the next untaken entry when it finishes one; after every step the group's lanes share their least
//     hit distance to cull with.
// `hit` holds the ray's best hit so far on entry (prim 0xFFFFFFFF: none) and the least (t, primitive
// index) pair over all primitive hits with t in [tmin, tbest] on exit, the result of trace(). An any-hit
// ray (`any`, group-uniform) stops as soon as one of its lanes records a hit. `has`: the group has a ray.
// Every lane of the wave must call it together. f_node / f_t: 512 LDS entries each (8 G per group).
// Returns true on a hit.
constexpr uint32_t kCoopEntries = 512;  // LDS frontier entries of a one-wave work-group (8 per lane)
template <bool MIXED, int LDS_DEPTH, int G>
__device__ __forceinline__ bool coop_trace(const DevScene& sc, const TRay& r_in, THit& hit, TStack<LDS_DEPTH, 64>& st,
                                           uint32_t* f_node, float* f_t, bool has, bool any) {
    static_assert(G == 16 || G == 32 || G == 64, "group of 16, 32 or 64 lanes");
    constexpr uint32_t FC = 4u * G;  // frontier capacity per buffer
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gbase = lane & ~(uint32_t)(G - 1), gl = lane - gbase;
    const uint64_t gmask = G == 64 ? ~0ull : (((1ull << G) - 1ull) << gbase);
    const uint32_t fb = gbase * 8u;  // the group's two frontier buffers
    const float kInfF = __int_as_float(0x7f800000);
    TState s;
    tstate_init(s, r_in);
    const TRay& r = s.r;
    // group-shared bound: the best hit distance found so far by any lane of the group (or the entry best)
    float bound = hit.prim != 0xFFFFFFFFu ? hit.t : s.tbest;
    uint32_t nf = has ? 1u : 0u, cur = 0;  // frontier size, current buffer (group-uniform)
    if (has && gl == 0) {
        f_node[fb] = 0;  // the root
        f_t[fb] = r.tmin;
    }
    __syncthreads();
    // ---- 1. breadth-first frontier ----
    bool expanding = has && nf < (uint32_t)G;
    while (__ballot(expanding) != 0) {
        bool live = false, inner = false;
        uint32_t node = 0;
        float tn = kInfF;
        if (expanding && gl < nf) {
            node = f_node[fb + cur * FC + gl];
            tn = f_t[fb + cur * FC + gl];
            live = tn <= bound;
            inner = live && !child_is_leaf(node);
        }
        const bool grow = expanding && (__ballot(inner) & gmask) != 0;  // group-uniform
        // children the ray enters before the bound (entry distances rounded down like the one-lane step)
        bool in[4] = {false, false, false, false};
        float ent[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        uint4 ch = make_uint4(0, 0, 0, 0);
        uint32_t cnt = grow && live ? 1u : 0u;
        if (grow && inner) {
            const float4* np = reinterpret_cast<const float4*>(sc.nodes + node);
            const float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
            ch = reinterpret_cast<const uint4*>(np)[6];
            const float lo[3][4] = {{lx.x, lx.y, lx.z, lx.w}, {ly.x, ly.y, ly.z, ly.w}, {lz.x, lz.y, lz.z, lz.w}};
            const float hi[3][4] = {{hx.x, hx.y, hx.z, hx.w}, {hy.x, hy.y, hy.z, hy.w}, {hz.x, hz.y, hz.z, hz.w}};
            cnt = 0;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const float nx = ((s.idx >= 0.0f ? lo[0][c] : hi[0][c]) - r.ox) * s.idx;
                const float fx = ((s.idx >= 0.0f ? hi[0][c] : lo[0][c]) - r.ox) * s.idx;
                const float ny = ((s.idy >= 0.0f ? lo[1][c] : hi[1][c]) - r.oy) * s.idy;
                const float fy = ((s.idy >= 0.0f ? hi[1][c] : lo[1][c]) - r.oy) * s.idy;
                const float nz = ((s.idz >= 0.0f ? lo[2][c] : hi[2][c]) - r.oz) * s.idz;
                const float fz = ((s.idz >= 0.0f ? hi[2][c] : lo[2][c]) - r.oz) * s.idz;
                const float tnear = fmaxf(fmaxf(nx, ny), fmaxf(nz, r.tmin));
                const float tfar = fminf(fminf(fminf(fx, fy), fz) * kBoxWiden, bound);
                in[c] = tnear <= tfar && pick4(ch, (uint32_t)c) != kEmptyChild;
                ent[c] = __uint_as_float(__float_as_uint(tnear) & ~3u);
                cnt += in[c] ? 1u : 0u;
            }
        }
        uint32_t total;
        const uint32_t off = group_scan7(cnt, gmask, total);
        const uint32_t nxt = cur ^ 1u;
        if (grow) {
            if (inner) {
                uint32_t k = off;
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    if (in[c]) {
                        f_node[fb + nxt * FC + k] = pick4(ch, (uint32_t)c);
                        f_t[fb + nxt * FC + k] = ent[c];
                        k++;
                    }
                }
            } else if (live) {
                f_node[fb + nxt * FC + off] = node;
                f_t[fb + nxt * FC + off] = tn;
            }
        }
        __syncthreads();
        if (grow) {
            cur = nxt;
            nf = total;
        }
        // only leaves (or culled entries) left, or enough entries: phase 2 takes the frontier as it is
        expanding = grow && nf < (uint32_t)G;
    }
    // ---- 2. depth-first traversal of the frontier entries, lanes taking entries in turn ----
    bool active = false, owns = false, any_hit = false;
    bool gdone = !has;   // group-uniform
    uint32_t taken = 0;  // group-uniform: entries handed out so far
    uint32_t obest = 0xFFFFFFFFu;
    float ot = kInfF, ou = 0.0f, ov = 0.0f;
    const uint64_t lt = ((1ull << lane) - 1ull) & gmask;
    for (;;) {
        const uint64_t idle = __ballot(!active) & gmask;
        if (!gdone && taken < nf) {
            const uint32_t rank = (uint32_t)__popcll(idle & lt);
            if (!active && taken + rank < nf) {
                const uint32_t e = taken + rank;
                const float tn = f_t[fb + cur * FC + e];
                if (tn <= bound) {
                    s.cur = f_node[fb + cur * FC + e];
                    s.sp = 0;
                    s.tbest = bound;
                    s.best = 0xFFFFFFFFu;
                    active = true;
                }
            }
            taken = min(nf, taken + (uint32_t)__popcll(idle));
        }
        const uint64_t gact = __ballot(active) & gmask;
        if (gact == 0 && taken >= nf) gdone = true;
        if (any && (__ballot(any_hit) & gmask) != 0) gdone = true;
        if (__ballot(!gdone) == 0) break;
        if (!gdone && active) {
            // adopt a closer bound found by another lane (the lane's own candidate no longer wins)
            if (bound < s.tbest) {
                s.tbest = bound;
                s.best = 0xFFFFFFFFu;
            }
            const bool done = tstate_step<false, MIXED, false, LDS_DEPTH, 64, 0>(sc, s, st, nullptr, true);
            if (s.best != 0xFFFFFFFFu) {  // the lane's least (t, index) pair so far
                if (!owns || s.tbest < ot || (s.tbest == ot && s.best < obest)) {
                    ot = s.tbest;
                    obest = s.best;
                    ou = s.bu;
                    ov = s.bv;
                }
                owns = true;
                any_hit = true;
            }
            if (done) active = false;
        }
        bound = fminf(bound, group_min_nonneg<G>(owns ? ot : kInfF));
    }
    // the least (t, index) pair over the group's lanes and the entry best
    const float wt = group_min_nonneg<G>(owns ? ot : kInfF);
    const uint32_t wp = group_min_u32<G>(owns && ot == wt ? obest : 0xFFFFFFFFu);
    const uint64_t holder = __ballot(owns && ot == wt && obest == wp) & gmask;
    const int hl = holder ? __ffsll((unsigned long long)holder) - 1 : (int)lane;
    const float hu = __shfl(ou, hl, 64), hv = __shfl(ov, hl, 64);
    if (any) {
        const bool found = wp != 0xFFFFFFFFu || hit.prim != 0xFFFFFFFFu;
        if (found) hit.prim = wp != 0xFFFFFFFFu ? wp : hit.prim;
        return found;
    }
    if (wp == 0xFFFFFFFFu) return hit.prim != 0xFFFFFFFFu;  // nothing closer than the entry best
    const bool better = hit.prim == 0xFFFFFFFFu || wt < hit.t || (wt == hit.t && wp < hit.prim);
    if (!better) return true;
    hit.t = wt;
    hit.prim = wp;
    hit.u = hu;
    hit.v = hv;
    return true;
}

}  // namespace prt
