// bdpt_device.h — float32 device bidirectional path tracer (pkg/integrator/bdpt.go, bdpt_mis.go).
//
// One thread per camera sample: camera subpath, light subpath, then every (s, t) strategy with its
// MIS weight (calculateMISWeight, bdpt_mis.go:13-48) and an any-hit visibility test. Subpath vertices
// live in global memory, vertex k of sample i at [k * stride + i] (coalesced across a wave). Light
// tracing (t = 1) contributions are splatted to the pixel the camera ray maps to (MapRayToPixel) with
// 64-bit fixed-point atomics, so the splat image is independent of the order of the additions.
#pragma once
#include "shading.h"
#include "traverse.h"

namespace prt {

// Draw slots of the counter stream: camera path bounce b: 1+b; light emission: kLightSlot; light path
// bounce b: kLightSlot+1+b; direct lighting at camera vertex t: kDirectSlot+t; light tracing from light
// vertex s: kSplatSlot+s (positions k within a slot).
constexpr uint32_t kLightSlot = 70, kDirectSlot = 140, kSplatSlot = 210;
constexpr int kBdptMaxDepth = 64;
constexpr float kSplatScale = 16777216.0f;  // fixed point: 2^24 units per radiance unit

// VF_AREA / VF_DELTA: a vertex on a light (light >= 0) whose type is an area light (LightTypeArea) /
// a point light (LightTypePoint), set where the vertex is made so that the MIS weights read no light.
// VF_CONN: the vertex counts in its subpath's MIS walk (store_walk_sums).
enum : uint32_t {
    VF_LIGHT = 1, VF_CAMERA = 2, VF_SPECULAR = 4, VF_INF = 8, VF_FRONT = 16, VF_AREA = 32, VF_DELTA = 64, VF_CONN = 128
};

// Vertex, bdpt.go:14-44 (Light is the index of a scene light, -1 when the vertex carries none)
struct BV {
    V3 p, n, wi, beta, Le;
    float pdf_fwd, pdf_rev;
    int mat, light;
    uint32_t flags;
    float u = 0.0f, v = 0.0f;  // surface UV (textured scenes)
    float q = 0.0f;            // MIS walk sum below this vertex (store_walk_sums)
};

// Subpath vertices in global memory as five planes of float4 (six with surface UVs in textured
// scenes), plane f of vertex k of sample i at [(f * kmax + k) * stride + i] (each plane coalesced
// across a wave). Plane 0 is the vertex record (pdf_fwd, pdf_rev, flags | (mat + 1) << 8, light);
// plane 4 holds (Le, q).
struct VertexStore {
    float4* data;
    uint32_t stride;  // samples per chunk
    uint32_t kmax;    // vertices per subpath
    uint32_t uv;      // 1: plane 5 holds (u, v)
    __device__ __forceinline__ float4* at(int f, int k, uint32_t i) const {
        return data + ((size_t)f * kmax + (size_t)k) * stride + i;
    }
    __device__ __forceinline__ void put(int k, uint32_t i, const BV& v) const {
        *at(0, k, i) = make_float4(v.pdf_fwd, v.pdf_rev, __uint_as_float(v.flags | ((uint32_t)(v.mat + 1) << 8)),
                                   __int_as_float(v.light));
        *at(1, k, i) = make_float4(v.p.x, v.p.y, v.p.z, v.beta.x);
        *at(2, k, i) = make_float4(v.n.x, v.n.y, v.n.z, v.beta.y);
        *at(3, k, i) = make_float4(v.wi.x, v.wi.y, v.wi.z, v.beta.z);
        *at(4, k, i) = make_float4(v.Le.x, v.Le.y, v.Le.z, v.q);
        if (uv) *at(5, k, i) = make_float4(v.u, v.v, 0.0f, 0.0f);
    }
    __device__ __forceinline__ float4 rec(int k, uint32_t i) const { return *at(0, k, i); }
    __device__ __forceinline__ BV get(int k, uint32_t i) const {
        float4 r = *at(0, k, i), a = *at(1, k, i), b = *at(2, k, i), c = *at(3, k, i), e = *at(4, k, i);
        BV v;
        v.pdf_fwd = r.x;
        v.pdf_rev = r.y;
        uint32_t fm = __float_as_uint(r.z);
        v.flags = fm & 0xFFu;
        v.mat = (int)(fm >> 8) - 1;
        v.light = __float_as_int(r.w);
        v.p = v3(a.x, a.y, a.z);
        v.n = v3(b.x, b.y, b.z);
        v.wi = v3(c.x, c.y, c.z);
        v.beta = v3(a.w, b.w, c.w);
        v.Le = v3(e.x, e.y, e.z);
        v.q = e.w;
        if (uv) {
            float4 t = *at(5, k, i);
            v.u = t.x;
            v.v = t.y;
        }
        return v;
    }
    __device__ __forceinline__ void set_pdf_rev(int k, uint32_t i, float pdf) const { at(0, k, i)->y = pdf; }
    __device__ __forceinline__ void set_pdf_fwd(int k, uint32_t i, float pdf) const { at(0, k, i)->x = pdf; }
};
__device__ __forceinline__ bool rec_specular(float4 r) { return (__float_as_uint(r.z) & VF_SPECULAR) != 0; }

__device__ __forceinline__ bool light_area(const DevScene& sc, int li) {  // LightTypeArea
    int t = sc.lights[li].type;
    return t == PRT_LIGHT_QUAD || t == PRT_LIGHT_SPHERE || t == PRT_LIGHT_DISC || t == PRT_LIGHT_DISC_SPOT;
}
__device__ __forceinline__ bool light_delta(const DevScene& sc, int li) {  // LightTypePoint
    return sc.lights[li].type == PRT_LIGHT_POINT_SPOT;
}
// flags of a vertex made on light li
__device__ __forceinline__ uint32_t light_vertex_flags(const DevScene& sc, int li) {
    const int t = sc.lights[li].type;
    return VF_LIGHT | (light_is_infinite(t) ? VF_INF : 0u) | (light_area(sc, li) ? VF_AREA : 0u) |
           (light_delta(sc, li) ? VF_DELTA : 0u);
}
__device__ __forceinline__ bool on_surface(const DevScene& sc, const BV& v) {  // IsOnSurface :33-41
    if ((v.flags & VF_LIGHT) && v.light >= 0) return (v.flags & VF_AREA) != 0;
    return v.mat >= 0;
}
__device__ __forceinline__ Surf surf_of(const BV& v) {
    Surf s;
    s.p = v.p;
    s.n = v.n;
    s.front = (v.flags & VF_FRONT) != 0;
    s.mat = v.mat;
    s.u = v.u;
    s.v = v.v;
    return s;
}

// convertSolidAngleToAreaPdf, bdpt_mis.go:341-362
__device__ __forceinline__ float convert_density(const DevScene& sc, const BV& from, const BV& next, float pdf) {
    if (next.flags & VF_INF) return pdf;
    V3 d = next.p - from.p;
    float d2 = dot(d, d);
    if (d2 == 0.0f) return 0.0f;
    float inv = 1.0f / d2;
    if (on_surface(sc, next)) pdf *= fabsf(dot(d * sqrtf(inv), next.n));
    return pdf * inv;
}
__device__ __forceinline__ float remap0(float f) { return f != 0.0f ? f : 1.0f; }

// Camera.CalculateRayPDFs / MapRayToPixel / SampleCameraFromPoint (camera.go:194-316)
__device__ __forceinline__ bool cam_map(const DevCamera& c, V3 o, V3 d, int& px, int& py) {
    V3 w = v3(c.w);
    V3 ipc = v3(c.center) - w * c.focus;
    float t = dot(ipc - o, w) / dot(d, w);
    if (!(t > 0.0f)) return false;
    V3 rel = (o + d * t) - ipc;
    float nx = (dot(rel, v3(c.u)) + c.vw / 2.0f) / c.vw, ny = (dot(rel, v3(c.v)) + c.vh / 2.0f) / c.vh;
    float fx = nx * (float)c.width, fy = ny * (float)c.height;
    if (!(fx > -1.0f && fx < (float)c.width + 1.0f && fy > -1.0f && fy < (float)c.height + 1.0f)) return false;
    px = (int)fx;  // truncation toward zero as Go's int()
    py = (int)fy;
    return px >= 0 && px < c.width && py >= 0 && py < c.height;
}
__device__ __forceinline__ float cam_dir_pdf(const DevCamera& c, V3 o, V3 d) {
    float ct = dot(d, v3(c.forward));
    if (ct <= c.cos_total_width) return 0.0f;
    int x, y;
    if (!cam_map(c, o, d, x, y)) return 0.0f;
    return 1.0f / (c.plane_area * ct * ct * ct);
}

// EmissionSample + Light.SampleEmission / PDF_Le (light_sampling.go:54-115, *_light.go)
struct ESample {
    V3 p, n, dir, emission;
    float area_pdf, dir_pdf;
};
__device__ __forceinline__ V3 sample_on_unit_sphere(float s0, float s1) {  // sampling.go:97-104
    float z = 1.0f - 2.0f * s0;
    float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float sp, cp;
    sincosf(2.0f * kPiF * s1, &sp, &cp);
    return v3(r * cp, r * sp, z);
}
__device__ __forceinline__ void infinite_disk(const DevScene& sc, V3 d, float s0, float s1, ESample& es) {
    V3 up = fabsf(d.x) > 0.9f ? v3(0, 1, 0) : v3(1, 0, 0);
    V3 right = normalize(cross(d, up));
    up = normalize(cross(right, d));
    V3 ds = sample_point_in_unit_disk(s0, s1);
    float R = sc.world_radius;
    V3 dp = v3(sc.world_center) + right * (ds.x * R) + up * (ds.y * R);
    es.p = dp + d * -R;
    es.dir = d;
    es.n = d * -1.0f;
    es.area_pdf = 1.0f / (kPiF * R * R);
}
__device__ __forceinline__ ESample light_sample_emission(const DevScene& sc, const DevLight& L, float p0, float p1,
                                                         float d0, float d1) {
    ESample es;
    if (L.type == PRT_LIGHT_QUAD) {
        es.p = v3(L.corner) + v3(L.u) * p0 + v3(L.v) * p1;
        es.n = v3(L.n);
        es.dir = sample_cosine_hemisphere(es.n, d0, d1);
        es.area_pdf = 1.0f / L.area;
        es.dir_pdf = dot(es.dir, es.n) / kPiF;
        es.emission = emit_dir(sc, L.mat, es.dir);
    } else if (L.type == PRT_LIGHT_DISC) {  // disc_light.go:104-111 + SampleEmissionDirection
        disc_sample_uniform(L, p0, p1, es.p, es.n);
        es.area_pdf = 1.0f / (kPiF * L.radius * L.radius);
        es.dir = sample_cosine_hemisphere(es.n, d0, d1);
        es.dir_pdf = dot(es.dir, es.n) / kPiF;
        es.emission = emit_dir(sc, L.mat, es.dir);
    } else if (L.type == PRT_LIGHT_DISC_SPOT) {  // disc_spot_light.go:170-192
        disc_sample_uniform(L, p0, p1, es.p, es.n);
        es.dir = sample_cone(v3(L.spot_dir), L.cos_total, d0, d1);
        es.emission = v3(L.emission) * spot_falloff(L, dot(es.dir, v3(L.spot_dir)));
        es.area_pdf = 1.0f / (kPiF * L.radius * L.radius);
        es.dir_pdf = uniform_cone_pdf(L.cos_total);
    } else if (L.type == PRT_LIGHT_POINT_SPOT) {  // point_spot_light.go:153-182
        es.p = v3(L.from);
        es.dir = sample_cone(v3(L.spot_dir), L.cos_total, d0, d1);
        es.emission = v3(L.emission) * spot_falloff(L, dot(es.dir, v3(L.spot_dir)));
        es.n = es.dir;
        es.area_pdf = 1.0f;
        es.dir_pdf = uniform_cone_pdf(L.cos_total);
    } else if (L.type == PRT_LIGHT_SPHERE) {
        V3 ld = sample_on_unit_sphere(p0, p1);
        es.p = v3(L.center) + ld * L.radius;
        es.n = ld;
        es.area_pdf = 1.0f / (4.0f * kPiF * L.radius * L.radius);
        es.dir = sample_cosine_hemisphere(ld, d0, d1);
        es.dir_pdf = dot(es.dir, ld) / kPiF;
        es.emission = emit_dir(sc, L.mat, es.dir);
    } else if (L.type == PRT_LIGHT_ENVMAP) {
        float pdf;
        V3 wi = env_sample(sc.env, d0, d1, pdf);
        infinite_disk(sc, wi * -1.0f, p0, p1, es);
        es.dir_pdf = pdf;
        es.emission = env_eval(sc.env, wi);
    } else {
        V3 d = sample_on_unit_sphere(d0, d1);
        infinite_disk(sc, d, p0, p1, es);
        es.dir_pdf = 1.0f / (4.0f * kPiF);
        if (L.type == PRT_LIGHT_UNIFORM_INFINITE) {
            es.emission = v3(L.emission);
        } else {
            float t = 0.5f * (d.y + 1.0f);
            es.emission = v3(L.bottom) * (1.0f - t) + v3(L.top) * t;
        }
    }
    return es;
}
__device__ __forceinline__ void light_pdf_le(const DevScene& sc, const DevLight& L, V3 p, V3 d, float& pos, float& dir) {
    pos = dir = 0.0f;
    if (L.type == PRT_LIGHT_QUAD) {  // quad_light.go:129-176
        V3 tp = p - v3(L.corner), U = v3(L.u), W = v3(L.v);
        float uu = dot(U, U), vv = dot(W, W), uv = dot(U, W);
        if (uu == 0.0f || vv == 0.0f) return;
        float det = uu * vv - uv * uv;
        if (fabsf(det) < 1e-8f) return;
        float tu = dot(tp, U), tv = dot(tp, W);
        float a = (vv * tu - uv * tv) / det, b = (uu * tv - uv * tu) / det;
        if (a < 0.0f || a > 1.0f || b < 0.0f || b > 1.0f) return;
        V3 rec = v3(L.corner) + U * a + W * b;
        if (length(rec - p) > 0.001f) return;
        pos = 1.0f / L.area;
        float ct = dot(d, v3(L.n));
        if (ct <= 0.0f) return;
        dir = ct / kPiF;
    } else if (L.type == PRT_LIGHT_DISC || L.type == PRT_LIGHT_DISC_SPOT) {  // disc_light.go:114-133
        V3 tp = p - v3(L.center), n = v3(L.n);  // validatePointOnDisc (:142-155), tolerance 0.001
        float hn = dot(tp, n);
        if (fabsf(hn) > 0.001f || length(tp - n * hn) > L.radius) return;
        float ct = dot(d, n);
        if (ct <= 0.0f) return;
        pos = 1.0f / (kPiF * L.radius * L.radius);
        dir = ct / kPiF;
        if (L.type == PRT_LIGHT_DISC_SPOT && dot(d, v3(L.spot_dir)) < L.cos_total) dir = 0.0f;  // :195-210
    } else if (L.type == PRT_LIGHT_POINT_SPOT) {  // point_spot_light.go:185-207
        if (length(p - v3(L.from)) > 0.001f) return;
        if (dot(d, v3(L.spot_dir)) < L.cos_total) return;
        pos = 1.0f;
        dir = uniform_cone_pdf(L.cos_total);
    } else if (L.type == PRT_LIGHT_SPHERE) {  // sphere_light.go:184-205
        V3 c = v3(L.center);
        if (!(fabsf(length(p - c) - L.radius) <= 0.001f)) return;
        V3 n = normalize(p - c);
        float ct = dot(d, n);
        if (ct <= 0.0f) return;
        pos = 1.0f / (4.0f * kPiF * L.radius * L.radius);
        dir = ct / kPiF;
    } else {
        float R = sc.world_radius;
        if (R <= 0.0f) return;
        pos = 1.0f / (kPiF * R * R);
        dir = L.type == PRT_LIGHT_ENVMAP ? env_pdf(sc.env, d * -1.0f) : 1.0f / (4.0f * kPiF);
    }
}

// InfiniteLightDensity, bdpt_mis.go:318-337
__device__ __forceinline__ float infinite_density(const DevScene& sc, V3 p, V3 n, V3 d) {
    float tot = 0.0f;
    for (int i = 0; i < sc.num_lights; i++)
        if (light_is_infinite(sc.lights[i].type)) tot += light_pdf(sc, sc.lights[i], p, n, d) * sc.light_weights[i];
    return tot;
}

// evaluateBRDF, bdpt.go:564-583
__device__ __forceinline__ V3 vertex_brdf(const DevScene& sc, const BV& v, V3 out, int mode) {
    if (v.flags & VF_INF) return v3(0, 0, 0);
    if ((v.flags & VF_LIGHT) && v.mat < 0) return v3(1, 1, 1);
    if (v.mat < 0) return v3(0, 0, 0);
    return brdf(sc, v.mat, v.wi, out, surf_of(v), mode);
}

// calculateLightPdf / calculateLightOriginPdf / calculateVertexPdf, bdpt_mis.go:190-315
__device__ __forceinline__ float vtx_light_pdf(const DevScene& sc, const BV& cur, const BV& to) {
    V3 w = to.p - cur.p;
    float inv = 1.0f / dot(w, w);
    w = w * sqrtf(inv);
    float pdf = 0.0f;
    if (cur.flags & VF_LIGHT) {
        if (cur.flags & VF_INF) {
            float R = sc.world_radius;
            if (R == 0.0f) R = 1.0f;
            pdf = 1.0f / (kPiF * R * R);
        } else if (cur.light >= 0) {
            float pos, dir;
            light_pdf_le(sc, sc.lights[cur.light], cur.p, w, pos, dir);
            pdf = dir * inv;
        }
    }
    if (on_surface(sc, to)) pdf *= fabsf(dot(to.n, w));
    return pdf;
}
__device__ __forceinline__ float vtx_light_origin_pdf(const DevScene& sc, const BV& lv, const BV& to) {
    V3 w = to.p - lv.p;
    if (dot(w, w) == 0.0f) return 0.0f;
    w = normalize(w);
    if (lv.flags & VF_INF) return infinite_density(sc, to.p, to.n, w * -1.0f);
    if (!(lv.flags & VF_LIGHT) || lv.light < 0 || sc.num_lights == 0) return 0.0f;
    float pos, dir;
    light_pdf_le(sc, sc.lights[lv.light], lv.p, w, pos, dir);
    return pos * sc.light_weights[lv.light];
}
__device__ __forceinline__ float vtx_pdf(const DevScene& sc, const DevCamera& cam, const BV& cur, const BV* prev,
                                         const BV& next) {
    if (cur.flags & VF_LIGHT) return vtx_light_pdf(sc, cur, next);
    V3 wn = next.p - cur.p;
    if (dot(wn, wn) == 0.0f) return 0.0f;
    wn = normalize(wn);
    V3 wp = v3(0, 0, 0);
    if (prev) {
        wp = prev->p - cur.p;
        if (dot(wp, wp) == 0.0f) return 0.0f;
        wp = normalize(wp);
    } else if (!(cur.flags & VF_CAMERA)) {
        return 0.0f;
    }
    float pdf;
    if (cur.flags & VF_CAMERA) {
        pdf = cam_dir_pdf(cam, cur.p, wn);
        if (pdf == 0.0f) return 0.0f;
    } else if (cur.mat >= 0) {
        bool delta;
        pdf = mat_pdf(sc, cur.mat, wp, wn, cur.n, delta);
        if (delta) return 0.0f;
    } else {
        return 0.0f;
    }
    return convert_density(sc, cur, next, pdf);
}

// A segment between two vertices as vtx_pdf uses it, computed once where several densities share it:
// n = normalize(v) (the direction), u = v * sqrt(1 / |v|^2) and inv = 1 / |v|^2 (the area conversion).
// The reversed segment negates n and u exactly.
struct Seg {
    V3 n, u;
    float inv;
    bool zero;
};
__device__ __forceinline__ Seg seg_of(V3 v) {
    Seg g;
    const float d2 = dot(v, v);
    g.zero = d2 == 0.0f;
    g.n = normalize(v);
    g.inv = 1.0f / d2;
    g.u = v * sqrtf(g.inv);
    return g;
}
__device__ __forceinline__ Seg seg_rev(const Seg& g) {
    Seg r = g;
    r.n = g.n * -1.0f;
    r.u = g.u * -1.0f;
    return r;
}
// vtx_pdf(cur, prev, next) for a previous vertex, with the segments cur->prev (sp) and cur->next (sn).
__device__ __forceinline__ float vtx_pdf_seg(const DevScene& sc, const DevCamera& cam, const BV& cur, const Seg& sp,
                                             const BV& next, const Seg& sn) {
    if (cur.flags & VF_LIGHT) {  // vtx_light_pdf
        float pdf = 0.0f;
        if (cur.flags & VF_INF) {
            float R = sc.world_radius;
            if (R == 0.0f) R = 1.0f;
            pdf = 1.0f / (kPiF * R * R);
        } else if (cur.light >= 0) {
            float pos, dir;
            light_pdf_le(sc, sc.lights[cur.light], cur.p, sn.u, pos, dir);
            pdf = dir * sn.inv;
        }
        if (on_surface(sc, next)) pdf *= fabsf(dot(next.n, sn.u));
        return pdf;
    }
    if (sn.zero || sp.zero) return 0.0f;
    float pdf;
    if (cur.flags & VF_CAMERA) {
        pdf = cam_dir_pdf(cam, cur.p, sn.n);
        if (pdf == 0.0f) return 0.0f;
    } else if (cur.mat >= 0) {
        bool delta;
        pdf = mat_pdf(sc, cur.mat, sp.n, sn.n, cur.n, delta);
        if (delta) return 0.0f;
    } else {
        return 0.0f;
    }
    if (next.flags & VF_INF) return pdf;  // convert_density
    if (sn.zero) return 0.0f;
    if (on_surface(sc, next)) pdf *= fabsf(dot(sn.u, next.n));
    return pdf * sn.inv;
}

// The part of a subpath's MIS walk that no connection changes, once per subpath (after its densities
// are final). With r_m = remap0(rev_m) / remap0(fwd_m) and conn_m the walk's rule for vertex m (neither m
// nor the vertex before it specular or on a point light; the camera walk ends at vertex 1, the light walk
// at vertex 0), P(m) = sum over k <= m of conn_k * r_k * ... * r_m = r_m * (conn_m + P(m - 1)). Vertex m
// keeps P(m - 1) as q and conn_m as VF_CONN.
__device__ __forceinline__ void store_walk_sums(const VertexStore& VS, uint32_t i, int n, bool light) {
    float P = 0.0f;
    bool prev_ok = true;
    for (int m = 0; m < n; m++) {
        float4* rp = VS.at(0, m, i);
        const float4 r = *rp;
        uint32_t fm = __float_as_uint(r.z);
        const bool ok = !(fm & (VF_SPECULAR | VF_DELTA));
        VS.at(4, m, i)->w = P;
        if (m > 0 || light) {
            const bool conn = ok && prev_ok;
            if (conn) {
                fm |= VF_CONN;
                rp->z = __uint_as_float(fm);
            }
            P = (remap0(r.y) / remap0(r.x)) * ((conn ? 1.0f : 0.0f) + P);
        }
        prev_ok = ok;
    }
}

// calculateMISWeight with the per-vertex rules of calculateMIS{Camera,Light}VertexPdfs (:52-187).
// ct1 = camera vertex t-1, ls1 = light vertex s-1 (s >= 1), sv = the sampled vertex of the s = 1 and
// t = 1 strategies. Only the two vertices next to the connection on each side are read.
__device__ __forceinline__ float mis_weight(const DevScene& sc, const DevCamera& cam, const VertexStore& CP,
                                           const VertexStore& LP, uint32_t i, const BV& sv, const BV& ct1,
                                           const BV& ls1, int s, int t) {
    if (s + t == 2) return 1.0f;
    BV ct2, ls2;
    // each side's walk past its second vertex: (conn + q) and that vertex's forward density
    float c_tail = 0.0f, c_fwd2 = 1.0f, l_tail = 0.0f, l_fwd2 = 1.0f;
    bool c_first = false;
    if (t > 1) {
        ct2 = CP.get(t - 2, i);
        c_first = !(ct2.flags & VF_SPECULAR);
        c_tail = ((ct2.flags & VF_CONN) ? 1.0f : 0.0f) + ct2.q;
        c_fwd2 = remap0(ct2.pdf_fwd);
    }
    if (s > 1) {
        ls2 = LP.get(s - 2, i);
        l_tail = ((ls2.flags & VF_CONN) ? 1.0f : 0.0f) + ls2.q;
        l_fwd2 = remap0(ls2.pdf_fwd);
    }
    // reverse densities of the vertices next to the connection
    float c_rev1 = 0.0f, c_rev2 = 0.0f, l_rev1 = 0.0f, l_rev2 = 0.0f;
    if (s == 0) {
        if (t > 1) c_rev1 = vtx_light_origin_pdf(sc, ct1, ct2);
        if (t > 2) c_rev2 = vtx_pdf(sc, cam, ct1, nullptr, ct2);
    } else if (t == 1) {
        if (s > 1) {
            l_rev1 = vtx_pdf(sc, cam, sv, nullptr, ls1);
            l_rev2 = vtx_pdf(sc, cam, ls1, &sv, ls2);
        }
    } else if (s == 1) {
        c_rev1 = vtx_pdf(sc, cam, sv, nullptr, ct1);
        c_rev2 = vtx_pdf(sc, cam, ct1, &sv, ct2);
        l_rev1 = vtx_pdf(sc, cam, ct1, &ct2, sv);
    } else {  // the three segments the four densities share: ct1->ls1, ct1->ct2, ls1->ls2
        const Seg a = seg_of(ls1.p - ct1.p), b = seg_of(ct2.p - ct1.p), c = seg_of(ls2.p - ls1.p);
        const Seg ar = seg_rev(a);
        c_rev1 = vtx_pdf_seg(sc, cam, ls1, c, ct1, ar);
        c_rev2 = vtx_pdf_seg(sc, cam, ct1, a, ct2, b);
        l_rev1 = vtx_pdf_seg(sc, cam, ct1, b, ls1, a);
        l_rev2 = vtx_pdf_seg(sc, cam, ls1, ar, ls2, c);
    }
    // The walks of calculateMISWeight (bdpt_mis.go:30-50): from the connection outwards, ri is the running
    // product of remap0(rev) / remap0(fwd) and connectible vertices add ri. Only the two vertices next to
    // the connection take the densities above; the rest of each walk is the stored sum q (store_walk_sums).
    float sum = 0.0f;
    if (t > 1) {  // camera walk, vertices t-1 .. 1
        float ri = remap0(c_rev1) / remap0(ct1.pdf_fwd);
        if (c_first) sum += ri;
        if (t > 2) sum += ri * (remap0(c_rev2) / c_fwd2) * c_tail;
    }
    if (s == 1) {  // light walk, the sampled light vertex only
        sum += remap0(l_rev1) / remap0(sv.pdf_fwd);
    } else if (s > 1) {  // light walk, vertices s-1 .. 0
        const float ri = remap0(l_rev1) / remap0(ls1.pdf_fwd);
        sum += ri + ri * (remap0(l_rev2) / l_fwd2) * l_tail;
    }
    return 1.0f / (1.0f + sum);
}

}  // namespace prt
