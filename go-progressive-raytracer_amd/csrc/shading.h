// shading.h — float32 device versions of the reference's camera, BSDF and light routines.
// Each function cites the reference routine it follows (paths relative to the reference tree).
#pragma once
#include "prt_common.h"
#include "traverse.h"

namespace prt {

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ V3 v3(const float* p) { return {p[0], p[1], p[2]}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 mulv(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 normalize(V3 a) {  // Vec3.Normalize: zero stays zero
    float l = length(a);
    if (l == 0.0f) return {0, 0, 0};
    return {a.x / l, a.y / l, a.z / l};
}
__device__ __forceinline__ float lum(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

constexpr float kPiF = 3.14159265358979323846f;

// ---- warps, pkg/core/sampling.go ------------------------------------------------------
__device__ __forceinline__ V3 sample_cosine_hemisphere(V3 n, float s0, float s1) {  // :42-67
    float a = 2.0f * kPiF * s0;
    float r = sqrtf(s1);
    float sa, ca;
    sincosf(a, &sa, &ca);
    float x = r * ca, y = r * sa, zc = sqrtf(1.0f - s1);
    V3 nt = fabsf(n.x) > 0.1f ? v3(0, 1, 0) : v3(1, 0, 0);
    V3 t = normalize(cross(nt, n));
    V3 b = cross(n, t);
    return t * x + b * y + n * zc;
}
__device__ __forceinline__ V3 sample_point_in_unit_disk(float s0, float s1) {  // :108-126
    float ox = 2.0f * s0 - 1.0f, oy = 2.0f * s1 - 1.0f;
    if (ox == 0.0f && oy == 0.0f) return {0, 0, 0};
    float theta, r;
    if (fabsf(ox) > fabsf(oy)) {
        r = ox;
        theta = kPiF / 4.0f * (oy / ox);
    } else {
        r = oy;
        theta = kPiF / 2.0f - kPiF / 4.0f * (ox / oy);
    }
    float st, ct;
    sincosf(theta, &st, &ct);
    return {r * ct, r * st, 0};
}
__device__ __forceinline__ V3 sample_point_in_unit_sphere(float s0, float s1, float s2) {  // :130-147
    float r = cbrtf(s0);
    float phi = 2.0f * kPiF * s1;
    float ct = 2.0f * s2 - 1.0f;
    float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float sp, cp;
    sincosf(phi, &sp, &cp);
    return {r * st * cp, r * st * sp, r * ct};
}

// ---- camera, pkg/geometry/camera.go:173-190 --------------------------------------------
__device__ __forceinline__ void camera_ray(const DevCamera& c, int i, int j, float l0, float l1, float j0, float j1,
                                           V3& o, V3& d) {
    float fx = (float)i + (j0 - 0.5f), fy = (float)j + (j1 - 0.5f);
    V3 ps = v3(c.pixel00) + v3(c.du) * fx + v3(c.dv) * fy;
    o = v3(c.center);
    if (c.has_lens) {
        V3 p = sample_point_in_unit_disk(l0, l1);
        o = o + v3(c.disk_u) * p.x + v3(c.disk_v) * p.y;
    }
    d = normalize(ps - o);
}

// ---- surface interaction ----------------------------------------------------------------
struct Surf {
    V3 p, n;
    bool front;
    int mat;
    float u = 0.0f, v = 0.0f;  // SurfaceInteraction.UV, set when the scene has textures
};

// In-plane frame of the reference's cap and body UVs: t = normalize(n x ref), b = n x t with
// ref = (0,1,0) unless |n.y| >= 0.9 (cylinder.go:190-200, 243-252; cone.go:236-246, 329-338).
__device__ __forceinline__ void uv_frame(V3 n, V3& t, V3& b) {
    V3 ref = fabsf(n.y) < 0.9f ? v3(0, 1, 0) : v3(1, 0, 0);
    t = normalize(cross(n, ref));
    b = cross(n, t);
}

// Hit record of a primitive hit (SetFaceNormal, material/interfaces.go:58-65).
// The point ray.At(t) is moved back onto the primitive's surface (plane projection for triangles and
// quads, radial for spheres): in float32, o + t*d far from the origin lies up to |p|*2^-24 off the
// surface, enough for a grazing continuation ray to re-hit the surface beyond tmin = 0.001, which
// the float64 reference never does.
__device__ __forceinline__ Surf make_surf(const DevScene& sc, V3 o, V3 d, const THit& h) {
    const float4* pp = sc.prims + 3 * (size_t)h.prim;
    float4 a = pp[0], b = pp[1];
    uint32_t tag = __float_as_uint(a.w);
    uint32_t kind = prim_kind(tag);
    Surf s;
    s.p = o + d * h.t;
    s.mat = (int)__float_as_uint(b.w);
    V3 outward;
    if (kind == kPrimTriangle) {
        float4 c = pp[2];
        V3 v0 = v3(a.x, a.y, a.z), v1 = v3(b.x, b.y, b.z), v2 = v3(c.x, c.y, c.z);
        V3 gn = normalize(cross(v1 - v0, v2 - v0));  // triangle.go:92-96
        s.p = s.p - gn * dot(s.p - v0, gn);
        const uint32_t side = prim_side(tag);
        outward = side ? v3(sc.tri_normals[side].x, sc.tri_normals[side].y, sc.tri_normals[side].z) : gn;
    } else if (kind == kPrimSphere) {
        V3 c = v3(a.x, a.y, a.z);
        V3 r = s.p - c;
        float len = length(r);
        if (len > 0.0f) s.p = c + r * (fabsf(b.x) / len);
        outward = (s.p - c) * (1.0f / b.x);  // sphere.go:63, sign of r kept
    } else if (kind == kPrimQuad) {
        const QuadExt& q = sc.quads[prim_side(tag)];
        outward = v3(q.nx, q.ny, q.nz);
        s.p = s.p - outward * (dot(s.p, outward) - q.D);
    } else if (kind == kPrimDisc) {  // SetFaceNormal with the disc normal (disc.go:85)
        outward = v3(b.x, b.y, b.z);
        s.p = s.p - outward * dot(s.p - v3(a.x, a.y, a.z), outward);
    } else {  // cylinder or cone: h.u holds the part hit (0 body, 1 base cap, 2 top cap)
        const float4 c = pp[2];
        const V3 base = v3(a.x, a.y, a.z), axis = v3(b.x, b.y, b.z);
        const bool cone = kind == kPrimCone;
        const float H = cone ? c.z : c.y;
        const int part = (int)h.u;
        if (part == 0) {
            float hh = dot(s.p - base, axis);
            V3 centre = base + axis * hh;
            V3 radial = s.p - centre;
            float rl = length(radial);
            if (!cone) {  // radial normal (cylinder.go:181-184)
                outward = rl > 0.0f ? radial * (1.0f / rl) : radial;
                s.p = centre + outward * c.x;
            } else {  // normalize(radial + axis (rb - rt) / H) with the unnormalized radial (cone.go:221-229)
                float rs = c.x - hh * ((c.x - c.y) / H);  // the surface's radius at this height
                if (rl > 0.0f) radial = radial * (rs / rl);
                s.p = centre + radial;
                outward = normalize(radial + axis * ((c.x - c.y) / H));
            }
        } else {
            outward = part == 1 ? axis * -1.0f : axis;
            V3 cc = part == 1 ? base : base + axis * H;
            s.p = s.p - axis * dot(s.p - cc, axis);
        }
    }
    s.front = dot(d, outward) < 0.0f;
    s.n = s.front ? outward : outward * -1.0f;
    if (sc.textured) {  // hit UVs of each shape (the reference's Hit functions)
        float u = h.u, v = h.v;  // triangle barycentrics (triangle.go:151-161), quad (alpha, beta) (quad.go:165)
        if (kind == kPrimTriangle) {
            if (sc.tri_uv) {
                const float2* tu = sc.tri_uv + 3 * (size_t)h.prim;
                const float2 t0 = tu[0];
                if (t0.x == t0.x) {  // not NaN: the mesh has vertex UVs
                    const float2 t1 = tu[1], t2 = tu[2];
                    float w = 1.0f - h.u - h.v;
                    u = t0.x * w + t1.x * h.u + t2.x * h.v;
                    v = t0.y * w + t1.y * h.u + t2.y * h.v;
                }
            }
        } else if (kind == kPrimSphere) {  // sphere.go:64-68
            float ny = fmaxf(-1.0f, fminf(1.0f, outward.y));
            u = (atan2f(-outward.z, outward.x) + kPiF) / (2.0f * kPiF);
            v = acosf(-ny) / kPiF;
        } else if (kind == kPrimDisc) {  // disc.go:76-80 with NewDisc's frame
            V3 n = v3(b.x, b.y, b.z);
            V3 right = normalize(cross(fabsf(n.x) > 0.1f ? v3(0, 1, 0) : v3(1, 0, 0), n));
            V3 up = normalize(cross(n, right));
            V3 cp = s.p - v3(a.x, a.y, a.z);
            float r = pp[2].x;
            u = (dot(cp, right) / r + 1.0f) * 0.5f;
            v = (dot(cp, up) / r + 1.0f) * 0.5f;
        } else if (kind == kPrimCylinder || kind == kPrimCone) {
            const float4 c = pp[2];
            const V3 base = v3(a.x, a.y, a.z), axis = v3(b.x, b.y, b.z);
            const bool cone = kind == kPrimCone;
            const float H = cone ? c.z : c.y;
            const int part = (int)h.u;
            V3 t, bt;
            if (part == 0) {
                float hh = dot(s.p - base, axis);
                V3 radial = s.p - (base + axis * hh);
                uv_frame(axis, t, bt);
                u = (atan2f(dot(radial, bt), dot(radial, t)) + kPiF) / (2.0f * kPiF);
                v = hh / H;
            } else {
                V3 n = part == 1 ? axis * -1.0f : axis;
                V3 cc = part == 1 ? base : base + axis * H;
                float rad = cone ? (part == 1 ? c.x : c.y) : c.x;
                uv_frame(n, t, bt);
                V3 lp = s.p - cc;
                u = (dot(lp, t) / rad + 1.0f) * 0.5f;
                v = (dot(lp, bt) / rad + 1.0f) * 0.5f;
            }
        }
        s.u = u;
        s.v = v;
    }
    return s;
}

// ImageTexture.Evaluate (image_texture.go:24-55): wrap by truncation, flip v, nearest texel, clamp.
__device__ __forceinline__ V3 texture_eval(const DevTexture& t, float u, float v) {
    u = u - truncf(u);
    v = v - truncf(v);
    if (u < 0.0f) u += 1.0f;
    if (v < 0.0f) v += 1.0f;
    int x = (int)fminf(fmaxf(u * (float)t.width, 0.0f), (float)(t.width - 1));
    int y = (int)fminf(fmaxf((1.0f - v) * (float)t.height, 0.0f), (float)(t.height - 1));
    const float* p = t.rgb + 3 * ((size_t)y * t.width + x);
    return v3(p[0], p[1], p[2]);
}
// Albedo.Evaluate(hit.UV, hit.Point): the solid colour or the material's texture
__device__ __forceinline__ V3 albedo_of(const DevScene& sc, const DevMaterial& M, const Surf& s) {
    return M.tex > 0 ? texture_eval(sc.textures[M.tex - 1], s.u, s.v) : v3(M.albedo);
}

// ---- materials, pkg/material ------------------------------------------------------------
struct ScatterOut {
    V3 dir, att;
    float pdf;
};

__device__ __forceinline__ V3 reflect(V3 v, V3 n) { return v - n * (2.0f * dot(v, n)); }
__device__ __forceinline__ V3 refract(V3 uv, V3 n, float eta) {  // dielectric.go:155-161
    float ct = fminf(-dot(uv, n), 1.0f);
    V3 perp = (uv + n * ct) * eta;
    V3 par = n * (-sqrtf(fabsf(1.0f - dot(perp, perp))));
    return perp + par;
}
__device__ __forceinline__ float reflectance(float c, float ratio) {  // dielectric.go:164-172
    float r0 = (1.0f - ratio) / (1.0f + ratio);
    r0 = r0 * r0;
    float x = 1.0f - c;
    float x2 = x * x;
    return r0 + (1.0f - r0) * (x2 * x2 * x);
}
__device__ __forceinline__ bool is_reflection_path(V3 in, V3 out, V3 n) {  // layered.go:110-119
    V3 inc = in * -1.0f;
    V3 refl = inc - n * (2.0f * dot(inc, n));
    return length(out - refl) < 0.1f;
}

// Scatter of a non-layered material; draws from stream positions base+k.
__device__ __forceinline__ bool scatter_basic(const DevScene& sc, const DevMaterial& M, V3 din, const Surf& s,
                                              uint32_t key, uint32_t pos, uint32_t& used, ScatterOut& out) {
    switch (M.type) {
        case 0: {  // Lambertian, lambertian.go:25-50
            float u0 = draw(key, pos), u1 = draw(key, pos + 1);
            used = 2;
            V3 d = sample_cosine_hemisphere(s.n, u0, u1);
            float ct = dot(normalize(d), s.n);
            if (ct < 0.0f) ct = 0.0f;
            out.dir = d;
            out.pdf = ct / kPiF;
            out.att = albedo_of(sc, M, s) * (1.0f / kPiF);
            return true;
        }
        case 1: {  // Metal, metal.go:38-59
            V3 r = reflect(normalize(din), s.n);
            used = 0;
            if (M.fuzz > 0.0f) {
                float u0 = draw(key, pos), u1 = draw(key, pos + 1), u2 = draw(key, pos + 2);
                used = 3;
                r = r + sample_point_in_unit_sphere(u0, u1, u2) * M.fuzz;
            }
            out.dir = r;
            out.att = albedo_of(sc, M, s);
            out.pdf = 0.0f;
            return dot(r, s.n) > 0.0f;
        }
        case 2: {  // Dielectric, dielectric.go:20-60
            float ratio = s.front ? 1.0f / M.ior : M.ior;
            V3 ud = normalize(din);
            float ct = fminf(-dot(ud, s.n), 1.0f);
            float st = sqrtf(1.0f - ct * ct);
            bool cannot = ratio * st > 1.0f;
            used = 0;
            bool refl = cannot;
            if (!cannot) {
                used = 1;
                refl = reflectance(ct, ratio) > draw(key, pos);
            }
            out.dir = refl ? reflect(ud, s.n) : refract(ud, s.n, ratio);
            out.att = v3(1, 1, 1);
            out.pdf = 0.0f;
            return true;
        }
        default:  // Emissive never scatters (emissive.go:24-27)
            used = 0;
            return false;
    }
}

// Material.Scatter including Layered (layered.go:24-69) and Mix (mix.go:30-37, one draw picks
// material2 with probability ratio). Draws start at pos.
__device__ __forceinline__ bool scatter(const DevScene& sc, int m, V3 din, const Surf& s, uint32_t key, uint32_t pos,
                                        ScatterOut& out) {
    const DevMaterial& M = sc.mats[m];
    uint32_t used = 0;
    if (M.type == kMatMix) {
        const int pick = draw(key, pos) < M.ratio ? M.inner : M.outer;
        return scatter_basic(sc, sc.mats[pick], din, s, key, pos + 1, used, out);
    }
    if (M.type != 4) return scatter_basic(sc, M, din, s, key, pos, used, out);
    ScatterOut o;
    if (!scatter_basic(sc, sc.mats[M.outer], din, s, key, pos, used, o)) return false;
    V3 sd = normalize(o.dir);
    if (!(dot(sd, s.n) < 0.0f)) {
        out = o;
        return true;
    }
    ScatterOut i;
    uint32_t used2 = 0;
    if (!scatter_basic(sc, sc.mats[M.inner], sd, s, key, pos + used, used2, i)) {
        out = o;
        return true;
    }
    out.dir = i.dir;
    out.att = mulv(o.att, i.att);
    out.pdf = i.pdf;
    return true;
}

__device__ __forceinline__ V3 brdf_basic(const DevScene& sc, const DevMaterial& M, V3 in, V3 outd, const Surf& s,
                                        int mode) {
    switch (M.type) {
        case 0:  // lambertian.go:53-63
            if (dot(outd, s.n) <= 0.0f) return {0, 0, 0};
            return albedo_of(sc, M, s) * (1.0f / kPiF);
        case 1: {  // metal.go:62-74
            V3 r = reflect(in * -1.0f, s.n);
            if (length(outd - r) < 0.001f) return albedo_of(sc, M, s);
            return {0, 0, 0};
        }
        case 2: {  // dielectric.go:63-143
            V3 wi = normalize(in), wo = normalize(outd), n = normalize(s.n);
            float wiN = dot(wi, n), woN = dot(wo, n);
            bool same = (wiN > 0.0f) == (woN > 0.0f);
            float eta = s.front ? 1.0f / M.ior : M.ior;
            if (!same) {
                V3 r = reflect(wi, n);
                if (length(wo - r) < 0.001f) {
                    float f = reflectance(fabsf(wiN), eta);
                    return {f, f, f};
                }
            } else {
                V3 rr = refract(wi, n, eta);
                if (length(wo - rr) < 0.001f) {
                    float tr = 1.0f - reflectance(fabsf(wiN), eta);
                    if (mode == 0) {
                        float etap = s.front ? M.ior : 1.0f / M.ior;
                        tr /= etap * etap;
                    }
                    return {tr, tr, tr};
                }
            }
            return {0, 0, 0};
        }
        default:
            return {0, 0, 0};
    }
}
__device__ __forceinline__ V3 brdf(const DevScene& sc, int m, V3 in, V3 outd, const Surf& s, int mode) {
    const DevMaterial& M = sc.mats[m];
    if (M.type == kMatMix)  // mix.go:40-46
        return brdf_basic(sc, sc.mats[M.outer], in, outd, s, mode) * (1.0f - M.ratio) +
               brdf_basic(sc, sc.mats[M.inner], in, outd, s, mode) * M.ratio;
    if (M.type != 4) return brdf_basic(sc, M, in, outd, s, mode);
    if (is_reflection_path(in, outd, s.n))
        return brdf_basic(sc, sc.mats[M.outer], in, outd, s, mode);  // layered.go:73-87
    return brdf_basic(sc, sc.mats[M.inner], in, outd, s, mode);
}
__device__ __forceinline__ float pdf_basic(const DevMaterial& M, V3 outd, V3 n, bool& delta) {
    if (M.type == 0) {  // lambertian.go:66-72
        delta = false;
        float ct = dot(outd, n);
        return ct <= 0.0f ? 0.0f : ct / kPiF;
    }
    delta = (M.type == 1 || M.type == 2);  // metal/dielectric are delta; emissive is not
    return 0.0f;
}
__device__ __forceinline__ float mat_pdf(const DevScene& sc, int m, V3 in, V3 outd, V3 n, bool& delta) {
    const DevMaterial& M = sc.mats[m];
    if (M.type == kMatMix) {  // mix.go:49-56: never delta
        bool d1, d2;
        float p = pdf_basic(sc.mats[M.outer], outd, n, d1) * (1.0f - M.ratio) +
                  pdf_basic(sc.mats[M.inner], outd, n, d2) * M.ratio;
        delta = false;
        return p;
    }
    if (M.type != 4) return pdf_basic(M, outd, n, delta);
    bool d2;
    float p = is_reflection_path(in, outd, n) ? pdf_basic(sc.mats[M.outer], outd, n, d2)
                                              : pdf_basic(sc.mats[M.inner], outd, n, d2);
    delta = false;  // layered.go:90-104
    return p;
}

// ---- lights, pkg/lights ------------------------------------------------------------------
struct LSample {
    V3 dir, emission;
    float dist, pdf;
    V3 p, n;  // LightSample.Point / Normal (used by BDPT's sampled light vertex)
};

// Emitter.Emit for a ray of direction d without a hit record (emissive.go:30-37): the material's
// emission; the disc spot light's material emits only against its spot direction (d . spot <= -0.3,
// disc_spot_light.go:26-40). Zero for non-emitters.
__device__ __forceinline__ V3 emit_dir(const DevScene& sc, int m, V3 d) {
    const DevMaterial& M = sc.mats[m];
    if (M.type == 3) return v3(M.emission);
    if (M.type == kMatSpotEmitter && dot(d, v3(M.spot_dir)) <= -0.3f) return v3(M.emission);
    return v3(0, 0, 0);
}
// ... with a hit record: the back face emits nothing
__device__ __forceinline__ V3 emit_hit(const DevScene& sc, int m, V3 d, bool front) {
    return front ? emit_dir(sc, m, d) : v3(0, 0, 0);
}

// sphere test for light routines (sphere.go:27-81), returns t or -1
__device__ __forceinline__ float sphere_hit_t(V3 c, float rad, V3 o, V3 d, float tmin, float tmax) {
    float t0, t1;
    if (!sphere_roots(o.x - c.x, o.y - c.y, o.z - c.z, d.x, d.y, d.z, rad, t0, t1)) return -1.0f;
    float root = t0;
    if (root < tmin || root > tmax) {
        root = t1;
        if (root < tmin || root > tmax) return -1.0f;
    }
    return root;
}

__device__ __forceinline__ void env_lookup(const DevEnv& e, V3 d, int& x, int& y) {
    d = normalize(d);
    float cy = fmaxf(-1.0f, fminf(1.0f, d.y));
    float theta = acosf(cy);
    float phi = atan2f(d.z, d.x);
    if (phi < 0.0f) phi += 2.0f * kPiF;
    x = min((int)(phi / (2.0f * kPiF) * e.width), e.width - 1);
    y = min((int)(theta / kPiF * e.height), e.height - 1);
    x = max(x, 0);
    y = max(y, 0);
}
__device__ __forceinline__ V3 env_eval(const DevEnv& e, V3 d) {
    int x, y;
    env_lookup(e, d, x, y);
    const float* p = e.rgb + ((size_t)y * e.width + x) * 3;
    return v3(p[0], p[1], p[2]);
}
__device__ __forceinline__ float env_pdf(const DevEnv& e, V3 d) {
    if (!(e.integral > 0.0f)) return 0.0f;
    int x, y;
    env_lookup(e, d, x, y);
    float st = sinf(((float)y + 0.5f) / e.height * kPiF);
    if (st <= 0.0f) return 0.0f;
    return (e.func[(size_t)y * e.width + x] / e.integral) / (2.0f * kPiF * kPiF * st);
}
__device__ __forceinline__ int find_interval(const float* cdf, int n, float u) {
    int lo = 0, hi = n;
    while (lo + 1 < hi) {
        int mid = (lo + hi) >> 1;
        if (cdf[mid] <= u) lo = mid;
        else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ V3 env_sample(const DevEnv& e, float s0, float s1, float& pdf) {
    int y = find_interval(e.row_cdf, e.height, s0);
    float dy = e.row_cdf[y + 1] - e.row_cdf[y];
    float fy = dy > 0.0f ? (s0 - e.row_cdf[y]) / dy : 0.5f;
    const float* cc = e.col_cdf + (size_t)y * (e.width + 1);
    int x = find_interval(cc, e.width, s1);
    float dx = cc[x + 1] - cc[x];
    float fx = dx > 0.0f ? (s1 - cc[x]) / dx : 0.5f;
    float u = ((float)x + fx) / e.width, v = ((float)y + fy) / e.height;
    float theta = v * kPiF, phi = u * 2.0f * kPiF;
    float st, ct, sp, cp;
    sincosf(theta, &st, &ct);
    sincosf(phi, &sp, &cp);
    float sc = sinf(((float)y + 0.5f) / e.height * kPiF);
    pdf = (sc > 0.0f && e.integral > 0.0f) ? (e.func[(size_t)y * e.width + x] / e.integral) / (2.0f * kPiF * kPiF * sc)
                                           : 0.0f;
    return v3(st * cp, ct, st * sp);
}

__device__ __forceinline__ bool light_is_infinite(int type) {
    return type == PRT_LIGHT_UNIFORM_INFINITE || type == PRT_LIGHT_GRADIENT_INFINITE || type == PRT_LIGHT_ENVMAP;
}

// falloff, point_spot_light.go:121-139 (the disc spot light uses the same curve)
__device__ __forceinline__ float spot_falloff(const DevLight& L, float cos_a) {
    if (cos_a < L.cos_total) return 0.0f;
    if (cos_a >= L.cos_falloff) return 1.0f;
    float d = (cos_a - L.cos_total) / (L.cos_falloff - L.cos_total);
    return d * d * d * d;
}
__device__ __forceinline__ V3 sample_cone(V3 w, float cos_total, float s0, float s1) {  // core/sampling.go:70-94
    V3 u = fabsf(w.x) > 0.1f ? v3(0, 1, 0) : v3(1, 0, 0);
    u = normalize(cross(u, w));
    V3 v = cross(w, u);
    float ct = 1.0f - s0 * (1.0f - cos_total);
    float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float sp, cp;
    sincosf(2.0f * kPiF * s1, &sp, &cp);
    return u * (st * cp) + v * (st * sp) + w * ct;
}
__device__ __forceinline__ float uniform_cone_pdf(float cos_total) {  // light_sampling.go:83-85
    return 1.0f / (2.0f * kPiF * (1.0f - cos_total));
}
__device__ __forceinline__ void disc_sample_uniform(const DevLight& L, float s0, float s1, V3& p, V3& n) {
    float r = sqrtf(s0) * L.radius;  // SampleUniform, disc.go:121-134
    float sp, cp;
    sincosf(2.0f * kPiF * s1, &sp, &cp);
    p = v3(L.center) + v3(L.right) * (r * cp) + v3(L.up) * (r * sp);
    n = v3(L.n);
}
// the disc of a disc light re-hit from p along dir (DiscLight.PDF, disc_light.go:82-101); t or -1
__device__ __forceinline__ float disc_hit_t(const DevLight& L, V3 p, V3 dir) {
    V3 n = v3(L.n);
    float den = dot(n, dir);
    if (fabsf(den) < 1e-6f) return -1.0f;
    float t = dot(n, v3(L.center) - p) / den;
    if (t < 0.001f) return -1.0f;
    V3 cp = p + dir * t - v3(L.center);
    return dot(cp, cp) > L.radius * L.radius ? -1.0f : t;
}

// Light.Emit of an infinite light for a ray leaving the scene
__device__ __forceinline__ V3 light_emit(const DevScene& sc, const DevLight& L, V3 d) {
    if (L.type == PRT_LIGHT_UNIFORM_INFINITE) return v3(L.emission);  // uniform_infinite_light.go:124-126
    if (L.type == PRT_LIGHT_GRADIENT_INFINITE) {                       // gradient_infinite_light.go:133-137
        V3 n = normalize(d);
        float t = 0.5f * (n.y + 1.0f);
        return v3(L.bottom) * (1.0f - t) + v3(L.top) * t;
    }
    if (L.type == PRT_LIGHT_ENVMAP) return env_eval(sc.env, d);
    return {0, 0, 0};
}

__device__ __forceinline__ LSample sphere_sample_uniform(const DevScene& sc, const DevLight& L, V3 p, float s0, float s1) {
    float z = 1.0f - 2.0f * s0;  // sphere_light.go:52-77
    float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float sp, cp;
    sincosf(2.0f * kPiF * s1, &sp, &cp);
    V3 ld = v3(r * cp, r * sp, z);
    V3 pt = v3(L.center) + ld * L.radius;
    V3 dir = pt - p;
    LSample ls;
    ls.dist = length(dir);
    ls.dir = normalize(dir);
    ls.pdf = 1.0f / (4.0f * kPiF * L.radius * L.radius);
    ls.emission = emit_dir(sc, L.mat, ls.dir);
    ls.p = pt;
    ls.n = ld;
    return ls;
}

// Light.Sample (interfaces.go:22-25) for every supported light type
__device__ __forceinline__ LSample light_sample(const DevScene& sc, const DevLight& L, V3 p, V3 n, float s0, float s1) {
    LSample ls;
    if (L.type == PRT_LIGHT_QUAD) {  // quad_light.go:36-84
        V3 spt = v3(L.corner) + v3(L.u) * s0 + v3(L.v) * s1;
        V3 tl = spt - p;
        float dist = length(tl);
        V3 dir = tl * (1.0f / dist);
        V3 qn = v3(L.n);
        float ct = fabsf(dot(qn, dir * -1.0f));
        ls.dir = dir;
        ls.dist = dist;
        ls.p = spt;
        ls.n = qn;
        if (ct < 1e-8f) {
            ls.emission = {0, 0, 0};
            ls.pdf = 0.0f;
            return ls;
        }
        ls.pdf = (1.0f / L.area) * dist * dist / ct;
        ls.emission = dot(dir, qn) < 0.0f ? emit_dir(sc, L.mat, dir) : v3(0, 0, 0);
        return ls;
    }
    if (L.type == PRT_LIGHT_SPHERE) {  // sphere_light.go:37-141
        V3 c = v3(L.center);
        V3 tc = c - p;
        float dc = length(tc);
        if (dc <= L.radius) return sphere_sample_uniform(sc, L, p, s0, s1);
        V3 w = normalize(tc);
        V3 u = fabsf(w.x) > 0.1f ? v3(0, 1, 0) : v3(1, 0, 0);
        u = normalize(cross(u, w));
        V3 v = cross(w, u);
        float stm = L.radius / dc;
        float ctm = sqrtf(fmaxf(0.0f, 1.0f - stm * stm));
        float ct = 1.0f - s0 * (1.0f - ctm);
        float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
        float sp, cp;
        sincosf(2.0f * kPiF * s1, &sp, &cp);
        V3 dir = u * (st * cp) + v * (st * sp) + w * ct;
        float t = sphere_hit_t(c, L.radius, p, dir, 0.001f, __int_as_float(0x7f800000));
        if (t < 0.0f) return sphere_sample_uniform(sc, L, p, s0, s1);
        ls.dir = dir;
        ls.dist = t;
        ls.pdf = 1.0f / (2.0f * kPiF * (1.0f - ctm));
        ls.emission = emit_dir(sc, L.mat, dir);
        ls.p = p + dir * t;
        V3 outward = (ls.p - c) * (1.0f / L.radius);
        ls.n = dot(dir, outward) < 0.0f ? outward : outward * -1.0f;  // SetFaceNormal of the re-hit
        return ls;
    }
    if (L.type == PRT_LIGHT_DISC || L.type == PRT_LIGHT_DISC_SPOT) {
        // DiscLight.Sample (disc_light.go:31-79: no front-face test), DiscSpotLight.Sample (:105-116)
        V3 spt, sn;
        disc_sample_uniform(L, s0, s1, spt, sn);
        V3 d = spt - p;
        float dist = length(d);
        ls.p = spt;
        ls.n = sn;
        if (dist == 0.0f) {
            ls.dir = v3(0, 1, 0);
            ls.dist = 0.0f;
            ls.emission = v3(0, 0, 0);
            ls.pdf = 1.0f;
            return ls;
        }
        V3 dn = normalize(d);
        float ct = fabsf(dot(sn, dn * -1.0f));
        ls.pdf = ct < 1e-6f ? 0.0f : (1.0f / (kPiF * L.radius * L.radius)) * dist * dist / ct;
        ls.dir = dn;
        ls.dist = dist;
        ls.emission = emit_dir(sc, L.mat, dn);
        if (L.type == PRT_LIGHT_DISC_SPOT) ls.emission = ls.emission * spot_falloff(L, dot(v3(L.spot_dir), normalize(p - spt)));
        return ls;
    }
    if (L.type == PRT_LIGHT_POINT_SPOT) {  // PointSpotLight.Sample, point_spot_light.go:47-92 (pdf 1)
        V3 tl = v3(L.from) - p;
        float dist = length(tl);
        ls.p = v3(L.from);
        ls.pdf = 1.0f;
        if (dist == 0.0f) {
            ls.n = v3(0, 1, 0);
            ls.dir = v3(0, 1, 0);
            ls.dist = 0.0f;
            ls.emission = v3(0, 0, 0);
            return ls;
        }
        V3 to_light = normalize(tl);
        float att = spot_falloff(L, dot(v3(L.spot_dir), to_light * -1.0f));
        ls.emission = v3(L.emission) * (att / (dist * dist));
        ls.n = to_light * -1.0f;
        ls.dir = to_light;
        ls.dist = dist;
        return ls;
    }
    if (L.type == PRT_LIGHT_ENVMAP) {
        float pdf;
        V3 dir = env_sample(sc.env, s0, s1, pdf);
        ls.dir = dir;
        ls.dist = __int_as_float(0x7f800000);
        ls.pdf = pdf;
        ls.emission = env_eval(sc.env, dir);
        ls.p = p + dir * 1e10f;
        ls.n = dir * -1.0f;
        return ls;
    }
    // uniform / gradient infinite: cosine hemisphere about the shading normal
    V3 dir = sample_cosine_hemisphere(n, s0, s1);
    ls.dir = dir;
    ls.dist = __int_as_float(0x7f800000);
    ls.pdf = dot(dir, n) / kPiF;
    ls.p = p + dir * 1e10f;
    ls.n = dir * -1.0f;
    if (L.type == PRT_LIGHT_UNIFORM_INFINITE) {
        ls.emission = v3(L.emission);
    } else {
        float t = 0.5f * (dir.y + 1.0f);
        ls.emission = v3(L.bottom) * (1.0f - t) + v3(L.top) * t;
    }
    return ls;
}

// Light.PDF (interfaces.go:27-29)
__device__ __forceinline__ float light_pdf(const DevScene& sc, const DevLight& L, V3 p, V3 n, V3 dir) {
    if (L.type == PRT_LIGHT_QUAD) {  // quad_light.go:87-104: re-hit the quad from p
        V3 qn = v3(L.n);
        float den = dot(dir, qn);
        if (fabsf(den) < 1e-8f) return 0.0f;
        float t = (L.D - dot(p, qn)) / den;
        if (t < 0.001f) return 0.0f;
        V3 hv = (p + dir * t) - v3(L.corner);
        float alpha = dot(v3(L.W), cross(hv, v3(L.v)));
        float beta = dot(v3(L.W), cross(v3(L.u), hv));
        if (alpha < 0.0f || alpha > 1.0f || beta < 0.0f || beta > 1.0f) return 0.0f;
        float ct = fabsf(dot(qn, dir * -1.0f));
        if (ct < 1e-8f) return 0.0f;
        return (1.0f / L.area) * t * t / ct;
    }
    if (L.type == PRT_LIGHT_SPHERE) {  // sphere_light.go:144-165
        V3 c = v3(L.center);
        if (sphere_hit_t(c, L.radius, p, dir, 0.001f, __int_as_float(0x7f800000)) < 0.0f) return 0.0f;
        float dc = length(c - p);
        if (dc <= L.radius) return 1.0f / (4.0f * kPiF * L.radius * L.radius);
        float stm = L.radius / dc;
        float ctm = sqrtf(fmaxf(0.0f, 1.0f - stm * stm));
        return 1.0f / (2.0f * kPiF * (1.0f - ctm));
    }
    if (L.type == PRT_LIGHT_DISC || L.type == PRT_LIGHT_DISC_SPOT) {  // disc_light.go:82-101
        float t = disc_hit_t(L, p, dir);
        if (t < 0.0f) return 0.0f;
        float ct = fabsf(dot(v3(L.n), dir * -1.0f));
        if (ct < 1e-6f) return 0.0f;
        return (1.0f / (kPiF * L.radius * L.radius)) * t * t / ct;
    }
    if (L.type == PRT_LIGHT_POINT_SPOT) {  // point_spot_light.go:96-118: 1 within 0.999 of the light
        V3 tl = v3(L.from) - p;
        if (dot(tl, tl) == 0.0f) return 0.0f;
        return dot(dir, normalize(tl)) > 0.999f ? 1.0f : 0.0f;
    }
    if (L.type == PRT_LIGHT_ENVMAP) return env_pdf(sc.env, dir);
    float ct = dot(dir, n);  // uniform/gradient: cosine hemisphere
    return ct <= 0.0f ? 0.0f : ct / kPiF;
}

// WeightedLightSampler.SampleLight: CDF walk with u <= cumulative (weighted_light_sampler.go:82-96)
__device__ __forceinline__ int select_light(const DevScene& sc, float u, float& prob) {
    float cum = 0.0f;
    for (int i = 0; i < sc.num_lights; i++) {
        cum += sc.light_weights[i];
        if (u <= cum) {
            prob = sc.light_weights[i];
            return i;
        }
    }
    prob = sc.light_weights[sc.num_lights - 1];
    return sc.num_lights - 1;
}

// CalculateLightPDF (light_sampling.go:11-25)
__device__ __forceinline__ float total_light_pdf(const DevScene& sc, V3 p, V3 n, V3 dir) {
    float tot = 0.0f;
    for (int i = 0; i < sc.num_lights; i++) tot += light_pdf(sc, sc.lights[i], p, n, dir) * sc.light_weights[i];
    return tot;
}

__device__ __forceinline__ float power_heuristic(float f, float g) {  // path_tracing.go:224-234
    if (f == 0.0f) return 0.0f;
    return (f * f) / (f * f + g * g);
}

}  // namespace prt
