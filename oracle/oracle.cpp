// oracle.cpp — float64 CPU restatement of the reference path-tracing hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP library and the timed
// CPU baseline ("port"); nothing in the product links, loads or calls it. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
//
// Every routine follows the reference function cited beside it (paths relative to the
// df07/go-progressive-raytracer tree): same branch structure, epsilons and estimator
// quirks, float64 throughout (pkg/core/vec3.go:10). Randomness comes from the counter
// stream of oracle_rng.h instead of Go's math/rand (which cannot be run here).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../include/prt.h"
#include "oracle_rng.h"

namespace orc {

static const double kPi = 3.14159265358979323846;
static const double kInf = INFINITY;

// ---- pkg/core/vec3.go -----------------------------------------------------------
struct Vec3 {
    double x = 0, y = 0, z = 0;
    Vec3() {}
    Vec3(double a, double b, double c) : x(a), y(b), z(c) {}
    Vec3 operator+(const Vec3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vec3 operator-(const Vec3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    Vec3 operator*(double s) const { return {x * s, y * s, z * s}; }
    Vec3 mul(const Vec3& o) const { return {x * o.x, y * o.y, z * o.z}; }  // MultiplyVec :113
    double dot(const Vec3& o) const { return x * o.x + y * o.y + z * o.z; }
    Vec3 cross(const Vec3& o) const {                                        // Cross :103
        return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x};
    }
    double length() const { return std::sqrt(x * x + y * y + z * z); }
    double length2() const { return x * x + y * y + z * z; }
    Vec3 normalize() const {  // Normalize :93-99 (zero stays zero)
        double l = length();
        if (l == 0) return {0, 0, 0};
        return {x / l, y / l, z / l};
    }
    Vec3 neg() const { return {-x, -y, -z}; }
    double luminance() const { return 0.2126 * x + 0.7152 * y + 0.0722 * z; }  // :125
    bool is_zero() const { return x == 0 && y == 0 && z == 0; }
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
static Vec3 V(const double* p) { return {p[0], p[1], p[2]}; }
struct Vec2 {
    double x = 0, y = 0;
};
struct Ray {
    Vec3 o, d;
    Vec3 at(double t) const { return o + d * t; }  // Ray.At vec3.go:216
};

// Rotate X->Y->Z, pkg/core/vec3.go:162-198
static Vec3 rotate(Vec3 v, Vec3 r) {
    if (r.x == 0 && r.y == 0 && r.z == 0) return v;
    Vec3 res = v;
    if (r.x != 0) {
        double c = std::cos(r.x), s = std::sin(r.x);
        res = Vec3(res.x, res.y * c - res.z * s, res.y * s + res.z * c);
    }
    if (r.y != 0) {
        double c = std::cos(r.y), s = std::sin(r.y);
        res = Vec3(res.x * c + res.z * s, res.y, -res.x * s + res.z * c);
    }
    if (r.z != 0) {
        double c = std::cos(r.z), s = std::sin(r.z);
        res = Vec3(res.x * c - res.y * s, res.x * s + res.y * c, res.z);
    }
    return res;
}

// ---- Sampler over the counter stream (interface of pkg/core/sampling.go:10-14) ----
struct Sampler {
    uint32_t key = 0;
    uint32_t slot = 0;
    uint32_t k = 0;
    void set_slot(uint32_t s) { slot = s; k = 0; }
    double get1d() { return draw(key, slot * 16u + (k++)); }
    Vec2 get2d() { Vec2 r; r.x = get1d(); r.y = get1d(); return r; }
    Vec3 get3d() { Vec3 r; r.x = get1d(); r.y = get1d(); r.z = get1d(); return r; }
};

// ---- warps, pkg/core/sampling.go -----------------------------------------------
static Vec3 sample_cosine_hemisphere(Vec3 n, Vec2 s) {  // :42-67
    double a = 2.0 * kPi * s.x;
    double z = s.y;
    double r = std::sqrt(z);
    double x = r * std::cos(a);
    double y = r * std::sin(a);
    double zc = std::sqrt(1.0 - z);
    Vec3 nt = std::fabs(n.x) > 0.1 ? Vec3(0, 1, 0) : Vec3(1, 0, 0);
    Vec3 t = nt.cross(n).normalize();
    Vec3 b = n.cross(t);
    return t * x + b * y + n * zc;
}
static Vec3 sample_on_unit_sphere(Vec2 s) {  // :97-104
    double z = 1.0 - 2.0 * s.x;
    double r = std::sqrt(std::max(0.0, 1.0 - z * z));
    double phi = 2.0 * kPi * s.y;
    return {r * std::cos(phi), r * std::sin(phi), z};
}
static Vec3 sample_point_in_unit_disk(Vec2 s) {  // :108-126
    double ox = 2 * s.x - 1, oy = 2 * s.y - 1;
    if (ox == 0 && oy == 0) return {0, 0, 0};
    double theta, r;
    if (std::fabs(ox) > std::fabs(oy)) {
        r = ox;
        theta = kPi / 4 * (oy / ox);
    } else {
        r = oy;
        theta = kPi / 2 - kPi / 4 * (ox / oy);
    }
    return {r * std::cos(theta), r * std::sin(theta), 0};
}
static Vec3 sample_point_in_unit_sphere(Vec3 s) {  // :130-147
    double r = std::pow(s.x, 1.0 / 3.0);
    double phi = 2 * kPi * s.y;
    double ct = 2 * s.z - 1;
    double st = std::sqrt(1 - ct * ct);
    return {r * st * std::cos(phi), r * st * std::sin(phi), r * ct};
}

// ---- AABB, pkg/geometry/aabb.go -------------------------------------------------
struct AABB {
    Vec3 mn, mx;
    bool hit(const Ray& ray, double tmin, double tmax) const {  // :43-95
        for (int a = 0; a < 3; a++) {
            double lo = mn[a], hi = mx[a], o = ray.o[a], d = ray.d[a];
            if (std::fabs(d) < 1e-8) {
                if (o < lo || o > hi) return false;
                continue;
            }
            double inv = 1.0 / d;
            double t1 = (lo - o) * inv, t2 = (hi - o) * inv;
            if (t1 > t2) std::swap(t1, t2);
            tmin = std::max(tmin, t1);
            tmax = std::min(tmax, t2);
            if (tmin > tmax) return false;
        }
        return true;
    }
    AABB uni(const AABB& o) const {  // Union :98-111
        return {{std::min(mn.x, o.mn.x), std::min(mn.y, o.mn.y), std::min(mn.z, o.mn.z)},
                {std::max(mx.x, o.mx.x), std::max(mx.y, o.mx.y), std::max(mx.z, o.mx.z)}};
    }
    Vec3 center() const { return (mn + mx) * 0.5; }
    int longest_axis() const {  // :126-135
        Vec3 s = mx - mn;
        if (s.x > s.y && s.x > s.z) return 0;
        if (s.y > s.z) return 1;
        return 2;
    }
};
static AABB aabb_from_points(const Vec3* p, int n) {  // NewAABBFromPoints :23-40
    AABB b{p[0], p[0]};
    for (int i = 1; i < n; i++) {
        b.mn = {std::min(b.mn.x, p[i].x), std::min(b.mn.y, p[i].y), std::min(b.mn.z, p[i].z)};
        b.mx = {std::max(b.mx.x, p[i].x), std::max(b.mx.y, p[i].y), std::max(b.mx.z, p[i].z)};
    }
    return b;
}

// ---- hit record, pkg/material/interfaces.go:48-65 -------------------------------
struct Hit {
    Vec3 p, n;
    double t = 0;
    bool front = false;
    int mat = -1;
    Vec2 uv;
    int64_t prim = -1;
    void set_face_normal(const Ray& r, Vec3 outward) {
        front = r.d.dot(outward) < 0;
        n = front ? outward : outward * -1.0;
    }
};

struct Counters {
    uint64_t box = 0, prim = 0;
};

struct Shape {
    virtual ~Shape() {}
    virtual bool hit(const Ray& r, double tmin, double tmax, Hit& h, Counters* c) const = 0;
    virtual AABB bbox() const = 0;
};

// ---- Triangle, pkg/geometry/triangle.go:92-175 ------------------------------------
struct Triangle : Shape {
    Vec3 v0, v1, v2, normal;
    AABB box;
    int mat;
    int64_t id;
    bool has_uv = false;  // NewTriangleWithUVs (triangle.go:51-69)
    Vec2 uv0, uv1, uv2;
    Triangle(Vec3 a, Vec3 b, Vec3 c, int m, int64_t i) : v0(a), v1(b), v2(c), mat(m), id(i) {
        normal = (v1 - v0).cross(v2 - v0).normalize();  // computeNormal :92-96
        Vec3 p[3] = {v0, v1, v2};
        box = aabb_from_points(p, 3);
    }
    bool hit(const Ray& r, double tmin, double tmax, Hit& h, Counters* c) const override {
        if (c) c->prim++;
        const double eps = 1e-8;
        Vec3 e1 = v1 - v0, e2 = v2 - v0;
        Vec3 hv = r.d.cross(e2);
        double a = e1.dot(hv);
        if (a > -eps && a < eps) return false;
        double f = 1.0 / a;
        Vec3 s = r.o - v0;
        double u = f * s.dot(hv);
        if (u < 0.0 || u > 1.0) return false;
        Vec3 q = s.cross(e1);
        double v = f * r.d.dot(q);
        if (v < 0.0 || u + v > 1.0) return false;
        double t = f * e2.dot(q);
        if (t < tmin || t > tmax) return false;
        h.t = t;
        h.p = r.at(t);
        h.mat = mat;
        if (has_uv) {  // barycentric interpolation of the vertex UVs, :151-161
            double w = 1.0 - u - v;
            h.uv = {uv0.x * w + uv1.x * u + uv2.x * v, uv0.y * w + uv1.y * u + uv2.y * v};
        } else {
            h.uv = {u, v};
        }
        h.prim = id;
        h.set_face_normal(r, normal);
        return true;
    }
    AABB bbox() const override { return box; }
};

// ---- Sphere, pkg/geometry/sphere.go:27-90 -----------------------------------------
struct Sphere : Shape {
    Vec3 c;
    double r;
    int mat;
    int64_t id;
    Sphere(Vec3 cc, double rr, int m, int64_t i) : c(cc), r(rr), mat(m), id(i) {}
    bool hit(const Ray& ray, double tmin, double tmax, Hit& h, Counters* cnt) const override {
        if (cnt) cnt->prim++;
        Vec3 oc = ray.o - c;
        double a = ray.d.dot(ray.d);
        double hb = oc.dot(ray.d);
        double cc = oc.dot(oc) - r * r;
        double disc = hb * hb - a * cc;
        if (disc < 0) return false;
        double sq = std::sqrt(disc);
        double root = (-hb - sq) / a;
        if (root < tmin || root > tmax) {
            root = (-hb + sq) / a;
            if (root < tmin || root > tmax) return false;
        }
        h.t = root;
        h.p = ray.at(root);
        Vec3 outward = (h.p - c) * (1.0 / r);
        double theta = std::acos(-outward.y);
        double phi = std::atan2(-outward.z, outward.x) + kPi;
        h.uv = {phi / (2.0 * kPi), theta / kPi};
        h.mat = mat;
        h.prim = id;
        h.set_face_normal(ray, outward);
        return true;
    }
    AABB bbox() const override {  // :83-90 (a negative radius yields an inverted box, as there)
        Vec3 rv(r, r, r);
        return {c - rv, c + rv};
    }
};

// ---- Quad, pkg/geometry/quad.go ----------------------------------------------------
struct Quad : Shape {
    Vec3 corner, u, v, normal, w;
    double D;
    int mat;
    int64_t id;
    Quad(Vec3 c, Vec3 uu, Vec3 vv, int m, int64_t i) : corner(c), u(uu), v(vv), mat(m), id(i) {
        normal = u.cross(v).normalize();  // NewQuad :98-121
        D = normal.dot(corner);
        Vec3 cr = u.cross(v);
        w = normal * (1.0 / normal.dot(cr));
    }
    bool hit(const Ray& r, double tmin, double tmax, Hit& h, Counters* c) const override {
        if (c) c->prim++;
        double den = r.d.dot(normal);  // :133-180
        if (std::fabs(den) < 1e-8) return false;
        double t = (D - r.o.dot(normal)) / den;
        if (t < tmin || t > tmax) return false;
        Vec3 p = r.at(t);
        Vec3 hv = p - corner;
        double alpha = w.dot(hv.cross(v));
        double beta = w.dot(u.cross(hv));
        if (alpha < 0 || alpha > 1 || beta < 0 || beta > 1) return false;
        h.t = t;
        h.p = p;
        h.uv = {alpha, beta};
        h.mat = mat;
        h.prim = id;
        h.set_face_normal(r, normal);
        return true;
    }
    AABB bbox() const override {  // BoundingBox :183-210 with the +-0.001 slab (:44-77)
        Vec3 cs[4] = {corner, corner + u, corner + v, corner + u + v};
        const double thr = 0.9999, tol = 0.0001, e = 0.001;
        Vec3 n = normal;
        int al = -1;
        if (std::fabs(n.x) > thr && std::fabs(n.y) < tol && std::fabs(n.z) < tol) al = 0;
        else if (std::fabs(n.y) > thr && std::fabs(n.x) < tol && std::fabs(n.z) < tol) al = 1;
        else if (std::fabs(n.z) > thr && std::fabs(n.x) < tol && std::fabs(n.y) < tol) al = 2;
        AABB b = aabb_from_points(cs, 4);
        if (al >= 0) {
            double fixed = cs[0][al];
            double lo[3] = {b.mn.x, b.mn.y, b.mn.z}, hi[3] = {b.mx.x, b.mx.y, b.mx.z};
            lo[al] = fixed - e;
            hi[al] = fixed + e;
            b.mn = {lo[0], lo[1], lo[2]};
            b.mx = {hi[0], hi[1], hi[2]};
        }
        return b;
    }
};

// ---- Disc, pkg/geometry/disc.go:11-134 ---------------------------------------------
struct Disc : Shape {
    Vec3 c, n, right, up;
    double r;
    int mat;
    int64_t id;
    Disc(Vec3 center, Vec3 normal, double radius, int m, int64_t i) : c(center), r(radius), mat(m), id(i) {
        n = normal.normalize();  // NewDisc :20-43
        right = std::fabs(n.x) > 0.1 ? Vec3(0, 1, 0) : Vec3(1, 0, 0);
        right = right.cross(n).normalize();
        up = n.cross(right).normalize();
    }
    bool hit(const Ray& ray, double tmin, double tmax, Hit& h, Counters* cnt) const override {  // :46-88
        if (cnt) cnt->prim++;
        double den = n.dot(ray.d);
        if (std::fabs(den) < 1e-6) return false;
        double t = n.dot(c - ray.o) / den;
        if (t < tmin || t > tmax) return false;
        Vec3 p = ray.at(t);
        Vec3 cp = p - c;
        if (cp.length2() > r * r) return false;
        h.t = t;
        h.p = p;
        h.uv = {(cp.dot(right) / r + 1.0) / 2.0, (cp.dot(up) / r + 1.0) / 2.0};
        h.mat = mat;
        h.prim = id;
        h.set_face_normal(ray, n);
        return true;
    }
    AABB bbox() const override {  // BoundingBox :91-118
        Vec3 re = right * r, ue = up * r;
        Vec3 cs[4] = {c + re + ue, c + re - ue, c - re + ue, c - re - ue};
        return aabb_from_points(cs, 4);
    }
    void sample_uniform(Vec2 s, Vec3& p, Vec3& nn) const {  // SampleUniform :121-134
        double rr = std::sqrt(s.x) * r;
        double th = 2.0 * kPi * s.y;
        p = c + right * (rr * std::cos(th)) + up * (rr * std::sin(th));
        nn = n;
    }
};

// Cap disc of a cylinder or cone (hitCap, cylinder.go:213-262 / cone.go:299-349): plane hit with
// |d.n| >= 1e-8, t in [tmin, tmax], |p - c| <= radius; uv from the (axis x ref) frame.
static bool hit_cap(const Ray& ray, Vec3 c, Vec3 n, double radius, double tmin, double tmax, int mat, int64_t id,
                    Hit& h) {
    double den = ray.d.dot(n);
    if (std::fabs(den) < 1e-8) return false;
    double t = (c - ray.o).dot(n) / den;
    if (t < tmin || t > tmax) return false;
    Vec3 p = ray.at(t);
    Vec3 lp = p - c;
    if (lp.length() > radius) return false;
    Vec3 ref = std::fabs(n.y) < 0.9 ? Vec3(0, 1, 0) : Vec3(1, 0, 0);
    Vec3 tg = n.cross(ref).normalize();
    Vec3 bt = n.cross(tg);
    h.t = t;
    h.p = p;
    h.uv = {(lp.dot(tg) / radius + 1.0) / 2.0, (lp.dot(bt) / radius + 1.0) / 2.0};
    h.mat = mat;
    h.prim = id;
    h.set_face_normal(ray, n);
    return true;
}
// u of a body hit: the angle of the radial vector in the (axis x ref, axis x tangent) frame
static double body_u(Vec3 axis, Vec3 radial) {
    Vec3 ref = std::fabs(axis.y) < 0.9 ? Vec3(0, 1, 0) : Vec3(1, 0, 0);
    Vec3 tg = axis.cross(ref).normalize();
    Vec3 bt = axis.cross(tg);
    return (std::atan2(radial.dot(bt), radial.dot(tg)) + kPi) / (2.0 * kPi);
}
// Bounding box of the base-top segment widened by `ext` on every axis the unit axis is not
// parallel to (|axis_i| > 0.9999): Cylinder/Cone.BoundingBox (cylinder.go:41-88, cone.go:83-137).
static AABB axis_box(Vec3 b, Vec3 t, Vec3 axis, double ext) {
    Vec3 mn(std::min(b.x, t.x), std::min(b.y, t.y), std::min(b.z, t.z));
    Vec3 mx(std::max(b.x, t.x), std::max(b.y, t.y), std::max(b.z, t.z));
    double ex = std::fabs(axis.x) > 0.9999 ? 0 : ext, ey = std::fabs(axis.y) > 0.9999 ? 0 : ext,
           ez = std::fabs(axis.z) > 0.9999 ? 0 : ext;
    return AABB{{mn.x - ex, mn.y - ey, mn.z - ez}, {mx.x + ex, mx.y + ey, mx.z + ez}};
}

// ---- Cylinder, pkg/geometry/cylinder.go:11-262 --------------------------------------
struct Cylinder : Shape {
    Vec3 base, top, axis;
    double r, height;
    bool capped;
    int mat;
    int64_t id;
    Cylinder(Vec3 b, Vec3 t, double radius, bool cap, int m, int64_t i)
        : base(b), top(t), r(radius), capped(cap), mat(m), id(i) {
        height = (t - b).length();
        axis = (t - b).normalize();
    }
    // hitBody :122-210: the first root (near, then far) with t in range and 0 <= h <= height
    bool body(const Ray& ray, double tmin, double tmax, Hit& h) const {
        Vec3 dl = ray.o - base;
        double dv = ray.d.dot(axis), lv = dl.dot(axis);
        double a = ray.d.length2() - dv * dv;
        double b = 2.0 * (dl.dot(ray.d) - lv * dv);
        double c = dl.length2() - lv * lv - r * r;
        if (std::fabs(a) < 1e-8) return false;
        double disc = b * b - 4 * a * c;
        if (disc < 0) return false;
        double sq = std::sqrt(disc);
        double roots[2] = {(-b - sq) / (2 * a), (-b + sq) / (2 * a)};
        for (double t : roots) {
            if (t < tmin || t > tmax) continue;
            Vec3 p = ray.at(t);
            double hh = (p - base).dot(axis);
            if (hh < 0 || hh > height) continue;
            Vec3 radial = p - (base + axis * hh);
            h.t = t;
            h.p = p;
            h.uv = {body_u(axis, radial), hh / height};
            h.mat = mat;
            h.prim = id;
            h.set_face_normal(ray, radial.normalize());
            return true;
        }
        return false;
    }
    bool hit(const Ray& ray, double tmin, double tmax, Hit& h, Counters* cnt) const override {  // :91-119
        if (cnt) cnt->prim++;
        bool any = false;
        double closest = tmax;
        Hit tmp;
        if (body(ray, tmin, closest, tmp)) {
            h = tmp;
            closest = tmp.t;
            any = true;
        }
        if (capped) {
            if (hit_cap(ray, base, axis * -1.0, r, tmin, closest, mat, id, tmp)) {
                h = tmp;
                closest = tmp.t;
                any = true;
            }
            if (hit_cap(ray, top, axis, r, tmin, closest, mat, id, tmp)) {
                h = tmp;
                any = true;
            }
        }
        return any;
    }
    AABB bbox() const override { return axis_box(base, top, axis, r); }
};

// ---- Cone, pkg/geometry/cone.go:11-349 ----------------------------------------------
struct Cone : Shape {
    Vec3 base, top, axis, apex;
    double rb, rt, height, tan_a;
    bool capped;
    int mat;
    int64_t id;
    Cone(Vec3 b, double base_r, Vec3 t, double top_r, bool cap, int m, int64_t i)
        : base(b), top(t), rb(base_r), rt(top_r), capped(cap), mat(m), id(i) {  // NewCone :27-80
        height = (t - b).length();
        axis = (t - b).normalize();
        tan_a = (rb - rt) / height;
        apex = rt == 0 ? t : t + axis * (rt * height / (rb - rt));
    }
    bool valid(const Ray& ray, double t, double tmin, double tmax) const {  // validateIntersection :265-296
        if (t < tmin || t > tmax) return false;
        Vec3 p = ray.at(t);
        double hh = (p - base).dot(axis);
        if (hh < -1e-8 || hh > height + 1e-8) return false;
        return (p - apex).dot(axis) <= 1e-8;
    }
    // hitBody :167-262: tries (-b - sqrt D) / 2a first, then (-b + sqrt D) / 2a, whatever the sign of a
    bool body(const Ray& ray, double tmin, double tmax, Hit& h) const {
        Vec3 co = ray.o - apex;
        double dv = ray.d.dot(axis), cv = co.dot(axis), k = tan_a * tan_a;
        double a = ray.d.length2() - (1 + k) * dv * dv;
        double b = 2.0 * (ray.d.dot(co) - (1 + k) * dv * cv);
        double c = co.length2() - (1 + k) * cv * cv;
        if (std::fabs(a) < 1e-8) return false;
        double disc = b * b - 4 * a * c;
        if (disc < 0) return false;
        double sq = std::sqrt(disc);
        double t = (-b - sq) / (2 * a);
        if (!valid(ray, t, tmin, tmax)) {
            t = (-b + sq) / (2 * a);
            if (!valid(ray, t, tmin, tmax)) return false;
        }
        Vec3 p = ray.at(t);
        double hh = (p - base).dot(axis);
        Vec3 radial = p - (base + axis * hh);
        h.t = t;
        h.p = p;
        h.uv = {body_u(axis, radial), hh / height};
        h.mat = mat;
        h.prim = id;
        // normalize(radial + axis (rb - rt) / height), with the unnormalized radial (:221-229)
        h.set_face_normal(ray, (radial + axis * ((rb - rt) / height)).normalize());
        return true;
    }
    bool hit(const Ray& ray, double tmin, double tmax, Hit& h, Counters* cnt) const override {  // :140-164
        if (cnt) cnt->prim++;
        bool any = false;
        double closest = tmax;
        Hit tmp;
        if (body(ray, tmin, closest, tmp)) {
            h = tmp;
            closest = tmp.t;
            any = true;
        }
        if (capped) {
            if (hit_cap(ray, base, axis * -1.0, rb, tmin, closest, mat, id, tmp)) {
                h = tmp;
                closest = tmp.t;
                any = true;
            }
            if (rt > 0 && hit_cap(ray, top, axis, rt, tmin, closest, mat, id, tmp)) {
                h = tmp;
                any = true;
            }
        }
        return any;
    }
    AABB bbox() const override { return axis_box(base, top, axis, rb); }
};

// ---- Box, pkg/geometry/box.go:34-140 ----------------------------------------------
struct Box : Shape {
    std::unique_ptr<Quad> faces[6];
    AABB box;
    Box(Vec3 center, Vec3 size, Vec3 rot, int m, int64_t first_id) {
        Vec3 c[8] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                     {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
        for (auto& p : c) {
            p = Vec3(p.x * size.x, p.y * size.y, p.z * size.z);
            p = rotate(p, rot);
            p = p + center;
        }
        faces[0].reset(new Quad(c[4], c[5] - c[4], c[7] - c[4], m, first_id + 0));
        faces[1].reset(new Quad(c[1], c[0] - c[1], c[2] - c[1], m, first_id + 1));
        faces[2].reset(new Quad(c[5], c[1] - c[5], c[6] - c[5], m, first_id + 2));
        faces[3].reset(new Quad(c[0], c[4] - c[0], c[3] - c[0], m, first_id + 3));
        faces[4].reset(new Quad(c[3], c[7] - c[3], c[2] - c[3], m, first_id + 4));
        faces[5].reset(new Quad(c[4], c[0] - c[4], c[5] - c[4], m, first_id + 5));
        box = aabb_from_points(c, 8);
    }
    bool hit(const Ray& r, double tmin, double tmax, Hit& h, Counters* cnt) const override {
        bool any = false;
        double closest = tmax;
        for (auto& f : faces) {
            Hit fh;
            if (f->hit(r, tmin, closest, fh, cnt)) {
                if (fh.t < closest) {  // strict, as box.go:131
                    closest = fh.t;
                    h = fh;
                    any = true;
                }
            }
        }
        return any;
    }
    AABB bbox() const override { return box; }
};

// ---- midpoint BVH, pkg/geometry/bvh.go ----------------------------------------------
struct BVHNode {
    AABB box;
    int left = -1, right = -1;
    std::vector<const Shape*> shapes;  // leaf when non-empty or both children absent
    bool leaf = false;
};
struct BVH : Shape {
    std::vector<BVHNode> nodes;
    int root = -1;
    Vec3 center;
    double radius = 0;
    explicit BVH(std::vector<const Shape*> shapes) {  // NewBVH :26-55
        if (shapes.empty()) {
            center = {0, 0, 0};
            radius = 0;
            return;
        }
        nodes.reserve(2 * shapes.size() / 4 + 8);
        root = build(shapes);
        center = nodes[root].box.center();
        radius = (nodes[root].box.mx - center).length();
    }
    int build(std::vector<const Shape*>& shapes) {  // buildBVH :64-109
        AABB b = shapes[0]->bbox();
        for (size_t i = 1; i < shapes.size(); i++) b = b.uni(shapes[i]->bbox());
        int id = (int)nodes.size();
        nodes.emplace_back();
        nodes[id].box = b;
        if (shapes.size() <= 8) {
            nodes[id].leaf = true;
            nodes[id].shapes = shapes;
            return id;
        }
        int axis = b.longest_axis();  // findBestSplitSimple :112-133
        double lo = b.mn[axis], hi = b.mx[axis];
        if (hi <= lo) {
            nodes[id].leaf = true;
            nodes[id].shapes = shapes;
            return id;
        }
        double split = (lo + hi) * 0.5;
        std::vector<const Shape*> L, R;  // partitionShapesSimple :136-160
        for (auto* s : shapes) {
            if (s->bbox().center()[axis] < split) L.push_back(s);
            else R.push_back(s);
        }
        if (L.empty() || R.empty()) {
            nodes[id].leaf = true;
            nodes[id].shapes = shapes;
            return id;
        }
        std::vector<const Shape*>().swap(shapes);
        int l = build(L);
        int r = build(R);
        nodes[id].left = l;
        nodes[id].right = r;
        return id;
    }
    bool hit_node(int ni, const Ray& r, double tmin, double tmax, Hit& h, Counters* c) const {
        const BVHNode& n = nodes[ni];  // hitNode :191-239
        if (c) c->box++;
        if (!n.box.hit(r, tmin, tmax)) return false;
        bool any = false;
        double closest = tmax;
        if (n.leaf) {
            for (auto* s : n.shapes) {
                Hit sh;
                if (s->hit(r, tmin, closest, sh, c)) {
                    any = true;
                    closest = sh.t;
                    h = sh;
                }
            }
            return any;
        }
        Hit sh;
        if (n.left >= 0 && hit_node(n.left, r, tmin, closest, sh, c)) {
            any = true;
            closest = sh.t;
            h = sh;
        }
        if (n.right >= 0 && hit_node(n.right, r, tmin, closest, sh, c)) {
            any = true;
            closest = sh.t;
            h = sh;
        }
        return any;
    }
    bool hit(const Ray& r, double tmin, double tmax, Hit& h, Counters* c) const override {
        if (root < 0) return false;
        return hit_node(root, r, tmin, tmax, h, c);
    }
    AABB bbox() const override { return root < 0 ? AABB{} : nodes[root].box; }
};

// ---- TriangleMesh, pkg/geometry/triangle_mesh.go:33-152 -----------------------------
struct Mesh : Shape {
    std::vector<std::unique_ptr<Triangle>> tris;
    std::unique_ptr<BVH> bvh;
    AABB box;
    // rotateVertex, triangle_mesh.go:170-199: X, then Y, then Z
    static Vec3 rotate(Vec3 v, Vec3 rot) {
        if (rot.x != 0) {
            double c = std::cos(rot.x), s = std::sin(rot.x);
            v = Vec3(v.x, v.y * c - v.z * s, v.y * s + v.z * c);
        }
        if (rot.y != 0) {
            double c = std::cos(rot.y), s = std::sin(rot.y);
            v = Vec3(v.x * c + v.z * s, v.y, -v.x * s + v.z * c);
        }
        if (rot.z != 0) {
            double c = std::cos(rot.z), s = std::sin(rot.z);
            v = Vec3(v.x * c - v.y * s, v.x * s + v.y * c, v.z);
        }
        return v;
    }
    Mesh(const prt_mesh& m, int64_t first_id) {
        tris.reserve(m.num_faces);
        std::vector<const Shape*> sh;
        sh.reserve(m.num_faces);
        std::vector<Vec3> verts(m.num_vertices);
        for (int64_t k = 0; k < m.num_vertices; k++) {  // NewTriangleMesh :57-70
            Vec3 v = V(m.vertices + 3 * k);
            if (m.rotation) {
                if (m.center) v = v - V(m.center);
                v = rotate(v, V(m.rotation));
                if (m.center) v = v + V(m.center);
            }
            verts[k] = v;
        }
        for (int64_t i = 0; i < m.num_faces; i++) {
            const int32_t* f = m.faces + 3 * i;
            int mat = m.face_materials ? m.face_materials[i] : m.material;
            tris.emplace_back(new Triangle(verts[f[0]], verts[f[1]], verts[f[2]], mat, first_id + i));
            if (m.normals) tris.back()->normal = V(m.normals + 3 * i).normalize();  // NewTriangleWithNormal :35-49
            if (m.uvs) {  // TriangleMeshOptions.VertexUVs, triangle_mesh.go:97-112
                Triangle& t = *tris.back();
                t.has_uv = true;
                t.uv0 = {m.uvs[2 * f[0]], m.uvs[2 * f[0] + 1]};
                t.uv1 = {m.uvs[2 * f[1]], m.uvs[2 * f[1] + 1]};
                t.uv2 = {m.uvs[2 * f[2]], m.uvs[2 * f[2] + 1]};
            }
            sh.push_back(tris.back().get());
        }
        bvh.reset(new BVH(sh));
        if (!tris.empty()) {
            box = tris[0]->bbox();
            for (size_t i = 1; i < tris.size(); i++) box = box.uni(tris[i]->bbox());
        }
    }
    bool hit(const Ray& r, double tmin, double tmax, Hit& h, Counters* c) const override {
        return bvh->hit(r, tmin, tmax, h, c);
    }
    AABB bbox() const override { return box; }
};

// ---- Camera, pkg/geometry/camera.go:70-190 ------------------------------------------
struct Camera {
    Vec3 center, pixel00, du, dv, disk_u, disk_v, u, v, w, forward;
    double focus = 0, lens_radius = 0, lens_area = 1, vw = 0, vh = 0, plane_area = 0;
    double cos_total_width = 0;
    int width = 0, height = 0;
    void init(const prt_camera& c) {
        Vec3 C = V(c.center), L = V(c.look_at), U = V(c.up);
        w = (C - L).normalize();
        u = w.cross(U).normalize();
        v = w.cross(u);
        focus = c.focus_distance;
        if (focus <= 0) focus = (C - L).length();
        lens_radius = c.aperture / 2;
        lens_area = kPi * lens_radius * lens_radius;
        if (c.aperture == 0) lens_area = 1.0;
        width = c.width;
        height = (int)((double)width / c.aspect_ratio);
        double theta = c.vfov * kPi / 180.0;
        double hh = std::tan(theta / 2.0);
        vh = 2.0 * hh * focus;
        vw = vh * c.aspect_ratio;
        double vh1 = 2.0 * std::tan(theta / 2.0);
        plane_area = (vh1 * c.aspect_ratio) * vh1;
        Vec3 VU = u * vw, VV = v * vh;
        du = VU * (1.0 / (double)width);
        dv = VV * (1.0 / (double)height);
        Vec3 ul = C - w * focus - VU * 0.5 - VV * 0.5;
        pixel00 = ul + (du + dv) * 0.5;
        disk_u = u * (c.aperture / 2);
        disk_v = v * (c.aperture / 2);
        forward = (L - C).normalize();
        Vec3 corner = Vec3(-vw / 2, -vh / 2, -focus).normalize();
        cos_total_width = corner.dot(Vec3(0, 0, -1));
        center = C;
    }
    Ray get_ray(int i, int j, Vec2 sp, Vec2 sj) const {  // GetRay :173-190
        Vec3 jit(sj.x - 0.5, sj.y - 0.5, 0);
        Vec3 ps = pixel00 + du * ((double)i + jit.x) + dv * ((double)j + jit.y);
        Vec3 o = center;
        if (disk_u.length() > 0) {
            Vec3 p = sample_point_in_unit_disk(sp);
            o = center + disk_u * p.x + disk_v * p.y;
        }
        return {o, (ps - o).normalize()};
    }
};

// ---- materials, pkg/material ----------------------------------------------------------
static const int kMatSpotEmitter = 5;  // discSpotLightMaterial, disc_spot_light.go:12-46
struct Material {
    int type = 0, outer = -1, inner = -1;
    int tex = 0;  // 0: solid albedo, k > 0: texture k - 1
    Vec3 albedo, emission;
    double fuzz = 0, ior = 1, ratio = 0;
    Vec3 spot_dir;  // kMatSpotEmitter: the spot direction
};
struct Texture {
    int w = 0, h = 0;
    std::vector<double> rgb;  // h*w*3, row 0 at the top
};
// ImageTexture.Evaluate (image_texture.go:24-55): wrap to [0, 1) by truncation, flip v, nearest
// texel, clamp.
static Vec3 texture_eval(const Texture& t, Vec2 uv) {
    double u = uv.x - (double)(int64_t)uv.x, v = uv.y - (double)(int64_t)uv.y;
    if (u < 0) u += 1.0;
    if (v < 0) v += 1.0;
    int64_t x = (int64_t)(u * t.w), y = (int64_t)((1.0 - v) * t.h);
    x = std::max<int64_t>(0, std::min<int64_t>(x, t.w - 1));
    y = std::max<int64_t>(0, std::min<int64_t>(y, t.h - 1));
    const double* p = &t.rgb[3 * (size_t)(y * t.w + x)];
    return {p[0], p[1], p[2]};
}
struct Scatter {
    Ray in, out;
    Vec3 att;
    double pdf = 0;
    bool specular() const { return pdf <= 0; }  // IsSpecular interfaces.go:42-44
};

static Vec3 reflect(Vec3 v, Vec3 n) { return v - n * (2 * v.dot(n)); }  // metal.go:87-90
static Vec3 refract(Vec3 uv, Vec3 n, double eta) {                      // dielectric.go:155-161
    double ct = std::min(-uv.dot(n), 1.0);
    Vec3 perp = (uv + n * ct) * eta;
    Vec3 par = n * (-std::sqrt(std::fabs(1.0 - perp.length2())));
    return perp + par;
}
static double reflectance(double cosine, double ratio) {  // Schlick, dielectric.go:164-172
    double r0 = (1 - ratio) / (1 + ratio);
    r0 = r0 * r0;
    return r0 + (1 - r0) * std::pow(1 - cosine, 5);
}
static bool is_reflection_path(Vec3 in, Vec3 out, Vec3 n) {  // layered.go:110-119
    Vec3 inc = in.neg();
    Vec3 refl = inc - n * (2 * inc.dot(n));
    return (out - refl).length() < 0.1;
}

struct Scene;
static bool mat_scatter(const Scene& S, int m, const Ray& in, const Hit& h, Sampler& smp, Scatter& out);
static Vec3 mat_brdf(const Scene& S, int m, Vec3 in, Vec3 out, const Hit& h, int mode);
static double mat_pdf(const Scene& S, int m, Vec3 in, Vec3 out, Vec3 n, bool& delta);

// ---- lights, pkg/lights ------------------------------------------------------------------
struct LightSample {
    Vec3 p, n, dir, emission;
    double dist = 0, pdf = 0;
};
struct EnvMap {
    int w = 0, h = 0;
    std::vector<double> rgb;       // h*w*3
    std::vector<double> row_cdf;   // h+1
    std::vector<double> col_cdf;   // h*(w+1)
    std::vector<double> func;      // h*w, luminance * sin(theta)
    double integral = 0;
};
struct Light {
    int type = 0, shape = -1;
    Vec3 emission, top, bottom;
    // quad light (quad_light.go) / sphere light (sphere_light.go) geometry copies
    std::unique_ptr<Quad> quad;
    double area = 0;
    std::unique_ptr<Sphere> sphere;
    int mat = -1;
    Vec3 world_center;
    double world_radius = 0;
    std::shared_ptr<EnvMap> env;
    // disc lights (disc_light.go), spot lights (disc_spot_light.go, point_spot_light.go)
    std::unique_ptr<Disc> disc;
    Vec3 from, spot_dir;
    double cos_total = 0, cos_falloff = 0;
};

struct Scene {
    std::vector<Material> mats;
    std::vector<std::unique_ptr<Shape>> owned;
    std::vector<const Shape*> shapes;
    std::unique_ptr<BVH> bvh;
    std::vector<Light> lights;
    std::vector<double> weights;
    std::vector<Texture> textures;
    Camera cam;
    int64_t num_prims = 0;
};

// Albedo.Evaluate(hit.UV, hit.Point): SolidColor or ImageTexture (color_source.go, image_texture.go)
static Vec3 albedo_at(const Scene& S, const Material& M, const Hit& h) {
    return M.tex > 0 ? texture_eval(S.textures[M.tex - 1], h.uv) : M.albedo;
}

// Emitter.Emit for the ray direction d: emissive.go:30-37; the disc spot light's material additionally
// emits only against its spot direction (rayIn.Direction . spot <= -0.3, disc_spot_light.go:26-40)
static Vec3 emissive_emit(const Scene& S, int m, const Hit* h, Vec3 d) {
    if (h && !h->front) return {0, 0, 0};
    const Material& M = S.mats[m];
    if (M.type == kMatSpotEmitter && d.dot(M.spot_dir) > -0.3) return {0, 0, 0};
    return M.emission;
}
static bool is_emitter(const Scene& S, int m) {
    return m >= 0 && (S.mats[m].type == PRT_MAT_EMISSIVE || S.mats[m].type == kMatSpotEmitter);
}

static bool mat_scatter(const Scene& S, int m, const Ray& in, const Hit& h, Sampler& smp, Scatter& out) {
    const Material& M = S.mats[m];
    switch (M.type) {
        case PRT_MAT_LAMBERTIAN: {  // lambertian.go:25-50
            Vec3 d = sample_cosine_hemisphere(h.n, smp.get2d());
            double ct = d.normalize().dot(h.n);
            if (ct < 0) ct = 0;
            out.in = in;
            out.out = {h.p, d};
            out.pdf = ct / kPi;
            out.att = albedo_at(S, M, h) * (1.0 / kPi);
            return true;
        }
        case PRT_MAT_METAL: {  // metal.go:38-59
            Vec3 r = reflect(in.d.normalize(), h.n);
            if (M.fuzz > 0) r = r + sample_point_in_unit_sphere(smp.get3d()) * M.fuzz;
            out.in = in;
            out.out = {h.p, r};
            out.att = albedo_at(S, M, h);
            out.pdf = 0;
            return r.dot(h.n) > 0;
        }
        case PRT_MAT_DIELECTRIC: {  // dielectric.go:20-60
            double ratio = h.front ? 1.0 / M.ior : M.ior;
            Vec3 ud = in.d.normalize();
            double ct = std::min(-ud.dot(h.n), 1.0);
            double st = std::sqrt(1.0 - ct * ct);
            bool cannot = ratio * st > 1.0;
            Vec3 d;
            if (cannot || reflectance(ct, ratio) > smp.get1d()) d = reflect(ud, h.n);
            else d = refract(ud, h.n, ratio);
            out.in = in;
            out.out = {h.p, d};
            out.att = {1, 1, 1};
            out.pdf = 0;
            return true;
        }
        case PRT_MAT_EMISSIVE:
        case kMatSpotEmitter:
            return false;  // emissive.go:24-27, disc_spot_light.go:21-23
        case PRT_MAT_LAYERED: {  // layered.go:24-69
            Hit oh = h;
            oh.mat = M.outer;
            Scatter o;
            if (!mat_scatter(S, M.outer, in, oh, smp, o)) return false;
            Vec3 sd = o.out.d.normalize();
            if (!(sd.dot(h.n) < 0)) {
                out = o;
                return true;
            }
            Ray ir{h.p, sd};
            Hit ih = h;
            ih.mat = M.inner;
            Scatter i;
            if (!mat_scatter(S, M.inner, ir, ih, smp, i)) {
                out = o;
                return true;
            }
            out.in = in;
            out.out = i.out;
            out.att = o.att.mul(i.att);
            out.pdf = i.pdf;
            return true;
        }
        case PRT_MAT_MIX: {  // mix.go:30-37: one draw picks the material
            Hit mh = h;
            mh.mat = smp.get1d() < M.ratio ? M.inner : M.outer;
            return mat_scatter(S, mh.mat, in, mh, smp, out);
        }
    }
    return false;
}

static Vec3 mat_brdf(const Scene& S, int m, Vec3 in, Vec3 outd, const Hit& h, int mode) {
    const Material& M = S.mats[m];
    switch (M.type) {
        case PRT_MAT_LAMBERTIAN: {  // lambertian.go:53-63
            if (outd.dot(h.n) <= 0) return {0, 0, 0};
            return albedo_at(S, M, h) * (1.0 / kPi);
        }
        case PRT_MAT_METAL: {  // metal.go:62-74
            Vec3 r = reflect(in.neg(), h.n);
            if ((outd - r).length() < 0.001) return albedo_at(S, M, h);
            return {0, 0, 0};
        }
        case PRT_MAT_DIELECTRIC: {  // dielectric.go:63-143
            Vec3 wi = in.normalize(), wo = outd.normalize(), n = h.n.normalize();
            double wiN = wi.dot(n), woN = wo.dot(n);
            bool same = (wiN > 0) == (woN > 0);
            if (!same) {
                Vec3 r = reflect(wi, n);
                if ((wo - r).length() < 0.001) {
                    double ct = std::fabs(wiN);
                    double eta = h.front ? 1.0 / M.ior : M.ior;
                    double f = reflectance(ct, eta);
                    return {f, f, f};
                }
            } else {
                double eta = h.front ? 1.0 / M.ior : M.ior;
                double ct = std::fabs(wiN);
                Vec3 rr = refract(wi, n, eta);
                if ((wo - rr).length() < 0.001) {
                    double tr = 1.0 - reflectance(ct, eta);
                    if (mode == 0) {
                        double etap = h.front ? M.ior : 1.0 / M.ior;
                        tr /= etap * etap;
                    }
                    return {tr, tr, tr};
                }
            }
            return {0, 0, 0};
        }
        case PRT_MAT_EMISSIVE:
        case kMatSpotEmitter:
            return {0, 0, 0};
        case PRT_MAT_LAYERED:  // layered.go:73-87
            if (is_reflection_path(in, outd, h.n)) return mat_brdf(S, M.outer, in, outd, h, mode);
            return mat_brdf(S, M.inner, in, outd, h, mode);
        case PRT_MAT_MIX:  // mix.go:40-46
            return mat_brdf(S, M.outer, in, outd, h, mode) * (1.0 - M.ratio) +
                   mat_brdf(S, M.inner, in, outd, h, mode) * M.ratio;
    }
    return {0, 0, 0};
}

static double mat_pdf(const Scene& S, int m, Vec3 in, Vec3 outd, Vec3 n, bool& delta) {
    const Material& M = S.mats[m];
    switch (M.type) {
        case PRT_MAT_LAMBERTIAN: {  // lambertian.go:66-72
            delta = false;
            double ct = outd.dot(n);
            if (ct <= 0) return 0;
            return ct / kPi;
        }
        case PRT_MAT_METAL:
        case PRT_MAT_DIELECTRIC:
            delta = true;
            return 0;
        case PRT_MAT_EMISSIVE:
        case kMatSpotEmitter:
            delta = false;
            return 0;
        case PRT_MAT_LAYERED: {  // layered.go:90-104
            bool d2;
            double p;
            if (is_reflection_path(in, outd, n)) p = mat_pdf(S, M.outer, in, outd, n, d2);
            else p = mat_pdf(S, M.inner, in, outd, n, d2);
            delta = false;
            return p;
        }
        case PRT_MAT_MIX: {  // mix.go:49-56: never delta
            bool d1, d2;
            double p = mat_pdf(S, M.outer, in, outd, n, d1) * (1.0 - M.ratio) + mat_pdf(S, M.inner, in, outd, n, d2) * M.ratio;
            delta = false;
            return p;
        }
    }
    delta = false;
    return 0;
}

// environment map (no reference counterpart; contract of pkg/lights/interfaces.go:17-43)
static void env_dir_to_uv(const EnvMap& e, Vec3 d, int& x, int& y) {
    d = d.normalize();
    double cy = std::max(-1.0, std::min(1.0, d.y));
    double theta = std::acos(cy);
    double phi = std::atan2(d.z, d.x);
    if (phi < 0) phi += 2 * kPi;
    double u = phi / (2 * kPi), v = theta / kPi;
    x = std::min((int)(u * e.w), e.w - 1);
    y = std::min((int)(v * e.h), e.h - 1);
    if (x < 0) x = 0;
    if (y < 0) y = 0;
}
static Vec3 env_eval(const EnvMap& e, Vec3 d) {
    int x, y;
    env_dir_to_uv(e, d, x, y);
    const double* p = &e.rgb[((size_t)y * e.w + x) * 3];
    return {p[0], p[1], p[2]};
}
static double env_pdf(const EnvMap& e, Vec3 d) {
    if (e.integral <= 0) return 0;
    int x, y;
    env_dir_to_uv(e, d, x, y);
    double st = std::sin((y + 0.5) / e.h * kPi);
    if (st <= 0) return 0;
    double p = e.func[(size_t)y * e.w + x] / e.integral;  // density on the unit square
    return p / (2 * kPi * kPi * st);
}
static int find_interval(const double* cdf, int n, double u) {  // largest i with cdf[i] <= u
    int lo = 0, hi = n;                                          // cdf has n+1 entries
    while (lo + 1 < hi) {
        int mid = (lo + hi) >> 1;
        if (cdf[mid] <= u) lo = mid;
        else hi = mid;
    }
    return lo;
}
static Vec3 env_sample(const EnvMap& e, Vec2 s, double& pdf) {
    int y = find_interval(e.row_cdf.data(), e.h, s.x);
    double dy = (e.row_cdf[y + 1] - e.row_cdf[y]);
    double fy = dy > 0 ? (s.x - e.row_cdf[y]) / dy : 0.5;
    const double* cc = &e.col_cdf[(size_t)y * (e.w + 1)];
    int x = find_interval(cc, e.w, s.y);
    double dx = cc[x + 1] - cc[x];
    double fx = dx > 0 ? (s.y - cc[x]) / dx : 0.5;
    double u = (x + fx) / e.w, v = (y + fy) / e.h;
    double theta = v * kPi, phi = u * 2 * kPi;
    double st = std::sin(theta);
    Vec3 d(st * std::cos(phi), std::cos(theta), st * std::sin(phi));
    double sc = std::sin((y + 0.5) / e.h * kPi);
    pdf = (sc > 0 && e.integral > 0) ? (e.func[(size_t)y * e.w + x] / e.integral) / (2 * kPi * kPi * sc) : 0;
    return d;
}

static Vec3 light_emit(const Scene& S, const Light& L, const Ray& r) {  // Light.Emit for infinite lights
    switch (L.type) {
        case PRT_LIGHT_UNIFORM_INFINITE: return L.emission;  // uniform_infinite_light.go:124-126
        case PRT_LIGHT_GRADIENT_INFINITE: {                   // gradient_infinite_light.go:133-137
            Vec3 d = r.d.normalize();
            double t = 0.5 * (d.y + 1.0);
            return L.bottom * (1.0 - t) + L.top * t;
        }
        case PRT_LIGHT_ENVMAP: return env_eval(*L.env, r.d);
        default: return {0, 0, 0};
    }
}
static bool light_is_infinite(const Light& L) {
    return L.type == PRT_LIGHT_UNIFORM_INFINITE || L.type == PRT_LIGHT_GRADIENT_INFINITE ||
           L.type == PRT_LIGHT_ENVMAP;
}

static LightSample sphere_sample_uniform(const Scene& S, const Light& L, Vec3 p, Vec2 s) {  // sphere_light.go:52-77
    const Sphere& sp = *L.sphere;
    double z = 1.0 - 2.0 * s.x;
    double r = std::sqrt(std::max(0.0, 1.0 - z * z));
    double phi = 2.0 * kPi * s.y;
    Vec3 ld(r * std::cos(phi), r * std::sin(phi), z);
    Vec3 sp_pt = sp.c + ld * sp.r;
    Vec3 dir = sp_pt - p;
    LightSample ls;
    ls.p = sp_pt;
    ls.n = ld;
    ls.dist = dir.length();
    ls.dir = dir.normalize();
    ls.pdf = 1.0 / (4.0 * kPi * sp.r * sp.r);
    ls.emission = emissive_emit(S, L.mat, nullptr, ls.dir);
    return ls;
}

// falloff, point_spot_light.go:121-139 and disc_spot_light.go (same curve)
static double spot_falloff(const Light& L, double cos_a) {
    if (cos_a < L.cos_total) return 0;
    if (cos_a >= L.cos_falloff) return 1;
    double d = (cos_a - L.cos_total) / (L.cos_falloff - L.cos_total);
    return d * d * d * d;
}
static Vec3 sample_cone(Vec3 w, double cos_total, Vec2 s) {  // SampleCone, core/sampling.go:70-94
    Vec3 u = std::fabs(w.x) > 0.1 ? Vec3(0, 1, 0) : Vec3(1, 0, 0);
    u = u.cross(w).normalize();
    Vec3 v = w.cross(u);
    double ct = 1.0 - s.x * (1.0 - cos_total);
    double st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
    double phi = 2.0 * kPi * s.y;
    return u * (st * std::cos(phi)) + v * (st * std::sin(phi)) + w * ct;
}
static double uniform_cone_pdf(double cos_total) { return 1.0 / (2.0 * kPi * (1.0 - cos_total)); }  // :83-85

// DiscLight.Sample, disc_light.go:31-79 (no front-face test: Emit is called without a hit record)
static LightSample disc_light_sample(const Scene& S, const Light& L, Vec3 p, Vec2 s) {
    LightSample ls;
    Vec3 sp, nn;
    L.disc->sample_uniform(s, sp, nn);
    Vec3 d = sp - p;
    double dist = d.length();
    Vec3 dn = d.normalize();
    ls.p = sp;
    ls.n = nn;
    if (dist == 0) {
        ls.dir = {0, 1, 0};
        ls.dist = 0;
        ls.pdf = 1.0;
        return ls;
    }
    double pdf = 1.0 / (kPi * L.disc->r * L.disc->r);
    double ct = std::fabs(nn.dot(dn * -1.0));
    ls.pdf = ct < 1e-6 ? 0.0 : pdf * dist * dist / ct;
    ls.dir = dn;
    ls.dist = dist;
    ls.emission = emissive_emit(S, L.mat, nullptr, dn);
    return ls;
}

static LightSample light_sample(const Scene& S, const Light& L, Vec3 p, Vec3 n, Vec2 s) {
    LightSample ls;
    switch (L.type) {
        case PRT_LIGHT_DISC:
            return disc_light_sample(S, L, p, s);
        case PRT_LIGHT_DISC_SPOT: {  // DiscSpotLight.Sample, disc_spot_light.go:105-116
            ls = disc_light_sample(S, L, p, s);
            Vec3 lp = (p - ls.p).normalize();
            ls.emission = ls.emission * spot_falloff(L, L.spot_dir.dot(lp));
            return ls;
        }
        case PRT_LIGHT_POINT_SPOT: {  // PointSpotLight.Sample, point_spot_light.go:47-92
            Vec3 tl = L.from - p;
            double dist = tl.length();
            ls.p = L.from;
            if (dist == 0) {
                ls.n = {0, 1, 0};
                ls.dir = {0, 1, 0};
                ls.dist = 0;
                ls.pdf = 1.0;
                return ls;
            }
            Vec3 to_light = tl.normalize();
            double att = spot_falloff(L, L.spot_dir.dot(to_light * -1.0));
            ls.emission = L.emission * (att / (dist * dist));
            ls.pdf = 1.0;
            ls.n = to_light * -1.0;
            ls.dir = to_light;
            ls.dist = dist;
            return ls;
        }
        case PRT_LIGHT_QUAD: {  // quad_light.go:36-84
            const Quad& q = *L.quad;
            Vec3 sp = q.corner + q.u * s.x + q.v * s.y;
            Vec3 tl = sp - p;
            double dist = tl.length();
            Vec3 dir = tl * (1.0 / dist);
            double pdf = 1.0 / L.area;
            double ct = std::fabs(q.normal.dot(dir * -1.0));
            ls.p = sp;
            ls.n = q.normal;
            ls.dir = dir;
            ls.dist = dist;
            if (ct < 1e-8) {
                ls.emission = {0, 0, 0};
                ls.pdf = 0;
                return ls;
            }
            ls.pdf = pdf * dist * dist / ct;
            bool front = dir.dot(q.normal) < 0;
            ls.emission = front ? emissive_emit(S, L.mat, nullptr, dir) : Vec3(0, 0, 0);
            return ls;
        }
        case PRT_LIGHT_SPHERE: {  // sphere_light.go:37-141
            const Sphere& sp = *L.sphere;
            Vec3 tc = sp.c - p;
            double dc = tc.length();
            if (dc <= sp.r) return sphere_sample_uniform(S, L, p, s);
            Vec3 w = tc.normalize();
            Vec3 u = std::fabs(w.x) > 0.1 ? Vec3(0, 1, 0) : Vec3(1, 0, 0);
            u = u.cross(w).normalize();
            Vec3 v = w.cross(u);
            double stm = sp.r / dc;
            double ctm = std::sqrt(std::max(0.0, 1.0 - stm * stm));
            double ct = 1.0 - s.x * (1.0 - ctm);
            double st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
            double phi = 2.0 * kPi * s.y;
            Vec3 dir = u * (st * std::cos(phi)) + v * (st * std::sin(phi)) + w * ct;
            Ray r{p, dir};
            Hit h;
            if (!sp.hit(r, 0.001, kInf, h, nullptr)) return sphere_sample_uniform(S, L, p, s);
            ls.p = h.p;
            ls.n = h.n;
            ls.dir = dir;
            ls.dist = h.t;
            ls.pdf = 1.0 / (2.0 * kPi * (1.0 - ctm));
            ls.emission = emissive_emit(S, L.mat, nullptr, dir);
            return ls;
        }
        case PRT_LIGHT_UNIFORM_INFINITE:
        case PRT_LIGHT_GRADIENT_INFINITE: {  // uniform_infinite_light.go:73-85, gradient :86-100
            Vec3 dir = sample_cosine_hemisphere(n, s);
            double ct = dir.dot(n);
            ls.p = p + dir * 1e10;
            ls.n = dir * -1.0;
            ls.dir = dir;
            ls.dist = kInf;
            ls.pdf = ct / kPi;
            if (L.type == PRT_LIGHT_UNIFORM_INFINITE) ls.emission = L.emission;
            else {
                double t = 0.5 * (dir.y + 1.0);
                ls.emission = L.bottom * (1.0 - t) + L.top * t;
            }
            return ls;
        }
        case PRT_LIGHT_ENVMAP: {
            double pdf;
            Vec3 dir = env_sample(*L.env, s, pdf);
            ls.p = p + dir * 1e10;
            ls.n = dir * -1.0;
            ls.dir = dir;
            ls.dist = kInf;
            ls.pdf = pdf;
            ls.emission = env_eval(*L.env, dir);
            return ls;
        }
    }
    return ls;
}

static double light_pdf(const Scene& S, const Light& L, Vec3 p, Vec3 n, Vec3 dir) {
    switch (L.type) {
        case PRT_LIGHT_DISC:
        case PRT_LIGHT_DISC_SPOT: {  // DiscLight.PDF, disc_light.go:82-101
            Ray r{p, dir};
            Hit h;
            if (!L.disc->hit(r, 0.001, kInf, h, nullptr)) return 0;
            double area = 1.0 / (kPi * L.disc->r * L.disc->r);
            double ct = std::fabs(L.disc->n.dot(dir * -1.0));
            if (ct < 1e-6) return 0;
            return area * h.t * h.t / ct;
        }
        case PRT_LIGHT_POINT_SPOT: {  // point_spot_light.go:96-118: 1 within 0.999 of the light, else 0
            Vec3 tl = L.from - p;
            if (tl.length() == 0) return 0;
            return dir.dot(tl.normalize()) > 0.999 ? 1.0 : 0.0;
        }
        case PRT_LIGHT_QUAD: {  // quad_light.go:87-104
            Ray r{p, dir};
            Hit h;
            if (!L.quad->hit(r, 0.001, kInf, h, nullptr)) return 0;
            double dist = h.t;
            double ct = std::fabs(L.quad->normal.dot(dir * -1.0));
            if (ct < 1e-8) return 0;
            return (1.0 / L.area) * dist * dist / ct;
        }
        case PRT_LIGHT_SPHERE: {  // sphere_light.go:144-165
            Ray r{p, dir};
            Hit h;
            const Sphere& sp = *L.sphere;
            if (!sp.hit(r, 0.001, kInf, h, nullptr)) return 0;
            double dc = (sp.c - p).length();
            if (dc <= sp.r) return 1.0 / (4.0 * kPi * sp.r * sp.r);
            double stm = sp.r / dc;
            double ctm = std::sqrt(std::max(0.0, 1.0 - stm * stm));
            return 1.0 / (2.0 * kPi * (1.0 - ctm));
        }
        case PRT_LIGHT_UNIFORM_INFINITE:
        case PRT_LIGHT_GRADIENT_INFINITE: {
            double ct = dir.dot(n);
            if (ct <= 0) return 0;
            return ct / kPi;
        }
        case PRT_LIGHT_ENVMAP: return env_pdf(*L.env, dir);
    }
    return 0;
}

// WeightedLightSampler.SampleLight, weighted_light_sampler.go:82-96
static int select_light(const Scene& S, double u, double& prob) {
    int n = (int)S.lights.size();
    double cum = 0;
    for (int i = 0; i < n; i++) {
        cum += S.weights[i];
        if (u <= cum) {
            prob = S.weights[i];
            return i;
        }
    }
    prob = S.weights[n - 1];
    return n - 1;
}
// CalculateLightPDF, light_sampling.go:11-25
static double total_light_pdf(const Scene& S, Vec3 p, Vec3 n, Vec3 dir) {
    if (S.lights.empty()) return 0;
    double tot = 0;
    for (size_t i = 0; i < S.lights.size(); i++) tot += light_pdf(S, S.lights[i], p, n, dir) * S.weights[i];
    return tot;
}

// ---- path tracing, pkg/integrator/path_tracing.go ------------------------------------
struct PTConfig {
    int max_depth = 50, rr_min = 20;
};

static double power_heuristic(int nf, double fp, int ng, double gp) {  // :224-234
    if (fp == 0) return 0;
    double f = nf * fp, g = ng * gp;
    return (f * f) / (f * f + g * g);
}

static Vec3 direct_light(const Scene& S, const Scatter& sc, const Hit& h, Sampler& smp) {  // :116-157
    if (S.lights.empty()) return {0, 0, 0};
    double prob;
    int li = select_light(S, smp.get1d(), prob);
    LightSample ls = light_sample(S, S.lights[li], h.p, h.n, smp.get2d());
    ls.pdf *= prob;
    if (ls.emission.luminance() <= 0 || ls.pdf <= 0) return {0, 0, 0};
    Ray shadow{h.p, ls.dir};
    Hit bh;
    if (S.bvh->hit(shadow, 0.001, ls.dist - 0.001, bh, nullptr)) return {0, 0, 0};
    double cosine = ls.dir.dot(h.n);
    if (cosine <= 0) return {0, 0, 0};
    bool delta;
    double mp = mat_pdf(S, h.mat, sc.in.d, ls.dir, h.n, delta);
    if (delta) return {0, 0, 0};
    double w = power_heuristic(1, ls.pdf, 1, mp);
    Vec3 f = mat_brdf(S, h.mat, sc.in.d, ls.dir, h, 0);
    return f.mul(ls.emission) * (cosine * w / ls.pdf);
}

static Vec3 ray_color(const Scene& S, const PTConfig& cfg, const Ray& ray, Sampler& smp, int depth,
                      Vec3 thr) {  // rayColorRecursive :35-85
    if (depth <= 0) return {0, 0, 0};
    int bounce = cfg.max_depth - depth;
    smp.set_slot((uint32_t)bounce + 1);
    // ApplyRussianRoulette :193-220 (the sample is drawn on every level)
    double u = smp.get1d();
    double rrc = 1.0;
    if (bounce >= cfg.rr_min) {
        double q = std::min(0.95, std::max(0.5, thr.luminance()));
        if (u > q) return {0, 0, 0};
        rrc = 1.0 / q;
    }
    Hit h;
    if (!S.bvh->hit(ray, 0.001, kInf, h, nullptr)) {
        Vec3 e;
        for (auto& L : S.lights)
            if (light_is_infinite(L)) e = e + light_emit(S, L, ray);
        return e * rrc;
    }
    Vec3 le = is_emitter(S, h.mat) ? emissive_emit(S, h.mat, &h, ray.d) : Vec3(0, 0, 0);  // getEmittedLight :100-105
    Scatter sc;
    if (!mat_scatter(S, h.mat, ray, h, smp, sc)) return le * rrc;
    Vec3 cs;
    if (sc.specular()) {  // calculateSpecularColor :88-97
        Vec3 nt = thr.mul(sc.att);
        Vec3 in = ray_color(S, cfg, sc.out, smp, depth - 1, nt);
        cs = sc.att.mul(in);
    } else {  // calculateDiffuseColor :100-106
        smp.set_slot((uint32_t)bounce + 1);
        smp.k = 8;  // NEE draws take stream positions 8..10 of this bounce
        Vec3 dl = direct_light(S, sc, h, smp);
        Vec3 il;
        // CalculateIndirectLighting :160-189
        if (sc.pdf > 0) {
            Vec3 sd = sc.out.d.normalize();
            double cosine = sd.dot(h.n);
            if (cosine > 0) {
                double lp = total_light_pdf(S, h.p, h.n, sd);
                double w = power_heuristic(1, sc.pdf, 1, lp);
                Vec3 nt = thr.mul(sc.att) * (cosine / sc.pdf);
                Vec3 in = ray_color(S, cfg, sc.out, smp, depth - 1, nt);
                il = (sc.att * (cosine * w / sc.pdf)).mul(in);
            }
        }
        cs = dl + il;
    }
    return (le + cs) * rrc;
}

static void build_env(EnvMap& e, const prt_light& pl) {
    e.w = pl.env_width;
    e.h = pl.env_height;
    size_t n = (size_t)e.w * e.h;
    e.rgb.resize(n * 3);
    for (size_t i = 0; i < n * 3; i++) e.rgb[i] = (double)pl.env_rgb[i];
    e.func.resize(n);
    e.col_cdf.resize((size_t)e.h * (e.w + 1));
    e.row_cdf.resize(e.h + 1);
    std::vector<double> row_sum(e.h);
    for (int y = 0; y < e.h; y++) {
        double st = std::sin((y + 0.5) / e.h * kPi);
        double* cc = &e.col_cdf[(size_t)y * (e.w + 1)];
        cc[0] = 0;
        for (int x = 0; x < e.w; x++) {
            const double* p = &e.rgb[((size_t)y * e.w + x) * 3];
            double l = 0.2126 * p[0] + 0.7152 * p[1] + 0.0722 * p[2];
            double f = std::max(l, 0.0) * st;
            e.func[(size_t)y * e.w + x] = f;
            cc[x + 1] = cc[x] + f;
        }
        row_sum[y] = cc[e.w];
        if (row_sum[y] > 0) {
            for (int x = 1; x <= e.w; x++) cc[x] /= row_sum[y];
        } else {
            for (int x = 1; x <= e.w; x++) cc[x] = (double)x / e.w;
        }
        cc[e.w] = 1.0;
    }
    e.row_cdf[0] = 0;
    for (int y = 0; y < e.h; y++) e.row_cdf[y + 1] = e.row_cdf[y] + row_sum[y];
    double tot = e.row_cdf[e.h];
    for (int y = 1; y <= e.h; y++) e.row_cdf[y] = tot > 0 ? e.row_cdf[y] / tot : (double)y / e.h;
    e.row_cdf[e.h] = 1.0;
    // density on the unit square: func / (tot / (w*h))
    e.integral = tot / ((double)e.w * e.h);
}

// =============================== BDPT (pkg/integrator/bdpt.go, bdpt_mis.go) ======================
// Draw slots of the counter stream (the reference draws from one per-tile stream in call order):
// camera path bounce b: slot 1+b; light selection/emission: slot kLightSlot; light path bounce b:
// slot kLightSlot+1+b; direct lighting at camera vertex t: slot kDirectSlot+t; light tracing from
// light vertex s: slot kSplatSlot+s.
static const uint32_t kLightSlot = 70, kDirectSlot = 140, kSplatSlot = 210;

// Camera.CalculateRayPDFs / MapRayToPixel / SampleCameraFromPoint / EvaluateRayImportance,
// pkg/geometry/camera.go:194-316
static bool cam_map_ray_to_pixel(const Camera& c, const Ray& r, int& px, int& py) {
    Vec3 ipc = c.center - c.w * c.focus;
    Vec3 tp = ipc - r.o;
    double t = tp.dot(c.w) / r.d.dot(c.w);
    if (t <= 0) return false;
    Vec3 hp = r.o + r.d * t;
    Vec3 rel = hp - ipc;
    double plx = rel.dot(c.u), ply = rel.dot(c.v);
    double nx = (plx + c.vw / 2) / c.vw, ny = (ply + c.vh / 2) / c.vh;
    px = (int)(nx * (double)c.width);
    py = (int)(ny * (double)c.height);
    return px >= 0 && px < c.width && py >= 0 && py < c.height;
}
static void cam_ray_pdfs(const Camera& c, const Ray& r, double& pos, double& dir) {
    double ct = r.d.dot(c.forward);
    pos = dir = 0;
    if (ct <= c.cos_total_width) return;
    int x, y;
    if (!cam_map_ray_to_pixel(c, r, x, y)) return;
    pos = 1.0 / c.lens_area;
    dir = 1.0 / (c.plane_area * std::pow(ct, 3));
}
static double cam_importance(const Camera& c, const Ray& r) {
    double ct = r.d.dot(c.forward);
    if (ct < c.cos_total_width) return 0;
    return 1.0 / (c.plane_area * c.lens_area * std::pow(ct, 4));
}
struct CamSample {
    Ray ray;
    double weight = 0, pdf = 0;
    bool ok = false;
};
static CamSample cam_sample_from_point(const Camera& c, Vec3 ref, Vec2 s) {
    CamSample cs;
    Vec3 lc = sample_point_in_unit_disk(s) * c.lens_radius;
    Vec3 lp = c.center + c.u * lc.x + c.v * lc.y;
    Vec3 dir = (ref - lp).normalize();
    double dist = (ref - lp).length();
    cs.ray = {lp, dir};
    double ct = std::fabs(c.forward.dot(dir));
    cs.pdf = (dist * dist) / (ct * c.lens_area);
    cs.weight = cam_importance(c, cs.ray);
    cs.ok = Vec3(cs.weight, cs.weight, cs.weight).luminance() != 0;
    return cs;
}

static bool light_is_area(const Light& L) {  // LightTypeArea lights
    return L.type == PRT_LIGHT_QUAD || L.type == PRT_LIGHT_SPHERE || L.type == PRT_LIGHT_DISC ||
           L.type == PRT_LIGHT_DISC_SPOT;
}
static bool light_is_delta(const Light& L) { return L.type == PRT_LIGHT_POINT_SPOT; }  // LightTypePoint

struct EmissionSample {
    Vec3 p, n, dir, emission;
    double area_pdf = 0, dir_pdf = 0;
};
// SampleInfiniteLight (light_sampling.go:89-115) for a given travel direction
static void infinite_disk(const Light& L, Vec3 direction, Vec2 sp, EmissionSample& es) {
    Vec3 up = std::fabs(direction.x) > 0.9 ? Vec3(0, 1, 0) : Vec3(1, 0, 0);
    Vec3 right = direction.cross(up).normalize();
    up = right.cross(direction).normalize();
    Vec3 ds = sample_point_in_unit_disk(sp);
    Vec3 dp = L.world_center + right * (ds.x * L.world_radius) + up * (ds.y * L.world_radius);
    es.p = dp + direction * -L.world_radius;
    es.dir = direction;
    es.n = direction * -1.0;
    es.area_pdf = 1.0 / (kPi * L.world_radius * L.world_radius);
}
// Light.SampleEmission (interfaces.go:31-33)
static EmissionSample light_sample_emission(const Scene& S, const Light& L, Vec2 sp, Vec2 sd) {
    EmissionSample es;
    switch (L.type) {
        case PRT_LIGHT_DISC: {  // disc_light.go:104-111 + SampleEmissionDirection light_sampling.go:57-80
            L.disc->sample_uniform(sp, es.p, es.n);
            es.area_pdf = 1.0 / (kPi * L.disc->r * L.disc->r);
            es.dir = sample_cosine_hemisphere(es.n, sd);
            es.dir_pdf = es.dir.dot(es.n) / kPi;
            es.emission = emissive_emit(S, L.mat, nullptr, es.dir);
            return es;
        }
        case PRT_LIGHT_DISC_SPOT: {  // disc_spot_light.go:170-192
            L.disc->sample_uniform(sp, es.p, es.n);
            es.dir = sample_cone(L.spot_dir, L.cos_total, sd);
            es.emission = L.emission * spot_falloff(L, es.dir.dot(L.spot_dir));
            es.area_pdf = 1.0 / (kPi * L.disc->r * L.disc->r);
            es.dir_pdf = uniform_cone_pdf(L.cos_total);
            return es;
        }
        case PRT_LIGHT_POINT_SPOT: {  // point_spot_light.go:153-182
            es.p = L.from;
            es.dir = sample_cone(L.spot_dir, L.cos_total, sd);
            es.emission = L.emission * spot_falloff(L, es.dir.dot(L.spot_dir));
            es.n = es.dir;
            es.area_pdf = 1.0;
            es.dir_pdf = uniform_cone_pdf(L.cos_total);
            return es;
        }
        case PRT_LIGHT_QUAD: {  // quad_light.go:107-126
            const Quad& q = *L.quad;
            es.p = q.corner + q.u * sp.x + q.v * sp.y;
            es.n = q.normal;
            es.dir = sample_cosine_hemisphere(q.normal, sd);
            es.area_pdf = 1.0 / L.area;
            es.dir_pdf = es.dir.dot(q.normal) / kPi;
            es.emission = emissive_emit(S, L.mat, nullptr, es.dir);
            return es;
        }
        case PRT_LIGHT_SPHERE: {  // sphere_light.go:168-181 + SampleEmissionDirection light_sampling.go:54-80
            const Sphere& s = *L.sphere;
            double z = 1.0 - 2.0 * sp.x;
            double r = std::sqrt(std::max(0.0, 1.0 - z * z));
            double phi = 2.0 * kPi * sp.y;
            Vec3 ld(r * std::cos(phi), r * std::sin(phi), z);
            es.p = s.c + ld * s.r;
            es.n = ld;
            es.area_pdf = 1.0 / (4.0 * kPi * s.r * s.r);
            es.dir = sample_cosine_hemisphere(ld, sd);
            es.dir_pdf = es.dir.dot(ld) / kPi;
            es.emission = is_emitter(S, L.mat) ? emissive_emit(S, L.mat, nullptr, es.dir) : Vec3(0, 0, 0);
            return es;
        }
        case PRT_LIGHT_UNIFORM_INFINITE:
        case PRT_LIGHT_GRADIENT_INFINITE: {  // uniform_infinite_light.go:101-112, gradient :103-115
            Vec3 d = sample_on_unit_sphere(sd);
            infinite_disk(L, d, sp, es);
            es.dir_pdf = 1.0 / (4.0 * kPi);
            if (L.type == PRT_LIGHT_UNIFORM_INFINITE) es.emission = L.emission;
            else {
                double t = 0.5 * (d.y + 1.0);
                es.emission = L.bottom * (1.0 - t) + L.top * t;
            }
            return es;
        }
        case PRT_LIGHT_ENVMAP: {  // importance-sampled direction toward the sky, travelling inward
            double pdf;
            Vec3 wi = env_sample(*L.env, sd, pdf);
            infinite_disk(L, wi * -1.0, sp, es);
            es.dir_pdf = pdf;
            es.emission = env_eval(*L.env, wi);
            return es;
        }
    }
    return es;
}
// Light.PDF_Le (interfaces.go:35-39)
static bool on_disc(const Disc& dc, Vec3 p, double tol) {  // validatePointOnDisc, disc_light.go:142-155
    Vec3 tp = p - dc.c;
    if (std::fabs(tp.dot(dc.n)) > tol) return false;
    return (tp - dc.n * tp.dot(dc.n)).length() <= dc.r;
}
static void light_pdf_le(const Light& L, Vec3 p, Vec3 d, double& pos, double& dir) {
    pos = dir = 0;
    switch (L.type) {
        case PRT_LIGHT_DISC:
        case PRT_LIGHT_DISC_SPOT: {  // disc_light.go:114-133, disc_spot_light.go:195-210
            if (!on_disc(*L.disc, p, 0.001)) return;
            double ct = d.dot(L.disc->n);
            if (ct <= 0) return;
            pos = 1.0 / (kPi * L.disc->r * L.disc->r);
            dir = ct / kPi;
            if (L.type == PRT_LIGHT_DISC_SPOT && d.dot(L.spot_dir) < L.cos_total) dir = 0;
            return;
        }
        case PRT_LIGHT_POINT_SPOT:  // point_spot_light.go:185-207
            if ((p - L.from).length() > 0.001) return;
            if (d.dot(L.spot_dir) < L.cos_total) return;
            pos = 1.0;
            dir = uniform_cone_pdf(L.cos_total);
            return;
        case PRT_LIGHT_QUAD: {  // quad_light.go:129-176
            const Quad& q = *L.quad;
            Vec3 tp = p - q.corner;
            double uu = q.u.dot(q.u), vv = q.v.dot(q.v), uv = q.u.dot(q.v);
            if (uu == 0 || vv == 0) return;
            double det = uu * vv - uv * uv;
            if (std::fabs(det) < 1e-8) return;
            double tu = tp.dot(q.u), tv = tp.dot(q.v);
            double a = (vv * tu - uv * tv) / det, b = (uu * tv - uv * tu) / det;
            if (a < 0 || a > 1 || b < 0 || b > 1) return;
            Vec3 rec = q.corner + q.u * a + q.v * b;
            if ((rec - p).length() > 0.001) return;
            pos = 1.0 / L.area;
            double ct = d.dot(q.normal);
            if (ct <= 0) return;
            dir = ct / kPi;
            return;
        }
        case PRT_LIGHT_SPHERE: {  // sphere_light.go:184-205
            const Sphere& s = *L.sphere;
            if (!(std::fabs((p - s.c).length() - s.r) <= 0.001)) return;
            Vec3 n = (p - s.c).normalize();
            if (d.dot(n) <= 0) return;
            pos = 1.0 / (4.0 * kPi * s.r * s.r);
            dir = d.dot(n) / kPi;
            return;
        }
        case PRT_LIGHT_UNIFORM_INFINITE:
        case PRT_LIGHT_GRADIENT_INFINITE:
            if (L.world_radius <= 0) return;
            pos = 1.0 / (kPi * L.world_radius * L.world_radius);
            dir = 1.0 / (4.0 * kPi);
            return;
        case PRT_LIGHT_ENVMAP:
            if (L.world_radius <= 0) return;
            pos = 1.0 / (kPi * L.world_radius * L.world_radius);
            dir = env_pdf(*L.env, d * -1.0);
            return;
    }
}

// Vertex, pkg/integrator/bdpt.go:14-44
struct BVertex {
    Vec3 p, n, wi, beta, Le;
    int mat = -1;  // -1: no material (camera, light sample, background)
    Vec2 uv;
    bool front = true;
    const Light* light = nullptr;
    int light_index = 0;
    double pdf_fwd = 0, pdf_rev = 0;
    bool is_light = false, is_camera = false, is_specular = false, is_inf = false;
    bool on_surface() const {  // IsOnSurface :33-41
        if (is_light && light) return light_is_area(*light);
        return mat >= 0;
    }
    Hit hit() const {
        Hit h;
        h.p = p;
        h.n = n;
        h.front = front;
        h.mat = mat;
        h.uv = uv;
        return h;
    }
};
struct BPath {
    std::vector<BVertex> v;
    int len() const { return (int)v.size(); }
};

// convertSolidAngleToAreaPdf, bdpt_mis.go:341-362
static double convert_density(const BVertex& from, const BVertex& next, double pdf) {
    if (next.is_inf) return pdf;
    Vec3 d = next.p - from.p;
    double d2 = d.length2();
    if (d2 == 0) return 0.0;
    double inv = 1.0 / d2;
    if (next.on_surface()) pdf *= std::fabs((d * std::sqrt(inv)).dot(next.n));
    return pdf * inv;
}
static double remap0(double f) { return f != 0 ? f : 1.0; }

struct BDPTCtx {
    const Scene& S;
    int max_depth;
};

static BVertex background_vertex(const Ray& r, Vec3 bg, Vec3 beta, double pdf_fwd) {  // bdpt.go:585-601
    BVertex v;
    v.p = r.o + r.d * 1000.0;
    v.n = r.d * -1.0;
    v.wi = r.d * -1.0;
    v.pdf_fwd = pdf_fwd;
    v.pdf_rev = 0;
    v.is_light = !bg.is_zero();
    v.is_inf = true;
    v.beta = beta;
    v.Le = bg;
    return v;
}

// extendPath, bdpt.go:206-291
static void extend_path(const BDPTCtx& C, BPath& path, Ray ray, Vec3 beta, double pdf_fwd, Sampler& smp, int max_bounces,
                        bool camera_path, uint32_t slot0) {
    const Scene& S = C.S;
    for (int b = 0; b < max_bounces; b++) {
        int prev_i = path.len() - 1;
        Hit h;
        if (!S.bvh->hit(ray, 0.001, kInf, h, nullptr)) {
            if (camera_path) {
                Vec3 e;
                for (auto& L : S.lights)
                    if (light_is_infinite(L)) e = e + light_emit(S, L, ray);
                path.v.push_back(background_vertex(ray, e, beta, pdf_fwd));
            }
            break;
        }
        BVertex v;
        v.p = h.p;
        v.n = h.n;
        v.front = h.front;
        v.mat = h.mat;
        v.uv = h.uv;
        v.wi = ray.d * -1.0;
        v.beta = beta;
        v.Le = is_emitter(S, h.mat) ? emissive_emit(S, h.mat, &h, ray.d) : Vec3(0, 0, 0);
        v.is_light = !v.Le.is_zero();
        v.pdf_fwd = convert_density(path.v[prev_i], v, pdf_fwd);
        smp.set_slot(slot0 + (uint32_t)b);
        smp.k = 1;
        Scatter sc;
        if (!mat_scatter(S, h.mat, ray, h, smp, sc)) {
            path.v.push_back(v);
            break;
        }
        pdf_fwd = sc.pdf;
        v.is_specular = sc.specular();
        double ct = std::fabs(sc.out.d.dot(h.n));
        if (sc.specular()) beta = beta.mul(sc.att);
        else beta = beta.mul(sc.att) * (ct / sc.pdf);
        bool rdelta;
        double pdf_rev = mat_pdf(S, h.mat, sc.out.d, ray.d * -1.0, h.n, rdelta);
        if (rdelta) {
            v.is_specular = true;
            pdf_rev = 0;
            pdf_fwd = 0;
        }
        path.v[prev_i].pdf_rev = convert_density(v, path.v[prev_i], pdf_rev);
        path.v.push_back(v);
        ray = sc.out;
    }
}

static double infinite_light_density(const Scene& S, Vec3 p, Vec3 n, Vec3 d) {  // bdpt_mis.go:318-337
    if (S.lights.empty()) return 0;
    double tot = 0;
    for (size_t i = 0; i < S.lights.size(); i++)
        if (light_is_infinite(S.lights[i])) tot += light_pdf(S, S.lights[i], p, n, d) * S.weights[i];
    return tot;
}

static BPath camera_path(const BDPTCtx& C, const Ray& r, Sampler& smp) {  // bdpt.go:106-135
    BPath path;
    double pos, dir;
    cam_ray_pdfs(C.S.cam, r, pos, dir);
    BVertex cv;
    cv.p = r.o;
    cv.n = r.d * -1.0;
    cv.is_camera = true;
    cv.beta = Vec3(1, 1, 1);
    path.v.push_back(cv);
    extend_path(C, path, r, Vec3(1, 1, 1), dir, smp, C.max_depth, true, 1);
    return path;
}

static BPath light_path(const BDPTCtx& C, Sampler& smp) {  // bdpt.go:139-203
    const Scene& S = C.S;
    BPath path;
    if (S.lights.empty()) return path;
    smp.set_slot(kLightSlot);
    double sel;
    int li = select_light(S, smp.get1d(), sel);
    Vec2 sp = smp.get2d();
    Vec2 sd = smp.get2d();
    const Light& L = S.lights[li];
    EmissionSample es = light_sample_emission(S, L, sp, sd);
    double ct = std::fabs(es.dir.dot(es.n));
    BVertex lv;
    lv.p = es.p;
    lv.n = es.n;
    lv.light = &L;
    lv.light_index = li;
    lv.pdf_fwd = es.area_pdf * sel;
    lv.is_light = true;
    lv.is_inf = light_is_infinite(L);
    lv.beta = es.emission;
    lv.Le = es.emission;
    path.v.push_back(lv);
    Ray ray{es.p, es.dir};
    Vec3 beta = es.emission * (ct / (sel * es.area_pdf * es.dir_pdf));
    extend_path(C, path, ray, beta, es.dir_pdf, smp, C.max_depth - 1, false, kLightSlot + 1);
    if (path.v[0].is_inf) {
        if (path.len() > 1) {
            BVertex& fb = path.v[1];
            fb.pdf_fwd = es.area_pdf * sel;
            if (fb.mat >= 0) fb.pdf_fwd *= std::fabs(ray.d.dot(fb.n));
        }
        path.v[0].pdf_fwd = infinite_light_density(S, es.p, es.n, es.dir);
    }
    return path;
}

// evaluateBRDF, bdpt.go:564-583
static Vec3 vertex_brdf(const Scene& S, const BVertex& v, Vec3 out, int mode) {
    if (v.is_inf) return {0, 0, 0};
    if (v.is_light && v.mat < 0) return {1, 1, 1};
    if (v.mat < 0) return {0, 0, 0};
    return mat_brdf(S, v.mat, v.wi, out, v.hit(), mode);
}

// calculateLightPdf / calculateLightOriginPdf / calculateVertexPdf, bdpt_mis.go:190-315
static double vtx_light_pdf(const Scene& S, const BVertex& cur, const BVertex& to) {
    Vec3 w = to.p - cur.p;
    double inv = 1.0 / w.length2();
    w = w * std::sqrt(inv);
    double pdf = 0;
    if (cur.is_light) {
        if (cur.is_inf) {
            double R = S.bvh->radius;
            if (R == 0.0) R = 1.0;
            pdf = 1.0 / (kPi * R * R);
        } else if (cur.light) {
            double pos, dir;
            light_pdf_le(*cur.light, cur.p, w, pos, dir);
            pdf = dir * inv;
        }
    }
    if (to.on_surface()) pdf *= std::fabs(to.n.dot(w));
    return pdf;
}
static double vtx_light_origin_pdf(const Scene& S, const BVertex& lv, const BVertex& to) {
    Vec3 w = to.p - lv.p;
    if (w.length2() == 0) return 0;
    w = w.normalize();
    if (lv.is_inf) return infinite_light_density(S, to.p, to.n, w * -1.0);
    if (!lv.is_light || !lv.light) return 0;
    if (S.lights.empty()) return 0;
    double choice = (lv.light_index >= 0 && lv.light_index < (int)S.weights.size()) ? S.weights[lv.light_index] : 0.0;
    double pos, dir;
    light_pdf_le(*lv.light, lv.p, w, pos, dir);
    return pos * choice;
}
static double vtx_pdf(const Scene& S, const BVertex& cur, const BVertex* prev, const BVertex& next) {
    if (cur.is_light) return vtx_light_pdf(S, cur, next);
    Vec3 wn = next.p - cur.p;
    if (wn.length2() == 0) return 0;
    wn = wn.normalize();
    Vec3 wp;
    if (prev) {
        wp = prev->p - cur.p;
        if (wp.length2() == 0) return 0;
        wp = wp.normalize();
    } else if (!cur.is_camera) {
        return 0;
    }
    double pdf;
    if (cur.is_camera) {
        double pos;
        cam_ray_pdfs(S.cam, Ray{cur.p, wn}, pos, pdf);
        if (pdf == 0) return 0;
    } else if (cur.mat >= 0) {
        bool delta;
        pdf = mat_pdf(S, cur.mat, wp, wn, cur.n, delta);
        if (delta) return 0;
    } else {
        return 0;
    }
    return convert_density(cur, next, pdf);
}

// calculateMISWeight, bdpt_mis.go:13-48 (with the per-vertex rules of :52-187)
static double mis_weight(const Scene& S, const BPath& cp, const BPath& lp, const BVertex* sv, int s, int t) {
    if (s + t == 2) return 1.0;
    double sum = 0, ri = 1.0;
    for (int i = t - 1; i > 0; i--) {
        const BVertex& v = cp.v[i];
        double fwd = v.pdf_fwd, rev = v.pdf_rev;
        bool conn = !v.is_specular;
        if (s == 0) {
            if (i == t - 1 && t > 1) {
                rev = vtx_light_origin_pdf(S, cp.v[t - 1], cp.v[t - 2]);
                conn = true;
            } else if (i == t - 2 && t > 2) {
                rev = vtx_pdf(S, cp.v[t - 1], nullptr, cp.v[t - 2]);
            }
        } else if (t == 1) {
            if (i == 0) {
                fwd = sv->pdf_fwd;
                rev = sv->pdf_rev;
                conn = true;
            }
        } else if (s == 1) {
            if (i == t - 1 && t > 0) {
                rev = vtx_pdf(S, *sv, nullptr, cp.v[t - 1]);
                conn = true;
            } else if (i == t - 2 && t > 1) {
                rev = vtx_pdf(S, cp.v[t - 1], sv, cp.v[t - 2]);
            }
        } else {
            if (i == t - 1) {
                rev = vtx_pdf(S, lp.v[s - 1], &lp.v[s - 2], cp.v[t - 1]);
                conn = true;
            } else if (i == t - 2 && t > 1) {
                rev = vtx_pdf(S, cp.v[t - 1], &lp.v[s - 1], cp.v[t - 2]);
            }
        }
        if (i > 0) conn = conn && !cp.v[i - 1].is_specular;
        ri *= remap0(rev) / remap0(fwd);
        if (conn) sum += ri;
    }
    ri = 1.0;
    for (int i = s - 1; i >= 0; i--) {
        const BVertex& v = lp.v[i];
        double fwd = v.pdf_fwd, rev = v.pdf_rev;
        // light vertices: connectible if not specular and not on a delta (point) light (bdpt_mis.go:129-139)
        auto delta = [](const BVertex& x) { return x.is_light && x.light && light_is_delta(*x.light); };
        bool conn = !v.is_specular && !delta(v);
        if (i > 0) conn = conn && !lp.v[i - 1].is_specular && !delta(lp.v[i - 1]);
        if (s == 0) {
        } else if (t == 1) {
            if (i == s - 1 && s > 1) {
                rev = vtx_pdf(S, *sv, nullptr, lp.v[s - 1]);
                conn = true;
            } else if (i == s - 2 && s > 1) {
                rev = vtx_pdf(S, lp.v[s - 1], sv, lp.v[s - 2]);
            }
        } else if (s == 1) {
            if (i == s - 1 && sv) {
                fwd = sv->pdf_fwd;
                rev = vtx_pdf(S, cp.v[t - 1], &cp.v[t - 2], *sv);
                conn = true;
            }
        } else {
            if (i == s - 1) {
                rev = vtx_pdf(S, cp.v[t - 1], &cp.v[t - 2], lp.v[s - 1]);
                conn = true;
            } else if (i == s - 2 && s > 1) {
                rev = vtx_pdf(S, lp.v[s - 1], &cp.v[t - 1], lp.v[s - 2]);
            }
        }
        ri *= remap0(rev) / remap0(fwd);
        if (conn) sum += ri;
    }
    return 1.0 / (1.0 + sum);
}

struct Splat {
    Ray ray;
    Vec3 color;
};

// RayColor, bdpt.go:73-102, with the strategies of :294-561
static Vec3 bdpt_ray_color(const BDPTCtx& C, const Ray& ray, Sampler& smp, std::vector<Splat>& splats) {
    const Scene& S = C.S;
    BPath cp = camera_path(C, ray, smp);
    BPath lp = light_path(C, smp);
    Vec3 total;
    for (int s = 0; s <= lp.len(); s++) {
        for (int t = 1; t <= cp.len(); t++) {
            if (s == 1 && t == 1) continue;
            Vec3 L;
            bool have_splat = false;
            Splat sp;
            BVertex sampled;
            const BVertex* sv = nullptr;
            if (s == 0) {  // evaluatePathTracingStrategy :316-330
                if (t < cp.len()) continue;
                const BVertex& last = cp.v[t - 1];
                L = last.Le.mul(last.beta);
            } else if (t == 1) {  // evaluateLightTracingStrategy :394-471
                if (s <= 1 || s > lp.len()) continue;
                const BVertex& lv = lp.v[s - 1];
                if (lv.is_specular) continue;
                smp.set_slot(kSplatSlot + (uint32_t)s);
                CamSample cs = cam_sample_from_point(S.cam, lv.p, smp.get2d());
                if (!cs.ok) continue;
                Vec3 f = vertex_brdf(S, lv, cs.ray.d * -1.0, 1);
                Vec3 cbeta = Vec3(cs.weight, cs.weight, cs.weight) * (1.0 / cs.pdf);
                double cosine = (cs.ray.d * -1.0).dot(lv.n);
                if (cosine <= 0) continue;
                Vec3 contrib = f.mul(cbeta).mul(lv.beta);
                if (lv.on_surface()) contrib = contrib * cosine;
                if (contrib.is_zero()) continue;
                double dist = (lv.p - cs.ray.o).length();
                Hit bh;
                if (S.bvh->hit(Ray{lv.p, cs.ray.d * -1.0}, 0.001, dist - 0.001, bh, nullptr)) continue;
                Vec3 to_light = (lv.p - cs.ray.o).normalize();
                double lpdf = total_light_pdf(S, cs.ray.o, cs.ray.d * -1.0, to_light);
                sampled.p = cs.ray.o;
                sampled.n = cs.ray.d * -1.0;
                sampled.is_camera = true;
                sampled.beta = cbeta;
                sampled.pdf_fwd = cs.pdf;
                sampled.pdf_rev = lpdf;
                sv = &sampled;
                sp.ray = cs.ray;
                sp.color = contrib;
                have_splat = true;
            } else if (s == 1) {  // evaluateDirectLightingStrategy :333-391
                const BVertex& cv = cp.v[t - 1];
                if (cv.is_specular || cv.mat < 0) continue;
                smp.set_slot(kDirectSlot + (uint32_t)t);
                double sel;
                int li = select_light(S, smp.get1d(), sel);
                LightSample ls = light_sample(S, S.lights[li], cv.p, cv.n, smp.get2d());
                ls.pdf *= sel;
                if (ls.emission.is_zero() || ls.pdf <= 0) continue;
                double ct = std::fabs(ls.dir.dot(cv.n));
                if (ct <= 0) continue;
                Vec3 f = mat_brdf(S, cv.mat, cv.wi, ls.dir, cv.hit(), 0);
                Vec3 lbeta = ls.emission * (1.0 / ls.pdf);
                Vec3 contrib = f.mul(cv.beta).mul(lbeta) * ct;
                if (contrib.is_zero()) continue;
                Hit bh;
                if (S.bvh->hit(Ray{cv.p, ls.dir}, 0.001, ls.dist - 0.001, bh, nullptr)) continue;
                const Light& L0 = S.lights[li];
                sampled.p = ls.p;  // LightSample.Point
                sampled.n = ls.n;
                sampled.light = &L0;
                sampled.light_index = li;
                sampled.pdf_fwd = ls.pdf;
                sampled.is_light = true;
                sampled.is_inf = light_is_infinite(L0);
                sampled.beta = lbeta;
                sampled.Le = ls.emission;
                sv = &sampled;
                L = contrib;
            } else {  // evaluateConnectionStrategy :489-561
                const BVertex& lv = lp.v[s - 1];
                const BVertex& cv = cp.v[t - 1];
                if (lv.is_specular || cv.is_specular) continue;
                Vec3 d = lv.p - cv.p;
                double dist = d.length();
                if (dist < 0.001) continue;
                d = d * (1.0 / dist);
                double cc = d.dot(cv.n), cl = (d * -1.0).dot(lv.n);
                if (cc <= 0 || cl <= 0) continue;
                double G = (cc * cl) / (dist * dist);
                Vec3 fc = vertex_brdf(S, cv, d, 0);
                Vec3 fl = vertex_brdf(S, lv, d * -1.0, 1);
                Vec3 contrib = lv.beta.mul(fl).mul(fc).mul(cv.beta) * G;
                if (contrib.is_zero()) continue;
                Hit bh;
                if (S.bvh->hit(Ray{cv.p, d}, 0.001, dist - 0.001, bh, nullptr)) continue;
                L = contrib;
            }
            if (L.is_zero() && !have_splat) continue;
            double w = mis_weight(S, cp, lp, sv, s, t);
            total = total + L * w;
            if (have_splat) {
                sp.color = sp.color * w;
                splats.push_back(sp);
            }
        }
    }
    return total;
}

}  // namespace orc

// =============================== C interface ===========================================
using namespace orc;

struct orc_scene {
    Scene s;
};

extern "C" {

orc_scene* orc_scene_create(const prt_scene_desc* d) {
    auto* os = new orc_scene();
    Scene& S = os->s;
    for (int i = 0; i < d->num_materials; i++) {
        const prt_material& pm = d->materials[i];
        Material m;
        m.type = pm.type;
        m.outer = pm.outer;
        m.inner = pm.inner;
        m.albedo = V(pm.albedo);
        m.emission = V(pm.emission);
        m.fuzz = std::min(1.0, std::max(0.0, pm.fuzz));
        m.ior = pm.ior;
        m.tex = pm.texture;
        m.ratio = std::min(1.0, std::max(0.0, pm.ratio));  // NewMix clamps (mix.go:17-27)
        S.mats.push_back(m);
    }
    for (int i = 0; i < d->num_textures; i++) {
        const prt_texture& pt = d->textures[i];
        Texture t;
        t.w = pt.width;
        t.h = pt.height;
        t.rgb.assign(pt.rgb, pt.rgb + (size_t)pt.width * pt.height * 3);
        S.textures.push_back(std::move(t));
    }
    int64_t id = 0;
    for (int i = 0; i < d->num_meshes; i++) {
        S.owned.emplace_back(new Mesh(d->meshes[i], id));
        id += d->meshes[i].num_faces;
    }
    for (int i = 0; i < d->num_spheres; i++) {
        const prt_sphere& p = d->spheres[i];
        S.owned.emplace_back(new Sphere(V(p.center), p.radius, p.material, id++));
    }
    for (int i = 0; i < d->num_quads; i++) {
        const prt_quad& q = d->quads[i];
        S.owned.emplace_back(new Quad(V(q.corner), V(q.u), V(q.v), q.material, id++));
    }
    for (int i = 0; i < d->num_boxes; i++) {
        const prt_box& b = d->boxes[i];
        S.owned.emplace_back(new Box(V(b.center), V(b.size), V(b.rotation), b.material, id));
        id += 6;
    }
    for (int i = 0; i < d->num_discs; i++) {
        const prt_disc& dc = d->discs[i];
        S.owned.emplace_back(new Disc(V(dc.center), V(dc.normal), dc.radius, dc.material, id++));
    }
    for (int i = 0; i < d->num_cylinders; i++) {
        const prt_cylinder& c = d->cylinders[i];
        S.owned.emplace_back(new Cylinder(V(c.base_center), V(c.top_center), c.radius, c.capped != 0, c.material, id++));
    }
    for (int i = 0; i < d->num_cones; i++) {
        const prt_cone& c = d->cones[i];
        S.owned.emplace_back(
            new Cone(V(c.base_center), c.base_radius, V(c.top_center), c.top_radius, c.capped != 0, c.material, id++));
    }
    // disc spot lights: an internal emitting material and their disc join the shapes (scene.go:117-122)
    // ... and a sphere light without a shape (shape -1: NewSphereLight without AddSphereLight) gets
    // its emissive material here too, in light order
    std::vector<int> spot_mat(d->num_lights, -1);
    for (int i = 0; i < d->num_lights; i++) {
        const prt_light& pl = d->lights[i];
        if (pl.type == PRT_LIGHT_SPHERE && pl.shape == -1) {
            Material m;
            m.type = PRT_MAT_EMISSIVE;
            m.emission = V(pl.emission);
            spot_mat[i] = (int)S.mats.size();
            S.mats.push_back(m);
            continue;
        }
        if (pl.type != PRT_LIGHT_DISC_SPOT) continue;
        Vec3 dir = (V(pl.to) - V(pl.from)).normalize();
        Material m;
        m.type = kMatSpotEmitter;
        m.emission = V(pl.emission);
        m.spot_dir = dir;
        spot_mat[i] = (int)S.mats.size();
        S.mats.push_back(m);
        S.owned.emplace_back(new Disc(V(pl.from), dir, pl.radius, spot_mat[i], id++));
    }
    S.num_prims = id;
    for (auto& o : S.owned) S.shapes.push_back(o.get());
    S.bvh.reset(new BVH(S.shapes));  // Scene.Preprocess scene.go:45-48
    for (int i = 0; i < d->num_lights; i++) {
        const prt_light& pl = d->lights[i];
        Light L;
        L.type = pl.type;
        L.shape = pl.shape;
        L.emission = V(pl.emission);
        L.top = V(pl.top);
        L.bottom = V(pl.bottom);
        L.world_center = S.bvh->center;
        L.world_radius = S.bvh->radius;
        if (pl.type == PRT_LIGHT_QUAD) {
            const prt_quad& q = d->quads[pl.shape];
            L.quad.reset(new Quad(V(q.corner), V(q.u), V(q.v), q.material, -1));
            L.area = V(q.u).cross(V(q.v)).length();  // NewQuadLight quad_light.go:19-29
            L.mat = q.material;
        } else if (pl.type == PRT_LIGHT_SPHERE && pl.shape == -1) {
            L.sphere.reset(new Sphere(V(pl.from), pl.radius, spot_mat[i], -1));
            L.mat = spot_mat[i];
        } else if (pl.type == PRT_LIGHT_SPHERE) {
            const prt_sphere& sp = d->spheres[pl.shape];
            L.sphere.reset(new Sphere(V(sp.center), sp.radius, sp.material, -1));
            L.mat = sp.material;
        } else if (pl.type == PRT_LIGHT_ENVMAP) {
            L.env = std::make_shared<EnvMap>();
            build_env(*L.env, pl);
        } else if (pl.type == PRT_LIGHT_DISC) {
            const prt_disc& dc = d->discs[pl.shape];
            L.disc.reset(new Disc(V(dc.center), V(dc.normal), dc.radius, dc.material, -1));
            L.mat = dc.material;
        } else if (pl.type == PRT_LIGHT_DISC_SPOT || pl.type == PRT_LIGHT_POINT_SPOT) {
            // NewDiscSpotLight / NewPointSpotLight: direction from -> to, cosines of the cone angles
            L.from = V(pl.from);
            L.spot_dir = (V(pl.to) - V(pl.from)).normalize();
            L.cos_total = std::cos(pl.cone_angle * kPi / 180.0);
            L.cos_falloff = std::cos((pl.cone_angle - pl.cone_delta) * kPi / 180.0);
            if (pl.type == PRT_LIGHT_DISC_SPOT) {
                L.disc.reset(new Disc(L.from, L.spot_dir, pl.radius, spot_mat[i], -1));
                L.mat = spot_mat[i];
            }
        }
        S.lights.push_back(std::move(L));
    }
    int nl = d->num_lights;  // NewUniformLightSampler / NewWeightedLightSampler
    S.weights.assign(nl, nl ? 1.0 / nl : 0.0);
    if (d->light_weights && nl) {
        double tot = 0;
        for (int i = 0; i < nl; i++) tot += d->light_weights[i];
        for (int i = 0; i < nl; i++) S.weights[i] = tot == 0 ? 1.0 / nl : d->light_weights[i] / tot;
    }
    S.cam.init(d->camera);
    return os;
}

void orc_scene_destroy(orc_scene* s) { delete s; }

void orc_scene_world(const orc_scene* s, double* center, double* radius, int64_t* nprims) {
    center[0] = s->s.bvh->center.x;
    center[1] = s->s.bvh->center.y;
    center[2] = s->s.bvh->center.z;
    *radius = s->s.bvh->radius;
    *nprims = s->s.num_prims;
}

static void parallel_for(int64_t n, int nthreads, const std::function<void(int64_t, int64_t)>& fn);

// Closest hit of each ray through the reference midpoint BVH (BVH.Hit).
// out_t: +inf on miss; out_prim: -1 on miss. brute != 0 tests every primitive instead.
void orc_intersect(const orc_scene* os, const prt_ray* rays, int64_t n, double* out_t, int64_t* out_prim,
                   double* out_u, double* out_v, int32_t brute, int32_t nthreads, uint64_t* counters) {
    const Scene& S = os->s;
    std::atomic<uint64_t> nbox{0}, nprim{0};
    parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
        Counters c;
        for (int64_t i = b; i < e; i++) {
            const prt_ray& pr = rays[i];
            Ray r{{pr.org[0], pr.org[1], pr.org[2]}, {pr.dir[0], pr.dir[1], pr.dir[2]}};
            Hit h;
            bool hit = false;
            if (brute) {
                double closest = pr.tmax;
                for (auto* sh : S.shapes) {
                    const Mesh* m = dynamic_cast<const Mesh*>(sh);
                    if (m) {
                        for (auto& t : m->tris) {
                            Hit th;
                            if (t->hit(r, pr.tmin, closest, th, nullptr)) {
                                hit = true;
                                closest = th.t;
                                h = th;
                            }
                        }
                    } else {
                        Hit th;
                        if (sh->hit(r, pr.tmin, closest, th, nullptr)) {
                            hit = true;
                            closest = th.t;
                            h = th;
                        }
                    }
                }
            } else {
                hit = S.bvh->hit(r, pr.tmin, pr.tmax, h, counters ? &c : nullptr);
            }
            out_t[i] = hit ? h.t : INFINITY;
            out_prim[i] = hit ? h.prim : -1;
            if (out_u) out_u[i] = hit ? h.uv.x : 0;
            if (out_v) out_v[i] = hit ? h.uv.y : 0;
        }
        nbox += c.box;
        nprim += c.prim;
    });
    if (counters) {
        counters[0] = nbox.load();
        counters[1] = nprim.load();
    }
}

// Path-traced radiance sums per pixel for samples [sample_offset, sample_offset + spp)
// over the pixel rectangle [x0,x1)x[y0,y1) of the camera image. Output arrays are
// (y1-y0)*(x1-x0) pixels: sum_rgb (x3), lum_sum, lum_sq_sum (PixelStats.AddSample, stats.go:41-47).
// sq_rgb (optional, x3): per-channel sums of squares of the camera-side samples, for per-pixel variances.
int orc_render(const orc_scene* os, const prt_render_params* p, int32_t spp, int32_t sample_offset,
               int32_t x0, int32_t y0, int32_t x1, int32_t y1, double* sum_rgb, double* lum_sum,
               double* lum_sq, int32_t nthreads, double* splat_rgb, double* sq_rgb) {
    const Scene& S = os->s;
    const bool bdpt = p->integrator == PRT_INTEGRATOR_BDPT;
    if (!bdpt && p->integrator != PRT_INTEGRATOR_PATH) return PRT_ERR_UNSUPPORTED;
    PTConfig cfg;
    cfg.max_depth = p->max_depth;
    cfg.rr_min = p->rr_min_bounces;
    BDPTCtx bc{S, p->max_depth};
    int W = S.cam.width;
    int tw = x1 - x0, th = y1 - y0;
    // 64x64 tiles over a worker pool, as pkg/renderer/worker_pool.go:49-131
    int ts = p->tile_size > 0 ? p->tile_size : 64;
    int ntx = (tw + ts - 1) / ts, nty = (th + ts - 1) / ts;
    int ntiles = ntx * nty;
    std::atomic<int> next{0};
    int nt_ = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
    const size_t full = (size_t)S.cam.width * S.cam.height * 3;
    std::vector<std::vector<double>> tsplat(bdpt && splat_rgb ? nt_ : 0, std::vector<double>(full, 0.0));
    auto worker = [&](int wid) {
        std::vector<Splat> splats;
        for (;;) {
            int t = next.fetch_add(1);
            if (t >= ntiles) break;
            int tx = t % ntx, ty = t / ntx;
            for (int j = y0 + ty * ts; j < std::min(y1, y0 + (ty + 1) * ts); j++)
                for (int i = x0 + tx * ts; i < std::min(x1, x0 + (tx + 1) * ts); i++) {
                    size_t o = (size_t)(j - y0) * tw + (i - x0);
                    double sr = 0, sg = 0, sb = 0, ls = 0, lq = 0, qr = 0, qg = 0, qb = 0;
                    for (int s = sample_offset; s < sample_offset + spp; s++) {
                        Sampler smp;
                        smp.key = path_key(p->seed, (uint32_t)(j * W + i), (uint32_t)s);
                        smp.set_slot(0);
                        Vec2 a = smp.get2d();
                        Vec2 b = smp.get2d();
                        Ray r = S.cam.get_ray(i, j, a, b);
                        Vec3 c;
                        if (bdpt) {
                            splats.clear();
                            c = bdpt_ray_color(bc, r, smp, splats);
                            if (!tsplat.empty())
                                for (auto& sp : splats) {  // SplatQueue.AddSplat via MapRayToPixel
                                    int sx, sy;
                                    if (cam_map_ray_to_pixel(S.cam, sp.ray, sx, sy)) {
                                        double* o = &tsplat[wid][((size_t)sy * S.cam.width + sx) * 3];
                                        o[0] += sp.color.x;
                                        o[1] += sp.color.y;
                                        o[2] += sp.color.z;
                                    }
                                }
                        } else {
                            c = ray_color(S, cfg, r, smp, cfg.max_depth, Vec3(1, 1, 1));
                        }
                        sr += c.x;
                        sg += c.y;
                        sb += c.z;
                        qr += c.x * c.x;
                        qg += c.y * c.y;
                        qb += c.z * c.z;
                        double l = c.luminance();
                        ls += l;
                        lq += l * l;
                    }
                    sum_rgb[o * 3 + 0] = sr;
                    sum_rgb[o * 3 + 1] = sg;
                    sum_rgb[o * 3 + 2] = sb;
                    if (lum_sum) lum_sum[o] = ls;
                    if (lum_sq) lum_sq[o] = lq;
                    if (sq_rgb) {
                        sq_rgb[o * 3 + 0] = qr;
                        sq_rgb[o * 3 + 1] = qg;
                        sq_rgb[o * 3 + 2] = qb;
                    }
                }
        }
    };
    std::vector<std::thread> pool;
    for (int i = 0; i < nt_; i++) pool.emplace_back(worker, i);
    for (auto& t : pool) t.join();
    if (!tsplat.empty()) {
        std::fill(splat_rgb, splat_rgb + full, 0.0);
        for (auto& b : tsplat)
            for (size_t k = 0; k < full; k++) splat_rgb[k] += b[k];
    }
    return PRT_OK;
}

// Radiance of individual samples: pixels[i] = (x, y), samples[i] = sample index.
int orc_render_samples(const orc_scene* os, const prt_render_params* p, const int32_t* pixels,
                       const int32_t* samples, int64_t n, double* out_rgb, int32_t nthreads, double* splat_rgb) {
    const Scene& S = os->s;
    const bool bdpt = p->integrator == PRT_INTEGRATOR_BDPT;
    PTConfig cfg;
    cfg.max_depth = p->max_depth;
    cfg.rr_min = p->rr_min_bounces;
    BDPTCtx bc{S, p->max_depth};
    int W = S.cam.width;
    const size_t full = (size_t)S.cam.width * S.cam.height * 3;
    std::vector<std::vector<Splat>> per(bdpt ? n : 0);
    parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
        for (int64_t k = b; k < e; k++) {
            int i = pixels[2 * k], j = pixels[2 * k + 1];
            Sampler smp;
            smp.key = path_key(p->seed, (uint32_t)(j * W + i), (uint32_t)samples[k]);
            smp.set_slot(0);
            Vec2 a = smp.get2d();
            Vec2 bb = smp.get2d();
            Ray r = S.cam.get_ray(i, j, a, bb);
            Vec3 c = bdpt ? bdpt_ray_color(bc, r, smp, per[k]) : ray_color(S, cfg, r, smp, cfg.max_depth, Vec3(1, 1, 1));
            out_rgb[3 * k] = c.x;
            out_rgb[3 * k + 1] = c.y;
            out_rgb[3 * k + 2] = c.z;
        }
    });
    if (bdpt && splat_rgb) {
        std::fill(splat_rgb, splat_rgb + full, 0.0);
        for (int64_t k = 0; k < n; k++)
            for (auto& sp : per[k]) {
                int sx, sy;
                if (cam_map_ray_to_pixel(S.cam, sp.ray, sx, sy)) {
                    double* o = &splat_rgb[((size_t)sy * S.cam.width + sx) * 3];
                    o[0] += sp.color.x;
                    o[1] += sp.color.y;
                    o[2] += sp.color.z;
                }
            }
    }
    return PRT_OK;
}

// MIS weight of strategy (s, t) for hand-built paths (calculateMISWeight, bdpt_mis.go:13-48).
// Vertex flags: 1 light, 2 camera, 4 specular, 8 infinite light. light_ref indexes the lights of
// `pool` (the light object a vertex carries, -1 for none); light_index is its index in `os`.
typedef struct orc_vertex {
    double p[3], n[3];
    int32_t mat, light_ref, light_index, flags;
    double pdf_fwd, pdf_rev;
} orc_vertex;

static BVertex from_orc(const orc_vertex& o, const orc_scene* pool) {
    BVertex v;
    v.p = V(o.p);
    v.n = V(o.n);
    v.mat = o.mat;
    v.light = (pool && o.light_ref >= 0) ? &pool->s.lights[o.light_ref] : nullptr;
    v.light_index = o.light_index;
    v.is_light = (o.flags & 1) != 0;
    v.is_camera = (o.flags & 2) != 0;
    v.is_specular = (o.flags & 4) != 0;
    v.is_inf = (o.flags & 8) != 0;
    v.pdf_fwd = o.pdf_fwd;
    v.pdf_rev = o.pdf_rev;
    return v;
}

double orc_bdpt_mis(const orc_scene* os, const orc_scene* pool, const orc_vertex* cam, int32_t ncam,
                    const orc_vertex* light, int32_t nlight, const orc_vertex* sampled, int32_t s, int32_t t) {
    BPath cp, lp;
    for (int i = 0; i < ncam; i++) cp.v.push_back(from_orc(cam[i], pool));
    for (int i = 0; i < nlight; i++) lp.v.push_back(from_orc(light[i], pool));
    BVertex sv;
    if (sampled) sv = from_orc(*sampled, pool);
    return mis_weight(os->s, cp, lp, sampled ? &sv : nullptr, s, t);
}

// Camera importance helpers (camera.go:194-316): pdfs of a ray and its pixel.
void orc_camera_ray_pdfs(const orc_scene* os, const double* o, const double* d, double* pos, double* dir,
                         int32_t* px, int32_t* py) {
    Ray r{V(o), V(d)};
    cam_ray_pdfs(os->s.cam, r, *pos, *dir);
    int x = -1, y = -1;
    if (!cam_map_ray_to_pixel(os->s.cam, r, x, y)) x = y = -1;
    *px = x;
    *py = y;
}

// ---- unit-level entry points used by the golden-vector tests ------------------------
double orc_draw(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t pos) {
    return draw(path_key(seed, pixel, sample), pos);
}
void orc_sample_cosine_hemisphere(const double* n, const double* s, double* out) {
    Vec3 r = sample_cosine_hemisphere(V(n), Vec2{s[0], s[1]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void orc_sample_on_unit_sphere(const double* s, double* out) {
    Vec3 r = sample_on_unit_sphere(Vec2{s[0], s[1]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void orc_sample_point_in_unit_disk(const double* s, double* out) {
    Vec3 r = sample_point_in_unit_disk(Vec2{s[0], s[1]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void orc_sample_point_in_unit_sphere(const double* s, double* out) {
    Vec3 r = sample_point_in_unit_sphere(Vec3(s[0], s[1], s[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
double orc_power_heuristic(int32_t nf, double f, int32_t ng, double g) { return power_heuristic(nf, f, ng, g); }
double orc_reflectance(double c, double r) { return reflectance(c, r); }
// Single-shape intersection: kind 0 triangle (9 doubles), 1 sphere (4), 2 quad (9),
// 3 disc (centre 3, normal 3, radius), 4 cylinder (base 3, top 3, radius, capped),
// 5 cone (base 3, base radius, top 3, top radius, capped).
// out: t, p(3), n(3), front, u, v.  Returns 1 on hit.
int32_t orc_shape_hit(int32_t kind, const double* g, const double* o, const double* d, double tmin,
                      double tmax, double* out) {
    Ray r{V(o), V(d)};
    Hit h;
    bool hit = false;
    if (kind == 0) hit = Triangle(V(g), V(g + 3), V(g + 6), 0, 0).hit(r, tmin, tmax, h, nullptr);
    else if (kind == 1) hit = Sphere(V(g), g[3], 0, 0).hit(r, tmin, tmax, h, nullptr);
    else if (kind == 2) hit = Quad(V(g), V(g + 3), V(g + 6), 0, 0).hit(r, tmin, tmax, h, nullptr);
    else if (kind == 3) hit = Disc(V(g), V(g + 3), g[6], 0, 0).hit(r, tmin, tmax, h, nullptr);
    else if (kind == 4) hit = Cylinder(V(g), V(g + 3), g[6], g[7] != 0, 0, 0).hit(r, tmin, tmax, h, nullptr);
    else if (kind == 5) hit = Cone(V(g), g[3], V(g + 4), g[7], g[8] != 0, 0, 0).hit(r, tmin, tmax, h, nullptr);
    if (!hit) return 0;
    out[0] = h.t;
    out[1] = h.p.x; out[2] = h.p.y; out[3] = h.p.z;
    out[4] = h.n.x; out[5] = h.n.y; out[6] = h.n.z;
    out[7] = h.front ? 1 : 0;
    out[8] = h.uv.x; out[9] = h.uv.y;
    return 1;
}
// Disc frame and box: out = right(3), up(3), box min(3), box max(3).
void orc_disc_frame(const double* g, double* out) {
    Disc dc(V(g), V(g + 3), g[6], 0, 0);
    AABB b = dc.bbox();
    out[0] = dc.right.x; out[1] = dc.right.y; out[2] = dc.right.z;
    out[3] = dc.up.x; out[4] = dc.up.y; out[5] = dc.up.z;
    out[6] = b.mn.x; out[7] = b.mn.y; out[8] = b.mn.z; out[9] = b.mx.x; out[10] = b.mx.y; out[11] = b.mx.z;
}
// ImageTexture.Evaluate of a width x height float RGB image at (u, v): out = rgb.
void orc_texture_eval(const float* rgb, int32_t width, int32_t height, double u, double v, double* out) {
    Texture t;
    t.w = width;
    t.h = height;
    t.rgb.assign(rgb, rgb + (size_t)width * height * 3);
    Vec3 c = texture_eval(t, Vec2{u, v});
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}
// Cylinder (kind 4) or cone (kind 5) box in the layout of orc_shape_hit: out = min(3), max(3).
void orc_axial_bbox(int32_t kind, const double* g, double* out) {
    AABB b = kind == 4 ? Cylinder(V(g), V(g + 3), g[6], g[7] != 0, 0, 0).bbox()
                       : Cone(V(g), g[3], V(g + 4), g[7], g[8] != 0, 0, 0).bbox();
    out[0] = b.mn.x; out[1] = b.mn.y; out[2] = b.mn.z; out[3] = b.mx.x; out[4] = b.mx.y; out[5] = b.mx.z;
}
void orc_quad_bbox(const double* g, double* out) {
    AABB b = Quad(V(g), V(g + 3), V(g + 6), 0, 0).bbox();
    out[0] = b.mn.x; out[1] = b.mn.y; out[2] = b.mn.z; out[3] = b.mx.x; out[4] = b.mx.y; out[5] = b.mx.z;
}
int32_t orc_aabb_hit(const double* box, const double* o, const double* d, double tmin, double tmax) {
    AABB b{V(box), V(box + 3)};
    return b.hit(Ray{V(o), V(d)}, tmin, tmax) ? 1 : 0;
}
// Camera: GetRay for pixel (i,j) with lens and jitter samples; also image height.
void orc_camera_get_ray(const orc_scene* os, int32_t i, int32_t j, const double* lens, const double* jit,
                        double* o, double* d) {
    Ray r = os->s.cam.get_ray(i, j, Vec2{lens[0], lens[1]}, Vec2{jit[0], jit[1]});
    o[0] = r.o.x; o[1] = r.o.y; o[2] = r.o.z;
    d[0] = r.d.x; d[1] = r.d.y; d[2] = r.d.z;
}
int32_t orc_camera_height(const orc_scene* os) { return os->s.cam.height; }
// NewCamera's derived quantities (camera.go:70-170): forward3, lens area, image-plane area at z = 1,
// cosTotalWidth, viewport width and height, focus distance, width, height.
void orc_camera_info(const orc_scene* os, double* out) {
    const Camera& c = os->s.cam;
    out[0] = c.forward.x; out[1] = c.forward.y; out[2] = c.forward.z;
    out[3] = c.lens_area; out[4] = c.plane_area; out[5] = c.cos_total_width;
    out[6] = c.vw; out[7] = c.vh; out[8] = c.focus; out[9] = c.width; out[10] = c.height;
}
// SampleCameraFromPoint (camera.go:226-262): out = ray origin3, direction3, weight (We), pdf; returns 1
// when the sample contributes (the reference returns nil otherwise).
int32_t orc_camera_sample_from_point(const orc_scene* os, const double* ref, const double* s, double* out) {
    CamSample cs = cam_sample_from_point(os->s.cam, V(ref), Vec2{s[0], s[1]});
    out[0] = cs.ray.o.x; out[1] = cs.ray.o.y; out[2] = cs.ray.o.z;
    out[3] = cs.ray.d.x; out[4] = cs.ray.d.y; out[5] = cs.ray.d.z;
    out[6] = cs.weight; out[7] = cs.pdf;
    return cs.ok ? 1 : 0;
}
// EvaluateRayImportance (camera.go:264-279): We of a ray (one channel; the reference's three are equal).
double orc_camera_importance(const orc_scene* os, const double* o, const double* d) {
    return cam_importance(os->s.cam, Ray{V(o), V(d)});
}
// Light queries on light `li` of the scene: Sample (out: dir3, dist, emission3, pdf), PDF.
void orc_light_sample(const orc_scene* os, int32_t li, const double* p, const double* n, const double* s,
                      double* out) {
    LightSample ls = light_sample(os->s, os->s.lights[li], V(p), V(n), Vec2{s[0], s[1]});
    out[0] = ls.dir.x; out[1] = ls.dir.y; out[2] = ls.dir.z; out[3] = ls.dist;
    out[4] = ls.emission.x; out[5] = ls.emission.y; out[6] = ls.emission.z; out[7] = ls.pdf;
}
double orc_light_pdf(const orc_scene* os, int32_t li, const double* p, const double* n, const double* d) {
    return light_pdf(os->s, os->s.lights[li], V(p), V(n), V(d));
}
// SampleEmission (out: p3, n3, dir3, emission3, area_pdf, dir_pdf) and PDF_Le (out: pos, dir).
void orc_light_sample_emission(const orc_scene* os, int32_t li, const double* sp, const double* sd, double* out) {
    EmissionSample es = light_sample_emission(os->s, os->s.lights[li], Vec2{sp[0], sp[1]}, Vec2{sd[0], sd[1]});
    const Vec3* v[4] = {&es.p, &es.n, &es.dir, &es.emission};
    for (int k = 0; k < 4; k++) {
        out[3 * k] = v[k]->x; out[3 * k + 1] = v[k]->y; out[3 * k + 2] = v[k]->z;
    }
    out[12] = es.area_pdf;
    out[13] = es.dir_pdf;
}
void orc_light_pdf_le(const orc_scene* os, int32_t li, const double* p, const double* d, double* out) {
    double pos = 0, dir = 0;
    light_pdf_le(os->s.lights[li], V(p), V(d), pos, dir);
    out[0] = pos;
    out[1] = dir;
}
// Material queries: scatter with explicit stream key (returns 1 if scattered;
// out: dir3, att3, pdf), BRDF (mode 0 radiance / 1 importance), PDF (out delta).
int32_t orc_material_scatter(const orc_scene* os, int32_t m, const double* ro, const double* rd, const double* hp,
                             const double* hn, int32_t front, uint32_t key, double* out) {
    Hit h;
    h.p = V(hp);
    h.n = V(hn);
    h.front = front != 0;
    h.mat = m;
    Sampler smp;
    smp.key = key;
    smp.set_slot(1);
    Scatter sc;
    bool ok = mat_scatter(os->s, m, Ray{V(ro), V(rd)}, h, smp, sc);
    out[0] = sc.out.d.x; out[1] = sc.out.d.y; out[2] = sc.out.d.z;
    out[3] = sc.att.x; out[4] = sc.att.y; out[5] = sc.att.z; out[6] = sc.pdf;
    return ok ? 1 : 0;
}
void orc_material_brdf(const orc_scene* os, int32_t m, const double* in, const double* outd, const double* hn,
                       int32_t front, int32_t mode, double* out) {
    Hit h;
    h.n = V(hn);
    h.front = front != 0;
    Vec3 f = mat_brdf(os->s, m, V(in), V(outd), h, mode);
    out[0] = f.x; out[1] = f.y; out[2] = f.z;
}
double orc_material_pdf(const orc_scene* os, int32_t m, const double* in, const double* outd, const double* n,
                        int32_t* delta) {
    bool d;
    double p = mat_pdf(os->s, m, V(in), V(outd), V(n), d);
    *delta = d ? 1 : 0;
    return p;
}
// Radiance of one explicit ray through the PT estimator (RayColor) with key/bounce slots.
void orc_ray_color(const orc_scene* os, const prt_render_params* p, const double* o, const double* d,
                   uint32_t key, double* out) {
    PTConfig cfg;
    cfg.max_depth = p->max_depth;
    cfg.rr_min = p->rr_min_bounces;
    Sampler smp;
    smp.key = key;
    Vec3 c = ray_color(os->s, cfg, Ray{V(o), V(d)}, smp, cfg.max_depth, Vec3(1, 1, 1));
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}
// Pass schedule, getSamplesForPass progressive.go:98-123
int32_t orc_samples_for_pass(int32_t initial, int32_t max_spp, int32_t max_passes, int32_t pass);

// shouldStopSampling, tile_renderer.go:77-101: no stop before max(1, int(target * AdaptiveMinSamples))
// samples; then stop when the luminance coefficient of variation is below AdaptiveThreshold, or for a
// (near) black pixel when the variance is below 1e-6.
static bool should_stop(double ls, double lq, int32_t count, int32_t target, double amin, double thr) {
    int32_t min_samples = std::max<int32_t>(1, (int32_t)((double)target * amin));
    if (count < min_samples) return false;
    double mean = ls / count, mean_sq = lq / count;
    double var = std::max(0.0, mean_sq - mean * mean);
    if (mean <= 1e-8) return var < 1e-6;
    return std::sqrt(var) / mean < thr;
}

// RenderPass for passes first_pass..last_pass (progressive.go:126-223) with the per-pixel adaptive loop
// of adaptiveSamplePixelWithSplats (tile_renderer.go:50-73): while count < target and not stopped, take
// the sample with index = count. Per-pixel state lives in the caller's arrays (full image: sum_rgb
// W*H*3, lum_sum, lum_sq, counts; splat_rgb W*H*3 for BDPT) and carries over between calls.
int orc_render_passes(const orc_scene* os, const prt_render_params* p, int32_t first_pass, int32_t last_pass,
                      int32_t nthreads, double* sum_rgb, double* lum_sum, double* lum_sq, int32_t* counts,
                      double* splat_rgb) {
    const Scene& S = os->s;
    const bool bdpt = p->integrator == PRT_INTEGRATOR_BDPT;
    if (!bdpt && p->integrator != PRT_INTEGRATOR_PATH) return PRT_ERR_UNSUPPORTED;
    PTConfig cfg;
    cfg.max_depth = p->max_depth;
    cfg.rr_min = p->rr_min_bounces;
    BDPTCtx bc{S, p->max_depth};
    const int W = S.cam.width, H = S.cam.height;
    const size_t full = (size_t)W * H * 3;
    const int nt_ = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
    for (int32_t pass = first_pass; pass <= last_pass; pass++) {
        const int32_t target = orc_samples_for_pass(p->initial_samples, p->max_samples_per_pixel, p->max_passes, pass);
        std::atomic<int64_t> next{0};
        std::vector<std::vector<double>> tsplat(bdpt && splat_rgb ? nt_ : 0, std::vector<double>(full, 0.0));
        auto worker = [&](int wid) {
            std::vector<Splat> splats;
            for (;;) {
                int64_t k0 = next.fetch_add(256);
                if (k0 >= (int64_t)W * H) break;
                for (int64_t k = k0; k < std::min<int64_t>(k0 + 256, (int64_t)W * H); k++) {
                    int i = (int)(k % W), j = (int)(k / W);
                    while (counts[k] < target && !should_stop(lum_sum[k], lum_sq[k], counts[k], target,
                                                              p->adaptive_min_samples, p->adaptive_threshold)) {
                        Sampler smp;
                        smp.key = path_key(p->seed, (uint32_t)(j * W + i), (uint32_t)counts[k]);
                        smp.set_slot(0);
                        Vec2 a = smp.get2d();
                        Vec2 b = smp.get2d();
                        Ray r = S.cam.get_ray(i, j, a, b);
                        Vec3 c;
                        if (bdpt) {
                            splats.clear();
                            c = bdpt_ray_color(bc, r, smp, splats);
                            if (!tsplat.empty())
                                for (auto& sp : splats) {
                                    int sx, sy;
                                    if (cam_map_ray_to_pixel(S.cam, sp.ray, sx, sy)) {
                                        double* o = &tsplat[wid][((size_t)sy * W + sx) * 3];
                                        o[0] += sp.color.x;
                                        o[1] += sp.color.y;
                                        o[2] += sp.color.z;
                                    }
                                }
                        } else {
                            c = ray_color(S, cfg, r, smp, cfg.max_depth, Vec3(1, 1, 1));
                        }
                        sum_rgb[3 * k] += c.x;
                        sum_rgb[3 * k + 1] += c.y;
                        sum_rgb[3 * k + 2] += c.z;
                        double l = c.luminance();
                        lum_sum[k] += l;
                        lum_sq[k] += l * l;
                        counts[k]++;
                    }
                }
            }
        };
        std::vector<std::thread> pool;
        for (int t = 0; t < nt_; t++) pool.emplace_back(worker, t);
        for (auto& t : pool) t.join();
        for (auto& b : tsplat)  // processSplats after the pass's tiles (progressive.go:195)
            for (size_t q = 0; q < full; q++) splat_rgb[q] += b[q];
    }
    return PRT_OK;
}

int32_t orc_samples_for_pass(int32_t initial, int32_t max_spp, int32_t max_passes, int32_t pass) {
    if (max_passes == 1) return max_spp;
    if (pass == 1) return initial;
    int32_t rem = max_spp - initial;
    int32_t rp = max_passes - 1;
    int32_t per = rem / rp;
    int32_t t = initial + (pass - 1) * per;
    if (pass == max_passes) t = max_spp;
    return t;
}

}  // extern "C"

static void parallel_for(int64_t n, int nthreads, const std::function<void(int64_t, int64_t)>& fn) {
    int nt = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
    if (nt <= 1 || n < 1024) {
        fn(0, n);
        return;
    }
    const int64_t chunk = 256;
    std::atomic<int64_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++)
        pool.emplace_back([&]() {
            for (;;) {
                int64_t b = next.fetch_add(chunk);
                if (b >= n) break;
                fn(b, std::min(n, b + chunk));
            }
        });
    for (auto& t : pool) t.join();
}
