// workload.cpp — deterministic generators of the benchmark inputs (see include/prt_workload.h).
#include <cmath>
#include <algorithm>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/prt_workload.h"
#include "prt_internal.h"

namespace {

using prt::mix32;

struct P3 {
    double x, y, z;
};
P3 add(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
P3 sub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
P3 mul(P3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
P3 unit(P3 a) {
    double l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return {a.x / l, a.y / l, a.z / l};
}

// lattice value in [-1, 1] at integer point (ix, iy, iz)
double lattice(int64_t ix, int64_t iy, int64_t iz, uint32_t seed) {
    uint32_t h = mix32((uint32_t)ix * 0x8da6b343U ^ mix32((uint32_t)iy * 0xd8163841U ^
                                                          mix32((uint32_t)iz * 0xcb1ab31fU ^ seed)));
    return (double)(h >> 8) * (2.0 / 16777216.0) - 1.0;
}
double fade(double t) { return t * t * (3.0 - 2.0 * t); }
double vnoise3(double x, double y, double z, uint32_t seed) {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    double tx = fade(x - fx), ty = fade(y - fy), tz = fade(z - fz);
    double c[2][2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int d = 0; d < 2; d++) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
    double x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx;
    double x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
    double x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx;
    double x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
    double y0 = x00 + (x10 - x00) * ty;
    double y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}
// 3 octaves, lacunarity 2, gain 0.5, normalised to [-1, 1]
double fbm3(P3 p, double base, uint32_t seed) {
    double sum = 0, amp = 1, freq = base, norm = 0;
    for (int o = 0; o < 3; o++) {
        sum += amp * vnoise3(p.x * freq, p.y * freq, p.z * freq, seed + 0x9E3779B9U * (uint32_t)o);
        norm += amp;
        amp *= 0.5;
        freq *= 2.0;
    }
    return sum / norm;
}

double f32(double v) { return (double)(float)v; }

}  // namespace

extern "C" {

int prt_workload_icosphere(int32_t f, double radius, const double* center, double amplitude, uint64_t seed,
                           double* vertices, int64_t* num_vertices, int32_t* faces, int64_t* num_faces) {
    if (f < 1 || f > 4096 || !num_vertices || !num_faces)
        return prt::fail(PRT_ERR_INVALID_ARGUMENT, "frequency must be in [1, 4096]");
    const int64_t nv = 10ll * f * f + 2, nf = 20ll * f * f;
    if (!vertices || !faces) {
        *num_vertices = nv;
        *num_faces = nf;
        return PRT_OK;
    }
    if (*num_vertices < nv || *num_faces < nf) return prt::fail(PRT_ERR_INVALID_ARGUMENT, "buffers too small");
    // icosahedron: 12 unit vertices, 20 outward-wound faces
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    P3 iv[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                 {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& p : iv) p = unit(p);
    const int IF[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                           {11, 10, 2}, {10, 7, 6}, {7, 1, 8},  {3, 9, 4},  {3, 4, 2},   {3, 2, 6}, {3, 6, 8},
                           {3, 8, 9},  {4, 9, 5},  {2, 4, 11}, {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
    std::vector<P3> dirs;
    dirs.reserve(nv);
    for (auto& p : iv) dirs.push_back(p);
    // edge vertices, generated from the lower-index endpoint so both faces share them
    std::unordered_map<uint32_t, int64_t> edge_base;  // key a*16+b (a<b) -> first interior vertex
    auto edge_key = [](int a, int b) { return (uint32_t)(a * 16 + b); };
    for (auto& fc : IF)
        for (int e = 0; e < 3; e++) {
            int a = fc[e], b = fc[(e + 1) % 3];
            if (a > b) std::swap(a, b);
            if (edge_base.count(edge_key(a, b))) continue;
            edge_base[edge_key(a, b)] = (int64_t)dirs.size();
            for (int k = 1; k < f; k++) dirs.push_back(unit(add(iv[a], mul(sub(iv[b], iv[a]), (double)k / f))));
        }
    // vertex index of grid point (i along a->b, j along a->c) of face fc
    std::vector<int64_t> grid((size_t)(f + 1) * (f + 1));
    int64_t fo = 0;
    for (auto& fc : IF) {
        const int a = fc[0], b = fc[1], c = fc[2];
        auto on_edge = [&](int p, int q, int k) -> int64_t {  // k-th point from p toward q
            if (k == 0) return p;
            if (k == f) return q;
            if (p < q) return edge_base[edge_key(p, q)] + (k - 1);
            return edge_base[edge_key(q, p)] + (f - k - 1);
        };
        for (int i = 0; i <= f; i++)
            for (int j = 0; j + i <= f; j++) {
                int64_t id;
                if (j == 0) id = on_edge(a, b, i);
                else if (i == 0) id = on_edge(a, c, j);
                else if (i + j == f) id = on_edge(b, c, j);
                else {
                    id = (int64_t)dirs.size();
                    P3 p = add(iv[a], add(mul(sub(iv[b], iv[a]), (double)i / f), mul(sub(iv[c], iv[a]), (double)j / f)));
                    dirs.push_back(unit(p));
                }
                grid[(size_t)i * (f + 1) + j] = id;
            }
        for (int i = 0; i < f; i++)
            for (int j = 0; j + i < f; j++) {
                int64_t p00 = grid[(size_t)i * (f + 1) + j], p10 = grid[(size_t)(i + 1) * (f + 1) + j],
                        p01 = grid[(size_t)i * (f + 1) + j + 1];
                faces[3 * fo + 0] = (int32_t)p00;
                faces[3 * fo + 1] = (int32_t)p10;
                faces[3 * fo + 2] = (int32_t)p01;
                fo++;
                if (i + j + 1 < f) {
                    int64_t p11 = grid[(size_t)(i + 1) * (f + 1) + j + 1];
                    faces[3 * fo + 0] = (int32_t)p10;
                    faces[3 * fo + 1] = (int32_t)p11;
                    faces[3 * fo + 2] = (int32_t)p01;
                    fo++;
                }
            }
    }
    if ((int64_t)dirs.size() != nv || fo != nf) return prt::fail(PRT_ERR_DEVICE, "icosphere count mismatch");
    const double cx = center ? center[0] : 0, cy = center ? center[1] : 0, cz = center ? center[2] : 0;
    for (int64_t i = 0; i < nv; i++) {
        P3 d = dirs[i];
        double r = radius * (1.0 + amplitude * fbm3(d, 4.0, (uint32_t)seed ^ (uint32_t)(seed >> 32)));
        vertices[3 * i + 0] = f32(cx + d.x * r);
        vertices[3 * i + 1] = f32(cy + d.y * r);
        vertices[3 * i + 2] = f32(cz + d.z * r);
    }
    *num_vertices = nv;
    *num_faces = nf;
    return PRT_OK;
}

int prt_workload_rays(const double* lo, const double* hi, double scale, int64_t n, uint64_t seed, prt_ray* rays) {
    if (!lo || !hi || (!rays && n > 0) || n < 0) return prt::fail(PRT_ERR_INVALID_ARGUMENT, "bad arguments");
    const double pi = 3.14159265358979323846;
    double c[3], e[3];
    for (int a = 0; a < 3; a++) {
        c[a] = 0.5 * (lo[a] + hi[a]);
        e[a] = (hi[a] - lo[a]) * scale;
    }
    auto body = [&](int64_t i) {
        uint32_t key = prt::path_key(seed, (uint32_t)(i & 0xFFFFFFFF), (uint32_t)(i >> 32));
        double u0 = prt::draw_f64(key, 0), u1 = prt::draw_f64(key, 1), u2 = prt::draw_f64(key, 2);
        double u3 = prt::draw_f64(key, 3), u4 = prt::draw_f64(key, 4);
        prt_ray& r = rays[i];
        r.org[0] = (float)(c[0] + (u0 - 0.5) * e[0]);
        r.org[1] = (float)(c[1] + (u1 - 0.5) * e[1]);
        r.org[2] = (float)(c[2] + (u2 - 0.5) * e[2]);
        double z = 1.0 - 2.0 * u3;
        double rr = std::sqrt(std::fmax(0.0, 1.0 - z * z));
        double phi = 2.0 * pi * u4;
        r.dir[0] = (float)(rr * std::cos(phi));
        r.dir[1] = (float)(rr * std::sin(phi));
        r.dir[2] = (float)z;
        r.tmin = 0.001f;
        r.tmax = INFINITY;
    };
    int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++)
        pool.emplace_back([&, t]() {
            for (int64_t i = t; i < n; i += nt) body(i);
        });
    for (auto& th : pool) th.join();
    return PRT_OK;
}

int prt_workload_sky(int32_t W, int32_t H, uint64_t seed, double sun_peak, float* rgb) {
    if (W < 2 || H < 2 || !rgb) return prt::fail(PRT_ERR_INVALID_ARGUMENT, "bad sky size");
    const double pi = 3.14159265358979323846;
    const uint32_t s32 = (uint32_t)seed ^ (uint32_t)(seed >> 32);
    // sun at the centre of texel (xs, ys): elevation ~35 degrees
    const int xs = W / 8, ys = (int)(H * 0.3);
    auto dir = [&](double x, double y) {
        double th = (y + 0.5) / H * pi, ph = (x + 0.5) / W * 2 * pi;
        return P3{std::sin(th) * std::cos(ph), std::cos(th), std::sin(th) * std::sin(ph)};
    };
    P3 sun = dir(xs, ys);
    const double sigma = 0.01;  // radians
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            P3 d = dir(x, y);
            double n = fbm3(d, 3.0, s32);
            double r, g, b;
            if (d.y >= 0) {
                double mu = 0.5 * (1.0 - d.y);
                double l = std::exp(mu + 0.5 * n);
                r = 0.35 * l;
                g = 0.55 * l;
                b = 1.0 * l;
            } else {
                double l = 0.05 * std::exp(0.3 * n);
                r = 0.6 * l;
                g = 0.5 * l;
                b = 0.4 * l;
            }
            double cosang = std::fmax(-1.0, std::fmin(1.0, d.x * sun.x + d.y * sun.y + d.z * sun.z));
            double ang = std::acos(cosang);
            double lobe = (x == xs && y == ys) ? 1.0 : std::exp(-(ang / sigma) * (ang / sigma));
            if (lobe > 1e-12) {
                r += sun_peak * lobe * 1.0;
                g += sun_peak * lobe * 0.96;
                b += sun_peak * lobe * 0.9;
            }
            if (x == xs && y == ys) g = sun_peak;  // the peak texel holds exactly sun_peak in green
            float* o = rgb + ((size_t)y * W + x) * 3;
            o[0] = (float)r;
            o[1] = (float)g;
            o[2] = (float)b;
        }
    return PRT_OK;
}

}  // extern "C"
