// sse.cpp — the reference web server's progressive-render event stream as text (include/prt_sse.h).
//
// Host code only: PNG encoding of RGBA8 tiles (zlib), base64, and the JSON of TileUpdate and the
// passComplete event (web/server/render.go:22-30, 285-368), formatted the way Go's encoding/json
// writes these structs (field order, integer and float formatting).
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <charconv>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/prt_sse.h"
#include "prt_internal.h"

using prt::fail;

namespace {

void put_be32(std::string& s, uint32_t v) {
    s.push_back((char)(v >> 24));
    s.push_back((char)(v >> 16));
    s.push_back((char)(v >> 8));
    s.push_back((char)v);
}

void put_chunk(std::string& out, const char* type, const uint8_t* data, size_t n) {
    put_be32(out, (uint32_t)n);
    size_t at = out.size();
    out.append(type, 4);
    out.append((const char*)data, n);
    uLong crc = crc32(0L, Z_NULL, 0);
    crc = crc32(crc, (const Bytef*)out.data() + at, (uInt)(n + 4));
    put_be32(out, (uint32_t)crc);
}

uint8_t paeth(int a, int b, int c) {
    int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return (uint8_t)a;
    return (uint8_t)(pb <= pc ? b : c);
}

// 8-bit RGB PNG; each row takes the filter with the smallest sum of |signed byte| (the usual adaptive
// heuristic, which Go's encoder also applies).
bool png_rgb(const uint8_t* rgba, int w, int h, std::string& out) {
    const size_t stride = (size_t)w * 3;
    std::vector<uint8_t> raw((stride + 1) * (size_t)h), prev(stride, 0), cur(stride), cand[5];
    for (auto& c : cand) c.resize(stride);
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++)
            for (int c = 0; c < 3; c++) cur[3 * x + c] = rgba[4 * ((size_t)y * w + x) + c];
        long best_cost = -1;
        int best = 0;
        for (int f = 0; f < 5; f++) {
            long cost = 0;
            for (size_t i = 0; i < stride; i++) {
                int a = i >= 3 ? cur[i - 3] : 0, b = prev[i], c = i >= 3 ? prev[i - 3] : 0;
                int pred = f == 0 ? 0 : f == 1 ? a : f == 2 ? b : f == 3 ? (a + b) / 2 : paeth(a, b, c);
                uint8_t v = (uint8_t)(cur[i] - pred);
                cand[f][i] = v;
                cost += std::abs((int)(int8_t)v);
            }
            if (best_cost < 0 || cost < best_cost) {
                best_cost = cost;
                best = f;
            }
        }
        uint8_t* row = &raw[(stride + 1) * (size_t)y];
        row[0] = (uint8_t)best;
        std::memcpy(row + 1, cand[best].data(), stride);
        std::swap(prev, cur);
    }
    uLongf zn = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zn);
    if (compress2(z.data(), &zn, raw.data(), (uLong)raw.size(), Z_DEFAULT_COMPRESSION) != Z_OK) return false;
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
    out.assign((const char*)sig, 8);
    uint8_t ihdr[13];
    for (int k = 0; k < 4; k++) {
        ihdr[k] = (uint8_t)((uint32_t)w >> (24 - 8 * k));
        ihdr[4 + k] = (uint8_t)((uint32_t)h >> (24 - 8 * k));
    }
    ihdr[8] = 8;   // bit depth
    ihdr[9] = 2;   // truecolour
    ihdr[10] = 0;  // deflate
    ihdr[11] = 0;  // adaptive filtering
    ihdr[12] = 0;  // no interlace
    put_chunk(out, "IHDR", ihdr, 13);
    put_chunk(out, "IDAT", z.data(), zn);
    put_chunk(out, "IEND", nullptr, 0);
    return true;
}

void base64(const std::string& in, std::string& out) {
    static const char* T = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    const uint8_t* p = (const uint8_t*)in.data();
    size_t n = in.size(), i = 0;
    out.reserve(out.size() + 4 * ((n + 2) / 3));
    for (; i + 3 <= n; i += 3) {
        uint32_t v = (uint32_t)p[i] << 16 | (uint32_t)p[i + 1] << 8 | p[i + 2];
        out.push_back(T[v >> 18]);
        out.push_back(T[(v >> 12) & 63]);
        out.push_back(T[(v >> 6) & 63]);
        out.push_back(T[v & 63]);
    }
    if (i < n) {
        uint32_t v = (uint32_t)p[i] << 16 | (i + 1 < n ? (uint32_t)p[i + 1] << 8 : 0u);
        out.push_back(T[v >> 18]);
        out.push_back(T[(v >> 12) & 63]);
        out.push_back(i + 1 < n ? T[(v >> 6) & 63] : '=');
        out.push_back('=');
    }
}

// encoding/json's float64: shortest round-trip digits, fixed notation unless |f| < 1e-6 or >= 1e21
// (then exponent form with "e-07" shortened to "e-7").
void json_float(std::string& out, double f) {
    char b[64];
    const double a = std::fabs(f);
    const bool sci = a != 0 && (a < 1e-6 || a >= 1e21);
    auto r = std::to_chars(b, b + sizeof(b), f, sci ? std::chars_format::scientific : std::chars_format::fixed);
    std::string s(b, r.ptr);
    if (sci) {
        size_t n = s.size();
        if (n >= 4 && s[n - 4] == 'e' && s[n - 3] == '-' && s[n - 2] == '0') s.erase(n - 2, 1);
    }
    out += s;
}

void json_int(std::string& out, long long v) { out += std::to_string(v); }

int give(const std::string& s, char** text, int64_t* length) {
    char* p = (char*)std::malloc(s.size() + 1);
    if (!p) return fail(PRT_ERR_OUT_OF_MEMORY, "out of host memory");
    std::memcpy(p, s.data(), s.size());
    p[s.size()] = 0;
    *text = p;
    *length = (int64_t)s.size();
    return PRT_OK;
}

// vec3ToColor (progressive.go:482-495): gamma 2, clamp to [0,1], uint8(255 * v) truncation. Go's
// min/max keep a NaN (from a negative or NaN mean) and uint8(NaN) is 0 on amd64, so those give 0.
void to_rgba8(const float* mean, int64_t npix, std::vector<uint8_t>& rgba) {
    rgba.resize(4 * (size_t)npix);
    for (int64_t i = 0; i < npix; i++) {
        for (int c = 0; c < 3; c++) {
            double v = std::sqrt((double)mean[3 * i + c]);
            v = v > 0.0 ? std::min(1.0, v) : 0.0;
            rgba[4 * i + c] = (uint8_t)(255.0 * v);
        }
        rgba[4 * i + 3] = 255;
    }
}

}  // namespace

extern "C" {

void prt_free(void* p) { std::free(p); }

int prt_png_encode_rgba8(const uint8_t* rgba, int32_t w, int32_t h, uint8_t** png, int64_t* length) {
    if (!rgba || !png || !length || w <= 0 || h <= 0) return fail(PRT_ERR_INVALID_ARGUMENT, "bad PNG arguments");
    std::string s;
    if (!png_rgb(rgba, w, h, s)) return fail(PRT_ERR_DEVICE, "zlib compression failed");
    uint8_t* p = (uint8_t*)std::malloc(s.size());
    if (!p) return fail(PRT_ERR_OUT_OF_MEMORY, "out of host memory");
    std::memcpy(p, s.data(), s.size());
    *png = p;
    *length = (int64_t)s.size();
    return PRT_OK;
}

int prt_sse_format_pass(int32_t w, int32_t h, const float* mean_rgb, const uint32_t* counts,
                        const prt_sse_pass_info* info, char** text, int64_t* length) {
    if (!mean_rgb || !counts || !info || !text || !length || w <= 0 || h <= 0 || info->tile_size <= 0)
        return fail(PRT_ERR_INVALID_ARGUMENT, "bad SSE pass arguments");
    const int64_t npix = (int64_t)w * h;
    std::vector<uint8_t> img;
    to_rgba8(mean_rgb, npix, img);
    const int ts = info->tile_size;
    const int tx = (w + ts - 1) / ts, ty = (h + ts - 1) / ts, ntiles = tx * ty;
    // one PNG per tile (NewTileGrid order, progressive.go:455-479), sent twice as RenderPass does;
    // tiles are encoded on up to 16 threads
    std::vector<std::string> b64(ntiles);
    std::atomic<int> next{0};
    std::atomic<bool> ok{true};
    auto work = [&]() {
        std::vector<uint8_t> tile;
        std::string png;
        for (int k; (k = next.fetch_add(1)) < ntiles;) {
            const int i = k % tx, j = k / tx;
            const int x0 = i * ts, y0 = j * ts, x1 = std::min(x0 + ts, w), y1 = std::min(y0 + ts, h);
            const int tw = x1 - x0, th = y1 - y0;
            tile.resize(4 * (size_t)tw * th);
            for (int y = 0; y < th; y++)
                std::memcpy(&tile[4 * (size_t)y * tw], &img[4 * ((size_t)(y0 + y) * w + x0)], 4 * (size_t)tw);
            if (!png_rgb(tile.data(), tw, th, png)) ok = false;
            else base64(png, b64[k]);
        }
    };
    const int nthreads = std::max(1, std::min({16, ntiles / 4, (int)std::thread::hardware_concurrency()}));
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; t++) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    if (!ok) return fail(PRT_ERR_DEVICE, "zlib compression failed");
    std::string out;
    for (int round = 0; round < 2; round++)
        for (int k = 0; k < ntiles; k++) {
            out += "event: tile\ndata: {\"tileX\":";
            json_int(out, k % tx);
            out += ",\"tileY\":";
            json_int(out, k / tx);
            out += ",\"imageData\":\"";
            out += b64[k];
            out += "\",\"passNumber\":";
            json_int(out, info->pass_number);
            out += ",\"tileNumber\":";
            json_int(out, round * ntiles + k + 1);
            out += ",\"totalTiles\":";
            json_int(out, ntiles);
            out += ",\"totalPasses\":";
            json_int(out, info->total_passes);
            out += "}\n\n";
        }
    // assembleCurrentImage's RenderStats (progressive.go:395-431) and CalculateAverageLuminance
    // (stats.go:59-92) of the RGBA8 image
    long long total = 0;
    int64_t mn = info->max_samples_per_pixel, mx = 0;
    double lum = 0;
    for (int64_t p = 0; p < npix; p++) {
        total += counts[p];
        mn = std::min<int64_t>(mn, counts[p]);
        mx = std::max<int64_t>(mx, counts[p]);
        const double r = img[4 * p] / 255.0, g = img[4 * p + 1] / 255.0, b = img[4 * p + 2] / 255.0;
        lum += 0.2126 * r + 0.7152 * g + 0.0722 * b;
    }
    out += "event: passComplete\ndata: {\"event\":\"passComplete\",\"passNumber\":";
    json_int(out, info->pass_number);
    out += ",\"totalPasses\":";
    json_int(out, info->total_passes);
    out += ",\"elapsedMs\":";
    json_int(out, info->elapsed_ms);
    out += ",\"totalPixels\":";
    json_int(out, npix);
    out += ",\"totalSamples\":";
    json_int(out, total);
    out += ",\"averageSamples\":";
    json_float(out, (double)total / (double)npix);
    out += ",\"maxSamples\":";
    json_int(out, info->target_samples);
    out += ",\"minSamples\":";
    json_int(out, mn);
    out += ",\"maxSamplesUsed\":";
    json_int(out, mx);
    out += ",\"primitiveCount\":";
    json_int(out, info->primitive_count);
    out += ",\"averageLuminance\":";
    json_float(out, lum / (double)npix);
    out += "}\n\n";
    return give(out, text, length);
}

int prt_sse_format_complete(char** text, int64_t* length) {
    if (!text || !length) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL argument");
    return give("event: complete\ndata: Rendering completed\n\n", text, length);
}

int prt_sse_format_error(const char* message, char** text, int64_t* length) {
    if (!text || !length) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL argument");
    std::string m = message ? message : "";
    for (char& c : m)  // one SSE data line
        if (c == '\n' || c == '\r') c = ' ';
    return give("event: error\ndata: " + m + "\n\n", text, length);
}

int prt_sse_server_event(prt_server* server, const prt_scene* scene, const prt_render_params* params,
                         int32_t tile_size, const prt_server_event* ev, char** text, int64_t* length) {
    if (!server || !scene || !params || !ev || !text || !length) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL argument");
    *text = nullptr;
    *length = 0;
    switch (ev->type) {
        case PRT_EVENT_PASS_COMPLETE: {
            const int32_t w = params->width > 0 ? params->width : scene->image_width;
            const int32_t h = params->height > 0 ? params->height : scene->image_height;
            std::vector<float> mean(3 * (size_t)w * h);
            std::vector<uint32_t> cnt((size_t)w * h);
            int rc = prt_server_read_pass(server, ev->request_id, ev->pass_number, mean.data(), cnt.data());
            if (rc != PRT_OK) return rc;
            prt_sse_pass_info info{};
            info.pass_number = ev->pass_number;
            info.total_passes = params->max_passes;
            info.target_samples = ev->target_samples;
            info.max_samples_per_pixel = params->max_samples_per_pixel;
            info.tile_size = tile_size;
            info.primitive_count = scene->info.reference_primitives;
            info.elapsed_ms = (int64_t)ev->t_ms;  // Duration.Milliseconds truncates
            return prt_sse_format_pass(w, h, mean.data(), cnt.data(), &info, text, length);
        }
        case PRT_EVENT_DONE:
            return prt_sse_format_complete(text, length);
        case PRT_EVENT_ERROR: {
            std::string m = "Rendering failed: error code " + std::to_string(ev->error_code);
            return prt_sse_format_error(m.c_str(), text, length);
        }
        default:
            return PRT_OK;
    }
}

}  // extern "C"
