// ply.cpp — PLY mesh loader (include/prt_loaders.h), the loader behind the reference's dragon scene
// (LoadPLY, pkg/loaders/ply.go:64-413; pkg/scene/dragon.go:172-219).
//
// The whole file is read into memory once; vertex records are fixed-size rows decoded by a column plan
// built from the header, faces are walked sequentially.
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/prt_loaders.h"
#include "prt_internal.h"

namespace {

using prt::fail;

struct Prop {
    std::string name, type;       // scalar property
    bool list = false;
    std::string count_type, item_type;  // list property
};

int type_size(const std::string& t) {  // getTypeSize, ply.go:541-555 (unknown types count 4 bytes)
    if (t == "double" || t == "float64") return 8;
    if (t == "short" || t == "int16" || t == "ushort" || t == "uint16") return 2;
    if (t == "char" || t == "int8" || t == "uchar" || t == "uint8") return 1;
    return 4;
}

bool known_type(const std::string& t) {
    static const char* k[] = {"float", "float32", "double", "float64", "int", "int32", "uint", "uint32",
                              "short", "int16", "ushort", "uint16", "char", "int8", "uchar", "uint8"};
    for (const char* s : k)
        if (t == s) return true;
    return false;
}

float rd_f32(const uint8_t* p) {
    float v;
    std::memcpy(&v, p, 4);
    return v;
}

struct Reader {
    const uint8_t* p;
    const uint8_t* end;
    bool take(size_t n, const uint8_t** out) {
        if ((size_t)(end - p) < n) return false;
        *out = p;
        p += n;
        return true;
    }
};

void free_data(prt_ply_data* d) {
    if (!d) return;
    std::free(d->vertices);
    std::free(d->faces);
    std::free(d->normals);
    std::free(d->colors);
    std::free(d->texcoords);
    std::free(d);
}

template <class T>
T* alloc(size_t n) {
    return (T*)std::calloc(n > 0 ? n : 1, sizeof(T));
}

}  // namespace

extern "C" {

int prt_ply_load(const char* path, prt_ply_data** out) {
    if (!path || !out) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return fail(PRT_ERR_INVALID_ARGUMENT, std::string("failed to open PLY file: ") + path);
    std::vector<uint8_t> buf;
    {
        std::fseek(fp, 0, SEEK_END);
        long sz = std::ftell(fp);
        std::fseek(fp, 0, SEEK_SET);
        if (sz < 0) {
            std::fclose(fp);
            return fail(PRT_ERR_INVALID_ARGUMENT, "failed to read PLY file");
        }
        buf.resize((size_t)sz);
        size_t got = sz > 0 ? std::fread(buf.data(), 1, (size_t)sz, fp) : 0;
        std::fclose(fp);
        if (got != (size_t)sz) return fail(PRT_ERR_INVALID_ARGUMENT, "failed to read PLY file");
    }

    // ---- header (parsePLYHeader, ply.go:113-212) ----
    std::string format;
    int64_t nv = 0, nf = 0;
    std::vector<Prop> vprops, fprops;
    std::string element;
    size_t pos = 0;
    bool ended = false;
    while (pos < buf.size()) {
        size_t eol = pos;
        while (eol < buf.size() && buf[eol] != '\n') eol++;
        std::string line((const char*)buf.data() + pos, eol - pos);
        pos = eol + 1;
        // trim, split on whitespace
        std::vector<std::string> parts;
        {
            size_t i = 0;
            while (i < line.size()) {
                while (i < line.size() && std::isspace((unsigned char)line[i])) i++;
                size_t j = i;
                while (j < line.size() && !std::isspace((unsigned char)line[j])) j++;
                if (j > i) parts.emplace_back(line.substr(i, j - i));
                i = j;
            }
        }
        if (parts.size() == 1 && parts[0] == "end_header") {
            ended = true;
            break;
        }
        if (parts.empty()) continue;
        if (parts[0] == "format" && parts.size() >= 3) {
            format = parts[1];
        } else if (parts[0] == "element" && parts.size() >= 3) {
            char* endp = nullptr;
            long long c = std::strtoll(parts[2].c_str(), &endp, 10);
            if (!endp || *endp != '\0' || c < 0)
                return fail(PRT_ERR_INVALID_ARGUMENT, "failed to parse PLY header: invalid element count: " + parts[2]);
            element = parts[1];
            if (element == "vertex") nv = c;
            if (element == "face") nf = c;
        } else if (parts[0] == "property") {
            Prop pr;
            if (parts.size() >= 2 && parts[1] == "list") {  // parsePLYProperty, ply.go:215-236
                if (parts.size() < 5)
                    return fail(PRT_ERR_INVALID_ARGUMENT,
                                "failed to parse PLY header: failed to parse property: invalid list property definition");
                pr.list = true;
                pr.count_type = parts[2];
                pr.item_type = parts[3];
                pr.name = parts[4];
            } else {
                if (parts.size() < 3)
                    return fail(PRT_ERR_INVALID_ARGUMENT,
                                "failed to parse PLY header: failed to parse property: invalid property definition");
                pr.type = parts[1];
                pr.name = parts[2];
            }
            if (element == "vertex") vprops.push_back(pr);
            if (element == "face") fprops.push_back(pr);
        }
    }
    if (!ended) return fail(PRT_ERR_INVALID_ARGUMENT, "failed to parse PLY header: no end_header");
    if (format == "binary_big_endian")
        return fail(PRT_ERR_UNSUPPORTED, "binary big-endian PLY format not yet implemented");
    if (format == "ascii") return fail(PRT_ERR_UNSUPPORTED, "ASCII PLY format not yet supported");
    if (format != "binary_little_endian") return fail(PRT_ERR_UNSUPPORTED, "unsupported PLY format: " + format);

    // ---- vertex column plan (parseVertexFromBytes, ply.go:573-668): positions, normals, texture
    // coordinates come from float32 columns only, colours from uchar columns only ----
    enum Col { X, Y, Z, NX, NY, NZ, R, G, B, U, V, NCOL };
    int off[NCOL];
    for (int& o : off) o = -1;
    bool has_n = false, has_c = false, has_t = false;
    int vsize = 0;
    for (const Prop& p : vprops) {
        if (p.list) continue;  // lists are not part of the fixed-size vertex row
        const bool f32 = p.type == "float" || p.type == "float32";
        const bool u8 = p.type == "uchar" || p.type == "uint8";
        const std::string& n = p.name;
        // header flags follow the names whatever the type (ply.go:170-205)
        if (n == "nx" || n == "ny" || n == "nz") has_n = true;
        if (n == "red" || n == "r" || n == "green" || n == "g" || n == "blue" || n == "b") has_c = true;
        if (n == "u" || n == "s" || n == "texture_u" || n == "v" || n == "t" || n == "texture_v") has_t = true;
        if (f32) {
            if (n == "x") off[X] = vsize;
            else if (n == "y") off[Y] = vsize;
            else if (n == "z") off[Z] = vsize;
            else if (n == "nx") off[NX] = vsize;
            else if (n == "ny") off[NY] = vsize;
            else if (n == "nz") off[NZ] = vsize;
            else if (n == "u" || n == "s" || n == "texture_u") off[U] = vsize;
            else if (n == "v" || n == "t" || n == "texture_v") off[V] = vsize;
        } else if (u8) {
            if (n == "red" || n == "r") off[R] = vsize;
            else if (n == "green" || n == "g") off[G] = vsize;
            else if (n == "blue" || n == "b") off[B] = vsize;
        }
        vsize += type_size(p.type);
    }

    prt_ply_data* d = alloc<prt_ply_data>(1);
    d->num_vertices = nv;
    d->num_faces = nf;
    d->vertices = alloc<double>((size_t)nv * 3);
    d->faces = alloc<int32_t>((size_t)nf * 3);
    if (has_n) d->normals = alloc<double>((size_t)nv * 3);
    if (has_c) d->colors = alloc<double>((size_t)nv * 3);
    if (has_t) d->texcoords = alloc<double>((size_t)nv * 2);
    auto bail = [&](int code, const std::string& msg) {
        free_data(d);
        return fail(code, msg);
    };
    if (!d->vertices || !d->faces || (has_n && !d->normals) || (has_c && !d->colors) || (has_t && !d->texcoords))
        return bail(PRT_ERR_OUT_OF_MEMORY, "out of memory reading the PLY file");

    Reader rd{buf.data() + pos, buf.data() + buf.size()};
    const uint8_t* row = nullptr;
    auto f = [&](int c) { return off[c] >= 0 ? (double)rd_f32(row + off[c]) : 0.0; };
    auto u = [&](int c) { return off[c] >= 0 ? (double)row[off[c]] / 255.0 : 0.0; };
    for (int64_t i = 0; i < nv; i++) {
        if (!rd.take((size_t)vsize, &row))
            return bail(PRT_ERR_INVALID_ARGUMENT, "failed to read PLY data: failed to read vertex data");
        d->vertices[3 * i] = f(X);
        d->vertices[3 * i + 1] = f(Y);
        d->vertices[3 * i + 2] = f(Z);
        if (has_n) {
            d->normals[3 * i] = f(NX);
            d->normals[3 * i + 1] = f(NY);
            d->normals[3 * i + 2] = f(NZ);
        }
        if (has_c) {
            d->colors[3 * i] = u(R);
            d->colors[3 * i + 1] = u(G);
            d->colors[3 * i + 2] = u(B);
        }
        if (has_t) {
            d->texcoords[2 * i] = f(U);
            d->texcoords[2 * i + 1] = f(V);
        }
    }

    // ---- faces (ply.go:302-375); faces without a vertex_indices list add no triangle ----
    int64_t nread = 0;
    for (int64_t i = 0; i < nf; i++) {
        bool got = false;
        for (const Prop& p : fprops) {
            const uint8_t* q = nullptr;
            if (p.list && p.name == "vertex_indices") {
                int64_t cnt;
                if (p.count_type == "uchar" || p.count_type == "uint8") {
                    if (!rd.take(1, &q)) return bail(PRT_ERR_INVALID_ARGUMENT, "failed to read face vertex count");
                    cnt = q[0];
                } else if (p.count_type == "int" || p.count_type == "int32") {
                    if (!rd.take(4, &q)) return bail(PRT_ERR_INVALID_ARGUMENT, "failed to read face vertex count");
                    int32_t c32;
                    std::memcpy(&c32, q, 4);
                    cnt = c32;
                } else {
                    return bail(PRT_ERR_UNSUPPORTED, "unsupported list count type: " + p.count_type);
                }
                if (cnt != 3)
                    return bail(PRT_ERR_UNSUPPORTED, "only triangular faces supported, got " + std::to_string(cnt) +
                                                         " vertices at face " + std::to_string(i));
                const bool sgn = p.item_type == "int" || p.item_type == "int32";
                const bool uns = p.item_type == "uint" || p.item_type == "uint32";
                if (!sgn && !uns) return bail(PRT_ERR_UNSUPPORTED, "unsupported face index data type: " + p.item_type);
                if (!rd.take(12, &q)) return bail(PRT_ERR_INVALID_ARGUMENT, "failed to read face indices");
                for (int k = 0; k < 3; k++) {
                    int64_t idx;
                    if (sgn) {
                        int32_t v;
                        std::memcpy(&v, q + 4 * k, 4);
                        idx = v;
                    } else {
                        uint32_t v;
                        std::memcpy(&v, q + 4 * k, 4);
                        idx = v;
                    }
                    if (idx > INT32_MAX) return bail(PRT_ERR_INVALID_ARGUMENT, "face index out of range");
                    d->faces[3 * nread + k] = (int32_t)idx;
                }
                got = true;
            } else if (p.list) {  // skipPropertyBuffered, ply.go:414-436: uchar counts only
                if (p.count_type != "uchar" && p.count_type != "uint8")
                    return bail(PRT_ERR_UNSUPPORTED, "unsupported list count type: " + p.count_type);
                if (!rd.take(1, &q)) return bail(PRT_ERR_INVALID_ARGUMENT, "failed to skip face property");
                if (!known_type(p.item_type)) return bail(PRT_ERR_UNSUPPORTED, "unsupported data type: " + p.item_type);
                if (!rd.take((size_t)q[0] * type_size(p.item_type), &q))
                    return bail(PRT_ERR_INVALID_ARGUMENT, "failed to skip face property");
            } else {
                if (!known_type(p.type)) return bail(PRT_ERR_UNSUPPORTED, "unsupported data type: " + p.type);
                if (!rd.take((size_t)type_size(p.type), &q))
                    return bail(PRT_ERR_INVALID_ARGUMENT, "failed to skip face property");
            }
        }
        if (got) nread++;
    }
    d->num_faces = nread;
    *out = d;
    return PRT_OK;
}

void prt_ply_free(prt_ply_data* data) { free_data(data); }

}  // extern "C"
