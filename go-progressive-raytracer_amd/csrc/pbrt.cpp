// pbrt.cpp — the reference's PBRT-subset scene loader (include/prt_loaders.h).
//
// Parsing follows loaders.ParsePBRT (pkg/loaders/pbrt.go): line-based statements with continuation
// lines, the reference's tokenizer (quoted strings and bracketed arrays as single tokens), the
// AttributeBegin/AttributeEnd block and graphics-state stacks, material indices and area-light
// marking. Conversion follows scene.NewPBRTScene (pkg/scene/pbrt_scene.go): default sampling and
// camera, diffuse/conductor/dielectric materials, sphere/bilinearPatch/trianglemesh/box shapes,
// point/distant/infinite/infinite-gradient lights, area lights from marked shapes. Transform
// statements are parsed and, as in the reference, not applied.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "../../include/prt_loaders.h"
#include "prt_internal.h"

using prt::fail;

namespace {

struct Param {
    std::string type;
    std::vector<std::string> values;
};
struct Stmt {
    std::string type, subtype;
    std::map<std::string, Param> params;
    int material_index = -1;
};
struct Block {
    std::vector<Stmt> materials, shapes, lights, transforms;
};
struct Parsed {
    std::optional<Stmt> camera, film, sampler, integrator;
    bool have_lookat = false;
    double lookat[9] = {};
    std::vector<Stmt> materials, shapes, lights, transforms;
    std::vector<Block> attributes;
};

struct Error {
    std::string msg;
};

std::string trim_space(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}
std::vector<std::string> fields(const std::string& s) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
        size_t j = i;
        while (j < s.size() && !std::isspace((unsigned char)s[j])) j++;
        if (j > i) out.push_back(s.substr(i, j - i));
        i = j;
    }
    return out;
}
bool starts(const std::string& s, const char* p) { return s.compare(0, std::strlen(p), p) == 0; }
bool ends(const std::string& s, const char* p) {
    size_t n = std::strlen(p);
    return s.size() >= n && s.compare(s.size() - n, n, p) == 0;
}
// strings.Trim(s, cutset)
std::string trim_set(const std::string& s, const char* cut) {
    size_t a = 0, b = s.size();
    while (a < b && std::strchr(cut, s[a])) a++;
    while (b > a && std::strchr(cut, s[b - 1])) b--;
    return s.substr(a, b - a);
}
// strconv.ParseFloat on a whole token
bool parse_float(const std::string& s, double& v) {
    if (s.empty() || std::isspace((unsigned char)s[0])) return false;
    errno = 0;
    char* end = nullptr;
    v = std::strtod(s.c_str(), &end);
    return end == s.c_str() + s.size() && errno != ERANGE;
}
// strconv.Atoi with the error ignored: 0 on a syntax error, clamped on overflow
long long atoi_ignoring_error(const std::string& s) {
    size_t i = 0;
    bool neg = false;
    if (i < s.size() && (s[i] == '+' || s[i] == '-')) neg = s[i++] == '-';
    if (i == s.size()) return 0;
    unsigned long long v = 0;
    bool over = false;
    for (; i < s.size(); i++) {
        if (s[i] < '0' || s[i] > '9') return 0;
        if (v > (9223372036854775807ULL - (unsigned)(s[i] - '0')) / 10) over = true;
        else v = v * 10 + (unsigned)(s[i] - '0');
    }
    if (over) return neg ? (-9223372036854775807LL - 1) : 9223372036854775807LL;
    return neg ? -(long long)v : (long long)v;
}

// tokenizePBRT (pbrt.go:513-574): quoted strings and [...] arrays are single tokens
std::vector<std::string> tokenize(const std::string& line) {
    std::vector<std::string> tok;
    std::string cur;
    bool q = false, br = false;
    for (char ch : line) {
        if (ch == '"') {
            cur.push_back(ch);
            if (!br) {
                if (q) {
                    tok.push_back(cur);
                    cur.clear();
                    q = false;
                } else {
                    q = true;
                }
            }
        } else if (ch == '[' && !q) {
            if (!cur.empty()) tok.push_back(cur);
            cur = "[";
            br = true;
        } else if (ch == ']' && !q && br) {
            cur.push_back(ch);
            tok.push_back(cur);
            cur.clear();
            br = false;
        } else if ((ch == ' ' || ch == '\t') && !q && !br) {
            if (!cur.empty()) tok.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(ch);
        }
    }
    if (!cur.empty()) tok.push_back(cur);
    return tok;
}

// parseStatement (pbrt.go:577-660)
Stmt parse_statement(const std::string& line) {
    Stmt st;
    if (starts(line, "LookAt")) {
        st.type = "LookAt";
        st.params["values"] = Param{"float", fields(line.substr(6))};
        return st;
    }
    for (const char* t : {"Translate", "Rotate", "Scale", "Transform"})
        if (starts(line, t)) {
            st.type = t;
            st.params["values"] = Param{"float", fields(line.substr(std::strlen(t)))};
            return st;
        }
    std::vector<std::string> parts = tokenize(line);
    if (parts.size() < 2) throw Error{"invalid statement format"};
    st.type = parts[0];
    size_t i = 1;
    if (starts(parts[1], "\"") && ends(parts[1], "\"")) {
        st.subtype = trim_set(parts[1], "\"");
        i = 2;
    }
    while (i < parts.size()) {
        if (!starts(parts[i], "\"")) {
            i++;
            continue;
        }
        std::vector<std::string> def = fields(trim_set(parts[i], "\""));
        if (def.size() != 2) {
            i++;
            continue;
        }
        i++;
        Param p{def[0], {}};
        if (i < parts.size()) {
            if (starts(parts[i], "[") && ends(parts[i], "]")) p.values = fields(trim_set(parts[i], "[] "));
            else p.values = {parts[i]};
            i++;
        }
        st.params[def[1]] = p;
    }
    return st;
}

bool statement_start(const std::string& line) {  // isStatementStart (pbrt.go:706-722)
    for (const char* t : {"Camera", "Film", "Sampler", "Integrator", "LookAt", "Material", "Shape", "LightSource",
                          "AreaLightSource", "Translate", "Rotate", "Scale", "Transform", "ReverseOrientation",
                          "Attribute"})
        if (line == t || starts(line, (std::string(t) + " ").c_str())) return true;
    return false;
}

struct Parser {  // PBRTParser (pbrt.go:60-438)
    Parsed scene;
    std::vector<Block> blocks;  // attribute stack
    struct State {
        int material_index;
        std::optional<Stmt> area;
    };
    std::vector<State> states;
    int material_index = -1;
    bool in_world = false;
    std::vector<std::string> pending;

    void mark_area_light(Stmt& st) {
        if (states.empty() || !states.back().area) return;
        st.params["_areaLight"] = Param{"bool", {"true"}};
        for (auto& kv : states.back().area->params)
            if (kv.first == "L" || kv.first == "power") st.params[kv.first] = kv.second;
    }
    void route(Stmt st) {
        if (st.type == "LookAt") {
            const auto& v = st.params["values"].values;
            if (st.params.size() != 1 || v.size() != 9) throw Error{"error parsing LookAt: LookAt requires 9 values"};
            for (int k = 0; k < 9; k++)
                if (!parse_float(v[k], scene.lookat[k])) throw Error{"error parsing LookAt: invalid value '" + v[k] + "'"};
            scene.have_lookat = true;
            return;
        }
        const bool transform = st.type == "Translate" || st.type == "Rotate" || st.type == "Scale" || st.type == "Transform";
        if (!blocks.empty()) {
            Block& b = blocks.back();
            if (st.type == "Material") {
                b.materials.push_back(st);
            } else if (st.type == "Shape") {
                const int local = (int)b.materials.size() - 1;
                st.material_index = local >= 0 ? local : material_index;
                mark_area_light(st);
                b.shapes.push_back(st);
            } else if (st.type == "LightSource") {
                b.lights.push_back(st);
            } else if (st.type == "AreaLightSource") {
                if (!states.empty()) states.back().area = st;
                b.lights.push_back(st);
            } else if (transform) {
                b.transforms.push_back(st);
            }
        } else if (!in_world) {
            if (st.type == "Camera") scene.camera = st;
            else if (st.type == "Film") scene.film = st;
            else if (st.type == "Sampler") scene.sampler = st;
            else if (st.type == "Integrator") scene.integrator = st;
        } else {
            if (st.type == "Material") {
                scene.materials.push_back(st);
                material_index = (int)scene.materials.size() - 1;
            } else if (st.type == "Shape") {
                st.material_index = material_index;
                mark_area_light(st);
                scene.shapes.push_back(st);
            } else if (st.type == "LightSource") {
                scene.lights.push_back(st);
            } else if (st.type == "AreaLightSource") {
                if (!states.empty()) states.back().area = st;
                scene.lights.push_back(st);
            } else if (transform) {
                scene.transforms.push_back(st);
            }
        }
    }
    void flush(const char* context) {
        if (pending.empty()) return;
        std::string full;
        for (size_t k = 0; k < pending.size(); k++) full += (k ? " " : "") + pending[k];
        try {
            route(parse_statement(full));
        } catch (Error& e) {
            if (starts(e.msg, "error parsing LookAt")) throw;
            throw Error{std::string("error parsing statement ") + context + " '" + full + "': " + e.msg};
        }
        pending.clear();
    }
    void line(const std::string& raw) {
        const std::string l = trim_space(raw);
        if (l.empty() || l[0] == '#') return;
        if (l == "WorldBegin") {
            flush("before WorldBegin");
            in_world = true;
        } else if (l == "WorldEnd") {
            flush("before WorldEnd");
            in_world = false;
        } else if (l == "AttributeBegin") {
            flush("before AttributeBegin");
            State s{material_index, std::nullopt};
            if (!states.empty()) s.area = states.back().area;
            states.push_back(s);
            blocks.emplace_back();
        } else if (l == "AttributeEnd") {
            flush("before AttributeEnd");
            if (!blocks.empty()) {
                scene.attributes.push_back(std::move(blocks.back()));
                blocks.pop_back();
            }
            if (!states.empty()) {
                material_index = states.back().material_index;
                states.pop_back();
            }
        } else if (statement_start(l)) {
            flush("");
            pending = {l};
        } else {
            if (pending.empty()) throw Error{"unexpected continuation line: " + l};
            pending.push_back(l);
        }
    }
};

// ParsePBRT: bufio.Scanner lines (64 KiB limit, "\r\n" endings accepted)
Parsed parse(const char* text, size_t n) {
    Parser p;
    size_t i = 0;
    while (i < n) {
        size_t j = i;
        while (j < n && text[j] != '\n') j++;
        size_t e = j;
        if (e > i && text[e - 1] == '\r') e--;
        if (j - i >= 64 * 1024) throw Error{"error reading input: bufio.Scanner: token too long"};
        p.line(std::string(text + i, e - i));
        i = j + 1;
    }
    p.flush("at end of file");
    return std::move(p.scene);
}

// ---- NewPBRTScene (pbrt_scene.go) ----------------------------------------------------------------
bool get_float(const Stmt& s, const char* name, double& v) {
    auto it = s.params.find(name);
    return it != s.params.end() && !it->second.values.empty() && parse_float(it->second.values[0], v);
}
bool get3(const Stmt& s, const char* name, double v[3]) {
    auto it = s.params.find(name);
    if (it == s.params.end() || it->second.values.size() < 3) return false;
    double t[3];
    for (int k = 0; k < 3; k++)
        if (!parse_float(it->second.values[k], t[k])) return false;
    std::copy(t, t + 3, v);
    return true;
}
bool is_area_light(const Stmt& s) {
    auto it = s.params.find("_areaLight");
    return it != s.params.end() && !it->second.values.empty() && it->second.values[0] == "true";
}
void set3(double* d, double x, double y, double z) {
    d[0] = x;
    d[1] = y;
    d[2] = z;
}

struct Built {  // owns every array the descriptor points into
    prt_pbrt_scene out{};
    std::vector<prt_material> mats;
    std::vector<prt_sphere> spheres;
    std::vector<prt_quad> quads;
    std::vector<prt_box> boxes;
    std::vector<prt_light> lights;
    std::vector<prt_mesh> meshes;
    std::vector<std::unique_ptr<std::vector<double>>> verts;
    std::vector<std::unique_ptr<std::vector<int32_t>>> faces;
    int64_t shapes = 0;

    int material(const Stmt& s) {  // convertMaterial (pbrt_scene.go:158-201)
        prt_material m{};
        if (s.subtype == "diffuse") {
            m.type = PRT_MAT_LAMBERTIAN;
            if (!get3(s, "reflectance", m.albedo)) set3(m.albedo, 0.7, 0.7, 0.7);
        } else if (s.subtype == "conductor") {
            m.type = PRT_MAT_METAL;
            if (!get3(s, "eta", m.albedo)) set3(m.albedo, 0.7, 0.6, 0.5);
            double r;
            if (get_float(s, "roughness", r)) {
                if (r < 0 || r > 1) throw Error{"invalid metal roughness: must be between 0 and 1"};
                m.fuzz = r;
            }
        } else if (s.subtype == "dielectric") {
            m.type = PRT_MAT_DIELECTRIC;
            m.ior = 1.5;
            double e;
            if (get_float(s, "eta", e)) {
                if (e <= 0) throw Error{"invalid dielectric IOR: must be positive"};
                m.ior = e;
            }
        } else {
            throw Error{"unsupported material type: " + s.subtype};
        }
        m.outer = m.inner = -1;
        mats.push_back(m);
        return (int)mats.size() - 1;
    }
    int emissive(const double L[3]) {
        prt_material m{};
        m.type = PRT_MAT_EMISSIVE;
        m.outer = m.inner = -1;
        std::copy(L, L + 3, m.emission);
        mats.push_back(m);
        return (int)mats.size() - 1;
    }
    // convertShape (pbrt_scene.go:204-305); returns the kind added ('s' sphere, 'q' quad, 'm', 'b')
    char shape(const Stmt& s, int mat) {
        if (s.subtype == "sphere") {
            prt_sphere sp{};
            sp.radius = 1.0;
            double r;
            if (get_float(s, "radius", r)) {
                if (r <= 0) throw Error{"invalid sphere radius: must be positive"};
                sp.radius = r;
            }
            if (!get3(s, "center", sp.center)) set3(sp.center, 0, 0, 0);
            sp.material = mat;
            spheres.push_back(sp);
            shapes++;
            return 's';
        }
        if (s.subtype == "bilinearPatch") {
            double p00[3], p01[3], p10[3], p11[3];
            if (!get3(s, "P00", p00) || !get3(s, "P01", p01) || !get3(s, "P10", p10) || !get3(s, "P11", p11))
                throw Error{"bilinearPatch missing corner points"};
            prt_quad q{};
            for (int a = 0; a < 3; a++) {
                q.corner[a] = p00[a];
                q.u[a] = p01[a] - p00[a];
                q.v[a] = p10[a] - p00[a];
            }
            q.material = mat;
            quads.push_back(q);
            shapes++;
            return 'q';
        }
        if (s.subtype == "trianglemesh") {
            auto P = s.params.find("P");
            if (P == s.params.end() || P->second.values.size() % 3 != 0)
                throw Error{"trianglemesh missing or invalid vertices"};
            auto v = std::make_unique<std::vector<double>>();
            for (const auto& x : P->second.values) {
                double d;
                if (!parse_float(x, d)) throw Error{"invalid vertex coordinate '" + x + "'"};
                v->push_back(d);
            }
            auto I = s.params.find("indices");
            if (I == s.params.end() || I->second.values.size() % 3 != 0)
                throw Error{"trianglemesh missing or invalid indices"};
            auto f = std::make_unique<std::vector<int32_t>>();
            const int64_t nv = (int64_t)v->size() / 3;
            for (const auto& x : I->second.values) {
                long long k = atoi_ignoring_error(x);
                if (k < 0 || k >= nv) throw Error{"trianglemesh index out of range"};  // the reference panics
                f->push_back((int32_t)k);
            }
            prt_mesh m{};
            m.vertices = v->data();
            m.num_vertices = nv;
            m.faces = f->data();
            m.num_faces = (int64_t)f->size() / 3;
            m.material = mat;
            meshes.push_back(m);
            verts.push_back(std::move(v));
            faces.push_back(std::move(f));
            shapes++;
            return 'm';
        }
        if (s.subtype == "box") {
            prt_box b{};
            if (!get3(s, "center", b.center)) set3(b.center, 0, 0, 0);
            if (!get3(s, "size", b.size)) set3(b.size, 1, 1, 1);
            if (!get3(s, "rotation", b.rotation)) set3(b.rotation, 0, 0, 0);
            b.material = mat;
            boxes.push_back(b);
            shapes++;
            return 'b';
        }
        throw Error{"unsupported shape type: " + s.subtype};
    }
    // convertLight (pbrt_scene.go:337-407)
    void light(const Stmt& s) {
        prt_light L{};
        L.shape = -1;
        if (s.subtype == "point") {  // a 0.1-radius sphere light that is not a shape
            L.type = PRT_LIGHT_SPHERE;
            if (!get3(s, "I", L.emission)) set3(L.emission, 10, 10, 10);
            if (!get3(s, "from", L.from)) set3(L.from, 0, 5, 0);
            L.radius = 0.1;
        } else if (s.subtype == "distant") {  // direction ignored: a uniform infinite light
            L.type = PRT_LIGHT_UNIFORM_INFINITE;
            if (!get3(s, "L", L.emission)) set3(L.emission, 3, 3, 3);
        } else if (s.subtype == "infinite") {
            L.type = PRT_LIGHT_UNIFORM_INFINITE;
            if (!get3(s, "L", L.emission)) set3(L.emission, 1, 1, 1);
        } else if (s.subtype == "infinite-gradient") {
            L.type = PRT_LIGHT_GRADIENT_INFINITE;
            if (!get3(s, "topColor", L.top)) set3(L.top, 0.5, 0.7, 1.0);
            if (!get3(s, "bottomColor", L.bottom)) set3(L.bottom, 1.0, 1.0, 1.0);
        } else if (s.subtype == "diffuse") {
            throw Error{"AreaLightSource should be handled as graphics state, not converted as standalone light"};
        } else {
            throw Error{"unsupported light type: " + s.subtype};
        }
        lights.push_back(L);
    }
};

void convert(const Parsed& ps, Built& B) {
    prt_pbrt_scene& o = B.out;
    // createDefaultPBRTSamplingConfig (pbrt_scene.go:77-88)
    o.width = 400;
    o.height = 400;
    o.samples_per_pixel = 150;
    o.max_depth = 5;
    o.rr_min_bounces = 16;
    o.adaptive_min_samples = 0.20;
    o.adaptive_threshold = 0.005;
    // convertCamera (:91-155)
    prt_camera& c = o.desc.camera;
    set3(c.center, 0, 0, 0);
    set3(c.look_at, 0, 0, -1);
    set3(c.up, 0, 1, 0);
    c.width = 400;
    c.aspect_ratio = 1.0;
    c.vfov = 90.0;
    c.aperture = 0.0;
    c.focus_distance = 1.0;
    if (ps.have_lookat) {
        std::copy(ps.lookat, ps.lookat + 3, c.center);
        std::copy(ps.lookat + 3, ps.lookat + 6, c.look_at);
        std::copy(ps.lookat + 6, ps.lookat + 9, c.up);
    }
    double v;
    if (ps.camera && ps.camera->subtype == "perspective" && get_float(*ps.camera, "fov", v)) {
        if (v <= 0 || v >= 180) throw Error{"failed to convert camera: invalid camera FOV: must be between 0 and 180 degrees"};
        c.vfov = v;
    }
    if (ps.film) {
        if (get_float(*ps.film, "xresolution", v)) {
            if (v <= 0 || v > 8192) throw Error{"failed to convert camera: invalid image width: must be between 1 and 8192"};
            c.width = (int32_t)v;
            o.width = (int32_t)v;
        }
        if (get_float(*ps.film, "yresolution", v)) {
            if (v <= 0 || v > 8192) throw Error{"failed to convert camera: invalid image height: must be between 1 and 8192"};
            o.height = (int32_t)v;
            c.aspect_ratio = (double)c.width / v;
        }
    }
    // materials, then the top-level shapes with their material index
    std::vector<int> global;
    for (const Stmt& m : ps.materials) global.push_back(B.material(m));
    for (const Stmt& s : ps.shapes) {
        if (s.material_index < 0 || s.material_index >= (int)global.size())
            throw Error{"shape has no valid material (MaterialIndex: " + std::to_string(s.material_index) + ")"};
        B.shape(s, global[s.material_index]);
    }
    for (const Stmt& l : ps.lights) B.light(l);
    // processAttributeBlock (:410-468), blocks in the order they closed
    for (const Block& blk : ps.attributes) {
        std::vector<int> local;
        for (const Stmt& m : blk.materials) local.push_back(B.material(m));
        std::vector<std::pair<char, int>> added;  // kind and index of each block shape
        for (const Stmt& s : blk.shapes) {
            int mat;
            if (s.material_index >= 0 && s.material_index < (int)local.size()) mat = local[s.material_index];
            else if (s.material_index >= 0 && s.material_index < (int)global.size()) mat = global[s.material_index];
            else throw Error{"shape has no valid material (MaterialIndex: " + std::to_string(s.material_index) + ")"};
            double L[3];
            if (is_area_light(s) && get3(s, "L", L)) mat = B.emissive(L);
            const char kind = B.shape(s, mat);
            int idx = kind == 's' ? (int)B.spheres.size() - 1 : kind == 'q' ? (int)B.quads.size() - 1 : -1;
            added.push_back({kind, idx});
        }
        for (size_t k = 0; k < blk.shapes.size(); k++) {  // convertAreaLight (:308-334)
            const Stmt& s = blk.shapes[k];
            if (!is_area_light(s)) continue;
            double L[3];
            if (!get3(s, "L", L)) throw Error{"failed to convert area light: area light missing emission parameter 'L'"};
            prt_light l{};
            if (added[k].first == 'q') l.type = PRT_LIGHT_QUAD;
            else if (added[k].first == 's') l.type = PRT_LIGHT_SPHERE;
            else throw Error{"failed to convert area light: shape type is not supported as an area light"};
            l.shape = added[k].second;
            B.lights.push_back(l);
        }
        for (const Stmt& l : blk.lights)
            if (l.type != "AreaLightSource") B.light(l);
    }
    prt_scene_desc& d = o.desc;
    d.materials = B.mats.data();
    d.num_materials = (int32_t)B.mats.size();
    d.spheres = B.spheres.data();
    d.num_spheres = (int32_t)B.spheres.size();
    d.quads = B.quads.data();
    d.num_quads = (int32_t)B.quads.size();
    d.boxes = B.boxes.data();
    d.num_boxes = (int32_t)B.boxes.size();
    d.meshes = B.meshes.data();
    d.num_meshes = (int32_t)B.meshes.size();
    d.lights = B.lights.data();
    d.num_lights = (int32_t)B.lights.size();
    o.reference_shapes = B.shapes;
}

// validateFilePath (pbrt.go:440-480): a .pbrt file under a scenes/ directory or the temp directory
std::string clean_path(const std::string& p) {  // filepath.Clean, lexically
    const bool abs = !p.empty() && p[0] == '/';
    std::vector<std::string> out;
    size_t i = 0;
    while (i <= p.size()) {
        size_t j = p.find('/', i);
        if (j == std::string::npos) j = p.size();
        std::string seg = p.substr(i, j - i);
        if (seg.empty() || seg == ".") {
        } else if (seg == "..") {
            if (!out.empty() && out.back() != "..") out.pop_back();
            else if (!abs) out.push_back("..");
        } else {
            out.push_back(seg);
        }
        i = j + 1;
    }
    std::string r = abs ? "/" : "";
    for (size_t k = 0; k < out.size(); k++) r += (k ? "/" : "") + out[k];
    return r.empty() ? "." : r;
}
void validate_path(const std::string& path) {
    if (path.empty()) throw Error{"filename cannot be empty"};
    const std::string c = clean_path(path);
    const char* td = std::getenv("TMPDIR");
    const std::string tmp = td && *td ? td : "/tmp";
    const bool in_scenes = starts(c, "scenes/") || c.find("scenes/") != std::string::npos;
    if (!in_scenes && !starts(c, tmp.c_str())) throw Error{"file path must be in scenes/ directory"};
    if (c.find("..") != std::string::npos && c.find("scenes/") == std::string::npos)
        throw Error{"invalid file path: directory traversal not allowed"};
    std::string low = c;
    for (char& ch : low) ch = (char)std::tolower((unsigned char)ch);
    if (!ends(low, ".pbrt")) throw Error{"invalid file type: only .pbrt files are allowed"};
    if (c.size() > 512) throw Error{"file path too long: maximum 512 characters allowed"};
}

int build(const char* text, size_t n, prt_pbrt_scene** out) {
    auto* B = new Built();
    try {
        Parsed ps = parse(text, n);
        convert(ps, *B);
    } catch (Error& e) {
        delete B;
        return fail(PRT_ERR_INVALID_ARGUMENT, e.msg.c_str());
    }
    B->out.impl = B;
    *out = &B->out;
    return PRT_OK;
}

}  // namespace

extern "C" {

int prt_pbrt_parse(const char* text, int64_t length, prt_pbrt_scene** out) {
    if (!text || !out || length < 0) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    return build(text, (size_t)length, out);
}

int prt_pbrt_load(const char* path, prt_pbrt_scene** out) {
    if (!path || !out) return fail(PRT_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    try {
        validate_path(path);
    } catch (Error& e) {
        return fail(PRT_ERR_INVALID_ARGUMENT, e.msg.c_str());
    }
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(PRT_ERR_INVALID_ARGUMENT, (std::string("failed to open PBRT file: ") + path).c_str());
    std::string text;
    char buf[65536];
    for (size_t n; (n = std::fread(buf, 1, sizeof(buf), f)) > 0;) text.append(buf, n);
    std::fclose(f);
    return build(text.data(), text.size(), out);
}

void prt_pbrt_free(prt_pbrt_scene* s) {
    if (s) delete static_cast<Built*>(s->impl);
}

}  // extern "C"
