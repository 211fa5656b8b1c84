// prt_common.h — data layouts and small math shared by the host builder and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PRT_HD __host__ __device__ __forceinline__

namespace prt {

// ---- counter RNG --------------------------------------------------------------------
// Keyed by (seed, pixel, sample); draw position pos = slot*16 + k, slot 0 = camera,
// slot b+1 = bounce b (positions 0: Russian roulette, 1..7: BSDF, 8..10: light sample).
PRT_HD uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}
PRT_HD uint32_t path_key(uint64_t seed, uint32_t pixel, uint32_t sample) {
    uint32_t k = mix32((uint32_t)seed ^ mix32((uint32_t)(seed >> 32) + 0x9E3779B9U));
    k = mix32(k ^ mix32(pixel + 0x632BE5ABU));
    k = mix32(k ^ mix32(sample + 0x85EBCA6BU));
    return k;
}
PRT_HD float draw(uint32_t key, uint32_t pos) {
    uint32_t h = mix32(key ^ mix32(pos * 0x9E3779B9U + 0x7F4A7C15U));
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}
PRT_HD double draw_f64(uint32_t key, uint32_t pos) {
    uint32_t h = mix32(key ^ mix32(pos * 0x9E3779B9U + 0x7F4A7C15U));
    return (double)(h >> 8) * (1.0 / 16777216.0);
}

// ---- BVH layout in HBM ---------------------------------------------------------------
// One 4-wide node = 128 B (one L2 line): child boxes in SoA order, then 4 child words.
// child word: 0xFFFFFFFF = empty slot; bit 31 set = leaf (bits 3..30 first primitive,
// bits 0..2 count 1..4); otherwise the index of the child node.
struct alignas(128) Node4 {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    uint32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(Node4) == 128, "Node4 must be one 128-byte line");

constexpr uint32_t kEmptyChild = 0xFFFFFFFFu;
constexpr uint32_t kLeafBit = 0x80000000u;
PRT_HD bool child_is_leaf(uint32_t c) { return (c & kLeafBit) != 0; }
PRT_HD uint32_t leaf_first(uint32_t c) { return (c & 0x7FFFFFFFu) >> 3; }
PRT_HD uint32_t leaf_count(uint32_t c) { return c & 7u; }
PRT_HD uint32_t make_leaf(uint32_t first, uint32_t count) { return kLeafBit | (first << 3) | count; }

// Primitive record = 3 x float4 (48 B), in BVH leaf order.
//   a = (v0.xyz | corner.xyz | center.xyz, tag)   tag: bits 30..31 kind, bits 0..29 side index
//   b = (v1.xyz | u.xyz      | radius,0,0,  material)
//   c = (v2.xyz | v.xyz      | 0,0,0,       original primitive id)
constexpr uint32_t kPrimTriangle = 0u;
constexpr uint32_t kPrimSphere = 1u;
constexpr uint32_t kPrimQuad = 2u;
constexpr uint32_t kPrimDisc = 3u;  // a = (centre, tag), b = (unit normal, material), c = (radius, r^2, 0, id)
// a = (base centre, tag with side = capped), b = (unit axis, material), c = (radius, height, 0, id)
constexpr uint32_t kPrimCylinder = 4u;
// a = (base centre, tag with side = capped), b = (unit axis, material), c = (base r, top r, height, id)
constexpr uint32_t kPrimCone = 5u;
constexpr uint32_t kKindShift = 29u;
constexpr uint32_t kSideMask = (1u << kKindShift) - 1u;
PRT_HD uint32_t prim_tag(uint32_t kind, uint32_t side) { return (kind << kKindShift) | side; }
PRT_HD uint32_t prim_kind(uint32_t tag) { return tag >> kKindShift; }
PRT_HD uint32_t prim_side(uint32_t tag) { return tag & kSideMask; }

// Quad side data (float32 copies of float64 precomputations, NewQuad quad.go:98-121).
// Quad side table: unit normal and plane offset, and the barycentric coordinates' planes: alpha = A . (p - corner),
// beta = B . (p - corner) with A = v x W, B = W x u, W = n / (n . (u x v)).
struct alignas(64) QuadExt {  // 64 B: a record never straddles two 128-B lines
    float nx, ny, nz, D;
    float ax, ay, az, pad0;
    float bx, by, bz, pad1;
    float pad2[4];
};

// ---- materials and lights (float32 device copies) ------------------------------------
constexpr int32_t kMatSpotEmitter = 5;  // the disc of a disc spot light (emits against spot_dir only)
constexpr int32_t kMatMix = 6;
struct DevMaterial {
    int32_t type, outer, inner;
    int32_t tex;  // lambertian / metal albedo: 0 solid, k > 0 texture k - 1
    float albedo[3];
    float fuzz;
    float ior;
    float emission[3];
    float spot_dir[3];
    float ratio;  // mix: probability of material2 (inner)
};

struct DevTexture {
    int32_t width, height;
    const float* rgb;  // height*width*3, row 0 at the top
};

struct DevLight {
    int32_t type;
    int32_t mat;  // emitting material of quad/sphere lights
    int32_t pad0, pad1;
    // quad light
    float corner[3], u[3], v[3], n[3], W[3];
    float D, area;
    // sphere light
    float center[3], radius;
    // infinite lights; point spot light emission
    float emission[3], top[3], bottom[3];
    // disc lights (centre and radius above; n = unit normal) and spot lights
    float right[3], up[3];
    float from[3], spot_dir[3];
    float cos_total, cos_falloff;
    float from_lo[3];  // from (float64) - from (float32): the position's low part
};

struct DevEnv {
    int32_t width, height;
    float integral;  // mean of func over the unit square
    float pad;
    const float* rgb;      // height*width*3
    const float* func;     // height*width
    const float* row_cdf;  // height+1
    const float* col_cdf;  // height*(width+1)
};

struct DevCamera {
    float center[3], pixel00[3], du[3], dv[3], disk_u[3], disk_v[3];
    float u[3], v[3], w[3], forward[3];
    float lens_radius, lens_area, vw, vh, plane_area, cos_total_width, focus;
    int32_t width, height, has_lens;
};

// Everything a kernel needs to read the scene.
struct DevScene {
    const Node4* nodes;
    const float4* prims;
    const QuadExt* quads;
    const float4* tri_normals;  // custom triangle normals; a triangle's side index k > 0 selects entry k
    const DevMaterial* mats;
    const DevLight* lights;
    const float* light_weights;  // normalised selection probabilities
    const DevTexture* textures;
    const float2* tri_uv;  // per leaf-ordered primitive 3 vertex UVs (x = NaN: none), or null
    DevEnv env;
    DevCamera cam;
    int32_t num_lights;
    int32_t num_infinite;
    int32_t mixed;  // 1 if any non-triangle primitive exists
    int32_t max_stack;
    int32_t num_nodes;
    int32_t textured;  // 1 if any material reads a texture: hits carry UVs
    float world_center[3];
    float world_radius;
    int32_t num_prims;  // primitives in leaf order (the flat traversal tests them all)
};

}  // namespace prt
