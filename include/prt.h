/*
 * prt.h — C ABI of the MI355X (gfx950) path-tracing core.
 *
 * Plain C: POD structs, caller-owned buffers, int status codes, no C++ types or
 * exceptions across the boundary, no retained caller pointers (cgo pointer rules).
 * Every function returns PRT_OK (0) or a PRT_ERR_* code; prt_last_error() gives the
 * thread-local message of the last failure on the calling thread.
 *
 * Reference seams replaced (paths relative to the df07/go-progressive-raytracer tree):
 *   - geometry.Shape.Hit / BVH.Hit            pkg/geometry/interfaces.go:9-12, pkg/geometry/bvh.go:183-239
 *       -> prt_intersect_closest / prt_intersect_any (batched: one call per ray batch)
 *   - scene.Scene + Scene.Preprocess           pkg/scene/scene.go:11-19, 45-78
 *       -> prt_scene_create (flattened shapes, materials, lights; BVH built here)
 *   - renderer.NewProgressiveRaytracer         pkg/renderer/progressive.go:62-96
 *       -> prt_film_create
 *   - ProgressiveRaytracer.getSamplesForPass   pkg/renderer/progressive.go:98-123
 *       -> prt_samples_for_pass
 *   - ProgressiveRaytracer.RenderPass          pkg/renderer/progressive.go:126-223
 *       -> prt_render_pass (+ prt_film_read / prt_film_read_rgba8 for the image)
 *   - integrator.Integrator.RayColor           pkg/integrator/integrator.go:15-19
 *       -> selected by prt_render_params.integrator (path tracing / BDPT), evaluated on device
 *
 * Threading: a prt_scene is immutable and may be shared by threads after creation.
 * A prt_film has one owner; calls on one film must not overlap.
 */
#ifndef PRT_H
#define PRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define PRT_OK 0
#define PRT_ERR_INVALID_ARGUMENT 1
#define PRT_ERR_DEVICE 2          /* HIP runtime error */
#define PRT_ERR_OUT_OF_MEMORY 3
#define PRT_ERR_NO_DEVICE 4       /* no usable GPU */
#define PRT_ERR_UNSUPPORTED 5
#define PRT_ERR_CANCELLED 6

/* ---- rays and hits (intersection level) ---------------------------------- */
/* 32 bytes. Direction need not be unit length (the reference keeps unnormalised
 * metal-fuzz directions, pkg/material/metal.go:44-48). A hit is any t in [tmin, tmax]. */
typedef struct prt_ray {
    float org[3];
    float tmin;
    float dir[3];
    float tmax;
} prt_ray;

/* 16 bytes. prim_id = -1 on a miss (t = +inf). u, v are the barycentric weights of
 * vertex 1 and vertex 2 for triangles (the Moller-Trumbore u, v of
 * pkg/geometry/triangle.go:107-175), the quad's (alpha, beta) for quads
 * (pkg/geometry/quad.go:133-180), 0 for spheres and discs, and for cylinders and cones the
 * part hit in u (0 body, 1 base cap, 2 top cap) with v = 0. */
typedef struct prt_hit {
    float t;
    int32_t prim_id;
    float u;
    float v;
} prt_hit;

/* Buffer location flags for the intersection calls. */
#define PRT_BUFFERS_HOST 0u   /* rays/hits are host pointers; the call is synchronous */
#define PRT_BUFFERS_DEVICE 1u /* rays/hits are device pointers on the scene's device */

/* ---- scene description (all float64, as the reference) --------------------- */
/* Material types: pkg/material/{lambertian,metal,dielectric,emissive,layered}.go */
#define PRT_MAT_LAMBERTIAN 0
#define PRT_MAT_METAL 1
#define PRT_MAT_DIELECTRIC 2
#define PRT_MAT_EMISSIVE 3
#define PRT_MAT_LAYERED 4
/* (type 5 is internal: the disc of a disc spot light, discSpotLightMaterial, disc_spot_light.go:12-46) */
#define PRT_MAT_MIX 6 /* Mix (mix.go:9-56): material2 with probability ratio, else material1 */

typedef struct prt_material {
    int32_t type;
    int32_t outer;      /* layered: index of the outer material; mix: material1 */
    int32_t inner;      /* layered: index of the inner material; mix: material2 */
    int32_t texture;    /* lambertian / metal: 0 = the solid albedo, k > 0 = textures[k - 1]
                           (NewTexturedLambertian lambertian.go:20-22, NewTexturedMetal metal.go:26-35) */
    double albedo[3];   /* lambertian / metal */
    double fuzz;        /* metal: clamped to [0,1] as NewMetal does (metal.go:19-29) */
    double ior;         /* dielectric */
    double emission[3]; /* emissive */
    double ratio;       /* mix: clamped to [0,1] as NewMix does (mix.go:17-27) */
} prt_material;

/* ImageTexture (pkg/material/image_texture.go:8-55): row-major RGB, row 0 at the top (v = 1),
 * sampled nearest-neighbour with wrapped UVs. NewCheckerboardTexture / NewUVDebugTexture /
 * NewGradientTexture (procedural_textures.go) build such images on the host. */
typedef struct prt_texture {
    int32_t width;
    int32_t height;
    const float* rgb; /* width * height * 3 */
} prt_texture;

typedef struct prt_sphere { /* pkg/geometry/sphere.go:11-16; a negative radius flips the normal */
    double center[3];
    double radius;
    int32_t material;
    int32_t pad;
} prt_sphere;

typedef struct prt_quad { /* pkg/geometry/quad.go:98-121 */
    double corner[3];
    double u[3];
    double v[3];
    int32_t material;
    int32_t pad;
} prt_quad;

typedef struct prt_disc { /* pkg/geometry/disc.go:11-43 (normal is normalized on use) */
    double center[3];
    double normal[3];
    double radius;
    int32_t material;
    int32_t pad;
} prt_disc;

typedef struct prt_cylinder { /* NewCylinder, pkg/geometry/cylinder.go:11-39 */
    double base_center[3];
    double top_center[3];
    double radius;
    int32_t capped;    /* non-zero: closed by two discs */
    int32_t material;
} prt_cylinder;

typedef struct prt_cone { /* NewCone, pkg/geometry/cone.go:11-80: base_radius > top_radius >= 0, height > 0 */
    double base_center[3];
    double base_radius;
    double top_center[3];
    double top_radius; /* 0: pointed cone, > 0: frustum */
    int32_t capped;    /* non-zero: base disc, and a top disc for a frustum */
    int32_t material;
} prt_cone;

typedef struct prt_box { /* pkg/geometry/box.go:9-145: size = half extents, rotation X->Y->Z radians */
    double center[3];
    double size[3];
    double rotation[3];
    int32_t material;
    int32_t pad;
} prt_box;

typedef struct prt_mesh { /* pkg/geometry/triangle_mesh.go:33-146; primitive id = face index */
    const double* vertices;        /* num_vertices * 3 */
    int64_t num_vertices;
    const int32_t* faces;          /* num_faces * 3 */
    int64_t num_faces;
    const int32_t* face_materials; /* optional, num_faces entries */
    int32_t material;
    int32_t pad;
    /* TriangleMeshOptions (triangle_mesh.go:20-26), all optional: */
    const double* normals;         /* num_faces * 3 custom triangle normals (normalized on use) */
    const double* rotation;        /* 3 radians, applied X then Y then Z (rotateVertex :170-199) */
    const double* center;          /* 3, rotation centre (default the origin) */
    const double* uvs;             /* num_vertices * 2 per-vertex texture coordinates (VertexUVs) */
} prt_mesh;

/* Light types: pkg/lights/{quad_light,sphere_light,uniform_infinite_light,
 * gradient_infinite_light}.go, plus an equirectangular environment map (no
 * reference counterpart; same Light contract, pkg/lights/interfaces.go:17-43). */
#define PRT_LIGHT_QUAD 0
#define PRT_LIGHT_SPHERE 1
#define PRT_LIGHT_UNIFORM_INFINITE 2
#define PRT_LIGHT_GRADIENT_INFINITE 3
#define PRT_LIGHT_ENVMAP 4
#define PRT_LIGHT_DISC 5       /* pkg/lights/disc_light.go: shape = index into discs[] */
#define PRT_LIGHT_DISC_SPOT 6  /* pkg/lights/disc_spot_light.go (AddSpotLight, scene.go:117-122): from, to,
                                  emission, cone_angle, cone_delta, radius; its disc joins the shapes */
#define PRT_LIGHT_POINT_SPOT 7 /* pkg/lights/point_spot_light.go: from, to, emission, cone_angle, cone_delta */

typedef struct prt_light {
    int32_t type;
    int32_t shape;       /* quad/sphere lights: index into quads[] / spheres[] (the emitting shape);
                            a sphere light with shape -1 has no shape (NewSphereLight without
                            AddSphereLight, as the PBRT "point" light, pbrt_scene.go:348-361): centre
                            `from`, `radius`, `emission` */
    double emission[3];  /* uniform infinite light */
    double top[3];       /* gradient light top colour */
    double bottom[3];    /* gradient light bottom colour */
    const float* env_rgb;/* envmap: env_height rows * env_width * 3 linear RGB, row 0 = +Y pole */
    int32_t env_width;
    int32_t env_height;
    double from[3];      /* spot lights: position */
    double to[3];        /* spot lights: aim point */
    double cone_angle;   /* spot lights: total cone angle, degrees */
    double cone_delta;   /* spot lights: falloff transition angle, degrees */
    double radius;       /* disc spot light: disc radius; shape-less sphere light: sphere radius */
} prt_light;

typedef struct prt_camera { /* pkg/geometry/camera.go:18-32 */
    double center[3];
    double look_at[3];
    double up[3];
    int32_t width;
    int32_t pad;
    double aspect_ratio;
    double vfov;           /* degrees */
    double aperture;       /* 0 = pinhole */
    double focus_distance; /* <= 0: |center - look_at| */
} prt_camera;

typedef struct prt_scene_desc {
    const prt_material* materials; int32_t num_materials;
    const prt_sphere* spheres;     int32_t num_spheres;
    const prt_quad* quads;         int32_t num_quads;
    const prt_box* boxes;          int32_t num_boxes;
    const prt_mesh* meshes;        int32_t num_meshes;
    const prt_light* lights;       int32_t num_lights;
    const double* light_weights;   /* NULL: uniform light sampler (scene.go:62-64) */
    prt_camera camera;
    const prt_disc* discs;         int32_t num_discs;
    const prt_cylinder* cylinders; int32_t num_cylinders;
    const prt_cone* cones;         int32_t num_cones;
    const prt_texture* textures;   int32_t num_textures;
} prt_scene_desc;

/* Global primitive numbering used by prt_hit.prim_id: all mesh triangles (mesh by
 * mesh, face order), then spheres, then quads, then 6 quads per box (box.go:34-89 face order),
 * then discs, cylinders, cones, then the discs of disc spot lights (in light order). */

typedef struct prt_scene_info {
    int64_t num_primitives;
    int64_t num_triangles;
    int64_t bvh_nodes;      /* 4-wide nodes, 128 B each */
    int32_t bvh_depth;      /* levels of 4-wide nodes */
    int32_t max_stack;      /* traversal stack bound */
    double world_center[3]; /* reference BVH root box centre (bvh.go:38-48) */
    double world_radius;
    int64_t device_bytes;
    int32_t trace_variant;  /* traversal kernel variant picked by the autotuner (-1: not tuned yet) */
    int32_t pt_trace_waves; /* resident waves per SIMD of the path tracer's trace kernel (its default
                               occupancy variant unless PRT_PT_TRACE_VARIANT pins another; 0: the scene
                               has rendered no path-tracing chunk yet) */
    int64_t reference_primitives; /* Scene.GetPrimitiveCount (scene.go:81-99): mesh triangles + 1 per
                                     other shape (a box counts once, a disc spot light's disc too) */
    int32_t pt_trace_leaf_gate;   /* leaf gate of the path tracer's traversals: a wave tests leaves once
                                     this many eighths of its running lanes stand at one (0: always); 9:
                                     flat traversal, every primitive tested without the BVH (scenes of at
                                     most 64 primitives). Picked by timing each choice on a sample of the
                                     second bounce's rays inside a render chunk of >= 2^20 paths that is not
                                     a film's one-sample first chunk; a later chunk reads the timings (-1:
                                     not tuned yet) */
    int32_t bdpt_trace_leaf_gate; /* the same for the BDPT kernels, timed on the subpath kernel inside a
                                     BDPT chunk of >= 2^18 paths that is not a film's one-sample first chunk
                                     (-1: not tuned yet). PRT_TRACE_LEAF_GATE=<0-8, 9> pins both choices of
                                     new scenes */
} prt_scene_info;

typedef struct prt_scene prt_scene;
typedef struct prt_film prt_film;

const char* prt_last_error(void);
int prt_version(void);
int prt_device_count(int32_t* count);

/* Builds the BVH (binned SAH, float64 bounds, 4-wide, leaf <= 4) and uploads the scene
 * to `device`. Replaces Scene.Preprocess (scene.go:45-78). */
int prt_scene_create(int32_t device, const prt_scene_desc* desc, prt_scene** out);
int prt_scene_info_get(const prt_scene* scene, prt_scene_info* info);
int prt_scene_destroy(prt_scene* scene);

/* Closest hit for n rays (BVH.Hit with the ray's [tmin, tmax], bvh.go:183-239).
 * stream: hipStream_t or NULL (default stream). Host buffers: synchronous. */
int prt_intersect_closest(const prt_scene* scene, const prt_ray* rays, int64_t n,
                          prt_hit* hits, uint32_t flags, void* stream);
/* Any hit (shadow test): occluded[i] = 1 if some primitive is hit in [tmin, tmax]. */
int prt_intersect_any(const prt_scene* scene, const prt_ray* rays, int64_t n,
                      uint8_t* occluded, uint32_t flags, void* stream);
/* Measurement helper (no reference counterpart): per ray, the number of 4-wide node
 * visits, ray/box slab tests and ray/primitive tests the closest-hit kernel performs. */
int prt_traversal_counts(const prt_scene* scene, const prt_ray* rays, int64_t n,
                         uint32_t* node_visits, uint32_t* box_tests, uint32_t* prim_tests,
                         uint32_t flags, void* stream);

/* Allocation-free closest hit. No counterpart in the reference: for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same hits as prt_intersect_closest with PRT_BUFFERS_DEVICE; rays and hits are device pointers on
 * the scene's device. The work counter and traversal overflow stack live in a static per-device
 * array of PRT_STATIC_SCRATCH_BYTES instead of scene memory (the kernel grid shrinks if the scene's
 * stack bound needs more than that; PRT_ERR_UNSUPPORTED if one work-group's stack does not fit).
 * The call only enqueues work on `stream` (NULL: default stream) and returns: it calls no hipMalloc,
 * hipFree, memset or synchronization and uses the variant the autotuner picked for the scene, if any.
 * Every call on one device shares that device's array: calls that overlap on one device race (order
 * them on one stream). */
#define PRT_STATIC_SCRATCH_BYTES (64u << 20)
int prt_intersect_closest_static(const prt_scene* scene, const prt_ray* rays, int64_t n, prt_hit* hits,
                                 void* stream);

/* ---- progressive rendering (render level) -------------------------------- */
#define PRT_INTEGRATOR_PATH 0 /* pkg/integrator/path_tracing.go */
#define PRT_INTEGRATOR_BDPT 1 /* pkg/integrator/bdpt.go */

typedef struct prt_render_params {
    int32_t integrator;
    int32_t width;                 /* SamplingConfig.Width (scene.go:22-30); > 0 overrides the camera's */
    int32_t height;                /* SamplingConfig.Height; > 0 overrides the camera's (the web server sets
                                      both: camera width and aspect width/height, server.go:237-241) */
    int32_t max_depth;             /* SamplingConfig.MaxDepth */
    int32_t rr_min_bounces;        /* SamplingConfig.RussianRouletteMinBounces */
    int32_t tile_size;             /* ProgressiveConfig.TileSize (64) */
    int32_t initial_samples;       /* ProgressiveConfig.InitialSamples (1) */
    int32_t max_samples_per_pixel; /* ProgressiveConfig.MaxSamplesPerPixel */
    int32_t max_passes;            /* ProgressiveConfig.MaxPasses */
    int32_t pad;
    double adaptive_min_samples;   /* 1.0 = exact spp (adaptive stop off) */
    double adaptive_threshold;
    uint64_t seed;                 /* counter-RNG key */
} prt_render_params;

typedef struct prt_pass_stats { /* RenderStats, pkg/renderer/stats.go:10-17 */
    int64_t total_pixels;
    int64_t total_samples;
    double average_samples;
    int32_t max_samples;
    int32_t min_samples;
    int32_t max_samples_used;
    int32_t target_samples;
    double device_ms;              /* GPU time of the pass (HIP events) */
} prt_pass_stats;

int prt_samples_for_pass(const prt_render_params* params, int32_t pass_number, int32_t* target);
int prt_film_create(const prt_scene* scene, const prt_render_params* params, prt_film** out);
/* Renders every pixel up to the pass target (RenderPass, progressive.go:126-223). */
int prt_render_pass(prt_film* film, int32_t pass_number, prt_pass_stats* stats);
/* Renders `spp` more samples per pixel on all pixels, independent of the pass schedule. */
int prt_render_samples(prt_film* film, int32_t spp, prt_pass_stats* stats);
/* Per-pixel mean colour (count > 0) as width*height*3 floats; optional luminance
 * first/second moment sums and sample counts (PixelStats, stats.go:20-55). */
int prt_film_read(const prt_film* film, float* mean_rgb, float* lum_sum, float* lum_sq_sum,
                  uint32_t* counts);
/* BDPT light-tracing (t = 1) splats accumulated so far, width*height*3 floats (the SplatQueue
 * contents applied by processSplats, progressive.go:370-390; zero for path tracing). The mean
 * colour of prt_film_read already includes them. */
int prt_film_read_splats(const prt_film* film, float* splat_rgb);
/* Gamma-2 RGBA8 image of pixel rectangle [x0,x1)x[y0,y1) (vec3ToColor, progressive.go:482-495). */
int prt_film_read_rgba8(const prt_film* film, int32_t x0, int32_t y0, int32_t x1, int32_t y1,
                        uint8_t* rgba);
int prt_film_reset(prt_film* film);
int prt_film_destroy(prt_film* film);
/* Image size of a film (the params' width/height where set, else the camera's). No counterpart in the
 * reference, whose callers hold the size in SamplingConfig: lets FFI callers size the buffers that
 * prt_film_read and prt_film_read_rgba8 fill. */
int prt_film_size(const prt_film* film, int32_t* width, int32_t* height);

/* Measurement helper (no reference counterpart): a film's kernel-level statistics for profiling and
 * benchmarks. Waits for the film's enqueued work, returns in *stats (may be NULL) what was gathered
 * since the previous call, clears it, and sets what later renders gather:
 *   PRT_PROBE_COUNT  path-tracing traces run a counting variant of the trace kernel (same results,
 *                    slower) that counts rays, 4-wide node visits, ray/box slab tests and ray/primitive
 *                    tests, separately for extension (closest-hit) and shadow (any-hit) rays;
 *   PRT_PROBE_TIME   every kernel launch is bracketed by an event pair and the device time is summed
 *                    per kernel kind.
 * flags = 0 turns both off. */
#define PRT_PROBE_COUNT 1u
#define PRT_PROBE_TIME 2u
#define PRT_PROBE_KIND_CAMERA 0
#define PRT_PROBE_KIND_TRACE 1
#define PRT_PROBE_KIND_SHADE 2
#define PRT_PROBE_KIND_FILM 3
#define PRT_PROBE_KIND_BDPT_PATHS 4
#define PRT_PROBE_KIND_BDPT_WALKS 5
#define PRT_PROBE_KIND_BDPT_CONNECT 6
typedef struct prt_film_probe_stats {
    uint64_t extension_rays, extension_node_visits, extension_box_tests, extension_prim_tests;
    uint64_t shadow_rays, shadow_node_visits, shadow_box_tests, shadow_prim_tests;
    uint64_t trace_launches;
    double camera_ms, trace_ms, shade_ms, film_ms; /* device time per kernel kind (PRT_PROBE_TIME) */
    double bdpt_paths_ms, bdpt_walks_ms, bdpt_connect_ms;
} prt_film_probe_stats;
int prt_film_probe(prt_film* film, uint32_t flags, prt_film_probe_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* PRT_H */
