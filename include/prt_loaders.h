/* prt_loaders.h — mesh file loading for scenes handed to prt_scene_create.
 *
 * Replaces loaders.LoadPLY (pkg/loaders/ply.go:64-110): binary little-endian PLY with float32
 * vertex positions, optional float32 normals (nx, ny, nz), uchar colours (red/r, green/g, blue/b),
 * float32 texture coordinates (u/s/texture_u, v/t/texture_v), and triangular faces given as a
 * `vertex_indices` list (uchar or int count, int or uint indices). As in the reference, ASCII and
 * big-endian files are rejected, and vertex properties of other types do not set positions.
 * The arrays are owned by the returned object; free it with prt_ply_free.
 */
#ifndef PRT_LOADERS_H
#define PRT_LOADERS_H

#include <stdint.h>

#include "prt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct prt_ply_data { /* PLYData, pkg/loaders/ply.go:54-72 */
    int64_t num_vertices;
    int64_t num_faces;
    double* vertices;  /* num_vertices * 3 */
    int32_t* faces;    /* num_faces * 3 */
    double* normals;   /* num_vertices * 3, NULL without nx/ny/nz */
    double* colors;    /* num_vertices * 3 in [0, 1], NULL without colours */
    double* texcoords; /* num_vertices * 2, NULL without texture coordinates */
} prt_ply_data;

/* Loads a PLY file (LoadPLY, pkg/loaders/ply.go:64-110). */
int prt_ply_load(const char* path, prt_ply_data** out);
void prt_ply_free(prt_ply_data* data);

/* A PBRT-subset scene converted as scene.NewPBRTScene does (pkg/scene/pbrt_scene.go:15-74): `desc`
 * is ready for prt_scene_create (its arrays are owned by this object), the sampling settings are the
 * reference's PBRT defaults with the Film resolution. Parsing follows loaders.ParsePBRT
 * (pkg/loaders/pbrt.go): statements and continuation lines, AttributeBegin/End blocks with their
 * material and area-light state, LookAt/Camera/Film; materials diffuse, conductor, dielectric;
 * shapes sphere, bilinearPatch, trianglemesh, box; lights point (a 0.1-radius sphere light that is
 * not a shape), distant and infinite (uniform), infinite-gradient, and area lights on bilinearPatch
 * and sphere shapes. Translate/Rotate/Scale/Transform are parsed and ignored, as in the reference.
 * Errors (unsupported types, invalid values, a shape without a material) fail with a message. */
typedef struct prt_pbrt_scene {
    prt_scene_desc desc;
    int32_t width, height;              /* SamplingConfig.Width / Height */
    int32_t samples_per_pixel;          /* 150 (the Sampler statement is not read) */
    int32_t max_depth;                  /* 5 */
    int32_t rr_min_bounces;             /* 16 */
    int32_t pad;
    double adaptive_min_samples;        /* 0.20 */
    double adaptive_threshold;          /* 0.005 */
    int64_t reference_shapes;           /* len(Scene.Shapes) */
    void* impl;
} prt_pbrt_scene;

/* LoadPBRT (pbrt.go:97-110): the path must name a .pbrt file under a scenes/ directory or the
 * temporary directory (validateFilePath :440-480). */
int prt_pbrt_load(const char* path, prt_pbrt_scene** out);
/* ParsePBRT (pbrt.go:72-95) on text in memory. */
int prt_pbrt_parse(const char* text, int64_t length, prt_pbrt_scene** out);
void prt_pbrt_free(prt_pbrt_scene* scene);

#ifdef __cplusplus
}
#endif

#endif /* PRT_LOADERS_H */
