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

#ifdef __cplusplus
}
#endif

#endif /* PRT_LOADERS_H */
