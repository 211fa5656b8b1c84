/*
 * prt_workload.h — deterministic synthetic inputs of the benchmark configurations.
 *
 * No reference counterpart: the reference loads meshes from PLY files that are not in its
 * tree (pkg/scene/dragon.go:118-141, .gitignore /models) and has no ray microbenchmark or
 * HDR sky. These generators define those inputs; every output is a pure function of the
 * arguments (integer hashing, IEEE sqrt and float64 libm calls, then rounding to float32).
 */
#ifndef PRT_WORKLOAD_H
#define PRT_WORKLOAD_H

#include <stdint.h>

#include "prt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Geodesic icosphere of the given frequency (20*f^2 triangles, 10*f^2+2 vertices), vertex
 * directions displaced radially by r = radius * (1 + amplitude * fbm(dir)), fbm = 3 octaves of
 * seeded value noise (base frequency 4, lacunarity 2, gain 0.5, normalised to [-1,1]).
 * Frequency 256 = an icosahedron subdivided 8 times (1,310,720 triangles); 300 = 1,800,000.
 * Vertices are float64 values exactly representable in float32. Call with NULL buffers to
 * get the counts. */
int prt_workload_icosphere(int32_t frequency, double radius, const double* center, double amplitude,
                           uint64_t seed, double* vertices, int64_t* num_vertices, int32_t* faces,
                           int64_t* num_faces);

/* n rays: origins uniform in the box [lo,hi] scaled by `scale` about its centre, directions
 * uniform on the sphere (z = 1-2u, phi = 2*pi*v, pkg/core/sampling.go:97-104), tmin = 0.001,
 * tmax = +inf. Ray i depends only on (seed, i). */
int prt_workload_rays(const double* lo, const double* hi, double scale, int64_t n, uint64_t seed,
                      prt_ray* rays);

/* Equirectangular HDR sky, width*height*3 floats, row 0 = +Y pole: log-normal sky radiance
 * (value-noise log-luminance), dim ground below the horizon, and a sun lobe whose peak texel
 * holds exactly `sun_peak` in the green channel. */
int prt_workload_sky(int32_t width, int32_t height, uint64_t seed, double sun_peak, float* rgb);

#ifdef __cplusplus
}
#endif
#endif /* PRT_WORKLOAD_H */
