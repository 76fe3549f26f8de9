/*
 * bvh_scenes.h — seeded synthetic workloads (C ABI, host only, libbvhhost.so).
 *
 * The reference generates random meshes with srand(time) (src/loader.cpp:77-98),
 * which cannot be reproduced; these generators replace it for benchmarking and
 * parity with fixed seeds.  RNG: xoshiro256** seeded through SplitMix64;
 * u01 = (x >> 40) * 2^-24 (an exact float in [0,1)).  All arrays are SoA:
 * soa[0..8] = v0x v0y v0z v1x v1y v1z v2x v2y v2z.  Rays are 6 floats each:
 * origin xyz, direction xyz.
 */
#ifndef BVH_SCENES_H
#define BVH_SCENES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The first n raw 64-bit outputs of the generators' RNG for `seed` (xoshiro256**
 * whose four state words are the first four SplitMix64 outputs of `seed`). */
int bvh_rng_u64(uint64_t seed, uint32_t n, uint64_t* out);

/* Uniform soup: centroid ~ U[0,1]^3, each vertex = centroid + U[-0.005,0.005]^3. */
int bvh_scene_uniform(uint64_t seed, uint32_t n, float* const* soa);

/* Blobs: 95% of triangles in 64 isotropic Gaussian blobs (centres U[0,1]^3,
 * sigma 0.02, vertex offsets U[-0.003,0.003]^3), 5% large triangles of edge
 * about 0.1 uniform in the cube, then a Fisher-Yates shuffle. */
int bvh_scene_blobs(uint64_t seed, uint32_t n, float* const* soa);

/* Random-walk strands: `strands` walks of `steps` segments of length 0.01
 * from U[0,1]^3, directions uniform on the sphere, reflected at the cube
 * walls; one thin triangle (width 0.0005) per segment. n = strands * steps. */
int bvh_scene_strands(uint64_t seed, uint32_t strands, uint32_t steps, float* const* soa);

/* Batch of `meshes` triangle counts, log-uniform in [lo, hi]. */
int bvh_scene_batch_counts(uint64_t seed, uint32_t meshes, uint32_t lo, uint32_t hi, uint32_t* counts);
/* Mesh i of a batch: uniform soup with seed (seed ^ i), scaled by U[0.1,10]
 * and offset by U[-100,100]^3 (scale/offset drawn from the batch seed). */
int bvh_scene_batch_mesh(uint64_t seed, uint32_t index, uint32_t n, float* const* soa);

/* Pinhole camera rays (one per pixel, row-major), following generateRay of
 * the reference (src/render_engine.cu:155-188) evaluated on the host. */
int bvh_rays_camera(const float* eye, const float* look_at, const float* up, float fov_y_deg, uint32_t width,
                    uint32_t height, float* rays);
/* Camera on the bounding sphere of box[6] (min xyz, max xyz): eye = centre +
 * (|diag|/2) * normalize(0.6, 0.3, 0.7), looking at the centre, up (0,1,0). */
int bvh_rays_sphere_camera(const float* box, float fov_y_deg, uint32_t width, uint32_t height, float* rays);
/* Origins uniform on the bounding sphere of box, directions toward a uniform
 * point inside the box. */
int bvh_rays_sphere_to_box(uint64_t seed, const float* box, uint32_t n, float* rays);
/* Origins uniform in [0,1]^3, directions uniform on the sphere. */
int bvh_rays_uniform_cube(uint64_t seed, uint32_t n, float* rays);

/* FNV-1a 64 over a byte range, continuing from h (start with 14695981039346656037). */
uint64_t bvh_fnv1a64(const void* data, size_t bytes, uint64_t h);

#ifdef __cplusplus
}
#endif
#endif /* BVH_SCENES_H */
