/*
 * bvh_render.h — C entry point of the renderer (libbvhhost.so).
 *
 * renderImage of include/render_engine.h (the reference's API,
 * src/render_engine.cu:711-853) for callers that cannot pass C++ vectors:
 * one primary ray per pixel from the pinhole camera, closest hit on the GPU
 * through bvh_trace, shading on the host as the reference defines it, a
 * binary PPM written to `path` (when not NULL).
 *   nodes   num_nodes 48-byte BVHNode records (host memory)
 *   soa     nine host arrays of ntri floats (v0x v0y v0z v1x v1y v1z v2x v2y v2z)
 *   camera  10 floats: eye xyz, look-at xyz, up xyz, vertical fov (degrees)
 *   mode    0 normal, 1 depth, 2 diffuse, 3 heatmap (ShadingMode)
 *   stats   optional, 5 floats: render ms, mean nodes visited, mean AABB
 *           tests, mean triangle tests (over rays that visited a node), max
 *           nodes visited
 */
#ifndef BVH_RENDER_H
#define BVH_RENDER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int bvh_render_ppm(const void* nodes, uint32_t num_nodes, const float* const* soa, uint32_t ntri, int32_t width,
                   int32_t height, const float* camera, int32_t mode, float fixed_heatmap_max, const char* path,
                   float* stats);

#ifdef __cplusplus
}
#endif
#endif /* BVH_RENDER_H */
