/*
 * bvh_obj.h — OBJ ingestion (libbvhhost.so, host C++ over the C ABI).
 *
 * Replaces loadOBJ of the reference (src/loader.cpp:10-74) with the same
 * result, bit for bit: "v" lines give vertices (coordinates read as
 * std::istream reads floats: a malformed coordinate is 0 and the rest of its
 * line stays 0), "f" lines give fans (v0, v1, v2), (v0, v2, v3), ... over
 * the index before the first '/' of each token (negative = relative to the
 * vertices read so far, positive = 1-based), a triangle naming a vertex
 * outside [0, vertices read so far) is skipped, every other line is ignored.
 * A face token with no leading integer is an error (the reference throws).
 * Parsing runs on several host threads (the file is split at line ends;
 * vertex and triangle counts are combined by prefix sums), and
 * bvh_mesh_upload_obj writes the triangles straight into pinned staging
 * memory that is copied to the device by DMA (no pageable bounce copy).
 */
#ifndef BVH_OBJ_H
#define BVH_OBJ_H

#include <stddef.h>
#include <stdint.h>

#include "bvh_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bvh_obj_rec* bvh_obj_t;

/* threads <= 0: one per hardware thread (at most 64).  Errors: BVH_E_ARG
 * (cannot open the file, or a face token without an index), BVH_E_NOMEM,
 * BVH_E_UNSUPPORTED (more than 2^30 triangles). */
int bvh_obj_parse(const char* path, int32_t threads, bvh_obj_t* out);
int bvh_obj_info(bvh_obj_t obj, uint32_t* triangles, uint64_t* vertices);
/* Writes the triangles into nine caller arrays (v0x v0y v0z v1x v1y v1z v2x v2y v2z). */
int bvh_obj_fill(bvh_obj_t obj, float* const* soa);
/* The message of the last failed bvh_obj_parse on this thread. */
const char* bvh_obj_last_error(void);
int bvh_obj_free(bvh_obj_t obj);

/* Parse `path` and upload it as a resident mesh (pinned staging + DMA). */
int bvh_mesh_upload_obj(bvh_ctx_t ctx, const char* path, int32_t threads, bvh_mesh_t* out);

#ifdef __cplusplus
}
#endif
#endif /* BVH_OBJ_H */
