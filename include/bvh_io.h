/*
 * bvh_io.h — persisting trees (libbvhhost.so, host C++ over the C ABI).
 *
 *   bvh_tree_export_viz   the reference's exportBVHToBinary (src/main.cpp:24-29
 *                         VizNode, :43-64): 32 bytes per node (min xyz, max xyz,
 *                         left, right), no header, node order as built
 *   bvh_tree_export_obj   the reference's exportBVHToOBJ (src/main.cpp:66-99):
 *                         each box as 8 vertices and 6 polylines; leaves_only
 *                         selects nodes by their leaf bit (the reference's index
 *                         test assumes leaves last and is wrong for PLOC trees)
 *   bvh_tree_save/load    a resident tree to/from a file, so a server can keep
 *                         BVHs across restarts (no reference counterpart)
 *   bvh_tree_file_*       the same file format on host arrays (no device)
 *
 * Tree file: "BVHAMD01" (8 bytes), uint32 version (1), uint32 num_nodes,
 * uint32 num_prims, uint32 reserved (0), then num_nodes 48-byte BVHNode
 * records, then num_prims uint32 primitive indices (little-endian).
 */
#ifndef BVH_IO_H
#define BVH_IO_H

#include <stddef.h>
#include <stdint.h>

#include "bvh_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host arrays <-> file.  file_read with nodes == NULL only reports the sizes. */
int bvh_tree_file_write(const char* path, const void* nodes, uint32_t num_nodes, const uint32_t* indices,
                        uint32_t num_prims);
int bvh_tree_file_read(const char* path, void* nodes, uint32_t* num_nodes, uint32_t* indices, uint32_t* num_prims);

/* Structure check of a host node array before it is uploaded: num_nodes =
 * 2 num_prims - 1 (when num_prims > 0), internal children in range and not
 * self-references, leaf primitives and indices below num_prims, n leaves and
 * n - 1 internal nodes.  BVH_E_STATE when the array is not such a tree. */
int bvh_tree_validate(const void* nodes, uint32_t num_nodes, const uint32_t* indices, uint32_t num_prims);

/* Device trees <-> file.  bvh_tree_load rejects a file whose header does not
 * match its size or whose tree fails bvh_tree_validate (BVH_E_STATE) before
 * anything is uploaded; host allocation failure gives BVH_E_NOMEM. */
int bvh_tree_save(bvh_ctx_t ctx, bvh_tree_t tree, const char* path);
int bvh_tree_load(bvh_ctx_t ctx, const char* path, bvh_tree_t tree);

/* OBJ wireframe of a host node array or of a device tree. */
int bvh_nodes_export_obj(const void* nodes, uint32_t num_nodes, const char* path, int32_t leaves_only);
int bvh_tree_export_obj(bvh_ctx_t ctx, bvh_tree_t tree, const char* path, int32_t leaves_only);

/* The reference's 32-byte VizNode export of a device tree. */
int bvh_tree_export_viz(bvh_ctx_t ctx, bvh_tree_t tree, const char* path);

#ifdef __cplusplus
}
#endif
#endif /* BVH_IO_H */
