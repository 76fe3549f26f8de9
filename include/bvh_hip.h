/*
 * bvh_hip.h — C ABI of the MI355X (gfx950) BVH build and trace engine.
 *
 * Plain C: opaque handles, plain pointers and sizes, integer status codes.
 * Nothing in here exits the process; every call returns BVH_OK (0) or an
 * error code, and bvh_last_error() gives the message of the last failure on a
 * context.  One context owns one HIP stream and a pooled device workspace;
 * use one context per host thread.
 *
 * What each entry point replaces in the reference (paths relative to the
 * reference repository):
 *   bvh_mesh_upload      the nine cudaMalloc + cudaMemcpy H2D uploads at the
 *                        start of every build (src/cuda/lbvh_builder.cu:334-374,
 *                        src/cuda/lbvh_plus_builder.cu:424-439,
 *                        src/cuda/ploc_builder.cu:422-435) and
 *                        src/render_engine.cu:452-479 (uploadMesh)
 *   bvh_build            LBVHBuilderCUDA::runCompute  (lbvh_builder.cu:376-463),
 *                        LBVHPlusBuilderCUDA::runCompute (lbvh_plus_builder.cu:441-508),
 *                        PLOCBuilderCUDA::build (ploc_builder.cu:403-586)
 *   bvh_tree_download    LBVHBuilderCUDA::downloadResults (lbvh_builder.cu:465-482),
 *                        LBVHPlusBuilderCUDA::downloadResults (lbvh_plus_builder.cu:510-527),
 *                        the D2H copies and host root swap at ploc_builder.cu:554-582
 *   bvh_tree_from_nodes  the node upload of renderImage (render_engine.cu:742-746)
 *   bvh_trace            kRenderKernel/traverseBVH (render_engine.cu:289-446)
 *                        over caller-supplied rays
 *   bvh_sort_pairs_u64   cub::DeviceRadixSort::SortPairs (lbvh_builder.cu:412-439) and
 *                        thrust::sort_by_key (lbvh_plus_builder.cu:465, ploc_builder.cu:459-461)
 *   bvh_batch_*          no reference counterpart: the reference builds one mesh
 *                        per builder->build call (src/batch_runner.cpp:72-120 loops
 *                        over models and iterations); a batch builds k meshes in
 *                        one captured launch sequence, each tree identical to
 *                        its own bvh_build
 */
#ifndef BVH_HIP_H
#define BVH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bvh_context* bvh_ctx_t;
typedef struct bvh_mesh_rec* bvh_mesh_t;
typedef struct bvh_tree_rec* bvh_tree_t;
typedef struct bvh_batch_rec* bvh_batch_t;

/* Status codes. */
#define BVH_OK 0
#define BVH_E_ARG 1         /* invalid argument (null handle, n == 0, bad radius ...) */
#define BVH_E_HIP 2         /* a HIP runtime call failed */
#define BVH_E_NOMEM 3       /* device allocation failed */
#define BVH_E_NONFINITE 4   /* a vertex or ray coordinate is NaN or infinite */
#define BVH_E_NOPROGRESS 5  /* PLOC made no merge in an iteration (degenerate input) */
#define BVH_E_BUDGET 6      /* request exceeds the device-memory budget */
#define BVH_E_NODEVICE 7    /* no HIP device */
#define BVH_E_STATE 8       /* call not valid in the current state */
#define BVH_E_UNSUPPORTED 9 /* configuration not supported by this build */

/* Builders (reference: LBVHBuilderCUDA, LBVHPlusBuilderCUDA, PLOCBuilderCUDA). */
#define BVH_ALGO_LBVH 0
#define BVH_ALGO_LBVH_PLUS 1
#define BVH_ALGO_PLOC 2

/* Build flags. */
#define BVH_BUILD_KEEP_KEYS 1u /* keep the sorted keys for bvh_debug_sorted_keys */
#define BVH_BUILD_TUNE 2u      /* batches: time candidate launch configurations at creation (digit width, sort
                                  tile, front-end chunk, PLOC search and merge grids, one at a time), keep the
                                  fastest; every candidate must give the same trees */

typedef struct bvh_build_params {
    int32_t algo;           /* BVH_ALGO_* */
    int32_t ploc_radius;    /* PLOC window radius, >= 1 (reference default 25) */
    int32_t treelet_rounds; /* LBVH+ optimisation rounds, >= 0 (reference: 1) */
    int32_t radix_bits;     /* sort digit width: 0 = measured default (8), or 8 / 10; result-invariant */
    uint32_t flags;         /* BVH_BUILD_* */
} bvh_build_params;

#define BVH_MAX_STAGES 16
typedef struct bvh_stage_times {
    int32_t count;                    /* number of valid entries in ms/names */
    float total_ms;                   /* device time of the whole build */
    float ms[BVH_MAX_STAGES];         /* per-stage device time (hipEvents on the context stream) */
    char names[BVH_MAX_STAGES][24];   /* stage names, as printed by getTimingBreakdown */
    int32_t ploc_iterations;          /* PLOC: merge iterations executed (0 otherwise) */
    int32_t reserved;
} bvh_stage_times;

/* ---- device and context ------------------------------------------------ */
int bvh_device_count(int* count);
int bvh_ctx_create(int device, bvh_ctx_t* out);
int bvh_ctx_destroy(bvh_ctx_t ctx);
const char* bvh_last_error(bvh_ctx_t ctx);
/* Free and total device memory of the context's device (hipMemGetInfo; free
 * includes what the library's device pool holds unused, see
 * bvh_device_cache_info). */
int bvh_device_memory(bvh_ctx_t ctx, uint64_t* free_bytes, uint64_t* total_bytes);
const char* bvh_status_string(int status);

/* ---- raw device memory (used by tests, the renderer and the server) ---- */
int bvh_device_alloc(bvh_ctx_t ctx, size_t bytes, void** dptr);
int bvh_device_free(bvh_ctx_t ctx, void* dptr);
int bvh_copy_h2d(bvh_ctx_t ctx, void* dst, const void* src, size_t bytes);
int bvh_copy_d2h(bvh_ctx_t ctx, void* dst, const void* src, size_t bytes);
/* Page-locked host memory (hipHostMalloc): copies from it to the device are
 * DMA transfers with no pageable staging (the OBJ loader parses into it). */
int bvh_host_alloc_pinned(size_t bytes, void** ptr);
int bvh_host_free_pinned(void* ptr);

/* ---- meshes ------------------------------------------------------------- */
/* soa[0..8] = v0x v0y v0z v1x v1y v1z v2x v2y v2z, each n floats (host memory). */
int bvh_mesh_upload(bvh_ctx_t ctx, const float* const* soa, uint32_t n, bvh_mesh_t* out);
int bvh_mesh_free(bvh_ctx_t ctx, bvh_mesh_t mesh);
/* Like bvh_mesh_free for a caller that knows no queued work on any context
 * uses the mesh any more (bvh_mesh_free waits for the whole device instead):
 * waits for ctx's own stream only and keeps the mesh's device blocks in a
 * per-device cache that later uploads and tree buffers of a similar size
 * reuse when a block is at most a quarter plus 2 MiB larger than asked for
 * (no reference counterpart: a serving loop evicting resident meshes is the
 * caller it is for).  An allocation that finds no device memory frees
 * the cache and retries; bvh_device_cache_trim frees it on demand. */
int bvh_mesh_release(bvh_ctx_t ctx, bvh_mesh_t mesh);
/* Device bytes a mesh or a tree holds (a reused block may be larger than its
 * contents need).  No reference counterpart: for callers that budget device
 * memory, such as the serving loop. */
int bvh_mesh_bytes(bvh_mesh_t mesh, uint64_t* bytes);
int bvh_tree_bytes(bvh_tree_t tree, uint64_t* bytes);
/* Cached bytes and blocks of a device, freeing them all, and the most the
 * cache keeps (default an eighth of the device's memory; a release beyond it
 * frees the blocks at once; 0 keeps nothing; *previous, if given, receives
 * the limit it replaces).  Freed blocks go back to a per-device memory pool
 * that keeps the memory for later allocations (no driver call per block);
 * bvh_device_cache_trim also hands the pool's unused memory back to the
 * device. */
int bvh_device_cache_info(int32_t device, uint64_t* bytes, uint32_t* blocks);
int bvh_device_cache_trim(int32_t device);
int bvh_device_cache_set_limit(int32_t device, uint64_t bytes, uint64_t* previous);
int bvh_mesh_size(bvh_mesh_t mesh, uint32_t* n);
/* BVH_E_NONFINITE when a vertex coordinate of the mesh is NaN or infinite
 * (found by the upload; every build of such a mesh is refused before any
 * launch with the same code), else BVH_OK.  The reference checks nothing. */
int bvh_mesh_check(bvh_mesh_t mesh);

/* ---- trees -------------------------------------------------------------- */
int bvh_tree_create(bvh_ctx_t ctx, bvh_tree_t* out);
int bvh_tree_free(bvh_ctx_t ctx, bvh_tree_t tree);
/* bvh_mesh_release's counterpart for trees: no queued work may use it. */
int bvh_tree_release(bvh_ctx_t ctx, bvh_tree_t tree);
/* num_nodes = 2n-1 for every builder; num_prims = n. */
int bvh_tree_info(bvh_tree_t tree, uint32_t* num_nodes, uint32_t* num_prims);
/* nodes: num_nodes x 48-byte BVHNode records; indices: num_prims uint32 (either may be NULL). */
int bvh_tree_download(bvh_ctx_t ctx, bvh_tree_t tree, void* nodes, uint32_t* indices);
/* Wraps a host node array (e.g. from another builder) into a device tree for tracing. */
int bvh_tree_from_nodes(bvh_ctx_t ctx, const void* nodes, uint32_t num_nodes, bvh_tree_t tree);
/* Sets the primitive index array of such a tree (num_nodes must be 2 * num_prims - 1). */
int bvh_tree_set_indices(bvh_ctx_t ctx, bvh_tree_t tree, const uint32_t* indices, uint32_t num_prims);

/* ---- build -------------------------------------------------------------- */
/* Synchronous: returns once the tree is complete; times may be NULL. */
int bvh_build(bvh_ctx_t ctx, bvh_mesh_t mesh, const bvh_build_params* params, bvh_tree_t tree,
              bvh_stage_times* times);

/* ---- asynchronous build and trace ---------------------------------------- */
/* The reference's build and render calls are synchronous
 * (LBVHBuilderCUDA::build, src/cuda/lbvh_builder.cu:484-488, ends in
 * cudaEventSynchronize at :454; renderImage waits at src/render_engine.cu:789).
 * These enqueue the same work on the context's stream and return without
 * waiting for the device, so a host thread can queue several builds and traces
 * back to back; results are valid, and failures are reported, only after
 * bvh_sync.  Work on one context runs in call order.
 *   bvh_build_async  the work of bvh_build (no stage times).  Argument errors
 *                    and allocation failures return at once; non-finite input,
 *                    PLOC failures and internal errors are reported by
 *                    bvh_sync.  A call that has to grow the context's
 *                    workspace or the tree's storage first waits for the work
 *                    already queued; PLOC returns once its host-driven global
 *                    iterations are done (their loop reads one 32-byte state
 *                    per batch of launches), with its LDS tail still queued.
 *   bvh_trace_async  the work of bvh_trace on device buffers (no time); a ray
 *                    needing more than 256 stack entries is reported by bvh_sync.
 *   bvh_sync         waits for everything queued on the context and returns the
 *                    first failure among the asynchronous calls since the last
 *                    bvh_sync, in call order (BVH_OK if none); it forgets them
 *                    either way.  At most 65,536 asynchronous calls may be
 *                    pending (then BVH_E_STATE until bvh_sync). */
int bvh_build_async(bvh_ctx_t ctx, bvh_mesh_t mesh, const bvh_build_params* params, bvh_tree_t tree);
int bvh_trace_async(bvh_ctx_t ctx, bvh_tree_t tree, bvh_mesh_t mesh, const float* d_rays, uint32_t nrays, float* d_t,
                    int32_t* d_prim, int32_t* d_counters);
int bvh_sync(bvh_ctx_t ctx);

/* ---- device-memory footprint --------------------------------------------- */
/* No reference counterpart (the reference allocates per build and never
 * budgets).  Device bytes the library allocates, growth slack included, for
 * the pieces of a request: admission control sizes requests with this. */
typedef struct bvh_footprint {
    uint64_t mesh;             /* bvh_mesh_upload of one mesh of ntri triangles */
    uint64_t tree;             /* its tree: nodes, indices and traversal records */
    uint64_t build_workspace;  /* context workspace of one build over `meshes` meshes totalling ntri
                                  triangles (an upper bound over how they are split) */
    uint64_t trace_workspace;  /* context buffers of one bvh_trace_host of nrays rays */
} bvh_footprint;
int bvh_footprint_estimate(int32_t algo, uint32_t ntri, uint32_t meshes, uint32_t nrays, bvh_footprint* out);
/* Device bytes a context holds now in its grow-only build workspace and in
 * its trace buffers (released by bvh_ctx_destroy). */
int bvh_ctx_held_bytes(bvh_ctx_t ctx, uint64_t* build_workspace, uint64_t* trace_workspace);

/* ---- build without allocation ------------------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation): an LBVH build (the same kernels and the same result as
 * bvh_build with BVH_ALGO_LBVH) whose working memory is a fixed scratch array
 * of 40 MiB that the library holds on every device it runs on, from context
 * creation until the process ends.  It calls no hipMalloc or hipFree.
 * d_soa: nine device pointers (v0x v0y v0z v1x v1y v1z v2x v2y v2z), n floats
 * each; d_nodes: (2n-1) 48-byte BVHNode records; d_indices: n uint32 — all
 * device memory owned by the caller.  1 <= n <= BVH_NOALLOC_MAX_TRIANGLES,
 * else BVH_E_ARG.  Synchronous on the context stream.  Every call on one GPU
 * shares that GPU's scratch: calls that overlap on one GPU (from several
 * contexts or threads) race and must be serialised by the caller. */
#define BVH_NOALLOC_MAX_TRIANGLES 262144u
int bvh_build_lbvh_noalloc(bvh_ctx_t ctx, const float* const* d_soa, uint32_t n, void* d_nodes,
                           uint32_t* d_indices);

/* ---- batched builds ------------------------------------------------------ */
/* A batch builds k resident meshes into k trees with one algorithm (LBVH or
 * PLOC) in one hipGraph-captured launch sequence.  Every tree is bit-identical
 * to the tree bvh_build gives for its mesh alone.  Meshes and trees stay owned
 * by the caller and must outlive the batch; a tree of a batch must not be
 * rebuilt by another call while the batch may run.  Limits: every mesh has
 * >= 1 triangle; ceil(log2(max n)) + 30 + ceil(log2(k)) <= 64 key bits. */
typedef struct bvh_batch_info {
    uint32_t meshes;
    uint32_t triangles;                /* total over the batch */
    uint32_t max_triangles;            /* largest mesh */
    int32_t algo;
    int32_t radix_bits;                /* digit width in use */
    int32_t sort_passes;
    int32_t key_bits;                  /* sorted key bits: morton30 + mesh id (index bits below are not sorted) */
    int32_t ploc_captured_iterations;  /* global PLOC iterations in the graph */
    int32_t ploc_iterations;           /* global PLOC iterations the last run needed */
    int32_t ploc_nn_grid;              /* persistent grid of the neighbour search */
    int32_t ploc_merge_grid;           /* persistent grid of merge + compaction */
    int32_t graph_nodes;               /* kernel nodes in the captured graph */
    int32_t runs;
    int32_t eager_continuations;       /* runs that needed iterations beyond the graph */
    int32_t sort_block;                /* threads per onesweep tile in use (256 or 512) */
    int32_t front_chunk;               /* triangles per front-end block in use */
    int32_t tune_count;                /* candidates timed by BVH_BUILD_TUNE */
    float tune_ms[16];                 /* their median device times */
    int32_t tune_radix_bits[16];       /* and their settings */
    int32_t tune_nn_grid[16];
    int32_t tune_sort_block[16];
    int32_t tune_front_chunk[16];
    int32_t tune_merge_grid[16];
} bvh_batch_info;

/* The same segmented build launched directly on the context stream (no
 * graph), for mesh sets that change from call to call (the serving loop).
 * times (optional) covers the whole build. */
int bvh_build_batched(bvh_ctx_t ctx, const bvh_mesh_t* meshes, const bvh_tree_t* trees, uint32_t k,
                      const bvh_build_params* params, bvh_stage_times* times);

int bvh_batch_create(bvh_ctx_t ctx, const bvh_mesh_t* meshes, const bvh_tree_t* trees, uint32_t k,
                     const bvh_build_params* params, bvh_batch_t* out);
/* Synchronous; ms (optional) = device time of the whole batch. */
int bvh_batch_run(bvh_batch_t batch, float* ms);
int bvh_batch_get_info(bvh_batch_t batch, bvh_batch_info* info);
int bvh_batch_free(bvh_batch_t batch);

/* ---- traversal ----------------------------------------------------------- */
/* Visit order of a context's traces (bvh_trace, bvh_trace_host,
 * bvh_trace_async), set by bvh_ctx_set_trace_order.
 *   BVH_TRACE_REFERENCE (the default): traverseBVH's order and predicates
 *     (render_engine.cu:289-356): left child first, a node entered when its
 *     slab entry distance is below the closest t, a hit taken when its t is
 *     strictly smaller; ids, t and the three counters are the reference's.
 *   BVH_TRACE_NEAREST: no reference counterpart, for callers that want the
 *     closest hit and not the reference's counters: of two children the one
 *     with the smaller entry distance is visited first (left on a tie), a
 *     node is entered when its entry distance is at most the closest t, and
 *     the closest hit is the least (t, triangle index) pair, so it does not
 *     depend on the visit order.  Counters are those of this order. */
#define BVH_TRACE_REFERENCE 0
#define BVH_TRACE_NEAREST 1
int bvh_ctx_set_trace_order(bvh_ctx_t ctx, int32_t order);
/* Builds the tree's 64-byte-per-node traversal records now (the first trace
 * after a build does it otherwise).  A tree may be traced from several
 * contexts at once only after this (or one trace) has returned. */
int bvh_tree_prepare(bvh_ctx_t ctx, bvh_tree_t tree);
/* Closest hit of nrays rays (6 floats each: origin xyz, direction xyz) against
 * `tree` over `mesh`, all buffers in device memory.  t_out[i] = hit distance
 * (1e30 when missed), prim_out[i] = original triangle index or -1;
 * counters (optional, 3 int32 per ray) = nodes visited, AABB tests, triangle
 * tests, as counted by the reference renderer.  ms (optional) = device time. */
int bvh_trace(bvh_ctx_t ctx, bvh_tree_t tree, bvh_mesh_t mesh, const float* d_rays, uint32_t nrays, float* d_t,
              int32_t* d_prim, int32_t* d_counters, float* ms);
/* Same with host buffers (staged through the context's workspace). */
int bvh_trace_host(bvh_ctx_t ctx, bvh_tree_t tree, bvh_mesh_t mesh, const float* rays, uint32_t nrays, float* t,
                   int32_t* prim, int32_t* counters, float* ms);

/* Debug views for parity tests.  Morton codes of the mesh in input order:
 * algo LBVH/PLOC give 30-bit codes, LBVH+ 60-bit codes (uint64 each). */
int bvh_debug_morton(bvh_ctx_t ctx, bvh_mesh_t mesh, int32_t algo, uint64_t* codes);
/* Sorted keys of the last build made with BVH_BUILD_KEEP_KEYS.  LBVH/PLOC keys
 * are (morton30 << 32 | index) with vals unused; LBVH+ keys are morton60 with
 * the index in vals. */
int bvh_debug_sorted_keys(bvh_ctx_t ctx, uint64_t* keys, uint32_t* vals, uint32_t n);

/* Stable LSD radix sort of (key, val) pairs by key bits [begin_bit, end_bit),
 * in place on device buffers; d_vals may be NULL. */
int bvh_sort_pairs_u64(bvh_ctx_t ctx, uint64_t* d_keys, uint32_t* d_vals, uint32_t n, int begin_bit,
                       int end_bit, int radix_bits, float* ms);

/* Plain 16-byte-vector copy kernel of `bytes` (multiple of 16) between device
 * buffers: the calibration point for HBM traffic counters.  No reference
 * counterpart (measurement support). */
int bvh_debug_stream_copy(bvh_ctx_t ctx, void* d_dst, const void* d_src, size_t bytes, float* ms);

#ifdef __cplusplus
}
#endif
#endif /* BVH_HIP_H */
