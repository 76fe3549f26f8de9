/*
 * bvh_server.h — serving loop over the C ABI of bvh_hip.h (libbvhhost.so).
 *
 * No reference counterpart: the reference builds and renders one mesh per
 * call, synchronously (src/main.cpp:404-453 build then renderImage;
 * src/batch_runner.cpp:72-120).  The server takes build+trace requests as
 * they arrive and runs them with continuous batching:
 *   - S launch slots, each a bvh context (own HIP stream, default priority,
 *     own workspace) driven by its own host thread;
 *   - a free slot takes the FIFO prefix of the queue that fits the device
 *     memory budget and the group limits, builds every new mesh of that
 *     group in ONE segmented build (bvh_build_batched; plain bvh_build when
 *     the group holds one mesh), then traces each request's rays against
 *     its own tree; groups on different slots overlap on the device;
 *   - meshes and BVHs submitted with a nonzero mesh_key stay resident for
 *     later trace requests (soa all NULL), evicted least-recently-used when
 *     a new request needs the memory;
 *   - every result is the result the request gets alone: trees are
 *     bit-identical whatever the grouping, and traversal is per ray.
 * Timestamps are seconds on the server clock (steady clock, 0 at creation).
 */
#ifndef BVH_SERVER_H
#define BVH_SERVER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bvh_server_rec* bvh_server_t;

typedef struct bvh_server_config {
    int32_t device;
    int32_t slots;            /* concurrent launch groups, >= 1 (default 4 when 0) */
    int32_t max_group;        /* requests coalesced into one group, >= 1 (1 = plain one-mesh-per-launch) */
    uint32_t max_group_tris;  /* triangles built by one group (0 = 16M); a larger single mesh runs alone */
    uint64_t memory_budget;   /* device bytes the server may hold (0 = 90% of the device's free memory);
                                 an eighth is the device cache of evicted meshes' blocks
                                 (bvh_device_cache_set_limit) and a sixteenth headroom for
                                 reused blocks larger than the estimates admitted (residents
                                 are counted at their real size once built): admission uses 13/16 */
    int32_t algo;             /* BVH_ALGO_LBVH or BVH_ALGO_PLOC */
    int32_t ploc_radius;      /* PLOC window radius */
    int32_t trace_order;      /* BVH_TRACE_REFERENCE (0) or BVH_TRACE_NEAREST (bvh_hip.h) for every trace */
} bvh_server_config;

typedef struct bvh_request {
    uint64_t id;              /* caller's id, returned in the completion */
    uint64_t mesh_key;        /* nonzero: keep mesh + BVH resident under this key */
    const float* soa[9];      /* host mesh; all NULL = trace against resident mesh_key */
    uint32_t ntri;
    uint32_t nrays;
    const float* rays;        /* host rays, 6 floats each (origin, direction) */
    float* t_out;             /* host outputs, nrays each: hit distance (1e30 = miss) */
    int32_t* prim_out;        /*   and original triangle index (-1 = miss) */
    double arrival;           /* arrival time on the server clock (bvh_server_now) */
} bvh_request;

typedef struct bvh_completion {
    uint64_t id;
    int32_t status;           /* BVH_OK or a bvh_hip.h status code */
    int32_t slot;             /* slot that ran it */
    uint32_t group_size;      /* requests in its launch group */
    uint32_t built;           /* 1 if its mesh was built for it, 0 if it was resident */
    double arrival, admitted, done;
} bvh_completion;

typedef struct bvh_server_stats {
    uint64_t budget;          /* bytes */
    uint64_t in_use;          /* resident meshes/BVHs + in-flight meshes/BVHs + every slot's
                                 context workspace (held or reserved) */
    uint64_t peak;
    uint64_t resident_bytes;
    uint32_t resident_meshes;
    uint32_t queued;
    uint32_t in_flight;
    uint32_t groups;          /* launch groups run */
    uint32_t completed;
    uint32_t evictions;
    uint32_t rejected;        /* requests that can never fit the budget */
    uint32_t max_group_seen;
} bvh_server_stats;

int bvh_server_create(const bvh_server_config* cfg, bvh_server_t* out);
/* Thread-safe and non-blocking.  Host buffers must stay valid until the
 * request's completion has been polled. */
int bvh_server_submit(bvh_server_t s, const bvh_request* req);
/* Up to `max` completions; waits up to wait_ms for the first one (0 = no wait). */
int bvh_server_poll(bvh_server_t s, bvh_completion* out, uint32_t max, uint32_t* count, int32_t wait_ms);
double bvh_server_now(bvh_server_t s);
int bvh_server_get_stats(bvh_server_t s, bvh_server_stats* st);
/* Waits until nothing is queued or in flight (BVH_E_STATE on time-out). */
int bvh_server_drain(bvh_server_t s, int32_t wait_ms);
const char* bvh_server_last_error(bvh_server_t s);
int bvh_server_destroy(bvh_server_t s);

/* Device bytes of one request from bvh_footprint_estimate: transient = the
 * build workspace of its mesh alone + its trace buffers (held by the slot's
 * context afterwards: context workspaces only grow); resident = its mesh +
 * BVH with traversal records (kept when it is keyed).  Admission counts each
 * slot's held workspace and what a group makes it grow by. */
void bvh_serve_footprint(uint32_t ntri, uint32_t nrays, int32_t algo, uint64_t* transient, uint64_t* resident);

/* The admission policy, pure (no device): given the queue in FIFO order
 * (bytes each request needs now, triangles it builds, whether it can run
 * now), returns how many requests from the head form the next group:
 * the longest prefix of runnable requests whose bytes fit `free_bytes`,
 * with at most max_group requests and max_group_tris triangles (a head
 * request larger than max_group_tris forms a group alone).  0 = the head
 * must wait. */
uint32_t bvh_serve_form_group(const uint64_t* bytes, const uint32_t* tris, const uint8_t* runnable, uint32_t pending,
                              uint64_t free_bytes, uint32_t max_group, uint32_t max_group_tris);

#ifdef __cplusplus
}
#endif
#endif /* BVH_SERVER_H */
