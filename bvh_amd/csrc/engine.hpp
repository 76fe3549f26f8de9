// engine.hpp — internal interface between the HIP kernel files and the C ABI.
// Host-side launchers only; no device code here.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace bvhk {

struct DNode;

// Nine SoA vertex arrays of one mesh, resident in HBM.
struct MeshView {
    const float* v[9];
    uint32_t n;
};

// ---- LBVH/LBVH+/PLOC front end (k_front.hip) ----------------------------
// Kernel geometry for the bounds and Morton passes.
int front_grid(uint32_t n);

// Resets the six order-mapped scene-bound words.
void launch_init_scene(uint32_t* scene_ord, hipStream_t s);
// Per-triangle boxes (2 x float4 per triangle) and the scene box reduction.
void launch_tri_bounds(const MeshView& m, float4* tri_boxes, uint32_t* scene_ord, hipStream_t s);

// Sort plan: digit width, pass count and shifts for a key bit range.
struct SortPlan {
    int bits;          // digit width (8, 10 or 11)
    int passes;        // number of LSD passes
    int shift[8];      // low bit of each pass
    uint32_t mask[8];  // digit mask of each pass (the last pass may be narrower)
    uint32_t tiles;    // onesweep tiles
    uint32_t tile;     // keys per tile
    int hist_blocks;   // blocks of the histogram producer
};
SortPlan make_sort_plan(uint32_t n, int begin_bit, int end_bit, int bits);
// Bytes of the zero-initialised lookback area a plan needs (status words + tile counters).
size_t sort_zero_bytes(const SortPlan& p);

// Morton keys + fused all-pass digit histograms + zeroing of `zero_words` words.
//   m60 = false: keys = morton30 << 32 | i (vals unused)
//   m60 = true : keys = morton60, vals = i
void launch_morton(const float4* tri_boxes, uint32_t n, const uint32_t* scene_ord, bool m60, uint64_t* keys,
                   uint32_t* vals, const SortPlan& plan, uint32_t* hist_part, uint32_t* zero_ptr,
                   size_t zero_words, hipStream_t s);
// Debug: Morton codes only (no histograms).
void launch_morton_codes(const float4* tri_boxes, uint32_t n, const uint32_t* scene_ord, bool m60,
                         uint64_t* codes, hipStream_t s);

// ---- radix sort (k_sort.hip) ---------------------------------------------
// Histogram of keys already in memory (used by the standalone sort).
void launch_sort_histogram(const uint64_t* keys, uint32_t n, const SortPlan& plan, uint32_t* hist_part,
                           hipStream_t s);
// Reduces the per-block histograms and scans them into per-pass bin starts.
// (Row 0 of hist_part is overwritten with the column totals.)
void launch_hist_scan(uint32_t* hist_part, const SortPlan& plan, uint32_t* bin_start, hipStream_t s);
void launch_fill_zero(uint32_t* p, size_t words, hipStream_t s);
void launch_stream_copy(const void* src, void* dst, size_t bytes, hipStream_t s);
// All onesweep passes.  Result ends in (*keys_out, *vals_out) which is one of
// the two buffers passed in.
void launch_onesweep_pass(const SortPlan& plan, int pass, uint32_t n, const uint64_t* kin, uint64_t* kout,
                          const uint32_t* vin, uint32_t* vout, const uint32_t* bin_start, uint32_t* lookback,
                          hipStream_t s);
void launch_onesweep(const SortPlan& plan, uint32_t n, uint64_t* keys_a, uint64_t* keys_b, uint32_t* vals_a,
                     uint32_t* vals_b, const uint32_t* bin_start, uint32_t* lookback, uint64_t** keys_out,
                     uint32_t** vals_out, hipStream_t s);

// ---- hierarchy (k_hier.hip) ------------------------------------------------
struct HierArgs {
    const uint64_t* keys;   // sorted keys
    const uint32_t* vals;   // sorted indices (60-bit keys) or nullptr (packed 30-bit keys)
    uint32_t n;
    const float4* tri_boxes;
    DNode* nodes;           // 2n-1
    uint32_t* indices;      // n
    uint32_t* parents;      // optional, 2n-1
    int2* ranges;           // optional, n-1 (leaf range of each internal node)
    uint64_t* handoff;      // n words, epoch-tagged
    uint32_t epoch;
    uint32_t* err;          // set to nonzero when a sorted key names no primitive
};
int hier_chunk();
void launch_hierarchy(const HierArgs& a, hipStream_t s);

// ---- traversal (k_trace.hip) --------------------------------------------------
struct TraceArgs {
    const DNode* nodes;
    uint32_t num_nodes;
    uint32_t num_prims;
    const float4* tri;          // 3 x float4 per triangle
    const float* rays;          // 6 floats per ray: origin, direction
    uint32_t nrays;
    float* t_out;
    int32_t* prim_out;
    int32_t* counters;          // optional: 3 per ray
    uint32_t* overflow_count;   // [0] queued rays, [1] error flag (cleared by the caller)
    uint32_t* overflow_list;    // nrays entries
};
void launch_tri_aos(const MeshView& m, float4* tri, hipStream_t s);
void launch_trace(const TraceArgs& a, hipStream_t s);
void launch_trace_deep(const TraceArgs& a, hipStream_t s);

// ---- LBVH+ treelets (k_treelet.hip) -------------------------------------------
struct TreeletArgs {
    uint32_t n;
    DNode* nodes;        // 2n-1
    uint32_t* parents;   // 2n-1
    int2* ranges;        // n-1: min/max sorted leaf position under each internal node
    uint32_t* counters;  // n-1 arrival words (2*epoch+1 tags, never cleared)
    uint32_t epoch;
};
void launch_treelet(const TreeletArgs& a, hipStream_t s);

// ---- PLOC (k_ploc.hip) -------------------------------------------------------
// Clusters left when one workgroup takes over all remaining iterations in LDS.
constexpr int kPlocSmall = 2048;
constexpr int kPlocMaxIter = 4096;

// Device-side loop state of one PLOC iteration (32 B).
struct PlocState {
    uint32_t m;          // live clusters entering the iteration
    uint32_t node_base;  // id of the next internal node (before the root swap)
    uint32_t parity;     // which cluster buffer holds the live clusters
    uint32_t iters;      // merge iterations executed so far
    uint32_t flags;      // 1 = no progress, 2 = invalid key, 4 = tail overflow
    uint32_t pad[3];
};

struct PlocArgs {
    uint32_t n;
    int radius;
    float4* box[2];     // cluster boxes (2 x float4 each), double-buffered
    uint32_t* id[2];    // cluster node ids, double-buffered
    int32_t* nn;        // nearest neighbour per cluster
    DNode* nodes;       // 2n-1
    PlocState* state;   // kPlocMaxIter + 2 slots
    uint64_t* status;   // look-back words, one per merge tile
    uint32_t* tile_ctr; // merge tile counter
    int final_slot;     // state slot the tail kernel writes
};
int ploc_small();
int ploc_nn_block();
int ploc_merge_tile();
void launch_ploc_init(const PlocArgs& a, const uint64_t* keys, const float4* tri_boxes, uint32_t* indices,
                      hipStream_t s);
void launch_ploc_nn(const PlocArgs& a, int k, uint32_t m_bound, hipStream_t s);
void launch_ploc_merge(const PlocArgs& a, int k, uint32_t m_bound, hipStream_t s);
void launch_ploc_tail(const PlocArgs& a, int k, hipStream_t s);

}  // namespace bvhk
