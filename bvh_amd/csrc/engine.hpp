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

// ---- segments ----------------------------------------------------------------
// One mesh of a build.  A standalone build is a batch of one mesh.
struct SegDesc {
    const float* v[9];   // vertex SoA arrays
    uint32_t n;          // triangles
    uint32_t start;      // first slot of this mesh in the batch-wide per-triangle arrays
    DNode* nodes;        // output: 2n-1 nodes
    uint32_t* indices;   // output: n primitive indices
};

// How a kernel's blocks map onto segments: block b covers chunk blocks[b].y of
// segment blocks[b].x; with blocks == nullptr, one segment (`single`) and
// chunk = blockIdx.x.
struct SegBlocks {
    const SegDesc* segs;
    const uint2* blocks;
    SegDesc single;
    uint32_t nblocks;
    uint32_t chunk;  // elements per chunk for the front kernels and PLOC init (unused by the hierarchy)
};

// Packed 64-bit sort key of LBVH/PLOC: seg << seg_shift | morton30 << idx_bits | index.
// A single mesh uses idx_bits = 32 (key = morton30 << 32 | index).
struct KeyFormat {
    uint32_t idx_bits;
    uint32_t seg_shift;
};

// ---- LBVH/LBVH+/PLOC front end (k_front.hip) ----------------------------
// Triangles per block of the front kernels: kFrontChunk, or kFrontChunkLarge
// for a single mesh of at least kFrontLargeMin triangles (fewer blocks, so
// fewer scene and digit-total atomics; at 4M blob triangles bounds + Morton
// take 103 us instead of 114, while below ~1M the smaller chunk keeps more CUs
// busy).
constexpr uint32_t kFrontChunk = 4096;
constexpr uint32_t kFrontChunkLarge = 8192;
constexpr uint32_t kFrontLargeMin = 2000000;
constexpr int kFrontUnroll = 4;          // triangles per thread whose loads are in flight together

// Resets the per-segment scene words (8 per segment), clears hist_words words
// of the digit-total table and ctl_words hierarchy control words (when given)
// and advances *epoch.
void launch_init_build(uint32_t* seg_scene, uint32_t k, uint32_t* epoch, hipStream_t s, uint32_t* hist_tot = nullptr,
                       uint32_t hist_words = 0, uint32_t* ctl = nullptr, uint32_t ctl_words = 0);
// Per-triangle boxes (2 x float4 per triangle) and each segment's scene box.
void launch_tri_bounds(const SegBlocks& sb, float4* tri_boxes, uint32_t* seg_scene, hipStream_t s);

// Onesweep tile: threads x keys per thread.  Key-only passes with 8-bit
// digits over at least kSortWideMinKeys keys use the wide tile (twice the
// keys, so half the tiles: half the tile tickets, which serialise on one
// counter, and a shorter look-back chain; 146 KB of LDS).  Below that the
// narrow tiles keep more CUs busy (measured crossover near 300 narrow tiles
// on 256 CUs).
constexpr int kSortThreads = 512;
constexpr int kSortThreadsWide = 1024;
constexpr int kSortThreadsSmall = 256;  // 4096-key tiles: a batch tuning option (no payload)
constexpr int kSortItems = 16;
constexpr uint32_t kSortWideMinKeys = 2500000;

// Sort plan: digit width, pass count and shifts for a key bit range.
struct SortPlan {
    int bits;          // digit width (8 or 10)
    int passes;        // number of LSD passes
    int shift[8];      // low bit of each pass
    uint32_t mask[8];  // digit mask of each pass (the last pass may be narrower)
    uint32_t tiles;    // onesweep tiles
    uint32_t tile;     // keys per tile
    int block;         // threads per tile (kSortThreads or kSortThreadsWide)
    int hist_blocks;   // blocks of the histogram producer
};
// hist_blocks <= 0 picks a default producer size; vals: the sort carries a
// 32-bit payload (narrow tiles); wide_ok = false keeps the narrow tile;
// block > 0 asks for that tile (kSortThreadsSmall or kSortThreads; payload
// sorts always use kSortThreads).
SortPlan make_sort_plan(uint32_t n, int begin_bit, int end_bit, int bits, int hist_blocks = 0, bool vals = false,
                        bool wide_ok = true, int block = 0);
// Bytes of the zero-initialised lookback area a plan needs (status words + tile counters).
size_t sort_zero_bytes(const SortPlan& p);

// Morton keys + all-pass digit totals added into hist_tot (cleared by
// launch_init_build) + zeroing of `zero_words` words.
//   m60 = false: keys = packed per KeyFormat (vals unused)
//   m60 = true : keys = morton60, vals = index (single segment only)
void launch_morton(const SegBlocks& sb, const float4* tri_boxes, const uint32_t* seg_scene, bool m60,
                   const KeyFormat& kf, uint64_t* keys, uint32_t* vals, const SortPlan& plan, uint32_t* hist_tot,
                   uint32_t* zero_ptr, size_t zero_words, hipStream_t s);
// Debug: Morton codes only (no histograms), in input order.
void launch_morton_codes(const SegBlocks& sb, const float4* tri_boxes, const uint32_t* seg_scene, bool m60,
                         uint64_t* codes, hipStream_t s);

// ---- radix sort (k_sort.hip) ---------------------------------------------
// Digit totals (passes x bins) of keys already in memory (standalone and ray
// sorts); clears the table first.
void launch_sort_histogram(const uint64_t* keys, uint32_t n, const SortPlan& plan, uint32_t* hist_tot,
                           hipStream_t s);
void launch_fill_zero(uint32_t* p, size_t words, hipStream_t s);
void launch_stream_copy(const void* src, void* dst, size_t bytes, hipStream_t s);
// All onesweep passes.  Result ends in (*keys_out, *vals_out) which is one of
// the two buffers passed in.  Both return false, launching nothing, when a
// payload is passed with a plan made for keys only (its tiles are too large
// for a payload's LDS staging).
bool launch_onesweep_pass(const SortPlan& plan, int pass, uint32_t n, const uint64_t* kin, uint64_t* kout,
                          const uint32_t* vin, uint32_t* vout, const uint32_t* hist_tot, uint32_t* lookback,
                          hipStream_t s);
bool launch_onesweep(const SortPlan& plan, uint32_t n, uint64_t* keys_a, uint64_t* keys_b, uint32_t* vals_a,
                     uint32_t* vals_b, const uint32_t* hist_tot, uint32_t* lookback, uint64_t** keys_out,
                     uint32_t** vals_out, hipStream_t s);

// ---- hierarchy (k_hier.hip) ------------------------------------------------
// A finished subtree handed from one hierarchy level to the next (48 B):
// leaf range, mesh, packed delimiting deltas, box.
struct __align__(16) HUnit {
    uint32_t l, r, seg, beta;
    float mnx, mny, mnz, mxx;
    float mxy, mxz, pad0, pad1;
};
struct HierArgs {
    SegBlocks sb;           // level-1 blocks over chunks of hier_chunk() leaves
    const uint64_t* keys;   // sorted keys (batch-wide)
    const uint32_t* vals;   // sorted indices (60-bit keys) or nullptr (packed keys)
    KeyFormat kf;
    const float4* tri_boxes;
    uint32_t* parents;      // optional, 2n-1 (single segment)
    uint64_t* handoff;      // batch-wide, epoch-tagged (final climbing pass)
    const uint32_t* epoch;  // device word advanced by launch_init_build
    uint32_t* err;          // set to nonzero when a sorted key names no primitive
    HUnit* units[2];        // [0]: level-1 chunk slots (chunks x hier_chunk()), later levels;
                            // [1]: one entry per triangle
    uint32_t* counts;       // level-1 exits per chunk
    const HUnit* units_in;  // set per launch
    HUnit* units_out;
    uint64_t* status;       // hier_max_levels() x status_stride look-back words, epoch-tagged
    uint32_t status_stride; // >= level-1 chunks + 1
    uint32_t* ctl;          // 2 x hier_max_levels() words cleared by launch_init_build:
                            // per-level chunk tickets, then per-level unit counts
    int levels;             // chunked levels (hier_levels); a climbing pass finishes the rest
    int fold1;              // set by launch_hierarchy: level 1 maps the leaf level's slots itself
};
int hier_chunk();
int hier_unit_bytes();
int hier_max_levels();
int hier_levels(uint32_t maxn, uint32_t meshes);
void launch_hierarchy(const HierArgs& a, hipStream_t s);

// ---- traversal (k_trace.hip) --------------------------------------------------
struct TraceArgs {
    const DNode* nodes;
    uint32_t num_nodes;
    uint32_t num_prims;
    const float4* tri;          // 3 x float4 per triangle
    const float4* rec;          // 4 x float4 per internal node (k_trace_layout)
    const float* rays;          // 6 floats per ray: origin, direction
    uint32_t nrays;
    const uint64_t* order;      // optional: sorted coherence keys (ray index in the low bits)
    uint64_t order_mask;
    float* t_out;
    int32_t* prim_out;
    int32_t* counters;          // optional: 3 per ray
    uint32_t* overflow_count;   // [0] queued rays, [1] error flag, [2] batch counter (cleared by the caller)
    uint32_t* overflow_list;    // nrays entries
    int2* spill;                // spill_slots x trace_spill_entries() stack entries
    uint32_t spill_slots;
    int32_t nearest;            // 0: the reference's visit order; 1: nearer child first, least (t, index) hit
};
// Stack entries of one spill slot (the part of a lane's stack beyond LDS).
int trace_spill_entries();
// Spill slots bvh_trace provides for n rays.
uint32_t trace_spill_slots(uint32_t nrays);
void launch_tri_aos(const MeshView& m, float4* tri, uint32_t* bad, hipStream_t s);
// Coherence keys of rays (direction bins, origin Morton code, ray index).
void launch_ray_keys(const float* rays, uint32_t n, const DNode* nodes, uint32_t idx_bits, uint64_t* keys,
                     hipStream_t s);
// Per-internal-node traversal records (64 B each, indexed by node id).
void launch_trace_layout(const DNode* nodes, uint32_t num_nodes, float4* rec, hipStream_t s);
void launch_trace(const TraceArgs& a, hipStream_t s);
void launch_trace_deep(const TraceArgs& a, hipStream_t s);

// ---- LBVH+ treelets (k_treelet.hip) -------------------------------------------
struct TreeletArgs {
    uint32_t n;
    DNode* nodes;        // 2n-1
    uint32_t* parents;   // 2n-1
    uint32_t* counters;  // n-1 arrival words (2*epoch+1 tags, never cleared)
    uint32_t* list;      // n-1: ready nodes, list 1, list 2, ... back to back
    uint32_t* cnt;       // levels + 2 list lengths (cleared by launch_treelet)
    uint32_t epoch;
    uint32_t level_min;  // lists at least this long run level-synchronously; the first shorter seeds the tail
    int levels;          // level launches per round
};
// One restructuring round: seed, `levels` level launches, the climbing tail.
void launch_treelet(const TreeletArgs& a, hipStream_t s);

// ---- PLOC (k_ploc.hip) -------------------------------------------------------
// Clusters per mesh at or below which one workgroup finishes the mesh in LDS.
constexpr int kPlocSmall = 2048;
constexpr int kPlocMaxIter = 4096;
constexpr uint32_t kPlocNoProgress = 1u;   // an iteration of a mesh with >= 2 clusters merged nothing
constexpr uint32_t kPlocUnfinished = 8u;   // the tail ran before every mesh was small

// Device-side loop state of one PLOC iteration (32 B).
struct PlocState {
    uint32_t m;          // clusters (all meshes) entering the iteration
    uint32_t maxm;       // largest per-mesh cluster count entering the iteration
    uint32_t parity;     // which cluster buffer holds the live clusters
    uint32_t iters;      // global iterations executed so far
    uint32_t flags;      // kPlocNoProgress
    uint32_t pad[3];
};

struct PlocArgs {
    SegBlocks sb;       // meshes; blocks over sb.chunk-triangle chunks (for the init kernel)
    uint32_t nseg;      // meshes
    int radius;
    uint32_t m0, maxm0; // total and largest triangle count
    float4* box[2];     // cluster boxes (2 x float4 each), double-buffered
    uint32_t* id[2];    // cluster node ids (mesh-local), double-buffered
    uint32_t* sid[2];   // cluster mesh index, double-buffered (nullptr for one mesh)
    int32_t* nn;        // nearest neighbour per cluster (global slot)
    uint4* seg[2];      // per mesh and parity: first cluster, count, next node id, iterations
    PlocState* state;   // kPlocMaxIter slots
    uint64_t* status;   // look-back words, one per merge tile
    uint32_t* tile_ctr; // merge tile counter
    uint4* out;         // per mesh, written by the tail: count, next node id, iterations, flags
};
int ploc_small();
int ploc_merge_tile();
void launch_ploc_init(const PlocArgs& a, const uint64_t* keys, const KeyFormat& kf, const float4* tri_boxes,
                      uint32_t* err, hipStream_t s);
// One global iteration (no-op once every mesh is small).  The persistent grid
// is sized for m_bound clusters and capped at max_grid blocks.
void launch_ploc_nn(const PlocArgs& a, int k, uint32_t m_bound, uint32_t max_grid, hipStream_t s);
void launch_ploc_merge(const PlocArgs& a, int k, uint32_t m_bound, uint32_t max_grid, hipStream_t s);
// One workgroup per mesh; reads state slot k.
void launch_ploc_tail(const PlocArgs& a, int k, hipStream_t s);

}  // namespace bvhk
