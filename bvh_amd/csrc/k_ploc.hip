// k_ploc.hip — PLOC (Meister & Bittner 2017) on gfx950, for one mesh or for a
// batch of meshes (segments) in the same launches.
//
// Replaces kGatherAABBs/kInitLeafNodes_ploc/kInitializeSequence,
// kFindNearestNeighbors_ploc, kMergeClusters_ploc, cub::DeviceScan +
// kCompactIndices/kCompactAABBs, the per-iteration host round trips and the
// host root swap of the reference (src/cuda/ploc_builder.cu:143-293,403-586).
//
// Semantics kept from the reference, per mesh:
//   * leaves are nodes 0..n-1 in Morton order, internal nodes are numbered
//     from n in creation order, and the root (last created) is swapped with
//     node 0 (ploc_builder.cu:563-582) — applied here as an index remap at
//     write time, so no host pass is needed;
//   * nearest neighbour of cluster i = argmin over j in [i-r, i+r] ∩ mesh,
//     j != i, of SA(box_i U box_j), ties to the lowest j, non-finite distances
//     never chosen (the strict '<' scan of ploc_builder.cu:190-205);
//   * mutual pairs merge into the lower slot, the higher slot is dropped.
// Deliberate difference: new internal nodes are numbered by rank of the
// merging slot (ascending) within each iteration, one legal serialisation of
// the reference's nondeterministic atomicAdd (ploc_builder.cu:245).
//
// Segments: the clusters of all meshes live in one array, each mesh's
// clusters contiguous and in mesh order.  A window never crosses a mesh
// boundary, compaction keeps every mesh contiguous, and node ids are ranked
// within the mesh by a segmented scan, so every mesh gets exactly the tree
// its own build gives (a single build is a batch of one).
//
// MI355X mapping:
//   * k_ploc_nn: persistent blocks; each tile stages 768 clusters plus an
//     r-halo on each side in LDS (SoA floats), every lane owns three
//     consecutive clusters and scans the union of their windows once, so each
//     box read from LDS (stride 3 across lanes: conflict-free) feeds three
//     candidates; candidates pack (distance bits << 32 | j) into one u64
//     minimum each; up to r = 128 each pair is evaluated once (the distance
//     is symmetric) and the later cluster's half arrives by 64-bit LDS
//     atomic minimum;
//   * k_ploc_merge: persistent blocks take tiles of 1024 clusters in order
//     and run mutual-pair detection, new nodes, and stream compaction of the
//     survivors behind one decoupled look-back scan whose 64-bit words pack
//     (segment-start flag, merges within the mesh, survivors overall); the
//     tile's global traffic is one lane per cluster, the scan runs on its
//     flags in LDS;
//   * loop control lives on the device: iteration k reads state[k] and
//     writes state[k+1] (cluster count, largest mesh, buffer parity; a ring of
//     kPlocMaxIter slots) and per-mesh records (first cluster, count, next
//     node id), so the host launches (or captures) fixed sequences;
//   * once every mesh has <= kPlocSmall clusters, one workgroup per mesh
//     finishes all remaining iterations inside LDS (k_ploc_tail).
#include "dev_common.hpp"
#include "engine.hpp"

namespace bvhk {

constexpr int kNnBlock = 256;
// Consecutive clusters per lane (odd: conflict-free LDS): every box read
// from LDS serves that many candidates and one LDS minimum per candidate j
// covers them all, against more registers per lane.  Measured: three at
// r = 25 (five: 3.77 against 3.73 ms per 4M build), five at r = 100 (32M
// strands 27.1 against 29.3 ms).
constexpr int kNnPer = 3;
constexpr int kNnPerWide = 5;
constexpr int kNnWideR = 64;                 // radius from which lanes take kNnPerWide clusters

constexpr uint32_t kNnFill = 1u << 18;        // lanes that keep every CU busy
constexpr int kNnSymMaxR = 128;               // symmetric search up to this radius (its halo pairs grow as r^2)
constexpr int kMergeBlock = 256;
constexpr uint32_t kMergeBigMin = 1u << 18;   // from here on, 4 clusters per thread per merge tile
constexpr int kTailBlock = 1024;
constexpr uint32_t kNnGridMax = 2048;
constexpr uint32_t kMergeGridMax = 1024;

// Look-back word: bits 63-62 status, 61 segment start, 60-31 merges within
// the current mesh, 30-0 survivors overall.
constexpr uint64_t kStA = 1ull << 62;
constexpr uint64_t kStP = 2ull << 62;
constexpr uint64_t kStFlags = 3ull << 62;
constexpr uint64_t kSegF = 1ull << 61;
constexpr uint64_t kVMask = (1ull << 31) - 1;
constexpr uint64_t kMMask = ((1ull << 30) - 1) << 31;

int ploc_small() { return kPlocSmall; }
int ploc_merge_tile() { return kMergeBlock; }  // smallest merge tile (sizes the look-back words)

// Merge tile of an iteration with m live clusters.
__host__ __device__ __forceinline__ uint32_t merge_tile(uint32_t m) {
    return m >= kMergeBigMin ? 4u * kMergeBlock : (uint32_t)kMergeBlock;
}

// Lanes per cluster in the neighbour search: 0 = three clusters per lane
// (large m), else G lanes split each window (small m: more lanes in flight).
// Measured: three per lane (with symmetric pairs) beats one per lane from
// kNnFill clusters on.
__device__ __forceinline__ int nn_lanes(uint32_t m) {
    if (m >= kNnFill) return 0;
    if (m >= kNnFill / 2) return 2;
    if (m >= kNnFill / 4) return 4;
    return 8;
}

__device__ __forceinline__ uint32_t remap_id(uint32_t id, uint32_t root) {
    return id == 0 ? root : (id == root ? 0u : id);
}

__device__ __forceinline__ const SegDesc& seg_desc(const SegBlocks& sb, uint32_t s) {
    return sb.blocks ? sb.segs[s] : sb.single;
}

// Union of two finite boxes with the hardware min/max, for areas only.
__device__ __forceinline__ Box merge_hw(const Box& a, const Box& b) {
    Box r;
    r.mnx = fminf(a.mnx, b.mnx);
    r.mny = fminf(a.mny, b.mny);
    r.mnz = fminf(a.mnz, b.mnz);
    r.mxx = fmaxf(a.mxx, b.mxx);
    r.mxy = fmaxf(a.mxy, b.mxy);
    r.mxz = fmaxf(a.mxz, b.mxz);
    return r;
}

// Packed neighbour candidate; non-finite distances are never candidates.
__device__ __forceinline__ uint64_t nn_key(float dist, uint32_t j) {
    uint32_t bits = __float_as_uint(dist);
    return ((bits & 0x7f800000u) == 0x7f800000u) ? ~0ull : (((uint64_t)bits << 32) | j);
}

// Segmented-scan operator on look-back values (a precedes b): survivors add,
// merges add unless b starts a mesh.
__device__ __forceinline__ uint64_t scomb(uint64_t a, uint64_t b) {
    const uint64_t v = ((a & kVMask) + (b & kVMask)) & kVMask;
    const uint64_t mm = (b & kSegF) ? (b & kMMask) : (((a & kMMask) + (b & kMMask)) & kMMask);
    return ((a | b) & kSegF) | mm | v;
}

__global__ __launch_bounds__(256) void k_ploc_init(PlocArgs a, const uint64_t* __restrict__ keys, KeyFormat kf,
                                                   const float4* __restrict__ tri_boxes, uint32_t* err) {
    uint32_t seg, chunk;
    SegDesc d;
    if (a.sb.blocks) {
        const uint2 e = a.sb.blocks[blockIdx.x];
        seg = e.x;
        chunk = e.y;
        d = a.sb.segs[seg];
    } else {
        seg = 0;
        chunk = blockIdx.x;
        d = a.sb.single;
    }
    const uint32_t n = d.n;
    const uint32_t root = 2 * n - 2;
    const uint64_t idx_mask = (kf.idx_bits >= 32) ? 0xFFFFFFFFull : ((1ull << kf.idx_bits) - 1);
    const uint32_t lo = chunk * a.sb.chunk;
    const uint32_t hi = (lo + a.sb.chunk < n) ? lo + a.sb.chunk : n;
    const float4* __restrict__ tb = tri_boxes + 2 * (size_t)d.start;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += 256) {
        const size_t g = (size_t)d.start + i;
        uint32_t prim = (uint32_t)(keys[g] & idx_mask);
        if (prim >= n) {  // cannot happen for a correct sort; never index out of bounds
            atomicOr(err, 1u);
            prim = 0;
        }
        Box b = load_box(tb + 2 * (size_t)prim);
        store_node(d.nodes, n > 1 ? remap_id(i, root) : i, b, (int32_t)(prim | 0x80000000u), -1);
        d.indices[i] = prim;
        store_box(a.box[0] + 2 * g, b);
        a.id[0][g] = i;
        if (a.sid[0]) a.sid[0][g] = seg;
    }
    if (chunk == 0 && threadIdx.x == 0) a.seg[0][seg] = make_uint4(d.start, n, n, 0u);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        PlocState s;
        s.m = a.m0;
        s.maxm = a.maxm0;
        s.parity = 0;
        s.iters = 0;
        s.flags = 0;
        s.pad[0] = s.pad[1] = s.pad[2] = 0;
        a.state[0] = s;
    }
}

__device__ __forceinline__ bool ploc_global_step(const PlocState& st) {
    return st.maxm > (uint32_t)kPlocSmall && st.flags == 0;
}

// Symmetric neighbour search of one staged tile (three clusters per lane).
// dist(i, j) = dist(j, i) bit for bit: the two unions differ at most in the
// sign of a zero coordinate, which changes no extent except a zero one's
// sign, and the area of such a box is never -0 (that needs all three
// products -0, i.e. an odd number of -0 extents in each of three pairs).  So
// every pair i < j in a window is evaluated once, by the lane owning i: the
// candidate (dist, j) goes into i's running minimum, and (dist, i) into j's
// LDS minimum with a 64-bit LDS atomic.  Pairs whose i lies in the r clusters
// before the tile (and j in it) are evaluated by the tile as well; pairs whose
// j lies past the tile belong to the next tile's backward half.  The minimum
// of (distance bits, index) over both halves is the reference's first minimum
// of its ascending strict-< scan.
template <bool SEG, int PER>
__device__ __forceinline__ void nn_tile_symmetric(const PlocArgs& a, const float* s_nn, int span, int64_t t0,
                                                  uint32_t m, int r, uint32_t par) {
    uint64_t* s_back = reinterpret_cast<uint64_t*>(const_cast<float*>(s_nn) + 6 * span);
    const uint4* __restrict__ segs = a.seg[par];
    const int64_t tend = t0 + (kNnBlock * PER);  // first cluster past the tile
    auto box_at = [&](int64_t j) {
        const int q = (int)(j - t0) + r;
        Box b;
        b.mnx = s_nn[0 * span + q]; b.mny = s_nn[1 * span + q]; b.mnz = s_nn[2 * span + q];
        b.mxx = s_nn[3 * span + q]; b.mxy = s_nn[4 * span + q]; b.mxz = s_nn[5 * span + q];
        return b;
    };
    auto back_min = [&](int64_t j, uint64_t key) {
        __hip_atomic_fetch_min(reinterpret_cast<unsigned long long*>(s_back + (j - t0)), (unsigned long long)key,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    const int64_t i0 = t0 + (int64_t)PER * threadIdx.x;
    uint64_t best[PER];
#pragma unroll
    for (int c = 0; c < PER; ++c) best[c] = ~0ull;
    if (i0 < (int64_t)m) {
        // Forward windows (i, hi[c]] clipped to the mesh; hi is non-decreasing
        // in c, so one pass over (i0, max hi] serves all three.
        Box bi[PER];
        int64_t hi[PER];
#pragma unroll
        for (int c = 0; c < PER; ++c) {
            const int64_t i = i0 + c;
            if (i < (int64_t)m) {
                const uint4 sg = segs[SEG ? a.sid[par][i] : 0u];
                const int64_t shi = (int64_t)sg.x + sg.y - 1;
                hi[c] = (i + r < shi) ? i + r : shi;
                bi[c] = box_at(i);
            } else {
                hi[c] = i;  // empty forward window
                bi[c] = bi[0];
            }
        }
        int64_t jhi = hi[0];
#pragma unroll
        for (int c = 1; c < PER; ++c) jhi = hi[c] > jhi ? hi[c] : jhi;
        // Running minima as (distance bits, index): distances are never
        // negative, so their bits order like the values; starting from the
        // bits of +inf keeps every non-finite distance out, as the packed
        // keys do; a strict < over ascending indices keeps the first minimum.
        // Area-only unions use the hardware min/max: on finite boxes they
        // differ from the selects at most in the sign of a zero extent, which
        // changes no area's bits (an area of -0 would need all three products
        // -0, i.e. an odd number of -0 extents in each of three pairs).
        constexpr uint32_t kInfBits = 0x7f800000u;
        uint32_t fd[PER], fj[PER];
#pragma unroll
        for (int c = 0; c < PER; ++c) {
            fd[c] = kInfBits;
            fj[c] = 0;
        }
        auto visit = [&](int64_t j, bool all) {
            const Box bj = box_at(j);
            uint32_t bd = kInfBits, bi_ = 0;
#pragma unroll
            for (int c = 0; c < PER; ++c) {
                if (all || (j > i0 + c && j <= hi[c])) {
                    const Box u = merge_hw(bi[c], bj);
                    const uint32_t d = __float_as_uint(area(u));
                    if (d < fd[c]) {
                        fd[c] = d;
                        fj[c] = (uint32_t)j;
                    }
                    if (d < bd) {
                        bd = d;
                        bi_ = (uint32_t)(i0 + c);
                    }
                }
            }
            if (bd < kInfBits && j < tend) back_min(j, ((uint64_t)bd << 32) | bi_);
        };
        // j = i0 + 1 .. i0 + PER - 1: the windows start; then every window
        // holds j up to hi[0] (hi is non-decreasing in c); past it they end.
        const int64_t a_hi = (i0 + PER - 1 < jhi) ? i0 + PER - 1 : jhi;
        for (int64_t j = i0 + 1; j <= a_hi; ++j) visit(j, false);
        for (int64_t j = i0 + PER; j <= hi[0]; ++j) visit(j, true);
        for (int64_t j = (hi[0] + 1 > i0 + PER ? hi[0] + 1 : i0 + PER); j <= jhi; ++j) visit(j, false);
#pragma unroll
        for (int c = 0; c < PER; ++c) best[c] = fd[c] < kInfBits ? (((uint64_t)fd[c] << 32) | fj[c]) : ~0ull;
    }
    // Pairs (i, j) with i in the r clusters before the tile and j in it:
    // i = t0 - u, j = t0 + v for 1 <= u <= r, 0 <= v <= r - u.
    for (int q = threadIdx.x; q < r * r; q += kNnBlock) {
        const int64_t u = q / r + 1, v = q % r;
        const int64_t i = t0 - u, j = t0 + v;
        if (u + v > r || i < 0 || j >= (int64_t)m || j >= tend) continue;
        const uint4 sg = segs[SEG ? a.sid[par][i] : 0u];
        if (j > (int64_t)sg.x + sg.y - 1) continue;  // j is in a later mesh
        const uint64_t kb = nn_key(area(merge(box_at(i), box_at(j))), (uint32_t)i);
        if (kb != ~0ull) back_min(j, kb);
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < PER; ++c) {
        const int64_t i = i0 + c;
        if (i < (int64_t)m) {
            const uint64_t b = s_back[i - t0];
            const uint64_t v = b < best[c] ? b : best[c];
            a.nn[i] = (v == ~0ull) ? -1 : (int32_t)(uint32_t)v;
        }
    }
}

template <bool SEG, int PER>
__global__ __launch_bounds__(kNnBlock) void k_ploc_nn(PlocArgs a, int k) {
    extern __shared__ float s_nn[];  // 6 SoA arrays of ((kNnBlock * PER) + 2r) floats
    const PlocState st = a.state[k];
    if (!ploc_global_step(st)) return;
    const uint32_t m = st.m;
    const uint32_t par = st.parity;
    const int G = nn_lanes(m);
    const uint32_t tc = G == 0 ? (uint32_t)(kNnBlock * PER) : (uint32_t)(kNnBlock / G);  // clusters per tile
    const uint32_t tiles = (m + tc - 1) / tc;
    const uint32_t mtiles = (m + merge_tile(m) - 1) / merge_tile(m);
    // Reset the merge kernel's look-back words, tile counter and next state.
    for (uint32_t t = blockIdx.x * kNnBlock + threadIdx.x; t < mtiles; t += gridDim.x * kNnBlock) a.status[t] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *a.tile_ctr = 0;
        PlocState ns;
        ns.m = 0;
        ns.maxm = 0;
        ns.parity = par ^ 1u;
        ns.iters = st.iters + 1;
        ns.flags = 0;
        ns.pad[0] = ns.pad[1] = ns.pad[2] = 0;
        a.state[(k + 1) % kPlocMaxIter] = ns;
    }
    const int r = a.radius;
    const int span = (int)tc + 2 * r;
    const float4* __restrict__ bx = a.box[par];
    const uint4* __restrict__ segs = a.seg[par];
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        __syncthreads();
        const int64_t t0 = (int64_t)tile * tc;
        if (G == 0 && r <= kNnSymMaxR) {
            uint64_t* s_back = reinterpret_cast<uint64_t*>(s_nn + 6 * span);
            for (int q = threadIdx.x; q < (kNnBlock * PER); q += kNnBlock) s_back[q] = ~0ull;
        }
        for (int q = threadIdx.x; q < span; q += kNnBlock) {
            int64_t p = t0 - r + q;
            if (p >= 0 && p < (int64_t)m) {
                Box b = load_box(bx + 2 * (size_t)p);
                s_nn[0 * span + q] = b.mnx;
                s_nn[1 * span + q] = b.mny;
                s_nn[2 * span + q] = b.mnz;
                s_nn[3 * span + q] = b.mxx;
                s_nn[4 * span + q] = b.mxy;
                s_nn[5 * span + q] = b.mxz;
            }
        }
        __syncthreads();
        if (G != 0) {
            // G lanes per cluster: lane g scans window positions lo+g, lo+g+G, ...
            // and the group takes the minimum packed key.
            const int c = threadIdx.x / G, g = threadIdx.x % G;
            const int64_t i = t0 + c;
            unsigned long long best = ~0ull;
            if (i < (int64_t)m) {
                const uint4 sg = segs[SEG ? a.sid[par][i] : 0u];
                const int64_t slo = sg.x, shi = (int64_t)sg.x + sg.y - 1;
                const int64_t lo = (i - r > slo) ? i - r : slo;
                const int64_t hi = (i + r < shi) ? i + r : shi;
                const int qi = (int)(i - t0) + r;
                Box bi;
                bi.mnx = s_nn[0 * span + qi]; bi.mny = s_nn[1 * span + qi]; bi.mnz = s_nn[2 * span + qi];
                bi.mxx = s_nn[3 * span + qi]; bi.mxy = s_nn[4 * span + qi]; bi.mxz = s_nn[5 * span + qi];
                for (int64_t j = lo + g; j <= hi; j += G) {
                    if (j == i) continue;
                    const int q = (int)(j - t0) + r;
                    Box bj;
                    bj.mnx = s_nn[0 * span + q]; bj.mny = s_nn[1 * span + q]; bj.mnz = s_nn[2 * span + q];
                    bj.mxx = s_nn[3 * span + q]; bj.mxy = s_nn[4 * span + q]; bj.mxz = s_nn[5 * span + q];
                    const unsigned long long key = nn_key(area(merge(bi, bj)), (uint32_t)j);
                    best = key < best ? key : best;
                }
            }
            for (int off = G / 2; off >= 1; off >>= 1) {
                const unsigned long long o = __shfl_xor(best, off, G);
                best = o < best ? o : best;
            }
            if (g == 0 && i < (int64_t)m) a.nn[i] = (best == ~0ull) ? -1 : (int32_t)(uint32_t)best;
            continue;
        }
        if (r <= kNnSymMaxR) {
            nn_tile_symmetric<SEG, PER>(a, s_nn, span, t0, m, r, par);
            continue;
        }
        // This lane's clusters i0, i0+1, i0+2 and their windows, clipped to
        // their meshes.  Window bounds are non-decreasing in i (meshes are in
        // order), so one pass over [lo[0], hi[PER-1]] covers all three and
        // every box read from LDS serves up to three candidates.
        const int64_t i0 = t0 + (int64_t)PER * threadIdx.x;
        if (i0 >= (int64_t)m) continue;
        Box bi[PER];
        int64_t lo[PER], hi[PER];
        uint64_t best[PER];
#pragma unroll
        for (int c = 0; c < PER; ++c) {
            const int64_t i = i0 + c;
            best[c] = ~0ull;
            if (i < (int64_t)m) {
                const uint4 sg = segs[SEG ? a.sid[par][i] : 0u];
                const int64_t slo = sg.x, shi = (int64_t)sg.x + sg.y - 1;
                lo[c] = (i - r > slo) ? i - r : slo;
                hi[c] = (i + r < shi) ? i + r : shi;
                const int qi = (int)(i - t0) + r;
                bi[c].mnx = s_nn[0 * span + qi]; bi[c].mny = s_nn[1 * span + qi]; bi[c].mnz = s_nn[2 * span + qi];
                bi[c].mxx = s_nn[3 * span + qi]; bi[c].mxy = s_nn[4 * span + qi]; bi[c].mxz = s_nn[5 * span + qi];
            } else {  // past the end: an empty window inside cluster i0's (c > 0 here)
                lo[c] = hi[0] + 1;
                hi[c] = hi[0];
                bi[c] = bi[0];
            }
        }
        // lo and hi are non-decreasing in c (an empty window included).
        const int64_t jlo = lo[0];
        int64_t jhi = hi[0];
#pragma unroll
        for (int c = 1; c < PER; ++c) jhi = hi[c] > jhi ? hi[c] : jhi;
        for (int64_t j = jlo; j <= jhi; ++j) {
            const int q = (int)(j - t0) + r;
            Box bj;
            bj.mnx = s_nn[0 * span + q]; bj.mny = s_nn[1 * span + q]; bj.mnz = s_nn[2 * span + q];
            bj.mxx = s_nn[3 * span + q]; bj.mxy = s_nn[4 * span + q]; bj.mxz = s_nn[5 * span + q];
#pragma unroll
            for (int c = 0; c < PER; ++c) {
                if (j >= lo[c] && j <= hi[c] && j != i0 + c) {
                    const uint64_t key = nn_key(area(merge(bi[c], bj)), (uint32_t)j);
                    best[c] = key < best[c] ? key : best[c];
                }
            }
        }
#pragma unroll
        for (int c = 0; c < PER; ++c) {
            const int64_t i = i0 + c;
            if (i < (int64_t)m) a.nn[i] = (best[c] == ~0ull) ? -1 : (int32_t)(uint32_t)best[c];
        }
    }
}

__device__ __forceinline__ uint64_t wave_incl_scan_seg(uint64_t v) {
    const uint32_t l = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint64_t w = __shfl_up(v, o, 64);
        if (l >= (uint32_t)o) v = scomb(w, v);
    }
    return v;
}

// Exclusive block scan under scomb; *total = the block's combined value.
template <int BLOCK>
__device__ __forceinline__ uint64_t block_excl_scan_seg(uint64_t v, uint64_t* s_w, uint64_t* total) {
    constexpr int WAVES = BLOCK / kWave;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const uint64_t incl = wave_incl_scan_seg(v);
    if (lane == kWave - 1) s_w[wave] = incl;
    __syncthreads();
    uint64_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
        const uint64_t x = s_w[w];
        if (w < wave) before = scomb(before, x);
        all = scomb(all, x);
    }
    __syncthreads();
    *total = all;
    uint64_t prev = __shfl_up(incl, 1, 64);
    if (lane == 0) prev = 0;
    return scomb(before, prev);
}

// Per-tile LDS of the merge: flags, partners, mesh ids, output slots and
// new node ids of the tile's clusters, plus the scan's scratch.
template <int ITEMS>
struct MergeLds {
    int32_t nj[kMergeBlock * ITEMS];
    uint32_t sid[kMergeBlock * ITEMS];
    uint32_t pos[kMergeBlock * ITEMS];
    uint32_t nid[kMergeBlock * ITEMS];
    uint8_t fl[kMergeBlock * ITEMS];
    uint64_t w[kMergeBlock / kWave];
    uint64_t prefix;
    uint32_t tile;
};

// Merge + compaction over tiles of kMergeBlock * ITEMS clusters (the tile
// size is a function of m, shared with the look-back reset in k_ploc_nn).
// Global reads and writes go one lane per cluster (coalesced); the scan,
// which needs each thread's clusters consecutive, works on the tile's flags
// in LDS and hands output slots back through LDS.
template <bool SEG, int ITEMS>
__device__ void merge_tiles(const PlocArgs& a, const PlocState& st, int kn, MergeLds<ITEMS>& sm) {
    constexpr int kMergeItems = ITEMS;
    constexpr uint32_t kMergeTile = kMergeBlock * ITEMS;
    const uint32_t m = st.m;
    const uint32_t par = st.parity;
    const uint32_t tiles = (m + kMergeTile - 1) / kMergeTile;
    const int32_t* __restrict__ nn = a.nn;
    const uint4* __restrict__ segs = a.seg[par];
    uint4* __restrict__ segs_out = a.seg[par ^ 1u];
    const float4* __restrict__ bin = a.box[par];
    const uint32_t* __restrict__ iin = a.id[par];
    float4* __restrict__ bout = a.box[par ^ 1u];
    uint32_t* __restrict__ iout = a.id[par ^ 1u];
    const int tid = threadIdx.x;
    while (true) {
        __syncthreads();
        if (tid == 0) sm.tile = atomicAdd(a.tile_ctr, 1u);
        __syncthreads();
        const uint32_t tile = sm.tile;
        if (tile >= tiles) break;
        const uint32_t tbase = tile * kMergeTile;

        // Phase A, one lane per cluster (coalesced): mutual pairs and flags.
#pragma unroll
        for (int u = 0; u < kMergeItems; ++u) {
            const uint32_t q = u * kMergeBlock + tid;
            const uint32_t i = tbase + q;
            uint32_t f = 0;
            if (i < m) {
                const uint32_t s = SEG ? a.sid[par][i] : 0u;
                const uint4 sg = segs[s];
                const int32_t j = nn[i];
                const bool mutual = (j >= 0) && nn[j] == (int32_t)i;
                const bool merging = mutual && (int32_t)i < j;
                const bool valid = !(mutual && j < (int32_t)i);
                const bool first = (i == sg.x);
                const bool last = (i == sg.x + sg.y - 1);
                f = (valid ? 1u : 0u) | (merging ? 2u : 0u) | (first ? 4u : 0u) | (last ? 8u : 0u);
                sm.nj[q] = j;
                if (SEG) sm.sid[q] = s;
            }
            sm.fl[q] = (uint8_t)f;
        }
        __syncthreads();

        // Phase B, ITEMS consecutive clusters per thread: segmented scan of
        // (mesh start, merges, survivors) across the tile and the look-back.
        const uint32_t base = tid * kMergeItems;
        uint32_t fl[kMergeItems];
        uint64_t agg = 0;
#pragma unroll
        for (int it = 0; it < kMergeItems; ++it) {
            fl[it] = sm.fl[base + it];
            if (tbase + base + it < m) {
                const uint32_t f = fl[it];
                const uint64_t e = ((f & 4u) ? kSegF : 0ull) | ((uint64_t)((f >> 1) & 1u) << 31) | (uint64_t)(f & 1u);
                agg = scomb(agg, e);
            }
        }
        uint64_t tagg;
        const uint64_t excl = block_excl_scan_seg<kMergeBlock>(agg, sm.w, &tagg);
        // Decoupled look-back by wave 0, 64 predecessors' words per step; the
        // window is combined in tile order (farther first) under scomb.
        if (tid < kWave) {
            const int lane = tid;
            uint64_t* sw = a.status + tile;
            if (lane == 0)
                __hip_atomic_store(sw, (tile == 0 ? kStP : kStA) | tagg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint64_t pre = 0;
            int top = (int)tile - 1;  // nearest predecessor not yet combined
            while (top >= 0) {
                const int j = top - lane;
                uint64_t w = kStP;  // before tile 0: an inclusive prefix of 0 (never combined past a P)
                if (j >= 0) w = __hip_atomic_load(a.status + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint64_t f = w & kStFlags;
                const uint64_t pmask = __ballot(f == kStP);
                const uint64_t nready = __ballot(f == 0);
                const int first_p = pmask ? __ffsll((long long)pmask) - 1 : kWave - 1;
                const uint64_t need = (first_p == kWave - 1) ? ~0ull : ((2ull << first_p) - 1);
                if (nready & need) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                uint64_t v = (lane <= first_p) ? (w & ~kStFlags) : 0ull;
#pragma unroll
                for (int o = 1; o < kWave; o <<= 1) {
                    const uint64_t other = __shfl_down(v, o, kWave);
                    if (lane + o < kWave) v = scomb(other, v);
                }
                pre = scomb(__shfl(v, 0, kWave), pre);
                if (pmask) break;
                top -= kWave;
            }
            if (lane == 0) {
                if (tile > 0) __hip_atomic_store(sw, kStP | scomb(pre, tagg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                sm.prefix = pre;
                if (tile == tiles - 1) a.state[kn].m = (uint32_t)(scomb(pre, tagg) & kVMask);
            }
        }
        __syncthreads();
        // Each cluster's output slot (survivors) and new node id (merging
        // ones) into LDS; mesh starts and ends update the mesh records.
        const uint64_t run0 = scomb(sm.prefix, excl);
        uint32_t pos = (uint32_t)(run0 & kVMask);
        uint32_t mrank = (uint32_t)((run0 & kMMask) >> 31);
#pragma unroll
        for (int it = 0; it < kMergeItems; ++it) {
            const uint32_t i = tbase + base + it;
            if (i >= m) continue;
            const uint32_t f = fl[it];
            const uint32_t s = SEG ? sm.sid[base + it] : 0u;
            if (f & 4u) {
                mrank = 0;
                segs_out[s].x = pos;
            }
            if (f & 1u) {
                sm.pos[base + it] = pos;
                if (f & 2u) {
                    sm.nid[base + it] = segs[s].z + mrank;
                    ++mrank;
                }
                ++pos;
            }
            if (f & 8u) {
                const uint4 sg = segs[s];
                const uint32_t nm = sg.y - mrank;
                segs_out[s].y = nm;
                segs_out[s].z = sg.z + mrank;
                segs_out[s].w = sg.w + (sg.y >= 2 ? 1u : 0u);
                atomicMax(&a.state[kn].maxm, nm);
                if (sg.y >= 2 && mrank == 0) atomicOr(&a.state[kn].flags, kPlocNoProgress);
            }
        }
        __syncthreads();

        // Phase C, one lane per cluster again (coalesced): new nodes and the
        // compacted survivors.
#pragma unroll
        for (int u = 0; u < kMergeItems; ++u) {
            const uint32_t q = u * kMergeBlock + tid;
            const uint32_t i = tbase + q;
            if (i >= m) continue;
            const uint32_t f = sm.fl[q];
            if (!(f & 1u)) continue;
            const uint32_t s = SEG ? sm.sid[q] : 0u;
            const uint32_t p = sm.pos[q];
            Box b = load_box(bin + 2 * (size_t)i);
            uint32_t cid = iin[i];
            if (f & 2u) {
                const SegDesc& d = seg_desc(a.sb, s);
                const uint32_t root = 2 * d.n - 2;
                const uint32_t j = (uint32_t)sm.nj[q];
                const Box bj = load_box(bin + 2 * (size_t)j);
                const uint32_t idj = iin[j];
                b = merge(b, bj);
                const uint32_t nid = sm.nid[q];
                if (nid <= root)
                    store_node(d.nodes, remap_id(nid, root), b, (int32_t)remap_id(cid, root),
                               (int32_t)remap_id(idj, root));
                cid = nid;
            }
            store_box(bout + 2 * (size_t)p, b);
            iout[p] = cid;
            if (SEG) a.sid[par ^ 1u][p] = s;
        }
    }
}

template <bool SEG>
__global__ __launch_bounds__(kMergeBlock) void k_ploc_merge(PlocArgs a, int k) {
    __shared__ MergeLds<4> s_m;  // also holds the one-cluster-per-thread tile
    const PlocState st = a.state[k];
    const int kn = (k + 1) % kPlocMaxIter;
    if (!ploc_global_step(st)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) a.state[kn] = st;
        return;
    }
    if (merge_tile(st.m) == 4u * kMergeBlock) merge_tiles<SEG, 4>(a, st, kn, s_m);
    else merge_tiles<SEG, 1>(a, st, kn, *reinterpret_cast<MergeLds<1>*>(&s_m));
}

// Finishes every remaining iteration of one mesh (block = mesh) in LDS.
__global__ __launch_bounds__(kTailBlock) void k_ploc_tail(PlocArgs a, int k) {
    constexpr int PER = (kPlocSmall + kTailBlock - 1) / kTailBlock;
    __shared__ float s_b[6][kPlocSmall];
    __shared__ uint32_t s_id[kPlocSmall];
    // Packed neighbour minimum per cluster; the neighbour ids (s_nn) reuse its
    // storage once the minima are read (keeps two workgroups per CU).
    __shared__ uint64_t s_best[kPlocSmall];
    int32_t* const s_nn = reinterpret_cast<int32_t*>(s_best);
    static_assert(kPlocSmall == 2 * kTailBlock, "each thread converts at most two minima");
    __shared__ uint64_t s_w[kTailBlock / kWave];
    const uint32_t seg = blockIdx.x;
    const PlocState st = a.state[k];
    const uint4 sg = a.seg[st.parity][seg];
    const SegDesc& d = seg_desc(a.sb, seg);
    uint32_t m = sg.y;
    const uint32_t root = 2 * d.n - 2;
    const int r = a.radius;
    uint32_t node_base = sg.z;
    uint32_t iters = sg.w;
    uint32_t flags = st.flags;
    if (st.maxm > (uint32_t)kPlocSmall) flags |= kPlocUnfinished;
    if (flags || m > (uint32_t)kPlocSmall) {
        if (threadIdx.x == 0) a.out[seg] = make_uint4(m, node_base, iters, flags);
        return;
    }
    const float4* __restrict__ bx = a.box[st.parity] + 2 * (size_t)sg.x;
    const uint32_t* __restrict__ ids = a.id[st.parity] + sg.x;
    for (uint32_t i = threadIdx.x; i < m; i += kTailBlock) {
        Box b = load_box(bx + 2 * (size_t)i);
        s_b[0][i] = b.mnx; s_b[1][i] = b.mny; s_b[2][i] = b.mnz;
        s_b[3][i] = b.mxx; s_b[4][i] = b.mxy; s_b[5][i] = b.mxz;
        s_id[i] = ids[i];
    }
    __syncthreads();
    while (m > (uint32_t)kWave) {
        // Lanes per cluster: as m shrinks, each window is split over up to 64
        // lanes (strided candidates, then a shuffle min over the lane group;
        // the packed key's minimum is unique, so the split is exact).  Each
        // pair i < j of a window is evaluated once, by i's lanes (the distance
        // is symmetric bit for bit, as in k_ploc_nn): (d, j) goes into i's
        // minimum, (d, i) into j's through a 64-bit LDS atomic minimum.
        uint32_t L = 1;
        while (L < (uint32_t)kWave && m * (2 * L) <= (uint32_t)kTailBlock) L <<= 1;
        const uint32_t sub = threadIdx.x & (L - 1);
        for (uint32_t i = threadIdx.x; i < m; i += kTailBlock) s_best[i] = ~0ull;
        __syncthreads();
        for (uint32_t i = threadIdx.x / L; i < m; i += kTailBlock / L) {
            Box bi;
            bi.mnx = s_b[0][i]; bi.mny = s_b[1][i]; bi.mnz = s_b[2][i];
            bi.mxx = s_b[3][i]; bi.mxy = s_b[4][i]; bi.mxz = s_b[5][i];
            const int hi = ((int)i + r < (int)m - 1) ? (int)i + r : (int)m - 1;
            uint64_t best = ~0ull;
            for (int j = (int)i + 1 + (int)sub; j <= hi; j += (int)L) {
                Box bj;
                bj.mnx = s_b[0][j]; bj.mny = s_b[1][j]; bj.mnz = s_b[2][j];
                bj.mxx = s_b[3][j]; bj.mxy = s_b[4][j]; bj.mxz = s_b[5][j];
                const float dist = area(merge(bi, bj));
                const uint64_t kf = nn_key(dist, (uint32_t)j);
                best = kf < best ? kf : best;
                const uint64_t kb = nn_key(dist, i);
                if (kb != ~0ull)
                    __hip_atomic_fetch_min(reinterpret_cast<unsigned long long*>(s_best + j), (unsigned long long)kb,
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            for (uint32_t o = L >> 1; o >= 1; o >>= 1) {
                const uint64_t other = __shfl_xor(best, (int)o, kWave);
                best = other < best ? other : best;
            }
            if (sub == 0 && best != ~0ull)
                __hip_atomic_fetch_min(reinterpret_cast<unsigned long long*>(s_best + i), (unsigned long long)best,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        __syncthreads();
        {
            const uint32_t i0 = threadIdx.x, i1 = threadIdx.x + kTailBlock;
            const uint64_t v0 = i0 < m ? s_best[i0] : ~0ull;
            const uint64_t v1 = i1 < m ? s_best[i1] : ~0ull;
            __syncthreads();  // every minimum read before s_nn overwrites them
            if (i0 < m) s_nn[i0] = (v0 == ~0ull) ? -1 : (int32_t)(uint32_t)v0;
            if (i1 < m) s_nn[i1] = (v1 == ~0ull) ? -1 : (int32_t)(uint32_t)v1;
        }
        __syncthreads();
        // Per thread PER consecutive slots; scan of packed (merges, survivors).
        uint32_t fl[PER];
        Box bb[PER];
        uint32_t cid[PER];
        uint32_t cv = 0, cm = 0;
#pragma unroll
        for (int it = 0; it < PER; ++it) {
            const uint32_t i = threadIdx.x * PER + it;
            fl[it] = 0;
            if (i < m) {
                const int32_t j = s_nn[i];
                const bool mutual = (j >= 0) && s_nn[j] == (int32_t)i;
                const bool merging = mutual && (int32_t)i < j;
                const bool valid = !(mutual && j < (int32_t)i);
                fl[it] = (valid ? 1u : 0u) | (merging ? 2u : 0u);
                cv += valid ? 1u : 0u;
                cm += merging ? 1u : 0u;
                bb[it].mnx = s_b[0][i]; bb[it].mny = s_b[1][i]; bb[it].mnz = s_b[2][i];
                bb[it].mxx = s_b[3][i]; bb[it].mxy = s_b[4][i]; bb[it].mxz = s_b[5][i];
                cid[it] = s_id[i];
                if (merging) {
                    Box bj;
                    bj.mnx = s_b[0][j]; bj.mny = s_b[1][j]; bj.mnz = s_b[2][j];
                    bj.mxx = s_b[3][j]; bj.mxy = s_b[4][j]; bj.mxz = s_b[5][j];
                    const uint32_t idj = s_id[j];
                    bb[it] = merge(bb[it], bj);
                    cid[it] = idj;  // temporarily the right child id
                    fl[it] |= (cm - 1) << 2;  // rank of this merge within the thread
                }
            }
        }
        // Plain (unsegmented) counts packed as merges << 31 | survivors.
        uint64_t tot;
        const uint64_t ex = block_excl_scan_seg<kTailBlock>(((uint64_t)cm << 31) | cv, s_w, &tot);
        uint32_t pos = (uint32_t)(ex & kVMask);
        const uint32_t mbase = (uint32_t)((ex & kMMask) >> 31);
        // All LDS reads of this iteration are done (the scan synchronised).
#pragma unroll
        for (int it = 0; it < PER; ++it) {
            const uint32_t i = threadIdx.x * PER + it;
            if (i >= m || !(fl[it] & 1u)) continue;
            uint32_t out_id = s_id[i];
            if (fl[it] & 2u) {
                const uint32_t nid = node_base + mbase + (fl[it] >> 2);
                if (nid <= root)
                    store_node(d.nodes, remap_id(nid, root), bb[it], (int32_t)remap_id(out_id, root),
                               (int32_t)remap_id(cid[it], root));
                out_id = nid;
            }
            cid[it] = out_id;
            fl[it] = (fl[it] & 3u) | (pos << 2);
            ++pos;
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < PER; ++it) {
            const uint32_t i = threadIdx.x * PER + it;
            if (i >= m || !(fl[it] & 1u)) continue;
            const uint32_t p = fl[it] >> 2;
            s_b[0][p] = bb[it].mnx; s_b[1][p] = bb[it].mny; s_b[2][p] = bb[it].mnz;
            s_b[3][p] = bb[it].mxx; s_b[4][p] = bb[it].mxy; s_b[5][p] = bb[it].mxz;
            s_id[p] = cid[it];
        }
        __syncthreads();
        const uint32_t merges = (uint32_t)((tot & kMMask) >> 31);
        m = (uint32_t)(tot & kVMask);
        node_base += merges;
        ++iters;
        if (merges == 0) {
            flags |= kPlocNoProgress;
            break;
        }
    }
    // At most 64 clusters left: wave 0 finishes alone, wave-synchronously (no
    // workgroup barriers).  Same order of merges, ids and survivors as above.
    if (threadIdx.x < kWave && !(flags & kPlocNoProgress)) {
        const uint32_t lane = threadIdx.x;
        while (m > 1) {
            uint32_t L = 1;
            while (L < (uint32_t)kWave && m * (2 * L) <= (uint32_t)kWave) L <<= 1;
            const uint32_t sub = lane & (L - 1);
            const uint32_t ci = lane / L;
            uint64_t best = ~0ull;
            if (ci < m) {
                Box bi;
                bi.mnx = s_b[0][ci]; bi.mny = s_b[1][ci]; bi.mnz = s_b[2][ci];
                bi.mxx = s_b[3][ci]; bi.mxy = s_b[4][ci]; bi.mxz = s_b[5][ci];
                const int lo = ((int)ci - r > 0) ? (int)ci - r : 0;
                const int hi = ((int)ci + r < (int)m - 1) ? (int)ci + r : (int)m - 1;
                for (int j = lo + (int)sub; j <= hi; j += (int)L) {
                    if (j == (int)ci) continue;
                    Box bj;
                    bj.mnx = s_b[0][j]; bj.mny = s_b[1][j]; bj.mnz = s_b[2][j];
                    bj.mxx = s_b[3][j]; bj.mxy = s_b[4][j]; bj.mxz = s_b[5][j];
                    const uint64_t key = nn_key(area(merge(bi, bj)), (uint32_t)j);
                    best = key < best ? key : best;
                }
            }
            for (uint32_t o = L >> 1; o >= 1; o >>= 1) {
                const uint64_t other = __shfl_xor(best, (int)o, kWave);
                best = other < best ? other : best;
            }
            if (ci < m && sub == 0) s_nn[ci] = (best == ~0ull) ? -1 : (int32_t)(uint32_t)best;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // One lane per cluster: mutual pairs merge into the lower slot.
            const uint32_t i = lane;
            bool valid = false, merging = false;
            Box bb;
            uint32_t id_i = 0, id_j = 0;
            if (i < m) {
                const int32_t j = s_nn[i];
                const bool mutual = (j >= 0) && s_nn[j] == (int32_t)i;
                merging = mutual && (int32_t)i < j;
                valid = !(mutual && j < (int32_t)i);
                bb.mnx = s_b[0][i]; bb.mny = s_b[1][i]; bb.mnz = s_b[2][i];
                bb.mxx = s_b[3][i]; bb.mxy = s_b[4][i]; bb.mxz = s_b[5][i];
                id_i = s_id[i];
                if (merging) {
                    Box bj;
                    bj.mnx = s_b[0][j]; bj.mny = s_b[1][j]; bj.mnz = s_b[2][j];
                    bj.mxx = s_b[3][j]; bj.mxy = s_b[4][j]; bj.mxz = s_b[5][j];
                    id_j = s_id[j];
                    bb = merge(bb, bj);
                }
            }
            const uint64_t vmask = __ballot(valid), mmask = __ballot(merging);
            const uint64_t below = lanemask_lt();
            const uint32_t pos = (uint32_t)__popcll(vmask & below);
            uint32_t out_id = id_i;
            if (merging) {
                const uint32_t nid = node_base + (uint32_t)__popcll(mmask & below);
                if (nid <= root)
                    store_node(d.nodes, remap_id(nid, root), bb, (int32_t)remap_id(id_i, root),
                               (int32_t)remap_id(id_j, root));
                out_id = nid;
            }
            // Every lane has read its inputs before any slot is overwritten.
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (valid) {
                s_b[0][pos] = bb.mnx; s_b[1][pos] = bb.mny; s_b[2][pos] = bb.mnz;
                s_b[3][pos] = bb.mxx; s_b[4][pos] = bb.mxy; s_b[5][pos] = bb.mxz;
                s_id[pos] = out_id;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const uint32_t merges = (uint32_t)__popcll(mmask);
            m = (uint32_t)__popcll(vmask);
            node_base += merges;
            ++iters;
            if (merges == 0) {
                flags |= kPlocNoProgress;
                break;
            }
        }
    }
    if (threadIdx.x == 0) a.out[seg] = make_uint4(m, node_base, iters, flags);
}

void launch_ploc_init(const PlocArgs& a, const uint64_t* keys, const KeyFormat& kf, const float4* tri_boxes,
                      uint32_t* err, hipStream_t s) {
    hipLaunchKernelGGL(k_ploc_init, dim3(a.sb.nblocks), dim3(256), 0, s, a, keys, kf, tri_boxes, err);
}

void launch_ploc_nn(const PlocArgs& a, int k, uint32_t m_bound, uint32_t max_grid, hipStream_t s) {
    uint32_t g = (m_bound + kNnBlock / 8 - 1) / (kNnBlock / 8);  // enough for the finest split
    if (g > max_grid) g = max_grid;
    g = g < 1 ? 1 : (g > kNnGridMax ? kNnGridMax : g);
    // Six SoA float arrays of the staged span, then the tile's backward
    // minima (symmetric search; 6 * span floats keep them 8-byte aligned).
    const int per = a.radius >= kNnWideR ? kNnPerWide : kNnPer;
    const size_t tile = (size_t)kNnBlock * per;
    size_t lds = (size_t)6 * (tile + 2 * a.radius) * sizeof(float) + tile * sizeof(uint64_t);
    if (per == kNnPerWide) {
        if (a.sid[0]) hipLaunchKernelGGL((k_ploc_nn<true, kNnPerWide>), dim3(g), dim3(kNnBlock), lds, s, a, k);
        else hipLaunchKernelGGL((k_ploc_nn<false, kNnPerWide>), dim3(g), dim3(kNnBlock), lds, s, a, k);
    } else {
        if (a.sid[0]) hipLaunchKernelGGL((k_ploc_nn<true, kNnPer>), dim3(g), dim3(kNnBlock), lds, s, a, k);
        else hipLaunchKernelGGL((k_ploc_nn<false, kNnPer>), dim3(g), dim3(kNnBlock), lds, s, a, k);
    }
}

void launch_ploc_merge(const PlocArgs& a, int k, uint32_t m_bound, uint32_t max_grid, hipStream_t s) {
    uint32_t g = (m_bound + kMergeBlock - 1) / kMergeBlock;
    if (g > max_grid) g = max_grid;
    g = g < 1 ? 1 : (g > kMergeGridMax ? kMergeGridMax : g);
    if (a.sid[0]) hipLaunchKernelGGL(k_ploc_merge<true>, dim3(g), dim3(kMergeBlock), 0, s, a, k);
    else hipLaunchKernelGGL(k_ploc_merge<false>, dim3(g), dim3(kMergeBlock), 0, s, a, k);
}

void launch_ploc_tail(const PlocArgs& a, int k, hipStream_t s) {
    hipLaunchKernelGGL(k_ploc_tail, dim3(a.nseg), dim3(kTailBlock), 0, s, a, k);
}

}  // namespace bvhk
