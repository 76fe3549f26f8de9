// k_hier.hip — LBVH topology, leaf initialisation and node boxes.
//
// Replaces kBuildInternalNodes_nt/_P + kInitLeafNodes_nt/_P + kRefitHierarchy_nt
// (reference lbvh_builder.cu:183-274, lbvh_plus_builder.cu:292-335,341-366 with
// doOptimize = false).
//
// The reference finds each internal node top-down (Karras 2012: direction,
// exponential + binary range search, binary split search; O(log n) scattered
// key reads per node), then refits bottom-up in a second kernel through
// per-node atomic counters and __threadfence.  The tree is the Cartesian tree
// of the boundary deltas (delta(j) = common-prefix length of keys j and j+1,
// -1 past either end of a mesh): the node whose split is boundary s covers
// the leaves between the nearest boundary on each side with a smaller delta
// (the minimum inside any node's range is unique), and it is the left child
// of its parent iff its right delimiter's delta is the larger one.  Karras'
// numbering follows: an internal node is numbered by its right end when it
// is a left child and by its left end otherwise (the root is 0); leaf k is
// node n-1+k.  A node's box is the reference's left-then-right merge of its
// children, i.e. the left fold of its leaves' boxes with select min/max;
// that fold is associative (it keeps the leftmost of equal values, +0/-0
// included), so any evaluation that keeps operands in leaf order gives the
// reference's bits.
//
// MI355X mapping, level by level.  A level sees a sequence of units (level 1:
// the sorted leaves; later levels: finished subtrees) and cuts it into chunks
// of 1024 units, one 1024-thread workgroup each, with everything in LDS:
//   * a sparse table of the chunk's boundary deltas: each inner boundary finds
//     its nearest smaller delta on both sides by binary lifting (11 LDS reads
//     per side, all threads at once, no climbing chains);
//   * a segment tree of the units' boxes in heap order: a node's box is a
//     left-to-right query over its unit range;
//   * a boundary whose nearest smaller deltas both lie in the chunk yields a
//     finished node (record written once, plain stores); the others (and the
//     mesh walls, delta -1) cut the chunk into segments, each a finished
//     subtree whose parent lies outside the chunk.  Those segments are the
//     next level's units, compacted in order through decoupled look-back over
//     epoch-tagged status words (no clearing, so a build is graph-capturable).
// Each level shrinks the unit count ~50-100x on Morton-ordered data (4M blob
// triangles: 41,822 units after level 1, 623 after level 2, none after level
// 3).  Whatever remains after the last level (pathological delta patterns,
// e.g. a caterpillar) is finished by a climbing pass: one thread per unit,
// epoch-tagged 64-bit handoff words in HBM, records written with agent-scope
// (sc1) stores and read with agent-scope loads after a full wait, so no L2
// writeback or invalidate is needed.  Level-1 chunks never span meshes; later
// chunks may, with the mesh walls (delta -1) keeping meshes apart, so a batch
// of meshes builds with the same launches.
#include "dev_common.hpp"
#include "engine.hpp"

namespace bvhk {

constexpr int kHierThreads = 1024;       // one unit per thread
constexpr int kChunk = kHierThreads;     // units per chunk
constexpr int kLog = 10;                 // log2(kChunk)
constexpr int kHierMaxLevels = 6;
constexpr uint32_t kFlagA = 1u << 30;    // look-back: aggregate published
constexpr uint32_t kFlagP = 2u << 30;    // look-back: inclusive prefix published
constexpr uint32_t kValMask = (1u << 30) - 1;
constexpr float kBig = 3.402823466e+38f;

int hier_chunk() { return kChunk; }
int hier_unit_bytes() { return (int)sizeof(HUnit); }
int hier_max_levels() { return kHierMaxLevels; }

template <bool K60>
__device__ __forceinline__ int key_delta(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                         uint32_t j) {
    uint64_t x = keys[j];
    uint64_t y = keys[j + 1];
    if (K60 && x == y) {
        uint32_t v = vals[j] ^ vals[j + 1];
        return 96 + (v == 0 ? 32 : __clz((int)v));
    }
    uint64_t z = x ^ y;
    return z == 0 ? 64 : __clzll((long long)z);
}

__device__ __forceinline__ SegDesc seg_desc(const SegBlocks& sb, uint32_t seg) {
    return sb.segs ? sb.segs[seg] : sb.single;
}

// Packed deltas of a unit: (delta before it + 1) | (delta after it + 1) << 8.
__device__ __forceinline__ uint32_t pack_beta(int bl, int br) {
    return (uint32_t)(bl + 1) | ((uint32_t)(br + 1) << 8);
}
__device__ __forceinline__ int beta_left(uint32_t w) { return (int)(w & 0xFFu) - 1; }
__device__ __forceinline__ int beta_right(uint32_t w) { return (int)((w >> 8) & 0xFFu) - 1; }

// Node index of a finished subtree covering leaves [l, r] with delimiting
// deltas bl (boundary l-1) and br (boundary r); never the root.
__device__ __forceinline__ uint32_t subtree_index(uint32_t n, uint32_t l, uint32_t r, int bl, int br) {
    if (l == r) return n - 1 + l;
    return (br > bl) ? r : l;
}

struct HierSmem {
    float box[6][2 * kChunk];           // segment tree of unit boxes, heap order, units at kChunk + u
    int8_t st[kLog + 1][kChunk + 1];    // st[k][q] = min of boundary deltas q .. q + 2^k - 1
    uint32_t ul[kChunk], ur[kChunk];    // unit leaf ranges
    uint32_t us[kChunk];                // unit mesh
    uint32_t start[kChunk / 32];        // bit u: a segment starts at unit u
    uint32_t src[kChunk];               // level 1 (first chunked level): slot of unit u in level 0's output
    uint32_t wsum[kHierThreads / kWave];
    uint32_t ticket, out_base;
};

__device__ __forceinline__ Box tree_box(const HierSmem& s, int h) {
    Box b;
    b.mnx = s.box[0][h]; b.mny = s.box[1][h]; b.mnz = s.box[2][h];
    b.mxx = s.box[3][h]; b.mxy = s.box[4][h]; b.mxz = s.box[5][h];
    return b;
}

// Box of units [x, y], merged left to right (the reference's fold order).
__device__ Box range_box(const HierSmem& s, int x, int y) {
    Box lo_acc, hi_acc;
    lo_acc.mnx = lo_acc.mny = lo_acc.mnz = kBig;
    lo_acc.mxx = lo_acc.mxy = lo_acc.mxz = -kBig;
    hi_acc = lo_acc;
    int lo = x + kChunk, hi = y + kChunk + 1;
    while (lo < hi) {
        if (lo & 1) lo_acc = merge(lo_acc, tree_box(s, lo++));
        if (hi & 1) hi_acc = merge(tree_box(s, --hi), hi_acc);
        lo >>= 1;
        hi >>= 1;
    }
    return merge(lo_acc, hi_acc);
}

// One chunk of one level; the block's ticket picks the chunk.  LEAF: the
// leaf level (units are the sorted leaves of one mesh chunk).  Otherwise
// `total_in` units come from p.units_in, compacted, except for a folded level
// 1 (p.fold1), which reads the leaf level's per-chunk slots (p.counts[c] exits
// at c * kChunk) through a slot map it builds itself (no compaction pass).
template <bool LEAF, bool K60>
__device__ void hier_chunk_body(const HierArgs& p, HierSmem& s, uint32_t t, uint32_t nchunks, int level,
                                uint32_t total_in = 0) {
    const int tid = threadIdx.x;
    const int u = tid;
    const uint32_t epoch = *p.epoch;

    // ---- units: leaf range, mesh, box, delimiting deltas --------------------
    uint32_t m;            // units in this chunk
    uint32_t seg0 = 0;     // level 1: the chunk's mesh
    SegDesc d0;            // level 1: its descriptor
    if (tid < kChunk / 32) s.start[tid] = 0;
    if (LEAF) {
        uint32_t chunk;
        if (p.sb.blocks) {
            const uint2 e = p.sb.blocks[t];
            seg0 = e.x;
            chunk = e.y;
        } else {
            chunk = t;
        }
        d0 = seg_desc(p.sb, seg0);
        const uint32_t n = d0.n;
        const uint32_t a = chunk * kChunk;
        m = (n - a < (uint32_t)kChunk) ? n - a : (uint32_t)kChunk;
        const uint64_t* __restrict__ keys = p.keys + d0.start;
        const uint32_t* __restrict__ vals = K60 ? p.vals + d0.start : nullptr;
        const uint64_t idx_mask = (p.kf.idx_bits >= 32) ? 0xFFFFFFFFull : ((1ull << p.kf.idx_bits) - 1);
        Box b;
        b.mnx = b.mny = b.mnz = kBig;
        b.mxx = b.mxy = b.mxz = -kBig;
        int dr = -1;
        if ((uint32_t)u < m) {
            const uint32_t i = a + u;
            uint32_t prim = K60 ? vals[i] : (uint32_t)(keys[i] & idx_mask);
            if (prim >= n) {  // cannot happen for a correct sort; never index out of bounds
                atomicOr(p.err, 1u);
                prim = 0;
            }
            d0.indices[i] = prim;
            b = load_box(p.tri_boxes + 2 * ((size_t)d0.start + prim));
            store_node(d0.nodes, n - 1 + i, b, (int32_t)(prim | 0x80000000u), -1);
            dr = (i + 1 < n) ? key_delta<K60>(keys, vals, i) : -1;
            s.ul[u] = i;
            s.ur[u] = i;
            s.us[u] = seg0;
        }
        s.st[0][u + 1] = (int8_t)dr;
        if (tid == 0) s.st[0][0] = (int8_t)((a > 0) ? key_delta<K60>(keys, vals, a - 1) : -1);
        s.box[0][kChunk + u] = b.mnx; s.box[1][kChunk + u] = b.mny; s.box[2][kChunk + u] = b.mnz;
        s.box[3][kChunk + u] = b.mxx; s.box[4][kChunk + u] = b.mxy; s.box[5][kChunk + u] = b.mxz;
    } else {
        const uint32_t cnt = total_in;
        const uint32_t base = t * kChunk;
        m = (cnt - base < (uint32_t)kChunk) ? cnt - base : (uint32_t)kChunk;
        if (level == 1 && p.fold1) {
            // Slots of units [base, base + m): each thread owns a run of
            // leaf-level chunks; an exclusive scan of their exit counts gives
            // the run's first unit index.
            const uint32_t nch = p.sb.nblocks;
            const uint32_t cp = (nch + kHierThreads - 1) / kHierThreads;
            const uint32_t c0 = (uint32_t)tid * cp < nch ? (uint32_t)tid * cp : nch;
            const uint32_t c1 = c0 + cp < nch ? c0 + cp : nch;
            uint32_t run = 0;
            for (uint32_t c = c0; c < c1; ++c) run += p.counts[c];
            const uint32_t incl = wave_incl_sum(run);
            if (lane_id() == kWave - 1) s.wsum[tid / kWave] = incl;
            __syncthreads();
            uint32_t off = incl - run;
            for (int w = 0; w < tid / kWave; ++w) off += s.wsum[w];
            for (uint32_t c = c0; c < c1 && off < base + m; ++c) {
                const uint32_t cc = p.counts[c];
                const uint32_t lo = off > base ? off : base;
                const uint32_t hi = (off + cc < base + m) ? off + cc : base + m;
                for (uint32_t g = lo; g < hi; ++g) s.src[g - base] = c * kChunk + (g - off);
                off += cc;
            }
            __syncthreads();
        }
        float4 w0 = make_float4(kBig, kBig, kBig, -kBig), w1 = make_float4(-kBig, -kBig, 0.0f, 0.0f);
        int dr = -1;
        if ((uint32_t)u < m) {
            const HUnit* src = p.units_in + ((level == 1 && p.fold1) ? s.src[u] : base + u);
            const uint4 h = *reinterpret_cast<const uint4*>(src);
            w0 = reinterpret_cast<const float4*>(src)[1];
            w1 = reinterpret_cast<const float4*>(src)[2];
            s.ul[u] = h.x;
            s.ur[u] = h.y;
            s.us[u] = h.z;
            dr = beta_right(h.w);
            if (u == 0) s.st[0][0] = (int8_t)beta_left(h.w);
        }
        s.st[0][u + 1] = (int8_t)dr;
        s.box[0][kChunk + u] = w0.x; s.box[1][kChunk + u] = w0.y; s.box[2][kChunk + u] = w0.z;
        s.box[3][kChunk + u] = w0.w; s.box[4][kChunk + u] = w1.x; s.box[5][kChunk + u] = w1.y;
    }
    __syncthreads();

    // ---- sparse table of deltas and segment tree of boxes: two levels per
    // round, 5 rounds.  Table entries whose window runs past boundary kChunk
    // are never read (the lifting below stays inside [0, m]); their inputs are
    // clamped to the row.
    static_assert(kLog % 2 == 0, "levels are built in pairs");
#pragma unroll 1
    for (int k = 1; k <= kLog; k += 2) {
        const int w = 1 << (k - 1);
        {
            const int q = tid;
            auto at = [&](int j) -> int8_t { return s.st[k - 1][(j < kChunk) ? j : kChunk]; };
            const int8_t x0 = at(q), x1 = at(q + w), x2 = at(q + 2 * w), x3 = at(q + 3 * w);
            const int8_t lo = (x1 < x0) ? x1 : x0;
            const int8_t hi = (x3 < x2) ? x3 : x2;
            s.st[k][q] = lo;
            s.st[k + 1][q] = (hi < lo) ? hi : lo;
        }
        const int lvl_n = kChunk >> (k + 1);  // nodes on level k + 1
        if (tid < lvl_n) {
            const int h = lvl_n + tid;
            const Box c0 = merge(tree_box(s, 4 * h), tree_box(s, 4 * h + 1));
            const Box c1 = merge(tree_box(s, 4 * h + 2), tree_box(s, 4 * h + 3));
            const Box c = merge(c0, c1);
            const int h0 = 2 * h, h1 = 2 * h + 1;
            s.box[0][h0] = c0.mnx; s.box[1][h0] = c0.mny; s.box[2][h0] = c0.mnz;
            s.box[3][h0] = c0.mxx; s.box[4][h0] = c0.mxy; s.box[5][h0] = c0.mxz;
            s.box[0][h1] = c1.mnx; s.box[1][h1] = c1.mny; s.box[2][h1] = c1.mnz;
            s.box[3][h1] = c1.mxx; s.box[4][h1] = c1.mxy; s.box[5][h1] = c1.mxz;
            s.box[0][h] = c.mnx; s.box[1][h] = c.mny; s.box[2][h] = c.mnz;
            s.box[3][h] = c.mxx; s.box[4][h] = c.mxy; s.box[5][h] = c.mxz;
        }
        __syncthreads();
    }

    // ---- inner boundary j = u (between units u and u+1) -----------------------
    // Nearest boundary with a smaller delta on each side, by binary lifting:
    // qL in [0, u] (or -1), qR in [u+2, m] (or m+1).  Boundary q lies before
    // unit q; boundary 0 is the chunk's left edge, m its right edge.
    const int p1 = u + 1;
    const bool inner = (uint32_t)u + 1 < m;
    const int v = inner ? s.st[0][p1] : -1;
    int qL = -1, qR = (int)m + 1;
    bool node = false;
    if (inner && v >= 0) {
        int pl = p1, pr = p1 + 1;
#pragma unroll
        for (int k = kLog; k >= 0; --k) {
            const int sz = 1 << k;
            if (pl >= sz && s.st[k][pl - sz] >= v) pl -= sz;
            if (pr + sz - 1 <= (int)m && s.st[k][pr] >= v) pr += sz;
        }
        qL = pl - 1;
        qR = pr;
        node = (qL >= 0 && qR <= (int)m);
    }
    // Segment starts: unit 0, and the unit after every boundary that is not a
    // node here (mesh walls included).
    if (inner && !node) atomicOr(&s.start[(u + 1) >> 5], 1u << ((u + 1) & 31));
    if (tid == 0) atomicOr(&s.start[0], 1u);
    __syncthreads();

    auto unit_index = [&](int x, uint32_t n) -> uint32_t {
        return subtree_index(n, s.ul[x], s.ur[x], s.st[0][x], s.st[0][x + 1]);
    };
    if (node) {
        const uint32_t seg = s.us[u];
        const SegDesc d = LEAF ? d0 : seg_desc(p.sb, seg);
        const uint32_t n = d.n;
        const uint32_t l = s.ul[qL], r = s.ur[qR - 1];
        const int bL = s.st[0][qL], bR = s.st[0][qR];
        const uint32_t idx = (bL < 0 && bR < 0) ? 0u : ((bR > bL) ? r : l);
        const uint32_t lc = (qL == u) ? unit_index(u, n) : s.ur[u];
        const uint32_t rc = (u + 1 == qR - 1) ? unit_index(u + 1, n) : s.ul[u + 1];
        const Box b = range_box(s, qL, qR - 1);
        store_node(d.nodes, idx, b, (int32_t)lc, (int32_t)rc);
        if (p.parents) {
            p.parents[lc] = idx;
            p.parents[rc] = idx;
        }
    }

    // ---- segments: the next level's units ------------------------------------
    bool exit_unit = false;
    int y = 0;
    if ((uint32_t)u < m && ((s.start[u >> 5] >> (u & 31)) & 1u)) {
        // Last unit of the segment: before the next start, or the chunk's last.
        y = (int)m - 1;
        uint32_t wbits = s.start[u >> 5] & ~((2u << (u & 31)) - 1u);  // starts after u in this word
        int wi = u >> 5;
        while (true) {
            if (wbits) {
                const int nxt = wi * 32 + __ffs((int)wbits) - 1;
                if (nxt - 1 < y) y = nxt - 1;
                break;
            }
            if (++wi >= kChunk / 32) break;
            wbits = s.start[wi];
        }
        const int bl = s.st[0][u], br = s.st[0][y + 1];
        exit_unit = !(bl < 0 && br < 0);  // both walls: a finished mesh
    }
    // Rank among the chunk's exits (block exclusive scan of the flags).
    const uint64_t bal = __ballot(exit_unit);
    const int lane = tid % kWave, wave = tid / kWave;
    if (lane == 0) s.wsum[wave] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kHierThreads / kWave; ++w) {
        const uint32_t c = s.wsum[w];
        before += (w < wave) ? c : 0u;
        total += c;
    }
    const uint32_t rank = before + (uint32_t)__popcll(bal & lanemask_lt());

    if (LEAF) {
        // Leaf level: exits go to the chunk's own slots (level 1 maps them
        // itself; with no chunked level above, k_hier_gather compacts them),
        // so no leaf-level block waits for another.
        if (tid == 0) {
            p.counts[t] = total;
            s.out_base = t * kChunk;
        }
        __syncthreads();
    } else {
        // Decoupled look-back over the level's chunks (ticket order = unit order),
        // one wave reading 64 predecessors' words per step: epoch-tagged words, so
        // stale words of earlier builds read as not ready.
        if (wave == 0) {
            uint64_t* st = p.status + (size_t)level * p.status_stride;
            const uint64_t tag = (uint64_t)epoch << 32;
            if (lane == 0)
                __hip_atomic_store(st + t, tag | (t == 0 ? kFlagP : kFlagA) | total, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            uint32_t prefix = 0;
            int top = (int)t - 1;  // nearest predecessor not yet accumulated
            while (top >= 0) {
                const int j = top - lane;
                uint32_t lo = kFlagP;  // before chunk 0: an inclusive prefix of 0
                if (j >= 0) {
                    const uint64_t w = __hip_atomic_load(st + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    lo = ((uint32_t)(w >> 32) == epoch) ? (uint32_t)w : 0u;
                }
                const uint64_t pmask = __ballot((lo & ~kValMask) == kFlagP);
                const uint64_t nready = __ballot((lo & ~kValMask) == 0);
                // Lanes 0 .. first P (or all 64) must be ready.
                const int first_p = pmask ? __ffsll((long long)pmask) - 1 : kWave - 1;
                const uint64_t need = (first_p == kWave - 1) ? ~0ull : ((2ull << first_p) - 1);
                if (nready & need) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                uint32_t v = (lane <= first_p) ? (lo & kValMask) : 0u;
    #pragma unroll
                for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, kWave);
                prefix += v;
                if (pmask) break;
                top -= kWave;
            }
            if (lane == 0) {
                if (t > 0)
                    __hip_atomic_store(st + t, tag | kFlagP | (prefix + total), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                s.out_base = prefix;
                if (t == nchunks - 1) p.ctl[kHierMaxLevels + level] = prefix + total;
            }
        }
        __syncthreads();
    }
    if (exit_unit) {
        const Box b = range_box(s, u, y);
        HUnit* dst = p.units_out + s.out_base + rank;
        reinterpret_cast<uint4*>(dst)[0] =
            make_uint4(s.ul[u], s.ur[y], s.us[u], pack_beta(s.st[0][u], s.st[0][y + 1]));
        reinterpret_cast<float4*>(dst)[1] = make_float4(b.mnx, b.mny, b.mnz, b.mxx);
        reinterpret_cast<float4*>(dst)[2] = make_float4(b.mxy, b.mxz, 0.0f, 0.0f);
    }
}

// Compacts level 1's exits (p.counts[c] units in the slots of chunk c) into
// one ordered sequence.  Block b owns chunks [b*kGather, (b+1)*kGather): it
// sums the counts of all earlier chunks, scans its own, and copies.
constexpr int kGather = 64;
constexpr uint32_t kFoldMaxChunks = 8192;  // leaf chunks up to which level 1 maps the slots itself
__global__ __launch_bounds__(kHierThreads) void k_hier_gather(HierArgs p) {
    __shared__ uint32_t s_off[kGather + 1];
    __shared__ uint32_t s_w[kHierThreads / kWave];
    const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
    const uint32_t nch = p.sb.nblocks;
    const uint32_t c0 = blockIdx.x * kGather;
    const uint32_t c1 = (c0 + kGather < nch) ? c0 + kGather : nch;
    // Units before chunk c0.
    uint32_t acc = 0;
    for (uint32_t c = tid; c < c0; c += kHierThreads) acc += p.counts[c];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, kWave);
    if (lane == 0) s_w[wave] = acc;
    __syncthreads();
    if (tid == 0) {
        uint32_t base = 0;
        for (int w = 0; w < kHierThreads / kWave; ++w) base += s_w[w];
        uint32_t run = base;
        for (uint32_t c = c0; c < c1; ++c) {
            s_off[c - c0] = run;
            run += p.counts[c];
        }
        s_off[c1 - c0] = run;
        if (c1 == nch) p.ctl[kHierMaxLevels + 0] = run;
    }
    __syncthreads();
    const uint32_t first = s_off[0], last = s_off[c1 - c0];
    for (uint32_t g = first + tid; g < last; g += kHierThreads) {
        // Chunk of output unit g: the last c with s_off[c - c0] <= g.
        int lo = 0, hi = (int)(c1 - c0) - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_off[mid] <= g) lo = mid;
            else hi = mid - 1;
        }
        const uint4* src = reinterpret_cast<const uint4*>(p.units[0] + (size_t)(c0 + lo) * kChunk + (g - s_off[lo]));
        uint4* dst = reinterpret_cast<uint4*>(p.units[1] + g);
        const uint4 a = src[0], b = src[1], c = src[2];
        dst[0] = a;
        dst[1] = b;
        dst[2] = c;
    }
}

template <bool LEAF, bool K60>
__global__ __launch_bounds__(kHierThreads) void k_hier_level(HierArgs p, int level) {
    __shared__ HierSmem s;
    if (LEAF) {  // one block per chunk, no ordering needed
        hier_chunk_body<LEAF, K60>(p, s, blockIdx.x, p.sb.nblocks, level);
        return;
    }
    // Units entering this level: a folded level 1 sums the leaf level's exit
    // counts.
    uint32_t total_in;
    if (level == 1 && p.fold1) {
        uint32_t acc = 0;
        for (uint32_t c = threadIdx.x; c < p.sb.nblocks; c += kHierThreads) acc += p.counts[c];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, kWave);
        if (lane_id() == 0) s.wsum[threadIdx.x / kWave] = acc;
        __syncthreads();
        total_in = 0;
        for (int w = 0; w < kHierThreads / kWave; ++w) total_in += s.wsum[w];
        __syncthreads();
    } else {
        total_in = p.ctl[kHierMaxLevels + level - 1];
    }
    const uint32_t nchunks = (total_in + kChunk - 1) / (uint32_t)kChunk;
    // Later levels take chunks in ticket order (the look-back needs every
    // smaller ticket to be running or done).
    while (true) {
        if (threadIdx.x == 0) s.ticket = atomicAdd(&p.ctl[level], 1u);
        __syncthreads();
        const uint32_t t = s.ticket;
        if (t >= nchunks) break;
        hier_chunk_body<LEAF, K60>(p, s, t, nchunks, level, total_in);
        __syncthreads();
    }
}

// Finishes whatever the chunked levels left: each unit climbs with
// epoch-tagged handoff words in HBM; the first arrival at a parent stops, the
// second merges (left box first) and climbs on.
template <bool K60>
__global__ __launch_bounds__(256) void k_hier_climb(HierArgs p, int level) {
    const uint32_t cnt = p.ctl[kHierMaxLevels + level];
    const uint32_t epoch = *p.epoch;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += gridDim.x * blockDim.x) {
        const HUnit* src = p.units_in + e;
        const uint4 h = *reinterpret_cast<const uint4*>(src);
        const float4 w0 = reinterpret_cast<const float4*>(src)[1];
        const float4 w1 = reinterpret_cast<const float4*>(src)[2];
        const SegDesc d = seg_desc(p.sb, h.z);
        const uint32_t n = d.n;
        const uint64_t* __restrict__ keys = p.keys + d.start;
        const uint32_t* __restrict__ vals = K60 ? p.vals + d.start : nullptr;
        uint64_t* __restrict__ handoff = p.handoff + d.start;
        DNode* __restrict__ nodes = d.nodes;
        auto dget = [&](int64_t j) -> int {
            if (j < 0 || j >= (int64_t)n - 1) return -1;
            return key_delta<K60>(keys, vals, (uint32_t)j);
        };
        uint32_t l = h.x, r = h.y;
        int dl = beta_left(h.w), dr = beta_right(h.w);
        Box box;
        box.mnx = w0.x; box.mny = w0.y; box.mnz = w0.z;
        box.mxx = w0.w; box.mxy = w1.x; box.mxz = w1.y;
        uint32_t my = subtree_index(n, l, r, dl, dr);
        int32_t lc = 0, rc = 0;
        bool stored = true;  // the unit's record was written by an earlier launch
        while (true) {
            if (l == 0 && r == n - 1) {
                store_node(nodes, my, box, lc, rc);
                break;
            }
            const bool isLeft = dr > dl;
            const uint32_t sp = isLeft ? r : l - 1;
            const uint32_t far = isLeft ? l : r;
            if (!stored) {
                store_node_coherent(nodes, my, box, lc, rc);
                __builtin_amdgcn_s_waitcnt(0);
            }
            stored = false;
            const uint64_t hw = ((uint64_t)epoch << 32) | far;
            const uint64_t old = __hip_atomic_exchange(&handoff[sp], hw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((uint32_t)(old >> 32) != epoch) break;
            const uint32_t other = (uint32_t)old;
            const uint32_t sib = isLeft ? ((sp + 1 == other) ? (n - 1 + other) : (sp + 1))
                                        : ((other == sp) ? (n - 1 + sp) : sp);
            const Box sbox = load_box_coherent(nodes, sib);
            const Box pb = isLeft ? merge(box, sbox) : merge(sbox, box);
            lc = (int32_t)(isLeft ? my : sib);
            rc = (int32_t)(isLeft ? sib : my);
            if (isLeft) r = other;
            else l = other;
            dl = dget((int64_t)l - 1);
            dr = dget((int64_t)r);
            const uint32_t pidx = (l == 0 && r == n - 1) ? 0u : ((dr > dl) ? r : l);
            if (p.parents) {
                p.parents[lc] = pidx;
                p.parents[rc] = pidx;
            }
            my = pidx;
            box = pb;
        }
    }
}

int hier_levels(uint32_t maxn, uint32_t meshes) {
    // Each level shrinks the units >= ~64x on Morton-ordered data; meshes that
    // straddle a chunk boundary of a later level may need one more.
    int levels = 1;
    uint64_t chunks = (maxn + kChunk - 1) / kChunk;
    while (chunks > 1 && levels < kHierMaxLevels - 1) {
        ++levels;
        chunks = (chunks + 63) / 64;
    }
    if (meshes > 1 && maxn > (uint32_t)kChunk && levels < kHierMaxLevels) ++levels;
    return levels;
}

void launch_hierarchy(const HierArgs& a0, hipStream_t s) {
    HierArgs a = a0;
    const bool k60 = a.vals != nullptr;
    // The leaf level writes per-chunk slots of units[0].  Folded (a chunked
    // level above it, at most kFoldMaxChunks leaf chunks): level 1 maps the
    // slots itself and writes units[1], level lv reads units[(lv + 1) & 1] and
    // writes units[lv & 1].  Otherwise k_hier_gather compacts the slots into
    // units[1] and level lv reads units[lv & 1] and writes the other.  (Every
    // folded level-1 block scans all leaf chunks' counts, which costs more
    // than the compaction pass beyond ~8M leaves.)
    a.fold1 = (a.levels > 1 && a.sb.nblocks <= kFoldMaxChunks) ? 1 : 0;
    a.units_in = nullptr;
    a.units_out = a.units[0];
    if (k60) hipLaunchKernelGGL((k_hier_level<true, true>), dim3(a.sb.nblocks), dim3(kHierThreads), 0, s, a, 0);
    else hipLaunchKernelGGL((k_hier_level<true, false>), dim3(a.sb.nblocks), dim3(kHierThreads), 0, s, a, 0);
    if (!a.fold1)
        hipLaunchKernelGGL(k_hier_gather, dim3((a.sb.nblocks + kGather - 1) / kGather), dim3(kHierThreads), 0, s, a);
    uint32_t grid = a.sb.nblocks;
    for (int lv = 1; lv < a.levels; ++lv) {
        a.units_in = a.fold1 ? a.units[(lv + 1) & 1] : a.units[lv & 1];
        a.units_out = a.fold1 ? a.units[lv & 1] : a.units[(lv + 1) & 1];
        grid = (grid + 31) / 32;
        if (grid < 1) grid = 1;
        if (k60) hipLaunchKernelGGL((k_hier_level<false, true>), dim3(grid), dim3(kHierThreads), 0, s, a, lv);
        else hipLaunchKernelGGL((k_hier_level<false, false>), dim3(grid), dim3(kHierThreads), 0, s, a, lv);
    }
    // The climb reads the last level's output.
    a.units_in = a.fold1 ? a.units[(a.levels - 1) & 1] : a.units[a.levels & 1];
    uint32_t cgrid = (a.sb.nblocks + 63) / 64;
    cgrid = cgrid < 1 ? 1 : (cgrid > 512 ? 512 : cgrid);
    if (k60) hipLaunchKernelGGL(k_hier_climb<true>, dim3(cgrid), dim3(256), 0, s, a, a.levels - 1);
    else hipLaunchKernelGGL(k_hier_climb<false>, dim3(cgrid), dim3(256), 0, s, a, a.levels - 1);
}

}  // namespace bvhk
