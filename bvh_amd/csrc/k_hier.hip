// k_hier.hip — fused topology, leaf initialisation and bottom-up refit.
//
// Replaces kBuildInternalNodes_nt/_P + kInitLeafNodes_nt/_P + kRefitHierarchy_nt
// (reference lbvh_builder.cu:247-338, lbvh_plus_builder.cu:292-335,341-366 with
// doOptimize = false).
//
// The reference finds each internal node top-down (Karras 2012: direction,
// exponential + binary range search, binary split search; O(log n) scattered
// key reads per node), then refits bottom-up in a second kernel through
// per-node atomic counters and __threadfence.  Here one pass climbs from the
// leaves: a node covering sorted leaves [l, r] is the left child of its
// parent iff delta(r) > delta(l-1) (delta = common-prefix length of adjacent
// keys, -1 past either end); the parent's split is then r (left child) or
// l-1 (right child).  The radix tree of unique keys is unique, so this
// yields exactly the reference's topology, and its node numbering follows
// from the same rule: an internal node is numbered by its right end when it
// is a left child and by its left end otherwise (the root is 0), and leaf k
// is node n-1+k, as in Karras' layout.  The two children of a split meet in a
// handoff word: the first arrival deposits its far range end and stops, the
// second merges the boxes (left box first, as the reference merges
// nodes[left] with nodes[right]) and climbs on.
//
// MI355X mapping: a workgroup of 512 threads owns a chunk of 1024 sorted
// leaves (two per thread).  Phase 1: every split whose parent lies entirely
// inside the chunk (decided once per boundary from prefix/suffix minima of
// the chunk's delta values, so both children agree, and packed beside the
// delta) hands off through an LDS word with LDS-only release/acquire; the
// word names the depositing chain, whose subtree box waits in LDS, so the
// second arrival never re-reads a node record from memory and no climb step
// waits on HBM.  Chains whose parent straddles the chunk boundary park in an
// exit list.  Phase 2 runs the exits, one per thread: those nodes (the top of
// the tree, roughly 2n/1024 of them) hand off through an epoch-tagged 64-bit
// word in HBM; their records are written with agent-scope (sc1) stores and
// read with agent-scope loads, with a full wait before the exchange, so no L2
// writeback or invalidate is ever needed.  Epoch tags mean the handoff array
// is never cleared.  Blocks map onto (mesh, chunk) pairs, so a batch of
// meshes builds in one launch with chunk and root boundaries at every mesh
// boundary.
#include "dev_common.hpp"
#include "engine.hpp"

namespace bvhk {

constexpr int kHierBlock = 512;
constexpr int kHierItems = 2;
constexpr int kChunk = kHierBlock * kHierItems;
// s_waitcnt encoding (gfx9 family) that waits for vmcnt(0) only.
constexpr int kWaitVmcnt0 = 0x0F70;

int hier_chunk() { return kChunk; }

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

__device__ __forceinline__ int wave_incl_min(int v) {
    uint32_t l = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int w = __shfl_up(v, o, 64);
        if (l >= (uint32_t)o) v = (w < v) ? w : v;
    }
    return v;
}

// Inclusive min-scan of src[0..cnt) into dst (forward, or backward when REV).
template <bool REV>
__device__ void block_min_scan(const int8_t* src, int8_t* dst, int cnt, int* s_w) {
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    const int per = (cnt + kHierBlock - 1) / kHierBlock;
    const int lo = tid * per;
    const int hi = (lo + per < cnt) ? lo + per : cnt;
    int m = 127;
    for (int q = lo; q < hi; ++q) {
        int v = src[REV ? cnt - 1 - q : q];
        m = v < m ? v : m;
    }
    int incl = wave_incl_min(m);
    int ex = __shfl_up(incl, 1, 64);
    if (lane == 0) ex = 127;
    if (lane == kWave - 1) s_w[wave] = incl;
    __syncthreads();
    for (int w = 0; w < wave; ++w) ex = s_w[w] < ex ? s_w[w] : ex;
    int run = ex;
    for (int q = lo; q < hi; ++q) {
        int qq = REV ? cnt - 1 - q : q;
        int v = src[qq];
        run = v < run ? v : run;
        dst[qq] = (int8_t)run;
    }
    __syncthreads();
}

// Node record written with agent-scope atomic stores (coherent across XCDs
// without flushing the L2), and the matching loads.
__device__ __forceinline__ void store_node_coherent(DNode* nodes, uint32_t i, const Box& b, int32_t l, int32_t r) {
    float* f = reinterpret_cast<float*>(nodes + i);
    int32_t* q = reinterpret_cast<int32_t*>(nodes + i);
    __hip_atomic_store(f + 0, b.mnx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(f + 1, b.mny, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(f + 2, b.mnz, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(f + 3, 0.0f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(f + 4, b.mxx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(f + 5, b.mxy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(f + 6, b.mxz, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(f + 7, 0.0f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 8, l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 9, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    q[10] = 0;
    q[11] = 0;
}

__device__ __forceinline__ Box load_box_coherent(const DNode* nodes, uint32_t i) {
    const float* f = reinterpret_cast<const float*>(nodes + i);
    Box b;
    b.mnx = __hip_atomic_load(f + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    b.mny = __hip_atomic_load(f + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    b.mnz = __hip_atomic_load(f + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    b.mxx = __hip_atomic_load(f + 4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    b.mxy = __hip_atomic_load(f + 5, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    b.mxz = __hip_atomic_load(f + 6, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return b;
}

template <bool K60>
__global__ __launch_bounds__(kHierBlock) void k_hierarchy(HierArgs p) {
    __shared__ int8_t s_d[kChunk + 1];    // delta of boundaries a-1 .. b-1
    // Prefix and suffix minima of s_d; once read they are replaced by the
    // packed per-boundary words s_dq: delta + 1 (low byte), and bit 8 set when
    // the split at that boundary has its parent inside the chunk.
    __shared__ union {
        struct {
            int8_t pre[kChunk + 1];
            int8_t suf[kChunk + 1];
        } m;
        uint16_t dq[kChunk + 1];
    } s_u;
    int8_t* const s_pre = s_u.m.pre;
    int8_t* const s_suf = s_u.m.suf;
    uint16_t* const s_dq = s_u.dq;
    __shared__ uint32_t s_slot[kChunk];   // LDS handoff words: (far end - a + 1) | depositing chain << 16
    __shared__ float s_box[6][kChunk];    // each chain's current box, read by its partner
    __shared__ int s_w[kHierBlock / kWave];
    __shared__ uint32_t s_nexit;
    static_assert(kChunk <= 1024, "exit entries pack three 10-bit chunk offsets");

    // This block's mesh (segment) and chunk; everything below is in the
    // mesh's local coordinates.
    SegDesc d;
    uint32_t chunk;
    if (p.sb.blocks) {
        const uint2 e = p.sb.blocks[blockIdx.x];
        d = p.sb.segs[e.x];
        chunk = e.y;
    } else {
        d = p.sb.single;
        chunk = blockIdx.x;
    }
    const uint32_t n = d.n;
    const uint64_t* __restrict__ keys = p.keys + d.start;
    const uint32_t* __restrict__ vals = K60 ? p.vals + d.start : nullptr;
    const float4* __restrict__ tboxes = p.tri_boxes + 2 * (size_t)d.start;
    uint64_t* __restrict__ handoff = p.handoff + d.start;
    const uint32_t epoch = *p.epoch;
    const uint64_t idx_mask = (p.kf.idx_bits >= 32) ? 0xFFFFFFFFull : ((1ull << p.kf.idx_bits) - 1);
    const uint32_t a = chunk * kChunk;
    const uint32_t b = (a + kChunk < n) ? a + kChunk : n;
    const int len = (int)(b - a);
    const int tid = threadIdx.x;

    // Every item's primitive and leaf box first: the random box gathers of all
    // items are in flight together (and overlap the delta scans below)
    // instead of one per climb.
    uint32_t pprim[kHierItems];
    Box pbox[kHierItems];
#pragma unroll
    for (int k = 0; k < kHierItems; ++k) {
        const uint32_t i = a + k * kHierBlock + tid;
        uint32_t prim = 0;
        if (i < b) {
            prim = K60 ? vals[i] : (uint32_t)(keys[i] & idx_mask);
            if (prim >= n) {  // cannot happen for a correct sort; never index out of bounds
                atomicOr(p.err, 1u);
                prim = 0;
            }
            d.indices[i] = prim;
            pbox[k] = load_box(tboxes + 2 * (size_t)prim);
        }
        pprim[k] = prim;
    }
    for (int q = tid; q <= len; q += kHierBlock) {
        int64_t j = (int64_t)a - 1 + q;
        s_d[q] = (j < 0 || j >= (int64_t)n - 1) ? (int8_t)-1 : (int8_t)key_delta<K60>(keys, vals, (uint32_t)j);
    }
    for (int q = tid; q < len; q += kHierBlock) s_slot[q] = 0;
    if (tid == 0) s_nexit = 0;
    __syncthreads();
    block_min_scan<false>(s_d, s_pre, len + 1, s_w);
    block_min_scan<true>(s_d, s_suf, len + 1, s_w);
    // A split's parent lies inside the chunk iff a smaller delta exists on
    // both sides within the chunk (prefix/suffix minima); both children of a
    // split evaluate this identically.
    constexpr int kPerThread = (kChunk + 1 + kHierBlock - 1) / kHierBlock;
    uint16_t dqv[kPerThread];
#pragma unroll
    for (int t = 0; t < kPerThread; ++t) {
        const int q = tid + t * kHierBlock;
        if (q <= len) {
            const int dv = s_d[q];
            const bool in = q >= 1 && q <= len - 1 && s_pre[q - 1] < dv && s_suf[q + 1] < dv;
            dqv[t] = (uint16_t)((dv + 1) | (in ? 0x100 : 0));
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kPerThread; ++t) {
        const int q = tid + t * kHierBlock;
        if (q <= len) s_dq[q] = dqv[t];
    }
    __syncthreads();

    // delta at boundary j (-1 beyond either end), from LDS inside the chunk.
    auto dget = [&](int64_t j) -> int {
        if (j < 0 || j >= (int64_t)n - 1) return -1;
        if (j >= (int64_t)a - 1 && j <= (int64_t)b - 1) return s_d[j - ((int64_t)a - 1)];
        return key_delta<K60>(keys, vals, (uint32_t)j);
    };

    DNode* __restrict__ nodes = d.nodes;
    // delta at boundary j1 - 1, for a <= j1 <= b (-1 beyond the mesh).
    auto dloc = [&](uint32_t j1) -> int { return s_d[j1 - a]; };
    // Every prefetched box is in registers before the climbs start.  gfx950
    // counts stores and loads in one vmcnt, so a load still pending inside
    // the climb loop would make each step wait for its own node store.
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);

    // Phase 1, inside the chunk: splits whose parent lies entirely in the
    // chunk (decided from the prefix/suffix minima, which both children
    // evaluate identically) meet in LDS.  These climbs touch no global memory
    // but their node stores and never wait for them.  A chain whose parent
    // straddles the chunk boundary stores its node record coherently, parks
    // its box in LDS and leaves an exit entry (range and chain, 10 bits each);
    // no wave waits on HBM here.
    uint32_t exitw[kHierItems];
    for (int k = 0; k < kHierItems; ++k) {
        exitw[k] = ~0u;
        const int chain = k * kHierBlock + tid;
        const uint32_t i = a + chain;
        if (i >= b) continue;
        uint32_t l = i, r = i;
        uint32_t my = n - 1 + i;
        Box box = pbox[k];
        int32_t lc = (int32_t)(pprim[k] | 0x80000000u), rc = -1;
        // Packed deltas (delta + 1 | in-chunk bit) of boundaries l - 1 and r.
        uint32_t dl = s_dq[l - a];
        uint32_t dr = s_dq[r + 1 - a];
        bool done = false;
        while (true) {
            const bool isLeft = (dr & 0xFFu) > (dl & 0xFFu);
            // The split's own in-chunk bit (never set at the chunk's end
            // boundaries, so the root and straddling parents leave here).
            if (!((isLeft ? dr : dl) & 0x100u)) break;
            const uint32_t s = isLeft ? r : l - 1;
            const uint32_t far = isLeft ? l : r;
            store_node(nodes, my, box, lc, rc);
            // The partner reads this chain's box from LDS, not from the node
            // record: LDS-only release/acquire around the handoff word.
            s_box[0][chain] = box.mnx; s_box[1][chain] = box.mny; s_box[2][chain] = box.mnz;
            s_box[3][chain] = box.mxx; s_box[4][chain] = box.mxy; s_box[5][chain] = box.mxz;
            const uint32_t word = (far - a + 1u) | ((uint32_t)chain << 16);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            const uint32_t old = __hip_atomic_exchange(&s_slot[s - a], word, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old == 0) {
                done = true;
                break;
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
            const uint32_t other = a + (old & 0xFFFFu) - 1u;
            const int pc = (int)(old >> 16);
            // Only the far boundary moves; its delta and the partner's box are
            // read together.
            const uint32_t nv = s_dq[(isLeft ? other + 1u : other) - a];
            Box sbox;
            sbox.mnx = s_box[0][pc]; sbox.mny = s_box[1][pc]; sbox.mnz = s_box[2][pc];
            sbox.mxx = s_box[3][pc]; sbox.mxy = s_box[4][pc]; sbox.mxz = s_box[5][pc];
            const uint32_t sib = isLeft ? ((s + 1 == other) ? (n - 1 + other) : (s + 1))
                                        : ((other == s) ? (n - 1 + s) : s);
            lc = (int32_t)(isLeft ? my : sib);
            rc = (int32_t)(isLeft ? sib : my);
            if (isLeft) {
                r = other;
                dr = nv;
            } else {
                l = other;
                dl = nv;
            }
            box = isLeft ? merge(box, sbox) : merge(sbox, box);
            const uint32_t pidx = (l == 0 && r == n - 1) ? 0u : (((dr & 0xFFu) > (dl & 0xFFu)) ? r : l);
            if (p.parents) {
                p.parents[lc] = pidx;
                p.parents[rc] = pidx;
            }
            if (p.ranges) p.ranges[pidx] = make_int2((int)l, (int)r);
            my = pidx;
        }
        if (done) continue;
        if (l == 0 && r == n - 1) {  // the chunk is the whole mesh: this is the root
            store_node(nodes, my, box, lc, rc);
            continue;
        }
        store_node_coherent(nodes, my, box, lc, rc);
        s_box[0][chain] = box.mnx; s_box[1][chain] = box.mny; s_box[2][chain] = box.mnz;
        s_box[3][chain] = box.mxx; s_box[4][chain] = box.mxy; s_box[5][chain] = box.mxz;
        exitw[k] = (l - a) | ((r - a) << 10) | ((uint32_t)chain << 20);
    }
    // The exit records are complete in memory before any thread publishes a
    // handoff word for them; then the exits are compacted into s_slot.
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kHierItems; ++k)
        if (exitw[k] != ~0u) s_slot[atomicAdd(&s_nexit, 1u)] = exitw[k];
    __syncthreads();
    const uint32_t nexit = s_nexit;

    // Phase 2, one exit per thread: splits whose parent straddles a chunk
    // boundary.  The partner may run on another XCD, so every record is
    // stored with agent-coherent stores, complete before the epoch-tagged
    // handoff word in HBM is swapped.
    for (uint32_t e = tid; e < nexit; e += kHierBlock) {
        const uint32_t w = s_slot[e];
        uint32_t l = a + (w & 1023u), r = a + ((w >> 10) & 1023u);
        const uint32_t chain = w >> 20;
        Box box;
        box.mnx = s_box[0][chain]; box.mny = s_box[1][chain]; box.mnz = s_box[2][chain];
        box.mxx = s_box[3][chain]; box.mxy = s_box[4][chain]; box.mxz = s_box[5][chain];
        int dl = dloc(l);
        int dr = dloc(r + 1);
        uint32_t my = (l == r) ? n - 1 + l : ((dr > dl) ? r : l);
        int32_t lc = 0, rc = 0;
        bool stored = true;  // the exit record was stored in phase 1
        while (true) {
            if (l == 0 && r == n - 1) {
                store_node(nodes, my, box, lc, rc);
                break;
            }
            const bool isLeft = dr > dl;
            const uint32_t s = isLeft ? r : l - 1;
            const uint32_t far = isLeft ? l : r;
            if (!stored) {
                store_node_coherent(nodes, my, box, lc, rc);
                __builtin_amdgcn_s_waitcnt(0);
            }
            stored = false;
            const uint64_t hw = ((uint64_t)epoch << 32) | far;
            const uint64_t old = __hip_atomic_exchange(&handoff[s], hw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((uint32_t)(old >> 32) != epoch) break;
            const uint32_t other = (uint32_t)old;
            const uint32_t sib = isLeft ? ((s + 1 == other) ? (n - 1 + other) : (s + 1))
                                        : ((other == s) ? (n - 1 + s) : s);
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
            if (p.ranges) p.ranges[pidx] = make_int2((int)l, (int)r);
            my = pidx;
            box = pb;
        }
    }
}

void launch_hierarchy(const HierArgs& a, hipStream_t s) {
    if (a.vals) hipLaunchKernelGGL(k_hierarchy<true>, dim3(a.sb.nblocks), dim3(kHierBlock), 0, s, a);
    else hipLaunchKernelGGL(k_hierarchy<false>, dim3(a.sb.nblocks), dim3(kHierBlock), 0, s, a);
}

}  // namespace bvhk
