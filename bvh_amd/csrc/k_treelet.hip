// k_treelet.hip — LBVH+ treelet restructuring (Karras & Aila 2013, 7 leaves).
//
// Replaces kRefitAndOptimize(doOptimize = true) + optimizeTreeletSerial of the
// reference (src/cuda/lbvh_plus_builder.cu:71-247,341-366): one bottom-up pass
// in which every internal node, once both of its children are final, is
// refitted (box = left box U right box) and then becomes the root of a treelet
// that is rebuilt optimally:
//   * formation: start from the root, repeatedly expand the treelet leaf of
//     largest surface area that is not a BVH leaf (ties: first in order),
//     until 7 leaves or nothing left to expand;
//   * dynamic programme over the 2^n - 1 leaf subsets: singletons cost 0,
//     cost(s) = min over partitions p of cost(p) + cost(s^p), plus area(s);
//     partitions enumerated in the reference's order with a strict '<' (first
//     minimum wins); box(s) = ascending-index union of the leaf boxes;
//   * reconstruction reuses the treelet's internal node ids in DFS pop order
//     (right child popped first) and assigns them breadth first.
// Every float operation matches the oracle's serial transcription in order, so
// the resulting trees are bit-identical to it.
//
// MI355X mapping: threads climb from the leaves as in the reference, each
// advancing its two to eight chains together; at each climbing step every chain
// that became a second arrival is ready, and the wave forms up to eight
// ready treelets at once into LDS slots (seven dependent rounds of node
// loads, issued by all owners together), runs their dynamic programmes side
// by side on equal lane groups (64 lanes for a lone treelet, 8 each for
// eight), every subset size spread over the group with partitions split
// across lanes and reduced in enumeration order, and each owner lane rebuilds
// its treelet from the group's partition table.  A workgroup owns 512-2048
// sorted leaves (fewer for small trees, so that at least 512 workgroups run);
// a node whose leaf span lies inside the chunk hands off
// through a workgroup-scope counter, only chunk-straddling nodes use agent
// scope.  Records are written with agent-coherent stores and read coherently
// only for straddling treelets, and every arrival waits for its own stores
// first, so no whole-L2 writeback or invalidate is issued.  Leaf spans are
// kept per internal node (min/max sorted leaf position) and updated when a
// treelet reassigns its internal nodes.
#include "dev_common.hpp"
#include "engine.hpp"

namespace bvhk {

constexpr int kTBlock = 256;
constexpr int kTItemsMax = 8;  // leaf chains per thread (2, 4 or 8: a chunk of kTBlock x items leaves)
constexpr int kTWaves = kTBlock / kWave;
constexpr int kTMax = 7;
constexpr int kGroups = 8;               // treelets per wave per round (groups of >= 8 lanes)
// dp_group initialises a treelet's single-leaf costs one lane per leaf, so a
// group needs at least kTMax lanes.
static_assert(kWave / kGroups >= kTMax, "a treelet group needs a lane per leaf");
constexpr float kCostInf = 1e30f;

// One treelet in LDS: formed by its owner lane, optimised by one 8-lane
// group, rebuilt by the owner.
struct Slot {
    float lb[kTMax][6];      // treelet leaf boxes, in treelet order
    float cost1[127];        // cost of leaf subset s (1..127) at cost1[s - 1]
    uint8_t part[128];
    int32_t act[kTMax];      // treelet leaves (node ids)
    int32_t lo[kTMax];       // their sorted-leaf spans
    int32_t hi[kTMax];
    int32_t exn[kTMax - 1];  // expanded internal nodes and their children, in order
    int32_t exl[kTMax - 1];
    int32_t exr[kTMax - 1];
    int32_t n;  // leaves; the treelet has n - 1 expanded (internal) nodes
};

struct WaveLds {
    Slot slot[kGroups];
};
// An odd slot stride in words puts equal offsets of the eight slots in eight
// different LDS banks (the groups walk identical subset sequences), and four
// waves' slots fit 32 KB (five workgroups per CU).
static_assert(sizeof(Slot) % 8 == 4, "slot stride must be an odd number of words");
static_assert(sizeof(WaveLds) * kTWaves + 127 <= 32768, "treelet LDS (slots + subset table) must stay within 32 KB");

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ Box lds_box(const float* f) {
    Box b;
    b.mnx = f[0]; b.mny = f[1]; b.mnz = f[2];
    b.mxx = f[3]; b.mxy = f[4]; b.mxz = f[5];
    return b;
}
__device__ __forceinline__ void lds_put(float* f, const Box& b) {
    f[0] = b.mnx; f[1] = b.mny; f[2] = b.mnz;
    f[3] = b.mxx; f[4] = b.mxy; f[5] = b.mxz;
}


__device__ __forceinline__ Box node_box_of(const DNode& d) {
    Box b;
    b.mnx = d.mn.x; b.mny = d.mn.y; b.mnz = d.mn.z;
    b.mxx = d.mx.x; b.mxy = d.mx.y; b.mxz = d.mx.z;
    return b;
}

__device__ __forceinline__ void nodes_refit(const TreeletArgs& a, int32_t root, const Box& b) {
    store_box_coherent(a.nodes, (uint32_t)root, b);
}


template <bool COH>
__device__ __forceinline__ int2 leaf_span(const TreeletArgs& a, int32_t node) {
    const uint32_t first_leaf = a.n - 1;
    if ((uint32_t)node >= first_leaf) return make_int2((int)(node - first_leaf), (int)(node - first_leaf));
    return COH ? load_int2_coherent(a.ranges + node) : a.ranges[node];
}

// A node record: plain loads inside the workgroup's own chunk, agent-coherent
// loads for treelets whose records other workgroups (possibly on other XCDs)
// wrote during this launch.
template <bool COH>
__device__ __forceinline__ DNode node_rec(const DNode* nodes, int32_t i) {
    return COH ? load_node_coherent(nodes, (uint32_t)i) : nodes[i];
}

// Union of the subset's leaf boxes in ascending index order from the empty
// box (the oracle's order: min/max by selects are order-sensitive only for
// signed zeros, and this keeps even those identical).  Merging a finite box
// into the empty box returns it bit for bit, so the fold starts at the first
// leaf and costs one merge per further leaf (leaf boxes read from LDS; the
// subsets of one DP step all have the same size, so the loop is uniform).
__device__ __forceinline__ Box subset_box(const Slot& sl, int s) {
    Box b = lds_box(sl.lb[__ffs(s) - 1]);
    for (int t = s & (s - 1); t != 0; t &= t - 1) b = merge(b, lds_box(sl.lb[__ffs(t) - 1]));
    return b;
}

// Small per-lane tables (formation, reconstruction) as 8-wide vector values: they
// live in VGPRs and lane-dependent lookups become selects (a plain local
// array would be turned into an indexed load and land in scratch or LDS).
typedef int32_t RegTab __attribute__((ext_vector_type(8)));

// The entry whose bit is set in the one-hot mask.
__device__ __forceinline__ int32_t reg_pick(RegTab v, uint32_t onehot) {
    int32_t r = v[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) r = ((onehot >> k) & 1u) ? v[k] : r;
    return r;
}
__device__ __forceinline__ RegTab reg_put(RegTab v, uint32_t onehot, int32_t x) {
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = ((onehot >> k) & 1u) ? x : v[k];
    return v;
}

// Greedy formation (reference lbvh_plus_builder.cu:79-99): expand the
// treelet leaf of largest area that is not a BVH leaf (first maximum wins)
// until 7 leaves.  Each act leaf's children come with its node record, so
// an expansion costs one round of loads.  The leaves' areas and child ids
// are kept in registers as well (the argmax and the next expansion's loads
// need no LDS round trip); LDS gets what the DP and the rebuild read.
// Returns the refitted root box (the reference's union of the root's
// children) and the leaf count.
template <bool COH>
__device__ int form(const TreeletArgs& a, int32_t root, Slot& sl, Box& rootbox) {
    typedef float AreaTab __attribute__((ext_vector_type(8)));
    const DNode* __restrict__ nodes = a.nodes;
    const int4 rlr = node_rec<COH>(nodes, root).lr;
    const int2 rsp = COH ? load_int2_coherent(a.ranges + root) : a.ranges[root];
    sl.act[0] = root;
    sl.lo[0] = rsp.x;
    sl.hi[0] = rsp.y;
    RegTab lcv = 0, rcv = 0;
    AreaTab ar = -1.0f;
    lcv[0] = rlr.x;
    rcv[0] = rlr.y;
    uint32_t leafmask = 0;
    int n = 1, nexp = 0;
    DNode cl = node_rec<COH>(nodes, rlr.x);
    DNode cr = node_rec<COH>(nodes, rlr.y);
    rootbox = merge(node_box_of(cl), node_box_of(cr));
    lds_put(sl.lb[0], rootbox);
    ar[0] = area(rootbox);
    int best = 0;
    for (int step = 0; step < kTMax - 1; ++step) {
        if (step > 0) {
            best = -1;
            float maxA = -1.0f;
#pragma unroll
            for (int i = 0; i < kTMax; ++i) {
                if (i < n && !((leafmask >> i) & 1u) && ar[i] > maxA) {
                    maxA = ar[i];
                    best = i;
                }
            }
            if (best < 0) break;
            cl = node_rec<COH>(nodes, reg_pick(lcv, 1u << best));
            cr = node_rec<COH>(nodes, reg_pick(rcv, 1u << best));
        } else if (!(ar[0] > -1.0f)) {
            break;  // the reference's argmax finds nothing (non-finite area)
        }
        const uint32_t bbit = 1u << best, nbit = 1u << n;
        const int32_t l = reg_pick(lcv, bbit), r = reg_pick(rcv, bbit);
        const int2 sp_l = leaf_span<COH>(a, l), sp_r = leaf_span<COH>(a, r);
        sl.exn[nexp] = sl.act[best];
        sl.exl[nexp] = l;
        sl.exr[nexp] = r;
        ++nexp;
        const Box bl = node_box_of(cl), br = node_box_of(cr);
        sl.act[best] = l;
        lds_put(sl.lb[best], bl);
        sl.lo[best] = sp_l.x;
        sl.hi[best] = sp_l.y;
        lcv = reg_put(lcv, bbit, cl.lr.x);
        rcv = reg_put(rcv, bbit, cl.lr.y);
        const float al = area(bl), arr = area(br);
#pragma unroll
        for (int k = 0; k < kTMax; ++k) {
            ar[k] = ((bbit >> k) & 1u) ? al : ar[k];
            ar[k] = ((nbit >> k) & 1u) ? arr : ar[k];
        }
        leafmask = (leafmask & ~bbit) | ((cl.lr.x < 0 ? 1u : 0u) << best);
        sl.act[n] = r;
        lds_put(sl.lb[n], br);
        sl.lo[n] = sp_r.x;
        sl.hi[n] = sp_r.y;
        lcv = reg_put(lcv, nbit, cr.lr.x);
        rcv = reg_put(rcv, nbit, cr.lr.y);
        leafmask |= (cr.lr.x < 0 ? 1u : 0u) << n;
        ++n;
    }
    sl.n = n;
    return n;
}

// Cost and best partition of subset s (reference lbvh_plus_builder.cu:131-160):
// partitions in the reference's enumeration order, strict '<'.
__device__ __forceinline__ void dp_subset(Slot& sl, int s) {
    const Box b = subset_box(sl, s);
    const float areaS = area(b);
    float minCost = kCostInf;
    int bestP = -1;
    const int delta = (s - 1) & s;
    int p = (-delta) & s;
    do {
        const float c = sl.cost1[p - 1] + sl.cost1[(s ^ p) - 1];
        if (c < minCost) {
            minCost = c;
            bestP = p;
        }
        p = (p - delta) & s;
    } while (p != 0);
    sl.cost1[s - 1] = minCost + 1.0f * areaS;
    sl.part[s] = (uint8_t)bestP;
}

// Subsets of 7 leaves grouped by size, increasing within a size (so the
// subsets of n < 7 leaves are a prefix of each group); copied into LDS by
// every workgroup.
__constant__ uint8_t c_sub[127] = {
    1, 2, 4, 8, 16, 32, 64, 3, 5, 6, 9, 10, 12, 17, 18, 20, 24, 33, 34, 36, 40, 48, 65, 66, 68, 72, 80, 96,
    7, 11, 13, 14, 19, 21, 22, 25, 26, 28, 35, 37, 38, 41, 42, 44, 49, 50, 52, 56, 67, 69, 70, 73, 74, 76, 81,
    82, 84, 88, 97, 98, 100, 104, 112, 15, 23, 27, 29, 30, 39, 43, 45, 46, 51, 53, 54, 57, 58, 60, 71, 75, 77,
    78, 83, 85, 86, 89, 90, 92, 99, 101, 102, 105, 106, 108, 113, 114, 116, 120, 31, 47, 55, 59, 61, 62, 79,
    87, 91, 93, 94, 103, 107, 109, 110, 115, 117, 118, 121, 122, 124, 63, 95, 111, 119, 123, 125, 126, 127};
// The t-th (0-based) nonzero submask of `mask` in increasing order: the
// reference's partition enumeration p -> (p - delta) & s visits exactly these.
__device__ __forceinline__ int submask(int t, int mask) {
    int v = t + 1, r = 0;
    for (int m = mask; m != 0; m &= m - 1) {
        if (v & 1) r |= m & -m;
        v >>= 1;
    }
    return r;
}

// The dynamic programme of one treelet on a group of G lanes (G a power of
// two, 8..64; sub = lane in group).  Subsets of one size are independent:
// each gets L lanes (the largest power of two with L * count <= G), which
// split its partitions by enumeration position and reduce by (cost,
// position) -- the reference's first-minimum rule; with L = 1 a lane takes
// whole subsets.
// Binomial C(n, k) for n <= 7 and the first index of the k-leaf subsets in
// c_sub, in ALU: per-lane reads of constant tables are vector loads whose
// vmcnt wait would also drain the previous round's outstanding node stores.
__device__ __forceinline__ int binom7(int n, int k) {
    int c = 1;
#pragma unroll
    for (int i = 1; i <= kTMax; ++i)
        if (i <= k) c = c * (n - k + i) / i;  // exact at every step
    return c;
}
__device__ __forceinline__ int subk7(int k) {  // sum of C(7, j) for 1 <= j < k
    return k <= 1 ? 0 : k == 2 ? 7 : k == 3 ? 28 : k == 4 ? 63 : k == 5 ? 98 : k == 6 ? 119 : k == 7 ? 126 : 127;
}

__device__ void dp_group(Slot& sl, int sub, int G, const uint8_t* s_sub) {
    const int n = sl.n;
    if (sub < n) sl.cost1[(1 << sub) - 1] = 0.0f;
    wave_sync();
    for (int k = 2; k <= n; ++k) {
        const int cnt = binom7(n, k);
        const int subk = subk7(k);
        int L = 1;
        while (L * 2 * cnt <= G) L *= 2;
        if (L == 1) {
            for (int q = sub; q < cnt; q += G) dp_subset(sl, s_sub[subk + q]);
        } else {
            const int q = sub / L, j = sub & (L - 1);
            float best = kCostInf;
            int bestP = -1, bestT = 1 << 30;
            int s = 0;
            float areaS = 0.0f;
            if (q < cnt) {
                s = s_sub[subk + q];
                areaS = area(subset_box(sl, s));
                const int delta = (s - 1) & s;
                const int m = (1 << __popc(delta)) - 1;
                for (int t = j; t < m; t += L) {
                    const int p = submask(t, delta);
                    const float c = sl.cost1[p - 1] + sl.cost1[(s ^ p) - 1];
                    if (c < best) {
                        best = c;
                        bestP = p;
                        bestT = t;
                    }
                }
            }
            for (int off = L / 2; off >= 1; off >>= 1) {
                const float ob = __shfl_xor(best, off, L);
                const int op = __shfl_xor(bestP, off, L);
                const int ot = __shfl_xor(bestT, off, L);
                if (ob < best || (ob == best && ot < bestT)) {
                    best = ob;
                    bestP = op;
                    bestT = ot;
                }
            }
            if (q < cnt && j == 0) {
                sl.cost1[s - 1] = best + 1.0f * areaS;
                sl.part[s] = (uint8_t)bestP;
            }
        }
        wave_sync();
    }
}

// Reconstruction (reference lbvh_plus_builder.cu:162-245): the treelet's
// internal node ids in DFS pop order (right child popped first), assigned
// breadth first from the partition table; boxes are subset unions.  The DFS
// order is computed in closed form from the expansion records (every node is
// expanded after its parent): an internal node's rank is its parent's + 1 for
// a right child and + 1 + (internal nodes under the right sibling) for a left
// child.  Ranks, the BFS queue and the leaves stay in registers; per BFS node
// only the partition entry and the leaf boxes come from LDS.
__device__ void rebuild(const TreeletArgs& a, Slot& sl, int32_t root) {
    constexpr int kI = kTMax - 1;  // internal nodes of a full treelet
    DNode* __restrict__ nodes = a.nodes;
    const int n = sl.n;
    const int nexp = n - 1;
    const int full = (1 << n) - 1;
    RegTab act = 0, en = INT32_MIN, el = 0, er = 0;
#pragma unroll
    for (int i = 0; i < kTMax; ++i) act[i] = sl.act[i];
#pragma unroll
    for (int e = 0; e < kI; ++e) {
        en[e] = (e < nexp) ? sl.exn[e] : INT32_MIN;
        el[e] = sl.exl[e];
        er[e] = sl.exr[e];
    }
    // Internal nodes under each expanded node (children come later).
    RegTab sz = 0;
#pragma unroll
    for (int e = kI - 1; e >= 0; --e) {
        int32_t c = 0;
        if (e < nexp) {
            c = 1;
#pragma unroll
            for (int f = e + 1; f < kI; ++f)
                if (en[f] == el[e] || en[f] == er[e]) c += sz[f];
        }
        sz[e] = c;
    }
    // Right-first preorder ranks (the root, expanded first, has rank 0).
    RegTab rk = 0;
#pragma unroll
    for (int e = 0; e < kI; ++e) {
        if (e >= nexp) continue;
        int32_t szr = 0;
#pragma unroll
        for (int f = e + 1; f < kI; ++f)
            if (en[f] == er[e]) szr = sz[f];
#pragma unroll
        for (int f = e + 1; f < kI; ++f) {
            if (en[f] == er[e]) rk[f] = rk[e] + 1;
            if (en[f] == el[e]) rk[f] = rk[e] + 1 + szr;
        }
    }
    RegTab avail = root;
#pragma unroll
    for (int r = 0; r < kI; ++r) {
        avail[r] = root;
#pragma unroll
        for (int e = 0; e < kI; ++e)
            if (e < nexp && rk[e] == r) avail[r] = en[e];
    }
    // The BFS takes internal ids in rank order: avail as a FIFO (shifted on
    // every take), the queue appended at the one-hot tail `qt`.
    RegTab qm = 0, qn = 0;
    qm[0] = full;
    qn[0] = avail[0];
    auto take = [&]() {
#pragma unroll
        for (int k = 0; k < kI - 1; ++k) avail[k] = avail[k + 1];
        return avail[0];
    };
    uint32_t qt = 2u;  // one-hot: next free queue slot (slot 0 holds the root)
#pragma unroll
    for (int qh = 0; qh < kI; ++qh) {
        if (!((qt >> qh) > 1u)) continue;  // slot qh not filled
        const int mask = qm[qh];
        const int32_t cur = qn[qh];
        const int lm = sl.part[mask];
        const int rm = mask ^ lm;
        int32_t ln, rn;
        if (__popc(lm) == 1) {
            ln = reg_pick(act, (uint32_t)lm);
        } else {
            ln = take();
            qm = reg_put(qm, qt, lm);
            qn = reg_put(qn, qt, ln);
            qt <<= 1;
        }
        if (__popc(rm) == 1) {
            rn = reg_pick(act, (uint32_t)rm);
        } else {
            rn = take();
            qm = reg_put(qm, qt, rm);
            qn = reg_put(qn, qt, rn);
            qt <<= 1;
        }
        Box b = empty_box();
        int lo = 0x7fffffff, hi = -1;
#pragma unroll
        for (int i = 0; i < kTMax; ++i)
            if (i < n && ((mask >> i) & 1)) {
                b = merge(b, lds_box(sl.lb[i]));
                lo = sl.lo[i] < lo ? sl.lo[i] : lo;
                hi = sl.hi[i] > hi ? sl.hi[i] : hi;
            }
        // Coherent stores: an ancestor's treelet may be formed by another
        // workgroup on another XCD later in this launch.
        store_node_coherent<false>(nodes, (uint32_t)cur, b, ln, rn);  // an existing internal node
        store_u32_coherent(a.parents + ln, (uint32_t)cur);
        store_u32_coherent(a.parents + rn, (uint32_t)cur);
        store_int2_coherent(a.ranges + cur, make_int2(lo, hi));
    }
}

template <int ITEMS>
__global__ __launch_bounds__(kTBlock) void k_treelet(TreeletArgs a) {
    constexpr int kTItems = ITEMS;
    constexpr int kTChunk = kTBlock * ITEMS;
    __shared__ WaveLds s_w[kTWaves];
    __shared__ uint8_t s_sub[127];
    if (threadIdx.x < 127) s_sub[threadIdx.x] = c_sub[threadIdx.x];
    __syncthreads();
    const int lane = threadIdx.x % kWave;
    WaveLds& w = s_w[threadIdx.x / kWave];
    const uint32_t n = a.n;
    const uint32_t ca = blockIdx.x * kTChunk;
    const uint32_t cb = (ca + kTChunk < n) ? ca + kTChunk : n;
    const uint32_t tag = 2 * a.epoch + 1;
    // Every thread advances its kTItems chains together, so a wave always has
    // every ready treelet of all its chains to fill its kGroups slots (one
    // chain after another would leave most rounds with one or two treelets).
    uint32_t cur[kTItems];
    uint32_t act = 0;
#pragma unroll
    for (int it = 0; it < kTItems; ++it) {
        const uint32_t i = ca + it * kTBlock + threadIdx.x;
        cur[it] = n - 1 + i;
        if (i < cb && n > 1) act |= 1u << it;
    }
    while (__ballot(act != 0) != 0) {
        // This lane's record stores (coherent) are complete before its
        // arrivals are counted; a second arrival then reads them with plain
        // loads inside the chunk (same CU) or coherent loads across chunks.
        // No acquire/release fence, so no whole-L2 writeback or invalidate.
        __builtin_amdgcn_s_waitcnt(0);
        uint32_t par[kTItems];
        uint32_t pend = 0, glob = 0;
#pragma unroll
        for (int it = 0; it < kTItems; ++it) {
            par[it] = 0;
            if (!((act >> it) & 1u)) continue;
            const uint32_t pp = a.parents[cur[it]];
            par[it] = pp;
            const int2 sp = a.ranges[pp];
            const bool local = sp.x >= (int)ca && sp.y < (int)cb;
            const uint32_t old =
                local ? __hip_atomic_fetch_max(a.counters + pp, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                      : __hip_atomic_fetch_max(a.counters + pp, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (old == tag) {
                pend |= 1u << it;
                if (!local) glob |= 1u << it;
            } else {
                act &= ~(1u << it);
            }
        }
        // Ready treelets, up to kGroups per round (a lane offers its lowest
        // pending chain): formed by their owners into LDS slots, optimised
        // by lane groups, rebuilt by owners.
        while (true) {
            const bool todo = pend != 0;
            const uint64_t tmask = __ballot(todo);
            if (tmask == 0) break;
            const int it0 = todo ? __ffs((int)pend) - 1 : 0;
            uint32_t root = 0;
            bool global = false;
#pragma unroll
            for (int it = 0; it < kTItems; ++it)
                if (it == it0) {
                    root = par[it];
                    global = (glob >> it) & 1u;
                }
            const int rank = todo ? __popcll(tmask & lanemask_lt()) : kGroups;
            const bool mine = rank < kGroups;
            Slot& ms = w.slot[mine ? rank : 0];
            Box rootbox;
            int tn = 0;
            if (mine) tn = global ? form<true>(a, (int32_t)root, ms, rootbox) : form<false>(a, (int32_t)root, ms, rootbox);
            wave_sync();
            // T treelets this round, each on a group of G lanes.
            const int T = __popcll(tmask) < kGroups ? __popcll(tmask) : kGroups;
            const int G = T == 1 ? 64 : (T == 2 ? 32 : (T <= 4 ? 16 : 8));
            const int gi = lane / G, sub = lane & (G - 1);
            if (gi < T && w.slot[gi].n >= 2) dp_group(w.slot[gi], sub, G, s_sub);
            wave_sync();
            if (mine) {
                if (tn >= 2) rebuild(a, ms, (int32_t)root);
                else nodes_refit(a, (int32_t)root, rootbox);  // nothing to optimise
                pend &= ~(1u << it0);
#pragma unroll
                for (int it = 0; it < kTItems; ++it)
                    if (it == it0) {
                        cur[it] = root;
                        if (root == 0) act &= ~(1u << it);
                    }
            }
            wave_sync();
        }
    }
}

void launch_treelet(const TreeletArgs& a, hipStream_t s) {
    // Large trees: eight chains per thread (longer chunks, fewer straddling
    // nodes); small ones: fewer, so that at least 512 workgroups share the CUs.
    int items = kTItemsMax;
    while (items > 2 && (a.n + kTBlock * items - 1) / (uint32_t)(kTBlock * items) < 512) items >>= 1;
    const uint32_t grid = (a.n + kTBlock * items - 1) / (uint32_t)(kTBlock * items);
    if (items == 8) hipLaunchKernelGGL(k_treelet<8>, dim3(grid), dim3(kTBlock), 0, s, a);
    else if (items == 4) hipLaunchKernelGGL(k_treelet<4>, dim3(grid), dim3(kTBlock), 0, s, a);
    else hipLaunchKernelGGL(k_treelet<2>, dim3(grid), dim3(kTBlock), 0, s, a);
}

}  // namespace bvhk
