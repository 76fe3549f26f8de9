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
//     minimum wins); box(s) = ascending-index union of the leaf boxes, starting
//     from the empty box;
//   * reconstruction reuses the treelet's internal node ids in DFS pop order
//     (right child popped first) and assigns them breadth first.
// Every float operation matches the oracle's serial transcription in order, so
// the resulting trees are bit-identical to it.
//
// Schedule.  In the reference a thread climbs from every leaf and the second
// arrival at a node processes it.  A node is processed once, after its two
// children, which are still its children from before the pass (only the
// treelet of an ancestor, processed later, rewires them), and treelets of
// disjoint subtrees commute, so any bottom-up order gives the same tree.  Here
// a pass (one round) is level-synchronous at the bottom, where nearly all
// nodes are, and climbs at the top:
//   * k_tl_seed, one thread per internal node: a node whose children are two
//     BVH leaves is a two-leaf treelet, processed at once; a node with one
//     leaf child counts that child's arrival.  Second arrivals (atomic
//     maximum of an epoch tag on the node's arrival word) append the node to
//     list 1.
//   * k_tl_level<NM>, one launch per list L = 1..levels while the list holds
//     at least level_min nodes: one treelet per lane.  A node of list L has
//     at most 2^(L+1) leaves below it, so list 1 needs 4-leaf treelets and
//     later lists 7.  Formation keeps the leaf boxes in registers and the ids
//     in a lane-private LDS record; the subset areas come from a depth-first
//     walk over ascending leaf additions (the fold order) and the dynamic
//     programme is unrolled at compile time over subsets in increasing order
//     (every proper subset of s is smaller than s), its costs in registers
//     and its partitions in the LDS record; the wave stops at the largest
//     treelet among its lanes.  Second arrivals at the parents go to list
//     L+1, collected per workgroup.  Records are written with plain stores:
//     the next launch reads them.
//   * k_tl_tail: the first list shorter than level_min (or list levels+1)
//     seeds climbing chains, spread over all waves, that finish the top of
//     the tree as in the reference: ready treelets are formed into LDS slots
//     (agent-coherent loads), up to eight per wave round, optimised and
//     written by lane groups (agent-coherent stores), and every arrival waits
//     for its wave's stores first, so no whole-L2 writeback or invalidate is
//     issued.
#include "dev_common.hpp"
#include "engine.hpp"

namespace bvhk {

constexpr int kTBlock = 256;
constexpr int kTWaves = kTBlock / kWave;
constexpr int kTMax = 7;
constexpr int kGroups = 8;               // treelets per wave per round (groups of >= 8 lanes)
// dp_group initialises a treelet's single-leaf costs one lane per leaf, so a
// group needs at least kTMax lanes.
static_assert(kWave / kGroups >= kTMax, "a treelet group needs a lane per leaf");
constexpr float kCostInf = 1e30f;
constexpr int kLvBlock = 256;            // level kernels: one treelet per thread
constexpr int kTailItems = 2;            // tail chains per thread per pass
constexpr uint32_t kTailGrid = 1024;     // tail workgroups

// ---- shared helpers ------------------------------------------------------------

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ Box node_box_of(const DNode& d) {
    Box b;
    b.mnx = d.mn.x; b.mny = d.mn.y; b.mnz = d.mn.z;
    b.mxx = d.mx.x; b.mxy = d.mx.y; b.mxz = d.mx.z;
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

// Internal-node ranks of a treelet in the reference's DFS pop order (right
// child popped first), in closed form from the expansion records (every node
// is expanded after its parent): a node's rank is its parent's + 1 for a
// right child and + 1 + (internal nodes under the right sibling) for a left
// child.  Returns avail[r] = the node id of rank r.
template <int KI>
__device__ __forceinline__ RegTab dfs_ranks(RegTab en, RegTab el, RegTab er, int nexp, int32_t root) {
    // Internal nodes under each expanded node (children come later).
    RegTab sz = 0;
#pragma unroll
    for (int e = KI - 1; e >= 0; --e) {
        int32_t c = 0;
        if (e < nexp) {
            c = 1;
#pragma unroll
            for (int f = e + 1; f < KI; ++f)
                if (en[f] == el[e] || en[f] == er[e]) c += sz[f];
        }
        sz[e] = c;
    }
    // Right-first preorder ranks (the root, expanded first, has rank 0).
    RegTab rk = 0;
#pragma unroll
    for (int e = 0; e < KI; ++e) {
        if (e >= nexp) continue;
        int32_t szr = 0;
#pragma unroll
        for (int f = e + 1; f < KI; ++f)
            if (en[f] == er[e]) szr = sz[f];
#pragma unroll
        for (int f = e + 1; f < KI; ++f) {
            if (en[f] == er[e]) rk[f] = rk[e] + 1;
            if (en[f] == el[e]) rk[f] = rk[e] + 1 + szr;
        }
    }
    RegTab avail = root;
#pragma unroll
    for (int r = 0; r < KI; ++r) {
        avail[r] = root;
#pragma unroll
        for (int e = 0; e < KI; ++e)
            if (e < nexp && rk[e] == r) avail[r] = en[e];
    }
    return avail;
}

// Breadth-first reconstruction (reference lbvh_plus_builder.cu:162-245):
// internal ids are taken in rank order (avail as a FIFO, shifted on every
// take), the queue is appended at the one-hot tail `qt`.  part(mask) reads
// the partition table, emit(node, mask, left, right) writes one node.
template <int KI, class Part, class Emit>
__device__ __forceinline__ void bfs_rebuild(RegTab act, RegTab avail, int n, Part part, Emit emit) {
    RegTab qm = 0, qn = 0;
    qm[0] = (1 << n) - 1;
    qn[0] = avail[0];
    auto take = [&]() {
#pragma unroll
        for (int k = 0; k < KI - 1; ++k) avail[k] = avail[k + 1];
        return avail[0];
    };
    uint32_t qt = 2u;  // one-hot: next free queue slot (slot 0 holds the root)
#pragma unroll
    for (int qh = 0; qh < KI; ++qh) {
        if (!((qt >> qh) > 1u)) continue;  // slot qh not filled
        const int mask = qm[qh];
        const int32_t cur = qn[qh];
        const int lm = part(mask);
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
        emit(cur, mask, ln, rn);
    }
}

// ---- lane path: one treelet per lane ----------------------------------------------

// Leaf boxes stay in registers; the ids go to the lane's LDS record (they
// are only read back by the reconstruction, and the dynamic programme needs
// the registers).
template <int NM>
struct LaneTreelet {
    float bx[NM][6];      // treelet leaf boxes, in treelet order
};
template <int NM>
struct LaneLds {
    uint8_t part[1 << NM];  // best partition of each subset
    int32_t act[NM];        // treelet leaves (node ids)
    int32_t exn[NM - 1];    // expanded internal nodes, in expansion order, and their children
    int32_t exl[NM - 1];
    int32_t exr[NM - 1];
};
// Lane records at an odd stride in words: the 64 lanes' equal offsets fall in
// distinct banks.
template <int NM>
constexpr int lane_words() {
    return (int)(sizeof(LaneLds<NM>) / 4) | 1;
}

template <int NM>
__device__ __forceinline__ Box lane_box(const LaneTreelet<NM>& t, int i) {
    Box b;
    b.mnx = t.bx[i][0]; b.mny = t.bx[i][1]; b.mnz = t.bx[i][2];
    b.mxx = t.bx[i][3]; b.mxy = t.bx[i][4]; b.mxz = t.bx[i][5];
    return b;
}
template <int NM>
__device__ __forceinline__ void lane_put(LaneTreelet<NM>& t, int i, const Box& b) {
    t.bx[i][0] = b.mnx; t.bx[i][1] = b.mny; t.bx[i][2] = b.mnz;
    t.bx[i][3] = b.mxx; t.bx[i][4] = b.mxy; t.bx[i][5] = b.mxz;
}

// Greedy formation (reference lbvh_plus_builder.cu:79-99) in registers: step
// k adds leaf k + 1, so every table index but the expanded leaf's is a
// compile-time constant; the expanded leaf's entries are one-hot selects.
// Returns the leaf count (1: nothing to optimise) and the refitted root box.
template <int NM>
__device__ __forceinline__ int lane_form(const DNode* __restrict__ nodes, int32_t root, int4 rlr, LaneTreelet<NM>& t,
                                         LaneLds<NM>& rec, Box& rootbox) {
    int32_t act[NM], lc[NM], rc[NM];
    float ar[NM];
    uint32_t leafmask = 0;
    DNode cl = nodes[rlr.x];
    DNode cr = nodes[rlr.y];
    rootbox = merge(node_box_of(cl), node_box_of(cr));
    act[0] = root;
    if (!(area(rootbox) > -1.0f)) return 1;  // the reference's argmax finds nothing (non-finite area)
    int32_t l = rlr.x, r = rlr.y, ex = root;
    int best = 0, n = 1;
#pragma unroll
    for (int step = 0; step < NM - 1; ++step) {
        if (step > 0) {
            best = -1;
            float maxA = -1.0f;
#pragma unroll
            for (int i = 0; i <= step; ++i)
                if (!((leafmask >> i) & 1u) && ar[i] > maxA) {
                    maxA = ar[i];
                    best = i;
                }
            if (best < 0) break;
            ex = act[0];
            l = lc[0];
            r = rc[0];
#pragma unroll
            for (int i = 1; i <= step; ++i)
                if (best == i) {
                    ex = act[i];
                    l = lc[i];
                    r = rc[i];
                }
            cl = nodes[l];
            cr = nodes[r];
        }
        rec.exn[step] = ex;
        rec.exl[step] = l;
        rec.exr[step] = r;
        const Box bl = node_box_of(cl), br = node_box_of(cr);
        const float al = area(bl), arr = area(br);
#pragma unroll
        for (int i = 0; i <= step; ++i)
            if (best == i) {
                act[i] = l;
                lane_put(t, i, bl);
                ar[i] = al;
                lc[i] = cl.lr.x;
                rc[i] = cl.lr.y;
                leafmask = (leafmask & ~(1u << i)) | ((cl.lr.x < 0 ? 1u : 0u) << i);
            }
        act[step + 1] = r;
        lane_put(t, step + 1, br);
        ar[step + 1] = arr;
        lc[step + 1] = cr.lr.x;
        rc[step + 1] = cr.lr.y;
        leafmask |= (cr.lr.x < 0 ? 1u : 0u) << (step + 1);
        n = step + 2;
    }
#pragma unroll
    for (int i = 0; i < NM; ++i)
        if (i < n) rec.act[i] = act[i];
    return n;
}

// Box union with hardware min/max, for boxes used only through their area:
// on finite coordinates (meshes with NaN or infinite vertices are rejected
// before any build) it differs from merge() at most in the sign of a zero,
// which changes no area's value, so every cost comparison of the dynamic
// programme comes out as with merge().
__device__ __forceinline__ Box merge_area(const Box& a, const Box& b) {
    Box r;
    r.mnx = fminf(a.mnx, b.mnx);
    r.mny = fminf(a.mny, b.mny);
    r.mnz = fminf(a.mnz, b.mnz);
    r.mxx = fmaxf(a.mxx, b.mxx);
    r.mxy = fmaxf(a.mxy, b.mxy);
    r.mxz = fmaxf(a.mxz, b.mxz);
    return r;
}

// Areas of every subset of the first nmax leaves with two or more leaves, at
// c[s]: depth first, a child subset adds a leaf above the parent's highest,
// so each box is its parent's box merged with one leaf -- the reference's
// ascending fold from the empty box, one merge per subset.
template <int NM, int S, int J>
struct AreaWalk {
    static __device__ __forceinline__ void run(float (&c)[1 << NM], const LaneTreelet<NM>& t, const Box& b, int nmax) {
        if constexpr (J < NM) {
            if (J < nmax) {
                constexpr int C = S | (1 << J);
                const Box bc = merge_area(b, lane_box(t, J));
                if constexpr ((C & (C - 1)) != 0) c[C] = area(bc);
                AreaWalk<NM, C, J + 1>::run(c, t, bc, nmax);
                AreaWalk<NM, S, J + 1>::run(c, t, b, nmax);
            }
        }
    }
};

// Partitions of subset S in the reference's order (p -> (p - delta) & S from
// (-delta) & S, delta = S without its lowest leaf), strict '<'.  The winner is
// kept as its enumeration position K (below 64: an inline constant, where
// partition masks above 64 would each take a register).
template <int NM, int S, int D, int P, int K>
struct DpParts {
    static __device__ __forceinline__ void run(const float (&c)[1 << NM], float& best, int& bk) {
        const float v = c[P] + c[S ^ P];
        if (v < best) {
            best = v;
            bk = K;
        }
        constexpr int NX = (P - D) & S;
        if constexpr (NX != 0) DpParts<NM, S, D, NX, K + 1>::run(c, best, bk);
    }
};

// Costs in increasing subset order: c[s] holds area(s) on entry and cost(s)
// after; singletons cost 0; the subsets of the first k leaves are exactly
// those below 2^k, so the walk stops at 2^nmax.  part[s] receives the best
// partition's enumeration position (submask() turns it back into the mask).
template <int NM, int S>
struct DpWalk {
    static __device__ __forceinline__ void run(float (&c)[1 << NM], uint8_t* part, int nmax) {
        if constexpr (S < (1 << NM)) {
            if constexpr ((S & (S - 1)) == 0) {
                if (S >= (1 << nmax)) return;
                c[S] = 0.0f;
            } else {
                constexpr int D = (S - 1) & S;
                float best = kCostInf;
                int bk = -1;
                DpParts<NM, S, D, ((-D) & S), 0>::run(c, best, bk);
                c[S] = best + 1.0f * c[S];
                part[S] = (uint8_t)bk;
            }
            DpWalk<NM, S + 1>::run(c, part, nmax);
        }
    }
};

template <int NM>
__device__ __forceinline__ void lane_rebuild(const TreeletArgs& a, const LaneTreelet<NM>& t, const LaneLds<NM>& rec,
                                             int n, int32_t root) {
    constexpr int kI = NM - 1;
    DNode* __restrict__ nodes = a.nodes;
    uint32_t* __restrict__ parents = a.parents;
    RegTab act = 0, en = INT32_MIN, el = 0, er = 0;
#pragma unroll
    for (int i = 0; i < NM; ++i) act[i] = (i < n) ? rec.act[i] : 0;
    const int nexp = n - 1;
#pragma unroll
    for (int e = 0; e < kI; ++e) {
        en[e] = (e < nexp) ? rec.exn[e] : INT32_MIN;
        el[e] = (e < nexp) ? rec.exl[e] : 0;
        er[e] = (e < nexp) ? rec.exr[e] : 0;
    }
    const RegTab avail = dfs_ranks<kI>(en, el, er, nexp, root);
    bfs_rebuild<kI>(
        act, avail, n, [&](int mask) { return submask((int)rec.part[mask], (mask - 1) & mask); },
        [&](int32_t cur, int mask, int32_t ln, int32_t rn) {
            Box b = empty_box();
#pragma unroll
            for (int i = 0; i < NM; ++i)
                if (i < n && ((mask >> i) & 1)) b = merge(b, lane_box(t, i));
            store_node(nodes, (uint32_t)cur, b, ln, rn);  // an existing internal node
            parents[ln] = (uint32_t)cur;
            parents[rn] = (uint32_t)cur;
        });
}

// One treelet per lane: formation, dynamic programme (the wave runs to its
// largest treelet), reconstruction or refit.
template <int NM>
__device__ __forceinline__ void lane_treelet(const TreeletArgs& a, bool have, int32_t u, int4 lr, LaneLds<NM>& rec) {
    LaneTreelet<NM> t;
#pragma unroll
    for (int i = 0; i < NM; ++i) lane_put(t, i, Box{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f});
    Box rootbox;
    int n = 0;
    if (have) n = lane_form<NM>(a.nodes, u, lr, t, rec, rootbox);
    const int nmax = (int)wave_max_u32((uint32_t)n);
    if (nmax >= 2) {
        float c[1 << NM];
        AreaWalk<NM, 0, 0>::run(c, t, empty_box(), nmax);
        DpWalk<NM, 1>::run(c, rec.part, nmax);
    }
    if (have) {
        if (n >= 2) {
            lane_rebuild<NM>(a, t, rec, n, u);
        } else {
            float4* p = reinterpret_cast<float4*>(a.nodes + u);
            store_box(p, rootbox);  // nothing to optimise: refit only
        }
    }
}

// Second arrivals of a workgroup collect in LDS (one LDS atomic per wave) and
// reach the global list in blocks, one counter atomic per kFlushIters
// iterations: every wave appending through the one list counter directly
// serialises tens of thousands of same-address atomics per launch.
constexpr int kFlushIters = 8;
constexpr int kAppendCap = kLvBlock * kFlushIters;
struct BlockList {
    uint32_t buf[kAppendCap];
    uint32_t n, base;
};

// Counts an arrival at `target`; the second arrival appends it to the
// workgroup's buffer.
__device__ __forceinline__ void arrive_local(const TreeletArgs& a, bool have, uint32_t target, uint32_t tag,
                                             BlockList& bl) {
    bool second = false;
    if (have) {
        const uint32_t old =
            __hip_atomic_fetch_max(a.counters + target, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        second = old == tag;
    }
    const uint64_t m = __ballot(second);
    if (m == 0) return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t pos = 0;
    if ((int)lane_id() == leader) pos = atomicAdd(&bl.n, (uint32_t)__popcll(m));
    pos = __shfl(pos, leader);
    if (second) bl.buf[pos + __popcll(m & lanemask_lt())] = target;
}

// Workgroup-uniform: moves the buffered nodes to the list at `out`.
__device__ __forceinline__ void flush_list(BlockList& bl, uint32_t* out_count, uint32_t* out) {
    __syncthreads();
    if (threadIdx.x == 0) bl.base = bl.n ? atomicAdd(out_count, bl.n) : 0;
    __syncthreads();
    const uint32_t k = bl.n, b = bl.base;
    for (uint32_t i = threadIdx.x; i < k; i += kLvBlock) out[b + i] = bl.buf[i];
    __syncthreads();
    if (threadIdx.x == 0) bl.n = 0;
    __syncthreads();
}

template <int NM>
__device__ __forceinline__ LaneLds<NM>& lane_rec(uint32_t* s_rec) {
    return *reinterpret_cast<LaneLds<NM>*>(s_rec + threadIdx.x * lane_words<NM>());
}

// Round start: every internal node once, kAppendCap per workgroup.  Two leaf
// children: a two-leaf treelet, processed now, then its arrival at the
// parent; one leaf child: that child's arrival at the node.  Second arrivals
// go to list 1.
__global__ __launch_bounds__(kLvBlock) void k_tl_seed(TreeletArgs a) {
    __shared__ uint32_t s_rec[kLvBlock * lane_words<2>()];
    __shared__ BlockList s_list;
    if (threadIdx.x == 0) s_list.n = 0;
    __syncthreads();
    const uint32_t first_leaf = a.n - 1;
    const uint32_t tag = 2 * a.epoch + 1;
    for (int it = 0; it < kFlushIters; ++it) {
        const uint32_t u = blockIdx.x * kAppendCap + it * kLvBlock + threadIdx.x;
        const bool valid = u < first_leaf;
        const int4 lr = valid ? a.nodes[u].lr : make_int4(0, 0, 0, 0);
        const int nl = valid ? (((uint32_t)lr.x >= first_leaf) ? 1 : 0) + (((uint32_t)lr.y >= first_leaf) ? 1 : 0) : 0;
        const bool two = nl == 2;
        if (__ballot(two) != 0) lane_treelet<2>(a, two, (int32_t)u, lr, lane_rec<2>(s_rec));
        uint32_t target = u;
        if (two) target = u == 0 ? 0 : a.parents[u];
        const bool arrives = nl == 1 || (two && u != 0);
        arrive_local(a, arrives, target, tag, s_list);
    }
    flush_list(s_list, a.cnt + 1, a.list);
}

// First index of list L: lists 1, 2, ... follow each other.
__device__ __forceinline__ uint32_t list_base(const TreeletArgs& a, int level) {
    uint32_t b = 0;
    for (int j = 1; j < level; ++j) b += a.cnt[j];
    return b;
}

// ---- group path: treelets in LDS slots, lane-group DP --------------------------

// One treelet of at most kTMax leaves in LDS: formed by its owner lane,
// optimised by one lane group, rebuilt by the owner.
struct Slot {
    float lb[kTMax][6];                 // treelet leaf boxes, in treelet order
    float cost1[(1 << kTMax) - 1];      // cost of leaf subset s at cost1[s - 1]
    uint8_t part[1 << kTMax];
    int32_t act[kTMax];                 // treelet leaves (node ids)
    int32_t exn[kTMax - 1];             // expanded internal nodes and their children, in order
    int32_t exl[kTMax - 1];
    int32_t exr[kTMax - 1];
    int32_t qn[kTMax - 1];              // reconstruction plan, in BFS order: node, leaf mask, children
    int32_t qm[kTMax - 1];
    int32_t ql[kTMax - 1];
    int32_t qr[kTMax - 1];
    int32_t n;  // leaves; the treelet has n - 1 expanded (internal) nodes
};
struct WaveLds {
    Slot slot[kGroups];
};
// An odd slot stride in words puts equal offsets of the eight slots in eight
// different LDS banks (the groups walk identical subset sequences).
static_assert(sizeof(Slot) % 8 == 4, "slot stride must be an odd number of words");
static_assert(sizeof(WaveLds) * kTWaves + 127 <= 64 * 1024, "tail LDS (slots + subset table) must stay within 64 KB");

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

// Greedy formation into an LDS slot with agent-coherent loads: other
// workgroups, maybe on other XCDs, wrote the records during this launch.  The leaves' areas and child ids
// are kept in registers as well (the argmax and the next expansion's loads
// need no LDS round trip); LDS gets what the DP and the rebuild read.
__device__ int slot_form(const TreeletArgs& a, int32_t root, Slot& sl, Box& rootbox) {
    typedef float AreaTab __attribute__((ext_vector_type(8)));
    const DNode* __restrict__ nodes = a.nodes;
    const int4 rlr = load_node_coherent(nodes, (uint32_t)root).lr;
    sl.act[0] = root;
    RegTab lcv = 0, rcv = 0;
    AreaTab ar = -1.0f;
    lcv[0] = rlr.x;
    rcv[0] = rlr.y;
    uint32_t leafmask = 0;
    int n = 1, nexp = 0;
    DNode cl = load_node_coherent(nodes, (uint32_t)rlr.x);
    DNode cr = load_node_coherent(nodes, (uint32_t)rlr.y);
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
            cl = load_node_coherent(nodes, (uint32_t)reg_pick(lcv, 1u << best));
            cr = load_node_coherent(nodes, (uint32_t)reg_pick(rcv, 1u << best));
        } else if (!(ar[0] > -1.0f)) {
            break;  // the reference's argmax finds nothing (non-finite area)
        }
        const uint32_t bbit = 1u << best, nbit = 1u << n;
        const int32_t l = reg_pick(lcv, bbit), r = reg_pick(rcv, bbit);
        sl.exn[nexp] = sl.act[best];
        sl.exl[nexp] = l;
        sl.exr[nexp] = r;
        ++nexp;
        const Box bl = node_box_of(cl), br = node_box_of(cr);
        sl.act[best] = l;
        lds_put(sl.lb[best], bl);
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
        lcv = reg_put(lcv, nbit, cr.lr.x);
        rcv = reg_put(rcv, nbit, cr.lr.y);
        leafmask |= (cr.lr.x < 0 ? 1u : 0u) << n;
        ++n;
    }
    sl.n = n;
    return n;
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

// The t-th (0-based) nonzero submask of `mask` (at most kTMax - 1 bits) in
// increasing order, branch-free: t + 1 with its bits placed at mask's bits.
__device__ __forceinline__ int deposit(int t, int mask) {
    int v = t + 1, r = 0;
#pragma unroll
    for (int j = 0; j < kTMax - 1; ++j) {
        r |= (v & 1) ? (mask & -mask) : 0;
        v >>= 1;
        mask &= mask - 1;
    }
    return r;
}

// The dynamic programme of one treelet on a group of G lanes (G a power of
// two, 8..64; sub = lane in group).  First every subset's area, lanes
// striding over the subsets, each box folded from the leaf boxes in
// ascending order from the empty box; singletons cost 0.  Then by subset
// size: subsets of one size are independent and each gets L lanes (the
// largest power of two with L * count <= G), which split its partitions by
// enumeration position, four loads in flight per lane, and reduce by (cost,
// position) -- the reference's first-minimum rule; with L = 1 a lane takes
// whole subsets.  cost1[s - 1] holds area(s) until cost(s) replaces it.
__device__ void dp_group(Slot& sl, int sub, int G, const uint8_t* s_sub) {
    const int n = sl.n;
    const int full = (1 << n) - 1;
    {
        Box lb[kTMax];
#pragma unroll
        for (int i = 0; i < kTMax; ++i) lb[i] = lds_box(sl.lb[i]);
        for (int s = sub + 1; s <= full; s += G) {
            Box b = empty_box();
#pragma unroll
            for (int i = 0; i < kTMax; ++i)
                if ((s >> i) & 1) b = merge_area(b, lb[i]);
            sl.cost1[s - 1] = (s & (s - 1)) ? area(b) : 0.0f;
        }
    }
    wave_sync();
    for (int k = 2; k <= n; ++k) {
        const int cnt = binom7(n, k);
        const int subk = subk7(k);
        const int m = (1 << (k - 1)) - 1;  // partitions of a k-leaf subset
        int L = 1;
        while (L * 2 * cnt <= G) L *= 2;
        const int j = sub & (L - 1);
        for (int q = sub / L; q < cnt; q += G / L) {  // L > 1: one subset per lane group
            const int s = s_sub[subk + q];
            const int delta = (s - 1) & s;
            float best = kCostInf;
            int bestT = 1 << 30;
            for (int t0 = j; t0 < m; t0 += 4 * L) {
                float c[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int t = t0 + u * L;
                    const int p = t < m ? deposit(t, delta) : delta;
                    c[u] = sl.cost1[p - 1] + sl.cost1[(s ^ p) - 1];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (t0 + u * L < m && c[u] < best) {
                        best = c[u];
                        bestT = t0 + u * L;
                    }
            }
            for (int off = L / 2; off >= 1; off >>= 1) {
                const float ob = __shfl_xor(best, off, L);
                const int ot = __shfl_xor(bestT, off, L);
                if (ob < best || (ob == best && ot < bestT)) {
                    best = ob;
                    bestT = ot;
                }
            }
            if (j == 0) {
                sl.cost1[s - 1] = best + 1.0f * sl.cost1[s - 1];
                sl.part[s] = bestT < m ? (uint8_t)deposit(bestT, delta) : (uint8_t)0xFF;
            }
        }
        wave_sync();
    }
}

// Reconstruction from the slot, by the group: the owner derives the plan
// (DFS ranks, breadth-first assignment from the partition table) into the
// slot, then lane q of the group folds the box of plan entry q and writes
// that node.
__device__ __forceinline__ void slot_plan(Slot& sl, int32_t root) {
    constexpr int kI = kTMax - 1;
    const int n = sl.n;
    const int nexp = n - 1;
    RegTab act = 0, en = INT32_MIN, el = 0, er = 0;
#pragma unroll
    for (int i = 0; i < kTMax; ++i) act[i] = sl.act[i];
#pragma unroll
    for (int e = 0; e < kI; ++e) {
        en[e] = (e < nexp) ? sl.exn[e] : INT32_MIN;
        el[e] = sl.exl[e];
        er[e] = sl.exr[e];
    }
    const RegTab avail = dfs_ranks<kI>(en, el, er, nexp, root);
    int q = 0;
    bfs_rebuild<kI>(
        act, avail, n, [&](int mask) { return (int)sl.part[mask]; },
        [&](int32_t cur, int mask, int32_t ln, int32_t rn) {
            sl.qn[q] = cur;
            sl.qm[q] = mask;
            sl.ql[q] = ln;
            sl.qr[q] = rn;
            ++q;
        });
}
__device__ __forceinline__ void slot_emit(const TreeletArgs& a, const Slot& sl, int q) {
    const int mask = sl.qm[q];
    Box b = empty_box();
#pragma unroll
    for (int i = 0; i < kTMax; ++i)
        if ((mask >> i) & 1) b = merge(b, lds_box(sl.lb[i]));
    const int32_t cur = sl.qn[q], ln = sl.ql[q], rn = sl.qr[q];
    // Coherent stores: an ancestor's treelet may be formed by another
    // workgroup, maybe on another XCD, later in this launch.
    store_node_coherent<false>(a.nodes, (uint32_t)cur, b, ln, rn);  // an existing internal node
    store_u32_coherent(a.parents + ln, (uint32_t)cur);
    store_u32_coherent(a.parents + rn, (uint32_t)cur);
}

// List L (at least level_min nodes; shorter lists go to the climbing tail):
// one treelet per lane, in contiguous per-workgroup ranges of whole
// iterations of kLvBlock.
template <int NM>
__global__ __launch_bounds__(kLvBlock) void k_tl_level(TreeletArgs a, int level) {
    __shared__ uint32_t s_rec[kLvBlock * lane_words<NM>()];
    __shared__ BlockList s_list;
    const uint32_t cnt = a.cnt[level];
    if (cnt < a.level_min) return;  // this list and everything above go to the climbing tail
    const uint32_t per = ((cnt + gridDim.x - 1) / gridDim.x + kLvBlock - 1) / kLvBlock * kLvBlock;
    const uint32_t lo = blockIdx.x * per;
    if (lo >= cnt) return;
    const uint32_t hi = min(cnt, lo + per);
    if (threadIdx.x == 0) s_list.n = 0;
    __syncthreads();
    const uint32_t base = list_base(a, level);
    const uint32_t tag = 2 * a.epoch + 1;
    uint32_t* out = a.list + base + cnt;
    LaneLds<NM>& rec = lane_rec<NM>(s_rec);
    int it = 0;
    for (uint32_t i0 = lo; i0 < hi; i0 += kLvBlock) {
        const uint32_t e = i0 + threadIdx.x;
        const bool have = e < hi;
        const int32_t u = have ? (int32_t)a.list[base + e] : 0;
        const int4 lr = have ? a.nodes[u].lr : make_int4(0, 0, 0, 0);
        lane_treelet<NM>(a, have, u, lr, rec);
        const bool up = have && u != 0;
        const uint32_t target = up ? a.parents[u] : 0;
        arrive_local(a, up, target, tag, s_list);
        if (++it == kFlushIters) {
            flush_list(s_list, a.cnt + level + 1, out);
            it = 0;
        }
    }
    if (it != 0) flush_list(s_list, a.cnt + level + 1, out);
}

// The top of the tree: chains seeded by the first list shorter than
// level_min (or list levels + 1), spread over all waves of the grid, climb as
// in the reference: each chain processes its ready treelet and moves up while
// it is the second arrival.  A wave round takes up to kGroups ready treelets
// (a lane offers its lowest pending chain), formed by their owners into LDS
// slots with coherent loads, optimised and written by lane groups.
__global__ __launch_bounds__(kTBlock) void k_tl_tail(TreeletArgs a) {
    __shared__ WaveLds s_w[kTWaves];
    __shared__ uint8_t s_sub[127];
    if (threadIdx.x < 127) s_sub[threadIdx.x] = c_sub[threadIdx.x];
    __syncthreads();
    int level = 1;
    uint32_t base = 0;
    while (level <= a.levels && a.cnt[level] >= a.level_min) base += a.cnt[level++];
    const uint32_t cnt = a.cnt[level];
    const int lane = (int)lane_id();
    WaveLds& w = s_w[threadIdx.x / kWave];
    const uint32_t waves = gridDim.x * kTWaves;
    const uint32_t gwave = blockIdx.x * kTWaves + threadIdx.x / kWave;
    const uint32_t tag = 2 * a.epoch + 1;
    for (uint32_t q = 0; q < cnt; q += waves * kWave * kTailItems) {
        // Entry k of this pass goes to wave k % waves: a short list puts one
        // or two chains on each wave, so every wave round has few treelets
        // and wide lane groups.
        uint32_t cur[kTailItems];
        uint32_t act = 0;
#pragma unroll
        for (int it = 0; it < kTailItems; ++it) {
            const uint32_t e = q + ((uint32_t)(it * kWave + lane) * waves + gwave);
            cur[it] = 0;
            if (e < cnt) {
                cur[it] = a.list[base + e];
                act |= 1u << it;
            }
        }
        uint32_t pend = act;  // list entries are ready treelet roots
        while (__ballot(act != 0) != 0) {
            while (true) {
                const bool todo = pend != 0;
                const uint64_t tmask = __ballot(todo);
                if (tmask == 0) break;
                const int it0 = todo ? __ffs((int)pend) - 1 : 0;
                uint32_t root = 0;
#pragma unroll
                for (int it = 0; it < kTailItems; ++it)
                    if (it == it0) root = cur[it];
                const int rank = todo ? __popcll(tmask & lanemask_lt()) : kGroups;
                const bool mine = rank < kGroups;
                Slot& ms = w.slot[mine ? rank : 0];
                Box rootbox;
                int tn = 0;
                if (mine) tn = slot_form(a, (int32_t)root, ms, rootbox);
                wave_sync();
                // T treelets this round, each on a group of G lanes.
                const int T = __popcll(tmask) < kGroups ? __popcll(tmask) : kGroups;
                const int G = T == 1 ? 64 : (T == 2 ? 32 : (T <= 4 ? 16 : 8));
                const int gi = lane / G, sub = lane & (G - 1);
                if (gi < T && w.slot[gi].n >= 2) dp_group(w.slot[gi], sub, G, s_sub);
                wave_sync();
                if (mine) {
                    if (tn >= 2) slot_plan(ms, (int32_t)root);
                    else store_box_coherent(a.nodes, root, rootbox);  // nothing to optimise
                    pend &= ~(1u << it0);
                }
                wave_sync();
                if (gi < T && sub < w.slot[gi].n - 1) slot_emit(a, w.slot[gi], sub);
                wave_sync();
            }
            // The wave's record stores (coherent) are complete before its
            // arrivals are counted; a second arrival then reads them with
            // coherent loads.  No acquire/release fence, so no whole-L2
            // writeback or invalidate.
            __builtin_amdgcn_s_waitcnt(0);
#pragma unroll
            for (int it = 0; it < kTailItems; ++it) {
                if (!((act >> it) & 1u)) continue;
                if (cur[it] == 0) {  // the root is done
                    act &= ~(1u << it);
                    continue;
                }
                const uint32_t pp = a.parents[cur[it]];
                const uint32_t old =
                    __hip_atomic_fetch_max(a.counters + pp, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (old == tag) {
                    cur[it] = pp;
                    pend |= 1u << it;
                } else {
                    act &= ~(1u << it);
                }
            }
        }
    }
}

void launch_treelet(const TreeletArgs& a, hipStream_t s) {
    launch_fill_zero(a.cnt, (size_t)a.levels + 2, s);
    const uint32_t first_leaf = a.n - 1;
    hipLaunchKernelGGL(k_tl_seed, dim3((first_leaf + kAppendCap - 1) / kAppendCap), dim3(kLvBlock), 0, s, a);
    for (int level = 1; level <= a.levels; ++level) {
        // A node of list L covers at least L + 2 leaves and the nodes of one
        // list are disjoint subtrees; longer lists give each workgroup more
        // iterations.
        const uint64_t bound = (uint64_t)a.n / (uint64_t)(level + 2) + 1;
        const uint32_t grid = (uint32_t)std::min<uint64_t>((bound + kLvBlock - 1) / kLvBlock, 2048);
        if (level == 1) hipLaunchKernelGGL(k_tl_level<4>, dim3(grid), dim3(kLvBlock), 0, s, a, level);
        else hipLaunchKernelGGL(k_tl_level<kTMax>, dim3(grid), dim3(kLvBlock), 0, s, a, level);
    }
    hipLaunchKernelGGL(k_tl_tail, dim3(kTailGrid), dim3(kTBlock), 0, s, a);
}

}  // namespace bvhk
