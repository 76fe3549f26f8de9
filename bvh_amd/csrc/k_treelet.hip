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
// MI355X mapping: threads climb from the leaves as in the reference, but a
// treelet is processed by its whole wavefront: the lanes that became second
// arrivals are served one at a time, each treelet's subset boxes, costs and
// partitions live in the wave's private LDS slice and every DP level runs with
// one lane per subset.  A workgroup owns 2048 sorted leaves; a node whose leaf
// span lies inside the chunk hands off through a workgroup-scope counter (no
// cross-XCD coherence traffic), only chunk-straddling nodes use agent scope.
// Leaf spans are kept per internal node (min/max sorted leaf position) and
// updated when a treelet reassigns its internal nodes.
#include "dev_common.hpp"
#include "engine.hpp"

namespace bvhk {

constexpr int kTBlock = 256;
constexpr int kTItems = 8;
constexpr int kTChunk = kTBlock * kTItems;
constexpr int kTWaves = kTBlock / kWave;
constexpr int kTMax = 7;

struct WaveLds {
    float box[128][6];   // subset boxes (only needed for reconstruction)
    float cost[128];
    uint8_t part[128];
    int32_t act[8];      // treelet leaves (node ids)
    float abox[8][6];    // their boxes
    int32_t alo[8];      // their leaf spans
    int32_t ahi[8];
    int32_t exp_node[8]; // expanded (internal) nodes and their children, in order
    int32_t exp_l[8];
    int32_t exp_r[8];
    int32_t avail[8];    // reconstruction scratch of the owner lane
    int32_t stack[10];
    int32_t qmask[8];
    int32_t qnode[8];
    int32_t cnt;
    int32_t nexp;
};

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

__device__ __forceinline__ int2 leaf_span(const TreeletArgs& a, int32_t node) {
    const uint32_t first_leaf = a.n - 1;
    if ((uint32_t)node >= first_leaf) return make_int2((int)(node - first_leaf), (int)(node - first_leaf));
    return a.ranges[node];
}

// Refit + optimise the treelet rooted at `root`.  Called by the whole wave
// (uniform control flow); `owner` is the lane that climbs on from `root`.
__device__ void process_treelet(const TreeletArgs& a, WaveLds& w, int32_t root, int lane, int owner) {
    DNode* __restrict__ nodes = a.nodes;
    // Refit (the reference's initial union before optimising).
    const int4 rlr = nodes[root].lr;
    const Box rb = merge(node_box(nodes, rlr.x), node_box(nodes, rlr.y));
    if (lane == owner) {
        nodes[root].mn = make_float4(rb.mnx, rb.mny, rb.mnz, 0.0f);
        nodes[root].mx = make_float4(rb.mxx, rb.mxy, rb.mxz, 0.0f);
    }
    // Formation: serial on lane 0, results in LDS.
    if (lane == 0) {
        w.act[0] = root;
        lds_put(w.abox[0], rb);
        int2 sp = a.ranges[root];
        w.alo[0] = sp.x;
        w.ahi[0] = sp.y;
        int cnt = 1, nexp = 0;
        uint32_t leafmask = 0;  // bit i: treelet leaf i is a BVH leaf
        for (int step = 0; step < kTMax - 1; ++step) {
            int best = -1;
            float maxA = -1.0f;
            for (int i = 0; i < cnt; ++i) {
                if (!((leafmask >> i) & 1u)) {
                    float ar = area(lds_box(w.abox[i]));
                    if (ar > maxA) {
                        maxA = ar;
                        best = i;
                    }
                }
            }
            if (best < 0) break;
            const int32_t ex = w.act[best];
            const int4 lr = nodes[ex].lr;
            w.exp_node[nexp] = ex;
            w.exp_l[nexp] = lr.x;
            w.exp_r[nexp] = lr.y;
            ++nexp;
            const DNode nl = nodes[lr.x];
            const DNode nr = nodes[lr.y];
            w.act[best] = lr.x;
            lds_put(w.abox[best], load_box(reinterpret_cast<const float4*>(&nl)));
            leafmask = (leafmask & ~(1u << best)) | (((nl.lr.x & 0x80000000) != 0 ? 1u : 0u) << best);
            int2 s1 = leaf_span(a, lr.x);
            w.alo[best] = s1.x;
            w.ahi[best] = s1.y;
            w.act[cnt] = lr.y;
            lds_put(w.abox[cnt], load_box(reinterpret_cast<const float4*>(&nr)));
            leafmask |= ((nr.lr.x & 0x80000000) != 0 ? 1u : 0u) << cnt;
            int2 s2 = leaf_span(a, lr.y);
            w.alo[cnt] = s2.x;
            w.ahi[cnt] = s2.y;
            ++cnt;
        }
        w.cnt = cnt;
        w.nexp = nexp;
    }
    wave_sync();
    const int n = w.cnt;
    if (n < 2) return;
    const int full = (1 << n) - 1;
    // Subset boxes in the reference's ascending-index union order, singleton costs.
    for (int s = lane + 1; s <= full; s += kWave) {
        Box b = empty_box();
        for (int i = 0; i < n; ++i)
            if ((s >> i) & 1) b = merge(b, lds_box(w.abox[i]));
        lds_put(w.box[s], b);
        w.cost[s] = (__popc(s) == 1) ? 0.0f : 1e30f;
    }
    wave_sync();
    for (int k = 2; k <= n; ++k) {
        for (int s = lane + 1; s <= full; s += kWave) {
            if (__popc(s) != k) continue;
            const float areaS = area(lds_box(w.box[s]));
            float minCost = 1e30f;
            int bestP = -1;
            const int delta = (s - 1) & s;
            int p = (-delta) & s;
            do {
                const float c = w.cost[p] + w.cost[s ^ p];
                if (c < minCost) {
                    minCost = c;
                    bestP = p;
                }
                p = (p - delta) & s;
            } while (p != 0);
            w.cost[s] = minCost + 1.0f * areaS;
            w.part[s] = (uint8_t)bestP;
        }
        wave_sync();
    }
    // Reconstruction on the owner lane.
    if (lane == owner) {
        int32_t* avail = w.avail;
        int navail = 0;
        int32_t* stack = w.stack;
        int top = 0;
        stack[top++] = root;
        while (top > 0) {
            const int32_t curr = stack[--top];
            bool is_tleaf = false;
            for (int i = 0; i < n; ++i)
                if (w.act[i] == curr) {
                    is_tleaf = true;
                    break;
                }
            if (is_tleaf) continue;
            avail[navail++] = curr;
            int32_t cl = -1, cr = -1;
            for (int e = 0; e < w.nexp; ++e)
                if (w.exp_node[e] == curr) {
                    cl = w.exp_l[e];
                    cr = w.exp_r[e];
                }
            stack[top++] = cl;
            stack[top++] = cr;
        }
        int32_t* qmask = w.qmask;
        int32_t* qnode = w.qnode;
        int qh = 0, qt = 0, used = 1;
        qmask[qt] = full;
        qnode[qt] = avail[0];
        ++qt;
        while (qh < qt) {
            const int mask = qmask[qh];
            const int32_t cur = qnode[qh];
            ++qh;
            const int lm = w.part[mask];
            const int rm = mask ^ lm;
            int32_t ln, rn;
            if (__popc(lm) == 1) {
                ln = w.act[__ffs(lm) - 1];
            } else {
                ln = avail[used++];
                qmask[qt] = lm;
                qnode[qt] = ln;
                ++qt;
            }
            if (__popc(rm) == 1) {
                rn = w.act[__ffs(rm) - 1];
            } else {
                rn = avail[used++];
                qmask[qt] = rm;
                qnode[qt] = rn;
                ++qt;
            }
            const Box b = lds_box(w.box[mask]);
            store_node(nodes, (uint32_t)cur, b, ln, rn);
            a.parents[ln] = (uint32_t)cur;
            a.parents[rn] = (uint32_t)cur;
            int lo = 0x7fffffff, hi = -1;
            for (int i = 0; i < n; ++i)
                if ((mask >> i) & 1) {
                    lo = w.alo[i] < lo ? w.alo[i] : lo;
                    hi = w.ahi[i] > hi ? w.ahi[i] : hi;
                }
            a.ranges[cur] = make_int2(lo, hi);
        }
    }
    wave_sync();
}

__global__ __launch_bounds__(kTBlock) void k_treelet(TreeletArgs a) {
    __shared__ WaveLds s_w[kTWaves];
    const int lane = threadIdx.x % kWave;
    WaveLds& w = s_w[threadIdx.x / kWave];
    const uint32_t n = a.n;
    const uint32_t ca = blockIdx.x * kTChunk;
    const uint32_t cb = (ca + kTChunk < n) ? ca + kTChunk : n;
    const uint32_t tag = 2 * a.epoch + 1;
    for (int item = 0; item < kTItems; ++item) {
        const uint32_t i = ca + item * kTBlock + threadIdx.x;
        bool active = (i < cb) && (n > 1);
        uint32_t cur = n - 1 + i;
        while (__ballot(active) != 0) {
            bool pending = false;
            bool global = false;
            uint32_t par = 0;
            if (active) {
                par = a.parents[cur];
                const int2 sp = a.ranges[par];
                const bool local = sp.x >= (int)ca && sp.y < (int)cb;
                uint32_t old;
                if (local) {
                    old = __hip_atomic_fetch_max(a.counters + par, tag, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
                } else {
                    old = __hip_atomic_fetch_max(a.counters + par, tag, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
                    global = true;
                }
                if (old == tag) pending = true;
                else active = false;
            }
            uint64_t mask = __ballot(pending);
            if (__ballot(global && pending) != 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            while (mask) {
                const int owner = __ffsll((unsigned long long)mask) - 1;
                mask &= mask - 1;
                const int32_t root = __shfl((int)par, owner, kWave);
                process_treelet(a, w, root, lane, owner);
            }
            if (pending) {
                cur = par;
                if (par == 0) active = false;
            }
        }
    }
}

void launch_treelet(const TreeletArgs& a, hipStream_t s) {
    uint32_t grid = (a.n + kTChunk - 1) / kTChunk;
    hipLaunchKernelGGL(k_treelet, dim3(grid), dim3(kTBlock), 0, s, a);
}

}  // namespace bvhk
