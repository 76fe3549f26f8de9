// oracle.cpp — TEST INFRASTRUCTURE ONLY.
//
// Serial C++ restatement of the reference BVH build and traversal algorithms,
// used as the correctness oracle for the HIP kernels (and, in its OpenMP
// variant, as the reported CPU baseline).  Nothing in the product links,
// loads or calls this file; only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py do.
//
// Compiled with -O2 -ffp-contract=off and without -ffast-math: every float
// operation is a single IEEE-754 operation in source order, the same
// sequence the gfx950 kernels execute, so results are compared bit for bit.
//
// min/max are selects: min(a,b) = (b < a) ? b : a, max(a,b) = (a < b) ? b : a
// (what std::min/std::max compute; reference include/common.h:56-62 host path).
//
// Function -> reference it follows -> what pins it (tests/test_oracle_*.py):
//   tri_box            src/cuda/lbvh_builder.cu:47-66       brute-force min/max per triangle, hand-derived
//                                                          signed-zero box fixture (test_oracle_pins.py)
//   scene_box          lbvh_builder.cu:73-165                brute-force min/max over all vertices
//   norm_centroid      lbvh_builder.cu:167-181               mesh-level known cells: exact cell-centre
//                                                          meshes over non-unit boxes, box-centroid vs
//                                                          vertex-mean triangle, the ext > 1e-6 rule at,
//                                                          below and above 1e-6 (test_oracle_pins.py)
//   morton30           src/cuda/lbvh_builder.cu:11-27,167-181  expandBits known answers, bit-loop
//                                                          interleave, cell boundaries in closed form
//   morton60           src/cuda/lbvh_plus_builder.cu:20-35,276-290  same, z-major 20-bit interleave
//   sort               lbvh_builder.cu:412-439 (stable)      numpy lexsort on (code, index)
//   karras             lbvh_builder.cu:33-43,183-229         Karras 2012 Fig. 3 8-key example,
//                      lbvh_plus_builder.cu:41-47,292-324     top-down radix-tree recursion, invariants
//   refit              lbvh_builder.cu:243-274                brute-force union over each subtree
//   treelet optimise   lbvh_plus_builder.cu:71-247,341-366    brute force over all binary trees on
//                                                          the treelet leaves (DP optimality), leaf
//                                                          multiset and parent links preserved
//   ploc               src/cuda/ploc_builder.cu:173-264,403-586  naive O(m^2) window transcription in
//                                                          Python on small n, progress invariant
//   trace              src/render_engine.cu:194-356          brute-force closest hit over all triangles
//   trace_nearest      (no reference counterpart; predicates  brute-force least (t, index) hit over all
//                      of render_engine.cu:194-273)           triangles; visit counts against a
//                                                          pure-Python transcription on small trees
//   sah                src/evaluator.cpp:38-88               closed forms: n identical boxes give
//                                                          Ct(n-1)+Ci*n; hand-computed 3-leaf tree at
//                                                          (3,2) and (1,1); zero-area root
//   depth_stats        src/evaluator.cpp:90-113              complete 2^k-leaf trees and caterpillars
//   *_omp variants     (CPU baseline only)                   equal to the serial functions above on
//                                                          blob, uniform and strand meshes
//                                                          (test_oracle_omp.py)
// The reference itself cannot be compiled here (its headers need the CUDA
// toolkit), so no function is checked against reference binaries.
// Known deviations from the reference, recorded as semantics (not bugs):
//   * PLOC numbers new nodes in ascending cluster order within an iteration
//     (one legal serialisation of the reference's atomicAdd, ploc_builder.cu:245).
//   * The traversal stack never drops entries (reference drops pushes beyond 128,
//     render_engine.cu:346-351); identical whenever the tree depth is < 128.
//   * IEEE division/sqrt (the reference builds with --use_fast_math).

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

inline float fmin_s(float a, float b) { return (b < a) ? b : a; }
inline float fmax_s(float a, float b) { return (a < b) ? b : a; }

struct Box {
    float mn[3];
    float mx[3];
};

inline Box empty_box() {
    Box b;
    for (int k = 0; k < 3; ++k) { b.mn[k] = FLT_MAX; b.mx[k] = -FLT_MAX; }
    return b;
}

// merge(a, b): reference include/common.h:46-64.
inline Box merge(const Box& a, const Box& b) {
    Box r;
    for (int k = 0; k < 3; ++k) { r.mn[k] = fmin_s(a.mn[k], b.mn[k]); r.mx[k] = fmax_s(a.mx[k], b.mx[k]); }
    return r;
}

// surfaceArea: reference include/common.h:41-44, evaluated in source order.
inline float area(const Box& b) {
    float dx = b.mx[0] - b.mn[0];
    float dy = b.mx[1] - b.mn[1];
    float dz = b.mx[2] - b.mn[2];
    float t1 = dx * dy;
    float t2 = dy * dz;
    float t3 = dz * dx;
    float s = t1 + t2;
    s = s + t3;
    return 2.0f * s;
}

// 48-byte record, byte-compatible with BVHNode (include/bvh_node.h).
struct ONode {
    float mn[3];
    float pad0;
    float mx[3];
    float pad1;
    int32_t left;
    int32_t right;
    int32_t pad2;
    int32_t pad3;
};
static_assert(sizeof(ONode) == 48, "node record must be 48 bytes");

inline Box node_box(const ONode& n) {
    Box b;
    for (int k = 0; k < 3; ++k) { b.mn[k] = n.mn[k]; b.mx[k] = n.mx[k]; }
    return b;
}
inline void set_node(ONode& n, const Box& b, int32_t l, int32_t r) {
    for (int k = 0; k < 3; ++k) { n.mn[k] = b.mn[k]; n.mx[k] = b.mx[k]; }
    n.pad0 = n.pad1 = 0.0f;
    n.left = l;
    n.right = r;
    n.pad2 = n.pad3 = 0;
}

struct Soa {
    const float* v[9];
    float get(int comp, int64_t i) const { return v[comp][i]; }
};

// Per-triangle box: min(v0, min(v1, v2)) per axis (lbvh_builder.cu:55-60).
inline Box tri_box(const Soa& m, int64_t i) {
    Box b;
    for (int k = 0; k < 3; ++k) {
        float a0 = m.v[k][i], a1 = m.v[3 + k][i], a2 = m.v[6 + k][i];
        b.mn[k] = fmin_s(a0, fmin_s(a1, a2));
        b.mx[k] = fmax_s(a0, fmax_s(a1, a2));
    }
    return b;
}

Box scene_box(const Soa& m, int64_t n) {
    Box s = empty_box();
    for (int64_t i = 0; i < n; ++i) s = merge(s, tri_box(m, i));
    return s;
}

// Normalised centroid coordinate (lbvh_builder.cu:174-176).
inline void norm_centroid(const Box& tb, const Box& scene, float out[3]) {
    for (int k = 0; k < 3; ++k) {
        float c = (tb.mn[k] + tb.mx[k]) * 0.5f;
        float ext = scene.mx[k] - scene.mn[k];
        float div = (ext > 1e-6f) ? ext : 1.0f;
        out[k] = (c - scene.mn[k]) / div;
    }
}

// expandBits_nt (lbvh_builder.cu:11-17): 10 bits -> every third bit of 30.
inline uint32_t expand10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// morton3D_nt (lbvh_builder.cu:19-27): x in the most significant bit of each triple.
inline uint32_t morton30(float x, float y, float z) {
    x = fmin_s(fmax_s(x * 1024.0f, 0.0f), 1023.0f);
    y = fmin_s(fmax_s(y * 1024.0f, 0.0f), 1023.0f);
    z = fmin_s(fmax_s(z * 1024.0f, 0.0f), 1023.0f);
    uint32_t xx = expand10((uint32_t)x);
    uint32_t yy = expand10((uint32_t)y);
    uint32_t zz = expand10((uint32_t)z);
    return xx * 4 + yy * 2 + zz;
}

// expandBits_P (lbvh_plus_builder.cu:20-28): 21 bits -> every third bit of 63.
inline uint64_t expand21(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

// morton3D_P (lbvh_plus_builder.cu:30-35): z in the most significant bit.
inline uint64_t morton60(float x, float y, float z) {
    x = fmin_s(fmax_s(x * 1048576.0f, 0.0f), 1048575.0f);
    y = fmin_s(fmax_s(y * 1048576.0f, 0.0f), 1048575.0f);
    z = fmin_s(fmax_s(z * 1048576.0f, 0.0f), 1048575.0f);
    return (expand21((uint64_t)z) << 2) | (expand21((uint64_t)y) << 1) | expand21((uint64_t)x);
}

inline int clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }
inline int clz64(uint64_t x) { return x == 0 ? 64 : __builtin_clzll(x); }

// delta_nt (lbvh_builder.cu:33-43) over sorted (code30, index).
struct Delta30 {
    const uint32_t* code;
    const uint32_t* idx;
    int64_t n;
    int operator()(int64_t i, int64_t j) const {
        if (j < 0 || j >= n) return -1;
        if (code[i] == code[j]) return 32 + clz32(idx[i] ^ idx[j]);
        return clz32(code[i] ^ code[j]);
    }
};

// delta_P (lbvh_plus_builder.cu:41-47) over sorted (code60, index).
struct Delta60 {
    const uint64_t* code;
    const uint32_t* idx;
    int64_t n;
    int operator()(int64_t i, int64_t j) const {
        if (j < 0 || j >= n) return -1;
        if (code[i] == code[j]) return 64 + clz64((uint64_t)idx[i] ^ (uint64_t)idx[j]);
        return clz64(code[i] ^ code[j]);
    }
};

// Working node of the LBVH builders (reference src/cuda/lbvh_builder.cuh:9-14).
struct LNode {
    Box box;
    uint32_t left;
    uint32_t right;
    uint32_t parent;
};

// kBuildInternalNodes_nt / _P for one internal node i (lbvh_builder.cu:183-229).
template <class D>
void karras_node(const D& delta, int64_t n, int64_t i, std::vector<LNode>& nodes) {
    int64_t d = (delta(i, i + 1) - delta(i, i - 1)) >= 0 ? 1 : -1;
    int min_delta = delta(i, i - d);
    int64_t l_max = 2;
    while (delta(i, i + l_max * d) > min_delta) l_max *= 2;
    int64_t l = 0;
    for (int64_t t = l_max / 2; t >= 1; t /= 2)
        if (delta(i, i + (l + t) * d) > min_delta) l += t;
    int64_t j = i + l * d;
    int delta_node = delta(i, j);
    int64_t s = 0;
    int64_t first = (d > 0) ? i : j;
    int64_t last = (d > 0) ? j : i;
    int64_t t = last - first;
    do {
        t = (t + 1) >> 1;
        if (delta(first, first + s + t) > delta_node) s += t;
    } while (t > 1);
    int64_t split = first + s;
    uint32_t leftIdx = (split == first) ? (uint32_t)((n - 1) + split) : (uint32_t)split;
    uint32_t rightIdx = (split + 1 == last) ? (uint32_t)((n - 1) + split + 1) : (uint32_t)(split + 1);
    nodes[i].left = leftIdx;
    nodes[i].right = rightIdx;
    nodes[leftIdx].parent = (uint32_t)i;
    nodes[rightIdx].parent = (uint32_t)i;
}

// kInitLeafNodes (lbvh_builder.cu:231-241).
void init_leaves(const Soa& m, int64_t n, const uint32_t* sorted_idx, std::vector<LNode>& nodes) {
    for (int64_t i = 0; i < n; ++i) {
        LNode& L = nodes[(n - 1) + i];
        L.box = tri_box(m, sorted_idx[i]);
        L.left = sorted_idx[i] | 0x80000000u;
        L.right = 0xFFFFFFFFu;
    }
}

// ---------------------------------------------------------------------------
// Treelet optimisation (lbvh_plus_builder.cu:71-247), one treelet rooted at rootIdx.
constexpr int kTreelet = 7;
constexpr float kInf = 1e30f;

void optimize_treelet(std::vector<LNode>& nodes, int rootIdx) {
    int act[kTreelet];
    Box actBox[kTreelet];
    act[0] = rootIdx;
    actBox[0] = nodes[rootIdx].box;
    int leafCount = 1;
    for (int step = 0; step < kTreelet - 1; ++step) {
        int bestIdx = -1;
        float maxArea = -1.0f;
        for (int i = 0; i < leafCount; ++i) {
            bool isLeaf = (nodes[act[i]].left & 0x80000000u) != 0;
            if (!isLeaf) {
                float a = area(actBox[i]);
                if (a > maxArea) { maxArea = a; bestIdx = i; }
            }
        }
        if (bestIdx == -1) break;
        const LNode ex = nodes[act[bestIdx]];
        act[bestIdx] = (int)ex.left;
        actBox[bestIdx] = nodes[ex.left].box;
        act[leafCount] = (int)ex.right;
        actBox[leafCount] = nodes[ex.right].box;
        leafCount++;
    }
    const int n = leafCount;
    if (n < 2) return;

    float cost[128];
    uint8_t partition[128];
    for (int s = 1; s < (1 << n); ++s) cost[s] = (__builtin_popcount(s) == 1) ? 0.0f : kInf;
    for (int k = 2; k <= n; ++k) {
        for (int s = 1; s < (1 << n); ++s) {
            if (__builtin_popcount(s) != k) continue;
            Box b = empty_box();
            for (int i = 0; i < n; ++i)
                if ((s >> i) & 1) b = merge(b, actBox[i]);
            float areaS = area(b);
            float minCost = kInf;
            int bestP = -1;
            int delta = (s - 1) & s;
            int p = (-delta) & s;
            do {
                int other = s ^ p;
                float c = cost[p] + cost[other];
                if (c < minCost) { minCost = c; bestP = p; }
                p = (p - delta) & s;
            } while (p != 0);
            cost[s] = minCost + 1.0f * areaS;
            partition[s] = (uint8_t)bestP;
        }
    }

    int avail[kTreelet];
    int availCount = 0;
    int stack[kTreelet];
    int top = 0;
    stack[top++] = rootIdx;
    while (top > 0) {
        int curr = stack[--top];
        bool inLeaves = false;
        for (int i = 0; i < n; ++i)
            if (act[i] == curr) { inLeaves = true; break; }
        if (!inLeaves) {
            avail[availCount++] = curr;
            stack[top++] = (int)nodes[curr].left;
            stack[top++] = (int)nodes[curr].right;
        }
    }

    struct Task { int mask; int node; };
    Task queue[kTreelet];
    int head = 0, tail = 0;
    queue[tail++] = {(1 << n) - 1, avail[0]};
    int used = 1;
    while (head < tail) {
        Task t = queue[head++];
        int leftMask = partition[t.mask];
        int rightMask = t.mask ^ leftMask;
        int leftNode, rightNode;
        if (__builtin_popcount(leftMask) == 1) {
            leftNode = act[__builtin_ffs(leftMask) - 1];
        } else {
            leftNode = avail[used++];
            queue[tail++] = {leftMask, leftNode};
        }
        if (__builtin_popcount(rightMask) == 1) {
            rightNode = act[__builtin_ffs(rightMask) - 1];
        } else {
            rightNode = avail[used++];
            queue[tail++] = {rightMask, rightNode};
        }
        LNode& node = nodes[t.node];
        node.left = (uint32_t)leftNode;
        node.right = (uint32_t)rightNode;
        Box b = empty_box();
        for (int i = 0; i < n; ++i)
            if ((t.mask >> i) & 1) b = merge(b, actBox[i]);
        node.box = b;
        nodes[leftNode].parent = (uint32_t)t.node;
        nodes[rightNode].parent = (uint32_t)t.node;
    }
}

// kRefitHierarchy_nt / kRefitAndOptimize, executed as one legal serial schedule:
// leaf threads in ascending order, each climbing until it is the first arrival
// (lbvh_builder.cu:243-274, lbvh_plus_builder.cu:341-366).
void refit(std::vector<LNode>& nodes, int64_t n, bool optimize) {
    std::vector<int> counter(2 * n - 1, 0);
    for (int64_t i = 0; i < n; ++i) {
        uint32_t idx = (uint32_t)((n - 1) + i);
        while (idx != 0) {
            uint32_t parent = nodes[idx].parent;
            if (counter[parent]++ == 0) break;
            nodes[parent].box = merge(nodes[nodes[parent].left].box, nodes[nodes[parent].right].box);
            if (optimize) optimize_treelet(nodes, (int)parent);
            idx = parent;
        }
    }
}

void lnodes_to_out(const std::vector<LNode>& nodes, ONode* out) {
    for (size_t i = 0; i < nodes.size(); ++i) set_node(out[i], nodes[i].box, (int32_t)nodes[i].left, (int32_t)nodes[i].right);
}

// ---------------------------------------------------------------------------
// Traversal (render_engine.cu:155-356).
struct Ray {
    float o[3];
    float d[3];
    float inv[3];
};

inline Ray make_ray(const float* r) {
    Ray ray;
    for (int k = 0; k < 3; ++k) {
        ray.o[k] = r[k];
        ray.d[k] = r[3 + k];
        float dk = ray.d[k];
        ray.inv[k] = (std::fabs(dk) > 1e-6f) ? 1.0f / dk : (dk >= 0 ? 1e30f : -1e30f);
    }
    return ray;
}

// intersectAABB (render_engine.cu:194-214).
inline bool hit_box(const Ray& ray, const ONode& nd, float tmax) {
    float tminv[3], tmaxv[3];
    for (int k = 0; k < 3; ++k) {
        float t1 = (nd.mn[k] - ray.o[k]) * ray.inv[k];
        float t2 = (nd.mx[k] - ray.o[k]) * ray.inv[k];
        tminv[k] = fmin_s(t1, t2);
        tmaxv[k] = fmax_s(t1, t2);
    }
    float tenter = fmax_s(fmax_s(tminv[0], tminv[1]), tminv[2]);
    float texit = fmin_s(fmin_s(tmaxv[0], tmaxv[1]), tmaxv[2]);
    return texit >= fmax_s(tenter, 0.0f) && tenter < tmax;
}

// intersectTriangle (render_engine.cu:220-273), Moller-Trumbore, no contraction.
inline bool hit_tri(const Ray& ray, const Soa& m, uint32_t p, float& out_t) {
    float v0x = m.v[0][p], v0y = m.v[1][p], v0z = m.v[2][p];
    float v1x = m.v[3][p], v1y = m.v[4][p], v1z = m.v[5][p];
    float v2x = m.v[6][p], v2y = m.v[7][p], v2z = m.v[8][p];
    float e1x = v1x - v0x, e1y = v1y - v0y, e1z = v1z - v0z;
    float e2x = v2x - v0x, e2y = v2y - v0y, e2z = v2z - v0z;
    const float dx = ray.d[0], dy = ray.d[1], dz = ray.d[2];
    float px = dy * e2z - dz * e2y;
    float py = dz * e2x - dx * e2z;
    float pz = dx * e2y - dy * e2x;
    float det = e1x * px + e1y * py + e1z * pz;
    if (std::fabs(det) < 1e-6f) return false;
    float inv_det = 1.0f / det;
    float tx = ray.o[0] - v0x, ty = ray.o[1] - v0y, tz = ray.o[2] - v0z;
    float u = (tx * px + ty * py + tz * pz) * inv_det;
    if (u < -1e-5f || u > 1.00001f) return false;
    float qx = ty * e1z - tz * e1y;
    float qy = tz * e1x - tx * e1z;
    float qz = tx * e1y - ty * e1x;
    float v = (dx * qx + dy * qy + dz * qz) * inv_det;
    if (v < -1e-5f || u + v > 1.00001f) return false;
    float t = (e2x * qx + e2y * qy + e2z * qz) * inv_det;
    if (t < 1e-6f) return false;
    out_t = t;
    return true;
}

// traverseBVH (render_engine.cu:279-356): DFS, right pushed before left,
// box test against the current closest t, strict '<' on hits.
inline void trace_one(const ONode* nodes, int64_t num_nodes, const Soa& m, const float* r, float& out_t,
                      int32_t& out_prim, int32_t* counters, std::vector<int32_t>& stack) {
    Ray ray = make_ray(r);
    float best = 1e30f;
    int32_t prim = -1;
    int32_t visited = 0, aabb = 0, tris = 0;
    stack.clear();
    stack.push_back(0);
    while (!stack.empty()) {
        int32_t idx = stack.back();
        stack.pop_back();
        if (idx < 0 || idx >= num_nodes) continue;
        const ONode& nd = nodes[idx];
        visited++;
        aabb++;
        if (!hit_box(ray, nd, best)) continue;
        bool leaf = (nd.left < 0);
        if (leaf) {
            uint32_t p = (uint32_t)nd.left & 0x7FFFFFFFu;
            tris++;
            float t;
            if (hit_tri(ray, m, p, t)) {
                if (t < best) { best = t; prim = (int32_t)p; }
            }
        } else {
            if (nd.right >= 0 && nd.right < num_nodes) stack.push_back(nd.right);
            if (nd.left >= 0 && nd.left < num_nodes) stack.push_back(nd.left);
        }
    }
    out_t = best;
    out_prim = prim;
    if (counters) { counters[0] = visited; counters[1] = aabb; counters[2] = tris; }
}

// Slab test of intersectAABB (render_engine.cu:194-214) without the closest-t
// cull: true when the ray's interval through the box is non-empty; te = the
// entry distance.
inline bool slab_te(const Ray& ray, const ONode& nd, float& te) {
    float tminv[3], tmaxv[3];
    for (int k = 0; k < 3; ++k) {
        float t1 = (nd.mn[k] - ray.o[k]) * ray.inv[k];
        float t2 = (nd.mx[k] - ray.o[k]) * ray.inv[k];
        tminv[k] = fmin_s(t1, t2);
        tmaxv[k] = fmax_s(t1, t2);
    }
    te = fmax_s(fmax_s(tminv[0], tminv[1]), tminv[2]);
    float texit = fmin_s(fmin_s(tmaxv[0], tmaxv[1]), tmaxv[2]);
    return texit >= fmax_s(te, 0.0f);
}

// Nearest-first closest hit (no reference counterpart; the engine's
// BVH_TRACE_NEAREST order): the same slab and triangle predicates as
// traverseBVH, but of two children that pass, the one with the smaller entry
// distance is visited first (left on a tie); an entry whose entry distance
// exceeds the closest t is skipped (kept when equal); the hit is the least
// (t, triangle index) pair.  Counters as traverseBVH counts: every child in
// range is a node visited and an AABB test when its parent is expanded (the
// root once), every leaf entered a triangle test.
inline void trace_one_nearest(const ONode* nodes, int64_t num_nodes, const Soa& m, const float* r, float& out_t,
                              int32_t& out_prim, int32_t* counters, std::vector<std::pair<int32_t, float>>& stack) {
    Ray ray = make_ray(r);
    float best = 1e30f;
    int32_t prim = -1;
    int32_t visited = 0, aabb = 0, tris = 0;
    stack.clear();
    if (num_nodes > 0) {
        visited++;
        aabb++;
        float te;
        if (slab_te(ray, nodes[0], te)) stack.push_back({0, te});
    }
    while (!stack.empty()) {
        const std::pair<int32_t, float> e = stack.back();
        stack.pop_back();
        if (e.second > best) continue;
        const ONode& nd = nodes[e.first];
        if (nd.left < 0) {
            uint32_t p = (uint32_t)nd.left & 0x7FFFFFFFu;
            tris++;
            float t;
            if (hit_tri(ray, m, p, t) && (t < best || (t == best && (int32_t)p < prim))) {
                best = t;
                prim = (int32_t)p;
            }
            continue;
        }
        const int32_t c[2] = {nd.left, nd.right};
        bool keep[2] = {false, false};
        float te[2] = {0.0f, 0.0f};
        for (int k = 0; k < 2; ++k) {
            if (c[k] < 0 || c[k] >= num_nodes) continue;
            visited++;
            aabb++;
            keep[k] = slab_te(ray, nodes[c[k]], te[k]) && te[k] <= best;
        }
        if (keep[0] && keep[1]) {
            const int nearer = te[1] < te[0] ? 1 : 0;
            stack.push_back({c[1 - nearer], te[1 - nearer]});
            stack.push_back({c[nearer], te[nearer]});
        } else if (keep[0] || keep[1]) {
            const int k = keep[0] ? 0 : 1;
            stack.push_back({c[k], te[k]});
        }
    }
    out_t = best;
    out_prim = prim;
    if (counters) { counters[0] = visited; counters[1] = aabb; counters[2] = tris; }
}

// computeSAHRecursive (evaluator.cpp:54-88) with parameterised costs.
float sah_rec(const ONode* nodes, int64_t num, int64_t idx, float rootArea, float ct, float ci) {
    if (idx < 0 || idx >= num) return 0.0f;
    const ONode& nd = nodes[idx];
    float nodeArea = area(node_box(nd));
    float prob = nodeArea / rootArea;
    if (nd.left < 0) return prob * ci * 1.0f;
    float cost = prob * ct;
    if (nd.left >= 0 && nd.left < num) cost += sah_rec(nodes, num, nd.left, rootArea, ct, ci);
    if (nd.right >= 0 && nd.right < num) cost += sah_rec(nodes, num, nd.right, rootArea, ct, ci);
    return cost;
}

void depth_rec(const ONode* nodes, int64_t num, int64_t idx, int depth, int& maxDepth, double& sumDepth,
               int64_t& leaves) {
    if (idx < 0 || idx >= num) return;
    maxDepth = std::max(maxDepth, depth);
    const ONode& nd = nodes[idx];
    if (nd.left < 0) { sumDepth += depth; leaves++; return; }
    depth_rec(nodes, num, nd.left, depth + 1, maxDepth, sumDepth, leaves);
    depth_rec(nodes, num, nd.right, depth + 1, maxDepth, sumDepth, leaves);
}

// Stable order of (code, index) pairs whose input is in index order.
template <class K>
void stable_sort_codes(int64_t n, const K* codes, std::vector<K>& sc, std::vector<uint32_t>& si) {
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return codes[a] < codes[b]; });
    sc.resize(n);
    si.resize(n);
    for (int64_t i = 0; i < n; ++i) { si[i] = perm[i]; sc[i] = codes[perm[i]]; }
}

}  // namespace

extern "C" {

int orc_version() { return 1; }

void orc_tri_boxes(const float* const* soa, int64_t n, float* out8) {
    Soa m;
    for (int k = 0; k < 9; ++k) m.v[k] = soa[k];
    for (int64_t i = 0; i < n; ++i) {
        Box b = tri_box(m, i);
        float* o = out8 + 8 * i;
        o[0] = b.mn[0]; o[1] = b.mn[1]; o[2] = b.mn[2]; o[3] = 0.0f;
        o[4] = b.mx[0]; o[5] = b.mx[1]; o[6] = b.mx[2]; o[7] = 0.0f;
    }
}

void orc_scene_box(const float* const* soa, int64_t n, float* out6) {
    Soa m;
    for (int k = 0; k < 9; ++k) m.v[k] = soa[k];
    Box s = scene_box(m, n);
    for (int k = 0; k < 3; ++k) { out6[k] = s.mn[k]; out6[3 + k] = s.mx[k]; }
}

uint32_t orc_expand10(uint32_t v) { return expand10(v); }
uint64_t orc_expand21(uint64_t v) { return expand21(v); }
uint32_t orc_morton30_point(float x, float y, float z) { return morton30(x, y, z); }
uint64_t orc_morton60_point(float x, float y, float z) { return morton60(x, y, z); }

// Morton codes in input order (kComputeMortonCodes_nt, lbvh_builder.cu:167-181).
void orc_morton30(const float* const* soa, int64_t n, uint32_t* codes) {
    Soa m;
    for (int k = 0; k < 9; ++k) m.v[k] = soa[k];
    Box s = scene_box(m, n);
    for (int64_t i = 0; i < n; ++i) {
        float c[3];
        norm_centroid(tri_box(m, i), s, c);
        codes[i] = morton30(c[0], c[1], c[2]);
    }
}

// 60-bit codes (kComputeMortonCodes_P, lbvh_plus_builder.cu:276-290).
void orc_morton60(const float* const* soa, int64_t n, uint64_t* codes) {
    Soa m;
    for (int k = 0; k < 9; ++k) m.v[k] = soa[k];
    Box s = scene_box(m, n);
    for (int64_t i = 0; i < n; ++i) {
        float c[3];
        norm_centroid(tri_box(m, i), s, c);
        codes[i] = morton60(c[0], c[1], c[2]);
    }
}

// Karras topology over caller-given sorted (code30, index) -> internal child
// links only (for the Karras 2012 Fig. 3 known answer).  left/right: n-1 entries.
void orc_karras30(const uint32_t* sorted_codes, const uint32_t* sorted_idx, int64_t n, uint32_t* left,
                  uint32_t* right, uint32_t* parent) {
    std::vector<LNode> nodes(2 * n - 1);
    Delta30 d{sorted_codes, sorted_idx, n};
    for (int64_t i = 0; i + 1 < n; ++i) karras_node(d, n, i, nodes);
    for (int64_t i = 0; i + 1 < n; ++i) { left[i] = nodes[i].left; right[i] = nodes[i].right; }
    if (parent)
        for (int64_t i = 0; i < 2 * n - 1; ++i) parent[i] = (i == 0) ? 0xFFFFFFFFu : nodes[i].parent;
}

// Full LBVH build (LBVHBuilderCUDA::runCompute + downloadResults).
// nodes: 2n-1 records; indices: n; sorted_keys (optional): (code << 32 | index).
void orc_build_lbvh(const float* const* soa, int64_t n, void* nodes_out, uint32_t* indices,
                    uint64_t* sorted_keys) {
    Soa m;
    for (int k = 0; k < 9; ++k) m.v[k] = soa[k];
    std::vector<uint32_t> codes(n);
    orc_morton30(soa, n, codes.data());
    std::vector<uint32_t> sc, si;
    stable_sort_codes<uint32_t>(n, codes.data(), sc, si);
    if (sorted_keys)
        for (int64_t i = 0; i < n; ++i) sorted_keys[i] = ((uint64_t)sc[i] << 32) | si[i];
    std::vector<LNode> nodes(2 * n - 1);
    Delta30 d{sc.data(), si.data(), n};
    for (int64_t i = 0; i + 1 < n; ++i) karras_node(d, n, i, nodes);
    init_leaves(m, n, si.data(), nodes);
    refit(nodes, n, false);
    lnodes_to_out(nodes, (ONode*)nodes_out);
    std::memcpy(indices, si.data(), n * sizeof(uint32_t));
}

// LBVH+ (LBVHPlusBuilderCUDA::runCompute): 60-bit codes, refit, then `rounds`
// refit+optimise passes (the reference runs one).
void orc_build_lbvh_plus(const float* const* soa, int64_t n, int rounds, void* nodes_out, uint32_t* indices,
                         uint64_t* sorted_codes) {
    Soa m;
    for (int k = 0; k < 9; ++k) m.v[k] = soa[k];
    std::vector<uint64_t> codes(n);
    orc_morton60(soa, n, codes.data());
    std::vector<uint64_t> sc;
    std::vector<uint32_t> si;
    stable_sort_codes<uint64_t>(n, codes.data(), sc, si);
    if (sorted_codes) std::memcpy(sorted_codes, sc.data(), n * sizeof(uint64_t));
    std::vector<LNode> nodes(2 * n - 1);
    Delta60 d{sc.data(), si.data(), n};
    for (int64_t i = 0; i + 1 < n; ++i) karras_node(d, n, i, nodes);
    init_leaves(m, n, si.data(), nodes);
    refit(nodes, n, false);
    for (int r = 0; r < rounds; ++r) refit(nodes, n, true);
    lnodes_to_out(nodes, (ONode*)nodes_out);
    std::memcpy(indices, si.data(), n * sizeof(uint32_t));
}

// One treelet optimisation on a caller-supplied LBVH node array (48-B records)
// plus parent links; used by the brute-force DP test.
void orc_optimize_treelet(void* nodes_io, uint32_t* parent_io, int64_t num_nodes, int root) {
    ONode* on = (ONode*)nodes_io;
    std::vector<LNode> nodes(num_nodes);
    for (int64_t i = 0; i < num_nodes; ++i) {
        nodes[i].box = node_box(on[i]);
        nodes[i].left = (uint32_t)on[i].left;
        nodes[i].right = (uint32_t)on[i].right;
        nodes[i].parent = parent_io[i];
    }
    optimize_treelet(nodes, root);
    lnodes_to_out(nodes, on);
    for (int64_t i = 0; i < num_nodes; ++i) parent_io[i] = nodes[i].parent;
}

// PLOC (PLOCBuilderCUDA::build, ploc_builder.cu:403-586) with deterministic
// node numbering.  Returns 0, or 5 when an iteration makes no merge.
// iter_m (optional, capacity max_iters): cluster count entering each iteration.
int orc_build_ploc(const float* const* soa, int64_t n, int radius, void* nodes_out, uint32_t* indices,
                   int32_t* iter_m, int32_t max_iters, int32_t* n_iters) {
    Soa m;
    for (int k = 0; k < 9; ++k) m.v[k] = soa[k];
    ONode* out = (ONode*)nodes_out;
    std::vector<uint32_t> codes(n);
    orc_morton30(soa, n, codes.data());
    std::vector<uint32_t> sc, si;
    stable_sort_codes<uint32_t>(n, codes.data(), sc, si);
    std::memcpy(indices, si.data(), n * sizeof(uint32_t));

    std::vector<Box> cb(n), cb2(n);
    std::vector<int32_t> cid(n), cid2(n), nn(n), valid(n);
    for (int64_t i = 0; i < n; ++i) {
        cb[i] = tri_box(m, si[i]);
        set_node(out[i], cb[i], (int32_t)(si[i] | 0x80000000u), -1);
        cid[i] = (int32_t)i;
    }
    int64_t cur = n;
    int64_t next = n;
    int32_t iters = 0;
    while (cur > 1) {
        if (iter_m && iters < max_iters) iter_m[iters] = (int32_t)cur;
        iters++;
        // kFindNearestNeighbors_ploc (ploc_builder.cu:173-209).
        for (int64_t i = 0; i < cur; ++i) {
            float best = FLT_MAX;
            int64_t bj = -1;
            int64_t lo = std::max<int64_t>(0, i - radius), hi = std::min<int64_t>(cur - 1, i + radius);
            for (int64_t j = lo; j <= hi; ++j) {
                if (j == i) continue;
                float dist = area(merge(cb[i], cb[j]));
                if (dist < best) { best = dist; bj = j; }
                else if (dist == best) { if (bj == -1 || j < bj) bj = j; }
            }
            nn[i] = (int32_t)bj;
        }
        // kMergeClusters_ploc (ploc_builder.cu:211-264), ascending i.
        int64_t merges = 0;
        for (int64_t i = 0; i < cur; ++i) {
            int32_t j = nn[i];
            bool mutual = (j != -1) && nn[j] == (int32_t)i;
            if (mutual && i < j) {
                Box u = merge(cb[i], cb[j]);
                int64_t id = next++;
                set_node(out[id], u, cid[i], cid[j]);
                cid[i] = (int32_t)id;
                cb[i] = u;
                valid[i] = 1;
                merges++;
            } else if (mutual && j < i) {
                valid[i] = 0;
            } else {
                valid[i] = 1;
            }
        }
        if (merges == 0) {
            if (n_iters) *n_iters = iters;
            return 5;
        }
        // Exclusive scan + compaction (ploc_builder.cu:524-540).
        int64_t w = 0;
        for (int64_t i = 0; i < cur; ++i)
            if (valid[i]) { cb2[w] = cb[i]; cid2[w] = cid[i]; ++w; }
        std::swap(cb, cb2);
        std::swap(cid, cid2);
        cur = w;
    }
    if (n_iters) *n_iters = iters;
    // Root fix (ploc_builder.cu:563-582): root (last allocated) <-> node 0.
    int64_t total = next;
    if (total > 1) {
        int64_t root = total - 1;
        std::swap(out[0], out[root]);
        for (int64_t k = 0; k < total; ++k) {
            ONode& nd = out[k];
            if (nd.left < 0) continue;
            auto fix = [&](int32_t& c) {
                if (!(c & 0x80000000)) {
                    if (c == 0) c = (int32_t)root;
                    else if (c == root) c = 0;
                }
            };
            fix(nd.left);
            fix(nd.right);
        }
    }
    return 0;
}

// Closest-hit traversal of rays (6 floats each: origin, direction).
// counters (optional): 3 int per ray (nodes visited, AABB tests, triangle tests).
void orc_trace(const void* nodes, int64_t num_nodes, const float* const* soa, const float* rays, int64_t nrays,
               float* t_out, int32_t* prim_out, int32_t* counters) {
    Soa m;
    for (int k = 0; k < 9; ++k) m.v[k] = soa[k];
    const ONode* nd = (const ONode*)nodes;
    std::vector<int32_t> stack;
    stack.reserve(256);
    for (int64_t r = 0; r < nrays; ++r)
        trace_one(nd, num_nodes, m, rays + 6 * r, t_out[r], prim_out[r], counters ? counters + 3 * r : nullptr,
                  stack);
}

void orc_trace_omp(const void* nodes, int64_t num_nodes, const float* const* soa, const float* rays,
                   int64_t nrays, float* t_out, int32_t* prim_out, int32_t* counters) {
    Soa m;
    for (int k = 0; k < 9; ++k) m.v[k] = soa[k];
    const ONode* nd = (const ONode*)nodes;
#pragma omp parallel
    {
        std::vector<int32_t> stack;
        stack.reserve(256);
#pragma omp for schedule(dynamic, 1024)
        for (int64_t r = 0; r < nrays; ++r)
            trace_one(nd, num_nodes, m, rays + 6 * r, t_out[r], prim_out[r],
                      counters ? counters + 3 * r : nullptr, stack);
    }
}

// Nearest-first traversal (trace_one_nearest) of rays; counters optional.
void orc_trace_nearest(const void* nodes, int64_t num_nodes, const float* const* soa, const float* rays,
                       int64_t nrays, float* t_out, int32_t* prim_out, int32_t* counters) {
    Soa m;
    for (int k = 0; k < 9; ++k) m.v[k] = soa[k];
    const ONode* nd = (const ONode*)nodes;
#pragma omp parallel
    {
        std::vector<std::pair<int32_t, float>> stack;
        stack.reserve(256);
#pragma omp for schedule(dynamic, 1024)
        for (int64_t r = 0; r < nrays; ++r)
            trace_one_nearest(nd, num_nodes, m, rays + 6 * r, t_out[r], prim_out[r],
                              counters ? counters + 3 * r : nullptr, stack);
    }
}

// Brute-force closest hit over every triangle (same predicates and strict '<'
// in index order); pins the traversal on small meshes.
void orc_trace_brute(const float* const* soa, int64_t n, const float* rays, int64_t nrays, float* t_out,
                     int32_t* prim_out) {
    Soa m;
    for (int k = 0; k < 9; ++k) m.v[k] = soa[k];
    for (int64_t r = 0; r < nrays; ++r) {
        Ray ray = make_ray(rays + 6 * r);
        float best = 1e30f;
        int32_t prim = -1;
        for (int64_t p = 0; p < n; ++p) {
            float t;
            if (hit_tri(ray, m, (uint32_t)p, t) && t < best) { best = t; prim = (int32_t)p; }
        }
        t_out[r] = best;
        prim_out[r] = prim;
    }
}

float orc_sah(const void* nodes, int64_t num, float ct, float ci) {
    const ONode* nd = (const ONode*)nodes;
    if (num == 0) return 0.0f;
    float rootArea = area(node_box(nd[0]));
    if (rootArea <= 0.0f) return 0.0f;
    return sah_rec(nd, num, 0, rootArea, ct, ci);
}

// Max depth, mean leaf depth, leaf count (evaluator.cpp:90-113).
void orc_depth_stats(const void* nodes, int64_t num, int32_t* max_depth, double* avg_leaf_depth,
                     int64_t* leaves) {
    int md = 0;
    double sd = 0;
    int64_t lv = 0;
    if (num > 0) depth_rec((const ONode*)nodes, num, 0, 0, md, sd, lv);
    *max_depth = md;
    *avg_leaf_depth = lv ? sd / (double)lv : 0.0;
    *leaves = lv;
}

// ---------------------------------------------------------------------------
// OpenMP LBVH (CPU baseline).  Same results as orc_build_lbvh: parallel boxes
// and Morton codes, LSD radix sort of (code << 32 | index) over bits 32..61,
// parallel Karras, atomic-counter refit.  stage_ms (optional, 5 entries).
int orc_omp_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_build_lbvh_omp(const float* const* soa, int64_t n, void* nodes_out, uint32_t* indices,
                        double* stage_ms) {
    Soa m;
    for (int k = 0; k < 9; ++k) m.v[k] = soa[k];
    auto now = []() {
#ifdef _OPENMP
        return omp_get_wtime() * 1e3;
#else
        return 0.0;
#endif
    };
    double t0 = now();
    std::vector<Box> tb(n);
    Box scene = empty_box();
#pragma omp parallel
    {
        Box loc = empty_box();
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) { tb[i] = tri_box(m, i); loc = merge(loc, tb[i]); }
#pragma omp critical
        scene = merge(scene, loc);
    }
    double t1 = now();
    std::vector<uint64_t> keys(n), tmp(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        float c[3];
        norm_centroid(tb[i], scene, c);
        keys[i] = ((uint64_t)morton30(c[0], c[1], c[2]) << 32) | (uint64_t)i;
    }
    double t2 = now();
    // LSD radix: 8-bit digits over bits 32..63 (codes use 30 bits).
    int nt = orc_omp_threads();
    std::vector<int64_t> hist((size_t)nt * 256);
    for (int shift = 32; shift < 62; shift += 8) {
        std::fill(hist.begin(), hist.end(), 0);
#pragma omp parallel num_threads(nt)
        {
            int tid = 0;
#ifdef _OPENMP
            tid = omp_get_thread_num();
#endif
            int64_t lo = n * tid / nt, hi = n * (tid + 1) / nt;
            int64_t* h = &hist[(size_t)tid * 256];
            for (int64_t i = lo; i < hi; ++i) h[(keys[i] >> shift) & 255]++;
#pragma omp barrier
#pragma omp single
            {
                int64_t run = 0;
                for (int b = 0; b < 256; ++b)
                    for (int t = 0; t < nt; ++t) { int64_t c = hist[(size_t)t * 256 + b]; hist[(size_t)t * 256 + b] = run; run += c; }
            }
            for (int64_t i = lo; i < hi; ++i) tmp[h[(keys[i] >> shift) & 255]++] = keys[i];
        }
        std::swap(keys, tmp);
    }
    double t3 = now();
    std::vector<uint32_t> sc(n), si(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) { sc[i] = (uint32_t)(keys[i] >> 32); si[i] = (uint32_t)keys[i]; }
    std::vector<LNode> nodes(2 * n - 1);
    Delta30 d{sc.data(), si.data(), n};
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n - 1; ++i) karras_node(d, n, i, nodes);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        LNode& L = nodes[(n - 1) + i];
        L.box = tb[si[i]];
        L.left = si[i] | 0x80000000u;
        L.right = 0xFFFFFFFFu;
    }
    double t4 = now();
    std::vector<std::atomic<int>> counter(n > 1 ? n - 1 : 1);
    for (auto& c : counter) c.store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        uint32_t idx = (uint32_t)((n - 1) + i);
        while (idx != 0) {
            uint32_t parent = nodes[idx].parent;
            if (counter[parent].fetch_add(1, std::memory_order_acq_rel) == 0) break;
            nodes[parent].box = merge(nodes[nodes[parent].left].box, nodes[nodes[parent].right].box);
            idx = parent;
        }
    }
    double t5 = now();
    lnodes_to_out(nodes, (ONode*)nodes_out);
    std::memcpy(indices, si.data(), n * sizeof(uint32_t));
    if (stage_ms) {
        stage_ms[0] = t1 - t0; stage_ms[1] = t2 - t1; stage_ms[2] = t3 - t2; stage_ms[3] = t4 - t3; stage_ms[4] = t5 - t4;
    }
}


// ---------------------------------------------------------------------------
// OpenMP variants of LBVH+ and PLOC, for the CPU-baseline timing only.  Each
// returns exactly what its serial counterpart returns (tests check it): the
// sorts are stable LSD radix sorts of (code, index) by code, treelet rounds
// follow the reference's climbing schedule on threads (second arrival
// through an atomic counter, disjoint subtrees commute), and PLOC numbers
// its merges by a prefix sum in ascending cluster order, as the serial loop.

}  // extern "C"

// Parallel triangle boxes and scene box.
static Box omp_tri_boxes(const Soa& m, int64_t n, std::vector<Box>& tb) {
    tb.resize(n);
    Box scene = empty_box();
#pragma omp parallel
    {
        Box loc = empty_box();
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) { tb[i] = tri_box(m, i); loc = merge(loc, tb[i]); }
#pragma omp critical
        scene = merge(scene, loc);
    }
    return scene;
}

// Stable LSD radix sort of (code, index) pairs by the low `bits` bits of the
// code, 8-bit digits, one histogram per thread.
template <class K>
static void omp_radix_pairs(std::vector<K>& code, std::vector<uint32_t>& idx, int bits) {
    const int64_t n = (int64_t)code.size();
    std::vector<K> tc(n);
    std::vector<uint32_t> ti(n);
    const int nt = orc_omp_threads();
    std::vector<int64_t> hist((size_t)nt * 256);
    for (int shift = 0; shift < bits; shift += 8) {
        std::fill(hist.begin(), hist.end(), 0);
#pragma omp parallel num_threads(nt)
        {
            int tid = 0;
#ifdef _OPENMP
            tid = omp_get_thread_num();
#endif
            const int64_t lo = n * tid / nt, hi = n * (tid + 1) / nt;
            int64_t* h = &hist[(size_t)tid * 256];
            for (int64_t i = lo; i < hi; ++i) h[(code[i] >> shift) & 255]++;
#pragma omp barrier
#pragma omp single
            {
                int64_t run = 0;
                for (int b = 0; b < 256; ++b)
                    for (int t = 0; t < nt; ++t) {
                        const int64_t c = hist[(size_t)t * 256 + b];
                        hist[(size_t)t * 256 + b] = run;
                        run += c;
                    }
            }
            for (int64_t i = lo; i < hi; ++i) {
                const int64_t d = h[(code[i] >> shift) & 255]++;
                tc[d] = code[i];
                ti[d] = idx[i];
            }
        }
        std::swap(code, tc);
        std::swap(idx, ti);
    }
}

// Exclusive prefix sum of 0/1 flags; returns the total.
static int64_t omp_scan_flags(const std::vector<uint8_t>& f, int64_t n, std::vector<int64_t>& pos) {
    const int nt = orc_omp_threads();
    std::vector<int64_t> part((size_t)nt + 1, 0);
#pragma omp parallel num_threads(nt)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        const int64_t lo = n * tid / nt, hi = n * (tid + 1) / nt;
        int64_t c = 0;
        for (int64_t i = lo; i < hi; ++i) c += f[i];
        part[(size_t)tid + 1] = c;
#pragma omp barrier
#pragma omp single
        for (int t = 0; t < nt; ++t) part[(size_t)t + 1] += part[(size_t)t];
        int64_t run = part[(size_t)tid];
        for (int64_t i = lo; i < hi; ++i) { pos[i] = run; run += f[i]; }
    }
    return part[(size_t)nt];
}

// Treelet rounds on threads: each leaf's thread climbs; the second arrival
// at a node (acquire-release counter) refits it and optimises its treelet.
static void refit_omp(std::vector<LNode>& nodes, int64_t n, bool optimize) {
    std::vector<std::atomic<int>> counter(n > 1 ? n - 1 : 1);
    for (auto& c : counter) c.store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < n; ++i) {
        uint32_t idx = (uint32_t)((n - 1) + i);
        while (idx != 0) {
            const uint32_t parent = nodes[idx].parent;
            if (counter[parent].fetch_add(1, std::memory_order_acq_rel) == 0) break;
            nodes[parent].box = merge(nodes[nodes[parent].left].box, nodes[nodes[parent].right].box);
            if (optimize) optimize_treelet(nodes, (int)parent);
            idx = parent;
        }
    }
}

extern "C" {

void orc_build_lbvh_plus_omp(const float* const* soa, int64_t n, int rounds, void* nodes_out, uint32_t* indices) {
    Soa m;
    for (int k = 0; k < 9; ++k) m.v[k] = soa[k];
    std::vector<Box> tb;
    const Box scene = omp_tri_boxes(m, n, tb);
    std::vector<uint64_t> sc(n);
    std::vector<uint32_t> si(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        float c[3];
        norm_centroid(tb[i], scene, c);
        sc[i] = morton60(c[0], c[1], c[2]);
        si[i] = (uint32_t)i;
    }
    omp_radix_pairs<uint64_t>(sc, si, 60);
    std::vector<LNode> nodes(2 * n - 1);
    Delta60 d{sc.data(), si.data(), n};
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n - 1; ++i) karras_node(d, n, i, nodes);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        LNode& L = nodes[(n - 1) + i];
        L.box = tb[si[i]];
        L.left = si[i] | 0x80000000u;
        L.right = 0xFFFFFFFFu;
    }
    refit_omp(nodes, n, false);
    for (int r = 0; r < rounds; ++r) refit_omp(nodes, n, true);
    lnodes_to_out(nodes, (ONode*)nodes_out);
    std::memcpy(indices, si.data(), n * sizeof(uint32_t));
}

int orc_build_ploc_omp(const float* const* soa, int64_t n, int radius, void* nodes_out, uint32_t* indices,
                       int32_t* iter_m, int32_t max_iters, int32_t* n_iters) {
    Soa m;
    for (int k = 0; k < 9; ++k) m.v[k] = soa[k];
    ONode* out = (ONode*)nodes_out;
    std::vector<Box> tb;
    const Box scene = omp_tri_boxes(m, n, tb);
    std::vector<uint32_t> sc(n), si(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        float c[3];
        norm_centroid(tb[i], scene, c);
        sc[i] = morton30(c[0], c[1], c[2]);
        si[i] = (uint32_t)i;
    }
    omp_radix_pairs<uint32_t>(sc, si, 32);
    std::memcpy(indices, si.data(), n * sizeof(uint32_t));
    std::vector<Box> cb(n), cb2(n);
    std::vector<int32_t> cid(n), cid2(n), nn(n);
    std::vector<uint8_t> lf(n), vf(n);
    std::vector<int64_t> lpos(n), vpos(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        cb[i] = tb[si[i]];
        set_node(out[i], cb[i], (int32_t)(si[i] | 0x80000000u), -1);
        cid[i] = (int32_t)i;
    }
    int64_t cur = n, next = n;
    int32_t iters = 0;
    while (cur > 1) {
        if (iter_m && iters < max_iters) iter_m[iters] = (int32_t)cur;
        iters++;
#pragma omp parallel for schedule(dynamic, 4096)
        for (int64_t i = 0; i < cur; ++i) {
            float best = FLT_MAX;
            int64_t bj = -1;
            const int64_t lo = std::max<int64_t>(0, i - radius), hi = std::min<int64_t>(cur - 1, i + radius);
            for (int64_t j = lo; j <= hi; ++j) {
                if (j == i) continue;
                const float dist = area(merge(cb[i], cb[j]));
                if (dist < best) { best = dist; bj = j; }
                else if (dist == best) { if (bj == -1 || j < bj) bj = j; }
            }
            nn[i] = (int32_t)bj;
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < cur; ++i) {
            const int32_t j = nn[i];
            const bool mutual = (j != -1) && nn[j] == (int32_t)i;
            lf[i] = (mutual && i < j) ? 1 : 0;
            vf[i] = (mutual && j < i) ? 0 : 1;
        }
        const int64_t merges = omp_scan_flags(lf, cur, lpos);
        if (merges == 0) {
            if (n_iters) *n_iters = iters;
            return 5;
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < cur; ++i) {
            if (!lf[i]) continue;
            const int32_t j = nn[i];
            const Box u = merge(cb[i], cb[j]);
            const int64_t id = next + lpos[i];
            set_node(out[id], u, cid[i], cid[j]);
            cid[i] = (int32_t)id;
            cb[i] = u;
        }
        next += merges;
        const int64_t w = omp_scan_flags(vf, cur, vpos);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < cur; ++i)
            if (vf[i]) { cb2[vpos[i]] = cb[i]; cid2[vpos[i]] = cid[i]; }
        std::swap(cb, cb2);
        std::swap(cid, cid2);
        cur = w;
    }
    if (n_iters) *n_iters = iters;
    const int64_t total = next;
    if (total > 1) {
        const int64_t root = total - 1;
        std::swap(out[0], out[root]);
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < total; ++k) {
            ONode& nd = out[k];
            if (nd.left < 0) continue;
            auto fix = [&](int32_t& c) {
                if (!(c & 0x80000000)) {
                    if (c == 0) c = (int32_t)root;
                    else if (c == root) c = 0;
                }
            };
            fix(nd.left);
            fix(nd.right);
        }
    }
    return 0;
}

}  // extern "C"
