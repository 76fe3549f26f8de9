// k_trace.hip — closest-hit ray traversal over the unified BVHNode array.
//
// Replaces traverseBVH / intersectAABB / intersectTriangle of the reference
// renderer (src/render_engine.cu:194-356) for caller-supplied rays.
//
// Two visit orders (a template switch, chosen per context):
//   * the reference's, exactly: depth-first, left child before right, a node
//     is entered only if its slab test passes with tenter < (current closest
//     t), Moller-Trumbore with the reference's tolerances, closest hit by
//     strict '<'.  Per-ray counters follow the reference's definitions
//     (nodes visited, AABB tests, triangle tests);
//   * nearest-first (no reference counterpart): the nearer of two children
//     first, a node entered while tenter <= closest t, the closest hit the
//     least (t, triangle index) pair; same predicates and counter rules.
//
// MI355X mapping:
//   * children's slab tests run when their parent is expanded and only
//     children whose slab interval is non-empty are pushed, with their entry
//     distance; the reference's pop-time test then reduces to tenter < t.
//     The visit order and every decision are unchanged;
//   * the traversal reads a 64-byte record per internal node (both
//     children's boxes and references, one cache-line-aligned load) built
//     from the node array once per tree (k_trace_layout), so expanding a
//     node costs one memory round trip instead of two dependent ones (node,
//     then its children's boxes); leaf children carry their primitive;
//   * the left child stays in registers and only right children are
//     pushed; the stack holds (node, tenter) packed in 64 bits: 16 entries
//     per lane in LDS ([depth][lane], conflict-free), the next 32 in the
//     lane's own spill slot in HBM; a ray that would need a deeper stack is
//     not truncated (the reference drops pushes past 128): it is queued and
//     re-traced by k_trace_deep with a 256-entry stack (64 entries in LDS,
//     the rest in scratch);
//   * k_trace is persistent (one counter hands out 64-ray batches), so spill
//     slots belong to resident lanes, not to rays;
//   * the slab tests use the hardware min/max (exact for finite inputs, see
//     slab_fast);
//   * triangles are read from a 48-byte-per-triangle AoS copy of the mesh
//     (three 16-byte loads instead of nine scattered 4-byte loads).
#include "dev_common.hpp"
#include "engine.hpp"

namespace bvhk {

constexpr int kTraceBlock = 64;
// Stack entries of a lane kept in LDS: 16 (8 KB per 64-lane block) measured
// fastest against 12, 20 and 24 (fewer LDS bytes per ray let more rays be in
// flight; fewer entries spill more often).
constexpr int kShortStack = 16;
constexpr int kSpill = 32;        // entries of a lane's HBM spill slot beyond LDS
constexpr int kMaxStack = 256;    // all stack entries of a lane (k_trace_deep)
constexpr int kDeepLds = 64;      // LDS entries of k_trace_deep (32 KB per 64-lane block)

struct RayState {
    float ox, oy, oz, dx, dy, dz, ix, iy, iz;
};

__device__ __forceinline__ float inv_dir(float d) {
    return (fabsf(d) > 1e-6f) ? 1.0f / d : (d >= 0.0f ? 1e30f : -1e30f);
}

// Child reference in a traversal record: node index (>= 0), leaf
// (0x80000000 | primitive) or absent (the reference skips out-of-range
// children without counting them).
constexpr int32_t kAbsent = 0x7FFFFFFF;

__device__ __forceinline__ int32_t child_ref(const DNode* nodes, int32_t c, int32_t nn) {
    if (c < 0 || c >= nn) return kAbsent;
    const int32_t l = nodes[c].lr.x;
    return l < 0 ? l : c;
}

// Record i (internal node i): {left min, left ref}, {left max, right ref},
// {right min, 0}, {right max, 0}.
__global__ void k_trace_layout(const DNode* __restrict__ nodes, int32_t nn, float4* __restrict__ rec) {
    const int32_t stride = gridDim.x * blockDim.x;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += stride) {
        const int4 lr = nodes[i].lr;
        if (lr.x < 0) continue;  // leaf: no record
        const int32_t lref = child_ref(nodes, lr.x, nn);
        const int32_t rref = child_ref(nodes, lr.y, nn);
        float4 lmn = make_float4(0.f, 0.f, 0.f, 0.f), lmx = lmn, rmn = lmn, rmx = lmn;
        if (lref != kAbsent) {
            lmn = nodes[lr.x].mn;
            lmx = nodes[lr.x].mx;
        }
        if (rref != kAbsent) {
            rmn = nodes[lr.y].mn;
            rmx = nodes[lr.y].mx;
        }
        float4* o = rec + 4 * (size_t)i;
        o[0] = make_float4(lmn.x, lmn.y, lmn.z, __int_as_float(lref));
        o[1] = make_float4(lmx.x, lmx.y, lmx.z, __int_as_float(rref));
        o[2] = make_float4(rmn.x, rmn.y, rmn.z, 0.f);
        o[3] = make_float4(rmx.x, rmx.y, rmx.z, 0.f);
    }
}

void launch_trace_layout(const DNode* nodes, uint32_t num_nodes, float4* rec, hipStream_t s) {
    uint32_t g = (num_nodes + 255) / 256;
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(k_trace_layout, dim3(g ? g : 1), dim3(256), 0, s, nodes, (int32_t)num_nodes, rec);
}

// intersectTriangle (render_engine.cu:220-273); no contraction.
__device__ __forceinline__ bool hit_tri(const RayState& r, const float4* tri, uint32_t p, float& out_t) {
    const float4 a = tri[3 * (size_t)p + 0];
    const float4 b = tri[3 * (size_t)p + 1];
    const float4 c = tri[3 * (size_t)p + 2];
    float e1x = b.x - a.x, e1y = b.y - a.y, e1z = b.z - a.z;
    float e2x = c.x - a.x, e2y = c.y - a.y, e2z = c.z - a.z;
    float px = r.dy * e2z - r.dz * e2y;
    float py = r.dz * e2x - r.dx * e2z;
    float pz = r.dx * e2y - r.dy * e2x;
    float det = e1x * px + e1y * py + e1z * pz;
    if (fabsf(det) < 1e-6f) return false;
    float inv_det = 1.0f / det;
    float tx = r.ox - a.x, ty = r.oy - a.y, tz = r.oz - a.z;
    float u = (tx * px + ty * py + tz * pz) * inv_det;
    if (u < -1e-5f || u > 1.00001f) return false;
    float qx = ty * e1z - tz * e1y;
    float qy = tz * e1x - tx * e1z;
    float qz = tx * e1y - ty * e1x;
    float v = (r.dx * qx + r.dy * qy + r.dz * qz) * inv_det;
    if (v < -1e-5f || u + v > 1.00001f) return false;
    float t = (e2x * qx + e2y * qy + e2z * qz) * inv_det;
    if (t < 1e-6f) return false;
    out_t = t;
    return true;
}

__device__ __forceinline__ RayState load_ray(const float* rays, uint32_t i) {
    const float2* r2 = reinterpret_cast<const float2*>(rays + 6 * (size_t)i);
    float2 a = r2[0], b = r2[1], c = r2[2];
    RayState r;
    r.ox = a.x; r.oy = a.y; r.oz = b.x;
    r.dx = b.y; r.dy = c.x; r.dz = c.y;
    r.ix = inv_dir(r.dx);
    r.iy = inv_dir(r.dy);
    r.iz = inv_dir(r.dz);
    return r;
}

// Stack entries are (node reference, entry distance) packed in 64 bits.
//
// LaneStack (k_trace): the first LDS entries of a lane in LDS ([depth][lane]
// rows of 8-byte entries, conflict-free), the next kSpill in the lane's own
// spill slot in HBM (k_trace is persistent, so a slot per resident lane
// serves every ray that lane traces; only deep rays touch it); a ray that
// needs more is queued for k_trace_deep.  DeepStack (k_trace_deep): the
// first kDeepLds entries in LDS, the rest in a private array (scratch).
template <int LDS>
struct LaneStack {
    int2 (*e)[kTraceBlock];
    int lane;
    int2* slot;  // kSpill entries
    static constexpr int kCap = LDS + kSpill;
    __device__ bool put(int sp, int32_t i, float t) {
        const int2 v = make_int2(i, __float_as_int(t));
        if (sp < LDS) e[sp][lane] = v;
        else slot[sp - LDS] = v;
        return true;
    }
    __device__ void get(int sp, int32_t& i, float& t) const {
        // Two loads of different types, so the compiler keeps an LDS read
        // and a global read on their own paths instead of merging them into
        // one generic (flat) load through a selected address.
        int2 v;
        if (sp < LDS) {
            v = e[sp][lane];
        } else {
            const uint64_t w = *reinterpret_cast<const uint64_t*>(slot + (sp - LDS));
            v = make_int2((int32_t)(uint32_t)w, (int32_t)(uint32_t)(w >> 32));
        }
        i = v.x;
        t = __int_as_float(v.y);
    }
};

struct DeepStack {
    int2 (*e)[kTraceBlock];
    int lane;
    int2 deep[kMaxStack - kDeepLds];
    static constexpr int kCap = kMaxStack;
    __device__ bool put(int sp, int32_t i, float t) {
        const int2 v = make_int2(i, __float_as_int(t));
        if (sp < kDeepLds) e[sp][lane] = v;
        else deep[sp - kDeepLds] = v;
        return true;
    }
    __device__ void get(int sp, int32_t& i, float& t) const {
        const int2 v = sp < kDeepLds ? e[sp][lane] : deep[sp - kDeepLds];
        i = v.x;
        t = __int_as_float(v.y);
    }
};

// Slab test with the hardware min/max: rays and boxes are finite (checked on
// input; an inverse direction component is at most 1e30 in magnitude), so
// no operand is NaN, and min/max may order -0 and +0 either way: no
// comparison below can tell the two apart, so every decision equals the
// reference's (its intersectAABB, render_engine.cu:194-214).
__device__ __forceinline__ float slab_fast(const RayState& r, const float4 mn, const float4 mx, bool& hit) {
    const float tx1 = (mn.x - r.ox) * r.ix, tx2 = (mx.x - r.ox) * r.ix;
    const float ty1 = (mn.y - r.oy) * r.iy, ty2 = (mx.y - r.oy) * r.iy;
    const float tz1 = (mn.z - r.oz) * r.iz, tz2 = (mx.z - r.oz) * r.iz;
    const float tenter = fmaxf(fmaxf(fminf(tx1, tx2), fminf(ty1, ty2)), fminf(tz1, tz2));
    const float texit = fminf(fminf(fmaxf(tx1, tx2), fmaxf(ty1, ty2)), fmaxf(tz1, tz2));
    hit = texit >= fmaxf(tenter, 0.0f);
    return tenter;
}

// One ray's traversal, advanced one stack pop at a time.
//
// Reference order (NEAREST = false), the reference's depth-first order:
// children pushed right then left, so the left one is popped next; a popped
// entry is skipped unless its entry distance is below the current closest t;
// a hit replaces the closest one when its t is strictly smaller.  The child
// popped next is kept in registers (`cur`) and only the other one goes
// through the stack: the same decisions in the same order.  Children are
// counted when their parent is expanded: the reference pops (and counts)
// every child it pushes.  A child whose entry distance is already not below
// the closest t is not pushed: the closest t only falls, so its pop would be
// skipped anyway.  (A "while-while" split into internal and leaf phases was
// measured 2.6x slower: without postponing leaves, which would change the
// culls, each phase waits for the wave's slowest lane.)
//
// Nearest order (NEAREST = true): of two children that pass, the one with
// the smaller entry distance is taken first (the left one on a tie); an
// entry is skipped when its entry distance exceeds the closest t; the
// closest hit is the least (t, triangle index) pair, so it does not depend
// on the order triangles are met in.  Counted as the reference order
// counts.
template <bool NEAREST>
struct Walker {
    RayState r;
    float best;
    int32_t prim, visited, tris;
    int sp;
    bool have;       // cur holds the next entry
    int32_t cur;
    float cur_te;
    int64_t budget;  // a tree pops each node at most once

    __device__ __forceinline__ bool keep(float te) const { return NEAREST ? te <= best : te < best; }

    __device__ void start(const TraceArgs& a, const RayState& ray) {
        r = ray;
        best = 1e30f;
        prim = -1;
        visited = 0;
        tris = 0;
        sp = 0;
        have = false;
        cur = 0;
        cur_te = 0.0f;
        budget = 2 * (int64_t)a.num_nodes + 2;
        if (a.num_nodes > 0) {
            bool h;
            cur_te = slab_fast(r, a.nodes[0].mn, a.nodes[0].mx, h);
            visited = 1;
            const int32_t l0 = a.nodes[0].lr.x;
            cur = l0 < 0 ? l0 : 0;
            have = h;
        }
    }
    __device__ bool done() const { return !have && sp == 0; }

    // One pop; false when the stack capacity or the pop budget ran out.
    template <class Stack>
    __device__ __forceinline__ bool step(const TraceArgs& a, Stack& st) {
        if (--budget < 0) return false;
        int32_t ref;
        float te;
        if (have) {
            ref = cur;
            te = cur_te;
            have = false;
        } else {
            --sp;
            st.get(sp, ref, te);
        }
        if (!keep(te)) return true;
        if (ref < 0) {
            ++tris;
            float t;
            const uint32_t p = (uint32_t)ref & 0x7FFFFFFFu;
            if (p < a.num_prims && hit_tri(r, a.tri, p, t) &&
                (t < best || (NEAREST && t == best && (int32_t)p < prim))) {
                best = t;
                prim = (int32_t)p;
            }
            return true;
        }
        const float4* q = a.rec + 4 * (size_t)ref;
        const float4 A = q[0], B = q[1], C = q[2], D = q[3];
        const int32_t lref = __float_as_int(A.w);
        const int32_t rref = __float_as_int(B.w);
        const bool lp = lref != kAbsent, rp = rref != kAbsent;
        bool hl, hr;
        const float tl = slab_fast(r, A, B, hl);
        const float tr = slab_fast(r, C, D, hr);
        visited += (lp ? 1 : 0) + (rp ? 1 : 0);
        hl = hl && lp && keep(tl);
        hr = hr && rp && keep(tr);
        if (hl && hr) {
            const bool sw = NEAREST && tr < tl;  // right child first
            if (sp >= Stack::kCap || !st.put(sp, sw ? lref : rref, sw ? tl : tr)) return false;
            ++sp;
            cur = sw ? rref : lref;
            cur_te = sw ? tr : tl;
        } else {
            cur = hl ? lref : rref;
            cur_te = hl ? tl : tr;
        }
        have = hl || hr;
        return true;
    }
};

// Whole traversal of one ray; false when the stack capacity was exceeded.
template <bool NEAREST, class Stack>
__device__ __forceinline__ bool traverse(const TraceArgs& a, const RayState& r, Stack& st, float& best,
                                         int32_t& prim, int32_t cnt[3]) {
    Walker<NEAREST> w;
    w.start(a, r);
    while (!w.done())
        if (!w.step(a, st)) return false;
    best = w.best;
    prim = w.prim;
    cnt[0] = w.visited;
    cnt[1] = w.visited;
    cnt[2] = w.tris;
    return true;
}

__device__ __forceinline__ void write_result(const TraceArgs& a, uint32_t i, float t, int32_t p, const int32_t c[3]) {
    a.t_out[i] = t;
    a.prim_out[i] = p;
    if (a.counters) {
        a.counters[3 * (size_t)i + 0] = c[0];
        a.counters[3 * (size_t)i + 1] = c[1];
        a.counters[3 * (size_t)i + 2] = c[2];
    }
}

// Ray processed by slot k: the k-th ray of the coherence order (sorted keys
// carry the ray index in their low bits), or ray k.
__device__ __forceinline__ uint32_t ray_at(const TraceArgs& a, uint32_t k) {
    return a.order ? (uint32_t)(a.order[k] & a.order_mask) : k;
}

// Persistent: each one-wave block takes batches of kTraceBlock consecutive
// rays (in coherence order) from one counter until none are left, so the
// rays of a batch are traced side by side and waves stay busy to the end.
// Measured and not kept: refilling idle lanes one by one from the counter
// (a wave then traces a sliding window of rays) and splitting the rays into
// one contiguous part per XCD (each XCD's L2 then serves one region); both
// were slower.
template <int LDS, bool NEAREST>
__global__ __launch_bounds__(kTraceBlock) void k_trace(TraceArgs a) {
    __shared__ int2 s_stack[LDS][kTraceBlock];
    LaneStack<LDS> st;
    st.e = s_stack;
    st.lane = (int)threadIdx.x;
    st.slot = a.spill + ((size_t)blockIdx.x * kTraceBlock + threadIdx.x) * kSpill;
    while (true) {
        uint32_t b = 0;
        if (threadIdx.x == 0) b = atomicAdd(a.overflow_count + 2, 1u);
        b = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl(b, 0));
        if ((uint64_t)b * kTraceBlock >= a.nrays) break;
        const uint64_t k = (uint64_t)b * kTraceBlock + threadIdx.x;
        if (k < a.nrays) {
            const uint32_t i = ray_at(a, (uint32_t)k);
            const RayState r = load_ray(a.rays, i);
            float t;
            int32_t p;
            int32_t c[3] = {0, 0, 0};
            if (traverse<NEAREST>(a, r, st, t, p, c)) {
                write_result(a, i, t, p, c);
            } else {
                const uint32_t q = atomicAdd(a.overflow_count, 1u);
                a.overflow_list[q] = i;
            }
        }
    }
}

// Rays whose stack outgrew LDS and the spill slot; grid-stride over
// the queue (its length is read on the device, so no host round trip).
template <bool NEAREST>
__global__ __launch_bounds__(kTraceBlock) void k_trace_deep(TraceArgs a) {
    __shared__ int2 s_stack[kDeepLds][kTraceBlock];
    const uint32_t count = *a.overflow_count;
    for (uint32_t k = blockIdx.x * kTraceBlock + threadIdx.x; k < count; k += gridDim.x * kTraceBlock) {
        const uint32_t i = a.overflow_list[k];
        const RayState r = load_ray(a.rays, i);
        DeepStack st;
        st.e = s_stack;
        st.lane = (int)threadIdx.x;
        float t;
        int32_t p;
        int32_t c[3] = {0, 0, 0};
        if (traverse<NEAREST>(a, r, st, t, p, c)) write_result(a, i, t, p, c);
        else atomicOr(a.overflow_count + 1, 1u);  // deeper than kMaxStack or not a tree: reported as an error
    }
}

// Coherence keys: a 40-bit Morton code of the ray in five dimensions (origin
// in the root box, 8 bits per axis; direction in octahedral coordinates, 8
// bits per axis) above the ray index, so that sorting them puts rays with
// nearby origins and directions in the same wave (camera rays, which share
// their origin, come out in pixel tiles).  Order only: every ray's result is
// independent of where it runs.
__device__ __forceinline__ uint64_t spread5(uint32_t v) {  // 8 bits -> every fifth bit of 40
    uint64_t r = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b) r |= (uint64_t)((v >> b) & 1u) << (5 * b);
    return r;
}

__device__ __forceinline__ uint32_t quant8(float x) {  // [0, 1] -> 0..255
    return (uint32_t)fmin_s(fmax_s(x * 256.0f, 0.0f), 255.0f);
}

__global__ void k_ray_keys(const float* __restrict__ rays, uint32_t n, const DNode* __restrict__ nodes,
                           uint32_t idx_bits, uint64_t* __restrict__ keys) {
    const float4 mn = nodes[0].mn, mx = nodes[0].mx;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float* r = rays + 6 * (size_t)i;
        const float ox = r[0], oy = r[1], oz = r[2], dx = r[3], dy = r[4], dz = r[5];
        const float ex = mx.x - mn.x, ey = mx.y - mn.y, ez = mx.z - mn.z;
        const float x = (ox - mn.x) / (ex > 1e-6f ? ex : 1.0f);
        const float y = (oy - mn.y) / (ey > 1e-6f ? ey : 1.0f);
        const float z = (oz - mn.z) / (ez > 1e-6f ? ez : 1.0f);
        // Octahedral map of the direction onto [-1, 1]^2, then [0, 1]^2.
        const float l1 = fabsf(dx) + fabsf(dy) + fabsf(dz);
        const float inv = l1 > 0.0f ? 1.0f / l1 : 0.0f;
        float u = dx * inv, v = dy * inv;
        if (dz * inv < 0.0f) {
            const float uu = (1.0f - fabsf(v)) * (u >= 0.0f ? 1.0f : -1.0f);
            const float vv = (1.0f - fabsf(u)) * (v >= 0.0f ? 1.0f : -1.0f);
            u = uu;
            v = vv;
        }
        const uint64_t m = (spread5(quant8(x)) << 4) | (spread5(quant8(y)) << 3) | (spread5(quant8(z)) << 2) |
                           (spread5(quant8(u * 0.5f + 0.5f)) << 1) | spread5(quant8(v * 0.5f + 0.5f));
        keys[i] = (m << idx_bits) | i;
    }
}

void launch_ray_keys(const float* rays, uint32_t n, const DNode* nodes, uint32_t idx_bits, uint64_t* keys,
                     hipStream_t s) {
    uint32_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_ray_keys, dim3(g ? g : 1), dim3(256), 0, s, rays, n, nodes, idx_bits, keys);
}

// Mesh SoA -> 3 x float4 per triangle (v0, v1, v2; w = 0).  *bad (cleared
// by the caller) becomes nonzero when any coordinate is NaN or infinite.
__global__ void k_tri_aos(MeshView m, float4* __restrict__ tri, uint32_t* bad) {
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t any = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m.n; i += stride) {
        float v[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            v[k] = m.v[k][i];
            any |= (__float_as_uint(v[k]) & 0x7f800000u) == 0x7f800000u ? 1u : 0u;
        }
        tri[3 * (size_t)i + 0] = make_float4(v[0], v[1], v[2], 0.0f);
        tri[3 * (size_t)i + 1] = make_float4(v[3], v[4], v[5], 0.0f);
        tri[3 * (size_t)i + 2] = make_float4(v[6], v[7], v[8], 0.0f);
    }
    if (__ballot(any != 0) != 0 && (threadIdx.x & (kWave - 1)) == 0) atomicOr(bad, 1u);
}

void launch_tri_aos(const MeshView& m, float4* tri, uint32_t* bad, hipStream_t s) {
    uint32_t g = (m.n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_tri_aos, dim3(g ? g : 1), dim3(256), 0, s, m, tri, bad);
}

template <int LDS, bool NEAREST>
uint32_t resident_blocks() {
    static const uint32_t nb = [] {
        int dev = 0, cus = 0, per = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_trace<LDS, NEAREST>, kTraceBlock, 0);
        const uint32_t v = (uint32_t)(cus > 0 ? cus : 256) * (uint32_t)(per > 0 ? per : 8);
        return v;
    }();
    return nb;
}

int trace_spill_entries() { return kSpill; }

// One slot per lane of the grid launch_trace uses for n rays (either
// order): at most one 64-lane block per 64 rays, at most the resident blocks.
uint32_t trace_spill_slots(uint32_t nrays) {
    const uint32_t batches = (nrays + kTraceBlock - 1) / kTraceBlock;
    const uint32_t g0 = resident_blocks<kShortStack, false>(), g1 = resident_blocks<kShortStack, true>();
    const uint32_t g = g0 > g1 ? g0 : g1;
    return (batches < g ? batches : g) * kTraceBlock;
}

template <int LDS, bool NEAREST>
void launch_trace_t(const TraceArgs& a, hipStream_t s) {
    const uint32_t batches = (a.nrays + kTraceBlock - 1) / kTraceBlock;
    uint32_t g = resident_blocks<LDS, NEAREST>();
    if (g > batches) g = batches;
    if (g > a.spill_slots / kTraceBlock) g = a.spill_slots / kTraceBlock;
    hipLaunchKernelGGL((k_trace<LDS, NEAREST>), dim3(g ? g : 1), dim3(kTraceBlock), 0, s, a);
}

void launch_trace(const TraceArgs& a, hipStream_t s) {
    if (a.nearest) launch_trace_t<kShortStack, true>(a, s);
    else launch_trace_t<kShortStack, false>(a, s);
}

void launch_trace_deep(const TraceArgs& a, hipStream_t s) {
    if (a.nearest) hipLaunchKernelGGL(k_trace_deep<true>, dim3(1024), dim3(kTraceBlock), 0, s, a);
    else hipLaunchKernelGGL(k_trace_deep<false>, dim3(1024), dim3(kTraceBlock), 0, s, a);
}

}  // namespace bvhk
