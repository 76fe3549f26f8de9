// k_trace.hip — closest-hit ray traversal over the unified BVHNode array.
//
// Replaces traverseBVH / intersectAABB / intersectTriangle of the reference
// renderer (src/render_engine.cu:194-356) for caller-supplied rays.
//
// Semantics are the reference's, exactly: depth-first, left child before
// right, a node is entered only if its slab test passes with
// tenter < (current closest t), Moller-Trumbore with the reference's
// tolerances, closest hit by strict '<'.  Per-ray counters follow the
// reference's definitions (nodes visited, AABB tests, triangle tests).
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
//   * the traversal stack is a short per-lane stack in LDS ([depth][lane],
//     conflict-free), 24 entries of (node, tenter); a ray that would need a
//     deeper stack is not truncated (the reference drops pushes past 128):
//     it is queued and re-traced by k_trace_deep with a 256-entry stack (24
//     entries in LDS, the rest in scratch);
//   * triangles are read from a 48-byte-per-triangle AoS copy of the mesh
//     (three 16-byte loads instead of nine scattered 4-byte loads).
#include "dev_common.hpp"
#include "engine.hpp"

namespace bvhk {

constexpr int kTraceBlock = 64;
constexpr int kShortStack = 24;
constexpr int kDeepStack = 256;

struct RayState {
    float ox, oy, oz, dx, dy, dz, ix, iy, iz;
};

__device__ __forceinline__ float inv_dir(float d) {
    return (fabsf(d) > 1e-6f) ? 1.0f / d : (d >= 0.0f ? 1e30f : -1e30f);
}

// Slab test part of intersectAABB (render_engine.cu:194-214); returns tenter,
// sets hit when texit >= max(tenter, 0).
__device__ __forceinline__ float slab(const RayState& r, const DNode* nodes, int32_t i, bool& hit) {
    const float4 mn = nodes[i].mn;
    const float4 mx = nodes[i].mx;
    float tx1 = (mn.x - r.ox) * r.ix, tx2 = (mx.x - r.ox) * r.ix;
    float ty1 = (mn.y - r.oy) * r.iy, ty2 = (mx.y - r.oy) * r.iy;
    float tz1 = (mn.z - r.oz) * r.iz, tz2 = (mx.z - r.oz) * r.iz;
    float tminx = fmin_s(tx1, tx2), tmaxx = fmax_s(tx1, tx2);
    float tminy = fmin_s(ty1, ty2), tmaxy = fmax_s(ty1, ty2);
    float tminz = fmin_s(tz1, tz2), tmaxz = fmax_s(tz1, tz2);
    float tenter = fmax_s(fmax_s(tminx, tminy), tminz);
    float texit = fmin_s(fmin_s(tmaxx, tmaxy), tmaxz);
    hit = texit >= fmax_s(tenter, 0.0f);
    return tenter;
}

__device__ __forceinline__ float slab_box(const RayState& r, const float4 mn, const float4 mx, bool& hit) {
    float tx1 = (mn.x - r.ox) * r.ix, tx2 = (mx.x - r.ox) * r.ix;
    float ty1 = (mn.y - r.oy) * r.iy, ty2 = (mx.y - r.oy) * r.iy;
    float tz1 = (mn.z - r.oz) * r.iz, tz2 = (mx.z - r.oz) * r.iz;
    float tminx = fmin_s(tx1, tx2), tmaxx = fmax_s(tx1, tx2);
    float tminy = fmin_s(ty1, ty2), tmaxy = fmax_s(ty1, ty2);
    float tminz = fmin_s(tz1, tz2), tmaxz = fmax_s(tz1, tz2);
    float tenter = fmax_s(fmax_s(tminx, tminy), tminz);
    float texit = fmin_s(fmin_s(tmaxx, tmaxy), tmaxz);
    hit = texit >= fmax_s(tenter, 0.0f);
    return tenter;
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

// Stack accessors: LDS short stack or private deep stack.
struct LdsStack {
    int32_t (*idx)[kTraceBlock];
    float (*tent)[kTraceBlock];
    int lane;
    static constexpr int kCap = kShortStack;
    __device__ void put(int sp, int32_t i, float t) { idx[sp][lane] = i; tent[sp][lane] = t; }
    __device__ void get(int sp, int32_t& i, float& t) const { i = idx[sp][lane]; t = tent[sp][lane]; }
};
// Deep stack for rays that outgrew the short stack: the first kShortStack
// entries in LDS, the rest in a private array (scratch memory).
struct DeepStack {
    int32_t (*idx)[kTraceBlock];
    float (*tent)[kTraceBlock];
    int lane;
    int32_t deep_idx[kDeepStack - kShortStack];
    float deep_tent[kDeepStack - kShortStack];
    static constexpr int kCap = kDeepStack;
    __device__ void put(int sp, int32_t i, float t) {
        if (sp < kShortStack) {
            idx[sp][lane] = i;
            tent[sp][lane] = t;
        } else {
            deep_idx[sp - kShortStack] = i;
            deep_tent[sp - kShortStack] = t;
        }
    }
    __device__ void get(int sp, int32_t& i, float& t) const {
        if (sp < kShortStack) {
            i = idx[sp][lane];
            t = tent[sp][lane];
        } else {
            i = deep_idx[sp - kShortStack];
            t = deep_tent[sp - kShortStack];
        }
    }
};

// Returns false when the stack capacity was exceeded (result not written).
template <class Stack>
__device__ __forceinline__ bool traverse(const TraceArgs& a, const RayState& r, Stack& st, float& best,
                                         int32_t& prim, int32_t cnt[3]) {
    const DNode* __restrict__ nodes = a.nodes;
    const int32_t nn = (int32_t)a.num_nodes;
    best = 1e30f;
    prim = -1;
    int32_t visited = 0, aabb = 0, tris = 0;
    int sp = 0;
    int64_t budget = 2 * (int64_t)nn + 2;  // a tree pops each node at most once
    if (nn > 0) {
        bool h;
        float te = slab(r, nodes, 0, h);
        ++visited;
        ++aabb;
        if (h) {
            const int32_t l0 = nodes[0].lr.x;
            st.put(0, l0 < 0 ? l0 : 0, te);
            sp = 1;
        }
    }
    const float4* __restrict__ rec = a.rec;
    while (sp > 0) {
        if (--budget < 0) return false;
        --sp;
        int32_t ref;
        float te;
        st.get(sp, ref, te);
        if (!(te < best)) continue;
        if (ref < 0) {
            ++tris;
            float t;
            const uint32_t p = (uint32_t)ref & 0x7FFFFFFFu;
            if (p < a.num_prims && hit_tri(r, a.tri, p, t) && t < best) {
                best = t;
                prim = (int32_t)p;
            }
            continue;
        }
        const float4* q = rec + 4 * (size_t)ref;
        const float4 A = q[0], B = q[1], C = q[2], D = q[3];
        const int32_t lref = __float_as_int(A.w);
        const int32_t rref = __float_as_int(B.w);
        bool hr = false, hl = false;
        float tr = 0.0f, tl = 0.0f;
        if (rref != kAbsent) {
            tr = slab_box(r, C, D, hr);
            ++visited;
            ++aabb;
        }
        if (lref != kAbsent) {
            tl = slab_box(r, A, B, hl);
            ++visited;
            ++aabb;
        }
        const int need = (hr ? 1 : 0) + (hl ? 1 : 0);
        if (sp + need > Stack::kCap) return false;
        if (hr) st.put(sp++, rref, tr);
        if (hl) st.put(sp++, lref, tl);
    }
    cnt[0] = visited;
    cnt[1] = aabb;
    cnt[2] = tris;
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

__global__ __launch_bounds__(kTraceBlock) void k_trace(TraceArgs a) {
    __shared__ int32_t s_idx[kShortStack][kTraceBlock];
    __shared__ float s_tent[kShortStack][kTraceBlock];
    const uint32_t k = blockIdx.x * kTraceBlock + threadIdx.x;
    if (k >= a.nrays) return;
    const uint32_t i = ray_at(a, k);
    const RayState r = load_ray(a.rays, i);
    LdsStack st{s_idx, s_tent, (int)threadIdx.x};
    float t;
    int32_t p;
    int32_t c[3] = {0, 0, 0};
    if (traverse(a, r, st, t, p, c)) {
        write_result(a, i, t, p, c);
    } else {
        const uint32_t slot = atomicAdd(a.overflow_count, 1u);
        a.overflow_list[slot] = i;
    }
}

// Rays whose stack outgrew the LDS short stack; grid-stride over the queue
// (its length is read on the device, so no host round trip is needed).
__global__ __launch_bounds__(kTraceBlock) void k_trace_deep(TraceArgs a) {
    __shared__ int32_t s_idx[kShortStack][kTraceBlock];
    __shared__ float s_tent[kShortStack][kTraceBlock];
    const uint32_t count = *a.overflow_count;
    for (uint32_t k = blockIdx.x * kTraceBlock + threadIdx.x; k < count; k += gridDim.x * kTraceBlock) {
        const uint32_t i = a.overflow_list[k];
        const RayState r = load_ray(a.rays, i);
        DeepStack st;
        st.idx = s_idx;
        st.tent = s_tent;
        st.lane = (int)threadIdx.x;
        float t;
        int32_t p;
        int32_t c[3] = {0, 0, 0};
        if (traverse(a, r, st, t, p, c)) write_result(a, i, t, p, c);
        else atomicOr(a.overflow_count + 1, 1u);  // deeper than 256 or not a tree: reported as an error
    }
}

// Coherence keys: 9 bits of direction (3 per axis, direction scaled by its
// largest component) above a 30-bit Morton code of the origin in the root box,
// above the ray index; sorting them makes neighbouring lanes trace similar
// rays.  Order only: every ray's result is independent of where it runs.
__global__ void k_ray_keys(const float* __restrict__ rays, uint32_t n, const DNode* __restrict__ nodes,
                           uint32_t idx_bits, uint64_t* __restrict__ keys) {
    const float4 mn = nodes[0].mn, mx = nodes[0].mx;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float* r = rays + 6 * (size_t)i;
        const float ox = r[0], oy = r[1], oz = r[2], dx = r[3], dy = r[4], dz = r[5];
        const float ex = mx.x - mn.x, ey = mx.y - mn.y, ez = mx.z - mn.z;
        float x = (ox - mn.x) / (ex > 1e-6f ? ex : 1.0f);
        float y = (oy - mn.y) / (ey > 1e-6f ? ey : 1.0f);
        float z = (oz - mn.z) / (ez > 1e-6f ? ez : 1.0f);
        x = fmin_s(fmax_s(x, 0.0f), 1.0f);
        y = fmin_s(fmax_s(y, 0.0f), 1.0f);
        z = fmin_s(fmax_s(z, 0.0f), 1.0f);
        const uint64_t m = morton30(x, y, z);
        float dm = fmax_s(fabsf(dx), fmax_s(fabsf(dy), fabsf(dz)));
        dm = dm > 0.0f ? dm : 1.0f;
        const uint32_t qx = (uint32_t)fmin_s(fmax_s((dx / dm + 1.0f) * 4.0f, 0.0f), 7.0f);
        const uint32_t qy = (uint32_t)fmin_s(fmax_s((dy / dm + 1.0f) * 4.0f, 0.0f), 7.0f);
        const uint32_t qz = (uint32_t)fmin_s(fmax_s((dz / dm + 1.0f) * 4.0f, 0.0f), 7.0f);
        const uint64_t dk = (uint64_t)((qx << 6) | (qy << 3) | qz);
        keys[i] = (((dk << 30) | m) << idx_bits) | i;
    }
}

void launch_ray_keys(const float* rays, uint32_t n, const DNode* nodes, uint32_t idx_bits, uint64_t* keys,
                     hipStream_t s) {
    uint32_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_ray_keys, dim3(g ? g : 1), dim3(256), 0, s, rays, n, nodes, idx_bits, keys);
}

// Mesh SoA -> 3 x float4 per triangle (v0, v1, v2; w = 0).
__global__ void k_tri_aos(MeshView m, float4* __restrict__ tri) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m.n; i += stride) {
        tri[3 * (size_t)i + 0] = make_float4(m.v[0][i], m.v[1][i], m.v[2][i], 0.0f);
        tri[3 * (size_t)i + 1] = make_float4(m.v[3][i], m.v[4][i], m.v[5][i], 0.0f);
        tri[3 * (size_t)i + 2] = make_float4(m.v[6][i], m.v[7][i], m.v[8][i], 0.0f);
    }
}

void launch_tri_aos(const MeshView& m, float4* tri, hipStream_t s) {
    uint32_t g = (m.n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_tri_aos, dim3(g ? g : 1), dim3(256), 0, s, m, tri);
}

void launch_trace(const TraceArgs& a, hipStream_t s) {
    uint32_t g = (a.nrays + kTraceBlock - 1) / kTraceBlock;
    hipLaunchKernelGGL(k_trace, dim3(g ? g : 1), dim3(kTraceBlock), 0, s, a);
}

void launch_trace_deep(const TraceArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_trace_deep, dim3(1024), dim3(kTraceBlock), 0, s, a);
}

}  // namespace bvhk
