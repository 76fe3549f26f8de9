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
//   * children's slab tests run when their parent is expanded (their boxes
//     are read together, 2 x 32 B) and only children whose slab interval is
//     non-empty are pushed, with their entry distance; the reference's
//     pop-time test then reduces to tenter < t.  The visit order and every
//     decision are unchanged;
//   * the traversal stack is a short per-lane stack in LDS ([depth][lane],
//     conflict-free), 24 entries of (node, tenter); a ray that would need a
//     deeper stack is not truncated (the reference drops pushes past 128):
//     it is queued and re-traced by k_trace_deep with a 256-entry stack;
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
struct DeepStack {
    int32_t idx[kDeepStack];
    float tent[kDeepStack];
    static constexpr int kCap = kDeepStack;
    __device__ void put(int sp, int32_t i, float t) { idx[sp] = i; tent[sp] = t; }
    __device__ void get(int sp, int32_t& i, float& t) const { i = idx[sp]; t = tent[sp]; }
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
            st.put(0, 0, te);
            sp = 1;
        }
    }
    while (sp > 0) {
        if (--budget < 0) return false;
        --sp;
        int32_t idx;
        float te;
        st.get(sp, idx, te);
        if (!(te < best)) continue;
        const int4 lr = nodes[idx].lr;
        if (lr.x < 0) {
            ++tris;
            float t;
            const uint32_t p = (uint32_t)lr.x & 0x7FFFFFFFu;
            if (p < a.num_prims && hit_tri(r, a.tri, p, t) && t < best) {
                best = t;
                prim = (int32_t)p;
            }
            continue;
        }
        const bool vr = lr.y >= 0 && lr.y < nn;
        const bool vl = lr.x >= 0 && lr.x < nn;
        bool hr = false, hl = false;
        float tr = 0.0f, tl = 0.0f;
        if (vr) {
            tr = slab(r, nodes, lr.y, hr);
            ++visited;
            ++aabb;
        }
        if (vl) {
            tl = slab(r, nodes, lr.x, hl);
            ++visited;
            ++aabb;
        }
        const int need = (hr ? 1 : 0) + (hl ? 1 : 0);
        if (sp + need > Stack::kCap) return false;
        if (hr) st.put(sp++, lr.y, tr);
        if (hl) st.put(sp++, lr.x, tl);
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

__global__ __launch_bounds__(kTraceBlock) void k_trace(TraceArgs a) {
    __shared__ int32_t s_idx[kShortStack][kTraceBlock];
    __shared__ float s_tent[kShortStack][kTraceBlock];
    const uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= a.nrays) return;
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
    const uint32_t count = *a.overflow_count;
    for (uint32_t k = blockIdx.x * kTraceBlock + threadIdx.x; k < count; k += gridDim.x * kTraceBlock) {
        const uint32_t i = a.overflow_list[k];
        const RayState r = load_ray(a.rays, i);
        DeepStack st;
        float t;
        int32_t p;
        int32_t c[3] = {0, 0, 0};
        if (traverse(a, r, st, t, p, c)) write_result(a, i, t, p, c);
        else atomicOr(a.overflow_count + 1, 1u);  // deeper than 256 or not a tree: reported as an error
    }
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
    hipLaunchKernelGGL(k_trace_deep, dim3(256), dim3(kTraceBlock), 0, s, a);
}

}  // namespace bvhk
