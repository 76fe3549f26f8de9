// dev_common.hpp — device helpers shared by the gfx950 kernels.
//
// Boxes travel as two float4 (min xyz + pad, max xyz + pad), the byte layout of
// AABB_cw, so a box is one 32-byte aligned pair of 16-byte loads or stores.
// All min/max are selects (b < a ? b : a), matching the host-side definition
// in include/common.h, so every box is bit-identical to the CPU oracle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bvhk {

constexpr int kWave = 64;

__device__ __forceinline__ float fmin_s(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float fmax_s(float a, float b) { return (a < b) ? b : a; }

struct Box {
    float mnx, mny, mnz;
    float mxx, mxy, mxz;
};

__device__ __forceinline__ Box empty_box() {
    Box b;
    b.mnx = b.mny = b.mnz = 3.402823466e+38f;
    b.mxx = b.mxy = b.mxz = -3.402823466e+38f;
    return b;
}

// AABB_cw::merge (reference include/common.h:46-64) as selects.
__device__ __forceinline__ Box merge(const Box& a, const Box& b) {
    Box r;
    r.mnx = fmin_s(a.mnx, b.mnx);
    r.mny = fmin_s(a.mny, b.mny);
    r.mnz = fmin_s(a.mnz, b.mnz);
    r.mxx = fmax_s(a.mxx, b.mxx);
    r.mxy = fmax_s(a.mxy, b.mxy);
    r.mxz = fmax_s(a.mxz, b.mxz);
    return r;
}

// AABB_cw::surfaceArea (reference include/common.h:41-44), source order, no FMA
// (the file is compiled with -ffp-contract=off).
__device__ __forceinline__ float area(const Box& b) {
    float dx = b.mxx - b.mnx;
    float dy = b.mxy - b.mny;
    float dz = b.mxz - b.mnz;
    float t1 = dx * dy;
    float t2 = dy * dz;
    float t3 = dz * dx;
    float s = t1 + t2;
    s = s + t3;
    return 2.0f * s;
}

__device__ __forceinline__ Box load_box(const float4* p) {
    float4 a = p[0];
    float4 c = p[1];
    Box b;
    b.mnx = a.x; b.mny = a.y; b.mnz = a.z;
    b.mxx = c.x; b.mxy = c.y; b.mxz = c.z;
    return b;
}

__device__ __forceinline__ void store_box(float4* p, const Box& b) {
    p[0] = make_float4(b.mnx, b.mny, b.mnz, 0.0f);
    p[1] = make_float4(b.mxx, b.mxy, b.mxz, 0.0f);
}

// Device image of BVHNode (48 B): box (32 B), left, right, 8 B padding.
struct __align__(16) DNode {
    float4 mn;
    float4 mx;
    int4 lr;  // x = leftChild, y = rightChild, z = w = 0
};
static_assert(sizeof(DNode) == 48, "DNode must match BVHNode");

__device__ __forceinline__ void store_node(DNode* nodes, uint32_t i, const Box& b, int32_t l, int32_t r) {
    DNode* d = nodes + i;
    d->mn = make_float4(b.mnx, b.mny, b.mnz, 0.0f);
    d->mx = make_float4(b.mxx, b.mxy, b.mxz, 0.0f);
    d->lr = make_int4(l, r, 0, 0);
}

__device__ __forceinline__ Box node_box(const DNode* nodes, uint32_t i) {
    return load_box(reinterpret_cast<const float4*>(nodes + i));
}

// Node record written with agent-scope atomic stores (coherent across XCDs
// without flushing the L2), and the matching loads.
// 64-bit halves of a record: agent-scope relaxed atomic stores of 8 bytes
// each (global_store_dwordx2 sc1), half the instructions of 32-bit ones.
__device__ __forceinline__ uint64_t pack2f(float a, float b) {
    return (uint64_t)__float_as_uint(a) | ((uint64_t)__float_as_uint(b) << 32);
}
__device__ __forceinline__ void store_u64_coherent(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// PAD = false leaves the record's last 8 (padding) bytes alone: for records
// that already exist with zero padding (treelet restructuring reassigns
// existing internal nodes).
template <bool PAD = true>
__device__ __forceinline__ void store_node_coherent(DNode* nodes, uint32_t i, const Box& b, int32_t l, int32_t r) {
    uint64_t* q = reinterpret_cast<uint64_t*>(nodes + i);
    store_u64_coherent(q + 0, pack2f(b.mnx, b.mny));
    store_u64_coherent(q + 1, pack2f(b.mnz, 0.0f));
    store_u64_coherent(q + 2, pack2f(b.mxx, b.mxy));
    store_u64_coherent(q + 3, pack2f(b.mxz, 0.0f));
    store_u64_coherent(q + 4, (uint64_t)(uint32_t)l | ((uint64_t)(uint32_t)r << 32));
    if (PAD) q[5] = 0;
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

// Agent-coherent loads and stores of whole records (64-bit atomic accesses):
// data shared between workgroups of one launch that may sit on different XCDs,
// without whole-L2 writebacks or invalidates.
__device__ __forceinline__ DNode load_node_coherent(const DNode* nodes, uint32_t i) {
    const uint64_t* q = reinterpret_cast<const uint64_t*>(nodes + i);
    uint64_t w[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) w[k] = __hip_atomic_load(q + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    DNode d;
    d.mn = make_float4(__uint_as_float((uint32_t)w[0]), __uint_as_float((uint32_t)(w[0] >> 32)),
                       __uint_as_float((uint32_t)w[1]), __uint_as_float((uint32_t)(w[1] >> 32)));
    d.mx = make_float4(__uint_as_float((uint32_t)w[2]), __uint_as_float((uint32_t)(w[2] >> 32)),
                       __uint_as_float((uint32_t)w[3]), __uint_as_float((uint32_t)(w[3] >> 32)));
    d.lr = make_int4((int)(uint32_t)w[4], (int)(uint32_t)(w[4] >> 32), (int)(uint32_t)w[5], (int)(uint32_t)(w[5] >> 32));
    return d;
}
__device__ __forceinline__ int2 load_int2_coherent(const int2* p) {
    const uint64_t w = __hip_atomic_load(reinterpret_cast<const uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return make_int2((int)(uint32_t)w, (int)(uint32_t)(w >> 32));
}
__device__ __forceinline__ void store_int2_coherent(int2* p, int2 v) {
    const uint64_t w = (uint64_t)(uint32_t)v.x | ((uint64_t)(uint32_t)v.y << 32);
    __hip_atomic_store(reinterpret_cast<uint64_t*>(p), w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_u32_coherent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_box_coherent(DNode* nodes, uint32_t i, const Box& b) {
    uint64_t* q = reinterpret_cast<uint64_t*>(nodes + i);
    store_u64_coherent(q + 0, pack2f(b.mnx, b.mny));
    store_u64_coherent(q + 1, pack2f(b.mnz, 0.0f));
    store_u64_coherent(q + 2, pack2f(b.mxx, b.mxy));
    store_u64_coherent(q + 3, pack2f(b.mxz, 0.0f));
}


// Order-preserving map float -> uint32 (used for the order-independent
// scene-bounds reduction through integer atomics).
__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t o) {
    uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    return __uint_as_float(u);
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t l = __lane_id();
    return (l == 0) ? 0ull : ((~0ull) >> (64 - l));
}

// Wave64 reductions through cross-lane shuffles (lowered to DPP/permute).
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        uint32_t w = __shfl_xor(v, o, 64);
        v = (w < v) ? w : v;
    }
    return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        uint32_t w = __shfl_xor(v, o, 64);
        v = (w > v) ? w : v;
    }
    return v;
}

// Inclusive wave64 prefix sum (Hillis-Steele over shuffles).
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v) {
    uint32_t l = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t w = __shfl_up(v, o, 64);
        if (l >= (uint32_t)o) v += w;
    }
    return v;
}

// 30-bit Morton code (reference lbvh_builder.cu:11-27): x in the top bit of each triple.
__device__ __forceinline__ uint32_t expand10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
__device__ __forceinline__ uint32_t morton30(float x, float y, float z) {
    x = fmin_s(fmax_s(x * 1024.0f, 0.0f), 1023.0f);
    y = fmin_s(fmax_s(y * 1024.0f, 0.0f), 1023.0f);
    z = fmin_s(fmax_s(z * 1024.0f, 0.0f), 1023.0f);
    return expand10((uint32_t)x) * 4 + expand10((uint32_t)y) * 2 + expand10((uint32_t)z);
}

// 60-bit Morton code (reference lbvh_plus_builder.cu:20-35): z in the top bit.
__device__ __forceinline__ uint64_t expand21(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}
__device__ __forceinline__ uint64_t morton60(float x, float y, float z) {
    x = fmin_s(fmax_s(x * 1048576.0f, 0.0f), 1048575.0f);
    y = fmin_s(fmax_s(y * 1048576.0f, 0.0f), 1048575.0f);
    z = fmin_s(fmax_s(z * 1048576.0f, 0.0f), 1048575.0f);
    return (expand21((uint64_t)z) << 2) | (expand21((uint64_t)y) << 1) | expand21((uint64_t)x);
}

// Normalised centroid coordinate (reference lbvh_builder.cu:167-181).
__device__ __forceinline__ float norm_coord(float mn, float mx, float smn, float smx) {
    float c = (mn + mx) * 0.5f;
    float ext = smx - smn;
    float div = (ext > 1e-6f) ? ext : 1.0f;
    return (c - smn) / div;
}

}  // namespace bvhk
