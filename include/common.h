#pragma once
// Core vector and box types shared by the host and the gfx950 kernels.
//
// Layout-compatible with the reference include/common.h:9-65: float3_cw is a
// 16-byte aligned xyz triple (12 B used), AABB_cw is two of them (32 B).  The
// header compiles with plain g++ (host-only consumers such as evaluator.cpp,
// batch_runner.cpp and main.cpp) and with hipcc (device code).
//
// min/max semantics: every merge in this engine is written as an explicit
// select, min(a,b) = (b < a) ? b : a and max(a,b) = (a < b) ? b : a, which is
// what std::min/std::max compute on the host.  The same select runs on the GPU,
// so boxes (including the sign of zero) are bit-identical on both sides.

#include <float.h>
#include <cmath>
#include <algorithm>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define BVH_HD __host__ __device__
#else
#define BVH_HD
#endif

#ifndef __align__
#define __align__(n) __attribute__((aligned(n)))
#endif

BVH_HD inline float bvh_fmin_sel(float a, float b) { return (b < a) ? b : a; }
BVH_HD inline float bvh_fmax_sel(float a, float b) { return (a < b) ? b : a; }

struct __align__(16) float3_cw {
    float x, y, z;

    BVH_HD float3_cw() : x(0), y(0), z(0) {}
    BVH_HD float3_cw(float a, float b, float c) : x(a), y(b), z(c) {}

    BVH_HD float3_cw operator+(const float3_cw& b) const { return float3_cw(x + b.x, y + b.y, z + b.z); }
    BVH_HD float3_cw operator-(const float3_cw& b) const { return float3_cw(x - b.x, y - b.y, z - b.z); }
    BVH_HD float3_cw operator*(float s) const { return float3_cw(x * s, y * s, z * s); }
};

struct __align__(16) AABB_cw {
    float3_cw min;
    float3_cw max;

    // Empty box: min = +FLT_MAX, max = -FLT_MAX (reference common.h:33-36).
    BVH_HD AABB_cw() {
        min = float3_cw(FLT_MAX, FLT_MAX, FLT_MAX);
        max = float3_cw(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    }
    BVH_HD AABB_cw(const float3_cw& min_, const float3_cw& max_) : min(min_), max(max_) {}

    // 2 * (dx*dy + dy*dz + dz*dx), evaluated left to right without contraction
    // (reference common.h:41-44).
    BVH_HD float surfaceArea() const {
        float3_cw d = max - min;
        return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x);
    }

    // Componentwise union (reference common.h:46-64), as selects.
    BVH_HD AABB_cw merge(const AABB_cw& o) const {
        AABB_cw r;
        r.min.x = bvh_fmin_sel(min.x, o.min.x);
        r.min.y = bvh_fmin_sel(min.y, o.min.y);
        r.min.z = bvh_fmin_sel(min.z, o.min.z);
        r.max.x = bvh_fmax_sel(max.x, o.max.x);
        r.max.y = bvh_fmax_sel(max.y, o.max.y);
        r.max.z = bvh_fmax_sel(max.z, o.max.z);
        return r;
    }
};

static_assert(sizeof(float3_cw) == 16, "float3_cw must stay 16 bytes");
static_assert(sizeof(AABB_cw) == 32, "AABB_cw must stay 32 bytes");
