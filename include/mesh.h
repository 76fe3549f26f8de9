#pragma once
// Structure-of-arrays triangle soup, declaration-compatible with the reference
// include/mesh.h:9-57 (nine fp32 arrays, one entry per triangle).

#include "common.h"
#include <vector>
#include <cstddef>
#include <algorithm>

struct TriangleMesh {
    std::vector<float> v0x, v0y, v0z;
    std::vector<float> v1x, v1y, v1z;
    std::vector<float> v2x, v2y, v2z;

    size_t size() const { return v0x.size(); }

    void resize(size_t n) {
        v0x.resize(n); v0y.resize(n); v0z.resize(n);
        v1x.resize(n); v1y.resize(n); v1z.resize(n);
        v2x.resize(n); v2y.resize(n); v2z.resize(n);
    }

    void clear() {
        v0x.clear(); v0y.clear(); v0z.clear();
        v1x.clear(); v1y.clear(); v1z.clear();
        v2x.clear(); v2y.clear(); v2z.clear();
    }

    void addTriangle(const float3_cw& v0, const float3_cw& v1, const float3_cw& v2) {
        v0x.push_back(v0.x); v0y.push_back(v0.y); v0z.push_back(v0.z);
        v1x.push_back(v1.x); v1y.push_back(v1.y); v1z.push_back(v1.z);
        v2x.push_back(v2.x); v2y.push_back(v2.y); v2z.push_back(v2.z);
    }

    AABB_cw getTriangleBounds(size_t i) const {
        AABB_cw b;
        b.min.x = bvh_fmin_sel(v0x[i], bvh_fmin_sel(v1x[i], v2x[i]));
        b.min.y = bvh_fmin_sel(v0y[i], bvh_fmin_sel(v1y[i], v2y[i]));
        b.min.z = bvh_fmin_sel(v0z[i], bvh_fmin_sel(v1z[i], v2z[i]));
        b.max.x = bvh_fmax_sel(v0x[i], bvh_fmax_sel(v1x[i], v2x[i]));
        b.max.y = bvh_fmax_sel(v0y[i], bvh_fmax_sel(v1y[i], v2y[i]));
        b.max.z = bvh_fmax_sel(v0z[i], bvh_fmax_sel(v1z[i], v2z[i]));
        return b;
    }
};

// Device-side view of the nine arrays (reference include/mesh.h:53-57).
struct TrianglesSoADevice {
    float *v0x, *v0y, *v0z;
    float *v1x, *v1y, *v1z;
    float *v2x, *v2y, *v2z;
};
