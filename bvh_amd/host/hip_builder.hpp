// hip_builder.hpp — shared implementation of the drop-in builder classes.
// Host C++ only: reaches the GPU exclusively through the C ABI (bvh_hip.h).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/bvh_builder.h"

namespace bvh_amd {

// One builder's device state: a context (stream + pooled workspace), the
// uploaded mesh, the tree, and the host copies returned by getNodes/getIndices.
class HipBuilderCore {
public:
    HipBuilderCore(int algo, int ploc_radius, int treelet_rounds);
    ~HipBuilderCore();
    HipBuilderCore(const HipBuilderCore&) = delete;
    HipBuilderCore& operator=(const HipBuilderCore&) = delete;

    void build(const TriangleMesh& mesh);   // throws std::runtime_error on failure
    void cleanup();                         // frees device memory and host results

    const std::vector<BVHNode>& nodes() const { return nodes_; }
    const std::vector<uint32_t>& indices() const { return indices_; }
    float lastBuildMs() const { return last_ms_; }
    std::string breakdown() const;          // "  Stage:   x.xxx ms" lines

private:
    void ensureContext();
    int algo_, radius_, rounds_;
    void* ctx_ = nullptr;
    void* tree_ = nullptr;
    std::vector<BVHNode> nodes_;
    std::vector<uint32_t> indices_;
    float last_ms_ = 0.0f;
    std::vector<std::pair<std::string, float>> stages_;
};

}  // namespace bvh_amd
