// ploc_builder.cuh — drop-in PLOC builder on the MI355X engine; replaces
// PLOCBuilderCUDA of the reference (src/cuda/ploc_builder.cuh:10-67), same
// constructor (radius, default 25) and name string (main.cpp:467 compares it).
// Forwards to the C ABI (include/bvh_hip.h).
#pragma once

#include <string>

#include "../../include/bvh_builder.h"
#include "../host/hip_builder.hpp"

class PLOCBuilderCUDA : public BVHBuilder {
public:
    explicit PLOCBuilderCUDA(int r = 25) : radius_(r), core_(2, r, 1) {}
    ~PLOCBuilderCUDA() override = default;

    std::string getName() const override { return "PLOC CUDA (r=" + std::to_string(radius_) + ")"; }
    void build(const TriangleMesh& mesh) override { core_.build(mesh); }
    const std::vector<BVHNode>& getNodes() const override { return core_.nodes(); }
    const std::vector<uint32_t>& getIndices() const override { return core_.indices(); }
    float getLastBuildTimeMS() const override { return core_.lastBuildMs(); }
    std::string getTimingBreakdown() const override { return core_.breakdown(); }

    void cleanup() { core_.cleanup(); }

private:
    int radius_;
    bvh_amd::HipBuilderCore core_;
};
