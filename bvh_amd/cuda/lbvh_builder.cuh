// lbvh_builder.cuh — drop-in LBVH builder on the MI355X engine.
//
// The include path and the class name are those consumers of the reference
// already use (src/main.cpp:7, src/batch_runner.cpp:3, dynamic_cast at
// batch_runner.cpp:87-95); the class replaces LBVHBuilderCUDA of
// src/cuda/lbvh_builder.cuh:16-65.  It contains no GPU API: it forwards to the
// C ABI of include/bvh_hip.h.
#pragma once

#include "../../include/bvh_builder.h"
#include "../host/hip_builder.hpp"

class LBVHBuilderCUDA : public BVHBuilder {
public:
    LBVHBuilderCUDA() : core_(0, 25, 1) {}
    ~LBVHBuilderCUDA() override = default;

    std::string getName() const override { return "LBVH"; }
    void build(const TriangleMesh& mesh) override { core_.build(mesh); }
    const std::vector<BVHNode>& getNodes() const override { return core_.nodes(); }
    const std::vector<uint32_t>& getIndices() const override { return core_.indices(); }
    float getLastBuildTimeMS() const override { return core_.lastBuildMs(); }
    std::string getTimingBreakdown() const override { return core_.breakdown(); }

    void cleanup() { core_.cleanup(); }

private:
    bvh_amd::HipBuilderCore core_;
};
