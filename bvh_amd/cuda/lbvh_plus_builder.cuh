// lbvh_plus_builder.cuh — drop-in LBVH+ builder (7-leaf treelet optimisation)
// on the MI355X engine; replaces LBVHPlusBuilderCUDA of the reference
// (src/cuda/lbvh_plus_builder.cuh:11-61).  The reference runs one optimisation
// round; `rounds` selects more.  Forwards to the C ABI (include/bvh_hip.h).
#pragma once

#include "../../include/bvh_builder.h"
#include "../host/hip_builder.hpp"

class LBVHPlusBuilderCUDA : public BVHBuilder {
public:
    explicit LBVHPlusBuilderCUDA(int rounds = 1) : core_(1, 25, rounds) {}
    ~LBVHPlusBuilderCUDA() override = default;

    std::string getName() const override { return "LBVH+"; }
    void build(const TriangleMesh& mesh) override { core_.build(mesh); }
    const std::vector<BVHNode>& getNodes() const override { return core_.nodes(); }
    const std::vector<uint32_t>& getIndices() const override { return core_.indices(); }
    float getLastBuildTimeMS() const override { return core_.lastBuildMs(); }
    std::string getTimingBreakdown() const override { return core_.breakdown(); }

    void cleanup() { core_.cleanup(); }

private:
    bvh_amd::HipBuilderCore core_;
};
