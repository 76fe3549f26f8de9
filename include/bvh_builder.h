#pragma once
// Abstract builder interface, declaration-identical to the reference
// include/bvh_builder.h:10-37, so evaluator.cpp, batch_runner.cpp and main.cpp
// consume the MI355X builders unchanged.

#include "mesh.h"
#include "bvh_node.h"
#include <vector>
#include <string>
#include <cstdint>

class BVHBuilder {
public:
    virtual ~BVHBuilder() = default;

    // Name of the algorithm ("LBVH", "LBVH+", "PLOC CUDA (r=25)").
    virtual std::string getName() const = 0;

    // Builds the BVH over the mesh.
    virtual void build(const TriangleMesh& mesh) = 0;

    // Flattened node array; node 0 is the root.
    virtual const std::vector<BVHNode>& getNodes() const = 0;

    // indices[i] = original triangle index of the i-th primitive in BVH order.
    virtual const std::vector<uint32_t>& getIndices() const = 0;

    // Device time of the last build in milliseconds (upload excluded).
    virtual float getLastBuildTimeMS() const = 0;

    // Newline-separated "  Stage:   x.xxx ms" lines.
    virtual std::string getTimingBreakdown() const { return ""; }
};
