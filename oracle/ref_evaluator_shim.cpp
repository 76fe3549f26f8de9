// ref_evaluator_shim.cpp — TEST INFRASTRUCTURE ONLY.
// C entry points over the reference's own BVHEvaluator (src/evaluator.cpp,
// compiled unchanged from the reference tree against this repository's
// drop-in headers by oracle/Makefile).  Used to pin the oracle's SAH and
// depth statistics (evaluator.cpp:38-113) on trees built here.
#include "evaluator.h"

extern "C" float ref_compute_sah(const void* nodes, long num_nodes) {
    std::vector<BVHNode> v((const BVHNode*)nodes, (const BVHNode*)nodes + num_nodes);
    std::vector<uint32_t> indices;
    TriangleMesh mesh;
    return BVHEvaluator::computeSAH(v, indices, mesh);
}
