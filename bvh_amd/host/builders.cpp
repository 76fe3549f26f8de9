// builders.cpp — HipBuilderCore: the BVHBuilder drop-ins over the C ABI.
#include "hip_builder.hpp"

#include <cstdio>
#include <cstring>
#include <iomanip>
#include <sstream>
#include <stdexcept>

#include "../../include/bvh_hip.h"

namespace bvh_amd {

namespace {
[[noreturn]] void fail(bvh_ctx_t ctx, int st, const char* what) {
    std::string msg = std::string("bvh_amd: ") + what + " failed (" + bvh_status_string(st) + ")";
    if (ctx) msg += std::string(": ") + bvh_last_error(ctx);
    throw std::runtime_error(msg);
}
}  // namespace

HipBuilderCore::HipBuilderCore(int algo, int ploc_radius, int treelet_rounds)
    : algo_(algo), radius_(ploc_radius), rounds_(treelet_rounds) {}

HipBuilderCore::~HipBuilderCore() {
    cleanup();
    if (ctx_) bvh_ctx_destroy(static_cast<bvh_ctx_t>(ctx_));
}

void HipBuilderCore::ensureContext() {
    if (ctx_) return;
    bvh_ctx_t c = nullptr;
    int st = bvh_ctx_create(0, &c);
    if (st != BVH_OK) fail(nullptr, st, "bvh_ctx_create");
    ctx_ = c;
}

void HipBuilderCore::build(const TriangleMesh& mesh) {
    ensureContext();
    bvh_ctx_t ctx = static_cast<bvh_ctx_t>(ctx_);
    const uint32_t n = (uint32_t)mesh.size();
    if (n == 0) throw std::runtime_error("bvh_amd: cannot build a BVH over an empty mesh");
    const float* soa[9] = {mesh.v0x.data(), mesh.v0y.data(), mesh.v0z.data(), mesh.v1x.data(), mesh.v1y.data(),
                           mesh.v1z.data(), mesh.v2x.data(), mesh.v2y.data(), mesh.v2z.data()};
    bvh_mesh_t dm = nullptr;
    int st = bvh_mesh_upload(ctx, soa, n, &dm);
    if (st != BVH_OK) fail(ctx, st, "mesh upload");
    if (!tree_) {
        bvh_tree_t t = nullptr;
        st = bvh_tree_create(ctx, &t);
        if (st != BVH_OK) {
            bvh_mesh_free(ctx, dm);
            fail(ctx, st, "tree create");
        }
        tree_ = t;
    }
    bvh_tree_t tree = static_cast<bvh_tree_t>(tree_);
    bvh_build_params p;
    p.algo = algo_;
    p.ploc_radius = radius_;
    p.treelet_rounds = rounds_;
    p.radix_bits = 0;
    p.flags = 0;
    bvh_stage_times times;
    st = bvh_build(ctx, dm, &p, tree, &times);
    if (st != BVH_OK) {
        bvh_mesh_free(ctx, dm);
        fail(ctx, st, "build");
    }
    uint32_t nn = 0, np = 0;
    bvh_tree_info(tree, &nn, &np);
    nodes_.resize(nn);
    indices_.resize(np);
    st = bvh_tree_download(ctx, tree, nodes_.data(), indices_.data());
    bvh_mesh_free(ctx, dm);
    if (st != BVH_OK) fail(ctx, st, "download");
    last_ms_ = times.total_ms;
    stages_.clear();
    for (int i = 0; i < times.count; ++i) stages_.emplace_back(times.names[i], times.ms[i]);
}

void HipBuilderCore::cleanup() {
    if (tree_ && ctx_) bvh_tree_free(static_cast<bvh_ctx_t>(ctx_), static_cast<bvh_tree_t>(tree_));
    tree_ = nullptr;
    nodes_.clear();
    nodes_.shrink_to_fit();
    indices_.clear();
    indices_.shrink_to_fit();
}

// The reference's stage lines (lbvh_builder.cu:490-498,
// lbvh_plus_builder.cu:535-543, ploc_builder.cu:393-400), filled from the
// finer stages the C ABI reports.  Fused stages: the hierarchy kernel builds
// topology and boxes in one pass (its time is the "Topology Build" line and
// "Bottom-Up Refit" reads 0), the treelet kernel refits while it optimises,
// and PLOC's iteration kernels search, merge and compact together (their time
// is the "NN Search" line; "Merging" and "Compaction" read 0).
std::string HipBuilderCore::breakdown() const {
    auto sum = [&](const char* prefix) {
        float t = 0.0f;
        for (const auto& st : stages_)
            if (st.first.compare(0, std::strlen(prefix), prefix) == 0) t += st.second;
        return t;
    };
    std::ostringstream oss;
    oss << std::fixed << std::setprecision(3);
    if (algo_ == BVH_ALGO_LBVH) {
        oss << "  Bounds/Centroids: " << sum("Bounds/Centroids") << " ms\n";
        oss << "  Morton Codes:     " << sum("Morton Codes") << " ms\n";
        oss << "  CUB Radix Sort:   " << sum("Sort Pass") << " ms\n";
        oss << "  Topology Build:   " << sum("Topology+Refit") << " ms\n";
        oss << "  Bottom-Up Refit:  " << 0.0f << " ms";
    } else if (algo_ == BVH_ALGO_LBVH_PLUS) {
        const float opt = sum("Treelet Round");
        oss << "  Bounds/Centroids: " << sum("Bounds/Centroids") << " ms\n";
        oss << "  Morton Codes:     " << sum("Morton Codes") << " ms\n";
        oss << "  Radix Sort:       " << sum("Sort Pass") << " ms\n";
        oss << "  Topology Build:   " << sum("Topology+Refit") << " ms\n";
        oss << "  Refit + Optimize: " << opt << " ms (opt: " << opt << " ms)";
    } else {
        oss << "  Initialization:   "
            << sum("Bounds/Centroids") + sum("Morton Codes") + sum("Sort Pass") + sum("PLOC Init") << " ms\n";
        oss << "  NN Search (Red):  " << sum("PLOC Global") + sum("PLOC LDS Tail") << " ms\n";
        oss << "  Merging (Green):  " << 0.0f << " ms\n";
        oss << "  Compaction (Mag): " << 0.0f << " ms";
    }
    return oss.str();
}

}  // namespace bvh_amd
