// builders.cpp — HipBuilderCore: the BVHBuilder drop-ins over the C ABI.
#include "hip_builder.hpp"

#include <cstdio>
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

std::string HipBuilderCore::breakdown() const {
    std::ostringstream oss;
    oss << std::fixed << std::setprecision(3);
    for (size_t i = 0; i < stages_.size(); ++i) {
        std::string label = stages_[i].first + ":";
        oss << "  " << std::left << std::setw(24) << label << stages_[i].second << " ms";
        if (i + 1 < stages_.size()) oss << "\n";
    }
    return oss.str();
}

}  // namespace bvh_amd
