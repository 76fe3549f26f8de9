// capi.hip — implementation of the C ABI declared in include/bvh_hip.h.
//
// A context owns one HIP stream (default priority), a set of timing events and
// a grow-only device workspace reused across builds, so steady-state builds
// allocate nothing.  Errors never exit: they set the context's message and
// return a status code.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bvh_hip.h"
#include "dev_common.hpp"
#include "engine.hpp"

using namespace bvhk;

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

struct bvh_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[BVH_MAX_STAGES + 1] = {};
    std::string err;
    uint32_t epoch = 0;
    // workspace
    DevBuf tri_boxes, keys_a, keys_b, vals_a, vals_b, hist_part, bin_start, lookback, scene, handoff, parents,
        ranges;
    DevBuf ploc_box[2], ploc_id[2], ploc_nn, ploc_state, ploc_status;
    DevBuf counters;  // LBVH+ arrival words
    DevBuf trace_ovf, trace_list, trace_io;  // traversal queue and host-call staging
    // last build's sorted keys (for bvh_debug_sorted_keys)
    uint64_t* last_keys = nullptr;
    uint32_t* last_vals = nullptr;
    uint32_t last_n = 0;
};

struct bvh_mesh_rec {
    float* base = nullptr;  // 9 x n floats, one allocation
    float4* tri = nullptr;  // 3 x float4 per triangle (traversal copy)
    uint32_t n = 0;
    size_t stride = 0;      // floats between arrays
    MeshView view() const {
        MeshView m;
        for (int k = 0; k < 9; ++k) m.v[k] = base + k * stride;
        m.n = n;
        return m;
    }
};

struct bvh_tree_rec {
    DevBuf nodes, indices;
    uint32_t num_nodes = 0;
    uint32_t num_prims = 0;
};

namespace {

int set_err(bvh_ctx_t ctx, int code, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

#define HIP_TRY(ctx, call)                                                                                   \
    do {                                                                                                     \
        hipError_t e_ = (call);                                                                              \
        if (e_ != hipSuccess) return set_err(ctx, BVH_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

int ensure(bvh_ctx_t ctx, DevBuf& b, size_t bytes, bool zero = false) {
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return BVH_OK;
    if (b.p) {
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        HIP_TRY(ctx, hipFree(b.p));
        b.p = nullptr;
        b.cap = 0;
    }
    size_t want = bytes + bytes / 8;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        return set_err(ctx, BVH_E_NOMEM, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    b.cap = want;
    if (zero) HIP_TRY(ctx, hipMemsetAsync(b.p, 0, want, ctx->stream));
    return BVH_OK;
}

void free_buf(DevBuf& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

template <class T>
T* P(DevBuf& b) {
    return reinterpret_cast<T*>(b.p);
}

int check_launch(bvh_ctx_t ctx, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_err(ctx, BVH_E_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
    return BVH_OK;
}

struct StageClock {
    bvh_ctx_t ctx;
    int k = 0;
    const char* names[BVH_MAX_STAGES];
    int begin() {
        k = 0;
        HIP_TRY(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
        return BVH_OK;
    }
    // Past the last slot, the final stage absorbs the remaining work (the
    // total stays exact).
    int mark(const char* name) {
        if (k < BVH_MAX_STAGES) ++k;
        names[k - 1] = name;
        HIP_TRY(ctx, hipEventRecord(ctx->ev[k], ctx->stream));
        return BVH_OK;
    }
    int finish(bvh_stage_times* t) {
        HIP_TRY(ctx, hipEventSynchronize(ctx->ev[k]));
        if (!t) return BVH_OK;
        std::memset(t, 0, sizeof(*t));
        t->count = k;
        float total = 0.0f;
        for (int i = 0; i < k; ++i) {
            float ms = 0.0f;
            HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev[i], ctx->ev[i + 1]));
            t->ms[i] = ms;
            std::snprintf(t->names[i], sizeof(t->names[i]), "%s", names[i]);
        }
        HIP_TRY(ctx, hipEventElapsedTime(&total, ctx->ev[0], ctx->ev[k]));
        t->total_ms = total;
        return BVH_OK;
    }
};

int pick_bits(int requested, int span) {
    if (requested == 8 || requested == 10) return requested;
    // Tuned default: 10-bit digits (3 passes over 30 bits, 6 over 60).
    (void)span;
    return 10;
}

int ensure_tree(bvh_ctx_t ctx, bvh_tree_t tree, uint32_t n) {
    int st;
    if ((st = ensure(ctx, tree->nodes, (size_t)(2 * (size_t)n - 1) * sizeof(DNode)))) return st;
    if ((st = ensure(ctx, tree->indices, (size_t)n * 4))) return st;
    tree->num_prims = n;
    tree->num_nodes = 2 * n - 1;
    return BVH_OK;
}

// Front end + sort shared by all builders.  m60 selects the LBVH+ key format.
int build_front(bvh_ctx_t ctx, const MeshView& m, bool m60, int radix_bits, StageClock& clk, uint64_t** keys_out,
                uint32_t** vals_out) {
    const uint32_t n = m.n;
    int st;
    const int begin_bit = m60 ? 0 : 32;
    const int end_bit = m60 ? 60 : 62;
    SortPlan plan = make_sort_plan(n, begin_bit, end_bit, pick_bits(radix_bits, end_bit - begin_bit));
    const size_t bins = 1u << plan.bits;
    const size_t zero_bytes = (sort_zero_bytes(plan) + 15) & ~(size_t)15;
    if ((st = ensure(ctx, ctx->tri_boxes, (size_t)n * 32))) return st;
    if ((st = ensure(ctx, ctx->keys_a, (size_t)n * 8))) return st;
    if ((st = ensure(ctx, ctx->keys_b, (size_t)n * 8))) return st;
    if (m60) {
        if ((st = ensure(ctx, ctx->vals_a, (size_t)n * 4))) return st;
        if ((st = ensure(ctx, ctx->vals_b, (size_t)n * 4))) return st;
    }
    if ((st = ensure(ctx, ctx->hist_part, (size_t)plan.hist_blocks * plan.passes * bins * 4))) return st;
    if ((st = ensure(ctx, ctx->bin_start, (size_t)plan.passes * bins * 4))) return st;
    if ((st = ensure(ctx, ctx->lookback, zero_bytes))) return st;
    if ((st = ensure(ctx, ctx->scene, 256))) return st;

    if ((st = clk.begin())) return st;
    launch_init_scene(P<uint32_t>(ctx->scene), ctx->stream);
    launch_tri_bounds(m, P<float4>(ctx->tri_boxes), P<uint32_t>(ctx->scene), ctx->stream);
    if ((st = check_launch(ctx, "k_tri_bounds"))) return st;
    if ((st = clk.mark("Bounds/Centroids"))) return st;
    launch_morton(P<float4>(ctx->tri_boxes), n, P<uint32_t>(ctx->scene), m60, P<uint64_t>(ctx->keys_a),
                  m60 ? P<uint32_t>(ctx->vals_a) : nullptr, plan, P<uint32_t>(ctx->hist_part),
                  P<uint32_t>(ctx->lookback), zero_bytes / 4, ctx->stream);
    launch_hist_scan(P<uint32_t>(ctx->hist_part), plan, P<uint32_t>(ctx->bin_start), ctx->stream);
    if ((st = check_launch(ctx, "k_morton"))) return st;
    if ((st = clk.mark("Morton Codes"))) return st;
    static const char* pass_names[8] = {"Sort Pass 1", "Sort Pass 2", "Sort Pass 3", "Sort Pass 4",
                                        "Sort Pass 5", "Sort Pass 6", "Sort Pass 7", "Sort Pass 8"};
    uint64_t* kin = P<uint64_t>(ctx->keys_a);
    uint64_t* kout = P<uint64_t>(ctx->keys_b);
    uint32_t* vin = m60 ? P<uint32_t>(ctx->vals_a) : nullptr;
    uint32_t* vout = m60 ? P<uint32_t>(ctx->vals_b) : nullptr;
    for (int p = 0; p < plan.passes; ++p) {
        launch_onesweep_pass(plan, p, n, kin, kout, vin, vout, P<uint32_t>(ctx->bin_start),
                             P<uint32_t>(ctx->lookback), ctx->stream);
        if ((st = check_launch(ctx, "k_onesweep"))) return st;
        if ((st = clk.mark(pass_names[p]))) return st;
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    *keys_out = kin;
    *vals_out = vin;
    return BVH_OK;
}

int check_input_flag(bvh_ctx_t ctx) {
    uint32_t flag[2] = {0, 0};
    HIP_TRY(ctx, hipMemcpyAsync(flag, P<uint32_t>(ctx->scene) + 6, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (flag[0]) return set_err(ctx, BVH_E_NONFINITE, "mesh contains NaN or infinite vertex coordinates");
    if (flag[1]) return set_err(ctx, BVH_E_STATE, "internal error: sorted keys name no primitive");
    return BVH_OK;
}

int build_lbvh(bvh_ctx_t ctx, bvh_mesh_t mesh, const bvh_build_params* prm, bvh_tree_t tree, bvh_stage_times* times) {
    const MeshView m = mesh->view();
    const uint32_t n = m.n;
    int st;
    if ((st = ensure_tree(ctx, tree, n))) return st;
    if ((st = ensure(ctx, ctx->handoff, (size_t)n * 8, true))) return st;
    StageClock clk{ctx};
    uint64_t* keys = nullptr;
    uint32_t* vals = nullptr;
    if ((st = build_front(ctx, m, false, prm->radix_bits, clk, &keys, &vals))) return st;
    HierArgs ha;
    ha.keys = keys;
    ha.vals = nullptr;
    ha.n = n;
    ha.tri_boxes = P<float4>(ctx->tri_boxes);
    ha.nodes = P<DNode>(tree->nodes);
    ha.indices = P<uint32_t>(tree->indices);
    ha.parents = nullptr;
    ha.ranges = nullptr;
    ha.handoff = P<uint64_t>(ctx->handoff);
    ha.epoch = ++ctx->epoch;
    ha.err = P<uint32_t>(ctx->scene) + 7;
    launch_hierarchy(ha, ctx->stream);
    if ((st = check_launch(ctx, "k_hierarchy"))) return st;
    if ((st = clk.mark("Topology+Refit"))) return st;
    if ((st = clk.finish(times))) return st;
    ctx->last_keys = keys;
    ctx->last_vals = nullptr;
    ctx->last_n = n;
    return check_input_flag(ctx);
}

int build_lbvh_plus(bvh_ctx_t ctx, bvh_mesh_t mesh, const bvh_build_params* prm, bvh_tree_t tree,
                    bvh_stage_times* times) {
    const MeshView m = mesh->view();
    const uint32_t n = m.n;
    const int rounds = prm->treelet_rounds;
    if (rounds < 0 || rounds > 64) return set_err(ctx, BVH_E_ARG, "treelet_rounds %d out of range [0, 64]", rounds);
    int st;
    if ((st = ensure_tree(ctx, tree, n))) return st;
    if ((st = ensure(ctx, ctx->handoff, (size_t)n * 8, true))) return st;
    if ((st = ensure(ctx, ctx->parents, (size_t)(2 * (size_t)n) * 4))) return st;
    if ((st = ensure(ctx, ctx->ranges, (size_t)n * 8))) return st;
    if ((st = ensure(ctx, ctx->counters, (size_t)n * 4, true))) return st;
    StageClock clk{ctx};
    uint64_t* keys = nullptr;
    uint32_t* vals = nullptr;
    if ((st = build_front(ctx, m, true, prm->radix_bits, clk, &keys, &vals))) return st;
    HierArgs ha;
    ha.keys = keys;
    ha.vals = vals;
    ha.n = n;
    ha.tri_boxes = P<float4>(ctx->tri_boxes);
    ha.nodes = P<DNode>(tree->nodes);
    ha.indices = P<uint32_t>(tree->indices);
    ha.parents = P<uint32_t>(ctx->parents);
    ha.ranges = P<int2>(ctx->ranges);
    ha.handoff = P<uint64_t>(ctx->handoff);
    ha.epoch = ++ctx->epoch;
    ha.err = P<uint32_t>(ctx->scene) + 7;
    launch_hierarchy(ha, ctx->stream);
    if ((st = check_launch(ctx, "k_hierarchy"))) return st;
    if ((st = clk.mark("Topology+Refit"))) return st;
    static const char* round_names[8] = {"Treelet Round 1", "Treelet Round 2", "Treelet Round 3",
                                         "Treelet Round 4", "Treelet Round 5", "Treelet Round 6",
                                         "Treelet Round 7", "Treelet Rounds 8+"};
    for (int r = 0; r < rounds && n > 1; ++r) {
        TreeletArgs ta;
        ta.n = n;
        ta.nodes = P<DNode>(tree->nodes);
        ta.parents = P<uint32_t>(ctx->parents);
        ta.ranges = P<int2>(ctx->ranges);
        ta.counters = P<uint32_t>(ctx->counters);
        ta.epoch = ++ctx->epoch;
        launch_treelet(ta, ctx->stream);
        if ((st = check_launch(ctx, "k_treelet"))) return st;
        if (r < 7 || r == rounds - 1)
            if ((st = clk.mark(round_names[r < 7 ? r : 7]))) return st;
    }
    if ((st = clk.finish(times))) return st;
    ctx->last_keys = keys;
    ctx->last_vals = vals;
    ctx->last_n = n;
    return check_input_flag(ctx);
}

int build_ploc(bvh_ctx_t ctx, bvh_mesh_t mesh, const bvh_build_params* prm, bvh_tree_t tree, bvh_stage_times* times) {
    const MeshView m = mesh->view();
    const uint32_t n = m.n;
    const int radius = prm->ploc_radius;
    if (radius < 1 || radius > 1024) return set_err(ctx, BVH_E_ARG, "PLOC radius %d out of range [1, 1024]", radius);
    int st;
    if ((st = ensure_tree(ctx, tree, n))) return st;
    StageClock clk{ctx};
    uint64_t* keys = nullptr;
    uint32_t* vals = nullptr;
    if ((st = build_front(ctx, m, false, prm->radix_bits, clk, &keys, &vals))) return st;
    const size_t nn = n;
    if ((st = ensure(ctx, ctx->ploc_box[0], nn * 32))) return st;
    if ((st = ensure(ctx, ctx->ploc_box[1], nn * 32))) return st;
    if ((st = ensure(ctx, ctx->ploc_id[0], nn * 4))) return st;
    if ((st = ensure(ctx, ctx->ploc_id[1], nn * 4))) return st;
    if ((st = ensure(ctx, ctx->ploc_nn, nn * 4))) return st;
    if ((st = ensure(ctx, ctx->ploc_state, (size_t)(kPlocMaxIter + 2) * sizeof(PlocState)))) return st;
    if ((st = ensure(ctx, ctx->ploc_status, (nn / 256 + 2) * 8 + 64))) return st;
    PlocArgs a;
    a.n = n;
    a.radius = radius;
    for (int b = 0; b < 2; ++b) {
        a.box[b] = P<float4>(ctx->ploc_box[b]);
        a.id[b] = P<uint32_t>(ctx->ploc_id[b]);
    }
    a.nn = P<int32_t>(ctx->ploc_nn);
    a.nodes = P<DNode>(tree->nodes);
    a.state = P<PlocState>(ctx->ploc_state);
    a.status = P<uint64_t>(ctx->ploc_status);
    a.tile_ctr = reinterpret_cast<uint32_t*>(a.status + (nn / 256 + 2));
    a.final_slot = kPlocMaxIter + 1;
    launch_ploc_init(a, keys, P<float4>(ctx->tri_boxes), P<uint32_t>(tree->indices), ctx->stream);
    if ((st = check_launch(ctx, "k_ploc_init"))) return st;
    if ((st = clk.mark("PLOC Init"))) return st;
    PlocState hs;
    std::memset(&hs, 0, sizeof(hs));
    hs.m = n;
    int k = 0;
    if (n > 1) {
        // Global iterations in batches; the device skips iterations once the
        // cluster count is small, so a batch may overshoot harmlessly.
        uint32_t m_bound = n;
        while (m_bound > (uint32_t)kPlocSmall) {
            double est = std::log((double)m_bound / kPlocSmall) / std::log(1.0 / 0.72);
            int batch = (int)std::ceil(est);
            batch = batch < 2 ? 2 : (batch > 48 ? 48 : batch);
            // State slots form a ring (a batch never spans more than 48 of them).
            for (int b = 0; b < batch; ++b, k = (k + 1) % kPlocMaxIter) {
                launch_ploc_nn(a, k, m_bound, ctx->stream);
                launch_ploc_merge(a, k, m_bound, ctx->stream);
            }
            if ((st = check_launch(ctx, "k_ploc_nn/merge"))) return st;
            HIP_TRY(ctx, hipMemcpyAsync(&hs, a.state + k, sizeof(hs), hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            if (hs.flags) break;
            m_bound = hs.m;
        }
        if ((st = clk.mark("PLOC Global Iterations"))) return st;
        launch_ploc_tail(a, k, ctx->stream);
        if ((st = check_launch(ctx, "k_ploc_tail"))) return st;
        if ((st = clk.mark("PLOC LDS Tail"))) return st;
    }
    if ((st = clk.finish(times))) return st;
    if (n > 1) {
        HIP_TRY(ctx, hipMemcpyAsync(&hs, a.state + a.final_slot, sizeof(hs), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (times) times->ploc_iterations = (int32_t)hs.iters;
    ctx->last_keys = keys;
    ctx->last_vals = nullptr;
    ctx->last_n = n;
    if ((st = check_input_flag(ctx))) return st;
    if (hs.flags & 2u) return set_err(ctx, BVH_E_STATE, "internal error: sorted keys name no primitive");
    if (hs.flags & 4u) return set_err(ctx, BVH_E_STATE, "internal error: PLOC tail received too many clusters");
    if (hs.flags & 1u)
        return set_err(ctx, BVH_E_NOPROGRESS, "PLOC iteration %u made no merge (%u clusters left)", hs.iters, hs.m);
    if (n > 1 && (hs.m != 1 || hs.node_base != 2 * n - 1))
        return set_err(ctx, BVH_E_STATE, "PLOC ended with %u clusters and %u nodes", hs.m, hs.node_base);
    return BVH_OK;
}

}  // namespace

extern "C" {

const char* bvh_status_string(int status) {
    switch (status) {
        case BVH_OK: return "ok";
        case BVH_E_ARG: return "invalid argument";
        case BVH_E_HIP: return "HIP runtime error";
        case BVH_E_NOMEM: return "out of device memory";
        case BVH_E_NONFINITE: return "non-finite input";
        case BVH_E_NOPROGRESS: return "no progress";
        case BVH_E_BUDGET: return "over memory budget";
        case BVH_E_NODEVICE: return "no HIP device";
        case BVH_E_STATE: return "invalid state";
        case BVH_E_UNSUPPORTED: return "unsupported";
        default: return "unknown status";
    }
}

int bvh_device_count(int* count) {
    if (!count) return BVH_E_ARG;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return BVH_OK;
}

int bvh_ctx_create(int device, bvh_ctx_t* out) {
    if (!out) return BVH_E_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return BVH_E_NODEVICE;
    if (device < 0 || device >= count) return BVH_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return BVH_E_HIP;
    bvh_context* c = new bvh_context();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return BVH_E_HIP;
    }
    for (auto& e : c->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            delete c;
            return BVH_E_HIP;
        }
    }
    *out = c;
    return BVH_OK;
}

int bvh_ctx_destroy(bvh_ctx_t ctx) {
    if (!ctx) return BVH_E_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    DevBuf* bufs[] = {&ctx->tri_boxes, &ctx->keys_a,    &ctx->keys_b, &ctx->vals_a,  &ctx->vals_b,  &ctx->hist_part,
                      &ctx->bin_start, &ctx->lookback,  &ctx->scene,  &ctx->handoff, &ctx->parents, &ctx->ranges};
    for (DevBuf* b : bufs) free_buf(*b);
    DevBuf* pbufs[] = {&ctx->ploc_box[0], &ctx->ploc_box[1], &ctx->ploc_id[0], &ctx->ploc_id[1],
                       &ctx->ploc_nn,     &ctx->ploc_state,  &ctx->ploc_status, &ctx->counters,
                       &ctx->trace_ovf,   &ctx->trace_list,  &ctx->trace_io};
    for (DevBuf* b : pbufs) free_buf(*b);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return BVH_OK;
}

const char* bvh_last_error(bvh_ctx_t ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int bvh_device_alloc(bvh_ctx_t ctx, size_t bytes, void** dptr) {
    if (!ctx || !dptr) return BVH_E_ARG;
    hipError_t e = hipMalloc(dptr, bytes ? bytes : 16);
    if (e != hipSuccess) return set_err(ctx, BVH_E_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return BVH_OK;
}

int bvh_device_free(bvh_ctx_t ctx, void* dptr) {
    if (!ctx) return BVH_E_ARG;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (dptr) HIP_TRY(ctx, hipFree(dptr));
    return BVH_OK;
}

int bvh_copy_h2d(bvh_ctx_t ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return BVH_E_ARG;
    HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BVH_OK;
}

int bvh_copy_d2h(bvh_ctx_t ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return BVH_E_ARG;
    HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BVH_OK;
}

int bvh_mesh_upload(bvh_ctx_t ctx, const float* const* soa, uint32_t n, bvh_mesh_t* out) {
    if (!ctx || !soa || !out) return BVH_E_ARG;
    *out = nullptr;
    if (n == 0) return set_err(ctx, BVH_E_ARG, "mesh has no triangles");
    if (n > (1u << 30)) return set_err(ctx, BVH_E_ARG, "mesh too large (%u > 2^30 triangles)", n);
    for (int k = 0; k < 9; ++k)
        if (!soa[k]) return set_err(ctx, BVH_E_ARG, "soa[%d] is null", k);
    bvh_mesh_rec* m = new bvh_mesh_rec();
    m->n = n;
    m->stride = ((size_t)n + 63) & ~(size_t)63;  // 256-byte aligned arrays
    hipError_t e = hipMalloc(&m->base, m->stride * 9 * sizeof(float));
    if (e != hipSuccess) {
        delete m;
        return set_err(ctx, BVH_E_NOMEM, "mesh allocation failed: %s", hipGetErrorString(e));
    }
    for (int k = 0; k < 9; ++k) {
        e = hipMemcpyAsync(m->base + k * m->stride, soa[k], (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) {
            (void)hipFree(m->base);
            delete m;
            return set_err(ctx, BVH_E_HIP, "mesh upload failed: %s", hipGetErrorString(e));
        }
    }
    e = hipMalloc(&m->tri, (size_t)n * 48);
    if (e != hipSuccess) {
        (void)hipFree(m->base);
        delete m;
        return set_err(ctx, BVH_E_NOMEM, "mesh allocation failed: %s", hipGetErrorString(e));
    }
    launch_tri_aos(m->view(), m->tri, ctx->stream);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        (void)hipFree(m->base);
        (void)hipFree(m->tri);
        delete m;
        return set_err(ctx, BVH_E_HIP, "mesh upload sync failed");
    }
    *out = m;
    return BVH_OK;
}

int bvh_mesh_free(bvh_ctx_t ctx, bvh_mesh_t mesh) {
    if (!ctx || !mesh) return BVH_E_ARG;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (mesh->base) HIP_TRY(ctx, hipFree(mesh->base));
    if (mesh->tri) HIP_TRY(ctx, hipFree(mesh->tri));
    delete mesh;
    return BVH_OK;
}

int bvh_mesh_size(bvh_mesh_t mesh, uint32_t* n) {
    if (!mesh || !n) return BVH_E_ARG;
    *n = mesh->n;
    return BVH_OK;
}

int bvh_tree_create(bvh_ctx_t ctx, bvh_tree_t* out) {
    if (!ctx || !out) return BVH_E_ARG;
    *out = new bvh_tree_rec();
    return BVH_OK;
}

int bvh_tree_free(bvh_ctx_t ctx, bvh_tree_t tree) {
    if (!ctx || !tree) return BVH_E_ARG;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    free_buf(tree->nodes);
    free_buf(tree->indices);
    delete tree;
    return BVH_OK;
}

int bvh_tree_info(bvh_tree_t tree, uint32_t* num_nodes, uint32_t* num_prims) {
    if (!tree) return BVH_E_ARG;
    if (num_nodes) *num_nodes = tree->num_nodes;
    if (num_prims) *num_prims = tree->num_prims;
    return BVH_OK;
}

int bvh_tree_download(bvh_ctx_t ctx, bvh_tree_t tree, void* nodes, uint32_t* indices) {
    if (!ctx || !tree) return BVH_E_ARG;
    if (tree->num_nodes == 0) return set_err(ctx, BVH_E_STATE, "tree is empty");
    if (nodes)
        HIP_TRY(ctx, hipMemcpyAsync(nodes, tree->nodes.p, (size_t)tree->num_nodes * sizeof(DNode),
                                    hipMemcpyDeviceToHost, ctx->stream));
    if (indices && tree->indices.p)
        HIP_TRY(ctx, hipMemcpyAsync(indices, tree->indices.p, (size_t)tree->num_prims * 4, hipMemcpyDeviceToHost,
                                    ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BVH_OK;
}

int bvh_tree_from_nodes(bvh_ctx_t ctx, const void* nodes, uint32_t num_nodes, bvh_tree_t tree) {
    if (!ctx || !nodes || !tree || num_nodes == 0) return BVH_E_ARG;
    int st;
    if ((st = ensure(ctx, tree->nodes, (size_t)num_nodes * sizeof(DNode)))) return st;
    HIP_TRY(ctx, hipMemcpyAsync(tree->nodes.p, nodes, (size_t)num_nodes * sizeof(DNode), hipMemcpyHostToDevice,
                                ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    tree->num_nodes = num_nodes;
    tree->num_prims = (num_nodes + 1) / 2;
    return BVH_OK;
}

int bvh_build(bvh_ctx_t ctx, bvh_mesh_t mesh, const bvh_build_params* params, bvh_tree_t tree,
              bvh_stage_times* times) {
    if (!ctx || !mesh || !params || !tree) return ctx ? set_err(ctx, BVH_E_ARG, "null argument") : BVH_E_ARG;
    if (mesh->n == 0) return set_err(ctx, BVH_E_ARG, "mesh has no triangles");
    if (params->radix_bits != 0 && params->radix_bits != 8 && params->radix_bits != 10)
        return set_err(ctx, BVH_E_UNSUPPORTED, "radix_bits must be 0, 8 or 10");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    switch (params->algo) {
        case BVH_ALGO_LBVH: return build_lbvh(ctx, mesh, params, tree, times);
        case BVH_ALGO_PLOC: return build_ploc(ctx, mesh, params, tree, times);
        case BVH_ALGO_LBVH_PLUS: return build_lbvh_plus(ctx, mesh, params, tree, times);
        default: return set_err(ctx, BVH_E_UNSUPPORTED, "algorithm %d not available", params->algo);
    }
}

int bvh_debug_morton(bvh_ctx_t ctx, bvh_mesh_t mesh, int32_t algo, uint64_t* codes) {
    if (!ctx || !mesh || !codes) return BVH_E_ARG;
    const MeshView m = mesh->view();
    int st;
    if ((st = ensure(ctx, ctx->tri_boxes, (size_t)m.n * 32))) return st;
    if ((st = ensure(ctx, ctx->keys_a, (size_t)m.n * 8))) return st;
    if ((st = ensure(ctx, ctx->scene, 256))) return st;
    launch_init_scene(P<uint32_t>(ctx->scene), ctx->stream);
    launch_tri_bounds(m, P<float4>(ctx->tri_boxes), P<uint32_t>(ctx->scene), ctx->stream);
    launch_morton_codes(P<float4>(ctx->tri_boxes), m.n, P<uint32_t>(ctx->scene), algo == BVH_ALGO_LBVH_PLUS,
                        P<uint64_t>(ctx->keys_a), ctx->stream);
    if ((st = check_launch(ctx, "k_morton_codes"))) return st;
    HIP_TRY(ctx, hipMemcpyAsync(codes, ctx->keys_a.p, (size_t)m.n * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->last_keys = nullptr;
    return BVH_OK;
}

int bvh_debug_sorted_keys(bvh_ctx_t ctx, uint64_t* keys, uint32_t* vals, uint32_t n) {
    if (!ctx || !keys) return BVH_E_ARG;
    if (!ctx->last_keys || n != ctx->last_n) return set_err(ctx, BVH_E_STATE, "no sorted keys of size %u kept", n);
    HIP_TRY(ctx, hipMemcpyAsync(keys, ctx->last_keys, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (vals && ctx->last_vals)
        HIP_TRY(ctx, hipMemcpyAsync(vals, ctx->last_vals, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BVH_OK;
}

int bvh_trace(bvh_ctx_t ctx, bvh_tree_t tree, bvh_mesh_t mesh, const float* d_rays, uint32_t nrays, float* d_t,
              int32_t* d_prim, int32_t* d_counters, float* ms) {
    if (!ctx || !tree || !mesh || (nrays && (!d_rays || !d_t || !d_prim))) return BVH_E_ARG;
    if (tree->num_nodes == 0) return set_err(ctx, BVH_E_STATE, "tree is empty");
    if (nrays == 0) return BVH_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int st;
    if ((st = ensure(ctx, ctx->trace_ovf, 64))) return st;
    if ((st = ensure(ctx, ctx->trace_list, (size_t)nrays * 4))) return st;
    TraceArgs a;
    a.nodes = P<DNode>(tree->nodes);
    a.num_nodes = tree->num_nodes;
    a.num_prims = mesh->n;
    a.tri = mesh->tri;
    a.rays = d_rays;
    a.nrays = nrays;
    a.t_out = d_t;
    a.prim_out = d_prim;
    a.counters = d_counters;
    a.overflow_count = P<uint32_t>(ctx->trace_ovf);
    a.overflow_list = P<uint32_t>(ctx->trace_list);
    HIP_TRY(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    launch_fill_zero(a.overflow_count, 4, ctx->stream);
    launch_trace(a, ctx->stream);
    launch_trace_deep(a, ctx->stream);
    if ((st = check_launch(ctx, "k_trace"))) return st;
    HIP_TRY(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    uint32_t ovf[2] = {0, 0};
    HIP_TRY(ctx, hipMemcpyAsync(ovf, a.overflow_count, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (ms) HIP_TRY(ctx, hipEventElapsedTime(ms, ctx->ev[0], ctx->ev[1]));
    if (ovf[1]) return set_err(ctx, BVH_E_STATE, "traversal needed more than 256 stack entries (or the nodes are not a tree)");
    return BVH_OK;
}

int bvh_trace_host(bvh_ctx_t ctx, bvh_tree_t tree, bvh_mesh_t mesh, const float* rays, uint32_t nrays, float* t,
                   int32_t* prim, int32_t* counters, float* ms) {
    if (!ctx || !tree || !mesh || (nrays && (!rays || !t || !prim))) return BVH_E_ARG;
    if (nrays == 0) return BVH_OK;
    const size_t rb = (size_t)nrays * 24, tb = (size_t)nrays * 4, cb = (size_t)nrays * 12;
    int st;
    if ((st = ensure(ctx, ctx->trace_io, rb + 2 * tb + cb + 256))) return st;
    char* base = P<char>(ctx->trace_io);
    float* d_rays = reinterpret_cast<float*>(base);
    float* d_t = reinterpret_cast<float*>(base + ((rb + 63) & ~(size_t)63));
    int32_t* d_prim = reinterpret_cast<int32_t*>(reinterpret_cast<char*>(d_t) + ((tb + 63) & ~(size_t)63));
    int32_t* d_cnt = counters ? reinterpret_cast<int32_t*>(reinterpret_cast<char*>(d_prim) + ((tb + 63) & ~(size_t)63))
                              : nullptr;
    HIP_TRY(ctx, hipMemcpyAsync(d_rays, rays, rb, hipMemcpyHostToDevice, ctx->stream));
    if ((st = bvh_trace(ctx, tree, mesh, d_rays, nrays, d_t, d_prim, d_cnt, ms))) return st;
    HIP_TRY(ctx, hipMemcpyAsync(t, d_t, tb, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(prim, d_prim, tb, hipMemcpyDeviceToHost, ctx->stream));
    if (counters) HIP_TRY(ctx, hipMemcpyAsync(counters, d_cnt, cb, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BVH_OK;
}

int bvh_debug_stream_copy(bvh_ctx_t ctx, void* d_dst, const void* d_src, size_t bytes, float* ms) {
    if (!ctx || !d_dst || !d_src || (bytes % 16) != 0) return BVH_E_ARG;
    HIP_TRY(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    launch_stream_copy(d_src, d_dst, bytes, ctx->stream);
    int st;
    if ((st = check_launch(ctx, "k_stream_copy"))) return st;
    HIP_TRY(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev[1]));
    if (ms) HIP_TRY(ctx, hipEventElapsedTime(ms, ctx->ev[0], ctx->ev[1]));
    return BVH_OK;
}

int bvh_sort_pairs_u64(bvh_ctx_t ctx, uint64_t* d_keys, uint32_t* d_vals, uint32_t n, int begin_bit, int end_bit,
                       int radix_bits, float* ms) {
    if (!ctx || !d_keys) return BVH_E_ARG;
    if (n == 0) return BVH_OK;
    if (n >= (1u << 30)) return set_err(ctx, BVH_E_ARG, "n too large");
    if (begin_bit < 0 || end_bit > 64 || begin_bit >= end_bit) return set_err(ctx, BVH_E_ARG, "bad bit range");
    if (radix_bits != 0 && radix_bits != 8 && radix_bits != 10)
        return set_err(ctx, BVH_E_UNSUPPORTED, "radix_bits must be 0, 8 or 10");
    const int bits = pick_bits(radix_bits, end_bit - begin_bit);
    SortPlan plan = make_sort_plan(n, begin_bit, end_bit, bits);
    if (plan.passes > 8) return set_err(ctx, BVH_E_ARG, "too many passes");
    const size_t bins = 1u << plan.bits;
    const size_t zero_bytes = (sort_zero_bytes(plan) + 15) & ~(size_t)15;
    int st;
    if ((st = ensure(ctx, ctx->keys_b, (size_t)n * 8))) return st;
    if (d_vals && (st = ensure(ctx, ctx->vals_b, (size_t)n * 4))) return st;
    if ((st = ensure(ctx, ctx->hist_part, (size_t)plan.hist_blocks * plan.passes * bins * 4))) return st;
    if ((st = ensure(ctx, ctx->bin_start, (size_t)plan.passes * bins * 4))) return st;
    if ((st = ensure(ctx, ctx->lookback, zero_bytes))) return st;
    HIP_TRY(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    launch_sort_histogram(d_keys, n, plan, P<uint32_t>(ctx->hist_part), ctx->stream);
    launch_hist_scan(P<uint32_t>(ctx->hist_part), plan, P<uint32_t>(ctx->bin_start), ctx->stream);
    launch_fill_zero(P<uint32_t>(ctx->lookback), zero_bytes / 4, ctx->stream);
    uint64_t* kout = nullptr;
    uint32_t* vout = nullptr;
    launch_onesweep(plan, n, d_keys, P<uint64_t>(ctx->keys_b), d_vals, d_vals ? P<uint32_t>(ctx->vals_b) : nullptr,
                    P<uint32_t>(ctx->bin_start), P<uint32_t>(ctx->lookback), &kout, &vout, ctx->stream);
    if ((st = check_launch(ctx, "k_onesweep"))) return st;
    HIP_TRY(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    if (kout != d_keys) {
        HIP_TRY(ctx, hipMemcpyAsync(d_keys, kout, (size_t)n * 8, hipMemcpyDeviceToDevice, ctx->stream));
        if (d_vals) HIP_TRY(ctx, hipMemcpyAsync(d_vals, vout, (size_t)n * 4, hipMemcpyDeviceToDevice, ctx->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (ms) HIP_TRY(ctx, hipEventElapsedTime(ms, ctx->ev[0], ctx->ev[1]));
    ctx->last_keys = nullptr;
    return BVH_OK;
}

}  // extern "C"
