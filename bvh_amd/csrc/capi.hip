// capi.hip — implementation of the C ABI declared in include/bvh_hip.h.
//
// A context owns one HIP stream (default priority), a set of timing events and
// a grow-only device workspace reused across builds, so steady-state builds
// allocate nothing.  A batch owns a workspace of its own plus a captured
// hipGraph of its whole launch sequence.  Errors never exit: they set the
// context's message and return a status code.
//
// Every build is expressed as a job over segments (meshes): a standalone build
// is a job of one segment, a batch is a job of k.  Both run the same kernels,
// so each batched tree is bit-identical to its standalone build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bvh_hip.h"
#include "dev_common.hpp"
#include "engine.hpp"

using namespace bvhk;

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    bool fixed = false;    // a region of the no-allocation scratch: never freed or grown
    bool measure = false;  // footprint estimation: record the capacity an allocation would get
    int dev = 0;           // device of p
};

namespace {
void dev_free(int device, void* p);
}  // namespace

// Device workspace of one build pipeline.
struct Workspace {
    DevBuf tri_boxes, keys_a, keys_b, vals_a, vals_b, hist_tot, lookback, scene, handoff, epoch;
    DevBuf hunits[2], hstatus, hctl, hcounts;  // LBVH/LBVH+ hierarchy levels
    DevBuf parents, tl_list, counters;  // LBVH+: parent links, ready lists + list lengths, arrival words
    DevBuf ploc_box[2], ploc_id[2], ploc_sid[2], ploc_seg[2], ploc_nn, ploc_state, ploc_status, ploc_out;
    template <class F>
    void each(F f) {
        DevBuf* all[] = {&tri_boxes,   &keys_a,      &keys_b,      &vals_a,      &vals_b,     &hist_tot,
                         &lookback,    &scene,       &handoff,     &epoch,       &hunits[0],
                         &hunits[1],   &hstatus,     &hctl,        &hcounts,     &parents,
                         &tl_list,     &counters,    &ploc_box[0], &ploc_box[1], &ploc_id[0], &ploc_id[1],
                         &ploc_sid[0], &ploc_sid[1], &ploc_seg[0], &ploc_seg[1], &ploc_nn,    &ploc_state,
                         &ploc_status, &ploc_out};
        for (DevBuf* b : all) f(*b);
    }
};

// Segment table and block tables of a segmented job.
struct SegTables {
    DevBuf segs, fblocks, hblocks;
    std::vector<SegDesc> host_segs;
    void release() {
        for (DevBuf* b : {&segs, &fblocks, &hblocks}) {
            if (b->p) dev_free(b->dev, b->p);
            b->p = nullptr;
            b->cap = 0;
        }
    }
};

// An asynchronous call whose status words are still on their way to the host:
// a stream-ordered copy lands them in a 16-word pinned slot, read by bvh_sync.
enum PendingKind { kPendBuild, kPendPloc, kPendTrace };
struct Pending {
    PendingKind kind;
    uint32_t n;  // triangles of the build (PLOC record check)
};
constexpr uint32_t kPinSlotWords = 16;     // build: 8 flag words + PLOC record; trace: 2 overflow words
constexpr uint32_t kPinChunkSlots = 256;   // slots per pinned allocation (never moved while in use)
constexpr uint32_t kMaxPending = 1u << 16; // asynchronous calls between two bvh_sync

struct bvh_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[BVH_MAX_STAGES + 1] = {};
    std::string err;
    uint32_t treelet_epoch = 0;  // LBVH+ arrival-word tags
    Workspace ws;
    SegTables tabs;  // eager segmented builds
    DevBuf trace_ovf, trace_list, trace_io;  // traversal queue and host-call staging
    DevBuf trace_spill;                      // spill slots of deep traversal stacks
    DevBuf ray_keys[2];                      // ray coherence keys (sorted before tracing)
    int32_t trace_order = BVH_TRACE_REFERENCE;  // visit order of this context's traces
    uint32_t* host_flags = nullptr;          // pinned: a single build's 8 flag words land here
    std::vector<uint32_t*> pin_chunks;       // pinned status slots of asynchronous calls
    char* stage[2] = {nullptr, nullptr};     // pinned staging buffers of host copies (kStageBytes each)
    uint32_t* pin_small = nullptr;           // pinned words for small synchronous readbacks
    hipEvent_t stage_ev[2] = {nullptr, nullptr};  // DMA from/to stage[i] done
    std::vector<Pending> pending;            // asynchronous calls since the last bvh_sync, in order
    // last build's sorted keys (for bvh_debug_sorted_keys)
    uint64_t* last_keys = nullptr;
    uint32_t* last_vals = nullptr;
    uint32_t last_n = 0;
};

struct bvh_mesh_rec {
    float* base = nullptr;  // 9 x n floats, one allocation
    float4* tri = nullptr;  // 3 x float4 per triangle (traversal copy)
    size_t base_cap = 0, tri_cap = 0;  // allocation sizes (blocks may come from the release cache)
    int device = 0;
    uint32_t n = 0;
    bool nonfinite = false;  // a coordinate is NaN or infinite: every build is refused
    size_t stride = 0;      // floats between arrays
    MeshView view() const {
        MeshView m;
        for (int k = 0; k < 9; ++k) m.v[k] = base + k * stride;
        m.n = n;
        return m;
    }
};

struct bvh_tree_rec {
    int device = -1;         // device of its storage (set when first allocated)
    DevBuf nodes, indices;
    DevBuf rec;              // traversal records, built on first trace after a build
    bool rec_valid = false;
    uint32_t num_nodes = 0;
    uint32_t num_prims = 0;
};

// A build job: the segments and how kernel blocks map onto them.
struct Job {
    SegBlocks fsb;        // front blocks (fsb.chunk triangles each)
    SegBlocks hsb;        // hierarchy blocks (hier_chunk() leaves each)
    uint32_t k = 1;       // segments
    uint32_t total = 0;   // triangles over all segments
    uint32_t maxn = 0;    // largest segment
    KeyFormat kf{32, 62};
    bool m60 = false;
    int begin_bit = 32, end_bit = 62;
};

struct bvh_batch_rec {
    bvh_ctx_t ctx = nullptr;
    int algo = BVH_ALGO_LBVH;
    int radius = 25;
    uint32_t k = 0;
    std::vector<bvh_mesh_t> meshes;
    std::vector<bvh_tree_t> trees;
    SegTables tabs;
    Workspace ws;
    Job job;
    int radix_bits = 10;
    int ploc_iters = 0;   // global PLOC iterations captured in the graph
    uint32_t nn_grid = 0, merge_grid = 0;
    int sort_block = 0;   // onesweep tile threads (kSortThreads or kSortThreadsSmall)
    uint32_t front_chunk = 0;  // triangles per front-end block (kFrontChunk or kFrontChunkLarge)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bvh_batch_info info;
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

#define TRY(call)                  \
    do {                           \
        int st_ = (call);          \
        if (st_) return st_;       \
    } while (0)

// ---- released device blocks ------------------------------------------------
// Meshes and trees given back with bvh_mesh_release / bvh_tree_release (their
// owner vouches that no queued work uses them) keep their device blocks here,
// per device, instead of going to hipFree, which waits for the whole device;
// a later mesh upload or new buffer takes the smallest cached block that is
// large enough and at most a quarter plus 2 MiB larger than asked for (a
// wide window, so that most released blocks find a taker; bvh_mesh_bytes
// and bvh_tree_bytes report what an object really holds).  An
// allocation that fails frees every cached block of its device and tries
// once more.  The cache holds at most `limit` bytes (default an eighth of the
// device's memory); a released block beyond it is freed at once.
struct BlockCache {
    std::mutex mu;
    std::multimap<size_t, void*> blocks;  // capacity -> block
    uint64_t bytes = 0;
    uint64_t limit = 0;  // 0: not set yet
};
constexpr int kCacheDevices = 64;

BlockCache& block_cache(int device) {
    static BlockCache caches[kCacheDevices];
    return caches[(unsigned)device % kCacheDevices];
}

// Device blocks come from one memory pool per device that keeps what is
// freed (release threshold: never), allocated and freed in order on a private
// stream of the device: a long-running process (a server, or one server after
// another) allocates and frees without the driver once the pool has grown.
// Through the driver, each allocation cost ~65 us in a fresh process and
// ~1 ms after some 9,000 allocate/free cycles.  bvh_device_cache_trim hands
// the pool's unused memory back; bvh_device_memory counts it as free.
struct DevPool {
    std::once_flag once;
    hipMemPool_t pool = nullptr;
    hipStream_t stream = nullptr;
    hipError_t err = hipSuccess;
};

DevPool* dev_pool(int device) {
    static DevPool pools[kCacheDevices];
    DevPool& d = pools[(unsigned)device % kCacheDevices];
    std::call_once(d.once, [&] {
        int prev = 0;
        (void)hipGetDevice(&prev);
        d.err = hipSetDevice(device);
        hipMemPoolProps props = {};
        props.allocType = hipMemAllocationTypePinned;
        props.handleTypes = hipMemHandleTypeNone;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = device;
        if (d.err == hipSuccess) d.err = hipMemPoolCreate(&d.pool, &props);
        if (d.err == hipSuccess) {
            uint64_t never = UINT64_MAX;
            d.err = hipMemPoolSetAttribute(d.pool, hipMemPoolAttrReleaseThreshold, &never);
        }
        if (d.err == hipSuccess) d.err = hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking);
        (void)hipSetDevice(prev);
    });
    return &d;
}

hipError_t dev_malloc(int device, void** p, size_t bytes) {
    DevPool* d = dev_pool(device);
    if (d->err != hipSuccess) return d->err;
    hipError_t e = hipMallocFromPoolAsync(p, bytes, d->pool, d->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    return e;
}

// The caller has waited for every stream that used p.
void dev_free(int device, void* p) {
    if (!p) return;
    DevPool* d = dev_pool(device);
    if (d->err == hipSuccess) (void)hipFreeAsync(p, d->stream);
}

// Unused bytes the pool holds (0 before its first allocation).
uint64_t pool_unused(int device) {
    DevPool* d = dev_pool(device);
    if (d->err != hipSuccess) return 0;
    uint64_t reserved = 0, used = 0;
    if (hipMemPoolGetAttribute(d->pool, hipMemPoolAttrReservedMemCurrent, &reserved) != hipSuccess ||
        hipMemPoolGetAttribute(d->pool, hipMemPoolAttrUsedMemCurrent, &used) != hipSuccess)
        return 0;
    return reserved > used ? reserved - used : 0;
}

void* cache_take(int device, size_t want, size_t* cap) {
    BlockCache& c = block_cache(device);
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.blocks.lower_bound(want);
    if (it == c.blocks.end() || it->first > want + want / 4 + ((size_t)2 << 20)) return nullptr;
    void* p = it->second;
    *cap = it->first;
    c.bytes -= it->first;
    c.blocks.erase(it);
    return p;
}

void cache_give(int device, void* p, size_t cap) {
    if (!p) return;
    BlockCache& c = block_cache(device);
    {
        std::lock_guard<std::mutex> lk(c.mu);
        if (c.limit == 0) {
            size_t fr = 0, tot = 0;
            c.limit = hipMemGetInfo(&fr, &tot) == hipSuccess ? tot / 8 : ((uint64_t)8 << 30);
        }
        if (c.bytes + cap <= c.limit) {
            c.blocks.emplace(cap, p);
            c.bytes += cap;
            return;
        }
    }
    dev_free(device, p);
}

// Frees every cached block of the device and hands the pool's unused memory
// back to the driver.
void cache_trim(int device) {
    BlockCache& c = block_cache(device);
    std::lock_guard<std::mutex> lk(c.mu);
    for (auto& kv : c.blocks) dev_free(device, kv.second);
    c.blocks.clear();
    c.bytes = 0;
    DevPool* d = dev_pool(device);
    if (d->err == hipSuccess && hipStreamSynchronize(d->stream) == hipSuccess) (void)hipMemPoolTrimTo(d->pool, 0);
}

// Runs f with `device` current, then restores the caller's device.
template <class F>
int on_device(int32_t device, F f) {
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(device) != hipSuccess) return BVH_E_HIP;
    f();
    return hipSetDevice(prev) == hipSuccess ? BVH_OK : BVH_E_HIP;
}

// Device memory of at least `want` bytes (*cap = its size): a cached block or
// a new allocation.
hipError_t dev_alloc(int device, void** p, size_t want, size_t* cap) {
    if ((*p = cache_take(device, want, cap)) != nullptr) return hipSuccess;
    hipError_t e = dev_malloc(device, p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        bool any;
        {
            BlockCache& c = block_cache(device);
            std::lock_guard<std::mutex> lk(c.mu);
            any = !c.blocks.empty();
        }
        if (any) {
            cache_trim(device);
            e = dev_malloc(device, p, want);
        }
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return e;
    }
    *cap = want;
    return hipSuccess;
}

// Grow-only allocation; a new buffer is zeroed when `zero` is set (never
// called while a stream is being captured).
int ensure(bvh_ctx_t ctx, DevBuf& b, size_t bytes, bool zero = false) {
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return BVH_OK;
    if (b.measure) {
        b.cap = ((bytes + bytes / 8) + 255) & ~(size_t)255;
        return BVH_OK;
    }
    if (b.fixed)
        return set_err(ctx, BVH_E_STATE, "internal error: no-allocation scratch region of %zu bytes needs %zu", b.cap,
                       bytes);
    if (b.p) {
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        dev_free(b.dev, b.p);
        b.p = nullptr;
        b.cap = 0;
    }
    size_t want = bytes + bytes / 8;
    want = (want + 255) & ~(size_t)255;
    size_t cap = 0;
    hipError_t e = dev_alloc(ctx->device, &b.p, want, &cap);
    if (e != hipSuccess) {
        b.p = nullptr;
        return set_err(ctx, BVH_E_NOMEM, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    b.cap = cap;
    b.dev = ctx->device;
    if (zero) HIP_TRY(ctx, hipMemsetAsync(b.p, 0, cap, ctx->stream));
    return BVH_OK;
}

void free_buf(DevBuf& b) {
    if (b.p) dev_free(b.dev, b.p);
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

// Per-stage HIP events, recorded only when the caller asked for times (a
// build without them enqueues its kernels and nothing else; every builder
// ends with check_input_flags, which synchronises the stream).
struct StageClock {
    bvh_ctx_t ctx;
    bool on = true;
    int k = 0;
    const char* names[BVH_MAX_STAGES];
    int begin() {
        k = 0;
        if (on) HIP_TRY(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
        return BVH_OK;
    }
    // Past the last slot, the final stage absorbs the remaining work (the
    // total stays exact).
    int mark(const char* name) {
        if (!on) return BVH_OK;
        if (k < BVH_MAX_STAGES) ++k;
        names[k - 1] = name;
        HIP_TRY(ctx, hipEventRecord(ctx->ev[k], ctx->stream));
        return BVH_OK;
    }
    int finish(bvh_stage_times* t) {
        if (!on) return BVH_OK;
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

// Stage mark on an optional clock (captured sequences carry none).
int mark(StageClock* clk, const char* name) { return clk ? clk->mark(name) : BVH_OK; }

int pick_bits(int requested) {
    if (requested == 8 || requested == 10) return requested;
    // Measured default: an 8-bit pass (256 bins) runs in about half the time
    // of a 10-bit pass, so 4 passes over 30 bits beat 3 (and 8 over 60 beat 6).
    return 8;
}

// A tree's storage lives on the device of the context that first allocates it.
int tree_on(bvh_ctx_t ctx, bvh_tree_t tree) {
    if (tree->device >= 0 && tree->device != ctx->device)
        return set_err(ctx, BVH_E_ARG, "tree belongs to device %d, context to device %d", tree->device, ctx->device);
    tree->device = ctx->device;
    return BVH_OK;
}

int ensure_tree(bvh_ctx_t ctx, bvh_tree_t tree, uint32_t n) {
    TRY(tree_on(ctx, tree));
    TRY(ensure(ctx, tree->nodes, (size_t)(2 * (size_t)n - 1) * sizeof(DNode)));
    TRY(ensure(ctx, tree->indices, (size_t)n * 4));
    tree->num_prims = n;
    tree->num_nodes = 2 * n - 1;
    tree->rec_valid = false;
    return BVH_OK;
}

SegDesc seg_of(const bvh_mesh_rec* mesh, bvh_tree_rec* tree, uint32_t start) {
    SegDesc d;
    for (int k = 0; k < 9; ++k) d.v[k] = mesh->base + k * mesh->stride;
    d.n = mesh->n;
    d.start = start;
    d.nodes = tree ? P<DNode>(tree->nodes) : nullptr;
    d.indices = tree ? P<uint32_t>(tree->indices) : nullptr;
    return d;
}

uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

int bit_width(uint32_t v) {  // bits needed for values 0..v
    int b = 0;
    while (b < 32 && (v >> b) != 0) ++b;
    return b;
}

// Standalone build: one segment, keys morton30 << 32 | index (or morton60).
Job single_job(const SegDesc& seg, bool m60) {
    const uint32_t n = seg.n;
    Job j;
    j.fsb.segs = nullptr;
    j.fsb.blocks = nullptr;
    j.fsb.single = seg;
    j.fsb.chunk = n >= kFrontLargeMin ? kFrontChunkLarge : kFrontChunk;
    j.fsb.nblocks = cdiv(n, j.fsb.chunk);
    j.hsb = j.fsb;
    j.hsb.nblocks = cdiv(n, (uint32_t)hier_chunk());
    j.k = 1;
    j.total = n;
    j.maxn = n;
    j.m60 = m60;
    j.kf = KeyFormat{32, 62};
    j.begin_bit = m60 ? 0 : 32;
    j.end_bit = m60 ? 60 : 62;
    return j;
}

Job single_job(const bvh_mesh_rec* mesh, bvh_tree_rec* tree, bool m60) { return single_job(seg_of(mesh, tree, 0), m60); }

SortPlan plan_for(const Job& j, int radix_bits) {
    return make_sort_plan(j.total, j.begin_bit, j.end_bit, pick_bits(radix_bits), (int)j.fsb.nblocks, j.m60);
}

size_t zero_bytes_of(const SortPlan& plan) { return (sort_zero_bytes(plan) + 15) & ~(size_t)15; }

// Allocations of the front end and the hierarchy (outside any capture).
int reserve_front(bvh_ctx_t ctx, Workspace& w, const Job& j, const SortPlan& plan, bool hier) {
    const size_t n = j.total;
    const size_t bins = 1u << plan.bits;
    TRY(ensure(ctx, w.tri_boxes, n * 32));
    TRY(ensure(ctx, w.keys_a, n * 8));
    TRY(ensure(ctx, w.keys_b, n * 8));
    if (j.m60) {
        TRY(ensure(ctx, w.vals_a, n * 4));
        TRY(ensure(ctx, w.vals_b, n * 4));
    }
    TRY(ensure(ctx, w.hist_tot, (size_t)plan.passes * bins * 4));
    TRY(ensure(ctx, w.lookback, zero_bytes_of(plan)));
    TRY(ensure(ctx, w.scene, (size_t)j.k * 32));
    TRY(ensure(ctx, w.handoff, n * 8, true));
    TRY(ensure(ctx, w.epoch, 16, true));
    TRY(ensure(ctx, w.hctl, (size_t)2 * hier_max_levels() * 4));
    if (hier) {
        TRY(ensure(ctx, w.hunits[0], (size_t)j.hsb.nblocks * hier_chunk() * hier_unit_bytes()));
        TRY(ensure(ctx, w.hunits[1], n * (size_t)hier_unit_bytes()));
        TRY(ensure(ctx, w.hcounts, (size_t)j.hsb.nblocks * 4));
        TRY(ensure(ctx, w.hstatus, (size_t)hier_max_levels() * (j.hsb.nblocks + 1) * 8, true));
    }
    return BVH_OK;
}

// Bounds, Morton keys + histograms, and the LSD passes.  Result in *keys/*vals.
int enqueue_front(bvh_ctx_t ctx, Workspace& w, const Job& j, const SortPlan& plan, StageClock* clk,
                  uint64_t** keys_out, uint32_t** vals_out) {
    hipStream_t s = ctx->stream;
    launch_init_build(P<uint32_t>(w.scene), j.k, P<uint32_t>(w.epoch), s, P<uint32_t>(w.hist_tot),
                      (uint32_t)plan.passes << plan.bits, P<uint32_t>(w.hctl), 2u * hier_max_levels());
    launch_tri_bounds(j.fsb, P<float4>(w.tri_boxes), P<uint32_t>(w.scene), s);
    TRY(check_launch(ctx, "k_tri_bounds"));
    TRY(mark(clk, "Bounds/Centroids"));
    launch_morton(j.fsb, P<float4>(w.tri_boxes), P<uint32_t>(w.scene), j.m60, j.kf, P<uint64_t>(w.keys_a),
                  j.m60 ? P<uint32_t>(w.vals_a) : nullptr, plan, P<uint32_t>(w.hist_tot), P<uint32_t>(w.lookback),
                  zero_bytes_of(plan) / 4, s);
    TRY(check_launch(ctx, "k_morton"));
    TRY(mark(clk, "Morton Codes"));
    static const char* pass_names[8] = {"Sort Pass 1", "Sort Pass 2", "Sort Pass 3", "Sort Pass 4",
                                        "Sort Pass 5", "Sort Pass 6", "Sort Pass 7", "Sort Pass 8"};
    uint64_t* kin = P<uint64_t>(w.keys_a);
    uint64_t* kout = P<uint64_t>(w.keys_b);
    uint32_t* vin = j.m60 ? P<uint32_t>(w.vals_a) : nullptr;
    uint32_t* vout = j.m60 ? P<uint32_t>(w.vals_b) : nullptr;
    for (int p = 0; p < plan.passes; ++p) {
        if (!launch_onesweep_pass(plan, p, j.total, kin, kout, vin, vout, P<uint32_t>(w.hist_tot),
                                  P<uint32_t>(w.lookback), s))
            return set_err(ctx, BVH_E_ARG, "sort plan made for keys only");
        TRY(check_launch(ctx, "k_onesweep"));
        TRY(mark(clk, pass_names[p]));
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    *keys_out = kin;
    *vals_out = vin;
    return BVH_OK;
}

HierArgs hier_args(Workspace& w, const Job& j, const uint64_t* keys, const uint32_t* vals, bool links) {
    HierArgs ha;
    ha.sb = j.hsb;
    ha.keys = keys;
    ha.vals = vals;
    ha.kf = j.kf;
    ha.tri_boxes = P<float4>(w.tri_boxes);
    ha.parents = links ? P<uint32_t>(w.parents) : nullptr;
    ha.handoff = P<uint64_t>(w.handoff);
    ha.epoch = P<uint32_t>(w.epoch);
    ha.err = P<uint32_t>(w.scene) + 7;
    ha.units[0] = P<HUnit>(w.hunits[0]);
    ha.units[1] = P<HUnit>(w.hunits[1]);
    ha.units_in = nullptr;
    ha.units_out = nullptr;
    ha.counts = P<uint32_t>(w.hcounts);
    ha.status = P<uint64_t>(w.hstatus);
    ha.status_stride = j.hsb.nblocks + 1;
    ha.ctl = P<uint32_t>(w.hctl);
    ha.levels = hier_levels(j.maxn, j.k);
    return ha;
}

// Status of a single build from its flag words (non-finite input, internal error).
int single_flags_status(bvh_ctx_t ctx, const uint32_t* sc) {
    if (sc[6]) return set_err(ctx, BVH_E_NONFINITE, "mesh contains NaN or infinite vertex coordinates");
    if (sc[7]) return set_err(ctx, BVH_E_STATE, "internal error: sorted keys name no primitive");
    return BVH_OK;
}

// Status of one mesh's PLOC record (clusters left, nodes made, iterations, flags).
int ploc_record_status(bvh_ctx_t ctx, const uint4& o, uint32_t s, uint32_t n) {
    if (o.w & kPlocUnfinished) return set_err(ctx, BVH_E_STATE, "internal error: PLOC tail ran before mesh %u was small", s);
    if (o.w & kPlocNoProgress)
        return set_err(ctx, BVH_E_NOPROGRESS, "PLOC iteration %u of mesh %u made no merge (%u clusters left)", o.z, s,
                       o.x);
    if (o.x != 1 || o.y != 2 * n - 1)
        return set_err(ctx, BVH_E_STATE, "PLOC of mesh %u ended with %u clusters and %u nodes", s, o.x, o.y);
    return BVH_OK;
}

// Pinned slot of the next asynchronous call (its index is its place in ctx->pending).
int pending_slot(bvh_ctx_t ctx, uint32_t** slot) {
    const size_t idx = ctx->pending.size();
    if (idx >= kMaxPending)
        return set_err(ctx, BVH_E_STATE, "%u asynchronous calls without bvh_sync", (unsigned)kMaxPending);
    const size_t chunk = idx / kPinChunkSlots;
    while (ctx->pin_chunks.size() <= chunk) {
        uint32_t* p = nullptr;
        HIP_TRY(ctx, hipHostMalloc((void**)&p, (size_t)kPinChunkSlots * kPinSlotWords * 4, hipHostMallocDefault));
        ctx->pin_chunks.push_back(p);
    }
    *slot = ctx->pin_chunks[chunk] + (idx % kPinChunkSlots) * kPinSlotWords;
    return BVH_OK;
}

// Enqueues the copy of a single build's status words (and, for PLOC, its
// record) into the call's pinned slot; bvh_sync checks them.
int defer_build_check(bvh_ctx_t ctx, Workspace& w, PendingKind kind, uint32_t n) {
    uint32_t* slot = nullptr;
    TRY(pending_slot(ctx, &slot));
    HIP_TRY(ctx, hipMemcpyAsync(slot, w.scene.p, 8 * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (kind == kPendPloc) HIP_TRY(ctx, hipMemcpyAsync(slot + 8, w.ploc_out.p, 16, hipMemcpyDeviceToHost, ctx->stream));
    ctx->pending.push_back(Pending{kind, n});
    return BVH_OK;
}

// Meshes whose upload found a NaN or infinite coordinate are never built.
int refuse_nonfinite(bvh_ctx_t ctx, const bvh_mesh_t* meshes, uint32_t k) {
    for (uint32_t s = 0; s < k; ++s)
        if (meshes[s]->nonfinite)
            return set_err(ctx, BVH_E_NONFINITE, "mesh %u of the batch contains NaN or infinite vertex coordinates", s);
    return BVH_OK;
}

// Non-finite input and internal-error words of every segment.
int check_input_flags(bvh_ctx_t ctx, Workspace& w, uint32_t k, uint32_t* host_dst = nullptr) {
    // Single builds read their flag words into a pinned buffer (a pageable
    // destination adds a staged copy to every build); multi-mesh builds,
    // where that copy is small beside the build, use a pageable one.  A
    // caller-provided destination (8k words) is used as it is.
    const size_t words = (size_t)k * 8;
    std::vector<uint32_t> pageable;
    uint32_t* dst = nullptr;
    if (host_dst) {
        dst = host_dst;
    } else if (k == 1) {
        if (!ctx->host_flags) HIP_TRY(ctx, hipHostMalloc((void**)&ctx->host_flags, words * 4, hipHostMallocDefault));
        dst = ctx->host_flags;
    } else {
        pageable.resize(words);
        dst = pageable.data();
    }
    const uint32_t* sc = dst;
    HIP_TRY(ctx, hipMemcpyAsync(dst, w.scene.p, words * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (uint32_t s = 0; s < k; ++s) {
        if (sc[8 * s + 6]) {
            if (k == 1) return set_err(ctx, BVH_E_NONFINITE, "mesh contains NaN or infinite vertex coordinates");
            return set_err(ctx, BVH_E_NONFINITE, "mesh %u of the batch contains NaN or infinite vertex coordinates", s);
        }
    }
    if (sc[7]) return set_err(ctx, BVH_E_STATE, "internal error: sorted keys name no primitive");
    return BVH_OK;
}

int build_lbvh(bvh_ctx_t ctx, bvh_mesh_t mesh, const bvh_build_params* prm, bvh_tree_t tree, bvh_stage_times* times,
               bool async) {
    TRY(ensure_tree(ctx, tree, mesh->n));
    Job j = single_job(mesh, tree, false);
    SortPlan plan = plan_for(j, prm->radix_bits);
    Workspace& w = ctx->ws;
    TRY(reserve_front(ctx, w, j, plan, true));
    StageClock clk{ctx, times != nullptr};
    TRY(clk.begin());
    uint64_t* keys = nullptr;
    uint32_t* vals = nullptr;
    TRY(enqueue_front(ctx, w, j, plan, &clk, &keys, &vals));
    launch_hierarchy(hier_args(w, j, keys, nullptr, false), ctx->stream);
    TRY(check_launch(ctx, "k_hierarchy"));
    TRY(clk.mark("Topology+Refit"));
    TRY(clk.finish(times));
    ctx->last_keys = keys;
    ctx->last_vals = nullptr;
    ctx->last_n = mesh->n;
    return async ? defer_build_check(ctx, w, kPendBuild, mesh->n) : check_input_flags(ctx, w, 1);
}

// LBVH+ round schedule: ready lists of at least kTreeletLevelMin nodes take
// one launch each with one treelet per lane, up to kTreeletLevels launches;
// the first shorter list seeds the climbing chains that finish the tree.
// The lists hold each internal node once; the list lengths follow at a
// 256-byte boundary.
constexpr int kTreeletLevels = 16;
constexpr uint32_t kTreeletLevelMin = 8192;
size_t treelet_list_words(uint32_t n) { return ((size_t)n + 63) & ~(size_t)63; }
size_t treelet_list_bytes(uint32_t n) { return (treelet_list_words(n) + kTreeletLevels + 2) * 4; }

int build_lbvh_plus(bvh_ctx_t ctx, bvh_mesh_t mesh, const bvh_build_params* prm, bvh_tree_t tree,
                    bvh_stage_times* times, bool async) {
    const uint32_t n = mesh->n;
    const int rounds = prm->treelet_rounds;
    if (rounds < 0 || rounds > 64) return set_err(ctx, BVH_E_ARG, "treelet_rounds %d out of range [0, 64]", rounds);
    TRY(ensure_tree(ctx, tree, n));
    Job j = single_job(mesh, tree, true);
    SortPlan plan = plan_for(j, prm->radix_bits);
    Workspace& w = ctx->ws;
    TRY(reserve_front(ctx, w, j, plan, true));
    TRY(ensure(ctx, w.parents, (size_t)(2 * (size_t)n) * 4));
    TRY(ensure(ctx, w.tl_list, treelet_list_bytes(n)));
    TRY(ensure(ctx, w.counters, (size_t)n * 4, true));
    StageClock clk{ctx, times != nullptr};
    TRY(clk.begin());
    uint64_t* keys = nullptr;
    uint32_t* vals = nullptr;
    TRY(enqueue_front(ctx, w, j, plan, &clk, &keys, &vals));
    launch_hierarchy(hier_args(w, j, keys, vals, true), ctx->stream);
    TRY(check_launch(ctx, "k_hierarchy"));
    TRY(clk.mark("Topology+Refit"));
    static const char* round_names[8] = {"Treelet Round 1", "Treelet Round 2", "Treelet Round 3",
                                         "Treelet Round 4", "Treelet Round 5", "Treelet Round 6",
                                         "Treelet Round 7", "Treelet Rounds 8+"};
    for (int r = 0; r < rounds && n > 1; ++r) {
        TreeletArgs ta;
        ta.n = n;
        ta.nodes = P<DNode>(tree->nodes);
        ta.parents = P<uint32_t>(w.parents);
        ta.counters = P<uint32_t>(w.counters);
        ta.list = P<uint32_t>(w.tl_list);
        ta.cnt = ta.list + treelet_list_words(n);
        ta.epoch = ++ctx->treelet_epoch;
        ta.level_min = kTreeletLevelMin;
        ta.levels = kTreeletLevels;
        launch_treelet(ta, ctx->stream);
        TRY(check_launch(ctx, "k_treelet"));
        if (r < 7 || r == rounds - 1) TRY(clk.mark(round_names[r < 7 ? r : 7]));
    }
    TRY(clk.finish(times));
    ctx->last_keys = keys;
    ctx->last_vals = vals;
    ctx->last_n = n;
    return async ? defer_build_check(ctx, w, kPendBuild, n) : check_input_flags(ctx, w, 1);
}

// ---- PLOC ---------------------------------------------------------------------

int reserve_ploc(bvh_ctx_t ctx, Workspace& w, const Job& j) {
    const size_t n = j.total;
    for (int b = 0; b < 2; ++b) {
        TRY(ensure(ctx, w.ploc_box[b], n * 32));
        TRY(ensure(ctx, w.ploc_id[b], n * 4));
        TRY(ensure(ctx, w.ploc_seg[b], (size_t)j.k * 16));
        if (j.k > 1) TRY(ensure(ctx, w.ploc_sid[b], n * 4));
    }
    TRY(ensure(ctx, w.ploc_nn, n * 4));
    TRY(ensure(ctx, w.ploc_state, (size_t)kPlocMaxIter * sizeof(PlocState)));
    TRY(ensure(ctx, w.ploc_status, (n / (size_t)ploc_merge_tile() + 2) * 8 + 64));
    TRY(ensure(ctx, w.ploc_out, (size_t)j.k * 16));
    return BVH_OK;
}

PlocArgs ploc_args(Workspace& w, const Job& j, int radius) {
    PlocArgs a;
    a.sb = j.fsb;
    a.nseg = j.k;
    a.radius = radius;
    a.m0 = j.total;
    a.maxm0 = j.maxn;
    for (int b = 0; b < 2; ++b) {
        a.box[b] = P<float4>(w.ploc_box[b]);
        a.id[b] = P<uint32_t>(w.ploc_id[b]);
        a.sid[b] = j.k > 1 ? P<uint32_t>(w.ploc_sid[b]) : nullptr;
        a.seg[b] = P<uint4>(w.ploc_seg[b]);
    }
    a.nn = P<int32_t>(w.ploc_nn);
    a.state = P<PlocState>(w.ploc_state);
    a.status = P<uint64_t>(w.ploc_status);
    a.tile_ctr = reinterpret_cast<uint32_t*>(a.status + (j.total / (uint32_t)ploc_merge_tile() + 2));
    a.out = P<uint4>(w.ploc_out);
    return a;
}

// Ray sets from this size on are traced in coherence order.
constexpr uint32_t kRaySortMin = 16384;

// Most global PLOC iterations captured into a batch graph (more run eagerly).
constexpr int kPlocCaptureMax = 512;

// Global iterations a mesh of m clusters needs before it is small (measured
// decay of the cluster count is 0.72-0.87 per iteration).
int ploc_iter_estimate(uint32_t m, double ratio) {
    if (m <= (uint32_t)kPlocSmall) return 0;
    return (int)std::ceil(std::log((double)m / kPlocSmall) / std::log(1.0 / ratio));
}

// Host-driven global iterations from ring slot *slot until every mesh is
// small (or a flag is raised): batches of launches, one 32-byte state read
// per batch.  m_bound bounds the clusters entering slot *slot.
int ploc_drive(bvh_ctx_t ctx, const PlocArgs& a, int* slot, uint32_t m_bound, uint32_t maxm, uint32_t nn_grid,
               uint32_t merge_grid, PlocState* last) {
    int k = *slot;
    PlocState hs;
    std::memset(&hs, 0, sizeof(hs));
    hs.m = m_bound;
    hs.maxm = maxm;
    // The first batch assumes the fastest measured decay; later ones the
    // decay the previous batch showed (fewer state reads: each is a host
    // round trip, while surplus iterations exit at once).
    double ratio = 0.80;
    while (hs.maxm > (uint32_t)kPlocSmall && hs.flags == 0) {
        int batch = ploc_iter_estimate(hs.maxm, ratio);
        batch = batch < 2 ? 2 : (batch > 48 ? 48 : batch);
        const uint32_t before = hs.maxm;
        for (int b = 0; b < batch; ++b, k = (k + 1) % kPlocMaxIter) {
            launch_ploc_nn(a, k, hs.m, nn_grid, ctx->stream);
            launch_ploc_merge(a, k, hs.m, merge_grid, ctx->stream);
        }
        TRY(check_launch(ctx, "k_ploc_nn/merge"));
        HIP_TRY(ctx, hipMemcpyAsync(&hs, a.state + k, sizeof(hs), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        if (hs.maxm > (uint32_t)kPlocSmall && hs.maxm < before) {
            const double r = std::pow((double)hs.maxm / (double)before, 1.0 / batch);
            ratio = r < 0.5 ? 0.5 : (r > 0.97 ? 0.97 : r);
        }
    }
    *slot = k;
    if (last) *last = hs;
    return BVH_OK;
}

// Checks the per-mesh records the tail wrote.  *unfinished is set when the
// tail ran before every mesh was small.
int ploc_check(bvh_ctx_t ctx, Workspace& w, const Job& j, const std::vector<uint32_t>& ns, bool* unfinished,
               uint32_t* max_iters) {
    std::vector<uint4> out(j.k);
    HIP_TRY(ctx, hipMemcpyAsync(out.data(), w.ploc_out.p, (size_t)j.k * 16, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    *unfinished = false;
    uint32_t mi = 0;
    for (uint32_t s = 0; s < j.k; ++s) {
        const uint4 o = out[s];
        if (o.w & kPlocUnfinished) {
            *unfinished = true;
            continue;
        }
        TRY(ploc_record_status(ctx, o, s, ns[s]));
        mi = std::max(mi, o.z);
    }
    if (max_iters) *max_iters = mi;
    return BVH_OK;
}

// PLOC's global iterations are driven from the host (one 32-byte state read
// per batch of launches), so an asynchronous PLOC build returns once they are
// enqueued and done; the LDS tail and the checks stay deferred.
int build_ploc(bvh_ctx_t ctx, bvh_mesh_t mesh, const bvh_build_params* prm, bvh_tree_t tree, bvh_stage_times* times,
               bool async) {
    const uint32_t n = mesh->n;
    const int radius = prm->ploc_radius;
    if (radius < 1 || radius > 1024) return set_err(ctx, BVH_E_ARG, "PLOC radius %d out of range [1, 1024]", radius);
    TRY(ensure_tree(ctx, tree, n));
    Job j = single_job(mesh, tree, false);
    SortPlan plan = plan_for(j, prm->radix_bits);
    Workspace& w = ctx->ws;
    TRY(reserve_front(ctx, w, j, plan, false));
    TRY(reserve_ploc(ctx, w, j));
    StageClock clk{ctx, times != nullptr};
    TRY(clk.begin());
    uint64_t* keys = nullptr;
    uint32_t* vals = nullptr;
    TRY(enqueue_front(ctx, w, j, plan, &clk, &keys, &vals));
    const PlocArgs a = ploc_args(w, j, radius);
    launch_ploc_init(a, keys, j.kf, P<float4>(w.tri_boxes), P<uint32_t>(w.scene) + 7, ctx->stream);
    TRY(check_launch(ctx, "k_ploc_init"));
    TRY(clk.mark("PLOC Init"));
    int slot = 0;
    TRY(ploc_drive(ctx, a, &slot, n, n, 2048, 1024, nullptr));
    TRY(clk.mark("PLOC Global Iterations"));
    launch_ploc_tail(a, slot, ctx->stream);
    TRY(check_launch(ctx, "k_ploc_tail"));
    TRY(clk.mark("PLOC LDS Tail"));
    TRY(clk.finish(times));
    ctx->last_keys = keys;
    ctx->last_vals = nullptr;
    ctx->last_n = n;
    if (async) return defer_build_check(ctx, w, kPendPloc, n);
    TRY(check_input_flags(ctx, w, 1));
    bool unfinished = false;
    uint32_t iters = 0;
    TRY(ploc_check(ctx, w, j, std::vector<uint32_t>{n}, &unfinished, &iters));
    if (unfinished) return set_err(ctx, BVH_E_STATE, "internal error: PLOC tail ran before the mesh was small");
    if (times) times->ploc_iterations = (int32_t)iters;
    return BVH_OK;
}

// ---- segmented jobs and batches ----------------------------------------------------

// Refreshes a segment table (tree storage may have moved) and uploads it.
int seg_tables(bvh_ctx_t ctx, SegTables& t, const bvh_mesh_t* meshes, const bvh_tree_t* trees, uint32_t k,
               bool force) {
    std::vector<SegDesc> segs(k);
    uint32_t start = 0;
    for (uint32_t s = 0; s < k; ++s) {
        TRY(ensure_tree(ctx, trees[s], meshes[s]->n));
        segs[s] = seg_of(meshes[s], trees[s], start);
        start += meshes[s]->n;
    }
    bool same = !force && t.host_segs.size() == segs.size() &&
                std::memcmp(t.host_segs.data(), segs.data(), segs.size() * sizeof(SegDesc)) == 0;
    if (same) return BVH_OK;
    t.host_segs = segs;
    TRY(ensure(ctx, t.segs, segs.size() * sizeof(SegDesc)));
    HIP_TRY(ctx, hipMemcpyAsync(t.segs.p, segs.data(), segs.size() * sizeof(SegDesc), hipMemcpyHostToDevice,
                                ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BVH_OK;
}

// Job of k meshes: block tables (front-end blocks of front_chunk triangles,
// at least kFrontChunk), segment table and key format.
int plan_job(bvh_ctx_t ctx, SegTables& t, const bvh_mesh_t* meshes, const bvh_tree_t* trees, uint32_t k, Job& j,
             uint32_t front_chunk = kFrontChunk) {
    uint64_t total = 0;
    for (uint32_t s = 0; s < k; ++s) {
        if (!meshes[s] || !trees[s]) return set_err(ctx, BVH_E_ARG, "mesh or tree %u is null", s);
        total += meshes[s]->n;
    }
    if (total >= (1ull << 31))
        return set_err(ctx, BVH_E_UNSUPPORTED, "batch of %llu triangles (limit 2^31 - 1)", (unsigned long long)total);
    std::vector<uint2> fb, hb;
    const uint32_t hc = (uint32_t)hier_chunk();
    j.total = 0;
    j.maxn = 0;
    for (uint32_t s = 0; s < k; ++s) {
        const uint32_t n = meshes[s]->n;
        for (uint32_t c = 0; c < cdiv(n, front_chunk); ++c) fb.push_back(make_uint2(s, c));
        for (uint32_t c = 0; c < cdiv(n, hc); ++c) hb.push_back(make_uint2(s, c));
        j.total += n;
        j.maxn = std::max(j.maxn, n);
    }
    const int idx_bits = std::max(1, bit_width(j.maxn - 1));
    const int seg_bits = bit_width(k - 1);
    if (idx_bits + 30 + seg_bits > 64)
        return set_err(ctx, BVH_E_UNSUPPORTED, "batch keys need %d bits (> 64): too many or too large meshes",
                       idx_bits + 30 + seg_bits);
    TRY(ensure(ctx, t.fblocks, fb.size() * sizeof(uint2)));
    TRY(ensure(ctx, t.hblocks, hb.size() * sizeof(uint2)));
    HIP_TRY(ctx, hipMemcpyAsync(t.fblocks.p, fb.data(), fb.size() * sizeof(uint2), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(t.hblocks.p, hb.data(), hb.size() * sizeof(uint2), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    TRY(seg_tables(ctx, t, meshes, trees, k, true));
    j.k = k;
    j.m60 = false;
    j.fsb.segs = P<SegDesc>(t.segs);
    j.fsb.blocks = P<uint2>(t.fblocks);
    std::memset(&j.fsb.single, 0, sizeof(j.fsb.single));
    j.fsb.nblocks = (uint32_t)fb.size();
    j.fsb.chunk = front_chunk;
    j.hsb = j.fsb;
    j.hsb.blocks = P<uint2>(t.hblocks);
    j.hsb.nblocks = (uint32_t)hb.size();
    j.kf = KeyFormat{(uint32_t)idx_bits, (uint32_t)(idx_bits + 30)};
    j.begin_bit = idx_bits;
    j.end_bit = idx_bits + 30 + seg_bits;
    return BVH_OK;
}

std::vector<uint32_t> mesh_sizes(const bvh_mesh_t* meshes, uint32_t k) {
    std::vector<uint32_t> ns(k);
    for (uint32_t s = 0; s < k; ++s) ns[s] = meshes[s]->n;
    return ns;
}

// Eager segmented build on the context stream (same kernels as a batch graph).
int build_batched_eager(bvh_ctx_t ctx, const bvh_mesh_t* meshes, const bvh_tree_t* trees, uint32_t k,
                        const bvh_build_params* prm, bvh_stage_times* times) {
    Job j;
    TRY(plan_job(ctx, ctx->tabs, meshes, trees, k, j));
    SortPlan plan = plan_for(j, prm->radix_bits);
    if (plan.passes > 8) return set_err(ctx, BVH_E_UNSUPPORTED, "batch keys need more than 8 sort passes");
    Workspace& w = ctx->ws;
    TRY(reserve_front(ctx, w, j, plan, prm->algo == BVH_ALGO_LBVH));
    if (prm->algo == BVH_ALGO_PLOC) TRY(reserve_ploc(ctx, w, j));
    StageClock clk{ctx, times != nullptr};
    TRY(clk.begin());
    uint64_t* keys = nullptr;
    uint32_t* vals = nullptr;
    TRY(enqueue_front(ctx, w, j, plan, &clk, &keys, &vals));
    ctx->last_keys = nullptr;
    if (prm->algo == BVH_ALGO_LBVH) {
        launch_hierarchy(hier_args(w, j, keys, nullptr, false), ctx->stream);
        TRY(check_launch(ctx, "k_hierarchy"));
        TRY(clk.mark("Topology+Refit"));
        TRY(clk.finish(times));
        return check_input_flags(ctx, w, k);
    }
    const PlocArgs a = ploc_args(w, j, prm->ploc_radius);
    launch_ploc_init(a, keys, j.kf, P<float4>(w.tri_boxes), P<uint32_t>(w.scene) + 7, ctx->stream);
    TRY(check_launch(ctx, "k_ploc_init"));
    TRY(clk.mark("PLOC Init"));
    int slot = 0;
    TRY(ploc_drive(ctx, a, &slot, j.total, j.maxn, 2048, 1024, nullptr));
    TRY(clk.mark("PLOC Global Iterations"));
    launch_ploc_tail(a, slot, ctx->stream);
    TRY(check_launch(ctx, "k_ploc_tail"));
    TRY(clk.mark("PLOC LDS Tail"));
    TRY(clk.finish(times));
    TRY(check_input_flags(ctx, w, k));
    bool unfinished = false;
    uint32_t iters = 0;
    TRY(ploc_check(ctx, w, j, mesh_sizes(meshes, k), &unfinished, &iters));
    if (unfinished) return set_err(ctx, BVH_E_STATE, "internal error: PLOC tail ran before every mesh was small");
    if (times) times->ploc_iterations = (int32_t)iters;
    return BVH_OK;
}

int batch_tables(bvh_batch_t b, bool force) {
    return seg_tables(b->ctx, b->tabs, b->meshes.data(), b->trees.data(), b->k, force);
}

int batch_plan(bvh_batch_t b) {
    return plan_job(b->ctx, b->tabs, b->meshes.data(), b->trees.data(), b->k, b->job, b->front_chunk);
}

void batch_destroy_graph(bvh_batch_t b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Enqueues the whole batch (captured into the graph).
int batch_enqueue(bvh_batch_t b, const SortPlan& plan) {
    bvh_ctx_t ctx = b->ctx;
    Workspace& w = b->ws;
    const Job& j = b->job;
    uint64_t* keys = nullptr;
    uint32_t* vals = nullptr;
    TRY(enqueue_front(ctx, w, j, plan, nullptr, &keys, &vals));
    if (b->algo == BVH_ALGO_LBVH) {
        launch_hierarchy(hier_args(w, j, keys, nullptr, false), ctx->stream);
        return check_launch(ctx, "k_hierarchy");
    }
    const PlocArgs a = ploc_args(w, j, b->radius);
    launch_ploc_init(a, keys, j.kf, P<float4>(w.tri_boxes), P<uint32_t>(w.scene) + 7, ctx->stream);
    for (int k = 0; k < b->ploc_iters; ++k) {
        launch_ploc_nn(a, k % kPlocMaxIter, j.total, b->nn_grid, ctx->stream);
        launch_ploc_merge(a, k % kPlocMaxIter, j.total, b->merge_grid, ctx->stream);
    }
    launch_ploc_tail(a, b->ploc_iters % kPlocMaxIter, ctx->stream);
    return check_launch(ctx, "k_ploc");
}

int batch_capture(bvh_batch_t b) {
    bvh_ctx_t ctx = b->ctx;
    batch_destroy_graph(b);
    // Captured graphs keep the 8192-key sort tile (the wide tile has not run
    // inside a captured graph).
    SortPlan plan = make_sort_plan(b->job.total, b->job.begin_bit, b->job.end_bit, b->radix_bits,
                                   (int)b->job.fsb.nblocks, false, false, b->sort_block);
    if (plan.passes > 8) return set_err(ctx, BVH_E_UNSUPPORTED, "batch keys need more than 8 sort passes");
    TRY(reserve_front(ctx, b->ws, b->job, plan, b->algo == BVH_ALGO_LBVH));
    if (b->algo == BVH_ALGO_PLOC) TRY(reserve_ploc(ctx, b->ws, b->job));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int st = batch_enqueue(b, plan);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return set_err(ctx, BVH_E_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    b->graph = g;
    HIP_TRY(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    size_t nodes = 0;
    HIP_TRY(ctx, hipGraphGetNodes(b->graph, nullptr, &nodes));
    b->info.graph_nodes = (int32_t)nodes;
    b->info.radix_bits = plan.bits;
    b->info.sort_passes = plan.passes;
    b->info.ploc_captured_iterations = b->algo == BVH_ALGO_PLOC ? b->ploc_iters : 0;
    b->info.ploc_nn_grid = (int32_t)b->nn_grid;
    b->info.ploc_merge_grid = (int32_t)b->merge_grid;
    b->info.sort_block = plan.block;
    b->info.front_chunk = (int32_t)b->job.fsb.chunk;
    return BVH_OK;
}

int batch_launch(bvh_batch_t b, float* ms) {
    bvh_ctx_t ctx = b->ctx;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    TRY(batch_tables(b, false));
    HIP_TRY(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    HIP_TRY(ctx, hipGraphLaunch(b->exec, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev[1]));
    float t = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[1]));
    TRY(check_input_flags(ctx, b->ws, b->k));
    b->info.runs++;
    if (b->algo == BVH_ALGO_PLOC) {
        std::vector<uint32_t> ns(b->k);
        for (uint32_t s = 0; s < b->k; ++s) ns[s] = b->meshes[s]->n;
        const PlocArgs a = ploc_args(b->ws, b->job, b->radius);
        bool unfinished = false;
        TRY(ploc_check(ctx, b->ws, b->job, ns, &unfinished, nullptr));
        PlocState hs;
        HIP_TRY(ctx, hipMemcpyAsync(&hs, a.state + b->ploc_iters % kPlocMaxIter, sizeof(hs), hipMemcpyDeviceToHost,
                                    ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        if (unfinished) {
            // The graph ran out of captured iterations: continue eagerly.
            b->info.eager_continuations++;
            HIP_TRY(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
            int slot = b->ploc_iters % kPlocMaxIter;
            TRY(ploc_drive(ctx, a, &slot, hs.m, hs.maxm, b->nn_grid, b->merge_grid, &hs));
            launch_ploc_tail(a, slot, ctx->stream);
            TRY(check_launch(ctx, "k_ploc_tail"));
            HIP_TRY(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
            HIP_TRY(ctx, hipEventSynchronize(ctx->ev[1]));
            float t2 = 0.0f;
            HIP_TRY(ctx, hipEventElapsedTime(&t2, ctx->ev[0], ctx->ev[1]));
            t += t2;
            TRY(ploc_check(ctx, b->ws, b->job, ns, &unfinished, nullptr));
            if (unfinished) return set_err(ctx, BVH_E_STATE, "internal error: PLOC batch did not finish");
        }
        b->info.ploc_iterations = (int32_t)hs.iters;
    }
    if (ms) *ms = t;
    return BVH_OK;
}

// FNV-1a over every tree of the batch (tuning must not change any bit).
int batch_digest(bvh_batch_t b, uint64_t* h) {
    bvh_ctx_t ctx = b->ctx;
    uint64_t x = 1469598103934665603ull;
    std::vector<unsigned char> buf;
    for (uint32_t s = 0; s < b->k; ++s) {
        bvh_tree_t t = b->trees[s];
        const size_t nb = (size_t)t->num_nodes * sizeof(DNode), ib = (size_t)t->num_prims * 4;
        buf.resize(nb + ib);
        HIP_TRY(ctx, hipMemcpyAsync(buf.data(), t->nodes.p, nb, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(buf.data() + nb, t->indices.p, ib, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        for (unsigned char c : buf) {
            x ^= c;
            x *= 1099511628211ull;
        }
    }
    *h = x;
    return BVH_OK;
}

// Times candidate configurations and keeps the fastest, one knob at a time
// from the defaults (coordinate descent): digit width, onesweep tile, front-end
// chunk, and for PLOC the persistent grids of the neighbour search and of
// merge + compaction.  Every candidate must produce the same trees bit for
// bit (FNV-1a over all trees).
constexpr int kTuneMax = 16;  // candidates recorded in bvh_batch_info

struct TuneCfg {
    int bits;
    int sort_block;
    uint32_t front_chunk, nn_grid, merge_grid;
};

int batch_apply(bvh_batch_t b, const TuneCfg& c) {
    const bool replan = c.front_chunk != b->front_chunk;
    b->radix_bits = c.bits;
    b->sort_block = c.sort_block;
    b->front_chunk = c.front_chunk;
    b->nn_grid = c.nn_grid;
    b->merge_grid = c.merge_grid;
    if (replan) TRY(batch_plan(b));
    return batch_capture(b);
}

int batch_tune(bvh_batch_t b) {
    TuneCfg best{b->radix_bits, b->sort_block, b->front_chunk, b->nn_grid, b->merge_grid};
    float best_ms = 1e30f;
    uint64_t ref = 0;
    int count = 0;
    auto trial = [&](const TuneCfg& c) -> int {
        TRY(batch_apply(b, c));
        float ms[3];
        for (float& m : ms) TRY(batch_launch(b, &m));
        std::sort(ms, ms + 3);
        uint64_t h = 0;
        TRY(batch_digest(b, &h));
        if (count == 0) ref = h;
        else if (h != ref)
            return set_err(b->ctx, BVH_E_STATE,
                           "internal error: tuned configurations disagree (digits %d, tile %d, chunk %u, grids %u/%u)",
                           c.bits, c.sort_block, c.front_chunk, c.nn_grid, c.merge_grid);
        if (count < kTuneMax) {
            b->info.tune_ms[count] = ms[1];
            b->info.tune_radix_bits[count] = c.bits;
            b->info.tune_nn_grid[count] = (int32_t)c.nn_grid;
            b->info.tune_sort_block[count] = c.sort_block;
            b->info.tune_front_chunk[count] = (int32_t)c.front_chunk;
            b->info.tune_merge_grid[count] = (int32_t)c.merge_grid;
        }
        ++count;
        if (ms[1] < best_ms) {
            best_ms = ms[1];
            best = c;
        }
        return BVH_OK;
    };
    TRY(trial(best));
    const bool ploc = b->algo == BVH_ALGO_PLOC;
    for (int knob = 0; knob < (ploc ? 5 : 3); ++knob) {
        const TuneCfg base = best;
        TuneCfg alt = base;
        switch (knob) {
            case 0: alt.bits = base.bits == 8 ? 10 : 8; break;
            case 1: alt.sort_block = base.sort_block == kSortThreads ? kSortThreadsSmall : kSortThreads; break;
            case 2: alt.front_chunk = base.front_chunk == kFrontChunk ? kFrontChunkLarge : kFrontChunk; break;
            case 3: alt.nn_grid = base.nn_grid == 2048 ? 1024 : 2048; break;
            default: alt.merge_grid = base.merge_grid == 1024 ? 512 : 1024; break;
        }
        TRY(trial(alt));
    }
    b->info.tune_count = count;
    return batch_apply(b, best);
}

// ---- host copies through pinned staging ---------------------------------------
// Copies between pageable host memory and the device go through two pinned
// buffers of the context: the calling thread copies one chunk into (or out
// of) a pinned buffer while the DMA engine moves the other, so a copy runs
// at the pinned DMA rate and does not depend on the runtime's pageable-copy
// path (whose copies measured twice as slow once another server had run in
// the same process).  Small copies go straight to hipMemcpyAsync.
constexpr size_t kStageBytes = (size_t)8 << 20;
constexpr size_t kStageMin = (size_t)256 << 10;

int stage_ready(bvh_ctx_t ctx) {
    for (int i = 0; i < 2; ++i) {
        if (!ctx->stage[i]) HIP_TRY(ctx, hipHostMalloc((void**)&ctx->stage[i], kStageBytes, hipHostMallocDefault));
        if (!ctx->stage_ev[i]) HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->stage_ev[i], hipEventDisableTiming));
    }
    return BVH_OK;
}

// Reads `words` (at most 16) 32-bit words from the device after the work
// queued on the context stream, through pinned memory.
int read_words(bvh_ctx_t ctx, uint32_t* dst, const void* src, int words) {
    if (!ctx->pin_small) HIP_TRY(ctx, hipHostMalloc((void**)&ctx->pin_small, 64, hipHostMallocDefault));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->pin_small, src, (size_t)words * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    std::memcpy(dst, ctx->pin_small, (size_t)words * 4);
    return BVH_OK;
}

// Enqueues dst <- src (host) on the context stream; src may be reused as
// soon as this returns.
int stage_h2d(bvh_ctx_t ctx, void* dst, const void* src, size_t bytes) {
    if (bytes < kStageMin) {
        HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        return BVH_OK;
    }
    TRY(stage_ready(ctx));
    int b = 0;
    for (size_t off = 0; off < bytes; off += kStageBytes, b ^= 1) {
        const size_t len = std::min(kStageBytes, bytes - off);
        HIP_TRY(ctx, hipEventSynchronize(ctx->stage_ev[b]));  // the previous DMA out of this buffer is done
        std::memcpy(ctx->stage[b], static_cast<const char*>(src) + off, len);
        HIP_TRY(ctx, hipMemcpyAsync(static_cast<char*>(dst) + off, ctx->stage[b], len, hipMemcpyHostToDevice,
                                    ctx->stream));
        HIP_TRY(ctx, hipEventRecord(ctx->stage_ev[b], ctx->stream));
    }
    return BVH_OK;
}

// dst (host) <- src after the work already queued on the context stream;
// returns when dst holds the data.
int stage_d2h(bvh_ctx_t ctx, void* dst, const void* src, size_t bytes) {
    if (bytes < kStageMin) {
        HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        return BVH_OK;
    }
    TRY(stage_ready(ctx));
    const size_t chunks = (bytes + kStageBytes - 1) / kStageBytes;
    auto issue = [&](size_t c) -> int {
        const size_t off = c * kStageBytes, len = std::min(kStageBytes, bytes - off);
        HIP_TRY(ctx, hipMemcpyAsync(ctx->stage[c & 1], static_cast<const char*>(src) + off, len,
                                    hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipEventRecord(ctx->stage_ev[c & 1], ctx->stream));
        return BVH_OK;
    };
    TRY(issue(0));
    for (size_t c = 0; c < chunks; ++c) {
        if (c + 1 < chunks) TRY(issue(c + 1));  // its buffer was emptied by chunk c - 1's copy below
        HIP_TRY(ctx, hipEventSynchronize(ctx->stage_ev[c & 1]));
        const size_t off = c * kStageBytes, len = std::min(kStageBytes, bytes - off);
        std::memcpy(static_cast<char*>(dst) + off, ctx->stage[c & 1], len);
    }
    return BVH_OK;
}

// Builds the tree's traversal records once per build (complete on return, so
// other contexts may trace the tree afterwards).
int prepare_trace(bvh_ctx_t ctx, bvh_tree_t tree) {
    if (tree->rec_valid) return BVH_OK;
    TRY(tree_on(ctx, tree));
    TRY(ensure(ctx, tree->rec, (size_t)tree->num_nodes * 64));
    launch_trace_layout(P<DNode>(tree->nodes), tree->num_nodes, P<float4>(tree->rec), ctx->stream);
    TRY(check_launch(ctx, "k_trace_layout"));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    tree->rec_valid = true;
    return BVH_OK;
}

// Enqueues one traversal on the context stream (ray ordering, traversal,
// deep re-trace); *ovf_dev = the device words whose second one flags a ray
// that needed more stack than the deep path has.
int trace_enqueue(bvh_ctx_t ctx, bvh_tree_t tree, bvh_mesh_t mesh, const float* d_rays, uint32_t nrays, float* d_t,
                  int32_t* d_prim, int32_t* d_counters, const uint32_t** ovf_dev) {
    int st;
    if ((st = ensure(ctx, ctx->trace_ovf, 64))) return st;
    if ((st = ensure(ctx, ctx->trace_list, (size_t)nrays * 4))) return st;
    if ((st = ensure(ctx, ctx->trace_spill, (size_t)trace_spill_slots(nrays) * trace_spill_entries() * 8))) return st;
    if ((st = prepare_trace(ctx, tree))) return st;
    TraceArgs a;
    a.order = nullptr;
    a.order_mask = 0;
    // Large ray sets are traced in coherence order (a 40-bit Morton code of
    // origin and direction): neighbouring lanes then walk similar paths.
    constexpr int kRayKeyBits = 40;
    const int idx_bits = std::max(1, bit_width(nrays - 1));
    if (nrays >= kRaySortMin && idx_bits + kRayKeyBits <= 64) {
        SortPlan plan = make_sort_plan(nrays, idx_bits, idx_bits + kRayKeyBits, 8);
        Workspace& w = ctx->ws;
        const size_t bins = 1u << plan.bits;
        TRY(ensure(ctx, ctx->ray_keys[0], (size_t)nrays * 8));
        TRY(ensure(ctx, ctx->ray_keys[1], (size_t)nrays * 8));
        TRY(ensure(ctx, w.hist_tot, (size_t)plan.passes * bins * 4));
        TRY(ensure(ctx, w.lookback, zero_bytes_of(plan)));
        launch_ray_keys(d_rays, nrays, P<DNode>(tree->nodes), (uint32_t)idx_bits, P<uint64_t>(ctx->ray_keys[0]),
                        ctx->stream);
        launch_sort_histogram(P<uint64_t>(ctx->ray_keys[0]), nrays, plan, P<uint32_t>(w.hist_tot), ctx->stream);
        launch_fill_zero(P<uint32_t>(w.lookback), zero_bytes_of(plan) / 4, ctx->stream);
        uint64_t* kout = nullptr;
        uint32_t* vout = nullptr;
        if (!launch_onesweep(plan, nrays, P<uint64_t>(ctx->ray_keys[0]), P<uint64_t>(ctx->ray_keys[1]), nullptr,
                             nullptr, P<uint32_t>(w.hist_tot), P<uint32_t>(w.lookback), &kout, &vout, ctx->stream))
            return set_err(ctx, BVH_E_ARG, "sort plan made for keys only");
        TRY(check_launch(ctx, "ray ordering"));
        a.order = kout;
        a.order_mask = (1ull << idx_bits) - 1;
    }
    a.nodes = P<DNode>(tree->nodes);
    a.rec = P<float4>(tree->rec);
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
    a.spill = P<int2>(ctx->trace_spill);
    a.spill_slots = trace_spill_slots(nrays);
    a.nearest = ctx->trace_order == BVH_TRACE_NEAREST ? 1 : 0;
    launch_fill_zero(a.overflow_count, 4, ctx->stream);
    launch_trace(a, ctx->stream);
    launch_trace_deep(a, ctx->stream);
    if ((st = check_launch(ctx, "k_trace"))) return st;
    *ovf_dev = a.overflow_count;
    return BVH_OK;
}

}  // namespace

namespace {

// Capacity (with the growth slack) of every buffer a build or trace would
// allocate: the same reserve functions, run on buffers in measuring mode.
size_t measured(Workspace& w) {
    size_t t = 0;
    w.each([&](DevBuf& b) { t += b.cap; });
    return t;
}

uint64_t footprint_build(int algo, uint32_t ntri, uint32_t meshes) {
    if (ntri == 0 || meshes == 0) return 0;
    Workspace w;
    w.each([](DevBuf& b) { b.measure = true; });
    Job j;
    uint64_t tables = 0;
    if (meshes == 1) {
        SegDesc d;
        std::memset(&d, 0, sizeof(d));
        d.n = ntri;
        j = single_job(d, algo == BVH_ALGO_LBVH_PLUS);
    } else {
        // Upper bounds over every split of ntri triangles into `meshes` meshes.
        j.k = meshes;
        j.total = ntri;
        j.maxn = ntri;
        j.m60 = false;
        j.fsb.nblocks = cdiv(ntri, kFrontChunk) + meshes;
        j.fsb.chunk = kFrontChunk;
        j.hsb = j.fsb;
        j.hsb.nblocks = cdiv(ntri, (uint32_t)hier_chunk()) + meshes;
        const int idx_bits = std::max(1, bit_width(ntri - 1));
        j.kf = KeyFormat{(uint32_t)idx_bits, (uint32_t)(idx_bits + 30)};
        j.begin_bit = idx_bits;
        j.end_bit = std::min(64, idx_bits + 30 + bit_width(meshes - 1));
        DevBuf seg, fb, hb;
        seg.measure = fb.measure = hb.measure = true;
        (void)ensure(nullptr, seg, (size_t)meshes * sizeof(SegDesc));
        (void)ensure(nullptr, fb, (size_t)j.fsb.nblocks * sizeof(uint2));
        (void)ensure(nullptr, hb, (size_t)j.hsb.nblocks * sizeof(uint2));
        tables = seg.cap + fb.cap + hb.cap;
    }
    const SortPlan plan = plan_for(j, 0);
    (void)reserve_front(nullptr, w, j, plan, algo != BVH_ALGO_PLOC);
    if (algo == BVH_ALGO_LBVH_PLUS) {
        (void)ensure(nullptr, w.parents, (size_t)(2 * (size_t)ntri) * 4);
        (void)ensure(nullptr, w.tl_list, treelet_list_bytes(ntri));
        (void)ensure(nullptr, w.counters, (size_t)ntri * 4);
    }
    if (algo == BVH_ALGO_PLOC) (void)reserve_ploc(nullptr, w, j);
    return measured(w) + tables;
}

uint64_t footprint_trace(uint32_t nrays) {
    if (nrays == 0) return 0;
    const size_t rb = (size_t)nrays * 24, tb = (size_t)nrays * 4, cb = (size_t)nrays * 12;
    Workspace w;
    w.each([](DevBuf& b) { b.measure = true; });
    DevBuf io, ovf, list, spill, k0, k1;
    for (DevBuf* b : {&io, &ovf, &list, &spill, &k0, &k1}) b->measure = true;
    (void)ensure(nullptr, io, rb + 2 * tb + cb + 256);
    (void)ensure(nullptr, ovf, 64);
    (void)ensure(nullptr, list, (size_t)nrays * 4);
    (void)ensure(nullptr, spill, (size_t)trace_spill_slots(nrays) * trace_spill_entries() * 8);
    if (nrays >= kRaySortMin) {
        (void)ensure(nullptr, k0, (size_t)nrays * 8);
        (void)ensure(nullptr, k1, (size_t)nrays * 8);
        const SortPlan plan = make_sort_plan(nrays, 0, 64, 8);
        (void)ensure(nullptr, w.hist_tot, (size_t)plan.passes * (1u << plan.bits) * 4);
        (void)ensure(nullptr, w.lookback, zero_bytes_of(plan));
    }
    return io.cap + ovf.cap + list.cap + spill.cap + k0.cap + k1.cap + measured(w);
}

}  // namespace

// ---- no-allocation LBVH build (bvh_build_lbvh_noalloc) ----------------------------
//
// Every buffer an LBVH build of up to BVH_NOALLOC_MAX_TRIANGLES triangles
// needs (triangle boxes, the two key buffers, digit totals, look-back words,
// scene words, handoff words, epoch, hierarchy units and status words) lives
// at a fixed offset of this array, sized for that maximum.  Offsets do not
// depend on the call's triangle count, so a region only ever holds its own
// kind of data: the epoch-tagged handoff and look-back words stay valid for
// calls of any size.  Each device holds its own copy until the process ends.
constexpr size_t kNoAllocScratchBytes = (size_t)40 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kNoAllocMaxDevices = 64;

struct NoAllocDevice {
    void* ptr = nullptr;    // device address of this device's copy of `scratch`
    bool ready = false;     // cleared and laid out
    int status = BVH_OK;    // setup result
    Workspace layout;       // fixed regions inside ptr
};
std::mutex g_noalloc_mu;
NoAllocDevice g_noalloc[kNoAllocMaxDevices];

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Lays the regions of an LBVH build of kNoAllocMax triangles out from base;
// returns the bytes used.
size_t noalloc_layout(Workspace& w, unsigned char* base) {
    SegDesc d;
    std::memset(&d, 0, sizeof(d));
    d.n = BVH_NOALLOC_MAX_TRIANGLES;
    const Job j = single_job(d, false);
    const SortPlan plan = plan_for(j, 0);
    const size_t n = j.total;
    size_t off = 0;
    auto region = [&](DevBuf& b, size_t bytes) {
        b.p = base + off;
        b.cap = align256(bytes);
        b.fixed = true;
        off += b.cap;
    };
    region(w.tri_boxes, n * 32);
    region(w.keys_a, n * 8);
    region(w.keys_b, n * 8);
    region(w.hist_tot, (size_t)plan.passes * (1u << plan.bits) * 4);
    region(w.lookback, zero_bytes_of(plan));
    region(w.scene, 32);
    region(w.handoff, n * 8);
    region(w.epoch, 16);
    region(w.hctl, (size_t)2 * hier_max_levels() * 4);
    region(w.hunits[0], (size_t)j.hsb.nblocks * hier_chunk() * hier_unit_bytes());
    region(w.hunits[1], n * (size_t)hier_unit_bytes());
    region(w.hcounts, (size_t)j.hsb.nblocks * 4);
    region(w.hstatus, (size_t)hier_max_levels() * (j.hsb.nblocks + 1) * 8);
    return off;
}

// Once per device, with that device current: the address of the device's
// copy of `scratch` (the array's name is only a host-side stand-in), a
// zero fill of the whole array (handoff, status and epoch words must start at
// zero) and the fixed layout.
int noalloc_setup(bvh_ctx_t ctx) {
    if (ctx->device >= kNoAllocMaxDevices) return BVH_OK;  // the entry point reports it
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    NoAllocDevice& nd = g_noalloc[ctx->device];
    if (nd.ready || nd.status != BVH_OK) return nd.status;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        nd.status = set_err(ctx, BVH_E_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        return nd.status;
    }
    const size_t used = noalloc_layout(nd.layout, static_cast<unsigned char*>(ptr));
    if (used > kNoAllocScratchBytes) {
        nd.status = set_err(ctx, BVH_E_STATE, "internal error: no-allocation layout needs %zu of %zu bytes", used,
                            kNoAllocScratchBytes);
        return nd.status;
    }
    launch_fill_zero(static_cast<uint32_t*>(ptr), kNoAllocScratchBytes / 4, ctx->stream);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        nd.status = set_err(ctx, BVH_E_HIP, "clearing the no-allocation scratch failed: %s", hipGetErrorString(e));
        return nd.status;
    }
    nd.ptr = ptr;
    nd.ready = true;
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
    // Errors of the no-allocation setup are reported by that entry point.
    (void)noalloc_setup(c);
    *out = c;
    return BVH_OK;
}

int bvh_ctx_destroy(bvh_ctx_t ctx) {
    if (!ctx) return BVH_E_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->ws.each([](DevBuf& b) { free_buf(b); });
    ctx->tabs.release();
    free_buf(ctx->trace_ovf);
    free_buf(ctx->trace_list);
    free_buf(ctx->trace_spill);
    free_buf(ctx->ray_keys[0]);
    free_buf(ctx->ray_keys[1]);
    free_buf(ctx->trace_io);
    if (ctx->host_flags) (void)hipHostFree(ctx->host_flags);
    for (uint32_t* p : ctx->pin_chunks) (void)hipHostFree(p);
    if (ctx->pin_small) (void)hipHostFree(ctx->pin_small);
    for (int i = 0; i < 2; ++i) {
        if (ctx->stage[i]) (void)hipHostFree(ctx->stage[i]);
        if (ctx->stage_ev[i]) (void)hipEventDestroy(ctx->stage_ev[i]);
    }
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return BVH_OK;
}

const char* bvh_last_error(bvh_ctx_t ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int bvh_device_memory(bvh_ctx_t ctx, uint64_t* free_bytes, uint64_t* total_bytes) {
    if (!ctx || !free_bytes || !total_bytes) return BVH_E_ARG;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    size_t f = 0, t = 0;
    HIP_TRY(ctx, hipMemGetInfo(&f, &t));
    // Memory the block pool holds unused is this library's to reuse.
    const uint64_t spare = pool_unused(ctx->device);
    *free_bytes = (f + spare < t) ? f + spare : t;
    *total_bytes = t;
    return BVH_OK;
}

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
    HIP_TRY(ctx, hipSetDevice(ctx->device));  // the mesh lives on the context's device
    bvh_mesh_rec* m = new bvh_mesh_rec();
    m->n = n;
    m->stride = ((size_t)n + 63) & ~(size_t)63;  // 256-byte aligned arrays
    m->device = ctx->device;
    hipError_t e = dev_alloc(ctx->device, reinterpret_cast<void**>(&m->base), m->stride * 9 * sizeof(float),
                             &m->base_cap);
    if (e != hipSuccess) {
        delete m;
        return set_err(ctx, BVH_E_NOMEM, "mesh allocation failed: %s", hipGetErrorString(e));
    }
    for (int k = 0; k < 9; ++k) {
        const int st = stage_h2d(ctx, m->base + k * m->stride, soa[k], (size_t)n * 4);
        if (st != BVH_OK) {
            (void)hipStreamSynchronize(ctx->stream);
            dev_free(ctx->device, m->base);
            delete m;
            return st;
        }
    }
    // The traversal copy is followed by one word that the copy kernel sets
    // when any coordinate is NaN or infinite.
    e = dev_alloc(ctx->device, reinterpret_cast<void**>(&m->tri), (size_t)n * 48 + 64, &m->tri_cap);
    if (e != hipSuccess) {
        dev_free(ctx->device, m->base);
        delete m;
        return set_err(ctx, BVH_E_NOMEM, "mesh allocation failed: %s", hipGetErrorString(e));
    }
    uint32_t* bad = reinterpret_cast<uint32_t*>(m->tri + 3 * (size_t)n);
    launch_fill_zero(bad, 1, ctx->stream);
    launch_tri_aos(m->view(), m->tri, bad, ctx->stream);
    uint32_t hbad = 0;
    if (read_words(ctx, &hbad, bad, 1) != BVH_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        dev_free(ctx->device, m->base);
        dev_free(ctx->device, m->tri);
        delete m;
        return set_err(ctx, BVH_E_HIP, "mesh upload sync failed");
    }
    m->nonfinite = hbad != 0;
    *out = m;
    return BVH_OK;
}

int bvh_mesh_release(bvh_ctx_t ctx, bvh_mesh_t mesh) {
    if (!ctx || !mesh) return BVH_E_ARG;
    if (mesh->device != ctx->device) return set_err(ctx, BVH_E_ARG, "mesh of another device");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    cache_give(ctx->device, mesh->base, mesh->base_cap);
    cache_give(ctx->device, mesh->tri, mesh->tri_cap);
    delete mesh;
    return BVH_OK;
}

int bvh_mesh_bytes(bvh_mesh_t mesh, uint64_t* bytes) {
    if (!mesh || !bytes) return BVH_E_ARG;
    *bytes = mesh->base_cap + mesh->tri_cap;
    return BVH_OK;
}

int bvh_tree_bytes(bvh_tree_t tree, uint64_t* bytes) {
    if (!tree || !bytes) return BVH_E_ARG;
    *bytes = tree->nodes.cap + tree->indices.cap + tree->rec.cap;
    return BVH_OK;
}

int bvh_device_cache_info(int32_t device, uint64_t* bytes, uint32_t* blocks) {
    if (device < 0 || device >= kCacheDevices) return BVH_E_ARG;
    BlockCache& c = block_cache(device);
    std::lock_guard<std::mutex> lk(c.mu);
    if (bytes) *bytes = c.bytes;
    if (blocks) *blocks = (uint32_t)c.blocks.size();
    return BVH_OK;
}

int bvh_device_cache_set_limit(int32_t device, uint64_t bytes, uint64_t* previous) {
    if (device < 0 || device >= kCacheDevices) return BVH_E_ARG;
    BlockCache& c = block_cache(device);
    std::vector<void*> drop;
    {
        std::lock_guard<std::mutex> lk(c.mu);
        if (previous) {
            if (c.limit == 0) {
                size_t fr = 0, tot = 0;
                c.limit = hipMemGetInfo(&fr, &tot) == hipSuccess ? tot / 8 : ((uint64_t)8 << 30);
            }
            *previous = c.limit;
        }
        c.limit = bytes ? bytes : 1;  // 0 would mean "not set": keep nothing instead
        while (c.bytes > c.limit && !c.blocks.empty()) {  // largest blocks first
            auto it = std::prev(c.blocks.end());
            c.bytes -= it->first;
            drop.push_back(it->second);
            c.blocks.erase(it);
        }
    }
    return on_device(device, [&] {
        for (void* p : drop) dev_free(device, p);
    });
}

int bvh_device_cache_trim(int32_t device) {
    if (device < 0 || device >= kCacheDevices) return BVH_E_ARG;
    return on_device(device, [&] { cache_trim(device); });
}

// Freed blocks return to the device's pool at once, so freeing waits for
// every stream of the device (other contexts may still read the object);
// the caller's current device is kept.
static int wait_device(bvh_ctx_t ctx) {
    int prev = 0;
    (void)hipGetDevice(&prev);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const hipError_t e = hipDeviceSynchronize();
    (void)hipSetDevice(prev);
    HIP_TRY(ctx, e);
    return BVH_OK;
}

int bvh_mesh_free(bvh_ctx_t ctx, bvh_mesh_t mesh) {
    if (!ctx || !mesh) return BVH_E_ARG;
    TRY(wait_device(ctx));
    dev_free(ctx->device, mesh->base);
    dev_free(ctx->device, mesh->tri);
    delete mesh;
    return BVH_OK;
}

int bvh_mesh_check(bvh_mesh_t mesh) {
    if (!mesh) return BVH_E_ARG;
    return mesh->nonfinite ? BVH_E_NONFINITE : BVH_OK;
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
    TRY(wait_device(ctx));
    free_buf(tree->nodes);
    free_buf(tree->indices);
    free_buf(tree->rec);
    delete tree;
    return BVH_OK;
}

int bvh_tree_release(bvh_ctx_t ctx, bvh_tree_t tree) {
    if (!ctx || !tree) return BVH_E_ARG;
    if (tree->device >= 0 && tree->device != ctx->device) return set_err(ctx, BVH_E_ARG, "tree of another device");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (DevBuf* b : {&tree->nodes, &tree->indices, &tree->rec}) {
        cache_give(ctx->device, b->p, b->cap);
        b->p = nullptr;
        b->cap = 0;
    }
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
    if ((st = tree_on(ctx, tree))) return st;
    if ((st = ensure(ctx, tree->nodes, (size_t)num_nodes * sizeof(DNode)))) return st;
    HIP_TRY(ctx, hipMemcpyAsync(tree->nodes.p, nodes, (size_t)num_nodes * sizeof(DNode), hipMemcpyHostToDevice,
                                ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    tree->num_nodes = num_nodes;
    tree->num_prims = (num_nodes + 1) / 2;
    tree->rec_valid = false;
    return BVH_OK;
}

int bvh_tree_set_indices(bvh_ctx_t ctx, bvh_tree_t tree, const uint32_t* indices, uint32_t num_prims) {
    if (!ctx || !tree || !indices || num_prims == 0) return BVH_E_ARG;
    if (tree->num_nodes != 2 * num_prims - 1)
        return set_err(ctx, BVH_E_ARG, "%u indices for a tree of %u nodes", num_prims, tree->num_nodes);
    int st;
    if ((st = ensure(ctx, tree->indices, (size_t)num_prims * 4))) return st;
    HIP_TRY(ctx, hipMemcpyAsync(tree->indices.p, indices, (size_t)num_prims * 4, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    tree->num_prims = num_prims;
    return BVH_OK;
}

int bvh_build(bvh_ctx_t ctx, bvh_mesh_t mesh, const bvh_build_params* params, bvh_tree_t tree,
              bvh_stage_times* times) {
    if (!ctx || !mesh || !params || !tree) return ctx ? set_err(ctx, BVH_E_ARG, "null argument") : BVH_E_ARG;
    if (mesh->n == 0) return set_err(ctx, BVH_E_ARG, "mesh has no triangles");
    if (mesh->nonfinite) return set_err(ctx, BVH_E_NONFINITE, "mesh contains NaN or infinite vertex coordinates");
    if (params->radix_bits != 0 && params->radix_bits != 8 && params->radix_bits != 10)
        return set_err(ctx, BVH_E_UNSUPPORTED, "radix_bits must be 0, 8 or 10");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    switch (params->algo) {
        case BVH_ALGO_LBVH: return build_lbvh(ctx, mesh, params, tree, times, false);
        case BVH_ALGO_PLOC: return build_ploc(ctx, mesh, params, tree, times, false);
        case BVH_ALGO_LBVH_PLUS: return build_lbvh_plus(ctx, mesh, params, tree, times, false);
        default: return set_err(ctx, BVH_E_UNSUPPORTED, "algorithm %d not available", params->algo);
    }
}

int bvh_build_async(bvh_ctx_t ctx, bvh_mesh_t mesh, const bvh_build_params* params, bvh_tree_t tree) {
    if (!ctx || !mesh || !params || !tree) return ctx ? set_err(ctx, BVH_E_ARG, "null argument") : BVH_E_ARG;
    if (mesh->n == 0) return set_err(ctx, BVH_E_ARG, "mesh has no triangles");
    if (mesh->nonfinite) return set_err(ctx, BVH_E_NONFINITE, "mesh contains NaN or infinite vertex coordinates");
    if (params->radix_bits != 0 && params->radix_bits != 8 && params->radix_bits != 10)
        return set_err(ctx, BVH_E_UNSUPPORTED, "radix_bits must be 0, 8 or 10");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    switch (params->algo) {
        case BVH_ALGO_LBVH: return build_lbvh(ctx, mesh, params, tree, nullptr, true);
        case BVH_ALGO_PLOC: return build_ploc(ctx, mesh, params, tree, nullptr, true);
        case BVH_ALGO_LBVH_PLUS: return build_lbvh_plus(ctx, mesh, params, tree, nullptr, true);
        default: return set_err(ctx, BVH_E_UNSUPPORTED, "algorithm %d not available", params->algo);
    }
}

int bvh_sync(bvh_ctx_t ctx) {
    if (!ctx) return BVH_E_ARG;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    std::vector<Pending> pending;
    pending.swap(ctx->pending);
    if (e != hipSuccess) return set_err(ctx, BVH_E_HIP, "hipStreamSynchronize failed: %s", hipGetErrorString(e));
    for (size_t i = 0; i < pending.size(); ++i) {
        const uint32_t* sc = ctx->pin_chunks[i / kPinChunkSlots] + (i % kPinChunkSlots) * kPinSlotWords;
        const Pending& p = pending[i];
        int st = BVH_OK;
        if (p.kind == kPendTrace) {
            if (sc[1])
                st = set_err(ctx, BVH_E_STATE,
                             "traversal needed more than 256 stack entries (or the nodes are not a tree)");
        } else {
            st = single_flags_status(ctx, sc);
            if (st == BVH_OK && p.kind == kPendPloc) st = ploc_record_status(ctx, make_uint4(sc[8], sc[9], sc[10], sc[11]), 0, p.n);
        }
        if (st != BVH_OK) return st;  // the first failure in call order
    }
    return BVH_OK;
}

int bvh_build_lbvh_noalloc(bvh_ctx_t ctx, const float* const* d_soa, uint32_t n, void* d_nodes,
                           uint32_t* d_indices) {
    if (!ctx) return BVH_E_ARG;
    if (!d_soa || !d_nodes || !d_indices) return set_err(ctx, BVH_E_ARG, "null argument");
    for (int k = 0; k < 9; ++k)
        if (!d_soa[k]) return set_err(ctx, BVH_E_ARG, "vertex array %d is null", k);
    if (n == 0) return set_err(ctx, BVH_E_ARG, "mesh has no triangles");
    if (n > BVH_NOALLOC_MAX_TRIANGLES)
        return set_err(ctx, BVH_E_ARG, "%u triangles exceed the no-allocation limit of %u", n,
                       (uint32_t)BVH_NOALLOC_MAX_TRIANGLES);
    if (ctx->device >= kNoAllocMaxDevices)
        return set_err(ctx, BVH_E_UNSUPPORTED, "no-allocation builds support devices 0..%d", kNoAllocMaxDevices - 1);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    Workspace w;
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mu);
        const NoAllocDevice& nd = g_noalloc[ctx->device];
        if (!nd.ready)
            return set_err(ctx, nd.status != BVH_OK ? nd.status : BVH_E_STATE,
                           "no-allocation scratch is not set up on device %d", ctx->device);
        w = nd.layout;
    }
    SegDesc d;
    for (int k = 0; k < 9; ++k) d.v[k] = d_soa[k];
    d.n = n;
    d.start = 0;
    d.nodes = static_cast<DNode*>(d_nodes);
    d.indices = d_indices;
    const Job j = single_job(d, false);
    const SortPlan plan = plan_for(j, 0);
    TRY(reserve_front(ctx, w, j, plan, true));  // size checks only: every region is fixed
    uint64_t* keys = nullptr;
    uint32_t* vals = nullptr;
    TRY(enqueue_front(ctx, w, j, plan, nullptr, &keys, &vals));
    launch_hierarchy(hier_args(w, j, keys, nullptr, false), ctx->stream);
    TRY(check_launch(ctx, "k_hierarchy"));
    uint32_t flags[8];
    return check_input_flags(ctx, w, 1, flags);
}

int bvh_footprint_estimate(int32_t algo, uint32_t ntri, uint32_t meshes, uint32_t nrays, bvh_footprint* out) {
    if (!out || (algo != BVH_ALGO_LBVH && algo != BVH_ALGO_LBVH_PLUS && algo != BVH_ALGO_PLOC)) return BVH_E_ARG;
    if (algo == BVH_ALGO_LBVH_PLUS && meshes > 1) return BVH_E_UNSUPPORTED;
    std::memset(out, 0, sizeof(*out));
    if (ntri) {
        const uint64_t stride = ((uint64_t)ntri + 63) & ~(uint64_t)63;
        out->mesh = stride * 9 * 4 + (uint64_t)ntri * 48;
        DevBuf nodes, idx, rec;
        nodes.measure = idx.measure = rec.measure = true;
        (void)ensure(nullptr, nodes, (2 * (size_t)ntri - 1) * sizeof(DNode));
        (void)ensure(nullptr, idx, (size_t)ntri * 4);
        (void)ensure(nullptr, rec, (2 * (size_t)ntri - 1) * 64);
        out->tree = nodes.cap + idx.cap + rec.cap;
    }
    out->build_workspace = footprint_build(algo, ntri, meshes ? meshes : 1);
    out->trace_workspace = footprint_trace(nrays);
    return BVH_OK;
}

int bvh_ctx_held_bytes(bvh_ctx_t ctx, uint64_t* build_workspace, uint64_t* trace_workspace) {
    if (!ctx) return BVH_E_ARG;
    uint64_t b = 0;
    ctx->ws.each([&](DevBuf& d) { b += d.cap; });
    for (DevBuf* d : {&ctx->tabs.segs, &ctx->tabs.fblocks, &ctx->tabs.hblocks}) b += d->cap;
    uint64_t t = 0;
    for (DevBuf* d : {&ctx->trace_ovf, &ctx->trace_list, &ctx->trace_io, &ctx->trace_spill, &ctx->ray_keys[0],
                      &ctx->ray_keys[1]})
        t += d->cap;
    if (build_workspace) *build_workspace = b;
    if (trace_workspace) *trace_workspace = t;
    return BVH_OK;
}

int bvh_debug_morton(bvh_ctx_t ctx, bvh_mesh_t mesh, int32_t algo, uint64_t* codes) {
    if (!ctx || !mesh || !codes) return BVH_E_ARG;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const Job j = single_job(mesh, nullptr, algo == BVH_ALGO_LBVH_PLUS);
    Workspace& w = ctx->ws;
    TRY(ensure(ctx, w.tri_boxes, (size_t)j.total * 32));
    TRY(ensure(ctx, w.keys_a, (size_t)j.total * 8));
    TRY(ensure(ctx, w.scene, 32));
    TRY(ensure(ctx, w.epoch, 16, true));
    launch_init_build(P<uint32_t>(w.scene), 1, P<uint32_t>(w.epoch), ctx->stream);
    launch_tri_bounds(j.fsb, P<float4>(w.tri_boxes), P<uint32_t>(w.scene), ctx->stream);
    launch_morton_codes(j.fsb, P<float4>(w.tri_boxes), P<uint32_t>(w.scene), j.m60, P<uint64_t>(w.keys_a),
                        ctx->stream);
    TRY(check_launch(ctx, "k_morton_codes"));
    HIP_TRY(ctx, hipMemcpyAsync(codes, w.keys_a.p, (size_t)j.total * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->last_keys = nullptr;
    return check_input_flags(ctx, w, 1);
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

int bvh_batch_create(bvh_ctx_t ctx, const bvh_mesh_t* meshes, const bvh_tree_t* trees, uint32_t k,
                     const bvh_build_params* params, bvh_batch_t* out) {
    if (!ctx || !meshes || !trees || !params || !out || k == 0)
        return ctx ? set_err(ctx, BVH_E_ARG, "null argument or empty batch") : BVH_E_ARG;
    *out = nullptr;
    if (params->algo != BVH_ALGO_LBVH && params->algo != BVH_ALGO_PLOC)
        return set_err(ctx, BVH_E_UNSUPPORTED, "batches build LBVH or PLOC (algorithm %d requested)", params->algo);
    if (params->algo == BVH_ALGO_PLOC && (params->ploc_radius < 1 || params->ploc_radius > 1024))
        return set_err(ctx, BVH_E_ARG, "PLOC radius %d out of range [1, 1024]", params->ploc_radius);
    if (params->radix_bits != 0 && params->radix_bits != 8 && params->radix_bits != 10)
        return set_err(ctx, BVH_E_UNSUPPORTED, "radix_bits must be 0, 8 or 10");
    for (uint32_t s = 0; s < k; ++s)
        if (!meshes[s] || !trees[s]) return set_err(ctx, BVH_E_ARG, "mesh or tree %u is null", s);
    TRY(refuse_nonfinite(ctx, meshes, k));
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    bvh_batch_rec* b = new bvh_batch_rec();
    std::memset(&b->info, 0, sizeof(b->info));
    b->ctx = ctx;
    b->algo = params->algo;
    b->radius = params->ploc_radius;
    b->k = k;
    b->meshes.assign(meshes, meshes + k);
    b->trees.assign(trees, trees + k);
    b->radix_bits = pick_bits(params->radix_bits);
    b->nn_grid = 2048;
    b->merge_grid = 1024;
    b->sort_block = kSortThreads;
    b->front_chunk = kFrontChunk;
    auto fail = [&](int st) {
        bvh_batch_free(b);
        return st;
    };
    int st = batch_plan(b);
    if (st) return fail(st);
    b->ploc_iters = b->algo == BVH_ALGO_PLOC ? std::min(ploc_iter_estimate(b->job.maxn, 0.85) + 2, kPlocCaptureMax) : 0;
    if ((st = batch_capture(b))) return fail(st);
    if (b->algo == BVH_ALGO_PLOC) {
        // One run finds the exact number of global iterations these meshes need.
        if ((st = batch_launch(b, nullptr))) return fail(st);
        b->ploc_iters = std::min(b->info.ploc_iterations, kPlocCaptureMax);
        if ((st = batch_capture(b))) return fail(st);
    }
    if ((params->flags & BVH_BUILD_TUNE) && (st = batch_tune(b))) return fail(st);
    b->info.runs = 0;  // counters describe the caller's runs
    b->info.eager_continuations = 0;
    b->info.meshes = k;
    b->info.triangles = b->job.total;
    b->info.max_triangles = b->job.maxn;
    b->info.algo = b->algo;
    b->info.key_bits = b->job.end_bit - b->job.begin_bit;
    *out = b;
    return BVH_OK;
}

int bvh_build_batched(bvh_ctx_t ctx, const bvh_mesh_t* meshes, const bvh_tree_t* trees, uint32_t k,
                      const bvh_build_params* params, bvh_stage_times* times) {
    if (!ctx || !meshes || !trees || !params || k == 0)
        return ctx ? set_err(ctx, BVH_E_ARG, "null argument or empty batch") : BVH_E_ARG;
    if (params->algo != BVH_ALGO_LBVH && params->algo != BVH_ALGO_PLOC)
        return set_err(ctx, BVH_E_UNSUPPORTED, "batches build LBVH or PLOC (algorithm %d requested)", params->algo);
    if (params->algo == BVH_ALGO_PLOC && (params->ploc_radius < 1 || params->ploc_radius > 1024))
        return set_err(ctx, BVH_E_ARG, "PLOC radius %d out of range [1, 1024]", params->ploc_radius);
    if (params->radix_bits != 0 && params->radix_bits != 8 && params->radix_bits != 10)
        return set_err(ctx, BVH_E_UNSUPPORTED, "radix_bits must be 0, 8 or 10");
    for (uint32_t s = 0; s < k; ++s)
        if (!meshes[s] || !trees[s]) return set_err(ctx, BVH_E_ARG, "mesh or tree %u is null", s);
    TRY(refuse_nonfinite(ctx, meshes, k));
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    return build_batched_eager(ctx, meshes, trees, k, params, times);
}

int bvh_batch_run(bvh_batch_t batch, float* ms) {
    if (!batch || !batch->exec) return BVH_E_ARG;
    return batch_launch(batch, ms);
}

int bvh_batch_get_info(bvh_batch_t batch, bvh_batch_info* info) {
    if (!batch || !info) return BVH_E_ARG;
    *info = batch->info;
    return BVH_OK;
}

int bvh_batch_free(bvh_batch_t batch) {
    if (!batch) return BVH_E_ARG;
    bvh_ctx_t ctx = batch->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    batch_destroy_graph(batch);
    batch->ws.each([](DevBuf& b) { free_buf(b); });
    batch->tabs.release();
    delete batch;
    return BVH_OK;
}

int bvh_ctx_set_trace_order(bvh_ctx_t ctx, int32_t order) {
    if (!ctx) return BVH_E_ARG;
    if (order != BVH_TRACE_REFERENCE && order != BVH_TRACE_NEAREST)
        return set_err(ctx, BVH_E_ARG, "unknown trace order");
    ctx->trace_order = order;
    return BVH_OK;
}

int bvh_tree_prepare(bvh_ctx_t ctx, bvh_tree_t tree) {
    if (!ctx || !tree) return BVH_E_ARG;
    if (tree->num_nodes == 0) return set_err(ctx, BVH_E_STATE, "tree is empty");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    return prepare_trace(ctx, tree);
}

int bvh_trace(bvh_ctx_t ctx, bvh_tree_t tree, bvh_mesh_t mesh, const float* d_rays, uint32_t nrays, float* d_t,
              int32_t* d_prim, int32_t* d_counters, float* ms) {
    if (!ctx || !tree || !mesh || (nrays && (!d_rays || !d_t || !d_prim))) return BVH_E_ARG;
    if (tree->num_nodes == 0) return set_err(ctx, BVH_E_STATE, "tree is empty");
    if (nrays == 0) return BVH_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    const uint32_t* ovf_dev = nullptr;
    TRY(trace_enqueue(ctx, tree, mesh, d_rays, nrays, d_t, d_prim, d_counters, &ovf_dev));
    HIP_TRY(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    uint32_t ovf[2] = {0, 0};
    TRY(read_words(ctx, ovf, ovf_dev, 2));
    if (ms) HIP_TRY(ctx, hipEventElapsedTime(ms, ctx->ev[0], ctx->ev[1]));
    if (ovf[1]) return set_err(ctx, BVH_E_STATE, "traversal needed more than 256 stack entries (or the nodes are not a tree)");
    return BVH_OK;
}

int bvh_trace_async(bvh_ctx_t ctx, bvh_tree_t tree, bvh_mesh_t mesh, const float* d_rays, uint32_t nrays, float* d_t,
                    int32_t* d_prim, int32_t* d_counters) {
    if (!ctx || !tree || !mesh || (nrays && (!d_rays || !d_t || !d_prim))) return BVH_E_ARG;
    if (tree->num_nodes == 0) return set_err(ctx, BVH_E_STATE, "tree is empty");
    if (nrays == 0) return BVH_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    uint32_t* slot = nullptr;
    TRY(pending_slot(ctx, &slot));
    const uint32_t* ovf_dev = nullptr;
    TRY(trace_enqueue(ctx, tree, mesh, d_rays, nrays, d_t, d_prim, d_counters, &ovf_dev));
    HIP_TRY(ctx, hipMemcpyAsync(slot, ovf_dev, 8, hipMemcpyDeviceToHost, ctx->stream));
    ctx->pending.push_back(Pending{kPendTrace, nrays});
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
    if ((st = stage_h2d(ctx, d_rays, rays, rb))) return st;
    if ((st = bvh_trace(ctx, tree, mesh, d_rays, nrays, d_t, d_prim, d_cnt, ms))) return st;
    if ((st = stage_d2h(ctx, t, d_t, tb))) return st;
    if ((st = stage_d2h(ctx, prim, d_prim, tb))) return st;
    if (counters && (st = stage_d2h(ctx, counters, d_cnt, cb))) return st;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BVH_OK;
}

int bvh_host_alloc_pinned(size_t bytes, void** ptr) {
    if (!ptr || bytes == 0) return BVH_E_ARG;
    *ptr = nullptr;
    return hipHostMalloc(ptr, bytes, hipHostMallocDefault) == hipSuccess ? BVH_OK : BVH_E_NOMEM;
}

int bvh_host_free_pinned(void* ptr) {
    if (!ptr) return BVH_E_ARG;
    return hipHostFree(ptr) == hipSuccess ? BVH_OK : BVH_E_HIP;
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
    SortPlan plan = make_sort_plan(n, begin_bit, end_bit, pick_bits(radix_bits), 0, d_vals != nullptr);
    if (plan.passes > 8) return set_err(ctx, BVH_E_ARG, "too many passes");
    const size_t bins = 1u << plan.bits;
    const size_t zero_bytes = zero_bytes_of(plan);
    Workspace& w = ctx->ws;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    TRY(ensure(ctx, w.keys_b, (size_t)n * 8));
    if (d_vals) TRY(ensure(ctx, w.vals_b, (size_t)n * 4));
    TRY(ensure(ctx, w.hist_tot, (size_t)plan.passes * bins * 4));
    TRY(ensure(ctx, w.lookback, zero_bytes));
    HIP_TRY(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    launch_sort_histogram(d_keys, n, plan, P<uint32_t>(w.hist_tot), ctx->stream);
    launch_fill_zero(P<uint32_t>(w.lookback), zero_bytes / 4, ctx->stream);
    uint64_t* kout = nullptr;
    uint32_t* vout = nullptr;
    if (!launch_onesweep(plan, n, d_keys, P<uint64_t>(w.keys_b), d_vals, d_vals ? P<uint32_t>(w.vals_b) : nullptr,
                         P<uint32_t>(w.hist_tot), P<uint32_t>(w.lookback), &kout, &vout, ctx->stream))
        return set_err(ctx, BVH_E_ARG, "sort plan made for keys only");
    TRY(check_launch(ctx, "k_onesweep"));
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
