// k_ploc.hip — PLOC (Meister & Bittner 2017) on gfx950.
//
// Replaces kGatherAABBs/kInitLeafNodes_ploc/kInitializeSequence,
// kFindNearestNeighbors_ploc, kMergeClusters_ploc, cub::DeviceScan +
// kCompactIndices/kCompactAABBs, the per-iteration host round trips and the
// host root swap of the reference (src/cuda/ploc_builder.cu:143-293,403-586).
//
// Semantics kept from the reference:
//   * leaves are nodes 0..n-1 in Morton order, internal nodes are numbered
//     from n in creation order, and the root (last created) is swapped with
//     node 0 (ploc_builder.cu:563-582) — applied here as an index remap at
//     write time, so no host pass is needed;
//   * nearest neighbour of cluster i = argmin over j in [i-r, i+r], j != i,
//     of SA(box_i U box_j), ties to the lowest j, non-finite distances never
//     chosen (the strict '<' scan of ploc_builder.cu:190-205);
//   * mutual pairs merge into the lower slot, the higher slot is dropped.
// Deliberate difference: new internal nodes are numbered by rank of the
// merging slot (ascending) within each iteration, one legal serialisation of
// the reference's nondeterministic atomicAdd (ploc_builder.cu:245).
//
// MI355X mapping:
//   * k_ploc_nn: a workgroup stages its 256 clusters plus an r-halo on each
//     side in LDS (SoA floats, conflict-free), every lane scans its window
//     from LDS and packs (distance bits << 32 | j) into one u64 minimum;
//   * k_ploc_merge: per tile of 1024 clusters, mutual-pair detection, new
//     nodes, and stream compaction of the survivors behind a single-pass
//     decoupled look-back scan of packed (merges, survivors) counts — the
//     reference's scan + 2 D2H copies + 2 compaction kernels in one pass;
//   * loop control lives on the device: iteration k reads state[k] and its
//     last tile writes state[k+1] (cluster count, node base, buffer parity;
//     a ring of kPlocMaxIter slots), so the host launches batches of
//     iterations and reads one 32-byte state per batch;
//   * once <= kPlocSmall clusters remain, one workgroup finishes every
//     remaining iteration inside LDS (k_ploc_tail).
#include "dev_common.hpp"
#include "engine.hpp"

namespace bvhk {

constexpr int kNnBlock = 256;
constexpr int kMergeBlock = 256;
constexpr int kMergeItems = 4;
constexpr int kMergeTile = kMergeBlock * kMergeItems;
constexpr int kTailBlock = 1024;
constexpr uint32_t kFlagNoProgress = 1u;
constexpr uint64_t kStA = 1ull << 62;
constexpr uint64_t kStP = 2ull << 62;
constexpr uint64_t kStFlags = 3ull << 62;

int ploc_small() { return kPlocSmall; }
int ploc_nn_block() { return kNnBlock; }
int ploc_merge_tile() { return kMergeTile; }

__device__ __forceinline__ uint32_t remap_id(uint32_t id, uint32_t root) {
    return id == 0 ? root : (id == root ? 0u : id);
}

// Packed neighbour candidate; non-finite distances are never candidates.
__device__ __forceinline__ uint64_t nn_key(float dist, uint32_t j) {
    uint32_t bits = __float_as_uint(dist);
    return ((bits & 0x7f800000u) == 0x7f800000u) ? ~0ull : (((uint64_t)bits << 32) | j);
}

__global__ void k_ploc_init(PlocArgs a, const uint64_t* __restrict__ keys, const float4* __restrict__ tri_boxes,
                            uint32_t* __restrict__ indices) {
    const uint32_t n = a.n;
    const uint32_t root = 2 * n - 2;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t prim = (uint32_t)keys[i];
        if (prim >= n) {
            atomicOr(&a.state[0].flags, 2u);
            prim = 0;
        }
        Box b = load_box(tri_boxes + 2 * (size_t)prim);
        store_node(a.nodes, n > 1 ? remap_id(i, root) : i, b, (int32_t)(prim | 0x80000000u), -1);
        indices[i] = prim;
        store_box(a.box[0] + 2 * (size_t)i, b);
        a.id[0][i] = i;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        PlocState s;
        s.m = n;
        s.node_base = n;
        s.parity = 0;
        s.iters = 0;
        s.flags = 0;
        s.pad[0] = s.pad[1] = s.pad[2] = 0;
        a.state[0] = s;
    }
}

__global__ __launch_bounds__(kNnBlock) void k_ploc_nn(PlocArgs a, int k) {
    extern __shared__ float s_nn[];  // 6 SoA arrays of (kNnBlock + 2r) floats
    const PlocState st = a.state[k];
    const uint32_t m = st.m;
    if (m <= (uint32_t)kPlocSmall || st.flags) return;
    const uint32_t tiles = (m + kNnBlock - 1) / kNnBlock;
    if (blockIdx.x >= tiles) return;
    const int r = a.radius;
    const int span = kNnBlock + 2 * r;
    if (threadIdx.x == 0) {
        a.status[blockIdx.x] = 0;  // look-back word of the merge tile with this index
        if (blockIdx.x == 0) *a.tile_ctr = 0;
    }
    const float4* __restrict__ bx = a.box[st.parity];
    const int64_t t0 = (int64_t)blockIdx.x * kNnBlock;
    for (int q = threadIdx.x; q < span; q += kNnBlock) {
        int64_t p = t0 - r + q;
        if (p >= 0 && p < (int64_t)m) {
            Box b = load_box(bx + 2 * (size_t)p);
            s_nn[0 * span + q] = b.mnx;
            s_nn[1 * span + q] = b.mny;
            s_nn[2 * span + q] = b.mnz;
            s_nn[3 * span + q] = b.mxx;
            s_nn[4 * span + q] = b.mxy;
            s_nn[5 * span + q] = b.mxz;
        }
    }
    __syncthreads();
    const int64_t i = t0 + threadIdx.x;
    if (i >= (int64_t)m) return;
    const int qi = threadIdx.x + r;
    Box bi;
    bi.mnx = s_nn[0 * span + qi]; bi.mny = s_nn[1 * span + qi]; bi.mnz = s_nn[2 * span + qi];
    bi.mxx = s_nn[3 * span + qi]; bi.mxy = s_nn[4 * span + qi]; bi.mxz = s_nn[5 * span + qi];
    const int64_t lo = (i - r > 0) ? i - r : 0;
    const int64_t hi = (i + r < (int64_t)m - 1) ? i + r : (int64_t)m - 1;
    uint64_t best = ~0ull;
    for (int64_t j = lo; j <= hi; ++j) {
        if (j == i) continue;
        const int q = (int)(j - t0) + r;
        Box bj;
        bj.mnx = s_nn[0 * span + q]; bj.mny = s_nn[1 * span + q]; bj.mnz = s_nn[2 * span + q];
        bj.mxx = s_nn[3 * span + q]; bj.mxy = s_nn[4 * span + q]; bj.mxz = s_nn[5 * span + q];
        uint64_t key = nn_key(area(merge(bi, bj)), (uint32_t)j);
        best = key < best ? key : best;
    }
    a.nn[i] = (best == ~0ull) ? -1 : (int32_t)(uint32_t)best;
}

__device__ __forceinline__ uint64_t wave_incl_sum64(uint64_t v) {
    uint32_t l = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint64_t w = __shfl_up(v, o, 64);
        if (l >= (uint32_t)o) v += w;
    }
    return v;
}

// Exclusive block scan of packed (merges << 31 | survivors) counts.
template <int BLOCK>
__device__ __forceinline__ uint64_t block_excl_scan64(uint64_t v, uint64_t* s_w, uint64_t* total) {
    constexpr int WAVES = BLOCK / kWave;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    uint64_t incl = wave_incl_sum64(v);
    if (lane == kWave - 1) s_w[wave] = incl;
    __syncthreads();
    uint64_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
        uint64_t x = s_w[w];
        before += (w < wave) ? x : 0ull;
        all += x;
    }
    __syncthreads();
    *total = all;
    return before + incl - v;
}

__global__ __launch_bounds__(kMergeBlock) void k_ploc_merge(PlocArgs a, int k) {
    __shared__ uint32_t s_tile;
    __shared__ uint64_t s_w[kMergeBlock / kWave];
    __shared__ uint64_t s_prefix;
    const PlocState st = a.state[k];
    const uint32_t m = st.m;
    if (m <= (uint32_t)kPlocSmall || st.flags) {
        if (blockIdx.x == 0 && threadIdx.x == 0) a.state[(k + 1) % kPlocMaxIter] = st;
        return;
    }
    const uint32_t tiles = (m + kMergeTile - 1) / kMergeTile;
    if (threadIdx.x == 0) s_tile = atomicAdd(a.tile_ctr, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= tiles) return;
    const uint32_t n = a.n;
    const uint32_t root = 2 * n - 2;
    const int32_t* __restrict__ nn = a.nn;
    const uint32_t base = tile * kMergeTile + threadIdx.x * kMergeItems;

    int32_t nj[kMergeItems];
    uint32_t fl[kMergeItems];  // bit0 survivor, bit1 merging
    uint32_t cv = 0, cm = 0;
#pragma unroll
    for (int it = 0; it < kMergeItems; ++it) {
        const uint32_t i = base + it;
        fl[it] = 0;
        nj[it] = -1;
        if (i < m) {
            const int32_t j = nn[i];
            nj[it] = j;
            const bool mutual = (j >= 0) && nn[j] == (int32_t)i;
            const bool merging = mutual && (int32_t)i < j;
            const bool valid = !(mutual && j < (int32_t)i);
            fl[it] = (valid ? 1u : 0u) | (merging ? 2u : 0u);
            cv += valid ? 1u : 0u;
            cm += merging ? 1u : 0u;
        }
    }
    const uint64_t mine = ((uint64_t)cm << 31) | cv;
    uint64_t agg;
    const uint64_t excl = block_excl_scan64<kMergeBlock>(mine, s_w, &agg);
    if (threadIdx.x == 0) {
        uint64_t* sw = a.status + tile;
        if (tile == 0) {
            __hip_atomic_store(sw, kStP | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_prefix = 0;
        } else {
            __hip_atomic_store(sw, kStA | agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint64_t pre = 0;
            int t = (int)tile - 1;
            while (true) {
                uint64_t w = __hip_atomic_load(a.status + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                uint64_t f = w & kStFlags;
                if (f == 0) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                pre += w & ~kStFlags;
                if (f == kStP) break;
                --t;
            }
            __hip_atomic_store(sw, kStP | (pre + agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_prefix = pre;
        }
        if (tile == tiles - 1) {
            const uint64_t tot = s_prefix + agg;
            const uint32_t survivors = (uint32_t)(tot & 0x7FFFFFFFu);
            const uint32_t merges = (uint32_t)(tot >> 31);
            PlocState ns;
            ns.m = survivors;
            ns.node_base = st.node_base + merges;
            ns.parity = st.parity ^ 1u;
            ns.iters = st.iters + 1;
            ns.flags = st.flags | (merges == 0 ? kFlagNoProgress : 0u);
            ns.pad[0] = ns.pad[1] = ns.pad[2] = 0;
            a.state[(k + 1) % kPlocMaxIter] = ns;
        }
    }
    __syncthreads();
    const uint64_t run0 = s_prefix + excl;
    uint32_t pos = (uint32_t)(run0 & 0x7FFFFFFFu);
    uint32_t mrank = (uint32_t)(run0 >> 31);
    const float4* __restrict__ bin = a.box[st.parity];
    const uint32_t* __restrict__ iin = a.id[st.parity];
    float4* __restrict__ bout = a.box[st.parity ^ 1u];
    uint32_t* __restrict__ iout = a.id[st.parity ^ 1u];
#pragma unroll
    for (int it = 0; it < kMergeItems; ++it) {
        const uint32_t i = base + it;
        if (!(fl[it] & 1u)) continue;
        Box b = load_box(bin + 2 * (size_t)i);
        uint32_t cid = iin[i];
        if (fl[it] & 2u) {
            const uint32_t j = (uint32_t)nj[it];
            const Box bj = load_box(bin + 2 * (size_t)j);
            const uint32_t idj = iin[j];
            b = merge(b, bj);
            const uint32_t nid = st.node_base + mrank;
            ++mrank;
            if (nid <= root)
                store_node(a.nodes, remap_id(nid, root), b, (int32_t)remap_id(cid, root),
                           (int32_t)remap_id(idj, root));
            cid = nid;
        }
        store_box(bout + 2 * (size_t)pos, b);
        iout[pos] = cid;
        ++pos;
    }
}

// Finishes every remaining iteration of one mesh inside one workgroup's LDS.
__global__ __launch_bounds__(kTailBlock) void k_ploc_tail(PlocArgs a, int k) {
    constexpr int PER = (kPlocSmall + kTailBlock - 1) / kTailBlock;
    __shared__ float s_b[6][kPlocSmall];
    __shared__ uint32_t s_id[kPlocSmall];
    __shared__ int32_t s_nn[kPlocSmall];
    __shared__ uint64_t s_w[kTailBlock / kWave];
    const PlocState st = a.state[k];
    uint32_t m = st.m;
    const uint32_t n = a.n;
    const uint32_t root = 2 * n - 2;
    const int r = a.radius;
    uint32_t node_base = st.node_base;
    uint32_t iters = st.iters;
    uint32_t flags = st.flags;
    if (flags || m > (uint32_t)kPlocSmall) {
        if (threadIdx.x == 0) {
            PlocState f = st;
            if (m > (uint32_t)kPlocSmall) f.flags |= 4u;
            a.state[a.final_slot] = f;
        }
        return;
    }
    const float4* __restrict__ bx = a.box[st.parity];
    const uint32_t* __restrict__ ids = a.id[st.parity];
    for (uint32_t i = threadIdx.x; i < m; i += kTailBlock) {
        Box b = load_box(bx + 2 * (size_t)i);
        s_b[0][i] = b.mnx; s_b[1][i] = b.mny; s_b[2][i] = b.mnz;
        s_b[3][i] = b.mxx; s_b[4][i] = b.mxy; s_b[5][i] = b.mxz;
        s_id[i] = ids[i];
    }
    __syncthreads();
    while (m > 1) {
        for (uint32_t i = threadIdx.x; i < m; i += kTailBlock) {
            Box bi;
            bi.mnx = s_b[0][i]; bi.mny = s_b[1][i]; bi.mnz = s_b[2][i];
            bi.mxx = s_b[3][i]; bi.mxy = s_b[4][i]; bi.mxz = s_b[5][i];
            const int lo = ((int)i - r > 0) ? (int)i - r : 0;
            const int hi = ((int)i + r < (int)m - 1) ? (int)i + r : (int)m - 1;
            uint64_t best = ~0ull;
            for (int j = lo; j <= hi; ++j) {
                if (j == (int)i) continue;
                Box bj;
                bj.mnx = s_b[0][j]; bj.mny = s_b[1][j]; bj.mnz = s_b[2][j];
                bj.mxx = s_b[3][j]; bj.mxy = s_b[4][j]; bj.mxz = s_b[5][j];
                uint64_t key = nn_key(area(merge(bi, bj)), (uint32_t)j);
                best = key < best ? key : best;
            }
            s_nn[i] = (best == ~0ull) ? -1 : (int32_t)(uint32_t)best;
        }
        __syncthreads();
        // Per thread PER consecutive slots; scan of packed (merges, survivors).
        uint32_t fl[PER];
        Box bb[PER];
        uint32_t cid[PER];
        uint32_t cv = 0, cm = 0;
#pragma unroll
        for (int it = 0; it < PER; ++it) {
            const uint32_t i = threadIdx.x * PER + it;
            fl[it] = 0;
            if (i < m) {
                const int32_t j = s_nn[i];
                const bool mutual = (j >= 0) && s_nn[j] == (int32_t)i;
                const bool merging = mutual && (int32_t)i < j;
                const bool valid = !(mutual && j < (int32_t)i);
                fl[it] = (valid ? 1u : 0u) | (merging ? 2u : 0u);
                cv += valid ? 1u : 0u;
                cm += merging ? 1u : 0u;
                bb[it].mnx = s_b[0][i]; bb[it].mny = s_b[1][i]; bb[it].mnz = s_b[2][i];
                bb[it].mxx = s_b[3][i]; bb[it].mxy = s_b[4][i]; bb[it].mxz = s_b[5][i];
                cid[it] = s_id[i];
                if (merging) {
                    Box bj;
                    bj.mnx = s_b[0][j]; bj.mny = s_b[1][j]; bj.mnz = s_b[2][j];
                    bj.mxx = s_b[3][j]; bj.mxy = s_b[4][j]; bj.mxz = s_b[5][j];
                    const uint32_t idj = s_id[j];
                    bb[it] = merge(bb[it], bj);
                    cid[it] = idj;  // temporarily the right child id
                    fl[it] |= (cm - 1) << 2;  // rank of this merge within the thread
                }
            }
        }
        uint64_t tot;
        const uint64_t ex = block_excl_scan64<kTailBlock>(((uint64_t)cm << 31) | cv, s_w, &tot);
        uint32_t pos = (uint32_t)(ex & 0x7FFFFFFFu);
        const uint32_t mbase = (uint32_t)(ex >> 31);
        // All LDS reads of this iteration are done (the scan synchronised).
#pragma unroll
        for (int it = 0; it < PER; ++it) {
            const uint32_t i = threadIdx.x * PER + it;
            if (i >= m || !(fl[it] & 1u)) continue;
            uint32_t out_id = s_id[i];
            if (fl[it] & 2u) {
                const uint32_t nid = node_base + mbase + (fl[it] >> 2);
                if (nid <= root)
                    store_node(a.nodes, remap_id(nid, root), bb[it], (int32_t)remap_id(out_id, root),
                               (int32_t)remap_id(cid[it], root));
                out_id = nid;
            }
            cid[it] = out_id;
            fl[it] = (fl[it] & 3u) | (pos << 2);
            ++pos;
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < PER; ++it) {
            const uint32_t i = threadIdx.x * PER + it;
            if (i >= m || !(fl[it] & 1u)) continue;
            const uint32_t p = fl[it] >> 2;
            s_b[0][p] = bb[it].mnx; s_b[1][p] = bb[it].mny; s_b[2][p] = bb[it].mnz;
            s_b[3][p] = bb[it].mxx; s_b[4][p] = bb[it].mxy; s_b[5][p] = bb[it].mxz;
            s_id[p] = cid[it];
        }
        __syncthreads();
        const uint32_t merges = (uint32_t)(tot >> 31);
        m = (uint32_t)(tot & 0x7FFFFFFFu);
        node_base += merges;
        ++iters;
        if (merges == 0) {
            flags |= kFlagNoProgress;
            break;
        }
    }
    if (threadIdx.x == 0) {
        PlocState f;
        f.m = m;
        f.node_base = node_base;
        f.parity = st.parity;
        f.iters = iters;
        f.flags = flags;
        f.pad[0] = f.pad[1] = f.pad[2] = 0;
        a.state[a.final_slot] = f;
    }
}

void launch_ploc_init(const PlocArgs& a, const uint64_t* keys, const float4* tri_boxes, uint32_t* indices,
                      hipStream_t s) {
    uint32_t g = (a.n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_ploc_init, dim3(g), dim3(256), 0, s, a, keys, tri_boxes, indices);
}

void launch_ploc_nn(const PlocArgs& a, int k, uint32_t m_bound, hipStream_t s) {
    uint32_t g = (m_bound + kNnBlock - 1) / kNnBlock;
    size_t lds = (size_t)6 * (kNnBlock + 2 * a.radius) * sizeof(float);
    hipLaunchKernelGGL(k_ploc_nn, dim3(g ? g : 1), dim3(kNnBlock), lds, s, a, k);
}

void launch_ploc_merge(const PlocArgs& a, int k, uint32_t m_bound, hipStream_t s) {
    uint32_t g = (m_bound + kMergeTile - 1) / kMergeTile;
    hipLaunchKernelGGL(k_ploc_merge, dim3(g ? g : 1), dim3(kMergeBlock), 0, s, a, k);
}

void launch_ploc_tail(const PlocArgs& a, int k, hipStream_t s) {
    hipLaunchKernelGGL(k_ploc_tail, dim3(1), dim3(kTailBlock), 0, s, a, k);
}

}  // namespace bvhk
