// k_sort.hip — stable LSD radix sort of 64-bit keys (with optional 32-bit
// payload), one kernel per digit pass, decoupled look-back between tiles.
//
// Replaces cub::DeviceRadixSort::SortPairs (reference lbvh_builder.cu:412-439)
// and thrust::sort_by_key (lbvh_plus_builder.cu:465, ploc_builder.cu:459-461).
//
// Design for gfx950:
//   * digit totals of all passes come precomputed from the Morton pass
//     (k_front.hip) or from k_sort_histogram: every producer block adds its
//     LDS histogram into one zeroed table with global atomics, and each
//     onesweep tile turns its pass's row into bin starts with one block scan
//     (no separate reduce or scan launch);
//   * each onesweep tile (512 threads x 16 keys; 1024 threads for large
//     single sorts, 256 as a batch tuning option) takes a tile id from an
//     atomic counter (ids follow start order, so every look-back target is
//     already resident), ranks keys into per-wave LDS histograms (stable:
//     items are consumed in input order; lanes sharing a digit meet in a
//     64-bit LDS match mask, or by wave64 ballots in the top pass), publishes
//     its per-digit counts, resolves its global digit offsets by decoupled
//     look-back over agent-scope status words (flag | count in one 32-bit
//     word, so no payload fence is needed; eight predecessors per step),
//     reorders the tile in LDS and writes each digit run contiguously.
// Key traffic per pass: 8 B read + 8 B write per key (+4 + 4 with payload).
#include "dev_common.hpp"
#include "engine.hpp"

namespace bvhk {

constexpr uint32_t kFlagA = 1u << 30;
constexpr uint32_t kFlagP = 2u << 30;
constexpr uint32_t kValMask = (1u << 30) - 1;
// Predecessor status words read per look-back step.
constexpr int kLookWindow = 8;

__global__ void k_fill_zero(uint4* p, size_t n4) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) p[i] = make_uint4(0, 0, 0, 0);
}

// Plain streaming copy (16-B vectors); the reference point for HBM counters.
__global__ void k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n4) {
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) dst[i] = src[i];
}

void launch_stream_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
    size_t n4 = bytes / 16;
    hipLaunchKernelGGL(k_stream_copy, dim3(2048), dim3(256), 0, s, reinterpret_cast<const uint4*>(src),
                       reinterpret_cast<uint4*>(dst), n4);
}

void launch_fill_zero(uint32_t* p, size_t words, hipStream_t s) {
    size_t n4 = (words + 3) / 4;
    size_t g = (n4 + 255) / 256;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_fill_zero, dim3((uint32_t)g), dim3(256), 0, s, reinterpret_cast<uint4*>(p), n4);
}

struct ShiftArr {
    int s[8];
    uint32_t m[8];
};

template <int BITS>
__global__ __launch_bounds__(256) void k_sort_histogram(const uint64_t* __restrict__ keys, uint32_t n, int passes,
                                                         ShiftArr sh, uint32_t* __restrict__ hist_tot) {
    constexpr int BINS = 1 << BITS;
    extern __shared__ uint32_t s_hist[];
    for (int i = threadIdx.x; i < passes * BINS; i += 256) s_hist[i] = 0;
    __syncthreads();
    const uint32_t stride = gridDim.x * 256;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
        uint64_t k = keys[i];
        for (int p = 0; p < passes; ++p) atomicAdd(&s_hist[p * BINS + ((uint32_t)(k >> sh.s[p]) & sh.m[p])], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * BINS; i += 256) {
        const uint32_t c = s_hist[i];
        if (c) atomicAdd(hist_tot + i, c);
    }
}

void launch_sort_histogram(const uint64_t* keys, uint32_t n, const SortPlan& plan, uint32_t* hist_tot,
                           hipStream_t s) {
    launch_fill_zero(hist_tot, (size_t)plan.passes * (1u << plan.bits), s);
    ShiftArr sh;
    for (int i = 0; i < 8; ++i) {
        sh.s[i] = plan.shift[i];
        sh.m[i] = plan.mask[i];
    }
    size_t lds = (size_t)plan.passes * (1u << plan.bits) * 4;
    if (plan.bits == 8)
        hipLaunchKernelGGL(k_sort_histogram<8>, dim3(plan.hist_blocks), dim3(256), lds, s, keys, n, plan.passes, sh,
                           hist_tot);
    else
        hipLaunchKernelGGL(k_sort_histogram<10>, dim3(plan.hist_blocks), dim3(256), lds, s, keys, n, plan.passes, sh,
                           hist_tot);
}

// Exclusive sums over a block of T threads of two 32-bit values at once
// (packed in one 64-bit word: the low sum stays below 2^32, so no carry
// crosses into the high one); returns the packed prefixes.
template <int T>
__device__ __forceinline__ uint64_t block_excl_sum2(uint32_t hi, uint32_t lo, uint64_t* s_w) {
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const uint64_t v = ((uint64_t)hi << 32) | lo;
    uint64_t incl = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t a = __shfl_up((uint32_t)incl, o, kWave), b = __shfl_up((uint32_t)(incl >> 32), o, kWave);
        if (lane >= o) incl += ((uint64_t)b << 32) | a;
    }
    if (lane == kWave - 1) s_w[wave] = incl;
    __syncthreads();
    uint64_t before = 0;
#pragma unroll
    for (int w = 0; w < T / kWave; ++w) before += (w < wave) ? s_w[w] : 0ull;
    __syncthreads();
    return before + incl - v;
}

template <int BLOCK, int BITS, bool VALS, bool LDS_MATCH>
__global__ __launch_bounds__(BLOCK) void k_onesweep(const uint64_t* __restrict__ kin, uint64_t* __restrict__ kout,
                                                        const uint32_t* __restrict__ vin, uint32_t* __restrict__ vout,
                                                        uint32_t n, int shift, uint32_t MASK,
                                                        const uint32_t* __restrict__ hist_tot, uint32_t* status,
                                                        uint32_t* tile_ctr) {
    constexpr int BINS = 1 << BITS;
    constexpr int WAVES = BLOCK / kWave;
    constexpr int TILE = BLOCK * kSortItems;
    // Bins per thread; with more threads than bins the first BINS threads own one each.
    constexpr int BPT = (BINS + BLOCK - 1) / BLOCK;
    __shared__ uint32_t s_whist[WAVES * BINS];
    __shared__ uint32_t s_local[BINS];
    __shared__ uint32_t s_global[BINS];
    __shared__ uint64_t s_keys[TILE];
    __shared__ uint32_t s_vals[VALS ? TILE : 1];
    __shared__ uint32_t s_tile;
    __shared__ uint64_t s_w[WAVES];

    // Per-wave digit match masks for the ranking, in the key staging array
    // (unused until the reorder).
    static_assert(WAVES * BINS <= TILE, "match masks must fit the key staging array");
    uint64_t* const s_match = s_keys;

    const int tid = threadIdx.x;
    const int wave = tid / kWave;
    const int lane = tid % kWave;
    const bool owner = tid * BPT < BINS;
    if (tid == 0) s_tile = atomicAdd(tile_ctr, 1u);
    for (int i = tid; i < WAVES * BINS; i += BLOCK) {
        s_whist[i] = 0;
        s_match[i] = 0;
    }
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= gridDim.x) return;  // tile counter not cleared: publish nothing, touch nothing
    const uint32_t base = tile * TILE;
    const uint32_t seg = base + wave * (kSortItems * kWave);

    uint64_t key[kSortItems];
    uint32_t val[kSortItems];
    uint32_t rk[kSortItems];
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        uint32_t idx = seg + it * kWave + lane;
        bool valid = idx < n;
        key[it] = valid ? kin[idx] : ~0ull;
        if (VALS) val[it] = valid ? vin[idx] : 0u;
    }
    // The pass's digit totals of this thread's bins (their loads overlap the
    // key loads above); their exclusive scan, the global start of the bins,
    // is taken together with the tile's own counts after the ranking, so each
    // wave starts ranking as soon as its own keys arrive.
    uint32_t tot[BPT];
    uint32_t cnt[BPT];
    uint32_t tot_sum = 0, tsum = 0;
#pragma unroll
    for (int j = 0; j < BPT; ++j) {
        const int b = tid * BPT + j;
        cnt[j] = 0;
        tot[j] = owner ? hist_tot[b] : 0u;
        tot_sum += tot[j];
    }

    // Stable ranking inside each wave (items consumed in input order) into
    // per-wave LDS histograms; the lanes holding one digit find each other
    // in one of two ways:
    //  * LDS_MATCH: through the digit's 64-bit LDS match mask: every lane ORs
    //    in its bit, then reads the mask back.  A wave's LDS instructions
    //    complete in order, so the read sees all of this item's ORs and none
    //    of the next item's; the digit's lowest lane clears the mask after
    //    the read.  One atomic and one read per item, but lanes sharing a
    //    digit serialise on its address, so this suits well-mixed digits;
    //  * otherwise one ballot per digit bit present in MASK, at a cost that
    //    does not depend on the digit distribution (the top pass, whose
    //    digits are few and skewed).
    uint32_t* wh = s_whist + wave * BINS;
    uint64_t* wm = s_match + wave * BINS;
    const uint64_t lt_mask = lanemask_lt();
    const uint64_t my_bit = 1ull << lane;
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        uint32_t idx = seg + it * kWave + lane;
        bool valid = idx < n;
        uint32_t d = (uint32_t)(key[it] >> shift) & MASK;
        uint32_t r = 0;
        if (LDS_MATCH) {
            if (valid) __hip_atomic_fetch_or(wm + d, my_bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __builtin_amdgcn_wave_barrier();
            if (valid) {
                const uint64_t peers = __hip_atomic_load(wm + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                const uint64_t lt = peers & lt_mask;
                const uint32_t old = wh[d];
                r = old + (uint32_t)__popcll(lt);
                if (lt == 0) {
                    wh[d] = old + (uint32_t)__popcll(peers);
                    __hip_atomic_store(wm + d, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
            __builtin_amdgcn_wave_barrier();
        } else {
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < BITS; ++b) {
                if (!((MASK >> b) & 1u)) continue;
                bool bit = (d >> b) & 1u;
                uint64_t bb = __ballot(bit);
                peers &= bit ? bb : ~bb;
            }
            if (valid) {
                uint64_t lt = peers & lt_mask;
                uint32_t old = wh[d];
                r = old + (uint32_t)__popcll(lt);
                if (lt == 0) wh[d] = old + (uint32_t)__popcll(peers);
            }
        }
        rk[it] = r;
    }
    __syncthreads();
    // Per bin: exclusive prefix over the waves (in place), and the tile's
    // counts published for later tiles' look-backs.
#pragma unroll
    for (int j = 0; j < BPT; ++j) {
        const int b = tid * BPT + j;
        if (!owner) continue;
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            uint32_t c = s_whist[w * BINS + b];
            s_whist[w * BINS + b] = run;
            run += c;
        }
        cnt[j] = run;
        tsum += run;
        __hip_atomic_store(status + (size_t)tile * BINS + b, (tile == 0 ? kFlagP : kFlagA) | run, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    const uint64_t starts = block_excl_sum2<BLOCK>(tot_sum, tsum, s_w);
    uint32_t gstart = (uint32_t)(starts >> 32);
    const uint32_t lrun0 = (uint32_t)starts;
    // The tile's local bin starts: all the reorder into LDS needs.
    {
        uint32_t lr = lrun0;
#pragma unroll
        for (int j = 0; j < BPT; ++j) {
            if (!owner) continue;
            s_local[tid * BPT + j] = lr;
            lr += cnt[j];
        }
    }
    __syncthreads();
    // Reorder of the tile into LDS by digit.  Waves that own no bin (8-bit
    // digits: all but the first four) do theirs while the owners walk back.
    auto reorder = [&]() {
#pragma unroll
        for (int it = 0; it < kSortItems; ++it) {
            uint32_t idx = seg + it * kWave + lane;
            if (idx < n) {
                uint32_t d = (uint32_t)(key[it] >> shift) & MASK;
                uint32_t pos = s_local[d] + wh[d] + rk[it];
                s_keys[pos] = key[it];
                if (VALS) s_vals[pos] = val[it];
            }
        }
    };
    const bool owner_wave = wave * kWave * BPT < BINS;
    if (!owner_wave) reorder();
    // Decoupled look-back: each step reads the status words of the next
    // kLookWindow predecessors of all the thread's BPT bins at once (their
    // loads overlap), then consumes them nearest first: aggregates are added
    // and the walk goes on, an inclusive prefix ends it, an unpublished word
    // stalls it there.  (Tile 0 always publishes an inclusive prefix, so no
    // window reaches below it.)
    uint32_t g[BPT];
    int tj[BPT];
    uint32_t pending = 0;
#pragma unroll
    for (int j = 0; j < BPT; ++j) {
        g[j] = 0;
        tj[j] = (int)tile - 1;
        if (tile > 0 && owner) pending |= 1u << j;
    }
    while (pending) {
        uint32_t w[BPT][kLookWindow];
#pragma unroll
        for (int j = 0; j < BPT; ++j)
#pragma unroll
            for (int q = 0; q < kLookWindow; ++q)
                w[j][q] = (((pending >> j) & 1u) && tj[j] - q >= 0)
                              ? __hip_atomic_load(status + (size_t)(tj[j] - q) * BINS + tid * BPT + j,
                                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                              : 0u;
        bool stalled = false;
#pragma unroll
        for (int j = 0; j < BPT; ++j) {
            if (!((pending >> j) & 1u)) continue;
            bool go = true;
#pragma unroll
            for (int q = 0; q < kLookWindow; ++q) {
                if (!go) continue;
                const uint32_t flag = w[j][q] & ~kValMask;
                if (flag == 0) {
                    stalled = true;
                    go = false;
                    continue;
                }
                g[j] += w[j][q] & kValMask;
                if (flag == kFlagP) {
                    pending &= ~(1u << j);
                    go = false;
                } else {
                    --tj[j];
                }
            }
        }
        if (stalled) __builtin_amdgcn_s_sleep(1);
    }
    uint32_t lrun = lrun0;
#pragma unroll
    for (int j = 0; j < BPT; ++j) {
        const int b = tid * BPT + j;
        if (!owner) continue;
        if (tile > 0)
            __hip_atomic_store(status + (size_t)tile * BINS + b, kFlagP | (g[j] + cnt[j]), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        s_global[b] = gstart + g[j] - lrun;
        lrun += cnt[j];
        gstart += tot[j];
    }
    if (owner_wave) reorder();
    __syncthreads();

    const uint32_t tile_n = (n - base < (uint32_t)TILE) ? (n - base) : (uint32_t)TILE;
    for (uint32_t j = tid; j < tile_n; j += BLOCK) {
        uint64_t k = s_keys[j];
        uint32_t d = (uint32_t)(k >> shift) & MASK;
        uint32_t o = s_global[d] + j;
        if (o < n) {
            kout[o] = k;
            if (VALS) vout[o] = s_vals[j];
        }
    }
}

template <int BITS, bool LDS_MATCH>
static void onesweep_pass(int block, bool vals, const uint64_t* kin, uint64_t* kout, const uint32_t* vin,
                          uint32_t* vout, uint32_t n, int shift, uint32_t mask, uint32_t tiles, const uint32_t* hist_tot,
                          uint32_t* status, uint32_t* tile_ctr, hipStream_t s) {
    if (vals)
        hipLaunchKernelGGL((k_onesweep<kSortThreads, BITS, true, LDS_MATCH>), dim3(tiles), dim3(kSortThreads), 0, s,
                           kin, kout, vin, vout, n, shift, mask, hist_tot, status, tile_ctr);
    else if (BITS == 8 && block == kSortThreadsWide)
        hipLaunchKernelGGL((k_onesweep<kSortThreadsWide, 8, false, LDS_MATCH>), dim3(tiles), dim3(kSortThreadsWide), 0,
                           s, kin, kout, vin, vout, n, shift, mask, hist_tot, status, tile_ctr);
    else if (block == kSortThreadsSmall)
        hipLaunchKernelGGL((k_onesweep<kSortThreadsSmall, BITS, false, LDS_MATCH>), dim3(tiles),
                           dim3(kSortThreadsSmall), 0, s, kin, kout, vin, vout, n, shift, mask, hist_tot, status,
                           tile_ctr);
    else
        hipLaunchKernelGGL((k_onesweep<kSortThreads, BITS, false, LDS_MATCH>), dim3(tiles), dim3(kSortThreads), 0, s,
                           kin, kout, vin, vout, n, shift, mask, hist_tot, status, tile_ctr);
}

bool launch_onesweep_pass(const SortPlan& plan, int p, uint32_t n, const uint64_t* kin, uint64_t* kout,
                          const uint32_t* vin, uint32_t* vout, const uint32_t* hist_tot, uint32_t* lookback,
                          hipStream_t s) {
    const uint32_t bins = 1u << plan.bits;
    const bool vals = vin != nullptr;
    if (vals && plan.block != kSortThreads) return false;
    uint32_t* st = lookback + (size_t)p * plan.tiles * bins;
    uint32_t* ctr = lookback + (size_t)plan.passes * plan.tiles * bins + p;
    // LDS matching for the lower digits (well mixed after the earlier passes);
    // ballots for the top digit, where a few hot values are typical.
    const bool lds_match = p + 1 < plan.passes;
    const int sh = plan.shift[p];
    const uint32_t mk = plan.mask[p];
    const uint32_t* ht = hist_tot + p * bins;
    if (plan.bits == 8) {
        if (lds_match) onesweep_pass<8, true>(plan.block, vals, kin, kout, vin, vout, n, sh, mk, plan.tiles, ht, st, ctr, s);
        else onesweep_pass<8, false>(plan.block, vals, kin, kout, vin, vout, n, sh, mk, plan.tiles, ht, st, ctr, s);
    } else {
        if (lds_match) onesweep_pass<10, true>(plan.block, vals, kin, kout, vin, vout, n, sh, mk, plan.tiles, ht, st, ctr, s);
        else onesweep_pass<10, false>(plan.block, vals, kin, kout, vin, vout, n, sh, mk, plan.tiles, ht, st, ctr, s);
    }
    return true;
}

bool launch_onesweep(const SortPlan& plan, uint32_t n, uint64_t* keys_a, uint64_t* keys_b, uint32_t* vals_a,
                     uint32_t* vals_b, const uint32_t* hist_tot, uint32_t* lookback, uint64_t** keys_out,
                     uint32_t** vals_out, hipStream_t s) {
    uint64_t* kin = keys_a;
    uint64_t* kout = keys_b;
    uint32_t* vin = vals_a;
    uint32_t* vout = vals_b;
    for (int p = 0; p < plan.passes; ++p) {
        if (!launch_onesweep_pass(plan, p, n, kin, kout, vin, vout, hist_tot, lookback, s)) return false;
        uint64_t* tk = kin; kin = kout; kout = tk;
        uint32_t* tv = vin; vin = vout; vout = tv;
    }
    *keys_out = kin;
    *vals_out = vin;
    return true;
}

}  // namespace bvhk
