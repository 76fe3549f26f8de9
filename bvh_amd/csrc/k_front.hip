// k_front.hip — front end of every builder: per-triangle boxes, each mesh's
// scene box, and Morton keys with all LSD digit histograms fused into the
// same pass.
//
// Reference stages replaced:
//   kComputeBoundsAndCentroids_nt/_P/_ploc  (lbvh_builder.cu:47-66,
//       lbvh_plus_builder.cu:255-274, ploc_builder.cu:33-57)
//   kReduceBounds_Step1/2 + D2H sync         (lbvh_builder.cu:73-165,385-398)
//   kComputeMortonCodes_nt / _P              (lbvh_builder.cu:167-181,
//       lbvh_plus_builder.cu:276-290, ploc_builder.cu:121-141)
//
// MI355X mapping: both passes stream HBM with coalesced loads (nine SoA fp32
// arrays in, one 32-byte box per triangle out; then box in, 8-byte key out).
// The scene box never leaves the device: blocks reduce in registers/DPP and
// fold into six order-mapped uint32 words with integer atomics, which is
// exact and order-independent.  The Morton pass also counts every radix digit
// of every sort pass in LDS, so the sort needs no separate histogram read.
//
// Every kernel works on segments (meshes): a block covers a chunk of one
// mesh, described by a SegDesc.  A single build is a batch of one mesh, so
// batched and standalone builds run the same per-element arithmetic.
#include "dev_common.hpp"
#include "engine.hpp"

namespace bvhk {

constexpr int kFrontBlock = 512;

__device__ __forceinline__ void block_segment(const SegBlocks& sb, uint32_t& seg, uint32_t& chunk, SegDesc& d) {
    if (sb.blocks) {
        const uint2 e = sb.blocks[blockIdx.x];
        seg = e.x;
        chunk = e.y;
        d = sb.segs[seg];
    } else {
        seg = 0;
        chunk = blockIdx.x;
        d = sb.single;
    }
}

// Per-segment scene block: words 0-2 order-mapped min, 3-5 max, 6 non-finite
// flag, 7 internal-error flag.  Also clears the sort's digit-total table and
// the hierarchy's control words, and advances the build epoch.
__global__ void k_init_build(uint32_t* seg_scene, uint32_t k, uint32_t* epoch, uint32_t* hist_tot,
                             uint32_t hist_words, uint32_t* ctl, uint32_t ctl_words) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < k * 8; i += stride) {
        const uint32_t w = i & 7;
        seg_scene[i] = (w < 3) ? 0xFFFFFFFFu : 0u;
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < hist_words; i += stride) hist_tot[i] = 0u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ctl_words; i += stride) ctl[i] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0 && epoch) *epoch += 1u;
}

void launch_init_build(uint32_t* seg_scene, uint32_t k, uint32_t* epoch, hipStream_t s, uint32_t* hist_tot,
                       uint32_t hist_words, uint32_t* ctl, uint32_t ctl_words) {
    uint32_t words = (k * 8 > hist_words) ? k * 8 : hist_words;
    if (ctl_words > words) words = ctl_words;
    uint32_t g = (words + 255) / 256;
    hipLaunchKernelGGL(k_init_build, dim3(g ? g : 1), dim3(256), 0, s, seg_scene, k, epoch, hist_tot,
                       hist_tot ? hist_words : 0u, ctl, ctl ? ctl_words : 0u);
}

__device__ __forceinline__ uint32_t nonfinite(float f) {
    return ((__float_as_uint(f) & 0x7f800000u) == 0x7f800000u) ? 1u : 0u;
}

__global__ __launch_bounds__(kFrontBlock) void k_tri_bounds(SegBlocks sb, float4* __restrict__ tri_boxes,
                                                            uint32_t* __restrict__ seg_scene) {
    uint32_t seg, chunk;
    SegDesc d;
    block_segment(sb, seg, chunk, d);
    uint32_t omn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
    uint32_t omx[3] = {0u, 0u, 0u};
    uint32_t bad = 0;
    const uint32_t lo = chunk * sb.chunk;
    const uint32_t hi = (lo + sb.chunk < d.n) ? lo + sb.chunk : d.n;
    float4* __restrict__ out = tri_boxes + 2 * (size_t)d.start;
    const float* __restrict__ v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = d.v[k];
    // kFrontUnroll triangles per thread per step: all their loads are issued
    // before any box is computed, so a thread waits on HBM once per step.
    for (uint32_t i0 = lo + threadIdx.x; i0 < hi; i0 += kFrontUnroll * kFrontBlock) {
        float c[kFrontUnroll][9];
#pragma unroll
        for (int q = 0; q < kFrontUnroll; ++q) {
            const uint32_t i = i0 + q * kFrontBlock;
#pragma unroll
            for (int k = 0; k < 9; ++k) c[q][k] = (i < hi) ? v[k][i] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < kFrontUnroll; ++q) {
            const uint32_t i = i0 + q * kFrontBlock;
            if (i >= hi) continue;
            const float a0 = c[q][0], a1 = c[q][1], a2 = c[q][2];
            const float b0 = c[q][3], b1 = c[q][4], b2 = c[q][5];
            const float c0 = c[q][6], c1 = c[q][7], c2 = c[q][8];
            bad |= nonfinite(a0) | nonfinite(a1) | nonfinite(a2) | nonfinite(b0) | nonfinite(b1) | nonfinite(b2) |
                   nonfinite(c0) | nonfinite(c1) | nonfinite(c2);
            Box b;
            b.mnx = fmin_s(a0, fmin_s(b0, c0));
            b.mny = fmin_s(a1, fmin_s(b1, c1));
            b.mnz = fmin_s(a2, fmin_s(b2, c2));
            b.mxx = fmax_s(a0, fmax_s(b0, c0));
            b.mxy = fmax_s(a1, fmax_s(b1, c1));
            b.mxz = fmax_s(a2, fmax_s(b2, c2));
            store_box(out + 2 * (size_t)i, b);
            uint32_t o;
            o = f2ord(b.mnx); omn[0] = o < omn[0] ? o : omn[0];
            o = f2ord(b.mny); omn[1] = o < omn[1] ? o : omn[1];
            o = f2ord(b.mnz); omn[2] = o < omn[2] ? o : omn[2];
            o = f2ord(b.mxx); omx[0] = o > omx[0] ? o : omx[0];
            o = f2ord(b.mxy); omx[1] = o > omx[1] ? o : omx[1];
            o = f2ord(b.mxz); omx[2] = o > omx[2] ? o : omx[2];
        }
    }
    __shared__ uint32_t red[kFrontBlock / kWave][6];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        uint32_t a = wave_min_u32(omn[k]);
        uint32_t b = wave_max_u32(omx[k]);
        if (lane == 0) { red[wave][k] = a; red[wave][3 + k] = b; }
    }
    __syncthreads();
    uint32_t* sc = seg_scene + 8 * (size_t)seg;
    if (threadIdx.x < 6) {
        int k = threadIdx.x;
        uint32_t v = red[0][k];
        for (int w = 1; w < kFrontBlock / kWave; ++w) {
            uint32_t x = red[w][k];
            v = (k < 3) ? (x < v ? x : v) : (x > v ? x : v);
        }
        if (k < 3) atomicMin(sc + k, v);
        else atomicMax(sc + k, v);
    }
    if (__ballot(bad != 0) != 0 && lane == 0) atomicOr(sc + 6, 1u);
}

void launch_tri_bounds(const SegBlocks& sb, float4* tri_boxes, uint32_t* seg_scene, hipStream_t s) {
    hipLaunchKernelGGL(k_tri_bounds, dim3(sb.nblocks), dim3(kFrontBlock), 0, s, sb, tri_boxes, seg_scene);
}

SortPlan make_sort_plan(uint32_t n, int begin_bit, int end_bit, int bits, int hist_blocks, bool vals, bool wide_ok,
                        int block) {
    SortPlan p;
    p.bits = bits;
    int span = end_bit - begin_bit;
    p.passes = (span + bits - 1) / bits;
    for (int i = 0; i < 8; ++i) {
        p.shift[i] = begin_bit + i * bits;
        int w = end_bit - p.shift[i];
        w = w < 0 ? 0 : (w > bits ? bits : w);
        p.mask[i] = (1u << w) - 1u;
    }
    p.block = (wide_ok && bits == 8 && !vals && n >= kSortWideMinKeys) ? kSortThreadsWide : kSortThreads;
    if (block == kSortThreadsSmall && !vals) p.block = kSortThreadsSmall;
    else if (block == kSortThreads) p.block = kSortThreads;
    p.tile = (uint32_t)p.block * kSortItems;
    p.tiles = (n + p.tile - 1) / p.tile;
    if (p.tiles < 1) p.tiles = 1;
    if (hist_blocks > 0) {
        p.hist_blocks = hist_blocks;
    } else {
        uint32_t hb = (n + 4095) / 4096;
        p.hist_blocks = (int)(hb < 1 ? 1 : (hb > 1024 ? 1024 : hb));
    }
    return p;
}

size_t sort_zero_bytes(const SortPlan& p) {
    // Per pass: tiles x bins status words; plus one tile counter per pass (padded).
    size_t words = (size_t)p.passes * p.tiles * (1u << p.bits) + 64;
    return words * 4;
}

struct Shifts {
    int s[8];
    uint32_t m[8];
};

// Morton keys + digit totals of all passes (added into a table cleared by
// k_init_build) + zeroing of the look-back words.
template <int BITS, bool M60>
__global__ __launch_bounds__(kFrontBlock) void k_morton(SegBlocks sb, const float4* __restrict__ tri_boxes,
                                                        const uint32_t* __restrict__ seg_scene, KeyFormat kf,
                                                        uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                        int passes, Shifts sh, uint32_t* __restrict__ hist_tot,
                                                        uint4* __restrict__ zero_ptr, size_t zero_vec4) {
    constexpr int BINS = 1 << BITS;
    extern __shared__ uint32_t s_hist[];  // passes * BINS
    for (int i = threadIdx.x; i < passes * BINS; i += kFrontBlock) s_hist[i] = 0;
    {
        const size_t gstride = (size_t)gridDim.x * kFrontBlock;
        for (size_t j = (size_t)blockIdx.x * kFrontBlock + threadIdx.x; j < zero_vec4; j += gstride)
            zero_ptr[j] = make_uint4(0, 0, 0, 0);
    }
    uint32_t seg, chunk;
    SegDesc d;
    block_segment(sb, seg, chunk, d);
    const uint32_t* sc = seg_scene + 8 * (size_t)seg;
    const float smnx = ord2f(sc[0]), smny = ord2f(sc[1]), smnz = ord2f(sc[2]);
    const float smxx = ord2f(sc[3]), smxy = ord2f(sc[4]), smxz = ord2f(sc[5]);
    const uint64_t segbits = (uint64_t)seg << kf.seg_shift;
    __syncthreads();
    const uint32_t lo = chunk * sb.chunk;
    const uint32_t hi = (lo + sb.chunk < d.n) ? lo + sb.chunk : d.n;
    const float4* __restrict__ bx = tri_boxes + 2 * (size_t)d.start;
    for (uint32_t i0 = lo + threadIdx.x; i0 < hi; i0 += kFrontUnroll * kFrontBlock) {
        Box bq[kFrontUnroll];
#pragma unroll
        for (int q = 0; q < kFrontUnroll; ++q) {
            const uint32_t i = i0 + q * kFrontBlock;
            if (i < hi) bq[q] = load_box(bx + 2 * (size_t)i);
        }
#pragma unroll
        for (int q = 0; q < kFrontUnroll; ++q) {
            const uint32_t i = i0 + q * kFrontBlock;
            if (i >= hi) continue;
            const Box b = bq[q];
            float x = norm_coord(b.mnx, b.mxx, smnx, smxx);
            float y = norm_coord(b.mny, b.mxy, smny, smxy);
            float z = norm_coord(b.mnz, b.mxz, smnz, smxz);
            uint64_t key;
            const size_t g = (size_t)d.start + i;
            if (M60) {
                key = morton60(x, y, z);
                vals[g] = i;
            } else {
                key = segbits | ((uint64_t)morton30(x, y, z) << kf.idx_bits) | (uint64_t)i;
            }
            keys[g] = key;
            for (int p = 0; p < passes; ++p) {
                uint32_t dg = (uint32_t)(key >> sh.s[p]) & sh.m[p];
                atomicAdd(&s_hist[p * BINS + dg], 1u);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * BINS; i += kFrontBlock) {
        const uint32_t c = s_hist[i];
        if (c) atomicAdd(hist_tot + i, c);
    }
}

template <int BITS, bool M60>
static void morton_launch(const SegBlocks& sb, const float4* tri_boxes, const uint32_t* seg_scene, const KeyFormat& kf,
                          uint64_t* keys, uint32_t* vals, const SortPlan& plan, uint32_t* hist_tot,
                          uint32_t* zero_ptr, size_t zero_words, hipStream_t s) {
    Shifts sh;
    for (int i = 0; i < 8; ++i) {
        sh.s[i] = plan.shift[i];
        sh.m[i] = plan.mask[i];
    }
    size_t lds = (size_t)plan.passes * (1u << BITS) * 4;
    hipLaunchKernelGGL((k_morton<BITS, M60>), dim3(sb.nblocks), dim3(kFrontBlock), lds, s, sb, tri_boxes, seg_scene, kf,
                       keys, vals, plan.passes, sh, hist_tot, reinterpret_cast<uint4*>(zero_ptr), zero_words / 4);
}

void launch_morton(const SegBlocks& sb, const float4* tri_boxes, const uint32_t* seg_scene, bool m60,
                   const KeyFormat& kf, uint64_t* keys, uint32_t* vals, const SortPlan& plan, uint32_t* hist_tot,
                   uint32_t* zero_ptr, size_t zero_words, hipStream_t s) {
    if (plan.bits == 8) {
        if (m60) morton_launch<8, true>(sb, tri_boxes, seg_scene, kf, keys, vals, plan, hist_tot, zero_ptr, zero_words, s);
        else morton_launch<8, false>(sb, tri_boxes, seg_scene, kf, keys, vals, plan, hist_tot, zero_ptr, zero_words, s);
    } else {
        if (m60) morton_launch<10, true>(sb, tri_boxes, seg_scene, kf, keys, vals, plan, hist_tot, zero_ptr, zero_words, s);
        else morton_launch<10, false>(sb, tri_boxes, seg_scene, kf, keys, vals, plan, hist_tot, zero_ptr, zero_words, s);
    }
}

template <bool M60>
__global__ __launch_bounds__(kFrontBlock) void k_morton_codes(SegBlocks sb, const float4* __restrict__ tri_boxes,
                                                              const uint32_t* __restrict__ seg_scene,
                                                              uint64_t* __restrict__ codes) {
    uint32_t seg, chunk;
    SegDesc d;
    block_segment(sb, seg, chunk, d);
    const uint32_t* sc = seg_scene + 8 * (size_t)seg;
    const float smnx = ord2f(sc[0]), smny = ord2f(sc[1]), smnz = ord2f(sc[2]);
    const float smxx = ord2f(sc[3]), smxy = ord2f(sc[4]), smxz = ord2f(sc[5]);
    const uint32_t lo = chunk * sb.chunk;
    const uint32_t hi = (lo + sb.chunk < d.n) ? lo + sb.chunk : d.n;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kFrontBlock) {
        Box b = load_box(tri_boxes + 2 * ((size_t)d.start + i));
        float x = norm_coord(b.mnx, b.mxx, smnx, smxx);
        float y = norm_coord(b.mny, b.mxy, smny, smxy);
        float z = norm_coord(b.mnz, b.mxz, smnz, smxz);
        codes[(size_t)d.start + i] = M60 ? morton60(x, y, z) : (uint64_t)morton30(x, y, z);
    }
}

void launch_morton_codes(const SegBlocks& sb, const float4* tri_boxes, const uint32_t* seg_scene, bool m60,
                         uint64_t* codes, hipStream_t s) {
    if (m60)
        hipLaunchKernelGGL(k_morton_codes<true>, dim3(sb.nblocks), dim3(kFrontBlock), 0, s, sb, tri_boxes, seg_scene,
                           codes);
    else
        hipLaunchKernelGGL(k_morton_codes<false>, dim3(sb.nblocks), dim3(kFrontBlock), 0, s, sb, tri_boxes, seg_scene,
                           codes);
}

}  // namespace bvhk
