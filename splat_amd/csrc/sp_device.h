// Device-side building blocks: wave and block scans, the exact exp polynomial
// of the reference blend, saturating conversions.
#pragma once

#include <hip/hip_runtime.h>

#include "sp_common.h"
#include "sp_reach.h"

namespace sp {

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Inclusive scan across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= (uint32_t)off) x += y;
    }
    return x;
}

// Exclusive scan of one u32 per thread over a block of NT threads (NT a
// multiple of 64). All threads must call it; *total receives the block sum.
template <int NT>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *total) {
    constexpr int NW = NT / 64;
    __shared__ uint32_t s_w[NW];
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t inc = wave_inclusive_scan(v);
    if (lane_id() == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        const uint32_t t = s_w[k];
        pre += (uint32_t)k < w ? t : 0u;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

// Radix sort table columns. The keys of a segment form tiles of tile_keys
// keys; a table column (one digit count per column, one workgroup task)
// covers `group` consecutive tiles: 1 for small sorts, up to 8 when there are
// enough tiles to keep 2048 columns, which cuts the scattered table traffic of
// large sorts.
__device__ __forceinline__ uint32_t radix_tiles(uint64_t n, uint32_t tile_keys) {
    return (uint32_t)((n + tile_keys - 1) / tile_keys);
}
__device__ __forceinline__ uint32_t radix_group(uint32_t tiles) { return min(8u, max(1u, tiles / 2048u)); }
__device__ __forceinline__ uint32_t radix_columns(uint64_t n, uint32_t tile_keys) {
    const uint32_t t = radix_tiles(n, tile_keys), g = radix_group(t);
    return (t + g - 1) / g;
}

// Wave votes over the active lanes. The compiler takes a ballot of a single
// compare straight from the compare's lane mask; a ballot of a combined
// predicate (and HIP's __any / __all, which go through an int) costs a
// v_cndmask and a second compare. Hot loops therefore ballot each compare on
// its own (lane_mask) and combine the masks on the scalar unit.
__device__ __forceinline__ uint64_t lane_mask(bool c) { return __builtin_amdgcn_ballot_w64(c); }
__device__ __forceinline__ bool wave_any(bool c) { return lane_mask(c) != 0ull; }
__device__ __forceinline__ bool wave_all(bool c) { return lane_mask(!c) == 0ull; }

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp_u(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROW_MASK, 0xF, false);
}

// Sum of v over the 64 lanes, left in lane 63 (the same DPP steps as
// wave_total_f). Every lane must be active.
__device__ __forceinline__ uint32_t wave_total_u(uint32_t v) {
    v += dpp_u<0xB1, 0xF>(v);
    v += dpp_u<0x4E, 0xF>(v);
    v += dpp_u<0x141, 0xF>(v);
    v += dpp_u<0x140, 0xF>(v);
    v += dpp_u<0x142, 0xA>(v);
    v += dpp_u<0x143, 0xC>(v);
    return v;
}

// Totals of N values over a 256-thread block, one barrier round; every
// thread of the block calls it with all lanes active, and every thread gets
// the totals.
template <int N>
__device__ __forceinline__ void block_totals(const uint32_t (&v)[N], uint32_t (&tot)[N]) {
    __shared__ uint32_t s_t[4][N];
    const uint32_t w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const uint32_t t = wave_total_u(v[k]);
        if (lane_id() == 63) s_t[w][k] = t;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; ++k) tot[k] = (s_t[0][k] + s_t[1][k]) + (s_t[2][k] + s_t[3][k]);
    __syncthreads();  // s_t is reused by the next call
}

// The view of a batch that owns global tile gt: the last view whose
// tile_base <= gt (bases ascend with the view index).
__device__ __forceinline__ uint32_t view_of_tile(const ViewParams *views, uint32_t nviews, uint32_t gt) {
    uint32_t lo = 0, hi = nviews;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (views[mid].tile_base <= gt)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// The tile of blend workgroup b. The dispatcher deals workgroups to the 8
// XCDs round robin (b % 8), whatever their cost, so the tiles of one XCD must
// carry an even share of the frame's work: XCD x takes the 4x4-tile
// superblocks x, x + 8, x + 16, ... of the batch (numbered as the views'
// sb_base numbers them), one tile per workgroup, the 16 tiles of a
// superblock on consecutive workgroups so that they share the XCD's L2.
// Workgroups of tiles past a frame's edge, or past the last superblock,
// return false. The grid is blend_grid(total_sbs) workgroups.
constexpr uint32_t BLEND_XCDS = 8;
__host__ __device__ constexpr uint32_t blend_grid(uint32_t total_sbs) {
    return BLEND_XCDS * SB * SB * ((total_sbs + BLEND_XCDS - 1) / BLEND_XCDS);
}
__host__ __device__ inline bool blend_tile_of(const ViewParams *views, uint32_t nviews, uint32_t total_sbs,
                                              uint32_t b, uint32_t *gt, uint32_t *vi) {
    const uint32_t j = b / BLEND_XCDS;
    const uint32_t sbi = b % BLEND_XCDS + BLEND_XCDS * (j / (SB * SB)), q = j % (SB * SB);
    if (sbi >= total_sbs) return false;
    uint32_t lo = 0, hi = nviews;  // the last view whose sb_base <= sbi
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (views[mid].sb_base <= sbi)
            lo = mid;
        else
            hi = mid;
    }
    const ViewCtx &v = views[lo].ctx;
    const uint32_t ls = sbi - views[lo].sb_base, sbx = (v.tiles_x + SB - 1) / SB;
    const uint32_t tx = (ls % sbx) * SB + q % SB, ty = (ls / sbx) * SB + q / SB;
    if (tx >= v.tiles_x || ty >= v.tiles_y) return false;
    *gt = views[lo].tile_base + ty * v.tiles_x + tx;
    *vi = lo;
    return true;
}

// Sum of v over the 64 lanes, left in lane 63, with six DPP adds (no LDS
// traffic): pairs and quads (quad_perm), eights (row_half_mirror), rows of 16
// (row_mirror), then rows 0+1 and 2+3 (row_bcast:15 into rows 1, 3) and all
// four (row_bcast:31 into rows 2, 3). Every lane must be active. The order of
// additions is fixed, so the result is deterministic.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_f(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROW_MASK, 0xF, false));
}

__device__ __forceinline__ float wave_total_f(float v) {
    v += dpp_f<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
    v += dpp_f<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
    v += dpp_f<0x141, 0xF>(v);  // row_half_mirror
    v += dpp_f<0x140, 0xF>(v);  // row_mirror
    v += dpp_f<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_f<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// Reach bits of a staged record over the tile's four quadrants (sp_reach.h).
__device__ __forceinline__ uint32_t quadrant_mask(const uint4 &r0, const uint4 &r1, const uint4 &r2, uint32_t ox,
                                                  uint32_t oy) {
    const uint32_t w0[4] = {r0.x, r0.y, r0.z, r0.w}, w1[4] = {r1.x, r1.y, r1.z, r1.w};
    const uint32_t w2[4] = {r2.x, r2.y, r2.z, r2.w};
    return quadrant_mask_words(w0, w1, w2, ox, oy);
}

// Rust `as u32` on f32 (saturating, NaN -> 0).
__device__ __forceinline__ uint32_t sat_u32(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)v;
}

// F4::exp_neg (src/render/simd.rs:315-331), one lane, x86 semantics: the
// reference's Cephes-style range reduction and degree-6 polynomial with
// unfused multiply-adds (self + a*b). Bit-identical to the CPU blend when
// compiled without contraction.
// The first reduction step x - n * C1 is one fused multiply-add: C1 =
// 0.693359375 = 355 / 2^9 has 9 significant bits and n is an integer of at
// most 8 bits here (x >= -87 gives n >= -126, and the callers' inputs are at
// most about 0), so n * C1 is exact and rounding x - n * C1 once is what the
// unfused form does.
__device__ __forceinline__ float exp_neg(float xin) {
    // max(xin, -87) as one median-of-three (fmaxf adds a canonicalising max
    // in IEEE mode); equal for every non-NaN input, and inputs are finite.
    const float x = __builtin_amdgcn_fmed3f(xin, -87.0f, 3.402823466e38f);
    const float n = rintf(x * 1.44269504088896340736f);
    const float f = fmaf(-n, 0.6933594f, x) - n * -2.1219444e-4f;
    float p = 1.9875691e-4f;
    p = 1.3981999e-3f + p * f;
    p = 8.333452e-3f + p * f;
    p = 4.1665796e-2f + p * f;
    p = 1.6666665e-1f + p * f;
    p = 0.5f + p * f;
    const float f2 = f * f;
    const float r = (f + 1.0f) + p * f2;
    return r * __int_as_float(((int)n + 127) << 23);
}

// exp_neg(x) * op, equal to the reference's product wherever it is a normal
// float: exp_neg's result r * 2^n is exact (n >= -126, and r >= 1 when
// n = -126, for clamped inputs), so scaling the rounded r * op by 2^n rounds
// the same value once.
// (A subnormal product is below ALPHA_MIN, where the blend masks the pixel
// either way.) One v_ldexp_f32 replaces building 2^n and one multiply.
// exp_neg_parts returns r and n (exp_neg(x) = r * 2^n) from the same
// operations.
__device__ __forceinline__ float exp_neg_parts(float xin, int &n_out) {
    const float x = __builtin_amdgcn_fmed3f(xin, -87.0f, 3.402823466e38f);
    const float n = rintf(x * 1.44269504088896340736f);
    const float f = fmaf(-n, 0.6933594f, x) - n * -2.1219444e-4f;
    float p = 1.9875691e-4f;
    p = 1.3981999e-3f + p * f;
    p = 8.333452e-3f + p * f;
    p = 4.1665796e-2f + p * f;
    p = 1.6666665e-1f + p * f;
    p = 0.5f + p * f;
    const float f2 = f * f;
    n_out = (int)n;
    return (f + 1.0f) + p * f2;
}
__device__ __forceinline__ float exp_neg_times(float xin, float op) {
    int n;
    const float r = exp_neg_parts(xin, n);
    return __builtin_amdgcn_ldexpf(r * op, n);
}

// to_u8 (src/render/mod.rs:282-285), one channel.
__device__ __forceinline__ uint32_t to_u8(float v) {
    if (v < 0.0f) v = 0.0f;
    if (v > 1.0f) v = 1.0f;
    const float s = v * 255.0f + 0.5f;
    if (!(s > 0.0f)) return 0u;
    if (s >= 255.0f) return 255u;
    return (uint32_t)s;
}

}  // namespace sp
