// Device-side building blocks: wave and block scans, the exact exp polynomial
// of the reference blend, saturating conversions.
#pragma once

#include <hip/hip_runtime.h>

#include "sp_common.h"

namespace sp {

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

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

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
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
__device__ __forceinline__ float exp_neg(float xin) {
    const float x = fmaxf(xin, -87.0f);
    const float n = rintf(x * 1.44269504088896340736f);
    const float f = (x - n * 0.6933594f) - n * -2.1219444e-4f;
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
