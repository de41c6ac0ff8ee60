// Device-side building blocks: lookback scans, block scans, the exact exp
// polynomial of the reference blend, saturating conversions.
#pragma once

#include <hip/hip_runtime.h>

#include "sp_common.h"

namespace sp {

// ---- epoch-tagged lookback status words -----------------------------------
// word = epoch(30) << 34 | flag(2) << 32 | value(32). flag 0 = not published,
// 1 = block aggregate, 2 = inclusive prefix. A word from another launch (other
// epoch) reads as not published, so the status array never needs clearing.
constexpr unsigned long long FLAG_AGG = 1ull, FLAG_INC = 2ull;

__device__ __forceinline__ unsigned long long pack_status(uint32_t ep, unsigned long long flag, uint32_t v) {
    return ((unsigned long long)(ep & 0x3FFFFFFFu) << 34) | (flag << 32) | (unsigned long long)v;
}
__device__ __forceinline__ bool status_ready(unsigned long long s, uint32_t ep) {
    return (uint32_t)(s >> 34) == (ep & 0x3FFFFFFFu) && ((s >> 32) & 3ull) != 0;
}
__device__ __forceinline__ unsigned long long ld_status(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(unsigned long long *p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct Ticket {
    uint32_t bid;
    uint32_t epoch;
};

// Dynamic block id: blocks take ids in the order they start, so a block only
// ever waits on blocks that are already resident (forward progress without
// assumptions about dispatch order).
__device__ __forceinline__ Ticket acquire_block(uint32_t *counters) {
    __shared__ uint32_t s_bid, s_ep;
    if (threadIdx.x == 0) {
        s_bid = __hip_atomic_fetch_add(&counters[CNT_TICKET], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_ep = __hip_atomic_load(&counters[CNT_EPOCH], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    return Ticket{s_bid, s_ep};
}

// Called by every block at its end. The last block to finish resets the
// ticket/done counters and advances the epoch, so the next launch (also a
// replay of a captured graph) starts clean. Returns true in that block.
__device__ __forceinline__ bool finish_block(uint32_t *counters, uint32_t ep) {
    __shared__ int s_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t d =
            __hip_atomic_fetch_add(&counters[CNT_DONE], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = d == gridDim.x - 1;
        if (s_last) {
            __hip_atomic_store(&counters[CNT_TICKET], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&counters[CNT_DONE], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&counters[CNT_EPOCH], (ep + 1u) & 0x3FFFFFFFu, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
    return s_last != 0;
}

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
// multiple of 64). All threads must call it.
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

// Decoupled lookback for one value per block, done by wave 0 (64
// predecessors per step). Returns the sum of `aggregate` over blocks with a
// smaller id, in every thread.
__device__ __forceinline__ uint32_t lookback_scan1(unsigned long long *status, uint32_t bid, uint32_t ep,
                                                   uint32_t aggregate) {
    __shared__ uint32_t s_excl;
    if (threadIdx.x < 64) {
        const uint32_t lane = lane_id();
        if (bid == 0) {
            if (lane == 0) {
                st_status(&status[0], pack_status(ep, FLAG_INC, aggregate));
                s_excl = 0;
            }
        } else {
            if (lane == 0) st_status(&status[bid], pack_status(ep, FLAG_AGG, aggregate));
            uint32_t excl = 0;
            int64_t j = (int64_t)bid - 1;
            for (;;) {
                const int64_t idx = j - (int64_t)lane;
                const unsigned long long s = idx >= 0 ? ld_status(&status[idx]) : pack_status(ep, FLAG_INC, 0);
                const bool ready = status_ready(s, ep);
                const bool inc = ready && ((s >> 32) & 3ull) == FLAG_INC;
                const unsigned long long incm = __ballot(inc);
                const unsigned long long notready = __ballot(!ready);
                const int fi = incm ? __ffsll((long long)incm) - 1 : 64;
                const unsigned long long need = fi >= 63 ? ~0ull : ((2ull << fi) - 1ull);
                if (notready & need) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                const uint32_t v = (int)lane <= fi ? (uint32_t)s : 0u;
                excl += wave_sum(v);
                if (fi < 64) break;
                j -= 64;
            }
            if (lane == 0) {
                st_status(&status[bid], pack_status(ep, FLAG_INC, excl + aggregate));
                s_excl = excl;
            }
        }
    }
    __syncthreads();
    return s_excl;
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
