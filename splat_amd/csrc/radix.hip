// (c) Stable LSD radix sort of u32 keys carrying u32 values.
//
// Per sort: one histogram kernel computes the digit histograms of every pass
// at once (per-wave LDS histograms, merged with global atomics; counts are
// order-independent, so the result is deterministic). Then one onesweep
// kernel per 8-bit digit:
//   * each block ranks a 4096-key tile with wave64 ballot multisplit (8
//     ballots per key give the mask of lanes sharing its digit; the rank is
//     the popcount of that mask below the lane, added to a per-wave LDS digit
//     counter, so ranks follow the key order: stable),
//   * publishes its 256 digit counts and resolves its global digit offsets
//     with a decoupled lookback (one thread per digit) over epoch-tagged
//     status words,
//   * exchanges the tile through LDS into digit order, then writes it out so
//     that runs of one digit land on consecutive addresses.
//
// The rasterizer uses it twice per view: on the visible splats keyed by depth
// bits (32 bits, 4 passes; stable, so equal depths keep scene-index order),
// then on the tile pairs keyed by tile id (ceil(log2 tiles / 8) passes). Sorting
// the depth digits first and the tile digits last is exactly an LSD sort of
// the 64-bit key (tile << 32 | depth bits) with the low 32-bit passes done
// once per splat instead of once per pair (the pairs of a splat share its
// depth). Result order: tile, then depth, then scene index, which is
// CpuTiled's bucket order after radix::sort_hi32 + fix_ties
// (src/render/radix.rs:13-57, src/render/cpu.rs:396-414).
#include <algorithm>

#include "sp_device.h"
#include "sp_launch.h"

namespace sp {
namespace {

constexpr int NWAVES = RADIX_THREADS / 64;
constexpr int WAVE_KEYS = RADIX_ITEMS * 64;  // 1024 keys per wave per tile
constexpr int HIST_ITEMS = 16;

__global__ __launch_bounds__(RADIX_THREADS) void k_radix_hist(const uint32_t *__restrict__ keys,
                                                              const uint32_t *__restrict__ n_dev, uint64_t n_cap,
                                                              int passes, uint32_t *__restrict__ hist) {
    __shared__ uint32_t s_h[NWAVES][4][RADIX_BINS];
    const uint32_t w = threadIdx.x >> 6;
    for (uint32_t k = threadIdx.x; k < NWAVES * 4 * RADIX_BINS; k += RADIX_THREADS) (&s_h[0][0][0])[k] = 0;
    __syncthreads();
    const uint64_t n = min((uint64_t)*n_dev, n_cap);
    const uint64_t stride = (uint64_t)gridDim.x * RADIX_THREADS;
    for (uint64_t idx = (uint64_t)blockIdx.x * RADIX_THREADS + threadIdx.x; idx < n; idx += stride) {
        const uint32_t key = keys[idx];
        for (int p = 0; p < passes; ++p) atomicAdd(&s_h[w][p][(key >> (8 * p)) & 0xFFu], 1u);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < (uint32_t)passes * RADIX_BINS; k += RADIX_THREADS) {
        uint32_t c = 0;
        for (int ww = 0; ww < NWAVES; ++ww) c += (&s_h[ww][0][0])[k];
        if (c) atomicAdd(&hist[k], c);
    }
}

__global__ __launch_bounds__(RADIX_THREADS) void k_onesweep(const uint32_t *__restrict__ kin,
                                                            const uint32_t *__restrict__ vin,
                                                            uint32_t *__restrict__ kout,
                                                            uint32_t *__restrict__ vout,
                                                            const uint32_t *__restrict__ n_dev, uint64_t n_cap,
                                                            int shift, const uint32_t *__restrict__ hist,
                                                            Lookback lb, uint32_t *__restrict__ hist_clear,
                                                            int hist_clear_n) {
    __shared__ uint32_t s_wh[NWAVES][RADIX_BINS];  // per-wave digit counters, then per-wave offsets
    __shared__ uint32_t s_lstart[RADIX_BINS];      // block-local start of each digit
    __shared__ uint32_t s_gbase[RADIX_BINS];       // global position minus local position
    __shared__ uint2 s_ex[RADIX_TILE];             // key/value exchange in digit order
    const Ticket tk = acquire_block(lb.counters);
    const uint64_t n = min((uint64_t)*n_dev, n_cap);
    const uint64_t base = (uint64_t)tk.bid * RADIX_TILE;
    if (base < n) {
        const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
        for (uint32_t k = threadIdx.x; k < NWAVES * RADIX_BINS; k += RADIX_THREADS) (&s_wh[0][0])[k] = 0;
        __syncthreads();
        uint32_t key[RADIX_ITEMS], val[RADIX_ITEMS], rank[RADIX_ITEMS];
        const uint64_t wbase = base + (uint64_t)w * WAVE_KEYS;
#pragma unroll
        for (int k = 0; k < RADIX_ITEMS; ++k) {
            const uint64_t idx = wbase + (uint64_t)k * 64 + lane;
            const bool valid = idx < n;
            key[k] = valid ? kin[idx] : 0u;
            val[k] = valid ? vin[idx] : 0u;
        }
#pragma unroll
        for (int k = 0; k < RADIX_ITEMS; ++k) {
            const uint64_t idx = wbase + (uint64_t)k * 64 + lane;
            const bool valid = idx < n;
            const uint32_t d = (key[k] >> shift) & 0xFFu;
            unsigned long long m = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const bool bit = (d >> b) & 1u;
                const unsigned long long bb = __ballot(bit);
                m &= bit ? bb : ~bb;
            }
            const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            const uint32_t cnt = (uint32_t)__popcll(m);
            uint32_t before = 0;
            if (valid) before = s_wh[w][d];
            rank[k] = before + below;
            if (valid && below == 0) s_wh[w][d] = before + cnt;
        }
        __syncthreads();
        // One thread per digit: per-wave offsets, block count, lookback.
        const uint32_t d = threadIdx.x;
        uint32_t bc = 0;
#pragma unroll
        for (int ww = 0; ww < NWAVES; ++ww) {
            const uint32_t c = s_wh[ww][d];
            s_wh[ww][d] = bc;
            bc += c;
        }
        unsigned long long *st = lb.status + (uint64_t)tk.bid * RADIX_BINS;
        if (tk.bid == 0)
            st_status(&st[d], pack_status(tk.epoch, FLAG_INC, bc));
        else
            st_status(&st[d], pack_status(tk.epoch, FLAG_AGG, bc));
        uint32_t tot;
        const uint32_t lstart = block_exclusive_scan<RADIX_THREADS>(bc, &tot);
        const uint32_t gofs = block_exclusive_scan<RADIX_THREADS>(hist[d], &tot);
        uint32_t excl = 0;
        if (tk.bid > 0) {
            int64_t j = (int64_t)tk.bid - 1;
            while (j >= 0) {
                const unsigned long long sw = ld_status(&lb.status[(uint64_t)j * RADIX_BINS + d]);
                if (!status_ready(sw, tk.epoch)) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                excl += (uint32_t)sw;
                if (((sw >> 32) & 3ull) == FLAG_INC) break;
                --j;
            }
            st_status(&st[d], pack_status(tk.epoch, FLAG_INC, excl + bc));
        }
        s_lstart[d] = lstart;
        s_gbase[d] = gofs + excl - lstart;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < RADIX_ITEMS; ++k) {
            const uint64_t idx = wbase + (uint64_t)k * 64 + lane;
            if (idx < n) {
                const uint32_t dk = (key[k] >> shift) & 0xFFu;
                s_ex[s_lstart[dk] + s_wh[w][dk] + rank[k]] = make_uint2(key[k], val[k]);
            }
        }
        __syncthreads();
        const uint32_t cnt = (uint32_t)min((uint64_t)RADIX_TILE, n - base);
        for (uint32_t j = threadIdx.x; j < cnt; j += RADIX_THREADS) {
            const uint2 e = s_ex[j];
            const uint32_t pos = s_gbase[(e.x >> shift) & 0xFFu] + j;
            if (pos < n) {  // always true for a consistent histogram; guards the buffer
                kout[pos] = e.x;
                vout[pos] = e.y;
            }
        }
    }
    const bool last = finish_block(lb.counters, tk.epoch);
    if (last && hist_clear)
        for (int k = threadIdx.x; k < hist_clear_n; k += RADIX_THREADS) hist_clear[k] = 0;
}

}  // namespace

hipError_t launch_radix_sort(hipStream_t s, uint32_t *keys[2], uint32_t *vals[2], const uint32_t *n_dev,
                             uint64_t n_cap, int bits, Lookback lb, uint32_t *hist, int *out_index) {
    const int passes = (bits + 7) / 8;
    int cur = 0;
    *out_index = 0;
    if (passes == 0 || n_cap == 0) return hipSuccess;
    const uint64_t hist_blocks_needed = (n_cap + (uint64_t)RADIX_THREADS * HIST_ITEMS - 1) / ((uint64_t)RADIX_THREADS * HIST_ITEMS);
    const uint32_t hist_grid = (uint32_t)std::min<uint64_t>(hist_blocks_needed, 2048);
    hipLaunchKernelGGL(k_radix_hist, dim3(hist_grid), dim3(RADIX_THREADS), 0, s, keys[0], n_dev, n_cap, passes, hist);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint32_t grid = (uint32_t)((n_cap + RADIX_TILE - 1) / RADIX_TILE);
    for (int p = 0; p < passes; ++p) {
        const bool lastp = p == passes - 1;
        hipLaunchKernelGGL(k_onesweep, dim3(grid), dim3(RADIX_THREADS), 0, s, keys[cur], vals[cur], keys[1 - cur],
                           vals[1 - cur], n_dev, n_cap, 8 * p, hist + RADIX_BINS * p, lb, lastp ? hist : nullptr,
                           passes * RADIX_BINS);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        cur = 1 - cur;
    }
    *out_index = cur;
    return hipSuccess;
}

}  // namespace sp
