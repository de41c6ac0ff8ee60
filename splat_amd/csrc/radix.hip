// (c) Stable LSD radix sort of u32 keys carrying u32 values, over independent
// segments (one per view for the depth sort, one for the tile sort).
//
// Per sort, one histogram kernel computes every pass's digit histogram of
// every segment (per-wave LDS histograms merged with global atomics; counts
// are order-independent, so the result is deterministic). Then, per 8-bit
// digit:
//   * upsweep: each block counts the digits of a 4096-key tile into a
//     (segment, digit, tile) table,
//   * digit-row scan: one workgroup per (segment, digit) row turns the counts
//     into global positions (row base = the histogram's exclusive prefix),
//   * downsweep: each block re-reads its tile, ranks it with a wave64 ballot
//     multisplit (8 ballots give the mask of lanes sharing the digit; the rank
//     is the popcount below the lane plus a per-wave LDS counter, so ranks
//     follow key order: stable), exchanges it through LDS into digit order and
//     writes runs of one digit to consecutive addresses.
// No block ever waits for another (no lookback chains across the 8 XCDs).
// Grids are bounded and blocks stride over the live tiles (the live key count
// is read on the device), so sorting a few keys in a large buffer is cheap.
//
// The rasterizer sorts twice per view: the visible splats by depth bits (32
// bits, 4 passes; stable, so equal depths keep scene-index order), then the
// tile pairs by tile id. Depth digits first and tile digits last is an LSD
// sort of the 64-bit key (tile << 32 | depth bits) whose low 32-bit passes
// run once per splat instead of once per pair (a splat's pairs share its
// depth). Result order: tile, then depth, then scene index — CpuTiled's
// bucket order after radix::sort_hi32 + fix_ties (src/render/radix.rs:13-57,
// src/render/cpu.rs:396-414).
#include <algorithm>

#include "sp_device.h"
#include "sp_launch.h"

namespace sp {
namespace {

constexpr int NWAVES = RADIX_THREADS / 64;
constexpr int WAVE_KEYS = RADIX_ITEMS * 64;  // 1024 keys per wave per tile
constexpr int HIST_ITEMS = 16;
constexpr uint32_t RADIX_GRID = 4096;  // bound on the upsweep / downsweep grid

__global__ __launch_bounds__(RADIX_THREADS) void k_radix_hist(const uint32_t *__restrict__ keys,
                                                              const uint32_t *__restrict__ seg_counts,
                                                              uint64_t seg_stride, uint64_t n_cap, int passes,
                                                              uint32_t hb, uint32_t *__restrict__ hist) {
    __shared__ uint32_t s_h[NWAVES][4][RADIX_BINS];
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t seg = blockIdx.x / hb, local = blockIdx.x % hb;
    for (uint32_t k = threadIdx.x; k < NWAVES * 4 * RADIX_BINS; k += RADIX_THREADS) (&s_h[0][0][0])[k] = 0;
    __syncthreads();
    const uint64_t n = min((uint64_t)seg_counts[seg], n_cap);
    const uint32_t *kseg = keys + seg * seg_stride;
    const uint64_t stride = (uint64_t)hb * RADIX_THREADS;
    for (uint64_t idx = (uint64_t)local * RADIX_THREADS + threadIdx.x; idx < n; idx += stride) {
        const uint32_t key = kseg[idx];
        for (int p = 0; p < passes; ++p) atomicAdd(&s_h[w][p][(key >> (8 * p)) & 0xFFu], 1u);
    }
    __syncthreads();
    uint32_t *hs = hist + (uint64_t)seg * passes * RADIX_BINS;
    for (uint32_t k = threadIdx.x; k < (uint32_t)passes * RADIX_BINS; k += RADIX_THREADS) {
        uint32_t c = 0;
        for (int ww = 0; ww < NWAVES; ++ww) c += (&s_h[ww][0][0])[k];
        if (c) atomicAdd(&hs[k], c);
    }
}

// Live 4096-key tiles of a segment.
__device__ __forceinline__ uint32_t live_tiles(uint64_t n) { return (uint32_t)((n + RADIX_TILE - 1) / RADIX_TILE); }

// table[(seg * 256 + d) * bps + t] = count of digit d in tile t of `seg`.
// gps workgroups per segment stride over the segment's live tiles, so the
// grid follows the capacity only up to a bound and idle work follows the
// live count.
__global__ __launch_bounds__(RADIX_THREADS) void k_upsweep(const uint32_t *__restrict__ kin,
                                                           const uint32_t *__restrict__ seg_counts,
                                                           uint64_t seg_stride, uint64_t n_cap, uint32_t bps,
                                                           uint32_t gps, int shift, uint32_t *__restrict__ table) {
    __shared__ uint32_t s_h[NWAVES][RADIX_BINS];
    const uint32_t seg = blockIdx.x / gps;
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint64_t n = min((uint64_t)seg_counts[seg], n_cap);
    const uint32_t tiles = live_tiles(n);
    const uint32_t *ks = kin + seg * seg_stride;
    for (uint32_t local = blockIdx.x % gps; local < tiles; local += gps) {
        for (uint32_t k = threadIdx.x; k < NWAVES * RADIX_BINS; k += RADIX_THREADS) (&s_h[0][0])[k] = 0;
        __syncthreads();
        const uint64_t wbase = (uint64_t)local * RADIX_TILE + (uint64_t)w * WAVE_KEYS;
#pragma unroll 4
        for (int k = 0; k < RADIX_ITEMS; ++k) {
            const uint64_t idx = wbase + (uint64_t)k * 64 + lane;
            if (idx < n) atomicAdd(&s_h[w][(ks[idx] >> shift) & 0xFFu], 1u);
        }
        __syncthreads();
        const uint32_t d = threadIdx.x;
        uint32_t c = 0;
#pragma unroll
        for (int ww = 0; ww < NWAVES; ++ww) c += s_h[ww][d];
        table[((uint64_t)seg * RADIX_BINS + d) * bps + local] = c;
        __syncthreads();
    }
}

__device__ __forceinline__ void downsweep_tile(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                               uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                               uint32_t seg, uint64_t seg_stride, uint64_t n, uint32_t bps,
                                               uint32_t local, int shift, const uint32_t *__restrict__ table,
                                               uint32_t (*s_wh)[RADIX_BINS], uint32_t *s_lstart, uint32_t *s_gbase,
                                               uint2 *s_ex) {
    const uint64_t base = (uint64_t)local * RADIX_TILE;
    const uint32_t *ks = kin + seg * seg_stride;
    const uint32_t *vs = vin + seg * seg_stride;
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    for (uint32_t k = threadIdx.x; k < NWAVES * RADIX_BINS; k += RADIX_THREADS) (&s_wh[0][0])[k] = 0;
    __syncthreads();
    uint32_t key[RADIX_ITEMS], val[RADIX_ITEMS], rank[RADIX_ITEMS];
    const uint64_t wbase = base + (uint64_t)w * WAVE_KEYS;
#pragma unroll
    for (int k = 0; k < RADIX_ITEMS; ++k) {
        const uint64_t idx = wbase + (uint64_t)k * 64 + lane;
        const bool valid = idx < n;
        key[k] = valid ? ks[idx] : 0u;
        val[k] = valid ? vs[idx] : 0u;
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
        const uint32_t below =
            __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        const uint32_t cnt = (uint32_t)__popcll(m);
        uint32_t before = 0;
        if (valid) before = s_wh[w][d];
        rank[k] = before + below;
        if (valid && below == 0) s_wh[w][d] = before + cnt;
    }
    __syncthreads();
    const uint32_t d = threadIdx.x;
    uint32_t bc = 0;
#pragma unroll
    for (int ww = 0; ww < NWAVES; ++ww) {
        const uint32_t c = s_wh[ww][d];
        s_wh[ww][d] = bc;
        bc += c;
    }
    uint32_t tot;
    const uint32_t lstart = block_exclusive_scan<RADIX_THREADS>(bc, &tot);
    s_lstart[d] = lstart;
    s_gbase[d] = table[((uint64_t)seg * RADIX_BINS + d) * bps + local] - lstart;
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
    uint32_t *ko = kout + seg * seg_stride;
    uint32_t *vo = vout + seg * seg_stride;
    const uint32_t cnt = (uint32_t)min((uint64_t)RADIX_TILE, n - base);
    for (uint32_t j = threadIdx.x; j < cnt; j += RADIX_THREADS) {
        const uint2 e = s_ex[j];
        const uint32_t pos = s_gbase[(e.x >> shift) & 0xFFu] + j;
        if (pos < n) {  // always true for consistent counts; guards the buffer
            ko[pos] = e.x;
            vo[pos] = e.y;
        }
    }
}

__global__ __launch_bounds__(RADIX_THREADS) void k_downsweep(const uint32_t *__restrict__ kin,
                                                             const uint32_t *__restrict__ vin,
                                                             uint32_t *__restrict__ kout,
                                                             uint32_t *__restrict__ vout,
                                                             const uint32_t *__restrict__ seg_counts,
                                                             uint64_t seg_stride, uint64_t n_cap, uint32_t bps,
                                                             uint32_t gps, int shift,
                                                             const uint32_t *__restrict__ table) {
    __shared__ uint32_t s_wh[NWAVES][RADIX_BINS];  // per-wave digit counters, then per-wave offsets
    __shared__ uint32_t s_lstart[RADIX_BINS];      // block-local start of each digit
    __shared__ uint32_t s_gbase[RADIX_BINS];       // segment position minus local position
    __shared__ uint2 s_ex[RADIX_TILE];             // key/value exchange in digit order
    const uint32_t seg = blockIdx.x / gps;
    const uint64_t n = min((uint64_t)seg_counts[seg], n_cap);
    const uint32_t tiles = live_tiles(n);
    for (uint32_t local = blockIdx.x % gps; local < tiles; local += gps) {
        downsweep_tile(kin, vin, kout, vout, seg, seg_stride, n, bps, local, shift, table, s_wh, s_lstart, s_gbase,
                       s_ex);
        __syncthreads();  // LDS is reused by the next tile
    }
}


}  // namespace

uint64_t radix_table_words(uint32_t nseg, uint64_t n_cap) {
    return (uint64_t)nseg * RADIX_BINS * ((n_cap + RADIX_TILE - 1) / RADIX_TILE);
}

hipError_t launch_radix_sort(hipStream_t s, uint32_t *keys[2], uint32_t *vals[2], const uint32_t *seg_counts,
                             uint32_t nseg, uint64_t seg_stride, uint64_t n_cap, int bits, uint32_t *hist,
                             uint32_t *table, int *out_index) {
    const int passes = (bits + 7) / 8;
    int cur = 0;
    *out_index = 0;
    if (passes == 0 || n_cap == 0 || nseg == 0) return hipSuccess;
    const uint32_t hist_seg_stride = (uint32_t)passes * RADIX_BINS;
    hipError_t e = launch_fill(s, hist, (uint64_t)nseg * hist_seg_stride, 0u);
    if (e != hipSuccess) return e;
    const uint64_t hb_need = (n_cap + (uint64_t)RADIX_THREADS * HIST_ITEMS - 1) / ((uint64_t)RADIX_THREADS * HIST_ITEMS);
    const uint32_t hb =
        (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(hb_need, std::max<uint64_t>(1, 2048 / nseg)));
    hipLaunchKernelGGL(k_radix_hist, dim3(hb * nseg), dim3(RADIX_THREADS), 0, s, keys[0], seg_counts, seg_stride,
                       n_cap, passes, hb, hist);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint32_t bps = (uint32_t)((n_cap + RADIX_TILE - 1) / RADIX_TILE);
    // Workgroups per segment: one per tile up to a grid bound; they stride over
    // the live tiles, so a capacity far above the live count costs little.
    const uint32_t gps = std::max<uint32_t>(1, std::min<uint32_t>(bps, std::max<uint32_t>(1, RADIX_GRID / nseg)));
    for (int p = 0; p < passes; ++p) {
        hipLaunchKernelGGL(k_upsweep, dim3(gps * nseg), dim3(RADIX_THREADS), 0, s, keys[cur], seg_counts,
                           seg_stride, n_cap, bps, gps, 8 * p, table);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = launch_scan_digit_rows(s, table, bps, nseg, hist + RADIX_BINS * p, hist_seg_stride, seg_counts,
                                       n_cap);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_downsweep, dim3(gps * nseg), dim3(RADIX_THREADS), 0, s, keys[cur], vals[cur],
                           keys[1 - cur], vals[1 - cur], seg_counts, seg_stride, n_cap, bps, gps, 8 * p, table);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        cur = 1 - cur;
    }
    *out_index = cur;
    return hipSuccess;
}

}  // namespace sp
