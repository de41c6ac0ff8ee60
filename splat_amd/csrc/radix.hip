// (c) Stable LSD radix sort of u32 keys carrying u32 values, over independent
// segments (one per view for the depth sort, one for the tile sort).
//
// Per digit (8 bits, 256 bins; or 9 bits, 512 bins, in a pass that saves a
// pass, e.g. a 9-bit sort in one pass instead of 8 + 1):
//   * upsweep: each block counts the digits of a tile (4096 keys, or 1024 for
//     sorts small enough that 4096-key tiles would leave CUs idle, 2048 in
//     their 512-bin passes) into a (segment, column, digit) table, a column
//     being 1-8 consecutive tiles,
//   * digit-row scan: one workgroup per (segment, digit) row turns the counts
//     of that digit across the columns into positions and the row's total,
//   * downsweep: each block scans its segment's row totals into digit
//     bases (so no histogram pass is needed), re-reads its tile, ranks it with a wave64 ballot
//     multisplit (8 ballots give the mask of lanes sharing the digit; the rank
//     is the popcount below the lane plus a per-wave LDS counter, so ranks
//     follow key order: stable), exchanges it through LDS into digit order and
//     writes runs of one digit to consecutive addresses.
// No block ever waits for another (no lookback chains across the 8 XCDs).
// Grids are bounded and blocks stride over the live tiles (the live key count
// is read on the device), so sorting a few keys in a large buffer is cheap.
//
// A sparse first pass (the depth sort's) reads a full-length key array in
// which FULL_CULLED marks absent entries, skips those, makes each value from
// the entry's position, and writes the segment's present count for the later
// passes: the sort compacts the visible splats as it goes.
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
constexpr uint32_t RADIX_GRID = 4096;  // bound on the upsweep / downsweep grid

// table[(seg * bps + c) * BINS + d] = count of digit d in table column c of
// `seg` (radix_group consecutive tiles of 64 * NWAVES * ITEMS keys): a
// column's counts are contiguous, so each workgroup writes (and the downsweep
// reads) one block of BINS words. gps workgroups per segment stride over the
// segment's live columns, so the grid follows the capacity only up to a bound
// and idle work follows the live count.
template <int ITEMS, int BINS, bool SPARSE>
__global__ __launch_bounds__(RADIX_THREADS) void k_upsweep(const uint32_t *__restrict__ kin,
                                                           const uint32_t *__restrict__ seg_counts,
                                                           uint64_t seg_stride, uint64_t n_cap, uint32_t bps,
                                                           uint32_t gps, int shift, int width,
                                                           uint32_t *__restrict__ table) {
    constexpr bool sparse = SPARSE;
    __shared__ uint32_t s_h[NWAVES][BINS];
    constexpr int DPT = BINS / RADIX_THREADS;  // digits per thread
    const uint32_t dmask = (1u << width) - 1u;
    const uint32_t seg = blockIdx.x / gps;
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint64_t n = sparse ? n_cap : min((uint64_t)seg_counts[seg], n_cap);
    constexpr uint32_t TILE = RADIX_THREADS * ITEMS, WAVE_KEYS = 64 * ITEMS;
    static_assert(ITEMS % 4 == 0, "whole-slice loads take four keys per lane");
    const uint32_t tiles = radix_tiles(n, TILE), group = radix_group(tiles), cols = (tiles + group - 1) / group;
    const uint32_t *ks = kin + seg * seg_stride;
    for (uint32_t col = blockIdx.x % gps; col < cols; col += gps) {
        for (uint32_t k = threadIdx.x; k < NWAVES * BINS; k += RADIX_THREADS) (&s_h[0][0])[k] = 0;
        __syncthreads();
        const uint32_t t_end = min(tiles, (col + 1) * group);
        for (uint32_t t = col * group; t < t_end; ++t) {
            const uint64_t base = (uint64_t)t * TILE;
            const uint32_t rem = (uint32_t)min((uint64_t)TILE, n - base);
            // This wave's slice of the tile: keys [w * WAVE_KEYS, +WAVE_KEYS),
            // of which wrem are in range. All of the wave's loads are issued
            // before the first count. Counting ignores key order, so a full,
            // 16-byte aligned slice loads as ITEMS / 4 contiguous 1 KB reads of
            // four keys per lane; other slices load one key per lane per read.
            const uint32_t *kw = ks + base + w * WAVE_KEYS;
            const uint32_t wrem = rem > w * WAVE_KEYS ? min(rem - w * WAVE_KEYS, WAVE_KEYS) : 0u;
            uint32_t key[ITEMS];
            bool in[ITEMS];
            if (wrem == WAVE_KEYS && (reinterpret_cast<uintptr_t>(kw) & 15u) == 0) {  // wave-uniform
#pragma unroll
                for (int q = 0; q < ITEMS / 4; ++q) {
                    const uint4 v = reinterpret_cast<const uint4 *>(kw)[q * 64 + lane];
                    key[4 * q] = v.x;
                    key[4 * q + 1] = v.y;
                    key[4 * q + 2] = v.z;
                    key[4 * q + 3] = v.w;
                }
#pragma unroll
                for (int k = 0; k < ITEMS; ++k) in[k] = true;
            } else {
#pragma unroll
                for (int k = 0; k < ITEMS; ++k) {
                    const uint32_t i = k * 64 + lane;
                    in[k] = i < wrem;
                    key[k] = in[k] ? kw[i] : FULL_CULLED;
                }
            }
            uint32_t dg[ITEMS];
#pragma unroll
            for (int k = 0; k < ITEMS; ++k) {
                const bool present = in[k] && !(sparse && key[k] == FULL_CULLED);
                dg[k] = present ? (key[k] >> shift) & dmask : BINS;  // BINS: no key
            }
#pragma unroll
            for (int k = 0; k < ITEMS; ++k) {
                const uint32_t d = dg[k];
                const bool valid = d < BINS;
                // Sorted or clustered keys often give a whole wave one digit: one
                // add then replaces 64 colliding LDS atomics (s_h[w] is this wave's own).
                const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
                const uint64_t vm = lane_mask(valid);  // all lanes vote
                const uint32_t nvalid = (uint32_t)__popcll(vm);
                if ((vm & ~lane_mask(d == d0)) == 0ull) {  // every key of the wave has digit d0
                    if (lane == 0 && nvalid) atomicAdd(&s_h[w][d0], nvalid);
                } else if (valid) {
                    atomicAdd(&s_h[w][d], 1u);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < DPT; ++j) {
            const uint32_t d = threadIdx.x * DPT + j;
            uint32_t c = 0;
#pragma unroll
            for (int ww = 0; ww < NWAVES; ++ww) c += s_h[ww][d];
            table[((uint64_t)seg * bps + col) * BINS + d] = c;
        }
        __syncthreads();
    }
}

// One tile of RADIX_THREADS * ITEMS keys: rank, exchange through LDS into
// digit order, write.
// run[j] is the next segment position of digit threadIdx.x * DPT + j (the
// column base plus the keys of that digit in the column's earlier tiles); it
// advances past this tile's keys of the digit.
// The exchange buffer has one pad entry per 32, so the runs of one digit
// (about 16 keys long when the keys spread over all digits, 128 bytes apart)
// start in different LDS banks; four large-tile workgroups still fit a CU.
__device__ __forceinline__ uint32_t ex_slot(uint32_t p) { return p + (p >> 5); }

template <int ITEMS, int BINS, bool SPARSE>
__device__ __forceinline__ void downsweep_tile(const uint32_t *__restrict__ ks, const uint32_t *__restrict__ vs,
                                               uint32_t *__restrict__ ko, uint32_t *__restrict__ vo, uint64_t n,
                                               uint32_t t, int shift, int width, uint32_t *run,
                                               uint32_t (*s_wh)[BINS], uint32_t *s_lstart, uint32_t *s_gbase,
                                               uint2 *s_ex) {
    constexpr bool sparse = SPARSE;
    constexpr int DPT = BINS / RADIX_THREADS;
    constexpr uint32_t TILE = RADIX_THREADS * ITEMS, WAVE_KEYS = 64 * ITEMS;
    const uint64_t base = (uint64_t)t * TILE;
    const uint32_t rem = (uint32_t)min((uint64_t)TILE, n - base);
    const uint32_t *kt = ks + base;
    const uint32_t *vt = vs + base;
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint32_t dmask = (1u << width) - 1u;
    for (uint32_t k = threadIdx.x; k < NWAVES * BINS; k += RADIX_THREADS) (&s_wh[0][0])[k] = 0;
    __syncthreads();
    uint32_t key[ITEMS], val[ITEMS], rank[ITEMS];
    bool present[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t i = w * WAVE_KEYS + k * 64 + lane;
        key[k] = i < rem ? kt[i] : 0u;
        // A sparse pass skips the FULL_CULLED entries and makes the values:
        // the entry's position in the segment.
        present[k] = i < rem && !(sparse && key[k] == FULL_CULLED);
        val[k] = present[k] ? (sparse ? (uint32_t)(base + i) : vt[i]) : 0u;
    }
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const bool valid = present[k];
        const uint32_t d = (key[k] >> shift) & dmask;
        // m = the lanes holding this lane's digit. A wave whose keys share
        // one digit needs no bit tests; otherwise one ballot per bit that the
        // pass carries (digits are the key's bits [shift, shift + width): bits
        // at and above `bits` do not order the keys and may carry data).
        unsigned long long m = lane_mask(valid);
        const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
        if ((m & ~lane_mask(d == d0)) != 0ull) {
            const int nd = (int)~d;
            for (int b = 0; b < width; ++b) {
                // x = all ones where this lane's bit b is 0; bb = lanes whose bit b is 1;
                // bb ^ x = lanes whose bit b matches this lane's.
                const uint32_t x = (uint32_t)__builtin_amdgcn_sbfe(nd, b, 1);
                const unsigned long long bb = lane_mask(x == 0u);
                m &= bb ^ (((unsigned long long)x << 32) | x);
            }
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
    // Per digit: the waves' counts become per-wave offsets; bc[j] = the tile's
    // keys of digit threadIdx.x * DPT + j. Consecutive digits per thread, so one
    // block scan of the threads' sums gives every digit's tile-local start.
    uint32_t bc[DPT], tsum = 0;
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
        const uint32_t d = threadIdx.x * DPT + j;
        uint32_t b = 0;
#pragma unroll
        for (int ww = 0; ww < NWAVES; ++ww) {
            const uint32_t c = s_wh[ww][d];
            s_wh[ww][d] = b;
            b += c;
        }
        bc[j] = b;
        tsum += b;
    }
    uint32_t tot;
    uint32_t lstart = block_exclusive_scan<RADIX_THREADS>(tsum, &tot);
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
        const uint32_t d = threadIdx.x * DPT + j;
        s_lstart[d] = lstart;
        s_gbase[d] = run[j] - lstart;
        run[j] += bc[j];
        lstart += bc[j];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        if (present[k]) {
            const uint32_t dk = (key[k] >> shift) & dmask;
            s_ex[ex_slot(s_lstart[dk] + s_wh[w][dk] + rank[k])] = make_uint2(key[k], val[k]);
        }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < tot; j += RADIX_THREADS) {  // tot: the tile's present keys
        const uint2 e = s_ex[ex_slot(j)];
        const uint32_t pos = s_gbase[(e.x >> shift) & dmask] + j;
        if (pos < n) {  // always true for consistent counts; guards the buffer
            ko[pos] = e.x;
            vo[pos] = e.y;
        }
    }
}

template <int ITEMS, int BINS, bool SPARSE>
__global__ __launch_bounds__(RADIX_THREADS) void k_downsweep(const uint32_t *__restrict__ kin,
                                                             const uint32_t *__restrict__ vin,
                                                             uint32_t *__restrict__ kout,
                                                             uint32_t *__restrict__ vout,
                                                             uint32_t *__restrict__ seg_counts,
                                                             uint64_t seg_stride, uint64_t n_cap, uint32_t bps,
                                                             uint32_t gps, int shift, int width,
                                                             const uint32_t *__restrict__ table,
                                                             const uint32_t *__restrict__ row_total) {
    constexpr bool sparse = SPARSE;
    __shared__ uint32_t s_wh[NWAVES][BINS];  // per-wave digit counters, then per-wave offsets
    __shared__ uint32_t s_lstart[BINS];      // block-local start of each digit
    __shared__ uint32_t s_gbase[BINS];       // segment position minus local position
    constexpr int DPT = BINS / RADIX_THREADS;
    constexpr uint32_t TILE = RADIX_THREADS * ITEMS;
    __shared__ uint2 s_ex[TILE + TILE / 32];       // key/value exchange in digit order (padded)
    const uint32_t seg = blockIdx.x / gps;
    const uint64_t n = sparse ? n_cap : min((uint64_t)seg_counts[seg], n_cap);
    const uint32_t tiles = radix_tiles(n, TILE), group = radix_group(tiles), cols = (tiles + group - 1) / group;
    if (blockIdx.x % gps >= cols) return;  // block-uniform
    // Segment position of each of this thread's digits' first key: the row
    // totals of the lower digits.
    uint32_t rt[DPT], rsum = 0;
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
        rt[j] = row_total[(uint64_t)seg * BINS + threadIdx.x * DPT + j];
        rsum += rt[j];
    }
    uint32_t seg_total;
    uint32_t row_base[DPT];
    row_base[0] = block_exclusive_scan<RADIX_THREADS>(rsum, &seg_total);
#pragma unroll
    for (int j = 1; j < DPT; ++j) row_base[j] = row_base[j - 1] + rt[j - 1];
    // A sparse pass publishes the segment's present keys: the later passes'
    // segment length (no block of this pass reads it).
    if (sparse && blockIdx.x % gps == 0 && threadIdx.x == 0) seg_counts[seg] = seg_total;
    const uint32_t *ks = kin + seg * seg_stride;
    const uint32_t *vs = vin + seg * seg_stride;
    uint32_t *ko = kout + seg * seg_stride;
    uint32_t *vo = vout + seg * seg_stride;
    for (uint32_t col = blockIdx.x % gps; col < cols; col += gps) {
        uint32_t run[DPT];
#pragma unroll
        for (int j = 0; j < DPT; ++j)
            run[j] = row_base[j] + table[((uint64_t)seg * bps + col) * BINS + threadIdx.x * DPT + j];
        const uint32_t t_end = min(tiles, (col + 1) * group);
        for (uint32_t t = col * group; t < t_end; ++t) {
            downsweep_tile<ITEMS, BINS, SPARSE>(ks, vs, ko, vo, n, t, shift, width, run, s_wh, s_lstart, s_gbase,
                                                s_ex);
            __syncthreads();  // LDS is reused by the next tile
        }
    }
}

}  // namespace

uint64_t radix_table_words(uint32_t nseg, uint64_t n_cap) {
    const uint32_t tk = radix_tile_keys(nseg, n_cap);
    return (uint64_t)nseg * RADIX_BINS_MAX * ((n_cap + tk - 1) / tk);
}

uint64_t radix_table_bound(uint32_t max_seg) {
    const uint64_t small_tile = (uint64_t)RADIX_THREADS * RADIX_ITEMS_SMALL;
    return (uint64_t)RADIX_BINS_MAX * (RADIX_SMALL_MAX / small_tile + max_seg);
}

int radix_plan(int bits, int *widths) {
    // 8-bit digits, except that passes of 9 bits are used when they need fewer
    // passes (9 -> 9, 17 -> 9+8, 18 -> 9+9, 25 -> 9+8+8, 26 -> 9+9+8, 27 -> 9+9+9).
    const int p8 = (bits + 7) / 8, p9 = (bits + 8) / 9;
    const int passes = std::min(p8, p9);
    int nine = p9 < p8 ? bits - 8 * passes : 0;
    int left = bits;
    for (int p = 0; p < passes; ++p) {
        widths[p] = std::min(left, nine > 0 ? 9 : 8);
        if (nine > 0) --nine;
        left -= widths[p];
    }
    return passes;
}

namespace {
// One pass: upsweep, digit-row scan, downsweep from keys[cur] / vals[cur] into the other buffers.
template <int ITEMS, int BINS, bool SPARSE>
hipError_t radix_pass(hipStream_t s, uint32_t *keys[2], uint32_t *vals[2], int cur, uint32_t *seg_counts,
                      uint32_t nseg, uint64_t seg_stride, uint64_t n_cap, int shift, int width, uint32_t *row_total,
                      uint32_t *table) {
    constexpr uint32_t TILE = RADIX_THREADS * ITEMS;
    // Table columns per segment (at the capacity) and workgroups per segment:
    // one per column up to a grid bound; they stride over the live columns, so
    // a capacity far above the live count costs little.
    const uint32_t bps = (uint32_t)((n_cap + TILE - 1) / TILE);
    const uint32_t gps = std::max<uint32_t>(1, std::min<uint32_t>(bps, std::max<uint32_t>(1, RADIX_GRID / nseg)));
    hipLaunchKernelGGL((k_upsweep<ITEMS, BINS, SPARSE>), dim3(gps * nseg), dim3(RADIX_THREADS), 0, s, keys[cur],
                       seg_counts, seg_stride, n_cap, bps, gps, shift, width, table);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = launch_scan_digit_rows(s, table, bps, nseg, BINS, SPARSE ? nullptr : seg_counts, n_cap, TILE, row_total);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_downsweep<ITEMS, BINS, SPARSE>), dim3(gps * nseg), dim3(RADIX_THREADS), 0, s, keys[cur],
                       vals[cur], keys[1 - cur], vals[1 - cur], seg_counts, seg_stride, n_cap, bps, gps, shift, width,
                       table, row_total);
    return hipGetLastError();
}

template <int ITEMS>
hipError_t radix_passes(hipStream_t s, uint32_t *keys[2], uint32_t *vals[2], uint32_t *seg_counts,
                        uint32_t nseg, uint64_t seg_stride, uint64_t n_cap, int bits, uint32_t *row_total,
                        uint32_t *table, int *out_index, bool sparse_first) {
    // 512-bin passes of small sorts take 2048-key tiles (the table, sized for
    // 1024-key tiles, has room for their columns).
    constexpr int WIDE = ITEMS == RADIX_ITEMS_SMALL ? RADIX_ITEMS_WIDE : ITEMS;
    int widths[8];
    const int passes = radix_plan(bits, widths);
    int cur = 0;
    int shift = 0;
    for (int p = 0; p < passes; ++p) {
        const int width = widths[p];  // digit bits this pass carries
        const uint32_t bins = width > 8 ? 512u : 256u;
        const bool sparse = sparse_first && p == 0;
        // One instantiation per (digit bins, sparse first pass); the common
        // passes carry no sparse code.
        hipError_t e;
        if (bins == 512)
            e = sparse ? radix_pass<WIDE, 512, true>(s, keys, vals, cur, seg_counts, nseg, seg_stride, n_cap, shift,
                                                      width, row_total, table)
                       : radix_pass<WIDE, 512, false>(s, keys, vals, cur, seg_counts, nseg, seg_stride, n_cap, shift,
                                                       width, row_total, table);
        else
            e = sparse ? radix_pass<ITEMS, 256, true>(s, keys, vals, cur, seg_counts, nseg, seg_stride, n_cap, shift,
                                                       width, row_total, table)
                       : radix_pass<ITEMS, 256, false>(s, keys, vals, cur, seg_counts, nseg, seg_stride, n_cap, shift,
                                                        width, row_total, table);
        if (e != hipSuccess) return e;
        shift += width;
        cur = 1 - cur;
    }
    *out_index = cur;
    return hipSuccess;
}
}  // namespace

hipError_t launch_radix_sort(hipStream_t s, uint32_t *keys[2], uint32_t *vals[2], uint32_t *seg_counts,
                             uint32_t nseg, uint64_t seg_stride, uint64_t n_cap, int bits, uint32_t *row_total,
                             uint32_t *table, int *out_index, bool sparse_first) {
    *out_index = 0;
    if (bits <= 0 || n_cap == 0 || nseg == 0) return hipSuccess;
    if (radix_tile_keys(nseg, n_cap) == RADIX_THREADS * RADIX_ITEMS_SMALL)
        return radix_passes<RADIX_ITEMS_SMALL>(s, keys, vals, seg_counts, nseg, seg_stride, n_cap, bits, row_total,
                                               table, out_index, sparse_first);
    return radix_passes<RADIX_ITEMS_LARGE>(s, keys, vals, seg_counts, nseg, seg_stride, n_cap, bits, row_total,
                                           table, out_index, sparse_first);
}

}  // namespace sp
