// Device-wide exclusive scans in reduce-then-scan form (no cross-block
// waiting): producers write one count per block into a row; these kernels
// scan the rows (one workgroup per row, all rows in parallel); consumers
// re-read their block's offset. Used by the pair emission, the binned tile
// lists and the radix sort.
#include <algorithm>

#include "sp_device.h"
#include "sp_launch.h"

namespace sp {
namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;  // entries per thread per chunk
constexpr int SCAN_CHUNK = SCAN_THREADS * SCAN_ITEMS;

// Block-wide exclusive scan of `v` that also returns the block's float sum of
// `f` (one barrier round for both).
__device__ __forceinline__ uint32_t block_scan_fsum(uint32_t v, float f, uint32_t *total, float *ftotal) {
    constexpr int NW = SCAN_THREADS / 64;
    __shared__ uint32_t s_w[NW];
    __shared__ float s_f[NW];
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t inc = wave_inclusive_scan(v);
    const float fw = wave_total_f(f);  // in lane 63
    if (lane_id() == 63) {
        s_w[w] = inc;
        s_f[w] = fw;
    }
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    float ft = 0.0f;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        const uint32_t t = s_w[k];
        pre += (uint32_t)k < w ? t : 0u;
        tot += t;
        ft += s_f[k];
    }
    __syncthreads();
    *total = tot;
    *ftotal = ft;
    return pre + inc - v;
}

// Block-wide exclusive scan of SCAN_ITEMS consecutive entries per thread of
// `row` starting at `off` (entries beyond `len` read as 0); returns the chunk
// total. Adds `carry` to every written prefix. With ftotal, also the chunk's
// float sum (for telling a total of 2^32 or more from its wrapped u32 sum).
__device__ __forceinline__ uint32_t scan_chunk(uint32_t *row, uint64_t off, uint64_t len, uint32_t carry,
                                               bool write, float *ftotal = nullptr) {
    uint32_t v[SCAN_ITEMS];
    uint32_t sum = 0;
    float fs = 0.0f;
    const uint64_t b = off + (uint64_t)threadIdx.x * SCAN_ITEMS;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = (b + k < len) ? row[b + k] : 0u;
        sum += v[k];
        if (ftotal) fs += (float)v[k];
    }
    uint32_t total;
    uint32_t pre = (ftotal ? block_scan_fsum(sum, fs, &total, ftotal) : block_exclusive_scan<SCAN_THREADS>(sum, &total)) +
                   carry;
    if (write) {
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k) {
            if (b + k < len) row[b + k] = pre;
            pre += v[k];
        }
    }
    return total;
}

// One workgroup per row of `row_len` entries: in-place exclusive scan, plus
// row_base[r] (when given) added to every entry; row_total[r] (when given)
// receives the row's sum.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_rows(uint32_t *__restrict__ data, uint64_t row_len,
                                                            const uint32_t *__restrict__ row_base,
                                                            uint32_t *__restrict__ row_total) {
    uint32_t *row = data + (uint64_t)blockIdx.x * row_len;
    uint32_t carry = row_base ? row_base[blockIdx.x] : 0u;
    const uint32_t base = carry;
    for (uint64_t off = 0; off < row_len; off += SCAN_CHUNK) carry += scan_chunk(row, off, row_len, carry, true);
    if (row_total && threadIdx.x == 0) row_total[blockIdx.x] = carry - base;
}

// Up to SCAN_MULTI arrays of `rows` rows each, scanned like k_scan_rows in
// one launch: block b takes row b % rows of array b / rows. A row total of
// 2^32 or more is written as 0xFFFFFFFF (saturated): pair counts of a frame
// can exceed 32 bits, and a wrapped count would pass for a small one. The
// exact total is C + k * 2^32, with C the u32 sum and k read off a float sum
// carried through the same scan (exact while the float's error stays below
// 2^31, i.e. for totals up to 2^36; beyond that the row is saturated anyway).
// The in-row prefixes of a saturated row wrap; the capacity clamps downstream
// keep every access in bounds, and the host rejects the frame.
constexpr int SCAN_MULTI = 3;
struct ScanArrays {
    uint32_t *data[SCAN_MULTI];
    uint32_t *row_total[SCAN_MULTI];
    // With one row per array: the row total goes here and row_total[k][0]
    // becomes 0, the exclusive scan of that single total (a caller that would
    // scan the row totals next needs no second launch).
    uint32_t *single_total[SCAN_MULTI];
};
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_rows_multi(ScanArrays a, uint64_t row_len, uint32_t rows) {
    const uint32_t k = blockIdx.x / rows, r = blockIdx.x % rows;
    uint32_t *row = a.data[k] + (uint64_t)r * row_len;
    uint32_t carry = 0;
    float fsum = 0.0f;
    for (uint64_t off = 0; off < row_len; off += SCAN_CHUNK) {
        float fc;
        carry += scan_chunk(row, off, row_len, carry, true, &fc);
        fsum += fc;
    }
    if (a.row_total[k] && threadIdx.x == 0) {
        const double wraps = rint(((double)fsum - (double)carry) * (1.0 / 4294967296.0));
        const uint32_t tot = (fsum >= 68719476736.0f || wraps >= 1.0) ? 0xFFFFFFFFu : carry;
        if (a.single_total[k]) {
            *a.single_total[k] = tot;
            a.row_total[k][r] = 0u;
        } else {
            a.row_total[k][r] = tot;
        }
    }
}

// Radix digit rows: row (segment g, digit d) holds the digit's count in each
// of the segment's columns, at data[(g * row_len + c) * bins + d] (a column's
// counts are contiguous), of which the first radix_columns(min(seg_counts[g],
// n_cap), tile_keys) are live. Each live prefix becomes its exclusive scan and
// row_total[g * bins + d] its sum (the sort's downsweep turns the row totals of
// a segment into digit bases itself, so no separate histogram pass is
// needed).
// A 128-byte line of a column holds 32 digits. Workgroups are dealt to the 8
// XCDs round-robin, so block b takes a digit on XCD b % 8's share of the
// digits (bins / 8 consecutive ones): each line is then read and written
// through one XCD's L2 only.
constexpr uint32_t XCDS = 8;
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_digit_rows(uint32_t *__restrict__ data, uint64_t row_len,
                                                                  uint32_t bins,
                                                                  const uint32_t *__restrict__ seg_counts,
                                                                  uint64_t n_cap, uint32_t tile_keys,
                                                                  uint32_t *__restrict__ row_total) {
    const uint32_t per = bins / XCDS, x = blockIdx.x % XCDS, j = blockIdx.x / XCDS;
    const uint32_t g = j / per, d = x * per + j % per;
    const uint64_t n = seg_counts ? min((uint64_t)seg_counts[g], n_cap) : n_cap;  // none: every segment holds n_cap
    const uint64_t live = min(row_len, (uint64_t)radix_columns(n, tile_keys));
    uint32_t *col0 = data + (uint64_t)g * row_len * bins + d;
    uint32_t carry = 0;
    for (uint64_t off = 0; off < live; off += SCAN_CHUNK) {
        uint32_t v[SCAN_ITEMS];
        uint32_t sum = 0;
        const uint64_t b = off + (uint64_t)threadIdx.x * SCAN_ITEMS;
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k) {
            v[k] = (b + k < live) ? col0[(b + k) * bins] : 0u;
            sum += v[k];
        }
        uint32_t total;
        uint32_t pre = block_exclusive_scan<SCAN_THREADS>(sum, &total) + carry;
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k) {
            if (b + k < live) col0[(b + k) * bins] = pre;
            pre += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0) row_total[(uint64_t)g * bins + d] = carry;
}

__global__ __launch_bounds__(256) void k_fill(uint32_t *__restrict__ p, uint64_t n, uint32_t value) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = value;
}

__global__ __launch_bounds__(256) void k_fill_f32(float *__restrict__ p, uint64_t n, float value) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = value;
}

}  // namespace

hipError_t launch_scan_rows(hipStream_t s, uint32_t *data, uint64_t row_len, uint32_t rows,
                            const uint32_t *row_base, uint32_t *row_total) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scan_rows, dim3(rows), dim3(SCAN_THREADS), 0, s, data, row_len, row_base, row_total);
    return hipGetLastError();
}

hipError_t launch_scan_rows_multi(hipStream_t s, int narrays, uint32_t *const *data, uint32_t *const *row_total,
                                  uint64_t row_len, uint32_t rows, uint32_t *const *single_total) {
    if (rows == 0 || narrays <= 0) return hipSuccess;
    if (narrays > SCAN_MULTI) return hipErrorInvalidValue;
    if (single_total && rows != 1) return hipErrorInvalidValue;
    ScanArrays a{};
    for (int k = 0; k < narrays; ++k) {
        a.data[k] = data[k];
        a.row_total[k] = row_total[k];
        a.single_total[k] = single_total && row_total[k] ? single_total[k] : nullptr;
    }
    hipLaunchKernelGGL(k_scan_rows_multi, dim3(rows * (uint32_t)narrays), dim3(SCAN_THREADS), 0, s, a, row_len, rows);
    return hipGetLastError();
}

hipError_t launch_scan_digit_rows(hipStream_t s, uint32_t *data, uint64_t row_len, uint32_t nseg, uint32_t bins,
                                  const uint32_t *seg_counts, uint64_t n_cap, uint32_t tile_keys,
                                  uint32_t *row_total) {
    if (nseg == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scan_digit_rows, dim3(nseg * bins), dim3(SCAN_THREADS), 0, s, data, row_len, bins,
                       seg_counts, n_cap, tile_keys, row_total);
    return hipGetLastError();
}

hipError_t launch_fill(hipStream_t s, uint32_t *p, uint64_t n, uint32_t value) {
    if (n == 0) return hipSuccess;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, s, p, n, value);
    return hipGetLastError();
}

hipError_t launch_fill_f32(hipStream_t s, float *p, uint64_t n, float value) {
    if (n == 0) return hipSuccess;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_f32, dim3(grid), dim3(256), 0, s, p, n, value);
    return hipGetLastError();
}

}  // namespace sp
