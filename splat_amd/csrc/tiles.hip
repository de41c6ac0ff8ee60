// (b) Tile-pair emission and (d) tile ranges.
//
// Emission walks the depth-sorted visible splats. Each thread turns one
// splat's clamped pixel box into its 16x16 tile rectangle (the tiles
// [x0/16, ceil(x1/16)) x [y0/16, ceil(y1/16)) of src/render/cpu.rs:266-276; an
// empty box gives no pairs, :262-264). A block-wide scan plus a decoupled
// lookback gives every splat its pair offset in one pass; the block then
// expands its pairs cooperatively (each thread finds its splat by binary
// search over the block's offsets), so a splat covering 1000 tiles costs no
// more than 1000 coalesced writes spread over the block. Pairs come out in
// (depth, scene index) order, ready for the stable tile-id sort.
//
// Ranges: bucket bounds tile_start[t] (src/render/cpu.rs:150-163) from the
// boundaries of the tile-sorted keys; empty tiles between two boundaries get
// the same start.
#include <algorithm>

#include "sp_device.h"
#include "sp_launch.h"

namespace sp {
namespace {

__global__ __launch_bounds__(EMIT_THREADS) void k_emit(const uint32_t *__restrict__ dval, uint32_t n_splats,
                                                       const uint4 *__restrict__ rec, uint32_t tiles_x,
                                                       uint32_t *__restrict__ pkey, uint32_t *__restrict__ pval,
                                                       uint64_t pair_cap, Lookback lb) {
    __shared__ uint32_t s_pre[EMIT_THREADS];
    __shared__ uint32_t s_org[EMIT_THREADS];  // first tile id of the rectangle
    __shared__ uint32_t s_rw[EMIT_THREADS];   // rectangle width in tiles
    __shared__ uint32_t s_id[EMIT_THREADS];   // scene index
    const Ticket tk = acquire_block(lb.counters);
    const uint32_t nvis = min(lb.counters[CNT_VISIBLE], n_splats);  // dval holds n_splats entries
    const uint32_t j = tk.bid * EMIT_THREADS + threadIdx.x;
    uint32_t cnt = 0, org = 0, rw = 1, id = 0;
    if (j < nvis) {
        id = dval[j];
        const uint4 r2 = rec[3 * (uint64_t)id + 2];
        const uint32_t x0 = r2.y & 0xFFFFu, y0 = r2.y >> 16, x1 = r2.z & 0xFFFFu, y1 = r2.z >> 16;
        if (x0 < x1 && y0 < y1) {
            const uint32_t tx0 = x0 / TILE, ty0 = y0 / TILE;
            const uint32_t tx1 = (x1 + TILE - 1) / TILE, ty1 = (y1 + TILE - 1) / TILE;
            rw = tx1 - tx0;
            cnt = rw * (ty1 - ty0);
            org = ty0 * tiles_x + tx0;
        }
    }
    uint32_t total;
    const uint32_t local = block_exclusive_scan<EMIT_THREADS>(cnt, &total);
    const uint32_t excl = lookback_scan1(lb.status, tk.bid, tk.epoch, total);
    s_pre[threadIdx.x] = local;
    s_org[threadIdx.x] = org;
    s_rw[threadIdx.x] = rw;
    s_id[threadIdx.x] = id;
    if (tk.bid == gridDim.x - 1 && threadIdx.x == 0) {
        const uint64_t p = (uint64_t)excl + total;
        lb.counters[CNT_PAIRS] = (uint32_t)p;
        lb.counters[CNT_OVERFLOW] = p > pair_cap ? 1u : 0u;
        lb.counters[CNT_STAGED] = 0u;  // the blend of this frame accumulates into it
    }
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < total; q += EMIT_THREADS) {
        // Largest s with s_pre[s] <= q (zero-count entries share offsets).
        uint32_t lo = 0, hi = EMIT_THREADS;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_pre[mid] <= q)
                lo = mid;
            else
                hi = mid;
        }
        const uint32_t loc = q - s_pre[lo];
        const uint32_t w = s_rw[lo];
        const uint32_t ry = loc / w, rx = loc - ry * w;
        const uint64_t pos = (uint64_t)excl + q;
        if (pos < pair_cap) {
            pkey[pos] = s_org[lo] + ry * tiles_x + rx;
            pval[pos] = s_id[lo];
        }
    }
    finish_block(lb.counters, tk.epoch);
}

__global__ __launch_bounds__(256) void k_ranges(const uint32_t *__restrict__ skey,
                                                const uint32_t *__restrict__ counters, uint64_t pair_cap,
                                                uint32_t num_tiles, uint32_t *__restrict__ tile_start) {
    const uint64_t np = min((uint64_t)counters[CNT_PAIRS], pair_cap);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (np == 0) {
        for (uint64_t t = g; t <= num_tiles; t += stride) tile_start[t] = 0;
        return;
    }
    for (uint64_t p = g; p < np; p += stride) {
        const uint32_t k = min(skey[p], num_tiles - 1);  // keys are tile ids < num_tiles
        if (p == 0) {
            for (uint32_t t = 0; t <= k; ++t) tile_start[t] = 0;
        } else {
            const uint32_t prev = min(skey[p - 1], num_tiles - 1);
            for (uint32_t t = prev + 1; t <= k; ++t) tile_start[t] = (uint32_t)p;
        }
        if (p == np - 1)
            for (uint32_t t = k + 1; t <= num_tiles; ++t) tile_start[t] = (uint32_t)np;
    }
}

}  // namespace

hipError_t launch_emit(hipStream_t s, const uint32_t *dval, uint64_t n_cap, const uint4 *rec, const ViewCtx &v,
                       uint32_t *pkey, uint32_t *pval, uint64_t pair_cap, Lookback lb) {
    const uint32_t grid = n_cap ? (uint32_t)((n_cap + EMIT_THREADS - 1) / EMIT_THREADS) : 1;
    hipLaunchKernelGGL(k_emit, dim3(grid), dim3(EMIT_THREADS), 0, s, dval, (uint32_t)n_cap, rec, v.tiles_x, pkey, pval,
                       pair_cap, lb);
    return hipGetLastError();
}

hipError_t launch_ranges(hipStream_t s, const uint32_t *skey, const uint32_t *counters, uint64_t pair_cap,
                         uint32_t num_tiles, uint32_t *tile_start) {
    const uint64_t work = pair_cap > num_tiles ? pair_cap : (uint64_t)num_tiles + 1;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((work + 255) / 256, 65536);
    hipLaunchKernelGGL(k_ranges, dim3(grid), dim3(256), 0, s, skey, counters, pair_cap, num_tiles, tile_start);
    return hipGetLastError();
}

}  // namespace sp
