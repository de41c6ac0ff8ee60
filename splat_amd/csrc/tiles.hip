// (b) Tile-pair emission and (d) tile ranges.
//
// Emission walks each view's depth-sorted visible splats. A thread turns one
// splat's clamped pixel box into its 16x16 tile rectangle (the tiles
// [x0/16, ceil(x1/16)) x [y0/16, ceil(y1/16)) of src/render/cpu.rs:266-276; an
// empty box gives no pairs, :262-264). Offsets come from a reduce-then-scan:
// a count pass writes each block's pair total, one scan workgroup per view
// turns them into block offsets (row totals = pairs per view), a one-row scan
// turns the view totals into view bases, and the expand pass re-derives the
// rectangles and writes the pairs. The block expands its pairs cooperatively
// (each thread finds its splat by binary search over the block's offsets), so
// a splat covering 1000 tiles costs no more than 1000 coalesced writes
// spread over the block. Pairs come out view by view in (depth, scene index)
// order, carrying a global tile id (view tile base + tile) and a record index
// (view * n + scene index), ready for the stable tile-id sort. The
// differentiable path sorts the emission index instead and keeps the record
// index per emission index (powner), so that per-pair gradients can be
// written back in emission order, where each splat's pairs are contiguous.
//
// Ranges: bucket bounds tile_start[t] (src/render/cpu.rs:150-163) from the
// boundaries of the tile-sorted keys; empty tiles between two boundaries get
// the same start.
#include <algorithm>

#include "sp_device.h"
#include "sp_launch.h"

namespace sp {
namespace {

struct Rect {
    uint32_t cnt, org, rw, id;
};

// Tile rectangle of the j-th depth-ordered splat of view v.
__device__ __forceinline__ Rect splat_rect(const uint32_t *__restrict__ dval, const uint4 *__restrict__ rec,
                                           uint64_t vbase, uint32_t j, uint32_t nvis, uint32_t tiles_x,
                                           uint32_t tile_base) {
    Rect r{0u, 0u, 1u, 0u};
    if (j < nvis) {
        r.id = (uint32_t)(vbase + dval[vbase + j]);
        const uint4 r2 = rec[3 * (uint64_t)r.id + 2];
        const uint32_t x0 = r2.y & 0xFFFFu, y0 = r2.y >> 16, x1 = r2.z & 0xFFFFu, y1 = r2.z >> 16;
        if (x0 < x1 && y0 < y1) {
            const uint32_t tx0 = x0 / TILE, ty0 = y0 / TILE;
            const uint32_t tx1 = (x1 + TILE - 1) / TILE, ty1 = (y1 + TILE - 1) / TILE;
            r.rw = tx1 - tx0;
            r.cnt = r.rw * (ty1 - ty0);
            r.org = tile_base + ty0 * tiles_x + tx0;
        }
    }
    return r;
}

__global__ __launch_bounds__(EMIT_THREADS) void k_emit_count(const uint32_t *__restrict__ dval, uint32_t n_splats,
                                                             uint32_t bpv, const uint32_t *__restrict__ vis_count,
                                                             const uint4 *__restrict__ rec,
                                                             const ViewParams *__restrict__ views,
                                                             uint32_t *__restrict__ bcnt) {
    const uint32_t v = blockIdx.x / bpv, local = blockIdx.x % bpv;
    const uint64_t vbase = (uint64_t)v * n_splats;
    const uint32_t nvis = min(vis_count[v], n_splats);  // dval holds n_splats entries per view
    const Rect r = splat_rect(dval, rec, vbase, local * EMIT_THREADS + threadIdx.x, nvis, views[v].ctx.tiles_x,
                              views[v].tile_base);
    uint32_t total;
    block_exclusive_scan<EMIT_THREADS>(r.cnt, &total);
    if (threadIdx.x == 0) bcnt[blockIdx.x] = total;
}

__global__ __launch_bounds__(EMIT_THREADS) void k_emit_expand(const uint32_t *__restrict__ dval, uint32_t n_splats,
                                                              uint32_t bpv, const uint32_t *__restrict__ vis_count,
                                                              const uint4 *__restrict__ rec,
                                                              const ViewParams *__restrict__ views,
                                                              const uint32_t *__restrict__ boff,
                                                              const uint32_t *__restrict__ vbase_pairs,
                                                              uint32_t *__restrict__ pkey, uint32_t *__restrict__ pval,
                                                              uint32_t *__restrict__ powner, uint64_t pair_cap,
                                                              uint32_t *__restrict__ counters) {
    __shared__ uint32_t s_pre[EMIT_THREADS];
    __shared__ uint32_t s_org[EMIT_THREADS];  // first global tile id of the rectangle
    __shared__ uint32_t s_rw[EMIT_THREADS];   // rectangle width in tiles
    __shared__ uint32_t s_id[EMIT_THREADS];   // record index (view * n + scene index)
    const uint32_t v = blockIdx.x / bpv, local = blockIdx.x % bpv;
    const uint64_t vbase = (uint64_t)v * n_splats;
    const uint32_t nvis = min(vis_count[v], n_splats);
    const uint32_t tiles_x = views[v].ctx.tiles_x;
    const Rect r =
        splat_rect(dval, rec, vbase, local * EMIT_THREADS + threadIdx.x, nvis, tiles_x, views[v].tile_base);
    uint32_t total;
    const uint32_t lscan = block_exclusive_scan<EMIT_THREADS>(r.cnt, &total);
    s_pre[threadIdx.x] = lscan;
    s_org[threadIdx.x] = r.org;
    s_rw[threadIdx.x] = r.rw;
    s_id[threadIdx.x] = r.id;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counters[CNT_OVERFLOW] = (uint64_t)counters[CNT_PAIRS] > pair_cap ? 1u : 0u;
        counters[CNT_STAGED] = 0u;  // the blend of this frame accumulates into it
    }
    __syncthreads();
    const uint64_t base = (uint64_t)vbase_pairs[v] + boff[blockIdx.x];
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
        const uint32_t rwl = s_rw[lo];
        const uint32_t ry = loc / rwl, rx = loc - ry * rwl;
        const uint64_t pos = base + q;
        if (pos < pair_cap) {
            pkey[pos] = s_org[lo] + ry * tiles_x + rx;
            if (powner) {
                // Differentiable path: the pair carries its emission index; the
                // record index stays behind in emission order.
                pval[pos] = (uint32_t)pos;
                powner[pos] = s_id[lo];
            } else {
                pval[pos] = s_id[lo];
            }
        }
    }
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

__global__ __launch_bounds__(64) void k_view_pairs(const ViewParams *__restrict__ views, uint32_t nviews,
                                                   uint32_t total_tiles, const uint32_t *__restrict__ tile_start,
                                                   uint32_t *__restrict__ view_pairs) {
    const uint32_t v = blockIdx.x * 64 + threadIdx.x;
    if (v >= nviews) return;
    const uint32_t b0 = min(views[v].tile_base, total_tiles);
    const uint32_t b1 = v + 1 < nviews ? min(views[v + 1].tile_base, total_tiles) : total_tiles;
    view_pairs[v] = tile_start[b1] - tile_start[b0];
}

}  // namespace

uint32_t emit_blocks(uint32_t n) { return n ? (n + EMIT_THREADS - 1) / EMIT_THREADS : 1; }

hipError_t launch_emit(hipStream_t s, const uint32_t *dval, uint32_t n_splats, uint32_t nviews,
                       const uint32_t *vis_count, const uint4 *rec, const ViewParams *views, uint32_t *bcnt,
                       uint32_t *vtot, uint32_t *pkey, uint32_t *pval, uint32_t *powner, uint64_t pair_cap,
                       uint32_t *counters) {
    const uint32_t bpv = emit_blocks(n_splats);
    hipLaunchKernelGGL(k_emit_count, dim3(bpv * nviews), dim3(EMIT_THREADS), 0, s, dval, n_splats, bpv, vis_count,
                       rec, views, bcnt);
    hipError_t e = hipGetLastError();
    // Per view: block offsets, and the view's pair total; then the view bases and the total.
    if (e == hipSuccess) e = launch_scan_rows(s, bcnt, bpv, nviews, nullptr, vtot);
    if (e == hipSuccess) e = launch_scan_rows(s, vtot, nviews, 1, nullptr, counters + CNT_PAIRS);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_emit_expand, dim3(bpv * nviews), dim3(EMIT_THREADS), 0, s, dval, n_splats, bpv, vis_count,
                       rec, views, bcnt, vtot, pkey, pval, powner, pair_cap, counters);
    return hipGetLastError();
}

hipError_t launch_ranges(hipStream_t s, const uint32_t *skey, const uint32_t *counters, uint64_t pair_cap,
                         uint32_t num_tiles, uint32_t *tile_start) {
    const uint64_t work = pair_cap > num_tiles ? pair_cap : (uint64_t)num_tiles + 1;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((work + 255) / 256, 65536);
    hipLaunchKernelGGL(k_ranges, dim3(grid), dim3(256), 0, s, skey, counters, pair_cap, num_tiles, tile_start);
    return hipGetLastError();
}

hipError_t launch_view_pairs(hipStream_t s, const ViewParams *views, uint32_t nviews, uint32_t total_tiles,
                             const uint32_t *tile_start, uint32_t *view_pairs) {
    if (nviews == 0) return hipSuccess;
    hipLaunchKernelGGL(k_view_pairs, dim3((nviews + 63) / 64), dim3(64), 0, s, views, nviews, total_tiles, tile_start,
                       view_pairs);
    return hipGetLastError();
}

}  // namespace sp
