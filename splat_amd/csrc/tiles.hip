// (b) Tile-pair emission and (d) tile ranges.
//
// Emission walks each view's depth-sorted visible splats. A thread turns one
// splat's clamped pixel box into its 16x16 tile rectangle (the tiles
// [x0/16, ceil(x1/16)) x [y0/16, ceil(y1/16)) of src/render/cpu.rs:266-276; an
// empty box gives no pairs, :262-264). Offsets come from a reduce-then-scan:
// a count pass writes each block's pair total, one scan workgroup per view
// turns them into block offsets (row totals = pairs per view), a one-row scan
// turns the view totals into view bases, and the expand pass re-derives the
// rectangles and writes the pairs, each thread finding its splat by binary
// search over in-block offsets. The dense expansion deals fixed-size chunks
// of pairs to workgroups (the nearest splats cover thousands of tiles, so
// per-splat-block work would be badly unbalanced); the sparse second phase
// keeps one workgroup per splat block. Pairs come out view by view in (depth, scene index)
// order, carrying a global tile id (view tile base + tile) and a record index
// (view * n + scene index), ready for the stable tile-id sort. The
// differentiable path sorts the emission index instead and keeps the record
// index per emission index (powner), so that per-pair gradients can be
// written back in emission order, where each splat's pairs are contiguous.
//
// Front-to-back split. Almost every tile saturates (all pixels reach
// T < 1e-4) within the first few percent of its depth-ordered list, so the
// plain path emits in two phases: FIRST emits the pairs of each view's
// nearest K1 splats (K1 = max(kmin, vis / 2^shift), capped at vis) for every
// tile; the blend then marks the tiles it finished. SECOND emits the pairs of
// the remaining splats for the unfinished tiles only, counted in O(rows) per
// splat from a per-row prefix count of unfinished tiles. A tile's SECOND list
// continues its FIRST list in the same (depth, index) order, so a pixel sees
// exactly the sequence of the one-phase path and the image bits are the same.
// FIRST also scans each view's unsplit pair count into view_pairs[v]
// (RenderStats.overlaps).
//
// Ranges: bucket bounds tile_start[t] (src/render/cpu.rs:150-163) from the
// boundaries of the tile-sorted keys; empty tiles between two boundaries get
// the same start.
#include <algorithm>

#include "sp_device.h"
#include "sp_launch.h"

namespace sp {
namespace {

constexpr uint32_t EXPAND_CHUNK = 4096;  // pairs per workgroup step of the expansion
constexpr uint32_t EXPAND_GRID = 2048;   // workgroups of the expansion (grid-stride over chunks)

struct Rect {
    uint32_t cnt, org, rw, id;
    uint32_t tx0, ty0, tx1, ty1;
};

// Tile rectangle of the j-th depth-ordered splat of view v, from the
// depth-ordered pixel boxes (box[vbase + j] = the record's packed x0|y0, x1|y1).
__device__ __forceinline__ Rect splat_rect(const uint32_t *__restrict__ dval, const uint2 *__restrict__ box,
                                           uint64_t vbase, uint32_t j, uint32_t nvis, uint32_t tiles_x,
                                           uint32_t tile_base) {
    Rect r{0u, 0u, 1u, 0u, 0u, 0u, 0u, 0u};
    if (j < nvis) {
        r.id = (uint32_t)(vbase + dval[vbase + j]);
        const uint2 b = box[vbase + j];
        const uint32_t x0 = b.x & 0xFFFFu, y0 = b.x >> 16, x1 = b.y & 0xFFFFu, y1 = b.y >> 16;
        if (x0 < x1 && y0 < y1) {
            r.tx0 = x0 / TILE;
            r.ty0 = y0 / TILE;
            r.tx1 = (x1 + TILE - 1) / TILE;
            r.ty1 = (y1 + TILE - 1) / TILE;
            r.rw = r.tx1 - r.tx0;
            r.cnt = r.rw * (r.ty1 - r.ty0);
            r.org = tile_base + r.ty0 * tiles_x + r.tx0;
        }
    }
    return r;
}

// Number of depth-ordered splats of a view the FIRST phase emits.
__device__ __forceinline__ uint32_t first_count(uint32_t vis, EmitSplit sp) {
    const uint32_t f = (uint32_t)(((uint64_t)vis + (1ull << sp.shift) - 1) >> sp.shift);
    return min(vis, max(sp.kmin, f));
}

// Unfinished tiles of row `row_base` (global id of the row's first tile) in
// columns [0, x]: the inclusive row prefix.
__device__ __forceinline__ uint32_t row_incl(const uint32_t *__restrict__ ip, uint32_t row_base, int x) {
    return x < 0 ? 0u : ip[row_base + (uint32_t)x];
}

// Unfinished tiles of the rectangle of r (SECOND phase).
__device__ __forceinline__ uint32_t open_in_rect(const uint32_t *__restrict__ ip, const Rect &r, uint32_t tile_base,
                                                 uint32_t tiles_x) {
    uint32_t c = 0;
    for (uint32_t ty = r.ty0; ty < r.ty1; ++ty) {
        const uint32_t rb = tile_base + ty * tiles_x;
        c += row_incl(ip, rb, (int)r.tx1 - 1) - row_incl(ip, rb, (int)r.tx0 - 1);
    }
    return c;
}

template <int PHASE>
__device__ __forceinline__ uint32_t phase_count(const Rect &r, uint32_t j, uint32_t k1,
                                                const uint32_t *__restrict__ ip, uint32_t tile_base,
                                                uint32_t tiles_x) {
    if (PHASE == EMIT_SECOND) return (j >= k1 && r.cnt) ? open_in_rect(ip, r, tile_base, tiles_x) : 0u;
    return j < k1 ? r.cnt : 0u;
}

template <int PHASE>
__global__ __launch_bounds__(EMIT_THREADS) void k_emit_count(const uint32_t *__restrict__ dval, uint32_t n_splats,
                                                             uint32_t bpv, const uint32_t *__restrict__ vis_count,
                                                             const uint2 *__restrict__ box,
                                                             const ViewParams *__restrict__ views, EmitSplit split,
                                                             const uint32_t *__restrict__ ip,
                                                             uint32_t *__restrict__ bcnt,
                                                             uint32_t *__restrict__ bcnt_all) {
    const uint32_t v = blockIdx.x / bpv, local = blockIdx.x % bpv;
    const uint64_t vbase = (uint64_t)v * n_splats;
    const uint32_t nvis = min(vis_count[v], n_splats);  // dval holds n_splats entries per view
    const uint32_t k1 = first_count(nvis, split);
    const uint32_t tiles_x = views[v].ctx.tiles_x, tile_base = views[v].tile_base;
    const uint32_t j = local * EMIT_THREADS + threadIdx.x;
    const Rect r = splat_rect(dval, box, vbase, j, nvis, tiles_x, tile_base);
    uint32_t total;
    block_exclusive_scan<EMIT_THREADS>(phase_count<PHASE>(r, j, k1, ip, tile_base, tiles_x), &total);
    if (threadIdx.x == 0) bcnt[blockIdx.x] = total;
    if (PHASE != EMIT_SECOND) {
        uint32_t all;
        block_exclusive_scan<EMIT_THREADS>(r.cnt, &all);
        if (threadIdx.x == 0) bcnt_all[blockIdx.x] = all;
    }
}

// box[v*n + j] = pixel box words of view v's j-th depth-ordered splat: one
// random 8-byte gather per visible splat, so that the emission passes read
// the boxes in depth order instead of gathering them from the 48-byte records.
__global__ __launch_bounds__(EMIT_THREADS) void k_gather_boxes(const uint32_t *__restrict__ dval, uint32_t n_splats,
                                                               uint32_t bpv, const uint32_t *__restrict__ vis_count,
                                                               const uint4 *__restrict__ rec,
                                                               uint2 *__restrict__ box) {
    const uint32_t v = blockIdx.x / bpv, local = blockIdx.x % bpv;
    const uint64_t vbase = (uint64_t)v * n_splats;
    const uint32_t nvis = min(vis_count[v], n_splats);
    const uint32_t j = local * EMIT_THREADS + threadIdx.x;
    if (j >= nvis) return;
    const uint2 *r = reinterpret_cast<const uint2 *>(rec + 3 * (vbase + dval[vbase + j]) + 2);
    const uint2 w01 = r[0];  // words 8-9 of the record: blue, x0|y0
    const uint2 w23 = r[1];  // words 10-11: x1|y1, threshold
    box[vbase + j] = make_uint2(w01.y, w23.x);
}

// Pair-balanced expansion: the phase's pairs (views concatenated) are cut
// into EXPAND_CHUNK-pair chunks dealt to the workgroups; a workgroup finds the
// splat block holding its chunk's first pair (binary search over the view
// bases and the view's block offsets), re-derives that block's 256
// rectangles and their in-block offsets, expands the part of the chunk they
// cover, and moves on to the next block until the chunk is done. The nearest
// splats of a view cover thousands of tiles each, so splat-major work would
// leave one workgroup with most of the pairs.
__global__ __launch_bounds__(EMIT_THREADS) void k_emit_expand_dense(const uint32_t *__restrict__ dval, uint32_t n_splats,
                                                              uint32_t bpv, uint32_t nviews,
                                                              const uint32_t *__restrict__ vis_count,
                                                              const uint2 *__restrict__ box,
                                                              const ViewParams *__restrict__ views, EmitSplit split,
                                                              const uint32_t *__restrict__ boff,
                                                              const uint32_t *__restrict__ vbase_pairs,
                                                              uint32_t *__restrict__ pkey, uint32_t *__restrict__ pval,
                                                              uint32_t *__restrict__ powner, uint64_t pair_cap,
                                                              uint32_t *__restrict__ counters) {
    __shared__ uint32_t s_pre[EMIT_THREADS];
    __shared__ uint32_t s_org[EMIT_THREADS];  // first global tile id of the rectangle
    __shared__ uint32_t s_rw[EMIT_THREADS];   // rectangle width in tiles
    __shared__ uint32_t s_id[EMIT_THREADS];   // record index (view * n + scene index)
    __shared__ uint32_t s_v, s_blk;
    const uint64_t np = counters[CNT_PAIRS];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counters[CNT_OVERFLOW] = np > pair_cap ? 1u : 0u;
        counters[CNT_STAGED] = 0u;  // the blends of this frame accumulate into it
    }
    const uint64_t nw = min(np, pair_cap);
    for (uint64_t c0 = (uint64_t)blockIdx.x * EXPAND_CHUNK; c0 < nw; c0 += (uint64_t)gridDim.x * EXPAND_CHUNK) {
        const uint64_t c1 = min(c0 + EXPAND_CHUNK, nw);
        uint64_t q = c0;
        bool search = true;
        while (q < c1) {  // block-uniform
            if (!search && threadIdx.x == 0) s_blk = s_blk + 1;  // the next splat block of the same view
            if (search && threadIdx.x == 0) {
                uint32_t lo = 0, hi = nviews;  // last view whose base <= q
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (vbase_pairs[mid] <= q)
                        lo = mid;
                    else
                        hi = mid;
                }
                const uint64_t local = q - vbase_pairs[lo];
                const uint32_t *bo = boff + (uint64_t)lo * bpv;
                uint32_t bl = 0, bh = bpv;  // last block whose offset <= local
                while (bh - bl > 1) {
                    const uint32_t mid = (bl + bh) >> 1;
                    if (bo[mid] <= local)
                        bl = mid;
                    else
                        bh = mid;
                }
                s_v = lo;
                s_blk = bl;
            }
            __syncthreads();
            const uint32_t v = s_v, blk = s_blk;
            const uint64_t vbase = (uint64_t)v * n_splats;
            const uint32_t nvis = min(vis_count[v], n_splats);
            const uint32_t k1 = first_count(nvis, split);
            const uint32_t tiles_x = views[v].ctx.tiles_x, tile_base = views[v].tile_base;
            const uint32_t j = blk * EMIT_THREADS + threadIdx.x;
            const Rect r = splat_rect(dval, box, vbase, j, nvis, tiles_x, tile_base);
            uint32_t total;
            const uint32_t lscan =
                block_exclusive_scan<EMIT_THREADS>(phase_count<EMIT_FIRST>(r, j, k1, nullptr, tile_base, tiles_x), &total);
            s_pre[threadIdx.x] = lscan;
            s_org[threadIdx.x] = r.org;
            s_rw[threadIdx.x] = r.rw;
            s_id[threadIdx.x] = r.id;
            __syncthreads();
            const uint64_t bbase = (uint64_t)vbase_pairs[v] + boff[(uint64_t)v * bpv + blk];
            const uint64_t hi = min(c1, bbase + total);
            for (uint64_t qq = q + threadIdx.x; qq < hi; qq += EMIT_THREADS) {
                const uint32_t lq = (uint32_t)(qq - bbase);
                // Largest s with s_pre[s] <= lq (zero-count entries share offsets).
                uint32_t lo = 0, hi2 = EMIT_THREADS;
                while (hi2 - lo > 1) {
                    const uint32_t mid = (lo + hi2) >> 1;
                    if (s_pre[mid] <= lq)
                        lo = mid;
                    else
                        hi2 = mid;
                }
                const uint32_t loc = lq - s_pre[lo];
                const uint32_t rwl = s_rw[lo];
                const uint32_t ry = loc / rwl, rx = loc - ry * rwl;
                const uint32_t tile = s_org[lo] + ry * tiles_x + rx;
                pkey[qq] = tile;
                if (powner) {
                    // Differentiable path: the pair carries its emission index; the
                    // record index stays behind in emission order.
                    pval[qq] = (uint32_t)qq;
                    powner[qq] = s_id[lo];
                } else {
                    pval[qq] = s_id[lo];
                }
            }
            q = hi;
            // Search again at a view's end (its last blocks may be empty) or past its last block.
            const uint64_t vend = v + 1 < nviews ? (uint64_t)vbase_pairs[v + 1] : np;
            search = q >= vend || blk + 1 >= bpv;
            __syncthreads();  // s_pre and friends are rewritten for the next block
        }
    }
}

// Splat-major expansion of the SECOND phase: one workgroup per 256
// depth-ordered splats (the nearest, largest splats are all in FIRST, so no
// block is heavy here); threads stride over the block's pairs and find their
// splat by binary search over the in-block offsets.
__global__ __launch_bounds__(EMIT_THREADS) void k_emit_expand_open(const uint32_t *__restrict__ dval, uint32_t n_splats,
                                                                   uint32_t bpv, const uint32_t *__restrict__ vis_count,
                                                                   const uint2 *__restrict__ box,
                                                                   const ViewParams *__restrict__ views,
                                                                   EmitSplit split, const uint32_t *__restrict__ ip,
                                                                   const uint32_t *__restrict__ boff,
                                                                   const uint32_t *__restrict__ vbase_pairs,
                                                                   uint32_t *__restrict__ pkey,
                                                                   uint32_t *__restrict__ pval, uint64_t pair_cap,
                                                                   uint32_t *__restrict__ counters) {
    __shared__ uint32_t s_pre[EMIT_THREADS];
    __shared__ uint32_t s_org[EMIT_THREADS];
    __shared__ uint32_t s_rw[EMIT_THREADS];
    __shared__ uint32_t s_id[EMIT_THREADS];
    __shared__ uint32_t s_rows[EMIT_THREADS];
    const uint32_t v = blockIdx.x / bpv, local = blockIdx.x % bpv;
    const uint64_t vbase = (uint64_t)v * n_splats;
    const uint32_t nvis = min(vis_count[v], n_splats);
    const uint32_t k1 = first_count(nvis, split);
    const uint32_t tiles_x = views[v].ctx.tiles_x, tile_base = views[v].tile_base;
    const uint32_t j = local * EMIT_THREADS + threadIdx.x;
    const Rect r = splat_rect(dval, box, vbase, j, nvis, tiles_x, tile_base);
    uint32_t total;
    const uint32_t lscan =
        block_exclusive_scan<EMIT_THREADS>(phase_count<EMIT_SECOND>(r, j, k1, ip, tile_base, tiles_x), &total);
    s_pre[threadIdx.x] = lscan;
    s_org[threadIdx.x] = r.org;
    s_rw[threadIdx.x] = r.rw;
    s_id[threadIdx.x] = r.id;
    s_rows[threadIdx.x] = r.ty1 - r.ty0;
    if (blockIdx.x == 0 && threadIdx.x == 0 && counters[CNT_PAIRS2] > pair_cap) counters[CNT_OVERFLOW] = 1u;
    __syncthreads();
    const uint64_t base = (uint64_t)vbase_pairs[v] + boff[blockIdx.x];
    for (uint32_t q = threadIdx.x; q < total; q += EMIT_THREADS) {
        uint32_t lo = 0, hi = EMIT_THREADS;  // largest s with s_pre[s] <= q
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_pre[mid] <= q)
                lo = mid;
            else
                hi = mid;
        }
        const uint32_t loc = q - s_pre[lo];
        const uint32_t rwl = s_rw[lo];
        // loc-th unfinished tile of the rectangle, rows top to bottom.
        const uint32_t org = s_org[lo] - tile_base;
        const uint32_t ty0 = org / tiles_x, tx0 = org % tiles_x;
        uint32_t rem = loc, ty = ty0;
        uint32_t rb = tile_base + ty * tiles_x, before = row_incl(ip, rb, (int)tx0 - 1);
        for (uint32_t k = 0; k < s_rows[lo]; ++k, ++ty) {
            rb = tile_base + ty * tiles_x;
            before = row_incl(ip, rb, (int)tx0 - 1);
            const uint32_t c = row_incl(ip, rb, (int)(tx0 + rwl) - 1) - before;
            if (rem < c) break;
            rem -= c;
        }
        // Smallest x in [tx0, tx0 + rwl) with ip(x) - before > rem.
        uint32_t xl = tx0, xh = tx0 + rwl - 1;
        while (xl < xh) {
            const uint32_t mid = (xl + xh) >> 1;
            if (ip[rb + mid] - before > rem)
                xh = mid;
            else
                xl = mid + 1;
        }
        const uint64_t pos = base + q;
        if (pos < pair_cap) {
            pkey[pos] = rb + xl;
            pval[pos] = s_id[lo];
        }
    }
}

// Inclusive per-row prefix counts of unfinished tiles: one wave per tile
// row (lanes over 64 columns at a time), blockIdx.y = view.
__global__ __launch_bounds__(256) void k_row_prefix(const ViewParams *__restrict__ views,
                                                    const uint32_t *__restrict__ done, uint32_t *__restrict__ ip) {
    const ViewParams &vp = views[blockIdx.y];
    const uint32_t tx_n = vp.ctx.tiles_x, ty_n = vp.ctx.tiles_y;
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    for (uint32_t ty = blockIdx.x * 4 + w; ty < ty_n; ty += gridDim.x * 4) {
        const uint32_t rb = vp.tile_base + ty * tx_n;
        uint32_t carry = 0;
        for (uint32_t x0 = 0; x0 < tx_n; x0 += 64) {
            const uint32_t x = x0 + lane;
            const uint32_t open = (x < tx_n && !done[rb + x]) ? 1u : 0u;
            const uint32_t inc = wave_inclusive_scan(open);
            if (x < tx_n) ip[rb + x] = carry + inc;
            carry += (uint32_t)__shfl((int)inc, 63);
        }
    }
}

__global__ __launch_bounds__(256) void k_ranges(const uint32_t *__restrict__ skey,
                                                const uint32_t *__restrict__ count, uint64_t pair_cap,
                                                uint32_t num_tiles, uint32_t *__restrict__ tile_start) {
    const uint64_t np = min((uint64_t)*count, pair_cap);
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

template <int PHASE>
hipError_t emit_phase(hipStream_t s, const uint32_t *dval, uint32_t n_splats, uint32_t nviews,
                      const uint32_t *vis_count, const uint4 *rec, uint2 *box, const ViewParams *views,
                      EmitSplit split,
                      const uint32_t *ip, uint32_t *bcnt, uint32_t *bcnt_all, uint32_t *vtot, uint32_t *view_pairs,
                      uint32_t *pkey, uint32_t *pval, uint32_t *powner, uint64_t pair_cap, uint32_t count_slot,
                      uint32_t *counters) {
    const uint32_t bpv = emit_blocks(n_splats);
    if (PHASE == EMIT_FIRST) {
        hipLaunchKernelGGL(k_gather_boxes, dim3(bpv * nviews), dim3(EMIT_THREADS), 0, s, dval, n_splats, bpv,
                           vis_count, rec, box);
        const hipError_t eg = hipGetLastError();
        if (eg != hipSuccess) return eg;
    }
    hipLaunchKernelGGL(k_emit_count<PHASE>, dim3(bpv * nviews), dim3(EMIT_THREADS), 0, s, dval, n_splats, bpv,
                       vis_count, box, views, split, ip, bcnt, bcnt_all);
    hipError_t e = hipGetLastError();
    // Per view: block offsets, and the view's pair total; then the view bases and the total.
    if (e == hipSuccess) e = launch_scan_rows(s, bcnt, bpv, nviews, nullptr, vtot);
    if (e == hipSuccess) e = launch_scan_rows(s, vtot, nviews, 1, nullptr, counters + count_slot);
    // Unsplit pair count of every view (RenderStats.overlaps).
    if (e == hipSuccess && PHASE != EMIT_SECOND) e = launch_scan_rows(s, bcnt_all, bpv, nviews, nullptr, view_pairs);
    if (e != hipSuccess) return e;
    if (PHASE == EMIT_SECOND) {
        hipLaunchKernelGGL(k_emit_expand_open, dim3(bpv * nviews), dim3(EMIT_THREADS), 0, s, dval, n_splats, bpv,
                           vis_count, box, views, split, ip, bcnt, vtot, pkey, pval, pair_cap, counters);
    } else {
        const uint64_t chunks = (pair_cap + EXPAND_CHUNK - 1) / EXPAND_CHUNK;
        const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(chunks, EXPAND_GRID));
        hipLaunchKernelGGL(k_emit_expand_dense, dim3(grid), dim3(EMIT_THREADS), 0, s, dval, n_splats, bpv, nviews,
                           vis_count, box, views, split, bcnt, vtot, pkey, pval, powner, pair_cap, counters);
    }
    return hipGetLastError();
}

}  // namespace

uint32_t emit_blocks(uint32_t n) { return n ? (n + EMIT_THREADS - 1) / EMIT_THREADS : 1; }

hipError_t launch_emit(hipStream_t s, int phase, const uint32_t *dval, uint32_t n_splats, uint32_t nviews,
                       const uint32_t *vis_count, const uint4 *rec, uint2 *box, const ViewParams *views,
                       EmitSplit split, const uint32_t *ip, uint32_t *bcnt, uint32_t *bcnt_all, uint32_t *vtot,
                       uint32_t *view_pairs, uint32_t *pkey, uint32_t *pval, uint32_t *powner, uint64_t pair_cap,
                       uint32_t *counters) {
    if (phase == EMIT_SECOND)
        return emit_phase<EMIT_SECOND>(s, dval, n_splats, nviews, vis_count, rec, box, views, split, ip, bcnt,
                                       bcnt_all, vtot, view_pairs, pkey, pval, powner, pair_cap, CNT_PAIRS2,
                                       counters);
    return emit_phase<EMIT_FIRST>(s, dval, n_splats, nviews, vis_count, rec, box, views, split, ip, bcnt, bcnt_all,
                                  vtot, view_pairs, pkey, pval, powner, pair_cap, CNT_PAIRS, counters);
}

hipError_t launch_row_prefix(hipStream_t s, const ViewParams *views, uint32_t nviews, const uint32_t *done,
                             uint32_t *ip) {
    if (nviews == 0) return hipSuccess;
    hipLaunchKernelGGL(k_row_prefix, dim3(16, nviews), dim3(256), 0, s, views, done, ip);
    return hipGetLastError();
}

hipError_t launch_ranges(hipStream_t s, const uint32_t *skey, const uint32_t *count, uint64_t pair_cap,
                         uint32_t num_tiles, uint32_t *tile_start) {
    const uint64_t work = pair_cap > num_tiles ? pair_cap : (uint64_t)num_tiles + 1;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((work + 255) / 256, 4096);  // grid-stride over the pairs
    hipLaunchKernelGGL(k_ranges, dim3(grid), dim3(256), 0, s, skey, count, pair_cap, num_tiles, tile_start);
    return hipGetLastError();
}

}  // namespace sp
