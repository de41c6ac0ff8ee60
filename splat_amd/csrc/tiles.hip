// (b) Tile-pair emission and (d) tile ranges.
//
// Emission walks each view's depth-sorted visible splats. A thread turns one
// splat's clamped pixel box into its 16x16 tile rectangle (the tiles
// [x0/16, ceil(x1/16)) x [y0/16, ceil(y1/16)) of src/render/cpu.rs:266-276; an
// empty box gives no pairs, :262-264). Offsets come from a reduce-then-scan:
// a count pass writes each block's pair totals, one scan workgroup per view
// turns them into block offsets (row totals = pairs per view), and a one-row
// scan turns the view totals into view bases.
// The first phase is binned (see "Binned first phase" below): a
// pair-balanced expansion writes one pair per (splat, 4x4-tile superblock)
// its rectangle meets, the superblock pairs are sorted stably by superblock,
// and per-superblock workgroups write each tile's list in (depth, scene
// index) order. The sparse second phase expands tile pairs with one
// workgroup per splat block, each carrying a global tile id (view tile base
// + tile) and a record index (view * n + scene index), ready for the stable
// tile-id sort.
// The differentiable path also numbers the tile pairs in emission order
// (each record's first slot, and the first slot and record of each
// depth-ordered splat), so that per-pair gradients can be written back in
// emission order, where each splat's pairs are contiguous (a pair's slot is
// its record's first slot plus the tile's row-major place in the record's
// rectangle).
//
// Front-to-back split. Almost every tile saturates (all pixels reach
// T < 1e-4) within the first few percent of its depth-ordered list, so the
// plain path emits in two phases: FIRST emits the pairs of each view's
// nearest K1 splats (K1 = max(kmin, vis / 2^shift), capped at vis) for every
// tile; the blend then marks the tiles it finished and sets the bits of the
// others in per-row open-tile masks. SECOND emits the pairs of the remaining
// splats for the unfinished tiles only, counted per splat with one masked
// popcount per row (and word) of its rectangle. A tile's SECOND list
// continues its FIRST list in the same (depth, index) order, so a pixel sees
// exactly the sequence of the one-phase path and the image bits are the same.
// FIRST also scans each view's unsplit pair count into view_pairs[v]
// (RenderStats.overlaps).
//
// Ranges: bucket bounds tile_start[t] (src/render/cpu.rs:150-163) from the
// tile-sorted keys (binary search per tile for long lists, key boundaries
// otherwise); empty tiles between two boundaries get the same start.
#include <algorithm>

#include "sp_device.h"
#include "sp_launch.h"
#include "sp_splat.h"

namespace sp {
namespace {

// Frames with at most this many rows of BIN_SCAN_ROW tiles (65536 tiles, a
// 4096x4096 frame) have k_bin_write sum the row totals itself instead of a
// scan launch.
constexpr uint32_t BIN_ROWS_SUMMED = 32;
// Superblock pairs per workgroup step of the expansion: the chunk follows the
// live count (512 to 4096 pairs), so few pairs still spread over many CUs.
constexpr uint32_t EXPAND_CHUNK_MAX = 4096;
constexpr uint32_t EXPAND_CHUNK_MIN = 512;
constexpr uint32_t EXPAND_GRID = 2048;   // workgroups of the expansion (grid-stride over chunks)

struct Rect {
    uint32_t cnt, org, rw, id;
    uint32_t tx0, ty0, tx1, ty1;
};

// Tile rectangle of the j-th depth-ordered splat of view v, from the
// depth-ordered pixel boxes (box[vbase + j] = the record's packed x0|y0, x1|y1);
// with WITH_ID, also its record index (the count passes do not need it).
// Tile rectangle of a packed pixel box (x0 | y0 << 16, x1 | y1 << 16).
__device__ __forceinline__ Rect box_rect(uint2 b, uint32_t tiles_x, uint32_t tile_base) {
    Rect r{0u, 0u, 1u, 0u, 0u, 0u, 0u, 0u};
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
    return r;
}

template <bool WITH_ID = true>
__device__ __forceinline__ Rect splat_rect(const uint32_t *__restrict__ dval, const uint2 *__restrict__ box,
                                           uint64_t vbase, uint32_t j, uint32_t nvis, uint32_t tiles_x,
                                           uint32_t tile_base) {
    if (j >= nvis) return Rect{0u, 0u, 1u, 0u, 0u, 0u, 0u, 0u};
    Rect r = box_rect(box[vbase + j], tiles_x, tile_base);
    if (WITH_ID) r.id = (uint32_t)(vbase + dval[vbase + j]);
    return r;
}

// Superblock rectangle of a tile rectangle: the superblocks (SB x SB tiles)
// it overlaps, as count, first global superblock id and width.
struct SbRect {
    uint32_t cnt, org, rw;
};
__device__ __forceinline__ SbRect sb_rect(const Rect &r, uint32_t sbx, uint32_t sb_base) {
    SbRect q{0u, 0u, 1u};
    if (r.cnt) {
        const uint32_t sx0 = r.tx0 / SB, sy0 = r.ty0 / SB;
        const uint32_t sx1 = (r.tx1 + SB - 1) / SB, sy1 = (r.ty1 + SB - 1) / SB;
        q.rw = sx1 - sx0;
        q.cnt = q.rw * (sy1 - sy0);
        q.org = sb_base + sy0 * sbx + sx0;
    }
    return q;
}

// Sub-rectangle of the tile rectangle [tx0, tx1) x [ty0, ty1) inside the
// superblock whose first tile is (bx, by), in 8 bits: first column, last
// column, first row, last row, 2 bits each (the rectangle meets the
// superblock).
__device__ __forceinline__ uint32_t sb_sub(uint32_t tx0, uint32_t ty0, uint32_t tx1, uint32_t ty1, uint32_t bx,
                                           uint32_t by) {
    const uint32_t i0 = tx0 > bx ? tx0 - bx : 0u, i1 = min(tx1 - bx, SB) - 1u;
    const uint32_t j0 = ty0 > by ? ty0 - by : 0u, j1 = min(ty1 - by, SB) - 1u;
    return i0 | (i1 << 2) | (j0 << 4) | (j1 << 6);
}

// The 16-bit tile mask (bit row * SB + column) of a sub-rectangle code.
__device__ __forceinline__ uint32_t sb_sub_mask(uint32_t sub) {
    const uint32_t i0 = sub & 3u, i1 = (sub >> 2) & 3u, j0 = (sub >> 4) & 3u, j1 = (sub >> 6) & 3u;
    const uint32_t row = ((2u << i1) - 1u) & ~((1u << i0) - 1u);
    uint32_t m = 0;
#pragma unroll
    for (uint32_t jj = 0; jj < SB; ++jj)
        if (jj >= j0 && jj <= j1) m |= row << (SB * jj);
    return m;
}

// Number of depth-ordered splats of a view the FIRST phase emits.
__device__ __forceinline__ uint32_t first_count(uint32_t vis, EmitSplit sp) {
    const uint32_t f = (uint32_t)(((uint64_t)vis + (1ull << sp.shift) - 1) >> sp.shift);
    return min(vis, max(sp.kmin, f));
}

// Unfinished tiles as bits: row ty of view v holds its tiles' bits in the
// words [tile_base + ty * W, tile_base + (ty + 1) * W), W = ceil(tiles_x / 32)
// (bit x & 31 of word x >> 5), set by the first blend for each tile it leaves
// open. A view needs at most its tile count of words.
__device__ __forceinline__ uint32_t mask_words(uint32_t tiles_x) { return (tiles_x + 31u) >> 5; }

// Unfinished tiles among columns [x0, x1) (x0 < x1) of one row's mask words.
__device__ __forceinline__ uint32_t row_open(const uint32_t *__restrict__ rm, uint32_t x0, uint32_t x1) {
    uint32_t c = 0;
    for (uint32_t w = x0 >> 5; w <= (x1 - 1u) >> 5; ++w) {
        uint32_t m = rm[w];
        const uint32_t lo = w << 5;
        if (x0 > lo) m &= ~0u << (x0 - lo);
        if (x1 < lo + 32u) m &= (1u << (x1 - lo)) - 1u;
        c += (uint32_t)__popc(m);
    }
    return c;
}

// Column of the rem-th (from 0) unfinished tile among columns [x0, x1) of a row
// that holds more than rem of them.
__device__ __forceinline__ uint32_t row_nth_open(const uint32_t *__restrict__ rm, uint32_t x0, uint32_t x1,
                                                 uint32_t rem) {
    for (uint32_t w = x0 >> 5; w <= (x1 - 1u) >> 5; ++w) {
        uint32_t m = rm[w];
        const uint32_t lo = w << 5;
        if (x0 > lo) m &= ~0u << (x0 - lo);
        if (x1 < lo + 32u) m &= (1u << (x1 - lo)) - 1u;
        const uint32_t c = (uint32_t)__popc(m);
        if (rem < c) {
            for (uint32_t i = 0; i < rem; ++i) m &= m - 1u;
            return lo + (uint32_t)__builtin_ctz(m);
        }
        rem -= c;
    }
    return x1 - 1u;  // not reached for a valid rem
}

// Unfinished tiles of the rectangle of r (SECOND phase).
__device__ __forceinline__ uint32_t open_in_rect(const uint32_t *__restrict__ openm, const Rect &r,
                                                 uint32_t tile_base, uint32_t tiles_x) {
    const uint32_t W = mask_words(tiles_x);
    uint32_t c = 0;
    for (uint32_t ty = r.ty0; ty < r.ty1; ++ty) c += row_open(openm + tile_base + ty * W, r.tx0, r.tx1);
    return c;
}

template <int PHASE>
__device__ __forceinline__ uint32_t phase_count(const Rect &r, uint32_t j, uint32_t k1,
                                                const uint32_t *__restrict__ openm, uint32_t tile_base,
                                                uint32_t tiles_x) {
    if (PHASE == EMIT_SECOND) return (j >= k1 && r.cnt) ? open_in_rect(openm, r, tile_base, tiles_x) : 0u;
    return j < k1 ? r.cnt : 0u;
}

// The colour of view v's record of scene splat i, written into the record
// (a stored scene's projection leaves it to the phase whose pairs first
// include the splat, so only splats that a blend reads are evaluated).
__device__ __forceinline__ void emit_colour(const DevScene &sc, const ViewCtx &vc, uint4 *rec, uint64_t rid,
                                            uint32_t i) {
    float col[3];
    row_colour(sc, i, vc.eye, col);
    float *rf = reinterpret_cast<float *>(rec + 3 * rid);
    rf[6] = col[0];  // record word 1: x, y, z = red, w = green
    rf[7] = col[1];
    rf[8] = col[2];  // record word 2: x = blue
}

template <int PHASE>
__global__ __launch_bounds__(EMIT_THREADS) void k_emit_count(const uint32_t *__restrict__ dval, uint32_t n_splats,
                                                             uint32_t bpv, const uint32_t *__restrict__ vis_count,
                                                             const uint2 *__restrict__ sbox, uint2 *__restrict__ box,
                                                             const ViewParams *__restrict__ views, EmitSplit split,
                                                             uint32_t *__restrict__ openm,
                                                             uint32_t *__restrict__ bcnt,
                                                             uint32_t *__restrict__ bcnt_all,
                                                             uint32_t *__restrict__ bsb,
                                                             uint32_t *__restrict__ vopen, DevScene sc,
                                                             uint4 *__restrict__ rec) {
    const uint32_t v = blockIdx.x / bpv, local = blockIdx.x % bpv;
    if (PHASE == EMIT_SECOND && vopen[v] == 0) {  // block-uniform: the first phase finished the view
        if (threadIdx.x == 0) bcnt[blockIdx.x] = 0;
        return;
    }
    if (PHASE == EMIT_FIRST && vopen) {
        // The first blend sets the bits of the tiles it leaves open and counts them.
        if (local == 0 && threadIdx.x == 0) vopen[v] = 0;
        const uint32_t nw = views[v].ctx.tiles_y * mask_words(views[v].ctx.tiles_x);
        for (uint32_t i = local * EMIT_THREADS + threadIdx.x; i < nw; i += bpv * EMIT_THREADS)
            openm[views[v].tile_base + i] = 0u;
    }
    const uint64_t vbase = (uint64_t)v * n_splats;
    const uint32_t nvis = min(vis_count[v], n_splats);  // dval holds n_splats entries per view
    const uint32_t k1 = first_count(nvis, split);
    const uint32_t tiles_x = views[v].ctx.tiles_x, tile_base = views[v].tile_base;
    const uint32_t j = local * EMIT_THREADS + threadIdx.x;
    Rect r;
    if (PHASE == EMIT_FIRST) {
        // The first pass over the depth order also gathers each splat's box
        // from the scene-order boxes the projection wrote (8 bytes per splat:
        // a view's array stays cache-resident, where the 48-byte records would
        // not), so that every later pass reads the boxes in depth order.
        r = Rect{0u, 0u, 1u, 0u, 0u, 0u, 0u, 0u};
        if (j < nvis) {
            const uint2 b = sbox[vbase + dval[vbase + j]];
            box[vbase + j] = b;
            r = box_rect(b, tiles_x, tile_base);
        }
    } else {
        r = splat_rect<false>(dval, box, vbase, j, nvis, tiles_x, tile_base);
    }
    const uint32_t c = phase_count<PHASE>(r, j, k1, openm, tile_base, tiles_x);
    // Colours of the splats whose pairs this phase emits (rec is null where
    // the projection wrote them).
    if (rec && j < nvis && (PHASE == EMIT_FIRST ? j < k1 : c > 0u)) {
        const uint32_t i = dval[vbase + j];
        emit_colour(sc, views[v].ctx, rec, vbase + i, i);
    }
    if (PHASE == EMIT_SECOND) {
        const uint32_t v1[1] = {c};
        uint32_t t1[1];
        block_totals<1>(v1, t1);
        if (threadIdx.x == 0) bcnt[blockIdx.x] = t1[0];
    } else {
        // Tile pairs of this phase, unsplit pairs (RenderStats.overlaps) and,
        // binned, superblock pairs of the phase's splats.
        const uint32_t sbc = bsb && j < k1 ? sb_rect(r, (tiles_x + SB - 1) / SB, views[v].sb_base).cnt : 0u;
        const uint32_t v3[3] = {c, r.cnt, sbc};
        uint32_t t3[3];
        block_totals<3>(v3, t3);
        if (threadIdx.x == 0) {
            bcnt[blockIdx.x] = t3[0];
            bcnt_all[blockIdx.x] = t3[1];
            if (bsb) bsb[blockIdx.x] = t3[2];
        }
    }
}

// Pair-balanced expansion of the first phase's superblock pairs: the
// phase's pairs (views concatenated) are cut into chunks of 512-4096 pairs
// dealt to the workgroups; a workgroup finds the splat block holding its
// chunk's first pair (binary search over the view bases and the view's block
// offsets), re-derives that block's 256 superblock rectangles and their
// in-block offsets, expands the part of the chunk they cover, and moves on to
// the next block until the chunk is done. The nearest splats of a view cover
// hundreds of superblocks each, so splat-major work would leave one workgroup
// with most of the pairs.
// A pair is key = global superblock id | sub-rectangle << SB_SUB_SHIFT (the
// part of the splat's tile rectangle inside the superblock, sb_sub), value =
// record index; the pair count comes from counters[CNT_SBPAIRS].
__global__ __launch_bounds__(EMIT_THREADS) void k_emit_expand_sb(const uint32_t *__restrict__ dval, uint32_t n_splats,
                                                              uint32_t bpv, uint32_t nviews,
                                                              const uint32_t *__restrict__ vis_count,
                                                              const uint2 *__restrict__ box,
                                                              const ViewParams *__restrict__ views, EmitSplit split,
                                                              const uint32_t *__restrict__ boff,
                                                              const uint32_t *__restrict__ vbase_pairs,
                                                              uint32_t *__restrict__ pkey, uint32_t *__restrict__ pval,
                                                              uint64_t pair_cap, uint32_t *__restrict__ counters) {
    __shared__ uint32_t s_pre[EMIT_THREADS];
    __shared__ uint32_t s_org[EMIT_THREADS];  // first global tile id of the rectangle
    __shared__ uint32_t s_rw[EMIT_THREADS];   // rectangle width in tiles
    __shared__ uint32_t s_id[EMIT_THREADS];   // record index (view * n + scene index)
    __shared__ uint32_t s_t0[EMIT_THREADS], s_t1[EMIT_THREADS];  // tile rectangle
    __shared__ uint32_t s_v, s_blk;
    const uint64_t np = counters[CNT_SBPAIRS];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // The tile pairs (which a binned list holds too) set the capacity need.
        counters[CNT_OVERFLOW] = counters[CNT_PAIRS] > pair_cap ? 1u : 0u;
        counters[CNT_STAGED] = 0u;  // the blends of this frame accumulate into it
    }
    const uint64_t nw = min(np, pair_cap);
    const uint32_t CHUNK = (uint32_t)min((uint64_t)EXPAND_CHUNK_MAX,
                                         max((uint64_t)EXPAND_CHUNK_MIN, ((nw / gridDim.x + 255) / 256) * 256));
    for (uint64_t c0 = (uint64_t)blockIdx.x * CHUNK; c0 < nw; c0 += (uint64_t)gridDim.x * CHUNK) {
        const uint64_t c1 = min(c0 + CHUNK, nw);
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
            const uint32_t gw = (tiles_x + SB - 1) / SB;  // superblock grid width
            const SbRect sq = sb_rect(r, gw, views[v].sb_base);
            const uint32_t cnt = j < k1 ? sq.cnt : 0u, org = sq.org, rw = sq.rw;
            uint32_t total;
            const uint32_t lscan = block_exclusive_scan<EMIT_THREADS>(cnt, &total);
            s_pre[threadIdx.x] = lscan;
            s_org[threadIdx.x] = org;
            s_rw[threadIdx.x] = rw;
            s_id[threadIdx.x] = r.id;
            s_t0[threadIdx.x] = r.tx0 | (r.ty0 << 16);
            s_t1[threadIdx.x] = r.tx1 | (r.ty1 << 16);
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
                const uint32_t tile = s_org[lo] + ry * gw + rx;
                const uint32_t t0 = s_t0[lo], t1 = s_t1[lo];
                const uint32_t bx = ((t0 & 0xFFFFu) / SB + rx) * SB, by = ((t0 >> 16) / SB + ry) * SB;
                pkey[qq] = tile | (sb_sub(t0 & 0xFFFFu, t0 >> 16, t1 & 0xFFFFu, t1 >> 16, bx, by) << SB_SUB_SHIFT);
                pval[qq] = s_id[lo];
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
                                                                   EmitSplit split,
                                                                   const uint32_t *__restrict__ openm,
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
    if (blockIdx.x == 0 && threadIdx.x == 0 && counters[CNT_PAIRS2] > pair_cap) counters[CNT_OVERFLOW] = 1u;
    // Blocks without pairs (most: the first phase finished their tiles) leave
    // before deriving any rectangle; the scanned offsets give the block total.
    {
        const uint32_t nv = gridDim.x / bpv;
        const uint32_t vend = v + 1 < nv ? vbase_pairs[v + 1] : counters[CNT_PAIRS2];
        const uint32_t bend = local + 1 < bpv ? boff[blockIdx.x + 1] : vend - vbase_pairs[v];
        if (bend == boff[blockIdx.x]) return;  // block-uniform
    }
    const uint64_t vbase = (uint64_t)v * n_splats;
    const uint32_t nvis = min(vis_count[v], n_splats);
    const uint32_t k1 = first_count(nvis, split);
    const uint32_t tiles_x = views[v].ctx.tiles_x, tile_base = views[v].tile_base;
    const uint32_t j = local * EMIT_THREADS + threadIdx.x;
    const Rect r = splat_rect(dval, box, vbase, j, nvis, tiles_x, tile_base);
    uint32_t total;
    const uint32_t lscan =
        block_exclusive_scan<EMIT_THREADS>(phase_count<EMIT_SECOND>(r, j, k1, openm, tile_base, tiles_x), &total);
    s_pre[threadIdx.x] = lscan;
    s_org[threadIdx.x] = r.org;
    s_rw[threadIdx.x] = r.rw;
    s_id[threadIdx.x] = r.id;
    s_rows[threadIdx.x] = r.ty1 - r.ty0;
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
        const uint32_t ty0 = org / tiles_x, tx0 = org % tiles_x, W = mask_words(tiles_x);
        uint32_t rem = loc, ty = ty0;
        for (uint32_t k = 0; k + 1 < s_rows[lo]; ++k, ++ty) {
            const uint32_t c = row_open(openm + tile_base + ty * W, tx0, tx0 + rwl);
            if (rem < c) break;
            rem -= c;
        }
        const uint32_t xl = row_nth_open(openm + tile_base + ty * W, tx0, tx0 + rwl, rem);
        const uint32_t rb = tile_base + ty * tiles_x;
        const uint64_t pos = base + q;
        if (pos < pair_cap) {
            pkey[pos] = rb + xl;
            pval[pos] = s_id[lo];
        }
    }
}

// ---- Binned first phase ----
// The pairs of the first phase are not sorted: each splat emits one pair per
// superblock (SB x SB tiles) its tile rectangle overlaps, those few pairs are
// sorted by superblock (stable, so each superblock's list stays in depth
// order), and one workgroup per superblock then writes the tile lists of its
// 16 tiles: each 256-entry chunk of its list gives every entry a 16-bit mask
// of the tiles it covers, one ballot per tile ranks the entries in list
// order, and the entries of one tile land at consecutive addresses. A count
// pass and a scan over the tiles place every list first, so the result is the
// tile-major (tile, depth, index) order of a full tile sort, with tile_start,
// written without moving the tile pairs through sort passes.

// The view of a batch that owns global superblock sb (sb_base ascends).
__device__ __forceinline__ uint32_t view_of_sb(const ViewParams *views, uint32_t nviews, uint32_t sb) {
    uint32_t lo = 0, hi = nviews;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (views[mid].sb_base <= sb)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

struct SbGeom {
    uint32_t v, bx, by, tiles_x, tiles_y, tile_base;
};
__device__ __forceinline__ SbGeom sb_geom(const ViewParams *views, uint32_t nviews, uint32_t sb) {
    SbGeom g;
    g.v = view_of_sb(views, nviews, sb);
    const ViewParams &vp = views[g.v];
    g.tiles_x = vp.ctx.tiles_x;
    g.tiles_y = vp.ctx.tiles_y;
    g.tile_base = vp.tile_base;
    const uint32_t sbx = (g.tiles_x + SB - 1) / SB, local = sb - vp.sb_base;
    g.bx = (local % sbx) * SB;
    g.by = (local / sbx) * SB;
    return g;
}

// A superblock's list is cut into BIN_SEGS segments (each a whole number of
// 256-entry chunks, at least one), one workgroup per (superblock, segment),
// so that long lists (the centre of a view) spread over several CUs.
__device__ __forceinline__ void bin_segment(uint32_t e0, uint32_t e1, uint32_t y, uint32_t &s0, uint32_t &s1) {
    const uint32_t len = e1 - e0;
    const uint32_t per = max(256u, (((len + BIN_SEGS - 1) / BIN_SEGS) + 255u) & ~255u);
    s0 = min(e1, e0 + y * per);
    s1 = min(e1, s0 + per);
}

// segc[(sb * BIN_SEGS + y) * 16 + b] = entries of segment y of superblock sb
// that cover its tile b (from the sub-rectangles in the sorted keys).
__global__ __launch_bounds__(256) void k_bin_count(const uint32_t *__restrict__ sb_start,
                                                   const uint32_t *__restrict__ skey, uint32_t *__restrict__ segc) {
    __shared__ uint32_t s_c[4][SB * SB];
    const uint32_t sb = blockIdx.x, y = blockIdx.y, w = threadIdx.x >> 6, lane = lane_id();
    uint32_t s0, s1;
    bin_segment(sb_start[sb], sb_start[sb + 1], y, s0, s1);
    uint32_t mine = 0;  // lane b < 16: this wave's entries covering tile b
    if (s0 < s1) {
        uint32_t key_n = s0 + threadIdx.x < s1 ? skey[s0 + threadIdx.x] : 0u;
        for (uint32_t c = s0; c < s1; c += 256) {  // block-uniform
            const uint32_t e = c + threadIdx.x;
            const uint32_t m = e < s1 ? sb_sub_mask(key_n >> SB_SUB_SHIFT) : 0u;
            if (e + 256 < s1) key_n = skey[e + 256];  // next chunk
#pragma unroll
            for (uint32_t b = 0; b < SB * SB; ++b) {
                const uint32_t k = (uint32_t)__popcll(__ballot((m >> b) & 1u));
                if (lane == b) mine += k;
            }
        }
    }
    if (lane < SB * SB) s_c[w][lane] = mine;
    __syncthreads();
    if (threadIdx.x < SB * SB)
        segc[((uint64_t)sb * BIN_SEGS + y) * (SB * SB) + threadIdx.x] =
            s_c[0][threadIdx.x] + s_c[1][threadIdx.x] + s_c[2][threadIdx.x] + s_c[3][threadIdx.x];
}

// One thread per (superblock, tile): segc becomes the exclusive prefix over
// the superblock's segments, tcount[gt] the tile's list length and
// sbtile[sb * 16 + b] the global tile id (~0 past the frame's edge); the
// padding tcount[total_tiles, tcount_len) is cleared for the row scan.
__global__ __launch_bounds__(256) void k_bin_prefix(const ViewParams *__restrict__ views, uint32_t nviews,
                                                    uint32_t total_sbs, uint32_t *__restrict__ segc,
                                                    uint32_t *__restrict__ tcount, uint32_t *__restrict__ sbtile,
                                                    uint32_t total_tiles, uint32_t tcount_len) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < (uint64_t)total_sbs * (SB * SB)) {
        const uint32_t sb = (uint32_t)(g / (SB * SB)), b = (uint32_t)(g % (SB * SB));
        uint32_t *c = segc + (uint64_t)sb * BIN_SEGS * (SB * SB) + b;
        uint32_t acc = 0;
        for (uint32_t y = 0; y < BIN_SEGS; ++y) {
            const uint32_t k = c[y * (SB * SB)];
            c[y * (SB * SB)] = acc;
            acc += k;
        }
        const SbGeom q = sb_geom(views, nviews, sb);
        const uint32_t tx = q.bx + b % SB, ty = q.by + b / SB;
        uint32_t gt = ~0u;
        if (tx < q.tiles_x && ty < q.tiles_y) {
            gt = q.tile_base + ty * q.tiles_x + tx;
            tcount[gt] = acc;
        }
        sbtile[g] = gt;
    }
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t t = total_tiles + g; t < tcount_len; t += stride) tcount[t] = 0;
}

// Writes tile_start for the superblock's tiles (segment 0; clamped to the
// capacity: an over-capacity frame is rendered again with larger buffers) and
// the segment's entries into the tile lists: pval = record index, pkey (when
// given) = tile id. tscan: in-row exclusive prefix of tcount (rows of
// BIN_SCAN_ROW tiles); trow: the row totals, or with prefixed_rows their
// exclusive prefix (frames of more than BIN_ROWS_SUMMED rows scan them in a
// launch of their own; fewer are summed here); segc: the segment offsets
// within each tile's list.
__global__ __launch_bounds__(256) void k_bin_write(const uint32_t *__restrict__ sbtile,
                                                   const uint32_t *__restrict__ sb_start,
                                                   const uint32_t *__restrict__ skey,
                                                   const uint32_t *__restrict__ sval,
                                                   const uint32_t *__restrict__ segc,
                                                   const uint32_t *__restrict__ tscan,
                                                   const uint32_t *__restrict__ trow, uint32_t prefixed_rows,
                                                   uint32_t total_tiles,
                                                   uint32_t *__restrict__ tile_start, uint32_t *__restrict__ pkey,
                                                   uint32_t *__restrict__ pval, uint64_t pair_cap,
                                                   const uint32_t *__restrict__ counters) {
    __shared__ uint32_t s_next[SB * SB];     // next list position of each tile
    __shared__ uint32_t s_gt[SB * SB];       // global tile id of each tile (~0: past the frame's edge)
    __shared__ uint32_t s_wc[4][SB * SB];    // this chunk's entries per (wave, tile)
    __shared__ uint32_t s_base[4][SB * SB];  // first position per (wave, tile) in this chunk
    __shared__ uint32_t s_tot[SB * SB];      // this chunk's entries per tile
    const uint32_t sb = blockIdx.x, y = blockIdx.y, w = threadIdx.x >> 6, lane = lane_id();
    uint32_t s0, s1;
    bin_segment(sb_start[sb], sb_start[sb + 1], y, s0, s1);
    if (sb == 0 && y == 0 && threadIdx.x == 0)
        tile_start[total_tiles] = (uint32_t)min((uint64_t)counters[CNT_PAIRS], pair_cap);
    if (s0 >= s1 && y > 0) return;  // block-uniform; segment 0 always writes the tile starts
    // The first chunk's entries load while the tile starts are set up.
    uint32_t key_n = 0, rid_n = 0;
    if (s0 + threadIdx.x < s1) {
        key_n = skey[s0 + threadIdx.x];
        rid_n = sval[s0 + threadIdx.x];
    }
    if (threadIdx.x < SB * SB) {
        const uint32_t gt = sbtile[(uint64_t)sb * (SB * SB) + threadIdx.x];
        uint32_t start = 0;
        if (gt != ~0u) {
            const uint32_t row = gt / BIN_SCAN_ROW;
            if (prefixed_rows) {
                start = trow[row];
            } else {
                for (uint32_t r = 0; r < row; ++r) start += trow[r];
            }
            start += tscan[gt];
            if (y == 0) tile_start[gt] = (uint32_t)min((uint64_t)start, pair_cap);
        }
        s_gt[threadIdx.x] = gt;
        s_next[threadIdx.x] = start + segc[((uint64_t)sb * BIN_SEGS + y) * (SB * SB) + threadIdx.x];
    }
    for (uint32_t c = s0; c < s1; c += 256) {  // block-uniform
        const uint32_t e = c + threadIdx.x;
        const uint32_t m = e < s1 ? sb_sub_mask(key_n >> SB_SUB_SHIFT) : 0u, rid = rid_n;
        if (e + 256 < s1) {  // next chunk
            key_n = skey[e + 256];
            rid_n = sval[e + 256];
        }
        unsigned long long bal[SB * SB];
        uint32_t cnt = 0;  // lane b < 16: this wave's entries covering tile b
#pragma unroll
        for (uint32_t b = 0; b < SB * SB; ++b) {
            bal[b] = __ballot((m >> b) & 1u);
            cnt = lane == b ? (uint32_t)__popcll(bal[b]) : cnt;
        }
        if (lane < SB * SB) s_wc[w][lane] = cnt;
        __syncthreads();  // s_wc complete; s_next holds this chunk's starts
        if (threadIdx.x < 4 * SB * SB) {
            const uint32_t ww = threadIdx.x / (SB * SB), b = threadIdx.x % (SB * SB);
            uint32_t base = s_next[b];
            for (uint32_t k = 0; k < ww; ++k) base += s_wc[k][b];
            s_base[ww][b] = base;
            if (ww == 0) s_tot[b] = s_wc[0][b] + s_wc[1][b] + s_wc[2][b] + s_wc[3][b];
        }
        __syncthreads();  // s_base complete
        const uint32_t basev = s_base[w][lane % (SB * SB)];
        if (threadIdx.x < SB * SB) s_next[threadIdx.x] += s_tot[threadIdx.x];  // read again after the next first barrier
#pragma unroll
        for (uint32_t b = 0; b < SB * SB; ++b) {
            const uint32_t base_b = __builtin_amdgcn_readlane(basev, b);
            if ((m >> b) & 1u) {
                const uint32_t pos = base_b + __builtin_amdgcn_mbcnt_hi((uint32_t)(bal[b] >> 32),
                                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)bal[b], 0u));
                if (pos < pair_cap) {
                    pval[pos] = rid;
                    if (pkey) pkey[pos] = s_gt[b];
                }
            }
        }
    }
}

// (d) Tile ranges: tile_start[t] = the first sorted position whose tile id is
// >= t (tile_start[num_tiles] = the pair count), so tile t's list is
// [tile_start[t], tile_start[t + 1]).
// Long sorted lists (more than RANGE_SEARCH_MIN pairs) are indexed by one
// binary search per tile, whose cost follows the tile count. Shorter ones by
// one thread per pair that finds a change of tile id and writes the run of
// tiles that starts there; runs longer than RANGE_SHORT (empty tiles, common
// in the continuation lists) go to an LDS queue that the whole workgroup
// writes, so a single thread never walks thousands of tiles.
constexpr uint64_t RANGE_SEARCH_MIN = 8ull << 20;
constexpr uint32_t RANGE_SHORT = 16;

__global__ __launch_bounds__(256) void k_ranges(const uint32_t *__restrict__ skey, uint32_t key_mask,
                                                const uint32_t *__restrict__ count, uint64_t pair_cap,
                                                uint32_t num_tiles, uint32_t *__restrict__ tile_start) {
    __shared__ uint32_t s_lo[260], s_hi[260], s_val[260];  // queued long runs: at most one per thread, plus the end
    __shared__ uint32_t s_n;
    const uint64_t np = min((uint64_t)*count, pair_cap);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t last_tile = num_tiles - 1;  // keys are tile ids < num_tiles
    if (np == 0) {
        for (uint64_t t = g; t <= num_tiles; t += stride) tile_start[t] = 0;
        return;
    }
    if (np > RANGE_SEARCH_MIN) {
        for (uint64_t t = g; t <= num_tiles; t += stride) {
            uint64_t lo = 0, hi = np;
            while (lo < hi) {
                const uint64_t mid = (lo + hi) >> 1;
                if (min(skey[mid] & key_mask, last_tile) < t)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            tile_start[t] = (uint32_t)lo;
        }
        return;
    }
    auto put_run = [&](uint32_t lo, uint32_t hi, uint32_t val) {  // tiles [lo, hi) start at val
        if (hi <= lo) return;
        if (hi - lo <= RANGE_SHORT) {
            for (uint32_t t = lo; t < hi; ++t) tile_start[t] = val;
        } else {
            const uint32_t j = atomicAdd(&s_n, 1u);
            s_lo[j] = lo;
            s_hi[j] = hi;
            s_val[j] = val;
        }
    };
    // Block-uniform loop: every thread reaches the barriers.
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < np; base += stride) {
        if (threadIdx.x == 0) s_n = 0;
        __syncthreads();
        const uint64_t p = base + threadIdx.x;
        if (p < np) {
            const uint32_t k = min(skey[p] & key_mask, last_tile);
            const uint32_t lo = p == 0 ? 0u : min(skey[p - 1] & key_mask, last_tile) + 1u;
            put_run(lo, k + 1u, (uint32_t)p);
            if (p == np - 1) put_run(k + 1u, num_tiles + 1u, (uint32_t)np);
        }
        __syncthreads();
        const uint32_t nr = s_n;
        for (uint32_t i = 0; i < nr; ++i) {
            const uint32_t hi = s_hi[i], val = s_val[i];
            for (uint32_t t = s_lo[i] + threadIdx.x; t < hi; t += blockDim.x) tile_start[t] = val;
        }
        __syncthreads();
    }
}

// Differentiable path: the emission numbering of the first phase's tile
// pairs (view by view, splat by splat in depth order, each rectangle
// row-major), one thread per depth-ordered splat: pbase[record] = the
// record's first slot, pslot[v * (n + 1) + j] = the first slot of view v's
// j-th depth-ordered splat (for j < its visible count, plus the view's end
// slot after the last one) and prid[v * n + j] = that splat's record. The
// slots come from the same block offsets (boff) and view bases (vbase_pairs)
// as the emission's; they are clamped to pair_cap (the frame is then
// rendered again with more room).
__global__ __launch_bounds__(EMIT_THREADS) void k_emit_slots(const uint32_t *__restrict__ dval, uint32_t n_splats,
                                                       uint32_t bpv, const uint32_t *__restrict__ vis_count,
                                                       const uint2 *__restrict__ box,
                                                       const ViewParams *__restrict__ views, EmitSplit split,
                                                       const uint32_t *__restrict__ boff,
                                                       const uint32_t *__restrict__ vbase_pairs,
                                                       uint32_t *__restrict__ pbase, uint32_t *__restrict__ pslot,
                                                       uint32_t *__restrict__ prid, uint64_t pair_cap) {
    const uint32_t v = blockIdx.x / bpv, blk = blockIdx.x % bpv;
    const uint64_t vbase = (uint64_t)v * n_splats;
    const uint32_t nvis = min(vis_count[v], n_splats);
    if (blk * EMIT_THREADS >= nvis) return;  // block-uniform, before the scan's barriers
    const uint32_t j = blk * EMIT_THREADS + threadIdx.x;
    const uint32_t tiles_x = views[v].ctx.tiles_x, tile_base = views[v].tile_base;
    const Rect r = splat_rect(dval, box, vbase, j, nvis, tiles_x, tile_base);
    const uint32_t c = phase_count<EMIT_FIRST>(r, j, first_count(nvis, split), nullptr, tile_base, tiles_x);
    uint32_t total;
    const uint32_t pre = block_exclusive_scan<EMIT_THREADS>(c, &total);
    // Slots are u32 (list entries are); clamped to the capacity.
    const uint64_t lim = min(pair_cap, (uint64_t)0xFFFFFFFFu);
    const uint64_t base = (uint64_t)vbase_pairs[v] + boff[blockIdx.x] + pre;
    if (j >= nvis) return;
    const uint32_t b = (uint32_t)min(base, lim);
    if (c > 0) pbase[r.id] = b;
    uint32_t *ps = pslot + (uint64_t)v * (n_splats + 1u);
    ps[j] = b;
    if (j == nvis - 1u) ps[nvis] = (uint32_t)min(base + c, lim);
    prid[vbase + j] = r.id;
}

template <int PHASE>
hipError_t emit_phase(hipStream_t s, const uint32_t *dval, uint32_t n_splats, uint32_t nviews,
                      const uint32_t *vis_count, const uint2 *sbox, uint2 *box, const ViewParams *views,
                      EmitSplit split,
                      uint32_t *openm, uint32_t *bcnt, uint32_t *bcnt_all, uint32_t *vtot, uint32_t *view_pairs,
                      uint32_t *bsb, uint32_t *vsb, uint32_t *vopen, uint32_t *pkey, uint32_t *pval,
                      const PairSlots *slots, uint64_t pair_cap, uint32_t count_slot, uint32_t *counters,
                      const DevScene &sc, uint4 *colour_rec) {
    const uint32_t bpv = emit_blocks(n_splats);
    const bool bin = PHASE == EMIT_FIRST && bsb != nullptr;
    hipLaunchKernelGGL(k_emit_count<PHASE>, dim3(bpv * nviews), dim3(EMIT_THREADS), 0, s, dval, n_splats, bpv,
                       vis_count, sbox, box, views, split, openm, bcnt, bcnt_all, bin ? bsb : nullptr, vopen, sc,
                       colour_rec);
    hipError_t e = hipGetLastError();
    // Per view: block offsets and the view's pair total (and, in FIRST, the
    // unsplit pair count of every view, RenderStats.overlaps, and the binned
    // superblock pairs); then the view bases and the totals.
    uint32_t *rows_data[3] = {bcnt, bcnt_all, bsb}, *rows_tot[3] = {vtot, view_pairs, vsb};
    const int nrow_arrays = PHASE == EMIT_SECOND ? 1 : (bin ? 3 : 2);
    if (nviews == 1) {
        // One view: its base is 0 and its totals are the frame's, written by the same launch.
        uint32_t *single[3] = {counters + count_slot, nullptr, counters + CNT_SBPAIRS};
        if (e == hipSuccess) e = launch_scan_rows_multi(s, nrow_arrays, rows_data, rows_tot, bpv, 1, single);
    } else {
        if (e == hipSuccess) e = launch_scan_rows_multi(s, nrow_arrays, rows_data, rows_tot, bpv, nviews);
        uint32_t *view_data[2] = {vtot, vsb}, *view_tot[2] = {counters + count_slot, counters + CNT_SBPAIRS};
        if (e == hipSuccess) e = launch_scan_rows_multi(s, bin ? 2 : 1, view_data, view_tot, nviews, 1);
    }
    if (e != hipSuccess) return e;
    if (bin) {
        const uint64_t chunks = (pair_cap + EXPAND_CHUNK_MIN - 1) / EXPAND_CHUNK_MIN;
        const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(chunks, EXPAND_GRID));
        hipLaunchKernelGGL(k_emit_expand_sb, dim3(grid), dim3(EMIT_THREADS), 0, s, dval, n_splats, bpv,
                           nviews, vis_count, box, views, split, bsb, vsb, pkey, pval, pair_cap, counters);
        if (slots) {  // the differentiable path also numbers the tile pairs in emission order
            const hipError_t eb = hipGetLastError();
            if (eb != hipSuccess) return eb;
            hipLaunchKernelGGL(k_emit_slots, dim3(bpv * nviews), dim3(EMIT_THREADS), 0, s, dval, n_splats, bpv,
                               vis_count, box, views, split, bcnt, vtot, slots->pbase, slots->pslot, slots->prid,
                               pair_cap);
        }
    } else if (PHASE == EMIT_SECOND) {
        hipLaunchKernelGGL(k_emit_expand_open, dim3(bpv * nviews), dim3(EMIT_THREADS), 0, s, dval, n_splats, bpv,
                           vis_count, box, views, split, openm, bcnt, vtot, pkey, pval, pair_cap, counters);
    }
    return hipGetLastError();
}

}  // namespace

uint32_t emit_blocks(uint32_t n) { return n ? (n + EMIT_THREADS - 1) / EMIT_THREADS : 1; }

hipError_t launch_emit(hipStream_t s, int phase, const uint32_t *dval, uint32_t n_splats, uint32_t nviews,
                       const uint32_t *vis_count, const uint2 *sbox, uint2 *box, const ViewParams *views,
                       EmitSplit split, uint32_t *openm, uint32_t *vopen, uint32_t *bcnt,
                       uint32_t *bcnt_all, uint32_t *vtot, uint32_t *view_pairs, uint32_t *bsb, uint32_t *vsb,
                       uint32_t *pkey, uint32_t *pval, const PairSlots *slots, uint64_t pair_cap,
                       uint32_t *counters, const DevScene &sc, uint4 *colour_rec) {
    if (phase == EMIT_SECOND)
        return emit_phase<EMIT_SECOND>(s, dval, n_splats, nviews, vis_count, sbox, box, views, split, openm, bcnt,
                                       bcnt_all, vtot, view_pairs, nullptr, nullptr, vopen, pkey, pval, nullptr,
                                       pair_cap, CNT_PAIRS2, counters, sc, colour_rec);
    if (!bsb || !vsb) return hipErrorInvalidValue;  // the first phase is always binned
    return emit_phase<EMIT_FIRST>(s, dval, n_splats, nviews, vis_count, sbox, box, views, split, openm, bcnt, bcnt_all,
                                  vtot, view_pairs, bsb, vsb, vopen, pkey, pval, slots, pair_cap, CNT_PAIRS,
                                  counters, sc, colour_rec);
}

hipError_t launch_bin_lists(hipStream_t s, const ViewParams *views, uint32_t nviews, uint32_t total_sbs,
                            uint32_t total_tiles, const uint32_t *sb_start, const uint32_t *skey,
                            const uint32_t *sval, uint32_t *segc, uint32_t *sbtile, uint32_t *tcount, uint32_t *trow,
                            uint32_t *tile_start, uint32_t *pkey, uint32_t *pval, uint64_t pair_cap,
                            const uint32_t *counters) {
    if (total_tiles == 0 || total_sbs == 0) return launch_fill(s, tile_start, 1, 0u);
    const uint32_t rows = bin_scan_rows(total_tiles);
    const dim3 grid(total_sbs, BIN_SEGS);
    hipLaunchKernelGGL(k_bin_count, grid, dim3(256), 0, s, sb_start, skey, segc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint64_t pthreads = std::max<uint64_t>((uint64_t)total_sbs * SB * SB, (uint64_t)rows * BIN_SCAN_ROW - total_tiles);
    const uint32_t pgrid = (uint32_t)std::min<uint64_t>((pthreads + 255) / 256, 65535);
    hipLaunchKernelGGL(k_bin_prefix, dim3(pgrid), dim3(256), 0, s, views, nviews, total_sbs, segc, tcount, sbtile,
                       total_tiles, rows * BIN_SCAN_ROW);
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_scan_rows(s, tcount, BIN_SCAN_ROW, rows, nullptr, trow);
    const uint32_t prefixed = rows > BIN_ROWS_SUMMED ? 1u : 0u;
    if (e == hipSuccess && prefixed) e = launch_scan_rows(s, trow, rows, 1, nullptr, nullptr);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_bin_write, grid, dim3(256), 0, s, sbtile, sb_start, skey, sval, segc, tcount, trow,
                       prefixed, total_tiles, tile_start, pkey, pval, pair_cap, counters);
    return hipGetLastError();
}

uint32_t bin_scan_rows(uint32_t total_tiles) { return (total_tiles + BIN_SCAN_ROW - 1) / BIN_SCAN_ROW; }

hipError_t launch_ranges(hipStream_t s, const uint32_t *skey, uint32_t key_mask, const uint32_t *count,
                         uint64_t pair_cap, uint32_t num_tiles, uint32_t *tile_start) {
    const uint64_t work = pair_cap > num_tiles ? pair_cap : (uint64_t)num_tiles + 1;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((work + 255) / 256, 4096);  // grid-stride over the pairs
    hipLaunchKernelGGL(k_ranges, dim3(grid), dim3(256), 0, s, skey, key_mask, count, pair_cap, num_tiles,
                       tile_start);
    return hipGetLastError();
}

}  // namespace sp
