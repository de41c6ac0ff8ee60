// Host-side launchers of the rasterizer kernels (one translation unit each).
// Every per-view array is laid out view-major with a stride of n (the scene
// size): view v's entry for scene index i is at v * n + i.
#pragma once

#include <hip/hip_runtime.h>

#include "sp_common.h"

namespace sp {

// ---- scan.hip: reduce-then-scan building blocks ----
// In-place exclusive scan of `rows` rows of row_len u32 (one workgroup per
// row); row_base[r] (optional) is added; row_total[r] (optional) gets the sum.
hipError_t launch_scan_rows(hipStream_t s, uint32_t *data, uint64_t row_len, uint32_t rows,
                            const uint32_t *row_base, uint32_t *row_total);
// The same for up to 3 arrays of `rows` rows each, in one launch. With
// single_total (rows == 1 only), array k's total is written to
// *single_total[k] (where non-null) and its row total entry set to 0.
hipError_t launch_scan_rows_multi(hipStream_t s, int narrays, uint32_t *const *data, uint32_t *const *row_total,
                                  uint64_t row_len, uint32_t rows, uint32_t *const *single_total = nullptr);
// Radix tables: per segment g, row_len columns of bins digit counts each
// (entry (g, c, d) at data[(g * row_len + c) * bins + d]); for every (g, d)
// only the live columns (radix_columns(min(seg_counts[g], n_cap), tile_keys),
// n_cap when seg_counts is null) are scanned across the columns, in place and
// from 0; row_total[g * bins + d] receives the sum.
hipError_t launch_scan_digit_rows(hipStream_t s, uint32_t *data, uint64_t row_len, uint32_t nseg, uint32_t bins,
                                  const uint32_t *seg_counts, uint64_t n_cap, uint32_t tile_keys,
                                  uint32_t *row_total);
// p[0..n) = value (a kernel, so it may be captured in a graph).
hipError_t launch_fill(hipStream_t s, uint32_t *p, uint64_t n, uint32_t value);

// ---- project.hip: (a) projection + SH colour ----
uint32_t project_blocks(uint32_t n);
// Projects every splat into each of `nviews` views: records rec[3*(v*n+i)..]
// for visible splats (with their colours when `colour` is set or the scene is
// STORE_PARAMS; otherwise the colours stay zero, for the emission to fill in
// those a blend reads, except in the debug output, which always gets them), their pixel boxes sbox[v*n + i], and full[v*n + i] =
// the depth bits less the view's key_base, or FULL_CULLED for a culled splat
// (the depth sort's sparse first pass reads it directly). A key that does not
// fit DEPTH_KEY_BITS under a nonzero key_base sets counters[CNT_DEPTH_RANGE]
// (counters may be null). dbg / dbg_bbox / dbg_valid (optional, n entries)
// get view 0's full per-splat projection.
hipError_t launch_project(hipStream_t s, const DevScene &sc, const ViewParams *views, uint32_t nviews, uint4 *rec,
                          uint2 *sbox, uint32_t *full, float *dbg, uint16_t *dbg_bbox, uint8_t *dbg_valid,
                          uint32_t *counters, bool colour);

// ---- prep.hip: scene layout on the device ----
// The caller's arrays of structures, on the device: positions [n][3], cov3d
// [n][6] (f32 storage) or scales [n][3] and unit quaternions [n][4] (16-bit
// storage), opacities [n], k [n], base colours [n][3], SH rest [n][rest][3].
struct PrepInput {
    const float *pos, *cov, *scl, *quat, *opac, *k, *col, *sh;
};
// Writes the scene's blocked columns (scene_at) in `storage`'s layout: soa
// (f32 columns) and, for 16-bit storage, soa16; and its colour rows (crow,
// or crow16 for 16-bit storage, colour_row entries per splat). Sets *extreme
// (cleared by the caller) when a base colour or SH coefficient is NaN,
// infinite or too large for a colour sum to stay finite: at least 1e36 (f32,
// BF16), or at least 65504 (FP16, where it may round to infinity). Otherwise
// |colour| <= 1e36 * (1 + 15 * 1.1) stays below FLT_MAX at every direction.
hipError_t launch_prep(hipStream_t s, int storage, const PrepInput &in, uint32_t n, uint32_t rest, float *soa,
                       uint16_t *soa16, float *crow, uint16_t *crow16, uint32_t *extreme);

// ---- radix.hip: (c) stable LSD radix sort ----
// Words of the per-block digit-count table for nseg segments of up to n_cap
// keys; radix_table_bound(max_seg) bounds it for every small sort (segments
// holding at most RADIX_SMALL_MAX keys together) of up to max_seg segments.
uint64_t radix_table_words(uint32_t nseg, uint64_t n_cap);
uint64_t radix_table_bound(uint32_t max_seg);
// Pass plan of a `bits`-bit sort: digit widths (8, or 9 where that saves a
// pass) into widths[], returns the number of passes.
int radix_plan(int bits, int *widths);
// Sorts u32 keys with u32 values by bits [0, bits) (stable; higher key bits
// travel with the key without ordering it), 8- or 9-bit digits, over
// `nseg` independent segments: segment g holds min(seg_counts[g], n_cap) keys
// at offset g * seg_stride (device counts; n_cap sizes the grid). Ping-pongs
// between keys[0]/keys[1]; *out_index names the buffer with the result.
// row_total holds nseg * RADIX_BINS_MAX u32, table radix_table_words(nseg, n_cap) u32.
// sparse_first: every segment holds n_cap entries in keys[0], of which those
// equal to FULL_CULLED are absent; the first pass sorts the others with values
// made from their segment positions (vals[0] is not read) and writes each
// segment's count of them into seg_counts, which the later passes read.
hipError_t launch_radix_sort(hipStream_t s, uint32_t *keys[2], uint32_t *vals[2], uint32_t *seg_counts,
                             uint32_t nseg, uint64_t seg_stride, uint64_t n_cap, int bits, uint32_t *row_total,
                             uint32_t *table, int *out_index, bool sparse_first = false);

// ---- tiles.hip: (b) emission, (d) ranges ----
uint32_t emit_blocks(uint32_t n);
// One (global tile id, record index) pair per overlapped 16x16 tile of each
// view's depth-ordered splats, views in order, for emission phase `phase`
// (EMIT_FIRST: the nearest K1 splats of each view, every tile; EMIT_SECOND:
// the rest, only the tiles whose bits are set in openm, the per-row open-tile
// masks of the first blend; vopen[v] = 0 skips view v). FIRST clears openm and
// vopen (when given) for that blend.
// bcnt, bcnt_all: nviews rows of emit_blocks(n) block counts; vtot: nviews
// entries. Sets counters[CNT_PAIRS] (FIRST) or counters[CNT_PAIRS2]
// (SECOND), saturated at 0xFFFFFFFF, and counters[CNT_OVERFLOW] when above pair_cap. FIRST
// also writes each view's unsplit pair count to view_pairs[v].
// box: n entries per view, gathered by FIRST from the scene-order boxes sbox (the
// depth-ordered pixel boxes) and read by SECOND.
// FIRST is binned (bsb / vsb required, shaped like bcnt / vtot): its pairs are
// (global superblock id | sub-rectangle << SB_SUB_SHIFT, record index)
// instead, counted in counters[CNT_SBPAIRS]; counters[CNT_PAIRS] still counts
// the tile pairs. With slots (FIRST only), the emission slots of the tile
// pairs too, numbered view by view, splat by splat in depth order, each
// splat's rectangle row-major (see PairSlots). With colour_rec, the phase
// also writes the colour of each splat whose pairs it emits into its record
// (sc's colour rows; FIRST: the phase's splats, SECOND: those with pairs in
// unfinished tiles), which a stored scene's projection leaves out.
struct PairSlots {
    uint32_t *pbase;  // views x n: a record's first slot (records with pairs)
    uint32_t *pslot;  // views x (n + 1): view v's j-th depth-ordered splat's first slot at v * (n + 1) + j,
                      // for j below the view's visible count, and the view's end slot after the last
    uint32_t *prid;   // views x n: view v's j-th depth-ordered splat's record at v * n + j
};
hipError_t launch_emit(hipStream_t s, int phase, const uint32_t *dval, uint32_t n_splats, uint32_t nviews,
                       const uint32_t *vis_count, const uint2 *sbox, uint2 *box, const ViewParams *views,
                       EmitSplit split, uint32_t *openm, uint32_t *vopen, uint32_t *bcnt,
                       uint32_t *bcnt_all, uint32_t *vtot, uint32_t *view_pairs, uint32_t *bsb, uint32_t *vsb,
                       uint32_t *pkey, uint32_t *pval, const PairSlots *slots, uint64_t pair_cap,
                       uint32_t *counters, const DevScene &sc, uint4 *colour_rec);
// Rows of the binned tile-count scan: tcount holds bin_scan_rows * BIN_SCAN_ROW
// entries, trow bin_scan_rows.
uint32_t bin_scan_rows(uint32_t total_tiles);
// Binned FIRST lists: from the superblock-sorted pairs (skey = superblock id
// | sub-rectangle << SB_SUB_SHIFT, sval = record index; ranges
// sb_start[0..total_sbs]), the tile lists in tile order (pval = record index,
// pkey = tile id when pkey is given) and tile_start[0..total_tiles]. segc:
// total_sbs * BIN_SEGS * SB * SB scratch counts; sbtile: total_sbs * SB * SB.
hipError_t launch_bin_lists(hipStream_t s, const ViewParams *views, uint32_t nviews, uint32_t total_sbs,
                            uint32_t total_tiles, const uint32_t *sb_start, const uint32_t *skey,
                            const uint32_t *sval, uint32_t *segc, uint32_t *sbtile, uint32_t *tcount, uint32_t *trow,
                            uint32_t *tile_start, uint32_t *pkey, uint32_t *pval, uint64_t pair_cap,
                            const uint32_t *counters);
// tile_start[0..num_tiles] from the sorted pair keys (tile or superblock id =
// key & key_mask); *count pairs.
hipError_t launch_ranges(hipStream_t s, const uint32_t *skey, uint32_t key_mask, const uint32_t *count,
                         uint64_t pair_cap, uint32_t num_tiles, uint32_t *tile_start);

// ---- blend.hip: (e) front-to-back blend of every tile of every view ----
// mode BM_RGB / BM_FIRST / BM_SECOND write RGB8 (ViewParams::out); BM_DIFF
// writes the float image (ViewParams::fimg) and the per-pixel final transmittance / last
// contributing list position (pix_T / pix_last, total_tiles * 256 each).
// BM_FIRST sets done[t] per tile and keeps unfinished tiles' pixel state in
// pstate (total_tiles * 256 * 4 floats); BM_SECOND resumes from it. BM_FIRST
// also sets each unfinished tile's bit in openm (row ty of view v in the words
// from tile_base + ty * ceil(tiles_x / 32), cleared by the first emission) and
// counts the view's unfinished tiles in vopen[v].
// batch = records staged through LDS per step (64, 128 or 256). exact: the
// scene may give a splat an infinite colour (launch_prep's flag), which the
// reference's 4-lane groups turn into NaN beside the lanes that blend; such
// a scene runs a blend that reproduces them (RGB modes; BM_DIFF ignores it).
hipError_t launch_blend(hipStream_t s, int mode, const ViewParams *views, uint32_t nviews, uint32_t total_tiles,
                        uint32_t total_sbs, const uint32_t *tile_start, const uint32_t *sval, const uint4 *rec,
                        uint32_t *counters, float *pix_T, uint32_t *pix_last, uint32_t *done, float *pstate,
                        uint32_t *openm, uint32_t *vopen, int batch, bool exact);

// ---- diff.hip: adjoint of the blend and of the projection ----
// Per-pair gradients (NG2D arrays of gstride floats, indexed by emission
// slot: pbase[record] + the tile's row-major place in the record's tile
// rectangle) of every tile's list (record indices, sval) up to its last used
// position; each pair written is marked in the bit array fresh (bit e % 32 of
// word e / 32, which the caller clears first), and pairs that no pixel took
// are left untouched.
hipError_t launch_blend_bwd(hipStream_t s, const ViewParams *views, uint32_t nviews, uint32_t total_tiles,
                            uint32_t total_sbs, const uint32_t *tile_start, const uint32_t *sval, const uint32_t *pbase,
                            const uint4 *rec, const float *pix_T, const uint32_t *pix_last, float *gpair,
                            uint64_t gstride, uint32_t *fresh);
// g2d[record * NG2D + c] = sum over the record's pairs of gpair where the
// fresh bit is set (others count as zero), for every visible record of the
// views (pslot / prid as the emission wrote them); record-major, so each
// record's sums are one 36-byte write. Other records are untouched: the
// caller zero-fills g2d.
hipError_t launch_pair_reduce(hipStream_t s, const float *gpair, uint64_t gstride, const uint32_t *fresh,
                              const uint32_t *pslot, const uint32_t *prid, const uint32_t *vis_count, uint32_t n,
                              uint32_t nviews, float *g2d);
// Per-gaussian parameter gradients (sc.storage == STORE_PARAMS), summed over views.
hipError_t launch_project_bwd(hipStream_t s, const DevScene &sc, const ViewParams *views, uint32_t nviews,
                              const float *g2d, float *g_means, float *g_log_scales,
                              float *g_quats, float *g_logits, float *g_colors, float *g_sh_rest);
// p[0..n) = value for float arrays.
hipError_t launch_fill_f32(hipStream_t s, float *p, uint64_t n, float value);

// ---- term.hip: terminal output ----
// dst (dw x dh RGB8) = src (w wide) box-filtered by factor >= 1 (term::downsample).
hipError_t launch_downsample(hipStream_t s, const uint8_t *src, uint32_t w, uint32_t factor, uint32_t dw,
                             uint32_t dh, uint8_t *dst);
// Half-block cells (cols x rows) of src (w wide) downsampled by factor >= 1:
// 6 bytes per cell (TERM_TRUECOLOR) or 2 (TERM_ANSI256).
hipError_t launch_cells(hipStream_t s, const uint8_t *src, uint32_t w, uint32_t factor, uint32_t cols, uint32_t rows,
                        int32_t mode, uint8_t *cells);

}  // namespace sp
