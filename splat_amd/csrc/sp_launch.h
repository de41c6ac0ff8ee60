// Host-side launchers of the rasterizer kernels (one translation unit each).
#pragma once

#include <hip/hip_runtime.h>

#include "sp_common.h"

namespace sp {

// (a) Projection + SH colour + visible-splat compaction. Writes records for
// visible splats, depth keys/values (scene index) in scene-index order, and
// counters[CNT_VISIBLE]. dbg / dbg_bbox / dbg_valid (optional, n entries) get
// the full per-splat projection for inspection.
hipError_t launch_project(hipStream_t s, const DevScene &sc, const ViewCtx &v, uint4 *rec, uint32_t *dkey,
                          uint32_t *dval, Lookback lb, float *dbg, uint16_t *dbg_bbox, uint8_t *dbg_valid);

// (c) Stable LSD radix sort of u32 keys with u32 values over bits [0, bits),
// 8-bit digits, onesweep passes with decoupled lookback. The element count is
// read from *n_dev; n_cap bounds it and sizes the grid. Ping-pongs between
// keys[0]/keys[1]; returns in *out_index which buffer holds the result.
// `hist` holds ceil(bits/8)*256 zeroed u32 and is zeroed again on return.
hipError_t launch_radix_sort(hipStream_t s, uint32_t *keys[2], uint32_t *vals[2], const uint32_t *n_dev,
                             uint64_t n_cap, int bits, Lookback lb, uint32_t *hist, int *out_index);

// (b) Tile-pair emission in depth order: for the depth-sorted splats (values
// dval, count counters[CNT_VISIBLE]) write one (tile id, scene index) pair per
// overlapped 16x16 tile, compacted by a single-pass scan. Sets
// counters[CNT_PAIRS] (exact) and counters[CNT_OVERFLOW] when above cap.
hipError_t launch_emit(hipStream_t s, const uint32_t *dval, uint64_t n_cap, const uint4 *rec, const ViewCtx &v,
                       uint32_t *pkey, uint32_t *pval, uint64_t pair_cap, Lookback lb);

// (d) Tile ranges: tile_start[0..num_tiles] from the tile-sorted pair keys.
hipError_t launch_ranges(hipStream_t s, const uint32_t *skey, const uint32_t *counters, uint64_t pair_cap,
                         uint32_t num_tiles, uint32_t *tile_start);

// (e) Front-to-back blend of every 16x16 tile into RGB8.
hipError_t launch_blend(hipStream_t s, const ViewCtx &v, const uint32_t *tile_start, const uint32_t *sval,
                        const uint4 *rec, uint8_t *rgb, uint32_t *counters);

}  // namespace sp
