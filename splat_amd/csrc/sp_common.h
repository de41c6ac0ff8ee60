// Shared host/device definitions of the HIP rasterizer.
#pragma once

#include <stdint.h>

#include "splat_hip.h"

namespace sp {

// Reference constants (src/render/mod.rs:23-33, src/render/cpu.rs:34).
constexpr float DILATION = 0.3f;
constexpr float ALPHA_MIN = 1.0f / 255.0f;
constexpr float ALPHA_MAX = 0.99f;
constexpr float T_MIN = 1e-4f;
constexpr uint32_t TILE = 16;
// Terminal cell layouts (SP_TERM_TRUECOLOR / SP_TERM_ANSI256 of the C ABI).
constexpr int32_t TERM_TRUECOLOR = 0;
constexpr int32_t TERM_ANSI256 = 1;
// Superblock: SB x SB tiles, the bin of the first emission phase (tiles.hip).
constexpr uint32_t SB = 4;
constexpr uint32_t BIN_SCAN_ROW = 2048;  // tiles per row of the binned tile-count scan
constexpr uint32_t BIN_SEGS = 4;         // workgroups per superblock list in the binned first phase
constexpr uint32_t SB_SUB_SHIFT = 24;    // superblock pair keys: superblock id below, sub-rectangle above

// Workgroup shapes.
constexpr int PROJECT_THREADS = 256;  // one splat per thread
constexpr int PROJECT_VCHUNK = 8;     // views projected per pass over a splat held in registers
constexpr int EMIT_THREADS = 256;     // one depth-ordered splat per thread
constexpr int BLEND_THREADS = 256;    // one pixel of a 16x16 tile per thread
constexpr int RADIX_THREADS = 256;    // one or two digits per thread
// Keys per thread of a radix tile: 16 (4096-key tiles) for large sorts, 4
// (1024-key tiles) for sorts whose segments hold at most RADIX_SMALL_MAX keys
// together, whose 4096-key tiles would leave most CUs idle. Their 512-bin
// passes take RADIX_ITEMS_WIDE (2048-key tiles): 1024 keys over 512 digits
// leave runs of about 2 keys, whose scattered writes barely coalesce.
constexpr int RADIX_ITEMS_LARGE = 16;
constexpr int RADIX_ITEMS_SMALL = 4;
constexpr int RADIX_ITEMS_WIDE = 8;
constexpr uint64_t RADIX_SMALL_MAX = 16ull << 20;
__host__ __device__ constexpr uint32_t radix_tile_keys(uint32_t nseg, uint64_t n_cap) {
    return (uint32_t)(RADIX_THREADS * ((uint64_t)nseg * n_cap <= RADIX_SMALL_MAX ? RADIX_ITEMS_SMALL : RADIX_ITEMS_LARGE));
}
// Radix digits are 8 bits (256 bins); a pass may carry a 9-bit digit (512
// bins) when that saves a pass (a 9-bit sort in one pass instead of 8 + 1),
// so tables are sized for 512.
constexpr int RADIX_BINS_MAX = 512;
constexpr uint32_t MAX_VIEWS = 4096;  // views per batch

// Device counter slots (uint32 each) in the per-handle counter block.
enum Counter : int {
    CNT_PAIRS = 1,     // pairs of the (first) emission phase, exact even when over capacity (0xFFFFFFFF: 2^32 or more)
    CNT_OVERFLOW = 2,  // 1 when an emission phase exceeded the pair capacity
    CNT_STAGED = 3,    // pair records the blends gathered (early termination skips the rest)
    CNT_PAIRS2 = 4,    // pairs of the second emission phase
    CNT_SBPAIRS = 5,   // (superblock, splat) pairs of the binned first phase
    CNT_DEPTH_RANGE = 6,  // set when a depth key reached 2^DEPTH_KEY_BITS (never cleared on the device)
    CNT_SLOTS = 16
};

// Blend modes (blend.hip).
constexpr int BM_RGB = 0;
constexpr int BM_DIFF = 1;
constexpr int BM_FIRST = 2;
constexpr int BM_SECOND = 3;

// Emission phases (front-to-back split of each view's depth order).
constexpr int EMIT_FIRST = 0;   // the nearest K1 splats of each view, every tile
constexpr int EMIT_SECOND = 1;  // the remaining splats, unfinished tiles only
// K1 = min(vis, max(kmin, ceil(vis / 2^shift))); shift 0 emits everything at once.
struct EmitSplit {
    uint32_t kmin;
    uint32_t shift;
};

// ProjectCtx (src/render/mod.rs:155-197) plus frame/tile geometry.
struct ViewCtx {
    float rot[3][3];  // rows: right, down, forward (world -> camera)
    float eye[3];
    float focal, cx, cy, width, height, near_z, lim_x, lim_y;
    uint32_t w, h, tiles_x, tiles_y;
    float bg[3];
    // Depth-sort key base: when nonzero, a visible splat's key is its depth
    // bits less key_base, below 2^DEPTH_KEY_BITS (the host bounds the depth
    // range of the batch's views from the scene's bounding sphere); 0 keeps
    // the full 32-bit depth bits.
    uint32_t key_base;
};

// One view of a batch, in device memory: its context, the first global tile id
// of its tiles (views' tiles are numbered consecutively), the first global
// superblock id of its ceil(tiles_x / SB) x ceil(tiles_y / SB) superblocks
// (numbered the same way) and its RGB8 output.
// The differentiable path writes a float image (fimg, w*h*3, C + T*bg before
// quantisation) instead and reads the loss gradient of that image (gimg).
struct ViewParams {
    ViewCtx ctx;
    uint32_t tile_base;
    uint32_t sb_base;
    uint8_t *out;
    float *fimg;
    const float *gimg;
};

// Device-resident scene as columns of splat fields, stored in blocks of 64
// splats: column c of splat i is at scene_at(ncols, c, i), so a wave's read
// of one field is 256 contiguous bytes (as from a plain column), and the
// fields of its 64 splats fill one contiguous block of ncols * 256 bytes
// (separate columns would be ncols far-apart streams: a probe read 59 of them
// at 5.2 TB/s, 59 blocked fields at 6.4 TB/s, a plain copy at 6.6 TB/s).
// stride = n rounded up to the block size; a layout of ncols columns takes
// ncols * stride entries. Columns: 0-2 position x,y,z; 3-8 cov3d
// xx,xy,xz,yy,yz,zz; 9 opacity; 10 k = sqrt(2 ln(o/ALPHA_MIN)); 11-13 base
// colour; 14.. SH rest coefficient c channel ch at 14 + 3c + ch.
//
// The colour is also stored as one contiguous row per splat with the
// position it is seen from (colour_row entries: position x, y, z as f32 bits
// (six entries of a 16-bit row), base colour, then SH rest coefficient c
// channel ch, then padding). A single view's projection reads no colour (the columns
// above up to k only), and the emission evaluates the colour of just the
// splats whose pairs it emits for a blend, each row one or two 128-byte lines
// of a gather. A batch's projection, whose time goes to writing its views'
// records, evaluates the colour from the columns as it projects.
//
// Reduced-precision storage (FP16 / BF16) keeps the f32 columns 0-2 position,
// 3 opacity, 4 k (a layout of L_COLS columns), and stores the covariance
// factors and colour as 16-bit columns in soa16: 0-2 scale, 3-6 quaternion
// x,y,z,w, 7-9 base colour, 10.. SH rest (coefficient c channel ch at
// 10 + 3c + ch), blocked the same way, and its colour rows in 16 bits. The
// covariance is rebuilt in f32 on the device.
//
// The differentiable path (storage STORE_PARAMS) reads raw, unactivated
// parameters instead, row-major per gaussian as the caller's tensors hold
// them: means [n][3], log-scales [n][3], quaternions [n][4] (x,y,z,w,
// normalised on the fly), opacity logits [n], base colours [n][3], SH rest
// [n][rest][3]. Activation (exp, sigmoid, normalisation, covariance, k) is
// computed per use, and the projection evaluates the colour itself.
constexpr uint32_t SCENE_BLOCK = 64;
__host__ __device__ constexpr uint64_t scene_at(uint32_t ncols, uint32_t c, uint32_t i) {
    return ((uint64_t)(i / SCENE_BLOCK) * ncols + c) * SCENE_BLOCK + (i % SCENE_BLOCK);
}
struct DevParams {
    const float *means, *log_scales, *quats, *logits, *colors, *sh_rest;
};
// Entries of a splat's colour row (position, base colour, SH rest), rounded
// up to 16 bytes; the base colour starts at entry colour_base.
__host__ __device__ constexpr uint32_t colour_base(bool half) { return half ? 6u : 3u; }
__host__ __device__ constexpr uint32_t colour_row(uint32_t rest, bool half) {
    return half ? (6u + 3u + 3u * rest + 7u) & ~7u : (3u + 3u + 3u * rest + 3u) & ~3u;
}
struct DevScene {
    const float *soa;
    const uint16_t *soa16;
    const float *crow;      // f32 storage: colour rows
    const uint16_t *crow16; // 16-bit storage: colour rows
    uint64_t stride;
    uint32_t n;
    uint32_t deg;
    uint32_t rest;
    uint32_t storage;  // SP_STORE_* or STORE_PARAMS
    DevParams prm;
};
constexpr int STORE_PARAMS = 3;
// Per-pair gradient of the loss with respect to a projected record, reduced
// over the pixels of the pair's tile: mean x, y; conic a, b, c; opacity;
// red, green, blue.
constexpr int NG2D = 9;
// Gradient floats per gaussian: mean 3, log-scale 3, quaternion 4, opacity
// logit 1, base colour 3, SH rest 3 * rest.
constexpr int NG3D_BASE = 14;
constexpr int A_POS = 0, A_COV = 3, A_OPAC = 9, A_K = 10, A_COL = 11, A_SH = 14;
constexpr int L_POS = 0, L_OPAC = 3, L_K = 4, L_COLS = 5;                // f32 arrays, 16-bit storage
constexpr int H_SCALE = 0, H_QUAT = 3, H_COL = 7, H_SH = 10;             // 16-bit arrays
// Batches of at most this many views leave the colour to the emission.
constexpr uint32_t DEFER_COLOUR_VIEWS = 1;
// det <= NEAR_SINGULAR * a * c marks a projected covariance whose f32
// determinant lost more than ~10 of its 24 bits to cancellation; the 16-bit
// storage paths recompute such determinants in FP64.
constexpr float NEAR_SINGULAR = 1e-3f;

// Projected splat record (48 bytes); view v's record of scene index i is at
// v * n + i. Layout:
//   r0 = x, y, -0.5*conic_a, -conic_b
//   r1 = -0.5*conic_c, opacity, red, green
//   r2 = blue, bits(x0 | y0 << 16), bits(x1 | y1 << 16), falloff threshold -k^2/2 - 0.02
// The pre-scaled conic terms are the exact factors raster_tile multiplies by
// (src/render/cpu.rs:662, :671-672), so precomputing them changes no bits.

// Depth-key marker of a culled (view, splat) in the uncompacted key array;
// visible depths are positive finite floats, whose bits never reach it.
constexpr uint32_t FULL_CULLED = 0xFFFFFFFFu;
// Width of the depth keys relative to ViewCtx::key_base: three 9-bit radix
// passes instead of four 8-bit ones. Depth bits from just above the near
// plane up to 2^27 higher span 16 binary orders of magnitude of depth.
constexpr int DEPTH_KEY_BITS = 27;

}  // namespace sp
