// Shared host/device definitions of the HIP rasterizer.
#pragma once

#include <stdint.h>

namespace sp {

// Reference constants (src/render/mod.rs:23-33, src/render/cpu.rs:34).
constexpr float DILATION = 0.3f;
constexpr float ALPHA_MIN = 1.0f / 255.0f;
constexpr float ALPHA_MAX = 0.99f;
constexpr float T_MIN = 1e-4f;
constexpr uint32_t TILE = 16;

// Workgroup shapes.
constexpr int PROJECT_THREADS = 256;  // one splat per thread
constexpr int EMIT_THREADS = 256;     // one depth-ordered splat per thread
constexpr int BLEND_THREADS = 256;    // one pixel of a 16x16 tile per thread
constexpr int RADIX_THREADS = 256;    // one digit per thread in the lookback
constexpr int RADIX_ITEMS = 16;       // keys per thread per onesweep tile
constexpr int RADIX_TILE = RADIX_THREADS * RADIX_ITEMS;  // 4096 keys
constexpr int RADIX_BINS = 256;       // 8-bit digits

// Device counter slots (uint32 each) in the per-handle counter block.
enum Counter : int {
    CNT_VISIBLE = 0,   // splats that survived culling (V)
    CNT_PAIRS = 1,     // (splat, tile) pairs (P), exact even when over capacity
    CNT_OVERFLOW = 2,  // 1 when P exceeded the pair capacity
    CNT_STAGED = 3,    // pair records the blend gathered (early termination skips the rest)
    CNT_TICKET = 4,    // dynamic block-id ticket of the running lookback kernel
    CNT_DONE = 5,      // finished blocks of the running lookback kernel
    CNT_EPOCH = 6,     // launch epoch tag for lookback status words
    CNT_SLOTS = 16
};

// ProjectCtx (src/render/mod.rs:155-197) plus frame/tile geometry, passed by
// value to every per-view kernel.
struct ViewCtx {
    float rot[3][3];  // rows: right, down, forward (world -> camera)
    float eye[3];
    float focal, cx, cy, width, height, near_z, lim_x, lim_y;
    uint32_t w, h, tiles_x, tiles_y;
    float bg[3];
    uint32_t pad;
};

// Device-resident scene in structure-of-arrays form: array `a` of n floats
// starts at soa + a * stride. Arrays: 0-2 position x,y,z; 3-8 cov3d
// xx,xy,xz,yy,yz,zz; 9 opacity; 10 k = sqrt(2 ln(o/ALPHA_MIN)); 11-13 base
// colour; 14.. SH rest coefficient c channel ch at 14 + 3c + ch.
struct DevScene {
    const float *soa;
    uint64_t stride;
    uint32_t n;
    uint32_t deg;
    uint32_t rest;
    uint32_t pad;
};
constexpr int A_POS = 0, A_COV = 3, A_OPAC = 9, A_K = 10, A_COL = 11, A_SH = 14;

// Projected splat record (48 bytes, indexed by scene index):
//   r0 = x, y, -0.5*conic_a, -conic_b
//   r1 = -0.5*conic_c, opacity, red, green
//   r2 = blue, bits(x0 | y0 << 16), bits(x1 | y1 << 16), depth
// The pre-scaled conic terms are the exact factors raster_tile multiplies by
// (src/render/cpu.rs:662, :671-672), so precomputing them changes no bits.
struct Rec {
    float r[12];
};

// Lookback-scan state shared by the single-pass scans and the onesweep sort.
struct Lookback {
    unsigned long long *status;  // epoch-tagged status words
    uint32_t *counters;          // CNT_* slots (ticket, done, epoch)
};

}  // namespace sp
