// Terminal output on the device: the viewer's supersampling box filter and
// its half-block cells, so that only the cells cross PCIe.
//
//   k_downsample  term::downsample (src/term.rs:75-101): each output pixel is
//                 the truncated mean of a factor x factor block, per channel;
//                 trailing rows and columns that do not fill a block are
//                 dropped.
//   k_cells       HalfBlockImage (src/term.rs:103-128) over the downsampled
//                 frame: cell (cx, cy) shows display pixel (cx, 2cy) as its
//                 foreground and (cx, 2cy+1) as its background, as 24-bit
//                 colour (6 bytes per cell) or as xterm-256 indices
//                 (rgb_to_ansi256, src/term.rs:48-72; 2 bytes per cell).
//
// Integer arithmetic only, so both are exact. One thread per output pixel or
// cell; consecutive threads take consecutive columns, so the block rows they
// read are contiguous runs of 3 * factor bytes. A 1080p frame is 6 MB: these
// are small, bandwidth-bound kernels.
#include "sp_device.h"
#include "sp_launch.h"

namespace sp {
namespace {

constexpr int TERM_THREADS = 256;

// Sum of the factor x factor block at display pixel (x, y) of a w-wide RGB8 frame.
__device__ __forceinline__ uint3 block_sum(const uint8_t *__restrict__ src, uint32_t w, uint32_t x, uint32_t y,
                                           uint32_t factor) {
    uint3 acc = make_uint3(0u, 0u, 0u);
    for (uint32_t sy = 0; sy < factor; ++sy) {
        const uint8_t *row = src + 3 * ((uint64_t)(y * factor + sy) * w + (uint64_t)x * factor);
        for (uint32_t sx = 0; sx < factor; ++sx) {
            acc.x += row[3 * sx];
            acc.y += row[3 * sx + 1];
            acc.z += row[3 * sx + 2];
        }
    }
    return acc;
}

__device__ __forceinline__ uint3 display_pixel(const uint8_t *__restrict__ src, uint32_t w, uint32_t x, uint32_t y,
                                               uint32_t factor) {
    if (factor == 1) {
        const uint8_t *p = src + 3 * ((uint64_t)y * w + x);
        return make_uint3(p[0], p[1], p[2]);
    }
    const uint3 a = block_sum(src, w, x, y, factor);
    const uint32_t n = factor * factor;
    return make_uint3(a.x / n, a.y / n, a.z / n);
}

// Index of the 6-level cube coordinate nearest v (rgb_to_ansi256's cube_idx).
__device__ __forceinline__ int cube_idx(int v) { return v < 48 ? 0 : v < 115 ? 1 : min((v - 35) / 40, 5); }

__device__ __forceinline__ uint32_t ansi256(uint3 c) {
    const int LEVELS[6] = {0, 95, 135, 175, 215, 255};
    const int r = (int)c.x, g = (int)c.y, b = (int)c.z;
    const int ri = cube_idx(r), gi = cube_idx(g), bi = cube_idx(b);
    const int avg = (r + g + b) / 3;
    const int gray_idx = min(max((avg - 3) / 10, 0), 23);  // division truncates toward zero
    const int gray = 8 + gray_idx * 10;
    const auto sq = [](int d) { return d * d; };
    const int dg = sq(gray - r) + sq(gray - g) + sq(gray - b);
    const int dc = sq(LEVELS[ri] - r) + sq(LEVELS[gi] - g) + sq(LEVELS[bi] - b);
    return dg < dc ? (uint32_t)(232 + gray_idx) : (uint32_t)(16 + 36 * ri + 6 * gi + bi);
}

__global__ __launch_bounds__(TERM_THREADS) void k_downsample(const uint8_t *__restrict__ src, uint32_t w,
                                                             uint32_t factor, uint32_t dw, uint32_t dh,
                                                             uint8_t *__restrict__ dst) {
    const uint64_t i = (uint64_t)blockIdx.x * TERM_THREADS + threadIdx.x;
    if (i >= (uint64_t)dw * dh) return;
    const uint32_t x = (uint32_t)(i % dw), y = (uint32_t)(i / dw);
    const uint3 p = display_pixel(src, w, x, y, factor);
    uint8_t *o = dst + 3 * i;
    o[0] = (uint8_t)p.x;
    o[1] = (uint8_t)p.y;
    o[2] = (uint8_t)p.z;
}

__global__ __launch_bounds__(TERM_THREADS) void k_cells(const uint8_t *__restrict__ src, uint32_t w, uint32_t factor,
                                                        uint32_t cols, uint32_t rows, int32_t mode,
                                                        uint8_t *__restrict__ cells) {
    const uint64_t i = (uint64_t)blockIdx.x * TERM_THREADS + threadIdx.x;
    if (i >= (uint64_t)cols * rows) return;
    const uint32_t cx = (uint32_t)(i % cols), cy = (uint32_t)(i / cols);
    const uint3 top = display_pixel(src, w, cx, 2 * cy, factor);
    const uint3 bot = display_pixel(src, w, cx, 2 * cy + 1, factor);
    if (mode == TERM_ANSI256) {
        cells[2 * i] = (uint8_t)ansi256(top);
        cells[2 * i + 1] = (uint8_t)ansi256(bot);
    } else {
        uint8_t *o = cells + 6 * i;
        o[0] = (uint8_t)top.x;
        o[1] = (uint8_t)top.y;
        o[2] = (uint8_t)top.z;
        o[3] = (uint8_t)bot.x;
        o[4] = (uint8_t)bot.y;
        o[5] = (uint8_t)bot.z;
    }
}

}  // namespace

hipError_t launch_downsample(hipStream_t s, const uint8_t *src, uint32_t w, uint32_t factor, uint32_t dw,
                             uint32_t dh, uint8_t *dst) {
    const uint64_t n = (uint64_t)dw * dh;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_downsample, dim3((uint32_t)((n + TERM_THREADS - 1) / TERM_THREADS)), dim3(TERM_THREADS), 0, s,
                       src, w, factor, dw, dh, dst);
    return hipGetLastError();
}

hipError_t launch_cells(hipStream_t s, const uint8_t *src, uint32_t w, uint32_t factor, uint32_t cols, uint32_t rows,
                        int32_t mode, uint8_t *cells) {
    const uint64_t n = (uint64_t)cols * rows;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_cells, dim3((uint32_t)((n + TERM_THREADS - 1) / TERM_THREADS)), dim3(TERM_THREADS), 0, s, src,
                       w, factor, cols, rows, mode, cells);
    return hipGetLastError();
}

}  // namespace sp
