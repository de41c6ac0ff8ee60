// (e) Per-tile front-to-back alpha blending.
//
// One 256-thread workgroup per 16x16 tile of any view of the batch (the views'
// tiles are numbered consecutively), one pixel per thread; wave k of the
// group owns tile rows 4k..4k+3. The tile's depth-ordered splat list is
// staged through LDS 256 records at a time (each thread gathers one 48-byte
// record), then every wave walks the batch: a record whose pixel box misses
// the wave's four rows is skipped by the whole wave, otherwise each live
// pixel inside the box evaluates the gaussian. A pixel stops at T < T_MIN; the
// tile stops when no pixel is live.
//
// Per-pixel arithmetic is raster_tile's (src/render/cpu.rs:609-719) in the
// same operation order, with pixel offsets built from the 16-aligned tile
// origin exactly as its 4-lane groups build them
// (dx = lane + (4*group + ((ox + 0.5) - x))), and exp_neg's polynomial, so
// the output bytes equal the CPU backend's.
#include "sp_device.h"
#include "sp_launch.h"

namespace sp {
namespace {

constexpr int XCDS = 8;  // MI355X: workgroups are dealt round-robin to 8 XCDs

__global__ __launch_bounds__(BLEND_THREADS) void k_blend(const ViewParams *__restrict__ views, uint32_t nviews,
                                                         uint32_t total_tiles, const uint32_t *__restrict__ tile_start,
                                                         const uint32_t *__restrict__ sval,
                                                         const uint4 *__restrict__ rec, uint32_t per_xcd,
                                                         uint32_t *__restrict__ counters) {
    __shared__ uint4 s_rec[BLEND_THREADS][3];
    // Consecutive tiles (rows of an image) on the same XCD share its L2.
    const uint32_t gt = (blockIdx.x % XCDS) * per_xcd + blockIdx.x / XCDS;
    if (gt >= total_tiles) return;
    // The view owning global tile gt: last view whose tile_base <= gt (bases ascend).
    uint32_t lo = 0, hi = nviews;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (views[mid].tile_base <= gt)
            lo = mid;
        else
            hi = mid;
    }
    const ViewCtx &v = views[lo].ctx;
    uint8_t *__restrict__ rgb = views[lo].out;
    const uint32_t t = gt - views[lo].tile_base;
    const uint32_t tx = t % v.tiles_x, ty = t / v.tiles_x;
    const uint32_t ox = tx * TILE, oy = ty * TILE;
    const uint32_t lx = threadIdx.x & 15u, ly = threadIdx.x >> 4;
    const uint32_t px = ox + lx, py = oy + ly;
    const uint32_t wy0 = oy + (threadIdx.x >> 6) * 4u, wy1 = wy0 + 4u;
    bool live = px < v.w && py < v.h;
    const float fx = (float)ox;
    const float gx = (float)(lx & ~3u);  // 4 * group
    const float lane = (float)(lx & 3u);
    const float pyc = ((float)oy + (float)ly) + 0.5f;
    float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    const uint32_t start = tile_start[gt], end = tile_start[gt + 1];
    uint32_t b = start;
    for (; b < end; b += BLEND_THREADS) {
        const uint32_t idx = b + threadIdx.x;
        if (idx < end) {
            const uint4 *r = rec + 3 * (uint64_t)sval[idx];
            s_rec[threadIdx.x][0] = r[0];
            s_rec[threadIdx.x][1] = r[1];
            s_rec[threadIdx.x][2] = r[2];
        }
        __syncthreads();
        const uint32_t nb = min(end - b, (uint32_t)BLEND_THREADS);
        if (__any(live)) {
            for (uint32_t k = 0; k < nb; ++k) {
                const uint4 r2 = s_rec[k][2];
                const uint32_t by0 = r2.y >> 16, by1 = r2.z >> 16;
                if (by1 <= wy0 || by0 >= wy1) continue;  // wave-uniform
                const uint32_t bx0 = r2.y & 0xFFFFu, bx1 = r2.z & 0xFFFFu;
                if (!(live && px >= bx0 && px < bx1 && py >= by0 && py < by1)) continue;
                const uint4 r0 = s_rec[k][0];
                const uint4 r1 = s_rec[k][1];
                const float sx = __uint_as_float(r0.x), sy = __uint_as_float(r0.y);
                const float na = __uint_as_float(r0.z), nb2 = __uint_as_float(r0.w);
                const float nc = __uint_as_float(r1.x), op = __uint_as_float(r1.y);
                const float cx0 = (fx + 0.5f) - sx;
                const float dx = lane + (gx + cx0);
                const float dy = pyc - sy;
                const float qx = (dx * dx) * na;
                const float qy = (nc * dy) * dy;
                const float bdy = nb2 * dy;
                const float pw = (qx + qy) + bdy * dx;
                const float alpha = fminf(exp_neg(pw) * op, ALPHA_MAX);
                if (pw <= 0.0f && alpha >= ALPHA_MIN) {
                    const float wgt = alpha * T;
                    cr = cr + __uint_as_float(r1.z) * wgt;
                    cg = cg + __uint_as_float(r1.w) * wgt;
                    cb = cb + __uint_as_float(r2.x) * wgt;
                    T = T * (1.0f - alpha);
                    if (T < T_MIN) live = false;
                }
            }
        }
        if (__syncthreads_count(live) == 0) {
            b += BLEND_THREADS;
            break;
        }
    }
    if (threadIdx.x == 0 && end > start) atomicAdd(&counters[CNT_STAGED], min(b, end) - start);
    if (px < v.w && py < v.h) {
        uint8_t *o = rgb + 3 * ((uint64_t)py * v.w + px);
        o[0] = (uint8_t)to_u8(cr + T * v.bg[0]);
        o[1] = (uint8_t)to_u8(cg + T * v.bg[1]);
        o[2] = (uint8_t)to_u8(cb + T * v.bg[2]);
    }
}

}  // namespace

hipError_t launch_blend(hipStream_t s, const ViewParams *views, uint32_t nviews, uint32_t total_tiles,
                        const uint32_t *tile_start, const uint32_t *sval, const uint4 *rec, uint32_t *counters) {
    if (total_tiles == 0) return hipSuccess;
    const uint32_t per_xcd = (total_tiles + XCDS - 1) / XCDS;
    hipLaunchKernelGGL(k_blend, dim3(per_xcd * XCDS), dim3(BLEND_THREADS), 0, s, views, nviews, total_tiles,
                       tile_start, sval, rec, per_xcd, counters);
    return hipGetLastError();
}

}  // namespace sp
