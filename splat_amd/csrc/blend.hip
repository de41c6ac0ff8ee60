// (e) Per-tile front-to-back alpha blending.
//
// One 256-thread workgroup per 16x16 tile of any view of the batch (the views'
// tiles are numbered consecutively), one pixel per thread; wave k of the
// group owns one 8x8 quadrant. The tile's depth-ordered splat list is staged
// through LDS BATCH records at a time (each thread gathers one 48-byte
// record and bounds its falloff over each quadrant), then every wave walks
// the batch: a record whose pixel box misses the wave's quadrant, or whose
// falloff cannot reach the splat's alpha threshold at any in-box pixel of the
// quadrant, is never visited (the wave ballots the reach bits of 64 staged
// records at a time and walks only the set bits, in list order); one whose
// falloff stays below the threshold at every live in-box pixel is skipped
// after the falloff test (those pixels are masked by the reference too).
// A pixel stops at T < T_MIN; a wave stops when its quadrant has no live
// pixel, the tile when no pixel is live.
//
// Per-pixel arithmetic is raster_tile's (src/render/cpu.rs:609-719) in the
// same operation order, with pixel offsets built from the 16-aligned tile
// origin exactly as its 4-lane groups build them
// (dx = lane + (4*group + ((ox + 0.5) - x))), and exp_neg's polynomial, so
// the output bytes equal the CPU backend's.
//
// Modes:
//   BM_RGB     one list per tile, RGB8 output.
//   BM_DIFF    the differentiable forward (lists of record indices, like
//              the others): the output is the float image
//              C + T*bg, and each pixel's final transmittance and the
//              tile-list position after its last contribution are kept for
//              the adjoint (pix_T / pix_last, 256 per tile in thread order).
//   BM_FIRST   the first list of the front-to-back split: a tile whose pixels
//              all saturate writes its RGB8 output and is marked done; any
//              other tile keeps (T, C) per pixel in pstate for BM_SECOND.
//   BM_SECOND  the continuation list of the tiles not done: resumes from
//              pstate and writes the RGB8 output.
#include "sp_device.h"
#include "sp_launch.h"

namespace sp {
namespace {

template <int BATCH, int MODE>
__global__ __launch_bounds__(BLEND_THREADS) void k_blend(const ViewParams *__restrict__ views, uint32_t nviews,
                                                         uint32_t total_tiles, const uint32_t *__restrict__ tile_start,
                                                         const uint32_t *__restrict__ sval,
                                                         const uint4 *__restrict__ rec, uint32_t total_sbs,
                                                         uint32_t *__restrict__ counters, float *__restrict__ pix_T,
                                                         uint32_t *__restrict__ pix_last, uint32_t *__restrict__ done,
                                                         float4 *__restrict__ pstate) {
    __shared__ uint4 s_rec[BATCH][3];
    __shared__ uint32_t s_qm[BATCH];  // reach bits (0-3) and full-cover bits (4-7) of each staged record
    uint32_t gt, lo;
    if (!blend_tile_of(views, nviews, total_sbs, blockIdx.x, &gt, &lo)) return;
    if (MODE == BM_SECOND && done[gt]) return;  // finished by the first list
    const ViewCtx &v = views[lo].ctx;
    uint8_t *__restrict__ rgb = views[lo].out;
    const uint32_t t = gt - views[lo].tile_base;
    const uint32_t tx = t % v.tiles_x, ty = t / v.tiles_x;
    const uint32_t ox = tx * TILE, oy = ty * TILE;
    // Wave w owns the 8x8 quadrant (w & 1, w >> 1) of the tile; lane = row * 8 + column.
    const uint32_t w = threadIdx.x >> 6, lane_i = lane_id();
    const uint32_t qx0 = ox + 8u * (w & 1u), qy0 = oy + 8u * (w >> 1);
    const uint32_t px = qx0 + (lane_i & 7u), py = qy0 + (lane_i >> 3);
    const uint32_t lx = px - ox, ly = py - oy;
    const bool inside = px < v.w && py < v.h;
    const float fx = (float)ox;
    const float gx = (float)(lx & ~3u);  // 4 * group of the reference's 4-lane groups
    const float lane = (float)(lx & 3u);
    const float pyc = ((float)oy + (float)ly) + 0.5f;
    const uint64_t pix = (uint64_t)gt * BLEND_THREADS + threadIdx.x;
    float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    if (MODE == BM_SECOND) {
        const float4 st = pstate[pix];
        T = st.x;
        cr = st.y;
        cg = st.z;
        cb = st.w;
    }
    // A pixel is live while inside the frame with T >= T_MIN (T only falls and
    // is never NaN, so this is the reference's stop at T < T_MIN); deriving it
    // from T keeps it a plain compare rather than a loop-carried mask.
    const auto live = [&]() { return inside && T >= T_MIN; };
    const uint64_t inside_mask = lane_mask(inside);
    const auto live_mask = [&]() { return lane_mask(T >= T_MIN) & inside_mask; };
    uint32_t last = 0;
    bool finished = false;
    const uint32_t start = tile_start[gt], end = tile_start[gt + 1];
    uint32_t b = start;
    for (; b < end; b += BATCH) {
        const uint32_t idx = b + threadIdx.x;
        if (threadIdx.x < BATCH && idx < end) {
            uint32_t id = sval[idx];
            const uint4 *r = rec + 3 * (uint64_t)id;
            uint4 r0 = r[0];
            const uint4 r1 = r[1], r2 = r[2];
            s_qm[threadIdx.x] = quadrant_mask(r0, r1, r2, ox, oy) | (cover_mask(r2, ox, oy) << 4);
            // The staged x is the tile's ((ox + 0.5) - x), the reference's
            // per-record term of every pixel offset, computed once here.
            r0.x = __float_as_uint((fx + 0.5f) - __uint_as_float(r0.x));
            s_rec[threadIdx.x][0] = r0;
            s_rec[threadIdx.x][1] = r1;
            s_rec[threadIdx.x][2] = r2;
        }
        __syncthreads();
        const uint32_t nb = __builtin_amdgcn_readfirstlane(min(end - b, (uint32_t)BATCH));
        // Records of the batch that can reach this wave's quadrant, 64 at a
        // time: one LDS read per lane and a ballot, then the set bits in order.
        bool open = live_mask() != 0ull;
        for (uint32_t g0 = 0; g0 < nb && open; g0 += 64) {
            const uint32_t qm = g0 + lane_i < nb ? s_qm[g0 + lane_i] : 0u;
            unsigned long long reach = lane_mask((qm >> w) & 1u);
            // Records whose box holds the whole quadrant: no box test.
            const unsigned long long cover = lane_mask((qm >> (4 + w)) & 1u);
            while (reach) {
                const uint64_t lm = live_mask();
                if (lm == 0ull) {  // every pixel of the quadrant saturated
                    open = false;
                    break;
                }
                const uint32_t kb = (uint32_t)__builtin_ctzll(reach);
                const uint32_t k = g0 + kb;
                reach &= reach - 1ull;
                const uint4 r0 = s_rec[k][0];
                const uint4 r1 = s_rec[k][1];
                const uint4 r2 = s_rec[k][2];
                // All 48 bytes are read at once from one address register (the
                // compiler would otherwise sink each field's read to its first
                // use and rebuild the address for each, one VALU move apiece).
                asm volatile("" ::"v"(r0.x), "v"(r0.y), "v"(r0.z), "v"(r0.w), "v"(r1.x), "v"(r1.y), "v"(r1.z),
                             "v"(r1.w), "v"(r2.x), "v"(r2.y), "v"(r2.z), "v"(r2.w));
                // Live in-box lanes as a wave mask (the per-lane conditions below
                // are all wave masks, combined on the scalar unit).
                uint64_t act_mask = lm;
                if (!((cover >> kb) & 1ull)) {  // wave-uniform
                    const uint32_t bx0 = r2.y & 0xFFFFu, by0 = r2.y >> 16, bx1 = r2.z & 0xFFFFu, by1 = r2.z >> 16;
                    act_mask &= lane_mask(px >= bx0) & lane_mask(px < bx1) & lane_mask(py >= by0) & lane_mask(py < by1);
                }
                const float cx0 = __uint_as_float(r0.x), sy = __uint_as_float(r0.y);
                const float na = __uint_as_float(r0.z), nb2 = __uint_as_float(r0.w);
                const float nc = __uint_as_float(r1.x);
                const float dx = lane + (gx + cx0);
                const float dy = pyc - sy;
                const float qx = (dx * dx) * na;
                const float qy = (nc * dy) * dy;
                const float bdy = nb2 * dy;
                const float pw = (qx + qy) + bdy * dx;
                // Lanes below the splat's threshold have alpha < ALPHA_MIN (the
                // reference masks them); a wave with none left skips the rest.
                if ((act_mask & lane_mask(pw >= __uint_as_float(r2.w))) == 0ull) continue;
                const float op = __uint_as_float(r1.y);
                const float alpha = fminf(exp_neg_times(pw, op), ALPHA_MAX);
                const uint64_t upd = act_mask & lane_mask(pw <= 0.0f) & lane_mask(alpha >= ALPHA_MIN);
                if (__builtin_amdgcn_inverse_ballot_w64(upd)) {
                    const float wgt = alpha * T;
                    cr = cr + __uint_as_float(r1.z) * wgt;
                    cg = cg + __uint_as_float(r1.w) * wgt;
                    cb = cb + __uint_as_float(r2.x) * wgt;
                    T = T * (1.0f - alpha);
                    if (MODE == BM_DIFF) last = (b - start) + k + 1;
                }
            }
        }
        if (__syncthreads_count(live()) == 0) {
            b += BATCH;
            finished = true;
            break;
        }
    }
    if (threadIdx.x == 0 && end > start) atomicAdd(&counters[CNT_STAGED], min(b, end) - start);
    if (MODE == BM_FIRST) {
        // `finished` is block-uniform, so the count below is reached by all threads or none.
        if (!finished) finished = __syncthreads_count(live()) == 0;
        if (!finished) {
            pstate[pix] = make_float4(T, cr, cg, cb);
            if (threadIdx.x == 0) done[gt] = 0u;
            return;
        }
        if (threadIdx.x == 0) done[gt] = 1u;
    }
    if (MODE == BM_DIFF) {
        pix_T[pix] = T;
        pix_last[pix] = last;
        if (inside) {
            float *o = views[lo].fimg + 3 * ((uint64_t)py * v.w + px);
            o[0] = cr + T * v.bg[0];
            o[1] = cg + T * v.bg[1];
            o[2] = cb + T * v.bg[2];
        }
    } else if (inside) {
        uint8_t *o = rgb + 3 * ((uint64_t)py * v.w + px);
        o[0] = (uint8_t)to_u8(cr + T * v.bg[0]);
        o[1] = (uint8_t)to_u8(cg + T * v.bg[1]);
        o[2] = (uint8_t)to_u8(cb + T * v.bg[2]);
    }
}

template <int MODE>
void launch_mode(hipStream_t s, const ViewParams *views, uint32_t nviews, uint32_t total_tiles, uint32_t total_sbs,
                 const uint32_t *tile_start, const uint32_t *sval, const uint4 *rec,
                 uint32_t *counters, float *pix_T, uint32_t *pix_last, uint32_t *done, float4 *pstate, int batch) {
    const dim3 grid(blend_grid(total_sbs)), block(BLEND_THREADS);
    switch (batch) {
    case 64:
        hipLaunchKernelGGL((k_blend<64, MODE>), grid, block, 0, s, views, nviews, total_tiles, tile_start, sval,
                           rec, total_sbs, counters, pix_T, pix_last, done, pstate);
        break;
    case 128:
        hipLaunchKernelGGL((k_blend<128, MODE>), grid, block, 0, s, views, nviews, total_tiles, tile_start, sval,
                           rec, total_sbs, counters, pix_T, pix_last, done, pstate);
        break;
    default:
        hipLaunchKernelGGL((k_blend<256, MODE>), grid, block, 0, s, views, nviews, total_tiles, tile_start, sval,
                           rec, total_sbs, counters, pix_T, pix_last, done, pstate);
        break;
    }
}

}  // namespace

hipError_t launch_blend(hipStream_t s, int mode, const ViewParams *views, uint32_t nviews, uint32_t total_tiles,
                        uint32_t total_sbs, const uint32_t *tile_start, const uint32_t *sval, const uint4 *rec,
                        uint32_t *counters, float *pix_T, uint32_t *pix_last, uint32_t *done, float *pstate,
                        int batch) {
    if (total_tiles == 0) return hipSuccess;
    float4 *ps = reinterpret_cast<float4 *>(pstate);
    switch (mode) {
    case BM_DIFF:
        launch_mode<BM_DIFF>(s, views, nviews, total_tiles, total_sbs, tile_start, sval, rec, counters, pix_T, pix_last,
                             done, ps, batch);
        break;
    case BM_FIRST:
        launch_mode<BM_FIRST>(s, views, nviews, total_tiles, total_sbs, tile_start, sval, rec, counters, pix_T,
                              pix_last, done, ps, batch);
        break;
    case BM_SECOND:
        launch_mode<BM_SECOND>(s, views, nviews, total_tiles, total_sbs, tile_start, sval, rec, counters, pix_T,
                               pix_last, done, ps, batch);
        break;
    default:
        launch_mode<BM_RGB>(s, views, nviews, total_tiles, total_sbs, tile_start, sval, rec, counters, pix_T, pix_last,
                            done, ps, batch);
        break;
    }
    return hipGetLastError();
}

}  // namespace sp
