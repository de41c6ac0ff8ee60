// (e) Per-tile front-to-back alpha blending.
//
// One 256-thread workgroup per 16x16 tile of any view of the batch (the views'
// tiles are numbered consecutively, and blend_tile_of deals whole 4x4-tile
// superblocks to the XCDs in turn), one pixel per thread; wave k of the group
// owns one 8x8 quadrant. The tile's depth-ordered splat list is staged through
// LDS BATCH records at a time: 256 / BATCH lanes gather each 48-byte record,
// write its per-column and per-row power terms for this tile and bound its
// falloff over each quadrant. Then every wave walks the batch: a record whose
// pixel box misses the wave's quadrant, or whose falloff cannot reach the
// splat's alpha threshold at any in-box pixel of the quadrant, is never
// visited (the wave ballots the reach bits of 64 staged records at a time and
// walks only the set bits, in list order). A pixel stops at T < T_MIN; a wave
// stops when its quadrant has no live pixel, the tile when no pixel is live.
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
// EXACT instances serve scenes that may give a splat an infinite colour
// (flagged at upload): the reference's 4-lane groups then make NaN channels
// beside the lanes that blend, which they reproduce; other scenes never reach
// that code.
#include "sp_launch.h"
#include "sp_stage.h"

namespace sp {
namespace {

template <int BATCH, int MODE, bool EXACT>
__global__ __launch_bounds__(BLEND_THREADS) __attribute__((amdgpu_waves_per_eu(BATCH == 64 ? 8 : 1))) void k_blend(const ViewParams *__restrict__ views, uint32_t nviews,
                                                         uint32_t total_tiles, const uint32_t *__restrict__ tile_start,
                                                         const uint32_t *__restrict__ sval,
                                                         const uint4 *__restrict__ rec, uint32_t total_sbs,
                                                         uint32_t *__restrict__ counters, float *__restrict__ pix_T,
                                                         uint32_t *__restrict__ pix_last, uint32_t *__restrict__ done,
                                                         float4 *__restrict__ pstate, uint32_t *__restrict__ openm,
                                                         uint32_t *__restrict__ vopen) {
    constexpr int P = BLEND_THREADS / BATCH;  // lanes staging each record
    static_assert(P * BATCH == BLEND_THREADS && (P == 1 || P == 2 || P == 4), "staging split");
    __shared__ StagedRec s_st[BATCH];
    __shared__ uint8_t s_qb[BATCH][4];  // reach bit of each staged record per quadrant
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
    const uint64_t pix = (uint64_t)gt * BLEND_THREADS + threadIdx.x;
    // Lanes outside the frame carry T = 0, so a pixel is live exactly when
    // T >= T_MIN (T only falls and is never NaN, so this is the reference's
    // stop at T < T_MIN for the frame's pixels): one compare, no loop-carried
    // mask. An outside lane's T is never written out as its own.
    float T = inside ? 1.0f : 0.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    if (MODE == BM_SECOND) {
        const float4 st = pstate[pix];
        T = inside ? st.x : 0.0f;
        cr = st.y;
        cg = st.z;
        cb = st.w;
    }
    const auto live = [&]() { return T >= T_MIN; };
    const auto live_mask = [&]() { return lane_mask(T >= T_MIN); };
    uint32_t last = 0;
    bool finished = false;
    const uint32_t start = tile_start[gt], end = tile_start[gt + 1];
    uint32_t b = start;
    for (; b < end; b += BATCH) {
        const uint32_t rr = threadIdx.x / P, idx = b + rr;
        if (idx < end) {
            const uint32_t id = sval[idx];
            const uint4 *r = rec + 3 * (uint64_t)id;
            const uint4 r0 = r[0], r1 = r[1], r2 = r[2];
            stage_part<P, EXACT>(s_st[rr], s_qb[rr], r0, r1, r2, ox, oy, threadIdx.x % P);
        }
        __syncthreads();
        const uint32_t nb = __builtin_amdgcn_readfirstlane(min(end - b, (uint32_t)BATCH));
        // Records of the batch that can reach this wave's quadrant, 64 at a
        // time: one LDS read per lane and a ballot, then the set bits in order.
        // Live lanes as a wave mask (the per-lane conditions below are all
        // wave masks, combined on the scalar unit); it changes only when a
        // record updates T, and the wave stops once its quadrant saturated.
        uint64_t lm = live_mask();
        for (uint32_t g0 = 0; g0 < nb && lm != 0ull; g0 += 64) {
            const uint32_t qm = g0 + lane_i < nb ? s_qb[g0 + lane_i][w] : 0u;
            unsigned long long reach = lane_mask(qm != 0u);
            if (reach == 0ull) continue;
            if (EXACT && lane_mask(qm > 1u) != 0ull) {
                // A record of the group has an infinite colour. The reference
                // blends 4-lane groups of a row (src/render/cpu.rs:673-690):
                // once any lane of a group is live and inside the record's
                // box, every lane of the group adds colour * (alpha * T) with
                // alpha zeroed where it does not blend, and inf * 0 = NaN. The
                // fast walk below changes only the lanes that blend, which is
                // the same for finite colours only; this walk reproduces the
                // groups. A staged term is NaN exactly outside the box.
                while (reach) {
                    const uint32_t kb = (uint32_t)__builtin_ctzll(reach);
                    asm("s_bitset0_b64 %0, %1" : "+s"(reach) : "s"(kb));
                    const StagedRec &sr = s_st[g0 + kb];
                    const float2 cv = sr.col[lx], rv = sr.row[ly];
                    const float4 c4 = sr.orgb;
                    const float pw = (cv.y + rv.x) + rv.y * cv.x;
                    const float alpha = fminf(exp_neg_times(pw, c4.x), ALPHA_MAX);
                    const uint64_t upd = lm & lane_mask(pw <= 0.0f) & lane_mask(alpha >= ALPHA_MIN);
                    const uint64_t act = lm & lane_mask(!__builtin_isnan(cv.y) && !__builtin_isnan(rv.x));
                    // Lanes whose 4-lane group (a nibble of the 8x8 quadrant's
                    // lane mask: 4 consecutive columns of a row, 4-aligned in
                    // the tile) holds an active lane.
                    const uint64_t grp = ((act | (act >> 1) | (act >> 2) | (act >> 3)) & 0x1111111111111111ull) * 0xFull;
                    if (__builtin_amdgcn_inverse_ballot_w64(grp)) {
                        const float a = __builtin_amdgcn_inverse_ballot_w64(upd) ? alpha : 0.0f;
                        const float wgt = a * T;
                        cr = cr + c4.y * wgt;
                        cg = cg + c4.z * wgt;
                        cb = cb + c4.w * wgt;
                        T = T * (1.0f - a);
                        if (MODE == BM_DIFF && __builtin_amdgcn_inverse_ballot_w64(upd))
                            last = (b - start) + g0 + kb + 1;
                    }
                    lm = live_mask();
                    if (lm == 0ull) break;
                }
                continue;
            }
            // The visits are software-pipelined over two register sets: each
            // reads the next set record's terms from LDS before its own
            // arithmetic, so the LDS latency overlaps the previous visit's
            // dependent chain, and the sets alternate (no copies). The scalar
            // work per visit is kept small, since the CU's scalar unit issues
            // one instruction per SIMD turn while the SIMD issues up to two
            // VALU: the next record's index is s_ff1 clamped to 63 (s_ff1 of
            // an empty mask is -1; record 63 of the group is an in-bounds
            // read whose values are then never used), the update is
            // branch-free (a lane that must not update blends alpha = 0,
            // which leaves its C and T unchanged: C + c * (0 * T) = C for a
            // finite colour and T * (1 - 0) = T exactly), and the
            // per-lane LDS byte offsets of the group's terms are set once per
            // group. 43 VALU and 9 scalar instructions per visit.
            const char *gbase = reinterpret_cast<const char *>(&s_st[g0]);
            const char *pcol = gbase + lx * sizeof(float2);
            const char *prow = gbase + offsetof(StagedRec, row) + ly * sizeof(float2);
            // A lane-varying register for the uniform offset too, so each read's
            // address is one vector add (a scalar add and a move otherwise).
            uint32_t orgb_off = (uint32_t)offsetof(StagedRec, orgb);
            asm("" : "+v"(orgb_off));
            const char *porgb = gbase + orgb_off;
            const auto load = [&](uint32_t k, float2 &cv, float2 &rv, float4 &c4) {
                const uint32_t off = k * (uint32_t)sizeof(StagedRec);
                cv = *reinterpret_cast<const float2 *>(pcol + off);
                rv = *reinterpret_cast<const float2 *>(prow + off);
                c4 = *reinterpret_cast<const float4 *>(porgb + off);
            };
            const auto first_set = [](uint64_t m) {
                uint32_t k;
                asm("s_ff1_i32_b64 %0, %1" : "=s"(k) : "s"(m));
                return min(k, 63u);
            };
            uint32_t kA = first_set(reach), kB = 0;
            // Clears the visited bit: one scalar instruction instead of the
            // three of a 64-bit reach &= reach - 1.
            asm("s_bitset0_b64 %0, %1" : "+s"(reach) : "s"(kA));
            float2 cvA, rvA, cvB, rvB;
            float4 c4A, c4B;
            load(kA, cvA, rvA, c4A);
            // Visits record kc (terms cv, rv, c4) after reading the terms of
            // the next set record kn into the other set. Returns whether the
            // walk goes on: a set record is left and a pixel of the quadrant
            // is live; then kn's bit is cleared.
            const auto visit = [&](const float2 &cv, const float2 &rv, const float4 &c4, uint32_t kc, float2 &cvn,
                                   float2 &rvn, float4 &c4n, uint32_t &kn) -> bool {
                kn = first_set(reach);
                load(kn, cvn, rvn, c4n);
                // The arithmetic starts from values this empty asm passes on,
                // after the reads above (its memory clobber keeps them ahead of
                // it): the compiler would otherwise issue the next record's
                // reads late in the visit, or sink them past the exit test.
                float cx = cv.x, cy = cv.y, rx = rv.x, ry = rv.y;
                asm volatile("" : "+v"(cx), "+v"(cy), "+v"(rx), "+v"(ry) : : "memory");
                const float pw = (cy + rx) + ry * cx;
                const float alpha = fminf(exp_neg_times(pw, c4.x), ALPHA_MAX);
                const uint64_t upd = lm & lane_mask(pw <= 0.0f) & lane_mask(alpha >= ALPHA_MIN);
                const float a = __builtin_amdgcn_inverse_ballot_w64(upd) ? alpha : 0.0f;
                const float wgt = a * T;
                cr = cr + c4.y * wgt;
                cg = cg + c4.z * wgt;
                cb = cb + c4.w * wgt;
                T = T * (1.0f - a);
                if (MODE == BM_DIFF && __builtin_amdgcn_inverse_ballot_w64(upd)) last = (b - start) + g0 + kc + 1;
                lm = live_mask();
                // One test ends the walk: a saturated quadrant empties the
                // reach mask (two scalar instructions; the compiler would
                // combine the two tests through five).
                asm("s_cmp_eq_u64 %1, 0\n\ts_cselect_b64 %0, 0, %0" : "+s"(reach) : "s"(lm) : "scc");
                if (reach == 0ull) return false;
                asm("s_bitset0_b64 %0, %1" : "+s"(reach) : "s"(kn));
                return true;
            };
            while (visit(cvA, rvA, c4A, kA, cvB, rvB, c4B, kB) && visit(cvB, rvB, c4B, kB, cvA, rvA, c4A, kA)) {
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
            if (threadIdx.x == 0) {
                done[gt] = 0u;
                // The tile's bit in its row's open-tile mask, and the view's count.
                atomicOr(&openm[views[lo].tile_base + ty * ((v.tiles_x + 31u) >> 5) + (tx >> 5)], 1u << (tx & 31u));
                atomicAdd(&vopen[lo], 1u);
            }
            return;
        }
        if (threadIdx.x == 0) done[gt] = 1u;
    }
    if (MODE == BM_DIFF) {
        pix_T[pix] = inside ? T : 1.0f;
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

template <int MODE, bool EXACT>
void launch_mode(hipStream_t s, const ViewParams *views, uint32_t nviews, uint32_t total_tiles, uint32_t total_sbs,
                 const uint32_t *tile_start, const uint32_t *sval, const uint4 *rec,
                 uint32_t *counters, float *pix_T, uint32_t *pix_last, uint32_t *done, float4 *pstate,
                 uint32_t *openm, uint32_t *vopen, int batch) {
    const dim3 grid(blend_grid(total_sbs)), block(BLEND_THREADS);
    switch (batch) {
    case 64:
        hipLaunchKernelGGL((k_blend<64, MODE, EXACT>), grid, block, 0, s, views, nviews, total_tiles, tile_start, sval,
                           rec, total_sbs, counters, pix_T, pix_last, done, pstate, openm, vopen);
        break;
    case 128:
        hipLaunchKernelGGL((k_blend<128, MODE, EXACT>), grid, block, 0, s, views, nviews, total_tiles, tile_start,
                           sval, rec, total_sbs, counters, pix_T, pix_last, done, pstate, openm, vopen);
        break;
    default:
        hipLaunchKernelGGL((k_blend<256, MODE, EXACT>), grid, block, 0, s, views, nviews, total_tiles, tile_start,
                           sval, rec, total_sbs, counters, pix_T, pix_last, done, pstate, openm, vopen);
        break;
    }
}

template <int MODE>
void launch_exact(hipStream_t s, bool exact, const ViewParams *views, uint32_t nviews, uint32_t total_tiles,
                  uint32_t total_sbs, const uint32_t *tile_start, const uint32_t *sval, const uint4 *rec,
                  uint32_t *counters, float *pix_T, uint32_t *pix_last, uint32_t *done, float4 *pstate,
                  uint32_t *openm, uint32_t *vopen, int batch) {
    if (exact)
        launch_mode<MODE, true>(s, views, nviews, total_tiles, total_sbs, tile_start, sval, rec, counters, pix_T,
                                pix_last, done, pstate, openm, vopen, batch);
    else
        launch_mode<MODE, false>(s, views, nviews, total_tiles, total_sbs, tile_start, sval, rec, counters, pix_T,
                                 pix_last, done, pstate, openm, vopen, batch);
}

}  // namespace

hipError_t launch_blend(hipStream_t s, int mode, const ViewParams *views, uint32_t nviews, uint32_t total_tiles,
                        uint32_t total_sbs, const uint32_t *tile_start, const uint32_t *sval, const uint4 *rec,
                        uint32_t *counters, float *pix_T, uint32_t *pix_last, uint32_t *done, float *pstate,
                        uint32_t *openm, uint32_t *vopen, int batch, bool exact) {
    if (total_tiles == 0) return hipSuccess;
    float4 *ps = reinterpret_cast<float4 *>(pstate);
    switch (mode) {
    case BM_DIFF:
        launch_mode<BM_DIFF, false>(s, views, nviews, total_tiles, total_sbs, tile_start, sval, rec, counters, pix_T,
                                    pix_last, done, ps, openm, vopen, batch);
        break;
    case BM_FIRST:
        launch_exact<BM_FIRST>(s, exact, views, nviews, total_tiles, total_sbs, tile_start, sval, rec, counters, pix_T,
                               pix_last, done, ps, openm, vopen, batch);
        break;
    case BM_SECOND:
        launch_exact<BM_SECOND>(s, exact, views, nviews, total_tiles, total_sbs, tile_start, sval, rec, counters,
                                pix_T, pix_last, done, ps, openm, vopen, batch);
        break;
    default:
        launch_exact<BM_RGB>(s, exact, views, nviews, total_tiles, total_sbs, tile_start, sval, rec, counters, pix_T,
                             pix_last, done, ps, openm, vopen, batch);
        break;
    }
    return hipGetLastError();
}

}  // namespace sp
