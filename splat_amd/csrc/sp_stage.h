// Per-tile staging of projected records for the blend and its adjoint: the
// power terms of each tile column and row, and the reach bits of each quadrant.
#pragma once

#include "sp_device.h"

namespace sp {

// A record as staged for one tile. The power at pixel (column c, row r) of the
// tile is (col[c].y + row[r].x) + row[r].y * col[c].x, the reference's
// (qx + qy) + (-b dy) dx with every factor rounded as it rounds them, so a
// visit needs three operations for it. A column or row outside the record's
// pixel box stores a NaN in place of qx or qy: the power is then NaN, which
// fails `pw <= 0`, so the pixel is masked as the reference masks it by the box
// (whatever the opacity).
// 272 bytes, an odd number of 16-byte units, so the 16-byte writes of eight
// consecutive staging lanes fall in distinct LDS bank groups.
struct StagedRec {
    float2 col[TILE];  // (dx, qx = (dx * dx) * (-a/2)) per tile column
    float2 row[TILE];  // (qy = ((-c/2) * dy) * dy, -b * dy) per tile row
    float4 orgb;       // opacity, colour
};
static_assert(sizeof(StagedRec) == 272, "staged record size");

// Stages part `part` of P of one record for the tile at (ox, oy): its
// columns and rows [part * 16/P, (part + 1) * 16/P) and the reach bits of its
// quadrants [part * 4/P, (part + 1) * 4/P) (one byte per quadrant in qb), so
// the P lanes of a record share the work and every wave of the tile stages.
// The offsets are the reference's (src/render/cpu.rs:658-672):
// cx0 = (ox + 0.5) - x, dx = lane + (4*group + cx0) per column,
// dy = ((oy + r) + 0.5) - y per row. Frame coordinates are below 2^16, so
// (oy + r0) + (j + 0.5) is exact and equals (oy + r) + 0.5 for r = r0 + j.
template <int P, bool EXACT>
__device__ __forceinline__ void stage_part(StagedRec &d, uint8_t *qb, const uint4 &r0, const uint4 &r1,
                                           const uint4 &r2, uint32_t ox, uint32_t oy, uint32_t part) {
    constexpr int NC = (int)TILE / P, NQ = 4 / P;
    const int c0 = (int)part * NC;
    const float sx = __uint_as_float(r0.x), sy = __uint_as_float(r0.y);
    const float na = __uint_as_float(r0.z), nb2 = __uint_as_float(r0.w), nc = __uint_as_float(r1.x);
    const float thr = __uint_as_float(r2.w);
    const uint32_t bx0 = r2.y & 0xFFFFu, by0 = r2.y >> 16, bx1 = r2.z & 0xFFFFu, by1 = r2.z >> 16;
    const int lx0 = (int)bx0 - (int)ox, lx1 = (int)bx1 - (int)ox;
    const int ly0 = (int)by0 - (int)oy, ly1 = (int)by1 - (int)oy;
    const float cx0 = ((float)ox + 0.5f) - sx;
    const float qnan = __builtin_nanf("");
#pragma unroll
    for (int gi = 0; gi < NC; gi += 4) {
        const int c = c0 + gi;             // a multiple of 4: the first column of a 4-lane group
        const float g = (float)c + cx0;    // 4 * group + cx0
#pragma unroll
        for (int j = 0; j < 4; j += 2) {
            const float dxa = (float)j + g, dxb = (float)(j + 1) + g;
            const float qa = (dxa * dxa) * na, qb2 = (dxb * dxb) * na;
            *reinterpret_cast<float4 *>(&d.col[c + j]) =
                make_float4(dxa, (c + j >= lx0 && c + j < lx1) ? qa : qnan, dxb,
                            (c + j + 1 >= lx0 && c + j + 1 < lx1) ? qb2 : qnan);
        }
    }
    const float fy = (float)(oy + (uint32_t)c0);
#pragma unroll
    for (int j = 0; j < NC; j += 2) {
        const int r = c0 + j;
        const float dya = (fy + ((float)j + 0.5f)) - sy, dyb = (fy + ((float)j + 1.5f)) - sy;
        const float qa = (nc * dya) * dya, qb2 = (nc * dyb) * dyb;
        *reinterpret_cast<float4 *>(&d.row[r]) = make_float4((r >= ly0 && r < ly1) ? qa : qnan, nb2 * dya,
                                                             (r + 1 >= ly0 && r + 1 < ly1) ? qb2 : qnan, nb2 * dyb);
    }
    if (part == 0)
        d.orgb = make_float4(__uint_as_float(r1.y), __uint_as_float(r1.z), __uint_as_float(r1.w),
                             __uint_as_float(r2.x));
    // Colours are >= 0 and never NaN. In the EXACT blend a record with an
    // infinite colour marks every quadrant its box overlaps with a second bit
    // (value 3 instead of 1), whether or not its falloff reaches it: there the
    // reference's masked lanes turn NaN, and the blend walks a 64-record group
    // that holds such a record on its exact path.
    const bool inf_col = EXACT && (__builtin_isinf(__uint_as_float(r1.z)) || __builtin_isinf(__uint_as_float(r1.w)) ||
                                   __builtin_isinf(__uint_as_float(r2.x)));
    const float ia = 1.0f / (2.0f * na), ic = 1.0f / (2.0f * nc);
    uint32_t bits = 0;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        const uint32_t q = part * NQ + k;
        if (inf_col ? quadrant_overlaps(bx0, by0, bx1, by1, ox, oy, q)
                    : quadrant_reaches(sx, sy, na, nb2, nc, ia, ic, thr, bx0, by0, bx1, by1, ox, oy, q))
            bits |= (inf_col ? 3u : 1u) << (8 * k);
    }
    if constexpr (NQ == 4) *reinterpret_cast<uint32_t *>(qb) = bits;
    else if constexpr (NQ == 2) *reinterpret_cast<uint16_t *>(qb + 2 * part) = (uint16_t)bits;
    else qb[part] = (uint8_t)bits;
}

}  // namespace sp
