// Reach bound of a projected splat over the 8x8 quadrants of a 16x16 tile,
// shared by the forward blend, the adjoint blend and a host instance that the
// CPU tests drive (sp_debug_reach).
//
// A clear quadrant bit must prove that the reference blend (raster_tile,
// src/render/cpu.rs:655-686) masks the record at every pixel of the quadrant:
// there alpha = min(exp_neg(pw) * opacity, 0.99) < ALPHA_MIN. The record's
// threshold thr = -k^2/2 - 0.02 already lies 0.02 below the exact crossover
// ln(ALPHA_MIN / opacity), far beyond the error of exp_neg and of the
// product. What remains is the rounding of pw itself: the reference computes
// it in f32 from f32 offsets, dx = lane + (4g + ((ox + 0.5) - x)) and
// dy = ((oy + ly) + 0.5) - y, as ((dx*dx)*na + (nc*dy)*dy) + (nb2*dy)*dx.
// With u = 2^-24, D = |dx|, E = |dy| and S = |na| D^2 + |nc| E^2 + |nb2| D E:
//   - the offsets are off by at most u (3D + 19) and u E, which moves the
//     power by at most u (6|na| D^2 + 2|nc| E^2 + 4|nb2| D E + 38|na| D +
//     19|nb2| E);
//   - the five products and two sums round by at most 4u S;
//   - the bound below is itself evaluated in f32 from offsets rounded the same
//     way (at most 7u S). Its edge vertices are located with one reciprocal
//     per axis instead of a division per edge: a vertex off by e only moves
//     the evaluation to another point of the edge, at most 3 |coef| e^2 below
//     the edge's maximum, and e <= 3u |d| makes that under 27 u^2 S.
// All of it stays under 2^-19 (S + 2 (|na| + |nb2| + |nc|) (D + E)) with D
// and E the largest offsets of the rectangle, so that margin plus 0.01 is
// subtracted from thr before a bit is cleared. For ordinary splats the margin
// is a few 1e-4; for a needle (minor axis near the dilation limit, major axis
// of hundreds of pixels) the terms reach 1e6 and cancel along its axis, and
// the margin grows with them.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define SP_HD __host__ __device__
#else
#define SP_HD
#endif

namespace sp {

// Upper bound of the falloff power over a rectangle of offsets
// [dxa, dxb] x [dya, dyb] from the splat centre: the power
// na*dx^2 + nc*dy^2 + nb2*dx*dy (na, nc < 0) is concave with its maximum 0 at
// the centre; off the rectangle the maximum lies on an edge, where it is the
// clamped vertex of a 1D parabola. ia = 1 / (2 na), ic = 1 / (2 nc).
SP_HD inline float max_power(float na, float nb2, float nc, float ia, float ic, float dxa, float dxb, float dya,
                             float dyb) {
    if (dxa <= 0.0f && dxb >= 0.0f && dya <= 0.0f && dyb >= 0.0f) return 0.0f;
    float m = -INFINITY;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const float c = e ? dxb : dxa;  // edge dx = c
        const float d = fminf(fmaxf((-nb2 * c) * ic, dya), dyb);
        m = fmaxf(m, (na * c * c + nc * d * d) + nb2 * c * d);
    }
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const float c = e ? dyb : dya;  // edge dy = c
        const float d = fminf(fmaxf((-nb2 * c) * ia, dxa), dxb);
        m = fmaxf(m, (na * d * d + nc * c * c) + nb2 * c * d);
    }
    return m;
}

// Rounding margin of the power over that rectangle (the bound derived above).
SP_HD inline float power_margin(float na, float nb2, float nc, float dxa, float dxb, float dya, float dyb) {
    const float D = fmaxf(fabsf(dxa), fabsf(dxb)), E = fmaxf(fabsf(dya), fabsf(dyb));
    const float a = fabsf(na), b = fabsf(nb2), c = fabsf(nc);
    const float s = (a * D) * D + (c * E) * E + (b * D) * E;
    const float g = 2.0f * (a + b + c) * (D + E);
    return 0.01f + 1.9073486328125e-6f * (s + g);  // 2^-19
}

// Whether the record can reach its threshold in quadrant q ((q & 1, q >> 1) of
// the tile at ox, oy): its pixel box overlaps the quadrant and its power can
// reach the threshold at some pixel centre of the overlap, within the margin
// above (NaN keeps the bit). ia = 1 / (2 na), ic = 1 / (2 nc); box corners in
// pixels, [bx0, bx1) x [by0, by1).
SP_HD inline bool quadrant_reaches(float sx, float sy, float na, float nb2, float nc, float ia, float ic, float thr,
                                   uint32_t bx0, uint32_t by0, uint32_t bx1, uint32_t by1, uint32_t ox, uint32_t oy,
                                   uint32_t q) {
    const uint32_t qx0 = ox + 8u * (q & 1u), qy0 = oy + 8u * (q >> 1);
    const uint32_t xa = bx0 > qx0 ? bx0 : qx0, xb = bx1 < qx0 + 8u ? bx1 : qx0 + 8u;
    const uint32_t ya = by0 > qy0 ? by0 : qy0, yb = by1 < qy0 + 8u ? by1 : qy0 + 8u;
    if (xa >= xb || ya >= yb) return false;
    const float dxa = ((float)xa + 0.5f) - sx, dxb = ((float)xb - 0.5f) - sx;
    const float dya = ((float)ya + 0.5f) - sy, dyb = ((float)yb - 0.5f) - sy;
    const float pm = max_power(na, nb2, nc, ia, ic, dxa, dxb, dya, dyb);
    return !(pm < thr - power_margin(na, nb2, nc, dxa, dxb, dya, dyb));
}

// Whether the pixel box [bx0, bx1) x [by0, by1) overlaps quadrant q of the
// tile at (ox, oy), with no bound on the falloff: the reference adds
// colour * 0 to the masked lanes of the 4-lane groups that such a record
// touches, which only an infinite colour makes visible (inf * 0 = NaN).
SP_HD inline bool quadrant_overlaps(uint32_t bx0, uint32_t by0, uint32_t bx1, uint32_t by1, uint32_t ox, uint32_t oy,
                                    uint32_t q) {
    const uint32_t qx0 = ox + 8u * (q & 1u), qy0 = oy + 8u * (q >> 1);
    const uint32_t xa = bx0 > qx0 ? bx0 : qx0, xb = bx1 < qx0 + 8u ? bx1 : qx0 + 8u;
    const uint32_t ya = by0 > qy0 ? by0 : qy0, yb = by1 < qy0 + 8u ? by1 : qy0 + 8u;
    return xa < xb && ya < yb;
}

// Reach bits of a record over the tile's four quadrants (bit q as above).
// Record words: w0 = x, y, -a/2, -b; w1 = -c/2, opacity, r, g; w2 = b, box lo
// (x | y << 16), box hi, threshold.
SP_HD inline uint32_t quadrant_mask_words(const uint32_t w0[4], const uint32_t w1[4], const uint32_t w2[4],
                                          uint32_t ox, uint32_t oy) {
    const uint32_t bx0 = w2[1] & 0xFFFFu, by0 = w2[1] >> 16, bx1 = w2[2] & 0xFFFFu, by1 = w2[2] >> 16;
    float sx, sy, na, nb2, nc, thr;
    __builtin_memcpy(&sx, &w0[0], 4);
    __builtin_memcpy(&sy, &w0[1], 4);
    __builtin_memcpy(&na, &w0[2], 4);
    __builtin_memcpy(&nb2, &w0[3], 4);
    __builtin_memcpy(&nc, &w1[0], 4);
    __builtin_memcpy(&thr, &w2[3], 4);
    const float ia = 1.0f / (2.0f * na), ic = 1.0f / (2.0f * nc);
    uint32_t m = 0;
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q)
        if (quadrant_reaches(sx, sy, na, nb2, nc, ia, ic, thr, bx0, by0, bx1, by1, ox, oy, q)) m |= 1u << q;
    return m;
}

}  // namespace sp
