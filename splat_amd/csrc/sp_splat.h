// Per-splat device math shared by the forward projection and its adjoint:
// SH colour (degrees 0-3), parameter loading for every storage format,
// covariance from scale and rotation, and render::project of one splat into
// one view. The arithmetic follows the reference's operation order with
// contraction disabled (render::project, src/render/mod.rs:201-272;
// sh::eval_rest, src/sh.rs:39-64; scene::covariance, src/scene.rs:170-178).
#pragma once

#include "sp_device.h"

namespace sp {
namespace {

const __device__ float SH_C1 = 0.4886025f;
const __device__ float SH_C2[5] = {1.0925485f, -1.0925485f, 0.31539157f, -1.0925485f, 0.54627424f};
const __device__ float SH_C3[7] = {-0.5900436f, 2.8906114f, -0.4570458f, 0.37317633f,
                                   -0.4570458f, 1.4453057f, -0.5900436f};

template <int DEG>
constexpr int rest_of() {
    return (DEG + 1) * (DEG + 1) - 1;
}

// colour += eval_rest(deg, rest, dir), with r(k, ch) the SH rest coefficient
// k of channel ch. deg may be a run-time value; with a constant the compiler
// folds the tests, and either way the operations are the same.
template <typename R>
__device__ __forceinline__ void sh_color_impl(int deg, const R &r, float x, float y, float z, float c[3]) {
    if (deg < 1) return;
    float e[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) e[ch] = SH_C1 * (((-y) * r(0, ch) + z * r(1, ch)) - x * r(2, ch));
    if (deg >= 2) {
        const float xx = x * x, yy = y * y, zz = z * z;
        const float xy = x * y, yz = y * z, xz = x * z;
        const float b3 = SH_C2[0] * xy, b4 = SH_C2[1] * yz, b5 = SH_C2[2] * ((2.0f * zz - xx) - yy);
        const float b6 = SH_C2[3] * xz, b7 = SH_C2[4] * (xx - yy);
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            float s = b3 * r(3, ch);
            s = s + b4 * r(4, ch);
            s = s + b5 * r(5, ch);
            s = s + b6 * r(6, ch);
            s = s + b7 * r(7, ch);
            e[ch] = e[ch] + s;
        }
        if (deg >= 3) {
            const float b8 = (SH_C3[0] * y) * (3.0f * xx - yy);
            const float b9 = (SH_C3[1] * xy) * z;
            const float b10 = (SH_C3[2] * y) * ((4.0f * zz - xx) - yy);
            const float b11 = (SH_C3[3] * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy);
            const float b12 = (SH_C3[4] * x) * ((4.0f * zz - xx) - yy);
            const float b13 = (SH_C3[5] * z) * (xx - yy);
            const float b14 = (SH_C3[6] * x) * (xx - 3.0f * yy);
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                float s = b8 * r(8, ch);
                s = s + b9 * r(9, ch);
                s = s + b10 * r(10, ch);
                s = s + b11 * r(11, ch);
                s = s + b12 * r(12, ch);
                s = s + b13 * r(13, ch);
                s = s + b14 * r(14, ch);
                e[ch] = e[ch] + s;
            }
        }
    }
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) c[ch] = c[ch] + e[ch];
}

// colour += eval_rest(deg, rest, dir); sh holds coefficient k channel ch at 3k+ch.
template <int DEG>
__device__ __forceinline__ void sh_color(const float *sh, float x, float y, float z, float c[3]) {
    sh_color_impl(DEG, [&](int k, int ch) { return sh[3 * k + ch]; }, x, y, z, c);
}

// The splat's view direction as render::project takes it: normalize_or_zero
// of d = p - eye (glam: d * (1/|d|) when that is finite and > 0, else 0).
__device__ __forceinline__ void view_dir(float d0, float d1, float d2, float u[3]) {
    const float len = sqrtf((d0 * d0 + d1 * d1) + d2 * d2);
    const float rcp = 1.0f / len;
    u[0] = u[1] = u[2] = 0.0f;
    if (__builtin_isfinite(rcp) && rcp > 0.0f) {
        u[0] = d0 * rcp;
        u[1] = d1 * rcp;
        u[2] = d2 * rcp;
    }
}

// The colour of scene splat i seen from `eye` (f32 or 16-bit storage): the
// base colour plus the SH rest at the view direction, clamped at 0, as
// render::project computes it (src/render/mod.rs:250-262), from the splat's
// colour row, which holds its position too (read in 16-byte pieces: a gather
// of one row per lane, two 128-byte lines).
__device__ __forceinline__ void row_colour(const DevScene &sc, uint32_t i, const float eye[3], float col[3]) {
    constexpr uint32_t MAXW = 52;  // f32 colour_row of SH degree 3 (position 3, colour 3, rest 45, padding 1)
    const bool half = sc.storage != SP_STORE_F32;
    const uint32_t cw = colour_row(sc.rest, half);
    // v: position, base colour, SH rest as f32.
    float v[MAXW];
    if (!half) {
        const float4 *row = reinterpret_cast<const float4 *>(sc.crow + (uint64_t)cw * i);
#pragma unroll
        for (uint32_t q = 0; q < MAXW / 4; ++q) {
            const float4 x = 4 * q < cw ? row[q] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            v[4 * q] = x.x;
            v[4 * q + 1] = x.y;
            v[4 * q + 2] = x.z;
            v[4 * q + 3] = x.w;
        }
    } else {
        // Entries: six halves of position bits, then the 16-bit values.
        const uint4 *row = reinterpret_cast<const uint4 *>(sc.crow16 + (uint64_t)cw * i);
        const bool bf16 = sc.storage == SP_STORE_BF16;
        uint32_t w[(MAXW + 4) / 2];  // 56 halves as 28 words
#pragma unroll
        for (uint32_t q = 0; q < (MAXW + 4) / 8; ++q) {
            const uint4 x = 8 * q < cw ? row[q] : make_uint4(0u, 0u, 0u, 0u);
            w[4 * q] = x.x;
            w[4 * q + 1] = x.y;
            w[4 * q + 2] = x.z;
            w[4 * q + 3] = x.w;
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = __uint_as_float(w[c]);
#pragma unroll
        for (uint32_t e = 3; e < MAXW; ++e) {
            const uint32_t hidx = e + 3;  // half index in the row
            const uint16_t hv = (uint16_t)(w[hidx >> 1] >> (16 * (hidx & 1)));
            v[e] = bf16 ? __uint_as_float((uint32_t)hv << 16) : (float)__builtin_bit_cast(_Float16, hv);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) col[c] = v[3 + c];
    if (sc.deg > 0) {
        float u[3];
        view_dir(v[0] - eye[0], v[1] - eye[1], v[2] - eye[2], u);
        sh_color_impl((int)sc.deg, [&](int k, int ch) { return v[6 + 3 * k + ch]; }, u[0], u[1], u[2], col);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) col[c] = fmaxf(col[c], 0.0f);
}

// A splat's parameters, loaded once.
template <int DEG>
struct Splat {
    float opacity, k, p[3], cv[6], col[3];
    float sh[3 * rest_of<DEG>() + 1];
};

// 16-bit storage decode.
template <int STORE>
__device__ __forceinline__ float h2f(uint16_t h) {
    if (STORE == SP_STORE_BF16) return __uint_as_float((uint32_t)h << 16);
    return (float)__builtin_bit_cast(_Float16, h);
}

// cov3d from scale and a (re-normalised) rotation, f32, in the operation order
// of scene::covariance (src/scene.rs:170-178) and Scene::finalize's
// normalisation (src/scene.rs:72-79).
__device__ __forceinline__ void covariance_f32(const float sc[3], float q[4], float cv[6]) {
    const float len = sqrtf((q[0] * q[0] + q[2] * q[2]) + (q[1] * q[1] + q[3] * q[3]));
    if (__builtin_isfinite(len) && len > 1e-12f) {
#pragma unroll
        for (int c = 0; c < 4; ++c) q[c] = q[c] / len;
    } else {
        q[0] = q[1] = q[2] = 0.0f;
        q[3] = 1.0f;
    }
    const float x2 = q[0] + q[0], y2 = q[1] + q[1], z2 = q[2] + q[2];
    const float xx = q[0] * x2, xy = q[0] * y2, xz = q[0] * z2;
    const float yy = q[1] * y2, yz = q[1] * z2, zz = q[2] * z2;
    const float wx = q[3] * x2, wy = q[3] * y2, wz = q[3] * z2;
    const float r[3][3] = {{1.0f - (yy + zz), xy + wz, xz - wy},
                           {xy - wz, 1.0f - (xx + zz), yz + wx},
                           {xz + wy, yz - wx, 1.0f - (xx + yy)}};  // r[col][row]
    float m[3][3];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            m[j][c] = (r[0][c] * (j == 0 ? sc[0] : 0.0f) + r[1][c] * (j == 1 ? sc[1] : 0.0f)) +
                      r[2][c] * (j == 2 ? sc[2] : 0.0f);
    float s[3][3];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int c = 0; c < 3; ++c) s[j][c] = (m[0][c] * m[0][j] + m[1][c] * m[1][j]) + m[2][c] * m[2][j];
    cv[0] = s[0][0];
    cv[1] = s[1][0];
    cv[2] = s[2][0];
    cv[3] = s[1][1];
    cv[4] = s[2][1];
    cv[5] = s[2][2];
}

// Activated opacity and bounding-box radius k = sqrt(2 ln(o / ALPHA_MIN)) of a
// raw opacity logit (the differentiable path activates on the device).
__device__ __forceinline__ float sigmoid_f(float x) { return 1.0f / (1.0f + expf(-x)); }
__device__ __forceinline__ float k_of_opacity(float o) {
    return o < ALPHA_MIN ? 0.0f : sqrtf(2.0f * logf(o / ALPHA_MIN));
}

// COLOUR: also the base colour and SH rest (always for STORE_PARAMS).
template <int DEG, int STORE, bool COLOUR = true>
__device__ __forceinline__ void load_splat(const DevScene &sc, uint32_t i, Splat<DEG> &s) {
    const float *A = sc.soa;
    if (STORE == STORE_PARAMS) {
        const DevParams &P = sc.prm;
        const uint64_t i3 = 3ull * i;
#pragma unroll
        for (int c = 0; c < 3; ++c) s.p[c] = P.means[i3 + c];
        float scl[3], q[4];
#pragma unroll
        for (int c = 0; c < 3; ++c) scl[c] = expf(P.log_scales[i3 + c]);
#pragma unroll
        for (int c = 0; c < 4; ++c) q[c] = P.quats[4ull * i + c];
        covariance_f32(scl, q, s.cv);
        s.opacity = sigmoid_f(P.logits[i]);
        s.k = k_of_opacity(s.opacity);
#pragma unroll
        for (int c = 0; c < 3; ++c) s.col[c] = P.colors[i3 + c];
        const uint64_t sb = 3ull * rest_of<DEG>() * i;
#pragma unroll
        for (int c = 0; c < 3 * rest_of<DEG>(); ++c) s.sh[c] = P.sh_rest[sb + c];
    } else if (STORE == SP_STORE_F32) {
        constexpr uint32_t NC = A_SH + 3 * rest_of<DEG>();
        s.opacity = A[scene_at(NC, A_OPAC, i)];
        s.k = A[scene_at(NC, A_K, i)];
#pragma unroll
        for (int c = 0; c < 3; ++c) s.p[c] = A[scene_at(NC, A_POS + c, i)];
#pragma unroll
        for (int c = 0; c < 6; ++c) s.cv[c] = A[scene_at(NC, A_COV + c, i)];
        if constexpr (COLOUR) {
#pragma unroll
            for (int c = 0; c < 3; ++c) s.col[c] = A[scene_at(NC, A_COL + c, i)];
#pragma unroll
            for (int c = 0; c < 3 * rest_of<DEG>(); ++c) s.sh[c] = A[scene_at(NC, A_SH + c, i)];
        }
    } else {
        const uint16_t *B = sc.soa16;
        constexpr uint32_t NH = H_SH + 3 * rest_of<DEG>();
        s.opacity = A[scene_at(L_COLS, L_OPAC, i)];
        s.k = A[scene_at(L_COLS, L_K, i)];
#pragma unroll
        for (int c = 0; c < 3; ++c) s.p[c] = A[scene_at(L_COLS, L_POS + c, i)];
        float scl[3], q[4];
#pragma unroll
        for (int c = 0; c < 3; ++c) scl[c] = h2f<STORE>(B[scene_at(NH, H_SCALE + c, i)]);
#pragma unroll
        for (int c = 0; c < 4; ++c) q[c] = h2f<STORE>(B[scene_at(NH, H_QUAT + c, i)]);
        covariance_f32(scl, q, s.cv);
        if constexpr (COLOUR) {
#pragma unroll
            for (int c = 0; c < 3; ++c) s.col[c] = h2f<STORE>(B[scene_at(NH, H_COL + c, i)]);
#pragma unroll
            for (int c = 0; c < 3 * rest_of<DEG>(); ++c) s.sh[c] = h2f<STORE>(B[scene_at(NH, H_SH + c, i)]);
        }
    }
}

// The clamped colour of a loaded splat (base + SH at the view direction).
template <int DEG>
__device__ __forceinline__ void splat_colour(const Splat<DEG> &s, float d0, float d1, float d2, float col[3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) col[c] = s.col[c];
    if (DEG > 0) {
        float u[3];
        view_dir(d0, d1, d2, u);
        sh_color<DEG>(s.sh, u[0], u[1], u[2], col);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) col[c] = fmaxf(col[c], 0.0f);
}

// render::project of one splat into one view. Writes the 48-byte record and
// returns true when the splat survives culling. colour(d0, d1, d2, col) gives
// the splat's clamped colour for d = p - eye (it may leave zeros where the
// emission fills the colour in later).
template <int DEG, bool PREC, typename COL>
__device__ __forceinline__ bool project_view(const Splat<DEG> &s, const ViewCtx &v, uint4 *o, float *depth,
                                             float *dbg, uint16_t *dbg_bbox, uint2 *box, const COL &colour) {
    const float d0 = s.p[0] - v.eye[0], d1 = s.p[1] - v.eye[1], d2 = s.p[2] - v.eye[2];
    // t = rot * (p - eye), glam column order.
    const float tx = (v.rot[0][0] * d0 + v.rot[0][1] * d1) + v.rot[0][2] * d2;
    const float ty = (v.rot[1][0] * d0 + v.rot[1][1] * d1) + v.rot[1][2] * d2;
    const float tz = (v.rot[2][0] * d0 + v.rot[2][1] * d1) + v.rot[2][2] * d2;
    bool ok = !(s.opacity < ALPHA_MIN) && tz > v.near_z;
    const float inv_z = 1.0f / tz;
    const float x = v.focal * tx * inv_z + v.cx;
    const float y = v.focal * ty * inv_z + v.cy;
    float jx = tx * inv_z, jy = ty * inv_z;  // f32::clamp
    if (jx < -v.lim_x) jx = -v.lim_x;
    if (jx > v.lim_x) jx = v.lim_x;
    if (jy < -v.lim_y) jy = -v.lim_y;
    if (jy > v.lim_y) jy = v.lim_y;
    const float f = v.focal * inv_z;
    const float j0z = -f * jx, j1z = -f * jy;
    // Rows of T = J * R: (rot^T * j)[c] = right[c]*j.x + down[c]*j.y + forward[c]*j.z.
    float t0[3], t1[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        t0[c] = (v.rot[0][c] * f + v.rot[1][c] * 0.0f) + v.rot[2][c] * j0z;
        t1[c] = (v.rot[0][c] * 0.0f + v.rot[1][c] * f) + v.rot[2][c] * j1z;
    }
    const float *cv = s.cv;
    const float s00 = (cv[0] * t0[0] + cv[1] * t0[1]) + cv[2] * t0[2];
    const float s01 = (cv[1] * t0[0] + cv[3] * t0[1]) + cv[4] * t0[2];
    const float s02 = (cv[2] * t0[0] + cv[4] * t0[1]) + cv[5] * t0[2];
    const float s10 = (cv[0] * t1[0] + cv[1] * t1[1]) + cv[2] * t1[2];
    const float s11 = (cv[1] * t1[0] + cv[3] * t1[1]) + cv[4] * t1[2];
    const float s12 = (cv[2] * t1[0] + cv[4] * t1[1]) + cv[5] * t1[2];
    const float a = ((t0[0] * s00 + t0[1] * s01) + t0[2] * s02) + DILATION;
    const float b = (t1[0] * s00 + t1[1] * s01) + t1[2] * s02;
    const float cc = ((t1[0] * s10 + t1[1] * s11) + t1[2] * s12) + DILATION;
    const float det = a * cc - b * b;
    float inv_det;
    if (PREC && !(det > NEAR_SINGULAR * (a * cc))) {
        // Near-singular: the f32 determinant is mostly cancellation error.
        const double d64 = (double)a * (double)cc - (double)b * (double)b;
        ok = ok && d64 > 0.0;
        inv_det = (float)(1.0 / d64);
    } else {
        ok = ok && det > 0.0f;
        inv_det = 1.0f / det;
    }
    const float ext_x = s.k * sqrtf(a);
    const float ext_y = s.k * sqrtf(cc);
    ok = ok && !(x + ext_x <= 0.0f || x - ext_x >= v.width || y + ext_y <= 0.0f || y - ext_y >= v.height);
    ok = ok && __builtin_isfinite(x) && __builtin_isfinite(y) && __builtin_isfinite(ext_x) &&
         __builtin_isfinite(ext_y);
    if (!ok) return false;
    float col[3];
    colour(d0, d1, d2, col);
    const float ca = cc * inv_det, cb = -b * inv_det, cc2 = a * inv_det;
    // cpu::pixel_bbox (src/render/cpu.rs:96-102).
    const uint32_t bx0 = sat_u32(fmaxf(floorf(x - ext_x), 0.0f));
    const uint32_t by0 = sat_u32(fmaxf(floorf(y - ext_y), 0.0f));
    const uint32_t bx1 = min(sat_u32(fmaxf(ceilf(x + ext_x), 0.0f)), v.w);
    const uint32_t by1 = min(sat_u32(fmaxf(ceilf(y + ext_y), 0.0f)), v.h);
    const uint32_t lo = min(bx0, bx1) | (min(by0, by1) << 16);
    const uint32_t hi = bx1 | (by1 << 16);
    o[0] = make_uint4(__float_as_uint(x), __float_as_uint(y), __float_as_uint(-0.5f * ca), __float_as_uint(-cb));
    o[1] = make_uint4(__float_as_uint(-0.5f * cc2), __float_as_uint(s.opacity), __float_as_uint(col[0]),
                      __float_as_uint(col[1]));
    // Falloff threshold: pw < -k^2/2 - 0.02 gives exp(pw) * opacity below
    // ALPHA_MIN by a 2% margin (k^2/2 = ln(opacity / ALPHA_MIN)), far beyond the
    // rounding of exp_neg and of the product, so the blend may skip such pixels.
    const float thr = -0.5f * (s.k * s.k) - 0.02f;
    o[2] = make_uint4(__float_as_uint(col[2]), lo, hi, __float_as_uint(thr));
    if (box) *box = make_uint2(lo, hi);
    *depth = tz;
    if (dbg) {
        const float g[12] = {x, y, ca, cb, cc2, col[0], col[1], col[2], s.opacity, tz, ext_x, ext_y};
#pragma unroll
        for (int q = 0; q < 12; ++q) dbg[q] = g[q];
        dbg_bbox[0] = (uint16_t)min(bx0, bx1);
        dbg_bbox[1] = (uint16_t)min(by0, by1);
        dbg_bbox[2] = (uint16_t)bx1;
        dbg_bbox[3] = (uint16_t)by1;
    }
    return true;
}

}  // namespace
}  // namespace sp
