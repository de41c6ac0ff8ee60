// (a) EWA projection + SH colour (degrees 0-3) + compaction of visible splats.
//
// One thread per splat, structure-of-arrays reads (every field of splat i is
// a 4-byte load at array_base + i, so a wave reads 256 contiguous bytes per
// field). The arithmetic is render::project (src/render/mod.rs:201-272) and
// sh::eval_rest (src/sh.rs:39-64) in the reference's operation order, with
// contraction disabled, so every field is bit-identical to the CPU backend.
#include "sp_device.h"
#include "sp_launch.h"

namespace sp {
namespace {

const __device__ float SH_C1 = 0.4886025f;
const __device__ float SH_C2[5] = {1.0925485f, -1.0925485f, 0.31539157f, -1.0925485f, 0.54627424f};
const __device__ float SH_C3[7] = {-0.5900436f, 2.8906114f, -0.4570458f, 0.37317633f,
                                   -0.4570458f, 1.4453057f, -0.5900436f};

template <int DEG>
__device__ __forceinline__ void sh_color(const float *A, uint64_t S, uint32_t i, float x, float y, float z,
                                         float c[3]) {
    // Coefficient k, channel ch of splat i.
    auto r = [&](int k, int ch) { return A[(uint64_t)(A_SH + 3 * k + ch) * S + i]; };
    float e[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) e[ch] = SH_C1 * (((-y) * r(0, ch) + z * r(1, ch)) - x * r(2, ch));
    if (DEG >= 2) {
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
        if (DEG >= 3) {
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

template <int DEG>
__global__ __launch_bounds__(PROJECT_THREADS) void k_project(DevScene sc, ViewCtx v, uint4 *__restrict__ rec,
                                                             uint32_t *__restrict__ dkey,
                                                             uint32_t *__restrict__ dval, Lookback lb,
                                                             float *__restrict__ dbg,
                                                             uint16_t *__restrict__ dbg_bbox,
                                                             uint8_t *__restrict__ dbg_valid) {
    const Ticket tk = acquire_block(lb.counters);
    const uint32_t i = tk.bid * PROJECT_THREADS + threadIdx.x;
    const float *A = sc.soa;
    const uint64_t S = sc.stride;
    bool vis = false;
    float depth = 0.0f;
    if (i < sc.n) {
        const float opacity = A[(uint64_t)A_OPAC * S + i];
        const float px = A[(uint64_t)(A_POS + 0) * S + i];
        const float py = A[(uint64_t)(A_POS + 1) * S + i];
        const float pz = A[(uint64_t)(A_POS + 2) * S + i];
        float cv[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) cv[k] = A[(uint64_t)(A_COV + k) * S + i];
        const float kk = A[(uint64_t)A_K * S + i];
        const float d0 = px - v.eye[0], d1 = py - v.eye[1], d2 = pz - v.eye[2];
        // t = rot * (p - eye), glam column order.
        const float tx = (v.rot[0][0] * d0 + v.rot[0][1] * d1) + v.rot[0][2] * d2;
        const float ty = (v.rot[1][0] * d0 + v.rot[1][1] * d1) + v.rot[1][2] * d2;
        const float tz = (v.rot[2][0] * d0 + v.rot[2][1] * d1) + v.rot[2][2] * d2;
        bool ok = !(opacity < ALPHA_MIN) && tz > v.near_z;
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
        ok = ok && det > 0.0f;
        const float inv_det = 1.0f / det;
        const float ext_x = kk * sqrtf(a);
        const float ext_y = kk * sqrtf(cc);
        ok = ok && !(x + ext_x <= 0.0f || x - ext_x >= v.width || y + ext_y <= 0.0f || y - ext_y >= v.height);
        ok = ok && __builtin_isfinite(x) && __builtin_isfinite(y) && __builtin_isfinite(ext_x) &&
             __builtin_isfinite(ext_y);
        if (ok) {
            float col[3] = {A[(uint64_t)(A_COL + 0) * S + i], A[(uint64_t)(A_COL + 1) * S + i],
                            A[(uint64_t)(A_COL + 2) * S + i]};
            if (DEG > 0) {
                // normalize_or_zero(p - eye) (glam): d * (1/|d|) when that is finite and > 0.
                const float len = sqrtf((d0 * d0 + d1 * d1) + d2 * d2);
                const float rcp = 1.0f / len;
                float ux = 0.0f, uy = 0.0f, uz = 0.0f;
                if (__builtin_isfinite(rcp) && rcp > 0.0f) {
                    ux = d0 * rcp;
                    uy = d1 * rcp;
                    uz = d2 * rcp;
                }
                sh_color<DEG>(A, S, i, ux, uy, uz, col);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) col[c] = fmaxf(col[c], 0.0f);
            const float ca = cc * inv_det, cb = -b * inv_det, cc2 = a * inv_det;
            // cpu::pixel_bbox (src/render/cpu.rs:96-102).
            const uint32_t bx0 = sat_u32(fmaxf(floorf(x - ext_x), 0.0f));
            const uint32_t by0 = sat_u32(fmaxf(floorf(y - ext_y), 0.0f));
            const uint32_t bx1 = min(sat_u32(fmaxf(ceilf(x + ext_x), 0.0f)), v.w);
            const uint32_t by1 = min(sat_u32(fmaxf(ceilf(y + ext_y), 0.0f)), v.h);
            const uint32_t lo = min(bx0, bx1) | (min(by0, by1) << 16);
            const uint32_t hi = bx1 | (by1 << 16);
            uint4 *o = rec + 3 * (uint64_t)i;
            o[0] = make_uint4(__float_as_uint(x), __float_as_uint(y), __float_as_uint(-0.5f * ca),
                              __float_as_uint(-cb));
            o[1] = make_uint4(__float_as_uint(-0.5f * cc2), __float_as_uint(opacity), __float_as_uint(col[0]),
                              __float_as_uint(col[1]));
            o[2] = make_uint4(__float_as_uint(col[2]), lo, hi, __float_as_uint(tz));
            vis = true;
            depth = tz;
            if (dbg) {
                float *g = dbg + 12 * (uint64_t)i;
                g[0] = x;
                g[1] = y;
                g[2] = ca;
                g[3] = cb;
                g[4] = cc2;
                g[5] = col[0];
                g[6] = col[1];
                g[7] = col[2];
                g[8] = opacity;
                g[9] = tz;
                g[10] = ext_x;
                g[11] = ext_y;
                uint16_t *bb = dbg_bbox + 4 * (uint64_t)i;
                bb[0] = (uint16_t)min(bx0, bx1);
                bb[1] = (uint16_t)min(by0, by1);
                bb[2] = (uint16_t)bx1;
                bb[3] = (uint16_t)by1;
            }
        }
        if (dbg_valid) dbg_valid[i] = vis ? 1 : 0;
    }
    uint32_t total;
    const uint32_t local = block_exclusive_scan<PROJECT_THREADS>(vis ? 1u : 0u, &total);
    const uint32_t excl = lookback_scan1(lb.status, tk.bid, tk.epoch, total);
    if (vis && excl + local < sc.n) {  // always in range for a consistent scan; guards the buffer
        dkey[excl + local] = __float_as_uint(depth);
        dval[excl + local] = i;
    }
    if (tk.bid == gridDim.x - 1 && threadIdx.x == 0) lb.counters[CNT_VISIBLE] = excl + total;
    finish_block(lb.counters, tk.epoch);
}

}  // namespace

hipError_t launch_project(hipStream_t s, const DevScene &sc, const ViewCtx &v, uint4 *rec, uint32_t *dkey,
                          uint32_t *dval, Lookback lb, float *dbg, uint16_t *dbg_bbox, uint8_t *dbg_valid) {
    const uint32_t grid = sc.n ? (sc.n + PROJECT_THREADS - 1) / PROJECT_THREADS : 1;
    switch (sc.deg) {
    case 0:
        hipLaunchKernelGGL(k_project<0>, dim3(grid), dim3(PROJECT_THREADS), 0, s, sc, v, rec, dkey, dval, lb, dbg,
                           dbg_bbox, dbg_valid);
        break;
    case 1:
        hipLaunchKernelGGL(k_project<1>, dim3(grid), dim3(PROJECT_THREADS), 0, s, sc, v, rec, dkey, dval, lb, dbg,
                           dbg_bbox, dbg_valid);
        break;
    case 2:
        hipLaunchKernelGGL(k_project<2>, dim3(grid), dim3(PROJECT_THREADS), 0, s, sc, v, rec, dkey, dval, lb, dbg,
                           dbg_bbox, dbg_valid);
        break;
    default:
        hipLaunchKernelGGL(k_project<3>, dim3(grid), dim3(PROJECT_THREADS), 0, s, sc, v, rec, dkey, dval, lb, dbg,
                           dbg_bbox, dbg_valid);
        break;
    }
    return hipGetLastError();
}

}  // namespace sp
