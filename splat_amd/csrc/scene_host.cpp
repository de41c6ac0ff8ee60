// Host-side scene, camera and synthetic-workload helpers (libsplat_scene.so).
//
// f32 arithmetic in the reference's operation order (glam 0.29 scalar Vec3 /
// Mat3 on x86-64; no fused multiply-add: build with -ffp-contract=off).
// Transcendentals (logf, expf, sinf, cosf, tanf) come from the host libm, as
// Rust's f32 methods do on Linux.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "half_conv.h"
#include "splat_scene.h"

namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTau = 6.28318530717958647692f;

struct V3 {
    float x, y, z;
};

inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// glam Vec3::cross (scalar path).
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
// glam Vec3::normalize: self * (1 / length).
inline V3 normalize(V3 a) {
    float r = 1.0f / std::sqrt(dot(a, a));
    return mul(a, r);
}

V3 up_vector(int32_t up) {
    switch (up) {
    case SP_UP_POS_Y: return {0.f, 1.f, 0.f};
    case SP_UP_NEG_Y: return {0.f, -1.f, 0.f};
    case SP_UP_POS_Z: return {0.f, 0.f, 1.f};
    default: return {0.f, 0.f, -1.f};
    }
}

// UpAxis::horizontal (src/camera.rs:56-62).
void horizontal(int32_t up_axis, V3 *a, V3 *b) {
    V3 up = up_vector(up_axis);
    V3 seed = std::fabs(up.y) > 0.5f ? V3{0.f, 0.f, 1.f} : V3{0.f, 1.f, 0.f};
    *a = normalize(cross(seed, up));
    *b = normalize(cross(up, *a));
}

// Rust f32::to_radians: self * (PI / 180).
inline float to_radians(float deg) { return deg * (kPi / 180.0f); }

inline float sigmoid(float x) { return 1.0f / (1.0f + std::exp(-x)); }

}  // namespace

extern "C" {

// scene::covariance (src/scene.rs:170-178): m = Mat3::from_quat(q) * diag(s),
// s = m * m^T, upper triangle. Mat3 products follow glam's column sums
// ((x_axis*v.x + y_axis*v.y) + z_axis*v.z), zero terms included.
void sp_covariance(const float *scales, const float *quats, uint64_t n, float *cov_out) {
    for (uint64_t i = 0; i < n; ++i) {
        const float sx = scales[3 * i], sy = scales[3 * i + 1], sz = scales[3 * i + 2];
        const float qx = quats[4 * i], qy = quats[4 * i + 1], qz = quats[4 * i + 2],
                    qw = quats[4 * i + 3];
        // Mat3::from_quat.
        const float x2 = qx + qx, y2 = qy + qy, z2 = qz + qz;
        const float xx = qx * x2, xy = qx * y2, xz = qx * z2;
        const float yy = qy * y2, yz = qy * z2, zz = qz * z2;
        const float wx = qw * x2, wy = qw * y2, wz = qw * z2;
        const float r[3][3] = {// r[col][row]
                               {1.0f - (yy + zz), xy + wz, xz - wy},
                               {xy - wz, 1.0f - (xx + zz), yz + wx},
                               {xz + wy, yz - wx, 1.0f - (xx + yy)}};
        // m = r * diag(s): column j = r * (s_j e_j).
        const float d[3][3] = {{sx, 0.f, 0.f}, {0.f, sy, 0.f}, {0.f, 0.f, sz}};
        float m[3][3];
        for (int j = 0; j < 3; ++j)
            for (int c = 0; c < 3; ++c)
                m[j][c] = (r[0][c] * d[j][0] + r[1][c] * d[j][1]) + r[2][c] * d[j][2];
        // s = m * m^T: column j = m * (m^T column j) = m * (m[0][j], m[1][j], m[2][j]).
        float s[3][3];
        for (int j = 0; j < 3; ++j)
            for (int c = 0; c < 3; ++c)
                s[j][c] = (m[0][c] * m[0][j] + m[1][c] * m[1][j]) + m[2][c] * m[2][j];
        float *o = cov_out + 6 * i;
        o[0] = s[0][0];  // xx
        o[1] = s[1][0];  // xy
        o[2] = s[2][0];  // xz
        o[3] = s[1][1];  // yy
        o[4] = s[2][1];  // yz
        o[5] = s[2][2];  // zz
    }
}

// Scene::finalize (src/scene.rs:62-94). Quat::length follows glam's SSE2
// 4-lane dot ((x^2 + z^2) + (y^2 + w^2)).
void sp_finalize(uint64_t n, const float *scales, float *quats, float *opacities, float *cov_out) {
    for (uint64_t i = 0; i < n; ++i) {
        float *q = quats + 4 * i;
        const float d = (q[0] * q[0] + q[2] * q[2]) + (q[1] * q[1] + q[3] * q[3]);
        const float len = std::sqrt(d);
        if (std::isfinite(len) && len > 1e-12f) {
            for (int c = 0; c < 4; ++c) q[c] = q[c] / len;
        } else {
            q[0] = q[1] = q[2] = 0.f;
            q[3] = 1.f;
        }
        float o = opacities[i];
        opacities[i] = std::isfinite(o) ? std::min(std::max(o, 0.0f), 1.0f) : 0.0f;
    }
    sp_covariance(scales, quats, n, cov_out);
}

// Scene::bounds (src/scene.rs:123-166).
void sp_bounds(const float *positions, uint64_t n, float center_out[3], float *radius_out) {
    center_out[0] = center_out[1] = center_out[2] = 0.f;
    *radius_out = 1.f;
    if (n == 0) return;
    const uint64_t step = std::max<uint64_t>(n / 200000, 1);
    std::vector<V3> pts;
    for (uint64_t i = 0; i < n; i += step) {
        V3 p{positions[3 * i], positions[3 * i + 1], positions[3 * i + 2]};
        if (std::isfinite(p.x) && std::isfinite(p.y) && std::isfinite(p.z)) pts.push_back(p);
    }
    if (pts.empty()) return;
    // f32::total_cmp on finite values is the plain order except -0 < +0.
    auto total_less = [](float a, float b) {
        int32_t ia, ib;
        std::memcpy(&ia, &a, 4);
        std::memcpy(&ib, &b, 4);
        ia ^= (int32_t)((uint32_t)(ia >> 31) >> 1);
        ib ^= (int32_t)((uint32_t)(ib >> 31) >> 1);
        return ia < ib;
    };
    auto median = [&](std::vector<float> v) {
        size_t mid = v.size() / 2;
        std::nth_element(v.begin(), v.begin() + mid, v.end(), total_less);
        return v[mid];
    };
    std::vector<float> tmp(pts.size());
    float c[3];
    for (int a = 0; a < 3; ++a) {
        for (size_t i = 0; i < pts.size(); ++i) tmp[i] = a == 0 ? pts[i].x : (a == 1 ? pts[i].y : pts[i].z);
        c[a] = median(tmp);
    }
    V3 center{c[0], c[1], c[2]};
    std::vector<float> d(pts.size());
    for (size_t i = 0; i < pts.size(); ++i) {
        V3 e = sub(pts[i], center);
        d[i] = std::sqrt(dot(e, e));
    }
    auto pct = [&](float q) {
        size_t idx = std::min((size_t)((float)d.size() * q), d.size() - 1);
        std::nth_element(d.begin(), d.begin() + idx, d.end(), total_less);
        return d[idx];
    };
    const float p90 = pct(0.9f);
    const float p50 = pct(0.5f);
    const float radius = std::min(p90, 2.0f * p50);
    center_out[0] = center.x;
    center_out[1] = center.y;
    center_out[2] = center.z;
    *radius_out = radius > 1e-6f ? radius : 1.0f;
}

// OrbitCamera::new (src/camera.rs:87-103).
void sp_orbit_init(sp_orbit *cam, const float center[3], float radius, int32_t up) {
    const float fov_y = to_radians(50.0f);
    const float distance = radius / std::tan(fov_y * 0.5f) * 1.1f;
    cam->target[0] = center[0];
    cam->target[1] = center[1];
    cam->target[2] = center[2];
    cam->distance = distance;
    cam->yaw = 0.0f;
    cam->pitch = to_radians(15.0f);
    cam->fov_y = fov_y;
    cam->up = up;
    cam->scene_radius = radius;
}

// OrbitCamera::rotate (src/camera.rs:113-116).
void sp_orbit_rotate(sp_orbit *cam, float dyaw, float dpitch) {
    float r = std::fmod(cam->yaw + dyaw, kTau);
    cam->yaw = r < 0.0f ? r + kTau : r;
    const float max_pitch = 89.0f * kPi / 180.0f;
    cam->pitch = std::min(std::max(cam->pitch + dpitch, -max_pitch), max_pitch);
}

// OrbitCamera::zoom (src/camera.rs:119-123).
void sp_orbit_zoom(sp_orbit *cam, float factor) {
    const float lo = cam->scene_radius * 1e-3f, hi = cam->scene_radius * 1e3f;
    cam->distance = std::min(std::max(cam->distance * factor, lo), hi);
}

// OrbitCamera::eye + view + near (src/camera.rs:132-158).
void sp_orbit_view(const sp_orbit *cam, sp_view *out) {
    V3 a, b;
    horizontal(cam->up, &a, &b);
    const V3 up = up_vector(cam->up);
    const float sy = std::sin(cam->yaw), cy = std::cos(cam->yaw);
    const float sp = std::sin(cam->pitch), cp = std::cos(cam->pitch);
    const V3 dir = add(mul(add(mul(a, sy), mul(b, cy)), cp), mul(up, sp));
    const V3 target{cam->target[0], cam->target[1], cam->target[2]};
    const V3 eye = add(target, mul(dir, cam->distance));
    const V3 forward = normalize(sub(target, eye));
    const V3 right = normalize(cross(forward, up));
    const V3 down = cross(forward, right);
    const float e[3] = {eye.x, eye.y, eye.z}, r[3] = {right.x, right.y, right.z},
                d[3] = {down.x, down.y, down.z}, f[3] = {forward.x, forward.y, forward.z};
    std::memcpy(out->eye, e, sizeof e);
    std::memcpy(out->right, r, sizeof r);
    std::memcpy(out->down, d, sizeof d);
    std::memcpy(out->forward, f, sizeof f);
    out->fov_y = cam->fov_y;
    out->near_z = std::max(cam->scene_radius * 0.05f, 1e-4f);
}

// gen::Rng (src/gen.rs:22-54): SplitMix64.
uint64_t sp_rng_next_u64(uint64_t *s) {
    *s += 0x9E3779B97F4A7C15ull;
    uint64_t z = *s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

float sp_rng_f(uint64_t *s) { return (float)(sp_rng_next_u64(s) >> 40) / (float)(1ull << 24); }

float sp_rng_normal(uint64_t *s) {
    const float u = std::max(sp_rng_f(s), 1e-7f);
    const float v = sp_rng_f(s);
    return std::sqrt(-2.0f * std::log(u)) * std::cos(kTau * v);
}

void sp_synth_random(const sp_random_spec *spec, float *positions, float *log_scales,
                     float *scales, float *quats, float *opacity_logits, float *opacities,
                     float *colors, float *sh_rest) {
    uint64_t s = spec->seed;
    const uint32_t rest = (spec->sh_degree + 1) * (spec->sh_degree + 1) - 1;
    for (uint64_t i = 0; i < spec->n; ++i) {
        float p[3];
        for (;;) {
            p[0] = sp_rng_normal(&s);
            p[1] = sp_rng_normal(&s);
            p[2] = sp_rng_normal(&s);
            if (std::fabs(p[0]) < spec->clip && std::fabs(p[1]) < spec->clip &&
                std::fabs(p[2]) < spec->clip)
                break;
        }
        for (int c = 0; c < 3; ++c) positions[3 * i + c] = p[c];
        for (int c = 0; c < 3; ++c) {
            const float u = sp_rng_f(&s);
            const float ls = spec->ls_lo + (spec->ls_hi - spec->ls_lo) * u;
            if (log_scales) log_scales[3 * i + c] = ls;
            scales[3 * i + c] = std::exp(ls);
        }
        // Rng::quat: normalize(N(0,1)^4) (glam Quat::normalize = self * (1/length)).
        float q[4];
        for (int c = 0; c < 4; ++c) q[c] = sp_rng_normal(&s);
        const float len = std::sqrt((q[0] * q[0] + q[2] * q[2]) + (q[1] * q[1] + q[3] * q[3]));
        const float rl = 1.0f / len;
        for (int c = 0; c < 4; ++c) quats[4 * i + c] = q[c] * rl;
        const float logit = spec->op_sigma * sp_rng_normal(&s);
        if (opacity_logits) opacity_logits[i] = logit;
        opacities[i] = sigmoid(logit);
        for (int c = 0; c < 3; ++c) colors[3 * i + c] = spec->dc_mean + spec->dc_sigma * sp_rng_normal(&s);
        for (uint32_t k = 0; k < 3 * rest; ++k) sh_rest[3 * rest * i + k] = spec->sh_sigma * sp_rng_normal(&s);
    }
}

void sp_synth_lattice(uint32_t nx, uint32_t ny, uint32_t nz, float scale, float opacity,
                      float *positions, float *scales, float *quats, float *opacities,
                      float *colors) {
    uint64_t i = 0;
    for (uint32_t a = 0; a < nx; ++a)
        for (uint32_t b = 0; b < ny; ++b)
            for (uint32_t c = 0; c < nz; ++c, ++i) {
                const float p[3] = {(float)a / (float)(nx - 1), (float)b / (float)(ny - 1),
                                    (float)c / (float)(nz - 1)};
                for (int k = 0; k < 3; ++k) {
                    positions[3 * i + k] = p[k];
                    colors[3 * i + k] = p[k];
                    scales[3 * i + k] = scale;
                }
                quats[4 * i] = quats[4 * i + 1] = quats[4 * i + 2] = 0.f;
                quats[4 * i + 3] = 1.f;
                opacities[i] = opacity;
            }
}

void sp_to_f16(const float *in, uint64_t n, uint16_t *out) {
    for (uint64_t i = 0; i < n; ++i) out[i] = sp::f32_to_f16(in[i]);
}

void sp_to_bf16(const float *in, uint64_t n, uint16_t *out) {
    for (uint64_t i = 0; i < n; ++i) out[i] = sp::f32_to_bf16(in[i]);
}

double sp_psnr(const uint8_t *a, const uint8_t *b, uint64_t n_pixels) {
    if (n_pixels == 0) return 0.0;
    double se = 0.0;
    for (uint64_t i = 0; i < 3 * n_pixels; ++i) {
        const double d = (double)a[i] - (double)b[i];
        se += d * d;
    }
    const double mse = se / (double)(n_pixels * 3);
    if (mse == 0.0) return std::numeric_limits<double>::infinity();
    return 10.0 * std::log10(255.0 * 255.0 / mse);
}

}  // extern "C"
