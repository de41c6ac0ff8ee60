// ORACLE — test infrastructure only.
//
// A CPU restatement of splat-tui's `cpu` and `cpu-naive` rasterizers (the
// reference is Rust and cannot be compiled in this environment: no cargo, no
// crates). Only the tests, the smoke check and the CPU-baseline leg of the
// benchmark load this library, as the checker and the CPU timing baseline;
// the HIP rasterizer never calls it.
//
// Followed files (splat-tui, file:line):
//   constants            src/render/mod.rs:23-33, src/render/cpu.rs:34-40
//   ProjectCtx::new      src/render/mod.rs:168-197
//   project              src/render/mod.rs:201-272 (== cpu.rs project_fast :419-476)
//   power, to_u8         src/render/mod.rs:276-285
//   sh::eval_rest        src/sh.rs:39-64 (constants :5-22)
//   CpuNaive::render     src/render/naive.rs:29-100
//   Prep::new, spread3   src/render/cpu.rs:322-394
//   CpuTiled::render     src/render/cpu.rs:109-222, project_group :235-319
//   fix_ties             src/render/cpu.rs:399-414
//   raster_tile          src/render/cpu.rs:609-719
//   radix::sort_hi32     src/render/radix.rs:13-57
//   F4::exp_neg          src/render/simd.rs:315-331 (x86 lanes, unfused mul_add :212-214)
//   Scene::bounds        src/scene.rs:123-166
//
// All f32 arithmetic keeps the reference's operation order and is compiled
// without contraction (-ffp-contract=off), as Rust never fuses.
//
// Parity of this restatement against the Rust binary is pinned by the
// reference's own known-answer tests and invariants (see tests/test_oracle.py);
// bit-level agreement with glam 0.29's arithmetic is not pinnable here.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <sched.h>
#include <cstdlib>
#include <vector>

#include "splat_hip.h"

namespace {

constexpr float DILATION = 0.3f;
constexpr float ALPHA_MIN = 1.0f / 255.0f;
constexpr float ALPHA_MAX = 0.99f;
constexpr float T_MIN = 1e-4f;
constexpr uint32_t TILE = 16;
constexpr int LANES = 4;
constexpr int ROW_GROUPS = TILE / LANES;
constexpr int TILE_GROUPS = TILE * TILE / LANES;
constexpr size_t MIN_GROUP = 1024;
constexpr size_t GROUPS_PER_THREAD = 8;
constexpr size_t SMALL = 192;

const float SH_C1 = 0.4886025f;
const float SH_C2[5] = {1.0925485f, -1.0925485f, 0.31539157f, -1.0925485f, 0.54627424f};
const float SH_C3[7] = {-0.5900436f, 2.8906114f, -0.4570458f, 0.37317633f,
                        -0.4570458f, 1.4453057f, -0.5900436f};

inline uint32_t f2u(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float u2f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
// Rust `as u32` on f32: saturating, NaN -> 0.
inline uint32_t sat_u32(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)v;
}

struct Ctx {
    float rot[3][3];  // rows: right, down, forward
    float eye[3];
    float focal, cx, cy, width, height, near_z, lim_x, lim_y;
};

// ProjectCtx::new (src/render/mod.rs:168-197) + View::focal/rotation (src/camera.rs:174-181).
Ctx make_ctx(const sp_view &v, uint32_t w, uint32_t h) {
    Ctx c;
    c.focal = (float)h * 0.5f / std::tan(v.fov_y * 0.5f);
    const float fw = (float)w, fh = (float)h;
    auto lim = [&](float half) {
        float l = 1.3f * (half / c.focal);
        return std::isfinite(l) ? l : 0.0f;
    };
    for (int k = 0; k < 3; ++k) {
        c.rot[0][k] = v.right[k];
        c.rot[1][k] = v.down[k];
        c.rot[2][k] = v.forward[k];
        c.eye[k] = v.eye[k];
    }
    c.cx = fw * 0.5f;
    c.cy = fh * 0.5f;
    c.width = fw;
    c.height = fh;
    c.near_z = v.near_z;
    c.lim_x = lim(fw * 0.5f);
    c.lim_y = lim(fh * 0.5f);
    return c;
}

struct Splat2D {
    float x, y, conic[3], color[3], opacity, depth, ext_x, ext_y;
};

struct SceneRef {
    size_t n;
    uint32_t deg;
    uint32_t rest;  // rest coefficients per splat
    const float *pos, *cov, *opac, *col, *sh;
};

SceneRef scene_ref(const sp_scene *s) {
    SceneRef r;
    r.n = (size_t)s->n;
    r.deg = s->sh_degree;
    r.rest = (s->sh_degree + 1) * (s->sh_degree + 1) - 1;
    r.pos = s->positions;
    r.cov = s->cov3d;
    r.opac = s->opacities;
    r.col = s->colors;
    r.sh = s->sh_rest;
    return r;
}

// sh::eval_rest (src/sh.rs:39-64); rest holds `rest` RGB triples.
void eval_rest(uint32_t degree, const float *rest, uint32_t nrest, const float dir[3], float out[3]) {
    out[0] = out[1] = out[2] = 0.0f;
    if (degree == 0 || nrest == 0) return;
    const float x = dir[0], y = dir[1], z = dir[2];
    auto r = [&](int k, int ch) { return rest[3 * k + ch]; };
    float c[3];
    for (int ch = 0; ch < 3; ++ch) c[ch] = SH_C1 * (((-y) * r(0, ch) + z * r(1, ch)) - x * r(2, ch));
    if (degree >= 2) {
        const float xx = x * x, yy = y * y, zz = z * z;
        const float xy = x * y, yz = y * z, xz = x * z;
        for (int ch = 0; ch < 3; ++ch) {
            float s = (SH_C2[0] * xy) * r(3, ch);
            s = s + (SH_C2[1] * yz) * r(4, ch);
            s = s + (SH_C2[2] * ((2.0f * zz - xx) - yy)) * r(5, ch);
            s = s + (SH_C2[3] * xz) * r(6, ch);
            s = s + (SH_C2[4] * (xx - yy)) * r(7, ch);
            c[ch] = c[ch] + s;
        }
        if (degree >= 3) {
            for (int ch = 0; ch < 3; ++ch) {
                float s = ((SH_C3[0] * y) * (3.0f * xx - yy)) * r(8, ch);
                s = s + ((SH_C3[1] * xy) * z) * r(9, ch);
                s = s + ((SH_C3[2] * y) * ((4.0f * zz - xx) - yy)) * r(10, ch);
                s = s + ((SH_C3[3] * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy)) * r(11, ch);
                s = s + ((SH_C3[4] * x) * ((4.0f * zz - xx) - yy)) * r(12, ch);
                s = s + ((SH_C3[5] * z) * (xx - yy)) * r(13, ch);
                s = s + ((SH_C3[6] * x) * (xx - 3.0f * yy)) * r(14, ch);
                c[ch] = c[ch] + s;
            }
        }
    }
    out[0] = c[0];
    out[1] = c[1];
    out[2] = c[2];
}

// Prep's k (src/render/cpu.rs:384-388) == project's k (src/render/mod.rs:244).
inline float k_of(float o) { return o < ALPHA_MIN ? 0.0f : std::sqrt(2.0f * std::log(o / ALPHA_MIN)); }

// cpu::splat_color (src/render/cpu.rs:480-488).
void splat_color(const SceneRef &s, size_t i, const Ctx &ctx, float out[3]) {
    float c[3] = {s.col[3 * i], s.col[3 * i + 1], s.col[3 * i + 2]};
    if (s.deg > 0) {
        float d[3] = {s.pos[3 * i] - ctx.eye[0], s.pos[3 * i + 1] - ctx.eye[1], s.pos[3 * i + 2] - ctx.eye[2]};
        // glam normalize_or_zero: rcp = 1/length; finite and > 0 ? d*rcp : 0.
        const float len = std::sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
        const float rcp = 1.0f / len;
        float dir[3] = {0.f, 0.f, 0.f};
        if (std::isfinite(rcp) && rcp > 0.0f)
            for (int k = 0; k < 3; ++k) dir[k] = d[k] * rcp;
        float e[3];
        eval_rest(s.deg, s.sh + (size_t)3 * s.rest * i, s.rest, dir, e);
        for (int k = 0; k < 3; ++k) c[k] = c[k] + e[k];
    }
    for (int k = 0; k < 3; ++k) out[k] = std::fmax(c[k], 0.0f);
}

// render::project (src/render/mod.rs:201-272) with k supplied.
bool project(const SceneRef &s, size_t i, const Ctx &ctx, float k, Splat2D *out) {
    const float opacity = s.opac[i];
    if (opacity < ALPHA_MIN) return false;
    const float d[3] = {s.pos[3 * i] - ctx.eye[0], s.pos[3 * i + 1] - ctx.eye[1], s.pos[3 * i + 2] - ctx.eye[2]};
    float t[3];
    for (int r = 0; r < 3; ++r) t[r] = (ctx.rot[r][0] * d[0] + ctx.rot[r][1] * d[1]) + ctx.rot[r][2] * d[2];
    if (!(t[2] > ctx.near_z)) return false;
    const float inv_z = 1.0f / t[2];
    const float x = ctx.focal * t[0] * inv_z + ctx.cx;
    const float y = ctx.focal * t[1] * inv_z + ctx.cy;
    auto clamp = [](float v, float lo, float hi) {
        if (v < lo) v = lo;
        if (v > hi) v = hi;
        return v;
    };
    const float tx = clamp(t[0] * inv_z, -ctx.lim_x, ctx.lim_x);
    const float ty = clamp(t[1] * inv_z, -ctx.lim_y, ctx.lim_y);
    const float f = ctx.focal * inv_z;
    const float j0[3] = {f, 0.0f, -f * tx};
    const float j1[3] = {0.0f, f, -f * ty};
    // rt = rot^T: (rt * v)[c] = right[c]*v.x + down[c]*v.y + forward[c]*v.z.
    float t0[3], t1[3];
    for (int c = 0; c < 3; ++c) {
        t0[c] = (ctx.rot[0][c] * j0[0] + ctx.rot[1][c] * j0[1]) + ctx.rot[2][c] * j0[2];
        t1[c] = (ctx.rot[0][c] * j1[0] + ctx.rot[1][c] * j1[1]) + ctx.rot[2][c] * j1[2];
    }
    const float *c = s.cov + 6 * i;
    auto sigma = [&](const float v[3], float o[3]) {
        o[0] = (c[0] * v[0] + c[1] * v[1]) + c[2] * v[2];
        o[1] = (c[1] * v[0] + c[3] * v[1]) + c[4] * v[2];
        o[2] = (c[2] * v[0] + c[4] * v[1]) + c[5] * v[2];
    };
    auto dot = [](const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; };
    float s0[3], s1[3];
    sigma(t0, s0);
    sigma(t1, s1);
    const float a = dot(t0, s0) + DILATION;
    const float b = dot(t1, s0);
    const float cc = dot(t1, s1) + DILATION;
    const float det = a * cc - b * b;
    if (!(det > 0.0f)) return false;
    const float inv_det = 1.0f / det;
    const float ext_x = k * std::sqrt(a);
    const float ext_y = k * std::sqrt(cc);
    if (x + ext_x <= 0.0f || x - ext_x >= ctx.width || y + ext_y <= 0.0f || y - ext_y >= ctx.height) return false;
    float color[3];
    splat_color(s, i, ctx, color);
    if (!(std::isfinite(x) && std::isfinite(y) && std::isfinite(ext_x) && std::isfinite(ext_y))) return false;
    out->x = x;
    out->y = y;
    out->conic[0] = cc * inv_det;
    out->conic[1] = -b * inv_det;
    out->conic[2] = a * inv_det;
    for (int q = 0; q < 3; ++q) out->color[q] = color[q];
    out->opacity = opacity;
    out->depth = t[2];
    out->ext_x = ext_x;
    out->ext_y = ext_y;
    return true;
}

// cpu::pixel_bbox (src/render/cpu.rs:96-102).
void pixel_bbox(const Splat2D &s, uint32_t w, uint32_t h, uint16_t bb[4]) {
    const uint32_t x0 = sat_u32(std::fmax(std::floor(s.x - s.ext_x), 0.0f));
    const uint32_t y0 = sat_u32(std::fmax(std::floor(s.y - s.ext_y), 0.0f));
    const uint32_t x1 = std::min(sat_u32(std::fmax(std::ceil(s.x + s.ext_x), 0.0f)), w);
    const uint32_t y1 = std::min(sat_u32(std::fmax(std::ceil(s.y + s.ext_y), 0.0f)), h);
    bb[0] = (uint16_t)std::min(x0, x1);
    bb[1] = (uint16_t)std::min(y0, y1);
    bb[2] = (uint16_t)x1;
    bb[3] = (uint16_t)y1;
}

// F4::exp_neg (src/render/simd.rs:315-331), one lane: Cephes-style range
// reduction plus a degree-6 polynomial, x86 unfused mul_add (self + a*b).
inline float exp_neg(float xin) {
    const float LOG2E = 1.44269504088896340736f;
    const float C1 = 0.6933594f;
    const float C2 = -2.1219444e-4f;
    const float x = std::fmax(xin, -87.0f);
    const float n = std::nearbyint(x * LOG2E);
    const float f = (x - n * C1) - n * C2;
    float p = 1.9875691e-4f;
    p = 1.3981999e-3f + p * f;
    p = 8.333452e-3f + p * f;
    p = 4.1665796e-2f + p * f;
    p = 1.6666665e-1f + p * f;
    p = 0.5f + p * f;
    const float f2 = f * f;
    const float r = (f + 1.0f) + p * f2;
    return r * u2f((uint32_t)((int32_t)n + 127) << 23);
}

// to_u8 (src/render/mod.rs:282-285), one channel.
inline uint8_t q8(float v) {
    if (v < 0.0f) v = 0.0f;
    if (v > 1.0f) v = 1.0f;
    const float s = v * 255.0f + 0.5f;
    if (!(s > 0.0f)) return 0;
    if (s >= 255.0f) return 255;
    return (uint8_t)s;
}

// radix::sort_hi32 (src/render/radix.rs:13-57).
void sort_hi32(uint64_t *keys, uint64_t *tmp, size_t n) {
    if (n <= SMALL) {
        std::sort(keys, keys + n);
        return;
    }
    uint32_t hist[4][256];
    std::memset(hist, 0, sizeof hist);
    for (size_t i = 0; i < n; ++i) {
        const uint32_t h = (uint32_t)(keys[i] >> 32);
        hist[0][h & 0xff]++;
        hist[1][(h >> 8) & 0xff]++;
        hist[2][(h >> 16) & 0xff]++;
        hist[3][h >> 24]++;
    }
    uint64_t *src = keys, *dst = tmp;
    bool in_tmp = false;
    for (int d = 0; d < 4; ++d) {
        const int shift = 32 + 8 * d;
        uint32_t *h = hist[d];
        if ((size_t)h[(src[0] >> shift) & 0xff] == n) continue;
        uint32_t sum = 0;
        for (int c = 0; c < 256; ++c) {
            const uint32_t v = h[c];
            h[c] = sum;
            sum += v;
        }
        for (size_t i = 0; i < n; ++i) {
            const uint64_t k = src[i];
            const uint32_t b = (uint32_t)((k >> shift) & 0xff);
            dst[h[b]++] = k;
        }
        std::swap(src, dst);
        in_tmp = !in_tmp;
    }
    if (in_tmp) std::memcpy(dst, src, n * sizeof(uint64_t));
}

// cpu::fix_ties (src/render/cpu.rs:399-414).
void fix_ties(uint64_t *keys, size_t n, const uint32_t *orig) {
    size_t i = 0;
    while (i + 1 < n) {
        const uint64_t d = keys[i] >> 32;
        if (keys[i + 1] >> 32 != d) {
            ++i;
            continue;
        }
        size_t j = i + 2;
        while (j < n && keys[j] >> 32 == d) ++j;
        std::sort(keys + i, keys + j, [&](uint64_t a, uint64_t b) { return orig[(uint32_t)a] < orig[(uint32_t)b]; });
        i = j;
    }
}

// Dynamic-scheduled parallel loop on `threads` std::threads.
template <class F>
void parallel_for(size_t n, int threads, F &&fn) {
    if (threads <= 1 || n <= 1) {
        for (size_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    const int t = (int)std::min<size_t>((size_t)threads, n);
    for (int k = 0; k < t; ++k)
        pool.emplace_back([&] {
            for (;;) {
                const size_t i = next.fetch_add(1);
                if (i >= n) break;
                fn(i);
            }
        });
    for (auto &th : pool) th.join();
}

// Default thread count: the CPUs this process may run on, capped by
// OMP_NUM_THREADS when it is set (a host's share of CPUs per GPU).
int resolve_threads(int threads) {
    if (threads > 0) return threads;
    int n = 0;
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof set, &set) == 0) n = CPU_COUNT(&set);
    if (n <= 0) n = (int)std::thread::hardware_concurrency();
    if (const char *e = std::getenv("OMP_NUM_THREADS")) {
        const int cap = std::atoi(e);
        if (cap > 0 && (n <= 0 || cap < n)) n = cap;
    }
    return n > 0 ? n : 1;
}

uint32_t spread3(uint32_t v) {
    uint32_t x = v & 0x3ff;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// Scene::bounds (src/scene.rs:123-166).
void bounds(const float *pos, size_t n, float center[3], float *radius) {
    center[0] = center[1] = center[2] = 0.f;
    *radius = 1.f;
    if (n == 0) return;
    const size_t step = std::max<size_t>(n / 200000, 1);
    std::vector<float> px, py, pz;
    for (size_t i = 0; i < n; i += step) {
        const float *p = pos + 3 * i;
        if (std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2])) {
            px.push_back(p[0]);
            py.push_back(p[1]);
            pz.push_back(p[2]);
        }
    }
    if (px.empty()) return;
    auto total_less = [](float a, float b) {
        int32_t ia = (int32_t)f2u(a), ib = (int32_t)f2u(b);
        ia ^= (int32_t)((uint32_t)(ia >> 31) >> 1);
        ib ^= (int32_t)((uint32_t)(ib >> 31) >> 1);
        return ia < ib;
    };
    auto median = [&](std::vector<float> v) {
        const size_t mid = v.size() / 2;
        std::nth_element(v.begin(), v.begin() + mid, v.end(), total_less);
        return v[mid];
    };
    const float c[3] = {median(px), median(py), median(pz)};
    std::vector<float> d(px.size());
    for (size_t i = 0; i < px.size(); ++i) {
        const float e[3] = {px[i] - c[0], py[i] - c[1], pz[i] - c[2]};
        d[i] = std::sqrt((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]);
    }
    auto pct = [&](float q) {
        const size_t idx = std::min((size_t)((float)d.size() * q), d.size() - 1);
        std::nth_element(d.begin(), d.begin() + idx, d.end(), total_less);
        return d[idx];
    };
    const float p90 = pct(0.9f), p50 = pct(0.5f);
    const float r = std::min(p90, 2.0f * p50);
    center[0] = c[0];
    center[1] = c[1];
    center[2] = c[2];
    *radius = r > 1e-6f ? r : 1.0f;
}

// Per-splat data the tiled rasterizer needs (src/render/cpu.rs:43-54).
struct RSplat {
    float x, y, conic[3], color[3], opacity;
    uint16_t bbox[4];
};

struct Group {
    std::vector<uint64_t> keys;
    std::vector<std::pair<uint32_t, uint32_t>> tiles;
    std::vector<uint32_t> counts;
};

}  // namespace

// ---- CpuTiled restatement -------------------------------------------------

struct or_cpu {
    // Prep (src/render/cpu.rs:322-394): Morton-reordered SoA copy.
    size_t n;
    uint32_t deg, rest;
    std::vector<float> pos, cov, opac, col, sh, k;
    std::vector<uint32_t> orig;
    // Reused per-frame scratch.
    std::vector<Group> groups;
    std::vector<uint64_t> keys, sort_tmp;
    std::vector<RSplat> splats;
    std::vector<uint32_t> tile_start;
    SceneRef ref() const {
        SceneRef r;
        r.n = n;
        r.deg = deg;
        r.rest = rest;
        r.pos = pos.data();
        r.cov = cov.data();
        r.opac = opac.data();
        r.col = col.data();
        r.sh = sh.data();
        return r;
    }
};

namespace {

void project_group(or_cpu *h, Group &g, size_t base, size_t len, const Ctx &ctx, uint32_t tiles_x,
                   uint32_t w, uint32_t hh, size_t num_tiles) {
    const SceneRef s = h->ref();
    g.keys.clear();
    g.tiles.clear();
    g.counts.assign(num_tiles, 0);
    for (size_t i = base; i < base + len; ++i) {
        Splat2D p;
        if (!project(s, i, ctx, h->k[i], &p)) continue;
        uint16_t bb[4];
        pixel_bbox(p, w, hh, bb);
        const uint32_t rec = (uint32_t)g.keys.size();
        RSplat &r = h->splats[i];
        r.x = p.x;
        r.y = p.y;
        for (int q = 0; q < 3; ++q) {
            r.conic[q] = p.conic[q];
            r.color[q] = p.color[q];
        }
        r.opacity = p.opacity;
        std::memcpy(r.bbox, bb, sizeof bb);
        g.keys.push_back(((uint64_t)f2u(p.depth) << 32) | (uint64_t)i);
        if (bb[0] >= bb[2] || bb[1] >= bb[3]) continue;
        const uint32_t x0 = bb[0] / TILE, y0 = bb[1] / TILE;
        const uint32_t x1 = (bb[2] + TILE - 1) / TILE, y1 = (bb[3] + TILE - 1) / TILE;
        for (uint32_t ty = y0; ty < y1; ++ty)
            for (uint32_t tx = x0; tx < x1; ++tx) {
                const uint32_t t = ty * tiles_x + tx;
                g.tiles.emplace_back(t, rec);
                g.counts[t] += 1;
            }
    }
}

// raster_tile (src/render/cpu.rs:609-719). `work` counts live in-box pixel evaluations.
void raster_tile(const RSplat *splats, const uint64_t *keys, size_t nk, uint32_t ox, uint32_t oy, uint32_t width,
                 uint32_t height, const float bg[3], uint8_t *out, uint64_t *work, size_t *processed = nullptr) {
    const uint32_t tw = std::min(TILE, width - ox);
    const uint32_t th = std::min(TILE, height - oy);
    float tr[TILE_GROUPS][LANES], cr[TILE_GROUPS][LANES], cg[TILE_GROUPS][LANES], cb[TILE_GROUPS][LANES];
    for (int i = 0; i < TILE_GROUPS; ++i)
        for (int l = 0; l < LANES; ++l) {
            tr[i][l] = 1.0f;
            cr[i][l] = cg[i][l] = cb[i][l] = 0.0f;
        }
    const uint32_t full_row = (uint32_t)((1ull << tw) - 1);
    uint32_t live[TILE] = {0};
    for (uint32_t y = 0; y < th; ++y) live[y] = full_row;
    uint32_t live_rows = th;
    const float fx = (float)ox, fy = (float)oy;
    float dxs[ROW_GROUPS][LANES], qx[ROW_GROUPS][LANES];
    uint64_t evals = 0;
    if (processed) *processed = nk;
    for (size_t ki = 0; ki < nk; ++ki) {
        const RSplat &s = splats[(uint32_t)keys[ki]];
        const uint32_t bx0 = s.bbox[0], by0 = s.bbox[1], bx1 = s.bbox[2], by1 = s.bbox[3];
        const uint32_t lx0 = bx0 > ox ? bx0 - ox : 0;
        const uint32_t lx1 = std::min(bx1 > ox ? bx1 - ox : 0, tw);
        const uint32_t ly0 = by0 > oy ? by0 - oy : 0;
        const uint32_t ly1 = std::min(by1 > oy ? by1 - oy : 0, th);
        if (lx0 >= lx1) continue;
        const uint32_t bbox_bits = (uint32_t)((1ull << lx1) - 1) & ~((1u << lx0) - 1);
        const uint32_t g0 = lx0 / LANES, g1 = (lx1 + LANES - 1) / LANES;
        const float a = s.conic[0], b = s.conic[1], c = s.conic[2];
        const float cx0 = fx + 0.5f - s.x;
        const float ha = -0.5f * a;
        for (uint32_t g = g0; g < g1; ++g)
            for (int l = 0; l < LANES; ++l) {
                const float dx = (float)l + ((float)(g * LANES) + cx0);
                dxs[g][l] = dx;
                qx[g][l] = dx * dx * ha;
            }
        const float op = s.opacity, sr = s.color[0], sg = s.color[1], sb = s.color[2];
        for (uint32_t ly = ly0; ly < ly1; ++ly) {
            const uint32_t lm = live[ly] & bbox_bits;
            if (lm == 0) continue;
            const float dy = fy + (float)ly + 0.5f - s.y;
            const float qy = -0.5f * c * dy * dy;
            const float bdy = -b * dy;
            uint32_t sat = 0;
            for (uint32_t g = g0; g < g1; ++g) {
                const uint32_t gm = (lm >> (g * LANES)) & 0xF;
                if (gm == 0) continue;
                const uint32_t i = ly * ROW_GROUPS + g;
                for (int l = 0; l < LANES; ++l) {
                    const bool inbox = (gm >> l) & 1;
                    evals += inbox;
                    const float pw = (qx[g][l] + qy) + bdy * dxs[g][l];
                    float alpha = std::fmin(exp_neg(pw) * op, ALPHA_MAX);
                    const bool m = inbox && pw <= 0.0f && alpha >= ALPHA_MIN;
                    alpha = m ? alpha : 0.0f;
                    const float t = tr[i][l];
                    const float wgt = alpha * t;
                    cr[i][l] = cr[i][l] + sr * wgt;
                    cg[i][l] = cg[i][l] + sg * wgt;
                    cb[i][l] = cb[i][l] + sb * wgt;
                    const float nt = t * (1.0f - alpha);
                    tr[i][l] = nt;
                    if (nt < T_MIN && m) sat |= 1u << (g * LANES + l);
                }
            }
            if (sat != 0) {
                live[ly] &= ~sat;
                if (live[ly] == 0) live_rows -= 1;
            }
        }
        if (live_rows == 0) {
            if (processed) *processed = ki + 1;
            break;
        }
    }
    if (work) *work = evals;
    if (!out) return;
    for (uint32_t ly = 0; ly < th; ++ly) {
        uint8_t *row = out + 3 * ((size_t)(oy + ly) * width + ox);
        for (uint32_t lx = 0; lx < tw; ++lx) {
            const uint32_t i = ly * ROW_GROUPS + lx / LANES, l = lx % LANES;
            const float t = tr[i][l];
            row[3 * lx + 0] = q8(cr[i][l] + t * bg[0]);
            row[3 * lx + 1] = q8(cg[i][l] + t * bg[1]);
            row[3 * lx + 2] = q8(cb[i][l] + t * bg[2]);
        }
    }
}

struct Frame3 {
    uint32_t tiles_x, num_tiles;
};

// Steps 1-2 of CpuTiled::render: project + emit, counting sort by tile.
// Returns false on more than 2^32 pairs.
bool cpu_bucket(or_cpu *h, const sp_view *view, uint32_t w, uint32_t hh, int threads, Frame3 *fr,
                sp_stats *st) {
    const size_t n = h->n;
    const Ctx ctx = make_ctx(*view, w, hh);
    const uint32_t tiles_x = (w + TILE - 1) / TILE, tiles_y = (hh + TILE - 1) / TILE;
    const size_t num_tiles = (size_t)tiles_x * tiles_y;
    fr->tiles_x = tiles_x;
    fr->num_tiles = (uint32_t)num_tiles;
    const size_t jobs = std::max<size_t>(std::min((n + MIN_GROUP - 1) / MIN_GROUP, (size_t)threads * GROUPS_PER_THREAD), 1);
    const size_t group = std::max<size_t>((n + jobs - 1) / jobs, 1);
    const size_t ngroups = (n + group - 1) / group;
    h->groups.resize(ngroups);
    if (h->splats.size() != n) h->splats.assign(n, RSplat{});
    parallel_for(ngroups, threads, [&](size_t gi) {
        const size_t base = gi * group;
        project_group(h, h->groups[gi], base, std::min(group, n - base), ctx, tiles_x, w, hh, num_tiles);
    });
    uint64_t visible = 0;
    for (auto &g : h->groups) visible += g.keys.size();
    h->tile_start.assign(num_tiles + 1, 0);
    uint64_t total = 0;
    for (size_t t = 0; t < num_tiles; ++t) {
        h->tile_start[t] = (uint32_t)total;
        for (auto &g : h->groups) {
            const uint32_t c = g.counts[t];
            g.counts[t] = (uint32_t)total;
            total += c;
            if (total > 0xFFFFFFFFull) return false;
        }
    }
    h->tile_start[num_tiles] = (uint32_t)total;
    h->keys.resize(total);
    if (h->sort_tmp.size() < total) h->sort_tmp.resize(total);
    parallel_for(h->groups.size(), threads, [&](size_t gi) {
        Group &g = h->groups[gi];
        for (auto &tr : g.tiles) h->keys[g.counts[tr.first]++] = g.keys[tr.second];
    });
    st->visible = visible;
    st->overlaps = total;
    return true;
}

}  // namespace

extern "C" {

int or_resolve_threads(int threads) { return resolve_threads(threads); }

// Prep::new (src/render/cpu.rs:349-393).
or_cpu *or_cpu_create(const sp_scene *scene) {
    or_cpu *h = new or_cpu();
    const SceneRef s = scene_ref(scene);
    h->n = s.n;
    h->deg = s.deg;
    h->rest = s.rest;
    const size_t n = s.n;
    float center[3], radius;
    bounds(s.pos, n, center, &radius);
    float lo[3];
    for (int k = 0; k < 3; ++k) lo[k] = center[k] - 4.0f * radius;
    const float scale = 1024.0f / (8.0f * radius);
    std::vector<uint64_t> order(n);
    for (size_t i = 0; i < n; ++i) {
        uint32_t q[3];
        for (int k = 0; k < 3; ++k) {
            float v = (s.pos[3 * i + k] - lo[k]) * scale;
            v = std::fmin(std::fmax(v, 0.0f), 1023.0f);  // Vec3::clamp = max(min).min(max)
            q[k] = sat_u32(v);
        }
        const uint32_t code = spread3(q[0]) | spread3(q[1]) << 1 | spread3(q[2]) << 2;
        order[i] = ((uint64_t)code << 32) | (uint64_t)i;
    }
    std::sort(order.begin(), order.end());
    h->orig.resize(n);
    for (size_t j = 0; j < n; ++j) h->orig[j] = (uint32_t)order[j];
    h->pos.resize(3 * n);
    h->cov.resize(6 * n);
    h->opac.resize(n);
    h->col.resize(3 * n);
    h->sh.resize((size_t)3 * s.rest * n);
    h->k.resize(n);
    for (size_t j = 0; j < n; ++j) {
        const size_t i = h->orig[j];
        std::memcpy(&h->pos[3 * j], s.pos + 3 * i, 12);
        std::memcpy(&h->cov[6 * j], s.cov + 6 * i, 24);
        h->opac[j] = s.opac[i];
        std::memcpy(&h->col[3 * j], s.col + 3 * i, 12);
        if (s.rest) std::memcpy(&h->sh[(size_t)3 * s.rest * j], s.sh + (size_t)3 * s.rest * i, 12 * s.rest);
        h->k[j] = k_of(h->opac[j]);
    }
    return h;
}

void or_cpu_destroy(or_cpu *h) { delete h; }

// CpuTiled::render (src/render/cpu.rs:109-222). `out` is w*h*3 bytes (may be
// NULL to only count), `work` (optional) receives the number of live in-box
// pixel evaluations of the blend. Returns 0, or -1 on >2^32 pairs.
int or_cpu_render(or_cpu *h, const sp_view *view, uint32_t w, uint32_t hh, const float bg[3], uint8_t *out,
                  sp_stats *st, int threads, uint64_t *work) {
    sp_stats local;
    if (!st) st = &local;
    std::memset(st, 0, sizeof *st);
    if (work) *work = 0;
    if (w == 0 || hh == 0) return 0;
    threads = resolve_threads(threads);
    Frame3 fr;
    if (!cpu_bucket(h, view, w, hh, threads, &fr, st)) return -1;
    std::vector<uint32_t> order(fr.num_tiles);
    for (uint32_t t = 0; t < fr.num_tiles; ++t) order[t] = t;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return h->tile_start[a + 1] - h->tile_start[a] > h->tile_start[b + 1] - h->tile_start[b];
    });
    std::vector<uint64_t> works(fr.num_tiles, 0);
    parallel_for(fr.num_tiles, threads, [&](size_t oi) {
        const uint32_t t = order[oi];
        const size_t b = h->tile_start[t], e = h->tile_start[t + 1];
        sort_hi32(h->keys.data() + b, h->sort_tmp.data() + b, e - b);
        fix_ties(h->keys.data() + b, e - b, h->orig.data());
        const uint32_t tx = t % fr.tiles_x, ty = t / fr.tiles_x;
        raster_tile(h->splats.data(), h->keys.data() + b, e - b, tx * TILE, ty * TILE, w, hh, bg, out, &works[t]);
    });
    if (work)
        for (uint64_t v : works) *work += v;
    return 0;
}

// Per-tile blend statistics of CpuTiled: lens[t] = pairs of tile t,
// processed[t] = pairs blended before every pixel of the tile saturated.
int or_cpu_tile_stats(or_cpu *h, const sp_view *view, uint32_t w, uint32_t hh, uint32_t *lens, uint32_t *processed,
                      int threads) {
    if (w == 0 || hh == 0) return 0;
    threads = resolve_threads(threads);
    Frame3 fr;
    sp_stats st;
    if (!cpu_bucket(h, view, w, hh, threads, &fr, &st)) return -1;
    const float bg[3] = {0.f, 0.f, 0.f};
    parallel_for(fr.num_tiles, threads, [&](size_t t) {
        const size_t b = h->tile_start[t], e = h->tile_start[t + 1];
        sort_hi32(h->keys.data() + b, h->sort_tmp.data() + b, e - b);
        fix_ties(h->keys.data() + b, e - b, h->orig.data());
        const uint32_t tx = (uint32_t)t % fr.tiles_x, ty = (uint32_t)t / fr.tiles_x;
        size_t done = 0;
        raster_tile(h->splats.data(), h->keys.data() + b, e - b, tx * TILE, ty * TILE, w, hh, bg, nullptr, nullptr,
                    &done);
        lens[t] = (uint32_t)(e - b);
        processed[t] = (uint32_t)done;
    });
    return 0;
}

// Canonical sorted pairs of CpuTiled for `view`: keys = tile<<32 | depth bits,
// vals = original scene index, tile_start[0..tiles]. Returns 0, -1 on overflow
// of 2^32 pairs, -5 when cap is too small (*n_pairs still set).
int or_cpu_pairs(or_cpu *h, const sp_view *view, uint32_t w, uint32_t hh, uint64_t *keys, uint32_t *vals,
                 uint32_t *tile_start, uint64_t cap, uint64_t *n_pairs, int threads) {
    *n_pairs = 0;
    if (w == 0 || hh == 0) return 0;
    threads = resolve_threads(threads);
    Frame3 fr;
    sp_stats st;
    if (!cpu_bucket(h, view, w, hh, threads, &fr, &st)) return -1;
    *n_pairs = st.overlaps;
    for (uint32_t t = 0; t <= fr.num_tiles; ++t) tile_start[t] = h->tile_start[t];
    if (st.overlaps > cap) return -5;
    parallel_for(fr.num_tiles, threads, [&](size_t t) {
        const size_t b = h->tile_start[t], e = h->tile_start[t + 1];
        sort_hi32(h->keys.data() + b, h->sort_tmp.data() + b, e - b);
        fix_ties(h->keys.data() + b, e - b, h->orig.data());
        for (size_t p = b; p < e; ++p) {
            keys[p] = ((uint64_t)t << 32) | (h->keys[p] >> 32);
            vals[p] = h->orig[(uint32_t)h->keys[p]];
        }
    });
    return 0;
}

// CpuNaive::render (src/render/naive.rs:29-100).
int or_naive_render(const sp_scene *scene, const sp_view *view, uint32_t w, uint32_t hh, const float bg[3],
                    uint8_t *out, sp_stats *st) {
    sp_stats local;
    if (!st) st = &local;
    std::memset(st, 0, sizeof *st);
    if (w == 0 || hh == 0) return 0;
    const SceneRef s = scene_ref(scene);
    const Ctx ctx = make_ctx(*view, w, hh);
    std::vector<Splat2D> splats;
    for (size_t i = 0; i < s.n; ++i) {
        Splat2D p;
        if (project(s, i, ctx, k_of(s.opac[i]), &p)) splats.push_back(p);
    }
    st->visible = splats.size();
    std::stable_sort(splats.begin(), splats.end(),
                     [](const Splat2D &a, const Splat2D &b) { return f2u(a.depth) < f2u(b.depth); });
    const size_t np = (size_t)w * hh;
    std::vector<float> color(3 * np, 0.0f), trans(np, 1.0f);
    for (const Splat2D &sp : splats) {
        const uint32_t x0 = sat_u32(std::fmax(std::floor(sp.x - sp.ext_x), 0.0f));
        const uint32_t y0 = sat_u32(std::fmax(std::floor(sp.y - sp.ext_y), 0.0f));
        const uint32_t x1 = std::min(sat_u32(std::fmax(std::ceil(sp.x + sp.ext_x), 0.0f)), w);
        const uint32_t y1 = std::min(sat_u32(std::fmax(std::ceil(sp.y + sp.ext_y), 0.0f)), hh);
        for (uint32_t py = y0; py < y1; ++py) {
            const float dy = (float)py + 0.5f - sp.y;
            for (uint32_t px = x0; px < x1; ++px) {
                const size_t i = (size_t)py * w + px;
                const float t = trans[i];
                if (t < T_MIN) continue;
                const float dx = (float)px + 0.5f - sp.x;
                // render::power (src/render/mod.rs:276-278).
                const float pw = -0.5f * (sp.conic[0] * dx * dx + sp.conic[2] * dy * dy) - sp.conic[1] * dx * dy;
                if (pw > 0.0f) continue;
                const float alpha = std::fmin(sp.opacity * std::exp(pw), ALPHA_MAX);
                if (alpha < ALPHA_MIN) continue;
                const float wgt = alpha * t;
                for (int c = 0; c < 3; ++c) color[3 * i + c] += sp.color[c] * wgt;
                trans[i] = t * (1.0f - alpha);
            }
        }
    }
    st->overlaps = splats.size();
    for (size_t i = 0; i < np; ++i) {
        const float t = trans[i];
        for (int c = 0; c < 3; ++c) out[3 * i + c] = q8(color[3 * i + c] + t * bg[c]);
    }
    return 0;
}

// render::project for every splat (k computed as project does). valid[i],
// proj[i*12..] = x, y, conic[3], color[3], opacity, depth, ext_x, ext_y; bbox[i*4..].
void or_project(const sp_scene *scene, const sp_view *view, uint32_t w, uint32_t hh, uint8_t *valid, float *proj,
                uint16_t *bbox) {
    const SceneRef s = scene_ref(scene);
    const Ctx ctx = make_ctx(*view, w, hh);
    for (size_t i = 0; i < s.n; ++i) {
        Splat2D p;
        valid[i] = project(s, i, ctx, k_of(s.opac[i]), &p) ? 1 : 0;
        float *o = proj + 12 * i;
        uint16_t *bb = bbox + 4 * i;
        if (!valid[i]) {
            std::memset(o, 0, 48);
            std::memset(bb, 0, 8);
            continue;
        }
        const float v[12] = {p.x, p.y, p.conic[0], p.conic[1], p.conic[2], p.color[0], p.color[1], p.color[2],
                             p.opacity, p.depth, p.ext_x, p.ext_y};
        std::memcpy(o, v, 48);
        pixel_bbox(p, w, hh, bb);
    }
}

// Single-function entry points for known-answer tests.
float or_exp_neg(float x) { return exp_neg(x); }
uint8_t or_to_u8(float v) { return q8(v); }
void or_sort_hi32(uint64_t *keys, uint64_t *tmp, uint64_t n) { sort_hi32(keys, tmp, (size_t)n); }
void or_fix_ties(uint64_t *keys, uint64_t n, const uint32_t *orig) { fix_ties(keys, (size_t)n, orig); }
void or_eval_sh(uint32_t degree, const float *rest, uint32_t nrest, const float dir[3], float out[3]) {
    eval_rest(degree, rest, nrest, dir, out);
}
void or_bounds(const float *pos, uint64_t n, float center[3], float *radius) { bounds(pos, (size_t)n, center, radius); }
uint32_t or_spread3(uint32_t v) { return spread3(v); }
void or_orig_order(const or_cpu *h, uint32_t *orig) {
    for (size_t j = 0; j < h->n; ++j) orig[j] = h->orig[j];
}

}  // extern "C"

// ---- FP64 reference (the precision variant's error yardstick) -------------
//
// The same mathematics as render::project + sh::eval_rest + the cpu-naive
// blend (src/render/mod.rs:201-278, src/sh.rs:39-64, src/render/naive.rs:45-93),
// evaluated in double from the f32 scene parameters: the covariance from
// scale/quat (src/scene.rs:170-178) when they are given, exp() instead of the
// polynomial. Culling and the blend's thresholds are the reference's, so every
// difference from it is arithmetic error of the path under test.

namespace {

struct P64 {
    double x, y, conic[3], color[3], opacity, depth, ext_x, ext_y;
    uint32_t bbox[4];
};

bool project64(const sp_scene *s, size_t i, const sp_view &v, uint32_t w, uint32_t h, P64 *o) {
    const double op = s->opacities[i];
    if (op < (double)ALPHA_MIN) return false;
    double cov[6];
    if (s->scales && s->rotations) {
        const double sx = s->scales[3 * i], sy = s->scales[3 * i + 1], sz = s->scales[3 * i + 2];
        const double qx = s->rotations[4 * i], qy = s->rotations[4 * i + 1], qz = s->rotations[4 * i + 2],
                     qw = s->rotations[4 * i + 3];
        const double R[3][3] = {{1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qw * qz), 2 * (qx * qz + qw * qy)},
                                {2 * (qx * qy + qw * qz), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qw * qx)},
                                {2 * (qx * qz - qw * qy), 2 * (qy * qz + qw * qx), 1 - 2 * (qx * qx + qy * qy)}};
        const double sc[3] = {sx, sy, sz};
        double S[3][3];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                double acc = 0;
                for (int k = 0; k < 3; ++k) acc += R[a][k] * sc[k] * sc[k] * R[b][k];
                S[a][b] = acc;
            }
        cov[0] = S[0][0];
        cov[1] = S[0][1];
        cov[2] = S[0][2];
        cov[3] = S[1][1];
        cov[4] = S[1][2];
        cov[5] = S[2][2];
    } else {
        for (int k = 0; k < 6; ++k) cov[k] = s->cov3d[6 * i + k];
    }
    const double d[3] = {s->positions[3 * i] - (double)v.eye[0], s->positions[3 * i + 1] - (double)v.eye[1],
                         s->positions[3 * i + 2] - (double)v.eye[2]};
    const double rot[3][3] = {{v.right[0], v.right[1], v.right[2]},
                              {v.down[0], v.down[1], v.down[2]},
                              {v.forward[0], v.forward[1], v.forward[2]}};
    double t[3];
    for (int r = 0; r < 3; ++r) t[r] = rot[r][0] * d[0] + rot[r][1] * d[1] + rot[r][2] * d[2];
    if (!(t[2] > (double)v.near_z)) return false;
    const double focal = (double)h * 0.5 / std::tan(0.5 * (double)v.fov_y);
    const double x = focal * t[0] / t[2] + 0.5 * w, y = focal * t[1] / t[2] + 0.5 * h;
    const double lx = 1.3 * (0.5 * w / focal), ly = 1.3 * (0.5 * h / focal);
    const double tx = std::min(std::max(t[0] / t[2], -lx), lx), ty = std::min(std::max(t[1] / t[2], -ly), ly);
    const double f = focal / t[2];
    const double j0[3] = {f, 0, -f * tx}, j1[3] = {0, f, -f * ty};
    double T0[3], T1[3];
    for (int c = 0; c < 3; ++c) {
        T0[c] = rot[0][c] * j0[0] + rot[1][c] * j0[1] + rot[2][c] * j0[2];
        T1[c] = rot[0][c] * j1[0] + rot[1][c] * j1[1] + rot[2][c] * j1[2];
    }
    auto sig = [&](const double u[3], double r[3]) {
        r[0] = cov[0] * u[0] + cov[1] * u[1] + cov[2] * u[2];
        r[1] = cov[1] * u[0] + cov[3] * u[1] + cov[4] * u[2];
        r[2] = cov[2] * u[0] + cov[4] * u[1] + cov[5] * u[2];
    };
    double s0[3], s1[3];
    sig(T0, s0);
    sig(T1, s1);
    const double a = T0[0] * s0[0] + T0[1] * s0[1] + T0[2] * s0[2] + (double)DILATION;
    const double b = T1[0] * s0[0] + T1[1] * s0[1] + T1[2] * s0[2];
    const double cc = T1[0] * s1[0] + T1[1] * s1[1] + T1[2] * s1[2] + (double)DILATION;
    const double det = a * cc - b * b;
    if (!(det > 0)) return false;
    const double k = std::sqrt(2.0 * std::log(op / (double)ALPHA_MIN));
    const double ex = k * std::sqrt(a), ey = k * std::sqrt(cc);
    if (x + ex <= 0 || x - ex >= w || y + ey <= 0 || y - ey >= h) return false;
    double col[3] = {s->colors[3 * i], s->colors[3 * i + 1], s->colors[3 * i + 2]};
    if (s->sh_degree > 0) {
        const double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const double X = d[0] / len, Y = d[1] / len, Z = d[2] / len;
        const uint32_t rest = (s->sh_degree + 1) * (s->sh_degree + 1) - 1;
        const float *r = s->sh_rest + (size_t)3 * rest * i;
        const double XX = X * X, YY = Y * Y, ZZ = Z * Z;
        const double basis[15] = {-0.4886025119029199 * Y, 0.4886025119029199 * Z, -0.4886025119029199 * X,
                                  1.0925484305920792 * X * Y, -1.0925484305920792 * Y * Z,
                                  0.31539156525252005 * (2 * ZZ - XX - YY), -1.0925484305920792 * X * Z,
                                  0.5462742152960396 * (XX - YY), -0.5900435899266435 * Y * (3 * XX - YY),
                                  2.890611442640554 * X * Y * Z, -0.4570457994644658 * Y * (4 * ZZ - XX - YY),
                                  0.3731763325901154 * Z * (2 * ZZ - 3 * XX - 3 * YY),
                                  -0.4570457994644658 * X * (4 * ZZ - XX - YY), 1.445305721320277 * Z * (XX - YY),
                                  -0.5900435899266435 * X * (XX - 3 * YY)};
        for (uint32_t q = 0; q < rest; ++q)
            for (int ch = 0; ch < 3; ++ch) col[ch] += basis[q] * (double)r[3 * q + ch];
    }
    o->x = x;
    o->y = y;
    o->conic[0] = cc / det;
    o->conic[1] = -b / det;
    o->conic[2] = a / det;
    for (int c = 0; c < 3; ++c) o->color[c] = std::max(col[c], 0.0);
    o->opacity = op;
    o->depth = t[2];
    o->ext_x = ex;
    o->ext_y = ey;
    const double fx0 = std::max(std::floor(x - ex), 0.0), fy0 = std::max(std::floor(y - ey), 0.0);
    const uint32_t x1 = (uint32_t)std::min(std::max(std::ceil(x + ex), 0.0), (double)w);
    const uint32_t y1 = (uint32_t)std::min(std::max(std::ceil(y + ey), 0.0), (double)h);
    o->bbox[0] = std::min((uint32_t)std::min(fx0, 4.0e9), x1);
    o->bbox[1] = std::min((uint32_t)std::min(fy0, 4.0e9), y1);
    o->bbox[2] = x1;
    o->bbox[3] = y1;
    return true;
}

}  // namespace

extern "C" {

// FP64 projection of every splat: valid[i], proj[i*12..] = x, y, conic a,b,c, rgb,
// opacity, depth, ext_x, ext_y (doubles).
void or_project_f64(const sp_scene *s, const sp_view *v, uint32_t w, uint32_t h, uint8_t *valid, double *proj,
                    int threads) {
    threads = resolve_threads(threads);
    const size_t n = (size_t)s->n;
    const size_t chunk = 4096;
    parallel_for((n + chunk - 1) / chunk, threads, [&](size_t c) {
        for (size_t i = c * chunk; i < std::min(n, (c + 1) * chunk); ++i) {
            P64 p;
            valid[i] = project64(s, i, *v, w, h, &p) ? 1 : 0;
            double *o = proj + 12 * i;
            if (!valid[i]) {
                for (int k = 0; k < 12; ++k) o[k] = 0;
                continue;
            }
            const double g[12] = {p.x, p.y, p.conic[0], p.conic[1], p.conic[2], p.color[0], p.color[1],
                                  p.color[2], p.opacity, p.depth, p.ext_x, p.ext_y};
            for (int k = 0; k < 12; ++k) o[k] = g[k];
        }
    });
}

// FP64 render with the cpu-naive semantics: rgb (h*w*3 doubles, before
// quantisation, background composited) and the quantised RGB8 image.
void or_render_f64(const sp_scene *s, const sp_view *v, uint32_t w, uint32_t h, const float bg[3], double *rgb,
                   uint8_t *out, int threads) {
    threads = resolve_threads(threads);
    const size_t n = (size_t)s->n;
    std::vector<P64> ps;
    std::vector<uint8_t> ok(n);
    std::vector<P64> all(n);
    const size_t chunk = 4096;
    parallel_for((n + chunk - 1) / chunk, threads, [&](size_t c) {
        for (size_t i = c * chunk; i < std::min(n, (c + 1) * chunk); ++i) ok[i] = project64(s, i, *v, w, h, &all[i]);
    });
    std::vector<uint32_t> idx;
    for (size_t i = 0; i < n; ++i)
        if (ok[i]) idx.push_back((uint32_t)i);
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return all[a].depth < all[b].depth; });
    const size_t np = (size_t)w * h;
    std::vector<double> C(3 * np, 0.0), T(np, 1.0);
    const uint32_t band = 8;
    const uint32_t nbands = (h + band - 1) / band;
    parallel_for(nbands, threads, [&](size_t bi) {
        const uint32_t r0 = (uint32_t)bi * band, r1 = std::min(h, r0 + band);
        for (uint32_t id : idx) {
            const P64 &p = all[id];
            const uint32_t y0 = std::max(p.bbox[1], r0), y1 = std::min(p.bbox[3], r1);
            for (uint32_t py = y0; py < y1; ++py) {
                const double dy = py + 0.5 - p.y;
                for (uint32_t px = p.bbox[0]; px < p.bbox[2]; ++px) {
                    const size_t q = (size_t)py * w + px;
                    const double t = T[q];
                    if (t < (double)T_MIN) continue;
                    const double dx = px + 0.5 - p.x;
                    const double pw = -0.5 * (p.conic[0] * dx * dx + p.conic[2] * dy * dy) - p.conic[1] * dx * dy;
                    if (pw > 0) continue;
                    const double al = std::min(p.opacity * std::exp(pw), (double)ALPHA_MAX);
                    if (al < (double)ALPHA_MIN) continue;
                    for (int c = 0; c < 3; ++c) C[3 * q + c] += p.color[c] * al * t;
                    T[q] = t * (1.0 - al);
                }
            }
        }
    });
    for (size_t q = 0; q < np; ++q)
        for (int c = 0; c < 3; ++c) {
            const double val = C[3 * q + c] + T[q] * (double)bg[c];
            if (rgb) rgb[3 * q + c] = val;
            if (out) {
                const double cl = std::min(std::max(val, 0.0), 1.0);
                out[3 * q + c] = (uint8_t)std::min(255.0, std::floor(cl * 255.0 + 0.5));
            }
        }
}

}  // extern "C"
