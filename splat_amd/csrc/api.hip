// C ABI of libsplat_hip.so and the per-view orchestration.
//
// One handle owns the device copy of a scene plus every per-view buffer, all
// reused across frames (as CpuTiled reuses its scratch, src/render/cpu.rs:68-79).
// A view is one stream-ordered sequence of kernels with a single host
// synchronisation at the end:
//   project (+compact) -> depth radix sort -> emit pairs -> tile radix sort
//   -> tile ranges -> blend
// Every data-dependent size (visible count, pair count) stays on the device;
// grids are sized from capacities and kernels read the live counts. When the
// pair count exceeds the capacity the frame is re-run with a larger buffer
// (the reference's gpu backend grows the same way, src/render/gpu.rs:1039-1056).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "renderer.hpp"
#include "sp_common.h"
#include "sp_launch.h"
#include "splat_hip.h"

using sp::ViewCtx;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return fail(SP_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Prep's k (src/render/cpu.rs:384-388), computed once per scene on the host so
// that the device never evaluates a logarithm that could differ in the last bit.
float k_of(float o) { return o < sp::ALPHA_MIN ? 0.0f : std::sqrt(2.0f * std::log(o / sp::ALPHA_MIN)); }

// ProjectCtx::new (src/render/mod.rs:168-197) + View::focal/rotation (src/camera.rs:174-181).
ViewCtx make_ctx(const sp_view &v, uint32_t w, uint32_t h, const float bg[3]) {
    ViewCtx c;
    std::memset(&c, 0, sizeof c);
    c.focal = (float)h * 0.5f / std::tan(v.fov_y * 0.5f);
    const float fw = (float)w, fh = (float)h;
    auto lim = [&](float half) {
        const float l = 1.3f * (half / c.focal);
        return std::isfinite(l) ? l : 0.0f;
    };
    for (int k = 0; k < 3; ++k) {
        c.rot[0][k] = v.right[k];
        c.rot[1][k] = v.down[k];
        c.rot[2][k] = v.forward[k];
        c.eye[k] = v.eye[k];
        c.bg[k] = bg[k];
    }
    c.cx = fw * 0.5f;
    c.cy = fh * 0.5f;
    c.width = fw;
    c.height = fh;
    c.near_z = v.near_z;
    c.lim_x = lim(fw * 0.5f);
    c.lim_y = lim(fh * 0.5f);
    c.w = w;
    c.h = h;
    c.tiles_x = (w + sp::TILE - 1) / sp::TILE;
    c.tiles_y = (h + sp::TILE - 1) / sp::TILE;
    return c;
}

uint32_t bits_for(uint32_t max_value) { return max_value == 0 ? 0u : 32u - (uint32_t)__builtin_clz(max_value); }

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct sp_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    sp_options opt{};
    uint32_t n = 0, deg = 0, rest = 0;
    uint64_t stride = 0;
    uint64_t bytes = 0;
    float *d_soa = nullptr;
    uint4 *d_rec = nullptr;
    uint32_t *d_dkey[2] = {nullptr, nullptr}, *d_dval[2] = {nullptr, nullptr};
    uint64_t pair_cap = 0;
    uint32_t *d_pkey[2] = {nullptr, nullptr}, *d_pval[2] = {nullptr, nullptr};
    uint32_t *d_tile_start = nullptr;
    uint64_t tile_cap = 0;
    uint8_t *d_rgb = nullptr;
    uint64_t rgb_cap = 0;
    uint32_t *d_counters = nullptr;
    unsigned long long *d_status = nullptr;
    uint64_t status_cap = 0;
    uint32_t *d_hist = nullptr;
    uint32_t *h_counters = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    sp_frame_info last{};
};

namespace {

int dalloc(sp_handle *h, void **p, uint64_t bytes) {
    *p = nullptr;
    if (h->opt.max_device_bytes && h->bytes + bytes > h->opt.max_device_bytes)
        return fail(SP_E_NOMEM, "device memory budget exceeded (" + std::to_string(h->bytes + bytes) + " > " +
                                    std::to_string(h->opt.max_device_bytes) + " bytes)");
    const hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(SP_E_NOMEM, std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    h->bytes += bytes;
    return SP_OK;
}

template <class T>
void dfree(sp_handle *h, T *&p, uint64_t bytes) {
    if (p) {
        (void)hipFree(p);
        h->bytes -= bytes;
    }
    p = nullptr;
}

uint64_t status_words(uint64_t n, uint64_t pair_cap) {
    const uint64_t scan1 = (n + 255) / 256 + 1;
    const uint64_t m = std::max(n, pair_cap);
    const uint64_t radix = ((m + sp::RADIX_TILE - 1) / sp::RADIX_TILE + 1) * sp::RADIX_BINS;
    return std::max(scan1, radix);
}

int ensure_status(sp_handle *h) {
    const uint64_t need = status_words(h->n, h->pair_cap);
    if (need <= h->status_cap) return SP_OK;
    dfree(h, h->d_status, h->status_cap * 8);
    h->status_cap = 0;
    int rc = dalloc(h, (void **)&h->d_status, need * 8);
    if (rc) return rc;
    // Zero words read as "not published" in every epoch. The clear is ordered on
    // the handle's stream (kernels run there) and completed before any launch:
    // recycled device memory may hold words of another buffer or handle.
    HIP_TRY(hipMemsetAsync(h->d_status, 0, need * 8, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->status_cap = need;
    return SP_OK;
}

int ensure_pairs(sp_handle *h, uint64_t cap) {
    if (cap <= h->pair_cap) return SP_OK;
    if (cap > 0xFFFFFFFFull) return fail(SP_E_CAPACITY, "more than 2^32 gaussian/tile pairs");
    for (int b = 0; b < 2; ++b) {
        dfree(h, h->d_pkey[b], h->pair_cap * 4);
        dfree(h, h->d_pval[b], h->pair_cap * 4);
    }
    h->pair_cap = 0;
    for (int b = 0; b < 2; ++b) {
        int rc = dalloc(h, (void **)&h->d_pkey[b], cap * 4);
        if (!rc) rc = dalloc(h, (void **)&h->d_pval[b], cap * 4);
        if (rc) {
            for (int c = 0; c < 2; ++c) {
                dfree(h, h->d_pkey[c], cap * 4);
                dfree(h, h->d_pval[c], cap * 4);
            }
            return rc;
        }
    }
    h->pair_cap = cap;
    return ensure_status(h);
}

int ensure_tiles(sp_handle *h, uint64_t tiles) {
    if (tiles + 1 <= h->tile_cap) return SP_OK;
    dfree(h, h->d_tile_start, h->tile_cap * 4);
    h->tile_cap = 0;
    const int rc = dalloc(h, (void **)&h->d_tile_start, (tiles + 1) * 4);
    if (rc) return rc;
    h->tile_cap = tiles + 1;
    return SP_OK;
}

int ensure_rgb(sp_handle *h, uint64_t bytes) {
    if (bytes <= h->rgb_cap) return SP_OK;
    dfree(h, h->d_rgb, h->rgb_cap);
    h->rgb_cap = 0;
    const int rc = dalloc(h, (void **)&h->d_rgb, bytes);
    if (rc) return rc;
    h->rgb_cap = bytes;
    return SP_OK;
}

sp::DevScene dev_scene(const sp_handle *h) {
    sp::DevScene s;
    s.soa = h->d_soa;
    s.stride = h->stride;
    s.n = h->n;
    s.deg = h->deg;
    s.rest = h->rest;
    s.pad = 0;
    return s;
}

int check_frame(uint32_t w, uint32_t hgt) {
    if (w > SP_MAX_DIM || hgt > SP_MAX_DIM)
        return fail(SP_E_PIXEL_LIMIT, "frame " + std::to_string(w) + "x" + std::to_string(hgt) + " exceeds the " +
                                          std::to_string(SP_MAX_DIM) + " pixel limit");
    return SP_OK;
}

// Run one view. d_out == nullptr stops after the tile ranges (inspection).
// On return *pair_buf names the ping-pong buffer that holds the sorted pairs.
int run_view(sp_handle *h, const sp_view &view, uint32_t w, uint32_t hgt, const float bg[3], uint8_t *d_out,
             sp_stats *st, int *pair_buf) {
    const double t0 = now_ms();
    const ViewCtx ctx = make_ctx(view, w, hgt, bg);
    const uint32_t num_tiles = ctx.tiles_x * ctx.tiles_y;
    int rc = ensure_tiles(h, num_tiles);
    if (rc) return rc;
    const sp::DevScene sc = dev_scene(h);
    const int tile_bits = (int)bits_for(num_tiles - 1);
    for (int attempt = 0;; ++attempt) {
        // Built per attempt: growing the pair buffers reallocates the status array.
        const sp::Lookback lb{h->d_status, h->d_counters};
        if (h->opt.timing) HIP_TRY(hipEventRecord(h->ev[0], h->stream));
        HIP_TRY(sp::launch_project(h->stream, sc, ctx, h->d_rec, h->d_dkey[0], h->d_dval[0], lb, nullptr, nullptr,
                                   nullptr));
        int di = 0;
        HIP_TRY(sp::launch_radix_sort(h->stream, h->d_dkey, h->d_dval, h->d_counters + sp::CNT_VISIBLE, h->n, 32, lb,
                                      h->d_hist, &di));
        HIP_TRY(sp::launch_emit(h->stream, h->d_dval[di], h->n, h->d_rec, ctx, h->d_pkey[0], h->d_pval[0],
                                h->pair_cap, lb));
        if (h->opt.timing) HIP_TRY(hipEventRecord(h->ev[1], h->stream));
        int pi = 0;
        HIP_TRY(sp::launch_radix_sort(h->stream, h->d_pkey, h->d_pval, h->d_counters + sp::CNT_PAIRS, h->pair_cap,
                                      tile_bits, lb, h->d_hist, &pi));
        HIP_TRY(sp::launch_ranges(h->stream, h->d_pkey[pi], h->d_counters, h->pair_cap, num_tiles, h->d_tile_start));
        if (h->opt.timing) HIP_TRY(hipEventRecord(h->ev[2], h->stream));
        if (d_out)
            HIP_TRY(sp::launch_blend(h->stream, ctx, h->d_tile_start, h->d_pval[pi], h->d_rec, d_out,
                                     h->d_counters));
        if (h->opt.timing) HIP_TRY(hipEventRecord(h->ev[3], h->stream));
        HIP_TRY(hipMemcpyAsync(h->h_counters, h->d_counters, sp::CNT_SLOTS * 4, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        *pair_buf = pi;
        if (h->h_counters[sp::CNT_OVERFLOW] && attempt < 3) {
            const uint64_t p = h->h_counters[sp::CNT_PAIRS];
            rc = ensure_pairs(h, p + p / 4 + 1024);
            if (rc) return rc;
            continue;
        }
        if (h->h_counters[sp::CNT_OVERFLOW]) return fail(SP_E_CAPACITY, "pair buffer overflow persisted");
        break;
    }
    if (st) {
        st->visible = h->h_counters[sp::CNT_VISIBLE];
        st->overlaps = h->h_counters[sp::CNT_PAIRS];
        if (h->opt.timing) {
            float a = 0, b = 0, c = 0;
            HIP_TRY(hipEventElapsedTime(&a, h->ev[0], h->ev[1]));
            HIP_TRY(hipEventElapsedTime(&b, h->ev[1], h->ev[2]));
            HIP_TRY(hipEventElapsedTime(&c, h->ev[2], h->ev[3]));
            st->project_ms = a;
            st->sort_ms = b;
            st->raster_ms = c;
        }
        st->total_ms = now_ms() - t0;
    }
    h->last.width = w;
    h->last.height = hgt;
    h->last.tiles = num_tiles;
    h->last.visible = h->h_counters[sp::CNT_VISIBLE];
    h->last.pairs = h->h_counters[sp::CNT_PAIRS];
    h->last.pairs_staged = d_out ? h->h_counters[sp::CNT_STAGED] : 0;
    h->last.pair_capacity = h->pair_cap;
    h->last.device_bytes = h->bytes;
    return SP_OK;
}

int setup(sp_handle *h) {
    HIP_TRY(hipSetDevice(h->device));
    return SP_OK;
}

}  // namespace

extern "C" {

const char *sp_last_error(void) { return g_err.c_str(); }

void sp_default_options(sp_options *opt) {
    std::memset(opt, 0, sizeof *opt);
    opt->device = 0;
    opt->storage = SP_STORE_F32;
    opt->timing = 1;
}

const char *sp_name(const sp_handle *) { return "hip"; }

void sp_destroy(sp_handle *h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    (void)hipFree(h->d_soa);
    (void)hipFree(h->d_rec);
    for (int b = 0; b < 2; ++b) {
        (void)hipFree(h->d_dkey[b]);
        (void)hipFree(h->d_dval[b]);
        (void)hipFree(h->d_pkey[b]);
        (void)hipFree(h->d_pval[b]);
    }
    (void)hipFree(h->d_tile_start);
    (void)hipFree(h->d_rgb);
    (void)hipFree(h->d_counters);
    (void)hipFree(h->d_status);
    (void)hipFree(h->d_hist);
    if (h->h_counters) (void)hipHostFree(h->h_counters);
    for (auto &e : h->ev)
        if (e) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

int sp_create(const sp_scene *scene, const sp_options *opt_in, sp_handle **out) {
    if (!out) return fail(SP_E_INVALID, "sp_create: out is NULL");
    *out = nullptr;
    if (!scene) return fail(SP_E_INVALID, "sp_create: scene is NULL");
    sp_options opt;
    if (opt_in)
        opt = *opt_in;
    else
        sp_default_options(&opt);
    if (scene->sh_degree > 3) return fail(SP_E_INVALID, "sh_degree must be 0..3");
    if (scene->n > 0x7FFFFFFFull) return fail(SP_E_INVALID, "more than 2^31 gaussians");
    const uint32_t n = (uint32_t)scene->n;
    const uint32_t rest = (scene->sh_degree + 1) * (scene->sh_degree + 1) - 1;
    if (n && (!scene->positions || !scene->cov3d || !scene->opacities || !scene->colors))
        return fail(SP_E_INVALID, "scene arrays missing");
    if (n && rest && !scene->sh_rest) return fail(SP_E_INVALID, "sh_rest missing for sh_degree > 0");
    if (opt.storage != SP_STORE_F32) return fail(SP_E_UNSUPPORTED, "storage precision not supported yet");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(SP_E_DEVICE, "no HIP device");
    if (opt.device < 0 || opt.device >= count) return fail(SP_E_DEVICE, "device ordinal out of range");

    sp_handle *h = new sp_handle();
    h->device = opt.device;
    h->opt = opt;
    h->n = n;
    h->deg = scene->sh_degree;
    h->rest = rest;
    auto bail = [&](int rc) {
        sp_destroy(h);
        return rc;
    };
    int rc = setup(h);
    if (rc) return bail(rc);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(SP_E_DEVICE, "hipStreamCreate failed"));
    for (auto &e : h->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail(fail(SP_E_DEVICE, "hipEventCreate failed"));

    // Structure-of-arrays layout (one 64-aligned array per field).
    h->stride = ((uint64_t)n + 63) / 64 * 64;
    if (h->stride == 0) h->stride = 64;
    const uint64_t ncols = (uint64_t)sp::A_SH + 3ull * rest;
    {
        std::vector<float> soa(ncols * h->stride, 0.0f);
        const uint64_t S = h->stride;
        for (uint32_t i = 0; i < n; ++i) {
            for (int c = 0; c < 3; ++c) soa[(sp::A_POS + c) * S + i] = scene->positions[3ull * i + c];
            for (int c = 0; c < 6; ++c) soa[(sp::A_COV + c) * S + i] = scene->cov3d[6ull * i + c];
            const float o = scene->opacities[i];
            soa[sp::A_OPAC * S + i] = o;
            soa[sp::A_K * S + i] = k_of(o);
            for (int c = 0; c < 3; ++c) soa[(sp::A_COL + c) * S + i] = scene->colors[3ull * i + c];
            for (uint32_t k = 0; k < 3 * rest; ++k)
                soa[(sp::A_SH + k) * S + i] = scene->sh_rest[(uint64_t)3 * rest * i + k];
        }
        rc = dalloc(h, (void **)&h->d_soa, soa.size() * 4);
        if (rc) return bail(rc);
        if (hipMemcpyAsync(h->d_soa, soa.data(), soa.size() * 4, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess)
            return bail(fail(SP_E_DEVICE, "scene upload failed"));
    }
    const uint64_t nn = std::max<uint64_t>(n, 1);
    rc = dalloc(h, (void **)&h->d_rec, nn * 48);
    for (int b = 0; b < 2 && !rc; ++b) {
        rc = dalloc(h, (void **)&h->d_dkey[b], nn * 4);
        if (!rc) rc = dalloc(h, (void **)&h->d_dval[b], nn * 4);
    }
    if (!rc) rc = dalloc(h, (void **)&h->d_counters, sp::CNT_SLOTS * 4);
    if (!rc) rc = dalloc(h, (void **)&h->d_hist, 4 * sp::RADIX_BINS * 4);
    if (rc) return bail(rc);
    // Each handle starts its lookback epochs at a distinct base, so status words
    // left in recycled memory by another handle never carry a live epoch tag.
    static std::atomic<uint32_t> g_handles{0};
    uint32_t init[sp::CNT_SLOTS] = {0};
    init[sp::CNT_EPOCH] = ((g_handles.fetch_add(1) * 0x9E3779B1u) >> 2 | 1u) & 0x3FFFFFFFu;
    if (hipMemcpyAsync(h->d_counters, init, sizeof init, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipMemsetAsync(h->d_hist, 0, 4 * sp::RADIX_BINS * 4, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return bail(fail(SP_E_DEVICE, "counter init failed"));
    if (hipHostMalloc((void **)&h->h_counters, sp::CNT_SLOTS * 4, hipHostMallocDefault) != hipSuccess)
        return bail(fail(SP_E_DEVICE, "hipHostMalloc failed"));
    const uint64_t cap0 = opt.initial_pairs ? opt.initial_pairs : std::max<uint64_t>(16ull * n, 1ull << 20);
    rc = ensure_pairs(h, cap0);
    if (!rc) rc = ensure_status(h);
    if (rc) return bail(rc);
    *out = h;
    return SP_OK;
}

int sp_render_device(sp_handle *h, const sp_view *view, uint32_t width, uint32_t height, const float background[3],
                     uint8_t *rgb_out_device, sp_stats *stats) {
    if (!h || !view || !background) return fail(SP_E_INVALID, "sp_render_device: NULL argument");
    if (stats) std::memset(stats, 0, sizeof *stats);
    int rc = check_frame(width, height);
    if (rc) return rc;
    if (width == 0 || height == 0) return SP_OK;
    if (!rgb_out_device) return fail(SP_E_INVALID, "sp_render_device: NULL output");
    rc = setup(h);
    if (rc) return rc;
    int pb = 0;
    return run_view(h, *view, width, height, background, rgb_out_device, stats, &pb);
}

int sp_render(sp_handle *h, const sp_view *view, uint32_t width, uint32_t height, const float background[3],
              uint8_t *rgb_out, sp_stats *stats) {
    if (!h || !view || !background) return fail(SP_E_INVALID, "sp_render: NULL argument");
    if (stats) std::memset(stats, 0, sizeof *stats);
    int rc = check_frame(width, height);
    if (rc) return rc;
    if (width == 0 || height == 0) return SP_OK;
    if (!rgb_out) return fail(SP_E_INVALID, "sp_render: NULL output");
    rc = setup(h);
    if (rc) return rc;
    const uint64_t bytes = 3ull * width * height;
    rc = ensure_rgb(h, bytes);
    if (rc) return rc;
    const double t0 = now_ms();
    int pb = 0;
    rc = run_view(h, *view, width, height, background, h->d_rgb, stats, &pb);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(rgb_out, h->d_rgb, bytes, hipMemcpyDeviceToHost));
    if (stats) stats->total_ms = now_ms() - t0;
    return SP_OK;
}

int sp_debug_project(sp_handle *h, const sp_view *view, uint32_t width, uint32_t height, uint8_t *valid,
                     float *proj, uint16_t *bbox) {
    if (!h || !view || !valid || !proj || !bbox) return fail(SP_E_INVALID, "sp_debug_project: NULL argument");
    int rc = check_frame(width, height);
    if (rc) return rc;
    rc = setup(h);
    if (rc) return rc;
    const uint64_t n = h->n;
    if (n == 0) return SP_OK;
    const float bg[3] = {0.f, 0.f, 0.f};
    const ViewCtx ctx = make_ctx(*view, width, height, bg);
    float *d_dbg = nullptr;
    uint16_t *d_bb = nullptr;
    uint8_t *d_valid = nullptr;
    HIP_TRY(hipMalloc(&d_dbg, n * 48));
    HIP_TRY(hipMalloc(&d_bb, n * 8));
    HIP_TRY(hipMalloc(&d_valid, n));
    HIP_TRY(hipMemsetAsync(d_dbg, 0, n * 48, h->stream));
    HIP_TRY(hipMemsetAsync(d_bb, 0, n * 8, h->stream));
    const sp::Lookback lb{h->d_status, h->d_counters};
    hipError_t e = sp::launch_project(h->stream, dev_scene(h), ctx, h->d_rec, h->d_dkey[0], h->d_dval[0], lb, d_dbg,
                                      d_bb, d_valid);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(proj, d_dbg, n * 48, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(bbox, d_bb, n * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(valid, d_valid, n, hipMemcpyDeviceToHost);
    (void)hipFree(d_dbg);
    (void)hipFree(d_bb);
    (void)hipFree(d_valid);
    if (e != hipSuccess) return fail(SP_E_DEVICE, std::string("sp_debug_project: ") + hipGetErrorString(e));
    return SP_OK;
}

int sp_debug_pairs(sp_handle *h, const sp_view *view, uint32_t width, uint32_t height, uint64_t *keys,
                   uint32_t *vals, uint32_t *tile_start, uint64_t cap, uint64_t *n_pairs) {
    if (!h || !view || !n_pairs || !tile_start) return fail(SP_E_INVALID, "sp_debug_pairs: NULL argument");
    *n_pairs = 0;
    int rc = check_frame(width, height);
    if (rc) return rc;
    if (width == 0 || height == 0) return SP_OK;
    rc = setup(h);
    if (rc) return rc;
    const float bg[3] = {0.f, 0.f, 0.f};
    sp_stats st;
    int pb = 0;
    rc = run_view(h, *view, width, height, bg, nullptr, &st, &pb);
    if (rc) return rc;
    const uint64_t np = st.overlaps;
    *n_pairs = np;
    const uint32_t tiles = ((width + 15) / 16) * ((height + 15) / 16);
    HIP_TRY(hipMemcpy(tile_start, h->d_tile_start, (tiles + 1ull) * 4, hipMemcpyDeviceToHost));
    if (np > cap) return fail(SP_E_CAPACITY, "sp_debug_pairs: capacity too small");
    if (np == 0) return SP_OK;
    if (!keys || !vals) return fail(SP_E_INVALID, "sp_debug_pairs: NULL keys/vals");
    std::vector<uint32_t> tk(np);
    std::vector<uint4> rec((uint64_t)h->n * 3);
    HIP_TRY(hipMemcpy(tk.data(), h->d_pkey[pb], np * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(vals, h->d_pval[pb], np * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(rec.data(), h->d_rec, rec.size() * 16, hipMemcpyDeviceToHost));
    for (uint64_t p = 0; p < np; ++p) keys[p] = ((uint64_t)tk[p] << 32) | rec[3ull * vals[p] + 2].w;
    return SP_OK;
}

int sp_last_frame_info(const sp_handle *h, sp_frame_info *info) {
    if (!h || !info) return fail(SP_E_INVALID, "sp_last_frame_info: NULL argument");
    *info = h->last;
    return SP_OK;
}

int sp_render_batch(sp_handle *h, const sp_view *views, const uint32_t *wh, uint32_t n_views,
                    const float background[3], uint8_t *const *outs, int32_t outs_on_device, sp_stats *stats) {
    if (!h || (!views && n_views) || (!wh && n_views) || !background || (!outs && n_views))
        return fail(SP_E_INVALID, "sp_render_batch: NULL argument");
    for (uint32_t v = 0; v < n_views; ++v) {
        const int rc = outs_on_device
                           ? sp_render_device(h, &views[v], wh[2 * v], wh[2 * v + 1], background, outs[v],
                                              stats ? &stats[v] : nullptr)
                           : sp_render(h, &views[v], wh[2 * v], wh[2 * v + 1], background, outs[v],
                                       stats ? &stats[v] : nullptr);
        if (rc) return rc;
    }
    return SP_OK;
}

}  // extern "C"

// ---- C++ Renderer layer ----------------------------------------------------

namespace splat {

void Frame::resize(uint32_t w, uint32_t h) {
    if (w > SP_MAX_DIM || h > SP_MAX_DIM)
        throw std::runtime_error("frame " + std::to_string(w) + "x" + std::to_string(h) + " exceeds the " +
                                 std::to_string(SP_MAX_DIM) + " pixel limit");
    width = w;
    height = h;
    pixels.resize(3ull * w * h);
}

HipRenderer::HipRenderer(const sp_scene &scene, const sp_options *opt) {
    if (sp_create(&scene, opt, &h_) != SP_OK) throw std::runtime_error(sp_last_error());
}

HipRenderer::~HipRenderer() { sp_destroy(h_); }

RenderStats HipRenderer::render(const View &view, uint32_t width, uint32_t height, const float background[3],
                                Frame &frame) {
    frame.resize(width, height);
    RenderStats st;
    std::memset(&st, 0, sizeof st);
    if (width == 0 || height == 0) return st;
    if (sp_render(h_, &view, width, height, background, frame.pixels.data(), &st) != SP_OK)
        throw std::runtime_error(sp_last_error());
    return st;
}

std::unique_ptr<Renderer> create(Backend backend, const sp_scene &scene) {
    switch (backend) {
    case Backend::Hip: return std::unique_ptr<Renderer>(new HipRenderer(scene));
    }
    throw std::runtime_error("unknown backend");
}

}  // namespace splat
