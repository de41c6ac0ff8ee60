// C ABI of libsplat_hip.so and the orchestration of a batch of views.
//
// One handle owns the device copy of a scene plus every per-view buffer, all
// reused across frames (as CpuTiled reuses its scratch, src/render/cpu.rs:68-79).
// A batch of views (one view for Renderer::render) is one stream-ordered
// sequence of launches, each covering every view:
//   project (+ per-view compaction) -> depth radix sort (one segment per view)
//   -> emit pairs (all views, one global tile numbering) -> tile radix sort
//   -> tile ranges -> blend
// with a single host synchronisation at the end. Every data-dependent size
// (visible counts, pair count) stays on the device; grids are sized from
// capacities and kernels read the live counts, and no kernel waits on
// another block, so a batch can be captured once into a HIP graph and
// replayed. When the pairs exceed the capacity the batch is re-run with a
// larger buffer (the reference's gpu backend grows the same way,
// src/render/gpu.rs:1039-1056).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "half_conv.h"
#include "renderer.hpp"
#include "sp_common.h"
#include "sp_launch.h"
#include "sp_reach.h"
#include "splat_hip.h"

using sp::ViewCtx;
using sp::ViewParams;

// Per-frame scratch of sp_render_device_static: one region per GPU, reserved
// when the library's code object loads on that device and held until the
// process ends. Host code must not use this name as an address (it is a host
// stand-in); it uses the per-device pointer from hipGetSymbolAddress.
static __device__ __align__(256) uint8_t scratch[SP_STATIC_SCRATCH_BYTES];

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

// The depth-sort key base of a view (ViewCtx::key_base): the near plane's
// bits when every visible depth of the scene in this view lies below the
// near plane's bits plus 2^DEPTH_KEY_BITS, else 0 (full 32-bit keys). A
// visible splat has near_z < depth <= |forward| * |p - eye| (up to the f32
// rounding of the projection, covered by a relative margin of 2^-16), and
// |p - eye| <= |c - eye| + r over the bounding sphere (c, r) of the scene's
// finite positions.
uint32_t depth_key_base(const sp_handle *h, const ViewCtx &c);

// What a batch's blend stage produces.
constexpr int BL_NONE = 0;  // stop after the tile ranges
constexpr int BL_RGB = 1;   // RGB8 images (Renderer::render)
constexpr int BL_DIFF = 2;  // float images plus the state of the adjoint

uint32_t bits_for(uint32_t max_value) { return max_value == 0 ? 0u : 32u - (uint32_t)__builtin_clz(max_value); }

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// What a captured batch graph depends on besides the contents of the
// (device-resident) view table.
struct GraphKey {
    uint32_t nviews = 0, total_tiles = 0, tile_bits = 0, with_blend = 0;
    uint32_t split_shift = 0, split_kmin = 0, blend_batch = 0, total_sbs = 0;
    uint32_t depth_bits = 0;
    uint64_t generation = 0;
    bool operator==(const GraphKey &o) const {
        return nviews == o.nviews && total_tiles == o.total_tiles && tile_bits == o.tile_bits &&
               depth_bits == o.depth_bits &&
               total_sbs == o.total_sbs &&
               with_blend == o.with_blend && split_shift == o.split_shift && split_kmin == o.split_kmin &&
               blend_batch == o.blend_batch && generation == o.generation;
    }
};

// A captured pipeline and the shape it was captured for.
struct GraphEntry {
    GraphKey key{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int pair_buf = 0;
    uint64_t used = 0;  // last use (handle clock), for least-recently-used eviction
};
// Shapes kept captured per handle (callers alternating a few resolutions
// replay their graphs instead of capturing again).
constexpr int GRAPH_CACHE = 8;

}  // namespace

struct sp_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    sp_options opt{};
    uint32_t n = 0, deg = 0, rest = 0;
    uint64_t stride = 0;
    uint64_t bytes = 0;
    uint64_t generation = 1;  // bumped whenever a device buffer moves
    float *d_soa = nullptr;
    uint16_t *d_soa16 = nullptr;
    void *d_crow = nullptr;  // colour rows (f32 or 16-bit entries, sp::colour_row per splat)
    uint64_t soa_bytes = 0, soa16_bytes = 0, crow_bytes = 0;
    // The scene may give a splat an infinite colour (launch_prep's flag): its
    // frames blend through the instance that reproduces the reference's
    // NaN channels beside such splats.
    bool extreme_colour = false;
    // Bounding sphere of the scene's finite positions (found at upload; not
    // for differentiable handles, whose positions come with each call): it
    // bounds each view's depths, so the depth sort can use DEPTH_KEY_BITS-bit
    // keys. depth_full is set for good if a frame ever found a key beyond it.
    bool bound_ok = false, depth_full = false;
    double bound_c[3] = {0.0, 0.0, 0.0}, bound_r = 0.0;
    int depth_bits = 32;  // key width of the last batch's depth sort
    // Per-view arrays, view_cap views of n entries each.
    uint32_t view_cap = 0;
    uint4 *d_rec = nullptr;
    uint2 *d_box = nullptr;  // depth-ordered pixel boxes, view_cap x n
    uint2 *d_sbox = nullptr; // scene-order pixel boxes written by the projection, view_cap x n
    uint32_t *d_dkey[2] = {nullptr, nullptr}, *d_dval[2] = {nullptr, nullptr};
    // One allocation: CNT_SLOTS frame counters (d_counters), then 5 x view_cap
    // view stats (d_vstat): visible, pairs, emission view bases, superblock-pair
    // bases, unfinished tiles after the first phase. The counters, visible
    // counts and pair counts are thus contiguous and come back in one copy.
    uint32_t *d_statblk = nullptr;
    uint32_t *d_vstat = nullptr;
    uint32_t *d_bcnt = nullptr;   // view_cap x block counts (projection, emission)
    uint32_t *d_bcnt_all = nullptr;  // view_cap x block counts of the unsplit emission
    uint32_t *d_bsb = nullptr;    // view_cap x block counts of superblock pairs (binned first phase)
    ViewParams *d_views = nullptr, *h_views = nullptr;
    uint32_t *d_rowtot = nullptr;  // radix digit-row totals: view_cap x RADIX_BINS_MAX digits
    uint32_t *d_table = nullptr;  // radix per-block digit counts
    uint64_t table_words = 0;
    // Pairs of all views.
    uint64_t pair_cap = 0;
    uint32_t *d_pkey[2] = {nullptr, nullptr}, *d_pval[2] = {nullptr, nullptr};
    uint32_t *d_tile_start = nullptr;
    uint64_t tile_cap = 0;
    // Binned first phase: superblock list bounds, per-tile counts and row totals.
    uint32_t *d_sb_start = nullptr, *d_tcount = nullptr, *d_trow = nullptr;
    uint32_t *d_segc = nullptr;  // (sb_cap - 1) x BIN_SEGS x 16 segment counts
    uint32_t *d_sbtile = nullptr;  // (sb_cap - 1) x 16 global tile ids
    uint64_t sb_cap = 0, tcount_cap = 0;
    uint8_t *d_rgb = nullptr;
    uint64_t rgb_cap = 0;
    uint32_t *d_counters = nullptr;
    uint32_t *h_counters = nullptr;  // pinned copy of d_statblk's first CNT_SLOTS + 2 * view_cap words
    hipEvent_t ev[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    // sp_render's read-back: a pinned staging buffer, filled in chunks whose
    // host copies overlap the next chunk's transfer.
    uint8_t *h_rgb_pin = nullptr;
    uint64_t rgb_pin_cap = 0;
    hipEvent_t ev_rb[4] = {nullptr, nullptr, nullptr, nullptr};
    sp_frame_info last{};
    int pair_buf = 0;
    int depth_buf = 0;  // ping-pong buffer holding the last depth sort's output
    int blend_batch = 64;  // records staged through LDS per blend step (64, 128 or 256)
    // Front-to-back split of the plain path: the first phase emits each view's
    // nearest max(split_kmin, vis >> split_shift) splats (split_shift 0: no split).
    uint32_t split_shift = 3, split_kmin = 16384;
    uint32_t *d_done = nullptr;  // split_cap tiles: finished by the first phase
    uint32_t *d_openm = nullptr;  // split_cap words: per-row bits of the tiles the first blend left open
    float *d_pstate = nullptr;   // split_cap * 256 pixels: (T, r, g, b) between the phases
    uint64_t split_cap = 0;
    // Differentiable path (sp_diff_create): raw parameters on the caller's
    // device memory, per-pair owners, per-pixel blend state, gradient scratch.
    bool diff = false;
    sp::DevParams prm{};
    // Emission slots of the differentiable handles' tile pairs (sp::PairSlots).
    uint32_t *d_pbase = nullptr;  // view_cap x n: first emission slot of each record
    uint32_t *d_pslot = nullptr;  // view_cap x (n + 1): first slot of each depth-ordered splat, views' end slots
    uint32_t *d_prid = nullptr;   // view_cap x n: record of each depth-ordered splat
    float *d_pixT = nullptr;       // pix_cap: final transmittance per pixel (256 per tile)
    uint32_t *d_pixlast = nullptr; // pix_cap: tile-list position after the last contribution
    uint64_t pix_cap = 0;
    float *d_gpair = nullptr;  // NG2D x gpair_cap per-pair gradients
    uint32_t *d_fresh = nullptr;  // (gpair_cap + 31) / 32 words: the pairs the running backward wrote
    uint64_t gpair_cap = 0;
    float *d_g2d = nullptr;  // g2d_cap x NG2D per-record gradients, record-major
    uint64_t g2d_cap = 0;
    uint32_t fwd_views = 0, fwd_tiles = 0, fwd_sbs = 0;  // shape of the last differentiable forward
    uint64_t fwd_pairs = 0;
    uint32_t last_sbs = 0;  // superblocks of the last rendered batch
    // Buffers carved from a fixed region (sp_render_device_static): nothing may
    // grow or be freed; a frame that does not fit is SP_E_CAPACITY.
    bool fixed = false;
    // Captured pipelines, by shape.
    GraphEntry graphs[GRAPH_CACHE];
    uint64_t graph_clock = 0;
};

namespace {

uint32_t depth_key_base(const sp_handle *h, const ViewCtx &c) {
    if (!h->bound_ok || h->depth_full || !(c.near_z > 0.0f) || !std::isfinite(c.near_z)) return 0u;
    double d2 = 0.0, f2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        d2 += ((double)c.eye[k] - h->bound_c[k]) * ((double)c.eye[k] - h->bound_c[k]);
        f2 += (double)c.rot[2][k] * (double)c.rot[2][k];
    }
    const double far = std::sqrt(f2) * (std::sqrt(d2) + h->bound_r) * (1.0 + 1.0 / 65536.0);
    if (!(far < 3.0e38)) return 0u;
    uint32_t near_bits, far_bits;
    const float nf = c.near_z, ff = std::nextafter((float)far, INFINITY);
    std::memcpy(&near_bits, &nf, 4);
    std::memcpy(&far_bits, &ff, 4);
    if (far_bits <= near_bits) return near_bits;  // every splat lies at or before the near plane
    return far_bits - near_bits < (1u << sp::DEPTH_KEY_BITS) ? near_bits : 0u;
}

int dalloc(sp_handle *h, void **p, uint64_t bytes) {
    *p = nullptr;
    if (h->opt.max_device_bytes && h->bytes + bytes > h->opt.max_device_bytes)
        return fail(SP_E_NOMEM, "device memory budget exceeded (" + std::to_string(h->bytes + bytes) + " > " +
                                    std::to_string(h->opt.max_device_bytes) + " bytes)");
    const hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(SP_E_NOMEM, std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    h->bytes += bytes;
    h->generation++;
    return SP_OK;
}

template <class T>
void dfree(sp_handle *h, T *&p, uint64_t bytes) {
    if (p) {
        (void)hipFree(p);
        h->bytes -= bytes;
        h->generation++;
    }
    p = nullptr;
}

// A fixed handle's buffers cannot grow.
int fixed_full(const char *what) {
    return fail(SP_E_CAPACITY, std::string("the frame's ") + what + " exceed the static scratch of sp_render_device_static (" +
                                   std::to_string(SP_STATIC_SCRATCH_BYTES >> 20) + " MiB)");
}

void drop_graph(GraphEntry &g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g = GraphEntry{};
}

void drop_graphs(sp_handle *h) {
    for (GraphEntry &g : h->graphs) drop_graph(g);
}

// The cached graph for `key`, or the entry to capture it into: an empty one,
// one captured before a buffer moved (it can never match again), or the least
// recently used.
GraphEntry &graph_slot(sp_handle *h, const GraphKey &key, bool *hit) {
    *hit = false;
    for (GraphEntry &g : h->graphs)
        if (g.exec && g.key == key) {
            *hit = true;
            return g;
        }
    GraphEntry *victim = &h->graphs[0];
    for (GraphEntry &g : h->graphs) {
        if (!g.exec || g.key.generation != h->generation) return g;
        if (g.used < victim->used) victim = &g;
    }
    return *victim;
}

// Radix count table large enough for the depth sort of view_cap views and the
// tile sort of pair_cap pairs.
// The table of the depth sort (one segment per view) and of a pair sort, the
// two sizes a fixed handle's single-view sorts use.
uint64_t table_need_single(uint64_t n, uint64_t pairs) {
    return std::max(sp::radix_table_words(1, std::max<uint64_t>(n, 1)), sp::radix_table_words(1, std::max<uint64_t>(pairs, 1)));
}

int ensure_table(sp_handle *h) {
    const uint32_t vc = std::max<uint32_t>(h->view_cap, 1);
    if (h->fixed) return table_need_single(h->n, h->pair_cap) <= h->table_words ? SP_OK : fixed_full("radix tables");
    const uint64_t need = std::max({sp::radix_table_words(vc, std::max<uint32_t>(h->n, 1)),
                                    sp::radix_table_words(1, std::max<uint64_t>(h->pair_cap, 1)),
                                    sp::radix_table_bound(vc)});
    if (need <= h->table_words) return SP_OK;
    if (h->opt.max_device_bytes && h->bytes - h->table_words * 4 + need * 4 > h->opt.max_device_bytes)
        return fail(SP_E_NOMEM, "device memory budget exceeded growing the radix table");
    dfree(h, h->d_table, h->table_words * 4);
    h->table_words = 0;
    const int rc = dalloc(h, (void **)&h->d_table, need * 4);
    if (rc) return rc;
    h->table_words = need;
    return SP_OK;
}

int ensure_pairs(sp_handle *h, uint64_t cap) {
    if (cap <= h->pair_cap) return SP_OK;
    if (h->fixed) return fixed_full("tile pairs");
    if (cap > 0xFFFFFFFFull) return fail(SP_E_CAPACITY, "more than 2^32 gaussian/tile pairs");
    // The budget is checked before anything is freed: a refused growth leaves
    // the handle's buffers (and its ability to render smaller frames) intact.
    const uint64_t per_pair = 4ull * 4;
    if (h->opt.max_device_bytes && h->bytes - h->pair_cap * per_pair + cap * per_pair > h->opt.max_device_bytes)
        return fail(SP_E_NOMEM, "device memory budget exceeded growing the pair buffers to " + std::to_string(cap) +
                                    " pairs");
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
    return ensure_table(h);
}

// Pair capacity of a new handle. The split plain path materialises only the
// first phase's pairs at once.
uint64_t initial_pairs(const sp_handle *h) {
    return h->opt.initial_pairs ? h->opt.initial_pairs
                                : std::max<uint64_t>((h->diff ? 16ull : 4ull) * h->n, 1ull << 20);
}

// Returns the tile-pair buffers to a new handle's capacity, so later frames
// grow them as a new handle's would. The pair sorts and range searches size
// their grids and radix tiles from the capacity, so a capacity left far above
// the need (after a frame that sorted every pair of an unsplit pipeline) slows
// every later frame of a smaller need.
int reset_pairs(sp_handle *h) {
    if (h->fixed) return SP_OK;
    for (int b = 0; b < 2; ++b) {
        dfree(h, h->d_pkey[b], h->pair_cap * 4);
        dfree(h, h->d_pval[b], h->pair_cap * 4);
    }
    h->pair_cap = 0;
    return ensure_pairs(h, initial_pairs(h));
}

// Per-view arrays for `views` views.
int ensure_views(sp_handle *h, uint32_t views) {
    if (views <= h->view_cap) return SP_OK;
    if (h->fixed) return fixed_full("views");
    const uint64_t nn = std::max<uint64_t>(h->n, 1);
    const uint64_t old = h->view_cap;
    const uint64_t blocks = std::max(sp::project_blocks(h->n), sp::emit_blocks(h->n));
    dfree(h, h->d_rec, old * nn * 48);
    dfree(h, h->d_box, old * nn * 8);
    dfree(h, h->d_sbox, old * nn * 8);
    if (h->diff) {
        dfree(h, h->d_pbase, old * nn * 4);
        dfree(h, h->d_pslot, old * (nn + 1) * 4);
        dfree(h, h->d_prid, old * nn * 4);
    }
    for (int b = 0; b < 2; ++b) {
        dfree(h, h->d_dkey[b], old * nn * 4);
        dfree(h, h->d_dval[b], old * nn * 4);
    }
    dfree(h, h->d_statblk, sp::CNT_SLOTS * 4 + old * 20);
    h->d_vstat = h->d_counters = nullptr;
    dfree(h, h->d_bcnt, old * blocks * 4);
    dfree(h, h->d_bcnt_all, old * blocks * 4);
    dfree(h, h->d_bsb, old * blocks * 4);
    dfree(h, h->d_views, old * sizeof(ViewParams));
    dfree(h, h->d_rowtot, old * sp::RADIX_BINS_MAX * 4);
    if (h->h_views) (void)hipHostFree(h->h_views);
    if (h->h_counters) (void)hipHostFree(h->h_counters);
    h->h_views = nullptr;
    h->h_counters = nullptr;
    h->view_cap = 0;
    const uint64_t v = views;
    int rc = dalloc(h, (void **)&h->d_rec, v * nn * 48);
    if (!rc) rc = dalloc(h, (void **)&h->d_box, v * nn * 8);
    if (!rc) rc = dalloc(h, (void **)&h->d_sbox, v * nn * 8);
    if (!rc && h->diff) rc = dalloc(h, (void **)&h->d_pbase, v * nn * 4);
    if (!rc && h->diff) rc = dalloc(h, (void **)&h->d_pslot, v * (nn + 1) * 4);
    if (!rc && h->diff) rc = dalloc(h, (void **)&h->d_prid, v * nn * 4);
    for (int b = 0; b < 2 && !rc; ++b) {
        rc = dalloc(h, (void **)&h->d_dkey[b], v * nn * 4);
        if (!rc) rc = dalloc(h, (void **)&h->d_dval[b], v * nn * 4);
    }
    if (!rc) rc = dalloc(h, (void **)&h->d_statblk, sp::CNT_SLOTS * 4 + v * 20);
    if (!rc) rc = dalloc(h, (void **)&h->d_bcnt, v * blocks * 4);
    if (!rc) rc = dalloc(h, (void **)&h->d_bcnt_all, v * blocks * 4);
    if (!rc) rc = dalloc(h, (void **)&h->d_bsb, v * blocks * 4);
    if (!rc) rc = dalloc(h, (void **)&h->d_views, v * sizeof(ViewParams));
    if (!rc) rc = dalloc(h, (void **)&h->d_rowtot, v * sp::RADIX_BINS_MAX * 4);
    if (rc) return rc;
    if (hipHostMalloc((void **)&h->h_views, v * sizeof(ViewParams), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&h->h_counters, (sp::CNT_SLOTS + 2 * v) * 4, hipHostMallocDefault) != hipSuccess)
        return fail(SP_E_NOMEM, "hipHostMalloc failed");
    // The frame counters start at 0.
    h->d_counters = h->d_statblk;
    h->d_vstat = h->d_statblk + sp::CNT_SLOTS;
    const uint32_t zero[sp::CNT_SLOTS] = {0};
    if (hipMemcpyAsync(h->d_counters, zero, sizeof zero, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return fail(SP_E_DEVICE, "counter init failed");
    h->view_cap = views;
    return ensure_table(h);
}

int ensure_tiles(sp_handle *h, uint64_t tiles) {
    if (tiles + 1 <= h->tile_cap) return SP_OK;
    if (h->fixed) return fixed_full("tiles");
    dfree(h, h->d_tile_start, h->tile_cap * 4);
    h->tile_cap = 0;
    const int rc = dalloc(h, (void **)&h->d_tile_start, (tiles + 1) * 4);
    if (rc) return rc;
    h->tile_cap = tiles + 1;
    return SP_OK;
}

// Superblock bounds and tile counts of the binned first phase.
int ensure_bin(sp_handle *h, uint64_t tiles, uint64_t sbs) {
    if (h->fixed && (sbs + 1 > h->sb_cap || sp::bin_scan_rows((uint32_t)tiles) * sp::BIN_SCAN_ROW > h->tcount_cap))
        return fixed_full("superblocks");
    if (sbs + 1 > h->sb_cap) {
        const uint64_t seg_words = (uint64_t)sp::BIN_SEGS * sp::SB * sp::SB;
        dfree(h, h->d_sb_start, h->sb_cap * 4);
        dfree(h, h->d_segc, (h->sb_cap ? h->sb_cap - 1 : 0) * seg_words * 4);
        dfree(h, h->d_sbtile, (h->sb_cap ? h->sb_cap - 1 : 0) * sp::SB * sp::SB * 4);
        h->sb_cap = 0;
        int rc = dalloc(h, (void **)&h->d_sb_start, (sbs + 1) * 4);
        if (!rc) rc = dalloc(h, (void **)&h->d_segc, sbs * seg_words * 4);
        if (!rc) rc = dalloc(h, (void **)&h->d_sbtile, sbs * sp::SB * sp::SB * 4);
        if (rc) return rc;
        h->sb_cap = sbs + 1;
    }
    const uint64_t rows = sp::bin_scan_rows((uint32_t)tiles);
    if (rows * sp::BIN_SCAN_ROW > h->tcount_cap) {
        const uint64_t old = h->tcount_cap;
        dfree(h, h->d_tcount, old * 4);
        dfree(h, h->d_trow, old / sp::BIN_SCAN_ROW * 4);
        h->tcount_cap = 0;
        int rc = dalloc(h, (void **)&h->d_tcount, rows * sp::BIN_SCAN_ROW * 4);
        if (!rc) rc = dalloc(h, (void **)&h->d_trow, rows * 4);
        if (rc) return rc;
        h->tcount_cap = rows * sp::BIN_SCAN_ROW;
    }
    return SP_OK;
}

// Per-pixel blend state of the differentiable forward (256 entries per tile).
int ensure_pix(sp_handle *h, uint64_t tiles) {
    const uint64_t need = tiles * sp::BLEND_THREADS;
    if (need <= h->pix_cap) return SP_OK;
    if (h->fixed) return fixed_full("pixels");
    dfree(h, h->d_pixT, h->pix_cap * 4);
    dfree(h, h->d_pixlast, h->pix_cap * 4);
    h->pix_cap = 0;
    int rc = dalloc(h, (void **)&h->d_pixT, need * 4);
    if (!rc) rc = dalloc(h, (void **)&h->d_pixlast, need * 4);
    if (rc) return rc;
    h->pix_cap = need;
    return SP_OK;
}

// Tile flags, row prefixes and pixel state of the front-to-back split.
int ensure_split(sp_handle *h, uint64_t tiles) {
    if (tiles <= h->split_cap) return SP_OK;
    if (h->fixed) return fixed_full("tile states");
    dfree(h, h->d_done, h->split_cap * 4);
    dfree(h, h->d_openm, h->split_cap * 4);
    dfree(h, h->d_pstate, h->split_cap * sp::BLEND_THREADS * 16);
    h->split_cap = 0;
    int rc = dalloc(h, (void **)&h->d_done, tiles * 4);
    if (!rc) rc = dalloc(h, (void **)&h->d_openm, tiles * 4);
    if (!rc) rc = dalloc(h, (void **)&h->d_pstate, tiles * sp::BLEND_THREADS * 16);
    if (rc) return rc;
    h->split_cap = tiles;
    return SP_OK;
}

// Gradient scratch of the backward: per-pair, per-record and carry arrays.
int ensure_grad(sp_handle *h, uint64_t pairs, uint64_t records) {
    if (h->fixed) return fixed_full("gradients");
    int rc = SP_OK;
    if (pairs > h->gpair_cap) {
        dfree(h, h->d_gpair, h->gpair_cap * sp::NG2D * 4);
        dfree(h, h->d_fresh, (h->gpair_cap + 31) / 32 * 4);
        h->gpair_cap = 0;
        rc = dalloc(h, (void **)&h->d_gpair, pairs * sp::NG2D * 4);
        if (!rc) rc = dalloc(h, (void **)&h->d_fresh, (pairs + 31) / 32 * 4);
        if (rc) return rc;
        h->gpair_cap = pairs;
    }
    if (records > h->g2d_cap) {
        dfree(h, h->d_g2d, h->g2d_cap * sp::NG2D * 4);
        h->g2d_cap = 0;
        rc = dalloc(h, (void **)&h->d_g2d, records * sp::NG2D * 4);
        if (rc) return rc;
        h->g2d_cap = records;
    }
    return SP_OK;
}

int ensure_rgb(sp_handle *h, uint64_t bytes) {
    if (bytes <= h->rgb_cap) return SP_OK;
    if (h->fixed) return fixed_full("frame bytes");
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
    s.soa16 = h->d_soa16;
    const bool half = h->opt.storage != SP_STORE_F32;
    s.crow = half ? nullptr : static_cast<const float *>(h->d_crow);
    s.crow16 = half ? static_cast<const uint16_t *>(h->d_crow) : nullptr;
    s.stride = h->stride;
    s.n = h->n;
    s.deg = h->deg;
    s.rest = h->rest;
    s.storage = h->diff ? (uint32_t)sp::STORE_PARAMS : (uint32_t)h->opt.storage;
    s.prm = h->prm;
    return s;
}

int check_frame(uint32_t w, uint32_t hgt) {
    if (w > SP_MAX_DIM || hgt > SP_MAX_DIM)
        return fail(SP_E_PIXEL_LIMIT, "frame " + std::to_string(w) + "x" + std::to_string(hgt) + " exceeds the " +
                                          std::to_string(SP_MAX_DIM) + " pixel limit");
    return SP_OK;
}

int setup(sp_handle *h) {
    HIP_TRY(hipSetDevice(h->device));
    return SP_OK;
}

// Enqueue the whole pipeline for the first `nviews` entries of h_views. The
// tile lists of the (last) phase end in ping-pong buffer *pair_buf. The first
// phase is binned (superblock pairs sorted, tile lists written per
// superblock); the differentiable path also numbers its tile pairs in
// emission order (sp::PairSlots) for the adjoint.
// BL_RGB with a split runs the emission, tile sort, ranges and blend twice:
// the first phase for each view's nearest splats over all tiles, the second
// for the rest over the tiles the first blend left unfinished.
// `after_first` (inspection only, never captured) runs after the first blend
// with the first phase's lists in d_pval[first list buffer].
int enqueue(sp_handle *h, uint32_t nviews, uint32_t total_tiles, uint32_t total_sbs, int tile_bits, int mode,
            bool events, int *pair_buf, const std::function<int(int)> *after_first = nullptr) {
    const sp::DevScene sc = dev_scene(h);
    hipStream_t s = h->stream;
    uint32_t *vis = h->d_vstat;
    uint32_t *view_pairs = h->d_vstat + h->view_cap;
    uint32_t *vtot = h->d_vstat + 2 * h->view_cap;
    const bool split = mode == BL_RGB && h->split_shift > 0;
    const sp::EmitSplit es = split ? sp::EmitSplit{h->split_kmin, h->split_shift} : sp::EmitSplit{0u, 0u};
    sp::PairSlots slots{h->d_pbase, h->d_pslot, h->d_prid};
    const sp::PairSlots *pslots = mode == BL_DIFF ? &slots : nullptr;
    // A single view's stored-scene colours are evaluated by the emission, for
    // the splats a blend reads; a batch's projection, bound by its record
    // writes, evaluates them itself, as does the differentiable projection. A
    // frame that stops before the blend needs none.
    const bool defer = nviews <= sp::DEFER_COLOUR_VIEWS && sc.storage != sp::STORE_PARAMS;
    uint4 *colour_rec = (defer && mode != BL_NONE) ? h->d_rec : nullptr;
    auto mark = [&](int k) { return events ? hipEventRecord(h->ev[k], s) : hipSuccess; };
    HIP_TRY(hipMemcpyAsync(h->d_views, h->h_views, nviews * sizeof(ViewParams), hipMemcpyHostToDevice, s));
    HIP_TRY(mark(0));
    // Depth keys of every (view, splat), FULL_CULLED for the culled ones: the
    // depth sort's sparse first pass compacts them (scene-index order within a
    // digit, so the sort stays stable) and writes each view's visible count.
    HIP_TRY(sp::launch_project(s, sc, h->d_views, nviews, h->d_rec, h->d_sbox, h->d_dkey[0], nullptr, nullptr,
                               nullptr, h->d_counters, !defer));
    if (h->n == 0) HIP_TRY(sp::launch_fill(s, vis, nviews, 0u));  // no sort runs to count them
    int di = 0;
    HIP_TRY(sp::launch_radix_sort(s, h->d_dkey, h->d_dval, vis, nviews, h->n, h->n, h->depth_bits, h->d_rowtot,
                                  h->d_table, &di, true));
    h->depth_buf = di;
    uint32_t *vsb = h->d_vstat + 3 * h->view_cap;
    uint32_t *vopen = h->d_vstat + 4 * h->view_cap;
    HIP_TRY(sp::launch_emit(s, sp::EMIT_FIRST, h->d_dval[di], h->n, nviews, vis, h->d_sbox, h->d_box, h->d_views,
                            es, split ? h->d_openm : nullptr, split ? vopen : nullptr, h->d_bcnt, h->d_bcnt_all,
                            vtot, view_pairs,
                            h->d_bsb, vsb, h->d_pkey[0], h->d_pval[0], pslots,
                            h->pair_cap, h->d_counters, sc, colour_rec));
    HIP_TRY(mark(1));
    int pi = 0;
    {
        int si = 0;
        const int sb_bits = (int)bits_for(total_sbs ? total_sbs - 1 : 0);
        HIP_TRY(sp::launch_radix_sort(s, h->d_pkey, h->d_pval, h->d_counters + sp::CNT_SBPAIRS, 1, 0, h->pair_cap,
                                      sb_bits, h->d_rowtot, h->d_table, &si));
        HIP_TRY(sp::launch_ranges(s, h->d_pkey[si], (1u << sp::SB_SUB_SHIFT) - 1u, h->d_counters + sp::CNT_SBPAIRS,
                                  h->pair_cap, total_sbs, h->d_sb_start));
        pi = 1 - si;
        // Tile ids are kept only when the lists are read back (no blend).
        HIP_TRY(sp::launch_bin_lists(s, h->d_views, nviews, total_sbs, total_tiles, h->d_sb_start, h->d_pkey[si],
                                     h->d_pval[si], h->d_segc, h->d_sbtile, h->d_tcount, h->d_trow,
                                     h->d_tile_start,
                                     mode == BL_NONE ? h->d_pkey[pi] : nullptr, h->d_pval[pi], h->pair_cap,
                                     h->d_counters));
    }
    HIP_TRY(mark(2));
    if (mode != BL_NONE) {
        const int bm = split ? sp::BM_FIRST : (mode == BL_DIFF ? sp::BM_DIFF : sp::BM_RGB);
        HIP_TRY(sp::launch_blend(s, bm, h->d_views, nviews, total_tiles, total_sbs, h->d_tile_start, h->d_pval[pi],
                                 h->d_rec, h->d_counters, h->d_pixT, h->d_pixlast, h->d_done, h->d_pstate,
                                 split ? h->d_openm : nullptr, split ? vopen : nullptr, h->blend_batch,
                                 h->extreme_colour));
    }
    HIP_TRY(mark(3));
    if (after_first) {
        const int rc = (*after_first)(pi);
        if (rc) return rc;
    }
    if (split) {
        HIP_TRY(sp::launch_emit(s, sp::EMIT_SECOND, h->d_dval[di], h->n, nviews, vis, h->d_sbox, h->d_box,
                                h->d_views, es, h->d_openm, vopen, h->d_bcnt, h->d_bcnt_all, vtot, view_pairs, nullptr,
                                nullptr, h->d_pkey[0], h->d_pval[0], nullptr, h->pair_cap, h->d_counters, sc,
                                colour_rec));
    }
    HIP_TRY(mark(4));
    if (split) {
        HIP_TRY(sp::launch_radix_sort(s, h->d_pkey, h->d_pval, h->d_counters + sp::CNT_PAIRS2, 1, 0, h->pair_cap,
                                      tile_bits, h->d_rowtot, h->d_table, &pi));
        HIP_TRY(sp::launch_ranges(s, h->d_pkey[pi], 0xFFFFFFFFu, h->d_counters + sp::CNT_PAIRS2, h->pair_cap,
                                  total_tiles, h->d_tile_start));
    }
    HIP_TRY(mark(5));
    if (split)
        HIP_TRY(sp::launch_blend(s, sp::BM_SECOND, h->d_views, nviews, total_tiles, total_sbs, h->d_tile_start,
                                 h->d_pval[pi],
                                 h->d_rec, h->d_counters, nullptr, nullptr, h->d_done, h->d_pstate, nullptr,
                                 nullptr, h->blend_batch, h->extreme_colour));
    HIP_TRY(mark(6));
    // Counters, visible counts and pair counts (view_cap apart) in one copy.
    HIP_TRY(hipMemcpyAsync(h->h_counters, h->d_counters, (sp::CNT_SLOTS + h->view_cap + nviews) * 4,
                           hipMemcpyDeviceToHost, s));
    *pair_buf = pi;
    return SP_OK;
}

// Render `nviews` views: mode BL_RGB writes outs[v] as device RGB8 buffers,
// BL_DIFF as device float images (the differentiable forward), BL_NONE stops
// after the tile ranges. use_graph: capture the batch once and replay it.
int run_views(sp_handle *h, uint32_t nviews, const sp_view *views, const uint32_t *wh, const float bg[3],
              void *const *outs, int mode, bool use_graph, sp_stats *stats,
              const std::function<int(int)> *after_first = nullptr) {
    const bool blend = mode != BL_NONE;
    const double t0 = now_ms();
    if (nviews == 0) return SP_OK;
    if (nviews > sp::MAX_VIEWS) return fail(SP_E_INVALID, "too many views in one batch");
    if ((uint64_t)nviews * std::max<uint32_t>(h->n, 1) > 0xFFFFFFFFull)
        return fail(SP_E_INVALID, "views x gaussians exceeds 2^32 records");
    int rc = ensure_views(h, nviews);
    // A failed growth (a device allocation that did not succeed) leaves no pair
    // buffer; the kernels always get one.
    if (!rc && h->pair_cap == 0) rc = ensure_pairs(h, 1ull << 20);
    if (!rc) rc = ensure_table(h);  // after a failed growth the pairs may have grown without their table
    if (rc) return rc;
    uint64_t total = 0, total_sb = 0;
    for (uint32_t v = 0; v < nviews; ++v) {
        ViewParams &p = h->h_views[v];
        std::memset(&p, 0, sizeof p);
        p.ctx = make_ctx(views[v], wh[2 * v], wh[2 * v + 1], bg);
        p.tile_base = (uint32_t)total;
        p.sb_base = (uint32_t)total_sb;
        total_sb += (uint64_t)((p.ctx.tiles_x + sp::SB - 1) / sp::SB) * ((p.ctx.tiles_y + sp::SB - 1) / sp::SB);
        if (outs && mode == BL_RGB) p.out = (uint8_t *)outs[v];
        if (outs && mode == BL_DIFF) p.fimg = (float *)outs[v];
        total += (uint64_t)p.ctx.tiles_x * p.ctx.tiles_y;
    }
    if (total >= 0xFFFFFFFFull) return fail(SP_E_INVALID, "too many tiles in one batch");
    const uint32_t total_tiles = (uint32_t)total;
    // Depth keys: DEPTH_KEY_BITS wide when every view's depths are bounded
    // close enough to its near plane, else the full 32 bits for the batch.
    const auto plan_depth = [&] {
        bool narrow = true;
        for (uint32_t v = 0; v < nviews; ++v) {
            h->h_views[v].ctx.key_base = depth_key_base(h, h->h_views[v].ctx);
            narrow = narrow && h->h_views[v].ctx.key_base != 0u;
        }
        if (!narrow)
            for (uint32_t v = 0; v < nviews; ++v) h->h_views[v].ctx.key_base = 0u;
        h->depth_bits = narrow ? sp::DEPTH_KEY_BITS : 32;
    };
    plan_depth();
    rc = ensure_tiles(h, total_tiles);
    if (!rc && mode == BL_DIFF) rc = ensure_pix(h, total_tiles);
    if (!rc && total_sb >= (1ull << sp::SB_SUB_SHIFT))
        return fail(SP_E_INVALID, "too many superblocks in one batch");
    if (!rc) rc = ensure_bin(h, total_tiles, total_sb);
    if (!rc && mode == BL_RGB && h->split_shift > 0) rc = ensure_split(h, total_tiles);
    if (rc) return rc;
    const int tile_bits = (int)bits_for(total_tiles ? total_tiles - 1 : 0);
    const bool events = h->opt.timing && !use_graph;
    for (int attempt = 0;; ++attempt) {
        int pb = 0;
        if (use_graph) {
            GraphKey key;
            key.nviews = nviews;
            key.total_tiles = total_tiles;
            key.tile_bits = (uint32_t)tile_bits;
            key.with_blend = (uint32_t)mode;
            key.split_shift = h->split_shift;
            key.split_kmin = h->split_kmin;
            key.blend_batch = (uint32_t)h->blend_batch;
            key.total_sbs = (uint32_t)total_sb;
            key.depth_bits = (uint32_t)h->depth_bits;
            key.generation = h->generation;
            bool hit = false;
            GraphEntry &ge = graph_slot(h, key, &hit);
            if (!hit) {
                drop_graph(ge);
                HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
                int gpb = 0;
                rc = enqueue(h, nviews, total_tiles, (uint32_t)total_sb, tile_bits, mode, false, &gpb);
                hipGraph_t g = nullptr;
                const hipError_t ce = hipStreamEndCapture(h->stream, &g);
                if (rc) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                if (ce != hipSuccess) return fail(SP_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(ce));
                ge.graph = g;
                HIP_TRY(hipGraphInstantiate(&ge.exec, g, nullptr, nullptr, 0));
                ge.key = key;
                ge.pair_buf = gpb;
            }
            ge.used = ++h->graph_clock;
            HIP_TRY(hipGraphLaunch(ge.exec, h->stream));
            pb = ge.pair_buf;
        } else {
            rc = enqueue(h, nviews, total_tiles, (uint32_t)total_sb, tile_bits, mode, events, &pb, after_first);
            if (rc) return rc;
        }
        HIP_TRY(hipStreamSynchronize(h->stream));
        h->pair_buf = pb;
        if (h->h_counters[sp::CNT_DEPTH_RANGE]) {
            // A depth key did not fit the planned width (the bound failed):
            // this handle sorts full 32-bit depth keys from now on.
            h->depth_full = true;
            HIP_TRY(sp::launch_fill(h->stream, h->d_counters + sp::CNT_DEPTH_RANGE, 1, 0u));
            plan_depth();
            continue;
        }
        // Pair counts saturate at 2^32 - 1 (the scans never wrap): more pairs
        // than a u32 index addresses cannot be rendered (the reference's CPU
        // backend stops there too, src/render/cpu.rs:158-160).
        const bool split_on = mode == BL_RGB && h->split_shift > 0;
        if (h->h_counters[sp::CNT_PAIRS] == 0xFFFFFFFFu || h->h_counters[sp::CNT_SBPAIRS] == 0xFFFFFFFFu ||
            (split_on && h->h_counters[sp::CNT_PAIRS2] == 0xFFFFFFFFu))
            return fail(SP_E_CAPACITY, "more than 2^32 (gaussian, tile) pairs in one phase of the frame");
        if (h->h_counters[sp::CNT_OVERFLOW] && attempt < 3) {
            uint64_t p = h->h_counters[sp::CNT_PAIRS];
            // The second phase's count means something only when the first
            // phase fitted: after a first phase that overflowed, its blend saw
            // partial lists and left nearly every tile open.
            if (mode == BL_RGB && h->split_shift > 0 && p <= h->pair_cap)
                p = std::max<uint64_t>(p, h->h_counters[sp::CNT_PAIRS2]);
            rc = ensure_pairs(h, p + p / 4 + 1024);
            if (rc) return rc;
            continue;
        }
        if (h->h_counters[sp::CNT_OVERFLOW]) return fail(SP_E_CAPACITY, "pair buffer overflow persisted");
        break;
    }
    float tp = 0, ts = 0, tr = 0;
    if (events) {
        float e[6];
        for (int k = 0; k < 6; ++k) HIP_TRY(hipEventElapsedTime(&e[k], h->ev[k], h->ev[k + 1]));
        tp = e[0] + e[3];  // projection, depth sort, both emission phases
        ts = e[1] + e[4];  // both tile sorts and ranges
        tr = e[2] + e[5];  // both blends
    }
    const double total_ms = now_ms() - t0;
    const uint32_t *vis = h->h_counters + sp::CNT_SLOTS;
    const uint32_t *vpairs = vis + h->view_cap;
    uint64_t vis_sum = 0;
    for (uint32_t v = 0; v < nviews; ++v) {
        vis_sum += vis[v];
        if (stats) {
            sp_stats &st = stats[v];
            st.visible = vis[v];
            st.overlaps = vpairs[v];
            // Stage times cover the whole launch sequence (all views of the batch).
            st.project_ms = tp;
            st.sort_ms = ts;
            st.raster_ms = tr;
            st.total_ms = total_ms;
        }
    }
    h->last.width = wh[0];
    h->last.height = wh[1];
    h->last.tiles = total_tiles;
    h->last_sbs = (uint32_t)total_sb;
    h->last.visible = vis_sum;
    uint64_t pair_sum = 0;
    for (uint32_t v = 0; v < nviews; ++v) pair_sum += vpairs[v];
    h->last.pairs = pair_sum;
    h->last.pairs_staged = blend ? h->h_counters[sp::CNT_STAGED] : 0;
    h->last.pair_capacity = h->pair_cap;
    h->last.device_bytes = h->bytes;
    h->last.depth_key_bits = (uint64_t)h->depth_bits;
    return SP_OK;
}

int render_one(sp_handle *h, const sp_view *view, uint32_t width, uint32_t height, const float bg[3],
               uint8_t *d_out, sp_stats *stats) {
    const uint32_t wh[2] = {width, height};
    void *outs[1] = {d_out};
    // Without stage timing the view replays a captured graph (one launch per
    // view); with it, the stages are enqueued with timing events between them.
    return run_views(h, 1, view, wh, bg, outs, d_out != nullptr ? BL_RGB : BL_NONE, !h->opt.timing, stats);
}

// Upload the scene in the handle's storage format. The caller's arrays go to
// the device as they are, one copy each, and a kernel builds the
// blocked column layout (launch_prep, scene_at); k is computed on host threads
// meanwhile, with the host's logf, as Prep::new does (src/render/cpu.rs:384-388).
int upload_scene(sp_handle *h, const sp_scene *scene) {
    const uint32_t n = h->n, rest = h->rest;
    const uint64_t S = h->stride;
    const bool f32 = h->opt.storage == SP_STORE_F32;
    h->soa_bytes = (f32 ? (uint64_t)sp::A_SH + 3ull * rest : (uint64_t)sp::L_COLS) * S * 4;
    h->soa16_bytes = f32 ? 0 : ((uint64_t)sp::H_SH + 3ull * rest) * S * 2;
    h->crow_bytes = (uint64_t)sp::colour_row(rest, !f32) * std::max<uint32_t>(n, 1) * (f32 ? 4 : 2);
    int rc = dalloc(h, (void **)&h->d_soa, h->soa_bytes);
    if (!rc && !f32) rc = dalloc(h, (void **)&h->d_soa16, h->soa16_bytes);
    if (!rc) rc = dalloc(h, &h->d_crow, h->crow_bytes);
    if (rc) return rc;
    // The last block is padded to S splats: the padding reads as zero.
    HIP_TRY(hipMemsetAsync(h->d_soa, 0, h->soa_bytes, h->stream));
    if (!f32) HIP_TRY(hipMemsetAsync(h->d_soa16, 0, h->soa16_bytes, h->stream));
    if (n == 0) {
        HIP_TRY(hipStreamSynchronize(h->stream));
        return SP_OK;
    }
    // k and the bounding box of the finite positions on host threads while
    // the copies run.
    std::vector<float> kv(n);
    const uint32_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const uint32_t nt = std::max(1u, std::min(hw, n / 65536u));
    std::vector<std::array<double, 6>> boxes(nt);
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < nt; ++t)
        pool.emplace_back([&, t] {
            const uint64_t b = (uint64_t)n * t / nt, e = (uint64_t)n * (t + 1) / nt;
            for (uint64_t i = b; i < e; ++i) kv[i] = k_of(scene->opacities[i]);
            std::array<double, 6> &bb = boxes[t];
            bb = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
            for (uint64_t i = b; i < e; ++i) {
                const float *q = scene->positions + 3 * i;
                if (!std::isfinite(q[0]) || !std::isfinite(q[1]) || !std::isfinite(q[2])) continue;
                for (int c = 0; c < 3; ++c) {
                    bb[c] = std::min(bb[c], (double)q[c]);
                    bb[3 + c] = std::max(bb[3 + c], (double)q[c]);
                }
            }
        });
    // Staging: the caller's arrays back to back.
    const uint64_t geo = f32 ? 6ull : 7ull;  // cov3d, or scale + quaternion
    const uint64_t per = 3 + geo + 1 + 1 + 3 + 3ull * rest;
    float *tmp = nullptr;
    rc = dalloc(h, (void **)&tmp, (per * n + 1) * 4);
    if (rc) {
        for (auto &th : pool) th.join();
        return rc;
    }
    sp::PrepInput in{};
    float *p = tmp;
    auto put = [&](const float *src, uint64_t count) -> const float * {
        float *dst = p;
        p += count;
        if (src && count && hipMemcpyAsync(dst, src, count * 4, hipMemcpyHostToDevice, h->stream) != hipSuccess)
            rc = SP_E_DEVICE;
        return dst;
    };
    in.pos = put(scene->positions, 3ull * n);
    if (f32) {
        in.cov = put(scene->cov3d, 6ull * n);
    } else {
        in.scl = put(scene->scales, 3ull * n);
        in.quat = put(scene->rotations, 4ull * n);
    }
    in.opac = put(scene->opacities, n);
    float *dk = p;
    in.k = put(nullptr, n);
    in.col = put(scene->colors, 3ull * n);
    in.sh = put(rest ? scene->sh_rest : nullptr, 3ull * rest * n);
    // One more word after the staged arrays: the extreme-colour flag.
    uint32_t *d_flag = reinterpret_cast<uint32_t *>(p);
    if (!rc && hipMemsetAsync(d_flag, 0, 4, h->stream) != hipSuccess) rc = SP_E_DEVICE;
    for (auto &th : pool) th.join();
    // Splats with a non-finite position are culled in every view, so the box
    // of the others bounds every visible depth.
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (const auto &bb : boxes)
        for (int c = 0; c < 3; ++c) {
            lo[c] = std::min(lo[c], bb[c]);
            hi[c] = std::max(hi[c], bb[3 + c]);
        }
    double r2 = 0.0;
    for (int c = 0; c < 3; ++c) {
        h->bound_c[c] = 0.5 * (lo[c] + hi[c]);
        r2 += 0.25 * (hi[c] - lo[c]) * (hi[c] - lo[c]);
    }
    h->bound_r = std::sqrt(r2);
    h->bound_ok = lo[0] <= hi[0] && std::isfinite(h->bound_r);
    if (!rc && hipMemcpyAsync(dk, kv.data(), 4ull * n, hipMemcpyHostToDevice, h->stream) != hipSuccess)
        rc = SP_E_DEVICE;
    hipError_t e = rc ? hipErrorUnknown
                      : sp::launch_prep(h->stream, h->opt.storage, in, n, rest, h->d_soa, h->d_soa16,
                                        f32 ? static_cast<float *>(h->d_crow) : nullptr,
                                        f32 ? nullptr : static_cast<uint16_t *>(h->d_crow), d_flag);
    uint32_t flag = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&flag, d_flag, 4, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    h->extreme_colour = flag != 0;
    dfree(h, tmp, (per * n + 1) * 4);
    if (rc || e != hipSuccess) return fail(SP_E_DEVICE, std::string("scene upload failed: ") + hipGetErrorString(e));
    return SP_OK;
}

// Device, stream and timing events of a new handle.
int open_handle(sp_handle *h) {
    int rc = setup(h);
    if (rc) return rc;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(SP_E_DEVICE, "hipStreamCreate failed");
    for (auto &e : h->ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(SP_E_DEVICE, "hipEventCreate failed");
    return SP_OK;
}

// Counters, one view's arrays and the initial pair capacity.
int finish_handle(sp_handle *h) {
    int rc = ensure_views(h, 1);
    if (!rc) rc = ensure_pairs(h, initial_pairs(h));
    return rc;
}

int device_ok(const sp_options &opt) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(SP_E_DEVICE, "no HIP device");
    if (opt.device < 0 || opt.device >= count) return fail(SP_E_DEVICE, "device ordinal out of range");
    return SP_OK;
}

// Device address of `scratch` on the current device (h->device), looked up
// once per device.
int static_scratch(int device, uint8_t **out) {
    constexpr int MAX_DEVICES = 64;
    static std::mutex mu;
    static uint8_t *ptr[MAX_DEVICES] = {};
    if (device < 0 || device >= MAX_DEVICES) return fail(SP_E_DEVICE, "device ordinal out of range for the static scratch");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SP_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        ptr[device] = (uint8_t *)p;
    }
    *out = ptr[device];
    return SP_OK;
}

// Bump allocation of 256-aligned pieces of a region; with base == nullptr it
// only measures.
struct Carve {
    uint8_t *base;
    uint64_t off = 0;
    template <class T>
    T *take(uint64_t bytes) {
        off = (off + 255) & ~255ull;
        T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
        off += std::max<uint64_t>(bytes, 16);
        return p;
    }
};

// Points a copy of the handle's per-frame buffers at pieces of `base` for one
// view of width x height with pair capacity `pairs`; returns the bytes used.
uint64_t carve_frame(sp_handle *t, uint8_t *base, uint32_t width, uint32_t height, uint64_t pairs) {
    Carve c{base};
    const uint64_t n = std::max<uint32_t>(t->n, 1);
    const uint64_t blocks = std::max(sp::project_blocks(t->n), sp::emit_blocks(t->n));
    const uint64_t tiles = (uint64_t)((width + 15) / 16) * ((height + 15) / 16);
    const uint64_t sbs = (uint64_t)((width + 16 * sp::SB - 1) / (16 * sp::SB)) * ((height + 16 * sp::SB - 1) / (16 * sp::SB));
    const uint64_t rows = sp::bin_scan_rows((uint32_t)tiles);
    t->view_cap = 1;
    t->d_rec = c.take<uint4>(n * 48);
    t->d_box = c.take<uint2>(n * 8);
    t->d_sbox = c.take<uint2>(n * 8);
    for (int b = 0; b < 2; ++b) {
        t->d_dkey[b] = c.take<uint32_t>(n * 4);
        t->d_dval[b] = c.take<uint32_t>(n * 4);
    }
    t->d_statblk = c.take<uint32_t>(sp::CNT_SLOTS * 4 + 20);
    t->d_counters = t->d_statblk;
    t->d_vstat = t->d_statblk ? t->d_statblk + sp::CNT_SLOTS : nullptr;
    t->d_bcnt = c.take<uint32_t>(blocks * 4);
    t->d_bcnt_all = c.take<uint32_t>(blocks * 4);
    t->d_bsb = c.take<uint32_t>(blocks * 4);
    t->d_views = c.take<ViewParams>(sizeof(ViewParams));
    t->d_rowtot = c.take<uint32_t>(sp::RADIX_BINS_MAX * 4);
    t->table_words = table_need_single(t->n, pairs);
    t->d_table = c.take<uint32_t>(t->table_words * 4);
    t->pair_cap = pairs;
    for (int b = 0; b < 2; ++b) {
        t->d_pkey[b] = c.take<uint32_t>(pairs * 4);
        t->d_pval[b] = c.take<uint32_t>(pairs * 4);
    }
    t->tile_cap = tiles + 1;
    t->d_tile_start = c.take<uint32_t>((tiles + 1) * 4);
    t->sb_cap = sbs + 1;
    t->d_sb_start = c.take<uint32_t>((sbs + 1) * 4);
    t->d_segc = c.take<uint32_t>(sbs * sp::BIN_SEGS * sp::SB * sp::SB * 4);
    t->d_sbtile = c.take<uint32_t>(sbs * sp::SB * sp::SB * 4);
    t->tcount_cap = rows * sp::BIN_SCAN_ROW;
    t->d_tcount = c.take<uint32_t>(rows * sp::BIN_SCAN_ROW * 4);
    t->d_trow = c.take<uint32_t>(rows * 4);
    t->split_cap = t->split_shift > 0 ? tiles : 0;
    t->d_done = t->split_shift > 0 ? c.take<uint32_t>(tiles * 4) : nullptr;
    t->d_openm = t->split_shift > 0 ? c.take<uint32_t>(tiles * 4) : nullptr;
    t->d_pstate = t->split_shift > 0 ? c.take<float>(tiles * sp::BLEND_THREADS * 16) : nullptr;
    t->rgb_cap = 0;
    t->d_rgb = nullptr;
    return c.off;
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
    drop_graphs(h);
    (void)hipFree(h->d_soa);
    (void)hipFree(h->d_soa16);
    (void)hipFree(h->d_crow);
    (void)hipFree(h->d_rec);
    (void)hipFree(h->d_box);
    (void)hipFree(h->d_sbox);
    for (int b = 0; b < 2; ++b) {
        (void)hipFree(h->d_dkey[b]);
        (void)hipFree(h->d_dval[b]);
        (void)hipFree(h->d_pkey[b]);
        (void)hipFree(h->d_pval[b]);
    }
    (void)hipFree(h->d_statblk);
    (void)hipFree(h->d_bcnt);
    (void)hipFree(h->d_bcnt_all);
    (void)hipFree(h->d_views);
    (void)hipFree(h->d_table);
    (void)hipFree(h->d_tile_start);
    (void)hipFree(h->d_sb_start);
    (void)hipFree(h->d_segc);
    (void)hipFree(h->d_sbtile);
    (void)hipFree(h->d_tcount);
    (void)hipFree(h->d_trow);
    (void)hipFree(h->d_bsb);
    (void)hipFree(h->d_rgb);
    (void)hipFree(h->d_rowtot);
    (void)hipFree(h->d_pbase);
    (void)hipFree(h->d_pslot);
    (void)hipFree(h->d_prid);
    (void)hipFree(h->d_pixT);
    (void)hipFree(h->d_pixlast);
    (void)hipFree(h->d_gpair);
    (void)hipFree(h->d_fresh);
    (void)hipFree(h->d_g2d);
    (void)hipFree(h->d_done);
    (void)hipFree(h->d_openm);
    (void)hipFree(h->d_pstate);
    if (h->h_counters) (void)hipHostFree(h->h_counters);
    if (h->h_views) (void)hipHostFree(h->h_views);
    if (h->h_rgb_pin) (void)hipHostFree(h->h_rgb_pin);
    for (auto &e : h->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : h->ev_rb)
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
    if (opt.storage != SP_STORE_F32 && opt.storage != SP_STORE_F16 && opt.storage != SP_STORE_BF16)
        return fail(SP_E_INVALID, "unknown storage precision");
    const uint32_t n = (uint32_t)scene->n;
    const uint32_t rest = (scene->sh_degree + 1) * (scene->sh_degree + 1) - 1;
    if (n && (!scene->positions || !scene->opacities || !scene->colors))
        return fail(SP_E_INVALID, "scene arrays missing");
    if (n && opt.storage == SP_STORE_F32 && !scene->cov3d) return fail(SP_E_INVALID, "cov3d missing");
    if (n && opt.storage != SP_STORE_F32 && (!scene->scales || !scene->rotations))
        return fail(SP_E_INVALID, "16-bit storage needs scales and rotations");
    if (n && rest && !scene->sh_rest) return fail(SP_E_INVALID, "sh_rest missing for sh_degree > 0");
    if (device_ok(opt)) return SP_E_DEVICE;

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
    int rc = open_handle(h);
    if (rc) return bail(rc);
    // Structure-of-arrays layout (one 64-aligned array per field).
    h->stride = std::max<uint64_t>(((uint64_t)n + 63) / 64 * 64, 64);
    rc = upload_scene(h, scene);
    if (!rc) rc = finish_handle(h);
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
    return render_one(h, view, width, height, background, rgb_out_device, stats);
}

int sp_render_device_static(sp_handle *h, const sp_view *view, uint32_t width, uint32_t height,
                            const float background[3], uint8_t *rgb_out_device, sp_stats *stats) {
    if (!h || !view || !background) return fail(SP_E_INVALID, "sp_render_device_static: NULL argument");
    if (stats) std::memset(stats, 0, sizeof *stats);
    int rc = check_frame(width, height);
    if (rc) return rc;
    if (width == 0 || height == 0) return SP_OK;
    if (!rgb_out_device) return fail(SP_E_INVALID, "sp_render_device_static: NULL output");
    if (h->diff) return fail(SP_E_INVALID, "sp_render_device_static: not for a differentiable handle");
    rc = setup(h);
    if (rc) return rc;
    uint8_t *base = nullptr;
    rc = static_scratch(h->device, &base);
    if (rc) return rc;
    // A copy of the handle whose per-frame buffers are pieces of the scratch:
    // the scene, stream, events, tuning and pinned host staging are the
    // handle's own (all made by sp_create), and nothing is captured.
    sp_handle t = *h;
    for (GraphEntry &g : t.graphs) g = GraphEntry{};
    t.fixed = true;
    // The pair capacity takes what the fixed buffers leave.
    const uint64_t fixed_bytes = carve_frame(&t, nullptr, width, height, 0);
    if (fixed_bytes + 256 * 4 > SP_STATIC_SCRATCH_BYTES)
        return fixed_full("per-splat and per-tile arrays");
    uint64_t pairs = (SP_STATIC_SCRATCH_BYTES - fixed_bytes) / 16;
    while (pairs > 0 && carve_frame(&t, nullptr, width, height, pairs) > SP_STATIC_SCRATCH_BYTES) pairs -= pairs / 32 + 1;
    pairs = std::min<uint64_t>(pairs, 0xFFFFFFFFull);
    if (pairs == 0) return fixed_full("tile pairs");
    carve_frame(&t, base, width, height, pairs);
    // The frame counters start at 0 (a fill kernel: the call may sit between
    // captured work of the caller).
    HIP_TRY(sp::launch_fill(t.stream, t.d_statblk, sp::CNT_SLOTS + 5, 0u));
    const uint32_t wh[2] = {width, height};
    void *outs[1] = {rgb_out_device};
    rc = run_views(&t, 1, view, wh, background, outs, BL_RGB, false, stats);
    h->last = t.last;
    h->last.device_bytes = h->bytes;
    return rc;
}

// Copies a frame from h->d_rgb into the caller's (pageable) buffer: through
// the handle's pinned staging buffer in up to four chunks, the host copying
// chunk i while chunk i + 1 is in flight (a pinned DMA plus a host copy take
// about 0.2 ms for a 1080p frame, a pageable hipMemcpy about 0.36 ms). Falls
// back to a plain hipMemcpy when no pinned buffer can be had.
static int read_back(sp_handle *h, uint8_t *dst, uint64_t bytes) {
    if (bytes > h->rgb_pin_cap) {
        if (h->h_rgb_pin) (void)hipHostFree(h->h_rgb_pin);
        h->h_rgb_pin = nullptr;
        h->rgb_pin_cap = 0;
        if (hipHostMalloc((void **)&h->h_rgb_pin, bytes, hipHostMallocDefault) == hipSuccess)
            h->rgb_pin_cap = bytes;
        else
            (void)hipGetLastError();
    }
    bool events = h->h_rgb_pin != nullptr;
    for (auto &e : h->ev_rb)
        if (events && !e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            e = nullptr;
            events = false;
        }
    if (!events) {
        HIP_TRY(hipMemcpy(dst, h->d_rgb, bytes, hipMemcpyDeviceToHost));
        return SP_OK;
    }
    const uint32_t nch = bytes >= (1u << 20) ? 4u : 1u;
    const uint64_t step = (((bytes + nch - 1) / nch) + 4095) & ~4095ull;
    for (uint32_t i = 0; i < nch; ++i) {
        const uint64_t a = std::min(bytes, (uint64_t)i * step), b = std::min(bytes, a + step);
        if (b > a) HIP_TRY(hipMemcpyAsync(h->h_rgb_pin + a, h->d_rgb + a, b - a, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipEventRecord(h->ev_rb[i], h->stream));
    }
    for (uint32_t i = 0; i < nch; ++i) {
        const uint64_t a = std::min(bytes, (uint64_t)i * step), b = std::min(bytes, a + step);
        HIP_TRY(hipEventSynchronize(h->ev_rb[i]));
        if (b > a) std::memcpy(dst + a, h->h_rgb_pin + a, b - a);
    }
    return SP_OK;
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
    rc = render_one(h, view, width, height, background, h->d_rgb, stats);
    if (rc) return rc;
    rc = read_back(h, rgb_out, bytes);
    if (rc) return rc;
    if (stats) stats->total_ms = now_ms() - t0;
    return SP_OK;
}

int sp_downsample_device(sp_handle *h, const uint8_t *src, uint32_t width, uint32_t height, uint32_t factor,
                         uint8_t *dst) {
    if (!h) return fail(SP_E_INVALID, "sp_downsample_device: NULL handle");
    factor = std::max(factor, 1u);
    const uint32_t dw = width / factor, dh = height / factor;
    if ((uint64_t)dw * dh == 0) return SP_OK;
    if (!src || !dst) return fail(SP_E_INVALID, "sp_downsample_device: NULL frame");
    int rc = setup(h);
    if (rc) return rc;
    HIP_TRY(sp::launch_downsample(h->stream, src, width, factor, dw, dh, dst));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return SP_OK;
}

int sp_render_cells(sp_handle *h, const sp_view *view, uint32_t cols, uint32_t rows, uint32_t factor,
                    const float background[3], int32_t mode, uint8_t *cells_out, sp_stats *stats) {
    if (!h || !view || !background) return fail(SP_E_INVALID, "sp_render_cells: NULL argument");
    if (stats) std::memset(stats, 0, sizeof *stats);
    if (mode != SP_TERM_TRUECOLOR && mode != SP_TERM_ANSI256) return fail(SP_E_INVALID, "sp_render_cells: bad mode");
    factor = std::max(factor, 1u);
    const uint64_t w = (uint64_t)cols * factor, hh = 2ull * rows * factor;
    if (w > SP_MAX_DIM || hh > SP_MAX_DIM) return fail(SP_E_PIXEL_LIMIT, "sp_render_cells: pixel limit");
    if (w == 0 || hh == 0) return SP_OK;
    if (!cells_out) return fail(SP_E_INVALID, "sp_render_cells: NULL output");
    int rc = setup(h);
    if (rc) return rc;
    // The frame and its cells share the handle's RGB buffer: cells after the frame.
    const uint64_t bytes = 3 * w * hh;
    const uint64_t cell_bytes = (mode == SP_TERM_ANSI256 ? 2ull : 6ull) * cols * rows;
    rc = ensure_rgb(h, bytes + cell_bytes);
    if (rc) return rc;
    const double t0 = now_ms();
    rc = render_one(h, view, (uint32_t)w, (uint32_t)hh, background, h->d_rgb, stats);
    if (rc) return rc;
    HIP_TRY(sp::launch_cells(h->stream, h->d_rgb, (uint32_t)w, factor, cols, rows, mode, h->d_rgb + bytes));
    HIP_TRY(hipMemcpyAsync(cells_out, h->d_rgb + bytes, cell_bytes, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
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
    ViewParams *p = h->h_views;
    std::memset(p, 0, sizeof *p);
    p->ctx = make_ctx(*view, width, height, bg);
    float *d_dbg = nullptr;
    uint16_t *d_bb = nullptr;
    uint8_t *d_valid = nullptr;
    HIP_TRY(hipMalloc(&d_dbg, n * 48));
    HIP_TRY(hipMalloc(&d_bb, n * 8));
    HIP_TRY(hipMalloc(&d_valid, n));
    hipError_t e = hipMemsetAsync(d_dbg, 0, n * 48, h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_bb, 0, n * 8, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(h->d_views, h->h_views, sizeof(ViewParams), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess)
        e = sp::launch_project(h->stream, dev_scene(h), h->d_views, 1, h->d_rec, h->d_sbox, h->d_dkey[0], d_dbg, d_bb,
                               d_valid, nullptr, false);
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
    rc = render_one(h, view, width, height, bg, nullptr, &st);
    if (rc) return rc;
    const uint64_t np = st.overlaps;
    *n_pairs = np;
    const uint32_t tiles = ((width + 15) / 16) * ((height + 15) / 16);
    HIP_TRY(hipMemcpy(tile_start, h->d_tile_start, (tiles + 1ull) * 4, hipMemcpyDeviceToHost));
    if (np > cap) return fail(SP_E_CAPACITY, "sp_debug_pairs: capacity too small");
    if (np == 0) return SP_OK;
    if (!keys || !vals) return fail(SP_E_INVALID, "sp_debug_pairs: NULL keys/vals");
    std::vector<uint32_t> tk(np);
    HIP_TRY(hipMemcpy(tk.data(), h->d_pkey[h->pair_buf], np * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(vals, h->d_pval[h->pair_buf], np * 4, hipMemcpyDeviceToHost));
    // Depth bits of each visible splat, from the depth sort's (depth, index) output.
    const uint32_t nv = st.visible;
    std::vector<uint32_t> dk(nv), dv(nv), depth(h->n, 0u);
    HIP_TRY(hipMemcpy(dk.data(), h->d_dkey[h->depth_buf], nv * 4ull, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(dv.data(), h->d_dval[h->depth_buf], nv * 4ull, hipMemcpyDeviceToHost));
    const uint32_t kb = h->h_views[0].ctx.key_base;  // the sort's keys are the depth bits less the view's base
    for (uint32_t q = 0; q < nv; ++q) depth[dv[q]] = dk[q] + kb;
    for (uint64_t p = 0; p < np; ++p) keys[p] = ((uint64_t)tk[p] << 32) | depth[vals[p]];
    return SP_OK;
}

int sp_debug_split_pairs(sp_handle *h, const sp_view *view, uint32_t width, uint32_t height, uint32_t *first_vals,
                         uint32_t *first_start, uint32_t *done, uint32_t *second_vals, uint32_t *second_start,
                         uint64_t cap, uint64_t *n_first, uint64_t *n_second) {
    if (!h || !view || !first_start || !done || !second_start || !n_first || !n_second)
        return fail(SP_E_INVALID, "sp_debug_split_pairs: NULL argument");
    *n_first = *n_second = 0;
    int rc = check_frame(width, height);
    if (rc) return rc;
    if (width == 0 || height == 0) return SP_OK;
    if (h->diff) return fail(SP_E_INVALID, "sp_debug_split_pairs: a differentiable handle has no split");
    rc = setup(h);
    if (rc) return rc;
    rc = ensure_rgb(h, 3ull * width * height);
    if (rc) return rc;
    const uint32_t tiles = ((width + 15) / 16) * ((height + 15) / 16);
    const bool split = h->split_shift > 0;
    std::vector<uint32_t> v1;
    // Copy the first phase's lists and tile flags out before the second phase
    // reuses the buffers.
    const std::function<int(int)> grab = [&](int pi) -> int {
        HIP_TRY(hipStreamSynchronize(h->stream));
        HIP_TRY(hipMemcpy(first_start, h->d_tile_start, (tiles + 1ull) * 4, hipMemcpyDeviceToHost));
        const uint64_t n1 = first_start[tiles];
        v1.resize(n1);
        if (n1) HIP_TRY(hipMemcpy(v1.data(), h->d_pval[pi], n1 * 4, hipMemcpyDeviceToHost));
        if (split)
            HIP_TRY(hipMemcpy(done, h->d_done, tiles * 4ull, hipMemcpyDeviceToHost));
        else
            std::memset(done, 0, tiles * 4ull);
        return SP_OK;
    };
    const uint32_t wh[2] = {width, height};
    const float bg[3] = {0.f, 0.f, 0.f};
    void *outs[1] = {h->d_rgb};
    rc = run_views(h, 1, view, wh, bg, outs, BL_RGB, false, nullptr, &grab);
    if (rc) return rc;
    *n_first = v1.size();
    if (split) {
        HIP_TRY(hipMemcpy(second_start, h->d_tile_start, (tiles + 1ull) * 4, hipMemcpyDeviceToHost));
        *n_second = second_start[tiles];
    } else {
        std::memset(second_start, 0, (tiles + 1ull) * 4);
    }
    if (*n_first > cap || *n_second > cap) return fail(SP_E_CAPACITY, "sp_debug_split_pairs: capacity too small");
    if ((*n_first && !first_vals) || (*n_second && !second_vals))
        return fail(SP_E_INVALID, "sp_debug_split_pairs: NULL value buffer");
    if (*n_first) std::memcpy(first_vals, v1.data(), *n_first * 4);
    if (*n_second) HIP_TRY(hipMemcpy(second_vals, h->d_pval[h->pair_buf], *n_second * 4, hipMemcpyDeviceToHost));
    return SP_OK;
}

int sp_debug_reach(const float *proj, const uint16_t *bbox, uint32_t width, uint32_t height, uint8_t *masks) {
    if (!proj || !bbox || !masks) return fail(SP_E_INVALID, "sp_debug_reach: NULL argument");
    int rc = check_frame(width, height);
    if (rc) return rc;
    // The record as the projection packs it (sp_splat.h): x, y, -a/2, -b;
    // -c/2, opacity, r, g; b, box lo, box hi, falloff threshold.
    auto u = [](float f) {
        uint32_t r;
        std::memcpy(&r, &f, 4);
        return r;
    };
    const float k = k_of(proj[8]);
    const uint32_t w0[4] = {u(proj[0]), u(proj[1]), u(-0.5f * proj[2]), u(-proj[3])};
    const uint32_t w1[4] = {u(-0.5f * proj[4]), u(proj[8]), u(proj[5]), u(proj[6])};
    const uint32_t w2[4] = {u(proj[7]), (uint32_t)bbox[0] | ((uint32_t)bbox[1] << 16),
                            (uint32_t)bbox[2] | ((uint32_t)bbox[3] << 16), u(-0.5f * (k * k) - 0.02f)};
    const uint32_t tx = (width + 15) / 16, ty = (height + 15) / 16;
    for (uint32_t y = 0; y < ty; ++y)
        for (uint32_t x = 0; x < tx; ++x)
            masks[(uint64_t)y * tx + x] = (uint8_t)sp::quadrant_mask_words(w0, w1, w2, 16 * x, 16 * y);
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
    if (stats)
        for (uint32_t v = 0; v < n_views; ++v) std::memset(&stats[v], 0, sizeof(sp_stats));
    if (n_views == 0) return SP_OK;
    int rc = setup(h);
    if (rc) return rc;
    // Zero-sized views render nothing (as Renderer::render); the others form the batch.
    std::vector<sp_view> bv;
    std::vector<uint32_t> bwh;
    std::vector<uint32_t> idx;
    uint64_t host_bytes = 0;
    for (uint32_t v = 0; v < n_views; ++v) {
        rc = check_frame(wh[2 * v], wh[2 * v + 1]);
        if (rc) return rc;
        if (wh[2 * v] == 0 || wh[2 * v + 1] == 0) continue;
        if (!outs[v]) return fail(SP_E_INVALID, "sp_render_batch: NULL output");
        bv.push_back(views[v]);
        bwh.push_back(wh[2 * v]);
        bwh.push_back(wh[2 * v + 1]);
        idx.push_back(v);
        host_bytes += 3ull * wh[2 * v] * wh[2 * v + 1];
    }
    const uint32_t nb = (uint32_t)idx.size();
    if (nb == 0) return SP_OK;
    std::vector<uint8_t *> douts(nb);
    if (outs_on_device) {
        for (uint32_t k = 0; k < nb; ++k) douts[k] = outs[idx[k]];
    } else {
        rc = ensure_rgb(h, host_bytes);
        if (rc) return rc;
        uint64_t o = 0;
        for (uint32_t k = 0; k < nb; ++k) {
            douts[k] = h->d_rgb + o;
            o += 3ull * bwh[2 * k] * bwh[2 * k + 1];
        }
    }
    std::vector<sp_stats> st(nb);
    std::vector<void *> vouts(douts.begin(), douts.end());
    rc = run_views(h, nb, bv.data(), bwh.data(), background, vouts.data(), BL_RGB, true, st.data());
    if (rc) return rc;
    if (!outs_on_device) {
        for (uint32_t k = 0; k < nb; ++k)
            HIP_TRY(hipMemcpy(outs[idx[k]], douts[k], 3ull * bwh[2 * k] * bwh[2 * k + 1], hipMemcpyDeviceToHost));
    }
    if (stats)
        for (uint32_t k = 0; k < nb; ++k) stats[idx[k]] = st[k];
    return SP_OK;
}

int sp_get_tuning(const sp_handle *h, sp_tuning *t) {
    if (!h || !t) return fail(SP_E_INVALID, "sp_get_tuning: NULL argument");
    std::memset(t, 0, sizeof *t);
    t->blend_batch = h->blend_batch;
    t->split_shift = h->split_shift;
    t->split_kmin = h->split_kmin;
    return SP_OK;
}

int sp_set_tuning(sp_handle *h, const sp_tuning *t) {
    if (!h || !t) return fail(SP_E_INVALID, "sp_set_tuning: NULL argument");
    if (t->blend_batch != 64 && t->blend_batch != 128 && t->blend_batch != 256)
        return fail(SP_E_INVALID, "blend_batch must be 64, 128 or 256");
    if (t->split_shift > 16) return fail(SP_E_INVALID, "split_shift must be 0..16");
    h->blend_batch = t->blend_batch;
    h->split_shift = t->split_shift;
    h->split_kmin = t->split_kmin;
    return SP_OK;
}

// Coordinate descent over the tunables on the caller's batch: the split depth
// first (the largest effect), then the LDS batch of the blend, then the
// minimum size of the first phase. Each candidate is timed as the median of
// `reps` captured batch replays after one warm-up (which also captures the
// graph). Every candidate renders the same bytes; only the time differs.
int sp_autotune(sp_handle *h, const sp_view *views, const uint32_t *wh, uint32_t n_views,
                const float background[3], int32_t reps, sp_tuning *best, double *best_ms) {
    if (!h || !views || !wh || !background || n_views == 0) return fail(SP_E_INVALID, "sp_autotune: bad argument");
    if (h->diff) return fail(SP_E_INVALID, "sp_autotune: the differentiable path has no tunables");
    int rc = setup(h);
    if (rc) return rc;
    std::vector<sp_view> bv;
    std::vector<uint32_t> bwh;
    uint64_t bytes = 0;
    for (uint32_t v = 0; v < n_views; ++v) {
        rc = check_frame(wh[2 * v], wh[2 * v + 1]);
        if (rc) return rc;
        if (wh[2 * v] == 0 || wh[2 * v + 1] == 0) continue;
        bv.push_back(views[v]);
        bwh.push_back(wh[2 * v]);
        bwh.push_back(wh[2 * v + 1]);
        bytes += 3ull * wh[2 * v] * wh[2 * v + 1];
    }
    const uint32_t nb = (uint32_t)bv.size();
    if (nb == 0) return fail(SP_E_INVALID, "sp_autotune: every view is empty");
    rc = ensure_rgb(h, bytes);
    if (rc) return rc;
    std::vector<void *> outs(nb);
    uint64_t o = 0;
    for (uint32_t k = 0; k < nb; ++k) {
        outs[k] = h->d_rgb + o;
        o += 3ull * bwh[2 * k] * bwh[2 * k + 1];
    }
    reps = std::max(1, std::min(reps, 50));
    sp_tuning cur;
    sp_get_tuning(h, &cur);
    // Each candidate starts from a new handle's pair capacity, which its
    // untimed first frame grows to its own need: a candidate that sorts every
    // pair unsplit would otherwise leave capacity-sized grids to the ones after.
    auto time_of = [&](const sp_tuning &t, double *ms) -> int {
        int r = sp_set_tuning(h, &t);
        if (r) return r;
        r = reset_pairs(h);
        if (r) return r;
        r = run_views(h, nb, bv.data(), bwh.data(), background, outs.data(), BL_RGB, true, nullptr);
        if (r) return r;
        std::vector<double> ts;
        for (int k = 0; k < reps; ++k) {
            const double t0 = now_ms();
            r = run_views(h, nb, bv.data(), bwh.data(), background, outs.data(), BL_RGB, true, nullptr);
            if (r) return r;
            ts.push_back(now_ms() - t0);
        }
        std::sort(ts.begin(), ts.end());
        *ms = ts[ts.size() / 2];
        return SP_OK;
    };
    double cur_ms = 0;
    rc = time_of(cur, &cur_ms);
    if (rc) return rc;
    auto consider = [&](sp_tuning t) -> int {
        double ms = 0;
        const int r = time_of(t, &ms);
        if (r) return r;
        if (ms < cur_ms) {
            cur = t;
            cur_ms = ms;
        }
        return SP_OK;
    };
    for (uint32_t sh : {0u, 1u, 2u, 3u, 4u, 5u}) {
        if (sh == cur.split_shift) continue;
        sp_tuning t = cur;
        t.split_shift = sh;
        if ((rc = consider(t))) return rc;
    }
    for (int bb : {64, 128, 256}) {
        if (bb == cur.blend_batch) continue;
        sp_tuning t = cur;
        t.blend_batch = bb;
        if ((rc = consider(t))) return rc;
    }
    if (cur.split_shift > 0)
        for (uint32_t km : {4096u, 16384u, 65536u, 98304u}) {
            if (km == cur.split_kmin) continue;
            sp_tuning t = cur;
            t.split_kmin = km;
            if ((rc = consider(t))) return rc;
        }
    rc = sp_set_tuning(h, &cur);
    if (rc) return rc;
    rc = reset_pairs(h);  // later frames grow them to the chosen setting's need
    if (rc) return rc;
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return SP_OK;
}

int sp_diff_create(uint64_t n, uint32_t sh_degree, const sp_options *opt_in, sp_handle **out) {
    if (!out) return fail(SP_E_INVALID, "sp_diff_create: out is NULL");
    *out = nullptr;
    sp_options opt;
    if (opt_in)
        opt = *opt_in;
    else
        sp_default_options(&opt);
    if (sh_degree > 3) return fail(SP_E_INVALID, "sh_degree must be 0..3");
    if (n > 0x7FFFFFFFull) return fail(SP_E_INVALID, "more than 2^31 gaussians");
    if (device_ok(opt)) return SP_E_DEVICE;
    sp_handle *h = new sp_handle();
    h->device = opt.device;
    h->opt = opt;
    h->opt.storage = SP_STORE_F32;
    h->diff = true;
    h->n = (uint32_t)n;
    h->deg = sh_degree;
    h->rest = (sh_degree + 1) * (sh_degree + 1) - 1;
    h->stride = std::max<uint64_t>((n + 63) / 64 * 64, 64);
    int rc = open_handle(h);
    if (!rc) rc = finish_handle(h);
    if (rc) {
        sp_destroy(h);
        return rc;
    }
    *out = h;
    return SP_OK;
}

int sp_diff_forward(sp_handle *h, const sp_params *params, const sp_view *views, const uint32_t *wh,
                    uint32_t n_views, const float background[3], float *const *images, sp_stats *stats) {
    if (!h || !params || !views || !wh || !background || !images)
        return fail(SP_E_INVALID, "sp_diff_forward: NULL argument");
    if (!h->diff) return fail(SP_E_INVALID, "sp_diff_forward: handle not made by sp_diff_create");
    if (h->n && (!params->means || !params->log_scales || !params->quats || !params->opacity_logits ||
                 !params->colors || (h->rest && !params->sh_rest)))
        return fail(SP_E_INVALID, "sp_diff_forward: parameter array missing");
    for (uint32_t v = 0; v < n_views; ++v) {
        const int rc = check_frame(wh[2 * v], wh[2 * v + 1]);
        if (rc) return rc;
        if (!images[v] && wh[2 * v] && wh[2 * v + 1]) return fail(SP_E_INVALID, "sp_diff_forward: NULL image");
    }
    int rc = setup(h);
    if (rc) return rc;
    h->prm = sp::DevParams{params->means, params->log_scales, params->quats, params->opacity_logits,
                           params->colors, params->sh_rest};
    h->fwd_views = 0;
    std::vector<void *> outs(images, images + n_views);
    rc = run_views(h, n_views, views, wh, background, outs.data(), BL_DIFF, false, stats);
    if (rc) return rc;
    h->fwd_views = n_views;
    h->fwd_tiles = (uint32_t)h->last.tiles;
    h->fwd_sbs = h->last_sbs;
    h->fwd_pairs = h->last.pairs;
    return SP_OK;
}

int sp_diff_backward(sp_handle *h, const sp_params *params, const float *const *grad_images,
                     const sp_param_grads *grads, double *backward_ms) {
    if (!h || !params || !grad_images || !grads) return fail(SP_E_INVALID, "sp_diff_backward: NULL argument");
    if (!h->diff) return fail(SP_E_INVALID, "sp_diff_backward: handle not made by sp_diff_create");
    if (h->fwd_views == 0) return fail(SP_E_INVALID, "sp_diff_backward: no preceding sp_diff_forward");
    if (h->n && (!grads->means || !grads->log_scales || !grads->quats || !grads->opacity_logits ||
                 !grads->colors || (h->rest && !grads->sh_rest)))
        return fail(SP_E_INVALID, "sp_diff_backward: gradient array missing");
    const uint32_t nv = h->fwd_views;
    for (uint32_t v = 0; v < nv; ++v)
        if (!grad_images[v] && h->h_views[v].ctx.w && h->h_views[v].ctx.h)
            return fail(SP_E_INVALID, "sp_diff_backward: NULL gradient image");
    int rc = setup(h);
    if (rc) return rc;
    const double t0 = now_ms();
    h->prm = sp::DevParams{params->means, params->log_scales, params->quats, params->opacity_logits,
                           params->colors, params->sh_rest};
    const uint64_t np = h->fwd_pairs, records = (uint64_t)nv * h->n;
    rc = ensure_grad(h, std::max<uint64_t>(np, 1), std::max<uint64_t>(records, 1));
    if (rc) return rc;
    for (uint32_t v = 0; v < nv; ++v) h->h_views[v].gimg = grad_images[v];
    hipStream_t s = h->stream;
    HIP_TRY(hipMemcpyAsync(h->d_views, h->h_views, nv * sizeof(ViewParams), hipMemcpyHostToDevice, s));
    // Per-pair gradients are not cleared: each backward marks the pairs it
    // writes in a bit array cleared here, and the reduction counts the others
    // as zero.
    HIP_TRY(sp::launch_fill(s, h->d_fresh, (np + 31) / 32, 0u));
    HIP_TRY(sp::launch_fill_f32(s, h->d_g2d, (uint64_t)nv * h->n * sp::NG2D, 0.0f));  // the records in use
    HIP_TRY(sp::launch_blend_bwd(s, h->d_views, nv, h->fwd_tiles, h->fwd_sbs, h->d_tile_start, h->d_pval[h->pair_buf],
                                 h->d_pbase, h->d_rec, h->d_pixT, h->d_pixlast, h->d_gpair, h->gpair_cap,
                                 h->d_fresh));
    HIP_TRY(sp::launch_pair_reduce(s, h->d_gpair, h->gpair_cap, h->d_fresh, h->d_pslot, h->d_prid, h->d_vstat,
                                   (uint32_t)h->n, nv, h->d_g2d));
    HIP_TRY(sp::launch_project_bwd(s, dev_scene(h), h->d_views, nv, h->d_g2d, grads->means,
                                   grads->log_scales, grads->quats, grads->opacity_logits, grads->colors,
                                   grads->sh_rest));
    HIP_TRY(hipStreamSynchronize(s));
    if (backward_ms) *backward_ms = now_ms() - t0;
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
