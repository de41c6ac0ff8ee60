/*
 * splat_hip.h — C ABI of libsplat_hip.so, the MI355X (gfx950) Gaussian-splat
 * rasterizer that plugs in behind splat-tui's `Renderer` trait as
 * `--backend hip`.
 *
 * Every entry point returns 0 on success and a negative SP_E_* code on
 * failure; sp_last_error() then holds a message for the calling thread.
 * No exception, panic or abort crosses this boundary. Host buffers are owned
 * by the caller; device buffers are owned by the handle. A handle is used by
 * one thread at a time (the trait is `Send`, not `Sync`).
 *
 * Reference interfaces replaced (splat-tui, file:line):
 *   sp_create / sp_destroy   render::create + CpuTiled::new
 *                            (src/render/mod.rs:115-121, src/render/cpu.rs:82-91)
 *   sp_render                Renderer::render (src/render/mod.rs:86-97),
 *                            implemented as CpuTiled::render (src/render/cpu.rs:109-222)
 *   sp_name                  Renderer::name (src/render/mod.rs:87)
 *   sp_stats                 RenderStats (src/render/mod.rs:73-83)
 *   sp_view                  camera::View (src/camera.rs:163-170)
 *   sp_scene                 scene::Scene after finalize (src/scene.rs:11-28, :62-94)
 */
#ifndef SPLAT_HIP_H
#define SPLAT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants shared with the reference (src/render/mod.rs:23-33, cpu.rs:34) ---- */
#define SP_TILE 16
#define SP_MAX_DIM 16384u

/* ---- error codes ---- */
#define SP_OK 0
#define SP_E_INVALID (-1)    /* bad argument (null pointer, inconsistent sizes) */
#define SP_E_PIXEL_LIMIT (-2) /* width or height above SP_MAX_DIM ("pixel limit") */
#define SP_E_DEVICE (-3)     /* HIP runtime error or no device */
#define SP_E_NOMEM (-4)      /* device allocation failed / over the handle's budget */
#define SP_E_CAPACITY (-5)   /* caller-supplied capacity too small (debug calls) */
#define SP_E_UNSUPPORTED (-6)
#define SP_E_FORMAT (-7)     /* malformed or unsupported scene file */
#define SP_E_IO (-8)         /* scene file cannot be read */

/* Storage precision of the per-splat parameters held on the device. */
#define SP_STORE_F32 0
#define SP_STORE_F16 1
#define SP_STORE_BF16 2

/*
 * Activated scene parameters, laid out exactly as splat-tui's Scene vectors
 * (glam Vec3 = 3 packed f32, Quat = x,y,z,w). `cov3d` is the upper triangle
 * (xx, xy, xz, yy, yz, zz) that Scene::finalize computes (src/scene.rs:87-92,
 * :170-178). `scales`/`rotations` may be NULL for SP_STORE_F32; the FP16/BF16
 * storage variants and the gradient entry points need them.
 */
typedef struct sp_scene {
    uint64_t n;               /* number of gaussians */
    uint32_t sh_degree;       /* 0..3 */
    uint32_t reserved;
    const float *positions;   /* [n][3] */
    const float *cov3d;       /* [n][6] */
    const float *opacities;   /* [n] in [0,1] */
    const float *colors;      /* [n][3] base colour (SH_C0*f_dc + 0.5) */
    const float *sh_rest;     /* [n][(deg+1)^2-1][3], NULL when sh_degree == 0 */
    const float *scales;      /* [n][3] or NULL */
    const float *rotations;   /* [n][4] unit quaternions (x,y,z,w) or NULL */
} sp_scene;

/* camera::View (src/camera.rs:163-170): world->camera rows are right/down/forward. */
typedef struct sp_view {
    float eye[3];
    float right[3];
    float down[3];
    float forward[3];
    float fov_y;  /* radians */
    float near_z; /* near plane distance (View::near) */
} sp_view;

/* RenderStats (src/render/mod.rs:73-83). `project_ms` covers projection, the
 * per-splat depth sort and tile-pair emission; `sort_ms` the stable tile sort
 * and tile ranges; `raster_ms` the blend. Device times from HIP events. */
typedef struct sp_stats {
    uint64_t visible;   /* gaussians that survived culling */
    uint64_t overlaps;  /* (gaussian, 16x16 tile) pairs */
    double project_ms;
    double sort_ms;
    double raster_ms;
    double total_ms;
} sp_stats;

typedef struct sp_options {
    int32_t device;            /* HIP device ordinal */
    int32_t storage;           /* SP_STORE_* */
    uint64_t initial_pairs;    /* initial pair capacity (0 = automatic) */
    uint64_t max_device_bytes; /* 0 = unlimited; above it allocation fails with SP_E_NOMEM */
    int32_t timing;            /* 1 = fill the *_ms fields of sp_stats from HIP events; 0 = no stage
                                  times, and a single view replays a captured graph */
    int32_t reserved;
} sp_options;

typedef struct sp_handle sp_handle;

/* Defaults: device 0, f32 storage, automatic capacity, no limit, timing on. */
void sp_default_options(sp_options *opt);

/* Upload `scene` to the device and build the per-scene data (layout transform
 * to structure-of-arrays, the bounding-box radius k = sqrt(2 ln(o/ALPHA_MIN)),
 * as Prep::new does once per scene, src/render/cpu.rs:384-388).
 * Replaces render::create(Backend::Hip, scene) (src/render/mod.rs:115-121). */
int sp_create(const sp_scene *scene, const sp_options *opt, sp_handle **out);
void sp_destroy(sp_handle *h);

/* "hip" (Renderer::name, src/render/mod.rs:87). */
const char *sp_name(const sp_handle *h);

/* Renderer::render (src/render/mod.rs:89-96): render `view` at width x height
 * over `background` into the caller's row-major RGB8 buffer of
 * width*height*3 bytes. Zero width or height renders nothing and returns
 * zeroed stats; a side above SP_MAX_DIM returns SP_E_PIXEL_LIMIT. */
int sp_render(sp_handle *h, const sp_view *view, uint32_t width, uint32_t height,
              const float background[3], uint8_t *rgb_out, sp_stats *stats);

/* As sp_render, but `rgb_out_device` is device memory (width*height*3 bytes)
 * and no image is copied to the host. Returns when the frame is complete.
 * The handle works on its own stream: work other streams queued on the buffer
 * must be complete before the call (as for every device buffer passed in). */
int sp_render_device(sp_handle *h, const sp_view *view, uint32_t width, uint32_t height,
                     const float background[3], uint8_t *rgb_out_device, sp_stats *stats);

/* Bytes of the static per-frame scratch of sp_render_device_static, reserved
 * on every device the library's code object loads on. */
#define SP_STATIC_SCRATCH_BYTES (128ull << 20)

/* sp_render_device without any device allocation. No counterpart in the
 * reference: for callers that cannot allocate (code between HIP graph
 * captures, tight serving loops, harnesses that forbid hipMalloc). The
 * per-frame buffers (records, sort keys, tile pairs, tables, tile state) are
 * pieces of one static region of SP_STATIC_SCRATCH_BYTES per device instead
 * of the handle's growable buffers; the scene stays the handle's (uploaded by
 * sp_create). The per-splat arrays take 80 bytes per splat and, with the
 * front-to-back split on, a 1920x1080 frame's tile state 34 MB; the tile
 * pairs of a phase get what is left, 16 bytes each. A frame that does not
 * fit returns SP_E_CAPACITY (the output's contents are then unspecified) and
 * the handle stays usable. Images and stats equal sp_render_device's. Every
 * call on a GPU shares that GPU's region: calls that overlap on one GPU (from
 * several threads or handles) race; each call returns after its frame is
 * complete, so calls in sequence are safe. */
int sp_render_device_static(sp_handle *h, const sp_view *view, uint32_t width, uint32_t height,
                            const float background[3], uint8_t *rgb_out_device, sp_stats *stats);

/* Last error message of the calling thread ("" if none). */
const char *sp_last_error(void);

/* ---------------------------------------------------------------------------
 * Terminal output on the device (the viewer's path: src/app.rs:190-230).
 * --------------------------------------------------------------------------- */

#define SP_TERM_TRUECOLOR 0 /* 6 bytes per cell: foreground r,g,b then background r,g,b */
#define SP_TERM_ANSI256 1   /* 2 bytes per cell: foreground, background xterm-256 index */

/* term::downsample (src/term.rs:75-101) on device frames: `src` (width*height*3)
 * box-filtered by an integer factor (0 acts as 1) into `dst`
 * ((width/factor)*(height/factor)*3), each channel the truncated mean of its
 * block. Returns when `dst` is complete. */
int sp_downsample_device(sp_handle *h, const uint8_t *src, uint32_t width, uint32_t height, uint32_t factor,
                         uint8_t *dst);

/* The viewer's frame as terminal cells: render the view at
 * (cols*factor) x (2*rows*factor), downsample by `factor`, and write the
 * half-block cells of the cols x (2*rows) result (HalfBlockImage,
 * src/term.rs:103-128; rgb_to_ansi256, src/term.rs:48-72) to the host,
 * row-major: cell (cx, cy) has display pixel (cx, 2cy) as its foreground and
 * (cx, 2cy+1) as its background, in the `mode` layout above. Only the cells
 * cross PCIe. A rendered side above SP_MAX_DIM returns SP_E_PIXEL_LIMIT. */
int sp_render_cells(sp_handle *h, const sp_view *view, uint32_t cols, uint32_t rows, uint32_t factor,
                    const float background[3], int32_t mode, uint8_t *cells_out, sp_stats *stats);

/* ---------------------------------------------------------------------------
 * Inspection entry points (the oracle comparison surface). They have no direct
 * counterpart in the reference's public API; they expose the intermediate
 * arrays of CpuTiled::render (src/render/cpu.rs:131-200) in canonical form.
 * ------------------------------------------------------------------------- */

/* Per-splat projection of `view`: for every scene index i, valid[i] = 1 when
 * the splat survives culling (render::project, src/render/mod.rs:201-272) and
 * then proj[i*12 .. i*12+12] = x, y, conic a, b, c, r, g, b, opacity, depth,
 * ext_x, ext_y, and bbox[i*4 .. +4] = the clamped pixel box x0, y0, x1, y1
 * (src/render/cpu.rs:96-102). Arrays are n long. */
int sp_debug_project(sp_handle *h, const sp_view *view, uint32_t width, uint32_t height,
                     uint8_t *valid, float *proj, uint16_t *bbox);

/* Sorted tile pairs of `view` in blend order: keys[p] = tile << 32 | depth
 * bits, vals[p] = original scene index, tile_start[0..tiles] = bucket bounds
 * (tiles = ceil(w/16)*ceil(h/16)). `cap` is the length of keys/vals; *n_pairs
 * receives the total even when it exceeds cap (then SP_E_CAPACITY). */
int sp_debug_pairs(sp_handle *h, const sp_view *view, uint32_t width, uint32_t height,
                   uint64_t *keys, uint32_t *vals, uint32_t *tile_start, uint64_t cap,
                   uint64_t *n_pairs);

/* Both tile lists of a real front-to-back split render of `view` (the
 * handle's tuning): first_vals / first_start[0..tiles] = the first phase's
 * lists (scene indices in blend order) over every tile, done[t] = 1 when the
 * first blend finished tile t, second_vals / second_start = the continuation
 * lists of the tiles not done. A tile's first list followed by its second
 * list is CpuTiled's bucket of that tile after sort_hi32 + fix_ties
 * (src/render/cpu.rs:149-214). With the split off (split_shift 0) the first
 * lists are the whole lists and the second ones are empty. `cap` is the
 * length of each value buffer; *n_first / *n_second receive the totals even
 * when one exceeds cap (then SP_E_CAPACITY). */
int sp_debug_split_pairs(sp_handle *h, const sp_view *view, uint32_t width, uint32_t height, uint32_t *first_vals,
                         uint32_t *first_start, uint32_t *done, uint32_t *second_vals, uint32_t *second_start,
                         uint64_t cap, uint64_t *n_first, uint64_t *n_second);

/* Reach bits of the blend for one projected splat: for every 16x16 tile of a
 * width x height frame (row-major), bit q is set unless the splat provably
 * blends no pixel of the tile's 8x8 quadrant (q & 1, q >> 1), i.e. unless
 * alpha < ALPHA_MIN there under raster_tile's f32 arithmetic
 * (src/render/cpu.rs:655-686). `proj` and `bbox` are one splat's
 * sp_debug_project fields. Host code only (the same source the blend runs on
 * the device); needs no device. */
int sp_debug_reach(const float *proj, const uint16_t *bbox, uint32_t width, uint32_t height, uint8_t *masks);

/* Work counters of the handle's last rendered view (no reference counterpart;
 * the roofline numerators of the benchmark). */
typedef struct sp_frame_info {
    uint32_t width, height;
    uint64_t tiles;          /* 16x16 tiles of the frame */
    uint64_t visible;        /* = sp_stats.visible */
    uint64_t pairs;          /* = sp_stats.overlaps */
    uint64_t pairs_staged;   /* pair records the blend gathered before its tiles saturated */
    uint64_t pair_capacity;  /* current pair buffer capacity */
    uint64_t device_bytes;   /* device memory held by the handle */
    uint64_t depth_key_bits; /* key width of the last depth sort: 27 (depths bounded near the views' near planes) or 32 */
} sp_frame_info;

int sp_last_frame_info(const sp_handle *h, sp_frame_info *info);

/* ---------------------------------------------------------------------------
 * Batched multi-view rendering (no reference counterpart; a new form of the
 * same Renderer::render contract for many cameras at once).
 * ------------------------------------------------------------------------- */

/* Render n_views views at their own resolutions (wh[2*v], wh[2*v+1]) in one
 * captured HIP graph. outs[v] receives width*height*3 bytes: host memory when
 * outs_on_device == 0, device memory otherwise. stats may be NULL or n_views
 * long. Images are identical to n_views calls of sp_render. */
int sp_render_batch(sp_handle *h, const sp_view *views, const uint32_t *wh, uint32_t n_views,
                    const float background[3], uint8_t *const *outs, int32_t outs_on_device,
                    sp_stats *stats);

/* ---------------------------------------------------------------------------
 * Tuning (no reference counterpart). The tunables change only the speed;
 * every setting renders the same bytes.
 * ------------------------------------------------------------------------- */
typedef struct sp_tuning {
    int32_t blend_batch;   /* records staged through LDS per blend step: 64, 128 or 256 */
    uint32_t split_shift;  /* front-to-back split: the first phase takes each view's nearest
                              max(split_kmin, visible >> split_shift) splats; 0 = no split */
    uint32_t split_kmin;
    uint32_t reserved;
} sp_tuning;

int sp_get_tuning(const sp_handle *h, sp_tuning *t);
int sp_set_tuning(sp_handle *h, const sp_tuning *t);

/* Autotuner: times candidate tunings on this batch of views (captured batch
 * replays into handle-owned scratch, median of `reps` after a warm-up,
 * coordinate descent over split_shift, blend_batch, split_kmin), keeps the
 * fastest on the handle and reports it with its batch time in milliseconds. */
int sp_autotune(sp_handle *h, const sp_view *views, const uint32_t *wh, uint32_t n_views,
                const float background[3], int32_t reps, sp_tuning *best, double *best_ms);

/* ---------------------------------------------------------------------------
 * Differentiable rendering (no reference counterpart: splat-tui renders only;
 * this is the adjoint of the same forward). A handle made by sp_diff_create
 * holds no scene: each sp_diff_forward reads raw, unactivated parameters from
 * the caller's device memory, activates them on the device (exp of the
 * log-scales, sigmoid of the logits, normalised quaternions, covariance,
 * k = sqrt(2 ln(o/ALPHA_MIN))) and renders float images C + T*bg (the value
 * to_u8 would quantise, src/render/mod.rs:282-285) with the `cpu` backend's
 * blend rules. sp_diff_backward then turns dL/dimage into dL/dparameters.
 * ------------------------------------------------------------------------- */

/* Device pointers, row-major per gaussian (the layout of contiguous
 * [n][k] float tensors). */
typedef struct sp_params {
    const float *means;          /* [n][3] */
    const float *log_scales;     /* [n][3] */
    const float *quats;          /* [n][4] x,y,z,w, any non-zero length */
    const float *opacity_logits; /* [n] */
    const float *colors;         /* [n][3] base colour (SH DC term) */
    const float *sh_rest;        /* [n][(deg+1)^2-1][3], NULL for degree 0 */
} sp_params;

/* Device pointers receiving dL/dparameter, same shapes as sp_params (overwritten). */
typedef struct sp_param_grads {
    float *means, *log_scales, *quats, *opacity_logits, *colors, *sh_rest;
} sp_param_grads;

int sp_diff_create(uint64_t n, uint32_t sh_degree, const sp_options *opt, sp_handle **out);

/* Render n_views views into images[v] (device, height*width*3 floats,
 * row-major RGB). The handle keeps what the backward needs; `params` must
 * stay valid and unchanged until the matching sp_diff_backward. */
int sp_diff_forward(sp_handle *h, const sp_params *params, const sp_view *views, const uint32_t *wh,
                    uint32_t n_views, const float background[3], float *const *images, sp_stats *stats);

/* Gradients of a loss L given dL/dimage (grad_images[v], device, same layout
 * as the forward's images) for the handle's last sp_diff_forward. Opacity
 * thresholds, the bounding boxes and the depth order are treated as
 * constants (they are piecewise constant in the parameters). The reduction
 * over pixels and pairs uses no floating-point atomics: results are
 * bit-reproducible. backward_ms (optional) receives the wall time. */
int sp_diff_backward(sp_handle *h, const sp_params *params, const float *const *grad_images,
                     const sp_param_grads *grads, double *backward_ms);

#ifdef __cplusplus
}
#endif

#endif /* SPLAT_HIP_H */
