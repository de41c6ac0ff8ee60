/*
 * splat_scene.h — host-side scene and camera helpers (libsplat_scene.so).
 *
 * These are the layers the rasterizer consumes but does not own: scene
 * finalisation (activations, covariance), the robust scene bounds, the orbit
 * camera, and the seeded synthetic workloads the benchmarks render. All
 * arithmetic is f32 in the operation order of the reference (glam 0.29 on
 * x86-64, no fused multiply-add), so results are reproducible on every host.
 *
 * Reference interfaces (splat-tui, file:line):
 *   sp_covariance        scene::covariance (src/scene.rs:170-178)
 *   sp_finalize          Scene::finalize (src/scene.rs:62-94)
 *   sp_bounds            Scene::bounds (src/scene.rs:123-166)
 *   sp_orbit_*           camera::OrbitCamera (src/camera.rs:69-159)
 *   sp_rng_*             gen::Rng, SplitMix64 + Box-Muller (src/gen.rs:22-54)
 *   sp_psnr              bench::psnr (src/bench.rs:153-170)
 */
#ifndef SPLAT_SCENE_H
#define SPLAT_SCENE_H

#include <stdint.h>

#include "splat_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Upper-triangle covariance R S S^T R^T for n splats: scales [n][3], quats [n][4] (x,y,z,w). */
void sp_covariance(const float *scales, const float *quats, uint64_t n, float *cov_out);

/* Scene::finalize: normalise quats in place (identity when not finite or ~0),
 * clamp opacities to [0,1] in place (non-finite -> 0), write cov3d [n][6]. */
void sp_finalize(uint64_t n, const float *scales, float *quats, float *opacities, float *cov_out);

/* Scene::bounds: per-axis median centre and min(p90, 2*p50) radius of a
 * (sub-sampled) point cloud. Empty or all-non-finite input gives (0, 1). */
void sp_bounds(const float *positions, uint64_t n, float center_out[3], float *radius_out);

/* Up axes of camera::UpAxis. */
#define SP_UP_POS_Y 0
#define SP_UP_NEG_Y 1
#define SP_UP_POS_Z 2
#define SP_UP_NEG_Z 3

typedef struct sp_orbit {
    float target[3];
    float distance;
    float yaw;
    float pitch;
    float fov_y;
    int32_t up;
    float scene_radius;
} sp_orbit;

void sp_orbit_init(sp_orbit *cam, const float center[3], float radius, int32_t up);
void sp_orbit_rotate(sp_orbit *cam, float dyaw, float dpitch);
void sp_orbit_zoom(sp_orbit *cam, float factor);
void sp_orbit_view(const sp_orbit *cam, sp_view *out);

/* gen::Rng. The state is one u64. */
uint64_t sp_rng_next_u64(uint64_t *state);
float sp_rng_f(uint64_t *state);      /* uniform [0,1) */
float sp_rng_normal(uint64_t *state); /* standard normal */

/*
 * Seeded random scene (the 1M / 3M / 500k workloads): means ~ N(0,I)
 * resampled until max|x_i| < clip, log-scales ~ U[ls_lo, ls_hi] per axis,
 * uniformly random unit quaternions, opacity = sigmoid(N(0, op_sigma)),
 * base colour ~ N(dc_mean, dc_sigma), SH rest coefficients ~ N(0, sh_sigma).
 * Arrays are caller-allocated; sh_rest may be NULL when sh_degree == 0.
 * log_scales / opacity_logits receive the pre-activation parameters.
 */
typedef struct sp_random_spec {
    uint64_t n;
    uint64_t seed;
    uint32_t sh_degree;
    float clip;
    float ls_lo, ls_hi;
    float op_sigma;
    float dc_mean, dc_sigma;
    float sh_sigma;
} sp_random_spec;

void sp_synth_random(const sp_random_spec *spec, float *positions, float *log_scales,
                     float *scales, float *quats, float *opacity_logits, float *opacities,
                     float *colors, float *sh_rest);

/* RGB lattice of nx*ny*nz isotropic gaussians on the unit cube:
 * position (i/(nx-1), j/(ny-1), k/(nz-1)) = colour, given scale and opacity. */
void sp_synth_lattice(uint32_t nx, uint32_t ny, uint32_t nz, float scale, float opacity,
                      float *positions, float *scales, float *quats, float *opacities,
                      float *colors);

/* f32 -> IEEE binary16 / bfloat16, round to nearest even (the conversions the
 * reduced-precision scene storage uses). */
void sp_to_f16(const float *in, uint64_t n, uint16_t *out);
void sp_to_bf16(const float *in, uint64_t n, uint16_t *out);

/* bench::psnr over two RGB8 images of n_pixels pixels (inf when identical, 0 when empty). */
double sp_psnr(const uint8_t *a, const uint8_t *b, uint64_t n_pixels);

#ifdef __cplusplus
}
#endif

#endif /* SPLAT_SCENE_H */
