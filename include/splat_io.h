/*
 * splat_io.h — scene file readers and writers (libsplat_scene.so).
 *
 * The two capture formats splat-tui loads, parsed on the host into the
 * activated structure of arrays that sp_create uploads:
 *
 *   PLY    the 3DGS trainer layout (x y z [nx ny nz] f_dc_0..2 [f_rest_*]
 *          opacity scale_0..2 rot_0..3) in ascii, binary_little_endian or
 *          binary_big_endian, any scalar type; plain coloured point clouds
 *          (red green blue) too.
 *   .splat 32-byte records: position 3 x f32, linear scale 3 x f32, RGBA u8,
 *          rotation (w, x, y, z) as u8 (q * 128 + 128).
 *
 * Activations (exp of log-scales, sigmoid of opacity logits) use the host
 * libm in f32, as the reference's Rust f32 methods do on Linux, so the arrays
 * equal the reference loader's bit for bit. Binary bodies are decoded by a
 * pool of host threads; files are never trusted for allocation sizes (the
 * body must hold every vertex the header declares before anything of that
 * size is allocated).
 *
 * Reference interfaces (splat-tui, file:line):
 *   sp_ply_parse     ply::parse (src/ply.rs:107-379: header, columns, body)
 *   sp_ply_write     ply::write (src/ply.rs:382-431)
 *   sp_splat_parse   splat_format::parse (src/splat_format.rs:18-44)
 *   sp_splat_write   splat_format::write (src/splat_format.rs:46-62)
 *   sp_scene_load    Scene::load (src/scene.rs:97-118): dispatch by
 *                    extension, finalize, reject empty scenes
 *   sp_io_finalize   Scene::finalize (src/scene.rs:62-94)
 */
#ifndef SPLAT_IO_H
#define SPLAT_IO_H

#include <stdint.h>

#include "splat_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A parsed scene; owns its arrays. */
typedef struct sp_io_scene sp_io_scene;

/* Views of a parsed scene's arrays (or, for the writers, of the caller's). */
typedef struct sp_io_arrays {
    uint64_t n;           /* gaussians */
    uint32_t sh_degree;   /* 0..3 */
    uint32_t rest;        /* SH rest coefficients per gaussian: (deg+1)^2 - 1 */
    float *positions;     /* [n][3] */
    float *scales;        /* [n][3] activated (linear) */
    float *rotations;     /* [n][4] quaternion (x, y, z, w); unit after finalize */
    float *opacities;     /* [n] activated */
    float *colors;        /* [n][3] base colour (SH_C0 * f_dc + 0.5) */
    float *sh_rest;       /* [n][rest][3], interleaved RGB per coefficient; NULL when rest == 0 */
    float *cov3d;         /* [n][6] after finalize, else NULL */
} sp_io_arrays;

/* Parse a PLY file held in memory. threads = 0: one per hardware thread.
 * Malformed or unsupported files return SP_E_FORMAT (message in
 * sp_io_last_error). The scene is not finalized (no cov3d). */
int sp_ply_parse(const uint8_t *data, uint64_t len, uint32_t threads, sp_io_scene **out);

/* Parse a .splat file held in memory (length a multiple of 32, else SP_E_FORMAT). */
int sp_splat_parse(const uint8_t *data, uint64_t len, uint32_t threads, sp_io_scene **out);

/* Scene::load: read `path`, parse by extension (.splat, .ply, any case; a file
 * starting with "ply" otherwise), finalize, and reject an empty scene
 * (SP_E_FORMAT). A file that cannot be read is SP_E_IO. */
int sp_scene_load(const char *path, uint32_t threads, sp_io_scene **out);

/* Scene::finalize on a parsed scene: unit quaternions (identity when not
 * finite or ~0), opacities clamped to [0,1] (non-finite -> 0), cov3d. */
int sp_io_finalize(sp_io_scene *scene);

/* Array views of a parsed scene (valid until sp_io_free). */
void sp_io_arrays_of(sp_io_scene *scene, sp_io_arrays *out);

void sp_io_free(sp_io_scene *scene);

/* Message of the last failed call on this thread ("" if none). */
const char *sp_io_last_error(void);

/* ply::write: the standard 3DGS binary little-endian layout. *written receives
 * the file size; when cap is smaller nothing is written and SP_E_CAPACITY is
 * returned (call with cap = 0 to size the buffer). */
int sp_ply_write(const sp_io_arrays *scene, uint8_t *out, uint64_t cap, uint64_t *written);

/* splat_format::write: 32 bytes per gaussian, same capacity contract. */
int sp_splat_write(const sp_io_arrays *scene, uint8_t *out, uint64_t cap, uint64_t *written);

#ifdef __cplusplus
}
#endif

#endif /* SPLAT_IO_H */
