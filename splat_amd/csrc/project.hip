// (a) EWA projection + SH colour (degrees 0-3) + per-view compaction of the
// visible splats, for every view of a batch in one launch.
//
// One thread per splat. The thread reads its splat once — structure-of-arrays
// loads (each field of splat i is a 4-byte load at array_base + i, so a wave
// reads 256 contiguous bytes per field) kept in registers, SH included — and
// projects it into every view, PROJECT_VCHUNK views per round. Each (view,
// splat) depth key goes to a full-length array, with a sentinel for the culled
// ones: the depth sort's first pass skips those, so no compaction pass runs.
//
// The arithmetic is render::project (src/render/mod.rs:201-272) and
// sh::eval_rest (src/sh.rs:39-64) in the reference's operation order, with
// contraction disabled, so every field is bit-identical to the CPU backend.
#include "sp_launch.h"
#include "sp_splat.h"

namespace sp {
namespace {

// Projects every splat into every view. Writes the records of visible
// splats with their pixel boxes also in sbox[v*n + i] (8 bytes, so that the
// depth-ordered box gather reads a compact array) and full[v*n + i] = depth
// bits (or FULL_CULLED).
// COLOUR: evaluate each visible splat's colour here (from the colour
// columns); otherwise the records' colours are left to the emission
// (emit_colour), which evaluates only those of the splats a blend reads.
template <int DEG, int STORE, bool COLOUR>
__global__ __launch_bounds__(PROJECT_THREADS) void k_project(DevScene sc, const ViewParams *__restrict__ views,
                                                             uint32_t nviews, uint4 *__restrict__ rec,
                                                             uint2 *__restrict__ sbox,
                                                             uint32_t *__restrict__ full,
                                                             float *__restrict__ dbg,
                                                             uint16_t *__restrict__ dbg_bbox,
                                                             uint8_t *__restrict__ dbg_valid,
                                                             uint32_t *__restrict__ counters) {
    const uint32_t i = blockIdx.x * PROJECT_THREADS + threadIdx.x;
    if (i >= sc.n) return;  // no barriers follow
    const uint64_t N = sc.n;
    Splat<DEG> s;
    load_splat<DEG, STORE, COLOUR>(sc, i, s);
    for (uint32_t v0 = 0; v0 < nviews; v0 += PROJECT_VCHUNK) {
#pragma unroll
        for (int c = 0; c < PROJECT_VCHUNK; ++c) {
            const uint32_t v = v0 + c;
            if (v < nviews) {
                float *g = (dbg && v == 0) ? dbg + 12 * (uint64_t)i : nullptr;
                uint16_t *gb = (dbg && v == 0) ? dbg_bbox + 4 * (uint64_t)i : nullptr;
                float depth = 0.0f;
                const auto colour = [&](float d0, float d1, float d2, float col[3]) {
                    if constexpr (COLOUR) {
                        splat_colour<DEG>(s, d0, d1, d2, col);
                    } else {
                        col[0] = col[1] = col[2] = 0.0f;  // the emission's (k_debug_colour's for dbg)
                    }
                };
                const bool vis = project_view<DEG, STORE != SP_STORE_F32>(
                    s, views[v].ctx, rec + 3 * ((uint64_t)v * N + i), &depth, g, gb, sbox + (uint64_t)v * N + i,
                    colour);
                // Visible depths lie above the near plane, so their bits exceed
                // key_base (the near plane's bits, or 0).
                const uint32_t kb = views[v].ctx.key_base, key = __float_as_uint(depth) - kb;
                full[(uint64_t)v * N + i] = vis ? key : FULL_CULLED;
                // The host's depth bound is conservative; should a key not fit
                // the planned width anyway, the flag makes it render the frame
                // again with full keys.
                if (vis && kb != 0u && (key >> DEPTH_KEY_BITS) != 0u && counters)
                    atomicOr(&counters[CNT_DEPTH_RANGE], 1u);
                if (dbg_valid && v == 0) dbg_valid[i] = vis ? 1 : 0;
            }
        }
    }
}

// The debug output's colours of a stored scene projected without them (view
// 0, visible splats).
__global__ __launch_bounds__(PROJECT_THREADS) void k_debug_colour(DevScene sc, const ViewParams *__restrict__ views,
                                                                  float *__restrict__ dbg,
                                                                  const uint8_t *__restrict__ valid) {
    const uint32_t i = blockIdx.x * PROJECT_THREADS + threadIdx.x;
    if (i >= sc.n || !valid[i]) return;
    float col[3];
    row_colour(sc, i, views[0].ctx.eye, col);
#pragma unroll
    for (int c = 0; c < 3; ++c) dbg[12 * (uint64_t)i + 5 + c] = col[c];
}

struct ProjectOut {
    uint4 *rec;
    uint2 *sbox;
    uint32_t *full;
    float *dbg;
    uint16_t *dbg_bbox;
    uint8_t *dbg_valid;
    uint32_t *counters;
};

template <int DEG, int STORE, bool COLOUR>
void launch_deg(hipStream_t s, uint32_t grid, const DevScene &sc, const ViewParams *views, uint32_t nviews,
                const ProjectOut &o) {
    hipLaunchKernelGGL((k_project<DEG, STORE, COLOUR>), dim3(grid), dim3(PROJECT_THREADS), 0, s, sc, views, nviews,
                       o.rec, o.sbox, o.full, o.dbg, o.dbg_bbox, o.dbg_valid, o.counters);
}

template <int STORE, bool COLOUR>
void launch_store(hipStream_t s, uint32_t grid, const DevScene &sc, const ViewParams *views, uint32_t nviews,
                  const ProjectOut &o) {
    switch (sc.deg) {
    case 0: launch_deg<0, STORE, COLOUR>(s, grid, sc, views, nviews, o); break;
    case 1: launch_deg<1, STORE, COLOUR>(s, grid, sc, views, nviews, o); break;
    case 2: launch_deg<2, STORE, COLOUR>(s, grid, sc, views, nviews, o); break;
    default: launch_deg<3, STORE, COLOUR>(s, grid, sc, views, nviews, o); break;
    }
}

template <int STORE>
void launch_stored(hipStream_t s, uint32_t grid, const DevScene &sc, const ViewParams *views, uint32_t nviews,
                   const ProjectOut &o, bool colour) {
    if (colour) launch_store<STORE, true>(s, grid, sc, views, nviews, o);
    else launch_store<STORE, false>(s, grid, sc, views, nviews, o);
}

}  // namespace

uint32_t project_blocks(uint32_t n) { return n ? (n + PROJECT_THREADS - 1) / PROJECT_THREADS : 1; }

hipError_t launch_project(hipStream_t s, const DevScene &sc, const ViewParams *views, uint32_t nviews, uint4 *rec,
                          uint2 *sbox, uint32_t *full, float *dbg, uint16_t *dbg_bbox, uint8_t *dbg_valid,
                          uint32_t *counters, bool colour) {
    if (sc.n == 0 || nviews == 0) return hipSuccess;
    const uint32_t grid = project_blocks(sc.n);
    const ProjectOut o{rec, sbox, full, dbg, dbg_bbox, dbg_valid, counters};
    switch (sc.storage) {
    case SP_STORE_F16: launch_stored<SP_STORE_F16>(s, grid, sc, views, nviews, o, colour); break;
    case SP_STORE_BF16: launch_stored<SP_STORE_BF16>(s, grid, sc, views, nviews, o, colour); break;
    case STORE_PARAMS: launch_store<STORE_PARAMS, true>(s, grid, sc, views, nviews, o); break;
    default: launch_stored<SP_STORE_F32>(s, grid, sc, views, nviews, o, colour); break;
    }
    if (dbg && dbg_valid && !colour && sc.storage != STORE_PARAMS) {
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_debug_colour, dim3(grid), dim3(PROJECT_THREADS), 0, s, sc, views, dbg, dbg_valid);
    }
    return hipGetLastError();
}

}  // namespace sp
