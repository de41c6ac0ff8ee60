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
template <int DEG, int STORE>
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
    load_splat<DEG, STORE>(sc, i, s);
    for (uint32_t v0 = 0; v0 < nviews; v0 += PROJECT_VCHUNK) {
#pragma unroll
        for (int c = 0; c < PROJECT_VCHUNK; ++c) {
            const uint32_t v = v0 + c;
            if (v < nviews) {
                float *g = (dbg && v == 0) ? dbg + 12 * (uint64_t)i : nullptr;
                uint16_t *gb = (dbg && v == 0) ? dbg_bbox + 4 * (uint64_t)i : nullptr;
                float depth = 0.0f;
                const bool vis = project_view<DEG, STORE != SP_STORE_F32>(
                    s, views[v].ctx, rec + 3 * ((uint64_t)v * N + i), &depth, g, gb, sbox + (uint64_t)v * N + i);
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

struct ProjectOut {
    uint4 *rec;
    uint2 *sbox;
    uint32_t *full;
    float *dbg;
    uint16_t *dbg_bbox;
    uint8_t *dbg_valid;
    uint32_t *counters;
};

template <int DEG, int STORE>
void launch_deg(hipStream_t s, uint32_t grid, const DevScene &sc, const ViewParams *views, uint32_t nviews,
                const ProjectOut &o) {
    hipLaunchKernelGGL((k_project<DEG, STORE>), dim3(grid), dim3(PROJECT_THREADS), 0, s, sc, views, nviews, o.rec,
                       o.sbox, o.full, o.dbg, o.dbg_bbox, o.dbg_valid, o.counters);
}

template <int STORE>
void launch_store(hipStream_t s, uint32_t grid, const DevScene &sc, const ViewParams *views, uint32_t nviews,
                  const ProjectOut &o) {
    switch (sc.deg) {
    case 0: launch_deg<0, STORE>(s, grid, sc, views, nviews, o); break;
    case 1: launch_deg<1, STORE>(s, grid, sc, views, nviews, o); break;
    case 2: launch_deg<2, STORE>(s, grid, sc, views, nviews, o); break;
    default: launch_deg<3, STORE>(s, grid, sc, views, nviews, o); break;
    }
}

}  // namespace

uint32_t project_blocks(uint32_t n) { return n ? (n + PROJECT_THREADS - 1) / PROJECT_THREADS : 1; }

hipError_t launch_project(hipStream_t s, const DevScene &sc, const ViewParams *views, uint32_t nviews, uint4 *rec,
                          uint2 *sbox, uint32_t *full, float *dbg, uint16_t *dbg_bbox, uint8_t *dbg_valid,
                          uint32_t *counters) {
    if (sc.n == 0 || nviews == 0) return hipSuccess;
    const uint32_t grid = project_blocks(sc.n);
    const ProjectOut o{rec, sbox, full, dbg, dbg_bbox, dbg_valid, counters};
    switch (sc.storage) {
    case SP_STORE_F16: launch_store<SP_STORE_F16>(s, grid, sc, views, nviews, o); break;
    case SP_STORE_BF16: launch_store<SP_STORE_BF16>(s, grid, sc, views, nviews, o); break;
    case STORE_PARAMS: launch_store<STORE_PARAMS>(s, grid, sc, views, nviews, o); break;
    default: launch_store<SP_STORE_F32>(s, grid, sc, views, nviews, o); break;
    }
    return hipGetLastError();
}

}  // namespace sp
