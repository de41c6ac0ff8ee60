// (a) EWA projection + SH colour (degrees 0-3) + per-view compaction of the
// visible splats, for every view of a batch in one launch.
//
// One thread per splat. The thread reads its splat once — structure-of-arrays
// loads (each field of splat i is a 4-byte load at array_base + i, so a wave
// reads 256 contiguous bytes per field) kept in registers, SH included — and
// projects it into every view, PROJECT_VCHUNK views per round. Each view's
// visible splats are then compacted in scene-index order: per-block counts
// (wave ballots), one scan workgroup per view over its block counts, and a
// compaction pass that places each visible splat at its block offset plus
// its in-block ballot rank.
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
// depth-ordered box gather reads a compact array), full[v*n + i] = depth bits
// (or FULL_CULLED) and, per view, the block's visible count
// bcnt[v * gridDim.x + block].
template <int DEG, int STORE>
__global__ __launch_bounds__(PROJECT_THREADS) void k_project(DevScene sc, const ViewParams *__restrict__ views,
                                                             uint32_t nviews, uint4 *__restrict__ rec,
                                                             uint2 *__restrict__ sbox,
                                                             uint32_t *__restrict__ full,
                                                             uint32_t *__restrict__ bcnt,
                                                             float *__restrict__ dbg,
                                                             uint16_t *__restrict__ dbg_bbox,
                                                             uint8_t *__restrict__ dbg_valid) {
    constexpr int NW = PROJECT_THREADS / 64;
    __shared__ uint32_t s_wcnt[PROJECT_VCHUNK][NW];
    const uint32_t i = blockIdx.x * PROJECT_THREADS + threadIdx.x;
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint64_t N = sc.n;
    Splat<DEG> s;
    const bool have = i < sc.n;
    if (have) load_splat<DEG, STORE>(sc, i, s);
    for (uint32_t v0 = 0; v0 < nviews; v0 += PROJECT_VCHUNK) {
#pragma unroll
        for (int c = 0; c < PROJECT_VCHUNK; ++c) {
            const uint32_t v = v0 + c;
            bool vis = false;
            if (have && v < nviews) {
                float *g = (dbg && v == 0) ? dbg + 12 * (uint64_t)i : nullptr;
                uint16_t *gb = (dbg && v == 0) ? dbg_bbox + 4 * (uint64_t)i : nullptr;
                float depth = 0.0f;
                vis = project_view<DEG, STORE != SP_STORE_F32>(s, views[v].ctx, rec + 3 * ((uint64_t)v * N + i),
                                                                &depth, g, gb, sbox + (uint64_t)v * N + i);
                full[(uint64_t)v * N + i] = vis ? __float_as_uint(depth) : FULL_CULLED;
                if (dbg_valid && v == 0) dbg_valid[i] = vis ? 1 : 0;
            }
            const unsigned long long m = __ballot(vis);
            if (lane == 0) s_wcnt[c][w] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        if (threadIdx.x < PROJECT_VCHUNK && v0 + threadIdx.x < nviews) {
            uint32_t t = 0;
#pragma unroll
            for (int ww = 0; ww < NW; ++ww) t += s_wcnt[threadIdx.x][ww];
            bcnt[(uint64_t)(v0 + threadIdx.x) * gridDim.x + blockIdx.x] = t;
        }
        __syncthreads();
    }
}

// Per-view compaction of the visible splats in scene-index order: block b of
// view v reads full[v*n + b*256 ..] and writes (depth bits, scene index) at
// the view's scanned block offset plus the in-block ballot rank.
__global__ __launch_bounds__(PROJECT_THREADS) void k_compact(const uint32_t *__restrict__ full, uint32_t n,
                                                             uint32_t nblocks, const uint32_t *__restrict__ boff,
                                                             uint32_t *__restrict__ dkey,
                                                             uint32_t *__restrict__ dval) {
    constexpr int NW = PROJECT_THREADS / 64;
    __shared__ uint32_t s_w[NW];
    const uint32_t v = blockIdx.x / nblocks, b = blockIdx.x % nblocks;
    const uint32_t i = b * PROJECT_THREADS + threadIdx.x;
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint64_t vb = (uint64_t)v * n;
    const uint32_t key = i < n ? full[vb + i] : FULL_CULLED;
    const bool vis = key != FULL_CULLED;
    const unsigned long long m = __ballot(vis);
    const uint32_t r = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    if (lane == 0) s_w[w] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t wo = 0;
#pragma unroll
    for (int ww = 0; ww < NW; ++ww) wo += (uint32_t)ww < w ? s_w[ww] : 0u;
    if (vis) {
        const uint32_t pos = boff[(uint64_t)v * nblocks + b] + wo + r;
        if (pos < n) {  // always in range for consistent counts; guards the buffer
            dkey[vb + pos] = key;
            dval[vb + pos] = i;
        }
    }
}

template <int DEG, int STORE>
void launch_deg(hipStream_t s, uint32_t grid, const DevScene &sc, const ViewParams *views, uint32_t nviews,
                uint4 *rec, uint2 *sbox, uint32_t *full, uint32_t *bcnt, float *dbg, uint16_t *dbg_bbox,
                uint8_t *dbg_valid) {
    hipLaunchKernelGGL((k_project<DEG, STORE>), dim3(grid), dim3(PROJECT_THREADS), 0, s, sc, views, nviews, rec,
                       sbox, full, bcnt, dbg, dbg_bbox, dbg_valid);
}

template <int STORE>
void launch_store(hipStream_t s, uint32_t grid, const DevScene &sc, const ViewParams *views, uint32_t nviews,
                  uint4 *rec, uint2 *sbox, uint32_t *full, uint32_t *bcnt, float *dbg, uint16_t *dbg_bbox,
                  uint8_t *dbg_valid) {
    switch (sc.deg) {
    case 0: launch_deg<0, STORE>(s, grid, sc, views, nviews, rec, sbox, full, bcnt, dbg, dbg_bbox, dbg_valid); break;
    case 1: launch_deg<1, STORE>(s, grid, sc, views, nviews, rec, sbox, full, bcnt, dbg, dbg_bbox, dbg_valid); break;
    case 2: launch_deg<2, STORE>(s, grid, sc, views, nviews, rec, sbox, full, bcnt, dbg, dbg_bbox, dbg_valid); break;
    default: launch_deg<3, STORE>(s, grid, sc, views, nviews, rec, sbox, full, bcnt, dbg, dbg_bbox, dbg_valid); break;
    }
}

}  // namespace

uint32_t project_blocks(uint32_t n) { return n ? (n + PROJECT_THREADS - 1) / PROJECT_THREADS : 1; }

hipError_t launch_project(hipStream_t s, const DevScene &sc, const ViewParams *views, uint32_t nviews, uint4 *rec,
                          uint2 *sbox, uint32_t *full, uint32_t *bcnt, uint32_t *vis_count, uint32_t *dkey,
                          uint32_t *dval, float *dbg, uint16_t *dbg_bbox, uint8_t *dbg_valid) {
    const uint32_t grid = project_blocks(sc.n);
    switch (sc.storage) {
    case SP_STORE_F16:
        launch_store<SP_STORE_F16>(s, grid, sc, views, nviews, rec, sbox, full, bcnt, dbg, dbg_bbox, dbg_valid);
        break;
    case SP_STORE_BF16:
        launch_store<SP_STORE_BF16>(s, grid, sc, views, nviews, rec, sbox, full, bcnt, dbg, dbg_bbox, dbg_valid);
        break;
    case STORE_PARAMS:
        launch_store<STORE_PARAMS>(s, grid, sc, views, nviews, rec, sbox, full, bcnt, dbg, dbg_bbox, dbg_valid);
        break;
    default:
        launch_store<SP_STORE_F32>(s, grid, sc, views, nviews, rec, sbox, full, bcnt, dbg, dbg_bbox, dbg_valid);
        break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // Per view: exclusive scan of the block counts; row totals are the visible counts.
    e = launch_scan_rows(s, bcnt, grid, nviews, nullptr, vis_count);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_compact, dim3(grid * nviews), dim3(PROJECT_THREADS), 0, s, full, sc.n, grid, bcnt, dkey,
                       dval);
    return hipGetLastError();
}

}  // namespace sp
