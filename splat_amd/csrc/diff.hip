// Differentiation: the adjoint of the blend and of the projection.
//
// The differentiable forward renders float images (C + T*bg) with the same
// binned pipeline as the plain path (no front-to-back split). Its emission
// also numbers the tile pairs in emission order, view by view and splat by
// splat in depth order, so the pairs of one (view, splat) record are
// contiguous: each record's first slot (pbase), and the first slot and
// record of each depth-ordered splat (pslot, prid); a pair's slot e is
// pbase[record] plus the tile's row-major place in the record's tile
// rectangle. The blend keeps, per pixel, its final transmittance and the
// tile-list position after its last contribution. The backward is three
// passes, none of which uses a floating-point atomic, so gradients are
// deterministic:
//
//   1. k_blend_bwd: one workgroup per tile walks the tile's list back to
//      front from the last position any of its pixels used, recovering each
//      pixel's transmittance before a contribution as T_i = T_{i+1} / (1 - a_i)
//      and carrying G . S, S being the colour accumulated behind it. Per
//      pixel and record:
//        dL/dc     = G * a * T_i
//        dL/da     = G . (c * T_i - S / (1 - a))
//        dL/dop    = dL/da * exp(pw),   dL/dpw = dL/da * a   (0 where a is capped)
//        dL/dmean2 = dL/dpw * (A dx + B dy, B dx + C dy)
//        dL/dconic = dL/dpw * (-dx^2 / 2, -dx dy, -dy^2 / 2)
//      with G = dL/dpixel. The nine values are summed over the tile's 256
//      pixels in a fixed order (one nine-value halving reduction within a
//      wave, then the four waves);
//      the mean and conic terms are summed as moments of dL/dpw (dx, dy,
//      dx^2, dx dy, dy^2) and multiplied by the conic once per record. They are
//      stored per pair at its emission slot. Records are staged as the
//      forward blend stages them (per-column and per-row power terms, NaN
//      outside the box, reach bits per quadrant), so every mask decision is
//      the forward's; records whose falloff cannot reach a quadrant are never
//      visited by its wave (one ballot over the staged records' reach bits).
//   2. k_pair_reduce: sums each (view, splat) record's pairs. The pairs
//      this backward wrote are marked in a bit array (cleared per call), so
//      the per-pair gradients themselves are never cleared. A workgroup takes
//      256 consecutive depth-ordered splats of a view, whose pairs are one
//      contiguous run of emission slots; it stages the run through LDS with
//      coalesced reads, and each thread adds its own splat's slots in order.
//   3. k_project_bwd: one thread per splat chains the per-view 2D gradients
//      through the EWA projection, the covariance (scale, rotation), the
//      opacity sigmoid and the SH colour, and sums the views.
#include "sp_launch.h"
#include "sp_splat.h"
#include "sp_stage.h"

namespace sp {
namespace {

constexpr int BWD_BATCH = 64;  // records per LDS step of the adjoint blend
constexpr int PR_THREADS = 256;          // depth-ordered splats per workgroup of the pair reduction
constexpr uint32_t PR_SERIAL = 4;        // marked pairs a thread sums alone
constexpr uint32_t PR_SERIAL_WORDS = 4;  // mark words a thread scans alone
constexpr int PBWD_THREADS = 128;

// Totals over the 64 lanes of the NG2D = 9 per-lane values g[] at once, by
// recursive halving. Each step pairs lanes that hold the same subset of the
// values; each lane adds its partner's copy of the whole subset and keeps the
// half that its lane bit selects, so later steps carry fewer values (9, 5, 3,
// 2, then 1 per lane): 34 VALU against 54 for nine one-value reductions. The
// pairs are lane 15 - i (row_mirror; lane bit 3 selects), 7 - i
// (row_half_mirror; bit 2), i ^ 2 and i ^ 1 (quad_perm; bits 1, 0); each lane
// of a row of 16 then holds one value's sum over its row, and swaps across
// rows (permlane16 / permlane32) add the four rows. Lane i < 16 ends with the
// total of value totals9_index(i) (9 of the 16 lanes hold one). Every lane
// must be active; the order of additions is fixed.
template <int CTRL>
__device__ __forceinline__ float halve_pair(float lo, float hi, bool upper) {
    const float s_lo = lo + dpp_f<CTRL, 0xF>(lo), s_hi = hi + dpp_f<CTRL, 0xF>(hi);
    return upper ? s_hi : s_lo;
}
// Pair (lo, nothing): the upper lanes' result is left as the sum too. It
// stands for nothing there, and every later value computed from it stands for
// nothing as well (partner lanes hold the same subset in each slot), so it
// only ever reaches lanes that hold no total (totals9_index gives NG2D).
template <int CTRL>
__device__ __forceinline__ float halve_last(float lo) {
    return lo + dpp_f<CTRL, 0xF>(lo);
}
__device__ __forceinline__ float wave_totals9(const float (&g)[NG2D], uint32_t lane) {
    static_assert(NG2D == 9, "the halving steps are laid out for nine values");
    const bool b3 = lane & 8u, b2 = lane & 4u, b1 = lane & 2u, b0 = lane & 1u;
    float a[5];
#pragma unroll
    for (int j = 0; j < 4; ++j) a[j] = halve_pair<0x140>(g[j], g[j + 5], b3);
    a[4] = halve_last<0x140>(g[4]);
    float b[3];
#pragma unroll
    for (int k = 0; k < 2; ++k) b[k] = halve_pair<0x141>(a[k], a[k + 3], b2);
    b[2] = halve_last<0x141>(a[2]);
    const float c0 = halve_pair<0x4E>(b[0], b[2], b1), c1 = halve_last<0x4E>(b[1]);
    float d = halve_pair<0xB1>(c0, c1, b0);
    const auto r16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(d), __float_as_uint(d), false, false);
    d = __uint_as_float(r16[0]) + __uint_as_float(r16[1]);  // rows 0+1 and 2+3
    const auto r32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(d), __float_as_uint(d), false, false);
    return __uint_as_float(r32[0]) + __uint_as_float(r32[1]);  // all four rows
}
// Which of the nine values lane `lane` (< 16) holds after wave_totals9, or
// NG2D when it holds none: the slot chosen at each halving, read back.
__device__ __forceinline__ uint32_t totals9_index(uint32_t lane) {
    const uint32_t cs = lane & 1u;                    // of c0, c1
    const uint32_t bs = (lane & 2u) ? 2u + cs : cs;   // of b[0..2] (3: none)
    const uint32_t as = (lane & 4u) ? 3u + bs : bs;   // of a[0..4]
    const uint32_t vi = (lane & 8u) ? 5u + as : as;   // of g[0..8]
    return (bs < 3u && as < 5u && vi < (uint32_t)NG2D) ? vi : (uint32_t)NG2D;
}
// The lane that holds value c's total after wave_totals9 (the inverse of
// totals9_index over lanes 0-15).
__device__ __forceinline__ uint32_t totals9_lane(int c) {
    constexpr uint32_t lanes[NG2D] = {0, 1, 2, 4, 5, 8, 9, 10, 12};
    return lanes[c];
}

template <int BATCH>
__global__ __launch_bounds__(BLEND_THREADS) void k_blend_bwd(
    const ViewParams *__restrict__ views, uint32_t nviews, uint32_t total_tiles,
    const uint32_t *__restrict__ tile_start, const uint32_t *__restrict__ sval, const uint32_t *__restrict__ pbase,
    const uint4 *__restrict__ rec, const float *__restrict__ pix_T, const uint32_t *__restrict__ pix_last,
    float *__restrict__ gpair, uint64_t gstride, uint32_t *__restrict__ fresh, uint32_t total_sbs) {
    constexpr int P = 4;  // lanes staging each record
    static_assert(BATCH <= 64 && BATCH * P <= BLEND_THREADS, "one reach ballot per batch");
    __shared__ StagedRec s_st[BATCH];
    __shared__ uint8_t s_qb[BATCH][4];  // reach bit of each staged record per quadrant
    __shared__ float4 s_geo[BATCH];     // y, -a/2, -b, -c/2 of each staged record
    __shared__ uint32_t s_e[BATCH];
    __shared__ float s_part[4][NG2D][BATCH];
    __shared__ uint64_t s_wrote[4];  // the records each wave left a partial sum for
    __shared__ uint32_t s_wmax[4];
    uint32_t gt, vi;
    if (!blend_tile_of(views, nviews, total_sbs, blockIdx.x, &gt, &vi)) return;
    const ViewCtx &v = views[vi].ctx;
    const uint32_t t = gt - views[vi].tile_base;
    const uint32_t tx = t % v.tiles_x, ty = t / v.tiles_x;
    const uint32_t ox = tx * TILE, oy = ty * TILE;
    // Same pixel <-> thread mapping as the forward blend (8x8 quadrant per wave).
    const uint32_t w = threadIdx.x >> 6, lane_i = lane_id();
    const uint32_t qx0 = ox + 8u * (w & 1u), qy0 = oy + 8u * (w >> 1);
    const uint32_t px = qx0 + (lane_i & 7u), py = qy0 + (lane_i >> 3);
    const uint32_t lx = px - ox, ly = py - oy;
    const bool inside = px < v.w && py < v.h;
    // The value this lane writes after each nine-value wave reduction.
    const uint32_t sum_slot = lane_i < 16u ? totals9_index(lane_i) : (uint32_t)NG2D;
    const float pyc = ((float)oy + (float)ly) + 0.5f;
    const uint64_t g = (uint64_t)gt * BLEND_THREADS + threadIdx.x;
    float T = pix_T[g];
    const uint32_t last = pix_last[g];
    float G0 = 0.0f, G1 = 0.0f, G2 = 0.0f;
    if (inside) {
        const float *gi = views[vi].gimg + 3 * ((uint64_t)py * v.w + px);
        G0 = gi[0];
        G1 = gi[1];
        G2 = gi[2];
    }
    // D = G . S, S being the colour composited behind the current record
    // (T * bg behind the last one); dL/da needs S only through D.
    float D = T * ((G0 * v.bg[0] + G1 * v.bg[1]) + G2 * v.bg[2]);
    uint32_t m = last;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off));
    if (lane_i == 0) s_wmax[w] = m;
    __syncthreads();
    const uint32_t lim = max(max(s_wmax[0], s_wmax[1]), max(s_wmax[2], s_wmax[3]));
    const uint32_t start = tile_start[gt];
    for (uint32_t hi = lim; hi > 0;) {
        const uint32_t lo = hi > (uint32_t)BATCH ? hi - BATCH : 0u;
        const uint32_t nb = hi - lo;
        // The forward's staging (same terms, same reach bits), plus what the
        // gradients need: the record's y and conic, and the pair's emission
        // slot (the record's first slot plus this tile's row-major place in
        // the record's tile rectangle, as the emission numbered them).
        const uint32_t rr = threadIdx.x / P, part = threadIdx.x % P;
        if (rr < nb) {
            const uint32_t id = sval[start + lo + rr];
            const uint4 *r = rec + 3 * (uint64_t)id;
            const uint4 r0 = r[0], r1 = r[1], r2 = r[2];
            stage_part<P, false>(s_st[rr], s_qb[rr], r0, r1, r2, ox, oy, part);
            if (part == 0) {
                const uint32_t tx0 = (r2.y & 0xFFFFu) / TILE, ty0 = (r2.y >> 16) / TILE;
                const uint32_t rw = ((r2.z & 0xFFFFu) + TILE - 1u) / TILE - tx0;
                s_e[rr] = pbase[id] + (ty - ty0) * rw + (tx - tx0);
                s_geo[rr] = make_float4(__uint_as_float(r0.y), __uint_as_float(r0.z), __uint_as_float(r0.w),
                                        __uint_as_float(r1.x));
            }
        }
        __syncthreads();
        // Records that can reach this wave's quadrant, walked back to front.
        unsigned long long reach = lane_mask(lane_i < nb && s_qb[min(lane_i, (uint32_t)BATCH - 1u)][w] != 0u);
        unsigned long long wrote = 0ull;
        if (reach != 0ull) {
            // Software-pipelined over two register sets, as the forward blend
            // walks: a visit reads the next set record's terms from LDS before
            // its own arithmetic, and the sets alternate. The next record is
            // the highest set bit (s_flbit clamped: an empty mask gives record
            // 0 of the batch, an in-bounds read whose values go unused).
            const char *gbase = reinterpret_cast<const char *>(s_st);
            const char *pcol = gbase + lx * sizeof(float2);
            const char *prow = gbase + offsetof(StagedRec, row) + ly * sizeof(float2);
            // Lane-varying registers for the uniform offsets too, so each
            // read's address is one vector add.
            uint32_t orgb_off = (uint32_t)offsetof(StagedRec, orgb), geo_off = 0u;
            asm("" : "+v"(orgb_off), "+v"(geo_off));
            const char *porgb = gbase + orgb_off;
            const char *pgeo = reinterpret_cast<const char *>(s_geo) + geo_off;
            struct Terms {
                float2 cv, rv;  // the record's (dx, qx) at this column, (qy, -b dy) at this row
                float4 c4;      // opacity, colour
                float sy;       // the record's y
            };
            const auto load = [&](uint32_t k, Terms &q) {
                const uint32_t off = k * (uint32_t)sizeof(StagedRec);
                q.cv = *reinterpret_cast<const float2 *>(pcol + off);
                q.rv = *reinterpret_cast<const float2 *>(prow + off);
                q.c4 = *reinterpret_cast<const float4 *>(porgb + off);
                q.sy = *reinterpret_cast<const float *>(pgeo + k * (uint32_t)sizeof(float4));
            };
            const auto last_set = [](uint64_t mk) {
                uint32_t z;
                asm("s_flbit_i32_b64 %0, %1" : "=s"(z) : "s"(mk));
                return 63u - min(z, 63u);
            };
            // Visits record kc (terms q) after reading the terms of the next
            // set record kn into qn. Returns whether a set record is left
            // (then kn's bit is cleared).
            const auto visit = [&](const Terms &q, uint32_t kc, Terms &qn, uint32_t &kn) -> bool {
                kn = last_set(reach);
                load(kn, qn);
                // The arithmetic starts from values this empty asm passes on,
                // after the reads above (its memory clobber keeps them ahead).
                float cx = q.cv.x, cy = q.cv.y, rx = q.rv.x, ry = q.rv.y;
                asm volatile("" : "+v"(cx), "+v"(cy), "+v"(rx), "+v"(ry) : : "memory");
                // A lane takes part where the forward's pixel was still live
                // at this position (before its last contribution), the power
                // is <= 0 (a NaN term marks a pixel outside the box) and
                // alpha reaches ALPHA_MIN: the forward's decisions, from the
                // forward's arithmetic on the same staged terms.
                const uint64_t am = lane_mask(lo + kc < last);
                if (am != 0ull) {
                    const float pw = (cy + rx) + ry * cx;
                    int n;
                    const float rp = exp_neg_parts(pw, n);
                    const float raw = __builtin_amdgcn_ldexpf(rp * q.c4.x, n);
                    const float alpha = fminf(raw, ALPHA_MAX);
                    const uint64_t cm = am & lane_mask(pw <= 0.0f) & lane_mask(alpha >= ALPHA_MIN);
                    // No lane taking the record: no gradient, T and D unchanged.
                    if (cm != 0ull) {
                        const bool con = __builtin_amdgcn_inverse_ballot_w64(cm);
                        // Branch-free over the wave: a lane without a
                        // contribution gets zero weights (aT, gs), so its
                        // gradients are zero and its T and D stay as they are.
                        // The gradient arithmetic uses fused multiply-adds:
                        // only the decisions above must repeat the forward.
                        const float ex = __builtin_amdgcn_ldexpf(rp, n);
                        const float dx = cx, dy = pyc - q.sy;
                        const float om = 1.0f - alpha;
                        // One hardware reciprocal (1 ulp) for both divisions:
                        // gradients need no bit-exactness, and T's recovery
                        // drifts by ulps only.
                        const float rom = __builtin_amdgcn_rcpf(om);
                        const float Ti = T * rom;
                        const float aT = con ? alpha * Ti : 0.0f;
                        const float Gc = fmaf(G2, q.c4.w, fmaf(G1, q.c4.z, G0 * q.c4.y));
                        float gv[NG2D];
                        gv[6] = G0 * aT;
                        gv[7] = G1 * aT;
                        gv[8] = G2 * aT;
                        const float ga = fmaf(Ti, Gc, -(D * rom));
                        D = fmaf(Gc, aT, D);
                        T = con ? Ti : T;
                        // Opacity and falloff terms vanish where alpha is capped.
                        const float gs = (con && raw < ALPHA_MAX) ? ga : 0.0f;
                        gv[5] = gs * ex;
                        // Moments of dL/dpw over the pixels; the record's conic
                        // coefficients are applied once per record after the sums.
                        const float gpw = gs * alpha;
                        const float u = gpw * dx, wdy = gpw * dy;
                        gv[0] = u;
                        gv[1] = wdy;
                        gv[2] = u * dx;
                        gv[3] = u * dy;
                        gv[4] = wdy * dy;
                        const float sum = wave_totals9(gv, lane_i);  // wave-uniform: all 64 lanes take part
                        if (sum_slot < (uint32_t)NG2D) s_part[w][sum_slot][kc] = sum;
                        wrote |= 1ull << kc;
                    }
                }
                if (reach == 0ull) return false;
                asm("s_bitset0_b64 %0, %1" : "+s"(reach) : "s"(kn));
                return true;
            };
            uint32_t kA = last_set(reach), kB = 0;
            asm("s_bitset0_b64 %0, %1" : "+s"(reach) : "s"(kA));
            Terms qA, qB;
            load(kA, qA);
            while (visit(qA, kA, qB, kB) && visit(qB, kB, qA, kA)) {
            }
        }
        if (lane_i == 0) s_wrote[w] = wrote;
        __syncthreads();
        // A record that no pixel of the tile took has a zero gradient here: its
        // pair is left unwritten and unmarked, and the reduction counts it as
        // zero.
        const uint64_t w0 = s_wrote[0], w1 = s_wrote[1], w2 = s_wrote[2], w3 = s_wrote[3];
        if (threadIdx.x < nb && (((w0 | w1 | w2 | w3) >> threadIdx.x) & 1ull)) {
            const uint32_t k = threadIdx.x;
            const uint32_t e = s_e[k];
            // A wave that left no partial for this record contributes zero.
            const auto part_of = [&](uint64_t wm, int ww, int c) {
                return ((wm >> k) & 1ull) ? s_part[ww][c][k] : 0.0f;
            };
            float sum[NG2D];
#pragma unroll
            for (int c = 0; c < NG2D; ++c)
                sum[c] = (part_of(w0, 0, c) + part_of(w1, 1, c)) + (part_of(w2, 2, c) + part_of(w3, 3, c));
            // dL/dmean = gpw (A dx + B dy, B dx + C dy) and dL/dconic =
            // gpw (-dx^2/2, -dx dy, -dy^2/2), from the moment sums
            // (A = -2 na, B = -nb2, C = -2 nc: the conic a, b, c).
            const float4 geo = s_geo[k];
            const float A = -2.0f * geo.y, B = -geo.z, C = -2.0f * geo.w;
            const float U = sum[0], W = sum[1];
            sum[0] = A * U + B * W;
            sum[1] = B * U + C * W;
            sum[2] = -0.5f * sum[2];
            sum[3] = -sum[3];
            sum[4] = -0.5f * sum[4];
#pragma unroll
            for (int c = 0; c < NG2D; ++c) gpair[c * gstride + e] = sum[c];
            atomicOr(&fresh[e >> 5], 1u << (e & 31u));  // this call wrote gpair[.., e]
        }
        __syncthreads();
        hi = lo;
    }
}

// Bits of fresh word wd that lie in the slot run [b, e).
__device__ __forceinline__ uint32_t run_bits(uint32_t wd, uint32_t b, uint32_t e) {
    const uint32_t w0 = wd * 32u;
    const uint32_t lo = b > w0 ? b - w0 : 0u, hi = e > w0 ? min(e - w0, 32u) : 0u;
    if (lo >= hi) return 0u;
    const uint32_t upto = hi == 32u ? 0xFFFFFFFFu : (1u << hi) - 1u;
    return upto & ~((1u << lo) - 1u);
}

// Sums each visible record's marked pairs, one thread per depth-ordered
// splat of a view. Few pairs are marked (a tile's backward stops where its
// pixels' contributions ended, and a pair no pixel took is not written), so a
// thread reads the marks of its splat's slot run and adds the marked pairs in
// slot order. A splat whose run spans more than PR_SERIAL_WORDS mark words,
// or holds more than PR_SERIAL marked pairs, is summed by its whole wave
// instead: the wave loads 64 mark words at once, and for each two words that
// hold marks the lanes take one slot each (lanes 0-31 the first word, 32-63
// the second), adding its pair to a per-lane sum; one nine-value wave
// reduction then gives the splat's totals. Either way the order of the
// additions is fixed, so the sums are deterministic.
__global__ __launch_bounds__(PR_THREADS) void k_pair_reduce(const float *__restrict__ gpair, uint64_t gstride,
                                                            const uint32_t *__restrict__ fresh,
                                                            const uint32_t *__restrict__ pslot,
                                                            const uint32_t *__restrict__ prid,
                                                            const uint32_t *__restrict__ vis_count, uint32_t n,
                                                            uint32_t bpv, float *__restrict__ g2d) {
    const uint32_t v = blockIdx.x / bpv, blk = blockIdx.x % bpv;
    const uint32_t nvis = min(vis_count[v], n);
    if (blk * PR_THREADS >= nvis) return;  // block-uniform: every lane of a live wave stays for the wave path
    const uint32_t j = blk * PR_THREADS + threadIdx.x, lane = lane_id();
    const bool valid = j < nvis;
    const uint32_t *ps = pslot + (uint64_t)v * (n + 1u);
    const uint32_t b = valid ? ps[j] : 0u, e = valid ? ps[j + 1] : 0u;
    const uint32_t wfirst = b >> 5, wend = (e + 31u) >> 5;  // mark words [wfirst, wend) cover the run
    const bool wide = wend - wfirst > PR_SERIAL_WORDS;
    static_assert(PR_SERIAL_WORDS == 4, "the slot pick below reads four words");
    // A narrow run's mark words, read at once.
    uint32_t wds[PR_SERIAL_WORDS];
    uint32_t nf = 0;
#pragma unroll
    for (uint32_t i = 0; i < PR_SERIAL_WORDS; ++i) {
        const uint32_t wd = wfirst + i;
        wds[i] = (!wide && wd < wend) ? fresh[wd] & run_bits(wd, b, e) : 0u;
        nf += __popc(wds[i]);
    }
    float sum[NG2D];
#pragma unroll
    for (int c = 0; c < NG2D; ++c) sum[c] = 0.0f;
    if (!wide && nf != 0u && nf <= PR_SERIAL) {
        // The marked slots in order, then all their reads at once, then the
        // sums in slot order (an absent slot adds zero, which changes nothing).
        uint32_t sl[PR_SERIAL];
#pragma unroll
        for (uint32_t k = 0; k < PR_SERIAL; ++k) {
            uint32_t i = 0;
#pragma unroll
            for (uint32_t q = PR_SERIAL_WORDS; q-- > 0;)
                if (wds[q]) i = q;
            const uint32_t w = wds[0] ? wds[0] : wds[1] ? wds[1] : wds[2] ? wds[2] : wds[3];
            sl[k] = (wfirst + i) * 32u + (w ? (uint32_t)__builtin_ctz(w) : 0u);
#pragma unroll
            for (uint32_t q = 0; q < PR_SERIAL_WORDS; ++q)
                if (q == i) wds[q] &= wds[q] - 1u;
        }
        float val[PR_SERIAL][NG2D];
#pragma unroll
        for (uint32_t k = 0; k < PR_SERIAL; ++k)
#pragma unroll
            for (int c = 0; c < NG2D; ++c) val[k][c] = k < nf ? gpair[c * gstride + sl[k]] : 0.0f;
#pragma unroll
        for (uint32_t k = 0; k < PR_SERIAL; ++k)
#pragma unroll
            for (int c = 0; c < NG2D; ++c) sum[c] += val[k][c];
    }
    bool any = nf != 0u;
    uint64_t big = lane_mask(wide || nf > PR_SERIAL);
    while (big) {
        const uint32_t l = (uint32_t)__builtin_ctzll(big);
        big &= big - 1ull;
        const uint32_t lb = (uint32_t)__builtin_amdgcn_readlane((int)b, (int)l);
        const uint32_t le = (uint32_t)__builtin_amdgcn_readlane((int)e, (int)l);
        const uint32_t lw0 = lb >> 5, lw1 = (le + 31u) >> 5;
        float acc[NG2D];
#pragma unroll
        for (int c = 0; c < NG2D; ++c) acc[c] = 0.0f;
        bool marked = false;
        for (uint32_t base = lw0; base < lw1; base += 64u) {
            const uint32_t wd = base + lane;
            const uint32_t word = wd < lw1 ? fresh[wd] & run_bits(wd, lb, le) : 0u;
            uint64_t wm = lane_mask(word != 0u);
            marked = marked || wm != 0ull;
            while (wm) {
                const uint32_t k1 = (uint32_t)__builtin_ctzll(wm);
                wm &= wm - 1ull;
                const uint32_t k2 = wm ? (uint32_t)__builtin_ctzll(wm) : k1;
                const bool second = wm != 0ull;
                if (second) wm &= wm - 1ull;
                const uint32_t k = lane < 32u ? k1 : k2;
                const uint32_t bits1 = (uint32_t)__builtin_amdgcn_readlane((int)word, (int)k1);
                const uint32_t bits2 = second ? (uint32_t)__builtin_amdgcn_readlane((int)word, (int)k2) : 0u;
                const uint32_t bits = lane < 32u ? bits1 : bits2;
                const uint32_t sl = (base + k) * 32u + (lane & 31u);
                if ((bits >> (lane & 31u)) & 1u)
#pragma unroll
                    for (int c = 0; c < NG2D; ++c) acc[c] += gpair[c * gstride + sl];
            }
        }
        if (!marked) continue;  // wave-uniform
        const float tot = wave_totals9(acc, lane);  // every lane of the wave takes part
#pragma unroll
        for (int c = 0; c < NG2D; ++c) {
            const float t = __uint_as_float(
                (uint32_t)__builtin_amdgcn_readlane((int)__float_as_uint(tot), (int)totals9_lane(c)));
            if (lane == l) sum[c] = t;
        }
        if (lane == l) any = true;
    }
    if (!valid || !any) return;  // records without marked pairs keep the caller's zeros
    float *dst = g2d + (uint64_t)prid[(uint64_t)v * n + j] * NG2D;
#pragma unroll
    for (int c = 0; c < NG2D; ++c) dst[c] = sum[c];
}

// Real SH basis of degree DEG (the order sh_color evaluates it) with its
// gradient: y[k] and dy[k] = d y[k] / d(x, y, z).
template <int DEG>
__device__ __forceinline__ void sh_basis(float x, float y, float z, float *yk, float (*dy)[3]) {
    const float c1 = SH_C1;
    yk[0] = -c1 * y; dy[0][0] = 0.0f; dy[0][1] = -c1; dy[0][2] = 0.0f;
    yk[1] = c1 * z;  dy[1][0] = 0.0f; dy[1][1] = 0.0f; dy[1][2] = c1;
    yk[2] = -c1 * x; dy[2][0] = -c1; dy[2][1] = 0.0f; dy[2][2] = 0.0f;
    if constexpr (DEG >= 2) {
        const float xx = x * x, yy = y * y, zz = z * z;
        const float *C2 = SH_C2;
        yk[3] = C2[0] * x * y; dy[3][0] = C2[0] * y; dy[3][1] = C2[0] * x; dy[3][2] = 0.0f;
        yk[4] = C2[1] * y * z; dy[4][0] = 0.0f; dy[4][1] = C2[1] * z; dy[4][2] = C2[1] * y;
        yk[5] = C2[2] * (2.0f * zz - xx - yy);
        dy[5][0] = -2.0f * C2[2] * x; dy[5][1] = -2.0f * C2[2] * y; dy[5][2] = 4.0f * C2[2] * z;
        yk[6] = C2[3] * x * z; dy[6][0] = C2[3] * z; dy[6][1] = 0.0f; dy[6][2] = C2[3] * x;
        yk[7] = C2[4] * (xx - yy); dy[7][0] = 2.0f * C2[4] * x; dy[7][1] = -2.0f * C2[4] * y; dy[7][2] = 0.0f;
        if constexpr (DEG >= 3) {
            const float *C3 = SH_C3;
            yk[8] = C3[0] * y * (3.0f * xx - yy);
            dy[8][0] = 6.0f * C3[0] * x * y; dy[8][1] = C3[0] * (3.0f * xx - 3.0f * yy); dy[8][2] = 0.0f;
            yk[9] = C3[1] * x * y * z;
            dy[9][0] = C3[1] * y * z; dy[9][1] = C3[1] * x * z; dy[9][2] = C3[1] * x * y;
            yk[10] = C3[2] * y * (4.0f * zz - xx - yy);
            dy[10][0] = -2.0f * C3[2] * x * y; dy[10][1] = C3[2] * (4.0f * zz - xx - 3.0f * yy);
            dy[10][2] = 8.0f * C3[2] * y * z;
            yk[11] = C3[3] * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
            dy[11][0] = -6.0f * C3[3] * x * z; dy[11][1] = -6.0f * C3[3] * y * z;
            dy[11][2] = C3[3] * (6.0f * zz - 3.0f * xx - 3.0f * yy);
            yk[12] = C3[4] * x * (4.0f * zz - xx - yy);
            dy[12][0] = C3[4] * (4.0f * zz - 3.0f * xx - yy); dy[12][1] = -2.0f * C3[4] * x * y;
            dy[12][2] = 8.0f * C3[4] * x * z;
            yk[13] = C3[5] * z * (xx - yy);
            dy[13][0] = 2.0f * C3[5] * x * z; dy[13][1] = -2.0f * C3[5] * y * z; dy[13][2] = C3[5] * (xx - yy);
            yk[14] = C3[6] * x * (xx - 3.0f * yy);
            dy[14][0] = C3[6] * (3.0f * xx - 3.0f * yy); dy[14][1] = -6.0f * C3[6] * x * y; dy[14][2] = 0.0f;
        }
    }
}

struct GradOut {
    float *means, *log_scales, *quats, *logits, *colors, *sh_rest;
};

template <int DEG>
__global__ __launch_bounds__(PBWD_THREADS) void k_project_bwd(DevScene sc, const ViewParams *__restrict__ views,
                                                              uint32_t nviews, const float *__restrict__ g2d,
                                                              GradOut go) {
    constexpr int R = (DEG + 1) * (DEG + 1) - 1;
    const uint32_t i = blockIdx.x * PBWD_THREADS + threadIdx.x;
    if (i >= sc.n) return;
    const DevParams &P = sc.prm;
    const uint64_t i3 = 3ull * i;
    float p[3], s[3], q[4], col[3], sh[3 * R + 1];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        p[c] = P.means[i3 + c];
        s[c] = expf(P.log_scales[i3 + c]);
        col[c] = P.colors[i3 + c];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) q[c] = P.quats[4ull * i + c];
#pragma unroll
    for (int c = 0; c < 3 * R; ++c) sh[c] = P.sh_rest[3ull * R * i + c];
    const float op = sigmoid_f(P.logits[i]);
    // Rotation of the normalised quaternion (as covariance_f32 builds it) and
    // the covariance the forward used.
    float cv[6];
    float qn[4] = {q[0], q[1], q[2], q[3]};
    covariance_f32(s, qn, cv);  // normalises qn in place
    const float qlen = sqrtf((q[0] * q[0] + q[2] * q[2]) + (q[1] * q[1] + q[3] * q[3]));
    const bool qok = __builtin_isfinite(qlen) && qlen > 1e-12f;
    const float x = qn[0], y = qn[1], z = qn[2], wq = qn[3];
    const float Rm[3][3] = {{1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y - wq * z), 2.0f * (x * z + wq * y)},
                            {2.0f * (x * y + wq * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z - wq * x)},
                            {2.0f * (x * z - wq * y), 2.0f * (y * z + wq * x), 1.0f - 2.0f * (x * x + y * y)}};
    const float Sg[3][3] = {{cv[0], cv[1], cv[2]}, {cv[1], cv[3], cv[4]}, {cv[2], cv[4], cv[5]}};
    float gp[3] = {0.0f, 0.0f, 0.0f}, gcol[3] = {0.0f, 0.0f, 0.0f}, gsh[3 * R + 1];
#pragma unroll
    for (int c = 0; c < 3 * R; ++c) gsh[c] = 0.0f;
    float Gs[3][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};  // sum of G + G^T
    float gop = 0.0f;
    // Each view's 2D gradients are read during the previous view's
    // arithmetic, so the loop does not wait for a load at every view.
    float gn[NG2D];
#pragma unroll
    for (int c = 0; c < NG2D; ++c) gn[c] = nviews ? g2d[(uint64_t)i * NG2D + c] : 0.0f;
    for (uint32_t vv = 0; vv < nviews; ++vv) {
        float g[NG2D];
        bool any = false;
#pragma unroll
        for (int c = 0; c < NG2D; ++c) {
            g[c] = gn[c];
            any = any || g[c] != 0.0f;
        }
        if (vv + 1 < nviews) {
            const uint64_t nrid = (uint64_t)(vv + 1) * sc.n + i;
#pragma unroll
            for (int c = 0; c < NG2D; ++c) gn[c] = g2d[nrid * NG2D + c];
        }
        if (!any) continue;
        const ViewCtx &v = views[vv].ctx;
        const float d0 = p[0] - v.eye[0], d1 = p[1] - v.eye[1], d2 = p[2] - v.eye[2];
        const float tx = (v.rot[0][0] * d0 + v.rot[0][1] * d1) + v.rot[0][2] * d2;
        const float ty = (v.rot[1][0] * d0 + v.rot[1][1] * d1) + v.rot[1][2] * d2;
        const float tz = (v.rot[2][0] * d0 + v.rot[2][1] * d1) + v.rot[2][2] * d2;
        const float inv_z = 1.0f / tz;
        const float jxr = tx * inv_z, jyr = ty * inv_z;
        const bool fx_free = jxr >= -v.lim_x && jxr <= v.lim_x;
        const bool fy_free = jyr >= -v.lim_y && jyr <= v.lim_y;
        const float jx = fminf(fmaxf(jxr, -v.lim_x), v.lim_x);
        const float jy = fminf(fmaxf(jyr, -v.lim_y), v.lim_y);
        const float f = v.focal * inv_z;
        float u0[3], u1[3], t0[3], t1[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            u0[c] = v.rot[0][c] - jx * v.rot[2][c];
            u1[c] = v.rot[1][c] - jy * v.rot[2][c];
            t0[c] = f * u0[c];
            t1[c] = f * u1[c];
        }
        float St0[3], St1[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            St0[r] = Sg[r][0] * t0[0] + Sg[r][1] * t0[1] + Sg[r][2] * t0[2];
            St1[r] = Sg[r][0] * t1[0] + Sg[r][1] * t1[1] + Sg[r][2] * t1[2];
        }
        const float a = (t0[0] * St0[0] + t0[1] * St0[1] + t0[2] * St0[2]) + DILATION;
        const float b = t0[0] * St1[0] + t0[1] * St1[1] + t0[2] * St1[2];
        const float cc = (t1[0] * St1[0] + t1[1] * St1[1] + t1[2] * St1[2]) + DILATION;
        const float det = a * cc - b * b;
        float inv_det;
        if (!(det > NEAR_SINGULAR * (a * cc)))
            inv_det = (float)(1.0 / ((double)a * (double)cc - (double)b * (double)b));
        else
            inv_det = 1.0f / det;
        const float id2 = inv_det * inv_det;
        const float gA = g[2], gB = g[3], gC = g[4];
        const float ga = (-cc * cc * gA + b * cc * gB - b * b * gC) * id2;
        const float gb = (2.0f * b * cc * gA - (a * cc + b * b) * gB + 2.0f * a * b * gC) * id2;
        const float gc = (-b * b * gA + a * b * gB - a * a * gC) * id2;
        float gt0[3], gt1[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            gt0[c] = 2.0f * ga * St0[c] + gb * St1[c];
            gt1[c] = gb * St0[c] + 2.0f * gc * St1[c];
        }
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                Gs[r][c] += 2.0f * ga * t0[r] * t0[c] + gb * (t0[r] * t1[c] + t1[r] * t0[c]) + 2.0f * gc * t1[r] * t1[c];
        const float gf = (gt0[0] * u0[0] + gt0[1] * u0[1] + gt0[2] * u0[2]) +
                         (gt1[0] * u1[0] + gt1[1] * u1[1] + gt1[2] * u1[2]);
        const float gjx = -f * (gt0[0] * v.rot[2][0] + gt0[1] * v.rot[2][1] + gt0[2] * v.rot[2][2]);
        const float gjy = -f * (gt1[0] * v.rot[2][0] + gt1[1] * v.rot[2][1] + gt1[2] * v.rot[2][2]);
        const float iz2 = inv_z * inv_z;
        float gtx = g[0] * v.focal * inv_z;
        float gty = g[1] * v.focal * inv_z;
        float gtz = -(g[0] * v.focal * tx + g[1] * v.focal * ty) * iz2 - gf * v.focal * iz2;
        if (fx_free) {
            gtx += gjx * inv_z;
            gtz -= gjx * tx * iz2;
        }
        if (fy_free) {
            gty += gjy * inv_z;
            gtz -= gjy * ty * iz2;
        }
        float gd[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) gd[c] = v.rot[0][c] * gtx + v.rot[1][c] * gty + v.rot[2][c] * gtz;
        // Colour: max(base + SH(dir), 0) with dir = normalize(p - eye).
        float pre[3] = {col[0], col[1], col[2]};
        float ux = 0.0f, uy = 0.0f, uz = 0.0f, rcp = 0.0f;
        if constexpr (DEG > 0) {
            const float len = sqrtf((d0 * d0 + d1 * d1) + d2 * d2);
            rcp = 1.0f / len;
            if (__builtin_isfinite(rcp) && rcp > 0.0f) {
                ux = d0 * rcp;
                uy = d1 * rcp;
                uz = d2 * rcp;
            } else {
                rcp = 0.0f;
            }
            sh_color<DEG>(sh, ux, uy, uz, pre);
        }
        float gcv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            gcv[c] = pre[c] > 0.0f ? g[6 + c] : 0.0f;
            gcol[c] += gcv[c];
        }
        if constexpr (DEG > 0) {
            float yk[R > 0 ? R : 1], dyk[R > 0 ? R : 1][3];
            sh_basis<DEG>(ux, uy, uz, yk, dyk);
            float gdir[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const float wk = gcv[0] * sh[3 * k] + gcv[1] * sh[3 * k + 1] + gcv[2] * sh[3 * k + 2];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    gsh[3 * k + c] += gcv[c] * yk[k];
                    gdir[c] += wk * dyk[k][c];
                }
            }
            const float dg = ux * gdir[0] + uy * gdir[1] + uz * gdir[2];
            gd[0] += (gdir[0] - ux * dg) * rcp;
            gd[1] += (gdir[1] - uy * dg) * rcp;
            gd[2] += (gdir[2] - uz * dg) * rcp;
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) gp[c] += gd[c];
        gop += g[5];
    }
    // Sigma = M M^T with M = R diag(s): dL/dM = (G + G^T) M.
    float gs[3], gR[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float acc = 0.0f;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const float gM = (Gs[r][0] * Rm[0][c] + Gs[r][1] * Rm[1][c] + Gs[r][2] * Rm[2][c]) * s[c];
            acc += gM * Rm[r][c];
            gR[r][c] = gM * s[c];
        }
        gs[c] = acc;
    }
    // d R / d (x, y, z, w) of the normalised quaternion.
    float gqn[4];
    gqn[0] = 2.0f * (y * (gR[0][1] + gR[1][0]) + z * (gR[0][2] + gR[2][0]) + wq * (gR[2][1] - gR[1][2])) -
             4.0f * x * (gR[1][1] + gR[2][2]);
    gqn[1] = 2.0f * (x * (gR[0][1] + gR[1][0]) + z * (gR[1][2] + gR[2][1]) + wq * (gR[0][2] - gR[2][0])) -
             4.0f * y * (gR[0][0] + gR[2][2]);
    gqn[2] = 2.0f * (x * (gR[0][2] + gR[2][0]) + y * (gR[1][2] + gR[2][1]) + wq * (gR[1][0] - gR[0][1])) -
             4.0f * z * (gR[0][0] + gR[1][1]);
    gqn[3] = 2.0f * (x * (gR[2][1] - gR[1][2]) + y * (gR[0][2] - gR[2][0]) + z * (gR[1][0] - gR[0][1]));
    const float qd = qn[0] * gqn[0] + qn[1] * gqn[1] + qn[2] * gqn[2] + qn[3] * gqn[3];
#pragma unroll
    for (int c = 0; c < 4; ++c) go.quats[4ull * i + c] = qok ? (gqn[c] - qn[c] * qd) / qlen : 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        go.means[i3 + c] = gp[c];
        go.log_scales[i3 + c] = gs[c] * s[c];
        go.colors[i3 + c] = gcol[c];
    }
    go.logits[i] = gop * op * (1.0f - op);
#pragma unroll
    for (int c = 0; c < 3 * R; ++c) go.sh_rest[3ull * R * i + c] = gsh[c];
}

template <int DEG>
void launch_pbwd(hipStream_t s, const DevScene &sc, const ViewParams *views, uint32_t nviews, const float *g2d,
                 const GradOut &go) {
    const uint32_t grid = (sc.n + PBWD_THREADS - 1) / PBWD_THREADS;
    hipLaunchKernelGGL(k_project_bwd<DEG>, dim3(grid), dim3(PBWD_THREADS), 0, s, sc, views, nviews, g2d, go);
}

}  // namespace

hipError_t launch_blend_bwd(hipStream_t s, const ViewParams *views, uint32_t nviews, uint32_t total_tiles,
                            uint32_t total_sbs, const uint32_t *tile_start, const uint32_t *sval,
                            const uint32_t *pbase,
                            const uint4 *rec, const float *pix_T, const uint32_t *pix_last, float *gpair,
                            uint64_t gstride, uint32_t *fresh) {
    if (total_tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(k_blend_bwd<BWD_BATCH>, dim3(blend_grid(total_sbs)), dim3(BLEND_THREADS), 0, s, views, nviews,
                       total_tiles, tile_start, sval, pbase, rec, pix_T, pix_last, gpair, gstride, fresh, total_sbs);
    return hipGetLastError();
}

hipError_t launch_pair_reduce(hipStream_t s, const float *gpair, uint64_t gstride, const uint32_t *fresh,
                              const uint32_t *pslot, const uint32_t *prid, const uint32_t *vis_count, uint32_t n,
                              uint32_t nviews, float *g2d) {
    if (n == 0 || nviews == 0) return hipSuccess;
    const uint32_t bpv = (n + PR_THREADS - 1) / PR_THREADS;
    hipLaunchKernelGGL(k_pair_reduce, dim3(bpv * nviews), dim3(PR_THREADS), 0, s, gpair, gstride, fresh, pslot, prid,
                       vis_count, n, bpv, g2d);
    return hipGetLastError();
}

hipError_t launch_project_bwd(hipStream_t s, const DevScene &sc, const ViewParams *views, uint32_t nviews,
                              const float *g2d, float *g_means, float *g_log_scales,
                              float *g_quats, float *g_logits, float *g_colors, float *g_sh_rest) {
    if (sc.n == 0) return hipSuccess;
    const GradOut go{g_means, g_log_scales, g_quats, g_logits, g_colors, g_sh_rest};
    switch (sc.deg) {
    case 0: launch_pbwd<0>(s, sc, views, nviews, g2d, go); break;
    case 1: launch_pbwd<1>(s, sc, views, nviews, g2d, go); break;
    case 2: launch_pbwd<2>(s, sc, views, nviews, g2d, go); break;
    default: launch_pbwd<3>(s, sc, views, nviews, g2d, go); break;
    }
    return hipGetLastError();
}

}  // namespace sp
