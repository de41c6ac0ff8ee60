// Scene preparation on the device (the upload half of Prep::new,
// src/render/cpu.rs:349-393): the caller's arrays-of-structures, copied to the
// device as they are, become the renderer's blocked column layout (scene_at:
// blocks of 64 splats, field-major inside a block) in one kernel, one thread
// per splat. Each output column of a block is written by consecutive threads
// (coalesced); each splat's own fields are read from its contiguous record.
//
// f32 storage: columns position x3, cov3d x6, opacity, k, base colour x3, SH
// rest; the colour again as rows.
// 16-bit storage: f32 columns position x3, opacity, k; 16-bit columns scale
// x3, quaternion x4, base colour x3 and SH rest; the colour again as 16-bit
// rows; all 16-bit values rounded to FP16 or BF16 (round to nearest even, the
// host's conversion code compiled for the device, so the bits are the host's).
// k = sqrt(2 ln(o / ALPHA_MIN)) arrives precomputed by the host, whose logf
// the reference's bits follow.
#include "half_conv.h"
#include "sp_device.h"
#include "sp_launch.h"

namespace sp {
namespace {

// Sets *flag once per wave that holds a splat with an extreme colour value
// (launch_prep).
__device__ __forceinline__ void flag_extreme(bool bad, uint32_t *flag) {
    if (wave_any(bad) && lane_id() == 0) atomicOr(flag, 1u);
}

__global__ __launch_bounds__(256) void k_prep_f32(const float *__restrict__ pos, const float *__restrict__ cov,
                                                  const float *__restrict__ opac, const float *__restrict__ kk,
                                                  const float *__restrict__ col, const float *__restrict__ sh,
                                                  uint32_t n, uint32_t rest, float *__restrict__ soa,
                                                  float *__restrict__ crow, uint32_t *__restrict__ extreme) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    bool bad = false;
    const uint32_t NC = A_SH + 3 * rest;
    if (i < n) {
#pragma unroll
        for (int c = 0; c < 3; ++c) soa[scene_at(NC, A_POS + c, i)] = pos[3ull * i + c];
#pragma unroll
        for (int c = 0; c < 6; ++c) soa[scene_at(NC, A_COV + c, i)] = cov[6ull * i + c];
        soa[scene_at(NC, A_OPAC, i)] = opac[i];
        soa[scene_at(NC, A_K, i)] = kk[i];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float v = col[3ull * i + c];
            bad |= !(fabsf(v) < 1e36f);
            soa[scene_at(NC, A_COL + c, i)] = v;
        }
        for (uint32_t c = 0; c < 3 * rest; ++c) {
            const float v = sh[(uint64_t)3 * rest * i + c];
            bad |= !(fabsf(v) < 1e36f);
            soa[scene_at(NC, A_SH + c, i)] = v;
        }
        // The row in 16-byte stores: position, base colour, SH rest, zero padding.
        const uint32_t cw = colour_row(rest, false);
        const auto val = [&](uint32_t q) {
            return q < 3u   ? pos[3ull * i + q]
                   : q < 6u ? col[3ull * i + (q - 3u)]
                   : q - 6u < 3u * rest ? sh[(uint64_t)3 * rest * i + (q - 6u)] : 0.0f;
        };
        float4 *row = reinterpret_cast<float4 *>(crow + (uint64_t)cw * i);
        for (uint32_t q = 0; q < cw; q += 4) row[q / 4] = make_float4(val(q), val(q + 1), val(q + 2), val(q + 3));
    }
    flag_extreme(bad, extreme);
}

template <bool BF16>
__device__ __forceinline__ uint16_t to16(float f) {
    return BF16 ? f32_to_bf16(f) : f32_to_f16(f);
}

template <bool BF16>
__global__ __launch_bounds__(256) void k_prep_16(const float *__restrict__ pos, const float *__restrict__ scl,
                                                 const float *__restrict__ quat, const float *__restrict__ opac,
                                                 const float *__restrict__ kk, const float *__restrict__ col,
                                                 const float *__restrict__ sh, uint32_t n, uint32_t rest,
                                                 float *__restrict__ soa, uint16_t *__restrict__ soa16,
                                                 uint16_t *__restrict__ crow16, uint32_t *__restrict__ extreme) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const float lim = BF16 ? 1e36f : 65504.0f;
    bool bad = false;
    const uint32_t NH = H_SH + 3 * rest;
    if (i < n) {
#pragma unroll
        for (int c = 0; c < 3; ++c) soa[scene_at(L_COLS, L_POS + c, i)] = pos[3ull * i + c];
        soa[scene_at(L_COLS, L_OPAC, i)] = opac[i];
        soa[scene_at(L_COLS, L_K, i)] = kk[i];
#pragma unroll
        for (int c = 0; c < 3; ++c) soa16[scene_at(NH, H_SCALE + c, i)] = to16<BF16>(scl[3ull * i + c]);
#pragma unroll
        for (int c = 0; c < 4; ++c) soa16[scene_at(NH, H_QUAT + c, i)] = to16<BF16>(quat[4ull * i + c]);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float v = col[3ull * i + c];
            bad |= !(fabsf(v) < lim);
            soa16[scene_at(NH, H_COL + c, i)] = to16<BF16>(v);
        }
        for (uint32_t c = 0; c < 3 * rest; ++c) {
            const float v = sh[(uint64_t)3 * rest * i + c];
            bad |= !(fabsf(v) < lim);
            soa16[scene_at(NH, H_SH + c, i)] = to16<BF16>(v);
        }
        // The row in 16-byte stores of eight values: the f32 position's bits
        // (low half first), base colour, SH rest, zero padding.
        const uint32_t cw = colour_row(rest, true);
        const auto val = [&](uint32_t q) -> uint32_t {
            return q < 6u   ? (__float_as_uint(pos[3ull * i + q / 2u]) >> (16u * (q & 1u))) & 0xFFFFu
                   : q < 9u ? to16<BF16>(col[3ull * i + (q - 6u)])
                   : q - 9u < 3u * rest ? to16<BF16>(sh[(uint64_t)3 * rest * i + (q - 9u)]) : 0u;
        };
        uint4 *row = reinterpret_cast<uint4 *>(crow16 + (uint64_t)cw * i);
        for (uint32_t q = 0; q < cw; q += 8)
            row[q / 8] = make_uint4(val(q) | val(q + 1) << 16, val(q + 2) | val(q + 3) << 16,
                                    val(q + 4) | val(q + 5) << 16, val(q + 6) | val(q + 7) << 16);
    }
    flag_extreme(bad, extreme);
}

}  // namespace

hipError_t launch_prep(hipStream_t s, int storage, const PrepInput &in, uint32_t n, uint32_t rest, float *soa,
                       uint16_t *soa16, float *crow, uint16_t *crow16, uint32_t *extreme) {
    if (n == 0) return hipSuccess;
    const uint32_t grid = (n + 255) / 256;
    if (storage == SP_STORE_F32)
        hipLaunchKernelGGL(k_prep_f32, dim3(grid), dim3(256), 0, s, in.pos, in.cov, in.opac, in.k, in.col, in.sh, n,
                           rest, soa, crow, extreme);
    else if (storage == SP_STORE_F16)
        hipLaunchKernelGGL(k_prep_16<false>, dim3(grid), dim3(256), 0, s, in.pos, in.scl, in.quat, in.opac, in.k,
                           in.col, in.sh, n, rest, soa, soa16, crow16, extreme);
    else
        hipLaunchKernelGGL(k_prep_16<true>, dim3(grid), dim3(256), 0, s, in.pos, in.scl, in.quat, in.opac, in.k,
                           in.col, in.sh, n, rest, soa, soa16, crow16, extreme);
    return hipGetLastError();
}

}  // namespace sp
