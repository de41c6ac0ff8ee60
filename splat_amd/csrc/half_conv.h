// f32 -> IEEE binary16 / bfloat16 conversion with round-to-nearest-even, for
// the reduced-precision scene storage.
#pragma once

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define SP_HALF_HD __host__ __device__
#else
#define SP_HALF_HD
#endif

namespace sp {

// The same code converts on the host (tests, rounded-scene references) and on
// the device (the scene upload), so both give the same bits.
SP_HALF_HD inline uint16_t f32_to_f16(float f) {
    uint32_t x;
    __builtin_memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t a = x & 0x7FFFFFFFu;
    if (a >= 0x7F800000u) return (uint16_t)(sign | (a > 0x7F800000u ? 0x7E00u : 0x7C00u));  // NaN / inf
    if (a >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u);  // rounds to >= 65520: inf
    if (a < 0x38800000u) {                                   // below 2^-14: half subnormal or zero
        if (a < 0x33000000u) return (uint16_t)sign;         // below 2^-25: rounds to zero
        const uint32_t e = a >> 23;
        const uint32_t m = (a & 0x7FFFFFu) | 0x800000u;
        const uint32_t shift = 126u - e;  // 14..24
        uint32_t q = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (q & 1u))) ++q;
        return (uint16_t)(sign | q);
    }
    uint32_t r = a - 0x38000000u;  // rebias the exponent (127 -> 15)
    r += 0xFFFu + ((r >> 13) & 1u);
    return (uint16_t)(sign | (r >> 13));
}

SP_HALF_HD inline uint16_t f32_to_bf16(float f) {
    uint32_t x;
    __builtin_memcpy(&x, &f, 4);
    if ((x & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((x >> 16) | 0x40u);  // quiet NaN
    x += 0x7FFFu + ((x >> 16) & 1u);
    return (uint16_t)(x >> 16);
}

}  // namespace sp
