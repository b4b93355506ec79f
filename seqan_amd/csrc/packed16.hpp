// packed16.hpp -- packed int16 helpers shared by the gfx950 kernels: two
// int16 lanes per 32-bit register (v_pk_add_u16 / v_pk_max_i16 /
// v_pk_sub_i16 with clamp), the halves holding two different pairs.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace sqmi {

typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ s16x2 as16(uint32_t x) { return __builtin_bit_cast(s16x2, x); }
__device__ __forceinline__ uint32_t asu(s16x2 x) { return __builtin_bit_cast(uint32_t, x); }
__device__ __forceinline__ s16x2 vmax(s16x2 a, s16x2 b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ s16x2 vmin(s16x2 a, s16x2 b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ s16x2 splat(int x)
{
    s16x2 r = {(short)x, (short)x};
    return r;
}
__device__ __forceinline__ s16x2 pack(int lo, int hi)
{
    s16x2 r = {(short)lo, (short)hi};
    return r;
}
__device__ __forceinline__ uint32_t bfi(uint32_t mask, uint32_t a, uint32_t b) { return (mask & a) | (~mask & b); }
// Saturating a + b per half (v_pk_add_i16 with clamp).
__device__ __forceinline__ s16x2 addSat(s16x2 a, s16x2 b) { return __builtin_elementwise_add_sat(a, b); }
// Saturating a - b per half: its sign bit is exactly (a < b), whatever the
// magnitudes (v_pk_sub_i16 with clamp).
__device__ __forceinline__ s16x2 subSat(s16x2 a, s16x2 b) { return __builtin_elementwise_sub_sat(a, b); }
// 0xffff in the halves where a > b.
__device__ __forceinline__ uint32_t gtMask(s16x2 a, s16x2 b) { return asu(subSat(b, a) >> splat(15)); }

// 0xffff in the halves where a == b.
__device__ __forceinline__ uint32_t eqMask(s16x2 a, s16x2 b)
{
    return ((uint16_t)a.x == (uint16_t)b.x ? 0xffffu : 0u) | ((uint16_t)a.y == (uint16_t)b.y ? 0xffff0000u : 0u);
}

}  // namespace sqmi
