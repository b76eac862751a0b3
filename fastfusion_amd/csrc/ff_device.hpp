// ff_device.hpp -- device-side arithmetic that reproduces the reference CPU's
// voxel-update numerics on gfx950.
//
// The reference updates voxels with the SSE/AVX rounding mode set to
// round-toward-zero (src/fusion/update_functions.cpp:1678-1679) and the
// RCPPS approximate reciprocal (update_functions.cpp:1246, 1373, 1386).
// Here every round-toward-zero operation is computed in the default
// round-to-nearest mode together with its exact rounding error (TwoSum,
// FMA-based TwoProduct); when the nearest result lies farther from zero than
// the exact one, it is stepped one ulp toward zero.  This gives the
// round-toward-zero result bit for bit without touching the MODE register.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FF_BRANCHINIT 4294967294u
#define FF_BL 8
#define FF_BRICK 512

namespace ff {

// r: round-to-nearest result; e: exact error (exact value minus r), or NaN
// when the operation involved an infinity.  Returns the round-toward-zero
// result.
__device__ __forceinline__ float rz_fix(float r, float e)
{
    const uint32_t rb = __float_as_uint(r);
    const bool adjust = (e != 0.0f) && (e == e) && (((__float_as_uint(e) ^ rb) & 0x80000000u) != 0u);
    return adjust ? __uint_as_float(rb - 1u) : r;
}

__device__ __forceinline__ float rz_mul(float a, float b)
{
    const float p = a * b;
    return rz_fix(p, __builtin_fmaf(a, b, -p));
}

__device__ __forceinline__ float rz_add(float a, float b)
{
    const float s = a + b;
    const float bb = s - a;
    const float e = (a - (s - bb)) + (b - bb);
    return rz_fix(s, e);
}

__device__ __forceinline__ float rz_sub(float a, float b) { return rz_add(a, -b); }

// sqrt rounded toward zero: start from the hardware estimate and correct it
// with exact FMA residuals until s*s <= x < next(s)^2.
__device__ __forceinline__ float rz_sqrt(float x)
{
    if (!(x > 0.0f) || x == __builtin_huge_valf()) return __builtin_sqrtf(x);
    float s = __builtin_sqrtf(x);
#pragma unroll
    for (int i = 0; i < 2; i++)
        if (__builtin_fmaf(s, s, -x) > 0.0f) s = __uint_as_float(__float_as_uint(s) - 1u);
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const float n = __uint_as_float(__float_as_uint(s) + 1u);
        if (__builtin_fmaf(n, n, -x) <= 0.0f) s = n;
    }
    return s;
}

// RCPPS from a 2048-entry table of its outputs for inputs 1.m (m = i << 12),
// held in LDS by the caller.
__device__ __forceinline__ float rcp_x86(float x, const uint32_t *tab)
{
    const uint32_t u = __float_as_uint(x);
    const uint32_t s = u & 0x80000000u, e = (u >> 23) & 0xffu, m = u & 0x7fffffu;
    if (e == 0u) return __uint_as_float(s | 0x7f800000u);
    if (e == 255u) return __uint_as_float(m ? (u | 0x00400000u) : s);
    const uint32_t t = tab[m >> 12];
    const int re = (int)((t >> 23) & 0xffu) + 127 - (int)e;
    if (re <= 0) return __uint_as_float(s);
    return __uint_as_float(s | ((uint32_t)re << 23) | (t & 0x7fffffu));
}

// CVTTPS2DQ / CVTPS2DQ-under-RZ semantics: truncation, and 0x80000000 for
// NaN and out-of-range inputs.
__device__ __forceinline__ int cvt_x86(float v)
{
    return (v >= -2147483648.0f && v < 2147483648.0f) ? (int)v : (int)0x80000000u;
}

// leafScaleFunc with SCALE_LINEAR (src/fusion/inline_functions.hpp:128-145)
__device__ __forceinline__ uint32_t leaf_scale_target(float depth)
{
    const float f = fmaxf(__fdiv_rn(depth, 1.0f), 1.0f);
    return (uint32_t)(uint16_t)floorf(__fmul_rn(f, 8.0f));
}

}  // namespace ff
