// ff_device.hpp -- device-side arithmetic that reproduces the reference CPU's
// voxel-update numerics on gfx950.
//
// The reference updates voxels with the SSE/AVX rounding mode set to
// round-toward-zero (src/fusion/update_functions.cpp:1678-1679) and the
// RCPPS approximate reciprocal (update_functions.cpp:1246, 1373, 1386).
// The voxel kernels do the same: their first statement sets the f32
// rounding mode of the MODE register to round-toward-zero (rz_mode_on), and
// rz_add / rz_sub / rz_mul are then single hardware operations, named so
// that the reference's operation order stays visible (the build does not
// contract them: -ffp-contract=off).  The MODE register is per wave, so the
// setting ends with the kernel's waves.  RCPPS is a table lookup and the
// x86 truncating conversions are restated below.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FF_BRANCHINIT 4294967294u
#define FF_BL 8
#define FF_BRICK 512

namespace ff {

// sets the f32 round mode of the calling wave to round-toward-zero: the
// first statement of every kernel that uses the functions below
__device__ __forceinline__ void rz_mode_on() { __builtin_amdgcn_s_setreg(2049 /* hwreg(MODE, 0, 2) */, 3); }

// valid only after rz_mode_on()
__device__ __forceinline__ float rz_mul(float a, float b) { return a * b; }
__device__ __forceinline__ float rz_add(float a, float b) { return a + b; }
__device__ __forceinline__ float rz_sub(float a, float b) { return a - b; }

// sqrt rounded toward zero: start from the hardware estimate and correct it
// with FMA residuals until s*s <= x < next(s)^2 (only the residuals' signs
// are used, which no rounding mode changes here).
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
