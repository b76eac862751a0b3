// ff_sdf.hpp -- the per-row part of the voxel kernels, shared by the
// integration update (update8AddLoopAVXSingleInteger,
// src/fusion/update_functions.cpp:1187-1444) and the loop-closure
// de-integration (update8AddLoopSSESingleDeleteloop,
// src/fusion/loopclosure.cpp:15-230), which build a row's origin, column
// step and thresholds the same way.  Valid only after rz_mode_on().
#pragma once
#include "ff_device.hpp"
#include "ff_internal.hpp"

namespace ff {

// The per-row part of update8AddLoopAVXSingleInteger: the row origin of
// brick row (zr, yr) accumulated by repeated adds in the reference's order,
// the column step d1, the truncations and the fall-off reciprocal.
struct SdfRow {
    float R[3], d1[3], td, tw, rfall;
};

__device__ __forceinline__ SdfRow sdf_row_setup(const ushort4 info, int zr, int yr, const SdfParams &P,
                                                const uint32_t *tab)
{
    const float EPS = 0.005f;
    SdfRow w;
    const float ls = (float)info.w;
    const float fls = rz_mul(ls, P.scale);
    const float ofx = (float)info.x, ofy = (float)info.y, ofz = (float)info.z;
    float d2[3], d3[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        w.R[r] = rz_add(rz_mul(rz_add(rz_add(rz_mul(P.m[4 * r], ofx), rz_mul(P.m[4 * r + 1], ofy)),
                                      rz_mul(P.m[4 * r + 2], ofz)),
                               P.scale),
                        P.m[4 * r + 3]);
        w.d1[r] = rz_mul(P.m[4 * r], fls);
        d2[r] = rz_mul(P.m[4 * r + 1], fls);
        d3[r] = rz_mul(P.m[4 * r + 2], fls);
    }
    w.td = rz_mul(P.dT, ls);
    w.tw = rz_mul(rz_mul(3.0f, P.dT), ls);
    for (int i = 0; i < zr; i++) {
        w.R[0] = rz_add(w.R[0], d3[0]); w.R[1] = rz_add(w.R[1], d3[1]); w.R[2] = rz_add(w.R[2], d3[2]);
    }
    for (int i = 0; i < yr; i++) {
        w.R[0] = rz_add(w.R[0], d2[0]); w.R[1] = rz_add(w.R[1], d2[1]); w.R[2] = rz_add(w.R[2], d2[2]);
    }
    w.rfall = rcp_x86(rz_sub(w.tw, EPS), tab);
    return w;
}

}  // namespace ff
