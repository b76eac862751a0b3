/*
 * ff_oracle_loop.c -- CPU restatement of fastfusion's loop-closure
 * de-integration (src/fusion/loopclosure.cpp).
 *
 * TEST INFRASTRUCTURE ONLY (the parity oracle); the product never links it.
 *
 * The reference subtracts frames whose pose a loop closure changed from a
 * partner map (FusionMipMapCPU(FusionMipMapCPU*, nLeavesTotalLoop),
 * geometryfusion_mipmap_cpu.cpp:645-717) whose bricks are numbered densely
 * in the order they are first touched: _leafMapLoop maps a brick of the
 * main map to its slot, _nLeavesMapped counts the slots handed out, and the
 * slot holds float distances, float weights (_weightsLoop) and colors.
 * Built with the flags of ff_oracle.c; the voxel loop runs with the SSE
 * rounding mode set to round-toward-zero like the reference
 * (loopclosure.cpp:271-272).
 */
#include <fenv.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ff_oracle.h"

#define BRICK 512
#define EPS_W 0.005f     /* DISTANCEWEIGHTEPSILON, definitions.h:251 */
#define WEIGHT_FACTOR 3.0f /* definitions.h:254 */
#define COLOR_MUL 256.0f /* COLOR_MULTIPLICATOR, definitions.h:56 */

struct ffo_loop {
    uint32_t map_n, total, nMapped;
    int use_color;
    uint32_t *map;   /* _leafMapLoop */
    float *dist;     /* _distance of the partner, 512 per slot */
    float *wgt;      /* _weightsLoop */
    uint16_t *col;   /* _color: R, G, B per voxel (colortype3 x, y, z) */
};

ffo_loop *ffo_loop_create(uint32_t map_n, uint32_t total, int use_color)
{
    ffo_loop *L = (ffo_loop *)calloc(1, sizeof(ffo_loop));
    L->map_n = map_n;
    L->total = total;
    L->use_color = use_color;
    L->map = (uint32_t *)malloc(sizeof(uint32_t) * map_n);
    L->dist = (float *)calloc((size_t)total * BRICK, sizeof(float));
    L->wgt = (float *)calloc((size_t)total * BRICK, sizeof(float));
    if (use_color) L->col = (uint16_t *)calloc((size_t)total * BRICK * 3, sizeof(uint16_t));
    ffo_loop_reset(L);
    return L;
}

void ffo_loop_destroy(ffo_loop *L)
{
    if (!L) return;
    free(L->map); free(L->dist); free(L->wgt); free(L->col);
    free(L);
}

/* updateArraysForLoopClosure, geometryfusion_mipmap_cpu.cpp:3927-3931: no
 * slot mapped, every brick unmapped.  The reference's memsets there cover
 * only the first _nLeavesTotal floats of the arrays; here every slot is
 * cleared, which is what those arrays hold after allocateMemoryCPU
 * (geometryfusion_mipmap_cpu.cpp:1025, 1034-1043). */
void ffo_loop_reset(ffo_loop *L)
{
    L->nMapped = 0;
    for (uint32_t i = 0; i < L->map_n; i++) L->map[i] = FFO_BRANCHINIT;
    memset(L->dist, 0, sizeof(float) * (size_t)L->total * BRICK);
    memset(L->wgt, 0, sizeof(float) * (size_t)L->total * BRICK);
    if (L->col) memset(L->col, 0, sizeof(uint16_t) * (size_t)L->total * BRICK * 3);
}

/* CVTPS2DQ under round-toward-zero and CVTTSS2SI: truncation, and the
 * integer-indefinite value for NaN and out-of-range inputs. */
static int cvt_rz(float v)
{
    if (!(v >= -2147483648.0f && v < 2147483648.0f)) return (int)0x80000000u;
    return (int)v;
}

/* update8AddLoopSSESingleDeleteloop, loopclosure.cpp:15-230, one SSE lane at
 * a time.  Differences from the integration kernel: float depth in metres
 * (NaN for no measurement), color from three planes, the x loop in two groups
 * of four lanes whose second group starts at the first plus 4 d11, no
 * maximum-distance test but a NaN test on the SDF, the fall-off term with
 * its sign reversed, and colors stored by plain conversion (no clamp).  Must
 * run with the rounding mode set to round-toward-zero. */
static void delete_brick_rz(const float *depth, const uint8_t *red, const uint8_t *green,
                            const uint8_t *blue, int W, int H, const float *m, float fx, float fy,
                            float cx, float cy, float scale, float dT, const uint16_t *o,
                            uint16_t leafScale, float *dist, float *wgt, uint16_t *col)
{
    volatile float vs = scale; /* keep the compiler from folding across fesetround */
    float s = vs;
    float fls = (float)leafScale * s;
    float ox = ((m[0] * (float)o[0] + m[1] * (float)o[1]) + m[2] * (float)o[2]) * s + m[3];
    float oy = ((m[4] * (float)o[0] + m[5] * (float)o[1]) + m[6] * (float)o[2]) * s + m[7];
    float oz = ((m[8] * (float)o[0] + m[9] * (float)o[1]) + m[10] * (float)o[2]) * s + m[11];
    float d11 = m[0] * fls, d21 = m[4] * fls, d31 = m[8] * fls;
    /* _mm_setr_ps(0.0f, d11, 2.0f*d11, 3.0f*d11), loopclosure.cpp:42-44 */
    float l1[4] = {0.0f, d11, 2.0f * d11, 3.0f * d11};
    float l2[4] = {0.0f, d21, 2.0f * d21, 3.0f * d21};
    float l3[4] = {0.0f, d31, 2.0f * d31, 3.0f * d31};
    float td = dT * (float)leafScale;
    float tw = (WEIGHT_FACTOR * dT) * (float)leafScale;
    d11 *= 4.0f; d21 *= 4.0f; d31 *= 4.0f;
    float d12 = m[1] * fls, d22 = m[5] * fls, d32 = m[9] * fls;
    float d13 = m[2] * fls, d23 = m[6] * fls, d33 = m[10] * fls;
    float rfall = ffo_rcp(tw - EPS_W);
    float pzx = ox, pzy = oy, pzz = oz;
    int idx = 0;
    for (int z = 0; z < 8; z++, pzx += d13, pzy += d23, pzz += d33) {
        float pyx = pzx, pyy = pzy, pyz = pzz;
        for (int y = 0; y < 8; y++, pyx += d12, pyy += d22, pyz += d32) {
            float pxx = pyx, pxy = pyy, pxz = pyz;
            for (int x = 0; x < 2; x++, pxx += d11, pxy += d21, pxz += d31) {
                for (int k = 0; k < 4; k++, idx++) {
                    float px = pxx + l1[k], py = pxy + l2[k], pz = pxz + l3[k];
                    float len = sqrtf((px * px + py * py) + pz * pz);
                    float r = ffo_rcp(pz);
                    int imx = cvt_rz(cx + fx * (px * r));
                    int imy = cvt_rz(cy + fy * (py * r));
                    int cyi = imy < H - 1 ? imy : H - 1; if (cyi < 0) cyi = 0;
                    int cxi = imx < W - 1 ? imx : W - 1; if (cxi < 0) cxi = 0;
                    int ii = W * cyi + cxi;
                    float h = depth[ii];
                    float dI = len - (len * r) * h;
                    float wA = wgt[idx], dA = dist[idx];
                    int mask = (pz > 0.0f) && (imy < H) && (imy > -1) && (imx < W) && (imx > -1) && (dI == dI);
                    float front = (dI < EPS_W) ? 1.0f : 0.0f;
                    float back = (dI >= EPS_W && dI < tw) ? 1.0f : 0.0f;
                    float fall = (dI - tw) * rfall;
                    float wI = mask ? (front + fall * back) : 0.0f;
                    float wN = wA + wI;
                    float f = ffo_rcp(wN + ((wN <= 0.0f) ? 1.0f : 0.0f));
                    float mn = (td < dI) ? td : dI;                 /* _mm_min_ps(td, dI) */
                    float mx = (-1.0f * td > mn) ? -1.0f * td : mn; /* _mm_max_ps(-td, mn) */
                    float add = mask ? mx * wI : 0.0f;
                    dist[idx] = ((dA * wA) + add) * f;
                    wgt[idx] = wN;
                    if (col) {
                        const uint8_t *plane[3] = {red, green, blue};
                        for (int ch = 0; ch < 3; ch++) {
                            float inc = (float)plane[ch][ii] * COLOR_MUL;
                            float acc = (float)col[idx * 3 + ch];
                            float v = ((acc * wA) + (inc * wI)) * f;
                            col[idx * 3 + ch] = (uint16_t)cvt_rz(v);
                        }
                    }
                }
            }
        }
    }
}

void ffo_delete_brick(const float *depth, const uint8_t *red, const uint8_t *green, const uint8_t *blue,
                      int w, int h, const float m[12], const float K[4], float scale,
                      float distance_threshold, const uint16_t pos[3], uint16_t leaf_scale,
                      float *dist, float *weight, uint16_t *color)
{
    int old = fegetround();
    fesetround(FE_TOWARDZERO);
    delete_brick_rz(depth, red, green, blue, w, h, m, K[0], K[1], K[2], K[3], scale, distance_threshold,
                    pos, leaf_scale, dist, weight, color);
    fesetround(old);
}

/* updateWrapper_deleteloop, loopclosure.cpp:232-311: the queued bricks in
 * order; a brick without a slot takes the next one, and when none is left
 * the call stops there (the reference returns false, leaving the bricks
 * before it updated).  The pose is CameraInfo camera->world, inverted and
 * combined with the map offset as in deleteOldLoopPoses
 * (geometryfusion_mipmap_cpu.cpp:3990-4010, the same construction as addMap's
 * pInv).  Returns 1, or 0 when the slots ran out; *applied = queue entries
 * updated. */
int ffo_loop_delete(ffo_loop *L, const ffo_ctx *c, const float *depth, const uint8_t *red,
                    const uint8_t *green, const uint8_t *blue, int w, int h, const double R[9],
                    const double t[3], const double K[4], const uint32_t *queue, uint32_t n,
                    uint32_t *applied)
{
    float m[12], Kf[4];
    ffo_camera_inverse(c, R, t, K, m, Kf);
    const uint16_t *pos = ffo_leaf_pos(c), *lsc = ffo_leaf_scale(c);
    float scale = ffo_scale(c), dT = ffo_threshold(c);
    int ok = 1;
    uint32_t l = 0;
    int old = fegetround();
    fesetround(FE_TOWARDZERO);
    for (; l < n; l++) {
        uint32_t b = queue[l];
        uint32_t slot = L->map[b];
        if (slot >= L->nMapped) {
            if (L->nMapped >= L->total) { ok = 0; break; }
            slot = L->map[b] = L->nMapped++;
        }
        size_t v = (size_t)slot * BRICK;
        delete_brick_rz(depth, red, green, blue, w, h, m, Kf[0], Kf[1], Kf[2], Kf[3], scale, dT,
                        pos + (size_t)b * 3, lsc[b], L->dist + v, L->wgt + v,
                        L->col ? L->col + v * 3 : NULL);
    }
    fesetround(old);
    if (applied) *applied = l;
    return ok;
}

uint32_t ffo_loop_mapped(const ffo_loop *L) { return L->nMapped; }
const uint32_t *ffo_loop_map(const ffo_loop *L) { return L->map; }
const float *ffo_loop_distance(const ffo_loop *L) { return L->dist; }
const float *ffo_loop_weights(const ffo_loop *L) { return L->wgt; }
const uint16_t *ffo_loop_color(const ffo_loop *L) { return L->col; }
