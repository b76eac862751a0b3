/*
 * ff_oracle.c -- CPU restatement of fastfusion's integration path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity oracle and the timed CPU baseline);
 * the product never links it.  Every function cites the reference lines it
 * follows; reference paths are relative to the reference repository root.
 *
 * Build: gcc -O2 -fPIC -shared -frounding-math -ffp-contract=off -msse2 -mno-fma
 * The voxel update switches the SSE rounding mode to round-toward-zero like
 * the reference (update_functions.cpp:1678-1679) so every float operation in
 * it rounds the same way; -frounding-math stops constant folding across that.
 */
#include "ff_oracle.h"

#include <fenv.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../fastfusion_amd/csrc/rcp_table.h"

#define BL 8             /* BRICKLENGTH, definitions.h:234 */
#define BRICK 512
#define EPS_W 0.005f     /* DISTANCEWEIGHTEPSILON, definitions.h:251 */
#define WEIGHT_FACTOR 3.0f
#define COLOR_MUL 256.0f /* COLOR_MULTIPLICATOR, definitions.h:62 */

struct ffo_ctx {
    float scale, dT;
    int use_color;
    uint16_t bandwidth;
    uint16_t n;
    float offset[3];
    uint32_t nLeavesTotal, nBranchesTotal, nLeavesUsed, nBranchesUsed;
    uint32_t *tree;
    uint16_t *leafPos, *leafScale;
    uint32_t *leafParent;
    float *dist, *wgt;
    uint16_t *col;
    uint32_t *queueIndexOfLeaf, *queue, nQueued;
    int imgW, imgH;
    float *pxp, *pyp;
    uint8_t *processed;
    int grows;
    const float *conf; /* per-pixel confidence for the next frames, or NULL */
};

static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* RCPPS reproduced from the captured table (update_functions.cpp:1246). */
float ffo_rcp(float x)
{
    uint32_t u = f2u(x), s = u & 0x80000000u, e = (u >> 23) & 0xffu, m = u & 0x7fffffu;
    if (e == 0) return u2f(s | 0x7f800000u);
    if (e == 255) return u2f(m ? (u | 0x00400000u) : s);
    uint32_t t = FF_RCP_TABLE[m >> 12];
    int re = (int)((t >> 23) & 0xffu) + 127 - (int)e;
    if (re <= 0) return u2f(s);
    return u2f(s | ((uint32_t)re << 23) | (t & 0x7fffffu));
}

/* leafScaleFunc with SCALE_LINEAR (inline_functions.hpp:128-145). */
uint16_t ffo_leaf_scale_func(float depth)
{
    float factor = fmaxf(depth / 1.0f, 1.0f);
    return (uint16_t)floorf(factor * (float)BL);
}

ffo_ctx *ffo_create(float scale, float distance_threshold, int use_color,
                    uint32_t n_leaves_total, uint32_t n_branches_total)
{
    ffo_ctx *c = (ffo_ctx *)calloc(1, sizeof(ffo_ctx));
    c->scale = scale;
    c->dT = distance_threshold;
    c->use_color = use_color;
    /* _bandwidth(_distanceThreshold / _scale*BANDWIDTHFACTOR), ctor
     * geometryfusion_mipmap_cpu.cpp:425: float quotient, double product, u16. */
    c->bandwidth = (uint16_t)((double)(distance_threshold / scale) * 1.0);
    c->nLeavesTotal = n_leaves_total;
    c->nBranchesTotal = n_branches_total;
    c->nBranchesUsed = 16; /* ctor 622 */
    c->tree = (uint32_t *)malloc(sizeof(uint32_t) * n_branches_total);
    for (uint32_t i = 0; i < n_branches_total; i++) c->tree[i] = FFO_BRANCHINIT;
    c->leafPos = (uint16_t *)calloc((size_t)n_leaves_total * 3, sizeof(uint16_t));
    c->leafScale = (uint16_t *)calloc(n_leaves_total, sizeof(uint16_t));
    c->leafParent = (uint32_t *)calloc(n_leaves_total, sizeof(uint32_t));
    c->dist = (float *)calloc((size_t)n_leaves_total * BRICK, sizeof(float));
    c->wgt = (float *)calloc((size_t)n_leaves_total * BRICK, sizeof(float));
    c->col = (uint16_t *)calloc((size_t)n_leaves_total * BRICK * 3, sizeof(uint16_t));
    c->queueIndexOfLeaf = (uint32_t *)malloc(sizeof(uint32_t) * n_leaves_total);
    c->queue = (uint32_t *)malloc(sizeof(uint32_t) * n_leaves_total);
    return c;
}

void ffo_destroy(ffo_ctx *c)
{
    if (!c) return;
    free(c->tree); free(c->leafPos); free(c->leafScale); free(c->leafParent);
    free(c->dist); free(c->wgt); free(c->col); free(c->queueIndexOfLeaf); free(c->queue);
    free(c->pxp); free(c->pyp); free(c->processed);
    ffo_mesh_forget(c);
    free(c);
}

/* camPamsFloat (definitions.h:70-89) */
typedef struct { float r[9], t[3], fx, fy, cx, cy; } campams;

/* p: world<-camera in voxel units, geometryfusion_mipmap_cpu.cpp:1617-1625 */
static campams make_p(const ffo_ctx *c, const double R[9], const double t[3], const double K[4])
{
    campams p;
    double s = (double)c->scale;
    for (int i = 0; i < 9; i++) p.r[i] = (float)(R[i] / s);
    for (int i = 0; i < 3; i++) p.t[i] = (float)((t[i] - (double)c->offset[i]) / s);
    p.fx = (float)K[0]; p.fy = (float)K[1]; p.cx = (float)K[2]; p.cy = (float)K[3];
    return p;
}

/* pInv: camera<-world with the offset folded in, 1771-1793, using
 * CameraInfo::getExtrinsicInverse (camerautils.cpp:153-160): R' = R^T,
 * t' = (-R^T) t, then t'' = R'_i0 ox + R'_i1 oy + R'_i2 oz + t'_i in double. */
static campams make_pinv(const ffo_ctx *c, const double R[9], const double t[3], const double K[4])
{
    campams q;
    double Rt[9], tt[3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Rt[i * 3 + j] = R[j * 3 + i];
    for (int i = 0; i < 3; i++)
        tt[i] = ((-Rt[i * 3 + 0]) * t[0] + (-Rt[i * 3 + 1]) * t[1]) + (-Rt[i * 3 + 2]) * t[2];
    for (int i = 0; i < 9; i++) q.r[i] = (float)Rt[i];
    double o0 = c->offset[0], o1 = c->offset[1], o2 = c->offset[2];
    for (int i = 0; i < 3; i++)
        q.t[i] = (float)(((Rt[i * 3 + 0] * o0 + Rt[i * 3 + 1] * o1) + Rt[i * 3 + 2] * o2) + tt[i]);
    q.fx = (float)K[0]; q.fy = (float)K[1]; q.cx = (float)K[2]; q.cy = (float)K[3];
    return q;
}

void ffo_camera_inverse(const ffo_ctx *c, const double R[9], const double t[3], const double K[4],
                        float m[12], float Kf[4])
{
    campams q = make_pinv(c, R, t, K);
    for (int r = 0; r < 3; r++) {
        m[4 * r] = q.r[3 * r]; m[4 * r + 1] = q.r[3 * r + 1]; m[4 * r + 2] = q.r[3 * r + 2];
        m[4 * r + 3] = q.t[r];
    }
    Kf[0] = q.fx; Kf[1] = q.fy; Kf[2] = q.cx; Kf[3] = q.cy;
}

/* computeBoundingboxIntCPU, u16 overload, geometryfusion_mipmap_cpu.cpp:1270-1303 */
static void bbox_first(const campams *p, const uint16_t *depth, float scaling, float maxcam,
                       int w, int h, int bw, int mn[3], int mx[3])
{
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            float pz = (float)depth[y * w + x] * scaling;
            if (pz > 0.0f && pz < maxcam) {
                float px = ((float)x - p->cx) / p->fx * pz;
                float py = ((float)y - p->cy) / p->fy * pz;
                int v[3];
                for (int k = 0; k < 3; k++)
                    v[k] = (int)(((p->r[k * 3] * px + p->r[k * 3 + 1] * py) + p->r[k * 3 + 2] * pz) + p->t[k]);
                for (int k = 0; k < 3; k++) {
                    if (mn[k] > v[k] - bw) mn[k] = v[k] - bw;
                    if (mx[k] < v[k] + bw) mx[k] = v[k] + bw;
                }
            }
        }
}

/* setInitialVolume, geometryfusion_mipmap_cpu.cpp:1186-1202 */
static int set_initial_volume(ffo_ctx *c, const int mn[3], const int mx[3])
{
    if (mn[0] >= mx[0] || mn[1] >= mx[1] || mn[2] >= mx[2]) return 0;
    uint16_t hi[3], high;
    for (int k = 0; k < 3; k++) hi[k] = (uint16_t)(mx[k] - mn[k]);
    high = hi[0] > hi[1] ? hi[0] : hi[1];
    if (hi[2] > high) high = hi[2];
    uint16_t n;
    for (n = 2 * BL; n <= high; n *= 2) ;
    c->n = n;
    for (int k = 0; k < 3; k++) c->offset[k] = (float)(mn[k] - (n - hi[k]) / 2) * c->scale;
    return 1;
}

static void queue_leaf(ffo_ctx *c, uint32_t leaf)
{
    if (c->queueIndexOfLeaf[leaf] >= c->nQueued) {
        c->queueIndexOfLeaf[leaf] = c->nQueued;
        c->queue[c->nQueued++] = leaf;
    }
}

/* queryPointDepthSingle_func_subtree + addLeaf, traversal_functions.cpp:308-395, 185-211 */
static void query_point(ffo_ctx *c, uint16_t x, uint16_t y, uint16_t z, uint16_t target)
{
    uint16_t ox = 0, oy = 0, oz = 0, half = c->n >> 1;
    uint32_t branch = 0, pl = FFO_BRANCHINIT;
    while ((uint16_t)(half >> 1) >= target) {
        uint32_t co = (x >= ox + half) * 1u + (y >= oy + half) * 2u + (z >= oz + half) * 4u;
        uint32_t ptr = branch + co + 8;
        uint32_t child = c->tree[ptr];
        uint32_t leaf = c->tree[ptr - 8];
        if (leaf < c->nLeavesTotal) {
            pl = leaf;
            queue_leaf(c, leaf);
        }
        if (child == FFO_BRANCHINIT) {
            c->tree[ptr] = child = c->nBranchesUsed;
            c->nBranchesUsed += 16; /* CHILDBRANCHSIZEFUNCTION */
        }
        ox += half * ((co & 1) != 0);
        oy += half * ((co & 2) != 0);
        oz += half * ((co & 4) != 0);
        branch = child;
        half >>= 1;
    }
    uint32_t co = (x >= ox + half) * 1u + (y >= oy + half) * 2u + (z >= oz + half) * 4u;
    uint32_t ptr = branch + co;
    uint32_t child = c->tree[ptr];
    if (child >= FFO_BRANCHINIT) {
        uint32_t li = c->nLeavesUsed;
        c->tree[ptr] = child = li;
        if (li < c->nLeavesTotal) {
            c->leafPos[li * 3 + 0] = ox + half * ((co & 1) != 0);
            c->leafPos[li * 3 + 1] = oy + half * ((co & 2) != 0);
            c->leafPos[li * 3 + 2] = oz + half * ((co & 4) != 0);
            c->leafScale[li] = half / BL;
            c->leafParent[li] = pl;
        }
        c->nLeavesUsed++;
    }
    if (child < c->nLeavesTotal) queue_leaf(c, child);
}

/* The de-dup recurrence of transformLoopSimPrecalculatedNeg_subtree
 * (traversal_wrapper_functions.cpp:601-609, 633-639, 688-690). */
void ffo_dedup(const int32_t *vx, const int32_t *vy, const int32_t *vz, const uint8_t *valid,
               int w, int h, uint8_t *processed)
{
    int *tx = (int *)malloc(sizeof(int) * 3 * (size_t)w);
    for (int i = 0; i < 3 * w; i++) tx[i] = -1;
    for (int y = 0; y < h; y++) {
        int lx = -1, ly = -1, lz = -1;
        for (int x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            int kx = vx[i] & -8, ky = vy[i] & -8, kz = vz[i] & -8;
            if (valid[i] && (kx != tx[3 * x] || ky != tx[3 * x + 1] || kz != tx[3 * x + 2]) &&
                (kx != lx || ky != ly || kz != lz)) {
                tx[3 * x] = lx = kx; tx[3 * x + 1] = ly = ky; tx[3 * x + 2] = lz = kz;
                processed[i] = 1;
            } else {
                tx[3 * x] = tx[3 * x + 1] = tx[3 * x + 2] = lx = ly = lz = -1;
                processed[i] = 0;
            }
        }
    }
    free(tx);
}

/* transformLoopSimPrecalculatedNeg_subtree, u16 overload,
 * traversal_wrapper_functions.cpp:576-694 */
static void traverse(ffo_ctx *c, const float *qx1, const float *qx2, const float *qx3,
                     const float *qy1, const float *qy2, const float *qy3, const float t[3],
                     const uint16_t *data, int w, int h, float scaling, float maxcam,
                     int bmin[3], int bmax[3])
{
    int *tx = (int *)malloc(sizeof(int) * 3 * (size_t)w);
    for (int i = 0; i < 3 * w; i++) tx[i] = -1;
    const float *qx[3] = {qx1, qx2, qx3}, *qy[3] = {qy1, qy2, qy3};
    int n = c->n, bw = c->bandwidth;
    for (int y = 0; y < h; y++) {
        int lx = -1, ly = -1, lz = -1;
        for (int x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            float D = (float)data[i];
            int v[3];
            for (int k = 0; k < 3; k++) v[k] = (int)((((qx[k][x] + qy[k][y]) * D) * scaling) + t[k]);
            float z = D * scaling;
            int kx = v[0] & -8, ky = v[1] & -8, kz = v[2] & -8;
            if ((z > 0.0f && z < maxcam) && (kx != tx[3 * x] || ky != tx[3 * x + 1] || kz != tx[3 * x + 2]) &&
                (kx != lx || ky != ly || kz != lz)) {
                tx[3 * x] = lx = kx; tx[3 * x + 1] = ly = ky; tx[3 * x + 2] = lz = kz;
                c->processed[i] = 1;
                for (int k = 0; k < 3; k++) {
                    int lo = v[k] - bw, hi = v[k] + bw;
                    if (lo < 0 && bmin[k] > lo) bmin[k] = lo;
                    if (hi >= n && bmax[k] < hi) bmax[k] = hi;
                }
                if (v[0] >= 0 && (uint16_t)v[0] < n && v[1] >= 0 && (uint16_t)v[1] < n &&
                    v[2] >= 0 && (uint16_t)v[2] < n)
                    query_point(c, (uint16_t)v[0], (uint16_t)v[1], (uint16_t)v[2], ffo_leaf_scale_func(z));
            } else {
                tx[3 * x] = tx[3 * x + 1] = tx[3 * x + 2] = lx = ly = lz = -1;
                c->processed[i] = 0;
            }
        }
    }
    free(tx);
}

/* highestPowerOf2, geometryfusion_mipmap_cpu.cpp:2233-2237: the largest power
 * of two <= v, and 1 for v == 0. */
static uint16_t highest_pow2(uint16_t v)
{
    uint16_t r = 1;
    while (r * 2 <= v) r *= 2;
    return r;
}

/* shiftBranchesMipMap, treeandbrick.cpp:419-437 */
static void shift_branches(uint32_t *tree, uint32_t index, uint16_t side, uint32_t shift)
{
    uint16_t half = side / 2;
    for (uint32_t i = 0; i < 8; i++)
        if (half > BL && tree[index + 8 + i] != FFO_BRANCHINIT) {
            shift_branches(tree, tree[index + 8 + i], half, shift);
            tree[index + 8 + i] += shift;
        }
}

/* grow, geometryfusion_mipmap_cpu.cpp:3071-3319 (mesh-cell bookkeeping excluded) */
static void grow(ffo_ctx *c, const int bmin[3], const int bmax[3])
{
    int n = c->n;
    uint16_t p[3], m[3];
    for (int k = 0; k < 3; k++) {
        p[k] = (uint16_t)((bmax[k] >= n) * ((bmax[k] - n) / n + 1));
        m[k] = (uint16_t)((bmin[k] < 0) * ((-bmin[k]) / n + 1));
    }
    for (int k = 0; k < 3; k++) {
        uint16_t cut = p[k] & m[k];
        if (cut) {
            if (m[k] > p[k]) m[k] = (uint16_t)(2 * highest_pow2(m[k]) + (m[k] & (uint16_t)~cut));
            else p[k] = (uint16_t)(2 * highest_pow2(p[k]) + (p[k] & (uint16_t)~cut));
        }
    }
    uint16_t maxAdd = 0;
    for (int k = 0; k < 3; k++) {
        if (maxAdd < highest_pow2(m[k])) maxAdd = highest_pow2(m[k]);
        if (maxAdd < highest_pow2(p[k])) maxAdd = highest_pow2(p[k]);
    }
    uint32_t add = (uint32_t)maxAdd * 16;
    if (add > c->nBranchesTotal - c->nBranchesUsed) return;
    shift_branches(c->tree, 0, c->n, add);
    memmove(c->tree + add, c->tree, sizeof(uint32_t) * c->nBranchesUsed);
    for (uint32_t i = 0; i < add; i++) c->tree[i] = FFO_BRANCHINIT;
    c->nBranchesUsed += add;
    uint32_t index = (maxAdd - 1) * 16;
    uint16_t addv[3] = {0, 0, 0};
    for (uint16_t level = 0; level < maxAdd; level++) {
        unsigned dec = 1u << level;
        uint32_t co = 0;
        for (int k = 0; k < 3; k++)
            if ((m[k] & dec) || (!(p[k] & dec) && (p[k] > m[k]))) {
                addv[k] += c->n;
                co += 1u << k;
            }
        c->tree[index + co + 8] = index + 16;
        c->n *= 2;
        index -= 16;
    }
    for (int k = 0; k < 3; k++) c->offset[k] -= (float)addv[k] * c->scale;
    for (uint32_t i = 0; i < c->nLeavesUsed && i < c->nLeavesTotal; i++) {
        c->leafPos[i * 3 + 0] += addv[0];
        c->leafPos[i * 3 + 1] += addv[1];
        c->leafPos[i * 3 + 2] += addv[2];
    }
    c->grows++;
}

/* Converts like CVTPS2DQ under round-toward-zero: truncation, and the
 * integer-indefinite value for NaN and out-of-range inputs. */
static int cvt_rz(float v)
{
    if (!(v >= -2147483648.0f && v < 2147483648.0f)) return (int)0x80000000u;
    return (int)v;
}

/* update8AddLoopAVXSingleInteger, update_functions.cpp:1187-1444, one lane at
 * a time (the AVX2 code is lane-wise).  Must run with the rounding mode set
 * to round-toward-zero. */
static void update_brick_rz(const uint16_t *depth, float scaling, float maxcam, const uint8_t *rgb,
                            int W, int H, const float *m, float fx, float fy, float cx, float cy,
                            float scale, float dT, const uint16_t *o, uint16_t leafScale,
                            float *dist, float *wgt, uint16_t *col, const float *conf)
{
    volatile float vs = scale; /* keep the compiler from folding across fesetround */
    float s = vs;
    float fls = (float)leafScale * s;
    float ox = ((m[0] * (float)o[0] + m[1] * (float)o[1]) + m[2] * (float)o[2]) * s + m[3];
    float oy = ((m[4] * (float)o[0] + m[5] * (float)o[1]) + m[6] * (float)o[2]) * s + m[7];
    float oz = ((m[8] * (float)o[0] + m[9] * (float)o[1]) + m[10] * (float)o[2]) * s + m[11];
    float d11 = m[0] * fls, d21 = m[4] * fls, d31 = m[8] * fls;
    float d12 = m[1] * fls, d22 = m[5] * fls, d32 = m[9] * fls;
    float d13 = m[2] * fls, d23 = m[6] * fls, d33 = m[10] * fls;
    float l1[8], l2[8], l3[8];
    l1[0] = l2[0] = l3[0] = 0.0f;
    for (int k = 1; k < 8; k++) {
        l1[k] = (float)k * d11;
        l2[k] = (float)k * d21;
        l3[k] = (float)k * d31;
    }
    float td = dT * (float)leafScale;
    float tw = (WEIGHT_FACTOR * dT) * (float)leafScale;
    float rfall = ffo_rcp(tw - EPS_W);
    float pzx = ox, pzy = oy, pzz = oz;
    int idx = 0;
    for (int z = 0; z < 8; z++, pzx += d13, pzy += d23, pzz += d33) {
        float pyx = pzx, pyy = pzy, pyz = pzz;
        for (int y = 0; y < 8; y++, pyx += d12, pyy += d22, pyz += d32) {
            for (int k = 0; k < 8; k++, idx++) {
                float px = pyx + l1[k], py = pyy + l2[k], pz = pyz + l3[k];
                float len = sqrtf((px * px + py * py) + pz * pz);
                float r = ffo_rcp(pz);
                int imx = cvt_rz(cx + fx * (px * r));
                int imy = cvt_rz(cy + fy * (py * r));
                int cyi = imy < H - 1 ? imy : H - 1; if (cyi < 0) cyi = 0;
                int cxi = imx < W - 1 ? imx : W - 1; if (cxi < 0) cxi = 0;
                int ii = W * cyi + cxi;
                float h8 = (float)depth[ii];
                float h = scaling * h8;
                float dI = len - (len * r) * h;
                float wA = wgt[idx], dA = dist[idx];
                int mask = (pz > 0.0f) && (imy < H) && (imy > -1) && (imx < W) && (imx > -1) && (h < maxcam);
                float front = (dI < EPS_W) ? 1.0f : 0.0f;
                float back = (dI >= EPS_W && dI < tw) ? 1.0f : 0.0f;
                float fall = (tw - dI) * rfall;
                float wI = mask ? (front + fall * back) : 0.0f;
                /* learned confidence: a factor on the reference's weight */
                if (conf) wI = wI * conf[ii];
                float wN = wA + wI;
                float f = ffo_rcp(wN + ((wN <= 0.0f) ? 1.0f : 0.0f));
                float mn = (td < dI) ? td : dI;         /* _mm256_min_ps(td, dI) */
                float mx = (-1.0f * td > mn) ? -1.0f * td : mn; /* _mm256_max_ps(-td, mn) */
                float add = mask ? mx * wI : 0.0f;
                dist[idx] = ((dA * wA) + add) * f;
                wgt[idx] = wN;
                if (col) {
                    const float cmax = 255.0f * COLOR_MUL;
                    for (int ch = 0; ch < 3; ch++) {
                        float inc = (float)rgb[ii * 3 + (2 - ch)] * COLOR_MUL;
                        float acc = (float)col[idx * 3 + ch];
                        float v = ((acc * wA) + (inc * wI)) * f;
                        v = (v < cmax) ? v : cmax;
                        col[idx * 3 + ch] = (uint16_t)cvt_rz(v);
                    }
                }
            }
        }
    }
}

void ffo_update_brick(const uint16_t *depth, float scaling, float maxcam, const uint8_t *bgr,
                      int w, int h, const float m[12], const float K[4], float scale,
                      float distance_threshold, const uint16_t pos[3], uint16_t leaf_scale,
                      float *dist, float *weight, uint16_t *color)
{
    int old = fegetround();
    fesetround(FE_TOWARDZERO);
    update_brick_rz(depth, scaling, maxcam, bgr, w, h, m, K[0], K[1], K[2], K[3], scale,
                    distance_threshold, pos, leaf_scale, dist, weight, color, NULL);
    fesetround(old);
}

/* FusionMipMapCPU::addMap, u16 overload, geometryfusion_mipmap_cpu.cpp:1603-2034
 * (synchronous: no distance-update thread; mesh-cell bookkeeping is separate). */
int ffo_add_map(ffo_ctx *c, const uint16_t *depth, const uint8_t *bgr, int w, int h,
                const double R[9], const double t[3], const double K[4],
                float scaling, float maxcam)
{
    campams p = make_p(c, R, t, K);
    for (uint32_t i = 0; i < c->nLeavesTotal; i++) c->queueIndexOfLeaf[i] = FFO_MAXLEAFNUMBER;
    c->nQueued = 0;

    if (!c->n) {
        int mn[3] = {1000000, 1000000, 1000000}, mx[3] = {-1000000, -1000000, -1000000};
        bbox_first(&p, depth, scaling, maxcam, w, h, c->bandwidth, mn, mx);
        set_initial_volume(c, mn, mx);
        p = make_p(c, R, t, K);
    }
    if (c->imgW != w || c->imgH != h) {
        c->imgW = w; c->imgH = h;
        free(c->pxp); free(c->pyp); free(c->processed);
        c->pxp = (float *)malloc(sizeof(float) * w);
        c->pyp = (float *)malloc(sizeof(float) * h);
        c->processed = (uint8_t *)calloc((size_t)w * h, 1);
        for (int x = 0; x < w; x++) c->pxp[x] = ((float)x - p.cx) / p.fx;
        for (int y = 0; y < h; y++) c->pyp[y] = ((float)y - p.cy) / p.fy;
    }
    int bmin[3] = {0, 0, 0}, bmax[3] = {c->n - 1, c->n - 1, c->n - 1};
    campams q = make_pinv(c, R, t, K);
    float mm[12] = {q.r[0], q.r[1], q.r[2], q.t[0], q.r[3], q.r[4], q.r[5], q.t[1],
                    q.r[6], q.r[7], q.r[8], q.t[2]};

    float *qx = (float *)malloc(sizeof(float) * 3 * w), *qy = (float *)malloc(sizeof(float) * 3 * h);
    for (int x = 0; x < w; x++) {
        qx[x] = p.r[0] * c->pxp[x];
        qx[w + x] = p.r[3] * c->pxp[x];
        qx[2 * w + x] = p.r[6] * c->pxp[x];
    }
    for (int y = 0; y < h; y++) {
        qy[y] = p.r[1] * c->pyp[y] + p.r[2];
        qy[h + y] = p.r[4] * c->pyp[y] + p.r[5];
        qy[2 * h + y] = p.r[7] * c->pyp[y] + p.r[8];
    }
    if (c->n)
        traverse(c, qx, qx + w, qx + 2 * w, qy, qy + h, qy + 2 * h, p.t, depth, w, h, scaling, maxcam,
                 bmin, bmax);
    else
        memset(c->processed, 0, (size_t)w * h);
    free(qx); free(qy);

    if (c->nLeavesUsed < c->nLeavesTotal && c->nBranchesUsed < c->nBranchesTotal &&
        (bmin[0] < 0 || bmin[1] < 0 || bmin[2] < 0 || bmax[0] >= c->n || bmax[1] >= c->n || bmax[2] >= c->n))
        grow(c, bmin, bmax);

    int old = fegetround();
    fesetround(FE_TOWARDZERO);
    for (uint32_t l = 0; l < c->nQueued; l++) {
        uint32_t b = c->queue[l];
        update_brick_rz(depth, scaling, maxcam, bgr, w, h, mm, q.fx, q.fy, q.cx, q.cy, c->scale, c->dT,
                        c->leafPos + (size_t)b * 3, c->leafScale[b], c->dist + (size_t)b * BRICK,
                        c->wgt + (size_t)b * BRICK, c->use_color ? c->col + (size_t)b * BRICK * 3 : NULL, c->conf);
    }
    fesetround(old);
    return (int)c->nQueued;
}

uint32_t ffo_n_leaves(const ffo_ctx *c) { return c->nLeavesUsed; }

/* Per-pixel confidence multiplying the weight of the following frames
 * (W x H floats that stay valid while used), or NULL for the plain method. */
void ffo_set_confidence(ffo_ctx *c, const float *conf) { c->conf = conf; }
uint32_t ffo_n_branches(const ffo_ctx *c) { return c->nBranchesUsed; }
uint32_t ffo_side(const ffo_ctx *c) { return c->n; }
void ffo_offset(const ffo_ctx *c, float out[3]) { memcpy(out, c->offset, sizeof(float) * 3); }
uint32_t ffo_n_queued(const ffo_ctx *c) { return c->nQueued; }
const uint32_t *ffo_queue(const ffo_ctx *c) { return c->queue; }
const uint32_t *ffo_tree(const ffo_ctx *c) { return c->tree; }
const uint16_t *ffo_leaf_pos(const ffo_ctx *c) { return c->leafPos; }
const uint16_t *ffo_leaf_scale(const ffo_ctx *c) { return c->leafScale; }
const uint32_t *ffo_leaf_parent(const ffo_ctx *c) { return c->leafParent; }
const float *ffo_distance(const ffo_ctx *c) { return c->dist; }
const float *ffo_weights(const ffo_ctx *c) { return c->wgt; }
const uint16_t *ffo_color(const ffo_ctx *c) { return c->col; }
const uint8_t *ffo_processed(const ffo_ctx *c) { return c->processed; }
int ffo_grow_count(const ffo_ctx *c) { return c->grows; }
uint32_t ffo_leaves_total(const ffo_ctx *c) { return c->nLeavesTotal; }
float ffo_scale(const ffo_ctx *c) { return c->scale; }
float ffo_threshold(const ffo_ctx *c) { return c->dT; }
