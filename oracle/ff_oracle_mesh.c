/*
 * ff_oracle_mesh.c -- CPU restatement of fastfusion's per-cell marching cubes.
 *
 * TEST INFRASTRUCTURE ONLY (parity oracle and CPU baseline).
 *
 * Cells.  The reference partitions space into "mesh cells"
 * (src/fusion/meshcelltraversalSplitReferenced.hpp:6-790, 2568-5913,
 * 9123-9282): for every octree node without a child branch ("terminal
 * node") an interior cell, and between neighbouring terminal nodes wall,
 * edge and corner ("middle") cells, split to the size of the finest node
 * they touch.  Each cell side k holds lastLeaf[k], the finest brick covering
 * that side (createMeshCellStructuresForLeaf, 9123-9282: a finer brick
 * replaces a coarser one).  This file restates that structure directly from
 * the tree: it enumerates terminal nodes, and for each node the cells it is
 * the canonical owner of -- the smallest node adjacent to the cell, ties
 * going to the lowest lastLeaf slot.  A cell (type, o, size) lies at o + size
 * on its thin axes and spans [o, o + size) on its full axes.  lastLeaf slot
 * bits: for walls bit 0 is the + side of the normal axis; for edges along x /
 * y / z the bits are (y, z) / (x, z) / (x, y); for corners (x, y, z).
 *
 * Meshing.  Each cell is meshed from scratch like
 * addBrick*_indexed_incremental (src/fusion/mesh_interleaved_meshcell.cpp):
 * samples taken from lastLeaf with a fallback along the leaf-parent chain
 * while the accumulated weight is 0; interior samples interpolated
 * trilinearly (weight of the lower corner only), wall samples bilinearly,
 * edge samples linearly (weight = max of the two), corner samples point-
 * wise; a cube is skipped when any corner weight is 0 (mesh.hpp:259-291);
 * vertices are placed with the double-precision VertexInterp
 * (mesh.cpp:1266-1339) and colored with its float twin; middle cells are
 * non-indexed and drop degenerate triangles (mesh_interleaved_meshcell.cpp:
 * 2480-2679).  Vertices are emitted per sample point in (z, y, x) raster
 * order, x-, y-, z-edge, faces per cube in raster order.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ff_oracle.h"
#include "../fastfusion_amd/csrc/mc_tables.h"

#define BL 8
#define BRICK 512

/* the deepest parent-chain level (1: the lastLeaf itself) any sample of the
 * cell being meshed read; the incremental update keeps it per cell */
static int g_sample_depth;
/* the most chain levels the incremental update checks: 16 (all) except
 * where a test plants a mistake (ffo_mesh_set_depth_cap) */
static int g_depth_cap = 16;
static void note_depth(int lv)
{
    if (lv > g_sample_depth) g_sample_depth = lv;
}

/* the context layout is private to ff_oracle.c; accessors are used instead */

typedef struct {
    int type;
    int o[3];
    int size;
    uint32_t ll[8];
} cell_t;

typedef struct {
    cell_t *cells;
    size_t n, cap;
} cellvec;

typedef struct {
    float *xyz;
    uint8_t *rgb;
    uint32_t *faces;
    size_t nv, nf, capv, capf;
} meshbuf;

typedef struct {
    const uint32_t *tree;
    const uint16_t *pos;
    const uint16_t *scl;
    const uint32_t *par;
    const float *dist, *wgt;
    const uint16_t *col;
    uint32_t nLeavesTotal;
    uint32_t n;
    float off[3];
    float scale;
} mapview;

static int tri_len(int ci) { return (int)strlen(FF_MC_TRI[ci]); }
static int tri_edge(int ci, int k)
{
    char ch = FF_MC_TRI[ci][k];
    return ch <= '9' ? ch - '0' : ch - 'a' + 10;
}

/* ---------------- tree queries ---------------- */

/* Locates the terminal node containing voxel p; returns its size (0 when p
 * is outside the volume), origin, and the deepest brick on the path. */
static uint32_t locate(const mapview *m, const int p[3], int org[3], uint32_t *leaf)
{
    *leaf = FFO_BRANCHINIT;
    if (p[0] < 0 || p[1] < 0 || p[2] < 0 || p[0] >= (int)m->n || p[1] >= (int)m->n || p[2] >= (int)m->n) return 0;
    uint32_t half = m->n >> 1, branch = 0;
    int o[3] = {0, 0, 0};
    for (;;) {
        uint32_t co = 0;
        for (int k = 0; k < 3; k++)
            if (p[k] >= o[k] + (int)half) { co |= 1u << k; o[k] += half; }
        uint32_t l = m->tree[branch + co];
        if (l < m->nLeavesTotal) *leaf = l;
        uint32_t child = m->tree[branch + co + 8];
        if (child == FFO_BRANCHINIT || half <= BL) {
            memcpy(org, o, sizeof o);
            return half;
        }
        branch = child;
        half >>= 1;
    }
}

static void push_cell(cellvec *v, const cell_t *c)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 1024;
        v->cells = (cell_t *)realloc(v->cells, v->cap * sizeof(cell_t));
    }
    v->cells[v->n++] = *c;
}

/* cells owned by terminal node (p, s) with covering brick lt */
static void node_cells(const mapview *m, const int p[3], int s, uint32_t lt, cellvec *out)
{
    cell_t c;
    int q[3], org[3];
    uint32_t lf;
    /* interior */
    memset(&c, 0, sizeof c);
    c.type = 0; memcpy(c.o, p, sizeof c.o); c.size = s;
    for (int k = 0; k < 8; k++) c.ll[k] = FFO_BRANCHINIT;
    c.ll[0] = lt;
    push_cell(out, &c);
    /* walls: normal axis a; the node is on side 1 for its - face, side 0 for its + face */
    for (int a = 0; a < 3; a++)
        for (int side = 0; side < 2; side++) {
            memcpy(q, p, sizeof q);
            q[a] = side ? p[a] + s : p[a] - 1;
            uint32_t ns = locate(m, q, org, &lf);
            if (!ns) continue;
            /* owner: smallest node, ties to slot 0 (the - side) */
            int own = side ? (ns >= (uint32_t)s) : (ns > (uint32_t)s);
            if (!own) continue;
            for (int k = 0; k < 8; k++) c.ll[k] = FFO_BRANCHINIT;
            c.type = 1 + a; c.size = s;
            memcpy(c.o, p, sizeof c.o);
            if (!side) c.o[a] = p[a] - s;
            c.ll[side ? 0 : 1] = lt;
            c.ll[side ? 1 : 0] = lf;
            push_cell(out, &c);
        }
    /* edges along axis e, thin axes (u, v); slot = ubit + 2 vbit */
    static const int eax[3][2] = {{1, 2}, {0, 2}, {0, 1}};
    for (int e = 0; e < 3; e++) {
        const int u = eax[e][0], v = eax[e][1];
        for (int cu = 0; cu < 2; cu++)
            for (int cv = 0; cv < 2; cv++) {
                /* the edge line at (p_u + cu s, p_v + cv s) */
                const int lu = p[u] + cu * s, lv = p[v] + cv * s;
                const int mine = (1 - cu) + 2 * (1 - cv);
                uint32_t sz[4], lfs[4];
                int ok = 1;
                for (int sl = 0; sl < 4 && ok; sl++) {
                    if (sl == mine) { sz[sl] = s; lfs[sl] = lt; continue; }
                    memcpy(q, p, sizeof q);
                    q[u] = lu - 1 + (sl & 1);
                    q[v] = lv - 1 + (sl >> 1);
                    sz[sl] = locate(m, q, org, &lfs[sl]);
                    if (!sz[sl]) ok = 0;
                }
                if (!ok) continue;
                uint32_t mn = sz[0];
                for (int sl = 1; sl < 4; sl++) if (sz[sl] < mn) mn = sz[sl];
                if (mn != (uint32_t)s) continue;
                int first = 0;
                while (sz[first] != mn) first++;
                if (first != mine) continue;
                for (int k = 0; k < 8; k++) c.ll[k] = FFO_BRANCHINIT;
                c.type = 4 + e; c.size = s;
                c.o[e] = p[e]; c.o[u] = lu - s; c.o[v] = lv - s;
                for (int sl = 0; sl < 4; sl++) c.ll[sl] = lfs[sl];
                push_cell(out, &c);
            }
    }
    /* corners: point p + s (cx, cy, cz); slot bits (x, y, z) = + side */
    for (int cz = 0; cz < 2; cz++)
        for (int cy = 0; cy < 2; cy++)
            for (int cx = 0; cx < 2; cx++) {
                const int pt[3] = {p[0] + cx * s, p[1] + cy * s, p[2] + cz * s};
                const int mine = (1 - cx) + 2 * (1 - cy) + 4 * (1 - cz);
                uint32_t sz[8], lfs[8];
                int ok = 1;
                for (int sl = 0; sl < 8 && ok; sl++) {
                    if (sl == mine) { sz[sl] = s; lfs[sl] = lt; continue; }
                    for (int k = 0; k < 3; k++) q[k] = pt[k] - 1 + ((sl >> k) & 1);
                    sz[sl] = locate(m, q, org, &lfs[sl]);
                    if (!sz[sl]) ok = 0;
                }
                if (!ok) continue;
                uint32_t mn = sz[0];
                for (int sl = 1; sl < 8; sl++) if (sz[sl] < mn) mn = sz[sl];
                if (mn != (uint32_t)s) continue;
                int first = 0;
                while (sz[first] != mn) first++;
                if (first != mine) continue;
                c.type = 7; c.size = s;
                for (int k = 0; k < 3; k++) c.o[k] = pt[k] - s;
                for (int sl = 0; sl < 8; sl++) c.ll[sl] = lfs[sl];
                push_cell(out, &c);
            }
}

static void walk(const mapview *m, uint32_t branch, const int o[3], uint32_t half, uint32_t inh, cellvec *out)
{
    for (uint32_t co = 0; co < 8; co++) {
        int p[3];
        for (int k = 0; k < 3; k++) p[k] = o[k] + (int)(((co >> k) & 1) * half);
        uint32_t l = m->tree[branch + co];
        uint32_t cov = l < m->nLeavesTotal ? l : inh;
        uint32_t child = m->tree[branch + co + 8];
        if (child == FFO_BRANCHINIT || half <= BL) {
            if (cov != FFO_BRANCHINIT) node_cells(m, p, (int)half, cov, out);
        } else {
            walk(m, child, p, half >> 1, cov, out);
        }
    }
}

static int required_present(const cell_t *c)
{
    static const int need[8] = {1, 2, 2, 2, 4, 4, 4, 8};
    for (int k = 0; k < need[c->type]; k++)
        if (c->ll[k] >= FFO_BRANCHINIT) return 0;
    return 1;
}

static uint64_t cell_key(const cell_t *c)
{
    int lg = 0;
    while ((1 << lg) < c->size) lg++;
    return ((uint64_t)(uint16_t)c->o[0] << 48) | ((uint64_t)(uint16_t)c->o[1] << 32) |
           ((uint64_t)(uint16_t)c->o[2] << 16) | ((uint64_t)lg << 3) | (uint64_t)c->type;
}

static int cmp_cells(const void *a, const void *b)
{
    uint64_t ka = cell_key((const cell_t *)a), kb = cell_key((const cell_t *)b);
    return ka < kb ? -1 : ka > kb ? 1 : 0;
}

/* ---------------- sampling ---------------- */

typedef struct { float d, w; uint16_t c[3]; } sample_t;

static int clampi(int v) { return v < 0 ? 0 : v > BL - 1 ? BL - 1 : v; }

/* axis decomposition of a sample coordinate inside a leaf (lo, ls):
 * low index, fraction, high index (mesh_interleaved_meshcell.cpp:130-147) */
static void split_axis(int X, int lo, int ls, int *il, float *r, int *ih)
{
    int l = (X - lo) / ls, rem = (X - lo) % ls;
    float rr = (float)rem / (float)ls;
    *il = clampi(l);
    *r = rr;
    *ih = clampi(l + (rr > 0.0 && l < BL - 1));
}

static uint16_t f2u16(float v) { return (uint16_t)(int)v; }

/* interior sample at voxel X (src/fusion/mesh_interleaved_meshcell.cpp:95-260) */
static void sample_interior(const mapview *m, uint32_t ll, const int X[3], sample_t *s)
{
    s->w = 0.0f; s->d = 0.0f; s->c[0] = s->c[1] = s->c[2] = 0;
    int lv = 0;
    for (uint32_t leaf = ll; leaf < FFO_BRANCHINIT && s->w <= 0.0f; leaf = m->par[leaf]) {
        lv++;
        const uint16_t *lo = m->pos + 3 * (size_t)leaf;
        int ls = m->scl[leaf];
        int il[3], ih[3];
        float r[3];
        for (int k = 0; k < 3; k++) split_axis(X[k], lo[k], ls, &il[k], &r[k], &ih[k]);
        float rxn = 1.0f - r[0], ryn = 1.0f - r[1], rzn = 1.0f - r[2];
        float f[8] = {rzn * ryn * rxn, rzn * ryn * r[0], rzn * r[1] * rxn, rzn * r[1] * r[0],
                      r[2] * ryn * rxn, r[2] * ryn * r[0], r[2] * r[1] * rxn, r[2] * r[1] * r[0]};
        int idx[8];
        for (int i = 0; i < 8; i++) {
            int ix = (i & 1) ? ih[0] : il[0], iy = (i & 2) ? ih[1] : il[1], iz = (i & 4) ? ih[2] : il[2];
            idx[i] = (iz * BL + iy) * BL + ix;
        }
        size_t st = (size_t)leaf * BRICK;
        float wf[8], df[8], cf[3][8];
        for (int i = 0; i < 8; i++) {
            wf[i] = m->wgt[st + idx[i]];
            df[i] = m->dist[st + idx[i]];
            for (int ch = 0; ch < 3; ch++) cf[ch][i] = (float)m->col[(st + idx[i]) * 3 + ch];
        }
        s->w += wf[0];
        /* SSE: s_i = f_i v_i + f_{i+4} v_{i+4}; hadd; (s0 + s1) + (s2 + s3) */
        float a0 = f[0] * df[0] + f[4] * df[4], a1 = f[1] * df[1] + f[5] * df[5];
        float a2 = f[2] * df[2] + f[6] * df[6], a3 = f[3] * df[3] + f[7] * df[7];
        s->d = (a0 + a1) + (a2 + a3);
        for (int ch = 0; ch < 3; ch++) {
            float b0 = f[0] * cf[ch][0] + f[4] * cf[ch][4], b1 = f[1] * cf[ch][1] + f[5] * cf[ch][5];
            float b2 = f[2] * cf[ch][2] + f[6] * cf[ch][6], b3 = f[3] * cf[ch][3] + f[7] * cf[ch][7];
            s->c[ch] = f2u16((b0 + b1) + (b2 + b3));
        }
    }
    note_depth(lv);
}

/* wall sample: thin axis a at X[a] (no interpolation), full axes u (inner)
 * and v (outer) interpolated bilinearly (mesh_interleaved_meshcell.cpp:700-760) */
static void sample_wall(const mapview *m, uint32_t ll, int a, int u, int v, const int X[3], sample_t *s)
{
    s->w = 0.0f; s->d = 0.0f; s->c[0] = s->c[1] = s->c[2] = 0;
    int lv = 0;
    for (uint32_t leaf = ll; leaf < FFO_BRANCHINIT && s->w <= 0.0f; leaf = m->par[leaf]) {
        lv++;
        const uint16_t *lo = m->pos + 3 * (size_t)leaf;
        int ls = m->scl[leaf];
        int la = clampi((X[a] - lo[a]) / ls);
        int ul, uh, vl, vh;
        float ru, rv;
        split_axis(X[u], lo[u], ls, &ul, &ru, &uh);
        split_axis(X[v], lo[v], ls, &vl, &rv, &vh);
        int id[4][3];
        for (int i = 0; i < 4; i++) {
            id[i][a] = la;
            id[i][u] = (i & 1) ? uh : ul;
            id[i][v] = (i & 2) ? vh : vl;
        }
        size_t st = (size_t)leaf * BRICK;
        float wf[4], df[4], cf[3][4];
        for (int i = 0; i < 4; i++) {
            size_t q = st + (size_t)((id[i][2] * BL + id[i][1]) * BL + id[i][0]);
            wf[i] = m->wgt[q]; df[i] = m->dist[q];
            for (int ch = 0; ch < 3; ch++) cf[ch][i] = (float)m->col[q * 3 + ch];
        }
        float ruI = 1.0f - ru, rvI = 1.0f - rv;
        s->w += rvI * ruI * wf[0] + rvI * ru * wf[1] + rv * ruI * wf[2] + rv * ru * wf[3];
        s->d = rvI * ruI * df[0] + rvI * ru * df[1] + rv * ruI * df[2] + rv * ru * df[3];
        for (int ch = 0; ch < 3; ch++)
            s->c[ch] = f2u16(rvI * ruI * cf[ch][0] + rvI * ru * cf[ch][1] + rv * ruI * cf[ch][2] + rv * ru * cf[ch][3]);
    }
    note_depth(lv);
}

/* edge sample: full axis e interpolated linearly, weight = max of the two
 * (mesh_interleaved_meshcell.cpp:1800-1860) */
static void sample_edge(const mapview *m, uint32_t ll, int e, const int X[3], sample_t *s)
{
    s->w = 0.0f; s->d = 0.0f; s->c[0] = s->c[1] = s->c[2] = 0;
    int lv = 0;
    for (uint32_t leaf = ll; leaf < FFO_BRANCHINIT && s->w <= 0.0f; leaf = m->par[leaf]) {
        lv++;
        const uint16_t *lo = m->pos + 3 * (size_t)leaf;
        int ls = m->scl[leaf];
        int el, eh;
        float re;
        split_axis(X[e], lo[e], ls, &el, &re, &eh);
        int id0[3], id1[3];
        for (int k = 0; k < 3; k++) id0[k] = id1[k] = clampi((X[k] - lo[k]) / ls);
        id0[e] = el; id1[e] = eh;
        size_t st = (size_t)leaf * BRICK;
        size_t q0 = st + (size_t)((id0[2] * BL + id0[1]) * BL + id0[0]);
        size_t q1 = st + (size_t)((id1[2] * BL + id1[1]) * BL + id1[0]);
        float w0 = m->wgt[q0], w1 = m->wgt[q1];
        s->w += (w0 < w1) ? w1 : w0; /* std::max */
        s->d = (1.0f - re) * m->dist[q0] + re * m->dist[q1];
        for (int ch = 0; ch < 3; ch++)
            s->c[ch] = f2u16((1.0f - re) * (float)m->col[q0 * 3 + ch] + re * (float)m->col[q1 * 3 + ch]);
    }
    note_depth(lv);
}

/* corner sample: point-wise (mesh_interleaved_meshcell.cpp:2550-2575) */
static void sample_point(const mapview *m, uint32_t ll, const int X[3], sample_t *s)
{
    s->w = 0.0f; s->d = 0.0f; s->c[0] = s->c[1] = s->c[2] = 0;
    int lv = 0;
    for (uint32_t leaf = ll; leaf < FFO_BRANCHINIT && s->w <= 0.0f; leaf = m->par[leaf]) {
        lv++;
        const uint16_t *lo = m->pos + 3 * (size_t)leaf;
        int ls = m->scl[leaf];
        size_t q = (size_t)leaf * BRICK;
        int id[3];
        for (int k = 0; k < 3; k++) id[k] = clampi((X[k] - lo[k]) / ls);
        q += (size_t)((id[2] * BL + id[1]) * BL + id[0]);
        s->w += m->wgt[q];
        s->d = m->dist[q];
        for (int ch = 0; ch < 3; ch++) s->c[ch] = m->col[q * 3 + ch];
    }
    note_depth(lv);
}

/* ---------------- interpolation (mesh.cpp:1266-1339) ---------------- */
static void vinterp(const float p1[3], const float p2[3], double v1, double v2, float out[3])
{
    if (fabs(0.0 - v1) < 1e-40) { memcpy(out, p1, 3 * sizeof(float)); return; }
    if (fabs(0.0 - v2) < 1e-40) { memcpy(out, p2, 3 * sizeof(float)); return; }
    if (fabs(v1 - v2) < 1e-40) { memcpy(out, p1, 3 * sizeof(float)); return; }
    double mu = (0.0 - v1) / (v2 - v1);
    for (int k = 0; k < 3; k++) out[k] = (float)((double)p1[k] + mu * ((double)p2[k] - (double)p1[k]));
}

static void cinterp(const uint16_t c1[3], const uint16_t c2[3], float v1, float v2, uint8_t out[3])
{
    uint16_t r[3];
    if ((double)fabsf(0.0f - v1) < 1e-40) memcpy(r, c1, sizeof r);
    else if ((double)fabsf(0.0f - v2) < 1e-40) memcpy(r, c2, sizeof r);
    else if ((double)fabsf(v1 - v2) < 1e-40) memcpy(r, c1, sizeof r);
    else {
        float mu = (0.0f - v1) / (v2 - v1);
        for (int k = 0; k < 3; k++) {
            float p = (float)c1[k] + mu * ((float)c2[k] - (float)c1[k]);
            r[k] = p <= 65280.0f ? (p >= 0.0f ? (uint16_t)(int)p : 0) : 65280;
        }
    }
    for (int k = 0; k < 3; k++) out[k] = (uint8_t)(r[k] / 256);
}

static void mb_vertex(meshbuf *mb, const float p[3], const uint8_t c[3])
{
    if (mb->nv == mb->capv) {
        mb->capv = mb->capv ? 2 * mb->capv : 4096;
        mb->xyz = (float *)realloc(mb->xyz, mb->capv * 3 * sizeof(float));
        mb->rgb = (uint8_t *)realloc(mb->rgb, mb->capv * 3);
    }
    memcpy(mb->xyz + 3 * mb->nv, p, 3 * sizeof(float));
    memcpy(mb->rgb + 3 * mb->nv, c, 3);
    mb->nv++;
}

static void mb_face(meshbuf *mb, uint32_t a, uint32_t b, uint32_t c)
{
    if (mb->nf == mb->capf) {
        mb->capf = mb->capf ? 2 * mb->capf : 4096;
        mb->faces = (uint32_t *)realloc(mb->faces, mb->capf * 3 * sizeof(uint32_t));
    }
    mb->faces[3 * mb->nf] = a; mb->faces[3 * mb->nf + 1] = b; mb->faces[3 * mb->nf + 2] = c;
    mb->nf++;
}

/* Bourke corner k of the cube at grid (i, j, l) -> grid offsets */
static const int CORNER[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
/* Bourke edge -> (grid offset, axis) */
static const int EDGE[12][4] = {{0, 0, 0, 0}, {1, 0, 0, 1}, {0, 1, 0, 0}, {0, 0, 0, 1}, {0, 0, 1, 0}, {1, 0, 1, 1},
                                {0, 1, 1, 0}, {0, 0, 1, 1}, {0, 0, 0, 2}, {1, 0, 0, 2}, {1, 1, 0, 2}, {0, 1, 0, 2}};
static const int EDGE_CORNERS[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6}, {6, 7}, {7, 4},
                                        {0, 4}, {1, 5}, {2, 6}, {3, 7}};

/* Meshes one cell; appends to mb; returns 0 (empty mesh on the reference's
 * early-outs: a leaf larger than the cell, or an interior cell that does not
 * fit its leaf). */
static void mesh_cell(const mapview *m, const cell_t *c, meshbuf *mb)
{
    const uint32_t *L = c->ll;
    const size_t cell_v0 = mb->nv; /* faces index vertices relative to the cell */
    int type = c->type, size = c->size;
    int sizeMin, stretch = 0;
    int n[3];          /* grid points per axis */
    int coord[3][8];   /* voxel coordinate of each grid index per axis */
    int elen[3];       /* edge length per axis */
    if (type == 0) {
        sizeMin = m->scl[L[0]];
        if (sizeMin > size || size / sizeMin > BL) return;
        const uint16_t *lp = m->pos + 3 * (size_t)L[0];
        int ext = m->scl[L[0]] * BL;
        for (int k = 0; k < 3; k++)
            if (c->o[k] + size > lp[k] + ext || c->o[k] < lp[k]) return;
        for (int k = 0; k < 3; k++) {
            n[k] = size / sizeMin;
            for (int i = 0; i < n[k]; i++) coord[k][i] = c->o[k] + i * sizeMin;
            elen[k] = sizeMin;
        }
    } else if (type <= 3) {
        int a = type - 1;
        sizeMin = m->scl[L[0]] < m->scl[L[1]] ? m->scl[L[0]] : m->scl[L[1]];
        stretch = m->scl[L[0]];
        if (stretch > size) return;
        for (int k = 0; k < 3; k++) {
            if (k == a) {
                n[k] = 2;
                coord[k][0] = c->o[k] + size - stretch;
                coord[k][1] = c->o[k] + size;
                elen[k] = stretch;
            } else {
                n[k] = size / sizeMin;
                for (int i = 0; i < n[k]; i++) coord[k][i] = c->o[k] + i * sizeMin;
                elen[k] = sizeMin;
            }
        }
    } else if (type <= 6) {
        int e = type - 4;
        stretch = m->scl[L[0]];
        if (m->scl[L[1]] > stretch) stretch = m->scl[L[1]];
        if (m->scl[L[2]] > stretch) stretch = m->scl[L[2]];
        sizeMin = m->scl[L[0]];
        for (int k = 1; k < 4; k++) if (m->scl[L[k]] < sizeMin) sizeMin = m->scl[L[k]];
        if (stretch > size) return;
        for (int k = 0; k < 3; k++) {
            if (k == e) {
                n[k] = size / sizeMin;
                for (int i = 0; i < n[k]; i++) coord[k][i] = c->o[k] + i * sizeMin;
                elen[k] = sizeMin;
            } else {
                n[k] = 2;
                coord[k][0] = c->o[k] + size - stretch;
                coord[k][1] = c->o[k] + size;
                elen[k] = stretch;
            }
        }
    } else {
        stretch = m->scl[L[0]];
        for (int k = 1; k < 7; k++) if (stretch < m->scl[L[k]]) stretch = m->scl[L[k]];
        if (stretch > size) return;
        for (int k = 0; k < 3; k++) {
            n[k] = 2;
            coord[k][0] = c->o[k] + size - stretch;
            coord[k][1] = c->o[k] + size;
            elen[k] = stretch;
        }
        sizeMin = stretch;
    }
    (void)sizeMin;
    /* samples */
    sample_t S[BRICK];
    for (int l = 0; l < n[2]; l++)
        for (int j = 0; j < n[1]; j++)
            for (int i = 0; i < n[0]; i++) {
                int X[3] = {coord[0][i], coord[1][j], coord[2][l]};
                sample_t *s = &S[(l * BL + j) * BL + i];
                int g[3] = {i, j, l};
                if (type == 0) {
                    sample_interior(m, L[0], X, s);
                } else if (type <= 3) {
                    int a = type - 1;
                    int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
                    sample_wall(m, L[g[a]], a, u, v, X, s);
                } else if (type <= 6) {
                    int e = type - 4;
                    int u = e == 0 ? 1 : 0, v = e == 2 ? 1 : 2;
                    sample_edge(m, L[g[u] + 2 * g[v]], e, X, s);
                } else {
                    sample_point(m, L[i + 2 * j + 4 * l], X, s);
                }
            }
    const float off[3] = {m->off[0], m->off[1], m->off[2]};
    if (type == 7) {
        /* getColoredTrianglesForCubePrecomputed (mesh.cpp:1931-2008) on one cube */
        sample_t *cs[8];
        float P[8][3];
        int ci = 0, veto = 0;
        for (int k = 0; k < 8; k++) {
            const int *g = CORNER[k];
            cs[k] = &S[(g[2] * BL + g[1]) * BL + g[0]];
            if (cs[k]->w == 0.0f) veto = 1;
            if (cs[k]->d < 0.0f) ci |= 1 << k;
            for (int a = 0; a < 3; a++) {
                float base = off[a] + (float)coord[a][0] * m->scale;
                P[k][a] = g[a] ? base + 1.0f * (m->scale * (float)stretch) : base + 0.0f * (m->scale * (float)stretch);
            }
        }
        if (veto) ci = 0;
        int nt = tri_len(ci) / 3;
        for (int t = 0; t < nt; t++) {
            float V[3][3];
            uint8_t C[3][3];
            for (int q = 0; q < 3; q++) {
                int e = tri_edge(ci, 3 * t + q);
                int a = EDGE_CORNERS[e][0], b = EDGE_CORNERS[e][1];
                vinterp(P[a], P[b], cs[a]->d, cs[b]->d, V[q]);
                cinterp(cs[a]->c, cs[b]->c, cs[a]->d, cs[b]->d, C[q]);
            }
            int deg = 0;
            for (int q = 0; q < 3; q++) {
                const float *A = V[q], *B = V[(q + 1) % 3];
                if (A[0] == B[0] && A[1] == B[1] && A[2] == B[2]) deg = 1;
            }
            if (deg) continue;
            uint32_t base = (uint32_t)(mb->nv - cell_v0);
            for (int q = 0; q < 3; q++) mb_vertex(mb, V[q], C[q]);
            mb_face(mb, base, base + 1, base + 2);
        }
        return;
    }
    /* cubes: mark the edges their triangles use */
    static uint8_t used[BRICK * 3];
    static uint8_t cube[BRICK];
    memset(used, 0, sizeof used);
    for (int l = 0; l + 1 < n[2]; l++)
        for (int j = 0; j + 1 < n[1]; j++)
            for (int i = 0; i + 1 < n[0]; i++) {
                int ci = 0, veto = 0;
                for (int k = 0; k < 8; k++) {
                    const sample_t *s = &S[((l + CORNER[k][2]) * BL + j + CORNER[k][1]) * BL + i + CORNER[k][0]];
                    if (s->w == 0.0f) veto = 1;
                    if (s->d < 0.0f) ci |= 1 << k;
                }
                if (veto) ci = 0;
                int id = (l * BL + j) * BL + i;
                cube[id] = (uint8_t)ci;
                int len = tri_len(ci);
                for (int t = 0; t < len; t++) {
                    int e = tri_edge(ci, t);
                    int gi = i + EDGE[e][0], gj = j + EDGE[e][1], gl = l + EDGE[e][2];
                    used[((gl * BL + gj) * BL + gi) * 3 + EDGE[e][3]] = 1;
                }
            }
    /* vertices in (z, y, x) raster order, x-, y-, z-edge */
    static uint32_t vidx[BRICK * 3];
    uint32_t base = (uint32_t)mb->nv;
    for (int l = 0; l < n[2]; l++)
        for (int j = 0; j < n[1]; j++)
            for (int i = 0; i < n[0]; i++)
                for (int a = 0; a < 3; a++) {
                    int id = (l * BL + j) * BL + i;
                    if (!used[id * 3 + a]) continue;
                    int g2[3] = {i, j, l};
                    g2[a]++;
                    const sample_t *s1 = &S[id], *s2 = &S[(g2[2] * BL + g2[1]) * BL + g2[0]];
                    float p1[3], p2[3];
                    int gg[3] = {i, j, l};
                    for (int k = 0; k < 3; k++) {
                        p1[k] = off[k] + (float)coord[k][gg[k]] * m->scale;
                        p2[k] = p1[k];
                    }
                    p2[a] = off[a] + (float)(coord[a][gg[a]] + elen[a]) * m->scale;
                    float V[3];
                    uint8_t C[3];
                    vinterp(p1, p2, s1->d, s2->d, V);
                    cinterp(s1->c, s2->c, s1->d, s2->d, C);
                    vidx[id * 3 + a] = (uint32_t)mb->nv - base;
                    mb_vertex(mb, V, C);
                }
    for (int l = 0; l + 1 < n[2]; l++)
        for (int j = 0; j + 1 < n[1]; j++)
            for (int i = 0; i + 1 < n[0]; i++) {
                int ci = cube[(l * BL + j) * BL + i];
                int len = tri_len(ci);
                for (int t = 0; t < len; t += 3) {
                    uint32_t f[3];
                    for (int q = 0; q < 3; q++) {
                        int e = tri_edge(ci, t + q);
                        int gi = i + EDGE[e][0], gj = j + EDGE[e][1], gl = l + EDGE[e][2];
                        f[q] = vidx[((gl * BL + gj) * BL + gi) * 3 + EDGE[e][3]];
                    }
                    mb_face(mb, f[0], f[1], f[2]);
                }
            }
}

/* ---------------- public entry points ---------------- */

typedef struct {
    cellvec cells;
    uint32_t *nv, *nf;
    meshbuf mb;
} meshout;

static meshout g_out;

static void make_view(const ffo_ctx *c, mapview *m)
{
    m->tree = ffo_tree(c);
    m->pos = ffo_leaf_pos(c);
    m->scl = ffo_leaf_scale(c);
    m->par = ffo_leaf_parent(c);
    m->dist = ffo_distance(c);
    m->wgt = ffo_weights(c);
    m->col = ffo_color(c);
    m->nLeavesTotal = ffo_leaves_total(c);
    m->n = ffo_side(c);
    ffo_offset(c, m->off);
    m->scale = ffo_scale(c);
}

/* Meshes every meshable cell of the current map from scratch.  Returns the
 * number of cells (including cells whose mesh is empty). */
int ffo_mesh_all(ffo_ctx *c)
{
    free(g_out.cells.cells); free(g_out.nv); free(g_out.nf);
    free(g_out.mb.xyz); free(g_out.mb.rgb); free(g_out.mb.faces);
    memset(&g_out, 0, sizeof g_out);
    mapview m;
    make_view(c, &m);
    if (!m.n) return 0;
    cellvec all = {0};
    int o[3] = {0, 0, 0};
    walk(&m, 0, o, m.n >> 1, FFO_BRANCHINIT, &all);
    /* keep cells whose required lastLeaf slots exist (pushMeshCellQueue,
     * geometryfusion_mipmap_cpu.cpp:2653-2667) */
    size_t k = 0;
    for (size_t i = 0; i < all.n; i++)
        if (required_present(&all.cells[i])) all.cells[k++] = all.cells[i];
    all.n = k;
    qsort(all.cells, all.n, sizeof(cell_t), cmp_cells);
    g_out.cells = all;
    g_out.nv = (uint32_t *)calloc(all.n + 1, sizeof(uint32_t));
    g_out.nf = (uint32_t *)calloc(all.n + 1, sizeof(uint32_t));
    for (size_t i = 0; i < all.n; i++) {
        size_t v0 = g_out.mb.nv, f0 = g_out.mb.nf;
        mesh_cell(&m, &all.cells[i], &g_out.mb);
        g_out.nv[i] = (uint32_t)(g_out.mb.nv - v0);
        g_out.nf[i] = (uint32_t)(g_out.mb.nf - f0);
    }
    return (int)all.n;
}

void ffo_mesh_counts(const ffo_ctx *c, size_t out[3])
{
    (void)c;
    out[0] = g_out.cells.n;
    out[1] = g_out.mb.nv;
    out[2] = g_out.mb.nf;
}

/* desc: 13 int32 per cell (type, ox, oy, oz, size, lastLeaf[8]); faces are
 * local to their cell's vertex range. */
void ffo_mesh_copy(const ffo_ctx *c, int32_t *desc, uint32_t *nv, uint32_t *nf, float *xyz, uint8_t *rgb,
                   uint32_t *faces)
{
    (void)c;
    for (size_t i = 0; i < g_out.cells.n; i++) {
        const cell_t *cl = &g_out.cells.cells[i];
        int32_t *d = desc + 13 * i;
        d[0] = cl->type; d[1] = cl->o[0]; d[2] = cl->o[1]; d[3] = cl->o[2]; d[4] = cl->size;
        for (int k = 0; k < 8; k++) d[5 + k] = (int32_t)cl->ll[k];
        nv[i] = g_out.nv[i];
        nf[i] = g_out.nf[i];
    }
    memcpy(xyz, g_out.mb.xyz, g_out.mb.nv * 3 * sizeof(float));
    memcpy(rgb, g_out.mb.rgb, g_out.mb.nv * 3);
    memcpy(faces, g_out.mb.faces, g_out.mb.nf * 3 * sizeof(uint32_t));
}

/* ---------------- incremental update (CPU baseline) ----------------
 * updateMeshes semantics for timing: re-mesh the cells whose samples read a
 * brick (a lastLeaf, or a brick of its parent chain up to the level the
 * samples reached) updated since the previous call, or that are new or
 * changed their lastLeafs; keep the other cells' meshes; then
 * concatenate all cell meshes into one mesh like meshWrapperInterleaved
 * (geometryfusion_mipmap_cpu.cpp:2820-2880). */
typedef struct {
    uint64_t key;
    uint32_t ll[8];
    float *xyz;
    uint8_t *rgb;
    uint32_t *faces;
    uint32_t nv, nf;
    int used;
    int depth; /* the deepest parent-chain level its samples read (0: no samples) */
} centry;

typedef struct {
    centry *tab;
    size_t cap, n;
    uint32_t *last_queue_frame; /* per leaf: update counter when it was last queued */
    uint32_t nleaves_cap;
    uint32_t epoch;
    meshbuf all;
    cellvec upd; /* the cells the last update re-meshed (getMeshCellsUpdate) */
    const ffo_ctx *owner; /* the context this state belongs to */
    uint32_t last_epoch;  /* epoch of the last update */
} inc_state;

static inc_state g_inc;

/* the incremental state is kept for one context at a time: a call for
 * another context starts it afresh */
static void inc_reset(const ffo_ctx *c)
{
    for (size_t i = 0; i < g_inc.cap; i++) {
        free(g_inc.tab[i].xyz); free(g_inc.tab[i].rgb); free(g_inc.tab[i].faces);
    }
    free(g_inc.tab);
    free(g_inc.last_queue_frame);
    free(g_inc.all.xyz); free(g_inc.all.rgb); free(g_inc.all.faces);
    free(g_inc.upd.cells);
    memset(&g_inc, 0, sizeof g_inc);
    g_inc.owner = c;
}

static void inc_own(const ffo_ctx *c)
{
    if (g_inc.owner != c) inc_reset(c);
}

void ffo_mesh_forget(const ffo_ctx *c)
{
    if (g_inc.owner == c) inc_reset(NULL);
}

static size_t inc_find(uint64_t key)
{
    size_t h = (size_t)((key * 0x9e3779b97f4a7c15ull) >> 17) & (g_inc.cap - 1);
    while (g_inc.tab[h].used && g_inc.tab[h].key != key) h = (h + 1) & (g_inc.cap - 1);
    return h;
}

static void inc_rehash(size_t newcap)
{
    centry *old = g_inc.tab;
    size_t oc = g_inc.cap;
    g_inc.tab = (centry *)calloc(newcap, sizeof(centry));
    g_inc.cap = newcap;
    for (size_t i = 0; i < oc; i++)
        if (old[i].used) g_inc.tab[inc_find(old[i].key)] = old[i];
    free(old);
}

/* a brick among the first `depth` (at most 16) of the parent chain of leaf
 * was queued after `since` */
static int chain_changed(const mapview *m, uint32_t leaf, uint32_t since, int depth)
{
    for (int d = 0; d < depth && d < 16 && leaf < FFO_BRANCHINIT; d++) {
        if (g_inc.last_queue_frame[leaf] > since) return 1;
        leaf = m->par[leaf];
    }
    return 0;
}

/* Records the bricks queued by the last add_map (call after each frame). */
void ffo_mesh_note_frame(ffo_ctx *c)
{
    inc_own(c);
    uint32_t cap = ffo_leaves_total(c);
    if (g_inc.nleaves_cap != cap) {
        free(g_inc.last_queue_frame);
        g_inc.last_queue_frame = (uint32_t *)calloc(cap, sizeof(uint32_t));
        g_inc.nleaves_cap = cap;
    }
    g_inc.epoch++;
    const uint32_t *q = ffo_queue(c);
    for (uint32_t i = 0; i < ffo_n_queued(c); i++) g_inc.last_queue_frame[q[i]] = g_inc.epoch;
}

/* Incremental updateMeshes; returns the number of cells re-meshed. */
int ffo_mesh_update(ffo_ctx *c)
{
    inc_own(c);
    uint32_t last_epoch = g_inc.last_epoch;
    mapview m;
    make_view(c, &m);
    if (!m.n) return 0;
    if (!g_inc.tab) {
        g_inc.cap = 1 << 16;
        g_inc.tab = (centry *)calloc(g_inc.cap, sizeof(centry));
    }
    cellvec all = {0};
    int o[3] = {0, 0, 0};
    walk(&m, 0, o, m.n >> 1, FFO_BRANCHINIT, &all);
    size_t k = 0;
    for (size_t i = 0; i < all.n; i++)
        if (required_present(&all.cells[i])) all.cells[k++] = all.cells[i];
    all.n = k;
    while (2 * all.n > g_inc.cap) inc_rehash(2 * g_inc.cap);
    int remeshed = 0;
    meshbuf tmp = {0};
    g_inc.upd.n = 0;
    for (size_t i = 0; i < all.n; i++) {
        const cell_t *cl = &all.cells[i];
        uint64_t key = cell_key(cl);
        size_t h = inc_find(key);
        centry *e = &g_inc.tab[h];
        /* a kept cell re-meshes when a brick its samples read changed: one
         * within the chain depth they reached (the samples stop at the same
         * bricks while those are unchanged) */
        int dirty = !e->used || memcmp(e->ll, cl->ll, sizeof e->ll) != 0;
        for (int s = 0; s < 8 && !dirty; s++)
            if (cl->ll[s] < FFO_BRANCHINIT &&
                chain_changed(&m, cl->ll[s], last_epoch, e->depth < g_depth_cap ? e->depth : g_depth_cap))
                dirty = 1;
        if (!dirty) continue;
        tmp.nv = tmp.nf = 0;
        g_sample_depth = 0;
        mesh_cell(&m, cl, &tmp);
        e->depth = g_sample_depth;
        if (!e->used) { e->used = 1; e->key = key; g_inc.n++; }
        memcpy(e->ll, cl->ll, sizeof e->ll);
        e->xyz = (float *)realloc(e->xyz, (tmp.nv ? tmp.nv : 1) * 3 * sizeof(float));
        e->rgb = (uint8_t *)realloc(e->rgb, (tmp.nv ? tmp.nv : 1) * 3);
        e->faces = (uint32_t *)realloc(e->faces, (tmp.nf ? tmp.nf : 1) * 3 * sizeof(uint32_t));
        memcpy(e->xyz, tmp.xyz, tmp.nv * 3 * sizeof(float));
        memcpy(e->rgb, tmp.rgb, tmp.nv * 3);
        memcpy(e->faces, tmp.faces, tmp.nf * 3 * sizeof(uint32_t));
        e->nv = (uint32_t)tmp.nv;
        e->nf = (uint32_t)tmp.nf;
        remeshed++;
        if (g_inc.upd.n == g_inc.upd.cap) {
            g_inc.upd.cap = g_inc.upd.cap ? 2 * g_inc.upd.cap : 1024;
            g_inc.upd.cells = (cell_t *)realloc(g_inc.upd.cells, g_inc.upd.cap * sizeof(cell_t));
        }
        g_inc.upd.cells[g_inc.upd.n++] = *cl;
    }
    free(tmp.xyz); free(tmp.rgb); free(tmp.faces);
    /* concatenate every current cell's mesh (the reference's per-update copy) */
    g_inc.all.nv = g_inc.all.nf = 0;
    for (size_t i = 0; i < all.n; i++) {
        centry *e = &g_inc.tab[inc_find(cell_key(&all.cells[i]))];
        uint32_t base = (uint32_t)g_inc.all.nv;
        for (uint32_t v = 0; v < e->nv; v++) mb_vertex(&g_inc.all, e->xyz + 3 * v, e->rgb + 3 * v);
        for (uint32_t f = 0; f < e->nf; f++)
            mb_face(&g_inc.all, base + e->faces[3 * f], base + e->faces[3 * f + 1], base + e->faces[3 * f + 2]);
    }
    free(all.cells);
    g_inc.last_epoch = g_inc.epoch;
    return remeshed;
}

size_t ffo_mesh_update_faces(void) { return g_inc.all.nf; }

/* the concatenated mesh of the last ffo_mesh_update: out = {vertices, faces};
 * copied when the pointers are not null */
void ffo_mesh_update_all(size_t out[2], float *xyz, uint8_t *rgb, uint32_t *faces)
{
    out[0] = g_inc.all.nv;
    out[1] = g_inc.all.nf;
    if (xyz) memcpy(xyz, g_inc.all.xyz, g_inc.all.nv * 3 * sizeof(float));
    if (rgb) memcpy(rgb, g_inc.all.rgb, g_inc.all.nv * 3);
    if (faces) memcpy(faces, g_inc.all.faces, g_inc.all.nf * 3 * sizeof(uint32_t));
}

/* Test hook: caps the chain depth the incremental update checks (a planted
 * mistake for the tests that show the depth rule matters). */
void ffo_mesh_set_depth_cap(int cap) { g_depth_cap = cap; }

/* CellUpdate (geometryfusion_mipmap_cpu.cpp:2961-3047): the cells the last
 * ffo_mesh_update re-meshed, in key order, each with its own mesh (faces
 * local to the cell).  out = {cells, vertices, faces}. */
void ffo_mesh_update_sizes(size_t out[3])
{
    qsort(g_inc.upd.cells, g_inc.upd.n, sizeof(cell_t), cmp_cells);
    out[0] = g_inc.upd.n;
    out[1] = out[2] = 0;
    for (size_t i = 0; i < g_inc.upd.n; i++) {
        const centry *e = &g_inc.tab[inc_find(cell_key(&g_inc.upd.cells[i]))];
        out[1] += e->nv;
        out[2] += e->nf;
    }
}

void ffo_mesh_update_copy(int32_t *desc13, uint32_t *nv, uint32_t *nf, float *xyz, uint8_t *rgb, uint32_t *faces)
{
    size_t v = 0, f = 0;
    for (size_t i = 0; i < g_inc.upd.n; i++) {
        const cell_t *cl = &g_inc.upd.cells[i];
        const centry *e = &g_inc.tab[inc_find(cell_key(cl))];
        int32_t *d = desc13 + 13 * i;
        d[0] = cl->type; d[1] = cl->o[0]; d[2] = cl->o[1]; d[3] = cl->o[2]; d[4] = cl->size;
        for (int k = 0; k < 8; k++) d[5 + k] = (int32_t)cl->ll[k];
        nv[i] = e->nv;
        nf[i] = e->nf;
        memcpy(xyz + 3 * v, e->xyz, e->nv * 3 * sizeof(float));
        memcpy(rgb + 3 * v, e->rgb, e->nv * 3);
        memcpy(faces + 3 * f, e->faces, e->nf * 3 * sizeof(uint32_t));
        v += e->nv;
        f += e->nf;
    }
}
