/*
 * synth.c -- seeded synthetic RGB-D frames of analytic indoor scenes.
 *
 * A scene is an axis-aligned room (the camera sits inside it), optional
 * interior walls, axis-aligned boxes and spheres.  Frames are ray-cast
 * exactly (closed-form ray/box and ray/sphere hits), quantised like a
 * Kinect-style sensor (u16 depth, 1/5000 m per unit, 0 = missing), with
 * Gaussian depth noise sigma(z) = a + b (z - 0.4)^2, dropout and optional
 * "flying pixels" at depth edges.  Color is BGR8 (byte 0 = blue), the
 * channel order the reference reads (src/fusion/update_functions.cpp:1274-1299).
 * The scene's signed distance function is exact for evaluation.
 *
 * Camera convention (the reference's CameraInfo, camerautils.cpp:50-56):
 * R, t map camera coordinates to world coordinates; the camera looks along
 * +z with x to the right and y down.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define FFS_MAX_PRIM 512

typedef struct { float lo[3], hi[3]; float col[3]; } ffs_box;
typedef struct { float c[3], r; float col[3]; } ffs_sphere;

typedef struct ffs_scene {
    float room[3];            /* room spans [0, room] on each axis */
    int nbox, nsph;
    ffs_box box[FFS_MAX_PRIM];
    ffs_sphere sph[FFS_MAX_PRIM];
} ffs_scene;

/* ---------------- random numbers (splitmix64) ---------------- */
static uint64_t sm64(uint64_t *s)
{
    uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static double uni(uint64_t *s) { return (double)(sm64(s) >> 11) * (1.0 / 9007199254740992.0); }
static double urange(uint64_t *s, double a, double b) { return a + (b - a) * uni(s); }
static double gauss(uint64_t *s)
{
    double u1 = uni(s), u2 = uni(s);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static void rand_col(uint64_t *s, float c[3])
{
    for (int k = 0; k < 3; k++) c[k] = (float)urange(s, 0.15, 0.95);
}

static void add_box(ffs_scene *sc, const float lo[3], const float hi[3], const float col[3])
{
    if (sc->nbox >= FFS_MAX_PRIM) return;
    ffs_box *b = &sc->box[sc->nbox++];
    memcpy(b->lo, lo, sizeof b->lo);
    memcpy(b->hi, hi, sizeof b->hi);
    memcpy(b->col, col, sizeof b->col);
}

static void add_sphere(ffs_scene *sc, const float c[3], float r, const float col[3])
{
    if (sc->nsph >= FFS_MAX_PRIM) return;
    ffs_sphere *p = &sc->sph[sc->nsph++];
    memcpy(p->c, c, sizeof p->c);
    p->r = r;
    memcpy(p->col, col, sizeof p->col);
}

/* Places n_boxes boxes standing on the floor or against walls, and n_spheres
 * spheres with radii in [rmin, rmax], inside the room with a margin. */
static void scatter(ffs_scene *sc, uint64_t *s, int n_boxes, int n_spheres, float rmin, float rmax)
{
    const float *L = sc->room;
    for (int i = 0; i < n_boxes; i++) {
        float lo[3], hi[3], col[3];
        for (int k = 0; k < 2; k++) {
            float ext = (float)urange(s, 0.2, fminf(1.0f, 0.25f * L[k]));
            float c = (float)urange(s, 0.1 + ext / 2, L[k] - 0.1 - ext / 2);
            lo[k] = c - ext / 2;
            hi[k] = c + ext / 2;
        }
        lo[2] = 0.0f;
        hi[2] = (float)urange(s, 0.3, fminf(1.6f, 0.6f * L[2]));
        rand_col(s, col);
        add_box(sc, lo, hi, col);
    }
    for (int i = 0; i < n_spheres; i++) {
        float c[3], col[3], r = (float)urange(s, rmin, rmax);
        for (int k = 0; k < 3; k++) c[k] = (float)urange(s, 0.1 + r, L[k] - 0.1 - r);
        rand_col(s, col);
        add_sphere(sc, c, r, col);
    }
}

ffs_scene *ffs_room_create(uint64_t seed, float lx, float ly, float lz, int n_boxes, int n_spheres,
                           float rmin, float rmax)
{
    ffs_scene *sc = (ffs_scene *)calloc(1, sizeof(ffs_scene));
    uint64_t s = seed * 0x2545f4914f6cdd1dull + 17;
    sc->room[0] = lx; sc->room[1] = ly; sc->room[2] = lz;
    scatter(sc, &s, n_boxes, n_spheres, rmin, rmax);
    return sc;
}

/* A floor of rooms: interior walls 0.1 m thick on a grid of cells_x x cells_y
 * rooms, each wall with a 1 m wide, 2.1 m high door gap, plus n_prims
 * primitives (boxes and spheres alternately). */
ffs_scene *ffs_floor_create(uint64_t seed, float lx, float ly, float lz, int cells_x, int cells_y,
                            int n_prims)
{
    ffs_scene *sc = (ffs_scene *)calloc(1, sizeof(ffs_scene));
    uint64_t s = seed * 0x2545f4914f6cdd1dull + 29;
    sc->room[0] = lx; sc->room[1] = ly; sc->room[2] = lz;
    float wc[3] = {0.85f, 0.85f, 0.8f};
    for (int i = 1; i < cells_x; i++) {
        float x = lx * i / cells_x;
        float door = (float)urange(&s, 0.3, ly / cells_y - 1.3);
        for (int j = 0; j < cells_y; j++) {
            float y0 = ly * j / cells_y, y1 = ly * (j + 1) / cells_y;
            float d0 = y0 + door, d1 = d0 + 1.0f;
            float lo1[3] = {x - 0.05f, y0, 0}, hi1[3] = {x + 0.05f, d0, lz};
            float lo2[3] = {x - 0.05f, d1, 0}, hi2[3] = {x + 0.05f, y1, lz};
            float lo3[3] = {x - 0.05f, d0, 2.1f}, hi3[3] = {x + 0.05f, d1, lz};
            add_box(sc, lo1, hi1, wc); add_box(sc, lo2, hi2, wc); add_box(sc, lo3, hi3, wc);
        }
    }
    for (int j = 1; j < cells_y; j++) {
        float y = ly * j / cells_y;
        float door = (float)urange(&s, 0.3, lx / cells_x - 1.3);
        for (int i = 0; i < cells_x; i++) {
            float x0 = lx * i / cells_x, x1 = lx * (i + 1) / cells_x;
            float d0 = x0 + door, d1 = d0 + 1.0f;
            float lo1[3] = {x0, y - 0.05f, 0}, hi1[3] = {d0, y + 0.05f, lz};
            float lo2[3] = {d1, y - 0.05f, 0}, hi2[3] = {x1, y + 0.05f, lz};
            float lo3[3] = {d0, y - 0.05f, 2.1f}, hi3[3] = {d1, y + 0.05f, lz};
            add_box(sc, lo1, hi1, wc); add_box(sc, lo2, hi2, wc); add_box(sc, lo3, hi3, wc);
        }
    }
    scatter(sc, &s, n_prims / 2, n_prims - n_prims / 2, 0.1f, 0.5f);
    return sc;
}

void ffs_destroy(ffs_scene *sc) { free(sc); }

int ffs_counts(const ffs_scene *sc, int out[2])
{
    out[0] = sc->nbox;
    out[1] = sc->nsph;
    return 0;
}

/* ---------------- signed distance (positive in free space) ---------------- */
static float sd_box(const ffs_box *b, const float p[3])
{
    float q[3], outside = 0.0f, inside = -1e30f;
    for (int k = 0; k < 3; k++) {
        float c = 0.5f * (b->lo[k] + b->hi[k]), e = 0.5f * (b->hi[k] - b->lo[k]);
        q[k] = fabsf(p[k] - c) - e;
        float m = fmaxf(q[k], 0.0f);
        outside += m * m;
        if (q[k] > inside) inside = q[k];
    }
    return sqrtf(outside) + fminf(inside, 0.0f);
}

float ffs_sdf(const ffs_scene *sc, float x, float y, float z)
{
    float p[3] = {x, y, z};
    float d = 1e30f;
    for (int k = 0; k < 3; k++) {
        d = fminf(d, p[k]);
        d = fminf(d, sc->room[k] - p[k]);
    }
    for (int i = 0; i < sc->nbox; i++) d = fminf(d, sd_box(&sc->box[i], p));
    for (int i = 0; i < sc->nsph; i++) {
        const ffs_sphere *s = &sc->sph[i];
        float dx = p[0] - s->c[0], dy = p[1] - s->c[1], dz = p[2] - s->c[2];
        d = fminf(d, sqrtf(dx * dx + dy * dy + dz * dz) - s->r);
    }
    return d;
}

void ffs_sdf_batch(const ffs_scene *sc, const float *xyz, size_t n, float *out)
{
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++) out[i] = ffs_sdf(sc, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

/* ---------------- ray casting ---------------- */
/* Returns the ray parameter of the first hit (t > 0) and the hit color. */
static double cast(const ffs_scene *sc, const double o[3], const double d[3], float col[3])
{
    double best = 1e30;
    int kind = -1, idx = -1;
    /* room interior: exit distance */
    for (int k = 0; k < 3; k++) {
        if (d[k] > 1e-12) {
            double t = (sc->room[k] - o[k]) / d[k];
            if (t > 0 && t < best) { best = t; kind = 0; idx = 2 * k + 1; }
        } else if (d[k] < -1e-12) {
            double t = (0.0 - o[k]) / d[k];
            if (t > 0 && t < best) { best = t; kind = 0; idx = 2 * k; }
        }
    }
    for (int i = 0; i < sc->nbox; i++) {
        const ffs_box *b = &sc->box[i];
        double tn = -1e30, tf = 1e30;
        int ok = 1;
        for (int k = 0; k < 3 && ok; k++) {
            if (fabs(d[k]) < 1e-12) {
                if (o[k] < b->lo[k] || o[k] > b->hi[k]) ok = 0;
            } else {
                double t1 = (b->lo[k] - o[k]) / d[k], t2 = (b->hi[k] - o[k]) / d[k];
                if (t1 > t2) { double tmp = t1; t1 = t2; t2 = tmp; }
                if (t1 > tn) tn = t1;
                if (t2 < tf) tf = t2;
                if (tn > tf) ok = 0;
            }
        }
        if (ok && tn > 0 && tn < best) { best = tn; kind = 1; idx = i; }
    }
    for (int i = 0; i < sc->nsph; i++) {
        const ffs_sphere *s = &sc->sph[i];
        double oc[3] = {o[0] - s->c[0], o[1] - s->c[1], o[2] - s->c[2]};
        double a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        double bb = oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2];
        double cc = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - (double)s->r * s->r;
        double disc = bb * bb - a * cc;
        if (disc < 0) continue;
        double t = (-bb - sqrt(disc)) / a;
        if (t > 0 && t < best) { best = t; kind = 2; idx = i; }
    }
    if (kind < 0) return -1.0;
    double h[3] = {o[0] + best * d[0], o[1] + best * d[1], o[2] + best * d[2]};
    const float *base;
    static const float wallcol[6][3] = {{0.8f, 0.7f, 0.6f}, {0.6f, 0.7f, 0.8f}, {0.7f, 0.8f, 0.6f},
                                        {0.75f, 0.6f, 0.7f}, {0.55f, 0.5f, 0.45f}, {0.9f, 0.9f, 0.9f}};
    if (kind == 0) base = wallcol[idx];
    else if (kind == 1) base = sc->box[idx].col;
    else base = sc->sph[idx].col;
    /* a 10 cm checker texture so color varies across every surface */
    int chk = ((int)floor(h[0] * 10.0) + (int)floor(h[1] * 10.0) + (int)floor(h[2] * 10.0)) & 1;
    float f = chk ? 1.0f : 0.7f;
    for (int k = 0; k < 3; k++) col[k] = base[k] * f;
    return best;
}

/* Renders one frame.  noise_mult scales the noise sigma, dropout is the
 * fraction of missing pixels, flying the fraction of depth-edge pixels
 * replaced by a blend of the two sides.  depth_gt (optional, may be NULL)
 * receives the noiseless metric depth. */
void ffs_render(const ffs_scene *sc, const double R[9], const double t[3], const double K[4],
                int w, int h, float depth_scale, float noise_mult, float dropout, float flying,
                uint64_t noise_seed, uint16_t *depth, uint8_t *bgr, float *depth_gt)
{
    float *z = (float *)malloc(sizeof(float) * (size_t)w * h);
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            double dc[3] = {(x - K[2]) / K[0], (y - K[3]) / K[1], 1.0};
            double dw[3];
            for (int k = 0; k < 3; k++) dw[k] = R[k * 3] * dc[0] + R[k * 3 + 1] * dc[1] + R[k * 3 + 2] * dc[2];
            float col[3] = {0, 0, 0};
            double tt = cast(sc, t, dw, col);
            size_t i = (size_t)y * w + x;
            z[i] = tt > 0 ? (float)tt : 0.0f; /* dc has unit z, so t is the camera-frame depth */
            bgr[3 * i + 0] = (uint8_t)lrintf(255.0f * col[2]);
            bgr[3 * i + 1] = (uint8_t)lrintf(255.0f * col[1]);
            bgr[3 * i + 2] = (uint8_t)lrintf(255.0f * col[0]);
        }
    }
    if (depth_gt) memcpy(depth_gt, z, sizeof(float) * (size_t)w * h);
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++) {
        uint64_t s = noise_seed * 0x9e3779b97f4a7c15ull + (uint64_t)y * 0x632be59bd9b4e019ull + 1;
        for (int x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            float zz = z[i];
            double g = gauss(&s), u = uni(&s), uf = uni(&s);
            if (zz <= 0.0f) { depth[i] = 0; continue; }
            if (flying > 0.0f && uf < flying * 10.0) {
                /* an edge pixel: a neighbour differs by > 10 cm */
                float zn = zz;
                if (x + 1 < w && fabsf(z[i + 1] - zz) > 0.1f) zn = z[i + 1];
                else if (y + 1 < h && fabsf(z[i + w] - zz) > 0.1f) zn = z[i + w];
                if (zn != zz && uf < flying) {
                    double a = uni(&s);
                    zz = (float)(zz * a + zn * (1.0 - a));
                }
            }
            double sigma = noise_mult * (0.0012 + 0.0019 * (zz - 0.4) * (zz - 0.4));
            double zn = zz + sigma * g;
            if (u < dropout || zn <= 0.0) { depth[i] = 0; continue; }
            double q = zn * depth_scale + 0.5;
            depth[i] = q >= 65535.0 ? 65535 : (uint16_t)q;
        }
    }
    free(z);
}

/* ---------------- trajectories ---------------- */
static void look_rot(const double f_in[3], double R[9])
{
    double f[3], n = sqrt(f_in[0] * f_in[0] + f_in[1] * f_in[1] + f_in[2] * f_in[2]);
    for (int k = 0; k < 3; k++) f[k] = f_in[k] / n;
    double up[3] = {0, 0, 1};
    double x[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    n = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    if (n < 1e-9) { x[0] = 1; x[1] = 0; x[2] = 0; n = 1; }
    for (int k = 0; k < 3; k++) x[k] /= n;
    double y[3] = {f[1] * x[2] - f[2] * x[1], f[2] * x[0] - f[0] * x[2], f[0] * x[1] - f[1] * x[0]};
    for (int k = 0; k < 3; k++) {
        R[k * 3 + 0] = x[k];
        R[k * 3 + 1] = y[k];
        R[k * 3 + 2] = f[k];
    }
}

/* Circle of the given radius around the room centre at 1.5 m height (or mid
 * height for low rooms), looking at the centre with a slight downward tilt. */
int ffs_trajectory_circle(const ffs_scene *sc, int n, float radius, double *R, double *t)
{
    double cx = 0.5 * sc->room[0], cy = 0.5 * sc->room[1];
    double hz = sc->room[2] > 2.0f ? 1.5 : 0.5 * sc->room[2];
    for (int i = 0; i < n; i++) {
        double a = 6.283185307179586 * i / (n > 0 ? n : 1);
        double p[3] = {cx + radius * cos(a), cy + radius * sin(a), hz};
        double f[3] = {cx - p[0], cy - p[1], -0.35 * radius};
        look_rot(f, R + 9 * i);
        memcpy(t + 3 * i, p, sizeof p);
    }
    return n;
}

/* Seeded random walk: each step moves step_m metres in a random direction
 * and turns yaw/pitch by up to step_deg, staying dmin..dmax metres from the
 * nearest surface (the scene SDF) and inside the room. */
int ffs_trajectory_walk(const ffs_scene *sc, uint64_t seed, int n, float dmin, float dmax,
                        float step_m, float step_deg, double *R, double *t)
{
    uint64_t s = seed * 0xd1342543de82ef95ull + 5;
    double p[3] = {0.5 * sc->room[0], 0.5 * sc->room[1], sc->room[2] > 2.0f ? 1.4 : 0.5 * sc->room[2]};
    /* find a start position that satisfies the distance bounds */
    for (int tries = 0; tries < 10000; tries++) {
        float d = ffs_sdf(sc, (float)p[0], (float)p[1], (float)p[2]);
        if (d >= dmin && d <= dmax) break;
        for (int k = 0; k < 3; k++) p[k] = urange(&s, 0.3, sc->room[k] - 0.3);
    }
    double yaw = urange(&s, 0, 6.283185307179586), pitch = -0.2;
    double srad = step_deg * 3.141592653589793 / 180.0;
    for (int i = 0; i < n; i++) {
        double f[3] = {cos(pitch) * cos(yaw), cos(pitch) * sin(yaw), sin(pitch)};
        look_rot(f, R + 9 * i);
        memcpy(t + 3 * i, p, sizeof p);
        for (int tries = 0; tries < 64; tries++) {
            double dir[3] = {gauss(&s), gauss(&s), 0.3 * gauss(&s)};
            double nn = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]) + 1e-12;
            double q[3];
            for (int k = 0; k < 3; k++) q[k] = p[k] + step_m * dir[k] / nn;
            float d = ffs_sdf(sc, (float)q[0], (float)q[1], (float)q[2]);
            if (d >= dmin && d <= dmax) { memcpy(p, q, sizeof q); break; }
        }
        yaw += urange(&s, -srad, srad);
        pitch += urange(&s, -srad, srad);
        if (pitch > 0.5) pitch = 0.5;
        if (pitch < -0.6) pitch = -0.6;
    }
    return n;
}
