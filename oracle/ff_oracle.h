/*
 * ff_oracle.h -- CPU restatement of fastfusion's per-frame integration path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle: the tests,
 * the smoke check and the CPU-baseline leg of the benchmark load it to check
 * or time the reference algorithm.  The product (fastfusion_amd) never links
 * or calls it.
 *
 * It restates, in plain scalar C, the reference's FusionMipMapCPU::addMap
 * u16 path (src/fusion/geometryfusion_mipmap_cpu.cpp:1603-2034), its brick
 * allocation (src/fusion/traversal_wrapper_functions.cpp:576-694,
 * src/fusion/traversal_functions.cpp:185-211, 308-395), the volume set-up and
 * growth (geometryfusion_mipmap_cpu.cpp:1186-1303, 3071-3319) and the AVX2
 * voxel update (src/fusion/update_functions.cpp:1187-1444, 1633-1732), which
 * runs in round-toward-zero with the RCPPS approximate reciprocal.
 */
#ifndef FF_ORACLE_H
#define FF_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FFO_BRANCHINIT 4294967294u
#define FFO_MAXLEAFNUMBER 10000000u

typedef struct ffo_ctx ffo_ctx;

ffo_ctx *ffo_create(float scale, float distance_threshold, int use_color,
                    uint32_t n_leaves_total, uint32_t n_branches_total);
void ffo_destroy(ffo_ctx *c);

/* One frame: depth u16 (row stride w), BGR8 (row stride 3w), pose camera->world
 * (R row-major, t), K = {fx, fy, cx, cy}.  Returns the number of queued bricks. */
int ffo_add_map(ffo_ctx *c, const uint16_t *depth, const uint8_t *bgr, int w, int h,
                const double R[9], const double t[3], const double K[4],
                float scaling, float maxcam);

/* Learned confidence (no counterpart in the reference): the weight of every
 * sample of the following frames is multiplied by conf[pixel] (W x H floats,
 * kept by the caller); NULL restores the plain method. */
void ffo_set_confidence(ffo_ctx *c, const float *conf);

/* state accessors */
uint32_t ffo_n_leaves(const ffo_ctx *c);
uint32_t ffo_n_branches(const ffo_ctx *c);
uint32_t ffo_side(const ffo_ctx *c);
void ffo_offset(const ffo_ctx *c, float out[3]);
uint32_t ffo_n_queued(const ffo_ctx *c);
const uint32_t *ffo_queue(const ffo_ctx *c);
const uint32_t *ffo_tree(const ffo_ctx *c);
const uint16_t *ffo_leaf_pos(const ffo_ctx *c);    /* 3 per leaf */
const uint16_t *ffo_leaf_scale(const ffo_ctx *c);
const uint32_t *ffo_leaf_parent(const ffo_ctx *c);
const float *ffo_distance(const ffo_ctx *c);       /* 512 per leaf */
const float *ffo_weights(const ffo_ctx *c);
const uint16_t *ffo_color(const ffo_ctx *c);       /* 3 per voxel: R, G, B (x, y, z) */
const uint8_t *ffo_processed(const ffo_ctx *c);    /* last frame's de-dup result, w*h */
int ffo_grow_count(const ffo_ctx *c);
uint32_t ffo_leaves_total(const ffo_ctx *c);
float ffo_scale(const ffo_ctx *c);

/* Meshing (ff_oracle_mesh.c): meshes every meshable cell of the current map
 * from scratch; cells are sorted by (ox, oy, oz, size, type). */
int ffo_mesh_all(ffo_ctx *c);
void ffo_mesh_counts(const ffo_ctx *c, size_t out[3]);
void ffo_mesh_copy(const ffo_ctx *c, int32_t *desc13, uint32_t *nv, uint32_t *nf, float *xyz, uint8_t *rgb,
                   uint32_t *faces);
/* Incremental meshing for the CPU baseline: ffo_mesh_note_frame after each
 * ffo_add_map records the queued bricks; ffo_mesh_update re-meshes the cells
 * they affect and concatenates all cell meshes (one context at a time). */
void ffo_mesh_note_frame(ffo_ctx *c);
int ffo_mesh_update(ffo_ctx *c);
size_t ffo_mesh_update_faces(void);
void ffo_mesh_update_all(size_t out[2], float *xyz, uint8_t *rgb, uint32_t *faces);
void ffo_mesh_set_depth_cap(int cap);
void ffo_mesh_update_sizes(size_t out[3]);
/* drops the incremental state if it belongs to c (c is being destroyed) */
void ffo_mesh_forget(const ffo_ctx *c);
void ffo_mesh_update_copy(int32_t *desc13, uint32_t *nv, uint32_t *nf, float *xyz, uint8_t *rgb, uint32_t *faces);

/* unit-level entry points used by the tests */
float ffo_rcp(float x);
uint16_t ffo_leaf_scale_func(float depth);
/* Processed-pixel mask of the de-dup recurrence for given per-pixel keys. */
void ffo_dedup(const int32_t *vx, const int32_t *vy, const int32_t *vz, const uint8_t *valid,
               int w, int h, uint8_t *processed);
/* Updates one brick (512 voxels) in place with the reference's voxel kernel. */
void ffo_update_brick(const uint16_t *depth, float scaling, float maxcam, const uint8_t *bgr,
                      int w, int h, const float m[12], const float K[4], float scale,
                      float distance_threshold, const uint16_t pos[3], uint16_t leaf_scale,
                      float *dist, float *weight, uint16_t *color);

float ffo_threshold(const ffo_ctx *c);
/* addMap's pInv (geometryfusion_mipmap_cpu.cpp:1771-1793) as 12 floats
 * (row r: R'_r0, R'_r1, R'_r2, t''_r) and the intrinsics fx, fy, cx, cy. */
void ffo_camera_inverse(const ffo_ctx *c, const double R[9], const double t[3], const double K[4],
                        float m[12], float Kf[4]);

/* Loop-closure de-integration (ff_oracle_loop.c, src/fusion/loopclosure.cpp):
 * a partner map of `total` slots for bricks numbered below map_n. */
typedef struct ffo_loop ffo_loop;
ffo_loop *ffo_loop_create(uint32_t map_n, uint32_t total, int use_color);
void ffo_loop_destroy(ffo_loop *L);
void ffo_loop_reset(ffo_loop *L);
/* Subtracts one frame (depth in metres, NaN = no measurement; color planes or
 * NULL) from the queued bricks of the main map c.  Returns 1, or 0 when the
 * slots ran out; *applied = queue entries updated. */
int ffo_loop_delete(ffo_loop *L, const ffo_ctx *c, const float *depth, const uint8_t *red,
                    const uint8_t *green, const uint8_t *blue, int w, int h, const double R[9],
                    const double t[3], const double K[4], const uint32_t *queue, uint32_t n,
                    uint32_t *applied);
/* One brick through the de-integration kernel, in place (color: R, G, B per voxel). */
void ffo_delete_brick(const float *depth, const uint8_t *red, const uint8_t *green, const uint8_t *blue,
                      int w, int h, const float m[12], const float K[4], float scale,
                      float distance_threshold, const uint16_t pos[3], uint16_t leaf_scale,
                      float *dist, float *weight, uint16_t *color);
uint32_t ffo_loop_mapped(const ffo_loop *L);
const uint32_t *ffo_loop_map(const ffo_loop *L);
const float *ffo_loop_distance(const ffo_loop *L);
const float *ffo_loop_weights(const ffo_loop *L);
const uint16_t *ffo_loop_color(const ffo_loop *L);

#ifdef __cplusplus
}
#endif
#endif
