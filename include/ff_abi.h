/*
 * ff_abi.h -- C ABI of the MI355X fusion library (libffgpu.so).
 *
 * The reference has no FFI layer: its integration path sits behind the C++
 * class FusionMipMapCPU (src/fusion/geometryfusion_mipmap_cpu.hpp:111-163),
 * a subclass of Fusion_AoS (src/fusion/geometryfusion_aos.hpp:20-48).  Each
 * entry point below replaces one member of that class; the comment names it.
 * All arguments are plain pointers and sizes; no exception crosses the ABI.
 * One context is used by one thread at a time, like the reference's fusion
 * thread.  Negative return values are FF_E* error codes; ff_last_error()
 * returns the message of the last failure on that context.
 */
#ifndef FF_ABI_H
#define FF_ABI_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FF_OK 0
#define FF_E_ARG -1        /* invalid argument */
#define FF_E_DEVICE -2     /* HIP runtime failure */
#define FF_E_POOL_FULL -3  /* brick or branch pool exhausted */
#define FF_E_STATE -4      /* call not valid in the current state */

typedef struct ff_ctx ff_ctx;

/* Constructor arguments of FusionMipMapCPU (geometryfusion_mipmap_cpu.hpp:114-116:
 * scale, distanceThreshold, color) plus the brick-pool capacity that the
 * reference derives from host memory (geometryfusion_mipmap_cpu.cpp:497-509)
 * and the HIP device ordinal. */
typedef struct ff_params {
    float scale;              /* voxel edge in metres */
    float distance_threshold; /* truncation in metres */
    int use_color;            /* fuse BGR color (the reference's `color`) */
    int incremental_meshing;  /* setIncrementalMeshing */
    uint32_t max_leaves;      /* brick capacity (_nLeavesTotal); 0 = default */
    uint32_t max_branches;    /* tree capacity in 32-bit slots (_nBranchesTotal); 0 = default */
    int device;               /* HIP device ordinal */
} ff_params;

/* State scalars after a frame, for inspection and parity checks. */
typedef struct ff_counts {
    uint32_t n_leaves_used;   /* _nLeavesUsed */
    uint32_t n_branches_used; /* _nBranchesUsed */
    uint32_t n_queued;        /* _nLeavesQueuedSurface: bricks updated by the last frame */
    uint32_t side;            /* _n: tree side length in voxels */
    float offset[3];          /* _offset: world position of voxel (0,0,0) */
    uint32_t frames;          /* _framesAdded */
    uint32_t grows;           /* number of grow() calls so far */
} ff_counts;

/* FusionMipMapCPU::FusionMipMapCPU (geometryfusion_mipmap_cpu.cpp:420-640). */
ff_ctx *ff_create(const ff_params *params);
/* FusionMipMapCPU::~FusionMipMapCPU */
void ff_destroy(ff_ctx *ctx);
const char *ff_last_error(const ff_ctx *ctx);

/* FusionMipMapCPU::addMap(const cv::Mat &depth, CameraInfo, const cv::Mat &rgb,
 * float scaling, float maxcamdistance), u16 overload
 * (geometryfusion_mipmap_cpu.cpp:1603-2034).  depth: CV_16UC1 host image with
 * row stride depth_stride (bytes); bgr: CV_8UC3 host image, row stride
 * bgr_stride (bytes), may be NULL when color is off; R, t: CameraInfo
 * rotation (row-major) and translation, camera->world; K = {fx, fy, cx, cy}.
 * Returns the number of bricks updated (the reference's return value) or FF_E*. */
int ff_add_map(ff_ctx *ctx, const uint16_t *depth, size_t depth_stride, const uint8_t *bgr,
               size_t bgr_stride, int width, int height, const double R[9], const double t[3],
               const double K[4], float scaling, float maxcam);

/* Same, with depth and bgr already resident in device memory (dense rows).
 * The call returns once the frame's counts are known, possibly while its
 * voxel update still reads the images: they must stay unchanged until the
 * next ff_add_map* call on the context returns (or any call that reads the
 * map, which waits for the frame).  Every later call sees the whole frame,
 * as after ff_add_map, which keeps no reference to the caller's host images
 * once it returns. */
int ff_add_map_device(ff_ctx *ctx, const uint16_t *depth_dev, const uint8_t *bgr_dev, int width,
                      int height, const double R[9], const double t[3], const double K[4],
                      float scaling, float maxcam);

/* No counterpart in the reference: ff_add_map_device for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The frame's image-sized buffers (bit
 * planes, processed bitmap, count-scan arrays, per-pixel parent records,
 * confidences) come from one per-device array of 40 MiB that the library
 * holds for the life of the process, instead of the context's own buffers;
 * the call allocates nothing (a mesh pool compaction or buffer growth due
 * from a previous updateMeshes waits for the next ff_update_meshes).  Images
 * whose buffers exceed the array (above 1280 x 960 pixels) give FF_E_ARG.
 * Results equal ff_add_map_device's.  Every call on a device shares that
 * device's array: calls that overlap in time on one device (from several
 * threads, on any contexts) race.  ff_copy_processed and ff_copy_confidence
 * read the array after such a call, until the next call on the device. */
int ff_add_map_device_noalloc(ff_ctx *ctx, const uint16_t *depth_dev, const uint8_t *bgr_dev, int width,
                              int height, const double R[9], const double t[3], const double K[4],
                              float scaling, float maxcam);

/* FusionMipMapCPU::updateMeshes (geometryfusion_mipmap_cpu.cpp:2884-2954):
 * re-meshes the cells touched by the frames since the last call, as the map
 * stands at the call.  The call records the request and returns (the
 * reference's updateMeshes returns true; its meshing thread also runs behind
 * the caller): the device work is enqueued by the next call on the context
 * -- the next ff_add_map*, right after that frame's first kernels (which
 * read only the new images), or any other call first -- so that it never
 * sits between a frame and the next frame's start.  Every mesh query below
 * sees its result.  Its host bookkeeping runs at the next mesh call or
 * query, which also returns its errors (FF_E_POOL_FULL: cell table
 * overflow).  Returns FF_OK or FF_E*. */
int ff_update_meshes(ff_ctx *ctx);

/* No counterpart in the reference: enqueues a requested mesh update and
 * waits until all of the context's device work has completed (after it the
 * device images of the last frame are no longer read). */
int ff_sync(ff_ctx *ctx);

/* The whole map meshed from scratch, the counterpart of the reference's
 * non-incremental getMeshRecursive (geometryfusion_mipmap_cpu.cpp:3658-3674):
 * every cell is re-meshed with the same per-cell marching cubes, whatever
 * changed, and the call returns when the meshes are complete; ff_mesh_sizes
 * and ff_copy_mesh then return the full mesh, which equals the union of the
 * incremental cell meshes of an up-to-date map. */
int ff_mesh_full(ff_ctx *ctx);

/* getMeshInterleavedMarchingCubes (geometryfusion_mipmap_cpu.cpp:3359-3374):
 * the MeshInterleaved layout -- vertices 3 x f32, colors 3 x u8 (R, G, B),
 * faces 3 x u32 per triangle.  ff_mesh_sizes returns the counts; ff_copy_mesh
 * fills caller buffers of those sizes (any pointer may be NULL to skip it). */
int ff_mesh_sizes(ff_ctx *ctx, size_t *n_vertices, size_t *n_faces);
int ff_copy_mesh(ff_ctx *ctx, float *xyz, uint8_t *rgb, uint32_t *faces);

/* Per-cell output (getMeshCellsUpdate, geometryfusion_mipmap_cpu.cpp:3006-3047):
 * the number of cells with a mesh, and for each cell its descriptor
 * {type, ox, oy, oz, size} and its vertex and face counts.  Cells are listed
 * in the order their geometry appears in ff_copy_mesh. */
int ff_cell_count(ff_ctx *ctx, size_t *n_cells);
int ff_copy_cells(ff_ctx *ctx, int32_t *desc5, uint32_t *n_vertices, uint32_t *n_faces);

/* getMeshCellsUpdate (geometryfusion_mipmap_cpu.cpp:2989-3047; CellUpdate /
 * InterleavedMeshCell, geometryfusion_mipmap_cpu.hpp:40-59): the cells the
 * last completed ff_update_meshes re-meshed, each with its own mesh
 * (vertices, colors and triangle indices local to the cell), empty meshes
 * included.  The reference names a cell by its index in its cell array;
 * here a cell is named by its descriptor {type, ox, oy, oz, size}, which
 * stays the same across updates.  ff_cell_update_sizes gives the number of
 * cells and their summed vertex and face counts; ff_copy_cell_update fills
 * per cell the descriptor and counts, and the meshes concatenated in that
 * order (any pointer may be NULL).  Cells are in key order. */
int ff_cell_update_sizes(ff_ctx *ctx, size_t *n_cells, size_t *n_vertices, size_t *n_faces);
int ff_copy_cell_update(ff_ctx *ctx, int32_t *desc5, uint32_t *n_vertices, uint32_t *n_faces, float *xyz,
                        uint8_t *rgb, uint32_t *faces);

/* setIncrementalMeshing (geometryfusion_mipmap_cpu.hpp:131). */
int ff_set_incremental_meshing(ff_ctx *ctx, int on);

/* Device memory knob with no counterpart in the reference (whose meshing
 * samples cells on the host): the capacity, in tiles of 64 samples (1 KiB),
 * of the buffer that holds the samples of the cells re-meshed by one
 * updateMeshes.  Cells beyond it are sampled in place by the marching cubes
 * pass itself and the buffer then grows; the default is sized from
 * max_leaves. */
int ff_set_mesh_sample_tiles(ff_ctx *ctx, uint32_t tiles);

/* Device memory knob with no counterpart in the reference (whose cells own
 * their meshes on the heap): the capacity, in vertices and in triangles, of
 * the pools the cells' meshes live in (at least 4096; rounded down to a
 * multiple of 64; split into 64 arenas), and compact_percent (1-100,
 * default 50): when an arena's space allocated since the last compaction
 * passes that share of it, the update compacts the live geometry into fresh
 * pools.  Setting it drops the current geometry: the next updateMeshes
 * meshes every cell.  The default capacity is sized from max_leaves. */
int ff_set_mesh_pool(ff_ctx *ctx, uint32_t vertices, uint32_t compact_percent);
/* The pool capacity and the number of compactions so far. */
int ff_mesh_pool_info(ff_ctx *ctx, uint32_t *vertices, uint32_t *compactions);

/* Map state, for inspection and parity (the reference's protected members
 * _tree, _leafPos, _leafScale, _leafParent, _distance, _weights, _color). */
int ff_get_counts(ff_ctx *ctx, ff_counts *out);
int ff_copy_tree(ff_ctx *ctx, uint32_t *out, size_t n_slots);
int ff_copy_leaves(ff_ctx *ctx, uint16_t *pos3, uint16_t *scale, uint32_t *parent, size_t n);
/* Copies TSDF data of the given leaves: 512 distances, 512 weights and
 * 3 x 512 color channels (R, G, B planes) per leaf; any output may be NULL. */
int ff_copy_bricks(ff_ctx *ctx, const uint32_t *leaves, size_t n, float *dist, float *weight,
                   uint16_t *color);
/* The queue of bricks updated by the last frame (_leafNumberSurface), in
 * ascending leaf order. */
int ff_copy_queue(ff_ctx *ctx, uint32_t *out, size_t cap);
/* De-dup result of the last frame: one byte per pixel, 1 = processed. */
int ff_copy_processed(ff_ctx *ctx, uint8_t *out, size_t n);

/* Per-phase GPU timing (no counterpart in the reference, which prints
 * cv::getTickCount averages, geometryfusion_mipmap_cpu.cpp:979-991).
 * ff_set_timing(ctx, 1) records HIP events around each frame's phases on
 * the context stream: 0 = allocation kernels, 1 = SDF update kernel,
 * 2 = meshing; ff_set_timing(ctx, 2) only around the SDF update (the least
 * event traffic: each event record costs launch time); 0 turns it off.
 * ff_get_timing returns per phase the summed milliseconds, the number of
 * timed launches, the summed algorithmic bytes and work units (processed
 * pixels, queued bricks, re-meshed cells); arrays have 3 entries. */
/* Launch mode, no counterpart in the reference: with capture on, a frame's
 * device work is captured once into HIP graphs and replayed; off (the
 * default, measured faster for one stream), every kernel is launched
 * directly. */
int ff_set_graph_capture(ff_ctx *ctx, int on);
int ff_set_timing(ff_ctx *ctx, int on);
/* Times the SDF update on one frame in n (default 1): its events are
 * recorded by the kernel's own dispatch, which costs launch time. */
int ff_set_timing_every(ff_ctx *ctx, uint32_t n);
int ff_get_timing(ff_ctx *ctx, double *ms, uint64_t *calls, double *bytes, uint64_t *units);
/* The marching-cubes kernel alone (no counterpart in the reference): with
 * timing on, its dispatch records events on one mesh update in n (the
 * ff_set_timing_every count): summed milliseconds and timed launches; with
 * all phases timed (ff_set_timing(ctx, 1)) also, on every update, its
 * algorithmic bytes (14 B per consumed cell sample, 15 B per vertex and 12 B
 * per triangle written), the number of updates counted and of re-meshed
 * cells. */
int ff_get_mc_timing(ff_ctx *ctx, double *ms, uint64_t *calls, double *bytes, uint64_t *byte_calls,
                     uint64_t *cells);
int ff_reset_timing(ff_ctx *ctx);

/* Device stream the context launches on (a hipStream_t), for callers that
 * time or order work around it. */
void *ff_stream(ff_ctx *ctx);

/* ---- Map dump and restore ----------------------------------------------
 * FusionMipMapCPU::getDescription / copyCPU2HDD / copyHDD2CPU
 * (geometryfusion_mipmap_cpu.cpp:1128-1184): the reference's raw format
 * (counts, tree, brick positions and scales, distances, weights, colors as
 * R, G, B per voxel), followed by a block the reference does not write
 * ("FFPARENT", count, the bricks' parents).  The reference names the file
 * resource/<description>.dat and restores into an object whose volume
 * already matches; ff_load_map takes the volume (offset, side n) instead
 * and needs a context created with the map's scale and color setting and
 * large enough pools.  Without the parents block, a brick's parent is the
 * deepest brick above it in the tree.  After a restore, the next
 * ff_update_meshes meshes every cell. */
int ff_map_description(const ff_ctx *ctx, char *buf, size_t n);
int ff_save_map(ff_ctx *ctx, const char *path);
int ff_load_map(ff_ctx *ctx, const char *path, const float offset[3], uint32_t n);

/* ---- Loop-closure de-integration ---------------------------------------
 * The reference's loop-closure partner map (the FusionMipMapCPU(
 * FusionMipMapCPU*, nLeavesTotalLoop) constructor,
 * geometryfusion_mipmap_cpu.cpp:645-717): a dense store of `max_leaves` slots
 * (0: MAX_LOOP_CLOSURE_LEAVES = 100000, definitions.h:124), each with 512
 * float distances, 512 float weights (_weightsLoop) and, when the map fuses
 * color, 512 colors; _leafMapLoop gives each brick of the map its slot.
 * The partner reads the brick positions, scales and volume offset of `ctx`
 * as they are at each call (the reference copies them into the partner,
 * :3912-3932); `ctx` must outlive it.  Colors are copied as R, G, B planes
 * of 512 per slot. */
typedef struct ff_loop ff_loop;
ff_loop *ff_loop_create(ff_ctx *ctx, uint32_t max_leaves);
void ff_loop_destroy(ff_loop *lp);
const char *ff_loop_last_error(const ff_loop *lp);
/* updateArraysForLoopClosure's reset (geometryfusion_mipmap_cpu.cpp:3927-3931):
 * no slot handed out, every slot cleared. */
int ff_loop_reset(ff_loop *lp);
/* One frame of deleteOldLoopPoses (geometryfusion_mipmap_cpu.cpp:3990-4019):
 * updateWrapper_deleteloop (loopclosure.cpp:232-311) over the frame's
 * recorded brick queue with update8AddLoopSSESingleDeleteloop
 * (loopclosure.cpp:15-230).  depth: W x H floats in metres, NaN where there
 * is no measurement (IMAGEINFINITE); red, green, blue: W x H byte planes
 * (required when the map fuses color, else ignored); R, t: CameraInfo
 * camera -> world; K = {fx, fy, cx, cy}; queue: brick indices of the map,
 * in the order they are handed slots.  *applied (may be NULL) gets the
 * number of queue entries applied.  Returns FF_OK, or FF_E_POOL_FULL when a
 * brick found no free slot: the entries before it were applied and the rest
 * not (the reference returns false there); FF_E_ARG (nothing applied) when
 * an entry names a brick beyond the map's used bricks. */
int ff_loop_delete_frame(ff_loop *lp, const float *depth, const uint8_t *red, const uint8_t *green,
                         const uint8_t *blue, int width, int height, const double R[9], const double t[3],
                         const double K[4], const uint32_t *queue, size_t n_queue, size_t *applied);
/* Same, with every array already in device memory (dense rows). */
int ff_loop_delete_frame_device(ff_loop *lp, const float *depth_dev, const uint8_t *red_dev,
                                const uint8_t *green_dev, const uint8_t *blue_dev, int width, int height,
                                const double R[9], const double t[3], const double K[4],
                                const uint32_t *queue_dev, size_t n_queue, size_t *applied);
/* deleteOldLoopPoses' loop (geometryfusion_mipmap_cpu.cpp:3940-4022) as one
 * submission: frames[0..n_frames) applied in order, each as
 * ff_loop_delete_frame_device, with one synchronisation at the end.
 * applied[k] (may be NULL) gets frame k's applied entries.  A frame naming
 * a brick beyond the map's used bricks is skipped (FF_E_ARG after the
 * others are applied); FF_E_POOL_FULL when a frame stopped early. */
typedef struct ff_loop_frame {
    const float *depth;                /* device, W x H metres, NaN = no measurement */
    const uint8_t *red, *green, *blue; /* device planes, or NULL without color */
    int width, height;
    double R[9], t[3], K[4];           /* camera -> world, intrinsics fx fy cx cy */
    const uint32_t *queue;             /* device, n_queue brick indices */
    size_t n_queue;
} ff_loop_frame;
int ff_loop_delete_frames(ff_loop *lp, const ff_loop_frame *frames, int n_frames, size_t *applied);
/* _nLeavesMapped and the slot capacity. */
int ff_loop_mapped(ff_loop *lp, uint32_t *n_mapped, uint32_t *capacity);
/* _leafMapLoop for bricks [0, n): the slot, or 4294967294 (BRANCHINIT). */
int ff_loop_copy_map(ff_loop *lp, uint32_t *out, size_t n);
/* The partner's _distance, _weightsLoop and _color of the given slots. */
int ff_loop_copy_bricks(ff_loop *lp, const uint32_t *slots, size_t n, float *dist, float *weight,
                        uint16_t *color);
/* GPU timing of the de-integration kernel (no counterpart in the
 * reference): summed milliseconds, calls, algorithmic bytes and applied
 * queue entries. */
int ff_loop_set_timing(ff_loop *lp, int on);
int ff_loop_get_timing(ff_loop *lp, double *ms, uint64_t *calls, double *bytes, uint64_t *units);

/* ---- Input side: TUM association files and PNG images ------------------
 * No member of FusionMipMapCPU: the reference's viewer reads them
 * (onlinefusionviewer_main.cpp:266-346, onlinefusionviewer.cpp:934-935).
 * Host functions; they need no GPU and no context. */
typedef struct ff_tum_frame {
    double R[9], t[3]; /* camera -> world, composed across files like the reference */
    char depth[512];   /* association file's directory + the listed depth image */
    char rgb[512];     /* the same for the color image ("" without color) */
} ff_tum_frame;
/* Lines "stamp tx ty tz qx qy qz qw stamp depth [stamp rgb]" of n_files
 * association files, each file's poses composed with the last pose of the
 * previous file.  Writes at most cap frames; returns the number of frames
 * (possibly above cap), or a negative error. */
int ff_tum_read(const char *const *assoc_paths, int n_files, int use_color, ff_tum_frame *out, size_t cap);
/* PNG images as cv::imread reads them: ff_png_read_depth the samples of a
 * single-channel 8- or 16-bit image unchanged (IMREAD_UNCHANGED);
 * ff_png_read_bgr any 8/16-bit gray, RGB, gray+alpha, RGBA or palette image
 * as 8-bit B, G, R (IMREAD_COLOR).  Non-interlaced images only. */
int ff_png_size(const char *path, int *width, int *height, int *channels, int *bitdepth);
int ff_png_read_depth(const char *path, uint16_t *out, size_t n);
int ff_png_read_bgr(const char *path, uint8_t *out, size_t n);
const char *ff_input_last_error(void);

/* ---- Learned per-pixel confidence -------------------------------------
 * No counterpart in the reference, whose per-sample weight is the fixed
 * linear fall-off of getDistanceWeight_AoS_variables_h
 * (inline_functions.hpp:83-124, update_functions.cpp:1355-1403).  With
 * learned weighting on, that weight is multiplied by c(pixel) in (0, 1) from
 * an MLP over 27 per-pixel features (5x5 depth patch relative to the centre
 * depth, depth / 4 m, incidence cosine): 27 -> 64 -> 64 -> 64 -> 1, ReLU,
 * sigmoid, BF16 weights and activations with FP32 accumulation.
 * Parameters are FP32, FF_MLP_NPARAMS of them, laid out as W1[64][27], b1[64],
 * W2[64][64], b2[64], W3[64][64], b3[64], w4[64], b4 (row = output unit). */
#define FF_MLP_NPARAMS 10177
#define FF_MLP_RES_SCALE 0.0025f /* s of the training loss c r^2 / (2 s^2) - log(c) / 2, metres */
#define FF_MLP_RES_OUT 0.3f      /* residual of a sample outside the converged map's band, metres */
#define FF_MLP_GATE 0.002f       /* confidences below this reject the pixel (weight 0) */

/* Turns learned weighting on with the given parameters, or off (params NULL:
 * the plain method, bit-identical to the reference's arithmetic). */
int ff_set_learned_weighting(ff_ctx *ctx, const float *params, size_t n_params);
/* The last frame's per-pixel confidence (W x H floats; 0 at invalid depth
 * and where the confidence is under FF_MLP_GATE). */
int ff_copy_confidence(ff_ctx *ctx, float *out, size_t n);

/* Trainer of the confidence network on the device: samples are the features
 * of valid pixels with the residual r = TSDF of a converged map at the
 * pixel's back-projected point when that lies inside the map's truncation
 * band (|TSDF| < 0.9 x the brick's truncation), else +-FF_MLP_RES_OUT (the
 * sign of the saturated TSDF; + where the map has no weight).  Training is Adam on the mean loss over
 * batches drawn uniformly from the samples. */
typedef struct ff_trainer ff_trainer;
ff_trainer *ff_trainer_create(int device, size_t capacity, const float *init_params, size_t n_params,
                              uint64_t seed);
void ff_trainer_destroy(ff_trainer *tr);
const char *ff_trainer_last_error(const ff_trainer *tr);
/* Adds samples from one frame against the converged map `map` (host depth,
 * dense rows; pose camera -> world); keeps each valid pixel with probability
 * keep.  Returns the number of samples added. */
int ff_trainer_collect(ff_trainer *tr, ff_ctx *map, const uint16_t *depth, int w, int h, const double R[9],
                       const double t[3], const double K[4], float scaling, float maxcam, float keep);
int ff_trainer_size(const ff_trainer *tr, size_t *n);
/* Samples as 32 BF16 features (bit patterns; 27 used) and a residual each. */
int ff_trainer_copy_samples(ff_trainer *tr, uint16_t *feat, float *resid, size_t n);
int ff_trainer_set_samples(ff_trainer *tr, const uint16_t *feat, const float *resid, size_t n);
/* steps of Adam (lr, beta 0.9 / 0.999) on batches of `batch` samples
 * (a multiple of 2048); last_loss gets the mean loss of the last batch. */
int ff_trainer_train(ff_trainer *tr, int steps, uint32_t batch, float lr, float *last_loss);
/* Gradient of the mean loss over the given samples (n a multiple of 2048),
 * without an update: the check of the training kernels. */
int ff_trainer_grad(ff_trainer *tr, const uint32_t *index, uint32_t n, float *grad, float *loss);
int ff_trainer_params(ff_trainer *tr, float *out, size_t n_params);

/* ---- Batched streams ---------------------------------------------------
 * No counterpart in the reference (one FusionMipMapCPU per map): a tick
 * advances n independent contexts on one device by one frame each, with one
 * SDF launch over all of their queued bricks.  Per context the outcome is
 * that of ff_add_map_device (its counts, ff_last_error on failure); the
 * return value is FF_OK or the first error.  Inputs are device memory,
 * dense rows, borrowed for the call. */
#define FF_BATCH_MAX 64
typedef struct ff_frame {
    const uint16_t *depth;  /* u16 depth, device memory */
    const uint8_t *bgr;     /* BGR8, device memory; NULL without color */
    int width, height;
    double R[9], t[3], K[4]; /* camera -> world, intrinsics fx fy cx cy */
    float scaling, maxcam;
} ff_frame;
int ff_add_map_batch(ff_ctx *const *ctxs, const ff_frame *frames, int n);
/* updateMeshes for the n contexts of a batch, with one marching-cubes
 * launch over all of their re-meshed cells. */
int ff_update_meshes_batch(ff_ctx *const *ctxs, int n);
/* Launch configuration of the batch led by ctxs[0]: threads per brick (512
 * or 256: one or two voxels per thread; 64: one wave per brick, a row of 8
 * voxels per lane), bricks per workgroup (per wave for 64), HIP-graph replay
 * of the tick's device work.
 * By default the autotuner times every candidate on the first ticks and
 * keeps the fastest; set_config fixes one, autotune restarts the tuning. */
int ff_batch_set_config(ff_ctx *lead, int threads_per_brick, int bricks_per_workgroup, int graphs);
int ff_batch_autotune(ff_ctx *lead, int ticks_per_candidate);
int ff_batch_get_config(const ff_ctx *lead, int *threads_per_brick, int *bricks_per_workgroup, int *graphs,
                        int *tuning);

#ifdef __cplusplus
}
#endif
#endif
