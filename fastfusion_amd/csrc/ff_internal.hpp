// ff_internal.hpp -- context and shared types of the fusion library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/ff_abi.h"

#define FF_HEMPTY 0xffffffffffffffffull
#define FF_FLAG_HASH_FULL 1u
#define FF_FLAG_POOL_FULL 2u
#define FF_FLAG_ORDER_BIG 4u  // more new nodes than k_order ranks: the host orders them by the per-pixel scan

// processed-pixel bitmap: 32 pixels per word, rows FF_PB_STRIDE words apart,
// FF_PB_PAD rows of padding above the image and 2 FF_PB_PAD below it (the
// de-dup sweep's skewed lanes store up to H + 158 rows past the top)
#define FF_PB_STRIDE 64
#define FF_PB_PAD 64
// image size limits (the de-dup ring has 64 words of columns; the ray
// tables are allocated with the context at these sizes)
#define FF_MAX_W 2048
#define FF_MAX_H 2048
inline size_t pbits_bytes(int H) { return sizeof(uint32_t) * FF_PB_STRIDE * (size_t)(H + 3 * FF_PB_PAD); }

namespace ff {

// camPamsFloat (src/fusion/definitions.h:70-89)
struct CamF {
    float r[9], t[3], fx, fy, cx, cy;
};

struct FrameCounters {
    uint32_t nQueued, listCount, newB, newL;
    int bmin[3], bmax[3];
    uint32_t flags, nProcessed;
    uint32_t sdfRan;  // set by the SDF update when it ran for the frame
    uint32_t pad[3];  // 64 bytes: a kernel reads them as four 16-byte words
};
static_assert(sizeof(FrameCounters) == 64, "FrameCounters is read as four 16-byte words");

// what the SDF update checks on the device before running: a frame whose
// node table overflowed, or that must grow the volume first, waits for the host
struct SdfGate {
    uint32_t listCap, leavesTotal, branchesTotal;
    int force;
};

// device views of the map arrays
struct MapView {
    uint32_t *tree;
    ushort4 *leafinfo;  // x, y, z, scale  (_leafPos, _leafScale)
    uint32_t *parent;   // _leafParent
    uint32_t *stamp;    // frame stamp of the last queueing
    uint32_t *queue;    // _leafNumberSurface
    float *dist;        // 512 per brick
    float *wgt;         // 512 per brick
    uint16_t *col;      // R, G, B planes of 512 per brick
    uint32_t nLeavesTotal, nBranchesTotal;
};

// nodes created during one frame
struct HashView {
    uint64_t *keys;
    uint32_t *minpix, *aux, *assigned, *list;
    uint32_t mask, listCap;
};

struct SdfParams {
    float m[12];
    float fx, fy, cx, cy, scale, dT, scaling, maxcam;
    int useColor;
    const float *conf;  // per-pixel confidence multiplying the weight, or nullptr
};

// per-frame values the kernels read from device memory, so that a frame's
// launches can be captured once and replayed from a graph
struct FrameDev {
    CamF p;                 // camera -> voxel coordinates (back-projection)
    SdfParams sp;           // SDF update (voxel -> camera, intrinsics, confidence)
    const uint16_t *depth;  // the frame in device memory
    const uint8_t *bgr;
    float scaling, maxcam, ifx, ify;  // ifx, ify: 1 / fx, 1 / fy
    uint32_t stamp, n, bandwidth, pad;
};

// the map's used counts, committed on the device at the end of each
// allocation (the host mirror follows after the frame's synchronisation)
struct MapCounters {
    uint32_t nBranchesUsed, nLeavesUsed, pad[2];
};
static_assert(sizeof(MapCounters) == 16, "MapCounters is read as one 16-byte word");

// one stream of a batched SDF launch
struct SdfStream {
    const uint32_t *queue;
    MapView m;
    const FrameDev *fd;
    FrameCounters *fc;
    const MapCounters *mapc;
    SdfGate gate;
    int W, H;
};

// one stream of a batched allocation phase: every buffer and size the
// allocation kernels of that stream take (grid z = stream)
struct AllocStream {
    const FrameDev *fd;
    const float *pxp, *pyp;
    FrameCounters *fc;
    uint32_t *planes, *cntB, *cntL, *firstD, *pbits, *pl0;
    uint64_t *pref, *bsum;
    MapCounters *mapc;
    ushort4 *binfo;  // mesh-side branch geometry (k_alloc_end)
    MapView m;
    HashView hv;
    int W, H, words;
    uint32_t binfoCap;
    uint32_t nLeavesUsed;  // host's count before the frame (k_queue's grid)
};

struct MeshState;
struct LearnedState;
struct BatchState;

}  // namespace ff

struct ff_ctx {
    // parameters
    float scale = 0, dT = 0;
    bool useColor = true, incremental = true;
    bool broken = false;  // a frame overflowed the per-frame node table: no further frames
    int device = 0;
    uint16_t bandwidth = 0;
    // map scalars (host mirror)
    uint16_t n = 0;
    float offset[3] = {0, 0, 0};
    uint32_t nLeavesTotal = 0, nBranchesTotal = 0, nLeavesUsed = 0, nBranchesUsed = 16, nQueued = 0;
    uint32_t frames = 0, grows = 0, stampCounter = 0;
    int imgW = 0, imgH = 0, capW = 0, capH = 0;
    std::vector<float> pxp, pyp;
    hipStream_t stream = nullptr;
    // a frame (bit planes, de-dup, descent, node commit, queue, SDF update)
    // runs on `side`, beside the previous mesh update on `stream`, and only
    // the SDF update waits: for the update's sampling (sequence word
    // FF_SEQ_SAMPLE), after which its marching cubes read the sample buffer
    // and no brick, unless some cells' samples did not fit the buffer (the
    // update's tile count above its capacity): then the SDF update stands
    // down for the host, the device and the host both deciding from that
    // count (pendingTiles, reported to h_rep).  The descent and the
    // allocation may run beside the update's partition and marking: the
    // update reads the tree only through the bricks and branch slots counted
    // when it was launched (a later frame's links point past them), and a
    // brick the next frame stamps can only make a cell dirty, never clean
    hipStream_t side = nullptr;
    // Order between the streams: sequence words in device memory, one per
    // kind of producer; the producing stream writes the word's next value
    // after its work (a one-thread kernel, or the first thread of the
    // kernel that follows the work there) and a consuming stream waits
    // until the word reaches it (hipStreamWaitValue32, ~2 us of host time;
    // ~4 us from the producer's end to the consumer's start, where an event
    // wait on a pending event costs ~16 us of host time and ~10 us).
    // seq[k]: the last value enqueued.
    uint32_t *d_seq = nullptr;
    uint32_t seq[4] = {0, 0, 0, 0};
    bool seqOk = false;  // the runtime takes stream memory operations
    // such a frame returns once its SDF update has started (the update's
    // first block writes the frame's counters, then the frame's stamp, to
    // mapped host memory); the next mesh update waits for FF_SEQ_SDF on
    // `stream`, and every other reader of the map joins the frame first
    // (frame_join); sdfInFlight until then
    bool sdfInFlight = false;
    // an SDF update timed by its dispatch's events whose frame returned
    // before it ended: its time is read at the next frame or join
    bool sdfTimedPending = false;
    double sdfTimedBytes = 0.0;
    uint32_t sdfTimedUnits = 0;
    // the learned confidence of an overlapped frame runs on `aux`, beside the
    // frame's de-dup and descent; its SDF update waits for FF_SEQ_CONF
    hipStream_t aux = nullptr;
    bool uploadOnSide = false;  // the frame in flight had its host images copied up on `side`
    // the last mesh update's sampling signalled FF_SEQ_SAMPLE and no frame
    // has waited for it yet; its tile count (device), the buffer's capacity
    // then, and its number
    bool samplePending = false;
    const uint32_t *pendingTiles = nullptr;
    uint32_t pendingTileCap = 0, pendingId = 0;
    // the SDF update of the frame in flight was gated on the tile count of
    // update sdfTileId (0: on none)
    uint32_t sdfTileId = 0, sdfTileCap = 0;
    // a frame is enqueueing a requested mesh update: the update leaves its
    // hash turnover for the frame to launch on `side` after its SDF update
    bool frameInFlush = false;
    // map arrays
    uint32_t *d_tree = nullptr, *d_treeTmp = nullptr, *d_parent = nullptr, *d_stamp = nullptr, *d_queue = nullptr;
    ushort4 *d_leafinfo = nullptr;
    float *d_dist = nullptr, *d_wgt = nullptr;
    uint16_t *d_col = nullptr;
    // frame buffers
    uint16_t *d_depth = nullptr;
    uint8_t *d_bgr = nullptr;
    float *d_pxp = nullptr, *d_pyp = nullptr;  // per-column / per-row ray tables (FF_MAX_W / FF_MAX_H entries)
    // where the last frame left its processed-pixel bitmap and confidences
    // (the context's buffers, or the device's scratch array after
    // ff_add_map_device_noalloc)
    const uint32_t *lastPbits = nullptr;
    const float *lastConf = nullptr;
    const float *prevConf = nullptr;  // the frame before's (its SDF update may still read it)
    bool directOnly = false;  // the frame in flight launches directly (no graph capture or replay)
    uint32_t *d_planes = nullptr, *d_pbits = nullptr, *d_cntB = nullptr, *d_cntL = nullptr, *d_firstD = nullptr,
             *d_pl0 = nullptr;
    uint64_t *d_pref = nullptr, *d_bsum = nullptr;
    ff::FrameCounters *d_fc = nullptr, *h_fc = nullptr;
    ff::FrameCounters *d_hfc = nullptr;  // h_fc (mapped host memory) as the device addresses it
    ff::FrameDev *d_fd = nullptr, *h_fd = nullptr;     // h_fd: pinned staging
    ff::MapCounters *d_mapc = nullptr, *h_mapc = nullptr;
    // the captured per-frame launches: [0] allocation, [1] confidence + SDF
    hipGraph_t frameGraph[2] = {nullptr, nullptr};
    hipGraphExec_t frameExec[2] = {nullptr, nullptr};
    uint64_t frameKey = 0;     // what the captured graph depends on
    uint32_t bufGen = 0;       // bumped when buffers a graph refers to are replaced
    bool useGraphs = false;    // direct launches measure faster for one stream
    uint32_t timeEvery = 1;    // the SDF update is timed on one frame in timeEvery
    bool sdfTimed = false;     // the frame in flight records the SDF phase's events
    ff::HashView hv{};
    ff::MeshState *mesh = nullptr;
    ff::LearnedState *learned = nullptr;  // learned confidence (off: nullptr)
    ff::BatchState *batch = nullptr;      // batched ticks led by this context
    std::string err;
    // per-phase timing (0 allocation, 1 SDF update, 2 meshing)
    bool timing = false;     // events around the SDF update
    bool timingAll = false;  // and around the allocation and meshing phases
    hipEvent_t ev[10] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double phaseMs[3] = {0, 0, 0}, phaseBytes[3] = {0, 0, 0};
    uint64_t phaseCalls[3] = {0, 0, 0}, phaseUnits[3] = {0, 0, 0};
    // the marching-cubes kernel alone: its dispatch records ev[8], ev[9] on
    // one update in timeEvery (durations); with all phases timed, its
    // algorithmic bytes and re-meshed cells on every update
    uint64_t updates = 0;
    double mcMs = 0, mcBytes = 0;
    uint64_t mcCalls = 0, mcByteCalls = 0, mcUnits = 0;

    ff::MapView mapView() const
    {
        ff::MapView m;
        m.tree = d_tree; m.leafinfo = d_leafinfo; m.parent = d_parent; m.stamp = d_stamp; m.queue = d_queue;
        m.dist = d_dist; m.wgt = d_wgt; m.col = d_col;
        m.nLeavesTotal = nLeavesTotal; m.nBranchesTotal = nBranchesTotal;
        return m;
    }
    void freeFrameBuffers();
};

namespace ff {
// learned confidence (ff_learned.hip): the per-pixel buffer for a W x H
// frame (allocated on first use), and the launch filling it from c->d_fd
void learned_free(ff_ctx *c);
const float *learned_prepare(ff_ctx *c, int W, int H, bool alternate);
void learned_launch(ff_ctx *c, int W, int H, hipStream_t st);
void learned_launch_direct(ff_ctx *c, int W, int H, hipStream_t st);
void learned_swap_conf(ff_ctx *c, float *buf, size_t cap, float **old, size_t *oldCap);
void sync_map_counters(ff_ctx *c);
void drop_frame_graphs(ff_ctx *c);
// a frame in phases (ff_fusion.hip): host preparation and parameter upload,
// device work (allocation, then confidence + SDF; or the empty-frame path),
// and the host part after the device work completed
// upload = false: the frame block is left in c->h_fd for the caller to copy
// overlap: the frame runs on the side stream (its confidences then go to
// the other of two buffers)
int frame_prepare(ff_ctx *c, const uint16_t *depth, const uint8_t *bgr, int W, int H, const double R[9],
                  const double t[3], const double K[4], float scaling, float maxcam, bool upload = true,
                  bool overlap = false);
void launch_alloc(ff_ctx *c, int W, int H);
// launch_alloc in two parts: the front reads the frame and writes only the
// bit planes, the processed-pixel bitmap and the frame counters (nothing of
// the map, nothing the meshing reads); the back descends and allocates
void launch_front(ff_ctx *c, int W, int H, hipStream_t st, bool carry);
void launch_back(ff_ctx *c, int W, int H, hipStream_t st);
void launch_sdf(ff_ctx *c, int W, int H);
void launch_empty(ff_ctx *c, int W, int H);
int frame_finish(ff_ctx *c, int W, int H, bool timed);
// the frame in flight joined (a no-op when none is): stream = true enqueues a
// wait for its SDF update on `stream` (the mesh update's path), false waits
// on the host; later work on `stream` and host reads see the whole frame
void frame_join(ff_ctx *c, bool stream = false);
// waits until the mapped host word *w equals want (a kernel's report);
// false when the stream drained or failed without writing it
bool await_word(const uint32_t *w, uint32_t want, hipStream_t st, bool atLeast = false);
// the sequence words (ff_ctx::d_seq): after the mesh update's sampling,
// after a frame's SDF update (and the hash turnover launched behind it),
// after the learned confidence, after a frame's host images went up on the
// side stream
enum { FF_SEQ_SAMPLE = 0, FF_SEQ_SDF = 1, FF_SEQ_CONF = 2, FF_SEQ_UP = 3 };
// enqueues on st: word k takes its next value; false (st drained instead)
// when the runtime refused
bool seq_signal(ff_ctx *c, int k, hipStream_t st);
// enqueues on st a wait until word k holds the last value signalled (one
// already signalled: a wait for work enqueued before)
void seq_wait(ff_ctx *c, int k, hipStream_t st);
int ensure_image(ff_ctx *c, int W, int H);
// addMap's pInv: voxel -> camera with the map offset folded in
// (geometryfusion_mipmap_cpu.cpp:1771-1793)
CamF camera_inverse(const ff_ctx *c, const double R[9], const double t[3], const double K[4]);
// batched SDF update over the concatenated brick lists of n streams
// (ff_fusion.hip): per-stream gates and counts into prefix[0..n], then one
// launch; threads per brick tpb (512 or 256), bricks per workgroup bpw
void launch_batch_prefix(hipStream_t st, const SdfStream *d_ss, int n, uint32_t *d_prefix);
void launch_sdf_batch(hipStream_t st, const SdfStream *d_ss, int n, uint32_t *d_prefix, int tpb, int bpw);
void launch_gather_fc(hipStream_t st, const SdfStream *d_ss, int n, FrameCounters *d_out);
void batch_free(ff_ctx *c);
// meshing module hooks (ff_mesh.hip)
int mesh_init(ff_ctx *c);
// the host part of an enqueued mesh update (a no-op when none is pending):
// every mesh query and every change of the map's structure calls it first.
// maintain = false (inside a frame) leaves its allocating part (buffer
// growth, pool compaction) to the next call with maintain = true
int mesh_complete(ff_ctx *c, bool maintain = true);
// enqueues a requested mesh update (updateMeshes records the request; the
// next frame enqueues it after its first kernels, any other call first)
int mesh_flush(ff_ctx *c);
// the tile count update `id` reported when its sampling began (false: the
// stream drained without the report)
bool mesh_sample_report(ff_ctx *c, uint32_t id, uint32_t *nTiles);
// the hash turnover an update enqueued by a frame left to it, on st
void mesh_launch_deferred_turnover(ff_ctx *c, hipStream_t st);
void mesh_free(ff_ctx *c);
void mesh_on_initial_volume(ff_ctx *c);
int mesh_on_load(ff_ctx *c, const std::vector<ushort4> &binfo);
void fill_tree_tail(ff_ctx *c, uint32_t from);
void mesh_on_grow(ff_ctx *c);
void mesh_after_frame(ff_ctx *c);
// the mesher's branch-geometry table (origin and half size per branch
// block), filled by the allocation phase
void mesh_binfo(const ff_ctx *c, ushort4 **binfo, uint32_t *cap);
// the batched allocation phase of n streams (descriptors in device memory
// d_as, host copy h_as for the grid sizes)
AllocStream alloc_stream(ff_ctx *c, int W, int H);
void launch_alloc_batch(hipStream_t st, const AllocStream *d_as, const AllocStream *h_as, int n);
void mesh_on_grow_blocks(ff_ctx *c, uint32_t oldBlocks, uint32_t addBlocks, const uint32_t addv[3],
                         const std::vector<ushort4> &chain);
}  // namespace ff
