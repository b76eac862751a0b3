// ff_batch.hip -- many independent map streams per tick on one GPU.
//
// The reference fuses one map per FusionMipMapCPU object; a tick here
// advances n such objects (contexts) by one frame each:
//   host     each stream's frame preparation (pose, first-frame volume);
//            the frame blocks and stream descriptors of all streams go to
//            the device in one copy
//   device   the allocation phase of all streams, one launch per kernel
//            (grid z = stream), the learned confidence of streams that use
//            it, then ONE SDF launch over the concatenation of every
//            stream's queued bricks (k_sdf_batch: a prefix over the
//            streams' gated counts maps worklist items to streams), and one
//            gather of the streams' frame counters
//   host     each stream's finish (used counts, errors, grow)
// The device part can be captured into a HIP graph and replayed per tick.
// Threads per brick (tpb), bricks per workgroup (bpw) and the launch mode are
// picked by an autotuner that times each candidate on the first ticks;
// every candidate computes the same maps.
#include "ff_device.hpp"
#include "ff_internal.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

namespace ff {

struct TuneCand {
    int tpb, bpw;
    bool graphs;
    double ms;
    int runs;
};

// per-tick device descriptors, one block so that a tick uploads them in one
// copy: frame blocks, allocation descriptors, SDF descriptors (cap each)
struct TickDesc {
    FrameDev *fd;
    AllocStream *as;
    SdfStream *ss;
};

static size_t desc_bytes(int cap, TickDesc *d, char *base)
{
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = (o + bytes + 255) & ~(size_t)255;
        return base ? base + at : nullptr;
    };
    char *a = take(sizeof(FrameDev) * cap);
    char *b = take(sizeof(AllocStream) * cap);
    char *c = take(sizeof(SdfStream) * cap);
    if (d) {
        d->fd = reinterpret_cast<FrameDev *>(a);
        d->as = reinterpret_cast<AllocStream *>(b);
        d->ss = reinterpret_cast<SdfStream *>(c);
    }
    return o;
}

struct BatchState {
    std::vector<ff_ctx *> ctxs;  // the batch (key)
    int cap = 0;
    char *d_desc = nullptr, *h_desc = nullptr;
    size_t descBytes = 0;
    TickDesc dd{}, hd{};  // device and host views of the descriptor block
    SdfStream *d_ss = nullptr, *h_ss = nullptr;
    uint32_t *d_prefix = nullptr;
    FrameCounters *d_fcs = nullptr, *h_fcs = nullptr;
    // launch configuration and its autotuning
    int tpb = 512, bpw = 1;
    bool graphs = false;
    bool tuning = true;
    int ticksPerCand = 2;
    std::vector<TuneCand> cands;
    size_t cur = 0;
    // captured device work of a tick
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t key = 0;
    uint64_t ticks = 0;
};

static void drop_graph(BatchState *b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void batch_free(ff_ctx *c)
{
    BatchState *b = c->batch;
    if (!b) return;
    drop_graph(b);
    hipFree(b->d_desc); hipHostFree(b->h_desc); hipFree(b->d_prefix); hipFree(b->d_fcs); hipHostFree(b->h_fcs);
    delete b;
    c->batch = nullptr;
}

static void default_cands(BatchState *b)
{
    b->cands.clear();
    for (int g = 0; g < 2; g++)
        for (int tpb : {512, 256, 64})
            for (int bpw : {1, 2, 4, 8}) b->cands.push_back({tpb, bpw, g == 1, 0.0, 0});
    b->cur = 0;
    b->tuning = true;
}

static BatchState *batch_for(ff_ctx *const *ctxs, int n)
{
    ff_ctx *lead = ctxs[0];
    BatchState *b = lead->batch;
    bool same = b && (int)b->ctxs.size() == n;
    for (int i = 0; same && i < n; i++) same = b->ctxs[i] == ctxs[i];
    if (same) return b;
    if (!b) {
        b = new BatchState();
        lead->batch = b;
        default_cands(b);
    }
    drop_graph(b);
    b->ctxs.assign(ctxs, ctxs + n);
    if (b->cap < n) {
        hipFree(b->d_desc); hipHostFree(b->h_desc); hipFree(b->d_prefix); hipFree(b->d_fcs); hipHostFree(b->h_fcs);
        b->d_desc = nullptr; b->h_desc = nullptr; b->d_prefix = nullptr; b->d_fcs = nullptr; b->h_fcs = nullptr;
        b->d_ss = nullptr; b->h_ss = nullptr;
        b->cap = 0;
        b->descBytes = desc_bytes(n, nullptr, nullptr);
        if (hipMalloc(&b->d_desc, b->descBytes) != hipSuccess ||
            hipHostMalloc(&b->h_desc, b->descBytes, hipHostMallocDefault) != hipSuccess ||
            hipMalloc(&b->d_prefix, sizeof(uint32_t) * (n + 1)) != hipSuccess ||
            hipMalloc(&b->d_fcs, sizeof(FrameCounters) * n) != hipSuccess ||
            hipHostMalloc(&b->h_fcs, sizeof(FrameCounters) * n, hipHostMallocDefault) != hipSuccess)
            return nullptr;
        desc_bytes(n, &b->dd, b->d_desc);
        desc_bytes(n, &b->hd, b->h_desc);
        b->d_ss = b->dd.ss;
        b->h_ss = b->hd.ss;
        b->cap = n;
    }
    return b;
}

struct TickPlan {
    BatchState *b;
    hipStream_t st;
    std::vector<int> dev;  // contexts with a volume (device work), index into ctxs
    std::vector<int> W, H;
};

static thread_local TickPlan *g_plan = nullptr;  // the plan being captured or launched

static void launch_tick(ff_ctx *lead, int, int)
{
    TickPlan &P = *g_plan;
    BatchState *b = P.b;
    const int nd = (int)P.dev.size();
    launch_alloc_batch(P.st, b->dd.as, b->hd.as, nd);
    for (int i : P.dev)
        if (b->ctxs[i]->learned) learned_launch(b->ctxs[i], P.W[i], P.H[i], b->ctxs[i]->stream);
    // events only in direct launches (lead != nullptr): the batched SDF time
    launch_batch_prefix(P.st, b->d_ss, nd, b->d_prefix);
    if (lead && lead->timing) hipEventRecord(lead->ev[2], P.st);
    launch_sdf_batch(P.st, b->d_ss, nd, b->d_prefix, b->tpb, b->bpw);
    if (lead && lead->timing) hipEventRecord(lead->ev[3], P.st);
    launch_gather_fc(P.st, b->d_ss, nd, b->d_fcs);
    hipMemcpyAsync(b->h_fcs, b->d_fcs, sizeof(FrameCounters) * nd, hipMemcpyDeviceToHost, P.st);
}

static uint64_t tick_key(const TickPlan &P)
{
    uint64_t k = 1469598103934665603ull;
    auto mix = [&](uint64_t v) { k = (k ^ v) * 1099511628211ull; };
    mix((uint64_t)P.b->tpb);
    mix((uint64_t)P.b->bpw);
    for (int i : P.dev) {
        const ff_ctx *c = P.b->ctxs[i];
        mix((uint64_t)i);
        mix((uint64_t)P.W[i] << 32 | (uint32_t)P.H[i]);
        mix((uint64_t)c->bufGen << 1 | (c->learned != nullptr));
    }
    return k;
}

}  // namespace ff

using namespace ff;

extern "C" {

int ff_add_map_batch(ff_ctx *const *ctxs, const ff_frame *frames, int n)
{
    if (!ctxs || !frames || n <= 0 || n > FF_BATCH_MAX) return FF_E_ARG;
    for (int i = 0; i < n; i++) {
        if (!ctxs[i]) return FF_E_ARG;
        for (int j = 0; j < i; j++)
            if (ctxs[j] == ctxs[i]) return FF_E_ARG;
        if (ctxs[i]->device != ctxs[0]->device) {
            ctxs[0]->err = "ff_add_map_batch: all contexts must be on one device";
            return FF_E_ARG;
        }
    }
    ff_ctx *lead = ctxs[0];
    hipSetDevice(lead->device);
    // a context's own stream may still hold its last mesh update: the
    // batch runs on the lead's stream
    for (int i = 0; i < n; i++) {
        const int rc = mesh_complete(ctxs[i]);
        if (rc != FF_OK) return rc;
    }
    BatchState *b = batch_for(ctxs, n);
    if (!b) {
        lead->err = "ff_add_map_batch: out of device memory";
        return FF_E_DEVICE;
    }
    const auto t0 = std::chrono::steady_clock::now();
    // every context runs on the lead's stream for the tick
    const hipStream_t st = lead->stream;
    std::vector<hipStream_t> own(n);
    for (int i = 0; i < n; i++) {
        own[i] = ctxs[i]->stream;
        if (own[i] != st) hipStreamSynchronize(own[i]);
        ctxs[i]->stream = st;
    }
    TickPlan P;
    P.b = b;
    P.st = st;
    P.W.assign(n, 0);
    P.H.assign(n, 0);
    std::vector<int> empty;
    int rc = FF_OK;
    for (int i = 0; i < n; i++) {
        ff_ctx *c = ctxs[i];
        const ff_frame &f = frames[i];
        if (!f.depth || (c->useColor && !f.bgr)) {
            c->err = "ff_add_map_batch: missing depth or color image";
            if (rc == FF_OK) rc = FF_E_ARG;
            continue;
        }
        int r = ensure_image(c, f.width, f.height);
        if (!r) r = frame_prepare(c, f.depth, c->useColor ? f.bgr : nullptr, f.width, f.height, f.R, f.t, f.K,
                                  f.scaling, f.maxcam, false);
        if (r) {
            if (rc == FF_OK) rc = r;
            continue;
        }
        P.W[i] = f.width;
        P.H[i] = f.height;
        if (c->n) {
            P.dev.push_back(i);
        } else {
            // no volume yet (an empty first frame): its own upload
            const hipError_t ec = hipMemcpyAsync(c->d_fd, c->h_fd, sizeof(FrameDev), hipMemcpyHostToDevice, st);
            if (ec != hipSuccess && rc == FF_OK) rc = FF_E_DEVICE;
            empty.push_back(i);
        }
    }
    // stream descriptors of the device work; each context reads its frame
    // block from the batch's block for the tick
    const int nd = (int)P.dev.size();
    std::vector<FrameDev *> ownFd(n, nullptr);
    for (int k = 0; k < nd; k++) {
        ff_ctx *c = ctxs[P.dev[k]];
        memcpy(&b->hd.fd[k], c->h_fd, sizeof(FrameDev));
        ownFd[P.dev[k]] = c->d_fd;
        c->d_fd = b->dd.fd + k;
        b->hd.as[k] = alloc_stream(c, P.W[P.dev[k]], P.H[P.dev[k]]);
        SdfStream &s = b->h_ss[k];
        s.queue = c->d_queue;
        s.m = c->mapView();
        s.fd = c->d_fd;
        s.fc = c->d_fc;
        s.mapc = c->d_mapc;
        s.gate.listCap = c->hv.listCap;
        s.gate.leavesTotal = c->nLeavesTotal;
        s.gate.branchesTotal = c->nBranchesTotal;
        s.gate.force = 0;
        s.W = P.W[P.dev[k]];
        s.H = P.H[P.dev[k]];
    }
    TuneCand *cand = nullptr;
    if (b->tuning && !b->cands.empty()) {
        cand = &b->cands[b->cur];
        b->tpb = cand->tpb;
        b->bpw = cand->bpw;
        b->graphs = cand->graphs;
    }
    if (nd) hipMemcpyAsync(b->d_desc, b->h_desc, b->descBytes, hipMemcpyHostToDevice, st);
    for (int i : empty) launch_empty(ctxs[i], P.W[i], P.H[i]);
    const auto t1 = std::chrono::steady_clock::now();
    g_plan = &P;
    bool timedSdf = false;
    if (nd) {
        bool launched = false;
        if (b->graphs) {
            const uint64_t key = tick_key(P);
            if (!b->exec || b->key != key) {
                drop_graph(b);
                hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
                if (e == hipSuccess) {
                    launch_tick(nullptr, 0, 0);
                    e = hipStreamEndCapture(st, &b->graph);
                }
                if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
                if (e != hipSuccess) {
                    (void)hipGetLastError();
                    drop_graph(b);
                } else {
                    b->key = key;
                }
            }
            if (b->exec && hipGraphLaunch(b->exec, st) == hipSuccess) launched = true;
        }
        if (!launched) {
            launch_tick(lead, 0, 0);
            timedSdf = lead->timing;
        }
    }
    g_plan = nullptr;
    hipError_t e = hipStreamSynchronize(st);
    const auto t2 = std::chrono::steady_clock::now();
    bool devOk = e == hipSuccess;
    if (!devOk) {
        lead->err = std::string("ff_add_map_batch: ") + hipGetErrorString(e);
        rc = FF_E_DEVICE;
    }
    if (devOk && timedSdf) {
        // the batched SDF update on the lead's timing: time, launches, bytes
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, lead->ev[2], lead->ev[3]) == hipSuccess) {
            lead->phaseMs[1] += ms;
            lead->phaseCalls[1]++;
            for (int k = 0; k < nd; k++) {
                const ff_ctx *c = ctxs[P.dev[k]];
                if (b->h_fcs[k].sdfRan)
                    lead->phaseBytes[1] += (double)b->h_fcs[k].nQueued * (c->useColor ? 14336.0 : 8192.0);
                lead->phaseUnits[1] += b->h_fcs[k].sdfRan ? b->h_fcs[k].nQueued : 0u;
            }
        }
    }
    if (devOk) {
        for (int k = 0; k < nd; k++) memcpy(ctxs[P.dev[k]]->h_fc, &b->h_fcs[k], sizeof(FrameCounters));
        for (int i = 0; i < n; i++) {
            if (!P.W[i]) continue;
            const int r = frame_finish(ctxs[i], P.W[i], P.H[i], false);
            if (r < 0 && rc == FF_OK) rc = r;
        }
    }
    for (int i = 0; i < n; i++) {
        ctxs[i]->stream = own[i];
        if (ownFd[i]) ctxs[i]->d_fd = ownFd[i];
    }
    // autotuning: the device part of each candidate's ticks, first tick of a
    // candidate (capture, warm-up) not counted
    if (cand && devOk) {
        const double ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
        if (cand->runs++ > 0) cand->ms += ms;
        if (cand->runs > b->ticksPerCand) {
            b->cur++;
            if (b->cur >= b->cands.size()) {
                size_t best = 0;
                for (size_t k = 1; k < b->cands.size(); k++)
                    if (b->cands[k].ms < b->cands[best].ms) best = k;
                b->tpb = b->cands[best].tpb;
                b->bpw = b->cands[best].bpw;
                b->graphs = b->cands[best].graphs;
                b->tuning = false;
            }
        }
    }
    (void)t0;
    b->ticks++;
    return rc;
}

int ff_batch_set_config(ff_ctx *lead, int tpb, int bpw, int graphs)
{
    if (!lead || (tpb != 512 && tpb != 256 && tpb != 64) || bpw < 1 || bpw > 64) return FF_E_ARG;
    if (!lead->batch) {
        lead->batch = new BatchState();
        default_cands(lead->batch);
    }
    BatchState *b = lead->batch;
    drop_graph(b);
    b->tpb = tpb;
    b->bpw = bpw;
    b->graphs = graphs != 0;
    b->tuning = false;
    return FF_OK;
}

int ff_batch_autotune(ff_ctx *lead, int ticks_per_candidate)
{
    if (!lead || ticks_per_candidate < 1) return FF_E_ARG;
    if (!lead->batch) lead->batch = new BatchState();
    BatchState *b = lead->batch;
    drop_graph(b);
    default_cands(b);
    b->ticksPerCand = ticks_per_candidate;
    return FF_OK;
}

int ff_batch_get_config(const ff_ctx *lead, int *tpb, int *bpw, int *graphs, int *tuning)
{
    if (!lead) return FF_E_ARG;
    const BatchState *b = lead->batch;
    if (tpb) *tpb = b ? b->tpb : 512;
    if (bpw) *bpw = b ? b->bpw : 1;
    if (graphs) *graphs = b ? (int)b->graphs : 0;
    if (tuning) *tuning = b ? (int)b->tuning : 1;
    return FF_OK;
}

}  // extern "C"
