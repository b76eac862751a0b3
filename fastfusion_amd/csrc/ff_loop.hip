// ff_loop.hip -- loop-closure de-integration on gfx950.
//
// The reference subtracts the frames whose pose a loop closure changed from
// a partner map (FusionMipMapCPU(FusionMipMapCPU*, nLeavesTotalLoop),
// src/fusion/geometryfusion_mipmap_cpu.cpp:645-717, driven by
// deleteOldLoopPoses, :3934-4024).  The partner numbers bricks densely in the
// order they are first touched: _leafMapLoop maps a brick of the main map to
// its slot, _nLeavesMapped counts the slots handed out, and a slot holds 512
// float distances, 512 float weights (_weightsLoop) and 512 colors.  Per
// frame, updateWrapper_deleteloop (src/fusion/loopclosure.cpp:232-311) walks
// the frame's recorded brick queue in order, hands a slot to each brick that
// has none (stopping when none is left), and runs
// update8AddLoopSSESingleDeleteloop (loopclosure.cpp:15-230) on it.
//
// Here the partner lives in HBM next to the main map, with the slot count
// in device memory, so that any number of frames (deleteOldLoopPoses' loop,
// :3940-4022) go out as one submission with one synchronisation at the end.
// Per frame, on the map's stream:
//   k_loop_map1    queues up to 8192 entries: one workgroup claims each
//                  brick's first entry, hands out slots in queue order, finds
//                  where the hand-out stops when the store is full and counts
//                  per brick the entries to apply (a brick queued twice is
//                  updated twice)
//   k_loop_claim / k_loop_flags / k_loop_assign
//                  the same for longer queues, over many workgroups
//   k_loop_delete  a 512-thread block per brick, one voxel per thread: the
//                  de-integration update in the hardware round-toward-zero
//                  mode with the RCPPS table, repeated per entry
// The brick positions, scales and the volume offset are read from the main
// map as they are at the call.
#include "ff_device.hpp"
#include "ff_internal.hpp"
#include "ff_sdf.hpp"
#include "rcp_table.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <string>

namespace ff {

__constant__ uint32_t c_rcp_loop[2048];

// the slot count (_nLeavesMapped), kept in device memory across frames
struct LoopState {
    uint32_t nMapped, pad[3];
};

// per frame
struct LoopCounters {
    uint32_t failpos;      // first queue entry not applied (set to the queue length before the frame)
    uint32_t bad;          // a queue entry names a brick beyond the map's used bricks (0 before)
    uint32_t nMappedNext;  // the slot count after the frame
    uint32_t pad;
};

// device pointers and sizes the loop kernels share
struct LoopArgs {
    const ushort4 *leafinfo;
    uint32_t *map, *first, *count;
    float *dist, *wgt;
    uint16_t *col;
    const float *depth;
    const uint8_t *red, *green, *blue;
    int W, H;
    uint32_t nLeavesUsed, total;
};

#define FF_LOOP_TPB 1024
#define FF_LOOP_SMALL 8192  // queues up to this length are mapped by one workgroup
#define FF_LOOP_GRID 8192   // blocks of k_loop_delete; each loops over queue entries
#define FF_LOOP_TIME_EVERY 8  // with timing on, frames of a submission timed: one in 8
#define FF_UNCLAIMED 0xffffffffu

// a claim made by another thread of the same kernel: read at L2, not from a
// possibly stale L1 line
__device__ __forceinline__ uint32_t load_claim(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_loop_claim(const uint32_t *__restrict__ queue, uint32_t n, LoopArgs A, LoopCounters *lc)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t b = queue[i];
    if (b >= A.nLeavesUsed) {
        atomicOr(&lc->bad, 1u);
        return;
    }
    atomicMin(&A.first[b], i);
}

// entry i takes a new slot: its brick's first entry, without a slot yet (the
// reference's test threadOffset >= _nLeavesMapped, loopclosure.cpp:278)
__device__ __forceinline__ bool loop_takes_slot(uint32_t i, uint32_t b, const LoopArgs &A, uint32_t nMapped)
{
    return b < A.nLeavesUsed && load_claim(&A.first[b]) == i && A.map[b] >= nMapped;
}

__global__ __launch_bounds__(FF_LOOP_TPB) void k_loop_flags(const uint32_t *__restrict__ queue, uint32_t n,
                                                            LoopArgs A, const LoopState *st,
                                                            uint32_t *__restrict__ bsum)
{
    const uint32_t i = blockIdx.x * FF_LOOP_TPB + threadIdx.x;
    const uint32_t nMapped = st->nMapped;
    const bool nw = i < n && loop_takes_slot(i, queue[i], A, nMapped);
    const int cnt = __syncthreads_count(nw);
    if (threadIdx.x == 0) bsum[blockIdx.x] = (uint32_t)cnt;
}

__device__ __forceinline__ uint32_t block_sum_1024(uint32_t v, uint32_t *red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();  // red[] free from an earlier use
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();  // the 16 wave sums written
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < FF_LOOP_TPB / 64; k++) s += red[k];
    return s;
}

__global__ __launch_bounds__(FF_LOOP_TPB) void k_loop_assign(const uint32_t *__restrict__ queue, uint32_t n,
                                                             LoopArgs A, const LoopState *st, LoopCounters *lc,
                                                             const uint32_t *__restrict__ bsum, uint32_t nb)
{
    __shared__ uint32_t red[FF_LOOP_TPB / 64], wsum[FF_LOOP_TPB / 64];
    const uint32_t nMapped = st->nMapped;
    if (lc->bad) {  // the call is rejected: nothing is mapped or applied
        if (blockIdx.x == 0 && threadIdx.x == 0) lc->nMappedNext = nMapped;
        return;
    }
    // new slots taken by the entries before this block, and by all entries
    uint32_t before = 0, all = 0;
    for (uint32_t j = threadIdx.x; j < nb; j += FF_LOOP_TPB) {
        const uint32_t v = bsum[j];
        all += v;
        if (j < blockIdx.x) before += v;
    }
    before = block_sum_1024(before, red);
    all = block_sum_1024(all, red);
    const uint32_t avail = A.total - nMapped;
    const uint32_t i = blockIdx.x * FF_LOOP_TPB + threadIdx.x;
    uint32_t b = 0;
    bool nw = false;
    if (i < n) {
        b = queue[i];
        nw = loop_takes_slot(i, b, A, nMapped);  // reads map[b] only for b's first entry, before it is set
    }
    const uint64_t bal = __ballot(nw);
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) wsum[wid] = (uint32_t)__popcll(bal);
    __syncthreads();  // wave counts written
    uint32_t woff = 0;
    for (uint32_t k = 0; k < wid; k++) woff += wsum[k];
    // P: new slots taken by the entries before i in queue order.  Entry i is
    // applied iff the hand-out has not stopped before it: fewer than `avail`
    // new slots before it, or exactly `avail` and it needs none.
    const uint32_t P = before + woff + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
    if (i < n) {
        if (P < avail || (P == avail && !nw)) {
            if (nw) A.map[b] = nMapped + P;
            atomicAdd(&A.count[b], 1u);
        } else if (nw && P == avail) {
            lc->failpos = i;  // the one entry that found the store full
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) lc->nMappedNext = nMapped + min(all, avail);
}

// The three kernels above in one workgroup, for queues of at most
// FF_LOOP_SMALL entries: claims, then the hand-out chunk by chunk with a
// running count of the slots taken.
__global__ __launch_bounds__(FF_LOOP_TPB) void k_loop_map1(const uint32_t *__restrict__ queue, uint32_t n,
                                                           LoopArgs A, const LoopState *st, LoopCounters *lc)
{
    __shared__ uint32_t wsum[FF_LOOP_TPB / 64];
    __shared__ uint32_t s_bad, s_fail;
    const uint32_t nMapped = st->nMapped;
    if (threadIdx.x == 0) {
        s_bad = 0u;
        s_fail = n;
    }
    __syncthreads();  // flags initialised
    for (uint32_t i = threadIdx.x; i < n; i += FF_LOOP_TPB) {
        const uint32_t b = queue[i];
        if (b >= A.nLeavesUsed)
            s_bad = 1u;
        else
            atomicMin(&A.first[b], i);
    }
    __threadfence();  // the claims are performed before any thread reads them
    __syncthreads();
    if (s_bad) {  // the frame is rejected; k_loop_delete releases the claims
        if (threadIdx.x == 0) {
            lc->bad = 1u;
            lc->failpos = 0u;
            lc->nMappedNext = nMapped;
        }
        return;
    }
    const uint32_t avail = A.total - nMapped;
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t running = 0;
    for (uint32_t base = 0; base < n; base += FF_LOOP_TPB) {
        const uint32_t i = base + threadIdx.x;
        uint32_t b = 0;
        bool nw = false;
        if (i < n) {
            b = queue[i];
            nw = loop_takes_slot(i, b, A, nMapped);
        }
        const uint64_t bal = __ballot(nw);
        if (lane == 0) wsum[wid] = (uint32_t)__popcll(bal);
        __syncthreads();  // wave counts of this chunk written
        uint32_t woff = 0, tot = 0;
#pragma unroll
        for (uint32_t k = 0; k < FF_LOOP_TPB / 64; k++) {
            const uint32_t v = wsum[k];
            woff += (k < wid) ? v : 0u;
            tot += v;
        }
        const uint32_t P = running + woff + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
        if (i < n) {
            if (P < avail || (P == avail && !nw)) {
                if (nw) A.map[b] = nMapped + P;
                atomicAdd(&A.count[b], 1u);
            } else if (nw && P == avail) {
                s_fail = i;  // the one entry that found the store full
            }
        }
        running += tot;
        __syncthreads();  // wsum read by every thread before the next chunk
    }
    if (threadIdx.x == 0) {
        lc->failpos = s_fail;
        lc->nMappedNext = nMapped + min(running, avail);
    }
}

// One voxel of update8AddLoopSSESingleDeleteloop (loopclosure.cpp:56-225):
// column x of row `w`, old (d, wt, c) -> new.  The x loop runs two groups of
// four SSE lanes; the second group starts at the row origin plus 4 d11
// (pxx += d11 after d11 *= 4) and lane k adds 0, d11, 2 d11, 3 d11.
__device__ __forceinline__ void loop_point(const SdfRow &w, int x, float &d, float &wt, uint16_t c[3],
                                           const LoopArgs &A, const SdfParams &P, const uint32_t *tab)
{
    const float EPS = 0.005f;
    const int g = x >> 2, k = x & 3;
    float bx = w.R[0], by = w.R[1], bz = w.R[2];
    if (g) {
        bx = rz_add(bx, rz_mul(w.d1[0], 4.0f));
        by = rz_add(by, rz_mul(w.d1[1], 4.0f));
        bz = rz_add(bz, rz_mul(w.d1[2], 4.0f));
    }
    const float fk = (float)k;
    const float px = rz_add(bx, k ? rz_mul(fk, w.d1[0]) : 0.0f);
    const float py = rz_add(by, k ? rz_mul(fk, w.d1[1]) : 0.0f);
    const float pz = rz_add(bz, k ? rz_mul(fk, w.d1[2]) : 0.0f);
    const float len = rz_sqrt(rz_add(rz_add(rz_mul(px, px), rz_mul(py, py)), rz_mul(pz, pz)));
    const float r = rcp_x86(pz, tab);
    const int W = A.W, H = A.H;
    const int imx = cvt_x86(rz_add(P.cx, rz_mul(P.fx, rz_mul(px, r))));
    const int imy = cvt_x86(rz_add(P.cy, rz_mul(P.fy, rz_mul(py, r))));
    const int cyi = max(min(imy, H - 1), 0), cxi = max(min(imx, W - 1), 0);
    const int ii = W * cyi + cxi;
    const float h = A.depth[ii];
    const float dI = rz_sub(len, rz_mul(rz_mul(len, r), h));
    // _mm_cmpord_ps(dInc, dInc): a NaN depth (no measurement) masks the voxel
    const bool mask = (pz > 0.0f) && (imy < H) && (imy > -1) && (imx < W) && (imx > -1) && (dI == dI);
    const float front = (dI < EPS) ? 1.0f : 0.0f;
    const float back = (dI >= EPS && dI < w.tw) ? 1.0f : 0.0f;
    const float fall = rz_mul(rz_sub(dI, w.tw), w.rfall);  // the reference's reversed sign
    const float wI = mask ? rz_add(front, rz_mul(fall, back)) : 0.0f;
    const float wA = wt, dA = d;
    const float wn = rz_add(wA, wI);
    const float f = rcp_x86(rz_add(wn, (wn <= 0.0f) ? 1.0f : 0.0f), tab);
    const float mn = (w.td < dI) ? w.td : dI;  // _mm_min_ps(td, dI)
    const float ntd = -w.td;
    const float mx = (ntd > mn) ? ntd : mn;    // _mm_max_ps(-td, mn)
    const float add = mask ? rz_mul(mx, wI) : 0.0f;
    d = rz_mul(rz_add(rz_mul(dA, wA), add), f);
    wt = wn;
    if (P.useColor) {
        const uint8_t *plane[3] = {A.red, A.green, A.blue};
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            const float inc = (float)plane[ch][ii] * 256.0f;
            const float v = rz_mul(rz_add(rz_mul((float)c[ch], wA), rz_mul(inc, wI)), f);
            c[ch] = (uint16_t)cvt_x86(v);  // colortype = float: truncation, low 16 bits, no clamp
        }
    }
}

__global__ __launch_bounds__(512) void k_loop_delete(const uint32_t *__restrict__ queue, uint32_t n, LoopArgs A,
                                                     SdfParams P, LoopState *st, const LoopCounters *lc)
{
    rz_mode_on();  // first: every float operation below rounds toward zero
    __shared__ uint32_t tab[2048];
    for (int i = threadIdx.x; i < 2048; i += 512) tab[i] = c_rcp_loop[i];
    const uint32_t failpos = lc->failpos;
    const bool bad = lc->bad != 0u;
    __syncthreads();  // rcp table ready; counters read
    if (blockIdx.x == 0 && threadIdx.x == 0) st->nMapped = lc->nMappedNext;
    const int v = threadIdx.x;
    for (uint32_t e = blockIdx.x; e < n; e += gridDim.x) {
        const uint32_t b = queue[e];
        // block-uniform: only the brick's first entry updates it, and only
        // this block resets that entry's claim (after the barrier below)
        if (b >= A.nLeavesUsed || A.first[b] != e) continue;
        const uint32_t reps = (!bad && e < failpos) ? A.count[b] : 0u;
        if (reps) {
            const size_t slot = A.map[b];
            const SdfRow w = sdf_row_setup(A.leafinfo[b], v >> 6, (v >> 3) & 7, P, tab);
            const size_t vi = slot * FF_BRICK + v;
            float d = A.dist[vi], wt = A.wgt[vi];
            uint16_t c[3] = {0, 0, 0};
            uint16_t *cb = A.col + slot * FF_BRICK * 3 + v;
            if (P.useColor) {
                c[0] = cb[0];
                c[1] = cb[FF_BRICK];
                c[2] = cb[2 * FF_BRICK];
            }
            for (uint32_t r = 0; r < reps; r++) loop_point(w, v & 7, d, wt, c, A, P, tab);
            A.dist[vi] = d;
            A.wgt[vi] = wt;
            if (P.useColor) {
                cb[0] = c[0];
                cb[FF_BRICK] = c[1];
                cb[2 * FF_BRICK] = c[2];
            }
        }
        __syncthreads();  // every thread has read first[b] and count[b]
        if (threadIdx.x == 0) {
            A.first[b] = FF_UNCLAIMED;
            A.count[b] = 0u;
        }
    }
}

__global__ void k_loop_fill(uint32_t *p, size_t n, uint32_t v)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace ff

using namespace ff;

#include <vector>

struct ff_loop {
    ff_ctx *ctx = nullptr;
    uint32_t total = 0, mapN = 0, nMapped = 0;
    bool color = false;
    uint32_t *d_map = nullptr, *d_first = nullptr, *d_count = nullptr, *d_bsum = nullptr, *d_queue = nullptr;
    float *d_dist = nullptr, *d_wgt = nullptr, *d_depth = nullptr;
    uint16_t *d_col = nullptr;
    uint8_t *d_planes = nullptr;
    size_t bsumCap = 0, queueCap = 0, imgCap = 0, lcCap = 0;
    LoopState *d_st = nullptr, *h_st = nullptr;
    LoopCounters *d_lc = nullptr, *h_lc = nullptr;  // one per frame of a submission
    std::vector<hipEvent_t> ev;                      // two per timed frame
    bool timing = false;
    double ms = 0, bytes = 0;
    uint64_t calls = 0, units = 0;
    std::string err;
};

static bool loop_ok(ff_loop *lp, hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    lp->err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}

#define LCHK(call)                                          \
    do {                                                    \
        if (!loop_ok(lp, (call), #call)) return FF_E_DEVICE; \
    } while (0)

static void loop_free(ff_loop *lp)
{
    hipFree(lp->d_map); hipFree(lp->d_first); hipFree(lp->d_count); hipFree(lp->d_bsum);
    hipFree(lp->d_queue); hipFree(lp->d_dist); hipFree(lp->d_wgt); hipFree(lp->d_col);
    hipFree(lp->d_depth); hipFree(lp->d_planes); hipFree(lp->d_lc); hipFree(lp->d_st);
    if (lp->h_lc) hipHostFree(lp->h_lc);
    if (lp->h_st) hipHostFree(lp->h_st);
    for (hipEvent_t e : lp->ev)
        if (e) hipEventDestroy(e);
}

static int loop_fill(ff_loop *lp, uint32_t *p, size_t n, uint32_t v)
{
    if (!n) return FF_OK;
    k_loop_fill<<<(unsigned)((n + 255) / 256), 256, 0, lp->ctx->stream>>>(p, n, v);
    LCHK(hipGetLastError());
    return FF_OK;
}

// updateArraysForLoopClosure (geometryfusion_mipmap_cpu.cpp:3927-3931): no
// slot handed out, every brick without one, and (unlike the reference,
// whose memsets there cover only the first _nLeavesTotal values) every slot
// cleared, as after allocateMemoryCPU
static int loop_clear(ff_loop *lp)
{
    frame_join(lp->ctx);
    hipStream_t st = lp->ctx->stream;
    const size_t v = (size_t)lp->total * FF_BRICK;
    int rc;
    if ((rc = loop_fill(lp, lp->d_map, lp->mapN, FF_BRANCHINIT)) != FF_OK) return rc;
    if ((rc = loop_fill(lp, lp->d_first, lp->mapN, FF_UNCLAIMED)) != FF_OK) return rc;
    if ((rc = loop_fill(lp, lp->d_count, lp->mapN, 0u)) != FF_OK) return rc;
    if ((rc = loop_fill(lp, reinterpret_cast<uint32_t *>(lp->d_dist), v, 0u)) != FF_OK) return rc;
    if ((rc = loop_fill(lp, reinterpret_cast<uint32_t *>(lp->d_wgt), v, 0u)) != FF_OK) return rc;
    if (lp->d_col && (rc = loop_fill(lp, reinterpret_cast<uint32_t *>(lp->d_col), v * 3 / 2, 0u)) != FF_OK) return rc;
    if ((rc = loop_fill(lp, reinterpret_cast<uint32_t *>(lp->d_st), sizeof(LoopState) / 4, 0u)) != FF_OK) return rc;
    LCHK(hipStreamSynchronize(st));
    lp->nMapped = 0;
    return FF_OK;
}

static int frame_check(ff_loop *lp, const ff_loop_frame &f, int k)
{
    const char *why = nullptr;
    if (f.width <= 0 || f.height <= 0 || !f.depth || (f.n_queue && !f.queue))
        why = "invalid argument";
    else if (lp->color && (!f.red || !f.green || !f.blue))
        why = "the map fuses color: red, green and blue planes are required";
    else if (f.n_queue > 0x7fffffffu)
        why = "queue longer than 2^31 entries";
    if (!why) return FF_OK;
    lp->err = std::string("frame ") + std::to_string(k) + ": " + why;
    return FF_E_ARG;
}

// deleteOldLoopPoses' loop over nf frames (geometryfusion_mipmap_cpu.cpp:
// 3940-4022) as one submission on the map's stream, then one
// synchronisation; inputs in device memory
static int loop_submit(ff_loop *lp, const ff_loop_frame *fr, int nf, size_t *applied)
{
    frame_join(lp->ctx);  // the map the frames are subtracted from is complete
    ff_ctx *c = lp->ctx;
    for (int k = 0; k < nf; k++) {
        if (applied) applied[k] = 0;
        const int rc = frame_check(lp, fr[k], k);
        if (rc != FF_OK) return rc;
    }
    if (nf <= 0) return FF_OK;
    hipStream_t st = c->stream;
    if ((size_t)nf > lp->lcCap) {
        hipFree(lp->d_lc);
        if (lp->h_lc) hipHostFree(lp->h_lc);
        lp->d_lc = nullptr; lp->h_lc = nullptr; lp->lcCap = 0;
        LCHK(hipMalloc(&lp->d_lc, sizeof(LoopCounters) * nf));
        LCHK(hipHostMalloc(&lp->h_lc, sizeof(LoopCounters) * nf, hipHostMallocDefault));
        lp->lcCap = nf;
    }
    uint32_t nbMax = 0;
    for (int k = 0; k < nf; k++) {
        const size_t n = fr[k].n_queue;
        lp->h_lc[k] = LoopCounters{(uint32_t)n, 0u, 0u, 0u};
        if (n > FF_LOOP_SMALL) nbMax = std::max(nbMax, (uint32_t)((n + FF_LOOP_TPB - 1) / FF_LOOP_TPB));
    }
    if (nbMax > lp->bsumCap) {
        hipFree(lp->d_bsum);
        lp->d_bsum = nullptr; lp->bsumCap = 0;
        LCHK(hipMalloc(&lp->d_bsum, sizeof(uint32_t) * nbMax));
        lp->bsumCap = nbMax;
    }
    if (lp->timing && lp->ev.size() < 2 * (size_t)nf) {
        const size_t have = lp->ev.size();
        lp->ev.resize(2 * (size_t)nf, nullptr);
        for (size_t i = have; i < lp->ev.size(); i++) LCHK(hipEventCreate(&lp->ev[i]));
    }
    LCHK(hipMemcpyAsync(lp->d_lc, lp->h_lc, sizeof(LoopCounters) * nf, hipMemcpyHostToDevice, st));
    for (int k = 0; k < nf; k++) {
        const ff_loop_frame &f = fr[k];
        const uint32_t n = (uint32_t)f.n_queue;
        if (!n) continue;
        const CamF q = camera_inverse(c, f.R, f.t, f.K);
        SdfParams P{};
        for (int r = 0; r < 3; r++) {
            P.m[4 * r] = q.r[3 * r];
            P.m[4 * r + 1] = q.r[3 * r + 1];
            P.m[4 * r + 2] = q.r[3 * r + 2];
            P.m[4 * r + 3] = q.t[r];
        }
        P.fx = q.fx; P.fy = q.fy; P.cx = q.cx; P.cy = q.cy;
        P.scale = c->scale;
        P.dT = c->dT;
        P.useColor = lp->color ? 1 : 0;
        LoopArgs A{};
        A.leafinfo = c->d_leafinfo;
        A.map = lp->d_map; A.first = lp->d_first; A.count = lp->d_count;
        A.dist = lp->d_dist; A.wgt = lp->d_wgt; A.col = lp->d_col;
        A.depth = f.depth; A.red = f.red; A.green = f.green; A.blue = f.blue;
        A.W = f.width; A.H = f.height;
        // (the used count passes the capacity when the pool ran out; the
        // slot map and the per-brick arrays have mapN entries)
        A.nLeavesUsed = std::min(c->nLeavesUsed, lp->mapN);
        A.total = lp->total;
        LoopCounters *lc = lp->d_lc + k;
        if (n <= FF_LOOP_SMALL) {
            k_loop_map1<<<1, FF_LOOP_TPB, 0, st>>>(f.queue, n, A, lp->d_st, lc);
        } else {
            const uint32_t nb = (n + FF_LOOP_TPB - 1) / FF_LOOP_TPB;
            k_loop_claim<<<(n + 255) / 256, 256, 0, st>>>(f.queue, n, A, lc);
            k_loop_flags<<<nb, FF_LOOP_TPB, 0, st>>>(f.queue, n, A, lp->d_st, lp->d_bsum);
            k_loop_assign<<<nb, FF_LOOP_TPB, 0, st>>>(f.queue, n, A, lp->d_st, lc, lp->d_bsum, nb);
        }
        // timed frames (one in FF_LOOP_TIME_EVERY): the dispatch itself
        // records the events, so they time the kernel alone
        const uint32_t grid = std::min<uint32_t>(n, FF_LOOP_GRID);
        if (lp->timing && k % FF_LOOP_TIME_EVERY == 0)
            hipExtLaunchKernelGGL(k_loop_delete, dim3(grid), dim3(512), 0, st, lp->ev[2 * k], lp->ev[2 * k + 1], 0,
                                  (const uint32_t *)f.queue, (uint32_t)n, A, P, lp->d_st, (const LoopCounters *)lc);
        else
            k_loop_delete<<<grid, 512, 0, st>>>(f.queue, n, A, P, lp->d_st, lc);
        LCHK(hipGetLastError());
    }
    LCHK(hipMemcpyAsync(lp->h_lc, lp->d_lc, sizeof(LoopCounters) * nf, hipMemcpyDeviceToHost, st));
    LCHK(hipMemcpyAsync(lp->h_st, lp->d_st, sizeof(LoopState), hipMemcpyDeviceToHost, st));
    LCHK(hipStreamSynchronize(st));
    lp->nMapped = lp->h_st->nMapped;
    int bad = -1, full = -1;
    for (int k = 0; k < nf; k++) {
        const LoopCounters &h = lp->h_lc[k];
        const size_t n = fr[k].n_queue;
        if (h.bad) {
            if (bad < 0) bad = k;
            continue;
        }
        if (applied) applied[k] = h.failpos;
        if (h.failpos < n && full < 0) full = k;
        if (lp->timing && n && k % FF_LOOP_TIME_EVERY == 0) {
            float ms = 0;
            LCHK(hipEventElapsedTime(&ms, lp->ev[2 * k], lp->ev[2 * k + 1]));
            lp->ms += ms;
            lp->calls++;
            lp->units += h.failpos;
            lp->bytes += (double)h.failpos * (lp->color ? 14336.0 : 8192.0) +
                         (double)fr[k].width * fr[k].height * (lp->color ? 7.0 : 4.0);
        }
    }
    if (bad >= 0) {
        lp->err = "frame " + std::to_string(bad) + ": a queue entry names a brick beyond the map's used bricks";
        return FF_E_ARG;
    }
    if (full >= 0) {
        lp->err = "the loop-closure map's " + std::to_string(lp->total) + " slots are taken: frame " +
                  std::to_string(full) + " applied the first " + std::to_string(lp->h_lc[full].failpos) + " of " +
                  std::to_string(fr[full].n_queue) + " queue entries";
        return FF_E_POOL_FULL;
    }
    return FF_OK;
}

static ff_loop_frame one_frame(const float *depth, const uint8_t *red, const uint8_t *green, const uint8_t *blue,
                               int W, int H, const double R[9], const double t[3], const double K[4],
                               const uint32_t *queue, size_t n)
{
    ff_loop_frame f{};
    f.depth = depth; f.red = red; f.green = green; f.blue = blue;
    f.width = W; f.height = H;
    if (R) std::copy(R, R + 9, f.R);
    if (t) std::copy(t, t + 3, f.t);
    if (K) std::copy(K, K + 4, f.K);
    f.queue = queue;
    f.n_queue = n;
    return f;
}

extern "C" {

ff_loop *ff_loop_create(ff_ctx *ctx, uint32_t max_leaves)
{
    if (!ctx) return nullptr;
    ff_loop *lp = new ff_loop();
    lp->ctx = ctx;
    lp->total = max_leaves ? max_leaves : 100000u;  // MAX_LOOP_CLOSURE_LEAVES, definitions.h:124
    lp->mapN = ctx->nLeavesTotal;
    lp->color = ctx->useColor;
    const size_t v = (size_t)lp->total * FF_BRICK;
    bool ok = hipSetDevice(ctx->device) == hipSuccess &&
              hipMalloc(&lp->d_map, sizeof(uint32_t) * lp->mapN) == hipSuccess &&
              hipMalloc(&lp->d_first, sizeof(uint32_t) * lp->mapN) == hipSuccess &&
              hipMalloc(&lp->d_count, sizeof(uint32_t) * lp->mapN) == hipSuccess &&
              hipMalloc(&lp->d_dist, sizeof(float) * v) == hipSuccess &&
              hipMalloc(&lp->d_wgt, sizeof(float) * v) == hipSuccess &&
              (!lp->color || hipMalloc(&lp->d_col, sizeof(uint16_t) * v * 3) == hipSuccess) &&
              hipMalloc(&lp->d_st, sizeof(LoopState)) == hipSuccess &&
              hipHostMalloc(&lp->h_st, sizeof(LoopState), hipHostMallocDefault) == hipSuccess &&
              hipMemcpyToSymbol(HIP_SYMBOL(c_rcp_loop), FF_RCP_TABLE, sizeof(FF_RCP_TABLE), 0,
                                hipMemcpyHostToDevice) == hipSuccess;
    if (!ok || loop_clear(lp) != FF_OK) {
        ctx->err = "ff_loop_create: device allocation failed";
        loop_free(lp);
        delete lp;
        return nullptr;
    }
    return lp;
}

void ff_loop_destroy(ff_loop *lp)
{
    if (!lp) return;
    frame_join(lp->ctx);
    hipStreamSynchronize(lp->ctx->stream);
    loop_free(lp);
    delete lp;
}

const char *ff_loop_last_error(const ff_loop *lp) { return lp ? lp->err.c_str() : "null loop-closure map"; }

int ff_loop_reset(ff_loop *lp)
{
    if (!lp) return FF_E_ARG;
    return loop_clear(lp);
}

int ff_loop_delete_frame(ff_loop *lp, const float *depth, const uint8_t *red, const uint8_t *green,
                         const uint8_t *blue, int width, int height, const double R[9], const double t[3],
                         const double K[4], const uint32_t *queue, size_t n_queue, size_t *applied)
{
    if (!lp) return FF_E_ARG;
    if (applied) *applied = 0;
    ff_loop_frame f = one_frame(depth, red, green, blue, width, height, R, t, K, queue, n_queue);
    int rc = frame_check(lp, f, 0);
    if (rc != FF_OK) return rc;
    if (!R || !t || !K) {
        lp->err = "frame 0: invalid argument";
        return FF_E_ARG;
    }
    hipStream_t st = lp->ctx->stream;
    const size_t npix = (size_t)width * height;
    if (npix > lp->imgCap) {
        hipFree(lp->d_depth); hipFree(lp->d_planes);
        lp->d_depth = nullptr; lp->d_planes = nullptr; lp->imgCap = 0;
        LCHK(hipMalloc(&lp->d_depth, sizeof(float) * npix));
        LCHK(hipMalloc(&lp->d_planes, 3 * npix));
        lp->imgCap = npix;
    }
    if (n_queue > lp->queueCap) {
        hipFree(lp->d_queue);
        lp->d_queue = nullptr; lp->queueCap = 0;
        LCHK(hipMalloc(&lp->d_queue, sizeof(uint32_t) * n_queue));
        lp->queueCap = n_queue;
    }
    LCHK(hipMemcpyAsync(lp->d_depth, depth, sizeof(float) * npix, hipMemcpyHostToDevice, st));
    const uint8_t *pl[3] = {red, green, blue};
    const uint8_t *dp[3] = {nullptr, nullptr, nullptr};
    for (int k = 0; k < 3; k++)
        if (pl[k] && lp->color) {
            dp[k] = lp->d_planes + k * npix;
            LCHK(hipMemcpyAsync(lp->d_planes + k * npix, pl[k], npix, hipMemcpyHostToDevice, st));
        }
    if (n_queue) LCHK(hipMemcpyAsync(lp->d_queue, queue, sizeof(uint32_t) * n_queue, hipMemcpyHostToDevice, st));
    f = one_frame(lp->d_depth, dp[0], dp[1], dp[2], width, height, R, t, K, lp->d_queue, n_queue);
    return loop_submit(lp, &f, 1, applied);
}

int ff_loop_delete_frame_device(ff_loop *lp, const float *depth_dev, const uint8_t *red_dev,
                                const uint8_t *green_dev, const uint8_t *blue_dev, int width, int height,
                                const double R[9], const double t[3], const double K[4],
                                const uint32_t *queue_dev, size_t n_queue, size_t *applied)
{
    if (!lp) return FF_E_ARG;
    if (!R || !t || !K) {
        lp->err = "frame 0: invalid argument";
        return FF_E_ARG;
    }
    const ff_loop_frame f = one_frame(depth_dev, red_dev, green_dev, blue_dev, width, height, R, t, K, queue_dev,
                                      n_queue);
    return loop_submit(lp, &f, 1, applied);
}

int ff_loop_delete_frames(ff_loop *lp, const ff_loop_frame *frames, int n_frames, size_t *applied)
{
    if (!lp) return FF_E_ARG;
    if (n_frames < 0 || (n_frames && !frames)) {
        lp->err = "invalid argument";
        return FF_E_ARG;
    }
    return loop_submit(lp, frames, n_frames, applied);
}

int ff_loop_mapped(ff_loop *lp, uint32_t *n_mapped, uint32_t *capacity)
{
    if (!lp) return FF_E_ARG;
    if (n_mapped) *n_mapped = lp->nMapped;
    if (capacity) *capacity = lp->total;
    return FF_OK;
}

int ff_loop_copy_map(ff_loop *lp, uint32_t *out, size_t n)
{
    if (!lp || !out || n > lp->mapN) return FF_E_ARG;
    LCHK(hipStreamSynchronize(lp->ctx->stream));
    LCHK(hipMemcpy(out, lp->d_map, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    return FF_OK;
}

int ff_loop_copy_bricks(ff_loop *lp, const uint32_t *slots, size_t n, float *dist, float *weight, uint16_t *color)
{
    if (!lp || (!slots && n)) return FF_E_ARG;
    LCHK(hipStreamSynchronize(lp->ctx->stream));
    for (size_t i = 0; i < n; i++) {
        const size_t s = slots[i];
        if (s >= lp->total) {
            lp->err = "slot beyond the loop-closure map's capacity";
            return FF_E_ARG;
        }
        if (dist) LCHK(hipMemcpy(dist + i * FF_BRICK, lp->d_dist + s * FF_BRICK, FF_BRICK * 4, hipMemcpyDeviceToHost));
        if (weight) LCHK(hipMemcpy(weight + i * FF_BRICK, lp->d_wgt + s * FF_BRICK, FF_BRICK * 4, hipMemcpyDeviceToHost));
        if (color && lp->d_col)
            LCHK(hipMemcpy(color + i * FF_BRICK * 3, lp->d_col + s * FF_BRICK * 3, FF_BRICK * 6, hipMemcpyDeviceToHost));
    }
    return FF_OK;
}

int ff_loop_set_timing(ff_loop *lp, int on)
{
    if (!lp) return FF_E_ARG;
    lp->timing = on != 0;
    return FF_OK;
}

int ff_loop_get_timing(ff_loop *lp, double *ms, uint64_t *calls, double *bytes, uint64_t *units)
{
    if (!lp) return FF_E_ARG;
    if (ms) *ms = lp->ms;
    if (calls) *calls = lp->calls;
    if (bytes) *bytes = lp->bytes;
    if (units) *units = lp->units;
    return FF_OK;
}

}  // extern "C"
