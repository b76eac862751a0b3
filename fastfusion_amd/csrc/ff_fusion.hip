// ff_fusion.hip -- per-frame depth integration on gfx950.
//
// One frame (the reference's FusionMipMapCPU::addMap, u16 overload,
// src/fusion/geometryfusion_mipmap_cpu.cpp:1603-2034) runs as:
//   k_planes     per pixel: brick key of the back-projected point; three bit
//                planes per row (valid, same key as left, same key as up)
//   k_dedup      the reference's order-dependent de-dup recurrence
//                (traversal_wrapper_functions.cpp:601-690), solved exactly by
//                one wavefront sweeping rows with a lane skew, 32 columns per
//                lane, bit-parallel within a lane
//   k_descend    per processed pixel: octree descent through the existing
//                tree, queueing the bricks on the path, recording every
//                missing node with the smallest pixel index that reaches it
//   k_resolve / scan / k_assign
//                new branches and bricks get indices in the order the
//                sequential reference creates them (pixel order, top-down)
//   k_link       writes the new tree links and bricks, with the parent brick
//                the reference sees at the creating pixel's turn
//   k_sdf        the per-voxel SDF update (update_functions.cpp:1187-1444)
//                with the reference's round-toward-zero + RCPPS arithmetic
#include "ff_device.hpp"
#include "ff_internal.hpp"
#include "ff_sdf.hpp"
#include "rcp_table.h"

#include <hip/hip_ext.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

using namespace ff;

namespace ff {

__constant__ uint32_t c_rcp_table[2048];

// ---------------------------------------------------------------------------
// hash table of nodes created this frame
// key: [type:1][depth:5][x:16][y:16][z:16]
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t node_key(uint32_t type, uint32_t depth, uint32_t x, uint32_t y,
                                             uint32_t z)
{
    return ((uint64_t)type << 53) | ((uint64_t)depth << 48) | ((uint64_t)x << 32) |
           ((uint64_t)y << 16) | (uint64_t)z;
}

__device__ __forceinline__ uint32_t key_hash(uint64_t k, uint32_t mask)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return (uint32_t)k & mask;
}

__device__ void hash_insert_min(const HashView &h, uint64_t key, uint32_t pix, uint32_t aux,
                                FrameCounters *fc)
{
    uint32_t s = key_hash(key, h.mask);
    for (uint32_t probe = 0; probe <= h.mask; probe++) {
        unsigned long long prev = atomicCAS((unsigned long long *)&h.keys[s], (unsigned long long)FF_HEMPTY,
                                            (unsigned long long)key);
        if (prev == FF_HEMPTY) {
            uint32_t li = atomicAdd(&fc->listCount, 1u);
            if (li < h.listCap) h.list[li] = s;
            h.aux[s] = aux;
            atomicMin(&h.minpix[s], pix);
            return;
        }
        if (prev == key) {
            atomicMin(&h.minpix[s], pix);
            return;
        }
        s = (s + 1) & h.mask;
    }
    atomicOr(&fc->flags, FF_FLAG_HASH_FULL);
}

__device__ __forceinline__ int hash_find(const HashView &h, uint64_t key)
{
    uint32_t s = key_hash(key, h.mask);
    for (uint32_t probe = 0; probe <= h.mask; probe++) {
        uint64_t k = h.keys[s];
        if (k == key) return (int)s;
        if (k == FF_HEMPTY) return -1;
        s = (s + 1) & h.mask;
    }
    return -1;
}

// ---------------------------------------------------------------------------
// first frame: integer bounding box (computeBoundingboxIntCPU, u16 overload,
// geometryfusion_mipmap_cpu.cpp:1270-1303)
// ---------------------------------------------------------------------------
__global__ void k_bbox_first(const uint16_t *__restrict__ depth, int W, int H, float scaling, float maxcam,
                             CamF p, int bw, FrameCounters *fc)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    int mn[3] = {1000000, 1000000, 1000000}, mx[3] = {-1000000, -1000000, -1000000};
    if (x < W) {
        const float pz = __fmul_rn((float)depth[y * W + x], scaling);
        if (pz > 0.0f && pz < maxcam) {
            const float px = __fmul_rn(__fdiv_rn(__fsub_rn((float)x, p.cx), p.fx), pz);
            const float py = __fmul_rn(__fdiv_rn(__fsub_rn((float)y, p.cy), p.fy), pz);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float s = __fadd_rn(__fadd_rn(__fadd_rn(__fmul_rn(p.r[3 * k], px), __fmul_rn(p.r[3 * k + 1], py)),
                                                    __fmul_rn(p.r[3 * k + 2], pz)),
                                          p.t[k]);
                const int v = cvt_x86(s);
                mn[k] = v - bw;
                mx[k] = v + bw;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        int a = mn[k], b = mx[k];
        for (int o = 32; o > 0; o >>= 1) {
            a = min(a, __shfl_xor(a, o));
            b = max(b, __shfl_xor(b, o));
        }
        if ((threadIdx.x & 63) == 0) {
            atomicMin(&fc->bmin[k], a);
            atomicMax(&fc->bmax[k], b);
        }
    }
}

// v_k = (int)((((qx_k[x] + qy_k[y]) * D) * scaling) + t_k), RN, no contraction
// (traversal_wrapper_functions.cpp:613-630), with the reference's per-frame
// ray tables qx_k = R_k0 * pxp[x], qy_k = R_k1 * pyp[y] + R_k2
// (geometryfusion_mipmap_cpu.cpp:1765-1790) formed in place: the same float
// operations, so the same values
// (px = pxp[x], py = pyp[y])
__device__ __forceinline__ void voxel_of(const CamF &p, float px, float py, float D, float scaling, int v[3])
{
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float qx = __fmul_rn(p.r[3 * k], px);                            // R * pxp
        const float qy = __fadd_rn(__fmul_rn(p.r[3 * k + 1], py), p.r[3 * k + 2]);  // R * pyp + R_z
        v[k] = cvt_x86(__fadd_rn(__fmul_rn(__fmul_rn(__fadd_rn(qx, qy), D), scaling), p.t[k]));
    }
}

// the frame counters' reset (bounds start as the whole volume)
__device__ __forceinline__ void frame_counters_reset(FrameCounters *fc, uint32_t n)
{
    FrameCounters z;
    z.nQueued = z.listCount = z.newB = z.newL = 0u;
    for (int k = 0; k < 3; k++) { z.bmin[k] = 0; z.bmax[k] = (int)n - 1; }
    z.flags = z.nProcessed = z.sdfRan = 0u;
    *fc = z;
}

__global__ void k_frame_begin(const FrameDev *__restrict__ fd, FrameCounters *fc)
{
    if (threadIdx.x == 0) frame_counters_reset(fc, fd->n);
}

// ---------------------------------------------------------------------------
// k_planes: per 64-pixel row chunk, ballot the three relations into bit
// planes: one uint4 per (row y, word w) at y * words + w holding
// {valid, key == left key, key == up key, 0}
// ---------------------------------------------------------------------------
// (block (0, 0) also resets the frame counters: no kernel before the
// descent reads them)
__device__ __forceinline__ void planes_body(const FrameDev &fd, int W, int H, int words,
                                            const float *__restrict__ pxp, const float *__restrict__ pyp,
                                            uint32_t *__restrict__ planes, FrameCounters *fc)
{
    const int lane = threadIdx.x, y = blockIdx.y, x = blockIdx.x * 64 + lane;
    const uint16_t *__restrict__ depth = fd.depth;
    const float scaling = fd.scaling, maxcam = fd.maxcam;
    const CamF &p = fd.p;
    if (blockIdx.x == 0 && y == 0 && lane == 0) frame_counters_reset(fc, fd.n);
    bool valid = false, el = false, eu = false;
    if (x < W) {
        const float D = (float)depth[y * W + x];
        const float z = __fmul_rn(D, scaling);
        valid = z > 0.0f && z < maxcam;
        const float px = pxp[x], py = pyp[y];
        int v[3];
        voxel_of(p, px, py, D, scaling, v);
        if (x > 0) {
            int u[3];
            voxel_of(p, pxp[x - 1], py, (float)depth[y * W + x - 1], scaling, u);
            el = ((u[0] & -8) == (v[0] & -8)) && ((u[1] & -8) == (v[1] & -8)) && ((u[2] & -8) == (v[2] & -8));
        }
        if (y > 0) {
            int u[3];
            voxel_of(p, px, pyp[y - 1], (float)depth[(y - 1) * W + x], scaling, u);
            eu = ((u[0] & -8) == (v[0] & -8)) && ((u[1] & -8) == (v[1] & -8)) && ((u[2] & -8) == (v[2] & -8));
        }
    }
    const uint64_t bv = __ballot(valid), bl = __ballot(el), bu = __ballot(eu);
    if (lane < 2) {
        const int w = blockIdx.x * 2 + lane;
        if (w < words) {
            const int sh = lane * 32;
            reinterpret_cast<uint4 *>(planes)[(size_t)y * words + w] =
                make_uint4((uint32_t)(bv >> sh), (uint32_t)(bl >> sh), (uint32_t)(bu >> sh), 0u);
        }
    }
}

__global__ __launch_bounds__(64) void k_planes(const FrameDev *__restrict__ fd, int W, int H, int words,
                                               const float *__restrict__ pxp, const float *__restrict__ pyp,
                                               uint32_t *__restrict__ planes, FrameCounters *fc)
{
    planes_body(*fd, W, H, words, pxp, pyp, planes, fc);
}

// The same with the frame's parameters as a kernel argument: the first
// thread also stores them where the frame's later kernels read them, in
// place of a separate upload.
__global__ __launch_bounds__(64) void k_planes_v(const FrameDev f, FrameDev *fdOut, int W, int H, int words,
                                                 const float *__restrict__ pxp, const float *__restrict__ pyp,
                                                 uint32_t *__restrict__ planes, FrameCounters *fc)
{
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *fdOut = f;
    planes_body(f, W, H, words, pxp, pyp, planes, fc);
}

// the batched forms of the allocation kernels: stream blockIdx.z of S, grid
// x (and y) sized for the largest stream; blocks outside a stream's range
// return at once
__global__ __launch_bounds__(64) void k_planes_b(const AllocStream *__restrict__ S)
{
    const AllocStream &a = S[blockIdx.z];
    if ((int)blockIdx.y >= a.H || (int)blockIdx.x * 64 >= a.W) return;
    planes_body(*a.fd, a.W, a.H, a.words, a.pxp, a.pyp, a.planes, a.fc);
}

// ---------------------------------------------------------------------------
// k_dedup: P(x,y) = valid & !(Eu & P(x,y-1)) & !(El & P(x-1,y)).
// Wave 0 of the block solves the recurrence: lane l owns columns
// [32 l, 32 l + 32) and handles row s - l at step s, so the left neighbour's
// result for the same row arrives from lane l-1's previous step through a
// one-lane DPP shift; rows of a lane are carried in a register.  Within a
// lane the horizontal chain is solved bit-parallel: maximal runs linked by
// El alternate 1,0,1,... from their start, so P is the run bits whose
// distance from the run's start is even.  One wave does every step, so a
// step costs its instruction count: the other waves stream the bit planes
// into an LDS ring laid out by diagonal (row + word), where step s of every
// lane is ring row s at the lane's column, read with constant offsets; rows
// outside the image are written as zeros, so no step tests bounds.  Four
// bands of R diagonals, written three bands ahead of the sweep, one block
// barrier per band.  The sweep's results go to a second LDS ring (128 steps
// of 64 lanes, one conflict-free store per step); after each band the other
// waves write the rows it completed to pbits, a row's words together.
// Storing a step's 64 words straight to pbits (64 rows, 64 lines) cost the
// computing wave about twice its instruction time.
// ---------------------------------------------------------------------------
#define FF_DD_NB 4
#define FF_DD_R 32
#define FF_DD_THREADS 1024
#define FF_DD_PSTEPS 128  // the result ring's steps (covers the rows a band completes and the next band)

// ring columns: 20, 40 or 64 words (images up to 640, 1280, 2048 pixels wide)
__host__ __device__ static inline int dedup_wmax(int words) { return words <= 20 ? 20 : (words <= 40 ? 40 : 64); }
// the plane ring, then the result ring
size_t dedup_lds_bytes(int words)
{
    return (size_t)FF_DD_NB * FF_DD_R * dedup_wmax(words) * sizeof(uint4) + FF_DD_PSTEPS * 64 * sizeof(uint32_t);
}
int dedup_plane_rows(int H) { return H; }

// the ring entries one loader thread moves per band
template <int WMAX>
struct DedupLoad {
    static constexpr int PER_BAND = FF_DD_R * WMAX;
    static constexpr int PT = (PER_BAND + FF_DD_THREADS - 65) / (FF_DD_THREADS - 64);
    uint4 v[PT];
    uint32_t in;  // bit j: entry j lies in the image (the others put zeros)
};

// diagonal band `band` of the planes: entry e = kk * WMAX + w holds the
// (valid, El, Eu) words of row band * R + kk - w, column word w, or zeros
template <int WMAX>
__device__ __forceinline__ void dedup_fetch(DedupLoad<WMAX> &L, const uint4 *__restrict__ planes, int H, int words,
                                            int band, int lt)
{
    // every loader issues its PT loads unconditionally (an entry outside the
    // image reads row 0, word 0 and is zeroed when it is put), so that the
    // put of the older register set waits for its own loads alone: loads
    // behind a branch make the wait before it cover every load in flight
    L.in = 0u;
#pragma unroll
    for (int j = 0; j < DedupLoad<WMAX>::PT; j++) {
        const int e = lt + j * (FF_DD_THREADS - 64);
        const int kk = e / WMAX, w = e - kk * WMAX;
        const int row = band * FF_DD_R + kk - w;
        const bool in = e < DedupLoad<WMAX>::PER_BAND && w < words && row >= 0 && row < H;
        L.v[j] = planes[in ? (size_t)row * words + w : 0];
        L.in |= (uint32_t)in << j;
    }
}

template <int WMAX>
__device__ __forceinline__ void dedup_put(const DedupLoad<WMAX> &L, uint4 *ring, int band, int lt)
{
    constexpr int RING = FF_DD_NB * FF_DD_R;
#pragma unroll
    for (int j = 0; j < DedupLoad<WMAX>::PT; j++) {
        const int e = lt + j * (FF_DD_THREADS - 64);
        if (e < DedupLoad<WMAX>::PER_BAND) {
            const int kk = e / WMAX, w = e - kk * WMAX;
            const uint32_t m = 0u - ((L.in >> j) & 1u);  // all ones for an entry in the image
            const uint4 v = L.v[j];
            ring[((band * FF_DD_R + kk) & (RING - 1)) * WMAX + w] = make_uint4(v.x & m, v.y & m, v.z & m, v.w & m);
        }
    }
}

// the rows band `it` completed (their last word, at step row + words - 1,
// in the band) from the result ring to pbits; rows outside the image are
// skipped.  A row's first word is at most 63 steps older (two bands back);
// while this runs the computing wave fills band it + 1, whose ring slots
// held band it - 3, and it reaches band it + 2 (band it - 2's slots) only
// after the next barrier, which waits for these LDS reads.
template <int WMAX>
__device__ __forceinline__ void dedup_writeout(const uint32_t *pring, uint32_t *__restrict__ pbits, int H, int words,
                                               int it, int lt)
{
    const int r0 = it * FF_DD_R - words + 1;
    for (int e = lt; e < FF_DD_R * WMAX; e += FF_DD_THREADS - 64) {
        const int j = e / WMAX, w = e - j * WMAX;
        const int r = r0 + j;
        if (w < words && r >= 0 && r < H)
            pbits[(size_t)(FF_PB_PAD + r) * FF_PB_STRIDE + w] = pring[((r + w) & (FF_DD_PSTEPS - 1)) * 64 + w];
    }
}

template <int WMAX>
__device__ __forceinline__ void dedup_body(uint4 *ring, const uint32_t *__restrict__ planes32, int H, int words,
                                           uint32_t *__restrict__ pbits, FrameCounters *fc)
{
    // (the result ring follows the plane ring)
    uint32_t *pring = reinterpret_cast<uint32_t *>(ring + FF_DD_NB * FF_DD_R * WMAX);
    constexpr int R = FF_DD_R, RING = FF_DD_NB * R;
    constexpr uint32_t EVEN = 0x55555555u;
    const uint4 *planes = reinterpret_cast<const uint4 *>(planes32);
    const int tid = threadIdx.x;
    const bool compute = __builtin_amdgcn_readfirstlane(tid >> 6) == 0;  // wave 0 (a wave-uniform branch)
    const int nsteps = H + words - 1;
    const int niter = (nsteps + R - 1) / R;
    const int lt = tid - 64;
    // the loaders keep two bands in registers: band b is fetched two
    // iterations before it goes into the ring, so a fetch has two
    // iterations of the sweep to arrive
    DedupLoad<WMAX> pre[2];
    if (!compute) {
        for (int b = 0; b < 3; b++) {
            dedup_fetch<WMAX>(pre[0], planes, H, words, b, lt);
            dedup_put<WMAX>(pre[0], ring, b, lt);
        }
        dedup_fetch<WMAX>(pre[1], planes, H, words, 3, lt);
        dedup_fetch<WMAX>(pre[0], planes, H, words, 4, lt);
    }
    __syncthreads();  // the first three bands are in the ring
    const int l = tid & 63;
    const int lw = l < WMAX ? l : WMAX - 1;  // lanes past the ring's columns read a column they ignore
    uint32_t *pw = pring + l;  // the lane's column of the result ring
    uint32_t pprev = 0u, pin = 0u, nproc = 0u;
    // The computing wave and the loaders run separate loops with one barrier
    // per iteration each, niter rounded up to even in both (the computing
    // wave skips the extra one), so that the loaders' loop has no
    // conditional half.  Kept apart and unconditional, each loop's memory
    // waits count its own operations: the loaders wait for the older band's
    // loads only, and a fetch has two iterations to arrive.  (In one shared
    // loop the waits also covered the computing wave's stores and came out
    // as waits for every load in flight: a memory round trip per iteration.)
    const int niter2 = (niter + 1) & ~1;
    if (compute) {
        for (int it = 0; it < niter2; it++) {
            if (it == niter) {
                __syncthreads();
                break;
            }
            // R diagonals
            const uint4 *q = ring + ((it * R) & (RING - 1)) * WMAX + lw;
            uint4 rq[R];
#pragma unroll
            for (int k = 0; k < R; k++) rq[k] = q[k * WMAX];
            uint32_t *o = pw + ((it * R) & (FF_DD_PSTEPS - 1)) * 64;
#pragma unroll
            for (int k = 0; k < R; k++) {
                const uint32_t va = rq[k].x, el = rq[k].y, eu = rq[k].z;
                const uint32_t a = va & ~(eu & pprev);
                // link j <- j-1 inside the word (a at j-1) and bit 0 <- the
                // left lane's bit 31 (linked only if that pixel was processed)
                const uint32_t a1 = __builtin_amdgcn_alignbit(a, pin, 31);
                const uint32_t cont = a & el & a1;
                const uint32_t ev = a & ~cont & EVEN;                  // runs starting at even bits
                const uint32_t t = (ev << 1) + cont;                   // carries run through each such run
                const uint32_t E = (~t & cont) | ev;                   // the bits of runs with an even start
                const uint32_t P = a & ~(E ^ EVEN);                    // even distance from the run's start
                o[k * 64] = P;
                nproc += (uint32_t)__popc(P);
                pprev = P;
                pin = (uint32_t)__builtin_amdgcn_mov_dpp((int)P, 0x138, 0xf, 0xf, true);  // wave_shr:1, lane 0 gets 0
            }
            __syncthreads();
        }
    } else {
        // band it + 3 into the ring (into the slot band it - 1 was read
        // from), then band it + 5 fetched into the registers it came from;
        // unrolled by two so that each register set keeps its registers
        // (band it + 3 is in pre[1] for even it, in pre[0] for odd it)
        for (int it = 0; it < niter2; it += 2) {
            dedup_put<WMAX>(pre[1], ring, it + 3, lt);
            dedup_fetch<WMAX>(pre[1], planes, H, words, it + 5, lt);
            __syncthreads();
            dedup_writeout<WMAX>(pring, pbits, H, words, it, lt);
            dedup_put<WMAX>(pre[0], ring, it + 4, lt);
            dedup_fetch<WMAX>(pre[0], planes, H, words, it + 6, lt);
            __syncthreads();
            dedup_writeout<WMAX>(pring, pbits, H, words, it + 1, lt);
        }
    }
    if (compute) {
        if (l >= words) nproc = 0u;
        for (int o = 32; o > 0; o >>= 1) nproc += __shfl_xor(nproc, o);
        if (l == 0) fc->nProcessed = nproc;
    }
}

template <int WMAX>
__global__ __launch_bounds__(FF_DD_THREADS) void k_dedup(const uint32_t *__restrict__ planes, int H, int words,
                                                         uint32_t *__restrict__ pbits, FrameCounters *fc)
{
    extern __shared__ uint4 ring[];
    dedup_body<WMAX>(ring, planes, H, words, pbits, fc);
}

// one block per stream whose ring width is WMAX
template <int WMAX>
__global__ __launch_bounds__(FF_DD_THREADS) void k_dedup_b(const AllocStream *__restrict__ S)
{
    extern __shared__ uint4 ring[];
    const AllocStream &a = S[blockIdx.z];
    if (dedup_wmax(a.words) != WMAX) return;
    dedup_body<WMAX>(ring, a.planes, a.H, a.words, a.pbits, a.fc);
}

// ---------------------------------------------------------------------------
// k_descend: queryPointDepthSingle_func_subtree (traversal_functions.cpp:308-395)
// read-only against the tree as it was before this frame.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void descend_body(const FrameDev *__restrict__ fd, int W, int H, int words,
                                             const float *__restrict__ pxp, const float *__restrict__ pyp,
                                             const uint32_t *__restrict__ pbits, const MapView &m, const HashView &hv,
                                             uint32_t *__restrict__ pl0, FrameCounters *fc)
{
    const uint16_t *__restrict__ depth = fd->depth;
    const float scaling = fd->scaling;
    const uint32_t n = fd->n, stamp = fd->stamp;
    const int bw = (int)fd->bandwidth;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    const int lane = threadIdx.x & 63;
    const bool proc = x < W && ((pbits[(size_t)(y + FF_PB_PAD) * FF_PB_STRIDE + (x >> 5)] >> (x & 31)) & 1u);
    const uint64_t pb = __ballot(proc);
    if (pb == 0) return;  // wave-uniform
    const uint32_t pix = (uint32_t)(y * W + x);
    int v[3] = {0, 0, 0};
    float D = 0.0f;
    if (proc) {
        D = (float)depth[pix];
        voxel_of(fd->p, pxp[x], pyp[y], D, scaling, v);
    }
    // bounding-box tracking (traversal_wrapper_functions.cpp:642-676), one
    // atomic per wave and axis
#pragma unroll
    for (int k = 0; k < 3; k++) {
        int lo = proc ? v[k] - bw : 0, hi = proc ? v[k] + bw : 0;
        for (int o = 32; o > 0; o >>= 1) {
            lo = min(lo, __shfl_xor(lo, o));
            hi = max(hi, __shfl_xor(hi, o));
        }
        if (lane == 0) {
            if (lo < 0) atomicMin(&fc->bmin[k], lo);
            if (hi >= (int)n) atomicMax(&fc->bmax[k], hi);
        }
    }
    const bool inb = proc && v[0] >= 0 && (uint32_t)(uint16_t)v[0] < n && v[1] >= 0 &&
                     (uint32_t)(uint16_t)v[1] < n && v[2] >= 0 && (uint32_t)(uint16_t)v[2] < n;
    const uint32_t X = (uint16_t)v[0], Y = (uint16_t)v[1], Z = (uint16_t)v[2];
    const uint32_t target = inb ? leaf_scale_target(__fmul_rn(D, scaling)) : 0u;
    // target node of this pixel; pixels of a wave with the same target node
    // walk the same path, so only the lowest lane (smallest pixel index) does
    uint32_t tsize = n >> 1, tdepth = 1;
    while ((tsize >> 1) >= target && tsize > 1) {
        tsize >>= 1;
        tdepth++;
    }
    const uint64_t tkey = inb ? node_key(1, tdepth, X & ~(tsize - 1u), Y & ~(tsize - 1u), Z & ~(tsize - 1u)) : 0ull;
    bool leader = false;
    uint64_t todo = __ballot(inb);
    while (todo) {
        const int l0 = __ffsll((long long)todo) - 1;
        const uint64_t k0 = __shfl(tkey, l0);
        const uint64_t same = __ballot(inb && tkey == k0);
        if (lane == l0) leader = true;
        todo &= ~same;
    }
    // Bricks on the path take the frame's stamp (plain stores: no result is
    // awaited, and the walk stays a chain of tree loads only); k_queue then
    // collects the stamped bricks.  Queueing here with atomics cost more than
    // the walk: every pixel's path shares the coarse bricks.
    if (!leader) return;
    uint32_t ox = 0, oy = 0, oz = 0, half = n >> 1, branch = 0, depthLvl = 1;
    uint32_t pl = FF_BRANCHINIT, pld = 0;
    bool missing = false;
    while ((half >> 1) >= target) {
        const uint32_t co = (X >= ox + half) * 1u + (Y >= oy + half) * 2u + (Z >= oz + half) * 4u;
        const uint32_t nx = ox + half * (co & 1u), ny = oy + half * ((co >> 1) & 1u), nz = oz + half * ((co >> 2) & 1u);
        if (!missing) {
            const uint32_t ptr = branch + co + 8;
            const uint32_t child = m.tree[ptr];
            const uint32_t leaf = m.tree[ptr - 8];
            if (leaf < m.nLeavesTotal) {
                pl = leaf;
                pld = depthLvl;
                m.stamp[leaf] = stamp;
            }
            if (child == FF_BRANCHINIT) {
                missing = true;
                hash_insert_min(hv, node_key(0, depthLvl, nx, ny, nz), pix, branch, fc);
            }
            branch = child;
        } else {
            hash_insert_min(hv, node_key(0, depthLvl, nx, ny, nz), pix, FF_BRANCHINIT, fc);
        }
        ox = nx; oy = ny; oz = nz;
        half >>= 1;
        depthLvl++;
    }
    const uint32_t co = (X >= ox + half) * 1u + (Y >= oy + half) * 2u + (Z >= oz + half) * 4u;
    const uint32_t nx = ox + half * (co & 1u), ny = oy + half * ((co >> 1) & 1u), nz = oz + half * ((co >> 2) & 1u);
    if (!missing) {
        const uint32_t leaf = m.tree[branch + co];
        if (leaf >= FF_BRANCHINIT)
            hash_insert_min(hv, node_key(1, depthLvl, nx, ny, nz), pix, branch, fc);
        else if (leaf < m.nLeavesTotal)
            m.stamp[leaf] = stamp;
    } else {
        hash_insert_min(hv, node_key(1, depthLvl, nx, ny, nz), pix, FF_BRANCHINIT, fc);
    }
    pl0[2 * pix] = pl;
    pl0[2 * pix + 1] = pld;
}

__global__ void k_descend(const FrameDev *__restrict__ fd, int W, int H, int words, const float *__restrict__ pxp,
                          const float *__restrict__ pyp, const uint32_t *__restrict__ pbits, MapView m, HashView hv,
                          uint32_t *__restrict__ pl0, FrameCounters *fc)
{
    descend_body(fd, W, H, words, pxp, pyp, pbits, m, hv, pl0, fc);
}

__global__ void k_descend_b(const AllocStream *__restrict__ S)
{
    const AllocStream &a = S[blockIdx.z];
    if ((int)blockIdx.y >= a.H || (int)(blockIdx.x * blockDim.x) >= a.W) return;
    descend_body(a.fd, a.W, a.H, a.words, a.pxp, a.pyp, a.pbits, a.m, a.hv, a.pl0, a.fc);
}

// ---------------------------------------------------------------------------
// Indices of the frame's new nodes.  The reference creates nodes pixel by
// pixel in raster order, each pixel's missing branches from the top down,
// then its brick; branch and brick indices are handed out in that order.
// So a new branch's index is nBranchesUsed + 16 x (its rank among the new
// branches by (creating pixel, depth)) and a new brick's is nLeavesUsed + its
// rank among the new bricks by creating pixel.  k_order sorts up to
// FF_ORDER_MAX nodes in one block (bitonic, in LDS); a frame with more (a
// first frame, a large jump) is flagged and ordered by the host through the
// per-pixel count scan below (k_resolve, k_scan_a/b, k_assign).
// ---------------------------------------------------------------------------
#define FF_ORDER_MAX 8192   // nodes ranked in LDS (64 KiB of keys)
#define FF_ORDER_THREADS 1024

// key: FF_ORDER_MAX words of LDS
__device__ __forceinline__ void order_body(const HashView &hv, FrameCounters *fc, const MapCounters *__restrict__ mapc,
                                           uint64_t *key)
{
    // key: type (63) | creating pixel (29..50, < 2^22) | depth (24..28) | list
    // index (0..12): sorted ascending, branches come first in (pixel,
    // depth) order, then bricks in pixel order
    __shared__ uint32_t nBr;
    const uint32_t n = min(fc->listCount, hv.listCap);
    if (n > FF_ORDER_MAX) {
        if (threadIdx.x == 0) atomicOr(&fc->flags, FF_FLAG_ORDER_BIG);
        return;
    }
    const uint32_t t = threadIdx.x;
    uint32_t N = 2;
    while (N < n) N <<= 1;
    if (t == 0) nBr = 0;
    __syncthreads();
    uint32_t myBr = 0;
    for (uint32_t i = t; i < N; i += FF_ORDER_THREADS) {
        uint64_t k = ~0ull;  // padding sorts last
        if (i < n) {
            const uint32_t s = hv.list[i];
            const uint64_t hk = hv.keys[s];
            const uint64_t type = (hk >> 53) & 1ull;
            k = (type << 63) | ((uint64_t)hv.minpix[s] << 29) | (((hk >> 48) & 31ull) << 24) | (uint64_t)i;
            myBr += type ? 0u : 1u;
        }
        key[i] = k;
    }
    if (myBr) atomicAdd(&nBr, myBr);
    __syncthreads();
    // bitonic sort of the N keys
    for (uint32_t k = 2; k <= N; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t p = t; p < N / 2; p += FF_ORDER_THREADS) {
                const uint32_t i = ((p & ~(j - 1u)) << 1) | (p & (j - 1u)), l = i + j;
                const uint64_t a = key[i], b = key[l];
                if ((a > b) == ((i & k) == 0u)) {
                    key[i] = b;
                    key[l] = a;
                }
            }
            __syncthreads();
        }
    const uint32_t nb = nBr;
    for (uint32_t r = t; r < n; r += FF_ORDER_THREADS) {
        const uint64_t k = key[r];
        const uint32_t s = hv.list[(uint32_t)k & 0x1fffu];
        hv.assigned[s] = (k >> 63) ? mapc->nLeavesUsed + (r - nb) : mapc->nBranchesUsed + 16u * r;
    }
    if (t == 0) {
        fc->newB = nb;
        fc->newL = n - nb;
    }
}

__global__ __launch_bounds__(FF_ORDER_THREADS) void k_order(HashView hv, FrameCounters *fc,
                                                        const MapCounters *__restrict__ mapc)
{
    __shared__ uint64_t key[FF_ORDER_MAX];
    order_body(hv, fc, mapc, key);
}

__global__ __launch_bounds__(FF_ORDER_THREADS) void k_order_b(const AllocStream *__restrict__ S)
{
    const AllocStream &a = S[blockIdx.z];
    __shared__ uint64_t key[FF_ORDER_MAX];
    order_body(a.hv, a.fc, a.mapc, key);
}

// per new node: its creating pixel counts it
__device__ __forceinline__ void resolve_body(const HashView &hv, const FrameCounters *fc, uint32_t *__restrict__ cntB,
                                             uint32_t *__restrict__ cntL, uint32_t *__restrict__ firstD)
{
    const uint32_t n = min(fc->listCount, hv.listCap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t s = hv.list[i];
        const uint64_t k = hv.keys[s];
        const uint32_t pix = hv.minpix[s];
        const uint32_t type = (uint32_t)(k >> 53) & 1u, d = (uint32_t)(k >> 48) & 31u;
        if (type == 0) {
            atomicAdd(&cntB[pix], 1u);
            atomicMin(&firstD[pix], d);
        } else {
            cntL[pix] = 1u;
        }
    }
}

__global__ void k_resolve(HashView hv, const FrameCounters *fc, uint32_t *__restrict__ cntB,
                          uint32_t *__restrict__ cntL, uint32_t *__restrict__ firstD)
{
    resolve_body(hv, fc, cntB, cntL, firstD);
}


// ---------------------------------------------------------------------------
// exclusive scan of (cntB, cntL) over pixels, 4096 pixels per block
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t u = __shfl_up(v, o);
        if (lane >= o) v += u;
    }
    return v;
}

__device__ uint64_t block_excl_scan(uint64_t v, uint64_t *total)
{
    __shared__ uint64_t ws[17];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint64_t inc = wave_incl_scan(v);
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    if (w == 0) {
        uint64_t t = lane < nw ? ws[lane] : 0;
        const uint64_t ti = wave_incl_scan(t);
        if (lane < nw) ws[lane] = ti - t;
        if (lane == nw - 1) ws[16] = ti;
    }
    __syncthreads();
    const uint64_t r = ws[w] + inc - v;
    *total = ws[16];
    __syncthreads();
    return r;
}

__device__ __forceinline__ void scan_a_body(const uint32_t *__restrict__ cntB, const uint32_t *__restrict__ cntL,
                                            uint32_t n, uint64_t *__restrict__ pref, uint64_t *__restrict__ bsum)
{
    const uint32_t base = blockIdx.x * 4096 + threadIdx.x * 4;
    uint64_t e[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t i = base + k;
        e[k] = i < n ? ((uint64_t)cntB[i] | ((uint64_t)cntL[i] << 32)) : 0ull;
        s += e[k];
    }
    uint64_t tot;
    uint64_t ex = block_excl_scan(s, &tot);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t i = base + k;
        if (i < n) pref[i] = ex;
        ex += e[k];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_a(const uint32_t *__restrict__ cntB, const uint32_t *__restrict__ cntL,
                                                 uint32_t n, uint64_t *__restrict__ pref, uint64_t *__restrict__ bsum)
{
    scan_a_body(cntB, cntL, n, pref, bsum);
}


__device__ __forceinline__ void scan_b_body(uint64_t *__restrict__ bsum, uint32_t nb, FrameCounters *fc)
{
    const uint32_t per = (nb + 1023) / 1024;
    uint64_t loc[8], s = 0;
    for (uint32_t k = 0; k < per && k < 8; k++) {
        const uint32_t i = threadIdx.x * per + k;
        loc[k] = i < nb ? bsum[i] : 0ull;
        s += loc[k];
    }
    uint64_t tot;
    uint64_t ex = block_excl_scan(s, &tot);
    for (uint32_t k = 0; k < per && k < 8; k++) {
        const uint32_t i = threadIdx.x * per + k;
        if (i < nb) bsum[i] = ex;
        ex += loc[k];
    }
    if (threadIdx.x == 0) {
        fc->newB = (uint32_t)tot;
        fc->newL = (uint32_t)(tot >> 32);
    }
}

__global__ __launch_bounds__(1024) void k_scan_b(uint64_t *__restrict__ bsum, uint32_t nb, FrameCounters *fc)
{
    scan_b_body(bsum, nb, fc);
}


__device__ __forceinline__ void assign_body(const HashView &hv, const FrameCounters *fc,
                                            const uint64_t *__restrict__ pref, const uint64_t *__restrict__ bsum,
                                            const uint32_t *__restrict__ firstD, const MapCounters *__restrict__ mapc)
{
    const uint32_t nBranchesUsed = mapc->nBranchesUsed, nLeavesUsed = mapc->nLeavesUsed;
    const uint32_t n = min(fc->listCount, hv.listCap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t s = hv.list[i];
        const uint64_t k = hv.keys[s];
        const uint32_t pix = hv.minpix[s];
        const uint32_t type = (uint32_t)(k >> 53) & 1u, d = (uint32_t)(k >> 48) & 31u;
        const uint64_t p = pref[pix] + bsum[pix >> 12];
        if (type == 0)
            hv.assigned[s] = nBranchesUsed + 16u * ((uint32_t)p + d - firstD[pix]);
        else
            hv.assigned[s] = nLeavesUsed + (uint32_t)(p >> 32);
    }
}

__global__ void k_assign(HashView hv, const FrameCounters *fc, const uint64_t *__restrict__ pref,
                         const uint64_t *__restrict__ bsum, const uint32_t *__restrict__ firstD,
                         const MapCounters *__restrict__ mapc)
{
    assign_body(hv, fc, pref, bsum, firstD, mapc);
}


// parent branch's tree index of a node created this frame
__device__ __forceinline__ uint32_t parent_branch(const HashView &hv, uint32_t aux, uint32_t d, uint32_t x,
                                                  uint32_t y, uint32_t z, uint32_t n)
{
    if (aux != FF_BRANCHINIT) return aux;
    const uint32_t psize = (n >> (d - 1));
    const uint32_t msk = ~(psize - 1u);
    const int s = hash_find(hv, node_key(0, d - 1, x & msk, y & msk, z & msk));
    return s >= 0 ? hv.assigned[s] : FF_BRANCHINIT;
}

// (slow: the host's ordering pass; otherwise a frame flagged for it skips)
__device__ __forceinline__ void link_body(const HashView &hv, const MapView &m, const FrameCounters *fc,
                                          const uint32_t *__restrict__ pl0, const FrameDev *__restrict__ fd,
                                          FrameCounters *fcw, bool slow)
{
    if (!slow && (fc->flags & FF_FLAG_ORDER_BIG)) return;
    // a frame whose node table or node list overflowed has an incomplete
    // node set: nothing of it enters the tree (nothing is committed either)
    if ((fc->flags & FF_FLAG_HASH_FULL) || fc->listCount > hv.listCap) return;
    const uint32_t n = fd->n, stamp = fd->stamp;
    const uint32_t cnt = min(fc->listCount, hv.listCap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
        const uint32_t s = hv.list[i];
        const uint64_t k = hv.keys[s];
        const uint32_t type = (uint32_t)(k >> 53) & 1u, d = (uint32_t)(k >> 48) & 31u;
        const uint32_t x = (uint32_t)(k >> 32) & 0xffffu, y = (uint32_t)(k >> 16) & 0xffffu, z = (uint32_t)k & 0xffffu;
        const uint32_t size = n >> d;
        const uint32_t co = ((x / size) & 1u) | (((y / size) & 1u) << 1) | (((z / size) & 1u) << 2);
        const uint32_t pb = parent_branch(hv, hv.aux[s], d, x, y, z, n);
        const uint32_t idx = hv.assigned[s];
        const bool pb_ok = pb != FF_BRANCHINIT && pb + 16u <= m.nBranchesTotal;
        if (type == 0) {
            if (pb_ok && idx + 16u <= m.nBranchesTotal) m.tree[pb + co + 8] = idx;
            else atomicOr(&fcw->flags, FF_FLAG_POOL_FULL);
            continue;
        }
        if (!pb_ok) {
            atomicOr(&fcw->flags, FF_FLAG_POOL_FULL);
            continue;
        }
        m.tree[pb + co] = idx;
        if (idx >= m.nLeavesTotal) continue;
        // parent brick at the creating pixel's turn: the deepest brick on the
        // path that existed before this frame or was created by an earlier pixel
        const uint32_t pix = hv.minpix[s];
        uint32_t pl = pl0[2 * pix], pld = pl0[2 * pix + 1];
        for (uint32_t dd = d - 1; dd > pld && dd >= 1; dd--) {
            const uint32_t as = n >> dd, am = ~(as - 1u);
            const int e = hash_find(hv, node_key(1, dd, x & am, y & am, z & am));
            if (e >= 0 && hv.minpix[e] < pix && hv.assigned[e] < m.nLeavesTotal) {
                pl = hv.assigned[e];
                break;
            }
        }
        m.leafinfo[idx] = make_ushort4((unsigned short)x, (unsigned short)y, (unsigned short)z,
                                       (unsigned short)(size / FF_BL));
        m.parent[idx] = pl;
        m.stamp[idx] = stamp;  // queued by k_queue
    }
}

__global__ void k_link(HashView hv, MapView m, const FrameCounters *fc, const uint32_t *__restrict__ pl0,
                       const FrameDev *__restrict__ fd, FrameCounters *fcw, int slow)
{
    link_body(hv, m, fc, pl0, fd, fcw, slow != 0);
}

__global__ void k_link_b(const AllocStream *__restrict__ S)
{
    const AllocStream &a = S[blockIdx.z];
    link_body(a.hv, a.m, a.fc, a.pl0, a.fd, a.fc, false);
}

// The end of the allocation phase, per new node: the branch's geometry for
// the mesher (origin, half size) and the node-table slot's reset; then the
// frame's new branches and bricks join the map's used counts (a frame whose
// node table overflowed commits nothing; the host reports it).  k_link,
// whose lookups read other nodes' slots, has completed.
__device__ __forceinline__ void alloc_end_body(const HashView &hv, const FrameCounters *__restrict__ fc,
                                               MapCounters *mapc, ushort4 *binfo, uint32_t binfoBlocks,
                                               const FrameDev *__restrict__ fd, bool slow)
{
    if (!slow && (fc->flags & FF_FLAG_ORDER_BIG)) return;
    const uint32_t n = fd->n;
    const uint32_t cnt = min(fc->listCount, hv.listCap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
        const uint32_t s = hv.list[i];
        const uint64_t k = hv.keys[s];
        if (!((k >> 53) & 1u)) {
            const uint32_t d = (uint32_t)(k >> 48) & 31u;
            const uint32_t idx = hv.assigned[s];
            if (idx / 16 < binfoBlocks)
                binfo[idx / 16] = make_ushort4((unsigned short)((k >> 32) & 0xffffu),
                                               (unsigned short)((k >> 16) & 0xffffu), (unsigned short)(k & 0xffffu),
                                               (unsigned short)((n >> d) >> 1));
        }
        hv.keys[s] = FF_HEMPTY;
        hv.minpix[s] = 0xffffffffu;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && !(fc->flags & FF_FLAG_HASH_FULL) && fc->listCount <= hv.listCap) {
        mapc->nBranchesUsed += 16u * fc->newB;
        mapc->nLeavesUsed += fc->newL;
    }
}

__global__ void k_alloc_end(HashView hv, const FrameCounters *__restrict__ fc, MapCounters *mapc, ushort4 *binfo,
                            uint32_t binfoBlocks, const FrameDev *__restrict__ fd, int slow)
{
    alloc_end_body(hv, fc, mapc, binfo, binfoBlocks, fd, slow != 0);
}

__global__ void k_alloc_end_b(const AllocStream *__restrict__ S)
{
    const AllocStream &a = S[blockIdx.z];
    alloc_end_body(a.hv, a.fc, a.mapc, a.binfo, a.binfoCap, a.fd, false);
}

// ---------------------------------------------------------------------------
// k_queue: the frame's queue (_leafNumberSurface) is every brick carrying the
// frame's stamp -- those on a processed pixel's path (k_descend) and the new
// ones (k_link).  A block takes FF_QUEUE_SPAN bricks at a time and appends
// them with one atomic.  The reference queues bricks in traversal order; the
// SDF update does not depend on the order.
// ---------------------------------------------------------------------------
#define FF_QUEUE_THREADS 1024
#define FF_QUEUE_SPAN (8 * FF_QUEUE_THREADS)

// (the used count can pass the capacity when the pools ran out; bricks
// exist only below nLeavesTotal)
__device__ __forceinline__ void queue_body(const uint32_t *__restrict__ stamp, const MapCounters *__restrict__ mapc,
                                           uint32_t nLeavesTotal, const FrameDev *__restrict__ fd,
                                           uint32_t *__restrict__ queue, FrameCounters *fc, bool slow)
{
    __shared__ uint32_t wsum[FF_QUEUE_THREADS / 64 + 1];
    if (!slow && (fc->flags & FF_FLAG_ORDER_BIG)) return;
    const uint32_t nb = min(mapc->nLeavesUsed, nLeavesTotal), st = fd->stamp;
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    for (uint32_t base = blockIdx.x * FF_QUEUE_SPAN; base < nb; base += gridDim.x * FF_QUEUE_SPAN) {
        uint32_t mask = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint32_t i = base + (uint32_t)k * FF_QUEUE_THREADS + t;
            if (i < nb && stamp[i] == st) mask |= 1u << k;
        }
        // block-exclusive prefix of the per-thread counts
        const uint32_t c = (uint32_t)__popc(mask);
        uint32_t incl = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u = __shfl_up(incl, o);
            if (lane >= (uint32_t)o) incl += u;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        if (t == 0) {
            uint32_t tot = 0;
            for (uint32_t k = 0; k < FF_QUEUE_THREADS / 64; k++) {
                const uint32_t x = wsum[k];
                wsum[k] = tot;
                tot += x;
            }
            wsum[FF_QUEUE_THREADS / 64] = tot ? atomicAdd(&fc->nQueued, tot) : 0u;
        }
        __syncthreads();
        uint32_t q = wsum[FF_QUEUE_THREADS / 64] + wsum[w] + incl - c;
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (mask & (1u << k)) queue[q++] = base + (uint32_t)k * FF_QUEUE_THREADS + t;
        __syncthreads();  // wsum is reused
    }
}

__global__ __launch_bounds__(FF_QUEUE_THREADS) void k_queue(const uint32_t *__restrict__ stamp,
                                                            const MapCounters *__restrict__ mapc, uint32_t nLeavesTotal,
                                                            const FrameDev *__restrict__ fd, uint32_t *__restrict__ queue,
                                                            FrameCounters *fc, int slow)
{
    queue_body(stamp, mapc, nLeavesTotal, fd, queue, fc, slow != 0);
}

__global__ __launch_bounds__(FF_QUEUE_THREADS) void k_queue_b(const AllocStream *__restrict__ S)
{
    const AllocStream &a = S[blockIdx.z];
    queue_body(a.m.stamp, a.mapc, a.m.nLeavesTotal, a.fd, a.m.queue, a.fc, false);
}

// The queue of a frame built by one block, ordered by the image row band
// its bricks project to (a counting sort over FF_QBANDS bands of the image
// height, by each brick centre's projection with the frame's camera): the
// SDF update hands contiguous eighths of the queue to the 8 XCDs, which then
// each read one horizontal strip of the image.  Order within a band is
// whatever the block's atomics give; the update does not depend on it.
#define FF_QBANDS 32

__device__ __forceinline__ uint32_t queue_band(const ushort4 li, const FrameDev *__restrict__ fd, int H)
{
    const SdfParams &P = fd->sp;
    const float h = 4.0f * (float)li.w;  // the brick's centre, in voxels
    const float c[3] = {(float)li.x + h, (float)li.y + h, (float)li.z + h};
    float q[3];
#pragma unroll
    for (int r = 0; r < 3; r++)
        q[r] = (P.m[4 * r] * c[0] + P.m[4 * r + 1] * c[1] + P.m[4 * r + 2] * c[2]) * P.scale + P.m[4 * r + 3];
    if (!(q[2] > 0.0f)) return 0u;
    const float v = P.cy + P.fy * (q[1] / q[2]);
    const int b = (int)(v * (float)FF_QBANDS / (float)H);
    return (uint32_t)min(max(b, 0), FF_QBANDS - 1);
}

// the frame's queue (built by queue_body in brick order) reordered by band:
// entries and bands staged in `stage` (at least FF_ORDER_MAX words of LDS),
// a counting sort back into the queue; a queue longer than the stage stays
// in brick order
__device__ __forceinline__ void queue_band_sort(const ushort4 *__restrict__ info, const FrameDev *__restrict__ fd,
                                                int H, uint32_t *queue, FrameCounters *fc, uint64_t *stage)
{
    __shared__ uint32_t cnt[FF_QBANDS];
    // the count was taken up by an atomic and the entries stored by other
    // waves: both read past the vector L1
    const uint32_t n = __hip_atomic_load(&fc->nQueued, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t t = threadIdx.x;
    if (n > FF_ORDER_MAX) return;  // (uniform)
    if (t < FF_QBANDS) cnt[t] = 0u;
    __syncthreads();
    for (uint32_t j = t; j < n; j += blockDim.x) {
        const uint32_t i = __hip_atomic_load(&queue[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t band = queue_band(info[i], fd, H);
        stage[j] = ((uint64_t)band << 32) | i;
        atomicAdd(&cnt[band], 1u);
    }
    __syncthreads();
    if (t == 0) {
        uint32_t tot = 0;
        for (int b = 0; b < FF_QBANDS; b++) {
            const uint32_t c = cnt[b];
            cnt[b] = tot;
            tot += c;
        }
    }
    __syncthreads();
    for (uint32_t j = t; j < n; j += blockDim.x) {
        const uint64_t e = stage[j];
        queue[atomicAdd(&cnt[(uint32_t)(e >> 32)], 1u)] = (uint32_t)e;
    }
}

// blocks of k_queue for a map of nLeavesUsed bricks (its grid-stride loop
// covers any count)
static inline uint32_t queue_grid(uint32_t nLeavesUsed)
{
    return std::min<uint32_t>(256u, nLeavesUsed / FF_QUEUE_SPAN + 2u);
}

// ---------------------------------------------------------------------------
// k_alloc_tail: k_order, k_link, k_alloc_end and (queue) k_queue as the
// passes of one block, each behind a block barrier instead of a launch.
// Their work is the frame's new nodes (at most FF_ORDER_MAX) and, for the
// queue, a map of at most FF_TAIL_SPANS spans of bricks.  Within the block
// plain stores are seen by the later passes; the ORDER_BIG flag is set by an
// atomic, so it is read with an atomic load.
// ---------------------------------------------------------------------------
#define FF_TAIL_SPANS 4

// The frame's counters, once its allocation is done, to mapped host memory
// (one thread): the first 48 bytes, then, after a system-scope fence, the
// frame's stamp in pad[0], so that a host that sees the stamp sees them.
// The host then returns from the frame: the SDF update's gate is evaluated
// on the host from the same counters.
__device__ __forceinline__ void frame_report(const FrameCounters *fc, FrameCounters *hostFc, uint32_t stamp)
{
    const uint4 *s = reinterpret_cast<const uint4 *>(fc);
    uint4 *d = reinterpret_cast<uint4 *>(hostFc);
    const uint4 a = s[0], b = s[1], c = s[2];
    d[0] = a;
    d[1] = b;
    d[2] = c;
    __threadfence_system();
    *reinterpret_cast<volatile uint32_t *>(&hostFc->pad[0]) = stamp;
}

// hostFc: where the frame's counters are reported (nullptr: not here -- a
// separate k_queue follows, and k_frame_report after it)
__global__ __launch_bounds__(FF_ORDER_THREADS) void k_alloc_tail(HashView hv, MapView m, FrameCounters *fc,
                                                             const uint32_t *__restrict__ pl0,
                                                             const FrameDev *__restrict__ fd, MapCounters *mapc,
                                                             ushort4 *binfo, uint32_t binfoBlocks, int queue, int H,
                                                             FrameCounters *hostFc)
{
    __shared__ uint32_t big;
    __shared__ uint64_t key[FF_ORDER_MAX];  // the ordering's sort keys, then the queue's band sort
    order_body(hv, fc, mapc, key);
    __syncthreads();
    if (threadIdx.x == 0)
        big = __hip_atomic_load(&fc->flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & FF_FLAG_ORDER_BIG;
    __syncthreads();
    if (big) {
        // the host orders the frame (it sees the flag in the report)
        if (threadIdx.x == 0 && hostFc) frame_report(fc, hostFc, fd->stamp);
        return;
    }
    link_body(hv, m, fc, pl0, fd, fc, true);
    __syncthreads();  // every lookup of the node table is done before its slots are reset
    alloc_end_body(hv, fc, mapc, binfo, binfoBlocks, fd, true);
    if (!queue) return;
    __syncthreads();  // the used counts are committed
    queue_body(m.stamp, mapc, m.nLeavesTotal, fd, m.queue, fc, true);
    __syncthreads();  // the queue and its count are written
    queue_band_sort(m.leafinfo, fd, H, m.queue, fc, key);
    __syncthreads();
    if (threadIdx.x == 0 && hostFc) frame_report(fc, hostFc, fd->stamp);
}

// the report after a separate k_queue
__global__ void k_frame_report(const FrameCounters *fc, FrameCounters *hostFc, const FrameDev *__restrict__ fd)
{
    if (threadIdx.x == 0) frame_report(fc, hostFc, fd->stamp);
}

// ---------------------------------------------------------------------------
// grow (geometryfusion_mipmap_cpu.cpp:3071-3319): shift child-branch pointers,
// move the tree, shift brick positions
// ---------------------------------------------------------------------------
__global__ void k_shift_children(uint32_t *tree, uint32_t nBranchesUsed, uint32_t shift)
{
    const uint32_t b = (blockIdx.x * blockDim.x + threadIdx.x);
    const uint32_t slot = (b / 8) * 16 + 8 + (b % 8);
    if (slot < nBranchesUsed && tree[slot] != FF_BRANCHINIT) tree[slot] += shift;
}

__global__ void k_copy_u32(const uint32_t *src, uint32_t *dst, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void k_fill_u32(uint32_t *dst, uint32_t v, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}

__global__ void k_shift_leaves(ushort4 *info, uint32_t n, uint32_t ax, uint32_t ay, uint32_t az)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        ushort4 v = info[i];
        v.x = (unsigned short)(v.x + ax);
        v.y = (unsigned short)(v.y + ay);
        v.z = (unsigned short)(v.z + az);
        info[i] = v;
    }
}

// ---------------------------------------------------------------------------
// k_sdf: update8AddLoopAVXSingleInteger (update_functions.cpp:1187-1444).
// One 512-thread block per queued brick, one voxel per thread.  Each thread
// rebuilds its row origin with the same sequential adds as the reference
// (z steps, then y steps) and adds k * d1j like the AVX lane k.
// ---------------------------------------------------------------------------
// one voxel (column k of its row): old (dA, wA, cA) -> new (dN, wN, cN)
__device__ __forceinline__ void sdf_point(const SdfRow &w, int k, float dA, float wA, const uint16_t cA[3],
                                          const uint16_t *__restrict__ depth, const uint8_t *__restrict__ bgr,
                                          int W, int H, const SdfParams &P, const uint32_t *tab, float &dN,
                                          float &wN, uint16_t cN[3])
{
    const float EPS = 0.005f;
    const float fk = (float)k;
    // lane 0 adds +0.0f like _mm256_setr_ps(0.0f, ...)
    const float px = rz_add(w.R[0], k ? rz_mul(fk, w.d1[0]) : 0.0f);
    const float py = rz_add(w.R[1], k ? rz_mul(fk, w.d1[1]) : 0.0f);
    const float pz = rz_add(w.R[2], k ? rz_mul(fk, w.d1[2]) : 0.0f);
    const float len = rz_sqrt(rz_add(rz_add(rz_mul(px, px), rz_mul(py, py)), rz_mul(pz, pz)));
    const float r = rcp_x86(pz, tab);
    const int imx = cvt_x86(rz_add(P.cx, rz_mul(P.fx, rz_mul(px, r))));
    const int imy = cvt_x86(rz_add(P.cy, rz_mul(P.fy, rz_mul(py, r))));
    const int cyi = max(min(imy, H - 1), 0), cxi = max(min(imx, W - 1), 0);
    const int ii = W * cyi + cxi;
    const float h8 = (float)depth[ii];
    const float h = rz_mul(P.scaling, h8);
    const float dI = rz_sub(len, rz_mul(rz_mul(len, r), h));
    const bool mask = (pz > 0.0f) && (imy < H) && (imy > -1) && (imx < W) && (imx > -1) && (h < P.maxcam);
    const float front = (dI < EPS) ? 1.0f : 0.0f;
    const float back = (dI >= EPS && dI < w.tw) ? 1.0f : 0.0f;
    const float fall = rz_mul(rz_sub(w.tw, dI), w.rfall);
    float wI = mask ? rz_add(front, rz_mul(fall, back)) : 0.0f;
    if (P.conf) wI = rz_mul(wI, P.conf[ii]);  // learned per-pixel confidence
    const float wn = rz_add(wA, wI);
    const float f = rcp_x86(rz_add(wn, (wn <= 0.0f) ? 1.0f : 0.0f), tab);
    const float mn = (w.td < dI) ? w.td : dI;             // _mm256_min_ps(td, dI)
    const float ntd = -w.td;
    const float mx = (ntd > mn) ? ntd : mn;               // _mm256_max_ps(-td, mn)
    const float add = mask ? rz_mul(mx, wI) : 0.0f;
    dN = rz_mul(rz_add(rz_mul(dA, wA), add), f);
    wN = wn;
    if (P.useColor) {
        const float cmax = 65280.0f;
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            const float inc = (float)bgr[ii * 3 + (2 - ch)] * 256.0f;
            float vv = rz_mul(rz_add(rz_mul((float)cA[ch], wA), rz_mul(inc, wI)), f);
            vv = (vv < cmax) ? vv : cmax;
            cN[ch] = (uint16_t)cvt_x86(vv);
        }
    }
}

// voxel v (z-major, 0..511) of brick `leaf`
__device__ __forceinline__ void sdf_voxel(uint32_t leaf, int v, const MapView &m, const uint16_t *__restrict__ depth,
                                          const uint8_t *__restrict__ bgr, int W, int H, const SdfParams &P,
                                          const uint32_t *tab)
{
    const size_t vi = (size_t)leaf * FF_BRICK + v;
    const float wA = m.wgt[vi], dA = m.dist[vi];
    uint16_t cA[3] = {0, 0, 0}, cN[3];
    uint16_t *cbase = m.col + (size_t)leaf * FF_BRICK * 3 + v;
    if (P.useColor) {
        cA[0] = cbase[0];
        cA[1] = cbase[FF_BRICK];
        cA[2] = cbase[2 * FF_BRICK];
    }
    const SdfRow w = sdf_row_setup(m.leafinfo[leaf], v >> 6, (v >> 3) & 7, P, tab);
    float dN, wN;
    sdf_point(w, v & 7, dA, wA, cA, depth, bgr, W, H, P, tab, dN, wN, cN);
    m.dist[vi] = dN;
    m.wgt[vi] = wN;
    if (P.useColor) {
        cbase[0] = cN[0];
        cbase[FF_BRICK] = cN[1];
        cbase[2 * FF_BRICK] = cN[2];
    }
}

// the 8 voxels of row `row` (zr * 8 + yr) of brick `leaf`: the row set-up
// once, 32-byte loads and stores of the row's distances and weights
__device__ __forceinline__ void sdf_row(uint32_t leaf, int row, const MapView &m, const uint16_t *__restrict__ depth,
                                        const uint8_t *__restrict__ bgr, int W, int H, const SdfParams &P,
                                        const uint32_t *tab)
{
    const size_t vb = (size_t)leaf * FF_BRICK + (size_t)row * 8;
    float4 *dp = reinterpret_cast<float4 *>(m.dist + vb);
    float4 *wp = reinterpret_cast<float4 *>(m.wgt + vb);
    const float4 d0 = dp[0], d1 = dp[1], w0 = wp[0], w1 = wp[1];
    const float dA[8] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w};
    const float wA[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
    uint16_t cA[3][8];
    uint4 *cp[3] = {nullptr, nullptr, nullptr};
    if (P.useColor) {
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            cp[ch] = reinterpret_cast<uint4 *>(m.col + (size_t)leaf * FF_BRICK * 3 + (size_t)ch * FF_BRICK + row * 8);
            const uint4 u = *cp[ch];
            const uint32_t q[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
            for (int j = 0; j < 4; j++) {
                cA[ch][2 * j] = (uint16_t)(q[j] & 0xffffu);
                cA[ch][2 * j + 1] = (uint16_t)(q[j] >> 16);
            }
        }
    }
    const SdfRow w = sdf_row_setup(m.leafinfo[leaf], row >> 3, row & 7, P, tab);
    float dN[8], wN[8];
    uint16_t cN[3][8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const uint16_t c3[3] = {P.useColor ? cA[0][k] : (uint16_t)0, P.useColor ? cA[1][k] : (uint16_t)0,
                                P.useColor ? cA[2][k] : (uint16_t)0};
        uint16_t o3[3] = {0, 0, 0};
        sdf_point(w, k, dA[k], wA[k], c3, depth, bgr, W, H, P, tab, dN[k], wN[k], o3);
        cN[0][k] = o3[0]; cN[1][k] = o3[1]; cN[2][k] = o3[2];
    }
    dp[0] = make_float4(dN[0], dN[1], dN[2], dN[3]);
    dp[1] = make_float4(dN[4], dN[5], dN[6], dN[7]);
    wp[0] = make_float4(wN[0], wN[1], wN[2], wN[3]);
    wp[1] = make_float4(wN[4], wN[5], wN[6], wN[7]);
    if (P.useColor) {
#pragma unroll
        for (int ch = 0; ch < 3; ch++)
            *cp[ch] = make_uint4((uint32_t)cN[ch][0] | ((uint32_t)cN[ch][1] << 16),
                                 (uint32_t)cN[ch][2] | ((uint32_t)cN[ch][3] << 16),
                                 (uint32_t)cN[ch][4] | ((uint32_t)cN[ch][5] << 16),
                                 (uint32_t)cN[ch][6] | ((uint32_t)cN[ch][7] << 16));
    }
}

#define FF_SDF_GRID 1024  // blocks of the SDF update; each loops over queued bricks

// A frame's counters and the map's used counts, each field of which the
// gate below tests: read as whole 16-byte words in flight together (field by
// field behind the tests' branches they were eight dependent round trips).
struct GateView {
    uint4 w[4];  // FrameCounters
    uint4 mc;    // MapCounters
};

__device__ __forceinline__ GateView gate_load(const FrameCounters *fc, const MapCounters *mapc)
{
    GateView v;
    const uint4 *p = reinterpret_cast<const uint4 *>(fc);
#pragma unroll
    for (int k = 0; k < 4; k++) v.w[k] = p[k];
    v.mc = *reinterpret_cast<const uint4 *>(mapc);
    return v;
}

// the map's used counts are the committed ones (after this frame); the
// words of FrameCounters: w[0] = nQueued, listCount, newB, newL;
// w[1] = bmin[0..2], bmax[0]; w[2] = bmax[1..2], flags, nProcessed
__device__ __forceinline__ bool sdf_gate_open(const GateView &v, uint32_t n, const SdfGate &g)
{
    const uint32_t listCount = v.w[0].y, flags = v.w[2].z;
    const int bmin0 = (int)v.w[1].x, bmin1 = (int)v.w[1].y, bmin2 = (int)v.w[1].z;
    const int bmax0 = (int)v.w[1].w, bmax1 = (int)v.w[2].x, bmax2 = (int)v.w[2].y;
    const bool held = ((flags & (FF_FLAG_HASH_FULL | FF_FLAG_ORDER_BIG)) != 0u) | (listCount > g.listCap);
    const bool out = (bmin0 < 0) | (bmin1 < 0) | (bmin2 < 0) | (bmax0 >= (int)n) | (bmax1 >= (int)n) |
                     (bmax2 >= (int)n);
    const bool grows = (v.mc.y < g.leavesTotal) & (v.mc.x < g.branchesTotal);  // nLeavesUsed, nBranchesUsed
    return g.force || !(held | (out & grows));
}

// tiles: the tile count of the mesh update this launch follows the sampling
// of (nullptr: it follows the whole update); above tileCap, the update's
// marching cubes read bricks, and the launch stands down for the host
__global__ __launch_bounds__(512) void k_sdf(const uint32_t *__restrict__ queue, MapView m,
                                             const FrameDev *__restrict__ fd, int W, int H, FrameCounters *fc,
                                             const MapCounters *__restrict__ mapc, SdfGate gate,
                                             const uint32_t *tiles, uint32_t tileCap)
{
    rz_mode_on();  // first: every float operation below rounds toward zero
    // the counters, the rcp table's entries and the first queue slot are
    // all requested before anything waits on them
    const GateView gv = gate_load(fc, mapc);
    const uint32_t nTiles = tiles ? *tiles : 0u;
    uint32_t tv[2048 / 512];
#pragma unroll
    for (int k = 0; k < 2048 / 512; k++) tv[k] = c_rcp_table[threadIdx.x + 512 * k];
    // blocks loop over the queued bricks; the count comes from the device
    const bool open = sdf_gate_open(gv, fd->n, gate) && nTiles <= tileCap;
    if (!open) return;
    const uint16_t *__restrict__ depth = fd->depth;
    const uint8_t *__restrict__ bgr = fd->bgr;
    const SdfParams P = fd->sp;
    const uint32_t nQ = gv.w[0].x;
    if (blockIdx.x == 0 && threadIdx.x == 0) fc->sdfRan = 1u;
    // Workgroups go to the 8 XCDs in turn (block b on XCD b % 8); XCD x
    // takes the x-th eighth of the queue.  The queue is in brick order, and
    // bricks made together lie near each other, so each XCD's L2 holds the
    // image lines of its own part of the view instead of all of it.
    const uint32_t x = blockIdx.x & 7u, per = gridDim.x >> 3;
    const uint32_t lo = (uint32_t)(((uint64_t)nQ * x) >> 3), hi = (uint32_t)(((uint64_t)nQ * (x + 1)) >> 3);
    uint32_t b = lo + (blockIdx.x >> 3);
    if (b >= hi) return;
    const uint32_t leaf0 = queue[b];
    __shared__ uint32_t tab[2048];
#pragma unroll
    for (int k = 0; k < 2048 / 512; k++) tab[threadIdx.x + 512 * k] = tv[k];
    __syncthreads();  // rcp table ready
    sdf_voxel(leaf0, threadIdx.x, m, depth, bgr, W, H, P, tab);
    for (b += per; b < hi; b += per) sdf_voxel(queue[b], threadIdx.x, m, depth, bgr, W, H, P, tab);
}

// ---------------------------------------------------------------------------
// batched SDF update: the queued bricks of n streams as one worklist
// ---------------------------------------------------------------------------
// per stream: its gate, its count (0 when the gate is closed), the exclusive
// prefix over streams (n <= 64: one wave)
__global__ void k_batch_prefix(const SdfStream *__restrict__ ss, int n, uint32_t *__restrict__ prefix)
{
    const int lane = threadIdx.x;
    uint32_t cnt = 0;
    if (lane < n) {
        const SdfStream &S = ss[lane];
        const GateView gv = gate_load(S.fc, S.mapc);
        const bool open = sdf_gate_open(gv, S.fd->n, S.gate);
        cnt = open ? gv.w[0].x : 0u;
        S.fc->sdfRan = open ? 1u : 0u;
    }
    uint32_t x = cnt;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane < n) prefix[lane] = x - cnt;
    if (lane == 63) prefix[n] = x;
}

// blocks take bpw consecutive bricks of the worklist at a time; tpb threads
// per brick, 512 / tpb voxels per thread
template <int TPB>
__global__ __launch_bounds__(TPB) void k_sdf_batch(const SdfStream *__restrict__ ss, int n,
                                                   const uint32_t *__restrict__ prefix, int bpw)
{
    rz_mode_on();
    const uint32_t total = prefix[n];
    if ((uint32_t)blockIdx.x * (uint32_t)bpw >= total) return;
    __shared__ uint32_t tab[2048];
    for (int i = threadIdx.x; i < 2048; i += TPB) tab[i] = c_rcp_table[i];
    __syncthreads();
    for (uint32_t chunk = blockIdx.x; chunk * (uint32_t)bpw < total; chunk += gridDim.x) {
        for (int k = 0; k < bpw; k++) {
            const uint32_t item = chunk * (uint32_t)bpw + (uint32_t)k;
            if (item >= total) break;
            int lo = 0, hi = n;  // prefix[lo] <= item < prefix[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (prefix[mid] <= item) lo = mid;
                else hi = mid;
            }
            const SdfStream &S = ss[lo];
            const FrameDev *fd = S.fd;
            const uint32_t leaf = S.queue[item - prefix[lo]];
#pragma unroll
            for (int v = threadIdx.x; v < FF_BRICK; v += TPB)
                sdf_voxel(leaf, v, S.m, fd->depth, fd->bgr, S.W, S.H, fd->sp, tab);
        }
    }
}

// one wave per brick, lane = brick row (8 voxels): the row set-up is shared by
// its 8 voxels; waves take bpw consecutive bricks of the worklist at a time
__global__ __launch_bounds__(256) void k_sdf_batch_rows(const SdfStream *__restrict__ ss, int n,
                                                        const uint32_t *__restrict__ prefix, int bpw)
{
    rz_mode_on();
    const uint32_t total = prefix[n];
    const uint32_t wavesPerBlock = blockDim.x >> 6;
    if ((uint32_t)blockIdx.x * wavesPerBlock * (uint32_t)bpw >= total) return;
    __shared__ uint32_t tab[2048];
    for (int i = threadIdx.x; i < 2048; i += blockDim.x) tab[i] = c_rcp_table[i];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t nWaves = gridDim.x * wavesPerBlock;
    for (uint32_t chunk = blockIdx.x * wavesPerBlock + (threadIdx.x >> 6); chunk * (uint32_t)bpw < total;
         chunk += nWaves) {
        for (int k = 0; k < bpw; k++) {
            const uint32_t item = chunk * (uint32_t)bpw + (uint32_t)k;
            if (item >= total) break;
            int lo = 0, hi = n;  // prefix[lo] <= item < prefix[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (prefix[mid] <= item) lo = mid;
                else hi = mid;
            }
            const SdfStream &S = ss[lo];
            const FrameDev *fd = S.fd;
            sdf_row(S.queue[item - prefix[lo]], (int)lane, S.m, fd->depth, fd->bgr, S.W, S.H, fd->sp, tab);
        }
    }
}

__global__ void k_gather_fc(const SdfStream *__restrict__ ss, int n, FrameCounters *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = *ss[i].fc;
}

#define FF_SDF_BATCH_GRID 2048

void launch_batch_prefix(hipStream_t st, const SdfStream *d_ss, int n, uint32_t *d_prefix)
{
    k_batch_prefix<<<1, 64, 0, st>>>(d_ss, n, d_prefix);
}

void launch_sdf_batch(hipStream_t st, const SdfStream *d_ss, int n, uint32_t *d_prefix, int tpb, int bpw)
{
    if (tpb == 64) k_sdf_batch_rows<<<FF_SDF_BATCH_GRID, 256, 0, st>>>(d_ss, n, d_prefix, bpw);
    else if (tpb == 256) k_sdf_batch<256><<<FF_SDF_BATCH_GRID, 256, 0, st>>>(d_ss, n, d_prefix, bpw);
    else k_sdf_batch<512><<<FF_SDF_BATCH_GRID, 512, 0, st>>>(d_ss, n, d_prefix, bpw);
}

void launch_gather_fc(hipStream_t st, const SdfStream *d_ss, int n, FrameCounters *d_out)
{
    k_gather_fc<<<1, 64, 0, st>>>(d_ss, n, d_out);
}

}  // namespace ff

// ===========================================================================
// host side
// ===========================================================================
namespace ff {

static bool hip_ok(ff_ctx *c, hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    c->err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}

#define HIPCHK(call)                                  \
    do {                                              \
        if (!hip_ok(c, (call), #call)) return FF_E_DEVICE; \
    } while (0)

static CamF make_p(const ff_ctx *c, const double R[9], const double t[3], const double K[4])
{
    // geometryfusion_mipmap_cpu.cpp:1617-1625
    CamF p;
    const double s = (double)c->scale;
    for (int i = 0; i < 9; i++) p.r[i] = (float)(R[i] / s);
    for (int i = 0; i < 3; i++) p.t[i] = (float)((t[i] - (double)c->offset[i]) / s);
    p.fx = (float)K[0]; p.fy = (float)K[1]; p.cx = (float)K[2]; p.cy = (float)K[3];
    return p;
}

static CamF make_pinv(const ff_ctx *c, const double R[9], const double t[3], const double K[4])
{
    // geometryfusion_mipmap_cpu.cpp:1771-1793 with CameraInfo::getExtrinsicInverse
    // (camerautils.cpp:153-160)
    CamF q;
    double Rt[9], tt[3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Rt[i * 3 + j] = R[j * 3 + i];
    for (int i = 0; i < 3; i++)
        tt[i] = ((-Rt[i * 3 + 0]) * t[0] + (-Rt[i * 3 + 1]) * t[1]) + (-Rt[i * 3 + 2]) * t[2];
    for (int i = 0; i < 9; i++) q.r[i] = (float)Rt[i];
    const double o0 = c->offset[0], o1 = c->offset[1], o2 = c->offset[2];
    for (int i = 0; i < 3; i++)
        q.t[i] = (float)(((Rt[i * 3 + 0] * o0 + Rt[i * 3 + 1] * o1) + Rt[i * 3 + 2] * o2) + tt[i]);
    q.fx = (float)K[0]; q.fy = (float)K[1]; q.cx = (float)K[2]; q.cy = (float)K[3];
    return q;
}

CamF camera_inverse(const ff_ctx *c, const double R[9], const double t[3], const double K[4])
{
    return make_pinv(c, R, t, K);
}

// the dynamic LDS of the de-dup sweeps (their ring of diagonal bands)
static hipError_t dedup_attributes()
{
    hipError_t e = hipSuccess;
#define FF_DD_ATTR(K, WM)                                                                                  \
    if (e == hipSuccess)                                                                                   \
        e = hipFuncSetAttribute((const void *)K<WM>, hipFuncAttributeMaxDynamicSharedMemorySize,         \
                                (int)dedup_lds_bytes(WM));
    FF_DD_ATTR(k_dedup, 20) FF_DD_ATTR(k_dedup, 40) FF_DD_ATTR(k_dedup, 64)
    FF_DD_ATTR(k_dedup_b, 20) FF_DD_ATTR(k_dedup_b, 40) FF_DD_ATTR(k_dedup_b, 64)
#undef FF_DD_ATTR
    return e;
}

// ---------------------------------------------------------------------------
// The per-device scratch array of ff_add_map_device_noalloc: a frame's
// image-sized buffers (bit planes, processed bitmap, the count scan's
// per-pixel arrays, the per-pixel parent record, the confidences) carved
// from one file-scope array instead of the context's own allocations.  In
// host code `scratch` names a host-side stand-in; its device address on
// each device comes from hipGetSymbolAddress, taken once per device with
// that device current (at the first context created on it).
// ---------------------------------------------------------------------------
#define FF_SCRATCH_BYTES (40u << 20)  // 40 MiB: images up to 1280 x 960 (about 32 B per pixel)
#define FF_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[FF_SCRATCH_BYTES];
static void *g_scratch[FF_MAX_DEVICES];
static std::mutex g_scratch_mu;

static hipError_t scratch_register(int device)
{
    if (device < 0 || device >= FF_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));  // the caller made `device` current
    if (e == hipSuccess) g_scratch[device] = ptr;
    return e;
}

static void *scratch_ptr(int device)
{
    if (device < 0 || device >= FF_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch[device];
}

// carves a W x H frame's buffers from the scratch array at base; returns the
// bytes used (more than FF_SCRATCH_BYTES: the image does not fit)
struct ScratchLayout {
    uint32_t *planes, *pbits, *cntB, *cntL, *firstD, *pl0;
    uint64_t *pref, *bsum;
    float *conf;
};
static size_t scratch_layout(unsigned char *base, int W, int H, ScratchLayout *L)
{
    const size_t npix = (size_t)W * H;
    const size_t words = (size_t)(W + 31) / 32;
    const size_t sizes[9] = {sizeof(uint4) * words * (size_t)dedup_plane_rows(H), pbits_bytes(H),
                             sizeof(uint32_t) * npix, sizeof(uint32_t) * npix, sizeof(uint32_t) * npix,
                             sizeof(uint32_t) * 2 * npix, sizeof(uint64_t) * npix,
                             sizeof(uint64_t) * ((npix + 4095) / 4096 + 1), sizeof(float) * npix};
    unsigned char *p[9];
    size_t off = 0;
    for (int k = 0; k < 9; k++) {
        p[k] = base + off;
        off += (sizes[k] + 255) & ~(size_t)255;
    }
    if (L) {
        L->planes = reinterpret_cast<uint32_t *>(p[0]);
        L->pbits = reinterpret_cast<uint32_t *>(p[1]);
        L->cntB = reinterpret_cast<uint32_t *>(p[2]);
        L->cntL = reinterpret_cast<uint32_t *>(p[3]);
        L->firstD = reinterpret_cast<uint32_t *>(p[4]);
        L->pl0 = reinterpret_cast<uint32_t *>(p[5]);
        L->pref = reinterpret_cast<uint64_t *>(p[6]);
        L->bsum = reinterpret_cast<uint64_t *>(p[7]);
        L->conf = reinterpret_cast<float *>(p[8]);
    }
    return off;
}

int ensure_image(ff_ctx *c, int W, int H)
{
    if (W <= 0 || H <= 0 || W > FF_MAX_W || H > FF_MAX_H) {
        c->err = "image width and height must be in [1, 2048]";
        return FF_E_ARG;
    }
    if (c->capW >= W && c->capH >= H) return FF_OK;
    const size_t npix = (size_t)W * H;
    const int words = (W + 31) / 32;
    frame_join(c);  // the frame in flight reads the buffers freed here
    c->freeFrameBuffers();
    HIPCHK(hipMalloc(&c->d_depth, npix * 2));
    HIPCHK(hipMalloc(&c->d_bgr, npix * 3));
    HIPCHK(hipMalloc(&c->d_planes, sizeof(uint4) * words * (size_t)dedup_plane_rows(H)));
    HIPCHK(hipMalloc(&c->d_pbits, pbits_bytes(H)));
    HIPCHK(hipMalloc(&c->d_cntB, sizeof(uint32_t) * npix));
    HIPCHK(hipMalloc(&c->d_cntL, sizeof(uint32_t) * npix));
    HIPCHK(hipMalloc(&c->d_firstD, sizeof(uint32_t) * npix));
    HIPCHK(hipMalloc(&c->d_pl0, sizeof(uint32_t) * 2 * npix));
    HIPCHK(hipMalloc(&c->d_pref, sizeof(uint64_t) * npix));
    HIPCHK(hipMalloc(&c->d_bsum, sizeof(uint64_t) * ((npix + 4095) / 4096 + 1)));
    HIPCHK(hipMemsetAsync(c->d_pbits, 0, pbits_bytes(H), c->stream));
    c->capW = W;
    c->capH = H;
    c->bufGen++;  // captured frame graphs refer to the old buffers
    return FF_OK;
}


void drop_frame_graphs(ff_ctx *c)
{
    for (int i = 0; i < 2; i++) {
        if (c->frameExec[i]) hipGraphExecDestroy(c->frameExec[i]);
        if (c->frameGraph[i]) hipGraphDestroy(c->frameGraph[i]);
        c->frameExec[i] = nullptr;
        c->frameGraph[i] = nullptr;
    }
}

// writes the host's used counts to the device copy (after a host-side change)
// tree slots [from, capacity) back to empty
void fill_tree_tail(ff_ctx *c, uint32_t from)
{
    const uint32_t cnt = c->nBranchesTotal - from;
    k_fill_u32<<<(cnt + 255) / 256, 256, 0, c->stream>>>(c->d_tree + from, FF_BRANCHINIT, cnt);
    hipStreamSynchronize(c->stream);
}

void sync_map_counters(ff_ctx *c)
{
    c->h_mapc->nBranchesUsed = c->nBranchesUsed;
    c->h_mapc->nLeavesUsed = c->nLeavesUsed;
    c->h_mapc->pad[0] = c->h_mapc->pad[1] = 0;
    hipMemcpyAsync(c->d_mapc, c->h_mapc, sizeof(MapCounters), hipMemcpyHostToDevice, c->stream);
    hipStreamSynchronize(c->stream);  // h_mapc is reused
}

static void launch_grow(ff_ctx *c, const int bmin[3], const int bmax[3])
{
    // grow (geometryfusion_mipmap_cpu.cpp:3071-3319), device part
    auto hp2 = [](uint16_t v) {
        uint16_t r = 1;
        while (r * 2 <= v) r *= 2;
        return r;
    };
    const int n = c->n;
    uint16_t p[3], mm[3];
    for (int k = 0; k < 3; k++) {
        p[k] = (uint16_t)((bmax[k] >= n) * ((bmax[k] - n) / n + 1));
        mm[k] = (uint16_t)((bmin[k] < 0) * ((-bmin[k]) / n + 1));
    }
    for (int k = 0; k < 3; k++) {
        const uint16_t cut = p[k] & mm[k];
        if (cut) {
            if (mm[k] > p[k]) mm[k] = (uint16_t)(2 * hp2(mm[k]) + (mm[k] & (uint16_t)~cut));
            else p[k] = (uint16_t)(2 * hp2(p[k]) + (p[k] & (uint16_t)~cut));
        }
    }
    uint16_t maxAdd = 0;
    for (int k = 0; k < 3; k++) {
        maxAdd = std::max(maxAdd, hp2(mm[k]));
        maxAdd = std::max(maxAdd, hp2(p[k]));
    }
    const uint32_t add = (uint32_t)maxAdd * 16u;
    if (add > c->nBranchesTotal - c->nBranchesUsed) return;
    const uint32_t used = c->nBranchesUsed;
    const uint32_t nb = (used / 16) * 8;
    k_shift_children<<<(nb + 255) / 256, 256, 0, c->stream>>>(c->d_tree, used, add);
    // move [0, used) to [add, add + used) through a scratch copy
    k_copy_u32<<<(used + 255) / 256, 256, 0, c->stream>>>(c->d_tree, c->d_treeTmp, used);
    k_copy_u32<<<(used + 255) / 256, 256, 0, c->stream>>>(c->d_treeTmp, c->d_tree + add, used);
    k_fill_u32<<<(add + 255) / 256, 256, 0, c->stream>>>(c->d_tree, FF_BRANCHINIT, add);
    c->nBranchesUsed += add;
    sync_map_counters(c);
    uint32_t index = (maxAdd - 1) * 16u;
    uint16_t addv[3] = {0, 0, 0};
    std::vector<std::pair<uint32_t, uint32_t>> links;
    for (uint16_t level = 0; level < maxAdd; level++) {
        const unsigned dec = 1u << level;
        uint32_t co = 0;
        for (int k = 0; k < 3; k++)
            if ((mm[k] & dec) || (!(p[k] & dec) && (p[k] > mm[k]))) {
                addv[k] += c->n;
                co += 1u << k;
            }
        links.push_back({index + co + 8, index + 16});
        c->n *= 2;
        index -= 16;
    }
    for (auto &l : links)
        hipMemcpyAsync(c->d_tree + l.first, &l.second, sizeof(uint32_t), hipMemcpyHostToDevice, c->stream);
    hipStreamSynchronize(c->stream);
    {
        // node geometry of the new root chain: the branch added at level L has
        // octant size n_old 2^L; origins accumulate from the new root down
        std::vector<ushort4> chain(maxAdd);
        uint32_t org[3] = {0, 0, 0};
        for (int L = (int)maxAdd - 1; L >= 0; L--) {
            const uint32_t octant = (uint32_t)(c->n >> (maxAdd - L));
            chain[maxAdd - 1 - L] = make_ushort4((unsigned short)org[0], (unsigned short)org[1], (unsigned short)org[2],
                                                 (unsigned short)octant);
            const uint32_t co = (uint32_t)(links[L].first - (maxAdd - 1 - L) * 16u - 8u);
            for (int k = 0; k < 3; k++) org[k] += octant * ((co >> k) & 1u);
        }
        const uint32_t addv32[3] = {addv[0], addv[1], addv[2]};
        mesh_on_grow_blocks(c, used / 16, add / 16, addv32, chain);
    }
    for (int k = 0; k < 3; k++) c->offset[k] -= (float)addv[k] * c->scale;
    const uint32_t nl = std::min(c->nLeavesUsed, c->nLeavesTotal);
    if (nl) k_shift_leaves<<<(nl + 255) / 256, 256, 0, c->stream>>>(c->d_leafinfo, nl, addv[0], addv[1], addv[2]);
    c->grows++;
}

}  // namespace ff

using namespace ff;

void ff_ctx::freeFrameBuffers()
{
    hipFree(d_depth); hipFree(d_bgr); hipFree(d_planes); hipFree(d_pbits);
    hipFree(d_cntB); hipFree(d_cntL); hipFree(d_firstD); hipFree(d_pl0); hipFree(d_pref); hipFree(d_bsum);
    d_depth = nullptr; d_bgr = nullptr; d_planes = d_pbits = nullptr;
    d_cntB = d_cntL = d_firstD = d_pl0 = nullptr; d_pref = d_bsum = nullptr;
    capW = capH = 0;
}

extern "C" {

ff_ctx *ff_create(const ff_params *prm)
{
    if (!prm || !(prm->scale > 0.0f) || !(prm->distance_threshold > 0.0f)) return nullptr;
    ff_ctx *c = new ff_ctx();
    c->scale = prm->scale;
    c->dT = prm->distance_threshold;
    c->useColor = prm->use_color != 0;
    c->incremental = prm->incremental_meshing != 0;
    c->device = prm->device;
    // _bandwidth(_distanceThreshold / _scale*BANDWIDTHFACTOR) (ctor, :425)
    c->bandwidth = (uint16_t)((double)(prm->distance_threshold / prm->scale) * 1.0);
    c->nLeavesTotal = prm->max_leaves ? prm->max_leaves : (1u << 20);
    c->nBranchesTotal = prm->max_branches ? prm->max_branches : (1u << 25);
    c->nBranchesUsed = 16;
    if (hipSetDevice(c->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return nullptr;
    }
    int waitValue = 0;
    c->seqOk = hipDeviceGetAttribute(&waitValue, hipDeviceAttributeCanUseStreamWaitValue, c->device) == hipSuccess &&
               waitValue != 0 && hipMalloc(&c->d_seq, sizeof(c->seq)) == hipSuccess &&
               hipMemset(c->d_seq, 0, sizeof(c->seq)) == hipSuccess;
    if (!c->seqOk || hipStreamCreate(&c->side) != hipSuccess) {
        // the frame then runs on the main stream
        if (c->side) hipStreamDestroy(c->side);
        c->side = nullptr;
        (void)hipGetLastError();
    }
    if (c->side && hipStreamCreate(&c->aux) != hipSuccess) {
        // the confidence then runs on the side stream with the frame
        c->aux = nullptr;
        (void)hipGetLastError();
    }
    const size_t NL = c->nLeavesTotal;
    bool ok = hipMalloc(&c->d_tree, sizeof(uint32_t) * c->nBranchesTotal) == hipSuccess &&
              hipMalloc(&c->d_treeTmp, sizeof(uint32_t) * c->nBranchesTotal) == hipSuccess &&
              hipMalloc(&c->d_leafinfo, sizeof(ushort4) * NL) == hipSuccess &&
              hipMalloc(&c->d_parent, sizeof(uint32_t) * NL) == hipSuccess &&
              hipMalloc(&c->d_stamp, sizeof(uint32_t) * NL) == hipSuccess &&
              hipMalloc(&c->d_queue, sizeof(uint32_t) * NL) == hipSuccess &&
              hipMalloc(&c->d_dist, sizeof(float) * NL * FF_BRICK) == hipSuccess &&
              hipMalloc(&c->d_wgt, sizeof(float) * NL * FF_BRICK) == hipSuccess &&
              hipMalloc(&c->d_fc, sizeof(FrameCounters)) == hipSuccess &&
              hipHostMalloc(&c->h_fc, sizeof(FrameCounters), hipHostMallocMapped | hipHostMallocCoherent) ==
                  hipSuccess &&
              hipHostGetDevicePointer(reinterpret_cast<void **>(&c->d_hfc), c->h_fc, 0) == hipSuccess &&
              hipMalloc(&c->d_fd, sizeof(FrameDev)) == hipSuccess &&
              hipHostMalloc(&c->h_fd, sizeof(FrameDev), hipHostMallocDefault) == hipSuccess &&
              hipMalloc(&c->d_mapc, sizeof(MapCounters)) == hipSuccess &&
              hipHostMalloc(&c->h_mapc, sizeof(MapCounters), hipHostMallocDefault) == hipSuccess &&
              hipMalloc(&c->d_pxp, sizeof(float) * FF_MAX_W) == hipSuccess &&
              hipMalloc(&c->d_pyp, sizeof(float) * FF_MAX_H) == hipSuccess &&
              dedup_attributes() == hipSuccess && scratch_register(c->device) == hipSuccess;
    if (ok && c->useColor) ok = hipMalloc(&c->d_col, sizeof(uint16_t) * NL * FF_BRICK * 3) == hipSuccess;
    const uint32_t hcap = 1u << 22;
    ok = ok && hipMalloc(&c->hv.keys, sizeof(uint64_t) * hcap) == hipSuccess &&
         hipMalloc(&c->hv.minpix, sizeof(uint32_t) * hcap) == hipSuccess &&
         hipMalloc(&c->hv.aux, sizeof(uint32_t) * hcap) == hipSuccess &&
         hipMalloc(&c->hv.assigned, sizeof(uint32_t) * hcap) == hipSuccess &&
         hipMalloc(&c->hv.list, sizeof(uint32_t) * (hcap / 2)) == hipSuccess;
    if (!ok) {
        ff_destroy(c);
        return nullptr;
    }
    memset(c->h_fc, 0, sizeof(FrameCounters));
    c->hv.mask = hcap - 1;
    c->hv.listCap = hcap / 2;
    hipMemsetAsync(c->hv.keys, 0xff, sizeof(uint64_t) * hcap, c->stream);
    hipMemsetAsync(c->hv.minpix, 0xff, sizeof(uint32_t) * hcap, c->stream);
    hipMemsetAsync(c->d_tree, 0xff, sizeof(uint32_t) * c->nBranchesTotal, c->stream);  // then fix to BRANCHINIT
    k_fill_u32<<<(c->nBranchesTotal + 255) / 256, 256, 0, c->stream>>>(c->d_tree, FF_BRANCHINIT, c->nBranchesTotal);
    hipMemsetAsync(c->d_stamp, 0, sizeof(uint32_t) * NL, c->stream);
    hipMemsetAsync(c->d_dist, 0, sizeof(float) * NL * FF_BRICK, c->stream);
    hipMemsetAsync(c->d_wgt, 0, sizeof(float) * NL * FF_BRICK, c->stream);
    if (c->d_col) hipMemsetAsync(c->d_col, 0, sizeof(uint16_t) * NL * FF_BRICK * 3, c->stream);
    hipMemcpyToSymbolAsync(HIP_SYMBOL(c_rcp_table), FF_RCP_TABLE, sizeof(FF_RCP_TABLE), 0, hipMemcpyHostToDevice,
                           c->stream);
    memset(c->h_fd, 0, sizeof(FrameDev));
    sync_map_counters(c);
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        ff_destroy(c);
        return nullptr;
    }
    if (mesh_init(c) != FF_OK) {
        ff_destroy(c);
        return nullptr;
    }
    return c;
}

void ff_destroy(ff_ctx *c)
{
    if (!c) return;
    hipSetDevice(c->device);
    // every stream drained before anything is freed (a frame's SDF update
    // may still run on the side stream)
    if (c->side) hipStreamSynchronize(c->side);
    if (c->aux) hipStreamSynchronize(c->aux);
    if (c->stream) hipStreamSynchronize(c->stream);
    batch_free(c);
    mesh_free(c);
    learned_free(c);
    c->freeFrameBuffers();
    hipFree(c->d_tree); hipFree(c->d_treeTmp); hipFree(c->d_leafinfo); hipFree(c->d_parent);
    hipFree(c->d_stamp); hipFree(c->d_queue); hipFree(c->d_dist); hipFree(c->d_wgt); hipFree(c->d_col);
    hipFree(c->d_fc); hipHostFree(c->h_fc);
    hipFree(c->d_fd); hipHostFree(c->h_fd);
    hipFree(c->d_mapc); hipHostFree(c->h_mapc);
    hipFree(c->d_pxp); hipFree(c->d_pyp);
    drop_frame_graphs(c);
    hipFree(c->hv.keys); hipFree(c->hv.minpix); hipFree(c->hv.aux); hipFree(c->hv.assigned); hipFree(c->hv.list);
    for (int i = 0; i < 10; i++)
        if (c->ev[i]) hipEventDestroy(c->ev[i]);
    if (c->side) {
        hipStreamSynchronize(c->side);
        hipStreamDestroy(c->side);
    }
    hipFree(c->d_seq);
    if (c->aux) {
        hipStreamSynchronize(c->aux);
        hipStreamDestroy(c->aux);
    }
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

const char *ff_last_error(const ff_ctx *c) { return c ? c->err.c_str() : "null context"; }

void *ff_stream(ff_ctx *c)
{
    if (!c) return nullptr;
    frame_join(c, true);  // work the caller enqueues on it sees the whole frame
    return (void *)c->stream;
}

}  // extern "C"

namespace ff {

// A frame's device work reads every per-frame value from c->d_fd, so that
// it can be captured once and replayed as graphs: the allocation (ending
// with the commit of the used counts), then the learned confidence when on,
// the gated SDF update and the counters' copy back.
void launch_front(ff_ctx *c, int W, int H, hipStream_t st, bool carry)
{
    const int words = (W + 31) / 32;
    if (carry)
        k_planes_v<<<dim3((W + 63) / 64, H), 64, 0, st>>>(*c->h_fd, c->d_fd, W, H, words, c->d_pxp, c->d_pyp,
                                                          c->d_planes, c->d_fc);
    else
        k_planes<<<dim3((W + 63) / 64, H), 64, 0, st>>>(c->d_fd, W, H, words, c->d_pxp, c->d_pyp, c->d_planes,
                                                        c->d_fc);
    const size_t lds = dedup_lds_bytes(words);
    switch (dedup_wmax(words)) {
    case 20: k_dedup<20><<<1, FF_DD_THREADS, lds, st>>>(c->d_planes, H, words, c->d_pbits, c->d_fc); break;
    case 40: k_dedup<40><<<1, FF_DD_THREADS, lds, st>>>(c->d_planes, H, words, c->d_pbits, c->d_fc); break;
    default: k_dedup<64><<<1, FF_DD_THREADS, lds, st>>>(c->d_planes, H, words, c->d_pbits, c->d_fc); break;
    }
}

void launch_alloc(ff_ctx *c, int W, int H)
{
    launch_front(c, W, H, c->stream, false);
    launch_back(c, W, H, c->stream);
}

void launch_back(ff_ctx *c, int W, int H, hipStream_t st)
{
    const int words = (W + 31) / 32;
    const FrameDev *fd = c->d_fd;
    const MapView mv = c->mapView();
    k_descend<<<dim3((W + 127) / 128, H), 128, 0, st>>>(fd, W, H, words, c->d_pxp, c->d_pyp, c->d_pbits, mv, c->hv,
                                                       c->d_pl0, c->d_fc);
    ushort4 *binfo;
    uint32_t binfoBlocks;
    mesh_binfo(c, &binfo, &binfoBlocks);
    // the tail in one block; the queue pass with it while the map (with
    // this frame's new bricks) is a few spans
    const bool small = (uint64_t)c->nLeavesUsed + FF_ORDER_MAX <= (uint64_t)FF_TAIL_SPANS * FF_QUEUE_SPAN;
    k_alloc_tail<<<1, FF_ORDER_THREADS, 0, st>>>(c->hv, mv, c->d_fc, c->d_pl0, fd, c->d_mapc, binfo, binfoBlocks,
                                                 small ? 1 : 0, H, small ? c->d_hfc : nullptr);
    if (!small) {
        k_queue<<<queue_grid(c->nLeavesUsed), FF_QUEUE_THREADS, 0, st>>>(c->d_stamp, c->d_mapc, c->nLeavesTotal, fd,
                                                                         c->d_queue, c->d_fc, 0);
        k_frame_report<<<1, 64, 0, st>>>(c->d_fc, c->d_hfc, fd);
    }
}

// A frame with more new nodes than k_order ranks (flagged; its link, commit
// and SDF update stood down): order them through the per-pixel count scan,
// then link and commit, and bring the counters back.
static int order_big_frame(ff_ctx *c, int W, int H)
{
    const uint32_t npix = (uint32_t)W * H;
    const FrameDev *fd = c->d_fd;
    HIPCHK(hipMemsetAsync(c->d_cntB, 0, sizeof(uint32_t) * npix, c->stream));
    HIPCHK(hipMemsetAsync(c->d_cntL, 0, sizeof(uint32_t) * npix, c->stream));
    HIPCHK(hipMemsetAsync(c->d_firstD, 0xff, sizeof(uint32_t) * npix, c->stream));
    k_resolve<<<256, 256, 0, c->stream>>>(c->hv, c->d_fc, c->d_cntB, c->d_cntL, c->d_firstD);
    const uint32_t nb = (npix + 4095) / 4096;
    k_scan_a<<<nb, 1024, 0, c->stream>>>(c->d_cntB, c->d_cntL, npix, c->d_pref, c->d_bsum);
    k_scan_b<<<1, 1024, 0, c->stream>>>(c->d_bsum, nb, c->d_fc);
    k_assign<<<256, 256, 0, c->stream>>>(c->hv, c->d_fc, c->d_pref, c->d_bsum, c->d_firstD, c->d_mapc);
    k_link<<<256, 256, 0, c->stream>>>(c->hv, c->mapView(), c->d_fc, c->d_pl0, fd, c->d_fc, 1);
    ushort4 *binfo;
    uint32_t binfoBlocks;
    mesh_binfo(c, &binfo, &binfoBlocks);
    k_alloc_end<<<256, 256, 0, c->stream>>>(c->hv, c->d_fc, c->d_mapc, binfo, binfoBlocks, fd, 1);
    k_queue<<<queue_grid(c->nLeavesUsed + c->hv.listCap), FF_QUEUE_THREADS, 0, c->stream>>>(
        c->d_stamp, c->d_mapc, c->nLeavesTotal, fd, c->d_queue, c->d_fc, 1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->h_fc, c->d_fc, sizeof(FrameCounters), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return FF_OK;
}

AllocStream alloc_stream(ff_ctx *c, int W, int H)
{
    AllocStream a;
    a.fd = c->d_fd;
    a.pxp = c->d_pxp; a.pyp = c->d_pyp;
    a.fc = c->d_fc;
    a.planes = c->d_planes; a.cntB = c->d_cntB; a.cntL = c->d_cntL; a.firstD = c->d_firstD;
    a.pbits = c->d_pbits; a.pl0 = c->d_pl0;
    c->lastPbits = c->d_pbits;
    a.pref = c->d_pref; a.bsum = c->d_bsum;
    a.mapc = c->d_mapc;
    mesh_binfo(c, &a.binfo, &a.binfoCap);
    a.m = c->mapView();
    a.hv = c->hv;
    a.W = W; a.H = H; a.words = (W + 31) / 32;
    a.nLeavesUsed = c->nLeavesUsed;
    return a;
}

// launch_alloc for n streams at once: every kernel one launch, grid z = stream
void launch_alloc_batch(hipStream_t st, const AllocStream *d_as, const AllocStream *h_as, int n)
{
    int maxW = 1, maxH = 1;
    uint32_t maxLeaves = 0;
    bool cls[3] = {false, false, false};
    for (int k = 0; k < n; k++) {
        const AllocStream &a = h_as[k];
        maxW = std::max(maxW, a.W);
        maxH = std::max(maxH, a.H);
        maxLeaves = std::max(maxLeaves, a.nLeavesUsed);
        const int wm = dedup_wmax(a.words);
        cls[wm == 20 ? 0 : (wm == 40 ? 1 : 2)] = true;
    }
    const int G = 64;  // blocks per stream of the node-list kernels
    k_planes_b<<<dim3((maxW + 63) / 64, maxH, n), 64, 0, st>>>(d_as);
    if (cls[0]) k_dedup_b<20><<<dim3(1, 1, n), FF_DD_THREADS, dedup_lds_bytes(20), st>>>(d_as);
    if (cls[1]) k_dedup_b<40><<<dim3(1, 1, n), FF_DD_THREADS, dedup_lds_bytes(40), st>>>(d_as);
    if (cls[2]) k_dedup_b<64><<<dim3(1, 1, n), FF_DD_THREADS, dedup_lds_bytes(64), st>>>(d_as);
    k_descend_b<<<dim3((maxW + 127) / 128, maxH, n), 128, 0, st>>>(d_as);
    k_order_b<<<dim3(1, 1, n), FF_ORDER_THREADS, 0, st>>>(d_as);
    k_link_b<<<dim3(G, 1, n), 256, 0, st>>>(d_as);
    k_alloc_end_b<<<dim3(G, 1, n), 256, 0, st>>>(d_as);
    k_queue_b<<<dim3(queue_grid(maxLeaves), 1, n), FF_QUEUE_THREADS, 0, st>>>(d_as);
}

// the SDF update waits for the host only when the node table overflowed or
// the volume must grow first (or, following a mesh update's sampling, when
// that update's marching cubes read bricks: tiles); direct: a direct launch,
// which with timing on has the dispatch itself record the SDF phase's events
// (they then time the kernel alone)
static void launch_sdf_only(ff_ctx *c, int W, int H, bool direct, hipStream_t st, const uint32_t *tiles = nullptr,
                            uint32_t tileCap = 0)
{
    SdfGate gate;
    gate.listCap = c->hv.listCap;
    gate.leavesTotal = c->nLeavesTotal;
    gate.branchesTotal = c->nBranchesTotal;
    gate.force = 0;
    // (timed on one frame in timeEvery)
    const bool timedNow = direct && c->timing && c->frames % c->timeEvery == 0;
    if (direct) c->sdfTimed = timedNow;
    if (timedNow)
        hipExtLaunchKernelGGL(k_sdf, dim3(FF_SDF_GRID), dim3(512), 0, st, c->ev[2], c->ev[3], 0,
                              (const uint32_t *)c->d_queue, c->mapView(), (const FrameDev *)c->d_fd, W, H, c->d_fc,
                              (const MapCounters *)c->d_mapc, gate, tiles, tileCap);
    else
        k_sdf<<<FF_SDF_GRID, 512, 0, st>>>(c->d_queue, c->mapView(), c->d_fd, W, H, c->d_fc, c->d_mapc, gate, tiles,
                                           tileCap);
}

// (the learned confidence first when on; the counters' copy back follows
// outside, after the timing event)
void launch_sdf(ff_ctx *c, int W, int H)
{
    if (c->learned) learned_launch(c, W, H, c->stream);
    launch_sdf_only(c, W, H, false, c->stream);
}

static bool capture(ff_ctx *c, void (*fn)(ff_ctx *, int, int), int W, int H, hipGraph_t *g, hipGraphExec_t *x)
{
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        fn(c, W, H);
        e = hipStreamEndCapture(c->stream, g);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(x, *g, nullptr, nullptr, 0);
    if (e != hipSuccess) (void)hipGetLastError();
    return e == hipSuccess;
}

// the frame's two launch sequences, from graphs captured for this
// configuration (timing events are recorded between them)
// whether a frame's allocation phase may run on the side stream: direct
// launches (a captured frame replays on one stream), no per-phase events
// (they would time the overlap), an existing volume and image tables of
// this size (the first frame and a new size prepare on the main stream)
static bool front_overlap(const ff_ctx *c, int W, int H)
{
    return c->side && c->seqOk && !c->useGraphs && !c->timingAll && c->n && c->imgW == W &&
           c->imgH == H;
}

static int run_frame(ff_ctx *c, int W, int H, bool overlap)
{
    if (overlap) {
        // the frame on the side stream, beside the mesh update still in
        // flight on the main stream: the front at once, the descent beside
        // the update's partition and marking (which read the tree as counted
        // at their launch), the SDF update once the update's sampling is
        // done (its marching cubes may still run); the update's hash
        // turnover follows the SDF update on the side stream
        // the learned confidence reads only the image: it starts at once on
        // the third stream, beside the de-dup and the descent
        bool confAside = c->learned && c->aux;
        if (confAside) {
            // (the previous frame's SDF update, which reads its
            // confidences, may still run: a wait when they share a buffer)
            if (c->lastConf == c->prevConf) seq_wait(c, FF_SEQ_SDF, c->aux);
            // the frame's host images go up on the side stream first
            if (c->uploadOnSide && seq_signal(c, FF_SEQ_UP, c->side)) seq_wait(c, FF_SEQ_UP, c->aux);
            learned_launch_direct(c, W, H, c->aux);
            confAside = seq_signal(c, FF_SEQ_CONF, c->aux);  // (else aux has drained)
        }
        launch_front(c, W, H, c->side, true);  // carries the frame's parameters up
        // a requested mesh update (of the map before this frame) goes on the
        // main stream now: the front reads only the new frame
        c->frameInFlush = true;
        const int rf = mesh_flush(c);
        c->frameInFlush = false;
        if (rf != FF_OK) return rf;
        launch_back(c, W, H, c->side);
        if (c->learned) {
            if (confAside) seq_wait(c, FF_SEQ_CONF, c->side);
            else if (!c->aux) learned_launch(c, W, H, c->side);
        }
        const uint32_t *tiles = nullptr;
        c->sdfTileId = 0;
        if (c->samplePending) {
            seq_wait(c, FF_SEQ_SAMPLE, c->side);
            tiles = c->pendingTiles;
            c->sdfTileId = c->pendingId;
            c->sdfTileCap = c->pendingTileCap;
            c->samplePending = false;
        }
        launch_sdf_only(c, W, H, true, c->side, tiles, c->sdfTileCap);
        mesh_launch_deferred_turnover(c, c->side);
        c->sdfInFlight = seq_signal(c, FF_SEQ_SDF, c->side);  // (else the side stream has drained)
        HIPCHK(hipGetLastError());
        return FF_OK;
    }
    const uint64_t key = ((uint64_t)W << 48) ^ ((uint64_t)H << 32) ^ ((uint64_t)c->bufGen << 1) ^
                         (uint64_t)(c->learned != nullptr);
    const bool graphs = c->useGraphs && !c->directOnly;
    if (graphs && (!c->frameExec[0] || c->frameKey != key)) {
        drop_frame_graphs(c);
        if (capture(c, launch_alloc, W, H, &c->frameGraph[0], &c->frameExec[0]) &&
            capture(c, launch_sdf, W, H, &c->frameGraph[1], &c->frameExec[1])) {
            c->frameKey = key;
        } else {
            drop_frame_graphs(c);  // capture is an optimisation: launch directly from now on
            c->useGraphs = false;
        }
    }
    const bool g = graphs && c->useGraphs && c->frameExec[0];
    if (c->timingAll) hipEventRecord(c->ev[0], c->stream);
    if (g) HIPCHK(hipGraphLaunch(c->frameExec[0], c->stream));
    else launch_alloc(c, W, H);
    if (c->timingAll) hipEventRecord(c->ev[1], c->stream);
    // the SDF phase's events bracket k_sdf alone when launched directly
    if (!g && c->learned) learned_launch(c, W, H, c->stream);
    if (g) {
        if (c->timing) hipEventRecord(c->ev[2], c->stream);
        HIPCHK(hipGraphLaunch(c->frameExec[1], c->stream));
        if (c->timing) hipEventRecord(c->ev[3], c->stream);
        c->sdfTimed = c->timing;
    } else {
        launch_sdf_only(c, W, H, true, c->stream);  // which leaves the counters in h_fc
    }
    HIPCHK(hipGetLastError());
    return FF_OK;
}

// Host part of a frame up to the upload of its parameters (first-frame
// volume, column/row tables, pose).  depth / bgr: the frame in device
// memory, dense rows.
int frame_prepare(ff_ctx *c, const uint16_t *depth, const uint8_t *bgr, int W, int H, const double R[9],
                  const double t[3], const double K[4], float scaling, float maxcam, bool upload, bool overlap)
{
    if (c->broken) {
        c->err = "a frame overflowed the per-frame node table: the context is unusable";
        return FF_E_STATE;
    }
    CamF p = make_p(c, R, t, K);
    FrameCounters *hf = c->h_fc;
    const uint32_t stamp = ++c->stampCounter;
    if (!c->n) {
        // first frame: initial volume (setInitialVolume, :1186-1202)
        memset(hf, 0, sizeof(FrameCounters));
        for (int k = 0; k < 3; k++) { hf->bmin[k] = 1000000; hf->bmax[k] = -1000000; }
        HIPCHK(hipMemcpyAsync(c->d_fc, hf, sizeof(FrameCounters), hipMemcpyHostToDevice, c->stream));
        k_bbox_first<<<dim3((W + 255) / 256, H), 256, 0, c->stream>>>(depth, W, H, scaling, maxcam, p,
                                                                       c->bandwidth, c->d_fc);
        HIPCHK(hipMemcpyAsync(hf, c->d_fc, sizeof(FrameCounters), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        const int *mn = hf->bmin, *mx = hf->bmax;
        if (!(mn[0] >= mx[0] || mn[1] >= mx[1] || mn[2] >= mx[2])) {
            uint16_t hi[3];
            for (int k = 0; k < 3; k++) hi[k] = (uint16_t)(mx[k] - mn[k]);
            const uint16_t high = std::max(std::max(hi[0], hi[1]), hi[2]);
            uint16_t n;
            for (n = 2 * FF_BL; n <= high; n *= 2) ;
            c->n = n;
            for (int k = 0; k < 3; k++) c->offset[k] = (float)(mn[k] - (n - hi[k]) / 2) * c->scale;
            mesh_on_initial_volume(c);
        }
        p = make_p(c, R, t, K);
    }
    if (c->imgW != W || c->imgH != H) {
        // column/row tables built with the frame that introduced the size (:1745-1764)
        c->imgW = W; c->imgH = H;
        c->pxp.resize(W); c->pyp.resize(H);
        for (int x = 0; x < W; x++) c->pxp[x] = ((float)x - p.cx) / p.fx;
        for (int y = 0; y < H; y++) c->pyp[y] = ((float)y - p.cy) / p.fy;
        HIPCHK(hipMemcpyAsync(c->d_pxp, c->pxp.data(), sizeof(float) * W, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->d_pyp, c->pyp.data(), sizeof(float) * H, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));  // the host tables are reused
    }
    const CamF q = make_pinv(c, R, t, K);
    // the frame's parameters, staged in pinned memory for the device
    FrameDev &f = *c->h_fd;
    f.p = p;
    for (int i = 0; i < 3; i++) {
        f.sp.m[4 * i] = q.r[3 * i]; f.sp.m[4 * i + 1] = q.r[3 * i + 1]; f.sp.m[4 * i + 2] = q.r[3 * i + 2];
        f.sp.m[4 * i + 3] = q.t[i];
    }
    f.sp.fx = q.fx; f.sp.fy = q.fy; f.sp.cx = q.cx; f.sp.cy = q.cy;
    f.sp.scale = c->scale; f.sp.dT = c->dT; f.sp.scaling = scaling; f.sp.maxcam = maxcam;
    f.sp.useColor = c->useColor ? 1 : 0;
    f.sp.conf = nullptr;
    if (c->learned) {
        f.sp.conf = learned_prepare(c, W, H, overlap);
        if (!f.sp.conf) return FF_E_DEVICE;
    }
    c->prevConf = c->lastConf;
    c->lastConf = f.sp.conf;
    f.depth = depth;
    f.bgr = bgr;
    f.scaling = scaling;
    f.maxcam = maxcam;
    f.ifx = 1.0f / (float)K[0];
    f.ify = 1.0f / (float)K[1];
    f.stamp = stamp;
    f.n = c->n;
    f.bandwidth = c->bandwidth;
    f.pad = 0;
    if (upload) HIPCHK(hipMemcpyAsync(c->d_fd, c->h_fd, sizeof(FrameDev), hipMemcpyHostToDevice, c->stream));
    return FF_OK;
}

// the device work of a frame without a volume (an empty first frame)
void launch_empty(ff_ctx *c, int W, int H)
{
    k_frame_begin<<<1, 64, 0, c->stream>>>(c->d_fd, c->d_fc);
    hipMemsetAsync(c->d_pbits, 0, pbits_bytes(H), c->stream);
    hipMemcpyAsync(c->h_fc, c->d_fc, sizeof(FrameCounters), hipMemcpyDeviceToHost, c->stream);
}

// Host part after the frame's device work completed and c->h_fc holds its
// counters: errors, used counts, grow with the deferred SDF update, bytes.
// algorithmic bytes of one SDF update: 512 voxels x 14 B (8 B without
// color) read and written per queued brick, and the BGR image
static double sdf_bytes(const ff_ctx *c, uint32_t nQueued, int W, int H)
{
    return (double)nQueued * (c->useColor ? 14336.0 : 8192.0) + (c->useColor ? 3.0 * W * H : 0.0);
}

int frame_finish(ff_ctx *c, int W, int H, bool timed)
{
    FrameCounters *hf = c->h_fc;
    if (timed && c->timing && c->n) {
        float ms = 0.0f;
        if (c->timingAll && hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess) {
            c->phaseMs[0] += ms;
            c->phaseCalls[0]++;
        }
        if (c->sdfTimed && hf->sdfRan) {
            // read once the update has ended (at the next frame or join)
            c->sdfTimedPending = true;
            c->sdfTimedBytes = sdf_bytes(c, hf->nQueued, W, H);
            c->sdfTimedUnits = hf->nQueued;
        }
    }
    c->sdfTimed = false;
    if (!hf->sdfRan || (hf->flags & (FF_FLAG_ORDER_BIG | FF_FLAG_HASH_FULL)) || hf->listCount > c->hv.listCap) {
        // the host's turn (an SDF update that stood down, a large or an
        // overflowing frame): the side stream drained first
        frame_join(c);
    }
    if (hf->flags & FF_FLAG_ORDER_BIG) {
        const int r = order_big_frame(c, W, H);
        if (r) return r;
    }
    if ((hf->flags & FF_FLAG_HASH_FULL) || hf->listCount > c->hv.listCap) {
        // the frame's nodes were dropped, but node-table entries beyond the
        // list may remain: the context takes no further frames
        c->broken = true;
        c->err = (hf->flags & FF_FLAG_HASH_FULL) ? "per-frame node table overflow: the context is unusable"
                                                 : "per-frame node list overflow: the context is unusable";
        return FF_E_POOL_FULL;
    }
    // the host mirror of the counts k_commit added on the device
    c->nBranchesUsed += 16u * hf->newB;
    c->nLeavesUsed += hf->newL;
    c->nQueued = hf->nQueued;
    if (hf->flags & FF_FLAG_POOL_FULL) {
        c->err = "branch pool exhausted";
        c->incremental = false;
    }
    // grow (:1874-1877), before the SDF update which keeps this frame's pInv
    if (c->nLeavesUsed < c->nLeavesTotal && c->nBranchesUsed < c->nBranchesTotal &&
        (hf->bmin[0] < 0 || hf->bmin[1] < 0 || hf->bmin[2] < 0 || hf->bmax[0] >= (int)c->n ||
         hf->bmax[1] >= (int)c->n || hf->bmax[2] >= (int)c->n)) {
        int bmin[3] = {hf->bmin[0], hf->bmin[1], hf->bmin[2]}, bmax[3] = {hf->bmax[0], hf->bmax[1], hf->bmax[2]};
        launch_grow(c, bmin, bmax);
        mesh_on_grow(c);
    }
    if (c->nQueued && !hf->sdfRan) {
        // the gated launch stood down (volume grown above): update now with
        // this frame's parameters, still in c->d_fd
        SdfGate gate;
        gate.listCap = c->hv.listCap;
        gate.leavesTotal = c->nLeavesTotal;
        gate.branchesTotal = c->nBranchesTotal;
        gate.force = 1;
        if (c->timing) hipEventRecord(c->ev[4], c->stream);
        k_sdf<<<FF_SDF_GRID, 512, 0, c->stream>>>(c->d_queue, c->mapView(), c->d_fd, W, H, c->d_fc, c->d_mapc, gate,
                                                  nullptr, 0u);
        if (c->timing) hipEventRecord(c->ev[5], c->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        float ms = 0.0f;
        if (c->timing && hipEventElapsedTime(&ms, c->ev[4], c->ev[5]) == hipSuccess) {
            c->phaseMs[1] += ms;
            c->phaseCalls[1]++;
            c->phaseBytes[1] += sdf_bytes(c, c->nQueued, W, H);
            c->phaseUnits[1] += c->nQueued;
        }
    }
    // algorithmic bytes of the allocation: images read once, 16 B per
    // processed pixel (the SDF update's are added with its timed launches;
    // batched ticks account their SDF launch on the lead context instead);
    // counted only while every phase is timed, like the phases' times
    if (timed && c->timingAll) {
        c->phaseBytes[0] += (double)W * H * 2 + 16.0 * hf->nProcessed;
        c->phaseUnits[0] += hf->nProcessed;
    }

    if (c->nLeavesUsed > c->nLeavesTotal || c->nBranchesUsed > c->nBranchesTotal) c->incremental = false;
    c->frames++;
    mesh_after_frame(c);
    return (int)c->nQueued;
}

bool await_word(const uint32_t *w, uint32_t want, hipStream_t st, bool atLeast)
{
    const volatile uint32_t *v = w;
    auto reached = [&]() { const uint32_t x = *v; return atLeast ? (int32_t)(x - want) >= 0 : x == want; };
    for (uint32_t k = 1;; k++) {
        if (reached()) break;
        if ((k & 255u) == 0u) {
            const hipError_t e = hipStreamQuery(st);
            if (e != hipErrorNotReady) {
                (void)hipGetLastError();
                if (!reached()) return false;
                break;
            }
        }
        __builtin_ia32_pause();
    }
    std::atomic_thread_fence(std::memory_order_acquire);  // what was written before the word
    return true;
}

// k_sdf's gate (sdf_gate_open), evaluated on the host from the counters the
// allocation reported: whether the frame's SDF update runs or stands down
// for the host (a node table that overflowed, a frame that must grow the
// volume first)
static bool sdf_gate_host(const ff_ctx *c)
{
    const FrameCounters &f = *c->h_fc;
    const bool held = (f.flags & (FF_FLAG_HASH_FULL | FF_FLAG_ORDER_BIG)) != 0u || f.listCount > c->hv.listCap;
    if (held) return false;
    const int n = (int)c->n;
    const bool out = f.bmin[0] < 0 || f.bmin[1] < 0 || f.bmin[2] < 0 || f.bmax[0] >= n || f.bmax[1] >= n ||
                     f.bmax[2] >= n;
    // the used counts the allocation committed
    const uint64_t nl = (uint64_t)c->nLeavesUsed + f.newL, nb = (uint64_t)c->nBranchesUsed + 16ull * f.newB;
    const bool grows = nl < c->nLeavesTotal && nb < c->nBranchesTotal;
    return !(out && grows);
}

// the SDF phase's time of a timed frame whose update had not ended when the
// frame returned (the caller knows it has ended by now)
static void sdf_timing_collect(ff_ctx *c)
{
    if (!c->sdfTimedPending) return;
    c->sdfTimedPending = false;
    float ms = 0.0f;
    if (hipEventSynchronize(c->ev[3]) == hipSuccess && hipEventElapsedTime(&ms, c->ev[2], c->ev[3]) == hipSuccess) {
        c->phaseMs[1] += ms;
        c->phaseCalls[1]++;
        c->phaseBytes[1] += c->sdfTimedBytes;
        c->phaseUnits[1] += c->sdfTimedUnits;
    } else {
        (void)hipGetLastError();
    }
}

// one thread writes a sequence word's next value: the stream starts the
// launch only once the work before it has ended.  A kernel launch costs the
// host ~2.5 us, the runtime's stream write operation several times that.
__global__ void k_seq_write(uint32_t *word, uint32_t v)
{
    if (threadIdx.x == 0) __hip_atomic_store(word, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

bool seq_signal(ff_ctx *c, int k, hipStream_t st)
{
    const uint32_t v = c->seq[k] + 1u;
    k_seq_write<<<1, 64, 0, st>>>(c->d_seq + k, v);
    if (hipGetLastError() == hipSuccess) {
        c->seq[k] = v;
        return true;
    }
    hipStreamSynchronize(st);
    return false;
}

void seq_wait(ff_ctx *c, int k, hipStream_t st)
{
    if (hipStreamWaitValue32(st, c->d_seq + k, c->seq[k], hipStreamWaitValueGte, 0xffffffffu) != hipSuccess) {
        (void)hipGetLastError();
        hipDeviceSynchronize();
    }
}

// the sequence words back to zero long before a value could wrap (with
// every stream of the context drained, no wait is pending)
static void seq_rebase(ff_ctx *c)
{
    if (c->seq[0] < 0x40000000u && c->seq[1] < 0x40000000u && c->seq[2] < 0x40000000u && c->seq[3] < 0x40000000u)
        return;
    frame_join(c);
    hipStreamSynchronize(c->stream);
    if (c->aux) hipStreamSynchronize(c->aux);
    hipMemset(c->d_seq, 0, sizeof(c->seq));
    for (uint32_t &v : c->seq) v = 0;
}

void frame_join(ff_ctx *c, bool stream)
{
    if (!c->sdfInFlight) return;
    c->sdfInFlight = false;
    if (stream) {
        seq_wait(c, FF_SEQ_SDF, c->stream);
        return;
    }
    hipStreamSynchronize(c->side);
    sdf_timing_collect(c);
}

static int add_map_impl(ff_ctx *c, const uint16_t *depth, const uint8_t *bgr, int W, int H, const double R[9],
                        const double t[3], const double K[4], float scaling, float maxcam, bool overlap)
{
    // a frame on the main stream waits there for the previous frame's SDF
    // update; an overlapped one follows it on the side stream, and the
    // previous update's timing events are read before they are recorded again
    if (!overlap) {
        // a frame on the main stream: a requested mesh update first, then a
        // wait there for the previous frame's SDF update
        const int rf = mesh_flush(c);
        if (rf != FF_OK) return rf;
        frame_join(c, true);
        c->samplePending = false;  // the frame follows the whole update on the stream
    }
    c->sdfTileId = 0;
    sdf_timing_collect(c);
    seq_rebase(c);
    int rc = frame_prepare(c, depth, bgr, W, H, R, t, K, scaling, maxcam, !overlap, overlap);
    if (rc) return rc;
    c->lastPbits = c->d_pbits;
    if (c->n) {
        rc = run_frame(c, W, H, overlap);
        if (rc) return rc;
    } else {
        launch_empty(c, W, H);
    }
    if (overlap && c->n) {
        // the frame ran on the side stream: its counters are in host memory
        // once its allocation is done (its SDF update, which follows the
        // previous mesh update's marching cubes, may not have started when
        // the call returns)
        if (!await_word(&c->h_fc->pad[0], c->stampCounter, c->side)) {
            c->sdfInFlight = false;
            HIPCHK(hipStreamSynchronize(c->side));
            c->err = "the frame's allocation did not report its counters";
            return FF_E_DEVICE;
        }
    } else {
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    if (c->n) {
        bool ran = sdf_gate_host(c);
        if (ran && c->sdfTileId) {
            // and the tile count of the mesh update it followed
            uint32_t nt = 0;
            if (!mesh_sample_report(c, c->sdfTileId, &nt)) {
                c->sdfInFlight = false;
                HIPCHK(hipStreamSynchronize(c->side));
                HIPCHK(hipStreamSynchronize(c->stream));
                c->err = "the mesh update did not report its tile count";
                return FF_E_DEVICE;
            }
            ran = nt <= c->sdfTileCap;
        }
        c->h_fc->sdfRan = ran ? 1u : 0u;
    }
    return frame_finish(c, W, H, true);
}

}  // namespace ff

extern "C" {

int ff_add_map(ff_ctx *c, const uint16_t *depth, size_t depth_stride, const uint8_t *bgr, size_t bgr_stride,
               int W, int H, const double R[9], const double t[3], const double K[4], float scaling, float maxcam)
{
    if (!c || !depth || !R || !t || !K) return FF_E_ARG;
    if (c->useColor && !bgr) {
        c->err = "color is enabled but no bgr image was given";
        return FF_E_ARG;
    }
    hipSetDevice(c->device);
    int rc = ensure_image(c, W, H);
    if (rc) return rc;
    if (!depth_stride) depth_stride = (size_t)W * 2;
    if (!bgr_stride) bgr_stride = (size_t)W * 3;
    // the images go up on the stream that first reads them
    const bool overlap = front_overlap(c, W, H);
    if (!overlap) frame_join(c, true);  // the previous frame may still read c->d_depth, c->d_bgr
    hipStream_t us = overlap ? c->side : c->stream;
    c->uploadOnSide = overlap;
    HIPCHK(hipMemcpy2DAsync(c->d_depth, (size_t)W * 2, depth, depth_stride, (size_t)W * 2, H, hipMemcpyHostToDevice,
                            us));
    if (c->useColor)
        HIPCHK(hipMemcpy2DAsync(c->d_bgr, (size_t)W * 3, bgr, bgr_stride, (size_t)W * 3, H, hipMemcpyHostToDevice, us));
    rc = add_map_impl(c, c->d_depth, c->useColor ? c->d_bgr : nullptr, W, H, R, t, K, scaling, maxcam, overlap);
    c->uploadOnSide = false;
    return rc;
}

int ff_add_map_device(ff_ctx *c, const uint16_t *depth_dev, const uint8_t *bgr_dev, int W, int H, const double R[9],
                      const double t[3], const double K[4], float scaling, float maxcam)
{
    if (!c || !depth_dev || !R || !t || !K) return FF_E_ARG;
    if (c->useColor && !bgr_dev) {
        c->err = "color is enabled but no bgr image was given";
        return FF_E_ARG;
    }
    hipSetDevice(c->device);
    int rc = ensure_image(c, W, H);
    if (rc) return rc;
    // read in place: the frame's SDF update may still read the images when
    // the call returns (until the next frame's counters are in)
    return add_map_impl(c, depth_dev, c->useColor ? bgr_dev : nullptr, W, H, R, t, K, scaling, maxcam,
                        front_overlap(c, W, H));
}

int ff_add_map_device_noalloc(ff_ctx *c, const uint16_t *depth_dev, const uint8_t *bgr_dev, int W, int H,
                              const double R[9], const double t[3], const double K[4], float scaling, float maxcam)
{
    if (!c || !depth_dev || !R || !t || !K) return FF_E_ARG;
    if (c->useColor && !bgr_dev) {
        c->err = "color is enabled but no bgr image was given";
        return FF_E_ARG;
    }
    if (W <= 0 || H <= 0 || W > FF_MAX_W || H > FF_MAX_H) {
        c->err = "image width and height must be in [1, 2048]";
        return FF_E_ARG;
    }
    ScratchLayout L;
    if (scratch_layout(nullptr, W, H, nullptr) > FF_SCRATCH_BYTES) {
        c->err = "image too large for the allocation-free entry point (its buffers must fit 40 MiB: up to "
                 "1280 x 960 pixels)";
        return FF_E_ARG;
    }
    hipSetDevice(c->device);
    unsigned char *base = static_cast<unsigned char *>(scratch_ptr(c->device));
    if (!base) {
        c->err = "no scratch array address for this device";
        return FF_E_STATE;
    }
    scratch_layout(base, W, H, &L);
    // the frame runs on the scratch array's buffers in place of the
    // context's (which keep their contents), with direct launches (a
    // captured graph refers to the context's buffers)
    uint32_t *sv[6] = {c->d_planes, c->d_pbits, c->d_cntB, c->d_cntL, c->d_firstD, c->d_pl0};
    uint64_t *sv64[2] = {c->d_pref, c->d_bsum};
    c->d_planes = L.planes; c->d_pbits = L.pbits; c->d_cntB = L.cntB; c->d_cntL = L.cntL;
    c->d_firstD = L.firstD; c->d_pl0 = L.pl0; c->d_pref = L.pref; c->d_bsum = L.bsum;
    float *confSave = nullptr;
    size_t confCapSave = 0;
    learned_swap_conf(c, L.conf, (size_t)W * H, &confSave, &confCapSave);
    c->directOnly = true;
    const int rc = add_map_impl(c, depth_dev, c->useColor ? bgr_dev : nullptr, W, H, R, t, K, scaling, maxcam,
                                front_overlap(c, W, H));
    c->directOnly = false;
    learned_swap_conf(c, confSave, confCapSave, nullptr, nullptr);
    c->d_planes = sv[0]; c->d_pbits = sv[1]; c->d_cntB = sv[2]; c->d_cntL = sv[3]; c->d_firstD = sv[4];
    c->d_pl0 = sv[5]; c->d_pref = sv64[0]; c->d_bsum = sv64[1];
    return rc;
}

int ff_set_graph_capture(ff_ctx *c, int on)
{
    if (!c) return FF_E_ARG;
    hipSetDevice(c->device);
    drop_frame_graphs(c);
    c->useGraphs = on != 0;
    return FF_OK;
}

int ff_set_timing(ff_ctx *c, int on)
{
    if (!c || on < 0 || on > 2) return FF_E_ARG;
    const int rc = mesh_complete(c);
    if (rc != FF_OK) return rc;
    if (on && !c->ev[0])
        for (int i = 0; i < 10; i++) HIPCHK(hipEventCreate(&c->ev[i]));
    c->timing = on != 0;
    c->timingAll = on == 1;
    return FF_OK;
}

int ff_set_timing_every(ff_ctx *c, uint32_t n)
{
    if (!c || n == 0) return FF_E_ARG;
    c->timeEvery = n;
    return FF_OK;
}

int ff_get_timing(ff_ctx *c, double *ms, uint64_t *calls, double *bytes, uint64_t *units)
{
    if (!c) return FF_E_ARG;
    const int rc = mesh_complete(c);
    if (rc != FF_OK) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->side) HIPCHK(hipStreamSynchronize(c->side));
    sdf_timing_collect(c);
    for (int i = 0; i < 3; i++) {
        if (ms) ms[i] = c->phaseMs[i];
        if (calls) calls[i] = c->phaseCalls[i];
        if (bytes) bytes[i] = c->phaseBytes[i];
        if (units) units[i] = c->phaseUnits[i];
    }
    return FF_OK;
}

int ff_reset_timing(ff_ctx *c)
{
    if (!c) return FF_E_ARG;
    const int rc = mesh_complete(c);  // its counts belong before the reset
    if (rc != FF_OK) return rc;
    if (c->side) HIPCHK(hipStreamSynchronize(c->side));
    c->sdfTimedPending = false;
    for (int i = 0; i < 3; i++) {
        c->phaseMs[i] = 0.0;
        c->phaseCalls[i] = 0;
        c->phaseBytes[i] = 0.0;
        c->phaseUnits[i] = 0;
    }
    c->mcMs = c->mcBytes = 0.0;
    c->mcCalls = c->mcByteCalls = c->mcUnits = 0;
    return FF_OK;
}

int ff_get_mc_timing(ff_ctx *c, double *ms, uint64_t *calls, double *bytes, uint64_t *byte_calls, uint64_t *cells)
{
    if (!c) return FF_E_ARG;
    const int rc = mesh_complete(c);
    if (rc != FF_OK) return rc;
    if (ms) *ms = c->mcMs;
    if (calls) *calls = c->mcCalls;
    if (bytes) *bytes = c->mcBytes;
    if (byte_calls) *byte_calls = c->mcByteCalls;
    if (cells) *cells = c->mcUnits;
    return FF_OK;
}

int ff_get_counts(ff_ctx *c, ff_counts *o)
{
    if (!c || !o) return FF_E_ARG;
    o->n_leaves_used = c->nLeavesUsed;
    o->n_branches_used = c->nBranchesUsed;
    o->n_queued = c->nQueued;
    o->side = c->n;
    for (int k = 0; k < 3; k++) o->offset[k] = c->offset[k];
    o->frames = c->frames;
    o->grows = c->grows;
    return FF_OK;
}

int ff_copy_tree(ff_ctx *c, uint32_t *out, size_t n)
{
    if (!c || !out || n > c->nBranchesTotal) return FF_E_ARG;
    frame_join(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(out, c->d_tree, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return FF_OK;
}

int ff_copy_leaves(ff_ctx *c, uint16_t *pos3, uint16_t *scale, uint32_t *parent, size_t n)
{
    if (!c || n > c->nLeavesTotal) return FF_E_ARG;
    frame_join(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<ushort4> info(n);
    HIPCHK(hipMemcpy(info.data(), c->d_leafinfo, n * sizeof(ushort4), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) {
        if (pos3) { pos3[3 * i] = info[i].x; pos3[3 * i + 1] = info[i].y; pos3[3 * i + 2] = info[i].z; }
        if (scale) scale[i] = info[i].w;
    }
    if (parent) HIPCHK(hipMemcpy(parent, c->d_parent, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return FF_OK;
}

int ff_copy_bricks(ff_ctx *c, const uint32_t *leaves, size_t n, float *dist, float *weight, uint16_t *color)
{
    if (!c || (!leaves && n)) return FF_E_ARG;
    frame_join(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; i++) {
        const uint32_t l = leaves[i];
        if (l >= c->nLeavesTotal) return FF_E_ARG;
        if (dist) HIPCHK(hipMemcpy(dist + i * FF_BRICK, c->d_dist + (size_t)l * FF_BRICK, FF_BRICK * 4, hipMemcpyDeviceToHost));
        if (weight) HIPCHK(hipMemcpy(weight + i * FF_BRICK, c->d_wgt + (size_t)l * FF_BRICK, FF_BRICK * 4, hipMemcpyDeviceToHost));
        if (color && c->d_col)
            HIPCHK(hipMemcpy(color + i * FF_BRICK * 3, c->d_col + (size_t)l * FF_BRICK * 3, FF_BRICK * 6,
                             hipMemcpyDeviceToHost));
    }
    return FF_OK;
}

int ff_copy_queue(ff_ctx *c, uint32_t *out, size_t cap)
{
    if (!c || !out) return FF_E_ARG;
    frame_join(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    const size_t n = std::min<size_t>(cap, c->nQueued);
    HIPCHK(hipMemcpy(out, c->d_queue, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    std::sort(out, out + n);
    return (int)n;
}

int ff_copy_processed(ff_ctx *c, uint8_t *out, size_t n)
{
    if (!c || !out || n < (size_t)c->imgW * c->imgH) return FF_E_ARG;
    frame_join(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    const int W = c->imgW, H = c->imgH, words = (W + 31) / 32;
    (void)words;
    if (!c->lastPbits) return FF_E_STATE;
    std::vector<uint32_t> bits((size_t)(H + 2 * FF_PB_PAD) * FF_PB_STRIDE);
    HIPCHK(hipMemcpy(bits.data(), c->lastPbits, bits.size() * 4, hipMemcpyDeviceToHost));
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++)
            out[(size_t)y * W + x] = (bits[(size_t)(y + FF_PB_PAD) * FF_PB_STRIDE + (x >> 5)] >> (x & 31)) & 1u;
    return FF_OK;
}

}  // extern "C"
