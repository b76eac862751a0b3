// ff_mesh.hip -- incremental marching cubes over mesh cells on gfx950.
//
// The reference keeps a "mesh cell" partition of space
// (src/fusion/meshcelltraversalSplitReferenced.hpp) and re-meshes the cells
// touched by each frame (updateMeshes, geometryfusion_mipmap_cpu.cpp:
// 2884-2954; MeshCell::updateMesh, treeandbrick_incremental.cpp:2867-2987).
// Here the partition is derived from the octree on every update:
//   k_terminal   octants without a child branch that some brick covers
//   k_cells      each such node emits the cells it owns: its interior, and
//                the wall / edge / corner cells between it and its
//                neighbours where it is the smallest adjacent node (ties to
//                the lowest lastLeaf slot); lastLeaf[k] = the finest brick
//                covering side k (located by descending the tree)
//   k_mark       a cell keeps last update's mesh when it existed then with
//                the same bricks and none of the bricks its samples read
//                (its lastLeafs and their parent chains up to the level the
//                samples reached) changed since; a dirty cell gets its
//                sampling grid and a run of 64-sample tiles
//   k_sample     one thread per sample of every dirty cell (reference
//                sampling rules per cell type, parent-chain fallback); each
//                tile raises its cell's chain depth
//   k_mc         one wave per dirty cell: samples into LDS, cube
//                classification, a bitmap prefix scan numbers the used edges
//                in raster order, vertices (double VertexInterp) and faces
//                are written to device pools
// Meshes stay on the device; ff_copy_mesh concatenates them on request.
#include "ff_device.hpp"
#include "ff_internal.hpp"
#include "mc_tables.h"

#include <hip/hip_ext.h>
#include <algorithm>
#include <cstring>
#include <vector>

namespace ff {

struct CellRec {
    uint64_t key;     // ox:16 oy:16 oz:16 log2size:5 type:3
    uint32_t ll[8];   // lastLeaf per side
    uint32_t v0, nv, f0, nf;
    uint32_t hslot;   // slot of the key in the lookup hash
    uint32_t pad[3];  // the update that last meshed it, chain depth, owner offset
};
static_assert(sizeof(CellRec) == 72, "a cell record is nine 8-byte words (k_cells stages it as such)");

struct MeshCounters {
    uint32_t nTerm, nCells, nDirty, flags, nTiles, pad[3];
};
static_assert(sizeof(MeshCounters) == 32, "the SDF update carries MeshCounters up as two 16-byte words");

#define FF_MESH_OVERFLOW 1u
#define FF_CELL_OVERFLOW 2u
#define FF_MESH_HALF 4u  // some arena passed half of its capacity

// The vertex and face pools are split into FF_ARENAS arenas with one
// allocation counter each (vertices in the low, faces in the high 32 bits),
// each on its own 128-byte line: cells spread their allocations over them,
// so no single counter serialises the meshing waves.
#define FF_ARENAS 64u
#define FF_ARENA_STRIDE 16u

struct Pools {
    float4 *v;                    // x, y, z, rgb packed
    uint32_t *f;                  // 3 cell-local indices per face
    unsigned long long *arena;    // FF_ARENAS counters, FF_ARENA_STRIDE apart
    uint32_t vA, fA;              // capacity of one arena
    uint32_t vH, fH;              // use of one arena that asks for a compaction
};

// reserves nv vertices and nf faces in arena k; false when they do not fit.
// In two steps, so that a caller can work while the atomic returns:
// pool_reserve issues it (old: the arena's use before), pool_place checks
// the reservation and gives the first vertex and face.
__device__ __forceinline__ unsigned long long pool_reserve(const Pools &P, uint32_t k, uint32_t nv, uint32_t nf)
{
    if (nv == 0 && nf == 0) return 0ull;
    return atomicAdd(&P.arena[k * FF_ARENA_STRIDE], ((unsigned long long)nf << 32) | nv);
}

__device__ __forceinline__ bool pool_place(const Pools &P, uint32_t k, uint32_t nv, uint32_t nf,
                                           unsigned long long old, uint32_t &v0, uint32_t &f0, MeshCounters *mc)
{
    v0 = f0 = 0;
    if (nv == 0 && nf == 0) return true;
    const uint32_t vo = (uint32_t)old, fo = (uint32_t)(old >> 32);
    if (vo + nv > P.vA || fo + nf > P.fA) {
        atomicOr(&mc->flags, FF_MESH_OVERFLOW);
        return false;
    }
    const uint32_t vh = P.vH, fh = P.fH;
    if ((vo <= vh && vo + nv > vh) || (fo <= fh && fo + nf > fh)) atomicOr(&mc->flags, FF_MESH_HALF);
    v0 = k * P.vA + vo;
    f0 = k * P.fA + fo;
    return true;
}

__device__ __forceinline__ bool pool_alloc(const Pools &P, uint32_t k, uint32_t nv, uint32_t nf, uint32_t &v0,
                                           uint32_t &f0, MeshCounters *mc)
{
    return pool_place(P, k, nv, nf, pool_reserve(P, k, nv, nf), v0, f0, mc);
}

// per cube index: its triangles' edges, 4 bits each in order (bits 0-59),
// and the triangle count (bits 60-63); one load per cube
__constant__ uint64_t c_tri64[256];

#define FF_TILE_CAP_MAX (1u << 22)  // 4 GiB of samples
#define FF_POOL_MAX (1u << 29)      // vertices (8 GiB) and triangles (6 GiB) the pools grow to

// read-only view of the map for sampling
struct MapRO {
    const ushort4 *info;     // pos, scale
    const uint32_t *parent;
    const float *dist, *wgt;
    const uint16_t *col;     // R, G, B planes of 512 per brick
    float off[3];
    float scale;
    int useColor;
};

// one stream of a batched mesh update: every buffer and size the per-stream
// mesh kernels take (grid z = stream)
struct MeshBatchStream {
    const uint32_t *tree;
    const ushort4 *binfo;
    uint32_t nBlocks, nLeaves, nSlots, n;  // the tree as launched (bricks, branch slots)
    uint4 *term;
    uint32_t termCap;
    MeshCounters *mc;
    CellRec *cells;       // the table being built
    const CellRec *prev;  // the previous table
    uint32_t cellCap, nPrev;
    uint64_t *hkeys;
    uint32_t *hval;
    uint32_t hmask;
    int havePrev;
    const uint32_t *stamp, *parent;
    uint32_t since;
    uint32_t updateId;  // recorded in the cells this update re-meshes
    struct CellGeom *geom;
    MapRO m;
    struct TileRec *tiles;
    uint32_t tileCap;
    uint4 *sbuf;
    const unsigned long long *arena;
    int arenaSum;  // totals of the arena counters into mc->pad (timing)
};

struct MeshState {
    ushort4 *d_binfo = nullptr;  // per branch block: origin of the node, octant size
    uint32_t binfoCap = 0;
    uint4 *d_term = nullptr;     // x, y, z, (size << 0) ; .w = covering leaf, size packed in x high bits
    uint4 *d_jump = nullptr;     // FF_JUMP_CELLS descent entries of the current tree
    uint32_t *d_termSize = nullptr;
    uint32_t termCap = 0;
    CellRec *d_cells[2] = {nullptr, nullptr};
    uint32_t cellCap = 0;
    // incremental partition (an update whose tree grew by a few bricks and
    // branches since the current table was built): the terminal nodes the
    // change touches, the cells they own (fresh), and counters: [0] touched
    // nodes, [1] fresh cells
    uint4 *d_termAff = nullptr;
    CellRec *d_fresh = nullptr;
    uint32_t *d_inc = nullptr;
    uint32_t nCells[2] = {0, 0};
    int cur = 0;
    struct CellGeom *d_geom = nullptr;  // per dirty cell
    struct TileRec *d_tiles = nullptr;  // per sample tile: its cell's geometry and bricks
    uint4 *d_sbuf = nullptr;            // 64 samples per tile: d, w, R | G << 16, B
    uint32_t tileCap = 0;
    // cell key -> index of the current table, in one of two hashes: an
    // update that builds a new table empties the current hash of the old
    // keys and fills the other with the new ones in one pass (k_hash_turnover)
    uint64_t *d_hkeysBuf[2] = {nullptr, nullptr};
    uint32_t *d_hvalBuf[2] = {nullptr, nullptr};
    int hIdx = 0;
    uint64_t *d_hkeys = nullptr;  // = d_hkeysBuf[hIdx]
    uint32_t *d_hval = nullptr;   // = d_hvalBuf[hIdx]
    uint32_t hmask = 0;
    float4 *d_vpool = nullptr;    // x, y, z, rgb packed
    uint32_t *d_fpool = nullptr;  // 3 local indices per face
    uint32_t vCap = 0, fCap = 0;  // FF_ARENAS arenas of vCap / FF_ARENAS, fCap / FF_ARENAS
    uint32_t compactions = 0;     // live-geometry compactions so far
    uint32_t compactPercent = 50; // an arena used this far since the last compaction asks for one
    unsigned long long *d_arena = nullptr;
    MeshCounters *d_mc = nullptr, *h_mc = nullptr;  // d_mc: the counters of the update in flight
    // two sets of device counters, alternating between single updates: the
    // first kernel of an update clears the set the next update counts in
    // (its last readers -- the previous update's kernels, compaction and
    // the SDF update that took its counters up -- are done by then), so
    // that no clearing launch is needed
    MeshCounters *d_mcBuf[2] = {nullptr, nullptr};
    int mcIdx = 0;
    bool mcNextZero = false, mcZeroNow = false;
    MeshCounters *d_hmc = nullptr;  // h_mc (mapped host memory) as the device addresses it
    uint32_t lastMeshStamp = 0;
    int nxt = 1;   // table being built by the update in flight
    MapRO m{};     // its map view
    // batched updates led by this context: stream descriptors and prefix
    struct McStream *d_ms = nullptr, *h_ms = nullptr;  // inside d_mdesc / h_mdesc
    MeshBatchStream *d_mbs = nullptr, *h_mbs = nullptr;
    char *d_mdesc = nullptr, *h_mdesc = nullptr;  // one block: one upload per update
    size_t mdescBytes = 0;
    MeshCounters *d_mcs = nullptr, *h_mcs = nullptr;  // the streams' counters, gathered
    uint32_t *d_msPrefix = nullptr;
    int msCap = 0;
    uint64_t lastVAlloc = 0, lastFAlloc = 0;  // arena totals after the previous update (timing)
    bool haveVolume = false;
    // the tree the current table was built for (used branch slots, bricks,
    // grows): while it is unchanged the partition is, and an update re-marks
    // the current table in place instead of building a new one
    uint32_t builtNB = 0, builtNL = 0, builtGrows = 0;
    bool reuse = false;  // the update in flight re-marks the current table
    bool mcTimed = false;  // the update in flight records k_mc's events (ev[8], ev[9])
    // updates numbered from 1: the cells an update re-meshes carry its
    // number (CellRec::pad[0]); lastUpdateId is the last completed one
    uint32_t updateSeq = 0, launchId = 0, lastUpdateId = 0;
    bool fullNext = false;  // the next update meshes every cell from scratch (ff_mesh_full)
    // an update whose device work is enqueued and whose host part
    // (mesh_finish) runs at the next mesh call, query or state change
    bool pending = false, pendingTimed = false;
    // an update requested (updateMeshes) but not yet enqueued: the next
    // frame enqueues it right after its first kernels, every other call on
    // the context first; reqStamp is the frame stamp it meshes
    bool requested = false;
    uint32_t reqStamp = 0;
    uint32_t stampFor = 0;  // the stamp of the update being enqueued
    // reports in mapped host memory after h_mc: h_rep[0] the update's tile
    // count and h_rep[1] its number (k_sample, when the marking is done),
    // h_rep[2] its number once its counters are in h_mc (k_mesh_report,
    // after the marching cubes); d_rep as the device addresses them
    uint32_t *h_rep = nullptr, *d_rep = nullptr;
    bool doneMapped = false;  // the update in flight reports its counters itself
    // a hash turnover of the update in flight that the next frame launches
    // on its side stream after its SDF update (off the main stream)
    bool turnPending = false;
    int turnTable = 0;
    // the map the update in flight meshes: the frame stamp and the tree
    // (frames fused before its host part runs must not be attributed to it)
    uint32_t launchStamp = 0, launchNB = 0, launchNL = 0, launchGrows = 0;
    // the allocating part of an update's host part (sample-buffer growth,
    // compaction into fresh pools, pool growth), left for the next mesh call
    // or query when the host part ran inside a frame (a grow): frames
    // allocate no device memory
    bool maintPending = false;
    uint32_t maintFlags = 0, maintTiles = 0;
    int maintTable = 0;
};

// both cell hashes emptied (every cell is meshed again by the next update)
static hipError_t hash_reset(MeshState *s, hipStream_t st)
{
    hipError_t e = hipSuccess;
    for (int k = 0; k < 2 && e == hipSuccess; k++)
        e = hipMemsetAsync(s->d_hkeysBuf[k], 0xff, sizeof(uint64_t) * (s->hmask + 1), st);
    return e;
}


// ---------------------------------------------------------------------------
// tree queries
// ---------------------------------------------------------------------------
// The mesh update reads the tree as it was when the update was launched:
// bricks below nLeaves and branch slots below nSlots.  Indices are handed
// out in increasing order and a slot, once linked, keeps its link, so a link
// to a brick or branch at or past those counts was made by a later frame
// (whose allocation may run beside the update) and reads as empty.
//
// Terminal node containing voxel (x, y, z): returns its size (0 outside the
// volume), origin, and the deepest brick on the path.
__device__ uint32_t locate(const uint32_t *__restrict__ tree, uint32_t nLeaves, uint32_t nSlots, uint32_t n, int x,
                           int y, int z, int org[3], uint32_t *leaf)
{
    *leaf = FF_BRANCHINIT;
    if (x < 0 || y < 0 || z < 0 || x >= (int)n || y >= (int)n || z >= (int)n) return 0u;
    uint32_t half = n >> 1, branch = 0;
    int ox = 0, oy = 0, oz = 0;
    for (;;) {
        uint32_t co = 0;
        if (x >= ox + (int)half) { co |= 1u; ox += half; }
        if (y >= oy + (int)half) { co |= 2u; oy += half; }
        if (z >= oz + (int)half) { co |= 4u; oz += half; }
        const uint32_t l = tree[branch + co];
        if (l < nLeaves) *leaf = l;
        const uint32_t child = tree[branch + co + 8];
        if (child >= nSlots || half <= FF_BL) {
            org[0] = ox; org[1] = oy; org[2] = oz;
            return half;
        }
        branch = child;
        half >>= 1;
    }
}

// Jump table for descents: the volume as a 16 x 16 x 16 grid of cells of side
// n / 16 (n >= 16, a power of two), each with what the first four levels of
// a descent to any of its voxels give: the branch to go on from (half size
// n / 32) and the deepest brick so far, or, for a descent that ends within
// them, no branch and the terminal node's half size.  Nodes of those levels
// are aligned to their size, so every voxel of a cell takes the same octants.
#define FF_JUMP_N 16
#define FF_JUMP_CELLS (FF_JUMP_N * FF_JUMP_N * FF_JUMP_N)

__device__ uint4 jump_entry(const uint32_t *__restrict__ tree, uint32_t nLeaves, uint32_t nSlots, uint32_t n,
                           uint32_t cell)
{
    const uint32_t cs = n / FF_JUMP_N;
    const int x = (int)((cell % FF_JUMP_N) * cs), y = (int)(((cell / FF_JUMP_N) % FF_JUMP_N) * cs),
              z = (int)((cell / (FF_JUMP_N * FF_JUMP_N)) * cs);
    uint32_t leaf = FF_BRANCHINIT, half = n >> 1, branch = 0;
    int ox = 0, oy = 0, oz = 0;
    for (int lv = 0; lv < 4; lv++) {
        uint32_t co = 0;
        if (x >= ox + (int)half) { co |= 1u; ox += half; }
        if (y >= oy + (int)half) { co |= 2u; oy += half; }
        if (z >= oz + (int)half) { co |= 4u; oz += half; }
        const uint32_t l = tree[branch + co];
        if (l < nLeaves) leaf = l;
        const uint32_t child = tree[branch + co + 8];
        if (child >= nSlots || half <= FF_BL) return make_uint4(FF_BRANCHINIT, leaf, half, 0u);
        branch = child;
        half >>= 1;
    }
    return make_uint4(branch, leaf, 0u, 0u);
}

// locate() starting from the jump table: the same result, four dependent
// tree loads fewer
__device__ uint32_t locate_jump(const uint4 *__restrict__ jump, const uint32_t *__restrict__ tree,
                                uint32_t nLeaves, uint32_t nSlots, uint32_t n, int x, int y, int z, int org[3],
                                uint32_t *leaf)
{
    *leaf = FF_BRANCHINIT;
    if (x < 0 || y < 0 || z < 0 || x >= (int)n || y >= (int)n || z >= (int)n) return 0u;
    const uint32_t cs = n / FF_JUMP_N;
    const uint4 e = jump[((uint32_t)z / cs * FF_JUMP_N + (uint32_t)y / cs) * FF_JUMP_N + (uint32_t)x / cs];
    *leaf = e.y;
    if (e.x == FF_BRANCHINIT) {
        const int m = ~((int)e.z - 1);
        org[0] = x & m; org[1] = y & m; org[2] = z & m;
        return e.z;
    }
    uint32_t half = n >> 5, branch = e.x;
    int ox = x & ~((int)cs - 1), oy = y & ~((int)cs - 1), oz = z & ~((int)cs - 1);
    for (;;) {
        uint32_t co = 0;
        if (x >= ox + (int)half) { co |= 1u; ox += half; }
        if (y >= oy + (int)half) { co |= 2u; oy += half; }
        if (z >= oz + (int)half) { co |= 4u; oz += half; }
        const uint32_t l = tree[branch + co];
        if (l < nLeaves) *leaf = l;
        const uint32_t child = tree[branch + co + 8];
        if (child >= nSlots || half <= FF_BL) {
            org[0] = ox; org[1] = oy; org[2] = oz;
            return half;
        }
        branch = child;
        half >>= 1;
    }
}

__device__ __forceinline__ uint64_t cell_key(int type, int ox, int oy, int oz, int size)
{
    const uint32_t lg = 31u - __clz((uint32_t)size);
    return ((uint64_t)(uint16_t)ox << 48) | ((uint64_t)(uint16_t)oy << 32) | ((uint64_t)(uint16_t)oz << 16) |
           ((uint64_t)lg << 3) | (uint64_t)type;
}

// Appends one element per participating lane with a single atomic per wave.
// Every lane of the wave must call it (want may be false).
// the set bits of m below this lane (mbcnt: no 64-bit lane mask is formed,
// which a kernel would otherwise hoist and hold in two registers)
__device__ __forceinline__ uint32_t lanes_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool want)
{
    const uint64_t m = __ballot(want);
    if (!m) return 0xffffffffu;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    return want ? base + lanes_below(m) : 0xffffffffu;
}

// The same for a whole block (every thread calls it; it holds two barriers):
// one atomic per block, since appends of many waves to one counter
// serialise on its address.
template <int NW>
__device__ __forceinline__ uint32_t block_append(uint32_t *counter, bool want)
{
    __shared__ uint32_t wc[NW + 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t m = __ballot(want);
    if (lane == 0) wc[w] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int k = 0; k < NW; k++) {
            const uint32_t c = wc[k];
            wc[k] = t;
            t += c;
        }
        wc[NW] = t ? atomicAdd(counter, t) : 0u;
    }
    __syncthreads();
    const uint32_t r = wc[NW] + wc[w] + lanes_below(m);
    __syncthreads();  // wc is reused by the next call
    return want ? r : 0xffffffffu;
}

// A block-wide exclusive rank of `want` among the block's threads and the
// block's total (every thread calls it; it holds two barriers).
template <int NW>
__device__ __forceinline__ uint32_t block_rank(bool want, uint32_t *total)
{
    __shared__ uint32_t wc[NW + 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t m = __ballot(want);
    if (lane == 0) wc[w] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int k = 0; k < NW; k++) {
            const uint32_t c = wc[k];
            wc[k] = t;
            t += c;
        }
        wc[NW] = t;
    }
    __syncthreads();
    const uint32_t r = wc[w] + lanes_below(m);
    *total = wc[NW];
    __syncthreads();  // wc is reused by the next call
    return r;
}

// ---------------------------------------------------------------------------
// incremental partition
// ---------------------------------------------------------------------------
// A cell depends only on its owner node p (size s) and on the terminal
// nodes and deepest bricks at p's option points, which lie in [p - 1, p + s]
// on each axis.  Bricks and branches are only ever added (a grow rebuilds the
// partition), so between two tables the tree changes by the bricks and
// branch blocks made since the older one: a node whose option box meets none
// of their boxes has the same cells in both tables, and a cell changes owner
// only between nodes whose option boxes meet a change.  An update with few
// changes keeps the cells of untouched owners and recomputes those of
// touched terminal nodes (their option points read as the full partition
// reads them).
#define FF_INC_BOXES 1024u  // changed bricks + branch blocks an incremental update takes
#define FF_INC_NODES 4096u  // touched terminal nodes an incremental update takes (27 cells each at most)

struct IncView {
    uint32_t l0, l1;  // bricks made since the table was built
    uint32_t b0, b1;  // branch blocks made since
};

// the change boxes (x, y, z, side) into LDS (every thread calls it; one
// barrier): bricks from their position and scale, branch blocks from the
// node geometry the allocation records
__device__ __forceinline__ uint32_t load_changes(const IncView &v, const ushort4 *__restrict__ info,
                                                 const ushort4 *__restrict__ binfo, uint4 *box)
{
    const uint32_t nl = v.l1 - v.l0, n = min(nl + (v.b1 - v.b0), FF_INC_BOXES);
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
        if (k < nl) {
            const ushort4 li = info[v.l0 + k];
            box[k] = make_uint4(li.x, li.y, li.z, (uint32_t)li.w * FF_BL);
        } else {
            const ushort4 bi = binfo[v.b0 + (k - nl)];
            box[k] = make_uint4(bi.x, bi.y, bi.z, 2u * bi.w);
        }
    }
    __syncthreads();
    return n;
}

// whether node (p, s) has an option point in one of the n change boxes
__device__ __forceinline__ bool node_touched(const uint4 *box, uint32_t n, int px, int py, int pz, int s)
{
    for (uint32_t k = 0; k < n; k++) {
        const uint4 b = box[k];
        if (px - 1 < (int)(b.x + b.w) && (int)b.x <= px + s && py - 1 < (int)(b.y + b.w) && (int)b.y <= py + s &&
            pz - 1 < (int)(b.z + b.w) && (int)b.z <= pz + s)
            return true;
    }
    return false;
}

// the owner node of a cell: the key's origin plus the recorded offsets
// (CellRec::pad[2], bit k: + size on axis k)
__device__ __forceinline__ bool cell_owner_touched(uint64_t key, uint32_t own, const uint4 *box, uint32_t n)
{
    const int s = 1 << ((key >> 3) & 31u);
    const int ox = (int)(key >> 48) + ((own & 1u) ? s : 0), oy = (int)((key >> 32) & 0xffffu) + ((own & 2u) ? s : 0),
              oz = (int)((key >> 16) & 0xffffu) + ((own & 4u) ? s : 0);
    return node_touched(box, n, ox, oy, oz, s);
}

// whether an update with these counts runs incrementally (the same on
// every kernel of the update)
__device__ __forceinline__ bool inc_mode(bool inc, const uint32_t *incCnt)
{
    return inc && incCnt[0] <= FF_INC_NODES;
}

// ---------------------------------------------------------------------------
// k_terminal: covered terminal nodes
// ---------------------------------------------------------------------------
#define FF_TERM_THREADS 256  // block size of k_terminal (every thread of a block appends together)
// box, nBox: the change boxes of an incremental update (nullptr: a full
// one), whose touched terminal nodes also go to termAff (count incCnt[0])
__device__ __forceinline__ void terminal_body(const uint32_t *__restrict__ tree, const ushort4 *__restrict__ binfo,
                                              uint32_t nBlocks, uint32_t nLeaves, uint32_t nSlots, uint32_t n,
                                              uint4 *__restrict__ term, uint32_t termCap, MeshCounters *mc,
                                              const uint4 *box = nullptr, uint32_t nBox = 0,
                                              uint4 *termAff = nullptr, uint32_t *incCnt = nullptr)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    bool want = false, touched = false;
    uint4 rec = make_uint4(0, 0, 0, 0);
    if (t < nBlocks * 8u) {
        const uint32_t b = t >> 3, co = t & 7u;
        const ushort4 bi = binfo[b];
        const uint32_t half = bi.w;
        if (tree[b * 16 + 8 + co] >= nSlots && half != 0) {
            const int x = bi.x + (int)(half * (co & 1u)), y = bi.y + (int)(half * ((co >> 1) & 1u)),
                      z = bi.z + (int)(half * ((co >> 2) & 1u));
            int org[3];
            uint32_t lt;
            const uint32_t sz = locate(tree, nLeaves, nSlots, n, x, y, z, org, &lt);
            if (sz == half && lt != FF_BRANCHINIT) {
                want = true;
                rec = make_uint4((uint32_t)x | ((uint32_t)y << 16), (uint32_t)z, half, lt);
                touched = box && node_touched(box, nBox, x, y, z, (int)half);
            }
        }
    }
    const uint32_t i = block_append<FF_TERM_THREADS / 64>(&mc->nTerm, want);
    if (want) {
        if (i < termCap) term[i] = rec;
        else atomicOr(&mc->flags, FF_CELL_OVERFLOW);
    }
    if (box) {
        // (more than FF_INC_NODES: the update runs in full, inc_mode)
        const uint32_t j = block_append<FF_TERM_THREADS / 64>(&incCnt[0], touched);
        if (touched && j < FF_INC_NODES) termAff[j] = rec;
    }
}

// (blocks past the terminal candidates fill the jump table, 256 cells each)
// inc: an incremental update (the change boxes from iv; incCnt zero)
__global__ __launch_bounds__(FF_TERM_THREADS) void k_terminal(const uint32_t *__restrict__ tree, const ushort4 *__restrict__ binfo, uint32_t nBlocks,
                           uint32_t nLeaves, uint32_t nSlots, uint32_t n, uint4 *__restrict__ term, uint32_t termCap,
                           MeshCounters *mc, uint4 *__restrict__ jump, MeshCounters *clearNext, bool inc,
                           IncView iv, const ushort4 *__restrict__ info, uint4 *termAff, uint32_t *incCnt)
{
    __shared__ uint4 box[FF_INC_BOXES];
    if (blockIdx.x == 0 && threadIdx.x == 0) *clearNext = MeshCounters{};  // the next update's counters
    const uint32_t tb = (nBlocks * 8u + blockDim.x - 1u) / blockDim.x;
    if (blockIdx.x >= tb) {
        const uint32_t cell = (blockIdx.x - tb) * blockDim.x + threadIdx.x;
        if (cell < FF_JUMP_CELLS) jump[cell] = jump_entry(tree, nLeaves, nSlots, n, cell);
        return;
    }
    if (!inc) {
        terminal_body(tree, binfo, nBlocks, nLeaves, nSlots, n, term, termCap, mc);
        return;
    }
    const uint32_t nBox = load_changes(iv, info, binfo, box);
    terminal_body(tree, binfo, nBlocks, nLeaves, nSlots, n, term, termCap, mc, box, nBox, termAff, incCnt);
}

// the batched forms of the per-stream mesh kernels: stream blockIdx.z of S
__global__ __launch_bounds__(FF_TERM_THREADS) void k_terminal_b(const MeshBatchStream *__restrict__ S)
{
    const MeshBatchStream &a = S[blockIdx.z];
    if (blockIdx.x * blockDim.x >= a.nBlocks * 8u) return;
    terminal_body(a.tree, a.binfo, a.nBlocks, a.nLeaves, a.nSlots, a.n, a.term, a.termCap, a.mc);
}

// ---------------------------------------------------------------------------
// k_cells: cells owned by each covered terminal node
// ---------------------------------------------------------------------------
// Every ownership test looks at terminal nodes containing points next to the
// node p (size s): per axis the coordinate is one of p - 1, p, p + s - 1,
// p + s (option 0..3).  A block of 512 threads takes 8 nodes at a time: one
// thread locates each of the 64 option points of a node (the 8 inside ones
// are the node itself), then one thread per (node, candidate) decides
// ownership from them; the block appends its cells with one atomic.
__device__ __forceinline__ int opt_code(int ox, int oy, int oz) { return ox + 4 * oy + 16 * oz; }

#define FF_CELL_STAGE 256  // cells a k_cells block stages before appending (8 nodes emit at most 216)

// the cells of the nt terminal nodes term[] appended to cells (count *count,
// capacity cap); each records its owner's offset from its origin (pad[2])
__device__ __forceinline__ void cells_body(const uint4 *__restrict__ jump, const uint32_t *__restrict__ tree,
                                           uint32_t nLeaves, uint32_t nSlots, uint32_t n,
                                           const uint4 *__restrict__ term, uint32_t nt, MeshCounters *mc,
                                           CellRec *cells, uint32_t cap, uint32_t *count)
{
    __shared__ uint2 nb[8][64];  // (size, lastLeaf) per node and option point
    // the block's cells are staged here and appended to the table together,
    // one atomic per FF_CELL_STAGE cells (appends of every block to one
    // counter serialise on its address)
    __shared__ uint2 stage[FF_CELL_STAGE][9];
    __shared__ uint32_t nstage, sbase;
    const int tid = threadIdx.x;
    if (tid == 0) nstage = 0u;
    // (the first iteration's barriers order this before any use)
    auto flush = [&]() {
        if (tid == 0) sbase = nstage ? atomicAdd(count, nstage) : 0u;
        __syncthreads();
        const uint32_t ns = nstage, b = sbase;
        for (uint32_t e = (uint32_t)tid; e < ns * 9u; e += blockDim.x) {
            const uint32_t k = e / 9u;
            if (b + k < cap) reinterpret_cast<uint2 *>(cells + b + k)[e % 9u] = stage[k][e % 9u];
        }
        if (tid == 0 && ns && b + ns > cap) atomicOr(&mc->flags, FF_CELL_OVERFLOW);
        __syncthreads();  // stage is read before it is refilled
        if (tid == 0) nstage = 0u;
        __syncthreads();
    };
    for (uint32_t n0 = blockIdx.x * 8u; n0 < nt; n0 += gridDim.x * 8u) {
        {
            const uint32_t node = n0 + (uint32_t)(tid >> 6);
            const int code = tid & 63;
            uint2 r = make_uint2(0u, FF_BRANCHINIT);
            if (node < nt) {
                const uint4 tn = term[node];
                const int p[3] = {(int)(tn.x & 0xffffu), (int)(tn.x >> 16), (int)tn.y};
                const int sz = (int)tn.z;
                const int op[3] = {code & 3, (code >> 2) & 3, code >> 4};
                const bool inside = (op[0] == 1 || op[0] == 2) && (op[1] == 1 || op[1] == 2) && (op[2] == 1 || op[2] == 2);
                if (inside) {
                    r = make_uint2((uint32_t)sz, tn.w);
                } else {
                    int q[3];
                    for (int k = 0; k < 3; k++)
                        q[k] = op[k] == 0 ? p[k] - 1 : (op[k] == 1 ? p[k] : (op[k] == 2 ? p[k] + sz - 1 : p[k] + sz));
                    int org[3];
                    uint32_t lf;
                    r.x = jump ? locate_jump(jump, tree, nLeaves, nSlots, n, q[0], q[1], q[2], org, &lf)
                               : locate(tree, nLeaves, nSlots, n, q[0], q[1], q[2], org, &lf);
                    r.y = lf;
                }
            }
            nb[tid >> 6][code] = r;
        }
        __syncthreads();
        {
            // candidates: 0 interior, 1-6 walls, 7-18 edges, 19-26 corners
            const uint32_t node = n0 + (uint32_t)(tid / 27);
            const int cand = tid % 27;
            bool emit = false;
            int type = 0, o[3] = {0, 0, 0}, s = 1, nll = 0;
            uint32_t ll[8], own = 0;  // own: bit k, the owner lies at o + s on axis k
            if (tid < 8 * 27 && node < nt) {
                const uint2 *N = nb[tid / 27];
                const uint4 tn = term[node];
                const int p[3] = {(int)(tn.x & 0xffffu), (int)(tn.x >> 16), (int)tn.y};
                s = (int)tn.z;
                const uint32_t lt = tn.w;
                if (cand == 0) {
                    ll[0] = lt;
                    type = 0; o[0] = p[0]; o[1] = p[1]; o[2] = p[2]; nll = 1;
                    emit = true;
                } else if (cand <= 6) {
                    const int a = (cand - 1) >> 1, side = (cand - 1) & 1;
                    int op[3] = {1, 1, 1};
                    op[a] = side ? 3 : 0;
                    const uint2 r = N[opt_code(op[0], op[1], op[2])];
                    const uint32_t ns = r.x;
                    if (ns && (side ? (ns >= (uint32_t)s) : (ns > (uint32_t)s))) {
                        o[0] = p[0]; o[1] = p[1]; o[2] = p[2];
                        if (!side) o[a] = p[a] - s;
                        own = side ? 0u : 1u << a;
                        ll[side ? 0 : 1] = lt;
                        ll[side ? 1 : 0] = r.y;
                        type = 1 + a; nll = 2;
                        emit = true;
                    }
                } else if (cand <= 18) {
                    const int e = (cand - 7) >> 2, cu = (cand - 7) & 1, cv = ((cand - 7) >> 1) & 1;
                    const int u = e == 0 ? 1 : 0, v = e == 2 ? 1 : 2;
                    const int mine = (1 - cu) + 2 * (1 - cv);
                    bool ok = true;
                    for (int sl = 0; sl < 4; sl++) {
                        int op[3];
                        op[e] = 1;
                        op[u] = 2 * cu + (sl & 1);
                        op[v] = 2 * cv + (sl >> 1);
                        const uint2 r = N[opt_code(op[0], op[1], op[2])];
                        ll[sl] = r.y;
                        // a smaller node, or an equal one in an earlier slot, owns the cell
                        if (sl != mine && (!r.x || r.x < (uint32_t)s || (sl < mine && r.x == (uint32_t)s))) ok = false;
                    }
                    if (ok) {
                        o[e] = p[e]; o[u] = p[u] + cu * s - s; o[v] = p[v] + cv * s - s;
                        own = (cu ? 0u : 1u << u) | (cv ? 0u : 1u << v);
                        type = 4 + e; nll = 4;
                        emit = true;
                    }
                } else {
                    const int cb = cand - 19, c3[3] = {cb & 1, (cb >> 1) & 1, cb >> 2};
                    const int mine = (1 - c3[0]) + 2 * (1 - c3[1]) + 4 * (1 - c3[2]);
                    bool ok = true;
                    for (int sl = 0; sl < 8; sl++) {
                        const uint2 r = N[opt_code(2 * c3[0] + (sl & 1), 2 * c3[1] + ((sl >> 1) & 1), 2 * c3[2] + (sl >> 2))];
                        ll[sl] = r.y;
                        if (sl != mine && (!r.x || r.x < (uint32_t)s || (sl < mine && r.x == (uint32_t)s))) ok = false;
                    }
                    if (ok) {
                        for (int k = 0; k < 3; k++) o[k] = p[k] + c3[k] * s - s;
                        own = (c3[0] ? 0u : 1u) | (c3[1] ? 0u : 2u) | (c3[2] ? 0u : 4u);
                        type = 7; nll = 8;
                        emit = true;
                    }
                }
                // pushMeshCellQueue needs every side's brick
                for (int k = 0; k < nll && emit; k++)
                    if (ll[k] >= FF_BRANCHINIT) emit = false;
            }
            // (its barriers also order this iteration's nb reads before the
            // next iteration's writes)
            uint32_t total;
            const uint32_t rank = block_rank<8>(emit, &total);
            if (nstage + total > FF_CELL_STAGE) flush();  // block-uniform
            if (emit) {
                CellRec r;
                r.key = cell_key(type, o[0], o[1], o[2], s);
                for (int k = 0; k < 8; k++) r.ll[k] = k < nll ? ll[k] : FF_BRANCHINIT;
                r.v0 = r.nv = r.f0 = r.nf = 0;
                r.hslot = 0xffffffffu;
                r.pad[0] = r.pad[1] = 0;
                r.pad[2] = own;
                const uint2 *q = reinterpret_cast<const uint2 *>(&r);
                for (int k = 0; k < 9; k++) stage[nstage + rank][k] = q[k];
            }
            __syncthreads();  // every stage write and nstage read are done
            if (tid == 0) nstage += total;
            __syncthreads();
        }
    }
    flush();
}

// inc: an incremental update, which (inc_mode) takes the touched nodes
// termAff and appends their cells to fresh (count incCnt[1])
__global__ __launch_bounds__(512) void k_cells(const uint32_t *__restrict__ tree, uint32_t nLeaves, uint32_t nSlots,
                                               uint32_t n,
                                               const uint4 *__restrict__ term, uint32_t termCap, MeshCounters *mc,
                                               CellRec *cells, uint32_t cap, const uint4 *__restrict__ jump, bool inc,
                                               const uint4 *__restrict__ termAff, CellRec *fresh, uint32_t *incCnt)
{
    if (inc_mode(inc, incCnt))
        cells_body(jump, tree, nLeaves, nSlots, n, termAff, incCnt[0], mc, fresh, FF_INC_NODES * 27u, &incCnt[1]);
    else
        cells_body(jump, tree, nLeaves, nSlots, n, term, min(mc->nTerm, termCap), mc, cells, cap, &mc->nCells);
}

__global__ __launch_bounds__(512) void k_cells_b(const MeshBatchStream *__restrict__ S)
{
    const MeshBatchStream &a = S[blockIdx.z];
    cells_body(nullptr, a.tree, a.nLeaves, a.nSlots, a.n, a.term, min(a.mc->nTerm, a.termCap), a.mc, a.cells,
               a.cellCap, &a.mc->nCells);
}

// ---------------------------------------------------------------------------
// k_mark: reuse the previous mesh of unchanged cells
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t hkey_hash(uint64_t k, uint32_t mask)
{
    k ^= k >> 31;
    k *= 0x7fb5d329728ea185ull;
    k ^= k >> 27;
    k *= 0x81dadef4bc2dd44dull;
    k ^= k >> 33;
    return (uint32_t)k & mask;
}

// a brick among the first `depth` (at most 16) bricks of the parent chain of
// any of the bricks ll[] updated after `since`; the chains advance together,
// so each level's loads are in flight at once.  depth is the deepest chain
// level the cell's samples read when it was last meshed: a brick further up
// was not read (every sample stopped at a brick with weight), and the bricks
// below it are unchanged, so the samples still stop there.
__device__ __forceinline__ bool chains_dirty(const uint32_t *__restrict__ stamp, const uint32_t *__restrict__ parent,
                                             const uint32_t ll[8], uint32_t since, uint32_t depth)
{
    uint32_t cur[8];
#pragma unroll
    for (int k = 0; k < 8; k++) cur[k] = ll[k];
    for (int d = 0; d < (int)min(depth, 16u); d++) {
        uint32_t st[8], pa[8];
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (cur[k] < FF_BRANCHINIT) {
                st[k] = stamp[cur[k]];
                pa[k] = parent[cur[k]];
            }
        bool any = false;
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (cur[k] < FF_BRANCHINIT) {
                if (st[k] > since) return true;
                cur[k] = pa[k];
                any = true;
            }
        if (!any) return false;
    }
    return false;
}

// ---------------------------------------------------------------------------
// sampling (per cell type)
// ---------------------------------------------------------------------------

struct Sample {
    float d, w;
    uint16_t c[3];
    uint16_t lv;  // bricks read along the parent chain (1: the lastLeaf alone)
};

// orders LDS traffic between the lanes of one wave (each wave owns its own
// LDS slice, so no block barrier is needed)
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int clamp7(int v) { return v < 0 ? 0 : (v > 7 ? 7 : v); }

__device__ __forceinline__ void split_axis(int X, int lo, int ls, int &il, float &r, int &ih)
{
    const int l = (X - lo) / ls, rem = (X - lo) % ls;
    const float rr = __fdiv_rn((float)rem, (float)ls);
    il = clamp7(l);
    r = rr;
    ih = clamp7(l + ((rr > 0.0f && l < FF_BL - 1) ? 1 : 0));
}

__device__ __forceinline__ void fetch(const MapRO &m, uint32_t leaf, int ix, int iy, int iz, float &w, float &d,
                                      float c[3])
{
    const size_t q = (size_t)leaf * FF_BRICK + (size_t)((iz * FF_BL + iy) * FF_BL + ix);
    w = m.wgt[q];
    d = m.dist[q];
    if (m.useColor) {
        const uint16_t *cb = m.col + (size_t)leaf * FF_BRICK * 3 + (size_t)((iz * FF_BL + iy) * FF_BL + ix);
        c[0] = (float)cb[0]; c[1] = (float)cb[FF_BRICK]; c[2] = (float)cb[2 * FF_BRICK];
    } else {
        c[0] = c[1] = c[2] = 0.0f;
    }
}

__device__ __forceinline__ uint16_t f2u16(float v) { return (uint16_t)(int)v; }

// interior: trilinear; weight of the lower corner (mesh_interleaved_meshcell.cpp:95-260)
__device__ void sample_interior(const MapRO &m, uint32_t ll, const int X[3], Sample &s)
{
    s.w = 0.0f; s.d = 0.0f; s.c[0] = s.c[1] = s.c[2] = 0;
    int guard = 0;
    for (uint32_t leaf = ll; leaf < FF_BRANCHINIT && s.w <= 0.0f && guard < 16; leaf = m.parent[leaf], guard++) {
        const ushort4 li = m.info[leaf];
        const int lo[3] = {li.x, li.y, li.z}, ls = li.w;
        int il[3], ih[3];
        float r[3];
        for (int k = 0; k < 3; k++) split_axis(X[k], lo[k], ls, il[k], r[k], ih[k]);
        const float rxn = 1.0f - r[0], ryn = 1.0f - r[1], rzn = 1.0f - r[2];
        const float f[8] = {rzn * ryn * rxn, rzn * ryn * r[0], rzn * r[1] * rxn, rzn * r[1] * r[0],
                            r[2] * ryn * rxn, r[2] * ryn * r[0], r[2] * r[1] * rxn, r[2] * r[1] * r[0]};
        float wf[8], df[8], cf[8][3];
#pragma unroll
        for (int i = 0; i < 8; i++)
            fetch(m, leaf, (i & 1) ? ih[0] : il[0], (i & 2) ? ih[1] : il[1], (i & 4) ? ih[2] : il[2], wf[i], df[i], cf[i]);
        s.w += wf[0];
        const float a0 = f[0] * df[0] + f[4] * df[4], a1 = f[1] * df[1] + f[5] * df[5];
        const float a2 = f[2] * df[2] + f[6] * df[6], a3 = f[3] * df[3] + f[7] * df[7];
        s.d = (a0 + a1) + (a2 + a3);
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            const float b0 = f[0] * cf[0][ch] + f[4] * cf[4][ch], b1 = f[1] * cf[1][ch] + f[5] * cf[5][ch];
            const float b2 = f[2] * cf[2][ch] + f[6] * cf[6][ch], b3 = f[3] * cf[3][ch] + f[7] * cf[7][ch];
            s.c[ch] = f2u16((b0 + b1) + (b2 + b3));
        }
    }
    s.lv = (uint16_t)guard;
}

// wall: thin axis a pointwise, u (inner) and v (outer) bilinear (:700-760)
__device__ void sample_wall(const MapRO &m, uint32_t ll, int a, int u, int v, const int X[3], Sample &s)
{
    s.w = 0.0f; s.d = 0.0f; s.c[0] = s.c[1] = s.c[2] = 0;
    int guard = 0;
    for (uint32_t leaf = ll; leaf < FF_BRANCHINIT && s.w <= 0.0f && guard < 16; leaf = m.parent[leaf], guard++) {
        const ushort4 li = m.info[leaf];
        const int lo[3] = {li.x, li.y, li.z}, ls = li.w;
        const int la = clamp7((X[a] - lo[a]) / ls);
        int ul, uh, vl, vh;
        float ru, rv;
        split_axis(X[u], lo[u], ls, ul, ru, uh);
        split_axis(X[v], lo[v], ls, vl, rv, vh);
        float wf[4], df[4], cf[4][3];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            int id[3];
            id[a] = la;
            id[u] = (i & 1) ? uh : ul;
            id[v] = (i & 2) ? vh : vl;
            fetch(m, leaf, id[0], id[1], id[2], wf[i], df[i], cf[i]);
        }
        const float ruI = 1.0f - ru, rvI = 1.0f - rv;
        s.w += rvI * ruI * wf[0] + rvI * ru * wf[1] + rv * ruI * wf[2] + rv * ru * wf[3];
        s.d = rvI * ruI * df[0] + rvI * ru * df[1] + rv * ruI * df[2] + rv * ru * df[3];
#pragma unroll
        for (int ch = 0; ch < 3; ch++)
            s.c[ch] = f2u16(rvI * ruI * cf[0][ch] + rvI * ru * cf[1][ch] + rv * ruI * cf[2][ch] + rv * ru * cf[3][ch]);
    }
    s.lv = (uint16_t)guard;
}

// edge: axis e linear, weight = max of the two (:1800-1860)
__device__ void sample_edge(const MapRO &m, uint32_t ll, int e, const int X[3], Sample &s)
{
    s.w = 0.0f; s.d = 0.0f; s.c[0] = s.c[1] = s.c[2] = 0;
    int guard = 0;
    for (uint32_t leaf = ll; leaf < FF_BRANCHINIT && s.w <= 0.0f && guard < 16; leaf = m.parent[leaf], guard++) {
        const ushort4 li = m.info[leaf];
        const int lo[3] = {li.x, li.y, li.z}, ls = li.w;
        int el, eh;
        float re;
        split_axis(X[e], lo[e], ls, el, re, eh);
        int i0[3], i1[3];
        for (int k = 0; k < 3; k++) i0[k] = i1[k] = clamp7((X[k] - lo[k]) / ls);
        i0[e] = el; i1[e] = eh;
        float w0, w1, d0, d1, c0[3], c1[3];
        fetch(m, leaf, i0[0], i0[1], i0[2], w0, d0, c0);
        fetch(m, leaf, i1[0], i1[1], i1[2], w1, d1, c1);
        s.w += (w0 < w1) ? w1 : w0;
        s.d = (1.0f - re) * d0 + re * d1;
#pragma unroll
        for (int ch = 0; ch < 3; ch++) s.c[ch] = f2u16((1.0f - re) * c0[ch] + re * c1[ch]);
    }
    s.lv = (uint16_t)guard;
}

// corner: pointwise (:2550-2575)
__device__ void sample_point(const MapRO &m, uint32_t ll, const int X[3], Sample &s)
{
    s.w = 0.0f; s.d = 0.0f; s.c[0] = s.c[1] = s.c[2] = 0;
    int guard = 0;
    for (uint32_t leaf = ll; leaf < FF_BRANCHINIT && s.w <= 0.0f && guard < 16; leaf = m.parent[leaf], guard++) {
        const ushort4 li = m.info[leaf];
        const int lo[3] = {li.x, li.y, li.z}, ls = li.w;
        float w, d, c[3];
        fetch(m, leaf, clamp7((X[0] - lo[0]) / ls), clamp7((X[1] - lo[1]) / ls), clamp7((X[2] - lo[2]) / ls), w, d, c);
        s.w += w;
        s.d = d;
        s.c[0] = (uint16_t)c[0]; s.c[1] = (uint16_t)c[1]; s.c[2] = (uint16_t)c[2];
    }
    s.lv = (uint16_t)guard;
}

__device__ __forceinline__ uint32_t wave_excl_scan_u32(uint32_t v, uint32_t *total)
{
    const int lane = threadIdx.x & 63;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    *total = __shfl(x, 63);
    return x - v;
}


// Geometry of a dirty cell, fixed when it is marked: origin, size, the
// sample spacing on full axes (sizeMin), the offset of the two sample planes
// on thin axes (stretch) and the sample grid n (0 when the cell meshes to
// nothing).  tile0 is the first 64-sample tile of its samples in the sample
// buffer, or FF_NO_TILE when the buffer was full and k_mc samples in place.
struct CellGeom {
    uint32_t cell;
    uint32_t tile0;
    uint16_t o[3];
    uint16_t size, sizeMin, stretch;
    uint8_t type, n[3];
    uint32_t pad;
};

#define FF_NO_TILE 0xffffffffu     // samples did not fit the buffer: sampled in place
#define FF_EMPTY_CELL 0xfffffffeu  // the cell meshes to nothing

// A 64-sample tile of a dirty cell, self-contained so that the sampling pass
// reads one record per tile (no cell or geometry lookups behind it): the
// cell's geometry, with g.pad = the tile's index within the cell and
// g.tile0 = FF_NO_TILE for a tile of a cell sampled in place, and its bricks.
struct TileRec {
    CellGeom g;  // 28 B
    uint32_t spare;
    uint32_t ll[8];
};
static_assert(sizeof(TileRec) == 64, "a tile record is four 16-byte words");

// The marking's slots for a whole block (every thread of the block calls it
// with its cell; it holds two barriers): the geometry slot of a dirty cell
// and the first of its nt tiles.  mc->nDirty counts the dirty cells.  split:
// cells of more than two tiles (interior cells) take geometry slots from
// the front of the array and the others from the back (mc->pad[2] and
// mc->pad[1] count them, taken together by one 64-bit atomic), so that the
// marching cubes, which takes dirty cells in slot order, starts the long
// cells first; otherwise slots in append order.  One atomic per counter
// per block, not per wave: the counters are shared by every block of the
// launch, and appends to one address serialise.
#define FF_MARK_THREADS 256  // block size of the marking kernels
// blocks of k_sample (256 threads each, striding over the tiles): a 1280x960
// frame at 2.5 mm leaves tens of thousands of tiles, whose sampling is a
// chain of dependent brick reads per lane, so more lanes in flight hide
// more of the latency (measured: 8192 blocks +3% frames/s over 4096 there
// and no change at 640x480; 1024 or 2048 lose 6-11% there, 16384 gains
// nothing more)
#define FF_SAMPLE_BLOCKS 8192

struct MarkSlot {
    uint32_t d, t0;
};

template <int NW>
__device__ __forceinline__ MarkSlot block_mark_slots(MeshCounters *mc, bool isDirty, uint32_t nt, uint32_t cap,
                                                     bool split)
{
    __shared__ uint32_t cnt[NW][3];  // per wave: long cells, short cells (or all dirty ones), tiles
    __shared__ uint32_t base[NW][3];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool big = split && isDirty && nt > 2u;
    const uint64_t mb = __ballot(big), ms = __ballot(isDirty && !big);
    uint32_t ntot;
    const uint32_t ex = wave_excl_scan_u32(nt, &ntot);
    if (lane == 0) {
        cnt[w][0] = (uint32_t)__popcll(mb);
        cnt[w][1] = (uint32_t)__popcll(ms);
        cnt[w][2] = ntot;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tb = 0, ts = 0, tt = 0;
        for (int k = 0; k < NW; k++) {
            base[k][0] = tb; base[k][1] = ts; base[k][2] = tt;
            tb += cnt[k][0]; ts += cnt[k][1]; tt += cnt[k][2];
        }
        uint32_t hi = 0, lo = 0, t = 0;
        if (tb + ts) {
            lo = atomicAdd(&mc->nDirty, tb + ts);  // append order when not split
            if (split) {
                const unsigned long long r = atomicAdd(reinterpret_cast<unsigned long long *>(&mc->pad[1]),
                                                       ((unsigned long long)tb << 32) | (unsigned long long)ts);
                lo = (uint32_t)r;
                hi = (uint32_t)(r >> 32);
            }
        }
        if (tt) t = atomicAdd(&mc->nTiles, tt);
        for (int k = 0; k < NW; k++) {
            base[k][0] += hi; base[k][1] += lo; base[k][2] += t;
        }
    }
    __syncthreads();
    MarkSlot r;
    r.t0 = base[w][2] + ex;
    if (!isDirty) r.d = 0xffffffffu;
    else if (big) r.d = base[w][0] + lanes_below(mb);
    else if (split) r.d = cap - 1u - (base[w][1] + lanes_below(ms));
    else r.d = base[w][1] + lanes_below(ms);
    __syncthreads();  // base is read before the next call writes it
    return r;
}

// the dirty cell of marching-cubes index di (slots as block_mark_slots gives them)
__device__ __forceinline__ uint32_t dirty_geom(uint32_t di, uint32_t nBig, uint32_t cap, bool split)
{
    return !split || di < nBig ? di : cap - 1u - (di - nBig);
}

// the nt tile records of a dirty cell from tile t0 (those within the buffer)
__device__ __forceinline__ void put_tiles(TileRec *tiles, uint32_t tileCap, uint32_t t0, uint32_t nt, const CellGeom &g,
                                          const uint32_t ll[8])
{
    for (uint32_t k = 0; k < nt && t0 + k < tileCap; k++) {
        TileRec r;
        r.g = g;
        r.g.pad = k;
        r.spare = 0;
        for (int j = 0; j < 8; j++) r.ll[j] = ll[j];
        tiles[t0 + k] = r;
    }
}

__device__ __forceinline__ bool axis_full(int type, int k)
{
    return type == 0 || (type <= 3 ? k != type - 1 : (type <= 6 ? k == type - 4 : false));
}

// sampling grid of a cell (per type, from the scales of its lastLeaf bricks;
// mesh_interleaved_meshcell.cpp interior / wall / edge / corner meshers)
template <typename Info>
__device__ __forceinline__ void cell_geom_t(uint64_t key, Info info, CellGeom &g)
{
    const int type = (int)(key & 7u);
    const int size = 1 << ((key >> 3) & 31u);
    const int o[3] = {(int)(key >> 48), (int)((key >> 32) & 0xffffu), (int)((key >> 16) & 0xffffu)};
    bool empty = false;
    int sizeMin = 1, stretch = 1;
    if (type == 0) {
        const ushort4 li = info(0);
        sizeMin = li.w;
        if (sizeMin > size || size / sizeMin > FF_BL) empty = true;
        const int lp[3] = {li.x, li.y, li.z}, ext = sizeMin * FF_BL;
        for (int k = 0; k < 3; k++)
            if (o[k] + size > lp[k] + ext || o[k] < lp[k]) empty = true;
    } else if (type <= 3) {
        const int s0 = info(0).w, s1 = info(1).w;
        sizeMin = min(s0, s1);
        stretch = s0;
        if (stretch > size) empty = true;
    } else if (type <= 6) {
        int sc[4];
        for (int k = 0; k < 4; k++) sc[k] = info(k).w;
        stretch = max(max(sc[0], sc[1]), sc[2]);
        sizeMin = min(min(sc[0], sc[1]), min(sc[2], sc[3]));
        if (stretch > size) empty = true;
    } else {
        stretch = info(0).w;
        for (int k = 1; k < 7; k++) stretch = max(stretch, (int)info(k).w);
        sizeMin = stretch;
        if (stretch > size) empty = true;
    }
    int n[3];
    for (int k = 0; k < 3; k++) {
        n[k] = axis_full(type, k) ? size / sizeMin : 2;
        if (n[k] > FF_BL) empty = true;  // cannot happen for a tree-derived cell; keeps LDS indexing in range
    }
    g.type = (uint8_t)type;
    for (int k = 0; k < 3; k++) {
        g.o[k] = (uint16_t)o[k];
        g.n[k] = empty ? 0 : (uint8_t)n[k];
    }
    g.size = (uint16_t)size;
    g.sizeMin = (uint16_t)sizeMin;
    g.stretch = (uint16_t)stretch;
}

// (the bricks' records read where they lie)
__device__ void cell_geom(const MapRO &m, uint64_t key, const uint32_t ll[8], CellGeom &g)
{
    cell_geom_t(key, [&](int k) { return m.info[ll[k]]; }, g);
}

__device__ __forceinline__ int geom_total(const CellGeom &g) { return (int)g.n[0] * g.n[1] * g.n[2]; }

// grid coordinate k of sample index gk: full axes step by sizeMin, thin axes
// take the two planes size - stretch and size past the origin
__device__ __forceinline__ int geom_coord(const CellGeom &g, int k, int gk)
{
    return axis_full(g.type, k) ? (int)g.o[k] + gk * g.sizeMin
                                : (gk ? (int)g.o[k] + g.size : (int)g.o[k] + g.size - g.stretch);
}

// sample s (raster order over n) of a cell; ll points at the cell's lastLeafs
__device__ __forceinline__ Sample sample_cell(const MapRO &m, const CellGeom &g, const uint32_t *ll, int s)
{
    const int n0 = g.n[0], n1 = g.n[1];
    const int i = s % n0, j = (s / n0) % n1, l = s / (n0 * n1);
    const int gg[3] = {i, j, l};
    int X[3];
    for (int k = 0; k < 3; k++) X[k] = geom_coord(g, k, gg[k]);
    Sample sm;
    const int type = g.type;
    if (type == 0) {
        const uint32_t l0 = ll[0];
        // full interior cell (the cell is its brick): the samples are the
        // brick's voxels; only zero-weight ones fall back along the chain
        if (g.size == g.sizeMin * FF_BL && n0 == FF_BL && n1 == FF_BL && g.n[2] == FF_BL) {
            const ushort4 li = m.info[l0];
            const size_t q = (size_t)l0 * FF_BRICK + s;
            const float w = m.wgt[q], d = m.dist[q];
            uint16_t c0 = 0, c1 = 0, c2 = 0;
            if (m.useColor) {
                const uint16_t *cb = m.col + (size_t)l0 * FF_BRICK * 3 + s;
                c0 = cb[0]; c1 = cb[FF_BRICK]; c2 = cb[2 * FF_BRICK];
            }
            if (li.x == g.o[0] && li.y == g.o[1] && li.z == g.o[2] && w > 0.0f) {
                sm.d = d; sm.w = w;
                sm.c[0] = c0; sm.c[1] = c1; sm.c[2] = c2;
                sm.lv = 1;
                return sm;
            }
        }
        sample_interior(m, l0, X, sm);
    } else if (type <= 3) {
        const int a = type - 1, u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
        sample_wall(m, ll[gg[a]], a, u, v, X, sm);
    } else if (type <= 6) {
        const int e = type - 4, u = e == 0 ? 1 : 0, v = e == 2 ? 1 : 2;
        sample_edge(m, ll[gg[u] + 2 * gg[v]], e, X, sm);
    } else {
        sample_point(m, ll[i + 2 * j + 4 * l], X, sm);
    }
    return sm;
}

__device__ __forceinline__ uint4 pack_sample(const Sample &s)
{
    return make_uint4(__float_as_uint(s.d), __float_as_uint(s.w), (uint32_t)s.c[0] | ((uint32_t)s.c[1] << 16),
                      (uint32_t)s.c[2]);
}

__device__ __forceinline__ void mark_body(CellRec *cells, const CellRec *__restrict__ prev,
                                          const uint64_t *__restrict__ hkeys, const uint32_t *__restrict__ hval,
                                          uint32_t hmask, bool havePrev, const uint32_t *__restrict__ stamp,
                                          const uint32_t *__restrict__ parent, uint32_t since, MeshCounters *mc,
                                          CellGeom *geom, uint32_t cap, const MapRO &m, TileRec *tiles,
                                          uint32_t tileCap, bool split, uint32_t updateId)
{
    const uint32_t nc = min(mc->nCells, cap);
    const uint32_t stride = gridDim.x * blockDim.x;
    // block-uniform iterations (the slots are assigned per block)
    for (uint32_t ib = blockIdx.x * blockDim.x; ib < nc; ib += stride) {
        const uint32_t i = ib + threadIdx.x;
        bool isDirty = false;
        CellGeom g;
        uint32_t ll[8];
        uint32_t nt = 0;
        if (i < nc) {
            CellRec &c = cells[i];
            int found = -1;
            if (havePrev) {
                uint32_t h = hkey_hash(c.key, hmask);
                for (uint32_t pr = 0; pr <= hmask; pr++) {
                    const uint64_t k = hkeys[h];
                    if (k == c.key) { found = (int)hval[h]; break; }
                    if (k == FF_HEMPTY) break;
                    h = (h + 1) & hmask;
                }
            }
            bool clean = found >= 0;
            if (clean) {
                const CellRec &p = prev[found];
                uint32_t cl[8];
                for (int k = 0; k < 8; k++) {
                    cl[k] = c.ll[k];
                    clean = clean && (p.ll[k] == cl[k]);
                }
                if (clean && chains_dirty(stamp, parent, cl, since, p.pad[1])) clean = false;
                if (clean) {
                    c.v0 = p.v0; c.nv = p.nv; c.f0 = p.f0; c.nf = p.nf;
                    c.pad[0] = p.pad[0];
                    c.pad[1] = p.pad[1];
                }
            }
            isDirty = !clean;
            if (isDirty) {
                c.pad[0] = updateId;  // the update that meshed it (CellUpdate)
                for (int k = 0; k < 8; k++) ll[k] = c.ll[k];
                cell_geom(m, c.key, ll, g);
                g.cell = i;
                g.pad = 0;
                nt = ((uint32_t)geom_total(g) + 63u) >> 6;
                c.pad[1] = nt ? 1u : 0u;  // chain depth: raised by the sampling
            }
        }
        const MarkSlot sl = block_mark_slots<FF_MARK_THREADS / 64>(mc, isDirty, nt, cap, split);
        const uint32_t d = sl.d;
        if (isDirty) {
            const uint32_t t0 = sl.t0;
            const bool fits = t0 + nt <= tileCap;
            g.tile0 = nt == 0 ? FF_EMPTY_CELL : (fits ? t0 : FF_NO_TILE);
            put_tiles(tiles, tileCap, t0, nt, g, ll);
            geom[d] = g;
        }
    }
}

__global__ __launch_bounds__(FF_MARK_THREADS) void k_mark(CellRec *cells, const CellRec *__restrict__ prev, const uint64_t *__restrict__ hkeys,
                       const uint32_t *__restrict__ hval, uint32_t hmask, bool havePrev,
                       const uint32_t *__restrict__ stamp, const uint32_t *__restrict__ parent, uint32_t since,
                       MeshCounters *mc, CellGeom *geom, uint32_t cap, MapRO m, TileRec *tiles, uint32_t tileCap,
                       uint32_t updateId)
{
    mark_body(cells, prev, hkeys, hval, hmask, havePrev, stamp, parent, since, mc, geom, cap, m, tiles, tileCap, true,
              updateId);
}

__global__ __launch_bounds__(FF_MARK_THREADS) void k_mark_b(const MeshBatchStream *__restrict__ S)
{
    const MeshBatchStream &a = S[blockIdx.z];
    mark_body(a.cells, a.prev, a.hkeys, a.hval, a.hmask, a.havePrev != 0, a.stamp, a.parent, a.since, a.mc, a.geom,
              a.cellCap, a.m, a.tiles, a.tileCap, false, a.updateId);
}

// An incremental update's table and marks: the previous table's cells whose
// owners no change touches, kept with their meshes and re-marked as
// k_mark_same re-marks them, then the fresh cells of the touched nodes,
// matched with the previous table by key as k_mark matches them; appended
// to the new table in that order (one atomic per block).  In full mode
// (inc_mode false: too many touched nodes) the table k_cells built, marked
// by k_mark's pass.
__global__ __launch_bounds__(FF_MARK_THREADS) void k_mark_inc(CellRec *cells, const CellRec *__restrict__ prev, uint32_t nPrev,
                           const uint64_t *__restrict__ hkeys, const uint32_t *__restrict__ hval, uint32_t hmask,
                           const uint32_t *__restrict__ stamp, const uint32_t *__restrict__ parent, uint32_t since,
                           MeshCounters *mc, CellGeom *geom, uint32_t cap, MapRO m, TileRec *tiles, uint32_t tileCap,
                           uint32_t updateId, IncView iv, const ushort4 *__restrict__ binfo,
                           const CellRec *__restrict__ fresh, const uint32_t *__restrict__ incCnt)
{
    if (!inc_mode(true, incCnt)) {
        mark_body(cells, prev, hkeys, hval, hmask, true, stamp, parent, since, mc, geom, cap, m, tiles, tileCap, true,
                  updateId);
        return;
    }
    __shared__ uint4 box[FF_INC_BOXES];
    const uint32_t nBox = load_changes(iv, m.info, binfo, box);
    const uint32_t nItems = nPrev + min(incCnt[1], FF_INC_NODES * 27u);
    const uint32_t stride = gridDim.x * blockDim.x;
    // block-uniform iterations (the table and marking slots are per block)
    for (uint32_t ib = blockIdx.x * blockDim.x; ib < nItems; ib += stride) {
        const uint32_t i = ib + threadIdx.x;
        bool keep = false, isDirty = false;
        CellRec c;
        // (the bricks' records, requested beside the stamps: a dirty cell's
        // geometry needs them)
        ushort4 li[8];
        if (i < nPrev) {
            c = prev[i];
#pragma unroll
            for (int k = 0; k < 8; k++) li[k] = m.info[c.ll[k] < FF_BRANCHINIT ? c.ll[k] : 0u];
            keep = !cell_owner_touched(c.key, c.pad[2], box, nBox);
            if (keep) isDirty = chains_dirty(stamp, parent, c.ll, since, c.pad[1]);
        } else if (i < nItems) {
            c = fresh[i - nPrev];
#pragma unroll
            for (int k = 0; k < 8; k++) li[k] = m.info[c.ll[k] < FF_BRANCHINIT ? c.ll[k] : 0u];
            keep = true;
            int found = -1;
            uint32_t h = hkey_hash(c.key, hmask);
            for (uint32_t pr = 0; pr <= hmask; pr++) {
                const uint64_t k = hkeys[h];
                if (k == c.key) { found = (int)hval[h]; break; }
                if (k == FF_HEMPTY) break;
                h = (h + 1) & hmask;
            }
            bool clean = found >= 0;
            if (clean) {
                const CellRec &p = prev[found];
                for (int k = 0; k < 8; k++) clean = clean && (p.ll[k] == c.ll[k]);
                if (clean && chains_dirty(stamp, parent, c.ll, since, p.pad[1])) clean = false;
                if (clean) {
                    c.v0 = p.v0; c.nv = p.nv; c.f0 = p.f0; c.nf = p.nf;
                    c.pad[0] = p.pad[0];
                    c.pad[1] = p.pad[1];
                }
            }
            isDirty = !clean;
        }
        const uint32_t j = block_append<FF_MARK_THREADS / 64>(&mc->nCells, keep);
        CellGeom g;
        uint32_t nt = 0;
        if (isDirty) {
            // a fresh cell's record has an empty mesh range; a kept one's is
            // cleared, like k_mark_same's (k_mc fills it unless the mesh is empty)
            c.v0 = c.nv = c.f0 = c.nf = 0u;
            c.pad[0] = updateId;  // the update that meshed it (CellUpdate)
            cell_geom_t(c.key, [&](int k) { return li[k]; }, g);
            g.cell = j;
            g.pad = 0;
            nt = ((uint32_t)geom_total(g) + 63u) >> 6;
            c.pad[1] = nt ? 1u : 0u;  // chain depth: raised by the sampling
        }
        if (keep) {
            if (j < cap) cells[j] = c;
            else atomicOr(&mc->flags, FF_CELL_OVERFLOW);
        }
        const MarkSlot sl = block_mark_slots<FF_MARK_THREADS / 64>(mc, isDirty && j < cap, nt, cap, true);
        if (isDirty && j < cap) {
            const uint32_t t0 = sl.t0;
            const bool fits = t0 + nt <= tileCap;
            g.tile0 = nt == 0 ? FF_EMPTY_CELL : (fits ? t0 : FF_NO_TILE);
            put_tiles(tiles, tileCap, t0, nt, g, c.ll);
            geom[sl.d] = g;
        }
    }
}

// The table of the previous update re-marked in place, when the tree has not
// changed since it was built: every cell then exists with the same bricks,
// so it is clean exactly when none of its bricks' parent chains was updated
// (k_mark's test with the previous table's own record).  A dirty cell's mesh
// range is cleared first, like the fresh record k_cells writes: k_mc fills
// it unless the new mesh is empty.
__global__ __launch_bounds__(FF_MARK_THREADS) void k_mark_same(CellRec *cells, const uint32_t *__restrict__ stamp, const uint32_t *__restrict__ parent,
                            uint32_t since, MeshCounters *mc, uint32_t nCells, MeshCounters *clearNext, CellGeom *geom,
                            uint32_t cap, MapRO m, TileRec *tiles, uint32_t tileCap, uint32_t updateId)
{
    const bool split = true;  // long cells first (block_mark_slots)
    // the counters start at zero; block 0 records the table's size and
    // clears the next update's counters
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        mc->nCells = nCells;
        *clearNext = MeshCounters{};
    }
    const uint32_t nc = min(nCells, cap);
    const uint32_t stride = gridDim.x * blockDim.x;
    // block-uniform iterations (the slots are assigned per block)
    for (uint32_t ib = blockIdx.x * blockDim.x; ib < nc; ib += stride) {
        const uint32_t i = ib + threadIdx.x;
        bool isDirty = false;
        CellGeom g;
        uint32_t ll[8];
        uint32_t nt = 0;
        if (i < nc) {
            CellRec &c = cells[i];
            for (int k = 0; k < 8; k++) ll[k] = c.ll[k];
            // the bricks' records requested beside their stamps (a dirty
            // cell's geometry needs them; behind the test they were one
            // more dependent round trip)
            ushort4 li[8];
#pragma unroll
            for (int k = 0; k < 8; k++) li[k] = m.info[ll[k] < FF_BRANCHINIT ? ll[k] : 0u];
            isDirty = chains_dirty(stamp, parent, ll, since, c.pad[1]);
            if (isDirty) {
                c.v0 = c.nv = c.f0 = c.nf = 0u;
                c.pad[0] = updateId;
                cell_geom_t(c.key, [&](int k) { return li[k]; }, g);
                g.cell = i;
                g.pad = 0;
                nt = ((uint32_t)geom_total(g) + 63u) >> 6;
                c.pad[1] = nt ? 1u : 0u;  // chain depth: raised by the sampling
            }
        }
        const MarkSlot sl = block_mark_slots<FF_MARK_THREADS / 64>(mc, isDirty, nt, cap, split);
        const uint32_t d = sl.d;
        if (isDirty) {
            const uint32_t t0 = sl.t0;
            const bool fits = t0 + nt <= tileCap;
            g.tile0 = nt == 0 ? FF_EMPTY_CELL : (fits ? t0 : FF_NO_TILE);
            put_tiles(tiles, tileCap, t0, nt, g, ll);
            geom[d] = g;
        }
    }
}

// arena (when not null): block 0 also writes the totals of the arena
// counters into mc->pad (vertices, faces allocated since the last compaction)
__device__ __forceinline__ void hash_build_body(CellRec *cells, MeshCounters *mc, uint32_t cap, uint64_t *hkeys,
                                                uint32_t *hval, uint32_t hmask, const unsigned long long *arena)
{
    if (arena && blockIdx.x == 0 && threadIdx.x < 64) {
        const int lane = threadIdx.x;
        const unsigned long long a = lane < (int)FF_ARENAS ? arena[lane * FF_ARENA_STRIDE] : 0ull;
        uint32_t v = (uint32_t)a, f = (uint32_t)(a >> 32);
        for (int o = 32; o > 0; o >>= 1) {
            v += __shfl_xor(v, o);
            f += __shfl_xor(f, o);
        }
        if (lane == 0) { mc->pad[0] = v; mc->pad[1] = f; }
    }
    const uint32_t nc = min(mc->nCells, cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x) {
        const uint64_t key = cells[i].key;
        uint32_t h = hkey_hash(key, hmask);
        for (uint32_t pr = 0; pr <= hmask; pr++) {
            const unsigned long long prevk =
                atomicCAS((unsigned long long *)&hkeys[h], (unsigned long long)FF_HEMPTY, (unsigned long long)key);
            if (prevk == FF_HEMPTY || prevk == key) {
                hval[h] = i;
                cells[i].hslot = h;
                break;
            }
            h = (h + 1) & hmask;
        }
    }
}


__global__ void k_hash_build_b(const MeshBatchStream *__restrict__ S)
{
    const MeshBatchStream &a = S[blockIdx.z];
    hash_build_body(a.cells, a.mc, a.cellCap, a.hkeys, a.hval, a.hmask, nullptr);
}

// empties the slots of a whole table (by recorded slot, so concurrent
// clears cannot cut each other's probe chains)
__device__ __forceinline__ void hash_clear_cells_body(const CellRec *__restrict__ cells, uint32_t nc, uint64_t *hkeys)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x) {
        const uint32_t h = cells[i].hslot;
        if (h != 0xffffffffu) hkeys[h] = FF_HEMPTY;
    }
}


// the turnover of an update that built a new table: the previous table's
// slots emptied in its hash, the new table's keys into the other (empty)
// hash; the two parts touch different memory, so one pass does both
// (incCnt: the incremental counters, zero for the next update)
__global__ void k_hash_turnover(const CellRec *__restrict__ prev, uint32_t nPrev, uint64_t *prevKeys, CellRec *cells,
                                MeshCounters *mc, uint32_t cap, uint64_t *hkeys, uint32_t *hval, uint32_t hmask,
                                const unsigned long long *arena, uint32_t *incCnt)
{
    if (blockIdx.x == 0 && threadIdx.x < 2) incCnt[threadIdx.x] = 0u;
    hash_clear_cells_body(prev, nPrev, prevKeys);
    hash_build_body(cells, mc, cap, hkeys, hval, hmask, arena);
}

// (the previous table of each stream)
__global__ void k_hash_clear_cells_b(const MeshBatchStream *__restrict__ S)
{
    const MeshBatchStream &a = S[blockIdx.z];
    hash_clear_cells_body(a.prev, a.nPrev, a.hkeys);
}


// ---------------------------------------------------------------------------
// k_sample: one thread per sample of the dirty cells' 64-sample tiles
// ---------------------------------------------------------------------------
// the deepest parent-chain level any sample of a tile read, raised into
// its cell's record (CellRec::pad[1], set to 1 by the marking)
__device__ __forceinline__ void raise_depth(CellRec *cells, uint32_t cell, uint32_t lv)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) lv = max(lv, (uint32_t)__shfl_xor((int)lv, o));
    if ((threadIdx.x & 63u) == 0u && lv > 1u) atomicMax(&cells[cell].pad[1], lv);
}

// (the grid's stride is a multiple of 64, so a wave's lanes are one tile)
__device__ __forceinline__ void sample_body(const CellGeom *__restrict__ geom, CellRec *__restrict__ cells,
                                            const TileRec *__restrict__ tiles, const MeshCounters *mc,
                                            uint32_t tileCap, const MapRO &m, uint4 *__restrict__ sbuf)
{
    const uint32_t nt = min(mc->nTiles, tileCap);
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nt * 64u; t += gridDim.x * blockDim.x) {
        // one record: the geometry and the bricks (read where they lie, by
        // the index the sample picks); the geometry's two 16-byte words are
        // requested together, before any test waits on a field (behind an
        // early exit for the in-place tiles they were three dependent round
        // trips)
        const TileRec *r = tiles + (t >> 6);
        const uint4 *gw = reinterpret_cast<const uint4 *>(r);
        const uint4 w0 = gw[0], w1 = gw[1];  // CellGeom (28 B) and the spare word
        // (an empty statement that takes every word: the compiler keeps the
        // two loads here instead of splitting them around the tests below)
        asm volatile("" ::"v"(w0.x), "v"(w0.y), "v"(w0.z), "v"(w0.w), "v"(w1.x), "v"(w1.y), "v"(w1.z));
        static_assert(sizeof(CellGeom) == 28, "the geometry is the record's first 28 bytes");
        CellGeom g;
        memcpy(&g, &w0, 16);
        memcpy(reinterpret_cast<unsigned char *>(&g) + 16, &w1, 12);
        const int s = (int)(g.pad * 64u + (t & 63u));
        uint32_t lv = 0;
        // (a tile of a cell sampled in place is skipped; both tests taken
        // together, so that both words are waited for once)
        if ((g.tile0 != FF_NO_TILE) & (s < geom_total(g))) {
            const Sample sm = sample_cell(m, g, r->ll, s);
            sbuf[t] = pack_sample(sm);
            lv = sm.lv;
        }
        raise_depth(cells, g.cell, lv);
    }
}

// rep: where the update's tile count and then its number go (mapped host
// memory), so that the host knows whether the next frame's SDF update,
// which follows this sampling, stands down (cells whose samples did not fit
// the buffer are sampled from the bricks by the marching cubes)
__global__ __launch_bounds__(256) void k_sample(const CellGeom *__restrict__ geom, CellRec *__restrict__ cells,
                                                const TileRec *__restrict__ tiles, const MeshCounters *mc,
                                                uint32_t tileCap, MapRO m, uint4 *__restrict__ sbuf,
                                                uint32_t *__restrict__ rep, uint32_t updateId)
{
    if (rep && blockIdx.x == 0 && threadIdx.x == 0) {
        rep[0] = mc->nTiles;
        __threadfence_system();
        *reinterpret_cast<volatile uint32_t *>(&rep[1]) = updateId;
    }
    sample_body(geom, cells, tiles, mc, tileCap, m, sbuf);
}

// the update's counters to mapped host memory once its marching cubes are
// done, then (after a system-scope fence) its number: a host that sees the
// number sees the counters
__global__ void k_mesh_report(const MeshCounters *__restrict__ mc, MeshCounters *hmc, uint32_t *seq, uint32_t id)
{
    if (threadIdx.x == 0) {
        const uint4 *src = reinterpret_cast<const uint4 *>(mc);
        uint4 *dst = reinterpret_cast<uint4 *>(hmc);
        const uint4 a = src[0], b = src[1];
        dst[0] = a;
        dst[1] = b;
        __threadfence_system();
        *reinterpret_cast<volatile uint32_t *>(seq) = id;
    }
}

__global__ __launch_bounds__(256) void k_sample_b(const MeshBatchStream *__restrict__ S)
{
    const MeshBatchStream &a = S[blockIdx.z];
    sample_body(a.geom, a.cells, a.tiles, a.mc, a.tileCap, a.m, a.sbuf);
}

// ---------------------------------------------------------------------------
// k_mc: one wave per dirty cell -- cube classification, edge numbering,
// vertices and faces
// ---------------------------------------------------------------------------
// VertexInterp, double (mesh.cpp:1266-1300)
__device__ __forceinline__ void vinterp(const float p1[3], const float p2[3], float v1f, float v2f, float out[3])
{
    const double v1 = v1f, v2 = v2f;
    if (fabs(0.0 - v1) < 1e-40) { out[0] = p1[0]; out[1] = p1[1]; out[2] = p1[2]; return; }
    if (fabs(0.0 - v2) < 1e-40) { out[0] = p2[0]; out[1] = p2[1]; out[2] = p2[2]; return; }
    if (fabs(v1 - v2) < 1e-40) { out[0] = p1[0]; out[1] = p1[1]; out[2] = p1[2]; return; }
    const double mu = __ddiv_rn(0.0 - v1, v2 - v1);
#pragma unroll
    for (int k = 0; k < 3; k++) out[k] = (float)__dadd_rn((double)p1[k], __dmul_rn(mu, __dsub_rn((double)p2[k], (double)p1[k])));
}

// VertexInterp for colors, float with clamping (mesh.cpp:1302-1323), then / 256
__device__ __forceinline__ uint32_t cinterp(const uint16_t c1[3], const uint16_t c2[3], float v1, float v2)
{
    uint16_t r[3];
    if ((double)fabsf(0.0f - v1) < 1e-40) { r[0] = c1[0]; r[1] = c1[1]; r[2] = c1[2]; }
    else if ((double)fabsf(0.0f - v2) < 1e-40) { r[0] = c2[0]; r[1] = c2[1]; r[2] = c2[2]; }
    else if ((double)fabsf(v1 - v2) < 1e-40) { r[0] = c1[0]; r[1] = c1[1]; r[2] = c1[2]; }
    else {
        const float mu = __fdiv_rn(0.0f - v1, v2 - v1);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float p = (float)c1[k] + mu * ((float)c2[k] - (float)c1[k]);
            r[k] = p <= 65280.0f ? (p >= 0.0f ? (uint16_t)(int)p : (uint16_t)0) : (uint16_t)65280;
        }
    }
    return (uint32_t)(r[0] / 256) | ((uint32_t)(r[1] / 256) << 8) | ((uint32_t)(r[2] / 256) << 16);
}

// edge e of a cube as an offset of used-edge slots from the cube's first
// sample: 3 (dz * 64 + dy * 8 + dx) + axis, where (dx, dy, dz) is the edge's
// lower end and axis its direction (Bourke edges e0-e11), 8 bits per edge,
// in registers rather than a table
__device__ __forceinline__ int edge_slot_off(int e)
{
    constexpr uint32_t E0 = 0u | (4u << 8) | (24u << 16) | (1u << 24);       // e0..e3
    constexpr uint32_t E1 = 192u | (196u << 8) | (216u << 16) | (193u << 24);  // e4..e7
    constexpr uint32_t E2 = 2u | (5u << 8) | (29u << 16) | (26u << 24);      // e8..e11
    const uint32_t w = e < 4 ? E0 : (e < 8 ? E1 : E2);
    return (int)((w >> (8 * (e & 3))) & 255u);
}

__device__ __forceinline__ int tri_edge(uint64_t T, int t) { return (int)((T >> (4 * t)) & 15u); }
__device__ __forceinline__ int tri_count(uint64_t T) { return (int)(T >> 60); }

// q / d for q < 512 and 1 <= d < 128 by a multiply with m = ceil(2^16 / d)
// (the error q (m d - 2^16) / (d 2^16) < 1/128 stays below the gap 1/d)
__device__ __forceinline__ uint32_t div_small(uint32_t q, uint32_t m) { return (q * m) >> 16; }

struct WaveLds {
    float d[FF_BRICK];
    uint16_t c[3][FF_BRICK];
    // sample flags as bit rows: byte (l * FF_BL + j) holds the row of 8
    // samples (., j, l), bit i for sample i; neg: distance < 0, zero: zero
    // weight (words 0-15 neg, 16-31 zero)
    uint32_t flags[32];
    uint32_t used[48];     // 1536 edge slots: (sample * 3 + axis)
    uint32_t wpref[48];    // vertex index of the first slot of each word
    uint16_t cube[344];    // cube index | triangle count << 8
    uint16_t foff[344];    // face offset per cube
    uint16_t list[1344];   // used slots in vertex order (at most 3 * 7 * 8 * 8 edges)
};

// sample s of a cell with n0 x n1 samples per layer, n0 = 1 << lg0 and
// n1 = 1 << lg1 (cell sample counts are powers of two)
__device__ __forceinline__ void put_sample(WaveLds &L, int lg0, int lg1, int s, float d, float w, uint32_t c01,
                                           uint32_t c2)
{
    const int i = s & ((1 << lg0) - 1), j = (s >> lg0) & ((1 << lg1) - 1), l = s >> (lg0 + lg1);
    const int id = (l * FF_BL + j) * FF_BL + i;
    L.d[id] = d;
    L.c[0][id] = (uint16_t)(c01 & 0xffffu);
    L.c[1][id] = (uint16_t)(c01 >> 16);
    L.c[2][id] = (uint16_t)c2;
    if (d < 0.0f) atomicOr(&L.flags[id >> 5], 1u << (id & 31));
    if (w == 0.0f) atomicOr(&L.flags[16 + (id >> 5)], 1u << (id & 31));
}

// byte r of the neg (zero = false) or zero-weight (zero = true) bit rows
__device__ __forceinline__ uint32_t flag_row(const WaveLds &L, bool zero, int r)
{
    return reinterpret_cast<const uint8_t *>(L.flags)[(zero ? 64 : 0) + r];
}

// Bourke corner k of a cube, coordinate a: x = bit 0 ^ bit 1, y = bit 1,
// z = bit 2
__device__ __forceinline__ int corner_c(int k, int a) { return a == 0 ? (k ^ (k >> 1)) & 1 : (k >> a) & 1; }
// sample index of corner k from the cube's first sample
__device__ __forceinline__ int corner_id(int k) { return (corner_c(k, 2) * FF_BL + corner_c(k, 1)) * FF_BL + corner_c(k, 0); }

// the two corners of Bourke edge e (0-11), 4 bits each in a constant
__device__ __forceinline__ int edge_corner(int e, int end)
{
    constexpr uint64_t EA = 0x321076543210ull;  // first corners of e11..e0, 4 bits each
    constexpr uint64_t EB = 0x765447650321ull;  // second corners
    return (int)(((end ? EB : EA) >> (4 * e)) & 15u);
}

// bits 0..7 of x moved to bits 0, 3, ..., 21
__device__ __forceinline__ uint32_t spread3(uint32_t x)
{
    uint32_t r = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) r |= ((x >> i) & 1u) << (3 * i);
    return r;
}

// The used-edge bitmap (slot = sample * 3 + axis, axes x, y, z) from the
// flag rows, lane r for sample row r = l * 8 + j: an edge is used by some
// triangle exactly when its two samples differ in sign and a cube that
// contains it has no zero-weight corner (each triangle-table row uses
// exactly the cut edges of its cube; a cube with a zero-weight corner
// emits nothing).  Row r's 24 slots are bits 24 r .. 24 r + 23; lane w
// assembles word w from rows (32 w) / 24 and the next.
__device__ __forceinline__ void used_edges(WaveLds &L, int lane, int nc0, int nc1, int nc2)
{
    const int r = lane, j = r & 7, l = r >> 3;
    // cubes of row (j, l) with no zero-weight corner, bits i < nc0
    uint32_t V = 0u;
    if (j < nc1 && l < nc2) {
        const uint32_t z = flag_row(L, true, r) | flag_row(L, true, r + 1) | flag_row(L, true, r + FF_BL) |
                           flag_row(L, true, r + FF_BL + 1);
        V = ~(z | (z >> 1)) & ((1u << nc0) - 1u);
    }
    const uint32_t Vy = __shfl_up(V, 1), Vz = __shfl_up(V, FF_BL), Vyz = __shfl_up(V, FF_BL + 1);
    const uint32_t vj = j > 0 ? Vy : 0u, vl = l > 0 ? Vz : 0u, vjl = (j > 0 && l > 0) ? Vyz : 0u;
    const uint32_t n0 = flag_row(L, false, r);
    const uint32_t ny = j < FF_BL - 1 ? flag_row(L, false, r + 1) : 0u;
    const uint32_t nz = l < FF_BL - 1 ? flag_row(L, false, r + FF_BL) : 0u;
    const uint32_t ux = (n0 ^ (n0 >> 1)) & (V | vj | vl | vjl);
    const uint32_t wy = V | vl, wz = V | vj;
    const uint32_t uy = (n0 ^ ny) & (wy | (wy << 1));
    const uint32_t uz = (n0 ^ nz) & (wz | (wz << 1));
    const uint32_t chunk = spread3(ux & 0xffu) | (spread3(uy & 0xffu) << 1) | (spread3(uz & 0xffu) << 2);
    const int w = lane < 48 ? lane : 47, r0 = (32 * w) / 24, o = 32 * w - 24 * r0;
    const uint32_t a = (uint32_t)__shfl((int)chunk, r0), b = (uint32_t)__shfl((int)chunk, min(r0 + 1, 63));
    if (lane < 48) L.used[lane] = (a >> o) | (b << (24 - o));
}

// One dirty cell, by one wave (lane 0..63) with its LDS slice: samples into
// LDS, cube classification, edge numbering, vertices and faces.  di is the
// cell's dirty index (it picks the pool arena); act false for a wave without
// a cell (it still takes part in the wave-wide steps).
// Cells whose samples k_sample wrote to the buffer read them from there;
// cells that did not fit the buffer are sampled by the wave itself.
__device__ void mc_cell(WaveLds &L, const uint64_t *tri, int lane, CellRec *cells, const CellGeom &g, MeshCounters *mc,
                        const uint4 *__restrict__ sbuf, const MapRO &m, const Pools &P, uint32_t di, bool act)
{
    // the wave's cell is uniform: skip cells of the other pass (empty
    // cells keep the zero counts k_cells gave them)
    if (g.tile0 == FF_EMPTY_CELL) return;
    const int type = g.type;
    const int n[3] = {g.n[0], g.n[1], g.n[2]};
    const int total = n[0] * n[1] * n[2];
    const int lg0 = __ffs(max(n[0], 1)) - 1, lg1 = __ffs(max(n[1], 1)) - 1;
    if (lane < 32) L.flags[lane] = 0u;
    wave_sync();  // the samples OR their flags in
    if (g.tile0 != FF_NO_TILE) {
        // 1 <= total <= 512 (a cell with tiles has samples; n <= FF_BL per
        // axis): every lane's loads are in flight together before the first
        // LDS store waits on one; lanes past the end re-read the last sample
        // so that no load sits behind a branch
        const uint4 *src = sbuf + (size_t)g.tile0 * 64u;
        uint4 v[FF_BRICK / 64];
#pragma unroll
        for (int k = 0; k < FF_BRICK / 64; k++) v[k] = src[min(lane + 64 * k, total - 1)];
        if (total == FF_BRICK) {
            // the full 8 x 8 x 8 grid: sample s is bit s of the flag rows, so
            // the wave's 64 samples give two words of each by a ballot (in
            // place of an LDS atomic per sample, 32 lanes on one word)
#pragma unroll
            for (int k = 0; k < FF_BRICK / 64; k++) {
                const int id = lane + 64 * k;
                const float d = __uint_as_float(v[k].x), w = __uint_as_float(v[k].y);
                L.d[id] = d;
                L.c[0][id] = (uint16_t)(v[k].z & 0xffffu);
                L.c[1][id] = (uint16_t)(v[k].z >> 16);
                L.c[2][id] = (uint16_t)v[k].w;
                const uint64_t nb = __ballot(d < 0.0f), zb = __ballot(w == 0.0f);
                if (lane == 0) {
                    L.flags[2 * k] = (uint32_t)nb;
                    L.flags[2 * k + 1] = (uint32_t)(nb >> 32);
                    L.flags[16 + 2 * k] = (uint32_t)zb;
                    L.flags[16 + 2 * k + 1] = (uint32_t)(zb >> 32);
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < FF_BRICK / 64; k++)
                if (lane + 64 * k < total)
                    put_sample(L, lg0, lg1, lane + 64 * k, __uint_as_float(v[k].x), __uint_as_float(v[k].y), v[k].z,
                               v[k].w);
        }
    } else {
        const uint32_t *ll = cells[g.cell].ll;
        uint32_t lv = 0;
        for (int s = lane; s < total; s += 64) {
            const Sample sm = sample_cell(m, g, ll, s);
            const uint4 v = pack_sample(sm);
            put_sample(L, lg0, lg1, s, __uint_as_float(v.x), __uint_as_float(v.y), v.z, v.w);
            lv = max(lv, (uint32_t)sm.lv);
        }
        if (act) raise_depth(cells, g.cell, lv);
    }
    if (lane < 48) L.used[lane] = 0u;
    wave_sync();
    CellRec *cp = act ? &cells[g.cell] : nullptr;
    const int sizeMin = g.sizeMin, stretch = g.stretch;
    bool full[3];
    for (int k = 0; k < 3; k++) full[k] = axis_full(type, k);
    const int elen[3] = {full[0] ? sizeMin : stretch, full[1] ? sizeMin : stretch, full[2] ? sizeMin : stretch};
    uint32_t nv = 0, nf = 0;
    if (type == 7) {
        // non-indexed corner cell with degenerate triangles dropped
        // (getColoredTrianglesForCubePrecomputed, mesh.cpp:1931-2008)
        int ci = 0;
        bool veto = false;
        for (int k = 0; k < 8; k++) {
            const int r = corner_c(k, 2) * FF_BL + corner_c(k, 1), x = corner_c(k, 0);
            if (total) {
                if ((flag_row(L, true, r) >> x) & 1u) veto = true;
                if ((flag_row(L, false, r) >> x) & 1u) ci |= 1 << k;
            }
        }
        float basev[3];
        const float sx = m.scale * (float)stretch;
        for (int a = 0; a < 3; a++) basev[a] = m.off[a] + (float)geom_coord(g, a, 0) * m.scale;
        if (veto || !total) ci = 0;
        const uint64_t T = tri[ci];
        const int ntri = tri_count(T);
        bool keep = false;
        float V[3][3];
        uint32_t C[3];
        if (lane < ntri) {
            for (int q = 0; q < 3; q++) {
                const int e = tri_edge(T, 3 * lane + q);
                const int a = edge_corner(e, 0), b = edge_corner(e, 1);
                const int ia = corner_id(a), ib = corner_id(b);
                float Pa[3], Pb[3];
                for (int k = 0; k < 3; k++) {
                    Pa[k] = corner_c(a, k) ? basev[k] + 1.0f * sx : basev[k] + 0.0f * sx;
                    Pb[k] = corner_c(b, k) ? basev[k] + 1.0f * sx : basev[k] + 0.0f * sx;
                }
                vinterp(Pa, Pb, L.d[ia], L.d[ib], V[q]);
                const uint16_t ca[3] = {L.c[0][ia], L.c[1][ia], L.c[2][ia]};
                const uint16_t cb[3] = {L.c[0][ib], L.c[1][ib], L.c[2][ib]};
                C[q] = cinterp(ca, cb, L.d[ia], L.d[ib]);
            }
            keep = true;
            for (int q = 0; q < 3; q++) {
                const float *A = V[q], *B = V[(q + 1) % 3];
                if (A[0] == B[0] && A[1] == B[1] && A[2] == B[2]) keep = false;
            }
        }
        const uint64_t kb = __ballot(keep);
        const uint32_t rank = lanes_below(kb);
        nf = __popcll(kb);
        nv = 3 * nf;
        uint32_t v0 = 0, f0 = 0, ok = 1;
        if (lane == 0) ok = pool_alloc(P, di % FF_ARENAS, nv, nf, v0, f0, mc) ? 1u : 0u;
        v0 = __shfl(v0, 0);
        f0 = __shfl(f0, 0);
        const bool fits = __shfl(ok, 0) != 0;
        if (keep && fits) {
            for (int q = 0; q < 3; q++) {
                P.v[v0 + 3 * rank + q] = make_float4(V[q][0], V[q][1], V[q][2], __uint_as_float(C[q]));
                P.f[3 * (f0 + rank) + q] = 3 * rank + q;
            }
        }
        if (lane == 0) {
            if (!fits) nv = nf = 0;
            cp->v0 = v0; cp->nv = nv; cp->f0 = f0; cp->nf = nf;
        }
        wave_sync();
        return;
    }
    // -- cubes: index, used edges, triangle counts
    const int nc0 = max(n[0] - 1, 0), nc1 = max(n[1] - 1, 0), nc2 = max(n[2] - 1, 0);
    const int ncubes = nc0 * nc1 * nc2;
    // q -> (i, j, l) by multiplies (nc0, nc1 are 1, 3 or 7; q < 343)
    const uint32_t m0 = nc0 ? (65536u + (uint32_t)nc0 - 1u) / (uint32_t)nc0 : 0u;
    const uint32_t m01 = nc0 * nc1 ? (65536u + (uint32_t)(nc0 * nc1) - 1u) / (uint32_t)(nc0 * nc1) : 0u;
    for (int q = lane; q < ncubes; q += 64) {
        const int jl = (int)div_small((uint32_t)q, m0), l = (int)div_small((uint32_t)q, m01);
        const int i = q - jl * nc0, j = jl - l * nc1;
        // the cube's corners are bits i, i + 1 of rows r, r + 1 (y + 1),
        // r + 8 (z + 1) and r + 9
        const int r = l * FF_BL + j;
        const uint32_t n00 = flag_row(L, false, r) >> i, n10 = flag_row(L, false, r + 1) >> i;
        const uint32_t n01 = flag_row(L, false, r + FF_BL) >> i, n11 = flag_row(L, false, r + FF_BL + 1) >> i;
        const uint32_t zw = (flag_row(L, true, r) | flag_row(L, true, r + 1) | flag_row(L, true, r + FF_BL) |
                             flag_row(L, true, r + FF_BL + 1)) >> i;
        int ci = (int)((n00 & 1u) | (n00 & 2u) | ((n10 & 2u) << 1) | ((n10 & 1u) << 3) | ((n01 & 1u) << 4) |
                       ((n01 & 2u) << 4) | ((n11 & 2u) << 5) | ((n11 & 1u) << 7));
        if (zw & 3u) ci = 0;
        const uint64_t T = tri[ci];
        const int nt = tri_count(T);
        L.cube[q] = (uint16_t)(ci | (nt << 8));
    }
    used_edges(L, lane, nc0, nc1, nc2);
    wave_sync();
    // -- vertex numbering over the used-edge bitmap, and the slot list
    {
        const uint32_t wd = lane < 48 ? L.used[lane] : 0u;
        uint32_t tot;
        const uint32_t ex = wave_excl_scan_u32((uint32_t)__popc(wd), &tot);
        if (lane < 48) {
            L.wpref[lane] = ex;
            uint32_t bits = wd, k = ex;
            while (bits) {
                const uint32_t b = (uint32_t)__ffs(bits) - 1u;
                L.list[k++] = (uint16_t)(lane * 32 + b);
                bits &= bits - 1u;
            }
        }
        nv = tot;
    }
    // face offsets per cube: contiguous chunks of 6 cubes per lane
    {
        const int per = 6;
        uint32_t loc = 0;
        for (int q = lane * per; q < min(ncubes, lane * per + per); q++) loc += L.cube[q] >> 8;
        uint32_t tot;
        uint32_t ex = wave_excl_scan_u32(loc, &tot);
        for (int q = lane * per; q < min(ncubes, lane * per + per); q++) {
            L.foff[q] = (uint16_t)ex;
            ex += L.cube[q] >> 8;
        }
        nf = tot;
    }
    wave_sync();
    // vertex vi: the used edge L.list[vi] interpolated (VertexInterp)
    auto vertex = [&](uint32_t vi) {
        const int slot = L.list[vi], a = slot % 3, id = slot / 3;
        const int gg[3] = {id & 7, (id >> 3) & 7, id >> 6};
        const int id2 = id + (a == 0 ? 1 : (a == 1 ? FF_BL : FF_BL * FF_BL));
        float p1[3], p2[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int ck = geom_coord(g, k, gg[k]);
            p1[k] = m.off[k] + (float)ck * m.scale;
            p2[k] = k == a ? m.off[k] + (float)(ck + elen[k]) * m.scale : p1[k];
        }
        float V[3];
        vinterp(p1, p2, L.d[id], L.d[id2], V);
        const uint16_t ca[3] = {L.c[0][id], L.c[1][id], L.c[2][id]};
        const uint16_t cb[3] = {L.c[0][id2], L.c[1][id2], L.c[2][id2]};
        const uint32_t rgb = cinterp(ca, cb, L.d[id], L.d[id2]);
        return make_float4(V[0], V[1], V[2], __uint_as_float(rgb));
    };
    // the pool reservation is issued first and each lane's first vertex is
    // computed while it returns
    unsigned long long old = 0ull;
    if (lane == 0) old = pool_reserve(P, di % FF_ARENAS, nv, nf);
    float4 first = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if ((uint32_t)lane < nv) first = vertex((uint32_t)lane);
    uint32_t v0 = 0, f0 = 0, ok = 1;
    if (lane == 0) ok = pool_place(P, di % FF_ARENAS, nv, nf, old, v0, f0, mc) ? 1u : 0u;
    v0 = __shfl(v0, 0);
    f0 = __shfl(f0, 0);
    const bool fits = __shfl(ok, 0) != 0;
    if (fits) {
        // vertices: one used edge per lane
        if ((uint32_t)lane < nv) P.v[v0 + lane] = first;
        for (uint32_t vi = lane + 64; vi < nv; vi += 64) P.v[v0 + vi] = vertex(vi);
        // faces
        for (int q = lane; q < ncubes; q += 64) {
            const int cq = L.cube[q];
            const int nt = cq >> 8;
            if (!nt) continue;
            const uint64_t T = tri[cq & 255];
            const int jl = (int)div_small((uint32_t)q, m0), l = (int)div_small((uint32_t)q, m01);
            const int i = q - jl * nc0, j = jl - l * nc1;
            const int b0 = (l * FF_BL + j) * FF_BL + i;
            const uint32_t fb = f0 + L.foff[q];
            for (int t = 0; t < 3 * nt; t++) {
                const int slot = b0 * 3 + edge_slot_off(tri_edge(T, t));
                const uint32_t wd = L.used[slot >> 5];
                P.f[3 * fb + t] = L.wpref[slot >> 5] + __popc(wd & ((1u << (slot & 31)) - 1u));
            }
        }
    }
    if (lane == 0) {
        if (!fits) nv = nf = 0;
        cp->v0 = v0; cp->nv = nv; cp->f0 = f0; cp->nf = nf;
    }
    wave_sync();
}

// (dirty cells in slot order as k_mark / k_mark_same split them: the long
// ones first).  sampleSeq: the stream starts this launch only once the
// sampling before it has ended, so its first thread writes sampleVal there,
// the sequence word a frame's SDF update waits for (in place of a stream
// memory operation, which cost ~13 us of host time on this stream)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_mc(CellRec *cells, const CellGeom *__restrict__ geom, MeshCounters *mc,
                                            const uint4 *__restrict__ sbuf, MapRO m, Pools P, uint32_t geomCap,
                                            uint32_t *sampleSeq, uint32_t sampleVal)
{
    if (sampleSeq && blockIdx.x == 0 && threadIdx.x == 0)
        __hip_atomic_store(sampleSeq, sampleVal, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __shared__ WaveLds lds[4];
    __shared__ uint64_t tri[256];
    tri[threadIdx.x] = c_tri64[threadIdx.x];  // 256 threads
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    WaveLds &L = lds[wave];
    const uint32_t nd = mc->nDirty, nBig = mc->pad[2];
    for (uint32_t base = blockIdx.x * 4; base < nd; base += gridDim.x * 4) {
        // waves of a block run independently: each owns lds[wave]
        const uint32_t di = base + wave;
        const bool act = di < nd;
        CellGeom g;
        if (act) {
            g = geom[dirty_geom(di, nBig, geomCap, true)];
        } else {
            g.cell = 0; g.tile0 = FF_EMPTY_CELL; g.type = 0;
            g.n[0] = g.n[1] = g.n[2] = 0;
            g.o[0] = g.o[1] = g.o[2] = 0;
            g.size = g.sizeMin = g.stretch = 1;
        }
        mc_cell(L, tri, lane, cells, g, mc, sbuf, m, P, di, act);
    }
}

// one stream of a batched marching-cubes launch
struct McStream {
    CellRec *cells;
    const CellGeom *geom;
    MeshCounters *mc;
    const uint4 *sbuf;
    MapRO m;
    Pools P;
};

// exclusive prefix of the streams' dirty-cell counts (n <= 64: one wave)
__global__ void k_mc_prefix(const McStream *__restrict__ ms, int n, uint32_t *__restrict__ prefix)
{
    const int lane = threadIdx.x;
    const uint32_t cnt = lane < n ? ms[lane].mc->nDirty : 0u;
    uint32_t x = cnt;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane < n) prefix[lane] = x - cnt;
    if (lane == 63) prefix[n] = x;
}

// the dirty cells of n streams as one worklist, one wave per cell
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_mc_batch(const McStream *__restrict__ ms, int n,
                                                  const uint32_t *__restrict__ prefix)
{
    __shared__ WaveLds lds[4];
    __shared__ uint64_t tri[256];
    tri[threadIdx.x] = c_tri64[threadIdx.x];  // 256 threads
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    WaveLds &L = lds[wave];
    const uint32_t total = prefix[n];
    for (uint32_t base = blockIdx.x * 4; base < total; base += gridDim.x * 4) {
        const uint32_t item = base + wave;
        const bool act = item < total;
        int lo = 0, hi = n;  // prefix[lo] <= item < prefix[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (prefix[mid] <= item) lo = mid;
            else hi = mid;
        }
        const McStream &S = ms[lo];
        const uint32_t di = act ? item - prefix[lo] : 0u;
        CellGeom g;
        if (act) {
            g = S.geom[di];
        } else {
            g.cell = 0; g.tile0 = FF_EMPTY_CELL; g.type = 0;
            g.n[0] = g.n[1] = g.n[2] = 0;
            g.o[0] = g.o[1] = g.o[2] = 0;
            g.size = g.sizeMin = g.stretch = 1;
        }
        mc_cell(L, tri, lane, S.cells, g, S.mc, S.sbuf, S.m, S.P, di, act);
    }
}

// live-geometry compaction: copies each cell's segments to fresh pools
__global__ void k_compact(CellRec *cells, MeshCounters *mc, uint32_t cap, const float4 *__restrict__ vsrc,
                          const uint32_t *__restrict__ fsrc, Pools dst)
{
    const uint32_t nc = min(mc->nCells, cap);
    const int lane = threadIdx.x & 63;
    for (uint32_t c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; c < nc; c += (gridDim.x * blockDim.x) >> 6) {
        CellRec &r = cells[c];
        const uint32_t nv = r.nv, nf = r.nf;
        uint32_t v0 = 0, f0 = 0, ok = 1;
        if (lane == 0) ok = pool_alloc(dst, c % FF_ARENAS, nv, nf, v0, f0, mc) ? 1u : 0u;
        v0 = __shfl(v0, 0);
        f0 = __shfl(f0, 0);
        if (__shfl(ok, 0)) {
            for (uint32_t i = lane; i < nv; i += 64) dst.v[v0 + i] = vsrc[r.v0 + i];
            for (uint32_t i = lane; i < 3 * nf; i += 64) dst.f[3 * f0 + i] = fsrc[3 * r.f0 + i];
        }
        if (lane == 0) {
            if (!ok) { r.nv = 0; r.nf = 0; }
            r.v0 = v0; r.f0 = f0;
        }
    }
}

// concatenation for the host: cell k's vertices go to vo[k], its faces to
// fo[k] with indices made global
__global__ void k_gather(const uint4 *__restrict__ segs, const uint2 *__restrict__ outOff, uint32_t n,
                         const float4 *__restrict__ vsrc, const uint32_t *__restrict__ fsrc, float4 *__restrict__ vdst,
                         uint32_t *__restrict__ fdst, uint32_t localFaces)
{
    const int lane = threadIdx.x & 63;
    for (uint32_t c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; c < n; c += (gridDim.x * blockDim.x) >> 6) {
        const uint4 sg = segs[c];  // v0, nv, f0, nf
        const uint2 oo = outOff[c];
        for (uint32_t i = lane; i < sg.y; i += 64) vdst[oo.x + i] = vsrc[sg.x + i];
        const uint32_t fo = localFaces ? 0u : oo.x;  // face indices local to the cell, or into the whole mesh
        for (uint32_t i = lane; i < 3 * sg.w; i += 64) fdst[3 * (size_t)oo.y + i] = fsrc[3 * (size_t)sg.z + i] + fo;
    }
}

__global__ void k_binfo_shift(const ushort4 *src, ushort4 *dst, uint32_t n, uint32_t ax, uint32_t ay, uint32_t az)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        ushort4 b = src[i];
        b.x = (unsigned short)(b.x + ax); b.y = (unsigned short)(b.y + ay); b.z = (unsigned short)(b.z + az);
        dst[i] = b;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static bool mok(ff_ctx *c, hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    c->err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}
#define MCHK(call)                                  \
    do {                                            \
        if (!mok(c, (call), #call)) return FF_E_DEVICE; \
    } while (0)
#define MCHK_CTX(ctx, call)                            \
    do {                                               \
        if (!mok((ctx), (call), #call)) return FF_E_DEVICE; \
    } while (0)

int mesh_init(ff_ctx *c)
{
    MeshState *s = new MeshState();
    c->mesh = s;
    // triangle tables into constant memory
    uint64_t tri[256];
    for (int i = 0; i < 256; i++) {
        const char *r = FF_MC_TRI[i];
        const int len = (int)strlen(r);  // at most 15 edges (5 triangles)
        uint64_t T = (uint64_t)(len / 3) << 60;
        for (int k = 0; k < len; k++) T |= (uint64_t)(r[k] <= '9' ? r[k] - '0' : r[k] - 'a' + 10) << (4 * k);
        tri[i] = T;
    }
    MCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_tri64), tri, sizeof tri));
    s->binfoCap = c->nBranchesTotal / 16 + 1;
    MCHK(hipMalloc(&s->d_binfo, sizeof(ushort4) * s->binfoCap));
    MCHK(hipMemset(s->d_binfo, 0, sizeof(ushort4) * s->binfoCap));
    s->termCap = std::max<uint32_t>(1u << 16, std::min<uint32_t>(c->nLeavesTotal * 2u, 1u << 24));
    MCHK(hipMalloc(&s->d_term, sizeof(uint4) * s->termCap));
    MCHK(hipMalloc(&s->d_jump, sizeof(uint4) * FF_JUMP_CELLS));
    s->cellCap = std::max<uint32_t>(1u << 18, std::min<uint32_t>(c->nLeavesTotal * 16u, 1u << 25));
    for (int k = 0; k < 2; k++) MCHK(hipMalloc(&s->d_cells[k], sizeof(CellRec) * s->cellCap));
    MCHK(hipMalloc(&s->d_termAff, sizeof(uint4) * FF_INC_NODES));
    MCHK(hipMalloc(&s->d_fresh, sizeof(CellRec) * FF_INC_NODES * 27u));
    MCHK(hipMalloc(&s->d_inc, sizeof(uint32_t) * 4));
    MCHK(hipMemset(s->d_inc, 0, sizeof(uint32_t) * 4));
    MCHK(hipMalloc(&s->d_geom, sizeof(CellGeom) * s->cellCap));
    s->tileCap = std::max<uint32_t>(1u << 16, std::min<uint32_t>(c->nLeavesTotal * 8u, 1u << 19));
    MCHK(hipMalloc(&s->d_tiles, sizeof(TileRec) * s->tileCap));
    MCHK(hipMalloc(&s->d_sbuf, sizeof(uint4) * 64 * (size_t)s->tileCap));
    uint32_t hcap = 1;
    while (hcap < 2 * s->cellCap) hcap <<= 1;
    s->hmask = hcap - 1;
    for (int k = 0; k < 2; k++) {
        MCHK(hipMalloc(&s->d_hkeysBuf[k], sizeof(uint64_t) * hcap));
        MCHK(hipMemset(s->d_hkeysBuf[k], 0xff, sizeof(uint64_t) * hcap));
        MCHK(hipMalloc(&s->d_hvalBuf[k], sizeof(uint32_t) * hcap));
    }
    s->hIdx = 0;
    s->d_hkeys = s->d_hkeysBuf[0];
    s->d_hval = s->d_hvalBuf[0];
    s->vCap = std::max<uint32_t>(1u << 22, std::min<uint32_t>(c->nLeavesTotal * 64u, 1u << 28));
    s->vCap -= s->vCap % FF_ARENAS;
    s->fCap = s->vCap;
    MCHK(hipMalloc(&s->d_vpool, sizeof(float4) * s->vCap));
    MCHK(hipMalloc(&s->d_fpool, sizeof(uint32_t) * 3 * (size_t)s->fCap));
    MCHK(hipMalloc(&s->d_arena, sizeof(unsigned long long) * FF_ARENAS * FF_ARENA_STRIDE));
    MCHK(hipMemset(s->d_arena, 0, sizeof(unsigned long long) * FF_ARENAS * FF_ARENA_STRIDE));
    for (int k = 0; k < 2; k++) {
        MCHK(hipMalloc(&s->d_mcBuf[k], sizeof(MeshCounters)));
        MCHK(hipMemset(s->d_mcBuf[k], 0, sizeof(MeshCounters)));
    }
    s->d_mc = s->d_mcBuf[0];
    // the counters, then the report words
    MCHK(hipHostMalloc(&s->h_mc, 2 * sizeof(MeshCounters), hipHostMallocMapped | hipHostMallocCoherent));
    MCHK(hipHostGetDevicePointer(reinterpret_cast<void **>(&s->d_hmc), s->h_mc, 0));
    memset(s->h_mc, 0, 2 * sizeof(MeshCounters));
    s->h_rep = reinterpret_cast<uint32_t *>(s->h_mc + 1);
    s->d_rep = reinterpret_cast<uint32_t *>(s->d_hmc + 1);

    return FF_OK;
}

void mesh_free(ff_ctx *c)
{
    MeshState *s = c->mesh;
    if (!s) return;
    hipFree(s->d_binfo); hipFree(s->d_term); hipFree(s->d_jump);
    hipFree(s->d_termAff); hipFree(s->d_fresh); hipFree(s->d_inc);
    hipFree(s->d_cells[0]); hipFree(s->d_cells[1]);
    hipFree(s->d_geom); hipFree(s->d_tiles); hipFree(s->d_sbuf);
    for (int k = 0; k < 2; k++) { hipFree(s->d_hkeysBuf[k]); hipFree(s->d_hvalBuf[k]); }
    hipFree(s->d_vpool); hipFree(s->d_fpool);
    hipFree(s->d_arena);
    hipFree(s->d_mdesc); hipHostFree(s->h_mdesc); hipFree(s->d_msPrefix);
    hipFree(s->d_mcs); hipHostFree(s->h_mcs);
    hipFree(s->d_mcBuf[0]); hipFree(s->d_mcBuf[1]); hipHostFree(s->h_mc);
    delete s;
    c->mesh = nullptr;
}

void mesh_on_initial_volume(ff_ctx *c)
{
    MeshState *s = c->mesh;
    mesh_complete(c, false);  // (inside a frame: no allocation)
    const ushort4 root = make_ushort4(0, 0, 0, (unsigned short)(c->n >> 1));
    hipMemcpyAsync(s->d_binfo, &root, sizeof root, hipMemcpyHostToDevice, c->stream);
    hipStreamSynchronize(c->stream);
    s->haveVolume = true;
}

// a restored map: its branch geometry, and no previous cell table (every
// cell is meshed by the next update)
int mesh_on_load(ff_ctx *c, const std::vector<ushort4> &binfo)
{
    MeshState *s = c->mesh;
    mesh_complete(c);
    const size_t nb = std::min<size_t>(binfo.size(), s->binfoCap);
    if (nb) MCHK(hipMemcpyAsync(s->d_binfo, binfo.data(), sizeof(ushort4) * nb, hipMemcpyHostToDevice, c->stream));
    MCHK(hash_reset(s, c->stream));
    MCHK(hipStreamSynchronize(c->stream));
    s->nCells[0] = s->nCells[1] = 0;
    s->cur = 0;
    s->lastMeshStamp = 0;
    s->haveVolume = true;
    return FF_OK;
}

// grow: branch blocks moved by addBlocks, origins shifted; the new root
// chain's node origins and octant sizes are given by the caller
void mesh_on_grow_blocks(ff_ctx *c, uint32_t oldBlocks, uint32_t addBlocks, const uint32_t addv[3],
                         const std::vector<ushort4> &chain)
{
    MeshState *s = c->mesh;
    mesh_complete(c, false);  // its table is the one the grow discards (inside a frame: no allocation)
    // the old geometry through the tree's scratch copy, free once the tree
    // has moved (4 B per tree slot, 0.5 B per slot needed here)
    ushort4 *tmp = reinterpret_cast<ushort4 *>(c->d_treeTmp);
    hipMemcpyAsync(tmp, s->d_binfo, sizeof(ushort4) * oldBlocks, hipMemcpyDeviceToDevice, c->stream);
    k_binfo_shift<<<(oldBlocks + 255) / 256, 256, 0, c->stream>>>(tmp, s->d_binfo + addBlocks, oldBlocks, addv[0],
                                                                  addv[1], addv[2]);
    hipMemcpyAsync(s->d_binfo, chain.data(), sizeof(ushort4) * chain.size(), hipMemcpyHostToDevice, c->stream);
    hipStreamSynchronize(c->stream);
    // every cell moved: drop the previous table so all cells re-mesh
    s->nCells[s->cur] = 0;
    (void)hash_reset(s, c->stream);
}

void mesh_on_grow(ff_ctx *) {}
void mesh_after_frame(ff_ctx *) {}

// the branch-geometry table of a context (for a batched allocation phase)
void mesh_binfo(const ff_ctx *c, ushort4 **binfo, uint32_t *cap)
{
    *binfo = c->mesh->d_binfo;
    *cap = c->mesh->binfoCap;
}

// totals of the arena counters (vertices, faces allocated since the last
// compaction) into pad[0], pad[1]
__global__ void k_arena_sum(const unsigned long long *arena, MeshCounters *mc)
{
    const int lane = threadIdx.x;
    const unsigned long long a = lane < (int)FF_ARENAS ? arena[lane * FF_ARENA_STRIDE] : 0ull;
    uint32_t v = (uint32_t)a, f = (uint32_t)(a >> 32);
    for (int o = 32; o > 0; o >>= 1) {
        v += __shfl_xor(v, o);
        f += __shfl_xor(f, o);
    }
    if (lane == 0) { mc->pad[0] = v; mc->pad[1] = f; }
}

__global__ void k_mesh_begin(MeshCounters *mc)
{
    if (threadIdx.x == 0) {
        MeshCounters z;
        z.nTerm = z.nCells = z.nDirty = z.flags = z.nTiles = 0u;
        z.pad[0] = z.pad[1] = z.pad[2] = 0u;
        *mc = z;
    }
}

// the counters of an update that re-marks the current table of nCells cells
__global__ void k_mesh_begin_same(MeshCounters *mc, uint32_t nCells)
{
    if (threadIdx.x == 0) {
        MeshCounters z;
        z.nTerm = z.nDirty = z.flags = z.nTiles = 0u;
        z.nCells = nCells;
        z.pad[0] = z.pad[1] = z.pad[2] = 0u;
        *mc = z;
    }
}

__global__ void k_mesh_begin_b(const MeshBatchStream *__restrict__ S)
{
    if (threadIdx.x == 0) {
        MeshCounters z;
        z.nTerm = z.nCells = z.nDirty = z.flags = z.nTiles = 0u;
        z.pad[0] = z.pad[1] = z.pad[2] = 0u;
        *S[blockIdx.z].mc = z;
    }
}

// every stream's counters into one array (one copy back), after the arena
// totals of streams that account bytes
__global__ void k_mesh_gather_b(const MeshBatchStream *__restrict__ S, MeshCounters *__restrict__ out)
{
    const MeshBatchStream &a = S[blockIdx.z];
    const int lane = threadIdx.x;
    uint32_t v = 0, f = 0;
    if (a.arenaSum) {
        const unsigned long long x = lane < (int)FF_ARENAS ? a.arena[lane * FF_ARENA_STRIDE] : 0ull;
        v = (uint32_t)x;
        f = (uint32_t)(x >> 32);
        for (int o = 32; o > 0; o >>= 1) {
            v += __shfl_xor(v, o);
            f += __shfl_xor(f, o);
        }
    }
    if (lane == 0) {
        MeshCounters r = *a.mc;
        if (a.arenaSum) {
            r.pad[0] = v;
            r.pad[1] = f;
            *a.mc = r;
        }
        out[blockIdx.z] = r;
    }
}

static Pools pools(const MeshState *s)
{
    Pools P;
    P.v = s->d_vpool;
    P.f = s->d_fpool;
    P.arena = s->d_arena;
    P.vA = s->vCap / FF_ARENAS;
    P.fA = s->fCap / FF_ARENAS;
    P.vH = (uint32_t)((uint64_t)P.vA * s->compactPercent / 100u);
    P.fH = (uint32_t)((uint64_t)P.fA * s->compactPercent / 100u);
    return P;
}

// update_meshes in phases (so that a batch can share the marching-cubes
// launch): pre -- cells, dirty marks, samples; the marching cubes (mc);
// post -- lookup hash, counters' copy back; finish -- after the device work.
// the counter set the update after this one counts in
static inline MeshCounters *mc_other(const MeshState *s) { return s->d_mcBuf[1 - s->mcIdx]; }

// after a single update's first kernel (which cleared the other set): the
// next update counts in that set
static inline void mc_flip(MeshState *s)
{
    s->mcIdx ^= 1;
    s->mcNextZero = true;
}

// host part of pre: the table to build and the map view (false: nothing
// to mesh)
static bool mesh_pre_host(ff_ctx *c, bool allowReuse)
{
    MeshState *s = c->mesh;
    if (!s->haveVolume || !c->n) return false;
    s->reuse = allowReuse && !s->fullNext && s->nCells[s->cur] > 0 && s->builtNB == c->nBranchesUsed &&
               s->builtNL == c->nLeavesUsed && s->builtGrows == c->grows;
    // this update's counter set (zero when the previous single update's
    // first kernel cleared it); a batched update counts in it too and
    // clears no other set
    s->d_mc = s->d_mcBuf[s->mcIdx];
    s->mcZeroNow = s->mcNextZero;
    s->mcNextZero = false;
    s->launchStamp = s->stampFor;
    s->launchId = ++s->updateSeq;
    s->launchNB = c->nBranchesUsed;
    s->launchNL = c->nLeavesUsed;
    s->launchGrows = c->grows;
    s->nxt = s->reuse ? s->cur : 1 - s->cur;
    MapRO m;
    m.info = c->d_leafinfo; m.parent = c->d_parent; m.dist = c->d_dist; m.wgt = c->d_wgt; m.col = c->d_col;
    for (int k = 0; k < 3; k++) m.off[k] = c->offset[k];
    m.scale = c->scale;
    m.useColor = c->useColor ? 1 : 0;
    s->m = m;
    return true;
}

// after the sampling the next frame's SDF update may start (the marching
// cubes read the sample buffer; when some cells did not fit it, which the
// update's tile count shows, the SDF update stands down): the update's tile
// count and number for that frame's gate (the marching cubes' launch
// signals the sampling's end)
static void mesh_samples_taken(ff_ctx *c)
{
    MeshState *s = c->mesh;
    c->pendingTiles = &s->d_mc->nTiles;
    c->pendingTileCap = s->tileCap;
    c->pendingId = s->launchId;
}

bool mesh_sample_report(ff_ctx *c, uint32_t id, uint32_t *nTiles)
{
    MeshState *s = c->mesh;
    if (!await_word(&s->h_rep[1], id, c->stream)) return false;
    *nTiles = s->h_rep[0];
    return true;
}

bool mesh_launch_pre(ff_ctx *c, bool timed)
{
    MeshState *s = c->mesh;
    if (!mesh_pre_host(c, true)) return false;
    if (timed && c->timingAll) hipEventRecord(c->ev[6], c->stream);
    const int nxt = s->nxt;
    if (s->reuse) {
        const MapRO m = s->m;
        const uint32_t nc = std::min(s->nCells[s->cur], s->cellCap);
        if (!s->mcZeroNow) k_mesh_begin<<<1, 64, 0, c->stream>>>(s->d_mc);
        k_mark_same<<<std::max(1u, std::min(4096u, (nc + FF_MARK_THREADS - 1) / FF_MARK_THREADS)), FF_MARK_THREADS, 0,
                      c->stream>>>(
            s->d_cells[nxt], c->d_stamp, c->d_parent, s->lastMeshStamp, s->d_mc, nc, mc_other(s), s->d_geom,
            s->cellCap, m, s->d_tiles, s->tileCap, s->launchId);
        k_sample<<<FF_SAMPLE_BLOCKS, 256, 0, c->stream>>>(s->d_geom, s->d_cells[nxt], s->d_tiles, s->d_mc, s->tileCap, m,
                                              s->d_sbuf, s->d_rep, s->launchId);
        mesh_samples_taken(c);
        mc_flip(s);
        return true;
    }
    if (!s->mcZeroNow) k_mesh_begin<<<1, 64, 0, c->stream>>>(s->d_mc);
    const uint32_t nBlocks = c->nBranchesUsed / 16;
    // the tree as of this update (the next frame's allocation may link new
    // nodes while it runs)
    const uint32_t nLeaves = std::min(c->nLeavesUsed, c->nLeavesTotal), nSlots = c->nBranchesUsed;
    // incremental when the current table was built for this tree less the
    // bricks and branch blocks made since, and those are few
    const bool havePrev = s->nCells[s->cur] > 0 && !s->fullNext;
    IncView iv;
    iv.l0 = std::min(s->builtNL, nLeaves);
    iv.l1 = nLeaves;
    iv.b0 = std::min(s->builtNB, nSlots) / 16u;
    iv.b1 = nBlocks;
    const bool inc = havePrev && s->builtGrows == c->grows && s->builtNL <= c->nLeavesUsed &&
                     s->builtNB <= nSlots && (iv.l1 - iv.l0) + (iv.b1 - iv.b0) <= FF_INC_BOXES;
    // (plus FF_JUMP_CELLS / 256 blocks that fill the jump table)
    k_terminal<<<(nBlocks * 8 + FF_TERM_THREADS - 1) / FF_TERM_THREADS + FF_JUMP_CELLS / FF_TERM_THREADS,
                 FF_TERM_THREADS, 0, c->stream>>>(
        c->d_tree, s->d_binfo, nBlocks, nLeaves, nSlots, c->n, s->d_term, s->termCap, s->d_mc, s->d_jump,
        mc_other(s), inc, iv, c->d_leafinfo, s->d_termAff, s->d_inc);
    // (an incremental update meshes a few touched nodes: a small grid, whose
    // blocks retire at once, lets the marking start sooner; should the
    // touched nodes exceed FF_INC_NODES, the same grid runs the full pass)
    k_cells<<<inc ? 64 : 1024, 512, 0, c->stream>>>(c->d_tree, nLeaves, nSlots, c->n, s->d_term, s->termCap, s->d_mc,
                                         s->d_cells[nxt], s->cellCap, s->d_jump, inc, s->d_termAff, s->d_fresh,
                                         s->d_inc);
    const MapRO m = s->m;
    if (inc) {
        const uint32_t nPrev = std::min(s->nCells[s->cur], s->cellCap);
        k_mark_inc<<<std::max(1024u, std::min(4096u, (nPrev + FF_MARK_THREADS - 1) / FF_MARK_THREADS + 16u)),
                     FF_MARK_THREADS, 0, c->stream>>>(
            s->d_cells[nxt], s->d_cells[s->cur], nPrev, s->d_hkeys, s->d_hval, s->hmask, c->d_stamp, c->d_parent,
            s->lastMeshStamp, s->d_mc, s->d_geom, s->cellCap, m, s->d_tiles, s->tileCap, s->launchId, iv,
            s->d_binfo, s->d_fresh, s->d_inc);
    } else {
        k_mark<<<1024, FF_MARK_THREADS, 0, c->stream>>>(s->d_cells[nxt], s->d_cells[s->cur], s->d_hkeys, s->d_hval,
                                                        s->hmask, havePrev, c->d_stamp, c->d_parent,
                                                        s->lastMeshStamp, s->d_mc, s->d_geom, s->cellCap, m,
                                                        s->d_tiles, s->tileCap, s->launchId);
    }
    k_sample<<<FF_SAMPLE_BLOCKS, 256, 0, c->stream>>>(s->d_geom, s->d_cells[nxt], s->d_tiles, s->d_mc, s->tileCap, m, s->d_sbuf,
                                          s->d_rep, s->launchId);
    mesh_samples_taken(c);
    mc_flip(s);
    return true;
}

// the descriptor of a context in a batched update (after mesh_pre_host)
static MeshBatchStream mesh_batch_stream(ff_ctx *c)
{
    MeshState *s = c->mesh;
    MeshBatchStream a;
    a.tree = c->d_tree;
    a.binfo = s->d_binfo;
    a.nBlocks = c->nBranchesUsed / 16;
    a.nLeaves = std::min(c->nLeavesUsed, c->nLeavesTotal);
    a.nSlots = c->nBranchesUsed;
    a.n = c->n;
    a.term = s->d_term;
    a.termCap = s->termCap;
    a.mc = s->d_mc;
    a.cells = s->d_cells[s->nxt];
    a.prev = s->d_cells[s->cur];
    a.cellCap = s->cellCap;
    a.nPrev = s->nCells[s->cur];
    a.hkeys = s->d_hkeys;
    a.hval = s->d_hval;
    a.hmask = s->hmask;
    a.havePrev = s->nCells[s->cur] > 0 && !s->fullNext;
    a.stamp = c->d_stamp;
    a.parent = c->d_parent;
    a.since = s->lastMeshStamp;
    a.updateId = s->launchId;
    a.geom = s->d_geom;
    a.m = s->m;
    a.tiles = s->d_tiles;
    a.tileCap = s->tileCap;
    a.sbuf = s->d_sbuf;
    a.arena = s->d_arena;
    a.arenaSum = c->timingAll ? 1 : 0;
    return a;
}

// grid x per stream of a batched launch: the single-stream grid spread
// over the streams (4x oversubscribed), within [lo, single]
static int batch_grid(int single, int n, int lo)
{
    return std::max(lo, std::min(single, single * 4 / std::max(n, 1)));
}

McStream mesh_stream(ff_ctx *c)
{
    MeshState *s = c->mesh;
    McStream ms;
    ms.cells = s->d_cells[s->nxt];
    ms.geom = s->d_geom;
    ms.mc = s->d_mc;
    ms.sbuf = s->d_sbuf;
    ms.m = s->m;
    ms.P = pools(s);
    return ms;
}

void mesh_launch_mc(ff_ctx *c)
{
    MeshState *s = c->mesh;
    // (timed on one update in timeEvery, by events its own dispatch records)
    s->mcTimed = c->timing && c->ev[8] && (c->updates++ % c->timeEvery) == 0;
    // the launch writes the sampling's sequence word as it starts
    uint32_t *sw = nullptr;
    uint32_t sv = 0;
    if (c->seqOk) {
        sw = c->d_seq + FF_SEQ_SAMPLE;
        sv = c->seq[FF_SEQ_SAMPLE] + 1u;
    }
    if (s->mcTimed)
        hipExtLaunchKernelGGL(k_mc, dim3(2048), dim3(256), 0, c->stream, c->ev[8], c->ev[9], 0,
                              (CellRec *)s->d_cells[s->nxt], (const CellGeom *)s->d_geom, s->d_mc,
                              (const uint4 *)s->d_sbuf, s->m, pools(s), s->cellCap, sw, sv);
    else
        k_mc<<<2048, 256, 0, c->stream>>>(s->d_cells[s->nxt], s->d_geom, s->d_mc, s->d_sbuf, s->m, pools(s),
                                          s->cellCap, sw, sv);
    // a frame waits for the word only once the launch is in the stream
    if (sw && hipPeekAtLastError() == hipSuccess) {
        c->seq[FF_SEQ_SAMPLE] = sv;
        c->samplePending = true;
    }
    // the counters are final (the hash turnover does not change them)
    s->doneMapped = !c->timingAll;
    if (s->doneMapped) k_mesh_report<<<1, 64, 0, c->stream>>>(s->d_mc, s->d_hmc, &s->d_rep[2], s->launchId);
}

// the hash turnover of an update that built a new table: the previous
// table's keys leave the lookup hash, the new ones enter the other one
static void launch_turnover(ff_ctx *c, int nxt, hipStream_t st)
{
    MeshState *s = c->mesh;
    const int b = 1 - s->hIdx;
    k_hash_turnover<<<1024, 256, 0, st>>>(s->d_cells[s->cur], s->nCells[s->cur], s->d_hkeys, s->d_cells[nxt], s->d_mc,
                                          s->cellCap, s->d_hkeysBuf[b], s->d_hvalBuf[b], s->hmask,
                                          c->timingAll ? s->d_arena : nullptr, s->d_inc);
    s->hIdx = b;  // the next update looks the new table up there
    s->d_hkeys = s->d_hkeysBuf[b];
    s->d_hval = s->d_hvalBuf[b];
}

void mesh_launch_deferred_turnover(ff_ctx *c, hipStream_t st)
{
    MeshState *s = c->mesh;
    if (!s || !s->turnPending) return;
    s->turnPending = false;
    launch_turnover(c, s->turnTable, st);
}

void mesh_launch_post(ff_ctx *c, bool timed)
{
    MeshState *s = c->mesh;
    const int nxt = s->nxt;
    if (s->reuse) {
        // same table, same keys: the lookup hash stands
        if (c->timingAll) k_arena_sum<<<1, 64, 0, c->stream>>>(s->d_arena, s->d_mc);
    } else if (c->frameInFlush) {
        // the frame that enqueues this update launches it on its side stream
        // after its SDF update (it reads only the tables, which the marching
        // cubes leave alone but for the mesh ranges); the next update waits
        // for that frame's SDF sequence word, which follows it
        s->turnPending = true;
        s->turnTable = nxt;
    } else {
        launch_turnover(c, nxt, c->stream);
    }
    if (timed && c->timingAll) hipEventRecord(c->ev[7], c->stream);
    if (c->timingAll) hipMemcpyAsync(s->h_mc, s->d_mc, sizeof(MeshCounters), hipMemcpyDeviceToHost, c->stream);
}

// the allocating part of an update's host part (see maintPending)
static int mesh_maintain(ff_ctx *c)
{
    MeshState *s = c->mesh;
    if (!s->maintPending) return FF_OK;
    s->maintPending = false;
    const int nxt = s->maintTable;
    const uint32_t flags = s->maintFlags;
    if (s->maintTiles > s->tileCap) {
        // some cells' samples did not fit the buffer (the marching cubes
        // sampled them in place): grow it for the next updates
        uint32_t cap = s->tileCap;
        while (cap < s->maintTiles && cap < FF_TILE_CAP_MAX) cap <<= 1;
        if (cap > s->tileCap) {
            TileRec *nt = nullptr;
            uint4 *nb = nullptr;
            if (hipMalloc(&nt, sizeof(TileRec) * cap) == hipSuccess &&
                hipMalloc(&nb, sizeof(uint4) * 64 * (size_t)cap) == hipSuccess) {
                MCHK(hipStreamSynchronize(c->stream));  // no update in flight reads the old buffer
                hipFree(s->d_tiles);
                hipFree(s->d_sbuf);
                s->d_tiles = nt;
                s->d_sbuf = nb;
                s->tileCap = cap;
            } else {
                hipFree(nt);  // keep the old buffer; in-place meshing covers the rest
                (void)hipGetLastError();
            }
        }
    }
    if (flags & (FF_MESH_OVERFLOW | FF_MESH_HALF)) {
        // compact live geometry into fresh pools (cell k to arena k mod FF_ARENAS)
        float4 *nv = nullptr;
        uint32_t *nf = nullptr;
        MCHK(hipMalloc(&nv, sizeof(float4) * s->vCap));
        if (!mok(c, hipMalloc(&nf, sizeof(uint32_t) * 3 * (size_t)s->fCap), "hipMalloc(face pool)")) {
            hipFree(nv);
            return FF_E_DEVICE;
        }
        MCHK(hipMemsetAsync(s->d_arena, 0, sizeof(unsigned long long) * FF_ARENAS * FF_ARENA_STRIDE, c->stream));
        Pools dst = pools(s);
        dst.v = nv;
        dst.f = nf;
        // the compaction's own overflow shows in the flags word, cleared first
        MCHK(hipMemsetAsync(&s->d_mc->flags, 0, sizeof(uint32_t), c->stream));
        k_compact<<<1024, 256, 0, c->stream>>>(s->d_cells[nxt], s->d_mc, s->cellCap, s->d_vpool, s->d_fpool, dst);
        if (c->timingAll) {
            k_arena_sum<<<1, 64, 0, c->stream>>>(s->d_arena, s->d_mc);
            MCHK(hipMemcpyAsync(s->h_mc, s->d_mc, sizeof(MeshCounters), hipMemcpyDeviceToHost, c->stream));
        }
        MCHK(hipStreamSynchronize(c->stream));
        uint32_t cflags = 0;
        MCHK(hipMemcpy(&cflags, &s->d_mc->flags, sizeof cflags, hipMemcpyDeviceToHost));
        if (c->timingAll) {
            s->lastVAlloc = s->h_mc->pad[0];
            s->lastFAlloc = s->h_mc->pad[1];
        }
        hipFree(s->d_vpool);
        hipFree(s->d_fpool);
        s->d_vpool = nv;
        s->d_fpool = nf;
        s->compactions++;
        const bool lost = (cflags & FF_MESH_OVERFLOW) != 0;
        if ((lost || (flags & FF_MESH_OVERFLOW)) && s->vCap < FF_POOL_MAX) {
            // the update's meshes or the live geometry itself did not fit an
            // arena: twice the pools (empty: the next update meshes every cell,
            // which also needs room beside the live geometry)
            const uint32_t cap = (uint32_t)std::min<uint64_t>((uint64_t)s->vCap * 2u, FF_POOL_MAX);
            float4 *gv = nullptr;
            uint32_t *gf = nullptr;
            if (hipMalloc(&gv, sizeof(float4) * cap) == hipSuccess &&
                hipMalloc(&gf, sizeof(uint32_t) * 3 * (size_t)cap) == hipSuccess) {
                hipFree(s->d_vpool);
                hipFree(s->d_fpool);
                s->d_vpool = gv;
                s->d_fpool = gf;
                s->vCap = s->fCap = cap;
                MCHK(hipMemset(s->d_arena, 0, sizeof(unsigned long long) * FF_ARENAS * FF_ARENA_STRIDE));
                s->lastVAlloc = s->lastFAlloc = 0;
            } else {
                hipFree(gv);
                (void)hipGetLastError();
            }
        }
        if ((flags & FF_MESH_OVERFLOW) || lost) {
            // cells that did not fit were stored empty; mesh everything again next time
            s->nCells[nxt] = 0;
            MCHK(hash_reset(s, c->stream));
        }
    }
    return FF_OK;
}

// after the device work of pre / mc / post completed
int mesh_finish(ff_ctx *c, bool timed)
{
    MeshState *s = c->mesh;
    const int nxt = s->nxt;
    if (timed && c->timingAll) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, c->ev[6], c->ev[7]) == hipSuccess) {
            c->phaseMs[2] += ms;
            c->phaseCalls[2]++;
        }
    }
    const MeshCounters &h = *s->h_mc;
    if (h.flags & FF_CELL_OVERFLOW) {
        c->err = "mesh cell table overflow";
        return FF_E_POOL_FULL;
    }
    s->cur = nxt;
    s->nCells[nxt] = std::min(h.nCells, s->cellCap);
    s->lastUpdateId = s->launchId;
    s->lastMeshStamp = s->launchStamp;
    if (!s->reuse) {
        s->builtNB = s->launchNB;
        s->builtNL = s->launchNL;
        s->builtGrows = s->launchGrows;
    }
    if (c->timingAll) c->phaseUnits[2] += h.nDirty;
    if (s->mcTimed) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, c->ev[8], c->ev[9]) == hipSuccess) {
            c->mcMs += ms;
            c->mcCalls++;
        }
        s->mcTimed = false;
    }
    // algorithmic bytes: the 512 source voxels (14 B) of each re-meshed cell,
    // 15 B per vertex and 12 B per triangle written (vertex and face counts
    // from the arena totals, which only grow between compactions)
    if (c->timingAll) {
        const uint64_t va = h.pad[0], fa = h.pad[1];
        const uint64_t dv = va >= s->lastVAlloc ? va - s->lastVAlloc : va, df = fa >= s->lastFAlloc ? fa - s->lastFAlloc : fa;
        c->phaseBytes[2] += (double)h.nDirty * 512.0 * 14.0 + 15.0 * (double)dv + 12.0 * (double)df;
        // k_mc alone: the re-meshed cells' samples it consumes (one 14 B
        // voxel record per sample, 64 per tile) and the geometry it writes
        c->mcBytes += 14.0 * 64.0 * (double)h.nTiles + 15.0 * (double)dv + 12.0 * (double)df;
        c->mcByteCalls++;
        c->mcUnits += h.nDirty;
        s->lastVAlloc = va;
        s->lastFAlloc = fa;
    }
    const uint32_t nDirty = h.nDirty;
    const bool tilesShort = h.nTiles > s->tileCap;
    if (tilesShort || (h.flags & (FF_MESH_OVERFLOW | FF_MESH_HALF))) {
        s->maintPending = true;
        s->maintFlags = h.flags & (FF_MESH_OVERFLOW | FF_MESH_HALF);
        s->maintTiles = tilesShort ? h.nTiles : 0u;
        s->maintTable = nxt;
    }
    return (int)nDirty;
}

static int update_meshes(ff_ctx *c, uint32_t stamp);

int mesh_flush(ff_ctx *c)
{
    MeshState *s = c->mesh;
    if (!s || !s->requested) return FF_OK;
    s->requested = false;
    return update_meshes(c, s->reqStamp);
}

// the host part of an enqueued update, once its device work completed
int mesh_complete(ff_ctx *c, bool maintain)
{
    frame_join(c);  // every caller reads or changes the map next
    MeshState *s = c->mesh;
    if (!s) return FF_OK;
    const int rf = mesh_flush(c);
    if (rf != FF_OK) return rf;
    if (!s->pending) return maintain ? mesh_maintain(c) : FF_OK;
    s->pending = false;
    // (a turnover left for a frame that did not launch it goes now)
    mesh_launch_deferred_turnover(c, c->stream);
    if (s->doneMapped) {
        // the counters are in host memory once the marching cubes are done:
        // no wait for the rest of the stream (later work on it stays
        // ordered after the update's hash turnover)
        s->doneMapped = false;
        if (!await_word(&s->h_rep[2], s->launchId, c->stream)) {
            MCHK(hipStreamSynchronize(c->stream));
            c->err = "the mesh update did not report its counters";
            return FF_E_DEVICE;
        }
    } else {
        MCHK(hipStreamSynchronize(c->stream));
    }
    const int r = mesh_finish(c, s->pendingTimed);
    if (r < 0) return r;
    return maintain ? mesh_maintain(c) : FF_OK;
}

// enqueues the update of the map as of frame stamp `stamp`; the previous
// one's host part runs first (a frame whose SDF update may still run is
// joined on the device: the update's kernels wait for it on the stream)
static int update_meshes(ff_ctx *c, uint32_t stamp)
{
    frame_join(c, true);
    const int rc = mesh_complete(c);
    if (rc != FF_OK) return rc;
    c->mesh->stampFor = stamp;
    if (!mesh_launch_pre(c, true)) return FF_OK;
    mesh_launch_mc(c);
    mesh_launch_post(c, true);
    MCHK(hipGetLastError());
    c->mesh->pending = true;
    c->mesh->pendingTimed = true;
    return FF_OK;
}

// host copy of the current cell table, sorted by key
static int host_cells(ff_ctx *c, std::vector<CellRec> &out)
{
    MeshState *s = c->mesh;
    const int rc = mesh_complete(c);
    if (rc != FF_OK) return rc;
    out.resize(std::min(s->nCells[s->cur], s->cellCap));
    MCHK(hipStreamSynchronize(c->stream));
    if (!out.empty())
        MCHK(hipMemcpy(out.data(), s->d_cells[s->cur], sizeof(CellRec) * out.size(), hipMemcpyDeviceToHost));
    std::sort(out.begin(), out.end(), [](const CellRec &a, const CellRec &b) { return a.key < b.key; });
    return FF_OK;
}

// the meshes of the given cells (in that order) into host buffers: vertices
// (xyz, rgb) concatenated, faces indexing the concatenation or (local) each
// cell's own vertices
static int gather_cells(ff_ctx *c, const std::vector<CellRec> &cells, bool local, float *xyz, uint8_t *rgb,
                        uint32_t *faces)
{
    MeshState *s = c->mesh;
    // device gather of the cells' segments into dense buffers
    std::vector<uint4> segs(cells.size());
    std::vector<uint2> offs(cells.size());
    size_t vt = 0, ft = 0;
    for (size_t k = 0; k < cells.size(); k++) {
        const CellRec &r = cells[k];
        segs[k] = make_uint4(r.v0, r.nv, r.f0, r.nf);
        offs[k] = make_uint2((uint32_t)vt, (uint32_t)ft);
        vt += r.nv;
        ft += r.nf;
    }
    if (vt >= 0xffffffffull || ft >= 0xffffffffull) {
        c->err = "mesh larger than 2^32 vertices";
        return FF_E_ARG;
    }
    std::vector<float4> vp(vt);
    std::vector<uint32_t> fp(3 * ft);
    if (!cells.empty() && (vt || ft)) {
        uint4 *d_segs = nullptr;
        uint2 *d_offs = nullptr;
        float4 *d_v = nullptr;
        uint32_t *d_f = nullptr;
        hipError_t e = hipMalloc(&d_segs, sizeof(uint4) * segs.size());
        if (e == hipSuccess) e = hipMalloc(&d_offs, sizeof(uint2) * offs.size());
        if (e == hipSuccess) e = hipMalloc(&d_v, sizeof(float4) * std::max<size_t>(vt, 1));
        if (e == hipSuccess) e = hipMalloc(&d_f, sizeof(uint32_t) * 3 * std::max<size_t>(ft, 1));
        if (e == hipSuccess) e = hipMemcpyAsync(d_segs, segs.data(), sizeof(uint4) * segs.size(), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(d_offs, offs.data(), sizeof(uint2) * offs.size(), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) {
            k_gather<<<1024, 256, 0, c->stream>>>(d_segs, d_offs, (uint32_t)cells.size(), s->d_vpool, s->d_fpool, d_v, d_f,
                                                  local ? 1u : 0u);
            e = hipGetLastError();
        }
        if (e == hipSuccess && vt) e = hipMemcpyAsync(vp.data(), d_v, sizeof(float4) * vt, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && ft) e = hipMemcpyAsync(fp.data(), d_f, sizeof(uint32_t) * 3 * ft, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        hipFree(d_segs); hipFree(d_offs); hipFree(d_v); hipFree(d_f);
        MCHK(e);
    }
    for (size_t i = 0; i < vt; i++) {
        const float4 v = vp[i];
        if (xyz) { xyz[3 * i] = v.x; xyz[3 * i + 1] = v.y; xyz[3 * i + 2] = v.z; }
        if (rgb) {
            uint32_t cc;
            memcpy(&cc, &v.w, sizeof cc);
            rgb[3 * i] = (uint8_t)(cc & 0xffu);
            rgb[3 * i + 1] = (uint8_t)((cc >> 8) & 0xffu);
            rgb[3 * i + 2] = (uint8_t)((cc >> 16) & 0xffu);
        }
    }
    if (faces) memcpy(faces, fp.data(), sizeof(uint32_t) * fp.size());
    return FF_OK;
}

// the cells the last completed update re-meshed, in key order
static int update_cells(ff_ctx *c, std::vector<CellRec> &out)
{
    std::vector<CellRec> all;
    const int rc = host_cells(c, all);
    if (rc) return rc;
    out.clear();
    const uint32_t id = c->mesh->lastUpdateId;
    if (!id) return FF_OK;
    for (const CellRec &r : all)
        if (r.pad[0] == id) out.push_back(r);
    return FF_OK;
}

static void cell_desc(const CellRec &r, int32_t *d)
{
    d[0] = (int32_t)(r.key & 7u);
    d[1] = (int32_t)((r.key >> 48) & 0xffffu);
    d[2] = (int32_t)((r.key >> 32) & 0xffffu);
    d[3] = (int32_t)((r.key >> 16) & 0xffffu);
    d[4] = 1 << ((r.key >> 3) & 31u);
}

}  // namespace ff

using namespace ff;

extern "C" {

int ff_update_meshes(ff_ctx *c)
{
    if (!c) return FF_E_ARG;
    hipSetDevice(c->device);
    if (!c->incremental) return 0;
    MeshState *s = c->mesh;
    // (a request not yet enqueued goes first)
    const int rc = mesh_flush(c);
    if (rc != FF_OK) return rc;
    s->requested = true;
    s->reqStamp = c->stampCounter;
    return FF_OK;
}

int ff_sync(ff_ctx *c)
{
    if (!c) return FF_E_ARG;
    hipSetDevice(c->device);
    const int rc = mesh_flush(c);
    if (rc != FF_OK) return rc;
    frame_join(c);
    if (c->aux) MCHK(hipStreamSynchronize(c->aux));
    MCHK(hipStreamSynchronize(c->stream));
    return FF_OK;
}

int ff_mesh_full(ff_ctx *c)
{
    if (!c) return FF_E_ARG;
    hipSetDevice(c->device);
    int rc = mesh_complete(c);
    if (rc != FF_OK) return rc;
    MeshState *s = c->mesh;
    s->fullNext = true;
    rc = update_meshes(c, c->stampCounter);
    s->fullNext = false;
    if (rc != FF_OK) return rc;
    return mesh_complete(c);
}

int ff_mesh_sizes(ff_ctx *c, size_t *nv, size_t *nf)
{
    if (!c) return FF_E_ARG;
    std::vector<CellRec> cells;
    int rc = host_cells(c, cells);
    if (rc) return rc;
    size_t v = 0, f = 0;
    for (const auto &r : cells) { v += r.nv; f += r.nf; }
    if (nv) *nv = v;
    if (nf) *nf = f;
    return FF_OK;
}

int ff_copy_mesh(ff_ctx *c, float *xyz, uint8_t *rgb, uint32_t *faces)
{
    if (!c) return FF_E_ARG;
    std::vector<CellRec> cells;
    int rc = host_cells(c, cells);
    if (rc) return rc;
    return gather_cells(c, cells, false, xyz, rgb, faces);
}

int ff_cell_update_sizes(ff_ctx *c, size_t *n_cells, size_t *n_vertices, size_t *n_faces)
{
    if (!c) return FF_E_ARG;
    std::vector<CellRec> cells;
    const int rc = update_cells(c, cells);
    if (rc) return rc;
    size_t v = 0, f = 0;
    for (const CellRec &r : cells) { v += r.nv; f += r.nf; }
    if (n_cells) *n_cells = cells.size();
    if (n_vertices) *n_vertices = v;
    if (n_faces) *n_faces = f;
    return FF_OK;
}

int ff_copy_cell_update(ff_ctx *c, int32_t *desc5, uint32_t *nv, uint32_t *nf, float *xyz, uint8_t *rgb,
                        uint32_t *faces)
{
    if (!c) return FF_E_ARG;
    std::vector<CellRec> cells;
    const int rc = update_cells(c, cells);
    if (rc) return rc;
    for (size_t i = 0; i < cells.size(); i++) {
        if (desc5) cell_desc(cells[i], desc5 + 5 * i);
        if (nv) nv[i] = cells[i].nv;
        if (nf) nf[i] = cells[i].nf;
    }
    return gather_cells(c, cells, true, xyz, rgb, faces);
}

int ff_cell_count(ff_ctx *c, size_t *n)
{
    if (!c || !n) return FF_E_ARG;
    const int rc = mesh_complete(c);
    if (rc != FF_OK) return rc;
    *n = c->mesh->nCells[c->mesh->cur];
    return FF_OK;
}

int ff_copy_cells(ff_ctx *c, int32_t *desc5, uint32_t *nv, uint32_t *nf)
{
    if (!c) return FF_E_ARG;
    std::vector<CellRec> cells;
    int rc = host_cells(c, cells);
    if (rc) return rc;
    for (size_t i = 0; i < cells.size(); i++) {
        const CellRec &r = cells[i];
        if (desc5) {
            desc5[5 * i] = (int32_t)(r.key & 7u);
            desc5[5 * i + 1] = (int32_t)(r.key >> 48);
            desc5[5 * i + 2] = (int32_t)((r.key >> 32) & 0xffffu);
            desc5[5 * i + 3] = (int32_t)((r.key >> 16) & 0xffffu);
            desc5[5 * i + 4] = 1 << ((r.key >> 3) & 31u);
        }
        if (nv) nv[i] = r.nv;
        if (nf) nf[i] = r.nf;
    }
    return FF_OK;
}

int ff_set_incremental_meshing(ff_ctx *c, int on)
{
    if (!c) return FF_E_ARG;
    const int rc = mesh_complete(c);
    if (rc != FF_OK) return rc;
    c->incremental = on != 0;
    return FF_OK;
}

int ff_set_mesh_sample_tiles(ff_ctx *c, uint32_t tiles)
{
    if (!c || !c->mesh || tiles == 0 || tiles > FF_TILE_CAP_MAX) {
        if (c) c->err = "ff_set_mesh_sample_tiles: tiles must be in [1, 4194304]";
        return FF_E_ARG;
    }
    hipSetDevice(c->device);
    MeshState *s = c->mesh;
    const int rc = mesh_complete(c);
    if (rc != FF_OK) return rc;
    MCHK(hipStreamSynchronize(c->stream));
    TileRec *nt = nullptr;
    uint4 *nb = nullptr;
    MCHK(hipMalloc(&nt, sizeof(TileRec) * tiles));
    if (!mok(c, hipMalloc(&nb, sizeof(uint4) * 64 * (size_t)tiles), "hipMalloc(sample buffer)")) {
        hipFree(nt);
        return FF_E_DEVICE;
    }
    hipFree(s->d_tiles);
    hipFree(s->d_sbuf);
    s->d_tiles = nt;
    s->d_sbuf = nb;
    s->tileCap = tiles;
    return FF_OK;
}

int ff_set_mesh_pool(ff_ctx *c, uint32_t vertices, uint32_t compact_percent)
{
    if (!c || !c->mesh || vertices < 64u * FF_ARENAS || compact_percent < 1 || compact_percent > 100) {
        if (c) c->err = "ff_set_mesh_pool: at least 4096 vertices, compact_percent in [1, 100]";
        return FF_E_ARG;
    }
    hipSetDevice(c->device);
    MeshState *s = c->mesh;
    const int rc = mesh_complete(c);
    if (rc != FF_OK) return rc;
    MCHK(hipStreamSynchronize(c->stream));
    const uint32_t cap = vertices - vertices % FF_ARENAS;
    float4 *nv = nullptr;
    uint32_t *nf = nullptr;
    MCHK(hipMalloc(&nv, sizeof(float4) * cap));
    if (!mok(c, hipMalloc(&nf, sizeof(uint32_t) * 3 * (size_t)cap), "hipMalloc(face pool)")) {
        hipFree(nv);
        return FF_E_DEVICE;
    }
    hipFree(s->d_vpool);
    hipFree(s->d_fpool);
    s->d_vpool = nv;
    s->d_fpool = nf;
    s->vCap = s->fCap = cap;
    s->compactPercent = compact_percent;
    MCHK(hipMemset(s->d_arena, 0, sizeof(unsigned long long) * FF_ARENAS * FF_ARENA_STRIDE));
    // the geometry of the old pools is gone: the next update meshes every cell
    s->nCells[s->cur] = 0;
    MCHK(hash_reset(s, c->stream));
    MCHK(hipStreamSynchronize(c->stream));
    s->lastVAlloc = s->lastFAlloc = 0;
    return FF_OK;
}

int ff_mesh_pool_info(ff_ctx *c, uint32_t *vertices, uint32_t *compactions)
{
    if (!c || !c->mesh) return FF_E_ARG;
    const int rc = mesh_complete(c);
    if (rc != FF_OK) return rc;
    if (vertices) *vertices = c->mesh->vCap;
    if (compactions) *compactions = c->mesh->compactions;
    return FF_OK;
}

int ff_update_meshes_batch(ff_ctx *const *ctxs, int n)
{
    if (!ctxs || n <= 0 || n > FF_BATCH_MAX) return FF_E_ARG;
    for (int i = 0; i < n; i++) {
        if (!ctxs[i] || !ctxs[i]->mesh) return FF_E_ARG;
        for (int j = 0; j < i; j++)
            if (ctxs[j] == ctxs[i]) return FF_E_ARG;
        if (ctxs[i]->device != ctxs[0]->device) return FF_E_ARG;
    }
    ff_ctx *lead = ctxs[0];
    hipSetDevice(lead->device);
    for (int i = 0; i < n; i++) {
        const int rc = mesh_complete(ctxs[i]);
        if (rc != FF_OK) return rc;
    }
    MeshState *ls = lead->mesh;
    if (ls->msCap < n) {
        hipFree(ls->d_mdesc); hipHostFree(ls->h_mdesc); hipFree(ls->d_msPrefix);
        hipFree(ls->d_mcs); hipHostFree(ls->h_mcs);
        ls->d_mdesc = ls->h_mdesc = nullptr;
        ls->d_ms = ls->h_ms = nullptr;
        ls->d_mbs = ls->h_mbs = nullptr;
        ls->d_msPrefix = nullptr;
        ls->d_mcs = ls->h_mcs = nullptr;
        ls->msCap = 0;
        const size_t a = (sizeof(McStream) * n + 255) & ~(size_t)255;
        ls->mdescBytes = a + sizeof(MeshBatchStream) * n;
        MCHK_CTX(lead, hipMalloc(&ls->d_mdesc, ls->mdescBytes));
        MCHK_CTX(lead, hipHostMalloc(&ls->h_mdesc, ls->mdescBytes, hipHostMallocDefault));
        MCHK_CTX(lead, hipMalloc(&ls->d_msPrefix, sizeof(uint32_t) * (n + 1)));
        MCHK_CTX(lead, hipMalloc(&ls->d_mcs, sizeof(MeshCounters) * n));
        MCHK_CTX(lead, hipHostMalloc(&ls->h_mcs, sizeof(MeshCounters) * n, hipHostMallocDefault));
        ls->d_ms = reinterpret_cast<McStream *>(ls->d_mdesc);
        ls->h_ms = reinterpret_cast<McStream *>(ls->h_mdesc);
        ls->d_mbs = reinterpret_cast<MeshBatchStream *>(ls->d_mdesc + a);
        ls->h_mbs = reinterpret_cast<MeshBatchStream *>(ls->h_mdesc + a);
        ls->msCap = n;
    }
    // every context runs on the lead's stream for the update
    const hipStream_t st = lead->stream;
    std::vector<hipStream_t> own(n);
    for (int i = 0; i < n; i++) {
        own[i] = ctxs[i]->stream;
        if (own[i] != st) hipStreamSynchronize(own[i]);
        ctxs[i]->stream = st;
    }
    std::vector<int> act;
    for (int i = 0; i < n; i++) {
        // each context is meshed as of its last frame (the stamp its next
        // update compares the bricks' stamps against)
        ctxs[i]->mesh->stampFor = ctxs[i]->stampCounter;
        if (ctxs[i]->incremental && mesh_pre_host(ctxs[i], false)) act.push_back(i);
    }
    const int na = (int)act.size();
    if (na) {
        // one upload of all descriptors, one launch per kernel, one copy back
        uint32_t maxT = 1;
        for (int k = 0; k < na; k++) {
            ls->h_mbs[k] = mesh_batch_stream(ctxs[act[k]]);
            ls->h_ms[k] = mesh_stream(ctxs[act[k]]);
            maxT = std::max(maxT, (ls->h_mbs[k].nBlocks * 8u + FF_TERM_THREADS - 1u) / FF_TERM_THREADS);
        }
        hipMemcpyAsync(ls->d_mdesc, ls->h_mdesc, ls->mdescBytes, hipMemcpyHostToDevice, st);
        const MeshBatchStream *S = ls->d_mbs;
        k_mesh_begin_b<<<dim3(1, 1, na), 64, 0, st>>>(S);
        k_terminal_b<<<dim3(maxT, 1, na), FF_TERM_THREADS, 0, st>>>(S);
        k_cells_b<<<dim3(batch_grid(1024, na, 32), 1, na), 512, 0, st>>>(S);
        k_mark_b<<<dim3(batch_grid(1024, na, 32), 1, na), FF_MARK_THREADS, 0, st>>>(S);
        k_sample_b<<<dim3(batch_grid(4096, na, 64), 1, na), 256, 0, st>>>(S);
        k_mc_prefix<<<1, 64, 0, st>>>(ls->d_ms, na, ls->d_msPrefix);
        k_mc_batch<<<2048, 256, 0, st>>>(ls->d_ms, na, ls->d_msPrefix);
        k_hash_clear_cells_b<<<dim3(batch_grid(1024, na, 32), 1, na), 256, 0, st>>>(S);
        k_hash_build_b<<<dim3(batch_grid(1024, na, 32), 1, na), 256, 0, st>>>(S);
        k_mesh_gather_b<<<dim3(1, 1, na), 64, 0, st>>>(S, ls->d_mcs);
        hipMemcpyAsync(ls->h_mcs, ls->d_mcs, sizeof(MeshCounters) * na, hipMemcpyDeviceToHost, st);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    int rc = FF_OK;
    if (e != hipSuccess) {
        lead->err = std::string("ff_update_meshes_batch: ") + hipGetErrorString(e);
        rc = FF_E_DEVICE;
    } else {
        for (int k = 0; k < na; k++) {
            ff_ctx *c = ctxs[act[k]];
            memcpy(c->mesh->h_mc, &ls->h_mcs[k], sizeof(MeshCounters));
            const int r = mesh_finish(c, false);
            if (r < 0 && rc == FF_OK) rc = r;
        }
    }
    for (int i = 0; i < n; i++) ctxs[i]->stream = own[i];
    return rc;
}

}  // extern "C"
