// ff_mapio.hip -- the map's dump and restore in the reference's raw format.
//
// FusionMipMapCPU::copyCPU2HDD / copyHDD2CPU (geometryfusion_mipmap_cpu.cpp:
// 1135-1184) write and read, in this order and without padding:
//   u32 nBranchesUsed, u32 nLeavesUsed,
//   u32 tree[nBranchesUsed],
//   u16 leafPos[nLeavesUsed][3], u16 leafScale[nLeavesUsed],
//   f32 distance[nLeavesUsed][512], f32 weights[nLeavesUsed][512],
//   (with color) u16 color[nLeavesUsed][512][3]   (colortype3: R, G, B)
// into resource/<offset.x>_<offset.y>_<offset.z>_<scale>_<n>.dat
// (getDescription, :1128-1133).  Here the bricks' colors are three planes of
// 512 per brick in HBM, converted to the reference's per-voxel triples on
// the way out and back.  The reference does not store the bricks' parents
// (_leafParent); a trailing block "FFPARENT" + u32 count + u32 parents[count]
// carries them, which a reader of the reference format never reaches.
// Without it, a brick's parent is taken as the deepest brick above it in the
// tree.
#include "ff_device.hpp"
#include "ff_internal.hpp"

#include <cstdio>
#include <cstring>
#include <vector>

namespace ff {

static const char kParentTag[8] = {'F', 'F', 'P', 'A', 'R', 'E', 'N', 'T'};
static const size_t kChunk = 1u << 14;  // bricks per transfer

static bool io_ok(ff_ctx *c, hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    c->err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}

// the deepest brick on the path from the root to the node of brick `leaf`
// (excluded), for maps restored without their parents
static void parents_from_tree(const std::vector<uint32_t> &tree, const std::vector<uint16_t> &pos,
                              const std::vector<uint16_t> &scale, uint32_t n, uint32_t nLeavesTotal,
                              std::vector<uint32_t> &parent)
{
    const size_t nl = scale.size();
    parent.assign(nl, FF_BRANCHINIT);
    for (size_t i = 0; i < nl; i++) {
        const uint32_t x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        const uint32_t size = (uint32_t)scale[i] * FF_BL;  // the brick's node size
        uint32_t branch = 0, half = n >> 1, ox = 0, oy = 0, oz = 0, last = FF_BRANCHINIT;
        while (half >= size && branch + 16 <= tree.size()) {
            uint32_t co = 0;
            if (x >= ox + half) { co |= 1u; ox += half; }
            if (y >= oy + half) { co |= 2u; oy += half; }
            if (z >= oz + half) { co |= 4u; oz += half; }
            const uint32_t l = tree[branch + co];
            if (half == size) break;  // the brick's own slot
            if (l < nLeavesTotal && l != i) last = l;
            const uint32_t child = tree[branch + co + 8];
            if (child == FF_BRANCHINIT) break;
            branch = child;
            half >>= 1;
        }
        parent[i] = last;
    }
}

// origin and octant size of every branch block, from the root down
static void branch_geometry(const std::vector<uint32_t> &tree, uint32_t n, std::vector<ushort4> &binfo)
{
    const size_t nb = tree.size() / 16;
    binfo.assign(nb + 1, make_ushort4(0, 0, 0, 0));
    if (!nb) return;
    struct Item { uint32_t branch, x, y, z, half; };
    std::vector<Item> stack{{0u, 0u, 0u, 0u, n >> 1}};
    while (!stack.empty()) {
        const Item it = stack.back();
        stack.pop_back();
        binfo[it.branch / 16] = make_ushort4((unsigned short)it.x, (unsigned short)it.y, (unsigned short)it.z,
                                             (unsigned short)it.half);
        for (uint32_t co = 0; co < 8; co++) {
            const uint32_t child = tree[it.branch + co + 8];
            if (child == FF_BRANCHINIT || child + 16 > tree.size()) continue;
            stack.push_back({child, it.x + it.half * (co & 1u), it.y + it.half * ((co >> 1) & 1u),
                             it.z + it.half * ((co >> 2) & 1u), it.half >> 1});
        }
    }
}

}  // namespace ff

using namespace ff;

extern "C" {

int ff_map_description(const ff_ctx *c, char *buf, size_t n)
{
    if (!c || !buf || !n) return FF_E_ARG;
    const int r = snprintf(buf, n, "%.10f_%.10f_%.10f_%.10f_%.7i", (double)c->offset[0], (double)c->offset[1],
                           (double)c->offset[2], (double)c->scale, (int)c->n);
    return (r < 0 || (size_t)r >= n) ? FF_E_ARG : r;
}

int ff_save_map(ff_ctx *c, const char *path)
{
    if (!c || !path) return FF_E_ARG;
    hipSetDevice(c->device);
    frame_join(c);
    if (!io_ok(c, hipStreamSynchronize(c->stream), "ff_save_map")) return FF_E_DEVICE;
    const uint32_t nB = std::min(c->nBranchesUsed, c->nBranchesTotal);
    const uint32_t nL = std::min(c->nLeavesUsed, c->nLeavesTotal);
    FILE *f = fopen(path, "wb");
    if (!f) {
        c->err = std::string("ff_save_map: cannot open ") + path;
        return FF_E_ARG;
    }
    bool ok = fwrite(&nB, 4, 1, f) == 1 && fwrite(&nL, 4, 1, f) == 1;
    std::vector<uint32_t> tree(nB), parent(nL);
    std::vector<ushort4> info(nL);
    ok = ok && io_ok(c, hipMemcpy(tree.data(), c->d_tree, 4 * (size_t)nB, hipMemcpyDeviceToHost), "tree") &&
         io_ok(c, hipMemcpy(info.data(), c->d_leafinfo, sizeof(ushort4) * nL, hipMemcpyDeviceToHost), "leaves") &&
         io_ok(c, hipMemcpy(parent.data(), c->d_parent, 4 * (size_t)nL, hipMemcpyDeviceToHost), "parents");
    if (ok && nB) ok = fwrite(tree.data(), 4, nB, f) == nB;
    if (ok && nL) {
        std::vector<uint16_t> pos(3 * (size_t)nL), sc(nL);
        for (uint32_t i = 0; i < nL; i++) {
            pos[3 * i] = info[i].x; pos[3 * i + 1] = info[i].y; pos[3 * i + 2] = info[i].z;
            sc[i] = info[i].w;
        }
        ok = fwrite(pos.data(), 2, pos.size(), f) == pos.size() && fwrite(sc.data(), 2, sc.size(), f) == sc.size();
        std::vector<float> buf(kChunk * FF_BRICK);
        for (int which = 0; which < 2 && ok; which++) {
            const float *src = which == 0 ? c->d_dist : c->d_wgt;
            for (size_t b0 = 0; b0 < nL && ok; b0 += kChunk) {
                const size_t nb = std::min(kChunk, (size_t)nL - b0);
                ok = io_ok(c, hipMemcpy(buf.data(), src + b0 * FF_BRICK, sizeof(float) * nb * FF_BRICK,
                                        hipMemcpyDeviceToHost), "voxels") &&
                     fwrite(buf.data(), sizeof(float), nb * FF_BRICK, f) == nb * FF_BRICK;
            }
        }
        if (ok && c->useColor) {
            std::vector<uint16_t> planes(kChunk * FF_BRICK * 3), aos(kChunk * FF_BRICK * 3);
            for (size_t b0 = 0; b0 < nL && ok; b0 += kChunk) {
                const size_t nb = std::min(kChunk, (size_t)nL - b0);
                ok = io_ok(c, hipMemcpy(planes.data(), c->d_col + b0 * FF_BRICK * 3, 2 * nb * FF_BRICK * 3,
                                        hipMemcpyDeviceToHost), "colors");
                for (size_t b = 0; b < nb && ok; b++)
                    for (int v = 0; v < FF_BRICK; v++)
                        for (int ch = 0; ch < 3; ch++)
                            aos[(b * FF_BRICK + v) * 3 + ch] = planes[(b * 3 + ch) * FF_BRICK + v];
                ok = ok && fwrite(aos.data(), 2, nb * FF_BRICK * 3, f) == nb * FF_BRICK * 3;
            }
        }
    }
    if (ok) ok = fwrite(kParentTag, 1, 8, f) == 8 && fwrite(&nL, 4, 1, f) == 1 &&
                 (!nL || fwrite(parent.data(), 4, nL, f) == nL);
    if (fclose(f) != 0) ok = false;
    if (!ok) {
        if (c->err.empty()) c->err = std::string("ff_save_map: write failed: ") + path;
        return FF_E_DEVICE;
    }
    return FF_OK;
}

int ff_load_map(ff_ctx *c, const char *path, const float offset[3], uint32_t n)
{
    if (!c || !path || !offset || n < 2 * FF_BL || n > 32768 || (n & (n - 1))) return FF_E_ARG;
    hipSetDevice(c->device);
    FILE *f = fopen(path, "rb");
    if (!f) {
        c->err = std::string("ff_load_map: cannot open ") + path;
        return FF_E_ARG;
    }
    auto fail = [&](const char *why) {
        fclose(f);
        c->err = std::string("ff_load_map: ") + why;
        return FF_E_ARG;
    };
    uint32_t nB = 0, nL = 0;
    if (fread(&nB, 4, 1, f) != 1 || fread(&nL, 4, 1, f) != 1) return fail("truncated header");
    if (nB > c->nBranchesTotal || nL > c->nLeavesTotal || nB % 16) return fail("map larger than the context's pools");
    std::vector<uint32_t> tree(nB);
    std::vector<uint16_t> pos(3 * (size_t)nL), sc(nL);
    if ((nB && fread(tree.data(), 4, nB, f) != nB) || (nL && fread(pos.data(), 2, pos.size(), f) != pos.size()) ||
        (nL && fread(sc.data(), 2, nL, f) != nL))
        return fail("truncated tree or bricks");
    frame_join(c);
    if (!io_ok(c, hipStreamSynchronize(c->stream), "ff_load_map")) {
        fclose(f);
        return FF_E_DEVICE;
    }
    drop_frame_graphs(c);
    // voxels straight to HBM, chunk by chunk
    {
        std::vector<float> buf(kChunk * FF_BRICK);
        for (int which = 0; which < 2; which++) {
            float *dst = which == 0 ? c->d_dist : c->d_wgt;
            for (size_t b0 = 0; b0 < nL; b0 += kChunk) {
                const size_t nb = std::min(kChunk, (size_t)nL - b0);
                if (fread(buf.data(), sizeof(float), nb * FF_BRICK, f) != nb * FF_BRICK) return fail("truncated voxels");
                if (!io_ok(c, hipMemcpy(dst + b0 * FF_BRICK, buf.data(), sizeof(float) * nb * FF_BRICK,
                                        hipMemcpyHostToDevice), "voxels")) {
                    fclose(f);
                    return FF_E_DEVICE;
                }
            }
        }
        if (c->useColor) {
            std::vector<uint16_t> aos(kChunk * FF_BRICK * 3), planes(kChunk * FF_BRICK * 3);
            for (size_t b0 = 0; b0 < nL; b0 += kChunk) {
                const size_t nb = std::min(kChunk, (size_t)nL - b0);
                if (fread(aos.data(), 2, nb * FF_BRICK * 3, f) != nb * FF_BRICK * 3) return fail("truncated colors");
                for (size_t b = 0; b < nb; b++)
                    for (int v = 0; v < FF_BRICK; v++)
                        for (int ch = 0; ch < 3; ch++)
                            planes[(b * 3 + ch) * FF_BRICK + v] = aos[(b * FF_BRICK + v) * 3 + ch];
                if (!io_ok(c, hipMemcpy(c->d_col + b0 * FF_BRICK * 3, planes.data(), 2 * nb * FF_BRICK * 3,
                                        hipMemcpyHostToDevice), "colors")) {
                    fclose(f);
                    return FF_E_DEVICE;
                }
            }
        }
    }
    std::vector<uint32_t> parent;
    char tag[8];
    uint32_t np = 0;
    if (fread(tag, 1, 8, f) == 8 && !memcmp(tag, kParentTag, 8) && fread(&np, 4, 1, f) == 1 && np == nL) {
        parent.resize(nL);
        if (nL && fread(parent.data(), 4, nL, f) != nL) return fail("truncated parents");
    } else {
        parents_from_tree(tree, pos, sc, n, c->nLeavesTotal, parent);
    }
    fclose(f);
    std::vector<ushort4> info(nL);
    for (uint32_t i = 0; i < nL; i++) info[i] = make_ushort4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], sc[i]);
    std::vector<ushort4> binfo;
    branch_geometry(tree, n, binfo);
    const bool ok =
        io_ok(c, hipMemsetAsync(c->d_stamp, 0, sizeof(uint32_t) * c->nLeavesTotal, c->stream), "stamps") &&
        io_ok(c, hipMemcpyAsync(c->d_tree, tree.data(), 4 * (size_t)nB, hipMemcpyHostToDevice, c->stream), "tree") &&
        io_ok(c, hipMemcpyAsync(c->d_leafinfo, info.data(), sizeof(ushort4) * nL, hipMemcpyHostToDevice, c->stream),
              "leaves") &&
        io_ok(c, hipMemcpyAsync(c->d_parent, parent.data(), 4 * (size_t)nL, hipMemcpyHostToDevice, c->stream),
              "parents") &&
        io_ok(c, hipStreamSynchronize(c->stream), "ff_load_map");
    if (!ok) return FF_E_DEVICE;
    // the unused part of the tree stays empty
    if (nB < c->nBranchesTotal) fill_tree_tail(c, nB);
    c->n = (uint16_t)n;
    for (int k = 0; k < 3; k++) c->offset[k] = offset[k];
    c->nBranchesUsed = nB;
    c->nLeavesUsed = nL;
    c->nQueued = 0;
    sync_map_counters(c);
    return mesh_on_load(c, binfo);
}

}  // extern "C"
