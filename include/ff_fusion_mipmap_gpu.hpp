// ff_fusion_mipmap_gpu.hpp -- header-only C++ adapter with the member names
// of the reference's FusionMipMapCPU (src/fusion/geometryfusion_mipmap_cpu.hpp:
// 111-163) over the C ABI of libffgpu.so (ff_abi.h).
//
// A caller of the reference swaps
//     FusionMipMapCPU *fusion = new FusionMipMapCPU(0,0,0, scale, threshold, 0, true);
//     fusion->addMap(depth, caminfo, rgb, 1.0f/5000.0f, maxcam);
//     fusion->updateMeshes();
//     MeshInterleaved mesh = fusion->getMeshInterleavedMarchingCubes();
// for
//     ff::FusionMipMapGPU *fusion = new ff::FusionMipMapGPU(0,0,0, scale, threshold, 0, true);
//     fusion->addMap(ff::DepthView{...}, pose, ff::ColorView{...}, 1.0f/5000.0f, maxcam);
//     fusion->updateMeshes();
//     ff::Mesh mesh = fusion->getMeshInterleavedMarchingCubes();
// where the views carry the cv::Mat data pointer, size and step, and the pose
// carries CameraInfo's rotation, translation (camera -> world) and intrinsics.
#pragma once
#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <algorithm>
#include <string>
#include <vector>

#include "ff_abi.h"

namespace ff {

struct DepthView {           // CV_16UC1
    const uint16_t *data;
    int cols, rows;
    size_t step;             // bytes per row
};

struct ColorView {           // CV_8UC3, BGR
    const uint8_t *data;
    int cols, rows;
    size_t step;
};

struct CameraPose {          // CameraInfo: R, t map camera to world
    double R[9];             // row-major
    double t[3];
    double fx, fy, cx, cy;
};

struct Mesh {                // MeshInterleaved (src/fusion/mesh_interleaved.hpp:35-69)
    std::vector<float> vertices;   // 3 per vertex
    std::vector<uint8_t> colors;   // R, G, B per vertex
    std::vector<uint32_t> faces;   // 3 per triangle

    // MeshInterleaved::writePLY (src/fusion/mesh_interleaved.cpp:139-200):
    // filename + ".ply", the reference's header lines
    bool writePLY(const std::string &filename, bool binary = true) const
    {
        std::ofstream file((filename + ".ply").c_str(), std::ios::out | std::ios::binary);
        if (!file.is_open()) return false;
        const size_t nv = vertices.size() / 3, nf = faces.size() / 3;
        file << "ply" << (binary ? "\nformat binary_little_endian 1.0" : "\nformat ascii 1.0");
        file << "\nelement vertex " << nv << "\nproperty float32 x\nproperty float32 y\nproperty float32 z";
        if (!colors.empty()) file << "\nproperty uchar red\nproperty uchar green\nproperty uchar blue";
        if (nf) file << "\nelement face " << nf << "\nproperty list uint8 int32 vertex_indices";
        file << "\nend_header";
        if (binary) file << "\n";
        for (size_t i = 0; i < nv; i++) {
            if (binary) file.write(reinterpret_cast<const char *>(&vertices[3 * i]), 3 * sizeof(float));
            else file << "\n" << vertices[3 * i] << " " << vertices[3 * i + 1] << " " << vertices[3 * i + 2];
            if (!colors.empty()) {
                if (binary) file.write(reinterpret_cast<const char *>(&colors[3 * i]), 3);
                else file << " " << (int)colors[3 * i] << " " << (int)colors[3 * i + 1] << " " << (int)colors[3 * i + 2];
            }
        }
        const unsigned char three = 3;
        for (size_t i = 0; i < nf; i++) {
            if (binary) {
                file.write(reinterpret_cast<const char *>(&three), 1);
                file.write(reinterpret_cast<const char *>(&faces[3 * i]), 3 * sizeof(uint32_t));
            } else {
                file << "\n" << 3 << " " << faces[3 * i] << " " << faces[3 * i + 1] << " " << faces[3 * i + 2];
            }
        }
        return bool(file);
    }

    // MeshInterleaved::writeOBJ (src/fusion/mesh_interleaved.cpp:19-137) for a
    // mesh without normals or textures: filename + ".obj"
    bool writeOBJ(const std::string &filename) const
    {
        std::ofstream file((filename + ".obj").c_str(), std::ios::out | std::ios::binary);
        if (!file.is_open()) return false;
        file << "\n#Written from Volume";
        const size_t nv = vertices.size() / 3, nf = faces.size() / 3;
        for (size_t i = 0; i < nv; i++) {
            file << "\nv " << vertices[3 * i] << " " << vertices[3 * i + 1] << " " << vertices[3 * i + 2];
            if (!colors.empty())
                file << " " << (float)colors[3 * i] / 255.0f << " " << (float)colors[3 * i + 1] / 255.0f << " "
                     << (float)colors[3 * i + 2] / 255.0f;
        }
        for (size_t i = 0; i < nf; i++)
            file << "\nf " << faces[3 * i] + 1 << " " << faces[3 * i + 1] + 1 << " " << faces[3 * i + 2] + 1;
        return bool(file);
    }
};

class FusionMipMapGPU {
public:
    // FusionMipMapCPU(offsetX, offsetY, offsetZ, scale, distanceThreshold, n, color):
    // as in the reference, the offset and n are set by the first frame.
    FusionMipMapGPU(float /*offsetX*/, float /*offsetY*/, float /*offsetZ*/, float scale, float distanceThreshold,
                    uint16_t /*n*/ = 0, bool color = true, uint32_t maxLeaves = 0, int device = 0)
    {
        ff_params p{};
        p.scale = scale;
        p.distance_threshold = distanceThreshold;
        p.use_color = color ? 1 : 0;
        p.incremental_meshing = 1;
        p.max_leaves = maxLeaves;
        p.device = device;
        ctx_ = ff_create(&p);
        if (!ctx_) throw std::runtime_error("ff_create failed");
    }
    ~FusionMipMapGPU() { ff_destroy(ctx_); }
    FusionMipMapGPU(const FusionMipMapGPU &) = delete;
    FusionMipMapGPU &operator=(const FusionMipMapGPU &) = delete;

    // int addMap(const cv::Mat &depth, CameraInfo caminfo, const cv::Mat &rgb,
    //            float scaling, float maxcamdistance)
    int addMap(const DepthView &depth, const CameraPose &cam, const ColorView &rgb, float scaling,
               float maxcamdistance = 10.0f)
    {
        const double K[4] = {cam.fx, cam.fy, cam.cx, cam.cy};
        return check(ff_add_map(ctx_, depth.data, depth.step, rgb.data, rgb.step, depth.cols, depth.rows, cam.R,
                                cam.t, K, scaling, maxcamdistance));
    }

    // bool updateMeshes(): records the request (the next call enqueues it on
    // the device and runs beside it); always true
    bool updateMeshes()
    {
        check(ff_update_meshes(ctx_));
        return true;
    }

    // MeshInterleaved getMeshInterleavedMarchingCubes()
    Mesh getMeshInterleavedMarchingCubes()
    {
        size_t nv = 0, nf = 0;
        check(ff_mesh_sizes(ctx_, &nv, &nf));
        Mesh m;
        m.vertices.resize(3 * nv);
        m.colors.resize(3 * nv);
        m.faces.resize(3 * nf);
        check(ff_copy_mesh(ctx_, m.vertices.data(), m.colors.data(), m.faces.data()));
        return m;
    }

    void setIncrementalMeshing(bool on) { check(ff_set_incremental_meshing(ctx_, on ? 1 : 0)); }

    // std::string getDescription(), bool copyCPU2HDD(), bool copyHDD2CPU()
    // (geometryfusion_mipmap_cpu.cpp:1128-1184): resource/<description>.dat
    std::string getDescription()
    {
        char buf[300];
        check(ff_map_description(ctx_, buf, sizeof buf));
        return buf;
    }
    bool copyCPU2HDD() { return ff_save_map(ctx_, ("resource/" + getDescription() + ".dat").c_str()) == FF_OK; }
    // the reference restores into an object whose volume already matches
    // the file: here the volume is given
    bool copyHDD2CPU(const std::string &path, const float offset[3], uint32_t n)
    {
        return ff_load_map(ctx_, path.c_str(), offset, n) == FF_OK;
    }
    void setThreadMeshing(bool) {}  // meshing is a device pass; there is no meshing thread
    // the reference stores the count clamped at 0 and never reads it (its
    // backward depth checks are commented out, :1629-1703)
    void setDepthChecks(int depthchecks) { depthChecks_ = depthchecks < 0 ? 0 : depthchecks; }
    int depthChecks() const { return depthChecks_; }

    ff_ctx *handle() { return ctx_; }

    // the bricks the last addMap updated (_leafNumberSurface, ascending), as
    // the reference records them per frame for loop closure (_leafQueues,
    // geometryfusion_mipmap_cpu.cpp:2165-2167)
    std::vector<uint32_t> lastQueue()
    {
        ff_counts c{};
        check(ff_get_counts(ctx_, &c));
        std::vector<uint32_t> q(c.n_queued);
        if (c.n_queued) q.resize((size_t)check(ff_copy_queue(ctx_, q.data(), q.size())));
        return q;
    }

private:
    int check(int rc)
    {
        if (rc < 0) throw std::runtime_error(std::string("fastfusion_amd: ") + ff_last_error(ctx_));
        return rc;
    }
    ff_ctx *ctx_ = nullptr;
    int depthChecks_ = 0;
};

// The loop-closure partner map (FusionMipMapCPU(FusionMipMapCPU*,
// nLeavesTotalLoop), geometryfusion_mipmap_cpu.cpp:645-717) of a
// FusionMipMapGPU, which must outlive it.
class LoopClosureGPU {
public:
    explicit LoopClosureGPU(FusionMipMapGPU &fusion, uint32_t nLeavesTotalLoop = 100000)
    {
        lp_ = ff_loop_create(fusion.handle(), nLeavesTotalLoop);
        if (!lp_) throw std::runtime_error("ff_loop_create failed");
    }
    ~LoopClosureGPU() { ff_loop_destroy(lp_); }
    LoopClosureGPU(const LoopClosureGPU &) = delete;
    LoopClosureGPU &operator=(const LoopClosureGPU &) = delete;

    // updateArraysForLoopClosure's reset (:3927-3931)
    void reset() { check(ff_loop_reset(lp_)); }

    // updateWrapper_deleteloop (loopclosure.cpp:232-311) for one frame:
    // depth in metres (NaN: no measurement), R, G, B planes (NULL without
    // color), the frame's recorded queue.  Returns false when the slots ran
    // out (the first *applied entries were applied), like the reference.
    bool deleteFrame(const float *depth, const uint8_t *red, const uint8_t *green, const uint8_t *blue, int width,
                     int height, const CameraPose &cam, const std::vector<uint32_t> &queue, size_t *applied = nullptr)
    {
        const double K[4] = {cam.fx, cam.fy, cam.cx, cam.cy};
        const int rc = ff_loop_delete_frame(lp_, depth, red, green, blue, width, height, cam.R, cam.t, K,
                                            queue.data(), queue.size(), applied);
        if (rc == FF_E_POOL_FULL) return false;
        check(rc);
        return true;
    }

    // deleteOldLoopPoses' loop (:3940-4022) over frames in device memory,
    // one submission
    bool deleteFrames(const std::vector<ff_loop_frame> &frames, std::vector<size_t> *applied = nullptr)
    {
        std::vector<size_t> a(frames.size());
        const int rc = ff_loop_delete_frames(lp_, frames.data(), (int)frames.size(), a.data());
        if (applied) *applied = a;
        if (rc == FF_E_POOL_FULL) return false;
        check(rc);
        return true;
    }

    uint32_t leavesMapped()
    {
        uint32_t n = 0;
        check(ff_loop_mapped(lp_, &n, nullptr));
        return n;
    }

    // startLoopClosure's pose map (geometryfusion_mipmap_cpu.cpp:4045-4054):
    // poses ordered by minus the squared translation change (double sum,
    // float key), equal keys in pose order, at most maxPoses of them
    static std::vector<size_t> poseOrder(const std::vector<CameraPose> &loop, const std::vector<CameraPose> &fused,
                                         size_t maxPoses = 50)
    {
        std::vector<std::pair<float, size_t>> key;
        for (size_t i = 0; i < loop.size() && i < fused.size(); i++) {
            const double dx = loop[i].t[0] - fused[i].t[0], dy = loop[i].t[1] - fused[i].t[1],
                         dz = loop[i].t[2] - fused[i].t[2];
            key.emplace_back((float)(-((dx * dx + dy * dy) + dz * dz)), i);
        }
        std::stable_sort(key.begin(), key.end(),
                         [](const std::pair<float, size_t> &a, const std::pair<float, size_t> &b) {
                             return a.first < b.first;
                         });
        std::vector<size_t> order;
        for (size_t i = 0; i < key.size() && i < maxPoses; i++) order.push_back(key[i].second);
        return order;
    }

private:
    int check(int rc)
    {
        if (rc < 0) throw std::runtime_error(std::string("fastfusion_amd: ") + ff_loop_last_error(lp_));
        return rc;
    }
    ff_loop *lp_ = nullptr;
};

}  // namespace ff
