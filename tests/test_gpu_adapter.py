"""The header-only C++ adapter (include/ff_fusion_mipmap_gpu.hpp, the class a
caller of the reference's FusionMipMapCPU swaps in) run end to end: a small
C++ program fuses frames written by the test through addMap / updateMeshes
and writes its counts and final mesh; the same frames fused through the
Python mirror must give the same counts and the same mesh, bit for bit."""
import os
import subprocess

import numpy as np
import pytest

import fastfusion_amd
from fastfusion_amd import FusionMipMapGPU, synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
K = (262.5, 262.5, 159.5, 119.5)
W, H = 320, 240

PROGRAM = r'''
#include "ff_fusion_mipmap_gpu.hpp"
#include <cstdio>
#include <fstream>
#include <vector>
int main(int argc, char **argv)
{
    const std::string dir = argv[1];
    const int n = std::atoi(argv[2]);
    std::ifstream poses(dir + "/poses.txt");
    ff::FusionMipMapGPU fusion(0, 0, 0, 0.01f, 0.03f, 0, true);
    fusion.setThreadMeshing(true);
    fusion.setDepthChecks(2);
    std::vector<uint16_t> depth(320 * 240);
    std::vector<uint8_t> bgr(320 * 240 * 3);
    for (int i = 0; i < n; i++) {
        ff::CameraPose p{};
        for (double &v : p.R) poses >> v;
        for (double &v : p.t) poses >> v;
        poses >> p.fx >> p.fy >> p.cx >> p.cy;
        std::ifstream(dir + "/depth" + std::to_string(i) + ".raw", std::ios::binary)
            .read(reinterpret_cast<char *>(depth.data()), depth.size() * 2);
        std::ifstream(dir + "/bgr" + std::to_string(i) + ".raw", std::ios::binary)
            .read(reinterpret_cast<char *>(bgr.data()), bgr.size());
        const int q = fusion.addMap(ff::DepthView{depth.data(), 320, 240, 640}, p,
                                    ff::ColorView{bgr.data(), 320, 240, 960}, 1.0f / 5000.0f, 4.0f);
        fusion.updateMeshes();
        std::printf("%d\n", q);
    }
    const ff::Mesh m = fusion.getMeshInterleavedMarchingCubes();
    std::ofstream out(dir + "/mesh.raw", std::ios::binary);
    const uint64_t nv = m.vertices.size() / 3, nf = m.faces.size() / 3;
    out.write(reinterpret_cast<const char *>(&nv), 8);
    out.write(reinterpret_cast<const char *>(&nf), 8);
    out.write(reinterpret_cast<const char *>(m.vertices.data()), m.vertices.size() * 4);
    out.write(reinterpret_cast<const char *>(m.colors.data()), m.colors.size());
    out.write(reinterpret_cast<const char *>(m.faces.data()), m.faces.size() * 4);
    return fusion.depthChecks() == 2 ? 0 : 3;
}
'''


def test_cpp_adapter_equals_python_mirror(tmp_path):
    sc, R, t, cfg = synth.config_c1(n_frames=50)
    n = 6
    frames = [sc.render(R[i].reshape(3, 3), t[i], K, W, H, noise=cfg["noise"], dropout=cfg["dropout"],
                        seed=1000 + i) for i in range(n)]
    with open(tmp_path / "poses.txt", "w") as f:
        for i in range(n):
            f.write(" ".join(repr(float(v)) for v in list(R[i]) + list(t[i]) + list(K)) + "\n")
    for i, (d, bgr) in enumerate(frames):
        np.ascontiguousarray(d, np.uint16).tofile(tmp_path / f"depth{i}.raw")
        np.ascontiguousarray(bgr, np.uint8).tofile(tmp_path / f"bgr{i}.raw")
    src = tmp_path / "adapter.cpp"
    src.write_text(PROGRAM)
    lib = os.path.dirname(fastfusion_amd.lib_path())
    exe = tmp_path / "adapter"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-L", lib,
                           "-lffgpu", "-Wl,-rpath," + lib, "-o", str(exe)])
    print("running the C++ adapter program", flush=True)
    run = subprocess.run([str(exe), str(tmp_path), str(n)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr[-2000:]
    queued_cpp = [int(x) for x in run.stdout.split()]

    gpu = FusionMipMapGPU(0, 0, 0, 0.01, 0.03, 0, True)
    queued_py = []
    for i, (d, bgr) in enumerate(frames):
        queued_py.append(gpu.addMap(d, R[i], t[i], K, bgr, 1 / 5000.0, 4.0))
        gpu.updateMeshes()
    assert queued_cpp == queued_py and min(queued_py) > 0
    m = gpu.getMeshInterleavedMarchingCubes()

    raw = (tmp_path / "mesh.raw").read_bytes()
    nv, nf = np.frombuffer(raw[:16], np.uint64)
    o = 16
    v = np.frombuffer(raw[o:o + 12 * int(nv)], np.float32).reshape(-1, 3)
    o += 12 * int(nv)
    c = np.frombuffer(raw[o:o + 3 * int(nv)], np.uint8).reshape(-1, 3)
    o += 3 * int(nv)
    fa = np.frombuffer(raw[o:o + 12 * int(nf)], np.uint32).reshape(-1, 3)
    assert nf > 100
    assert np.array_equal(v.view(np.uint32), m.vertices.view(np.uint32))
    assert np.array_equal(c, m.colors)
    assert np.array_equal(fa, m.faces)
