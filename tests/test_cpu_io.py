"""Output formats without a GPU: the reference's raw map format (restated in
oracle/mapio.py, geometryfusion_mipmap_cpu.cpp:1128-1184) and the mesh
writers of the Python and C++ mirrors (MeshInterleaved::writePLY /
writeOBJ, mesh_interleaved.cpp:19-200), which must write the same bytes."""
import os
import subprocess

import numpy as np

from fastfusion_amd import MeshInterleaved, synth
from oracle import mapio
from oracle.pyoracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
K = (262.5, 262.5, 159.5, 119.5)


def test_map_format_layout():
    rng = np.random.default_rng(3)
    nb, nl = 48, 5
    tree = rng.integers(0, 2 ** 32, nb, dtype=np.uint64).astype(np.uint32)
    pos = rng.integers(0, 4096, (nl, 3)).astype(np.uint16)
    sc = rng.integers(1, 9, nl).astype(np.uint16)
    d = rng.normal(size=(nl, 512)).astype(np.float32)
    w = rng.random((nl, 512)).astype(np.float32)
    c = rng.integers(0, 65280, (nl, 512, 3)).astype(np.uint16)
    par = rng.integers(0, nl, nl).astype(np.uint32)
    raw = mapio.encode(tree, pos, sc, d, w, c)
    # the reference's sizes: counts, tree, sidetype3 + sidetype, two float
    # bricks, colortype3 per voxel
    assert len(raw) == 8 + 4 * nb + nl * (6 + 2 + 2048 + 2048 + 3072)
    assert np.frombuffer(raw, "<u4", 2).tolist() == [nb, nl]
    assert raw[8:8 + 4 * nb] == tree.tobytes()
    o = 8 + 4 * nb + 8 * nl
    assert raw[o:o + 4] == d[0, :1].tobytes()
    ext = mapio.encode(tree, pos, sc, d, w, c, par)
    assert ext.startswith(raw) and ext[len(raw):len(raw) + 8] == b"FFPARENT"
    back = mapio.decode(ext, True)
    assert back["ref_bytes"] == len(raw) and back["end"] == len(ext)
    for k, v in (("tree", tree), ("pos", pos), ("scale", sc), ("dist", d), ("wgt", w), ("color", c),
                 ("parents", par)):
        assert np.array_equal(back[k], v), k


def test_map_description_matches_reference_format():
    # "%.10f_%.10f_%.10f_%.10f_%.7i" of float offset, float scale, int n
    assert mapio.description((1.5, -2.25, 0.0), 0.005, 1024) == \
        "1.5000000000_-2.2500000000_0.0000000000_0.0049999999_0001024"


def test_oracle_map_round_trip():
    sc = synth.Scene.room_scene(9, (4.0, 4.0, 3.0), 3, 3, 0.1, 0.4)
    R, t = sc.walk(3, 10, 0.5, 3.0, 0.02, 2.0)
    o = Oracle(0.01, 0.03, True, 1 << 16, 1 << 20)
    for i in range(3):
        d, bgr = sc.render(R[i].reshape(3, 3), t[i], K, 320, 240, seed=70 + i)
        o.add_map(d, R[i], t[i], K, bgr, 1 / 5000.0, 4.0)
    raw = mapio.from_oracle(o)
    back = mapio.decode(raw, True)
    pos, scl, par = o.leaves()
    n = len(pos)
    dd, ww, cc = o.bricks(np.arange(n))
    assert n > 100
    assert np.array_equal(back["tree"], o.tree())
    assert np.array_equal(back["pos"], pos) and np.array_equal(back["scale"], scl)
    assert np.array_equal(back["parents"], par)
    assert np.array_equal(back["dist"].view(np.uint32), dd.view(np.uint32))
    assert np.array_equal(back["wgt"].view(np.uint32), ww.view(np.uint32))
    assert np.array_equal(back["color"], cc.transpose(0, 2, 1))


def _mesh():
    rng = np.random.default_rng(5)
    v = rng.normal(0, 2, (40, 3)).astype(np.float32)
    v[0] = [1e-5, 123456.7, -0.0]
    v[1] = [0.1, 1.0 / 3.0, -7.25]
    c = rng.integers(0, 256, (40, 3)).astype(np.uint8)
    f = rng.integers(0, 40, (60, 3)).astype(np.uint32)
    return MeshInterleaved(v, c, f)


def test_ply_header_is_the_reference_header(tmp_path):
    m = _mesh()
    m.writePLY(str(tmp_path / "m"))
    data = (tmp_path / "m.ply").read_bytes()
    hdr = data[:data.index(b"end_header") + len(b"end_header") + 1]
    assert hdr == (b"ply\nformat binary_little_endian 1.0\nelement vertex 40\nproperty float32 x\n"
                   b"property float32 y\nproperty float32 z\nproperty uchar red\nproperty uchar green\n"
                   b"property uchar blue\nelement face 60\nproperty list uint8 int32 vertex_indices\nend_header\n")
    body = data[len(hdr):]
    assert len(body) == 40 * 15 + 60 * 13
    vv = np.frombuffer(body[:600], dtype=[("p", "<f4", 3), ("c", "u1", 3)])
    assert np.array_equal(vv["p"], m.vertices) and np.array_equal(vv["c"], m.colors)


def test_python_and_cpp_writers_write_the_same_bytes(tmp_path):
    m = _mesh()
    raw = tmp_path / "mesh.bin"
    with open(raw, "wb") as f:
        f.write(np.array([len(m.vertices), len(m.faces)], "<u4").tobytes())
        f.write(m.vertices.tobytes())
        f.write(m.colors.tobytes())
        f.write(m.faces.tobytes())
    src = tmp_path / "writers.cpp"
    src.write_text(
        '#include "ff_fusion_mipmap_gpu.hpp"\n'
        "#include <cstdio>\n"
        "int main(int, char **argv) {\n"
        "  FILE *f = fopen(argv[1], \"rb\"); uint32_t n[2];\n"
        "  if (fread(n, 4, 2, f) != 2) return 1;\n"
        "  ff::Mesh m; m.vertices.resize(3 * n[0]); m.colors.resize(3 * n[0]); m.faces.resize(3 * n[1]);\n"
        "  if (fread(m.vertices.data(), 4, 3 * n[0], f) != 3 * n[0]) return 1;\n"
        "  if (fread(m.colors.data(), 1, 3 * n[0], f) != 3 * n[0]) return 1;\n"
        "  if (fread(m.faces.data(), 4, 3 * n[1], f) != 3 * n[1]) return 1;\n"
        "  fclose(f);\n"
        "  std::string out = argv[2];\n"
        "  return (m.writePLY(out + \"_bin\", true) && m.writePLY(out + \"_asc\", false) && m.writeOBJ(out)) ? 0 : 2;\n"
        "}\n")
    exe = tmp_path / "writers"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    subprocess.check_call([str(exe), str(raw), str(tmp_path / "cpp")])
    m.writePLY(str(tmp_path / "py_bin"), binary=True)
    m.writePLY(str(tmp_path / "py_asc"), binary=False)
    m.writeOBJ(str(tmp_path / "py"))
    for a, b in (("cpp_bin.ply", "py_bin.ply"), ("cpp_asc.ply", "py_asc.ply"), ("cpp.obj", "py.obj")):
        assert (tmp_path / a).read_bytes() == (tmp_path / b).read_bytes(), a
    obj = (tmp_path / "py.obj").read_text().split("\n")
    assert obj[:2] == ["", "#Written from Volume"]
    assert obj[2].startswith("v 1e-05 123457 -0 ")
    assert obj[-1].startswith("f ") and min(int(x) for x in obj[-1].split()[1:]) >= 1
