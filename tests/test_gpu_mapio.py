"""Map dump and restore on the GPU (copyCPU2HDD / copyHDD2CPU,
geometryfusion_mipmap_cpu.cpp:1128-1184): the file the library writes equals,
byte for byte, the oracle's map in the reference's format (oracle/mapio.py);
a map restored into a fresh context has the same tree, bricks and meshes, and
fusing on from it gives the same map as fusing on from the original."""
import numpy as np
import pytest

from fastfusion_amd import FusionMipMapGPU, synth
from oracle import mapio
from tests.helpers import make_pair

pytestmark = pytest.mark.gpu

K = (262.5, 262.5, 159.5, 119.5)


def _frames(n):
    sc = synth.Scene.room_scene(31, (4.0, 4.0, 3.0), 4, 3, 0.1, 0.45)
    R, t = sc.walk(n, 32, 0.5, 3.0, 0.02, 2.0)
    return [(R[i], t[i]) + sc.render(R[i].reshape(3, 3), t[i], K, 320, 240, seed=300 + i) for i in range(n)]


def _same_map(a, b, tag):
    ca, cb = a.counts(), b.counts()
    for k in ("n_leaves_used", "n_branches_used", "side"):
        assert ca[k] == cb[k], f"{tag}: {k}"
    assert np.array_equal(ca["offset"], cb["offset"])
    assert np.array_equal(a.tree(), b.tree()), f"{tag}: tree"
    for x, y in zip(a.leaves(), b.leaves()):
        assert np.array_equal(x, y), f"{tag}: leaves"
    ids = np.arange(ca["n_leaves_used"], dtype=np.uint32)
    for x, y in zip(a.bricks(ids), b.bricks(ids)):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8)), f"{tag}: bricks"


def _triangles(m):
    v = m.vertices[m.faces.astype(np.int64)].reshape(len(m.faces), 9)
    return np.unique(v.view(np.uint32), axis=0)


def test_dump_equals_reference_format(tmp_path):
    fr = _frames(6)
    gpu, ora = make_pair(0.01, 0.03)
    for R, t, d, bgr in fr:
        gpu.addMap(d, R, t, K, bgr, 1 / 5000.0, 4.0)
        ora.add_map(d, R, t, K, bgr, 1 / 5000.0, 4.0)
    path = gpu.copyCPU2HDD(str(tmp_path))
    off = gpu.counts()["offset"]
    assert path.endswith(mapio.description(off, 0.01, gpu.counts()["side"]) + ".dat")
    data = open(path, "rb").read()
    assert data == mapio.from_oracle(ora)


def test_restore_then_fuse_on(tmp_path):
    fr = _frames(9)
    a = FusionMipMapGPU(0, 0, 0, 0.01, 0.03, 0, True, max_leaves=1 << 17)
    for R, t, d, bgr in fr[:6]:
        a.addMap(d, R, t, K, bgr, 1 / 5000.0, 4.0)
    a.updateMeshes()
    path = a.copyCPU2HDD(str(tmp_path))
    b = FusionMipMapGPU(0, 0, 0, 0.01, 0.03, 0, True, max_leaves=1 << 17)
    b.copyHDD2CPU(path)
    _same_map(a, b, "restored")
    b.updateMeshes()
    ta, tb = _triangles(a.getMeshInterleavedMarchingCubes()), _triangles(b.getMeshInterleavedMarchingCubes())
    assert len(ta) > 1000 and np.array_equal(ta, tb)
    for i, (R, t, d, bgr) in enumerate(fr[6:]):
        assert a.addMap(d, R, t, K, bgr, 1 / 5000.0, 4.0) == b.addMap(d, R, t, K, bgr, 1 / 5000.0, 4.0)
        a.updateMeshes()
        b.updateMeshes()
        _same_map(a, b, f"frame {6 + i}")
    ta, tb = _triangles(a.getMeshInterleavedMarchingCubes()), _triangles(b.getMeshInterleavedMarchingCubes())
    assert np.array_equal(ta, tb)


def test_restore_without_parent_block(tmp_path):
    """A file in the bare reference format (no parents block): the parents
    become the deepest bricks above in the tree; everything else restores."""
    fr = _frames(5)
    a = FusionMipMapGPU(0, 0, 0, 0.01, 0.03, 0, False, max_leaves=1 << 17)
    for R, t, d, bgr in fr:
        a.addMap(d, R, t, K, None, 1 / 5000.0, 4.0)
    path = a.copyCPU2HDD(str(tmp_path))
    raw = open(path, "rb").read()
    dec = mapio.decode(raw, False)
    bare = tmp_path / "bare.dat"
    bare.write_bytes(raw[:dec["ref_bytes"]])
    b = FusionMipMapGPU(0, 0, 0, 0.01, 0.03, 0, False, max_leaves=1 << 17)
    b.loadMap(str(bare), a.counts()["offset"], a.counts()["side"])
    assert np.array_equal(a.tree(), b.tree())
    pa, sa, _ = a.leaves()
    pb, sb, parb = b.leaves()
    assert np.array_equal(pa, pb) and np.array_equal(sa, sb)
    # every reconstructed parent is a brick strictly above its child
    tree = b.tree()
    for i in np.flatnonzero(parb < len(pa))[:200]:
        p = parb[i]
        ext = int(sb[p]) * 8
        assert sb[p] > sb[i] and np.all(pb[p] <= pb[i]) and np.all(pb[i] < pb[p].astype(np.int64) + ext)
    b.updateMeshes()
    assert len(b.getMeshInterleavedMarchingCubes().faces) > 1000
    del tree
