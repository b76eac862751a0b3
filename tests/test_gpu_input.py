"""The reference's input path end to end on the GPU: a synthetic sequence
written as a TUM RGB-D directory (16-bit depth PNGs, color PNGs, an
association file with quaternion poses) and fused through
fastfusion_amd.tum equals the same frames fused from memory."""
import numpy as np
import pytest

from fastfusion_amd import FusionMipMapGPU, synth, tum

pytestmark = pytest.mark.gpu
PIL = pytest.importorskip("PIL.Image")


def _quat(R):
    """(qx, qy, qz, qw) of a rotation matrix (Shepperd's method)."""
    m = R
    tr = np.trace(m)
    if tr > 0:
        s = np.sqrt(tr + 1.0) * 2
        return np.array([(m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s, (m[1, 0] - m[0, 1]) / s, 0.25 * s])
    i = int(np.argmax(np.diag(m)))
    j, k = (i + 1) % 3, (i + 2) % 3
    s = np.sqrt(1.0 + m[i, i] - m[j, j] - m[k, k]) * 2
    q = np.zeros(4)
    q[i] = 0.25 * s
    q[j] = (m[j, i] + m[i, j]) / s
    q[k] = (m[k, i] + m[i, k]) / s
    q[3] = (m[k, j] - m[j, k]) / s
    return q


def test_tum_directory_equals_memory(tmp_path):
    sc = synth.Scene.room_scene(41, (4.0, 4.0, 3.0), 4, 3, 0.1, 0.45)
    n = 8
    R, t = sc.walk(n, 42, 0.5, 3.0, 0.02, 2.0)
    (tmp_path / "depth").mkdir()
    (tmp_path / "rgb").mkdir()
    lines, mem = [], []
    for i in range(n):
        d, bgr = sc.render(R[i].reshape(3, 3), t[i], tum.K_DEFAULT, 640, 480, seed=500 + i)
        PIL.fromarray(d).save(tmp_path / "depth" / f"{i}.png")
        PIL.fromarray(np.ascontiguousarray(bgr[:, :, ::-1]), "RGB").save(tmp_path / "rgb" / f"{i}.png")
        q = _quat(R[i].reshape(3, 3))
        lines.append(f"{i}.0 {t[i][0]:.7f} {t[i][1]:.7f} {t[i][2]:.7f} {q[0]:.8f} {q[1]:.8f} {q[2]:.8f} {q[3]:.8f} "
                     f"{i}.0 depth/{i}.png {i}.0 rgb/{i}.png")
        mem.append((d, bgr))
    (tmp_path / "associate.txt").write_text("\n".join(lines) + "\n")
    frames = tum.read_associations(str(tmp_path / "associate.txt"))
    assert len(frames) == n
    for i, (Rf, tf, dp, rp) in enumerate(frames):
        assert np.allclose(Rf, R[i].reshape(3, 3), atol=1e-6) and np.allclose(tf, t[i], atol=1e-6)
        assert np.array_equal(tum.read_depth_png(dp), mem[i][0])
        assert np.array_equal(tum.read_color_png(rp), mem[i][1])
    a = FusionMipMapGPU(0, 0, 0, 0.01, 0.03, 0, True, max_leaves=1 << 17)
    b = FusionMipMapGPU(0, 0, 0, 0.01, 0.03, 0, True, max_leaves=1 << 17)
    nq = tum.fuse(frames, a, max_cam_distance=4.0)
    for i, (Rf, tf, _, _) in enumerate(frames):
        assert b.addMap(mem[i][0], Rf, tf, tum.K_DEFAULT, mem[i][1], 1 / 5000.0, 4.0) == nq[i]
        b.updateMeshes()
    assert np.array_equal(a.tree(), b.tree())
    ids = np.arange(a.counts()["n_leaves_used"], dtype=np.uint32)
    for x, y in zip(a.bricks(ids), b.bricks(ids)):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))
    ma, mb = a.getMeshInterleavedMarchingCubes(), b.getMeshInterleavedMarchingCubes()

    def tris(m):  # the cell order of the export may differ between contexts
        return np.unique(m.vertices[m.faces.astype(np.int64)].reshape(len(m.faces), 9).view(np.uint32), axis=0)

    assert len(ma.faces) > 1000 and np.array_equal(tris(ma), tris(mb))
