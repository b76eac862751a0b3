"""The allocation-free frame entry point (ff_add_map_device_noalloc) against
ff_add_map_device: two contexts fed the same frames, one through each entry
point, must hold bit-identical maps after every call -- a 640x480 frame,
then a smaller 320x240 one (the second call reuses the scratch array with a
different layout).  Images whose buffers do not fit the array are rejected
on the host."""
import numpy as np
import pytest

from fastfusion_amd import FFError, FusionMipMapGPU, hiprt, synth
from tests.helpers import compare_state

pytestmark = pytest.mark.gpu

K320 = (262.5, 262.5, 159.5, 119.5)


def _same(a, b, w, h, tag):
    ca, cb = a.counts(), b.counts()
    for k in ("n_leaves_used", "n_branches_used", "n_queued", "side", "grows"):
        assert ca[k] == cb[k], f"{tag}: {k} {ca[k]} vs {cb[k]}"
    assert np.array_equal(a.processed(w, h), b.processed(w, h)), f"{tag}: processed pixels"
    assert np.array_equal(a.tree(), b.tree()), f"{tag}: tree"
    for x, y in zip(a.leaves(), b.leaves()):
        assert np.array_equal(x, y), f"{tag}: leaves"
    assert np.array_equal(a.queue(), b.queue()), f"{tag}: queue"
    ids = np.arange(ca["n_leaves_used"], dtype=np.uint32)
    da, wa, cla = a.bricks(ids)
    db, wb, clb = b.bricks(ids)
    assert np.array_equal(da.view(np.uint32), db.view(np.uint32)), f"{tag}: distances"
    assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32)), f"{tag}: weights"
    assert np.array_equal(cla, clb), f"{tag}: colors"


def test_noalloc_equals_add_map_device():
    sc, R, t, cfg = synth.config_c2(n_frames=20)
    plain = FusionMipMapGPU(0, 0, 0, cfg["scale"], cfg["threshold"], 0, True, max_leaves=1 << 17)
    noalloc = FusionMipMapGPU(0, 0, 0, cfg["scale"], cfg["threshold"], 0, True, max_leaves=1 << 17)
    calls = [(0, 640, 480, synth.K_DEFAULT), (6, 320, 240, K320)]
    bufs = []
    for i, w, h, K in calls:
        d, bgr = sc.render(R[i].reshape(3, 3), t[i], K, w, h, seed=1000 + i)
        db, cb = hiprt.DeviceBuffer.from_array(d), hiprt.DeviceBuffer.from_array(bgr)
        bufs += [db, cb]
        n_a = plain.addMapDevice(db.addr(), cb.addr(), w, h, R[i], t[i], K, 1 / 5000.0, cfg["maxcam"])
        n_b = noalloc.addMapDeviceNoAlloc(db.addr(), cb.addr(), w, h, R[i], t[i], K, 1 / 5000.0, cfg["maxcam"])
        assert n_a == n_b and n_a > 0, (n_a, n_b)
        _same(plain, noalloc, w, h, f"frame {i} at {w}x{h}")
    # both maps keep meshing alike
    plain.updateMeshes()
    noalloc.updateMeshes()
    ma, mb = plain.getMeshInterleavedMarchingCubes(), noalloc.getMeshInterleavedMarchingCubes()
    assert len(ma.faces) > 0
    assert np.array_equal(ma.vertices.view(np.uint32), mb.vertices.view(np.uint32))
    assert np.array_equal(ma.faces, mb.faces) and np.array_equal(ma.colors, mb.colors)


def test_noalloc_learned_mixed_with_add_map_device():
    """Learned weighting on, each frame's mesh update requested: one context
    takes every frame through ff_add_map_device, the other alternates the
    allocation-free entry point (its confidences in the scratch array) with
    ff_add_map_device (its own two alternating confidence buffers); maps and
    meshes must stay bit-identical."""
    from fastfusion_amd import learned
    sc, R, t, cfg = synth.config_c2(n_frames=20)
    w = learned.load_weights()
    a = FusionMipMapGPU(0, 0, 0, cfg["scale"], cfg["threshold"], 0, True, max_leaves=1 << 17)
    b = FusionMipMapGPU(0, 0, 0, cfg["scale"], cfg["threshold"], 0, True, max_leaves=1 << 17)
    a.setLearnedWeighting(w)
    b.setLearnedWeighting(w)
    bufs = []
    for k, i in enumerate(range(0, 12, 2)):
        d, bgr = sc.render(R[i].reshape(3, 3), t[i], K320, 320, 240, seed=1000 + i)
        db, cb = hiprt.DeviceBuffer.from_array(d), hiprt.DeviceBuffer.from_array(bgr)
        bufs += [db, cb]
        args = (db.addr(), cb.addr(), 320, 240, R[i], t[i], K320, 1 / 5000.0, cfg["maxcam"])
        n_a = a.addMapDevice(*args)
        n_b = b.addMapDeviceNoAlloc(*args) if k % 2 == 0 else b.addMapDevice(*args)
        assert n_a == n_b and n_a > 0, (k, n_a, n_b)
        a.updateMeshes()
        b.updateMeshes()
        if k % 2 == 1:
            _same(a, b, 320, 240, f"frame {i}")
    ma, mb = a.getMeshInterleavedMarchingCubes(), b.getMeshInterleavedMarchingCubes()
    assert len(ma.faces) > 0
    assert np.array_equal(ma.vertices.view(np.uint32), mb.vertices.view(np.uint32))
    assert np.array_equal(ma.faces, mb.faces) and np.array_equal(ma.colors, mb.colors)


def test_noalloc_rejects_oversize_image():
    fz = FusionMipMapGPU(0, 0, 0, 0.01, 0.03, 0, True, max_leaves=1 << 12)
    d = hiprt.DeviceBuffer(2048 * 2048 * 2)
    c = hiprt.DeviceBuffer(2048 * 2048 * 3)
    with pytest.raises(FFError, match="too large"):
        fz.addMapDeviceNoAlloc(d.addr(), c.addr(), 2048, 2048, np.eye(3), np.zeros(3), (1000.0, 1000.0, 1023.5, 1023.5))
    assert fz.counts()["frames"] == 0
