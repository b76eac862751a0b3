"""Batched streams: maps advanced together by ff_add_map_batch (one SDF
launch over all queued bricks) and ff_update_meshes_batch (one
marching-cubes launch) must equal the same maps fused one at a time through
addMap / updateMeshes, bit for bit, for every launch configuration the
autotuner can pick (threads per brick, bricks per workgroup, graph replay)."""
import numpy as np
import pytest

from fastfusion_amd import Batch, FusionMipMapGPU, hiprt, synth

pytestmark = pytest.mark.gpu

K320 = (262.5, 262.5, 159.5, 119.5)

# (room seed, width, height, K, voxel, truncation, color)
STREAMS = [(21, 320, 240, K320, 0.010, 0.030, True),
           (22, 640, 480, synth.K_DEFAULT, 0.005, 0.015, True),
           (23, 320, 240, K320, 0.008, 0.024, False)]


def _make(cfg):
    seed, w, h, K, vox, tau, color = cfg
    return FusionMipMapGPU(0, 0, 0, vox, tau, 0, color, max_leaves=1 << 17)


def _frames(n_ticks):
    out = []
    for seed, w, h, K, vox, tau, color in STREAMS:
        sc = synth.Scene.room_scene(seed, (4.0, 4.0, 3.0), 6, 4, 0.1, 0.5)
        R, t = sc.walk(n_ticks, seed + 100, 0.5, 3.0, 0.02, 2.0)
        fr = [sc.render(R[i].reshape(3, 3), t[i], K, w, h, seed=seed * 1000 + i) for i in range(n_ticks)]
        out.append((R, t, fr))
    return out


def _same_map(a, b, tag):
    ca, cb = a.counts(), b.counts()
    for k in ("n_leaves_used", "n_branches_used", "n_queued", "side", "grows"):
        assert ca[k] == cb[k], f"{tag}: {k} {ca[k]} vs {cb[k]}"
    assert np.array_equal(a.tree(), b.tree()), f"{tag}: tree"
    la, lb = a.leaves(), b.leaves()
    for x, y in zip(la, lb):
        assert np.array_equal(x, y), f"{tag}: leaves"
    ids = np.arange(ca["n_leaves_used"], dtype=np.uint32)
    da, wa, cla = a.bricks(ids)
    db, wb, clb = b.bricks(ids)
    assert np.array_equal(da.view(np.uint32), db.view(np.uint32)), f"{tag}: distances"
    assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32)), f"{tag}: weights"
    assert np.array_equal(cla, clb), f"{tag}: colors"


def _same_mesh(a, b, tag):
    ma, mb = a.getMeshInterleavedMarchingCubes(), b.getMeshInterleavedMarchingCubes()
    assert np.array_equal(ma.vertices.view(np.uint32), mb.vertices.view(np.uint32)), f"{tag}: vertices"
    assert np.array_equal(ma.colors, mb.colors), f"{tag}: colors"
    assert np.array_equal(ma.faces, mb.faces), f"{tag}: faces"
    return len(ma.faces)


@pytest.mark.parametrize("config", [(512, 1, False), (256, 4, False), (64, 2, False), (512, 2, True), None])
def test_batch_equals_single_streams(config):
    n_ticks = 5
    data = _frames(n_ticks)
    single = [_make(c) for c in STREAMS]
    batched = [_make(c) for c in STREAMS]
    batch = Batch(batched)
    if config is None:
        batch.autotune(1)      # cycles through the candidates during the ticks
    else:
        batch.set_config(*config)
    bufs = []
    for tick in range(n_ticks):
        frames = []
        for s, (seed, w, h, K, vox, tau, color) in enumerate(STREAMS):
            R, t, fr = data[s]
            d, bgr = fr[tick]
            single[s].addMap(d, R[tick], t[tick], K, bgr if color else None, 1 / 5000.0, 4.0)
            db, cb = hiprt.DeviceBuffer.from_array(d), hiprt.DeviceBuffer.from_array(bgr)
            bufs += [db, cb]
            frames.append((db.addr(), cb.addr() if color else None, w, h, R[tick], t[tick], K, 1 / 5000.0, 4.0))
        batch.addMaps(frames)
        for s in range(len(STREAMS)):
            single[s].updateMeshes()
        batch.updateMeshes()
        for s in range(len(STREAMS)):
            _same_map(single[s], batched[s], f"tick {tick} stream {s}")
            # the batched update re-meshes the cells the single one does (a
            # stale stamp would re-mesh every cell: equal meshes, more work)
            ua, ub = single[s].getMeshCellsUpdate(), batched[s].getMeshCellsUpdate()
            assert np.array_equal(ua.desc, ub.desc), \
                f"tick {tick} stream {s}: {len(ub.desc)} cells re-meshed batched, {len(ua.desc)} single"
    nf = [_same_mesh(single[s], batched[s], f"stream {s}") for s in range(len(STREAMS))]
    assert min(nf) > 100
