"""The benchmark configurations against the oracle on the GPU.

* The 640x480, 5 mm random walk, all 1000 frames in order: the map state
  (tree, bricks, parents, queue, de-dup mask, TSDF, weights, colors) bit for
  bit every 25 frames and on the last, and the incremental meshes (updated
  after every frame) against the oracle's from-scratch meshes every 100
  frames and on the last.
* The 24 batched streams (320x240 to 1280x960, 4-10 mm), advanced by one
  batched tick per frame: every stream against its own oracle after each
  tick, and its meshes after the last.
"""
import numpy as np
import pytest

from fastfusion_amd import Batch, FusionMipMapGPU, hiprt, synth
from oracle.pyoracle import Oracle
from tests.helpers import compare_state, make_pair
from tests.test_gpu_mesh import compare_mesh

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(900)
def test_c2_walk_all_frames():
    sc, R, t, cfg = synth.config_c2(n_frames=1000)
    gpu, ora = make_pair(cfg["scale"], cfg["threshold"], max_leaves=1 << 18, max_branches=1 << 23)
    w, h, K = cfg["w"], cfg["h"], synth.K_DEFAULT
    for i in range(1000):
        d, bgr = sc.render(R[i].reshape(3, 3), t[i], K, w, h, noise=cfg["noise"], dropout=cfg["dropout"],
                           seed=1000 + i)
        ng = gpu.addMap(d, R[i], t[i], K, bgr, 1 / 5000.0, cfg["maxcam"])
        no = ora.add_map(d, R[i], t[i], K, bgr, 1 / 5000.0, cfg["maxcam"])
        assert ng == no, f"frame {i}: gpu queued {ng}, oracle {no}"
        gpu.updateMeshes()
        if i % 25 == 0 or i == 999:
            compare_state(gpu, ora, w, h, tag=f"frame {i}")
        if i % 100 == 0 or i == 999:
            ncell, nface = compare_mesh(gpu, ora, cfg["scale"], f"frame {i}")
            print(f"frame {i}: {gpu.counts()['n_leaves_used']} bricks, {ncell} cells, {nface} triangles",
                  flush=True)
    assert gpu.counts()["n_leaves_used"] > 3000


def test_c3_24_streams_against_oracles():
    cfgs = synth.stream_configs(24)
    assert {(c["w"], c["h"]) for c in cfgs} == {(320, 240), (640, 480), (1280, 960)}
    n_ticks = 3
    fus = [FusionMipMapGPU(0, 0, 0, c["scale"], c["threshold"], 0, True, max_leaves=1 << 16) for c in cfgs]
    oras = [Oracle(c["scale"], c["threshold"], True, max_leaves=1 << 16, max_branches=1 << 21) for c in cfgs]
    walks = [c["scene"].walk(n_ticks, c["seed"] + 7, 0.5, 3.0, 0.02, 2.0) for c in cfgs]
    batch = Batch(fus)
    batch.autotune(1)  # the candidates cycle through the ticks
    for tick in range(n_ticks):
        frames, imgs, bufs = [], [], []
        for k, c in enumerate(cfgs):
            R, t = walks[k]
            d, bgr = c["scene"].render(R[tick].reshape(3, 3), t[tick], c["K"], c["w"], c["h"],
                                       seed=c["seed"] * 10007 + tick)
            db, cb = hiprt.DeviceBuffer.from_array(d), hiprt.DeviceBuffer.from_array(bgr)
            bufs += [db, cb]
            imgs.append((d, bgr))
            frames.append((db.addr(), cb.addr(), c["w"], c["h"], R[tick], t[tick], c["K"], 1 / 5000.0,
                           c["maxcam"]))
        batch.addMaps(frames)
        batch.updateMeshes()
        for k, c in enumerate(cfgs):
            R, t = walks[k]
            d, bgr = imgs[k]
            no = oras[k].add_map(d, R[tick], t[tick], c["K"], bgr, 1 / 5000.0, c["maxcam"])
            assert fus[k].counts()["n_queued"] == no, f"tick {tick} stream {k}"
            compare_state(fus[k], oras[k], c["w"], c["h"], tag=f"tick {tick} stream {k} ({c['w']}x{c['h']})")
    faces = [compare_mesh(fus[k], oras[k], c["scale"], f"stream {k}")[1] for k, c in enumerate(cfgs)]
    assert min(faces) > 0
