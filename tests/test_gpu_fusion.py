"""GPU parity of the integration path (allocation + SDF update) against the
oracle, bit for bit: allocated tree, brick positions/scales/parents, queued
set, de-dup mask, TSDF distances, weights and colors."""
import numpy as np
import pytest

from fastfusion_amd import synth
from tests.helpers import make_pair, run_pair

pytestmark = pytest.mark.gpu


def test_first_frame_c1_640():
    sc, R, t, cfg = synth.config_c1(n_frames=200)
    gpu, ora = make_pair(cfg["scale"], cfg["threshold"])
    # the reference's two settings without effect on the fused map: the
    # depth-check count is clamped at 0 and stored, the meshing thread flag
    # is accepted (geometryfusion_mipmap_cpu.cpp:1110-1120)
    gpu.setDepthChecks(-2)
    gpu.setThreadMeshing(True)
    assert gpu.depthChecks == 0
    run_pair(sc, R, t, [0], cfg, gpu, ora)


def test_c1_sequence_640():
    sc, R, t, cfg = synth.config_c1(n_frames=200)
    gpu, ora = make_pair(cfg["scale"], cfg["threshold"])
    run_pair(sc, R, t, list(range(0, 24)), cfg, gpu, ora, check_every=4)


def test_c2_walk_5mm():
    sc, R, t, cfg = synth.config_c2(n_frames=60)
    gpu, ora = make_pair(cfg["scale"], cfg["threshold"], max_leaves=1 << 18)
    run_pair(sc, R, t, list(range(0, 30)), cfg, gpu, ora, check_every=5)


def test_small_image_no_color_320():
    sc, R, t, cfg = synth.config_c1(n_frames=50)
    cfg = dict(cfg, w=320, h=240)
    K = (262.5, 262.5, 159.5, 119.5)
    gpu, ora = make_pair(cfg["scale"], cfg["threshold"], color=False)
    from tests import helpers
    old = helpers.K
    helpers.K = K
    try:
        run_pair(sc, R, t, list(range(0, 8)), cfg, gpu, ora, check_every=2)
    finally:
        helpers.K = old


def test_grow_parity():
    """A first view close to a wall, then views across the room: the tree
    grows upward (grow(), geometryfusion_mipmap_cpu.cpp:3071-3319), and the
    GPU's tree, bricks and TSDF stay bit-identical to the oracle's."""
    from tests.test_cpu_oracle import grow_sequence
    sc, R, t = grow_sequence()
    gpu, ora = make_pair(0.02, 0.06)
    cfg = dict(w=320, h=240, maxcam=6.0)
    from tests import helpers
    old = helpers.K
    helpers.K = (262.5, 262.5, 159.5, 119.5)
    try:
        run_pair(sc, R, t, list(range(len(R))), cfg, gpu, ora)
    finally:
        helpers.K = old
    assert gpu.counts()["grows"] >= 1


def test_graph_capture_parity():
    """Frames replayed from captured HIP graphs (per-frame values read from
    device memory) give the same map as the oracle, through a grow."""
    from tests.test_cpu_oracle import grow_sequence
    sc, R, t = grow_sequence()
    gpu, ora = make_pair(0.02, 0.06)
    gpu.setGraphCapture(True)
    cfg = dict(w=320, h=240, maxcam=6.0)
    from tests import helpers
    old = helpers.K
    helpers.K = (262.5, 262.5, 159.5, 119.5)
    try:
        run_pair(sc, R, t, list(range(len(R))), cfg, gpu, ora)
    finally:
        helpers.K = old
    assert gpu.counts()["grows"] >= 1
