"""GPU parity on large frames: the 1280x960, 2.5 mm floor workload (several
thousand bricks queued per frame, more than one per block of the SDF
launch's grid) and a first frame creating more nodes than the one-block
ordering pass ranks (the per-pixel count-scan path), both bit for bit
against the oracle; and the floor workload with learned weighting on."""
import numpy as np
import pytest

from fastfusion_amd import learned, synth
from tests.helpers import make_pair, run_pair

pytestmark = pytest.mark.gpu

# bricks per SDF launch grid (FF_SDF_GRID in ff_internal.hpp): frames above
# twice this run the launch's grid-stride loop at least twice per block
SDF_GRID = 1024
# new tree nodes one frame may create before the ordering falls back to the
# per-pixel count scan (FF_ORDER_MAX)
ORDER_MAX = 8192


def test_floor_1280_frames_bit_exact():
    """First frames of the 12x12x3 m floor at 1280x960, 2.5 mm voxels, 10 mm
    truncation: each frame queues ~3 800 bricks.  Every queued brick is
    updated once per frame (update_functions.cpp:1683-1727)."""
    sc, R, t, cfg = synth.config_c5(n_frames=8)
    gpu, ora = make_pair(cfg["scale"], cfg["threshold"], max_leaves=1 << 16, max_branches=1 << 22)
    queued = run_pair(sc, R, t, list(range(4)), cfg, gpu, ora)
    assert max(queued) > 2 * SDF_GRID, queued


def test_floor_1280_learned_bit_exact():
    """The same frames with the committed confidence network on: the oracle
    is fed the GPU's per-pixel confidences and must give the same map."""
    sc, R, t, cfg = synth.config_c5(n_frames=8)
    gpu, ora = make_pair(cfg["scale"], cfg["threshold"], max_leaves=1 << 16, max_branches=1 << 22)
    gpu.setLearnedWeighting(learned.load_weights())
    queued = run_pair(sc, R, t, list(range(3)), cfg, gpu, ora, learned=True)
    assert max(queued) > 2 * SDF_GRID, queued


def test_big_first_frame_count_scan_path():
    """A 1280x960 first frame at 1.5 mm voxels creates ~10 700 bricks and
    ~2 700 branches in one frame, beyond the one-block ordering: tree slots,
    brick indices and parents from the per-pixel count scan, and the SDF
    update that follows it, equal the oracle's; then two more frames."""
    sc, R, t, cfg = synth.config_c5(n_frames=8)
    cfg = dict(cfg, scale=0.0015, threshold=0.006)
    gpu, ora = make_pair(cfg["scale"], cfg["threshold"], max_leaves=1 << 16, max_branches=1 << 22)
    run_pair(sc, R, t, [0], cfg, gpu, ora)
    c = gpu.counts()
    assert c["n_leaves_used"] + c["n_branches_used"] // 16 > ORDER_MAX, c
    run_pair(sc, R, t, [1, 2], cfg, gpu, ora)
