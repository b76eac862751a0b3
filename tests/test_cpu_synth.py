"""Synthetic scene generator: exact ray casting and signed distance."""
import numpy as np

from fastfusion_amd import synth


def test_render_flat_wall_depth():
    sc = synth.Scene.room_scene(1, (4.0, 4.0, 3.0), 0, 0)
    # camera at (1, 2, 1.5) looking along -x at the x = 0 wall
    R = np.array([[0, 0, -1], [1, 0, 0], [0, -1, 0]], np.float64)  # columns: x_c, y_c, z_c in world
    R = np.stack([[0, 1, 0], [0, 0, -1], [-1, 0, 0]], 1).astype(np.float64)
    d, bgr, gt = sc.render(R, np.array([1.0, 2.0, 1.5]), (262.5, 262.5, 159.5, 119.5), 320, 240, noise=0.0,
                           dropout=0.0, with_gt=True)
    assert abs(gt[120, 160] - 1.0) < 1e-5
    assert d[120, 160] == 5000
    # depth along camera z is constant on a fronto-parallel plane
    assert np.allclose(gt, 1.0, atol=1e-4)


def test_sdf_room_and_sphere():
    sc = synth.Scene.room_scene(2, (4.0, 4.0, 3.0), 0, 0)
    assert abs(sc.sdf([[1.0, 2.0, 1.5]])[0] - 1.0) < 1e-6
    assert abs(sc.sdf([[2.0, 2.0, 0.25]])[0] - 0.25) < 1e-6


def test_noise_model_statistics():
    sc = synth.Scene.room_scene(3, (4.0, 4.0, 3.0), 0, 0)
    R = np.stack([[0, 1, 0], [0, 0, -1], [-1, 0, 0]], 1).astype(np.float64)
    d, _, gt = sc.render(R, np.array([2.4, 2.0, 1.5]), (262.5, 262.5, 159.5, 119.5), 320, 240, noise=1.0,
                         dropout=0.005, seed=5, with_gt=True)
    ok = d > 0
    err = d[ok] / 5000.0 - gt[ok]
    sigma = 0.0012 + 0.0019 * (2.4 - 0.4) ** 2
    assert abs(err.std() - sigma) < 0.25 * sigma
    assert 0.002 < (~ok).mean() < 0.01
