"""Seeded synthetic RGB-D workloads (host side) over lib/libffsynth.so.

Scenes, trajectories and sensor noise for the benchmark configurations:
analytic rooms (walls, axis boxes, spheres), a multi-room floor, circular
and random-walk trajectories, Kinect-style u16 depth (1/5000 m) with
sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2 noise, dropout and flying pixels.
"""
import ctypes
import os

import numpy as np

from . import LIB_DIR, FFError

K_DEFAULT = (525.0, 525.0, 319.5, 239.5)
_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    path = os.path.join(LIB_DIR, "libffsynth.so")
    if not os.path.exists(path):
        raise FFError(f"{path} is missing: run `python -m fastfusion_amd.build`")
    lib = ctypes.CDLL(path)
    P, D = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
    lib.ffs_room_create.restype = P
    lib.ffs_room_create.argtypes = [ctypes.c_uint64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float]
    lib.ffs_floor_create.restype = P
    lib.ffs_floor_create.argtypes = [ctypes.c_uint64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.ffs_destroy.argtypes = [P]
    lib.ffs_render.argtypes = [P, D, D, D, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                               ctypes.c_float, ctypes.c_float, ctypes.c_uint64, P, P, P]
    lib.ffs_trajectory_circle.argtypes = [P, ctypes.c_int, ctypes.c_float, P, P]
    lib.ffs_trajectory_walk.argtypes = [P, ctypes.c_uint64, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                        ctypes.c_float, ctypes.c_float, P, P]
    lib.ffs_sdf.restype = ctypes.c_float
    lib.ffs_sdf.argtypes = [P, ctypes.c_float, ctypes.c_float, ctypes.c_float]
    lib.ffs_sdf_batch.argtypes = [P, P, ctypes.c_size_t, P]
    _lib = lib
    return lib


def _dp(a):
    return np.ascontiguousarray(a, np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Scene:
    """An analytic scene; ``room`` = (lx, ly, lz) in metres."""

    def __init__(self, handle, room):
        self._lib = _load()
        self._h = handle
        self.room = room

    @classmethod
    def room_scene(cls, seed, size=(4.0, 4.0, 3.0), n_boxes=6, n_spheres=4, rmin=0.1, rmax=0.5):
        lib = _load()
        h = lib.ffs_room_create(int(seed), *[float(s) for s in size], int(n_boxes), int(n_spheres),
                                float(rmin), float(rmax))
        return cls(h, size)

    @classmethod
    def floor_scene(cls, seed, size=(12.0, 12.0, 3.0), cells=(3, 3), n_prims=60):
        lib = _load()
        h = lib.ffs_floor_create(int(seed), *[float(s) for s in size], int(cells[0]), int(cells[1]),
                                 int(n_prims))
        return cls(h, size)

    def __del__(self):
        try:
            self._lib.ffs_destroy(self._h)
        except Exception:
            pass

    def render(self, R, t, K=K_DEFAULT, w=640, h=480, noise=1.0, dropout=0.005, flying=0.0, seed=0,
               depth_scale=5000.0, with_gt=False):
        depth = np.zeros((h, w), np.uint16)
        bgr = np.zeros((h, w, 3), np.uint8)
        gt = np.zeros((h, w), np.float32) if with_gt else None
        self._lib.ffs_render(self._h, _dp(R), _dp(t), _dp(K), int(w), int(h), float(depth_scale),
                             float(noise), float(dropout), float(flying), int(seed),
                             depth.ctypes.data_as(ctypes.c_void_p), bgr.ctypes.data_as(ctypes.c_void_p),
                             gt.ctypes.data_as(ctypes.c_void_p) if gt is not None else None)
        return (depth, bgr, gt) if with_gt else (depth, bgr)

    def circle(self, n, radius=1.5):
        R = np.zeros((n, 9), np.float64)
        t = np.zeros((n, 3), np.float64)
        self._lib.ffs_trajectory_circle(self._h, int(n), float(radius), R.ctypes.data_as(ctypes.c_void_p),
                                        t.ctypes.data_as(ctypes.c_void_p))
        return R, t

    def walk(self, n, seed, dmin=0.5, dmax=3.0, step_m=0.02, step_deg=2.0):
        R = np.zeros((n, 9), np.float64)
        t = np.zeros((n, 3), np.float64)
        self._lib.ffs_trajectory_walk(self._h, int(seed), int(n), float(dmin), float(dmax), float(step_m),
                                      float(step_deg), R.ctypes.data_as(ctypes.c_void_p),
                                      t.ctypes.data_as(ctypes.c_void_p))
        return R, t

    def sdf(self, xyz):
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        out = np.zeros(len(xyz), np.float32)
        self._lib.ffs_sdf_batch(self._h, xyz.ctypes.data_as(ctypes.c_void_p), len(xyz),
                                out.ctypes.data_as(ctypes.c_void_p))
        return out


def config_c1(n_frames=200, seed_scene=1, w=640, h=480):
    """Plain method, CPU-runnable: 4x4x3 m room, 6 boxes, 4 spheres (0.1-0.5 m),
    circular trajectory of radius 1.5 m, 0.5 % dropout; voxel 10 mm, 3-voxel
    truncation, max depth 4 m."""
    sc = Scene.room_scene(seed_scene, (4.0, 4.0, 3.0), 6, 4, 0.1, 0.5)
    R, t = sc.circle(n_frames, 1.5)
    return sc, R, t, dict(scale=0.010, threshold=0.030, maxcam=4.0, w=w, h=h, dropout=0.005, noise=1.0)


def config_c2(n_frames=1000, seed_scene=11, seed_traj=12, w=640, h=480):
    """Plain method at full size: the same room generator, seeded random walk
    0.5-3 m from surfaces in 2 cm / 2 degree steps; voxel 5 mm, 3-voxel
    truncation."""
    sc = Scene.room_scene(seed_scene, (4.0, 4.0, 3.0), 6, 4, 0.1, 0.5)
    R, t = sc.walk(n_frames, seed_traj, 0.5, 3.0, 0.02, 2.0)
    return sc, R, t, dict(scale=0.005, threshold=0.015, maxcam=4.0, w=w, h=h, dropout=0.005, noise=1.0)


def config_c5(n_frames=3000, seed=3000, w=1280, h=960):
    """Large map: a 12x12x3 m floor of 3x3 rooms (interior walls with door
    gaps) and 60 primitives, a seeded random walk 0.5-3 m from surfaces in
    2 cm / 2 degree steps, 1280x960 depth (fx = fy = 1050); voxel 2.5 mm,
    4-voxel truncation (10 mm), max depth 4 m."""
    sc = Scene.floor_scene(seed, (12.0, 12.0, 3.0), (3, 3), 60)
    R, t = sc.walk(n_frames, seed + 1, 0.5, 3.0, 0.02, 2.0)
    f = 525.0 * w / 640.0
    K = (f, f, (w - 1) / 2.0, (h - 1) / 2.0)
    return sc, R, t, dict(scale=0.0025, threshold=0.010, maxcam=4.0, w=w, h=h, dropout=0.005, noise=1.0, K=K)


def stream_configs(n):
    """The batched-streams workload: stream i (seed 100 + i) has a room of
    2-8 m per horizontal side and 2.5-3.5 m height with 3-20 primitives, a
    depth resolution from {320x240, 640x480, 1280x960} (fx = fy = 525 w / 640),
    4-10 mm voxels, 3-voxel truncation, a random walk 0.5-3 m from surfaces
    (walk seed = stream seed + 7), frame i rendered with noise seed
    seed * 10007 + i."""
    out = []
    for i in range(n):
        rng = np.random.default_rng(100 + i)
        size = (float(rng.uniform(2, 8)), float(rng.uniform(2, 8)), float(rng.uniform(2.5, 3.5)))
        nprim = int(rng.integers(3, 21))
        nb = int(rng.integers(0, nprim + 1))
        w, h = [(320, 240), (640, 480), (1280, 960)][int(rng.integers(0, 3))]
        f = 525.0 * w / 640.0
        vox = float(rng.uniform(0.004, 0.010))
        sc = Scene.room_scene(100 + i, size, nb, nprim - nb, 0.1, 0.5)
        out.append(dict(scene=sc, size=size, prims=nprim, w=w, h=h, K=(f, f, (w - 1) / 2.0, (h - 1) / 2.0),
                        scale=vox, threshold=3 * vox, seed=100 + i, maxcam=4.0))
    return out
