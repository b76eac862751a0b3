"""The input side of the reference's fusion loop: TUM RGB-D association
files and PNG images (onlinefusionviewer_main.cpp:266-346,
onlinefusionviewer.cpp:896-959), over the library's host functions
(ff_tum_read, ff_png_read_depth, ff_png_read_bgr in include/ff_abi.h).

    frames = read_associations(["rgbd_dataset_freiburg1_xyz/associate.txt"])
    fuse(frames, fusion)            # addMap + updateMeshes per frame
"""
import ctypes

import numpy as np

from . import FFError, load_library

# the reference's defaults (onlinefusionviewer_main.cpp:234-237; definitions.h:202, 302)
K_DEFAULT = (525.0, 525.0, 319.5, 239.5)
DEPTH_SCALE = 5000.0
MAX_CAM_DISTANCE = 10.0


class _TumFrame(ctypes.Structure):
    _fields_ = [("R", ctypes.c_double * 9), ("t", ctypes.c_double * 3), ("depth", ctypes.c_char * 512),
                ("rgb", ctypes.c_char * 512)]


def _err(lib, rc, what):
    if rc < 0:
        raise FFError(f"{what}: {lib.ff_input_last_error().decode()}")
    return rc


def read_associations(paths, use_color=True):
    """Frames of association files: a list of (R 3x3 camera -> world, t, depth
    path, rgb path), poses of each further file composed with the last pose
    of the previous one, as the reference's viewer builds its trajectory."""
    lib = load_library()
    if isinstance(paths, (str, bytes)):
        paths = [paths]
    arr = (ctypes.c_char_p * len(paths))(*[p.encode() if isinstance(p, str) else p for p in paths])
    n = _err(lib, lib.ff_tum_read(arr, len(paths), int(bool(use_color)), None, 0), "ff_tum_read")
    out = (_TumFrame * max(n, 1))()
    _err(lib, lib.ff_tum_read(arr, len(paths), int(bool(use_color)), out, n), "ff_tum_read")
    return [(np.array(f.R[:]).reshape(3, 3), np.array(f.t[:]), f.depth.decode(), f.rgb.decode()) for f in out[:n]]


def png_size(path):
    lib = load_library()
    w, h, c, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _err(lib, lib.ff_png_size(path.encode(), ctypes.byref(w), ctypes.byref(h), ctypes.byref(c), ctypes.byref(b)), path)
    return w.value, h.value, c.value, b.value


def read_depth_png(path):
    """cv::imread(path, -1) of a single-channel depth PNG: (H, W) uint16."""
    lib = load_library()
    w, h, _, _ = png_size(path)
    out = np.zeros((h, w), np.uint16)
    _err(lib, lib.ff_png_read_depth(path.encode(), out.ctypes.data, out.size), path)
    return out


def read_color_png(path):
    """cv::imread(path) of a color PNG: (H, W, 3) uint8, B, G, R."""
    lib = load_library()
    w, h, _, _ = png_size(path)
    out = np.zeros((h, w, 3), np.uint8)
    _err(lib, lib.ff_png_read_bgr(path.encode(), out.ctypes.data, out.size), path)
    return out


def fuse(frames, fusion, K=K_DEFAULT, depth_scale=DEPTH_SCALE, max_cam_distance=MAX_CAM_DISTANCE, mesh=True,
         first=0, count=None):
    """The reference's per-frame loop (onlinefusionviewer.cpp:934-952):
    addMap(depth, pose, rgb, 1 / imageDepthScale, maxCamDistance), then
    updateMeshes.  Returns the number of updated bricks per frame."""
    out = []
    end = len(frames) if count is None else min(len(frames), first + count)
    for R, t, dpath, rpath in frames[first:end]:
        depth = read_depth_png(dpath)
        rgb = read_color_png(rpath) if (rpath and fusion.color) else None
        out.append(fusion.addMap(depth, R, t, K, rgb, 1.0 / depth_scale, max_cam_distance))
        if mesh:
            fusion.updateMeshes()
    return out
