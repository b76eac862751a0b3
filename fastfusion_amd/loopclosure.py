"""Loop-closure de-integration: the reference's loop-closure partner map.

When a loop closure moves the poses of frames already fused, the reference
subtracts those frames from a partner map (FusionMipMapCPU(FusionMipMapCPU*,
nLeavesTotalLoop), geometryfusion_mipmap_cpu.cpp:645-717):

* ``startLoopClosure`` (:4033-4067) orders the old poses by how far the new
  trajectory moved them (``pose_order``);
* ``deleteOldLoopPoses`` (:3934-4024) re-reads each such frame's depth (u16
  PNG scaled by 1/5000, zero -> NaN: ``depth_metres``) and color planes, and
  runs ``updateWrapper_deleteloop`` (loopclosure.cpp:232-311) over the brick
  queue recorded when the frame was fused (``_leafQueues``), which hands
  each brick a slot of the partner on first touch and applies
  ``update8AddLoopSSESingleDeleteloop`` (loopclosure.cpp:15-230) there.

``LoopClosure`` is that partner, in HBM, over the C ABI of libffgpu.so
(``ff_loop_*`` in include/ff_abi.h); every de-integration runs in the HIP
library.  ``addNewLoopPoses`` is empty in the reference (:4026-4029) and
has no counterpart here.
"""
import ctypes

import numpy as np

from . import FFError, _LoopFrame, _SZ, _dptr, _ptr, load_library

MAX_LOOP_CLOSURE_LEAVES = 100000  # definitions.h:124
MAX_NUMBER_POSES_PER_LOOP = 50    # definitions.h:125
BRANCHINIT = 4294967294           # definitions.h:130
FF_E_POOL_FULL = -3


def depth_metres(depth, scale=1.0 / 5000.0):
    """A u16 depth image as deleteOldLoopPoses reads it back
    (geometryfusion_mipmap_cpu.cpp:3950-3951): convertTo(CV_32FC1, 1/5000),
    i.e. float(d) * float(1/5000), then IMAGEINFINITE (quiet NaN) where the
    value is 0."""
    d = np.asarray(depth)
    m = d.astype(np.float32) * np.float32(scale)
    m[m == 0.0] = np.nan
    return m


def planes(bgr):
    """(red, green, blue) byte planes of a BGR8 image, as contiguous arrays."""
    bgr = np.asarray(bgr, np.uint8)
    return tuple(np.ascontiguousarray(bgr[:, :, k]) for k in (2, 1, 0))


def pose_order(t_loop, t_fused, max_poses=None):
    """The order in which deleteOldLoopPoses visits the poses: the
    std::multimap<float, size_t> of startLoopClosure
    (geometryfusion_mipmap_cpu.cpp:4045-4054), keyed by minus the squared
    change of the translation (summed in double, stored as float), ascending,
    equal keys in pose order; at most max_poses of them
    (_maxPosesPerLoop, :3940)."""
    a = np.asarray(t_loop, np.float64).reshape(-1, 3)
    b = np.asarray(t_fused, np.float64).reshape(-1, 3)
    if len(a) != len(b):
        raise ValueError("the new and the fused trajectory must have the same length")
    d = a - b
    key = (-((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])).astype(np.float32)
    order = np.argsort(key, kind="stable").tolist()
    return order if max_poses is None else order[:max_poses]


class LoopClosure:
    """The loop-closure partner of a FusionMipMapGPU map: ``max_leaves`` slots
    of 512 float distances, float weights and colors."""

    def __init__(self, fusion, max_leaves=MAX_LOOP_CLOSURE_LEAVES):
        self._lib = load_library()
        self.fusion = fusion  # the partner reads the map's bricks: keep it alive
        self.color = fusion.color
        self._h = self._lib.ff_loop_create(fusion.handle, int(max_leaves))
        if not self._h:
            raise FFError("ff_loop_create failed: " + self._lib.ff_last_error(fusion.handle).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ff_loop_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self, rc):
        return FFError(self._lib.ff_loop_last_error(self._h).decode() or f"error {rc}")

    def _check(self, rc):
        if rc < 0:
            raise self._err(rc)
        return rc

    def reset(self):
        """updateArraysForLoopClosure's reset (:3927-3931)."""
        self._check(self._lib.ff_loop_reset(self._h))

    def delete_frame(self, depth, color, R, t, K, queue):
        """Subtracts one frame from the queued bricks (updateWrapper_deleteloop).

        depth: H x W float32 metres with NaN for no measurement, or u16 raw
        depth (converted by ``depth_metres``); color: BGR8 image, a (red,
        green, blue) tuple of planes, or None for a map without color.
        Returns (ok, applied): ok is False when a brick found no free slot,
        and then only the first ``applied`` queue entries were applied."""
        d = np.asarray(depth)
        if d.dtype == np.uint16:
            d = depth_metres(d)
        d = np.ascontiguousarray(d, np.float32)
        h, w = d.shape
        pl = (None, None, None)
        if color is not None and self.color:
            if isinstance(color, (tuple, list)):
                pl = tuple(np.ascontiguousarray(p, np.uint8) for p in color)
            else:
                pl = planes(color)
            if any(p.shape != (h, w) for p in pl):
                raise FFError("color planes must be H x W like depth")
        q = np.ascontiguousarray(queue, np.uint32)
        applied = _SZ()
        rc = self._lib.ff_loop_delete_frame(self._h, _ptr(d), _ptr(pl[0]), _ptr(pl[1]), _ptr(pl[2]), w, h,
                                            _dptr(R), _dptr(t), _dptr(K), _ptr(q) if len(q) else None, len(q),
                                            ctypes.byref(applied))
        if rc == FF_E_POOL_FULL:
            return False, applied.value
        self._check(rc)
        return True, applied.value

    def delete_frame_device(self, depth_ptr, red_ptr, green_ptr, blue_ptr, w, h, R, t, K, queue_ptr, n):
        """Same, with device pointers (float depth, byte planes, u32 queue)."""
        applied = _SZ()
        vp = ctypes.c_void_p
        rc = self._lib.ff_loop_delete_frame_device(self._h, vp(depth_ptr), vp(red_ptr), vp(green_ptr), vp(blue_ptr),
                                                   int(w), int(h), _dptr(R), _dptr(t), _dptr(K), vp(queue_ptr),
                                                   int(n), ctypes.byref(applied))
        if rc == FF_E_POOL_FULL:
            return False, applied.value
        self._check(rc)
        return True, applied.value

    @staticmethod
    def prepare_frames(frames):
        """The ff_loop_frame array of frames = [(depth_ptr, red_ptr, green_ptr,
        blue_ptr, w, h, R, t, K, queue_ptr, n_queue)] (device pointers), for
        delete_frames."""
        n = len(frames)
        arr = (_LoopFrame * max(n, 1))()
        for k, (dp, rp, gp, bp, w, h, R, t, K, qp, nq) in enumerate(frames):
            f = arr[k]
            f.depth, f.red, f.green, f.blue = dp, rp or None, gp or None, bp or None
            f.width, f.height = int(w), int(h)
            f.R[:] = [float(x) for x in np.asarray(R, np.float64).ravel()]
            f.t[:] = [float(x) for x in np.asarray(t, np.float64).ravel()]
            f.K[:] = [float(x) for x in np.asarray(K, np.float64).ravel()]
            f.queue, f.n_queue = qp or None, int(nq)
        return arr, n

    def delete_frames(self, frames):
        """deleteOldLoopPoses' loop as one submission (ff_loop_delete_frames)
        over a list as for prepare_frames, or its result.  Returns (ok,
        applied per frame); a frame naming an invalid brick raises after the
        others ran."""
        arr, n = frames if isinstance(frames, tuple) else self.prepare_frames(frames)
        applied = np.zeros(max(n, 1), np.uintp)
        rc = self._lib.ff_loop_delete_frames(self._h, ctypes.cast(arr, ctypes.c_void_p), n, _ptr(applied))
        if rc == FF_E_POOL_FULL:
            return False, applied[:n].astype(np.int64)
        self._check(rc)
        return True, applied[:n].astype(np.int64)

    def mapped(self):
        """_nLeavesMapped."""
        n, cap = ctypes.c_uint32(), ctypes.c_uint32()
        self._check(self._lib.ff_loop_mapped(self._h, ctypes.byref(n), ctypes.byref(cap)))
        return n.value

    def capacity(self):
        n, cap = ctypes.c_uint32(), ctypes.c_uint32()
        self._check(self._lib.ff_loop_mapped(self._h, ctypes.byref(n), ctypes.byref(cap)))
        return cap.value

    def slot_map(self, n=None):
        """_leafMapLoop for the map's first n bricks (default: the used ones)."""
        n = self.fusion.counts()["n_leaves_used"] if n is None else int(n)
        out = np.zeros(max(n, 1), np.uint32)
        self._check(self._lib.ff_loop_copy_map(self._h, _ptr(out), n))
        return out[:n]

    def bricks(self, slots):
        """(dist (n, 512), weights (n, 512), colors (n, 3, 512) R, G, B or None)."""
        s = np.ascontiguousarray(slots, np.uint32)
        n = len(s)
        d = np.zeros((n, 512), np.float32)
        w = np.zeros((n, 512), np.float32)
        c = np.zeros((n, 3, 512), np.uint16) if self.color else None
        self._check(self._lib.ff_loop_copy_bricks(self._h, _ptr(s), n, _ptr(d), _ptr(w), _ptr(c)))
        return d, w, c

    def set_timing(self, on=True):
        self._check(self._lib.ff_loop_set_timing(self._h, int(bool(on))))

    def timing(self):
        """(milliseconds, calls, algorithmic bytes, applied queue entries) of the
        de-integration kernel since timing was turned on."""
        ms, calls, by, units = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_double(), ctypes.c_uint64()
        self._check(self._lib.ff_loop_get_timing(self._h, ctypes.byref(ms), ctypes.byref(calls), ctypes.byref(by),
                                                 ctypes.byref(units)))
        return ms.value, calls.value, by.value, units.value

    def delete_old_poses(self, frame, poses_loop, poses_fused, queues, K, max_poses=MAX_NUMBER_POSES_PER_LOOP):
        """deleteOldLoopPoses (geometryfusion_mipmap_cpu.cpp:3934-4024): for the
        poses in ``pose_order``, subtracts frame p = ``frame(p)`` -> (u16 depth,
        BGR8 or None) at pose ``poses_loop[p]`` = (R, t) over ``queues[p]``,
        then records the pose as fused (``poses_fused[p] = poses_loop[p]``).
        Returns the poses visited."""
        t_loop = [np.asarray(p[1], np.float64) for p in poses_loop]
        t_fused = [np.asarray(p[1], np.float64) for p in poses_fused]
        order = pose_order(t_loop, t_fused, max_poses)
        for p in order:
            depth, bgr = frame(p)
            R, t = poses_loop[p]
            self.delete_frame(depth_metres(depth), bgr, R, t, K, queues[p])
            poses_fused[p] = poses_loop[p]
        return order
