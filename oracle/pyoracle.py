"""ctypes binding of the oracle (oracle/lib/libfforacle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    path = os.path.join(HERE, "lib", "libfforacle.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run `python oracle/build.py`")
    lib = ctypes.CDLL(path)
    P, U32 = ctypes.c_void_p, ctypes.c_uint32
    D = ctypes.POINTER(ctypes.c_double)
    lib.ffo_create.restype = P
    lib.ffo_create.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_int, U32, U32]
    lib.ffo_destroy.argtypes = [P]
    lib.ffo_add_map.restype = ctypes.c_int
    lib.ffo_add_map.argtypes = [P, P, P, ctypes.c_int, ctypes.c_int, D, D, D, ctypes.c_float, ctypes.c_float]
    for name in ("ffo_n_leaves", "ffo_n_branches", "ffo_side", "ffo_n_queued"):
        getattr(lib, name).restype = U32
        getattr(lib, name).argtypes = [P]
    lib.ffo_grow_count.restype = ctypes.c_int
    lib.ffo_grow_count.argtypes = [P]
    lib.ffo_offset.argtypes = [P, P]
    for name in ("ffo_queue", "ffo_tree", "ffo_leaf_pos", "ffo_leaf_scale", "ffo_leaf_parent",
                 "ffo_distance", "ffo_weights", "ffo_color", "ffo_processed"):
        getattr(lib, name).restype = P
        getattr(lib, name).argtypes = [P]
    lib.ffo_rcp.restype = ctypes.c_float
    lib.ffo_rcp.argtypes = [ctypes.c_float]
    lib.ffo_leaf_scale_func.restype = ctypes.c_uint16
    lib.ffo_leaf_scale_func.argtypes = [ctypes.c_float]
    lib.ffo_dedup.argtypes = [P, P, P, P, ctypes.c_int, ctypes.c_int, P]
    lib.ffo_update_brick.argtypes = [P, ctypes.c_float, ctypes.c_float, P, ctypes.c_int, ctypes.c_int, P, P,
                                     ctypes.c_float, ctypes.c_float, P, ctypes.c_uint16, P, P, P]
    lib.ffo_set_confidence.argtypes = [P, P]
    _mesh_sigs(lib)
    _loop_sigs(lib)
    _lib = lib
    return lib


def _mesh_sigs(lib):
    P = ctypes.c_void_p
    lib.ffo_mesh_all.restype = ctypes.c_int
    lib.ffo_mesh_all.argtypes = [P]
    lib.ffo_mesh_counts.argtypes = [P, P]
    lib.ffo_mesh_copy.argtypes = [P, P, P, P, P, P, P]
    lib.ffo_mesh_note_frame.argtypes = [P]
    lib.ffo_mesh_update.restype = ctypes.c_int
    lib.ffo_mesh_update.argtypes = [P]
    lib.ffo_mesh_update_faces.restype = ctypes.c_size_t
    lib.ffo_mesh_update_faces.argtypes = []
    lib.ffo_mesh_update_sizes.argtypes = [P]
    lib.ffo_mesh_update_all.argtypes = [P, P, P, P]
    lib.ffo_mesh_set_depth_cap.argtypes = [ctypes.c_int]
    lib.ffo_mesh_update_copy.argtypes = [P, P, P, P, P, P]


def _loop_sigs(lib):
    P, U32, I = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int
    D = ctypes.POINTER(ctypes.c_double)
    lib.ffo_loop_create.restype = P
    lib.ffo_loop_create.argtypes = [U32, U32, I]
    lib.ffo_loop_destroy.argtypes = [P]
    lib.ffo_loop_reset.argtypes = [P]
    lib.ffo_loop_delete.restype = I
    lib.ffo_loop_delete.argtypes = [P, P, P, P, P, P, I, I, D, D, D, P, U32, P]
    lib.ffo_delete_brick.argtypes = [P, P, P, P, I, I, P, P, ctypes.c_float, ctypes.c_float, P, ctypes.c_uint16,
                                     P, P, P]
    lib.ffo_camera_inverse.argtypes = [P, D, D, D, P, P]
    lib.ffo_loop_mapped.restype = U32
    lib.ffo_loop_mapped.argtypes = [P]
    for name in ("ffo_loop_map", "ffo_loop_distance", "ffo_loop_weights", "ffo_loop_color"):
        getattr(lib, name).restype = P
        getattr(lib, name).argtypes = [P]
    lib.ffo_leaves_total.restype = U32
    lib.ffo_leaves_total.argtypes = [P]


def _vp(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _dp(a):
    return np.ascontiguousarray(a, np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _arr(ptr, dtype, n):
    if n == 0:
        return np.zeros(0, dtype)
    buf = (ctypes.c_char * (n * np.dtype(dtype).itemsize)).from_address(ptr)
    return np.frombuffer(buf, dtype=dtype, count=n).copy()


class Oracle:
    """Sequential CPU restatement of FusionMipMapCPU's integration path."""

    def __init__(self, scale, threshold, color=True, max_leaves=1 << 16, max_branches=1 << 21):
        self.lib = load()
        self.h = self.lib.ffo_create(float(scale), float(threshold), int(bool(color)), int(max_leaves),
                                     int(max_branches))
        self.color = color

    def __del__(self):
        try:
            self.lib.ffo_destroy(self.h)
        except Exception:
            pass

    def add_map(self, depth, R, t, K, bgr=None, scaling=1.0 / 5000.0, maxcam=10.0, conf=None):
        """One frame; conf: optional H x W per-pixel confidence multiplying the weight."""
        depth = np.ascontiguousarray(depth, np.uint16)
        h, w = depth.shape
        bgr = np.ascontiguousarray(bgr, np.uint8) if bgr is not None else np.zeros((h, w, 3), np.uint8)
        cf = None
        if conf is not None:
            cf = np.ascontiguousarray(conf, np.float32).reshape(h, w)
        self.lib.ffo_set_confidence(self.h, _vp(cf) if cf is not None else None)
        nq = self.lib.ffo_add_map(self.h, _vp(depth), _vp(bgr), w, h, _dp(R), _dp(t), _dp(K),
                                  float(scaling), float(maxcam))
        self.lib.ffo_set_confidence(self.h, None)
        self.lib.ffo_mesh_note_frame(self.h)
        return nq

    def counts(self):
        off = np.zeros(3, np.float32)
        self.lib.ffo_offset(self.h, _vp(off))
        return {"n_leaves_used": self.lib.ffo_n_leaves(self.h), "n_branches_used": self.lib.ffo_n_branches(self.h),
                "n_queued": self.lib.ffo_n_queued(self.h), "side": self.lib.ffo_side(self.h), "offset": off,
                "grows": self.lib.ffo_grow_count(self.h)}

    def queue(self):
        n = self.lib.ffo_n_queued(self.h)
        return _arr(self.lib.ffo_queue(self.h), np.uint32, n)

    def tree(self):
        return _arr(self.lib.ffo_tree(self.h), np.uint32, self.lib.ffo_n_branches(self.h))

    def leaves(self):
        n = self.lib.ffo_n_leaves(self.h)
        pos = _arr(self.lib.ffo_leaf_pos(self.h), np.uint16, 3 * n).reshape(n, 3)
        sc = _arr(self.lib.ffo_leaf_scale(self.h), np.uint16, n)
        par = _arr(self.lib.ffo_leaf_parent(self.h), np.uint32, n)
        return pos, sc, par

    def bricks(self, leaves):
        n = self.lib.ffo_n_leaves(self.h)
        d = _arr(self.lib.ffo_distance(self.h), np.float32, 512 * n).reshape(n, 512)
        w = _arr(self.lib.ffo_weights(self.h), np.float32, 512 * n).reshape(n, 512)
        c = _arr(self.lib.ffo_color(self.h), np.uint16, 3 * 512 * n).reshape(n, 512, 3).transpose(0, 2, 1)
        idx = np.asarray(leaves, np.int64)
        return d[idx], w[idx], c[idx]

    def processed(self, w, h):
        return _arr(self.lib.ffo_processed(self.h), np.uint8, w * h).reshape(h, w)

    def update_meshes(self):
        """Incremental re-mesh of the cells affected since the last call plus the
        concatenation of all cell meshes (CPU-baseline timing path)."""
        return self.lib.ffo_mesh_update(self.h)

    def mesh_incremental(self):
        """The concatenated mesh the last update_meshes produced: (xyz, rgb, faces)."""
        cnt = np.zeros(2, np.uint64)
        self.lib.ffo_mesh_update_all(_vp(cnt), None, None, None)
        nv, nf = (int(v) for v in cnt)
        xyz = np.zeros((nv, 3), np.float32)
        rgb = np.zeros((nv, 3), np.uint8)
        faces = np.zeros((nf, 3), np.uint32)
        self.lib.ffo_mesh_update_all(_vp(cnt), _vp(xyz), _vp(rgb), _vp(faces))
        return xyz, rgb, faces

    def cell_update(self):
        """The cells the last update_meshes re-meshed (CellUpdate), in key
        order: (desc (n, 13), nv, nf, xyz, rgb, faces local per cell)."""
        cnt = np.zeros(3, np.uint64)
        self.lib.ffo_mesh_update_sizes(_vp(cnt))
        nc, nv, nf = (int(v) for v in cnt)
        desc = np.zeros((nc, 13), np.int32)
        cv = np.zeros(nc, np.uint32)
        cf = np.zeros(nc, np.uint32)
        xyz = np.zeros((nv, 3), np.float32)
        rgb = np.zeros((nv, 3), np.uint8)
        faces = np.zeros((nf, 3), np.uint32)
        self.lib.ffo_mesh_update_copy(_vp(desc), _vp(cv), _vp(cf), _vp(xyz), _vp(rgb), _vp(faces))
        return desc, cv, cf, xyz, rgb, faces

    def mesh_all(self):
        """Meshes every meshable cell from scratch.  Returns (desc (n, 13): type,
        ox, oy, oz, size, lastLeaf[8]; nv; nf; xyz; rgb; faces local per cell)."""
        self.lib.ffo_mesh_all(self.h)
        cnt = np.zeros(3, np.uint64)
        self.lib.ffo_mesh_counts(self.h, _vp(cnt))
        nc, nv, nf = (int(v) for v in cnt)
        desc = np.zeros((nc, 13), np.int32)
        cv = np.zeros(nc, np.uint32)
        cf = np.zeros(nc, np.uint32)
        xyz = np.zeros((nv, 3), np.float32)
        rgb = np.zeros((nv, 3), np.uint8)
        faces = np.zeros((nf, 3), np.uint32)
        self.lib.ffo_mesh_copy(self.h, _vp(desc), _vp(cv), _vp(cf), _vp(xyz), _vp(rgb), _vp(faces))
        return desc, cv, cf, xyz, rgb, faces


def rcp(x):
    lib = load()
    return np.array([lib.ffo_rcp(float(v)) for v in np.asarray(x, np.float32).ravel()], np.float32)


def leaf_scale_func(z):
    return load().ffo_leaf_scale_func(float(z))


def dedup(vx, vy, vz, valid):
    h, w = valid.shape
    out = np.zeros((h, w), np.uint8)
    load().ffo_dedup(_vp(np.ascontiguousarray(vx, np.int32)), _vp(np.ascontiguousarray(vy, np.int32)),
                     _vp(np.ascontiguousarray(vz, np.int32)), _vp(np.ascontiguousarray(valid, np.uint8)),
                     w, h, _vp(out))
    return out


def update_brick(depth, bgr, scaling, maxcam, m12, K, scale, threshold, pos, leaf_scale, dist, weight, color):
    """One brick through the oracle's voxel update; arrays are updated in place."""
    depth = np.ascontiguousarray(depth, np.uint16)
    h, w = depth.shape
    bgr = np.ascontiguousarray(bgr, np.uint8)
    m = np.ascontiguousarray(m12, np.float32)
    Kf = np.ascontiguousarray(K, np.float32)
    p = np.ascontiguousarray(pos, np.uint16)
    load().ffo_update_brick(_vp(depth), float(scaling), float(maxcam), _vp(bgr), w, h, _vp(m), _vp(Kf),
                            float(scale), float(threshold), _vp(p), int(leaf_scale), _vp(dist), _vp(weight),
                            _vp(color))


class LoopOracle:
    """The loop-closure partner map (ff_oracle_loop.c): `total` slots for the
    bricks of the oracle map `main`, handed out in queue order."""

    def __init__(self, main, total):
        self.lib = load()
        self.main = main
        self.map_n = self.lib.ffo_leaves_total(main.h)
        self.total = int(total)
        self.h = self.lib.ffo_loop_create(self.map_n, self.total, int(bool(main.color)))
        self.color = main.color

    def __del__(self):
        try:
            self.lib.ffo_loop_destroy(self.h)
        except Exception:
            pass

    def reset(self):
        self.lib.ffo_loop_reset(self.h)

    def delete_frame(self, depth_m, rgb_planes, R, t, K, queue):
        """depth_m: H x W float32 metres (NaN = no measurement); rgb_planes:
        (red, green, blue) H x W uint8 or None.  Returns (ok, applied)."""
        d = np.ascontiguousarray(depth_m, np.float32)
        h, w = d.shape
        pl = [np.ascontiguousarray(p, np.uint8) for p in rgb_planes] if rgb_planes is not None else [None] * 3
        q = np.ascontiguousarray(queue, np.uint32)
        applied = np.zeros(1, np.uint32)
        ok = self.lib.ffo_loop_delete(self.h, self.main.h, _vp(d), _vp(pl[0]), _vp(pl[1]), _vp(pl[2]), w, h,
                                      _dp(R), _dp(t), _dp(K), _vp(q), len(q), _vp(applied))
        return bool(ok), int(applied[0])

    def mapped(self):
        return self.lib.ffo_loop_mapped(self.h)

    def slot_map(self):
        return _arr(self.lib.ffo_loop_map(self.h), np.uint32, self.map_n)

    def bricks(self, slots):
        n = self.total
        idx = np.asarray(slots, np.int64)
        d = _arr(self.lib.ffo_loop_distance(self.h), np.float32, 512 * n).reshape(n, 512)[idx]
        w = _arr(self.lib.ffo_loop_weights(self.h), np.float32, 512 * n).reshape(n, 512)[idx]
        c = None
        if self.color:
            c = _arr(self.lib.ffo_loop_color(self.h), np.uint16, 3 * 512 * n).reshape(n, 512, 3)
            c = c[idx].transpose(0, 2, 1)
        return d, w, c


def camera_inverse(main, R, t, K):
    """addMap's pInv for the oracle map `main`: (m (12,), K (4,)) float32."""
    m = np.zeros(12, np.float32)
    Kf = np.zeros(4, np.float32)
    load().ffo_camera_inverse(main.h, _dp(R), _dp(t), _dp(K), _vp(m), _vp(Kf))
    return m, Kf


def delete_brick(depth_m, planes, m12, K, scale, threshold, pos, leaf_scale, dist, weight, color):
    """One brick through the oracle's de-integration kernel, in place; color
    is (512, 3) R, G, B per voxel or None."""
    d = np.ascontiguousarray(depth_m, np.float32)
    h, w = d.shape
    pl = [np.ascontiguousarray(p, np.uint8) for p in planes] if planes is not None else [None] * 3
    load().ffo_delete_brick(_vp(d), _vp(pl[0]), _vp(pl[1]), _vp(pl[2]), w, h,
                            _vp(np.ascontiguousarray(m12, np.float32)), _vp(np.ascontiguousarray(K, np.float32)),
                            float(scale), float(threshold), _vp(np.ascontiguousarray(pos, np.uint16)),
                            int(leaf_scale), _vp(dist), _vp(weight), _vp(color) if color is not None else None)
