"""fastfusion_amd -- MI355X-native TSDF fusion + incremental meshing.

Host-side mirror of the reference's ``FusionMipMapCPU`` interface
(src/fusion/geometryfusion_mipmap_cpu.hpp:111-163) over the C ABI of
``lib/libffgpu.so`` (declared in include/ff_abi.h).  Every method maps to one
ABI call; all compute runs in the HIP library.  There is no CPU fallback:
importing the binding raises if the library has not been built.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
_lib_handle = None


class FFError(RuntimeError):
    pass


class _Params(ctypes.Structure):
    _fields_ = [("scale", ctypes.c_float), ("distance_threshold", ctypes.c_float),
                ("use_color", ctypes.c_int), ("incremental_meshing", ctypes.c_int),
                ("max_leaves", ctypes.c_uint32), ("max_branches", ctypes.c_uint32),
                ("device", ctypes.c_int)]


class _Counts(ctypes.Structure):
    _fields_ = [("n_leaves_used", ctypes.c_uint32), ("n_branches_used", ctypes.c_uint32),
                ("n_queued", ctypes.c_uint32), ("side", ctypes.c_uint32),
                ("offset", ctypes.c_float * 3), ("frames", ctypes.c_uint32),
                ("grows", ctypes.c_uint32)]


class _Frame(ctypes.Structure):
    _fields_ = [("depth", ctypes.c_void_p), ("bgr", ctypes.c_void_p), ("width", ctypes.c_int),
                ("height", ctypes.c_int), ("R", ctypes.c_double * 9), ("t", ctypes.c_double * 3),
                ("K", ctypes.c_double * 4), ("scaling", ctypes.c_float), ("maxcam", ctypes.c_float)]


class _LoopFrame(ctypes.Structure):
    _fields_ = [("depth", ctypes.c_void_p), ("red", ctypes.c_void_p), ("green", ctypes.c_void_p),
                ("blue", ctypes.c_void_p), ("width", ctypes.c_int), ("height", ctypes.c_int),
                ("R", ctypes.c_double * 9), ("t", ctypes.c_double * 3), ("K", ctypes.c_double * 4),
                ("queue", ctypes.c_void_p), ("n_queue", ctypes.c_size_t)]


# name -> (restype, argtypes); exactly the symbols of include/ff_abi.h
_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_DP = ctypes.POINTER(ctypes.c_double)
ABI = {
    "ff_create": (_P, [ctypes.POINTER(_Params)]),
    "ff_destroy": (None, [_P]),
    "ff_last_error": (ctypes.c_char_p, [_P]),
    "ff_add_map": (ctypes.c_int, [_P, _P, _SZ, _P, _SZ, ctypes.c_int, ctypes.c_int, _DP, _DP, _DP,
                                  ctypes.c_float, ctypes.c_float]),
    "ff_add_map_device": (ctypes.c_int, [_P, _P, _P, ctypes.c_int, ctypes.c_int, _DP, _DP, _DP,
                                         ctypes.c_float, ctypes.c_float]),
    "ff_add_map_device_noalloc": (ctypes.c_int, [_P, _P, _P, ctypes.c_int, ctypes.c_int, _DP, _DP, _DP,
                                                 ctypes.c_float, ctypes.c_float]),
    "ff_update_meshes": (ctypes.c_int, [_P]),
    "ff_sync": (ctypes.c_int, [_P]),
    "ff_mesh_full": (ctypes.c_int, [_P]),
    "ff_mesh_sizes": (ctypes.c_int, [_P, ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)]),
    "ff_copy_mesh": (ctypes.c_int, [_P, _P, _P, _P]),
    "ff_cell_count": (ctypes.c_int, [_P, ctypes.POINTER(_SZ)]),
    "ff_cell_update_sizes": (ctypes.c_int, [_P, ctypes.POINTER(_SZ), ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)]),
    "ff_copy_cell_update": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "ff_copy_cells": (ctypes.c_int, [_P, _P, _P, _P]),
    "ff_set_incremental_meshing": (ctypes.c_int, [_P, ctypes.c_int]),
    "ff_set_mesh_sample_tiles": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "ff_set_mesh_pool": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32]),
    "ff_mesh_pool_info": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]),
    "ff_get_counts": (ctypes.c_int, [_P, ctypes.POINTER(_Counts)]),
    "ff_copy_tree": (ctypes.c_int, [_P, _P, _SZ]),
    "ff_copy_leaves": (ctypes.c_int, [_P, _P, _P, _P, _SZ]),
    "ff_copy_bricks": (ctypes.c_int, [_P, _P, _SZ, _P, _P, _P]),
    "ff_copy_queue": (ctypes.c_int, [_P, _P, _SZ]),
    "ff_copy_processed": (ctypes.c_int, [_P, _P, _SZ]),
    "ff_stream": (_P, [_P]),
    "ff_set_graph_capture": (ctypes.c_int, [_P, ctypes.c_int]),
    "ff_set_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "ff_set_timing_every": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "ff_get_timing": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "ff_reset_timing": (ctypes.c_int, [_P]),
    "ff_get_mc_timing": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "ff_tum_read": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _P, _SZ]),
    "ff_png_size": (ctypes.c_int, [ctypes.c_char_p, _P, _P, _P, _P]),
    "ff_png_read_depth": (ctypes.c_int, [ctypes.c_char_p, _P, _SZ]),
    "ff_png_read_bgr": (ctypes.c_int, [ctypes.c_char_p, _P, _SZ]),
    "ff_input_last_error": (ctypes.c_char_p, []),
    "ff_map_description": (ctypes.c_int, [_P, ctypes.c_char_p, _SZ]),
    "ff_save_map": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "ff_load_map": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_float), ctypes.c_uint32]),
    "ff_loop_create": (_P, [_P, ctypes.c_uint32]),
    "ff_loop_destroy": (None, [_P]),
    "ff_loop_last_error": (ctypes.c_char_p, [_P]),
    "ff_loop_reset": (ctypes.c_int, [_P]),
    "ff_loop_delete_frame": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_int, ctypes.c_int, _DP, _DP, _DP,
                                            _P, _SZ, ctypes.POINTER(_SZ)]),
    "ff_loop_delete_frame_device": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_int, ctypes.c_int, _DP, _DP,
                                                   _DP, _P, _SZ, ctypes.POINTER(_SZ)]),
    "ff_loop_delete_frames": (ctypes.c_int, [_P, _P, ctypes.c_int, _P]),
    "ff_loop_mapped": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]),
    "ff_loop_copy_map": (ctypes.c_int, [_P, _P, _SZ]),
    "ff_loop_copy_bricks": (ctypes.c_int, [_P, _P, _SZ, _P, _P, _P]),
    "ff_loop_set_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "ff_loop_get_timing": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "ff_set_learned_weighting": (ctypes.c_int, [_P, _P, _SZ]),
    "ff_copy_confidence": (ctypes.c_int, [_P, _P, _SZ]),
    "ff_trainer_create": (_P, [ctypes.c_int, _SZ, _P, _SZ, ctypes.c_uint64]),
    "ff_trainer_destroy": (None, [_P]),
    "ff_trainer_last_error": (ctypes.c_char_p, [_P]),
    "ff_trainer_collect": (ctypes.c_int, [_P, _P, _P, ctypes.c_int, ctypes.c_int, _DP, _DP, _DP,
                                          ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    "ff_trainer_size": (ctypes.c_int, [_P, ctypes.POINTER(_SZ)]),
    "ff_trainer_copy_samples": (ctypes.c_int, [_P, _P, _P, _SZ]),
    "ff_trainer_set_samples": (ctypes.c_int, [_P, _P, _P, _SZ]),
    "ff_trainer_train": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_uint32, ctypes.c_float,
                                        ctypes.POINTER(ctypes.c_float)]),
    "ff_trainer_grad": (ctypes.c_int, [_P, _P, ctypes.c_uint32, _P, ctypes.POINTER(ctypes.c_float)]),
    "ff_trainer_params": (ctypes.c_int, [_P, _P, _SZ]),
    "ff_add_map_batch": (ctypes.c_int, [_P, _P, ctypes.c_int]),
    "ff_update_meshes_batch": (ctypes.c_int, [_P, ctypes.c_int]),
    "ff_batch_set_config": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "ff_batch_autotune": (ctypes.c_int, [_P, ctypes.c_int]),
    "ff_batch_get_config": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                           ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
}

MLP_NPARAMS = 10177       # FF_MLP_NPARAMS
MLP_RES_SCALE = 0.0025    # FF_MLP_RES_SCALE
MLP_GATE = 0.002          # FF_MLP_GATE


def lib_path():
    return os.path.join(LIB_DIR, "libffgpu.so")


def load_library():
    """Loads libffgpu.so; raises FFError if it has not been built."""
    global _lib_handle
    if _lib_handle is not None:
        return _lib_handle
    path = lib_path()
    if not os.path.exists(path):
        raise FFError(f"{path} is missing: build it with `python -m fastfusion_amd.build` "
                      "(the HIP library is the only implementation; there is no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in ABI.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib_handle = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _dptr(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_DP)


class MeshInterleaved:
    """The reference's MeshInterleaved (src/fusion/mesh_interleaved.hpp:35-69):
    vertices (N, 3) float32, colors (N, 3) uint8 as R, G, B, faces (M, 3) uint32."""

    def __init__(self, vertices, colors, faces):
        self.vertices = vertices
        self.colors = colors
        self.faces = faces

    def writePLY(self, filename, binary=True):
        """MeshInterleaved::writePLY (src/fusion/mesh_interleaved.cpp:139-200):
        writes filename + ".ply", binary little-endian or ASCII, with the
        reference's header lines."""
        v, c, f = self.vertices, self.colors, self.faces
        nv, nf = len(v), len(f)
        hdr = "ply" + ("\nformat binary_little_endian 1.0" if binary else "\nformat ascii 1.0")
        hdr += f"\nelement vertex {nv}\nproperty float32 x\nproperty float32 y\nproperty float32 z"
        if len(c):
            hdr += "\nproperty uchar red\nproperty uchar green\nproperty uchar blue"
        if nf:
            hdr += f"\nelement face {nf}\nproperty list uint8 int32 vertex_indices"
        hdr += "\nend_header"
        with open(filename + ".ply", "wb") as out:
            if binary:
                out.write((hdr + "\n").encode())
                rec = [("p", "<f4", 3)] + ([("c", "u1", 3)] if len(c) else [])
                vv = np.zeros(nv, dtype=rec)
                vv["p"] = v
                if len(c):
                    vv["c"] = c
                out.write(vv.tobytes())
                ff = np.zeros(nf, dtype=[("n", "u1"), ("i", "<u4", 3)])
                ff["n"] = 3
                ff["i"] = f
                out.write(ff.tobytes())
            else:
                lines = [hdr]
                for i in range(nv):
                    line = f"{_g(v[i][0])} {_g(v[i][1])} {_g(v[i][2])}"
                    if len(c):
                        line += f" {int(c[i][0])} {int(c[i][1])} {int(c[i][2])}"
                    lines.append(line)
                for t in f:
                    lines.append(f"3 {int(t[0])} {int(t[1])} {int(t[2])}")
                out.write("\n".join(lines).encode())

    def writeOBJ(self, filename):
        """MeshInterleaved::writeOBJ (src/fusion/mesh_interleaved.cpp:19-137)
        for a mesh without normals and textures: filename + ".obj", vertices
        with their colors as floats in [0, 1], 1-based faces."""
        v, c, f = self.vertices, self.colors, self.faces
        lines = ["", "#Written from Volume"]
        cf = c.astype(np.float32) / np.float32(255.0) if len(c) else None
        for i in range(len(v)):
            line = f"v {_g(v[i][0])} {_g(v[i][1])} {_g(v[i][2])}"
            if cf is not None:
                line += f" {_g(cf[i][0])} {_g(cf[i][1])} {_g(cf[i][2])}"
            lines.append(line)
        for t in f:
            lines.append(f"f {int(t[0]) + 1} {int(t[1]) + 1} {int(t[2]) + 1}")
        with open(filename + ".obj", "wb") as out:
            out.write("\n".join(lines).encode())


def _g(x):
    """A float as C++ streams print it by default (printf %g, 6 digits)."""
    return "%g" % float(x)


class FusionMipMapGPU:
    """Drop-in for the reference's FusionMipMapCPU hot path.

    ``FusionMipMapCPU(offsetX, offsetY, offsetZ, scale, distanceThreshold, n, color)``
    (geometryfusion_mipmap_cpu.cpp:420-640): like the reference, the offset
    and ``n`` arguments are ignored (the volume is set by the first frame).
    """

    def __init__(self, offsetX=0.0, offsetY=0.0, offsetZ=0.0, scale=0.005, distanceThreshold=0.005,
                 n=0, color=True, max_leaves=0, max_branches=0, device=0, incremental=True):
        self._lib = load_library()
        p = _Params(float(scale), float(distanceThreshold), int(bool(color)), int(bool(incremental)),
                    int(max_leaves), int(max_branches), int(device))
        self._ctx = self._lib.ff_create(ctypes.byref(p))
        if not self._ctx:
            raise FFError("ff_create failed (no HIP device or allocation failure)")
        self.color = bool(color)
        # a frame's R (9), t (3) and K (4), staged in one buffer the library
        # reads during the call (far cheaper per frame than converting each
        # argument to a fresh array)
        self._pose = (ctypes.c_double * 16)()
        self._posev = np.frombuffer(self._pose, np.float64)
        a = ctypes.addressof(self._pose)
        self._pR, self._pt, self._pK = (ctypes.cast(a + 8 * o, _DP) for o in (0, 9, 12))

    def _pose_args(self, R, t, K):
        v = self._posev
        v[0:9] = np.ravel(R)
        v[9:12] = np.ravel(t)
        v[12:16] = np.ravel(K)
        return self._pR, self._pt, self._pK

    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.ff_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc < 0:
            raise FFError(self._lib.ff_last_error(self._ctx).decode() or f"error {rc}")
        return rc

    # -- FusionMipMapCPU::addMap (u16 overload), :1603-2034
    def addMap(self, depth, R, t, K, rgb=None, scaling=1.0 / 5000.0, maxcamdistance=10.0):
        depth = np.ascontiguousarray(depth, dtype=np.uint16)
        h, w = depth.shape
        if rgb is not None:
            rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
            if rgb.shape != (h, w, 3):
                raise FFError("rgb must be (h, w, 3) BGR8 like depth")
        pR, pt, pK = self._pose_args(R, t, K)
        rc = self._lib.ff_add_map(self._ctx, _ptr(depth), w * 2, _ptr(rgb), w * 3, w, h, pR, pt, pK, float(scaling),
                                  float(maxcamdistance))
        return self._check(rc)

    def addMapDevice(self, depth_ptr, rgb_ptr, w, h, R, t, K, scaling=1.0 / 5000.0, maxcamdistance=10.0):
        pR, pt, pK = self._pose_args(R, t, K)
        rc = self._lib.ff_add_map_device(self._ctx, depth_ptr, rgb_ptr, int(w), int(h), pR, pt, pK, float(scaling),
                                         float(maxcamdistance))
        return self._check(rc)

    def addMapDeviceNoAlloc(self, depth_ptr, rgb_ptr, w, h, R, t, K, scaling=1.0 / 5000.0, maxcamdistance=10.0):
        """addMapDevice through ff_add_map_device_noalloc: the frame's
        image-sized buffers from the library's per-device scratch array, no
        device allocation (no counterpart in the reference)."""
        pR, pt, pK = self._pose_args(R, t, K)
        rc = self._lib.ff_add_map_device_noalloc(self._ctx, depth_ptr, rgb_ptr, int(w), int(h), pR, pt, pK,
                                                 float(scaling), float(maxcamdistance))
        return self._check(rc)

    # -- updateMeshes, :2884-2954
    def updateMeshes(self):
        """Requests the update of the map as it stands; its device work is
        enqueued by the next call on the context (the next frame's, right
        after that frame's first kernels)."""
        return self._check(self._lib.ff_update_meshes(self._ctx))

    def sync(self):
        """Enqueues a requested update and waits for all of the context's
        device work (no counterpart in the reference)."""
        return self._check(self._lib.ff_sync(self._ctx))

    def meshFull(self):
        """Every cell meshed from scratch now (getMeshRecursive's counterpart,
        geometryfusion_mipmap_cpu.cpp:3658-3674); returns when complete."""
        return self._check(self._lib.ff_mesh_full(self._ctx))

    def setIncrementalMeshing(self, on=True):
        self._check(self._lib.ff_set_incremental_meshing(self._ctx, int(bool(on))))

    def setThreadMeshing(self, on=True):
        """setThreadMeshing (geometryfusion_mipmap_cpu.cpp:1110-1114): accepted
        and ignored; the meshing is a device pass that runs behind the caller."""
        self.threadMeshing = bool(on)

    def setDepthChecks(self, depthchecks):
        """setDepthChecks (geometryfusion_mipmap_cpu.cpp:1116-1120): the
        reference stores the number clamped at 0 and never reads it (its
        backward depth checks in addMap are commented out, :1629-1703), so
        the value is kept and has no effect."""
        self.depthChecks = max(int(depthchecks), 0)

    def setMeshSampleTiles(self, tiles):
        """Capacity of the meshing sample buffer in 64-sample tiles (no
        counterpart in the reference; cells beyond it are meshed in place)."""
        self._check(self._lib.ff_set_mesh_sample_tiles(self._ctx, int(tiles)))

    def setMeshPool(self, vertices, compact_percent=50):
        """Capacity of the mesh pools in vertices and triangles, and the
        share of an arena whose use since the last compaction makes an update
        compact (no counterpart in the reference; the next update meshes
        every cell)."""
        self._check(self._lib.ff_set_mesh_pool(self._ctx, int(vertices), int(compact_percent)))

    def meshPoolInfo(self):
        """(pool capacity, compactions so far)."""
        v, n = ctypes.c_uint32(), ctypes.c_uint32()
        self._check(self._lib.ff_mesh_pool_info(self._ctx, ctypes.byref(v), ctypes.byref(n)))
        return v.value, n.value

    # -- getMeshInterleavedMarchingCubes, :3359-3374
    def getMeshInterleavedMarchingCubes(self):
        nv, nf = _SZ(), _SZ()
        self._check(self._lib.ff_mesh_sizes(self._ctx, ctypes.byref(nv), ctypes.byref(nf)))
        v = np.zeros((nv.value, 3), np.float32)
        c = np.zeros((nv.value, 3), np.uint8)
        f = np.zeros((nf.value, 3), np.uint32)
        self._check(self._lib.ff_copy_mesh(self._ctx, _ptr(v), _ptr(c), _ptr(f)))
        return MeshInterleaved(v, c, f)

    def getMeshCells(self):
        """Per-cell descriptors {type, ox, oy, oz, size} with vertex / face counts."""
        n = _SZ()
        self._check(self._lib.ff_cell_count(self._ctx, ctypes.byref(n)))
        d = np.zeros((n.value, 5), np.int32)
        nv = np.zeros(n.value, np.uint32)
        nf = np.zeros(n.value, np.uint32)
        self._check(self._lib.ff_copy_cells(self._ctx, _ptr(d), _ptr(nv), _ptr(nf)))
        return d, nv, nf

    def getMeshCellsUpdate(self):
        """getMeshCellsUpdate (geometryfusion_mipmap_cpu.cpp:2989-3047): the
        cells re-meshed by the last updateMeshes, each with its own mesh."""
        nc, nv, nf = _SZ(), _SZ(), _SZ()
        self._check(self._lib.ff_cell_update_sizes(self._ctx, ctypes.byref(nc), ctypes.byref(nv), ctypes.byref(nf)))
        u = CellUpdate()
        u.desc = np.zeros((nc.value, 5), np.int32)
        u.n_vertices = np.zeros(nc.value, np.uint32)
        u.n_faces = np.zeros(nc.value, np.uint32)
        u.vertices = np.zeros((nv.value, 3), np.float32)
        u.colors = np.zeros((nv.value, 3), np.uint8)
        u.faces = np.zeros((nf.value, 3), np.uint32)
        self._check(self._lib.ff_copy_cell_update(self._ctx, _ptr(u.desc), _ptr(u.n_vertices), _ptr(u.n_faces),
                                                  _ptr(u.vertices), _ptr(u.colors), _ptr(u.faces)))
        return u

    # -- state export
    def counts(self):
        c = _Counts()
        self._check(self._lib.ff_get_counts(self._ctx, ctypes.byref(c)))
        return {"n_leaves_used": c.n_leaves_used, "n_branches_used": c.n_branches_used,
                "n_queued": c.n_queued, "side": c.side, "offset": np.array(c.offset[:], np.float32),
                "frames": c.frames, "grows": c.grows}

    def tree(self, n_slots=None):
        n = self.counts()["n_branches_used"] if n_slots is None else n_slots
        out = np.zeros(n, np.uint32)
        self._check(self._lib.ff_copy_tree(self._ctx, _ptr(out), n))
        return out

    def leaves(self, n=None):
        n = self.counts()["n_leaves_used"] if n is None else n
        pos = np.zeros((n, 3), np.uint16)
        sc = np.zeros(n, np.uint16)
        par = np.zeros(n, np.uint32)
        self._check(self._lib.ff_copy_leaves(self._ctx, _ptr(pos), _ptr(sc), _ptr(par), n))
        return pos, sc, par

    def bricks(self, leaves):
        leaves = np.ascontiguousarray(leaves, dtype=np.uint32)
        n = len(leaves)
        d = np.zeros((n, 512), np.float32)
        w = np.zeros((n, 512), np.float32)
        c = np.zeros((n, 3, 512), np.uint16)
        self._check(self._lib.ff_copy_bricks(self._ctx, _ptr(leaves), n, _ptr(d), _ptr(w),
                                             _ptr(c) if self.color else None))
        return d, w, c

    def queue(self):
        n = self.counts()["n_queued"]
        out = np.zeros(max(n, 1), np.uint32)
        k = self._check(self._lib.ff_copy_queue(self._ctx, _ptr(out), n))
        return out[:k]

    def processed(self, w, h):
        out = np.zeros((h, w), np.uint8)
        self._check(self._lib.ff_copy_processed(self._ctx, _ptr(out), w * h))
        return out

    def stream(self):
        return self._lib.ff_stream(self._ctx)

    def setGraphCapture(self, on=True):
        """Replay each frame's launches from captured HIP graphs, or launch directly (default)."""
        self._check(self._lib.ff_set_graph_capture(self._ctx, int(bool(on))))

    def set_timing(self, on=True, sdf_only=False, every=1):
        """HIP-event timing of every phase, or (sdf_only) of the SDF update
        alone; the SDF update on one frame in `every`."""
        self._check(self._lib.ff_set_timing(self._ctx, (2 if sdf_only else 1) if on else 0))
        self._check(self._lib.ff_set_timing_every(self._ctx, int(every)))

    def reset_timing(self):
        self._check(self._lib.ff_reset_timing(self._ctx))

    def timing(self):
        """Per phase (allocation, sdf, mesh): ms, launches, algorithmic bytes, units."""
        ms = np.zeros(3, np.float64)
        calls = np.zeros(3, np.uint64)
        by = np.zeros(3, np.float64)
        un = np.zeros(3, np.uint64)
        self._check(self._lib.ff_get_timing(self._ctx, _ptr(ms), _ptr(calls), _ptr(by), _ptr(un)))
        names = ("alloc", "sdf", "mesh")
        return {n: {"ms": float(ms[i]), "calls": int(calls[i]), "bytes": float(by[i]), "units": int(un[i])}
                for i, n in enumerate(names)}

    def mc_timing(self):
        """The marching-cubes kernel alone: summed ms and timed launches
        (events its dispatch records), and with all phases timed the summed
        algorithmic bytes over `byte_calls` updates and the re-meshed cells."""
        ms, by = ctypes.c_double(), ctypes.c_double()
        calls, bcalls, cells = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self._lib.ff_get_mc_timing(self._ctx, ctypes.byref(ms), ctypes.byref(calls), ctypes.byref(by),
                                               ctypes.byref(bcalls), ctypes.byref(cells)))
        return {"ms": ms.value, "calls": calls.value, "bytes": by.value, "byte_calls": bcalls.value,
                "cells": cells.value}

    # -- learned confidence (no counterpart in the reference)
    def getDescription(self):
        """FusionMipMapCPU::getDescription (geometryfusion_mipmap_cpu.cpp:1128-1133):
        offset, scale and side as the map's file stem."""
        buf = ctypes.create_string_buffer(300)
        self._check(self._lib.ff_map_description(self._ctx, buf, 300))
        return buf.value.decode()

    def saveMap(self, path):
        """The map in the reference's raw format (copyCPU2HDD) at `path`."""
        self._check(self._lib.ff_save_map(self._ctx, os.fsencode(path)))

    def loadMap(self, path, offset, n):
        """Restores a map written in the reference's raw format into this
        context (same scale and color setting), with its volume offset and side."""
        off = (ctypes.c_float * 3)(*[float(x) for x in offset])
        self._check(self._lib.ff_load_map(self._ctx, os.fsencode(path), off, int(n)))

    def copyCPU2HDD(self, directory="resource"):
        """copyCPU2HDD (geometryfusion_mipmap_cpu.cpp:1135-1157): writes
        <directory>/<getDescription()>.dat; returns its path."""
        path = os.path.join(directory, self.getDescription() + ".dat")
        self.saveMap(path)
        return path

    def copyHDD2CPU(self, path):
        """copyHDD2CPU (geometryfusion_mipmap_cpu.cpp:1160-1184) for a file
        named by getDescription(): the volume is read from the name."""
        stem = os.path.splitext(os.path.basename(path))[0].split("_")
        if len(stem) != 5:
            raise ValueError(f"{path}: expected <x>_<y>_<z>_<scale>_<n>.dat")
        self.loadMap(path, [float(stem[0]), float(stem[1]), float(stem[2])], int(stem[4]))

    def setLearnedWeighting(self, params):
        """Turns learned per-pixel confidence on (params: MLP_NPARAMS floats)
        or off (None: the plain method)."""
        if params is None:
            self._check(self._lib.ff_set_learned_weighting(self._ctx, None, 0))
            return
        p = np.ascontiguousarray(params, np.float32).ravel()
        self._check(self._lib.ff_set_learned_weighting(self._ctx, _ptr(p), p.size))

    def confidence(self, w, h):
        out = np.zeros(w * h, np.float32)
        self._check(self._lib.ff_copy_confidence(self._ctx, _ptr(out), out.size))
        return out.reshape(h, w)

    @property
    def handle(self):
        return self._ctx


class Trainer:
    """Device trainer of the confidence network (ff_trainer_* in ff_abi.h)."""

    def __init__(self, init_params, capacity=1 << 22, seed=1, device=0):
        self._lib = load_library()
        p = np.ascontiguousarray(init_params, np.float32).ravel()
        if p.size != MLP_NPARAMS:
            raise FFError(f"expected {MLP_NPARAMS} parameters")
        self._h = self._lib.ff_trainer_create(int(device), int(capacity), _ptr(p), p.size, int(seed))
        if not self._h:
            raise FFError("ff_trainer_create failed")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ff_trainer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc < 0:
            raise FFError(f"fastfusion_amd trainer error {rc}: "
                          f"{self._lib.ff_trainer_last_error(self._h).decode(errors='replace')}")
        return rc

    def collect(self, fusion, depth, R, t, K, scaling=1.0 / 5000.0, maxcamdistance=10.0, keep=1.0):
        depth = np.ascontiguousarray(depth, np.uint16)
        h, w = depth.shape
        return self._check(self._lib.ff_trainer_collect(self._h, fusion.handle, _ptr(depth), w, h,
                                                        _dptr(np.asarray(R, np.float64).reshape(9)),
                                                        _dptr(np.asarray(t, np.float64).reshape(3)),
                                                        _dptr(np.asarray(K, np.float64).reshape(4)),
                                                        float(scaling), float(maxcamdistance), float(keep)))

    def size(self):
        n = _SZ()
        self._check(self._lib.ff_trainer_size(self._h, ctypes.byref(n)))
        return n.value

    def samples(self, n=None):
        n = self.size() if n is None else n
        feat = np.zeros((n, 32), np.uint16)
        res = np.zeros(n, np.float32)
        self._check(self._lib.ff_trainer_copy_samples(self._h, _ptr(feat), _ptr(res), n))
        return feat, res

    def set_samples(self, feat_bits, resid):
        feat_bits = np.ascontiguousarray(feat_bits, np.uint16)
        resid = np.ascontiguousarray(resid, np.float32)
        self._check(self._lib.ff_trainer_set_samples(self._h, _ptr(feat_bits), _ptr(resid), len(resid)))

    def train(self, steps, batch=65536, lr=1e-3):
        loss = ctypes.c_float(0.0)
        self._check(self._lib.ff_trainer_train(self._h, int(steps), int(batch), float(lr), ctypes.byref(loss)))
        return loss.value

    def grad(self, index):
        index = np.ascontiguousarray(index, np.uint32)
        g = np.zeros(MLP_NPARAMS, np.float32)
        loss = ctypes.c_float(0.0)
        self._check(self._lib.ff_trainer_grad(self._h, _ptr(index), len(index), _ptr(g), ctypes.byref(loss)))
        return g, loss.value

    def params(self):
        out = np.zeros(MLP_NPARAMS, np.float32)
        self._check(self._lib.ff_trainer_params(self._h, _ptr(out), out.size))
        return out


class CellUpdate:
    """CellUpdate (geometryfusion_mipmap_cpu.hpp:40-59): per re-meshed cell its
    descriptor {type, ox, oy, oz, size} and its own mesh.  Arrays hold the
    cells' meshes concatenated; `cells()` yields (desc, vertices, colors,
    indices) per cell with indices local to the cell, like
    InterleavedMeshCell."""

    def cells(self):
        v0 = np.concatenate([[0], np.cumsum(self.n_vertices)]).astype(np.int64)
        f0 = np.concatenate([[0], np.cumsum(self.n_faces)]).astype(np.int64)
        for i in range(len(self.desc)):
            yield (self.desc[i], self.vertices[v0[i]:v0[i + 1]], self.colors[v0[i]:v0[i + 1]],
                   self.faces[f0[i]:f0[i + 1]])


class Batch:
    """A fixed set of FusionMipMapGPU maps advanced one frame per tick
    (ff_add_map_batch / ff_update_meshes_batch): one SDF launch and one
    marching-cubes launch for all of them."""

    def __init__(self, fusions):
        self._lib = load_library()
        self.fusions = list(fusions)
        self._ctxs = (ctypes.c_void_p * len(self.fusions))(*[f.handle for f in self.fusions])
        self._frames = (_Frame * len(self.fusions))()

    def _check(self, rc):
        if rc < 0:
            msgs = [self._lib.ff_last_error(f.handle).decode(errors="replace") for f in self.fusions]
            raise FFError(f"batch error {rc}: " + "; ".join(m for m in msgs if m))
        return rc

    def addMaps(self, frames):
        """frames: per map (depth_ptr, bgr_ptr, w, h, R, t, K, scaling, maxcam), device pointers."""
        for i, (dp, cp, w, h, R, t, K, sc, mc) in enumerate(frames):
            f = self._frames[i]
            f.depth, f.bgr, f.width, f.height = dp, cp, int(w), int(h)
            f.R[:] = [float(v) for v in np.asarray(R, np.float64).ravel()]
            f.t[:] = [float(v) for v in np.asarray(t, np.float64).ravel()]
            f.K[:] = [float(v) for v in K]
            f.scaling, f.maxcam = float(sc), float(mc)
        return self._check(self._lib.ff_add_map_batch(self._ctxs, self._frames, len(self.fusions)))

    def updateMeshes(self):
        return self._check(self._lib.ff_update_meshes_batch(self._ctxs, len(self.fusions)))

    def set_config(self, threads_per_brick=512, bricks_per_workgroup=1, graphs=False):
        self._check(self._lib.ff_batch_set_config(self._ctxs[0], int(threads_per_brick),
                                                  int(bricks_per_workgroup), int(bool(graphs))))

    def autotune(self, ticks_per_candidate=2):
        self._check(self._lib.ff_batch_autotune(self._ctxs[0], int(ticks_per_candidate)))

    def config(self):
        v = [ctypes.c_int() for _ in range(4)]
        self._check(self._lib.ff_batch_get_config(self._ctxs[0], *[ctypes.byref(x) for x in v]))
        return {"threads_per_brick": v[0].value, "bricks_per_workgroup": v[1].value, "graphs": bool(v[2].value),
                "tuning": bool(v[3].value)}
