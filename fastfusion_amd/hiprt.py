"""Minimal ctypes access to the HIP runtime (device buffers, copies, sync).

Used to keep benchmark inputs resident in HBM without importing torch.
"""
import ctypes
import ctypes.util

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is not None:
        return _hip
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            _hip = ctypes.CDLL(name)
            break
        except OSError:
            continue
    if _hip is None:
        raise RuntimeError("libamdhip64.so not found")
    P = ctypes.c_void_p
    _hip.hipMalloc.argtypes = [ctypes.POINTER(P), ctypes.c_size_t]
    _hip.hipFree.argtypes = [P]
    _hip.hipMemcpy.argtypes = [P, P, ctypes.c_size_t, ctypes.c_int]
    _hip.hipDeviceSynchronize.argtypes = []
    _hip.hipSetDevice.argtypes = [ctypes.c_int]
    _hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
    _hip.hipStreamSynchronize.argtypes = [P]
    return _hip


H2D, D2H, D2D = 1, 2, 3


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


def device_count():
    n = ctypes.c_int(0)
    rc = hip().hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


class DeviceBuffer:
    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        self.nbytes = int(nbytes)
        check(hip().hipMalloc(ctypes.byref(self.ptr), self.nbytes), "hipMalloc")

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        check(hip().hipMemcpy(b.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D), "hipMemcpy H2D")
        return b

    def addr(self, offset=0):
        return (self.ptr.value or 0) + int(offset)

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def copy_to_device(addr, a):
    """Copies the host array a to device address addr (synchronous)."""
    a = np.ascontiguousarray(a)
    check(hip().hipMemcpy(ctypes.c_void_p(addr), a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D), "hipMemcpy H2D")


def synchronize():
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
