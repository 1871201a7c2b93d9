"""Device buffers for GPU tests, through the HIP runtime the library itself
is linked against (loaded by name after the library, so the dynamic loader
hands back the same runtime).  Torch is kept out of the tests' processes:
its wheel bundles another libamdhip64.so.7, and once torch is loaded the
library would bind to that runtime instead of its own."""
import ctypes

import numpy as np

import spatial_cell as sc

_hip = None
H2D, D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        sc.load()
        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipSetDevice.argtypes = [ctypes.c_int]
    return _hip


def _ok(e, what):
    if e != 0:
        raise RuntimeError(f"{what} failed: hipError {e}")


class DeviceArray:
    """n elements of dtype on device 0; from a numpy array or filled with
    a byte pattern."""

    def __init__(self, n, dtype, init=None, fill_byte=None):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = max(self.n * self.dtype.itemsize, 16)
        h = hip()
        _ok(h.hipSetDevice(0), "hipSetDevice")
        p = ctypes.c_void_p()
        _ok(h.hipMalloc(ctypes.byref(p), self.nbytes), "hipMalloc")
        self.ptr = p.value
        if init is not None:
            a = np.ascontiguousarray(init, dtype=self.dtype)
            assert a.size * self.dtype.itemsize <= self.nbytes
            if a.size:
                _ok(h.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy")
        elif fill_byte is not None:
            b = np.full(self.nbytes, fill_byte, np.uint8)
            _ok(h.hipMemcpy(self.ptr, b.ctypes.data, self.nbytes, H2D), "hipMemcpy")

    def get(self, n=None):
        n = self.n if n is None else n
        out = np.zeros(n, self.dtype)
        if n:
            _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, n * self.dtype.itemsize, D2H), "hipMemcpy")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
