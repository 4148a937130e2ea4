"""Device buffers for tests that hand device pointers to the library:
hipMalloc / hipMemcpy / hipFree of the HIP runtime liblighter_amd.so is
linked against (libamdhip64.so.7), through ctypes.  (torch ships its own
HIP runtime; its device pointers belong to that runtime, not this one.)"""
import ctypes as C

import numpy as np

H2D, D2H = 1, 2
_hip = None


def hip():
    global _hip
    if _hip is None:
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipGetErrorString.restype = C.c_char_p
        _hip = h
    return _hip


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


class DeviceBuffer:
    """nbytes of device memory on the current GPU."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        _check(hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @classmethod
    def of(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    @property
    def addr(self):
        return self.ptr.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy")

    def download(self, dtype, n):
        out = np.empty(n, dtype)
        assert out.nbytes <= self.nbytes
        _check(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy")
        return out

    def free(self):
        if self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
