"""Device buffers, copies, streams and events through the HIP runtime libqmpaint itself links
(libamdhip64.so.7, loaded once per process), so that the painter, the bench and the tests share
one runtime and nothing else initialises the device."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import painter

H2D, D2H, D2D = 1, 2, 3
_hip = None


def runtime():
    global _hip
    if _hip is None:
        painter.library()  # loads libamdhip64.so.7 as libqmpaint's dependency
        hip = C.CDLL("libamdhip64.so.7")
        vp, sz = C.c_void_p, C.c_size_t
        hip.hipMalloc.argtypes = [C.POINTER(vp), sz]
        hip.hipFree.argtypes = [vp]
        hip.hipMemcpy.argtypes = [vp, vp, sz, C.c_int]
        hip.hipMemset.argtypes = [vp, C.c_int, sz]
        hip.hipDeviceSynchronize.argtypes = []
        hip.hipEventCreate.argtypes = [C.POINTER(vp)]
        hip.hipEventRecord.argtypes = [vp, vp]
        hip.hipEventSynchronize.argtypes = [vp]
        hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), vp, vp]
        hip.hipEventDestroy.argtypes = [vp]
        hip.hipStreamCreate.argtypes = [C.POINTER(vp)]
        hip.hipStreamDestroy.argtypes = [vp]
        hip.hipStreamSynchronize.argtypes = [vp]
        hip.hipGetErrorString.argtypes = [C.c_int]
        hip.hipGetErrorString.restype = C.c_char_p
        _hip = hip
    return _hip


def check(rc):
    if rc != 0:
        raise RuntimeError(f"HIP error {rc}: {runtime().hipGetErrorString(rc).decode()}")


class DeviceBuffer:
    """Device memory of `nbytes`, freed on close or collection."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(runtime().hipMalloc(C.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, array):
        a = np.ascontiguousarray(array)
        buf = cls(a.nbytes)
        check(runtime().hipMemcpy(buf.ptr, a.ctypes.data, a.nbytes, H2D))
        return buf

    def to_array(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        check(runtime().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H))
        return out

    def close(self):
        if self.ptr:
            runtime().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = C.c_void_p()
        check(runtime().hipStreamCreate(C.byref(s)))
        self.handle = s.value

    def synchronize(self):
        check(runtime().hipStreamSynchronize(self.handle))

    def close(self):
        if self.handle:
            runtime().hipStreamDestroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(runtime().hipEventCreate(C.byref(e)))
        self.handle = e.value

    def record(self, stream=None):
        check(runtime().hipEventRecord(self.handle, stream.handle if stream else None))

    def synchronize(self):
        check(runtime().hipEventSynchronize(self.handle))

    def seconds_since(self, start: "Event") -> float:
        ms = C.c_float()
        check(runtime().hipEventElapsedTime(C.byref(ms), start.handle, self.handle))
        return ms.value / 1e3


def synchronize():
    check(runtime().hipDeviceSynchronize())
