"""Minimal HIP runtime plumbing over ctypes: device buffers and events.

Used by the benchmark and the device-pointer tests so that inputs sit in HBM
and HIP events are recorded on the very stream the library launches on.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        _hip.hipEventDestroy.argtypes = [C.c_void_p]
        _hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _hip.hipEventSynchronize.argtypes = [C.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        _hip.hipStreamSynchronize.argtypes = [C.c_void_p]
        _hip.hipStreamDestroy.argtypes = [C.c_void_p]
        _hip.hipSetDevice.argtypes = [C.c_int]
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


H2D, D2H = 1, 2


class DeviceArray:
    """float32 / uint32 device buffer with host copies."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(p), max(self.nbytes, 4)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.asfortranarray(a)
        d = cls(a.shape, a.dtype)
        _ok(hip().hipMemcpy(C.c_void_p(d.ptr), a.ctypes.data_as(C.c_void_p), d.nbytes, H2D), "hipMemcpy")
        return d

    def to_host(self):
        a = np.zeros(self.shape, dtype=self.dtype, order="F")
        _ok(hip().hipMemcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes, D2H), "hipMemcpy")
        return a

    def free(self):
        if self.ptr:
            hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = C.c_void_p()
        _ok(hip().hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        _ok(hip().hipStreamSynchronize(C.c_void_p(self.handle)), "hipStreamSynchronize")


class Event:
    def __init__(self):
        e = C.c_void_p()
        _ok(hip().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream: Stream | None):
        _ok(hip().hipEventRecord(C.c_void_p(self.handle), C.c_void_p(stream.handle if stream else None)),
            "hipEventRecord")

    def elapsed_ms(self, end: "Event") -> float:
        _ok(hip().hipEventSynchronize(C.c_void_p(end.handle)), "hipEventSynchronize")
        t = C.c_float()
        _ok(hip().hipEventElapsedTime(C.byref(t), C.c_void_p(self.handle), C.c_void_p(end.handle)),
            "hipEventElapsedTime")
        return float(t.value)


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
