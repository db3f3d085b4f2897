"""Minimal HIP runtime access (device buffers, streams, events) through the same
libamdhip64 the native library links, so device pointers, streams and events are
shared with VHSDecode.Hip.Native without a second runtime in the process.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib

_hip = None

H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        _lib.load()  # the native library first: its libamdhip64 is the one we bind
        path = "/opt/rocm/lib/libamdhip64.so.7"
        _hip = ctypes.CDLL(path if os.path.exists(path) else "libamdhip64.so")
        _hip.hipGetErrorString.restype = ctypes.c_char_p
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                        ctypes.c_void_p]
        _hip.hipMemsetAsync.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p]
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        _hip.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
        _hip.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_uint]
        _hip.hipHostFree.argtypes = [ctypes.c_void_p]
    return _hip


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"HIP error {rc}: {hip().hipGetErrorString(rc).decode()}")


def device_count() -> int:
    n = ctypes.c_int(0)
    rc = hip().hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    _check(hip().hipDeviceSynchronize())


def copy_device(dst: int, src: int, nbytes: int):
    """Device-to-device copy (the runtime's blit kernel); synchronous."""
    _check(hip().hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), int(nbytes), 3))


def download_ptr(ptr: int, dtype, count: int) -> np.ndarray:
    """Copy `count` elements from device address `ptr` (memory another owner holds) to a
    new numpy array; synchronous."""
    out = np.empty(count, dtype=dtype)
    _check(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), out.nbytes, D2H))
    return out


class PinnedArray:
    """Page-locked host memory (hipHostMalloc) viewed as a numpy array; DMA reads and
    writes it directly."""

    def __init__(self, count: int, dtype=np.int16):
        self.dtype = np.dtype(dtype)
        self.nbytes = int(count) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(hip().hipHostMalloc(ctypes.byref(p), max(self.nbytes, 1), 0))
        self.ptr = p.value
        buf = (ctypes.c_char * max(self.nbytes, 1)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=self.dtype, count=int(count))

    def free(self):
        if getattr(self, "ptr", None):
            self.array = None
            _check(hip().hipHostFree(ctypes.c_void_p(self.ptr)))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceBuffer:
    """Raw HBM allocation (hipMalloc)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray, offset: int = 0):
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        _check(hip().hipMemcpy(ctypes.c_void_p(self.ptr + offset), a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D))

    def download(self, dtype, count: int, offset: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        _check(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr + offset), out.nbytes,
                               D2H))
        return out

    def fill(self, byte_value: int):
        """Set every byte (synchronous, so work on other streams sees it)."""
        _check(hip().hipMemsetAsync(ctypes.c_void_p(self.ptr), int(byte_value), self.nbytes, ctypes.c_void_p(0)))
        synchronize()

    def zero(self, stream=None):
        _check(hip().hipMemsetAsync(ctypes.c_void_p(self.ptr), 0, self.nbytes, ctypes.c_void_p(stream or 0)))

    def free(self):
        if self.ptr:
            hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        _check(hip().hipStreamCreateWithFlags(ctypes.byref(s), 1))  # non-blocking, default priority
        self.handle = s.value

    def synchronize(self):
        _check(hip().hipStreamSynchronize(ctypes.c_void_p(self.handle)))

    def __del__(self):
        try:
            if self.handle:
                hip().hipStreamDestroy(ctypes.c_void_p(self.handle))
        except Exception:
            pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        _check(hip().hipEventCreate(ctypes.byref(e)))
        self.handle = e.value

    def record(self, stream=None):
        _check(hip().hipEventRecord(ctypes.c_void_p(self.handle), ctypes.c_void_p(stream or 0)))

    def synchronize(self):
        _check(hip().hipEventSynchronize(ctypes.c_void_p(self.handle)))

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        _check(hip().hipEventElapsedTime(ctypes.byref(ms), ctypes.c_void_p(self.handle), ctypes.c_void_p(end.handle)))
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                hip().hipEventDestroy(ctypes.c_void_p(self.handle))
        except Exception:
            pass
