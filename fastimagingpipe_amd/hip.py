"""Minimal HIP runtime plumbing over ctypes: device buffers, streams and events.

Used by bench.py and the GPU tests to hold inputs resident in HBM and time the
library's launches on a stream of their own.  Nothing here computes.
"""
from __future__ import annotations

import ctypes
import ctypes.util

import numpy as np

_CANDIDATES = ["libamdhip64.so", "libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"]


def _load():
    for name in _CANDIDATES:
        try:
            return ctypes.CDLL(name)
        except OSError:
            continue
    path = ctypes.util.find_library("amdhip64")
    if path:
        return ctypes.CDLL(path)
    raise OSError("libamdhip64 not found")


_hip = _load()
_vp = ctypes.c_void_p
_hip.hipMalloc.argtypes = [ctypes.POINTER(_vp), ctypes.c_size_t]
_hip.hipFree.argtypes = [_vp]
_hip.hipMemcpy.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int]
_hip.hipMemset.argtypes = [_vp, ctypes.c_int, ctypes.c_size_t]
_hip.hipDeviceSynchronize.argtypes = []
_hip.hipStreamCreate.argtypes = [ctypes.POINTER(_vp)]
_hip.hipStreamDestroy.argtypes = [_vp]
_hip.hipStreamSynchronize.argtypes = [_vp]
_hip.hipEventCreate.argtypes = [ctypes.POINTER(_vp)]
_hip.hipEventDestroy.argtypes = [_vp]
_hip.hipEventRecord.argtypes = [_vp, _vp]
_hip.hipEventSynchronize.argtypes = [_vp]
_hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), _vp, _vp]
_hip.hipGetErrorString.argtypes = [ctypes.c_int]
_hip.hipGetErrorString.restype = ctypes.c_char_p
_hip.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]

H2D, D2H, D2D = 1, 2, 3


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: {_hip.hipGetErrorString(rc).decode()}")


def synchronize() -> None:
    _check(_hip.hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info():
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    _check(_hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return free.value, total.value


class DeviceArray:
    """A device allocation holding a copy of a numpy array (or zeros)."""

    def __init__(self, host: np.ndarray | None = None, nbytes: int | None = None):
        if host is not None:
            host = np.ascontiguousarray(host)
            nbytes = host.nbytes
        self.nbytes = int(nbytes or 0)
        self.ptr = _vp()
        _check(_hip.hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")
        if host is not None:
            _check(_hip.hipMemcpy(self.ptr, host.ctypes.data_as(_vp), self.nbytes, H2D), "hipMemcpy H2D")

    @property
    def addr(self) -> int:
        return int(self.ptr.value or 0)

    def upload(self, host: np.ndarray) -> None:
        host = np.ascontiguousarray(host)
        if host.nbytes > self.nbytes:
            raise ValueError("upload larger than the allocation")
        _check(_hip.hipMemcpy(self.ptr, host.ctypes.data_as(_vp), host.nbytes, H2D), "hipMemcpy H2D")

    def to_host(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        _check(_hip.hipMemcpy(out.ctypes.data_as(_vp), self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self) -> None:
        if self.ptr:
            _hip.hipFree(self.ptr)
            self.ptr = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.handle = _vp()
        _check(_hip.hipStreamCreate(ctypes.byref(self.handle)), "hipStreamCreate")

    @property
    def addr(self) -> int:
        return int(self.handle.value or 0)

    def synchronize(self) -> None:
        _check(_hip.hipStreamSynchronize(self.handle), "hipStreamSynchronize")


class Event:
    def __init__(self):
        self.handle = _vp()
        _check(_hip.hipEventCreate(ctypes.byref(self.handle)), "hipEventCreate")

    def record(self, stream: Stream | None = None) -> None:
        _check(_hip.hipEventRecord(self.handle, stream.handle if stream else None), "hipEventRecord")

    def synchronize(self) -> None:
        _check(_hip.hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        _check(_hip.hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return float(ms.value)

