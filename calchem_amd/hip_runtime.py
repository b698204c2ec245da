"""Minimal ctypes access to the HIP runtime (device buffers, streams, events).

Used by the benchmark and the GPU tests to keep inputs resident in HBM and to
time work with HIP events on the stream the kernels run on.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_hip = None


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipGetErrorString.restype = C.c_char_p
        _hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [C.c_void_p]
        _hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _hip.hipEventSynchronize.argtypes = [C.c_void_p]
        _hip.hipEventDestroy.argtypes = [C.c_void_p]
        _hip.hipStreamDestroy.argtypes = [C.c_void_p]
    return _hip


def _ok(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()}")


H2D, D2H = 1, 2


class DeviceArray:
    def __init__(self, nbytes: int):
        self.ptr = C.c_void_p()
        self.nbytes = int(nbytes)
        _ok(hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 16)), "hipMalloc")

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        _ok(hip().hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return d

    def copy_from_numpy(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError(f"{a.nbytes} bytes do not fit a {self.nbytes}-byte device array")
        _ok(hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")

    def to_numpy(self, shape, dtype=np.float64) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    @property
    def address(self) -> int:
        return self.ptr.value

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.handle = C.c_void_p()
        _ok(hip().hipStreamCreate(C.byref(self.handle)), "hipStreamCreate")

    def synchronize(self):
        _ok(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    @property
    def address(self) -> int:
        return self.handle.value or 0


class Event:
    def __init__(self):
        self.handle = C.c_void_p()
        _ok(hip().hipEventCreate(C.byref(self.handle)), "hipEventCreate")

    def record(self, stream: Stream):
        _ok(hip().hipEventRecord(self.handle, stream.handle), "hipEventRecord")

    def synchronize(self):
        _ok(hip().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        _ok(hip().hipEventElapsedTime(C.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return float(ms.value)


def device_synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def device_name() -> str:
    buf = C.create_string_buffer(256)
    try:
        hip().hipDeviceGetName(buf, 256, 0)
        return buf.value.decode(errors="replace")
    except Exception:
        return "unknown"
