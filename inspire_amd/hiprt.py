"""Device buffers and events through the HIP runtime libpirgpu.so links
(/opt/rocm/lib/libamdhip64.so.7), for callers of the device-memory entry
points and for timing on a stream. Plumbing only: no kernel runs here.

A second HIP runtime in the process (e.g. one bundled with another framework)
does not share devices or pointers with this one, so device addresses handed
to libpirgpu must come from this runtime."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_PATHS = ["/opt/rocm/lib/libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"]
_rt = None


def rt() -> ctypes.CDLL:
    global _rt
    if _rt is None:
        for p in _PATHS:
            if os.path.exists(p):
                _rt = ctypes.CDLL(p)
                break
        if _rt is None:
            raise ImportError("libamdhip64 not found under /opt/rocm/lib")
    return _rt


def _ok(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


class DeviceBuffer:
    """hipMalloc'd bytes, freed on close()."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        _ok(rt().hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(self.nbytes, 1))),
            "hipMalloc")

    @property
    def addr(self) -> int:
        return int(self.ptr.value)

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _ok(rt().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p),
                           ctypes.c_size_t(a.nbytes), 1), "hipMemcpy H2D")

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _ok(rt().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr,
                           ctypes.c_size_t(out.nbytes), 2), "hipMemcpy D2H")
        return out

    def fill(self, byte: int) -> None:
        _ok(rt().hipMemset(self.ptr, ctypes.c_int(byte), ctypes.c_size_t(self.nbytes)),
            "hipMemset")

    def close(self) -> None:
        if self.ptr.value:
            rt().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def synchronize() -> None:
    _ok(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Event:
    """hipEvent on the default (null) stream unless a stream handle is given."""

    def __init__(self):
        self.ev = ctypes.c_void_p()
        _ok(rt().hipEventCreate(ctypes.byref(self.ev)), "hipEventCreate")

    def record(self, stream: int = 0) -> None:
        _ok(rt().hipEventRecord(self.ev, ctypes.c_void_p(stream)), "hipEventRecord")

    def elapsed_ms(self, later: "Event") -> float:
        _ok(rt().hipEventSynchronize(later.ev), "hipEventSynchronize")
        ms = ctypes.c_float()
        _ok(rt().hipEventElapsedTime(ctypes.byref(ms), self.ev, later.ev), "hipEventElapsedTime")
        return float(ms.value)

    def close(self) -> None:
        if self.ev.value:
            rt().hipEventDestroy(self.ev)
            self.ev = ctypes.c_void_p()
