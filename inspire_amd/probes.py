"""Calibration kernels for measurement (libpirgpu_probes.so, declared in
include/pirgpu_probes.h): a plain copy that scales HBM byte counters, a
streaming read for the achievable read bandwidth and a register-only int8
MFMA stream for the achievable MFMA rate. Not part of the answer path."""
from __future__ import annotations

import ctypes
import os

from . import PirError

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpirgpu_probes.so")
_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m inspire_amd.build`")
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def copy_probe(nbytes: int, repeats: int = 10) -> float:
    """Average ms of a plain nbytes device copy kernel (counter calibration)."""
    ms = ctypes.c_double()
    rc = lib().pirg_copy_probe(ctypes.c_size_t(nbytes), repeats, ctypes.byref(ms))
    if rc != 0:
        raise PirError(rc, "copy probe failed")
    return ms.value


def read_probe(nbytes: int, repeats: int = 10) -> float:
    """Average ms of one sweep of a plain streaming-read kernel over nbytes."""
    ms = ctypes.c_double()
    rc = lib().pirg_read_probe(ctypes.c_size_t(nbytes), repeats, ctypes.byref(ms))
    if rc != 0:
        raise PirError(rc, "read probe failed")
    return ms.value


def mfma_probe(blocks: int = 1024, iters: int = 4096):
    """(ms, TOP/s) of a register-only int8 MFMA stream (achievable MFMA rate)."""
    ms = ctypes.c_double()
    tops = ctypes.c_double()
    rc = lib().pirg_mfma_probe(blocks, iters, ctypes.byref(ms), ctypes.byref(tops))
    if rc != 0:
        raise PirError(rc, "mfma probe failed")
    return ms.value, tops.value


def mfma_probe_random(blocks: int = 512, iters: int = 16384):
    """(ms, TOP/s) of the register-only int8 MFMA stream on random operands
    (the rate a product over random data can approach)."""
    ms = ctypes.c_double()
    tops = ctypes.c_double()
    rc = lib().pirg_mfma_probe_random(blocks, iters, ctypes.byref(ms), ctypes.byref(tops))
    if rc != 0:
        raise PirError(rc, "mfma probe failed")
    return ms.value, tops.value
