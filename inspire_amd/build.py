"""Builds libpirgpu.so (HIP kernels + the C ABI of include/pirgpu.h) and
libpirgpu_probes.so (the calibration kernels of include/pirgpu_probes.h,
used by measurements only) in-tree.

    python -m inspire_amd.build        # or inspire_amd.build.build()

hipcc cross-compiles for gfx950 without a GPU. Every HIP translation unit is
compiled with -ffp-contract=off (the FFT kernels must not fuse multiply-adds:
their rounding has to match crypto/fft.h operation for operation) and with
packed FP32 disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD_DIR = os.path.join(HERE, "_build")
LIB_PATH = os.path.join(HERE, "libpirgpu.so")
PROBES_LIB_PATH = os.path.join(HERE, "libpirgpu_probes.so")

SOURCES = ["limb_kernels.hip", "fft_kernels.hip", "codec.hip", "server.cpp", "serve.cpp",
           "pack.cpp", "loader.cpp", "proto_loader.cpp", "vpir.cpp", "preprocess_kernels.hip",
           "preprocess.cpp", "noalloc.cpp"]
# Per-source flags. The scan kernels keep up to 224 int32 accumulators per
# lane; LLVM's default AGPR form for their MFMAs makes the register allocator
# shuttle accumulators between AGPRs and VGPRs every k-block, while the VGPR
# form keeps them in place.
EXTRA_FLAGS = {"limb_kernels.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}
# Host-only sources built with g++ (see twiddles.cpp for why).
GXX_SOURCES = ["twiddles.cpp"]
ARCH = os.environ.get("PIRG_OFFLOAD_ARCH", "gfx950")

COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    f"-I{os.path.join(REPO, 'include')}",
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"command failed: {' '.join(cmd)}")
    return proc


def build(verbose: bool = False) -> str:
    """Compiles every source and links libpirgpu.so; returns its path."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    cc = hipcc()
    objs = []
    for src in SOURCES:
        obj = os.path.join(BUILD_DIR, src.replace(".", "_") + ".o")
        # Host .cpp files launch kernels and use the HIP runtime: compile all as HIP.
        lang = ["-x", "hip"]
        cmd = [cc, *COMMON_FLAGS, *EXTRA_FLAGS.get(src, []), *lang, "-c",
               os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    for src in GXX_SOURCES:
        obj = os.path.join(BUILD_DIR, src.replace(".", "_") + ".o")
        cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = LIB_PATH + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, LIB_PATH)
    # The calibration probes: their own library, so the answer path's ABI
    # holds only the reference's interface and its batching / serving forms.
    tmp = PROBES_LIB_PATH + ".tmp"
    cmd = [cc, *COMMON_FLAGS, "-x", "hip", "-shared", os.path.join(CSRC, "probes.hip"),
           "-o", tmp]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(tmp, PROBES_LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
