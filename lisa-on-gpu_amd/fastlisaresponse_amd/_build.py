"""Build libflr_hip.so, the gfx950 HIP library behind the C ABI in include/flr.h.

hipcc compiles for gfx950 without a GPU present, so this runs on any machine with ROCm. The
library is written next to this file (lib/libflr_hip.so) so that it travels with the tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
SRC_DIR = os.path.join(os.path.dirname(PKG_DIR), "csrc")
REPO_ROOT = os.path.dirname(os.path.dirname(PKG_DIR))
INCLUDE_DIR = os.path.join(REPO_ROOT, "include")
LIB_DIR = os.path.join(PKG_DIR, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libflr_hip.so")

SOURCES = ["flr_kernels.hip", "flr_capi.hip"]
HEADERS = ["flr_internal.hpp", "flr_device.hpp"]

# -ffp-contract=off: the time/delay/index arithmetic must round exactly like the reference CPU
# path; the fast stencils request their fused multiply-adds explicitly.
# -packed-fp32-ops off: no packed FP32 instructions are emitted anywhere.
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found: ROCm is required to build libflr_hip.so")
    return exe


def build(verbose: bool = False) -> str:
    """Compile every HIP source into lib/libflr_hip.so and return its path."""
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = [os.path.join(SRC_DIR, s) for s in SOURCES]
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", INCLUDE_DIR, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{proc.stdout}\n{proc.stderr}")
    if verbose and proc.stderr.strip():
        print(proc.stderr, file=sys.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
