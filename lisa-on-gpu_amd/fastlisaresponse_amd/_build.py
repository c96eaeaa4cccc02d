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

SOURCES = ["flr_kernels.hip", "flr_fused_fp64.hip", "flr_fused_mixed.hip", "flr_fused_q31.hip",
           "flr_fused_c128.hip",
           "flr_waveform.hip", "flr_capi.hip"]
HEADERS = ["flr_internal.hpp", "flr_device.hpp", "flr_tdi_stock.hpp", "flr_paths.hpp",
           "flr_fused.hpp"]
BUILD_DIR = os.path.join(LIB_DIR, "obj")

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


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)


def build(verbose: bool = False) -> str:
    """Compile every HIP source (in parallel, one object each) and link lib/libflr_hip.so;
    returns its path."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        procs.append((src, _run([hipcc(), *compile_flags, "-I", INCLUDE_DIR, "-c", "-o", obj,
                                 os.path.join(SRC_DIR, src)], verbose)))
    errors = []
    for src, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            errors.append(f"{src}: hipcc failed ({p.returncode}):\n{out}\n{err}")
        elif verbose and err.strip():
            print(err, file=sys.stderr)
    if errors:
        raise RuntimeError("\n".join(errors))
    tmp = LIB_PATH + ".tmp"
    p = _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    out, err = p.communicate()
    if p.returncode != 0:
        raise RuntimeError(f"link failed ({p.returncode}):\n{out}\n{err}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
