"""Builds libshtk_mi355x.so (the HIP kernels and the C ABI) for gfx950 in-tree.

The library is compiled with hipcc directly; no torch extension machinery is
involved, because the product boundary is the plain C ABI of include/shtk.h.
Each source compiles to its own object (in parallel) under ``_build/``, then
one hipcc link step writes the shared library.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
OBJ_DIR = os.path.join(PKG_DIR, "_build")
LIB_NAME = "libshtk_mi355x.so"
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)
SOURCES = ["shtk_plan.cpp", "shtk_legendre.hip", "shtk_fft.hip", "shtk_split.hip", "shtk_ml.hip",
           "shtk_mfma32.hip"]
ARCH = os.environ.get("SHTK_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 instructions are disabled for all device code; the host pass
# warns that it does not know the feature, which is harmless.
HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the ROCm toolchain is required to build libshtk_mi355x.so")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed building {LIB_NAME} (exit {res.returncode}): {' '.join(cmd[-3:])}")


def build(verbose: bool = False) -> str:
    """Compile every source into LIB_PATH (always recompiles)."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    cc = hipcc()
    inc = ["-I", os.path.join(PKG_DIR, "..", "include")]
    objs = [os.path.join(OBJ_DIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    cmds = [[cc, *HIPCC_FLAGS, *inc, "-c", "-o", o, os.path.join(CSRC, s)] for s, o in zip(SOURCES, objs)]
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for fut in [ex.submit(_run, c, verbose) for c in cmds]:
            fut.result()
    tmp = LIB_PATH + ".tmp"
    _run([cc, *HIPCC_FLAGS, "-shared", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
