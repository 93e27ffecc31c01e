"""Build the MI355X library (libcholmod_mi355x.so) in-tree with hipcc for gfx950.

    python -m suitesparse_amd.build
"""
from __future__ import annotations

import os
import pathlib
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libcholmod_mi355x.so"
BUILD = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wno-unused-function"]

C_SOURCES = ["symbolic.c", "amd.c"]
HIP_SOURCES = ["mxc_device.hip"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True)


def build(verbose: bool = False) -> pathlib.Path:
    BUILD.mkdir(exist_ok=True)
    objs = []
    for src in C_SOURCES:
        o = BUILD / (src + ".o")
        _run(["gcc", *C_FLAGS, "-c", CSRC / src, "-o", o], verbose)
        objs.append(o)
    for src in HIP_SOURCES:
        o = BUILD / (src + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", CSRC / src, "-o", o], verbose)
        objs.append(o)
    tmp = OUT.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
