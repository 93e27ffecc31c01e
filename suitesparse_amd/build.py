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
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]

C_SOURCES = ["symbolic.c", "amd.c", "metis_nd.c"]
HIP_SOURCES = ["mxc_device.hip"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True)


# The CHOLMOD drop-in (libcholmod_mi355x_dropin.so) compiles against the host
# CHOLMOD's headers: cholmod.h fixes the struct layouts and prototypes it
# implements.  SUITESPARSE_ROOT names a SuiteSparse source tree that holds them.
SUITESPARSE_ROOT = pathlib.Path(os.environ.get("SUITESPARSE_ROOT", "/root/reference"))
DROPIN = PKG / "libcholmod_mi355x_dropin.so"


def cholmod_headers():
    inc = [SUITESPARSE_ROOT / "CHOLMOD" / "Include", SUITESPARSE_ROOT / "SuiteSparse_config"]
    return inc if all((d / h).exists() for d, h in zip(inc, ("cholmod.h", "SuiteSparse_config.h"))) else None


def build_dropin(verbose: bool = False):
    """CHOLMOD's cholmod_l_analyze / factorize / solve over this library, for
    linking in front of a host CHOLMOD (skipped when its headers are absent)."""
    inc = cholmod_headers()
    if inc is None:
        return None
    BUILD.mkdir(exist_ok=True)
    o = BUILD / "cholmod_dropin.c.o"
    _run(["gcc", *C_FLAGS, *[f"-I{d}" for d in inc], "-c", CSRC / "cholmod_dropin.c", "-o", o], verbose)
    tmp = DROPIN.with_suffix(".so.tmp")
    _run(["gcc", "-shared", "-fPIC", o, "-o", tmp, f"-L{PKG}", "-lcholmod_mi355x", "-Wl,-rpath,$ORIGIN",
          "-ldl", "-lpthread"], verbose)
    os.replace(tmp, DROPIN)
    return DROPIN


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
    build_dropin(verbose)
    return OUT


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
