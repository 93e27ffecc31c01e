"""ctypes binding of libcholmod_mi355x.so (the C ABI in include/cholmod_mi355x.h).

The library is built in-tree by ``suitesparse_amd.build`` (hipcc, gfx950).  There
is no fallback: if the shared object is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import re

PKG = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG / "libcholmod_mi355x.so"
HEADER = PKG.parent / "include" / "cholmod_mi355x.h"

MXC_OK, MXC_NOT_INSTALLED, MXC_OUT_OF_MEMORY, MXC_TOO_LARGE = 0, -1, -2, -3
MXC_INVALID, MXC_GPU_PROBLEM, MXC_NOT_POSDEF = -4, -5, 1
ORDER_NATURAL, ORDER_GIVEN, ORDER_AMD, ORDER_METIS, ORDER_DEFAULT = 0, 1, 2, 3, -1
FP64, FP32 = 0, 1
HOST, DEVICE = 0, 1

STATUS_NAMES = {0: "CHOLMOD_OK", -1: "CHOLMOD_NOT_INSTALLED", -2: "CHOLMOD_OUT_OF_MEMORY",
                -3: "CHOLMOD_TOO_LARGE", -4: "CHOLMOD_INVALID", -5: "CHOLMOD_GPU_PROBLEM",
                1: "CHOLMOD_NOT_POSDEF"}

_lib = None
vp, i64, f64, i32 = C.c_void_p, C.c_int64, C.c_double, C.c_int
pi64, pf64 = C.POINTER(C.c_int64), C.POINTER(C.c_double)

_PROTOS = {
    "mxc_analyze": (i32, [i64, vp, vp, i32, vp, C.POINTER(vp)]),
    "mxc_symbolic_info": (i32, [vp, vp, vp]),
    "mxc_symbolic_export": (i32, [vp, vp, vp, vp, vp, vp, vp]),
    "mxc_symbolic_import": (i32, [i64, i64, vp, vp, vp, vp, vp, C.POINTER(vp)]),
    "mxc_symbolic_free": (None, [vp]),
    "mxc_amd_order": (i32, [i64, vp, vp, vp]),
    "mxc_amd_order_stats": (i32, [i64, vp, vp, vp, vp]),
    "mxc_metis_order": (i32, [i64, vp, vp, C.c_double, C.c_double, vp]),
    "mxc_symbolic_strategy": (i32, [vp, vp, vp, vp, vp]),
    "mxc_device_init": (i32, [i32]),
    "mxc_device_info": (i32, [vp]),
    "mxc_factor_create": (i32, [vp, vp, vp, i32, C.POINTER(vp)]),
    "mxc_factor_create_in": (i32, [vp, vp, vp, i32, vp, C.POINTER(vp)]),
    "mxc_pool_create": (i32, [C.c_size_t, C.POINTER(vp)]),
    "mxc_pool_info": (i32, [vp, vp]),
    "mxc_pool_free": (i32, [vp]),
    "mxc_memory_info": (i32, [vp, vp]),
    "mxc_factor_stats": (i32, [vp, vp, vp]),
    "mxc_factor_free": (None, [vp]),
    "mxc_factor_info": (i32, [vp, vp, vp]),
    "mxc_factorize": (i32, [vp, vp, i32, f64, pi64]),
    "mxc_factor_get_Lx": (i32, [vp, vp]),
    "mxc_solve": (i32, [vp, vp, i64, vp, i32]),
    "mxc_super_lsolve": (i32, [vp, vp, i64, i32]),
    "mxc_super_ltsolve": (i32, [vp, vp, i64, i32]),
    "mxc_solve_bset": (i32, [vp, i32, vp, vp, i64, vp, vp, vp]),
    "mxc_lsolve_pattern": (i32, [vp, i32, vp, i64, vp, vp]),
    "mxc_residual": (i32, [vp, vp, vp, i64, vp, i32]),
    "mxc_solve_refine": (i32, [vp, vp, i64, vp, i32, f64, i32, vp]),
    "mxc_solve_refine_segments": (i32, [vp, vp, i64, vp, i32, f64, i32, i64, vp, vp]),
    "mxc_autotune": (i32, [vp, i32, vp]),
    "mxc_set_tiles": (i32, [vp, vp]),
    "mxc_get_tiles": (i32, [vp, vp]),
    "mxc_rcond": (i32, [vp, vp]),
    "mxc_rcond_noalloc": (i32, [vp, vp]),
    "mxc_updown": (i32, [vp, i32, i64, vp, vp, vp, f64, vp, vp]),
    "mxc_synchronize": (i32, [vp]),
    "mxc_set_option": (i32, [vp, i32, i64]),
    "mxc_set_timing": (i32, [vp, i32]),
    "mxc_get_timing": (i32, [vp, vp, vp]),
    "mxc_last_error": (C.c_char_p, []),
}


def header_symbols() -> list[str]:
    """Every function the public header declares."""
    txt = HEADER.read_text()
    return sorted(set(re.findall(r"\b(mxc_[a-z0-9_]+)\s*\(", txt)))


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built: run python -m suitesparse_amd.build "
                               "(there is no CPU fallback)")
        # torch ships its own HIP runtime, which shares the process's HSA
        # runtime with this library's: loaded after this library, it finds
        # ROCm's HSA already in place and sees no GPU.  Load torch first.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _PROTOS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return (lib().mxc_last_error() or b"").decode(errors="replace")


class MxcError(RuntimeError):
    def __init__(self, what: str, status: int):
        self.status = status
        msg = f"{what}: {STATUS_NAMES.get(status, status)}"
        err = last_error() if _lib is not None else ""
        if err:
            msg += f" ({err})"
        super().__init__(msg)


def check(what: str, status: int) -> int:
    if status < 0:
        raise MxcError(what, status)
    return status
