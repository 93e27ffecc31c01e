"""The C-ABI library loads and exports every function include/cholmod_mi355x.h
declares; host-side argument checks answer with CHOLMOD's status codes.  No
device calls."""
import ctypes as C

from suitesparse_amd import _lib


def test_exports_every_declared_symbol():
    L = _lib.lib()
    names = _lib.header_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_status_codes_match_cholmod():
    assert (_lib.MXC_OK, _lib.MXC_NOT_INSTALLED, _lib.MXC_OUT_OF_MEMORY, _lib.MXC_TOO_LARGE,
            _lib.MXC_INVALID, _lib.MXC_GPU_PROBLEM, _lib.MXC_NOT_POSDEF) == (0, -1, -2, -3, -4, -5, 1)


def test_null_arguments_rejected():
    L = _lib.lib()
    h = C.c_void_p()
    assert L.mxc_analyze(-1, None, None, 0, None, C.byref(h)) == _lib.MXC_INVALID
    assert L.mxc_analyze(3, None, None, 0, None, None) == _lib.MXC_INVALID
    assert L.mxc_factorize(None, None, 0, 0.0, None) == _lib.MXC_INVALID
    assert L.mxc_solve(None, None, 1, None, 0) == _lib.MXC_INVALID
    assert L.mxc_factor_create(None, None, None, 0, C.byref(h)) == _lib.MXC_INVALID
