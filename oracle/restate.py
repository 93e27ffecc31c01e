"""ctypes wrapper over the oracle's CPU restatement (oracle/supernodal_oracle.c).

ORACLE / TEST INFRASTRUCTURE ONLY: imported by tests/ and the smoke check.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

LIB_PATH = pathlib.Path(__file__).resolve().parent / "_build" / "liboracle.so"
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} not built (make -C oracle)")
        L = C.CDLL(str(LIB_PATH))
        vp, i64 = C.c_void_p, C.c_int64
        L.oracle_super_numeric.argtypes = [i64, i64] + [vp] * 7 + [C.c_double, vp, vp]
        L.oracle_super_numeric.restype = C.c_int
        for fn in (L.oracle_super_lsolve, L.oracle_super_ltsolve):
            fn.argtypes = [i64, i64, vp, vp, vp, vp, vp, vp, i64]
            fn.restype = None
        L.oracle_solve.argtypes = [i64, i64, vp, vp, vp, vp, vp, vp, vp, vp, i64]
        L.oracle_solve.restype = None
        L.oracle_permute_lower.argtypes = [i64, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_permute_lower.restype = C.c_int
        L.oracle_updown.argtypes = [i64, i64, vp, vp, vp, vp, vp, C.c_int, i64, vp, vp, vp, C.c_double, vp, vp]
        L.oracle_updown.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def permute_lower(A, perm):
    """S = tril(P A P') (CSC, sorted rows) from tril(A)."""
    n = A.n
    nnz = int(A.p[-1])
    Sp = np.zeros(n + 1, dtype=np.int64)
    Si = np.zeros(max(nnz, 1), dtype=np.int64)
    Sx = np.zeros(max(nnz, 1))
    lib().oracle_permute_lower(n, _p(_i64(A.p)), _p(_i64(A.i)), _p(np.ascontiguousarray(A.x)),
                               _p(_i64(perm)), _p(Sp), _p(Si), _p(Sx))
    return Sp, Si[:nnz], Sx[:nnz]


def super_numeric(A, sym: dict, beta: float = 0.0):
    """Left-looking supernodal LL' of P A P' + beta I.  Returns (Lx, status, minor)."""
    Sp, Si, Sx = permute_lower(A, sym["Perm"])
    Lx = np.zeros(int(sym["Lpx"][-1]) if len(sym["Lpx"]) else 1)
    minor = C.c_int64(0)
    ns = len(sym["Super"]) - 1
    keep = [_i64(sym[k]) for k in ("Super", "Lpi", "Lpx", "Ls")]
    st = lib().oracle_super_numeric(A.n, ns, *[_p(a) for a in keep], _p(Sp), _p(Si),
                                    _p(np.ascontiguousarray(Sx)), float(beta), _p(Lx),
                                    C.byref(minor))
    return Lx, int(st), int(minor.value)


def solve(sym: dict, Lx: np.ndarray, B: np.ndarray) -> np.ndarray:
    """X = A \\ B through the supernodal factor (B n x nrhs)."""
    B2 = np.asarray(B, dtype=np.float64)
    vec = B2.ndim == 1
    B2 = B2.reshape(len(sym["Perm"]), -1)
    n, nrhs = B2.shape
    b = np.ascontiguousarray(B2.T).reshape(-1)
    x = np.zeros_like(b)
    ns = len(sym["Super"]) - 1
    keep = [_i64(sym[k]) for k in ("Super", "Lpi", "Lpx", "Ls")]
    lib().oracle_solve(n, ns, *[_p(a) for a in keep], _p(np.ascontiguousarray(Lx)),
                       _p(_i64(sym["Perm"])), _p(b), _p(x), nrhs)
    X = x.reshape(nrhs, n).T
    return X[:, 0].copy() if vec else X.copy()


def lsolve(sym, Lx, X):
    """In-place L \\ X in the permuted order (X n x nrhs, returned as a new array)."""
    X2 = np.asarray(X, dtype=np.float64).reshape(len(sym["Perm"]), -1)
    n, nrhs = X2.shape
    x = np.ascontiguousarray(X2.T).reshape(-1).copy()
    ns = len(sym["Super"]) - 1
    keep = [_i64(sym[k]) for k in ("Super", "Lpi", "Lpx", "Ls")]
    lib().oracle_super_lsolve(n, ns, *[_p(a) for a in keep], _p(np.ascontiguousarray(Lx)), _p(x), nrhs)
    return x.reshape(nrhs, n).T.copy()


def ltsolve(sym, Lx, X):
    X2 = np.asarray(X, dtype=np.float64).reshape(len(sym["Perm"]), -1)
    n, nrhs = X2.shape
    x = np.ascontiguousarray(X2.T).reshape(-1).copy()
    ns = len(sym["Super"]) - 1
    keep = [_i64(sym[k]) for k in ("Super", "Lpi", "Lpx", "Ls")]
    lib().oracle_super_ltsolve(n, ns, *[_p(a) for a in keep], _p(np.ascontiguousarray(Lx)), _p(x), nrhs)
    return x.reshape(nrhs, n).T.copy()


def dense_L(sym: dict, Lx: np.ndarray) -> np.ndarray:
    """Densify the supernodal factor (lower triangle, n x n) for small checks."""
    n = len(sym["Perm"])
    L = np.zeros((n, n))
    Super, Lpi, Lpx, Ls = sym["Super"], sym["Lpi"], sym["Lpx"], sym["Ls"]
    for s in range(len(Super) - 1):
        k1, k2 = int(Super[s]), int(Super[s + 1])
        rows = Ls[Lpi[s]:Lpi[s + 1]]
        nsrow = len(rows)
        blk = Lx[Lpx[s]:Lpx[s] + nsrow * (k2 - k1)].reshape(k2 - k1, nsrow).T
        for c in range(k2 - k1):
            r = rows[c:]
            L[r, k1 + c] = blk[c:, c]
    return L


def updown(sym: dict, Lx: np.ndarray, update: bool, Cp, Ci, Cx, dbound: float = 0.0) -> dict:
    """Rank-k update (update=True) or downdate of the supernodal LL' factor Lx by
    C (CSC, rows in the factor's order); returns the new values, modfl, the first
    column whose new pivot is not positive (n if none) and the bounded pivots."""
    out = np.ascontiguousarray(Lx, dtype=np.float64).copy()
    Cp = _i64(Cp)
    Ci = _i64(Ci) if len(Ci) else np.zeros(1, np.int64)
    Cx = np.ascontiguousarray(Cx, dtype=np.float64) if len(Cx) else np.zeros(1)
    ns = len(sym["Super"]) - 1
    keep = [_i64(sym[k]) for k in ("Super", "Lpi", "Lpx", "Ls")]
    fl = C.c_double(0)
    info = np.zeros(2, dtype=np.int64)
    rc = lib().oracle_updown(len(sym["Perm"]), ns, *[_p(a) for a in keep], _p(out), int(bool(update)),
                             len(Cp) - 1, _p(Cp), _p(Ci), _p(Cx), float(dbound), C.byref(fl), _p(info))
    return {"Lx": out, "rc": int(rc), "modfl": fl.value, "minor": int(info[0]), "ndbounds_hit": int(info[1])}
