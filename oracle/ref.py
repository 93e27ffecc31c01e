"""ctypes wrapper over the reference CHOLMOD compiled into oracle/_ref/libssref.so.

ORACLE / TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py.  The product never imports it.

The library is the reference's own CPU path (CHOLMOD 5.3 supernodal, AMD,
METIS, MKL BLAS) built by oracle/ref.mk from the sources under /root/reference,
plus the glue in oracle/ref_driver.c.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE / "_ref" / "libssref.so"

GIVEN, AMD, METIS, NATURAL, DEFAULT = 0, 1, 2, 3, 4
_lib = None

_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def available() -> bool:
    return LIB_PATH.exists()


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} not built (make -f oracle/ref.mk)")
        os.environ.setdefault("MKL_THREADING_LAYER", "GNU")
        os.environ["MKL_THREADING_LAYER"] = "GNU"
        L = C.CDLL(str(LIB_PATH))
        L.ref_set_threads.argtypes = [C.c_int]
        L.ref_set_threads.restype = C.c_int
        L.ref_analyze.argtypes = [C.c_int64, _i64p, _i64p, C.c_void_p, C.c_int, C.c_void_p,
                                  C.c_int, C.POINTER(C.c_double)]
        L.ref_analyze.restype = C.c_void_p
        L.ref_info.argtypes = [C.c_void_p, _i64p, _f64p]
        L.ref_get_symbolic.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        L.ref_factorize.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_double),
                                    C.POINTER(C.c_int64)]
        L.ref_factorize_p.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.POINTER(C.c_double),
                                      C.POINTER(C.c_int64)]
        L.ref_get_Lx.argtypes = [C.c_void_p, _f64p]
        L.ref_solve.argtypes = [C.c_void_p, _f64p, C.c_int64, _f64p, C.POINTER(C.c_double)]
        L.ref_residual.argtypes = [C.c_void_p, _f64p, _f64p, C.c_int64, _f64p]
        L.ref_solve2_bset.argtypes = [C.c_void_p, C.c_int, _f64p, _i64p, C.c_int64, _f64p, _i64p,
                                      C.POINTER(C.c_int64)]
        L.ref_strategy.argtypes = [C.c_void_p, C.POINTER(C.c_int64), _f64p, _f64p,
                                   C.POINTER(C.c_double)]
        L.ref_order.argtypes = [C.c_int64, _i64p, _i64p, C.c_int, _i64p, _f64p]
        L.ref_order.restype = C.c_int
        L.ref_updown.argtypes = [C.c_void_p, C.c_int, C.c_int64, _i64p, _i64p, _f64p, C.c_double, _f64p,
                                 C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        L.ref_updown.restype = C.c_int
        L.ref_free.argtypes = [C.c_void_p]
        L.ref_free.restype = None
        L.ref_read_sparse.argtypes = [C.c_char_p] + [C.c_void_p] * 4
        L.ref_read_sparse.restype = C.c_int
        _lib = L
    return _lib


def set_threads(n: int) -> int:
    return lib().ref_set_threads(int(n))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def read_sparse(path) -> dict:
    """cholmod_l_read_sparse2 on a file: nrow, ncol, stype, xtype and CSC p, i, x."""
    dims = np.zeros(5, dtype=np.int64)
    b = str(path).encode()
    if lib().ref_read_sparse(b, _ptr(dims), None, None, None) != 0:
        raise RuntimeError(f"reference reader failed on {path}")
    nrow, ncol, nnz, stype, xtype = (int(v) for v in dims)
    p = np.zeros(ncol + 1, dtype=np.int64)
    i = np.zeros(max(nnz, 1), dtype=np.int64)
    x = np.zeros(max(nnz, 1))
    lib().ref_read_sparse(b, _ptr(dims), _ptr(p), _ptr(i), _ptr(x))
    return {"nrow": nrow, "ncol": ncol, "stype": stype, "xtype": xtype, "p": p, "i": i[:nnz],
            "x": x[:nnz]}


def order(n: int, p, i, which: int):
    """The raw cholmod_l_amd (which = AMD) or cholmod_l_metis (which = METIS,
    no postorder) permutation of tril(A), and Common->fl, lnz, anz after it."""
    perm = np.zeros(n, dtype=np.int64)
    stats = np.zeros(3)
    rc = lib().ref_order(int(n), np.ascontiguousarray(p, dtype=np.int64),
                         np.ascontiguousarray(i, dtype=np.int64), int(which), perm, stats)
    if rc != 0:
        raise RuntimeError(f"reference ordering failed ({rc})")
    return perm, stats


class RefFactor:
    """Reference analyze -> factorize -> solve on tril(A) (stype = -1)."""

    def __init__(self, A, ordering: int = AMD, perm=None, single: bool = False):
        L = lib()
        self.A = A
        self.single = single
        t = C.c_double(0)
        up = None if perm is None else np.ascontiguousarray(perm, dtype=np.int64)
        self._h = L.ref_analyze(A.n, np.ascontiguousarray(A.p), np.ascontiguousarray(A.i),
                                _ptr(np.ascontiguousarray(A.x)), int(ordering), _ptr(up),
                                int(single), C.byref(t))
        if not self._h:
            raise RuntimeError("reference analyze failed")
        self.t_analyze = t.value
        info = np.zeros(10, dtype=np.int64)
        dinfo = np.zeros(2)
        L.ref_info(self._h, info, dinfo)
        (self.n, self.nsuper, self.ssize, self.xsize, self.maxcsize, self.maxesize,
         self.is_super, self.ordering, self.status, self.minor) = (int(v) for v in info)
        self.fl, self.lnz = float(dinfo[0]), float(dinfo[1])

    def strategy(self) -> dict:
        """Common->selected, method[0..2].fl / lnz and anz left by the analysis."""
        sel, anz = C.c_int64(0), C.c_double(0)
        fl, lnz = np.zeros(3), np.zeros(3)
        lib().ref_strategy(self._h, C.byref(sel), fl, lnz, C.byref(anz))
        return {"selected": int(sel.value), "method_fl": tuple(fl), "method_lnz": tuple(lnz),
                "anz": float(anz.value)}

    def symbolic(self) -> dict:
        n, ns = self.n, self.nsuper
        out = {k: np.zeros(sz, dtype=np.int64) for k, sz in
               (("Perm", n), ("ColCount", n), ("Super", ns + 1), ("Lpi", ns + 1),
                ("Lpx", ns + 1), ("Ls", self.ssize))}
        rc = lib().ref_get_symbolic(self._h, *(_ptr(out[k]) for k in
                                               ("Perm", "ColCount", "Super", "Lpi", "Lpx", "Ls")))
        if rc != 0:
            raise RuntimeError("factor is not supernodal")
        out.update(n=n, nsuper=ns, ssize=self.ssize, xsize=self.xsize, maxcsize=self.maxcsize,
                   maxesize=self.maxesize, fl=self.fl, lnz=self.lnz)
        return out

    def factorize(self, x=None):
        t = C.c_double(0)
        minor = C.c_int64(0)
        xs = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        st = lib().ref_factorize(self._h, _ptr(xs), C.byref(t), C.byref(minor))
        self.t_factor = t.value
        self.minor = minor.value
        self.status = st
        return st

    def factorize_p(self, beta: float, x=None):
        """cholmod_l_factorize_p: factor A + beta*I (fset NULL)."""
        t = C.c_double(0)
        minor = C.c_int64(0)
        xs = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        st = lib().ref_factorize_p(self._h, _ptr(xs), float(beta), C.byref(t), C.byref(minor))
        self.t_factor = t.value
        self.minor = minor.value
        self.status = st
        return st

    def Lx(self) -> np.ndarray:
        out = np.zeros(self.xsize)
        if lib().ref_get_Lx(self._h, out) != 0:
            raise RuntimeError("no numeric supernodal factor")
        return out

    def solve(self, B: np.ndarray) -> np.ndarray:
        B2 = np.asarray(B, dtype=np.float64)
        if B2.ndim == 1:
            B2 = B2[:, None]
        nrhs = B2.shape[1]
        b = np.ascontiguousarray(B2.T).reshape(-1)      # column-major n x nrhs
        x = np.zeros_like(b)
        t = C.c_double(0)
        rc = lib().ref_solve(self._h, b, nrhs, x, C.byref(t))
        if rc != 0:
            raise RuntimeError(f"reference solve failed ({rc})")
        self.t_solve = t.value
        return x.reshape(nrhs, self.n).T.copy()

    def solve2_bset(self, sys: int, b: np.ndarray, bset) -> tuple[np.ndarray, np.ndarray]:
        """cholmod_l_solve2 with Bset: (X, Xset); X is NaN off Xset.  Leaves the
        factor simplicial (the reference converts it), so call it last."""
        bb = np.ascontiguousarray(np.asarray(b, dtype=np.float64).reshape(-1))
        bs = np.ascontiguousarray(bset, dtype=np.int64).reshape(-1)
        x = np.full(self.n, np.nan)
        xs = np.zeros(self.n, dtype=np.int64)
        ln = C.c_int64(0)
        rc = lib().ref_solve2_bset(self._h, int(sys), bb, bs if bs.size else np.zeros(1, np.int64),
                                   bs.size, x, xs, C.byref(ln))
        if rc != 0:
            raise RuntimeError(f"reference solve2 failed ({rc})")
        return x, xs[:ln.value].copy()

    def residual(self, B: np.ndarray, X: np.ndarray) -> np.ndarray:
        B2 = np.atleast_2d(np.asarray(B, dtype=np.float64).T).T
        X2 = np.atleast_2d(np.asarray(X, dtype=np.float64).T).T
        nrhs = B2.shape[1]
        b = np.ascontiguousarray(B2.T).reshape(-1)
        x = np.ascontiguousarray(X2.T).reshape(-1)
        r = np.zeros_like(b)
        lib().ref_residual(self._h, b, x, nrhs, r)
        return r.reshape(nrhs, self.n).T.copy()

    def updown(self, update: bool, Cp, Ci, Cx, dbound: float = 0.0) -> dict:
        """cholmod_l_updown with C (CSC, rows in the factor's order).  The result
        comes back as LL' values in the supernodal layout of the factor before the
        first call (the reference keeps a simplicial LDL' factor afterwards)."""
        Cp = np.ascontiguousarray(Cp, dtype=np.int64)
        ncol = len(Cp) - 1
        Ci = np.ascontiguousarray(Ci, dtype=np.int64)
        Cx = np.ascontiguousarray(Cx, dtype=np.float64)
        if Ci.size == 0:
            Ci, Cx = np.zeros(1, np.int64), np.zeros(1)
        Lx = np.zeros(self.xsize)
        fl, nb, t = C.c_double(0), C.c_int64(0), C.c_double(0)
        st = lib().ref_updown(self._h, int(bool(update)), ncol, Cp, Ci, Cx, float(dbound), Lx,
                              C.byref(fl), C.byref(nb), C.byref(t))
        if st < 0:
            raise RuntimeError(f"reference updown failed ({st})")
        self.t_updown = t.value
        return {"Lx": Lx, "status": st, "modfl": fl.value, "ndbounds_hit": nb.value, "t": t.value}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.ref_free(h)
            self._h = None
