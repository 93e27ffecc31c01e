"""CHOLMOD-style host interface over the MI355X supernodal path.

Mirrors the reference's int64 API for this path -- names, argument meaning and
error behaviour -- on top of the C ABI (include/cholmod_mi355x.h):

  ``analyze(A, cm)``                cholmod_l_analyze     CHOLMOD/Include/cholmod.h:2714
  ``analyze_p(A, perm, cm)``        cholmod_l_analyze_p   cholmod.h:2734
  ``factorize(A, L, cm)``           cholmod_l_factorize   cholmod.h:2775
  ``factorize_p(A, beta, L, cm)``   cholmod_l_factorize_p cholmod.h:2794
  ``solve(sys, L, B, cm)``          cholmod_l_solve       cholmod.h:2823
  ``updown(update, C, L, cm)``      cholmod_l_updown      cholmod.h:3465
  ``Common``                        cholmod_common: status, fl, lnz, nmethods, method[0].ordering,
                                    selected, method[0..2].fl/lnz, anz, useGPU, print
  ``Factor``                        cholmod_factor: n, minor, Perm, ColCount, super, pi,
                                    px, s, x, nsuper, ssize, xsize, maxcsize, maxesize

As in CHOLMOD, a failing call returns False/None and leaves the reason in
``cm.status`` (the same integer codes, cholmod.h:833-840); a matrix that is not
positive definite still returns True with ``cm.status == CHOLMOD_NOT_POSDEF``
and ``L.minor`` set to the failing column (cholmod_super_numeric.c:37-41).

Matrices are :class:`suitesparse_amd.problems.SymMatrix` (lower triangle, CSC,
int64), i.e. CHOLMOD's ``stype = -1`` storage.  Right-hand sides are numpy
arrays (host) or ``torch`` CUDA tensors (device-resident, no copies).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import threading

import numpy as np

from . import _lib
from ._lib import check, lib

# status codes (cholmod.h:833-840)
CHOLMOD_OK = 0
CHOLMOD_NOT_INSTALLED = -1
CHOLMOD_OUT_OF_MEMORY = -2
CHOLMOD_TOO_LARGE = -3
CHOLMOD_INVALID = -4
CHOLMOD_GPU_PROBLEM = -5
CHOLMOD_NOT_POSDEF = 1
CHOLMOD_DSMALL = 2

# orderings (cholmod.h:448-460)
CHOLMOD_NATURAL = 0
CHOLMOD_GIVEN = 1
CHOLMOD_AMD = 2
CHOLMOD_METIS = 3
CHOLMOD_COLAMD = 5

# solve systems (cholmod.h:2805-2816)
CHOLMOD_A, CHOLMOD_LDLt, CHOLMOD_LD, CHOLMOD_DLt = 0, 1, 2, 3
CHOLMOD_L, CHOLMOD_Lt, CHOLMOD_D, CHOLMOD_P, CHOLMOD_Pt = 4, 5, 6, 7, 8

PRECISIONS = {"fp64": _lib.FP64, "fp32": _lib.FP32}


@dataclasses.dataclass
class Common:
    """The parts of cholmod_common this path reads or writes."""
    status: int = CHOLMOD_OK
    fl: float = -1.0
    lnz: float = -1.0
    nmethods: int = 0                  # 0: default strategy (GIVEN if perm, AMD, then METIS)
    ordering: int = CHOLMOD_AMD        # method[0].ordering, used when nmethods = 1
    # left by analyze: Common->selected, method[0..2].fl / lnz, anz
    selected: int = -1
    method_fl: tuple = (-1.0, -1.0, -1.0)
    method_lnz: tuple = (-1.0, -1.0, -1.0)
    anz: float = -1.0
    useGPU: int = 1
    device: int = 0
    precision: str = "fp64"            # numeric factor: "fp64" or "fp32" (+ FP64 refinement)
    print: int = 0
    # update/downdate: Common->dbound (0: off), and what updown leaves
    dbound: float = 0.0
    modfl: float = 0.0
    ndbounds_hit: int = 0

    def __post_init__(self):
        self._started = False

    def start(self):
        if not self._started:
            st = lib().mxc_device_init(int(self.device))
            if st != 0:
                self.status = st
                return False
            self._started = True
        return True


def _ptr(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(C.c_void_p)


def _is_device(a) -> bool:
    return hasattr(a, "is_cuda") and bool(a.is_cuda)


def _device_ok(t, cm: Common) -> bool:
    """A device tensor the C layer may read or write as doubles: float64 and on
    the Common's device (the C layer takes a raw pointer and a length)."""
    import torch
    return t.dtype == torch.float64 and t.device.index in (None, cm.device) and t.device.type == "cuda"


def _same_pattern(A, L) -> bool:
    """A's pattern is the one L was analyzed with: the device plan places each
    value by its position in that pattern (the reference instead scatters A
    through its own pattern, so a different pattern must not be accepted)."""
    Ap, Ai = L.A_pattern
    p, i = A.p, A.i
    if p is Ap and i is Ai:
        return True
    p = np.asarray(p)
    i = np.asarray(i)
    return (p.shape == Ap.shape and i.shape == Ai.shape and np.array_equal(p, Ap)
            and np.array_equal(i, Ai))


class Factor:
    """Symbolic + (device-resident) numeric supernodal factor."""

    def __init__(self, sym_handle, A, cm: Common):
        self._free_lock = threading.Lock()
        self._sym = sym_handle
        self._num = None
        self._num_precision = None
        self.A_pattern = (np.ascontiguousarray(A.p, dtype=np.int64),
                          np.ascontiguousarray(A.i, dtype=np.int64))
        info = np.zeros(8, dtype=np.int64)
        dinfo = np.zeros(2)
        check("mxc_symbolic_info", lib().mxc_symbolic_info(sym_handle, _ptr(info), _ptr(dinfo)))
        (self.n, self.nsuper, self.ssize, self.xsize, self.maxcsize,
         self.maxesize, self.ordering) = (int(v) for v in info[:7])
        self.fl, self.lnz = float(dinfo[0]), float(dinfo[1])
        n, ns = self.n, self.nsuper
        self.Perm = np.zeros(n, dtype=np.int64)
        self.ColCount = np.zeros(n, dtype=np.int64)
        self.super = np.zeros(ns + 1, dtype=np.int64)
        self.pi = np.zeros(ns + 1, dtype=np.int64)
        self.px = np.zeros(ns + 1, dtype=np.int64)
        self.s = np.zeros(self.ssize, dtype=np.int64)
        check("mxc_symbolic_export", lib().mxc_symbolic_export(
            sym_handle, _ptr(self.Perm), _ptr(self.ColCount), _ptr(self.super), _ptr(self.pi),
            _ptr(self.px), _ptr(self.s)))
        self.is_super = True
        self.is_ll = True
        self.minor = n
        self.is_numeric = False

    # -- numeric object management
    def _numeric(self, cm: Common):
        prec = PRECISIONS[cm.precision]
        if self._num is not None and self._num_precision == prec:
            return self._num
        if self._num is not None:
            lib().mxc_factor_free(self._num)
            self._num = None
        h = C.c_void_p()
        Ap, Ai = self.A_pattern
        st = lib().mxc_factor_create(self._sym, _ptr(Ap), _ptr(Ai), prec, C.byref(h))
        if st != 0:
            cm.status = st
            return None
        self._num = h
        self._num_precision = prec
        return h

    @property
    def x(self) -> np.ndarray:
        """L->x (copied from the device), double."""
        if not self.is_numeric:
            return None
        out = np.zeros(self.xsize)
        check("mxc_factor_get_Lx", lib().mxc_factor_get_Lx(self._num, _ptr(out)))
        return out

    def info(self) -> dict:
        info = np.zeros(10, dtype=np.int64)
        dinfo = np.zeros(6)
        check("mxc_factor_info", lib().mxc_factor_info(self._num, _ptr(info), _ptr(dinfo)))
        keys = ("n", "nsuper", "xsize", "levels", "pairs", "update_tiles", "device_bytes", "precision",
                "fp64_blocks", "escalated_fp64")
        return dict(zip(keys, (int(v) for v in info)))

    def set_timing(self, on, classes=None):
        """Time kernel classes with HIP events (classes: iterable of class indices
        0..9 as listed in timing(); default all)."""
        mask = 0
        if on:
            mask = 1 if classes is None else sum(1 << c for c in classes) | (1 << 30)
        check("mxc_set_timing", lib().mxc_set_timing(self._num, int(mask)))

    def set_pivot_rho(self, rho: float):
        """Pivot-ratio threshold of an FP32 factor: a diagonal block with a pivot at
        or below rho * a_jj is refactored in FP64 (0 = CHOLMOD semantics)."""
        check("mxc_set_option", lib().mxc_set_option(self._num, 2, int(round(rho * 1e9))))

    def set_lookahead(self, on: bool):
        """Overlap each panel's far trailing update with the next panel (default on)."""
        check("mxc_set_option", lib().mxc_set_option(self._num, 3, int(bool(on))))

    def set_split16(self, on: bool):
        """FP32 factors: split-FP16 products on the FP16 MFMA (default on)."""
        check("mxc_set_option", lib().mxc_set_option(self._num, 4, int(bool(on))))

    def set_solve_chain(self, on: bool):
        """Solve sweeps: one chained launch per etree level (default on); off: one
        launch per 64-column step."""
        check("mxc_set_option", lib().mxc_set_option(self._num, 5, int(bool(on))))

    def set_deterministic(self, on: bool):
        """Bitwise-repeatable factorization: descendant updates in rounds, one pair
        per target supernode per launch, applied in descendant order (default off)."""
        check("mxc_set_option", lib().mxc_set_option(self._num, 6, int(bool(on))))

    def set_fold_trsm(self, on: bool):
        """FP64: the trailing launch that factors the next diagonal block also
        applies its TRSM (default off: measured slower); off: one TRSM launch per step."""
        check("mxc_set_option", lib().mxc_set_option(self._num, 7, int(bool(on))))

    def set_t2_grid(self, per_cu: int):
        """FP64: workgroups per CU of the look-ahead's second-stream trailing update,
        each walking its share of the tiles (default 3; 0: one workgroup per tile)."""
        check("mxc_set_option", lib().mxc_set_option(self._num, 8, int(per_cu)))

    def set_small_fused(self, max_per_level):
        """FP64: factor the supernodes of at most 16 columns and 64 rows (those
        with descendant updates) whole, a workgroup each, in one launch per level,
        in every level that has at most max_per_level of them (True: every level,
        False or 0: none)."""
        v = (1 << 62) if max_per_level is True else int(max_per_level)
        check("mxc_set_option", lib().mxc_set_option(self._num, 9, v))

    def set_graphs(self, on: bool):
        """Replay the factorization / solve sweeps from captured hipGraphs (default on)."""
        check("mxc_set_option", lib().mxc_set_option(self._num, 1, int(bool(on))))

    def autotune(self, reps: int = 3) -> list:
        """Pick the tile edge per (rows, cols) bucket by timing; returns the table.
        Invalidates the numeric factor (factorize again afterwards)."""
        t = np.zeros(225, dtype=np.int32)
        check("mxc_autotune", lib().mxc_autotune(self._num, int(reps), _ptr(t)))
        self.is_numeric = False
        return t.tolist()

    def set_tiles(self, table):
        t = np.ascontiguousarray(table, dtype=np.int32)
        check("mxc_set_tiles", lib().mxc_set_tiles(self._num, _ptr(t)))

    def tiles(self) -> list:
        t = np.zeros(225, dtype=np.int32)
        check("mxc_get_tiles", lib().mxc_get_tiles(self._num, _ptr(t)))
        return t.tolist()

    def kernel_flops(self) -> dict:
        """Algorithmic flops of one factorization by kernel class."""
        info = np.zeros(10, dtype=np.int64)
        dinfo = np.zeros(6)
        check("mxc_factor_info", lib().mxc_factor_info(self._num, _ptr(info), _ptr(dinfo)))
        return {"update": float(dinfo[2]), "potrf": float(dinfo[3]), "trsm": float(dinfo[4]),
                "trailing": float(dinfo[5])}

    def timing(self) -> dict:
        t = np.zeros(10)
        c = np.zeros(5, dtype=np.int64)
        check("mxc_get_timing", lib().mxc_get_timing(self._num, _ptr(t), _ptr(c)))
        keys = ("assembly", "updates", "potrf", "trsm", "trailing", "forward", "backward",
                "residual", "factor_total", "solve_total")
        out = {k: float(v) for k, v in zip(keys, t)}
        out["launches"] = dict(zip(("update", "potrf", "trsm", "trailing", "solve"), (int(v) for v in c)))
        return out

    def free(self):
        """Release the device factor and the symbolic factor (once; safe to call
        from any thread and more than once)."""
        lock = getattr(self, "_free_lock", None)
        if lock is None:
            return
        with lock:
            num, self._num = self._num, None
            sym, self._sym = self._sym, None
        if num is not None:
            lib().mxc_factor_free(num)
        if sym is not None:
            lib().mxc_symbolic_free(sym)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# analyze
# ---------------------------------------------------------------------------

def analyze_p(A, perm, cm: Common) -> Factor | None:
    """Order and analyze tril(A) as cholmod_l_analyze_p does (cholmod_analyze.c:384-909).

    nmethods = 0: the default strategy -- the given perm (if any), AMD, then
    METIS unless AMD's prediction is already good; the smallest nnz(L) wins.
    nmethods = 1: cm.ordering alone (GIVEN uses perm, and falls back to AMD
    without one; COLAMD is AMD for a symmetric matrix)."""
    cm.status = CHOLMOD_OK
    h = C.c_void_p()
    p = np.ascontiguousarray(A.p, dtype=np.int64)
    i = np.ascontiguousarray(A.i, dtype=np.int64)
    if perm is not None:
        perm = np.ascontiguousarray(perm, dtype=np.int64)
        if perm.shape != (A.n,):
            cm.status = CHOLMOD_INVALID
            return None
    if cm.nmethods == 0:
        ordering = _lib.ORDER_DEFAULT
    elif cm.nmethods == 1:
        ordering = cm.ordering
        if ordering == CHOLMOD_COLAMD or (ordering == CHOLMOD_GIVEN and perm is None):
            ordering = CHOLMOD_AMD
        if ordering not in (CHOLMOD_NATURAL, CHOLMOD_GIVEN, CHOLMOD_AMD, CHOLMOD_METIS):
            cm.status = CHOLMOD_NOT_INSTALLED
            return None
    else:
        cm.status = CHOLMOD_INVALID
        return None
    st = lib().mxc_analyze(int(A.n), _ptr(p), _ptr(i), int(ordering),
                           _ptr(perm) if ordering in (CHOLMOD_GIVEN, _lib.ORDER_DEFAULT) else None,
                           C.byref(h))
    if st != 0:
        cm.status = st
        return None
    L = Factor(h, A, cm)
    cm.fl, cm.lnz = L.fl, L.lnz
    sel = C.c_int64()
    mfl, mlnz, anz = np.zeros(3), np.zeros(3), C.c_double()
    check("mxc_symbolic_strategy", lib().mxc_symbolic_strategy(h, C.byref(sel), _ptr(mfl), _ptr(mlnz),
                                                              C.byref(anz)))
    cm.selected = int(sel.value)
    cm.method_fl, cm.method_lnz = tuple(float(v) for v in mfl), tuple(float(v) for v in mlnz)
    cm.anz = float(anz.value)
    return L


def analyze(A, cm: Common) -> Factor | None:
    return analyze_p(A, None, cm)


def import_symbolic(A, sym: dict, cm: Common) -> Factor | None:
    """Adopt a supernodal symbolic factor computed elsewhere (cholmod_factor fields)."""
    cm.status = CHOLMOD_OK
    h = C.c_void_p()
    arrs = [np.ascontiguousarray(sym[k], dtype=np.int64) for k in ("Perm", "Super", "Lpi", "Lpx", "Ls")]
    st = lib().mxc_symbolic_import(int(A.n), int(len(arrs[1]) - 1), *[_ptr(a) for a in arrs], C.byref(h))
    if st != 0:
        cm.status = st
        return None
    L = Factor(h, A, cm)
    L.fl = float(sym.get("fl", L.fl))
    return L


# ---------------------------------------------------------------------------
# factorize
# ---------------------------------------------------------------------------

def factorize_p(A, beta, L: Factor, cm: Common, values=None) -> bool:
    """Numeric factorization of A + beta*I with L's pattern.  ``values`` (a numpy
    array or CUDA tensor of tril(A)'s values) overrides A.x."""
    cm.status = CHOLMOD_OK
    if L is None or A.n != L.n or not _same_pattern(A, L):
        cm.status = CHOLMOD_INVALID
        return False
    nnz_a = int(L.A_pattern[0][-1]) if L.n > 0 else 0
    x = A.x if values is None else values
    loc = _lib.DEVICE if _is_device(x) else _lib.HOST
    if loc == _lib.DEVICE:
        if not (_device_ok(x, cm) and x.dim() == 1 and x.is_contiguous() and x.numel() == nnz_a):
            cm.status = CHOLMOD_INVALID
            return False
    else:
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        if x.size != nnz_a:
            cm.status = CHOLMOD_INVALID
            return False
    if not cm.start():
        return False
    h = L._numeric(cm)
    if h is None:
        return False
    b = float(beta[0] if np.ndim(beta) else beta)
    minor = C.c_int64(0)
    st = lib().mxc_factorize(h, _ptr(x), loc, b, C.byref(minor))
    L.minor = int(minor.value)
    if st < 0:
        cm.status = st
        return False
    cm.status = st          # CHOLMOD_OK or CHOLMOD_NOT_POSDEF (a warning: returns True)
    L.is_numeric = True
    return True


def factorize(A, L: Factor, cm: Common, values=None) -> bool:
    return factorize_p(A, 0.0, L, cm, values)


# ---------------------------------------------------------------------------
# solve
# ---------------------------------------------------------------------------

def solve(sys: int, L: Factor, B, cm: Common):
    """X = A\\B (sys CHOLMOD_A / LDLt), L\\B (CHOLMOD_L), L'\\B (CHOLMOD_Lt),
    P*B (CHOLMOD_P) or P'*B (CHOLMOD_Pt).  Returns a new array / tensor."""
    cm.status = CHOLMOD_OK
    if L is None or not L.is_numeric:
        cm.status = CHOLMOD_INVALID
        return None
    dev = _is_device(B)
    if dev:
        import torch
        if not _device_ok(B, cm) or B.dim() not in (1, 2):
            cm.status = CHOLMOD_INVALID
            return None
        Bc = B if B.dim() == 2 else B.reshape(-1, 1)
        Bc = Bc.t().contiguous().t() if not Bc.t().is_contiguous() else Bc   # column-major
        X = torch.empty_like(Bc.t().contiguous()).t()
        if Bc.shape[0] != L.n:
            cm.status = CHOLMOD_INVALID
            return None
    else:
        B2 = np.asarray(B, dtype=np.float64)
        vec = B2.ndim == 1
        B2 = B2.reshape(L.n, -1) if vec else B2
        if B2.shape[0] != L.n:
            cm.status = CHOLMOD_INVALID
            return None
        Bc = np.asfortranarray(B2)
        X = np.zeros_like(Bc, order="F")
    if not _solve_into(sys, L, Bc, X, cm):
        return None
    if not dev and np.ndim(B) == 1:
        return X[:, 0].copy()
    return X


def _solve_into(sys: int, L: Factor, Bc, X, cm: Common) -> bool:
    """The body of solve / solve2: Bc and X column-major n x nrhs, both host
    arrays or both device tensors."""
    dev = _is_device(Bc)
    nrhs = Bc.shape[1]
    bptr, xptr, loc = Bc, X, (_lib.DEVICE if dev else _lib.HOST)
    h = L._num
    if sys in (CHOLMOD_A, CHOLMOD_LDLt):
        st = lib().mxc_solve(h, _ptr(bptr), nrhs, _ptr(xptr), loc)
    elif sys in (CHOLMOD_L, CHOLMOD_LD, CHOLMOD_Lt, CHOLMOD_DLt):
        if dev:
            X.copy_(Bc)
        else:
            X[...] = Bc
        fn = lib().mxc_super_lsolve if sys in (CHOLMOD_L, CHOLMOD_LD) else lib().mxc_super_ltsolve
        st = fn(h, _ptr(xptr), nrhs, loc)
    elif sys == CHOLMOD_D:
        st = 0
        if dev:
            X.copy_(Bc)
        else:
            X[...] = Bc
    elif sys in (CHOLMOD_P, CHOLMOD_Pt):
        st = 0
        perm = L.Perm
        if dev:
            import torch
            pt = torch.as_tensor(perm, device=Bc.device)
            if sys == CHOLMOD_P:
                X.copy_(Bc[pt])
            else:
                X[pt] = Bc
        else:
            if sys == CHOLMOD_P:
                X[...] = Bc[perm]
            else:
                X[perm] = Bc
    else:
        cm.status = CHOLMOD_INVALID
        return False
    if st != 0:
        cm.status = st
        return False
    return True


def solve2(sys: int, L: Factor, B, cm: Common, Bset=None, X=None):
    """cholmod_l_solve2 (cholmod.h:2845; CHOLMOD/Cholesky/cholmod_solve.c:130-700).

    Without Bset: as solve(), except that X is reused when it is an array (or
    device tensor) of the right shape, float64 and column-major -- the
    reference's ensure_dense -- so a sequence of solves allocates no new X.
    The reference's workspaces Y and E live in the device factor and are
    reused by every call.  Returns (X, None).

    With Bset (row indices; B a single column, of which only B[Bset] is read):
    the sparse right-hand-side solve, on host arrays.  Returns (X, Xset): Xset
    holds the rows where X is defined, in the order the reference returns
    them; X's other entries are not written (a reused X keeps its values, a
    new X holds zeros).  Bset is ignored for CHOLMOD_P / CHOLMOD_Pt, as in the
    reference.  On failure returns (None, None) with cm.status set.
    """
    cm.status = CHOLMOD_OK
    if L is None or not L.is_numeric:
        cm.status = CHOLMOD_INVALID
        return None, None
    if sys in (CHOLMOD_P, CHOLMOD_Pt):
        Bset = None
    dev = _is_device(B)
    if dev and (not _device_ok(B, cm) or B.dim() not in (1, 2)):
        cm.status = CHOLMOD_INVALID
        return None, None
    if Bset is None:
        if dev:
            Bc = B if B.dim() == 2 else B.reshape(-1, 1)
            Bc = Bc.t().contiguous().t() if not Bc.t().is_contiguous() else Bc
            reuse = (X is not None and _is_device(X) and tuple(X.shape) == tuple(Bc.shape)
                     and X.dtype == Bc.dtype and X.t().is_contiguous())
            if not reuse:
                import torch
                X = torch.empty_like(Bc.t().contiguous()).t()
        else:
            B2 = np.asarray(B, dtype=np.float64)
            B2 = B2.reshape(L.n, -1) if B2.ndim == 1 else B2
            Bc = np.asfortranarray(B2)
            reuse = (isinstance(X, np.ndarray) and X.shape == Bc.shape and X.dtype == np.float64
                     and X.flags.f_contiguous)
            if not reuse:
                X = np.zeros_like(Bc, order="F")
        if Bc.shape[0] != L.n:
            cm.status = CHOLMOD_INVALID
            return None, None
        if not _solve_into(sys, L, Bc, X, cm):
            return None, None
        return X, None
    if dev:
        cm.status = CHOLMOD_INVALID          # the reference's Bset path is host-only
        return None, None
    b = np.ascontiguousarray(np.asarray(B, dtype=np.float64).reshape(L.n, -1))
    if b.shape[1] != 1:
        cm.status = CHOLMOD_INVALID          # "Bset requires a single right-hand side"
        return None, None
    bs = np.ascontiguousarray(Bset, dtype=np.int64).reshape(-1)
    if not (isinstance(X, np.ndarray) and X.dtype == np.float64 and X.size == L.n
            and X.flags.c_contiguous):
        X = np.zeros(L.n)
    xset = np.zeros(L.n, dtype=np.int64)
    xlen = C.c_int64(0)
    st = lib().mxc_solve_bset(L._num, int(sys), _ptr(b), _ptr(bs), bs.size, _ptr(X), _ptr(xset),
                              C.byref(xlen))
    if st != 0:
        cm.status = st
        return None, None
    return X, xset[:xlen.value].copy()


def lsolve_pattern(L: Factor, Bset, sys: int = CHOLMOD_A) -> np.ndarray:
    """The Xset that solve2(sys, L, b, cm, Bset) returns, from the symbolic
    factor alone (host; cholmod_l_lsolve_pattern, cholmod_rowfac.c:383, on the
    simplicial form of L, permuted back for CHOLMOD_A as cholmod_l_solve2 does)."""
    bs = np.ascontiguousarray(Bset, dtype=np.int64).reshape(-1)
    xset = np.zeros(max(L.n, 1), dtype=np.int64)
    xlen = C.c_int64(0)
    check("mxc_lsolve_pattern", lib().mxc_lsolve_pattern(L._sym, int(sys), _ptr(bs), bs.size, _ptr(xset),
                                                         C.byref(xlen)))
    return xset[:xlen.value].copy()


def solve_refine(L: Factor, B, cm: Common, tol: float = 1e-10, maxiter: int = 10, segments=None):
    """x = A\\b with FP64 iterative refinement through the (FP32 or FP64) factor.
    ``segments`` (row offsets of independent systems) makes the stopping test
    per system.  Returns (X, info) with info = {iterations, relres, relres0}."""
    cm.status = CHOLMOD_OK
    if L is None or not L.is_numeric:
        cm.status = CHOLMOD_INVALID
        return None, None
    dev = _is_device(B)
    if dev:
        import torch
        if not _device_ok(B, cm) or B.dim() not in (1, 2) or B.shape[0] != L.n:
            cm.status = CHOLMOD_INVALID
            return None, None
        Bc = B if B.dim() == 2 else B.reshape(-1, 1)
        Bc = Bc.t().contiguous().t()
        X = torch.empty_like(Bc.t().contiguous()).t()
        loc = _lib.DEVICE
    else:
        B2 = np.asarray(B, dtype=np.float64)
        if B2.shape[0] != L.n or B2.ndim not in (1, 2):
            cm.status = CHOLMOD_INVALID
            return None, None
        Bc = np.asfortranarray(B2.reshape(L.n, -1))
        X = np.zeros_like(Bc, order="F")
        loc = _lib.HOST
    out = np.zeros(4)
    if segments is None:
        st = lib().mxc_solve_refine(L._num, _ptr(Bc), Bc.shape[1], _ptr(X), loc, float(tol),
                                    int(maxiter), _ptr(out))
    else:
        seg = np.ascontiguousarray(segments, dtype=np.int64)
        st = lib().mxc_solve_refine_segments(L._num, _ptr(Bc), Bc.shape[1], _ptr(X), loc, float(tol),
                                             int(maxiter), len(seg) - 1, _ptr(seg), _ptr(out))
    if st != 0:
        cm.status = st
        return None, None
    return X, {"iterations": int(out[0]), "relres": float(out[1]), "relres0": float(out[2]),
               "fp64_fallback": bool(out[3])}


def rcond(L: Factor, cm: Common) -> float:
    """cholmod_l_rcond: (min diag(L) / max diag(L))^2 of the current factor."""
    cm.status = CHOLMOD_OK
    if L is None or not L.is_numeric:
        cm.status = CHOLMOD_INVALID
        return -1.0
    out = C.c_double(0)
    st = lib().mxc_rcond(L._num, C.byref(out))
    if st != 0:
        cm.status = st
        return -1.0
    return out.value


def _csc(Cm, n: int):
    """(p, i, x) of an n-row sparse C: a scipy.sparse matrix, a dense array or a
    (p, i, x) triple of CSC arrays (rows sorted within each column)."""
    if isinstance(Cm, tuple):
        p, i, x = Cm
        return (np.ascontiguousarray(p, dtype=np.int64), np.ascontiguousarray(i, dtype=np.int64),
                np.ascontiguousarray(x, dtype=np.float64))
    if hasattr(Cm, "tocsc"):
        Cs = Cm.tocsc()
        if Cs.shape[0] != n:
            raise ValueError("C and L dimensions do not match")
        Cs.sort_indices()
        return (np.ascontiguousarray(Cs.indptr, dtype=np.int64), np.ascontiguousarray(Cs.indices, dtype=np.int64),
                np.ascontiguousarray(Cs.data, dtype=np.float64))
    D = np.asarray(Cm, dtype=np.float64).reshape(n, -1)
    nz = D != 0
    p = np.concatenate([[0], np.cumsum(nz.sum(axis=0))]).astype(np.int64)
    i = np.nonzero(nz.T)[1].astype(np.int64)
    return p, i, np.ascontiguousarray(D.T[nz.T])


def updown(update: bool, Cm, L: Factor, cm: Common) -> bool:
    """cholmod_l_updown: L L' + C C' (update) or L L' - C C' (downdate) in place
    on the device.  C's rows are in L's order (L.Perm is not applied, as in the
    reference); every entry of C(:, c) must lie in the pattern of L(:, k), k the
    column's first row, else CHOLMOD_INVALID (the modification needs new fill).
    Leaves cm.modfl and cm.ndbounds_hit; a pivot raised to cm.dbound gives
    CHOLMOD_DSMALL.  A downdate that loses positive definiteness returns False
    with CHOLMOD_NOT_POSDEF and L.minor (L then holds no values: the device
    factor is LL', where the reference keeps an indefinite LDL')."""
    cm.status = CHOLMOD_OK
    cm.modfl = 0.0
    cm.ndbounds_hit = 0
    if L is None or not L.is_numeric or L._num is None:
        cm.status = CHOLMOD_INVALID
        return False
    try:
        p, i, x = _csc(Cm, L.n)
    except ValueError:
        cm.status = CHOLMOD_INVALID
        return False
    ncol = len(p) - 1
    if i.size == 0:
        i, x = np.zeros(1, np.int64), np.zeros(1)
    fl = C.c_double(0)
    info = np.zeros(2, dtype=np.int64)
    st = lib().mxc_updown(L._num, int(bool(update)), ncol, _ptr(p), _ptr(i), _ptr(x), float(cm.dbound),
                          C.byref(fl), _ptr(info))
    cm.modfl = fl.value
    cm.ndbounds_hit = int(info[1])
    if st == _lib.MXC_NOT_POSDEF:
        cm.status = CHOLMOD_NOT_POSDEF
        L.minor = int(info[0])
        L.is_numeric = False
        return False
    if st != 0:
        cm.status = st
        return False
    if cm.ndbounds_hit:
        cm.status = CHOLMOD_DSMALL
    return True


def residual(L: Factor, B, X):
    """R = B - A X in FP64 on the device (values of the last factorization)."""
    B2 = np.asfortranarray(np.asarray(B, dtype=np.float64).reshape(L.n, -1))
    X2 = np.asfortranarray(np.asarray(X, dtype=np.float64).reshape(L.n, -1))
    R = np.zeros_like(B2, order="F")
    check("mxc_residual", lib().mxc_residual(L._num, _ptr(B2), _ptr(X2), B2.shape[1], _ptr(R), _lib.HOST))
    return R


def device_info() -> dict:
    info = np.zeros(4, dtype=np.int64)
    check("mxc_device_info", lib().mxc_device_info(_ptr(info)))
    return {"cus": int(info[0]), "hbm_bytes": int(info[1]), "arch": f"gfx{int(info[2])}",
            "clock_khz": int(info[3])}
