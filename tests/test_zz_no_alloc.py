"""mxc_rcond_noalloc (no hipMalloc / hipFree: its per-block partials live in a
static device array per GPU) against mxc_rcond (which allocates them), through
the C ABI, on two calls in a row: the second on a smaller factor, so partials
left in the shared array by the first call must not leak into the second.
The bar is bitwise equality: both run the same kernel over the same grid."""
import ctypes as C

import numpy as np
import pytest

from suitesparse_amd import _lib, cholmod, problems

pytestmark = pytest.mark.gpu


def _call(fn, L):
    out = C.c_double(-1.0)
    st = fn(L._num, C.byref(out))
    assert st == 0, (st, _lib.lib().mxc_last_error())
    return out.value


def _factor(A, precision):
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_AMD, precision=precision)
    L = cholmod.analyze(A, cm)
    assert L is not None and cholmod.factorize(A, L, cm), cm.status
    return L, cm


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_rcond_noalloc_matches_rcond(gpu, precision):
    lib = _lib.lib()
    rng = np.random.default_rng(7)
    for side in (70, 20):          # second call: a smaller input
        A = problems.grid2d_lognormal(side, rng)
        L, cm = _factor(A, precision)
        a = _call(lib.mxc_rcond, L)
        b = _call(lib.mxc_rcond_noalloc, L)
        print(f"{precision} side {side}: rcond {a!r} noalloc {b!r}", flush=True)
        assert 0.0 < a <= 1.0
        assert a == b
        L.free()


def test_rcond_noalloc_trivial_cases(gpu):
    """Not positive definite: 0 (no device work); a 1 x 1 system: 1."""
    lib = _lib.lib()
    A = problems.laplace2d_5pt(12)
    x = A.x.copy()
    x[A.p[:-1][40]] = -1.0
    A2 = problems.SymMatrix(n=A.n, p=A.p, i=A.i, x=x)
    L, cm = _factor(A2, "fp64")
    assert cm.status == cholmod.CHOLMOD_NOT_POSDEF
    assert _call(lib.mxc_rcond_noalloc, L) == _call(lib.mxc_rcond, L) == 0.0
    one = problems.SymMatrix(n=1, p=np.array([0, 1], dtype=np.int64), i=np.array([0], dtype=np.int64),
                             x=np.array([4.0]))
    L1, _ = _factor(one, "fp64")
    assert _call(lib.mxc_rcond_noalloc, L1) == _call(lib.mxc_rcond, L1) == 1.0
