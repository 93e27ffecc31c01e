"""The reference's own test matrices (CHOLMOD Demo/Matrix and Tcov/Matrix,
copied as data into tests/golden/cholmod_matrices) through the GPU path:
symbolic analysis bit-identical to the reference's (AMD), L within 1e-12 of
the reference factor, x within 1e-10 of the reference solution for the demo's
right-hand side b_i = 1 + i/n (cholmod_dl_demo.c:243-249), the demo's
residual test (cholmod_dl_demo.c:769), FP32 + FP64 refinement to 1e-10, and
the not-positive-definite status and minor of the indefinite ones."""
import pathlib

import numpy as np
import pytest

from suitesparse_amd import cholmod, io

FIX = pathlib.Path(__file__).resolve().parent / "golden" / "cholmod_matrices"
SPD = ["bcsstk01.tri", "bcsstk02.tri", "can___24.mtx", "k01up", "r5lo", "diag", "two.tri",
       "2diag.tri"]
INDEFINITE = ["2lo.tri", "2up.tri"]


def _load(name):
    return io.to_symmetric(io.read_sparse(FIX / name), name=name)


@pytest.mark.gpu
@pytest.mark.parametrize("name", SPD)
def test_reference_matrix_fp64_and_fp32ir(gpu, ref_lib, name):
    A = _load(name)
    n = A.n
    R = ref_lib.RefFactor(A, ref_lib.AMD)
    assert R.factorize() == 0
    sym = R.symbolic()
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_AMD)
    L = cholmod.analyze(A, cm)
    for k, mine in (("Perm", L.Perm), ("Super", L.super), ("Lpi", L.pi), ("Lpx", L.px),
                    ("Ls", L.s)):
        assert np.array_equal(mine, sym[k]), k
    assert cholmod.factorize(A, L, cm) and cm.status == cholmod.CHOLMOD_OK
    Lr = R.Lx()
    assert np.max(np.abs(L.x - Lr)) <= 1e-12 * np.max(np.abs(Lr))
    b = (1.0 + np.arange(n) / n)[:, None]
    X = cholmod.solve(cholmod.CHOLMOD_A, L, b, cm)
    Xr = R.solve(b)
    assert np.linalg.norm(X - Xr) <= 1e-10 * np.linalg.norm(Xr)
    Af = A.to_scipy_full()
    anorm = abs(Af).sum(axis=1).max()
    resid = np.abs(Af @ X - b).max() / (anorm * np.abs(X).max() + np.abs(b).max())
    assert resid < 1e-7          # the demo's pass criterion
    assert resid < 1e-13
    cm32 = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_AMD, precision="fp32")
    L32 = cholmod.analyze(A, cm32)
    assert cholmod.factorize(A, L32, cm32) and cm32.status == cholmod.CHOLMOD_OK
    X32, info = cholmod.solve_refine(L32, b, cm32, tol=1e-10, maxiter=20)
    assert info["relres"] <= 1e-10
    # a 1e-10 residual bounds the forward error by cond(A) * 1e-10
    kappa = np.linalg.cond(Af.toarray())
    assert np.linalg.norm(X32 - Xr) <= 10 * kappa * 1e-10 * np.linalg.norm(Xr) + 1e-14


@pytest.mark.gpu
@pytest.mark.parametrize("name", INDEFINITE)
def test_reference_matrix_not_posdef(gpu, ref_lib, name):
    A = _load(name)
    R = ref_lib.RefFactor(A, ref_lib.AMD)
    st = R.factorize()
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_AMD)
    L = cholmod.analyze(A, cm)
    assert cholmod.factorize(A, L, cm)
    assert cm.status == st == cholmod.CHOLMOD_NOT_POSDEF
    assert L.minor == R.minor
