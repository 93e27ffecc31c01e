"""solve2: X reuse, and the sparse right-hand side (Bset) against the reference.

cholmod_l_solve2 (CHOLMOD/Cholesky/cholmod_solve.c:130-700) with Bset solves
with the entries b[Bset] only and defines X on Xset.  The device factor stays
supernodal; X on Xset must equal the reference's solve2 result (FP64: within
1e-10 relative to the largest |x| on Xset; FP32 factor: 1e-4), Xset must be
the reference's, and X off Xset must be left as it was.  Without Bset, solve2
writes into a given X of the right shape and matches solve().
"""
import numpy as np
import pytest

from suitesparse_amd import cholmod, problems

pytestmark = pytest.mark.gpu

SYSTEMS = [cholmod.CHOLMOD_A, cholmod.CHOLMOD_LDLt, cholmod.CHOLMOD_LD, cholmod.CHOLMOD_DLt,
           cholmod.CHOLMOD_L, cholmod.CHOLMOD_Lt, cholmod.CHOLMOD_D]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("name,make,ordering", [
    ("lap2d_40_amd", lambda: problems.laplace2d_5pt(40), 1),
    ("st27_12_metis", lambda: problems.stencil3d_27pt(12, seed=4), 2),
], ids=["lap2d_40_amd", "st27_12_metis"])
def test_bset_matches_reference(name, make, ordering, prec, gpu, ref_lib):
    A = make()
    print(f"case {name} n={A.n} {prec}", flush=True)
    R = ref_lib.RefFactor(A, ordering)
    sym = R.symbolic()
    assert R.factorize() == 0
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision=prec)
    L = cholmod.analyze_p(A, sym["Perm"], cm)
    assert L is not None and cholmod.factorize(A, L, cm), cm.status
    rng = np.random.default_rng(2)
    b = rng.uniform(-1, 1, A.n)
    tol = 1e-10 if prec == "fp64" else 1e-4
    for sysc in SYSTEMS:
        for bset in (np.array([A.n // 3]), rng.choice(A.n, 5, replace=False), np.arange(A.n)):
            x_ref, xs_ref = R.solve2_bset(sysc, b, bset)
            X0 = np.full(A.n, 7.0)
            X, xs = cholmod.solve2(sysc, L, b, cm, Bset=bset, X=X0)
            assert cm.status == cholmod.CHOLMOD_OK and X is X0
            assert np.array_equal(xs, xs_ref), (sysc, bset[:5])
            off = np.ones(A.n, bool)
            off[xs] = False
            assert np.all(X[off] == 7.0)
            scale = np.abs(x_ref[xs]).max()
            err = np.abs(X[xs] - x_ref[xs]).max() / scale
            assert err <= tol, (sysc, len(bset), err)
    # sys A with every row in Bset is the dense solve
    X, xs = cholmod.solve2(cholmod.CHOLMOD_A, L, b, cm, Bset=np.arange(A.n))
    Xd = cholmod.solve(cholmod.CHOLMOD_A, L, b, cm)
    assert np.abs(X - Xd).max() <= (1e-12 if prec == "fp64" else 1e-6) * np.abs(Xd).max()


def test_solve2_reuses_x(gpu):
    A = problems.stencil3d_27pt(10, seed=1)
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_AMD)
    L = cholmod.analyze(A, cm)
    assert cholmod.factorize(A, L, cm)
    rng = np.random.default_rng(3)
    X = None
    for k in range(3):
        B = np.asfortranarray(rng.uniform(-1, 1, (A.n, 3)))
        X1, xs = cholmod.solve2(cholmod.CHOLMOD_A, L, B, cm, X=X)
        assert xs is None and cm.status == cholmod.CHOLMOD_OK
        if X is not None:
            assert X1 is X
        X = X1
        Xs = cholmod.solve(cholmod.CHOLMOD_A, L, B, cm)
        # the forward sweep adds the rows below a supernode with atomics: the
        # summation order, not the result, differs between two solves
        assert np.abs(X - Xs).max() <= 1e-12 * np.abs(Xs).max()
    # a wrongly shaped X is replaced, as the reference's ensure_dense does
    X2, _ = cholmod.solve2(cholmod.CHOLMOD_A, L, B[:, :1], cm, X=X)
    assert X2 is not X and X2.shape == (A.n, 1)
    # Bset with two right-hand sides is invalid
    Xb, xsb = cholmod.solve2(cholmod.CHOLMOD_A, L, B, cm, Bset=[0])
    assert Xb is None and cm.status == cholmod.CHOLMOD_INVALID
