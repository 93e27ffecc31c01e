"""Chained solve sweeps against the per-step sweeps and the reference solve.

The chained sweeps take one launch per etree level; inside it, each 64-column
block of a supernode waits through a per-supernode counter for the blocks it
depends on (MXC_OPT_SOLVE_CHAIN, on by default).  The per-step sweeps take one
launch per 64-column step of a level (option off).  Both solve with the same
factor, so they differ only in summation order: FP64 within 1e-12 relative of
each other and within 1e-10 of the reference CHOLMOD solve (oracle/_ref); FP32
within 1e-5 relative of each other.  The cases hold supernodes of 1 to 14
blocks, rows below the supernodes, and 1 to 6 right-hand sides (one or two
passes of four).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from suitesparse_amd import cholmod, problems

pytestmark = pytest.mark.gpu


def _dense(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n))
    A = M @ M.T + n * np.eye(n)
    T = sp.csc_matrix(np.tril(A))
    return problems.SymMatrix(n=n, p=T.indptr.astype(np.int64), i=T.indices.astype(np.int64),
                              x=T.data.astype(np.float64), name=f"dense_{n}")


CASES = [
    ("st27_17_metis", lambda: problems.stencil3d_27pt(17, seed=3), 2),
    ("grid3d_16_metis", lambda: problems.grid3d_lognormal(16, np.random.default_rng(7)), 2),
    ("lap2d_60_amd", lambda: problems.laplace2d_5pt(60), 1),
    ("dense_900", lambda: _dense(900, 5), 3),
]


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("nrhs", [1, 2, 4, 6])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("name,make,ordering", CASES, ids=[c[0] for c in CASES])
def test_chain_matches_steps(name, make, ordering, prec, nrhs, gpu, ref_lib):
    A = make()
    print(f"case {name} n={A.n} {prec} nrhs={nrhs}", flush=True)
    R = ref_lib.RefFactor(A, ordering)
    sym = R.symbolic()
    assert R.factorize() == 0
    rng = np.random.default_rng(nrhs)
    B = rng.uniform(-1, 1, (A.n, nrhs))
    Xref = R.solve(B)

    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision=prec)
    L = cholmod.analyze_p(A, sym["Perm"], cm)
    assert L is not None and cholmod.factorize(A, L, cm), cm.status
    L.set_solve_chain(False)
    X0 = cholmod.solve(cholmod.CHOLMOD_A, L, B, cm)
    Y0 = cholmod.solve(cholmod.CHOLMOD_L, L, B, cm)
    Z0 = cholmod.solve(cholmod.CHOLMOD_Lt, L, B, cm)
    L.set_solve_chain(True)
    for rep in range(2):          # the second call replays the captured sweeps
        X1 = cholmod.solve(cholmod.CHOLMOD_A, L, B, cm)
        assert cm.status == cholmod.CHOLMOD_OK
        d = _rel(X1, X0)
        print(f"  rep {rep}: chain vs steps {d:.2e}, vs reference {_rel(X1, Xref):.2e}", flush=True)
        assert d <= (1e-12 if prec == "fp64" else 1e-5)
        if prec == "fp64":
            assert _rel(X1, Xref) <= 1e-10
    Y1 = cholmod.solve(cholmod.CHOLMOD_L, L, B, cm)
    Z1 = cholmod.solve(cholmod.CHOLMOD_Lt, L, B, cm)
    tol = 1e-12 if prec == "fp64" else 1e-5
    assert _rel(Y1, Y0) <= tol and _rel(Z1, Z0) <= tol


def test_chain_refine_fp32(gpu):
    """FP32 factor + FP64 refinement through the chained sweeps reaches 1e-10."""
    A = problems.stencil3d_27pt(20, seed=9)
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_AMD, precision="fp32")
    L = cholmod.analyze(A, cm)
    assert cholmod.factorize(A, L, cm)
    rng = np.random.default_rng(4)
    Xt = rng.uniform(-1, 1, (A.n, 4))
    Af = sp.csc_matrix((A.x, A.i, A.p), shape=(A.n, A.n))
    Af = Af + sp.tril(Af, -1).T
    B = Af @ Xt
    X, info = cholmod.solve_refine(L, B, cm, tol=1e-10)
    print(f"  refine: {info}", flush=True)
    assert info["relres"] <= 1e-10 and not info["fp64_fallback"]
