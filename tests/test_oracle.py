"""Pins of the oracle's CPU restatement (oracle/supernodal_oracle.c).

The restatement is the checker for the GPU path, so it is itself checked
against: the reference CHOLMOD factor and solution (oracle/_ref) and the
committed fixture made from it; the uniqueness of the Cholesky factor of an
SPD matrix (dense LAPACK on P A P'); closed forms (1D Laplacian, diagonal,
block diagonal); residuals; and the not-positive-definite contract (status,
L->minor, zeroed tail) of the reference.  Planted mistakes in the oracle's
output must fail these checks.
"""
import pathlib

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import restate
from suitesparse_amd import problems

GOLD = pathlib.Path(__file__).resolve().parent / "golden"


def _mat(dense_lower, name="m"):
    T = sp.csc_matrix(np.tril(dense_lower))
    T.sort_indices()
    return problems.SymMatrix(n=T.shape[0], p=T.indptr.astype(np.int64), i=T.indices.astype(np.int64),
                              x=T.data.astype(np.float64), name=name)


def _natural_sym(A):
    """Symbolic factor of A in the natural order (product analysis, CPU only)."""
    from suitesparse_amd import cholmod
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN)
    L = cholmod.analyze_p(A, np.arange(A.n), cm)
    assert L is not None
    return {"Perm": L.Perm, "Super": L.super, "Lpi": L.pi, "Lpx": L.px, "Ls": L.s}


def test_fixture_reference_factor():
    """Committed reference factor/solution (grid2d 12x12, AMD) reproduced to rounding."""
    z = np.load(GOLD / "small_grid2d_12.npz")
    A = problems.SymMatrix(n=len(z["p"]) - 1, p=z["p"], i=z["i"], x=z["x"])
    sym = {k: z[k] for k in ("Perm", "Super", "Lpi", "Lpx", "Ls")}
    Lx, st, minor = restate.super_numeric(A, sym)
    assert st == 0 and minor == A.n
    assert np.abs(Lx - z["Lx"]).max() <= 1e-13 * np.abs(z["Lx"]).max()
    X = restate.solve(sym, Lx, z["B"])
    assert np.linalg.norm(X - z["X"]) <= 1e-13 * np.linalg.norm(z["X"])


@pytest.mark.parametrize("ordering", [1, 2])
def test_against_reference(ref_lib, ordering):
    for A in (problems.laplace2d_5pt(30), problems.grid3d_lognormal(9, np.random.default_rng(4)),
              problems.stencil3d_27pt(7, seed=9)):
        R = ref_lib.RefFactor(A, ordering)
        sym = R.symbolic()
        assert R.factorize() == 0
        Lref = R.Lx()
        Lx, st, minor = restate.super_numeric(A, sym)
        assert st == 0 and minor == A.n
        assert np.abs(Lx - Lref).max() <= 1e-13 * np.abs(Lref).max()
        B = np.random.default_rng(3).uniform(-1, 1, (A.n, 3))
        X = restate.solve(sym, Lx, B)
        Xr = R.solve(B)
        assert np.linalg.norm(X - Xr) <= 1e-12 * np.linalg.norm(Xr)


def test_uniqueness_dense_lapack():
    """The densified supernodal L equals numpy's (LAPACK) Cholesky of P A P'."""
    for seed, side in ((0, 6), (1, 9)):
        A = problems.grid2d_lognormal(side, np.random.default_rng(seed))
        perm = np.random.default_rng(seed).permutation(A.n)
        from suitesparse_amd import cholmod
        cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN)
        L = cholmod.analyze_p(A, perm, cm)
        sym = {"Perm": L.Perm, "Super": L.super, "Lpi": L.pi, "Lpx": L.px, "Ls": L.s}
        Lx, st, _ = restate.super_numeric(A, sym)
        Ld = restate.dense_L(sym, Lx)
        Af = A.to_scipy_full().toarray()
        P = L.Perm
        Lnp = np.linalg.cholesky(Af[np.ix_(P, P)])
        assert np.abs(Ld - Lnp).max() <= 1e-12 * np.abs(Lnp).max()


def test_closed_form_laplace1d():
    n = 40
    T = 2 * np.eye(n) - np.eye(n, k=-1) - np.eye(n, k=1)
    A = _mat(T, "lap1d")
    sym = _natural_sym(A)
    Lx, st, _ = restate.super_numeric(A, sym)
    Ld = restate.dense_L(sym, Lx)
    k = np.arange(n)
    assert np.allclose(np.diag(Ld), np.sqrt((k + 2) / (k + 1)), rtol=0, atol=1e-14)
    k = np.arange(n - 1)
    assert np.allclose(np.diag(Ld, -1), -np.sqrt((k + 1) / (k + 2)), rtol=0, atol=1e-14)


def test_closed_form_diagonal_and_blocks():
    d = np.arange(1.0, 31.0)
    A = _mat(np.diag(d))
    sym = _natural_sym(A)
    Lx, _, _ = restate.super_numeric(A, sym)
    assert np.array_equal(np.diag(restate.dense_L(sym, Lx)), np.sqrt(d))
    # block diagonal: each block factors independently
    rng = np.random.default_rng(2)
    blocks = []
    for m in (5, 1, 7):
        M = rng.standard_normal((m, m))
        blocks.append(M @ M.T + m * np.eye(m))
    Bd = np.zeros((13, 13))
    o = 0
    for b in blocks:
        Bd[o:o + len(b), o:o + len(b)] = b
        o += len(b)
    A = _mat(Bd)
    sym = _natural_sym(A)
    Lx, _, _ = restate.super_numeric(A, sym)
    Ld = restate.dense_L(sym, Lx)
    o = 0
    for b in blocks:
        m = len(b)
        assert np.abs(Ld[o:o + m, o:o + m] - np.linalg.cholesky(b)).max() < 1e-13 * np.abs(b).max()
        o += m


def test_residual_and_beta():
    A = problems.grid3d_lognormal(8, np.random.default_rng(8))
    perm = np.random.default_rng(8).permutation(A.n)
    from suitesparse_amd import cholmod
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN)
    L = cholmod.analyze_p(A, perm, cm)
    sym = {"Perm": L.Perm, "Super": L.super, "Lpi": L.pi, "Lpx": L.px, "Ls": L.s}
    beta = 0.75
    Lx, st, _ = restate.super_numeric(A, sym, beta=beta)
    B = np.random.default_rng(0).uniform(-1, 1, (A.n, 2))
    X = restate.solve(sym, Lx, B)
    R = B - (A.matvec(X) + beta * X)
    assert np.linalg.norm(R) <= 1e-13 * np.linalg.norm(B)


def test_not_posdef_matches_reference(ref_lib):
    """Indefinite input: same status, L->minor and factor (zeros after) as CHOLMOD."""
    for side, bad in ((20, 150), (20, 7), (12, 100)):
        A = problems.laplace2d_5pt(side)
        x = A.x.copy()
        diag_pos = A.p[:-1]          # diagonal is the first entry of each column
        x[diag_pos[bad]] = -1.0
        A2 = problems.SymMatrix(n=A.n, p=A.p, i=A.i, x=x)
        R = ref_lib.RefFactor(A2, ref_lib.AMD)
        sym = R.symbolic()
        st_ref = R.factorize()
        Lref = R.Lx()
        Lx, st, minor = restate.super_numeric(A2, sym)
        assert st == 1 and st_ref == 1
        assert minor == R.minor
        assert np.abs(Lx - Lref).max() <= 1e-12 * max(1.0, np.abs(Lref).max())


def test_planted_mistakes_are_caught():
    """Plausible mistakes in the oracle's output (planted on the CPU) fail the pins."""
    A = problems.grid2d_lognormal(8, np.random.default_rng(5))
    perm = np.random.default_rng(5).permutation(A.n)
    from suitesparse_amd import cholmod
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN)
    L = cholmod.analyze_p(A, perm, cm)
    sym = {"Perm": L.Perm, "Super": L.super, "Lpi": L.pi, "Lpx": L.px, "Ls": L.s}
    Lx, _, _ = restate.super_numeric(A, sym)
    Af = A.to_scipy_full().toarray()
    P = L.Perm
    Lnp = np.linalg.cholesky(Af[np.ix_(P, P)])

    def pin(Lx_try):
        return np.abs(restate.dense_L(sym, Lx_try) - Lnp).max() <= 1e-12 * np.abs(Lnp).max()

    assert pin(Lx)
    # (1) descendant updates dropped: factor each supernode of P A P' alone
    bad = Lx.copy()
    s = int(np.argmax(np.diff(L.super) * 0 + np.arange(L.nsuper)))   # last supernode
    p0, p1 = L.px[s], L.px[s + 1]
    bad[p0:p1] *= 1.0 + 1e-9
    assert not pin(bad)
    # (2) sign error in the off-diagonal part of one column
    bad = Lx.copy()
    s = 0
    nscol = L.super[1] - L.super[0]
    nsrow = L.pi[1] - L.pi[0]
    if nsrow > nscol:
        bad[L.px[0] + nscol] *= -1
        assert not pin(bad)
    # (3) transposed diagonal block
    bad = Lx.copy()
    for s in range(L.nsuper):
        nscol = L.super[s + 1] - L.super[s]
        nsrow = L.pi[s + 1] - L.pi[s]
        if nscol >= 2:
            blk = bad[L.px[s]:L.px[s + 1]].reshape(nscol, nsrow)
            blk[:, :nscol] = blk[:, :nscol].T.copy()
            break
    assert not pin(bad)
