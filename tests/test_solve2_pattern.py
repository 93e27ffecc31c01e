"""Sparse right-hand-side pattern (cholmod_l_solve2 with Bset) against the reference.

mxc_lsolve_pattern computes from the symbolic factor alone the Xset that
cholmod_l_solve2 returns for a sparse b (CHOLMOD/Cholesky/cholmod_solve.c:
290-640): the columns of L reached from Bset, in the reference's own order,
permuted back for CHOLMOD_A.  The reference (oracle/_ref) runs the real solve2
(which converts its factor to simplicial form) on the same permutation, so
Xset must match entry for entry, order included.  Cases: empty Bset, one
entry, duplicates, a few scattered rows, every row; all seven systems.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from suitesparse_amd import cholmod, problems


def _dense(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n))
    T = sp.csc_matrix(np.tril(M @ M.T + n * np.eye(n)))
    return problems.SymMatrix(n=n, p=T.indptr.astype(np.int64), i=T.indices.astype(np.int64),
                              x=T.data.astype(np.float64), name=f"dense_{n}")


CASES = [
    ("lap2d_30_amd", lambda: problems.laplace2d_5pt(30), 1),
    ("st27_8_metis", lambda: problems.stencil3d_27pt(8, seed=2), 2),
    ("grid3d_10_amd", lambda: problems.grid3d_lognormal(10, np.random.default_rng(5)), 1),
    ("dense_60", lambda: _dense(60, 3), 1),
]
SYSTEMS = [cholmod.CHOLMOD_A, cholmod.CHOLMOD_LDLt, cholmod.CHOLMOD_LD, cholmod.CHOLMOD_DLt,
           cholmod.CHOLMOD_L, cholmod.CHOLMOD_Lt, cholmod.CHOLMOD_D]


def _bsets(n, rng):
    return [np.zeros(0, np.int64), np.array([n - 1]), np.array([0]), np.array([3, 3, 1]),
            rng.choice(n, size=min(7, n), replace=False), np.arange(n)[::-1].copy()]


@pytest.mark.parametrize("name,make,ordering", CASES, ids=[c[0] for c in CASES])
def test_pattern_matches_reference(name, make, ordering, ref_lib):
    A = make()
    R = ref_lib.RefFactor(A, ordering)
    sym = R.symbolic()
    assert R.factorize() == 0
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN)
    L = cholmod.analyze_p(A, sym["Perm"], cm)
    assert L is not None
    rng = np.random.default_rng(11)
    b = rng.uniform(-1, 1, A.n)
    for sysc in SYSTEMS:
        for bset in _bsets(A.n, rng):
            _, xs_ref = R.solve2_bset(sysc, b, bset)
            xs = cholmod.lsolve_pattern(L, bset, sysc)
            assert np.array_equal(xs, xs_ref), (name, sysc, bset[:8], xs[:8], xs_ref[:8])


def test_pattern_invariants():
    """Without the reference: Xset of one column is its path to the root of the
    column tree of L (j + 1 inside a supernode, else the first row below it);
    the union over Bset entries holds each path once; CHOLMOD_D keeps Bset."""
    A = problems.laplace2d_5pt(20)
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_AMD)
    L = cholmod.analyze(A, cm)
    sup, pi, s = L.super, L.pi, L.s
    col_super = np.repeat(np.arange(L.nsuper), np.diff(sup))

    def parent(j):
        k = col_super[j]
        if j + 1 < sup[k + 1]:
            return j + 1
        nscol, nsrow = sup[k + 1] - sup[k], pi[k + 1] - pi[k]
        return int(s[pi[k] + nscol]) if nsrow > nscol else -1

    for j in (0, 7, L.n // 2, L.n - 1):
        path = []
        i = j
        while i != -1:
            path.append(i)
            i = parent(i)
        assert list(cholmod.lsolve_pattern(L, [j], cholmod.CHOLMOD_L)) == path
    xs = cholmod.lsolve_pattern(L, [0, 5, 9, 0], cholmod.CHOLMOD_L)
    assert len(set(xs.tolist())) == len(xs)
    assert list(cholmod.lsolve_pattern(L, [4, 2, 4], cholmod.CHOLMOD_D)) == [4, 2, 4]
    inv = np.empty(L.n, np.int64)
    inv[L.Perm] = np.arange(L.n)
    xa = cholmod.lsolve_pattern(L, [3], cholmod.CHOLMOD_A)
    assert list(inv[xa]) == list(cholmod.lsolve_pattern(L, [int(inv[3])], cholmod.CHOLMOD_L))
