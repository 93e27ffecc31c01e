"""The oracle's rank-k update/downdate (oracle_updown), pinned on the CPU:

* against the mathematics: Lnew is the Cholesky factor of L L' +- C C'
  (numpy on the densified factor), which is unique, so a wrong sign, a
  dropped rank or a misrouted row shows;
* against the reference: cholmod_l_updown (Modify/cholmod_updown.c:74) on the
  same factor gives the same values to rounding (it orders merged ranks
  differently), the same Common->modfl and ndbounds_hit;
* against closed forms: modfl of a path graph, the first non-positive pivot
  of a downdate of the identity, the bounded pivot of dbound;
* an update followed by the same downdate gives back the factor.

Factors come from the oracle's own supernodal factorization of the product's
analysis (AMD / METIS); the reference analyses the same matrices with the
same orderings and reaches the same symbolic factor (tests/test_metis.py).
"""
import numpy as np
import pytest

from oracle import restate
from suitesparse_amd import cholmod, problems

from updown_cases import dense_c, pattern_columns


def _factor(A, ordering):
    cm = cholmod.Common(nmethods=1, ordering=ordering)
    L = cholmod.analyze(A, cm)
    sym = {"Perm": L.Perm, "Super": L.super, "Lpi": L.pi, "Lpx": L.px, "Ls": L.s}
    Lx, st, _ = restate.super_numeric(A, sym)
    assert st == 0
    return sym, Lx


def _cases():
    return [
        ("lap2d_12_amd", lambda: problems.laplace2d_5pt(12), cholmod.CHOLMOD_AMD),
        ("grid3d_7_amd", lambda: problems.grid3d_lognormal(7, np.random.default_rng(4)), cholmod.CHOLMOD_AMD),
        ("diff3d_10_metis", lambda: problems.diffusion3d_7pt(10, seed=5), cholmod.CHOLMOD_METIS),
        ("st27_7_metis", lambda: problems.stencil3d_27pt(7, seed=6), cholmod.CHOLMOD_METIS),
    ]


RANKS = [(1, True, 1.0), (3, True, 1.0), (8, True, 0.7), (11, True, 0.5), (2, False, 0.05), (9, False, 0.03)]


@pytest.mark.parametrize("name,make,ordering", _cases(), ids=[c[0] for c in _cases()])
def test_oracle_is_the_cholesky_factor(name, make, ordering):
    A = make()
    sym, Lx0 = _factor(A, ordering)
    rng = np.random.default_rng(7)
    L0 = restate.dense_L(sym, Lx0)
    for k, update, scale in RANKS:
        Cp, Ci, Cx = pattern_columns(sym, rng.integers(0, A.n, k), rng, scale=scale, dup=(k == 3))
        o = restate.updown(sym, Lx0, update, Cp, Ci, Cx)
        assert o["rc"] == 0 and o["minor"] == A.n
        Cd = dense_c(A.n, Cp, Ci, Cx)
        Le = np.linalg.cholesky(L0 @ L0.T + (1 if update else -1) * Cd @ Cd.T)
        Lo = restate.dense_L(sym, o["Lx"])
        assert np.abs(Lo - Le).max() <= 1e-13 * np.abs(Le).max(), (name, k, update)


@pytest.mark.parametrize("name,make,ordering", _cases(), ids=[c[0] for c in _cases()])
def test_oracle_equals_reference(ref_lib, name, make, ordering):
    A = make()
    sym, Lx0 = _factor(A, ordering)
    rng = np.random.default_rng(8)
    ref_order = ref_lib.AMD if ordering == cholmod.CHOLMOD_AMD else ref_lib.METIS
    for k, update, scale in RANKS + [(0, True, 1.0)]:
        Cp, Ci, Cx = pattern_columns(sym, rng.integers(0, A.n, k), rng, scale=scale, dup=(k == 3))
        if k == 8:                                     # two empty columns among them
            Cp = np.concatenate([Cp[:3], Cp[2:3], Cp[3:], Cp[-1:]])
        R = ref_lib.RefFactor(A, ref_order)
        R.factorize()
        rs = R.symbolic()
        assert all(np.array_equal(rs[a], sym[a]) for a in ("Perm", "Super", "Lpi", "Lpx", "Ls"))
        Lr0 = R.Lx()                                   # both start from the reference's factor
        r = R.updown(update, Cp, Ci, Cx)
        o = restate.updown(sym, Lr0, update, Cp, Ci, Cx)
        assert r["status"] == 0
        assert o["modfl"] == r["modfl"], (name, k)
        scale_l = np.abs(o["Lx"]).max()
        assert np.abs(o["Lx"] - r["Lx"]).max() <= 1e-14 * scale_l * max(1, k), (name, k, update)


def test_update_then_downdate_gives_back_the_factor():
    A = problems.diffusion3d_7pt(9, seed=9)
    sym, Lx0 = _factor(A, cholmod.CHOLMOD_METIS)
    rng = np.random.default_rng(10)
    Cp, Ci, Cx = pattern_columns(sym, rng.integers(0, A.n, 6), rng)
    up = restate.updown(sym, Lx0, True, Cp, Ci, Cx)
    back = restate.updown(sym, up["Lx"], False, Cp, Ci, Cx)
    assert back["minor"] == A.n
    assert np.abs(back["Lx"] - Lx0).max() <= 1e-12 * np.abs(Lx0).max()
    assert not np.array_equal(up["Lx"], Lx0)


def test_modfl_closed_forms():
    """modfl (t_cholmod_updown_worker.c:784) sums 6 + 4 |L(:, j)| over each
    path.  A dense matrix is one supernode: the path of column k is k..n-1 and
    |L(:, j)| = n - j, so a rank-1 modification from k costs
    6 (n-k) + 2 (n-k)(n-k+1).  The identity has one column per supernode and no
    parents: 10 per column of C, and an empty column of C counts as one entry
    at n-1 (worker.c:205-209)."""
    n = 12
    ii, jj = np.tril_indices(n)
    order = np.lexsort((ii, jj))
    ii, jj = ii[order], jj[order]
    p = np.concatenate([[0], np.cumsum(np.bincount(jj, minlength=n))]).astype(np.int64)
    A = problems.SymMatrix(n=n, p=p, i=ii.astype(np.int64), x=np.where(ii == jj, float(n), 1.0))
    sym, Lx0 = _factor(A, cholmod.CHOLMOD_NATURAL)
    assert len(sym["Super"]) == 2
    for k in (0, 5, n - 1):
        m = n - k
        o = restate.updown(sym, Lx0, True, [0, 1], [k], [0.5])
        assert o["modfl"] == 6 * m + 2 * m * (m + 1)
    I = _identity(7)
    sym, Lx0 = _factor(I, cholmod.CHOLMOD_NATURAL)
    assert restate.updown(sym, Lx0, True, [0, 1, 2, 2], [3, 5], [1.0, 2.0])["modfl"] == 30


def _identity(n):
    return problems.SymMatrix(n=n, p=np.arange(n + 1, dtype=np.int64), i=np.arange(n, dtype=np.int64),
                              x=np.ones(n))


def test_downdate_that_loses_positive_definiteness(ref_lib):
    """I - 4 e_2 e_2' has d_2 = -3: the oracle names column 2; the reference
    keeps the negative d (sqrt gives NaN in its LL' view)."""
    A = _identity(5)
    sym, Lx0 = _factor(A, cholmod.CHOLMOD_NATURAL)
    o = restate.updown(sym, Lx0, False, [0, 1], [2], [2.0])
    assert o["minor"] == 2
    R = ref_lib.RefFactor(A, ref_lib.NATURAL)
    R.factorize()
    r = R.updown(False, [0, 1], [2], [2.0])
    assert np.isnan(r["Lx"][sym["Lpx"][2]])


def test_dbound(ref_lib):
    """I - 0.998 e_1 e_1': d_1 = 0.002 is raised to dbound = 0.01, once, as
    cholmod_dbound does (Utility/t_cholmod_bound.c)."""
    A = _identity(4)
    sym, Lx0 = _factor(A, cholmod.CHOLMOD_NATURAL)
    c = np.sqrt(0.998)
    o = restate.updown(sym, Lx0, False, [0, 1], [1], [c], dbound=0.01)
    assert o["ndbounds_hit"] == 1 and o["minor"] == 4
    assert o["Lx"][sym["Lpx"][1]] == np.sqrt(0.01)
    R = ref_lib.RefFactor(A, ref_lib.NATURAL)
    R.factorize()
    r = R.updown(False, [0, 1], [1], [c], dbound=0.01)
    assert r["ndbounds_hit"] == 1
    assert r["Lx"][sym["Lpx"][1]] == o["Lx"][sym["Lpx"][1]]


def test_entry_outside_the_pattern_is_refused():
    A = problems.laplace2d_5pt(6)
    sym, Lx0 = _factor(A, cholmod.CHOLMOD_AMD)
    cs = np.repeat(np.arange(len(sym["Super"]) - 1), np.diff(sym["Super"]))
    # column 0's pattern: find a row > 0 outside it
    s = cs[0]
    rows = set(sym["Ls"][sym["Lpi"][s]:sym["Lpi"][s + 1]].tolist())
    out = next(r for r in range(1, A.n) if r not in rows)
    o = restate.updown(sym, Lx0, True, [0, 2], [0, out], [1.0, 1.0])
    assert o["rc"] == -1
