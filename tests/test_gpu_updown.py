"""Rank-k update/downdate of the device factor (mxc_updown, cholmod.updown)
against the oracle (oracle_updown, run on the device factor's own values) and
the reference's cholmod_l_updown (CHOLMOD/Modify/cholmod_updown.c:74).

FP64: the device performs the oracle's operations in the oracle's order with
no contraction, so L must agree bit for bit; modfl exactly.  Against the
reference (which orders merged ranks by its own path walk) the values agree
to 1e-14 relative.  The renewed inverse blocks are checked through solves.
"""
import numpy as np
import pytest

from oracle import restate
from suitesparse_amd import cholmod, problems

from updown_cases import pattern_columns, sym_of

pytestmark = pytest.mark.gpu


def _factor(A, ordering, precision="fp64"):
    cm = cholmod.Common(nmethods=1, ordering=ordering, precision=precision)
    L = cholmod.analyze(A, cm)
    assert L is not None
    assert cholmod.factorize(A, L, cm) and cm.status == 0
    return L, cm


CASES = [
    ("lap2d_40_amd", lambda: problems.laplace2d_5pt(40), cholmod.CHOLMOD_AMD),
    ("diff3d_20_metis", lambda: problems.diffusion3d_7pt(20, seed=3), cholmod.CHOLMOD_METIS),
    ("st27_12_metis", lambda: problems.stencil3d_27pt(12, seed=4), cholmod.CHOLMOD_METIS),
]
RANKS = [(1, True, 1.0), (5, True, 1.0), (8, True, 0.7), (13, True, 0.5), (3, False, 0.05)]


@pytest.mark.parametrize("name,make,ordering", CASES, ids=[c[0] for c in CASES])
def test_updown_equals_oracle(gpu, name, make, ordering):
    A = make()
    L, cm = _factor(A, ordering)
    sym = sym_of(L)
    rng = np.random.default_rng(21)
    b = rng.uniform(-1, 1, A.n)
    for k, update, scale in RANKS:
        Lx0 = L.x
        Cp, Ci, Cx = pattern_columns(sym, rng.integers(0, A.n, k), rng, scale=scale, dup=(k == 5))
        print(f"  {name} rank {k} {'update' if update else 'downdate'}", flush=True)
        o = restate.updown(sym, Lx0, update, Cp, Ci, Cx)
        assert o["rc"] == 0 and o["minor"] == A.n
        assert cholmod.updown(update, (Cp, Ci, Cx), L, cm), cm.status
        assert cm.status == 0 and cm.modfl == o["modfl"]
        Lx = L.x
        d = np.abs(Lx - o["Lx"]).max() / np.abs(o["Lx"]).max()
        print(f"    max rel diff vs oracle {d:.3e} bitwise {np.array_equal(Lx, o['Lx'])}", flush=True)
        assert np.array_equal(Lx, o["Lx"])
        # the solve runs through the renewed diagonal-block inverses
        x = cholmod.solve(cholmod.CHOLMOD_A, L, b, cm)
        xo = restate.solve(sym, o["Lx"], b)
        assert np.abs(x - xo).max() <= 1e-11 * np.abs(xo).max(), name


def test_updown_equals_reference(gpu, ref_lib):
    A = problems.diffusion3d_7pt(16, seed=7)
    rng = np.random.default_rng(22)
    for k, update, scale in ((1, True, 1.0), (8, True, 1.0), (4, False, 0.05)):
        L, cm = _factor(A, cholmod.CHOLMOD_METIS)
        sym = sym_of(L)
        R = ref_lib.RefFactor(A, ref_lib.METIS)
        R.factorize()
        Cp, Ci, Cx = pattern_columns(sym, rng.integers(0, A.n, k), rng, scale=scale)
        Lg0, Lr0 = L.x, R.Lx()
        assert np.abs(Lg0 - Lr0).max() <= 1e-13 * np.abs(Lr0).max()
        r = R.updown(update, Cp, Ci, Cx)
        assert cholmod.updown(update, (Cp, Ci, Cx), L, cm)
        assert cm.modfl == r["modfl"]
        Lx = L.x
        assert np.abs(Lx - r["Lx"]).max() <= 1e-13 * np.abs(r["Lx"]).max(), k
        b = rng.uniform(-1, 1, A.n)
        x = cholmod.solve(cholmod.CHOLMOD_A, L, b, cm)
        xr = R.solve(b)[:, 0]
        assert np.linalg.norm(x - xr) <= 1e-10 * np.linalg.norm(xr)


def test_update_then_downdate_and_refactor(gpu):
    A = problems.diffusion3d_7pt(14, seed=8)
    L, cm = _factor(A, cholmod.CHOLMOD_METIS)
    sym = sym_of(L)
    rng = np.random.default_rng(23)
    Lx0 = L.x
    b = rng.uniform(-1, 1, A.n)
    x0 = cholmod.solve(cholmod.CHOLMOD_A, L, b, cm)
    Cp, Ci, Cx = pattern_columns(sym, rng.integers(0, A.n, 6), rng)
    assert cholmod.updown(True, (Cp, Ci, Cx), L, cm)
    # the stored A no longer matches: refinement is refused until a factorization
    X, info = cholmod.solve_refine(L, b, cm)
    assert X is None and cm.status == cholmod.CHOLMOD_INVALID
    assert cholmod.updown(False, (Cp, Ci, Cx), L, cm)
    assert np.abs(L.x - Lx0).max() <= 1e-12 * np.abs(Lx0).max()
    x1 = cholmod.solve(cholmod.CHOLMOD_A, L, b, cm)
    assert np.abs(x1 - x0).max() <= 1e-10 * np.abs(x0).max()
    assert cholmod.factorize(A, L, cm)
    assert np.array_equal(L.x, Lx0) or np.abs(L.x - Lx0).max() <= 1e-12 * np.abs(Lx0).max()
    X, info = cholmod.solve_refine(L, b, cm)
    assert X is not None and info["relres"] <= 1e-10


def _identity(n, d=1.0):
    return problems.SymMatrix(n=n, p=np.arange(n + 1, dtype=np.int64), i=np.arange(n, dtype=np.int64),
                              x=np.full(n, d))


def test_downdate_not_posdef_and_dbound(gpu):
    A = _identity(300)
    L, cm = _factor(A, cholmod.CHOLMOD_NATURAL)
    # I - 0.998 e_7 e_7': the pivot 0.002 is raised to dbound
    cm.dbound = 0.01
    assert cholmod.updown(False, ([0, 1], [7], [np.sqrt(0.998)]), L, cm)
    assert cm.status == cholmod.CHOLMOD_DSMALL and cm.ndbounds_hit == 1
    assert L.x[L.px[7]] == np.sqrt(0.01)
    cm.dbound = 0.0
    # I - 4 e_150 e_150': not positive definite at column 150
    assert not cholmod.updown(False, ([0, 1], [150], [2.0]), L, cm)
    assert cm.status == cholmod.CHOLMOD_NOT_POSDEF and L.minor == 150
    assert not cholmod.updown(True, ([0, 1], [3], [1.0]), L, cm) and cm.status == cholmod.CHOLMOD_INVALID
    assert cholmod.factorize(A, L, cm) and cm.status == 0
    assert np.array_equal(L.x[L.px[:-1]], np.ones(300))


def test_refused_inputs_leave_L_unchanged(gpu):
    A = problems.laplace2d_5pt(20)
    L, cm = _factor(A, cholmod.CHOLMOD_AMD)
    sym = sym_of(L)
    Lx0 = L.x
    cs = np.repeat(np.arange(L.nsuper), np.diff(L.super))
    s = cs[0]
    rows = set(L.s[L.pi[s]:L.pi[s + 1]].tolist())
    out = next(r for r in range(1, A.n) if r not in rows)
    assert not cholmod.updown(True, ([0, 2], [0, out], [1.0, 1.0]), L, cm)
    assert cm.status == cholmod.CHOLMOD_INVALID
    assert not cholmod.updown(True, ([0, 2], [5, 3], [1.0, 1.0]), L, cm)     # unsorted column
    assert cm.status == cholmod.CHOLMOD_INVALID
    assert np.array_equal(L.x, Lx0)
    # no columns, and empty columns only: nothing changes, modfl as the reference counts it
    assert cholmod.updown(True, ([0], [], []), L, cm) and cm.modfl == 0
    assert cholmod.updown(True, ([0, 0, 0], [], []), L, cm)
    assert cm.modfl == restate.updown(sym, Lx0, True, [0, 0, 0], [], [])["modfl"]
    assert np.array_equal(L.x, Lx0)


def test_updown_fp32_factor(gpu):
    A = problems.diffusion3d_7pt(16, seed=9)
    L, cm = _factor(A, cholmod.CHOLMOD_METIS, precision="fp32")
    sym = sym_of(L)
    rng = np.random.default_rng(24)
    Lx0 = L.x                                        # FP32 values, as double
    Cp, Ci, Cx = pattern_columns(sym, rng.integers(0, A.n, 8), rng)
    o = restate.updown(sym, Lx0, True, Cp, Ci, Cx)   # the same update in double
    assert cholmod.updown(True, (Cp, Ci, Cx), L, cm)
    assert cm.modfl == o["modfl"]
    assert np.abs(L.x - o["Lx"]).max() <= 2e-5 * np.abs(o["Lx"]).max()
    b = rng.uniform(-1, 1, A.n)
    x = cholmod.solve(cholmod.CHOLMOD_A, L, b, cm)
    xo = restate.solve(sym, o["Lx"], b)
    assert np.linalg.norm(x - xo) <= 1e-4 * np.linalg.norm(xo)
