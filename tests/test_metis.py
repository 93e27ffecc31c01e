"""Orderings inside the product's analyze, checked against the reference.

* mxc_metis_order must return cholmod_l_metis's permutation (METIS_NodeND on
  A + A', no postorder) element for element; the graphs cover what changes
  METIS's path: plain 2D/3D grids, multi-DOF grids whose identical
  neighbourhoods trigger compression (and two separators per level), graphs
  with isolated vertices and with several components, skewed degrees, and
  sizes on both sides of the 5000-vertex and 120-vertex switches.
* mxc_amd_order_stats must leave cholmod_l_amd's Common->fl, lnz, anz.
* CHOLMOD's default strategy (nmethods = 0: GIVEN, AMD, then METIS unless
  AMD's prediction is already good) must select the same method and produce
  the same symbolic factor, with and without a user permutation.
* Without the reference: CHOLMOD's default strategy on the 64^3 workload
  picks METIS and reproduces the committed symbolic digests.

The reference is oracle/_ref/libssref.so (CHOLMOD with its vendored METIS,
built from the sources); tests that need it skip when it is absent.
"""
import hashlib
import json
import pathlib

import numpy as np
import pytest

from suitesparse_amd import _lib, cholmod, problems

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
FIELDS = (("Perm", "Perm"), ("ColCount", "ColCount"), ("Super", "super"), ("Lpi", "pi"),
          ("Lpx", "px"), ("Ls", "s"))


def _from_edges(n, a, b, name):
    a = np.asarray(a, dtype=np.int64)
    b = np.asarray(b, dtype=np.int64)
    keep = a != b
    lo, hi = np.maximum(a[keep], b[keep]), np.minimum(a[keep], b[keep])
    key = np.unique(hi * n + lo)
    ej, ei = key // n, key % n
    w = np.ones(len(ei))
    deg = np.bincount(ei, minlength=n) + np.bincount(ej, minlength=n)
    return problems._lower_csc(n, ei, ej, w, deg + 1.0, name)


def _edges(A):
    cols = np.repeat(np.arange(A.n), np.diff(A.p))
    m = A.i != cols
    return A.i[m], cols[m]


def _multi_dof(A, dof):
    """Every vertex becomes `dof` vertices coupled to all copies of its neighbours."""
    ei, ej = _edges(A)
    a, b = [], []
    for s in range(dof):
        for t in range(dof):
            a.append(ei * dof + s)
            b.append(ej * dof + t)
        for t in range(s):
            a.append(np.arange(A.n) * dof + s)
            b.append(np.arange(A.n) * dof + t)
    return _from_edges(A.n * dof, np.concatenate(a), np.concatenate(b), f"{A.name}x{dof}")


def _random(n, deg, seed, isolated=0.0):
    rng = np.random.default_rng(seed)
    m = n * deg // 2
    a, b = rng.integers(0, n, m), rng.integers(0, n, m)
    if isolated:
        iso = rng.random(n) < isolated
        keep = ~(iso[a] | iso[b])
        a, b = a[keep], b[keep]
    return _from_edges(n, a, b, f"random{n}")


def _components(k, side):
    parts = [problems.laplace2d_5pt(side) for _ in range(k)]
    a, b, off = [], [], 0
    for P in parts:
        ei, ej = _edges(P)
        a.append(ei + off)
        b.append(ej + off)
        off += P.n
    return _from_edges(off, np.concatenate(a), np.concatenate(b), f"components{k}x{side}")


def _power_law(n, seed):
    rng = np.random.default_rng(seed)
    w = rng.pareto(1.5, n) + 1.0
    a = rng.choice(n, 4 * n, p=w / w.sum())
    b = rng.integers(0, n, 4 * n)
    return _from_edges(n, a, b, f"powerlaw{n}")


def _graphs():
    rng = np.random.default_rng(11)
    return [
        problems.laplace2d_5pt(3), problems.laplace2d_5pt(11), problems.laplace2d_5pt(64),
        problems.grid2d_lognormal(90, rng),
        problems.grid3d_lognormal(8, rng), problems.diffusion3d_7pt(20, seed=5),
        problems.stencil3d_27pt(12, seed=6), problems.stencil3d_27pt(24, seed=7),
        _multi_dof(problems.laplace2d_5pt(20), 2), _multi_dof(problems.laplace2d_5pt(24), 3),
        _multi_dof(problems.diffusion3d_7pt(12, seed=8), 2),
        _random(50, 4, 1), _random(3000, 8, 2), _random(9000, 6, 3, isolated=0.05),
        _components(6, 30), _components(40, 8), _power_law(5000, 4),
    ]


@pytest.fixture(scope="module")
def graphs():
    return _graphs()


def _metis(A):
    perm = np.zeros(A.n, dtype=np.int64)
    p = np.ascontiguousarray(A.p, dtype=np.int64)
    i = np.ascontiguousarray(A.i, dtype=np.int64)
    _lib.check("mxc_metis_order", _lib.lib().mxc_metis_order(A.n, cholmod._ptr(p), cholmod._ptr(i),
                                                              3000.0, 0.66, cholmod._ptr(perm)))
    return perm


def _amd_stats(A):
    perm = np.zeros(A.n, dtype=np.int64)
    st = np.zeros(3)
    p = np.ascontiguousarray(A.p, dtype=np.int64)
    i = np.ascontiguousarray(A.i, dtype=np.int64)
    _lib.check("mxc_amd_order_stats", _lib.lib().mxc_amd_order_stats(
        A.n, cholmod._ptr(p), cholmod._ptr(i), cholmod._ptr(perm), cholmod._ptr(st)))
    return perm, st


def test_metis_equals_reference(ref_lib, graphs):
    for A in graphs:
        rp, _ = ref_lib.order(A.n, A.p, A.i, ref_lib.METIS)
        mp = _metis(A)
        assert np.array_equal(mp, rp), A.name


def test_metis_is_a_permutation_and_deterministic(graphs):
    for A in graphs[:6]:
        p1, p2 = _metis(A), _metis(A)
        assert np.array_equal(np.sort(p1), np.arange(A.n))
        assert np.array_equal(p1, p2)


def test_metis_identity_cases():
    # no off-diagonal entries: identity, as cholmod_l_metis returns
    n = 7
    A = problems.SymMatrix(n=n, p=np.arange(n + 1, dtype=np.int64), i=np.arange(n, dtype=np.int64),
                           x=np.ones(n))
    assert np.array_equal(_metis(A), np.arange(n))
    # dense and larger than nswitch: identity (cholmod_metis.c:724-733)
    perm = np.zeros(4, dtype=np.int64)
    B = _from_edges(4, [1, 2, 3, 2, 3, 3], [0, 0, 0, 1, 1, 2], "k4")
    p, i = np.ascontiguousarray(B.p), np.ascontiguousarray(B.i)
    _lib.check("mxc_metis_order", _lib.lib().mxc_metis_order(4, cholmod._ptr(p), cholmod._ptr(i), 2.0, 0.5,
                                                              cholmod._ptr(perm)))
    assert np.array_equal(perm, np.arange(4))


def test_amd_statistics_equal_reference(ref_lib, graphs):
    for A in graphs:
        rp, rst = ref_lib.order(A.n, A.p, A.i, ref_lib.AMD)
        mp, mst = _amd_stats(A)
        assert np.array_equal(mp, rp), A.name
        assert np.array_equal(mst, rst), (A.name, mst, rst)


def test_amd_statistics_closed_form():
    """A path 0-1-...-(n-1) ordered from one end has no fill: lnz = 2n - 1,
    one division per entry of L below the diagonal and one multiply-subtract
    per column, and AMD's counts are exact there."""
    n = 30
    A = _from_edges(n, np.arange(1, n), np.arange(n - 1), "path")
    _, st = _amd_stats(A)
    assert st[1] == 2 * n - 1              # lnz
    assert st[2] == 2 * n - 1              # anz
    assert st[0] == (n - 1) + 2 * (n - 1) + n


def _check_same(L, cm, R):
    sym = R.symbolic()
    for k, a in FIELDS:
        assert np.array_equal(getattr(L, a), sym[k]), k
    assert (L.nsuper, L.ssize, L.xsize, L.maxcsize, L.maxesize) == \
        (R.nsuper, R.ssize, R.xsize, R.maxcsize, R.maxesize)
    assert L.fl == R.fl and L.lnz == R.lnz
    st = R.strategy()
    assert cm.selected == st["selected"]
    assert L.ordering == R.ordering
    for m in range(3):
        if cm.method_lnz[m] >= 0:
            assert cm.method_lnz[m] == st["method_lnz"][m], m
    assert cm.method_fl[1] == st["method_fl"][1]
    assert cm.anz == st["anz"]


@pytest.mark.parametrize("make", [
    lambda: problems.laplace2d_5pt(40),                       # AMD is good enough: stop
    lambda: problems.grid3d_lognormal(14, np.random.default_rng(3)),
    lambda: problems.stencil3d_27pt(16, seed=4),              # METIS tried
    lambda: _multi_dof(problems.diffusion3d_7pt(10, seed=9), 3),
    lambda: _random(4000, 10, 5),
], ids=["2d", "3d7", "3d27", "3d7x3dof", "random"])
def test_default_strategy_equals_reference(ref_lib, make):
    A = make()
    R = ref_lib.RefFactor(A, ref_lib.DEFAULT)
    cm = cholmod.Common()                      # nmethods = 0
    L = cholmod.analyze(A, cm)
    assert L is not None and cm.status == 0
    _check_same(L, cm, R)


def test_default_strategy_with_user_perm(ref_lib):
    rng = np.random.default_rng(12)
    for A, perm in ((problems.stencil3d_27pt(14, seed=10), None),
                    (problems.grid2d_lognormal(30, rng), None)):
        for up in (np.arange(A.n), rng.permutation(A.n)):
            R = ref_lib.RefFactor(A, ref_lib.DEFAULT, perm=up)
            cm = cholmod.Common()
            L = cholmod.analyze_p(A, up, cm)
            assert L is not None
            _check_same(L, cm, R)


def test_metis_analysis_equals_reference(ref_lib):
    for A in (problems.diffusion3d_7pt(18, seed=11), _multi_dof(problems.laplace2d_5pt(30), 2)):
        R = ref_lib.RefFactor(A, ref_lib.METIS)
        cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_METIS)
        L = cholmod.analyze(A, cm)
        assert L is not None and L.ordering == cholmod.CHOLMOD_METIS
        sym = R.symbolic()
        for k, a in FIELDS:
            assert np.array_equal(getattr(L, a), sym[k]), k


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.int64).tobytes()).hexdigest()


def test_default_strategy_workload_digest():
    """CHOLMOD's default strategy on the 64^3 workload tries AMD, finds its
    prediction poor, runs METIS and keeps it; the result matches the digests
    written from the reference's analysis (the METIS-only analyses of both
    3D workloads are checked in test_symbolic.py)."""
    name = "diffusion3d_64_metis"
    gold = json.loads((GOLD / "symbolic_golden.json").read_text())[name]
    A = problems.diffusion3d_7pt(64, seed=1)
    cm = cholmod.Common()
    L = cholmod.analyze(A, cm)
    assert L is not None and L.ordering == cholmod.CHOLMOD_METIS and cm.selected == 2
    assert cm.method_lnz[1] > cm.method_lnz[2] > 0
    for k, a in FIELDS:
        assert _digest(getattr(L, a)) == gold["sha256"][k], k
    assert (L.nsuper, L.xsize, L.fl, L.lnz) == (gold["nsuper"], gold["xsize"], gold["fl"], gold["lnz"])
