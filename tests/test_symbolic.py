"""Bit-exactness of the product's CPU symbolic analysis with the reference.

For a given ordering the product must reproduce cholmod_l_analyze_p's Perm,
ColCount, Super, Lpi, Lpx, Ls and the scalars (nsuper, ssize, xsize,
maxcsize, maxesize, fl, lnz) exactly: against the compiled reference where it
is built, and against committed SHA-256 digests for the benchmark workloads
(ordered by the product's own AMD and METIS).
"""
import hashlib
import json
import pathlib

import numpy as np
import pytest

from suitesparse_amd import cholmod, problems

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
FIELDS = (("Perm", "Perm"), ("ColCount", "ColCount"), ("Super", "super"), ("Lpi", "pi"),
          ("Lpx", "px"), ("Ls", "s"))


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.int64).tobytes()).hexdigest()


def _check_equal(L, R, sym):
    for k, a in FIELDS:
        assert np.array_equal(getattr(L, a), sym[k]), k
    assert (L.nsuper, L.ssize, L.xsize, L.maxcsize, L.maxesize) == \
        (R.nsuper, R.ssize, R.xsize, R.maxcsize, R.maxesize)
    assert L.fl == R.fl and L.lnz == R.lnz


def _matrices():
    rng = np.random.default_rng(21)
    return [problems.laplace2d_5pt(37), problems.grid2d_lognormal(50, rng),
            problems.grid3d_lognormal(13, rng), problems.stencil3d_27pt(11, seed=3),
            problems.diffusion3d_7pt(10, seed=4)]


@pytest.mark.parametrize("ordering", [1, 2, 3], ids=["amd", "metis", "natural"])
def test_bitexact_vs_reference(ref_lib, ordering):
    for A in _matrices():
        R = ref_lib.RefFactor(A, ordering)
        sym = R.symbolic()
        cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN)
        L = cholmod.analyze_p(A, sym["Perm"], cm)
        assert L is not None and cm.status == 0
        _check_equal(L, R, sym)


def test_bitexact_random_orderings(ref_lib):
    """GIVEN orderings that are not postorders exercise the postorder fold."""
    rng = np.random.default_rng(5)
    for A in _matrices()[:3]:
        perm = rng.permutation(A.n)
        R = ref_lib.RefFactor(A, ref_lib.GIVEN, perm=perm)
        sym = R.symbolic()
        cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN)
        L = cholmod.analyze_p(A, perm, cm)
        _check_equal(L, R, sym)


@pytest.mark.parametrize("name,make,ordering", [
    ("laplace2d_100_amd", lambda: problems.laplace2d_5pt(100), cholmod.CHOLMOD_AMD),
    ("diffusion3d_64_metis", lambda: problems.diffusion3d_7pt(64, seed=1), cholmod.CHOLMOD_METIS),
    ("stencil27_96_metis", lambda: problems.stencil3d_27pt(96, seed=2), cholmod.CHOLMOD_METIS),
])
def test_golden_digests(name, make, ordering):
    """The product's own ordering (AMD or METIS) and analysis reproduce the
    digests written from the reference's analysis of the benchmark workloads."""
    gold = json.loads((GOLD / "symbolic_golden.json").read_text())[name]
    A = make()
    cm = cholmod.Common(nmethods=1, ordering=ordering)
    L = cholmod.analyze(A, cm)
    assert L is not None and L.ordering == ordering
    for k, a in FIELDS:
        assert _digest(getattr(L, a)) == gold["sha256"][k], k
    assert (L.n, L.nsuper, L.ssize, L.xsize, L.maxcsize, L.maxesize) == \
        (gold["n"], gold["nsuper"], gold["ssize"], gold["xsize"], gold["maxcsize"], gold["maxesize"])
    assert L.fl == gold["fl"] and L.lnz == gold["lnz"]


def test_invariants_small():
    """Structural invariants against a brute-force symbolic elimination."""
    A = problems.grid2d_lognormal(9, np.random.default_rng(1))
    perm = np.random.default_rng(1).permutation(A.n)
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN)
    L = cholmod.analyze_p(A, perm, cm)
    P = L.Perm
    assert np.array_equal(np.sort(P), np.arange(A.n))
    # brute-force pattern of L for P A P'
    Af = A.to_scipy_full().toarray()[np.ix_(P, P)] != 0
    n = A.n
    pat = np.tril(Af).copy()
    for k in range(n):
        rows = np.nonzero(pat[k + 1:, k])[0] + k + 1
        for i in rows:
            pat[rows, i] |= True
    pat = np.tril(pat)
    assert np.array_equal(L.ColCount, pat.sum(axis=0))
    assert L.lnz == pat.sum() and L.fl == float((pat.sum(axis=0) ** 2).sum())
    # every column of L lies in its supernode's row list
    for s in range(L.nsuper):
        rows = set(L.s[L.pi[s]:L.pi[s + 1]].tolist())
        for k in range(L.super[s], L.super[s + 1]):
            assert set(np.nonzero(pat[:, k])[0].tolist()) <= rows


def test_import_roundtrip(ref_lib):
    A = problems.grid3d_lognormal(8, np.random.default_rng(2))
    R = ref_lib.RefFactor(A, ref_lib.METIS)
    sym = R.symbolic()
    cm = cholmod.Common()
    L = cholmod.import_symbolic(A, sym, cm)
    assert L is not None and cm.status == 0
    for k, a in FIELDS[:1] + FIELDS[2:]:
        assert np.array_equal(getattr(L, a), sym[k])
    assert L.xsize == R.xsize and L.maxesize == R.maxesize


def test_invalid_inputs():
    A = problems.laplace2d_5pt(5)
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN)
    assert cholmod.analyze_p(A, np.zeros(A.n, dtype=np.int64), cm) is None
    assert cm.status == cholmod.CHOLMOD_INVALID
    assert cholmod.analyze_p(A, np.arange(A.n - 1), cm) is None
    assert cm.status == cholmod.CHOLMOD_INVALID
    bad = {"Perm": np.arange(A.n), "Super": np.array([0, A.n + 1]), "Lpi": np.array([0, 1]),
           "Lpx": np.array([0, 1]), "Ls": np.array([0])}
    assert cholmod.import_symbolic(A, bad, cm) is None and cm.status == cholmod.CHOLMOD_INVALID


def _random_pattern(n, density, seed, dense_rows=0):
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=density, random_state=seed, format="csc")
    M = M + M.T
    if dense_rows:
        rows = rng.choice(n, dense_rows, replace=False)
        D = sp.lil_matrix((n, n))
        for r in rows:
            D[r, :] = 1.0
            D[:, r] = 1.0
        M = M + D.tocsc()
    M = (abs(M) + sp.eye(n) * n).tocsc()
    T = sp.tril(M).tocsc()
    T.sort_indices()
    return problems.SymMatrix(n=n, p=T.indptr.astype(np.int64), i=T.indices.astype(np.int64),
                              x=T.data.astype(np.float64), name=f"rand{n}_{seed}")


def test_amd_bitexact_vs_reference(ref_lib):
    """The product's AMD + analysis reproduces cholmod_l_analyze with CHOLMOD_AMD."""
    mats = _matrices() + [problems.laplace2d_5pt(200),
                          _random_pattern(300, 0.02, 1), _random_pattern(500, 0.01, 2, dense_rows=3),
                          _random_pattern(64, 0.3, 3), _random_pattern(1000, 0.004, 4, dense_rows=1)]
    for A in mats:
        R = ref_lib.RefFactor(A, ref_lib.AMD)
        sym = R.symbolic()
        cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_AMD)
        L = cholmod.analyze(A, cm)
        assert L is not None and cm.status == 0
        _check_equal(L, R, sym)


def test_amd_golden_laplace100():
    """Without the reference: AMD on the 100x100 Laplacian reproduces the committed digests."""
    gold = json.loads((GOLD / "symbolic_golden.json").read_text())["laplace2d_100_amd"]
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_AMD)
    L = cholmod.analyze(problems.laplace2d_5pt(100), cm)
    for k, a in FIELDS:
        assert _digest(getattr(L, a)) == gold["sha256"][k], k
