"""GPU: the fused small-supernode launch (k_small_level, FP64 factors).

Supernodes of at most 16 columns and 64 rows that have descendant updates
are factored whole, a workgroup each, in one launch per etree level, in the
levels with at most a threshold of them (an option, off by default).  On the
same inputs, the factor with
the launch on must match the factor with it off (the level's update, POTRF
and TRSM launches) to 1e-13 * max|L| (the update sums and the TRSM differ in
order of operations only), both must match the reference's factor to the
1e-12 bar of the other parity tests, and their solutions the reference's to
1e-10.  A not-positive-definite matrix whose failing column lies in a small
supernode must give the reference's status, L->minor and factor either way.
"""
import numpy as np
import pytest

from suitesparse_amd import _lib, batch, cholmod, problems

pytestmark = pytest.mark.gpu


def _launches(L):
    counts, ms = np.zeros(8, dtype=np.int64), np.zeros(4)
    _lib.check("mxc_factor_stats", _lib.lib().mxc_factor_stats(L._num, cholmod._ptr(counts), cholmod._ptr(ms)))
    return int(counts[4])


def _cases():
    return [
        ("lap2d_100_amd", lambda: problems.laplace2d_5pt(100), 1),
        ("grid2d_60_amd", lambda: problems.grid2d_lognormal(60, np.random.default_rng(4)), 1),
        ("grid3d_14_metis", lambda: problems.grid3d_lognormal(14, np.random.default_rng(7)), 2),
        ("st27_10_metis", lambda: problems.stencil3d_27pt(10, seed=5), 2),
    ]


@pytest.mark.parametrize("name,make,ordering", _cases(), ids=[c[0] for c in _cases()])
def test_fused_matches_unfused_and_reference(name, make, ordering, gpu, ref_lib):
    A = make()
    print(f"case {name} n={A.n}", flush=True)
    R = ref_lib.RefFactor(A, ordering)
    sym = R.symbolic()
    assert R.factorize() == 0
    Lref = R.Lx()
    B = np.random.default_rng(1).uniform(-1, 1, (A.n, 2))
    Xref = R.solve(B)
    xs, launches = {}, {}
    for fused in (True, False):
        cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision="fp64")
        L = cholmod.analyze_p(A, sym["Perm"], cm)
        L._numeric(cm)
        L.set_small_fused(fused)
        launches[fused] = _launches(L)
        for rep in range(2):                  # graph capture, then replay
            assert cholmod.factorize(A, L, cm) and cm.status == cholmod.CHOLMOD_OK and L.minor == A.n
        xs[fused] = L.x
        err = np.abs(xs[fused] - Lref).max() / np.abs(Lref).max()
        X = cholmod.solve(cholmod.CHOLMOD_A, L, B, cm)
        fe = np.linalg.norm(X - Xref) / np.linalg.norm(Xref)
        print(f"  fused={fused} launches={launches[fused]} L err vs ref {err:.2e} x err {fe:.2e}", flush=True)
        assert err <= 1e-12 and fe <= 1e-10
    d = np.abs(xs[True] - xs[False]).max() / np.abs(xs[False]).max()
    print(f"  fused vs unfused {d:.2e}", flush=True)
    assert d <= 1e-13
    # the option changes the computation: the fused rows below the diagonal
    # block come from elimination, the TRSM launch multiplies by the inverse
    assert d > 0


@pytest.mark.parametrize("threshold", [8, 64])
def test_fused_threshold_matches_unfused(threshold, gpu):
    """Only the levels with at most `threshold` small supernodes are fused; the
    others keep the separate launches in the same factorization."""
    A = problems.grid2d_lognormal(80, np.random.default_rng(9))
    xs = {}
    for v in (0, threshold):
        cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_AMD, precision="fp64")
        L = cholmod.analyze(A, cm)
        L._numeric(cm)
        L.set_small_fused(v)
        assert cholmod.factorize(A, L, cm) and cm.status == cholmod.CHOLMOD_OK
        xs[v] = L.x
    d = np.abs(xs[threshold] - xs[0]).max() / np.abs(xs[0]).max()
    print(f"threshold {threshold}: vs unfused {d:.2e}", flush=True)
    assert d <= 1e-13


def test_fused_batch_matches_reference(gpu, ref_lib):
    rng = np.random.default_rng(21)
    mats = [problems.grid2d_lognormal(int(s), rng, shift=1e-3) for s in rng.integers(6, 40, size=60)]
    cm = cholmod.Common(precision="fp64")
    b = batch.analyze_batch(mats, cm)
    b.L._numeric(cm)
    b.L.set_small_fused(True)
    assert batch.factorize_batch(b), cm.status
    B = np.random.default_rng(2).uniform(-1, 1, (b.A.n, 3))
    X = batch.solve_batch(b, B)
    worst = 0.0
    for j, A in enumerate(mats):
        sl = b.system_slice(j)
        R = ref_lib.RefFactor(A, ref_lib.AMD)
        R.factorize()
        Xr = R.solve(B[sl])
        worst = max(worst, np.linalg.norm(X[sl] - Xr) / np.linalg.norm(Xr))
    print(f"batch of {len(mats)}: worst fwd err {worst:.2e}", flush=True)
    assert worst <= 1e-10


@pytest.mark.parametrize("side,bad", [(30, 5), (30, 400), (20, 17)])
def test_fused_not_posdef_matches_reference(side, bad, gpu, ref_lib):
    A = problems.laplace2d_5pt(side)
    x = A.x.copy()
    x[A.p[:-1][bad]] = -1.0
    A2 = problems.SymMatrix(n=A.n, p=A.p, i=A.i, x=x)
    R = ref_lib.RefFactor(A2, ref_lib.AMD)
    sym = R.symbolic()
    assert R.factorize() == 1
    Lref = R.Lx()
    print(f"case side={side} bad={bad} ref minor={R.minor}", flush=True)
    for fused in (True, False):
        cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision="fp64")
        L = cholmod.analyze_p(A2, sym["Perm"], cm)
        L._numeric(cm)
        L.set_small_fused(fused)
        assert cholmod.factorize(A2, L, cm)
        assert cm.status == cholmod.CHOLMOD_NOT_POSDEF
        assert L.minor == R.minor
        err = np.abs(L.x - Lref).max() / np.abs(Lref).max()
        print(f"  fused={fused} minor={L.minor} err {err:.2e}", flush=True)
        assert err <= 1e-12
