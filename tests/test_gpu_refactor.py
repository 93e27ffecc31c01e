"""GPU: same-pattern refactorization through captured hipGraphs, not-positive-
definite handling, and device-resident inputs.

Graph replays are checked against the eager (graphs off) path on every
replay, with eager work and synchronizations between replays; the not-posdef
case must reproduce the reference's status, L->minor and factor (zeros after
the failing column)."""
import numpy as np
import pytest

from suitesparse_amd import cholmod, problems

pytestmark = pytest.mark.gpu


def test_graph_replay_matches_eager(gpu, ref_lib):
    rng = np.random.default_rng(3)
    A0 = problems.grid3d_lognormal(14, rng)
    R = ref_lib.RefFactor(A0, ref_lib.METIS)
    perm = R.symbolic()["Perm"]
    cmg = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision="fp64")
    cme = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision="fp64")
    Lg = cholmod.analyze_p(A0, perm, cmg)
    Le = cholmod.analyze_p(A0, perm, cme)
    B = rng.uniform(-1, 1, (A0.n, 2))
    for rep in range(4):
        A = problems.revalue(A0, np.random.default_rng(100 + rep))
        print(f"replay {rep}", flush=True)
        assert cholmod.factorize(A, Lg, cmg)
        if rep == 0:
            Le._numeric(cme)
            Le.set_graphs(False)
        assert cholmod.factorize(A, Le, cme)
        xg, xe = Lg.x, Le.x
        assert np.abs(xg - xe).max() <= 1e-12 * np.abs(xe).max()
        Xg = cholmod.solve(cholmod.CHOLMOD_A, Lg, B, cmg)
        Xe = cholmod.solve(cholmod.CHOLMOD_A, Le, B, cme)
        assert np.linalg.norm(Xg - Xe) <= 1e-12 * np.linalg.norm(Xe)
        R.factorize(A.x)
        Xr = R.solve(B)
        assert np.linalg.norm(Xg - Xr) <= 1e-10 * np.linalg.norm(Xr)


def test_not_posdef_matches_reference(gpu, ref_lib):
    for side, bad in ((30, 400), (30, 5), (12, 143)):
        A = problems.laplace2d_5pt(side)
        x = A.x.copy()
        x[A.p[:-1][bad]] = -1.0
        A2 = problems.SymMatrix(n=A.n, p=A.p, i=A.i, x=x)
        R = ref_lib.RefFactor(A2, ref_lib.AMD)
        sym = R.symbolic()
        assert R.factorize() == 1
        Lref = R.Lx()
        print(f"case side={side} bad={bad} ref minor={R.minor}", flush=True)
        for prec in ("fp64",):
            cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision=prec)
            L = cholmod.analyze_p(A2, sym["Perm"], cm)
            assert cholmod.factorize(A2, L, cm)
            assert cm.status == cholmod.CHOLMOD_NOT_POSDEF
            assert L.minor == R.minor
            err = np.abs(L.x - Lref).max() / np.abs(Lref).max()
            assert err <= 1e-12, err


def test_not_posdef_inside_a_supernode(gpu, ref_lib):
    """A dense 320-column supernode (five 64-column steps, natural order) whose
    leading minor turns indefinite in step 0, 2 or 4.  From step 1 on, a step's
    diagonal block is factored at the end of the previous step's trailing
    launch (the workgroup of the block's tile); the breakdown it finds must give
    the reference's status, L->minor and factor (zeros after the failing
    column), in FP64 and, with the FP32 pivot safeguard off, in FP32."""
    import scipy.sparse as sp
    n = 320
    rng = np.random.default_rng(21)
    M = rng.standard_normal((n, n))
    A = M @ M.T + n * np.eye(n)
    for bad in (30, 150, 290):
        A2 = A.copy()
        A2[bad, bad] = -1.0
        T = sp.csc_matrix(np.tril(A2))
        As = problems.SymMatrix(n=n, p=T.indptr.astype(np.int64), i=T.indices.astype(np.int64),
                                x=T.data.astype(np.float64))
        R = ref_lib.RefFactor(As, ref_lib.NATURAL)
        sym = R.symbolic()
        assert R.factorize() == 1
        Lref = R.Lx()
        print(f"case bad={bad} ref minor={R.minor}", flush=True)
        for prec in ("fp64", "fp32"):
            cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision=prec)
            L = cholmod.analyze_p(As, sym["Perm"], cm)
            L._numeric(cm)
            if prec == "fp32":
                L.set_pivot_rho(0.0)
            assert cholmod.factorize(As, L, cm)
            assert cm.status == cholmod.CHOLMOD_NOT_POSDEF
            assert L.minor == R.minor
            err = np.abs(L.x - Lref).max() / np.abs(Lref).max()
            assert err <= (1e-12 if prec == "fp64" else 1e-4), err


def test_device_resident_inputs(gpu, ref_lib):
    import torch
    A = problems.stencil3d_27pt(10, seed=8)
    R = ref_lib.RefFactor(A, ref_lib.METIS)
    sym = R.symbolic()
    R.factorize()
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision="fp64")
    L = cholmod.analyze_p(A, sym["Perm"], cm)
    ax = torch.as_tensor(A.x, device="cuda:0")
    assert cholmod.factorize(A, L, cm, values=ax)
    B = np.random.default_rng(4).uniform(-1, 1, (A.n, 3))
    Bd = torch.as_tensor(np.asfortranarray(B), device="cuda:0").t().contiguous().t()
    X = cholmod.solve(cholmod.CHOLMOD_A, L, Bd, cm)
    Xh = X.cpu().numpy()
    Xr = R.solve(B)
    assert np.linalg.norm(Xh - Xr) <= 1e-10 * np.linalg.norm(Xr)
    # L \ b and L' \ b in the permuted order against the oracle restatement
    from oracle import restate
    Lx = L.x
    Y = cholmod.solve(cholmod.CHOLMOD_L, L, B, cm)
    Yo = restate.lsolve(sym, Lx, B)
    assert np.linalg.norm(Y - Yo) <= 1e-12 * np.linalg.norm(Yo)
    Z = cholmod.solve(cholmod.CHOLMOD_Lt, L, B, cm)
    Zo = restate.ltsolve(sym, Lx, B)
    assert np.linalg.norm(Z - Zo) <= 1e-12 * np.linalg.norm(Zo)


def test_rcond_matches_reference_definition(gpu, ref_lib):
    """rcond = (min diag L / max diag L)^2 as cholmod_l_rcond computes it, on the
    reference's own factor; 0 for a not-positive-definite factorization."""
    A = problems.grid2d_lognormal(30, np.random.default_rng(9))
    R = ref_lib.RefFactor(A, ref_lib.AMD)
    sym = R.symbolic()
    R.factorize()
    Lref = R.Lx()
    d = np.concatenate([Lref[sym["Lpx"][s] + np.arange(sym["Super"][s + 1] - sym["Super"][s]) *
                             (sym["Lpi"][s + 1] - sym["Lpi"][s] + 1)] for s in range(len(sym["Super"]) - 1)])
    expect = (d.min() / d.max()) ** 2
    # FP32: the definition is checked exactly on the device factor's own
    # diagonal, the value against the reference to FP32-factor accuracy (1e-4,
    # the tolerance of every FP32 factor entry in these tests)
    for prec, tol in (("fp64", 1e-12), ("fp32", 1e-4)):
        cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision=prec)
        L = cholmod.analyze_p(A, sym["Perm"], cm)
        assert cholmod.factorize(A, L, cm)
        rc = cholmod.rcond(L, cm)
        assert abs(rc - expect) <= tol * expect, (prec, rc, expect)
        Lx = L.x
        dd = np.concatenate([Lx[sym["Lpx"][s] + np.arange(sym["Super"][s + 1] - sym["Super"][s]) *
                                (sym["Lpi"][s + 1] - sym["Lpi"][s] + 1)] for s in range(len(sym["Super"]) - 1)])
        own = (dd.min() / dd.max()) ** 2
        assert abs(rc - own) <= 1e-6 * own, (prec, rc, own)
    x = A.x.copy()
    x[A.p[:-1][10]] = -1.0
    A2 = problems.SymMatrix(n=A.n, p=A.p, i=A.i, x=x)
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN)
    L = cholmod.analyze_p(A2, sym["Perm"], cm)
    assert cholmod.factorize(A2, L, cm) and cm.status == cholmod.CHOLMOD_NOT_POSDEF
    assert cholmod.rcond(L, cm) == 0.0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_deterministic_factor_is_bitwise_repeatable(gpu, ref_lib, precision):
    """MXC_OPT_DETERMINISTIC: descendant updates in rounds (one pair per target
    supernode per launch, in descendant order), so every entry of L receives
    its updates in a fixed order.  Factorizations repeated from the captured
    graph and eagerly give the same L bit for bit, and it is the reference's
    factor to the usual tolerance."""
    A = problems.grid3d_lognormal(22, np.random.default_rng(9))
    R = ref_lib.RefFactor(A, ref_lib.AMD)
    sym = R.symbolic()
    assert R.factorize() == 0
    Lref = R.Lx()
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision=precision)
    L = cholmod.analyze_p(A, sym["Perm"], cm)
    L._numeric(cm)
    L.set_deterministic(True)
    xs = []
    for graphs in (True, True, True, False):
        L.set_graphs(graphs)
        assert cholmod.factorize(A, L, cm) and cm.status == 0
        xs.append(L.x)
    for x in xs[1:]:
        assert np.array_equal(x, xs[0])
    err = np.abs(xs[0] - Lref).max() / np.abs(Lref).max()
    print(f"{precision}: deterministic factor vs reference {err:.3e}", flush=True)
    assert err <= (1e-12 if precision == "fp64" else 1e-4)
    # the default (atomic) assembly gives the same factor to rounding
    L.set_deterministic(False)
    L.set_graphs(True)
    assert cholmod.factorize(A, L, cm) and cm.status == 0
    d = np.abs(L.x - xs[0]).max() / np.abs(xs[0]).max()
    print(f"  atomic vs deterministic assembly: {d:.3e}", flush=True)
    assert d <= (1e-13 if precision == "fp64" else 1e-5)


@pytest.mark.parametrize("kind", ["diffusion3d_30_metis", "stencil27_18_amd"])
def test_folded_trsm_is_bitwise_the_separate_trsm(gpu, ref_lib, kind):
    """MXC_OPT_FOLD_TRSM (FP64): the trailing launch whose tile (0, 0) factors
    the next diagonal block also applies that block's TRSM in its first tile
    column, after waiting for the block's inverse.  Its MFMA sequence is the
    TRSM tile's, so with the deterministic assembly the factor is bit for bit
    the one with separate TRSM launches, replayed or eager; the fold takes
    launches away; and the factor is the reference's to 1e-12."""
    from suitesparse_amd import _lib
    if kind.startswith("diffusion"):
        A = problems.diffusion3d_7pt(30, seed=12)
        ordering = cholmod.CHOLMOD_METIS
    else:
        A = problems.stencil3d_27pt(18, seed=13)
        ordering = cholmod.CHOLMOD_AMD
    print(f"case {kind} n={A.n}", flush=True)
    cm = cholmod.Common(nmethods=1, ordering=ordering, precision="fp64")
    L = cholmod.analyze(A, cm)
    assert L is not None
    L._numeric(cm)
    L.set_deterministic(True)

    def trsm_launches():
        counts, ms = np.zeros(8, dtype=np.int64), np.zeros(4)
        _lib.check("mxc_factor_stats", _lib.lib().mxc_factor_stats(L._num, cholmod._ptr(counts),
                                                                    cholmod._ptr(ms)))
        return int(counts[2])

    xs, launches = {}, {}
    for fold in (False, True):
        L.set_fold_trsm(fold)
        launches[fold] = trsm_launches()
        runs = []
        for graphs in (True, True, False):
            L.set_graphs(graphs)
            assert cholmod.factorize(A, L, cm) and cm.status == 0
            runs.append(L.x)
        for x in runs[1:]:
            assert np.array_equal(x, runs[0])
        xs[fold] = runs[0]
    print(f"  TRSM launches: separate {launches[False]}, folded {launches[True]}", flush=True)
    assert launches[True] < launches[False]
    assert np.array_equal(xs[True], xs[False])
    R = ref_lib.RefFactor(A, ref_lib.GIVEN, perm=L.Perm)
    assert R.factorize() == 0
    Lref = R.Lx()
    err = np.abs(xs[True] - Lref).max() / np.abs(Lref).max()
    print(f"  folded factor vs reference {err:.3e}", flush=True)
    assert err <= 1e-12
