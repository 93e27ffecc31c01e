"""Batched factorization: the block-diagonal batch's symbolic factor is the
concatenation of each system's own reference analysis (CPU), and on the GPU
every system's solution matches the reference's for that system alone."""
import numpy as np
import pytest

from suitesparse_amd import batch, cholmod, problems


def _systems(k, seed, lo=6, hi=14):
    rng = np.random.default_rng(seed)
    sides = rng.integers(lo, hi + 1, size=k)
    return [problems.grid2d_lognormal(int(s), rng, shift=1e-3) for s in sides]


def test_batch_symbolic_is_concatenation(ref_lib):
    mats = _systems(12, 0)
    cm = cholmod.Common()
    b = batch.analyze_batch(mats, cm)
    assert b is not None and b.orderings <= 9
    L = b.L
    sup_off = 0
    lpi_base = 0
    lpx_base = 0
    for j, A in enumerate(mats):
        R = ref_lib.RefFactor(A, ref_lib.AMD)
        sym = R.symbolic()
        off = b.offsets[j]
        ns = R.nsuper
        assert np.array_equal(L.super[sup_off:sup_off + ns + 1], sym["Super"] + off)
        assert np.array_equal(L.pi[sup_off:sup_off + ns + 1], sym["Lpi"] + lpi_base)
        assert np.array_equal(L.px[sup_off:sup_off + ns + 1], sym["Lpx"] + lpx_base)
        assert np.array_equal(L.s[lpi_base:lpi_base + R.ssize], sym["Ls"] + off)
        assert np.array_equal(L.Perm[b.system_slice(j)], sym["Perm"] + off)
        sup_off += ns
        lpi_base += R.ssize
        lpx_base += R.xsize


@pytest.mark.gpu
def test_batch_solve_matches_reference(gpu, ref_lib):
    mats = _systems(40, 1, lo=8, hi=40)
    for prec in ("fp64", "fp32"):
        cm = cholmod.Common(precision=prec)
        b = batch.analyze_batch(mats, cm)
        assert batch.factorize_batch(b), cm.status
        B = np.random.default_rng(2).uniform(-1, 1, (b.A.n, 4))
        if prec == "fp64":
            X = batch.solve_batch(b, B)
        else:
            X, info = batch.solve_refine_batch(b, B, tol=1e-10)
            print(f"batch fp32 IR {info}", flush=True)
            assert info["relres"] <= 1e-10
        worst_fe, worst_rr = 0.0, 0.0
        for j, A in enumerate(mats):
            sl = b.system_slice(j)
            R = ref_lib.RefFactor(A, ref_lib.AMD)
            R.factorize()
            Xr = R.solve(B[sl])
            fe = np.linalg.norm(X[sl] - Xr) / np.linalg.norm(Xr)
            rr = np.max(np.linalg.norm(B[sl] - A.matvec(X[sl]), axis=0) / np.linalg.norm(B[sl], axis=0))
            worst_fe, worst_rr = max(worst_fe, fe), max(worst_rr, rr)
        print(f"batch {prec}: worst fwd err {worst_fe:.2e}, worst relres {worst_rr:.2e}", flush=True)
        assert worst_rr <= 1e-10
        if prec == "fp64":
            assert worst_fe <= 1e-10
