"""GPU parity at the benchmark workloads' full sizes, against the reference.

The reference CHOLMOD (oracle/_ref, compiled from the reference sources, MKL
BLAS, FP64) factors and solves the same seeded inputs on the host cores:

* 3D 7-point diffusion 64^3 (n = 262,144), METIS, FP64, 1 RHS: the product's
  own METIS ordering and analysis equal the reference's (Perm and the whole
  symbolic factor); every entry of
  L->x is within 1e-12 * max|L| of the reference's; ||x - x_ref|| / ||x_ref||
  <= 1e-10.
* 3D 27-point 96^3 (n = 884,736), METIS, 4 RHS: FP64 as above (L compared in
  chunks); FP32 factor + FP64 refinement reaches relative residual <= 1e-10,
  and its forward error against the reference's FP64 x_ref is stated and held
  to kappa_bound * relres, where kappa_bound = 2 max(a_ii) / 1e-2 bounds the
  condition number (A = S + 1e-2 I with S diagonally dominant with nonnegative
  diagonal, so lambda_min >= 1e-2; Gershgorin gives lambda_max <= 2 max a_ii).
* 4,096 independent 2D systems (side U{16..128}), AMD, 4 RHS: every system's
  relative residual <= 1e-10 in FP64 and in FP32 + refinement; forward error
  against the reference (each system alone) on a seeded sample of 64 systems,
  <= 1e-10 for FP64.
* 64^3 again: a rank-8 update of the resident factor (cholmod_l_updown) equals
  the oracle's bit for bit, with the reference's modfl and updated solution.
* cholmod_l_factorize_p with beta != 0 (A + beta I), including a matrix that is
  not positive definite until beta is added and one that stays indefinite
  (status, L->minor and L against the reference).
"""
import os

import numpy as np
import pytest

from suitesparse_amd import batch, cholmod, problems

pytestmark = pytest.mark.gpu


def _threads():
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return os.cpu_count() or 1


def _max_rel_diff_chunked(a, b, chunk=1 << 26):
    scale = 0.0
    diff = 0.0
    for k in range(0, a.size, chunk):
        aa, bb = a[k:k + chunk], b[k:k + chunk]
        scale = max(scale, float(np.abs(bb).max()))
        diff = max(diff, float(np.abs(aa - bb).max()))
    return diff / scale


def _relres_cols(A, B, X):
    R = B - A.matvec(X)
    return np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)


def test_diffusion3d_64_fp64_full(gpu, ref_lib):
    A, B, xt = problems.workload_diffusion3d_64()
    print(f"case diffusion3d_64 n={A.n}", flush=True)
    ref_lib.set_threads(_threads())
    R = ref_lib.RefFactor(A, ref_lib.METIS)
    sym = R.symbolic()
    assert R.factorize() == 0
    Xref = R.solve(B)
    Lref = R.Lx()
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_METIS, precision="fp64")
    L = cholmod.analyze(A, cm)
    assert L is not None, cm.status
    for k_ref, k_own in (("Perm", "Perm"), ("Super", "super"), ("Lpi", "pi"), ("Lpx", "px"), ("Ls", "s")):
        assert np.array_equal(sym[k_ref], getattr(L, k_own)), k_ref
    assert L.fl == R.fl and L.lnz == R.lnz
    assert cholmod.factorize(A, L, cm) and cm.status == 0 and L.minor == A.n
    err = _max_rel_diff_chunked(L.x, Lref)
    print(f"  L max rel err {err:.3e}", flush=True)
    assert err <= 1e-12
    del Lref
    X = cholmod.solve(cholmod.CHOLMOD_A, L, B, cm)
    fe = np.linalg.norm(X - Xref) / np.linalg.norm(Xref)
    fe_true = np.linalg.norm(X - xt) / np.linalg.norm(xt)
    rr = _relres_cols(A, B, X).max()
    print(f"  fwd err vs ref {fe:.3e}, vs x_true {fe_true:.3e}, relres {rr:.3e}", flush=True)
    assert fe <= 1e-10 and rr <= 1e-12
    # a rank-8 update of the resident factor (cholmod_l_updown): bit for bit the
    # oracle's on the same values, modfl and the updated solution as the reference's
    from oracle import restate
    from updown_cases import pattern_columns, sym_of
    rng = np.random.default_rng(31)
    Cp, Ci, Cx = pattern_columns(sym_of(L), rng.integers(0, A.n, 8), rng, dens=0.3)
    o = restate.updown(sym_of(L), L.x, True, Cp, Ci, Cx)
    assert cholmod.updown(True, (Cp, Ci, Cx), L, cm) and cm.modfl == o["modfl"]
    assert np.array_equal(L.x, o["Lx"])
    del o
    r = R.updown(True, Cp, Ci, Cx)
    assert r["modfl"] == cm.modfl
    Xu = cholmod.solve(cholmod.CHOLMOD_A, L, B, cm)
    Xu_ref = R.solve(B)
    fe_u = np.linalg.norm(Xu - Xu_ref) / np.linalg.norm(Xu_ref)
    print(f"  rank-8 update: modfl {cm.modfl:.6g}, fwd err vs ref {fe_u:.3e}, reference updown "
          f"{r['t']:.3f} s", flush=True)
    assert fe_u <= 1e-10
    L.free()


def test_stencil27_96_fp64_and_fp32ir_full(gpu, ref_lib):
    A, B, _ = problems.workload_stencil27_96()
    print(f"case stencil27_96 n={A.n} nrhs={B.shape[1]}", flush=True)
    # the product's own METIS ordering; the reference analyzes it as given
    # (its own METIS yields the same permutation: test_symbolic.py checks the
    # digests of this workload)
    cmo = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_METIS)
    Lo = cholmod.analyze(A, cmo)
    assert Lo is not None and Lo.ordering == cholmod.CHOLMOD_METIS
    perm = Lo.Perm.copy()
    Lo.free()
    ref_lib.set_threads(_threads())
    R = ref_lib.RefFactor(A, ref_lib.GIVEN, perm=perm)
    sym = R.symbolic()
    assert R.factorize() == 0
    Xref = R.solve(B)
    print(f"  reference factor {R.t_factor:.1f} s", flush=True)

    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision="fp64")
    L = cholmod.analyze_p(A, perm, cm)
    assert L is not None, cm.status
    assert np.array_equal(L.s, sym["Ls"]) and np.array_equal(L.px, sym["Lpx"])
    assert cholmod.factorize(A, L, cm) and cm.status == 0
    Lx = L.x
    Lref = R.Lx()
    err = _max_rel_diff_chunked(Lx, Lref)
    del Lx, Lref
    print(f"  FP64 L max rel err {err:.3e}", flush=True)
    assert err <= 1e-12
    X64 = cholmod.solve(cholmod.CHOLMOD_A, L, B, cm)
    fe64 = np.linalg.norm(X64 - Xref) / np.linalg.norm(Xref)
    rr64 = _relres_cols(A, B, X64).max()
    print(f"  FP64 fwd err vs ref {fe64:.3e}, relres {rr64:.3e}", flush=True)
    assert fe64 <= 1e-10
    L.free()
    del R

    cm32 = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision="fp32")
    L32 = cholmod.analyze_p(A, perm, cm32)
    assert cholmod.factorize(A, L32, cm32) and cm32.status == 0
    X32, info = cholmod.solve_refine(L32, B, cm32, tol=1e-10, maxiter=10)
    rr32 = _relres_cols(A, B, X32).max()
    fe32 = np.linalg.norm(X32 - Xref) / np.linalg.norm(Xref)
    col = np.repeat(np.arange(A.n), np.diff(A.p))
    dmax = float(A.x[A.i == col].max())
    kappa_bound = 2.0 * dmax / 1e-2
    print(f"  FP32+IR {info}; relres {rr32:.3e}, fwd err vs FP64 ref {fe32:.3e} "
          f"(kappa bound {kappa_bound:.3e})", flush=True)
    assert info["relres"] <= 1e-10 and rr32 <= 1e-10
    assert not info["fp64_fallback"]
    assert fe32 <= kappa_bound * rr32
    L32.free()


def test_batch4096_full(gpu, ref_lib):
    sides, rng = problems.batch_sides(4096)
    mats = [problems.grid2d_lognormal(int(s), rng) for s in sides]
    nt = sum(m.n for m in mats)
    B = np.asfortranarray(rng.uniform(-1.0, 1.0, (nt, 4)))
    print(f"case batch4096 n_total={nt}", flush=True)
    sample = np.sort(np.random.default_rng(31).choice(len(mats), 64, replace=False))
    ref_lib.set_threads(min(8, _threads()))
    refs = {}
    for j in sample:
        R = ref_lib.RefFactor(mats[j], ref_lib.AMD)
        assert R.factorize() == 0
        refs[int(j)] = R
    cm = cholmod.Common(precision="fp64")
    b = batch.analyze_batch(mats, cm)
    assert b is not None
    seg = b.offsets
    for prec in ("fp64", "fp32"):
        cmp = cholmod.Common(precision=prec)
        b.cm = cmp
        assert batch.factorize_batch(b) and cmp.status == 0
        if prec == "fp64":
            X = batch.solve_batch(b, B)
        else:
            X, info = batch.solve_refine_batch(b, B, tol=1e-10, maxiter=20)
            print(f"  fp32 IR {info}", flush=True)
        Rm = B - b.A.matvec(X)
        rn = np.sqrt(np.add.reduceat(Rm * Rm, seg[:-1], axis=0))
        bn = np.sqrt(np.add.reduceat(B * B, seg[:-1], axis=0))
        rel = rn / bn
        worst_fe = 0.0
        for j, R in refs.items():
            sl = b.system_slice(j)
            Xr = R.solve(np.ascontiguousarray(B[sl]))
            worst_fe = max(worst_fe, float(np.linalg.norm(X[sl] - Xr) / np.linalg.norm(Xr)))
        print(f"  {prec}: worst per-system relres {rel.max():.3e}, worst fwd err vs ref "
              f"(64 systems) {worst_fe:.3e}", flush=True)
        assert rel.max() <= 1e-10
        if prec == "fp64":
            assert worst_fe <= 1e-10
        else:
            assert worst_fe <= 1e-8
        del X, Rm
    b.L.free()


def _shifted(A, sigma):
    col = np.repeat(np.arange(A.n), np.diff(A.p))
    x = A.x.copy()
    x[A.i == col] -= sigma
    return problems.SymMatrix(n=A.n, p=A.p, i=A.i, x=x, name=A.name + f"_m{sigma:g}")


@pytest.mark.parametrize("case", ["spd_beta", "indefinite_made_spd", "still_indefinite",
                                  "grid3d_beta", "dense_beta"])
def test_factorize_p_beta(case, gpu, ref_lib):
    """cholmod_l_factorize_p(A, beta, ...) factors A + beta*I
    (CHOLMOD/Cholesky/cholmod_factorize.c:93-418): status, minor, L and x
    against the reference for beta != 0."""
    import scipy.sparse as sp
    ordering = ref_lib.AMD
    if case == "spd_beta":
        A, beta = problems.grid2d_lognormal(60, np.random.default_rng(41)), 0.75
    elif case == "indefinite_made_spd":
        # 5-point Laplacian has lambda_min ~ 2 pi^2 / side^2; shift it below 0, beta lifts it back
        A, beta = _shifted(problems.laplace2d_5pt(40), 0.5), 0.6
    elif case == "still_indefinite":
        A, beta = _shifted(problems.laplace2d_5pt(40), 1.0), 0.3
    elif case == "grid3d_beta":
        A, beta = problems.grid3d_lognormal(20, np.random.default_rng(43)), 2.5
        ordering = ref_lib.METIS
    else:
        n = 300
        M = np.random.default_rng(44).standard_normal((n, n))
        D = M @ M.T
        T = sp.csc_matrix(np.tril(D))
        A = problems.SymMatrix(n=n, p=T.indptr.astype(np.int64), i=T.indices.astype(np.int64),
                               x=T.data.astype(np.float64), name="dense300")
        beta, ordering = 40.0, ref_lib.NATURAL
    print(f"case {case} n={A.n} beta={beta}", flush=True)
    R = ref_lib.RefFactor(A, ordering)
    sym = R.symbolic()
    st = R.factorize_p(beta)
    Lref = R.Lx()
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision="fp64")
    L = cholmod.analyze_p(A, sym["Perm"], cm)
    assert cholmod.factorize_p(A, [beta, 0.0], L, cm)
    print(f"  ref status {st} minor {R.minor}; gpu status {cm.status} minor {L.minor}", flush=True)
    assert cm.status == st and L.minor == R.minor
    err = np.abs(L.x - Lref).max() / np.abs(Lref).max()
    print(f"  L max rel err {err:.3e}", flush=True)
    assert err <= 1e-12
    if st == 0:
        Bm = np.random.default_rng(5).uniform(-1, 1, (A.n, 2))
        X = cholmod.solve(cholmod.CHOLMOD_A, L, Bm, cm)
        Xr = R.solve(Bm)
        fe = np.linalg.norm(X - Xr) / np.linalg.norm(Xr)
        print(f"  x fwd err vs ref {fe:.3e}", flush=True)
        assert fe <= 1e-10
