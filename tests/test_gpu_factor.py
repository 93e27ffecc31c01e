"""GPU parity of the supernodal factorization and solve against the reference.

The reference CHOLMOD (oracle/_ref, compiled from the reference sources, MKL
BLAS, FP64) supplies the ordering, the factor L->x and x_ref on the same
seeded inputs.  Bars: FP64 factor entries within 1e-12 * max|L| of the
reference's (summation order differs, so not bitwise); ||x - x_ref|| / ||x_ref||
<= 1e-10; FP32 factor within 1e-4 * max|L|, and FP32 + FP64 refinement reaches
relative residual <= 1e-10.
"""
import numpy as np
import pytest

from suitesparse_amd import cholmod, problems

pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(11)
    return [
        ("lap2d_100_amd", lambda: problems.laplace2d_5pt(100), 1),
        ("grid3d_16_metis", lambda: problems.grid3d_lognormal(16, np.random.default_rng(7)), 2),
        ("st27_12_metis", lambda: problems.stencil3d_27pt(12, seed=5), 2),
        ("grid2d_40_amd", lambda: problems.grid2d_lognormal(40, rng), 1),
        ("dense_150", lambda: _dense(150, 3), 3),
        ("dense_600", lambda: _dense(600, 6), 3),   # 3 panels: merged second-stream update
        ("tiny_1", lambda: _dense(1, 4), 3),
        ("tiny_3", lambda: _dense(3, 4), 3),
    ]


def _dense(n, seed):
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n))
    A = M @ M.T + n * np.eye(n)
    T = sp.csc_matrix(np.tril(A))
    return problems.SymMatrix(n=n, p=T.indptr.astype(np.int64), i=T.indices.astype(np.int64),
                              x=T.data.astype(np.float64), name=f"dense_{n}")


@pytest.mark.parametrize("name,make,ordering", _cases(), ids=[c[0] for c in _cases()])
def test_factor_solve_fp64(name, make, ordering, gpu, ref_lib):
    A = make()
    print(f"case {name} n={A.n}", flush=True)
    R = ref_lib.RefFactor(A, ordering)
    sym = R.symbolic()
    assert R.factorize() == 0
    Lref = R.Lx()
    rng = np.random.default_rng(1)
    B = rng.uniform(-1, 1, (A.n, 2))
    Xref = R.solve(B)

    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision="fp64")
    L = cholmod.analyze_p(A, sym["Perm"], cm)
    assert L is not None and cm.status == 0
    assert np.array_equal(L.s, sym["Ls"]) and np.array_equal(L.px, sym["Lpx"])
    assert cholmod.factorize(A, L, cm), cm.status
    assert cm.status == cholmod.CHOLMOD_OK and L.minor == A.n
    Lx = L.x
    scale = np.abs(Lref).max()
    err = np.abs(Lx - Lref).max() / scale
    print(f"  L max rel err {err:.3e}", flush=True)
    assert err <= 1e-12
    X = cholmod.solve(cholmod.CHOLMOD_A, L, B, cm)
    fe = np.linalg.norm(X - Xref) / np.linalg.norm(Xref)
    print(f"  x fwd err vs ref {fe:.3e}", flush=True)
    assert fe <= 1e-10


@pytest.mark.parametrize("name,make,ordering", _cases()[:4], ids=[c[0] for c in _cases()[:4]])
def test_factor_fp32_refine(name, make, ordering, gpu, ref_lib):
    A = make()
    print(f"case {name} fp32 n={A.n}", flush=True)
    R = ref_lib.RefFactor(A, ordering)
    sym = R.symbolic()
    R.factorize()
    Lref = R.Lx()
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision="fp32")
    L = cholmod.analyze_p(A, sym["Perm"], cm)
    assert cholmod.factorize(A, L, cm), cm.status
    err = np.abs(L.x - Lref).max() / np.abs(Lref).max()
    print(f"  L32 max rel err {err:.3e}", flush=True)
    assert err <= 1e-4
    B = np.random.default_rng(2).uniform(-1, 1, (A.n, 3))
    X, info = cholmod.solve_refine(L, B, cm, tol=1e-10, maxiter=10)
    print(f"  IR {info}", flush=True)
    assert info["relres"] <= 1e-10
    r = B - A.matvec(X)
    assert np.linalg.norm(r, axis=0).max() / np.linalg.norm(B, axis=0).min() <= 1e-9


def _near_singular_pairs(m, delta, seed=0):
    """Block-diagonal SPD matrix of 2x2 blocks [[1, 1-delta], [1-delta, 1]] (pivot
    ~2*delta: cond ~ 1/delta), weakly coupled along a chain and symmetrically scaled."""
    import scipy.sparse as sp
    n = 2 * m
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for b in range(m):
        i = 2 * b
        rows += [i, i + 1, i + 1]
        cols += [i, i, i + 1]
        vals += [1.0, 1.0 - delta, 1.0]
        if b + 1 < m:                       # tiny coupling to the next block
            rows.append(i + 2)
            cols.append(i)
            vals.append(1e-3 * delta)
    T = sp.csc_matrix((vals, (rows, cols)), shape=(n, n))
    d = np.exp(rng.uniform(-1, 1, n))
    T = sp.diags(d) @ T @ sp.diags(d)
    T = sp.tril(T).tocsc()
    T.sort_indices()
    return problems.SymMatrix(n=n, p=T.indptr.astype(np.int64), i=T.indices.astype(np.int64),
                              x=T.data, name=f"pairs{m}")


def test_fp32_breakdown_refactored_in_fp64(gpu):
    """Near-singular 2x2 pairs (true pivot ratio ~1e-9 < rho = 1e-6; in FP32
    the computed pivots scatter around 0, and 1 - 1e-9 itself rounds to 1):
    every breaking block is refactored in FP64 on the device from A's FP64
    values, and refinement through the FP32 factor then reaches 1e-10 with no
    whole-matrix FP64 factor."""
    A = _near_singular_pairs(200, 1e-9)
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision="fp32")
    L = cholmod.analyze_p(A, np.arange(A.n), cm)
    assert cholmod.factorize(A, L, cm) and cm.status == 0 and L.minor == A.n
    inf = L.info()
    print(inf, flush=True)
    assert inf["fp64_blocks"] > 0 and inf["escalated_fp64"] == 0
    # B = A X_true with X_true = O(1): then ||x|| ~ ||b||, so a backward-stable
    # solve reaches a small residual (a random B instead has ||x|| ~ 1e9 ||b||)
    Xt = np.random.default_rng(1).uniform(-1, 1, (A.n, 2))
    B = np.asfortranarray(A.matvec(Xt))
    X, info = cholmod.solve_refine(L, B, cm, tol=1e-10, maxiter=10)
    print(info, flush=True)
    assert not info["fp64_fallback"]
    assert info["relres"] <= 1e-10
    R = B - A.matvec(X)
    assert np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)) <= 1e-10


def test_fp32_without_redo_falls_back_to_fp64(gpu):
    """The same matrix with the FP64 redo off (rho = 0, CHOLMOD's semantics in
    FP32): either the FP32 factorization meets a pivot <= 0 and reports it, or
    it completes with pivots that carry no accuracy, refinement through it
    stalls, and mxc_solve_refine builds, factors and refines with an FP64
    factor of the same values."""
    A = _near_singular_pairs(200, 1e-9)
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision="fp32")
    L = cholmod.analyze_p(A, np.arange(A.n), cm)
    L._numeric(cm)
    L.set_pivot_rho(0.0)
    assert cholmod.factorize(A, L, cm)
    inf = L.info()
    print(cm.status, L.minor, inf, flush=True)
    assert inf["fp64_blocks"] == 0
    if cm.status == cholmod.CHOLMOD_NOT_POSDEF:
        assert L.minor < A.n
        return
    Xt = np.random.default_rng(1).uniform(-1, 1, (A.n, 2))
    B = np.asfortranarray(A.matvec(Xt))
    X, info = cholmod.solve_refine(L, B, cm, tol=1e-10, maxiter=5)
    print(info, flush=True)
    assert info["fp64_fallback"] and info["relres"] <= 1e-10


def _pair_blocks(sizes, delta, seed):
    """Block-diagonal SPD matrix of dense blocks of the given sizes, each built
    from near-singular 2x2 pairs [[1, 1-delta], [1-delta, 1]] (an odd size ends
    with a 1 on the diagonal) plus a dense symmetric coupling of norm ~1e-3 delta
    inside the block, symmetrically scaled: true pivot ratios ~delta, and in
    FP32 1 - delta rounds to 1, so FP32 pivots carry no information."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    blocks = []
    for m in sizes:
        B = np.eye(m)
        for k in range(0, m - 1, 2):
            B[k + 1, k] = B[k, k + 1] = 1.0 - delta
        E = rng.standard_normal((m, m))
        B += 1e-3 * delta / m * (E + E.T)
        d = np.exp(rng.uniform(-1, 1, m))
        blocks.append(d[:, None] * B * d[None, :])
    T = sp.tril(sp.block_diag(blocks, format="csc")).tocsc()
    T.sort_indices()
    return problems.SymMatrix(n=T.shape[0], p=T.indptr.astype(np.int64), i=T.indices.astype(np.int64),
                              x=T.data.astype(np.float64), name=f"pairs{len(sizes)}")


@pytest.mark.parametrize("case", ["dense64", "dense200", "small12", "tiny5"])
def test_fp32_breakdown_refactors_block_in_fp64(gpu, case):
    """Dense blocks of near-singular pairs (pivot ratios ~1e-9 < rho): each
    breaking diagonal block is refactored in FP64 on the device and stored in
    FP32, the factorization continues in FP32 with no whole-matrix FP64 factor,
    and refinement through it reaches 1e-10.  Covered: a 64-column block
    (k_potrf_diag), a 200-column supernode (steps 1..3 factored as tails of the
    trailing launches), blocks of 12 columns (k_potrf_small) and of 5
    (k_potrf_tiny).  Against the FP64 factorization of the same matrix, the
    FP32 factor agrees to FP32 rounding of L's entries."""
    sizes = {"dense64": [64], "dense200": [200], "small12": [12] * 24, "tiny5": [5] * 40}[case]
    A = _pair_blocks(sizes, 1e-9, seed=len(sizes))
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision="fp32")
    L = cholmod.analyze_p(A, np.arange(A.n), cm)
    assert cholmod.factorize(A, L, cm) and cm.status == 0 and L.minor == A.n
    inf = L.info()
    print(case, inf, flush=True)
    assert inf["fp64_blocks"] > 0 and inf["escalated_fp64"] == 0
    Xt = np.random.default_rng(3).uniform(-1, 1, (A.n, 2))
    B = np.asfortranarray(A.matvec(Xt))
    X, info = cholmod.solve_refine(L, B, cm, tol=1e-10, maxiter=10)
    print(info, flush=True)
    assert not info["fp64_fallback"]
    assert info["relres"] <= 1e-10
    R = B - A.matvec(X)
    assert np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)) <= 1e-10
    cm64 = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision="fp64")
    L64 = cholmod.analyze_p(A, np.arange(A.n), cm64)
    assert cholmod.factorize(A, L64, cm64) and cm64.status == 0
    err = np.abs(L.x - L64.x).max() / np.abs(L64.x).max()
    print(f"  FP32 factor vs FP64 factor: {err:.3e}", flush=True)
    assert err <= 1e-5
