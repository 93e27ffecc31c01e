"""The tile autotuner (mxc_autotune): it times the 32/64/128 tile edges on a
sample of each (rows, cols, depth) bucket's real blocks, installs the fastest
per bucket and rebuilds the launch lists.  Checked: the table it returns holds
only valid edges and is what mxc_get_tiles reads back; a saved table installs
through mxc_set_tiles; and the factor after tuning (FP64 and FP32) is the
reference's, to the same tolerances as without tuning (the edge changes the
summation order only)."""
import numpy as np
import pytest

from suitesparse_amd import cholmod, problems

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_autotune_then_factor_matches_reference(gpu, ref_lib, precision):
    A = problems.grid3d_lognormal(24, np.random.default_rng(5))
    R = ref_lib.RefFactor(A, ref_lib.AMD)
    sym = R.symbolic()
    assert R.factorize() == 0
    Lref = R.Lx()
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision=precision)
    L = cholmod.analyze_p(A, sym["Perm"], cm)
    L._numeric(cm)
    table = L.autotune(reps=1)
    print(precision, "edges used:", {e: table.count(e) for e in set(table)}, flush=True)
    assert len(table) == 225 and set(table) <= {32, 64, 128}
    assert L.tiles() == table
    assert cholmod.factorize(A, L, cm) and cm.status == 0
    err = np.abs(L.x - Lref).max() / np.abs(Lref).max()
    print(f"  L vs reference after tuning: {err:.3e}", flush=True)
    assert err <= (1e-12 if precision == "fp64" else 1e-4)
    B = np.random.default_rng(6).uniform(-1, 1, (A.n, 2))
    if precision == "fp64":
        X = cholmod.solve(cholmod.CHOLMOD_A, L, B, cm)
        Xr = R.solve(B)
        assert np.linalg.norm(X - Xr) <= 1e-10 * np.linalg.norm(Xr)
    else:
        X, info = cholmod.solve_refine(L, B, cm, tol=1e-10, maxiter=20)
        assert info["relres"] <= 1e-10
    # a saved table installs and reproduces the factor
    other = [64 if e == 32 else 32 if e == 64 else e for e in table]
    L.set_tiles(other)
    assert L.tiles() == other
    assert cholmod.factorize(A, L, cm) and cm.status == 0
    err2 = np.abs(L.x - Lref).max() / np.abs(Lref).max()
    assert err2 <= (1e-12 if precision == "fp64" else 1e-4)
