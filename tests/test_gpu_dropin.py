"""CHOLMOD's own API through the MI355X drop-in, against the host CHOLMOD (the
reference compiled from its sources) in the same process: the caller
tests/dropin/dropin_flow.c runs the reference demo's sequence
(CHOLMOD/Demo/cholmod_dl_demo.c:287-759: read, analyze, factorize three times,
rcond, solve, solve2 with reused workspaces, solve2 with a sparse right-hand
side, one step of refinement), then a rank-2 update and downdate
(cholmod_l_updown) of a fresh factor and an update that needs new fill, once
with Common->useGPU = 1 and once with 0.

Checked per matrix: the same ordering (Perm, the method CHOLMOD's default
strategy selected, its fl / lnz and the per-method lnz), status and L->minor; the
factors within 1e-12 * max|L| of each other; x within 1e-10 (relative); the
demo's residual below 1e-13 (its own pass bar is 1e-7); the Bset solutions
equal to the full solve on Xset (relative to max |x|); L left simplicial after a Bset solve, as the
reference leaves it; and the GPU path really ran (kernel launches counted in
Common->gpuNumKernelLaunches, a device pool in Common->dev_mempool) whenever
the analysis is supernodal."""
import pathlib

import numpy as np
import pytest

from test_dropin import FIX, FLOW, run_flow  # noqa: E402

pytestmark = pytest.mark.gpu

SPD = ["bcsstk01.tri", "bcsstk02.tri", "can___24.mtx", "k01up", "r5lo", "diag", "two.tri", "2diag.tri"]


def _check(v, supernodal_expected=None):
    assert v["gpu_status"] == v["host_status"], (v["gpu_status"], v["host_status"])
    assert v["gpu_minor"] == v["host_minor"]
    assert v["perm_equal"] == "1"
    # the same ordering strategy: method chosen, fill, work and what the
    # orderings left in Common (METIS included, cholmod_analyze.c:554-782)
    for k in ("ordering", "selected", "called_nd", "anz", "method_lnz", "fl", "lnz"):
        assert v[f"gpu_{k}"] == v[f"host_{k}"], (k, v[f"gpu_{k}"], v[f"host_{k}"])
    assert v["gpu_is_super"] == v["host_is_super"]
    assert v["gpu_nsuper"] == v["host_nsuper"]
    if supernodal_expected is not None:
        assert v["gpu_is_super"] == ("1" if supernodal_expected else "0")
    if v["gpu_is_super"] == "1":
        assert int(v["gpu_gpu_launches"]) > 0 and float(v["gpu_pool_bytes"]) > 0
        assert float(v["L_rel_diff"]) <= 1e-12, v["L_rel_diff"]
    assert int(v["host_gpu_launches"]) == 0
    if v["host_status"] == "0":
        assert float(v["x_rel_diff"]) <= 1e-10, v["x_rel_diff"]
        assert float(v["gpu_resid"]) <= 1e-13 and float(v["gpu_resid_ir"]) <= 1e-13
        # errors relative to max |x|; solve and solve2 agree to rounding (the
        # forward sweep adds ancestor contributions atomically, so two solves
        # need not agree bit for bit)
        assert float(v["gpu_bset_err"]) <= 1e-12
        assert float(v["gpu_solve2_err"]) <= 1e-13
        assert v["gpu_is_super_after_bset"] == v["host_is_super_after_bset"]
        rg, rh = float(v["gpu_rcond"]), float(v["host_rcond"])
        assert abs(rg - rh) <= 1e-10 * rh
    if "gpu_ud_status" in v or "host_ud_status" in v:
        # cholmod_l_updown: the device keeps L supernodal LL', the host turns it
        # simplicial LDL'; the same modfl, solutions of A + P'CC'P, and the same
        # downdate back to A.  C that needs fill goes to the host on both sides.
        assert v["gpu_ud_status"] == v["host_ud_status"] == "0"
        assert v["gpu_ud_modfl"] == v["host_ud_modfl"]
        assert v["gpu_ud_is_super"] == "1" and v["host_ud_is_super"] == "0"
        assert float(v["gpu_ud_resid"]) <= 1e-13 and float(v["xup_rel_diff"]) <= 1e-10
        assert float(v["gpu_back_err"]) <= 1e-10
        assert v["gpu_fill_status"] == v["host_fill_status"]
        if v["host_fill_status"] == "0":
            assert v["gpu_fill_modfl"] == v["host_fill_modfl"]
            assert v["gpu_fill_is_super"] == "0"
            assert float(v["gpu_fill_resid"]) <= 1e-13 and float(v["xfill_rel_diff"]) <= 1e-10


@pytest.mark.parametrize("mode", ["supernodal", "auto"])
@pytest.mark.parametrize("name", SPD)
def test_demo_flow_through_dropin(gpu, name, mode):
    assert FLOW.exists(), "tests/dropin/_build/dropin_flow not built"
    print(f"case {name} {mode}", flush=True)
    out, v = run_flow(FIX / name, *(["supernodal"] if mode == "supernodal" else []))
    print(out.stdout[-1500:], flush=True)
    assert out.returncode == 0, out.stderr[-2000:]
    _check(v, supernodal_expected=True if mode == "supernodal" else None)


@pytest.mark.parametrize("name", ["2lo.tri", "2up.tri"])
def test_not_posdef_through_dropin(gpu, name):
    print(f"case {name}", flush=True)
    out, v = run_flow(FIX / name, "supernodal")
    assert out.returncode == 0, out.stderr[-2000:]
    assert v["host_status"] == "1"
    _check(v)


def _write_mm(path, A, upper=False):
    """Matrix Market 'coordinate real symmetric' with the lower (or upper) triangle."""
    col = np.repeat(np.arange(A.n), np.diff(A.p))
    r, c = (col, A.i) if upper else (A.i, col)
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n")
        f.write(f"{A.n} {A.n} {len(A.x)}\n")
        np.savetxt(f, np.column_stack([r + 1, c + 1, A.x]), fmt="%d %d %.17g")


@pytest.mark.parametrize("kind", ["grid3d_22_lower", "grid2d_150_upper", "stencil27_22_lower"])
def test_larger_matrix_through_dropin(gpu, tmp_path, kind):
    """The default strategy stops after AMD on the first two; on the 27-point
    grid AMD predicts more than 500 flops per entry of L, so METIS runs too."""
    from suitesparse_amd import problems
    rng = np.random.default_rng(17)
    if kind.startswith("grid3d"):
        A = problems.grid3d_lognormal(22, rng)
    elif kind.startswith("stencil27"):
        A = problems.stencil3d_27pt(22, seed=18)
    else:
        A = problems.grid2d_lognormal(150, rng)
    path = tmp_path / f"{kind}.mtx"
    _write_mm(path, A, upper=kind.endswith("upper"))
    print(f"case {kind} n={A.n}", flush=True)
    out, v = run_flow(path)
    print(out.stdout[-1500:], flush=True)
    assert out.returncode == 0, out.stderr[-2000:]
    _check(v, supernodal_expected=True)
    if kind.startswith("stencil27"):
        assert v["host_called_nd"] == "1"
