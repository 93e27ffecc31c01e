"""The CHOLMOD drop-in (suitesparse_amd/libcholmod_mi355x_dropin.so) on the
CPU: it exports CHOLMOD's entry points with CHOLMOD's own names (compiled
against cholmod.h, so the prototypes are that header's), and linked in front
of a host CHOLMOD it forwards every call to the host when the GPU is not
requested (Common->useGPU = 0, cholmod_super_symbolic.c:250-289): the caller
in tests/dropin/dropin_flow.c then runs the reference demo's sequence
(cholmod_dl_demo.c:287-759) to the demo's own pass criterion without touching
a device."""
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
DROPIN = ROOT / "suitesparse_amd" / "libcholmod_mi355x_dropin.so"
FLOW = ROOT / "tests" / "dropin" / "_build" / "dropin_flow"
FIX = ROOT / "tests" / "golden" / "cholmod_matrices"
ENTRY_POINTS = ["cholmod_l_analyze", "cholmod_l_analyze_p", "cholmod_l_factorize", "cholmod_l_factorize_p",
                "cholmod_l_solve", "cholmod_l_solve2", "cholmod_l_free_factor", "cholmod_l_finish",
                "cholmod_l_gpu_stats", "cholmod_l_updown"]


def run_flow(path, *args, env=None):
    import os
    e = dict(os.environ, MKL_THREADING_LAYER="GNU")
    e.update(env or {})
    out = subprocess.run([str(FLOW), str(path), *args], capture_output=True, text=True, timeout=600, env=e)
    vals = {}
    for line in out.stdout.splitlines():
        k, _, v = line.partition(" ")
        if k and v and (k.startswith(("gpu_", "host_")) or k in ("L_rel_diff", "x_rel_diff", "xup_rel_diff", "xfill_rel_diff", "perm_equal")):
            vals[k] = v
    return out, vals


def _need(p):
    if not p.exists():
        pytest.skip(f"{p.name} not built (needs the host CHOLMOD headers and oracle/_ref)")


def test_dropin_exports_cholmod_entry_points():
    _need(DROPIN)
    out = subprocess.run(["nm", "-D", "--defined-only", str(DROPIN)], capture_output=True, text=True).stdout
    names = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    assert set(ENTRY_POINTS) <= names
    # only CHOLMOD's entry points: nothing else of the host library is shadowed
    assert {n for n in names if n.startswith("cholmod_")} == set(ENTRY_POINTS)


@pytest.mark.parametrize("name", ["bcsstk01.tri", "bcsstk02.tri", "can___24.mtx", "k01up", "r5lo"])
@pytest.mark.parametrize("mode", ["auto", "supernodal"])
def test_host_forwarding_runs_demo_flow(name, mode):
    _need(FLOW)
    args = ["host"] + (["supernodal"] if mode == "supernodal" else [])
    out, v = run_flow(FIX / name, *args)
    assert out.returncode == 0, out.stderr[-2000:]
    assert v["host_status"] == "0" and int(v["host_gpu_launches"]) == 0
    assert float(v["host_resid"]) < 1e-7 and float(v["host_resid_ir"]) < 1e-7   # the demo's criterion
    assert float(v["host_bset_err"]) < 1e-10
    if mode == "supernodal":
        assert v["host_is_super"] == "1" and v["host_is_super_after_bset"] == "0"
