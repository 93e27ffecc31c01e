"""Serving loop on the GPU: requests admitted while others run (several slots,
each its own HIP stream), plans reused across same-pattern requests, memory
admission and eviction; every result is the same alone and in the mix.

Tolerance: every request reaches relative residual <= 1e-10 (FP32 factor +
FP64 refinement) both alone and in the mix, and the two solutions differ by
at most 1e-8 relative (both are converged solutions of the same system)."""
import numpy as np
import pytest

from suitesparse_amd import serving

pytestmark = pytest.mark.gpu


def _reqs(seed=11, count=14):
    return serving.make_requests(count, seed=seed, side2d=(16, 48), side3d=(8, 14),
                                 refactor_frac=0.35, nrhs=(1, 4))


def _run(reqs, slots, budget, rate=None):
    srv = serving.Server(slots=slots, hbm_budget=budget, precision="fp32", tol=1e-10)
    arrivals = (np.zeros(len(reqs)) if rate is None
                else serving.poisson_arrivals(len(reqs), rate, seed=4))
    stats = serving.run_stream(srv, reqs, arrivals)
    return {r.rid: r for r in srv.completed}, stats


def test_same_alone_and_in_mix(gpu):
    alone, st1 = _run(_reqs(), slots=1, budget=16 << 30)
    mixed, st2 = _run(_reqs(), slots=4, budget=16 << 30, rate=2000.0)
    print("alone", st1, "\nmixed", st2, flush=True)
    assert st1["ok"] == st2["ok"] == len(alone)
    assert st2["plan_reuse"] > 0
    for rid, ra in alone.items():
        rm = mixed[rid]
        for r in (ra, rm):
            R = r.B - r.A.matvec(r.X)
            rel = np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(r.B, axis=0))
            assert rel <= 1e-10, (rid, rel)
        d = np.linalg.norm(rm.X - ra.X) / np.linalg.norm(ra.X)
        assert d <= 1e-8, (rid, d)


def test_oversized_request_rejected(gpu):
    """A request that cannot fit the budget at all fails with CHOLMOD_OUT_OF_MEMORY
    while the others complete."""
    reqs = _reqs(seed=13, count=5)
    srv = serving.Server(slots=2, hbm_budget=3 << 20, precision="fp32")
    big = serving.make_requests(1, seed=99, side2d=(200, 200), side3d=(40, 40), refactor_frac=0.0)[0]
    big.rid = 100
    stats = serving.run_stream(srv, reqs + [big], np.zeros(6))
    done = {r.rid: r for r in srv.completed}
    assert done[100].status == -2
    assert stats["requests"] == 6


def test_admission_under_tight_budget(gpu):
    """A budget that fits one or two plans forces head-of-line waits and LRU
    eviction; every request still completes correctly."""
    reqs = _reqs(seed=12, count=10)
    res, st = _run(reqs, slots=3, budget=8 << 20)
    print(st, flush=True)
    assert st["ok"] == len(reqs)
    assert st["evictions"] > 0
    for r in res.values():
        R = r.B - r.A.matvec(r.X)
        assert np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(r.B, axis=0)) <= 1e-10


def test_background_metis_plan_replaces_amd(gpu):
    """A 3D pattern is first served from an AMD plan; its METIS plan, built in
    the background, replaces it for the later requests of the pattern, and
    every request still converges."""
    from suitesparse_amd import problems
    A0 = problems.grid3d_lognormal(12, np.random.default_rng(3))
    rng = np.random.default_rng(8)
    reqs = [serving.Request(rid=k, A=problems.revalue(A0, np.random.default_rng(50 + k)),
                            B=rng.uniform(-1, 1, (A0.n, 2)), key="3d_12") for k in range(10)]
    srv = serving.Server(slots=2, hbm_budget=16 << 30, precision="fp32", tol=1e-10)
    stats = serving.run_stream(srv, reqs, np.arange(len(reqs)) * 0.1)
    print(stats, flush=True)
    assert stats["ok"] == len(reqs)
    assert srv.upgrades == 1 and srv.builds == 1
    for r in srv.completed:
        R = r.B - r.A.matvec(r.X)
        assert np.max(np.linalg.norm(R, axis=0) / np.linalg.norm(r.B, axis=0)) <= 1e-10
