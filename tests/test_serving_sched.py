"""Serving scheduler on the CPU, with a stub backend in place of the device.

Checks the admission bookkeeping that the device runs rely on:
* requests for a pattern whose plan is still being built wait for it: one
  build per pattern, however many same-pattern requests are admitted at once
  from concurrent producers;
* no two requests run on one plan at the same time, and no plan is freed while
  a request uses it;
* reserved bytes always equal the bytes of the plans in the table, a failed
  build gives its reservation back, and every plan is freed exactly once
  (explicitly, by the server).
"""
import random
import threading
import time

import numpy as np
import pytest

from suitesparse_amd import serving
from suitesparse_amd.problems import SymMatrix


class StubBackend:
    def __init__(self, fail_keys=(), plan_bytes=1000, build_s=0.01, run_s=0.002):
        self.fail_keys = set(fail_keys)
        self.plan_bytes = plan_bytes
        self.build_s, self.run_s = build_s, run_s
        self.lock = threading.Lock()
        self.builds = {}
        self.freed = {}
        self.active = {}
        self.errors = []
        self.live = set()

    def init(self):
        pass

    def estimate(self, req):
        return 2 * self.plan_bytes          # the estimate is above the real size, as on the device

    def build(self, req):
        time.sleep(self.build_s * random.random())
        if req.key in self.fail_keys:
            raise RuntimeError("planted build failure")
        with self.lock:
            self.builds[req.key] = self.builds.get(req.key, 0) + 1
            h = (req.key, self.builds[req.key])
            self.live.add(h)
        return h, self.plan_bytes

    def run(self, req, h):
        with self.lock:
            if h not in self.live:
                self.errors.append(f"run on a freed plan {h}")
            if self.active.get(h):
                self.errors.append(f"two requests on plan {h}")
            self.active[h] = True
        time.sleep(self.run_s * random.random())
        with self.lock:
            self.active[h] = False
        req.X = req.B.copy()
        req.info = {"relres": 0.0}
        req.status = 0

    def free(self, h):
        with self.lock:
            if self.active.get(h):
                self.errors.append(f"plan {h} freed while in use")
            self.freed[h] = self.freed.get(h, 0) + 1
            self.live.discard(h)


def _requests(count, nkeys, seed=0):
    rng = np.random.default_rng(seed)
    A = SymMatrix(n=1, p=np.array([0, 1]), i=np.array([0]), x=np.array([1.0]), name="one")
    return [serving.Request(rid=r, A=A, B=np.ones((1, 1)), key=f"k{int(rng.integers(nkeys))}")
            for r in range(count)]


def _submit_concurrently(srv, reqs, producers=8):
    chunks = [reqs[k::producers] for k in range(producers)]
    ths = [threading.Thread(target=lambda c=c: [srv.submit(r) for r in c]) for c in chunks]
    for t in ths:
        t.start()
    for t in ths:
        t.join()


def _wait_done(srv, count, timeout=60.0):
    t0 = time.time()
    while len(srv.completed) < count:
        assert time.time() - t0 < timeout, "stream did not drain"
        time.sleep(0.01)


def _check_reserved(srv):
    with srv._cv:
        assert srv._reserved == sum(p.nbytes + (p.pending[1] if p.pending else 0) for p in srv._plans.values())


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_one_build_per_pattern_under_concurrent_admission(seed):
    random.seed(seed)
    be = StubBackend()
    srv = serving.Server(slots=8, hbm_budget=1 << 40, backend=be)
    reqs = _requests(400, 12, seed)
    srv.start()
    _submit_concurrently(srv, reqs)
    _wait_done(srv, len(reqs))
    _check_reserved(srv)
    assert not be.errors, be.errors[:5]
    keys = {r.key for r in reqs}
    assert be.builds == {k: 1 for k in keys}          # no duplicate plan for any pattern
    assert srv.builds == len(keys)
    assert all(r.status == 0 for r in srv.completed)
    assert sum(r.reused_plan for r in srv.completed) == len(reqs) - len(keys)
    srv.drain()
    assert srv.reserved_bytes() == 0
    assert all(v == 1 for v in be.freed.values()) and len(be.freed) == len(keys)
    assert not be.live


def test_failed_build_releases_its_reservation():
    random.seed(3)
    be = StubBackend(fail_keys={"k1", "k4"})
    srv = serving.Server(slots=6, hbm_budget=1 << 40, backend=be)
    reqs = _requests(200, 6, seed=3)
    srv.start()
    _submit_concurrently(srv, reqs)
    _wait_done(srv, len(reqs))
    _check_reserved(srv)
    bad = [r for r in srv.completed if r.key in be.fail_keys]
    assert bad and all(r.status == -5 for r in bad)
    assert all(r.status == 0 for r in srv.completed if r.key not in be.fail_keys)
    with srv._cv:
        assert not any(k in srv._plans for k in be.fail_keys)
    srv.drain()
    assert srv.reserved_bytes() == 0
    assert not be.errors, be.errors[:5]


def test_tight_budget_evicts_idle_plans_only():
    random.seed(5)
    be = StubBackend(plan_bytes=1000)
    # room for two placeholders (estimate 2000 each) or four resident plans
    srv = serving.Server(slots=4, hbm_budget=4000, backend=be)
    reqs = _requests(300, 10, seed=5)
    srv.start()
    _submit_concurrently(srv, reqs, producers=4)
    _wait_done(srv, len(reqs))
    _check_reserved(srv)
    assert srv.evictions > 0
    assert not be.errors, be.errors[:5]
    assert all(r.status == 0 for r in srv.completed)
    srv.drain()
    assert srv.reserved_bytes() == 0
    assert all(v == 1 for v in be.freed.values())
    assert not be.live


def test_oversized_request_fails_without_blocking():
    be = StubBackend(plan_bytes=1000)
    srv = serving.Server(slots=2, hbm_budget=1500, backend=be)   # estimate 2000 > budget
    reqs = _requests(5, 2)
    srv.start()
    for r in reqs:
        srv.submit(r)
    _wait_done(srv, len(reqs))
    assert all(r.status == -2 for r in srv.completed)
    srv.drain()
    assert srv.reserved_bytes() == 0


class UpgradingStub(StubBackend):
    """Patterns k0..k2 have a final ordering other than their first one: the
    server builds their second plan in the background and swaps it in."""
    UPGRADED = ("k0", "k1", "k2")

    def first_ordering(self, key):
        return 0

    def final_ordering(self, key):
        return 1 if key in self.UPGRADED else 0

    def build(self, req, ordering=None):
        return super().build(req)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_background_plan_upgrade(seed):
    random.seed(seed)
    be = UpgradingStub(build_s=0.02)
    srv = serving.Server(slots=6, hbm_budget=1 << 40, backend=be)
    reqs = _requests(400, 6, seed)
    srv.start()
    _submit_concurrently(srv, reqs)
    _wait_done(srv, len(reqs))
    srv.drain()                                   # joins the background builds too
    assert not be.errors, be.errors[:5]
    assert all(r.status == 0 for r in srv.completed)
    keys = {r.key for r in reqs}
    up = [k for k in keys if k in UpgradingStub.UPGRADED]
    assert 1 <= srv.upgrades <= len(up)
    for k in keys:                                # a second plan only for the upgraded patterns
        assert be.builds[k] == (2 if k in up else 1) or (k in up and be.builds[k] == 1)
    assert srv.reserved_bytes() == 0
    assert all(v == 1 for v in be.freed.values()) and not be.live
