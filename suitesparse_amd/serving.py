"""Request loop: continuous batching of sparse SPD solves on one GPU.

Requests (a matrix tril(A), right-hand sides, a pattern key) arrive over time.
The server keeps ``slots`` worker threads; each takes the oldest waiting
request that fits the device-memory budget, factors it and solves with FP64
iterative refinement (FP32 factor by default), then frees its slot.  Every
plan (symbolic analysis + device plan + numeric factor storage) lives on its
own HIP stream, so the work of different requests overlaps on the device;
requests arriving while others run are admitted as soon as a slot and memory
are free -- not only when the device goes idle.

Admission is by the device bytes a plan holds (the factor, the inverse
diagonal blocks, the update plan and the solve vectors), estimated before the
plan is built from nnz(L) = xsize and checked against a fixed HBM budget.
Plans stay resident after a request finishes and are reused by later requests
with the same pattern (a refactorization: new values, the captured hipGraph is
replayed).  When a new plan does not fit, least-recently-used idle plans are
evicted.  Two requests with the same pattern share one plan, so they run one
after the other: the first one admitted registers a placeholder for its
pattern, and later ones wait for it instead of building a second plan.
Plans are freed explicitly by the server (eviction, a failed build, stop),
never by a finalizer on an arbitrary thread.

The ordering is AMD for 2D grids and METIS for 3D grids (the product's own
restatements, bit-exact with cholmod_l_analyze's AMD and METIS); the request's
pattern key names its kind.  A new 3D pattern is first served from an AMD plan
(its analysis takes about a quarter of METIS's on one core), while a
background thread orders the same pattern with METIS and builds that plan;
the first request to take the plan after it is ready swaps it in (the METIS
factor has fewer entries and flops), so only the first requests of a pattern
wait for an analysis, and never for METIS's.
"""
from __future__ import annotations

import collections
import dataclasses
import threading
import time

import numpy as np

from . import cholmod
from .problems import SymMatrix


@dataclasses.dataclass
class Request:
    rid: int
    A: SymMatrix
    B: np.ndarray                  # n x nrhs
    key: str                       # pattern identity (same key => same pattern)
    arrival: float = 0.0           # seconds from the start of the stream
    spec: tuple | None = None      # (kind, side, nrhs): drawn when submitted (request_specs)
    nrhs: int = 0                  # kept when the matrices are dropped after completion
    # results
    X: np.ndarray | None = None
    info: dict | None = None
    start: float = 0.0
    done: float = 0.0
    status: int = 0
    reused_plan: bool = False


@dataclasses.dataclass
class _Plan:
    """A resident plan, or the placeholder of one being built.  While
    ``state == "building"`` the plan holds the admission estimate in
    ``nbytes`` and one user (the slot building it), so requests with the same
    key wait for it instead of building a second plan."""
    key: str
    handle: object = None          # backend plan (DeviceBackend: cholmod.Factor + Common)
    nbytes: int = 0
    state: str = "building"        # "building" -> "ready"
    last_used: float = 0.0
    users: int = 0
    pending: object = None         # (handle, nbytes) of the final-ordering plan, built in the background


def plan_bytes_estimate(xsize: int, nsuper_cols_sq: int, nnz: int, n: int, nrhs: int,
                        precision: str) -> int:
    """Device bytes of a plan: factor + inverse blocks in the working type, the
    value/CSR maps (int64), solve and refinement vectors (FP64)."""
    t = 4 if precision == "fp32" else 8
    return int(t * (xsize + nsuper_cols_sq) + 40 * nnz + 8 * n * (4 * nrhs + 2) + (1 << 20))


class DeviceBackend:
    """Plans on the MI355X: analysis (AMD for 2D patterns, METIS for 3D) +
    device plan per pattern; each request is a refactorization (graph replay
    after the first) and a solve with FP64 iterative refinement."""

    def __init__(self, precision: str = "fp32", tol: float = 1e-10, maxiter: int = 20):
        self.precision, self.tol, self.maxiter = precision, tol, maxiter

    def init(self):
        cm = cholmod.Common()
        if not cm.start():
            raise RuntimeError(f"device init failed: {cm.status}")

    def estimate(self, req: Request) -> int:
        # before analysis the factor size is unknown: a conservative fill bound
        # from nnz(A) (replaced by the plan's real size once it exists)
        nnz = int(req.A.p[-1])
        xguess = 40 * nnz + 64 * req.A.n
        return plan_bytes_estimate(xguess, 64 * req.A.n, nnz, req.A.n, req.B.shape[1], self.precision)

    def final_ordering(self, key: str) -> int:
        return cholmod.CHOLMOD_METIS if key.startswith("3d") else cholmod.CHOLMOD_AMD

    def first_ordering(self, key: str) -> int:
        """The ordering of a pattern's first plan: AMD (a METIS plan follows
        in the background when final_ordering differs)."""
        return cholmod.CHOLMOD_AMD

    def build(self, req: Request, ordering: int | None = None):
        if ordering is None:
            ordering = self.first_ordering(req.key)
        cm = cholmod.Common(nmethods=1, ordering=ordering, precision=self.precision)
        L = cholmod.analyze(req.A, cm)
        if L is None:
            raise RuntimeError(f"analyze failed: {cm.status}")
        if L._numeric(cm) is None:
            L.free()
            raise RuntimeError(f"plan failed: {cm.status}")
        nbytes = L.info()["device_bytes"] + 8 * L.n * (4 * req.B.shape[1] + 2)
        return (L, cm), nbytes

    def run(self, req: Request, handle):
        L, cm = handle
        if not cholmod.factorize(req.A, L, cm) or cm.status != cholmod.CHOLMOD_OK:
            req.status = cm.status
            return
        X, info = cholmod.solve_refine(L, req.B, cm, tol=self.tol, maxiter=self.maxiter)
        req.X = X
        req.info = dict(info or {}, reserved=(req.info or {}).get("reserved", 0))
        req.status = cm.status

    def free(self, handle):
        handle[0].free()


class Server:
    def __init__(self, slots: int = 8, hbm_budget: int = 64 << 30, precision: str = "fp32",
                 tol: float = 1e-10, maxiter: int = 20, keep_results: bool = True, backend=None):
        self.slots = slots
        self.keep_results = keep_results   # False: drop A, B and X once a request is done
        self.budget = int(hbm_budget)
        self.backend = backend or DeviceBackend(precision, tol, maxiter)
        self._cv = threading.Condition()
        self._queue: collections.deque[Request] = collections.deque()
        self._plans: dict[str, _Plan] = {}
        self._reserved = 0              # bytes of resident plans and placeholders
        self._stop = False
        self._threads = []
        self._t0 = None
        self.completed: list[Request] = []
        self.evictions = 0
        self.builds = 0
        self.upgrades = 0               # plans replaced by their final-ordering plan
        self._upgraders: list[threading.Thread] = []
        self.backend.init()

    # -------------------------------------------------------------- lifecycle
    def start(self):
        self._t0 = time.perf_counter()
        for s in range(self.slots):
            th = threading.Thread(target=self._worker, name=f"slot{s}", daemon=True)
            th.start()
            self._threads.append(th)

    def now(self) -> float:
        return time.perf_counter() - self._t0

    def submit(self, req: Request):
        with self._cv:
            self._queue.append(req)
            self._cv.notify_all()

    def drain(self):
        with self._cv:
            while self._queue or any(p.users for p in self._plans.values()):
                self._cv.wait(0.05)
        self.stop()

    def stop(self):
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for th in self._threads:
            th.join()
        for th in list(self._upgraders):
            th.join()
        with self._cv:
            for p in list(self._plans.values()):
                self._release(p)

    def reserved_bytes(self) -> int:
        with self._cv:
            return self._reserved

    # -------------------------------------------------------------- admission
    # Every method below runs with self._cv held.
    def _take(self) -> tuple[Request, _Plan] | None:
        """Oldest waiting request whose plan is idle, or whose new plan fits
        the budget (a placeholder is registered for it at once)."""
        idx = 0
        while idx < len(self._queue):
            req = self._queue[idx]
            p = self._plans.get(req.key)
            if p is not None:
                if p.users == 0 and p.state == "ready":
                    del self._queue[idx]
                    if p.pending is not None:      # the final-ordering plan is ready: swap it in
                        old, (p.handle, nbytes) = p.handle, p.pending
                        p.pending = None
                        self._reserved -= p.nbytes
                        p.nbytes = nbytes
                        self.backend.free(old)
                        self.upgrades += 1
                    p.users += 1
                    req.reused_plan = True
                    return req, p
                idx += 1          # its plan is busy or still being built: wait for it
                continue
            need = self.backend.estimate(req)
            if need > self.budget:
                # can never fit: fail it (CHOLMOD_OUT_OF_MEMORY) instead of waiting forever
                del self._queue[idx]
                req.status = -2
                req.info = {"error": "plan larger than the HBM budget"}
                req.start = req.done = self.now()
                self.completed.append(req)
                self._cv.notify_all()
                continue
            if self._reserved + need > self.budget:
                self._evict(self._reserved + need - self.budget)
            if self._reserved + need <= self.budget:
                del self._queue[idx]
                self._reserved += need
                req.info = {"reserved": need}
                p = _Plan(key=req.key, nbytes=need, users=1)
                self._plans[req.key] = p
                return req, p
            return None          # head-of-line: keep arrival order for new patterns
        return None

    def _release(self, p: _Plan):
        """Drop a plan from the table and free its device memory (explicitly,
        here, never from a finalizer on some other thread)."""
        if self._plans.get(p.key) is p:
            del self._plans[p.key]
        self._reserved -= p.nbytes
        p.nbytes = 0
        if p.handle is not None:
            h, p.handle = p.handle, None
            self.backend.free(h)
        if p.pending is not None:
            (h, nb), p.pending = p.pending, None
            self._reserved -= nb
            self.backend.free(h)

    def _evict(self, nbytes: int):
        idle = sorted((p for p in self._plans.values() if p.users == 0 and p.state == "ready"),
                      key=lambda p: p.last_used)
        freed = 0
        for p in idle:
            if freed >= nbytes:
                break
            freed += p.nbytes
            self._release(p)
            self.evictions += 1

    # -------------------------------------------------------------- worker
    def _worker(self):
        while True:
            with self._cv:
                got = None
                while not self._stop:
                    got = self._take()
                    if got is not None:
                        break
                    self._cv.wait(0.01)
                if got is None:
                    return
            req, plan = got
            req.start = self.now()
            try:
                if plan.state == "building":
                    self._build(req, plan)
                self.backend.run(req, plan.handle)
            except Exception as e:      # report, never kill the slot
                req.status = -5
                req.info = {"error": repr(e)}
            req.done = self.now()
            req.nrhs = req.B.shape[1] if req.B is not None else 0
            if not self.keep_results:
                req.A = req.B = req.X = None
            with self._cv:
                plan.users -= 1
                plan.last_used = req.done
                if plan.state == "building":
                    self._release(plan)  # the build failed: give back its reservation
                self.completed.append(req)
                self._cv.notify_all()

    def _build(self, req: Request, plan: _Plan):
        handle, nbytes = self.backend.build(req)
        with self._cv:
            # swap the admission estimate for the plan's real size
            self._reserved += nbytes - plan.nbytes
            plan.nbytes = nbytes
            plan.handle = handle
            plan.state = "ready"
            self.builds += 1
            upgrade = (hasattr(self.backend, "final_ordering") and not self._stop and
                       self.backend.final_ordering(req.key) != self.backend.first_ordering(req.key))
            if upgrade:
                # the pattern's matrices, held here: the request may drop them once it is done
                pattern = Request(rid=req.rid, A=req.A, B=req.B, key=req.key)
                th = threading.Thread(target=self._upgrade, args=(pattern, plan), name=f"upgrade-{req.key}",
                                      daemon=True)
                self._upgraders.append(th)
                th.start()

    def _upgrade(self, pattern: Request, plan: _Plan):
        """Build the pattern's final-ordering plan beside the serving one and
        leave it in plan.pending; its admission estimate is reserved first (no
        upgrade when it does not fit)."""
        req = pattern
        with self._cv:
            need = self.backend.estimate(pattern)
            if self._stop or self._plans.get(plan.key) is not plan or self._reserved + need > self.budget:
                return
            self._reserved += need
        try:
            handle, nbytes = self.backend.build(pattern, self.backend.final_ordering(req.key))
        except Exception:
            with self._cv:
                self._reserved -= need
            return
        with self._cv:
            self._reserved += nbytes - need
            if self._plans.get(plan.key) is plan and plan.state == "ready" and not self._stop:
                plan.pending = (handle, nbytes)
            else:                               # evicted or stopped meanwhile
                self._reserved -= nbytes
                self.backend.free(handle)
            self._cv.notify_all()


# ----------------------------------------------------------------------------
# the serving workload: Poisson arrivals over a mix of 2D and 3D grids
# ----------------------------------------------------------------------------

def make_requests(count: int, seed: int = 4, side2d=(32, 512), side3d=(16, 64),
                  refactor_frac: float = 0.3, nrhs=(1, 8), sigma: float = 0.5):
    """Requests without arrival times: half 2D 5-point (side U{32..512}), half
    3D 7-point (side U{16..64}), log-normal(sigma) weights; a refactor_frac
    share reuse the pattern of an earlier request with redrawn values; nrhs
    U{1..8}, right-hand sides U(-1, 1)."""
    from . import problems
    rng = np.random.default_rng(seed)
    reqs = []
    seen = []
    for rid in range(count):
        if seen and rng.random() < refactor_frac:
            kind, side = seen[rng.integers(len(seen))]
        elif rng.random() < 0.5:
            kind, side = "2d", int(rng.integers(side2d[0], side2d[1] + 1))
        else:
            kind, side = "3d", int(rng.integers(side3d[0], side3d[1] + 1))
        seen.append((kind, side))
        A = (problems.grid2d_lognormal(side, rng, sigma=sigma) if kind == "2d"
             else problems.grid3d_lognormal(side, rng, sigma=sigma))
        k = int(rng.integers(nrhs[0], nrhs[1] + 1))
        B = rng.uniform(-1.0, 1.0, (A.n, k))
        reqs.append(Request(rid=rid, A=A, B=np.asfortranarray(B), key=f"{kind}{side}"))
    return reqs


def poisson_arrivals(count: int, rate: float, seed: int = 4) -> np.ndarray:
    rng = np.random.default_rng(seed + 1000)
    return np.cumsum(rng.exponential(1.0 / rate, size=count))


class PatternBank:
    """One matrix per (kind, side) pattern; each request redraws the values
    (log-normal edge weights, diagonal = row sum + shift) on that pattern.
    The index arrays of a pattern are built once."""

    def __init__(self, sigma: float = 0.5, shift: float = 1e-3):
        self.sigma, self.shift = sigma, shift
        self._pats = {}
        self._lock = threading.Lock()

    def _pattern(self, kind: str, side: int):
        key = (kind, side)
        with self._lock:
            return self._pattern_locked(key, kind, side)

    def _pattern_locked(self, key, kind: str, side: int):
        if key not in self._pats:
            from . import problems
            gen = problems.grid2d_lognormal if kind == "2d" else problems.grid3d_lognormal
            A = gen(side, np.random.default_rng(0), sigma=self.sigma, shift=self.shift)
            col = np.repeat(np.arange(A.n), np.diff(A.p))
            off = np.flatnonzero(A.i != col)
            dpos = np.flatnonzero(A.i == col)
            self._pats[key] = (A, off, dpos, A.i[off], col[off])
        return self._pats[key]

    def draw(self, kind: str, side: int, rng: np.random.Generator) -> SymMatrix:
        A, off, dpos, ri, ci = self._pattern(kind, side)
        w = np.exp(self.sigma * rng.standard_normal(off.size))
        x = np.empty(A.x.size)
        x[off] = -w
        x[dpos] = self.shift + np.bincount(ri, w, A.n) + np.bincount(ci, w, A.n)
        return SymMatrix(n=A.n, p=A.p, i=A.i, x=x, name=A.name)


def request_specs(count: int, seed: int = 4, side2d=(32, 512), side3d=(16, 64),
                  refactor_frac: float = 0.3, nrhs=(1, 8)):
    """The serving workload without its matrices: half 2D 5-point (side
    U{32..512}), half 3D 7-point (side U{16..64}); a refactor_frac share
    reuses the pattern of an earlier request; nrhs U{1..8}.  Values are drawn
    when the request is submitted (materialize)."""
    rng = np.random.default_rng(seed)
    reqs, seen = [], []
    for rid in range(count):
        if seen and rng.random() < refactor_frac:
            kind, side = seen[rng.integers(len(seen))]
        elif rng.random() < 0.5:
            kind, side = "2d", int(rng.integers(side2d[0], side2d[1] + 1))
        else:
            kind, side = "3d", int(rng.integers(side3d[0], side3d[1] + 1))
        seen.append((kind, side))
        k = int(rng.integers(nrhs[0], nrhs[1] + 1))
        reqs.append(Request(rid=rid, A=None, B=None, key=f"{kind}{side}", spec=(kind, side, k)))
    return reqs


def materialize(req: Request, bank: PatternBank, seed: int = 4):
    """Draw the request's values and right-hand sides (per-request seed)."""
    kind, side, k = req.spec
    rng = np.random.default_rng([seed, req.rid])
    req.A = bank.draw(kind, side, rng)
    req.B = np.asfortranarray(rng.uniform(-1.0, 1.0, (req.A.n, k)))


def run_stream(server: Server, reqs, arrivals, bank: PatternBank | None = None,
               lookahead: int = 64, producers: int = 8, progress: int = 0) -> dict:
    """Feed requests at their arrival times; return latency / throughput stats.
    Requests built by request_specs are materialized ahead of their arrival
    by a pool of producer threads, in arrival order, at most `lookahead` ahead.
    progress > 0 prints a line (to stderr) each time that many more requests were submitted."""
    import concurrent.futures
    import faulthandler
    import queue
    import sys
    if not faulthandler.is_enabled():
        # a crash in native code then prints every thread's Python stack
        faulthandler.enable(file=sys.stderr, all_threads=True)
    ready: "queue.Queue" = queue.Queue(maxsize=lookahead)
    bank = bank or PatternBank()

    def one(req):
        if req.A is None:
            materialize(req, bank)
        return req

    def produce():
        with concurrent.futures.ThreadPoolExecutor(max_workers=producers) as ex:
            for c0 in range(0, len(reqs), 16):
                for req in ex.map(one, reqs[c0:c0 + 16]):
                    ready.put(req)

    prod = threading.Thread(target=produce, daemon=True)
    prod.start()
    server.start()
    late = 0
    for t in arrivals:
        req = ready.get()
        req.arrival = float(t)
        delay = req.arrival - server.now()
        if delay > 0:
            time.sleep(delay)
        elif delay < -0.005:
            late += 1                     # producer behind the arrival clock
        server.submit(req)
        if progress and (req.rid + 1) % progress == 0:
            print(f"[serving] {req.rid + 1} submitted, {len(server.completed)} done at t = "
                  f"{server.now():.1f} s", file=sys.stderr, flush=True)
    prod.join()
    server.drain()
    done = sorted(server.completed, key=lambda r: r.rid)
    lat = np.array([r.done - r.arrival for r in done]) * 1e3
    span = max(r.done for r in done) - min(r.arrival for r in done)
    ok = [r for r in done if r.status == 0]
    nr = [r.nrhs for r in ok]
    return {
        "requests": len(done), "ok": len(ok),
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()), "solves_per_s": len(ok) / span,
        "rhs_per_s": sum(nr) / span,
        "plan_reuse": sum(r.reused_plan for r in done) / max(1, len(done)),
        "evictions": server.evictions,
        "plan_upgrades": server.upgrades,
        "max_relres": max((r.info or {}).get("relres", 0.0) for r in ok) if ok else None,
        "fp64_fallbacks": sum(bool((r.info or {}).get("fp64_fallback")) for r in ok),
        "late_submissions": late,
        "wall_s": span,
    }
