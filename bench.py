"""Benchmark: supernodal Cholesky factor + solve on one MI355X.

One "step" = one numeric factorization + one solve (+ FP64 refinement for the
mixed form) of the same system, inputs resident in HBM.

  * headline (the metric's workload): 3D 27-point stencil on 96^3
    (n = 884,736), METIS ordering (the product's own, identical to the
    reference's), 4 RHS:
    FP32 factor + FP64 iterative refinement to a 1e-10 relative residual
    against plain FP64 factor + solve.  value = FP32+IR factor+solve GFLOP/s =
    (fl + 4 lnz nrhs) / time-to-solution; the FP64 figure, the speedup,
    solves/s and p50/p99 step latency beside it; forward error against the
    reference's FP64 x; roofline of the dominant kernel class from HIP events.
  * 64^3 7-point diffusion, FP64, 1 RHS (secondary record).
  * batched: 4,096 2D systems (side 16..128), FP64 and FP32+IR.
  * serving: a Poisson stream of 20,000 requests (--serving N), 2D 5-point
    side 32..512 and 3D 7-point side 16..64, FP32+IR, run in a child process.
  * cpu_baseline: the reference CHOLMOD (compiled from the reference sources,
    MKL BLAS, OpenMP) on the host cores this job may use, same matrix and
    ordering: FP64, and its single-precision factor + FP64 refinement.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batch]
                    [--no-64] [--serving N] [--only-headline] [--only-serving]

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

MFMA_PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def product_ordering(A, ordering: int) -> tuple[np.ndarray, float]:
    """The product's own fill-reducing ordering and analysis of A (METIS for
    the 3D workloads, as cholmod_l_analyze computes it): the final permutation,
    which the later analyses of the same matrix take as given, and the time."""
    from suitesparse_amd import cholmod
    cm = cholmod.Common(nmethods=1, ordering=ordering)
    t0 = time.perf_counter()
    L = cholmod.analyze(A, cm)
    assert L is not None and L.ordering == ordering, cm.status
    perm = L.Perm.copy()
    L.free()
    return perm, time.perf_counter() - t0


def relres(A, B, X) -> float:
    R = B - A.matvec(X)
    return float((np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)).max())


def backward_error(A, B, X) -> float:
    """|Ax - b| / (|A| |x| + |b|) in infinity norms (largest absolute row sum),
    the residual cholmod_dl_demo.c:614-617 reports; |A| of the full symmetric A."""
    B2 = np.asarray(B).reshape(A.n, -1)
    X2 = np.asarray(X).reshape(A.n, -1)
    R = B2 - A.matvec(X2)
    col = np.repeat(np.arange(A.n), np.diff(A.p))
    ax = np.abs(A.x)
    off = A.i != col
    anorm = (np.bincount(A.i, ax, A.n) + np.bincount(col[off], ax[off], A.n)).max()

    def inf(M):
        return float(np.abs(M).sum(axis=1).max())
    return float(inf(R) / (anorm * inf(X2) + inf(B2)))


TUNING = ROOT / "suitesparse_amd" / "tuning"
PROFILES = ROOT / "profiles"
TRAFFIC_FILE = PROFILES / "traffic_update_diffusion3d_64_fp64.md"
UPDATE_KERNEL = "k_tile<double, 0,"          # descendant update + scatter, FP64, every tile edge
CALIB_KERNEL = "__amd_rocclr_copyBuffer"


def calib_copy(mib: int):
    """One device-to-device copy of exactly `mib` MiB (read mib, write mib): the
    reference dispatch that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def _counter_rows(d: pathlib.Path, counter: str):
    """(kernel name, value) per dispatch from a rocprofv3 --pmc output folder."""
    import csv
    files = sorted(d.rglob("*counter_collection.csv"))
    if not files:
        raise FileNotFoundError(f"no counter_collection.csv under {d}")
    rows = []
    for f in files:
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") == counter:
                    rows.append((r["Kernel_Name"], float(r["Counter_Value"]), int(r["Dispatch_Id"])))
    return rows


def traffic_summary(fetch_dir: str, write_dir: str, calib_mib: int, cmd: str) -> dict:
    """HBM bytes per launch of the update kernel from two --pmc runs: the average
    of FETCH_SIZE (WRITE_SIZE) over its dispatches, scaled by what the counter
    reads for the calibration copy of calib_mib MiB (known: calib_mib * 1024 KiB
    read and written)."""
    res = {}
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows = _counter_rows(pathlib.Path(d), counter)
        cal = max((v for k, v, _ in rows if k.startswith(CALIB_KERNEL)), default=0.0)
        upd = [v for k, v, _ in rows if UPDATE_KERNEL in k]
        if cal <= 0 or not upd:
            raise RuntimeError(f"{counter}: calibration copy or update dispatches missing")
        scale = calib_mib * 1024.0 / cal
        res[counter] = {"calib_kib": cal, "scale": scale, "dispatches": len(upd),
                        "avg_kib": sum(upd) / len(upd),
                        "bytes_per_launch": sum(upd) / len(upd) * scale * 1024.0}
    total = res["FETCH_SIZE"]["bytes_per_launch"] + res["WRITE_SIZE"]["bytes_per_launch"]
    lines = ["# HBM traffic of the descendant-update kernel, 64^3 7-point diffusion, FP64", "",
             "Two counter passes over the same command (repository root, one MI355X):", "", "```"]
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        lines.append(f"timeout -s KILL 300 rocprofv3 --pmc {counter} --output-format csv -d {d} -o run -- {cmd}")
    lines += ["```", "",
              f"Calibration: one device-to-device copy of {calib_mib} MiB ({CALIB_KERNEL}) in the same "
              "process; scale = known KiB / counter KiB.", "",
              "| counter | calibration copy (KiB) | scale | update dispatches | avg per dispatch (KiB) | bytes per launch (scaled) |",
              "|---|---|---|---|---|---|"]
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        r = res[counter]
        lines.append(f"| {counter} | {r['calib_kib']:.0f} | {r['scale']:.4f} | {r['dispatches']} | "
                     f"{r['avg_kib']:.1f} | {r['bytes_per_launch']:.0f} |")
    lines += ["", f"traffic_bytes_per_launch: {total:.0f}", ""]
    TRAFFIC_FILE.write_text("\n".join(lines))
    return {"traffic_bytes_per_launch": total, **res}


# kernel classes of the counter summaries: (label, test on the kernel name, algorithmic note)
COUNTER_CLASSES = [
    ("descendant update + scatter", lambda k: "k_tile<" in k and ", 0," in k or "k_tile_h<0" in k),
    ("in-supernode trailing update", lambda k: "k_tile<" in k and ", 1," in k or "k_tile_h<1" in k),
    ("panel TRSM", lambda k: "k_tile<" in k and ", 2," in k or "k_tile_h<2" in k or "k_trsm_tiny" in k),
    ("diagonal POTRF", lambda k: "k_potrf" in k),
    ("assembly (fill, scatter A, beta)", lambda k: any(t in k for t in ("k_fill", "k_scatter_a", "k_add_diag",
                                                                         "k_gather_vals"))),
    ("solve sweeps", lambda k: any(t in k for t in ("k_chain_", "k_sweep_", "k_perm_"))),
    ("residual + norms", lambda k: any(t in k for t in ("k_residual", "k_sumsq"))),
    ("update/downdate", lambda k: "k_updown" in k),
    ("small supernodes, fused", lambda k: "k_small_level" in k),
]


def counter_summary(dirs, calib_mib: int, title: str, cmds, out: pathlib.Path, cus: int = 256, xcds: int = 8):
    """Per kernel class, from rocprofv3 --pmc folders of the same command: dispatches,
    device time (the dispatches' own timestamps), HBM bytes (FETCH_SIZE, WRITE_SIZE
    scaled by the calibration copy) and GB/s, MFMA busy share and MFMA flops
    (512 * SQ_INSTS_VALU_MFMA_MOPS_*) and TFLOP/s.  GRBM_GUI_ACTIVE is summed over
    the XCDs (it reads xcds * clock * duration), so the busy share is
    SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE / xcds * CUs * 4)."""
    import csv
    per = {}          # (dir, dispatch) -> {name, t, counters}
    for d in dirs:
        for f in sorted(pathlib.Path(d).rglob("*counter_collection.csv")):
            with open(f) as fh:
                for r in csv.DictReader(fh):
                    key = (d, int(r["Dispatch_Id"]))
                    e = per.setdefault(key, {"name": r["Kernel_Name"], "t": 0.0, "c": {}})
                    e["t"] = max(e["t"], (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9)
                    e["c"][r["Counter_Name"]] = e["c"].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    scale = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        cal = max((e["c"][ctr] for e in per.values() if ctr in e["c"] and e["name"].startswith(CALIB_KERNEL)),
                  default=0.0)
        if cal > 0:
            scale[ctr] = calib_mib * 1024.0 / cal
    agg = {}
    for e in per.values():
        if e["name"].startswith(CALIB_KERNEL):
            continue
        cls = next((lab for lab, f in COUNTER_CLASSES if f(e["name"])), None)
        if cls is None:
            continue
        a = agg.setdefault(cls, {"n": {}, "t": {}, "c": {}})
        for ctr, v in e["c"].items():
            a["c"][ctr] = a["c"].get(ctr, 0.0) + v
            a["n"][ctr] = a["n"].get(ctr, 0) + 1
            a["t"][ctr] = a["t"].get(ctr, 0.0) + e["t"]
    lines = [f"# Hardware counters per kernel class — {title}", "", "Counter passes, each a run of its own "
             "(repository root, one MI355X):", "", "```"] + list(cmds) + ["```", "",
             f"HBM bytes: FETCH_SIZE and WRITE_SIZE (KiB per dispatch) times the factor that makes the calibration "
             f"copy of {calib_mib} MiB in the same process read and write exactly {calib_mib} MiB "
             f"(FETCH x{scale.get('FETCH_SIZE', float('nan')):.4f}, WRITE x{scale.get('WRITE_SIZE', float('nan')):.4f}). "
             "GB/s = bytes / the dispatches' own duration in that pass (dispatches are serialized while counters "
             "are read). MFMA busy = SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE / 8 x 256 CUs x 4 SIMDs), summed "
             "over the class's dispatches; GRBM_GUI_ACTIVE is summed over the 8 XCDs (it reads 16-19 GHz of "
             "dispatch time here). MFMA flops = 512 x SQ_INSTS_VALU_MFMA_MOPS_<type>; peaks: FP64 78.6 TFLOP/s, "
             "FP32 157.3, FP16 2,500 (dense).", "",
             "| class | dispatches | fetch GB | write GB | HBM GB/s | MFMA busy | MFMA TFLOP (F64 / F32 / F16) | MFMA TFLOP/s |",
             "|---|---|---|---|---|---|---|---|"]
    for lab, _ in COUNTER_CLASSES:
        if lab not in agg:
            continue
        a = agg[lab]
        c = a["c"]
        fetch = c["FETCH_SIZE"] * scale["FETCH_SIZE"] * 1024 / 1e9 if "FETCH_SIZE" in c and "FETCH_SIZE" in scale else None
        write = c["WRITE_SIZE"] * scale["WRITE_SIZE"] * 1024 / 1e9 if "WRITE_SIZE" in c and "WRITE_SIZE" in scale else None
        gbs = None
        if fetch is not None and write is not None and a["t"].get("FETCH_SIZE") and a["t"].get("WRITE_SIZE"):
            gbs = fetch / a["t"]["FETCH_SIZE"] + write / a["t"]["WRITE_SIZE"]
        busy = (c["SQ_VALU_MFMA_BUSY_CYCLES"] / (c["GRBM_GUI_ACTIVE"] / xcds * cus * 4)
                if c.get("GRBM_GUI_ACTIVE") and "SQ_VALU_MFMA_BUSY_CYCLES" in c else None)
        fl = [c.get(f"SQ_INSTS_VALU_MFMA_MOPS_{t}") for t in ("F64", "F32", "F16")]
        fl = [None if v is None else 512 * v / 1e12 for v in fl]
        tmf = a["t"].get("SQ_VALU_MFMA_BUSY_CYCLES")
        rate = sum(v for v in fl if v) / tmf if tmf and any(fl) else None
        n = max(a["n"].values())
        fmt = lambda v, f: "–" if v is None else f.format(v)
        lines.append(f"| {lab} | {n} | {fmt(fetch, '{:.2f}')} | {fmt(write, '{:.2f}')} | {fmt(gbs, '{:.0f}')} | "
                     f"{fmt(busy, '{:.1%}')} | {' / '.join(fmt(v, '{:.3f}') for v in fl)} | {fmt(rate, '{:.1f}')} |")
    out.write_text("\n".join(lines) + "\n")
    return agg


def read_traffic():
    """Bytes per update launch from the committed counter summary, or None."""
    if not TRAFFIC_FILE.exists():
        return None
    for line in TRAFFIC_FILE.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def tile_table(name: str, precision: str, mode: str, L):
    """Install the tile table for this workload: 'auto' runs the autotuner and
    saves its table; otherwise a saved table is loaded when one exists."""
    path = TUNING / f"{name}_{precision}.json"
    if mode == "auto":
        t0 = time.perf_counter()
        table = L.autotune(reps=3)
        path.write_text(json.dumps({"table": table, "tune_s": time.perf_counter() - t0}))
        return "autotuned"
    if mode == "saved" and path.exists():
        L.set_tiles(json.loads(path.read_text())["table"])
        return "saved"
    return "default"


def gpu_factor_solve(A, perm, B, precision, steps, warmup, refine=False, tol=1e-10,
                     tune=("", "none"), roof_steps=0):
    """Run `steps` timed factor+solve steps (graph replay, inputs resident in HBM);
    then `roof_steps` more steps with HIP events around the update (class 1)
    and trailing (class 4) launches, for the roofline of the dominant kernel
    class.  Returns timings, the last X and the factor."""
    import torch
    from suitesparse_amd import cholmod
    cm = cholmod.Common(nmethods=1, ordering=cholmod.CHOLMOD_GIVEN, precision=precision)
    t0 = time.perf_counter()
    L = cholmod.analyze_p(A, perm, cm)
    assert L is not None, cm.status
    t_an = time.perf_counter() - t0
    assert cm.start()
    dev = torch.device("cuda:0")
    Ax = torch.as_tensor(A.x, device=dev)
    Bd = torch.as_tensor(np.asfortranarray(B), device=dev)
    Bd = Bd.t().contiguous().t()
    t0 = time.perf_counter()
    L._numeric(cm)
    t_plan = time.perf_counter() - t0
    tiles = tile_table(tune[0], precision, tune[1], L) if tune[0] else "default"
    info = None

    def step():
        nonlocal info
        ok = cholmod.factorize(A, L, cm, values=Ax)
        assert ok and cm.status == 0, cm.status
        if refine:
            X, info = cholmod.solve_refine(L, Bd, cm, tol=tol, maxiter=20)
        else:
            X = cholmod.solve(cholmod.CHOLMOD_A, L, Bd, cm)
        return X

    for _ in range(warmup):
        step()
    lat = []
    torch.cuda.synchronize()
    tstart = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        X = step()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t1)
    torch.cuda.synchronize()
    ttot = time.perf_counter() - tstart
    classes = None
    if roof_steps > 0:
        L.set_timing(True, classes=[1, 4])    # HIP events on the factor's stream
        for _ in range(roof_steps):
            step()
        torch.cuda.synchronize()
        classes = L.timing()
        classes["steps"] = roof_steps
        L.set_timing(False)
    # split factor / solve times on one extra untimed step
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    cholmod.factorize(A, L, cm, values=Ax)
    torch.cuda.synchronize()
    tf = time.perf_counter() - t1
    t1 = time.perf_counter()
    if refine:
        cholmod.solve_refine(L, Bd, cm, tol=tol, maxiter=20)
    else:
        cholmod.solve(cholmod.CHOLMOD_A, L, Bd, cm)
    torch.cuda.synchronize()
    ts = time.perf_counter() - t1
    Xh = X.t().contiguous().t().cpu().numpy() if hasattr(X, "cpu") else X
    return {
        "L": L, "lat": lat, "t_total": ttot, "t_factor": tf, "t_solve": ts,
        "t_analyze": t_an, "t_plan": t_plan, "X": np.asarray(Xh), "classes": classes,
        "refine": info, "info": L.info(), "tiles": tiles,
    }


def class_roofline(g, precision: str) -> dict | None:
    """Roofline of the dominant MFMA kernel class (descendant updates or
    in-supernode trailing updates, whichever takes longer per factorization):
    achieved = that class's algorithmic flops per launch / its average launch
    time from HIP events; peak = the FP64 or FP32 matrix peak."""
    c = g["classes"]
    if not c:
        return None
    L = g["L"]
    kf = L.kernel_flops()
    steps = c["steps"]
    cand = {"update": (c["updates"], c["launches"]["update"], kf["update"]),
            "trailing": (c["trailing"], c["launches"]["trailing"], kf["trailing"])}
    name = max(cand, key=lambda k: cand[k][0])
    ms, nl, fl = cand[name]
    ms_step, nl_step = ms / steps, nl / steps
    avg_ms = ms_step / max(nl_step, 1)
    achieved = fl / (ms_step / 1e3) / 1e12 if ms_step > 0 else 0.0
    peak = MFMA_PEAK_TFLOPS["f64" if precision == "fp64" else "f32"]
    kern = {("fp64", "update"): "k_tile<double, 0, *> (descendant update + scatter)",
            ("fp64", "trailing"): "k_tile<double, 1, *> (in-supernode trailing update)",
            ("fp32", "update"): "k_tile_h<0, *> / k_tile<float, 0, 32> (descendant update + scatter)",
            ("fp32", "trailing"): "k_tile_h<1, *> / k_tile<float, 1, 32> (in-supernode trailing update)"}
    other = [k for k in cand if k != name][0]
    return {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": None,
            "kernel": kern[(precision, name)], "class": name,
            "launches_per_step": nl_step, "avg_launch_ms": round(avg_ms, 4),
            "flops_per_step": fl, "flops_per_launch": fl / max(nl_step, 1),
            "class_ms_per_step": round(ms_step, 3),
            f"{other}_ms_per_step": round(cand[other][0] / steps, 3),
            f"{other}_tflops": round(cand[other][2] / (cand[other][0] / steps / 1e3) / 1e12, 3)
            if cand[other][0] > 0 else None}


def host_cpu_info() -> dict:
    """The host CPU the baseline ran on: lscpu model / sockets / cores, the CPUs
    this process may use (affinity) and the cgroup CPU quota, if any."""
    info = {}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k] = v
    except Exception as e:          # lscpu missing
        info["lscpu_error"] = str(e)
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity_cpus"] = os.cpu_count()
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()
        info["cgroup_cpu_quota"] = None if q == "max" else round(int(q) / int(per), 2)
    except Exception:
        info["cgroup_cpu_quota"] = None
    return info


def baseline_threads(cpu: dict) -> int:
    """Every CPU this job may use: OMP_NUM_THREADS when set (the host's share
    for this GPU), else the affinity set capped by the cgroup quota."""
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    if env > 0:
        return env
    n = cpu.get("affinity_cpus") or 1
    q = cpu.get("cgroup_cpu_quota")
    return max(1, min(n, int(q))) if q else n


def cpu_baseline(A, perm, B, fl, lnz, threads, reps=3, budget_s=20.0, fp32=True):
    """Reference CHOLMOD (FP64, MKL, OpenMP) on the host: factor + solve, and its
    single-precision supernodal factor + FP64 refinement."""
    from oracle import ref
    ref.set_threads(threads)
    F = ref.RefFactor(A, ref.GIVEN, perm=perm)
    tf, ts = [], []
    t0 = time.perf_counter()
    X = None
    while len(tf) < reps and (time.perf_counter() - t0 < budget_s or not tf):
        F.factorize()
        tf.append(F.t_factor)
        X = F.solve(B)
        ts.append(F.t_solve)
    t = float(np.median(tf)) + float(np.median(ts))
    nrhs = B.shape[1]
    res = {
        "value": (fl + 4.0 * lnz * nrhs) / t / 1e9, "unit": "GFLOP/s", "cores": threads,
        "kind": "reference",
        "sample": f"{len(tf)} factor+solve run(s) of the full system, same ordering, median; "
                  f"analyze excluded (t_analyze={F.t_analyze:.2f}s)",
        "t_factor_s": float(np.median(tf)), "t_solve_s": float(np.median(ts)),
        "time_to_solution_s": t, "factor_gflops": fl / float(np.median(tf)) / 1e9,
        "relres": relres(A, B, X), "X": X,
    }
    del F
    if fp32:
        res["fp32_ir"] = cpu_fp32_ir(A, perm, B, ref)
        res["fp32_ir"]["gflops"] = (fl + 4.0 * lnz * nrhs) / res["fp32_ir"]["time_to_solution_s"] / 1e9
    return res


def cpu_fp32_ir(A, perm, B, ref, tol=1e-10, maxiter=10):
    """The reference's single-precision supernodal factor (rs_ worker,
    cholmod_super_numeric.c:297-299) + FP64 refinement on the host: r = b - A x
    in double (the sdmult of the demo's refinement, cholmod_dl_demo.c:627-663),
    correction solves through the single factor.  One timed factorization."""
    F = ref.RefFactor(A, ref.GIVEN, perm=perm, single=True)
    F.factorize()
    t_f = F.t_factor
    t0 = time.perf_counter()
    X = F.solve(B)
    rr0 = rr = relres(A, B, X)
    it = 0
    while rr > tol and it < maxiter:
        X = X + F.solve(B - A.matvec(X))
        rr = relres(A, B, X)
        it += 1
    t_s = time.perf_counter() - t0
    return {"t_factor_s": t_f, "t_solve_refine_s": t_s, "time_to_solution_s": t_f + t_s,
            "iterations": it, "relres0": rr0, "relres": rr}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--roof-steps", type=int, default=2,
                    help="instrumented steps after the timed ones (per-class HIP events)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--side", type=int, default=96, help="side of the 27-point headline grid")
    ap.add_argument("--only-headline", action="store_true",
                    help="only the 96^3 FP64 vs FP32+IR headline (profiling)")
    ap.add_argument("--only-64", action="store_true", help="only the 64^3 FP64 record (profiling)")
    ap.add_argument("--no-64", action="store_true", help="skip the 64^3 FP64 record")
    ap.add_argument("--tiles", choices=["saved", "auto", "default"], default="saved",
                    help="tile table: the saved autotuned one (default), retune, or built-in")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched workload (4,096 2D systems, FP64 and FP32+IR)")
    ap.add_argument("--only-batch", action="store_true", help="run only the batched workload")
    ap.add_argument("--only-serving", action="store_true", help="run only the serving stream (--serving N)")
    ap.add_argument("--serving", type=int, default=20000, metavar="N",
                    help="serving stream of N requests (FP32+IR; 0: none), in a child process")
    ap.add_argument("--serving-rate", type=float, default=100.0, help="Poisson arrival rate (1/s)")
    ap.add_argument("--serving-slots", type=int, default=8)
    ap.add_argument("--serving-budget-gib", type=float, default=64.0)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run one device copy of this size first (counter calibration)")
    ap.add_argument("--traffic-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="no GPU: turn two rocprofv3 --pmc outputs into the traffic summary")
    ap.add_argument("--counter-summary", nargs="+", metavar="PMC_DIR",
                    help="no GPU: turn rocprofv3 --pmc outputs of one command into a per-class counter table")
    ap.add_argument("--counter-title", default="")
    ap.add_argument("--counter-cmd", action="append", default=[], help="a command line to quote (repeatable)")
    ap.add_argument("--counter-out", default="")
    ap.add_argument("--traffic-cmd", default="python3 bench.py --steps 1 --warmup 1 --no-cpu "
                    "--only-64 --calib-copy-mib 1024")
    args = ap.parse_args()
    if args.counter_summary:
        counter_summary(args.counter_summary, args.calib_copy_mib or 1024, args.counter_title, args.counter_cmd,
                        pathlib.Path(args.counter_out))
        return
    if args.traffic_summary:
        r = traffic_summary(args.traffic_summary[0], args.traffic_summary[1],
                            args.calib_copy_mib or 1024, args.traffic_cmd)
        print(json.dumps(r))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(the factorization would shard along the elimination tree: independent subtrees per GPU, "
              "then one point-to-point exchange of the separator updates)",
              file=sys.stderr)
        sys.exit(2)

    if args.calib_copy_mib:
        calib_copy(args.calib_copy_mib)

    if args.only_batch:
        out = {"metric": "batched factor+solve (4,096 2D systems)", "value": None}
        batch_section(args, out)
        print(json.dumps(out))
        return
    if args.only_serving:
        out = {"metric": "serving: p50/p99 latency and solves/s (FP32+IR to 1e-10)", "value": None}
        serving_section(args, out)
        print(json.dumps(out))
        return
    if args.only_64:
        out = {"metric": "factor+solve GFLOP/s (FP64, 3D 7-pt 64^3, METIS)", "value": None}
        out["diffusion3d_64_fp64"] = section_64(args)
        print(json.dumps(out))
        return

    out = headline_section(args)
    if not args.only_headline:
        if not args.no_64:
            out["diffusion3d_64_fp64"] = section_64(args)
        if not args.no_batch:
            batch_section(args, out)
        if args.serving:
            serving_child(args, out)
    print(json.dumps(out))


def headline_section(args) -> dict:
    """The metric's own workload: 3D 27-point 96^3 (n = 884,736), METIS, 4 RHS.
    Plain FP64 factor + solve against FP32 factor + FP64 refinement to a 1e-10
    relative residual; value = FP32+IR factor+solve GFLOP/s (fl + 4 lnz nrhs per
    time-to-solution), beside the FP64 figure, the speedup, solves/s and the
    p50/p99 step latency."""
    from suitesparse_amd import problems
    side = args.side
    log(f"[bench] generating {side}^3 27-point stencil")
    A3, B3, _ = problems.workload_stencil27_96(side)
    from suitesparse_amd import cholmod
    log(f"[bench] {side}^3 METIS ordering + analysis (CPU)")
    perm3, t_order = product_ordering(A3, cholmod.CHOLMOD_METIS)
    nrhs = B3.shape[1]
    res = {}
    for prec in ("fp64", "fp32"):
        log(f"[bench] {side}^3 {prec}")
        r = gpu_factor_solve(A3, perm3, B3, prec, args.steps, args.warmup, refine=(prec == "fp32"),
                             tune=(f"stencil27_{side}", args.tiles), roof_steps=args.roof_steps)
        L = r["L"]
        tts = r["t_total"] / args.steps
        lat = np.array(r["lat"]) * 1e3
        res[prec] = {"time_to_solution_ms": round(tts * 1e3, 3),
                     "t_factor_ms": round(r["t_factor"] * 1e3, 3),
                     "t_solve_ms": round(r["t_solve"] * 1e3, 3),
                     "factor_gflops": round(L.fl / r["t_factor"] / 1e9, 1),
                     "factor_solve_gflops": round((L.fl + 4 * L.lnz * nrhs) / tts / 1e9, 1),
                     "solves_per_s": round(1.0 / tts, 3),
                     "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                                    "p99": round(float(np.percentile(lat, 99)), 3)},
                     "relres": relres(A3, B3, r["X"]), "backward_error": backward_error(A3, B3, r["X"]),
                     "refine": r["refine"], "tiles": r["tiles"], "t_analyze_s": round(r["t_analyze"], 2),
                     "t_plan_s": round(r["t_plan"], 2), "roofline": class_roofline(r, prec)}
        res[prec]["_X"] = r["X"]
        fl, lnz = L.fl, L.lnz
        log(f"[bench]   {({k: v for k, v in res[prec].items() if k != '_X'})}")
        L.free()
        del r, L
        import gc
        gc.collect()
    f64, f32 = res["fp64"], res["fp32"]
    speedup = f64["time_to_solution_ms"] / f32["time_to_solution_ms"]
    out = {
        "metric": "factor+solve GFLOP/s & solves/s: FP64 vs FP32+IR at 1e-10 residual; p99 latency",
        "value": f32["factor_solve_gflops"], "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": f32["time_to_solution_ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64ir", "data": "synthetic",
        "config": {"workload": f"stencil3d_27pt_{side}^3_metis_4rhs_fp32ir_vs_fp64", "n": A3.n,
                   "nnz": A3.nnz_full, "fl": fl, "lnz": lnz, "nrhs": nrhs},
        "t_order_metis_s": round(t_order, 2),
        "solves_per_s": f32["solves_per_s"], "latency_ms": f32["latency_ms"],
        "fp64_value": f64["factor_solve_gflops"], "fp64_ms_per_step": f64["time_to_solution_ms"],
        "fp64_solves_per_s": f64["solves_per_s"], "fp64_latency_ms": f64["latency_ms"],
        "speedup_fp32ir_vs_fp64": round(speedup, 3),
        "relres": f32["relres"], "fp64_relres": f64["relres"],
        "roofline": f32["roofline"], "fp64_roofline": f64["roofline"],
    }
    X64, X32 = res["fp64"].pop("_X"), res["fp32"].pop("_X")
    out["fwd_err_fp32ir_vs_gpu_fp64"] = float(np.linalg.norm(X32 - X64) / np.linalg.norm(X64))
    out["fp32ir"] = f32
    out["fp64"] = f64
    if not args.no_cpu:
        cpu = host_cpu_info()
        thr = baseline_threads(cpu)
        log(f"[bench] reference CPU baseline, {side}^3 ({thr} threads, {cpu})")
        try:
            cb = cpu_baseline(A3, perm3, B3, fl, lnz, thr, reps=1, budget_s=0.0, fp32=True)
            Xr = cb.pop("X")
            out["fwd_err_vs_ref"] = float(np.linalg.norm(X32 - Xr) / np.linalg.norm(Xr))
            out["fp64_fwd_err_vs_ref"] = float(np.linalg.norm(X64 - Xr) / np.linalg.norm(Xr))
            cb["host"] = cpu
            # the baseline's headline value is the reference's own FP32 + IR path
            cb["fp64_value"] = cb["value"]
            cb["value"] = cb["fp32_ir"]["gflops"]
            cb["sample"] = ("one factor+solve of the full 96^3 system per precision (reference FP64 "
                            "cholmod_l_factorize + cholmod_l_solve, and its single-precision factor + "
                            "FP64 refinement to 1e-10); value = the FP32+IR path's (fl + 4 lnz nrhs) / "
                            "time-to-solution, analyze excluded")
            out["cpu_baseline"] = cb
            out["speedup_vs_cpu"] = round(out["value"] / cb["value"], 2)
            log(f"[bench] cpu: {cb['value']:.1f} GFLOP/s FP32+IR, {cb['fp64_value']:.1f} FP64")
        except Exception as e:   # reference library missing on this host
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    return out


def section_64(args) -> dict:
    """3D 7-point diffusion 64^3 (n = 262,144), METIS, FP64, 1 RHS."""
    from suitesparse_amd import problems
    log("[bench] generating 64^3 7-point diffusion")
    A, B, xt = problems.workload_diffusion3d_64()
    from suitesparse_amd import cholmod
    log("[bench] 64^3 METIS ordering + analysis (CPU)")
    perm, t_order = product_ordering(A, cholmod.CHOLMOD_METIS)
    log("[bench] FP64 factor+solve on GPU")
    g = gpu_factor_solve(A, perm, B, "fp64", args.steps, args.warmup,
                         tune=("diffusion3d_64", args.tiles), roof_steps=args.roof_steps)
    L = g["L"]
    fl, lnz, nrhs = L.fl, L.lnz, B.shape[1]
    ms = g["t_total"] / args.steps * 1e3
    value = (fl + 4.0 * lnz * nrhs) / (g["t_total"] / args.steps) / 1e9
    lat = np.array(g["lat"]) * 1e3
    roof = class_roofline(g, "fp64")
    sec = {"workload": "diffusion3d_7pt_64^3_metis_fp64_1rhs", "n": A.n, "nnz": A.nnz_full,
           "fl": fl, "lnz": lnz, "nrhs": nrhs, "t_order_metis_s": round(t_order, 2),
           "value_gflops": round(value, 2), "ms_per_step": round(ms, 3),
           "solves_per_s": round(1e3 / ms, 3),
           "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                          "p99": round(float(np.percentile(lat, 99)), 3)},
           "t_factor_ms": round(g["t_factor"] * 1e3, 3), "t_solve_ms": round(g["t_solve"] * 1e3, 3),
           "factor_gflops": round(fl / g["t_factor"] / 1e9, 2), "tiles": g["tiles"],
           "relres": relres(A, B, g["X"]), "fwd_err_vs_x_true": float(np.linalg.norm(g["X"] - xt) /
                                                                    np.linalg.norm(xt)),
           "backward_error": backward_error(A, B, g["X"]), "roofline": roof}
    if roof is not None:
        c = g["classes"]
        kf = L.kernel_flops()
        upd_ms = c["updates"] / c["steps"]
        sec["update_kernel"] = {"tflops": round(kf["update"] / (upd_ms / 1e3) / 1e12, 3) if upd_ms else None,
                                "traffic_bytes_per_launch": read_traffic(),
                                "launches_per_step": c["launches"]["update"] / c["steps"]}
    log(f"[bench] 64^3 fp64: {ms:.1f} ms/step, {value:.1f} GFLOP/s")
    sec["modify"] = modify_record(A, perm, L, args)
    L.free()
    if not args.no_cpu:
        cpu = host_cpu_info()
        thr = baseline_threads(cpu)
        try:
            cb = cpu_baseline(A, perm, B, fl, lnz, thr, reps=3, budget_s=20.0, fp32=False)
            Xr = cb.pop("X")
            sec["fwd_err_vs_ref"] = float(np.linalg.norm(g["X"] - Xr) / np.linalg.norm(Xr))
            sec["cpu_baseline"] = cb
            sec["speedup_vs_cpu"] = round(value / cb["value"], 2)
        except Exception as e:
            sec["cpu_baseline"] = {"value": None, "error": str(e)}
    return sec


def modify_columns(A, perm, ks, scale=0.1):
    """C whose column c is column ks[c] of tril(P A P') times scale: inside the
    pattern of L(:, ks[c]), so the device modifies the resident factor."""
    import scipy.sparse as sp
    Al = sp.csc_matrix((A.x, A.i, A.p), shape=(A.n, A.n))
    Af = (Al + sp.tril(Al, -1).T).tocsc()
    Pa = sp.tril(Af[perm][:, perm]).tocsc()
    Pa.sort_indices()
    cols = [Pa[:, int(k)] for k in ks]
    C = sp.hstack(cols).tocsc() * scale
    C.sort_indices()
    return (C.indptr.astype(np.int64), C.indices.astype(np.int64), C.data.astype(np.float64))


def modify_record(A, perm, L, args) -> dict:
    """cholmod_l_updown on the resident 64^3 factor: rank-1 and rank-8 update
    followed by the same downdate (the factor comes back to A's), wall time per
    call (host plan + kernels; the call returns when L is modified).  The
    reference times the same calls on the host cores (its first call also
    converts the supernodal factor to simplicial LDL')."""
    import torch
    from suitesparse_amd import cholmod
    cm = cholmod.Common()
    rng = np.random.default_rng(31)
    rec = {"workload": "update then downdate of the 64^3 factor by columns of tril(PAP') x 0.1, "
                       "C in L's order, FP64"}
    Cs = {k: modify_columns(A, perm, rng.integers(0, A.n // 2, k)) for k in (1, 8)}
    for k, C in Cs.items():
        ts = []
        for _ in range(3):
            for upd in (True, False):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                ok = cholmod.updown(upd, C, L, cm)
                ts.append(time.perf_counter() - t0)
                assert ok and cm.status == 0, cm.status
        rec[f"rank{k}"] = {"ms_per_call": round(float(np.median(ts)) * 1e3, 3), "modfl": cm.modfl,
                           "gflops": round(cm.modfl / float(np.median(ts)) / 1e9, 2)}
    log(f"[bench] 64^3 updown: rank-1 {rec['rank1']['ms_per_call']} ms, rank-8 {rec['rank8']['ms_per_call']} ms")
    if not args.no_cpu:
        try:
            from oracle import ref
            cpu = host_cpu_info()
            ref.set_threads(baseline_threads(cpu))
            R = ref.RefFactor(A, ref.GIVEN, perm=perm)
            R.factorize()
            first = R.updown(True, *Cs[1])["t"]
            t1 = R.updown(False, *Cs[1])["t"]
            t8 = [R.updown(u, *Cs[8])["t"] for u in (True, False)]
            rec["cpu_baseline"] = {"kind": "reference", "rank1_first_call_s": round(first, 4),
                                   "rank1_s": round(t1, 4), "rank8_s": round(float(np.median(t8)), 4),
                                   "sample": "the same C on the same factor; the first call converts L to "
                                             "simplicial LDL'"}
            rec["speedup_rank8_vs_cpu"] = round(float(np.median(t8)) / (rec["rank8"]["ms_per_call"] / 1e3), 2)
        except Exception as e:
            rec["cpu_baseline"] = {"error": str(e)}
    return rec


def batch_section(args, out):
    """4,096 independent 2D 5-point systems (side U{16..128}, log-normal weights,
    AMD), 4 RHS each, factored and solved as one block-diagonal forest: FP64,
    and FP32 + per-system FP64 refinement to 1e-10."""
    import gc
    import torch
    from suitesparse_amd import batch, cholmod, problems
    log("[bench] batched: generating 4096 2D systems")
    sides, rng = problems.batch_sides(4096)
    mats = [problems.grid2d_lognormal(int(s), rng) for s in sides]
    nt = sum(m.n for m in mats)
    B = np.asfortranarray(rng.uniform(-1.0, 1.0, (nt, 4)))
    t0 = time.perf_counter()
    cm = cholmod.Common(precision="fp64")
    b = batch.analyze_batch(mats, cm)
    t_an = time.perf_counter() - t0
    assert cm.start()
    L = b.L
    dev = torch.device("cuda:0")
    Ax = torch.as_tensor(b.A.x, device=dev)
    Bd = torch.as_tensor(B, device=dev).t().contiguous().t()
    seg = b.offsets
    res = {"workload": "batch_4096_grid2d_side16..128_amd_4rhs", "systems": len(mats), "n_total": nt,
           "fl": L.fl, "lnz": L.lnz, "t_analyze_s": round(t_an, 2)}
    steps = max(1, args.steps)
    for prec in ("fp64", "fp32"):
        cmp = cholmod.Common(precision=prec)
        cmp.start()
        t0 = time.perf_counter()
        L._numeric(cmp)
        t_plan = time.perf_counter() - t0

        def step():
            ok = cholmod.factorize(b.A, L, cmp, values=Ax)
            assert ok and cmp.status == 0, cmp.status
            if prec == "fp32":
                return cholmod.solve_refine(L, Bd, cmp, tol=1e-10, maxiter=20, segments=seg)
            return cholmod.solve(cholmod.CHOLMOD_A, L, Bd, cmp), None

        step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            X, info = step()
        torch.cuda.synchronize()
        tt = (time.perf_counter() - t0) / steps
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        cholmod.factorize(b.A, L, cmp, values=Ax)
        torch.cuda.synchronize()
        tf = time.perf_counter() - t1
        Xh = X.t().contiguous().t().cpu().numpy()
        R = B - b.A.matvec(Xh)
        rn = np.sqrt(np.add.reduceat(R * R, seg[:-1], axis=0))
        bn = np.sqrt(np.add.reduceat(B * B, seg[:-1], axis=0))
        rel = float((rn / bn).max())
        res[prec] = {"time_ms": round(tt * 1e3, 2), "t_factor_ms": round(tf * 1e3, 2),
                     "t_solve_ms": round((tt - tf) * 1e3, 2), "t_plan_s": round(t_plan, 2),
                     "systems_per_s": round(len(mats) / tt, 1),
                     "factor_gflops": round(L.fl / tf / 1e9, 1),
                     "max_relres_per_system": rel, "refine": info}
        if prec == "fp64":
            # the same factorization with each level's small supernodes factored
            # whole, one wave each, in one launch per level (k_small_level)
            L.set_small_fused(True)
            cholmod.factorize(b.A, L, cmp, values=Ax)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for _ in range(steps):
                cholmod.factorize(b.A, L, cmp, values=Ax)
            torch.cuda.synchronize()
            res[prec]["t_factor_ms_small_fused"] = round((time.perf_counter() - t1) / steps * 1e3, 2)
            L.set_small_fused(False)
        log(f"[bench]   batch {prec}: {res[prec]}")
        del X, Xh, R
        gc.collect()
    res["speedup_fp32ir_vs_fp64"] = round(res["fp64"]["time_ms"] / res["fp32"]["time_ms"], 3)
    out["batched"] = res


def serving_child(args, out):
    """The serving stream in a child process (this script with --only-serving),
    so that whatever happens there cannot take the bench line with it."""
    import subprocess
    cmd = [sys.executable, "-X", "faulthandler", str(ROOT / "bench.py"), "--only-serving",
           "--serving", str(args.serving), "--serving-rate", str(args.serving_rate),
           "--serving-slots", str(args.serving_slots), "--serving-budget-gib", str(args.serving_budget_gib)]
    log(f"[bench] serving child: {' '.join(cmd[2:])}")
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, text=True)
    try:
        out["serving"] = json.loads(p.stdout.strip().splitlines()[-1])["serving"]
    except Exception as e:
        out["serving"] = {"error": f"serving child exited {p.returncode}: {e!r}"}


def serving_section(args, out):
    """Poisson stream of requests (2D 5-point side 32..512 / 3D 7-point side
    16..64, log-normal weights, 30% same-pattern refactorizations, 1..8 RHS),
    FP32 + FP64 refinement to 1e-10, admission by device bytes against a
    fixed HBM budget; latency is arrival -> result on the device."""
    from suitesparse_amd import serving
    n = args.serving
    log(f"[bench] serving: {n} requests at {args.serving_rate}/s, {args.serving_slots} slots")
    reqs = serving.request_specs(n)
    arrivals = serving.poisson_arrivals(n, args.serving_rate)
    srv = serving.Server(slots=args.serving_slots, hbm_budget=int(args.serving_budget_gib * (1 << 30)),
                         precision="fp32", tol=1e-10, keep_results=False)
    st = serving.run_stream(srv, reqs, arrivals, progress=1000)
    st.update({"workload": f"serving_{n}req_poisson{args.serving_rate:g}_2d32..512amd_3d16..64metis_fp32ir",
               "rate_per_s": args.serving_rate, "slots": args.serving_slots,
               "hbm_budget_gib": args.serving_budget_gib})
    log(f"[bench]   serving: {st}")
    out["serving"] = st


if __name__ == "__main__":
    main()
