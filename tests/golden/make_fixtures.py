"""Write the golden fixtures under tests/golden/ from the reference CHOLMOD.

Runs in the build container (needs oracle/_ref/libssref.so, built from the
reference sources by oracle/ref.mk).  For each benchmark workload it records
the symbolic scalars and the SHA-256 digests of Perm/ColCount/Super/Lpi/Lpx/Ls
of the reference's analysis (AMD or METIS), which the product's own ordering
and analysis must reproduce.  Small cases also keep the reference's factor and
solution.

    python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import hashlib
import json
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from oracle import ref  # noqa: E402
from suitesparse_amd import problems  # noqa: E402


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.int64).tobytes()).hexdigest()


def symbolic_record(F) -> dict:
    s = F.symbolic()
    return {
        "n": F.n, "nsuper": F.nsuper, "ssize": F.ssize, "xsize": F.xsize,
        "maxcsize": F.maxcsize, "maxesize": F.maxesize, "fl": F.fl, "lnz": F.lnz,
        "sha256": {k: digest(s[k]) for k in ("Perm", "ColCount", "Super", "Lpi", "Lpx", "Ls")},
    }, s


def main():
    ref.set_threads(8)
    out = {}
    work = [
        ("laplace2d_100_amd", problems.laplace2d_5pt(100), ref.AMD),
        ("diffusion3d_64_metis", problems.diffusion3d_7pt(64, seed=1), ref.METIS),
        ("stencil27_96_metis", problems.stencil3d_27pt(96, seed=2), ref.METIS),
    ]
    for name, A, ordering in work:
        print("analyzing", name, flush=True)
        F = ref.RefFactor(A, ordering)
        rec, s = symbolic_record(F)
        rec["ordering"] = {ref.AMD: "amd", ref.METIS: "metis"}[ordering]
        rec["t_analyze_s"] = F.t_analyze
        out[name] = rec
        del F
    (HERE / "symbolic_golden.json").write_text(json.dumps(out, indent=1))

    # small numeric fixtures: reference factor and solution
    rng = np.random.default_rng(0)
    A = problems.grid2d_lognormal(12, rng)
    F = ref.RefFactor(A, ref.AMD)
    s = F.symbolic()
    F.factorize()
    B = np.random.default_rng(1).uniform(-1, 1, (A.n, 2))
    X = F.solve(B)
    np.savez_compressed(HERE / "small_grid2d_12.npz", p=A.p, i=A.i, x=A.x, B=B, X=X, Lx=F.Lx(),
                        **{k: s[k] for k in ("Perm", "Super", "Lpi", "Lpx", "Ls")})
    print("done")


if __name__ == "__main__":
    main()
