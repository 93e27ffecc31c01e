"""Update/downdate cases shared by tests/test_updown.py (oracle, CPU) and
tests/test_gpu_updown.py (device): matrices C whose entries stay inside the
pattern of the factor, rows in the factor's order."""
import numpy as np


def sym_of(L) -> dict:
    """The symbolic arrays of a cholmod.Factor in the oracle's keys."""
    return {"Perm": L.Perm, "Super": L.super, "Lpi": L.pi, "Lpx": L.px, "Ls": L.s}


def col_super(sym) -> np.ndarray:
    Super = np.asarray(sym["Super"])
    return np.repeat(np.arange(len(Super) - 1), np.diff(Super))


def pattern_columns(sym, cols, rng, dens=0.6, scale=1.0, dup=False):
    """One column of C per entry of cols: its first row k = cols[c], then a
    random subset of the rows of L(:, k); values N(0, scale^2).  dup repeats
    some rows (a later value wins, as cholmod_updown reads C)."""
    Super, Lpi, Ls = (np.asarray(sym[k]) for k in ("Super", "Lpi", "Ls"))
    cs = col_super(sym)
    Cp, Ci, Cx = [0], [], []
    for k in cols:
        k = int(k)
        s = cs[k]
        rows = Ls[Lpi[s] + (k - Super[s]):Lpi[s + 1]]
        keep = [k] + [int(r) for r in rows[1:] if rng.random() < dens]
        if dup and len(keep) > 2:
            keep = sorted(keep + list(rng.choice(keep[1:], 2)))
        Ci += keep
        Cx += list(scale * rng.standard_normal(len(keep)))
        Cp.append(len(Ci))
    return np.array(Cp, np.int64), np.array(Ci, np.int64), np.array(Cx)


def dense_c(n, Cp, Ci, Cx) -> np.ndarray:
    """C as a dense n x k array (a later repeated entry wins)."""
    k = len(Cp) - 1
    D = np.zeros((n, k))
    for c in range(k):
        for p in range(Cp[c], Cp[c + 1]):
            D[Ci[p], c] = Cx[p]
    return D
