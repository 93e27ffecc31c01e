"""Seeded SPD workload generators (grid stencils) for the supernodal Cholesky path.

Every generator returns the LOWER triangle of a symmetric positive definite
matrix in compressed-column form with int64 indices -- the storage CHOLMOD's
int64 interface takes for ``stype = -1`` (CHOLMOD/Include/cholmod.h:1168-1206,
``cholmod_sparse``) -- plus the right-hand sides the workload prescribes.

Node numbering is lexicographic: ``id = x + s*y`` (2D) and ``id = x + s*y + s*s*z``
(3D).  Randomness comes from ``numpy.random.default_rng(seed)`` (PCG64), so the
same seed yields the same matrix on every machine.

Workloads (sizes as named in the benchmark table):
  * ``laplace2d_5pt(100)``        2D 5-point Laplacian, A_ii = 4, A_ij = -1,
                                  Dirichlet (missing neighbours dropped), seed 0.
  * ``diffusion3d_7pt(64)``       3D 7-point variable-coefficient diffusion,
                                  k = exp(g), g ~ N(0,1) per cell (seed 1), face
                                  weight = harmonic mean 2 k_i k_j / (k_i + k_j);
                                  a Dirichlet boundary face adds k_i to A_ii.
  * ``stencil3d_27pt(96)``        3D 27-point, one weight w = exp(N(0, 0.5^2)) per
                                  undirected edge (seed 2), A_ij = -w,
                                  A_ii = sum_j |A_ij| + 1e-2.
  * ``grid2d_lognormal(side,...)`` 2D 5-point, w = exp(N(0, 0.5^2)), A_ii = sum + shift.
  * ``grid3d_lognormal(side,...)`` 3D 7-point, same weights.
"""
from __future__ import annotations

import dataclasses
import numpy as np


@dataclasses.dataclass
class SymMatrix:
    """Lower triangle (i >= j) of a symmetric matrix, CSC, int64 indices."""
    n: int
    p: np.ndarray      # int64[n+1]
    i: np.ndarray      # int64[nnz]
    x: np.ndarray      # float64[nnz]
    name: str = ""

    @property
    def nnz_lower(self) -> int:
        return int(self.p[-1])

    @property
    def nnz_full(self) -> int:
        return 2 * int(self.p[-1]) - self.n

    def to_scipy_full(self):
        import scipy.sparse as sp
        L = sp.csc_matrix((self.x, self.i, self.p), shape=(self.n, self.n))
        return (L + L.T - sp.diags(L.diagonal())).tocsc()

    def matvec(self, X: np.ndarray) -> np.ndarray:
        """A @ X for the full symmetric matrix (X is n or n x k)."""
        return self.to_scipy_full() @ X


def _lower_csc(n: int, ei: np.ndarray, ej: np.ndarray, w: np.ndarray,
               diag: np.ndarray, name: str) -> SymMatrix:
    """Assemble tril(A) from off-diagonal edges (ei > ej, value -w) and a diagonal."""
    rows = np.concatenate([np.arange(n, dtype=np.int64), ei.astype(np.int64)])
    cols = np.concatenate([np.arange(n, dtype=np.int64), ej.astype(np.int64)])
    vals = np.concatenate([diag.astype(np.float64), -w.astype(np.float64)])
    order = np.lexsort((rows, cols))
    rows, cols, vals = rows[order], cols[order], vals[order]
    counts = np.bincount(cols, minlength=n)
    p = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=p[1:])
    return SymMatrix(n=n, p=p, i=rows, x=vals, name=name)


def _grid_edges(shape: tuple[int, ...], offsets: list[tuple[int, ...]]):
    """Edges (hi, lo) between grid nodes for each 'forward' offset, in offset order.

    ``shape`` is (sx, sy[, sz]); an offset is (dx, dy[, dz]).  Only offsets whose
    neighbour id is larger are passed, so each undirected edge appears once.
    """
    dims = len(shape)
    coords = np.indices(shape[::-1]).reshape(dims, -1)[::-1]   # coords[0] = x fastest
    strides = [1]
    for s in shape[:-1]:
        strides.append(strides[-1] * s)
    ids = np.arange(int(np.prod(shape)), dtype=np.int64)
    his, los = [], []
    for off in offsets:
        ok = np.ones(ids.shape, dtype=bool)
        nb = ids.copy()
        for d in range(dims):
            c = coords[d] + off[d]
            ok &= (c >= 0) & (c < shape[d])
            nb += off[d] * strides[d]
        lo = ids[ok]
        hi = nb[ok]
        assert np.all(hi > lo)
        his.append(hi)
        los.append(lo)
    return his, los


def _forward_offsets(dims: int, full: bool) -> list[tuple[int, ...]]:
    """Neighbour offsets with a positive linear id (x fastest), in a fixed order."""
    if dims == 2:
        cand = [(1, 0), (0, 1)] if not full else [(dx, dy) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
    else:
        if not full:
            cand = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
        else:
            cand = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
    out = []
    for off in cand:
        # linear id increment sign: z dominates, then y, then x
        key = off[::-1]
        if any(key) and next(v for v in key if v != 0) > 0:
            out.append(off)
    return out


def laplace2d_5pt(side: int = 100) -> SymMatrix:
    """2D 5-point Laplacian, A_ii = 4, A_ij = -1, Dirichlet boundary."""
    n = side * side
    his, los = _grid_edges((side, side), _forward_offsets(2, False))
    ei, ej = np.concatenate(his), np.concatenate(los)
    return _lower_csc(n, ei, ej, np.ones(ei.size), np.full(n, 4.0), f"laplace2d_5pt_{side}")


def diffusion3d_7pt(side: int = 64, seed: int = 1) -> SymMatrix:
    """3D 7-point variable-coefficient diffusion with harmonic-mean face weights."""
    rng = np.random.default_rng(seed)
    n = side ** 3
    k = np.exp(rng.standard_normal(n))
    his, los = _grid_edges((side, side, side), _forward_offsets(3, False))
    ei, ej = np.concatenate(his), np.concatenate(los)
    w = 2.0 * k[ei] * k[ej] / (k[ei] + k[ej])
    diag = np.zeros(n)
    np.add.at(diag, ei, w)
    np.add.at(diag, ej, w)
    # Dirichlet: each of the 6 faces of a cell with no neighbour adds k_i
    nnb = np.zeros(n, dtype=np.int64)
    np.add.at(nnb, ei, 1)
    np.add.at(nnb, ej, 1)
    diag += (6 - nnb) * k
    return _lower_csc(n, ei, ej, w, diag, f"diffusion3d_7pt_{side}")


def stencil3d_27pt(side: int = 96, seed: int = 2, sigma: float = 0.5,
                   shift: float = 1e-2) -> SymMatrix:
    """3D 27-point SPD stencil, w = exp(N(0, sigma^2)) per edge, diag = sum|offdiag| + shift."""
    rng = np.random.default_rng(seed)
    n = side ** 3
    his, los = _grid_edges((side, side, side), _forward_offsets(3, True))
    ei, ej = np.concatenate(his), np.concatenate(los)
    w = np.exp(sigma * rng.standard_normal(ei.size))
    diag = np.full(n, shift)
    np.add.at(diag, ei, w)
    np.add.at(diag, ej, w)
    return _lower_csc(n, ei, ej, w, diag, f"stencil3d_27pt_{side}")


def grid2d_lognormal(side: int, rng: np.random.Generator, sigma: float = 0.5,
                     shift: float = 1e-3) -> SymMatrix:
    """2D 5-point grid, w = exp(N(0, sigma^2)) per edge, diag = sum + shift."""
    n = side * side
    his, los = _grid_edges((side, side), _forward_offsets(2, False))
    ei, ej = np.concatenate(his), np.concatenate(los)
    w = np.exp(sigma * rng.standard_normal(ei.size))
    diag = np.full(n, shift)
    np.add.at(diag, ei, w)
    np.add.at(diag, ej, w)
    return _lower_csc(n, ei, ej, w, diag, f"grid2d_lognormal_{side}")


def grid3d_lognormal(side: int, rng: np.random.Generator, sigma: float = 0.5,
                     shift: float = 1e-3) -> SymMatrix:
    """3D 7-point grid, w = exp(N(0, sigma^2)) per edge, diag = sum + shift."""
    n = side ** 3
    his, los = _grid_edges((side, side, side), _forward_offsets(3, False))
    ei, ej = np.concatenate(his), np.concatenate(los)
    w = np.exp(sigma * rng.standard_normal(ei.size))
    diag = np.full(n, shift)
    np.add.at(diag, ei, w)
    np.add.at(diag, ej, w)
    return _lower_csc(n, ei, ej, w, diag, f"grid3d_lognormal_{side}")


def revalue(A: SymMatrix, rng: np.random.Generator, sigma: float = 0.5,
            shift: float = 1e-3) -> SymMatrix:
    """Same pattern, redrawn log-normal off-diagonal weights (refactorization)."""
    n = A.n
    col = np.repeat(np.arange(n), np.diff(A.p))
    off = A.i != col
    x = np.empty_like(A.x)
    w = np.exp(sigma * rng.standard_normal(int(off.sum())))
    x[off] = -w
    x[~off] = shift + np.bincount(A.i[off], w, n) + np.bincount(col[off], w, n)
    return SymMatrix(n=n, p=A.p, i=A.i, x=x, name=A.name + "_rev")


# ---------------------------------------------------------------------------
# The benchmark workloads with their right-hand sides
# ---------------------------------------------------------------------------

def workload_laplace2d_100():
    """2D 5-point 100x100 (n=10,000); one RHS b = A x_true, x_true ~ U(-1,1), seed 0."""
    A = laplace2d_5pt(100)
    rng = np.random.default_rng(0)
    xt = rng.uniform(-1.0, 1.0, A.n)
    return A, A.matvec(xt)[:, None], xt[:, None]


def workload_diffusion3d_64():
    """3D 7-point 64^3 diffusion (n=262,144); one RHS b = A x_true, seed 1 (after k)."""
    rng = np.random.default_rng(1)
    n = 64 ** 3
    k = np.exp(rng.standard_normal(n))   # same draws diffusion3d_7pt makes
    A = diffusion3d_7pt(64, seed=1)
    assert A.n == n and k.size == n
    xt = rng.uniform(-1.0, 1.0, n)
    return A, A.matvec(xt)[:, None], xt[:, None]


def workload_stencil27_96(side: int = 96, nrhs: int = 4):
    """3D 27-point 96^3 (n=884,736), 4 RHS ~ U(-1,1) drawn after the edge weights (seed 2)."""
    A = stencil3d_27pt(side, seed=2)
    rng = np.random.default_rng(2)
    nedge = A.nnz_lower - A.n
    rng.standard_normal(nedge)            # skip the edge-weight draws
    B = rng.uniform(-1.0, 1.0, (A.n, nrhs))
    return A, np.asfortranarray(B), None


def batch_sides(count: int = 4096, seed: int = 3, lo: int = 16, hi: int = 128):
    """Side lengths for the batched workload: side ~ U{lo..hi}."""
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi + 1, size=count), rng
