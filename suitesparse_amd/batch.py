"""Batched factorization of many independent SPD systems on one GPU.

A batch of k systems is factored as ONE block-diagonal matrix
blockdiag(A_1, ..., A_k).  Its supernodal elimination tree is the forest of
the k systems' trees, so the level schedule of the device plan already
interleaves every system: each level is one update launch per tile class
and one POTRF/TRSM/trailing launch per 64-column step, across all systems at
once (the persistent-kernel role: one grid covers the work of every system at
that level).  Same-pattern refactorization of the whole batch replays one
captured hipGraph.

The symbolic factor of the block-diagonal matrix is the concatenation of the
per-system factors (no supernode can span two systems: amalgamation only
merges a supernode with its parent, and the weighted postorder visits the
trees in order), so each system's Super/Lpi/Lpx/Ls are exactly what
cholmod_l_analyze gives for that system alone (tests/test_batch.py).

The ordering of each distinct pattern is computed once (AMD by default) and
reused for every system with that pattern.
"""
from __future__ import annotations

import dataclasses
import hashlib

import numpy as np

from . import _lib, cholmod
from .problems import SymMatrix


@dataclasses.dataclass
class Batch:
    A: SymMatrix                 # block-diagonal tril(A)
    offsets: np.ndarray          # row offset of each system (k+1)
    nnz_offsets: np.ndarray      # offset of each system's values in A.x (k+1)
    L: cholmod.Factor            # symbolic (+ numeric after factorize)
    cm: cholmod.Common
    orderings: int               # distinct patterns ordered

    @property
    def k(self) -> int:
        return len(self.offsets) - 1

    def system_slice(self, j: int) -> slice:
        return slice(int(self.offsets[j]), int(self.offsets[j + 1]))

    def values(self, mats) -> np.ndarray:
        """Concatenate the values of k matrices with the batch's patterns."""
        return np.concatenate([np.asarray(m.x, dtype=np.float64) for m in mats])


def _pattern_key(A: SymMatrix) -> str:
    h = hashlib.sha1()
    h.update(np.int64(A.n).tobytes())
    h.update(np.ascontiguousarray(A.p, dtype=np.int64).tobytes())
    h.update(np.ascontiguousarray(A.i, dtype=np.int64).tobytes())
    return h.hexdigest()


def ordering_for(A: SymMatrix, method: int = cholmod.CHOLMOD_AMD) -> np.ndarray:
    """Fill-reducing ordering of one pattern (product AMD, or natural)."""
    if method == cholmod.CHOLMOD_NATURAL:
        return np.arange(A.n, dtype=np.int64)
    if method != cholmod.CHOLMOD_AMD:
        raise ValueError("batched analysis supports AMD or natural orderings")
    perm = np.zeros(A.n, dtype=np.int64)
    p = np.ascontiguousarray(A.p, dtype=np.int64)
    i = np.ascontiguousarray(A.i, dtype=np.int64)
    st = _lib.lib().mxc_amd_order(A.n, cholmod._ptr(p), cholmod._ptr(i), cholmod._ptr(perm))
    _lib.check("mxc_amd_order", st)
    return perm


def block_diagonal(mats) -> tuple[SymMatrix, np.ndarray, np.ndarray]:
    ns = np.array([m.n for m in mats], dtype=np.int64)
    nnzs = np.array([int(m.p[-1]) for m in mats], dtype=np.int64)
    off = np.zeros(len(mats) + 1, dtype=np.int64)
    np.cumsum(ns, out=off[1:])
    noff = np.zeros(len(mats) + 1, dtype=np.int64)
    np.cumsum(nnzs, out=noff[1:])
    p = np.empty(off[-1] + 1, dtype=np.int64)
    i = np.empty(noff[-1], dtype=np.int64)
    x = np.empty(noff[-1], dtype=np.float64)
    for j, m in enumerate(mats):
        p[off[j]:off[j + 1]] = np.asarray(m.p[:-1], dtype=np.int64) + noff[j]
        i[noff[j]:noff[j + 1]] = np.asarray(m.i, dtype=np.int64) + off[j]
        x[noff[j]:noff[j + 1]] = m.x
    p[-1] = noff[-1]
    return SymMatrix(n=int(off[-1]), p=p, i=i, x=x, name=f"batch{len(mats)}"), off, noff


def analyze_batch(mats, cm: cholmod.Common, ordering: int = cholmod.CHOLMOD_AMD) -> Batch | None:
    """Order every distinct pattern once, then analyze the block-diagonal matrix."""
    cache = {}
    perms = []
    for m in mats:
        key = _pattern_key(m)
        if key not in cache:
            cache[key] = ordering_for(m, ordering)
        perms.append(cache[key])
    A, off, noff = block_diagonal(mats)
    perm = np.concatenate([pm + off[j] for j, pm in enumerate(perms)]) if perms else np.zeros(0, np.int64)
    cm.nmethods, cm.ordering = 1, cholmod.CHOLMOD_GIVEN      # the block permutation as given
    L = cholmod.analyze_p(A, perm, cm)
    if L is None:
        return None
    return Batch(A=A, offsets=off, nnz_offsets=noff, L=L, cm=cm, orderings=len(cache))


def factorize_batch(b: Batch, values=None) -> bool:
    """Factor every system (values: concatenated tril values, host or device)."""
    return cholmod.factorize(b.A, b.L, b.cm, values=values)


def solve_batch(b: Batch, B):
    """X = A_j \\ B_j for every system j (B stacked: sum(n_j) x nrhs)."""
    return cholmod.solve(cholmod.CHOLMOD_A, b.L, B, b.cm)


def solve_refine_batch(b: Batch, B, tol: float = 1e-10, maxiter: int = 10):
    """FP64 iterative refinement over the whole batch until every system's
    relative residual is <= tol."""
    return cholmod.solve_refine(b.L, B, b.cm, tol=tol, maxiter=maxiter, segments=b.offsets)
