"""Matrix files: the triplet and Matrix Market reader of CHOLMOD's Check module.

read_sparse(path) returns what cholmod_l_read_sparse2 (CHOLMOD/Check/cholmod_read.c,
documented at :40-125) returns for a real or pattern-only sparse matrix, with
the reader's default prefer_binary = FALSE:

  * lines starting with '%' are comments, except a '%%MatrixMarket' header,
    whose storage word sets the symmetry ("general" -> stype 0, "symmetric" /
    "hermitian" -> stype -1);
  * the first data line holds nrow ncol nnz [stype]; a fourth integer
    overrides the header's stype, further tokens are ignored;
  * every entry line holds 2 (pattern), 3 (real) or 4 (complex) tokens;
  * indices are one-based unless a zero index appears, then zero-based;
  * an unknown stype is inferred from the entries (both strict triangles ->
    0, only the lower -> -1, otherwise 1);
  * a symmetric pattern-only matrix gets off-diagonals -1 and, where the
    diagonal is present, diagonal = 1 + degree (an SPD matrix); an
    unsymmetric pattern-only matrix gets ones;
  * for a symmetric matrix an entry of the other triangle is moved across,
    duplicates are summed, and the result is stored as the upper triangle
    (stype 1, Common->prefer_upper).

Complex matrices are not read (the numeric path here is real).  to_symmetric
turns a symmetric result into the tril(A) CSC form that analyze / factorize
take, and aat_plus_beta forms tril(A A' + beta I) for an unsymmetric one, as
cholmod_dl_demo.c:300-305 factors it.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .problems import SymMatrix


@dataclasses.dataclass
class SparseMatrix:
    nrow: int
    ncol: int
    p: np.ndarray          # int64, ncol + 1
    i: np.ndarray          # int64, sorted within each column
    x: np.ndarray          # float64
    stype: int             # 0 unsymmetric, -1 lower stored, 1 upper stored


def _tokens(path):
    """(header storage word or None, list of data lines as token lists)."""
    storage = None
    lines = []
    with open(path) as f:
        for raw in f:
            s = raw.strip()
            if not s:
                continue
            if s.startswith("%"):
                if s.lower().startswith("%%matrixmarket") and storage is None:
                    words = s.lower().split()
                    storage = words[-1] if len(words) >= 5 else ""
                continue
            lines.append(s.split())
    return storage, lines


def read_sparse(path) -> SparseMatrix:
    storage, lines = _tokens(path)
    if not lines:
        raise ValueError(f"{path}: no data")
    head = lines[0]
    if len(head) < 3:
        raise ValueError(f"{path}: first data line must hold nrow ncol nnz")
    nrow, ncol, nnz = int(head[0]), int(head[1]), int(head[2])
    stype = None
    if storage is not None:
        if storage == "general":
            stype = 0
        elif storage in ("symmetric", "hermitian"):
            stype = -1
        else:
            raise ValueError(f"{path}: storage '{storage}' not supported")
    if len(head) >= 4:
        st = int(head[3])
        stype = -1 if st < 0 else (1 if st > 0 else 0)
    ent = lines[1:1 + nnz]
    if len(ent) < nnz:
        raise ValueError(f"{path}: {nnz} entries announced, {len(ent)} found")
    ntok = len(ent[0]) if ent else 3
    if ntok == 4:
        raise ValueError(f"{path}: complex matrices are not supported")
    if ntok not in (2, 3) or any(len(e) != ntok for e in ent):
        raise ValueError(f"{path}: inconsistent entry lines")
    pattern = ntok == 2
    r = np.array([int(e[0]) for e in ent], dtype=np.int64)
    c = np.array([int(e[1]) for e in ent], dtype=np.int64)
    v = np.ones(nnz) if pattern else np.array([float(e[2]) for e in ent])
    if nnz and (r.min() == 0 or c.min() == 0):
        base = 0
    else:
        base = 1
    r -= base
    c -= base
    if nnz and (r.min() < 0 or c.min() < 0 or r.max() >= nrow or c.max() >= ncol):
        raise ValueError(f"{path}: index out of range")
    if stype is None:
        # cholmod_read.c:742-766: no strictly upper entry -> "lower", no strictly
        # lower entry -> "upper"; both (a diagonal) -> upper; neither -> unsymmetric
        is_lower = not np.any(r < c)
        is_upper = not np.any(r > c)
        if nrow != ncol:
            stype = 0
        elif is_upper:
            stype = 1
        elif is_lower:
            stype = -1
        else:
            stype = 0
    if stype != 0 and nrow != ncol:
        raise ValueError(f"{path}: a symmetric matrix must be square")
    if pattern:
        if stype == 0:
            v = np.ones(nnz)
        else:
            # cholmod_read.c:866-900: degrees from the entries strictly inside the
            # stored triangle, counted for both a(i,j) and a(j,i); diagonal
            # 1 + degree, off-diagonals -1
            inside = (r > c) if stype < 0 else (r < c)
            deg = np.bincount(r[inside], minlength=nrow) + np.bincount(c[inside], minlength=nrow)
            v = np.where(r == c, 1.0 + deg[r].astype(np.float64), -1.0)
    if stype != 0:
        # t_cholmod_triplet_to_sparse.c:115-127: an entry of the other triangle
        # is moved across; cholmod_read.c:1249-1255 then returns the upper form
        r, c = np.minimum(r, c), np.maximum(r, c)
        stype = 1
    # sum duplicates, sort by (column, row)
    key = c * max(nrow, 1) + r
    order = np.argsort(key, kind="stable")
    key, v = key[order], v[order]
    uniq, start = np.unique(key, return_index=True)
    v = np.add.reduceat(v, start) if v.size else v
    rows = uniq % max(nrow, 1)
    cols = uniq // max(nrow, 1)
    p = np.zeros(ncol + 1, dtype=np.int64)
    np.cumsum(np.bincount(cols, minlength=ncol), out=p[1:])
    return SparseMatrix(nrow=nrow, ncol=ncol, p=p, i=rows.astype(np.int64), x=v.astype(np.float64),
                        stype=stype)


def to_symmetric(M: SparseMatrix, name: str = "") -> SymMatrix:
    """tril(A) CSC of a symmetric matrix read with stype -1 or 1."""
    if M.stype == 0:
        raise ValueError("unsymmetric matrix: use aat_plus_beta")
    import scipy.sparse as sp
    A = sp.csc_matrix((M.x, M.i, M.p), shape=(M.nrow, M.ncol))
    if M.stype > 0:
        A = A.T.tocsc()
    A = sp.tril(A).tocsc()
    A.sort_indices()
    return SymMatrix(n=M.nrow, p=A.indptr.astype(np.int64), i=A.indices.astype(np.int64),
                     x=A.data.astype(np.float64), name=name)


def aat_plus_beta(M: SparseMatrix, beta: float, name: str = "") -> SymMatrix:
    """tril(A A' + beta I) of an unsymmetric matrix (the demo's factorize_p case)."""
    import scipy.sparse as sp
    A = sp.csc_matrix((M.x, M.i, M.p), shape=(M.nrow, M.ncol))
    C = (A @ A.T + beta * sp.identity(M.nrow)).tocsc()
    C = sp.tril(C).tocsc()
    C.sort_indices()
    return SymMatrix(n=M.nrow, p=C.indptr.astype(np.int64), i=C.indices.astype(np.int64),
                     x=C.data.astype(np.float64), name=name)
