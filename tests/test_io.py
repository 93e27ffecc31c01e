"""The matrix-file reader (suitesparse_amd/io.py) against the reference's own
cholmod_l_read_sparse2 (CHOLMOD/Check/cholmod_read.c), bit for bit, on the
matrices that CHOLMOD's Demo/Matrix and Tcov/Matrix directories hold (copied
as data into tests/golden/cholmod_matrices): Matrix Market and triplet
headers, one- and zero-based indices, lower / upper / inferred storage,
diagonal-only, pattern-only (Laplacian values) and empty matrices."""
import pathlib

import numpy as np
import pytest

from suitesparse_amd import io

FIX = pathlib.Path(__file__).resolve().parent / "golden" / "cholmod_matrices"
FILES = sorted(p.name for p in FIX.iterdir())


@pytest.mark.parametrize("name", FILES)
def test_reader_matches_reference(ref_lib, name):
    ref = ref_lib.read_sparse(FIX / name)
    M = io.read_sparse(FIX / name)
    assert (M.nrow, M.ncol, M.stype) == (ref["nrow"], ref["ncol"], ref["stype"])
    assert np.array_equal(M.p, ref["p"])
    assert np.array_equal(M.i, ref["i"])
    assert np.array_equal(M.x, ref["x"])          # bit-exact values


def test_symmetric_form_is_the_same_matrix():
    """bcsstk01 stored lower (Demo .tri) and upper (Tcov k01up) is one matrix."""
    a = io.to_symmetric(io.read_sparse(FIX / "bcsstk01.tri"))
    b = io.to_symmetric(io.read_sparse(FIX / "k01up"))
    assert a.n == b.n == 48
    assert np.array_equal(a.p, b.p) and np.array_equal(a.i, b.i)
    assert np.allclose(a.x, b.x, rtol=1e-15, atol=0)


def test_pattern_matrix_is_spd_laplacian():
    """can___24 is pattern-only: off-diagonals -1, diagonal 1 + degree."""
    A = io.to_symmetric(io.read_sparse(FIX / "can___24.mtx")).to_scipy_full().toarray()
    off = A - np.diag(np.diag(A))
    assert set(np.unique(off)) <= {-1.0, 0.0}
    assert np.array_equal(np.diag(A), 1 + (off != 0).sum(axis=1))
    assert np.linalg.eigvalsh(A).min() > 0


def test_reader_errors(tmp_path):
    bad = tmp_path / "complex.tri"
    bad.write_text("2 2 2 -1\n1 1 1.0 0.0\n2 2 1.0 0.0\n")
    with pytest.raises(ValueError):
        io.read_sparse(bad)
    short = tmp_path / "short.tri"
    short.write_text("3 3 4 -1\n1 1 1.0\n2 2 1.0\n")
    with pytest.raises(ValueError):
        io.read_sparse(short)
    oob = tmp_path / "oob.tri"
    oob.write_text("2 2 1 -1\n3 1 1.0\n")
    with pytest.raises(ValueError):
        io.read_sparse(oob)
    uns = io.read_sparse(FIX / "one.tri")
    with pytest.raises(ValueError):
        io.to_symmetric(uns)
    C = io.aat_plus_beta(uns, 1e-6)
    assert C.n == 1 and np.allclose(C.x, [1.0 + 1e-6])
