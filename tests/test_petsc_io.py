"""PETSc binary reader/writer (ligmg_amd.petsc_io) on known bytes and round
trips.  The layout is PETSc's documented binary format: big-endian class id
1211216 (matrix) / 1211214 (vector), sizes, row lengths, columns, values."""
import struct

import numpy as np
import pytest
import scipy.sparse as sp

from ligmg_amd import petsc_io


def test_matrix_known_bytes(tmp_path):
    A = sp.csr_matrix(np.array([[1.0, 0.0, 2.0], [0.0, 0.0, 0.0]]))
    p = tmp_path / "a.petsc"
    petsc_io.write_mat(p, A)
    expect = (struct.pack(">4i", 1211216, 2, 3, 2) + struct.pack(">2i", 2, 0) + struct.pack(">2i", 0, 2)
              + struct.pack(">2d", 1.0, 2.0))
    assert p.read_bytes() == expect
    assert expect[:4] == bytes([0x00, 0x12, 0x7B, 0x50])
    B = petsc_io.read_mat(p)
    assert B.shape == (2, 3) and (B != A).nnz == 0


def test_vector_known_bytes(tmp_path):
    p = tmp_path / "v.petsc"
    petsc_io.write_vec(p, [0.5, -3.0])
    assert p.read_bytes() == struct.pack(">2i", 1211214, 2) + struct.pack(">2d", 0.5, -3.0)
    assert np.array_equal(petsc_io.read_vec(p), [0.5, -3.0])


def test_round_trip_and_errors(tmp_path):
    rng = np.random.default_rng(0)
    A = sp.random(300, 200, density=0.03, random_state=1, format="csr")
    A.sort_indices()
    p = tmp_path / "m.petsc"
    petsc_io.write_mat(p, A)
    B = petsc_io.read_mat(p)
    assert np.array_equal(B.indptr, A.indptr) and np.array_equal(B.indices, A.indices)
    assert np.array_equal(B.data, A.data)
    x = rng.standard_normal(77)
    petsc_io.write_vec(tmp_path / "x.petsc", x)
    assert np.array_equal(petsc_io.read_vec(tmp_path / "x.petsc"), x)
    raw = p.read_bytes()
    (tmp_path / "t.petsc").write_bytes(raw[:-8])
    with pytest.raises(ValueError):
        petsc_io.read_mat(tmp_path / "t.petsc")
    with pytest.raises(ValueError):
        petsc_io.read_vec(p)  # a matrix is not a vector
