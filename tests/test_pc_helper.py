"""The "ligmg" PETSc preconditioner's logic without PETSc (ligmg_pc_*,
petsc_shim.hpp:18-74): the rows of the preconditioning matrix go in one at a
time as PCSetUp's MatGetRow loop hands them over, the solver is built with
default options, and PCApply is a full solve.  CPU tests check the input
handling (nothing reaches the GPU); the GPU test drives the helper from a
row iterator over a scipy matrix and compares with ligmg_solve."""
import ctypes as C

import numpy as np
import pytest

import ligmg_amd
from ligmg_amd import _ffi, graphs

P = C.POINTER


def new_pc():
    h = C.c_void_p()
    assert _ffi.solver_lib().ligmg_pc_create(C.byref(h)) == 0
    return h


def add_row(h, i, cols, vals):
    cols = np.ascontiguousarray(cols, np.int64)
    vals = np.ascontiguousarray(vals, np.float64)
    return _ffi.solver_lib().ligmg_pc_add_row(h, i, len(cols), cols.ctypes.data_as(P(C.c_int64)),
                                              vals.ctypes.data_as(P(C.c_double)))


def test_pc_helper_rejects_bad_input():
    lib = _ffi.solver_lib()
    h = new_pc()
    try:
        assert add_row(h, 0, [0], [1.0]) == _ffi.ERR_INVALID  # no begin
        assert lib.ligmg_pc_begin(h, 3) == 0
        assert add_row(h, 1, [1], [1.0]) == _ffi.ERR_INVALID  # out of order
        assert b"in order" in lib.ligmg_last_error()
        assert add_row(h, 0, [0, 3], [1.0, 2.0]) == _ffi.ERR_INVALID  # column out of range
        assert add_row(h, 0, [1, 0, 1], [1.0, 2.0, 3.0]) == _ffi.ERR_INVALID  # duplicate
        assert b"duplicate" in lib.ligmg_last_error()
        assert add_row(h, 0, [1, 0], [-1.0, 1.0]) == 0  # any column order
        assert lib.ligmg_pc_setup(h) == _ffi.ERR_INVALID  # rows 1, 2 missing
        x = np.zeros(3)
        assert lib.ligmg_pc_apply(h, x.ctypes.data_as(P(C.c_double)), x.ctypes.data_as(P(C.c_double)),
                                  None) == _ffi.ERR_INVALID  # not set up
        assert lib.ligmg_pc_begin(h, 0) == _ffi.ERR_INVALID
    finally:
        assert lib.ligmg_pc_destroy(h) == 0


@pytest.mark.gpu
def test_pc_helper_matches_solve():
    """MatGetRow order, columns shuffled within each row as an arbitrary
    matrix type may return them: the PC's apply gives ligmg_solve's bits
    with default options; a second PCSetUp replaces the solver."""
    lib = _ffi.solver_lib()
    rng = np.random.default_rng(0)
    h = new_pc()
    try:
        for g in (graphs.rmat_config(12, "lognormal", pendants=True), graphs.grid2d(40, 33)):
            A = g.to_scipy()
            assert lib.ligmg_pc_begin(h, g.n) == 0
            for i in range(g.n):
                a, b = A.indptr[i], A.indptr[i + 1]
                p = rng.permutation(b - a)
                assert add_row(h, i, A.indices[a:b][p], A.data[a:b][p]) == 0
            assert lib.ligmg_pc_setup(h) == 0, lib.ligmg_last_error()
            rhs = graphs.rhs(g.n, 1, 4)
            x = np.empty(g.n)
            info = _ffi.ligmg_info()
            assert lib.ligmg_pc_apply(h, rhs.ctypes.data_as(P(C.c_double)), x.ctypes.data_as(P(C.c_double)),
                                      C.byref(info)) == 0
            ref = ligmg_amd.Solver(g).solve(rhs)
            print(f"n={g.n} iters={info.iters} ref iters={ref.iters()}")
            assert info.iters == ref.iters() and info.rel_residual <= 1e-8
            assert np.array_equal(x, ref.solution)
    finally:
        lib.ligmg_pc_destroy(h)
