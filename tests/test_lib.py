"""CPU-side checks of the native libraries: they load, export every entry
point the public headers declare, and the host-only parts (options,
generators) behave.  No GPU compute is called here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from ligmg_amd import _ffi, graphs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ligmg_[a-z0-9_]+)\s*\(", txt)))


def exported(lib):
    out = []
    for name in dir(lib):
        pass
    return out


@pytest.mark.parametrize("header,libpath", [("ligmg_amd.h", _ffi.LIB_SOLVER),
                                            ("ligmg_graphs.h", _ffi.LIB_GRAPHS)])
def test_library_exports_every_declared_symbol(header, libpath):
    assert os.path.exists(libpath), f"{libpath} not built"
    lib = C.CDLL(libpath)
    names = declared(header)
    assert names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_options_defaults_match_reference():
    """options.hpp:80-134 defaults."""
    import ligmg_amd
    o = ligmg_amd.Options()
    assert (o.max_levels, o.min_size, o.elim_degree, o.tol, o.thresh) == (50, 4000, 4, 1e-8, 1e-8)
    assert (o.agg_iters, o.iters, o.pre_smooth, o.post_smooth) == (10, 500, 2, 2)
    assert (o.num_tvs, o.tvs_smooth, o.kcycle, o.kcycle_iters) == (10, 4, 0, 2)
    assert (o.votes_to_seed, o.max_agg_size, o.agg_iter_factor, o.agg_init_factor) == (1, -1, 0.5, 0.5)
    assert (o.cheby_lower, o.cheby_upper, o.elim_factor, o.elim_iters) == (0.3, 1.1, 0.9, 1)
    assert o.outer_ksp_type == _ffi.KSP_CG and o.kcycle_ksp_type == _ffi.KSP_FCG
    assert o.aff_scaling == _ffi.AFF_MAX and o.smoother_type == _ffi.SMOOTHER_CHEBYSHEV
    assert (o.neighbor_of_neighbor, o.rejection_ratio, o.rejection_to_seed) == (0, -1.0, 0)
    with pytest.raises(AttributeError):
        o.not_an_option = 1


def test_generators_are_laplacians():
    for g in [graphs.grid2d(7, 5), graphs.grid3d(3, 4, 5, wkind=_ffi.W_UNIFORM, p0=.1, p1=10),
              graphs.rmat_config(9, "lognormal", pendants=True),
              graphs.barabasi_albert(500, 5, seed=1), graphs.rgg(800, 0.08, seed=2)]:
        A = g.to_scipy()
        assert abs(A - A.T).max() == 0
        assert np.abs(A @ np.ones(g.n)).max() < 1e-9 * np.abs(g.vals).max()
        rows = np.repeat(np.arange(g.n), np.diff(g.rowptr))
        assert np.all(g.vals[rows != g.colind] < 0)
        assert np.all(np.diff(g.colind)[np.diff(rows) == 0] > 0)  # sorted, unique


def test_grid_sizes_and_rmat_lcc():
    g = graphs.grid2d(256, 256)
    assert (g.n, g.nnz) == (65536, 326656)
    g = graphs.grid3d(16, 16, 16)
    assert g.nnz == 16 ** 3 + 2 * 3 * 15 * 16 * 16
    r = graphs.rmat_config(10, "uniform")
    import scipy.sparse.csgraph as cg
    assert cg.connected_components(r.to_scipy(), directed=False)[0] == 1


def test_generators_deterministic_across_threads():
    a = graphs.rmat(11, 16, wkind=_ffi.W_LOGNORMAL, p0=0, p1=1.5, pendant_frac=0.1, seed=7, nthreads=1)
    b = graphs.rmat(11, 16, wkind=_ffi.W_LOGNORMAL, p0=0, p1=1.5, pendant_frac=0.1, seed=7, nthreads=4)
    assert np.array_equal(a.colind, b.colind) and np.array_equal(a.vals, b.vals)


def test_rhs_mean_zero_per_block():
    cp = np.array([0, 100, 250, 400])
    B = graphs.rhs(400, 3, 0, comp_ptr=cp)
    for c in range(3):
        assert np.all(np.abs(B[cp[c]:cp[c + 1]].sum(0)) < 1e-12)


def test_solve_rejects_wrong_shapes_before_the_library():
    """A right-hand side of the wrong length must raise ValueError in Python,
    before ligmg_solve (which reads n values of b and writes n of x) is
    called, for numpy arrays and torch tensors alike."""
    import torch
    from ligmg_amd.solver import Solver

    class NoLib:
        def __getattr__(self, name):
            raise AssertionError(f"library called: {name}")

    s = Solver.__new__(Solver)
    s._lib, s._h, s.n = NoLib(), None, 10
    for b in (np.zeros(5), torch.zeros(5, dtype=torch.float64), torch.zeros(11),
              torch.zeros(10, 2), torch.zeros(2, 5)):
        with pytest.raises(ValueError):
            s.solve(b)
    for B in (np.zeros((5, 2)), torch.zeros(5, 2), torch.zeros(10), torch.zeros(10, 65),
              torch.zeros(10, 0)):
        with pytest.raises(ValueError):
            s.solve_batch(B)


def test_dense_pinv_rejects_bad_input_before_any_device_work():
    """ligmg_dense_pinv checks its CSR on the host and fails with a status
    code and a message, never touching the device for a bad matrix."""
    lib = _ffi.solver_lib()
    P32, PD = C.POINTER(C.c_int32), C.POINTER(C.c_double)
    out = np.zeros(16)

    def call(n, rp, ci, v):
        rp, ci, v = (np.asarray(rp, np.int32), np.asarray(ci, np.int32), np.asarray(v, np.float64))
        return lib.ligmg_dense_pinv(n, rp.ctypes.data_as(P32), ci.ctypes.data_as(P32), v.ctypes.data_as(PD),
                                    out.ctypes.data_as(PD))

    assert call(0, [0], [0], [0.0]) == _ffi.ERR_INVALID
    assert call(9000, [0], [0], [0.0]) == _ffi.ERR_UNSUPPORTED
    assert b"8192" in lib.ligmg_last_error()
    assert call(2, [1, 2, 3], [0, 1, 1], [1.0, 1.0, 1.0]) == _ffi.ERR_INVALID       # rp[0] != 0
    assert call(2, [0, 2, 1], [0, 1, 1], [1.0, -1.0, 1.0]) == _ffi.ERR_INVALID      # rp decreasing
    assert call(2, [0, 2, 4], [0, 1, 0, 2], [1.0, -1.0, -1.0, 1.0]) == _ffi.ERR_INVALID  # column 2 of 2
    assert b"column" in lib.ligmg_last_error()
