"""Dense pseudo-inverse of the coarsest level (the reference's coarse direct
solve, multigrid.hpp:37-84 and 217-257).  Coarse levels of at least 256
vertices are inverted on the device (blocked Gauss-Jordan, FP64 MFMA
updates); smaller ones on the host.  The device result is checked by what
fixes a pseudo-inverse of a Laplacian: z = L^+ r solves L z = r - (the
component means of r) and has zero mean on every connected component."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

import ligmg_amd
import oracle_ffi as O
from ligmg_amd import _ffi, graphs

pytestmark = pytest.mark.gpu


def coarse_level(d):
    lc = d.num_levels() - 1
    k, n, nnz, _ = d.level_info(lc)
    assert k == _ffi.LEVEL_COARSE
    rp, ci, v = d.level_matrix(lc)
    return lc, sp.csr_matrix((v, ci, rp), shape=(n, n))


def check_pinv(d, seed):
    lc, L = coarse_level(d)
    n = L.shape[0]
    ncomp, lab = csg.connected_components(L, directed=False)
    r = np.random.default_rng(seed).standard_normal(n)
    z = d.apply_cycle(lc, r)
    cnt = np.bincount(lab, minlength=ncomp)
    rm = r - (np.bincount(lab, r, ncomp) / cnt)[lab]
    res = np.linalg.norm(L @ z - rm) / np.linalg.norm(rm)
    zmean = np.abs(np.bincount(lab, z, ncomp) / cnt).max() / np.abs(z).max()
    print(f"coarse n={n} components={ncomp} residual {res:.2e} mean {zmean:.1e}", flush=True)
    assert res <= 1e-9
    assert zmean <= 1e-12
    return n


@pytest.mark.parametrize("shape", [(24, 24, 24), (30, 30, 9)])
def test_device_pinv_grid(shape):
    g = graphs.grid3d(*shape, graphs.W_UNIFORM, 0.5, 1.5, seed=1)
    d = ligmg_amd.Solver(g, ligmg_amd.Options(max_levels=2))
    assert check_pinv(d, seed=2) >= 256  # the device path


def test_device_pinv_components():
    """Several components in one coarse level: each gets its own nullspace
    term."""
    g = graphs.union([graphs.grid2d(40, 30, graphs.W_UNIFORM, 0.5, 1.5, seed=s) for s in range(3)])
    d = ligmg_amd.Solver(g, ligmg_amd.Options(max_levels=2))
    _, L = coarse_level(d)
    assert csg.connected_components(L, directed=False)[0] == 3
    assert check_pinv(d, seed=4) >= 256


def test_device_pinv_solve_matches_oracle():
    """The whole solve with a device-inverted coarse level against the
    oracle's (host Cholesky) pseudo-inverse: the same preconditioner up to
    rounding."""
    g = graphs.grid3d(20, 20, 20, graphs.W_UNIFORM, 0.5, 1.5, seed=3)
    kw = dict(max_levels=2)
    d = ligmg_amd.Solver(g, ligmg_amd.Options(**kw))
    r = O.OracleSolver(g, O.default_options(**kw))
    assert coarse_level(d)[1].shape[0] >= 256
    b = graphs.rhs(g.n, 1, 5)
    info = d.solve(b)
    xr, rinfo, rhist = r.solve(b)
    assert info.rel_residual() <= 1e-8
    assert abs(info.iters() - rinfo.iters) <= 1
    k = min(6, len(rhist), len(info.residuals))
    assert np.allclose(np.asarray(info.residuals)[:k], rhist[:k], rtol=1e-7)
    x = info.solution
    assert np.linalg.norm((x - x.mean()) - (xr - xr.mean())) <= 1e-6 * np.linalg.norm(xr - xr.mean())


def test_host_pinv_below_threshold():
    """Coarse levels under 256 vertices keep the host factorisation."""
    g = graphs.grid2d(12, 12, graphs.W_UNIFORM, 0.5, 1.5, seed=1)
    d = ligmg_amd.Solver(g, ligmg_amd.Options(max_levels=2))
    assert check_pinv(d, seed=6) < 256


@pytest.mark.parametrize("k", [1, 5, 16, 33, 64])
def test_coarse_block_apply_matches_columns(k):
    """The batched solves' coarse apply (matrix-core tiles from 16 columns)
    against the single-vector apply of every column."""
    g = graphs.grid3d(24, 24, 24, graphs.W_UNIFORM, 0.5, 1.5, seed=1)
    d = ligmg_amd.Solver(g, ligmg_amd.Options(max_levels=2))
    lc, L = coarse_level(d)
    n = L.shape[0]
    X = np.random.default_rng(k).standard_normal((n, k))
    Y = np.empty_like(X)
    P = C.POINTER(C.c_double)
    _check = ligmg_amd.solver._check
    _check(_ffi.solver_lib().ligmg_coarse_apply_block(d._h, k, X.ctypes.data_as(P), Y.ctypes.data_as(P)),
           "ligmg_coarse_apply_block")
    for c in range(k):
        z = d.apply_cycle(lc, X[:, c].copy())
        assert np.linalg.norm(Y[:, c] - z) <= 1e-12 * np.linalg.norm(z), c


@pytest.mark.parametrize("shape", [(24, 24, 24), (26, 25, 23)])
def test_device_pinv_matches_numpy_pinv(shape):
    """The device-inverted pseudo-inverse against numpy's SVD pseudo-inverse
    of the same coarse matrix, applied to 64 random columns: coarse levels of
    about a thousand vertices, several 128-wide Gauss-Jordan blocks and a
    padded last block."""
    g = graphs.grid3d(*shape, graphs.W_UNIFORM, 0.5, 1.5, seed=2)
    d = ligmg_amd.Solver(g, ligmg_amd.Options(max_levels=2))
    lc, L = coarse_level(d)
    n = L.shape[0]
    assert n >= 256 and n % 128 != 0, n
    X = np.random.default_rng(n).standard_normal((n, 64))
    Y = np.empty_like(X)
    P = C.POINTER(C.c_double)
    ligmg_amd.solver._check(
        _ffi.solver_lib().ligmg_coarse_apply_block(d._h, 64, X.ctypes.data_as(P), Y.ctypes.data_as(P)),
        "ligmg_coarse_apply_block")
    ref = np.linalg.pinv(L.toarray()) @ X
    err = np.linalg.norm(Y - ref) / np.linalg.norm(ref)
    print(f"coarse n={n}: |device - numpy pinv| {err:.2e}", flush=True)
    assert err <= 1e-11


def _ring_laplacian(n, seed, comps):
    """Laplacian of `comps` weighted rings with random chords (CSR, sorted
    columns, diagonal stored)."""
    rng = np.random.default_rng(seed)
    sizes = [n // comps + (1 if c < n % comps else 0) for c in range(comps)]
    rows, cols = [], []
    off = 0
    for m in sizes:
        i = np.arange(m)
        rows += list(off + i)
        cols += list(off + (i + 1) % m)
        a = rng.integers(0, m, m // 2)
        b = (a + 1 + rng.integers(1, m - 2, m // 2)) % m
        rows += list(off + a)
        cols += list(off + b)
        off += m
    w = rng.uniform(0.5, 1.5, len(rows))
    W = sp.coo_matrix((w, (rows, cols)), shape=(n, n)).tocsr()
    W = (W + W.T).tocsr()
    L = (sp.diags(np.asarray(W.sum(axis=1)).ravel()) - W).tocsr()
    L.sum_duplicates()
    L.sort_indices()
    return L


@pytest.mark.parametrize("n,comps", [(255, 1), (256, 1), (257, 1), (383, 1), (384, 1), (385, 1), (640, 1),
                                     (300, 2), (513, 3)])
def test_dense_pinv_block_boundaries(n, comps):
    """The coarse pseudo-inverse at sizes around the host/device threshold
    (256 vertices) and the 128-wide Gauss-Jordan blocks (a full last block,
    one more vertex, one fewer), connected and with several components,
    against numpy's SVD pseudo-inverse."""
    L = _ring_laplacian(n, seed=n, comps=comps)
    assert csg.connected_components(L, directed=False)[0] == comps
    rp = L.indptr.astype(np.int32)
    ci = L.indices.astype(np.int32)
    v = L.data.astype(np.float64)
    out = np.empty((n, n))
    P32, PD = C.POINTER(C.c_int32), C.POINTER(C.c_double)
    ligmg_amd.solver._check(
        _ffi.solver_lib().ligmg_dense_pinv(n, rp.ctypes.data_as(P32), ci.ctypes.data_as(P32),
                                           v.ctypes.data_as(PD), out.ctypes.data_as(PD)), "ligmg_dense_pinv")
    ref = np.linalg.pinv(L.toarray())
    err = np.abs(out - ref).max() / np.abs(ref).max()
    print(f"n={n} components={comps}: |device - numpy pinv| {err:.2e}", flush=True)
    assert err <= 1e-10
