"""Seeded synthetic graph Laplacians (bindings of include/ligmg_graphs.h).

Each generator returns a :class:`Graph` holding host numpy arrays of the
Laplacian in CSR form (int32 indices, float64 values, sorted columns, the
diagonal stored) plus the block structure of block-diagonal unions.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _ffi
from ._ffi import W_GAUSS_DIST, W_LOGNORMAL, W_UNIFORM, W_UNIT  # noqa: F401


@dataclass
class Graph:
    n: int
    rowptr: np.ndarray
    colind: np.ndarray
    vals: np.ndarray
    comp_ptr: np.ndarray

    @property
    def nnz(self):
        return int(self.rowptr[-1])

    @property
    def ncomp(self):
        return len(self.comp_ptr) - 1

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.vals, self.colind, self.rowptr), shape=(self.n, self.n))

    def upper_edges(self):
        """(i, j, w) for stored off-diagonals with i < j, in row order."""
        rows = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.rowptr))
        m = self.colind > rows
        return rows[m], self.colind[m].astype(np.int64), -self.vals[m]


class _Owner:
    """Frees a C-allocated ligmg_graph when the last numpy view of it dies."""

    def __init__(self, g):
        self.g = g
        # bound now: at interpreter exit module globals may already be gone
        self._free = _ffi.graphs_lib().ligmg_graph_free
        self._byref = C.byref

    def __del__(self):
        self._free(self._byref(self.g))


def _view(owner, addr, ctype, count, dtype):
    buf = (ctype * max(count, 1)).from_address(addr)
    buf._owner = owner  # the ctypes buffer keeps the allocation alive
    return np.frombuffer(buf, dtype=dtype, count=count)


def _take(g):
    """Wraps the generator's arrays without copying them."""
    n, nnz = g.n, g.nnz
    owner = _Owner(g)
    rp = _view(owner, C.addressof(g.rowptr.contents), C.c_int32, n + 1, np.int32)
    ci = _view(owner, C.addressof(g.colind.contents), C.c_int32, nnz, np.int32)
    vv = _view(owner, C.addressof(g.vals.contents), C.c_double, nnz, np.float64)
    cp = np.ctypeslib.as_array(g.comp_ptr, shape=(g.ncomp + 1,)).copy()
    return Graph(n=int(n), rowptr=rp, colind=ci, vals=vv, comp_ptr=cp)


def _call(fn, *args):
    g = _ffi.ligmg_graph()
    rc = fn(*args, C.byref(g))
    if rc != 0:
        raise RuntimeError(f"graph generator failed ({rc})")
    return _take(g)


def read_mtx(path):
    """MatrixMarket coordinate file -> Graph holding the matrix as stored
    (symmetric files mirrored, duplicates summed, rows sorted).  An
    adjacency matrix becomes a Laplacian with prep.adjacency_to_laplacian."""
    g = _ffi.ligmg_graph()
    rc = _ffi.graphs_lib().ligmg_graph_read_mtx(str(path).encode(), C.byref(g))
    if rc != 0:
        raise ValueError(f"cannot read MatrixMarket file {path} (code {rc})")
    return _take(g)


def grid2d(nx, ny, wkind=W_UNIT, p0=0.0, p1=0.0, seed=0):
    return _call(_ffi.graphs_lib().ligmg_gen_grid2d, nx, ny, wkind, p0, p1, seed)


def grid3d(nx, ny, nz, wkind=W_UNIT, p0=0.0, p1=0.0, seed=0):
    return _call(_ffi.graphs_lib().ligmg_gen_grid3d, nx, ny, nz, wkind, p0, p1, seed)


def rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, wkind=W_UNIT, p0=0.0, p1=0.0,
         pendant_frac=0.0, pendant_max=3, seed=0, nthreads=0):
    return _call(_ffi.graphs_lib().ligmg_gen_rmat, scale, edge_factor, a, b, c, wkind, p0, p1,
                 pendant_frac, pendant_max, seed, nthreads)


def barabasi_albert(n, m=5, wkind=W_UNIT, p0=0.0, p1=0.0, seed=0):
    return _call(_ffi.graphs_lib().ligmg_gen_ba, n, m, wkind, p0, p1, seed)


def rgg(n, radius, wkind=W_GAUSS_DIST, p0=0.0, p1=0.0, seed=0):
    return _call(_ffi.graphs_lib().ligmg_gen_rgg, n, radius, wkind, p0, p1, seed)


def from_edges(n, u, v, w):
    u = np.ascontiguousarray(u, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.int32)
    w = np.ascontiguousarray(w, dtype=np.float64)
    P = C.POINTER
    return _call(_ffi.graphs_lib().ligmg_graph_from_edges, n, len(u),
                 u.ctypes.data_as(P(C.c_int32)), v.ctypes.data_as(P(C.c_int32)),
                 w.ctypes.data_as(P(C.c_double)))


def from_scipy(A):
    """Graph from a scipy sparse Laplacian (sorted indices, int32)."""
    A = A.tocsr()
    A.sort_indices()
    return Graph(n=A.shape[0], rowptr=A.indptr.astype(np.int32),
                 colind=A.indices.astype(np.int32), vals=A.data.astype(np.float64),
                 comp_ptr=np.array([0, A.shape[0]], dtype=np.int64))


def union(graphs):
    arr = (_ffi.ligmg_graph * len(graphs))()
    keep = []
    for i, g in enumerate(graphs):
        rp = np.ascontiguousarray(g.rowptr, dtype=np.int32)
        ci = np.ascontiguousarray(g.colind, dtype=np.int32)
        vv = np.ascontiguousarray(g.vals, dtype=np.float64)
        cp = np.ascontiguousarray(g.comp_ptr, dtype=np.int64)
        keep += [rp, ci, vv, cp]
        arr[i].n = g.n
        arr[i].nnz = g.nnz
        arr[i].rowptr = rp.ctypes.data_as(C.POINTER(C.c_int32))
        arr[i].colind = ci.ctypes.data_as(C.POINTER(C.c_int32))
        arr[i].vals = vv.ctypes.data_as(C.POINTER(C.c_double))
        arr[i].ncomp = g.ncomp
        arr[i].comp_ptr = cp.ctypes.data_as(C.POINTER(C.c_int64))
    return _call(_ffi.graphs_lib().ligmg_graph_union, arr, len(graphs))


def rhs(n, k=1, seed=0, comp_ptr=None):
    """k mean-zero N(0,1) right-hand sides, column-major (n, k) float64
    (returned as an (n,) vector when k == 1)."""
    out = np.empty((k, n), dtype=np.float64)
    if comp_ptr is None:
        cp, nc = None, 1
    else:
        cp = np.ascontiguousarray(comp_ptr, dtype=np.int64)
        nc = len(cp) - 1
    rc = _ffi.graphs_lib().ligmg_gen_rhs(
        n, k, seed, nc, None if cp is None else cp.ctypes.data_as(C.POINTER(C.c_int64)),
        out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise RuntimeError("rhs generator failed")
    return out[0] if k == 1 else out.T.copy(order="F")


def grid2d_config(n=256):
    """2-D 5-point grid n x n, unit weights (the CPU-runnable configuration)."""
    return grid2d(n, n)


def rmat_config(scale, weights="uniform", pendants=False, seed=0, nthreads=0):
    """R-MAT (a,b,c,d) = (.57,.19,.19,.05), edge factor 16, symmetrised, LCC;
    weights uniform[0.5,1.5] or lognormal(0,1.5); optional pendant paths of
    length uniform{1,2,3} at 10% of the vertices."""
    if weights == "uniform":
        wk, p0, p1 = W_UNIFORM, 0.5, 1.5
    elif weights == "lognormal":
        wk, p0, p1 = W_LOGNORMAL, 0.0, 1.5
    else:
        wk, p0, p1 = W_UNIT, 0.0, 0.0
    return rmat(scale, 16, 0.57, 0.19, 0.19, wk, p0, p1, 0.1 if pendants else 0.0, 3, seed,
                nthreads)
