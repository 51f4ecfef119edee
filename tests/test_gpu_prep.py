"""Device graph preprocessing through the C ABI against the oracle, bit for
bit: component labels, the largest component's renumbering and submatrix,
and the adjacency -> Laplacian conversion (canonical-order row sums)."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_ffi as O
import ligmg_amd
from ligmg_amd import _ffi, graphs, prep
from test_prep import random_components

pytestmark = pytest.mark.gpu


def adjacency_of(g):
    """The weighted adjacency matrix behind a generated Laplacian."""
    A = -g.to_scipy()
    A.setdiag(0)
    A.eliminate_zeros()
    A = A.tocsr()
    A.sort_indices()
    return A


def cases():
    yield "components7", random_components(11)
    yield "components_big", random_components(12, parts=40, isolated=300)
    g = graphs.rmat_config(12, "lognormal", pendants=True)
    yield "rmat12", adjacency_of(g)
    u = graphs.union([graphs.barabasi_albert(n, 3, seed=s) for s, n in enumerate([900, 300, 900])])
    yield "ba_union_tie", adjacency_of(u)


@pytest.mark.parametrize("mem", ["host", "device"])
@pytest.mark.parametrize("name,A", list(cases()))
def test_prep_matches_oracle(name, A, mem):
    import torch
    print(f"case={name} n={A.shape[0]} nnz={A.nnz} mem={mem}", flush=True)
    rp, ci, v = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)
    if mem == "device":
        dev = torch.device("cuda:0")
        args = (torch.from_numpy(rp).to(dev), torch.from_numpy(ci).to(dev), torch.from_numpy(v).to(dev))
        back = lambda t: t.cpu().numpy()  # noqa: E731
    else:
        args = (rp, ci, v)
        back = lambda t: t  # noqa: E731
    lab, k = prep.components(args)
    rlab, rk = O.components(rp, ci)
    assert k == rk and np.array_equal(back(lab), rlab)
    orp, oci, ov, nid = prep.largest_component(args)
    r = O.largest_component(rp, ci, v)
    for a, b in zip((orp, oci, ov, nid), r):
        assert np.array_equal(back(a), b)
    lrp, lci, lv = prep.adjacency_to_laplacian(args)
    r = O.adjacency_to_laplacian(rp, ci, v)
    for a, b in zip((lrp, lci, lv), r):
        assert np.array_equal(back(a), b)


def test_prep_full_size_rmat_pipeline():
    """R-MAT 20 adjacency (as generated, LCC already taken): components,
    largest component and Laplacian on the device against the oracle, and
    the resulting Laplacian solves."""
    g = graphs.rmat_config(20, "uniform")
    A = adjacency_of(g)
    # add isolated vertices and a small separate component at the end
    n0 = A.shape[0]
    extra = sp.coo_matrix((np.ones(2), ([n0 + 1, n0 + 2], [n0 + 2, n0 + 1])), shape=(n0 + 5, n0 + 5))
    A = (sp.block_diag([A, sp.csr_matrix((5, 5))]) + extra).tocsr()
    A.sort_indices()
    rp, ci, v = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data
    lab, k = prep.components((rp, ci, v))
    rlab, rk = O.components(rp, ci)
    assert k == rk == 5 and np.array_equal(lab, rlab)
    orp, oci, ov, nid = prep.largest_component((rp, ci, v))
    assert len(orp) - 1 == n0 and np.array_equal(nid[:n0], np.arange(n0))
    lrp, lci, lv = prep.adjacency_to_laplacian((orp, oci, ov))
    rl = O.adjacency_to_laplacian(orp, oci, ov)
    assert all(np.array_equal(a, b) for a, b in zip((lrp, lci, lv), rl))
    s = ligmg_amd.Solver((lrp, lci, lv))
    b = graphs.rhs(n0, 1, 1)
    info = s.solve(b)
    assert info.rel_residual() <= 1e-8


def test_invalid_csr_is_rejected_before_use():
    """Malformed input becomes an error code: row pointers are checked
    before any column is read, columns before any gather."""
    lib = _ffi.solver_lib()
    import ctypes as C
    rp = np.array([0, 2, 1, 3], np.int32)  # decreasing
    ci = np.array([0, 1, 2], np.int32)
    lab = np.empty(3, np.int32)
    k = C.c_int64()
    assert lib.ligmg_components(3, rp.ctypes.data, ci.ctypes.data, _ffi.MEM_HOST, lab.ctypes.data,
                                C.byref(k)) == _ffi.ERR_INVALID
    rp = np.array([0, 2, 3, 3], np.int32)
    ci = np.array([0, 7, 1], np.int32)  # column out of range
    assert lib.ligmg_components(3, rp.ctypes.data, ci.ctypes.data, _ffi.MEM_HOST, lab.ctypes.data,
                                C.byref(k)) == _ffi.ERR_INVALID
    assert b"out of range" in lib.ligmg_last_error()
    # the solver checks the same and also wants sorted rows
    g = graphs.grid2d(4, 4)
    ci2 = g.colind.copy()
    ci2[1], ci2[2] = ci2[2], ci2[1]
    with pytest.raises(ligmg_amd.LigmgError):
        ligmg_amd.Solver((g.rowptr, ci2, g.vals))
