"""Graph ingestion and preprocessing: the MatrixMarket reader against
scipy.io.mmread, and the oracle's restatements of connected_components /
largest_cc (cc.hpp:22-105) and the adjacency -> Laplacian step
(main.cpp:79-110) against scipy.sparse.csgraph and numpy."""
import math
import zlib

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

import oracle_ffi as O
from ligmg_amd import graphs


def scipy_csr(path):
    A = scipy.io.mmread(str(path)).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    return A


def test_read_mtx_reference_fixture(golden_dir):
    path = golden_dir / "my_matrix.mtx"
    g = graphs.read_mtx(path)
    A = scipy_csr(path)
    assert g.n == A.shape[0]
    assert np.array_equal(g.rowptr, A.indptr) and np.array_equal(g.colind, A.indices)
    assert np.array_equal(g.vals, A.data)


@pytest.mark.parametrize("field,symmetry", [("real", "general"), ("real", "symmetric"),
                                            ("integer", "general"), ("pattern", "symmetric"),
                                            ("pattern", "general")])
def test_read_mtx_formats(tmp_path, field, symmetry):
    rng = np.random.default_rng(zlib.crc32(f"{field}{symmetry}".encode()))
    n = 57
    i = rng.integers(0, n, 300)
    j = rng.integers(0, n, 300)
    if symmetry == "symmetric":
        i, j = np.maximum(i, j), np.minimum(i, j)  # lower triangle
    v = rng.integers(-9, 10, 300) if field == "integer" else rng.standard_normal(300)
    path = tmp_path / "a.mtx"
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {field} {symmetry}\n% a comment\n%\n")
        f.write(f"{n} {n} {len(i)}\n")
        for a, b, c in zip(i, j, v):  # duplicates included on purpose
            f.write(f"{a + 1} {b + 1}\n" if field == "pattern" else f"{a + 1} {b + 1} {float(c):.17g}\n")
    g = graphs.read_mtx(path)
    vals = np.ones(len(i)) if field == "pattern" else v.astype(np.float64)
    ii, jj, vv = i, j, vals
    if symmetry == "symmetric":
        off = i != j
        ii, jj, vv = np.concatenate([i, j[off]]), np.concatenate([j, i[off]]), np.concatenate([vals, vals[off]])
    A = sp.coo_matrix((vv, (ii, jj)), shape=(n, n)).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    assert np.array_equal(g.rowptr, A.indptr) and np.array_equal(g.colind, A.indices)
    assert np.allclose(g.vals, A.data, rtol=0, atol=1e-12 * np.abs(A.data).max())


def test_read_mtx_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.mtx"
    p.write_text("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n")
    with pytest.raises(ValueError):
        graphs.read_mtx(p)
    p.write_text("%%MatrixMarket matrix coordinate real general\n2 3 1\n1 1 1.0\n")
    with pytest.raises(ValueError):
        graphs.read_mtx(p)
    p.write_text("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n")
    with pytest.raises(ValueError):
        graphs.read_mtx(p)


def random_components(seed, n=400, parts=7, isolated=9):
    """A random symmetric adjacency pattern with several components and
    isolated vertices, vertex ids shuffled."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(5, 90, parts)
    base = np.concatenate([[0], np.cumsum(sizes)])
    us, vs = [], []
    for p in range(parts):
        k = sizes[p]
        # a random spanning path plus extra edges inside the part
        order = rng.permutation(k)
        us += list(base[p] + order[:-1])
        vs += list(base[p] + order[1:])
        e = rng.integers(0, k, (k, 2))
        us += list(base[p] + e[:, 0])
        vs += list(base[p] + e[:, 1])
    total = base[-1] + isolated
    perm = rng.permutation(total)
    u, v = perm[np.array(us)], perm[np.array(vs)]
    keep = u != v
    u, v = u[keep], v[keep]
    w = rng.uniform(0.5, 2.0, len(u))
    A = sp.coo_matrix((np.concatenate([w, w]), (np.concatenate([u, v]), np.concatenate([v, u]))),
                      shape=(total, total)).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    return A


@pytest.mark.parametrize("seed", range(5))
def test_oracle_components_match_scipy(seed):
    A = random_components(seed)
    lab, k = O.components(A.indptr, A.indices)
    kr, ref = connected_components(A, directed=False)
    assert k == kr
    for c in range(kr):
        members = np.where(ref == c)[0]
        assert np.all(lab[members] == members.max())  # the reference's label: largest index


@pytest.mark.parametrize("seed", range(3))
def test_oracle_largest_component_and_extract(seed):
    A = random_components(seed)
    orp, oci, ov, new_id = O.largest_component(A.indptr, A.indices, A.data)
    kr, ref = connected_components(A, directed=False)
    sizes = np.bincount(ref)
    big = np.flatnonzero(sizes == sizes.max())
    # among equally large components the one with the smallest label wins
    labels = [np.where(ref == c)[0].max() for c in big]
    keep = np.sort(np.where(ref == big[int(np.argmin(labels))])[0])
    assert np.array_equal(np.where(new_id >= 0)[0], keep)
    assert np.array_equal(new_id[keep], np.arange(len(keep)))
    S = A[keep][:, keep].tocsr()
    S.sort_indices()
    assert np.array_equal(orp, S.indptr) and np.array_equal(oci, S.indices) and np.array_equal(ov, S.data)


def test_oracle_largest_component_tie_goes_to_smallest_label():
    # two triangles {0,1,2} and {3,4,5}: labels 2 and 5, the first wins
    e = [(0, 1), (1, 2), (0, 2), (3, 4), (4, 5), (3, 5)]
    u = np.array([a for a, b in e] + [b for a, b in e])
    v = np.array([b for a, b in e] + [a for a, b in e])
    A = sp.coo_matrix((np.ones(len(u)), (u, v)), shape=(7, 7)).tocsr()
    A.sort_indices()
    _, _, _, new_id = O.largest_component(A.indptr, A.indices, A.data)
    assert list(new_id) == [0, 1, 2, -1, -1, -1, -1]


def test_oracle_laplacian_matches_numpy():
    A = random_components(3)
    # self loops on a few rows: L_ii = deg_i - A_ii with deg counting them
    A = (A + sp.diags(np.r_[np.zeros(A.shape[0] - 3), [1.5, 0.25, 4.0]])).tocsr()
    A.sort_indices()
    orp, oci, ov = O.adjacency_to_laplacian(A.indptr, A.indices, A.data)
    n = A.shape[0]
    deg = np.array([math.fsum(A.data[A.indptr[i]:A.indptr[i + 1]]) for i in range(n)])
    L = sp.csr_matrix((ov, oci, orp), shape=(n, n))
    ref = (sp.diags(deg) - A).tocsr()
    assert np.allclose(L.toarray(), ref.toarray(), rtol=0, atol=1e-13)
    # pattern: A's pattern plus every diagonal, columns sorted
    P = (A != 0).astype(np.int8) + sp.eye(n, dtype=np.int8)
    P = P.tocsr()
    P.sort_indices()
    assert np.array_equal(orp, P.indptr) and np.array_equal(oci, P.indices)
    # integer weights: sums are exact, so rows of L sum to zero exactly
    B = A.copy()
    B.data = np.round(B.data * 4)
    orp, oci, ov = O.adjacency_to_laplacian(B.indptr, B.indices, B.data)
    L = sp.csr_matrix((ov, oci, orp), shape=(n, n))
    assert np.all(np.asarray(L.sum(axis=1)).ravel() == 0)
