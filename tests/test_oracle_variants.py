"""The reference's optional aggregation variants in the oracle
(aggregation.hpp:452-466 neighbour-of-neighbour voting, 622-670
reject_aggregates, 563-620 + 752-776 conductance rejection), each checked by
the property the variant exists to guarantee, on the final aggregation
states of every aggregation level."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_ffi as O
from ligmg_amd import _ffi, graphs

ST_SEED, ST_SEEDNEIGHBOR, ST_DECIDED, ST_DECIDEDROOT, ST_UNDECIDED = range(5)
AGGREGATED = (ST_SEED, ST_SEEDNEIGHBOR, ST_DECIDED, ST_DECIDEDROOT)


def agg_levels(s):
    for l in range(s.num_levels()):
        if s.level_info(l)[0] == _ffi.LEVEL_AGG:
            yield l


def graph_cases():
    yield "rmat10", graphs.rmat_config(10, "lognormal", pendants=True)
    yield "grid_w", graphs.grid2d(24, 21, wkind=_ffi.W_UNIFORM, p0=0.1, p1=10.0, seed=2)
    yield "ba", graphs.barabasi_albert(900, 4, seed=3)


@pytest.mark.parametrize("m", [2, 4])
@pytest.mark.parametrize("name,g", list(graph_cases()))
def test_max_agg_size_bounds_every_aggregate(name, g, m):
    s = O.OracleSolver(g, O.default_options(max_agg_size=m))
    base = O.OracleSolver(g)
    changed = False
    for l in agg_levels(s):
        agg = s.level_map(l)
        sizes = np.bincount(agg)
        assert sizes.max() <= m, (l, sizes.max())
        ind, st = s.level_states(l)
        assert np.all(ind >= 0)
        changed |= l < base.num_levels() and (base.level_info(l)[0] != _ffi.LEVEL_AGG
                                               or not np.array_equal(base.level_map(l), agg))
    assert changed
    x, info, _ = s.solve(graphs.rhs(g.n, 1, 0))
    assert info.status == _ffi.CONVERGED


def edge_split(s, l):
    """per vertex: external and internal off-diagonal edge weight under the
    final states, and the weight to neighbours that now stand alone
    (undecided or seed with ind = self: candidates for the final round's
    rejected vertices)"""
    rp, ci, v = s.level_matrix(l)
    n = len(rp) - 1
    ind, st = s.level_states(l)
    rows = np.repeat(np.arange(n), np.diff(rp))
    off = rows != ci
    agg_r = np.isin(st[rows], AGGREGATED)
    agg_c = np.isin(st[ci], AGGREGATED)
    inside = off & agg_r & agg_c & (ind[rows] == ind[ci])
    outside = off & agg_r & ~inside
    alone = off & agg_r & np.isin(st[ci], (ST_UNDECIDED, ST_SEED)) & (ind[ci] == ci)
    w = -v
    ext = np.bincount(rows[outside], w[outside], minlength=n)
    inn = np.bincount(rows[inside], w[inside], minlength=n)
    y = np.bincount(rows[alone], w[alone], minlength=n)
    return ext, inn, y, st


@pytest.mark.parametrize("to_seed", [0, 1])
@pytest.mark.parametrize("r", [0.8, 2.0])
@pytest.mark.parametrize("name,g", list(graph_cases()))
def test_conductance_rejection_bounds_kept_members(name, g, r, to_seed):
    """The last voting round ends with a rejection that keeps only joined
    vertices whose external / internal weight, on the states before it, is
    at most r.  Undoing that round's rejections moves weight from external
    to internal, so (ext - Y) / (int + Y), with Y the weight to every
    neighbour that now stands alone, bounds the kept vertex's ratio from
    below: it must be at most r."""
    s = O.OracleSolver(g, O.default_options(rejection_ratio=r, rejection_to_seed=to_seed))
    base = O.OracleSolver(g)
    worst_base = 0.0
    for l in agg_levels(s):
        ext, inn, y, st = edge_split(s, l)
        joined = np.isin(st, (ST_DECIDED, ST_SEEDNEIGHBOR))
        low = (ext - y)[joined] / (inn + y)[joined]
        assert np.all(low <= r * (1 + 1e-12)), (l, low.max())
        if l < base.num_levels() and base.level_info(l)[0] == _ffi.LEVEL_AGG:
            eb, ib, _, sb = edge_split(base, l)
            jb = np.isin(sb, (ST_DECIDED, ST_SEEDNEIGHBOR))
            with np.errstate(divide="ignore"):
                worst_base = max(worst_base, (eb[jb] / ib[jb]).max(initial=0.0))
    assert worst_base > r  # the option had something to reject
    x, info, _ = s.solve(graphs.rhs(g.n, 1, 0))
    assert info.status == _ffi.CONVERGED


def test_rejection_to_seed_turns_rejected_vertices_into_seeds():
    g = graphs.rmat_config(10, "lognormal", pendants=True)
    a = O.OracleSolver(g, O.default_options(rejection_ratio=0.5))
    b = O.OracleSolver(g, O.default_options(rejection_ratio=0.5, rejection_to_seed=1))
    la = next(agg_levels(a))
    assert la == next(agg_levels(b))
    assert a.level_info(la)[3] != b.level_info(la)[3]
    x, info, _ = b.solve(graphs.rhs(g.n, 1, 0))
    assert info.status == _ffi.CONVERGED


@pytest.mark.parametrize("name,g", list(graph_cases()))
def test_neighbor_of_neighbor_states(name, g):
    s = O.OracleSolver(g, O.default_options(neighbor_of_neighbor=1))
    seen = False
    for l in agg_levels(s):
        ind, st = s.level_states(l)
        assert not np.any(st == ST_DECIDED)  # joined vertices are SeedNeighbors
        sn = np.flatnonzero(st == ST_SEEDNEIGHBOR)
        seen |= len(sn) > 0
        roots = ind[sn]
        assert np.all(np.isin(st[roots], (ST_SEED, ST_DECIDEDROOT))) and np.all(ind[roots] == roots)
        # every aggregate is connected in the level graph
        rp, ci, v = s.level_matrix(l)
        agg = s.level_map(l)
        n = len(rp) - 1
        A = sp.csr_matrix((np.ones(len(ci)), ci, rp), shape=(n, n))
        rows = np.repeat(np.arange(n), np.diff(rp))
        same = agg[rows] == agg[ci]
        B = sp.csr_matrix((np.ones(same.sum()), (rows[same], ci[same])), shape=(n, n))
        k, lab = sp.csgraph.connected_components(B, directed=False)
        assert k == agg.max() + 1
    assert seen
    x, info, _ = s.solve(graphs.rhs(g.n, 1, 0))
    assert info.status == _ffi.CONVERGED


def test_rejection_needs_the_level_matrix():
    g = graphs.grid2d(10, 10)
    rp, ci, _ = O.csr_args(g)
    S = np.ones(g.nnz)
    o = O.default_options(rejection_ratio=1.0)
    out = [np.empty(g.n, np.int32) for _ in range(3)]
    nc = O.i64()
    rc = O.lib().oracle_aggregate_with_strength(g.n, O.ptr(rp, O.i32), O.ptr(ci, O.i32), O.ptr(S, O.f64),
                                                O.C.byref(o), *[O.ptr(a, O.i32) for a in out],
                                                O.C.byref(nc))
    assert rc == _ffi.ERR_UNSUPPORTED


def test_stalled_coarsening_is_reported():
    """Rejection that removes nearly every member stalls coarsening until
    max_levels; a coarse level too large for the dense solve is reported
    (LIGMG_ERR_UNSUPPORTED) instead of factorised for minutes."""
    g = graphs.barabasi_albert(12000, 5, seed=3)
    with pytest.raises(RuntimeError, match=rf"\({_ffi.ERR_UNSUPPORTED}\)"):
        O.OracleSolver(g, O.default_options(rejection_ratio=0.1, max_levels=4))
