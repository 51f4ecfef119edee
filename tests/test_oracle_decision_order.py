"""The reference's one-rank order for the sums that decide aggregation,
pinned in the oracle.

tally_votes adds voter weights left to right in ascending voter index
(aggregation.hpp:539-547); col_sum is Reduce(Column, plus, 0.0), each column
folded in ascending row order (aggregation.hpp:863-864).  The tests check
the oracle against literal restatements of those loops and on hand-built
cases (tests/decision_cases.py) where the order decides a seed, with the
wrong orders planted in the oracle to show that each case tells them apart.
"""
import numpy as np
import pytest

import decision_cases as D
import oracle_ffi as O
from ligmg_amd import _ffi, graphs

ST_SEED, ST_DECIDED, ST_DECIDEDROOT, ST_UNDECIDED = 0, 2, 3, 4


@pytest.fixture
def plant():
    yield lambda mode: O.lib().oracle_set_decision_order(mode)
    O.lib().oracle_set_decision_order(0)


def oracle_aggregate(g, S, **kw):
    rp, ci, _ = O.csr_args(g)
    S = np.ascontiguousarray(S, np.float64)
    agg, ind, st = (np.empty(g.n, np.int32) for _ in range(3))
    nc = O.i64()
    o = O.default_options(**kw)
    assert O.lib().oracle_aggregate_with_strength(
        g.n, O.ptr(rp, O.i32), O.ptr(ci, O.i32), O.ptr(S, O.f64), O.C.byref(o), O.ptr(agg, O.i32),
        O.ptr(ind, O.i32), O.ptr(st, O.i32), O.C.byref(nc)) == 0
    return agg, ind, st, nc.value


def oracle_strength(g, scaling):
    rp, ci, v = O.csr_args(g)
    T = 10
    X = np.empty(g.n * T)
    assert O.lib().oracle_tv_init(g.n, 7, 0, T, O.ptr(X, O.f64)) == 0
    S = np.empty(g.nnz)
    cs = np.empty(g.n)
    assert O.lib().oracle_strength(g.n, O.ptr(rp, O.i32), O.ptr(ci, O.i32), O.ptr(v, O.f64),
                                   O.ptr(X, O.f64), T, 4, _ffi.STRENGTH_AFFINITY, scaling, 1e-8,
                                   O.ptr(S, O.f64), O.ptr(cs, O.f64)) == 0
    return S, cs


def reduce_column(g, S):
    """Reduce(Column, plus, 0.0) on one rank: column by column, each column's
    stored entries in ascending row order, added to an accumulator that
    starts at 0.0.  Written as the reference's DCSC traversal: rows visited
    in ascending order, each adding into its columns' accumulators, which
    gives every column its entries in ascending row."""
    cs = [0.0] * g.n
    for r in range(g.n):
        for k in range(g.rowptr[r], g.rowptr[r + 1]):
            s = float(S[k])
            if s != 0.0:  # the pruned matrix stores no zero
                cs[g.colind[k]] += s
    return np.array(cs)


def test_seq_sum_is_left_fold():
    rng = np.random.default_rng(0)
    for m in [0, 1, 63, 64, 65, 257, 4097, 10000]:
        a = rng.standard_normal(m) * np.exp(rng.uniform(-20, 20, m))
        a = np.ascontiguousarray(a)
        assert O.lib().oracle_seq_sum(O.ptr(a, O.f64), m) == D._fold(a)


@pytest.mark.parametrize("scaling", [_ffi.AFF_MAX, _ffi.AFF_ROW])
def test_col_sum_is_reference_reduce(scaling, plant):
    """col_sum equals the reference's Reduce(Column) bit for bit on a graph
    with hub columns of up to several hundred entries; the canonical split
    planted in the oracle does not."""
    g = graphs.rmat(11, 16, seed=3)
    assert np.diff(g.rowptr).max() > 256
    S, cs = oracle_strength(g, scaling)
    ref = reduce_column(g, S)
    assert np.array_equal(cs, ref)
    plant(1)
    _, cs_canon = oracle_strength(g, scaling)
    assert not np.array_equal(cs_canon, ref)


@pytest.mark.parametrize("m", [100, 300, 5000])
def test_knife_edge_seeds_centre(m, plant):
    """Every neighbour of c votes for it at iteration 0 and c casts no vote:
    the reference's order makes votes[c] == col_sum[c] and c a Seed.  The
    canonical split and a descending-voter tally, planted in the oracle,
    both leave c undecided on the same case."""
    g, S, c, seed = D.flipping_case(m)
    print(f"m={m} seed={seed} centre={c}")
    _, ind, st, _ = oracle_aggregate(g, S, agg_iters=1)
    assert st[c] == ST_SEED and ind[c] == c
    for mode in (1, 2):
        plant(mode)
        _, ind2, st2, _ = oracle_aggregate(g, S, agg_iters=1)
        assert st2[c] == ST_UNDECIDED, mode
    plant(0)
    # all ten rounds: the neighbours join c's aggregate one round later
    agg, ind, st, nc = oracle_aggregate(g, S)
    nb = g.colind[g.rowptr[c]:g.rowptr[c + 1]]
    nb = nb[nb != c]
    assert st[c] == ST_DECIDEDROOT and np.all(ind[nb] == c) and np.all(agg[nb] == agg[c])


def test_knife_edge_two_rounds_differ(plant):
    """With two voting rounds the seeding delay under a wrong order changes
    the aggregates themselves: c's neighbours stay singletons."""
    g, S, c, _ = D.flipping_case(300)
    agg, _, _, nc = oracle_aggregate(g, S, agg_iters=2)
    plant(1)
    agg1, _, _, nc1 = oracle_aggregate(g, S, agg_iters=2)
    assert nc1 > nc and not np.array_equal(agg, agg1)


def test_oracle_results_do_not_depend_on_threads():
    """The oracle's OpenMP loops (the CPU baseline's multi-threaded leg) give
    the single-threaded bits: same hierarchy, same solution."""
    g = graphs.rmat_config(12, "lognormal", pendants=True)
    b = graphs.rhs(g.n, 1, 0)
    out = []
    for t in (1, 8):
        assert O.lib().oracle_set_threads(t) == t
        r = O.OracleSolver(g)
        x, info, hist = r.solve(b)
        mats = [r.level_matrix(l) for l in range(r.num_levels())]
        out.append((x, hist, mats))
    O.lib().oracle_set_threads(1)
    (x1, h1, m1), (x8, h8, m8) = out
    assert np.array_equal(x1, x8) and np.array_equal(h1, h8)
    for a, c in zip(m1, m8):
        for u, v in zip(a, c):
            assert np.array_equal(u, v)
