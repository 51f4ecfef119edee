"""Hand-built strength inputs where the summation order of the two sums that
decide aggregation changes the outcome.

mark_nodes (aggregation.hpp:742-748) seeds an undecided vertex i when
votes[i] >= votes_to_seed * col_sum[i].  tally_votes adds each voter's weight
in ascending voter index (aggregation.hpp:539-547) and col_sum folds column i
in ascending row order (Reduce(Column), aggregation.hpp:863-864).  When every
stored strength entry of column i comes from a voter for i in the same
iteration, the reference adds the same numbers in the same order on both
sides, so the comparison is an exact equality and i becomes a Seed.  Any
other order (a strided tree over the row including the zero diagonal, or the
voters in another order) can round one side differently and leave i
undecided.

The case: a centre c with m neighbours j, each j with one private partner
p_j, all at random indices.  Strengths (injected on the symmetric pattern):
  S[j, c] in [0.5, 1)   every neighbour's best entry, at or above theta_0;
  S[c, j] in [0.01, 0.4) so c's own vote is below theta_0 and c casts no
                          vote: the mutual-vote rule (aggregation.hpp:722)
                          does not apply to it;
  S[j, p_j] = 0.1, S[p_j, j] = 0.3  partners neither vote at iteration 0 nor
                          pull j away from c.
"""
import numpy as np
import scipy.sparse as sp

from ligmg_amd import graphs


def knife_edge(m, seed):
    """Returns (graph, S, c): a unit-weight Laplacian on the case's pattern,
    the strength values on that pattern (CSR order) and the centre's index."""
    rng = np.random.default_rng(seed)
    n = 2 * m + 1
    perm = rng.permutation(n)
    c = int(perm[0])
    nb = perm[1:m + 1].astype(np.int64)
    pt = perm[m + 1:].astype(np.int64)
    u = np.concatenate([np.full(m, c), nb])
    v = np.concatenate([nb, pt])
    A = sp.coo_matrix((np.ones(2 * m), (u, v)), shape=(n, n)).tocsr()
    A = A + A.T
    L = sp.diags(np.asarray(A.sum(axis=1)).ravel()) - A
    g = graphs.from_scipy(L.tocsr())
    # strength values by (row, col)
    want = {}
    w_jc = rng.uniform(0.5, 1.0, m)
    w_cj = rng.uniform(0.01, 0.4, m)
    for t in range(m):
        j, p = int(nb[t]), int(pt[t])
        want[(j, c)] = w_jc[t]
        want[(c, j)] = w_cj[t]
        want[(j, p)] = 0.1
        want[(p, j)] = 0.3
    rows = np.repeat(np.arange(g.n), np.diff(g.rowptr))
    S = np.array([want.get((int(r), int(k)), 0.0) for r, k in zip(rows, g.colind)])
    return g, S, c


def column_in_row_order(g, S, c):
    """The values of column c in ascending row order (what Reduce(Column)
    folds and, when every one of them votes, what tally_votes adds)."""
    rows = np.repeat(np.arange(g.n), np.diff(g.rowptr))
    sel = g.colind == c
    order = np.argsort(rows[sel], kind="stable")
    return S[sel][order]


def _fold(vals):
    s = 0.0
    for x in vals:
        s += float(x)
    return s


def _canon(vals):
    import oracle_ffi as O
    return O.canon_sum(np.asarray(vals, dtype=np.float64))


def flipping_case(m, max_seed=200):
    """The first seed whose case the reference's order seeds while both
    planted orders would not: the canonical split (tally over the m voters,
    col_sum over the m + 1 row positions with the zero diagonal at c's rank)
    and a descending-voter tally each come out below the ascending col_sum."""
    for seed in range(max_seed):
        g, S, c = knife_edge(m, seed)
        col = column_in_row_order(g, S, c)
        voters = col[col != 0.0]
        asc = _fold(col)
        desc = _fold(voters[::-1])
        k = int(np.sum(g.colind[g.rowptr[c]:g.rowptr[c + 1]] < c))
        with_diag = np.concatenate([voters[:k], [0.0], voters[k:]])
        if _canon(voters) < _canon(with_diag) and desc < asc:
            return g, S, c, seed
    raise RuntimeError("no flipping seed found")
