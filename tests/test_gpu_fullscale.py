"""The headline workload at full size: R-MAT 23 with pendant paths and
lognormal(0, 1.5) weights (5,530,432 vertices, 266,041,316 stored entries).
The device hierarchy equals the oracle's on every level: elimination sets,
aggregates, states, strengths, coarse matrices (pattern and values) and the
Chebyshev bounds.  The oracle runs its OpenMP loops, whose bits do not depend
on the thread count."""
import time

import pytest

import oracle_ffi as O
import ligmg_amd
from ligmg_amd import graphs
from test_gpu_parity import assert_same_hierarchy

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(1200)
def test_rmat23_hierarchy_bit_exact():
    t = time.perf_counter()
    g = graphs.rmat_config(23, "lognormal", pendants=True)
    print(f"rmat23 n={g.n} nnz={g.nnz} generated in {time.perf_counter() - t:.1f}s", flush=True)
    d = ligmg_amd.Solver(g)
    print(f"device setup {d.setup_time:.2f}s, {d.num_levels()} levels", flush=True)
    t = time.perf_counter()
    r = O.OracleSolver(g, O.default_options())
    print(f"oracle setup {time.perf_counter() - t:.1f}s", flush=True)
    t = time.perf_counter()
    assert_same_hierarchy(d, r)
    print(f"compared in {time.perf_counter() - t:.1f}s", flush=True)
