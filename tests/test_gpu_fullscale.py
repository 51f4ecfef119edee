"""The headline workload at full size: R-MAT 23 with pendant paths and
lognormal(0, 1.5) weights (5,530,432 vertices, 266,041,316 stored entries).
The device hierarchy equals the oracle's on every level: elimination sets,
aggregates, states, strengths, coarse matrices (pattern and values) and the
Chebyshev bounds; the FP64 solve (V-cycle CG to 1e-8) takes the same number
of iterations with the same residual history up to the rounding of
differently ordered inner products.  The oracle runs its OpenMP loops, whose
bits do not depend on the thread count."""
import time

import numpy as np
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
    b = graphs.rhs(g.n, 1, 0)
    info = d.solve(b)
    t = time.perf_counter()
    xr, rinfo, rhist = r.solve(b)
    print(f"device solve {info.time():.3f}s {info.iters()} iterations; oracle solve {time.perf_counter() - t:.1f}s "
          f"{rinfo.iters} iterations", flush=True)
    assert info.rel_residual() <= 1e-8 and abs(info.iters() - rinfo.iters) <= 1
    k = min(10, len(rhist), len(info.residuals))
    assert np.allclose(np.asarray(info.residuals)[:k], rhist[:k], rtol=1e-7)
    x = info.solution
    xm, xrm = x - x.mean(), xr - xr.mean()
    assert np.linalg.norm(xm - xrm) <= 1e-6 * np.linalg.norm(xrm)
