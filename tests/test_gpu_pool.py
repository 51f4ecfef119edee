"""Pooled device memory: setups and solves that reuse blocks released by
earlier solvers (parked until the device has finished with them) give the
same bits as the first, and returning the cache to HIP between them changes
nothing."""
import numpy as np
import pytest

import ligmg_amd
from ligmg_amd import _ffi, graphs

pytestmark = pytest.mark.gpu


def run(g, kw, rhs):
    s = ligmg_amd.Solver(g, ligmg_amd.Options(**kw))
    info = s.solve(rhs)
    out = (s.num_levels(), [s.level_matrix(l) for l in range(s.num_levels())], info.solution, info.residuals)
    s.close()
    return out


def same(a, b):
    assert a[0] == b[0]
    for x, y in zip(a[1], b[1]):
        for u, v in zip(x, y):
            assert np.array_equal(u, v)
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])


@pytest.mark.parametrize("kw", [{}, {"kcycle": 1, "outer_ksp_type": _ffi.KSP_FCG, "precision": _ffi.PREC_MIXED}],
                         ids=["vcycle", "kcycle_mixed"])
def test_repeated_setups_reuse_memory(kw):
    g = graphs.rmat_config(16, "lognormal", pendants=True)
    small = graphs.grid2d(70, 60)
    rhs, rhs_small = graphs.rhs(g.n, 1, 0), graphs.rhs(small.n, 1, 0)
    first = run(g, kw, rhs)
    print("first setup done", flush=True)
    for rep in range(3):
        run(small, kw, rhs_small)  # other sizes take and return blocks in between
        same(first, run(g, kw, rhs))
        print(f"rep {rep} same", flush=True)
    ligmg_amd.trim_memory()
    same(first, run(g, kw, rhs))


def test_solver_kept_while_others_come_and_go():
    """Blocks of a live solver are never handed out again."""
    g = graphs.rmat_config(14, "uniform")
    rhs = graphs.rhs(g.n, 1, 3)
    keep = ligmg_amd.Solver(g)
    ref = keep.solve(rhs).solution
    for rep in range(3):
        run(graphs.barabasi_albert(4000 + 1000 * rep, 4, seed=rep), {}, graphs.rhs(4000 + 1000 * rep, 1, 0))
        ligmg_amd.trim_memory() if rep == 1 else None
        assert np.array_equal(keep.solve(rhs).solution, ref)


def test_concurrent_handles_with_graph_capture():
    """Setups on one thread (allocating from the pool) while another thread's
    handle captures and replays its preconditioner graph: both get the bits
    of a sequential run."""
    import threading
    ga = graphs.rmat_config(14, "lognormal", pendants=True)
    gb = graphs.grid2d(120, 110, wkind=_ffi.W_UNIFORM, p0=0.1, p1=10.0, seed=4)
    ra, rb = graphs.rhs(ga.n, 1, 2), graphs.rhs(gb.n, 1, 3)
    ref_a = ligmg_amd.Solver(ga).solve(ra).solution
    ref_b = ligmg_amd.Solver(gb, ligmg_amd.Options(use_graphs=1)).solve(rb).solution
    errors = []

    def setups():
        try:
            for _ in range(4):
                s = ligmg_amd.Solver(ga)
                if not np.array_equal(s.solve(ra).solution, ref_a):
                    errors.append("setup thread: different bits")
                s.close()
        except Exception as e:  # noqa: BLE001
            errors.append(repr(e))

    def solves():
        try:
            s = ligmg_amd.Solver(gb, ligmg_amd.Options(use_graphs=1))
            for _ in range(6):
                if not np.array_equal(s.solve(rb).solution, ref_b):
                    errors.append("graph thread: different bits")
            s.close()
        except Exception as e:  # noqa: BLE001
            errors.append(repr(e))

    ts = [threading.Thread(target=setups), threading.Thread(target=solves)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in ts)
    assert not errors, errors
