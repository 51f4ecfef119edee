"""The fused coarse tail (the levels whose operators hold at most tail_nnz
entries run as one single-workgroup launch per coarse-grid correction)
against the per-operator launches of the same hierarchy: every variant of
the cycle must give the same bits - solutions, residual histories, cycles.
"""
import numpy as np
import pytest

import ligmg_amd
from ligmg_amd import _ffi, graphs

pytestmark = pytest.mark.gpu

FCG = {"outer_ksp_type": _ffi.KSP_FCG}
VARIANTS = {
    "vcycle": {},
    "kcycle_fcg": {"kcycle": 1, **FCG},
    "kcycle_cg_inner": {"kcycle": 1, "kcycle_ksp_type": _ffi.KSP_CG, **FCG},
    "kcycle_3_iters": {"kcycle": 1, "kcycle_iters": 3, **FCG},
    "jacobi_kcycle": {"smoother_type": _ffi.SMOOTHER_JACOBI, "kcycle": 1, **FCG},
    "constant_correction": {"constant_correction": 1, **FCG},
    "smooth_3_1": {"pre_smooth": 3, "post_smooth": 1, "kcycle": 1, **FCG},
    "no_pre_smooth": {"pre_smooth": 0, "post_smooth": 2, **FCG},
    "mixed_kcycle": {"precision": _ffi.PREC_MIXED, "kcycle": 1, **FCG},
    "mixed_vcycle": {"precision": _ffi.PREC_MIXED, **FCG},
}


def pair(g, tail_nnz, comp_ptr=None, **kw):
    on = ligmg_amd.Solver(g, ligmg_amd.Options(tail_nnz=tail_nnz, iters=60, **kw), comp_ptr=comp_ptr)
    off = ligmg_amd.Solver(g, ligmg_amd.Options(tail_nnz=0, iters=60, **kw), comp_ptr=comp_ptr)
    assert off.tail_start() == off.num_levels()
    assert on.tail_start() < on.num_levels(), "nothing fused"
    return on, off


def same_solve(on, off, b):
    a, c = on.solve(b), off.solve(b)
    assert a.iters() == c.iters()
    assert np.array_equal(a.residuals, c.residuals, equal_nan=True)
    assert np.array_equal(a.solution, c.solution, equal_nan=True)
    return a


@pytest.mark.parametrize("variant", list(VARIANTS))
def test_tail_same_bits(variant):
    g = graphs.rmat_config(13, "lognormal", pendants=True)
    kw = VARIANTS[variant]
    on, off = pair(g, 1 << 15, **kw)
    print(f"variant={variant} levels={on.num_levels()} tail_start={on.tail_start()}", flush=True)
    info = same_solve(on, off, graphs.rhs(g.n, 1, 0))
    print(f"iters={info.iters()} rel={info.rel_residual():.2e}")
    base = 1 if on.level_info(0)[0] == _ffi.LEVEL_ELIM else 0
    u = graphs.rhs(on.level_info(base)[1], 1, 7)
    precs = [_ffi.PREC_FP64] + ([_ffi.PREC_MIXED] if kw.get("precision") == _ffi.PREC_MIXED else [])
    for p in precs:
        assert np.array_equal(on.apply_cycle(base, u, p), off.apply_cycle(base, u, p))


@pytest.mark.parametrize("prec", [_ffi.PREC_FP64, _ffi.PREC_MIXED])
def test_tail_long_rows_whole_hierarchy(prec):
    """Everything below the base level fused, including levels with rows
    longer than one SpMV unit (chunked partial sums)."""
    g = graphs.rmat_config(16, "lognormal", pendants=True)
    on, off = pair(g, 1 << 30, precision=prec, kcycle=1, **FCG)
    base = 1 if on.level_info(0)[0] == _ffi.LEVEL_ELIM else 0
    print(f"levels={on.num_levels()} tail_start={on.tail_start()} base={base}", flush=True)
    assert on.tail_start() == base + 1
    longest = 0
    for l in range(on.tail_start(), on.num_levels()):
        rp = on.level_matrix(l)[0]
        longest = max(longest, int(np.diff(rp).max()))
    assert longest > 2048  # the chunked path is exercised
    info = same_solve(on, off, graphs.rhs(g.n, 1, 2))
    assert info.rel_residual() <= 1e-8


def test_tail_many_blocks():
    """Per-block Krylov coefficients and nullspace projection inside the
    tail: a union of graphs of different sizes, one block each."""
    gs = [graphs.barabasi_albert(n, 4, seed=s) for s, n in enumerate([3000, 800, 5000, 120, 2000])]
    u = graphs.union(gs)
    for kw in ({"kcycle": 1, **FCG}, {}, {"precision": _ffi.PREC_MIXED, "kcycle": 1, **FCG}):
        on, off = pair(u, 1 << 15, comp_ptr=u.comp_ptr, **kw)
        print(f"opts={kw} levels={on.num_levels()} tail_start={on.tail_start()}", flush=True)
        info = same_solve(on, off, graphs.rhs(u.n, 1, 5, comp_ptr=u.comp_ptr))
        assert info.rel_residual() <= 1e-8


def test_tail_graph_replays_and_repeat_solves():
    """Captured preconditioner replays and repeated solves reuse the tail's
    buffers: the same bits every time."""
    g = graphs.rmat_config(14, "uniform")
    on, off = pair(g, 1 << 15, kcycle=1, **FCG)
    b = graphs.rhs(g.n, 1, 9)
    first = same_solve(on, off, b)
    for _ in range(2):
        again = on.solve(b)
        assert np.array_equal(again.solution, first.solution)
        assert np.array_equal(again.residuals, first.residuals)
    eager = ligmg_amd.Solver(g, ligmg_amd.Options(tail_nnz=1 << 15, use_graphs=0, kcycle=1, **FCG))
    assert np.array_equal(eager.solve(b).solution, first.solution)
