"""Solver::solve(b, from) (ligmg.hpp:36-38 -> ligmg::solve,
multigrid.hpp:315-354): the system of any hierarchy level, with that level's
elimination applied once, against the oracle's solve of the same level."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_ffi as O
import ligmg_amd
from ligmg_amd import _ffi, graphs

pytestmark = pytest.mark.gpu

VARIANTS = {
    "vcycle_cg": {},
    "kcycle_fcg": {"kcycle": 1, "outer_ksp_type": _ffi.KSP_FCG},
    "mixed_fcg": {"precision": _ffi.PREC_MIXED, "outer_ksp_type": _ffi.KSP_FCG},
    "fgmres": {"outer_ksp_type": _ffi.KSP_FGMRES},
}


def level_system(s, l, seed):
    rp, ci, v = s.level_matrix(l)
    n = len(rp) - 1
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    b = np.random.default_rng(seed).standard_normal(n)
    return A, b - b.mean()


@pytest.fixture(scope="module")
def pend():
    return graphs.rmat_config(12, "lognormal", pendants=True)


def pick_levels(kinds):
    """The first elimination and the first aggregation level below the top
    (the two ways solve(b, from) starts)."""
    el = next(l for l in range(1, len(kinds)) if kinds[l] == _ffi.LEVEL_ELIM)
    ag = next(l for l in range(1, len(kinds)) if kinds[l] == _ffi.LEVEL_AGG)
    return el, ag


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("which", ["elim", "agg"])
def test_solve_from_matches_oracle(pend, which, variant):
    kw = VARIANTS[variant]
    d = ligmg_amd.Solver(pend, ligmg_amd.Options(**kw))
    r = O.OracleSolver(pend, O.default_options(**{k: v for k, v in kw.items() if k != "precision"}))
    kinds = [d.level_info(l)[0] for l in range(d.num_levels())]
    el, ag = pick_levels(kinds)
    l = el if which == "elim" else ag
    A, b = level_system(d, l, seed=l)
    print(f"variant={variant} level={l} kind={_ffi.LEVEL_NAMES[kinds[l]]} n={A.shape[0]}", flush=True)
    info = d.solve(b, level=l)
    xr, rinfo, rhist = r.solve(b, level=l)
    assert info.status == _ffi.CONVERGED and info.rel_residual() <= 1e-8
    assert info.base_level == (l + 1 if kinds[l] == _ffi.LEVEL_ELIM else l) == rinfo.base_level
    # the same preconditioner: residual histories coincide until rounding in
    # the differently ordered inner products builds up (FP32 cycles: r0 only)
    k = 1 if "mixed" in variant else min(6, len(rhist), len(info.residuals))
    assert np.allclose(np.asarray(info.residuals)[:k], rhist[:k], rtol=1e-7)
    assert abs(info.iters() - rinfo.iters) <= 2
    x = info.solution
    xm, xrm = x - x.mean(), xr - xr.mean()
    assert np.linalg.norm(xm - xrm) <= 1e-6 * np.linalg.norm(xrm)
    full = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert info.full_rel_residual == pytest.approx(full, rel=1e-6, abs=1e-14)
    assert info.cycle_complexity() == pytest.approx(rinfo.cycle_complexity, rel=1e-12)


def test_solve_from_elimination_level_without_top_elim(pend):
    """top_elim off: the outer Krylov solve runs on the elimination level's
    own matrix, which the cycles never apply."""
    d = ligmg_amd.Solver(pend, ligmg_amd.Options(top_elim=0))
    r = O.OracleSolver(pend, O.default_options(top_elim=0))
    kinds = [d.level_info(l)[0] for l in range(d.num_levels())]
    l, _ = pick_levels(kinds)
    A, b = level_system(d, l, seed=3)
    info = d.solve(b, level=l)
    xr, rinfo, rhist = r.solve(b, level=l)
    assert info.base_level == l == rinfo.base_level
    assert info.status == _ffi.CONVERGED and info.rel_residual() <= 1e-8
    k = min(6, len(rhist), len(info.residuals))
    assert np.allclose(np.asarray(info.residuals)[:k], rhist[:k], rtol=1e-7)
    full = np.linalg.norm(b - A @ info.solution) / np.linalg.norm(b)
    assert info.full_rel_residual == pytest.approx(full, rel=1e-6, abs=1e-14)


def test_solve_from_level0_is_solve(pend):
    d = ligmg_amd.Solver(pend)
    b = graphs.rhs(pend.n, 1, 3)
    a, c = d.solve(b), d.solve(b, level=0)
    assert np.array_equal(a.solution, c.solution) and np.array_equal(a.residuals, c.residuals)
    with pytest.raises(ligmg_amd.LigmgError):
        d.solve(np.zeros(3), level=d.num_levels())


def test_solve_from_renumbered_level():
    """A level of 2^16 vertices or more is stored in degree order for the
    solve; solve(b, from) takes and returns vectors in the level's setup
    order either way."""
    g = graphs.rmat_config(18, "lognormal", pendants=True)
    on = ligmg_amd.Solver(g, ligmg_amd.Options(solve_layout=1))
    off = ligmg_amd.Solver(g, ligmg_amd.Options(solve_layout=0))
    kinds = [on.level_info(l)[0] for l in range(on.num_levels())]
    l = next(l for l in range(1, len(kinds)) if kinds[l] == _ffi.LEVEL_AGG)
    assert on.level_info(l)[1] >= 1 << 16
    A, b = level_system(on, l, seed=9)
    ia, ib = on.solve(b, level=l), off.solve(b, level=l)
    print(f"level={l} n={A.shape[0]} iters {ia.iters()} / {ib.iters()}")
    assert ia.rel_residual() <= 1e-8 and ib.rel_residual() <= 1e-8
    assert abs(ia.iters() - ib.iters()) <= 1
    x, y = ia.solution - ia.solution.mean(), ib.solution - ib.solution.mean()
    assert np.linalg.norm(x - y) <= 1e-7 * np.linalg.norm(y)
    full = np.linalg.norm(b - A @ ia.solution) / np.linalg.norm(b)
    assert ia.full_rel_residual == pytest.approx(full, rel=1e-6, abs=1e-14)
