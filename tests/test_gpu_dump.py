"""Hierarchy dumps (dump_hierarchy, ligmg.hpp:47-52), loads and the CSV
stats line (dump_csv, ligmg.hpp:42-45) through the C ABI.

A dump holds every level in PETSc binary format, in setup order; its files
must agree with the level introspection bit for bit, and a solver loaded
from it must solve exactly as the dumped one (same bits), including levels
stored in the solve layout, the degree input order and many blocks."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import ligmg_amd
from ligmg_amd import _ffi, graphs, petsc_io

pytestmark = pytest.mark.gpu


def check_files(s, d):
    nl = s.num_levels()
    for l in range(nl):
        kind, n, nnz, nc = s.level_info(l)
        pre = os.path.join(d, f"{l}_")
        A = petsc_io.read_mat(pre + "A.petsc")
        rp, ci, v = s.level_matrix(l)
        assert np.array_equal(A.indptr, rp) and np.array_equal(A.indices, ci) and np.array_equal(A.data, v)
        if kind == _ffi.LEVEL_COARSE:
            assert not os.path.exists(pre + "P.petsc")
            continue
        P = petsc_io.read_mat(pre + "P.petsc")
        R = petsc_io.read_mat(pre + "R.petsc")
        assert P.shape == (n, nc) and R.shape == (nc, n)
        assert (R != P.T).nnz == 0
        m = s.level_map(l)
        if kind == _ffi.LEVEL_ELIM:
            Q = petsc_io.read_mat(pre + "Q.petsc")
            elim = m < 0
            assert np.array_equal(np.diff(Q.indptr), elim.astype(int))
            kept = np.flatnonzero(~elim)
            Pk = P[kept]
            assert np.array_equal(Pk.indices, m[kept]) and np.all(Pk.data == 1.0)
            # eliminated rows interpolate from their (kept) neighbours with
            # (1 / a_ee) * -a_ej; Q holds 1 / a_ee
            L = sp.csr_matrix((v, ci, rp), shape=(n, n))
            diag = L.diagonal()
            for i in np.flatnonzero(elim)[:50]:
                row = L.getrow(i)
                nb = [(m[j], (1.0 / diag[i]) * -a) for j, a in zip(row.indices, row.data) if j != i]
                got = P.getrow(i)
                assert list(zip(got.indices, got.data)) == nb
                assert Q[i, i] == 1.0 / diag[i]
        else:
            assert np.array_equal(P.indices, m) and np.all(P.data == 1.0)
            S = petsc_io.read_mat(pre + "strength.petsc")
            assert np.array_equal(S.indptr, rp) and np.array_equal(S.indices, ci)
            assert np.array_equal(S.data, s.level_strength(l))
            ind, state = s.level_states(l)
            assert np.array_equal(petsc_io.read_vec(pre + "agg_ind.petsc"), ind)
            assert np.array_equal(petsc_io.read_vec(pre + "agg_state.petsc"), state)


def same_solver(a, b, rhs):
    assert a.num_levels() == b.num_levels()
    for l in range(a.num_levels()):
        assert a.level_info(l) == b.level_info(l)
        for x, y in zip(a.level_matrix(l), b.level_matrix(l)):
            assert np.array_equal(x, y)
        if a.level_info(l)[0] == _ffi.LEVEL_AGG:
            assert a.level_cheby(l) == b.level_cheby(l)
    assert np.array_equal(a.ordering(), b.ordering())
    ia, ib = a.solve(rhs), b.solve(rhs)
    assert ia.iters() == ib.iters()
    assert np.array_equal(ia.residuals, ib.residuals)
    assert np.array_equal(ia.solution, ib.solution)
    return ia


CASES = {
    "rmat12_pend": (lambda: graphs.rmat_config(12, "lognormal", pendants=True), {}),
    "rmat18_layout_mixed": (lambda: graphs.rmat_config(18, "lognormal", pendants=True),
                            {"precision": _ffi.PREC_MIXED, "outer_ksp_type": _ffi.KSP_FCG}),
    "grid_kcycle": (lambda: graphs.grid2d(60, 50, wkind=_ffi.W_UNIFORM, p0=0.1, p1=10.0, seed=2),
                    {"kcycle": 1, "outer_ksp_type": _ffi.KSP_FCG}),
    "ba_degree_order": (lambda: graphs.barabasi_albert(5000, 4, seed=3), {"ordering": _ffi.ORDER_DEGREE}),
}


@pytest.mark.parametrize("case", list(CASES))
def test_dump_files_and_reload(case, tmp_path):
    make, kw = CASES[case]
    g = make()
    print(f"case={case} n={g.n}", flush=True)
    s = ligmg_amd.Solver(g, ligmg_amd.Options(**kw))
    d = tmp_path / "h"
    s.dump_hierarchy(d)
    check_files(s, str(d))
    t = ligmg_amd.Solver.load(d, ligmg_amd.Options(**kw))
    info = same_solver(s, t, graphs.rhs(g.n, 1, 4))
    assert info.rel_residual() <= 1e-8
    if case == "rmat18_layout_mixed":  # levels stored in the solve layout: elimination and aggregation
        assert s.level_info(0)[0] == _ffi.LEVEL_ELIM and s.level_info(1)[0] == _ffi.LEVEL_AGG
        assert s.level_info(1)[1] >= 1 << 16


def test_reload_with_other_solve_options(tmp_path):
    """The dump fixes the hierarchy; solve options may change on load."""
    g = graphs.rmat_config(13, "uniform")
    s = ligmg_amd.Solver(g)
    s.dump_hierarchy(tmp_path / "h")
    kw = dict(kcycle=1, outer_ksp_type=_ffi.KSP_FCG, precision=_ffi.PREC_MIXED)
    fresh = ligmg_amd.Solver(g, ligmg_amd.Options(**kw))
    loaded = ligmg_amd.Solver.load(tmp_path / "h", ligmg_amd.Options(**kw))
    same_solver(fresh, loaded, graphs.rhs(g.n, 1, 1))


def test_blocks_survive_the_dump(tmp_path):
    gs = [graphs.barabasi_albert(n, 4, seed=s) for s, n in enumerate([3000, 800, 5000])]
    u = graphs.union(gs)
    s = ligmg_amd.Solver(u, comp_ptr=u.comp_ptr)
    s.dump_hierarchy(tmp_path / "h")
    assert np.array_equal(petsc_io.read_vec(str(tmp_path / "h" / "blocks.petsc")), u.comp_ptr)
    t = ligmg_amd.Solver.load(tmp_path / "h")
    same_solver(s, t, graphs.rhs(u.n, 1, 5, comp_ptr=u.comp_ptr))


def test_load_rejects_bad_dumps(tmp_path):
    with pytest.raises(ligmg_amd.LigmgError) as e:
        ligmg_amd.Solver.load(tmp_path / "missing")
    assert e.value.code == _ffi.ERR_IO
    g = graphs.grid2d(60, 60)
    s = ligmg_amd.Solver(g)
    assert s.num_levels() >= 3
    d = tmp_path / "h"
    s.dump_hierarchy(d)
    raw = (d / "1_A.petsc").read_bytes()
    (d / "1_A.petsc").write_bytes(raw[:-16])  # truncated
    with pytest.raises(ligmg_amd.LigmgError) as e:
        ligmg_amd.Solver.load(d)
    assert e.value.code == _ffi.ERR_IO
    s.dump_hierarchy(d)
    A = petsc_io.read_mat(str(d / "2_A.petsc"))
    petsc_io.write_mat(str(d / "2_A.petsc"), A[:-1, :-1])  # does not match the level above
    with pytest.raises(ligmg_amd.LigmgError) as e:
        ligmg_amd.Solver.load(d)
    assert e.value.code == _ffi.ERR_IO


def test_csv_line(tmp_path):
    g = graphs.grid2d(40, 40)
    o = ligmg_amd.Options(kcycle=1, outer_ksp_type=_ffi.KSP_FCG, pre_smooth=3)
    s = ligmg_amd.Solver(g, o)
    info = s.solve(graphs.rhs(g.n, 1, 0))
    text = s.dump_csv(info, graph="grid40")
    head, line = text.strip().split("\n")
    cols, vals = head.split(","), line.split(",")
    assert len(cols) == len(vals) == 46
    row = dict(zip(cols, vals))
    assert row["StrengthMetric"] == "affinity" and row["Smoother"] == "chebyshev"
    assert row["KCycle"] == "1" and row["OuterKSPType"] == "fcg" and row["KCycleKSPType"] == "fcg"
    assert row["PreSmoothing"] == "3" and row["Tolerance"] == "1e-08" and row["AffScaling"] == "Max"
    assert row["graph"] == "grid40" and row["solver"] == "ligmg" and row["nodes"] == "1"
    assert int(row["iters"]) == info.iters() and int(row["nnz"]) == s.level_info(0)[2]
    assert float(row["rel_error"]) == pytest.approx(info.rel_residual(), rel=1e-5)
    assert float(row["cycle_complexity"]) == pytest.approx(info.cycle_complexity(), rel=1e-5)
