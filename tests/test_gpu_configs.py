"""The five benchmark workloads at their full sizes, on the GPU.

Where the oracle finishes in seconds (the 256x256 grid) the HIP path is
compared with it directly.  The large workloads are checked through
properties that do not depend on size:
  - the residual ||b - L x|| / ||b|| recomputed on the host with scipy from
    the returned solution (an SpMV independent of the library);
  - agreement of two different methods (FP64 V-cycle/CG against the FP32
    preconditioner inside FP64 flexible CG, or against the K-cycle) within
    1e-6 after mean removal: both converge to the minimum-norm solution;
  - per-(graph, column) convergence of batched solves;
  - directional finite differences of the autograd gradient.
"""
import numpy as np
import pytest

import oracle_ffi as O
import ligmg_amd
from ligmg_amd import _ffi, graphs

pytestmark = pytest.mark.gpu


def centered_rel(x, y):
    x = x - x.mean()
    y = y - y.mean()
    return float(np.linalg.norm(x - y) / np.linalg.norm(y))


def host_residual(g, b, x):
    A = g.to_scipy()
    return float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))


def test_grid256_vcycle_matches_oracle():
    """2-D 5-point grid 256x256, unit weights, Chebyshev degree 2, V-cycle CG."""
    g = graphs.grid2d_config(256)
    assert g.n == 65536 and (g.nnz - g.n) // 2 == 130560
    b = graphs.rhs(g.n, 1, 0)
    d = ligmg_amd.Solver(g)
    r = O.OracleSolver(g)
    assert d.num_levels() == r.num_levels()
    for l in range(d.num_levels()):
        assert d.level_info(l) == r.level_info(l), f"level {l}"
    info = d.solve(b)
    xr, rinfo, rhist = r.solve(b)
    print(f"grid256 levels={d.num_levels()} iters={info.iters()} oracle={rinfo.iters} "
          f"rel={info.rel_residual():.2e} setup={d.setup_time:.3f}s solve={info.time():.4f}s")
    assert info.status == _ffi.CONVERGED and info.rel_residual() <= 1e-8
    # same method: the residual histories coincide early on; rounding in
    # the (differently ordered) inner products moves the last iterations
    h = np.asarray(info.residuals)
    assert np.allclose(h[:10], rhist[:10], rtol=1e-6)
    assert abs(info.iters() - rinfo.iters) <= 2
    assert centered_rel(info.solution, xr) <= 1e-6
    assert host_residual(g, b, info.solution) <= 1e-7


def test_rmat22_kcycle():
    """R-MAT scale 22, uniform[0.5,1.5] weights, LCC, K-cycle to 1e-8."""
    g = graphs.rmat_config(22, "uniform")
    print(f"rmat22 n={g.n} nnz={g.nnz}", flush=True)
    b = graphs.rhs(g.n, 1, 0)
    k = ligmg_amd.Solver(g, ligmg_amd.Options(kcycle=1, outer_ksp_type=_ffi.KSP_FCG))
    ik = k.solve(b)
    print(f"kcycle setup={k.setup_time:.2f}s solve={ik.time():.3f}s iters={ik.iters()} "
          f"rel={ik.rel_residual():.2e} full={ik.full_rel_residual:.2e}", flush=True)
    assert ik.status == _ffi.CONVERGED and ik.rel_residual() <= 1e-8
    assert host_residual(g, b, ik.solution) <= 1e-7
    xk = ik.solution
    k.close()
    v = ligmg_amd.Solver(g)
    iv = v.solve(b)
    print(f"vcycle iters={iv.iters()} solve={iv.time():.3f}s", flush=True)
    assert iv.rel_residual() <= 1e-8
    assert centered_rel(xk, iv.solution) <= 1e-6
    # same hierarchy for the same input and seed: bitwise reproducible
    assert np.array_equal(v.solve(b).solution, iv.solution)


def test_rmat23_mixed_against_fp64():
    """R-MAT scale 23 with pendant paths and lognormal(0,1.5) weights: the
    FP32-preconditioned solve and the FP64 solve both reach 1e-8 and agree
    within 1e-6."""
    g = graphs.rmat_config(23, "lognormal", pendants=True)
    print(f"rmat23 n={g.n} nnz={g.nnz}", flush=True)
    b = graphs.rhs(g.n, 1, 0)
    s64 = ligmg_amd.Solver(g)
    i64 = s64.solve(b)
    x64 = i64.solution
    print(f"fp64 setup={s64.setup_time:.2f}s solve={i64.time():.3f}s iters={i64.iters()}", flush=True)
    s64.close()
    s32 = ligmg_amd.Solver(g, ligmg_amd.Options(precision=_ffi.PREC_MIXED, outer_ksp_type=_ffi.KSP_FCG))
    i32 = s32.solve(b)
    print(f"mixed setup={s32.setup_time:.2f}s solve={i32.time():.3f}s iters={i32.iters()}", flush=True)
    assert i64.rel_residual() <= 1e-8 and i32.rel_residual() <= 1e-8
    assert centered_rel(i32.solution, x64) <= 1e-6
    r64, r32 = host_residual(g, b, x64), host_residual(g, b, i32.solution)
    print(f"host residuals fp64={r64:.2e} mixed={r32:.2e}")
    assert r64 <= 1e-7 and r32 <= 1e-7


def ba48():
    ns = np.random.default_rng(0).integers(20000, 400001, 48)
    return graphs.union([graphs.barabasi_albert(int(n), 5, seed=i) for i, n in enumerate(ns)])


def test_batched_48_ba_graphs_16_rhs():
    """48 independent Barabasi-Albert graphs (m=5, n in [20k, 400k], seeds
    0..47), 16 mean-zero right-hand sides each, one hierarchy; every
    (graph, column) pair converges on its own."""
    u = ba48()
    k = 16
    print(f"ba48 n={u.n} nnz={u.nnz}", flush=True)
    B = graphs.rhs(u.n, k, 0, comp_ptr=u.comp_ptr).reshape(u.n, k, order="F")
    s = ligmg_amd.Solver(u, ligmg_amd.Options(precision=_ffi.PREC_MIXED, outer_ksp_type=_ffi.KSP_FCG),
                         comp_ptr=u.comp_ptr)
    X, info, its = s.solve_batch(B)
    print(f"setup={s.setup_time:.2f}s batch solve={info.time:.3f}s max iters={int(its.max())}",
          flush=True)
    assert info.status == _ffi.CONVERGED
    A = u.to_scipy()
    R = B - A @ X
    cp = u.comp_ptr
    worst = 0.0
    for blk in range(48):
        lo, hi = cp[blk], cp[blk + 1]
        nr = np.linalg.norm(R[lo:hi], axis=0) / np.linalg.norm(B[lo:hi], axis=0)
        worst = max(worst, float(nr.max()))
    print(f"worst per-(graph, column) residual {worst:.2e}")
    assert worst <= 1e-7
    one = s.solve(np.ascontiguousarray(B[:, 5]))
    for blk in (0, 17, 47):
        lo, hi = cp[blk], cp[blk + 1]
        assert centered_rel(X[lo:hi, 5], one.solution[lo:hi]) <= 1e-6


def test_batched_grid3d_160_64_rhs():
    """3-D 7-point grid 160^3, weights uniform[0.1, 10], 64 right-hand sides."""
    g = graphs.grid3d(160, 160, 160, wkind=_ffi.W_UNIFORM, p0=0.1, p1=10.0, seed=0)
    k = 64
    B = graphs.rhs(g.n, k, 1).reshape(g.n, k, order="F")
    s = ligmg_amd.Solver(g, ligmg_amd.Options(precision=_ffi.PREC_MIXED, outer_ksp_type=_ffi.KSP_FCG))
    X, info, its = s.solve_batch(B)
    print(f"grid3d160 setup={s.setup_time:.2f}s batch solve={info.time:.3f}s "
          f"per rhs={info.time / k * 1e3:.2f}ms iters={int(its.min())}..{int(its.max())}", flush=True)
    assert info.status == _ffi.CONVERGED
    R = B - g.to_scipy() @ X
    nr = np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)
    assert nr.max() <= 1e-7
    one = s.solve(np.ascontiguousarray(B[:, 63]))
    assert centered_rel(X[:, 63], one.solution) <= 1e-6


def test_autograd_rgg_1m():
    """Random geometric graph, 1,000,000 vertices, mean degree 12, LCC,
    w = exp(-d^2/r^2) in FP32; b = +1/-1 at 256 source/sink pairs; loss =
    sum over 10k vertex pairs of (x_i - x_j)^2.  The directional derivatives
    of the loss along the gradients (w and b) match central finite
    differences through the same solver."""
    import torch
    from ligmg_amd.autograd import EdgePattern, laplacian_solve
    from test_gradients import rgg_instance
    g, ei, ej, w, b, pi, pj = rgg_instance(1000000, seed=0)
    print(f"rgg n={g.n} edges={len(w)}", flush=True)
    dev = torch.device("cuda:0")
    pat = EdgePattern(g.n, ei, ej, dev)
    tpi, tpj = torch.tensor(pi, device=dev), torch.tensor(pj, device=dev)
    wt = torch.tensor(w, device=dev, requires_grad=True)
    bt = torch.tensor(b, device=dev, requires_grad=True)

    def loss_of(wv, bv):
        x = laplacian_solve(wv, bv, pat)
        return ((x[tpi] - x[tpj]) ** 2).sum()

    loss = loss_of(wt, bt)
    loss.backward()
    gw = wt.grad.detach().to(torch.float64)
    gb = bt.grad.detach()
    # directions along the gradient (scaled so no weight moves by more than
    # 0.1%): the signal is sum w g^2 / max|g|, far above the solver's noise
    w64 = torch.tensor(w, device=dev, dtype=torch.float64)
    b64 = torch.tensor(b, device=dev)
    dw = w64 * gw / gw.abs().max()
    eps = 1e-3
    with torch.no_grad():
        lp = loss_of(w64 + eps * dw, b64).item()
        lm = loss_of(w64 - eps * dw, b64).item()
    fd = (lp - lm) / (2 * eps)
    an = float((gw * dw).sum())
    print(f"loss={loss.item():.6g} directional: autograd={an:.6g} fd={fd:.6g}")
    assert an == pytest.approx(fd, rel=2e-3)
    db = gb - gb.mean()
    epsb = 1e-2 * float(b64.norm() / db.norm())
    with torch.no_grad():
        lp = loss_of(w64, b64 + epsb * db).item()
        lm = loss_of(w64, b64 - epsb * db).item()
    fdb = (lp - lm) / (2 * epsb)
    anb = float((gb * db).sum())
    print(f"grad b directional: autograd={anb:.6g} fd={fdb:.6g}")
    assert anb == pytest.approx(fdb, rel=1e-4)
