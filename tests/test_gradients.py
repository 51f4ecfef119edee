"""Gradients of a loss of x = L(w)^+ b.

CPU part: the closed form dloss/dw_e = -(lam_i - lam_j)(x_i - x_j),
lam = L^+ dloss/dx, against central finite differences with a dense FP64
pseudo-inverse (pins the formula the HIP kernel implements).
GPU part: the torch.autograd.Function through the C ABI against that dense
FP64 reference and against finite differences, on the configuration's
2,000-vertex random geometric graph."""
import math

import numpy as np
import pytest

from ligmg_amd import graphs


def rgg_instance(n=2000, seed=0):
    """Random geometric graph in the unit square, radius for mean degree 12,
    LCC, w = exp(-d^2/r^2) (FP32); b = +1/-1 at 256 source/sink pairs (fewer
    when the graph is small); loss = sum over random vertex pairs of
    (x_i - x_j)^2."""
    r = math.sqrt(12 / (math.pi * n))
    g = graphs.rgg(n, r, seed=seed)
    ei, ej, w = g.upper_edges()
    rng = np.random.default_rng(seed)
    npairs = min(256, g.n // 4)
    perm = rng.permutation(g.n)
    b = np.zeros(g.n)
    b[perm[:npairs]] = 1.0
    b[perm[npairs:2 * npairs]] = -1.0
    pi = rng.integers(0, g.n, 10000)
    pj = rng.integers(0, g.n, 10000)
    return g, ei, ej, w.astype(np.float32), b, pi, pj


def dense_L(n, ei, ej, w):
    L = np.zeros((n, n))
    np.add.at(L, (ei, ej), -w)
    np.add.at(L, (ej, ei), -w)
    L[np.arange(n), np.arange(n)] = -L.sum(1)
    return L


def loss_and_grads(n, ei, ej, w, b, pi, pj):
    Lp = np.linalg.pinv(dense_L(n, ei, ej, w.astype(np.float64)))
    x = Lp @ (b - b.mean())
    d = x[pi] - x[pj]
    loss = np.sum(d * d)
    gx = np.zeros(n)
    np.add.at(gx, pi, 2 * d)
    np.add.at(gx, pj, -2 * d)
    lam = Lp @ (gx - gx.mean())
    gw = -(lam[ei] - lam[ej]) * (x[ei] - x[ej])
    return loss, x, gw, lam


def test_closed_form_gradient_matches_finite_differences():
    g, ei, ej, w, b, pi, pj = rgg_instance(300, seed=1)
    w = w.astype(np.float64)
    loss, x, gw, lam = loss_and_grads(g.n, ei, ej, w, b, pi, pj)
    rng = np.random.default_rng(0)
    for e in rng.choice(len(w), 12, replace=False):
        h = 1e-5 * w[e]
        wp, wm = w.copy(), w.copy()
        wp[e] += h
        wm[e] -= h
        fd = (loss_and_grads(g.n, ei, ej, wp, b, pi, pj)[0] -
              loss_and_grads(g.n, ei, ej, wm, b, pi, pj)[0]) / (2 * h)
        assert fd == pytest.approx(gw[e], rel=1e-5, abs=1e-8 * abs(loss))
    # gradient w.r.t. b is lam (projected): check one coordinate
    for i in rng.choice(g.n, 5, replace=False):
        h = 1e-4
        bp, bm = b.copy(), b.copy()
        bp[i] += h
        bm[i] -= h
        fd = (loss_and_grads(g.n, ei, ej, w, bp, pi, pj)[0] -
              loss_and_grads(g.n, ei, ej, w, bm, pi, pj)[0]) / (2 * h)
        assert fd == pytest.approx(lam[i], rel=1e-5, abs=1e-8 * abs(loss))


@pytest.mark.gpu
def test_autograd_matches_dense_reference_and_fd():
    import torch
    from ligmg_amd.autograd import EdgePattern, laplacian_solve
    g, ei, ej, w, b, pi, pj = rgg_instance(2000, seed=0)
    dev = torch.device("cuda:0")
    pat = EdgePattern(g.n, ei, ej, dev)
    wt = torch.tensor(w, device=dev, requires_grad=True)
    bt = torch.tensor(b, device=dev, requires_grad=True)
    x = laplacian_solve(wt, bt, pat)
    tpi, tpj = torch.tensor(pi, device=dev), torch.tensor(pj, device=dev)
    loss = ((x[tpi] - x[tpj]) ** 2).sum()
    loss.backward()
    rl, rx, rgw, rlam = loss_and_grads(g.n, pat.ei.cpu().numpy(), pat.ej.cpu().numpy(), w, b, pi, pj)
    xg = x.detach().cpu().numpy()
    print(f"n={g.n} edges={len(w)} loss={loss.item():.6g} ref={rl:.6g}")
    assert np.linalg.norm(xg - rx) <= 1e-6 * np.linalg.norm(rx)
    assert loss.item() == pytest.approx(rl, rel=1e-6)
    gw = wt.grad.detach().cpu().numpy().astype(np.float64)
    gb = bt.grad.detach().cpu().numpy()
    err_w = np.linalg.norm(gw - rgw) / np.linalg.norm(rgw)
    err_b = np.linalg.norm(gb - rlam) / np.linalg.norm(rlam)
    print(f"grad w rel err vs dense FP64: {err_w:.2e}; grad b: {err_b:.2e}")
    assert err_w <= 1e-5 and err_b <= 1e-6  # w grads are returned in FP32
    # central finite differences through the GPU solver on a few edges
    rng = np.random.default_rng(3)
    for e in rng.choice(len(w), 4, replace=False):
        h = 1e-3 * float(w[e])
        vals = []
        for sgn in (1, -1):
            w2 = torch.tensor(w, device=dev, dtype=torch.float64)
            w2[e] += sgn * h
            x2 = laplacian_solve(w2, torch.tensor(b, device=dev), pat, None)
            vals.append(((x2[tpi] - x2[tpj]) ** 2).sum().item())
        fd = (vals[0] - vals[1]) / (2 * h)
        assert fd == pytest.approx(rgw[e], rel=1e-3, abs=1e-6 * rl)


@pytest.mark.gpu
def test_autograd_pendant_graph_shift_dependent_loss():
    """A graph with a top elimination level (pendant paths on R-MAT): the
    solver's raw solution there is P x_c + Q b, not mean-zero.  The forward
    must still be x = L^+ b and the gradients those of x for a loss that a
    constant shift of x changes, sum(c * x) + sum(x^2)."""
    import torch
    import ligmg_amd
    from ligmg_amd import _ffi
    from ligmg_amd.autograd import EdgePattern, laplacian_solve
    g = graphs.rmat_config(10, "lognormal", pendants=True)
    ei, ej, w = g.upper_edges()
    s = ligmg_amd.Solver(g)
    assert s.level_info(0)[0] == _ffi.LEVEL_ELIM, "the case must have a top elimination level"
    rng = np.random.default_rng(5)
    b = rng.standard_normal(g.n)
    c = rng.standard_normal(g.n)
    dev = torch.device("cuda:0")
    pat = EdgePattern(g.n, ei, ej, dev)
    wt = torch.tensor(w, device=dev, dtype=torch.float64, requires_grad=True)
    bt = torch.tensor(b, device=dev, requires_grad=True)
    ct = torch.tensor(c, device=dev)
    x = laplacian_solve(wt, bt, pat)
    loss = (ct * x).sum() + (x * x).sum()
    loss.backward()
    # dense FP64 reference on the pattern's edge order
    pei, pej = pat.ei.cpu().numpy(), pat.ej.cpu().numpy()
    wmap = {(int(a), int(bq)): float(v) for a, bq, v in zip(ei, ej, w)}
    wp = np.array([wmap[(int(a), int(bq))] for a, bq in zip(pei, pej)])
    Lp = np.linalg.pinv(dense_L(g.n, pei, pej, wp))
    rx = Lp @ (b - b.mean())
    gx = c + 2 * rx
    lam = Lp @ (gx - gx.mean())
    rgw = -(lam[pei] - lam[pej]) * (rx[pei] - rx[pej])
    xg = x.detach().cpu().numpy()
    gw = wt.grad.detach().cpu().numpy()
    gb = bt.grad.detach().cpu().numpy()
    err_x = np.linalg.norm(xg - rx) / np.linalg.norm(rx)
    err_w = np.linalg.norm(gw - rgw) / np.linalg.norm(rgw)
    err_b = np.linalg.norm(gb - lam) / np.linalg.norm(lam)
    print(f"n={g.n} x err {err_x:.2e} grad w err {err_w:.2e} grad b err {err_b:.2e}")
    # a solve to 1e-8 relative residual; an unprojected solution would be
    # off by its mean shift, an O(1) error in x, grad b and the c . x term
    assert err_x <= 1e-6 and err_w <= 1e-5 and err_b <= 1e-6
    assert abs(xg.mean()) <= 1e-12 * np.abs(xg).max()
