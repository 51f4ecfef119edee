"""Differentiable Laplacian solves: x = L(w)^+ b as a torch.autograd.Function.

Forward assembles L(w) on the device from edge weights, builds the hierarchy
and solves.  Backward needs one adjoint solve, reusing the same hierarchy
because L is symmetric:

    lam = L^+ (dloss/dx)
    dloss/db   = lam
    dloss/dw_e = -(lam_i - lam_j) (x_i - x_j)      (HIP kernel, ligmg_edge_grad)

The last line follows from dL = sum_e dw_e (u_i - u_j)(u_i - u_j)^T and
d(L^+) b = -L^+ dL L^+ b for a connected graph (the nullspace, the
constants, does not move with w).  The graph must be connected; b and the
incoming gradient are projected to mean zero, and so are the solutions: the
solver returns L^+ b plus a multiple of the constant vector (with a top
elimination level, x = P x_c + Q b is not mean-zero), and L^+ b is the
mean-zero one, so the forward value and both gradients are exactly those of
x = L^+ b for any loss, shift-invariant or not.
"""
import ctypes as C

import numpy as np
import torch

from . import _ffi
from .solver import Options, Solver, _check


class EdgePattern:
    """Fixed undirected edge set (ei[e] < ej[e]) and its Laplacian CSR
    pattern on a device; eid maps each stored entry to its edge (-1 on the
    diagonal)."""

    def __init__(self, n, ei, ej, device="cuda"):
        ei = np.asarray(ei, dtype=np.int64)
        ej = np.asarray(ej, dtype=np.int64)
        if np.any(ei == ej):
            raise ValueError("self loops are not allowed")
        lo, hi = np.minimum(ei, ej), np.maximum(ei, ej)
        m = len(lo)
        rows = np.concatenate([lo, hi, np.arange(n)])
        cols = np.concatenate([hi, lo, np.arange(n)])
        eids = np.concatenate([np.arange(m), np.arange(m), np.full(n, -1)])
        order = np.lexsort((cols, rows))
        rows, cols, eids = rows[order], cols[order], eids[order]
        if np.any((rows[1:] == rows[:-1]) & (cols[1:] == cols[:-1])):
            raise ValueError("duplicate edges")
        rowptr = np.zeros(n + 1, dtype=np.int32)
        np.add.at(rowptr, rows + 1, 1)
        rowptr = np.cumsum(rowptr).astype(np.int32)
        self.n, self.m = n, m
        dev = torch.device(device)
        self.rowptr = torch.from_numpy(rowptr).to(dev)
        self.colind = torch.from_numpy(cols.astype(np.int32)).to(dev)
        self.eid = torch.from_numpy(eids.astype(np.int32)).to(dev)
        self.ei = torch.from_numpy(lo.astype(np.int32)).to(dev)
        self.ej = torch.from_numpy(hi.astype(np.int32)).to(dev)

    def assemble(self, w):
        """L(w) values on the pattern (FP64, device)."""
        w = w.detach().to(torch.float64).contiguous()
        vals = torch.empty(self.colind.shape[0], dtype=torch.float64, device=w.device)
        torch.cuda.current_stream().synchronize()
        _check(_ffi.solver_lib().ligmg_assemble_laplacian(
            self.n, self.rowptr.data_ptr(), self.colind.data_ptr(), self.eid.data_ptr(),
            w.data_ptr(), vals.data_ptr()), "ligmg_assemble_laplacian")
        return vals


class LaplacianSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w, b, pattern, opts):
        vals = pattern.assemble(w)
        solver = Solver((pattern.rowptr, pattern.colind, vals), opts)
        bb = b.detach().to(torch.float64)
        bb = bb - bb.mean()
        x = solver.solve(bb).solution
        x = x - x.mean()
        ctx.solver = solver
        ctx.pattern = pattern
        ctx.save_for_backward(x)
        ctx.dtypes = (w.dtype, b.dtype)
        return x.to(b.dtype)

    @staticmethod
    def backward(ctx, gx):
        (x,) = ctx.saved_tensors
        g = gx.detach().to(torch.float64)
        g = g - g.mean()
        lam = ctx.solver.solve(g).solution
        lam = lam - lam.mean()
        P = ctx.pattern
        gw = torch.empty(P.m, dtype=torch.float64, device=x.device)
        torch.cuda.current_stream().synchronize()
        _check(_ffi.solver_lib().ligmg_edge_grad(P.m, P.ei.data_ptr(), P.ej.data_ptr(), x.data_ptr(),
                                                 lam.data_ptr(), gw.data_ptr()), "ligmg_edge_grad")
        wdt, bdt = ctx.dtypes
        return gw.to(wdt), lam.to(bdt), None, None


def laplacian_solve(w, b, pattern, opts=None):
    """x = L(w)^+ b, differentiable in w (edge weights, ordered as the
    pattern's edges) and b."""
    return LaplacianSolve.apply(w, b, pattern, opts if opts is not None else Options())
