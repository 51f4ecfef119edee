"""Python facade over the C ABI, mirroring the reference's C++ API.

    solver = ligmg_amd.Solver(A, opts)      # ligmg::Solver(A, opts), ligmg.hpp:17-28
    info = solver.solve(b)                  # Solver::solve(b), ligmg.hpp:31-33
    x = info.solution                       # ConvergenceInfo (util.hpp:143-170)

``A`` is a graph Laplacian in CSR form: a :class:`ligmg_amd.graphs.Graph`, a
scipy sparse matrix, or a ``(rowptr, colind, vals)`` triple of numpy arrays
or torch tensors (host or ``cuda`` device memory; device tensors are read in
place).  All compute runs in the HIP library; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import _ffi

_OPT_FIELDS = [f for f, _ in _ffi.ligmg_options._fields_]


class LigmgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ligmg error {code}: {msg}")
        self.code = code


def _check(rc, what):
    if rc != 0:
        msg = _ffi.solver_lib().ligmg_last_error()
        raise LigmgError(rc, f"{what}: {msg.decode() if msg else ''}")


class Options:
    """ligmg::Options (options.hpp:80-134): same field names and defaults.

    ``Options(kcycle=1, outer_ksp_type=KSP_FCG)`` overrides fields."""

    def __init__(self, **kw):
        object.__setattr__(self, "_c", _ffi.ligmg_options())
        _ffi.solver_lib().ligmg_options_default(C.byref(self._c))
        for k, v in kw.items():
            setattr(self, k, v)

    def __getattr__(self, k):
        if k in _OPT_FIELDS:
            return getattr(self._c, k)
        raise AttributeError(k)

    def __setattr__(self, k, v):
        if k not in _OPT_FIELDS:
            raise AttributeError(f"unknown option {k}")
        setattr(self._c, k, v)

    def as_dict(self):
        return {k: getattr(self._c, k) for k in _OPT_FIELDS}


def default_options():
    """default_options() (options.hpp:136)."""
    return Options()


class ConvergenceInfo:
    """ConvergenceInfo (util.hpp:143-170) plus the full-system residual the
    reference only prints (multigrid.hpp:357-366), status and setup time."""

    def __init__(self, residuals, info, solution):
        self.residuals = residuals
        self._iters = info.iters
        self._time = info.time
        self._work = info.work
        self.solution = solution
        self.status = info.status
        self.base_level = info.base_level
        self.full_rel_residual = info.full_rel_residual
        self.setup_time = info.setup_time
        self._info = info

    def iters(self):
        return self._iters

    def time(self):
        return self._time

    def work(self):
        return self._work

    def rel_residual(self):
        return self._info.rel_residual

    def cycle_complexity(self):
        return self._info.cycle_complexity

    def cf(self):
        return self._info.cf

    def ecf(self):
        return self._info.ecf

    def wda(self):
        return self._info.wda

    def __str__(self):
        # util.hpp:172-178
        return ("iters    error  comp  conv econv   wda     time\n"
                f"{self.iters():5d} {self.rel_residual():.2e} {self.cycle_complexity():5.2f} "
                f"{self.cf():.3f} {self.ecf():.3f} {self.wda():5.2f} {self.time():.2e}")


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _csr_args(A):
    """Returns (n, rowptr, colind, vals, mem_kind, keepalive)."""
    if hasattr(A, "rowptr") and hasattr(A, "colind"):
        rp, ci, v, n = A.rowptr, A.colind, A.vals, A.n
    elif isinstance(A, tuple) and len(A) == 3:
        rp, ci, v = A
        n = len(rp) - 1
    else:  # scipy sparse
        A = A.tocsr()
        A.sort_indices()
        rp, ci, v, n = A.indptr, A.indices, A.data, A.shape[0]
    if _is_torch(rp):
        import torch
        rp = rp.to(torch.int32).contiguous()
        ci = ci.to(torch.int32).contiguous()
        v = v.to(torch.float64).contiguous()
        mem = _ffi.MEM_DEVICE if rp.is_cuda else _ffi.MEM_HOST
        if rp.is_cuda:
            # the library runs on its own stream: torch's pending writes first
            torch.cuda.current_stream().synchronize()
        return n, rp.data_ptr(), ci.data_ptr(), v.data_ptr(), mem, (rp, ci, v)
    rp = np.ascontiguousarray(rp, dtype=np.int32)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.float64)
    return n, rp.ctypes.data, ci.ctypes.data, v.ctypes.data, _ffi.MEM_HOST, (rp, ci, v)


class Solver:
    """ligmg::Solver<T,F,DER> (ligmg.hpp:12-60).  The constructor runs the
    whole setup on the GPU."""

    def __init__(self, A, opts=None, comp_ptr=None):
        self._lib = _ffi.solver_lib()
        self.opts = opts if opts is not None else Options()
        n, rp, ci, v, mem, keep = _csr_args(A)
        self.n = n
        cp = None
        if comp_ptr is None and hasattr(A, "comp_ptr") and len(A.comp_ptr) > 2:
            comp_ptr = A.comp_ptr
        if comp_ptr is not None:
            cp = np.ascontiguousarray(comp_ptr, dtype=np.int64)
        self.comp_ptr = cp
        h = _ffi.HANDLE()
        rc = self._lib.ligmg_create(
            n, rp, ci, v, mem, C.byref(self.opts._c), 0 if cp is None else len(cp) - 1,
            None if cp is None else cp.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(h))
        _check(rc, "ligmg_create")
        self._h = h
        t = C.c_double()
        self._lib.ligmg_setup_time(self._h, C.byref(t))
        self.setup_time = t.value

    @classmethod
    def load(cls, path, opts=None):
        """A solver from a hierarchy written by dump_hierarchy (new): the
        levels are read back, the solve-phase set-up runs for ``opts``."""
        self = cls.__new__(cls)
        self._lib = _ffi.solver_lib()
        self.opts = opts if opts is not None else Options()
        h = _ffi.HANDLE()
        _check(self._lib.ligmg_load_hierarchy(os.fsencode(str(path)), C.byref(self.opts._c), C.byref(h)),
               "ligmg_load_hierarchy")
        self._h = h
        self.comp_ptr = None
        self.n = self.level_info(0)[1]
        t = C.c_double()
        self._lib.ligmg_setup_time(self._h, C.byref(t))
        self.setup_time = t.value
        return self

    def dump_hierarchy(self, path):
        """Solver::dump_hierarchy (ligmg.hpp:47-52): one directory of PETSc
        binary files, <level>_A.petsc, _P, _R, _Q / _strength ..."""
        _check(self._lib.ligmg_dump_hierarchy(self._h, os.fsencode(str(path))), "ligmg_dump_hierarchy")

    def dump_csv(self, info, graph=""):
        """Solver::dump_csv (ligmg.hpp:42-45): the reference's CSV header
        and result line for the solve ``info`` (a ConvergenceInfo)."""
        buf = C.create_string_buffer(1 << 14)
        _check(self._lib.ligmg_dump_csv(self._h, graph.encode(), C.byref(info._info), buf, len(buf)),
               "ligmg_dump_csv")
        return buf.value.decode()

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ligmg_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # --------------------------------------------------------------- solve
    def solve(self, b, history_cap=1024, level=0):
        """Solve L x = b to opts.tol relative residual.  ``b`` is a numpy
        array (host) or a torch tensor (a cuda one stays on the device); the
        solution has the same kind and shape.  ``level`` > 0 solves the
        system of that hierarchy level (Solver::solve(b, from),
        ligmg.hpp:36-38): b then has the level's size and is numbered as
        level_matrix(level) reports it."""
        info = _ffi.ligmg_info()
        hist = np.zeros(history_cap, dtype=np.float64)
        n = self.n if level == 0 else self.level_info(level)[1]
        if _is_torch(b):
            import torch
            if tuple(b.shape) not in ((n,), (n, 1)):
                raise ValueError(f"b must have shape ({n},) or ({n}, 1), not {tuple(b.shape)}")
            bb = b.to(torch.float64).contiguous()
            x = torch.empty_like(bb)
            mem = _ffi.MEM_DEVICE if bb.is_cuda else _ffi.MEM_HOST
            if bb.is_cuda:
                torch.cuda.current_stream().synchronize()
            rc = self._lib.ligmg_solve_from(self._h, level, bb.data_ptr(), x.data_ptr(), mem, C.byref(info),
                                            hist.ctypes.data_as(C.POINTER(C.c_double)), history_cap)
        else:
            bb = np.ascontiguousarray(b, dtype=np.float64)
            if bb.shape != (n,):
                raise ValueError(f"b must have shape ({n},)")
            x = np.empty_like(bb)
            rc = self._lib.ligmg_solve_from(self._h, level, bb.ctypes.data, x.ctypes.data, _ffi.MEM_HOST,
                                            C.byref(info), hist.ctypes.data_as(C.POINTER(C.c_double)),
                                            history_cap)
        _check(rc, "ligmg_solve_from")
        return ConvergenceInfo(hist[:info.history_len].copy(), info, x)

    def solve_batch(self, B):
        """Solve for the k columns of B (n x k, k <= 64) in one hierarchy
        traversal per cycle.  Returns (X, info, iters_per_column); every
        (graph block, column) pair has its own Krylov coefficients."""
        info = _ffi.ligmg_info()
        if _is_torch(B):
            import torch
            if B.dim() != 2 or B.shape[0] != self.n or not 1 <= B.shape[1] <= 64:
                raise ValueError(f"B must have shape ({self.n}, k) with 1 <= k <= 64, not {tuple(B.shape)}")
            k = B.shape[1]
            Bc = B.to(torch.float64).t().contiguous()  # column-major storage
            Xc = torch.empty_like(Bc)
            mem = _ffi.MEM_DEVICE if Bc.is_cuda else _ffi.MEM_HOST
            if Bc.is_cuda:
                torch.cuda.current_stream().synchronize()
            its = np.zeros(k, dtype=np.int32)
            rc = self._lib.ligmg_solve_batch(self._h, k, Bc.data_ptr(), Xc.data_ptr(), mem,
                                             C.byref(info), its.ctypes.data_as(C.POINTER(C.c_int32)))
            X = Xc.t()
        else:
            Bc = np.asfortranarray(B, dtype=np.float64)
            if Bc.ndim != 2 or Bc.shape[0] != self.n or not 1 <= Bc.shape[1] <= 64:
                raise ValueError(f"B must have shape ({self.n}, k) with 1 <= k <= 64")
            k = Bc.shape[1]
            X = np.empty_like(Bc, order="F")
            its = np.zeros(k, dtype=np.int32)
            rc = self._lib.ligmg_solve_batch(self._h, k, Bc.ctypes.data, X.ctypes.data, _ffi.MEM_HOST,
                                             C.byref(info), its.ctypes.data_as(C.POINTER(C.c_int32)))
        _check(rc, "ligmg_solve_batch")
        return X, info, its

    # -------------------------------------------------------- introspection
    def ordering(self):
        """Internal vertex order: entry i is the caller's index of internal
        vertex i (identity under ORDER_NATURAL).  Level introspection
        reports matrices and maps in this order."""
        p = np.empty(self.n, np.int32)
        _check(self._lib.ligmg_ordering(self._h, p.ctypes.data), "ligmg_ordering")
        return p

    def num_levels(self):
        n = C.c_int32()
        _check(self._lib.ligmg_num_levels(self._h, C.byref(n)), "ligmg_num_levels")
        return n.value

    def tail_start(self):
        """First level of the fused coarse tail (num_levels() if none)."""
        n = C.c_int32()
        _check(self._lib.ligmg_tail_start(self._h, C.byref(n)), "ligmg_tail_start")
        return n.value

    def level_info(self, l):
        k, n, nnz, nc = C.c_int32(), C.c_int64(), C.c_int64(), C.c_int64()
        _check(self._lib.ligmg_level_info(self._h, l, C.byref(k), C.byref(n), C.byref(nnz),
                                          C.byref(nc)), "ligmg_level_info")
        return k.value, n.value, nnz.value, nc.value

    def level_matrix(self, l):
        k, n, nnz, nc = self.level_info(l)
        rp = np.empty(n + 1, np.int32)
        ci = np.empty(nnz, np.int32)
        v = np.empty(nnz, np.float64)
        P = C.POINTER
        _check(self._lib.ligmg_level_matrix(self._h, l, rp.ctypes.data_as(P(C.c_int32)),
                                            ci.ctypes.data_as(P(C.c_int32)),
                                            v.ctypes.data_as(P(C.c_double))), "ligmg_level_matrix")
        return rp, ci, v

    def level_map(self, l):
        k, n, nnz, nc = self.level_info(l)
        m = np.empty(n, np.int32)
        _check(self._lib.ligmg_level_map(self._h, l, m.ctypes.data_as(C.POINTER(C.c_int32))),
               "ligmg_level_map")
        return m

    def level_states(self, l):
        k, n, nnz, nc = self.level_info(l)
        ind = np.empty(n, np.int32)
        st = np.empty(n, np.int32)
        P = C.POINTER
        _check(self._lib.ligmg_level_states(self._h, l, ind.ctypes.data_as(P(C.c_int32)),
                                            st.ctypes.data_as(P(C.c_int32))), "ligmg_level_states")
        return ind, st

    def level_strength(self, l):
        k, n, nnz, nc = self.level_info(l)
        s = np.empty(nnz, np.float64)
        _check(self._lib.ligmg_level_strength(self._h, l, s.ctypes.data_as(C.POINTER(C.c_double))),
               "ligmg_level_strength")
        return s

    def level_cheby(self, l):
        lo, hi = C.c_double(), C.c_double()
        _check(self._lib.ligmg_level_cheby(self._h, l, C.byref(lo), C.byref(hi)), "ligmg_level_cheby")
        return lo.value, hi.value

    def level_spmv(self, l, x, precision=_ffi.PREC_FP64):
        """y = A_l x with the solve-phase kernel of level l (host arrays)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        P = C.POINTER
        _check(self._lib.ligmg_level_spmv(self._h, l, precision, x.ctypes.data_as(P(C.c_double)),
                                          y.ctypes.data_as(P(C.c_double))), "ligmg_level_spmv")
        return y

    def apply_cycle(self, l, r, precision=_ffi.PREC_FP64):
        """z = M r: one preconditioner cycle from level l (host arrays)."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty_like(r)
        P = C.POINTER
        _check(self._lib.ligmg_apply_cycle(self._h, l, precision, r.ctypes.data_as(P(C.c_double)),
                                           z.ctypes.data_as(P(C.c_double))), "ligmg_apply_cycle")
        return z

    def print(self):
        """print_hierarchy (multigrid.hpp:265-292)."""
        print(f"{'Level':>5}  {'Size':>12}  {'NNZ':>14}  {'Type':>6}")
        for l in range(self.num_levels()):
            k, n, nnz, nc = self.level_info(l)
            print(f"{l:5d}  {n:12,d}  {nnz:14,d}  {_ffi.LEVEL_NAMES[k]:>6}")


# README.md:64 spells the type ligmg::ligmg
ligmg = Solver


def aggregate_with_strength(g, S, opts=None):
    """Voting aggregation on the GPU driven by given strength values on the
    pattern of g (aggregation.hpp:838-964).  Returns (agg, ind, state, n_c)."""
    lib = _ffi.solver_lib()
    o = opts if opts is not None else Options()
    rp = np.ascontiguousarray(g.rowptr, np.int32)
    ci = np.ascontiguousarray(g.colind, np.int32)
    S = np.ascontiguousarray(S, np.float64)
    agg = np.empty(g.n, np.int32)
    ind = np.empty(g.n, np.int32)
    st = np.empty(g.n, np.int32)
    nc = C.c_int64()
    P = C.POINTER
    rc = lib.ligmg_aggregate_with_strength(
        g.n, rp.ctypes.data_as(P(C.c_int32)), ci.ctypes.data_as(P(C.c_int32)),
        S.ctypes.data_as(P(C.c_double)), C.byref(o._c), agg.ctypes.data_as(P(C.c_int32)),
        ind.ctypes.data_as(P(C.c_int32)), st.ctypes.data_as(P(C.c_int32)), C.byref(nc))
    _check(rc, "ligmg_aggregate_with_strength")
    return agg, ind, st, nc.value


def trim_memory():
    """Return the device memory the library caches for reuse (temporaries of
    setup, buffers of closed solvers) to HIP."""
    _check(_ffi.solver_lib().ligmg_trim_memory(), "ligmg_trim_memory")
