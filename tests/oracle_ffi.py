"""ctypes access to the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: it is only ever used here, as the checker.
"""
import ctypes as C
import os

import numpy as np

from ligmg_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")

i32, i64, u64, f64, P = C.c_int32, C.c_int64, C.c_uint64, C.c_double, C.POINTER
H = C.c_void_p
PROTOS = {
    "oracle_canon_sum": (f64, [P(f64), i64]),
    "oracle_seq_sum": (f64, [P(f64), i64]),
    "oracle_set_decision_order": (None, [i32]),
    "oracle_set_threads": (i32, [i32]),
    "oracle_default_options": (None, [P(_ffi.ligmg_options)]),
    "oracle_create": (C.c_int, [i64, P(i32), P(i32), P(f64), P(_ffi.ligmg_options), i32, P(i64),
                                P(H)]),
    "oracle_destroy": (C.c_int, [H]),
    "oracle_components": (C.c_int, [i64, C.c_void_p, C.c_void_p, C.c_void_p, P(i64)]),
    "oracle_largest_component": (C.c_int, [i64, C.c_void_p, C.c_void_p, C.c_void_p, P(i64), P(i64)]),
    "oracle_extract": (C.c_int, [i64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, i64, C.c_void_p,
                                 C.c_void_p, C.c_void_p]),
    "oracle_adjacency_to_laplacian": (C.c_int, [i64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                C.c_void_p, C.c_void_p]),
    "oracle_ordering": (C.c_int, [H, C.c_void_p]),
    "oracle_solve": (C.c_int, [H, P(f64), P(f64), P(_ffi.ligmg_info), P(f64), i32]),
    "oracle_solve_from": (C.c_int, [H, i32, P(f64), P(f64), P(_ffi.ligmg_info), P(f64), i32]),
    "oracle_num_levels": (C.c_int, [H, P(i32)]),
    "oracle_level_info": (C.c_int, [H, i32, P(i32), P(i64), P(i64), P(i64)]),
    "oracle_level_matrix": (C.c_int, [H, i32, P(i32), P(i32), P(f64)]),
    "oracle_level_map": (C.c_int, [H, i32, P(i32)]),
    "oracle_level_states": (C.c_int, [H, i32, P(i32), P(i32)]),
    "oracle_level_strength": (C.c_int, [H, i32, P(f64)]),
    "oracle_level_cheby": (C.c_int, [H, i32, P(f64), P(f64)]),
    "oracle_level_pinv": (C.c_int, [H, i32, P(f64)]),
    "oracle_apply_cycle": (C.c_int, [H, i32, P(f64), P(f64)]),
    "oracle_degrees": (C.c_int, [i64, P(i32), P(i32), P(f64), P(i32)]),
    "oracle_elim_select": (C.c_int, [i64, P(i32), P(i32), P(f64), i32, P(i32), P(i64)]),
    "oracle_tv_init": (C.c_int, [i64, u64, i32, i32, P(f64)]),
    "oracle_strength": (C.c_int, [i64, P(i32), P(i32), P(f64), P(f64), i32, i32, i32, i32, f64,
                                  P(f64), P(f64)]),
    "oracle_aggregate_with_strength": (C.c_int, [i64, P(i32), P(i32), P(f64),
                                                 P(_ffi.ligmg_options), P(i32), P(i32), P(i32),
                                                 P(i64)]),
    "oracle_cheby_apply": (C.c_int, [i64, P(i32), P(i32), P(f64), f64, f64, i32, P(f64), P(f64)]),
    "oracle_eig_estimate": (C.c_int, [i64, P(i32), P(i32), P(f64), u64, i32, i32, P(f64)]),
    "oracle_pinv": (C.c_int, [i64, P(i32), P(i32), P(f64), P(f64)]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle not built: make -C oracle")
        _lib = _ffi.bind(C.CDLL(LIB), PROTOS)
    return _lib


def ptr(a, t):
    return a.ctypes.data_as(P(t))


def csr_args(g):
    rp = np.ascontiguousarray(g.rowptr, dtype=np.int32)
    ci = np.ascontiguousarray(g.colind, dtype=np.int32)
    v = np.ascontiguousarray(g.vals, dtype=np.float64)
    return rp, ci, v


def default_options(**kw):
    o = _ffi.ligmg_options()
    lib().oracle_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def canon_sum(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return lib().oracle_canon_sum(ptr(a, f64), len(a))


class OracleSolver:
    """Hierarchy + solve of the oracle on a Graph."""

    def __init__(self, g, opts=None, comp_ptr=None):
        self.g = g
        self.opts = opts if opts is not None else default_options()
        rp, ci, v = csr_args(g)
        self.h = H()
        cp = None if comp_ptr is None else np.ascontiguousarray(comp_ptr, dtype=np.int64)
        nc = 0 if cp is None else len(cp) - 1
        rc = lib().oracle_create(g.n, ptr(rp, i32), ptr(ci, i32), ptr(v, f64),
                                 C.byref(self.opts), nc, None if cp is None else ptr(cp, i64),
                                 C.byref(self.h))
        if rc != 0:
            raise RuntimeError(f"oracle_create failed ({rc})")

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def num_levels(self):
        n = i32()
        lib().oracle_num_levels(self.h, C.byref(n))
        return n.value

    def ordering(self):
        p = np.empty(self.g.n, np.int32)
        assert lib().oracle_ordering(self.h, p.ctypes.data) == 0
        return p

    def level_info(self, l):
        k, n, nnz, nc = i32(), i64(), i64(), i64()
        rc = lib().oracle_level_info(self.h, l, C.byref(k), C.byref(n), C.byref(nnz), C.byref(nc))
        assert rc == 0
        return k.value, n.value, nnz.value, nc.value

    def level_matrix(self, l):
        k, n, nnz, nc = self.level_info(l)
        rp = np.empty(n + 1, np.int32)
        ci = np.empty(nnz, np.int32)
        v = np.empty(nnz, np.float64)
        lib().oracle_level_matrix(self.h, l, ptr(rp, i32), ptr(ci, i32), ptr(v, f64))
        return rp, ci, v

    def level_map(self, l):
        k, n, nnz, nc = self.level_info(l)
        m = np.empty(n, np.int32)
        assert lib().oracle_level_map(self.h, l, ptr(m, i32)) == 0
        return m

    def level_states(self, l):
        k, n, nnz, nc = self.level_info(l)
        ind = np.empty(n, np.int32)
        st = np.empty(n, np.int32)
        assert lib().oracle_level_states(self.h, l, ptr(ind, i32), ptr(st, i32)) == 0
        return ind, st

    def level_strength(self, l):
        k, n, nnz, nc = self.level_info(l)
        s = np.empty(nnz, np.float64)
        assert lib().oracle_level_strength(self.h, l, ptr(s, f64)) == 0
        return s

    def level_cheby(self, l):
        lo, hi = f64(), f64()
        assert lib().oracle_level_cheby(self.h, l, C.byref(lo), C.byref(hi)) == 0
        return lo.value, hi.value

    def level_pinv(self, l):
        k, n, nnz, nc = self.level_info(l)
        out = np.empty((n, n), np.float64)
        assert lib().oracle_level_pinv(self.h, l, ptr(out, f64)) == 0
        return out

    def apply_cycle(self, l, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty_like(r)
        assert lib().oracle_apply_cycle(self.h, l, ptr(r, f64), ptr(z, f64)) == 0
        return z

    def solve(self, b, cap=1024, level=0):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(b)
        info = _ffi.ligmg_info()
        hist = np.empty(cap, np.float64)
        rc = lib().oracle_solve_from(self.h, level, ptr(b, f64), ptr(x, f64), C.byref(info),
                                     ptr(hist, f64), cap)
        if rc != 0:
            raise RuntimeError(f"oracle_solve failed ({rc})")
        return x, info, hist[:info.history_len].copy()


# ---- graph preprocessing (cc.hpp, main.cpp:79-110) -------------------------
def components(rp, ci):
    rp, ci = np.ascontiguousarray(rp, np.int32), np.ascontiguousarray(ci, np.int32)
    n = len(rp) - 1
    lab = np.empty(n, np.int32)
    k = i64()
    assert lib().oracle_components(n, rp.ctypes.data, ci.ctypes.data, lab.ctypes.data, C.byref(k)) == 0
    return lab, k.value


def largest_component(rp, ci, v):
    rp, ci = np.ascontiguousarray(rp, np.int32), np.ascontiguousarray(ci, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    n = len(rp) - 1
    new_id = np.empty(n, np.int32)
    no, mo = i64(), i64()
    assert lib().oracle_largest_component(n, rp.ctypes.data, ci.ctypes.data, new_id.ctypes.data,
                                          C.byref(no), C.byref(mo)) == 0
    orp = np.empty(no.value + 1, np.int32)
    oci = np.empty(max(mo.value, 1), np.int32)
    ov = np.empty(max(mo.value, 1), np.float64)
    assert lib().oracle_extract(n, rp.ctypes.data, ci.ctypes.data, v.ctypes.data, new_id.ctypes.data, no.value,
                                orp.ctypes.data, oci.ctypes.data, ov.ctypes.data) == 0
    return orp, oci[:mo.value], ov[:mo.value], new_id


def adjacency_to_laplacian(rp, ci, v):
    rp, ci = np.ascontiguousarray(rp, np.int32), np.ascontiguousarray(ci, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    n = len(rp) - 1
    has = np.array([np.any(ci[rp[i]:rp[i + 1]] == i) for i in range(n)])
    m = len(ci) + int((~has).sum())
    orp = np.empty(n + 1, np.int32)
    oci = np.empty(max(m, 1), np.int32)
    ov = np.empty(max(m, 1), np.float64)
    assert lib().oracle_adjacency_to_laplacian(n, rp.ctypes.data, ci.ctypes.data, v.ctypes.data, orp.ctypes.data,
                                               oci.ctypes.data, ov.ctypes.data) == 0
    return orp, oci[:m], ov[:m]
