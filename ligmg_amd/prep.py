"""Graph preprocessing on the GPU: connected components, the largest
component, adjacency -> Laplacian (the reference's driver steps,
cc.hpp:22-105 and main.cpp:79-110).

Every function takes a CSR either as host numpy arrays or as torch tensors
on the GPU and returns results of the same kind.  There is no CPU
fallback: the HIP library must be present.
"""
import ctypes as C

import numpy as np

from . import _ffi
from .graphs import Graph
from .solver import _check


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _csr(A):
    if isinstance(A, Graph):
        return A.rowptr, A.colind, A.vals
    return A


def _mem(rp):
    if _is_torch(rp):
        import torch
        assert rp.is_cuda, "torch inputs must live on the GPU"
        torch.cuda.current_stream().synchronize()
        return _ffi.MEM_DEVICE
    return _ffi.MEM_HOST


def _ptr(a):
    return a.data_ptr() if _is_torch(a) else a.ctypes.data


def _empty(like, n, dtype):
    if _is_torch(like):
        import torch
        return torch.empty(n, dtype={np.int32: torch.int32, np.float64: torch.float64}[dtype], device=like.device)
    return np.empty(n, dtype=dtype)


def _host_ints(rp, ci):
    if _is_torch(rp):
        return rp, ci
    return np.ascontiguousarray(rp, np.int32), np.ascontiguousarray(ci, np.int32)


def components(A):
    """label[i] = the largest vertex index of i's component; returns
    (label, ncomp)."""
    rp, ci = _host_ints(*_csr(A)[:2])
    n = len(rp) - 1
    lab = _empty(rp, n, np.int32)
    k = C.c_int64()
    _check(_ffi.solver_lib().ligmg_components(n, _ptr(rp), _ptr(ci), _mem(rp), _ptr(lab), C.byref(k)),
           "ligmg_components")
    return lab, k.value


def largest_component(A):
    """The submatrix on the largest connected component, vertices renumbered
    in ascending order.  Returns (rowptr, colind, vals, new_id)."""
    rp, ci, v = _csr(A)
    rp, ci = _host_ints(rp, ci)
    if not _is_torch(v):
        v = np.ascontiguousarray(v, np.float64)
    n = len(rp) - 1
    mem = _mem(rp)
    lib = _ffi.solver_lib()
    new_id = _empty(rp, n, np.int32)
    no, mo = C.c_int64(), C.c_int64()
    _check(lib.ligmg_largest_component(n, _ptr(rp), _ptr(ci), mem, _ptr(new_id), C.byref(no), C.byref(mo)),
           "ligmg_largest_component")
    orp = _empty(rp, no.value + 1, np.int32)
    oci = _empty(rp, max(mo.value, 1), np.int32)
    ov = _empty(rp, max(mo.value, 1), np.float64)
    _check(lib.ligmg_extract(n, _ptr(rp), _ptr(ci), _ptr(v), _ptr(new_id), no.value, mem, _ptr(orp), _ptr(oci),
                             _ptr(ov)), "ligmg_extract")
    return orp, oci[:mo.value], ov[:mo.value], new_id


def adjacency_to_laplacian(A):
    """L = D - A with D the row sums of A; returns (rowptr, colind, vals)."""
    rp, ci, v = _csr(A)
    rp, ci = _host_ints(rp, ci)
    if not _is_torch(v):
        v = np.ascontiguousarray(v, np.float64)
    n = len(rp) - 1
    mem = _mem(rp)
    lib = _ffi.solver_lib()
    m = C.c_int64()
    _check(lib.ligmg_laplacian_nnz(n, _ptr(rp), _ptr(ci), mem, C.byref(m)), "ligmg_laplacian_nnz")
    orp = _empty(rp, n + 1, np.int32)
    oci = _empty(rp, max(m.value, 1), np.int32)
    ov = _empty(rp, max(m.value, 1), np.float64)
    _check(lib.ligmg_adjacency_to_laplacian(n, _ptr(rp), _ptr(ci), _ptr(v), mem, _ptr(orp), _ptr(oci), _ptr(ov)),
           "ligmg_adjacency_to_laplacian")
    return orp, oci[:m.value], ov[:m.value]
