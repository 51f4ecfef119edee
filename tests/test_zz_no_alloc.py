"""ligmg_components_noalloc (no allocation, static per-GPU scratch) against
ligmg_components through the C ABI: the same labels and component count on
two calls in a row, the second on a smaller graph, so scratch left over from
the first call must not leak into the second."""
import ctypes as C

import numpy as np
import pytest

from ligmg_amd import _ffi, graphs

pytestmark = pytest.mark.gpu


def labels(g, noalloc):
    import torch
    lib = _ffi.solver_lib()
    rp = torch.tensor(np.asarray(g.rowptr, np.int32), device="cuda")
    ci = torch.tensor(np.asarray(g.colind, np.int32), device="cuda")
    lab = torch.full((g.n,), -7, dtype=torch.int32, device="cuda")
    k = C.c_int64(-1)
    torch.cuda.synchronize()
    if noalloc:
        rc = lib.ligmg_components_noalloc(g.n, rp.data_ptr(), ci.data_ptr(), lab.data_ptr(), C.byref(k))
    else:
        rc = lib.ligmg_components(g.n, rp.data_ptr(), ci.data_ptr(), _ffi.MEM_DEVICE, lab.data_ptr(), C.byref(k))
    assert rc == 0, _ffi.solver_lib().ligmg_last_error()
    return lab.cpu().numpy(), k.value


def test_noalloc_components_match_components():
    big = graphs.union([graphs.rmat(15, 8, seed=1), graphs.grid2d(120, 90),
                        graphs.barabasi_albert(30000, 3, seed=2), graphs.rmat(12, 4, seed=5)])
    small = graphs.union([graphs.grid2d(30, 20), graphs.rmat(10, 2, seed=7)])
    for g in (big, small):
        print(f"n={g.n} nnz={g.nnz}", flush=True)
        la, ka = labels(g, noalloc=True)
        lb, kb = labels(g, noalloc=False)
        assert ka == kb and ka > 1
        assert np.array_equal(la, lb)


def test_noalloc_components_reject_oversize():
    import torch
    n = _ffi.NOALLOC_CC_MAX_N + 1
    # an edgeless CSR of that size with a full-size label array: the call
    # must return the size error before it reads or writes anything
    rp = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    ci = torch.zeros(1, dtype=torch.int32, device="cuda")
    lab = torch.zeros(n, dtype=torch.int32, device="cuda")
    k = C.c_int64(-1)
    torch.cuda.synchronize()
    rc = _ffi.solver_lib().ligmg_components_noalloc(n, rp.data_ptr(), ci.data_ptr(), lab.data_ptr(), C.byref(k))
    assert rc == _ffi.ERR_INVALID and b"maximum" in _ffi.solver_lib().ligmg_last_error()
