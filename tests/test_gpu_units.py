"""SpMV work units built on the device equal the host greedy packing (the
same unit rows, row counts, entry offsets, long rows and partial slots), so
every SpMV sums each row in the same order whichever builder ran."""
import ctypes as C

import numpy as np
import pytest

from ligmg_amd import _ffi, graphs
from ligmg_amd.solver import _check

pytestmark = pytest.mark.gpu

UNIT_NNZ = 2048
UNIT_ROWS = 512


def check(lengths):
    rp = np.zeros(len(lengths) + 1, dtype=np.int32)
    np.cumsum(lengths, out=rp[1:])
    same, nu = C.c_int32(0), C.c_int64(0)
    _check(_ffi.solver_lib().ligmg_check_units(len(lengths), rp.ctypes.data_as(C.POINTER(C.c_int32)),
                                               C.byref(same), C.byref(nu)), "ligmg_check_units")
    return same.value, nu.value


@pytest.mark.parametrize("n,seed", [(65536, 1), (300000, 2), (1 << 20, 3)])
def test_units_heavy_tailed_rows(n, seed):
    rng = np.random.default_rng(seed)
    L = np.minimum(rng.pareto(1.2, n) * 8, 20000).astype(np.int64)
    L[rng.random(n) < 0.1] = 0  # empty rows
    L[rng.integers(0, n, 20)] = UNIT_NNZ  # exactly one unit
    L[rng.integers(0, n, 20)] = UNIT_NNZ + 1  # one entry over: two chunks
    same, nu = check(L)
    print(f"n={n} nnz={L.sum()} units={nu}", flush=True)
    assert same == 1


@pytest.mark.parametrize("kind", ["empty", "full", "over", "alternating"])
def test_units_regular_rows(kind):
    n = 100000
    L = {"empty": np.zeros(n, np.int64), "full": np.full(n, UNIT_NNZ // 4, np.int64),
         "over": np.full(n, UNIT_NNZ + 1, np.int64)[: n // 10],
         "alternating": np.tile([1, UNIT_NNZ - 1, 3 * UNIT_NNZ], n // 3)}[kind]
    if kind == "over":
        L = np.concatenate([L, np.ones(n - len(L), np.int64)])
    same, nu = check(L)
    if kind == "empty":
        assert nu == -(-n // UNIT_ROWS)  # UNIT_ROWS empty rows per unit
    if kind == "full":
        assert nu == n // 4
    assert same == 1


def test_units_rmat_rows():
    g = graphs.rmat_config(17, "lognormal", pendants=True)
    same, _ = check(np.diff(g.rowptr.astype(np.int64)))
    assert same == 1
