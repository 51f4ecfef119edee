"""PETSc binary files, the format of hierarchy dumps (the reference writes
them with write_petsc, util.hpp:589-599 for matrices and
combblas_extra.hpp:26-43 for vectors).  Everything is big-endian:

    matrix: int32 1211216, rows, cols, nnz, int32[rows] row lengths,
            int32[nnz] column indices, float64[nnz] values
    vector: int32 1211214, n, float64[n]

Reading needs no PETSc; scipy holds the matrices.
"""
import numpy as np
import scipy.sparse as sp

MAT_CLASSID = 1211216
VEC_CLASSID = 1211214
_I4 = np.dtype(">i4")
_F8 = np.dtype(">f8")


def read_mat(path):
    """scipy CSR matrix from a PETSc binary matrix file."""
    raw = np.fromfile(path, dtype=np.uint8)
    hdr = raw[:16].view(_I4)
    if len(hdr) < 4 or hdr[0] != MAT_CLASSID:
        raise ValueError(f"{path}: not a PETSc binary matrix")
    m, n, nnz = (int(x) for x in hdr[1:4])
    off = 16
    lens = raw[off:off + 4 * m].view(_I4).astype(np.int64)
    off += 4 * m
    ci = raw[off:off + 4 * nnz].view(_I4).astype(np.int32)
    off += 4 * nnz
    v = raw[off:off + 8 * nnz].view(_F8).astype(np.float64)
    if len(lens) != m or len(ci) != nnz or len(v) != nnz or lens.sum() != nnz:
        raise ValueError(f"{path}: truncated or inconsistent")
    rp = np.zeros(m + 1, np.int64)
    np.cumsum(lens, out=rp[1:])
    return sp.csr_matrix((v, ci, rp), shape=(m, n))


def read_vec(path):
    """float64 array from a PETSc binary vector file."""
    raw = np.fromfile(path, dtype=np.uint8)
    hdr = raw[:8].view(_I4)
    if len(hdr) < 2 or hdr[0] != VEC_CLASSID:
        raise ValueError(f"{path}: not a PETSc binary vector")
    n = int(hdr[1])
    x = raw[8:8 + 8 * n].view(_F8).astype(np.float64)
    if len(x) != n:
        raise ValueError(f"{path}: truncated")
    return x


def write_mat(path, A):
    """Write a scipy sparse matrix (rows in stored order)."""
    A = sp.csr_matrix(A)
    m, n = A.shape
    with open(path, "wb") as f:
        f.write(np.array([MAT_CLASSID, m, n, A.nnz], _I4).tobytes())
        f.write(np.diff(A.indptr).astype(_I4).tobytes())
        f.write(A.indices.astype(_I4).tobytes())
        f.write(A.data.astype(_F8).tobytes())


def write_vec(path, x):
    x = np.asarray(x, dtype=np.float64)
    with open(path, "wb") as f:
        f.write(np.array([VEC_CLASSID, len(x)], _I4).tobytes())
        f.write(x.astype(_F8).tobytes())
