"""The smoother's lambda_max estimate restated in numpy: 10 steps of
preconditioned CG (D^-1 A, Jacobi) from the seeded start vector, the largest
eigenvalue of the Lanczos tridiagonal (smoothing.hpp:128-172 estimates the
same quantity through PETSc).  With left-to-right inner products it follows
the oracle's loops exactly; other summation orders measure how much rounding
alone moves the estimate on a given matrix."""
import math

import numpy as np
import scipy.sparse as sp

STREAM_EIG = 0x4549470000000000


def _mix64(z):
    z = z + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def start_vector(n, seed, level):
    """u01(seed, STREAM_EIG + level, i): the splitmix64 finaliser applied
    twice to (seed, stream, index), 53 bits."""
    with np.errstate(over="ignore"):
        s = _mix64(np.uint64(seed) ^ _mix64(np.uint64(STREAM_EIG + level)))
        h = _mix64(s + np.arange(n, dtype=np.uint64))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def seq_dot(a, b):
    return float(sum((a * b).tolist()))


ORDERS = {
    "left to right": seq_dot,
    "right to left": lambda a, b: float(sum((a * b)[::-1].tolist())),
    "pairwise": lambda a, b: float(np.dot(a, b)),
    "exact": lambda a, b: math.fsum((a * b).tolist()),
}


def lambda_max(rp, ci, v, seed, level, iters=10, dot=seq_dot):
    n = len(rp) - 1
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    dg = A.diagonal()
    off = abs(A).sum(1).A1 - abs(dg)
    dinv = np.where((dg != 0) & (off > 0), 1 / np.where(dg != 0, dg, 1), 0.0)
    x0 = start_vector(n, seed, level)
    r = x0 - float(sum(x0.tolist())) / n
    z = dinv * r
    p = z.copy()
    gam = dot(r, z)
    al, be = [], []
    for _ in range(iters):
        if not gam > 0:
            break
        q = A @ p
        de = dot(p, q)
        if not de > 0:
            break
        a = gam / de
        r = r - a * q
        z = dinv * r
        gn = dot(r, z)
        al.append(a)
        be.append(gn / gam)
        p = z + be[-1] * p
        gam = gn
    m = len(al)
    if m == 0:
        return 1.0
    T = np.zeros((m, m))
    for j in range(m):
        T[j, j] = 1 / al[j] + (be[j - 1] / al[j - 1] if j else 0)
        if j + 1 < m:
            T[j, j + 1] = T[j + 1, j] = math.sqrt(be[j]) / al[j]
    return float(np.linalg.eigvalsh(T).max())


def order_spread(rp, ci, v, seed, level, iters=10):
    """Relative spread of the estimate over the summation orders of its inner
    products."""
    lams = [lambda_max(rp, ci, v, seed, level, iters, f) for f in ORDERS.values()]
    return (max(lams) - min(lams)) / abs(lams[0])
