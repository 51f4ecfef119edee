// CPU restatement of LigMG's setup + solve path.  TEST INFRASTRUCTURE ONLY:
// the product (ligmg_amd/) never links, loads or calls this file.
//
// Reference citations are relative to the reference repository root.
//
// Deliberate, documented deviations from the reference (all are places where
// the reference's result depends on an unpinned third-party order or on
// rand() seeded with the clock):
//  * the sums that CombBLAS Reduce or LigMG's own loops compute on one rank
//    (col_sum, the vote tally, the Row/Column scaling sums, weighted
//    degrees) are sequential left folds from 0.0 in ascending index, as the
//    reference computes them (seq_sum below); every other order-dependent
//    floating-point sum (SpMV-shaped sums whose order CombBLAS and PETSc
//    leave unpinned) uses canon_sum();
//  * test vectors come from a counter-based RNG (ligmg_rng::tv_bits) with an
//    exact integer mean instead of rand()/RAND_MAX (combblas_extra.hpp:16-22);
//  * P entries of an elimination level use row e of L instead of column e
//    (identical for a bitwise-symmetric L); the Schur complement is formed
//    directly as L_ab + sum_e L_ea * P(e,b) instead of two SpGEMMs;
//  * the Chebyshev interval comes from a deterministic CG-Lanczos estimate,
//    one per level, shared by pre- and post-smoothing (smoothing.hpp:128-172
//    uses PETSc's estimate per smoother object);
//  * the coarsest level is a dense pseudo-inverse (multigrid.hpp:217-257 uses
//    PETSc CG + SOR);
//  * Krylov inner products and the constant-nullspace projection are taken
//    per block of comp_ptr (one block = the reference).
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../ligmg_amd/csrc/rng.hpp"

#ifdef _OPENMP
#include <omp.h>
#include <parallel/algorithm>
#endif

namespace {

// ---------------------------------------------------------------------------
// Canonical summation order.  Sequences of at most 64 terms are summed left
// to right from 0.0; up to 4096 terms they are split into 64 strided
// partial sums (term t goes to partial t % 64, each partial summed left to
// right from 0.0) combined by the tree p[l] += p[l + off], off = 32..1;
// longer ones are cut into consecutive 4096-term chunks, each summed by the
// previous rule, and the chunk results are added left to right from 0.0.
// Rows are summed over all their stored positions, masked positions
// contributing +0.0, so the split depends only on the row length.
// ---------------------------------------------------------------------------
double canon_sum_chunk(const double *a, int64_t m) {
  if (m <= 64) {
    double s = 0.0;
    for (int64_t i = 0; i < m; i++) s += a[i];
    return s;
  }
  double p[64];
  for (int l = 0; l < 64; l++) p[l] = 0.0;
  for (int64_t i = 0; i < m; i++) p[i & 63] += a[i];
  for (int off = 32; off >= 1; off >>= 1)
    for (int l = 0; l < off; l++) p[l] = p[l] + p[l + off];
  return p[0];
}

double canon_sum(const double *a, int64_t m) {
  if (m <= 4096) return canon_sum_chunk(a, m);
  double s = 0.0;
  for (int64_t c = 0; c < m; c += 4096) s += canon_sum_chunk(a + c, std::min<int64_t>(4096, m - c));
  return s;
}

// The reference's one-rank order for the two sums that decide aggregation
// and for the other CombBLAS reductions: a left fold from 0.0 over the terms
// in ascending index.  tally_votes adds each voter's weight into
// new_votes[ind] for i = 0, 1, ... (aggregation.hpp:539-547); the mpi
// allreduce that follows is the identity on one rank (549-550).
// SpParMat::Reduce(Column | Row, plus, 0.0) folds each column (row) from the
// identity in ascending row (column) order (aggregation.hpp:353-359,
// 863-864; main.cpp:86).
//
// decision_order is a test hook that plants the wrong order on purpose
// (1: the canonical split of canon_sum; 2: the tally in descending voter
// order), so tests can show that a hand-built case tells the orders apart.
int decision_order = 0;

double seq_sum(const double *a, int64_t m) {
  if (decision_order == 1) return canon_sum(a, m);
  double s = 0.0;
  for (int64_t i = 0; i < m; i++) s += a[i];
  return s;
}

double tally_sum(const double *a, int64_t m) {
  if (decision_order == 2) {
    double s = 0.0;
    for (int64_t i = m - 1; i >= 0; i--) s += a[i];
    return s;
  }
  return seq_sum(a, m);
}

struct Csr {
  int64_t n = 0;
  std::vector<int32_t> rp, ci;
  std::vector<double> v;
  int64_t nnz() const { return (int64_t)ci.size(); }
};

Csr make_csr(int64_t n, const int32_t *rp, const int32_t *ci, const double *v) {
  Csr A;
  A.n = n;
  A.rp.assign(rp, rp + n + 1);
  A.ci.assign(ci + rp[0], ci + rp[n]);
  A.v.assign(v + rp[0], v + rp[n]);
  if (rp[0] != 0)
    for (auto &x : A.rp) x -= rp[0];
  return A;
}

// ---- vertex order (new; the reference only permutes, randomly, across
// ranks: options.hpp:119, main.cpp:52-63).  Descending stored-entry count
// inside each block, ties by ascending index; the permuted matrix keeps
// its columns sorted.
std::vector<int32_t> degree_order(const Csr &A, const std::vector<int64_t> &blocks) {
  std::vector<int32_t> perm(A.n);
  std::iota(perm.begin(), perm.end(), 0);
  auto len = [&](int32_t i) { return A.rp[i + 1] - A.rp[i]; };
  for (size_t c = 0; c + 1 < blocks.size(); c++)
    std::stable_sort(perm.begin() + blocks[c], perm.begin() + blocks[c + 1],
                     [&](int32_t a, int32_t b) { return len(a) > len(b); });
  return perm;
}

Csr permute_csr(const Csr &A, const std::vector<int32_t> &perm) {
  std::vector<int32_t> inv(A.n);
  for (int64_t i = 0; i < A.n; i++) inv[perm[i]] = (int32_t)i;
  Csr B;
  B.n = A.n;
  B.rp.assign(A.n + 1, 0);
  for (int64_t i = 0; i < A.n; i++) B.rp[i + 1] = B.rp[i] + (A.rp[perm[i] + 1] - A.rp[perm[i]]);
  B.ci.resize(A.ci.size());
  B.v.resize(A.v.size());
  std::vector<std::pair<int32_t, double>> row;
  for (int64_t i = 0; i < A.n; i++) {
    const int32_t o = perm[i];
    row.clear();
    for (int32_t k = A.rp[o]; k < A.rp[o + 1]; k++) row.push_back({inv[A.ci[k]], A.v[k]});
    std::sort(row.begin(), row.end(),
              [](const std::pair<int32_t, double> &a, const std::pair<int32_t, double> &b) {
                return a.first < b.first;
              });
    for (size_t t = 0; t < row.size(); t++) {
      B.ci[B.rp[i] + t] = row[t].first;
      B.v[B.rp[i] + t] = row[t].second;
    }
  }
  return B;
}

// A row with no nonzero off-diagonal is an isolated vertex: a connected
// block reduced to one vertex, whose diagonal is the round-off (or exact 0)
// of a Galerkin sum that cancels.  Its value is a nullspace component, so
// the smoother and the elimination give it 0 instead of 1 / (round-off).
bool isolated(const Csr &A, int64_t i) {
  for (int32_t k = A.rp[i]; k < A.rp[i + 1]; k++)
    if (A.ci[k] != i && A.v[k] != 0) return false;
  return true;
}

double diag_of(const Csr &A, int64_t i) {
  for (int32_t k = A.rp[i]; k < A.rp[i + 1]; k++)
    if (A.ci[k] == i) return A.v[k];
  return 0.0;
}

// y = A x.  Loops marked "omp" run their iterations on several threads when
// the oracle is built with OpenMP; each iteration writes its own outputs and
// keeps its own summation order, so the results do not depend on the thread
// count (the CPU baseline's multi-threaded leg).
void spmv(const Csr &A, const double *x, double *y) {
#pragma omp parallel for schedule(dynamic, 512)
  for (int64_t i = 0; i < A.n; i++) {
    double s = 0.0;
    for (int32_t k = A.rp[i]; k < A.rp[i + 1]; k++) s += A.v[k] * x[A.ci[k]];
    y[i] = s;
  }
}

// ---------------------------------------------------------------------------
// Levels
// ---------------------------------------------------------------------------
enum { ST_SEED = 0, ST_SEEDNEIGHBOR = 1, ST_DECIDED = 2, ST_DECIDEDROOT = 3, ST_UNDECIDED = 4 };
enum { V_SEED = 0, V_UNDECIDED = 1, V_DECIDED = 2 };

struct Level {
  int kind = LIGMG_LEVEL_AGG;
  Csr A;
  int64_t n_coarse = 0;
  // elimination (elimination.hpp:20-68)
  std::vector<int32_t> new_id;  // -1 for eliminated
  Csr P;                        // n x n_kept
  Csr R;                        // n_kept x n
  std::vector<double> q;        // Q diagonal (0 for kept rows)
  int64_t nnzP = 0, nnzQ = 0;
  // aggregation (aggregation.hpp:176-274)
  std::vector<int32_t> agg, ind, state;
  std::vector<double> strength;
  std::vector<int32_t> mem_ptr, members;  // aggregate member lists
  std::vector<double> dinv;
  double cheb_lo = 0, cheb_hi = 0, lam = 0;
  // coarse
  std::vector<double> pinv;
};

struct Solver {
  ligmg_options opts;
  std::vector<Level> H;
  std::vector<int64_t> comp_ptr;  // blocks of level 0
  std::vector<int32_t> perm;      // internal position -> caller's vertex (empty: identity)
  double setup_time = 0;
};

// ---- a1: compute_degrees (elimination.hpp:70-80) -------------------------
std::vector<int32_t> degrees(const Csr &A) {
  std::vector<int32_t> d(A.n);
  for (int64_t i = 0; i < A.n; i++) {
    int32_t c = 0;
    for (int32_t k = A.rp[i]; k < A.rp[i + 1]; k++) c += (A.v[k] == 0) ? 0 : 1;
    d[i] = c - 1;
  }
  return d;
}

// ---- a2/a3: candidates, IndxHashMin SpMV, accept test, renumber
//      (elimination.hpp:186-261).  std::hash<int> is the identity, so the
//      semiring picks the smallest candidate index over entries A_ij != 0.
std::vector<int32_t> elim_select(const Csr &A, int max_degree, int64_t &n_kept) {
  auto deg = degrees(A);
  std::vector<int32_t> elim(A.n, 0);
  for (int64_t i = 0; i < A.n; i++) {
    int64_t best = -1;
    for (int32_t k = A.rp[i]; k < A.rp[i + 1]; k++) {
      int32_t j = A.ci[k];
      if (A.v[k] != 0 && deg[j] <= max_degree && (best < 0 || j < best)) best = j;
    }
    elim[i] = (best == i) ? 1 : 0;
  }
  n_kept = A.n - std::accumulate(elim.begin(), elim.end(), (int64_t)0);
  return elim;
}

struct Entry {
  int32_t r, c;
  double v;
};

// Contract a list of (row, col, value) triples into CSR: stable order by
// (row, col), values of equal keys combined with canon_sum in list order.
Csr contract(int64_t nrows, std::vector<Entry> &E) {
  auto by_key = [](const Entry &a, const Entry &b) { return a.r != b.r ? a.r < b.r : a.c < b.c; };
  // a stable sort has one result, whatever algorithm or thread count
#ifdef _OPENMP
  __gnu_parallel::stable_sort(E.begin(), E.end(), by_key);
#else
  std::stable_sort(E.begin(), E.end(), by_key);
#endif
  Csr C;
  C.n = nrows;
  C.rp.assign(nrows + 1, 0);
  std::vector<double> buf;
  size_t s = 0;
  while (s < E.size()) {
    size_t e = s;
    while (e < E.size() && E[e].r == E[s].r && E[e].c == E[s].c) e++;
    buf.resize(e - s);
    for (size_t t = s; t < e; t++) buf[t - s] = E[t].v;
    C.ci.push_back(E[s].c);
    C.v.push_back(canon_sum(buf.data(), (int64_t)buf.size()));
    C.rp[E[s].r + 1]++;
    s = e;
  }
  for (int64_t i = 0; i < nrows; i++) C.rp[i + 1] += C.rp[i];
  return C;
}

// ---- a4/a6: P, Q, R assembly (elimination.hpp:269-340) and the Galerkin
//      product R L P (multigrid.hpp:141-142) as the exact Schur complement.
bool single_elimination(const Csr &A, const ligmg_options &o, Level &L, Csr &coarse) {
  int64_t n_kept;
  auto elim = elim_select(A, o.elim_degree, n_kept);
  if ((double)n_kept > (double)A.n * o.elim_factor) return false;  // :226-228
  L.kind = LIGMG_LEVEL_ELIM;
  L.A = A;
  L.n_coarse = n_kept;
  L.new_id.assign(A.n, -1);
  int32_t c = 0;
  for (int64_t i = 0; i < A.n; i++)
    if (!elim[i]) L.new_id[i] = c++;
  L.q.assign(A.n, 0.0);
  L.P.n = A.n;
  L.P.rp.assign(A.n + 1, 0);
  std::vector<Entry> E;
  for (int64_t j = 0; j < A.n; j++) {
    if (!elim[j]) {
      L.P.ci.push_back(L.new_id[j]);
      L.P.v.push_back(1.0);
      for (int32_t k = A.rp[j]; k < A.rp[j + 1]; k++) {
        int32_t i = A.ci[k];
        if (!elim[i]) E.push_back({L.new_id[j], L.new_id[i], A.v[k]});
      }
    } else {
      const double d = diag_of(A, j);
      const double inv = isolated(A, j) ? 0.0 : 1 / d;
      L.q[j] = inv;
      size_t first = L.P.ci.size();
      for (int32_t k = A.rp[j]; k < A.rp[j + 1]; k++) {
        int32_t i = A.ci[k];
        if (i == j) continue;
        if (L.new_id[i] < 0) throw std::runtime_error("adjacent eliminated vertices");
        L.P.ci.push_back(L.new_id[i]);
        L.P.v.push_back(inv * -A.v[k]);
      }
      // Schur contributions: (a, b) += L_ea * P(e, b)
      size_t last = L.P.ci.size();
      size_t kk = first;
      for (int32_t k = A.rp[j]; k < A.rp[j + 1]; k++) {
        if (A.ci[k] == j) continue;
        double wa = A.v[k];
        int32_t na = L.P.ci[kk++];
        for (size_t t = first; t < last; t++) E.push_back({na, L.P.ci[t], wa * L.P.v[t]});
      }
    }
    L.P.rp[j + 1] = (int32_t)L.P.ci.size();
  }
  L.nnzP = L.P.nnz();
  L.nnzQ = A.n - n_kept;
  // R = P^T (n_kept x n), columns ascending
  L.R.n = n_kept;
  L.R.rp.assign(n_kept + 1, 0);
  for (int64_t j = 0; j < A.n; j++)
    for (int32_t k = L.P.rp[j]; k < L.P.rp[j + 1]; k++) L.R.rp[L.P.ci[k] + 1]++;
  for (int64_t a = 0; a < n_kept; a++) L.R.rp[a + 1] += L.R.rp[a];
  L.R.ci.resize(L.P.nnz());
  L.R.v.resize(L.P.nnz());
  {
    std::vector<int32_t> f(L.R.rp.begin(), L.R.rp.end() - 1);
    for (int64_t j = 0; j < A.n; j++)
      for (int32_t k = L.P.rp[j]; k < L.P.rp[j + 1]; k++) {
        int32_t a = L.P.ci[k];
        L.R.ci[f[a]] = (int32_t)j;
        L.R.v[f[a]++] = L.P.v[k];
      }
  }
  coarse = contract(n_kept, E);
  return true;
}

// ---- a7: create_tvs + rand_vec + orthagonalize (aggregation.hpp:24-36,
//      combblas_extra.hpp:16-22, util.hpp:19-24) with an exact mean.
void tv_init(int64_t n, uint64_t seed, int level, int T, double *X) {
  for (int t = 0; t < T; t++) {
    uint64_t S = 0;
    for (int64_t i = 0; i < n; i++) S += ligmg_rng::tv_bits(seed, level, t, i);
    double mean = ((double)S / (double)n) * 0x1p-32;
    for (int64_t i = 0; i < n; i++)
      X[i * T + t] = (double)ligmg_rng::tv_bits(seed, level, t, i) * 0x1p-32 - mean;
  }
}

// ---- a8: jacobi with IgnoreDiagSR, b = 0 (smoothing.hpp:11-32, 78-98) ------
void tv_jacobi(const Csr &A, double *X, int T, int sweeps) {
  const double w = 2.0 / 3.0;
  std::vector<double> Y((size_t)A.n * T);
  for (int s = 0; s < sweeps; s++) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < A.n; i++) {
      int32_t len = A.rp[i + 1] - A.rp[i];
      std::vector<double> buf(len);
      double d = 0;
      bool has_d = false;
      for (int32_t k = A.rp[i]; k < A.rp[i + 1]; k++)
        if (A.ci[k] == i) {
          d = A.v[k];
          has_d = true;
        }
      for (int t = 0; t < T; t++) {
        for (int32_t k = 0; k < len; k++) {
          double a = A.v[A.rp[i] + k];
          buf[k] = (a < 0) ? a * X[(int64_t)A.ci[A.rp[i] + k] * T + t] : 0.0;
        }
        double Rx = canon_sum(buf.data(), len);
        Y[i * T + t] = has_d ? w * (0.0 - Rx) / d + (1 - w) * X[i * T + t] : X[i * T + t];
      }
    }
    std::copy(Y.begin(), Y.end(), X);
  }
}

// ---- a9/a10: affinity_matrix (aggregation.hpp:276-404) + col_sum (:863-864)
// S holds the scaled strength at every stored position, 0.0 where pruned.
double affinity_raw(int metric, const double *X, int T, const double *nrm, int64_t r, int64_t c,
                    double a_rc) {
  if (r == c) return metric == LIGMG_STRENGTH_NONE ? (a_rc < 0 ? -a_rc : 0) : 0.0;
  if (metric == LIGMG_STRENGTH_AFFINITY) {
    double top = 0;
    for (int t = 0; t < T; t++) top += X[c * T + t] * X[r * T + t];
    return top * top / (nrm[c] * nrm[c] * nrm[r] * nrm[r]);
  } else if (metric == LIGMG_STRENGTH_ALGEBRAIC_DISTANCE) {
    double m = 0;
    for (int t = 0; t < T; t++) m = std::max(m, std::fabs(X[c * T + t] - X[r * T + t]));
    return 1 / m;
  }
  return a_rc < 0 ? -a_rc : 0;
}

// position of entry (c, r) in row c (the pattern is structurally symmetric)
int32_t transpose_pos(const Csr &A, int64_t r, int32_t c) {
  const int32_t *b = A.ci.data() + A.rp[c], *e = A.ci.data() + A.rp[c + 1];
  const int32_t *p = std::lower_bound(b, e, (int32_t)r);
  if (p == e || *p != r) throw std::runtime_error("pattern not symmetric");
  return (int32_t)(p - A.ci.data());
}

// column sums of values on A's pattern, Reduce(Column, plus, 0.0): column
// c's entries V(r, c) for r in row c's columns, folded in ascending r.  A
// pruned position holds +0.0, which leaves a sequential sum unchanged, so
// summing over the stored pattern equals summing over the pruned matrix.
void column_sums(const Csr &A, const std::vector<double> &V, std::vector<double> &out) {
  out.assign(A.n, 0.0);
#pragma omp parallel for schedule(dynamic, 512)
  for (int64_t c = 0; c < A.n; c++) {
    const int32_t len = A.rp[c + 1] - A.rp[c];
    std::vector<double> buf(len, 0.0);
    for (int32_t k = A.rp[c]; k < A.rp[c + 1]; k++) buf[k - A.rp[c]] = V[transpose_pos(A, c, A.ci[k])];
    out[c] = seq_sum(buf.data(), len);
  }
}

void strength(const Csr &A, double *X, int T, int sweeps, int metric, int scaling, double thresh,
              std::vector<double> &S, std::vector<double> &col_sum) {
  if (metric != LIGMG_STRENGTH_NONE) tv_jacobi(A, X, T, sweeps);
  std::vector<double> nrm(A.n, 0.0);
  for (int64_t i = 0; i < A.n; i++) {
    double s = 0;
    for (int t = 0; t < T; t++) s += X[i * T + t] * X[i * T + t];
    nrm[i] = s;
  }
  std::vector<double> raw(A.nnz()), rowmax(A.n, 0.0), colmax(A.n, 0.0);
#pragma omp parallel for schedule(dynamic, 512)
  for (int64_t r = 0; r < A.n; r++)
    for (int32_t k = A.rp[r]; k < A.rp[r + 1]; k++) raw[k] = affinity_raw(metric, X, T, nrm.data(), r, A.ci[k], A.v[k]);
  for (int64_t r = 0; r < A.n; r++)
    for (int32_t k = A.rp[r]; k < A.rp[r + 1]; k++) {
      rowmax[r] = std::max(rowmax[r], raw[k]);
      colmax[A.ci[k]] = std::max(colmax[A.ci[k]], raw[k]);
    }
  // Row / Column / RowColumn / Sym scaling (aggregation.hpp:350-366): row
  // sums, then column sums (ascending row), then (S^T + S) / 2.  A zero sum
  // leaves its (zero) entries at 0.
  if (scaling == LIGMG_AFF_ROW || scaling == LIGMG_AFF_ROWCOLUMN || scaling == LIGMG_AFF_SYM) {
    for (int64_t r = 0; r < A.n; r++) {
      const double rs = seq_sum(raw.data() + A.rp[r], A.rp[r + 1] - A.rp[r]);
      for (int32_t k = A.rp[r]; k < A.rp[r + 1]; k++) raw[k] = rs > 0 ? raw[k] / rs : 0.0;
    }
  }
  if (scaling == LIGMG_AFF_COLUMN || scaling == LIGMG_AFF_ROWCOLUMN) {
    std::vector<double> cs;
    column_sums(A, raw, cs);
    for (int64_t r = 0; r < A.n; r++)
      for (int32_t k = A.rp[r]; k < A.rp[r + 1]; k++) raw[k] = cs[A.ci[k]] > 0 ? raw[k] / cs[A.ci[k]] : 0.0;
  }
  if (scaling == LIGMG_AFF_SYM) {
    std::vector<double> sym(raw.size());
    for (int64_t r = 0; r < A.n; r++)
      for (int32_t k = A.rp[r]; k < A.rp[r + 1]; k++) sym[k] = (raw[transpose_pos(A, r, A.ci[k])] + raw[k]) / 2;
    raw.swap(sym);
  }
  S.assign(A.nnz(), 0.0);
  for (int64_t r = 0; r < A.n; r++)
    for (int32_t k = A.rp[r]; k < A.rp[r + 1]; k++) {
      double s = raw[k];
      if (scaling == LIGMG_AFF_MAX) {
        double den = std::max(rowmax[r], colmax[A.ci[k]]);
        s = den > 0 ? s / den : 0.0;
      }
      S[k] = (s < thresh || s != s) ? 0.0 : s;  // Prune (aggregation.hpp:401)
    }
  column_sums(A, S, col_sum);
}

// ---- a11: AggSR / VoteAgg / vote (aggregation.hpp:406-518) ---------------
struct Vote {
  int st = V_DECIDED;
  int32_t ind = -1;
  double w = -1;
};
bool gt(int a, int b) {
  if (a == V_SEED) return true;
  if (b == V_SEED) return false;
  if (a == V_UNDECIDED) return true;
  if (b == V_UNDECIDED) return false;
  return true;
}
Vote vadd(const Vote &a, const Vote &b) {
  if (a.st == b.st) return (a.w > b.w) ? a : b;
  return gt(a.st, b.st) ? a : b;
}
Vote vmul(int state, int32_t ind, double w, double thresh, bool nn) {
  if (state == ST_SEED || (nn && state == ST_SEEDNEIGHBOR)) return {w > thresh ? V_SEED : V_DECIDED, ind, w};
  if (state == ST_DECIDED || state == ST_DECIDEDROOT) return {V_DECIDED, ind, w};
  if (state == ST_UNDECIDED) return {V_UNDECIDED, ind, w};
  return {V_DECIDED, ind, w};
}

bool aggregated(int st) {
  return st == ST_SEED || st == ST_DECIDED || st == ST_DECIDEDROOT || st == ST_SEEDNEIGHBOR;
}

// reject_aggregates (aggregation.hpp:622-670): an aggregate larger than
// max_size loses its first |members|^2 / (size - 1) members (Decided or
// SeedNeighbor, ascending index), which return to the default AggState
// {-1, Undecided}; its root becomes DecidedRoot.
void reject_aggregates(std::vector<int32_t> &state, std::vector<int32_t> &ind, int64_t max_size) {
  const int64_t n = (int64_t)state.size();
  std::vector<int64_t> sizes(n, 0);
  for (int64_t i = 0; i < n; i++)
    if (aggregated(state[i])) sizes[ind[i]]++;
  std::vector<std::vector<int32_t>> members(n);
  for (int64_t i = 0; i < n; i++)
    if (state[i] == ST_DECIDED || state[i] == ST_SEEDNEIGHBOR) members[ind[i]].push_back((int32_t)i);
  for (int64_t r = 0; r < n; r++) {
    const auto &c = members[r];
    if (c.empty() || sizes[r] <= max_size) continue;
    const int64_t cnt = (int64_t)c.size() * (int64_t)c.size() / (sizes[r] - 1);
    for (int64_t t = 0; t < cnt; t++) {
      state[c[t]] = ST_UNDECIDED;
      ind[c[t]] = -1;
    }
    state[r] = ST_DECIDEDROOT;
  }
}

// ext_int_ratio + conductance_rejection (aggregation.hpp:563-620, 752-776):
// per aggregated vertex, the weight of its off-diagonal edges leaving its
// aggregate over the weight staying inside, each folded from 0 in ascending
// column order as the reference's dcsc_global_local_iter loop does
// (aggregation.hpp:583-606; "sum -= val" equals "sum += -val"); Decided and
// SeedNeighbor vertices above rejection_ratio become {i, Undecided} (or
// {i, Seed} with rejection_to_seed).
void conductance_rejection(const Csr &A, std::vector<int32_t> &state, std::vector<int32_t> &ind,
                           const ligmg_options &o) {
  const int64_t n = A.n;
  std::vector<double> ratio(n), ext, in;
  for (int64_t i = 0; i < n; i++) {
    const int32_t len = A.rp[i + 1] - A.rp[i];
    ext.assign(len, 0.0);
    in.assign(len, 0.0);
    if (aggregated(state[i]))
      for (int32_t k = A.rp[i]; k < A.rp[i + 1]; k++) {
        const int32_t j = A.ci[k];
        if (j == i) continue;
        if (aggregated(state[j]) && ind[j] == ind[i]) in[k - A.rp[i]] = -A.v[k];
        else ext[k - A.rp[i]] = -A.v[k];
      }
    ratio[i] = seq_sum(ext.data(), len) / seq_sum(in.data(), len);
  }
  for (int64_t i = 0; i < n; i++)
    if (ratio[i] > o.rejection_ratio && (state[i] == ST_DECIDED || state[i] == ST_SEEDNEIGHBOR)) {
      state[i] = o.rejection_to_seed ? ST_SEED : ST_UNDECIDED;
      ind[i] = (int32_t)i;
    }
}

// ---- a12-a14: tally_votes, mark_nodes, aggregation driver
//      (aggregation.hpp:522-561, 702-750, 838-964)
int64_t aggregate(const Csr &A, const std::vector<double> &S, const std::vector<double> &col_sum,
                  const ligmg_options &o, std::vector<int32_t> &agg, std::vector<int32_t> &ind,
                  std::vector<int32_t> &state) {
  const int64_t n = A.n;
  ind.resize(n);
  state.assign(n, ST_UNDECIDED);
  for (int64_t i = 0; i < n; i++) ind[i] = (int32_t)i;
  std::vector<double> votes(n, 0.0);
  std::vector<Vote> va(n);
  std::vector<int32_t> vf(n);
  std::vector<std::vector<double>> lists(n);
  for (int iter = 0; iter < o.agg_iters; iter++) {
    double th = o.agg_init_factor * std::pow(o.agg_iter_factor, iter);
#pragma omp parallel for schedule(dynamic, 512)
    for (int64_t i = 0; i < n; i++) {
      Vote y;
      for (int32_t k = A.rp[i]; k < A.rp[i + 1]; k++) {
        if (S[k] == 0.0) continue;  // pruned entry: not in the strength matrix
        int32_t j = A.ci[k];
        y = vadd(y, vmul(state[j], ind[j], S[k], th, o.neighbor_of_neighbor != 0));
      }
      va[i] = y;
    }
    for (auto &l : lists) l.clear();
    for (int64_t i = 0; i < n; i++) {
      vf[i] = -1;
      if (state[i] == ST_UNDECIDED && va[i].ind >= 0 && va[i].w >= th && va[i].ind != i) {
        lists[va[i].ind].push_back(va[i].w);
        vf[i] = va[i].ind;
      }
    }
    // lists[t] holds t's voters' weights in ascending voter index
    for (int64_t t = 0; t < n; t++) {
      double nv = tally_sum(lists[t].data(), (int64_t)lists[t].size());
      votes[t] = nv + votes[t];
    }
    std::vector<int32_t> ns = state, ni = ind;
    const int joined = o.neighbor_of_neighbor ? ST_SEEDNEIGHBOR : ST_DECIDED;
    for (int64_t i = 0; i < n; i++) {
      int st = state[i];
      if (st == ST_DECIDED || st == ST_DECIDEDROOT || st == ST_SEEDNEIGHBOR) {
      } else if (st == ST_SEED) {
        if (votes[i] > (double)o.max_agg_size) ns[i] = ST_DECIDEDROOT;
      } else if (vf[i] >= 0 && iter == 0 && vf[vf[i]] == i) {
        if (i < vf[i]) {
          ns[i] = ST_SEED;
          ni[i] = (int32_t)i;
        } else {
          ns[i] = joined;
          ni[i] = vf[i];
        }
      } else if (va[i].st == V_SEED) {
        ns[i] = joined;
        ni[i] = va[i].ind;
      } else if (votes[i] >= (double)o.votes_to_seed * col_sum[i]) {
        ns[i] = ST_SEED;
        ni[i] = (int32_t)i;
      }
    }
    state.swap(ns);
    ind.swap(ni);
    if (o.max_agg_size > 0) reject_aggregates(state, ind, o.max_agg_size);
    if (iter >= o.agg_iters / 2 && o.rejection_ratio >= 0) conductance_rejection(A, state, ind, o);
  }
  // an unaggregated vertex is its own aggregate; a vertex a rejection left
  // undecided with no target (ind -1, which the reference's
  // piecewise_constant_restriction would use as a row index) is one too
  for (int64_t i = 0; i < n; i++)
    if (ind[i] < 0) ind[i] = (int32_t)i;
  // piecewise_constant_restriction + remove_zero_rows (aggregation.hpp:950-951)
  std::vector<int32_t> root(n, 0);
  for (int64_t i = 0; i < n; i++) root[ind[i]] = 1;
  std::vector<int32_t> cid(n, -1);
  int64_t nc = 0;
  for (int64_t i = 0; i < n; i++)
    if (root[i]) cid[i] = (int32_t)nc++;
  agg.resize(n);
  for (int64_t i = 0; i < n; i++) agg[i] = cid[ind[i]];
  return nc;
}

// ---- a17: Chebyshev smoother (smoothing.hpp:128-188) -----------------------
// Textbook Chebyshev acceleration of Jacobi (Saad, Iterative Methods for
// Sparse Linear Systems, 2nd ed., Alg. 12.1) for D^-1 L on [lo, hi], in the
// three-term form x_{k+1} = x_k + alpha_k (x_k - x_{k-1})
//                          + beta_k D^-1 (b - L x_k).
void cheby_coeffs(double lo, double hi, int steps, std::vector<double> &al, std::vector<double> &be) {
  double theta = (hi + lo) / 2, delta = (hi - lo) / 2, sigma = theta / delta;
  double rho = 1 / sigma;
  al.assign(steps, 0.0);
  be.assign(steps, 0.0);
  be[0] = 1 / theta;
  for (int k = 1; k < steps; k++) {
    double rn = 1 / (2 * sigma - rho);
    al[k] = rn * rho;
    be[k] = 2 * rn / delta;
    rho = rn;
  }
}

void cheby_apply(const Csr &A, const std::vector<double> &dinv, double lo, double hi, int steps,
                 double *x, const double *b) {
  std::vector<double> al, be;
  cheby_coeffs(lo, hi, steps, al, be);
  std::vector<double> prev(x, x + A.n), Ax(A.n), nx(A.n);
  for (int k = 0; k < steps; k++) {
    spmv(A, x, Ax.data());
    for (int64_t i = 0; i < A.n; i++)
      nx[i] = x[i] + al[k] * (x[i] - prev[i]) + be[k] * dinv[i] * (b[i] - Ax[i]);
    std::copy(x, x + A.n, prev.begin());
    std::copy(nx.begin(), nx.end(), x);
  }
}

// reference jacobi smoother: one sweep (smoothing.hpp:206-211)
void jacobi_apply(const Csr &A, double *x, const double *b) {
  const double w = 2.0 / 3.0;
  std::vector<double> y(A.n);
  for (int64_t i = 0; i < A.n; i++) {
    double Rx = 0, d = 0;
    for (int32_t k = A.rp[i]; k < A.rp[i + 1]; k++) {
      if (A.ci[k] == i) d = A.v[k];
      if (A.v[k] < 0) Rx += A.v[k] * x[A.ci[k]];
    }
    y[i] = d != 0 ? w * (b[i] - Rx) / d + (1 - w) * x[i] : x[i];
  }
  std::copy(y.begin(), y.end(), x);
}

// Largest eigenvalue of a symmetric tridiagonal matrix by bisection on the
// Sturm sequence.
double tridiag_max_eig(const std::vector<double> &d, const std::vector<double> &e) {
  int m = (int)d.size();
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < m; i++) {
    double r = (i > 0 ? std::fabs(e[i - 1]) : 0) + (i + 1 < m ? std::fabs(e[i]) : 0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  auto count_below = [&](double x) {  // eigenvalues < x
    int c = 0;
    double q = d[0] - x;
    if (q < 0) c++;
    for (int i = 1; i < m; i++) {
      double qq = (q == 0) ? 1e-300 : q;
      q = d[i] - x - e[i - 1] * e[i - 1] / qq;
      if (q < 0) c++;
    }
    return c;
  };
  for (int it = 0; it < 200; it++) {
    double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (count_below(mid) == m) hi = mid;
    else lo = mid;
  }
  return hi;
}

// lambda_max(D^-1 L) by `iters` steps of Jacobi-preconditioned CG from a
// seeded mean-zero start vector and the Lanczos tridiagonal of its
// coefficients (the estimate PETSc's Chebyshev uses with a CG estimator,
// smoothing.hpp:158-165).
double eig_estimate(const Csr &A, const std::vector<double> &dinv, uint64_t seed, int level,
                    int iters) {
  const int64_t n = A.n;
  if (n == 0) return 1.0;
  std::vector<double> r(n), z(n), p(n), q(n);
  double mean = 0;
  for (int64_t i = 0; i < n; i++) {
    r[i] = ligmg_rng::u01(seed, ligmg_rng::STREAM_EIG + level, i);
    mean += r[i];
  }
  mean /= (double)n;
  for (int64_t i = 0; i < n; i++) r[i] -= mean;
  for (int64_t i = 0; i < n; i++) z[i] = dinv[i] * r[i];
  p = z;
  double gamma = 0;
  for (int64_t i = 0; i < n; i++) gamma += r[i] * z[i];
  std::vector<double> al, be;
  for (int it = 0; it < iters && gamma > 0; it++) {
    spmv(A, p.data(), q.data());
    double delta = 0;
    for (int64_t i = 0; i < n; i++) delta += p[i] * q[i];
    if (!(delta > 0)) break;
    double alpha = gamma / delta;
    for (int64_t i = 0; i < n; i++) r[i] -= alpha * q[i];
    for (int64_t i = 0; i < n; i++) z[i] = dinv[i] * r[i];
    double gn = 0;
    for (int64_t i = 0; i < n; i++) gn += r[i] * z[i];
    double beta = gn / gamma;
    al.push_back(alpha);
    be.push_back(beta);
    for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    gamma = gn;
  }
  int m = (int)al.size();
  if (m == 0) return 1.0;
  std::vector<double> d(m), e(m > 1 ? m - 1 : 0);
  for (int j = 0; j < m; j++) {
    d[j] = 1 / al[j] + (j > 0 ? be[j - 1] / al[j - 1] : 0);
    if (j + 1 < m) e[j] = std::sqrt(be[j]) / al[j];
  }
  return tridiag_max_eig(d, e);
}

// ---- a20: coarse level as a dense pseudo-inverse ---------------------------
// For each connected component c of the coarse graph,
//   L_c^+ = (L_c + 1 1^T / n_c)^-1 - 1 1^T / n_c.
// largest coarse level the dense solve takes: the hierarchy stops at
// min_size entries, so only coarsening that stalls until max_levels exceeds it
constexpr int64_t COARSE_DENSE_MAX = 8192;
struct CoarseTooLarge : std::runtime_error {
  using std::runtime_error::runtime_error;
};

std::vector<double> dense_pinv(const Csr &A) {
  const int64_t n = A.n;
  if (n > COARSE_DENSE_MAX) throw CoarseTooLarge("coarse level too large for the dense solve");
  std::vector<int32_t> comp(n, -1);
  std::vector<int64_t> csize;
  for (int64_t s = 0; s < n; s++) {
    if (comp[s] >= 0) continue;
    int32_t cid = (int32_t)csize.size();
    csize.push_back(0);
    std::vector<int64_t> stack{s};
    comp[s] = cid;
    while (!stack.empty()) {
      int64_t u = stack.back();
      stack.pop_back();
      csize[cid]++;
      for (int32_t k = A.rp[u]; k < A.rp[u + 1]; k++) {
        int32_t v = A.ci[k];
        if (A.v[k] != 0 && comp[v] < 0) {
          comp[v] = cid;
          stack.push_back(v);
        }
      }
    }
  }
  std::vector<double> M(n * n, 0.0);
  for (int64_t i = 0; i < n; i++)
    for (int32_t k = A.rp[i]; k < A.rp[i + 1]; k++) M[i * n + A.ci[k]] += A.v[k];
  for (int64_t i = 0; i < n; i++)
    for (int64_t j = 0; j < n; j++)
      if (comp[i] == comp[j]) M[i * n + j] += 1.0 / (double)csize[comp[i]];
  // symmetrise (coarse matrices are symmetric up to summation order)
  for (int64_t i = 0; i < n; i++)
    for (int64_t j = i + 1; j < n; j++) {
      double s = 0.5 * (M[i * n + j] + M[j * n + i]);
      M[i * n + j] = M[j * n + i] = s;
    }
  // Cholesky M = G G^T (lower)
  std::vector<double> G(n * n, 0.0);
  for (int64_t j = 0; j < n; j++) {
    double s = M[j * n + j];
    for (int64_t k = 0; k < j; k++) s -= G[j * n + k] * G[j * n + k];
    if (!(s > 0)) throw std::runtime_error("coarse matrix not positive semidefinite");
    double gjj = std::sqrt(s);
    G[j * n + j] = gjj;
    for (int64_t i = j + 1; i < n; i++) {
      double t = M[i * n + j];
      for (int64_t k = 0; k < j; k++) t -= G[i * n + k] * G[j * n + k];
      G[i * n + j] = t / gjj;
    }
  }
  // inverse via solves with unit vectors
  std::vector<double> inv(n * n, 0.0), y(n);
  for (int64_t c = 0; c < n; c++) {
    for (int64_t i = 0; i < n; i++) {
      double t = (i == c) ? 1.0 : 0.0;
      for (int64_t k = 0; k < i; k++) t -= G[i * n + k] * y[k];
      y[i] = t / G[i * n + i];
    }
    for (int64_t i = n - 1; i >= 0; i--) {
      double t = y[i];
      for (int64_t k = i + 1; k < n; k++) t -= G[k * n + i] * inv[k * n + c];
      inv[i * n + c] = t / G[i * n + i];
    }
  }
  for (int64_t i = 0; i < n; i++)
    for (int64_t j = 0; j < n; j++)
      if (comp[i] == comp[j]) inv[i * n + j] -= 1.0 / (double)csize[comp[i]];
  return inv;
}

// ---- a14 (tail) + a24: add_aggregation_level / make_hierarchy
//      (multigrid.hpp:132-263)
void aggregation_level(const Csr &A, const ligmg_options &o, int level_idx, Level &L, Csr &coarse) {
  L.kind = LIGMG_LEVEL_AGG;
  L.A = A;
  const int T = o.num_tvs;
  std::vector<double> X((size_t)A.n * T), col_sum;
  tv_init(A.n, o.seed, level_idx, T, X.data());
  strength(A, X.data(), T, o.tvs_smooth, o.strength_metric, o.aff_scaling, o.thresh, L.strength,
           col_sum);
  L.n_coarse = aggregate(A, L.strength, col_sum, o, L.agg, L.ind, L.state);
  std::vector<Entry> E;
  E.reserve(A.nnz());
  for (int64_t i = 0; i < A.n; i++)
    for (int32_t k = A.rp[i]; k < A.rp[i + 1]; k++) E.push_back({L.agg[i], L.agg[A.ci[k]], A.v[k]});
  coarse = contract(L.n_coarse, E);
  // member lists
  L.mem_ptr.assign(L.n_coarse + 1, 0);
  for (int64_t i = 0; i < A.n; i++) L.mem_ptr[L.agg[i] + 1]++;
  for (int64_t c = 0; c < L.n_coarse; c++) L.mem_ptr[c + 1] += L.mem_ptr[c];
  L.members.resize(A.n);
  std::vector<int32_t> f(L.mem_ptr.begin(), L.mem_ptr.end() - 1);
  for (int64_t i = 0; i < A.n; i++) L.members[f[L.agg[i]]++] = (int32_t)i;
  // smoother
  L.dinv.resize(A.n);
  for (int64_t i = 0; i < A.n; i++) {
    double d = diag_of(A, i);
    L.dinv[i] = (d != 0 && !isolated(A, i)) ? 1.0 / d : 0.0;
  }
  if (o.smoother_type == LIGMG_SMOOTHER_CHEBYSHEV) {
    L.lam = eig_estimate(A, L.dinv, o.seed, level_idx, o.cheby_eig_iters);
    L.cheb_lo = o.cheby_lower * L.lam;
    L.cheb_hi = o.cheby_upper * L.lam;
  }
}

void make_hierarchy(Solver &S, const Csr &A0) {
  const ligmg_options &o = S.opts;
  Csr cur = A0;
  while ((int)S.H.size() < o.max_levels &&
         (S.H.empty() || S.H.back().A.nnz() > o.min_size)) {
    if (o.always_elim) {
      Level L;
      Csr c;
      if (single_elimination(cur, o, L, c)) {
        S.H.push_back(std::move(L));
        cur = std::move(c);
      }
    }
    Level L;
    Csr c;
    aggregation_level(cur, o, (int)S.H.size(), L, c);
    S.H.push_back(std::move(L));
    cur = std::move(c);
  }
  Level C;
  C.kind = LIGMG_LEVEL_COARSE;
  C.A = cur;
  C.pinv = dense_pinv(cur);
  S.H.push_back(std::move(C));
}

// ---- a15/a16/a20/a21: do_level, vcycle, kcycle (aggregation.hpp:209-241,
//      elimination.hpp:33-52, multigrid.hpp:49-70, 86-130)
struct Blocks {  // blocks of a vector for inner products / projection
  std::vector<int64_t> ptr;
};

void project_mean(std::vector<double> &z, const Blocks &B) {
  for (size_t c = 0; c + 1 < B.ptr.size(); c++) {
    int64_t a = B.ptr[c], b = B.ptr[c + 1];
    if (b <= a) continue;
    double s = 0;
    for (int64_t i = a; i < b; i++) s += z[i];
    double m = s / (double)(b - a);
    for (int64_t i = a; i < b; i++) z[i] -= m;
  }
}

void cycle(const Solver &S, size_t l, std::vector<double> &x, const std::vector<double> &b);

void smooth(const Level &L, const ligmg_options &o, int steps, std::vector<double> &x,
            const std::vector<double> &b) {
  if (o.smoother_type == LIGMG_SMOOTHER_CHEBYSHEV)
    cheby_apply(L.A, L.dinv, L.cheb_lo, L.cheb_hi, steps, x.data(), b.data());
  else
    jacobi_apply(L.A, x.data(), b.data());
}

// Krylov (cg.hpp:18-116; PETSc KSPCG / KSPFCG).  Per-block coefficients and
// convergence tests; returns iterations, appends block-0 norms to hist.
template <class Prec>
int krylov(const Csr &A, std::vector<double> &x, const std::vector<double> &b, const Blocks &B,
           Prec prec, int type, double rtol, int maxit, std::vector<double> *hist, int *status) {
  const int64_t n = A.n;
  const size_t nb = B.ptr.size() - 1;
  auto bdot = [&](const std::vector<double> &u, const std::vector<double> &v, std::vector<double> &out) {
    out.assign(nb, 0.0);
    for (size_t c = 0; c < nb; c++) {
      double s = 0;
      for (int64_t i = B.ptr[c]; i < B.ptr[c + 1]; i++) s += u[i] * v[i];
      out[c] = s;
    }
  };
  std::vector<double> r(n), z(n), p(n), q(n), tmp(n);
  spmv(A, x.data(), tmp.data());
  for (int64_t i = 0; i < n; i++) r[i] = b[i] - tmp[i];
  std::vector<double> bn, rn, gamma, delta, t1;
  bdot(b, b, bn);
  bdot(r, r, rn);
  std::vector<char> active(nb, 1);
  auto total = [&](const std::vector<double> &v) {
    double s = 0;
    for (double a : v) s += a;
    return std::sqrt(s);
  };
  if (hist) hist->push_back(total(rn));
  for (size_t c = 0; c < nb; c++)
    if (!(std::sqrt(rn[c]) > rtol * std::sqrt(bn[c]))) active[c] = 0;
  int it = 0;
  if (status) *status = LIGMG_CONVERGED;
  bool any = std::any_of(active.begin(), active.end(), [](char a) { return a; });
  if (!any) return 0;
  prec(r, z);
  p = z;
  bdot(r, z, gamma);
  while (it < maxit) {
    spmv(A, p.data(), q.data());
    bdot(p, q, delta);
    std::vector<double> alpha(nb, 0.0);
    if (type == LIGMG_KSP_FCG || type == LIGMG_KSP_PIPEFCG) {
      bdot(p, r, t1);
      for (size_t c = 0; c < nb; c++) alpha[c] = active[c] && delta[c] != 0 ? t1[c] / delta[c] : 0;
    } else {
      for (size_t c = 0; c < nb; c++) alpha[c] = active[c] && delta[c] != 0 ? gamma[c] / delta[c] : 0;
    }
    for (size_t c = 0; c < nb; c++)
      for (int64_t i = B.ptr[c]; i < B.ptr[c + 1]; i++) {
        x[i] += alpha[c] * p[i];
        r[i] -= alpha[c] * q[i];
      }
    it++;
    bdot(r, r, rn);
    if (hist) hist->push_back(total(rn));
    bool nan = false;
    for (size_t c = 0; c < nb; c++) {
      if (rn[c] != rn[c]) nan = true;
      if (active[c] && !(std::sqrt(rn[c]) > rtol * std::sqrt(bn[c]))) active[c] = 0;
    }
    if (nan) {
      if (status) *status = LIGMG_DIVERGED_NAN;
      return it;
    }
    if (!std::any_of(active.begin(), active.end(), [](char a) { return a; })) return it;
    prec(r, z);
    if (type == LIGMG_KSP_FCG || type == LIGMG_KSP_PIPEFCG) {
      bdot(z, q, t1);
      for (size_t c = 0; c < nb; c++) {
        double beta = active[c] && delta[c] != 0 ? -t1[c] / delta[c] : 0;
        for (int64_t i = B.ptr[c]; i < B.ptr[c + 1]; i++) p[i] = z[i] + beta * p[i];
      }
    } else {
      std::vector<double> gn;
      bdot(r, z, gn);
      for (size_t c = 0; c < nb; c++) {
        double beta = active[c] && gamma[c] != 0 ? gn[c] / gamma[c] : 0;
        for (int64_t i = B.ptr[c]; i < B.ptr[c + 1]; i++) p[i] = z[i] + beta * p[i];
      }
      gamma = gn;
    }
  }
  if (status) *status = LIGMG_DIVERGED_ITS;
  return it;
}

// FGMRES(m) (Saad, Iterative Methods for Sparse Linear Systems, 2nd ed.,
// Alg. 9.6; PETSc KSPFGMRES defaults: restart 30, classical Gram-Schmidt,
// right preconditioning, unpreconditioned residual norm), with an Arnoldi
// process, Givens rotations and a convergence test per block.  The history
// holds the norm at the start of every restart cycle (the true residual, as
// PETSc logs it) and the Givens estimate after every iteration.
constexpr int FGMRES_RESTART = 30;

template <class Prec>
int fgmres(const Csr &A, std::vector<double> &x, const std::vector<double> &b, const Blocks &B, Prec prec,
           double rtol, int maxit, std::vector<double> *hist, int *status) {
  const int64_t n = A.n;
  const size_t nb = B.ptr.size() - 1;
  const int m = FGMRES_RESTART;
  auto bdot = [&](const std::vector<double> &u, const std::vector<double> &v, std::vector<double> &out) {
    out.assign(nb, 0.0);
    for (size_t c = 0; c < nb; c++) {
      double s = 0;
      for (int64_t i = B.ptr[c]; i < B.ptr[c + 1]; i++) s += u[i] * v[i];
      out[c] = s;
    }
  };
  auto none = [](const std::vector<char> &f) { return !std::any_of(f.begin(), f.end(), [](char a) { return a; }); };
  std::vector<double> bn, rn, r(n), w(n), tmp(n);
  bdot(b, b, bn);
  if (status) *status = LIGMG_CONVERGED;
  int it = 0;
  std::vector<char> active(nb, 1);
  while (true) {
    // cycle start: the true residual, logged (KSPGMRESCycle)
    spmv(A, x.data(), tmp.data());
    for (int64_t i = 0; i < n; i++) r[i] = b[i] - tmp[i];
    bdot(r, r, rn);
    double t = 0;
    bool nan = false;
    for (size_t c = 0; c < nb; c++) {
      t += rn[c];
      nan |= rn[c] != rn[c];
      if (!(std::sqrt(rn[c]) > rtol * std::sqrt(bn[c]))) active[c] = 0;
    }
    if (hist) hist->push_back(std::sqrt(t));
    if (nan) {
      if (status) *status = LIGMG_DIVERGED_NAN;
      return it;
    }
    if (none(active)) return it;
    if (it >= maxit) {
      if (status) *status = LIGMG_DIVERGED_ITS;
      return it;
    }
    std::vector<std::vector<double>> V(1, std::vector<double>(n, 0.0)), Z;
    std::vector<std::vector<double>> Hc(nb, std::vector<double>((size_t)(m + 1) * m, 0.0)), g(nb), cs(nb), sn(nb);
    std::vector<int> cols(nb, 0);
    std::vector<char> live = active;
    std::vector<double> est(nb);
    for (size_t c = 0; c < nb; c++) {
      const double beta = std::sqrt(rn[c]);
      est[c] = active[c] ? beta : 0.0;
      g[c].assign(m + 1, 0.0);
      g[c][0] = beta;
      cs[c].assign(m, 0.0);
      sn[c].assign(m, 0.0);
      const double s = (live[c] && beta > 0) ? 1.0 / beta : 0.0;
      for (int64_t i = B.ptr[c]; i < B.ptr[c + 1]; i++) V[0][i] = r[i] * s;
    }
    bool done = false;
    for (int j = 0; j < m; j++) {
      Z.emplace_back(n, 0.0);
      prec(V[j], Z[j]);
      spmv(A, Z[j].data(), w.data());
      std::vector<std::vector<double>> h(j + 1);
      for (int i = 0; i <= j; i++) bdot(w, V[i], h[i]);  // classical Gram-Schmidt
      for (size_t c = 0; c < nb; c++) {
        if (!live[c]) continue;
        for (int64_t k = B.ptr[c]; k < B.ptr[c + 1]; k++) {
          double s = w[k];
          for (int i = 0; i <= j; i++) s -= h[i][c] * V[i][k];
          w[k] = s;
        }
      }
      std::vector<double> hn;
      bdot(w, w, hn);
      V.emplace_back(n, 0.0);
      for (size_t c = 0; c < nb; c++) {
        if (!live[c]) continue;
        const double nrm = std::sqrt(hn[c]);
        const double s = nrm > 0 ? 1.0 / nrm : 0.0;
        for (int64_t k = B.ptr[c]; k < B.ptr[c + 1]; k++) V[j + 1][k] = w[k] * s;
        double *col = Hc[c].data() + (size_t)j * (m + 1);  // column j, rows 0..j+1
        for (int i = 0; i <= j; i++) col[i] = h[i][c];
        col[j + 1] = nrm;
        for (int i = 0; i < j; i++) {  // the previous rotations
          const double a = col[i], bb = col[i + 1];
          col[i] = cs[c][i] * a + sn[c][i] * bb;
          col[i + 1] = -sn[c][i] * a + cs[c][i] * bb;
        }
        const double a = col[j], bb = col[j + 1];
        const double d = std::hypot(a, bb);
        cs[c][j] = d > 0 ? a / d : 1.0;
        sn[c][j] = d > 0 ? bb / d : 0.0;
        col[j] = d;
        col[j + 1] = 0.0;
        g[c][j + 1] = -sn[c][j] * g[c][j];
        g[c][j] = cs[c][j] * g[c][j];
        cols[c] = j + 1;
        est[c] = std::fabs(g[c][j + 1]);
        if (!(est[c] > rtol * std::sqrt(bn[c]))) {  // converged (includes a happy breakdown)
          live[c] = 0;
          active[c] = 0;
        } else if (nrm == 0) {
          live[c] = 0;
        }
      }
      it++;
      double t2 = 0;
      for (size_t c = 0; c < nb; c++) t2 += est[c] * est[c];
      done = none(active) || it >= maxit;
      if (j + 1 == m && !done) break;  // restart: the next cycle logs the true residual
      if (hist) hist->push_back(std::sqrt(t2));
      if (done || none(live)) break;
    }
    // x += Z y, H y = g (upper triangular, per block)
    for (size_t c = 0; c < nb; c++) {
      const int k = cols[c];
      std::vector<double> y(k, 0.0);
      for (int i = k - 1; i >= 0; i--) {
        double s = g[c][i];
        for (int q = i + 1; q < k; q++) s -= Hc[c][(size_t)q * (m + 1) + i] * y[q];
        const double d = Hc[c][(size_t)i * (m + 1) + i];
        y[i] = d != 0 ? s / d : 0.0;
      }
      for (int64_t e = B.ptr[c]; e < B.ptr[c + 1]; e++) {
        double s = x[e];
        for (int q = 0; q < k; q++) s += y[q] * Z[q][e];
        x[e] = s;
      }
    }
    if (done) {
      if (status) *status = none(active) ? LIGMG_CONVERGED : LIGMG_DIVERGED_ITS;
      return it;
    }
  }
}

Blocks level_blocks(const Solver &S, size_t l) {
  // map level-0 blocks down through the hierarchy: only the first block
  // structure matters for the outer solve; inner levels use one block per
  // level-0 block mapped through elimination/aggregation ids.
  Blocks B;
  std::vector<int64_t> ptr = S.comp_ptr;
  for (size_t k = 0; k < l; k++) {
    const Level &L = S.H[k];
    std::vector<int64_t> np(ptr.size(), 0);
    for (size_t c = 0; c + 1 < ptr.size(); c++) {
      int64_t cnt = 0;
      if (L.kind == LIGMG_LEVEL_ELIM) {
        for (int64_t i = ptr[c]; i < ptr[c + 1]; i++) cnt += L.new_id[i] >= 0;
      } else {
        for (int64_t i = ptr[c]; i < ptr[c + 1]; i++) cnt += (L.ind[i] == i);
      }
      np[c + 1] = np[c] + cnt;
    }
    ptr = np;
  }
  B.ptr = ptr;
  return B;
}

void do_level(const Solver &S, size_t l, std::vector<double> &x, const std::vector<double> &b,
              bool k);

void recurse(const Solver &S, size_t l, std::vector<double> &x, const std::vector<double> &b,
             bool kc) {
  if (!kc || S.H[l + 1].kind == LIGMG_LEVEL_ELIM) {
    do_level(S, l + 1, x, b, kc);
    return;
  }
  // K-cycle: kcycle_iters steps of FCG on the next level (multigrid.hpp:113-126)
  Blocks B = level_blocks(S, l + 1);
  const Level &N = S.H[l + 1];
  krylov(
      N.A, x, b, B,
      [&](const std::vector<double> &r, std::vector<double> &z) {
        std::vector<double> xx(r.size(), 0.0);
        do_level(S, l + 1, xx, r, kc);
        z = xx;
        project_mean(z, B);
      },
      S.opts.kcycle_ksp_type, S.opts.tol, S.opts.kcycle_iters, nullptr, nullptr);
}

void do_level(const Solver &S, size_t l, std::vector<double> &x, const std::vector<double> &b,
              bool kc) {
  const Level &L = S.H[l];
  const ligmg_options &o = S.opts;
  if (L.kind == LIGMG_LEVEL_COARSE) {
    const int64_t n = L.A.n;
    for (int64_t i = 0; i < n; i++) {
      double s = 0;
      for (int64_t j = 0; j < n; j++) s += L.pinv[i * n + j] * b[j];
      x[i] = s;
    }
    return;
  }
  if (L.kind == LIGMG_LEVEL_ELIM) {
    std::vector<double> r(L.n_coarse), xc(L.n_coarse, 0.0);
    spmv(L.R, b.data(), r.data());
    recurse(S, l, xc, r, kc);
    spmv(L.P, xc.data(), x.data());
    for (int64_t i = 0; i < L.A.n; i++) x[i] += L.q[i] * b[i];
    return;
  }
  smooth(L, o, o.pre_smooth, x, b);
  std::vector<double> Ax(L.A.n), bc(L.n_coarse, 0.0), xc(L.n_coarse, 0.0);
  spmv(L.A, x.data(), Ax.data());
  for (int64_t c = 0; c < L.n_coarse; c++) {
    double s = 0;
    for (int32_t k = L.mem_ptr[c]; k < L.mem_ptr[c + 1]; k++) {
      int32_t i = L.members[k];
      s += b[i] - Ax[i];
    }
    bc[c] = s;
  }
  recurse(S, l, xc, bc, kc);
  for (int64_t i = 0; i < L.A.n; i++) x[i] += xc[L.agg[i]];
  smooth(L, o, o.post_smooth, x, b);
  if (o.constant_correction)
    for (auto &v : x) v = v * 4.0 / 3.0;
}

void cycle(const Solver &S, size_t l, std::vector<double> &x, const std::vector<double> &b) {
  do_level(S, l, x, b, S.opts.kcycle != 0);
}

// ---- a25: cycle_complexity (multigrid.hpp:372-387) + Level::work
double level_work(const Level &L, const ligmg_options &o) {
  if (L.kind == LIGMG_LEVEL_AGG)
    return (double)((o.pre_smooth + o.post_smooth + 1) * L.A.nnz() + 2 * L.A.n + L.A.n);
  if (L.kind == LIGMG_LEVEL_ELIM) return (double)(2 * L.nnzP + L.nnzQ);
  return (double)L.A.nnz();
}

double cycle_complexity(const Solver &S, size_t start) {
  double sum = 0;
  double it = 1;
  for (size_t i = start; i < S.H.size(); i++) {
    sum += it * level_work(S.H[i], S.opts);
    bool exact = S.H[i].kind == LIGMG_LEVEL_ELIM;
    if (!exact && S.opts.kcycle) it *= S.opts.kcycle_iters;
  }
  return sum / (double)S.H[0].A.nnz();
}

// ---- a22/a23: ligmg::solve (multigrid.hpp:315-370) -------------------------
// The system of level `from` (Solver::solve(b, from), ligmg.hpp:36-38): an
// elimination level is applied once around the outer Krylov solve on
// from + 1; the full residual is taken on H[from].
int solve(const Solver &S, const double *bin, double *xout, ligmg_info *info, double *hist,
          int cap, size_t from = 0) {
  auto t0 = std::chrono::steady_clock::now();
  const Level &L0 = S.H[from];
  const int64_t n = L0.A.n;
  std::vector<double> b(bin, bin + n);
  bool do_once = S.opts.top_elim && L0.kind == LIGMG_LEVEL_ELIM;
  size_t base = do_once ? from + 1 : from;
  std::vector<double> rhs;
  if (do_once) {
    rhs.resize(L0.n_coarse);
    spmv(L0.R, b.data(), rhs.data());
  } else {
    rhs = b;
  }
  Blocks B = level_blocks(S, base);
  project_mean(rhs, B);
  std::vector<double> x(rhs.size(), 0.0), h;
  int status = 0;
  int outer = S.opts.outer_ksp_type;
  auto prec = [&](const std::vector<double> &r, std::vector<double> &z) {
    std::vector<double> xx(r.size(), 0.0);
    cycle(S, base, xx, r);
    z = xx;
    project_mean(z, B);
  };
  int iters = outer == LIGMG_KSP_FGMRES
                  ? fgmres(S.H[base].A, x, rhs, B, prec, S.opts.tol, S.opts.iters, &h, &status)
                  : krylov(S.H[base].A, x, rhs, B, prec, outer, S.opts.tol, S.opts.iters, &h, &status);
  std::vector<double> sol(n);
  if (do_once) {
    spmv(L0.P, x.data(), sol.data());
    for (int64_t i = 0; i < n; i++) sol[i] += L0.q[i] * b[i];
  } else {
    sol = x;
  }
  auto t1 = std::chrono::steady_clock::now();
  std::copy(sol.begin(), sol.end(), xout);
  std::vector<double> r(n);
  spmv(L0.A, sol.data(), r.data());
  double rr = 0, bb = 0;
  for (int64_t i = 0; i < n; i++) {
    rr += (b[i] - r[i]) * (b[i] - r[i]);
    bb += b[i] * b[i];
  }
  if (info) {
    double cc = cycle_complexity(S, base);
    info->iters = iters;
    info->status = status;
    info->base_level = (int)base;
    info->history_len = (int)std::min<size_t>(h.size(), cap > 0 ? cap : 0);
    info->rel_residual = h.back() / h.front();
    info->full_rel_residual = bb > 0 ? std::sqrt(rr / bb) : 0;
    info->time = std::chrono::duration<double>(t1 - t0).count();
    info->work = iters * (1 + cc);
    info->cycle_complexity = iters > 0 ? info->work / iters : 0;
    info->cf = iters > 0 ? std::pow(info->rel_residual, 1.0 / iters) : 0;
    info->ecf = iters > 0 ? std::pow(info->rel_residual, 1.0 / (iters * info->cycle_complexity)) : 0;
    info->wda = iters > 0 ? -1.0 / std::log10(info->ecf) : 0;
    info->setup_time = S.setup_time;
  }
  if (hist)
    for (int i = 0; i < (int)h.size() && i < cap; i++) hist[i] = h[i];
  return 0;
}

bool check_opts(const ligmg_options *o) {
  return o->elim_iters == 1 && o->num_tvs >= 1 && o->num_tvs <= 16 &&
         o->aff_scaling >= LIGMG_AFF_NONE && o->aff_scaling <= LIGMG_AFF_MAX;
}

}  // namespace

struct oracle_solver : Solver {};

extern "C" {

double oracle_canon_sum(const double *a, int64_t m) { return canon_sum(a, m); }
double oracle_seq_sum(const double *a, int64_t m) { return seq_sum(a, m); }
void oracle_set_decision_order(int32_t mode) { decision_order = mode; }
int32_t oracle_set_threads(int32_t nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
  return omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

void oracle_default_options(ligmg_options *o) {
  memset(o, 0, sizeof(*o));
  o->verbose = 0;
  o->always_elim = 1;
  o->top_elim = 1;
  o->constant_correction = 0;
  o->strength_metric = LIGMG_STRENGTH_AFFINITY;
  o->max_levels = 50;
  o->min_size = 4000;
  o->elim_degree = 4;
  o->tol = 1e-8;
  o->thresh = 1e-8;
  o->agg_iters = 10;
  o->iters = 500;
  o->pre_smooth = 2;
  o->post_smooth = 2;
  o->num_tvs = 10;
  o->tvs_smooth = 4;
  o->smoother_type = LIGMG_SMOOTHER_CHEBYSHEV;
  o->kcycle = 0;
  o->outer_ksp_type = LIGMG_KSP_CG;
  o->kcycle_ksp_type = LIGMG_KSP_FCG;
  o->kcycle_iters = 2;
  o->votes_to_seed = 1;
  o->max_agg_size = -1;
  o->agg_iter_factor = 0.5;
  o->agg_init_factor = 0.5;
  o->cheby_lower = 0.3;
  o->cheby_upper = 1.1;
  o->aff_scaling = LIGMG_AFF_MAX;
  o->elim_factor = 0.9;
  o->elim_iters = 1;
  o->neighbor_of_neighbor = 0;
  o->seed = 0;
  o->cheby_eig_iters = 10;
  o->precision = LIGMG_PREC_FP64;
  o->use_graphs = 1;
  o->spmv_slice = 0;
  o->ordering = LIGMG_ORDER_NATURAL;
  o->solve_layout = 1;
  o->tail_nnz = 4096;
  o->rejection_ratio = -1;
  o->rejection_to_seed = 0;
  o->spmm_tune = 1;
}

int oracle_create(int64_t n, const int32_t *rp, const int32_t *ci, const double *v,
                  const ligmg_options *opts, int32_t ncomp, const int64_t *comp_ptr,
                  oracle_handle *out) {
  try {
    auto *S = new oracle_solver();
    if (opts) S->opts = *opts;
    else oracle_default_options(&S->opts);
    if (!check_opts(&S->opts)) {
      delete S;
      return LIGMG_ERR_UNSUPPORTED;
    }
    if (ncomp > 0 && comp_ptr) S->comp_ptr.assign(comp_ptr, comp_ptr + ncomp + 1);
    else S->comp_ptr = {0, n};
    auto t0 = std::chrono::steady_clock::now();
    Csr A = make_csr(n, rp, ci, v);
    if (S->opts.ordering == LIGMG_ORDER_DEGREE) {
      S->perm = degree_order(A, S->comp_ptr);
      A = permute_csr(A, S->perm);
    }
    make_hierarchy(*S, std::move(A));
    S->setup_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    *out = S;
    return 0;
  } catch (const CoarseTooLarge &) {
    return LIGMG_ERR_UNSUPPORTED;
  } catch (...) {
    return LIGMG_ERR_NUMERIC;
  }
}

int oracle_destroy(oracle_handle h) {
  delete h;
  return 0;
}

int oracle_solve(oracle_handle h, const double *b, double *x, ligmg_info *info, double *hist,
                 int32_t cap) {
  try {
    if (h->perm.empty()) return solve(*h, b, x, info, hist, cap);
    const int64_t n = (int64_t)h->perm.size();
    std::vector<double> bp(n), xp(n);
    for (int64_t i = 0; i < n; i++) bp[i] = b[h->perm[i]];
    int rc = solve(*h, bp.data(), xp.data(), info, hist, cap);
    for (int64_t i = 0; i < n; i++) x[h->perm[i]] = xp[i];
    return rc;
  } catch (...) {
    return LIGMG_ERR_NUMERIC;
  }
}

int oracle_solve_from(oracle_handle h, int32_t level, const double *b, double *x, ligmg_info *info,
                      double *hist, int32_t cap) {
  if (level < 0 || level >= (int)h->H.size()) return LIGMG_ERR_INVALID;
  if (level == 0) return oracle_solve(h, b, x, info, hist, cap);
  try {
    return solve(*h, b, x, info, hist, cap, (size_t)level);
  } catch (...) {
    return LIGMG_ERR_NUMERIC;
  }
}

int oracle_ordering(oracle_handle h, int32_t *perm) {
  const int64_t n = h->H[0].A.n;
  for (int64_t i = 0; i < n; i++) perm[i] = h->perm.empty() ? (int32_t)i : h->perm[i];
  return 0;
}

int oracle_num_levels(oracle_handle h, int32_t *out) {
  *out = (int32_t)h->H.size();
  return 0;
}

int oracle_level_info(oracle_handle h, int32_t l, int32_t *kind, int64_t *n, int64_t *nnz,
                      int64_t *nc) {
  if (l < 0 || l >= (int)h->H.size()) return LIGMG_ERR_INVALID;
  const Level &L = h->H[l];
  *kind = L.kind;
  *n = L.A.n;
  *nnz = L.A.nnz();
  *nc = L.n_coarse;
  return 0;
}

int oracle_level_matrix(oracle_handle h, int32_t l, int32_t *rp, int32_t *ci, double *v) {
  if (l < 0 || l >= (int)h->H.size()) return LIGMG_ERR_INVALID;
  const Csr &A = h->H[l].A;
  std::copy(A.rp.begin(), A.rp.end(), rp);
  std::copy(A.ci.begin(), A.ci.end(), ci);
  std::copy(A.v.begin(), A.v.end(), v);
  return 0;
}

int oracle_level_map(oracle_handle h, int32_t l, int32_t *map) {
  if (l < 0 || l >= (int)h->H.size()) return LIGMG_ERR_INVALID;
  const Level &L = h->H[l];
  if (L.kind == LIGMG_LEVEL_ELIM) std::copy(L.new_id.begin(), L.new_id.end(), map);
  else if (L.kind == LIGMG_LEVEL_AGG) std::copy(L.agg.begin(), L.agg.end(), map);
  else return LIGMG_ERR_INVALID;
  return 0;
}

int oracle_level_states(oracle_handle h, int32_t l, int32_t *ind, int32_t *state) {
  if (l < 0 || l >= (int)h->H.size() || h->H[l].kind != LIGMG_LEVEL_AGG) return LIGMG_ERR_INVALID;
  std::copy(h->H[l].ind.begin(), h->H[l].ind.end(), ind);
  std::copy(h->H[l].state.begin(), h->H[l].state.end(), state);
  return 0;
}

int oracle_level_strength(oracle_handle h, int32_t l, double *s) {
  if (l < 0 || l >= (int)h->H.size() || h->H[l].kind != LIGMG_LEVEL_AGG) return LIGMG_ERR_INVALID;
  std::copy(h->H[l].strength.begin(), h->H[l].strength.end(), s);
  return 0;
}

int oracle_level_cheby(oracle_handle h, int32_t l, double *lo, double *hi) {
  if (l < 0 || l >= (int)h->H.size() || h->H[l].kind != LIGMG_LEVEL_AGG) return LIGMG_ERR_INVALID;
  *lo = h->H[l].cheb_lo;
  *hi = h->H[l].cheb_hi;
  return 0;
}

int oracle_level_pinv(oracle_handle h, int32_t l, double *out) {
  if (l < 0 || l >= (int)h->H.size() || h->H[l].kind != LIGMG_LEVEL_COARSE)
    return LIGMG_ERR_INVALID;
  std::copy(h->H[l].pinv.begin(), h->H[l].pinv.end(), out);
  return 0;
}

int oracle_apply_cycle(oracle_handle h, int32_t l, const double *r, double *z) {
  if (l < 0 || l >= (int)h->H.size()) return LIGMG_ERR_INVALID;
  const int64_t n = h->H[l].A.n;
  std::vector<double> rr(r, r + n), x(n, 0.0);
  cycle(*h, l, x, rr);
  std::copy(x.begin(), x.end(), z);
  return 0;
}

int oracle_degrees(int64_t n, const int32_t *rp, const int32_t *ci, const double *v, int32_t *deg) {
  auto d = degrees(make_csr(n, rp, ci, v));
  std::copy(d.begin(), d.end(), deg);
  return 0;
}

int oracle_elim_select(int64_t n, const int32_t *rp, const int32_t *ci, const double *v,
                       int32_t max_degree, int32_t *flag, int64_t *n_kept) {
  auto f = elim_select(make_csr(n, rp, ci, v), max_degree, *n_kept);
  std::copy(f.begin(), f.end(), flag);
  return 0;
}

int oracle_tv_init(int64_t n, uint64_t seed, int32_t level, int32_t T, double *X) {
  tv_init(n, seed, level, T, X);
  return 0;
}

int oracle_strength(int64_t n, const int32_t *rp, const int32_t *ci, const double *v, double *X,
                    int32_t T, int32_t sweeps, int32_t metric, int32_t scaling, double thresh,
                    double *S, double *col_sum) {
  try {
    std::vector<double> s, cs;
    strength(make_csr(n, rp, ci, v), X, T, sweeps, metric, scaling, thresh, s, cs);
    std::copy(s.begin(), s.end(), S);
    std::copy(cs.begin(), cs.end(), col_sum);
    return 0;
  } catch (...) {
    return LIGMG_ERR_INVALID;
  }
}

int oracle_aggregate_with_strength(int64_t n, const int32_t *rp, const int32_t *ci,
                                   const double *S, const ligmg_options *opts, int32_t *agg,
                                   int32_t *ind, int32_t *state, int64_t *n_coarse) {
  if (opts->rejection_ratio >= 0) return LIGMG_ERR_UNSUPPORTED;  // needs the Laplacian's values
  try {
    Csr A = make_csr(n, rp, ci, S);
    std::vector<double> s(S + rp[0], S + rp[n]), cs(n);
    std::vector<double> buf;
    for (int64_t c = 0; c < n; c++) {
      int32_t len = A.rp[c + 1] - A.rp[c];
      buf.assign(len, 0.0);
      for (int32_t k = A.rp[c]; k < A.rp[c + 1]; k++) {
        int32_t r = A.ci[k];
        const int32_t *b = A.ci.data() + A.rp[r], *e = A.ci.data() + A.rp[r + 1];
        const int32_t *p = std::lower_bound(b, e, (int32_t)c);
        if (p == e || *p != c) return LIGMG_ERR_INVALID;
        buf[k - A.rp[c]] = s[p - A.ci.data()];
      }
      cs[c] = seq_sum(buf.data(), len);
    }
    std::vector<int32_t> a, i2, st;
    *n_coarse = aggregate(A, s, cs, *opts, a, i2, st);
    std::copy(a.begin(), a.end(), agg);
    std::copy(i2.begin(), i2.end(), ind);
    std::copy(st.begin(), st.end(), state);
    return 0;
  } catch (...) {
    return LIGMG_ERR_INVALID;
  }
}

int oracle_cheby_apply(int64_t n, const int32_t *rp, const int32_t *ci, const double *v, double lo,
                       double hi, int32_t steps, double *x, const double *b) {
  Csr A = make_csr(n, rp, ci, v);
  std::vector<double> dinv(n);
  for (int64_t i = 0; i < n; i++) {
    double d = diag_of(A, i);
    dinv[i] = (d != 0 && !isolated(A, i)) ? 1.0 / d : 0.0;
  }
  cheby_apply(A, dinv, lo, hi, steps, x, b);
  return 0;
}

int oracle_eig_estimate(int64_t n, const int32_t *rp, const int32_t *ci, const double *v,
                        uint64_t seed, int32_t level, int32_t iters, double *lam) {
  Csr A = make_csr(n, rp, ci, v);
  std::vector<double> dinv(n);
  for (int64_t i = 0; i < n; i++) {
    double d = diag_of(A, i);
    dinv[i] = (d != 0 && !isolated(A, i)) ? 1.0 / d : 0.0;
  }
  *lam = eig_estimate(A, dinv, seed, level, iters);
  return 0;
}

// ---- graph preprocessing (cc.hpp:22-105, main.cpp:79-110) ---------------
// connected_components: cc = iota; repeat cc_i = max(cc_i, max over stored
// j of row i of cc_j) (the MaxSRing SpMV and the max with the previous
// round, cc.hpp:42-48) until nothing changes.
int oracle_components(int64_t n, const int32_t *rp, const int32_t *ci, int32_t *label, int64_t *ncomp) {
  std::vector<int32_t> cc(n), nx(n);
  std::iota(cc.begin(), cc.end(), 0);
  while (true) {
    for (int64_t i = 0; i < n; i++) {
      int32_t m = std::numeric_limits<int32_t>::min();
      for (int32_t k = rp[i]; k < rp[i + 1]; k++) m = std::max(m, cc[ci[k]]);
      nx[i] = std::max(m, cc[i]);
    }
    if (nx == cc) break;
    cc.swap(nx);
  }
  std::vector<char> seen(n, 0);
  int64_t k = 0;
  for (int64_t i = 0; i < n; i++) {
    label[i] = cc[i];
    if (!seen[cc[i]]) {
      seen[cc[i]] = 1;
      k++;
    }
  }
  *ncomp = k;
  return 0;
}

// largest_cc: counts per label, the first maximum in label order, vertices
// of that component renumbered in ascending order (cc.hpp:60-84); the kept
// entries are those with both ends in it (cc.hpp:90-99).
int oracle_largest_component(int64_t n, const int32_t *rp, const int32_t *ci, int32_t *new_id, int64_t *n_out,
                             int64_t *nnz_out) {
  std::vector<int32_t> cc(n);
  int64_t nc;
  oracle_components(n, rp, ci, cc.data(), &nc);
  std::vector<uint64_t> counts(n, 0);
  for (int64_t i = 0; i < n; i++) counts[cc[i]]++;
  const int64_t id = std::max_element(counts.begin(), counts.end()) - counts.begin();
  int32_t next = 0;
  for (int64_t i = 0; i < n; i++) new_id[i] = (cc[i] == id) ? next++ : -1;
  int64_t m = 0;
  for (int64_t i = 0; i < n; i++)
    if (new_id[i] >= 0)
      for (int32_t k = rp[i]; k < rp[i + 1]; k++) m += new_id[ci[k]] >= 0;
  *n_out = next;
  *nnz_out = m;
  return 0;
}

int oracle_extract(int64_t n, const int32_t *rp, const int32_t *ci, const double *v, const int32_t *new_id,
                   int64_t n_out, int32_t *orp, int32_t *oci, double *ov) {
  std::vector<int64_t> row_of(n_out, -1);
  for (int64_t i = 0; i < n; i++)
    if (new_id[i] >= 0) row_of[new_id[i]] = i;
  orp[0] = 0;
  int64_t o = 0;
  for (int64_t r = 0; r < n_out; r++) {
    const int64_t i = row_of[r];
    for (int32_t k = rp[i]; k < rp[i + 1]; k++)
      if (new_id[ci[k]] >= 0) {
        oci[o] = new_id[ci[k]];
        if (ov) ov[o] = v[k];
        o++;
      }
    orp[r + 1] = (int32_t)o;
  }
  return 0;
}

// L = D - A: D the row sums of A (Reduce(Row): every stored entry in
// ascending column order from 0.0), the
// union of A's pattern and the diagonal; EWiseApply gives d - a on the
// diagonal and 0 - a elsewhere (main.cpp:82-104).
int oracle_adjacency_to_laplacian(int64_t n, const int32_t *rp, const int32_t *ci, const double *v,
                                  int32_t *orp, int32_t *oci, double *ov) {
  orp[0] = 0;
  int64_t o = 0;
  for (int64_t i = 0; i < n; i++) {
    const double d = seq_sum(v + rp[i], rp[i + 1] - rp[i]);
    bool done = false;
    for (int32_t k = rp[i]; k < rp[i + 1]; k++) {
      if (!done && ci[k] > i) {
        oci[o] = (int32_t)i;
        ov[o++] = d;
        done = true;
      }
      if (ci[k] == i) {
        oci[o] = ci[k];
        ov[o++] = d - v[k];
        done = true;
      } else {
        oci[o] = ci[k];
        ov[o++] = 0.0 - v[k];
      }
    }
    if (!done) {
      oci[o] = (int32_t)i;
      ov[o++] = d;
    }
    orp[i + 1] = (int32_t)o;
  }
  return 0;
}

int oracle_pinv(int64_t n, const int32_t *rp, const int32_t *ci, const double *v, double *dense) {
  try {
    auto p = dense_pinv(make_csr(n, rp, ci, v));
    std::copy(p.begin(), p.end(), dense);
    return 0;
  } catch (...) {
    return LIGMG_ERR_NUMERIC;
  }
}

}  // extern "C"
