/*
 * oracle.h - CPU restatement of LigMG's setup + solve (test infrastructure).
 *
 * THIS IS THE CHECKER, NOT THE PRODUCT.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  It restates the reference's
 * algorithm (include/ligmg/{elimination,aggregation,smoothing,multigrid,cg,
 * util}.hpp) in plain single-threaded C++ with the summation orders fixed
 * (seq_sum: the reference's own one-rank order for col_sum, the vote tally
 * and the other CombBLAS Reduce sums; canon_sum: a fixed order for the sums
 * the reference leaves to CombBLAS/PETSc internals), so that the device
 * implementation can be
 * compared with it bit for bit on elimination sets, aggregates and coarse
 * matrices.
 *
 * Parity pins (tests/test_oracle_*.py): scipy/numpy closed forms for the
 * Galerkin products, Schur complement, strength formula, Chebyshev
 * polynomial, pseudo-inverse and Krylov solutions; brute force on tiny
 * graphs for the elimination selection; hand-derived outcomes for voting.
 * The reference itself cannot be built here (CombBLAS/PETSc/MPI absent), so
 * no function is pinned against reference binaries.
 */
#ifndef LIGMG_ORACLE_H
#define LIGMG_ORACLE_H
#include <stdint.h>

#include "../include/ligmg_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_solver *oracle_handle;

double oracle_canon_sum(const double *a, int64_t m);
/* left fold from 0.0 in index order: the reference's one-rank order for
 * col_sum, the vote tally and the other CombBLAS Reduce sums */
double oracle_seq_sum(const double *a, int64_t m);
/* test hook: 0 = the reference's order (default); 1 = plant the canonical
 * split in the sequential sums; 2 = plant a descending-voter tally */
void oracle_set_decision_order(int32_t mode);
/* threads of the OpenMP loops (nthreads <= 0: unchanged); returns the
 * count now in effect (1 without OpenMP).  Results do not depend on it. */
int32_t oracle_set_threads(int32_t nthreads);
void oracle_default_options(ligmg_options *opts);

int oracle_create(int64_t n, const int32_t *rowptr, const int32_t *colind,
                  const double *vals, const ligmg_options *opts, int32_t ncomp,
                  const int64_t *comp_ptr, oracle_handle *out);
int oracle_destroy(oracle_handle h);
int oracle_ordering(oracle_handle h, int32_t *perm);
int oracle_solve(oracle_handle h, const double *b, double *x, ligmg_info *info,
                 double *history, int32_t history_cap);
/* the system of level `level` (b, x in that level's numbering) */
int oracle_solve_from(oracle_handle h, int32_t level, const double *b,
                      double *x, ligmg_info *info, double *history,
                      int32_t history_cap);
int oracle_num_levels(oracle_handle h, int32_t *out);
int oracle_level_info(oracle_handle h, int32_t level, int32_t *kind,
                      int64_t *n, int64_t *nnz, int64_t *n_coarse);
int oracle_level_matrix(oracle_handle h, int32_t level, int32_t *rowptr,
                        int32_t *colind, double *vals);
int oracle_level_map(oracle_handle h, int32_t level, int32_t *map);
int oracle_level_states(oracle_handle h, int32_t level, int32_t *ind,
                        int32_t *state);
int oracle_level_strength(oracle_handle h, int32_t level, double *s);
int oracle_level_cheby(oracle_handle h, int32_t level, double *lo,
                       double *hi);
/* coarse level only: dense pseudo-inverse (n*n, row-major) */
int oracle_level_pinv(oracle_handle h, int32_t level, double *out);
/* apply one V-/K-cycle of the hierarchy from the base level to r: z = M r */
int oracle_apply_cycle(oracle_handle h, int32_t level, const double *r,
                       double *z);

/* ---- stage functions -------------------------------------------------- */
int oracle_degrees(int64_t n, const int32_t *rp, const int32_t *ci,
                   const double *v, int32_t *deg);
/* 1 if eliminated, 0 otherwise; returns number kept in *n_kept */
int oracle_elim_select(int64_t n, const int32_t *rp, const int32_t *ci,
                       const double *v, int32_t max_degree, int32_t *flag,
                       int64_t *n_kept);
/* test vectors of a level before smoothing, layout X[i*T + t] */
int oracle_tv_init(int64_t n, uint64_t seed, int32_t level, int32_t T,
                   double *X);
/* Jacobi-smooth X in place (tvs_smooth sweeps), then strength per stored
 * entry (0 when pruned) and col_sum. */
int oracle_strength(int64_t n, const int32_t *rp, const int32_t *ci,
                    const double *v, double *X, int32_t T, int32_t sweeps,
                    int32_t metric, int32_t scaling, double thresh,
                    double *S, double *col_sum);
int oracle_aggregate_with_strength(int64_t n, const int32_t *rp,
                                   const int32_t *ci, const double *S,
                                   const ligmg_options *opts, int32_t *agg,
                                   int32_t *ind, int32_t *state,
                                   int64_t *n_coarse);
/* Chebyshev smoothing (steps) of x for L x = b on interval [lo, hi] of
 * D^-1 L, in place */
int oracle_cheby_apply(int64_t n, const int32_t *rp, const int32_t *ci,
                       const double *v, double lo, double hi, int32_t steps,
                       double *x, const double *b);
/* lambda_max estimate of D^-1 L used by the smoother */
int oracle_eig_estimate(int64_t n, const int32_t *rp, const int32_t *ci,
                        const double *v, uint64_t seed, int32_t level,
                        int32_t iters, double *lam);
int oracle_components(int64_t n, const int32_t *rp, const int32_t *ci,
                      int32_t *label, int64_t *ncomp);
int oracle_largest_component(int64_t n, const int32_t *rp, const int32_t *ci,
                             int32_t *new_id, int64_t *n_out,
                             int64_t *nnz_out);
int oracle_extract(int64_t n, const int32_t *rp, const int32_t *ci,
                   const double *v, const int32_t *new_id, int64_t n_out,
                   int32_t *orp, int32_t *oci, double *ov);
int oracle_adjacency_to_laplacian(int64_t n, const int32_t *rp,
                                  const int32_t *ci, const double *v,
                                  int32_t *orp, int32_t *oci, double *ov);
int oracle_pinv(int64_t n, const int32_t *rp, const int32_t *ci,
                const double *v, double *dense);

#ifdef __cplusplus
}
#endif
#endif
