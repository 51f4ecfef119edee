/*
 * ligmg_amd.h - C ABI of the MI355X graph-Laplacian AMG solver.
 *
 * This is the drop-in boundary for LigMG's setup + solve path.  The
 * reference exposes it as a C++ template class and a PETSc PC:
 *
 *   ligmg::Solver<T,F,DER>(A, opts)      include/ligmg/ligmg.hpp:17-28
 *   Solver::solve(b) / solve(b, from)    include/ligmg/ligmg.hpp:31-38
 *   ConvergenceInfo<T,F>                 include/ligmg/util.hpp:143-170
 *   Options / default_options()          include/ligmg/options.hpp:80-136
 *   PCCreate_LIGMG / PCSetUp / PCApply   include/ligmg/petsc_shim.hpp:18-74
 *
 * Here a solver is an opaque handle built from one CSR matrix (int32 row
 * pointers and column indices, double values, sorted columns, diagonal
 * stored) that may live in host or device memory.  All entry points return
 * an int status (0 = LIGMG_OK) and never throw; ligmg_last_error() gives a
 * thread-local message for the last failure.  A handle owns all device
 * memory of its hierarchy; calls on one handle must be serialised.
 */
#ifndef LIGMG_AMD_H
#define LIGMG_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
enum {
  LIGMG_OK = 0,
  LIGMG_ERR_INVALID = 1,     /* bad argument / not a Laplacian-shaped CSR   */
  LIGMG_ERR_UNSUPPORTED = 2, /* option value not implemented                */
  LIGMG_ERR_DEVICE = 3,      /* HIP runtime failure                          */
  LIGMG_ERR_NOMEM = 4,       /* allocation failed                            */
  LIGMG_ERR_NUMERIC = 5,     /* NaN/Inf in input or a singular coarse level  */
  LIGMG_ERR_IO = 6           /* file missing, unreadable or malformed        */
};

/* Krylov methods (the KSPType values parse_options.hpp:373-407 accepts).
 * FCG is flexible CG with one previous direction (Notay's FCG(1)).
 * PIPEFCG, PETSc's pipelined FCG, produces the same iterates up to rounding
 * (it reorders the reductions to overlap them with the preconditioner) and
 * runs as FCG here.  FGMRES is restarted flexible GMRES with PETSc's defaults
 * (restart 30, classical Gram-Schmidt, right preconditioning, unpreconditioned
 * residual norm), for the outer solve of ligmg_solve; K-cycle inner solves
 * and batched solves use CG / FCG / PIPEFCG. */
enum { LIGMG_KSP_CG = 0, LIGMG_KSP_FCG = 1, LIGMG_KSP_PIPEFCG = 2, LIGMG_KSP_FGMRES = 3 };
/* options.hpp:10-12 */
enum { LIGMG_SMOOTHER_JACOBI = 0, LIGMG_SMOOTHER_CHEBYSHEV = 1 };
enum {
  LIGMG_AFF_NONE = 0, LIGMG_AFF_ROW = 1, LIGMG_AFF_COLUMN = 2,
  LIGMG_AFF_ROWCOLUMN = 3, LIGMG_AFF_SYM = 4, LIGMG_AFF_MAX = 5
};
/* options.hpp:61 */
enum { LIGMG_STRENGTH_AFFINITY = 0, LIGMG_STRENGTH_ALGEBRAIC_DISTANCE = 1,
       LIGMG_STRENGTH_NONE = 2 };
/* arithmetic of the preconditioner (the outer Krylov loop is always FP64) */
enum { LIGMG_PREC_FP64 = 0, LIGMG_PREC_MIXED = 1 };
/* where the caller's arrays live */
enum { LIGMG_MEM_HOST = 0, LIGMG_MEM_DEVICE = 1 };
/* level kinds (levels.hpp:8-28 name(): "elim", "agg", "coarse") */
enum { LIGMG_LEVEL_ELIM = 0, LIGMG_LEVEL_AGG = 1, LIGMG_LEVEL_COARSE = 2 };
/* solve status (the reference ignores the KSP converged reason, cg.hpp:98) */
enum {
  LIGMG_CONVERGED = 0,
  LIGMG_DIVERGED_ITS = 1,
  LIGMG_DIVERGED_BREAKDOWN = 2,
  LIGMG_DIVERGED_NAN = 3
};

/* Vertex order of the hierarchy (new).  The reference randomly permutes
 * the input when it runs on more than one rank, for load balance
 * (Options::rand_perm, options.hpp:119; main.cpp:52-63), and keeps the
 * input's order on one rank: NATURAL, the default, does the same.  DEGREE
 * renumbers vertices by descending stored-entry count (ties by ascending
 * index) inside each connected block of comp_ptr, so the most-gathered
 * entries of every vector share a few MiB of L2 (faster SpMVs); it also
 * changes the aggregation's tie-breaks and therefore the hierarchy and
 * the iteration count.  Right-hand sides and solutions stay in the
 * caller's order; level introspection reports the internal order, and
 * ligmg_ordering returns the permutation. */
enum { LIGMG_ORDER_NATURAL = 0, LIGMG_ORDER_DEGREE = 1 };

/* Mirrors ligmg::Options (options.hpp:80-134) field names and defaults for
 * every field that affects setup or solve.  Fields marked (new) have no
 * counterpart in the reference. */
typedef struct ligmg_options {
  int32_t verbose;              /* false                                     */
  int32_t always_elim;          /* true: try an elimination level each time  */
  int32_t top_elim;             /* true: apply level-0 elimination once      */
  int32_t constant_correction;  /* false                                     */
  int32_t strength_metric;      /* LIGMG_STRENGTH_AFFINITY                   */
  int32_t max_levels;           /* 50                                        */
  int64_t min_size;             /* 4000 (nnz bound of the last agg level)    */
  int32_t elim_degree;          /* 4                                         */
  double tol;                   /* 1e-8                                      */
  double thresh;                /* 1e-8 strength prune threshold             */
  int32_t agg_iters;            /* 10                                        */
  int32_t iters;                /* 500                                       */
  int32_t pre_smooth;           /* 2                                         */
  int32_t post_smooth;          /* 2                                         */
  int32_t num_tvs;              /* 10 (at most 16 here)                      */
  int32_t tvs_smooth;           /* 4                                         */
  int32_t smoother_type;        /* LIGMG_SMOOTHER_CHEBYSHEV                  */
  int32_t kcycle;               /* false                                     */
  int32_t outer_ksp_type;       /* LIGMG_KSP_CG                              */
  int32_t kcycle_ksp_type;      /* LIGMG_KSP_FCG                             */
  int32_t kcycle_iters;         /* 2                                         */
  int64_t votes_to_seed;        /* 1                                         */
  int64_t max_agg_size;         /* -1: no aggregate-size rejection           */
  double agg_iter_factor;       /* 0.5                                       */
  double agg_init_factor;       /* 0.5                                       */
  double cheby_lower;           /* 0.3                                       */
  double cheby_upper;           /* 1.1                                       */
  int32_t aff_scaling;          /* LIGMG_AFF_MAX                             */
  double elim_factor;           /* 0.9                                       */
  int32_t elim_iters;           /* 1 (only 1 supported)                      */
  int32_t neighbor_of_neighbor; /* false                                     */
  /* (new) */
  uint64_t seed;                /* 0: seed of test vectors / eigen estimate  */
  int32_t cheby_eig_iters;      /* 10: CG-Lanczos steps for lambda_max       */
  int32_t precision;            /* LIGMG_PREC_FP64                           */
  int32_t use_graphs;           /* 1: replay cycles from captured hipGraphs  */
  int32_t spmv_slice;           /* 0: no column slicing of the solve SpMVs;
                                   >0: slice width (cols); -1: autotune     */
  int32_t ordering;             /* LIGMG_ORDER_NATURAL: vertex order used for
                                   setup and solve (see enum)               */
  int32_t solve_layout;         /* 1: the solve-phase copy of every large
                                   level is stored in degree order (gather
                                   locality; the hierarchy is unchanged);
                                   0: setup order                           */
  int32_t tail_nnz;             /* 4096: the coarse levels whose operators
                                   hold at most this many stored entries
                                   (with every coarser level) run as one
                                   single-workgroup launch per coarse-grid
                                   correction, same results; 0: off         */
  double rejection_ratio;       /* -1: conductance rejection off; >= 0: in the
                                   second half of the voting rounds, joined
                                   vertices whose external / internal edge
                                   weight exceeds it leave their aggregate  */
  int32_t rejection_to_seed;    /* false: a rejected vertex becomes undecided;
                                   true: a seed                             */
  int32_t spmm_tune;            /* 1: batched solves time the SpMM tiles
                                   (right-hand sides per lane, gathers in
                                   flight) of every level operator at the
                                   first solve of each batch width and keep
                                   the fastest; 0: the fixed tile.  Results
                                   are the same bits either way. (new)      */
} ligmg_options;

/* ConvergenceInfo (util.hpp:143-170) plus the fields the reference only
 * prints (full-system residual, multigrid.hpp:357-366) and status. */
typedef struct ligmg_info {
  int32_t iters;
  int32_t status;              /* LIGMG_CONVERGED / ...                        */
  int32_t history_len;         /* entries written to the history buffer        */
  int32_t base_level;          /* 1 when the level-0 elimination was applied   */
  double rel_residual;         /* last / first of the reduced-system history   */
  double full_rel_residual;    /* ||b - L x|| / ||b|| on the caller's system   */
  double time;                 /* solve wall time (s)                          */
  double work;                 /* iters * (1 + cycle complexity)               */
  double cycle_complexity;     /* work / iters                                 */
  double cf, ecf, wda;
  double setup_time;           /* s, from the handle                           */
} ligmg_info;

typedef struct ligmg_solver *ligmg_handle;

/* default_options() (options.hpp:136) */
void ligmg_options_default(ligmg_options *opts);

/* Solver(A, opts): builds the hierarchy on the device (ligmg.hpp:17-28 ->
 * make_hierarchy, multigrid.hpp:161-263).  opts may be NULL (defaults).
 * ncomp/comp_ptr (may be 0/NULL) name contiguous vertex blocks that are
 * independent graphs: Krylov inner products and the constant-nullspace
 * projection are then taken per block (batched solves). */
int ligmg_create(int64_t n, const int32_t *rowptr, const int32_t *colind,
                 const double *vals, int mem_kind, const ligmg_options *opts,
                 int32_t ncomp, const int64_t *comp_ptr, ligmg_handle *out);

/* Solver::solve(b) (ligmg.hpp:31-33 -> ligmg::solve, multigrid.hpp:315-370).
 * b and x have length n in mem_kind memory.  history (host, may be NULL)
 * receives up to history_cap reduced-system residual norms (incl. r0). */
int ligmg_solve(ligmg_handle h, const double *b, double *x, int mem_kind,
                ligmg_info *info, double *history, int32_t history_cap);

/* Solver::solve(b, from) (ligmg.hpp:36-38 -> ligmg::solve, multigrid.hpp:
 * 315-370 with from = level): solves the system of hierarchy level `level`.
 * b and x have that level's n entries (ligmg_level_info) in mem_kind memory,
 * numbered as ligmg_level_matrix reports the level (the caller's order for
 * level 0, where this is ligmg_solve).  When the level is an elimination
 * level and top_elim is set, it is applied once around the outer Krylov
 * solve on level + 1; info->full_rel_residual is taken on the level's own
 * matrix and info->base_level names the outer solve's level.  The cycle
 * complexity divides by the nnz of level 0, as the reference's does. */
int ligmg_solve_from(ligmg_handle h, int32_t level, const double *b, double *x,
                     int mem_kind, ligmg_info *info, double *history,
                     int32_t history_cap);

/* k right-hand sides at once (column-major, leading dimension n), one
 * hierarchy traversal per cycle with SpMM kernels; every (block, column)
 * pair runs its own flexible-CG coefficients and convergence test.
 * info receives the worst column; iters_out (host, may be NULL) the
 * per-column iteration counts. */
int ligmg_solve_batch(ligmg_handle h, int32_t k, const double *B, double *X,
                      int mem_kind, ligmg_info *info, int32_t *iters_out);

/* Edge-weight gradient of a loss of x = L(w)^+ b given the adjoint
 * lam = L^+ (dloss/dx): g_e = -(lam_i - lam_j)(x_i - x_j) for the m edges
 * (ei[e], ej[e]).  All arrays are device pointers.  No counterpart in the
 * reference (which has no differentiation); used by the torch autograd
 * binding. */
int ligmg_edge_grad(int64_t m, const int32_t *ei, const int32_t *ej,
                    const double *x, const double *lam, double *g);

/* Laplacian values on a fixed CSR pattern from edge weights: vals[k] =
 * -w[eid[k]] off the diagonal (eid[k] = edge of stored entry k, -1 on the
 * diagonal) and the diagonal = sum of the row's weights in column order.
 * Device pointers.  Rebuilds L(w) for the autograd binding. */
int ligmg_assemble_laplacian(int64_t n, const int32_t *rowptr,
                             const int32_t *colind, const int32_t *eid,
                             const double *w, double *vals);

/* Solver destructor / PCDestroy_LIGMG (petsc_shim.hpp:31-38). */
int ligmg_destroy(ligmg_handle h);

/* The internal vertex order (new; see LIGMG_ORDER_*): perm (host, length n)
 * receives, for every internal position, the caller's vertex index. */
int ligmg_ordering(ligmg_handle h, int32_t *perm);

/* (new) Device memory the library keeps for reuse: temporaries of setup and
 * of freed solvers are cached by size class (at most 32 GiB) instead of
 * returned to HIP, which makes repeated setups in one process as fast as
 * the first.  This returns all of it to HIP; the library never needs it.
 * Returns LIGMG_ERR_INVALID, releasing nothing, while another thread
 * captures a handle's preconditioner graph (a device synchronisation would
 * end that capture). */
int ligmg_trim_memory(void);

/* Thread-local message of the last failed call ("" if none). */
const char *ligmg_last_error(void);

/* ---- hierarchy introspection (print_hierarchy, multigrid.hpp:265-292;
 *      dump_hierarchy, ligmg.hpp:47-52) ---------------------------------- */
int ligmg_num_levels(ligmg_handle h, int32_t *out);
/* dump_hierarchy (ligmg.hpp:47-52): creates directory dir and writes each
 * level in PETSc binary format, in setup order, as the reference's dump
 * methods name them: <dir>/<level>_A.petsc for every level; P, R, Q for an
 * elimination level (elimination.hpp:62-67); P, R and strength for an
 * aggregation level (aggregation.hpp:254-272), whose AggState history the
 * reference writes per voting round and this library as the final
 * agg_ind.petsc / agg_state.petsc vectors.  blocks.petsc holds comp_ptr and
 * ordering.petsc the input order (LIGMG_ORDER_DEGREE). */
int ligmg_dump_hierarchy(ligmg_handle h, const char *dir);
/* A solver from a dump (new; the reference only writes dumps): the levels
 * are read back instead of constructed, then the solve-phase set-up runs as
 * in ligmg_create (smoothers, eigen-estimates, FP32 copies, layout), so
 * solves match those of the dumped solver for the same solve options.
 * Returns LIGMG_ERR_IO for missing or malformed files. */
int ligmg_load_hierarchy(const char *dir, const ligmg_options *opts, ligmg_handle *out);
/* dump_csv (ligmg.hpp:42-45): the header line of print_csv_header
 * (util.hpp:187-193, options.hpp:139-147) and the result line of print_csv
 * (util.hpp:195-205, options.hpp:149-164) for this solver and the solve
 * described by info, NUL-terminated in buf.  graph names the graph (the
 * reference prints the input file name without directory and extension);
 * options without a field here are printed with the reference's defaults. */
int ligmg_dump_csv(ligmg_handle h, const char *graph, const ligmg_info *info, char *buf, int64_t cap);
/* First level of the fused coarse tail (new; see tail_nnz): levels from
 * *out down run in one launch per coarse-grid correction; *out equals the
 * number of levels when nothing is fused. */
int ligmg_tail_start(ligmg_handle h, int32_t *out);
int ligmg_setup_time(ligmg_handle h, double *out);
int ligmg_level_info(ligmg_handle h, int32_t level, int32_t *kind, int64_t *n,
                     int64_t *nnz, int64_t *n_coarse);
/* matrix of a level copied to host arrays sized from ligmg_level_info */
int ligmg_level_matrix(ligmg_handle h, int32_t level, int32_t *rowptr,
                       int32_t *colind, double *vals);
/* elimination level: new id or -1 per vertex; aggregation level: coarse id
 * per vertex.  Host array of length n. */
int ligmg_level_map(ligmg_handle h, int32_t level, int32_t *map);
/* aggregation level: final AggState per vertex (ind, state), host arrays. */
int ligmg_level_states(ligmg_handle h, int32_t level, int32_t *ind,
                       int32_t *state);
/* aggregation level: strength value per stored entry of the level matrix
 * (0 where pruned or diagonal), host array of length nnz. */
int ligmg_level_strength(ligmg_handle h, int32_t level, double *s);
/* Chebyshev interval [lo, hi] of a level's smoother */
int ligmg_level_cheby(ligmg_handle h, int32_t level, double *lo, double *hi);

/* Times `iters` back-to-back launches of level `level`'s SpMV y = A x
 * (precision LIGMG_PREC_FP64: FP64 values and vectors; LIGMG_PREC_MIXED:
 * FP32 copies) on the solver's stream with HIP events.  Returns the average
 * milliseconds per launch and the algorithmic bytes of one launch,
 * nnz*(s_v + 4) + (n+1)*4 + 2*n*s_x.  Stands in for the reference's SpMV
 * micro-benchmark (src/bench_spmv.cpp:1-50). */
int ligmg_bench_spmv(ligmg_handle h, int32_t level, int32_t precision,
                     int32_t iters, double *avg_ms, int64_t *bytes);

/* (new) Times `iters` launches of one fused Chebyshev smoothing step on
 * aggregation level `level`, x_new = x + a (x - x_prev) + b D^-1 (r - A x),
 * in the given precision (the dominant kernel of a V-cycle).  Algorithmic
 * bytes of one launch: nnz*(s_v + 4) + (n+1)*4 + 5*n*s_x, plus n*8 for the
 * FP64 diagonal of the FP32 Laplacian form. */
int ligmg_bench_smoother(ligmg_handle h, int32_t level, int32_t precision,
                         int32_t iters, double *avg_ms, int64_t *bytes);

/* (new) ligmg_bench_spmv for a block of k <= 64 right-hand sides (k rounded
 * up to a power of two, kp; row-major n x kp blocks, the batched solves'
 * SpMM with its tuned tile).  Algorithmic bytes of one launch:
 * nnz*(s_v + 4) + (n+1)*4 + 2*n*kp*s_x. */
int ligmg_bench_spmm(ligmg_handle h, int32_t level, int32_t precision,
                     int32_t k, int32_t iters, double *avg_ms, int64_t *bytes);

/* (new) Times `iters` launches of the coarsest level's dense pseudo-inverse
 * apply to a block of k <= 64 right-hand sides (k rounded up to a power of
 * two, kp; FP64, on the matrix cores when kp >= 16, a matrix-vector product
 * when kp = 1) on the solver's stream.  Returns the average milliseconds
 * per launch and its floating-point operations, 2 * n_c^2 * kp. */
int ligmg_bench_coarse(ligmg_handle h, int32_t k, int32_t iters,
                       double *avg_ms, double *flops);

/* (new) Diagnostic: builds the SpMV work units of the row pointers rp
 * (host array of n + 1 non-decreasing offsets from 0) with the device
 * builder and with the host loop, and sets *same to 1 when every table
 * (unit rows, row counts, offsets, long rows, partial slots) agrees, else
 * 0.  Also reports the unit count. */
int ligmg_check_units(int64_t n, const int32_t *rp, int32_t *same, int64_t *nunits);

/* (new) Re-factors the coarsest level's dense pseudo-inverse as setup does
 * (on the device from 256 vertices: blocked Gauss-Jordan on the matrix
 * cores) and reports the wall milliseconds of the whole step (host matrix
 * assembly, upload, inversion), the milliseconds of the device inversion
 * kernels (HIP events; 0 on the host path) and the floating-point
 * operations of inverting an n_c x n_c matrix, 2 * n_c^3 (the padding to a
 * multiple of the block width is not counted; 0 on the host path). */
int ligmg_bench_pinv(ligmg_handle h, double *total_ms, double *gj_ms, double *flops);

/* (new) Diagnostic: Y = P X with the coarsest level's dense pseudo-inverse
 * P (n_c x n_c) for a block of k <= 64 right-hand sides, by the batched
 * solves' kernel (k rounded up to a power of two; the matrix cores from 16
 * columns).  X and Y are host arrays, n_c x k, row-major. */
int ligmg_coarse_apply_block(ligmg_handle h, int32_t k, const double *X, double *Y);

/* (new) Diagnostic: the dense pseudo-inverse that setup forms for a
 * coarsest level, here for any graph Laplacian given as CSR in host memory
 * (n <= 8192 vertices; symmetric, every column in [0, n)): the device
 * blocked Gauss-Jordan from 256 vertices, the host Cholesky below.  Writes
 * the n x n result row-major to the host array pinv. */
int ligmg_dense_pinv(int64_t n, const int32_t *rp, const int32_t *ci, const double *v, double *pinv);

/* z = M r: one cycle of the preconditioner from level `level` down (the
 * V- or K-cycle of multigrid.hpp:86-130 with a zero initial guess) in the
 * given precision.  r and z are host arrays of length n in FP64.  Mirrors
 * the oracle's apply_cycle for parity tests. */
int ligmg_apply_cycle(ligmg_handle h, int32_t level, int32_t precision,
                      const double *r, double *z);

/* y = A x with level `level`'s solve-phase operator (the same kernels and
 * layout the cycles use: column slices, FP32 Laplacian form for
 * LIGMG_PREC_MIXED).  x and y are host arrays of length n in FP64. */
int ligmg_level_spmv(ligmg_handle h, int32_t level, int32_t precision,
                     const double *x, double *y);

/* ---- the "ligmg" PETSc preconditioner without PETSc ---------------------
 * PCCreate_LIGMG / PCSetUp_LIGMG / PCApply_LIGMG / PCDestroy_LIGMG
 * (petsc_shim.hpp:18-74) as plain calls; the PETSc shim
 * (csrc/petsc/ligmg_pc.cpp) is a thin layer over them.  Set-up: begin(n),
 * then every row of the preconditioning matrix in order 0..n-1 (columns in
 * any order, no duplicates; host arrays), then setup(), which builds the
 * solver with default options as the reference's PCSetUp does (an earlier
 * solver is destroyed).  apply() is one full solve to the default tolerance
 * (PCApply, :40-62), b and x host arrays of length n; info may be NULL. */
typedef struct ligmg_pc *ligmg_pc_handle;
int ligmg_pc_create(ligmg_pc_handle *out);
int ligmg_pc_begin(ligmg_pc_handle pc, int64_t n);
int ligmg_pc_add_row(ligmg_pc_handle pc, int64_t row, int64_t ncols,
                     const int64_t *cols, const double *vals);
int ligmg_pc_setup(ligmg_pc_handle pc);
int ligmg_pc_apply(ligmg_pc_handle pc, const double *b, double *x,
                   ligmg_info *info);
int ligmg_pc_destroy(ligmg_pc_handle pc);

/* ---- graph preprocessing on the device (main.cpp:79-110, cc.hpp) ------ */
/* All arrays live in mem_kind memory (LIGMG_MEM_HOST or LIGMG_MEM_DEVICE);
 * the CSR has sorted columns; outputs are caller-allocated.               */

/* connected_components (cc.hpp:22-51): label[i] = the largest vertex index
 * of i's component (the fixed point of the reference's max-label SpMV).
 * The pattern is taken as undirected (explicit zeros count as edges, as in
 * the reference's SpMV).  *ncomp receives the number of components. */
int ligmg_components(int64_t n, const int32_t *rowptr, const int32_t *colind,
                     int mem_kind, int32_t *label, int64_t *ncomp);

/* ligmg_components without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (serving loops,
 * harnesses that forbid hipMalloc).  rowptr, colind and label are device
 * pointers; n may be at most LIGMG_NOALLOC_CC_MAX_N.  The work space (three
 * words per vertex plus flags, 48 MiB) is a static array that each GPU the
 * library runs on holds until the process ends; every call on a GPU shares
 * it, so calls that overlap on one GPU (two host threads) race.  Work runs
 * on the calling thread's default stream (hipStreamPerThread) and is
 * finished when the call returns.  Same results as ligmg_components. */
#define LIGMG_NOALLOC_CC_MAX_N 4194304
int ligmg_components_noalloc(int64_t n, const int32_t *rowptr,
                             const int32_t *colind, int32_t *label,
                             int64_t *ncomp);

/* largest_cc (cc.hpp:55-105), first half: new_id[i] = rank of i among the
 * vertices of the largest component (ascending original order), -1
 * elsewhere; ties between equally large components go to the smallest
 * label.  *n_out / *nnz_out size the submatrix for ligmg_extract. */
int ligmg_largest_component(int64_t n, const int32_t *rowptr,
                            const int32_t *colind, int mem_kind,
                            int32_t *new_id, int64_t *n_out, int64_t *nnz_out);

/* largest_cc second half (cc.hpp:90-104): the entries whose row and
 * column both have new_id >= 0, renumbered (columns stay sorted). */
int ligmg_extract(int64_t n, const int32_t *rowptr, const int32_t *colind,
                  const double *vals, const int32_t *new_id, int64_t n_out,
                  int mem_kind, int32_t *out_rowptr, int32_t *out_colind,
                  double *out_vals);

/* Adjacency matrix -> Laplacian L = D - A (main.cpp:79-110): D holds the
 * row sums of A (stored entries, self loops included, summed in the
 * canonical order); the pattern is A's plus the diagonal.  The first call
 * returns the output nnz (nnz(A) + rows without a stored diagonal). */
int ligmg_laplacian_nnz(int64_t n, const int32_t *rowptr,
                        const int32_t *colind, int mem_kind, int64_t *nnz_out);
int ligmg_adjacency_to_laplacian(int64_t n, const int32_t *rowptr,
                                 const int32_t *colind, const double *vals,
                                 int mem_kind, int32_t *out_rowptr,
                                 int32_t *out_colind, double *out_vals);

/* ---- stage entry points for parity tests (device execution) ---------- */
/* Voting aggregation (aggregation.hpp:838-964) driven by caller-supplied
 * strength values on the pattern of a CSR (host arrays).  Writes the
 * coarse id per vertex and the final AggState; returns n_coarse. */
int ligmg_aggregate_with_strength(int64_t n, const int32_t *rowptr,
                                  const int32_t *colind, const double *s,
                                  const ligmg_options *opts, int32_t *agg,
                                  int32_t *ind, int32_t *state,
                                  int64_t *n_coarse);

#ifdef __cplusplus
}
#endif
#endif
