/*
 * ligmg_graphs.h - seeded synthetic graph-Laplacian generators (host C++).
 *
 * The reference reads MatrixMarket files and converts adjacency matrices to
 * Laplacians in its CLI driver (src/main.cpp:33-110) and takes the largest
 * connected component with label propagation (include/ligmg/cc.hpp:56-106).
 * These generators produce the benchmark inputs directly as Laplacian CSR
 * (sorted column indices, diagonal stored, int32 indices), deterministically
 * from a seed with a counter-based RNG, so every run and every consumer
 * (oracle, device solver, tests) sees identical bits.
 *
 * All returned arrays are malloc'ed by the library and released with
 * ligmg_graph_free().  Every function returns 0 on success, nonzero on error
 * (invalid arguments or out of memory).
 */
#ifndef LIGMG_GRAPHS_H
#define LIGMG_GRAPHS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Edge-weight distributions. */
enum {
  LIGMG_W_UNIT = 0,      /* w = 1                                  */
  LIGMG_W_UNIFORM = 1,   /* w ~ U[p0, p1]                          */
  LIGMG_W_LOGNORMAL = 2, /* w = exp(p0 + p1 * N(0,1))              */
  LIGMG_W_GAUSS_DIST = 3 /* RGG only: w = (float)exp(-d^2 / r^2)   */
};

typedef struct ligmg_graph {
  int64_t n;        /* vertices                                           */
  int64_t nnz;      /* stored entries of the Laplacian (incl. diagonal)   */
  int32_t *rowptr;  /* n+1                                                */
  int32_t *colind;  /* nnz, ascending within each row                     */
  double *vals;     /* nnz, L = D - W (generators) or the file's values   */
  int32_t ncomp;    /* number of blocks (components) laid out contiguously */
  int64_t *comp_ptr;/* ncomp+1 vertex offsets of the blocks               */
} ligmg_graph;

void ligmg_graph_free(ligmg_graph *g);

/* MatrixMarket coordinate file (real / integer / pattern, general /
 * symmetric) into CSR as stored: the matrix itself, not a Laplacian of it
 * (convert an adjacency matrix with ligmg_adjacency_to_laplacian).  The
 * stored triangle of a symmetric file is mirrored, duplicates are summed
 * in file order, columns are sorted; ncomp = 1.  Returns 3 when the file
 * cannot be read, 4 when it is not a square coordinate matrix. */
int ligmg_graph_read_mtx(const char *path, ligmg_graph *out);

/* 2-D 5-point grid nx*ny, row-major ids (id = y*nx + x). */
int ligmg_gen_grid2d(int64_t nx, int64_t ny, int wkind, double p0, double p1,
                     uint64_t seed, ligmg_graph *out);

/* 3-D 7-point grid nx*ny*nz, id = (z*ny + y)*nx + x. */
int ligmg_gen_grid3d(int64_t nx, int64_t ny, int64_t nz, int wkind, double p0,
                     double p1, uint64_t seed, ligmg_graph *out);

/* R-MAT scale/edge_factor with quadrant probabilities a,b,c (d = 1-a-b-c);
 * symmetrised, self loops and duplicates removed, largest connected
 * component kept (ids relabelled in ascending original order).  When
 * pendant_frac > 0, a path of uniform{1..pendant_max} new vertices is
 * attached to each LCC vertex selected with probability pendant_frac. */
int ligmg_gen_rmat(int scale, int edge_factor, double a, double b, double c,
                   int wkind, double p0, double p1, double pendant_frac,
                   int pendant_max, uint64_t seed, int nthreads,
                   ligmg_graph *out);

/* Barabasi-Albert preferential attachment, m edges per new vertex, seeded
 * with an (m+1)-clique.  Always connected. */
int ligmg_gen_ba(int64_t n, int m, int wkind, double p0, double p1,
                 uint64_t seed, ligmg_graph *out);

/* Random geometric graph in the unit square: n points, radius r, largest
 * connected component, weights by wkind (LIGMG_W_GAUSS_DIST for exp(-d^2/r^2)
 * rounded to float). */
int ligmg_gen_rgg(int64_t n, double radius, int wkind, double p0, double p1,
                  uint64_t seed, ligmg_graph *out);

/* Block-diagonal union of graphs (in order); comp_ptr marks the blocks. */
int ligmg_graph_union(const ligmg_graph *gs, int count, ligmg_graph *out);

/* Laplacian from an undirected edge list (u != v, duplicates summed),
 * no component extraction. */
int ligmg_graph_from_edges(int64_t n, int64_t m, const int32_t *u,
                           const int32_t *v, const double *w,
                           ligmg_graph *out);

/* k right-hand sides (column-major, ld = n): i.i.d. N(0,1) from the seed,
 * then the mean of each column over each block of comp_ptr removed
 * (comp_ptr == NULL: one block). */
int ligmg_gen_rhs(int64_t n, int k, uint64_t seed, int ncomp,
                  const int64_t *comp_ptr, double *out);

/* Counter-based uniform in [0,1) and standard normal used by the generators,
 * exported so that tests can reproduce any drawn value. */
double ligmg_rng_u01(uint64_t seed, uint64_t stream, uint64_t idx);
double ligmg_rng_normal(uint64_t seed, uint64_t stream, uint64_t idx);

#ifdef __cplusplus
}
#endif
#endif
