// Device-resident multigrid hierarchy.
//
// The reference keeps a vector of polymorphic levels holding CombBLAS
// matrices (levels.hpp:8-28, multigrid.hpp:33-35) with P, R, Q as explicit
// sparse matrices.  Here every level is a flat record of device arrays:
// CSR with int32 indices, the inverse diagonal, and implicit transfer
// operators (aggregate ids / member lists for aggregation, kept ids plus a
// short P row per eliminated vertex for elimination).
#pragma once
#include <memory>
#include <vector>

#include "../../include/ligmg_amd.h"
#include "common.hpp"

namespace lg {

// Work units of the row-block SpMV: a unit is either a run of complete rows
// holding at most UNIT_NNZ stored entries, or one UNIT_NNZ-chunk of a longer
// row whose partial sum is combined afterwards in chunk order.
constexpr int UNIT_NNZ = 2048;
constexpr int UNIT_ROWS = 512;
constexpr int SPMV_BLOCK = 256;
// units per SpMV workgroup: the colind/value loads of the next unit are in
// flight while the rows of the current one are reduced.  Operators of few
// units take one unit per workgroup, so that they spread over more CUs.
constexpr int SPMV_UPW = 2;
constexpr int64_t SPMV_UPW_MIN_UNITS = 32768;
inline int spmv_upw(int64_t nunits) { return nunits >= SPMV_UPW_MIN_UNITS ? SPMV_UPW : 1; }

// Lanes that reduce one row of a complete-row unit: the smallest power of
// two K <= 64 with items * K >= the unit's mean row length.  K = 1 sums a
// row left to right; otherwise lane g of the row's group sums positions
// g, g + K, ... left to right and the K partials meet in an xor butterfly
// (off = K/2..1).
__host__ __device__ inline int row_lanes(int64_t cnt, int64_t nr, int items) {
  const int64_t avg = (cnt + nr - 1) / nr;
  int K = 1;
  while (K < 64 && items * (int64_t)K < avg) K <<= 1;
  return K;
}

// Entries per lane of a matrix's row reduction.  Large operators fill the
// GPU many times over and run fastest with long per-lane runs (R-MAT 23
// base level, FP32 Chebyshev step: 4 -> 1.353 ms, 8 -> 1.302, 16 -> 1.283,
// 32 -> 1.296); small ones launch few workgroups and are latency-bound,
// where short runs win (R-MAT 22 K-cycle: 16 everywhere is 23% slower).
constexpr int64_t ROW_ITEMS_BIG_NNZ = (int64_t)16 << 20;
inline int row_items_for(int64_t nnz) { return nnz >= ROW_ITEMS_BIG_NNZ ? 16 : 4; }

struct Units {
  int64_t nunits = 0, nlong = 0, nparts = 0;
  DevBuf<int32_t> row0;   // first row
  DevBuf<int32_t> nrows;  // > 0: complete rows; <= 0: -(partial slot) of a chunk
  DevBuf<int32_t> koff;   // first stored position
  DevBuf<int32_t> lrow;   // long rows
  DevBuf<int32_t> lpart;  // nlong + 1 slot offsets
  DevBuf<double> parts;   // partial sums (also used as float storage)
  int32_t row_items = 4;  // entries per lane of the row reduction (row_items_for)
  // batched SpMM tile picked per precision (0 FP64, 1 FP32) and log2 of the
  // block width; -1 until tuned (the fixed tile is used)
  int8_t spmm_cfg[2][8] = {{-1, -1, -1, -1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1, -1, -1, -1}};
};

struct HugeRows;  // rows too long for one wavefront of the setup row kernels

struct DCsr {
  int64_t n = 0, ncols = 0, nnz = 0;
  DevBuf<int32_t> rp, ci;
  DevBuf<double> v;   // empty for pattern-only operators (unit values)
  DevBuf<float> v32;  // FP32 copy (mixed precision)
  // FP32 Laplacian form: v32 holds the rounded off-diagonals (0 on the
  // diagonal) and d32 the FP64 diagonal -sum of those rounded entries, so
  // the FP32 operator keeps the constant nullspace exactly.
  DevBuf<double> d32;
  Units U;
  // Column-sliced copies for the solve (index 0: FP64 values, 1: FP32
  // values): slice k holds the entries with columns in [col0[k], col0[k+1])
  // as its own CSR over all rows, so one pass gathers from an x window that
  // stays in L2.  Empty when the operator is used unsliced.
  std::vector<std::unique_ptr<DCsr>> slices[2];
  DevBuf<double> acc;  // FP64 row accumulator across slice passes
  // long-row table of the setup row kernels, built on first use from rp
  // (a DCsr's pattern does not change once built)
  mutable std::shared_ptr<HugeRows> huge;
};

struct DLevel {
  int kind = LIGMG_LEVEL_AGG;
  DCsr A;
  int64_t n = 0, nc = 0;
  // elimination: P (n x nc) and R = P^T (nc x n) with values, Q diagonal
  DCsr P, R;
  DevBuf<int32_t> new_id;
  DevBuf<double> q;
  DevBuf<float> q32;
  int64_t nnzP = 0, nnzQ = 0;
  // aggregation: coarse id per vertex, members CSR (nc x n, unit values)
  DevBuf<int32_t> agg, ind, state;
  DevBuf<double> strength;
  // positions of A's transposed entries, kept from the strength computation
  // until the solve layout has used them
  DevBuf<int32_t> tpos;
  DCsr M;
  DevBuf<double> dinv;
  DevBuf<float> dinv32;
  double lam = 0, cheb_lo = 0, cheb_hi = 0;
  // coarse: dense pseudo-inverse, row-major n x n
  DevBuf<double> pinv;
  // Solve layout: the solve-phase copies (A, dinv, agg, M, P, R, q) of a
  // large level are renumbered by descending degree inside blocks; sperm maps
  // a solve position to the setup position (empty: setup order kept),
  // siperm is its inverse.  agg_setup / new_id keep the setup-order maps for
  // introspection; new_id_s is new_id in solve order.
  DevBuf<int32_t> sperm, siperm, agg_setup, new_id_s;
  std::vector<int64_t> blocks;  // contiguous vertex blocks (independent graphs)
};

}  // namespace lg
