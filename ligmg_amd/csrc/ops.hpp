// Solve-phase device operators (row-block SpMV family, vector ops,
// deterministic reductions).  T is the vector element type (double or
// float); matrix values are read in the same precision (A.v or A.v32), or
// as 1 for pattern-only operators.
#pragma once
#include "hierarchy.hpp"

namespace lg {

// Build the SpMV work units of A from its row pointers: the greedy packing
// of complete rows and UNIT_NNZ-chunks, on the device for operators of at
// least UNITS_DEVICE_MIN_ROWS rows.  build_units_host makes the same tables
// with a loop on the host.
constexpr int64_t UNITS_DEVICE_MIN_ROWS = 65536;
void build_units(DCsr &A, hipStream_t s);
void build_units_host(DCsr &A, hipStream_t s);
// Column-sliced copy of A for precision prec (0: FP64, 1: FP32) with
// slices of `width` columns (width >= ncols: no slicing).
void build_slices(DCsr &A, int prec, int64_t width, hipStream_t s);
// Milliseconds per y = A x launch (HIP events, `iters` launches), for tuning.
double time_spmv(const DCsr &A, int prec, int iters, hipStream_t s);

// FP32 copy of the values.  laplacian: store the operator in Laplacian form
// (rounded off-diagonals + FP64 diagonal = -their sum).
void make_v32(DCsr &A, hipStream_t s, bool laplacian = false);

// y = A x
template <class T> void op_spmv(const DCsr &A, const T *x, T *y, hipStream_t s);
// r = b - A x
template <class T> void op_residual(const DCsr &A, const T *x, const T *b, T *r, hipStream_t s);
// One Chebyshev step: xn = xc + alpha (xc - xp) + beta dinv (b - A xc).
// With xc == nullptr the current iterate is zero: xn = beta dinv b.
template <class T>
void op_cheb_step(const DCsr &A, const T *dinv, const T *b, const T *xc, const T *xp, T *xn,
                  double alpha, double beta, hipStream_t s);
// x = P xc + q b (elimination interpolation, P has values)
template <class T>
void op_interp(const DCsr &P, const T *q, const T *xc, const T *b, T *x, hipStream_t s);
// x[i] += xc[agg[i]]
template <class T>
void op_prolong_add(const int32_t *agg, int64_t n, const T *xc, T *x, hipStream_t s);
// y = pinv x (dense n x n, FP64 matrix; T vectors)
template <class T>
void op_dense_apply(const double *Pinv, int64_t n, const T *x, T *y, hipStream_t s);

void op_d2f(const double *a, float *b, int64_t n, hipStream_t s);
void op_f2d(const float *a, double *b, int64_t n, hipStream_t s);

// Plan of a deterministic blocked reduction over vectors of length n split
// into contiguous blocks (independent graphs).  Stage 1 sums fixed chunks
// clipped to block boundaries (one workgroup each, fixed tree); stage 2
// adds the chunks of each block in order, so results are run-to-run
// reproducible.
struct RedPlan {
  int64_t n = 0;
  int nb = 1, nseg = 0;
  DevBuf<int64_t> seg_lo, seg_hi;  // nseg
  DevBuf<int32_t> bseg;            // nb + 1
  DevBuf<int64_t> bptr;            // nb + 1 block offsets
  mutable DevBuf<double> part;     // nseg * 4
  std::vector<int64_t> hbptr;
  void build(int64_t n, const std::vector<int64_t> &blocks, hipStream_t s);
  // out[c*np + j] = sum over block c of a_j[i] * b_j[i] (b_j == nullptr:
  // 1), j < np (np <= 4), accumulated in FP64.
  template <class T>
  void dots(int np, const T *const *a, const T *const *b, double *out, hipStream_t s) const;
  // Krylov form (np <= 2): block c's results go to the scalar slots
  // sc[c * SC_N + s0] (and s1), then the scalar step `post` (KrPost) runs on
  // them in the same launch.  Same bits as dots + the separate scalar kernels.
  template <class T>
  void dots_sc(int np, const T *const *a, const T *const *b, double *sc, int s0, int s1, int post, int fcg,
               double rtol, hipStream_t s) const;
  // out[c] = sum over block c of a[i]
  template <class T> void sums(const T *a, double *out, hipStream_t s) const;
  // block vectors (n x kp row-major): out[(blk * np + j) * kp + c] = sum over
  // block blk of a_j[i, c] * b_j[i, c]  (np <= 2, b_j == nullptr: 1)
  template <class T>
  void bdots(int np, int kp, const T *a0, const T *b0, const T *a1, const T *b1, double *out,
             DevBuf<double> &bpart, hipStream_t s) const;
  // the second stage of bdots alone, from segment partials a fused kernel
  // wrote to bpart in bdots' layout
  void bdots_finish(int np, int kp, double *out, const DevBuf<double> &bpart, hipStream_t s) const;
};

// a -= mean of a over each block, means taken from sums[c] (device)
template <class T> void op_sub_block_mean(T *a, const RedPlan &P, const double *sums, hipStream_t s);

// ---- block (k right-hand sides) operators, batch.hip.  Block vectors are
// row-major n x kp, kp a power of two <= 64.
template <class T>
void op_bspmv(const DCsr &A, const T *X, T *Y, int kp, DevBuf<double> &parts, hipStream_t s);
// picks A's batched SpMM tile for this precision and block width by timing
// the candidates once (kept in A.U); returns the tile id
template <class T>
int tune_spmm(DCsr &A, int kp, DevBuf<double> &parts, hipStream_t s);
// right-hand sides per lane and gathers in flight of tile id cfg
void spmm_tile(int cfg, int *vw, int *u);
template <class T>
void op_bresidual(const DCsr &A, const T *X, const T *B, T *R, int kp, DevBuf<double> &parts,
                  hipStream_t s);
template <class T>
void op_bcheb_step(const DCsr &A, const T *dinv, const T *B, const T *Xc, const T *Xp, T *Xn,
                   double alpha, double beta, int kp, DevBuf<double> &parts, hipStream_t s);
template <class T>
void op_binterp(const DCsr &P, const T *q, const T *Xc, const T *B, T *X, int kp,
                DevBuf<double> &parts, hipStream_t s);
template <class T>
void op_bprolong_add(const int32_t *agg, int64_t n, const T *Xc, T *X, int kp, hipStream_t s);
// Y = Pm X with FP64 MFMA tiles (kp >= 16)
template <class T>
void op_bdense_apply(const double *Pm, int64_t n, const T *X, T *Y, int kp, hipStream_t s);
template <class T>
void op_bsub_block_mean(T *a, int kp, const RedPlan &P, const double *sums, hipStream_t s);
// z = c (FP32) minus the block means of c from sums, in FP64: the same bits as
// converting c and then subtracting the means.  With w set, also the segment
// partials of z.w in bpart, as RedPlan::bdots forms them (finish them with
// bdots_finish).
void op_bsub_block_mean_f2d(double *z, const float *c, int kp, const RedPlan &P, const double *sums, hipStream_t s,
                            const double *w = nullptr, DevBuf<double> *bpart = nullptr);
// column-major n x k caller block <-> row-major n x kp internal block; the
// caller's vertex of internal row i is perm[i] (identity when perm is null)
void op_to_block(const double *B, int64_t n, int k, int kp, double *X, const int32_t *perm, hipStream_t s);
void op_from_block(const double *X, int64_t n, int k, int kp, double *B, const int32_t *perm,
                   hipStream_t s);

}  // namespace lg
