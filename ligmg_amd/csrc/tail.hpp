// Fused coarse tail of the cycle: the levels below a size threshold run as
// one single-workgroup launch per coarse-grid correction.
//
// The reference's K-cycle (multigrid.hpp:101-130) solves every non-exact
// coarse level with kcycle_iters steps of flexible CG preconditioned by the
// next cycle, so a level d aggregation steps below the top is visited
// kcycle_iters^d times per outer iteration.  On the small coarse levels each
// visit costs a few microseconds of launch latency per operation (SpMV, dot,
// scalar update), dozens of them per Krylov step.
//
// The correction below a level has no data-dependent control flow: the
// inner Krylov solves run a fixed number of steps with per-block
// convergence flags on the device.  So the host records it once, at setup,
// as a straight-line program of TailInstr (the same operations, operands
// and buffers the per-operator path launches), and one workgroup of
// TAIL_THREADS threads executes the program with a workgroup barrier where
// the per-operator path has a kernel boundary.
//
// Every instruction keeps the summation order of the kernel it stands for
// (solve.hip, solver.hip): rows the row-block SpMV sums with one thread are
// summed by one thread, wavefront rows by one wavefront with the same
// butterfly, and the 256-thread trees of long-row chunks and of the blocked
// dot products run on 256-thread groups of the workgroup.  The fused tail
// therefore produces the same bits as the separate launches.
#pragma once
#include "ops.hpp"

namespace lg {

// 256 threads: the tail's levels are tiny, and a barrier of 4 wavefronts
// costs less than one of 16 (K-cycle solve 0.485 -> 0.471 s on R-MAT 22)
constexpr int TAIL_THREADS = 256;
// classes of rows reduced by K = 2, 4, ..., 64 lanes
constexpr int TAIL_GCLASSES = 6;

// One CSR operator with the row schedule of its SpMV units: rows of units
// reduced by one thread, rows reduced by one wavefront, and the
// UNIT_NNZ-chunks of long rows with their partial-sum slots.
template <class T>
struct TailOp {
  const int32_t *rp = nullptr, *ci = nullptr;
  const T *v = nullptr;       // null: unit values
  const double *d = nullptr;  // FP64 diagonal term of an FP32 Laplacian-form operator
  // rows reduced by one thread, and rows reduced by groups of K = 2 << c
  // lanes (grow[gofs[c] .. gofs[c + 1]]), as row_lanes() assigns them
  const int32_t *trow = nullptr, *grow = nullptr;
  const int32_t *chunk = nullptr;  // (row, first position, slot) per chunk
  const int32_t *lrow = nullptr, *lpart = nullptr;
  double *parts = nullptr;
  int32_t ntrow = 0, nchunk = 0, nlong = 0;
  int32_t gofs[TAIL_GCLASSES + 1] = {};
};

// Per-level data of the Krylov instructions (RedPlan + KWork).
template <class T>
struct TailLevel {
  int32_t n = 0, nb = 0, nseg = 0;
  const int64_t *seg_lo = nullptr, *seg_hi = nullptr, *bptr = nullptr;
  const int32_t *bseg = nullptr;
  double *part = nullptr, *sc = nullptr, *bsum = nullptr;
  T *kr = nullptr, *kz = nullptr, *kp = nullptr, *kq = nullptr;
};

enum TailOpcode : int32_t {
  TI_SPMV = 0,   // ops[i0] applied to p0, epilogue i1 (TE_*)
  TI_CHEB0,      // p2[i] = beta p0[i] p1[i] (first Chebyshev step), n = i0
  TI_ZERO,       // p0[0..i0) = 0
  TI_COPY,       // p0[0..i0) = p1
  TI_PROLONG,    // p2[i] += p1[p0[i]], i < i0
  TI_SCALE,      // p0[0..i0) *= f0
  TI_DENSE,      // p2 = p0 (dense i0 x i0) p1
  TI_KR_INIT,    // Krylov scalars of level lvl
  TI_DOTS,       // i0 products (a0 b0, a1 b1) = (p0 p1, p2 p3) -> slots i1, i2, then scalar step i3
  TI_UPDATE,     // x (p0) += alpha p, r -= alpha q
  TI_PUPDATE,    // p = z + beta p
  TI_SUBMEAN,    // p0 -= block means (block sums from the previous TI_DOTS with i1 < 0)
};
enum TailEpi : int32_t {
  TE_STORE = 0,  // p1 = y
  TE_RESID,      // p1 = b, p2 = r
  TE_CHEB,       // p1 = dinv, p2 = b, p3 = xp (may be null), p4 = xn; f0 alpha, f1 beta
  TE_INTERP,     // p1 = q, p2 = b, p3 = x
};
enum TailPost : int32_t { TP_NONE = 0, TP_ALPHA, TP_CHECK, TP_BETA };

struct TailInstr {
  int32_t op = 0, lvl = 0, i0 = 0, i1 = 0, i2 = 0, i3 = 0, pad0 = 0, pad1 = 0;
  const void *p[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  double f[2] = {0, 0};
};

// Scalar parameters of the Krylov instructions.
struct TailParams {
  int32_t fcg = 0;
  double rtol = 0;
};

// Host-side owner of an operator's row schedule.
struct TailSched {
  DevBuf<int32_t> trow, grow, chunk;
  int32_t ntrow = 0, nchunk = 0;
  int32_t gofs[TAIL_GCLASSES + 1] = {};
};

// Row schedule of A from its SpMV units (build_units must have run).
void tail_schedule(const DCsr &A, TailSched &S, hipStream_t s);

// Runs a recorded program (len instructions) in one workgroup.  ops and
// lv are device arrays indexed as the program's instructions index them.
template <class T>
void tail_launch(const TailInstr *prog, int64_t len, const TailOp<T> *ops, const TailLevel<T> *lv,
                 const TailParams &P, hipStream_t s);

}  // namespace lg
