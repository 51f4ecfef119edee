// Device setup: low-degree elimination, affinity strength, voting
// aggregation, Galerkin products by contraction (sort + canonical segmented
// sum), smoother set-up and the coarse pseudo-inverse.
//
// Summation orders.  The sums that the reference computes on one rank in a
// fixed order of its own are left folds from 0.0 in ascending index
// (run_rows_seq, segmented_seq_sum): col_sum and the Row/Column scaling
// sums (CombBLAS Reduce), the vote tally (tally_votes' loop over voters)
// and the conductance sums (ext_int_ratio's loop).  Every other
// floating-point sum whose order could change a result (SpMV-shaped sums
// whose order the reference leaves to CombBLAS/PETSc) uses the canonical
// order: a row (or run of equal keys) of at most 64 terms is summed left to
// right by one lane; a longer one by a whole wavefront as 64 strided partial
// sums combined by the tree p[l] += p[l+off], off = 32..1.  Masked row
// positions contribute +0.0 (which changes neither kind of sum).  Files
// compiled with -ffp-contract=off, so products and sums round exactly as
// written.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <type_traits>

#include "../../include/ligmg_amd.h"
#include "ops.hpp"
#include "prims.hpp"
#include "rng.hpp"
#include "setup.hpp"

namespace lg {

namespace {

enum { ST_SEED = 0, ST_SEEDNEIGHBOR = 1, ST_DECIDED = 2, ST_DECIDEDROOT = 3, ST_UNDECIDED = 4 };
// vote priorities: Seed > Undecided > Decided (aggregation.hpp:408-417)
enum { PRIO_DECIDED = 0, PRIO_UNDECIDED = 1, PRIO_SEED = 2 };

constexpr int MAXT = 16;

// --------------------------------------------------------------------------
// Generic canonical row kernel: one lane per row, rows of 65..4096 entries
// done cooperatively by the wavefront.  Rows longer than CHUNK are left to
// k_rows_chunks (one wavefront per 4096-entry chunk) and k_rows_huge.
// --------------------------------------------------------------------------
constexpr int CHUNK = HUGE_ROW;

// An Op with skip(row) leaves such rows out entirely (no add, no finish):
// for operators whose results are only read for the other rows.
template <class Op, class = void>
struct HasSkip : std::false_type {};
template <class Op>
struct HasSkip<Op, std::void_t<decltype(std::declval<const Op &>().skip(0))>> : std::true_type {};
template <class Op>
__device__ __forceinline__ bool row_skip(const Op &op, int row) {
  if constexpr (HasSkip<Op>::value) return op.skip(row);
  else return false;
}

// Rows of at most WAVE entries, one lane each, summed left to right (in
// row order: neighbouring lanes read neighbouring rows.  Grouping the rows
// by length, so that the lanes of a wavefront loop about equally long, was
// measured slower: 25 -> 34 ms of voting, 14 -> 22 ms of affinities per
// R-MAT 23 setup).
template <class Op>
__global__ void __launch_bounds__(256) k_rows(int n, const int32_t *__restrict__ rp, Op op) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n || row_skip(op, row)) return;
  const int a = rp[row], b = rp[row + 1];
  if (b - a > WAVE) return;
  typename Op::Acc acc = op.zero();
  // unrolled: the gathers of several entries are in flight at once; the
  // adds keep their order
#pragma unroll 4
  for (int k = a; k < b; k++) op.add(acc, row, k);
  op.finish(row, acc);
}

// Rows of WAVE + 1 .. CHUNK entries (mrow).  Wavefront w takes the rows
// j = w, w + W, w + 2W, ... of the list (W wavefronts in the grid) in
// batches of 64, lane l holding the l-th row of a batch: its bounds are
// loaded for all lanes at once, the wavefront sums the batch's rows one
// after another (64 strided partial sums and the fixed tree each; lane l
// keeps row l's result) and the lanes finish their rows together.  Striding
// spreads the long rows evenly over the wavefronts wherever they sit (hubs
// are neighbours in vertex order on R-MAT graphs and in the solve layout).
template <class Op>
__global__ void __launch_bounds__(256) k_rows_mid(int64_t nmid, const int32_t *__restrict__ mrow,
                                                  const int32_t *__restrict__ rp, Op op) {
  using Acc = typename Op::Acc;
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x / WAVE);
  const int64_t w = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE;
  for (int64_t base = w; base < nmid; base += nw * WAVE) {
    const int64_t j = base + (int64_t)lane * nw;
    int row = -1, a = 0, b = 0;
    if (j < nmid) {
      row = mrow[j];
      if (row_skip(op, row)) {
        row = -1;
      } else {
        a = rp[row];
        b = rp[row + 1];
      }
    }
    Acc acc = op.zero();
    uint64_t mask = __ballot(row >= 0);
    while (mask) {
      const int L = __ffsll((long long)mask) - 1;
      mask &= mask - 1;
      const int la = __shfl(a, L, WAVE), lb = __shfl(b, L, WAVE), lr = __shfl(row, L, WAVE);
      Acc p = op.zero();
#pragma unroll 4
      for (int k = la + lane; k < lb; k += WAVE) op.add(p, lr, k);
      for (int off = 32; off >= 1; off >>= 1) {
        Acc o = op.shfl_down(p, off);
        p = op.merge(p, o);
      }
      p = op.shfl0(p);
      if (lane == L) acc = p;
    }
    if (row >= 0) op.finish(row, acc);
  }
}

// wavefronts of k_rows_mid: one per row up to about 8 per SIMD of the GPU
// (each then sums a batch of up to 64 rows, so up to 2^19 rows take one
// batch per wavefront)
constexpr int64_t MID_WAVES = 256 * 4 * 8;

// one wavefront per chunk of a huge row: the chunk's canonical sum
template <class Op>
__global__ void __launch_bounds__(256) k_rows_chunks(int64_t nchunks, const int32_t *cstart,
                                                     const int32_t *cend, const int32_t *crow, Op op,
                                                     typename Op::Acc *cacc) {
  using Acc = typename Op::Acc;
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t c = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE;
  if (c >= nchunks) return;
  const int la = cstart[c], lb = cend[c], lr = crow[c];
  if (row_skip(op, lr)) return;  // the whole wavefront: one chunk per wavefront
  Acc p = op.zero();
  if (lb - la <= WAVE) {
    if (lane == 0)
      for (int k = la; k < lb; k++) op.add(p, lr, k);
  } else {
#pragma unroll 4
    for (int k = la + lane; k < lb; k += WAVE) op.add(p, lr, k);
    for (int off = 32; off >= 1; off >>= 1) {
      Acc o = op.shfl_down(p, off);
      p = op.merge(p, o);
    }
  }
  if (lane == 0) cacc[c] = p;
}

// chunk results of each huge row combined left to right
template <class Op>
__global__ void k_rows_huge(int64_t nrows, const int32_t *row, const int32_t *cptr, Op op,
                            const typename Op::Acc *cacc) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= nrows || row_skip(op, row[j])) return;
  typename Op::Acc acc = op.zero();
  for (int c = cptr[j]; c < cptr[j + 1]; c++) acc = op.merge(acc, cacc[c]);
  op.finish(row[j], acc);
}

}  // namespace

namespace {

// rows with lo < length <= hi
__global__ void k_huge_flags(int64_t n, const int32_t *rp, int32_t lo, int32_t hi, int32_t *flag) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t len = rp[i + 1] - rp[i];
    flag[i] = (len > lo && len <= hi) ? 1 : 0;
  }
}
__global__ void k_huge_nchunks(int64_t nh, const int32_t *row, const int32_t *rp, int32_t *cnt) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nh; j += (int64_t)gridDim.x * blockDim.x) {
    const int32_t len = rp[row[j] + 1] - rp[row[j]];
    cnt[j] = (len + CHUNK - 1) / CHUNK;
  }
}
// the chunks of huge row j: [rp[r] + c * CHUNK, min(.. + CHUNK, rp[r + 1]))
__global__ void k_huge_fill(int64_t nh, const int32_t *row, const int32_t *rp, const int32_t *cptr,
                            int32_t *cstart, int32_t *cend, int32_t *crow) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nh; j += (int64_t)gridDim.x * blockDim.x) {
    const int32_t r = row[j];
    const int64_t b = rp[r + 1];
    int32_t c = cptr[j];
    for (int64_t k = rp[r]; k < b; k += CHUNK, c++) {
      cstart[c] = (int32_t)k;
      cend[c] = (int32_t)(b - k > CHUNK ? k + CHUNK : b);
      crow[c] = r;
    }
  }
}

}  // namespace

// Rows longer than CHUNK, in ascending order, and their chunks, built on the
// device (two counts come back to size the tables).
void find_huge(int64_t n, const int32_t *rp, HugeRows &Hr, hipStream_t s) {
  Hr.nrows = 0;
  Hr.nchunks = 0;
  Hr.nlong = 0;
  Hr.nmid = 0;
  if (n <= 0) return;
  DevBuf<int32_t> flag(n), idx(n);
  constexpr int32_t ANY = 0x7fffffff;
  hipLaunchKernelGGL(k_huge_flags, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, n, rp, (int32_t)WAVE,
                     (int32_t)CHUNK, flag.p);
  LG_LAUNCH_CHECK();
  Hr.nmid = compact_flags(flag.p, n, idx.p, s);
  if (Hr.nmid) {
    Hr.mrow.alloc(Hr.nmid);
    d2d(Hr.mrow.p, idx.p, Hr.nmid, s);
  }
  hipLaunchKernelGGL(k_huge_flags, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, n, rp, (int32_t)SEQ_LANE,
                     ANY, flag.p);
  LG_LAUNCH_CHECK();
  Hr.nlong = compact_flags(flag.p, n, idx.p, s);
  if (Hr.nlong == 0) return;  // no row is longer than CHUNK either
  Hr.lrow.alloc(Hr.nlong);
  d2d(Hr.lrow.p, idx.p, Hr.nlong, s);
  hipLaunchKernelGGL(k_huge_flags, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, n, rp, (int32_t)CHUNK, ANY,
                     flag.p);
  LG_LAUNCH_CHECK();
  const int64_t nh = compact_flags(flag.p, n, idx.p, s);
  if (nh == 0) return;
  Hr.nrows = nh;
  Hr.row.alloc(nh);
  d2d(Hr.row.p, idx.p, nh, s);
  Hr.cptr.alloc(nh + 1);
  hipLaunchKernelGGL(k_huge_nchunks, dim3(grid_for(nh, 256, 16384)), dim3(256), 0, s, nh, Hr.row.p, rp, Hr.cptr.p);
  LG_LAUNCH_CHECK();
  fill_i32(Hr.cptr.p + nh, 0, 1, s);
  Hr.nchunks = exclusive_scan_i32(Hr.cptr.p, Hr.cptr.p, nh + 1, s);
  Hr.cstart.alloc(Hr.nchunks);
  Hr.cend.alloc(Hr.nchunks);
  Hr.crow.alloc(Hr.nchunks);
  hipLaunchKernelGGL(k_huge_fill, dim3(grid_for(nh, 256, 16384)), dim3(256), 0, s, nh, Hr.row.p, rp, Hr.cptr.p,
                     Hr.cstart.p, Hr.cend.p, Hr.crow.p);
  LG_LAUNCH_CHECK();
  Hr.cacc.alloc((size_t)Hr.nchunks * 192);  // room for the largest Acc
}

// the long-row table of A, built once per matrix
const HugeRows &huge_of(const DCsr &A, hipStream_t s) {
  if (!A.huge) {
    auto h = std::make_shared<HugeRows>();
    find_huge(A.n, A.rp.p, *h, s);
    A.huge = std::move(h);
  }
  return *A.huge;
}

namespace {

template <class Op>
void run_rows(int64_t n, const int32_t *rp, const Op &op, const HugeRows &Hr, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_rows<Op>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (int)n, rp, op);
  LG_LAUNCH_CHECK();
  if (Hr.nmid) {
    const int64_t waves = std::min<int64_t>(Hr.nmid, MID_WAVES);
    hipLaunchKernelGGL(k_rows_mid<Op>, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, Hr.nmid, Hr.mrow.p,
                       rp, op);
    LG_LAUNCH_CHECK();
  }
  if (Hr.nrows) {
    using Acc = typename Op::Acc;
    static_assert(sizeof(Acc) <= 192, "Acc too large for the chunk buffer");
    Acc *cacc = (Acc *)Hr.cacc.p;
    hipLaunchKernelGGL(k_rows_chunks<Op>, dim3((unsigned)((Hr.nchunks + 3) / 4)), dim3(256), 0, s,
                       Hr.nchunks, Hr.cstart.p, Hr.cend.p, Hr.crow.p, op, cacc);
    LG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_rows_huge<Op>, dim3((unsigned)((Hr.nrows + 63) / 64)), dim3(64), 0, s, Hr.nrows,
                       Hr.row.p, Hr.cptr.p, op, (const Acc *)cacc);
    LG_LAUNCH_CHECK();
  }
}

// Sequential row sums for DblSum operators: each row's terms folded from
// 0.0 in stored (ascending column) order.  Rows of at most SEQ_LANE entries
// take one lane each; longer ones one wavefront each (wave_seq_fold).  The
// term of position k is op.add applied to 0.0, which is the term itself
// (a -0.0 term becomes +0.0, which no fold from +0.0 can tell apart).
template <class Op>
__global__ void __launch_bounds__(256) k_rows_seq(int n, const int32_t *__restrict__ rp, Op op) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n || row_skip(op, row)) return;
  const int a = rp[row], b = rp[row + 1];
  if (b - a > SEQ_LANE) return;
  double acc = 0.0;
#pragma unroll 8
  for (int k = a; k < b; k++) op.add(acc, row, k);
  op.finish(row, acc);
}

template <class Op>
__global__ void __launch_bounds__(256) k_rows_seq_long(int64_t nrows, const int32_t *__restrict__ lrow,
                                                       const int32_t *__restrict__ rp, Op op) {
  __shared__ double buf[4 * SEQ_BLOCK];
  const int64_t j = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE;
  if (j >= nrows) return;
  const int row = lrow[j];
  if (row_skip(op, row)) return;  // the whole wavefront: one row per wavefront
  const double acc = wave_seq_fold(rp[row], rp[row + 1], buf + (threadIdx.x / WAVE) * SEQ_BLOCK, [&](int64_t k) {
    double t = 0.0;
    op.add(t, row, (int)k);
    return t;
  });
  if ((threadIdx.x & (WAVE - 1)) == 0) op.finish(row, acc);
}

template <class Op>
void run_rows_seq(int64_t n, const int32_t *rp, const Op &op, const HugeRows &Hr, hipStream_t s) {
  static_assert(std::is_same<typename Op::Acc, double>::value, "sequential rows are double sums");
  if (n <= 0) return;
  hipLaunchKernelGGL(k_rows_seq<Op>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (int)n, rp, op);
  LG_LAUNCH_CHECK();
  if (Hr.nlong) {
    hipLaunchKernelGGL(k_rows_seq_long<Op>, dim3((unsigned)((Hr.nlong + 3) / 4)), dim3(256), 0, s, Hr.nlong,
                       Hr.lrow.p, rp, op);
    LG_LAUNCH_CHECK();
  }
}

struct IntSum {
  using Acc = int;
  __device__ Acc zero() const { return 0; }
  __device__ Acc merge(Acc a, Acc b) const { return a + b; }
  __device__ Acc shfl_down(Acc a, int off) const { return __shfl_down(a, off, WAVE); }
  __device__ Acc shfl0(Acc a) const { return __shfl(a, 0, WAVE); }
};
struct DblSum {
  using Acc = double;
  __device__ Acc zero() const { return 0.0; }
  __device__ Acc merge(Acc a, Acc b) const { return a + b; }
  __device__ Acc shfl_down(Acc a, int off) const { return __shfl_down(a, off, WAVE); }
  __device__ Acc shfl0(Acc a) const { return __shfl(a, 0, WAVE); }
};
struct DblMax {  // std::max semantics: (a < b) ? b : a
  using Acc = double;
  __device__ Acc zero() const { return 0.0; }
  __device__ Acc merge(Acc a, Acc b) const { return (a < b) ? b : a; }
  __device__ Acc shfl_down(Acc a, int off) const { return __shfl_down(a, off, WAVE); }
  __device__ Acc shfl0(Acc a) const { return __shfl(a, 0, WAVE); }
};
struct NoAcc {
  using Acc = int;
  __device__ Acc zero() const { return 0; }
  __device__ Acc merge(Acc a, Acc) const { return a; }
  __device__ Acc shfl_down(Acc a, int) const { return a; }
  __device__ Acc shfl0(Acc a) const { return a; }
};

// ---- a1: compute_degrees (elimination.hpp:70-80) -----------------------------
struct OpDeg : IntSum {
  const double *v;
  int32_t *deg;
  __device__ void add(int &acc, int, int k) const { acc += (v[k] == 0) ? 0 : 1; }
  __device__ void finish(int row, int acc) const { deg[row] = acc - 1; }
};

// ---- a2: IndxHashMin SpMV: smallest candidate index over A_ij != 0
//      (elimination.hpp:122-154, 201-202)
struct OpElimMin {
  using Acc = int;
  const int32_t *ci, *deg;
  const double *v;
  int maxd;
  int32_t *elim;
  __device__ Acc zero() const { return 0x7fffffff; }
  __device__ Acc merge(Acc a, Acc b) const { return a < b ? a : b; }
  __device__ Acc shfl_down(Acc a, int off) const { return __shfl_down(a, off, WAVE); }
  __device__ Acc shfl0(Acc a) const { return __shfl(a, 0, WAVE); }
  __device__ void add(int &acc, int, int k) const {
    int j = ci[k];
    if (v[k] != 0 && deg[j] <= maxd && j < acc) acc = j;
  }
  __device__ void finish(int row, int acc) const { elim[row] = (acc == row) ? 1 : 0; }
};

// Schur contraction entry count per row: kept rows keep their kept
// columns; an eliminated row with m off-diagonal entries emits m*m.
struct OpSchurCount : IntSum {
  const int32_t *rp, *ci, *newid;
  int32_t *cnt;
  __device__ void add(int &acc, int, int k) const { acc += (newid[ci[k]] >= 0) ? 1 : 0; }
  __device__ void finish(int row, int acc) const {
    if (newid[row] >= 0) {
      cnt[row] = acc;
    } else {
      int m = 0;
      for (int k = rp[row]; k < rp[row + 1]; k++) m += (ci[k] != row);
      cnt[row] = m * m;
    }
  }
};

__global__ void k_kept_flags(int64_t n, const int32_t *elim, int32_t *kept) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    kept[i] = 1 - elim[i];
}
__global__ void k_mask_ids(int64_t n, const int32_t *kept, int32_t *id) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if (!kept[i]) id[i] = -1;
}

// every off-diagonal neighbour of an eliminated vertex must be kept
__global__ void k_elim_check(int n, const int32_t *rp, const int32_t *ci, const int32_t *newid,
                             int32_t *bad) {
  int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n || newid[row] >= 0) return;
  for (int k = rp[row]; k < rp[row + 1]; k++)
    if (ci[k] != row && newid[ci[k]] < 0) atomicOr(bad, 1);
}

// P row lengths: kept 1, eliminated = off-diagonal entries
__global__ void k_p_count(int n, const int32_t *rp, const int32_t *ci, const int32_t *newid,
                          int32_t *cnt) {
  int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  if (newid[row] >= 0) {
    cnt[row] = 1;
    return;
  }
  int m = 0;
  for (int k = rp[row]; k < rp[row + 1]; k++) m += (ci[k] != row);
  cnt[row] = m;
}

__device__ __forceinline__ double row_diag(const int32_t *rp, const int32_t *ci, const double *v,
                                           int row, bool *found) {
  int lo = rp[row], hi = rp[row + 1] - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    int c = ci[mid];
    if (c == row) {
      *found = true;
      return v[mid];
    }
    if (c < row) lo = mid + 1;
    else hi = mid - 1;
  }
  *found = false;
  return 0.0;
}

// ---- a4: P, Q assembly (elimination.hpp:287-331); P(e, new i) = 1/d * -L_ei
__global__ void k_p_fill(int n, const int32_t *rp, const int32_t *ci, const double *v,
                         const int32_t *newid, const int32_t *prp, int32_t *pci, double *pv,
                         double *q) {
  int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  int o = prp[row];
  if (newid[row] >= 0) {
    pci[o] = newid[row];
    pv[o] = 1.0;
    q[row] = 0.0;
    return;
  }
  bool f;
  const double d = row_diag(rp, ci, v, row, &f);
  // an isolated vertex (no nonzero off-diagonal: a block reduced to one
  // vertex) gets 0: its value is a nullspace component
  bool iso = true;
  for (int k = rp[row]; k < rp[row + 1]; k++)
    if (ci[k] != row && v[k] != 0) iso = false;
  const double inv = iso ? 0.0 : 1 / d;
  q[row] = inv;
  for (int k = rp[row]; k < rp[row + 1]; k++) {
    int i = ci[k];
    if (i == row) continue;
    pci[o] = newid[i];
    pv[o] = inv * -v[k];
    o++;
  }
}

// FP32 elimination operators from the FP32 Laplacian form of the level
// matrix (off-diagonals v32, diagonal d32 = -sum of them): q32 = 1 / d32,
// P32(e, b) = -v32(e, b) / d32(e), so every P32 row sums to 1 again and the
// FP32 cycle eliminates exactly the operator it smooths with.
__global__ void k_p_fill32(int n, const int32_t *rp, const int32_t *ci, const float *v32, const double *d32,
                           const int32_t *newid, const int32_t *prp, float *pv32, float *q32) {
  int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  int o = prp[row];
  if (newid[row] >= 0) {
    pv32[o] = 1.0f;
    q32[row] = 0.0f;
    return;
  }
  const double d = d32[row];
  const double inv = d != 0 ? 1 / d : 0.0;
  q32[row] = (float)inv;
  for (int k = rp[row]; k < rp[row + 1]; k++) {
    if (ci[k] == row) continue;
    pv32[o++] = (float)(inv * -(double)v32[k]);
  }
}

__global__ void k_p_transpose_keys32(int n, const int32_t *prp, const int32_t *pci, const float *pv32,
                                     int rb, uint64_t *keys, uint64_t *vals) {
  int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  for (int k = prp[row]; k < prp[row + 1]; k++) {
    keys[k] = ((uint64_t)pci[k] << rb) | (uint64_t)row;
    vals[k] = (uint64_t)__float_as_uint(pv32[k]);
  }
}

__global__ void k_unpack_vals32(const uint64_t *vals, int64_t m, float *v32) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x)
    v32[i] = __uint_as_float((uint32_t)vals[i]);
}

// inverse diagonal for the smoother; isolated rows get 0 (see k_p_fill).
// One wavefront per row (hub rows have hundreds of thousands of entries).
// The kept-column entries of positions [la, lb) of kept row `row`, in
// stored order from list position o, by the whole wavefront.
__device__ __forceinline__ void schur_copy_kept(int la, int lb, int o, int row, const int32_t *ci,
                                                const double *v, const int32_t *newid, int cb, uint64_t *keys,
                                                uint64_t *vals) {
  const int lane = threadIdx.x & (WAVE - 1);
  const uint64_t r = (uint64_t)newid[row] << cb;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int base = la; base < lb; base += WAVE) {
    const int k = base + lane;
    const int col = (k < lb) ? newid[ci[k]] : -1;
    const uint64_t m = __ballot(col >= 0);
    if (col >= 0) {
      const int pos = o + __popcll(m & lt);
      keys[pos] = r | (uint64_t)col;
      vals[pos] = (uint64_t)__double_as_longlong(v[k]);
    }
    o += __popcll(m);
  }
}

// ---- a6: Schur contraction entries, in the oracle's list order: rows in
//      ascending order; a kept row lists its kept columns; an eliminated row
//      e lists (a, b) -> L_ea * P(e, b) over its off-diagonal pairs.
__global__ void __launch_bounds__(256) k_schur_fill(int n, const int32_t *rp, const int32_t *ci,
                                                    const double *v, const int32_t *newid,
                                                    const int32_t *off, const int32_t *prp,
                                                    const double *pv, int cb, uint64_t *keys,
                                                    uint64_t *vals) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = row < n;
  int a = 0, b = 0;
  bool kept = false;
  if (valid) {
    a = rp[row];
    b = rp[row + 1];
    kept = newid[row] >= 0;
  }
  // kept rows of 17 .. WAVE entries are copied by the whole wavefront
  // (coalesced), shorter ones and eliminated rows by their own lane; longer
  // kept rows are left to k_schur_rows_fill and k_schur_chunk_fill
  const bool wide = valid && kept && (b - a > WAVE);
  const bool lng = valid && kept && !wide && (b - a > 16);
  if (valid && !lng && !wide) {
    int o = off[row];
    if (kept) {
      uint64_t r = (uint64_t)newid[row] << cb;
      for (int k = a; k < b; k++) {
        int c = newid[ci[k]];
        if (c >= 0) {
          keys[o] = r | (uint64_t)c;
          vals[o] = (uint64_t)__double_as_longlong(v[k]);
          o++;
        }
      }
    } else {
      int pk = prp[row];
      for (int k = a; k < b; k++) {
        if (ci[k] == row) continue;
        double wa = v[k];
        uint64_t na = (uint64_t)newid[ci[k]] << cb;
        int pt = prp[row];
        for (int k2 = a; k2 < b; k2++) {
          if (ci[k2] == row) continue;
          keys[o] = na | (uint64_t)newid[ci[k2]];
          vals[o] = (uint64_t)__double_as_longlong(wa * pv[pt]);
          o++;
          pt++;
        }
        pk++;
      }
    }
  }
  uint64_t mask = __ballot(lng);
  while (mask) {
    const int L = __ffsll((long long)mask) - 1;
    mask &= mask - 1;
    const int la = __shfl(a, L, WAVE), lb = __shfl(b, L, WAVE), lr = __shfl(row, L, WAVE);
    schur_copy_kept(la, lb, off[lr], lr, ci, v, newid, cb, keys, vals);
  }
}

// Kept entries of each CHUNK-entry chunk of a hub row, one wavefront per
// chunk.  Hub rows belong to kept vertices (an eliminated vertex has a few
// neighbours at most); a chunk of any other row counts 0.
__global__ void __launch_bounds__(256) k_schur_chunk_count(int64_t nchunks, const int32_t *cstart,
                                                           const int32_t *cend, const int32_t *crow,
                                                           const int32_t *ci, const int32_t *newid,
                                                           int32_t *ccnt) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t c = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE;
  if (c >= nchunks) return;  // the whole wavefront
  int cnt = 0;
  if (newid[crow[c]] >= 0)
    for (int k = cstart[c] + lane; k < cend[c]; k += WAVE) cnt += (newid[ci[k]] >= 0) ? 1 : 0;
  for (int off = 32; off >= 1; off >>= 1) cnt += __shfl_xor(cnt, off, WAVE);
  if (lane == 0) ccnt[c] = cnt;
}

// Kept rows of WAVE + 1 .. CHUNK entries (mrow), one wavefront each.
__global__ void __launch_bounds__(256) k_schur_rows_fill(int64_t nmid, const int32_t *mrow, const int32_t *rp,
                                                         const int32_t *ci, const double *v,
                                                         const int32_t *newid, const int32_t *off, int cb,
                                                         uint64_t *keys, uint64_t *vals) {
  const int64_t j = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE;
  if (j >= nmid) return;  // the whole wavefront
  const int32_t row = mrow[j];
  if (newid[row] < 0) return;
  schur_copy_kept(rp[row], rp[row + 1], off[row], row, ci, v, newid, cb, keys, vals);
}

// One chunk of a hub row per wavefront, at the row's list offset plus the
// kept entries of the row's earlier chunks (cpos: exclusive scan of the
// chunk counts; the row's first chunk sits (cstart - rp[row]) / CHUNK
// chunks back).
__global__ void __launch_bounds__(256) k_schur_chunk_fill(int64_t nchunks, const int32_t *cstart,
                                                          const int32_t *cend, const int32_t *crow,
                                                          const int32_t *rp, const int32_t *ci,
                                                          const double *v, const int32_t *newid,
                                                          const int32_t *off, const int32_t *cpos, int cb,
                                                          uint64_t *keys, uint64_t *vals) {
  const int64_t c = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE;
  if (c >= nchunks) return;  // the whole wavefront
  const int32_t row = crow[c], la = cstart[c];
  if (newid[row] < 0) return;
  const int64_t c0 = c - (la - rp[row]) / CHUNK;
  schur_copy_kept(la, cend[c], off[row] + (cpos[c] - cpos[c0]), row, ci, v, newid, cb, keys, vals);
}

// CSR from sorted unique keys: rows counted, columns = low bits.
// Row pointers of a CSR whose entries are keys (row << cb | col) sorted
// ascending: rp[r] = first position whose row is >= r.  Position i writes
// rp for the rows between its predecessor's row and its own, so every
// rp entry is written exactly once (i = m closes the trailing rows).
__global__ void k_sorted_rowptr(const uint64_t *keys, int64_t m, int cb, int64_t nrows, int32_t *rp) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= m;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = (i < m) ? (int64_t)(keys[i] >> cb) : nrows + 1;
    const int64_t pr = (i > 0) ? (int64_t)(keys[i - 1] >> cb) : -1;
    for (int64_t q = pr + 1; q <= r && q <= nrows; q++) rp[q] = (int32_t)i;
  }
}

__global__ void k_key_cols(const uint64_t *keys, int64_t m, int cb, int32_t *ci) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x)
    ci[i] = (int32_t)(keys[i] & (((uint64_t)1 << cb) - 1));
}

// transpose of P: keys (col << rb) | row
__global__ void k_p_transpose_keys(int n, const int32_t *prp, const int32_t *pci, const double *pv,
                                   int rb, uint64_t *keys, uint64_t *vals) {
  int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  for (int k = prp[row]; k < prp[row + 1]; k++) {
    keys[k] = ((uint64_t)pci[k] << rb) | (uint64_t)row;
    vals[k] = (uint64_t)__double_as_longlong(pv[k]);
  }
}

__global__ void k_unpack_pairs(const uint64_t *keys, const uint64_t *vals, int64_t m, int rb, int32_t *ci,
                               double *v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    ci[i] = (int32_t)(keys[i] & (((uint64_t)1 << rb) - 1));
    v[i] = __longlong_as_double((long long)vals[i]);
  }
}

// ---- a7: test vectors (aggregation.hpp:24-36) ------------------------------
__global__ void k_tv_sum(int n, uint64_t seed, int level, int T, unsigned long long *sums) {
  // per-block partial sums in registers, one atomic per block and t
  for (int t = 0; t < T; t++) {
    unsigned long long s = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
      s += ligmg_rng::tv_bits(seed, level, t, i);
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, WAVE);
    if ((threadIdx.x & (WAVE - 1)) == 0) atomicAdd(&sums[t], s);
  }
}

// Test vectors are stored vertex-major with a row stride TS (the power of
// two >= T): a vertex's T values never straddle a cache line, so each
// gather of a neighbour's row costs one line.  Slots t >= T stay unused.
__global__ void k_tv_fill(int n, uint64_t seed, int level, int T, int TS, const unsigned long long *sums,
                          double *X) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int t = 0; t < T; t++) {
    double mean = ((double)sums[t] / (double)n) * 0x1p-32;
    X[(int64_t)i * TS + t] = (double)ligmg_rng::tv_bits(seed, level, t, i) * 0x1p-32 - mean;
  }
}

// ---- a8: Jacobi on the test vectors, IgnoreDiagSR, b = 0
//      (smoothing.hpp:11-32, 78-98)
// accumulators of N test vectors (N >= T; lanes t >= T stay unused)
template <int N>
struct TVAcc {
  double v[N];
};
// One weighted Jacobi sweep on the T test vectors, N = the accumulator width
// chosen for T at the launch (each vector's sum has the same order for any N).
template <int N>
struct OpJacobiTV {
  using Acc = TVAcc<N>;
  const int32_t *ci;
  const double *v, *X, *diag;
  const uint8_t *hasd;
  double *Y;
  int T, TS;  // vector count, row stride
  __device__ Acc zero() const {
    Acc a;
#pragma unroll
    for (int t = 0; t < N; t++) a.v[t] = 0.0;
    return a;
  }
  __device__ Acc merge(const Acc &a, const Acc &b) const {
    Acc r;
#pragma unroll
    for (int t = 0; t < N; t++) r.v[t] = a.v[t] + b.v[t];
    return r;
  }
  __device__ Acc shfl_down(const Acc &a, int off) const {
    Acc r;
#pragma unroll
    for (int t = 0; t < N; t++) r.v[t] = __shfl_down(a.v[t], off, WAVE);
    return r;
  }
  __device__ Acc shfl0(const Acc &a) const {
    Acc r;
#pragma unroll
    for (int t = 0; t < N; t++) r.v[t] = __shfl(a.v[t], 0, WAVE);
    return r;
  }
  __device__ void add(Acc &acc, int, int k) const {
    const double a = v[k];
    const double *xj = X + (int64_t)ci[k] * TS;
    if (TS & 1) {
#pragma unroll
      for (int t = 0; t < N; t++)
        if (t < T) acc.v[t] += (a < 0) ? a * xj[t] : 0.0;
    } else {
      // an even row stride keeps every row 16-byte aligned: the row is read
      // two vectors per load (the odd half of a last pair may be padding
      // inside the row, never used); the products and sums are unchanged
#pragma unroll
      for (int t = 0; t < N; t += 2)
        if (t < T) {
          const double2 p = *(const double2 *)(xj + t);
          acc.v[t] += (a < 0) ? a * p.x : 0.0;
          if (t + 1 < N && t + 1 < T) acc.v[t + 1] += (a < 0) ? a * p.y : 0.0;
        }
    }
  }
  __device__ void finish(int row, const Acc &acc) const {
    const double w = 2.0 / 3.0;
    const double d = diag[row];
#pragma unroll
    for (int t = 0; t < N; t++) {
      if (t >= T) break;
      double x = X[(int64_t)row * TS + t];
      Y[(int64_t)row * TS + t] = hasd[row] ? w * (0.0 - acc.v[t]) / d + (1 - w) * x : x;
    }
  }
};

__global__ void k_diag(int n, const int32_t *rp, const int32_t *ci, const double *v, double *diag,
                       uint8_t *hasd) {
  int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  bool f;
  double d = row_diag(rp, ci, v, row, &f);
  diag[row] = d;
  hasd[row] = f ? 1 : 0;
}

__global__ void k_tv_norm(int n, int T, int TS, const double *X, double *nrm) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0;
  for (int t = 0; t < T; t++) s += X[(int64_t)i * TS + t] * X[(int64_t)i * TS + t];
  nrm[i] = s;
}

// ---- a9: affinity_matrix (aggregation.hpp:276-404) -------------------------
struct OpAffinity : DblMax {
  const int32_t *ci;
  const double *v, *X, *nrm;
  int T, TS, metric;  // vector count, row stride
  double *raw, *rowmax;
  __device__ void add(double &acc, int row, int k) const {
    const int c = ci[k];
    double r;
    if (c == row) {
      r = (metric == LIGMG_STRENGTH_NONE) ? (v[k] < 0 ? -v[k] : 0.0) : 0.0;
    } else if (metric == LIGMG_STRENGTH_AFFINITY) {
      const double *xc = X + (int64_t)c * TS, *xr = X + (int64_t)row * TS;
      double top = 0;
      if (TS & 1) {
        for (int t = 0; t < T; t++) top += xc[t] * xr[t];
      } else {  // 16-byte rows reads, as in OpJacobiTV; same order of sums
#pragma unroll
        for (int t = 0; t < MAXT; t += 2)
          if (t < T) {
            const double2 a = *(const double2 *)(xc + t), b = *(const double2 *)(xr + t);
            top += a.x * b.x;
            if (t + 1 < T) top += a.y * b.y;
          }
      }
      r = top * top / (nrm[c] * nrm[c] * nrm[row] * nrm[row]);
    } else if (metric == LIGMG_STRENGTH_ALGEBRAIC_DISTANCE) {
      const double *xc = X + (int64_t)c * TS, *xr = X + (int64_t)row * TS;
      double m = 0;
      for (int t = 0; t < T; t++) {
        double d = fabs(xc[t] - xr[t]);
        m = (m < d) ? d : m;
      }
      r = 1 / m;
    } else {
      r = v[k] < 0 ? -v[k] : 0.0;
    }
    raw[k] = r;
    acc = (acc < r) ? r : acc;
  }
  __device__ void finish(int row, double acc) const { rowmax[row] = acc; }
};

// position of the transposed entry: tpos[k] for k = (r, c) is the index of
// (c, r) in row c (the pattern is structurally symmetric)
struct OpTpos : NoAcc {
  const int32_t *rp, *ci;
  int32_t *tpos, *bad;
  __device__ void add(int &, int row, int k) const {
    int c = ci[k];
    int lo = rp[c], hi = rp[c + 1] - 1, f = -1;
    while (lo <= hi) {
      int mid = (lo + hi) >> 1;
      int x = ci[mid];
      if (x == row) {
        f = mid;
        break;
      }
      if (x < row) lo = mid + 1;
      else hi = mid - 1;
    }
    tpos[k] = f;
    if (f < 0) atomicOr(bad, 1);
  }
  __device__ void finish(int, int) const {}
};

// (column, row | index) of every stored entry, for transpose_positions
struct OpTposKeys : NoAcc {
  const int32_t *ci;
  uint64_t *keys, *vals;
  __device__ void add(int &, int row, int k) const {
    keys[k] = (uint64_t)(uint32_t)ci[k];
    vals[k] = ((uint64_t)(uint32_t)row << 32) | (uint32_t)k;
  }
  __device__ void finish(int, int) const {}
};

// Sorted by column (stable), the entries of column c sit at rp[c] ..
// rp[c + 1] in ascending row order - the positions of their transposes -
// when the pattern is structurally symmetric; anything else sets *bad.
// The item at sorted position p is entry k = (row, c), and position p holds
// (c, row): p and k are each other's transposed positions, so the table is
// written at p, in order (tpos is its own inverse).
__global__ void k_tpos_scatter(int64_t m, const int32_t *rp, const int32_t *ci, const uint64_t *keys,
                               const uint64_t *vals, int32_t *tpos, int32_t *bad) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < m; p += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c = (int32_t)keys[p];
    const int32_t row = (int32_t)(vals[p] >> 32);
    const int32_t k = (int32_t)(vals[p] & 0xffffffffu);
    tpos[p] = k;
    if (p < rp[c] || p >= rp[c + 1] || ci[p] != row) atomicOr(bad, 1);
  }
}

struct OpColMax : DblMax {
  const int32_t *tpos;
  const double *raw;
  double *colmax;
  __device__ void add(double &acc, int, int k) const {
    double r = raw[tpos[k]];
    acc = (acc < r) ? r : acc;
  }
  __device__ void finish(int row, double acc) const { colmax[row] = acc; }
};

struct OpScale : NoAcc {
  const int32_t *ci;
  const double *raw, *rowmax, *colmax;
  double thresh;
  int scaling;
  double *S;
  __device__ void add(int &, int row, int k) const {
    double s = raw[k];
    if (scaling == LIGMG_AFF_MAX) {
      double a = rowmax[row], b = colmax[ci[k]];
      double den = (a < b) ? b : a;
      s = den > 0 ? s / den : 0.0;
    }
    S[k] = (s < thresh || s != s) ? 0.0 : s;
  }
  __device__ void finish(int, int) const {}
};

// ---- a10: col_sum = Reduce(Column, plus, 0.0): each column folded in
// ascending row order (aggregation.hpp:863-864), read through the
// transposed-entry positions.  Row sums (Reduce(Row), the weighted degree
// of an adjacency matrix, main.cpp:86) fold each row in ascending column
// order.  Both run through run_rows_seq.
struct OpRowSum : DblSum {
  const double *v;
  double *out;
  __device__ void add(double &acc, int, int k) const { acc += v[k]; }
  __device__ void finish(int row, double acc) const { out[row] = acc; }
};

struct OpColSum : DblSum {
  const int32_t *tpos;
  const double *S;
  double *cs;
  __device__ void add(double &acc, int, int k) const { acc += S[tpos[k]]; }
  __device__ void finish(int row, double acc) const { cs[row] = acc; }
};

// ---- a11: vote with AggSR (aggregation.hpp:406-518).  The left fold with
//      ties to the later operand equals the maximum of (priority, weight,
//      position), which any reduction order reproduces.
struct VoteAcc {
  int prio, pos;  // pos: the stored entry of the best neighbour so far
  double w;
};
struct OpVote {
  using Acc = VoteAcc;
  const int32_t *ci, *state, *ind;
  const double *S;
  double thresh;
  int nn;  // neighbor_of_neighbor: a SeedNeighbor attracts like its seed
  int8_t *vst;
  int32_t *vind;
  double *vw;
  __device__ Acc zero() const { return {-1, -1, -1.0}; }
  // only Undecided vertices vote; the vote of any other row is never read
  // (k_voters, k_mark) in the iteration that computes it
  __device__ bool skip(int row) const { return state[row] != ST_UNDECIDED; }
  __device__ static bool better(const Acc &b, const Acc &a) {  // b beats a
    if (b.prio != a.prio) return b.prio > a.prio;
    if (b.w != a.w) return b.w > a.w;
    return b.pos > a.pos;
  }
  __device__ Acc merge(const Acc &a, const Acc &b) const { return better(b, a) ? b : a; }
  __device__ Acc shfl_down(const Acc &a, int off) const {
    return {__shfl_down(a.prio, off, WAVE), __shfl_down(a.pos, off, WAVE), __shfl_down(a.w, off, WAVE)};
  }
  __device__ Acc shfl0(const Acc &a) const {
    return {__shfl(a.prio, 0, WAVE), __shfl(a.pos, 0, WAVE), __shfl(a.w, 0, WAVE)};
  }
  __device__ void add(Acc &acc, int, int k) const {
    const double s = S[k];
    if (s == 0.0) return;
    const int j = ci[k];
    const int st = state[j];
    int p;
    if (st == ST_SEED || (nn && st == ST_SEEDNEIGHBOR)) p = (s > thresh) ? PRIO_SEED : PRIO_DECIDED;
    else if (st == ST_UNDECIDED) p = PRIO_UNDECIDED;
    else p = PRIO_DECIDED;
    Acc m = {p, k, s};
    if (better(m, acc)) acc = m;
  }
  __device__ void finish(int row, const Acc &acc) const {
    if (acc.prio < 0) {
      vst[row] = PRIO_DECIDED;
      vind[row] = -1;
      vw[row] = -1.0;
    } else {
      vst[row] = (int8_t)acc.prio;
      vind[row] = ind[ci[acc.pos]];  // the winner's aggregate index, read once
      vw[row] = acc.w;
    }
  }
};

// ---- a12: tally_votes (aggregation.hpp:522-561) ----------------------------
__global__ void k_voters(int n, const int32_t *state, const int32_t *vind, const double *vw,
                         double thresh, int32_t *flag, int32_t *vf) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool ok = state[i] == ST_UNDECIDED && vind[i] >= 0 && vw[i] >= thresh && vind[i] != i;
  flag[i] = ok ? 1 : 0;
  vf[i] = ok ? vind[i] : -1;
}

__global__ void k_voter_pairs(int64_t m, const int32_t *idx, const int32_t *vind, const double *vw,
                              uint64_t *keys, uint64_t *vals) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < m;
       t += (int64_t)gridDim.x * blockDim.x) {
    int i = idx[t];
    keys[t] = (uint64_t)vind[i];
    vals[t] = (uint64_t)__double_as_longlong(vw[i]);
  }
}

__global__ void k_scatter_sums(int64_t m, const uint64_t *keys, const double *sums, double *nv) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < m;
       t += (int64_t)gridDim.x * blockDim.x)
    nv[keys[t]] = sums[t];
}

__global__ void k_add_votes(int n, const double *nv, double *votes) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) votes[i] = nv[i] + votes[i];
}

// ---- a13: mark_nodes (aggregation.hpp:702-750) -----------------------------
__global__ void k_mark(int n, int first, const int32_t *state, const int32_t *ind, const int8_t *vst,
                       const int32_t *vind, const int32_t *vf, const double *votes,
                       const double *cs, double votes_to_seed, double max_agg, int nn, int32_t *nstate,
                       int32_t *nind) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int joined = nn ? ST_SEEDNEIGHBOR : ST_DECIDED;
  int st = state[i], id = ind[i];
  if (st == ST_DECIDED || st == ST_DECIDEDROOT || st == ST_SEEDNEIGHBOR) {
  } else if (st == ST_SEED) {
    if (votes[i] > max_agg) st = ST_DECIDEDROOT;
  } else if (vf[i] >= 0 && first && vf[vf[i]] == i) {
    if (i < vf[i]) {
      st = ST_SEED;
      id = i;
    } else {
      st = joined;
      id = vf[i];
    }
  } else if (vst[i] == PRIO_SEED) {
    st = joined;
    id = vind[i];
  } else if (votes[i] >= votes_to_seed * cs[i]) {
    st = ST_SEED;
    id = i;
  }
  nstate[i] = st;
  nind[i] = id;
}

__device__ __forceinline__ bool aggregated(int st) {
  return st == ST_SEED || st == ST_DECIDED || st == ST_DECIDEDROOT || st == ST_SEEDNEIGHBOR;
}

// ---- reject_aggregates (aggregation.hpp:622-670).  An aggregate's size
// counts its root (a Seed or DecidedRoot) and its members, so the
// reference's |members|^2 / (size - 1) first members are all of them: every
// member of an aggregate above max_size returns to {-1, Undecided} and the
// root becomes DecidedRoot.
__global__ void k_agg_sizes(int n, const int32_t *state, const int32_t *ind, int32_t *sizes) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && aggregated(state[i])) atomicAdd(&sizes[ind[i]], 1);
}
__global__ void k_reject_large(int n, const int32_t *sizes, int64_t max_size, int32_t *state, int32_t *ind) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int st = state[i];
  if ((st == ST_DECIDED || st == ST_SEEDNEIGHBOR) && sizes[ind[i]] > max_size) {
    state[i] = ST_UNDECIDED;
    ind[i] = -1;
  } else if ((st == ST_SEED || st == ST_DECIDEDROOT) && ind[i] == i && sizes[i] > max_size) {
    state[i] = ST_DECIDEDROOT;
  }
}

// ---- ext_int_ratio (aggregation.hpp:563-620): weight of an aggregated
// vertex's off-diagonal edges leaving (ext) / staying in (!ext) its
// aggregate, sequential row sums (the reference's loop runs over each row in
// ascending column order) with the other entries +0
struct OpEdgeSplit : DblSum {
  const int32_t *ci, *state, *ind;
  const double *v;
  int ext;
  double *out;
  __device__ void add(double &acc, int row, int k) const {
    const int j = ci[k];
    double t = 0.0;
    if (aggregated(state[row]) && j != row) {
      const bool inside = aggregated(state[j]) && ind[j] == ind[row];
      if (inside != (ext != 0)) t = -v[k];
    }
    acc += t;
  }
  __device__ void finish(int row, double acc) const { out[row] = acc; }
};
// conductance_rejection (aggregation.hpp:752-776)
__global__ void k_conductance(int n, const double *ext, const double *in, double ratio, int to_seed,
                              int32_t *state, int32_t *ind) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int st = state[i];
  if (ext[i] / in[i] > ratio && (st == ST_DECIDED || st == ST_SEEDNEIGHBOR)) {
    state[i] = to_seed ? ST_SEED : ST_UNDECIDED;
    ind[i] = i;
  }
}
// a vertex a rejection left undecided with no target is its own aggregate
__global__ void k_self_unassigned(int n, int32_t *ind) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && ind[i] < 0) ind[i] = i;
}

// ---- a14: R with zero rows removed -> coarse id = rank of the root
__global__ void k_mark_roots(int n, const int32_t *ind, int32_t *root) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) root[ind[i]] = 1;
}
__global__ void k_agg_ids(int n, const int32_t *ind, const int32_t *cid, int32_t *agg) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) agg[i] = cid[ind[i]];
}

__global__ void k_init_states(int n, int32_t *state, int32_t *ind) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    state[i] = ST_UNDECIDED;
    ind[i] = i;
  }
}

// ---- aggregation Galerkin entries (agg_i, agg_j, L_ij) in CSR order
struct OpGalAgg : NoAcc {
  const int32_t *ci, *agg;
  const double *v;
  int cb;
  uint64_t *keys, *vals;
  __device__ void add(int &, int row, int k) const {
    keys[k] = ((uint64_t)agg[row] << cb) | (uint64_t)agg[ci[k]];
    vals[k] = (uint64_t)__double_as_longlong(v[k]);
  }
  __device__ void finish(int, int) const {}
};

__global__ void k_member_pairs(int n, const int32_t *agg, uint64_t *keys, uint64_t *vals) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    keys[i] = (uint64_t)agg[i];
    vals[i] = (uint64_t)i;
  }
}

__global__ void k_members_unpack(int n, const uint64_t *vals, int32_t *mem) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) mem[i] = (int32_t)vals[i];
}

unsigned g1(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

}  // namespace

void row_sums_seq(int64_t n, const int32_t *rp, const double *v, double *out, hipStream_t s) {
  HugeRows Hr;
  find_huge(n, rp, Hr, s);
  run_rows_seq(n, rp, OpRowSum{{}, v, out}, Hr, s);
}

// ----------------------------------------------------------------------------
// R = P^T (rows of R = columns of P, in ascending row order of P), with the
// FP32 values too when P has them.
// ----------------------------------------------------------------------------
void transpose_p(const DCsr &P, int64_t ncols, DCsr &R, hipStream_t s) {
  const int64_t n = P.n, nnzP = P.nnz;
  DevBuf<uint64_t> keys(nnzP), vals(nnzP);
  const int rb = std::max(1, bits_for((uint64_t)n));
  const int cb = std::max(1, bits_for((uint64_t)ncols));
  hipLaunchKernelGGL(k_p_transpose_keys, dim3(g1(n)), dim3(256), 0, s, (int)n, P.rp.p, P.ci.p, P.v.p, rb, keys.p,
                     vals.p);
  LG_LAUNCH_CHECK();
  DevBuf<uint64_t> keys32, vals32;
  if (P.v32.p) {
    keys32.alloc(nnzP);
    vals32.alloc(nnzP);
    hipLaunchKernelGGL(k_p_transpose_keys32, dim3(g1(n)), dim3(256), 0, s, (int)n, P.rp.p, P.ci.p, P.v32.p, rb,
                       keys32.p, vals32.p);
    LG_LAUNCH_CHECK();
    radix_sort_pairs(keys32, vals32, nnzP, rb + cb, s);
  }
  radix_sort_pairs(keys, vals, nnzP, rb + cb, s);
  R.n = ncols;
  R.ncols = n;
  R.nnz = nnzP;
  R.rp.alloc(ncols + 1);
  R.ci.alloc(std::max<int64_t>(nnzP, 1));
  R.v.alloc(std::max<int64_t>(nnzP, 1));
  hipLaunchKernelGGL(k_sorted_rowptr, dim3(grid_for(nnzP + 1, 256, 8192)), dim3(256), 0, s, keys.p, nnzP, rb, ncols,
                     R.rp.p);
  LG_LAUNCH_CHECK();
  if (nnzP) {
    hipLaunchKernelGGL(k_unpack_pairs, dim3(grid_for(nnzP, 256, 8192)), dim3(256), 0, s, keys.p, vals.p, nnzP, rb,
                       R.ci.p, R.v.p);
    LG_LAUNCH_CHECK();
  }
  if (P.v32.p) {
    R.v32.alloc(std::max<int64_t>(nnzP, 1));
    if (nnzP) {
      hipLaunchKernelGGL(k_unpack_vals32, dim3(grid_for(nnzP, 256, 8192)), dim3(256), 0, s, vals32.p, nnzP,
                         R.v32.p);
      LG_LAUNCH_CHECK();
    }
  }
}

// Member lists of the aggregates (nc x n, unit values): row c lists the
// vertices i with agg[i] = c in ascending order.
void build_members(const int32_t *agg, int64_t n, int64_t nc, DCsr &M, hipStream_t s) {
  DevBuf<uint64_t> keys(n), vals(n);
  hipLaunchKernelGGL(k_member_pairs, dim3(g1(n)), dim3(256), 0, s, (int)n, agg, keys.p, vals.p);
  LG_LAUNCH_CHECK();
  radix_sort_pairs(keys, vals, n, std::max(1, bits_for((uint64_t)nc)), s);
  M.n = nc;
  M.ncols = n;
  M.nnz = n;
  M.rp.alloc(nc + 1);
  M.ci.alloc(std::max<int64_t>(n, 1));
  hipLaunchKernelGGL(k_sorted_rowptr, dim3(grid_for(n + 1, 256, 8192)), dim3(256), 0, s, keys.p, n, 0, nc, M.rp.p);
  LG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_members_unpack, dim3(g1(n)), dim3(256), 0, s, (int)n, vals.p, M.ci.p);
  LG_LAUNCH_CHECK();
}

// ----------------------------------------------------------------------------
// Contraction: (key, value) list -> CSR of nrows rows.
// ----------------------------------------------------------------------------
void contract_to_csr(DevBuf<uint64_t> &keys, DevBuf<uint64_t> &vals, int64_t m, int64_t nrows,
                     int64_t ncols, DCsr &out, hipStream_t s) {
  const int cb = std::max(1, bits_for((uint64_t)ncols));
  const int rb = std::max(1, bits_for((uint64_t)nrows));
  radix_sort_pairs(keys, vals, m, cb + rb, s);
  DevBuf<uint64_t> uk(m);
  DevBuf<double> us(m);
  int64_t nu = segmented_canon_sum(keys.p, vals.p, m, uk.p, us.p, s);
  keys.release();
  vals.release();
  out.n = nrows;
  out.ncols = ncols;
  out.nnz = nu;
  out.rp.alloc(nrows + 1);
  out.ci.alloc(std::max<int64_t>(nu, 1));
  out.v.alloc(std::max<int64_t>(nu, 1));
  hipLaunchKernelGGL(k_sorted_rowptr, dim3(grid_for(nu + 1, 256, 8192)), dim3(256), 0, s, uk.p, nu, cb, nrows,
                     out.rp.p);
  LG_LAUNCH_CHECK();
  if (nu) {
    hipLaunchKernelGGL(k_key_cols, dim3(grid_for(nu, 256, 8192)), dim3(256), 0, s, uk.p, nu, cb, out.ci.p);
    LG_LAUNCH_CHECK();
  }
  d2d(out.v.p, us.p, nu, s);
}

// ----------------------------------------------------------------------------
// Elimination level (elimination.hpp:156-345, multigrid.hpp:132-149)
// ----------------------------------------------------------------------------
bool setup_elimination(const DCsr &A, const ligmg_options &o, DLevel &L, DCsr &coarse,
                       hipStream_t s) {
  const int64_t n = A.n;
  const HugeRows &Hr = huge_of(A, s);
  DevBuf<int32_t> deg(n), elim(n);
  run_rows(n, A.rp.p, OpDeg{{}, A.v.p, deg.p}, Hr, s);
  run_rows(n, A.rp.p, OpElimMin{A.ci.p, deg.p, A.v.p, o.elim_degree, elim.p}, Hr, s);
  // kept flags -> new ids (-1 for an eliminated vertex)
  DevBuf<int32_t> kept(n), newid(n);
  hipLaunchKernelGGL(k_kept_flags, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, n, elim.p, kept.p);
  LG_LAUNCH_CHECK();
  int64_t n_kept = exclusive_scan_i32(kept.p, newid.p, n, s);
  if ((double)n_kept > (double)n * o.elim_factor) return false;  // elimination.hpp:226-228
  hipLaunchKernelGGL(k_mask_ids, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, n, kept.p, newid.p);
  LG_LAUNCH_CHECK();
  {
    DevBuf<int32_t> bad(1);
    fill_i32(bad.p, 0, 1, s);
    hipLaunchKernelGGL(k_elim_check, dim3(g1(n)), dim3(256), 0, s, (int)n, A.rp.p, A.ci.p, newid.p,
                       bad.p);
    LG_LAUNCH_CHECK();
    int32_t hb = 0;
    to_host(&hb, bad.p, 1, s);
    LG_HIP(hipStreamSynchronize(s));
    LG_CHECK(hb == 0, LIGMG_ERR_NUMERIC,
             "eliminated vertices are adjacent (zero-valued stored entries in a row)");
  }
  L.kind = LIGMG_LEVEL_ELIM;
  L.n = n;
  L.nc = n_kept;
  L.new_id = std::move(newid);
  // P and q
  DevBuf<int32_t> pcnt(n + 1);
  fill_i32(pcnt.p, 0, n + 1, s);
  hipLaunchKernelGGL(k_p_count, dim3(g1(n)), dim3(256), 0, s, (int)n, A.rp.p, A.ci.p, L.new_id.p, pcnt.p);
  LG_LAUNCH_CHECK();
  int64_t nnzP = exclusive_scan_i32(pcnt.p, pcnt.p, n + 1, s);
  L.P.n = n;
  L.P.ncols = n_kept;
  L.P.nnz = nnzP;
  L.P.rp = std::move(pcnt);
  L.P.ci.alloc(nnzP);
  L.P.v.alloc(nnzP);
  L.q.alloc(n);
  hipLaunchKernelGGL(k_p_fill, dim3(g1(n)), dim3(256), 0, s, (int)n, A.rp.p, A.ci.p, A.v.p, L.new_id.p,
                     L.P.rp.p, L.P.ci.p, L.P.v.p, L.q.p);
  LG_LAUNCH_CHECK();
  L.nnzP = nnzP;
  L.nnzQ = n - n_kept;
  // R = P^T
  transpose_p(L.P, n_kept, L.R, s);
  // Schur complement by contraction
  DevBuf<int32_t> cnt(n + 1);
  fill_i32(cnt.p, 0, n + 1, s);
  run_rows(n, A.rp.p, OpSchurCount{{}, A.rp.p, A.ci.p, L.new_id.p, cnt.p}, Hr, s);
  int64_t m = exclusive_scan_i32(cnt.p, cnt.p, n + 1, s);
  DevBuf<uint64_t> keys(m), vals(m);
  const int cb = std::max(1, bits_for((uint64_t)n_kept));
  hipLaunchKernelGGL(k_schur_fill, dim3(g1(n)), dim3(256), 0, s, (int)n, A.rp.p, A.ci.p, A.v.p,
                     L.new_id.p, cnt.p, L.P.rp.p, L.P.v.p, cb, keys.p, vals.p);
  LG_LAUNCH_CHECK();
  if (Hr.nmid) {
    hipLaunchKernelGGL(k_schur_rows_fill, dim3((unsigned)((Hr.nmid + 3) / 4)), dim3(256), 0, s, Hr.nmid,
                       Hr.mrow.p, A.rp.p, A.ci.p, A.v.p, L.new_id.p, cnt.p, cb, keys.p, vals.p);
    LG_LAUNCH_CHECK();
  }
  if (Hr.nchunks) {
    // hub rows: one wavefront per chunk
    DevBuf<int32_t> ccnt(Hr.nchunks + 1), cpos(Hr.nchunks + 1);
    const unsigned gw = (unsigned)((Hr.nchunks + 3) / 4);
    hipLaunchKernelGGL(k_schur_chunk_count, dim3(gw), dim3(256), 0, s, Hr.nchunks, Hr.cstart.p, Hr.cend.p,
                       Hr.crow.p, A.ci.p, L.new_id.p, ccnt.p);
    LG_LAUNCH_CHECK();
    fill_i32(ccnt.p + Hr.nchunks, 0, 1, s);
    exclusive_scan_i32(ccnt.p, cpos.p, Hr.nchunks + 1, s);
    hipLaunchKernelGGL(k_schur_chunk_fill, dim3(gw), dim3(256), 0, s, Hr.nchunks, Hr.cstart.p, Hr.cend.p,
                       Hr.crow.p, A.rp.p, A.ci.p, A.v.p, L.new_id.p, cnt.p, cpos.p, cb, keys.p, vals.p);
    LG_LAUNCH_CHECK();
  }
  contract_to_csr(keys, vals, m, n_kept, n_kept, coarse, s);
  return true;
}

void make_elim32(DLevel &L, hipStream_t s) {
  const DCsr &A = L.A;
  LG_CHECK(A.v32.p && A.d32.p, LIGMG_ERR_INVALID, "make_elim32: level matrix has no FP32 Laplacian form");
  const int64_t n = A.n, nnzP = L.P.nnz, n_kept = L.nc;
  L.P.v32.alloc(std::max<int64_t>(nnzP, 1));
  L.q32.alloc(std::max<int64_t>(n, 1));
  const int32_t *newid = L.new_id_s.p ? L.new_id_s.p : L.new_id.p;  // in the order of A's rows
  hipLaunchKernelGGL(k_p_fill32, dim3(g1(n)), dim3(256), 0, s, (int)n, A.rp.p, A.ci.p, A.v32.p, A.d32.p, newid,
                     L.P.rp.p, L.P.v32.p, L.q32.p);
  LG_LAUNCH_CHECK();
  // R32 = P32^T with the same (stable) transposition that built R
  DevBuf<uint64_t> keys(nnzP), vals(nnzP);
  const int rb = std::max(1, bits_for((uint64_t)n));
  const int cb = std::max(1, bits_for((uint64_t)n_kept));
  hipLaunchKernelGGL(k_p_transpose_keys32, dim3(g1(n)), dim3(256), 0, s, (int)n, L.P.rp.p, L.P.ci.p,
                     L.P.v32.p, rb, keys.p, vals.p);
  LG_LAUNCH_CHECK();
  radix_sort_pairs(keys, vals, nnzP, rb + cb, s);
  L.R.v32.alloc(std::max<int64_t>(nnzP, 1));
  if (nnzP) {
    hipLaunchKernelGGL(k_unpack_vals32, dim3(grid_for(nnzP, 256, 8192)), dim3(256), 0, s, vals.p, nnzP,
                       L.R.v32.p);
    LG_LAUNCH_CHECK();
  }
}

// ----------------------------------------------------------------------------
// Strength of connection (affinity_matrix + col_sum)
// ----------------------------------------------------------------------------
namespace {
// v[k] / (its row's sum), 0 where the sum is not positive
struct OpDivRows : NoAcc {
  const double *rs;
  double *v;
  __device__ void add(int &, int row, int k) const {
    const double d = rs[row];
    v[k] = d > 0 ? v[k] / d : 0.0;
  }
  __device__ void finish(int, int) const {}
};
__global__ void k_div_cols(int64_t nnz, const int32_t *ci, const double *cs, double *v) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x) {
    const double d = cs[ci[k]];
    v[k] = d > 0 ? v[k] / d : 0.0;
  }
}
// (S^T + S) / 2 on a structurally symmetric pattern; tpos[k] is the
// position of the transposed entry
__global__ void k_sym_half(int64_t nnz, const int32_t *tpos, const double *v, double *out) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x)
    out[k] = (v[tpos[k]] + v[k]) / 2;
}
}  // namespace

namespace {
// tpos[k] = position of the transposed entry of k.  Large matrices: one
// stable sort of the entries by column (a few passes) instead of a binary
// search per entry; both give the same table.  Returns false when the
// pattern is not structurally symmetric.
bool transpose_positions(const DCsr &A, const HugeRows &Hr, int32_t *tpos, hipStream_t s) {
  const int64_t n = A.n, m = A.nnz;
  DevBuf<int32_t> bad(1);
  fill_i32(bad.p, 0, 1, s);
  if (m >= ((int64_t)1 << 20)) {
    DevBuf<uint64_t> keys(m), vals(m);
    run_rows(n, A.rp.p, OpTposKeys{{}, A.ci.p, keys.p, vals.p}, Hr, s);
    radix_sort_pairs(keys, vals, m, std::max(1, bits_for((uint64_t)n)), s);
    hipLaunchKernelGGL(k_tpos_scatter, dim3(grid_for(m, 256, 16384)), dim3(256), 0, s, m, A.rp.p, A.ci.p,
                       keys.p, vals.p, tpos, bad.p);
    LG_LAUNCH_CHECK();
  } else {
    run_rows(n, A.rp.p, OpTpos{{}, A.rp.p, A.ci.p, tpos, bad.p}, Hr, s);
  }
  int32_t hb = 0;
  to_host(&hb, bad.p, 1, s);
  LG_HIP(hipStreamSynchronize(s));
  return hb == 0;
}

// One accumulator width for every T <= MAXT.  Narrower ones (N = T) lower
// the register count and raise occupancy, but measured slower (7.6% at
// T = 10 on R-MAT 23): more gathers of the test-vector array (n * TS * 8
// bytes, larger than the last-level cache) in flight thrash the caches.
void jacobi_tv_sweep(int64_t n, const DCsr &A, const double *X, const double *diag, const uint8_t *hasd,
                     double *Y, int T, int TS, const HugeRows &Hr, hipStream_t s) {
  run_rows(n, A.rp.p, OpJacobiTV<MAXT>{A.ci.p, A.v.p, X, diag, hasd, Y, T, TS}, Hr, s);
}
}  // namespace

void compute_strength(const DCsr &A, const ligmg_options &o, int level_idx, DevBuf<double> &S,
                      DevBuf<double> &cs, DevBuf<int32_t> &tpos, hipStream_t s) {
  const int64_t n = A.n, nnz = A.nnz;
  const int T = o.num_tvs;
  const HugeRows &Hr = huge_of(A, s);
  int TS = 1;
  while (TS < T) TS <<= 1;
  DevBuf<double> X((size_t)n * TS), Y((size_t)n * TS), nrm(n), diag(n), raw(nnz), rowmax(n), colmax(n);
  DevBuf<uint8_t> hasd(n);
  tpos.alloc(nnz);
  hipLaunchKernelGGL(k_diag, dim3(g1(n)), dim3(256), 0, s, (int)n, A.rp.p, A.ci.p, A.v.p, diag.p, hasd.p);
  LG_LAUNCH_CHECK();
  if (o.strength_metric != LIGMG_STRENGTH_NONE) {
    DevBuf<unsigned long long> sums(T);
    LG_HIP(hipMemsetAsync(sums.p, 0, sizeof(unsigned long long) * T, s));
    hipLaunchKernelGGL(k_tv_sum, dim3(std::min<unsigned>(g1(n), 1024)), dim3(256), 0, s, (int)n, o.seed,
                       level_idx, T, sums.p);
    LG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tv_fill, dim3(g1(n)), dim3(256), 0, s, (int)n, o.seed, level_idx, T, TS, sums.p, X.p);
    LG_LAUNCH_CHECK();
    for (int sw = 0; sw < o.tvs_smooth; sw++) {
      jacobi_tv_sweep(n, A, X.p, diag.p, hasd.p, Y.p, T, TS, Hr, s);
      std::swap(X.p, Y.p);
    }
    hipLaunchKernelGGL(k_tv_norm, dim3(g1(n)), dim3(256), 0, s, (int)n, T, TS, X.p, nrm.p);
    LG_LAUNCH_CHECK();
  }
  run_rows(n, A.rp.p, OpAffinity{{}, A.ci.p, A.v.p, X.p, nrm.p, T, TS, o.strength_metric, raw.p, rowmax.p},
           Hr, s);
  LG_CHECK(transpose_positions(A, Hr, tpos.p, s), LIGMG_ERR_INVALID,
           "matrix pattern is not structurally symmetric");
  // Row / Column / RowColumn / Sym scaling (aggregation.hpp:350-366): row
  // sums, then column sums (ascending row), then (S^T + S) / 2; a zero sum
  // leaves its (zero) entries at 0
  const int sc = o.aff_scaling;
  if (sc == LIGMG_AFF_ROW || sc == LIGMG_AFF_ROWCOLUMN || sc == LIGMG_AFF_SYM) {
    DevBuf<double> rs(n);
    run_rows_seq(n, A.rp.p, OpRowSum{{}, raw.p, rs.p}, Hr, s);
    run_rows(n, A.rp.p, OpDivRows{{}, rs.p, raw.p}, Hr, s);
  }
  if (sc == LIGMG_AFF_COLUMN || sc == LIGMG_AFF_ROWCOLUMN) {
    DevBuf<double> csum(n);
    run_rows_seq(n, A.rp.p, OpColSum{{}, tpos.p, raw.p, csum.p}, Hr, s);
    hipLaunchKernelGGL(k_div_cols, dim3(grid_for(nnz, 256)), dim3(256), 0, s, nnz, A.ci.p, csum.p, raw.p);
    LG_LAUNCH_CHECK();
  }
  if (sc == LIGMG_AFF_SYM) {
    DevBuf<double> sym(std::max<int64_t>(nnz, 1));
    hipLaunchKernelGGL(k_sym_half, dim3(grid_for(nnz, 256)), dim3(256), 0, s, nnz, tpos.p, raw.p, sym.p);
    LG_LAUNCH_CHECK();
    raw = std::move(sym);
  }
  if (sc == LIGMG_AFF_MAX) run_rows(n, A.rp.p, OpColMax{{}, tpos.p, raw.p, colmax.p}, Hr, s);
  S.alloc(nnz);
  run_rows(n, A.rp.p, OpScale{{}, A.ci.p, raw.p, rowmax.p, colmax.p, o.thresh, o.aff_scaling, S.p}, Hr,
           s);
  cs.alloc(n);
  run_rows_seq(n, A.rp.p, OpColSum{{}, tpos.p, S.p, cs.p}, Hr, s);
}

// ----------------------------------------------------------------------------
// Voting (aggregation.hpp:838-964)
// ----------------------------------------------------------------------------
int64_t run_voting(const DCsr &A, const double *S, const double *cs, const ligmg_options &o,
                   DevBuf<int32_t> &agg, DevBuf<int32_t> &ind, DevBuf<int32_t> &state, hipStream_t s) {
  const int64_t n = A.n;
  const HugeRows &Hr = huge_of(A, s);
  state.alloc(n);
  ind.alloc(n);
  DevBuf<int32_t> nstate(n), nind(n), vind(n), vf(n), flag(n), idx(n);
  DevBuf<int8_t> vst(n);
  DevBuf<double> vw(n), votes(n), nv(n);
  hipLaunchKernelGGL(k_init_states, dim3(g1(n)), dim3(256), 0, s, (int)n, state.p, ind.p);
  LG_LAUNCH_CHECK();
  fill_f64(votes.p, 0.0, n, s);
  const int kb = std::max(1, bits_for((uint64_t)n));
  for (int iter = 0; iter < o.agg_iters; iter++) {
    const double th = o.agg_init_factor * std::pow(o.agg_iter_factor, iter);
    run_rows(n, A.rp.p, OpVote{A.ci.p, state.p, ind.p, S, th, o.neighbor_of_neighbor ? 1 : 0, vst.p, vind.p, vw.p},
             Hr, s);
    hipLaunchKernelGGL(k_voters, dim3(g1(n)), dim3(256), 0, s, (int)n, state.p, vind.p, vw.p, th, flag.p,
                       vf.p);
    LG_LAUNCH_CHECK();
    int64_t m = compact_flags(flag.p, n, idx.p, s);
    fill_f64(nv.p, 0.0, n, s);
    if (m > 0) {
      DevBuf<uint64_t> keys(m), vals(m), uk(m);
      DevBuf<double> us(m);
      hipLaunchKernelGGL(k_voter_pairs, dim3(grid_for(m, 256, 8192)), dim3(256), 0, s, m, idx.p, vind.p,
                         vw.p, keys.p, vals.p);
      LG_LAUNCH_CHECK();
      // stable sort: each target's voters stay in ascending voter index,
      // the order in which tally_votes adds them (aggregation.hpp:539-547)
      radix_sort_pairs(keys, vals, m, kb, s);
      int64_t nu = segmented_seq_sum(keys.p, vals.p, m, uk.p, us.p, s);
      hipLaunchKernelGGL(k_scatter_sums, dim3(grid_for(nu, 256, 8192)), dim3(256), 0, s, nu, uk.p, us.p,
                         nv.p);
      LG_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_add_votes, dim3(g1(n)), dim3(256), 0, s, (int)n, nv.p, votes.p);
    LG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_mark, dim3(g1(n)), dim3(256), 0, s, (int)n, iter == 0 ? 1 : 0, state.p, ind.p,
                       vst.p, vind.p, vf.p, votes.p, cs, (double)o.votes_to_seed,
                       (double)o.max_agg_size, o.neighbor_of_neighbor ? 1 : 0, nstate.p, nind.p);
    LG_LAUNCH_CHECK();
    std::swap(state.p, nstate.p);
    std::swap(ind.p, nind.p);
    if (o.max_agg_size > 0) {
      DevBuf<int32_t> sizes(n);
      fill_i32(sizes.p, 0, n, s);
      hipLaunchKernelGGL(k_agg_sizes, dim3(g1(n)), dim3(256), 0, s, (int)n, state.p, ind.p, sizes.p);
      LG_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_reject_large, dim3(g1(n)), dim3(256), 0, s, (int)n, sizes.p, (int64_t)o.max_agg_size,
                         state.p, ind.p);
      LG_LAUNCH_CHECK();
    }
    if (iter >= o.agg_iters / 2 && o.rejection_ratio >= 0) {
      LG_CHECK(A.v.p, LIGMG_ERR_UNSUPPORTED, "conductance rejection needs the level matrix values");
      DevBuf<double> ext(n), in(n);
      run_rows_seq(n, A.rp.p, OpEdgeSplit{{}, A.ci.p, state.p, ind.p, A.v.p, 1, ext.p}, Hr, s);
      run_rows_seq(n, A.rp.p, OpEdgeSplit{{}, A.ci.p, state.p, ind.p, A.v.p, 0, in.p}, Hr, s);
      hipLaunchKernelGGL(k_conductance, dim3(g1(n)), dim3(256), 0, s, (int)n, ext.p, in.p, o.rejection_ratio,
                         o.rejection_to_seed ? 1 : 0, state.p, ind.p);
      LG_LAUNCH_CHECK();
    }
  }
  hipLaunchKernelGGL(k_self_unassigned, dim3(g1(n)), dim3(256), 0, s, (int)n, ind.p);
  LG_LAUNCH_CHECK();
  DevBuf<int32_t> root(n), cid(n);
  fill_i32(root.p, 0, n, s);
  hipLaunchKernelGGL(k_mark_roots, dim3(g1(n)), dim3(256), 0, s, (int)n, ind.p, root.p);
  LG_LAUNCH_CHECK();
  int64_t nc = exclusive_scan_i32(root.p, cid.p, n, s);
  agg.alloc(n);
  hipLaunchKernelGGL(k_agg_ids, dim3(g1(n)), dim3(256), 0, s, (int)n, ind.p, cid.p, agg.p);
  LG_LAUNCH_CHECK();
  return nc;
}

void col_sums_of(const DCsr &A, const double *S, DevBuf<double> &cs, hipStream_t s) {
  const int64_t n = A.n;
  const HugeRows &Hr = huge_of(A, s);
  DevBuf<int32_t> tpos(A.nnz);
  LG_CHECK(transpose_positions(A, Hr, tpos.p, s), LIGMG_ERR_INVALID, "matrix pattern is not structurally symmetric");
  cs.alloc(n);
  run_rows_seq(n, A.rp.p, OpColSum{{}, tpos.p, S, cs.p}, Hr, s);
}

// ----------------------------------------------------------------------------
// Aggregation level (aggregation.hpp:838-964, multigrid.hpp:151-159)
// ----------------------------------------------------------------------------
void setup_aggregation(const DCsr &A, const ligmg_options &o, int level_idx, DLevel &L, DCsr &coarse,
                       hipStream_t s) {
  const int64_t n = A.n;
  L.kind = LIGMG_LEVEL_AGG;
  L.n = n;
  DevBuf<double> cs;
  compute_strength(A, o, level_idx, L.strength, cs, L.tpos, s);
  L.nc = run_voting(A, L.strength.p, cs.p, o, L.agg, L.ind, L.state, s);
  const int64_t nc = L.nc;
  // Galerkin R A P
  {
    DevBuf<uint64_t> keys(A.nnz), vals(A.nnz);
    const int cb = std::max(1, bits_for((uint64_t)nc));
    const HugeRows &Hr = huge_of(A, s);
    run_rows(n, A.rp.p, OpGalAgg{{}, A.ci.p, L.agg.p, A.v.p, cb, keys.p, vals.p}, Hr, s);
    contract_to_csr(keys, vals, A.nnz, nc, nc, coarse, s);
  }
  // member lists (restriction operator, unit values)
  build_members(L.agg.p, n, nc, L.M, s);
}

// Inverse diagonal of A (0 where the diagonal is missing or zero).
namespace {
// 1/diag per row, 0 for an isolated vertex (no nonzero off-diagonal) or a
// zero diagonal: the row machinery spreads hub rows over many wavefronts.
// The diagonal sits at one position, so the merged sum is that value.
struct DiagAcc {
  double d;
  int off;
};
struct OpDinv {
  using Acc = DiagAcc;
  const int32_t *ci;
  const double *v;
  double *dinv;
  float *dinv32;
  __device__ Acc zero() const { return {0.0, 0}; }
  __device__ Acc merge(const Acc &a, const Acc &b) const { return {a.d + b.d, a.off | b.off}; }
  __device__ Acc shfl_down(const Acc &a, int off) const {
    return {__shfl_down(a.d, off, WAVE), __shfl_down(a.off, off, WAVE)};
  }
  __device__ Acc shfl0(const Acc &a) const { return {__shfl(a.d, 0, WAVE), __shfl(a.off, 0, WAVE)}; }
  __device__ void add(Acc &acc, int row, int k) const {
    if (ci[k] == row) acc.d = v[k];
    else if (v[k] != 0) acc.off = 1;
  }
  __device__ void finish(int row, const Acc &acc) const {
    const double r = (acc.d != 0 && acc.off) ? 1.0 / acc.d : 0.0;
    dinv[row] = r;
    if (dinv32) dinv32[row] = (float)r;
  }
};

// FP32 Laplacian form: off-diagonals rounded to FP32 (the diagonal slot 0),
// the diagonal -(sum of the rounded off-diagonals) in FP64, so the FP32
// operator keeps the constant nullspace exactly.  Canonical row sums.
struct OpLap32 : DblSum {
  const int32_t *ci;
  const double *v;
  float *v32;
  double *d32;
  __device__ void add(double &acc, int row, int k) const {
    if (ci[k] == row) {
      v32[k] = 0.0f;
    } else {
      const float f = (float)v[k];
      v32[k] = f;
      acc += (double)f;
    }
  }
  __device__ void finish(int row, double acc) const { d32[row] = -acc; }
};
}  // namespace

void setup_dinv(const DCsr &A, DLevel &L, bool fp32, hipStream_t s) {
  const int64_t n = A.n;
  L.dinv.alloc(n);
  if (fp32) L.dinv32.alloc(n);
  run_rows(n, A.rp.p, OpDinv{A.ci.p, A.v.p, L.dinv.p, fp32 ? L.dinv32.p : nullptr}, huge_of(A, s), s);
}

void lap32_form(DCsr &A, hipStream_t s) {
  A.v32.alloc(std::max<int64_t>(A.nnz, 1));
  A.d32.alloc(std::max<int64_t>(A.n, 1));
  run_rows(A.n, A.rp.p, OpLap32{{}, A.ci.p, A.v.p, A.v32.p, A.d32.p}, huge_of(A, s), s);
}

// ----------------------------------------------------------------------------
// lambda_max(D^-1 A) by Jacobi-preconditioned CG + Lanczos tridiagonal
// (the CG eigen-estimator of the reference's Chebyshev, smoothing.hpp:158-165)
// ----------------------------------------------------------------------------
namespace {
// start vector entry of setup position perm[i] (perm null: i), so a level
// stored in solve layout starts from the same vector, permuted
__global__ void k_eig_init(int n, uint64_t seed, uint64_t stream, const int32_t *perm, double *r) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = ligmg_rng::u01(seed, stream, perm ? perm[i] : i);
}
__global__ void k_axpy_dev(int n, const double *alpha, double sign, const double *x, double *y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += sign * alpha[0] * x[i];
}
__global__ void k_mul(int n, const double *a, const double *b, double *c) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = a[i] * b[i];
}
__global__ void k_xpby_dev(int n, const double *x, const double *beta, double *p) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = x[i] + beta[0] * p[i];
}

// Scalars of the estimator.  They stay on the device; the recorded alpha and
// beta are read once, after the last step.  The stop rules are those of the
// loop `while gamma > 0: ...; if !(delta > 0) break`.
enum { EG_GAMMA = 0, EG_DELTA, EG_GN, EG_ALPHA, EG_BETA, EG_ACTIVE, EG_M, EG_N };

// after gamma = (r, z)
__global__ void k_eig_start(double *sc) {
  sc[EG_ACTIVE] = sc[EG_GAMMA] > 0 ? 1.0 : 0.0;
  sc[EG_M] = 0;
}
// after delta = (p, A p): alpha = gamma / delta, or stop
__global__ void k_eig_alpha(double *sc) {
  double a = 0;
  if (sc[EG_ACTIVE] != 0) {
    if (sc[EG_DELTA] > 0) a = sc[EG_GAMMA] / sc[EG_DELTA];
    else sc[EG_ACTIVE] = 0;
  }
  sc[EG_ALPHA] = a;
}
// after gn = (r, z): record (alpha, gn / gamma), gamma = gn, stop unless gn > 0
__global__ void k_eig_beta(double *sc, double *al, double *be) {
  double b = 0;
  if (sc[EG_ACTIVE] != 0) {
    const int m = (int)sc[EG_M];
    b = sc[EG_GN] / sc[EG_GAMMA];
    al[m] = sc[EG_ALPHA];
    be[m] = b;
    sc[EG_M] = m + 1;
    sc[EG_GAMMA] = sc[EG_GN];
    if (!(sc[EG_GN] > 0)) sc[EG_ACTIVE] = 0;
  }
  sc[EG_BETA] = b;
}
}  // namespace

double tridiag_max_eig(const std::vector<double> &d, const std::vector<double> &e) {
  const int m = (int)d.size();
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < m; i++) {
    double r = (i > 0 ? std::fabs(e[i - 1]) : 0) + (i + 1 < m ? std::fabs(e[i]) : 0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  auto below = [&](double x) {
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
    if (below(mid) == m) hi = mid;
    else lo = mid;
  }
  return hi;
}

double eig_estimate(const DCsr &A, const double *dinv, uint64_t seed, int level_idx, int iters,
                    const int32_t *perm, hipStream_t s) {
  const int64_t n = A.n;
  if (n == 0) return 1.0;
  RedPlan R;
  R.build(n, {0, n}, s);
  DevBuf<double> r(n), z(n), p(n), q(n), sc(EG_N), al(std::max(iters, 1)), be(std::max(iters, 1));
  hipLaunchKernelGGL(k_eig_init, dim3(g1(n)), dim3(256), 0, s, (int)n, seed,
                     ligmg_rng::STREAM_EIG + (uint64_t)level_idx, perm, r.p);
  LG_LAUNCH_CHECK();
  R.sums(r.p, sc.p, s);
  op_sub_block_mean(r.p, R, sc.p, s);
  hipLaunchKernelGGL(k_mul, dim3(g1(n)), dim3(256), 0, s, (int)n, dinv, r.p, z.p);
  d2d(p.p, z.p, n, s);
  auto dot = [&](const double *a, const double *b, int slot) {
    const double *A_[1] = {a}, *B_[1] = {b};
    R.dots(1, A_, B_, sc.p + slot, s);
  };
  dot(r.p, z.p, EG_GAMMA);
  hipLaunchKernelGGL(k_eig_start, dim3(1), dim3(1), 0, s, sc.p);
  LG_LAUNCH_CHECK();
  // a stopped estimator runs the remaining steps with alpha = beta = 0 and
  // records nothing more
  for (int it = 0; it < iters; it++) {
    op_spmv<double>(A, p.p, q.p, s);
    dot(p.p, q.p, EG_DELTA);
    hipLaunchKernelGGL(k_eig_alpha, dim3(1), dim3(1), 0, s, sc.p);
    hipLaunchKernelGGL(k_axpy_dev, dim3(g1(n)), dim3(256), 0, s, (int)n, sc.p + EG_ALPHA, -1.0, q.p, r.p);
    hipLaunchKernelGGL(k_mul, dim3(g1(n)), dim3(256), 0, s, (int)n, dinv, r.p, z.p);
    dot(r.p, z.p, EG_GN);
    hipLaunchKernelGGL(k_eig_beta, dim3(1), dim3(1), 0, s, sc.p, al.p, be.p);
    hipLaunchKernelGGL(k_xpby_dev, dim3(g1(n)), dim3(256), 0, s, (int)n, z.p, sc.p + EG_BETA, p.p);
    LG_LAUNCH_CHECK();
  }
  double h_sc[EG_N];
  std::vector<double> ha(al.n), hb(be.n);
  to_host(h_sc, sc.p, EG_N, s);
  to_host(ha.data(), al.p, al.n, s);
  to_host(hb.data(), be.p, be.n, s);
  LG_HIP(hipStreamSynchronize(s));
  const int m = (int)h_sc[EG_M];  // entries beyond m were never written
  if (m == 0) return 1.0;
  std::vector<double> d(m), e(m > 1 ? m - 1 : 0);
  for (int j = 0; j < m; j++) {
    d[j] = 1 / ha[j] + (j > 0 ? hb[j - 1] / ha[j - 1] : 0);
    if (j + 1 < m) e[j] = std::sqrt(hb[j]) / ha[j];
  }
  return tridiag_max_eig(d, e);
}

// ----------------------------------------------------------------------------
// Coarse level: dense pseudo-inverse per connected component,
// L_c^+ = (L_c + 1 1^T / n_c)^-1 - 1 1^T / n_c  (host, FP64 Cholesky).
// ----------------------------------------------------------------------------
std::vector<double> coarse_dense_spd(int64_t n, const std::vector<int32_t> &rp, const std::vector<int32_t> &ci,
                                     const std::vector<double> &v, std::vector<int32_t> &comp,
                                     std::vector<int64_t> &csize) {
  comp.assign(n, -1);
  csize.clear();
  std::vector<int64_t> stack;
  for (int64_t s0 = 0; s0 < n; s0++) {
    if (comp[s0] >= 0) continue;
    int32_t cid = (int32_t)csize.size();
    csize.push_back(0);
    stack.assign(1, s0);
    comp[s0] = cid;
    while (!stack.empty()) {
      int64_t u = stack.back();
      stack.pop_back();
      csize[cid]++;
      for (int32_t k = rp[u]; k < rp[u + 1]; k++) {
        int32_t w = ci[k];
        if (v[k] != 0 && comp[w] < 0) {
          comp[w] = cid;
          stack.push_back(w);
        }
      }
    }
  }
  std::vector<double> M((size_t)n * n, 0.0);
  for (int64_t i = 0; i < n; i++)
    for (int32_t k = rp[i]; k < rp[i + 1]; k++) M[i * n + ci[k]] += v[k];
  for (int64_t i = 0; i < n; i++)
    for (int64_t j = 0; j < n; j++)
      if (comp[i] == comp[j]) M[i * n + j] += 1.0 / (double)csize[comp[i]];
  for (int64_t i = 0; i < n; i++)
    for (int64_t j = i + 1; j < n; j++) {
      double a = 0.5 * (M[i * n + j] + M[j * n + i]);
      M[i * n + j] = M[j * n + i] = a;
    }
  return M;
}

std::vector<double> dense_pinv_host(int64_t n, const std::vector<int32_t> &rp,
                                    const std::vector<int32_t> &ci, const std::vector<double> &v) {
  // the hierarchy stops at min_size entries, so only coarsening that stalls
  // until max_levels leaves a larger coarse level
  LG_CHECK(n <= COARSE_DENSE_MAX, LIGMG_ERR_UNSUPPORTED,
           "the coarsest level has " + std::to_string(n) + " vertices; the dense coarse solve takes at most " +
               std::to_string(COARSE_DENSE_MAX) + " (coarsening stalled: check the aggregation options)");
  std::vector<int32_t> comp;
  std::vector<int64_t> csize;
  std::vector<double> M = coarse_dense_spd(n, rp, ci, v, comp, csize);
  // in-place Cholesky (lower), then inverse from G^-T G^-1
  std::vector<double> &G = M;
  for (int64_t j = 0; j < n; j++) {
    double sjj = G[j * n + j];
    for (int64_t k = 0; k < j; k++) sjj -= G[j * n + k] * G[j * n + k];
    LG_CHECK(sjj > 0, LIGMG_ERR_NUMERIC, "coarse matrix is not positive semidefinite");
    double g = std::sqrt(sjj);
    G[j * n + j] = g;
    for (int64_t i = j + 1; i < n; i++) {
      double t = G[i * n + j];
      for (int64_t k = 0; k < j; k++) t -= G[i * n + k] * G[j * n + k];
      G[i * n + j] = t / g;
    }
  }
  // W = G^-1 (lower triangular)
  std::vector<double> W((size_t)n * n, 0.0);
  for (int64_t c = 0; c < n; c++) {
    for (int64_t i = c; i < n; i++) {
      double t = (i == c) ? 1.0 : 0.0;
      for (int64_t k = c; k < i; k++) t -= G[i * n + k] * W[k * n + c];
      W[i * n + c] = t / G[i * n + i];
    }
  }
  // inv = W^T W
  std::vector<double> inv((size_t)n * n, 0.0);
  for (int64_t i = 0; i < n; i++)
    for (int64_t j = 0; j <= i; j++) {
      double t = 0;
      for (int64_t k = i; k < n; k++) t += W[k * n + i] * W[k * n + j];
      inv[i * n + j] = inv[j * n + i] = t;
    }
  for (int64_t i = 0; i < n; i++)
    for (int64_t j = 0; j < n; j++)
      if (comp[i] == comp[j]) inv[i * n + j] -= 1.0 / (double)csize[comp[i]];
  return inv;
}

}  // namespace lg
