// Dense pseudo-inverse of the coarsest level on the device: blocked
// Gauss-Jordan inversion of an SPD matrix, its rank-b updates on FP64 MFMA.
//
// The coarse operator L (a graph Laplacian, one nullspace vector per
// connected component C) is made SPD by adding 1_C 1_C^T / |C| for every
// component; the pseudo-inverse is the inverse of that matrix minus the same
// terms.  The reference solves its coarse level with a PETSc direct solver
// (multigrid.hpp:37-84, 217-257); this library applies a dense pseudo-inverse,
// factored once at setup.
//
// Gauss-Jordan by GB x GB blocks (A padded to np = GB * nb rows with an
// identity block), for k = 0 .. nb-1, with D = A_kk^-1:
//   A_kj <- D A_kj             (j != k)        row panel
//   A_ij <- A_ij - A_ik A_kj   (i, j != k)     trailing update (old A_ik)
//   A_ik <- -A_ik D            (i != k)        column panel
//   A_kk <- D
// leaves A^-1 in place.  No pivoting: the matrix is SPD, so every diagonal
// block met on the way is SPD (a Schur complement of an SPD matrix).  The
// trailing update reads and writes the whole matrix once per step, so the
// block width sets the update's flops per byte (GB / 8): at 128 the matrix
// cores, not HBM, bound it.
#include <cmath>
#include <vector>

#include "prims.hpp"
#include "setup.hpp"

namespace lg {

namespace {

constexpr int GB = COARSE_GJ_BLOCK;  // Gauss-Jordan block
constexpr int TL = 64;               // output tile of the panel and update kernels
constexpr int NQ = GB / 4;           // K slots of v_mfma_f64_16x16x4f64 per block
constexpr int DT = 512;              // threads of k_gj_diag
constexpr int DR = GB * GB / DT;     // rows of the diagonal block held per thread
constexpr int DQ = 2;                // pivots per step of k_gj_diag
static_assert(GB % TL == 0 && DT % GB == 0 && DR == GB / (DT / GB) && (DT / GB) % DQ == 0, "block shapes");
typedef double v4d __attribute__((ext_vector_type(4)));

// D = A_kk^-1 by Gauss-Jordan on DQ pivots a step (the DQ x DQ pivot block
// inverted in registers, then a rank-DQ update).  Thread t holds column
// t % GB of rows t / GB + (DT / GB) r (r = 0 .. DR-1) in registers for the
// whole inversion; each step publishes its pivot rows and columns through
// LDS (double-buffered: one barrier a step).  The steps are latency-bound
// (publish, barrier, pivot inverse), so fewer, wider steps are faster.  D is
// written to A_kk and to dinv.
__global__ void __launch_bounds__(DT) k_gj_diag(double *A, int64_t np, int k, double *dinv) {
  constexpr int RS = DT / GB;  // row phases
  __shared__ double prow[2][DQ][GB], pcol[2][DQ][GB];
  const int t = threadIdx.x, j = t % GB, w = t / GB;
  const int64_t o = (int64_t)k * GB;
  double m[DR];
#pragma unroll
  for (int r = 0; r < DR; r++) m[r] = A[(o + w + RS * r) * np + o + j];
  // the slot loop is unrolled (the pivot rows of a step share slot r0, a
  // static register index); the steps inside a slot are not
#pragma unroll
  for (int r0 = 0; r0 < DR; r0++) {
#pragma unroll 1
    for (int h = 0; h < RS / DQ; h++) {
      const int p = RS * r0 + DQ * h, b = (r0 * (RS / DQ) + h) & 1;
      const int wq = w - DQ * h, jq = j - p;  // my pivot row / column in this step, when in [0, DQ)
      if (wq >= 0 && wq < DQ) prow[b][wq][j] = m[r0];
      if (jq >= 0 && jq < DQ) {
#pragma unroll
        for (int r = 0; r < DR; r++) {
          pcol[b][jq][w * DR + r] = m[r];
          m[r] = 0.0;  // M[i][p + jq] <- 0 before the update below
        }
      }
      __syncthreads();
      // E = inverse of the pivot block (unpivoted: it is SPD)
      double E[DQ][DQ];
#pragma unroll
      for (int a = 0; a < DQ; a++)
#pragma unroll
        for (int c = 0; c < DQ; c++) E[a][c] = prow[b][a][p + c];
#pragma unroll
      for (int q = 0; q < DQ; q++) {
        const double rp = 1.0 / E[q][q];
#pragma unroll
        for (int c = 0; c < DQ; c++) E[q][c] = (c == q ? 1.0 : E[q][c]) * rp;
#pragma unroll
        for (int a = 0; a < DQ; a++) {
          if (a == q) continue;
          const double f = E[a][q];
#pragma unroll
          for (int c = 0; c < DQ; c++) E[a][c] = fma(-f, E[q][c], c == q ? 0.0 : E[a][c]);
        }
      }
      // the new pivot rows at my column: E times the old ones, E itself on
      // the pivot columns
      double pr[DQ];
#pragma unroll
      for (int a = 0; a < DQ; a++) {
        double s = 0.0, e = 0.0;
#pragma unroll
        for (int c = 0; c < DQ; c++) {
          s = fma(E[a][c], prow[b][c][j], s);
          e = (jq == c) ? E[a][c] : e;
        }
        pr[a] = (jq >= 0 && jq < DQ) ? e : s;
      }
      // every row minus its pivot-column entries times the new pivot rows,
      // then the pivot rows replaced by the new ones
#pragma unroll
      for (int r = 0; r < DR; r++) {
        double v = m[r];
#pragma unroll
        for (int a = 0; a < DQ; a++) v = fma(-pcol[b][a][w * DR + r], pr[a], v);
        m[r] = v;
      }
      if (wq >= 0 && wq < DQ) {
        double y = pr[0];
#pragma unroll
        for (int a = 1; a < DQ; a++) y = (wq == a) ? pr[a] : y;
        m[r0] = y;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < DR; r++) {
    const int i = w + RS * r;
    A[(o + i) * np + o + j] = m[r];
    dinv[i * GB + j] = m[r];
  }
}

// Operand layout of v_mfma_f64_16x16x4f64 (measured on gfx950): lane
// l = 16 lk + li supplies X[li][lk] and Y[lk][li] of the 16 x 4 / 4 x 16
// slices and holds C[lk + 4v][li], v = 0..3.  The kernels below give lane
// (li, lk) the K indices lk * NQ + q, q = 0 .. NQ-1, so each lane's X values
// are one contiguous run of a row; any assignment of K indices to slots gives
// the same sum up to rounding order.
//
// Y (GB rows x 16 NT columns, row stride ld, global) into LDS in slot order:
// S[(q * NT + tj) * 64 + 16 lk + li] = Y[lk * NQ + q][16 tj + li], which each
// MFMA step reads as 64 consecutive doubles.
template <int NT>
__device__ __forceinline__ void stage_y(double *S, const double *Y, int64_t ld) {
  for (int e = threadIdx.x; e < GB * 16 * NT; e += blockDim.x) {
    const int r = e / (16 * NT), c = e % (16 * NT);
    const int lk = r / NQ, q = r % NQ, tj = c / 16, li = c % 16;
    S[(q * NT + tj) * 64 + 16 * lk + li] = Y[(int64_t)r * ld + c];
  }
}
// this lane's run of X (the wave's 16 rows from X, row stride ld), times sign
__device__ __forceinline__ void load_x(double a[NQ], const double *X, int64_t ld, double sign) {
  const int lane = threadIdx.x & (WAVE - 1), li = lane & 15, lk = lane >> 4;
  const double *x = X + (int64_t)li * ld + lk * NQ;
#pragma unroll
  for (int q = 0; q < NQ; q++) a[q] = sign * x[q];
}
template <int NT>
__device__ __forceinline__ void mfma_slots(const double a[NQ], const double *S, v4d acc[NT]) {
  const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
  for (int q = 0; q < NQ; q++)
#pragma unroll
    for (int tj = 0; tj < NT; tj++)
      acc[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], S[(q * NT + tj) * 64 + lane], acc[tj], 0, 0, 0);
}
// the same products in groups of 8 slots that the scheduler keeps apart, so
// it does not hoist every LDS read of Y to the top (and run out of registers)
template <int NT>
__device__ __forceinline__ void mfma_slots_grouped(const double a[NQ], const double *S, v4d acc[NT]) {
  const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
  for (int g = 0; g < NQ; g += 8) {
#pragma unroll
    for (int q = g; q < g + 8; q++)
#pragma unroll
      for (int tj = 0; tj < NT; tj++)
        acc[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], S[(q * NT + tj) * 64 + lane], acc[tj], 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
  }
}
// C (16 rows x 16 NT columns at G, row stride ld) from / to the accumulators
template <int NT>
__device__ __forceinline__ void load_c(v4d acc[NT], const double *G, int64_t ld) {
  const int lane = threadIdx.x & (WAVE - 1), li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int tj = 0; tj < NT; tj++)
#pragma unroll
    for (int v = 0; v < 4; v++) acc[tj][v] = G[(int64_t)(lk + 4 * v) * ld + 16 * tj + li];
}
template <int NT>
__device__ __forceinline__ void store_c(const v4d acc[NT], double *G, int64_t ld) {
  const int lane = threadIdx.x & (WAVE - 1), li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int tj = 0; tj < NT; tj++)
#pragma unroll
    for (int v = 0; v < 4; v++) G[(int64_t)(lk + 4 * v) * ld + 16 * tj + li] = acc[tj][v];
}

// A_kj <- D A_kj for the TL-column tile blockIdx.x outside block column k:
// the GB x TL panel is staged in LDS (the workgroup overwrites it), D is read
// from dinv; wavefront w writes rows 32w .. 32w+31.
__global__ void __launch_bounds__(256) k_gj_rowpanel(double *A, int64_t np, int k, const double *dinv) {
  __shared__ double S[GB * TL];
  const int jt = blockIdx.x;
  if (jt / (GB / TL) == k) return;
  double *P = A + (int64_t)k * GB * np + (int64_t)jt * TL;
  stage_y<TL / 16>(S, P, np);
  __syncthreads();
  const int w = threadIdx.x / WAVE;
  for (int h = 0; h < GB / 64; h++) {
    const int r0 = 16 * (GB / 16 / 4) * w + 16 * h;  // GB / 4 rows per wavefront
    double a[NQ];
    load_x(a, dinv + (int64_t)r0 * GB, GB, 1.0);
    v4d acc[TL / 16] = {};
    mfma_slots<TL / 16>(a, S, acc);
    store_c<TL / 16>(acc, P + (int64_t)r0 * np, np);
  }
}

// A_ij <- A_ij - A_ik A_kj for the TL x TL tiles (it, jt) outside block row
// and column k with jt = c0 + blockIdx.x and it = r0 + blockIdx.y + s rstep
// < r1, leaving out the tiles of block (la, la) when la >= 0.  The workgroup
// keeps the GB x TL panel A_kj of column tile jt in LDS and walks its row
// tiles; wavefront w updates rows 16w .. 16w+15 of each, loading the next
// tile's operands while the matrix cores work on the current one.  The
// workgroups of one blockIdx.y walk the same row tiles at about the same
// time, so the column panel A_ik they read stays in L2.  This launch writes
// neither panel.
__global__ void __launch_bounds__(256, 2) k_gj_update(double *A, int64_t np, int k, int c0, int r0, int r1,
                                                      int rstep, int la) {
  __shared__ double S[GB * TL];
  constexpr int KT = GB / TL;  // tiles per block
  const int jt = c0 + blockIdx.x;
  if (jt / KT == k) return;
  const bool colla = la >= 0 && jt / KT == la;
  auto next_tile = [&](int it) {
    while (it < r1 && (it / KT == k || (colla && it / KT == la))) it += rstep;
    return it;
  };
  int it = next_tile(r0 + blockIdx.y);
  if (it >= r1) return;
  stage_y<TL / 16>(S, A + (int64_t)k * GB * np + (int64_t)jt * TL, np);
  __syncthreads();
  const int w = threadIdx.x / WAVE;
  const double *Xk = A + (int64_t)k * GB + (int64_t)16 * w * np;  // column panel, wavefront rows
  double *Cj = A + (int64_t)jt * TL + (int64_t)16 * w * np;
  // the next tile's X run and C tile are loaded before the current tile's
  // products: loads complete in order, so anything issued later would hold
  // up the wait for them
  double a[NQ];
  v4d acc[TL / 16];
  load_x(a, Xk + (int64_t)it * TL * np, np, -1.0);
  load_c<TL / 16>(acc, Cj + (int64_t)it * TL * np, np);
  for (;;) {
    const int nx = next_tile(it + rstep);
    double an[NQ];
    v4d accn[TL / 16];
    if (nx < r1) {
      load_x(an, Xk + (int64_t)nx * TL * np, np, -1.0);
      load_c<TL / 16>(accn, Cj + (int64_t)nx * TL * np, np);
    }
    mfma_slots_grouped<TL / 16>(a, S, acc);
    store_c<TL / 16>(acc, Cj + (int64_t)it * TL * np, np);
    if (nx >= r1) break;
    it = nx;
#pragma unroll
    for (int q = 0; q < NQ; q++) a[q] = an[q];
#pragma unroll
    for (int tj = 0; tj < TL / 16; tj++) acc[tj] = accn[tj];
  }
}

// A_ik <- -A_ik D for the TL-row tile blockIdx.x outside block row k, D
// staged in LDS: wavefront w reads and then writes rows 16w .. 16w+15 of the
// TL x GB panel only.
__global__ void __launch_bounds__(256) k_gj_colpanel(double *A, int64_t np, int k, const double *dinv) {
  __shared__ double S[GB * GB];
  const int it = blockIdx.x;
  if (it / (GB / TL) == k) return;
  stage_y<GB / 16>(S, dinv, GB);
  __syncthreads();
  const int w = threadIdx.x / WAVE;
  double *X = A + (int64_t)(it * TL + 16 * w) * np + (int64_t)k * GB;
  double a[NQ];
  load_x(a, X, np, -1.0);
  v4d acc[GB / 16] = {};
  mfma_slots<GB / 16>(a, S, acc);
  store_c<GB / 16>(acc, X, np);
}

// The SPD matrix of the coarse level, padded to np x np: the stored entries
// of each row (in row order), then 1 / |C| on the pairs of one component C,
// then (A + A^T) / 2; identity on the padding.  The same steps as the host
// assembly (coarse_dense_spd).
__global__ void k_spd_rows(int64_t n, int64_t np, const int32_t *rp, const int32_t *ci, const double *v,
                           double *A) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
    if (i >= n) {
      A[i * np + i] = 1.0;
      continue;
    }
    for (int32_t k = rp[i]; k < rp[i + 1]; k++) A[i * np + ci[k]] += v[k];
  }
}
__global__ void k_spd_comp(int64_t n, int64_t np, const int32_t *comp, const double *cinv, double *A) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / n, j = e % n;
    if (comp[i] == comp[j]) A[i * np + j] += cinv[comp[i]];
  }
}
__global__ void k_spd_sym(int64_t n, int64_t np, double *A) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / n, j = e % n;
    if (j <= i) continue;
    const double a = 0.5 * (A[i * np + j] + A[j * np + i]);
    A[i * np + j] = a;
    A[j * np + i] = a;
  }
}

// pinv (n x n) = inverse (padded, stride np) minus 1 / |C| on the pairs of
// one component
__global__ void k_pinv_compact(const double *Ainv, int64_t np, int64_t n, const int32_t *comp,
                               const double *cinv, double *pinv) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / n, j = e % n;
    const double c = (comp[i] == comp[j]) ? cinv[comp[i]] : 0.0;
    pinv[e] = Ainv[i * np + j] - c;
  }
}

}  // namespace

void coarse_pinv(int64_t n, const std::vector<int32_t> &rp, const std::vector<int32_t> &ci,
                 const std::vector<double> &v, DevBuf<double> &pinv, hipStream_t s, double *gj_ms) {
  if (gj_ms) *gj_ms = 0.0;
  if (n < COARSE_DEVICE_MIN) {
    const std::vector<double> h = dense_pinv_host(n, rp, ci, v);
    pinv.alloc(std::max<size_t>(h.size(), 1));
    to_device(pinv.p, h.data(), h.size(), s);
    LG_HIP(hipStreamSynchronize(s));
    return;
  }
  LG_CHECK(n <= COARSE_DENSE_MAX, LIGMG_ERR_UNSUPPORTED,
           "the coarsest level has " + std::to_string(n) + " vertices; the dense coarse solve takes at most " +
               std::to_string(COARSE_DENSE_MAX) + " (coarsening stalled: check the aggregation options)");
  // components (host: the coarse level's entries are few), then the matrix
  // assembled on the device
  std::vector<int32_t> comp(n, -1);
  std::vector<int64_t> csize;
  {
    std::vector<int64_t> stack;
    for (int64_t s0 = 0; s0 < n; s0++) {
      if (comp[s0] >= 0) continue;
      const int32_t cid = (int32_t)csize.size();
      csize.push_back(0);
      stack.assign(1, s0);
      comp[s0] = cid;
      while (!stack.empty()) {
        const int64_t u = stack.back();
        stack.pop_back();
        csize[cid]++;
        for (int32_t k = rp[u]; k < rp[u + 1]; k++)
          if (v[k] != 0 && comp[ci[k]] < 0) {
            comp[ci[k]] = cid;
            stack.push_back(ci[k]);
          }
      }
    }
  }
  std::vector<double> cinv(csize.size());
  for (size_t c = 0; c < csize.size(); c++) cinv[c] = 1.0 / (double)csize[c];
  const int64_t nb = (n + GB - 1) / GB, np = nb * GB, nnz = rp[n];
  DevBuf<int32_t> drp(n + 1), dci(std::max<int64_t>(nnz, 1)), dcomp(n);
  DevBuf<double> dv(std::max<int64_t>(nnz, 1)), dcinv(cinv.size());
  to_device(drp.p, rp.data(), n + 1, s);
  to_device(dci.p, ci.data(), nnz, s);
  to_device(dv.p, v.data(), nnz, s);
  to_device(dcomp.p, comp.data(), n, s);
  to_device(dcinv.p, cinv.data(), cinv.size(), s);
  DevBuf<double> A((size_t)np * np), dinv(GB * GB);
  fill_f64(A.p, 0.0, np * np, s);
  hipLaunchKernelGGL(k_spd_rows, dim3(grid_for(np, 256, 16384)), dim3(256), 0, s, n, np, (const int32_t *)drp.p,
                     (const int32_t *)dci.p, (const double *)dv.p, A.p);
  LG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_spd_comp, dim3(grid_for(n * n, 256, 16384)), dim3(256), 0, s, n, np,
                     (const int32_t *)dcomp.p, (const double *)dcinv.p, A.p);
  LG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_spd_sym, dim3(grid_for(n * n, 256, 16384)), dim3(256), 0, s, n, np, A.p);
  LG_LAUNCH_CHECK();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (gj_ms) {
    LG_HIP(hipEventCreate(&e0));
    LG_HIP(hipEventCreate(&e1));
    LG_HIP(hipEventRecord(e0, s));
  }
  const int nt = (int)(np / TL), KT = GB / TL, mt = nt - KT;
  // row strips of the update: as many as keep every workgroup resident at
  // once, two per CU (the update's LDS and registers allow two).  A second
  // round of workgroups measured slower: the rounds no longer walk the same
  // rows of the column panel together.
  int dev = 0, cus = 0;
  LG_HIP(hipGetDevice(&dev));
  LG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int nstrip = (int)std::max<int64_t>(1, std::min<int64_t>(mt, 2 * (int64_t)std::max(cus, 1) / nt));
  // Look-ahead: while step k's trailing update runs on this stream, a
  // second stream updates block (k+1, k+1) alone and inverts it.  The
  // diagonal kernel's workgroup (8 wavefronts, 104 registers each) fits on a
  // CU beside one update workgroup, and the update leaves such CUs.  The two
  // D buffers alternate with k.
  DevBuf<double> dinv2(GB * GB);
  double *dbuf[2] = {dinv.p, dinv2.p};
  struct Side {
    hipStream_t st = nullptr;
    hipEvent_t rp = nullptr, blk = nullptr, diag = nullptr;
    ~Side() {
      if (rp) (void)hipEventDestroy(rp);
      if (blk) (void)hipEventDestroy(blk);
      if (diag) (void)hipEventDestroy(diag);
      if (st) (void)hipStreamDestroy(st);
    }
  } side;
  LG_HIP(hipStreamCreateWithFlags(&side.st, hipStreamNonBlocking));
  LG_HIP(hipEventCreateWithFlags(&side.rp, hipEventDisableTiming));
  LG_HIP(hipEventCreateWithFlags(&side.blk, hipEventDisableTiming));
  LG_HIP(hipEventCreateWithFlags(&side.diag, hipEventDisableTiming));
  hipLaunchKernelGGL(k_gj_diag, dim3(1), dim3(DT), 0, s, A.p, np, 0, dbuf[0]);
  LG_LAUNCH_CHECK();
  for (int k = 0; k < nb; k++) {
    const double *D = dbuf[k & 1];
    const int la = (k + 1 < nb) ? k + 1 : -1;
    if (k > 0) LG_HIP(hipStreamWaitEvent(s, side.diag, 0));
    hipLaunchKernelGGL(k_gj_rowpanel, dim3((unsigned)nt), dim3(256), 0, s, A.p, np, k, D);
    LG_LAUNCH_CHECK();
    if (la >= 0) {
      // side: block (la, la) of step k's update, then its inverse
      LG_HIP(hipEventRecord(side.rp, s));
      LG_HIP(hipStreamWaitEvent(side.st, side.rp, 0));
      hipLaunchKernelGGL(k_gj_update, dim3((unsigned)KT), dim3(256), 0, side.st, A.p, np, k, la * KT, la * KT,
                         la * KT + KT, 1, -1);
      LG_LAUNCH_CHECK();
      LG_HIP(hipEventRecord(side.blk, side.st));
      hipLaunchKernelGGL(k_gj_diag, dim3(1), dim3(DT), 0, side.st, A.p, np, la, dbuf[la & 1]);
      LG_LAUNCH_CHECK();
      LG_HIP(hipEventRecord(side.diag, side.st));
    }
    hipLaunchKernelGGL(k_gj_update, dim3((unsigned)nt, (unsigned)nstrip), dim3(256), 0, s, A.p, np, k, 0, 0, nt,
                       nstrip, la);
    LG_LAUNCH_CHECK();
    // the column panel overwrites A_(la)k, which the side's block update reads
    if (la >= 0) LG_HIP(hipStreamWaitEvent(s, side.blk, 0));
    hipLaunchKernelGGL(k_gj_colpanel, dim3((unsigned)nt), dim3(256), 0, s, A.p, np, k, D);
    LG_LAUNCH_CHECK();
  }
  if (gj_ms) {
    LG_HIP(hipEventRecord(e1, s));
    LG_HIP(hipEventSynchronize(e1));
    float ms = 0;
    LG_HIP(hipEventElapsedTime(&ms, e0, e1));
    *gj_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  pinv.alloc((size_t)n * n);
  hipLaunchKernelGGL(k_pinv_compact, dim3(grid_for(n * n, 256, 16384)), dim3(256), 0, s, (const double *)A.p, np,
                     n, (const int32_t *)dcomp.p, (const double *)dcinv.p, pinv.p);
  LG_LAUNCH_CHECK();
  LG_HIP(hipStreamSynchronize(s));
}

}  // namespace lg
