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
// Gauss-Jordan by 64 x 64 blocks (A padded to np = 64 * nb rows with an
// identity block), for k = 0 .. nb-1, with D = A_kk^-1:
//   A_kj <- D A_kj             (j != k)        row panel
//   A_ij <- A_ij - A_ik A_kj   (i, j != k)     trailing update (old A_ik)
//   A_ik <- -A_ik D            (i != k)        column panel
//   A_kk <- D
// leaves A^-1 in place.  No pivoting: the matrix is SPD, so every diagonal
// block met on the way is SPD (a Schur complement of an SPD matrix).
#include <cmath>
#include <vector>

#include "prims.hpp"
#include "setup.hpp"

namespace lg {

namespace {

constexpr int GB = 64;       // block size
constexpr int GLD = GB + 1;  // LDS row stride (doubles): rows fall in different banks
typedef double v4d __attribute__((ext_vector_type(4)));

// D = A_kk^-1 by unblocked Gauss-Jordan in LDS (one workgroup of 256 lanes,
// 16 elements each); D is written to A_kk and to dinv.
__global__ void __launch_bounds__(256) k_gj_diag(double *A, int64_t np, int k, double *dinv) {
  __shared__ double M[GB * GLD];
  __shared__ double f[GB];
  const int t = threadIdx.x;
  const int64_t o = (int64_t)k * GB;
  for (int e = t; e < GB * GB; e += 256) M[(e / GB) * GLD + e % GB] = A[(o + e / GB) * np + o + e % GB];
  __syncthreads();
  for (int p = 0; p < GB; p++) {
    const double piv = M[p * GLD + p];
    __syncthreads();
    // pivot row: M[p][p] <- 1, then the row divided by the pivot
    if (t < GB) M[p * GLD + t] = ((t == p) ? 1.0 : M[p * GLD + t]) / piv;
    if (t >= GB && t < 2 * GB) f[t - GB] = M[(t - GB) * GLD + p];
    __syncthreads();
    // other rows: M[i][p] <- 0, then row i minus f_i times the pivot row
    for (int e = t; e < GB * GB; e += 256) {
      const int i = e / GB, j = e % GB;
      if (i == p) continue;
      const double m = (j == p) ? 0.0 : M[i * GLD + j];
      M[i * GLD + j] = m - f[i] * M[p * GLD + j];
    }
    __syncthreads();
  }
  for (int e = t; e < GB * GB; e += 256) {
    const double d = M[(e / GB) * GLD + e % GB];
    A[(o + e / GB) * np + o + e % GB] = d;
    dinv[e] = d;
  }
}

// one 64 x 64 block from global (row stride ld) into LDS (stride GLD)
__device__ __forceinline__ void load_block(double *S, const double *G, int64_t ld) {
  for (int e = threadIdx.x; e < GB * GB; e += 256) S[(e / GB) * GLD + e % GB] = G[(int64_t)(e / GB) * ld + e % GB];
}

// C (64 x 64, four 16 x 16 tiles per wavefront: rows 16w .. 16w+15) =
// sign * X Y + C, X (stride ldx) and Y (stride ldy) read in place (global
// memory or LDS).  Operand layout of v_mfma_f64_16x16x4f64 (measured on
// gfx950): lane l supplies X[l%16][l/16] and Y[l/16][l%16] of the 16 x 4 /
// 4 x 16 slices and holds C[l/16 + 4v][l%16], v = 0..3.  Wavefront w reads
// rows 16w .. 16w+15 of X only.
__device__ __forceinline__ void mfma_block(const double *X, int64_t ldx, const double *Y, int64_t ldy, double sign,
                                           v4d acc[4]) {
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  const int li = lane & 15, lk = lane >> 4;
  for (int kk = 0; kk < GB; kk += 4) {
    const double a = sign * X[(16 * w + li) * ldx + kk + lk];
#pragma unroll
    for (int tj = 0; tj < 4; tj++) {
      const double b = Y[(kk + lk) * ldy + 16 * tj + li];
      acc[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[tj], 0, 0, 0);
    }
  }
}
__device__ __forceinline__ void store_block(const v4d acc[4], double *G, int64_t ld) {
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int tj = 0; tj < 4; tj++)
#pragma unroll
    for (int v = 0; v < 4; v++) G[(int64_t)(16 * w + lk + 4 * v) * ld + 16 * tj + li] = acc[tj][v];
}
__device__ __forceinline__ void load_acc(v4d acc[4], const double *G, int64_t ld) {
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int tj = 0; tj < 4; tj++)
#pragma unroll
    for (int v = 0; v < 4; v++) acc[tj][v] = G[(int64_t)(16 * w + lk + 4 * v) * ld + 16 * tj + li];
}

// A_kj <- D A_kj for block column j = blockIdx.x (skipping k).  Every
// wavefront reads all of A_kj, which the workgroup overwrites: A_kj is
// staged in LDS first.
__global__ void __launch_bounds__(256) k_gj_rowpanel(double *A, int64_t np, int k, const double *dinv) {
  __shared__ double Y[GB * GLD];
  const int j = blockIdx.x;
  if (j == k) return;
  double *Akj = A + (int64_t)k * GB * np + (int64_t)j * GB;
  load_block(Y, Akj, np);
  __syncthreads();
  v4d acc[4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
  mfma_block(dinv, GB, Y, GLD, 1.0, acc);
  store_block(acc, Akj, np);
}

// A_ij <- A_ij - A_ik A_kj for i, j != k (blockIdx.y = i, blockIdx.x = j):
// operands read in place (this launch writes neither panel)
__global__ void __launch_bounds__(256) k_gj_update(double *A, int64_t np, int k) {
  const int i = blockIdx.y, j = blockIdx.x;
  if (i == k || j == k) return;
  const double *Aik = A + (int64_t)i * GB * np + (int64_t)k * GB;
  const double *Akj = A + (int64_t)k * GB * np + (int64_t)j * GB;
  double *Aij = A + (int64_t)i * GB * np + (int64_t)j * GB;
  v4d acc[4];
  load_acc(acc, Aij, np);
  mfma_block(Aik, np, Akj, np, -1.0, acc);
  store_block(acc, Aij, np);
}

// A_ik <- -A_ik D for block row i = blockIdx.x (skipping k): wavefront w
// reads and then writes rows 16w .. 16w+15 of A_ik only
__global__ void __launch_bounds__(256) k_gj_colpanel(double *A, int64_t np, int k, const double *dinv) {
  const int i = blockIdx.x;
  if (i == k) return;
  double *Aik = A + (int64_t)i * GB * np + (int64_t)k * GB;
  v4d acc[4] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
  mfma_block(Aik, np, dinv, GB, -1.0, acc);
  store_block(acc, Aik, np);
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
  for (int k = 0; k < nb; k++) {
    hipLaunchKernelGGL(k_gj_diag, dim3(1), dim3(256), 0, s, A.p, np, k, dinv.p);
    LG_LAUNCH_CHECK();
    if (nb == 1) break;
    hipLaunchKernelGGL(k_gj_rowpanel, dim3((unsigned)nb), dim3(256), 0, s, A.p, np, k, (const double *)dinv.p);
    LG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gj_update, dim3((unsigned)nb, (unsigned)nb), dim3(256), 0, s, A.p, np, k);
    LG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gj_colpanel, dim3((unsigned)nb), dim3(256), 0, s, A.p, np, k, (const double *)dinv.p);
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
