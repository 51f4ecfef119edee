// Block (multi right-hand side) kernels for batched solves.
//
// Block vectors are row-major n x kp (kp = number of right-hand sides
// rounded up to a power of two, at most 64): the kp values of a row are
// contiguous, so one gather of a column index fetches a whole row of X.  A
// group of kp lanes handles one matrix row, lane c handling right-hand
// side c.  Row sums accumulate in FP64.
#include <algorithm>
#include <type_traits>

#include "../../include/ligmg_amd.h"
#include "ops.hpp"
#include "prims.hpp"

namespace lg {

namespace {

template <class T>
struct BValOne {
  __device__ __forceinline__ T operator()(int64_t) const { return (T)1; }
};
template <class T>
struct BValArr {
  const T *v;
  __device__ __forceinline__ T operator()(int64_t k) const { return v[k]; }
};

// ----------------------------------------------------------- epilogues (i, c)
template <class T>
struct BStore {
  T *y;
  int kp;
  __device__ __forceinline__ void operator()(int64_t i, int c, double s) const {
    y[i * kp + c] = (T)s;
  }
};
template <class T>
struct BResid {
  const T *b;
  T *r;
  int kp;
  __device__ __forceinline__ void operator()(int64_t i, int c, double s) const {
    const int64_t o = i * kp + c;
    r[o] = (T)((double)b[o] - s);
  }
};
template <class T>
struct BCheb {
  const T *dinv, *b, *xc, *xp;
  T *xn;
  double alpha, beta;
  int kp;
  __device__ __forceinline__ void operator()(int64_t i, int c, double s) const {
    const int64_t o = i * kp + c;
    double cur = (double)xc[o];
    double d = alpha * (cur - (xp ? (double)xp[o] : 0.0));  // null xp: zero previous iterate
    xn[o] = (T)(cur + d + beta * (double)dinv[i] * ((double)b[o] - s));
  }
};
template <class T>
struct BInterp {
  const T *q, *b;
  T *x;
  int kp;
  __device__ __forceinline__ void operator()(int64_t i, int c, double s) const {
    const int64_t o = i * kp + c;
    x[o] = (T)(s + (double)q[i] * (double)b[o]);
  }
};
template <class T, class Epi>
struct BDiag {  // FP32 Laplacian form: + d[i] x[i, c]
  const double *d;
  const T *x;
  Epi e;
  int kp;
  __device__ __forceinline__ void operator()(int64_t i, int c, double s) const {
    e(i, c, s + d[i] * (double)x[i * kp + c]);
  }
};

// One workgroup per unit (the SpMV units).  Complete-row units: group g of
// kp / VW lanes handles rows g, g + G, ...; lane t of a group handles the
// VW adjacent right-hand sides t*VW .. t*VW + VW - 1 with one wide load per
// gathered row of X, and keeps U entries' gathers in flight.  Long-row
// chunks: groups of kp lanes (one right-hand side each, whatever VW) split
// the chunk's entries and reduce through LDS into kp partial sums.  Every
// (row, right-hand side) sum of a complete row is one lane adding the row's
// entries in stored order, so the tile (VW, U) does not change any bit.
template <class T, int VW>
struct BVec {
  typedef T type __attribute__((ext_vector_type(VW)));
};
template <class T>
struct BVec<T, 1> {
  typedef T type;
};
template <class T, int VW>
__device__ __forceinline__ double bvec_at(const typename BVec<T, VW>::type &x, int v) {
  if constexpr (VW == 1) return (double)x;
  else return (double)x[v];
}

template <class T, int VW, int U, class Val, class Epi>
__global__ void __launch_bounds__(256)
    k_spmm_units(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci, Val val,
                 const T *__restrict__ X, int kp, int lk, const int32_t *__restrict__ urow0,
                 const int32_t *__restrict__ unrows, const int32_t *__restrict__ ukoff,
                 double *__restrict__ parts, Epi epi) {
  typedef typename BVec<T, VW>::type V;
  const int u = blockIdx.x;
  const int row0 = urow0[u], nr = unrows[u], k0 = ukoff[u];
  if (nr > 0) {
    const int lg = lk - (VW == 1 ? 0 : VW == 2 ? 1 : VW == 4 ? 2 : 3);  // log2(lanes per row)
    const int G = 256 >> lg;
    const int g = threadIdx.x >> lg, c0 = (threadIdx.x & ((1 << lg) - 1)) * VW;
    for (int r = g; r < nr; r += G) {
      const int row = row0 + r;
      const int a = rp[row], b = rp[row + 1];
      double s[VW];
#pragma unroll
      for (int v = 0; v < VW; v++) s[v] = 0;
      int j = a;
      for (; j + U - 1 < b; j += U) {  // U independent gathers in flight
        int cc[U];
        double vv[U];
        V xx[U];
#pragma unroll
        for (int t = 0; t < U; t++) {
          cc[t] = ci[j + t];
          vv[t] = (double)val(j + t);
        }
#pragma unroll
        for (int t = 0; t < U; t++) xx[t] = *(const V *)(X + (int64_t)cc[t] * kp + c0);
#pragma unroll
        for (int t = 0; t < U; t++)
#pragma unroll
          for (int v = 0; v < VW; v++) s[v] += vv[t] * bvec_at<T, VW>(xx[t], v);
      }
      for (; j < b; j++) {
        const double vj = (double)val(j);
        const V xj = *(const V *)(X + (int64_t)ci[j] * kp + c0);
#pragma unroll
        for (int v = 0; v < VW; v++) s[v] += vj * bvec_at<T, VW>(xj, v);
      }
#pragma unroll
      for (int v = 0; v < VW; v++) epi(row, c0 + v, s[v]);
    }
  } else {
    const int G = 256 >> lk;
    const int g = threadIdx.x >> lk, c = threadIdx.x & (kp - 1);
    __shared__ double red[256];
    const int row = row0;
    const int cnt = min(k0 + UNIT_NNZ, rp[row + 1]) - k0;
    double s = 0;
    for (int t = g; t < cnt; t += G) {
      const int j = k0 + t;
      s += (double)val(j) * (double)X[(int64_t)ci[j] * kp + c];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    if (g == 0) {
      for (int gg = 1; gg < G; gg++) s += red[gg * kp + c];
      parts[(int64_t)(-nr) * kp + c] = s;
    }
  }
}

template <class Epi>
__global__ void k_spmm_long(const int32_t *lrow, const int32_t *lpart, int64_t nlong, int kp,
                            const double *parts, Epi epi) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nlong * kp;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = t / kp;
    const int c = (int)(t % kp);
    double s = 0;
    for (int p = lpart[j]; p < lpart[j + 1]; p++) s += parts[(int64_t)p * kp + c];
    epi(lrow[j], c, s);
  }
}

int log2i(int k) {
  int l = 0;
  while ((1 << l) < k) l++;
  return l;
}

template <class T>
const T *bvals(const DCsr &A);
template <>
const double *bvals<double>(const DCsr &A) {
  return A.v.p;
}
template <>
const float *bvals<float>(const DCsr &A) {
  return A.v32.p;
}

// SpMM tiles (right-hand sides per lane VW, gathers in flight U) by id;
// id 0 is the fixed tile.  VW never exceeds the block width kp.
constexpr int SPMM_NCFG = 6;
constexpr int SPMM_VW[SPMM_NCFG] = {1, 1, 2, 2, 4, 4};
constexpr int SPMM_U[SPMM_NCFG] = {2, 4, 2, 4, 2, 4};

template <class T>
constexpr int prec_slot() {
  return std::is_same<T, float>::value ? 1 : 0;
}

template <class T, int VW, int U, class Val, class Epi>
void launch_spmm_tile(const DCsr &A, Val val, const T *X, int kp, Epi epi, double *parts, hipStream_t s) {
  hipLaunchKernelGGL((k_spmm_units<T, VW, U, Val, Epi>), dim3((unsigned)A.U.nunits), dim3(256), 0, s, A.rp.p,
                     A.ci.p, val, X, kp, log2i(kp), A.U.row0.p, A.U.nrows.p, A.U.koff.p, parts, epi);
  LG_LAUNCH_CHECK();
}

template <class T, class Val, class Epi>
void launch_spmm_cfg(int cfg, const DCsr &A, Val val, const T *X, int kp, Epi epi, DevBuf<double> &parts,
                     hipStream_t s) {
  switch (cfg) {
    case 1: launch_spmm_tile<T, 1, 4>(A, val, X, kp, epi, parts.p, s); break;
    case 2: launch_spmm_tile<T, 2, 2>(A, val, X, kp, epi, parts.p, s); break;
    case 3: launch_spmm_tile<T, 2, 4>(A, val, X, kp, epi, parts.p, s); break;
    case 4: launch_spmm_tile<T, 4, 2>(A, val, X, kp, epi, parts.p, s); break;
    case 5: launch_spmm_tile<T, 4, 4>(A, val, X, kp, epi, parts.p, s); break;
    default: launch_spmm_tile<T, 1, 2>(A, val, X, kp, epi, parts.p, s); break;
  }
}

template <class T, class Val, class Epi>
void launch_spmm(const DCsr &A, Val val, const T *X, int kp, Epi epi, DevBuf<double> &parts,
                 hipStream_t s) {
  if (A.U.nunits == 0) return;
  LG_CHECK(parts.n >= (size_t)std::max<int64_t>(A.U.nparts, 1) * kp, LIGMG_ERR_INVALID,
           "spmm partial-sum buffer too small");
  int cfg = A.U.spmm_cfg[prec_slot<T>()][log2i(kp)];
  if (cfg < 0 || SPMM_VW[cfg] > kp) cfg = 0;
  launch_spmm_cfg<T>(cfg, A, val, X, kp, epi, parts, s);
  if (A.U.nlong) {
    hipLaunchKernelGGL((k_spmm_long<Epi>), dim3(grid_for(A.U.nlong * kp, 256)), dim3(256), 0, s,
                       A.U.lrow.p, A.U.lpart.p, A.U.nlong, kp, (const double *)parts.p, epi);
    LG_LAUNCH_CHECK();
  }
}

template <class T, class Epi>
void dispatch_spmm(const DCsr &A, const T *X, int kp, Epi epi, DevBuf<double> &parts, hipStream_t s) {
  const T *v = bvals<T>(A);
  if (std::is_same<T, float>::value && A.d32.p)
    launch_spmm<T>(A, BValArr<T>{v}, X, kp, BDiag<T, Epi>{A.d32.p, X, epi, kp}, parts, s);
  else if (v) launch_spmm<T>(A, BValArr<T>{v}, X, kp, epi, parts, s);
  else launch_spmm<T>(A, BValOne<T>{}, X, kp, epi, parts, s);
}

// ------------------------------------------------------------- elementwise
template <class T>
__global__ void k_bcheb_zero(const T *dinv, const T *b, T *xn, double beta, int64_t n, int lk) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < (n << lk);
       t += (int64_t)gridDim.x * blockDim.x)
    xn[t] = (T)(beta * (double)dinv[t >> lk] * (double)b[t]);
}

template <class T>
__global__ void k_bprolong(const int32_t *agg, int64_t n, int kp, int lk, const T *xc, T *x) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < (n << lk);
       t += (int64_t)gridDim.x * blockDim.x)
    x[t] += xc[(int64_t)agg[t >> lk] * kp + (t & (kp - 1))];
}

// column-major (ld n) <-> row-major n x kp (padding columns zero)
// column-major caller block (vertex order given by perm, identity when
// null) <-> row-major internal block
__global__ void k_to_block(const double *B, int64_t n, int k, int kp, const int32_t *perm, double *X) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n * kp;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / kp;
    const int c = (int)(t % kp);
    const int64_t src = perm ? perm[i] : i;
    X[t] = c < k ? B[(int64_t)c * n + src] : 0.0;
  }
}
__global__ void k_from_block(const double *X, int64_t n, int k, int kp, const int32_t *perm, double *B) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n * k;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t / n, i = t % n;
    const int64_t dst = perm ? perm[i] : i;
    B[c * n + dst] = X[i * kp + c];
  }
}

// ------------------------------------------------------- coarse dense apply
// Y (n x kp) = Pm (n x n, FP64, row-major) * X (n x kp) with FP64 MFMA
// 16x16x4 tiles.  Operand layout of v_mfma_f64_16x16x4f64 (measured on
// gfx950): lane l = 16 lk + li supplies A[li][lk] and B[lk][li] of the
// 16x4 / 4x16 slices and holds D[lk + 4v][li], v = 0..3.
//
// One workgroup per 16 rows of Y and all kp = 16 NT columns, so Pm is read
// once.  Its four wavefronts split the K range: lane (li, lk) of wavefront w
// takes K = (4 w + lk) Q + q, q = 0 .. Q-1 (Q = ceil(n / 16)), one contiguous
// run of its row of Pm; two accumulator sets alternate on q so consecutive
// products into one tile do not wait on each other.  The four partial tiles
// are added in LDS in wavefront order.
template <class T, int NT>
__global__ void __launch_bounds__(256) k_dense_mfma(const double *__restrict__ Pm, int n,
                                                    const T *__restrict__ X, int kp, int Q,
                                                    T *__restrict__ Y) {
  typedef double v4d __attribute__((ext_vector_type(4)));
  __shared__ double part[4][16 * 16 * NT];
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  const int li = lane & 15, lk = lane >> 4;
  const int i0 = blockIdx.x * 16, row = i0 + li;
  const double *prow = Pm + (int64_t)(row < n ? row : n - 1) * n;
  const int kb = (4 * w + lk) * Q;
  v4d acc[2][NT];
#pragma unroll
  for (int h = 0; h < 2; h++)
#pragma unroll
    for (int tj = 0; tj < NT; tj++) acc[h][tj] = v4d{0, 0, 0, 0};
  // K slots go in blocks of QB (16 products), each block's operands loaded
  // one block ahead
  constexpr int QB = 16 / NT;
  auto load_blk = [&](int q0, double *aa, double (*bb)[NT]) {
#pragma unroll
    for (int u = 0; u < QB; u++) {
      const int kk = kb + q0 + u;
      const bool ok = q0 + u < Q && kk < n;
      aa[u] = (ok && row < n) ? prow[kk] : 0.0;
#pragma unroll
      for (int tj = 0; tj < NT; tj++) bb[u][tj] = ok ? (double)X[(int64_t)kk * kp + 16 * tj + li] : 0.0;
    }
  };
  double a[QB], b[QB][NT];
  load_blk(0, a, b);
  for (int q0 = 0; q0 < Q; q0 += QB) {
    double an[QB], bn[QB][NT];
    if (q0 + QB < Q) load_blk(q0 + QB, an, bn);
#pragma unroll
    for (int u = 0; u < QB; u++)
#pragma unroll
      for (int tj = 0; tj < NT; tj++)
        acc[u & 1][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u][tj], acc[u & 1][tj], 0, 0, 0);
#pragma unroll
    for (int u = 0; u < QB; u++) {
      a[u] = an[u];
#pragma unroll
      for (int tj = 0; tj < NT; tj++) b[u][tj] = bn[u][tj];
    }
  }
#pragma unroll
  for (int tj = 0; tj < NT; tj++)
#pragma unroll
    for (int v = 0; v < 4; v++) part[w][(lk + 4 * v) * 16 * NT + 16 * tj + li] = acc[0][tj][v] + acc[1][tj][v];
  __syncthreads();
  for (int e = threadIdx.x; e < 16 * 16 * NT; e += blockDim.x) {
    const int r = e / (16 * NT), c = e % (16 * NT);
    if (i0 + r < n) Y[(int64_t)(i0 + r) * kp + c] = (T)(((part[0][e] + part[1][e]) + part[2][e]) + part[3][e]);
  }
}

template <class T>
__global__ void k_dense_small(const double *__restrict__ Pm, int64_t n, const T *__restrict__ X,
                              int kp, T *__restrict__ Y) {
  // kp < 16: one thread per output entry
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n * kp;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / kp;
    const int c = (int)(t % kp);
    double s = 0;
    for (int64_t j = 0; j < n; j++) s += Pm[i * n + j] * (double)X[j * kp + c];
    Y[t] = (T)s;
  }
}

// -------------------------------------------------- per-(block, column) dots
// part[(s * np + j) * kp + c] for segment s; out[(blk * np + j) * kp + c]
template <class T>
__global__ void __launch_bounds__(256) k_bred1(int np, int kp, int lk, const T *a0, const T *b0,
                                               const T *a1, const T *b1, const int64_t *lo,
                                               const int64_t *hi, double *part) {
  const int s = blockIdx.x;
  const int G = 256 >> lk;
  const int g = threadIdx.x >> lk, c = threadIdx.x & (kp - 1);
  double acc0 = 0, acc1 = 0;
  for (int64_t i = lo[s] + g; i < hi[s]; i += G) {
    const int64_t o = i * kp + c;
    acc0 += (double)a0[o] * (b0 ? (double)b0[o] : 1.0);
    if (np > 1) acc1 += (double)a1[o] * (b1 ? (double)b1[o] : 1.0);
  }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = acc0;
  red[1][threadIdx.x] = acc1;
  __syncthreads();
  if (g == 0) {
    for (int j = 0; j < np; j++) {
      double t = 0;
      for (int gg = 0; gg < G; gg++) t += red[j][gg * kp + c];
      part[((int64_t)s * np + j) * kp + c] = t;
    }
  }
}

// out[blk, j, c] = sum of the segment partials of block blk: one workgroup
// per (block, j); thread (g, c) adds segments g, g + G, ... (G = 256 / kp)
// in order, and the G partials of a column are added in g order.
__global__ void __launch_bounds__(256) k_bred2(int np, int kp, int lk, const int32_t *bseg, const double *part,
                                               double *out) {
  const int blk = blockIdx.x, j = blockIdx.y;
  const int G = 256 >> lk;
  const int g = threadIdx.x >> lk, c = threadIdx.x & (kp - 1);
  double acc = 0;
  for (int sg = bseg[blk] + g; sg < bseg[blk + 1]; sg += G) acc += part[((int64_t)sg * np + j) * kp + c];
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  if (g == 0) {
    double t = 0;
    for (int gg = 0; gg < G; gg++) t += red[gg * kp + c];
    out[((int64_t)blk * np + j) * kp + c] = t;
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_bsub_mean(T *a, int kp, int lk, const int64_t *lo,
                                                   const int64_t *hi, const int32_t *segblk,
                                                   const int64_t *bptr, const double *sums) {
  const int s = blockIdx.x;
  const int blk = segblk[s];
  const int G = 256 >> lk;
  const int g = threadIdx.x >> lk, c = threadIdx.x & (kp - 1);
  const T m = (T)(sums[(int64_t)blk * kp + c] / (double)(bptr[blk + 1] - bptr[blk]));
  for (int64_t i = lo[s] + g; i < hi[s]; i += G) a[i * kp + c] -= m;
}

// z = c - (block mean), in FP64; with w set, also the segment partials of
// z.w exactly as the block dots form them (part[s * kp + c])
__global__ void __launch_bounds__(256) k_bsub_mean_f2d(double *z, const float *cv, int kp, int lk,
                                                       const int64_t *lo, const int64_t *hi,
                                                       const int32_t *segblk, const int64_t *bptr,
                                                       const double *sums, const double *w, double *part) {
  const int s = blockIdx.x;
  const int blk = segblk[s];
  const int G = 256 >> lk;
  const int g = threadIdx.x >> lk, c = threadIdx.x & (kp - 1);
  const double m = sums[(int64_t)blk * kp + c] / (double)(bptr[blk + 1] - bptr[blk]);
  double acc = 0;
  for (int64_t i = lo[s] + g; i < hi[s]; i += G) {
    const int64_t o = i * kp + c;
    const double zv = (double)cv[o] - m;
    z[o] = zv;
    if (w) acc += zv * w[o];
  }
  if (!w) return;  // uniform over the workgroup
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  if (g == 0) {
    double t = 0;
    for (int gg = 0; gg < G; gg++) t += red[gg * kp + c];
    part[(int64_t)s * kp + c] = t;
  }
}

}  // namespace

// ---------------------------------------------------------------- host side
// Times every tile that fits kp on A (the operator's own values and
// Laplacian-form diagonal, X of zeros, a plain store) and records the
// fastest for this precision and width; a tile already picked is kept.
template <class T>
int tune_spmm(DCsr &A, int kp, DevBuf<double> &parts, hipStream_t s) {
  int8_t &pick = A.U.spmm_cfg[prec_slot<T>()][log2i(kp)];
  if (pick >= 0 || A.U.nunits == 0) return pick < 0 ? 0 : pick;
  DevBuf<T> X((size_t)std::max<int64_t>(A.ncols, 1) * kp), Y((size_t)std::max<int64_t>(A.n, 1) * kp);
  fill_u8((uint8_t *)X.p, 0, (int64_t)X.bytes(), s);
  hipEvent_t e0, e1;
  LG_HIP(hipEventCreate(&e0));
  LG_HIP(hipEventCreate(&e1));
  float best = 0;
  int besti = 0;
  for (int c = 0; c < SPMM_NCFG; c++) {
    if (SPMM_VW[c] > kp) continue;
    A.U.spmm_cfg[prec_slot<T>()][log2i(kp)] = (int8_t)c;
    op_bspmv<T>(A, X.p, Y.p, kp, parts, s);  // warm
    LG_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < 3; r++) op_bspmv<T>(A, X.p, Y.p, kp, parts, s);
    LG_HIP(hipEventRecord(e1, s));
    LG_HIP(hipEventSynchronize(e1));
    float ms = 0;
    LG_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (c == 0 || ms < best) {
      best = ms;
      besti = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  pick = (int8_t)besti;
  return besti;
}

void spmm_tile(int cfg, int *vw, int *u) {
  *vw = SPMM_VW[cfg];
  *u = SPMM_U[cfg];
}

template <class T>
void op_bspmv(const DCsr &A, const T *X, T *Y, int kp, DevBuf<double> &parts, hipStream_t s) {
  dispatch_spmm<T>(A, X, kp, BStore<T>{Y, kp}, parts, s);
}
template <class T>
void op_bresidual(const DCsr &A, const T *X, const T *B, T *R, int kp, DevBuf<double> &parts,
                  hipStream_t s) {
  dispatch_spmm<T>(A, X, kp, BResid<T>{B, R, kp}, parts, s);
}
template <class T>
void op_bcheb_step(const DCsr &A, const T *dinv, const T *B, const T *Xc, const T *Xp, T *Xn,
                   double alpha, double beta, int kp, DevBuf<double> &parts, hipStream_t s) {
  if (!Xc) {
    hipLaunchKernelGGL(k_bcheb_zero<T>, dim3(grid_for(A.n * kp, 256, 16384)), dim3(256), 0, s, dinv, B,
                       Xn, beta, A.n, log2i(kp));
    LG_LAUNCH_CHECK();
    return;
  }
  dispatch_spmm<T>(A, Xc, kp, BCheb<T>{dinv, B, Xc, Xp, Xn, alpha, beta, kp}, parts, s);
}
template <class T>
void op_binterp(const DCsr &P, const T *q, const T *Xc, const T *B, T *X, int kp,
                DevBuf<double> &parts, hipStream_t s) {
  dispatch_spmm<T>(P, Xc, kp, BInterp<T>{q, B, X, kp}, parts, s);
}
template <class T>
void op_bprolong_add(const int32_t *agg, int64_t n, const T *Xc, T *X, int kp, hipStream_t s) {
  hipLaunchKernelGGL(k_bprolong<T>, dim3(grid_for(n * kp, 256, 16384)), dim3(256), 0, s, agg, n, kp,
                     log2i(kp), Xc, X);
  LG_LAUNCH_CHECK();
}
template <class T>
void op_bdense_apply(const double *Pm, int64_t n, const T *X, T *Y, int kp, hipStream_t s) {
  if (kp >= 16) {
    LG_CHECK(kp == 16 || kp == 32 || kp == 64, LIGMG_ERR_INVALID, "dense apply: kp must be 16, 32 or 64");
    const unsigned g = (unsigned)((n + 15) / 16);
    const int Q = (int)((n + 15) / 16);
    if (kp == 16) hipLaunchKernelGGL((k_dense_mfma<T, 1>), dim3(g), dim3(256), 0, s, Pm, (int)n, X, kp, Q, Y);
    else if (kp == 32) hipLaunchKernelGGL((k_dense_mfma<T, 2>), dim3(g), dim3(256), 0, s, Pm, (int)n, X, kp, Q, Y);
    else hipLaunchKernelGGL((k_dense_mfma<T, 4>), dim3(g), dim3(256), 0, s, Pm, (int)n, X, kp, Q, Y);
  } else {
    hipLaunchKernelGGL(k_dense_small<T>, dim3(grid_for(n * kp, 256, 16384)), dim3(256), 0, s, Pm, n, X,
                       kp, Y);
  }
  LG_LAUNCH_CHECK();
}

void op_to_block(const double *B, int64_t n, int k, int kp, double *X, const int32_t *perm, hipStream_t s) {
  hipLaunchKernelGGL(k_to_block, dim3(grid_for(n * kp, 256, 16384)), dim3(256), 0, s, B, n, k, kp, perm, X);
  LG_LAUNCH_CHECK();
}
void op_from_block(const double *X, int64_t n, int k, int kp, double *B, const int32_t *perm,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_from_block, dim3(grid_for(n * k, 256, 16384)), dim3(256), 0, s, X, n, k, kp, perm,
                     B);
  LG_LAUNCH_CHECK();
}

template <class T>
void RedPlan::bdots(int np, int kp, const T *a0, const T *b0, const T *a1, const T *b1, double *out,
                    DevBuf<double> &bpart, hipStream_t s) const {
  if (nseg == 0) return;
  LG_CHECK(bpart.n >= (size_t)nseg * np * kp, LIGMG_ERR_INVALID, "block reduction buffer too small");
  hipLaunchKernelGGL(k_bred1<T>, dim3(nseg), dim3(256), 0, s, np, kp, log2i(kp), a0, b0, a1, b1, seg_lo.p,
                     seg_hi.p, bpart.p);
  LG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bred2, dim3((unsigned)nb, (unsigned)np), dim3(256), 0, s, np, kp, log2i(kp), bseg.p, bpart.p,
                     out);
  LG_LAUNCH_CHECK();
}

void RedPlan::bdots_finish(int np, int kp, double *out, const DevBuf<double> &bpart, hipStream_t s) const {
  if (nseg == 0) return;
  LG_CHECK(bpart.n >= (size_t)nseg * np * kp, LIGMG_ERR_INVALID, "block reduction buffer too small");
  hipLaunchKernelGGL(k_bred2, dim3((unsigned)nb, (unsigned)np), dim3(256), 0, s, np, kp, log2i(kp), bseg.p, bpart.p,
                     out);
  LG_LAUNCH_CHECK();
}

void op_bsub_block_mean_f2d(double *z, const float *c, int kp, const RedPlan &P, const double *sums, hipStream_t s,
                            const double *w, DevBuf<double> *bpart) {
  if (P.nseg == 0) return;
  if (w) LG_CHECK(bpart && bpart->n >= (size_t)P.nseg * kp, LIGMG_ERR_INVALID, "block reduction buffer too small");
  hipLaunchKernelGGL(k_bsub_mean_f2d, dim3(P.nseg), dim3(256), 0, s, z, c, kp, log2i(kp), P.seg_lo.p, P.seg_hi.p,
                     P.bseg.p + P.nb + 1, P.bptr.p, sums, w, w ? bpart->p : nullptr);
  LG_LAUNCH_CHECK();
}

template <class T>
void op_bsub_block_mean(T *a, int kp, const RedPlan &P, const double *sums, hipStream_t s) {
  if (P.nseg == 0) return;
  hipLaunchKernelGGL(k_bsub_mean<T>, dim3(P.nseg), dim3(256), 0, s, a, kp, log2i(kp), P.seg_lo.p,
                     P.seg_hi.p, P.bseg.p + P.nb + 1, P.bptr.p, sums);
  LG_LAUNCH_CHECK();
}

#define LG_BINST(T)                                                                                  \
  template void op_bspmv<T>(const DCsr &, const T *, T *, int, DevBuf<double> &, hipStream_t);       \
  template int tune_spmm<T>(DCsr &, int, DevBuf<double> &, hipStream_t);                             \
  template void op_bresidual<T>(const DCsr &, const T *, const T *, T *, int, DevBuf<double> &,     \
                                hipStream_t);                                                        \
  template void op_bcheb_step<T>(const DCsr &, const T *, const T *, const T *, const T *, T *,     \
                                 double, double, int, DevBuf<double> &, hipStream_t);                \
  template void op_binterp<T>(const DCsr &, const T *, const T *, const T *, T *, int,              \
                              DevBuf<double> &, hipStream_t);                                        \
  template void op_bprolong_add<T>(const int32_t *, int64_t, const T *, T *, int, hipStream_t);      \
  template void op_bdense_apply<T>(const double *, int64_t, const T *, T *, int, hipStream_t);       \
  template void RedPlan::bdots<T>(int, int, const T *, const T *, const T *, const T *, double *,    \
                                  DevBuf<double> &, hipStream_t) const;                              \
  template void op_bsub_block_mean<T>(T *, int, const RedPlan &, const double *, hipStream_t);
LG_BINST(double)
LG_BINST(float)
#undef LG_BINST

}  // namespace lg
