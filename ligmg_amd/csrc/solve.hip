// Solve-phase kernels: row-block SpMV family with fused epilogues, transfer
// operators, dense coarse apply, deterministic blocked reductions.
#include <algorithm>
#include <type_traits>

#include "ops.hpp"
#include "prims.hpp"
#include "setup.hpp"
#include "solve_dev.hpp"

namespace lg {

namespace {

using namespace sdev;

// slice passes: first writes the FP64 accumulator, middle ones add to it,
// the last one hands accumulator + its own sum to the real epilogue
struct EpiAccSet {
  double *acc;
  __device__ __forceinline__ void operator()(int64_t i, double s) const { acc[i] = s; }
};
struct EpiAccAdd {
  double *acc;
  __device__ __forceinline__ void operator()(int64_t i, double s) const { acc[i] += s; }
};
template <class Epi>
struct EpiAccFinal {
  const double *acc;
  Epi e;
  __device__ __forceinline__ void operator()(int64_t i, double s) const { e(i, acc[i] + s); }
};

// x[col] through a buffer resource of the vector's bytes: one 32-bit offset
// per lane instead of a 64-bit address, and col = -1 (a masked lane) falls
// outside the buffer and reads 0.
template <class T>
__device__ __forceinline__ double buf_gather(__amdgpu_buffer_rsrc_t r, int col) {
  if constexpr (sizeof(T) == 4)
    return (double)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (uint32_t)col * 4u, 0, 0));
  else
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (uint32_t)col * 8u, 0, 0));
}

// Each workgroup takes upw consecutive units (spmv_upw).  A complete-row unit
// stages every product in LDS with coalesced loads of colind/values (the x
// gathers are the only scattered accesses); its rows are then reduced from
// LDS by groups of row_lanes() lanes while the next unit's colind/values
// loads are in flight.  A long-row chunk writes a partial sum (8 products
// per thread added in order, butterfly per wavefront, the wavefront sums
// added from 0) that k_spmv_long combines.
// BUF: x is read through a buffer resource of xbytes (< 2^31) bytes.
template <class T, class Val, class Epi, bool BUF>
__global__ void __launch_bounds__(SPMV_BLOCK)
    k_spmv_units(const int32_t *__restrict__ rp, const int32_t *__restrict__ ci, Val val,
                 const T *__restrict__ x, int32_t xbytes, const int32_t *__restrict__ urow0,
                 const int32_t *__restrict__ unrows, const int32_t *__restrict__ ukoff, int64_t nunits,
                 int upw, int row_items, double *__restrict__ parts, Epi epi) {
  constexpr int IT = UNIT_NNZ / SPMV_BLOCK;
  __shared__ double prod[UNIT_NNZ];
  __shared__ double ws[SPMV_BLOCK / WAVE];
  const int64_t ubeg = (int64_t)blockIdx.x * upw;
  const int64_t uend = min(ubeg + upw, nunits);
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  int col[IT];
  T vv[IT];
  auto fetch = [&](int64_t u) {
    const int row0 = urow0[u], nr = unrows[u], k0 = ukoff[u];
    const int cnt = nr > 0 ? rp[row0 + nr] - k0 : min(k0 + UNIT_NNZ, rp[row0 + 1]) - k0;
#pragma unroll
    for (int j = 0; j < IT; j++) {
      const int t = threadIdx.x + j * SPMV_BLOCK;
      col[j] = (t < cnt) ? __builtin_nontemporal_load(ci + k0 + t) : -1;
    }
#pragma unroll
    for (int j = 0; j < IT; j++) {
      const int t = threadIdx.x + j * SPMV_BLOCK;
      vv[j] = (t < cnt) ? val(k0 + t) : (T)0;
    }
  };
  fetch(ubeg);
  for (int64_t u = ubeg; u < uend; u++) {
    const int row0 = urow0[u], nr = unrows[u], k0 = ukoff[u];
    const int cnt = nr > 0 ? rp[row0 + nr] - k0 : min(k0 + UNIT_NNZ, rp[row0 + 1]) - k0;
    // all gathers of a lane in flight before any product
    double pr[IT];
    {
      double xx[IT];
      if constexpr (BUF) {
        // word 3 = 0x00020000: the raw-buffer format of gfx9 targets
        const __amdgpu_buffer_rsrc_t xr =
            __builtin_amdgcn_make_buffer_rsrc((void *)x, (short)0, xbytes, 0x00020000);
#pragma unroll
        for (int j = 0; j < IT; j++) xx[j] = buf_gather<T>(xr, col[j]);
      } else {
#pragma unroll
        for (int j = 0; j < IT; j++) xx[j] = (col[j] >= 0) ? (double)x[col[j]] : 0.0;
      }
#pragma unroll
      for (int j = 0; j < IT; j++) pr[j] = (double)vv[j] * xx[j];
    }
    if (nr > 0) {
      if (u > ubeg) __syncthreads();  // the previous unit's rows are read
#pragma unroll
      for (int j = 0; j < IT; j++) {
        const int t = threadIdx.x + j * SPMV_BLOCK;
        if (t < cnt) prod[t] = pr[j];
      }
      if (u + 1 < uend) fetch(u + 1);
      __syncthreads();
      const int K = row_lanes(cnt, nr, row_items);
      if (K == 1) {
        for (int r = threadIdx.x; r < nr; r += SPMV_BLOCK) {
          const int a = rp[row0 + r] - k0, b = rp[row0 + r + 1] - k0;
          double s = 0;
          for (int t = a; t < b; t++) s += prod[t];
          epi(row0 + r, s);
        }
      } else {
        // the whole workgroup runs the same trip count: butterflies stay converged
        const int g = threadIdx.x & (K - 1), ng = SPMV_BLOCK / K;
        for (int r0 = 0; r0 < nr; r0 += ng) {
          const int r = r0 + threadIdx.x / K;
          double s = 0;
          if (r < nr) {
            const int a = rp[row0 + r] - k0, b = rp[row0 + r + 1] - k0;
            for (int t = a + g; t < b; t += K) s += prod[t];
          }
          s = group_sum(s, K);
          if (g == 0 && r < nr) epi(row0 + r, s);
        }
      }
    } else {
      double s = 0;
#pragma unroll
      for (int j = 0; j < IT; j++) s += pr[j];
      if (u + 1 < uend) fetch(u + 1);
      s = wave_sum(s);
      __syncthreads();
      if (lane == 0) ws[w] = s;
      __syncthreads();
      if (threadIdx.x == 0) {
        double t = 0;
        for (int i = 0; i < SPMV_BLOCK / WAVE; i++) t += ws[i];
        parts[-nr] = t;
      }
    }
  }
}

// One wavefront per long row: its chunk partials added left to right from
// 0.0 (wave_seq_fold keeps a block of partials in flight; one lane per row
// waited on each load in turn).
template <class Epi>
__global__ void __launch_bounds__(256) k_spmv_long(const int32_t *lrow, const int32_t *lpart, int64_t nlong,
                                                   const double *parts, Epi epi) {
  __shared__ double buf[(256 / WAVE) * SEQ_BLOCK];
  const int64_t j = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE;
  if (j >= nlong) return;  // the whole wavefront
  const double s = wave_seq_fold(lpart[j], lpart[j + 1], buf + (threadIdx.x / WAVE) * SEQ_BLOCK,
                                 [&](int64_t k) { return parts[k]; });
  if ((threadIdx.x & (WAVE - 1)) == 0) epi(lrow[j], s);
}

template <class T, class Val, class Epi>
void launch_spmv(const DCsr &A, Val val, const T *x, Epi epi, hipStream_t s) {
  if (A.U.nunits == 0) return;
  double *parts = A.U.parts.p;
  const int upw = spmv_upw(A.U.nunits);
  const int64_t grid = (A.U.nunits + upw - 1) / upw;
  const int64_t xbytes = A.ncols * (int64_t)sizeof(T);
  if (xbytes < ((int64_t)1 << 31))
    hipLaunchKernelGGL((k_spmv_units<T, Val, Epi, true>), dim3((unsigned)grid), dim3(SPMV_BLOCK), 0, s, A.rp.p,
                       A.ci.p, val, x, (int32_t)xbytes, A.U.row0.p, A.U.nrows.p, A.U.koff.p, A.U.nunits, upw,
                       A.U.row_items, parts, epi);
  else
    hipLaunchKernelGGL((k_spmv_units<T, Val, Epi, false>), dim3((unsigned)grid), dim3(SPMV_BLOCK), 0, s,
                       A.rp.p, A.ci.p, val, x, (int32_t)0, A.U.row0.p, A.U.nrows.p, A.U.koff.p, A.U.nunits,
                       upw, A.U.row_items, parts, epi);
  LG_LAUNCH_CHECK();
  if (A.U.nlong) {
    hipLaunchKernelGGL((k_spmv_long<Epi>), dim3((unsigned)((A.U.nlong + 3) / 4)), dim3(256), 0, s, A.U.lrow.p,
                       A.U.lpart.p, A.U.nlong, (const double *)parts, epi);
    LG_LAUNCH_CHECK();
  }
}

template <class T>
const T *vals_of(const DCsr &A);
template <>
const double *vals_of<double>(const DCsr &A) {
  return A.v.p;
}
template <>
const float *vals_of<float>(const DCsr &A) {
  return A.v32.p;
}

template <class T>
constexpr int prec_index() {
  return std::is_same<T, float>::value ? 1 : 0;
}

template <class T, class Epi>
void dispatch_spmv(const DCsr &A, const T *x, Epi epi, hipStream_t s) {
  const auto &SL = A.slices[prec_index<T>()];
  if (SL.size() > 1) {
    const size_t ns = SL.size();
    for (size_t k = 0; k < ns; k++) {
      const DCsr &P = *SL[k];
      ValArr<T> val{vals_of<T>(P)};
      if (k == 0) {
        launch_spmv<T>(P, val, x, EpiAccSet{A.acc.p}, s);
      } else if (k + 1 < ns) {
        launch_spmv<T>(P, val, x, EpiAccAdd{A.acc.p}, s);
      } else if (std::is_same<T, float>::value && A.d32.p) {
        launch_spmv<T>(P, val, x, EpiAccFinal<EpiDiag<T, Epi>>{A.acc.p, {A.d32.p, x, epi}}, s);
      } else {
        launch_spmv<T>(P, val, x, EpiAccFinal<Epi>{A.acc.p, epi}, s);
      }
    }
    return;
  }
  const T *v = vals_of<T>(A);
  if (std::is_same<T, float>::value && A.d32.p)
    launch_spmv<T>(A, ValArr<T>{v}, x, EpiDiag<T, Epi>{A.d32.p, x, epi}, s);
  else if (v) launch_spmv<T>(A, ValArr<T>{v}, x, epi, s);
  else launch_spmv<T>(A, ValOne<T>{}, x, epi, s);
}

// ------------------------------------------------------------- elementwise
template <class T>
__global__ void k_cheb_zero(const T *dinv, const T *b, T *xn, double beta, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    xn[i] = cheb_first<T>(beta, dinv[i], b[i]);
}

template <class T>
__global__ void k_prolong_add(const int32_t *agg, int64_t n, const T *xc, T *x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] += xc[agg[i]];
}

// Laplacian-form FP32 copy: off-diagonals rounded, diagonal slot 0, and the
// FP64 diagonal = -(sum of the rounded off-diagonals); one wavefront per row
__global__ void k_d2f(const double *a, float *b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}
__global__ void k_f2d(const float *a, double *b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (double)a[i];
}

// y = Pinv x, one wavefront per row of the dense row-major matrix.
template <class T>
__global__ void k_dense_apply(const double *__restrict__ Pm, int64_t n, const T *__restrict__ x,
                              T *__restrict__ y) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t nw = (int64_t)gridDim.x * blockDim.x / WAVE;
  for (int64_t row = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE; row < n; row += nw) {
    double s = 0;
    for (int64_t j = lane; j < n; j += WAVE) s += Pm[row * n + j] * (double)x[j];
    s = wave_sum(s);
    if (lane == 0) y[row] = (T)s;
  }
}

// ---------------------------------------------------------------- reductions
// Dot-product segment: one workgroup each.  4096 (not 16384) gives mid-size
// vectors enough workgroups (R-MAT 22 K-cycle 0.408 -> 0.359 s, batched 160^3
// grid 61 -> 50 ms per right-hand side).
constexpr int RED_CHUNK = 4096;

template <class T>
__global__ void __launch_bounds__(256) k_red_stage1(int np, const T *a0, const T *b0, const T *a1,
                                                    const T *b1, const T *a2, const T *b2,
                                                    const T *a3, const T *b3, const int64_t *lo,
                                                    const int64_t *hi, double *part) {
  const int s = blockIdx.x;
  const int64_t l = lo[s], h = hi[s];
  double acc[4] = {0, 0, 0, 0};
  const T *A[4] = {a0, a1, a2, a3};
  const T *B[4] = {b0, b1, b2, b3};
  for (int64_t i = l + threadIdx.x; i < h; i += 256) {
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (j < np) acc[j] += (double)A[j][i] * (B[j] ? (double)B[j][i] : 1.0);
  }
  __shared__ double ws[4][4];
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    double v = wave_sum(acc[j]);
    if (lane == 0) ws[j][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4 && (int)threadIdx.x < np) {
    int j = threadIdx.x;
    part[s * 4 + j] = ((ws[j][0] + ws[j][1]) + ws[j][2]) + ws[j][3];
  }
}

// Where a block's dot results go: out[c * np + j], or the Krylov scalar
// slots of block c followed by its scalar step.
struct RedSink {
  double *out = nullptr;
  double *sc = nullptr;
  int s0 = 0, s1 = 0, post = KP_NONE, fcg = 0;
  double rtol = 0;
  __device__ __forceinline__ void put(int c, int np, const double *v) const {
    if (!sc) {
      for (int j = 0; j < np; j++) out[c * np + j] = v[j];
      return;
    }
    double *S = sc + c * SC_N;
    S[s0] = v[0];
    if (np > 1) S[s1] = v[1];
    kr_post(S, post, fcg, rtol);
  }
};

// Both stages in one workgroup for plans of at most RED_FUSE_SEGS segments:
// four 256-thread groups reduce segments g, g + 4, ... with the tree of
// k_red_stage1, then one thread per block adds its segment sums in order as
// k_red_stage2 does.  Same bits as the two launches.  Only for small
// vectors: one workgroup reads at one CU's bandwidth, which loses to
// nseg parallel workgroups beyond two segments (measured on the K-cycle).
constexpr int RED_FUSE_SEGS = 2;
template <class T>
__global__ void __launch_bounds__(1024) k_red_fused(int np, int nseg, int nb, const T *a0, const T *b0,
                                                    const T *a1, const T *b1, const T *a2, const T *b2,
                                                    const T *a3, const T *b3, const int64_t *lo,
                                                    const int64_t *hi, const int32_t *bseg, RedSink sink) {
  __shared__ double part[RED_FUSE_SEGS * 4];
  __shared__ double ws[4][4][4];  // group, product, wavefront
  const int grp = threadIdx.x / 256, vt = threadIdx.x % 256;
  const int lane = threadIdx.x & (WAVE - 1), w = vt / WAVE;
  const T *A[4] = {a0, a1, a2, a3};
  const T *B[4] = {b0, b1, b2, b3};
  for (int s0 = 0; s0 < nseg; s0 += 4) {
    const int s = s0 + grp;
    double acc[4] = {0, 0, 0, 0};
    if (s < nseg) {
      const int64_t l = lo[s], h = hi[s];
      for (int64_t i = l + vt; i < h; i += 256) {
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (j < np) acc[j] += (double)A[j][i] * (B[j] ? (double)B[j][i] : 1.0);
      }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      double v = wave_sum(acc[j]);
      if (lane == 0) ws[grp][j][w] = v;
    }
    __syncthreads();
    if (vt < 4 && vt < np && s < nseg) {
      const int j = vt;
      part[s * 4 + j] = ((ws[grp][j][0] + ws[grp][j][1]) + ws[grp][j][2]) + ws[grp][j][3];
    }
    __syncthreads();
  }
  if ((int)threadIdx.x < nb) {
    const int c = threadIdx.x;
    double v[4];
    for (int j = 0; j < np; j++) {
      double s = 0;
      for (int g = bseg[c]; g < bseg[c + 1]; g++) s += part[g * 4 + j];
      v[j] = s;
    }
    sink.put(c, np, v);
  }
}

// One wavefront per block: lane l adds the block's segment sums
// l*G .. (l+1)*G - 1 in order (G = ceil(segments / 64)), then lane 0 adds
// the 64 lane sums in order.  With at most 64 segments per block this is the
// plain in-order sum of the segment sums (the fused kernel's second stage).
__global__ void __launch_bounds__(WAVE) k_red_stage2(int np, const int32_t *bseg, const double *part,
                                                    RedSink sink) {
  __shared__ double q[4][WAVE];
  const int c = blockIdx.x, lane = threadIdx.x;
  const int g0 = bseg[c], g1 = bseg[c + 1];
  const int G = (g1 - g0 + WAVE - 1) / WAVE;
  const int a = min(g0 + lane * G, g1), b = min(a + G, g1);
  for (int j = 0; j < np; j++) {
    double s = 0;
    for (int g = a; g < b; g++) s += part[g * 4 + j];
    q[j][lane] = s;
  }
  __syncthreads();
  if (lane != 0) return;
  double v[4];
  for (int j = 0; j < np; j++) {
    double s = 0;
    for (int l = 0; l < WAVE; l++) s += q[j][l];
    v[j] = s;
  }
  sink.put(c, np, v);
}

template <class T>
__global__ void __launch_bounds__(256) k_sub_mean(T *a, const int64_t *lo, const int64_t *hi,
                                                  const int32_t *segblk, const int64_t *bptr,
                                                  const double *sums) {
  const int s = blockIdx.x;
  const int c = segblk[s];
  const T m = (T)(sums[c] / (double)(bptr[c + 1] - bptr[c]));
  for (int64_t i = lo[s] + threadIdx.x; i < hi[s]; i += 256) a[i] -= m;
}

}  // namespace

// ---------------------------------------------------------------- host side

namespace {

// The greedy packing walks the rows from 0: a unit starting at a row of at
// most UNIT_NNZ entries takes it and the following rows while there are at
// most UNIT_ROWS of them and their entries fit in UNIT_NNZ; a longer row is
// cut into UNIT_NNZ-chunks.  unit_next(r) is the row after the unit that
// starts at r.  rp is non-decreasing, so "rows r..r1-1 fit" is monotone in
// r1 and a binary search finds the largest r1 (any row longer than UNIT_NNZ
// already breaks the fit, so it needs no test of its own).
__host__ __device__ inline int64_t unit_next(const int32_t *rp, int64_t n, int64_t r) {
  const int32_t base = rp[r];
  if (rp[r + 1] - base > UNIT_NNZ) return r + 1;
  int64_t lo = r + 1, hi = (n < r + UNIT_ROWS) ? n : r + UNIT_ROWS;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (rp[mid] - base <= UNIT_NNZ) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// jump table 0: next[r] for r < n, next[n] = n
__global__ void k_unit_next(int64_t n, const int32_t *rp, int32_t *J) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r <= n; r += (int64_t)gridDim.x * blockDim.x)
    J[r] = (int32_t)(r < n ? unit_next(rp, n, r) : n);
}
// table k+1 = table k applied twice
__global__ void k_unit_jump(int64_t n, const int32_t *Jk, int32_t *Jk1) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r <= n; r += (int64_t)gridDim.x * blockDim.x)
    Jk1[r] = Jk[Jk[r]];
}
// number of unit starts on the chain from row 0 (n > 0): the largest number
// of steps that stays below n, plus the start itself
__global__ void k_unit_chain_len(int64_t n, const int32_t *J, int K, int64_t stride, int64_t *out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int64_t pos = 0, m = 0;
  for (int k = K - 1; k >= 0; k--) {
    const int32_t q = J[(int64_t)k * stride + pos];
    if (q < n) {
      pos = q;
      m += (int64_t)1 << k;
    }
  }
  out[0] = m + 1;
}
// start m of the chain, its unit count and, for a long row, its chunk count
__global__ void k_unit_chain(int64_t M, const int32_t *J, int K, int64_t stride, const int32_t *rp,
                             int32_t *start, int32_t *ucnt, int32_t *pcnt, int32_t *islong) {
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < M; m += (int64_t)gridDim.x * blockDim.x) {
    int64_t pos = 0;
    for (int k = 0; k < K; k++)
      if ((m >> k) & 1) pos = J[(int64_t)k * stride + pos];
    const int32_t len = rp[pos + 1] - rp[pos];
    const bool lng = len > UNIT_NNZ;
    const int32_t chunks = (len + UNIT_NNZ - 1) / UNIT_NNZ;
    start[m] = (int32_t)pos;
    ucnt[m] = lng ? chunks : 1;
    pcnt[m] = lng ? chunks : 0;
    islong[m] = lng ? 1 : 0;
  }
}
__global__ void k_unit_fill(int64_t M, const int32_t *start, const int32_t *J0, const int32_t *rp,
                            const int32_t *uoff, const int32_t *poff, const int32_t *loff, const int32_t *islong,
                            int32_t *row0, int32_t *nrows, int32_t *koff, int32_t *lrow, int32_t *lpart) {
  for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < M; m += (int64_t)gridDim.x * blockDim.x) {
    const int32_t r = start[m], u = uoff[m];
    if (islong[m]) {
      const int32_t a = rp[r], b = rp[r + 1];
      int32_t j = 0;
      for (int32_t k = a; k < b; k += UNIT_NNZ, j++) {
        row0[u + j] = r;
        nrows[u + j] = -(poff[m] + j);
        koff[u + j] = k;
      }
      lrow[loff[m]] = r;
      lpart[loff[m]] = poff[m];
    } else {
      row0[u] = r;
      nrows[u] = J0[r] - r;
      koff[u] = rp[r];
    }
  }
}

// the unit tables from host vectors
void set_units(Units &U, int64_t nnz, const std::vector<int32_t> &row0, const std::vector<int32_t> &nrows,
               const std::vector<int32_t> &koff, const std::vector<int32_t> &lrow,
               const std::vector<int32_t> &lpart, int64_t parts, hipStream_t s) {
  U.nunits = (int64_t)row0.size();
  U.row_items = row_items_for(nnz);
  U.nlong = (int64_t)lrow.size();
  U.nparts = parts;
  U.row0.alloc(U.nunits);
  U.nrows.alloc(U.nunits);
  U.koff.alloc(U.nunits);
  to_device(U.row0.p, row0.data(), U.nunits, s);
  to_device(U.nrows.p, nrows.data(), U.nunits, s);
  to_device(U.koff.p, koff.data(), U.nunits, s);
  U.lrow.alloc(std::max<int64_t>(U.nlong, 1));
  U.lpart.alloc(U.nlong + 1);
  to_device(U.lrow.p, lrow.data(), U.nlong, s);
  to_device(U.lpart.p, lpart.data(), U.nlong + 1, s);
  U.parts.alloc(std::max<int64_t>(parts, 1));
  LG_HIP(hipStreamSynchronize(s));
}

}  // namespace

// The greedy packing on the host, row by row.
void build_units_host(DCsr &A, hipStream_t s) {
  const int64_t n = A.n;
  std::vector<int32_t> rp = download(A.rp.p, n + 1, s);
  std::vector<int32_t> row0, nrows, koff, lrow, lpart;
  int64_t r = 0, parts = 0;
  while (r < n) {
    const int64_t r1 = unit_next(rp.data(), n, r);
    if (rp[r + 1] - rp[r] > UNIT_NNZ) {
      lrow.push_back((int32_t)r);
      lpart.push_back((int32_t)parts);
      for (int64_t k = rp[r]; k < rp[r + 1]; k += UNIT_NNZ) {
        row0.push_back((int32_t)r);
        nrows.push_back(-(int32_t)parts);
        koff.push_back((int32_t)k);
        parts++;
      }
    } else {
      row0.push_back((int32_t)r);
      nrows.push_back((int32_t)(r1 - r));
      koff.push_back(rp[r]);
    }
    r = r1;
  }
  lpart.push_back((int32_t)parts);
  set_units(A.U, rp[n] - rp[0], row0, nrows, koff, lrow, lpart, parts, s);
}

// The same packing on the device.  The unit starts are the chain 0,
// next(0), next(next(0)), ... below n; binary lifting (tables of next applied
// 2^k times) gives the m-th start directly, so every start is found in
// parallel.  Small operators take the host pass, whose download and loop
// cost less than the launches.
void build_units(DCsr &A, hipStream_t s) {
  const int64_t n = A.n;
  if (n < UNITS_DEVICE_MIN_ROWS) {
    build_units_host(A, s);
    return;
  }
  int K = 1;
  while (((int64_t)1 << K) < n + 1) K++;
  const int64_t stride = n + 1;
  DevBuf<int32_t> J((size_t)K * stride);
  hipLaunchKernelGGL(k_unit_next, dim3(grid_for(stride, 256, 16384)), dim3(256), 0, s, n, A.rp.p, J.p);
  LG_LAUNCH_CHECK();
  for (int k = 0; k + 1 < K; k++) {
    hipLaunchKernelGGL(k_unit_jump, dim3(grid_for(stride, 256, 16384)), dim3(256), 0, s, n, J.p + k * stride,
                       J.p + (k + 1) * stride);
    LG_LAUNCH_CHECK();
  }
  DevBuf<int64_t> Md(1);
  hipLaunchKernelGGL(k_unit_chain_len, dim3(1), dim3(64), 0, s, n, J.p, K, stride, Md.p);
  LG_LAUNCH_CHECK();
  int64_t M = 0;
  to_host(&M, Md.p, 1, s);
  LG_HIP(hipStreamSynchronize(s));
  DevBuf<int32_t> start(M), ucnt(M + 1), pcnt(M + 1), islong(M + 1), uoff(M + 1), poff(M + 1), loff(M + 1);
  hipLaunchKernelGGL(k_unit_chain, dim3(grid_for(M, 256, 16384)), dim3(256), 0, s, M, J.p, K, stride, A.rp.p,
                     start.p, ucnt.p, pcnt.p, islong.p);
  LG_LAUNCH_CHECK();
  fill_i32(ucnt.p + M, 0, 1, s);
  fill_i32(pcnt.p + M, 0, 1, s);
  fill_i32(islong.p + M, 0, 1, s);
  const int64_t nunits = exclusive_scan_i32(ucnt.p, uoff.p, M + 1, s);
  const int64_t parts = exclusive_scan_i32(pcnt.p, poff.p, M + 1, s);
  const int64_t nlong = exclusive_scan_i32(islong.p, loff.p, M + 1, s);
  Units &U = A.U;
  int32_t ends[2];
  to_host(&ends[0], A.rp.p, 1, s);
  to_host(&ends[1], A.rp.p + n, 1, s);
  LG_HIP(hipStreamSynchronize(s));
  U.nunits = nunits;
  U.row_items = row_items_for((int64_t)ends[1] - ends[0]);
  U.nlong = nlong;
  U.nparts = parts;
  U.row0.alloc(nunits);
  U.nrows.alloc(nunits);
  U.koff.alloc(nunits);
  U.lrow.alloc(std::max<int64_t>(nlong, 1));
  U.lpart.alloc(nlong + 1);
  U.parts.alloc(std::max<int64_t>(parts, 1));
  hipLaunchKernelGGL(k_unit_fill, dim3(grid_for(M, 256, 16384)), dim3(256), 0, s, M, start.p, J.p, A.rp.p,
                     uoff.p, poff.p, loff.p, islong.p, U.row0.p, U.nrows.p, U.koff.p, U.lrow.p, U.lpart.p);
  LG_LAUNCH_CHECK();
  fill_i32(U.lpart.p + nlong, (int32_t)parts, 1, s);
  LG_HIP(hipStreamSynchronize(s));
}

namespace {
__device__ __forceinline__ int lower_bound_row(const int32_t *ci, int lo, int hi, int c) {
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (ci[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
__global__ void k_slice_count(int64_t n, const int32_t *rp, const int32_t *ci, int c0, int c1,
                              int32_t *cnt) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int a = lower_bound_row(ci, rp[i], rp[i + 1], c0);
    int b = lower_bound_row(ci, a, rp[i + 1], c1);
    cnt[i] = b - a;
  }
}
template <class V>
__global__ void k_slice_copy(int64_t n, const int32_t *rp, const int32_t *ci, const V *v, int c0,
                             const int32_t *srp, int32_t *sci, V *sv) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int a = lower_bound_row(ci, rp[i], rp[i + 1], c0);
    int len = srp[i + 1] - srp[i];
    for (int t = 0; t < len; t++) {
      sci[srp[i] + t] = ci[a + t];
      sv[srp[i] + t] = v[a + t];
    }
  }
}
}  // namespace

void build_slices(DCsr &A, int prec, int64_t width, hipStream_t s) {
  auto &SL = A.slices[prec];
  SL.clear();
  if (width <= 0 || width >= A.ncols) return;
  const int64_t n = A.n;
  const int64_t ns = (A.ncols + width - 1) / width;
  DevBuf<int32_t> cnt(n + 1);
  for (int64_t k = 0; k < ns; k++) {
    const int c0 = (int)(k * width), c1 = (int)std::min<int64_t>((k + 1) * width, A.ncols);
    std::unique_ptr<DCsr> P(new DCsr());
    P->n = n;
    P->ncols = A.ncols;
    hipLaunchKernelGGL(k_slice_count, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, n, A.rp.p, A.ci.p,
                       c0, c1, cnt.p);
    LG_LAUNCH_CHECK();
    LG_HIP(hipMemsetAsync(cnt.p + n, 0, sizeof(int32_t), s));
    P->rp.alloc(n + 1);
    P->nnz = exclusive_scan_i32(cnt.p, P->rp.p, n + 1, s);
    P->ci.alloc(std::max<int64_t>(P->nnz, 1));
    if (prec == 0) {
      P->v.alloc(std::max<int64_t>(P->nnz, 1));
      hipLaunchKernelGGL(k_slice_copy<double>, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, n, A.rp.p,
                         A.ci.p, (const double *)A.v.p, c0, (const int32_t *)P->rp.p, P->ci.p, P->v.p);
    } else {
      P->v32.alloc(std::max<int64_t>(P->nnz, 1));
      hipLaunchKernelGGL(k_slice_copy<float>, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, n, A.rp.p,
                         A.ci.p, (const float *)A.v32.p, c0, (const int32_t *)P->rp.p, P->ci.p,
                         P->v32.p);
    }
    LG_LAUNCH_CHECK();
    build_units(*P, s);
    SL.push_back(std::move(P));
  }
  A.acc.ensure(std::max<int64_t>(n, 1));
  LG_HIP(hipStreamSynchronize(s));
}

double time_spmv(const DCsr &A, int prec, int iters, hipStream_t s) {
  const int64_t n = A.n;
  DevBuf<double> x(prec == 0 ? n : 0), y(prec == 0 ? n : 0);
  DevBuf<float> x32(prec == 1 ? n : 0), y32(prec == 1 ? n : 0);
  if (prec == 0) fill_f64(x.p, 1.0, n, s);
  else {
    DevBuf<double> t(n);
    fill_f64(t.p, 1.0, n, s);
    op_d2f(t.p, x32.p, n, s);
    LG_HIP(hipStreamSynchronize(s));
  }
  auto run = [&]() {
    if (prec == 0) op_spmv<double>(A, x.p, y.p, s);
    else op_spmv<float>(A, x32.p, y32.p, s);
  };
  run();
  hipEvent_t e0, e1;
  LG_HIP(hipEventCreate(&e0));
  LG_HIP(hipEventCreate(&e1));
  LG_HIP(hipEventRecord(e0, s));
  for (int i = 0; i < iters; i++) run();
  LG_HIP(hipEventRecord(e1, s));
  LG_HIP(hipEventSynchronize(e1));
  float ms = 0;
  LG_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / iters;
}

void make_v32(DCsr &A, hipStream_t s, bool laplacian) {
  if (!A.v.p) return;
  if (laplacian) {
    lap32_form(A, s);
  } else {
    A.v32.alloc(std::max<int64_t>(A.nnz, 1));
    op_d2f(A.v.p, A.v32.p, A.nnz, s);
  }
}

template <class T>
void op_spmv(const DCsr &A, const T *x, T *y, hipStream_t s) {
  dispatch_spmv<T>(A, x, EpiStore<T>{y}, s);
}

template <class T>
void op_residual(const DCsr &A, const T *x, const T *b, T *r, hipStream_t s) {
  dispatch_spmv<T>(A, x, EpiResid<T>{b, r}, s);
}

template <class T>
void op_cheb_step(const DCsr &A, const T *dinv, const T *b, const T *xc, const T *xp, T *xn,
                  double alpha, double beta, hipStream_t s) {
  if (!xc) {
    hipLaunchKernelGGL(k_cheb_zero<T>, dim3(grid_for(A.n, 256, 8192)), dim3(256), 0, s, dinv, b, xn,
                       beta, A.n);
    LG_LAUNCH_CHECK();
    return;
  }
  dispatch_spmv<T>(A, xc, EpiCheb<T>{dinv, b, xc, xp, xn, alpha, beta}, s);
}

template <class T>
void op_interp(const DCsr &P, const T *q, const T *xc, const T *b, T *x, hipStream_t s) {
  dispatch_spmv<T>(P, xc, EpiInterp<T>{q, b, x}, s);
}

template <class T>
void op_prolong_add(const int32_t *agg, int64_t n, const T *xc, T *x, hipStream_t s) {
  hipLaunchKernelGGL(k_prolong_add<T>, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, agg, n, xc, x);
  LG_LAUNCH_CHECK();
}

template <class T>
void op_dense_apply(const double *Pm, int64_t n, const T *x, T *y, hipStream_t s) {
  hipLaunchKernelGGL(k_dense_apply<T>, dim3(grid_for(n * WAVE, 256)), dim3(256), 0, s, Pm, n, x, y);
  LG_LAUNCH_CHECK();
}

void op_d2f(const double *a, float *b, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_d2f, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, a, b, n);
  LG_LAUNCH_CHECK();
}
void op_f2d(const float *a, double *b, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_f2d, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, a, b, n);
  LG_LAUNCH_CHECK();
}

#define LG_INST(T)                                                                                \
  template void op_spmv<T>(const DCsr &, const T *, T *, hipStream_t);                            \
  template void op_residual<T>(const DCsr &, const T *, const T *, T *, hipStream_t);             \
  template void op_cheb_step<T>(const DCsr &, const T *, const T *, const T *, const T *, T *,    \
                                double, double, hipStream_t);                                     \
  template void op_interp<T>(const DCsr &, const T *, const T *, const T *, T *, hipStream_t);    \
  template void op_prolong_add<T>(const int32_t *, int64_t, const T *, T *, hipStream_t);         \
  template void op_dense_apply<T>(const double *, int64_t, const T *, T *, hipStream_t);
LG_INST(double)
LG_INST(float)
#undef LG_INST

void RedPlan::build(int64_t n_, const std::vector<int64_t> &blocks, hipStream_t s) {
  n = n_;
  hbptr = blocks;
  if (hbptr.size() < 2) hbptr = {0, n};
  nb = (int)hbptr.size() - 1;
  std::vector<int64_t> lo, hi;
  std::vector<int32_t> bs(nb + 1, 0), segblk;
  for (int c = 0; c < nb; c++) {
    bs[c] = (int32_t)lo.size();
    for (int64_t a = hbptr[c]; a < hbptr[c + 1]; a += RED_CHUNK) {
      lo.push_back(a);
      hi.push_back(std::min<int64_t>(a + RED_CHUNK, hbptr[c + 1]));
      segblk.push_back(c);
    }
  }
  bs[nb] = (int32_t)lo.size();
  nseg = (int)lo.size();
  seg_lo.alloc(std::max(nseg, 1));
  seg_hi.alloc(std::max(nseg, 1));
  bseg.alloc(nb + 1 + std::max(nseg, 1));  // bseg then segment->block map
  bptr.alloc(nb + 1);
  part.alloc(std::max(nseg, 1) * 4);
  to_device(seg_lo.p, lo.data(), nseg, s);
  to_device(seg_hi.p, hi.data(), nseg, s);
  to_device(bseg.p, bs.data(), nb + 1, s);
  to_device(bseg.p + nb + 1, segblk.data(), nseg, s);
  to_device(bptr.p, hbptr.data(), nb + 1, s);
  LG_HIP(hipStreamSynchronize(s));
}

namespace {
template <class T>
void run_dots(const RedPlan &R, int np, const T *const *a, const T *const *b, const RedSink &sink,
              hipStream_t s) {
  if (R.nseg == 0) return;
  const T *A[4] = {nullptr, nullptr, nullptr, nullptr}, *B[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int j = 0; j < np; j++) {
    A[j] = a[j];
    B[j] = b[j];
  }
  if (R.nseg <= RED_FUSE_SEGS && R.nb <= 1024) {  // one launch; a thread per block for the second stage
    hipLaunchKernelGGL(k_red_fused<T>, dim3(1), dim3(1024), 0, s, np, R.nseg, R.nb, A[0], B[0], A[1], B[1],
                       A[2], B[2], A[3], B[3], R.seg_lo.p, R.seg_hi.p, R.bseg.p, sink);
    LG_LAUNCH_CHECK();
    return;
  }
  hipLaunchKernelGGL(k_red_stage1<T>, dim3(R.nseg), dim3(256), 0, s, np, A[0], B[0], A[1], B[1], A[2],
                     B[2], A[3], B[3], R.seg_lo.p, R.seg_hi.p, R.part.p);
  LG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_red_stage2, dim3(R.nb), dim3(WAVE), 0, s, np, R.bseg.p, R.part.p, sink);
  LG_LAUNCH_CHECK();
}
}  // namespace

template <class T>
void RedPlan::dots(int np, const T *const *a, const T *const *b, double *out, hipStream_t s) const {
  RedSink sink;
  sink.out = out;
  run_dots<T>(*this, np, a, b, sink, s);
}

template <class T>
void RedPlan::dots_sc(int np, const T *const *a, const T *const *b, double *sc, int s0, int s1, int post,
                      int fcg, double rtol, hipStream_t s) const {
  LG_CHECK(np >= 1 && np <= 2, 1, "dots_sc: one or two products");
  RedSink sink;
  sink.sc = sc;
  sink.s0 = s0;
  sink.s1 = s1;
  sink.post = post;
  sink.fcg = fcg;
  sink.rtol = rtol;
  run_dots<T>(*this, np, a, b, sink, s);
}

template <class T>
void RedPlan::sums(const T *a, double *out, hipStream_t s) const {
  const T *A[1] = {a}, *B[1] = {nullptr};
  dots<T>(1, A, B, out, s);
}

template <class T>
void op_sub_block_mean(T *a, const RedPlan &P, const double *sums, hipStream_t s) {
  if (P.nseg == 0) return;
  hipLaunchKernelGGL(k_sub_mean<T>, dim3(P.nseg), dim3(256), 0, s, a, P.seg_lo.p, P.seg_hi.p,
                     P.bseg.p + P.nb + 1, P.bptr.p, sums);
  LG_LAUNCH_CHECK();
}

template void RedPlan::dots<double>(int, const double *const *, const double *const *, double *,
                                    hipStream_t) const;
template void RedPlan::dots<float>(int, const float *const *, const float *const *, double *,
                                   hipStream_t) const;
template void RedPlan::dots_sc<double>(int, const double *const *, const double *const *, double *, int, int,
                                       int, int, double, hipStream_t) const;
template void RedPlan::dots_sc<float>(int, const float *const *, const float *const *, double *, int, int, int,
                                      int, double, hipStream_t) const;
template void RedPlan::sums<double>(const double *, double *, hipStream_t) const;
template void RedPlan::sums<float>(const float *, double *, hipStream_t) const;
template void op_sub_block_mean<double>(double *, const RedPlan &, const double *, hipStream_t);
template void op_sub_block_mean<float>(float *, const RedPlan &, const double *, hipStream_t);

}  // namespace lg
