// Fused coarse tail: one workgroup executes a recorded program of solve
// operations.  Each instruction replays the per-operator kernel named in
// its comment with the same summation order.
#include <vector>

#include "../../include/ligmg_amd.h"
#include "solve_dev.hpp"
#include "tail.hpp"

namespace lg {

namespace {

using namespace sdev;

constexpr int NW = TAIL_THREADS / WAVE;          // wavefronts
constexpr int NG = TAIL_THREADS / SPMV_BLOCK;    // 256-thread groups
constexpr int GW = SPMV_BLOCK / WAVE;            // wavefronts per group
constexpr int IT = UNIT_NNZ / SPMV_BLOCK;
static_assert(SPMV_BLOCK == 256, "the replayed trees are those of 256-thread workgroups");

// ------------------------------------------------------------------ SpMV
// k_spmv_units + k_spmv_long.  Thread rows: one thread sums the row left to
// right (entries loaded four at a time, added in order).  Group rows: K
// strided partials, xor butterfly over the K lanes.  Long rows: each
// UNIT_NNZ-chunk is the tree of a 256-thread workgroup (8 products per
// thread added in order, butterfly per wavefront, the 4 wavefront sums
// added from 0) on one group; the chunk sums are then added in order.
template <class T, class Val, class Epi>
__device__ __forceinline__ void t_spmv(const TailOp<T> &A, Val val, const T *__restrict__ x, Epi epi,
                                       double *ws) {
  const int32_t *__restrict__ rp = A.rp;
  const int32_t *__restrict__ ci = A.ci;
  for (int j = threadIdx.x; j < A.ntrow; j += TAIL_THREADS) {
    const int r = A.trow[j];
    const int b = rp[r + 1];
    double s = 0;
    int k = rp[r];
    for (; k + 4 <= b; k += 4) {
      const int c0 = ci[k], c1 = ci[k + 1], c2 = ci[k + 2], c3 = ci[k + 3];
      const T v0 = val(k), v1 = val(k + 1), v2 = val(k + 2), v3 = val(k + 3);
      const double x0 = (double)x[c0], x1 = (double)x[c1], x2 = (double)x[c2], x3 = (double)x[c3];
      s += (double)v0 * x0;
      s += (double)v1 * x1;
      s += (double)v2 * x2;
      s += (double)v3 * x3;
    }
    for (; k < b; k++) s += (double)val(k) * (double)x[ci[k]];
    epi(r, s);
  }
  for (int c = 0; c < TAIL_GCLASSES; c++) {
    const int K = 2 << c, lo = A.gofs[c], hi = A.gofs[c + 1];
    const int g = threadIdx.x & (K - 1), ng = TAIL_THREADS / K;
    // uniform trip count: the butterflies stay converged
    for (int j0 = lo; j0 < hi; j0 += ng) {
      const int j = j0 + threadIdx.x / K;
      double s = 0;
      int r = 0;
      if (j < hi) {
        r = A.grow[j];
        for (int k = rp[r] + g; k < rp[r + 1]; k += K) s += (double)val(k) * (double)x[ci[k]];
      }
      s = group_sum(s, K);
      if (g == 0 && j < hi) epi(r, s);
    }
  }
  const int lane = threadIdx.x & (WAVE - 1);
  if (A.nlong == 0) return;
  const int grp = threadIdx.x / SPMV_BLOCK, vt = threadIdx.x % SPMV_BLOCK, w = vt / WAVE;
  for (int j0 = 0; j0 < A.nchunk; j0 += NG) {
    const int j = j0 + grp;
    double sv = 0;
    int slot = 0;
    if (j < A.nchunk) {
      const int row = A.chunk[3 * j], k0 = A.chunk[3 * j + 1];
      slot = A.chunk[3 * j + 2];
      const int cnt = min(k0 + UNIT_NNZ, rp[row + 1]) - k0;
      double pr[IT];
#pragma unroll
      for (int i = 0; i < IT; i++) {
        const int t = vt + i * SPMV_BLOCK;
        pr[i] = (t < cnt) ? (double)val(k0 + t) * (double)x[ci[k0 + t]] : 0.0;
      }
#pragma unroll
      for (int i = 0; i < IT; i++) sv += pr[i];
    }
    sv = wave_sum(sv);
    if (lane == 0) ws[grp * GW + w] = sv;
    __syncthreads();
    if (vt == 0 && j < A.nchunk) {
      double s = 0;
      for (int q = 0; q < GW; q++) s += ws[grp * GW + q];
      A.parts[slot] = s;
    }
    __syncthreads();
  }
  for (int j = threadIdx.x; j < A.nlong; j += TAIL_THREADS) {
    double s = 0;
    for (int p = A.lpart[j]; p < A.lpart[j + 1]; p++) s += A.parts[p];
    epi(A.lrow[j], s);
  }
}

// dispatch_spmv: FP32 Laplacian-form operators add their FP64 diagonal
template <class T, class Epi>
__device__ __forceinline__ void t_spmv_any(const TailOp<T> &A, const T *x, Epi epi, double *ws) {
  if (A.d) t_spmv<T>(A, ValArr<T>{A.v}, x, EpiDiag<T, Epi>{A.d, x, epi}, ws);
  else if (A.v) t_spmv<T>(A, ValArr<T>{A.v}, x, epi, ws);
  else t_spmv<T>(A, ValOne<T>{}, x, epi, ws);
}

// ------------------------------------------------------------- reductions
// RedPlan::dots_sc: k_red_stage1 (one chunk per 256-thread group) +
// k_red_stage2 (chunk sums of a block added in order, stored into the
// scalar slots, then the scalar step).
// s0 < 0: the block sums go to bsum (nullspace projection).
template <class T>
__device__ __forceinline__ void t_dots(const TailLevel<T> &L, const TailParams &P, int np, const T *a0,
                                       const T *b0, const T *a1, const T *b1, int s0, int s1, int post,
                                       double *ws) {
  const int grp = threadIdx.x / SPMV_BLOCK, vt = threadIdx.x % SPMV_BLOCK, w = vt / WAVE;
  const int lane = threadIdx.x & (WAVE - 1);
  for (int g0 = 0; g0 < L.nseg; g0 += NG) {
    const int g = g0 + grp;
    double c0 = 0, c1 = 0;
    if (g < L.nseg) {
      const int64_t hi = L.seg_hi[g];
      for (int64_t i = L.seg_lo[g] + vt; i < hi; i += SPMV_BLOCK) {
        c0 += (double)a0[i] * (b0 ? (double)b0[i] : 1.0);
        if (np > 1) c1 += (double)a1[i] * (b1 ? (double)b1[i] : 1.0);
      }
    }
    c0 = wave_sum(c0);
    c1 = wave_sum(c1);
    if (lane == 0) {
      ws[(grp * GW + w) * 2] = c0;
      ws[(grp * GW + w) * 2 + 1] = c1;
    }
    __syncthreads();
    if (vt < np && g < L.nseg) {
      const double *q = ws + grp * GW * 2 + vt;
      L.part[g * 4 + vt] = ((q[0] + q[2]) + q[4]) + q[6];
    }
    __syncthreads();
  }
  for (int c = threadIdx.x; c < L.nb; c += TAIL_THREADS) {
    double r0 = 0, r1 = 0;
    for (int g = L.bseg[c]; g < L.bseg[c + 1]; g++) r0 += L.part[g * 4];
    if (np > 1)
      for (int g = L.bseg[c]; g < L.bseg[c + 1]; g++) r1 += L.part[g * 4 + 1];
    if (s0 < 0) {
      L.bsum[c] = r0;
      continue;
    }
    double *S = L.sc + (int64_t)c * SC_N;
    S[s0] = r0;
    if (np > 1) S[s1] = r1;
    if (post == TP_ALPHA) kr_alpha(S, P.fcg);
    else if (post == TP_CHECK) kr_check(S, P.rtol);
    else if (post == TP_BETA) kr_beta(S, P.fcg);
  }
}

// block of vertex i (blocks are contiguous ranges of bptr)
template <class T>
__device__ __forceinline__ int t_block(const TailLevel<T> &L, int64_t i) {
  int lo = 0, hi = L.nb;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (L.bptr[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// wavefront-uniform copies (scalar registers) of values read from LDS
__device__ __forceinline__ int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ const void *uni(const void *p) {
  const uint64_t u = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  return (const void *)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double uni(double d) {
  const uint64_t u = __double_as_longlong(d);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// instructions staged through LDS in chunks of TAIL_CHUNK
constexpr int TAIL_CHUNK = 64;
static_assert(sizeof(TailInstr) % 16 == 0, "instructions are copied in 16-byte pieces");

template <class T>
__global__ void __launch_bounds__(TAIL_THREADS)
    k_tail(const TailInstr *__restrict__ prog, int64_t len, const TailOp<T> *__restrict__ ops,
           const TailLevel<T> *__restrict__ lv, TailParams P) {
  __shared__ double ws[NW * 2];
  constexpr int PIECES = (int)(sizeof(TailInstr) / 16);
  __shared__ uint4 stage_raw[TAIL_CHUNK * PIECES];
  const TailInstr *stage = (const TailInstr *)stage_raw;
  const int lane = threadIdx.x & (WAVE - 1);
  for (int64_t k = 0; k < len; k++) {
    const int slot = (int)(k % TAIL_CHUNK);
    if (slot == 0) {
      const int cnt = (int)min<int64_t>(TAIL_CHUNK, len - k) * PIECES;
      const uint4 *src = (const uint4 *)(prog + k);
      for (int t = threadIdx.x; t < cnt; t += TAIL_THREADS) stage_raw[t] = src[t];
      __syncthreads();
    }
    TailInstr I;
    {
      const TailInstr &S = stage[slot];
      I.op = uni(S.op);
      I.lvl = uni(S.lvl);
      I.i0 = uni(S.i0);
      I.i1 = uni(S.i1);
      I.i2 = uni(S.i2);
      I.i3 = uni(S.i3);
      for (int q = 0; q < 6; q++) I.p[q] = uni(S.p[q]);
      I.f[0] = uni(S.f[0]);
      I.f[1] = uni(S.f[1]);
    }
    const int n = I.i0;
    switch (I.op) {
      case TI_SPMV: {
        const TailOp<T> &A = ops[I.i0];
        const T *x = (const T *)I.p[0];
        if (I.i1 == TE_STORE) {
          t_spmv_any<T>(A, x, EpiStore<T>{(T *)I.p[1]}, ws);
        } else if (I.i1 == TE_RESID) {
          t_spmv_any<T>(A, x, EpiResid<T>{(const T *)I.p[1], (T *)I.p[2]}, ws);
        } else if (I.i1 == TE_CHEB) {
          t_spmv_any<T>(A, x,
                        EpiCheb<T>{(const T *)I.p[1], (const T *)I.p[2], x, (const T *)I.p[3], (T *)I.p[4],
                                   I.f[0], I.f[1]},
                        ws);
        } else {
          t_spmv_any<T>(A, x, EpiInterp<T>{(const T *)I.p[1], (const T *)I.p[2], (T *)I.p[3]}, ws);
        }
        break;
      }
      case TI_CHEB0: {  // k_cheb_zero
        const T *dinv = (const T *)I.p[0], *b = (const T *)I.p[1];
        T *xn = (T *)I.p[2];
        for (int i = threadIdx.x; i < n; i += TAIL_THREADS) xn[i] = cheb_first<T>(I.f[0], dinv[i], b[i]);
        break;
      }
      case TI_ZERO: {  // k_zero
        T *p = (T *)I.p[0];
        for (int i = threadIdx.x; i < n; i += TAIL_THREADS) p[i] = (T)0;
        break;
      }
      case TI_COPY: {  // hipMemcpyAsync device to device
        T *d = (T *)I.p[0];
        const T *s = (const T *)I.p[1];
        for (int i = threadIdx.x; i < n; i += TAIL_THREADS) d[i] = s[i];
        break;
      }
      case TI_PROLONG: {  // k_prolong_add
        const int32_t *agg = (const int32_t *)I.p[0];
        const T *xc = (const T *)I.p[1];
        T *x = (T *)I.p[2];
        for (int i = threadIdx.x; i < n; i += TAIL_THREADS) x[i] += xc[agg[i]];
        break;
      }
      case TI_SCALE: {  // k_scale
        T *p = (T *)I.p[0];
        const T a = (T)I.f[0];
        for (int i = threadIdx.x; i < n; i += TAIL_THREADS) p[i] *= a;
        break;
      }
      case TI_DENSE: {  // k_dense_apply: one wavefront per row
        const double *Pm = (const double *)I.p[0];
        const T *x = (const T *)I.p[1];
        T *y = (T *)I.p[2];
        for (int row = threadIdx.x / WAVE; row < n; row += NW) {
          double s = 0;
          for (int j = lane; j < n; j += WAVE) s += Pm[(int64_t)row * n + j] * (double)x[j];
          s = wave_sum(s);
          if (lane == 0) y[row] = (T)s;
        }
        break;
      }
      case TI_KR_INIT: {  // k_kr_init
        const TailLevel<T> &L = lv[I.lvl];
        for (int c = threadIdx.x; c < L.nb; c += TAIL_THREADS) kr_init(L.sc + (int64_t)c * SC_N);
        break;
      }
      case TI_DOTS: {
        const TailLevel<T> &L = lv[I.lvl];
        t_dots<T>(L, P, I.i0, (const T *)I.p[0], (const T *)I.p[1], (const T *)I.p[2], (const T *)I.p[3],
                  I.i1, I.i2, I.i3, ws);
        break;
      }
      case TI_UPDATE: {  // k_kr_update (blocks with alpha 0 untouched)
        const TailLevel<T> &L = lv[I.lvl];
        T *x = (T *)I.p[0];
        for (int i = threadIdx.x; i < L.n; i += TAIL_THREADS) {
          const double a = L.sc[(int64_t)t_block(L, i) * SC_N + SC_ALPHA];
          if (a == 0.0) continue;
          const T at = (T)a;
          x[i] += at * L.kp[i];
          L.kr[i] -= at * L.kq[i];
        }
        break;
      }
      case TI_PUPDATE: {  // k_kr_pupdate
        const TailLevel<T> &L = lv[I.lvl];
        for (int i = threadIdx.x; i < L.n; i += TAIL_THREADS) {
          const T b = (T)L.sc[(int64_t)t_block(L, i) * SC_N + SC_BETA];
          L.kp[i] = L.kz[i] + b * L.kp[i];
        }
        break;
      }
      case TI_SUBMEAN: {  // k_sub_mean
        const TailLevel<T> &L = lv[I.lvl];
        T *a = (T *)I.p[0];
        for (int i = threadIdx.x; i < L.n; i += TAIL_THREADS) {
          const int c = t_block(L, i);
          const T m = (T)(L.bsum[c] / (double)(L.bptr[c + 1] - L.bptr[c]));
          a[i] -= m;
        }
        break;
      }
      default:
        break;
    }
    // the last instruction of a chunk also ends its reads of the stage
    __syncthreads();
  }
}

}  // namespace

void tail_schedule(const DCsr &A, TailSched &S, hipStream_t s) {
  const Units &U = A.U;
  const std::vector<int32_t> rp = download(A.rp.p, A.n + 1, s);
  const std::vector<int32_t> row0 = download(U.row0.p, U.nunits, s);
  const std::vector<int32_t> nrows = download(U.nrows.p, U.nunits, s);
  const std::vector<int32_t> koff = download(U.koff.p, U.nunits, s);
  std::vector<int32_t> tr, ch;
  std::vector<int32_t> gr[TAIL_GCLASSES];
  for (int64_t u = 0; u < U.nunits; u++) {
    const int32_t nr = nrows[u];
    if (nr > 0) {
      // the lanes k_spmv_units gives each row of the unit
      const int K = row_lanes(rp[row0[u] + nr] - koff[u], nr, U.row_items);
      int c = 0;
      while ((2 << c) < K) c++;
      auto &dst = K == 1 ? tr : gr[c];
      for (int32_t r = 0; r < nr; r++) dst.push_back(row0[u] + r);
    } else {
      ch.push_back(row0[u]);
      ch.push_back(koff[u]);
      ch.push_back(-nr);
    }
  }
  std::vector<int32_t> all;
  for (int c = 0; c < TAIL_GCLASSES; c++) {
    S.gofs[c] = (int32_t)all.size();
    all.insert(all.end(), gr[c].begin(), gr[c].end());
  }
  S.gofs[TAIL_GCLASSES] = (int32_t)all.size();
  S.ntrow = (int32_t)tr.size();
  S.nchunk = (int32_t)(ch.size() / 3);
  S.trow.alloc(tr.size());
  S.grow.alloc(all.size());
  S.chunk.alloc(ch.size());
  to_device(S.trow.p, tr.data(), tr.size(), s);
  to_device(S.grow.p, all.data(), all.size(), s);
  to_device(S.chunk.p, ch.data(), ch.size(), s);
  LG_HIP(hipStreamSynchronize(s));
}

template <class T>
void tail_launch(const TailInstr *prog, int64_t len, const TailOp<T> *ops, const TailLevel<T> *lv,
                 const TailParams &P, hipStream_t s) {
  if (len <= 0) return;
  hipLaunchKernelGGL(k_tail<T>, dim3(1), dim3(TAIL_THREADS), 0, s, prog, len, ops, lv, P);
  LG_LAUNCH_CHECK();
}

template void tail_launch<double>(const TailInstr *, int64_t, const TailOp<double> *, const TailLevel<double> *,
                                  const TailParams &, hipStream_t);
template void tail_launch<float>(const TailInstr *, int64_t, const TailOp<float> *, const TailLevel<float> *,
                                 const TailParams &, hipStream_t);

}  // namespace lg
