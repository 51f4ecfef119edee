// Device code shared by the per-operator solve kernels (solve.hip,
// solver.hip) and the fused coarse-tail kernel (tail.hip): value access,
// SpMV epilogues, the wavefront sum and the per-block CG / FCG scalar
// logic.  Both paths evaluate the same expressions in the same order, so a
// level solved inside the fused tail produces the bits the separate
// launches produce.
#pragma once
#include "common.hpp"

namespace lg {
namespace sdev {

// ------------------------------------------------------------ value access
// Matrix data is streamed once per product: non-temporal loads keep it from
// displacing the gathered vector in L2.
template <class T>
struct ValOne {
  __device__ __forceinline__ T operator()(int64_t) const { return (T)1; }
};
template <class T>
struct ValArr {
  const T *v;
  __device__ __forceinline__ T operator()(int64_t k) const { return __builtin_nontemporal_load(v + k); }
};

// --------------------------------------------------------------- epilogues
// Row sums are always accumulated in FP64 (a product of two FP32 numbers
// is exact in FP64), so the FP32 path only rounds storage, not the
// heavily cancelling Laplacian row sums of hub vertices.
template <class T>
struct EpiStore {
  T *y;
  __device__ __forceinline__ void operator()(int64_t r, double s) const { y[r] = (T)s; }
};
template <class T>
struct EpiResid {
  const T *b;
  T *r;
  __device__ __forceinline__ void operator()(int64_t i, double s) const {
    r[i] = (T)((double)b[i] - s);
  }
};
template <class T>
struct EpiCheb {
  const T *dinv, *b, *xc, *xp;
  T *xn;
  double alpha, beta;
  __device__ __forceinline__ void operator()(int64_t i, double s) const {
    double c = (double)xc[i];
    // a null xp is a zero previous iterate (second step from a zero guess);
    // every other caller passes alpha = 0 with it
    double d = alpha * (c - (xp ? (double)xp[i] : 0.0));
    xn[i] = (T)(c + d + beta * (double)dinv[i] * ((double)b[i] - s));
  }
};
template <class T>
struct EpiInterp {
  const T *q, *b;
  T *x;
  __device__ __forceinline__ void operator()(int64_t i, double s) const {
    x[i] = (T)(s + (double)q[i] * (double)b[i]);
  }
};
// adds d[i] x[i] (FP64) before the wrapped epilogue: the diagonal term of an
// FP32 operator stored in Laplacian form
template <class T, class Epi>
struct EpiDiag {
  const double *d;
  const T *x;
  Epi e;
  __device__ __forceinline__ void operator()(int64_t i, double s) const {
    e(i, s + d[i] * (double)x[i]);
  }
};

// first Chebyshev step from a zero iterate: xn = beta dinv b
template <class T>
__device__ __forceinline__ T cheb_first(double beta, T dinv, T b) {
  return (T)(beta * (double)dinv * (double)b);
}

// xor butterfly: every lane ends with the same bits
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, WAVE);
  return v;
}

// xor butterfly over aligned groups of K lanes (K a power of two <= 64)
__device__ __forceinline__ double group_sum(double v, int K) {
  for (int off = K >> 1; off >= 1; off >>= 1) v += __shfl_xor(v, off, WAVE);
  return v;
}

// ------------------------------------------------- CG / FCG scalar logic
// One set of scalars per (block, column): the Krylov coefficients of each
// independent graph are kept apart.
enum { SC_BN2 = 0, SC_RN2, SC_GAMMA, SC_DELTA, SC_T1, SC_ALPHA, SC_BETA, SC_ACTIVE, SC_GN, SC_N };

__device__ __forceinline__ void kr_init(double *S) {
  for (int j = 0; j < SC_N; j++) S[j] = 0;
  S[SC_ACTIVE] = 1;
}
__device__ __forceinline__ void kr_alpha(double *S, int fcg) {
  double num = fcg ? S[SC_T1] : S[SC_GAMMA];
  S[SC_ALPHA] = (S[SC_ACTIVE] != 0 && S[SC_DELTA] != 0) ? num / S[SC_DELTA] : 0.0;
}
__device__ __forceinline__ void kr_check(double *S, double rtol) {
  if (S[SC_ACTIVE] != 0 && !(sqrt(S[SC_RN2]) > rtol * sqrt(S[SC_BN2]))) S[SC_ACTIVE] = 0;
}
__device__ __forceinline__ void kr_beta(double *S, int fcg) {
  if (fcg) {
    S[SC_BETA] = (S[SC_ACTIVE] != 0 && S[SC_DELTA] != 0) ? -S[SC_T1] / S[SC_DELTA] : 0.0;
  } else {
    S[SC_BETA] = (S[SC_ACTIVE] != 0 && S[SC_GAMMA] != 0) ? S[SC_GN] / S[SC_GAMMA] : 0.0;
    S[SC_GAMMA] = S[SC_GN];
  }
}

// the scalar step that follows a dot product in CG / FCG
enum KrPost { KP_NONE = 0, KP_ALPHA, KP_CHECK, KP_BETA };
__device__ __forceinline__ void kr_post(double *S, int post, int fcg, double rtol) {
  if (post == KP_ALPHA) kr_alpha(S, fcg);
  else if (post == KP_CHECK) kr_check(S, rtol);
  else if (post == KP_BETA) kr_beta(S, fcg);
}

}  // namespace sdev
}  // namespace lg
