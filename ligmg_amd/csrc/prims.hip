// Deterministic device primitives (see prims.hpp).
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include "prims.hpp"

namespace lg {

namespace {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix of the calling thread and the block total in *total.
template <class T>
__device__ T block_exclusive_scan(T v, T *total) {
  __shared__ T wsum[SCAN_BLOCK / WAVE];
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  T incl = v;
  for (int off = 1; off < WAVE; off <<= 1) {
    T o = __shfl_up(incl, off, WAVE);
    if (lane >= off) incl += o;
  }
  if (lane == WAVE - 1) wsum[w] = incl;
  __syncthreads();
  T woff = 0, tot = 0;
  for (int i = 0; i < SCAN_BLOCK / WAVE; i++) {
    if (i < w) woff += wsum[i];
    tot += wsum[i];
  }
  __syncthreads();
  *total = tot;
  return woff + incl - v;
}

template <class T>
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_reduce(const T *in, int64_t n, T *bsum) {
  int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  T s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; i++)
    if (base + i < n) s += in[base + i];
  T tot;
  block_exclusive_scan<T>(s, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

template <class T>
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_apply(const T *in, T *out, int64_t n,
                                                          const T *boff, int64_t *total) {
  int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  T v[SCAN_ITEMS];
  T s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; i++) {
    v[i] = (base + i < n) ? in[base + i] : (T)0;
    s += v[i];
  }
  T tot;
  T pre = block_exclusive_scan<T>(s, &tot) + (boff ? boff[blockIdx.x] : (T)0);
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; i++) {
    if (base + i < n) out[base + i] = pre;
    if (total && base + i == n - 1) *total = (int64_t)pre + (int64_t)v[i];  // inclusive total
    pre += v[i];
  }
}

// Stream-ordered scan; the element that holds in[n-1] also writes the
// inclusive total to *total when total is not null.
template <class T>
void exclusive_scan_async(const T *in, T *out, int64_t n, int64_t *total, hipStream_t s) {
  if (n <= 0) return;
  const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb == 1) {
    hipLaunchKernelGGL(k_scan_apply<T>, dim3(1), dim3(SCAN_BLOCK), 0, s, in, out, n, (const T *)nullptr,
                       total);
    LG_LAUNCH_CHECK();
    return;
  }
  DevBuf<T> bsum(nb);
  hipLaunchKernelGGL(k_scan_reduce<T>, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, s, in, n, bsum.p);
  LG_LAUNCH_CHECK();
  exclusive_scan_async<T>(bsum.p, bsum.p, nb, nullptr, s);
  hipLaunchKernelGGL(k_scan_apply<T>, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, s, in, out, n,
                     (const T *)bsum.p, total);
  LG_LAUNCH_CHECK();
}

template <class T>
int64_t exclusive_scan_impl(const T *in, T *out, int64_t n, hipStream_t s) {
  if (n <= 0) return 0;
  DevBuf<int64_t> tot(1);
  exclusive_scan_async<T>(in, out, n, tot.p, s);
  int64_t h = 0;
  to_host(&h, tot.p, 1, s);
  LG_HIP(hipStreamSynchronize(s));
  return h;
}

// ------------------------------------------------------ segmented canon sum
__global__ void k_seg_heads(const uint64_t *keys, int64_t n, int32_t *flags) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// 64 segments per wave: a segment of at most 64 terms is summed left to
// right by its own lane; longer ones by the whole wave, one at a time, in
// 4096-term chunks (each with 64 strided partials and the fixed tree, a
// final chunk of <= 64 terms left to right) whose results are added left to
// right (the oracle's canon_sum).
__global__ void __launch_bounds__(256) k_seg_sum(const uint64_t *keys, const uint64_t *vals,
                                                 const int32_t *start, int64_t nseg, int64_t n,
                                                 uint64_t *out_keys, double *out_sums) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t wid = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE;
  const int64_t seg = wid * WAVE + lane;
  const bool valid = seg < nseg;
  int64_t a = 0, b = 0;
  if (valid) {
    a = start[seg];
    b = (seg + 1 < nseg) ? start[seg + 1] : n;
  }
  double res = 0.0;
  bool longseg = valid && (b - a > 64);
  if (valid && !longseg) {
    double s = 0.0;
    for (int64_t i = a; i < b; i++) s += __longlong_as_double((long long)vals[i]);
    res = s;
  }
  uint64_t mask = __ballot(longseg);
  while (mask) {
    int L = __ffsll((long long)mask) - 1;
    mask &= mask - 1;
    int64_t la = __shfl(a, L, WAVE), lb = __shfl(b, L, WAVE);
    double tot = 0.0;
    for (int64_t ca = la; ca < lb; ca += 4096) {
      const int64_t cb = min(ca + (int64_t)4096, lb);
      double p = 0.0;
      if (cb - ca <= WAVE) {
        for (int64_t i = ca; i < cb; i++) p += __longlong_as_double((long long)vals[i]);
      } else {
        for (int64_t i = ca + lane; i < cb; i += WAVE) p += __longlong_as_double((long long)vals[i]);
        p = wave_tree_sum(p);
        p = __shfl(p, 0, WAVE);
      }
      tot += p;
    }
    if (lane == L) res = tot;
  }
  if (valid) {
    out_keys[seg] = keys[a];
    out_sums[seg] = res;
  }
}

__global__ void k_compact_write(const int32_t *flags, const int32_t *pos, int64_t n, int32_t *out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    if (flags[i]) out[pos[i]] = (int32_t)i;
}

template <class T>
__global__ void k_fill(T *p, T v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace

int64_t exclusive_scan_i32(const int32_t *in, int32_t *out, int64_t n, hipStream_t s) {
  return exclusive_scan_impl<int32_t>(in, out, n, s);
}
int64_t exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t s) {
  return exclusive_scan_impl<int64_t>(in, out, n, s);
}

// rocPRIM's onesweep LSD radix sort.  A stable sort has one result, so any
// correct stable sort gives the same order; this one reads the keys for all
// digit histograms once and then moves every pair once per digit (no
// histogram pass per digit): 260 M pairs of 42-bit keys sort in 14.2 ms,
// against 24.3 ms for a reduce-then-scan sort with per-pass histograms.
void radix_sort_pairs(DevBuf<uint64_t> &keys, DevBuf<uint64_t> &vals, int64_t n, int bits, hipStream_t s) {
  if (n <= 1 || bits <= 0) return;
  LG_CHECK(n < ((int64_t)1 << 31), 1, "radix_sort_pairs: too many items");
  LG_CHECK((size_t)n <= keys.n && (size_t)n <= vals.n, 1, "radix_sort_pairs: buffers shorter than n");
  DevBuf<uint64_t> k2(n), v2(n);
  rocprim::double_buffer<uint64_t> kb(keys.p, k2.p), vb(vals.p, v2.p);
  size_t tb = 0;
  LG_HIP(rocprim::radix_sort_pairs(nullptr, tb, kb, vb, (size_t)n, 0, bits, s));
  DevBuf<uint8_t> tmp(std::max<size_t>(tb, 1));
  LG_HIP(rocprim::radix_sort_pairs((void *)tmp.p, tb, kb, vb, (size_t)n, 0, bits, s));
  // the sorted pairs end in one of the two buffers: hand the caller that one
  if (kb.current() != keys.p) std::swap(keys, k2);
  if (vb.current() != vals.p) std::swap(vals, v2);
  // no synchronisation: the temporaries go back to the device pool, which
  // reuses them only after the device has finished this work
}

int64_t segmented_canon_sum(const uint64_t *keys, const uint64_t *vals, int64_t n,
                            uint64_t *out_keys, double *out_sums, hipStream_t s) {
  if (n <= 0) return 0;
  LG_CHECK(n < ((int64_t)1 << 31), 1, "segmented_canon_sum: too many items");
  DevBuf<int32_t> flags(n), pos(n);
  hipLaunchKernelGGL(k_seg_heads, dim3(grid_for(n, 256)), dim3(256), 0, s, keys, n, flags.p);
  LG_LAUNCH_CHECK();
  int64_t nseg = exclusive_scan_i32(flags.p, pos.p, n, s);
  DevBuf<int32_t> start(nseg);
  hipLaunchKernelGGL(k_compact_write, dim3(grid_for(n, 256)), dim3(256), 0, s, flags.p, pos.p, n,
                     start.p);
  LG_LAUNCH_CHECK();
  int64_t waves = (nseg + WAVE - 1) / WAVE;
  hipLaunchKernelGGL(k_seg_sum, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, keys, vals,
                     start.p, nseg, n, out_keys, out_sums);
  LG_LAUNCH_CHECK();
  return nseg;
}

namespace {
// One lane per segment of at most SEQ_LANE items, summed left to right; a
// longer segment is appended to the long list (in any order: each is summed
// on its own by k_seg_seq_long).
__global__ void __launch_bounds__(256) k_seg_seq(const uint64_t *keys, const uint64_t *vals,
                                                 const int32_t *start, int64_t nseg, int64_t n,
                                                 uint64_t *out_keys, double *out_sums,
                                                 int32_t *nlong, int32_t *longs) {
  const int64_t seg = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (seg >= nseg) return;
  const int64_t a = start[seg], b = (seg + 1 < nseg) ? start[seg + 1] : n;
  out_keys[seg] = keys[a];
  if (b - a > SEQ_LANE) {
    longs[atomicAdd(nlong, 1)] = (int32_t)seg;
    return;
  }
  double acc = 0.0;
#pragma unroll 8
  for (int64_t i = a; i < b; i++) acc += __longlong_as_double((long long)vals[i]);
  out_sums[seg] = acc;
}

// One wavefront per long segment; the grid covers the most long segments
// there can be, and waves past the count exit.
__global__ void __launch_bounds__(256) k_seg_seq_long(const uint64_t *vals, const int32_t *start,
                                                      int64_t nseg, int64_t n, const int32_t *nlong,
                                                      const int32_t *longs, double *out_sums) {
  __shared__ double buf[4 * SEQ_BLOCK];
  const int64_t j = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE;
  if (j >= *nlong) return;
  const int64_t seg = longs[j];
  const int64_t a = start[seg], b = (seg + 1 < nseg) ? start[seg + 1] : n;
  const double acc = wave_seq_fold(a, b, buf + (threadIdx.x / WAVE) * SEQ_BLOCK,
                                   [&](int64_t i) { return __longlong_as_double((long long)vals[i]); });
  if ((threadIdx.x & (WAVE - 1)) == 0) out_sums[seg] = acc;
}
}  // namespace

int64_t segmented_seq_sum(const uint64_t *keys, const uint64_t *vals, int64_t n,
                          uint64_t *out_keys, double *out_sums, hipStream_t s) {
  if (n <= 0) return 0;
  LG_CHECK(n < ((int64_t)1 << 31), 1, "segmented_seq_sum: too many items");
  DevBuf<int32_t> flags(n), pos(n);
  hipLaunchKernelGGL(k_seg_heads, dim3(grid_for(n, 256)), dim3(256), 0, s, keys, n, flags.p);
  LG_LAUNCH_CHECK();
  int64_t nseg = exclusive_scan_i32(flags.p, pos.p, n, s);
  DevBuf<int32_t> start(nseg);
  hipLaunchKernelGGL(k_compact_write, dim3(grid_for(n, 256)), dim3(256), 0, s, flags.p, pos.p, n,
                     start.p);
  LG_LAUNCH_CHECK();
  const int64_t maxlong = n / (SEQ_LANE + 1);
  DevBuf<int32_t> nlong(1), longs(std::max<int64_t>(maxlong, 1));
  fill_i32(nlong.p, 0, 1, s);
  hipLaunchKernelGGL(k_seg_seq, dim3((unsigned)((nseg + 255) / 256)), dim3(256), 0, s, keys, vals,
                     start.p, nseg, n, out_keys, out_sums, nlong.p, longs.p);
  LG_LAUNCH_CHECK();
  if (maxlong > 0) {
    hipLaunchKernelGGL(k_seg_seq_long, dim3((unsigned)((maxlong + 3) / 4)), dim3(256), 0, s, vals,
                       start.p, nseg, n, nlong.p, longs.p, out_sums);
    LG_LAUNCH_CHECK();
  }
  return nseg;
}

int64_t compact_flags(const int32_t *flags, int64_t n, int32_t *out, hipStream_t s) {
  if (n <= 0) return 0;
  DevBuf<int32_t> pos(n);
  int64_t cnt = exclusive_scan_i32(flags, pos.p, n, s);
  hipLaunchKernelGGL(k_compact_write, dim3(grid_for(n, 256)), dim3(256), 0, s, flags, pos.p, n, out);
  LG_LAUNCH_CHECK();
  return cnt;
}

void fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<int32_t>, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, p, v, n);
  LG_LAUNCH_CHECK();
}
void fill_f64(double *p, double v, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, p, v, n);
  LG_LAUNCH_CHECK();
}
void fill_u8(uint8_t *p, uint8_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<uint8_t>, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, p, v, n);
  LG_LAUNCH_CHECK();
}

}  // namespace lg
