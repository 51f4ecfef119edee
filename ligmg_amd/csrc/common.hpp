// Shared HIP helpers: error handling, device buffers, wavefront utilities.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace lg {

// Error carrying a C-ABI status code (include/ligmg_amd.h).
// message of the last failed C-ABI call on this thread
extern thread_local std::string g_last_error;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define LG_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw ::lg::Error(3, std::string(#call) + ": " + hipGetErrorString(e_) + " at " + \
                               __FILE__ + ":" + std::to_string(__LINE__));             \
  } while (0)

#define LG_CHECK(cond, code, msg)             \
  do {                                                \
    if (!(cond)) throw ::lg::Error((code), (msg)); \
  } while (0)

#define LG_LAUNCH_CHECK() LG_HIP(hipGetLastError())

constexpr int WAVE = 64;

// Pooled device memory (pool.hip): blocks come back from dev_free as
// pending and are reused after the queued work that may read them is done;
// dev_pool_trim returns every cached block to HIP.
void *dev_alloc(size_t bytes);
void dev_free(void *p);
void dev_pool_trim();
// bracket a graph capture: the pool then never synchronises the device
void pool_capture_begin();
void pool_capture_end();
struct PoolCaptureScope {
  PoolCaptureScope() { pool_capture_begin(); }
  ~PoolCaptureScope() { pool_capture_end(); }
  PoolCaptureScope(const PoolCaptureScope &) = delete;
  PoolCaptureScope &operator=(const PoolCaptureScope &) = delete;
};

// Owning device allocation (pooled).
template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) {
      try {
        p = (T *)dev_alloc(count * sizeof(T));
      } catch (...) {
        p = nullptr;
        n = 0;
        throw;
      }
    }
  }
  void ensure(size_t count) {
    if (count > n) alloc(count);
  }
  void release() {
    if (p) dev_free(p);
    p = nullptr;
    n = 0;
  }
  T *get() const { return p; }
  size_t bytes() const { return n * sizeof(T); }
};

template <class T>
inline void to_device(T *dst, const T *src, size_t count, hipStream_t s) {
  if (count) LG_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, s));
}
template <class T>
inline void to_host(T *dst, const T *src, size_t count, hipStream_t s) {
  if (count) LG_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost, s));
}
template <class T>
inline void d2d(T *dst, const T *src, size_t count, hipStream_t s) {
  if (count) LG_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToDevice, s));
}
template <class T>
inline std::vector<T> download(const T *src, size_t count, hipStream_t s) {
  std::vector<T> v(count);
  to_host(v.data(), src, count, s);
  LG_HIP(hipStreamSynchronize(s));
  return v;
}

inline unsigned grid_for(int64_t work, int block, int64_t cap = 1 << 20) {
  int64_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

inline int bits_for(uint64_t v) {  // number of bits to hold values < v (v >= 1)
  int b = 0;
  while (b < 64 && ((uint64_t)1 << b) < v) b++;
  return b;
}

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ double shfl_down_d(double v, int off) { return __shfl_down(v, off, WAVE); }
__device__ __forceinline__ float shfl_down_f(float v, int off) { return __shfl_down(v, off, WAVE); }

// The canonical tree used after 64 strided partial sums: p[l] += p[l+off]
// for off = 32..1.  Lane 0 holds the result.
__device__ __forceinline__ double wave_tree_sum(double v) {
  for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_down(v, off, WAVE);
  return v;
}
__device__ __forceinline__ float wave_tree_sum(float v) {
  for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_down(v, off, WAVE);
  return v;
}
__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }

// Orders the LDS accesses of one wavefront around this point: LDS executes a
// wavefront's instructions in order, so after the fence every lane sees
// what the other lanes stored before it.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Rows (segments) longer than this are folded by a whole wavefront
// (wave_seq_fold); shorter ones by one lane each.
constexpr int SEQ_LANE = 256;

// Left fold from 0.0 of term(k), k = a, a+1, ..., b-1, by one wavefront:
// the exact sequential sum, computed by every lane.  The terms are read in
// blocks of SEQ_BLOCK = 512 (8 per lane, coalesced), and the next block's
// reads are in flight while the current one, staged in buf (this
// wavefront's SEQ_BLOCK-entry LDS slot), is added in index order.  A block
// that deep hides the latency of gathered terms: one 64-term chunk at a
// time left a 257 000-term hub row waiting ~4000 times for a gather.
constexpr int SEQ_CHUNKS = 8;
constexpr int SEQ_BLOCK = SEQ_CHUNKS * WAVE;
template <class Term>
__device__ __forceinline__ double wave_seq_fold(int64_t a, int64_t b, double *buf, const Term &term) {
  const int lane = threadIdx.x & (WAVE - 1);
  double nx[SEQ_CHUNKS];
  auto fetch = [&](int64_t base) {
#pragma unroll
    for (int c = 0; c < SEQ_CHUNKS; c++) {
      const int64_t k = base + c * WAVE + lane;
      nx[c] = (k < b) ? term(k) : 0.0;
    }
  };
  fetch(a);
  double acc = 0.0;
  for (int64_t base = a; base < b; base += SEQ_BLOCK) {
#pragma unroll
    for (int c = 0; c < SEQ_CHUNKS; c++) buf[c * WAVE + lane] = nx[c];
    wave_lds_sync();
    if (base + SEQ_BLOCK < b) fetch(base + SEQ_BLOCK);
    if (b - base >= SEQ_BLOCK) {
#pragma unroll 64
      for (int t = 0; t < SEQ_BLOCK; t++) acc += buf[t];
    } else {
      const int cnt = (int)(b - base);
      for (int t = 0; t < cnt; t++) acc += buf[t];
    }
    wave_lds_sync();
  }
  return acc;
}
#endif

}  // namespace lg
