// Graph preprocessing on the device: input validation, connected
// components, largest-component extraction and adjacency -> Laplacian.
//
// The reference does these steps in its driver on CombBLAS matrices:
// connected_components (cc.hpp:22-51) iterates a max-label SpMV to a fixed
// point, largest_cc (cc.hpp:55-105) keeps the biggest component renumbered
// in ascending order, and main.cpp:79-110 forms L = D - A.  Here components
// come from hook-and-shortcut passes (every root is hooked below its
// smallest neighbouring root, then every vertex jumps to its root), which
// need O(log n) full passes instead of O(diameter); the reference's labels
// (largest index of the component) are then read off the final forest, so
// the result is the same.  Every kernel is a bounded data-parallel pass: the
// host repeats them until nothing changes.
#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/ligmg_amd.h"
#include "prims.hpp"
#include "setup.hpp"

namespace lg {

namespace {

// pass 1 reads only the row pointers: rp[0] = 0, non-decreasing, rp[n] = m
__global__ void k_rowptr_check(int64_t n, int64_t m, const int32_t *rp, int *bad) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t a = rp[i];
    if ((i == 0 && a != 0) || (i == n && a != m) || (i < n && rp[i + 1] < a)) atomicOr(bad, 1);
  }
}

// pass 2, once the row pointers are known good: columns in [0, ncols)
__global__ void k_colind_check(int64_t m, int64_t ncols, const int32_t *ci, int *bad) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < m;
       k += (int64_t)gridDim.x * blockDim.x)
    if (ci[k] < 0 || ci[k] >= ncols) atomicOr(bad, 2);
}

__global__ void k_cc_init(int64_t n, int32_t *parent) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    parent[i] = (int32_t)i;
}

// every parent is a vertex of the same component with a smaller or equal
// index, so hooking a root below a smaller root (atomicMin) never closes a
// cycle, whatever order the edges are seen in
__global__ void k_cc_hook(int64_t n, const int32_t *rp, const int32_t *ci, int32_t *parent, int *changed) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t nw = (int64_t)gridDim.x * blockDim.x / WAVE;
  for (int64_t u = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE; u < n; u += nw) {
    const int32_t pu = parent[u];
    bool ch = false;
    for (int32_t k = rp[u] + lane; k < rp[u + 1]; k += WAVE) {
      const int32_t pv = parent[ci[k]];
      if (pv < pu) {
        if (atomicMin(&parent[pu], pv) > pv) ch = true;
      } else if (pu < pv) {
        if (atomicMin(&parent[pv], pu) > pu) ch = true;
      }
    }
    if (__any(ch) && lane == 0) atomicOr(changed, 1);
  }
}

// shortcut: every vertex points at its root (parents strictly decrease
// along a path, so the walk ends)
__global__ void k_cc_jump(int64_t n, int32_t *parent) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int32_t p = parent[i];
    while (true) {
      const int32_t q = parent[p];
      if (q >= p) break;
      p = q;
    }
    parent[i] = p;
  }
}

// largest vertex index and size of every component (indexed by its root)
__global__ void k_cc_stats(int64_t n, const int32_t *parent, int32_t *maxv, int32_t *count) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t r = parent[i];
    atomicMax(&maxv[r], (int32_t)i);
    atomicAdd(&count[r], 1);
  }
}

__global__ void k_cc_label(int64_t n, const int32_t *parent, const int32_t *maxv, int32_t *label) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    label[i] = maxv[parent[i]];
}

__global__ void k_keep_flags(int64_t n, const int32_t *parent, int32_t root, int32_t *flag) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= n;
       i += (int64_t)gridDim.x * blockDim.x)
    flag[i] = (i < n && parent[i] == root) ? 1 : 0;
}

__global__ void k_keep_ids(int64_t n, const int32_t *flag, const int32_t *scan, int32_t *new_id) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    new_id[i] = flag[i] ? scan[i] : -1;
}

// kept entries per kept row, in the order of the rows' new ids
__global__ void k_sub_count(int64_t n, const int32_t *rp, const int32_t *ci, const int32_t *new_id,
                            int32_t *cnt) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t nw = (int64_t)gridDim.x * blockDim.x / WAVE;
  for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE; r < n; r += nw) {
    const int32_t nr = new_id[r];
    if (nr < 0) continue;
    int c = 0;
    for (int32_t k = rp[r] + lane; k < rp[r + 1]; k += WAVE) c += new_id[ci[k]] >= 0;
    for (int off = WAVE / 2; off >= 1; off >>= 1) c += __shfl_xor(c, off, WAVE);
    if (lane == 0) cnt[nr] = c;
  }
}

__global__ void k_sub_fill(int64_t n, const int32_t *rp, const int32_t *ci, const double *v,
                           const int32_t *new_id, const int32_t *orp, int32_t *oci, double *ov) {
  const int lane = threadIdx.x & (WAVE - 1);
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t nw = (int64_t)gridDim.x * blockDim.x / WAVE;
  for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE; r < n; r += nw) {
    const int32_t nr = new_id[r];
    if (nr < 0) continue;
    int32_t o = orp[nr];
    for (int32_t base = rp[r]; base < rp[r + 1]; base += WAVE) {
      const int32_t k = base + lane;
      const int32_t c = (k < rp[r + 1]) ? new_id[ci[k]] : -1;
      const uint64_t m = __ballot(c >= 0);
      if (c >= 0) {
        const int32_t pos = o + __popcll(m & lt);
        oci[pos] = c;
        if (ov) ov[pos] = v[k];
      }
      o += __popcll(m);
    }
  }
}

// Laplacian row lengths: A's row plus a diagonal slot when A has none
__global__ void k_lap_len(int64_t n, const int32_t *rp, const int32_t *ci, int32_t *len) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t nw = (int64_t)gridDim.x * blockDim.x / WAVE;
  for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE; r <= n; r += nw) {
    if (r == n) {
      if (lane == 0) len[n] = 0;
      continue;
    }
    bool has = false;
    for (int32_t k = rp[r] + lane; k < rp[r + 1]; k += WAVE) has |= ci[k] == r;
    const bool any = __any(has);  // whole wavefront, before any lane-dependent branch
    if (lane == 0) len[r] = (rp[r + 1] - rp[r]) + (any ? 0 : 1);
  }
}

// L_ij = -A_ij (j != i), L_ii = deg_i - A_ii, the diagonal inserted in
// column order when A has none
__global__ void k_lap_fill(int64_t n, const int32_t *rp, const int32_t *ci, const double *v, const double *deg,
                           const int32_t *orp, int32_t *oci, double *ov) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    int32_t o = orp[r];
    bool done = false;
    for (int32_t k = rp[r]; k < rp[r + 1]; k++) {
      const int32_t c = ci[k];
      if (!done && c > r) {
        oci[o] = (int32_t)r;
        ov[o++] = deg[r];
        done = true;
      }
      if (c == r) {
        oci[o] = c;
        ov[o++] = deg[r] - v[k];
        done = true;
      } else {
        oci[o] = c;
        ov[o++] = 0.0 - v[k];
      }
    }
    if (!done) {
      oci[o] = (int32_t)r;
      ov[o] = deg[r];
    }
  }
}

unsigned gs(int64_t work) { return grid_for(work, 256, 16384); }

// device copies of caller arrays (or views of device ones)
template <class T>
struct In {
  DevBuf<T> own;
  const T *p = nullptr;
  In(const T *src, int64_t cnt, int mem, hipStream_t s) {
    if (mem == LIGMG_MEM_DEVICE || !src) {
      p = src;
    } else {
      own.alloc(std::max<int64_t>(cnt, 1));
      to_device(own.p, src, cnt, s);
      p = own.p;
    }
  }
};

template <class T>
void out_copy(T *dst, const T *src, int64_t cnt, int mem, hipStream_t s) {
  if (cnt <= 0) return;
  if (mem == LIGMG_MEM_DEVICE) d2d(dst, src, cnt, s);
  else to_host(dst, src, cnt, s);
}

int64_t nnz_of(const int32_t *rp, int64_t n, int mem, hipStream_t s) {
  int32_t e[2];
  if (mem == LIGMG_MEM_DEVICE) {
    to_host(&e[0], rp, 1, s);
    to_host(&e[1], rp + n, 1, s);
    LG_HIP(hipStreamSynchronize(s));
  } else {
    e[0] = rp[0];
    e[1] = rp[n];
  }
  LG_CHECK(e[0] == 0 && e[1] >= 0, LIGMG_ERR_INVALID, "rowptr[0] must be 0 and rowptr[n] >= 0");
  return e[1];
}

// roots (count > 0) of the forest: the number of components
__global__ void k_cc_count_roots(int64_t n, const int32_t *count, int *ncomp) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    if (count[i] > 0) atomicAdd(ncomp, 1);
}

// hooks + shortcuts to a forest of stars whose roots are the smallest index
// of every component; changed is one word of device scratch
void cc_forest_into(int64_t n, const int32_t *rp, const int32_t *ci, int32_t *parent, int *changed,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_cc_init, dim3(gs(n)), dim3(256), 0, s, n, parent);
  LG_LAUNCH_CHECK();
  for (int round = 0;; round++) {
    LG_CHECK(round < 256, LIGMG_ERR_NUMERIC, "connected components did not converge");
    fill_i32(changed, 0, 1, s);
    hipLaunchKernelGGL(k_cc_hook, dim3(gs(n * WAVE)), dim3(256), 0, s, n, rp, ci, parent, changed);
    LG_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_cc_jump, dim3(gs(n)), dim3(256), 0, s, n, parent);
    LG_LAUNCH_CHECK();
    int h = 0;
    to_host(&h, changed, 1, s);
    LG_HIP(hipStreamSynchronize(s));
    if (!h) break;
  }
}

void cc_forest(int64_t n, const int32_t *rp, const int32_t *ci, DevBuf<int32_t> &parent, hipStream_t s) {
  parent.alloc(n);
  DevBuf<int> changed(1);
  cc_forest_into(n, rp, ci, parent.p, changed.p, s);
}

// connected_components on a validated device CSR, into caller buffers:
// parent, maxv and count hold n words each, flags 2 (the forest's change
// flag, the component count); label (device, n) receives the labels.
// Returns the number of components.  Every buffer is written before it is
// read, so scratch left over from an earlier call does not matter.
int64_t cc_labels_into(int64_t n, const int32_t *rp, const int32_t *ci, int32_t *parent, int32_t *maxv,
                       int32_t *count, int *flags, int32_t *label, hipStream_t s) {
  cc_forest_into(n, rp, ci, parent, flags, s);
  fill_i32(maxv, -1, n, s);
  fill_i32(count, 0, n, s);
  fill_i32(flags + 1, 0, 1, s);
  hipLaunchKernelGGL(k_cc_stats, dim3(gs(n)), dim3(256), 0, s, n, (const int32_t *)parent, maxv, count);
  LG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_cc_label, dim3(gs(n)), dim3(256), 0, s, n, (const int32_t *)parent, (const int32_t *)maxv,
                     label);
  LG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_cc_count_roots, dim3(gs(n)), dim3(256), 0, s, n, (const int32_t *)count, flags + 1);
  LG_LAUNCH_CHECK();
  int k = 0;
  to_host(&k, flags + 1, 1, s);
  LG_HIP(hipStreamSynchronize(s));
  return k;
}

}  // namespace

// pass 3 (optional): strictly increasing columns within every row
__global__ void k_sorted_check(int64_t n, const int32_t *rp, const int32_t *ci, int *bad) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t nw = (int64_t)gridDim.x * blockDim.x / WAVE;
  for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE; r < n; r += nw) {
    bool ok = true;
    for (int32_t k = rp[r] + 1 + lane; k < rp[r + 1]; k += WAVE) ok &= ci[k - 1] < ci[k];
    if (!__all(ok) && lane == 0) atomicOr(bad, 4);
  }
}

// bad: one word of device scratch
void validate_csr_into(int64_t n, int64_t ncols, const int32_t *rp, const int32_t *ci, hipStream_t s, bool sorted,
                       int *bad) {
  int32_t m = 0;
  to_host(&m, rp + n, 1, s);
  LG_HIP(hipStreamSynchronize(s));
  LG_CHECK(m >= 0, LIGMG_ERR_INVALID, "rowptr[n] must be non-negative");
  fill_i32(bad, 0, 1, s);
  hipLaunchKernelGGL(k_rowptr_check, dim3(gs(n + 1)), dim3(256), 0, s, n, (int64_t)m, rp, bad);
  LG_LAUNCH_CHECK();
  int h = 0;
  to_host(&h, bad, 1, s);
  LG_HIP(hipStreamSynchronize(s));
  LG_CHECK(!(h & 1), LIGMG_ERR_INVALID, "rowptr must start at 0 and be non-decreasing");
  hipLaunchKernelGGL(k_colind_check, dim3(gs(m)), dim3(256), 0, s, (int64_t)m, ncols, ci, bad);
  LG_LAUNCH_CHECK();
  to_host(&h, bad, 1, s);
  LG_HIP(hipStreamSynchronize(s));
  LG_CHECK(!(h & 2), LIGMG_ERR_INVALID, "column index out of range");
  if (sorted) {
    hipLaunchKernelGGL(k_sorted_check, dim3(gs(n * WAVE)), dim3(256), 0, s, n, rp, ci, bad);
    LG_LAUNCH_CHECK();
    to_host(&h, bad, 1, s);
    LG_HIP(hipStreamSynchronize(s));
    LG_CHECK(!(h & 4), LIGMG_ERR_INVALID, "columns must be sorted and unique within every row");
  }
}

void validate_csr(int64_t n, int64_t ncols, const int32_t *rp, const int32_t *ci, hipStream_t s, bool sorted) {
  DevBuf<int> bad(1);
  validate_csr_into(n, ncols, rp, ci, s, sorted, bad.p);
}

}  // namespace lg

// Scratch of ligmg_components_noalloc: parent, per-root maximum and per-root
// count (LIGMG_NOALLOC_CC_MAX_N words each), then flag words.  Each GPU the
// library runs on holds its own copy until the process ends.  Host code
// never uses this name as an address: scratch_on_device() looks the
// device's copy up once per device.
constexpr int64_t NOALLOC_CC_N = LIGMG_NOALLOC_CC_MAX_N;
constexpr int64_t NOALLOC_CC_WORDS = 3 * NOALLOC_CC_N + 64;
static __device__ __align__(256) int32_t scratch[NOALLOC_CC_WORDS];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
int32_t *scratch_addr[MAX_DEVICES] = {};

// this device's copy of scratch, looked up on the first call on the device
int32_t *scratch_on_device() {
  int dev = 0;
  LG_HIP(hipGetDevice(&dev));
  LG_CHECK(dev >= 0 && dev < MAX_DEVICES, LIGMG_ERR_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_addr[dev]) {
    void *p = nullptr;
    LG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    scratch_addr[dev] = (int32_t *)p;
  }
  return scratch_addr[dev];
}

thread_local hipStream_t prep_stream = nullptr;
hipStream_t prep_s() {
  if (!prep_stream) LG_HIP(hipStreamCreateWithFlags(&prep_stream, hipStreamNonBlocking));
  return prep_stream;
}
int prep_fail(const lg::Error &e) {
  lg::g_last_error = e.what();
  return e.code;
}
}  // namespace

extern "C" {

int ligmg_components(int64_t n, const int32_t *rowptr, const int32_t *colind, int mem_kind, int32_t *label,
                     int64_t *ncomp) {
  if (n <= 0 || n >= INT32_MAX || !rowptr || !colind || !label) return LIGMG_ERR_INVALID;
  try {
    hipStream_t s = prep_s();
    const int64_t m = lg::nnz_of(rowptr, n, mem_kind, s);
    lg::In<int32_t> rp(rowptr, n + 1, mem_kind, s), ci(colind, m, mem_kind, s);
    lg::validate_csr(n, n, rp.p, ci.p, s, false);
    lg::DevBuf<int32_t> parent(n), maxv(n), count(n), lab(n);
    lg::DevBuf<int> flags(2);
    const int64_t k = lg::cc_labels_into(n, rp.p, ci.p, parent.p, maxv.p, count.p, flags.p, lab.p, s);
    lg::out_copy(label, (const int32_t *)lab.p, n, mem_kind, s);
    if (ncomp) *ncomp = k;
    LG_HIP(hipStreamSynchronize(s));
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return prep_fail(e);
  }
}

int ligmg_components_noalloc(int64_t n, const int32_t *rowptr, const int32_t *colind, int32_t *label,
                             int64_t *ncomp) {
  if (n <= 0 || !rowptr || !colind || !label) return LIGMG_ERR_INVALID;
  if (n > NOALLOC_CC_N) {
    lg::g_last_error = "ligmg_components_noalloc: n = " + std::to_string(n) + " is above the maximum of " +
                       std::to_string(NOALLOC_CC_N) + " vertices (use ligmg_components)";
    return LIGMG_ERR_INVALID;
  }
  try {
    // the calling thread's default stream of the current device: no stream
    // is created either
    hipStream_t s = hipStreamPerThread;
    int32_t *sc = scratch_on_device();
    int32_t *parent = sc, *maxv = sc + NOALLOC_CC_N, *count = sc + 2 * NOALLOC_CC_N;
    int *flags = (int *)(sc + 3 * NOALLOC_CC_N);
    lg::nnz_of(rowptr, n, LIGMG_MEM_DEVICE, s);
    lg::validate_csr_into(n, n, rowptr, colind, s, false, flags + 2);
    const int64_t k = lg::cc_labels_into(n, rowptr, colind, parent, maxv, count, flags, label, s);
    if (ncomp) *ncomp = k;
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return prep_fail(e);
  }
}

int ligmg_largest_component(int64_t n, const int32_t *rowptr, const int32_t *colind, int mem_kind,
                            int32_t *new_id, int64_t *n_out, int64_t *nnz_out) {
  if (n <= 0 || n >= INT32_MAX || !rowptr || !colind || !new_id || !n_out || !nnz_out) return LIGMG_ERR_INVALID;
  try {
    hipStream_t s = prep_s();
    const int64_t m = lg::nnz_of(rowptr, n, mem_kind, s);
    lg::In<int32_t> rp(rowptr, n + 1, mem_kind, s), ci(colind, m, mem_kind, s);
    lg::validate_csr(n, n, rp.p, ci.p, s, false);
    lg::DevBuf<int32_t> parent, maxv(n), count(n);
    lg::cc_forest(n, rp.p, ci.p, parent, s);
    lg::fill_i32(maxv.p, -1, n, s);
    lg::fill_i32(count.p, 0, n, s);
    hipLaunchKernelGGL(lg::k_cc_stats, dim3(lg::gs(n)), dim3(256), 0, s, n, (const int32_t *)parent.p, maxv.p,
                       count.p);
    LG_LAUNCH_CHECK();
    // the largest count; among equals the smallest label (cc.hpp:71-72:
    // max_element over counts indexed by label returns the first maximum)
    auto c = lg::download(count.p, n, s);
    auto mx = lg::download(maxv.p, n, s);
    int32_t best = -1;
    for (int64_t r = 0; r < n; r++) {
      if (c[r] <= 0) continue;
      if (best < 0 || c[r] > c[best] || (c[r] == c[best] && mx[r] < mx[best])) best = (int32_t)r;
    }
    lg::DevBuf<int32_t> flag(n + 1), scan(n + 1), ids(n), cnt(n + 1);
    hipLaunchKernelGGL(lg::k_keep_flags, dim3(lg::gs(n + 1)), dim3(256), 0, s, n, (const int32_t *)parent.p, best,
                       flag.p);
    LG_LAUNCH_CHECK();
    const int64_t nk = lg::exclusive_scan_i32(flag.p, scan.p, n + 1, s);
    hipLaunchKernelGGL(lg::k_keep_ids, dim3(lg::gs(n)), dim3(256), 0, s, n, (const int32_t *)flag.p,
                       (const int32_t *)scan.p, ids.p);
    LG_LAUNCH_CHECK();
    lg::fill_i32(cnt.p, 0, nk + 1, s);
    hipLaunchKernelGGL(lg::k_sub_count, dim3(lg::gs(n * lg::WAVE)), dim3(256), 0, s, n, rp.p, ci.p,
                       (const int32_t *)ids.p, cnt.p);
    LG_LAUNCH_CHECK();
    lg::DevBuf<int32_t> cs(nk + 1);
    const int64_t mk = lg::exclusive_scan_i32(cnt.p, cs.p, nk + 1, s);
    lg::out_copy(new_id, (const int32_t *)ids.p, n, mem_kind, s);
    LG_HIP(hipStreamSynchronize(s));
    *n_out = nk;
    *nnz_out = mk;
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return prep_fail(e);
  }
}

int ligmg_extract(int64_t n, const int32_t *rowptr, const int32_t *colind, const double *vals,
                  const int32_t *new_id, int64_t n_out, int mem_kind, int32_t *out_rowptr, int32_t *out_colind,
                  double *out_vals) {
  if (n <= 0 || n >= INT32_MAX || !rowptr || !colind || !new_id || n_out < 0 || !out_rowptr || !out_colind)
    return LIGMG_ERR_INVALID;
  try {
    hipStream_t s = prep_s();
    const int64_t m = lg::nnz_of(rowptr, n, mem_kind, s);
    lg::In<int32_t> rp(rowptr, n + 1, mem_kind, s), ci(colind, m, mem_kind, s), ids(new_id, n, mem_kind, s);
    lg::In<double> v(vals, m, mem_kind, s);
    lg::validate_csr(n, n, rp.p, ci.p, s, false);
    // new ids must be -1 or a rank below n_out, each rank used once
    {
      auto h = mem_kind == LIGMG_MEM_DEVICE ? lg::download(new_id, n, s) : std::vector<int32_t>(new_id, new_id + n);
      std::vector<char> seen(n_out, 0);
      for (int64_t i = 0; i < n; i++) {
        LG_CHECK(h[i] >= -1 && h[i] < n_out, LIGMG_ERR_INVALID, "ligmg_extract: new_id out of range");
        if (h[i] >= 0) {
          LG_CHECK(!seen[h[i]], LIGMG_ERR_INVALID, "ligmg_extract: new_id repeats a rank");
          seen[h[i]] = 1;
        }
      }
    }
    lg::DevBuf<int32_t> cnt(n_out + 1), orp(n_out + 1);
    lg::fill_i32(cnt.p, 0, n_out + 1, s);
    hipLaunchKernelGGL(lg::k_sub_count, dim3(lg::gs(n * lg::WAVE)), dim3(256), 0, s, n, rp.p, ci.p, ids.p, cnt.p);
    LG_LAUNCH_CHECK();
    const int64_t mk = lg::exclusive_scan_i32(cnt.p, orp.p, n_out + 1, s);
    lg::DevBuf<int32_t> oci(std::max<int64_t>(mk, 1));
    lg::DevBuf<double> ov(v.p ? std::max<int64_t>(mk, 1) : 0);
    hipLaunchKernelGGL(lg::k_sub_fill, dim3(lg::gs(n * lg::WAVE)), dim3(256), 0, s, n, rp.p, ci.p, v.p, ids.p,
                       (const int32_t *)orp.p, oci.p, ov.p);
    LG_LAUNCH_CHECK();
    lg::out_copy(out_rowptr, (const int32_t *)orp.p, n_out + 1, mem_kind, s);
    lg::out_copy(out_colind, (const int32_t *)oci.p, mk, mem_kind, s);
    if (out_vals && v.p) lg::out_copy(out_vals, (const double *)ov.p, mk, mem_kind, s);
    LG_HIP(hipStreamSynchronize(s));
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return prep_fail(e);
  }
}

int ligmg_laplacian_nnz(int64_t n, const int32_t *rowptr, const int32_t *colind, int mem_kind, int64_t *nnz_out) {
  if (n <= 0 || n >= INT32_MAX || !rowptr || !colind || !nnz_out) return LIGMG_ERR_INVALID;
  try {
    hipStream_t s = prep_s();
    const int64_t m = lg::nnz_of(rowptr, n, mem_kind, s);
    lg::In<int32_t> rp(rowptr, n + 1, mem_kind, s), ci(colind, m, mem_kind, s);
    lg::validate_csr(n, n, rp.p, ci.p, s, false);
    lg::DevBuf<int32_t> len(n + 1), orp(n + 1);
    hipLaunchKernelGGL(lg::k_lap_len, dim3(lg::gs((n + 1) * lg::WAVE)), dim3(256), 0, s, n, rp.p, ci.p, len.p);
    LG_LAUNCH_CHECK();
    const int64_t tot = lg::exclusive_scan_i32(len.p, orp.p, n + 1, s);
    LG_CHECK(tot < INT32_MAX, LIGMG_ERR_INVALID, "Laplacian too large for int32 indices");
    *nnz_out = tot;
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return prep_fail(e);
  }
}

int ligmg_adjacency_to_laplacian(int64_t n, const int32_t *rowptr, const int32_t *colind, const double *vals,
                                 int mem_kind, int32_t *out_rowptr, int32_t *out_colind, double *out_vals) {
  if (n <= 0 || n >= INT32_MAX || !rowptr || !colind || !vals || !out_rowptr || !out_colind || !out_vals)
    return LIGMG_ERR_INVALID;
  try {
    hipStream_t s = prep_s();
    const int64_t m = lg::nnz_of(rowptr, n, mem_kind, s);
    lg::In<int32_t> rp(rowptr, n + 1, mem_kind, s), ci(colind, m, mem_kind, s);
    lg::In<double> v(vals, m, mem_kind, s);
    lg::validate_csr(n, n, rp.p, ci.p, s, true);  // the diagonal is inserted in column order
    lg::DevBuf<int32_t> len(n + 1), orp(n + 1);
    hipLaunchKernelGGL(lg::k_lap_len, dim3(lg::gs((n + 1) * lg::WAVE)), dim3(256), 0, s, n, rp.p, ci.p, len.p);
    LG_LAUNCH_CHECK();
    const int64_t tot = lg::exclusive_scan_i32(len.p, orp.p, n + 1, s);
    lg::DevBuf<double> deg(n);
    lg::row_sums_seq(n, rp.p, v.p, deg.p, s);
    lg::DevBuf<int32_t> oci(std::max<int64_t>(tot, 1));
    lg::DevBuf<double> ov(std::max<int64_t>(tot, 1));
    hipLaunchKernelGGL(lg::k_lap_fill, dim3(lg::gs(n)), dim3(256), 0, s, n, rp.p, ci.p, v.p,
                       (const double *)deg.p, (const int32_t *)orp.p, oci.p, ov.p);
    LG_LAUNCH_CHECK();
    lg::out_copy(out_rowptr, (const int32_t *)orp.p, n + 1, mem_kind, s);
    lg::out_copy(out_colind, (const int32_t *)oci.p, tot, mem_kind, s);
    lg::out_copy(out_vals, (const double *)ov.p, tot, mem_kind, s);
    LG_HIP(hipStreamSynchronize(s));
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return prep_fail(e);
  }
}

}  // extern "C"
