// Device memory pool behind DevBuf.
//
// Setup allocates and frees thousands of temporaries (sort buffers, maps,
// per-level scratch).  hipFree synchronises and unmaps; once a process has
// released a large solver, each such free costs up to tens of milliseconds,
// which doubled the second setup's time.  Released blocks are kept here by
// size class instead.  A released block may still be read by work already
// queued on some stream, so it is parked as pending and becomes reusable
// only after one device-wide synchronisation - the guarantee hipFree's own
// synchronisation gave.  No allocation happens inside a captured graph, and
// while any handle captures one (on any thread) the pool takes fresh blocks
// instead of synchronising the device, which would end that capture.
#include <atomic>
#include <iterator>
#include <map>
#include <mutex>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/ligmg_amd.h"
#include "common.hpp"

namespace lg {

namespace {

// bytes kept for reuse at most; above it released blocks go back to HIP
constexpr size_t CACHE_MAX = (size_t)32 << 30;

struct Pool {
  std::mutex m;
  using Key = std::pair<int, size_t>;  // (device, class bytes)
  std::map<Key, std::vector<void *>> ready, pending;
  std::unordered_map<void *, Key> live;
  size_t cached = 0;
};

// graph captures of the library in progress (all threads)
std::atomic<int> g_capturing{0};

// never destroyed: blocks live until the process ends, after the HIP
// runtime may already be gone
Pool &pool() {
  static Pool *p = new Pool();
  return *p;
}

// classes 2^k and 3 * 2^(k-2): at most a third of a block unused
size_t size_class(size_t b) {
  size_t p2 = 256;
  while (p2 < b) p2 <<= 1;
  const size_t three_quarters = (p2 >> 1) + (p2 >> 2);
  return (p2 > 256 && b <= three_quarters) ? three_quarters : p2;
}

void flush_locked(Pool &P) {
  (void)hipDeviceSynchronize();
  for (auto *lists : {&P.ready, &P.pending})
    for (auto &kv : *lists)
      for (void *q : kv.second) (void)hipFree(q);
  P.ready.clear();
  P.pending.clear();
  P.cached = 0;
}

}  // namespace

void *dev_alloc(size_t bytes) {
  if (bytes == 0) return nullptr;
  int dev = 0;
  LG_HIP(hipGetDevice(&dev));
  const size_t cls = size_class(bytes);
  const Pool::Key key{dev, cls};
  Pool &P = pool();
  std::lock_guard<std::mutex> g(P.m);
  auto it = P.ready.find(key);
  if (it == P.ready.end() || it->second.empty()) {
    auto pt = P.pending.find(key);
    if (pt != P.pending.end() && !pt->second.empty() && g_capturing.load() == 0) {
      // this device's pending blocks are free once its queued work is done
      LG_HIP(hipDeviceSynchronize());
      for (auto q = P.pending.begin(); q != P.pending.end();) {
        if (q->first.first != dev) {
          ++q;
          continue;
        }
        auto &dst = P.ready[q->first];
        dst.insert(dst.end(), q->second.begin(), q->second.end());
        q = P.pending.erase(q);
      }
      it = P.ready.find(key);
    }
  }
  void *p = nullptr;
  if (it != P.ready.end() && !it->second.empty()) {
    p = it->second.back();
    it->second.pop_back();
    P.cached -= cls;
  } else {
    hipError_t e = hipMalloc(&p, cls);
    if (e != hipSuccess && g_capturing.load() == 0) {  // give the cache back and try once more
      (void)hipGetLastError();
      flush_locked(P);
      e = hipMalloc(&p, cls);
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw Error(LIGMG_ERR_NOMEM, "hipMalloc of " + std::to_string(cls) + " bytes failed");
    }
  }
  P.live[p] = key;
  return p;
}

void dev_free(void *p) {
  if (!p) return;
  Pool &P = pool();
  std::lock_guard<std::mutex> g(P.m);
  auto it = P.live.find(p);
  if (it == P.live.end()) return;
  const Pool::Key key = it->second;
  P.live.erase(it);
  // hipFree synchronises, so above the cap a block goes back to HIP, except
  // during a capture, when it is parked even above the cap.  The first free
  // outside a capture that finds the cache above the cap also returns ready
  // blocks (no queued work reads them) until the cache is under the cap;
  // blocks still pending follow once an allocation has synchronised them
  // into the ready lists and a later free finds the cache above the cap.
  if (P.cached + key.second > CACHE_MAX && g_capturing.load() == 0) {
    (void)hipFree(p);
    for (auto q = P.ready.begin(); q != P.ready.end() && P.cached > CACHE_MAX;) {
      auto &blocks = q->second;
      while (!blocks.empty() && P.cached > CACHE_MAX) {
        (void)hipFree(blocks.back());
        blocks.pop_back();
        P.cached -= q->first.second;
      }
      q = blocks.empty() ? P.ready.erase(q) : std::next(q);
    }
    return;
  }
  P.pending[key].push_back(p);
  P.cached += key.second;
}

// Taken under the pool lock, so a synchronisation or free that the pool has
// already started ends before the capture begins.
void pool_capture_begin() {
  Pool &P = pool();
  std::lock_guard<std::mutex> g(P.m);
  g_capturing.fetch_add(1);
}
void pool_capture_end() { g_capturing.fetch_sub(1); }

void dev_pool_trim() {
  Pool &P = pool();
  std::lock_guard<std::mutex> g(P.m);
  if (g_capturing.load() != 0)
    throw Error(LIGMG_ERR_INVALID, "ligmg_trim_memory: a preconditioner graph capture is in progress");
  flush_locked(P);
}

}  // namespace lg

int ligmg_trim_memory(void) {
  try {
    lg::dev_pool_trim();
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    lg::g_last_error = e.what();
    return e.code;
  }
}
