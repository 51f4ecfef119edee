// Seeded synthetic graph-Laplacian generators (host C++, OpenMP).
//
// Stands in for the reference's input path: MatrixMarket read, adjacency to
// Laplacian conversion (src/main.cpp:80-110: L = D - A with D the weighted
// degree) and largest connected component (include/ligmg/cc.hpp:56-106).
// Every random draw is a pure function of (seed, stream, index), so results
// do not depend on the thread count.
#include "../../include/ligmg_graphs.h"
#include "rng.hpp"

#include <omp.h>
#include <sys/mman.h>

#include <algorithm>
#include <cctype>
#include <climits>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

// Phase timing on stderr when LIGMG_GEN_VERBOSE is set.
struct Phase {
  bool on = getenv("LIGMG_GEN_VERBOSE") != nullptr;
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  void operator()(const char *what) {
    if (!on) return;
    auto now = std::chrono::steady_clock::now();
    fprintf(stderr, "[gen] %-14s %.2f s\n", what, std::chrono::duration<double>(now - t).count());
    t = now;
  }
};

using ligmg_rng::ctr_hash;
using ligmg_rng::u01;

// Large buffers: 2 MiB aligned and advised for transparent huge pages, so
// that first touch costs one fault per 2 MiB instead of per 4 KiB.  Freed
// with free().
void *huge_alloc(size_t bytes) {
  const size_t align = (size_t)2 << 20;
  size_t sz = (bytes + align - 1) / align * align;
  if (sz == 0) sz = align;
  void *p = aligned_alloc(align, sz);
  if (p) madvise(p, sz, MADV_HUGEPAGE);
  return p;
}

template <class T>
struct HugeArr {
  T *p = nullptr;
  size_t n = 0;
  explicit HugeArr(size_t count) : p((T *)huge_alloc(sizeof(T) * count)), n(count) {}
  HugeArr(const HugeArr &) = delete;
  HugeArr &operator=(const HugeArr &) = delete;
  ~HugeArr() { free(p); }
  T &operator[](size_t i) { return p[i]; }
  T *data() { return p; }
};

const uint64_t S_RMAT = 0x524d4154ull, S_WEIGHT = 0x57474854ull, S_PEND = 0x50454e44ull,
               S_PLEN = 0x504c454eull, S_BA = 0x42410000ull, S_RGGX = 0x52474758ull,
               S_RGGY = 0x52474759ull;

double normal01(uint64_t seed, uint64_t stream, uint64_t idx) {
  double a = u01(seed, stream, 2 * idx), b = u01(seed, stream, 2 * idx + 1);
  return std::sqrt(-2.0 * std::log(1.0 - a)) * std::cos(6.283185307179586 * b);
}

double edge_weight(int wkind, double p0, double p1, uint64_t seed, uint64_t key) {
  switch (wkind) {
    case LIGMG_W_UNIT: return 1.0;
    case LIGMG_W_UNIFORM: return p0 + (p1 - p0) * u01(seed, S_WEIGHT, key);
    case LIGMG_W_LOGNORMAL: return std::exp(p0 + p1 * normal01(seed, S_WEIGHT, key));
    default: return 1.0;
  }
}

struct Edges {
  std::vector<int32_t> u, v;  // u < v, unique
  std::vector<double> w;
};

struct ColVal {
  int32_t c;
  double v;
  bool operator<(const ColVal &o) const { return c < o.c; }
};

// Laplacian CSR from unique undirected edges (u < v).  Rows are sorted by
// column; the diagonal is the sequential sum of the row's weights in
// ascending column order.
int build_laplacian(int64_t n, const Edges &E, ligmg_graph *out) {
  const int64_t m = (int64_t)E.u.size();
  if (n <= 0 || 2 * m + n >= (int64_t)INT32_MAX) return 1;
  Phase ph;
  std::vector<int64_t> deg(n + 1, 0);
#pragma omp parallel for
  for (int64_t e = 0; e < m; e++) {
    __atomic_fetch_add(&deg[E.u[e]], 1, __ATOMIC_RELAXED);
    __atomic_fetch_add(&deg[E.v[e]], 1, __ATOMIC_RELAXED);
  }
  int64_t nnz = 2 * m + n;
  out->n = n;
  out->nnz = nnz;
  out->rowptr = (int32_t *)huge_alloc(sizeof(int32_t) * (n + 1));
  out->colind = (int32_t *)huge_alloc(sizeof(int32_t) * nnz);
  out->vals = (double *)huge_alloc(sizeof(double) * nnz);
  out->ncomp = 1;
  out->comp_ptr = (int64_t *)malloc(sizeof(int64_t) * 2);
  if (!out->rowptr || !out->colind || !out->vals || !out->comp_ptr) return 2;
  out->comp_ptr[0] = 0;
  out->comp_ptr[1] = n;
  std::vector<int64_t> start(n + 1, 0);
  for (int64_t i = 0; i < n; i++) start[i + 1] = start[i] + deg[i] + 1;
  for (int64_t i = 0; i <= n; i++) out->rowptr[i] = (int32_t)start[i];
  // the diagonal takes the first slot of every row; edges fill the rest in
  // any order (rows are sorted afterwards, so the result is deterministic)
  ph("  lap:deg");
  std::vector<int64_t> fill(n);
  ColVal *cv = (ColVal *)huge_alloc(sizeof(ColVal) * nnz);
  if (!cv) return 2;
  // first touch of the large arrays in parallel, in address order
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < nnz; k++) {
    cv[k].c = 0;
    out->colind[k] = 0;
    out->vals[k] = 0.0;
  }
#pragma omp parallel for
  for (int64_t i = 0; i < n; i++) {
    cv[start[i]] = {(int32_t)i, 0.0};
    fill[i] = start[i] + 1;
  }
#pragma omp parallel for
  for (int64_t e = 0; e < m; e++) {
    int32_t a = E.u[e], b = E.v[e];
    double w = E.w[e];
    cv[__atomic_fetch_add(&fill[a], 1, __ATOMIC_RELAXED)] = {b, -w};
    cv[__atomic_fetch_add(&fill[b], 1, __ATOMIC_RELAXED)] = {a, -w};
  }
  ph("  lap:scatter");
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < n; i++) {
    ColVal *r = cv + start[i];
    int64_t len = start[i + 1] - start[i];
    std::sort(r, r + len);
    double d = 0.0;
    int64_t dpos = -1;
    for (int64_t k = 0; k < len; k++) {
      if (r[k].c == i) dpos = k;
      else d += -r[k].v;
    }
    r[dpos].v = d;
    for (int64_t k = 0; k < len; k++) {
      out->colind[start[i] + k] = r[k].c;
      out->vals[start[i] + k] = r[k].v;
    }
  }
  ph("  lap:sort");
  free(cv);
  return 0;
}

// Union-find with path halving.
struct DSU {
  std::vector<int32_t> p;
  explicit DSU(int64_t n) : p(n) { std::iota(p.begin(), p.end(), 0); }
  int32_t find(int32_t x) {
    while (p[x] != x) {
      p[x] = p[p[x]];
      x = p[x];
    }
    return x;
  }
  void unite(int32_t a, int32_t b) {
    a = find(a);
    b = find(b);
    if (a == b) return;
    if (a < b) std::swap(a, b);
    p[a] = b;  // root is the smaller id
  }
};

// Keep the largest connected component (ties: the one holding the smallest
// vertex id); relabel kept vertices in ascending original order.
int64_t keep_lcc(int64_t n, Edges &E, std::vector<int32_t> *old_of_new = nullptr) {
  DSU d(n);
  for (size_t e = 0; e < E.u.size(); e++) d.unite(E.u[e], E.v[e]);
  std::vector<int64_t> size(n, 0);
  for (int64_t i = 0; i < n; i++) size[d.find((int32_t)i)]++;
  int64_t best = 0;
  for (int64_t i = 1; i < n; i++)
    if (size[i] > size[best]) best = i;
  std::vector<int32_t> nid(n, -1);
  int64_t cnt = 0;
  for (int64_t i = 0; i < n; i++)
    if (d.find((int32_t)i) == best) nid[i] = (int32_t)cnt++;
  if (old_of_new) {
    old_of_new->assign(cnt, 0);
    for (int64_t i = 0; i < n; i++)
      if (nid[i] >= 0) (*old_of_new)[nid[i]] = (int32_t)i;
  }
  size_t k = 0;
  for (size_t e = 0; e < E.u.size(); e++) {
    int32_t a = nid[E.u[e]], b = nid[E.v[e]];
    if (a < 0) continue;
    E.u[k] = a;
    E.v[k] = b;
    E.w[k] = E.w[e];
    k++;
  }
  E.u.resize(k);
  E.v.resize(k);
  E.w.resize(k);
  return cnt;
}

// Sort + unique undirected pairs (u < v) given as parallel arrays.
void dedup_pairs(int64_t n, std::vector<int32_t> &us, std::vector<int32_t> &vs) {
  const int64_t m = (int64_t)us.size();
  std::vector<std::atomic<int64_t>> cnt(n + 1);
  for (auto &c : cnt) c.store(0, std::memory_order_relaxed);
#pragma omp parallel for
  for (int64_t e = 0; e < m; e++) cnt[us[e]].fetch_add(1, std::memory_order_relaxed);
  std::vector<int64_t> start(n + 1, 0);
  for (int64_t i = 0; i < n; i++) start[i + 1] = start[i] + cnt[i].load();
  std::vector<std::atomic<int64_t>> pos(n);
  for (int64_t i = 0; i < n; i++) pos[i].store(start[i], std::memory_order_relaxed);
  HugeArr<int32_t> col(m);
#pragma omp parallel for
  for (int64_t e = 0; e < m; e++) col[pos[us[e]].fetch_add(1, std::memory_order_relaxed)] = vs[e];
  std::vector<int64_t> ulen(n, 0);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t i = 0; i < n; i++) {
    int32_t *b = col.data() + start[i], *e = col.data() + start[i + 1];
    std::sort(b, e);
    ulen[i] = std::unique(b, e) - b;
  }
  std::vector<int64_t> ostart(n + 1, 0);
  for (int64_t i = 0; i < n; i++) ostart[i + 1] = ostart[i] + ulen[i];
  us.assign(ostart[n], 0);
  vs.assign(ostart[n], 0);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t i = 0; i < n; i++)
    for (int64_t k = 0; k < ulen[i]; k++) {
      us[ostart[i] + k] = (int32_t)i;
      vs[ostart[i] + k] = col[start[i] + k];
    }
}

void assign_weights(Edges &E, int wkind, double p0, double p1, uint64_t seed) {
  E.w.resize(E.u.size());
#pragma omp parallel for
  for (int64_t e = 0; e < (int64_t)E.u.size(); e++)
    E.w[e] = edge_weight(wkind, p0, p1, seed, ((uint64_t)E.u[e] << 32) | (uint32_t)E.v[e]);
}

}  // namespace

extern "C" {

void ligmg_graph_free(ligmg_graph *g) {
  if (!g) return;
  free(g->rowptr);
  free(g->colind);
  free(g->vals);
  free(g->comp_ptr);
  memset(g, 0, sizeof(*g));
}

double ligmg_rng_u01(uint64_t seed, uint64_t stream, uint64_t idx) { return u01(seed, stream, idx); }
double ligmg_rng_normal(uint64_t seed, uint64_t stream, uint64_t idx) {
  return normal01(seed, stream, idx);
}

int ligmg_gen_grid2d(int64_t nx, int64_t ny, int wkind, double p0, double p1, uint64_t seed,
                     ligmg_graph *out) {
  if (nx <= 0 || ny <= 0 || !out) return 1;
  memset(out, 0, sizeof(*out));
  Edges E;
  for (int64_t y = 0; y < ny; y++)
    for (int64_t x = 0; x < nx; x++) {
      int32_t id = (int32_t)(y * nx + x);
      if (x + 1 < nx) { E.u.push_back(id); E.v.push_back(id + 1); }
      if (y + 1 < ny) { E.u.push_back(id); E.v.push_back((int32_t)(id + nx)); }
    }
  assign_weights(E, wkind, p0, p1, seed);
  return build_laplacian(nx * ny, E, out);
}

int ligmg_gen_grid3d(int64_t nx, int64_t ny, int64_t nz, int wkind, double p0, double p1,
                     uint64_t seed, ligmg_graph *out) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !out) return 1;
  memset(out, 0, sizeof(*out));
  Edges E;
  E.u.reserve(3 * nx * ny * nz);
  E.v.reserve(3 * nx * ny * nz);
  for (int64_t z = 0; z < nz; z++)
    for (int64_t y = 0; y < ny; y++)
      for (int64_t x = 0; x < nx; x++) {
        int32_t id = (int32_t)((z * ny + y) * nx + x);
        if (x + 1 < nx) { E.u.push_back(id); E.v.push_back(id + 1); }
        if (y + 1 < ny) { E.u.push_back(id); E.v.push_back((int32_t)(id + nx)); }
        if (z + 1 < nz) { E.u.push_back(id); E.v.push_back((int32_t)(id + nx * ny)); }
      }
  assign_weights(E, wkind, p0, p1, seed);
  return build_laplacian(nx * ny * nz, E, out);
}

int ligmg_gen_rmat(int scale, int edge_factor, double a, double b, double c, int wkind, double p0,
                   double p1, double pendant_frac, int pendant_max, uint64_t seed, int nthreads,
                   ligmg_graph *out) {
  if (scale < 1 || scale > 30 || edge_factor < 1 || !out) return 1;
  memset(out, 0, sizeof(*out));
  if (nthreads > 0) omp_set_num_threads(nthreads);
  Phase ph;
  const int64_t n0 = (int64_t)1 << scale;
  const int64_t E0 = (int64_t)edge_factor << scale;
  std::vector<int32_t> us(E0), vs(E0);
  const double ab = a + b, abc = a + b + c;
  // u01(seed, S_RMAT, idx) with the per-stream key hoisted out of the loop
  const uint64_t key = ligmg_rng::mix64(seed ^ ligmg_rng::mix64(S_RMAT));
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < E0; e++) {
    uint64_t x = 0, y = 0;
    for (int l = 0; l < scale; l++) {
      double r = (double)(ligmg_rng::mix64(key + (uint64_t)e * 64 + l) >> 11) *
                 (1.0 / 9007199254740992.0);
      uint64_t bx = 0, by = 0;
      if (r < a) {
      } else if (r < ab) {
        by = 1;
      } else if (r < abc) {
        bx = 1;
      } else {
        bx = 1;
        by = 1;
      }
      x = (x << 1) | bx;
      y = (y << 1) | by;
    }
    int32_t lo = (int32_t)std::min(x, y), hi = (int32_t)std::max(x, y);
    us[e] = lo;
    vs[e] = (lo == hi) ? -1 : hi;
  }
  // drop self loops
  int64_t k = 0;
  for (int64_t e = 0; e < E0; e++)
    if (vs[e] >= 0) {
      us[k] = us[e];
      vs[k] = vs[e];
      k++;
    }
  us.resize(k);
  vs.resize(k);
  ph("edges");
  dedup_pairs(n0, us, vs);
  ph("dedup");
  Edges E;
  E.u.swap(us);
  E.v.swap(vs);
  E.w.assign(E.u.size(), 0.0);
  int64_t n = keep_lcc(n0, E);
  ph("lcc");
  // pendant paths
  if (pendant_frac > 0 && pendant_max > 0) {
    int64_t next = n;
    for (int64_t v = 0; v < n; v++) {
      if (u01(seed, S_PEND, v) >= pendant_frac) continue;
      int len = 1 + (int)(u01(seed, S_PLEN, v) * pendant_max);
      if (len > pendant_max) len = pendant_max;
      int32_t prev = (int32_t)v;
      for (int l = 0; l < len; l++) {
        E.u.push_back(prev);
        E.v.push_back((int32_t)next);
        prev = (int32_t)next;
        next++;
      }
    }
    n = next;
  }
  ph("pendants");
  assign_weights(E, wkind, p0, p1, seed);
  ph("weights");
  int rc = build_laplacian(n, E, out);
  ph("laplacian");
  return rc;
}

int ligmg_gen_ba(int64_t n, int m, int wkind, double p0, double p1, uint64_t seed,
                 ligmg_graph *out) {
  if (m < 1 || n < m + 1 || !out) return 1;
  memset(out, 0, sizeof(*out));
  Edges E;
  std::vector<int32_t> rep;
  rep.reserve(2 * (size_t)n * m);
  for (int i = 0; i <= m; i++)
    for (int j = i + 1; j <= m; j++) {
      E.u.push_back(i);
      E.v.push_back(j);
      rep.push_back(i);
      rep.push_back(j);
    }
  uint64_t draw = 0;
  std::vector<int32_t> tg;
  for (int64_t v = m + 1; v < n; v++) {
    tg.clear();
    while ((int)tg.size() < m) {
      uint64_t h = ctr_hash(seed, S_BA, draw++);
      int32_t t = rep[h % rep.size()];
      if (std::find(tg.begin(), tg.end(), t) == tg.end()) tg.push_back(t);
    }
    for (int32_t t : tg) {
      E.u.push_back(t);
      E.v.push_back((int32_t)v);
      rep.push_back(t);
      rep.push_back((int32_t)v);
    }
  }
  assign_weights(E, wkind, p0, p1, seed);
  return build_laplacian(n, E, out);
}

int ligmg_gen_rgg(int64_t n, double radius, int wkind, double p0, double p1, uint64_t seed,
                  ligmg_graph *out) {
  if (n < 2 || radius <= 0 || !out) return 1;
  memset(out, 0, sizeof(*out));
  std::vector<double> X(n), Y(n);
  for (int64_t i = 0; i < n; i++) {
    X[i] = u01(seed, S_RGGX, i);
    Y[i] = u01(seed, S_RGGY, i);
  }
  int64_t nc = std::max<int64_t>(1, (int64_t)std::floor(1.0 / radius));
  std::vector<int64_t> cstart(nc * nc + 1, 0);
  auto cell = [&](int64_t i) {
    int64_t cx = std::min<int64_t>(nc - 1, (int64_t)(X[i] * nc));
    int64_t cy = std::min<int64_t>(nc - 1, (int64_t)(Y[i] * nc));
    return cy * nc + cx;
  };
  for (int64_t i = 0; i < n; i++) cstart[cell(i) + 1]++;
  for (int64_t c = 0; c < nc * nc; c++) cstart[c + 1] += cstart[c];
  std::vector<int32_t> members(n);
  {
    std::vector<int64_t> f(cstart.begin(), cstart.end() - 1);
    for (int64_t i = 0; i < n; i++) members[f[cell(i)]++] = (int32_t)i;
  }
  const double r2 = radius * radius;
  std::vector<Edges> per(omp_get_max_threads());
#pragma omp parallel
  {
    Edges &L = per[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 1024)
    for (int64_t i = 0; i < n; i++) {
      int64_t c = cell(i), cx = c % nc, cy = c / nc;
      for (int64_t dy = -1; dy <= 1; dy++)
        for (int64_t dx = -1; dx <= 1; dx++) {
          int64_t x2 = cx + dx, y2 = cy + dy;
          if (x2 < 0 || y2 < 0 || x2 >= nc || y2 >= nc) continue;
          int64_t cc = y2 * nc + x2;
          for (int64_t q = cstart[cc]; q < cstart[cc + 1]; q++) {
            int32_t j = members[q];
            if (j <= i) continue;
            double ddx = X[i] - X[j], ddy = Y[i] - Y[j], d2 = ddx * ddx + ddy * ddy;
            if (d2 >= r2) continue;
            L.u.push_back((int32_t)i);
            L.v.push_back(j);
            L.w.push_back(wkind == LIGMG_W_GAUSS_DIST ? (double)(float)std::exp(-d2 / r2)
                                                      : d2);
          }
        }
    }
  }
  Edges E;
  for (auto &L : per) {
    E.u.insert(E.u.end(), L.u.begin(), L.u.end());
    E.v.insert(E.v.end(), L.v.begin(), L.v.end());
    E.w.insert(E.w.end(), L.w.begin(), L.w.end());
  }
  // canonical edge order (thread partition independent)
  std::vector<int64_t> ord(E.u.size());
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int64_t p, int64_t q) {
    return E.u[p] != E.u[q] ? E.u[p] < E.u[q] : E.v[p] < E.v[q];
  });
  Edges S;
  S.u.resize(ord.size());
  S.v.resize(ord.size());
  S.w.resize(ord.size());
  for (size_t t = 0; t < ord.size(); t++) {
    S.u[t] = E.u[ord[t]];
    S.v[t] = E.v[ord[t]];
    S.w[t] = E.w[ord[t]];
  }
  int64_t nl = keep_lcc(n, S);
  if (wkind != LIGMG_W_GAUSS_DIST) assign_weights(S, wkind, p0, p1, seed);
  return build_laplacian(nl, S, out);
}

int ligmg_graph_union(const ligmg_graph *gs, int count, ligmg_graph *out) {
  if (!gs || count < 1 || !out) return 1;
  memset(out, 0, sizeof(*out));
  int64_t n = 0, nnz = 0;
  for (int g = 0; g < count; g++) {
    n += gs[g].n;
    nnz += gs[g].nnz;
  }
  if (nnz >= (int64_t)INT32_MAX || n >= (int64_t)INT32_MAX) return 1;
  out->n = n;
  out->nnz = nnz;
  out->rowptr = (int32_t *)huge_alloc(sizeof(int32_t) * (n + 1));
  out->colind = (int32_t *)huge_alloc(sizeof(int32_t) * nnz);
  out->vals = (double *)huge_alloc(sizeof(double) * nnz);
  out->ncomp = count;
  out->comp_ptr = (int64_t *)malloc(sizeof(int64_t) * (count + 1));
  if (!out->rowptr || !out->colind || !out->vals || !out->comp_ptr) return 2;
  int64_t vo = 0, eo = 0;
  out->comp_ptr[0] = 0;
  for (int g = 0; g < count; g++) {
    const ligmg_graph &G = gs[g];
    for (int64_t i = 0; i < G.n; i++) out->rowptr[vo + i] = (int32_t)(eo + G.rowptr[i]);
    for (int64_t e = 0; e < G.nnz; e++) {
      out->colind[eo + e] = (int32_t)(G.colind[e] + vo);
      out->vals[eo + e] = G.vals[e];
    }
    vo += G.n;
    eo += G.nnz;
    out->comp_ptr[g + 1] = vo;
  }
  out->rowptr[n] = (int32_t)nnz;
  return 0;
}

int ligmg_graph_from_edges(int64_t n, int64_t m, const int32_t *u, const int32_t *v,
                           const double *w, ligmg_graph *out) {
  if (n <= 0 || m < 0 || !out) return 1;
  memset(out, 0, sizeof(*out));
  // canonicalise and sum duplicates in (lo, hi) order
  std::vector<int64_t> ord(m);
  std::vector<int32_t> lo(m), hi(m);
  for (int64_t e = 0; e < m; e++) {
    if (u[e] == v[e] || u[e] < 0 || v[e] < 0 || u[e] >= n || v[e] >= n) return 1;
    lo[e] = std::min(u[e], v[e]);
    hi[e] = std::max(u[e], v[e]);
  }
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int64_t p, int64_t q) {
    return lo[p] != lo[q] ? lo[p] < lo[q] : hi[p] < hi[q];
  });
  Edges E;
  for (int64_t t = 0; t < m; t++) {
    int64_t e = ord[t];
    if (!E.u.empty() && E.u.back() == lo[e] && E.v.back() == hi[e]) E.w.back() += w[e];
    else {
      E.u.push_back(lo[e]);
      E.v.push_back(hi[e]);
      E.w.push_back(w[e]);
    }
  }
  return build_laplacian(n, E, out);
}

// MatrixMarket coordinate reader (the reference reads its input with
// CombBLAS ParallelReadMM, src/main.cpp:33-40).  Fields real, integer or
// pattern (value 1); symmetry general or symmetric (the stored triangle is
// mirrored).  Indices are 1-based in the file.  Duplicates are summed in
// file order; every row's columns come out sorted.  Square only.
int ligmg_graph_read_mtx(const char *path, ligmg_graph *out) {
  if (!path || !out) return 1;
  memset(out, 0, sizeof(*out));
  FILE *f = fopen(path, "rb");
  if (!f) return 3;
  fseek(f, 0, SEEK_END);
  const long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  if (sz <= 0) {
    fclose(f);
    return 4;
  }
  std::vector<char> buf((size_t)sz + 1);
  const size_t got = fread(buf.data(), 1, (size_t)sz, f);
  fclose(f);
  if (got != (size_t)sz) return 3;
  buf[sz] = 0;
  char *p = buf.data(), *end = buf.data() + sz;
  auto next_line = [&](char *q) {
    while (q < end && *q != '\n') q++;
    return q < end ? q + 1 : end;
  };
  // header: %%MatrixMarket matrix coordinate <field> <symmetry>
  char obj[64] = {0}, fmt[64] = {0}, field[64] = {0}, sym[64] = {0};
  if (sscanf(p, "%%%%MatrixMarket %63s %63s %63s %63s", obj, fmt, field, sym) != 4) return 4;
  for (char *c : {obj, fmt, field, sym})
    for (; *c; c++) *c = (char)tolower(*c);
  if (strcmp(obj, "matrix") || strcmp(fmt, "coordinate")) return 4;
  const bool pattern = !strcmp(field, "pattern");
  if (!pattern && strcmp(field, "real") && strcmp(field, "integer")) return 4;
  const bool symmetric = !strcmp(sym, "symmetric");
  if (!symmetric && strcmp(sym, "general")) return 4;
  p = next_line(p);
  while (p < end && (*p == '%' || *p == '\n' || *p == '\r')) p = next_line(p);
  long long nr = 0, nc = 0, ne = 0;
  if (sscanf(p, "%lld %lld %lld", &nr, &nc, &ne) != 3 || nr != nc || nr <= 0 || nr >= INT32_MAX || ne < 0)
    return 4;
  p = next_line(p);
  // entries: one chunk per thread, cut at line boundaries
  const int T = std::max(1, omp_get_max_threads());
  std::vector<char *> cut(T + 1);
  cut[0] = p;
  cut[T] = end;
  for (int t = 1; t < T; t++) {
    char *q = p + (end - p) * t / T;
    cut[t] = std::max(cut[t - 1], (q > p && q[-1] == '\n') ? q : next_line(q));
  }
  std::vector<std::vector<int32_t>> I(T), J(T);
  std::vector<std::vector<double>> V(T);
  std::vector<int> bad(T, 0);
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num();
    char *q = cut[t];
    while (q < cut[t + 1]) {
      char *e2;
      const long long i = strtoll(q, &e2, 10);
      if (e2 == q) {  // blank line
        q = next_line(q);
        continue;
      }
      q = e2;
      const long long j = strtoll(q, &e2, 10);
      if (e2 == q) {
        bad[t] = 1;
        break;
      }
      q = e2;
      double v = 1.0;
      if (!pattern) {
        v = strtod(q, &e2);
        if (e2 == q) {
          bad[t] = 1;
          break;
        }
        q = e2;
      }
      if (i < 1 || j < 1 || i > nr || j > nr) {
        bad[t] = 1;
        break;
      }
      I[t].push_back((int32_t)(i - 1));
      J[t].push_back((int32_t)(j - 1));
      V[t].push_back(v);
      q = next_line(q);
    }
  }
  for (int t = 0; t < T; t++)
    if (bad[t]) return 4;
  const int64_t n = nr;
  std::vector<int64_t> base(T + 1, 0);
  for (int t = 0; t < T; t++) base[t + 1] = base[t] + (int64_t)I[t].size();
  if (base[T] != ne) return 4;
  // COO (file order, mirrored entries right after their originals) -> CSR
  std::vector<int64_t> cnt(n + 1, 0);
  for (int t = 0; t < T; t++)
    for (size_t k = 0; k < I[t].size(); k++) {
      cnt[I[t][k] + 1]++;
      if (symmetric && I[t][k] != J[t][k]) cnt[J[t][k] + 1]++;
    }
  for (int64_t r = 0; r < n; r++) cnt[r + 1] += cnt[r];
  const int64_t m = cnt[n];
  if (m >= INT32_MAX) return 4;
  std::vector<int32_t> col(m);
  std::vector<double> val(m);
  std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
  for (int t = 0; t < T; t++)
    for (size_t k = 0; k < I[t].size(); k++) {
      const int32_t i = I[t][k], j = J[t][k];
      col[pos[i]] = j;
      val[pos[i]++] = V[t][k];
      if (symmetric && i != j) {
        col[pos[j]] = i;
        val[pos[j]++] = V[t][k];
      }
    }
  // sort each row by column (stable: duplicates keep file order) and sum them
  std::vector<int64_t> rlen(n, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t r = 0; r < n; r++) {
    const int64_t a = cnt[r], b = cnt[r + 1];
    std::vector<std::pair<int32_t, int64_t>> e;
    e.reserve(b - a);
    for (int64_t k = a; k < b; k++) e.push_back({col[k], k});
    std::stable_sort(e.begin(), e.end(),
                     [](const std::pair<int32_t, int64_t> &x, const std::pair<int32_t, int64_t> &y) {
                       return x.first < y.first;
                     });
    std::vector<int32_t> c2;
    std::vector<double> v2;
    for (auto &x : e) {
      if (!c2.empty() && c2.back() == x.first) v2.back() += val[x.second];
      else {
        c2.push_back(x.first);
        v2.push_back(val[x.second]);
      }
    }
    for (size_t k = 0; k < c2.size(); k++) {
      col[a + k] = c2[k];
      val[a + k] = v2[k];
    }
    rlen[r] = (int64_t)c2.size();
  }
  int64_t nnz = 0;
  for (int64_t r = 0; r < n; r++) nnz += rlen[r];
  out->n = n;
  out->nnz = nnz;
  out->rowptr = (int32_t *)huge_alloc(sizeof(int32_t) * (n + 1));
  out->colind = (int32_t *)huge_alloc(sizeof(int32_t) * std::max<int64_t>(nnz, 1));
  out->vals = (double *)huge_alloc(sizeof(double) * std::max<int64_t>(nnz, 1));
  out->ncomp = 1;
  out->comp_ptr = (int64_t *)malloc(sizeof(int64_t) * 2);
  if (!out->rowptr || !out->colind || !out->vals || !out->comp_ptr) {
    ligmg_graph_free(out);
    return 2;
  }
  out->comp_ptr[0] = 0;
  out->comp_ptr[1] = n;
  out->rowptr[0] = 0;
  for (int64_t r = 0; r < n; r++) out->rowptr[r + 1] = out->rowptr[r] + (int32_t)rlen[r];
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; r++) {
    const int64_t a = cnt[r], d = out->rowptr[r];
    for (int64_t k = 0; k < rlen[r]; k++) {
      out->colind[d + k] = col[a + k];
      out->vals[d + k] = val[a + k];
    }
  }
  return 0;
}

int ligmg_gen_rhs(int64_t n, int k, uint64_t seed, int ncomp, const int64_t *comp_ptr,
                  double *out) {
  if (n <= 0 || k <= 0 || !out) return 1;
  int64_t one[2] = {0, n};
  if (!comp_ptr) {
    comp_ptr = one;
    ncomp = 1;
  }
  for (int c = 0; c < k; c++) {
    double *x = out + (int64_t)c * n;
#pragma omp parallel for
    for (int64_t i = 0; i < n; i++) x[i] = normal01(seed, (uint64_t)c, i);
    for (int g = 0; g < ncomp; g++) {
      double s = 0.0;
      for (int64_t i = comp_ptr[g]; i < comp_ptr[g + 1]; i++) s += x[i];
      double mean = s / (double)(comp_ptr[g + 1] - comp_ptr[g]);
      for (int64_t i = comp_ptr[g]; i < comp_ptr[g + 1]; i++) x[i] -= mean;
    }
  }
  return 0;
}

}  // extern "C"
