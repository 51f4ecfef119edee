// Host orchestration of the device hierarchy and the C ABI
// (include/ligmg_amd.h).
//
//   ligmg_create  <- ligmg::Solver(A, opts)        ligmg.hpp:17-28
//                     make_hierarchy               multigrid.hpp:161-263
//   ligmg_solve   <- Solver::solve(b)              ligmg.hpp:31-33
//                     ligmg::solve                 multigrid.hpp:315-370
//   cycles        <- vcycle / kcycle / do_level    multigrid.hpp:86-130,
//                     aggregation.hpp:209-241, elimination.hpp:33-52
//   Krylov        <- petsc_fcg (PETSc KSPCG/FCG)   cg.hpp:18-116
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cerrno>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <sys/stat.h>

#include "../../include/ligmg_amd.h"
#include "ops.hpp"
#include "prims.hpp"
#include "setup.hpp"
#include "solve_dev.hpp"
#include "tail.hpp"
#include "petsc_io.hpp"

namespace lg {

thread_local std::string g_last_error;

namespace {

// ------------------------------------------------------------ Krylov kernels
using namespace sdev;

// per-block scalar logic of CG / FCG (one thread per block)
__global__ void k_kr_alpha(int nb, int fcg, double *sc) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nb) kr_alpha(sc + c * SC_N, fcg);
}
__global__ void k_kr_check(int nb, double rtol, double *sc) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nb) kr_check(sc + c * SC_N, rtol);
}
__global__ void k_kr_beta(int nb, int fcg, double *sc) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nb) kr_beta(sc + c * SC_N, fcg);
}
__global__ void k_kr_init(int nb, double *sc) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nb) kr_init(sc + c * SC_N);
}

// x += alpha_c p ; r -= alpha_c q   (segment-per-workgroup, block-aware)
template <class T>
__global__ void __launch_bounds__(256) k_kr_update(T *x, T *r, const T *p, const T *q,
                                                   const int64_t *lo, const int64_t *hi,
                                                   const int32_t *segblk, const double *sc) {
  const int s = blockIdx.x;
  const double a = sc[segblk[s] * SC_N + SC_ALPHA];
  if (a == 0.0) return;
  const T at = (T)a;
  for (int64_t i = lo[s] + threadIdx.x; i < hi[s]; i += 256) {
    x[i] += at * p[i];
    r[i] -= at * q[i];
  }
}
// p = z + beta_c p
template <class T>
__global__ void __launch_bounds__(256) k_kr_pupdate(T *p, const T *z, const int64_t *lo,
                                                    const int64_t *hi, const int32_t *segblk,
                                                    const double *sc) {
  const int s = blockIdx.x;
  const T b = (T)sc[segblk[s] * SC_N + SC_BETA];
  for (int64_t i = lo[s] + threadIdx.x; i < hi[s]; i += 256) p[i] = z[i] + b * p[i];
}

// FGMRES: y += sum_t coef[blk * K + t] V_t (V_t = V + t ld), added in t order
__global__ void __launch_bounds__(256) k_blk_lincomb(double *y, const double *V, int64_t ld, int k,
                                                     const double *coef, int K, const int64_t *lo,
                                                     const int64_t *hi, const int32_t *segblk) {
  const int s = blockIdx.x;
  const double *cf = coef + (int64_t)segblk[s] * K;
  for (int64_t i = lo[s] + threadIdx.x; i < hi[s]; i += 256) {
    double acc = y[i];
    for (int t = 0; t < k; t++) acc += cf[t] * V[t * ld + i];
    y[i] = acc;
  }
}
// FGMRES: y = x * sc[blk]
__global__ void __launch_bounds__(256) k_blk_scale(double *y, const double *x, const double *sc,
                                                   const int64_t *lo, const int64_t *hi, const int32_t *segblk) {
  const int s = blockIdx.x;
  const double a = sc[segblk[s]];
  for (int64_t i = lo[s] + threadIdx.x; i < hi[s]; i += 256) y[i] = x[i] * a;
}

inline bool fcg_like(int t) { return t == LIGMG_KSP_FCG || t == LIGMG_KSP_PIPEFCG; }

// batch: dot results out[(blk * np + j) * kp + c] -> scalar sets of (blk, c)
__global__ void k_kr_bstore(int nb, int kp, int np, const double *out, int s0, int s1, double *sc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nb * kp) return;
  const int blk = t / kp, c = t % kp;
  sc[t * SC_N + s0] = out[((int64_t)blk * np) * kp + c];
  if (np > 1) sc[t * SC_N + s1] = out[((int64_t)blk * np + 1) * kp + c];
}
// x += alpha p, r -= alpha q per (block, column); alpha = 0 marks a finished
// column, whose x and r stay.  Also writes the segment partials of r.r
// exactly as the block dots form them (part[s * kp + c]: the same loop and
// combining order), and r in FP32 when r32 is set (the mixed outer
// preconditioner's input).
template <class T>
__global__ void __launch_bounds__(256) k_kr_bupdate(T *x, T *r, const T *p, const T *q, int kp,
                                                    int lk, const int64_t *lo, const int64_t *hi,
                                                    const int32_t *segblk, const double *sc, double *part,
                                                    float *r32) {
  const int s = blockIdx.x;
  const int G = 256 >> lk, g = threadIdx.x >> lk, c = threadIdx.x & (kp - 1);
  const double a = sc[((int64_t)segblk[s] * kp + c) * SC_N + SC_ALPHA];
  const T at = (T)a;
  double acc = 0;
  for (int64_t i = lo[s] + g; i < hi[s]; i += G) {
    const int64_t o = i * kp + c;
    T rv = r[o];
    if (a != 0.0) {
      x[o] += at * p[o];
      rv -= at * q[o];
      r[o] = rv;
    }
    acc += (double)rv * (double)rv;
    if (r32) r32[o] = (float)rv;
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  if (g == 0) {
    double t = 0;
    for (int gg = 0; gg < G; gg++) t += red[gg * kp + c];
    part[(int64_t)s * kp + c] = t;
  }
}
template <class T>
__global__ void __launch_bounds__(256) k_kr_bpupdate(T *p, const T *z, int kp, int lk,
                                                     const int64_t *lo, const int64_t *hi,
                                                     const int32_t *segblk, const double *sc) {
  const int s = blockIdx.x;
  const int G = 256 >> lk, g = threadIdx.x >> lk, c = threadIdx.x & (kp - 1);
  const T b = (T)sc[((int64_t)segblk[s] * kp + c) * SC_N + SC_BETA];
  for (int64_t i = lo[s] + g; i < hi[s]; i += G) {
    const int64_t o = i * kp + c;
    p[o] = z[o] + b * p[o];
  }
}

template <class T>
struct KWork {
  DevBuf<T> r, z, p, q, x;
  DevBuf<double> sc, out;
  void alloc(int64_t n, int nb) {
    r.alloc(n);
    z.alloc(n);
    p.alloc(n);
    q.alloc(n);
    x.alloc(n);
    sc.alloc((size_t)nb * SC_N);
    out.alloc((size_t)nb * 4);
  }
};

std::vector<double> cheb_alpha(double lo, double hi, int steps, std::vector<double> &be) {
  // Saad, Alg. 12.1, three-term form (see the oracle for the derivation)
  double theta = (hi + lo) / 2, delta = (hi - lo) / 2, sigma = theta / delta;
  double rho = 1 / sigma;
  std::vector<double> al(std::max(steps, 0), 0.0);
  be.assign(std::max(steps, 0), 0.0);
  if (steps > 0) be[0] = 1 / theta;
  for (int k = 1; k < steps; k++) {
    double rn = 1 / (2 * sigma - rho);
    al[k] = rn * rho;
    be[k] = 2 * rn / delta;
    rho = rn;
  }
  return al;
}

}  // namespace

// Per-level work vectors of one precision.
template <class T>
struct LevelWork {
  DevBuf<T> B[3], R, b;
  KWork<T> kw;
};

// Per-level block work vectors (n x kp) of one precision.
template <class T>
struct BLevelWork {
  DevBuf<T> B[3], R, b;
  KWork<T> kw;  // K-cycle: vectors n*kp, scalars nb*kp sets
};

inline int ilog2(int k) {
  int l = 0;
  while ((1 << l) < k) l++;
  return l;
}

struct Solver {
  ligmg_options o;
  hipStream_t s = nullptr;
  std::vector<DLevel> H;
  std::vector<RedPlan> red;
  std::vector<LevelWork<double>> W64;
  std::vector<LevelWork<float>> W32;
  std::vector<BLevelWork<double>> BW64;
  std::vector<BLevelWork<float>> BW32;
  int bkp = 0;  // block width the batch buffers are sized for
  DevBuf<double> bparts, bpart_red, bsc_out;
  KWork<double> bouter;
  std::vector<std::vector<double>> al_pre, be_pre, al_post, be_post;
  KWork<double> outer;
  DevBuf<double> b0, sol, tmp0, dsc;
  // solve position of level 0 -> caller's vertex (empty: identity), and
  // the setup order's permutation of the caller's vertices (ordering)
  DevBuf<int32_t> perm, perm_input;
  DevBuf<double> rhs_base;  // reduced right-hand side of the outer solve
  double setup_time = 0;
  bool mixed = false;
  // captured preconditioner applications, keyed by (level, r, z)
  struct GraphKey {
    size_t lvl;
    const void *r, *z;
  };
  std::vector<std::pair<GraphKey, hipGraphExec_t>> graphs;
  // fused coarse tail: levels tail_l0.. (H.size(): none), the row
  // schedules of their operators (A, P, R, M per level), the device tables
  // the programs index, and one recorded program per entry level and
  // precision
  size_t tail_l0 = 0;
  std::vector<TailSched> tail_sched;
  DevBuf<TailOp<double>> tail_ops64;
  DevBuf<TailOp<float>> tail_ops32;
  DevBuf<TailLevel<double>> tail_lv64;
  DevBuf<TailLevel<float>> tail_lv32;
  struct TailProg {
    DevBuf<TailInstr> code;
    int64_t len = 0;
    const void *in = nullptr, *out = nullptr;
  };
  std::vector<TailProg> tail_prog[2];
  TailParams tail_p;
  static constexpr int64_t TAIL_MAX_INSTR = 1 << 18;

  ~Solver() {
    for (auto &g : graphs) (void)hipGraphExecDestroy(g.second);
    if (s) (void)hipStreamDestroy(s);
  }

  template <class T>
  LevelWork<T> &work(size_t l);

  // ---------------------------------------------------------------- setup
  void build(int64_t n, const int32_t *rp, const int32_t *ci, const double *v, int mem,
             const std::vector<int64_t> &blocks0) {
    LG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    mixed = o.precision == LIGMG_PREC_MIXED;
    auto t0 = std::chrono::steady_clock::now();
    DCsr cur;
    cur.n = n;
    cur.ncols = n;
    int32_t nnz32 = 0;
    if (mem == LIGMG_MEM_HOST) nnz32 = rp[n] - rp[0];
    else {
      int32_t ends[2];
      to_host(&ends[0], rp, 1, s);
      to_host(&ends[1], rp + n, 1, s);
      LG_HIP(hipStreamSynchronize(s));
      nnz32 = ends[1] - ends[0];
      LG_CHECK(ends[0] == 0, LIGMG_ERR_INVALID, "rowptr[0] must be 0");
    }
    cur.nnz = nnz32;
    cur.rp.alloc(n + 1);
    cur.ci.alloc(std::max<int64_t>(cur.nnz, 1));
    cur.v.alloc(std::max<int64_t>(cur.nnz, 1));
    if (mem == LIGMG_MEM_HOST) {
      LG_CHECK(rp[0] == 0, LIGMG_ERR_INVALID, "rowptr[0] must be 0");
      to_device(cur.rp.p, rp, n + 1, s);
      to_device(cur.ci.p, ci, cur.nnz, s);
      to_device(cur.v.p, v, cur.nnz, s);
    } else {
      d2d(cur.rp.p, rp, n + 1, s);
      d2d(cur.ci.p, ci, cur.nnz, s);
      d2d(cur.v.p, v, cur.nnz, s);
    }
    LG_HIP(hipStreamSynchronize(s));
    validate_csr(n, n, cur.rp.p, cur.ci.p, s, true);
    if (o.ordering == LIGMG_ORDER_DEGREE) degree_order(cur, blocks0, perm_input, s);
    auto last = t0;
    auto lap = [&](const char *what, size_t l, int64_t nn, int64_t nz) {
      if (!o.verbose) return;
      auto now = std::chrono::steady_clock::now();
      fprintf(stderr, "[ligmg] level %zu %-6s n=%lld nnz=%lld  %.3f s\n", l, what, (long long)nn,
              (long long)nz, std::chrono::duration<double>(now - last).count());
      last = now;
    };
    lap("input", 0, cur.n, cur.nnz);
    // make_hierarchy (multigrid.hpp:191-213)
    while ((int)H.size() < o.max_levels && (H.empty() || H.back().A.nnz > o.min_size)) {
      if (o.always_elim) {
        DLevel L;
        DCsr c;
        if (setup_elimination(cur, o, L, c, s)) {
          L.A = std::move(cur);
          H.push_back(std::move(L));
          cur = std::move(c);
          lap("elim", H.size() - 1, H.back().A.n, H.back().A.nnz);
        }
      }
      DLevel L;
      DCsr c;
      setup_aggregation(cur, o, (int)H.size(), L, c, s);
      L.A = std::move(cur);
      H.push_back(std::move(L));
      cur = std::move(c);
      lap("agg", H.size() - 1, H.back().A.n, H.back().A.nnz);
    }
    {
      DLevel C;
      C.kind = LIGMG_LEVEL_COARSE;
      C.n = cur.n;
      auto hrp = download(cur.rp.p, cur.n + 1, s);
      auto hci = download(cur.ci.p, cur.nnz, s);
      auto hv = download(cur.v.p, cur.nnz, s);
      coarse_pinv(cur.n, hrp, hci, hv, C.pinv, s);
      C.A = std::move(cur);
      H.push_back(std::move(C));
      lap("coarse", H.size() - 1, H.back().A.n, H.back().A.nnz);
    }
    H[0].blocks = blocks0;
    finish_setup();
    lap("solve-prep", H.size(), 0, 0);
    setup_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  // Everything after the hierarchy itself: per-level blocks, the solve
  // layout, SpMV units, smoothers and eigen-estimates, FP32 copies, work
  // vectors, slicing, the fused tail.  Shared by build and load.
  void finish_setup() {
    // with o.verbose, the time of each solve-preparation phase
    auto t_last = std::chrono::steady_clock::now();
    auto phase = [&](const char *what) {
      if (!o.verbose) return;
      LG_HIP(hipStreamSynchronize(s));
      const auto now = std::chrono::steady_clock::now();
      fprintf(stderr, "[ligmg]   solve-prep %-14s %.3f s\n", what,
              std::chrono::duration<double>(now - t_last).count());
      t_last = now;
    };
    double t_units = 0, t_eig = 0, t_red = 0;
    auto timed = [&](double &acc, auto &&f) {
      if (!o.verbose) {
        f();
        return;
      }
      LG_HIP(hipStreamSynchronize(s));
      const auto a = std::chrono::steady_clock::now();
      f();
      LG_HIP(hipStreamSynchronize(s));
      acc += std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
    };
    // blocks per level: level-0 blocks mapped through kept ids / roots
    for (size_t l = 0; l + 1 < H.size(); l++) {
      const DLevel &L = H[l];
      std::vector<int64_t> nb(L.blocks.size(), 0);
      if (L.blocks.size() == 2) {  // one block: all of the next level's vertices
        nb[1] = H[l + 1].A.n;
        H[l + 1].blocks = nb;
        continue;
      }
      std::vector<int32_t> keep;
      if (L.kind == LIGMG_LEVEL_ELIM) keep = download(L.new_id.p, L.n, s);
      else {
        keep = download(L.ind.p, L.n, s);
        for (int64_t i = 0; i < L.n; i++) keep[i] = (keep[i] == i) ? 1 : -1;
      }
      for (size_t c = 0; c + 1 < L.blocks.size(); c++) {
        int64_t cnt = 0;
        for (int64_t i = L.blocks[c]; i < L.blocks[c + 1]; i++) cnt += keep[i] >= 0;
        nb[c + 1] = nb[c] + cnt;
      }
      H[l + 1].blocks = nb;
    }
    phase("blocks");
    if (o.solve_layout && o.spmv_slice == 0) apply_solve_layout();
    for (auto &L : H) L.tpos.release();
    phase("layout");
    // caller's vertex of each level-0 solve position
    if (H[0].sperm.p && perm_input.p) {
      perm.alloc(H[0].A.n);
      relabel_map(H[0].A.n, H[0].sperm.p, perm_input.p, nullptr, perm.p, s);
    } else if (H[0].sperm.p) {
      perm.alloc(H[0].A.n);
      d2d(perm.p, H[0].sperm.p, H[0].A.n, s);
    } else if (perm_input.p) {
      perm.alloc(H[0].A.n);
      d2d(perm.p, perm_input.p, H[0].A.n, s);
    }
    // smoothers, SpMV units, FP32 copies, work vectors
    red.resize(H.size());
    W64.resize(H.size());
    if (mixed) W32.resize(H.size());
    al_pre.resize(H.size());
    be_pre.resize(H.size());
    al_post.resize(H.size());
    be_post.resize(H.size());
    for (size_t l = 0; l < H.size(); l++) {
      DLevel &L = H[l];
      timed(t_red, [&] { red[l].build(L.A.n, L.blocks, s); });
      if (L.kind != LIGMG_LEVEL_ELIM || l == 0) timed(t_units, [&] { build_units(L.A, s); });
      if (L.kind == LIGMG_LEVEL_AGG) {
        timed(t_units, [&] { build_units(L.M, s); });
        setup_dinv(L.A, L, mixed, s);
        if (o.smoother_type == LIGMG_SMOOTHER_CHEBYSHEV) {
          timed(t_eig, [&] {
            L.lam = eig_estimate(L.A, L.dinv.p, o.seed, (int)l, o.cheby_eig_iters, L.sperm.p, s);
          });
          L.cheb_lo = o.cheby_lower * L.lam;
          L.cheb_hi = o.cheby_upper * L.lam;
          al_pre[l] = cheb_alpha(L.cheb_lo, L.cheb_hi, o.pre_smooth, be_pre[l]);
          al_post[l] = cheb_alpha(L.cheb_lo, L.cheb_hi, o.post_smooth, be_post[l]);
        }
      } else if (L.kind == LIGMG_LEVEL_ELIM) {
        timed(t_units, [&] {
          build_units(L.P, s);
          build_units(L.R, s);
        });
      }
      if (mixed) {
        make_v32(L.A, s, true);
        if (L.kind == LIGMG_LEVEL_AGG) set_dinv32_from_lap(L);
        if (L.kind == LIGMG_LEVEL_ELIM) make_elim32(L, s);
      }
      alloc_work<double>(l);
      if (mixed) alloc_work<float>(l);
    }
    if (o.verbose)
      fprintf(stderr, "[ligmg]   solve-prep of which: SpMV units %.3f s, eigen estimates %.3f s, "
              "reduction plans %.3f s\n", t_units, t_eig, t_red);
    phase("levels");
    // SpMV layout: operators applied in the solve get column slices when
    // their vector does not fit in L2 (tuned or fixed by o.spmv_slice)
    const size_t base = (o.top_elim && H[0].kind == LIGMG_LEVEL_ELIM) ? 1 : 0;
    for (size_t l = base; l < H.size(); l++) {
      DLevel &L = H[l];
      if (L.kind == LIGMG_LEVEL_ELIM) continue;
      if (L.kind == LIGMG_LEVEL_COARSE && !o.kcycle) continue;
      if (!mixed || l == base) tune_slices(L.A, 0, l);
      if (mixed) tune_slices(L.A, 1, l);
    }
    const int64_t n0 = H[0].A.n;
    b0.alloc(n0);
    sol.alloc(n0);
    tmp0.alloc(n0);
    int maxnb = 1;
    for (auto &r : red) maxnb = std::max(maxnb, r.nb);
    dsc.alloc((size_t)8 * maxnb);
    phase("slices");
    setup_tail();
    LG_HIP(hipStreamSynchronize(s));
    phase("tail");
  }

  // ---------------------------------------------------------- dump / load
  // An operator stored with solve-order rows (row i is setup row rperm[i])
  // and columns (column c is setup column cperm[c]) back in setup order,
  // rows sorted by column as setup left them.  Null perms: identity.
  pio::HostCsr to_setup(const DCsr &A, const int32_t *rperm, const int32_t *cperm, bool vals) {
    pio::HostCsr H;
    const int64_t m = A.n;
    H.m = m;
    H.ncols = A.ncols;
    auto prp = download(A.rp.p, m + 1, s);
    auto pci = download(A.ci.p, A.nnz, s);
    std::vector<double> pv = vals && A.v.p ? download(A.v.p, A.nnz, s) : std::vector<double>(A.nnz, 1.0);
    std::vector<int32_t> rp_ = rperm ? download(rperm, m, s) : std::vector<int32_t>();
    std::vector<int32_t> cp_ = cperm ? download(cperm, A.ncols, s) : std::vector<int32_t>();
    std::vector<int32_t> len(m);
    for (int64_t i = 0; i < m; i++) len[rperm ? rp_[i] : i] = prp[i + 1] - prp[i];
    H.rp.assign(m + 1, 0);
    for (int64_t i = 0; i < m; i++) H.rp[i + 1] = H.rp[i] + len[i];
    H.ci.resize(A.nnz);
    H.v.resize(A.nnz);
    std::vector<std::pair<int32_t, double>> row;
    for (int64_t i = 0; i < m; i++) {
      row.clear();
      for (int32_t k = prp[i]; k < prp[i + 1]; k++) row.push_back({cperm ? cp_[pci[k]] : pci[k], pv[k]});
      std::stable_sort(row.begin(), row.end(),
                       [](const std::pair<int32_t, double> &a, const std::pair<int32_t, double> &b) {
                         return a.first < b.first;
                       });
      int32_t d = H.rp[rperm ? rp_[i] : i];
      for (const auto &e : row) {
        H.ci[d] = e.first;
        H.v[d++] = e.second;
      }
    }
    return H;
  }

  static void write_host(const std::string &p, const pio::HostCsr &A) {
    pio::write_mat(p, A.m, A.ncols, A.rp.data(), A.ci.data(), A.v.data());
  }

  // dump_hierarchy (ligmg.hpp:47-52) with the level files of the dump
  // methods (elimination.hpp:62-67, aggregation.hpp:254-272,
  // multigrid.hpp:81-83); the aggregation states are the final ones.
  void dump(const std::string &dir) {
    if (mkdir(dir.c_str(), 0777) != 0 && errno != EEXIST)
      throw Error(pio::ERR_IO, "cannot create directory " + dir);
    const size_t nl = H.size();
    for (size_t l = 0; l < nl; l++) {
      const DLevel &L = H[l];
      const std::string pre = dir + "/" + std::to_string(l) + "_";
      const int32_t *sp = L.sperm.p, *nsp = (l + 1 < nl) ? H[l + 1].sperm.p : nullptr;
      const pio::HostCsr A = to_setup(L.A, sp, sp, true);
      write_host(pre + "A.petsc", A);
      const int64_t n = L.A.n;
      if (L.kind == LIGMG_LEVEL_ELIM) {
        write_host(pre + "P.petsc", to_setup(L.P, sp, nsp, true));
        write_host(pre + "R.petsc", to_setup(L.R, nsp, sp, true));
        const auto nid = download(L.new_id.p, n, s);
        auto qs = download(L.q.p, n, s);
        std::vector<double> q(n);
        if (sp) {
          const auto spv = download(sp, n, s);
          for (int64_t i = 0; i < n; i++) q[spv[i]] = qs[i];
        } else {
          q = qs;
        }
        std::vector<int32_t> qrp(n + 1, 0), qci;
        std::vector<double> qv;
        for (int64_t i = 0; i < n; i++) {
          if (nid[i] < 0) {
            qci.push_back((int32_t)i);
            qv.push_back(q[i]);
          }
          qrp[i + 1] = (int32_t)qci.size();
        }
        pio::write_mat(pre + "Q.petsc", n, n, qrp.data(), qci.data(), qv.data());
      } else if (L.kind == LIGMG_LEVEL_AGG) {
        const auto agg = download(L.agg_setup.p ? L.agg_setup.p : L.agg.p, n, s);
        std::vector<int32_t> prp(n + 1), rrp(L.nc + 1, 0), rci(n);
        for (int64_t i = 0; i <= n; i++) prp[i] = (int32_t)i;
        for (int64_t i = 0; i < n; i++) rrp[agg[i] + 1]++;
        for (int64_t c = 0; c < L.nc; c++) rrp[c + 1] += rrp[c];
        std::vector<int32_t> fill(rrp.begin(), rrp.end() - 1);
        for (int64_t i = 0; i < n; i++) rci[fill[agg[i]]++] = (int32_t)i;
        pio::write_mat(pre + "P.petsc", n, L.nc, prp.data(), agg.data(), nullptr);
        pio::write_mat(pre + "R.petsc", L.nc, n, rrp.data(), rci.data(), nullptr);
        const auto st = download(L.strength.p, L.A.nnz, s);
        pio::write_mat(pre + "strength.petsc", n, n, A.rp.data(), A.ci.data(), st.data());
        const auto ind = download(L.ind.p, n, s), state = download(L.state.p, n, s);
        std::vector<double> di(ind.begin(), ind.end()), ds(state.begin(), state.end());
        pio::write_vec(pre + "agg_ind.petsc", di.data(), n);
        pio::write_vec(pre + "agg_state.petsc", ds.data(), n);
      }
    }
    std::vector<double> b(H[0].blocks.begin(), H[0].blocks.end());
    pio::write_vec(dir + "/blocks.petsc", b.data(), (int64_t)b.size());
    if (perm_input.p) {
      const auto pi = download(perm_input.p, H[0].A.n, s);
      std::vector<double> d(pi.begin(), pi.end());
      pio::write_vec(dir + "/ordering.petsc", d.data(), (int64_t)d.size());
    }
  }

  static std::vector<int32_t> as_ints(const std::vector<double> &x, int64_t n, const std::string &what) {
    LG_CHECK((int64_t)x.size() == n, pio::ERR_IO, what + ": wrong length");
    std::vector<int32_t> r(n);
    for (int64_t i = 0; i < n; i++) {
      LG_CHECK(x[i] == (double)(int32_t)x[i], pio::ERR_IO, what + ": not an integer vector");
      r[i] = (int32_t)x[i];
    }
    return r;
  }

  void upload_csr(const pio::HostCsr &A, DCsr &D, bool vals) {
    D.n = A.m;
    D.ncols = A.ncols;
    D.nnz = (int64_t)A.ci.size();
    D.rp.alloc(A.m + 1);
    D.ci.alloc(std::max<int64_t>(D.nnz, 1));
    to_device(D.rp.p, A.rp.data(), A.m + 1, s);
    to_device(D.ci.p, A.ci.data(), D.nnz, s);
    if (vals) {
      D.v.alloc(std::max<int64_t>(D.nnz, 1));
      to_device(D.v.p, A.v.data(), D.nnz, s);
    }
  }

  // A solver from a dumped hierarchy: the levels are read back, not
  // constructed; the solve-phase set-up runs as after build.
  void load(const std::string &dir) {
    LG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    mixed = o.precision == LIGMG_PREC_MIXED;
    auto t0 = std::chrono::steady_clock::now();
    for (size_t l = 0;; l++) {
      const std::string pre = dir + "/" + std::to_string(l) + "_";
      if (!pio::exists(pre + "A.petsc")) {
        LG_CHECK(l > 0, pio::ERR_IO, "no hierarchy dump in " + dir);
        break;
      }
      const pio::HostCsr A = pio::read_mat(pre + "A.petsc");
      const int64_t n = A.m;
      LG_CHECK(n > 0 && A.ncols == n, pio::ERR_IO, pre + "A.petsc: not a square matrix");
      LG_CHECK(l == 0 || H.back().nc == n, pio::ERR_IO, pre + "A.petsc: size does not match the level above");
      LG_CHECK(H.empty() || H.back().kind != LIGMG_LEVEL_COARSE, pio::ERR_IO, "levels after the coarse level");
      DLevel L;
      L.n = n;
      upload_csr(A, L.A, true);
      validate_csr(n, n, L.A.rp.p, L.A.ci.p, s, true);
      if (pio::exists(pre + "Q.petsc")) {
        const pio::HostCsr P = pio::read_mat(pre + "P.petsc"), Q = pio::read_mat(pre + "Q.petsc");
        LG_CHECK(P.m == n && Q.m == n && Q.ncols == n, pio::ERR_IO, pre + "P/Q: wrong shape");
        std::vector<int32_t> nid(n);
        std::vector<double> q(n, 0.0);
        int64_t kept = 0;
        for (int64_t i = 0; i < n; i++) {
          const int32_t a = Q.rp[i], b = Q.rp[i + 1];
          LG_CHECK(b - a <= 1 && (b == a || Q.ci[a] == i), pio::ERR_IO, pre + "Q.petsc: not diagonal");
          if (b > a) {
            nid[i] = -1;
            q[i] = Q.v[a];
          } else {
            nid[i] = (int32_t)kept++;
            LG_CHECK(P.rp[i + 1] - P.rp[i] == 1 && P.ci[P.rp[i]] == nid[i] && P.v[P.rp[i]] == 1.0, pio::ERR_IO,
                     pre + "P.petsc: a kept row is not a unit row");
          }
        }
        LG_CHECK(P.ncols == kept, pio::ERR_IO, pre + "P.petsc: column count differs from the kept vertices");
        L.kind = LIGMG_LEVEL_ELIM;
        L.nc = kept;
        L.new_id.alloc(n);
        to_device(L.new_id.p, nid.data(), n, s);
        L.q.alloc(n);
        to_device(L.q.p, q.data(), n, s);
        upload_csr(P, L.P, true);
        L.nnzP = L.P.nnz;
        L.nnzQ = n - kept;
        transpose_p(L.P, kept, L.R, s);
      } else if (pio::exists(pre + "P.petsc")) {
        const pio::HostCsr P = pio::read_mat(pre + "P.petsc");
        LG_CHECK(P.m == n, pio::ERR_IO, pre + "P.petsc: wrong shape");
        std::vector<int32_t> agg(n);
        for (int64_t i = 0; i < n; i++) {
          LG_CHECK(P.rp[i + 1] - P.rp[i] == 1 && P.v[P.rp[i]] == 1.0, pio::ERR_IO,
                   pre + "P.petsc: not a piecewise-constant interpolation");
          agg[i] = P.ci[P.rp[i]];
        }
        const pio::HostCsr S = pio::read_mat(pre + "strength.petsc");
        LG_CHECK(S.m == n && S.rp == A.rp && S.ci == A.ci, pio::ERR_IO,
                 pre + "strength.petsc: pattern differs from A");
        const auto ind = as_ints(pio::read_vec(pre + "agg_ind.petsc"), n, pre + "agg_ind.petsc");
        const auto state = as_ints(pio::read_vec(pre + "agg_state.petsc"), n, pre + "agg_state.petsc");
        L.kind = LIGMG_LEVEL_AGG;
        L.nc = P.ncols;
        L.agg.alloc(n);
        L.ind.alloc(n);
        L.state.alloc(n);
        to_device(L.agg.p, agg.data(), n, s);
        to_device(L.ind.p, ind.data(), n, s);
        to_device(L.state.p, state.data(), n, s);
        L.strength.alloc(std::max<size_t>(S.v.size(), 1));
        to_device(L.strength.p, S.v.data(), S.v.size(), s);
        build_members(L.agg.p, n, L.nc, L.M, s);
      } else {
        L.kind = LIGMG_LEVEL_COARSE;
        coarse_pinv(n, A.rp, A.ci, A.v, L.pinv, s);
      }
      H.push_back(std::move(L));
    }
    LG_CHECK(H.back().kind == LIGMG_LEVEL_COARSE, pio::ERR_IO, "the dump does not end with a coarse level");
    const int64_t n0 = H[0].A.n;
    std::vector<int64_t> blocks = {0, n0};
    if (pio::exists(dir + "/blocks.petsc")) {
      const auto b = pio::read_vec(dir + "/blocks.petsc");
      blocks.assign(b.begin(), b.end());
      LG_CHECK(blocks.size() >= 2 && blocks.front() == 0 && blocks.back() == n0, pio::ERR_IO,
               "blocks.petsc must span [0, n]");
      for (size_t c = 0; c + 1 < blocks.size(); c++)
        LG_CHECK(blocks[c] < blocks[c + 1], pio::ERR_IO, "blocks.petsc: empty block");
    }
    if (pio::exists(dir + "/ordering.petsc")) {
      const auto pi = as_ints(pio::read_vec(dir + "/ordering.petsc"), n0, "ordering.petsc");
      perm_input.alloc(n0);
      to_device(perm_input.p, pi.data(), n0, s);
    }
    H[0].blocks = blocks;
    finish_setup();
    setup_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  // print_csv_header / print_csv (util.hpp:187-205, options.hpp:139-164):
  // the reference's columns; options this library has no field for are
  // printed with the reference's defaults.
  static std::string csv_header() {
    return "StrengthMetric,Redistribute,MaxLevels,MinSize,ElimDegree,AlwaysElim,TopElim,Tolerance,Threshold,"
           "AggregationIterations,MaxIterations,RedistributionFactor,RedistributionBound,PreSmoothing,"
           "PostSmoothing,ConstantCorrection,Smoother,NumberTVs,TVSmoothingIterations,KCycle,KCycleIterations,"
           "OuterKSPType,KCycleKSPType,VotesToSeed,MaxAggSize,AggIterFactor,ChebyLower,ChebyUpper,AffScaling,"
           "RandPerm,RejectionRatio,CRThreshold,CRTVs,CRIters,CRSmoothingIters,CRNorm,"
           "graph,solver,iters,rel_error,cycle_complexity,wda,setup_time,solve_time,nnz,nodes\n";
  }
  std::string csv_line(const std::string &graph, const ligmg_info &info) const {
    auto g = [](double x) {
      char b[64];
      snprintf(b, sizeof b, "%g", x);  // std::ostream's default for doubles
      return std::string(b);
    };
    auto ksp = [](int t) {
      static const char *names[] = {"cg", "fcg", "pipefcg", "fgmres"};
      return std::string(names[std::min(std::max(t, 0), 3)]);
    };
    static const char *aff[] = {"None", "combblas::Row", "Column", "RowColumn", "Sym", "Max"};
    static const char *sm[] = {"affinity", "algebraic-distance", "none"};
    std::string r;
    r += std::string(sm[std::min(std::max(o.strength_metric, 0), 2)]) + ",1," + std::to_string(o.max_levels) + ",";
    r += std::to_string(o.min_size) + "," + std::to_string(o.elim_degree) + "," +
         std::to_string(o.always_elim ? 1 : 0) + "," + std::to_string(o.top_elim ? 1 : 0) + ",";
    r += g(o.tol) + "," + g(o.thresh) + "," + std::to_string(o.agg_iters) + "," + std::to_string(o.iters) + ",";
    r += g(10) + "," + g(10000) + "," + std::to_string(o.pre_smooth) + "," + std::to_string(o.post_smooth) + ",";
    r += std::to_string(o.constant_correction ? 1 : 0) + "," +
         (o.smoother_type == LIGMG_SMOOTHER_JACOBI ? "jacobi" : "chebyshev") + ",";
    r += std::to_string(o.num_tvs) + "," + std::to_string(o.tvs_smooth) + "," + std::to_string(o.kcycle ? 1 : 0) +
         "," + std::to_string(o.kcycle_iters) + ",";
    r += ksp(o.outer_ksp_type) + "," + ksp(o.kcycle_ksp_type) + "," + std::to_string(o.votes_to_seed) + "," +
         std::to_string(o.max_agg_size) + ",";
    r += g(o.agg_iter_factor) + "," + g(o.cheby_lower) + "," + g(o.cheby_upper) + "," +
         aff[std::min(std::max(o.aff_scaling, 0), 5)] + ",";
    r += "1," + g(-1) + "," + g(0.1) + ",4,2,4,inf,";
    r += graph + ",ligmg," + std::to_string(info.iters) + "," + g(info.rel_residual) + "," +
         g(info.cycle_complexity) + "," + g(info.wda) + "," + g(setup_time) + "," + g(info.time) + ",";
    r += std::to_string(H[0].A.nnz) + ",1\n";
    return r;
  }

  // Solve layout (see DLevel::sperm): levels of at least SOLVE_PERM_MIN
  // vertices other than the coarsest get a degree permutation; the level
  // matrix keeps each row's entries in their order (same row sums), the
  // transfer operators are relabeled on both sides.
  static constexpr int64_t SOLVE_PERM_MIN = 1 << 16;
  void apply_solve_layout() {
    const size_t nl = H.size();
    for (size_t l = 0; l + 1 < nl; l++)
      if (H[l].A.n >= SOLVE_PERM_MIN) degree_perm(H[l].A, H[l].blocks, H[l].sperm, H[l].siperm, s);
    for (size_t l = 0; l + 1 < nl; l++) {
      DLevel &V = H[l];
      const DLevel &N = H[l + 1];
      const int32_t *rp = V.sperm.p, *nip = N.siperm.p;
      if (!rp && !nip) continue;
      if (V.kind == LIGMG_LEVEL_AGG) {
        DevBuf<int32_t> a(V.n);
        relabel_map(V.n, rp, V.agg.p, nip, a.p, s);
        V.agg_setup = std::move(V.agg);
        V.agg = std::move(a);
        build_members(V.agg.p, V.n, V.nc, V.M, s);
      } else if (V.kind == LIGMG_LEVEL_ELIM) {
        DCsr P;
        permute_rows(V.P, rp, nip, V.nc, P, s);
        V.P = std::move(P);
        if (rp) {
          DevBuf<double> q(V.n);
          gather_vec(V.n, rp, V.q.p, q.p, s);
          V.q = std::move(q);
          V.new_id_s.alloc(V.n);
          relabel_map(V.n, rp, V.new_id.p, nullptr, V.new_id_s.p, s);
        }
        transpose_p(V.P, V.nc, V.R, s);
      }
    }
    // aggregation levels (the smoothed matrices) get their columns sorted
    // in the new order; an elimination level keeps its rows' entry order,
    // which its P rows follow
    for (size_t l = 0; l + 1 < nl; l++) {
      DLevel &V = H[l];
      if (!V.sperm.p) continue;
      DCsr A;
      if (V.kind == LIGMG_LEVEL_AGG) permute_sorted(V.A, V.sperm.p, V.siperm.p, V.tpos.p, A, s);
      else permute_rows(V.A, V.sperm.p, V.siperm.p, V.A.ncols, A, s);
      V.A = std::move(A);
    }
    LG_HIP(hipStreamSynchronize(s));
  }

  // ------------------------------------------------------ fused coarse tail
  // A level joins the tail when its solve operators hold at most
  // o.tail_nnz entries; the tail is the longest run of such levels that
  // ends at the coarsest one.
  bool tail_ok(const DLevel &L) const {
    const int64_t lim = o.tail_nnz;
    auto unsliced = [](const DCsr &A) { return A.slices[0].size() <= 1 && A.slices[1].size() <= 1; };
    if (L.kind == LIGMG_LEVEL_ELIM) return L.P.nnz <= lim && L.R.nnz <= lim && L.n <= lim;
    if (L.kind == LIGMG_LEVEL_COARSE) return L.A.nnz <= lim && L.n * L.n <= 8 * lim && unsliced(L.A);
    return L.A.nnz <= lim && L.n <= lim && unsliced(L.A);
  }

  template <class T>
  static constexpr int pidx() {
    return std::is_same<T, float>::value ? 1 : 0;
  }
  template <class T>
  const TailOp<T> *tail_ops() const {
    if constexpr (std::is_same<T, double>::value) return tail_ops64.p;
    else return tail_ops32.p;
  }
  template <class T>
  const TailLevel<T> *tail_lv() const {
    if constexpr (std::is_same<T, double>::value) return tail_lv64.p;
    else return tail_lv32.p;
  }

  template <class T>
  TailOp<T> tail_op(const DCsr &A, const TailSched &S) const {
    TailOp<T> op;
    op.rp = A.rp.p;
    op.ci = A.ci.p;
    // the value / diagonal choice of dispatch_spmv
    if constexpr (std::is_same<T, double>::value) {
      op.v = A.v.p;
    } else {
      op.v = A.v32.p;
      op.d = A.d32.p;
    }
    op.trow = S.trow.p;
    op.grow = S.grow.p;
    for (int c = 0; c <= TAIL_GCLASSES; c++) op.gofs[c] = S.gofs[c];
    op.chunk = S.chunk.p;
    op.ntrow = S.ntrow;
    op.nchunk = S.nchunk;
    op.lrow = A.U.lrow.p;
    op.lpart = A.U.lpart.p;
    op.nlong = (int32_t)A.U.nlong;
    op.parts = A.U.parts.p;
    return op;
  }

  // ---- recording: the same operations, operands and buffers as cycle /
  // recurse / krylov / precond above, as instructions
  static TailInstr ti(int op, int64_t n = 0) {
    TailInstr I;
    I.op = op;
    I.i0 = (int32_t)n;
    return I;
  }
  template <class T>
  void rec_spmv(std::vector<TailInstr> &c, size_t l, int which, const T *x, int epi, const void *p1,
                const void *p2 = nullptr, const void *p3 = nullptr, const void *p4 = nullptr, double f0 = 0,
                double f1 = 0) {
    TailInstr I = ti(TI_SPMV, (int64_t)(4 * l + which));
    I.i1 = epi;
    I.p[0] = x;
    I.p[1] = p1;
    I.p[2] = p2;
    I.p[3] = p3;
    I.p[4] = p4;
    I.f[0] = f0;
    I.f[1] = f1;
    c.push_back(I);
  }
  template <class T>
  void rec_dots(std::vector<TailInstr> &c, size_t l, int np, const T *a0, const T *b0, const T *a1,
                const T *b1, int s0, int s1, int post) {
    TailInstr I = ti(TI_DOTS, np);
    I.lvl = (int32_t)l;
    I.i1 = s0;
    I.i2 = s1;
    I.i3 = post;
    I.p[0] = a0;
    I.p[1] = b0;
    I.p[2] = a1;
    I.p[3] = b1;
    c.push_back(I);
  }
  template <class T>
  void rec_cheb(std::vector<TailInstr> &c, size_t l, const T *dinv, const T *b, const T *xc, const T *xp,
                T *xn, double alpha, double beta) {
    rec_spmv<T>(c, l, 0, xc, TE_CHEB, dinv, b, xp, xn, alpha, beta);
  }
  template <class T>
  int rec_smooth_from_zero(std::vector<TailInstr> &c, size_t l, const T *dinv, const T *b) {
    LevelWork<T> &w = work<T>(l);
    const int64_t n = H[l].n;
    const int m = o.pre_smooth;
    const bool jac = o.smoother_type == LIGMG_SMOOTHER_JACOBI;
    if (m <= 0 || jac) {
      TailInstr I = ti(TI_ZERO, n);
      I.p[0] = w.B[0].p;
      c.push_back(I);
      if (!jac) return 0;
      rec_cheb<T>(c, l, dinv, b, w.B[0].p, nullptr, w.B[1].p, 0.0, 2.0 / 3.0);
      return 1;
    }
    TailInstr I = ti(TI_CHEB0, n);
    I.p[0] = dinv;
    I.p[1] = b;
    I.p[2] = w.B[1].p;
    I.f[0] = be_pre[l][0];
    c.push_back(I);
    for (int k = 1; k < m; k++) {
      const T *xp = (k == 1) ? nullptr : w.B[(k - 1) % 3].p;
      rec_cheb<T>(c, l, dinv, b, w.B[k % 3].p, xp, w.B[(k + 1) % 3].p, al_pre[l][k], be_pre[l][k]);
    }
    return m % 3;
  }
  template <class T>
  int rec_smooth_from(std::vector<TailInstr> &c, size_t l, const T *dinv, const T *b, int s0) {
    LevelWork<T> &w = work<T>(l);
    const int m = o.post_smooth;
    if (o.smoother_type == LIGMG_SMOOTHER_JACOBI) {
      const int s1 = (s0 + 1) % 3;
      rec_cheb<T>(c, l, dinv, b, w.B[s0].p, nullptr, w.B[s1].p, 0.0, 2.0 / 3.0);
      return s1;
    }
    if (m <= 0) return s0;
    for (int k = 0; k < m; k++) {
      const T *xp = (k == 0) ? nullptr : w.B[(s0 + k - 1) % 3].p;
      rec_cheb<T>(c, l, dinv, b, w.B[(s0 + k) % 3].p, xp, w.B[(s0 + k + 1) % 3].p, al_post[l][k],
                  be_post[l][k]);
    }
    return (s0 + m) % 3;
  }
  template <class T>
  const T *rec_cycle(std::vector<TailInstr> &c, size_t l, const T *b) {
    DLevel &L = H[l];
    LevelWork<T> &w = work<T>(l);
    if (L.kind == LIGMG_LEVEL_COARSE) {
      TailInstr I = ti(TI_DENSE, L.n);
      I.p[0] = L.pinv.p;
      I.p[1] = b;
      I.p[2] = w.B[0].p;
      c.push_back(I);
      return w.B[0].p;
    }
    LevelWork<T> &wn = work<T>(l + 1);
    if (L.kind == LIGMG_LEVEL_ELIM) {
      rec_spmv<T>(c, l, 2, b, TE_STORE, wn.b.p);
      const T *xc = rec_recurse<T>(c, l, wn.b.p);
      rec_spmv<T>(c, l, 1, xc, TE_INTERP, qv<T>(L), b, w.B[0].p);
      return w.B[0].p;
    }
    const T *dinv = dinvv<T>(L);
    int cur = rec_smooth_from_zero<T>(c, l, dinv, b);
    rec_spmv<T>(c, l, 0, w.B[cur].p, TE_RESID, b, w.R.p);
    rec_spmv<T>(c, l, 3, w.R.p, TE_STORE, wn.b.p);
    const T *xc = rec_recurse<T>(c, l, wn.b.p);
    TailInstr I = ti(TI_PROLONG, L.n);
    I.p[0] = L.agg.p;
    I.p[1] = xc;
    I.p[2] = w.B[cur].p;
    c.push_back(I);
    cur = rec_smooth_from<T>(c, l, dinv, b, cur);
    if (o.constant_correction) {
      TailInstr S = ti(TI_SCALE, L.n);
      S.p[0] = w.B[cur].p;
      S.f[0] = 4.0 / 3.0;
      c.push_back(S);
    }
    return w.B[cur].p;
  }
  template <class T>
  const T *rec_recurse(std::vector<TailInstr> &c, size_t l, const T *bc) {
    const DLevel &N = H[l + 1];
    if (!o.kcycle || N.kind == LIGMG_LEVEL_ELIM) return rec_cycle<T>(c, l + 1, bc);
    KWork<T> &K = work<T>(l + 1).kw;
    TailInstr I = ti(TI_ZERO, N.A.n);
    I.p[0] = K.x.p;
    c.push_back(I);
    rec_krylov<T>(c, l + 1, K.x.p, bc);
    return K.x.p;
  }
  template <class T>
  void rec_precond(std::vector<TailInstr> &c, size_t lvl, const T *r, T *z) {
    const int64_t n = H[lvl].A.n;
    const T *cy = rec_cycle<T>(c, lvl, r);
    TailInstr I = ti(TI_COPY, n);
    I.p[0] = z;
    I.p[1] = cy;
    c.push_back(I);
    rec_dots<T>(c, lvl, 1, z, nullptr, nullptr, nullptr, -1, -1, TP_NONE);
    TailInstr M = ti(TI_SUBMEAN, n);
    M.lvl = (int32_t)lvl;
    M.p[0] = z;
    c.push_back(M);
  }
  template <class T>
  void rec_krylov(std::vector<TailInstr> &c, size_t lvl, T *x, const T *b) {
    KWork<T> &K = work<T>(lvl).kw;
    const int64_t n = H[lvl].A.n;
    const bool fcg = fcg_like(o.kcycle_ksp_type);
    TailInstr I = ti(TI_KR_INIT);
    I.lvl = (int32_t)lvl;
    c.push_back(I);
    rec_spmv<T>(c, lvl, 0, x, TE_RESID, b, K.r.p);
    rec_dots<T>(c, lvl, 2, b, b, K.r.p, K.r.p, SC_BN2, SC_RN2, TP_CHECK);
    rec_precond<T>(c, lvl, K.r.p, K.z.p);
    TailInstr C = ti(TI_COPY, n);
    C.p[0] = K.p.p;
    C.p[1] = K.z.p;
    c.push_back(C);
    rec_dots<T>(c, lvl, 1, K.r.p, K.z.p, nullptr, nullptr, SC_GAMMA, SC_GAMMA, TP_NONE);
    int it = 0;
    while (it < o.kcycle_iters) {
      rec_spmv<T>(c, lvl, 0, K.p.p, TE_STORE, K.q.p);
      if (fcg) rec_dots<T>(c, lvl, 2, K.p.p, K.q.p, K.p.p, K.r.p, SC_DELTA, SC_T1, TP_ALPHA);
      else rec_dots<T>(c, lvl, 1, K.p.p, K.q.p, nullptr, nullptr, SC_DELTA, SC_DELTA, TP_ALPHA);
      TailInstr U = ti(TI_UPDATE);
      U.lvl = (int32_t)lvl;
      U.p[0] = x;
      c.push_back(U);
      it++;
      rec_dots<T>(c, lvl, 1, K.r.p, K.r.p, nullptr, nullptr, SC_RN2, SC_RN2, TP_CHECK);
      if (it >= o.kcycle_iters) break;
      rec_precond<T>(c, lvl, K.r.p, K.z.p);
      if (fcg) rec_dots<T>(c, lvl, 1, K.z.p, K.q.p, nullptr, nullptr, SC_T1, SC_T1, TP_BETA);
      else rec_dots<T>(c, lvl, 1, K.r.p, K.z.p, nullptr, nullptr, SC_GN, SC_GN, TP_BETA);
      TailInstr B = ti(TI_PUPDATE);
      B.lvl = (int32_t)lvl;
      c.push_back(B);
    }
  }

  template <class T>
  void make_tail(DevBuf<TailOp<T>> &ops, DevBuf<TailLevel<T>> &lv) {
    const size_t nl = H.size();
    std::vector<TailOp<T>> ho(4 * nl);
    std::vector<TailLevel<T>> hl(nl);
    for (size_t l = tail_l0; l < nl; l++) {
      const DLevel &L = H[l];
      if (L.kind == LIGMG_LEVEL_ELIM) {
        ho[4 * l + 1] = tail_op<T>(L.P, tail_sched[4 * l + 1]);
        ho[4 * l + 2] = tail_op<T>(L.R, tail_sched[4 * l + 2]);
      } else {
        ho[4 * l] = tail_op<T>(L.A, tail_sched[4 * l]);
      }
      if (L.kind == LIGMG_LEVEL_AGG) ho[4 * l + 3] = tail_op<T>(L.M, tail_sched[4 * l + 3]);
      const RedPlan &R = red[l];
      const KWork<T> &K = work<T>(l).kw;
      TailLevel<T> &t = hl[l];
      t.n = (int32_t)L.A.n;
      t.nb = R.nb;
      t.nseg = R.nseg;
      t.seg_lo = R.seg_lo.p;
      t.seg_hi = R.seg_hi.p;
      t.bptr = R.bptr.p;
      t.bseg = R.bseg.p;
      t.part = R.part.p;
      t.sc = K.sc.p;
      t.bsum = dsc.p;
      t.kr = K.r.p;
      t.kz = K.z.p;
      t.kp = K.p.p;
      t.kq = K.q.p;
    }
    ops.alloc(4 * nl);
    lv.alloc(nl);
    to_device(ops.p, ho.data(), ho.size(), s);
    to_device(lv.p, hl.data(), hl.size(), s);
    // one program per entry level: the correction recurse(e) computes
    auto &progs = tail_prog[pidx<T>()];
    progs.clear();
    progs.resize(nl);
    for (size_t e = tail_l0 - 1; e + 1 < nl; e++) {
      std::vector<TailInstr> code;
      const T *out = rec_recurse<T>(code, e, work<T>(e + 1).b.p);
      TailProg &P = progs[e];
      P.code.alloc(code.size());
      to_device(P.code.p, code.data(), code.size(), s);
      P.len = (int64_t)code.size();
      P.in = work<T>(e + 1).b.p;
      P.out = out;
    }
  }

  // instructions of the correction below level e (growing with
  // kcycle_iters^depth for a K-cycle), without recording it
  int64_t tail_count(size_t e) const {
    const size_t nl = H.size();
    std::vector<int64_t> cyc(nl + 1, 0), rec(nl + 1, 0);
    const int m = o.kcycle_iters;
    for (size_t l = nl; l-- > e + 1;) {
      const DLevel &L = H[l];
      const int pre = o.pre_smooth, post = o.post_smooth;
      if (L.kind == LIGMG_LEVEL_COARSE) cyc[l] = 1;
      else if (L.kind == LIGMG_LEVEL_ELIM) cyc[l] = 2 + rec[l];
      else cyc[l] = 3 + std::max(pre, 1) + std::max(post, 1) + 1 + rec[l];
      // recurse(l - 1)
      const bool k = o.kcycle && L.kind != LIGMG_LEVEL_ELIM;
      rec[l - 1] = k ? 1 + 6 + (int64_t)m * (6 + cyc[l] + 3) : cyc[l];
    }
    return rec[e];
  }

  void setup_tail() {
    const size_t nl = H.size();
    tail_l0 = nl;
    if (o.tail_nnz <= 0) return;
    const size_t base = (o.top_elim && H[0].kind == LIGMG_LEVEL_ELIM) ? 1 : 0;
    size_t l0 = nl;
    while (l0 > base + 1 && tail_ok(H[l0 - 1])) l0--;
    // bounded program size (a deep K-cycle tail grows geometrically)
    while (l0 < nl && tail_count(l0 - 1) > TAIL_MAX_INSTR) l0++;
    if (l0 >= nl) return;
    tail_l0 = l0;
    tail_sched.clear();
    tail_sched.resize(4 * nl);
    for (size_t l = l0; l < nl; l++) {
      const DLevel &L = H[l];
      if (L.kind == LIGMG_LEVEL_ELIM) {
        tail_schedule(L.P, tail_sched[4 * l + 1], s);
        tail_schedule(L.R, tail_sched[4 * l + 2], s);
        continue;
      }
      tail_schedule(L.A, tail_sched[4 * l], s);
      if (L.kind == LIGMG_LEVEL_AGG) tail_schedule(L.M, tail_sched[4 * l + 3], s);
    }
    make_tail<double>(tail_ops64, tail_lv64);
    if (mixed) make_tail<float>(tail_ops32, tail_lv32);
    tail_p.fcg = fcg_like(o.kcycle_ksp_type);
    tail_p.rtol = o.tol;
    LG_HIP(hipStreamSynchronize(s));
    if (o.verbose)
      fprintf(stderr, "[ligmg] fused tail: levels %zu-%zu, %lld instructions per correction\n", l0, nl - 1,
              (long long)tail_prog[0][l0 - 1].len);
  }

  void tune_slices(DCsr &A, int prec, size_t l) {
    const int64_t es = prec ? 4 : 8;
    if (o.spmv_slice == 0) return;
    if (o.spmv_slice > 0) {
      if (A.ncols > o.spmv_slice) build_slices(A, prec, o.spmv_slice, s);
      return;
    }
    if (A.n * es <= ((int64_t)2 << 20)) return;
    double best_t = time_spmv(A, prec, 3, s);
    int64_t best_w = 0;
    for (int64_t mb : {1, 2, 4}) {
      int64_t w = (mb << 20) / es;
      if (w >= A.ncols) continue;
      build_slices(A, prec, w, s);
      double t = time_spmv(A, prec, 3, s);
      if (t < best_t) {
        best_t = t;
        best_w = w;
      }
    }
    build_slices(A, prec, best_w > 0 ? best_w : A.ncols, s);
    if (o.verbose)
      fprintf(stderr, "[ligmg] level %zu fp%d spmv: slice width %lld (%.3f ms)\n", l, prec ? 32 : 64,
              (long long)best_w, best_t);
  }

  // the FP32 smoother uses the diagonal of the FP32 Laplacian-form operator
  void set_dinv32_from_lap(DLevel &L);

  template <class T>
  void alloc_work(size_t l) {
    const DLevel &L = H[l];
    LevelWork<T> &w = work<T>(l);
    const int64_t n = L.A.n;
    for (auto &B : w.B) B.alloc(std::max<int64_t>(n, 1));
    if (L.kind == LIGMG_LEVEL_AGG) w.R.alloc(n);
    w.b.alloc(std::max<int64_t>(n, 1));
    bool need_kw = o.kcycle && l > 0 && L.kind != LIGMG_LEVEL_ELIM;
    if (need_kw) w.kw.alloc(n, red[l].nb);
  }

  // ---------------------------------------------------------------- cycles
  // Applies level l's cycle to b with a zero initial guess and returns the
  // buffer holding the result.  The buffer sequence depends only on the
  // options, so replays (and graph captures) see the same pointers.
  template <class T>
  const T *cycle(size_t l, const T *b) {
    DLevel &L = H[l];
    LevelWork<T> &w = work<T>(l);
    if (L.kind == LIGMG_LEVEL_COARSE) {
      op_dense_apply<T>(L.pinv.p, L.n, b, w.B[0].p, s);
      return w.B[0].p;
    }
    if (L.kind == LIGMG_LEVEL_ELIM) {
      LevelWork<T> &wn = work<T>(l + 1);
      op_spmv<T>(L.R, b, wn.b.p, s);
      const T *xc = recurse<T>(l, wn.b.p);
      op_interp<T>(L.P, qv<T>(L), xc, b, w.B[0].p, s);
      return w.B[0].p;
    }
    // aggregation level
    const T *dinv = dinvv<T>(L);
    int cur = smooth_from_zero<T>(L, w, dinv, b, o.pre_smooth, l);
    LevelWork<T> &wn = work<T>(l + 1);
    op_residual<T>(L.A, w.B[cur].p, b, w.R.p, s);
    op_spmv<T>(L.M, w.R.p, wn.b.p, s);
    const T *xc = recurse<T>(l, wn.b.p);
    op_prolong_add<T>(L.agg.p, L.n, xc, w.B[cur].p, s);
    cur = smooth_from<T>(L, w, dinv, b, cur, o.post_smooth, l);
    if (o.constant_correction) scale_vec<T>(w.B[cur].p, L.n, 4.0 / 3.0);
    return w.B[cur].p;
  }

  template <class T>
  int smooth_from_zero(DLevel &L, LevelWork<T> &w, const T *dinv, const T *b, int m, size_t l) {
    // the reference's Jacobi smoother is one sweep whatever the step count
    // (create_smoother, smoothing.hpp:206-211)
    if (o.smoother_type == LIGMG_SMOOTHER_JACOBI) {
      fill_zero<T>(w.B[0].p, L.n);
      return jacobi<T>(L, w, dinv, b, 0);
    }
    if (m <= 0) {
      fill_zero<T>(w.B[0].p, L.n);
      return 0;
    }
    // x_1 = beta_0 D^-1 b in B[1]; x_{k+1} in B[(k+1) % 3]
    op_cheb_step<T>(L.A, dinv, b, nullptr, nullptr, w.B[1].p, 0.0, be_pre[l][0], s);
    for (int k = 1; k < m; k++) {
      const T *xp = (k == 1) ? nullptr : w.B[(k - 1) % 3].p;
      op_cheb_step<T>(L.A, dinv, b, w.B[k % 3].p, xp, w.B[(k + 1) % 3].p, al_pre[l][k], be_pre[l][k],
                      s);
    }
    return m % 3;
  }

  template <class T>
  int smooth_from(DLevel &L, LevelWork<T> &w, const T *dinv, const T *b, int s0, int m, size_t l) {
    if (o.smoother_type == LIGMG_SMOOTHER_JACOBI) return jacobi<T>(L, w, dinv, b, s0);
    if (m <= 0) return s0;
    for (int k = 0; k < m; k++) {
      const T *xp = (k == 0) ? nullptr : w.B[(s0 + k - 1) % 3].p;
      op_cheb_step<T>(L.A, dinv, b, w.B[(s0 + k) % 3].p, xp, w.B[(s0 + k + 1) % 3].p, al_post[l][k],
                      be_post[l][k], s);
    }
    return (s0 + m) % 3;
  }

  // reference jacobi smoother: one damped sweep (smoothing.hpp:78-98, 206-211)
  template <class T>
  int jacobi(DLevel &L, LevelWork<T> &w, const T *dinv, const T *b, int s0) {
    int s1 = (s0 + 1) % 3;
    op_cheb_step<T>(L.A, dinv, b, w.B[s0].p, nullptr, w.B[s1].p, 0.0, 2.0 / 3.0, s);
    return s1;
  }

  template <class T>
  const T *recurse(size_t l, const T *bc) {
    const DLevel &N = H[l + 1];
    if (l + 1 >= tail_l0) {
      const TailProg &P = tail_prog[pidx<T>()][l];
      LG_CHECK(P.in == bc, LIGMG_ERR_INVALID, "fused tail entered with an unrecorded right-hand side");
      tail_launch<T>(P.code.p, P.len, tail_ops<T>(), tail_lv<T>(), tail_p, s);
      return (const T *)P.out;
    }
    if (!o.kcycle || N.kind == LIGMG_LEVEL_ELIM) return cycle<T>(l + 1, bc);
    // K-cycle: kcycle_iters steps of FCG on the next level (multigrid.hpp:113-126)
    LevelWork<T> &wn = work<T>(l + 1);
    KWork<T> &K = wn.kw;
    fill_zero<T>(K.x.p, N.A.n);
    krylov<T>(l + 1, N.A, K.x.p, bc, K, o.kcycle_ksp_type, o.tol, o.kcycle_iters, false, nullptr,
              nullptr);
    return K.x.p;
  }

  // ---------------------------------------------------------------- Krylov
  // CG / FCG on level lvl's matrix with per-block coefficients.  With
  // host_check the residual norms come back every iteration (outer loop,
  // history, stopping test); otherwise exactly maxit iterations run with
  // device-side convergence flags (K-cycle inner solves, no host sync).
  template <class T>
  int krylov(size_t lvl, const DCsr &A, T *x, const T *b, KWork<T> &K, int type, double rtol,
             int maxit, bool host_check, std::vector<double> *hist, int *status,
             bool outer_prec = false) {
    const RedPlan &R = red[lvl];
    const int nb = R.nb;
    const int64_t n = A.n;
    const bool fcg = fcg_like(type);
    hipLaunchKernelGGL(k_kr_init, dim3((nb + 63) / 64), dim3(64), 0, s, nb, K.sc.p);
    // r = b - A x
    op_residual<T>(A, x, b, K.r.p, s);
    dot2<T>(R, b, b, K.r.p, K.r.p, K, SC_BN2, SC_RN2, KP_CHECK, 0, rtol);
    std::vector<double> hsc;
    if (status) *status = LIGMG_CONVERGED;
    if (host_check) {
      hsc = download(K.sc.p, (size_t)nb * SC_N, s);
      if (hist) hist->push_back(total_norm(hsc, nb));
      if (!any_active(hsc, nb)) return 0;
    }
    apply_prec<T>(lvl, K.r.p, K.z.p, outer_prec);
    d2d(K.p.p, K.z.p, n, s);
    dot1<T>(R, K.r.p, K.z.p, K, SC_GAMMA);
    int it = 0;
    while (it < maxit) {
      op_spmv<T>(A, K.p.p, K.q.p, s);
      if (fcg) dot2<T>(R, K.p.p, K.q.p, K.p.p, K.r.p, K, SC_DELTA, SC_T1, KP_ALPHA, 1);
      else dot1<T>(R, K.p.p, K.q.p, K, SC_DELTA, KP_ALPHA, 0);
      hipLaunchKernelGGL(k_kr_update<T>, dim3(R.nseg), dim3(256), 0, s, x, K.r.p, K.p.p, K.q.p,
                         R.seg_lo.p, R.seg_hi.p, R.bseg.p + R.nb + 1, K.sc.p);
      LG_LAUNCH_CHECK();
      it++;
      dot1<T>(R, K.r.p, K.r.p, K, SC_RN2, KP_CHECK, 0, rtol);
      if (host_check) {
        hsc = download(K.sc.p, (size_t)nb * SC_N, s);
        double tn = total_norm(hsc, nb);
        if (hist) hist->push_back(tn);
        if (tn != tn) {
          if (status) *status = LIGMG_DIVERGED_NAN;
          return it;
        }
        if (!any_active(hsc, nb)) return it;
      }
      if (it >= maxit) break;
      apply_prec<T>(lvl, K.r.p, K.z.p, outer_prec);
      if (fcg) dot1<T>(R, K.z.p, K.q.p, K, SC_T1, KP_BETA, 1);
      else dot1<T>(R, K.r.p, K.z.p, K, SC_GN, KP_BETA, 0);
      hipLaunchKernelGGL(k_kr_pupdate<T>, dim3(R.nseg), dim3(256), 0, s, K.p.p, K.z.p, R.seg_lo.p,
                         R.seg_hi.p, R.bseg.p + R.nb + 1, K.sc.p);
      LG_LAUNCH_CHECK();
    }
    if (host_check && status) *status = LIGMG_DIVERGED_ITS;
    return it;
  }

  template <class T>
  void apply_prec(size_t lvl, const T *r, T *z, bool outer_prec) {
    if constexpr (std::is_same<T, double>::value) {
      if (outer_prec) {
        precond_outer(lvl, r, z);
        return;
      }
    }
    precond<T>(lvl, r, z);
  }

  int krylov_outer(size_t base, const double *rhs, std::vector<double> &h, int *status) {
    if (o.outer_ksp_type == LIGMG_KSP_FGMRES) return fgmres(base, outer.x.p, rhs, &h, status);
    return krylov<double>(base, H[base].A, outer.x.p, rhs, outer, o.outer_ksp_type, o.tol, o.iters,
                          true, &h, status, true);
  }

  // FGMRES(m) outer solve: restarted flexible GMRES with PETSc KSPFGMRES's
  // defaults (restart 30, classical Gram-Schmidt, right preconditioning,
  // unpreconditioned residual norm).  The Arnoldi process, Givens rotations
  // and convergence test are per block; the Hessenberg matrices live on the
  // host, the basis V (m + 1 vectors) and the preconditioned basis Z (m) on
  // the device.  The history logs the true residual at each cycle start and
  // the Givens estimate after each iteration, as PETSc's KSPGMRESCycle.
  static constexpr int FGMRES_RESTART = 30;
  DevBuf<double> gm_V, gm_Z, gm_coef;

  int fgmres(size_t lvl, double *x, const double *b, std::vector<double> *hist, int *status) {
    const RedPlan &R = red[lvl];
    const DCsr &A = H[lvl].A;
    const int nb = R.nb, m = FGMRES_RESTART;
    const int64_t n = A.n;
    const double rtol = o.tol;
    const int maxit = o.iters;
    gm_V.ensure((size_t)(m + 1) * n);
    gm_Z.ensure((size_t)m * n);
    gm_coef.ensure((size_t)nb * (m + 1));
    KWork<double> &K = outer;
    const int32_t *segblk = R.bseg.p + R.nb + 1;
    auto V = [&](int j) { return gm_V.p + (size_t)j * n; };
    auto Z = [&](int j) { return gm_Z.p + (size_t)j * n; };
    auto dots = [&](int np, const double *const *a, const double *const *bb) {
      R.dots<double>(np, a, bb, K.out.p, s);
      return download(K.out.p, (size_t)nb * np, s);
    };
    auto upload = [&](const std::vector<double> &c) {
      to_device(gm_coef.p, c.data(), c.size(), s);
      LG_HIP(hipStreamSynchronize(s));
    };
    auto lincomb = [&](double *y, const double *base, int k, const std::vector<double> &coef, int kk) {
      upload(coef);
      hipLaunchKernelGGL(k_blk_lincomb, dim3(R.nseg), dim3(256), 0, s, y, base, n, k, (const double *)gm_coef.p,
                         kk, R.seg_lo.p, R.seg_hi.p, segblk);
      LG_LAUNCH_CHECK();
    };
    auto scale = [&](double *y, const std::vector<double> &sc) {
      upload(sc);
      hipLaunchKernelGGL(k_blk_scale, dim3(R.nseg), dim3(256), 0, s, y, (const double *)y,
                         (const double *)gm_coef.p, R.seg_lo.p, R.seg_hi.p, segblk);
      LG_LAUNCH_CHECK();
    };
    auto none = [](const std::vector<char> &f) { return !std::any_of(f.begin(), f.end(), [](char a) { return a; }); };
    const double *bb1[1] = {b};
    const std::vector<double> bn = dots(1, bb1, bb1);
    if (status) *status = LIGMG_CONVERGED;
    int it = 0;
    std::vector<char> active(nb, 1);
    while (true) {
      op_residual<double>(A, x, b, V(0), s);
      const double *v0[1] = {V(0)};
      const std::vector<double> rn = dots(1, v0, v0);
      double t = 0;
      bool nan = false;
      for (int c = 0; c < nb; c++) {
        t += rn[c];
        nan |= rn[c] != rn[c];
        if (!(std::sqrt(rn[c]) > rtol * std::sqrt(bn[c]))) active[c] = 0;
      }
      if (hist) hist->push_back(std::sqrt(t));
      if (nan) {
        if (status) *status = LIGMG_DIVERGED_NAN;
        return it;
      }
      if (none(active)) return it;
      if (it >= maxit) {
        if (status) *status = LIGMG_DIVERGED_ITS;
        return it;
      }
      std::vector<double> Hh((size_t)nb * (m + 1) * m, 0.0), g((size_t)nb * (m + 1), 0.0), cs((size_t)nb * m, 0.0),
          sn((size_t)nb * m, 0.0), est(nb), sc(nb);
      std::vector<int> cols(nb, 0);
      std::vector<char> live = active;
      for (int c = 0; c < nb; c++) {
        const double beta = std::sqrt(rn[c]);
        est[c] = active[c] ? beta : 0.0;
        g[(size_t)c * (m + 1)] = beta;
        sc[c] = (live[c] && beta > 0) ? 1.0 / beta : 0.0;
      }
      scale(V(0), sc);
      bool done = false;
      for (int j = 0; j < m; j++) {
        d2d(K.r.p, V(j), n, s);
        apply_prec<double>(lvl, K.r.p, K.z.p, true);
        d2d(Z(j), K.z.p, n, s);
        op_spmv<double>(A, Z(j), V(j + 1), s);
        // classical Gram-Schmidt: h_i = <w, v_i>, i <= j, four at a time
        std::vector<double> h((size_t)(j + 1) * nb);
        for (int i0 = 0; i0 <= j; i0 += 4) {
          const int np = std::min(4, j + 1 - i0);
          const double *a[4] = {V(j + 1), V(j + 1), V(j + 1), V(j + 1)};
          const double *vv[4] = {nullptr, nullptr, nullptr, nullptr};
          for (int q = 0; q < np; q++) vv[q] = V(i0 + q);
          const std::vector<double> out = dots(np, a, vv);
          for (int q = 0; q < np; q++)
            for (int c = 0; c < nb; c++) h[(size_t)(i0 + q) * nb + c] = out[(size_t)c * np + q];
        }
        std::vector<double> coef((size_t)nb * (j + 1), 0.0);
        for (int c = 0; c < nb; c++)
          if (live[c])
            for (int i = 0; i <= j; i++) coef[(size_t)c * (j + 1) + i] = -h[(size_t)i * nb + c];
        lincomb(V(j + 1), V(0), j + 1, coef, j + 1);
        const double *w1[1] = {V(j + 1)};
        const std::vector<double> hn = dots(1, w1, w1);
        for (int c = 0; c < nb; c++) {
          const double nrm = std::sqrt(hn[c]);
          sc[c] = (live[c] && nrm > 0) ? 1.0 / nrm : 0.0;
        }
        scale(V(j + 1), sc);
        for (int c = 0; c < nb; c++) {
          if (!live[c]) continue;
          const double nrm = std::sqrt(hn[c]);
          double *col = Hh.data() + ((size_t)c * m + j) * (m + 1);  // column j, rows 0..j+1
          double *gc = g.data() + (size_t)c * (m + 1);
          double *csc = cs.data() + (size_t)c * m, *snc = sn.data() + (size_t)c * m;
          for (int i = 0; i <= j; i++) col[i] = h[(size_t)i * nb + c];
          col[j + 1] = nrm;
          for (int i = 0; i < j; i++) {
            const double a = col[i], bq = col[i + 1];
            col[i] = csc[i] * a + snc[i] * bq;
            col[i + 1] = -snc[i] * a + csc[i] * bq;
          }
          const double a = col[j], bq = col[j + 1];
          const double d = std::hypot(a, bq);
          csc[j] = d > 0 ? a / d : 1.0;
          snc[j] = d > 0 ? bq / d : 0.0;
          col[j] = d;
          col[j + 1] = 0.0;
          gc[j + 1] = -snc[j] * gc[j];
          gc[j] = csc[j] * gc[j];
          cols[c] = j + 1;
          est[c] = std::fabs(gc[j + 1]);
          if (!(est[c] > rtol * std::sqrt(bn[c]))) {
            live[c] = 0;
            active[c] = 0;
          } else if (nrm == 0) {
            live[c] = 0;
          }
        }
        it++;
        double t2 = 0;
        for (int c = 0; c < nb; c++) t2 += est[c] * est[c];
        done = none(active) || it >= maxit;
        if (j + 1 == m && !done) break;
        if (hist) hist->push_back(std::sqrt(t2));
        if (done || none(live)) break;
      }
      // x += Z y with H y = g, per block
      int kmax = 0;
      for (int c = 0; c < nb; c++) kmax = std::max(kmax, cols[c]);
      std::vector<double> y((size_t)nb * m, 0.0);
      for (int c = 0; c < nb; c++) {
        const int k = cols[c];
        const double *Hc = Hh.data() + (size_t)c * m * (m + 1);
        const double *gc = g.data() + (size_t)c * (m + 1);
        double *yc = y.data() + (size_t)c * m;
        for (int i = k - 1; i >= 0; i--) {
          double sacc = gc[i];
          for (int q = i + 1; q < k; q++) sacc -= Hc[(size_t)q * (m + 1) + i] * yc[q];
          const double d = Hc[(size_t)i * (m + 1) + i];
          yc[i] = d != 0 ? sacc / d : 0.0;
        }
      }
      if (kmax > 0) lincomb(x, Z(0), kmax, y, m);
      if (done) {
        if (status) *status = none(active) ? LIGMG_CONVERGED : LIGMG_DIVERGED_ITS;
        return it;
      }
    }
  }

  static double total_norm(const std::vector<double> &hsc, int nb) {
    double t = 0;
    for (int c = 0; c < nb; c++) t += hsc[c * SC_N + SC_RN2];
    return std::sqrt(t);
  }
  static bool any_active(const std::vector<double> &hsc, int nb) {
    for (int c = 0; c < nb; c++)
      if (hsc[c * SC_N + SC_ACTIVE] != 0) return true;
    return false;
  }

  template <class T>
  // dot products into the scalar slots, followed in the same launch by the
  // scalar step post (KrPost)
  void dot1(const RedPlan &R, const T *a, const T *b, KWork<T> &K, int slot, int post = KP_NONE, int fcg = 0,
            double rtol = 0) {
    const T *A_[1] = {a}, *B_[1] = {b};
    R.dots_sc<T>(1, A_, B_, K.sc.p, slot, slot, post, fcg, rtol, s);
  }
  template <class T>
  void dot2(const RedPlan &R, const T *a0, const T *b0, const T *a1, const T *b1, KWork<T> &K, int s0,
            int s1, int post = KP_NONE, int fcg = 0, double rtol = 0) {
    const T *A_[2] = {a0, a1}, *B_[2] = {b0, b1};
    R.dots_sc<T>(2, A_, B_, K.sc.p, s0, s1, post, fcg, rtol, s);
  }

  // z = M r with the constant removed per block (PETSc removes the attached
  // nullspace from preconditioned vectors; util.hpp:530-534)
  template <class T>
  void precond(size_t lvl, const T *r, T *z) {
    const int64_t n = H[lvl].A.n;
    const T *c = cycle<T>(lvl, r);
    d2d(z, c, n, s);
    const RedPlan &R = red[lvl];
    R.sums<T>(z, dsc.p, s);
    op_sub_block_mean<T>(z, R, dsc.p, s);
  }

  // outer preconditioner in FP64, cycle in FP32 when mixed.  With
  // o.use_graphs the whole application (hundreds of launches on the coarse
  // levels) is captured once per (level, r, z) and replayed as a hipGraph;
  // it contains no host synchronisation and no memset nodes.
  void precond_outer(size_t lvl, const double *r, double *z) {
    if (o.use_graphs) {
      for (auto &g : graphs)
        if (g.first.lvl == lvl && g.first.r == r && g.first.z == z) {
          LG_HIP(hipGraphLaunch(g.second, s));
          return;
        }
      hipGraph_t graph;
      {
        PoolCaptureScope cap;
        LG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
          precond_outer_eager(lvl, r, z);
        } catch (...) {
          (void)hipStreamEndCapture(s, &graph);
          throw;
        }
        LG_HIP(hipStreamEndCapture(s, &graph));
      }
      hipGraphExec_t exec;
      LG_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      (void)hipGraphDestroy(graph);
      graphs.push_back({{lvl, r, z}, exec});
      LG_HIP(hipGraphLaunch(exec, s));
      return;
    }
    precond_outer_eager(lvl, r, z);
  }

  void precond_outer_eager(size_t lvl, const double *r, double *z) {
    const int64_t n = H[lvl].A.n;
    if (mixed) {
      LevelWork<float> &w = W32[lvl];
      op_d2f(r, w.b.p, n, s);
      const float *c = cycle<float>(lvl, w.b.p);
      op_f2d(c, z, n, s);
    } else {
      const double *c = cycle<double>(lvl, r);
      d2d(z, c, n, s);
    }
    const RedPlan &R = red[lvl];
    R.sums<double>(z, dsc.p, s);
    op_sub_block_mean<double>(z, R, dsc.p, s);
  }

  template <class T>
  void fill_zero(T *p, int64_t n);
  template <class T>
  void scale_vec(T *p, int64_t n, double a);
  template <class T>
  const T *dinvv(const DLevel &L);
  template <class T>
  const T *qv(const DLevel &L);

  // ---------------------------------------------------------------- batch
  template <class T>
  BLevelWork<T> &bwork(size_t l);

  template <class T>
  void balloc(int kp) {
    auto &W = [&]() -> std::vector<BLevelWork<T>> & {
      if constexpr (std::is_same<T, double>::value) return BW64;
      else return BW32;
    }();
    W.clear();
    W.resize(H.size());
    for (size_t l = 0; l < H.size(); l++) {
      const DLevel &L = H[l];
      const int64_t n = std::max<int64_t>(L.A.n, 1);
      for (auto &B : W[l].B) B.alloc(n * kp);
      if (L.kind == LIGMG_LEVEL_AGG) W[l].R.alloc(n * kp);
      W[l].b.alloc(n * kp);
      if (o.kcycle && l > 0 && L.kind != LIGMG_LEVEL_ELIM) W[l].kw.alloc(n * kp, red[l].nb * kp);
    }
  }

  void batch_buffers(int kp) {
    if (bkp == kp) return;
    balloc<double>(kp);
    if (mixed) balloc<float>(kp);
    // scratch sized once for every operator, so no buffer is reallocated
    // while kernels that use it are queued
    int64_t maxparts = 1, maxseg = 1;
    for (size_t l = 0; l < H.size(); l++) {
      const DLevel &L = H[l];
      for (const DCsr *M : {&L.A, &L.M, &L.P, &L.R}) maxparts = std::max(maxparts, M->U.nparts);
      maxseg = std::max<int64_t>(maxseg, red[l].nseg);
    }
    bparts.alloc((size_t)maxparts * kp);
    bpart_red.alloc((size_t)maxseg * 2 * kp);
    bkp = kp;
    if (o.spmm_tune) {
      // every operator a batched solve applies: the FP64 ones (outer loop,
      // FP64 cycles) and, in mixed precision, their FP32 forms
      for (size_t l = 0; l < H.size(); l++) {
        DLevel &L = H[l];
        int t64[4] = {-1, -1, -1, -1}, t32[4] = {-1, -1, -1, -1}, i = 0;
        for (DCsr *M : {&L.A, &L.M, &L.P, &L.R}) {
          if (M->U.nunits) {
            t64[i] = tune_spmm<double>(*M, kp, bparts, s);
            if (mixed) t32[i] = tune_spmm<float>(*M, kp, bparts, s);
          }
          i++;
        }
        if (o.verbose) {
          std::string line;
          const char *nm[4] = {"A", "M", "P", "R"};
          for (int j = 0; j < 4; j++) {
            int vw, u;
            if (t64[j] >= 0) {
              spmm_tile(t64[j], &vw, &u);
              line += std::string(" ") + nm[j] + " fp64 " + std::to_string(vw) + "x" + std::to_string(u);
            }
            if (t32[j] >= 0) {
              spmm_tile(t32[j], &vw, &u);
              line += std::string(" ") + nm[j] + " fp32 " + std::to_string(vw) + "x" + std::to_string(u);
            }
          }
          if (!line.empty())
            fprintf(stderr, "[ligmg] spmm tiles k=%d level %zu (rhs per lane x gathers in flight):%s\n", kp, l,
                    line.c_str());
        }
      }
    }
  }

  // block version of cycle(): same buffer discipline
  template <class T>
  const T *bcycle(size_t l, const T *b, int kp) {
    DLevel &L = H[l];
    BLevelWork<T> &w = bwork<T>(l);
    if (L.kind == LIGMG_LEVEL_COARSE) {
      op_bdense_apply<T>(L.pinv.p, L.n, b, w.B[0].p, kp, s);
      return w.B[0].p;
    }
    if (L.kind == LIGMG_LEVEL_ELIM) {
      BLevelWork<T> &wn = bwork<T>(l + 1);
      op_bspmv<T>(L.R, b, wn.b.p, kp, bparts, s);
      const T *xc = brecurse<T>(l, wn.b.p, kp);
      op_binterp<T>(L.P, qv<T>(L), xc, b, w.B[0].p, kp, bparts, s);
      return w.B[0].p;
    }
    const T *dinv = dinvv<T>(L);
    int cur;
    const int m = o.pre_smooth;
    const bool jac = o.smoother_type == LIGMG_SMOOTHER_JACOBI;  // one sweep whatever m
    if (jac || m <= 0) {
      fill_zero<T>(w.B[0].p, L.n * kp);
      cur = 0;
      if (jac) {
        op_bcheb_step<T>(L.A, dinv, b, w.B[0].p, nullptr, w.B[1].p, 0.0, 2.0 / 3.0, kp, bparts, s);
        cur = 1;
      }
    } else {
      op_bcheb_step<T>(L.A, dinv, b, nullptr, nullptr, w.B[1].p, 0.0, be_pre[l][0], kp, bparts, s);
      for (int k = 1; k < m; k++) {
        const T *xp = (k == 1) ? nullptr : w.B[(k - 1) % 3].p;
        op_bcheb_step<T>(L.A, dinv, b, w.B[k % 3].p, xp, w.B[(k + 1) % 3].p, al_pre[l][k],
                         be_pre[l][k], kp, bparts, s);
      }
      cur = m % 3;
    }
    BLevelWork<T> &wn = bwork<T>(l + 1);
    op_bresidual<T>(L.A, w.B[cur].p, b, w.R.p, kp, bparts, s);
    op_bspmv<T>(L.M, w.R.p, wn.b.p, kp, bparts, s);
    const T *xc = brecurse<T>(l, wn.b.p, kp);
    op_bprolong_add<T>(L.agg.p, L.n, xc, w.B[cur].p, kp, s);
    const int mp = o.post_smooth;
    if (o.smoother_type == LIGMG_SMOOTHER_JACOBI) {
      int s1 = (cur + 1) % 3;
      op_bcheb_step<T>(L.A, dinv, b, w.B[cur].p, nullptr, w.B[s1].p, 0.0, 2.0 / 3.0, kp, bparts, s);
      cur = s1;
    } else {
      for (int k = 0; k < mp; k++) {
        const T *xp = (k == 0) ? nullptr : w.B[(cur + k - 1) % 3].p;
        op_bcheb_step<T>(L.A, dinv, b, w.B[(cur + k) % 3].p, xp, w.B[(cur + k + 1) % 3].p,
                         al_post[l][k], be_post[l][k], kp, bparts, s);
      }
      cur = (cur + mp) % 3;
    }
    if (o.constant_correction) scale_vec<T>(w.B[cur].p, L.n * kp, 4.0 / 3.0);
    return w.B[cur].p;
  }

  template <class T>
  const T *brecurse(size_t l, const T *bc, int kp) {
    const DLevel &N = H[l + 1];
    if (!o.kcycle || N.kind == LIGMG_LEVEL_ELIM) return bcycle<T>(l + 1, bc, kp);
    BLevelWork<T> &wn = bwork<T>(l + 1);
    KWork<T> &K = wn.kw;
    fill_zero<T>(K.x.p, N.A.n * kp);
    bkrylov<T>(l + 1, N.A, K.x.p, bc, K, kp, o.kcycle_ksp_type, o.tol, o.kcycle_iters, false, nullptr,
               nullptr, false);
    return K.x.p;
  }

  // z = (cycle applied to r) minus its block means.  r32_ready: the FP32
  // copy of r for the mixed outer preconditioner is already in place.  On
  // that path, with w set, the segment partials of z.w are also left in
  // bpart_red (returns true); otherwise returns false.
  template <class T>
  bool bprec(size_t lvl, const T *r, T *z, int kp, bool outer_prec, bool r32_ready = false,
             const T *w = nullptr) {
    const int64_t n = H[lvl].A.n;
    if constexpr (std::is_same<T, double>::value) {
      if (outer_prec && mixed) {
        // the block sums come from the FP32 result itself (the FP64 copy has
        // the same values), and one pass converts and subtracts the means
        BLevelWork<float> &bw = BW32[lvl];
        if (!r32_ready) op_d2f(r, bw.b.p, n * kp, s);
        const float *c = bcycle<float>(lvl, bw.b.p, kp);
        const RedPlan &R = red[lvl];
        R.bdots<float>(1, kp, c, nullptr, nullptr, nullptr, bsc_out.p, bpart_red, s);
        op_bsub_block_mean_f2d(z, c, kp, R, bsc_out.p, s, w, &bpart_red);
        return w != nullptr;
      } else {
        d2d(z, bcycle<T>(lvl, r, kp), n * kp, s);
      }
    } else {
      d2d(z, bcycle<T>(lvl, r, kp), n * kp, s);
    }
    const RedPlan &R = red[lvl];
    R.bdots<T>(1, kp, z, nullptr, nullptr, nullptr, bsc_out.p, bpart_red, s);
    op_bsub_block_mean<T>(z, kp, R, bsc_out.p, s);
    return false;
  }

  template <class T>
  void bdot(const RedPlan &R, int kp, const T *a0, const T *b0, const T *a1, const T *b1,
            KWork<T> &K, int s0, int s1) {
    const int np = a1 ? 2 : 1;
    R.bdots<T>(np, kp, a0, b0, a1, b1, K.out.p, bpart_red, s);
    hipLaunchKernelGGL(k_kr_bstore, dim3((R.nb * kp + 63) / 64), dim3(64), 0, s, R.nb, kp, np, K.out.p,
                       s0, s1, K.sc.p);
  }

  // per-(block, column) CG / FCG on block vectors; iters_col (host, may be
  // null) receives for every scalar set the iteration it converged at
  template <class T>
  int bkrylov(size_t lvl, const DCsr &A, T *x, const T *b, KWork<T> &K, int kp, int type,
              double rtol, int maxit, bool host_check, std::vector<int> *iters_set, int *status,
              bool outer_prec) {
    const RedPlan &R = red[lvl];
    const int ns = R.nb * kp, lk = ilog2(kp);
    const bool fcg = fcg_like(type);
    auto grid1 = dim3((ns + 63) / 64);
    hipLaunchKernelGGL(k_kr_init, grid1, dim3(64), 0, s, ns, K.sc.p);
    op_bresidual<T>(A, x, b, K.r.p, kp, bparts, s);
    bdot<T>(R, kp, b, b, K.r.p, K.r.p, K, SC_BN2, SC_RN2);
    hipLaunchKernelGGL(k_kr_check, grid1, dim3(64), 0, s, ns, rtol, K.sc.p);
    if (status) *status = LIGMG_CONVERGED;
    std::vector<double> hsc;
    std::vector<char> was_active(ns, 0);
    if (iters_set) iters_set->assign(ns, 0);
    if (host_check) {
      hsc = download(K.sc.p, (size_t)ns * SC_N, s);
      for (int t = 0; t < ns; t++) was_active[t] = hsc[t * SC_N + SC_ACTIVE] != 0;
      if (!any_active(hsc, ns)) return 0;
    }
    const int64_t nk = A.n * kp;
    bprec<T>(lvl, K.r.p, K.z.p, kp, outer_prec);
    d2d(K.p.p, K.z.p, nk, s);
    bdot<T>(R, kp, K.r.p, K.z.p, (const T *)nullptr, (const T *)nullptr, K, SC_GAMMA, SC_GAMMA);
    int it = 0;
    while (it < maxit) {
      op_bspmv<T>(A, K.p.p, K.q.p, kp, bparts, s);
      if (fcg) bdot<T>(R, kp, K.p.p, K.q.p, K.p.p, K.r.p, K, SC_DELTA, SC_T1);
      else bdot<T>(R, kp, K.p.p, K.q.p, (const T *)nullptr, (const T *)nullptr, K, SC_DELTA, SC_DELTA);
      hipLaunchKernelGGL(k_kr_alpha, grid1, dim3(64), 0, s, ns, fcg ? 1 : 0, K.sc.p);
      // the update also forms r.r's segment partials and, for the mixed
      // outer preconditioner, the FP32 copy of r
      float *r32 = nullptr;
      if constexpr (std::is_same<T, double>::value)
        if (outer_prec && mixed) r32 = BW32[lvl].b.p;
      LG_CHECK(bpart_red.n >= (size_t)R.nseg * kp, LIGMG_ERR_INVALID, "block reduction buffer too small");
      hipLaunchKernelGGL(k_kr_bupdate<T>, dim3(R.nseg), dim3(256), 0, s, x, K.r.p, K.p.p, K.q.p, kp, lk,
                         R.seg_lo.p, R.seg_hi.p, R.bseg.p + R.nb + 1, K.sc.p, bpart_red.p, r32);
      LG_LAUNCH_CHECK();
      it++;
      R.bdots_finish(1, kp, K.out.p, bpart_red, s);
      hipLaunchKernelGGL(k_kr_bstore, dim3((R.nb * kp + 63) / 64), dim3(64), 0, s, R.nb, kp, 1, K.out.p, SC_RN2,
                         SC_RN2, K.sc.p);
      hipLaunchKernelGGL(k_kr_check, grid1, dim3(64), 0, s, ns, rtol, K.sc.p);
      if (host_check) {
        hsc = download(K.sc.p, (size_t)ns * SC_N, s);
        bool nan = false;
        for (int t = 0; t < ns; t++) {
          double r2 = hsc[t * SC_N + SC_RN2];
          if (r2 != r2) nan = true;
          const bool act = hsc[t * SC_N + SC_ACTIVE] != 0;
          if (iters_set && (was_active[t] || act)) (*iters_set)[t] = it;
          was_active[t] = act;
        }
        if (nan) {
          if (status) *status = LIGMG_DIVERGED_NAN;
          return it;
        }
        if (!any_active(hsc, ns)) return it;
      }
      if (it >= maxit) break;
      // z.q (FCG) or r.z (CG): formed by the preconditioner's last pass
      // where it can, else a block dot
      if (bprec<T>(lvl, K.r.p, K.z.p, kp, outer_prec, true, fcg ? K.q.p : K.r.p)) {
        R.bdots_finish(1, kp, K.out.p, bpart_red, s);
        hipLaunchKernelGGL(k_kr_bstore, dim3((R.nb * kp + 63) / 64), dim3(64), 0, s, R.nb, kp, 1, K.out.p,
                           fcg ? SC_T1 : SC_GN, fcg ? SC_T1 : SC_GN, K.sc.p);
      } else if (fcg) {
        bdot<T>(R, kp, K.z.p, K.q.p, (const T *)nullptr, (const T *)nullptr, K, SC_T1, SC_T1);
      } else {
        bdot<T>(R, kp, K.r.p, K.z.p, (const T *)nullptr, (const T *)nullptr, K, SC_GN, SC_GN);
      }
      hipLaunchKernelGGL(k_kr_beta, grid1, dim3(64), 0, s, ns, fcg ? 1 : 0, K.sc.p);
      hipLaunchKernelGGL(k_kr_bpupdate<T>, dim3(R.nseg), dim3(256), 0, s, K.p.p, K.z.p, kp, lk,
                         R.seg_lo.p, R.seg_hi.p, R.bseg.p + R.nb + 1, K.sc.p);
      LG_LAUNCH_CHECK();
    }
    if (host_check && status) *status = LIGMG_DIVERGED_ITS;
    return it;
  }

  int solve_batch(int k, const double *B, double *X, int mem, ligmg_info *info, int32_t *iters_out);

  // ---------------------------------------------------------------- solve
  int solve(const double *b, double *x, int mem, ligmg_info *info, double *hist, int cap, int from = 0);
};

template <>
BLevelWork<double> &Solver::bwork<double>(size_t l) {
  return BW64[l];
}
template <>
BLevelWork<float> &Solver::bwork<float>(size_t l) {
  return BW32[l];
}

int Solver::solve_batch(int k, const double *Bin, double *Xout, int mem, ligmg_info *info,
                        int32_t *iters_out) {
  LG_CHECK(k >= 1 && k <= 64, LIGMG_ERR_INVALID, "solve_batch: k must be in [1, 64]");
  LG_CHECK(o.outer_ksp_type != LIGMG_KSP_FGMRES, LIGMG_ERR_UNSUPPORTED,
           "solve_batch: batched solves use CG / FCG / PIPEFCG");
  int kp = 1;
  while (kp < k) kp <<= 1;
  const DLevel &L0 = H[0];
  const int64_t n = L0.A.n;
  batch_buffers(kp);
  const bool do_once = o.top_elim && L0.kind == LIGMG_LEVEL_ELIM;
  const size_t base = do_once ? 1 : 0;
  const int64_t nbse = H[base].A.n;
  int maxnb = 1;
  for (auto &r : red) maxnb = std::max(maxnb, r.nb);
  bsc_out.ensure((size_t)maxnb * kp * 2);
  if (bouter.x.n < (size_t)(nbse * kp) || bouter.sc.n < (size_t)(red[base].nb * kp * SC_N))
    bouter.alloc(nbse * kp, red[base].nb * kp);
  DevBuf<double> bin(n * k), b0b(n * kp), rhs(nbse * kp), solb(n * kp), tmp(n * kp);
  if (mem == LIGMG_MEM_HOST) to_device(bin.p, Bin, n * k, s);
  else d2d(bin.p, Bin, n * k, s);
  LG_HIP(hipStreamSynchronize(s));
  auto t0 = std::chrono::steady_clock::now();
  op_to_block(bin.p, n, k, kp, b0b.p, perm.p, s);
  if (do_once) op_bspmv<double>(L0.R, b0b.p, rhs.p, kp, bparts, s);
  else d2d(rhs.p, b0b.p, n * kp, s);
  red[base].bdots<double>(1, kp, rhs.p, nullptr, nullptr, nullptr, bsc_out.p, bpart_red, s);
  op_bsub_block_mean<double>(rhs.p, kp, red[base], bsc_out.p, s);
  fill_zero<double>(bouter.x.p, nbse * kp);
  std::vector<int> its;
  int status = 0;
  int iters = bkrylov<double>(base, H[base].A, bouter.x.p, rhs.p, bouter, kp, o.outer_ksp_type, o.tol,
                              o.iters, true, &its, &status, true);
  if (do_once) op_binterp<double>(L0.P, L0.q.p, bouter.x.p, b0b.p, solb.p, kp, bparts, s);
  else d2d(solb.p, bouter.x.p, n * kp, s);
  op_from_block(solb.p, n, k, kp, bin.p, perm.p, s);
  LG_HIP(hipStreamSynchronize(s));
  auto t1 = std::chrono::steady_clock::now();
  // full residual per column (worst column reported)
  op_bresidual<double>(L0.A, solb.p, b0b.p, tmp.p, kp, bparts, s);
  red[0].bdots<double>(2, kp, tmp.p, tmp.p, b0b.p, b0b.p, bsc_out.p, bpart_red, s);
  std::vector<double> d = download(bsc_out.p, (size_t)red[0].nb * 2 * kp, s);
  double worst = 0;
  for (int c = 0; c < k; c++) {
    double rr = 0, bb = 0;
    for (int blk = 0; blk < red[0].nb; blk++) {
      rr += d[((size_t)blk * 2) * kp + c];
      bb += d[((size_t)blk * 2 + 1) * kp + c];
    }
    if (bb > 0) worst = std::max(worst, std::sqrt(rr / bb));
  }
  if (mem == LIGMG_MEM_HOST) to_host(Xout, bin.p, n * k, s);
  else d2d(Xout, bin.p, n * k, s);
  LG_HIP(hipStreamSynchronize(s));
  if (iters_out) {
    const int nb = red[base].nb;
    for (int c = 0; c < k; c++) {
      int m = 0;
      for (int blk = 0; blk < nb; blk++) m = std::max(m, its[(size_t)blk * kp + c]);
      iters_out[c] = m;
    }
  }
  if (info) {
    memset(info, 0, sizeof(*info));
    info->iters = iters;
    info->status = status;
    info->base_level = (int)base;
    info->full_rel_residual = worst;
    info->time = std::chrono::duration<double>(t1 - t0).count();
    info->setup_time = setup_time;
    // worst reduced-system relative residual over (block, column) sets
    std::vector<double> hsc = download(bouter.sc.p, (size_t)red[base].nb * kp * SC_N, s);
    double rel = 0;
    for (int t = 0; t < red[base].nb * kp; t++) {
      double bn = hsc[t * SC_N + SC_BN2];
      if (bn > 0) rel = std::max(rel, std::sqrt(hsc[t * SC_N + SC_RN2] / bn));
    }
    info->rel_residual = rel;
  }
  return LIGMG_OK;
}

template <>
LevelWork<double> &Solver::work<double>(size_t l) {
  return W64[l];
}
template <>
LevelWork<float> &Solver::work<float>(size_t l) {
  return W32[l];
}
template <>
const double *Solver::dinvv<double>(const DLevel &L) {
  return L.dinv.p;
}
template <>
const float *Solver::dinvv<float>(const DLevel &L) {
  return L.dinv32.p;
}
template <>
const double *Solver::qv<double>(const DLevel &L) {
  return L.q.p;
}
template <>
const float *Solver::qv<float>(const DLevel &L) {
  return L.q32.p;
}

namespace {
__global__ void k_dinv_from_d(int64_t n, const double *d, float *dinv32) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dinv32[i] = d[i] != 0 ? (float)(1.0 / d[i]) : 0.0f;
}
template <class T>
__global__ void k_scale(T *p, int64_t n, T a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] *= a;
}
template <class T>
__global__ void k_zero(T *p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (T)0;
}
}  // namespace

void Solver::set_dinv32_from_lap(DLevel &L) {
  if (!L.dinv32.p) L.dinv32.alloc(std::max<int64_t>(L.n, 1));
  hipLaunchKernelGGL(k_dinv_from_d, dim3(grid_for(L.n, 256, 8192)), dim3(256), 0, s, L.n, L.A.d32.p,
                     L.dinv32.p);
  LG_LAUNCH_CHECK();
}

template <class T>
void Solver::fill_zero(T *p, int64_t n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero<T>, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, p, n);
  LG_LAUNCH_CHECK();
}
template <class T>
void Solver::scale_vec(T *p, int64_t n, double a) {
  hipLaunchKernelGGL(k_scale<T>, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, p, n, (T)a);
  LG_LAUNCH_CHECK();
}

// solve (multigrid.hpp:315-370) of the system of level `from`: b and x have
// H[from].n entries, in the caller's order for level 0 and in the level's
// setup order (the order of its introspection) below it.  An elimination
// level `from` is applied once (R b, outer Krylov on from + 1, x = P x + Q b);
// the full residual is taken on H[from].A.
int Solver::solve(const double *b, double *x, int mem, ligmg_info *info, double *hist, int cap, int from) {
  LG_CHECK(from >= 0 && from < (int)H.size(), LIGMG_ERR_INVALID,
           "ligmg_solve_from: level " + std::to_string(from) + " out of range (" + std::to_string(H.size()) +
               " levels)");
  DLevel &L0 = H[from];
  const int64_t n = L0.A.n;
  // the matrix of an elimination level below the top is not applied by the
  // cycles, so it has no SpMV units until a solve starts from it (full
  // residual, or the outer Krylov solve when top_elim is off)
  if (L0.A.U.nunits == 0 && L0.A.nnz > 0) build_units(L0.A, s);
  // caller order -> solve layout: the input permutation and level 0's
  // layout for level 0, the level's own layout below it
  const int32_t *ip = from == 0 ? perm.p : L0.sperm.p;
  if (ip) {
    const double *src = b;
    if (mem == LIGMG_MEM_HOST) {
      to_device(tmp0.p, b, n, s);
      src = tmp0.p;
    }
    gather_vec(n, ip, src, b0.p, s);
  } else if (mem == LIGMG_MEM_HOST) {
    to_device(b0.p, b, n, s);
  } else {
    d2d(b0.p, b, n, s);
  }
  LG_HIP(hipStreamSynchronize(s));
  auto t0 = std::chrono::steady_clock::now();
  const bool do_once = o.top_elim && L0.kind == LIGMG_LEVEL_ELIM;
  const size_t base = do_once ? (size_t)from + 1 : (size_t)from;
  const DLevel &B = H[base];
  const int64_t nb_ = B.A.n;
  if (outer.x.n < (size_t)nb_ || outer.sc.n < (size_t)red[base].nb * SC_N) outer.alloc(nb_, red[base].nb);
  rhs_base.ensure(std::max<int64_t>(nb_, 1));
  double *rhs = rhs_base.p;
  if (do_once) op_spmv<double>(L0.R, b0.p, rhs, s);
  else d2d(rhs, b0.p, n, s);
  red[base].sums<double>(rhs, dsc.p, s);
  op_sub_block_mean<double>(rhs, red[base], dsc.p, s);
  fill_zero<double>(outer.x.p, nb_);
  std::vector<double> h;
  int status = 0;
  // the outer loop always runs in FP64 (custom preconditioner bridge)
  int iters = krylov_outer(base, rhs, h, &status);
  if (do_once) op_interp<double>(L0.P, L0.q.p, outer.x.p, b0.p, sol.p, s);
  else d2d(sol.p, outer.x.p, n, s);
  LG_HIP(hipStreamSynchronize(s));
  auto t1 = std::chrono::steady_clock::now();
  // full-system residual (multigrid.hpp:357-366)
  op_residual<double>(L0.A, sol.p, b0.p, tmp0.p, s);
  const double *A_[2] = {tmp0.p, b0.p}, *B_[2] = {tmp0.p, b0.p};
  RedPlan &R0 = red[from];
  R0.dots<double>(2, A_, B_, dsc.p, s);
  std::vector<double> d = download(dsc.p, 2 * R0.nb, s);
  double rr = 0, bb = 0;
  for (int c = 0; c < R0.nb; c++) {
    rr += d[c * 2];
    bb += d[c * 2 + 1];
  }
  if (ip) {
    if (mem == LIGMG_MEM_HOST) {
      scatter_vec(n, ip, sol.p, tmp0.p, s);
      to_host(x, tmp0.p, n, s);
    } else {
      scatter_vec(n, ip, sol.p, x, s);
    }
  } else if (mem == LIGMG_MEM_HOST) {
    to_host(x, sol.p, n, s);
  } else {
    d2d(x, sol.p, n, s);
  }
  LG_HIP(hipStreamSynchronize(s));
  if (info) {
    // cycle_complexity (multigrid.hpp:372-387): the work of the levels from
    // the base level down, over the nnz of level 0 whatever `from` is;
    // ConvergenceInfo (util.hpp:143-170)
    double sum = 0, mult = 1;
    for (size_t l = base; l < H.size(); l++) {
      const DLevel &L = H[l];
      double wk;
      if (L.kind == LIGMG_LEVEL_AGG)
        wk = (double)((o.pre_smooth + o.post_smooth + 1) * L.A.nnz + 3 * L.A.n);
      else if (L.kind == LIGMG_LEVEL_ELIM)
        wk = (double)(2 * L.nnzP + L.nnzQ);
      else
        wk = (double)L.A.nnz;
      sum += mult * wk;
      if (L.kind != LIGMG_LEVEL_ELIM && o.kcycle) mult *= o.kcycle_iters;
    }
    double cc = sum / (double)H[0].A.nnz;
    memset(info, 0, sizeof(*info));
    info->iters = iters;
    info->status = status;
    info->base_level = (int)base;
    info->history_len = (int)std::min<size_t>(h.size(), cap > 0 ? (size_t)cap : 0);
    info->rel_residual = h.empty() || h.front() == 0 ? 0 : h.back() / h.front();
    info->full_rel_residual = bb > 0 ? std::sqrt(rr / bb) : 0;
    info->time = std::chrono::duration<double>(t1 - t0).count();
    info->work = iters * (1 + cc);
    info->cycle_complexity = iters > 0 ? info->work / iters : 0;
    info->cf = iters > 0 ? std::pow(info->rel_residual, 1.0 / iters) : 0;
    info->ecf = iters > 0 ? std::pow(info->rel_residual, 1.0 / (iters * info->cycle_complexity)) : 0;
    info->wda = iters > 0 ? -1.0 / std::log10(info->ecf) : 0;
    info->setup_time = setup_time;
  }
  if (hist)
    for (int i = 0; i < (int)h.size() && i < cap; i++) hist[i] = h[i];
  return LIGMG_OK;
}

}  // namespace lg

// ------------------------------------------------------------------ C ABI
struct ligmg_solver {
  lg::Solver S;
};

namespace {
int fail(const lg::Error &e) {
  lg::g_last_error = e.what();
  return e.code;
}
int fail_unknown(const char *what) {
  lg::g_last_error = what;
  return LIGMG_ERR_INVALID;
}
}  // namespace

extern "C" {

void ligmg_options_default(ligmg_options *o) {
  memset(o, 0, sizeof(*o));
  o->verbose = 0;
  o->always_elim = 1;
  o->top_elim = 1;
  o->constant_correction = 0;
  o->strength_metric = LIGMG_STRENGTH_AFFINITY;
  o->max_levels = 50;
  o->min_size = 4000;
  o->elim_degree = 4;
  o->tol = 1e-8;
  o->thresh = 1e-8;
  o->agg_iters = 10;
  o->iters = 500;
  o->pre_smooth = 2;
  o->post_smooth = 2;
  o->num_tvs = 10;
  o->tvs_smooth = 4;
  o->smoother_type = LIGMG_SMOOTHER_CHEBYSHEV;
  o->kcycle = 0;
  o->outer_ksp_type = LIGMG_KSP_CG;
  o->kcycle_ksp_type = LIGMG_KSP_FCG;
  o->kcycle_iters = 2;
  o->votes_to_seed = 1;
  o->max_agg_size = -1;
  o->agg_iter_factor = 0.5;
  o->agg_init_factor = 0.5;
  o->cheby_lower = 0.3;
  o->cheby_upper = 1.1;
  o->aff_scaling = LIGMG_AFF_MAX;
  o->elim_factor = 0.9;
  o->elim_iters = 1;
  o->neighbor_of_neighbor = 0;
  o->seed = 0;
  o->cheby_eig_iters = 10;
  o->precision = LIGMG_PREC_FP64;
  o->use_graphs = 1;
  o->spmv_slice = 0;
  o->ordering = LIGMG_ORDER_NATURAL;
  o->solve_layout = 1;
  o->tail_nnz = 4096;
  o->rejection_ratio = -1;
  o->rejection_to_seed = 0;
  o->spmm_tune = 1;
}

const char *ligmg_last_error(void) { return lg::g_last_error.c_str(); }

static int check_options(const ligmg_options &o) {
  if (o.elim_iters != 1) return LIGMG_ERR_UNSUPPORTED;
  if (o.outer_ksp_type < LIGMG_KSP_CG || o.outer_ksp_type > LIGMG_KSP_FGMRES) return LIGMG_ERR_INVALID;
  if (o.kcycle_ksp_type < LIGMG_KSP_CG || o.kcycle_ksp_type > LIGMG_KSP_FGMRES) return LIGMG_ERR_INVALID;
  // inner K-cycle solves run on the device without host steps
  if (o.kcycle_ksp_type == LIGMG_KSP_FGMRES) return LIGMG_ERR_UNSUPPORTED;
  if (o.num_tvs < 1 || o.num_tvs > 16) return LIGMG_ERR_UNSUPPORTED;
  if (o.aff_scaling < LIGMG_AFF_NONE || o.aff_scaling > LIGMG_AFF_MAX) return LIGMG_ERR_INVALID;
  if (o.pre_smooth < 0 || o.post_smooth < 0 || o.iters < 0 || o.kcycle_iters < 1 || o.tail_nnz < 0)
    return LIGMG_ERR_INVALID;
  return LIGMG_OK;
}

int ligmg_create(int64_t n, const int32_t *rowptr, const int32_t *colind, const double *vals,
                 int mem_kind, const ligmg_options *opts, int32_t ncomp, const int64_t *comp_ptr,
                 ligmg_handle *out) {
  lg::g_last_error.clear();
  if (!out || n <= 0 || n >= INT32_MAX || !rowptr || !colind || !vals) {
    lg::g_last_error = "ligmg_create: invalid arguments";
    return LIGMG_ERR_INVALID;
  }
  *out = nullptr;
  try {
    std::unique_ptr<ligmg_solver> h(new ligmg_solver());
    if (opts) h->S.o = *opts;
    else ligmg_options_default(&h->S.o);
    int rc = check_options(h->S.o);
    if (rc) {
      lg::g_last_error = "ligmg_create: unsupported option value";
      return rc;
    }
    std::vector<int64_t> blocks;
    if (ncomp > 0 && comp_ptr) {
      blocks.assign(comp_ptr, comp_ptr + ncomp + 1);
      LG_CHECK(blocks.front() == 0 && blocks.back() == n, LIGMG_ERR_INVALID, "comp_ptr must span [0, n]");
      for (int c = 0; c < ncomp; c++)
        LG_CHECK(blocks[c] < blocks[c + 1], LIGMG_ERR_INVALID, "comp_ptr blocks must be non-empty");
    } else {
      blocks = {0, n};
    }
    h->S.build(n, rowptr, colind, vals, mem_kind, blocks);
    *out = h.release();
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  } catch (const std::exception &e) {
    return fail_unknown(e.what());
  }
}

int ligmg_load_hierarchy(const char *dir, const ligmg_options *opts, ligmg_handle *out) {
  lg::g_last_error.clear();
  if (!out || !dir) {
    lg::g_last_error = "ligmg_load_hierarchy: invalid arguments";
    return LIGMG_ERR_INVALID;
  }
  *out = nullptr;
  try {
    std::unique_ptr<ligmg_solver> h(new ligmg_solver());
    if (opts) h->S.o = *opts;
    else ligmg_options_default(&h->S.o);
    int rc = check_options(h->S.o);
    if (rc) {
      lg::g_last_error = "ligmg_load_hierarchy: unsupported option value";
      return rc;
    }
    h->S.load(dir);
    *out = h.release();
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  } catch (const std::exception &e) {
    return fail_unknown(e.what());
  }
}

int ligmg_dump_hierarchy(ligmg_handle h, const char *dir) {
  if (!h || !dir) return LIGMG_ERR_INVALID;
  try {
    h->S.dump(dir);
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  } catch (const std::exception &e) {
    return fail_unknown(e.what());
  }
}

int ligmg_dump_csv(ligmg_handle h, const char *graph, const ligmg_info *info, char *buf, int64_t cap) {
  if (!h || !graph || !info || !buf || cap <= 0) return LIGMG_ERR_INVALID;
  const std::string t = lg::Solver::csv_header() + h->S.csv_line(graph, *info);
  if ((int64_t)t.size() + 1 > cap) {
    lg::g_last_error = "ligmg_dump_csv: buffer too small (" + std::to_string(t.size() + 1) + " bytes needed)";
    return LIGMG_ERR_INVALID;
  }
  memcpy(buf, t.c_str(), t.size() + 1);
  return LIGMG_OK;
}

int ligmg_solve(ligmg_handle h, const double *b, double *x, int mem_kind, ligmg_info *info,
                double *history, int32_t cap) {
  if (!h || !b || !x) return LIGMG_ERR_INVALID;
  try {
    return h->S.solve(b, x, mem_kind, info, history, cap);
  } catch (const lg::Error &e) {
    return fail(e);
  } catch (const std::exception &e) {
    return fail_unknown(e.what());
  }
}

int ligmg_solve_from(ligmg_handle h, int32_t level, const double *b, double *x, int mem_kind, ligmg_info *info,
                     double *history, int32_t cap) {
  if (!h || !b || !x) return LIGMG_ERR_INVALID;
  try {
    return h->S.solve(b, x, mem_kind, info, history, cap, level);
  } catch (const lg::Error &e) {
    return fail(e);
  } catch (const std::exception &e) {
    return fail_unknown(e.what());
  }
}

int ligmg_destroy(ligmg_handle h) {
  delete h;
  return LIGMG_OK;
}

int ligmg_ordering(ligmg_handle h, int32_t *perm) {
  if (!h || !perm) return LIGMG_ERR_INVALID;
  try {
    lg::Solver &S = h->S;
    const int64_t n = S.H[0].A.n;
    if (S.perm_input.p) {
      lg::to_host(perm, S.perm_input.p, n, S.s);
      LG_HIP(hipStreamSynchronize(S.s));
    } else {
      for (int64_t i = 0; i < n; i++) perm[i] = (int32_t)i;
    }
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  }
}

int ligmg_num_levels(ligmg_handle h, int32_t *out) {
  if (!h || !out) return LIGMG_ERR_INVALID;
  *out = (int32_t)h->S.H.size();
  return LIGMG_OK;
}

int ligmg_tail_start(ligmg_handle h, int32_t *out) {
  if (!h || !out) return LIGMG_ERR_INVALID;
  *out = (int32_t)h->S.tail_l0;
  return LIGMG_OK;
}

int ligmg_setup_time(ligmg_handle h, double *out) {
  if (!h || !out) return LIGMG_ERR_INVALID;
  *out = h->S.setup_time;
  return LIGMG_OK;
}

int ligmg_level_info(ligmg_handle h, int32_t l, int32_t *kind, int64_t *n, int64_t *nnz,
                     int64_t *nc) {
  if (!h || l < 0 || l >= (int)h->S.H.size()) return LIGMG_ERR_INVALID;
  const lg::DLevel &L = h->S.H[l];
  *kind = L.kind;
  *n = L.A.n;
  *nnz = L.A.nnz;
  *nc = L.nc;
  return LIGMG_OK;
}

int ligmg_level_matrix(ligmg_handle h, int32_t l, int32_t *rp, int32_t *ci, double *v) {
  if (!h || l < 0 || l >= (int)h->S.H.size()) return LIGMG_ERR_INVALID;
  try {
    const lg::DLevel &L = h->S.H[l];
    // the solve layout's copy back in setup order
    const lg::pio::HostCsr A = h->S.to_setup(L.A, L.sperm.p, L.sperm.p, true);
    memcpy(rp, A.rp.data(), A.rp.size() * sizeof(int32_t));
    memcpy(ci, A.ci.data(), A.ci.size() * sizeof(int32_t));
    memcpy(v, A.v.data(), A.v.size() * sizeof(double));
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  }
}

int ligmg_level_map(ligmg_handle h, int32_t l, int32_t *map) {
  if (!h || l < 0 || l >= (int)h->S.H.size()) return LIGMG_ERR_INVALID;
  try {
    const lg::DLevel &L = h->S.H[l];
    hipStream_t s = h->S.s;
    if (L.kind == LIGMG_LEVEL_ELIM) lg::to_host(map, L.new_id.p, L.n, s);
    else if (L.kind == LIGMG_LEVEL_AGG) lg::to_host(map, L.agg_setup.p ? L.agg_setup.p : L.agg.p, L.n, s);
    else return LIGMG_ERR_INVALID;
    LG_HIP(hipStreamSynchronize(s));
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  }
}

int ligmg_level_states(ligmg_handle h, int32_t l, int32_t *ind, int32_t *state) {
  if (!h || l < 0 || l >= (int)h->S.H.size() || h->S.H[l].kind != LIGMG_LEVEL_AGG)
    return LIGMG_ERR_INVALID;
  try {
    const lg::DLevel &L = h->S.H[l];
    hipStream_t s = h->S.s;
    lg::to_host(ind, L.ind.p, L.n, s);
    lg::to_host(state, L.state.p, L.n, s);
    LG_HIP(hipStreamSynchronize(s));
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  }
}

int ligmg_level_strength(ligmg_handle h, int32_t l, double *out) {
  if (!h || l < 0 || l >= (int)h->S.H.size() || h->S.H[l].kind != LIGMG_LEVEL_AGG)
    return LIGMG_ERR_INVALID;
  try {
    const lg::DLevel &L = h->S.H[l];
    lg::to_host(out, L.strength.p, L.A.nnz, h->S.s);
    LG_HIP(hipStreamSynchronize(h->S.s));
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  }
}

int ligmg_level_cheby(ligmg_handle h, int32_t l, double *lo, double *hi) {
  if (!h || l < 0 || l >= (int)h->S.H.size() || h->S.H[l].kind != LIGMG_LEVEL_AGG)
    return LIGMG_ERR_INVALID;
  *lo = h->S.H[l].cheb_lo;
  *hi = h->S.H[l].cheb_hi;
  return LIGMG_OK;
}

int ligmg_aggregate_with_strength(int64_t n, const int32_t *rowptr, const int32_t *colind,
                                  const double *sv, const ligmg_options *opts, int32_t *agg,
                                  int32_t *ind, int32_t *state, int64_t *n_coarse) {
  if (n <= 0 || !rowptr || !colind || !sv || !agg || !ind || !state || !n_coarse)
    return LIGMG_ERR_INVALID;
  hipStream_t s = nullptr;
  try {
    ligmg_options o;
    if (opts) o = *opts;
    else ligmg_options_default(&o);
    // conductance rejection weighs the Laplacian's edges, which this entry
    // point does not receive
    if (o.rejection_ratio >= 0) {
      lg::g_last_error = "ligmg_aggregate_with_strength: rejection_ratio needs the level matrix";
      return LIGMG_ERR_UNSUPPORTED;
    }
    LG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    lg::DCsr A;
    A.n = n;
    A.nnz = rowptr[n] - rowptr[0];
    A.rp.alloc(n + 1);
    A.ci.alloc(std::max<int64_t>(A.nnz, 1));
    lg::DevBuf<double> S(std::max<int64_t>(A.nnz, 1));
    lg::to_device(A.rp.p, rowptr, n + 1, s);
    lg::to_device(A.ci.p, colind, A.nnz, s);
    lg::to_device(S.p, sv, A.nnz, s);
    lg::DevBuf<double> cs;
    lg::col_sums_of(A, S.p, cs, s);
    lg::DevBuf<int32_t> a, i2, st;
    *n_coarse = lg::run_voting(A, S.p, cs.p, o, a, i2, st, s);
    lg::to_host(agg, a.p, n, s);
    lg::to_host(ind, i2.p, n, s);
    lg::to_host(state, st.p, n, s);
    LG_HIP(hipStreamSynchronize(s));
    (void)hipStreamDestroy(s);
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    if (s) (void)hipStreamDestroy(s);
    return fail(e);
  }
}

int ligmg_bench_spmv(ligmg_handle h, int32_t l, int32_t prec, int32_t iters, double *avg_ms,
                     int64_t *bytes) {
  if (!h || l < 0 || l >= (int)h->S.H.size() || iters < 1 || !avg_ms || !bytes)
    return LIGMG_ERR_INVALID;
  try {
    lg::Solver &S = h->S;
    lg::DCsr &A = S.H[l].A;
    hipStream_t s = S.s;
    if (A.U.nunits == 0) lg::build_units(A, s);
    const bool f32 = prec == LIGMG_PREC_MIXED;
    if (f32 && !A.v32.p) lg::make_v32(A, s, true);
    const int64_t n = A.n;
    lg::DevBuf<double> x(f32 ? 0 : n), y(f32 ? 0 : n);
    lg::DevBuf<float> x32(f32 ? n : 0), y32(f32 ? n : 0);
    if (f32) {
      lg::DevBuf<double> t(n);
      lg::fill_f64(t.p, 1.0, n, s);
      lg::op_d2f(t.p, x32.p, n, s);
      LG_HIP(hipStreamSynchronize(s));
    } else {
      lg::fill_f64(x.p, 1.0, n, s);
    }
    auto run = [&]() {
      if (f32) lg::op_spmv<float>(A, x32.p, y32.p, s);
      else lg::op_spmv<double>(A, x.p, y.p, s);
    };
    run();  // warm-up
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
    *avg_ms = ms / iters;
    const int64_t sv = f32 ? 4 : 8, sx = f32 ? 4 : 8;
    // values + columns + row pointers + x read once + y written once; the
    // FP32 Laplacian form also reads its FP64 diagonal
    *bytes = A.nnz * (sv + 4) + (n + 1) * 4 + 2 * n * sx + (f32 && A.d32.p ? n * 8 : 0);
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  }
}

extern "C++" {
namespace {
template <class T>
double time_cheb(lg::Solver &S, lg::DLevel &L, const T *dinv, int iters, hipStream_t s) {
  const int64_t n = L.A.n;
  lg::DevBuf<T> b(n), x0(n), x1(n), x2(n);
  {
    lg::DevBuf<double> t(n);
    lg::fill_f64(t.p, 1.0, n, s);
    for (T *dst : {b.p, x0.p, x1.p}) {
      if constexpr (std::is_same<T, float>::value) lg::op_d2f(t.p, dst, n, s);
      else lg::d2d(dst, t.p, n, s);
    }
    LG_HIP(hipStreamSynchronize(s));
  }
  const double al = 0.5, be = 0.5 / std::max(L.cheb_hi, 1e-300);
  // rotate three buffers as the smoother does
  T *B[3] = {x0.p, x1.p, x2.p};
  auto step = [&](int k) {
    lg::op_cheb_step<T>(L.A, dinv, b.p, B[(k + 1) % 3], B[k % 3], B[(k + 2) % 3], al, be, s);
  };
  step(0);
  hipEvent_t e0, e1;
  LG_HIP(hipEventCreate(&e0));
  LG_HIP(hipEventCreate(&e1));
  LG_HIP(hipEventRecord(e0, s));
  for (int i = 0; i < iters; i++) step(i + 1);
  LG_HIP(hipEventRecord(e1, s));
  LG_HIP(hipEventSynchronize(e1));
  float ms = 0;
  LG_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / iters;
}
}  // namespace
}

int ligmg_bench_smoother(ligmg_handle h, int32_t l, int32_t prec, int32_t iters, double *avg_ms,
                         int64_t *bytes) {
  if (!h || l < 0 || l >= (int)h->S.H.size() || iters < 1 || !avg_ms || !bytes)
    return LIGMG_ERR_INVALID;
  try {
    lg::Solver &S = h->S;
    lg::DLevel &L = S.H[l];
    LG_CHECK(L.kind == LIGMG_LEVEL_AGG, LIGMG_ERR_INVALID, "ligmg_bench_smoother: not an aggregation level");
    const bool f32 = prec == LIGMG_PREC_MIXED;
    LG_CHECK(!f32 || (L.A.v32.p && L.dinv32.p), LIGMG_ERR_INVALID,
             "ligmg_bench_smoother: FP32 operators exist only with precision = LIGMG_PREC_MIXED");
    const int64_t n = L.A.n;
    if (f32) *avg_ms = time_cheb<float>(S, L, L.dinv32.p, iters, S.s);
    else *avg_ms = time_cheb<double>(S, L, L.dinv.p, iters, S.s);
    const int64_t sv = f32 ? 4 : 8, sx = f32 ? 4 : 8;
    *bytes = L.A.nnz * (sv + 4) + (n + 1) * 4 + 5 * n * sx + (f32 && L.A.d32.p ? n * 8 : 0);
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  }
}

int ligmg_bench_spmm(ligmg_handle h, int32_t l, int32_t prec, int32_t k, int32_t iters, double *avg_ms,
                     int64_t *bytes) {
  if (!h || l < 0 || l >= (int)h->S.H.size() || k < 1 || k > 64 || iters < 1 || !avg_ms || !bytes)
    return LIGMG_ERR_INVALID;
  try {
    lg::Solver &S = h->S;
    lg::DCsr &A = S.H[l].A;
    hipStream_t s = S.s;
    if (A.U.nunits == 0) lg::build_units(A, s);
    const bool f32 = prec == LIGMG_PREC_MIXED;
    if (f32 && !A.v32.p) lg::make_v32(A, s, true);
    int kp = 1;
    while (kp < k) kp <<= 1;
    const int64_t n = A.n;
    lg::DevBuf<double> parts((size_t)std::max<int64_t>(A.U.nparts, 1) * kp);
    lg::DevBuf<double> x(f32 ? 0 : n * kp), y(f32 ? 0 : n * kp);
    lg::DevBuf<float> x32(f32 ? n * kp : 0), y32(f32 ? n * kp : 0);
    if (f32) {
      lg::DevBuf<double> t(n * kp);
      lg::fill_f64(t.p, 1.0, n * kp, s);
      lg::op_d2f(t.p, x32.p, n * kp, s);
      LG_HIP(hipStreamSynchronize(s));
    } else {
      lg::fill_f64(x.p, 1.0, n * kp, s);
    }
    // the batched solve's tile for this operator and width
    if (S.o.spmm_tune) {
      if (f32) lg::tune_spmm<float>(A, kp, parts, s);
      else lg::tune_spmm<double>(A, kp, parts, s);
    }
    auto run = [&]() {
      if (f32) lg::op_bspmv<float>(A, x32.p, y32.p, kp, parts, s);
      else lg::op_bspmv<double>(A, x.p, y.p, kp, parts, s);
    };
    run();  // warm-up
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
    *avg_ms = ms / iters;
    const int64_t sv = f32 ? 4 : 8, sx = f32 ? 4 : 8;
    *bytes = A.nnz * (sv + 4) + (n + 1) * 4 + 2 * n * kp * sx;
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  }
}

int ligmg_bench_coarse(ligmg_handle h, int32_t k, int32_t iters, double *avg_ms, double *flops) {
  if (!h || k < 1 || k > 64 || iters < 1 || !avg_ms || !flops) return LIGMG_ERR_INVALID;
  try {
    lg::Solver &S = h->S;
    lg::DLevel &L = S.H.back();
    LG_CHECK(L.kind == LIGMG_LEVEL_COARSE && L.pinv.p, LIGMG_ERR_INVALID, "ligmg_bench_coarse: no dense coarse level");
    int kp = 1;
    while (kp < k) kp <<= 1;
    const int64_t n = L.n;
    lg::DevBuf<double> X((size_t)std::max<int64_t>(n, 1) * kp), Y((size_t)std::max<int64_t>(n, 1) * kp);
    lg::fill_f64(X.p, 1.0, (int64_t)X.n, S.s);
    auto once = [&] {
      if (kp == 1) lg::op_dense_apply<double>(L.pinv.p, n, X.p, Y.p, S.s);
      else lg::op_bdense_apply<double>(L.pinv.p, n, X.p, Y.p, kp, S.s);
    };
    once();
    hipEvent_t e0, e1;
    LG_HIP(hipEventCreate(&e0));
    LG_HIP(hipEventCreate(&e1));
    LG_HIP(hipEventRecord(e0, S.s));
    for (int i = 0; i < iters; i++) once();
    LG_HIP(hipEventRecord(e1, S.s));
    LG_HIP(hipEventSynchronize(e1));
    float ms = 0;
    LG_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = ms / iters;
    *flops = 2.0 * (double)n * (double)n * (double)kp;
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  }
}

int ligmg_bench_pinv(ligmg_handle h, double *total_ms, double *gj_ms, double *flops) {
  if (!h || !total_ms || !gj_ms || !flops) return LIGMG_ERR_INVALID;
  try {
    lg::Solver &S = h->S;
    lg::DLevel &L = S.H.back();
    LG_CHECK(L.kind == LIGMG_LEVEL_COARSE, LIGMG_ERR_INVALID, "ligmg_bench_pinv: no coarse level");
    const int64_t n = L.A.n;
    const auto rp = lg::download(L.A.rp.p, n + 1, S.s);
    const auto ci = lg::download(L.A.ci.p, L.A.nnz, S.s);
    const auto v = lg::download(L.A.v.p, L.A.nnz, S.s);
    lg::DevBuf<double> pinv;
    const auto t0 = std::chrono::steady_clock::now();
    lg::coarse_pinv(n, rp, ci, v, pinv, S.s, gj_ms);
    *total_ms = 1e3 * std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // Gauss-Jordan inversion of the n x n matrix: 2 n^3 flops (the padding's
    // extra work is not counted)
    *flops = (n >= lg::COARSE_DEVICE_MIN) ? 2.0 * (double)n * n * n : 0.0;
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  }
}

int ligmg_coarse_apply_block(ligmg_handle h, int32_t k, const double *X, double *Y) {
  if (!h || k < 1 || k > 64 || !X || !Y) return LIGMG_ERR_INVALID;
  try {
    lg::Solver &S = h->S;
    lg::DLevel &L = S.H.back();
    LG_CHECK(L.kind == LIGMG_LEVEL_COARSE && L.pinv.p, LIGMG_ERR_INVALID,
             "ligmg_coarse_apply_block: no dense coarse level");
    int kp = 1;
    while (kp < k) kp <<= 1;
    const int64_t n = L.n;
    std::vector<double> hx((size_t)n * kp, 0.0), hy((size_t)n * kp);
    for (int64_t i = 0; i < n; i++)
      for (int c = 0; c < k; c++) hx[i * kp + c] = X[i * k + c];
    lg::DevBuf<double> dx(std::max<size_t>(hx.size(), 1)), dy(std::max<size_t>(hx.size(), 1));
    lg::to_device(dx.p, hx.data(), hx.size(), S.s);
    if (kp == 1) lg::op_dense_apply<double>(L.pinv.p, n, dx.p, dy.p, S.s);
    else lg::op_bdense_apply<double>(L.pinv.p, n, dx.p, dy.p, kp, S.s);
    lg::to_host(hy.data(), dy.p, hy.size(), S.s);
    LG_HIP(hipStreamSynchronize(S.s));
    for (int64_t i = 0; i < n; i++)
      for (int c = 0; c < k; c++) Y[i * k + c] = hy[i * kp + c];
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  }
}

int ligmg_apply_cycle(ligmg_handle h, int32_t l, int32_t prec, const double *r, double *z) {
  if (!h || l < 0 || l >= (int)h->S.H.size() || !r || !z) return LIGMG_ERR_INVALID;
  try {
    lg::Solver &S = h->S;
    const bool f32 = prec == LIGMG_PREC_MIXED;
    LG_CHECK(!f32 || S.mixed, LIGMG_ERR_INVALID,
             "ligmg_apply_cycle: FP32 cycles exist only with precision = LIGMG_PREC_MIXED");
    const int64_t n = S.H[l].A.n;
    const int32_t *sp = S.H[l].sperm.p;  // setup-order vectors in and out
    lg::DevBuf<double> dr(n), dz(n), tmp(n);
    lg::to_device(tmp.p, r, n, S.s);
    if (sp) lg::gather_vec(n, sp, tmp.p, dr.p, S.s);
    else lg::d2d(dr.p, tmp.p, n, S.s);
    if (f32) {
      lg::DevBuf<float> fr(n);
      lg::op_d2f(dr.p, fr.p, n, S.s);
      const float *out = S.cycle<float>((size_t)l, fr.p);
      lg::op_f2d(out, dz.p, n, S.s);
    } else {
      const double *out = S.cycle<double>((size_t)l, dr.p);
      lg::d2d(dz.p, out, n, S.s);
    }
    if (sp) lg::scatter_vec(n, sp, dz.p, tmp.p, S.s);
    else lg::d2d(tmp.p, dz.p, n, S.s);
    lg::to_host(z, tmp.p, n, S.s);
    LG_HIP(hipStreamSynchronize(S.s));
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  }
}

int ligmg_level_spmv(ligmg_handle h, int32_t l, int32_t prec, const double *x, double *y) {
  if (!h || l < 0 || l >= (int)h->S.H.size() || !x || !y) return LIGMG_ERR_INVALID;
  try {
    lg::Solver &S = h->S;
    lg::DCsr &A = S.H[l].A;
    hipStream_t s = S.s;
    if (A.U.nunits == 0) lg::build_units(A, s);
    const int64_t n = A.n;
    const int32_t *sp = S.H[l].sperm.p;  // setup-order vectors in and out
    lg::DevBuf<double> dx(n), dy(n), tmp(n);
    lg::to_device(tmp.p, x, n, s);
    if (sp) lg::gather_vec(n, sp, tmp.p, dx.p, s);
    else lg::d2d(dx.p, tmp.p, n, s);
    if (prec == LIGMG_PREC_MIXED) {
      if (!A.v32.p) lg::make_v32(A, s, true);
      lg::DevBuf<float> fx(n), fy(n);
      lg::op_d2f(dx.p, fx.p, n, s);
      lg::op_spmv<float>(A, fx.p, fy.p, s);
      lg::op_f2d(fy.p, dy.p, n, s);
    } else {
      lg::op_spmv<double>(A, dx.p, dy.p, s);
    }
    if (sp) lg::scatter_vec(n, sp, dy.p, tmp.p, s);
    else lg::d2d(tmp.p, dy.p, n, s);
    lg::to_host(y, tmp.p, n, s);
    LG_HIP(hipStreamSynchronize(s));
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  }
}

int ligmg_solve_batch(ligmg_handle h, int32_t k, const double *B, double *X, int mem_kind,
                      ligmg_info *info, int32_t *iters_out) {
  if (!h || !B || !X || k < 1 || k > 64) return LIGMG_ERR_INVALID;
  try {
    return h->S.solve_batch(k, B, X, mem_kind, info, iters_out);
  } catch (const lg::Error &e) {
    return fail(e);
  } catch (const std::exception &e) {
    return fail_unknown(e.what());
  }
}

namespace {
__global__ void k_edge_grad(int64_t m, const int32_t *ei, const int32_t *ej, const double *x,
                            const double *lam, double *g) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = ei[e], j = ej[e];
    g[e] = -(lam[i] - lam[j]) * (x[i] - x[j]);
  }
}
__global__ void k_assemble(int64_t n, const int32_t *rp, const int32_t *ci, const int32_t *eid,
                           const double *w, double *v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double d = 0.0;
    int dpos = -1;
    for (int k = rp[i]; k < rp[i + 1]; k++) {
      const int e = eid[k];
      if (e < 0) {
        dpos = k;
      } else {
        v[k] = -w[e];
        d += w[e];
      }
    }
    if (dpos >= 0) v[dpos] = d;
  }
}
}  // namespace

int ligmg_edge_grad(int64_t m, const int32_t *ei, const int32_t *ej, const double *x,
                    const double *lam, double *g) {
  if (m < 0 || (m > 0 && (!ei || !ej || !x || !lam || !g))) return LIGMG_ERR_INVALID;
  if (m == 0) return LIGMG_OK;
  try {
    hipLaunchKernelGGL(k_edge_grad, dim3(lg::grid_for(m, 256, 16384)), dim3(256), 0, 0, m, ei, ej, x,
                       lam, g);
    LG_LAUNCH_CHECK();
    LG_HIP(hipStreamSynchronize(0));
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  }
}

int ligmg_assemble_laplacian(int64_t n, const int32_t *rp, const int32_t *ci, const int32_t *eid,
                             const double *w, double *v) {
  if (n <= 0 || !rp || !ci || !eid || !w || !v) return LIGMG_ERR_INVALID;
  try {
    hipLaunchKernelGGL(k_assemble, dim3(lg::grid_for(n, 256, 16384)), dim3(256), 0, 0, n, rp, ci, eid,
                       w, v);
    LG_LAUNCH_CHECK();
    LG_HIP(hipStreamSynchronize(0));
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  }
}

int ligmg_dense_pinv(int64_t n, const int32_t *rp, const int32_t *ci, const double *v, double *pinv) {
  if (n <= 0 || !rp || !ci || !v || !pinv) return LIGMG_ERR_INVALID;
  try {
    LG_CHECK(n <= lg::COARSE_DENSE_MAX, LIGMG_ERR_UNSUPPORTED,
             "ligmg_dense_pinv: at most " + std::to_string(lg::COARSE_DENSE_MAX) + " vertices");
    LG_CHECK(rp[0] == 0, LIGMG_ERR_INVALID, "ligmg_dense_pinv: rp[0] must be 0");
    for (int64_t i = 0; i < n; i++)
      LG_CHECK(rp[i + 1] >= rp[i], LIGMG_ERR_INVALID, "ligmg_dense_pinv: rp must be non-decreasing");
    for (int64_t k = 0; k < rp[n]; k++)
      LG_CHECK(ci[k] >= 0 && ci[k] < n, LIGMG_ERR_INVALID, "ligmg_dense_pinv: column out of range");
    const std::vector<int32_t> hrp(rp, rp + n + 1), hci(ci, ci + rp[n]);
    const std::vector<double> hv(v, v + rp[n]);
    hipStream_t s = nullptr;
    LG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{s};
    lg::DevBuf<double> P;
    lg::coarse_pinv(n, hrp, hci, hv, P, s);
    lg::to_host(pinv, P.p, (size_t)n * n, s);
    LG_HIP(hipStreamSynchronize(s));
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  }
}

int ligmg_check_units(int64_t n, const int32_t *rp, int32_t *same, int64_t *nunits) {
  if (n <= 0 || !rp || !same || !nunits) return LIGMG_ERR_INVALID;
  try {
    LG_CHECK(rp[0] == 0, LIGMG_ERR_INVALID, "ligmg_check_units: rp[0] must be 0");
    for (int64_t i = 0; i < n; i++)
      LG_CHECK(rp[i + 1] >= rp[i], LIGMG_ERR_INVALID, "ligmg_check_units: rp must be non-decreasing");
    hipStream_t s = 0;
    lg::DCsr A, B;
    A.n = B.n = n;
    A.rp.alloc(n + 1);
    B.rp.alloc(n + 1);
    lg::to_device(A.rp.p, rp, n + 1, s);
    lg::to_device(B.rp.p, rp, n + 1, s);
    lg::build_units(A, s);
    lg::build_units_host(B, s);
    const lg::Units &a = A.U, &b = B.U;
    bool eq = a.nunits == b.nunits && a.nlong == b.nlong && a.nparts == b.nparts && a.row_items == b.row_items;
    if (eq) {
      eq = lg::download(a.row0.p, a.nunits, s) == lg::download(b.row0.p, b.nunits, s) &&
           lg::download(a.nrows.p, a.nunits, s) == lg::download(b.nrows.p, b.nunits, s) &&
           lg::download(a.koff.p, a.nunits, s) == lg::download(b.koff.p, b.nunits, s) &&
           lg::download(a.lrow.p, a.nlong, s) == lg::download(b.lrow.p, b.nlong, s) &&
           lg::download(a.lpart.p, a.nlong + 1, s) == lg::download(b.lpart.p, b.nlong + 1, s);
    }
    *same = eq ? 1 : 0;
    *nunits = a.nunits;
    return LIGMG_OK;
  } catch (const lg::Error &e) {
    return fail(e);
  }
}

}  // extern "C"
