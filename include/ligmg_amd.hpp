// ligmg_amd.hpp - header-only C++ facade over the C ABI (ligmg_amd.h),
// mirroring the reference's public C++ API:
//
//   ligmg::Solver<T,F,DER>(A, opts) / Solver(A)   include/ligmg/ligmg.hpp:17-28
//   Solver::solve(b)                              include/ligmg/ligmg.hpp:31-33
//   Solver::solve(b, from)                        include/ligmg/ligmg.hpp:36-38
//   Solver::print()                               include/ligmg/ligmg.hpp:40
//   Solver::dump_csv(os, info)                    include/ligmg/ligmg.hpp:42-45
//   Solver::dump_hierarchy(dir)                   include/ligmg/ligmg.hpp:47-52
//   ConvergenceInfo<T,F> (residuals, iters(), time(), work(), solution,
//                         wda(), ecf(), cf(), cycle_complexity(),
//                         rel_residual())         include/ligmg/util.hpp:143-170
//   Options / default_options()                   include/ligmg/options.hpp:80-136
//
// The matrix is a CSR graph Laplacian (int32 indices, double values) in host
// or device memory instead of a CombBLAS SpParMat; vectors are std::vector
// (host) or raw device pointers.  Errors from the library are thrown as
// ligmg::Error, as the reference throws std::runtime_error on PETSc errors
// (include/ligmg/util.hpp:651-660).
#pragma once
#include <cmath>
#include <cstdio>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ligmg_amd.h"

namespace ligmg {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const char *what) {
  if (rc != LIGMG_OK) throw Error(rc, std::string(what) + ": " + ligmg_last_error());
}

using Options = ligmg_options;

inline Options default_options() {
  Options o;
  ligmg_options_default(&o);
  return o;
}

struct ConvergenceInfo {
  std::vector<double> residuals;  // reduced-system residual history incl. r0
  int _iters = 0;
  double _time = 0, _work = 0;
  std::vector<double> solution;  // full system (host solves)
  ligmg_info info{};

  double wda() const { return info.wda; }
  double cycle_complexity() const { return info.cycle_complexity; }
  double work() const { return _work; }
  double ecf() const { return info.ecf; }
  double rel_residual() const { return info.rel_residual; }
  double cf() const { return info.cf; }
  double time() const { return _time; }
  int iters() const { return _iters; }
  double full_rel_residual() const { return info.full_rel_residual; }
  bool converged() const { return info.status == LIGMG_CONVERGED; }
};

class Solver {
 public:
  Options opts;
  double setup_time = 0;

  // Solver(A, opts): host CSR arrays of an n x n graph Laplacian.
  Solver(int64_t n, const int32_t *rowptr, const int32_t *colind, const double *vals,
         const Options &o, int mem_kind = LIGMG_MEM_HOST)
      : opts(o), n_(n) {
    check(ligmg_create(n, rowptr, colind, vals, mem_kind, &opts, 0, nullptr, &h_), "ligmg_create");
    ligmg_setup_time(h_, &setup_time);
  }
  Solver(int64_t n, const int32_t *rowptr, const int32_t *colind, const double *vals)
      : Solver(n, rowptr, colind, vals, default_options()) {}
  Solver(const Solver &) = delete;
  Solver &operator=(const Solver &) = delete;
  Solver(Solver &&o) noexcept : opts(o.opts), setup_time(o.setup_time), h_(o.h_), n_(o.n_) {
    o.h_ = nullptr;
  }
  ~Solver() {
    if (h_) ligmg_destroy(h_);
  }

  // solve(b): host right-hand side, returns the solution in the info.
  ConvergenceInfo solve(const std::vector<double> &b) const {
    if ((int64_t)b.size() != n_) throw Error(LIGMG_ERR_INVALID, "solve: b has the wrong length");
    ConvergenceInfo ci;
    ci.solution.resize(n_);
    std::vector<double> hist(opts.iters + 2);
    check(ligmg_solve(h_, b.data(), ci.solution.data(), LIGMG_MEM_HOST, &ci.info, hist.data(),
                      (int32_t)hist.size()),
          "ligmg_solve");
    hist.resize(ci.info.history_len);
    ci.residuals = std::move(hist);
    ci._iters = ci.info.iters;
    ci._time = ci.info.time;
    ci._work = ci.info.work;
    return ci;
  }

  // solve(b, from): the system of hierarchy level `from`; b has that
  // level's size and is numbered as ligmg_level_matrix reports the level.
  ConvergenceInfo solve(const std::vector<double> &b, int from) const {
    int32_t kind;
    int64_t n, nnz, nc;
    check(ligmg_level_info(h_, from, &kind, &n, &nnz, &nc), "ligmg_level_info");
    if ((int64_t)b.size() != n) throw Error(LIGMG_ERR_INVALID, "solve: b has the wrong length for the level");
    ConvergenceInfo ci;
    ci.solution.resize(n);
    std::vector<double> hist(opts.iters + 2);
    check(ligmg_solve_from(h_, from, b.data(), ci.solution.data(), LIGMG_MEM_HOST, &ci.info, hist.data(),
                           (int32_t)hist.size()),
          "ligmg_solve_from");
    hist.resize(ci.info.history_len);
    ci.residuals = std::move(hist);
    ci._iters = ci.info.iters;
    ci._time = ci.info.time;
    ci._work = ci.info.work;
    return ci;
  }

  // solve with device-resident b and x (length n)
  ConvergenceInfo solve_device(const double *b, double *x) const {
    ConvergenceInfo ci;
    std::vector<double> hist(opts.iters + 2);
    check(ligmg_solve(h_, b, x, LIGMG_MEM_DEVICE, &ci.info, hist.data(), (int32_t)hist.size()),
          "ligmg_solve");
    hist.resize(ci.info.history_len);
    ci.residuals = std::move(hist);
    ci._iters = ci.info.iters;
    ci._time = ci.info.time;
    ci._work = ci.info.work;
    return ci;
  }

  // print_hierarchy (multigrid.hpp:265-292)
  void print() const {
    int32_t nl = 0;
    ligmg_num_levels(h_, &nl);
    std::printf("%5s  %12s  %14s  %6s\n", "Level", "Size", "NNZ", "Type");
    static const char *names[] = {"elim", "agg", "coarse"};
    for (int32_t l = 0; l < nl; l++) {
      int32_t kind;
      int64_t n, nnz, nc;
      ligmg_level_info(h_, l, &kind, &n, &nnz, &nc);
      std::printf("%5d  %12lld  %14lld  %6s\n", l, (long long)n, (long long)nnz, names[kind]);
    }
  }

  // dump_csv: header and result line of a finished solve
  void dump_csv(std::ostream &os, const ConvergenceInfo &cv, const std::string &graph = "") const {
    std::vector<char> buf(1 << 14);
    check(ligmg_dump_csv(h_, graph.c_str(), &cv.info, buf.data(), (int64_t)buf.size()), "ligmg_dump_csv");
    os << buf.data();
  }

  // dump_hierarchy: PETSc binary files of every level under dir
  void dump_hierarchy(const std::string &dir) const {
    check(ligmg_dump_hierarchy(h_, dir.c_str()), "ligmg_dump_hierarchy");
  }

  // a solver from a dump (new)
  static Solver load(const std::string &dir, const Options &o = default_options()) {
    Solver S;
    S.opts = o;
    check(ligmg_load_hierarchy(dir.c_str(), &S.opts, &S.h_), "ligmg_load_hierarchy");
    int32_t kind;
    int64_t nnz, nc;
    ligmg_level_info(S.h_, 0, &kind, &S.n_, &nnz, &nc);
    ligmg_setup_time(S.h_, &S.setup_time);
    return S;
  }

  ligmg_handle handle() const { return h_; }
  int64_t size() const { return n_; }

 private:
  Solver() = default;
  ligmg_handle h_ = nullptr;
  int64_t n_ = 0;
};

// README.md:64 spells the type ligmg::ligmg
using ligmg = Solver;

}  // namespace ligmg
