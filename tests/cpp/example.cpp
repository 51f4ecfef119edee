// The reference's usage example (examples/main.cpp:1-27) against the C++
// facade: read a MatrixMarket Laplacian, build the solver, solve a
// mean-zero right-hand side, print the ConvergenceInfo line.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>
#include <algorithm>

#include "ligmg_amd.hpp"

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: example matrix.mtx [dump-dir]\n");
    return 2;
  }
  std::ifstream in(argv[1]);
  std::string line;
  std::getline(in, line);
  bool symmetric = line.find("symmetric") != std::string::npos;
  while (std::getline(in, line) && line[0] == '%') {
  }
  long n, m, nz;
  std::istringstream(line) >> n >> m >> nz;
  std::vector<std::tuple<int, int, double>> t;
  for (long k = 0; k < nz; k++) {
    long i, j;
    double v;
    in >> i >> j >> v;
    t.emplace_back((int)i - 1, (int)j - 1, v);
    if (symmetric && i != j) t.emplace_back((int)j - 1, (int)i - 1, v);
  }
  std::sort(t.begin(), t.end());
  std::vector<int32_t> rp(n + 1, 0), ci;
  std::vector<double> vals;
  for (auto &e : t) {
    rp[std::get<0>(e) + 1]++;
    ci.push_back(std::get<1>(e));
    vals.push_back(std::get<2>(e));
  }
  for (long i = 0; i < n; i++) rp[i + 1] += rp[i];

  // random rhs, mean removed (examples/main.cpp:13-15 with a fixed seed)
  std::vector<double> rhs(n);
  unsigned s = 12345;
  double mean = 0;
  for (auto &x : rhs) {
    s = s * 1103515245u + 12345u;
    x = (double)(s >> 8) / (double)(1u << 24);
    mean += x;
  }
  mean /= n;
  for (auto &x : rhs) x -= mean;

  ligmg::Solver solver(n, rp.data(), ci.data(), vals.data(), ligmg::default_options());
  solver.print();
  auto result = solver.solve(rhs);
  std::printf("iters    error  comp  conv econv   wda     time\n");
  std::printf("%5d %.2e %5.2f %.3f %.3f %5.2f %.2e\n", result.iters(), result.rel_residual(),
              result.cycle_complexity(), result.cf(), result.ecf(), result.wda(), result.time());
  // check the returned solution against the system
  double rr = 0, bb = 0;
  for (long i = 0; i < n; i++) {
    double y = 0;
    for (int k = rp[i]; k < rp[i + 1]; k++) y += vals[k] * result.solution[ci[k]];
    rr += (rhs[i] - y) * (rhs[i] - y);
    bb += rhs[i] * rhs[i];
  }
  std::printf("full residual %.3e\n", std::sqrt(rr / bb));
  // --csv and --dump of the reference's driver (main.cpp:119-127)
  solver.dump_csv(std::cout, result, "my_matrix");
  if (argc > 2) {
    solver.dump_hierarchy(argv[2]);
    ligmg::Solver again = ligmg::Solver::load(argv[2]);
    auto r2 = again.solve(rhs);
    bool same = r2.iters() == result.iters() && r2.solution == result.solution;
    std::printf("reloaded hierarchy: %s\n", same ? "same solution" : "DIFFERENT solution");
    if (!same) return 1;
  }
  return (result.converged() && std::sqrt(rr / bb) < 1e-7) ? 0 : 1;
}
