// The PETSc-independent part of the "ligmg" preconditioner
// (include/ligmg/petsc_shim.hpp:18-74): PCSetUp's row-by-row assembly of
// the preconditioning matrix into CSR and the solver built from it with
// default options, PCApply's full solve, PCDestroy.  The PETSc shim
// (csrc/petsc/ligmg_pc.cpp) only moves rows and arrays between PETSc and
// these calls, so everything it decides is exercised without PETSc.
// Host code only.
#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/ligmg_amd.h"
#include "common.hpp"

struct ligmg_pc {
  int64_t n = -1;       // rows expected by the assembly in progress
  int64_t next = 0;     // the next row to add
  std::vector<int32_t> rp, ci;
  std::vector<double> v;
  ligmg_handle h = nullptr;
};

namespace {
int pc_fail(const std::string &m) {
  lg::g_last_error = m;
  return LIGMG_ERR_INVALID;
}
}  // namespace

extern "C" {

int ligmg_pc_create(ligmg_pc_handle *out) {
  if (!out) return LIGMG_ERR_INVALID;
  *out = new ligmg_pc();
  return LIGMG_OK;
}

int ligmg_pc_begin(ligmg_pc_handle pc, int64_t n) {
  if (!pc) return LIGMG_ERR_INVALID;
  if (n <= 0 || n >= INT32_MAX) return pc_fail("ligmg_pc_begin: n out of range");
  pc->n = n;
  pc->next = 0;
  pc->rp.assign(1, 0);
  pc->ci.clear();
  pc->v.clear();
  return LIGMG_OK;
}

int ligmg_pc_add_row(ligmg_pc_handle pc, int64_t row, int64_t ncols, const int64_t *cols, const double *vals) {
  if (!pc) return LIGMG_ERR_INVALID;
  if (pc->n < 0) return pc_fail("ligmg_pc_add_row: no assembly in progress (ligmg_pc_begin)");
  if (row != pc->next) return pc_fail("ligmg_pc_add_row: rows must come in order 0, 1, ..., n-1");
  if (ncols < 0 || (ncols > 0 && (!cols || !vals))) return pc_fail("ligmg_pc_add_row: bad row arrays");
  if ((int64_t)pc->ci.size() + ncols >= INT32_MAX) return pc_fail("ligmg_pc_add_row: too many entries");
  // a row in ascending column order (an AIJ row already is), duplicates refused
  std::vector<int64_t> ord(ncols);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return cols[a] < cols[b]; });
  for (int64_t k = 0; k < ncols; k++) {
    const int64_t c = cols[ord[k]];
    if (c < 0 || c >= pc->n) return pc_fail("ligmg_pc_add_row: column out of range");
    if (k > 0 && c == cols[ord[k - 1]]) return pc_fail("ligmg_pc_add_row: duplicate column");
  }
  for (int64_t k = 0; k < ncols; k++) {
    pc->ci.push_back((int32_t)cols[ord[k]]);
    pc->v.push_back(vals[ord[k]]);
  }
  pc->rp.push_back((int32_t)pc->ci.size());
  pc->next++;
  return LIGMG_OK;
}

int ligmg_pc_setup(ligmg_pc_handle pc) {
  if (!pc) return LIGMG_ERR_INVALID;
  if (pc->n < 0 || pc->next != pc->n) return pc_fail("ligmg_pc_setup: the assembly is missing rows");
  if (pc->h) {
    ligmg_destroy(pc->h);
    pc->h = nullptr;
  }
  // the reference PC always builds with default_options() (petsc_shim.hpp:31)
  ligmg_options o;
  ligmg_options_default(&o);
  const int rc = ligmg_create(pc->n, pc->rp.data(), pc->ci.data(), pc->v.data(), LIGMG_MEM_HOST, &o, 0, nullptr,
                              &pc->h);
  if (rc != LIGMG_OK) pc->h = nullptr;
  return rc;
}

int ligmg_pc_apply(ligmg_pc_handle pc, const double *b, double *x, ligmg_info *info) {
  if (!pc || !b || !x) return LIGMG_ERR_INVALID;
  if (!pc->h) return pc_fail("ligmg_pc_apply: not set up (ligmg_pc_setup)");
  ligmg_info local;
  return ligmg_solve(pc->h, b, x, LIGMG_MEM_HOST, info ? info : &local, nullptr, 0);
}

int ligmg_pc_destroy(ligmg_pc_handle pc) {
  if (!pc) return LIGMG_OK;
  if (pc->h) ligmg_destroy(pc->h);
  delete pc;
  return LIGMG_OK;
}

}  // extern "C"
