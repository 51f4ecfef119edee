// PETSc preconditioner "ligmg" over the C ABI - the drop-in for the
// reference's PC (include/ligmg/petsc_shim.hpp:18-74).  Built only where
// PETSc is installed (ligmg_amd/build.py, PETSC_DIR); PETSc is not part of
// this image, so this file is compiled and tested on PETSc installations
// only.
//
// Usage, as with the reference (README.md:37-50):
//   PCRegister("ligmg", PCCreate_LIGMG);
//   -ksp_type preonly -pc_type ligmg
//
// Differences from the reference shim: pc->pmat may be any assembled AIJ
// matrix (its rows are extracted to CSR and uploaded once at PCSetUp); the
// reference requires a MatShell wrapping a CombBLAS SpParMat.  Apply runs a
// full solve to the options' tolerance, as the reference does (:40-62).
#include <petsc/private/pcimpl.h>
#include <petscksp.h>

#include <vector>

#include "../../../include/ligmg_amd.h"

namespace {

struct LigmgPC {
  ligmg_handle h = nullptr;
  PetscInt n = 0;
};

PetscErrorCode PCSetUp_LIGMG(PC pc) {
  PetscFunctionBegin;
  LigmgPC *ctx = (LigmgPC *)pc->data;
  if (ctx->h) {
    ligmg_destroy(ctx->h);
    ctx->h = nullptr;
  }
  PetscInt m, n;
  PetscCall(MatGetSize(pc->pmat, &m, &n));
  PetscCheck(m == n, PETSC_COMM_SELF, PETSC_ERR_ARG_WRONG, "ligmg needs a square matrix");
  std::vector<int32_t> rp(m + 1, 0), ci;
  std::vector<double> v;
  for (PetscInt i = 0; i < m; i++) {
    PetscInt nc;
    const PetscInt *cols;
    const PetscScalar *vals;
    PetscCall(MatGetRow(pc->pmat, i, &nc, &cols, &vals));
    for (PetscInt k = 0; k < nc; k++) {
      ci.push_back((int32_t)cols[k]);
      v.push_back((double)PetscRealPart(vals[k]));
    }
    rp[i + 1] = (int32_t)ci.size();
    PetscCall(MatRestoreRow(pc->pmat, i, &nc, &cols, &vals));
  }
  ligmg_options o;
  ligmg_options_default(&o);  // the reference PC always uses default_options()
  int rc = ligmg_create(m, rp.data(), ci.data(), v.data(), LIGMG_MEM_HOST, &o, 0, nullptr, &ctx->h);
  PetscCheck(rc == LIGMG_OK, PETSC_COMM_SELF, PETSC_ERR_LIB, "ligmg_create: %s", ligmg_last_error());
  ctx->n = m;
  PetscFunctionReturn(PETSC_SUCCESS);
}

PetscErrorCode PCApply_LIGMG(PC pc, Vec x, Vec y) {
  PetscFunctionBegin;
  LigmgPC *ctx = (LigmgPC *)pc->data;
  const PetscScalar *b;
  PetscScalar *out;
  PetscCall(VecGetArrayRead(x, &b));
  PetscCall(VecGetArray(y, &out));
  ligmg_info info;
  int rc = ligmg_solve(ctx->h, (const double *)b, (double *)out, LIGMG_MEM_HOST, &info, nullptr, 0);
  PetscCall(VecRestoreArray(y, &out));
  PetscCall(VecRestoreArrayRead(x, &b));
  PetscCheck(rc == LIGMG_OK, PETSC_COMM_SELF, PETSC_ERR_LIB, "ligmg_solve: %s", ligmg_last_error());
  PetscFunctionReturn(PETSC_SUCCESS);
}

PetscErrorCode PCDestroy_LIGMG(PC pc) {
  PetscFunctionBegin;
  LigmgPC *ctx = (LigmgPC *)pc->data;
  if (ctx) {
    if (ctx->h) ligmg_destroy(ctx->h);
    delete ctx;
  }
  pc->data = nullptr;
  PetscFunctionReturn(PETSC_SUCCESS);
}

}  // namespace

extern "C" PetscErrorCode PCCreate_LIGMG(PC pc) {
  PetscFunctionBegin;
  pc->data = new LigmgPC();
  pc->ops->apply = PCApply_LIGMG;
  pc->ops->applytranspose = PCApply_LIGMG;  // L is symmetric
  pc->ops->setup = PCSetUp_LIGMG;
  pc->ops->destroy = PCDestroy_LIGMG;
  PetscFunctionReturn(PETSC_SUCCESS);
}
