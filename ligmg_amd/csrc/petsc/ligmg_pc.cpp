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
// Everything but the PETSc calls lives in the library's ligmg_pc_* helper
// (csrc/pc.hip), which tests drive without PETSc.
#include <petsc/private/pcimpl.h>
#include <petscksp.h>

#include <vector>

#include "../../../include/ligmg_amd.h"

namespace {

PetscErrorCode check(int rc, const char *what) {
  PetscCheck(rc == LIGMG_OK, PETSC_COMM_SELF, PETSC_ERR_LIB, "%s: %s", what, ligmg_last_error());
  return PETSC_SUCCESS;
}

// rows of pc->pmat in order into the helper, then the solver with default
// options (ligmg_pc_setup)
PetscErrorCode PCSetUp_LIGMG(PC pc) {
  PetscFunctionBegin;
  ligmg_pc_handle h = (ligmg_pc_handle)pc->data;
  PetscInt m, n;
  PetscCall(MatGetSize(pc->pmat, &m, &n));
  PetscCheck(m == n, PETSC_COMM_SELF, PETSC_ERR_ARG_WRONG, "ligmg needs a square matrix");
  PetscCall(check(ligmg_pc_begin(h, m), "ligmg_pc_begin"));
  std::vector<int64_t> c64;
  std::vector<double> v64;
  for (PetscInt i = 0; i < m; i++) {
    PetscInt nc;
    const PetscInt *cols;
    const PetscScalar *vals;
    PetscCall(MatGetRow(pc->pmat, i, &nc, &cols, &vals));
    c64.assign(cols, cols + nc);
    v64.resize(nc);
    for (PetscInt k = 0; k < nc; k++) v64[k] = (double)PetscRealPart(vals[k]);
    const int rc = ligmg_pc_add_row(h, i, nc, c64.data(), v64.data());
    PetscCall(MatRestoreRow(pc->pmat, i, &nc, &cols, &vals));
    PetscCall(check(rc, "ligmg_pc_add_row"));
  }
  PetscCall(check(ligmg_pc_setup(h), "ligmg_pc_setup"));
  PetscFunctionReturn(PETSC_SUCCESS);
}

PetscErrorCode PCApply_LIGMG(PC pc, Vec x, Vec y) {
  PetscFunctionBegin;
  const PetscScalar *b;
  PetscScalar *out;
  PetscCall(VecGetArrayRead(x, &b));
  PetscCall(VecGetArray(y, &out));
  const int rc = ligmg_pc_apply((ligmg_pc_handle)pc->data, (const double *)b, (double *)out, nullptr);
  PetscCall(VecRestoreArray(y, &out));
  PetscCall(VecRestoreArrayRead(x, &b));
  PetscCall(check(rc, "ligmg_pc_apply"));
  PetscFunctionReturn(PETSC_SUCCESS);
}

PetscErrorCode PCDestroy_LIGMG(PC pc) {
  PetscFunctionBegin;
  ligmg_pc_destroy((ligmg_pc_handle)pc->data);
  pc->data = nullptr;
  PetscFunctionReturn(PETSC_SUCCESS);
}

}  // namespace

extern "C" PetscErrorCode PCCreate_LIGMG(PC pc) {
  PetscFunctionBegin;
  ligmg_pc_handle h = nullptr;
  PetscCall(check(ligmg_pc_create(&h), "ligmg_pc_create"));
  pc->data = (void *)h;
  pc->ops->apply = PCApply_LIGMG;
  pc->ops->applytranspose = PCApply_LIGMG;  // L is symmetric
  pc->ops->setup = PCSetUp_LIGMG;
  pc->ops->destroy = PCDestroy_LIGMG;
  PetscFunctionReturn(PETSC_SUCCESS);
}
