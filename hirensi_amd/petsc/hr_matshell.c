/*
 * PETSc MatShell adapter: AK, AC, AG of the reference's explicit solver as
 * shells over the HIP operator of include/hirensi.h.
 *
 * The reference creates the three matrices with MatCreateAIJ
 * (PETSc_Global_Objects.F90:46-51), fills them element by element
 * (Assemble_PETSc_3D.F90:6-436) and then only calls on them
 *   MatDiagonalScale, MatScale          (Solver_PETSc_Explicit.F90:815-824, 1222-1229)
 *   MatMultAdd, MatMultTransposeAdd     (Solver_PETSc_Explicit.F90:1005-1034, 1323-1352)
 *   MatZeroEntries, MatAssemblyBegin/End (Program_Main.F90:692-697, Solve_Forward.F90:63-85)
 * which is exactly the operation set this shell provides.  Vectors keep the
 * reference's numbering; hr_apply reads and writes them in place, in host
 * memory or, for VECHIP vectors and a device-pointer context, in HBM.
 *
 * Builds only where PETSc is installed (PETSC_DIR); single process (one
 * rank owns all equations, as the one-GPU context does).
 */
#include <petscmat.h>
#include "hirensi.h"

typedef struct {
  hr_ctx* ctx;
  int which;
  int device;   /* the context takes device pointers (HR_DEVICE_PTRS): vectors stay in HBM */
} HrShell;

/* Array of a vector in the memory the context takes: with `device`, the
 * vector must live on the GPU (VECHIP) and its device array is passed without
 * a copy; otherwise the host array (the library stages it). */
static PetscErrorCode HrGetRead(const HrShell* s, Vec v, const PetscScalar** a) {
  PetscMemType mt;
  PetscErrorCode ierr;
  ierr = VecGetArrayReadAndMemType(v, a, &mt); CHKERRQ(ierr);
  if (PetscMemTypeDevice(mt) != (s->device != 0))
    SETERRQ(PETSC_COMM_SELF, PETSC_ERR_ARG_WRONG, "hirensi: vector memory does not match the context");
  return 0;
}

static PetscErrorCode HrGet(const HrShell* s, Vec v, PetscScalar** a) {
  PetscMemType mt;
  PetscErrorCode ierr;
  ierr = VecGetArrayAndMemType(v, a, &mt); CHKERRQ(ierr);
  if (PetscMemTypeDevice(mt) != (s->device != 0))
    SETERRQ(PETSC_COMM_SELF, PETSC_ERR_ARG_WRONG, "hirensi: vector memory does not match the context");
  return 0;
}

#define HRCHK(ctx, call)                                                               \
  do {                                                                                 \
    int _rc = (call);                                                                  \
    if (_rc) SETERRQ(PETSC_COMM_SELF, PETSC_ERR_LIB, "hirensi: %s", hr_last_error(ctx)); \
  } while (0)

static PetscErrorCode HrMultAddImpl(Mat A, Vec x, Vec y, Vec z, int transpose) {
  HrShell* s;
  const PetscScalar *xa, *ya;
  PetscScalar* za;
  PetscErrorCode ierr;
  ierr = MatShellGetContext(A, (void**)&s); CHKERRQ(ierr);
  ierr = HrGetRead(s, x, &xa); CHKERRQ(ierr);
  if (y == z) {
    ierr = HrGet(s, z, &za); CHKERRQ(ierr);
    HRCHK(s->ctx, hr_apply(s->ctx, s->which, transpose, xa, za, za));
  } else {
    ierr = HrGetRead(s, y, &ya); CHKERRQ(ierr);
    ierr = HrGet(s, z, &za); CHKERRQ(ierr);
    HRCHK(s->ctx, hr_apply(s->ctx, s->which, transpose, xa, ya, za));
  }
  /* device arrays: the product is queued on the library's stream; PETSc may
   * use z on its own stream next */
  if (s->device) HRCHK(s->ctx, hr_synchronize(s->ctx));
  if (y != z) { ierr = VecRestoreArrayReadAndMemType(y, &ya); CHKERRQ(ierr); }
  ierr = VecRestoreArrayAndMemType(z, &za); CHKERRQ(ierr);
  ierr = VecRestoreArrayReadAndMemType(x, &xa); CHKERRQ(ierr);
  return 0;
}

static PetscErrorCode HrMultAdd(Mat A, Vec x, Vec y, Vec z) { return HrMultAddImpl(A, x, y, z, 0); }
static PetscErrorCode HrMultTransposeAdd(Mat A, Vec x, Vec y, Vec z) { return HrMultAddImpl(A, x, y, z, 1); }

static PetscErrorCode HrMult(Mat A, Vec x, Vec z) {
  PetscErrorCode ierr;
  ierr = VecSet(z, 0.0); CHKERRQ(ierr);
  return HrMultAddImpl(A, x, z, z, 0);
}

static PetscErrorCode HrMultTranspose(Mat A, Vec x, Vec z) {
  PetscErrorCode ierr;
  ierr = VecSet(z, 0.0); CHKERRQ(ierr);
  return HrMultAddImpl(A, x, z, z, 1);
}

static PetscErrorCode HrDiagonalScale(Mat A, Vec l, Vec r) {
  HrShell* s;
  const PetscScalar *la = NULL, *ra = NULL;
  PetscErrorCode ierr;
  ierr = MatShellGetContext(A, (void**)&s); CHKERRQ(ierr);
  if (l) { ierr = HrGetRead(s, l, &la); CHKERRQ(ierr); }
  if (r) { ierr = HrGetRead(s, r, &ra); CHKERRQ(ierr); }
  HRCHK(s->ctx, hr_diagonal_scale(s->ctx, s->which, la, ra));
  if (s->device) HRCHK(s->ctx, hr_synchronize(s->ctx));
  if (l) { ierr = VecRestoreArrayReadAndMemType(l, &la); CHKERRQ(ierr); }
  if (r) { ierr = VecRestoreArrayReadAndMemType(r, &ra); CHKERRQ(ierr); }
  return 0;
}

static PetscErrorCode HrScale(Mat A, PetscScalar a) {
  HrShell* s;
  PetscErrorCode ierr;
  ierr = MatShellGetContext(A, (void**)&s); CHKERRQ(ierr);
  HRCHK(s->ctx, hr_scale(s->ctx, s->which, (double)a));
  return 0;
}

static PetscErrorCode HrZeroEntries(Mat A) {
  HrShell* s;
  PetscErrorCode ierr;
  ierr = MatShellGetContext(A, (void**)&s); CHKERRQ(ierr);
  HRCHK(s->ctx, hr_reset_scaling(s->ctx, s->which));
  return 0;
}

static PetscErrorCode HrDestroy(Mat A) {
  HrShell* s;
  PetscErrorCode ierr;
  ierr = MatShellGetContext(A, (void**)&s); CHKERRQ(ierr);
  ierr = PetscFree(s); CHKERRQ(ierr);
  return 0;
}

/* Create the shell for operator `which` (HR_K, HR_C, HR_G) of `ctx`.
 * device: ctx was created with HR_DEVICE_PTRS and the driver's vectors are
 * VECHIP (their device arrays go to the GPU operator without a copy); 0: host
 * vectors. */
PetscErrorCode HrMatCreateShellMem(MPI_Comm comm, hr_ctx* ctx, int which, int device, Mat* A) {
  long long neq, nnode, nstore;
  HrShell* s;
  PetscErrorCode ierr;
  if (hr_get_sizes(ctx, &neq, &nnode, &nstore)) SETERRQ(comm, PETSC_ERR_LIB, "hirensi: bad context");
  ierr = PetscNew(&s); CHKERRQ(ierr);
  s->ctx = ctx;
  s->which = which;
  s->device = device;
  ierr = MatCreateShell(comm, (PetscInt)neq, (PetscInt)neq, (PetscInt)neq, (PetscInt)neq, s, A); CHKERRQ(ierr);
  ierr = MatShellSetOperation(*A, MATOP_MULT, (void (*)(void))HrMult); CHKERRQ(ierr);
  ierr = MatShellSetOperation(*A, MATOP_MULT_ADD, (void (*)(void))HrMultAdd); CHKERRQ(ierr);
  ierr = MatShellSetOperation(*A, MATOP_MULT_TRANSPOSE, (void (*)(void))HrMultTranspose); CHKERRQ(ierr);
  ierr = MatShellSetOperation(*A, MATOP_MULT_TRANSPOSE_ADD, (void (*)(void))HrMultTransposeAdd); CHKERRQ(ierr);
  ierr = MatShellSetOperation(*A, MATOP_DIAGONAL_SCALE, (void (*)(void))HrDiagonalScale); CHKERRQ(ierr);
  ierr = MatShellSetOperation(*A, MATOP_SCALE, (void (*)(void))HrScale); CHKERRQ(ierr);
  ierr = MatShellSetOperation(*A, MATOP_ZERO_ENTRIES, (void (*)(void))HrZeroEntries); CHKERRQ(ierr);
  ierr = MatShellSetOperation(*A, MATOP_DESTROY, (void (*)(void))HrDestroy); CHKERRQ(ierr);
  return 0;
}

/* Host vectors (the reference's VECMPI/VECSEQ on one rank). */
PetscErrorCode HrMatCreateShell(MPI_Comm comm, hr_ctx* ctx, int which, Mat* A) {
  return HrMatCreateShellMem(comm, ctx, which, 0, A);
}

/* Fortran binding: CALL HrMatCreateShellF(ctx, which, A, ierr) (host vectors),
 * CALL HrMatCreateShellDevF(ctx, which, A, ierr) (VECHIP vectors). */
void hrmatcreateshellf_(hr_ctx** ctx, int* which, Mat* A, PetscErrorCode* ierr) {
  *ierr = HrMatCreateShellMem(PETSC_COMM_WORLD, *ctx, *which, 0, A);
}
void hrmatcreateshelldevf_(hr_ctx** ctx, int* which, Mat* A, PetscErrorCode* ierr) {
  *ierr = HrMatCreateShellMem(PETSC_COMM_WORLD, *ctx, *which, 1, A);
}
