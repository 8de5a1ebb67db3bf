/*
 * st.c -- spectral transformation STSHIFT with sigma = 0
 * (src/sys/classes/st/impls/shift/shift.c:79-97, st/interface/stsolve.c:16-60,
 * 244-359).  With sigma = 0 and a standard problem the operator is A itself
 * (STMatMAXPY_Private leaves T[0] = A, stsolve.c:598-610), so STGetOperator
 * hands out A and STApply is MatMult(A); eigenvalues need no back-transform.
 */
#include "slepcimpl.h"

PetscErrorCode STCreate(ST *st)
{
  PetscCall(PetscMallocZero(sizeof(struct _p_ST), st));
  return PETSC_SUCCESS;
}

PetscErrorCode STSetMatrices(ST st, PetscInt n, Mat A[])
{
  PetscCheck(st, PETSC_ERR_ARG_NULL, "null ST");
  PetscCheck(n == 1 && A && A[0], PETSC_ERR_SUP, "only standard eigenproblems (one matrix) are supported");
  st->A = A[0];
  return PETSC_SUCCESS;
}

PetscErrorCode STGetOperator(ST st, Mat *Op)
{
  PetscCheck(st && st->A, PETSC_ERR_ORDER, "STSetMatrices must be called first");
  PetscCheck(st->sigma == 0.0, PETSC_ERR_SUP, "shifted operators are not implemented");
  *Op = st->A;
  return PETSC_SUCCESS;
}

PetscErrorCode STRestoreOperator(ST st, Mat *Op)
{
  (void)st;
  *Op = NULL;
  return PETSC_SUCCESS;
}

/* STApply_Generic (stsolve.c:16-25): y = Op x */
PetscErrorCode STApply(ST st, Vec x, Vec y)
{
  PetscCheck(st && st->A, PETSC_ERR_ORDER, "STSetMatrices must be called first");
  return MatMult(st->A, x, y);
}

PetscErrorCode STDestroy(ST *st)
{
  if (!st || !*st) return PETSC_SUCCESS;
  free(*st);
  *st = NULL;
  return PETSC_SUCCESS;
}
