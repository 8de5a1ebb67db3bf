/*
 * mat.c -- the PETSc Mat pieces on the path: the registry (MatRegister /
 * MatSetType), MatMult, a host MATSEQDENSE for the projected-problem matrices
 * handed to BVMultInPlace (src/sys/classes/bv/interface/bvops.c:220-245 requires
 * a sequential dense Q), and the device CSR type MATHIPCSR whose MatMult is the
 * adaptive SpMV of libks_hip (replacing MatMult_SeqAIJHIPSPARSE).
 */
#include "slepcimpl.h"

static SlepcFunctionList MatList;

PetscErrorCode MatRegister(const char name[], PetscErrorCode (*create)(Mat))
{
  return SlepcFunctionListAdd(&MatList, name, (void *)create);
}

PetscErrorCode MatCreate(Mat *A)
{
  PetscCall(SlepcCheckInitialized());
  PetscCall(PetscMallocZero(sizeof(struct _p_Mat), A));
  PetscCall(PetscHeaderInit(&(*A)->hdr, "Mat"));
  return PETSC_SUCCESS;
}

PetscErrorCode MatSetSizes(Mat A, PetscInt m, PetscInt n, PetscInt M, PetscInt N)
{
  PetscCheck(A, PETSC_ERR_ARG_NULL, "null matrix");
  A->m = M >= 0 ? M : m;
  A->n = N >= 0 ? N : n;
  PetscCheck(A->m >= 0 && A->n >= 0, PETSC_ERR_ARG_OUTOFRANGE, "bad matrix sizes");
  return PETSC_SUCCESS;
}

PetscErrorCode MatSetType(Mat A, MatType type)
{
  PetscErrorCode (*create)(Mat);
  PetscCheck(A && type, PETSC_ERR_ARG_NULL, "null argument");
  if (!strcmp(A->hdr.type_name, type)) return PETSC_SUCCESS;
  create = (PetscErrorCode(*)(Mat))SlepcFunctionListFind(&MatList, type);
  if (!create && !strcmp(type, MATSEQDENSE)) SETERRQ(PETSC_ERR_ARG_WRONG, "use MatCreateSeqDense for %s", type);
  PetscCheck(create, PETSC_ERR_ARG_WRONG, "unknown Mat type %s", type);
  if (A->ops->destroy) PetscCall(A->ops->destroy(A));
  memset(A->ops, 0, sizeof(struct _MatOps));
  snprintf(A->hdr.type_name, sizeof(A->hdr.type_name), "%s", type);
  PetscCall(create(A));
  return PETSC_SUCCESS;
}

PetscErrorCode MatGetType(Mat A, MatType *type)
{
  PetscCheck(A, PETSC_ERR_ARG_NULL, "null matrix");
  *type = A->hdr.type_name;
  return PETSC_SUCCESS;
}

PetscErrorCode MatGetSize(Mat A, PetscInt *M, PetscInt *N)
{
  PetscCheck(A, PETSC_ERR_ARG_NULL, "null matrix");
  if (M) *M = A->m;
  if (N) *N = A->n;
  return PETSC_SUCCESS;
}

PetscErrorCode MatMult(Mat A, Vec x, Vec y)
{
  PetscCheck(A && x && y, PETSC_ERR_ARG_NULL, "null argument");
  PetscCheck(A->ops->mult, PETSC_ERR_SUP, "Mat type %s has no MatMult", A->hdr.type_name);
  PetscCheck(x->n == A->n && y->n == A->m, PETSC_ERR_ARG_SIZ, "MatMult sizes: A %dx%d, x %d, y %d", A->m, A->n,
             x->n, y->n);
  PetscCheck(x != y, PETSC_ERR_ARG_WRONG, "x and y must be different vectors");
  return A->ops->mult(A, x, y);
}

PetscErrorCode MatMultTranspose(Mat A, Vec x, Vec y)
{
  PetscCheck(A && x && y, PETSC_ERR_ARG_NULL, "null argument");
  PetscCheck(A->ops->multtranspose, PETSC_ERR_SUP, "Mat type %s has no MatMultTranspose", A->hdr.type_name);
  PetscCheck(x->n == A->m && y->n == A->n, PETSC_ERR_ARG_SIZ, "MatMultTranspose sizes: A %dx%d, x %d, y %d", A->m,
             A->n, x->n, y->n);
  PetscCheck(x != y, PETSC_ERR_ARG_WRONG, "x and y must be different vectors");
  return A->ops->multtranspose(A, x, y);
}

PetscErrorCode MatDestroy(Mat *A)
{
  if (!A || !*A) return PETSC_SUCCESS;
  if ((*A)->ops->destroy) PetscCall((*A)->ops->destroy(*A));
  free(*A);
  *A = NULL;
  return PETSC_SUCCESS;
}

/* ---------------------------------------------------------------- MATSEQDENSE (host) */
static PetscErrorCode MatDestroy_SeqDense(Mat A)
{
  Mat_SeqDense *d = (Mat_SeqDense *)A->data;
  if (d && d->owns) free(d->v);
  free(d);
  A->data = NULL;
  return PETSC_SUCCESS;
}

PetscErrorCode MatCreateSeqDenseView(PetscInt m, PetscInt n, PetscScalar *data, PetscInt lda, Mat *A)
{
  Mat_SeqDense *d;
  PetscCall(MatCreate(A));
  PetscCall(PetscMallocZero(sizeof(Mat_SeqDense), &d));
  (*A)->m = m;
  (*A)->n = n;
  snprintf((*A)->hdr.type_name, sizeof((*A)->hdr.type_name), "%s", MATSEQDENSE);
  d->lda = lda;
  if (data) d->v = data;
  else {
    PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)lda * (size_t)(n ? n : 1), &d->v));
    d->owns = PETSC_TRUE;
  }
  (*A)->data = d;
  (*A)->ops->destroy = MatDestroy_SeqDense;
  return PETSC_SUCCESS;
}

PetscErrorCode MatCreateSeqDense(PetscInt m, PetscInt n, PetscScalar *data, Mat *A)
{
  return MatCreateSeqDenseView(m, n, data, m, A);
}

PetscErrorCode MatDenseGetArray(Mat A, PetscScalar **array)
{
  PetscCheck(A && !strcmp(A->hdr.type_name, MATSEQDENSE), PETSC_ERR_ARG_WRONG, "not a seqdense matrix");
  *array = ((Mat_SeqDense *)A->data)->v;
  return PETSC_SUCCESS;
}

PetscErrorCode MatDenseRestoreArray(Mat A, PetscScalar **array)
{
  (void)A;
  *array = NULL;
  return PETSC_SUCCESS;
}

PetscErrorCode MatDenseGetLDA(Mat A, PetscInt *lda)
{
  PetscCheck(A && !strcmp(A->hdr.type_name, MATSEQDENSE), PETSC_ERR_ARG_WRONG, "not a seqdense matrix");
  *lda = ((Mat_SeqDense *)A->data)->lda;
  return PETSC_SUCCESS;
}

/* ---------------------------------------------------------------- MATHIPCSR (device) */
/* FP32 copy of FP64 values, built on the device the first time an FP32 vector
   is multiplied (the mixed-precision phase reads half the matrix bytes) */
PetscErrorCode MatHIPCSRGetF32(Mat A, ks_csr *out)
{
  Mat_HIPCSR *h = (Mat_HIPCSR *)A->data;
  if (h->vprec == SLEPC_AMD_F32) { *out = h->csr; return PETSC_SUCCESS; }
  if (!h->csr32) KSCall(ks_csr_convert(h->csr, SlepcAMDContext(), KS_F32, &h->csr32));
  *out = h->csr32;
  return PETSC_SUCCESS;
}

static PetscErrorCode MatMult_HIPCSR(Mat A, Vec x, Vec y)
{
  Mat_HIPCSR *h = (Mat_HIPCSR *)A->data;
  ks_csr csr = h->csr;
  PetscCheck(h->csr, PETSC_ERR_ORDER, "MatSeqAIJSetPreallocationCSR must be called first");
  PetscCheck(x->prec == y->prec, PETSC_ERR_ARG_WRONG, "x and y precisions differ");
  if (x->prec == SLEPC_AMD_F32) PetscCall(MatHIPCSRGetF32(A, &csr));
  KSCall(ks_spmv(SlepcAMDContext(), csr, x->d, y->d, x->prec == SLEPC_AMD_F64 ? KS_F64 : KS_F32, ((Mat_HIPCSR *)A->data)->gate));
  return PETSC_SUCCESS;
}

/* y = A^T x with the explicit transpose, formed on the device the first time */
static PetscErrorCode MatMultTranspose_HIPCSR(Mat A, Vec x, Vec y)
{
  Mat_HIPCSR *h = (Mat_HIPCSR *)A->data;
  PetscCheck(h->csr, PETSC_ERR_ORDER, "MatSeqAIJSetPreallocationCSR must be called first");
  PetscCheck(x->prec == y->prec && x->prec == h->vprec, PETSC_ERR_SUP,
             "MatMultTranspose needs vectors in the precision of the matrix values");
  if (!h->csrT) KSCall(ks_csr_transpose(h->csr, SlepcAMDContext(), &h->csrT));
  KSCall(ks_spmv(SlepcAMDContext(), h->csrT, x->d, y->d, x->prec == SLEPC_AMD_F64 ? KS_F64 : KS_F32, 0));
  return PETSC_SUCCESS;
}

/* A new MATHIPCSR holding A^T, formed on the device.  The two-sided Krylov-Schur
   runs its left Arnoldi process on it; the reference wraps the operator in
   MatCreateHermitianTranspose instead (ks-twosided.c:138). */
PetscErrorCode MatHIPCSRCreateTranspose(Mat A, Mat *AT)
{
  Mat_HIPCSR *h, *t;
  PetscCheck(A && AT && !strcmp(A->hdr.type_name, MATHIPCSR), PETSC_ERR_ARG_WRONG, "not a %s matrix", MATHIPCSR);
  h = (Mat_HIPCSR *)A->data;
  PetscCheck(h->csr, PETSC_ERR_ORDER, "matrix has no values");
  PetscCall(MatCreate(AT));
  PetscCall(MatSetSizes(*AT, A->n, A->m, A->n, A->m));
  PetscCall(MatSetType(*AT, MATHIPCSR));
  t = (Mat_HIPCSR *)(*AT)->data;
  KSCall(ks_csr_transpose(h->csr, SlepcAMDContext(), &t->csr));
  t->vprec = h->vprec;
  return PETSC_SUCCESS;
}

static PetscErrorCode MatDestroy_HIPCSR(Mat A)
{
  Mat_HIPCSR *h = (Mat_HIPCSR *)A->data;
  if (h && h->csr) KSCall(ks_csr_destroy(h->csr));
  if (h && h->csr32) KSCall(ks_csr_destroy(h->csr32));
  if (h && h->csrT) KSCall(ks_csr_destroy(h->csrT));
  free(h);
  A->data = NULL;
  return PETSC_SUCCESS;
}

PetscErrorCode MatCreate_HIPCSR(Mat A)
{
  Mat_HIPCSR *h;
  PetscCall(PetscMallocZero(sizeof(Mat_HIPCSR), &h));
  A->data = h;
  A->ops->mult = MatMult_HIPCSR;
  A->ops->multtranspose = MatMultTranspose_HIPCSR;
  A->ops->destroy = MatDestroy_HIPCSR;
  return PETSC_SUCCESS;
}

static PetscErrorCode MatHIPCSRSetHost(Mat A, const PetscInt i[], const PetscInt j[], const void *v,
                                       SlepcAMDPrecision vprec)
{
  Mat_HIPCSR *h;
  PetscCheck(A && i, PETSC_ERR_ARG_NULL, "null argument");
  PetscCheck(!strcmp(A->hdr.type_name, MATHIPCSR), PETSC_ERR_ARG_WRONG, "matrix type is %s, expected %s", A->hdr.type_name, MATHIPCSR);
  PetscCheck(A->m > 0 && A->n > 0, PETSC_ERR_ORDER, "MatSetSizes must be called first");
  h = (Mat_HIPCSR *)A->data;
  if (h->csr) { KSCall(ks_csr_destroy(h->csr)); h->csr = NULL; }
  if (h->csr32) { KSCall(ks_csr_destroy(h->csr32)); h->csr32 = NULL; }
  if (h->csrT) { KSCall(ks_csr_destroy(h->csrT)); h->csrT = NULL; }
  KSCall(ks_csr_create(&h->csr, SlepcAMDContext(), A->m, A->n, (int64_t)i[A->m], i, j, v,
                       vprec == SLEPC_AMD_F64 ? KS_F64 : KS_F32));
  h->vprec = vprec;
  return PETSC_SUCCESS;
}

PetscErrorCode MatSeqAIJSetPreallocationCSR(Mat A, const PetscInt i[], const PetscInt j[], const PetscScalar v[])
{
  return MatHIPCSRSetHost(A, i, j, v, SLEPC_AMD_F64);
}

PetscErrorCode MatHIPCSRSetArraysF32(Mat A, const PetscInt i[], const PetscInt j[], const float v[])
{
  return MatHIPCSRSetHost(A, i, j, v, SLEPC_AMD_F32);
}

PetscErrorCode MatHIPCSRSetDeviceArrays(Mat A, PetscInt nnz, const PetscInt *d_i, const PetscInt *d_j,
                                        const void *d_v, SlepcAMDPrecision vprec)
{
  Mat_HIPCSR *h;
  PetscCheck(A && d_i, PETSC_ERR_ARG_NULL, "null argument");
  PetscCheck(!strcmp(A->hdr.type_name, MATHIPCSR), PETSC_ERR_ARG_WRONG, "matrix type is %s, expected %s", A->hdr.type_name, MATHIPCSR);
  h = (Mat_HIPCSR *)A->data;
  if (h->csr) { KSCall(ks_csr_destroy(h->csr)); h->csr = NULL; }
  if (h->csr32) { KSCall(ks_csr_destroy(h->csr32)); h->csr32 = NULL; }
  if (h->csrT) { KSCall(ks_csr_destroy(h->csrT)); h->csrT = NULL; }
  KSCall(ks_csr_create_device(&h->csr, SlepcAMDContext(), A->m, A->n, nnz, d_i, d_j, d_v,
                              vprec == SLEPC_AMD_F64 ? KS_F64 : KS_F32));
  h->vprec = vprec;
  return PETSC_SUCCESS;
}

PetscErrorCode MatHIPCSRGetInfo(Mat A, PetscInt *nnz, PetscInt *max_row, PetscInt *variant, PetscInt *nblocks)
{
  Mat_HIPCSR *h;
  int64_t z;
  int mr, var, nb;
  PetscCheck(A && !strcmp(A->hdr.type_name, MATHIPCSR), PETSC_ERR_ARG_WRONG, "not a %s matrix", MATHIPCSR);
  h = (Mat_HIPCSR *)A->data;
  PetscCheck(h->csr, PETSC_ERR_ORDER, "matrix has no values");
  KSCall(ks_csr_info(h->csr, NULL, &z, &mr, &var, &nb));
  if (nnz) *nnz = (PetscInt)z;
  if (max_row) *max_row = mr;
  if (variant) *variant = var;
  if (nblocks) *nblocks = nb;
  return PETSC_SUCCESS;
}

PetscErrorCode MatCreateHIPCSR(PetscInt n, const PetscInt i[], const PetscInt j[], const PetscScalar v[], Mat *A)
{
  PetscCall(MatCreate(A));
  PetscCall(MatSetSizes(*A, n, n, n, n));
  PetscCall(MatSetType(*A, MATHIPCSR));
  PetscCall(MatSeqAIJSetPreallocationCSR(*A, i, j, v));
  return PETSC_SUCCESS;
}
