/*
 * ds.c -- the dense projected problem, FP64 on the host with LAPACK.
 *
 *   interface    src/sys/classes/ds/interface/dsops.c:63-160 (state, dimensions),
 *                :232-275 (DSTruncate), :329-344 (DSGetTruncateSize_Default),
 *                :723-1000 (solve, sort, vectors, extra row);
 *                dspriv.c:172-270 (eigenvalue sorting, column permutation)
 *   DSHEP        src/sys/classes/ds/impls/hep/dshep.c:137-426,643-670
 *                (compact arrowhead + tridiagonal storage, Rutishauser
 *                reduction DSArrowTridiag, dsteqr, sort, extra row, truncation)
 *   DSNHEP       src/sys/classes/ds/impls/dsutil.c:21-171 (gehrd/orghr/hseqr,
 *                trexc selection sort), dsnhep.c:101-226,318-425 (trevc vectors,
 *                extra row, truncation)
 */
#include "slepcimpl.h"

#include <float.h>

/* LAPACK/BLAS, Fortran ABI, 32-bit integers */
extern void   dgemv_(const char *, const int *, const int *, const double *, const double *, const int *,
                     const double *, const int *, const double *, double *, const int *);
extern double dnrm2_(const int *, const double *, const int *);
extern void   dgetrf_(const int *, const int *, double *, const int *, int *, int *);
extern void   dgetrs_(const char *, const int *, const int *, const double *, const int *, const int *, double *,
                      const int *, int *);
extern void   dlartg_(double *, double *, double *, double *, double *);
extern void   drot_(const int *, double *, const int *, double *, const int *, const double *, const double *);
extern void   dsteqr_(const char *, const int *, double *, double *, double *, const int *, double *, int *);
extern void   dgehrd_(const int *, const int *, const int *, double *, const int *, double *, double *, const int *, int *);
extern void   dorghr_(const int *, const int *, const int *, double *, const int *, const double *, double *,
                      const int *, int *);
extern void   dhseqr_(const char *, const char *, const int *, const int *, const int *, double *, const int *,
                      double *, double *, double *, const int *, double *, const int *, int *);
extern void   dtrexc_(const char *, const int *, double *, const int *, double *, const int *, int *, int *, double *,
                      int *);
extern void   dtrevc_(const char *, const char *, int *, const int *, const double *, const int *, double *,
                      const int *, double *, const int *, const int *, int *, double *, int *);
extern void   dlascl_(const char *, const int *, const int *, const double *, const double *, const int *,
                      const int *, double *, const int *, int *);

#define LAPACK_CHECK(name, info) PetscCheck((info) == 0, PETSC_ERR_LIB, "Error in LAPACK subroutine %s: info=%d", name, (int)(info))

static SlepcFunctionList DSList;

PetscErrorCode DSRegister(const char name[], PetscErrorCode (*create)(DS))
{
  return SlepcFunctionListAdd(&DSList, name, (void *)create);
}

PetscErrorCode DSCreate(DS *ds)
{
  PetscCall(PetscMallocZero(sizeof(struct _p_DS), ds));
  (*ds)->sc.which = EPS_LARGEST_MAGNITUDE;
  return PETSC_SUCCESS;
}

PetscErrorCode DSSetType(DS ds, DSType type)
{
  PetscErrorCode (*create)(DS);
  PetscCheck(ds && type, PETSC_ERR_ARG_NULL, "null argument");
  if (!strcmp(ds->type, type)) return PETSC_SUCCESS;
  create = (PetscErrorCode(*)(DS))SlepcFunctionListFind(&DSList, type);
  PetscCheck(create, PETSC_ERR_ARG_WRONG, "unknown DS type %s", type);
  memset(&ds->ops, 0, sizeof(ds->ops));
  strcpy(ds->type, type);
  return create(ds);
}

static void DSFreeMats(DS ds)
{
  for (int i = 0; i < DS_NUM_MAT; i++) {
    if (ds->matview[i]) MatDestroy(&ds->matview[i]);
    PetscFree(ds->mat[i]);
  }
  PetscFree(ds->work);
  PetscFree(ds->rwork);
  PetscFree(ds->iwork);
  PetscFree(ds->perm);
  PetscFree(ds->wr2);
  PetscFree(ds->wi2);
}

PetscErrorCode DSAllocate(DS ds, PetscInt ld)
{
  PetscCheck(ds, PETSC_ERR_ARG_NULL, "null DS");
  PetscCheck(ld >= 1, PETSC_ERR_ARG_OUTOFRANGE, "leading dimension must be at least one");
  DSFreeMats(ds);
  ds->ld = ld;
  size_t sq = (size_t)ld * (size_t)ld;
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * sq, &ds->mat[DS_MAT_A]));
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)ld * 3, &ds->mat[DS_MAT_T]));
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * sq, &ds->mat[DS_MAT_Q]));
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * sq, &ds->mat[DS_MAT_X]));
  if (!strcmp(ds->type, DSNHEPTS)) {   /* the second problem of the two-sided variant */
    PetscCall(PetscMallocZero(sizeof(PetscScalar) * sq, &ds->mat[DS_MAT_B]));
    PetscCall(PetscMallocZero(sizeof(PetscScalar) * sq, &ds->mat[DS_MAT_Z]));
    PetscCall(PetscMallocZero(sizeof(PetscScalar) * sq, &ds->mat[DS_MAT_Y]));
    PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)ld, &ds->wr2));
    PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)ld, &ds->wi2));
  }
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (sq + 8 * (size_t)ld), &ds->work));
  PetscCall(PetscMallocZero(sizeof(PetscReal) * (2 * (size_t)ld + 2), &ds->rwork));
  PetscCall(PetscMallocZero(sizeof(int) * (size_t)ld, &ds->iwork));
  PetscCall(PetscMallocZero(sizeof(PetscInt) * (size_t)ld, &ds->perm));
  ds->n = ds->l = ds->k = ds->t = 0;
  ds->state = DS_STATE_RAW;
  return PETSC_SUCCESS;
}

PetscErrorCode DSGetLeadingDimension(DS ds, PetscInt *ld)
{
  *ld = ds->ld;
  return PETSC_SUCCESS;
}

/* DSSetDimensions (dsops.c:130-160) */
PetscErrorCode DSSetDimensions(DS ds, PetscInt n, PetscInt l, PetscInt k)
{
  PetscCheck(ds && ds->ld, PETSC_ERR_ORDER, "DSAllocate must be called first");
  if (n == PETSC_DETERMINE) ds->n = ds->extrarow ? ds->ld - 1 : ds->ld;
  else if (n != PETSC_CURRENT) {
    PetscCheck(n >= 0 && n <= ds->ld, PETSC_ERR_ARG_OUTOFRANGE, "Illegal value of n. Must be between 0 and ld");
    PetscCheck(!ds->extrarow || n < ds->ld, PETSC_ERR_ARG_OUTOFRANGE, "A value of n equal to ld leaves no room for extra row");
    ds->n = n;
  }
  ds->t = ds->n;
  if (l == PETSC_DETERMINE) ds->l = 0;
  else if (l != PETSC_CURRENT) {
    PetscCheck(l >= 0 && l <= ds->n, PETSC_ERR_ARG_OUTOFRANGE, "Illegal value of l. Must be between 0 and n");
    ds->l = l;
  }
  if (k == PETSC_DETERMINE) ds->k = ds->n / 2;
  else if (k != PETSC_CURRENT) ds->k = k;
  return PETSC_SUCCESS;
}

PetscErrorCode DSGetDimensions(DS ds, PetscInt *n, PetscInt *l, PetscInt *k, PetscInt *t)
{
  if (n) *n = ds->n;
  if (l) *l = ds->l;
  if (k) *k = ds->k;
  if (t) *t = ds->t;
  return PETSC_SUCCESS;
}

PetscErrorCode DSSetState(DS ds, DSStateType state)
{
  PetscCheck(state >= DS_STATE_RAW && state <= DS_STATE_TRUNCATED, PETSC_ERR_ARG_WRONG, "Wrong state");
  ds->state = state;
  return PETSC_SUCCESS;
}

PetscErrorCode DSGetState(DS ds, DSStateType *state)
{
  *state = ds->state;
  return PETSC_SUCCESS;
}

PetscErrorCode DSSetCompact(DS ds, PetscBool comp)
{
  ds->compact = comp;
  return PETSC_SUCCESS;
}

PetscErrorCode DSSetExtraRow(DS ds, PetscBool ext)
{
  ds->extrarow = ext;
  return PETSC_SUCCESS;
}

PetscErrorCode DSSetSortCriterion(DS ds, int which, PetscScalar target)
{
  ds->sc.which = which;
  ds->sc.target = target;
  return PETSC_SUCCESS;
}

PetscErrorCode DSGetArray(DS ds, DSMatType m, PetscScalar **a)
{
  PetscCheck(m >= 0 && m < DS_NUM_MAT && ds->mat[m], PETSC_ERR_ARG_WRONG, "matrix %d not available", (int)m);
  *a = ds->mat[m];
  return PETSC_SUCCESS;
}

PetscErrorCode DSRestoreArray(DS ds, DSMatType m, PetscScalar **a)
{
  (void)ds; (void)m;
  *a = NULL;
  return PETSC_SUCCESS;
}

/* DSGetMat (dsops.c:417-440): a seqdense view of the leading rows x cols */
PetscErrorCode DSGetMat(DS ds, DSMatType m, Mat *A)
{
  PetscInt rows, cols;
  PetscCheck(m >= 0 && m < DS_NUM_MAT && ds->mat[m], PETSC_ERR_ARG_WRONG, "matrix %d not available", (int)m);
  if (ds->state == DS_STATE_TRUNCATED && m >= DS_MAT_Q) rows = ds->t;
  else rows = ((m == DS_MAT_A || m == DS_MAT_B) && ds->extrarow) ? ds->n + 1 : ds->n;
  cols = (m == DS_MAT_T) ? 3 : ds->n;
  PetscCheck(rows && cols, PETSC_ERR_ORDER, "Must call DSSetDimensions() first");
  if (ds->matview[m]) PetscCall(MatDestroy(&ds->matview[m]));
  PetscCall(MatCreateSeqDenseView(rows, cols, ds->mat[m], ds->ld, &ds->matview[m]));
  *A = ds->matview[m];
  return PETSC_SUCCESS;
}

PetscErrorCode DSRestoreMat(DS ds, DSMatType m, Mat *A)
{
  (void)ds; (void)m;
  *A = NULL;
  return PETSC_SUCCESS;
}

/* The projected problems are at most a few hundred wide: their LAPACK/BLAS
   calls run on the calling thread only.  A threaded BLAS spends far longer
   forking and joining its team than computing at these sizes (measured: a
   48-wide DSHEP solve takes 0.3 ms on one thread, 2 ms on eight).  MKL's
   per-thread setting is used (by value), so other MKL users in the process
   keep theirs. */
extern int MKL_Set_Num_Threads_Local(int);
#define DS_SEQ_BEGIN int _ds_prev_threads = MKL_Set_Num_Threads_Local(1)
#define DS_SEQ_END (void)MKL_Set_Num_Threads_Local(_ds_prev_threads)

PetscErrorCode DSSolve(DS ds, PetscScalar eigr[], PetscScalar eigi[])
{
  PetscErrorCode ierr;
  PetscCheck(ds->ops.solve, PETSC_ERR_ORDER, "DS type not set");
  if (ds->state >= DS_STATE_CONDENSED) return PETSC_SUCCESS;
  DS_SEQ_BEGIN;
  ierr = ds->ops.solve(ds, eigr, eigi);
  DS_SEQ_END;
  PetscCall(ierr);
  ds->state = DS_STATE_CONDENSED;
  return PETSC_SUCCESS;
}

PetscErrorCode DSSort(DS ds, PetscScalar eigr[], PetscScalar eigi[], PetscScalar rr[], PetscScalar ri[], PetscInt *k)
{
  PetscErrorCode ierr;
  PetscCheck(ds->state >= DS_STATE_CONDENSED, PETSC_ERR_ORDER, "Must call DSSolve() first");
  PetscCheck(ds->state < DS_STATE_TRUNCATED, PETSC_ERR_ORDER, "Cannot sort a truncated DS");
  for (PetscInt i = 0; i < ds->n; i++) ds->perm[i] = i;
  DS_SEQ_BEGIN;
  ierr = ds->ops.sort(ds, eigr, eigi, rr, ri, k);
  DS_SEQ_END;
  return ierr;
}

PetscErrorCode DSUpdateExtraRow(DS ds)
{
  PetscErrorCode ierr;
  PetscCheck(ds->extrarow, PETSC_ERR_ARG_WRONGSTATE, "Should have called DSSetExtraRow");
  DS_SEQ_BEGIN;
  ierr = ds->ops.update(ds);
  DS_SEQ_END;
  return ierr;
}

PetscErrorCode DSVectors(DS ds, DSMatType mat, PetscInt *j, PetscReal *rnorm)
{
  PetscErrorCode ierr;
  PetscCheck(mat == DS_MAT_X || mat == DS_MAT_Y, PETSC_ERR_SUP, "only eigenvectors (DS_MAT_X, DS_MAT_Y) are implemented");
  PetscCheck(!rnorm || j, PETSC_ERR_ORDER, "Must give a value of j");
  DS_SEQ_BEGIN;
  ierr = ds->ops.vectors(ds, mat, j, rnorm);
  DS_SEQ_END;
  return ierr;
}

PetscErrorCode DSGetTruncateSize(DS ds, PetscInt l, PetscInt n, PetscInt *k)
{
  if (!ds->ops.gettruncatesize) return PETSC_SUCCESS;
  return ds->ops.gettruncatesize(ds, l ? l : ds->l, n ? n : ds->n, k);
}

PetscErrorCode DSTruncate(DS ds, PetscInt n, PetscBool trim)
{
  PetscErrorCode ierr;
  PetscCheck(n >= ds->l && n <= ds->n, PETSC_ERR_ARG_OUTOFRANGE, "Illegal value of n (%d). Must be between l (%d) and n (%d)", n, ds->l, ds->n);
  DS_SEQ_BEGIN;
  ierr = ds->ops.truncate(ds, n, trim);
  DS_SEQ_END;
  PetscCall(ierr);
  ds->state = trim ? DS_STATE_RAW : DS_STATE_TRUNCATED;
  return PETSC_SUCCESS;
}

PetscErrorCode DSTranslateHarmonic(DS ds, PetscScalar tau, PetscReal beta, PetscBool recover, PetscScalar *g,
                                   PetscReal *gamma)
{
  PetscErrorCode ierr;
  PetscCheck(ds->ops.transharm, PETSC_ERR_SUP, "DS type %s has no harmonic translation", ds->type);
  DS_SEQ_BEGIN;
  ierr = ds->ops.transharm(ds, tau, beta, recover, g, gamma);
  DS_SEQ_END;
  PetscCall(ierr);
  ds->state = DS_STATE_RAW;
  return PETSC_SUCCESS;
}

PetscErrorCode DSDestroy(DS *ds)
{
  if (!ds || !*ds) return PETSC_SUCCESS;
  DSFreeMats(*ds);
  free(*ds);
  *ds = NULL;
  return PETSC_SUCCESS;
}

/* DSSetIdentity (dspriv.c:377-395) */
static void DSSetIdentity(DS ds, PetscScalar *M)
{
  memset(M + (size_t)ds->l * ds->ld, 0, sizeof(PetscScalar) * (size_t)ds->ld * (size_t)(ds->n - ds->l));
  for (PetscInt i = ds->l; i < ds->n; i++) M[i + (size_t)i * ds->ld] = 1.0;
}

/* DSSortEigenvaluesReal_Private (dspriv.c:224-246) */
static void DSSortEigenvaluesReal(DS ds, const PetscReal *eig, PetscInt *perm)
{
  PetscInt n = ds->t, l = ds->l;
  for (PetscInt i = l + 1; i < n; i++) {
    PetscScalar re = eig[perm[i]];
    PetscInt j = i - 1;
    int result = SlepcSCCompare(&ds->sc, re, 0.0, eig[perm[j]], 0.0);
    while (result < 0 && j >= l) {
      PetscInt tmp = perm[j]; perm[j] = perm[j + 1]; perm[j + 1] = tmp; j--;
      if (j >= l) result = SlepcSCCompare(&ds->sc, re, 0.0, eig[perm[j]], 0.0);
    }
  }
}

/* DSPermuteColumns_Private (dspriv.c:248-270) */
static void DSPermuteColumns(DS ds, PetscInt istart, PetscInt iend, PetscInt n, PetscScalar *Q, PetscInt *perm)
{
  PetscInt ld = ds->ld;
  for (PetscInt i = istart; i < iend; i++) {
    PetscInt p = perm[i];
    if (p != i) {
      PetscInt j = i + 1;
      while (perm[j] != i) j++;
      perm[j] = p;
      perm[i] = i;
      for (PetscInt k = 0; k < n; k++) {
        PetscScalar r = Q[k + (size_t)p * ld];
        Q[k + (size_t)p * ld] = Q[k + (size_t)i * ld];
        Q[k + (size_t)i * ld] = r;
      }
    }
  }
}

/* ================================================================ DSHEP */
/* plane rotation of two columns of length m: (x, y) <- (c x + s y, c y - s x),
   the operation of BLAS drot with unit strides, inline for these short columns */
static inline void RotCols(int m, PetscScalar *x, PetscScalar *y, PetscReal c, PetscReal s)
{
  for (int i = 0; i < m; i++) {
    const PetscScalar xi = x[i], yi = y[i];
    x[i] = c * xi + s * yi;
    y[i] = c * yi - s * xi;
  }
}

/* DSArrowTridiag (dshep.c:221-262), Rutishauser's reduction of an arrowhead
   matrix to tridiagonal form with Givens rotations accumulated into Q. */
static void DSArrowTridiag(int n, PetscReal *d, PetscReal *e, PetscScalar *Q, int ld)
{
  PetscReal c, s, p, off, temp;
  if (n <= 2) return;
  for (int j = 0; j < n - 2; j++) {
    temp = e[j + 1];
    dlartg_(&temp, &e[j], &c, &s, &e[j + 1]);
    s = -s;
    temp = d[j + 1];
    e[j] = c * s * (temp - d[j]);
    d[j + 1] = s * s * d[j] + c * c * temp;
    d[j] = c * c * d[j] + s * s * temp;
    int j2 = j + 2;
    RotCols(j2, Q + (size_t)j * ld, Q + (size_t)(j + 1) * ld, c, s);
    for (int i = j - 1; i >= 0; i--) {
      off = -s * e[i];
      e[i] = c * e[i];
      temp = e[i + 1];
      dlartg_(&temp, &off, &c, &s, &e[i + 1]);
      s = -s;
      temp = (d[i] - d[i + 1]) * s - 2.0 * c * e[i];
      p = s * temp;
      d[i + 1] += p;
      d[i] -= p;
      e[i] = -e[i] - c * temp;
      RotCols(j2, Q + (size_t)i * ld, Q + (size_t)(i + 1) * ld, c, s);
    }
  }
}

/* DSSolve_HEP_QR (dshep.c:383-426) with the compact DSIntermediate_HEP (dshep.c:267-321) */
static PetscErrorCode DSSolve_HEP(DS ds, PetscScalar *wr, PetscScalar *wi)
{
  int n = ds->n, l = ds->l, ld = ds->ld, n1, info;
  PetscReal *d = ds->mat[DS_MAT_T], *e = d + ld;
  PetscScalar *Q = ds->mat[DS_MAT_Q];
  PetscCheck(ds->compact, PETSC_ERR_SUP, "DSHEP is implemented with compact storage");
  DSSetIdentity(ds, Q);
  n1 = PetscMax(0, ds->k - l + 1);
  if (ds->state < DS_STATE_INTERMEDIATE) DSArrowTridiag(n1, d + l, e + l, Q + l + (size_t)l * ld, ld);
  for (int i = 0; i < l; i++) wr[i] = d[i];
  n1 = n - l;
  dsteqr_("V", &n1, d + l, e + l, Q + l + (size_t)l * ld, &ld, ds->rwork, &info);
  LAPACK_CHECK("steqr", info);
  for (int i = l; i < n; i++) wr[i] = d[i];
  for (int i = 0; i < n - 1; i++) e[i] = 0.0;
  if (wi) for (int i = l; i < n; i++) wi[i] = 0.0;
  return PETSC_SUCCESS;
}

/* DSSort_HEP (dshep.c:323-347) */
static PetscErrorCode DSSort_HEP(DS ds, PetscScalar *wr, PetscScalar *wi, PetscScalar *rr, PetscScalar *ri, PetscInt *k)
{
  PetscInt n = ds->n, l = ds->l;
  PetscReal *d = ds->mat[DS_MAT_T];
  (void)wi; (void)ri; (void)k;
  PetscCheck(!rr, PETSC_ERR_SUP, "arbitrary selection is not implemented");
  DSSortEigenvaluesReal(ds, d, ds->perm);
  for (PetscInt i = l; i < n; i++) wr[i] = d[ds->perm[i]];
  DSPermuteColumns(ds, l, n, n, ds->mat[DS_MAT_Q], ds->perm);
  for (PetscInt i = l; i < n; i++) d[i] = wr[i];
  return PETSC_SUCCESS;
}

/* DSUpdateExtraRow_HEP, compact (dshep.c:349-381) */
static PetscErrorCode DSUpdateExtraRow_HEP(DS ds)
{
  PetscInt n = ds->n, ld = ds->ld;
  PetscReal *e = ds->mat[DS_MAT_T] + ld, beta = e[n - 1];
  const PetscScalar *Q = ds->mat[DS_MAT_Q];
  for (PetscInt i = 0; i < n; i++) e[i] = beta * Q[n - 1 + (size_t)i * ld];
  ds->k = n;
  return PETSC_SUCCESS;
}

/* DSVectors_HEP (dshep.c:137-172) */
static PetscErrorCode DSVectors_HEP(DS ds, DSMatType mat, PetscInt *j, PetscReal *rnorm)
{
  PetscInt ld = ds->ld;
  PetscScalar *X = ds->mat[mat], *Q = ds->mat[DS_MAT_Q];
  if (j) {
    if (ds->state >= DS_STATE_CONDENSED) {
      memcpy(X + (size_t)(*j) * ld, Q + (size_t)(*j) * ld, sizeof(PetscScalar) * (size_t)ld);
      if (rnorm) *rnorm = fabs(Q[ds->n - 1 + (size_t)(*j) * ld]);
    } else {
      memset(X + (size_t)(*j) * ld, 0, sizeof(PetscScalar) * (size_t)ld);
      X[*j + (size_t)(*j) * ld] = 1.0;
      if (rnorm) *rnorm = 0.0;
    }
  } else {
    if (ds->state >= DS_STATE_CONDENSED) memcpy(X, Q, sizeof(PetscScalar) * (size_t)ld * (size_t)ld);
    else DSSetIdentity(ds, X);
  }
  return PETSC_SUCCESS;
}

/* DSTruncate_HEP (dshep.c:643-670); compact storage keeps the arrow in T */
static PetscErrorCode DSTruncate_HEP(DS ds, PetscInt n, PetscBool trim)
{
  if (trim) {
    ds->l = 0;
    ds->k = 0;
    ds->n = n;
    ds->t = n;
  } else {
    ds->k = ds->extrarow ? n : 0;
    ds->t = ds->n;
    ds->n = n;
  }
  return PETSC_SUCCESS;
}

PetscErrorCode DSCreate_HEP(DS ds)
{
  ds->ops.solve = DSSolve_HEP;
  ds->ops.sort = DSSort_HEP;
  ds->ops.update = DSUpdateExtraRow_HEP;
  ds->ops.vectors = DSVectors_HEP;
  ds->ops.truncate = DSTruncate_HEP;
  ds->ops.gettruncatesize = NULL;
  return PETSC_SUCCESS;
}

/* ================================================================ DSNHEP */
/* The real-Schur routines take the matrix pair they work on, so that DSNHEPTS
   can run them on (A, Q) and on (B, Z). */

/* DSSolve_NHEP_Private (dsutil.c:21-87): Hessenberg reduction when the state is
   RAW, then the real Schur form with dhseqr */
static PetscErrorCode NHEPSolve_Private(DS ds, DSMatType mA, DSMatType mQ, PetscScalar *wr, PetscScalar *wi)
{
  int n = ds->n, ld = ds->ld, ilo = ds->l + 1, lwork = 6 * ld, info;
  PetscScalar *A = ds->mat[mA], *Q = ds->mat[mQ], *tau = ds->work, *work = ds->work + ld;
  PetscCheck(wi, PETSC_ERR_ARG_NULL, "DSNHEP needs the imaginary parts array");
  memset(Q, 0, sizeof(PetscScalar) * (size_t)ld * (size_t)ld);
  for (int i = 0; i < n; i++) Q[i + (size_t)i * ld] = 1.0;
  if (n == 1) {
    wr[0] = A[0];
    wi[0] = 0.0;
    return PETSC_SUCCESS;
  }
  if (ds->state < DS_STATE_INTERMEDIATE) {
    dgehrd_(&n, &ilo, &n, A, &ld, tau, work, &lwork, &info);
    LAPACK_CHECK("gehrd", info);
    for (int j = 0; j < n - 1; j++)
      for (int i = j + 2; i < n; i++) {
        Q[i + (size_t)j * ld] = A[i + (size_t)j * ld];
        A[i + (size_t)j * ld] = 0.0;
      }
    dorghr_(&n, &ilo, &n, Q, &ld, tau, work, &lwork, &info);
    LAPACK_CHECK("orghr", info);
  }
  dhseqr_("S", "V", &n, &ilo, &n, A, &ld, wr, wi, Q, &ld, work, &lwork, &info);
  for (int j = 0; j < ds->l; j++) {
    if (j == n - 1 || A[j + 1 + (size_t)j * ld] == 0.0) {
      wr[j] = A[j + (size_t)j * ld];
      wi[j] = 0.0;
    } else {
      wr[j] = A[j + (size_t)j * ld];
      wr[j + 1] = A[j + (size_t)j * ld];
      wi[j] = sqrt(fabs(A[j + 1 + (size_t)j * ld])) * sqrt(fabs(A[j + (size_t)(j + 1) * ld]));
      wi[j + 1] = -wi[j];
      j++;
    }
  }
  LAPACK_CHECK("hseqr", info);
  return PETSC_SUCCESS;
}

/* eigenvalues from the diagonal blocks of a real Schur form, rows j0..n-1 */
static void SchurEigenvalues(DS ds, const PetscScalar *T, PetscInt j0, PetscScalar *wr, PetscScalar *wi)
{
  PetscInt n = ds->n, ld = ds->ld;
  for (PetscInt j = j0; j < n; j++) {
    wr[j] = T[j + (size_t)j * ld];
    if (j < n - 1 && T[j + 1 + (size_t)j * ld] != 0.0) {
      wi[j] = sqrt(fabs(T[j + 1 + (size_t)j * ld])) * sqrt(fabs(T[j + (size_t)(j + 1) * ld]));
      wr[j + 1] = wr[j];
      wi[j + 1] = -wi[j];
      j++;
    } else wi[j] = 0.0;
  }
}

/* DSSort_NHEP_Total (dsutil.c:93-171): selection sort with dtrexc swaps */
static PetscErrorCode NHEPSortTotal_Private(DS ds, DSMatType mT, DSMatType mQ, PetscScalar *wr, PetscScalar *wi)
{
  int n = ds->n, ld = ds->ld, info;
  PetscScalar *T = ds->mat[mT], *Q = ds->mat[mQ];
  for (int i = ds->l; i < n - 1; i++) {
    PetscScalar re = wr[i], im = wi[i];
    int pos = 0, j = i + 1;
    if (im != 0) j = i + 2;
    for (; j < n; j++) {
      if (SlepcSCCompare(&ds->sc, re, im, wr[j], wi[j]) > 0) {
        re = wr[j];
        im = wi[j];
        pos = j;
      }
      if (wi[j] != 0) j++;
    }
    if (pos) {
      int ifst = pos + 1, ilst = i + 1;
      dtrexc_("V", &n, T, &ld, Q, &ld, &ifst, &ilst, ds->work, &info);
      LAPACK_CHECK("trexc", info);
      SchurEigenvalues(ds, T, i, wr, wi);
    }
    if (wi[i] != 0) i++;
  }
  return PETSC_SUCCESS;
}

/* DSSortWithPermutation_NHEP_Private (dsutil.c:177-236): move the block at perm[i]
   to position i, for i = l, l+1, ... */
static PetscErrorCode NHEPSortPermutation_Private(DS ds, PetscInt *perm, DSMatType mT, DSMatType mQ, PetscScalar *wr,
                                                  PetscScalar *wi)
{
  int n = ds->n, ld = ds->ld, info;
  PetscScalar *T = ds->mat[mT], *Q = ds->mat[mQ];
  for (PetscInt i = ds->l; i < n - 1; i++) {
    PetscInt pos = perm[i];
    int inc = (pos < n - 1 && T[pos + 1 + (size_t)pos * ld] != 0.0) ? 2 : 1;
    if (pos != i) {
      PetscCheck((pos == 0 || T[pos + (size_t)(pos - 1) * ld] == 0.0 || perm[i + 1] == pos - 1) &&
                     (T[pos + 1 + (size_t)pos * ld] == 0.0 || perm[i + 1] == pos + 1),
                 PETSC_ERR_FP, "Invalid permutation due to a 2x2 block at position %d", pos);
      int ifst = pos + 1, ilst = i + 1;
      dtrexc_("V", &n, T, &ld, Q, &ld, &ifst, &ilst, ds->work, &info);
      LAPACK_CHECK("trexc", info);
      for (PetscInt j = i + 1; j < n; j++)
        if (perm[j] >= i && perm[j] < pos) perm[j] += inc;
      perm[i] = i;
      if (inc == 2) perm[i + 1] = i + 1;
    }
    if (inc == 2) i++;
  }
  SchurEigenvalues(ds, T, ds->l, wr, wi);
  return PETSC_SUCCESS;
}

/* DSUpdateExtraRow_NHEP (dsnhep.c:318-340): extra row <- extra row * Q */
static void NHEPUpdateExtraRow_Private(DS ds, DSMatType mA, DSMatType mQ)
{
  int n = ds->n, ld = ds->ld, inc = 1;
  double one = 1.0, zero = 0.0, *x = ds->work, *y = ds->work + ld;
  PetscScalar *A = ds->mat[mA];
  for (int i = 0; i < n; i++) x[i] = A[n + (size_t)i * ld];
  dgemv_("C", &n, &n, &one, ds->mat[mQ], &ld, x, &inc, &zero, y, &inc);
  for (int i = 0; i < n; i++) A[n + (size_t)i * ld] = y[i];
}

/* DSVectors_NHEP_Eigen_Some (dsnhep.c:101-167): eigenvector k of the Schur form
   mA, back-transformed with mQ into column k of mX */
static PetscErrorCode NHEPVectorsSome_Private(DS ds, DSMatType mA, DSMatType mQ, DSMatType mX, PetscInt *k,
                                              PetscReal *rnorm)
{
  int n = ds->n, ld = ds->ld, mm, mout, info, inc = 1, cols = 1, zero = 0, iscomplex = 0;
  double sone = 1.0, szero = 0.0, norm, done = 1.0;
  int *select = ds->iwork;
  PetscScalar *A = ds->mat[mA], *Y = ds->mat[mX] + (size_t)(*k) * ld;
  for (int i = 0; i < n; i++) select[i] = 0;
  select[*k] = 1;
  if (*k < n - 1 && A[*k + 1 + (size_t)(*k) * ld] != 0.0) iscomplex = 1;
  mm = iscomplex ? 2 : 1;
  if (iscomplex) select[*k + 1] = 1;
  dtrevc_("R", "S", select, &n, A, &ld, Y, &ld, Y, &ld, &mm, &mout, ds->work, &info);
  LAPACK_CHECK("trevc", info);
  PetscCheck(mout == mm, PETSC_ERR_ARG_WRONG, "Inconsistent arguments");
  if (ds->state >= DS_STATE_CONDENSED) {
    double *w = ds->work;
    memcpy(w, Y, sizeof(double) * (size_t)mout * (size_t)ld);
    dgemv_("N", &n, &n, &sone, ds->mat[mQ], &ld, w, &inc, &szero, Y, &inc);
    if (iscomplex) dgemv_("N", &n, &n, &sone, ds->mat[mQ], &ld, w + ld, &inc, &szero, Y + ld, &inc);
    norm = dnrm2_(&n, Y, &inc);
    if (iscomplex) {
      norm = SlepcAbsEigenvalue(norm, dnrm2_(&n, Y + ld, &inc));
      cols = 2;
    }
    dlascl_("G", &zero, &zero, &norm, &done, &n, &cols, Y, &ld, &info);
    LAPACK_CHECK("lascl", info);
  }
  if (iscomplex) (*k)++;
  if (rnorm) *rnorm = iscomplex ? SlepcAbsEigenvalue(Y[n - 1], Y[n - 1 + ld]) : fabs(Y[n - 1]);
  return PETSC_SUCCESS;
}

/* DSVectors_NHEP_Eigen_All (dsnhep.c:168-226) */
static PetscErrorCode NHEPVectorsAll_Private(DS ds, DSMatType mA, DSMatType mQ, DSMatType mX)
{
  int n = ds->n, ld = ds->ld, mout, info, inc = 1, cols, zero = 0;
  double norm, done = 1.0;
  PetscScalar *A = ds->mat[mA], *X = ds->mat[mX];
  const char *back = "A";
  if (ds->state >= DS_STATE_CONDENSED) {
    back = "B";
    memcpy(X, ds->mat[mQ], sizeof(PetscScalar) * (size_t)ld * (size_t)ld);
  }
  dtrevc_("R", back, NULL, &n, A, &ld, NULL, &ld, X, &ld, &n, &mout, ds->work, &info);
  LAPACK_CHECK("trevc", info);
  for (int i = 0; i < n; i++) {
    int iscomplex = (i < n - 1 && A[i + 1 + (size_t)i * ld] != 0.0);
    cols = 1;
    norm = dnrm2_(&n, X + (size_t)i * ld, &inc);
    if (iscomplex) {
      norm = SlepcAbsEigenvalue(norm, dnrm2_(&n, X + (size_t)(i + 1) * ld, &inc));
      cols = 2;
    }
    dlascl_("G", &zero, &zero, &norm, &done, &n, &cols, X + (size_t)i * ld, &ld, &info);
    LAPACK_CHECK("lascl", info);
    if (iscomplex) i++;
  }
  return PETSC_SUCCESS;
}

static PetscErrorCode DSSolve_NHEP(DS ds, PetscScalar *wr, PetscScalar *wi)
{
  return NHEPSolve_Private(ds, DS_MAT_A, DS_MAT_Q, wr, wi);
}

static PetscErrorCode DSSort_NHEP(DS ds, PetscScalar *wr, PetscScalar *wi, PetscScalar *rr, PetscScalar *ri, PetscInt *kk)
{
  (void)ri; (void)kk;
  PetscCheck(!rr || rr == wr, PETSC_ERR_SUP, "arbitrary selection is not implemented");
  return NHEPSortTotal_Private(ds, DS_MAT_A, DS_MAT_Q, wr, wi);
}

static PetscErrorCode DSUpdateExtraRow_NHEP(DS ds)
{
  NHEPUpdateExtraRow_Private(ds, DS_MAT_A, DS_MAT_Q);
  ds->k = ds->n;
  return PETSC_SUCCESS;
}

static PetscErrorCode DSVectors_NHEP(DS ds, DSMatType mat, PetscInt *j, PetscReal *rnorm)
{
  PetscCheck(mat == DS_MAT_X, PETSC_ERR_SUP, "DSNHEP computes right vectors only");
  if (j) return NHEPVectorsSome_Private(ds, DS_MAT_A, DS_MAT_Q, DS_MAT_X, j, rnorm);
  return NHEPVectorsAll_Private(ds, DS_MAT_A, DS_MAT_Q, DS_MAT_X);
}

/* DSGetTruncateSize_Default (dsops.c:329-344): do not split a 2x2 block */
static PetscErrorCode DSGetTruncateSize_NHEP(DS ds, PetscInt l, PetscInt n, PetscInt *k)
{
  PetscScalar val = ds->mat[DS_MAT_A][(l + *k) + (size_t)(l + *k - 1) * ds->ld];
  if (val != 0.0) {
    if (l + *k < n - 1) (*k)++;
    else (*k)--;
  }
  return PETSC_SUCCESS;
}

/* DSTruncate_NHEP (dsnhep.c:394-425); DSTruncate_NHEPTS (dsnhepts.c:369-398) does the
   same to B */
static PetscErrorCode TruncateNHEP_Private(DS ds, PetscInt n, PetscBool trim, PetscBool withB)
{
  PetscInt ld = ds->ld, l = ds->l;
  PetscScalar *A = ds->mat[DS_MAT_A], *B = withB ? ds->mat[DS_MAT_B] : NULL;
  if (trim) {
    if (ds->extrarow)
      for (PetscInt i = l; i < ds->n; i++) {
        A[ds->n + (size_t)i * ld] = 0.0;
        if (B) B[ds->n + (size_t)i * ld] = 0.0;
      }
    ds->l = 0;
    ds->k = 0;
    ds->n = n;
    ds->t = n;
  } else {
    if (ds->extrarow && ds->k == ds->n) {
      for (PetscInt i = l; i < n; i++) {
        A[n + (size_t)i * ld] = A[ds->n + (size_t)i * ld];
        if (B) B[n + (size_t)i * ld] = B[ds->n + (size_t)i * ld];
      }
      for (PetscInt i = l; i < ds->n; i++) {
        A[ds->n + (size_t)i * ld] = 0.0;
        if (B) B[ds->n + (size_t)i * ld] = 0.0;
      }
    }
    ds->k = ds->extrarow ? n : 0;
    ds->t = ds->n;
    ds->n = n;
  }
  return PETSC_SUCCESS;
}

static PetscErrorCode DSTruncate_NHEP(DS ds, PetscInt n, PetscBool trim)
{
  return TruncateNHEP_Private(ds, n, trim, PETSC_FALSE);
}

/* DSTranslateHarmonic_NHEP (dsnhep.c:466-535).  Translate: g = (A - tau I)^-T b
   with b = beta e_n, A(:, n-1) += beta g.  Recover (after the sort, with l = the
   converged count and k = the kept count): ghat = -Q(:, 0:l+k)^T g,
   A(0:l+k, l:l+k) += ghat Q(n-1, l:l+k) beta, g += Q(:, 0:l+k) ghat.  Both
   return gamma = sqrt(1 + g^T g); recover scales the extra row by it. */
static PetscErrorCode DSTranslateHarmonic_NHEP(DS ds, PetscScalar tau, PetscReal beta, PetscBool recover,
                                               PetscScalar *gin, PetscReal *gammaout)
{
  int n = ds->n, ld = ds->ld, one = 1, info;
  PetscScalar *A = ds->mat[DS_MAT_A], *g = gin ? gin : ds->work;
  PetscReal gamma = 1.0;
  if (!recover) {
    double *B = NULL;
    int *ipiv = ds->iwork;
    PetscCall(PetscMallocZero(sizeof(double) * (size_t)ld * (size_t)ld, &B));
    memcpy(B, A, sizeof(double) * (size_t)ld * (size_t)ld);
    for (int i = 0; i < n; i++) g[i] = 0.0;
    g[n - 1] = beta;
    for (int i = 0; i < n; i++) B[i + (size_t)i * ld] -= tau;
    dgetrf_(&n, &n, B, &ld, ipiv, &info);
    if (info) { free(B); LAPACK_CHECK("getrf", info); }
    dgetrs_("C", &n, &one, B, &ld, ipiv, g, &ld, &info);
    free(B);
    LAPACK_CHECK("getrs", info);
    for (int i = 0; i < n; i++) A[i + (size_t)(n - 1) * ld] += g[i] * beta;
  } else {
    const PetscScalar *Q = ds->mat[DS_MAT_Q];
    double *ghat = ds->work + 2 * (size_t)ld, dmone = -1.0, done = 1.0, dzero = 0.0;
    int ncol = ds->l + ds->k;
    dgemv_("C", &n, &ncol, &dmone, Q, &ld, g, &one, &dzero, ghat, &one);
    for (PetscInt i = 0; i < ds->l + ds->k; i++)
      for (PetscInt j = ds->l; j < ds->l + ds->k; j++)
        A[i + (size_t)j * ld] += ghat[i] * Q[n - 1 + (size_t)j * ld] * beta;
    dgemv_("N", &n, &ncol, &done, Q, &ld, ghat, &one, &done, g, &one);
  }
  if (gammaout || (recover && ds->extrarow)) gamma = SlepcAbsEigenvalue(1.0, dnrm2_(&n, g, &one));
  if (gammaout) *gammaout = gamma;
  if (recover && ds->extrarow)
    for (PetscInt j = ds->l; j < ds->l + ds->k; j++) A[ds->n + (size_t)j * ld] *= gamma;
  return PETSC_SUCCESS;
}

PetscErrorCode DSCreate_NHEP(DS ds)
{
  ds->ops.transharm = DSTranslateHarmonic_NHEP;
  ds->ops.solve = DSSolve_NHEP;
  ds->ops.sort = DSSort_NHEP;
  ds->ops.update = DSUpdateExtraRow_NHEP;
  ds->ops.vectors = DSVectors_NHEP;
  ds->ops.truncate = DSTruncate_NHEP;
  ds->ops.gettruncatesize = DSGetTruncateSize_NHEP;
  return PETSC_SUCCESS;
}

/* ================================================================ DSNHEPTS */
/* DSNHEPTS (dsnhepts.c): two related problems, A X = X Lambda from the Arnoldi
   process on the operator and B Y = Y Lambda' from the one on its transpose.
   After the solve A and B hold real Schur forms A Q = Q T, B Z = Z S; the sort
   orders A's eigenvalues by the criterion, B's the same way, then permutes B's
   blocks so that B's eigenvalue i is the one nearest A's eigenvalue i. */

/* DSSolve_NHEPTS (dsnhepts.c:241-252) */
static PetscErrorCode DSSolve_NHEPTS(DS ds, PetscScalar *wr, PetscScalar *wi)
{
  PetscCall(NHEPSolve_Private(ds, DS_MAT_A, DS_MAT_Q, wr, wi));
  return NHEPSolve_Private(ds, DS_MAT_B, DS_MAT_Z, ds->wr2, ds->wi2);
}

/* DSSort_NHEPTS (dsnhepts.c:188-224) */
static PetscErrorCode DSSort_NHEPTS(DS ds, PetscScalar *wr, PetscScalar *wi, PetscScalar *rr, PetscScalar *ri, PetscInt *kk)
{
  PetscInt n = ds->n, cont = 0, id = 0, *p = NULL, *idx = NULL, *idx2 = NULL;
  const PetscReal tol = sqrt(DBL_EPSILON);
  PetscScalar *wr2 = ds->wr2, *wi2 = ds->wi2;
  PetscErrorCode e;
  (void)ri; (void)kk;
  PetscCheck(!rr || rr == wr, PETSC_ERR_SUP, "arbitrary selection is not implemented");
  PetscCall(NHEPSortTotal_Private(ds, DS_MAT_A, DS_MAT_Q, wr, wi));
  PetscCall(NHEPSortTotal_Private(ds, DS_MAT_B, DS_MAT_Z, wr2, wi2));
  PetscCall(PetscMallocZero(sizeof(PetscInt) * (size_t)ds->ld, &p));
  PetscCall(PetscMallocZero(sizeof(PetscInt) * (size_t)ds->ld, &idx));
  PetscCall(PetscMallocZero(sizeof(PetscInt) * (size_t)ds->ld, &idx2));
  for (PetscInt i = 0; i < n; i++) {
    if (SlepcAbsEigenvalue(wr2[i] - wr[i], wi2[i] - wi[i]) > tol) { idx2[cont] = i; idx[cont++] = i; }
    p[i] = -1;
  }
  e = PETSC_SUCCESS;
  if (cont) {
    for (PetscInt i = 0; i < cont; i++) {
      PetscReal t = DBL_MAX, s;
      for (PetscInt j = 0; j < cont; j++)
        if (idx2[j] != -1 && (s = SlepcAbsEigenvalue(wr2[idx[j]] - wr[idx[i]], wi2[idx[j]] - wi[idx[i]])) < t) {
          id = j;
          t = s;
        }
      p[idx[i]] = idx[id];
      idx2[id] = -1;
    }
    for (PetscInt i = 0; i < n; i++)
      if (p[i] == -1) p[i] = i;
    e = NHEPSortPermutation_Private(ds, p, DS_MAT_B, DS_MAT_Z, wr2, wi2);
  }
  free(p);
  free(idx);
  free(idx2);
  return e;
}

/* DSUpdateExtraRow_NHEPTS (dsnhepts.c:226-239) */
static PetscErrorCode DSUpdateExtraRow_NHEPTS(DS ds)
{
  NHEPUpdateExtraRow_Private(ds, DS_MAT_A, DS_MAT_Q);
  NHEPUpdateExtraRow_Private(ds, DS_MAT_B, DS_MAT_Z);
  ds->k = ds->n;
  return PETSC_SUCCESS;
}

/* DSVectors_NHEPTS (dsnhepts.c:162-186): X from (A, Q), Y from (B, Z) */
static PetscErrorCode DSVectors_NHEPTS(DS ds, DSMatType mat, PetscInt *j, PetscReal *rnorm)
{
  const PetscBool left = mat == DS_MAT_Y;
  PetscCheck(mat == DS_MAT_X || mat == DS_MAT_Y, PETSC_ERR_ARG_OUTOFRANGE, "Invalid mat parameter");
  if (j) return NHEPVectorsSome_Private(ds, left ? DS_MAT_B : DS_MAT_A, left ? DS_MAT_Z : DS_MAT_Q, mat, j, rnorm);
  return NHEPVectorsAll_Private(ds, left ? DS_MAT_B : DS_MAT_A, left ? DS_MAT_Z : DS_MAT_Q, mat);
}

/* DSGetTruncateSize_NHEPTS (dsnhepts.c:349-367) */
static PetscErrorCode DSGetTruncateSize_NHEPTS(DS ds, PetscInt l, PetscInt n, PetscInt *k)
{
  const size_t at = (size_t)(l + *k) + (size_t)(l + *k - 1) * ds->ld;
  if (ds->mat[DS_MAT_A][at] != 0.0 || ds->mat[DS_MAT_B][at] != 0.0) {
    if (l + *k < n - 1) (*k)++;
    else (*k)--;
  }
  return PETSC_SUCCESS;
}

static PetscErrorCode DSTruncate_NHEPTS(DS ds, PetscInt n, PetscBool trim)
{
  return TruncateNHEP_Private(ds, n, trim, PETSC_TRUE);
}

PetscErrorCode DSCreate_NHEPTS(DS ds)
{
  ds->ops.solve = DSSolve_NHEPTS;
  ds->ops.sort = DSSort_NHEPTS;
  ds->ops.update = DSUpdateExtraRow_NHEPTS;
  ds->ops.vectors = DSVectors_NHEPTS;
  ds->ops.truncate = DSTruncate_NHEPTS;
  ds->ops.gettruncatesize = DSGetTruncateSize_NHEPTS;
  return PETSC_SUCCESS;
}
