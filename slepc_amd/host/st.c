/*
 * st.c -- spectral transformations STSHIFT (sigma = 0) and STSINVERT
 * (src/sys/classes/st/impls/shift/shift.c:79-97, st/impls/sinvert/sinvert.c:16-105,
 * st/interface/stsolve.c:16-60, 244-359).
 *
 * STSHIFT with sigma = 0 on a standard problem hands out A itself
 * (STMatMAXPY_Private leaves T[0] = A, stsolve.c:598-610); eigenvalues need no
 * back-transform.  STSINVERT's operator is (A - sigma I)^-1, applied by an inner
 * solve (KSP) on the device: by default, as in the reference (stsles.c:47-54),
 * the direct solve KSPPREONLY + PCLU (ks_lu_create / ks_lu_solve, a dense LU
 * with partial pivoting); or preconditioned conjugate gradients (ks_pcg_solve,
 * KSPCG) or BiCGStab (ks_bcgs_solve, KSPBCGS, for a shift inside the
 * spectrum), with a Jacobi or no preconditioner, or with BiCGStab the ILU(0)
 * factors of A - sigma I (ks_ilu0_create, PCILU).  Its eigenvalues theta map
 * back to lambda = sigma + 1/theta (STBackTransform_Sinvert, sinvert.c:32-50).
 * The default shift is the EPS target (STSetDefaultShift, from EPSSetTarget).
 */
#include "slepcimpl.h"

PetscErrorCode STCreate(ST *st)
{
  PetscCall(PetscMallocZero(sizeof(struct _p_ST), st));
  strcpy((*st)->type, STSHIFT);
  return PETSC_SUCCESS;
}

PetscErrorCode STSetMatrices(ST st, PetscInt n, Mat A[])
{
  PetscCheck(st, PETSC_ERR_ARG_NULL, "null ST");
  PetscCheck(n == 1 && A && A[0], PETSC_ERR_SUP, "only standard eigenproblems (one matrix) are supported");
  st->A = A[0];
  st->setup = PETSC_FALSE;
  return PETSC_SUCCESS;
}

PetscErrorCode STSetType(ST st, STType type)
{
  PetscCheck(st && type, PETSC_ERR_ARG_NULL, "null argument");
  PetscCheck(!strcmp(type, STSHIFT) || !strcmp(type, STSINVERT), PETSC_ERR_SUP, "ST type %s is not available", type);
  if (strcmp(st->type, type)) st->setup = PETSC_FALSE;
  strcpy(st->type, type);
  return PETSC_SUCCESS;
}

PetscErrorCode STGetType(ST st, STType *type)
{
  *type = st->type;
  return PETSC_SUCCESS;
}

/* STSetShift (stfunc.c:612-642) */
PetscErrorCode STSetShift(ST st, PetscScalar shift)
{
  if (st->sigma != shift || !st->sigma_set) st->setup = PETSC_FALSE;
  st->sigma = shift;
  st->sigma_set = PETSC_TRUE;
  return PETSC_SUCCESS;
}

PetscErrorCode STGetShift(ST st, PetscScalar *shift)
{
  *shift = st->sigma;
  return PETSC_SUCCESS;
}

/* STSetDefaultShift (stfunc.c:704-714): the shift taken when none was set */
PetscErrorCode STSetDefaultShift(ST st, PetscScalar defaultshift)
{
  if (st->defsigma != defaultshift) st->setup = PETSC_FALSE;
  st->defsigma = defaultshift;
  return PETSC_SUCCESS;
}

/* ------------------------------------------------------------ KSP / PC */
PetscErrorCode STGetKSP(ST st, KSP *ksp)
{
  if (!st->ksp) {
    PetscCall(PetscMallocZero(sizeof(struct _p_KSP), &st->ksp));
    strcpy(st->ksp->type, KSPPREONLY);   /* STSetDefaultKSP_Default (stsles.c:47-54) */
    strcpy(st->ksp->pc.type, PCLU);
    st->ksp->rtol = SLEPC_DEFAULT_TOL;   /* stsles.c:407 */
    st->ksp->abstol = 1e-50;
    st->ksp->maxits = 10000;
  }
  *ksp = st->ksp;
  return PETSC_SUCCESS;
}

PetscErrorCode KSPSetType(KSP ksp, KSPType type)
{
  PetscCheck(ksp && type, PETSC_ERR_ARG_NULL, "null argument");
  PetscCheck(!strcmp(type, KSPCG) || !strcmp(type, KSPBCGS) || !strcmp(type, KSPPREONLY), PETSC_ERR_SUP,
             "KSP type %s is not available (the device inner solvers are %s, %s and %s)", type, KSPPREONLY, KSPCG,
             KSPBCGS);
  strcpy(ksp->type, type);
  /* the LU factors serve only as the whole solve here (KSPPREONLY): an
     iterative solver left with the default PCLU takes Jacobi instead */
  if (!ksp->pc.set) strcpy(ksp->pc.type, !strcmp(type, KSPPREONLY) ? PCLU : PCJACOBI);
  return PETSC_SUCCESS;
}

PetscErrorCode KSPGetPC(KSP ksp, PC *pc)
{
  *pc = &ksp->pc;
  return PETSC_SUCCESS;
}

PetscErrorCode PCSetType(PC pc, PCType type)
{
  PetscCheck(pc && type, PETSC_ERR_ARG_NULL, "null argument");
  PetscCheck(!strcmp(type, PCJACOBI) || !strcmp(type, PCNONE) || !strcmp(type, PCILU) || !strcmp(type, PCLU),
             PETSC_ERR_SUP, "PC type %s is not available (lu, jacobi, ilu or none)", type);
  strcpy(pc->type, type);
  pc->set = PETSC_TRUE;
  return PETSC_SUCCESS;
}

/* KSPSetTolerances: values < 0 keep the current ones (PETSC_CURRENT) */
PetscErrorCode KSPSetTolerances(KSP ksp, PetscReal rtol, PetscReal abstol, PetscReal dtol, PetscInt maxits)
{
  (void)dtol;
  if (rtol >= 0) ksp->rtol = rtol;
  if (abstol >= 0) ksp->abstol = abstol;
  if (maxits > 0) ksp->maxits = maxits;
  return PETSC_SUCCESS;
}

PetscErrorCode KSPGetTotalIterations(KSP ksp, PetscInt *its)
{
  *its = (PetscInt)ksp->its_total;
  return PETSC_SUCCESS;
}

static void KSPFreeWork(KSP ksp)
{
  if (ksp->ilu) (void)ks_ilu0_destroy(ksp->ilu);
  ksp->ilu = NULL;
  if (ksp->lu) {
    if (ks_lu_check(SlepcAMDContext(), ksp->lu)) fprintf(stderr, "slepc_amd: a direct inner solve did not complete\n");
    (void)ks_lu_destroy(ksp->lu);
  }
  ksp->lu = NULL;
  if (ksp->d_dinv) (void)ks_free(ksp->d_dinv);
  if (ksp->d_work) (void)ks_free(ksp->d_work);
  ksp->d_dinv = ksp->d_work = NULL;
  ksp->n = 0;
  ksp->nwork = 0;
}

/* ------------------------------------------------------------ STSINVERT operator */
/* y = (A - sigma I)^-1 x: KSPSolve of STApply_Sinvert (sinvert.c:16-30) */
static PetscErrorCode MatMult_STSinvert(Mat P, Vec x, Vec y)
{
  ST st = (ST)P->data;
  KSP ksp = st->ksp;
  Mat_HIPCSR *h = (Mat_HIPCSR *)st->A->data;
  int its = 0, reason = 0;
  double rnorm = 0.0;
  PetscCheck(x->prec == SLEPC_AMD_F64 && y->prec == SLEPC_AMD_F64, PETSC_ERR_SUP,
             "shift-and-invert runs on FP64 vectors");
  const PetscBool cg = !strcmp(ksp->type, KSPCG) ? PETSC_TRUE : PETSC_FALSE;
  if (!strcmp(ksp->type, KSPPREONLY)) {   /* KSPSolve_PREONLY: one application of the LU factors */
    PetscCheck(ksp->lu, PETSC_ERR_ORDER, "KSP type %s with PC type %s needs STSetUp after they are set", KSPPREONLY,
               PCLU);
    KSCall(ks_lu_solve(SlepcAMDContext(), ksp->lu, (const double *)x->d, (double *)y->d));
    ksp->its_total += 1;
    ksp->nsolves++;
    /* the solves run asynchronously; every 64th confirms their bounded waits were met */
    if (ksp->nsolves % 64 == 0) KSCall(ks_lu_check(SlepcAMDContext(), ksp->lu));
    return PETSC_SUCCESS;
  }
  PetscCheck(strcmp(ksp->pc.type, PCILU) || (!cg && ksp->ilu), PETSC_ERR_ORDER,
             "PC type %s needs KSP type %s set before STSetUp", PCILU, KSPBCGS);
  if (cg)
    KSCall(ks_pcg_solve(SlepcAMDContext(), h->csr, st->sigma, ksp->d_dinv, (const double *)x->d, (double *)y->d,
                        ksp->rtol, ksp->abstol, (int)ksp->maxits, ksp->d_work, &its, &rnorm, &reason));
  else if (ksp->ilu)
    KSCall(ks_bcgs_solve_ilu(SlepcAMDContext(), h->csr, st->sigma, ksp->ilu, (const double *)x->d, (double *)y->d,
                             ksp->rtol, ksp->abstol, (int)ksp->maxits, ksp->d_work, &its, &rnorm, &reason));
  else
    KSCall(ks_bcgs_solve(SlepcAMDContext(), h->csr, st->sigma, ksp->d_dinv, (const double *)x->d, (double *)y->d,
                         ksp->rtol, ksp->abstol, (int)ksp->maxits, ksp->d_work, &its, &rnorm, &reason));
  ksp->its_total += its;
  ksp->nsolves++;
  PetscCheck(reason > 0, PETSC_ERR_NOT_CONVERGED, "KSP %s did not converge (reason %d, %d iterations, residual norm %g)%s",
             ksp->type, reason, its, rnorm,
             cg && reason == -8 ? "; A - sigma I is not positive definite: use KSPBCGS for a shift inside the spectrum"
                                : "");
  return PETSC_SUCCESS;
}

/* STSetUp for sinvert (sinvert.c:88-105): the shift, the preconditioner and the work vectors */
PetscErrorCode STSetUp(ST st)
{
  Mat_HIPCSR *h;
  KSP ksp;
  if (st->setup) return PETSC_SUCCESS;
  PetscCheck(st->A, PETSC_ERR_ORDER, "STSetMatrices must be called first");
  if (!st->sigma_set) st->sigma = strcmp(st->type, STSINVERT) ? 0.0 : st->defsigma;
  if (!strcmp(st->type, STSINVERT)) {
    PetscCheck(!strcmp(st->A->hdr.type_name, MATHIPCSR), PETSC_ERR_SUP, "shift-and-invert needs a %s matrix", MATHIPCSR);
    h = (Mat_HIPCSR *)st->A->data;
    PetscCheck(h->csr && h->vprec == SLEPC_AMD_F64, PETSC_ERR_SUP, "shift-and-invert needs FP64 matrix values");
    PetscCall(STGetKSP(st, &ksp));
    {
      /* KSPSetFromOptions of the ST's solver (prefix st_) */
      char v[SLEPC_NAME_LEN];
      PetscBool set;
      PetscCall(PetscOptionsGetString("-st_ksp_type", v, sizeof(v), &set));
      if (set) PetscCall(KSPSetType(ksp, v));
      PetscCall(PetscOptionsGetString("-st_pc_type", v, sizeof(v), &set));
      if (set) PetscCall(PCSetType(&ksp->pc, v));
    }
    /* BiCGStab's r rhat p v s t (CG uses the first four), then the solver
       state (64 doubles); sized for either, so KSPSetType may follow STSetUp */
    const size_t nwork = 6 * (size_t)st->A->m + 64;
    if (ksp->n != st->A->m || ksp->nwork < nwork) {
      KSPFreeWork(ksp);
      KSCall(ks_malloc((void **)&ksp->d_dinv, sizeof(double) * (size_t)st->A->m));
      KSCall(ks_malloc((void **)&ksp->d_work, sizeof(double) * nwork));
      ksp->n = st->A->m;
      ksp->nwork = nwork;
    }
    KSCall(ks_pcg_jacobi(SlepcAMDContext(), h->csr, st->sigma, !strcmp(ksp->pc.type, PCJACOBI), ksp->d_dinv));
    if (ksp->ilu) (void)ks_ilu0_destroy(ksp->ilu);
    ksp->ilu = NULL;
    if (ksp->lu) (void)ks_lu_destroy(ksp->lu);
    ksp->lu = NULL;
    PetscCheck(!strcmp(ksp->type, KSPPREONLY) == !strcmp(ksp->pc.type, PCLU), PETSC_ERR_SUP,
               "the device inner solvers pair KSP type %s with PC type %s only (given %s with %s)", KSPPREONLY, PCLU,
               ksp->type, ksp->pc.type);
    if (!strcmp(ksp->type, KSPPREONLY)) {
      PetscCheck(st->A->m <= KS_LU_MAXN, PETSC_ERR_SUP,
                 "the direct inner solve (%s + %s) holds A - sigma I dense and takes n <= %d (n = %d): choose %s or %s "
                 "with -st_ksp_type",
                 KSPPREONLY, PCLU, KS_LU_MAXN, (int)st->A->m, KSPCG, KSPBCGS);
      KSCall(ks_lu_create(SlepcAMDContext(), h->csr, st->sigma, &ksp->lu));
    }
    if (!strcmp(ksp->pc.type, PCILU)) {
      PetscCheck(!strcmp(ksp->type, KSPBCGS), PETSC_ERR_SUP, "PC type %s is available with KSP type %s", PCILU,
                 KSPBCGS);
      KSCall(ks_ilu0_create(SlepcAMDContext(), h->csr, st->sigma, &ksp->ilu));
    }
    ksp->its_total = 0;
    ksp->nsolves = 0;
    if (!st->P) {
      PetscCall(PetscMallocZero(sizeof(struct _p_Mat), &st->P));
      snprintf(st->P->hdr.type_name, sizeof(st->P->hdr.type_name), "%s", "stsinvert");
      st->P->ops->mult = MatMult_STSinvert;
    }
    st->P->m = st->P->n = st->A->m;
    st->P->data = st;
  }
  st->setup = PETSC_TRUE;
  return PETSC_SUCCESS;
}

PetscErrorCode STGetOperator(ST st, Mat *Op)
{
  PetscCheck(st && st->A, PETSC_ERR_ORDER, "STSetMatrices must be called first");
  PetscCall(STSetUp(st));
  if (!strcmp(st->type, STSINVERT)) {
    *Op = st->P;
    return PETSC_SUCCESS;
  }
  PetscCheck(st->sigma == 0.0, PETSC_ERR_SUP, "shifted operators are not implemented for %s", STSHIFT);
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
  Mat Op;
  PetscCall(STGetOperator(st, &Op));
  return MatMult(Op, x, y);
}

/* STBackTransform (stsolve.c:244-262): sinvert maps theta to sigma + 1/theta */
PetscErrorCode STBackTransform(ST st, PetscInt n, PetscScalar *eigr, PetscScalar *eigi)
{
  if (strcmp(st->type, STSINVERT)) return PETSC_SUCCESS;
  for (PetscInt j = 0; j < n; j++) {
    PetscScalar re = eigr[j], im = eigi ? eigi[j] : 0.0, nrm = re * re + im * im;
    if (!im) eigr[j] = 1.0 / re + st->sigma;
    else {
      eigr[j] = re / nrm + st->sigma;
      eigi[j] = -im / nrm;
    }
  }
  return PETSC_SUCCESS;
}

PetscBool STIsSinvert(ST st)
{
  return (st && !strcmp(st->type, STSINVERT)) ? PETSC_TRUE : PETSC_FALSE;
}

PetscErrorCode STDestroy(ST *st)
{
  if (!st || !*st) return PETSC_SUCCESS;
  if ((*st)->ksp) {
    KSPFreeWork((*st)->ksp);
    free((*st)->ksp);
  }
  free((*st)->P);
  free(*st);
  *st = NULL;
  return PETSC_SUCCESS;
}
