/*
 * eps.c -- EPS interface (src/eps/interface): object lifecycle and options
 * (epsbasic.c:48-233, epsopts.c), EPSSetUp (epssetup.c:286-418),
 * EPSSetDimensions_Default (epssetup.c:648-671), EPSAllocateSolution
 * (epssetup.c:692-732), EPSSolve (epssolve.c:119-208), eigenpair retrieval
 * (epssolve.c:18-25,406-560), EPSComputeResidualNorm_Private / EPSComputeError
 * (epssolve.c:666-815), EPSGetStartVector (epssolve.c:841-873) and the Krylov
 * convergence test (src/eps/impls/krylov/epskrylov.c:207-295,
 * epsdefault.c:224-302).
 */
#include "slepcimpl.h"

static SlepcFunctionList EPSList;

PetscErrorCode EPSRegister(const char name[], PetscErrorCode (*create)(EPS))
{
  return SlepcFunctionListAdd(&EPSList, name, (void *)create);
}

/* EPSCreate (epsbasic.c:48-158) defaults */
PetscErrorCode EPSCreate(EPS *out)
{
  EPS eps;
  PetscCall(SlepcCheckInitialized());
  PetscCall(PetscMallocZero(sizeof(struct _p_EPS), &eps));
  eps->nev = 1;
  eps->ncv = PETSC_DETERMINE;
  eps->mpd = PETSC_DETERMINE;
  eps->max_it = PETSC_DETERMINE;
  eps->tol = PETSC_DETERMINE;
  eps->which = (EPSWhich)0;
  eps->target = 0.0;
  eps->problem_type = (EPSProblemType)0;
  eps->reason = EPS_CONVERGED_ITERATING;
  eps->state = EPS_STATE_INITIAL;
  PetscCall(STCreate(&eps->st));
  PetscCall(DSCreate(&eps->ds));
  *out = eps;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSSetType(EPS eps, EPSType type)
{
  PetscErrorCode (*create)(EPS);
  PetscCheck(eps && type, PETSC_ERR_ARG_NULL, "null argument");
  if (!strcmp(eps->type, type)) return PETSC_SUCCESS;
  create = (PetscErrorCode(*)(EPS))SlepcFunctionListFind(&EPSList, type);
  PetscCheck(create, PETSC_ERR_ARG_WRONG, "unknown EPS type %s", type);
  if (eps->ops.destroy) PetscCall(eps->ops.destroy(eps));
  memset(&eps->ops, 0, sizeof(eps->ops));
  free(eps->data);
  eps->data = NULL;
  strcpy(eps->type, type);
  eps->state = EPS_STATE_INITIAL;
  return create(eps);
}

PetscErrorCode EPSGetType(EPS eps, EPSType *type)
{
  *type = eps->type;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSSetOperators(EPS eps, Mat A, Mat B)
{
  PetscCheck(eps && A, PETSC_ERR_ARG_NULL, "null argument");
  PetscCheck(!B, PETSC_ERR_SUP, "generalized eigenproblems are not implemented");
  PetscCheck(A->m == A->n, PETSC_ERR_ARG_WRONG, "A must be square");
  eps->A = A;
  PetscCall(STSetMatrices(eps->st, 1, &A));
  eps->state = EPS_STATE_INITIAL;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSSetProblemType(EPS eps, EPSProblemType type)
{
  switch (type) {
  case EPS_HEP: eps->ishermitian = PETSC_TRUE; eps->isgeneralized = PETSC_FALSE; break;
  case EPS_NHEP: eps->ishermitian = PETSC_FALSE; eps->isgeneralized = PETSC_FALSE; break;
  default: SETERRQ(PETSC_ERR_SUP, "only EPS_HEP and EPS_NHEP are supported");
  }
  eps->problem_type = type;
  eps->state = EPS_STATE_INITIAL;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSSetWhichEigenpairs(EPS eps, EPSWhich which)
{
  PetscCheck(which >= EPS_LARGEST_MAGNITUDE && which <= EPS_TARGET_REAL, PETSC_ERR_ARG_OUTOFRANGE, "invalid 'which' value");
  if (eps->which != which) eps->state = EPS_STATE_INITIAL;
  eps->which = which;
  return PETSC_SUCCESS;
}

/* EPSSetTarget (epsbasic.c:378-386): also the ST's default shift */
PetscErrorCode EPSSetTarget(EPS eps, PetscScalar target)
{
  if (eps->target != target) eps->state = EPS_STATE_INITIAL;
  eps->target = target;
  PetscCall(STSetDefaultShift(eps->st, target));
  return PETSC_SUCCESS;
}

PetscErrorCode EPSGetTarget(EPS eps, PetscScalar *target)
{
  *target = eps->target;
  return PETSC_SUCCESS;
}

/* EPSSetExtraction (epsopts.c:984-991); the solver's setup rejects what it cannot do */
PetscErrorCode EPSSetExtraction(EPS eps, EPSExtraction extr)
{
  PetscCheck(extr >= EPS_RITZ && extr <= EPS_REFINED_HARMONIC, PETSC_ERR_ARG_OUTOFRANGE, "invalid extraction type");
  if (eps->extraction != extr) eps->state = EPS_STATE_INITIAL;
  eps->extraction = extr;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSGetExtraction(EPS eps, EPSExtraction *extr)
{
  *extr = eps->extraction;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSSetDimensions(EPS eps, PetscInt nev, PetscInt ncv, PetscInt mpd)
{
  if (nev != PETSC_DETERMINE && nev != PETSC_CURRENT) {
    PetscCheck(nev > 0, PETSC_ERR_ARG_OUTOFRANGE, "Illegal value of nev. Must be > 0");
    eps->nev = nev;
  }
  if (ncv == PETSC_DETERMINE) eps->ncv = PETSC_DETERMINE;
  else if (ncv != PETSC_CURRENT) {
    PetscCheck(ncv > 0, PETSC_ERR_ARG_OUTOFRANGE, "Illegal value of ncv. Must be > 0");
    eps->ncv = ncv;
  }
  if (mpd == PETSC_DETERMINE) eps->mpd = PETSC_DETERMINE;
  else if (mpd != PETSC_CURRENT) {
    PetscCheck(mpd > 0, PETSC_ERR_ARG_OUTOFRANGE, "Illegal value of mpd. Must be > 0");
    eps->mpd = mpd;
  }
  eps->state = EPS_STATE_INITIAL;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSGetDimensions(EPS eps, PetscInt *nev, PetscInt *ncv, PetscInt *mpd)
{
  if (nev) *nev = eps->nev;
  if (ncv) *ncv = eps->ncv;
  if (mpd) *mpd = eps->mpd;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSSetTolerances(EPS eps, PetscReal tol, PetscInt maxits)
{
  if (tol == PETSC_DETERMINE) eps->tol = PETSC_DETERMINE;
  else if (tol != PETSC_CURRENT) {
    PetscCheck(tol > 0.0, PETSC_ERR_ARG_OUTOFRANGE, "Illegal value of tol. Must be > 0");
    eps->tol = tol;
  }
  if (maxits == PETSC_DETERMINE) eps->max_it = PETSC_DETERMINE;
  else if (maxits != PETSC_CURRENT) {
    PetscCheck(maxits > 0, PETSC_ERR_ARG_OUTOFRANGE, "Illegal value of maxits. Must be > 0");
    eps->max_it = maxits;
  }
  eps->state = EPS_STATE_INITIAL;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSGetTolerances(EPS eps, PetscReal *tol, PetscInt *maxits)
{
  if (tol) *tol = eps->tol;
  if (maxits) *maxits = eps->max_it;
  return PETSC_SUCCESS;
}

static void EPSFreeBasis(Vec **basis, PetscInt *n)
{
  if (*basis) {
    for (PetscInt i = 0; i < (*n < 0 ? -*n : *n); i++) VecDestroy(&(*basis)[i]);
    free(*basis);
    *basis = NULL;
  }
  *n = 0;
}

static PetscErrorCode EPSCopyBasis(PetscInt n, Vec *src, Vec **dst)
{
  PetscCall(PetscMallocZero(sizeof(Vec) * (size_t)n, dst));
  for (PetscInt i = 0; i < n; i++) {
    PetscCall(VecDuplicate(src[i], &(*dst)[i]));
    PetscCall(VecCopy(src[i], (*dst)[i]));
  }
  return PETSC_SUCCESS;
}

/* EPSSetInitialSpace (epssetup.c:592): stored until EPSSetUp inserts it */
PetscErrorCode EPSSetInitialSpace(EPS eps, PetscInt n, Vec is[])
{
  PetscCheck(n >= 0, PETSC_ERR_ARG_OUTOFRANGE, "Argument n cannot be negative");
  EPSFreeBasis(&eps->IS, &eps->nini);
  if (n > 0) {
    PetscCall(EPSCopyBasis(n, is, &eps->IS));
    eps->nini = -n;
  }
  if (n > 0) eps->state = EPS_STATE_INITIAL;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSSetDeflationSpace(EPS eps, PetscInt n, Vec v[])
{
  PetscCheck(n >= 0, PETSC_ERR_ARG_OUTOFRANGE, "Argument n cannot be negative");
  EPSFreeBasis(&eps->defl, &eps->nds);
  if (n > 0) {
    PetscCall(EPSCopyBasis(n, v, &eps->defl));
    eps->nds = -n;
  }
  eps->state = EPS_STATE_INITIAL;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSSetBV(EPS eps, BV bv)
{
  if (eps->V) PetscCall(BVDestroy(&eps->V));
  eps->V = bv;
  eps->state = EPS_STATE_INITIAL;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSGetBV(EPS eps, BV *bv)
{
  if (!eps->V) PetscCall(BVCreate(&eps->V));
  *bv = eps->V;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSGetDS(EPS eps, DS *ds)
{
  *ds = eps->ds;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSGetST(EPS eps, ST *st)
{
  *st = eps->st;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSMonitorSet(EPS eps, EPSMonitorFn fn, void *ctx)
{
  eps->monitor = fn;
  eps->monitorctx = ctx;
  return PETSC_SUCCESS;
}

/* EPSSetUp (epssetup.c:286-418) */
PetscErrorCode EPSSetUp(EPS eps)
{
  PetscInt k;
  PetscCheck(eps, PETSC_ERR_ARG_NULL, "null EPS");
  if (eps->state) return PETSC_SUCCESS;
  eps->reason = EPS_CONVERGED_ITERATING;
  if (!eps->type[0]) PetscCall(EPSSetType(eps, EPSKRYLOVSCHUR));
  PetscCheck(eps->A, PETSC_ERR_ARG_WRONGSTATE, "EPSSetOperators must be called first");
  eps->n = eps->A->m;
  if (!eps->problem_type) PetscCall(EPSSetProblemType(eps, EPS_NHEP));
  if (eps->nev > eps->n) eps->nev = eps->n;
  if (eps->ncv > eps->n) eps->ncv = eps->n;
  PetscCheck(!eps->ishermitian || (eps->which != EPS_LARGEST_IMAGINARY && eps->which != EPS_SMALLEST_IMAGINARY),
             PETSC_ERR_SUP, "Sorting the eigenvalues along the imaginary axis is not allowed when all eigenvalues are real");
  PetscCall(eps->ops.setup(eps));
  if (eps->tol == (PetscReal)PETSC_DETERMINE) eps->tol = SLEPC_DEFAULT_TOL;
  /* EPSSetUpSort_Default (epssetup.c:223-240): same criterion for EPS and DS */
  eps->sc.which = eps->which;
  eps->sc.target = eps->target;
  eps->sc.map = PETSC_FALSE;
  PetscCall(DSSetSortCriterion(eps->ds, eps->which, eps->target));
  if (STIsSinvert(eps->st)) {
    /* EPSCheckCompatibleST (epssetup.c:116-120) and the DS criterion on sigma + 1/theta */
    PetscCheck(eps->which == EPS_TARGET_MAGNITUDE || eps->which == EPS_TARGET_REAL, PETSC_ERR_USER_INPUT,
               "Shift-and-invert requires a target 'which' (see EPSSetWhichEigenpairs), for instance -st_type "
               "sinvert -eps_target 0 -eps_target_magnitude");
    PetscCall(STSetUp(eps->st));
    eps->ds->sc.map = PETSC_TRUE;
    eps->ds->sc.sigma = eps->st->sigma;
  } else eps->ds->sc.map = PETSC_FALSE;
  /* deflation and initial vectors */
  if (eps->nds < 0) {
    k = -eps->nds;
    PetscCall(BVInsertConstraints(eps->V, &k, eps->defl));
    PetscInt nin = -eps->nds;
    EPSFreeBasis(&eps->defl, &nin);
    eps->nds = k;
  }
  if (eps->nini < 0) {
    k = -eps->nini;
    PetscCheck(k <= eps->ncv, PETSC_ERR_USER_INPUT, "The number of initial vectors is larger than ncv");
    PetscCall(BVInsertVecs(eps->V, 0, &k, eps->IS, PETSC_TRUE));
    PetscInt nin = -eps->nini;
    EPSFreeBasis(&eps->IS, &nin);
    eps->nini = k;
  }
  eps->state = EPS_STATE_SETUP;
  return PETSC_SUCCESS;
}

/* the stream the basis' work runs on */
static ks_ctx BasisContext(BV V)
{
  ks_ctx c = NULL;
  if (V && V->data && !strcmp(BVTypeName(V), BVHIPKS) && !BVHIPKSGetContext(V, &c)) return c;
  return SlepcAMDContext();
}

/* EPSSolve (epssolve.c:119-208), in three parts so a caller can interleave
   several solves: EPSSolveBegin (set-up, start of the restart loop),
   EPSSolveStep (one restart; EPSSolveReady says whether it would wait for the
   device) and EPSSolveEnd (ordering of the converged pairs).  EPSSolve runs
   them back to back. */
PetscErrorCode EPSSolveBegin(EPS eps)
{
  PetscInt i;
  PetscCheck(eps, PETSC_ERR_ARG_NULL, "null EPS");
  PetscCheck(eps->state < EPS_STATE_SOLVED, PETSC_ERR_ARG_WRONGSTATE, "EPS already solved; change its settings first");
  PetscCall(EPSSetUp(eps));
  eps->t_begin = SlepcWallTime();
  eps->nconv = 0;
  eps->its = 0;
  eps->nmatmult = 0;
  eps->t_ds = 0.0;
  PetscCall(BVHIPKSResetPasses(eps->V));
  eps->stepping = PETSC_FALSE;
  eps->stepV = eps->V;
  for (i = 0; i < eps->ncv; i++) {
    eps->eigr[i] = 0.0;
    eps->eigi[i] = 0.0;
    eps->errest[i] = 0.0;
    eps->perm[i] = i;
  }
  if (eps->ops.solvebegin) {
    PetscCall(eps->ops.solvebegin(eps));
    eps->stepping = PETSC_TRUE;
  } else {
    PetscCall(eps->ops.solve(eps));
  }
  return PETSC_SUCCESS;
}

PetscErrorCode EPSSolveReady(EPS eps, PetscBool *ready)
{
  int idle = 1;
  PetscCheck(eps && ready, PETSC_ERR_ARG_NULL, "null argument");
  if (eps->stepping) {
    BV V = eps->stepV ? eps->stepV : eps->V;
    PetscBool pending;
    PetscCall(BVHIPKSGroupPending(V, &pending));
    if (pending) idle = 0;   /* grouped expansion still being enqueued by its owner */
    else KSCall(ks_ctx_query(BasisContext(V), &idle));
  }
  *ready = idle ? PETSC_TRUE : PETSC_FALSE;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSSolveStep(EPS eps, PetscBool *done)
{
  PetscBool d = PETSC_TRUE;
  PetscCheck(eps && done, PETSC_ERR_ARG_NULL, "null argument");
  if (eps->stepping) {
    PetscCall(eps->ops.solvestep(eps, &d));
    if (d) eps->stepping = PETSC_FALSE;
  }
  *done = d;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSSolveEnd(EPS eps)
{
  PetscInt i;
  PetscCheck(eps, PETSC_ERR_ARG_NULL, "null EPS");
  PetscCheck(!eps->stepping, PETSC_ERR_ARG_WRONGSTATE, "EPSSolveStep() has not finished");
  PetscCheck(eps->reason, PETSC_ERR_PLIB, "Internal error, solver returned without setting converged reason");
  eps->state = EPS_STATE_SOLVED;
  PetscCall(BVSetActiveColumns(eps->V, 0, eps->nconv));
  /* EPSComputeValues (epssolve.c:27-59): back-transform theta to lambda */
  PetscCall(STBackTransform(eps->st, eps->nconv, eps->eigr, eps->eigi));
  /* reorder conjugate eigenvalues (positive imaginary first) */
  for (i = 0; i < eps->nconv - 1; i++) {
    if (eps->eigi[i] != 0) {
      if (eps->eigi[i] < 0) {
        eps->eigi[i] = -eps->eigi[i];
        eps->eigi[i + 1] = -eps->eigi[i + 1];
        PetscCall(EPSComputeVectors(eps));
        PetscCall(BVScaleColumn(eps->V, i + 1, -1.0));
      }
      i++;
    }
  }
  SlepcSortEigenvalues(&eps->sc, eps->nconv, eps->eigr, eps->eigi, eps->perm);
  KSCall(ks_ctx_sync(BasisContext(eps->V)));
  eps->t_solve = SlepcWallTime() - eps->t_begin;
  /* remove deflation and initial subspaces */
  if (eps->nds) eps->nds = 0;
  eps->nini = 0;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSSolve(EPS eps)
{
  PetscBool done = PETSC_FALSE;
  PetscCheck(eps, PETSC_ERR_ARG_NULL, "null EPS");
  if (eps->state >= EPS_STATE_SOLVED) return PETSC_SUCCESS;
  PetscCall(EPSSolveBegin(eps));
  while (!done) PetscCall(EPSSolveStep(eps, &done));
  return EPSSolveEnd(eps);
}

PetscErrorCode EPSComputeVectors(EPS eps)
{
  PetscCheck(eps->state >= EPS_STATE_SOLVED, PETSC_ERR_ARG_WRONGSTATE, "Must call EPSSolve() first");
  if (eps->state == EPS_STATE_SOLVED && eps->ops.computevectors) PetscCall(eps->ops.computevectors(eps));
  eps->state = EPS_STATE_EIGENVECTORS;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSGetConverged(EPS eps, PetscInt *nconv)
{
  *nconv = eps->nconv;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSGetIterationNumber(EPS eps, PetscInt *its)
{
  *its = eps->its;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSGetConvergedReason(EPS eps, EPSConvergedReason *reason)
{
  *reason = eps->reason;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSGetEigenvalue(EPS eps, PetscInt i, PetscScalar *eigr, PetscScalar *eigi)
{
  PetscCheck(eps->state >= EPS_STATE_SOLVED, PETSC_ERR_ARG_WRONGSTATE, "Must call EPSSolve() first");
  PetscCheck(i >= 0 && i < eps->nconv, PETSC_ERR_ARG_OUTOFRANGE, "Argument 2 out of range");
  PetscInt k = eps->perm[i];
  if (eigr) *eigr = eps->eigr[k];
  if (eigi) *eigi = eps->eigi[k];
  return PETSC_SUCCESS;
}

/* EPSGetEigenvector -> BV_GetEigenvector (bvimpl.h:421-444) */
PetscErrorCode EPSGetEigenvector(EPS eps, PetscInt i, Vec Vr, Vec Vi)
{
  PetscInt k;
  PetscScalar ei;
  PetscCheck(eps->state >= EPS_STATE_SOLVED, PETSC_ERR_ARG_WRONGSTATE, "Must call EPSSolve() first");
  PetscCheck(i >= 0 && i < eps->nconv, PETSC_ERR_ARG_OUTOFRANGE, "Argument 2 out of range");
  PetscCall(EPSComputeVectors(eps));
  k = eps->perm[i];
  ei = eps->eigi[k];
  if (ei > 0.0) {
    if (Vr) PetscCall(BVCopyVec(eps->V, k, Vr));
    if (Vi) PetscCall(BVCopyVec(eps->V, k + 1, Vi));
  } else if (ei < 0.0) {
    if (Vr) PetscCall(BVCopyVec(eps->V, k - 1, Vr));
    if (Vi) {
      PetscCall(BVCopyVec(eps->V, k, Vi));
      PetscCall(VecScale(Vi, -1.0));
    }
  } else {
    if (Vr) PetscCall(BVCopyVec(eps->V, k, Vr));
    if (Vi) PetscCall(VecSet(Vi, 0.0));
  }
  return PETSC_SUCCESS;
}

/* EPSGetLeftEigenvector (epssolve.c:567-611): the left eigenvector from W, for
   solves with EPSSetTwoSided */
PetscErrorCode EPSGetLeftEigenvector(EPS eps, PetscInt i, Vec Wr, Vec Wi)
{
  PetscInt k;
  PetscScalar ei;
  PetscCheck(eps->state >= EPS_STATE_SOLVED, PETSC_ERR_ARG_WRONGSTATE, "Must call EPSSolve() first");
  PetscCheck(i >= 0 && i < eps->nconv, PETSC_ERR_ARG_OUTOFRANGE, "Argument 2 out of range");
  PetscCheck(eps->twosided && eps->W, PETSC_ERR_ARG_WRONGSTATE, "Must request left vectors with EPSSetTwoSided");
  PetscCall(EPSComputeVectors(eps));
  k = eps->perm[i];
  ei = eps->eigi[k];
  if (ei > 0.0) {
    if (Wr) PetscCall(BVCopyVec(eps->W, k, Wr));
    if (Wi) PetscCall(BVCopyVec(eps->W, k + 1, Wi));
  } else if (ei < 0.0) {
    if (Wr) PetscCall(BVCopyVec(eps->W, k - 1, Wr));
    if (Wi) {
      PetscCall(BVCopyVec(eps->W, k, Wi));
      PetscCall(VecScale(Wi, -1.0));
    }
  } else {
    if (Wr) PetscCall(BVCopyVec(eps->W, k, Wr));
    if (Wi) PetscCall(VecSet(Wi, 0.0));
  }
  return PETSC_SUCCESS;
}

/* EPSSetTwoSided (epsopts.c): also compute left eigenvectors, by the two-sided
   Krylov-Schur variant (ks-twosided.c) */
PetscErrorCode EPSSetTwoSided(EPS eps, PetscBool twosided)
{
  if (eps->twosided != twosided) eps->state = EPS_STATE_INITIAL;
  eps->twosided = twosided ? PETSC_TRUE : PETSC_FALSE;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSGetTwoSided(EPS eps, PetscBool *twosided)
{
  *twosided = eps->twosided;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSGetEigenpair(EPS eps, PetscInt i, PetscScalar *eigr, PetscScalar *eigi, Vec Vr, Vec Vi)
{
  PetscCall(EPSGetEigenvector(eps, i, Vr, Vi));
  PetscCall(EPSGetEigenvalue(eps, i, eigr, eigi));
  return PETSC_SUCCESS;
}

PetscErrorCode EPSGetErrorEstimate(EPS eps, PetscInt i, PetscReal *errest)
{
  PetscCheck(eps->state >= EPS_STATE_SOLVED, PETSC_ERR_ARG_WRONGSTATE, "Must call EPSSolve() first");
  PetscCheck(i >= 0 && i < eps->nconv, PETSC_ERR_ARG_OUTOFRANGE, "Argument 2 out of range");
  *errest = eps->errest[eps->perm[i]];
  return PETSC_SUCCESS;
}

static PetscErrorCode EPSSetWorkVecs(EPS eps, PetscInt nw)
{
  for (PetscInt i = 0; i < nw; i++) {
    if (eps->work[i] && (eps->work[i]->n != eps->n || eps->work[i]->prec != BVPrecision(eps->V)))
      PetscCall(VecDestroy(&eps->work[i]));
    if (!eps->work[i]) PetscCall(VecCreateHIP(eps->n, BVPrecision(eps->V), &eps->work[i]));
  }
  return PETSC_SUCCESS;
}

/* EPSComputeResidualNorm_Private (epssolve.c:666-718), standard problem */
static PetscErrorCode EPSComputeResidualNorm(EPS eps, PetscScalar kr, PetscScalar ki, Vec xr, Vec xi, Vec *z,
                                             PetscReal *norm)
{
  Vec u = z[0], w = z[2], v = z[1];
  PetscReal nr, ni;
  if (ki == 0 || fabs(ki) < fabs(kr * PETSC_MACHINE_EPSILON)) {
    PetscCall(MatMult(eps->A, xr, u));
    if (fabs(kr) > PETSC_MACHINE_EPSILON) {
      PetscCall(VecCopy(xr, w));
      PetscCall(VecAXPY(u, -kr, w));
    }
    PetscCall(VecNorm(u, NORM_2, norm));
  } else {
    PetscCall(MatMult(eps->A, xr, u));
    if (SlepcAbsEigenvalue(kr, ki) > PETSC_MACHINE_EPSILON) {
      PetscCall(VecCopy(xr, v));
      PetscCall(VecAXPY(u, -kr, v));
      PetscCall(VecCopy(xi, w));
      PetscCall(VecAXPY(u, ki, w));
    }
    PetscCall(VecNorm(u, NORM_2, &nr));
    PetscCall(MatMult(eps->A, xi, u));
    if (SlepcAbsEigenvalue(kr, ki) > PETSC_MACHINE_EPSILON) {
      PetscCall(VecAXPY(u, -kr, w));
      PetscCall(VecAXPY(u, -ki, v));
    }
    PetscCall(VecNorm(u, NORM_2, &ni));
    *norm = SlepcAbsEigenvalue(nr, ni);
  }
  return PETSC_SUCCESS;
}

/* EPSComputeError (epssolve.c:742-815) */
PetscErrorCode EPSComputeError(EPS eps, PetscInt i, EPSErrorType type, PetscReal *error)
{
  PetscScalar kr, ki;
  Vec w[3];
  PetscCheck(eps->state >= EPS_STATE_SOLVED, PETSC_ERR_ARG_WRONGSTATE, "Must call EPSSolve() first");
  PetscCall(EPSSetWorkVecs(eps, 5));
  w[0] = eps->work[1];
  w[1] = eps->work[4];
  w[2] = eps->work[2];
  PetscCall(EPSGetEigenpair(eps, i, &kr, &ki, eps->work[0], eps->work[3]));
  PetscCall(EPSComputeResidualNorm(eps, kr, ki, eps->work[0], eps->work[3], w, error));
  switch (type) {
  case EPS_ERROR_ABSOLUTE: break;
  case EPS_ERROR_RELATIVE: *error /= SlepcAbsEigenvalue(kr, ki); break;
  default: SETERRQ(PETSC_ERR_SUP, "only absolute and relative errors are implemented");
  }
  return PETSC_SUCCESS;
}

/* EPSGetStartVector (epssolve.c:841-873) */
PetscErrorCode EPSGetStartVector(EPS eps, PetscInt i, PetscBool *breakdown)
{
  PetscReal norm;
  PetscBool lindep;
  if (i > 0 || eps->nini == 0) PetscCall(BVSetRandomColumn(eps->V, i));
  PetscCall(BVOrthogonalizeColumn(eps->V, i, NULL, &norm, &lindep));
  if (breakdown) *breakdown = lindep;
  else if (lindep || norm == 0.0) {
    PetscCheck(i, PETSC_ERR_PLIB, "Initial vector is zero or belongs to the deflation space");
    SETERRQ(PETSC_ERR_CONV_FAILED, "Unable to generate more start vectors");
  }
  PetscCall(BVScaleColumn(eps->V, i, 1.0 / norm));
  return PETSC_SUCCESS;
}

/* EPSKrylovConvergence (epskrylov.c:207-295) for STSHIFT sigma=0, trivial
   region and the relative convergence test EPSConvergedRelative (epsdefault.c:224-232) */
PetscErrorCode EPSKrylovConvergence(EPS eps, PetscBool getall, PetscInt kini, PetscInt nits, PetscReal beta,
                                    PetscReal betat, PetscReal corrf, PetscInt *kout)
{
  PetscInt k, newk, newk2, marker = -1;
  PetscReal resnorm, w, lerrest;
  for (k = kini; k < kini + nits; k++) {
    PetscScalar re = eps->eigr[k], im = eps->eigi[k];
    newk = k;
    PetscCall(DSVectors(eps->ds, DS_MAT_X, &newk, &resnorm));
    resnorm *= beta * corrf;
    w = SlepcAbsEigenvalue(re, im);
    eps->errest[k] = (w != 0.0) ? resnorm / w : PETSC_MAX_REAL;
    if (marker == -1 && eps->errest[k] >= eps->tol) marker = k;
    if (eps->twosided) {   /* the left residual must converge too (epskrylov.c:269-276) */
      newk2 = k;
      PetscCall(DSVectors(eps->ds, DS_MAT_Y, &newk2, &resnorm));
      resnorm *= betat;
      lerrest = (w != 0.0) ? resnorm / w : PETSC_MAX_REAL;
      eps->errest[k] = PetscMax(eps->errest[k], lerrest);
      if (marker == -1 && lerrest >= eps->tol) marker = k;
    }
    if (newk == k + 1) {
      eps->errest[k + 1] = eps->errest[k];
      k++;
    }
    if (marker != -1 && !getall) break;
  }
  if (marker != -1) k = marker;
  *kout = k;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSGetSolveStats(EPS eps, int64_t *nmatmult, int64_t *gs_passes, PetscInt *its, PetscReal *t_ds,
                                PetscReal *t_solve)
{
  if (nmatmult) *nmatmult = eps->nmatmult;
  if (gs_passes) PetscCall(BVHIPKSGetPasses(eps->V, gs_passes));
  if (its) *its = eps->its;
  if (t_ds) *t_ds = eps->t_ds;
  if (t_solve) *t_solve = eps->t_solve;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSDestroy(EPS *eps)
{
  EPS e;
  if (!eps || !*eps) return PETSC_SUCCESS;
  e = *eps;
  if (e->ops.destroy) PetscCall(e->ops.destroy(e));
  free(e->data);
  for (int i = 0; i < 5; i++) PetscCall(VecDestroy(&e->work[i]));
  EPSFreeBasis(&e->IS, &e->nini);
  EPSFreeBasis(&e->defl, &e->nds);
  PetscCall(BVDestroy(&e->V));
  PetscCall(BVDestroy(&e->W));
  PetscCall(DSDestroy(&e->ds));
  PetscCall(STDestroy(&e->st));
  free(e->eigr);
  free(e->eigi);
  free(e->errest);
  free(e->perm);
  free(e);
  *eps = NULL;
  return PETSC_SUCCESS;
}
