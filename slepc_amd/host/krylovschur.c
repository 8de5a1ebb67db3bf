/*
 * krylovschur.c -- EPSKRYLOVSCHUR, default variant
 * (src/eps/impls/krylov/krylovschur/krylovschur.c:93-194 setup, :227-337 solve,
 * :1579-1627 create) with Ritz extraction: Lanczos + DSHEP for Hermitian
 * problems, Arnoldi + DSNHEP otherwise.  The restart loop, the truncation and
 * the locking logic are the reference's; the n-length work goes to the BV and
 * Mat types (BVHIPKS and MATHIPCSR on the MI355X).
 */
#include "slepcimpl.h"

/* EPSSetDimensions_Default (epssetup.c:648-671) */
static PetscErrorCode EPSSetDimensions_Default(EPS eps, PetscInt nev, PetscInt *ncv, PetscInt *mpd)
{
  if (*ncv != PETSC_DETERMINE) {
    PetscCheck(*ncv >= nev + 1 || (*ncv == nev && *ncv == eps->n), PETSC_ERR_USER_INPUT, "The value of ncv must be at least nev+1");
  } else if (*mpd != PETSC_DETERMINE) {
    *ncv = PetscMin(eps->n, nev + (*mpd));
  } else {
    if (nev < 500) *ncv = PetscMin(eps->n, PetscMax(2 * nev, nev + 15));
    else {
      *mpd = 500;
      *ncv = PetscMin(eps->n, nev + (*mpd));
    }
  }
  if (*mpd == PETSC_DETERMINE) *mpd = *ncv;
  return PETSC_SUCCESS;
}

/* EPSAllocateSolution (epssetup.c:692-732), extra = 1 */
static PetscErrorCode EPSAllocateSolution(EPS eps, PetscInt extra)
{
  PetscInt requested = eps->ncv + extra;
  BV V;
  free(eps->eigr);
  free(eps->eigi);
  free(eps->errest);
  free(eps->perm);
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)requested, &eps->eigr));
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)requested, &eps->eigi));
  PetscCall(PetscMallocZero(sizeof(PetscReal) * (size_t)requested, &eps->errest));
  PetscCall(PetscMallocZero(sizeof(PetscInt) * (size_t)requested, &eps->perm));
  PetscCall(EPSGetBV(eps, &V));
  if (V->data && (V->m != requested || V->n != eps->n || V->nc)) {
    /* BVResize without copy: rebuild the storage for the new sizes */
    PetscCall(V->ops.destroy(V));
    V->data = NULL;
    V->nc = 0;
  }
  if (!V->type[0]) PetscCall(BVSetType(V, BVHIPKS));
  if (!V->data) {
    V->n = eps->n;
    V->N = eps->n;
    V->m = requested;
    V->l = 0;
    V->k = requested;
    V->ci[0] = V->ci[1] = BV_NOCOL;
    PetscCall(V->ops.create(V));
  }
  return PETSC_SUCCESS;
}

/* EPSSetUp_KrylovSchur (krylovschur.c:93-194), default/symmetric variants */
static PetscErrorCode EPSSetUp_KrylovSchur(EPS eps)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  PetscCall(EPSSetDimensions_Default(eps, eps->nev, &eps->ncv, &eps->mpd));
  PetscCheck(eps->ncv <= eps->nev + eps->mpd, PETSC_ERR_USER_INPUT, "The value of ncv must not be larger than nev+mpd");
  if (eps->max_it == PETSC_DETERMINE) eps->max_it = PetscMax(100, 2 * eps->n / eps->ncv);
  if (!eps->which) eps->which = EPS_LARGEST_MAGNITUDE;   /* EPSSetWhichEigenpairs_Default with STSHIFT */
  PetscCheck(ctx->lock || eps->mpd >= eps->ncv, PETSC_ERR_SUP, "Should not use mpd parameter in non-locking variant");
  if (!ctx->keep) ctx->keep = 0.5;
  PetscCall(EPSAllocateSolution(eps, 1));
  if (eps->ishermitian) {
    PetscCall(DSSetType(eps->ds, DSHEP));
    PetscCall(DSSetCompact(eps->ds, PETSC_TRUE));
  } else {
    PetscCall(DSSetType(eps->ds, DSNHEP));
    PetscCall(DSSetCompact(eps->ds, PETSC_FALSE));
  }
  PetscCall(DSSetExtraRow(eps->ds, PETSC_TRUE));
  PetscCall(DSAllocate(eps->ds, eps->ncv + 1));
  return PETSC_SUCCESS;
}

/* EPSSolve_KrylovSchur_Default (krylovschur.c:227-337) */
static PetscErrorCode EPSSolve_KrylovSchur_Default(EPS eps)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  PetscInt k = 0, l, nv, ld, nconv;
  Mat U, Op, H, T;
  PetscReal beta = 0.0;
  PetscBool breakdown = PETSC_FALSE, hermitian = eps->ishermitian;
  double t0;

  PetscCall(DSGetLeadingDimension(eps->ds, &ld));
  /* get the starting Arnoldi vector */
  PetscCall(EPSGetStartVector(eps, 0, NULL));
  l = 0;

  /* restart loop */
  while (eps->reason == EPS_CONVERGED_ITERATING) {
    eps->its++;

    /* compute an nv-step Arnoldi factorization */
    nv = PetscMin(eps->nconv + eps->mpd, eps->ncv);
    PetscCall(DSSetDimensions(eps->ds, nv, eps->nconv, eps->nconv + l));
    PetscCall(STGetOperator(eps->st, &Op));
    eps->nmatmult += nv - (eps->nconv + l);
    if (hermitian) {
      PetscCall(DSGetMat(eps->ds, DS_MAT_T, &T));
      PetscCall(BVMatLanczos(eps->V, Op, T, eps->nconv + l, &nv, &beta, &breakdown));
      PetscCall(DSRestoreMat(eps->ds, DS_MAT_T, &T));
    } else {
      PetscCall(DSGetMat(eps->ds, DS_MAT_A, &H));
      PetscCall(BVMatArnoldi(eps->V, Op, H, eps->nconv + l, &nv, &beta, &breakdown));
      PetscCall(DSRestoreMat(eps->ds, DS_MAT_A, &H));
    }
    PetscCall(STRestoreOperator(eps->st, &Op));
    PetscCall(DSSetDimensions(eps->ds, nv, eps->nconv, eps->nconv + l));
    PetscCall(DSSetState(eps->ds, l ? DS_STATE_RAW : DS_STATE_INTERMEDIATE));
    PetscCall(BVSetActiveColumns(eps->V, eps->nconv, nv));

    /* solve the projected problem */
    t0 = SlepcWallTime();
    PetscCall(DSSolve(eps->ds, eps->eigr, eps->eigi));
    PetscCall(DSSort(eps->ds, eps->eigr, eps->eigi, NULL, NULL, NULL));
    PetscCall(DSUpdateExtraRow(eps->ds));

    /* check convergence */
    PetscCall(EPSKrylovConvergence(eps, PETSC_FALSE, eps->nconv, nv - eps->nconv, beta, 1.0, &k));
    eps->t_ds += SlepcWallTime() - t0;
    /* EPSStoppingBasic (epsdefault.c:290-302) */
    if (k >= eps->nev) eps->reason = EPS_CONVERGED_TOL;
    else if (eps->its >= eps->max_it) eps->reason = EPS_DIVERGED_ITS;
    nconv = k;

    /* update l */
    if (eps->reason != EPS_CONVERGED_ITERATING || breakdown || k == nv) l = 0;
    else {
      l = PetscMax(1, (PetscInt)((nv - k) * ctx->keep));
      if (!hermitian) PetscCall(DSGetTruncateSize(eps->ds, k, nv, &l));
    }
    if (!ctx->lock && l > 0) { l += k; k = 0; }   /* non-locking variant: reset no. of converged pairs */

    if (eps->reason == EPS_CONVERGED_ITERATING) {
      if (breakdown || k == nv) {
        /* start a new Arnoldi factorization */
        if (k < eps->nev) {
          PetscCall(EPSGetStartVector(eps, k, &breakdown));
          if (breakdown) eps->reason = EPS_DIVERGED_BREAKDOWN;
        }
      } else {
        /* prepare the Rayleigh quotient for restart */
        PetscCall(DSTruncate(eps->ds, k + l, PETSC_FALSE));
      }
    }
    /* update the corresponding vectors V(:,idx) = V*Q(:,idx) */
    PetscCall(DSGetMat(eps->ds, DS_MAT_Q, &U));
    PetscCall(BVMultInPlace(eps->V, U, eps->nconv, k + l));
    PetscCall(DSRestoreMat(eps->ds, DS_MAT_Q, &U));

    if (eps->reason == EPS_CONVERGED_ITERATING && !breakdown) PetscCall(BVCopyColumn(eps->V, nv, k + l));
    eps->nconv = k;
    if (eps->monitor) PetscCall(eps->monitor(eps, eps->its, nconv, eps->eigr, eps->eigi, eps->errest, nv, eps->monitorctx));
  }
  PetscCall(DSTruncate(eps->ds, eps->nconv, PETSC_TRUE));
  return PETSC_SUCCESS;
}

/* EPSComputeVectors_Schur (epsdefault.c:105-169) / _Hermitian (epsdefault.c:28-50) */
static PetscErrorCode EPSComputeVectors_KrylovSchur(EPS eps)
{
  Mat Z;
  if (eps->ishermitian) return PETSC_SUCCESS;   /* Ritz vectors are already in V */
  if (!eps->nconv) return PETSC_SUCCESS;
  PetscCall(DSVectors(eps->ds, DS_MAT_X, NULL, NULL));
  PetscCall(DSGetMat(eps->ds, DS_MAT_X, &Z));
  PetscCall(BVMultInPlace(eps->V, Z, 0, eps->nconv));
  PetscCall(DSRestoreMat(eps->ds, DS_MAT_X, &Z));
  return PETSC_SUCCESS;
}

PetscErrorCode EPSKrylovSchurSetRestart(EPS eps, PetscReal keep)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  PetscCheck(!strcmp(eps->type, EPSKRYLOVSCHUR), PETSC_ERR_ARG_WRONG, "not a Krylov-Schur solver");
  if (keep == PETSC_DEFAULT || keep == PETSC_DETERMINE) ctx->keep = 0.5;
  else {
    PetscCheck(keep >= 0.1 && keep <= 0.9, PETSC_ERR_ARG_OUTOFRANGE, "The keep argument %g must be in the range [0.1,0.9]", keep);
    ctx->keep = keep;
  }
  return PETSC_SUCCESS;
}

PetscErrorCode EPSKrylovSchurSetLocking(EPS eps, PetscBool lock)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  PetscCheck(!strcmp(eps->type, EPSKRYLOVSCHUR), PETSC_ERR_ARG_WRONG, "not a Krylov-Schur solver");
  ctx->lock = lock;
  return PETSC_SUCCESS;
}

/* EPSCreate_KrylovSchur (krylovschur.c:1579-1627) */
PetscErrorCode EPSCreate_KrylovSchur(EPS eps)
{
  EPS_KRYLOVSCHUR *ctx;
  PetscCall(PetscMallocZero(sizeof(EPS_KRYLOVSCHUR), &ctx));
  eps->data = ctx;
  ctx->lock = PETSC_TRUE;
  ctx->keep = 0.0;
  eps->ops.setup = EPSSetUp_KrylovSchur;
  eps->ops.solve = EPSSolve_KrylovSchur_Default;
  eps->ops.computevectors = EPSComputeVectors_KrylovSchur;
  eps->ops.destroy = NULL;
  return PETSC_SUCCESS;
}
