/*
 * krylovschur.c -- EPSKRYLOVSCHUR, default variant
 * (src/eps/impls/krylov/krylovschur/krylovschur.c:93-194 setup, :227-337 solve,
 * :1579-1627 create) with Ritz extraction (Lanczos + DSHEP for Hermitian
 * problems, Arnoldi + DSNHEP otherwise) or harmonic extraction (Arnoldi +
 * DSNHEP with the translated projected matrix, krylovschur.c:270-319).  The restart loop, the truncation and
 * the locking logic are the reference's; the n-length work goes to the BV and
 * Mat types (BVHIPKS and MATHIPCSR on the MI355X).
 */
#include "slepcimpl.h"

extern void dgetrf_(const int *, const int *, double *, const int *, int *, int *);
extern void dgetrs_(const char *, const int *, const int *, const double *, const int *, const int *, double *,
                    const int *, int *);
extern void dgemm_(const char *, const char *, const int *, const int *, const int *, const double *, const double *,
                   const int *, const double *, const int *, const double *, double *, const int *);

/* the left basis draws its random columns from a stream of its own */
#define KS_LEFT_SEED_OFFSET 1000000007ULL

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
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  PetscInt requested = eps->ncv + extra;
  BV V;
  uint64_t seed;
  free(eps->eigr);
  free(eps->eigi);
  free(eps->errest);
  free(eps->perm);
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)requested, &eps->eigr));
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)requested, &eps->eigi));
  PetscCall(PetscMallocZero(sizeof(PetscReal) * (size_t)requested, &eps->errest));
  PetscCall(PetscMallocZero(sizeof(PetscInt) * (size_t)requested, &eps->perm));
  PetscCall(EPSGetBV(eps, &V));
  if (!BVTypeName(V)[0]) PetscCall(BVSetType(V, BVHIPKS));
  /* BVResize without copy: the storage is rebuilt for the new sizes */
  if (!V->data || V->m != requested || V->n != eps->n || V->nc) PetscCall(BVResizeDiscard_Private(V, eps->n, requested));
  /* the device-path settings of this solver on its basis */
  if (!strcmp(BVTypeName(V), BVHIPKS)) {
    PetscCall(BVHIPKSSetStepGraphs(V, ctx->step_graphs));
    PetscCall(BVHIPKSSetGrouped(V, ctx->grouped));
  }
  /* W: a basis like V for the two-sided variant (epssetup.c:726-731) */
  if (eps->W && (!eps->twosided || eps->W->m != requested || eps->W->n != eps->n || strcmp(eps->W->vtype, V->vtype)))
    PetscCall(BVDestroy(&eps->W));
  if (eps->twosided && !eps->W) {
    PetscCall(BVCreate(&eps->W));
    PetscCall(BVSetVecType(eps->W, V->vtype));
    PetscCall(BVSetOrthogonalization(eps->W, V->orthog_type, V->orthog_ref, V->orthog_eta, V->orthog_block));
    PetscCall(BVGetRandomSeed(V, &seed));
    PetscCall(BVSetRandomSeed(eps->W, seed + KS_LEFT_SEED_OFFSET));
    PetscCall(BVSetSizes(eps->W, eps->n, eps->n, requested));
    PetscCall(BVSetType(eps->W, BVTypeName(V)));
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
  /* EPSSetWhichEigenpairs_Default (epsdefault.c:210-219) */
  if (!eps->which) eps->which = STIsSinvert(eps->st) ? EPS_TARGET_MAGNITUDE : EPS_LARGEST_MAGNITUDE;
  if (STIsSinvert(eps->st)) {
    PetscCheck(eps->extraction == EPS_RITZ && !eps->twosided && ctx->mixed_tol <= 0.0, PETSC_ERR_SUP,
               "shift-and-invert runs with Ritz extraction, one-sided, in FP64");
    PetscCheck(!eps->V || BVPrecision(eps->V) == SLEPC_AMD_F64, PETSC_ERR_SUP, "shift-and-invert needs an FP64 basis");
  }
  PetscCheck(ctx->lock || eps->mpd >= eps->ncv, PETSC_ERR_SUP, "Should not use mpd parameter in non-locking variant");
  PetscCheck(eps->extraction == EPS_RITZ || eps->extraction == EPS_HARMONIC, PETSC_ERR_SUP, "Unsupported extraction type");
  PetscCheck(!eps->twosided || !eps->ishermitian, PETSC_ERR_SUP, "Two-sided methods are not intended for symmetric problems");
  PetscCheck(!eps->twosided || ctx->mixed_tol <= 0.0, PETSC_ERR_SUP,
             "The two-sided variant is not available in the mixed-precision form");
  /* the two-sided variant takes precedence over the extraction (krylovschur.c:143-151) */
  ctx->harmonic = eps->extraction == EPS_HARMONIC && !eps->twosided;
  PetscCheck(!ctx->harmonic || ctx->mixed_tol <= 0.0, PETSC_ERR_SUP,
             "Harmonic extraction is not available in the mixed-precision form");
  if (!ctx->keep) ctx->keep = 0.5;
  PetscCall(EPSAllocateSolution(eps, 1));
  free(ctx->g);
  ctx->g = NULL;
  if (ctx->harmonic) PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)(eps->ncv + 1), &ctx->g));
  PetscCall(MatDestroy(&ctx->AT));
  free(ctx->M);
  ctx->M = NULL;
  if (eps->twosided) {
    Mat Op;
    PetscCall(STGetOperator(eps->st, &Op));
    PetscCall(MatHIPCSRCreateTranspose(Op, &ctx->AT));
    PetscCall(STRestoreOperator(eps->st, &Op));
    PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)eps->ncv * (size_t)eps->ncv, &ctx->M));
    PetscCall(DSSetType(eps->ds, DSNHEPTS));
    PetscCall(DSSetCompact(eps->ds, PETSC_FALSE));
  } else if (eps->ishermitian && !ctx->harmonic) {
    /* harmonic extraction runs the default (Arnoldi + DSNHEP) variant for Hermitian problems too */
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

/* EPSSolve_KrylovSchur_Default (krylovschur.c:227-337), written as a
   resumable state machine so that several solves can be interleaved:
     KSExpand      -- top of the restart loop: dimensions, then enqueue the
                      nv-step Arnoldi/Lanczos expansion on the device;
     KSAfterExpand -- wait for it, solve the projected problem, test
                      convergence, truncate, enqueue the restart GEMM and the
                      residual-vector copy, then expand again or finish.
   EPSSolve runs them back to back; a serving loop polls the device between. */
static PetscErrorCode KSExpand(EPS eps)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  Mat Op;
  PetscInt nv;
  eps->its++;
  nv = PetscMin(eps->nconv + eps->mpd, eps->ncv);
  PetscCall(DSSetDimensions(eps->ds, nv, eps->nconv, eps->nconv + ctx->l));
  PetscCall(STGetOperator(eps->st, &Op));
  eps->nmatmult += nv - (eps->nconv + ctx->l);
  PetscCall(BVGetKrylovOps(eps->V, &ctx->kry_ext));
  if (ctx->kry_ext) PetscCall(ctx->kry_ext->matkrylovbegin(eps->V, Op, eps->nconv + ctx->l, nv));
  else {   /* the reference's per-step loop, finished here; KSAfterExpand collects its results */
    ctx->kry_m = nv;
    PetscCall(BVMatKrylov(eps->V, Op, eps->nconv + ctx->l, &ctx->kry_m, &ctx->kry_beta, &ctx->kry_breakdown));
  }
  PetscCall(STRestoreOperator(eps->st, &Op));
  ctx->nv = nv;
  ctx->phase = KS_PHASE_EXPANDING;
  return PETSC_SUCCESS;
}

/* The hand-over of the mixed-precision form.  At the first phase's last
   restart the kept columns satisfy the Krylov-Schur relation
       A V_m = V_m B + v_{m+1} b^T      (B: m x m Schur form, b: the extra row)
   and span(V_m) is the Krylov space of one vector: reduced to Arnoldi form
   (Stewart, "A Krylov-Schur algorithm for large eigenproblems", SIAM J. Matrix
   Anal. Appl. 23 (2001), Sec. 2), A V_m Q = V_m Q H + beta v_{m+1} e_m^T with H
   upper Hessenberg, and that vector is u = V_m Q e_1.  One reflector takes b to
   beta e_m, then for rows r = m-1..2 a reflector on the leading r indices
   zeroes B(r, 0:r-2) without touching index m-1.  q1 = Q e_1 is stored; the
   FP64 phase restarts from u, so its Krylov space holds the whole retained
   subspace, not only the converged vectors. */
static void HouseApply(PetscInt m, PetscInt r, const PetscScalar *v, PetscScalar *B, PetscScalar *Q)
{
  /* P = I - 2 v v^T / (v^T v) on indices 0..r-1: B = P B P, Q = Q P */
  PetscScalar vv = 0.0;
  for (PetscInt i = 0; i < r; i++) vv += v[i] * v[i];
  if (vv == 0.0) return;
  for (PetscInt j = 0; j < m; j++) {   /* B(0:r, j) = P B(0:r, j) */
    PetscScalar t = 0.0;
    for (PetscInt i = 0; i < r; i++) t += v[i] * B[i + j * m];
    t *= 2.0 / vv;
    for (PetscInt i = 0; i < r; i++) B[i + j * m] -= t * v[i];
  }
  for (PetscInt i = 0; i < m; i++) {   /* B(i, 0:r) = B(i, 0:r) P, Q(i, 0:r) = Q(i, 0:r) P */
    PetscScalar t = 0.0, u = 0.0;
    for (PetscInt j = 0; j < r; j++) t += B[i + j * m] * v[j], u += Q[i + j * m] * v[j];
    t *= 2.0 / vv;
    u *= 2.0 / vv;
    for (PetscInt j = 0; j < r; j++) B[i + j * m] -= t * v[j], Q[i + j * m] -= u * v[j];
  }
}

/* v = x + sign(x_p) |x| e_p (the reflector taking x to a multiple of e_p) */
static void HouseVector(PetscInt r, const PetscScalar *x, PetscInt p, PetscScalar *v)
{
  PetscScalar nx = 0.0;
  for (PetscInt i = 0; i < r; i++) v[i] = x[i], nx += x[i] * x[i];
  nx = sqrt(nx);
  if (nx == 0.0) return;
  v[p] += x[p] >= 0.0 ? nx : -nx;
}

static PetscErrorCode KSHandoverVector(EPS eps, PetscInt m, PetscInt nv)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  DS ds = eps->ds;
  const PetscInt ld = ds->ld;
  PetscScalar *B = NULL, *Q = NULL, *v = NULL, *x = NULL;
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)m * (size_t)m * 2 + sizeof(PetscScalar) * 2 * (size_t)(m + 1), &B));
  Q = B + (size_t)m * m;
  v = Q + (size_t)m * m;
  x = v + m + 1;
  if (eps->ishermitian) {   /* DSHEP: B = diag(theta), b = the updated arrow (column 1 of T) */
    const PetscReal *e = ds->mat[DS_MAT_T] + ld;
    for (PetscInt i = 0; i < m; i++) B[i + i * m] = eps->eigr[i], x[i] = e[i];
  } else {                  /* DSNHEP: B = the leading Schur block, b = row nv of A */
    const PetscScalar *A = ds->mat[DS_MAT_A];
    for (PetscInt j = 0; j < m; j++) {
      for (PetscInt i = 0; i < m; i++) B[i + j * m] = A[i + (size_t)j * ld];
      x[j] = A[nv + (size_t)j * ld];
    }
  }
  for (PetscInt i = 0; i < m; i++) Q[i + i * m] = 1.0;
  HouseVector(m, x, m - 1, v);
  HouseApply(m, m, v, B, Q);
  for (PetscInt r = m - 1; r >= 2; r--) {
    for (PetscInt j = 0; j < r; j++) x[j] = B[r + j * m];
    HouseVector(r, x, r - 1, v);
    HouseApply(m, r, v, B, Q);
  }
  free(ctx->ho_q1);
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)m, &ctx->ho_q1));
  for (PetscInt i = 0; i < m; i++) ctx->ho_q1[i] = Q[i];
  ctx->ho_m = m;
  free(B);
  return PETSC_SUCCESS;
}

static PetscErrorCode KSAfterExpand(EPS eps)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  PetscInt k = 0, l = ctx->l, nv = ctx->nv, nconv;
  Mat U, H, T;
  PetscReal beta = 0.0;
  PetscReal gamma = 1.0;
  PetscBool breakdown = PETSC_FALSE, hermitian = eps->ishermitian && !ctx->harmonic, handover;
  double t0;

  if (ctx->kry_ext) PetscCall(ctx->kry_ext->matkrylovend(eps->V, eps->nconv + l, &nv, &beta, &breakdown));
  else {
    nv = ctx->kry_m;
    beta = ctx->kry_beta;
    breakdown = ctx->kry_breakdown;
  }
  if (hermitian) {
    PetscCall(DSGetMat(eps->ds, DS_MAT_T, &T));
    PetscCall(BVCopyTridiagonal(eps->V, T, eps->nconv + l, nv));
    PetscCall(DSRestoreMat(eps->ds, DS_MAT_T, &T));
  } else {
    PetscCall(DSGetMat(eps->ds, DS_MAT_A, &H));
    PetscCall(BVCopyHessenberg(eps->V, H, eps->nconv + l, nv));
    PetscCall(DSRestoreMat(eps->ds, DS_MAT_A, &H));
  }
  PetscCall(DSSetDimensions(eps->ds, nv, eps->nconv, eps->nconv + l));
  PetscCall(DSSetState(eps->ds, l ? DS_STATE_RAW : DS_STATE_INTERMEDIATE));
  PetscCall(BVSetActiveColumns(eps->V, eps->nconv, nv));

  /* translation of the Krylov decomposition for harmonic extraction */
  t0 = SlepcWallTime();
  if (ctx->harmonic) PetscCall(DSTranslateHarmonic(eps->ds, eps->target, beta, PETSC_FALSE, ctx->g, &gamma));

  /* solve the projected problem */
  PetscCall(DSSolve(eps->ds, eps->eigr, eps->eigi));
  PetscCall(DSSort(eps->ds, eps->eigr, eps->eigi, NULL, NULL, NULL));
  PetscCall(DSUpdateExtraRow(eps->ds));

  /* check convergence */
  PetscCall(EPSKrylovConvergence(eps, PETSC_FALSE, eps->nconv, nv - eps->nconv, beta, 0.0, gamma, &k));
  eps->t_ds += SlepcWallTime() - t0;
  /* EPSStoppingBasic (epsdefault.c:290-302) */
  if (k >= eps->nev) eps->reason = EPS_CONVERGED_TOL;
  else if (eps->its >= eps->max_it) eps->reason = EPS_DIVERGED_ITS;
  nconv = k;
  /* the first phase of the mixed form restarts once more at convergence and
     hands its kept subspace over (KSHandoverVector) */
  handover = ctx->handover && eps->reason == EPS_CONVERGED_TOL && !breakdown && k < nv;

  /* update l */
  if ((eps->reason != EPS_CONVERGED_ITERATING && !handover) || breakdown || k == nv) l = 0;
  else {
    l = PetscMax(1, (PetscInt)((nv - k) * ctx->keep));
    if (!hermitian) PetscCall(DSGetTruncateSize(eps->ds, k, nv, &l));
  }
  if (!ctx->lock && l > 0) { l += k; k = 0; }   /* non-locking variant: reset no. of converged pairs */
  if (handover) {
    PetscCall(KSHandoverVector(eps, k + l, nv));
    ctx->ho_nconv = nconv;
  }

  if (eps->reason == EPS_CONVERGED_ITERATING || handover) {
    if (breakdown || k == nv) {
      /* start a new Arnoldi factorization */
      if (k < eps->nev) {
        PetscCall(EPSGetStartVector(eps, k, &breakdown));
        if (breakdown) eps->reason = EPS_DIVERGED_BREAKDOWN;
      }
    } else {
      if (ctx->harmonic) {
        /* undo the translation: gamma u^ = u - V g~ (krylovschur.c:310-319) */
        PetscCall(DSSetDimensions(eps->ds, nv, k, l));
        PetscCall(DSTranslateHarmonic(eps->ds, 0.0, beta, PETSC_TRUE, ctx->g, &gamma));
        PetscCall(BVSetActiveColumns(eps->V, 0, nv));
        PetscCall(BVMultColumn(eps->V, -1.0, 1.0, nv, ctx->g));
        PetscCall(BVScaleColumn(eps->V, nv, 1.0 / gamma));
        PetscCall(BVSetActiveColumns(eps->V, eps->nconv, nv));
        PetscCall(DSSetDimensions(eps->ds, nv, k, nv));
      }
      /* prepare the Rayleigh quotient for restart */
      PetscCall(DSTruncate(eps->ds, k + l, PETSC_FALSE));
    }
  }
  /* update the corresponding vectors V(:,idx) = V*Q(:,idx) */
  PetscCall(DSGetMat(eps->ds, DS_MAT_Q, &U));
  PetscCall(BVMultInPlace(eps->V, U, eps->nconv, k + l));
  PetscCall(DSRestoreMat(eps->ds, DS_MAT_Q, &U));

  if ((eps->reason == EPS_CONVERGED_ITERATING || handover) && !breakdown) PetscCall(BVCopyColumn(eps->V, nv, k + l));
  eps->nconv = k;
  ctx->l = l;
  if (eps->monitor) PetscCall(eps->monitor(eps, eps->its, nconv, eps->eigr, eps->eigi, eps->errest, nv, eps->monitorctx));

  if (eps->reason == EPS_CONVERGED_ITERATING) PetscCall(KSExpand(eps));
  else {
    PetscCall(DSTruncate(eps->ds, eps->nconv, PETSC_TRUE));
    ctx->phase = KS_PHASE_DONE;
  }
  return PETSC_SUCCESS;
}

/* the start of the restart loop: starting vector, first expansion */
static PetscErrorCode KSStart(EPS eps)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  PetscCall(EPSGetStartVector(eps, 0, NULL));
  ctx->l = 0;
  return KSExpand(eps);
}

/* Mixed-precision first phase: the same Krylov-Schur on an FP32 basis (FP32
   SpMV on an FP32 copy of the operator values, FP32 Gram-Schmidt products with
   FP64 DGKS scalars and an FP64 projected problem) down to ctx->mixed_tol.  At
   convergence it restarts once more and its kept Krylov-Schur decomposition,
   reduced to Arnoldi form, gives the one vector u whose Krylov space is the
   kept subspace; u becomes the FP64 start vector of the main solve (column 0
   of V, used as the initial space), so the FP64 restarts continue from the
   whole retained subspace at the FP32 accuracy.
   The phase is itself stepped (MixedBegin, then EPSSolveStep of the FP32 solve
   until it is done, then MixedEnd), so a serving loop never blocks on it. */
static PetscErrorCode MixedBegin(EPS eps)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data, *c32;
  EPS e32;
  BV bv32;
  Vec v64 = NULL, *defl = NULL;

  ctx->t32_start = SlepcWallTime();
  PetscCall(EPSCreate(&e32));
  PetscCall(EPSSetOperators(e32, eps->A, NULL));
  PetscCall(EPSSetProblemType(e32, eps->problem_type));
  PetscCall(EPSSetWhichEigenpairs(e32, eps->which));
  PetscCall(EPSSetTarget(e32, eps->target));
  PetscCall(EPSSetDimensions(e32, eps->nev, eps->ncv, eps->mpd));
  PetscCall(EPSSetTolerances(e32, ctx->mixed_tol, eps->max_it));
  PetscCall(EPSSetType(e32, EPSKRYLOVSCHUR));
  c32 = (EPS_KRYLOVSCHUR *)e32->data;
  c32->keep = ctx->keep;
  c32->lock = ctx->lock;
  {
    /* the start vector the FP64 phase gets: "sum" of the converged Ritz vectors,
       or "arnoldi", the start of the Arnoldi form of the last restart's kept
       subspace (non-locking, so that decomposition couples every kept column) */
    char ho[16] = "sum";
    PetscBool set;
    PetscCall(PetscOptionsGetString("-eps_krylovschur_mixed_handover", ho, sizeof(ho), &set));
    PetscCheck(!strcmp(ho, "sum") || !strcmp(ho, "arnoldi"), PETSC_ERR_ARG_WRONG, "unknown hand-over %s", ho);
    if (!strcmp(ho, "arnoldi")) {
      c32->lock = PETSC_FALSE;
      c32->handover = PETSC_TRUE;
    }
  }
  PetscCall(EPSGetBV(e32, &bv32));
  PetscCall(BVSetType(bv32, BVHIPKS));
  PetscCall(BVHIPKSSetPrecision(bv32, SLEPC_AMD_F32));
  bv32->orthog_ref = eps->V->orthog_ref;
  bv32->orthog_eta = eps->V->orthog_eta;
  {
    uint64_t seed;
    PetscCall(BVGetRandomSeed(eps->V, &seed));
    PetscCall(BVSetRandomSeed(bv32, seed));
  }
  c32->step_graphs = ctx->step_graphs;
  /* same start: the FP64 initial vector inserted by EPSSetUp, rounded to FP32 */
  if (eps->nini > 0) {
    PetscCall(VecCreateHIP(eps->n, SLEPC_AMD_F32, &v64));
    PetscCall(BVCopyVec(eps->V, 0, v64));
    PetscCall(EPSSetInitialSpace(e32, 1, &v64));
    PetscCall(VecDestroy(&v64));
  }
  /* same deflation space */
  if (eps->nds > 0) {
    PetscCall(PetscMallocZero(sizeof(Vec) * (size_t)eps->nds, &defl));
    for (PetscInt i = 0; i < eps->nds; i++) {
      PetscCall(VecCreateHIP(eps->n, SLEPC_AMD_F32, &defl[i]));
      PetscCall(BVCopyVec(eps->V, i - eps->nds, defl[i]));
    }
    PetscCall(EPSSetDeflationSpace(e32, eps->nds, defl));
    for (PetscInt i = 0; i < eps->nds; i++) PetscCall(VecDestroy(&defl[i]));
    free(defl);
  }
  PetscCall(EPSSolveBegin(e32));
  ctx->e32 = e32;
  ctx->phase = KS_PHASE_MIXED32;
  eps->stepV = e32->V;
  return PETSC_SUCCESS;
}

static PetscErrorCode MixedEnd(EPS eps)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data, *c32;
  EPS e32 = ctx->e32;
  Vec u = NULL, xr = NULL, xi = NULL;
  PetscInt nuse;

  PetscCall(EPSSolveEnd(e32));
  c32 = (EPS_KRYLOVSCHUR *)e32->data;
  ctx->its32 = e32->its;
  ctx->nconv32 = c32->ho_q1 ? c32->ho_nconv : e32->nconv;
  ctx->nmatmult32 = e32->nmatmult;
  PetscCall(BVHIPKSGetPasses(e32->V, &ctx->passes32));
  nuse = PetscMin(e32->nconv, eps->nev);
  if (c32->ho_q1) {   /* u = V32(:, 0:m) q1, the start of the Arnoldi form, in FP64 */
    PetscCall(BVSetActiveColumns(e32->V, 0, c32->ho_m));
    PetscCall(VecCreateHIP(eps->n, SLEPC_AMD_F32, &u));
    PetscCall(BVMultVec(e32->V, 1.0, 0.0, u, c32->ho_q1));
    PetscCall(BVInsertVec(eps->V, 0, u));   /* FP32 -> FP64 conversion */
    eps->nini = 1;
  } else if (nuse > 0) {   /* the sum of the converged Ritz vectors, a conjugate pair by both parts */
    PetscCall(VecCreateHIP(eps->n, SLEPC_AMD_F32, &u));
    PetscCall(VecCreateHIP(eps->n, SLEPC_AMD_F32, &xr));
    PetscCall(VecCreateHIP(eps->n, SLEPC_AMD_F32, &xi));
    for (PetscInt i = 0; i < nuse; i++) {
      PetscScalar kr, ki;
      PetscCall(EPSGetEigenpair(e32, i, &kr, &ki, xr, xi));
      PetscCall(VecAXPY(u, 1.0, xr));
      if (ki > 0.0) PetscCall(VecAXPY(u, 1.0, xi));   /* the pair's real and imaginary parts once */
      if (ki != 0.0) i++;
    }
    PetscCall(BVInsertVec(eps->V, 0, u));   /* FP32 -> FP64 conversion */
    eps->nini = 1;
  }
  PetscCall(VecDestroy(&u));
  PetscCall(VecDestroy(&xr));
  PetscCall(VecDestroy(&xi));
  PetscCall(EPSDestroy(&ctx->e32));
  ctx->t32 = SlepcWallTime() - ctx->t32_start;
  eps->stepV = eps->V;
  return PETSC_SUCCESS;
}

/* ------------------------------------------------------------ two-sided variant */
/* EPSGetLeftStartVector (epssolve.c:879-900): a random column of W, orthonormalized */
PetscErrorCode EPSGetLeftStartVector(EPS eps, PetscInt i, PetscBool *breakdown)
{
  PetscReal norm;
  PetscBool lindep;
  PetscCall(BVSetRandomColumn(eps->W, i));
  PetscCall(BVOrthogonalizeColumn(eps->W, i, NULL, &norm, &lindep));
  if (breakdown) *breakdown = lindep;
  else if (lindep || norm == 0.0) {
    PetscCheck(i, PETSC_ERR_PLIB, "Left initial vector is zero");
    SETERRQ(PETSC_ERR_CONV_FAILED, "Unable to generate more left start vectors");
  }
  PetscCall(BVScaleColumn(eps->W, i, 1.0 / norm));
  return PETSC_SUCCESS;
}

/* EPSTwoSidedRQUpdate1 (ks-twosided.c:25-77): oblique correction of the two residual
   vectors, u -= V M^-1 W^T u and w -= W M^-T V^T w, folded into S and T */
static PetscErrorCode TwoSidedRQUpdate1(EPS eps, PetscInt nv, PetscReal beta, PetscReal betat)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  PetscInt ld, l, nnv, ncv = eps->ncv;
  int n_ = (int)nv, ncv_ = (int)ncv, one = 1, info = 0, *p = NULL;
  PetscScalar *A = NULL, *w = NULL, *S;
  Vec u = NULL;
  PetscErrorCode e = PETSC_SUCCESS;
  PetscCall(DSGetLeadingDimension(eps->ds, &ld));
  PetscCall(PetscMallocZero(sizeof(int) * (size_t)(nv > 0 ? nv : 1), &p));
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)ncv * (size_t)ncv, &A));
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)ncv, &w));
  PetscCall(BVGetActiveColumns(eps->V, &l, &nnv));
  PetscCall(BVSetActiveColumns(eps->V, 0, nv));
  PetscCall(BVSetActiveColumns(eps->W, 0, nv));
  e = BVGetColumn(eps->V, nv, &u);
  if (!e) e = BVDotVec(eps->W, u, w);
  if (u) BVRestoreColumn(eps->V, nv, &u);
  if (!e) {
    memcpy(A, ctx->M, sizeof(PetscScalar) * (size_t)ncv * (size_t)ncv);
    dgetrf_(&n_, &n_, A, &ncv_, p, &info);
    if (!info) dgetrs_("N", &n_, &one, A, &ncv_, p, w, &ncv_, &info);
    if (info) e = PETSC_ERR_LIB;
  }
  if (!e) e = BVMultColumn(eps->V, -1.0, 1.0, nv, w);
  if (!e) {
    S = eps->ds->mat[DS_MAT_A];
    for (PetscInt i = 0; i < nv; i++) S[(size_t)(nv - 1) * ld + i] += beta * w[i];
    u = NULL;
    e = BVGetColumn(eps->W, nv, &u);
    if (!e) e = BVDotVec(eps->V, u, w);
    if (u) BVRestoreColumn(eps->W, nv, &u);
  }
  if (!e) {
    dgetrs_("C", &n_, &one, A, &ncv_, p, w, &ncv_, &info);
    if (info) e = PETSC_ERR_LIB;
  }
  if (!e) e = BVMultColumn(eps->W, -1.0, 1.0, nv, w);
  if (!e) {
    S = eps->ds->mat[DS_MAT_B];
    for (PetscInt i = 0; i < nv; i++) S[(size_t)(nv - 1) * ld + i] += betat * w[i];
  }
  free(p);
  free(A);
  free(w);
  PetscCall(BVSetActiveColumns(eps->V, l, nnv));
  PetscCall(BVSetActiveColumns(eps->W, l, nnv));
  PetscCheck(!e || e != PETSC_ERR_LIB, PETSC_ERR_LIB, "LAPACK getrf/getrs failed on the projection matrix (info %d)", info);
  return e;
}

/* EPSTwoSidedRQUpdate2 (ks-twosided.c:79-136): re-orthogonalize the kept residual
   columns k of V and W, fold the coefficients into S and T, and M = Z^T M Q */
static PetscErrorCode TwoSidedRQUpdate2(EPS eps, PetscInt k)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  PetscInt ld, l, nv, ncv = eps->ncv;
  PetscReal norm;
  PetscScalar *c = NULL, *w = NULL, *A, zero = 0.0, sone = 1.0;
  int n_, ncv_ = (int)ncv, ld_;
  PetscCall(DSGetLeadingDimension(eps->ds, &ld));
  PetscCall(BVGetActiveColumns(eps->V, &l, &nv));
  PetscCall(BVSetActiveColumns(eps->V, 0, nv));
  PetscCall(BVSetActiveColumns(eps->W, 0, nv));
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)ncv * (size_t)ncv, &w));
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)(ncv + 1), &c));
  for (int side = 0; side < 2; side++) {
    BV X = side ? eps->W : eps->V;
    PetscCall(BVOrthogonalizeColumn(X, k, c, &norm, NULL));
    PetscCall(BVScaleColumn(X, k, 1.0 / norm));
    A = eps->ds->mat[side ? DS_MAT_B : DS_MAT_A];
    for (PetscInt j = l; j < k; j++) {
      for (PetscInt i = 0; i < k; i++) A[i + (size_t)j * ld] += c[i] * A[k + (size_t)j * ld];
      A[k + (size_t)j * ld] *= norm;
    }
  }
  n_ = (int)nv;
  ld_ = (int)ld;
  dgemm_("N", "N", &n_, &n_, &n_, &sone, ctx->M, &ncv_, eps->ds->mat[DS_MAT_Q], &ld_, &zero, w, &ncv_);
  dgemm_("C", "N", &n_, &n_, &n_, &sone, eps->ds->mat[DS_MAT_Z], &ld_, w, &ncv_, &zero, ctx->M, &ncv_);
  free(w);
  free(c);
  PetscCall(BVSetActiveColumns(eps->V, l, nv));
  PetscCall(BVSetActiveColumns(eps->W, l, nv));
  return PETSC_SUCCESS;
}

/* EPSSolve_KrylovSchur_TwoSided (ks-twosided.c:138-241), as a resumable state
   machine like the default variant's: TSExpand enqueues the nv-step Arnoldi
   expansion of V with Op on the device; TSAfterRight collects it and enqueues
   the expansion of W with A^T up to the nv V reached (the reference expands W
   after V, to V's nv); TSAfterLeft collects that and does the oblique
   Rayleigh-quotient updates, the projected problem, the convergence test and
   the restart, then expands again or finishes.  A serving loop polls the
   device between the three. */
static PetscErrorCode TSBegin(EPS eps, BV V, Mat Op, PetscInt k, PetscInt nv, const BVKrylovOps **ext, PetscInt *m,
                              PetscReal *beta, PetscBool *breakdown)
{
  PetscCall(BVGetKrylovOps(V, ext));
  if (*ext) PetscCall((*ext)->matkrylovbegin(V, Op, k, nv));
  else {   /* the reference's per-step loop, run here; the next phase collects its results */
    *m = nv;
    PetscCall(BVMatKrylov(V, Op, k, m, beta, breakdown));
  }
  eps->nmatmult += nv - k;
  return PETSC_SUCCESS;
}

static PetscErrorCode TSCollect(BV V, DS ds, DSMatType mat, PetscInt k, const BVKrylovOps *ext, PetscInt m,
                                PetscReal mbeta, PetscBool mbreakdown, PetscInt *nv, PetscReal *beta,
                                PetscBool *breakdown)
{
  Mat H;
  if (ext) PetscCall(ext->matkrylovend(V, k, nv, beta, breakdown));
  else {
    *nv = m;
    *beta = mbeta;
    *breakdown = mbreakdown;
  }
  PetscCall(DSGetMat(ds, mat, &H));
  PetscCall(BVCopyHessenberg(V, H, k, *nv));
  PetscCall(DSRestoreMat(ds, mat, &H));
  return PETSC_SUCCESS;
}

static PetscErrorCode TSExpand(EPS eps)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  PetscInt nv, k0 = eps->nconv + ctx->l;
  eps->its++;
  nv = PetscMin(eps->nconv + eps->mpd, eps->ncv);
  PetscCall(DSSetDimensions(eps->ds, nv, eps->nconv, k0));
  PetscCall(TSBegin(eps, eps->V, ctx->ts_Op, k0, nv, &ctx->kry_ext, &ctx->kry_m, &ctx->kry_beta, &ctx->kry_breakdown));
  ctx->nv = nv;
  ctx->phase = KS_PHASE_TS_RIGHT;
  return PETSC_SUCCESS;
}

static PetscErrorCode TSAfterRight(EPS eps)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  PetscInt nv = ctx->nv, k0 = eps->nconv + ctx->l;
  PetscCall(TSCollect(eps->V, eps->ds, DS_MAT_A, k0, ctx->kry_ext, ctx->kry_m, ctx->kry_beta, ctx->kry_breakdown,
                      &nv, &ctx->ts_beta, &ctx->ts_breakdown));
  PetscCall(DSSetDimensions(eps->ds, nv, eps->nconv, k0));
  PetscCall(TSBegin(eps, eps->W, ctx->AT, k0, nv, &ctx->tkry_ext, &ctx->tkry_m, &ctx->tkry_beta,
                    &ctx->tkry_breakdown));
  ctx->nv = nv;
  ctx->phase = KS_PHASE_TS_LEFT;
  return PETSC_SUCCESS;
}

static PetscErrorCode TSStart(EPS eps)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  PetscCall(EPSGetStartVector(eps, 0, NULL));
  PetscCall(EPSGetLeftStartVector(eps, 0, NULL));
  ctx->l = 0;
  memset(ctx->M, 0, sizeof(PetscScalar) * (size_t)eps->ncv * (size_t)eps->ncv);
  PetscCall(MatCreateSeqDenseView(eps->ncv, eps->ncv, ctx->M, eps->ncv, &ctx->ts_Mm));
  PetscCall(STGetOperator(eps->st, &ctx->ts_Op));
  return TSExpand(eps);
}

static PetscErrorCode TSFinish(EPS eps)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  PetscCall(STRestoreOperator(eps->st, &ctx->ts_Op));
  ctx->ts_Op = NULL;
  PetscCall(MatDestroy(&ctx->ts_Mm));
  PetscCall(DSTruncate(eps->ds, eps->nconv, PETSC_TRUE));
  ctx->phase = KS_PHASE_DONE;
  return PETSC_SUCCESS;
}

static PetscErrorCode TSAfterLeft(EPS eps)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  Mat U;
  PetscReal norm, norm2, beta = ctx->ts_beta, betat = 0.0;
  PetscInt l = ctx->l, nv = ctx->nv, nvt = ctx->nv, k = 0, dsn, dsk, k0 = eps->nconv + ctx->l;
  PetscBool breakdownt = PETSC_FALSE, breakdown = ctx->ts_breakdown, breakdownl = PETSC_FALSE;
  double t0;

  PetscCall(TSCollect(eps->W, eps->ds, DS_MAT_B, k0, ctx->tkry_ext, ctx->tkry_m, ctx->tkry_beta,
                      ctx->tkry_breakdown, &nvt, &betat, &breakdownt));
  nv = PetscMin(nv, nvt);
  PetscCall(DSSetDimensions(eps->ds, nv, eps->nconv, k0));
  PetscCall(DSSetState(eps->ds, l == 0 ? DS_STATE_INTERMEDIATE : DS_STATE_RAW));
  breakdown = (breakdown || breakdownt) ? PETSC_TRUE : PETSC_FALSE;

  /* update M, modify the Rayleigh quotients S and T */
  PetscCall(BVSetActiveColumns(eps->V, k0, nv));
  PetscCall(BVSetActiveColumns(eps->W, k0, nv));
  PetscCall(BVMatProject(eps->V, NULL, eps->W, ctx->ts_Mm));
  PetscCall(TwoSidedRQUpdate1(eps, nv, beta, betat));

  /* solve the projected problem */
  t0 = SlepcWallTime();
  PetscCall(DSSolve(eps->ds, eps->eigr, eps->eigi));
  PetscCall(DSSort(eps->ds, eps->eigr, eps->eigi, NULL, NULL, NULL));
  PetscCall(DSUpdateExtraRow(eps->ds));

  /* check convergence */
  PetscCall(BVNormColumn(eps->V, nv, NORM_2, &norm));
  PetscCall(BVNormColumn(eps->W, nv, NORM_2, &norm2));
  PetscCall(EPSKrylovConvergence(eps, PETSC_FALSE, eps->nconv, nv - eps->nconv, beta * norm, betat * norm2, 1.0, &k));
  eps->t_ds += SlepcWallTime() - t0;
  if (k >= eps->nev) eps->reason = EPS_CONVERGED_TOL;
  else if (eps->its >= eps->max_it) eps->reason = EPS_DIVERGED_ITS;

  /* update l */
  if (eps->reason != EPS_CONVERGED_ITERATING || breakdown || k == nv) l = 0;
  else {
    l = PetscMax(1, (PetscInt)((nv - k) * ctx->keep));
    PetscCall(DSGetTruncateSize(eps->ds, k, nv, &l));
  }
  if (!ctx->lock && l > 0) { l += k; k = 0; }

  /* V(:,idx) = V*Q(:,idx), W(:,idx) = W*Z(:,idx) */
  PetscCall(BVSetActiveColumns(eps->V, eps->nconv, nv));
  PetscCall(BVSetActiveColumns(eps->W, eps->nconv, nv));
  PetscCall(DSGetMat(eps->ds, DS_MAT_Q, &U));
  PetscCall(BVMultInPlace(eps->V, U, eps->nconv, k + l));
  PetscCall(DSRestoreMat(eps->ds, DS_MAT_Q, &U));
  PetscCall(DSGetMat(eps->ds, DS_MAT_Z, &U));
  PetscCall(BVMultInPlace(eps->W, U, eps->nconv, k + l));
  PetscCall(DSRestoreMat(eps->ds, DS_MAT_Z, &U));
  if (eps->reason == EPS_CONVERGED_ITERATING && !breakdown) {
    PetscCall(BVCopyColumn(eps->V, nv, k + l));
    PetscCall(BVCopyColumn(eps->W, nv, k + l));
  }
  if (eps->reason == EPS_CONVERGED_ITERATING) {
    if (breakdown || k == nv) {
      if (k < eps->nev) {
        PetscCall(EPSGetStartVector(eps, k, &breakdown));
        PetscCall(EPSGetLeftStartVector(eps, k, &breakdownl));
        if (breakdown || breakdownl) eps->reason = EPS_DIVERGED_BREAKDOWN;
      }
    } else {
      PetscCall(DSGetDimensions(eps->ds, &dsn, NULL, &dsk, NULL));
      PetscCall(DSSetDimensions(eps->ds, dsn, k, dsk));
      PetscCall(DSTruncate(eps->ds, k + l, PETSC_FALSE));
    }
    PetscCall(TwoSidedRQUpdate2(eps, k + l));
  }
  eps->nconv = k;
  ctx->l = l;
  if (eps->monitor) PetscCall(eps->monitor(eps, eps->its, k, eps->eigr, eps->eigi, eps->errest, nv, eps->monitorctx));
  if (eps->reason != EPS_CONVERGED_ITERATING) return TSFinish(eps);
  return TSExpand(eps);
}

static PetscErrorCode EPSSolveBegin_KrylovSchur(EPS eps)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  ctx->its32 = ctx->nconv32 = 0;
  ctx->nmatmult32 = ctx->passes32 = 0;
  ctx->t32 = 0.0;
  if (eps->twosided) return TSStart(eps);
  if (ctx->mixed_tol > 0.0 && ctx->mixed_tol > eps->tol && BVPrecision(eps->V) == SLEPC_AMD_F64) return MixedBegin(eps);
  return KSStart(eps);
}

/* one restart: collects the expansion in flight and enqueues the next one */
static PetscErrorCode EPSSolveStep_KrylovSchur(EPS eps, PetscBool *done)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  if (ctx->phase == KS_PHASE_MIXED32) {
    PetscBool d32;
    PetscCall(EPSSolveStep(ctx->e32, &d32));
    if (d32) {
      PetscCall(MixedEnd(eps));
      PetscCall(KSStart(eps));
    }
  } else if (ctx->phase == KS_PHASE_EXPANDING) PetscCall(KSAfterExpand(eps));
  else if (ctx->phase == KS_PHASE_TS_RIGHT) PetscCall(TSAfterRight(eps));
  else if (ctx->phase == KS_PHASE_TS_LEFT) PetscCall(TSAfterLeft(eps));
  *done = ctx->phase == KS_PHASE_DONE;
  return PETSC_SUCCESS;
}

static PetscErrorCode EPSSolve_KrylovSchur(EPS eps)
{
  PetscBool done = PETSC_FALSE;
  PetscCall(EPSSolveBegin_KrylovSchur(eps));
  while (!done) PetscCall(EPSSolveStep_KrylovSchur(eps, &done));
  return PETSC_SUCCESS;
}

PetscErrorCode EPSKrylovSchurSetMixedPrecision(EPS eps, PetscReal tol32)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  PetscCheck(!strcmp(eps->type, EPSKRYLOVSCHUR), PETSC_ERR_ARG_WRONG, "not a Krylov-Schur solver");
  ctx->mixed_tol = tol32 > 0.0 ? tol32 : 0.0;
  eps->state = EPS_STATE_INITIAL;
  return PETSC_SUCCESS;
}

/* replay each Arnoldi step from a captured graph (the basis' setting, applied at set-up) */
PetscErrorCode EPSKrylovSchurSetStepGraphs(EPS eps, PetscBool flg)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  PetscCheck(!strcmp(eps->type, EPSKRYLOVSCHUR), PETSC_ERR_ARG_WRONG, "not a Krylov-Schur solver");
  ctx->step_graphs = flg;
  if (eps->V && eps->V->data && !strcmp(BVTypeName(eps->V), BVHIPKS)) PetscCall(BVHIPKSSetStepGraphs(eps->V, flg));
  return PETSC_SUCCESS;
}

/* the serving loop's grouped expansions (applied to the basis at set-up) */
PetscErrorCode EPSKrylovSchurSetGrouped_Private(EPS eps, PetscBool flg)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  PetscCheck(!strcmp(eps->type, EPSKRYLOVSCHUR), PETSC_ERR_ARG_WRONG, "not a Krylov-Schur solver");
  ctx->grouped = flg;
  return PETSC_SUCCESS;
}

PetscErrorCode EPSKrylovSchurGetMixedStats(EPS eps, PetscInt *its32, PetscInt *nconv32, int64_t *nmatmult32,
                                           int64_t *passes32, PetscReal *t32)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  PetscCheck(!strcmp(eps->type, EPSKRYLOVSCHUR), PETSC_ERR_ARG_WRONG, "not a Krylov-Schur solver");
  if (its32) *its32 = ctx->its32;
  if (nconv32) *nconv32 = ctx->nconv32;
  if (nmatmult32) *nmatmult32 = ctx->nmatmult32;
  if (passes32) *passes32 = ctx->passes32;
  if (t32) *t32 = ctx->t32;
  return PETSC_SUCCESS;
}

/* EPSComputeVectors_Schur (epsdefault.c:105-169), which for Hermitian problems
   is EPSComputeVectors_Hermitian (epsdefault.c:53-72) */
static PetscErrorCode EPSComputeVectors_KrylovSchur(EPS eps)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  Mat Z;
  PetscInt n;
  if (eps->ishermitian && !ctx->harmonic) return PETSC_SUCCESS;   /* DSHEP: Ritz vectors are already in V */
  if (!eps->nconv) return PETSC_SUCCESS;
  PetscCall(DSVectors(eps->ds, DS_MAT_X, NULL, NULL));
  PetscCall(DSGetMat(eps->ds, DS_MAT_X, &Z));
  if (eps->ishermitian) {
    PetscCall(DSGetDimensions(eps->ds, &n, NULL, NULL, NULL));
    PetscCall(BVMultInPlace(eps->V, Z, 0, n));
    PetscCall(DSRestoreMat(eps->ds, DS_MAT_X, &Z));
    PetscCall(BVNormalize(eps->V, eps->eigi));
  } else {
    PetscCall(BVMultInPlace(eps->V, Z, 0, eps->nconv));
    PetscCall(DSRestoreMat(eps->ds, DS_MAT_X, &Z));
  }
  if (eps->twosided) {   /* left eigenvectors: W = W * Y (epsdefault.c:141-148) */
    PetscCall(DSVectors(eps->ds, DS_MAT_Y, NULL, NULL));
    PetscCall(BVSetActiveColumns(eps->W, 0, eps->nconv));
    PetscCall(DSGetMat(eps->ds, DS_MAT_Y, &Z));
    PetscCall(BVMultInPlace(eps->W, Z, 0, eps->nconv));
    PetscCall(DSRestoreMat(eps->ds, DS_MAT_Y, &Z));
  }
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

/* releases a first-phase FP32 solve abandoned mid-way (the data itself is freed by EPSDestroy) */
static PetscErrorCode EPSDestroy_KrylovSchur(EPS eps)
{
  EPS_KRYLOVSCHUR *ctx = (EPS_KRYLOVSCHUR *)eps->data;
  if (ctx && ctx->e32) PetscCall(EPSDestroy(&ctx->e32));
  if (ctx) {
    free(ctx->g);
    ctx->g = NULL;
    free(ctx->ho_q1);
    ctx->ho_q1 = NULL;
    free(ctx->M);
    ctx->M = NULL;
    PetscCall(MatDestroy(&ctx->AT));
  }
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
  eps->ops.solve = EPSSolve_KrylovSchur;
  eps->ops.solvebegin = EPSSolveBegin_KrylovSchur;
  eps->ops.solvestep = EPSSolveStep_KrylovSchur;
  eps->ops.computevectors = EPSComputeVectors_KrylovSchur;
  eps->ops.destroy = EPSDestroy_KrylovSchur;
  return PETSC_SUCCESS;
}
