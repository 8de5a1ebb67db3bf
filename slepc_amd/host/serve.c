/*
 * serve.c -- a request loop over the Krylov-Schur solver: eigenproblems of
 * uneven size arrive while others run, are admitted against an HBM budget,
 * and every admitted solve keeps its CSR matrix, Krylov basis and projected
 * problem resident until it finishes.
 *
 * No reference counterpart: SLEPc solves one EPS at a time per process.  Each
 * solve here is the unchanged EPSKRYLOVSCHUR on BVHIPKS/MATHIPCSR, driven in
 * parts (EPSSolveBegin / EPSSolveReady / EPSSolveStep / EPSSolveEnd), so a
 * request gives the same result alone and in the mix.
 *
 * Device side: one context (own HIP stream, own DGKS status, own reduction
 * workspace) per slot, created once; all request memory comes from one device
 * arena (ks_pool) sized to the budget, so admission and retirement never call
 * hipMalloc/hipFree.  The expansions of all active requests run concurrently
 * on their streams while the host solves the projected problems of the ones
 * whose expansions have finished.
 */
#include "slepcimpl.h"

#define KS_GROUP_MAX_MEMBERS 16   /* members of one grouped step (the kernel library's limit) */
/* a kernel-library call inside SlepcServerProgress: on failure record it and leave by `out` */
#define KSGO(x) do { int _k = (x); if (_k) { \
      ierr = SlepcError(__LINE__, __func__, __FILE__, PETSC_ERR_GPU, "device library: %s (%d) in %s", \
                        ks_error_string(_k), _k, #x); goto out; } } while (0)

typedef enum { REQ_QUEUED = 0, REQ_ACTIVE = 1, REQ_DONE = 2, REQ_FAILED = 3 } ReqState;

typedef struct {
  SlepcRequest in;
  ReqState state;
  PetscBool joined;           /* its pending grouped expansion waits on its stream already */
  size_t bytes;               /* reserved HBM */
  int slot;
  Mat A;
  EPS eps;
  SlepcResult res;
  PetscScalar *eigr, *eigi;   /* nev each */
  PetscReal *err;             /* relative residual of each returned pair */
  PetscScalar *X;             /* n x nev eigenvectors (want_vectors) */
} Request;

struct _p_SlepcServer {
  size_t budget, reserved, peak_reserved;
  PetscInt max_active, nactive;
  ks_ctx *slot_ctx;
  PetscInt *slot_req;         /* request id per slot, -1 free */
  ks_pool pool;
  ks_group group;             /* grouped Arnoldi steps of all expanding requests */
  PetscBool grouped;
  PetscInt group_depth;       /* grouped steps allowed in flight */
  int64_t group_steps, group_members;
  double t_admit, t_feed, t_step, t_retire, t_query;   /* host seconds in each part of the loop */
  Request *req;
  PetscInt nreq, capreq, next_queued;   /* FIFO admission cursor */
  PetscInt nfinished;
};

/* HBM a request holds while active: CSR (values, indices, row pointers, row
   blocks), the (ncv+1)-column basis, the coefficient buffer, the restart
   matrix, the start vector and the residual-check work vectors; each block
   rounded to the pool alignment. */
static size_t RequestBytes(const SlepcRequest *r)
{
  const size_t al = 256;
  size_t n = (size_t)r->n, nnz = (size_t)r->rowptr[r->n], ld = (size_t)BVHIPKSLeadingDimension(r->n, 8);
  size_t ld32 = (size_t)BVHIPKSLeadingDimension(r->n, 4);
  size_t m = (size_t)r->ncv + 1, b = 0, vs = r->val32 ? 4 : 8;
  int mixed = r->mixed_tol > 0.0 && r->mixed_tol > r->tol;
#define ROUND(x) (((x) + al - 1) / al * al)
  b += ROUND(4 * (n + 1)) + ROUND(4 * (nnz ? nnz : 1)) + ROUND(vs * (nnz ? nnz : 1)) + ROUND(4 * (n + 2));
  b += ROUND(8 * ld * m) + ROUND(8 * m * m) + ROUND(8 * (m + 2)) + ROUND(8 * m * m);
  b += 7 * ROUND(8 * n);
  if (mixed) {   /* the FP32 phase: its basis, buffers and vectors, and an FP32 copy of FP64 values */
    b += ROUND(4 * ld32 * m) + ROUND(8 * m * m) + ROUND(8 * (m + 2)) + ROUND(8 * m * m) + 7 * ROUND(4 * n);
    if (!r->val32) b += ROUND(4 * (n + 1)) + ROUND(4 * (nnz ? nnz : 1)) * 2 + ROUND(4 * (n + 2));
  }
  if (r->twosided) {   /* A^T (CSR with its row-block schedule), the left basis, buffers and vectors, M */
    b += ROUND(4 * (n + 1)) + ROUND(4 * (nnz ? nnz : 1)) + ROUND(8 * (nnz ? nnz : 1)) + ROUND(4 * (n + 2));
    b += ROUND(8 * ld * m) + ROUND(8 * m * m) + ROUND(8 * (m + 2)) + ROUND(8 * m * m) + 7 * ROUND(8 * n);
  }
#undef ROUND
  return b + 64 * 1024;
}

PetscErrorCode SlepcServerCreate(size_t hbm_budget, PetscInt max_active, SlepcServer *out)
{
  SlepcServer s;
  PetscCall(SlepcCheckInitialized());
  PetscCheck(out && hbm_budget > 0 && max_active > 0, PETSC_ERR_ARG_OUTOFRANGE, "bad server arguments");
  PetscCall(PetscMallocZero(sizeof(*s), &s));
  s->budget = hbm_budget;
  s->max_active = max_active;
  PetscCall(PetscMallocZero(sizeof(ks_ctx) * (size_t)max_active, &s->slot_ctx));
  PetscCall(PetscMallocZero(sizeof(PetscInt) * (size_t)max_active, &s->slot_req));
  for (PetscInt i = 0; i < max_active; i++) {
    KSCall(ks_ctx_create(&s->slot_ctx[i]));
    s->slot_req[i] = -1;
  }
  /* the arena holds the budget plus per-slot pinned staging of the restart
     matrices; 1/16 extra absorbs fragmentation */
  KSCall(ks_pool_create(&s->pool, hbm_budget + hbm_budget / 16, (size_t)max_active * 4 * 1024 * 1024));
  KSCall(ks_group_create(&s->group));
  s->grouped = PETSC_TRUE;
  s->group_depth = 4;
  *out = s;
  return PETSC_SUCCESS;
}

PetscErrorCode SlepcServerSubmit(SlepcServer s, const SlepcRequest *r, PetscInt *id)
{
  Request *q;
  PetscCheck(s && r && id, PETSC_ERR_ARG_NULL, "null argument");
  PetscCheck(r->n > 0 && r->rowptr && r->colidx && (r->val || r->val32), PETSC_ERR_ARG_NULL, "request without a matrix");
  PetscCheck(r->nev >= 1 && r->ncv > r->nev && r->ncv <= r->n && r->ncv + 1 <= KS_MAX_COLS, PETSC_ERR_ARG_OUTOFRANGE,
             "request nev=%d ncv=%d n=%d not supported", (int)r->nev, (int)r->ncv, (int)r->n);
  PetscCheck(!r->twosided || (r->val && !r->val32 && !(r->mixed_tol > r->tol)), PETSC_ERR_SUP,
             "a two-sided request needs FP64 values and no FP32 phase");
  if (s->nreq == s->capreq) {
    PetscInt cap = s->capreq ? 2 * s->capreq : 64;
    Request *nr = (Request *)realloc(s->req, sizeof(Request) * (size_t)cap);
    PetscCheck(nr, PETSC_ERR_MEM, "out of host memory");
    memset(nr + s->capreq, 0, sizeof(Request) * (size_t)(cap - s->capreq));
    s->req = nr;
    s->capreq = cap;
  }
  q = &s->req[s->nreq];
  memset(q, 0, sizeof(*q));
  q->in = *r;
  if (q->in.t_arrival <= 0.0) q->in.t_arrival = SlepcWallTime();
  q->bytes = RequestBytes(r);
  PetscCheck(q->bytes <= s->budget, PETSC_ERR_ARG_OUTOFRANGE,
             "request needs %zu bytes of HBM, more than the whole budget (%zu)", q->bytes, s->budget);
  q->state = REQ_QUEUED;
  q->slot = -1;
  q->res.t_arrival = q->in.t_arrival;
  *id = s->nreq++;
  return PETSC_SUCCESS;
}

/* builds the request's objects on its slot and starts the solve */
static PetscErrorCode RequestAdmit(SlepcServer s, PetscInt id, int slot)
{
  Request *q = &s->req[id];
  Vec v0;
  BV V;
  q->slot = slot;
  q->res.t_admit = SlepcWallTime();
  if (q->in.val32) {
    PetscCall(MatCreate(&q->A));
    PetscCall(MatSetSizes(q->A, q->in.n, q->in.n, q->in.n, q->in.n));
    PetscCall(MatSetType(q->A, MATHIPCSR));
    PetscCall(MatHIPCSRSetArraysF32(q->A, q->in.rowptr, q->in.colidx, q->in.val32));
  } else PetscCall(MatCreateHIPCSR(q->in.n, q->in.rowptr, q->in.colidx, q->in.val, &q->A));
  PetscCall(EPSCreate(&q->eps));
  PetscCall(EPSSetOperators(q->eps, q->A, NULL));
  PetscCall(EPSSetProblemType(q->eps, q->in.hermitian ? EPS_HEP : EPS_NHEP));
  PetscCall(EPSSetWhichEigenpairs(q->eps, q->in.which ? q->in.which : EPS_LARGEST_MAGNITUDE));
  PetscCall(EPSSetDimensions(q->eps, q->in.nev, q->in.ncv, PETSC_DETERMINE));
  PetscCall(EPSSetTolerances(q->eps, q->in.tol > 0 ? q->in.tol : 1e-10, q->in.max_it > 0 ? q->in.max_it : PETSC_DETERMINE));
  PetscCall(EPSSetType(q->eps, EPSKRYLOVSCHUR));
  if (q->in.mixed_tol > 0.0) PetscCall(EPSKrylovSchurSetMixedPrecision(q->eps, q->in.mixed_tol));
  PetscCall(EPSGetBV(q->eps, &V));
  PetscCall(BVSetType(V, BVHIPKS));
  /* the two-sided variant expands two bases per restart; it runs on its own stream */
  if (q->in.twosided) PetscCall(EPSSetTwoSided(q->eps, PETSC_TRUE));
  PetscCall(EPSKrylovSchurSetGrouped_Private(q->eps, s->grouped && !q->in.twosided));
  if (q->in.v0) {
    PetscCall(VecCreateHIP(q->in.n, SLEPC_AMD_F64, &v0));
    PetscCall(VecSetValuesHost(v0, q->in.v0));
    PetscCall(EPSSetInitialSpace(q->eps, 1, &v0));
    PetscCall(VecDestroy(&v0));
  }
  PetscCall(EPSSolveBegin(q->eps));
  q->state = REQ_ACTIVE;
  s->slot_req[slot] = id;
  s->nactive++;
  return PETSC_SUCCESS;
}

/* a request whose admission failed: whatever RequestAdmit built is destroyed,
   its reserved bytes go back to the budget and it finishes with the error code
   in its result, so the queue moves on */
static PetscErrorCode RequestAbort(SlepcServer s, PetscInt id, PetscErrorCode code)
{
  Request *q = &s->req[id];
  if (q->slot >= 0) KSCall(ks_ctx_sync(s->slot_ctx[q->slot]));
  PetscCall(EPSDestroy(&q->eps));
  PetscCall(MatDestroy(&q->A));
  if (q->slot >= 0 && s->slot_req[q->slot] == id) {
    s->slot_req[q->slot] = -1;
    s->nactive--;
  }
  q->slot = -1;
  s->reserved -= q->bytes;
  q->state = REQ_FAILED;
  q->res.error = code;
  q->res.reason = EPS_DIVERGED_BREAKDOWN;
  q->res.t_done = SlepcWallTime();
  s->nfinished++;
  q->in.rowptr = q->in.colidx = NULL;
  q->in.val = NULL;
  q->in.val32 = NULL;
  q->in.v0 = NULL;
  return PETSC_SUCCESS;
}

/* collects eigenvalues, residuals (and vectors), then frees the device state */
static PetscErrorCode RequestRetire(SlepcServer s, PetscInt id)
{
  Request *q = &s->req[id];
  PetscInt nev = q->in.nev, nret;
  Vec xr = NULL, xi = NULL;
  PetscCall(EPSSolveEnd(q->eps));
  q->res.nconv = q->eps->nconv;
  q->res.its = q->eps->its;
  q->res.reason = q->eps->reason;
  q->res.nmatmult = q->eps->nmatmult;
  PetscCall(BVHIPKSGetPasses(q->eps->V, &q->res.gs_passes));
  q->res.t_host_ds = q->eps->t_ds;
  nret = PetscMin(nev, q->eps->nconv);
  q->res.nret = nret;
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)nev, &q->eigr));
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)nev, &q->eigi));
  PetscCall(PetscMallocZero(sizeof(PetscReal) * (size_t)nev, &q->err));
  if (q->in.want_vectors) {
    PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)q->in.n * (size_t)nev, &q->X));
    PetscCall(VecCreateHIP(q->in.n, SLEPC_AMD_F64, &xr));
    PetscCall(VecCreateHIP(q->in.n, SLEPC_AMD_F64, &xi));
  }
  q->res.ngood = 0;
  for (PetscInt i = 0; i < nret; i++) {
    PetscCall(EPSGetEigenpair(q->eps, i, &q->eigr[i], &q->eigi[i], xr, xi));
    PetscCall(EPSComputeError(q->eps, i, EPS_ERROR_RELATIVE, &q->err[i]));
    if (q->err[i] <= q->eps->tol) q->res.ngood++;
    if (xr) PetscCall(VecGetValuesHost(xr, q->X + (size_t)i * (size_t)q->in.n));
  }
  if (xr) PetscCall(VecDestroy(&xr));
  if (xi) PetscCall(VecDestroy(&xi));
  PetscCall(EPSDestroy(&q->eps));
  PetscCall(MatDestroy(&q->A));
  KSCall(ks_ctx_sync(s->slot_ctx[q->slot]));
  q->res.t_done = SlepcWallTime();
  q->state = REQ_DONE;
  s->slot_req[q->slot] = -1;
  q->slot = -1;
  s->nactive--;
  s->reserved -= q->bytes;
  s->nfinished++;
  /* the matrix belongs to the caller again */
  q->in.rowptr = q->in.colidx = NULL;
  q->in.val = NULL;
  q->in.val32 = NULL;
  q->in.v0 = NULL;
  return PETSC_SUCCESS;
}

/* Enqueue grouped Arnoldi steps -- one launch per kernel kind for the current
   column of every request whose expansion is pending -- until `depth` steps
   are in flight.  Called between the host's restart work so the device
   keeps running while the projected problems are solved. */
static PetscErrorCode GroupFeed(SlepcServer s, PetscBool *moved, PetscInt depth)
{
  PetscErrorCode ierr = PETSC_SUCCESS;
  if (!s->grouped) return PETSC_SUCCESS;
  for (;;) {
    ks_group_member mem[KS_GROUP_MAX_MEMBERS];
    PetscInt who[KS_GROUP_MAX_MEMBERS], nm = 0;
    int inflight = 0, refine = -1;
    double eta = 0.0;
    KSGO(ks_group_inflight(s->group, &inflight));
    if (inflight >= depth) break;
    for (int i = 0; i < s->max_active && nm < KS_GROUP_MAX_MEMBERS; i++) {
      PetscInt id = s->slot_req[i];
      BV V;
      if (id < 0) continue;
      PetscBool pending;
      V = s->req[id].eps->V;
      ierr = BVHIPKSGroupPending(V, &pending);
      if (ierr) goto out;
      if (!pending) continue;
      if (refine < 0) { refine = (int)V->orthog_ref; eta = V->orthog_eta; }
      if ((int)V->orthog_ref != refine || V->orthog_eta != eta) continue;   /* next round */
      if (!s->req[id].joined) {
        KSGO(ks_group_wait_ctx(s->group, s->slot_ctx[i]));   /* after its restart work */
        s->req[id].joined = PETSC_TRUE;
      }
      ierr = BVHIPKSGroupMember(V, &mem[nm]);
      if (ierr) goto out;
      who[nm++] = id;
    }
    if (!nm) break;
    KSGO(ks_group_step(s->group, (int)nm, mem, (ks_refine)refine, eta));
    s->group_steps++;
    s->group_members += nm;
    *moved = PETSC_TRUE;
    for (PetscInt t = 0; t < nm; t++) {
      Request *q = &s->req[who[t]];
      PetscBool fin;
      ierr = BVHIPKSGroupAdvance(q->eps->V, &fin);
      if (ierr) goto out;
      if (fin) {   /* its stream picks up after the group's last step */
        KSGO(ks_ctx_wait_group(s->slot_ctx[q->slot], s->group));
        q->joined = PETSC_FALSE;
        ierr = BVHIPKSPostExpansion(q->eps->V);
        if (ierr) goto out;
      }
    }
  }
out:
  return ierr;
}

/* One pass of the loop, never waiting on the device: admit queued requests in
   arrival order while a slot and the budget allow, then advance every active
   solve whose device work has finished by one restart, retiring the finished
   ones.  *progressed says whether anything moved. */
PetscErrorCode SlepcServerProgress(SlepcServer s, PetscBool *progressed, PetscInt *nactive, PetscInt *nqueued,
                                   PetscInt *nfinished)
{
  ks_ctx prev_ctx;
  ks_pool prev_pool;
  PetscErrorCode ierr = PETSC_SUCCESS;
  PetscBool moved = PETSC_FALSE;
  double t0;
  PetscCheck(s, PETSC_ERR_ARG_NULL, "null server");
  prev_pool = ks_pool_use(s->pool);
  prev_ctx = SlepcAMDSetCurrentContext(NULL);
  /* admission, FIFO: a large request at the head waits for memory rather than
     being overtaken indefinitely */
  while (s->next_queued < s->nreq && s->nactive < s->max_active) {
    Request *q = &s->req[s->next_queued];
    int slot = -1;
    if (s->reserved + q->bytes > s->budget) break;
    for (int i = 0; i < s->max_active; i++)
      if (s->slot_req[i] < 0) { slot = i; break; }
    if (slot < 0) break;
    s->reserved += q->bytes;
    if (s->reserved > s->peak_reserved) s->peak_reserved = s->reserved;
    (void)SlepcAMDSetCurrentContext(s->slot_ctx[slot]);
    t0 = SlepcWallTime();
    ierr = RequestAdmit(s, s->next_queued, slot);
    s->t_admit += SlepcWallTime() - t0;
    if (ierr) {   /* the request fails alone: its objects and its budget are released */
      ierr = RequestAbort(s, s->next_queued, ierr);
      if (ierr) goto out;
    }
    s->next_queued++;
    moved = PETSC_TRUE;
  }
  t0 = SlepcWallTime();
  ierr = GroupFeed(s, &moved, s->group_depth);
  s->t_feed += SlepcWallTime() - t0;
  if (ierr) goto out;
  for (int i = 0; i < s->max_active; i++) {
    PetscInt id = s->slot_req[i];
    PetscBool ready, done;
    if (id < 0) continue;
    (void)SlepcAMDSetCurrentContext(s->slot_ctx[i]);
    t0 = SlepcWallTime();
    ierr = EPSSolveReady(s->req[id].eps, &ready);
    s->t_query += SlepcWallTime() - t0;
    if (ierr) goto out;
    if (!ready) continue;
    t0 = SlepcWallTime();
    ierr = EPSSolveStep(s->req[id].eps, &done);
    s->t_step += SlepcWallTime() - t0;
    if (ierr) goto out;
    moved = PETSC_TRUE;
    /* keep the device fed while the host works through the restarts, but
       only when the group has run dry: fuller steps cost the host less */
    t0 = SlepcWallTime();
    ierr = GroupFeed(s, &moved, 1);
    s->t_feed += SlepcWallTime() - t0;
    if (ierr) goto out;
    if (done) {
      t0 = SlepcWallTime();
      ierr = RequestRetire(s, id);
      s->t_retire += SlepcWallTime() - t0;
      if (ierr) goto out;
    }
  }
out:
  (void)SlepcAMDSetCurrentContext(prev_ctx);
  (void)ks_pool_use(prev_pool);
  if (ierr) return SlepcError(__LINE__, __func__, __FILE__, ierr, NULL);
  if (progressed) *progressed = moved;
  if (nactive) *nactive = s->nactive;
  if (nqueued) *nqueued = s->nreq - s->next_queued;
  if (nfinished) *nfinished = s->nfinished;
  return PETSC_SUCCESS;
}

PetscErrorCode SlepcServerGetResult(SlepcServer s, PetscInt id, PetscBool *done, SlepcResult *res,
                                    PetscScalar *eigr, PetscScalar *eigi, PetscReal *err, PetscScalar *X)
{
  Request *q;
  PetscCheck(s && id >= 0 && id < s->nreq, PETSC_ERR_ARG_OUTOFRANGE, "no request %d", (int)id);
  q = &s->req[id];
  if (done) *done = (q->state == REQ_DONE || q->state == REQ_FAILED) ? PETSC_TRUE : PETSC_FALSE;
  if (q->state == REQ_FAILED) {   /* no pairs: the error code says why */
    if (res) *res = q->res;
    return PETSC_SUCCESS;
  }
  if (q->state != REQ_DONE) return PETSC_SUCCESS;
  if (res) *res = q->res;
  if (eigr) memcpy(eigr, q->eigr, sizeof(PetscScalar) * (size_t)q->res.nret);
  if (eigi) memcpy(eigi, q->eigi, sizeof(PetscScalar) * (size_t)q->res.nret);
  if (err) memcpy(err, q->err, sizeof(PetscReal) * (size_t)q->res.nret);
  if (X) {
    PetscCheck(q->X, PETSC_ERR_ARG_WRONG, "request %d was submitted without want_vectors", (int)id);
    memcpy(X, q->X, sizeof(PetscScalar) * (size_t)q->in.n * (size_t)q->res.nret);
  }
  return PETSC_SUCCESS;
}

/* A whole stream served from C: request i is submitted once SlepcServerClock()
   passes the call's start plus req[i].t_arrival (an offset in seconds; the
   offsets ascend), the loop runs until every request has finished, and
   result i lands in res[i] and the arrays eigr[i], eigi[i], err[i] (nev
   entries each) and X[i] (n x nev, or NULL).  The latency clock of each
   request starts at its arrival. */
PetscErrorCode SlepcServerRun(SlepcServer s, PetscInt nreq, const SlepcRequest *req, SlepcResult *res,
                              PetscScalar **eigr, PetscScalar **eigi, PetscReal **err, PetscScalar **X, double *wall)
{
  PetscErrorCode ierr = PETSC_SUCCESS;
  PetscInt *ids = NULL, next = 0, ndone = 0, seen = 0;
  PetscBool *got = NULL;
  double t0;
  PetscCheck(s && (nreq == 0 || (req && res && eigr && eigi && err)), PETSC_ERR_ARG_NULL, "null argument");
  for (PetscInt i = 1; i < nreq; i++)
    PetscCheck(req[i].t_arrival >= req[i - 1].t_arrival, PETSC_ERR_ARG_WRONG, "arrival offsets must ascend");
  PetscCall(PetscMallocZero(sizeof(PetscInt) * (size_t)(nreq ? nreq : 1), &ids));
  ierr = PetscMallocZero(sizeof(PetscBool) * (size_t)(nreq ? nreq : 1), &got);
  if (ierr) { free(ids); return ierr; }
  t0 = SlepcServerClock();
  while (ndone < nreq) {
    const double now = SlepcServerClock();
    PetscInt nf = 0;
    while (next < nreq && t0 + req[next].t_arrival <= now) {
      SlepcRequest r = req[next];
      r.t_arrival = t0 + req[next].t_arrival;
      ierr = SlepcServerSubmit(s, &r, &ids[next]);
      if (ierr) goto out;
      next++;
    }
    ierr = SlepcServerProgress(s, NULL, NULL, NULL, &nf);
    if (ierr) goto out;
    if (nf == seen) continue;
    seen = nf;
    for (PetscInt i = 0; i < next; i++) {
      PetscBool d = PETSC_FALSE;
      if (got[i]) continue;
      ierr = SlepcServerGetResult(s, ids[i], &d, &res[i], eigr[i], eigi[i], err[i], X ? X[i] : NULL);
      if (ierr) goto out;
      if (d) { got[i] = PETSC_TRUE; ndone++; }
    }
  }
out:
  if (wall) *wall = SlepcServerClock() - t0;
  free(ids);
  free(got);
  return ierr;
}

PetscErrorCode SlepcServerSetGrouped(SlepcServer s, PetscBool grouped)
{
  PetscCheck(s, PETSC_ERR_ARG_NULL, "null server");
  PetscCheck(s->nreq == 0, PETSC_ERR_ORDER, "set the grouping before the first request");
  s->grouped = grouped;
  return PETSC_SUCCESS;
}

PetscErrorCode SlepcServerGetGroupStats(SlepcServer s, int64_t *steps, int64_t *member_steps)
{
  PetscCheck(s, PETSC_ERR_ARG_NULL, "null server");
  if (steps) *steps = s->group_steps;
  if (member_steps) *member_steps = s->group_members;
  return PETSC_SUCCESS;
}

PetscErrorCode SlepcServerGetHostTimes(SlepcServer s, double t[5])
{
  PetscCheck(s && t, PETSC_ERR_ARG_NULL, "null argument");
  t[0] = s->t_admit;
  t[1] = s->t_feed;
  t[2] = s->t_step;
  t[3] = s->t_retire;
  t[4] = s->t_query;
  return PETSC_SUCCESS;
}

PetscErrorCode SlepcServerGetStats(SlepcServer s, size_t *peak_reserved, size_t *pool_peak, size_t *budget)
{
  PetscCheck(s, PETSC_ERR_ARG_NULL, "null server");
  if (peak_reserved) *peak_reserved = s->peak_reserved;
  if (budget) *budget = s->budget;
  if (pool_peak) KSCall(ks_pool_info(s->pool, NULL, pool_peak, NULL, NULL));
  return PETSC_SUCCESS;
}

PetscErrorCode SlepcServerDestroy(SlepcServer *ps)
{
  SlepcServer s;
  if (!ps || !*ps) return PETSC_SUCCESS;
  s = *ps;
  {
    ks_pool prev = ks_pool_use(s->pool);
    for (int i = 0; i < s->max_active; i++) {
      PetscInt id = s->slot_req[i];
      if (id < 0) continue;
      Request *q = &s->req[id];
      ks_ctx prev_ctx = SlepcAMDSetCurrentContext(s->slot_ctx[i]);
      KSCall(ks_ctx_sync(s->slot_ctx[i]));
      PetscCall(EPSDestroy(&q->eps));
      PetscCall(MatDestroy(&q->A));
      (void)SlepcAMDSetCurrentContext(prev_ctx);
      q->state = REQ_FAILED;
    }
    (void)ks_pool_use(prev);
  }
  for (PetscInt i = 0; i < s->nreq; i++) {
    free(s->req[i].eigr);
    free(s->req[i].eigi);
    free(s->req[i].err);
    free(s->req[i].X);
  }
  free(s->req);
  KSCall(ks_group_destroy(s->group));
  KSCall(ks_pool_destroy(s->pool));
  for (int i = 0; i < s->max_active; i++) KSCall(ks_ctx_destroy(s->slot_ctx[i]));
  free(s->slot_ctx);
  free(s->slot_req);
  free(s);
  *ps = NULL;
  return PETSC_SUCCESS;
}

double SlepcServerClock(void) { return SlepcWallTime(); }
