/*
 * bvhipks.c -- BVHIPKS, the MI355X basis-vector type (registered with
 * BVRegister("hipks", BVCreate_HIPKS)).  It replaces the reference's BVMAT op
 * table (src/sys/classes/bv/impls/mat/bvmat.c:385-524) and its HIP variant
 * (mathip.hip.cpp / bvhip.hip.cpp, hipBLAS calls) with the kernels of
 * libks_hip.so:
 *
 *   storage   V: ld x (nc+m) column-major in HBM, ld = n rounded up to 64
 *             coefficient buffer: (nc+m) x m FP64 in HBM (host mirror refreshed
 *             once per expansion), Q staging for BVMultInPlace in HBM
 *   ops       dotvec/multvec (tall-skinny kernels), scale/norm/copy tails,
 *             multinplace (FP64 MFMA, no n x (e-s) temporary), random columns,
 *             orthonormalizecolumn (all CGS passes + DGKS + scaling on the device),
 *             matkrylovbegin/end (a whole Arnoldi/Lanczos expansion enqueued
 *             at once, one host sync when its results are collected).
 *
 * Storage precision is FP64 by default; BVHIPKSSetPrecision(bv, SLEPC_AMD_F32)
 * keeps V in FP32 (mixed-precision phase) while coefficients stay FP64.
 */
#include "slepcimpl.h"

typedef struct {
  void *dV;              /* ld x (nc+m) */
  double *d_buf;         /* (nc+m) x m coefficient buffer */
  double *d_q;           /* small coefficient vector for dotvec/multvec */
  double *d_Q;           /* ldq x ldq staging of the dense Q of BVMultInPlace */
  double *h_Q;           /* pinned host staging of Q */
  size_t Qcap;           /* elements of d_Q / h_Q */
  double *d_G;           /* block dot result */
  size_t Gcap;
  PetscInt ncols;        /* nc + m columns allocated */
  /* captured Arnoldi steps: step[j] makes column j+1 from column j; valid for
     the matrix step_csr and the orthogonalization settings they were made with */
  ks_graph *step;
  ks_csr step_csr;
  int step_refine;
  double step_eta;
  PetscBool step_failed;   /* a capture failed: steps run as plain launches */
} BV_HIPKS;

static void StepGraphsClear(BV_HIPKS *h, PetscInt m)
{
  if (!h->step) return;
  for (PetscInt j = 0; j < m; j++) (void)ks_graph_destroy(h->step[j]);
  free(h->step);
  h->step = NULL;
}

static ks_prec kp(BV bv) { return bv->prec == SLEPC_AMD_F64 ? KS_F64 : KS_F32; }
static size_t esz(BV bv) { return SlepcPrecSize(bv->prec); }
static void *colptr(BV bv, PetscInt j)
{
  BV_HIPKS *h = (BV_HIPKS *)bv->data;
  return (char *)h->dV + (size_t)(bv->nc + j) * (size_t)bv->ld * esz(bv);
}

static PetscErrorCode BVGetColumn_HIPKS(BV bv, PetscInt j, Vec *v)
{
  if (!*v) PetscCall(VecCreateHIPWithArray(bv->n, bv->prec, colptr(bv, j), v));
  else (*v)->d = colptr(bv, j), (*v)->n = bv->n, (*v)->prec = bv->prec;
  return PETSC_SUCCESS;
}

static PetscErrorCode BVScale_HIPKS(BV bv, PetscInt j, PetscScalar alpha)
{
  KSCall(ks_bv_scale(bv->ctx, kp(bv), bv->n, colptr(bv, j), alpha));
  return PETSC_SUCCESS;
}

static PetscErrorCode BVNorm_HIPKS(BV bv, PetscInt j, NormType type, PetscReal *val)
{
  PetscCheck(type == NORM_2 || type == NORM_FROBENIUS, PETSC_ERR_SUP, "only the 2-norm of a column is implemented");
  KSCall(ks_bv_norm(bv->ctx, kp(bv), bv->n, colptr(bv, j), val));
  return PETSC_SUCCESS;
}

static PetscErrorCode BVCopyColumn_HIPKS(BV bv, PetscInt j, PetscInt i)
{
  KSCall(ks_bv_copy(bv->ctx, kp(bv), bv->n, colptr(bv, j), colptr(bv, i)));
  return PETSC_SUCCESS;
}

/* BVSetRandomColumn: counter-based uniform(-1,1), seed advanced per call */
static PetscErrorCode BVSetRandomColumn_HIPKS(BV bv, PetscInt j)
{
  bv->rcount++;
  KSCall(ks_bv_random(bv->ctx, kp(bv), bv->n, colptr(bv, j), bv->rseed + 1000003ULL * bv->rcount));
  return PETSC_SUCCESS;
}

/* q = V(:,l:k)^T y (columns counted after constraints; l may be -nc) */
static PetscErrorCode BVDotVec_HIPKS(BV bv, Vec y, PetscScalar *q)
{
  BV_HIPKS *h = (BV_HIPKS *)bv->data;
  PetscInt k = bv->k - bv->l;
  PetscCheck(y->prec == bv->prec, PETSC_ERR_ARG_WRONG, "vector precision differs from the basis");
  if (k <= 0) return PETSC_SUCCESS;
  KSCall(ks_bv_dotvec(bv->ctx, kp(bv), bv->n, colptr(bv, bv->l), bv->ld, k, y->d, h->d_q));
  KSCall(ks_memcpy_dtoh(bv->ctx, q, h->d_q, sizeof(double) * (size_t)k));
  return PETSC_SUCCESS;
}

/* y = beta y + alpha V(:,l:k) q */
static PetscErrorCode BVMultVec_HIPKS(BV bv, PetscScalar alpha, PetscScalar beta, Vec y, PetscScalar *q)
{
  BV_HIPKS *h = (BV_HIPKS *)bv->data;
  PetscInt k = bv->k - bv->l;
  PetscCheck(y->prec == bv->prec, PETSC_ERR_ARG_WRONG, "vector precision differs from the basis");
  if (k > 0) KSCall(ks_memcpy_htod(bv->ctx, h->d_q, q, sizeof(double) * (size_t)k));
  KSCall(ks_bv_multvec(bv->ctx, kp(bv), bv->n, alpha, colptr(bv, bv->l), bv->ld, k > 0 ? k : 0, h->d_q, beta, y->d));
  return PETSC_SUCCESS;
}

/* BVMultInPlace: V(:,s:e) = V(:,l:k) Q(l:k,s:e); Q staged to HBM asynchronously */
static PetscErrorCode BVMultInPlace_HIPKS(BV bv, Mat Q, PetscInt s, PetscInt e)
{
  BV_HIPKS *h = (BV_HIPKS *)bv->data;
  Mat_SeqDense *qd = (Mat_SeqDense *)Q->data;
  PetscInt ldq = qd->lda, K = bv->k - bv->l;
  size_t need = (size_t)ldq * (size_t)e;
  if (s >= e || !bv->n) return PETSC_SUCCESS;
  PetscCheck(K >= 0, PETSC_ERR_ARG_OUTOFRANGE, "empty active column range");
  if (need > h->Qcap) {
    KSCall(ks_ctx_sync(bv->ctx));
    KSCall(ks_free(h->d_Q));
    KSCall(ks_host_free(h->h_Q));
    h->Qcap = need;
    KSCall(ks_malloc((void **)&h->d_Q, sizeof(double) * need));
    KSCall(ks_host_alloc((void **)&h->h_Q, sizeof(double) * need));
  }
  /* the pinned staging may still feed the previous copy: wait for the stream */
  KSCall(ks_ctx_sync(bv->ctx));
  memcpy(h->h_Q, qd->v, sizeof(double) * need);
  KSCall(ks_memcpy_htod_async(bv->ctx, h->d_Q, h->h_Q, sizeof(double) * need));
  KSCall(ks_bv_multinplace(bv->ctx, kp(bv), bv->n, colptr(bv, bv->l), bv->ld, K,
                           h->d_Q + (size_t)bv->l * ldq + bv->l, ldq, s - bv->l, e - bv->l));
  return PETSC_SUCCESS;
}

/* G(0:ye-ys, 0:xe-xs) = Y(:, ys:ye)^T X(:, xs:xe) (the block dot, ks_bv_gram),
   through a device array of the basis X, copied into host memory */
static PetscErrorCode BVGram_HIPKS(BV X, PetscInt xs, PetscInt xe, BV Y, PetscInt ys, PetscInt ye, PetscScalar *G,
                                   PetscInt ldg)
{
  BV_HIPKS *h = (BV_HIPKS *)X->data;
  const PetscInt p = xe - xs, q = ye - ys;
  size_t need = (size_t)p * (size_t)q;
  double *tmp;
  PetscCheck(!strcmp(Y->type, X->type) && Y->prec == X->prec && Y->n == X->n, PETSC_ERR_ARG_INCOMP,
             "block dot of bases of different type, precision or length");
  PetscCheck(xs >= -X->nc && xe <= X->m && ys >= -Y->nc && ye <= Y->m, PETSC_ERR_ARG_OUTOFRANGE, "column range out of bounds");
  if (p <= 0 || q <= 0) return PETSC_SUCCESS;
  if (need > h->Gcap) {
    KSCall(ks_ctx_sync(X->ctx));
    KSCall(ks_free(h->d_G));
    h->d_G = NULL;
    KSCall(ks_malloc((void **)&h->d_G, sizeof(double) * need));
    h->Gcap = need;
  }
  /* Y's columns live in Y's storage; both bases are on X's stream */
  KSCall(ks_bv_gram(X->ctx, kp(X), X->n, colptr(X, xs), X->ld, p, colptr(Y, ys), Y->ld, q, h->d_G, q));
  PetscCall(PetscMallocZero(sizeof(double) * need, &tmp));
  {
    int e = ks_memcpy_dtoh(X->ctx, tmp, h->d_G, sizeof(double) * need);
    if (e) { free(tmp); KSCall(e); }
  }
  for (PetscInt j = 0; j < p; j++) memcpy(G + (size_t)j * ldg, tmp + (size_t)j * q, sizeof(double) * (size_t)q);
  free(tmp);
  return PETSC_SUCCESS;
}

static PetscErrorCode BVGetBuffer_HIPKS(BV bv)
{
  BV_HIPKS *h = (BV_HIPKS *)bv->data;
  KSCall(ks_memcpy_dtoh(bv->ctx, bv->buffer, h->d_buf, sizeof(double) * (size_t)h->ncols * (size_t)bv->m));
  return PETSC_SUCCESS;
}

static PetscErrorCode StatusCheck(const ks_status *st)
{
  PetscCheck(st->error != KS_ERR_INNER, PETSC_ERR_USER_INPUT, "Invalid inner product (negative self dot product)");
  PetscCheck(!st->error, PETSC_ERR_GPU, "device status error %d", st->error);
  return PETSC_SUCCESS;
}

/* Whole-column orthonormalization on the device: BVOrthogonalizeGS with
   BVOrthogonalizeCGS1 passes (bvorthog.c:91-217), then the 1/norm scaling of
   BVOrthonormalizeColumn (bvorthog.c:416-422) when `scale`. */
static PetscErrorCode BVOrthonormalizeColumn_HIPKS(BV bv, PetscInt j, PetscBool scale, PetscBool safe,
                                                   PetscBool async, PetscReal *norm, PetscBool *lindep)
{
  BV_HIPKS *h = (BV_HIPKS *)bv->data;
  ks_status st;
  PetscCheck(bv->nc + j < KS_MAX_COLS, PETSC_ERR_SUP, "at most %d columns are supported", KS_MAX_COLS);
  if (!async) KSCall(ks_status_reset(bv->ctx));
  KSCall(ks_bv_orthonormalize_column(bv->ctx, kp(bv), bv->n, h->dV, bv->ld, bv->nc + j, h->d_buf, h->ncols, j, j,
                                     (ks_refine)bv->orthog_ref, bv->orthog_eta, scale ? 1 : 0, safe ? 1 : 0));
  if (async) return PETSC_SUCCESS;
  KSCall(ks_status_get(bv->ctx, &st));
  PetscCall(StatusCheck(&st));
  bv->gs_passes += st.total_passes;
  if (norm) *norm = safe ? 0.0 : st.nrm;
  if (lindep) *lindep = st.lindep ? PETSC_TRUE : PETSC_FALSE;
  return PETSC_SUCCESS;
}

/* Arnoldi/Lanczos expansion k..m (bvkrylov.c:88-97) with everything stream
   ordered: SpMV, the CGS passes, the DGKS decisions, lindep and the scaling
   stay on the device; the host waits once at the end for the status and the
   new coefficient columns. */
/* a grouped expansion leaves the steps to the caller that owns the group
   (the serving loop): FP64 basis and matrix, at most 128 columns, and no step
   that reaches the full-basis case */
static PetscBool GroupEligible(BV V, Mat A, PetscInt m)
{
  Mat_HIPCSR *h;
  if (!V->grouped || V->prec != SLEPC_AMD_F64 || strcmp(A->type, MATHIPCSR)) return PETSC_FALSE;
  h = (Mat_HIPCSR *)A->data;
  if (!h || !h->csr || h->vprec != SLEPC_AMD_F64) return PETSC_FALSE;
  return (V->nc + m <= 128 && m <= V->N - 1) ? PETSC_TRUE : PETSC_FALSE;
}

/* the end of an expansion: its status and coefficient columns k+1..m go to
   pinned host memory in stream order, so BVMatKrylovEnd waits once for the
   stream instead of making two copy round trips through a busy device */
PetscErrorCode BVHIPKSPostExpansion(BV V)
{
  BV_HIPKS *h = (BV_HIPKS *)V->data;
  const PetscInt k = V->kry_k, m = V->kry_m;
  KSCall(ks_status_post(V->ctx));
  KSCall(ks_memcpy_dtoh_async(V->ctx, V->buffer + (size_t)(k + 1) * h->ncols, h->d_buf + (size_t)(k + 1) * h->ncols,
                              sizeof(double) * (size_t)h->ncols * (size_t)(m - k)));
  V->kry_posted = PETSC_TRUE;
  return PETSC_SUCCESS;
}

/* one Arnoldi step: column j+1 = A column j, then its CGS+DGKS and scaling */
static PetscErrorCode ArnoldiStep(BV V, Mat A, PetscInt j)
{
  Vec vj = NULL, vj1 = NULL;
  PetscErrorCode e;
  PetscCall(BVGetColumn(V, j, &vj));
  PetscCall(BVGetColumn(V, j + 1, &vj1));
  e = MatMult(A, vj, vj1);
  PetscCall(BVRestoreColumn(V, j, &vj));
  PetscCall(BVRestoreColumn(V, j + 1, &vj1));
  if (e) return e;
  if (j == V->N - 1) return BVOrthonormalizeColumn_HIPKS(V, j + 1, PETSC_FALSE, PETSC_TRUE, PETSC_TRUE, NULL, NULL);
  return BVOrthonormalizeColumn_HIPKS(V, j + 1, PETSC_TRUE, PETSC_FALSE, PETSC_TRUE, NULL, NULL);
}

/* the same step replayed from a graph captured the first time column j is
   expanded (with this matrix and these orthogonalization settings); the
   step's launches depend on j only, and the status gates make the replay
   follow each restart's own DGKS decisions */
static PetscErrorCode ArnoldiStepGraph(BV V, Mat A, ks_csr csr, PetscInt j, PetscBool *done)
{
  BV_HIPKS *h = (BV_HIPKS *)V->data;
  ks_graph g = NULL;
  PetscErrorCode e;
  int ce;
  *done = PETSC_FALSE;
  if (h->step && (h->step_csr != csr || h->step_refine != (int)V->orthog_ref || h->step_eta != V->orthog_eta))
    StepGraphsClear(h, V->m);
  if (!h->step) {
    h->step = (ks_graph *)calloc((size_t)V->m, sizeof(ks_graph));
    if (!h->step) return PETSC_SUCCESS;
    h->step_csr = csr;
    h->step_refine = (int)V->orthog_ref;
    h->step_eta = V->orthog_eta;
  }
  if (!h->step[j]) {
    if (ks_graph_begin(V->ctx)) { h->step_failed = PETSC_TRUE; return PETSC_SUCCESS; }   /* events on: plain launches */
    e = ArnoldiStep(V, A, j);
    ce = ks_graph_end(V->ctx, &g);
    if (e || ce || !g) {   /* nothing of the step ran: it is redone with plain launches */
      (void)ks_graph_destroy(g);
      h->step_failed = PETSC_TRUE;
      return PETSC_SUCCESS;
    }
    h->step[j] = g;
  }
  KSCall(ks_graph_launch(V->ctx, h->step[j]));
  *done = PETSC_TRUE;
  return PETSC_SUCCESS;
}

static PetscErrorCode BVMatKrylovBegin_HIPKS(BV V, Mat A, PetscInt k, PetscInt m)
{
  PetscBool gate = !strcmp(A->type, MATHIPCSR);
  BV_HIPKS *h = (BV_HIPKS *)V->data;
  ks_csr csr = NULL;
  PetscErrorCode e = PETSC_SUCCESS;
  KSCall(ks_status_reset(V->ctx));
  V->kry_posted = PETSC_FALSE;
  V->kry_k = k;
  if (k < m && GroupEligible(V, A, m)) {
    V->kry_csr = ((Mat_HIPCSR *)A->data)->csr;
    V->kry_j = k;
    V->kry_m = m;
    V->kry_pending = PETSC_TRUE;
    return PETSC_SUCCESS;
  }
  if (gate) A->gate = PETSC_TRUE;
  /* graphs: a hipcsr matrix whose SpMV runs on this basis' stream, its FP32
     copy made beforehand (no allocation inside a capture) */
  if (V->step_graphs && !h->step_failed && gate && SlepcAMDContext() == V->ctx) {
    if (V->prec == SLEPC_AMD_F32) e = MatHIPCSRGetF32(A, &csr);
    else csr = ((Mat_HIPCSR *)A->data)->csr;
  }
  for (PetscInt j = k; j < m && !e; j++) {
    PetscBool done = PETSC_FALSE;
    if (csr && !h->step_failed && j != V->N - 1) e = ArnoldiStepGraph(V, A, csr, j, &done);
    if (!e && !done) e = ArnoldiStep(V, A, j);
  }
  A->gate = PETSC_FALSE;
  V->kry_m = m;
  if (!e && k < m) e = BVHIPKSPostExpansion(V);
  return e;
}

/* the next step of a grouped expansion: column nc+j+1 = A column nc+j, then its
   CGS+DGKS against the nc+j+1 columns before it and the scaling */
PetscErrorCode BVHIPKSGroupMember(BV V, ks_group_member *mem)
{
  BV_HIPKS *h = (BV_HIPKS *)V->data;
  PetscCheck(V->kry_pending, PETSC_ERR_ORDER, "no grouped expansion pending");
  mem->ctx = V->ctx;
  mem->A = V->kry_csr;
  mem->V = h->dV;
  mem->ld = V->ld;
  mem->n = V->n;
  mem->k = V->nc + V->kry_j + 1;
  mem->buf = h->d_buf;
  mem->ldbuf = h->ncols;
  mem->hcol = V->kry_j + 1;
  mem->col = V->kry_j + 1;
  return PETSC_SUCCESS;
}

PetscErrorCode BVHIPKSGroupAdvance(BV V, PetscBool *finished)
{
  V->kry_j++;
  *finished = V->kry_j >= V->kry_m ? PETSC_TRUE : PETSC_FALSE;
  if (*finished) V->kry_pending = PETSC_FALSE;
  return PETSC_SUCCESS;
}

static PetscErrorCode BVMatKrylovEnd_HIPKS(BV V, PetscInt k, PetscInt *m, PetscReal *beta, PetscBool *breakdown)
{
  BV_HIPKS *h = (BV_HIPKS *)V->data;
  ks_status st;
  PetscCheck(!V->kry_pending, PETSC_ERR_ORDER, "grouped expansion not fully enqueued");
  *m = V->kry_m;
  if (V->kry_posted) KSCall(ks_status_read(V->ctx, &st));
  else KSCall(ks_status_get(V->ctx, &st));
  PetscCall(StatusCheck(&st));
  V->gs_passes += st.total_passes;
  if (st.breakdown_col >= 0) {
    *m = st.breakdown_col;
    *breakdown = PETSC_TRUE;
  } else *breakdown = PETSC_FALSE;
  *beta = st.beta_last;
  /* new coefficient columns k+1..m into the host mirror (already there when posted) */
  if (!V->kry_posted || k != V->kry_k)
    KSCall(ks_memcpy_dtoh(V->ctx, V->buffer + (size_t)(k + 1) * h->ncols, h->d_buf + (size_t)(k + 1) * h->ncols,
                          sizeof(double) * (size_t)h->ncols * (size_t)(*m - k)));
  V->kry_posted = PETSC_FALSE;
  return PETSC_SUCCESS;
}

static PetscErrorCode BVDestroy_HIPKS(BV bv)
{
  BV_HIPKS *h = (BV_HIPKS *)bv->data;
  if (!h) return PETSC_SUCCESS;
  KSCall(ks_ctx_sync(bv->ctx));
  StepGraphsClear(h, bv->m);
  for (int l = 0; l < 2; l++) if (bv->cv[l]) PetscCall(VecDestroy(&bv->cv[l]));
  KSCall(ks_free(h->dV));
  KSCall(ks_free(h->d_buf));
  KSCall(ks_free(h->d_q));
  KSCall(ks_free(h->d_Q));
  KSCall(ks_free(h->d_G));
  KSCall(ks_host_free(h->h_Q));
  free(h);
  bv->data = NULL;
  KSCall(ks_host_free(bv->buffer));
  bv->buffer = NULL;
  return PETSC_SUCCESS;
}

/* storage allocation, run once sizes and type are known (BVCreate_Mat, bvmat.c:385-450) */
static PetscErrorCode BVAllocate_HIPKS(BV bv)
{
  BV_HIPKS *h;
  PetscCall(PetscMallocZero(sizeof(BV_HIPKS), &h));
  bv->data = h;
  bv->ld = ((bv->n + 63) / 64) * 64;
  h->ncols = bv->nc + bv->m;
  KSCall(ks_malloc(&h->dV, esz(bv) * (size_t)bv->ld * (size_t)h->ncols));
  KSCall(ks_memset_zero(bv->ctx, h->dV, esz(bv) * (size_t)bv->ld * (size_t)h->ncols));
  KSCall(ks_malloc((void **)&h->d_buf, sizeof(double) * (size_t)h->ncols * (size_t)bv->m));
  KSCall(ks_memset_zero(bv->ctx, h->d_buf, sizeof(double) * (size_t)h->ncols * (size_t)bv->m));
  KSCall(ks_malloc((void **)&h->d_q, sizeof(double) * (size_t)(h->ncols + 2)));
  free(bv->buffer);
  bv->buffer = NULL;
  /* pinned, so the expansion's coefficient columns can be copied into it in stream order */
  KSCall(ks_host_alloc((void **)&bv->buffer, sizeof(double) * (size_t)h->ncols * (size_t)bv->m));
  memset(bv->buffer, 0, sizeof(double) * (size_t)h->ncols * (size_t)bv->m);
  return PETSC_SUCCESS;
}

PetscErrorCode BVCreate_HIPKS(BV bv)
{
  bv->ops.getcolumn = BVGetColumn_HIPKS;
  bv->ops.restorecolumn = NULL;
  bv->ops.scale = BVScale_HIPKS;
  bv->ops.norm = BVNorm_HIPKS;
  bv->ops.copycolumn = BVCopyColumn_HIPKS;
  bv->ops.setrandomcolumn = BVSetRandomColumn_HIPKS;
  bv->ops.dotvec = BVDotVec_HIPKS;
  bv->ops.multvec = BVMultVec_HIPKS;
  bv->ops.multinplace = BVMultInPlace_HIPKS;
  bv->ops.gram = BVGram_HIPKS;
  bv->ops.getbuffer = BVGetBuffer_HIPKS;
  bv->ops.orthonormalizecolumn = BVOrthonormalizeColumn_HIPKS;
  bv->ops.matkrylovbegin = BVMatKrylovBegin_HIPKS;
  bv->ops.matkrylovend = BVMatKrylovEnd_HIPKS;
  bv->ops.destroy = BVDestroy_HIPKS;
  bv->ops.create = BVCreate_HIPKS;
  if (bv->m > 0 && bv->n > 0 && !bv->data) PetscCall(BVAllocate_HIPKS(bv));
  return PETSC_SUCCESS;
}

PetscErrorCode BVHIPKSSetPrecision(BV bv, SlepcAMDPrecision prec)
{
  PetscCheck(bv, PETSC_ERR_ARG_NULL, "null BV");
  PetscCheck(!bv->data, PETSC_ERR_ORDER, "precision must be set before the storage is allocated");
  bv->prec = prec;
  return PETSC_SUCCESS;
}

PetscErrorCode BVHIPKSSetStepGraphs(BV bv, PetscBool flg)
{
  PetscCheck(bv, PETSC_ERR_ARG_NULL, "null BV");
  bv->step_graphs = flg;
  if (!flg && bv->data) StepGraphsClear((BV_HIPKS *)bv->data, bv->m);
  return PETSC_SUCCESS;
}

PetscErrorCode BVHIPKSGetPrecision(BV bv, SlepcAMDPrecision *prec)
{
  PetscCheck(bv, PETSC_ERR_ARG_NULL, "null BV");
  *prec = bv->prec;
  return PETSC_SUCCESS;
}
