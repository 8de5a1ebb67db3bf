/*
 * bvhipks.c -- BVHIPKS, the MI355X basis-vector type (registered with
 * BVRegister("hipks", BVCreate_HIPKS)).  It plays the part of the reference's
 * BVMAT type (src/sys/classes/bv/impls/mat/bvmat.c:385-524) and its HIP
 * variant (bvhip.hip.cpp, hipBLAS calls), with the kernels of libks_hip.so.
 *
 * The type touches only members of the reference's struct _p_BV and fills only
 * members of its struct _BVOps (include/slepc/private/bvimpl.h:25-113); all of
 * its own state is behind bv->data:
 *
 *   storage   V: ld x (nc+m) column-major in HBM, ld = n rounded up to 64 (plus 2 KB when
 *             that stride is a multiple of 4 KB, BVHIPKSLeadingDimension); the
 *             coefficient buffer is the interface's bv->buffer (a device Vec
 *             with a pinned host copy); Q staging for BVMultInPlace in HBM
 *   ops       dotvec/multvec (tall-skinny kernels), scale/norm/copycolumn,
 *             multinplace (FP64 MFMA, no n x (e-s) temporary), dot (block
 *             Gram), gramschmidt (one CGS pass: the hook the reference's
 *             BVOrthogonalizeGS dispatches to, bvorthog.c:134)
 *   extension "BVGetKrylovOps_C" (on by default, -bv_hipks_krylov 0 turns it
 *             off): all CGS passes + DGKS + scaling of a column on the device,
 *             and a whole Arnoldi/Lanczos expansion enqueued at once with one
 *             host wait when its results are collected.
 *
 * The column vector type is VECHIP (FP64) or VECHIPF32 (FP32 storage, the
 * mixed-precision phase); coefficients are FP64 either way.
 */
#include "slepcimpl.h"

typedef struct {
  ks_ctx ctx;            /* stream and reduction workspace of every launch */
  SlepcAMDPrecision prec;
  void *dV;              /* ld x (nc+m) */
  double *d_q;           /* coefficient vectors of dotvec/multvec/gramschmidt (3 x (ncols+2)) */
  double *d_Q;           /* ldq x ldq staging of the dense Q of BVMultInPlace */
  double *h_Q;           /* pinned host staging of Q */
  size_t Qcap;           /* elements of d_Q / h_Q */
  double *d_G;           /* block dot result */
  size_t Gcap;
  PetscInt ncols;        /* nc + m columns allocated */
  int64_t gs_passes;     /* executed CGS passes */
  /* the device Krylov expansion (BVGetKrylovOps_C) */
  PetscBool ext;
  PetscInt kry_k, kry_m;
  Mat kry_A;             /* the operator of the expansion in flight */
  PetscBool kry_posted;  /* status and buffer copied to pinned host memory in stream order */
  int64_t resumes;       /* expansions a lean step stopped and the host resumed */
  /* grouped expansions (serving): the steps k..m-1 are enqueued by the server,
     step kry_j next, while kry_pending */
  PetscBool grouped, kry_pending;
  PetscInt kry_j;
  ks_csr kry_csr;
  /* captured Arnoldi steps: step[j] makes column j+1 from column j; valid for
     the matrix step_csr and the orthogonalization settings they were made with */
  PetscBool step_graphs;
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

static ks_prec kp(BV bv) { return ((BV_HIPKS *)bv->data)->prec == SLEPC_AMD_F64 ? KS_F64 : KS_F32; }
static size_t esz(BV bv) { return SlepcPrecSize(((BV_HIPKS *)bv->data)->prec); }
static ks_ctx kctx(BV bv) { return ((BV_HIPKS *)bv->data)->ctx; }
static void *colptr(BV bv, PetscInt j)
{
  BV_HIPKS *h = (BV_HIPKS *)bv->data;
  return (char *)h->dV + (size_t)(bv->nc + j) * (size_t)bv->ld * esz(bv);
}

/* device array of the coefficient buffer, for writing */
static PetscErrorCode BufferDevice(BV bv, double **d)
{
  Vec buf;
  PetscCall(BVGetBufferVec(bv, &buf));
  PetscCall(VecHIPGetArray(buf, (void **)d));
  return PETSC_SUCCESS;
}

static PetscErrorCode BVGetColumn_HIPKS(BV bv, PetscInt j, Vec *v)
{
  BV_HIPKS *h = (BV_HIPKS *)bv->data;
  if (!*v) PetscCall(VecCreateHIPWithArray(bv->n, h->prec, colptr(bv, j), v));
  else (*v)->d = colptr(bv, j), (*v)->n = bv->n, (*v)->prec = h->prec, (*v)->mask = PETSC_OFFLOAD_GPU;
  return PETSC_SUCCESS;
}

static PetscErrorCode BVScale_HIPKS(BV bv, PetscInt j, PetscScalar alpha)
{
  KSCall(ks_bv_scale(kctx(bv), kp(bv), bv->n, colptr(bv, j), alpha));
  return PETSC_SUCCESS;
}

static PetscErrorCode BVNorm_HIPKS(BV bv, PetscInt j, NormType type, PetscReal *val)
{
  PetscCheck(type == NORM_2 || type == NORM_FROBENIUS, PETSC_ERR_SUP, "only the 2-norm of a column is implemented");
  KSCall(ks_bv_norm(kctx(bv), kp(bv), bv->n, colptr(bv, j), val));
  return PETSC_SUCCESS;
}

static PetscErrorCode BVCopyColumn_HIPKS(BV bv, PetscInt j, PetscInt i)
{
  KSCall(ks_bv_copy(kctx(bv), kp(bv), bv->n, colptr(bv, j), colptr(bv, i)));
  return PETSC_SUCCESS;
}

/* q = V(:,l:k)^T y (columns counted after constraints; l may be -nc) */
static PetscErrorCode BVDotVec_HIPKS(BV bv, Vec y, PetscScalar *q)
{
  BV_HIPKS *h = (BV_HIPKS *)bv->data;
  PetscInt k = bv->k - bv->l;
  PetscCheck(y->prec == h->prec, PETSC_ERR_ARG_WRONG, "vector precision differs from the basis");
  if (k <= 0) return PETSC_SUCCESS;
  KSCall(ks_bv_dotvec(h->ctx, kp(bv), bv->n, colptr(bv, bv->l), bv->ld, k, y->d, h->d_q));
  KSCall(ks_memcpy_dtoh(h->ctx, q, h->d_q, sizeof(double) * (size_t)k));
  return PETSC_SUCCESS;
}

/* y = beta y + alpha V(:,l:k) q */
static PetscErrorCode BVMultVec_HIPKS(BV bv, PetscScalar alpha, PetscScalar beta, Vec y, PetscScalar *q)
{
  BV_HIPKS *h = (BV_HIPKS *)bv->data;
  PetscInt k = bv->k - bv->l;
  void *yd;
  PetscCheck(y->prec == h->prec, PETSC_ERR_ARG_WRONG, "vector precision differs from the basis");
  PetscCall(VecHIPGetArray(y, &yd));
  if (k > 0) KSCall(ks_memcpy_htod(h->ctx, h->d_q, q, sizeof(double) * (size_t)k));
  KSCall(ks_bv_multvec(h->ctx, kp(bv), bv->n, alpha, colptr(bv, bv->l), bv->ld, k > 0 ? k : 0, h->d_q, beta, yd));
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
    KSCall(ks_ctx_sync(h->ctx));
    KSCall(ks_free(h->d_Q));
    KSCall(ks_host_free(h->h_Q));
    h->Qcap = need;
    KSCall(ks_malloc((void **)&h->d_Q, sizeof(double) * need));
    KSCall(ks_host_alloc((void **)&h->h_Q, sizeof(double) * need));
  }
  /* the pinned staging may still feed the previous copy: wait for the stream */
  KSCall(ks_ctx_sync(h->ctx));
  memcpy(h->h_Q, qd->v, sizeof(double) * need);
  KSCall(ks_memcpy_htod_async(h->ctx, h->d_Q, h->h_Q, sizeof(double) * need));
  KSCall(ks_bv_multinplace(h->ctx, kp(bv), bv->n, colptr(bv, bv->l), bv->ld, K,
                           h->d_Q + (size_t)bv->l * ldq + bv->l, ldq, s - bv->l, e - bv->l));
  return PETSC_SUCCESS;
}

/* BVDot: M(ly:ky, lx:kx) = Y(:, ly:ky)^T X(:, lx:kx) by the block Gram kernel,
   through a device array, copied into the host matrix */
static PetscErrorCode BVDot_HIPKS(BV X, BV Y, Mat M)
{
  BV_HIPKS *h = (BV_HIPKS *)X->data;
  Mat_SeqDense *md = (Mat_SeqDense *)M->data;
  const PetscInt p = X->k - X->l, q = Y->k - Y->l;
  size_t need = (size_t)p * (size_t)q;
  double *tmp;
  PetscCheck(!strcmp(BVTypeName(Y), BVTypeName(X)) && ((BV_HIPKS *)Y->data)->prec == h->prec && Y->n == X->n,
             PETSC_ERR_ARG_INCOMP, "block dot of bases of different type, precision or length");
  if (p <= 0 || q <= 0) return PETSC_SUCCESS;
  if (need > h->Gcap) {
    KSCall(ks_ctx_sync(h->ctx));
    KSCall(ks_free(h->d_G));
    h->d_G = NULL;
    KSCall(ks_malloc((void **)&h->d_G, sizeof(double) * need));
    h->Gcap = need;
  }
  /* Y's columns live in Y's storage; both bases are on X's stream */
  KSCall(ks_bv_gram(h->ctx, kp(X), X->n, colptr(X, X->l), X->ld, p, colptr(Y, Y->l), Y->ld, q, h->d_G, q));
  PetscCall(PetscMallocZero(sizeof(double) * need, &tmp));
  {
    int e = ks_memcpy_dtoh(h->ctx, tmp, h->d_G, sizeof(double) * need);
    if (e) { free(tmp); KSCall(e); }
  }
  for (PetscInt j = 0; j < p; j++)
    memcpy(md->v + (size_t)(X->l + j) * md->lda + Y->l, tmp + (size_t)j * q, sizeof(double) * (size_t)q);
  free(tmp);
  return PETSC_SUCCESS;
}

/* The gramschmidt op: one CGS pass (BVOrthogonalizeCGS1, bvorthog.c:91-132) of
   column j (v == NULL) or of v against the nc+j columns before it.  The dots
   include the vector's own (|v|^2 in the same sweep), the update runs on the
   device, and |v'| is estimated from |v| as the reference does, with the
   explicit norm when the estimate is not positive.  For a column the
   coefficients go to the buffer on the device (c to column 0, h += c in
   column j); for v to the host arrays c and h. */
static PetscErrorCode BVGramSchmidt_HIPKS(BV bv, PetscInt j, Vec v, PetscBool *which, PetscScalar *h, PetscScalar *c,
                                          PetscReal *onorm, PetscReal *norm)
{
  BV_HIPKS *hk = (BV_HIPKS *)bv->data;
  const PetscInt k = bv->nc + j, ldb = bv->nc + bv->m;
  PetscScalar *cc = NULL;
  double *dc = hk->d_q, *dbuf = NULL;
  PetscReal beta, sum = 0.0;
  (void)which;
  PetscCheck(k < KS_MAX_COLS, PETSC_ERR_SUP, "at most %d columns are supported", KS_MAX_COLS);
  PetscCheck(!v || v->prec == hk->prec, PETSC_ERR_ARG_WRONG, "vector precision differs from the basis");
  bv->k = j;
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)(k + 1), &cc));
  /* c = V(:, 0:k)^T w and w^T w */
  if (!v) {
    KSCall(ks_bv_dotvec(hk->ctx, kp(bv), bv->n, colptr(bv, -bv->nc), bv->ld, k + 1, colptr(bv, j), dc));
  } else {
    if (k) KSCall(ks_bv_dotvec(hk->ctx, kp(bv), bv->n, colptr(bv, -bv->nc), bv->ld, k, v->d, dc));
    KSCall(ks_bv_dotvec(hk->ctx, kp(bv), bv->n, v->d, bv->ld, 1, v->d, dc + k));
  }
  {
    int e = ks_memcpy_dtoh(hk->ctx, cc, dc, sizeof(double) * (size_t)(k + 1));
    if (e) { free(cc); KSCall(e); }
  }
  PetscCheck(cc[k] >= -bv->deftol, PETSC_ERR_USER_INPUT, "Invalid inner product (negative self dot product)");
  beta = cc[k] < 0.0 ? 0.0 : sqrt(cc[k]);   /* BV_SafeSqrt (bvimpl.h:121-141), definite */
  /* q = v - V c */
  if (k) {
    void *w = colptr(bv, j);
    int e;
    if (v) (void)VecHIPGetArray(v, &w);
    e = ks_bv_multvec(hk->ctx, kp(bv), bv->n, -1.0, colptr(bv, -bv->nc), bv->ld, k, dc, 1.0, w);
    if (e) { free(cc); KSCall(e); }
  }
  if (onorm) *onorm = beta;
  if (norm) {   /* estimate |v'| from |v| */
    for (PetscInt i = 0; i < k; i++) sum += cc[i] * cc[i];
    *norm = beta * beta - sum;
    if (*norm <= 0.0) {
      PetscErrorCode e = v ? VecNorm(v, NORM_2, norm) : BVNorm_HIPKS(bv, j, NORM_2, norm);
      if (e) { free(cc); return e; }
    } else *norm = sqrt(*norm);
  }
  /* BV_AddCoefficients */
  if (!v) {
    PetscErrorCode e = BufferDevice(bv, &dbuf);
    if (!e && k) {
      int ke = ks_memcpy_dtod(hk->ctx, dbuf, dc, sizeof(double) * (size_t)k);
      if (!ke) ke = ks_bv_axpy(hk->ctx, KS_F64, k, 1.0, dc, dbuf + (size_t)j * ldb);
      if (ke) { free(cc); KSCall(ke); }
    }
    if (e) { free(cc); return e; }
  } else {
    for (PetscInt i = 0; i < k; i++) {
      if (c) c[i] = cc[i];
      if (h) h[i] += cc[i];
    }
  }
  free(cc);
  hk->gs_passes++;
  return PETSC_SUCCESS;
}

static PetscErrorCode StatusCheck(const ks_status *st)
{
  PetscCheck(st->error != KS_ERR_INNER, PETSC_ERR_USER_INPUT, "Invalid inner product (negative self dot product)");
  PetscCheck(!st->error, PETSC_ERR_GPU, "device status error %d", st->error);
  return PETSC_SUCCESS;
}

/* ---------------------------------------------------------------- the Krylov extension */
/* Whole-column orthonormalization on the device: BVOrthogonalizeGS with
   BVOrthogonalizeCGS1 passes (bvorthog.c:91-217), then the 1/norm scaling of
   BVOrthonormalizeColumn (bvorthog.c:416-422) when `scale`.  `async` leaves the
   outputs on the device (the status). */
static PetscErrorCode OrthonormalizeColumn(BV bv, PetscInt j, PetscBool scale, PetscBool safe, PetscBool async,
                                           PetscReal *norm, PetscBool *lindep)
{
  BV_HIPKS *h = (BV_HIPKS *)bv->data;
  double *dbuf;
  ks_status st;
  PetscCheck(bv->nc + j < KS_MAX_COLS, PETSC_ERR_SUP, "at most %d columns are supported", KS_MAX_COLS);
  PetscCall(BufferDevice(bv, &dbuf));
  if (async) {   /* inside an expansion: the lean sequence, which may stop it (BVMatKrylovEnd resumes) */
    KSCall(ks_bv_orthonormalize_column_lean(h->ctx, kp(bv), bv->n, h->dV, bv->ld, bv->nc + j, dbuf, h->ncols, j, j,
                                            (ks_refine)bv->orthog_ref, bv->orthog_eta, scale ? 1 : 0, safe ? 1 : 0));
    return PETSC_SUCCESS;
  }
  KSCall(ks_status_reset(h->ctx));
  KSCall(ks_bv_orthonormalize_column(h->ctx, kp(bv), bv->n, h->dV, bv->ld, bv->nc + j, dbuf, h->ncols, j, j,
                                     (ks_refine)bv->orthog_ref, bv->orthog_eta, scale ? 1 : 0, safe ? 1 : 0));
  KSCall(ks_status_get(h->ctx, &st));
  PetscCall(StatusCheck(&st));
  h->gs_passes += st.total_passes;
  if (norm) *norm = safe ? 0.0 : st.nrm;
  if (lindep) *lindep = st.lindep ? PETSC_TRUE : PETSC_FALSE;
  return PETSC_SUCCESS;
}

static PetscErrorCode BVOrthonormalizeColumn_HIPKS(BV bv, PetscInt j, PetscBool scale, PetscBool safe,
                                                   PetscReal *norm, PetscBool *lindep)
{
  return OrthonormalizeColumn(bv, j, scale, safe, PETSC_FALSE, norm, lindep);
}

/* a grouped expansion leaves the steps to the caller that owns the group
   (the serving loop): FP64 basis and matrix, at most 128 columns, and no step
   that reaches the full-basis case */
static PetscBool GroupEligible(BV V, Mat A, PetscInt m)
{
  BV_HIPKS *hk = (BV_HIPKS *)V->data;
  Mat_HIPCSR *h;
  if (!hk->grouped || hk->prec != SLEPC_AMD_F64 || !MatTypeIs(A, MATHIPCSR)) return PETSC_FALSE;
  h = (Mat_HIPCSR *)A->data;
  if (!h || !h->csr || h->vprec != SLEPC_AMD_F64) return PETSC_FALSE;
  return (V->nc + m <= 128 && m <= V->N - 1) ? PETSC_TRUE : PETSC_FALSE;
}

/* the end of an expansion: its status and the coefficient buffer go to pinned
   host memory in stream order, so BVMatKrylovEnd waits once for the stream
   instead of making two copy round trips through a busy device */
PetscErrorCode BVHIPKSPostExpansion(BV V)
{
  BV_HIPKS *h = (BV_HIPKS *)V->data;
  Vec buf;
  PetscScalar *hbuf;
  double *dbuf;
  PetscCall(BVGetBufferVec(V, &buf));
  PetscCall(VecHIPGetHostCopy(buf, &hbuf));
  PetscCall(BufferDevice(V, &dbuf));
  KSCall(ks_status_post(h->ctx));
  KSCall(ks_memcpy_dtoh_async(h->ctx, hbuf, dbuf, sizeof(double) * (size_t)h->ncols * (size_t)V->m));
  h->kry_posted = PETSC_TRUE;
  return PETSC_SUCCESS;
}

static PetscErrorCode SetGate(Mat A, PetscBool on)
{
  if (MatTypeIs(A, MATHIPCSR) && A->data) ((Mat_HIPCSR *)A->data)->gate = on;
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
  if (j == V->N - 1) return OrthonormalizeColumn(V, j + 1, PETSC_FALSE, PETSC_TRUE, PETSC_TRUE, NULL, NULL);
  return OrthonormalizeColumn(V, j + 1, PETSC_TRUE, PETSC_FALSE, PETSC_TRUE, NULL, NULL);
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
    if (ks_graph_begin(h->ctx)) { h->step_failed = PETSC_TRUE; return PETSC_SUCCESS; }   /* events on: plain launches */
    e = ArnoldiStep(V, A, j);
    ce = ks_graph_end(h->ctx, &g);
    if (e || ce || !g) {   /* nothing of the step ran: it is redone with plain launches */
      (void)ks_graph_destroy(g);
      h->step_failed = PETSC_TRUE;
      return PETSC_SUCCESS;
    }
    h->step[j] = g;
  }
  KSCall(ks_graph_launch(h->ctx, h->step[j]));
  *done = PETSC_TRUE;
  return PETSC_SUCCESS;
}

/* Arnoldi/Lanczos expansion k..m (bvkrylov.c:88-97) with everything stream
   ordered: SpMV, the CGS passes, the DGKS decisions, lindep and the scaling
   stay on the device; the host waits once at the end for the status and the
   new coefficient columns. */
/* steps j0..m-1 of the expansion, then the posted results */
static PetscErrorCode KrylovSteps(BV V, Mat A, PetscInt j0, PetscInt m)
{
  BV_HIPKS *h = (BV_HIPKS *)V->data;
  PetscBool hipcsr = MatTypeIs(A, MATHIPCSR);
  ks_csr csr = NULL;
  PetscErrorCode e = PETSC_SUCCESS;
  if (hipcsr) PetscCall(SetGate(A, PETSC_TRUE));
  /* graphs: a hipcsr matrix whose SpMV runs on this basis' stream, its FP32
     copy made beforehand (no allocation inside a capture) */
  if (h->step_graphs && !h->step_failed && hipcsr && SlepcAMDContext() == h->ctx) {
    if (h->prec == SLEPC_AMD_F32) e = MatHIPCSRGetF32(A, &csr);
    else csr = ((Mat_HIPCSR *)A->data)->csr;
  }
  for (PetscInt j = j0; j < m && !e; j++) {
    PetscBool done = PETSC_FALSE;
    if (csr && !h->step_failed && j != V->N - 1) e = ArnoldiStepGraph(V, A, csr, j, &done);
    if (!e && !done) e = ArnoldiStep(V, A, j);
  }
  if (hipcsr) PetscCall(SetGate(A, PETSC_FALSE));
  if (!e && j0 < m) e = BVHIPKSPostExpansion(V);
  return e;
}

static PetscErrorCode BVMatKrylovBegin_HIPKS(BV V, Mat A, PetscInt k, PetscInt m)
{
  BV_HIPKS *h = (BV_HIPKS *)V->data;
  KSCall(ks_status_reset(h->ctx));
  h->kry_posted = PETSC_FALSE;
  h->kry_k = k;
  h->kry_m = m;
  h->kry_A = A;
  if (k < m && GroupEligible(V, A, m)) {
    h->kry_csr = ((Mat_HIPCSR *)A->data)->csr;
    h->kry_j = k;
    h->kry_pending = PETSC_TRUE;
    return PETSC_SUCCESS;
  }
  return KrylovSteps(V, A, k, m);
}

/* the next step of a grouped expansion: column nc+j+1 = A column nc+j, then its
   CGS+DGKS against the nc+j+1 columns before it and the scaling */
PetscErrorCode BVHIPKSGroupMember(BV V, ks_group_member *mem)
{
  BV_HIPKS *h = (BV_HIPKS *)V->data;
  double *dbuf;
  PetscCheck(h->kry_pending, PETSC_ERR_ORDER, "no grouped expansion pending");
  PetscCall(BufferDevice(V, &dbuf));
  mem->ctx = h->ctx;
  mem->A = h->kry_csr;
  mem->V = h->dV;
  mem->ld = V->ld;
  mem->n = V->n;
  mem->k = V->nc + h->kry_j + 1;
  mem->buf = dbuf;
  mem->ldbuf = h->ncols;
  mem->hcol = h->kry_j + 1;
  mem->col = h->kry_j + 1;
  return PETSC_SUCCESS;
}

PetscErrorCode BVHIPKSGroupAdvance(BV V, PetscBool *finished)
{
  BV_HIPKS *h = (BV_HIPKS *)V->data;
  h->kry_j++;
  *finished = h->kry_j >= h->kry_m ? PETSC_TRUE : PETSC_FALSE;
  if (*finished) h->kry_pending = PETSC_FALSE;
  return PETSC_SUCCESS;
}

static PetscErrorCode BVMatKrylovEnd_HIPKS(BV V, PetscInt k, PetscInt *m, PetscReal *beta, PetscBool *breakdown)
{
  BV_HIPKS *h = (BV_HIPKS *)V->data;
  ks_status st;
  Vec buf;
  PetscCheck(!h->kry_pending, PETSC_ERR_ORDER, "grouped expansion not fully enqueued");
  *m = h->kry_m;
  for (;;) {
    if (h->kry_posted) KSCall(ks_status_read(h->ctx, &st));
    else KSCall(ks_status_get(h->ctx, &st));
    PetscCall(StatusCheck(&st));
    if (st.halt_col < 0) break;
    /* a lean step stopped the expansion at column c: complete that column, then
       enqueue the steps after it again (the status gates kept them from running) */
    {
      const PetscInt c = st.halt_col;
      double *dbuf;
      PetscCheck(h->kry_A && c >= 1 && c <= h->kry_m, PETSC_ERR_PLIB, "expansion stopped at column %d", (int)c);
      PetscCall(BufferDevice(V, &dbuf));
      KSCall(ks_bv_orthonormalize_resume(h->ctx, kp(V), V->n, h->dV, V->ld, V->nc + c, dbuf, h->ncols, c, c,
                                         (ks_refine)V->orthog_ref, V->orthog_eta));
      h->kry_posted = PETSC_FALSE;
      h->resumes++;
      PetscCall(KrylovSteps(V, h->kry_A, c, h->kry_m));
      if (c == h->kry_m) PetscCall(BVHIPKSPostExpansion(V));
    }
  }
  h->gs_passes += st.total_passes;
  if (st.breakdown_col >= 0) {
    *m = st.breakdown_col;
    *breakdown = PETSC_TRUE;
  } else *breakdown = PETSC_FALSE;
  *beta = st.beta_last;
  /* the posted copy of the buffer is the current host copy (the stream has
     passed it); otherwise the next host read copies it */
  PetscCall(BVGetBufferVec(V, &buf));
  if (h->kry_posted && k == h->kry_k) PetscCall(VecHIPSetHostCurrent(buf));
  h->kry_posted = PETSC_FALSE;
  return PETSC_SUCCESS;
}

static const BVKrylovOps KrylovOps_HIPKS = {BVOrthonormalizeColumn_HIPKS, BVMatKrylovBegin_HIPKS,
                                            BVMatKrylovEnd_HIPKS};

static PetscErrorCode BVGetKrylovOps_HIPKS(BV bv, const BVKrylovOps **ops)
{
  BV_HIPKS *h = (BV_HIPKS *)bv->data;
  *ops = (h && h->ext) ? &KrylovOps_HIPKS : NULL;
  return PETSC_SUCCESS;
}

/* ---------------------------------------------------------------- lifecycle */
static PetscErrorCode BVSetFromOptions_HIPKS(BV bv, PetscOptionItems *items)
{
  BV_HIPKS *h = (BV_HIPKS *)bv->data;
  PetscBool flg = h->ext, set;
  (void)items;
  PetscCall(PetscOptionsGetBool("-bv_hipks_krylov", &flg, &set));
  if (set) h->ext = flg;
  return PETSC_SUCCESS;
}

static PetscErrorCode BVDestroy_HIPKS(BV bv)
{
  BV_HIPKS *h = (BV_HIPKS *)bv->data;
  if (!h) return PETSC_SUCCESS;
  KSCall(ks_ctx_sync(h->ctx));
  StepGraphsClear(h, bv->m);
  KSCall(ks_free(h->dV));
  KSCall(ks_free(h->d_q));
  KSCall(ks_free(h->d_Q));
  KSCall(ks_free(h->d_G));
  KSCall(ks_host_free(h->h_Q));
  free(h);
  bv->data = NULL;
  PetscCall(PetscObjectComposeFunction(bv, "BVGetKrylovOps_C", NULL));
  return PETSC_SUCCESS;
}

/* BVCreate_HIPKS (cf. BVCreate_Mat, bvmat.c:385-524): storage for the sizes and
   vector type set on the BV, allocated on the current context's device */
PetscErrorCode BVCreate_HIPKS(BV bv)
{
  BV_HIPKS *h;
  PetscCheck(bv->m > 0 && bv->n > 0, PETSC_ERR_ORDER, "BVHIPKS needs the sizes first");
  PetscCall(PetscMallocZero(sizeof(BV_HIPKS), &h));
  bv->data = h;
  h->ctx = SlepcAMDContext();
  h->prec = BVPrecision(bv);
  h->ext = PETSC_TRUE;
  bv->hip = PETSC_TRUE;
  bv->ld = BVHIPKSLeadingDimension(bv->n, esz(bv));
  h->ncols = bv->nc + bv->m;
  KSCall(ks_malloc(&h->dV, esz(bv) * (size_t)bv->ld * (size_t)h->ncols));
  KSCall(ks_memset_zero(h->ctx, h->dV, esz(bv) * (size_t)bv->ld * (size_t)h->ncols));
  KSCall(ks_malloc((void **)&h->d_q, sizeof(double) * 3 * (size_t)(h->ncols + 2)));
  bv->ops->getcolumn = BVGetColumn_HIPKS;
  bv->ops->scale = BVScale_HIPKS;
  bv->ops->norm = BVNorm_HIPKS;
  bv->ops->copycolumn = BVCopyColumn_HIPKS;
  bv->ops->dotvec = BVDotVec_HIPKS;
  bv->ops->multvec = BVMultVec_HIPKS;
  bv->ops->multinplace = BVMultInPlace_HIPKS;
  bv->ops->dot = BVDot_HIPKS;
  bv->ops->gramschmidt = BVGramSchmidt_HIPKS;
  bv->ops->setfromoptions = BVSetFromOptions_HIPKS;
  bv->ops->destroy = BVDestroy_HIPKS;
  bv->ops->create = BVCreate_HIPKS;
  PetscCall(PetscObjectComposeFunction(bv, "BVGetKrylovOps_C", BVGetKrylovOps_HIPKS));
  /* the options database applies to every BVHIPKS (PETSC_OPTIONS) */
  PetscCall(BVSetFromOptions_HIPKS(bv, NULL));
  return PETSC_SUCCESS;
}

/* ---------------------------------------------------------------- type-specific functions */
static PetscErrorCode CheckHIPKS(BV bv)
{
  PetscCheck(bv, PETSC_ERR_ARG_NULL, "null BV");
  PetscCheck(!strcmp(BVTypeName(bv), BVHIPKS), PETSC_ERR_ARG_WRONG, "BV is not of type %s", BVHIPKS);
  PetscCheck(bv->data, PETSC_ERR_ORDER, "BV storage not allocated yet (set sizes first)");
  return PETSC_SUCCESS;
}

/* the FP64 storage of a BVHIPKS (column c of the storage, constraints first,
   at d + c ld) and its context, for the interface's device TSQR */
PetscErrorCode BVHIPKSGetDeviceArray_Private(BV bv, double **d, PetscInt *ld, ks_ctx *ctx)
{
  PetscCall(CheckHIPKS(bv));
  BV_HIPKS *h = (BV_HIPKS *)bv->data;
  PetscCheck(h->prec == SLEPC_AMD_F64, PETSC_ERR_SUP, "TSQR runs on FP64 bases");
  *d = (double *)h->dV;
  *ld = bv->ld;
  *ctx = h->ctx;
  return PETSC_SUCCESS;
}

PetscErrorCode BVHIPKSSetPrecision(BV bv, SlepcAMDPrecision prec)
{
  PetscCheck(bv, PETSC_ERR_ARG_NULL, "null BV");
  return BVSetVecType(bv, prec == SLEPC_AMD_F32 ? VECHIPF32 : VECHIP);
}

PetscErrorCode BVHIPKSGetPrecision(BV bv, SlepcAMDPrecision *prec)
{
  PetscCheck(bv, PETSC_ERR_ARG_NULL, "null BV");
  *prec = BVPrecision(bv);
  return PETSC_SUCCESS;
}

PetscErrorCode BVHIPKSSetStepGraphs(BV bv, PetscBool flg)
{
  BV_HIPKS *h;
  PetscCall(CheckHIPKS(bv));
  h = (BV_HIPKS *)bv->data;
  h->step_graphs = flg;
  if (!flg) StepGraphsClear(h, bv->m);
  return PETSC_SUCCESS;
}

PetscErrorCode BVHIPKSGetStepGraphs(BV bv, PetscBool *flg)
{
  PetscCall(CheckHIPKS(bv));
  *flg = ((BV_HIPKS *)bv->data)->step_graphs;
  return PETSC_SUCCESS;
}

PetscErrorCode BVHIPKSSetKrylovExtension(BV bv, PetscBool flg)
{
  PetscCall(CheckHIPKS(bv));
  ((BV_HIPKS *)bv->data)->ext = flg;
  return PETSC_SUCCESS;
}

PetscErrorCode BVHIPKSGetKrylovExtension(BV bv, PetscBool *flg)
{
  PetscCall(CheckHIPKS(bv));
  *flg = ((BV_HIPKS *)bv->data)->ext;
  return PETSC_SUCCESS;
}

PetscErrorCode BVHIPKSGetContext(BV bv, ks_ctx *ctx)
{
  PetscCall(CheckHIPKS(bv));
  *ctx = ((BV_HIPKS *)bv->data)->ctx;
  return PETSC_SUCCESS;
}

PetscErrorCode BVHIPKSGetPasses(BV bv, int64_t *passes)
{
  *passes = (bv && bv->data && !strcmp(BVTypeName(bv), BVHIPKS)) ? ((BV_HIPKS *)bv->data)->gs_passes : 0;
  return PETSC_SUCCESS;
}

PetscErrorCode BVHIPKSResetPasses(BV bv)
{
  if (bv && bv->data && !strcmp(BVTypeName(bv), BVHIPKS)) ((BV_HIPKS *)bv->data)->gs_passes = 0;
  return PETSC_SUCCESS;
}

PetscErrorCode BVHIPKSSetGrouped(BV bv, PetscBool grouped)
{
  PetscCall(CheckHIPKS(bv));
  ((BV_HIPKS *)bv->data)->grouped = grouped;
  return PETSC_SUCCESS;
}

PetscErrorCode BVHIPKSGroupPending(BV bv, PetscBool *pending)
{
  *pending = (bv && bv->data && !strcmp(BVTypeName(bv), BVHIPKS)) ? ((BV_HIPKS *)bv->data)->kry_pending
                                                                     : PETSC_FALSE;
  return PETSC_SUCCESS;
}
