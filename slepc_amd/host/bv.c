/*
 * bv.c -- BV interface (src/sys/classes/bv/interface): registry and type
 * selection (bvbasic.c:35-64, bvfunc.c:646-652), sizes, active columns, column
 * views, options, the coefficient buffer and its helpers
 * (include/slepc/private/bvimpl.h:284-415, bvbasic.c:755-786), Gram-Schmidt
 * orthogonalization of one column (bvorthog.c:91-427), block orthogonalization
 * (bvorthog.c:489-767), Krylov factorizations (bvkrylov.c:56-226) and the
 * products (bvops.c, bvglobal.c).  The numerical work is done by the type's ops
 * table (struct _BVOps, the reference's members only); a type may also offer a
 * device-resident Krylov expansion through the composed "BVGetKrylovOps_C".
 */
#include "slepcimpl.h"

/* LAPACK, Fortran ABI, 32-bit integers */
extern void dpotrf_(const char *, const int *, double *, const int *, int *);
extern void dtrtri_(const char *, const char *, const int *, double *, const int *, int *);
extern void dsyev_(const char *, const char *, const int *, double *, const int *, double *, double *, const int *,
                   int *);

static SlepcFunctionList BVList;

PetscErrorCode BVRegister(const char name[], PetscErrorCode (*create)(BV))
{
  return SlepcFunctionListAdd(&BVList, name, (void *)create);
}

/* BVCreate defaults (bvfunc.c:158-224): CGS, refine if needed, eta = 0.7071 */
PetscErrorCode BVCreate(BV *out)
{
  BV bv;
  PetscCall(SlepcCheckInitialized());
  PetscCall(PetscMallocZero(sizeof(struct _p_BV), &bv));
  PetscCall(PetscHeaderInit(&bv->hdr, "BV"));
  bv->vtype = VECHIP;
  bv->orthog_type = BV_ORTHOG_CGS;
  bv->orthog_ref = BV_ORTHOG_REFINE_IFNEEDED;
  bv->orthog_eta = 0.7071;
  bv->orthog_block = BV_ORTHOG_BLOCK_GS;
  bv->vmm = BV_MATMULT_MAT;
  bv->deftol = 10 * PETSC_MACHINE_EPSILON;
  bv->ci[0] = bv->ci[1] = BV_NOCOL;
  bv->lsplit = 0;
  PetscCall(PetscRandomCreate(&bv->rand));
  *out = bv;
  return PETSC_SUCCESS;
}

SlepcAMDPrecision BVPrecision(BV bv)
{
  return (bv->vtype && !strcmp(bv->vtype, VECHIPF32)) ? SLEPC_AMD_F32 : SLEPC_AMD_F64;
}

static PetscErrorCode BVCreateData(BV bv)
{
  if (bv->ops->create && bv->m > 0 && bv->n > 0 && !bv->data) PetscCall(bv->ops->create(bv));
  return PETSC_SUCCESS;
}

/* the type's storage, the buffer and the workspace go; type, sizes and settings stay */
static PetscErrorCode BVReset_Private(BV bv)
{
  PetscCheck(bv->ci[0] == BV_NOCOL && bv->ci[1] == BV_NOCOL, PETSC_ERR_ARG_WRONGSTATE,
             "columns of the BV are still fetched");
  if (bv->data && bv->ops->destroy) PetscCall(bv->ops->destroy(bv));
  bv->data = NULL;
  for (int l = 0; l < 2; l++) PetscCall(VecDestroy(&bv->cv[l]));
  PetscCall(VecDestroy(&bv->buffer));
  PetscFree(bv->h);
  PetscFree(bv->c);
  PetscFree(bv->work);
  bv->lwork = 0;
  return PETSC_SUCCESS;
}

PetscErrorCode BVSetSizes(BV bv, PetscInt n, PetscInt N, PetscInt m)
{
  PetscCheck(bv, PETSC_ERR_ARG_NULL, "null BV");
  PetscCheck(!bv->data, PETSC_ERR_ORDER, "sizes cannot be changed after storage is allocated");
  PetscCheck(m > 0, PETSC_ERR_ARG_OUTOFRANGE, "number of columns %d must be positive", m);
  bv->n = n >= 0 ? n : N;
  bv->N = N >= 0 ? N : n;
  bv->m = m;
  bv->k = m;
  bv->l = 0;
  PetscCall(BVCreateData(bv));
  return PETSC_SUCCESS;
}

/* new sizes (and no constraints) for a BV that may hold storage: the storage is
   rebuilt, its contents are not kept */
PetscErrorCode BVResizeDiscard_Private(BV bv, PetscInt n, PetscInt m)
{
  PetscCall(BVReset_Private(bv));
  bv->nc = 0;
  return BVSetSizes(bv, n, n, m);
}

PetscErrorCode BVSetSizesFromVec(BV bv, Vec t, PetscInt m)
{
  PetscCheck(t, PETSC_ERR_ARG_NULL, "null template vector");
  return BVSetSizes(bv, t->n, t->n, m);
}

PetscErrorCode BVGetSizes(BV bv, PetscInt *n, PetscInt *N, PetscInt *m)
{
  PetscCheck(bv, PETSC_ERR_ARG_NULL, "null BV");
  if (n) *n = bv->n;
  if (N) *N = bv->N;
  if (m) *m = bv->m;
  return PETSC_SUCCESS;
}

/* BVSetType (bvbasic.c:35-64): creation deferred until sizes are set */
PetscErrorCode BVSetType(BV bv, BVType type)
{
  PetscErrorCode (*create)(BV);
  PetscCheck(bv && type, PETSC_ERR_ARG_NULL, "null argument");
  if (!strcmp(BVTypeName(bv), type)) return PETSC_SUCCESS;
  create = (PetscErrorCode(*)(BV))SlepcFunctionListFind(&BVList, type);
  PetscCheck(create, PETSC_ERR_ARG_WRONG, "unknown BV type %s", type);
  PetscCall(BVReset_Private(bv));
  memset(bv->ops, 0, sizeof(struct _BVOps));
  bv->hdr.nfn = 0;
  snprintf(bv->hdr.type_name, sizeof(bv->hdr.type_name), "%s", type);
  bv->ops->create = create;
  PetscCall(BVCreateData(bv));
  return PETSC_SUCCESS;
}

PetscErrorCode BVGetType(BV bv, BVType *type)
{
  PetscCheck(bv, PETSC_ERR_ARG_NULL, "null BV");
  *type = BVTypeName(bv);
  return PETSC_SUCCESS;
}

/* BVSetVecType (bvbasic.c): the column vector type, before the storage exists */
PetscErrorCode BVSetVecType(BV bv, VecType vtype)
{
  PetscCheck(bv && vtype, PETSC_ERR_ARG_NULL, "null argument");
  PetscCheck(!bv->data, PETSC_ERR_ORDER, "the vector type must be set before the storage is allocated");
  if (!strcmp(vtype, VECHIP)) bv->vtype = VECHIP;
  else if (!strcmp(vtype, VECHIPF32)) bv->vtype = VECHIPF32;
  else SETERRQ(PETSC_ERR_SUP, "vector type %s is not available (%s or %s)", vtype, VECHIP, VECHIPF32);
  return PETSC_SUCCESS;
}

PetscErrorCode BVGetVecType(BV bv, VecType *vtype)
{
  PetscCheck(bv && vtype, PETSC_ERR_ARG_NULL, "null argument");
  *vtype = bv->vtype;
  return PETSC_SUCCESS;
}

/* BVSetOrthogonalization (bvbasic.c): only classical Gram-Schmidt runs on the device */
PetscErrorCode BVSetOrthogonalization(BV bv, BVOrthogType type, BVOrthogRefineType refine, PetscReal eta,
                                      BVOrthogBlockType block)
{
  PetscCheck(bv, PETSC_ERR_ARG_NULL, "null BV");
  PetscCheck(block >= BV_ORTHOG_BLOCK_GS && block <= BV_ORTHOG_BLOCK_SVQB, PETSC_ERR_ARG_OUTOFRANGE,
             "invalid block type %d", (int)block);
  PetscCheck(type == BV_ORTHOG_CGS, PETSC_ERR_SUP, "only classical Gram-Schmidt is implemented on the device");
  PetscCheck(refine >= BV_ORTHOG_REFINE_IFNEEDED && refine <= BV_ORTHOG_REFINE_ALWAYS, PETSC_ERR_ARG_OUTOFRANGE,
             "invalid refinement type %d", (int)refine);
  bv->orthog_block = block;
  bv->orthog_type = type;
  bv->orthog_ref = refine;
  if (eta != PETSC_DEFAULT && eta != PETSC_DETERMINE) {
    PetscCheck(eta > 0.0 && eta <= 1.0, PETSC_ERR_ARG_OUTOFRANGE, "invalid eta %g", eta);
    bv->orthog_eta = eta;
  }
  return PETSC_SUCCESS;
}

PetscErrorCode BVGetOrthogonalization(BV bv, BVOrthogType *type, BVOrthogRefineType *refine, PetscReal *eta,
                                      BVOrthogBlockType *block)
{
  PetscCheck(bv, PETSC_ERR_ARG_NULL, "null BV");
  if (type) *type = bv->orthog_type;
  if (refine) *refine = bv->orthog_ref;
  if (eta) *eta = bv->orthog_eta;
  if (block) *block = bv->orthog_block;
  return PETSC_SUCCESS;
}

static PetscErrorCode OptionsGetEnum(const char *name, const char *const *list, int nlist, int *val, PetscBool *set)
{
  char v[64];
  PetscCall(PetscOptionsGetString(name, v, sizeof(v), set));
  if (!*set) return PETSC_SUCCESS;
  for (int i = 0; i < nlist; i++)
    if (list[i] && !strcmp(list[i], v)) { *val = i; return PETSC_SUCCESS; }
  SETERRQ(PETSC_ERR_ARG_WRONG, "unknown value %s for option %s", v, name);
}

/* BVSetFromOptions (bvbasic.c) */
PetscErrorCode BVSetFromOptions(BV bv)
{
  static const char *const otype[] = {"cgs", "mgs"};
  static const char *const oref[] = {"ifneeded", "never", "always"};
  static const char *const oblock[] = {"gs", "chol", "tsqr", "tsqrchol", "svqb"};
  char v[64];
  PetscBool set;
  int t = (int)bv->orthog_type, r = (int)bv->orthog_ref, b = (int)bv->orthog_block;
  PetscReal eta = bv->orthog_eta;
  PetscCheck(bv, PETSC_ERR_ARG_NULL, "null BV");
  PetscCall(PetscOptionsGetString("-bv_type", v, sizeof(v), &set));
  if (set) PetscCall(BVSetType(bv, v));
  else if (!BVTypeName(bv)[0]) PetscCall(BVSetType(bv, BVHIPKS));
  PetscCall(OptionsGetEnum("-bv_orthog_type", otype, 2, &t, &set));
  PetscCall(OptionsGetEnum("-bv_orthog_refine", oref, 3, &r, &set));
  PetscCall(OptionsGetEnum("-bv_orthog_block", oblock, 5, &b, &set));
  PetscCall(PetscOptionsGetString("-bv_orthog_eta", v, sizeof(v), &set));
  if (set) eta = strtod(v, NULL);
  PetscCall(BVSetOrthogonalization(bv, (BVOrthogType)t, (BVOrthogRefineType)r, eta, (BVOrthogBlockType)b));
  if (bv->ops->setfromoptions) PetscCall(bv->ops->setfromoptions(bv, NULL));
  bv->sfocalled = PETSC_TRUE;
  return PETSC_SUCCESS;
}

PetscErrorCode BVSetActiveColumns(BV bv, PetscInt l, PetscInt k)
{
  PetscCheck(bv, PETSC_ERR_ARG_NULL, "null BV");
  PetscCheck(k >= 0 && k <= bv->m, PETSC_ERR_ARG_OUTOFRANGE, "bad k %d (m=%d)", k, bv->m);
  PetscCheck(l >= 0 && l <= k, PETSC_ERR_ARG_OUTOFRANGE, "bad l %d (k=%d)", l, k);
  bv->l = l;
  bv->k = k;
  return PETSC_SUCCESS;
}

PetscErrorCode BVGetActiveColumns(BV bv, PetscInt *l, PetscInt *k)
{
  PetscCheck(bv, PETSC_ERR_ARG_NULL, "null BV");
  if (l) *l = bv->l;
  if (k) *k = bv->k;
  return PETSC_SUCCESS;
}

static PetscErrorCode BVCheckReady(BV bv)
{
  PetscCheck(bv, PETSC_ERR_ARG_NULL, "null BV");
  PetscCheck(bv->data, PETSC_ERR_ORDER, "BV needs a type and sizes first");
  return PETSC_SUCCESS;
}

/* BVGetColumn (bvbasic.c:1116-1137): at most two columns lent at a time */
PetscErrorCode BVGetColumn(BV bv, PetscInt j, Vec *v)
{
  PetscInt l;
  PetscCall(BVCheckReady(bv));
  PetscCheck(j >= -bv->nc && j < bv->m, PETSC_ERR_ARG_OUTOFRANGE, "column %d out of range", j);
  PetscCheck(j != bv->ci[0] && j != bv->ci[1], PETSC_ERR_ARG_WRONG, "column %d already fetched", j);
  PetscCheck(bv->ci[0] == BV_NOCOL || bv->ci[1] == BV_NOCOL, PETSC_ERR_SUP, "too many columns requested");
  l = (bv->ci[0] == BV_NOCOL) ? 0 : 1;
  bv->ci[l] = j;
  PetscCall(bv->ops->getcolumn(bv, j, &bv->cv[l]));
  *v = bv->cv[l];
  return PETSC_SUCCESS;
}

PetscErrorCode BVRestoreColumn(BV bv, PetscInt j, Vec *v)
{
  PetscInt l;
  PetscCall(BVCheckReady(bv));
  PetscCheck(j == bv->ci[0] || j == bv->ci[1], PETSC_ERR_ARG_WRONG, "column %d has not been fetched", j);
  l = (j == bv->ci[0]) ? 0 : 1;
  PetscCheck(v && *v == bv->cv[l], PETSC_ERR_ARG_WRONG, "vector does not match the fetched column");
  if (bv->ops->restorecolumn) PetscCall(bv->ops->restorecolumn(bv, j, &bv->cv[l]));
  bv->ci[l] = BV_NOCOL;
  *v = NULL;
  return PETSC_SUCCESS;
}

/* BVSetRandomColumn (bvops.c:360-379, BVSetRandomColumn_Private): the column's
   values from the BV's generator (the same for any process count, so rrandom
   needs no separate path) */
PetscErrorCode BVSetRandomColumn(BV bv, PetscInt j)
{
  Vec x = NULL;
  PetscErrorCode e;
  PetscCall(BVCheckReady(bv));
  PetscCheck(j >= 0 && j < bv->m, PETSC_ERR_ARG_OUTOFRANGE, "column %d out of range", j);
  PetscCall(BVGetColumn(bv, j, &x));
  e = VecSetRandom(x, bv->rand);
  PetscCall(BVRestoreColumn(bv, j, &x));
  return e;
}

PetscErrorCode BVSetRandomSeed(BV bv, uint64_t seed)
{
  PetscCheck(bv, PETSC_ERR_ARG_NULL, "null BV");
  bv->rand->seed = seed;
  return PETSC_SUCCESS;
}

PetscErrorCode BVGetRandomSeed(BV bv, uint64_t *seed)
{
  PetscCheck(bv && seed, PETSC_ERR_ARG_NULL, "null argument");
  *seed = bv->rand->seed;
  return PETSC_SUCCESS;
}

PetscErrorCode BVScaleColumn(BV bv, PetscInt j, PetscScalar alpha)
{
  PetscCall(BVCheckReady(bv));
  PetscCheck(j >= 0 && j < bv->m, PETSC_ERR_ARG_OUTOFRANGE, "column %d out of range", j);
  if (alpha == 1.0) return PETSC_SUCCESS;
  return bv->ops->scale(bv, j, alpha);
}

PetscErrorCode BVNormColumn(BV bv, PetscInt j, NormType type, PetscReal *norm)
{
  PetscCall(BVCheckReady(bv));
  PetscCheck(j >= -bv->nc && j < bv->m, PETSC_ERR_ARG_OUTOFRANGE, "column %d out of range", j);
  return bv->ops->norm(bv, j, type, norm);
}

PetscErrorCode BVCopyColumn(BV bv, PetscInt j, PetscInt i)
{
  PetscCall(BVCheckReady(bv));
  PetscCheck(j >= 0 && j < bv->m && i >= 0 && i < bv->m, PETSC_ERR_ARG_OUTOFRANGE, "columns %d,%d out of range", j, i);
  if (j == i) return PETSC_SUCCESS;
  return bv->ops->copycolumn(bv, j, i);
}

PetscErrorCode BVCopyVec(BV bv, PetscInt j, Vec w)
{
  Vec z;
  PetscCall(BVGetColumn(bv, j, &z));
  PetscErrorCode e = VecCopy(z, w);
  PetscCall(BVRestoreColumn(bv, j, &z));
  return e;
}

PetscErrorCode BVInsertVec(BV bv, PetscInt j, Vec w)
{
  Vec z;
  PetscCall(BVGetColumn(bv, j, &z));
  PetscErrorCode e = VecCopy(w, z);
  PetscCall(BVRestoreColumn(bv, j, &z));
  return e;
}

PetscErrorCode BVGetColumnValuesHost(BV bv, PetscInt j, PetscScalar *values)
{
  Vec z;
  PetscCall(BVGetColumn(bv, j, &z));
  PetscErrorCode e = VecGetValuesHost(z, values);
  PetscCall(BVRestoreColumn(bv, j, &z));
  return e;
}

PetscErrorCode BVSetColumnValuesHost(BV bv, PetscInt j, const PetscScalar *values)
{
  Vec z;
  PetscCall(BVGetColumn(bv, j, &z));
  PetscErrorCode e = VecSetValuesHost(z, values);
  PetscCall(BVRestoreColumn(bv, j, &z));
  return e;
}

/* BVInsertVecs (bvfunc.c:340-374) */
PetscErrorCode BVInsertVecs(BV bv, PetscInt s, PetscInt *m, Vec *W, PetscBool orth)
{
  PetscInt ndep = 0;
  PetscReal norm;
  PetscBool lindep;
  PetscCall(BVCheckReady(bv));
  PetscCheck(m && *m > 0 && W, PETSC_ERR_ARG_OUTOFRANGE, "bad number of vectors");
  PetscCheck(s >= 0 && s + *m <= bv->m, PETSC_ERR_ARG_OUTOFRANGE, "too many vectors for %d columns", bv->m);
  for (PetscInt i = 0; i < *m; i++) {
    PetscCheck(W[i] && W[i]->n == bv->n, PETSC_ERR_ARG_SIZ, "vector %d has the wrong size", i);
    PetscCall(BVInsertVec(bv, s + i - ndep, W[i]));
    if (orth) {
      PetscCall(BVOrthogonalizeColumn(bv, s + i - ndep, NULL, &norm, &lindep));
      if (norm == 0.0 || lindep) ndep++;
      else PetscCall(BVScaleColumn(bv, s + i - ndep, 1.0 / norm));
    }
  }
  *m -= ndep;
  return PETSC_SUCCESS;
}

/* BVInsertConstraints (bvfunc.c:411-439): the storage is rebuilt with nc
   leading constraint columns, which are inserted and orthonormalized. */
PetscErrorCode BVInsertConstraints(BV bv, PetscInt *nc, Vec *C)
{
  PetscInt msave, ncin;
  PetscCall(BVCheckReady(bv));
  PetscCheck(nc && *nc >= 0, PETSC_ERR_ARG_OUTOFRANGE, "bad number of constraints");
  PetscCheck(!bv->nc, PETSC_ERR_ARG_WRONGSTATE, "constraints already present in this BV object");
  PetscCheck(bv->ci[0] == BV_NOCOL && bv->ci[1] == BV_NOCOL, PETSC_ERR_SUP, "cannot call BVInsertConstraints after BVGetColumn");
  if (!*nc) return PETSC_SUCCESS;
  ncin = *nc;
  msave = bv->m;
  PetscCall(BVReset_Private(bv));
  bv->m = ncin + msave;
  PetscCall(BVCreateData(bv));
  bv->l = 0;
  bv->k = bv->m;
  PetscCall(BVInsertVecs(bv, 0, nc, C, PETSC_TRUE));
  /* keep the storage: the first nc columns become the constraints; the buffer
     is made again for the (nc+m) x m shape */
  PetscCall(VecDestroy(&bv->buffer));
  bv->nc = *nc;
  bv->m = msave;
  bv->l = 0;
  bv->k = msave;
  bv->ci[0] = bv->ci[1] = BV_NOCOL;
  return PETSC_SUCCESS;
}

PetscErrorCode BVGetNumConstraints(BV bv, PetscInt *nc)
{
  PetscCheck(bv, PETSC_ERR_ARG_NULL, "null BV");
  *nc = bv->nc;
  return PETSC_SUCCESS;
}

/* BVDotVec (bvglobal.c): m = V(:,l:k)^T y */
PetscErrorCode BVDotVec(BV bv, Vec y, PetscScalar m[])
{
  PetscCall(BVCheckReady(bv));
  PetscCheck(y && y->n == bv->n, PETSC_ERR_ARG_SIZ, "vector size does not match");
  return bv->ops->dotvec(bv, y, m);
}

/* BVMultVec (bvops.c): y = beta y + alpha V(:,l:k) q */
PetscErrorCode BVMultVec(BV bv, PetscScalar alpha, PetscScalar beta, Vec y, PetscScalar q[])
{
  PetscCall(BVCheckReady(bv));
  PetscCheck(y && y->n == bv->n, PETSC_ERR_ARG_SIZ, "vector size does not match");
  return bv->ops->multvec(bv, alpha, beta, y, q);
}

/* BVMultColumn (bvops.c:165-192): column j = beta * column j + alpha * V(:, l:j) q,
   with the active columns set to l..j for the product */
PetscErrorCode BVMultColumn(BV bv, PetscScalar alpha, PetscScalar beta, PetscInt j, PetscScalar q[])
{
  PetscInt ksave;
  Vec y = NULL;
  PetscErrorCode e;
  PetscCall(BVCheckReady(bv));
  PetscCheck(j >= 0 && j < bv->m, PETSC_ERR_ARG_OUTOFRANGE, "Index j=%d but BV only has %d columns", j, bv->m);
  PetscCheck(q || j <= bv->l, PETSC_ERR_ARG_NULL, "coefficients are required");
  ksave = bv->k;
  bv->k = j;
  e = BVGetColumn(bv, j, &y);
  if (!e) e = bv->ops->multvec(bv, alpha, beta, y, q);
  if (y) BVRestoreColumn(bv, j, &y);
  bv->k = ksave;
  return e;
}

/* BVNormalize (bvglobal.c:829-849) as BVNormalize_LAPACK_Private (bvlapack.c:88-130):
   active columns to unit norm, a conjugate pair (eigi[i] != 0) by the Frobenius
   norm of its two columns */
PetscErrorCode BVNormalize(BV bv, PetscScalar eigi[])
{
  PetscCall(BVCheckReady(bv));
  for (PetscInt j = bv->l; j < bv->k; j++) {
    PetscReal nr, ni = 0.0;
    PetscBool pair = eigi && eigi[j] != 0.0 && j + 1 < bv->k;
    PetscCall(BVNormColumn(bv, j, NORM_2, &nr));
    if (pair) PetscCall(BVNormColumn(bv, j + 1, NORM_2, &ni));
    nr = SlepcAbsEigenvalue(nr, ni);
    PetscCheck(nr > 0.0, PETSC_ERR_FP, "zero column %d cannot be normalized", j);
    PetscCall(BVScaleColumn(bv, j, 1.0 / nr));
    if (pair) PetscCall(BVScaleColumn(bv, ++j, 1.0 / nr));
  }
  return PETSC_SUCCESS;
}

/* BVMultInPlace (bvops.c:220-245) */
PetscErrorCode BVMultInPlace(BV bv, Mat Q, PetscInt s, PetscInt e)
{
  PetscInt m, n;
  PetscCall(BVCheckReady(bv));
  PetscCheck(Q && MatTypeIs(Q, MATSEQDENSE), PETSC_ERR_ARG_WRONG, "Q must be a sequential dense matrix");
  PetscCheck(s >= bv->l && s <= bv->m, PETSC_ERR_ARG_OUTOFRANGE, "Argument s has wrong value %d, should be between %d and %d", s, bv->l, bv->m);
  PetscCheck(e >= bv->l && e <= bv->m, PETSC_ERR_ARG_OUTOFRANGE, "Argument e has wrong value %d, should be between %d and %d", e, bv->l, bv->m);
  PetscCall(MatGetSize(Q, &m, &n));
  PetscCheck(m >= bv->k, PETSC_ERR_ARG_SIZ, "Mat argument has %d rows, should have at least %d", m, bv->k);
  PetscCheck(e <= n, PETSC_ERR_ARG_SIZ, "Mat argument only has %d columns, the requested value of e is larger: %d", n, e);
  return bv->ops->multinplace(bv, Q, s, e);
}

/* BVDot (bvglobal.c:86-120): M(ly:ky, lx:kx) = Y(:, ly:ky)^T X(:, lx:kx) */
PetscErrorCode BVDot(BV X, BV Y, Mat M)
{
  PetscInt m, n;
  PetscCall(BVCheckReady(X));
  PetscCall(BVCheckReady(Y));
  PetscCheck(M && MatTypeIs(M, MATSEQDENSE), PETSC_ERR_ARG_WRONG, "M must be a sequential dense matrix");
  PetscCall(MatGetSize(M, &m, &n));
  PetscCheck(m >= Y->k, PETSC_ERR_ARG_SIZ, "Mat argument has %d rows, should have at least %d", m, Y->k);
  PetscCheck(n >= X->k, PETSC_ERR_ARG_SIZ, "Mat argument has %d columns, should have at least %d", n, X->k);
  PetscCheck(X->n == Y->n, PETSC_ERR_ARG_INCOMP, "Mismatching local dimension X %d, Y %d", X->n, Y->n);
  PetscCheck(!strcmp(BVTypeName(X), BVTypeName(Y)) && !strcmp(X->vtype, Y->vtype), PETSC_ERR_ARG_INCOMP,
             "X and Y must have the same type and vector type");
  if (X->l == X->k || Y->l == Y->k) return PETSC_SUCCESS;
  PetscCheck(X->ops->dot, PETSC_ERR_SUP, "BV type %s has no block dot", BVTypeName(X));
  return X->ops->dot(X, Y, M);
}

/* G(0:ye-ys, 0:xe-xs) = Y(:, ys:ye)^T X(:, xs:xe) into host memory with leading
   dimension ldg, through BVDot over the given column ranges */
static PetscErrorCode BVDotRange(BV X, PetscInt xs, PetscInt xe, BV Y, PetscInt ys, PetscInt ye, PetscScalar *G,
                                 PetscInt ldg)
{
  PetscInt xl = X->l, xk = X->k, yl = Y->l, yk = Y->k;
  struct _p_BV ycopy;
  Mat M = NULL;
  PetscErrorCode e;
  if (xe <= xs || ye <= ys) return PETSC_SUCCESS;
  /* two ranges of one basis: Y is a view of the same storage with its own
     active columns (the part BVGetSplit plays in bvorthog.c:489-502) */
  if (Y == X) {
    memcpy(&ycopy, Y, sizeof(ycopy));
    Y = &ycopy;
  }
  PetscCall(MatCreateSeqDense(ye, xe, NULL, &M));
  X->l = xs; X->k = xe;
  Y->l = ys; Y->k = ye;
  e = BVDot(X, Y, M);
  X->l = xl; X->k = xk;
  if (Y != &ycopy) { Y->l = yl; Y->k = yk; }
  if (!e) {
    Mat_SeqDense *d = (Mat_SeqDense *)M->data;
    for (PetscInt j = xs; j < xe; j++)
      memcpy(G + (size_t)(j - xs) * ldg, d->v + (size_t)j * d->lda + ys, sizeof(PetscScalar) * (size_t)(ye - ys));
  }
  MatDestroy(&M);
  return e;
}

static PetscErrorCode DenseAt(Mat A, PetscScalar **a, PetscInt *lda)
{
  Mat_SeqDense *d = (Mat_SeqDense *)A->data;
  *a = d->v;
  *lda = d->lda;
  return PETSC_SUCCESS;
}

/* Block Gram-Schmidt against the leading columns (bvorthog.c:489-502):
   R12 = V(:,0:l)^T V(:,l:k), V(:,l:k) -= V(:,0:l) R12; R12 goes to rows 0..l-1
   of the columns l..k-1 of W (k x k, column-major, leading dimension ldw) */
static PetscErrorCode OrthoBlockGS(BV V, PetscScalar *W, PetscInt ldw)
{
  const PetscInt l = V->l, k = V->k;
  Mat Q;
  PetscScalar *q;
  PetscInt ldq;
  PetscErrorCode e;
  PetscCall(BVDotRange(V, l, k, V, 0, l, W + l * ldw, ldw));
  /* V(:, l:k) = V(:, 0:k) [-R12; I] */
  PetscCall(MatCreateSeqDense(k, k, NULL, &Q));
  PetscCall(DenseAt(Q, &q, &ldq));
  for (PetscInt j = l; j < k; j++) {
    for (PetscInt i = 0; i < l; i++) q[i + j * ldq] = -W[i + j * ldw];
    q[j + j * ldq] = 1.0;
  }
  V->l = 0;
  e = BVMultInPlace(V, Q, l, k);
  V->l = l;
  MatDestroy(&Q);
  return e;
}

/* V(:, l:k) = V(:, l:k) R^-1 for the upper triangular R in W(l:k, l:k)
   (BVMatTriInv_LAPACK_Private, bvlapack.c:206-246, then BVMultInPlace) */
static PetscErrorCode OrthoTriInv(BV V, PetscScalar *W, PetscInt ldw)
{
  const PetscInt l = V->l, k = V->k, n = k - l;
  int n_ = (int)n, info = 0;
  Mat S;
  PetscScalar *sv;
  PetscInt lds;
  PetscCall(MatCreateSeqDense(k, k, NULL, &S));
  PetscCall(DenseAt(S, &sv, &lds));
  for (PetscInt j = 0; j < n; j++) memcpy(sv + (l + j) * lds + l, W + (l + j) * ldw + l, sizeof(PetscScalar) * (size_t)n);
  {
    int lds_ = (int)lds;
    dtrtri_("U", "N", &n_, sv + l * lds + l, &lds_, &info);
  }
  if (info) { MatDestroy(&S); SETERRQ(PETSC_ERR_LIB, "Error in LAPACK subroutine trtri: info=%d", info); }
  {
    PetscErrorCode e = BVMultInPlace(V, S, l, k);
    MatDestroy(&S);
    PetscCall(e);
  }
  return PETSC_SUCCESS;
}

/* Cholesky QR (bvorthog.c:581-596, BVMatCholInv_LAPACK_Private bvlapack.c:136-202):
   G = V2^T V2, upper Cholesky G = R^T R (retried once on G + 50 eps I), V2 = V2 R^-1 */
static PetscErrorCode OrthoChol(BV V, PetscScalar *W, PetscInt ldw)
{
  const PetscInt l = V->l, k = V->k, n = k - l;
  int n_ = (int)n, ld_ = (int)ldw, info = 0;
  PetscScalar *g, *save;
  PetscCall(BVDotRange(V, l, k, V, l, k, W + l * ldw + l, ldw));   /* W(l:k, l:k) = V2^T V2 */
  g = W + l * ldw + l;
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)n * (size_t)n, &save));
  for (PetscInt j = 0; j < n; j++) memcpy(save + j * n, g + j * ldw, sizeof(PetscScalar) * (size_t)n);
  dpotrf_("U", &n_, g, &ld_, &info);
  if (info) {   /* retry on the diagonally perturbed matrix */
    for (PetscInt j = 0; j < n; j++) {
      memcpy(g + j * ldw, save + j * n, sizeof(PetscScalar) * (size_t)n);
      g[j + j * ldw] += 50.0 * PETSC_MACHINE_EPSILON;
    }
    dpotrf_("U", &n_, g, &ld_, &info);
  }
  free(save);
  PetscCheck(info == 0, PETSC_ERR_LIB, "Error in LAPACK subroutine potrf: info=%d", info);
  for (PetscInt j = 0; j < n; j++)   /* zero below the diagonal */
    for (PetscInt i = j + 1; i < n; i++) g[i + j * ldw] = 0.0;
  return OrthoTriInv(V, W, ldw);
}

/* TSQR and TSQRCHOL (bvorthog.c:615-663): the device TSQR of the panel
   V(:, l:k) (ks_bv_tsqr); TSQR forms Q in place, TSQRCHOL keeps the factor
   only and applies R^-1 (BVMatTriInv_LAPACK_Private, bvlapack.c:206-246) */
static PetscErrorCode OrthoTSQR(BV V, PetscScalar *W, PetscInt ldw, PetscBool formq)
{
  const PetscInt l = V->l, k = V->k;
  double *d;
  PetscInt ld;
  ks_ctx ctx;
  PetscCall(BVHIPKSGetDeviceArray_Private(V, &d, &ld, &ctx));
  KSCall(ks_bv_tsqr(ctx, (int)V->n, d + (size_t)(V->nc + l) * (size_t)ld, (int)ld, (int)(k - l), formq ? 1 : 0,
                    W + l * ldw + l, (int)ldw));
  return formq ? PETSC_SUCCESS : OrthoTriInv(V, W, ldw);
}

/* SVQB (bvorthog.c:646-663, BVMatSVQB_LAPACK_Private bvlapack.c:261-345):
   G = V2^T V2, D = diag(G)^-1/2, D G D = U L U^T, V2 = V2 D U L^-1/2, R = L^1/2 U^T D^-1 */
static PetscErrorCode OrthoSVQB(BV V, PetscScalar *W, PetscInt ldw)
{
  const PetscInt l = V->l, k = V->k, n = k - l;
  int n_ = (int)n, info = 0, lwork = -1;
  Mat S;
  PetscScalar *sv, *D, *eig, *work, *U, a = 0.0;
  PetscInt lds;
  PetscCall(BVDotRange(V, l, k, V, l, k, W + l * ldw + l, ldw));   /* W(l:k, l:k) = V2^T V2 */
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)n * (size_t)(n + 2), &U));
  D = U + (size_t)n * n;
  eig = D + n;
  for (PetscInt i = 0; i < n; i++) D[i] = 1.0 / sqrt(W[(l + i) + (l + i) * ldw]);
  for (PetscInt j = 0; j < n; j++)
    for (PetscInt i = 0; i < n; i++) U[i + j * n] = W[(l + i) + (l + j) * ldw] * D[i] * D[j];
  dsyev_("V", "L", &n_, U, &n_, eig, &a, &lwork, &info);
  lwork = (int)a;
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)(lwork > 1 ? lwork : 1), &work));
  dsyev_("V", "L", &n_, U, &n_, eig, work, &lwork, &info);
  free(work);
  if (info) { free(U); SETERRQ(PETSC_ERR_LIB, "Error in LAPACK subroutine syev: info=%d", info); }
  /* S(l:k, l:k) = D U L^-1/2;  W(l:k, l:k) = L^1/2 U^T D^-1 */
  PetscCall(MatCreateSeqDense(k, k, NULL, &S));
  PetscCall(DenseAt(S, &sv, &lds));
  for (PetscInt j = 0; j < n; j++)
    for (PetscInt i = 0; i < n; i++) {
      sv[(l + i) + (l + j) * lds] = D[i] * U[i + j * n] / sqrt(eig[j]);
      W[(l + i) + (l + j) * ldw] = sqrt(eig[i]) * U[j + i * n] / D[j];
    }
  free(U);
  {
    PetscErrorCode e = BVMultInPlace(V, S, l, k);
    MatDestroy(&S);
    PetscCall(e);
  }
  return PETSC_SUCCESS;
}

/* BVOrthogonalize (bvorthog.c:726-767) */
PetscErrorCode BVOrthogonalize(BV V, Mat R)
{
  PetscInt m, n, l, k, ldw;
  PetscScalar *W = NULL;
  PetscErrorCode e = PETSC_SUCCESS;
  PetscCall(BVCheckReady(V));
  l = V->l;
  k = V->k;
  if (R) {
    PetscCheck(MatTypeIs(R, MATSEQDENSE), PETSC_ERR_ARG_WRONG, "R must be a sequential dense matrix");
    PetscCall(MatGetSize(R, &m, &n));
    PetscCheck(m == n, PETSC_ERR_ARG_SIZ, "Mat argument is not square, it has %d rows and %d columns", m, n);
    PetscCheck(n >= k, PETSC_ERR_ARG_SIZ, "Mat size %d is smaller than the number of BV active columns %d", n, k);
  }
  if (l >= k) return PETSC_SUCCESS;
  if (V->orthog_block == BV_ORTHOG_BLOCK_GS) {   /* column by column (bvorthog.c:507-547) */
    PetscScalar *r = NULL;
    PetscInt ldr = 0;
    if (R) PetscCall(DenseAt(R, &r, &ldr));
    for (PetscInt j = l; j < k && !e; j++) {
      PetscReal norm = 0.0;
      e = BVOrthogonalizeColumn(V, j, NULL, &norm, NULL);
      if (e) break;
      if (r) {   /* every coefficient of the column against the regular columns, from column 0
                    (bvorthog.c:531-535 starts at -nc, the same column when there are no constraints) */
        PetscInt lsave = V->l;
        V->l = 0;
        e = BV_StoreCoefficients(V, j, NULL, r + j * ldr);
        V->l = lsave;
        if (e) break;
        r[j + j * ldr] = norm;
      }
      if (norm == 0.0) { e = PETSC_ERR_CONV_FAILED; fprintf(stderr, "Breakdown in BVOrthogonalize due to a linearly dependent column\n"); break; }
      e = BVScaleColumn(V, j, 1.0 / norm);
    }
    return e;
  }
  PetscCheck(V->ops->dot, PETSC_ERR_SUP, "BV type %s has no block dot", BVTypeName(V));
  /* Constraint columns, which the reference refuses here (bvorthog.c:743), are
     taken as further leading columns: the block Gram-Schmidt pass against the
     leading columns then orthogonalizes against them too.  The view is shifted
     so that storage column c is column c: nc = 0, l and k moved up by nc. */
  const PetscInt nc = V->nc, msave = V->m;
  if (nc) {
    V->nc = 0;
    V->m = msave + nc;
    V->l = l + nc;
    V->k = k + nc;
  }
  ldw = V->k;
  W = (PetscScalar *)calloc((size_t)ldw * (size_t)ldw, sizeof(PetscScalar));
  if (!W) e = PETSC_ERR_MEM;
  if (!e && V->l > 0) e = OrthoBlockGS(V, W, ldw);
  if (!e) switch (V->orthog_block) {
      case BV_ORTHOG_BLOCK_CHOL: e = OrthoChol(V, W, ldw); break;
      case BV_ORTHOG_BLOCK_TSQR: e = OrthoTSQR(V, W, ldw, PETSC_TRUE); break;
      case BV_ORTHOG_BLOCK_TSQRCHOL: e = OrthoTSQR(V, W, ldw, PETSC_FALSE); break;
      case BV_ORTHOG_BLOCK_SVQB: e = OrthoSVQB(V, W, ldw); break;
      default: e = PETSC_ERR_ARG_OUTOFRANGE; break;
    }
  if (nc) {
    V->nc = nc;
    V->m = msave;
    V->l = l;
    V->k = k;
  }
  if (!e && R) {   /* BV_StoreCoeffsBlock_Default: columns l..k-1, rows 0..j (triangular) or 0..k-1 (SVQB) */
    PetscScalar *r;
    PetscInt ldr;
    PetscCall(DenseAt(R, &r, &ldr));
    for (PetscInt j = l; j < k; j++) {
      const PetscInt rows = V->orthog_block == BV_ORTHOG_BLOCK_SVQB ? k : j + 1;
      for (PetscInt i = 0; i < rows; i++) r[i + j * ldr] = W[(i + nc) + (j + nc) * ldw];
    }
  }
  free(W);
  return e;
}

/* BVMatProject (bvglobal.c:1093-1160): every entry of M(0:ky, 0:kx) but the leading
   ly x lx block, M = Y^T A X (A = NULL: X holds A X already) */
PetscErrorCode BVMatProject(BV X, Mat A, BV Y, Mat M)
{
  PetscInt m, n, ldm, lx, kx, ly, ky;
  PetscScalar *mv, *g = NULL;
  BV W = NULL;
  PetscErrorCode e = PETSC_SUCCESS;
  PetscCall(BVCheckReady(X));
  PetscCall(BVCheckReady(Y));
  PetscCheck(M && MatTypeIs(M, MATSEQDENSE), PETSC_ERR_ARG_WRONG, "M must be a sequential dense matrix");
  PetscCall(MatGetSize(M, &m, &n));
  PetscCheck(m >= Y->k, PETSC_ERR_ARG_SIZ, "Matrix M has %d rows, should have at least %d", m, Y->k);
  PetscCheck(n >= X->k, PETSC_ERR_ARG_SIZ, "Matrix M has %d columns, should have at least %d", n, X->k);
  PetscCheck(X->n == Y->n, PETSC_ERR_ARG_INCOMP, "Mismatching local dimension X %d, Y %d", X->n, Y->n);
  PetscCheck(X->ops->dot, PETSC_ERR_SUP, "BV type %s has no block dot", BVTypeName(X));
  lx = X->l; kx = X->k; ly = Y->l; ky = Y->k;
  if (!kx || !ky) return PETSC_SUCCESS;
  PetscCall(DenseAt(M, &mv, &ldm));
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)ky * (size_t)kx, &g));
  if (A) {   /* W = A X(:, 0:kx), one product per column */
    e = BVCreate(&W);
    if (!e) e = BVSetSizes(W, X->n, X->N, kx);
    if (!e) e = BVSetVecType(W, X->vtype);
    if (!e) e = BVSetType(W, BVTypeName(X));
    for (PetscInt j = 0; j < kx && !e; j++) {
      Vec x = NULL, w = NULL;
      e = BVGetColumn(X, j, &x);
      if (!e) e = BVGetColumn(W, j, &w);
      if (!e) e = MatMult(A, x, w);
      if (w) BVRestoreColumn(W, j, &w);
      if (x) BVRestoreColumn(X, j, &x);
    }
  }
  if (!e) e = BVDotRange(A ? W : X, 0, kx, Y, 0, ky, g, ky);
  if (!e)
    for (PetscInt j = 0; j < kx; j++)
      for (PetscInt i = 0; i < ky; i++)
        if (i >= ly || j >= lx) mv[i + j * ldm] = g[i + j * ky];
  if (W) BVDestroy(&W);
  free(g);
  return e;
}

/* ---------------------------------------------------------------- coefficients
 * The (nc+m) x m buffer of the orthogonalization (bvimpl.h:284-415): column 0
 * is scratch for the coefficients c of a pass, column j accumulates h and gets
 * |q| at row nc+j.  On a device BV (bv->hip) the buffer is written in place on
 * the device, as the reference's _HIP variants do (bvhip.hip.cpp:343-420). */
PetscErrorCode BVGetBufferVec(BV bv, Vec *buffer)
{
  PetscCall(BVCheckReady(bv));
  if (!bv->buffer) PetscCall(VecCreateHIP((bv->nc + bv->m) * bv->m, SLEPC_AMD_F64, &bv->buffer));
  *buffer = bv->buffer;
  return PETSC_SUCCESS;
}

/* BV_CleanCoefficients: entries 0..nc+j-1 of column j (or of h) to zero */
PetscErrorCode BV_CleanCoefficients(BV bv, PetscInt j, PetscScalar *h)
{
  const PetscInt ldb = bv->nc + bv->m;
  if (h) {
    for (PetscInt i = 0; i < bv->nc + j; i++) h[i] = 0.0;
    return PETSC_SUCCESS;
  }
  if (!bv->buffer) PetscCall(BVGetBufferVec(bv, &bv->buffer));
  if (bv->hip) {
    double *d;
    PetscCall(VecHIPGetArray(bv->buffer, (void **)&d));
    KSCall(ks_memset_zero(SlepcAMDContext(), d + (size_t)j * ldb, sizeof(double) * (size_t)(bv->nc + j)));
  } else {
    PetscScalar *a;
    PetscCall(VecGetArray(bv->buffer, &a));
    for (PetscInt i = 0; i < bv->nc + j; i++) a[(size_t)j * ldb + i] = 0.0;
    PetscCall(VecRestoreArray(bv->buffer, &a));
  }
  return PETSC_SUCCESS;
}

/* BV_SetValue: row nc+j of column k (or of h) */
PetscErrorCode BV_SetValue(BV bv, PetscInt j, PetscInt k, PetscScalar *h, PetscScalar value)
{
  const PetscInt ldb = bv->nc + bv->m;
  if (h) {
    h[bv->nc + j] = value;
    return PETSC_SUCCESS;
  }
  if (!bv->buffer) PetscCall(BVGetBufferVec(bv, &bv->buffer));
  if (bv->hip) {
    double *d;
    PetscCall(VecHIPGetArray(bv->buffer, (void **)&d));
    KSCall(ks_memcpy_htod(SlepcAMDContext(), d + (size_t)k * ldb + bv->nc + j, &value, sizeof(double)));
  } else {
    PetscScalar *a;
    PetscCall(VecGetArray(bv->buffer, &a));
    a[(size_t)k * ldb + bv->nc + j] = value;
    PetscCall(VecRestoreArray(bv->buffer, &a));
  }
  return PETSC_SUCCESS;
}

/* BV_StoreCoefficients: values l..j-1 of column j (or of h) into dest */
PetscErrorCode BV_StoreCoefficients(BV bv, PetscInt j, PetscScalar *h, PetscScalar *dest)
{
  const PetscScalar *a = NULL, *hh = h;
  if (!h) {
    if (!bv->buffer) PetscCall(BVGetBufferVec(bv, &bv->buffer));
    PetscCall(VecGetArrayRead(bv->buffer, &a));
    hh = a + (size_t)j * (bv->nc + bv->m);
  }
  for (PetscInt i = bv->l; i < j; i++) dest[i - bv->l] = hh[bv->nc + i];
  if (!h) PetscCall(VecRestoreArrayRead(bv->buffer, &a));
  return PETSC_SUCCESS;
}

/* BV_AllocateCoeffs (bvimpl.h): host h and c for BVOrthogonalizeVec */
static PetscErrorCode BV_AllocateCoeffs(BV bv)
{
  if (!bv->h) PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)(bv->nc + bv->m + 1), &bv->h));
  if (!bv->c) PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)(bv->nc + bv->m + 1), &bv->c));
  return PETSC_SUCCESS;
}

/* BV_NormVecOrColumn (bvorthog.c:20-30) */
static PetscErrorCode BV_NormVecOrColumn(BV bv, PetscInt j, Vec v, PetscReal *nrm)
{
  if (v) return VecNorm(v, NORM_2, nrm);
  return BVNormColumn(bv, j, NORM_2, nrm);
}

/* the extension hook (see slepcimpl.h) */
PetscErrorCode BVGetKrylovOps(BV bv, const BVKrylovOps **ops)
{
  PetscErrorCode (*f)(BV, const BVKrylovOps **) = NULL;
  *ops = NULL;
  PetscCall(PetscObjectQueryFunction(bv, "BVGetKrylovOps_C", &f));
  if (f) PetscCall(f(bv, ops));
  return PETSC_SUCCESS;
}

/* ---------------------------------------------------------------- Gram-Schmidt */
/* one CGS pass (BVOrthogonalizeGS1 of bvorthog.c:134): the type's gramschmidt op */
static PetscErrorCode BVOrthogonalizeGS1(BV bv, PetscInt j, Vec v, PetscBool *which, PetscScalar *h, PetscScalar *c,
                                         PetscReal *onorm, PetscReal *norm)
{
  PetscCheck(bv->ops->gramschmidt, PETSC_ERR_SUP, "BV type %s has no gramschmidt op", BVTypeName(bv));
  return bv->ops->gramschmidt(bv, j, v, which, h, c, onorm, norm);
}

/* BVOrthogonalizeGS (bvorthog.c:145-217): CGS with the refinement policy */
static PetscErrorCode BVOrthogonalizeGS(BV bv, PetscInt j, Vec v, PetscBool *which, PetscReal *norm,
                                        PetscBool *lindep)
{
  PetscScalar *h, *c;
  PetscReal onrm = 0.0, nrm = 0.0;
  PetscInt k, l;
  PetscBool dolindep = lindep ? PETSC_TRUE : PETSC_FALSE;
  if (v) {
    k = bv->k;
    h = bv->h;
    c = bv->c;
  } else {
    k = j;
    h = NULL;
    c = NULL;
  }
  PetscCall(BV_CleanCoefficients(bv, k, h));
  switch (bv->orthog_ref) {
  case BV_ORTHOG_REFINE_IFNEEDED:
    PetscCall(BVOrthogonalizeGS1(bv, k, v, which, h, c, &onrm, &nrm));
    /* repeat if ||q|| < eta ||h|| */
    l = 1;
    while (l < 3 && nrm != 0.0 && PetscAbsReal(nrm) < bv->orthog_eta * PetscAbsReal(onrm)) {
      l++;
      PetscCall(BVOrthogonalizeGS1(bv, k, v, which, h, c, &onrm, &nrm));
    }
    /* linear dependence check: criterion not satisfied in the last iteration */
    if (dolindep) *lindep = !(nrm != 0.0 && PetscAbsReal(nrm) >= bv->orthog_eta * PetscAbsReal(onrm));
    break;
  case BV_ORTHOG_REFINE_NEVER:
    PetscCall(BVOrthogonalizeGS1(bv, k, v, which, h, c, NULL, NULL));
    /* compute ||v|| */
    if (norm || dolindep) PetscCall(BV_NormVecOrColumn(bv, k, v, &nrm));
    /* linear dependence check: just test for exactly zero norm */
    if (dolindep) *lindep = !(nrm != 0.0);
    break;
  case BV_ORTHOG_REFINE_ALWAYS:
    PetscCall(BVOrthogonalizeGS1(bv, k, v, which, h, c, NULL, NULL));
    PetscCall(BVOrthogonalizeGS1(bv, k, v, which, h, c, dolindep ? &onrm : NULL, (norm || dolindep) ? &nrm : NULL));
    /* linear dependence check: criterion not satisfied in the second iteration */
    if (dolindep) *lindep = !(nrm != 0.0 && PetscAbsReal(nrm) >= bv->orthog_eta * PetscAbsReal(onrm));
    break;
  }
  if (norm) {
    *norm = nrm;
    if (!v) {   /* store norm value next to the orthogonalization coefficients */
      if (dolindep && *lindep) PetscCall(BV_SetValue(bv, k, k, h, 0.0));
      else PetscCall(BV_SetValue(bv, k, k, h, nrm));
    }
  }
  return PETSC_SUCCESS;
}

/* BVOrthogonalizeVec (bvorthog.c:236-266): v against the active columns and the constraints */
PetscErrorCode BVOrthogonalizeVec(BV bv, Vec v, PetscScalar H[], PetscReal *norm, PetscBool *lindep)
{
  PetscInt ksave, lsave;
  PetscErrorCode e;
  PetscCall(BVCheckReady(bv));
  PetscCheck(v && v->n == bv->n, PETSC_ERR_ARG_SIZ, "vector size does not match");
  PetscCall(BV_AllocateCoeffs(bv));
  ksave = bv->k;
  lsave = bv->l;
  bv->l = -bv->nc;
  e = BVOrthogonalizeGS(bv, 0, v, NULL, norm, lindep);
  bv->k = ksave;
  bv->l = lsave;
  PetscCall(e);
  if (H) for (PetscInt i = bv->l; i < bv->k; i++) H[i - bv->l] = bv->h[bv->nc + i];
  return PETSC_SUCCESS;
}

/* BVOrthogonalizeColumn (bvorthog.c:285-340) */
PetscErrorCode BVOrthogonalizeColumn(BV bv, PetscInt j, PetscScalar H[], PetscReal *norm, PetscBool *lindep)
{
  PetscInt ksave, lsave;
  const BVKrylovOps *ext;
  PetscErrorCode e;
  PetscCall(BVCheckReady(bv));
  PetscCheck(j >= 0 && j < bv->m, PETSC_ERR_ARG_OUTOFRANGE, "Index j=%d but BV only has %d columns", j, bv->m);
  PetscCall(BVGetKrylovOps(bv, &ext));
  ksave = bv->k;
  lsave = bv->l;
  bv->l = -bv->nc;   /* must also orthogonalize against constraints and leading columns */
  if (!bv->buffer) PetscCall(BVGetBufferVec(bv, &bv->buffer));
  if (ext) e = ext->orthonormalizecolumn(bv, j, PETSC_FALSE, PETSC_FALSE, norm, lindep);
  else e = BVOrthogonalizeGS(bv, j, NULL, NULL, norm, lindep);
  bv->k = ksave;
  bv->l = lsave;
  PetscCall(e);
  if (H) PetscCall(BV_StoreCoefficients(bv, j, NULL, H));
  return PETSC_SUCCESS;
}

/* BV_OrthogonalizeColumn_Safe (bvimpl.h:446-465): one pass, no refinement, the
   column reported as dependent */
PetscErrorCode BVOrthogonalizeColumn_Safe(BV bv, PetscInt j, PetscReal *norm, PetscBool *lindep)
{
  const BVKrylovOps *ext;
  PetscCall(BVGetKrylovOps(bv, &ext));
  if (ext) {
    PetscInt lsave = bv->l;
    bv->l = -bv->nc;
    PetscErrorCode e = ext->orthonormalizecolumn(bv, j, PETSC_FALSE, PETSC_TRUE, NULL, NULL);
    bv->l = lsave;
    PetscCall(e);
  } else {
    BVOrthogRefineType orthog_ref = bv->orthog_ref;
    bv->orthog_ref = BV_ORTHOG_REFINE_NEVER;   /* avoid refinement */
    PetscErrorCode e = BVOrthogonalizeColumn(bv, j, NULL, NULL, NULL);
    bv->orthog_ref = orthog_ref;
    PetscCall(e);
  }
  if (norm) *norm = 0.0;
  if (lindep) *lindep = PETSC_TRUE;
  return PETSC_SUCCESS;
}

/* BVOrthonormalizeColumn (bvorthog.c:380-427) */
PetscErrorCode BVOrthonormalizeColumn(BV bv, PetscInt j, PetscBool replace, PetscReal *norm, PetscBool *lindep)
{
  PetscReal nrm = 0.0;
  PetscInt ksave, lsave;
  PetscBool lndep = PETSC_FALSE;
  const BVKrylovOps *ext;
  PetscErrorCode e;
  PetscCall(BVCheckReady(bv));
  PetscCheck(j >= 0 && j < bv->m, PETSC_ERR_ARG_OUTOFRANGE, "Index j=%d but BV only has %d columns", j, bv->m);
  PetscCall(BVGetKrylovOps(bv, &ext));
  ksave = bv->k;
  lsave = bv->l;
  bv->l = -bv->nc;   /* must also orthogonalize against constraints and leading columns */
  if (!bv->buffer) PetscCall(BVGetBufferVec(bv, &bv->buffer));
  if (ext) {   /* all passes and the scaling on the device */
    e = ext->orthonormalizecolumn(bv, j, PETSC_TRUE, PETSC_FALSE, &nrm, &lndep);
    if (!e && replace && (nrm == 0.0 || lndep)) {
      e = BVSetRandomColumn(bv, j);
      if (!e) e = ext->orthonormalizecolumn(bv, j, PETSC_TRUE, PETSC_FALSE, &nrm, &lndep);
      if (!e && (nrm == 0.0 || lndep)) {   /* yet another attempt */
        e = BVSetRandomColumn(bv, j);
        if (!e) e = ext->orthonormalizecolumn(bv, j, PETSC_TRUE, PETSC_FALSE, &nrm, &lndep);
      }
    }
    bv->k = ksave;
    bv->l = lsave;
    PetscCall(e);
  } else {
    e = BVOrthogonalizeGS(bv, j, NULL, NULL, &nrm, &lndep);
    if (!e && replace && (nrm == 0.0 || lndep)) {
      e = BVSetRandomColumn(bv, j);
      if (!e) e = BVOrthogonalizeGS(bv, j, NULL, NULL, &nrm, &lndep);
      if (!e && (nrm == 0.0 || lndep)) {   /* yet another attempt */
        e = BVSetRandomColumn(bv, j);
        if (!e) e = BVOrthogonalizeGS(bv, j, NULL, NULL, &nrm, &lndep);
      }
    }
    bv->k = ksave;
    bv->l = lsave;
    PetscCall(e);
    /* scale */
    if (nrm != 1.0 && nrm != 0.0) PetscCall(bv->ops->scale(bv, j, 1.0 / nrm));
  }
  if (norm) *norm = nrm;
  if (lindep) *lindep = lndep;
  return PETSC_SUCCESS;
}

/* BVMatMultColumn (bvops.c:862-885) */
PetscErrorCode BVMatMultColumn(BV bv, Mat A, PetscInt j)
{
  Vec vj, vj1;
  PetscCall(BVCheckReady(bv));
  PetscCheck(j >= 0 && j + 1 < bv->m, PETSC_ERR_ARG_OUTOFRANGE, "Result should go in index j+1=%d but BV only has %d columns", j + 1, bv->m);
  PetscCall(BVGetColumn(bv, j, &vj));
  PetscCall(BVGetColumn(bv, j + 1, &vj1));
  PetscErrorCode e = MatMult(A, vj, vj1);
  PetscCall(BVRestoreColumn(bv, j, &vj));
  PetscCall(BVRestoreColumn(bv, j + 1, &vj1));
  return e;
}

PetscErrorCode BVGetBufferHost(BV bv, const PetscScalar **a, PetscInt *ld)
{
  Vec buf;
  PetscCall(BVGetBufferVec(bv, &buf));
  PetscCall(VecGetArrayRead(buf, a));
  if (ld) *ld = bv->nc + bv->m;
  return PETSC_SUCCESS;
}

/* ---------------------------------------------------------------- Krylov factorizations */
/* the expansion loop of BVMatArnoldi/BVMatLanczos (bvkrylov.c:88-97) */
static PetscErrorCode BVMatKrylovLoop(BV V, Mat A, PetscInt k, PetscInt *m, PetscReal *beta, PetscBool *breakdown)
{
  PetscBool lindep = PETSC_FALSE;
  for (PetscInt j = k; j < *m; j++) {
    PetscCall(BVMatMultColumn(V, A, j));
    if (j == V->N - 1) PetscCall(BVOrthogonalizeColumn_Safe(V, j + 1, beta, &lindep));   /* full basis requested */
    else PetscCall(BVOrthonormalizeColumn(V, j + 1, PETSC_FALSE, beta, &lindep));
    if (lindep) {
      *m = j + 1;
      break;
    }
  }
  *breakdown = lindep;
  return PETSC_SUCCESS;
}

static PetscErrorCode BVMatKrylovCheck(BV V, PetscInt k, PetscInt m)
{
  PetscCall(BVCheckReady(V));
  PetscCheck(k >= 0 && k <= V->m, PETSC_ERR_ARG_OUTOFRANGE, "Argument k has wrong value %d, should be between 0 and %d", k, V->m);
  PetscCheck(m > 0 && m <= V->m, PETSC_ERR_ARG_OUTOFRANGE, "Argument m has wrong value %d, should be between 1 and %d", m, V->m);
  PetscCheck(m > k, PETSC_ERR_ARG_OUTOFRANGE, "Argument m should be at least equal to k+1");
  return PETSC_SUCCESS;
}

/* The expansion split in two through the type's Krylov extension: Begin
   enqueues steps k..m-1 on the device, End waits once and returns the
   breakdown position and beta, with the new coefficient columns in the
   buffer's host copy. */
PetscErrorCode BVMatKrylovBegin(BV V, Mat A, PetscInt k, PetscInt m)
{
  const BVKrylovOps *ext;
  PetscCall(BVMatKrylovCheck(V, k, m));
  PetscCall(BVGetKrylovOps(V, &ext));
  PetscCheck(ext, PETSC_ERR_SUP, "BV type %s offers no device Krylov expansion", BVTypeName(V));
  return ext->matkrylovbegin(V, A, k, m);
}

PetscErrorCode BVMatKrylovEnd(BV V, PetscInt k, PetscInt *m, PetscReal *beta, PetscBool *breakdown)
{
  const BVKrylovOps *ext;
  PetscCall(BVGetKrylovOps(V, &ext));
  PetscCheck(ext, PETSC_ERR_SUP, "BV type %s offers no device Krylov expansion", BVTypeName(V));
  return ext->matkrylovend(V, k, m, beta, breakdown);
}

/* steps k..m-1 with either path, waiting for the result */
PetscErrorCode BVMatKrylov(BV V, Mat A, PetscInt k, PetscInt *m, PetscReal *beta, PetscBool *breakdown)
{
  const BVKrylovOps *ext;
  PetscCall(BVMatKrylovCheck(V, k, *m));
  PetscCall(BVGetKrylovOps(V, &ext));
  *beta = 0.0;
  *breakdown = PETSC_FALSE;
  if (!ext) return BVMatKrylovLoop(V, A, k, m, beta, breakdown);
  PetscCall(ext->matkrylovbegin(V, A, k, *m));
  return ext->matkrylovend(V, k, m, beta, breakdown);
}

/* H columns k..m-1 from the coefficient buffer (bvkrylov.c:100-109) */
PetscErrorCode BVCopyHessenberg(BV V, Mat H, PetscInt k, PetscInt m)
{
  Mat_SeqDense *hd = (Mat_SeqDense *)H->data;
  PetscScalar *h = hd->v;
  PetscInt ldh = hd->lda, ldb = V->nc + V->m;
  const PetscScalar *a;
  Vec buf;
  PetscCheck(MatTypeIs(H, MATSEQDENSE), PETSC_ERR_ARG_WRONG, "H must be seqdense");
  PetscCheck(H->m >= m && H->n >= m, PETSC_ERR_ARG_SIZ, "Matrix H is too small");
  PetscCall(BVGetBufferVec(V, &buf));
  PetscCall(VecGetArrayRead(buf, &a));
  for (PetscInt j = k; j < m - 1; j++) memcpy(h + (size_t)j * ldh, a + V->nc + (size_t)(j + 1) * ldb, sizeof(PetscScalar) * (size_t)(j + 2));
  memcpy(h + (size_t)(m - 1) * ldh, a + V->nc + (size_t)m * ldb, sizeof(PetscScalar) * (size_t)m);
  if (ldh > m) h[m + (size_t)(m - 1) * ldh] = a[V->nc + m + (size_t)m * ldb];
  PetscCall(VecRestoreArrayRead(buf, &a));
  return PETSC_SUCCESS;
}

/* alpha, beta of steps k..m-1 (bvkrylov.c:210-222): T stores alpha in column 0, beta in column 1 */
PetscErrorCode BVCopyTridiagonal(BV V, Mat T, PetscInt k, PetscInt m)
{
  Mat_SeqDense *td = (Mat_SeqDense *)T->data;
  PetscReal *alpha = td->v, *betat = td->v + td->lda;
  PetscInt ldb = V->nc + V->m;
  const PetscScalar *a;
  Vec buf;
  PetscCheck(MatTypeIs(T, MATSEQDENSE), PETSC_ERR_ARG_WRONG, "T must be seqdense");
  PetscCheck(T->m >= m && T->n >= 2, PETSC_ERR_ARG_SIZ, "Matrix T is too small");
  PetscCall(BVGetBufferVec(V, &buf));
  PetscCall(VecGetArrayRead(buf, &a));
  for (PetscInt j = k; j < m; j++) {
    alpha[j] = a[V->nc + j + (size_t)(j + 1) * ldb];
    betat[j] = a[V->nc + j + 1 + (size_t)(j + 1) * ldb];
  }
  PetscCall(VecRestoreArrayRead(buf, &a));
  return PETSC_SUCCESS;
}

/* BVMatArnoldi (bvkrylov.c:56-113) */
PetscErrorCode BVMatArnoldi(BV V, Mat A, Mat H, PetscInt k, PetscInt *m, PetscReal *beta, PetscBool *breakdown)
{
  PetscBool bd = PETSC_FALSE;
  PetscReal b = 0.0;
  PetscCall(BVMatKrylov(V, A, k, m, &b, &bd));
  if (breakdown) *breakdown = bd;
  if (beta) *beta = b;
  if (H) PetscCall(BVCopyHessenberg(V, H, k, *m));
  return PETSC_SUCCESS;
}

/* BVMatLanczos (bvkrylov.c:165-226) */
PetscErrorCode BVMatLanczos(BV V, Mat A, Mat T, PetscInt k, PetscInt *m, PetscReal *beta, PetscBool *breakdown)
{
  PetscBool bd = PETSC_FALSE;
  PetscReal b = 0.0;
  PetscCall(BVMatKrylov(V, A, k, m, &b, &bd));
  if (breakdown) *breakdown = bd;
  if (beta) *beta = b;
  if (T) PetscCall(BVCopyTridiagonal(V, T, k, *m));
  return PETSC_SUCCESS;
}

PetscErrorCode BVDestroy(BV *bv)
{
  if (!bv || !*bv) return PETSC_SUCCESS;
  PetscCall(BVReset_Private(*bv));
  PetscCall(PetscRandomDestroy(&(*bv)->rand));
  free(*bv);
  *bv = NULL;
  return PETSC_SUCCESS;
}
