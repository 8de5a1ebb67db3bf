/*
 * vec.c -- device vectors (the PETSc Vec operations the Krylov-Schur path uses:
 * VecCopy/VecAXPY/VecScale/VecNorm in EPSComputeResidualNorm_Private,
 * src/eps/interface/epssolve.c:666-718).  Storage lives in HBM; host values
 * move only through VecSetValuesHost/VecGetValuesHost.
 */
#include "slepcimpl.h"

size_t SlepcPrecSize(SlepcAMDPrecision p) { return p == SLEPC_AMD_F64 ? 8 : 4; }

static ks_prec kprec(SlepcAMDPrecision p) { return p == SLEPC_AMD_F64 ? KS_F64 : KS_F32; }

PetscErrorCode VecCreateHIP(PetscInt n, SlepcAMDPrecision prec, Vec *v)
{
  Vec x;
  PetscCall(SlepcCheckInitialized());
  PetscCheck(n >= 0, PETSC_ERR_ARG_OUTOFRANGE, "negative vector size %d", n);
  PetscCall(PetscMallocZero(sizeof(struct _p_Vec), &x));
  x->n = n;
  x->prec = prec;
  x->owns = PETSC_TRUE;
  KSCall(ks_malloc(&x->d, SlepcPrecSize(prec) * (size_t)(n ? n : 1)));
  KSCall(ks_memset_zero(SlepcAMDContext(), x->d, SlepcPrecSize(prec) * (size_t)n));
  *v = x;
  return PETSC_SUCCESS;
}

PetscErrorCode VecCreateHIPWithArray(PetscInt n, SlepcAMDPrecision prec, void *d, Vec *v)
{
  Vec x;
  PetscCall(PetscMallocZero(sizeof(struct _p_Vec), &x));
  x->n = n;
  x->prec = prec;
  x->d = d;
  x->owns = PETSC_FALSE;
  *v = x;
  return PETSC_SUCCESS;
}

PetscErrorCode VecDuplicate(Vec v, Vec *w)
{
  PetscCheck(v, PETSC_ERR_ARG_NULL, "null vector");
  return VecCreateHIP(v->n, v->prec, w);
}

PetscErrorCode VecDestroy(Vec *v)
{
  if (!v || !*v) return PETSC_SUCCESS;
  if ((*v)->owns) KSCall(ks_free((*v)->d));
  if ((*v)->h) KSCall(ks_host_free((*v)->h));
  free(*v);
  *v = NULL;
  return PETSC_SUCCESS;
}

PetscErrorCode VecGetSize(Vec v, PetscInt *n)
{
  PetscCheck(v, PETSC_ERR_ARG_NULL, "null vector");
  *n = v->n;
  return PETSC_SUCCESS;
}

/* the device copy changed: a host copy, if any, is out of date */
static void VecDeviceModified(Vec v)
{
  if (v->mask == PETSC_OFFLOAD_BOTH) v->mask = PETSC_OFFLOAD_GPU;
}

PetscErrorCode VecSetRandomSeeded(Vec v, uint64_t seed)
{
  PetscCheck(v, PETSC_ERR_ARG_NULL, "null vector");
  KSCall(ks_bv_random(SlepcAMDContext(), kprec(v->prec), v->n, v->d, seed));
  VecDeviceModified(v);
  return PETSC_SUCCESS;
}

/* VecSetRandom with the generator's next stream (PetscRandom, see slepcimpl.h) */
PetscErrorCode VecSetRandom(Vec v, PetscRandom r)
{
  PetscCheck(v && r, PETSC_ERR_ARG_NULL, "null argument");
  r->count++;
  return VecSetRandomSeeded(v, r->seed + 1000003ULL * r->count);
}

/* device pointer with write access: the host copy becomes stale */
PetscErrorCode VecHIPGetArray(Vec v, void **d)
{
  PetscCheck(v, PETSC_ERR_ARG_NULL, "null vector");
  *d = v->d;
  VecDeviceModified(v);
  return PETSC_SUCCESS;
}

PetscErrorCode VecHIPGetArrayRead(Vec v, const void **d)
{
  PetscCheck(v, PETSC_ERR_ARG_NULL, "null vector");
  *d = v->d;
  return PETSC_SUCCESS;
}

static PetscErrorCode VecHostAllocate(Vec v)
{
  PetscCheck(v->prec == SLEPC_AMD_F64, PETSC_ERR_SUP, "host access to FP32 vectors goes through VecGetValuesHost");
  if (v->h) return PETSC_SUCCESS;
  /* pinned, so copies into it can be enqueued in stream order */
  KSCall(ks_host_alloc((void **)&v->h, sizeof(PetscScalar) * (size_t)(v->n ? v->n : 1)));
  v->mask = PETSC_OFFLOAD_GPU;
  return PETSC_SUCCESS;
}

static PetscErrorCode VecToHost(Vec v)
{
  PetscCall(VecHostAllocate(v));
  if (v->mask != PETSC_OFFLOAD_BOTH) {
    if (v->n) KSCall(ks_memcpy_dtoh(SlepcAMDContext(), v->h, v->d, sizeof(PetscScalar) * (size_t)v->n));
    v->mask = PETSC_OFFLOAD_BOTH;
  }
  return PETSC_SUCCESS;
}

PetscErrorCode VecGetArrayRead(Vec v, const PetscScalar **a)
{
  PetscCheck(v && a, PETSC_ERR_ARG_NULL, "null argument");
  PetscCall(VecToHost(v));
  *a = v->h;
  return PETSC_SUCCESS;
}

PetscErrorCode VecRestoreArrayRead(Vec v, const PetscScalar **a)
{
  PetscCheck(v && a && *a == v->h, PETSC_ERR_ARG_WRONG, "array does not belong to the vector");
  *a = NULL;
  return PETSC_SUCCESS;
}

PetscErrorCode VecGetArray(Vec v, PetscScalar **a)
{
  PetscCheck(v && a, PETSC_ERR_ARG_NULL, "null argument");
  PetscCall(VecToHost(v));
  *a = v->h;
  return PETSC_SUCCESS;
}

/* the host values go back to the device (the host copy stays current) */
PetscErrorCode VecRestoreArray(Vec v, PetscScalar **a)
{
  PetscCheck(v && a && *a == v->h, PETSC_ERR_ARG_WRONG, "array does not belong to the vector");
  if (v->n) KSCall(ks_memcpy_htod(SlepcAMDContext(), v->d, v->h, sizeof(PetscScalar) * (size_t)v->n));
  v->mask = PETSC_OFFLOAD_BOTH;
  *a = NULL;
  return PETSC_SUCCESS;
}

PetscErrorCode VecHIPGetHostCopy(Vec v, PetscScalar **h)
{
  PetscCheck(v && h, PETSC_ERR_ARG_NULL, "null argument");
  PetscCall(VecHostAllocate(v));
  *h = v->h;
  return PETSC_SUCCESS;
}

PetscErrorCode VecHIPSetHostCurrent(Vec v)
{
  PetscCheck(v && v->h, PETSC_ERR_ARG_WRONGSTATE, "vector has no host copy");
  v->mask = PETSC_OFFLOAD_BOTH;
  return PETSC_SUCCESS;
}

PetscErrorCode VecSet(Vec v, PetscScalar alpha)
{
  PetscCheck(v, PETSC_ERR_ARG_NULL, "null vector");
  if (alpha == 0.0) {
    KSCall(ks_memset_zero(SlepcAMDContext(), v->d, SlepcPrecSize(v->prec) * (size_t)v->n));
    VecDeviceModified(v);
    return PETSC_SUCCESS;
  }
  /* rare path: build on the host */
  PetscScalar *h;
  PetscCall(PetscMallocZero(sizeof(PetscScalar) * (size_t)(v->n ? v->n : 1), &h));
  for (PetscInt i = 0; i < v->n; i++) h[i] = alpha;
  PetscErrorCode e = VecSetValuesHost(v, h);
  free(h);
  return e;
}

PetscErrorCode VecCopy(Vec x, Vec y)
{
  PetscCheck(x && y, PETSC_ERR_ARG_NULL, "null vector");
  PetscCheck(x->n == y->n, PETSC_ERR_ARG_SIZ, "vector sizes %d and %d differ", x->n, y->n);
  if (x->prec == y->prec) KSCall(ks_bv_copy(SlepcAMDContext(), kprec(x->prec), x->n, x->d, y->d));
  else KSCall(ks_bv_convert(SlepcAMDContext(), x->n, kprec(x->prec), x->d, kprec(y->prec), y->d));
  VecDeviceModified(y);
  return PETSC_SUCCESS;
}

PetscErrorCode VecAXPY(Vec y, PetscScalar alpha, Vec x)
{
  PetscCheck(x && y, PETSC_ERR_ARG_NULL, "null vector");
  PetscCheck(x->n == y->n && x->prec == y->prec, PETSC_ERR_ARG_SIZ, "incompatible vectors");
  KSCall(ks_bv_axpy(SlepcAMDContext(), kprec(x->prec), x->n, alpha, x->d, y->d));
  VecDeviceModified(y);
  return PETSC_SUCCESS;
}

PetscErrorCode VecScale(Vec x, PetscScalar alpha)
{
  PetscCheck(x, PETSC_ERR_ARG_NULL, "null vector");
  KSCall(ks_bv_scale(SlepcAMDContext(), kprec(x->prec), x->n, x->d, alpha));
  VecDeviceModified(x);
  return PETSC_SUCCESS;
}

PetscErrorCode VecNorm(Vec x, NormType type, PetscReal *nrm)
{
  PetscCheck(x, PETSC_ERR_ARG_NULL, "null vector");
  PetscCheck(type == NORM_2 || type == NORM_FROBENIUS, PETSC_ERR_SUP, "only the 2-norm is implemented");
  if (!x->n) { *nrm = 0.0; return PETSC_SUCCESS; }
  KSCall(ks_bv_norm(SlepcAMDContext(), kprec(x->prec), x->n, x->d, nrm));
  return PETSC_SUCCESS;
}

PetscErrorCode VecSetValuesHost(Vec v, const PetscScalar *values)
{
  PetscCheck(v && (values || !v->n), PETSC_ERR_ARG_NULL, "null argument");
  if (!v->n) return PETSC_SUCCESS;
  VecDeviceModified(v);
  if (v->prec == SLEPC_AMD_F64) {
    KSCall(ks_memcpy_htod(SlepcAMDContext(), v->d, values, sizeof(double) * (size_t)v->n));
  } else {
    float *f;
    PetscCall(PetscMallocZero(sizeof(float) * (size_t)v->n, &f));
    for (PetscInt i = 0; i < v->n; i++) f[i] = (float)values[i];
    int e = ks_memcpy_htod(SlepcAMDContext(), v->d, f, sizeof(float) * (size_t)v->n);
    free(f);
    KSCall(e);
  }
  return PETSC_SUCCESS;
}

PetscErrorCode VecGetValuesHost(Vec v, PetscScalar *values)
{
  PetscCheck(v && (values || !v->n), PETSC_ERR_ARG_NULL, "null argument");
  if (!v->n) return PETSC_SUCCESS;
  if (v->prec == SLEPC_AMD_F64) {
    KSCall(ks_memcpy_dtoh(SlepcAMDContext(), values, v->d, sizeof(double) * (size_t)v->n));
  } else {
    float *f;
    PetscCall(PetscMallocZero(sizeof(float) * (size_t)v->n, &f));
    int e = ks_memcpy_dtoh(SlepcAMDContext(), f, v->d, sizeof(float) * (size_t)v->n);
    for (PetscInt i = 0; i < v->n && !e; i++) values[i] = f[i];
    free(f);
    KSCall(e);
  }
  return PETSC_SUCCESS;
}
