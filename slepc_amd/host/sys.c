/*
 * sys.c -- initialization, errors, registries, timers and the eigenvalue
 * sorting criteria (src/sys/slepcsc.c:40-190).
 */
#include <stdarg.h>
#include <time.h>

#include "slepcimpl.h"

static char g_last_error[1024];
static ks_ctx g_ctx = NULL;
static PetscBool g_initialized = PETSC_FALSE;

PetscErrorCode SlepcError(int line, const char *func, const char *file, PetscErrorCode code, const char *fmt, ...)
{
  char msg[768];
  if (fmt) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, sizeof(msg), fmt, ap);
    va_end(ap);
    snprintf(g_last_error, sizeof(g_last_error), "[%d] %s (%s:%d): %s", code, func, file, line, msg);
    fprintf(stderr, "slepc_amd error %d in %s() at %s:%d: %s\n", code, func, file, line, msg);
  } else {
    fprintf(stderr, "slepc_amd:   called from %s() at %s:%d\n", func, file, line);
  }
  return code;
}

const char *SlepcAMDGetLastError(void) { return g_last_error; }

PetscErrorCode PetscMallocZero(size_t bytes, void *ptr)
{
  void **p = (void **)ptr;
  *p = calloc(1, bytes ? bytes : 1);
  if (!*p) SETERRQ(PETSC_ERR_MEM, "out of host memory (%zu bytes)", bytes);
  return PETSC_SUCCESS;
}

double SlepcWallTime(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

PetscErrorCode SlepcFunctionListAdd(SlepcFunctionList *fl, const char *name, void *create)
{
  for (int i = 0; i < fl->n; i++) {
    if (!strcmp(fl->entry[i].name, name)) { fl->entry[i].create = create; return PETSC_SUCCESS; }
  }
  PetscCheck(fl->n < 16, PETSC_ERR_SUP, "registry full");
  PetscCheck(strlen(name) < SLEPC_NAME_LEN, PETSC_ERR_ARG_WRONG, "type name too long: %s", name);
  strcpy(fl->entry[fl->n].name, name);
  fl->entry[fl->n].create = create;
  fl->n++;
  return PETSC_SUCCESS;
}

void *SlepcFunctionListFind(SlepcFunctionList *fl, const char *name)
{
  for (int i = 0; i < fl->n; i++)
    if (!strcmp(fl->entry[i].name, name)) return fl->entry[i].create;
  return NULL;
}

/* MKL service (mkl_rt), by value.  3 selects its GNU OpenMP threading layer:
   the layer is fixed for the process at MKL's first call, and Intel-threaded
   MKL called from inside a GNU OpenMP region (a host application's, or the
   CPU checks') computes wrong results once a second libgomp is loaded. */
extern int MKL_Set_Threading_Layer(int);

/* SlepcInitialize: selects the device, creates the shared stream context and
   registers the built-in types (dlregisslepc.c registers the reference's). */
PetscErrorCode SlepcInitialize(int device)
{
  if (g_initialized) return PETSC_SUCCESS;
  (void)MKL_Set_Threading_Layer(3);
  KSCall(ks_init(device));
  KSCall(ks_ctx_create(&g_ctx));
  PetscCall(MatRegister(MATHIPCSR, MatCreate_HIPCSR));
  PetscCall(BVRegister(BVHIPKS, BVCreate_HIPKS));
  PetscCall(DSRegister(DSHEP, DSCreate_HEP));
  PetscCall(DSRegister(DSNHEP, DSCreate_NHEP));
  PetscCall(DSRegister(DSNHEPTS, DSCreate_NHEPTS));
  PetscCall(EPSRegister(EPSKRYLOVSCHUR, EPSCreate_KrylovSchur));
  g_initialized = PETSC_TRUE;
  return PETSC_SUCCESS;
}

PetscErrorCode SlepcFinalize(void)
{
  if (!g_initialized) return PETSC_SUCCESS;
  KSCall(ks_ctx_destroy(g_ctx));
  g_ctx = NULL;
  g_initialized = PETSC_FALSE;
  return PETSC_SUCCESS;
}

PetscErrorCode SlepcInitialized(PetscBool *flg)
{
  *flg = g_initialized;
  return PETSC_SUCCESS;
}

PetscErrorCode SlepcCheckInitialized(void)
{
  PetscCheck(g_initialized, PETSC_ERR_ORDER, "SlepcInitialize() must be called first");
  return PETSC_SUCCESS;
}

/* The context new objects bind to and Vec/Mat operations run on.  The serving
   loop makes a request's own context current while it works on that request. */
static ks_ctx g_current = NULL;
ks_ctx SlepcAMDContext(void) { return g_current ? g_current : g_ctx; }
ks_ctx SlepcAMDSetCurrentContext(ks_ctx ctx)
{
  ks_ctx prev = g_current;
  g_current = ctx;
  return prev;
}
void *SlepcAMDGetDefaultContext(void) { return (void *)g_ctx; }

/* ---------------------------------------------------------------- sorting criteria */
/* SlepcCompare* (slepcsc.c:152-190): < 0 means a goes first */
int SlepcSCCompare(const SlepcSC *sc, PetscScalar ar, PetscScalar ai, PetscScalar br, PetscScalar bi)
{
  PetscReal a, b;
  if (sc->map) {   /* SlepcMap_ST (slepcsc.c:118-126): compare the back-transformed values */
    PetscScalar na = ar * ar + ai * ai, nb = br * br + bi * bi;
    if (!ai) ar = 1.0 / ar + sc->sigma;
    else { ar = ar / na + sc->sigma; ai = -ai / na; }
    if (!bi) br = 1.0 / br + sc->sigma;
    else { br = br / nb + sc->sigma; bi = -bi / nb; }
  }
  switch (sc->which) {
  case EPS_LARGEST_MAGNITUDE:  a = SlepcAbsEigenvalue(ar, ai); b = SlepcAbsEigenvalue(br, bi); return a < b ? 1 : (a > b ? -1 : 0);
  case EPS_SMALLEST_MAGNITUDE: a = SlepcAbsEigenvalue(ar, ai); b = SlepcAbsEigenvalue(br, bi); return a > b ? 1 : (a < b ? -1 : 0);
  case EPS_LARGEST_REAL:       a = ar; b = br; return a < b ? 1 : (a > b ? -1 : 0);
  case EPS_SMALLEST_REAL:      a = ar; b = br; return a > b ? 1 : (a < b ? -1 : 0);
  case EPS_LARGEST_IMAGINARY:  a = fabs(ai); b = fabs(bi); return a < b ? 1 : (a > b ? -1 : 0);
  case EPS_SMALLEST_IMAGINARY: a = fabs(ai); b = fabs(bi); return a > b ? 1 : (a < b ? -1 : 0);
  case EPS_TARGET_MAGNITUDE:
    a = SlepcAbsEigenvalue(ar - sc->target, ai); b = SlepcAbsEigenvalue(br - sc->target, bi);
    return a > b ? 1 : (a < b ? -1 : 0);
  case EPS_TARGET_REAL:
    a = fabs(ar - sc->target); b = fabs(br - sc->target);
    return a > b ? 1 : (a < b ? -1 : 0);
  }
  return 0;
}

/* SlepcSortEigenvalues (slepcsc.c:89-150): insertion sort keeping conjugate pairs together */
void SlepcSortEigenvalues(const SlepcSC *sc, PetscInt n, const PetscScalar *eigr, const PetscScalar *eigi,
                          PetscInt *perm)
{
  for (PetscInt i = n - 1; i >= 0; i--) {
    PetscScalar re = eigr[perm[i]], im = eigi[perm[i]];
    PetscInt j = i + 1, tmp;
    if (im != 0) { i--; im = eigi[perm[i]]; }
    while (j < n) {
      if (SlepcSCCompare(sc, re, im, eigr[perm[j]], eigi[perm[j]]) <= 0) break;
      if (!im) {
        if (eigi[perm[j]] == 0.0) {
          tmp = perm[j - 1]; perm[j - 1] = perm[j]; perm[j] = tmp; j++;
        } else {
          tmp = perm[j - 1]; perm[j - 1] = perm[j]; perm[j] = perm[j + 1]; perm[j + 1] = tmp; j += 2;
        }
      } else {
        if (eigi[perm[j]] == 0.0) {
          tmp = perm[j - 2]; perm[j - 2] = perm[j]; perm[j] = perm[j - 1]; perm[j - 1] = tmp; j++;
        } else {
          tmp = perm[j - 2]; perm[j - 2] = perm[j]; perm[j] = tmp;
          tmp = perm[j - 1]; perm[j - 1] = perm[j + 1]; perm[j + 1] = tmp; j += 2;
        }
      }
    }
  }
}
