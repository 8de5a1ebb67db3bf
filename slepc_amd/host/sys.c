/*
 * sys.c -- initialization, errors, registries, timers and the eigenvalue
 * sorting criteria (src/sys/slepcsc.c:40-190).
 */
#include <stdarg.h>
#include <strings.h>
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

/* ---------------------------------------------------------------- objects */
static PetscObjectId g_next_id = 1;

PetscErrorCode PetscHeaderInit(PetscObject obj, const char *class_name)
{
  memset(obj, 0, sizeof(*obj));
  snprintf(obj->class_name, sizeof(obj->class_name), "%s", class_name);
  obj->id = g_next_id++;
  return PETSC_SUCCESS;
}

/* a NULL function removes the entry */
PetscErrorCode PetscObjectComposeFunction_Private(PetscObject obj, const char name[], PetscVoidFn fn)
{
  PetscCheck(obj && name, PETSC_ERR_ARG_NULL, "null argument");
  PetscCheck(strlen(name) < sizeof(obj->fname[0]), PETSC_ERR_ARG_WRONG, "composed name too long: %s", name);
  for (int i = 0; i < obj->nfn; i++) {
    if (strcmp(obj->fname[i], name)) continue;
    if (fn) obj->fn[i] = fn;
    else {
      obj->nfn--;
      memmove(obj->fname[i], obj->fname[obj->nfn], sizeof(obj->fname[0]));
      obj->fn[i] = obj->fn[obj->nfn];
    }
    return PETSC_SUCCESS;
  }
  if (!fn) return PETSC_SUCCESS;
  PetscCheck(obj->nfn < SLEPC_MAX_COMPOSED, PETSC_ERR_SUP, "too many composed functions");
  strcpy(obj->fname[obj->nfn], name);
  obj->fn[obj->nfn++] = fn;
  return PETSC_SUCCESS;
}

PetscErrorCode PetscObjectQueryFunction_Private(PetscObject obj, const char name[], PetscVoidFn *fn)
{
  PetscCheck(obj && name && fn, PETSC_ERR_ARG_NULL, "null argument");
  *fn = NULL;
  for (int i = 0; i < obj->nfn; i++)
    if (!strcmp(obj->fname[i], name)) { *fn = obj->fn[i]; break; }
  return PETSC_SUCCESS;
}

PetscErrorCode PetscRandomCreate(PetscRandom *r)
{
  PetscCall(PetscMallocZero(sizeof(struct _p_PetscRandom), r));
  return PETSC_SUCCESS;
}

PetscErrorCode PetscRandomDestroy(PetscRandom *r)
{
  if (r && *r) PetscFree(*r);
  return PETSC_SUCCESS;
}

/* ---------------------------------------------------------------- options database */
#define SLEPC_MAX_OPTIONS 128
static struct {
  char name[64];
  char value[128];
} g_opt[SLEPC_MAX_OPTIONS];
static int g_nopt = 0;

PetscErrorCode PetscOptionsSetValue(const char name[], const char value[])
{
  int i;
  PetscCheck(name && name[0] == '-', PETSC_ERR_ARG_WRONG, "option names start with '-'");
  PetscCheck(strlen(name) < sizeof(g_opt[0].name) && (!value || strlen(value) < sizeof(g_opt[0].value)),
             PETSC_ERR_ARG_WRONG, "option %s too long", name);
  for (i = 0; i < g_nopt; i++)
    if (!strcmp(g_opt[i].name, name)) break;
  if (i == g_nopt) {
    PetscCheck(g_nopt < SLEPC_MAX_OPTIONS, PETSC_ERR_SUP, "options database full");
    g_nopt++;
  }
  strcpy(g_opt[i].name, name);
  snprintf(g_opt[i].value, sizeof(g_opt[i].value), "%s", value ? value : "");
  return PETSC_SUCCESS;
}

PetscErrorCode PetscOptionsClearValue(const char name[])
{
  for (int i = 0; i < g_nopt; i++) {
    if (strcmp(g_opt[i].name, name)) continue;
    g_opt[i] = g_opt[--g_nopt];
    return PETSC_SUCCESS;
  }
  return PETSC_SUCCESS;
}

/* "-a 1 -b -c x": a name followed by a word not starting with '-' takes it as value */
PetscErrorCode PetscOptionsInsertString(const char in_str[])
{
  char buf[2048], *tok[256];
  int nt = 0;
  if (!in_str) return PETSC_SUCCESS;
  PetscCheck(strlen(in_str) < sizeof(buf), PETSC_ERR_ARG_WRONG, "options string too long");
  strcpy(buf, in_str);
  for (char *t = strtok(buf, " \t\n"); t && nt < 256; t = strtok(NULL, " \t\n")) tok[nt++] = t;
  for (int i = 0; i < nt; i++) {
    PetscCheck(tok[i][0] == '-', PETSC_ERR_ARG_WRONG, "unexpected option value %s", tok[i]);
    if (i + 1 < nt && tok[i + 1][0] != '-') {
      PetscCall(PetscOptionsSetValue(tok[i], tok[i + 1]));
      i++;
    } else PetscCall(PetscOptionsSetValue(tok[i], NULL));
  }
  return PETSC_SUCCESS;
}

PetscErrorCode PetscOptionsGetString(const char name[], char *value, size_t len, PetscBool *set)
{
  *set = PETSC_FALSE;
  for (int i = 0; i < g_nopt; i++) {
    if (strcmp(g_opt[i].name, name)) continue;
    snprintf(value, len, "%s", g_opt[i].value);
    *set = PETSC_TRUE;
    break;
  }
  return PETSC_SUCCESS;
}

/* PetscOptionsGetBool: an empty value means true; 1/0, true/false, yes/no */
PetscErrorCode PetscOptionsGetBool(const char name[], PetscBool *flg, PetscBool *set)
{
  char v[128];
  PetscBool s;
  PetscCall(PetscOptionsGetString(name, v, sizeof(v), &s));
  if (set) *set = s;
  if (!s) return PETSC_SUCCESS;
  if (!v[0] || !strcmp(v, "1") || !strcasecmp(v, "true") || !strcasecmp(v, "yes")) *flg = PETSC_TRUE;
  else if (!strcmp(v, "0") || !strcasecmp(v, "false") || !strcasecmp(v, "no")) *flg = PETSC_FALSE;
  else SETERRQ(PETSC_ERR_ARG_WRONG, "option %s: %s is not a boolean", name, v);
  return PETSC_SUCCESS;
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
  PetscCall(PetscOptionsInsertString(getenv("PETSC_OPTIONS")));
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
