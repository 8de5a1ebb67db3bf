/*
 * slepcimpl.h -- private object layouts and helpers of libslepc_amd.so.
 *
 * The layouts follow the reference's private headers where the Krylov-Schur
 * path reads them: struct _p_BV / struct _BVOps (include/slepc/private/bvimpl.h:25-113),
 * struct _p_DS (include/slepc/private/dsimpl.h), struct _p_EPS / struct _EPSOps
 * (include/slepc/private/epsimpl.h:26-39 and following), EPS_KRYLOVSCHUR
 * (src/eps/impls/krylov/krylovschur/krylovschur.h:76-105).
 */
#ifndef SLEPCIMPL_H
#define SLEPCIMPL_H

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/ks_hip.h"
#include "../../include/slepc_amd.h"

#define PETSC_MACHINE_EPSILON 2.220446049250313e-16
#define PETSC_MAX_REAL 1.7976931348623157e308
#define SLEPC_DEFAULT_TOL 1e-8
#define SLEPC_NAME_LEN 32
#define BV_NOCOL (-2147483647 - 1)   /* free slot of the column views */

/* ---------------------------------------------------------------- errors */
PetscErrorCode SlepcError(int line, const char *func, const char *file, PetscErrorCode code, const char *fmt, ...);
#define SETERRQ(code, ...) return SlepcError(__LINE__, __func__, __FILE__, (code), __VA_ARGS__)
#define PetscCheck(cond, code, ...) do { if (!(cond)) SETERRQ((code), __VA_ARGS__); } while (0)
#define PetscCall(x) do { PetscErrorCode _ierr = (x); if (_ierr) { \
      SlepcError(__LINE__, __func__, __FILE__, _ierr, NULL); return _ierr; } } while (0)
#define KSCall(x) do { int _kerr = (x); if (_kerr) \
      SETERRQ(PETSC_ERR_GPU, "device library: %s (%d) in %s", ks_error_string(_kerr), _kerr, #x); } while (0)
#define PetscMax(a, b) ((a) > (b) ? (a) : (b))
#define PetscMin(a, b) ((a) < (b) ? (a) : (b))
#define PetscAbsReal(a) fabs(a)
#define SlepcAbsEigenvalue(x, y) hypot((x), (y))

PetscErrorCode PetscMallocZero(size_t bytes, void *ptr);
#define PetscFree(p) do { free(p); (p) = NULL; } while (0)
double SlepcWallTime(void);

/* registry of (name, constructor) */
typedef struct {
  char name[SLEPC_NAME_LEN];
  void *create;
} SlepcRegEntry;
typedef struct {
  SlepcRegEntry entry[16];
  int n;
} SlepcFunctionList;
PetscErrorCode SlepcFunctionListAdd(SlepcFunctionList *fl, const char *name, void *create);
void *SlepcFunctionListFind(SlepcFunctionList *fl, const char *name);

/* the device context objects bind to (default: the one made at initialization) */
ks_ctx SlepcAMDContext(void);
ks_ctx SlepcAMDSetCurrentContext(ks_ctx ctx);   /* NULL restores the default; returns the previous */
PetscErrorCode SlepcCheckInitialized(void);

/* ---------------------------------------------------------------- sorting criterion (slepcsc.c) */
typedef struct {
  int which;
  PetscScalar target;
  PetscBool map;             /* compare sigma + 1/theta (SlepcMap_ST of a shift-and-invert ST) */
  PetscScalar sigma;
} SlepcSC;
int SlepcSCCompare(const SlepcSC *sc, PetscScalar ar, PetscScalar ai, PetscScalar br, PetscScalar bi);
void SlepcSortEigenvalues(const SlepcSC *sc, PetscInt n, const PetscScalar *eigr, const PetscScalar *eigi,
                          PetscInt *perm);

/* ---------------------------------------------------------------- objects
 * PETSc's object header (petsc/private/petscimpl.h, outside the reference):
 * class and type names, a state counter, and the composed functions through
 * which a type offers methods beyond its ops table
 * (PetscObjectComposeFunction / PetscObjectQueryFunction). */
typedef void (*PetscVoidFn)(void);
typedef int64_t PetscObjectId;
typedef int64_t PetscObjectState;
#define SLEPC_MAX_COMPOSED 8
typedef struct _p_PetscObject {
  char class_name[SLEPC_NAME_LEN];
  char type_name[SLEPC_NAME_LEN];
  PetscObjectId id;
  PetscObjectState state;
  int nfn;
  char fname[SLEPC_MAX_COMPOSED][64];
  PetscVoidFn fn[SLEPC_MAX_COMPOSED];
} _p_PetscObject;
typedef struct _p_PetscObject *PetscObject;
#define PETSCHEADER(ObjectOps) \
  _p_PetscObject hdr;          \
  ObjectOps ops[1]
PetscErrorCode PetscHeaderInit(PetscObject obj, const char *class_name);
PetscErrorCode PetscObjectComposeFunction_Private(PetscObject obj, const char name[], PetscVoidFn fn);
PetscErrorCode PetscObjectQueryFunction_Private(PetscObject obj, const char name[], PetscVoidFn *fn);
#define PetscObjectComposeFunction(obj, name, f) \
  PetscObjectComposeFunction_Private((PetscObject)(obj), (name), (PetscVoidFn)(f))
#define PetscObjectQueryFunction(obj, name, fptr) \
  PetscObjectQueryFunction_Private((PetscObject)(obj), (name), (PetscVoidFn *)(fptr))

/* the opaque PETSc types the reference's BV layout names but this path never uses */
typedef struct _p_PetscLayout *PetscLayout;
typedef struct _p_IS *IS;
typedef struct _p_PetscViewer *PetscViewer;
typedef struct _p_PetscOptionItems PetscOptionItems;

/* PetscRandom: the counter generator of ks_bv_random; every draw advances the
   counter, draw c of seed s uses the stream s + 1000003 c */
struct _p_PetscRandom {
  uint64_t seed, count;
};
typedef struct _p_PetscRandom *PetscRandom;
PetscErrorCode PetscRandomCreate(PetscRandom *r);
PetscErrorCode PetscRandomDestroy(PetscRandom *r);

/* ---------------------------------------------------------------- options database
 * PETSc's, reduced to what the types here read: "-name value" pairs from the
 * PETSC_OPTIONS environment variable (read at SlepcInitialize) and from
 * PetscOptionsInsertString. */
PetscErrorCode PetscOptionsGetBool(const char name[], PetscBool *flg, PetscBool *set);
PetscErrorCode PetscOptionsGetString(const char name[], char *value, size_t len, PetscBool *set);

/* ---------------------------------------------------------------- Vec */
/* where the current values of a vector with a host copy are (PetscOffloadMask) */
typedef enum { PETSC_OFFLOAD_UNALLOCATED = 0, PETSC_OFFLOAD_GPU = 1, PETSC_OFFLOAD_BOTH = 3 } PetscOffloadMask;
struct _p_Vec {
  PetscInt n;
  SlepcAMDPrecision prec;
  void *d;               /* device storage */
  PetscBool owns;        /* device storage allocated by this Vec */
  PetscScalar *h;        /* pinned host copy (FP64 vectors), made by the first VecGetArray */
  PetscOffloadMask mask;
};
size_t SlepcPrecSize(SlepcAMDPrecision p);
PetscErrorCode VecCreateHIPWithArray(PetscInt n, SlepcAMDPrecision prec, void *d, Vec *v);
/* host access (FP64 vectors): the device values are copied to the host copy when
   they are newer; VecRestoreArray copies the host values back */
PetscErrorCode VecGetArray(Vec v, PetscScalar **a);
PetscErrorCode VecRestoreArray(Vec v, PetscScalar **a);
PetscErrorCode VecGetArrayRead(Vec v, const PetscScalar **a);
PetscErrorCode VecRestoreArrayRead(Vec v, const PetscScalar **a);
PetscErrorCode VecHIPGetArrayRead(Vec v, const void **d);
/* the host copy, made current by a copy the caller enqueued and waited for itself */
PetscErrorCode VecHIPGetHostCopy(Vec v, PetscScalar **h);
PetscErrorCode VecHIPSetHostCurrent(Vec v);
PetscErrorCode VecSetRandom(Vec v, PetscRandom r);

/* ---------------------------------------------------------------- Mat */
struct _MatOps {        /* the members of PETSc's struct _MatOps this path uses */
  PetscErrorCode (*mult)(Mat, Vec, Vec);
  PetscErrorCode (*multtranspose)(Mat, Vec, Vec);
  PetscErrorCode (*destroy)(Mat);
};
struct _p_Mat {
  PETSCHEADER(struct _MatOps);
  PetscInt m, n;
  void *data;
};
typedef struct {         /* MATSEQDENSE, host column-major */
  PetscScalar *v;
  PetscInt lda;
  PetscBool owns;
} Mat_SeqDense;
typedef struct {         /* MATHIPCSR */
  ks_csr csr;
  ks_csr csr32;          /* FP32 copy of FP64 values for FP32 vectors (built on demand) */
  ks_csr csrT;           /* A^T for MatMultTranspose (built on the device on demand) */
  SlepcAMDPrecision vprec;
  PetscBool gate;        /* skip the product once an expansion broke down */
} Mat_HIPCSR;
#define MatTypeIs(A, t) (!strcmp((A)->hdr.type_name, (t)))
/* the FP32-valued copy of a hipcsr matrix, made on first use */
PetscErrorCode MatHIPCSRGetF32(Mat A, ks_csr *out);
PetscErrorCode MatCreateSeqDenseView(PetscInt m, PetscInt n, PetscScalar *data, PetscInt lda, Mat *A);

/* ---------------------------------------------------------------- BV
 * struct _BVOps and struct _p_BV have exactly the members of the reference's
 * include/slepc/private/bvimpl.h:25-61 and :63-113, in the same order: a BV type
 * written against them (BVHIPKS) drops into an unmodified SLEPc.  The type's own
 * state lives behind bv->data; what it offers beyond the ops table goes through
 * composed functions. */
typedef enum { BV_MATMULT_VECS = 0, BV_MATMULT_MAT = 1, BV_MATMULT_MAT_SAVE = 2 } BVMatMultType;

typedef struct _BVOps *BVOps;

struct _BVOps {
  PetscErrorCode (*mult)(BV, PetscScalar, PetscScalar, BV, Mat);
  PetscErrorCode (*multvec)(BV, PetscScalar, PetscScalar, Vec, PetscScalar *);
  PetscErrorCode (*multinplace)(BV, Mat, PetscInt, PetscInt);
  PetscErrorCode (*multinplacetrans)(BV, Mat, PetscInt, PetscInt);
  PetscErrorCode (*dot)(BV, BV, Mat);
  PetscErrorCode (*dotvec)(BV, Vec, PetscScalar *);
  PetscErrorCode (*dotvec_local)(BV, Vec, PetscScalar *);
  PetscErrorCode (*dotvec_begin)(BV, Vec, PetscScalar *);
  PetscErrorCode (*dotvec_end)(BV, Vec, PetscScalar *);
  PetscErrorCode (*scale)(BV, PetscInt, PetscScalar);
  PetscErrorCode (*norm)(BV, PetscInt, NormType, PetscReal *);
  PetscErrorCode (*norm_local)(BV, PetscInt, NormType, PetscReal *);
  PetscErrorCode (*norm_begin)(BV, PetscInt, NormType, PetscReal *);
  PetscErrorCode (*norm_end)(BV, PetscInt, NormType, PetscReal *);
  PetscErrorCode (*normalize)(BV, PetscScalar *);
  PetscErrorCode (*matmult)(BV, Mat, BV);
  PetscErrorCode (*copy)(BV, BV);
  PetscErrorCode (*copycolumn)(BV, PetscInt, PetscInt);
  PetscErrorCode (*resize)(BV, PetscInt, PetscBool);
  PetscErrorCode (*getcolumn)(BV, PetscInt, Vec *);
  PetscErrorCode (*restorecolumn)(BV, PetscInt, Vec *);
  PetscErrorCode (*getarray)(BV, PetscScalar **);
  PetscErrorCode (*restorearray)(BV, PetscScalar **);
  PetscErrorCode (*getarrayread)(BV, const PetscScalar **);
  PetscErrorCode (*restorearrayread)(BV, const PetscScalar **);
  PetscErrorCode (*restoresplit)(BV, BV *, BV *);
  PetscErrorCode (*restoresplitrows)(BV, IS, IS, BV *, BV *);
  PetscErrorCode (*gramschmidt)(BV, PetscInt, Vec, PetscBool *, PetscScalar *, PetscScalar *, PetscReal *, PetscReal *);
  PetscErrorCode (*getmat)(BV, Mat *);
  PetscErrorCode (*restoremat)(BV, Mat *);
  PetscErrorCode (*duplicate)(BV, BV);
  PetscErrorCode (*create)(BV);
  PetscErrorCode (*setfromoptions)(BV, PetscOptionItems *);
  PetscErrorCode (*view)(BV, PetscViewer);
  PetscErrorCode (*destroy)(BV);
};

struct _p_BV {
  PETSCHEADER(struct _BVOps);
  /*------------------------- User parameters --------------------------*/
  PetscLayout        map;          /* layout of columns */
  VecType            vtype;        /* vector type */
  PetscInt           n, N;         /* dimensions of vectors (local, global) */
  PetscInt           m;            /* number of vectors */
  PetscInt           l;            /* number of leading columns */
  PetscInt           k;            /* number of active columns */
  PetscInt           nc;           /* number of constraints */
  PetscInt           ld;           /* leading dimension */
  BVOrthogType       orthog_type;  /* the method of vector orthogonalization */
  BVOrthogRefineType orthog_ref;   /* refinement method */
  PetscReal          orthog_eta;   /* refinement threshold */
  BVOrthogBlockType  orthog_block; /* the method of block orthogonalization */
  Mat                matrix;       /* inner product matrix */
  PetscBool          indef;        /* matrix is indefinite */
  BVMatMultType      vmm;          /* version of matmult operation */
  PetscBool          rrandom;      /* reproducible random vectors */
  PetscReal          deftol;       /* tolerance for BV_SafeSqrt */

  /*---------------------- Cached data and workspace -------------------*/
  Vec                buffer;       /* buffer vector used in orthogonalization */
  Mat                Abuffer;      /* auxiliary seqdense matrix that wraps the buffer */
  Vec                Bx;           /* result of matrix times a vector x */
  PetscObjectId      xid;          /* object id of vector x */
  PetscObjectState   xstate;       /* state of vector x */
  Vec                cv[2];        /* column vectors obtained with BVGetColumn() */
  PetscInt           ci[2];        /* column indices of obtained vectors */
  PetscObjectState   st[2];        /* state of obtained vectors */
  PetscObjectId      id[2];        /* object id of obtained vectors */
  PetscScalar        *h, *c;       /* orthogonalization coefficients */
  Vec                omega;        /* signature matrix values for indefinite case */
  PetscBool          defersfo;     /* deferred call to setfromoptions */
  BV                 cached;       /* cached BV to store result of matrix times BV */
  PetscObjectState   bvstate;      /* state of BV when BVApplyMatrixBV() was called */
  BV                 L, R;         /* BV objects obtained with BVGetSplit/Rows() */
  PetscObjectState   lstate, rstate; /* state of L and R when BVGetSplit/Rows() was called */
  PetscInt           lsplit;       /* value of l when BVGetSplit() was called (-1 if BVGetSplitRows()) */
  PetscInt           issplit;      /* !=0 if BV is from split (1=left, 2=right, -1=top, -2=bottom) */
  BV                 splitparent;  /* my parent if I am a split BV */
  PetscRandom        rand;         /* random number generator */
  Mat                Acreate;      /* matrix given at BVCreateFromMat() */
  Mat                Aget;         /* matrix returned for BVGetMat() */
  PetscBool          cuda;         /* true if NVIDIA GPU must be used */
  PetscBool          hip;          /* true if AMD GPU must be used */
  PetscBool          sfocalled;    /* setfromoptions has been called */
  PetscScalar        *work;
  PetscInt           lwork;
  void               *data;
};
#define BVTypeName(bv) (((PetscObject)(bv))->type_name)

/* BVGetBufferVec (bvbasic.c:755-786): the (nc+m) x m coefficient buffer, made on first use */
PetscErrorCode BVGetBufferVec(BV bv, Vec *buffer);
PetscErrorCode BV_CleanCoefficients(BV bv, PetscInt j, PetscScalar *h);
PetscErrorCode BV_SetValue(BV bv, PetscInt j, PetscInt k, PetscScalar *h, PetscScalar value);
PetscErrorCode BV_StoreCoefficients(BV bv, PetscInt j, PetscScalar *h, PetscScalar *dest);
PetscErrorCode BVOrthogonalizeColumn_Safe(BV bv, PetscInt j, PetscReal *norm, PetscBool *lindep);
PetscErrorCode BVMatKrylovBegin(BV V, Mat A, PetscInt k, PetscInt m);
PetscErrorCode BVMatKrylovEnd(BV V, PetscInt k, PetscInt *m, PetscReal *beta, PetscBool *breakdown);
PetscErrorCode BVCopyHessenberg(BV V, Mat H, PetscInt k, PetscInt m);
PetscErrorCode BVCopyTridiagonal(BV V, Mat T, PetscInt k, PetscInt m);
PetscErrorCode BVMatKrylov(BV V, Mat A, PetscInt k, PetscInt *m, PetscReal *beta, PetscBool *breakdown);
PetscErrorCode BVResizeDiscard_Private(BV bv, PetscInt n, PetscInt m);
PetscErrorCode BVSetRandomSeed(BV bv, uint64_t seed);
PetscErrorCode BVGetRandomSeed(BV bv, uint64_t *seed);

/* The one hook through which a BV type offers a device-resident Krylov
 * expansion (no reference counterpart).  A type composes
 * "BVGetKrylovOps_C" -> PetscErrorCode f(BV, const BVKrylovOps **); a NULL table
 * means the type (or this object) leaves orthonormalization to the per-pass
 * gramschmidt op and the expansion to the reference's per-step loop.
 *   orthonormalizecolumn  all CGS passes, the DGKS decisions, lindep and the
 *                         1/norm scaling of one column on the device
 *                         (BVOrthogonalizeGS bvorthog.c:145-217 + the scaling of
 *                         BVOrthonormalizeColumn bvorthog.c:416-422);
 *   matkrylovbegin/end    the steps k..m-1 of BVMatArnoldi/BVMatLanczos
 *                         (bvkrylov.c:88-97) enqueued at once, then one wait
 *                         for the breakdown position, beta and the new
 *                         coefficient columns (in the buffer's host copy). */
typedef struct {
  PetscErrorCode (*orthonormalizecolumn)(BV, PetscInt j, PetscBool scale, PetscBool safe, PetscReal *norm,
                                         PetscBool *lindep);
  PetscErrorCode (*matkrylovbegin)(BV, Mat, PetscInt k, PetscInt m);
  PetscErrorCode (*matkrylovend)(BV, PetscInt k, PetscInt *m, PetscReal *beta, PetscBool *breakdown);
} BVKrylovOps;
PetscErrorCode BVHIPKSGetDeviceArray_Private(BV bv, double **d, PetscInt *ld, ks_ctx *ctx);
PetscErrorCode BVGetKrylovOps(BV bv, const BVKrylovOps **ops);

/* BVHIPKS functions used by the solver and the serving loop */
PetscErrorCode BVHIPKSGetContext(BV bv, ks_ctx *ctx);
PetscErrorCode BVHIPKSGetPasses(BV bv, int64_t *passes);
PetscErrorCode BVHIPKSResetPasses(BV bv);
PetscErrorCode BVHIPKSSetGrouped(BV bv, PetscBool grouped);
PetscErrorCode BVHIPKSGroupPending(BV bv, PetscBool *pending);
PetscErrorCode BVHIPKSGroupMember(BV V, ks_group_member *mem);
PetscErrorCode BVHIPKSGroupAdvance(BV V, PetscBool *finished);
PetscErrorCode BVHIPKSPostExpansion(BV V);
PetscErrorCode BVHIPKSGetStepGraphs(BV bv, PetscBool *flg);
SlepcAMDPrecision BVPrecision(BV bv);

/* ---------------------------------------------------------------- DS */
typedef struct {
  PetscErrorCode (*solve)(DS, PetscScalar *, PetscScalar *);
  PetscErrorCode (*sort)(DS, PetscScalar *, PetscScalar *, PetscScalar *, PetscScalar *, PetscInt *);
  PetscErrorCode (*update)(DS);
  PetscErrorCode (*vectors)(DS, DSMatType, PetscInt *, PetscReal *);
  PetscErrorCode (*truncate)(DS, PetscInt, PetscBool);
  PetscErrorCode (*gettruncatesize)(DS, PetscInt, PetscInt, PetscInt *);
  PetscErrorCode (*transharm)(DS, PetscScalar, PetscReal, PetscBool, PetscScalar *, PetscReal *);
} DSOps;
struct _p_DS {
  char type[SLEPC_NAME_LEN];
  DSOps ops;
  PetscInt ld, n, l, k, t;
  DSStateType state;
  PetscBool compact, extrarow;
  PetscScalar *mat[DS_NUM_MAT];   /* ld x ld (T: ld x 3) */
  PetscScalar *work;
  PetscReal *rwork;
  int *iwork;
  PetscInt *perm;
  SlepcSC sc;
  Mat matview[DS_NUM_MAT];
  PetscScalar *wr2, *wi2;         /* DSNHEPTS: eigenvalues of B */
};
PetscErrorCode DSCreate_HEP(DS ds);
PetscErrorCode DSCreate_NHEP(DS ds);
PetscErrorCode DSCreate_NHEPTS(DS ds);

/* ---------------------------------------------------------------- ST */
struct _p_PC {
  char type[SLEPC_NAME_LEN];
  PetscBool set;             /* PCSetType was called (else KSPSetType picks the PC that goes with it) */
};
struct _p_KSP {              /* the inner solver of STSINVERT (device LU, CG or BiCGStab) */
  char type[SLEPC_NAME_LEN];
  struct _p_PC pc;
  PetscReal rtol, abstol;
  PetscInt maxits, n;
  double *d_dinv, *d_work;   /* Jacobi diagonal, the solver's vectors + its state */
  ks_ilu ilu;                /* PCILU: the ILU(0) factors of A - sigma I */
  ks_lu lu;                  /* KSPPREONLY + PCLU: the LU factors of A - sigma I */
  size_t nwork;              /* doubles in d_work */
  int64_t its_total, nsolves;
};
struct _p_ST {
  char type[SLEPC_NAME_LEN];
  Mat A;
  PetscScalar sigma, defsigma;
  PetscBool sigma_set, setup;
  Mat P;                     /* STSINVERT: the operator (A - sigma I)^-1 */
  KSP ksp;
};
PetscErrorCode STSetUp(ST st);
PetscBool STIsSinvert(ST st);

/* ---------------------------------------------------------------- EPS */
typedef struct {
  PetscErrorCode (*setup)(EPS);
  PetscErrorCode (*solve)(EPS);
  PetscErrorCode (*solvebegin)(EPS);               /* optional resumable form of solve */
  PetscErrorCode (*solvestep)(EPS, PetscBool *);
  PetscErrorCode (*computevectors)(EPS);
  PetscErrorCode (*destroy)(EPS);
} EPSOps;
typedef enum { EPS_STATE_INITIAL = 0, EPS_STATE_SETUP = 1, EPS_STATE_SOLVED = 2, EPS_STATE_EIGENVECTORS = 3 } EPSStateType;

struct _p_EPS {
  char type[SLEPC_NAME_LEN];
  EPSOps ops;
  void *data;
  Mat A;
  ST st;
  BV V;
  BV W;                      /* left basis of the two-sided variant */
  DS ds;
  PetscInt n, nev, ncv, mpd, max_it, nconv, its, nini, nds;
  PetscReal tol;
  EPSWhich which;
  PetscScalar target;
  EPSExtraction extraction;
  PetscBool twosided;
  EPSProblemType problem_type;
  PetscBool ishermitian, isgeneralized;
  EPSConvergedReason reason;
  EPSStateType state;
  PetscScalar *eigr, *eigi;
  PetscReal *errest;
  PetscInt *perm;
  SlepcSC sc;
  Vec *IS, *defl;
  EPSMonitorFn monitor;
  void *monitorctx;
  Vec work[5];
  int64_t nmatmult;
  PetscReal t_ds, t_solve;
  double t_begin;
  PetscBool stepping;        /* between EPSSolveBegin and the last EPSSolveStep */
  BV stepV;                  /* the basis whose device work the next step collects */
};

typedef enum { KS_PHASE_IDLE = 0, KS_PHASE_EXPANDING = 1, KS_PHASE_DONE = 2, KS_PHASE_MIXED32 = 3,
               KS_PHASE_TS_RIGHT = 4, KS_PHASE_TS_LEFT = 5 } KSPhase;

typedef struct {            /* EPS_KRYLOVSCHUR (krylovschur.h:76-105), default variant fields */
  PetscReal keep;
  PetscBool lock;
  /* restart-loop state kept between the halves of an iteration */
  KSPhase phase;
  PetscInt l, nv;
  /* mixed-precision form: FP32 basis down to mixed_tol, then FP64 restarts */
  PetscReal mixed_tol;       /* <= 0: off */
  PetscInt  its32, nconv32;
  int64_t   nmatmult32, passes32;
  PetscReal t32;
  EPS       e32;             /* the FP32 solve while the first phase runs */
  double    t32_start;
  /* harmonic extraction: the translation vector g (ncv + 1 entries) */
  PetscBool harmonic;
  PetscScalar *g;
  /* two-sided variant: A^T and the oblique projection M = W^T V (ncv x ncv);
     while its restart loop runs: the operator, a view of M, the right
     expansion's results, and the left expansion in flight (tkry_*) */
  Mat AT;
  PetscScalar *M;
  Mat ts_Op, ts_Mm;
  PetscReal ts_beta;
  PetscBool ts_breakdown;
  const BVKrylovOps *tkry_ext;
  PetscInt tkry_m;
  PetscReal tkry_beta;
  PetscBool tkry_breakdown;
  /* the expansion in flight: through the basis' Krylov extension, or (NULL)
     already run by the per-step loop, with its results here */
  const BVKrylovOps *kry_ext;
  PetscInt kry_m;
  PetscReal kry_beta;
  PetscBool kry_breakdown;
  /* device-path settings applied to the basis at set-up */
  PetscBool step_graphs, grouped;
  /* first phase of the mixed form: hand the kept subspace over at convergence
     as q1, the start of its Arnoldi form in the kept columns 0..ho_m-1 */
  PetscBool handover;
  PetscInt ho_m, ho_nconv;
  PetscScalar *ho_q1;
} EPS_KRYLOVSCHUR;
PetscErrorCode EPSKrylovSchurSetGrouped_Private(EPS eps, PetscBool flg);

PetscErrorCode EPSKrylovConvergence(EPS eps, PetscBool getall, PetscInt kini, PetscInt nits, PetscReal beta,
                                    PetscReal betat, PetscReal corrf, PetscInt *kout);
PetscErrorCode EPSGetLeftStartVector(EPS eps, PetscInt i, PetscBool *breakdown);
PetscErrorCode EPSGetStartVector(EPS eps, PetscInt i, PetscBool *breakdown);
PetscErrorCode EPSComputeVectors(EPS eps);
PetscErrorCode EPSMatMult(EPS eps, Mat A, Vec x, Vec y);

/* Leading dimension of a device basis of n rows with esz-byte entries: n
   rounded up to 64 entries (512-byte columns for FP64), plus 2 KB when the
   column stride would be a multiple of 4 KB.  Strides of 4 KB multiples map
   the same row of every column to the same HBM channels, and the sweeps read
   a dozen or more columns at one row offset at once: on the 128^3 basis
   (16 MB columns) the padding took the Gram-Schmidt sweeps from 101-115 to
   92-106 us. */
static inline PetscInt BVHIPKSLeadingDimension(PetscInt n, size_t esz)
{
  PetscInt ld = ((n + 63) / 64) * 64;
  if (((size_t)ld * esz) % 4096 == 0) ld += (PetscInt)(2048 / esz);
  return ld;
}

#endif
