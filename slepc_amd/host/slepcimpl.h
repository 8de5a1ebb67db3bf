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

/* ---------------------------------------------------------------- Vec */
struct _p_Vec {
  PetscInt n;
  SlepcAMDPrecision prec;
  void *d;               /* device storage */
  PetscBool owns;        /* device storage allocated by this Vec */
};
size_t SlepcPrecSize(SlepcAMDPrecision p);
PetscErrorCode VecCreateHIPWithArray(PetscInt n, SlepcAMDPrecision prec, void *d, Vec *v);

/* ---------------------------------------------------------------- Mat */
typedef struct {
  PetscErrorCode (*mult)(Mat, Vec, Vec);
  PetscErrorCode (*multtranspose)(Mat, Vec, Vec);
  PetscErrorCode (*destroy)(Mat);
} MatOps;
struct _p_Mat {
  char type[SLEPC_NAME_LEN];
  MatOps ops;
  PetscInt m, n;
  void *data;
  PetscBool gate;        /* hipcsr: skip the product once an expansion broke down */
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
} Mat_HIPCSR;
/* the FP32-valued copy of a hipcsr matrix, made on first use */
PetscErrorCode MatHIPCSRGetF32(Mat A, ks_csr *out);
PetscErrorCode MatCreateSeqDenseView(PetscInt m, PetscInt n, PetscScalar *data, PetscInt lda, Mat *A);

/* ---------------------------------------------------------------- BV */
typedef struct _BVOps {
  PetscErrorCode (*multvec)(BV, PetscScalar, PetscScalar, Vec, PetscScalar *);
  PetscErrorCode (*multinplace)(BV, Mat, PetscInt, PetscInt);
  PetscErrorCode (*dotvec)(BV, Vec, PetscScalar *);
  PetscErrorCode (*scale)(BV, PetscInt, PetscScalar);
  PetscErrorCode (*norm)(BV, PetscInt, NormType, PetscReal *);
  PetscErrorCode (*copycolumn)(BV, PetscInt, PetscInt);
  PetscErrorCode (*getcolumn)(BV, PetscInt, Vec *);
  PetscErrorCode (*restorecolumn)(BV, PetscInt, Vec *);
  PetscErrorCode (*setrandomcolumn)(BV, PetscInt);
  PetscErrorCode (*getbuffer)(BV);       /* refresh bv->buffer (host mirror) */
  PetscErrorCode (*destroy)(BV);
  PetscErrorCode (*create)(BV);          /* deferred until sizes are known */
  /* Whole-column orthonormalization (all CGS passes, DGKS decisions, lindep and
     the scaling) done inside the type.  Optional: the interface falls back to
     the pass-by-pass loop when it is NULL.  `async` leaves the outputs on the
     device (norm/lindep may be NULL); otherwise they are returned. */
  PetscErrorCode (*orthonormalizecolumn)(BV, PetscInt j, PetscBool scale, PetscBool safe, PetscBool async,
                                         PetscReal *norm, PetscBool *lindep);
  /* Whole Arnoldi/Lanczos expansion k..m split in two (optional): begin enqueues
     every step on the device without waiting, end waits once and returns the
     breakdown position, beta and the new coefficient columns (host mirror). */
  PetscErrorCode (*matkrylovbegin)(BV, Mat, PetscInt k, PetscInt m);
  PetscErrorCode (*matkrylovend)(BV, PetscInt k, PetscInt *m, PetscReal *beta, PetscBool *breakdown);
  /* G(0:ye-ys, 0:xe-xs) = Y(:, ys:ye)^T X(:, xs:xe) into host memory (leading
     dimension ldg): the block dot under BVDot, BVOrthogonalize, BVMatProject */
  PetscErrorCode (*gram)(BV X, PetscInt xs, PetscInt xe, BV Y, PetscInt ys, PetscInt ye, PetscScalar *G, PetscInt ldg);
} BVOps;

struct _p_BV {
  char type[SLEPC_NAME_LEN];
  BVOps ops;
  void *data;
  PetscInt n, N, m;          /* local/global rows, columns (without constraints) */
  PetscInt l, k, nc, ld;     /* active columns, constraints, leading dimension */
  BVOrthogType orthog_type;
  BVOrthogRefineType orthog_ref;
  BVOrthogBlockType orthog_block;   /* method of BVOrthogonalize */
  PetscReal orthog_eta;
  PetscReal deftol;
  PetscScalar *buffer;       /* host mirror of the (nc+m) x m coefficient buffer */
  Vec cv[2];                 /* column views */
  PetscInt ci[2];
  unsigned long long rseed, rcount;
  SlepcAMDPrecision prec;
  ks_ctx ctx;
  int64_t gs_passes;         /* executed CGS passes */
  /* results of an expansion done synchronously by BVMatKrylovBegin */
  PetscInt kry_m;
  PetscReal kry_beta;
  PetscBool kry_breakdown;
  /* grouped expansions (serving): the steps k..m-1 are enqueued by the server,
     step kry_j next, while kry_pending */
  PetscBool grouped, kry_pending;
  PetscInt kry_k, kry_j;
  ks_csr kry_csr;
  /* the status and the new coefficient columns were copied to pinned host
     memory in stream order at the end of the expansion */
  PetscBool kry_posted;
  /* replay each Arnoldi step of an expansion from a captured graph */
  PetscBool step_graphs;
};
PetscErrorCode BVOrthogonalizeColumn_Safe(BV bv, PetscInt j, PetscReal *norm, PetscBool *lindep);
PetscErrorCode BVMatKrylovBegin(BV V, Mat A, PetscInt k, PetscInt m);
PetscErrorCode BVMatKrylovEnd(BV V, PetscInt k, PetscInt *m, PetscReal *beta, PetscBool *breakdown);
PetscErrorCode BVCopyHessenberg(BV V, Mat H, PetscInt k, PetscInt m);
PetscErrorCode BVCopyTridiagonal(BV V, Mat T, PetscInt k, PetscInt m);
PetscErrorCode BVHIPKSGroupMember(BV V, ks_group_member *mem);
PetscErrorCode BVHIPKSGroupAdvance(BV V, PetscBool *finished);
PetscErrorCode BVHIPKSPostExpansion(BV V);

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
};
struct _p_KSP {              /* the inner solver of STSINVERT (device PCG) */
  char type[SLEPC_NAME_LEN];
  struct _p_PC pc;
  PetscReal rtol, abstol;
  PetscInt maxits, n;
  double *d_dinv, *d_work;   /* Jacobi diagonal, r z p q + scalars */
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

typedef enum { KS_PHASE_IDLE = 0, KS_PHASE_EXPANDING = 1, KS_PHASE_DONE = 2, KS_PHASE_MIXED32 = 3 } KSPhase;

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
  /* two-sided variant: A^T and the oblique projection M = W^T V (ncv x ncv) */
  Mat AT;
  PetscScalar *M;
} EPS_KRYLOVSCHUR;

PetscErrorCode EPSKrylovConvergence(EPS eps, PetscBool getall, PetscInt kini, PetscInt nits, PetscReal beta,
                                    PetscReal betat, PetscReal corrf, PetscInt *kout);
PetscErrorCode EPSGetLeftStartVector(EPS eps, PetscInt i, PetscBool *breakdown);
PetscErrorCode EPSGetStartVector(EPS eps, PetscInt i, PetscBool *breakdown);
PetscErrorCode EPSComputeVectors(EPS eps);
PetscErrorCode EPSMatMult(EPS eps, Mat A, Vec x, Vec y);

#endif
