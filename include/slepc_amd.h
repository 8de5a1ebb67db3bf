/*
 * slepc_amd.h -- SLEPc-shaped host API of libslepc_amd.so (C, no HIP types).
 *
 * The object model mirrors SLEPc 3.22: EPS (solver), BV (basis vectors), DS
 * (dense projected problem), ST (spectral transformation, shift with sigma=0),
 * plus the PETSc Mat/Vec pieces the Krylov-Schur path touches.  Types are
 * selected through the same registries: EPSRegister / BVRegister / MatRegister
 * / DSRegister with EPSSetType / BVSetType / MatSetType / DSSetType.
 *
 * The MI355X pieces are the plugins registered at SlepcInitialize():
 *   BVHIPKS   ("hipks")   BVCreate_HIPKS  -- basis vectors resident in HBM, fused
 *                                            CGS+DGKS, FP64-MFMA restart GEMM
 *   MATHIPCSR ("hipcsr")  MatCreate_HIPCSR -- device CSR with the adaptive SpMV
 * Function names, argument meaning and error behaviour follow the reference;
 * every function returns a PetscErrorCode (0 on success).
 */
#ifndef SLEPC_AMD_H
#define SLEPC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int    PetscErrorCode;
typedef int    PetscInt;
typedef double PetscScalar;
typedef double PetscReal;
typedef int    PetscBool;
#define PETSC_TRUE  1
#define PETSC_FALSE 0
#define PETSC_DETERMINE (-1)
#define PETSC_DECIDE    (-1)
#define PETSC_CURRENT   (-2)
#define PETSC_DEFAULT   (-2)

/* error codes (PETSc numbering) */
#define PETSC_SUCCESS               0
#define PETSC_ERR_MEM               55
#define PETSC_ERR_SUP               56
#define PETSC_ERR_ORDER             58
#define PETSC_ERR_LIB               76
#define PETSC_ERR_PLIB              77
#define PETSC_ERR_ARG_SIZ           60
#define PETSC_ERR_ARG_WRONG         62
#define PETSC_ERR_ARG_OUTOFRANGE    63
#define PETSC_ERR_FP                72
#define PETSC_ERR_ARG_WRONGSTATE    73
#define PETSC_ERR_ARG_INCOMP        75
#define PETSC_ERR_ARG_NULL          85
#define PETSC_ERR_USER_INPUT        90
#define PETSC_ERR_CONV_FAILED       91
#define PETSC_ERR_NOT_CONVERGED     PETSC_ERR_CONV_FAILED
#define PETSC_ERR_GPU               97
#define PETSC_ERR_FILE_OPEN         65
#define PETSC_ERR_FILE_READ         66
#define PETSC_ERR_FILE_UNEXPECTED   79

typedef struct _p_Vec *Vec;
typedef struct _p_Mat *Mat;
typedef struct _p_BV  *BV;
typedef struct _p_DS  *DS;
typedef struct _p_ST  *ST;
typedef struct _p_EPS *EPS;
typedef const char *VecType;
typedef const char *MatType;
typedef const char *BVType;
typedef const char *DSType;
typedef const char *STType;
typedef const char *EPSType;

#define VECHIP         "hip"       /* device vectors, FP64 */
#define VECHIPF32      "hipf32"    /* device vectors, FP32 (the mixed-precision basis) */
#define MATHIPCSR      "hipcsr"
#define MATSEQDENSE    "seqdense"
#define BVHIPKS        "hipks"
#define DSHEP          "hep"
#define DSNHEP         "nhep"
#define DSNHEPTS       "nhepts"
#define STSHIFT        "shift"
#define EPSKRYLOVSCHUR "krylovschur"

/* value precision of device data (FP64 default; FP32 for the mixed form) */
typedef enum { SLEPC_AMD_F64 = 0, SLEPC_AMD_F32 = 1 } SlepcAMDPrecision;

typedef enum { NORM_1 = 0, NORM_2 = 1, NORM_FROBENIUS = 2, NORM_INFINITY = 3 } NormType;

/* ------------------------------------------------------------ system */
PetscErrorCode SlepcInitialize(int device);
/* the options database: "-name value" pairs (PETSC_OPTIONS is read by SlepcInitialize) */
PetscErrorCode PetscOptionsInsertString(const char in_str[]);
PetscErrorCode PetscOptionsSetValue(const char name[], const char value[]);
PetscErrorCode PetscOptionsClearValue(const char name[]);
PetscErrorCode SlepcFinalize(void);
PetscErrorCode SlepcInitialized(PetscBool *flg);
/* last error message recorded by the library */
const char    *SlepcAMDGetLastError(void);

/* ------------------------------------------------------------ Vec (device, FP64 or FP32) */
PetscErrorCode VecCreateHIP(PetscInt n, SlepcAMDPrecision prec, Vec *v);
PetscErrorCode VecDuplicate(Vec v, Vec *w);
PetscErrorCode VecDestroy(Vec *v);
PetscErrorCode VecGetSize(Vec v, PetscInt *n);
PetscErrorCode VecSet(Vec v, PetscScalar alpha);
PetscErrorCode VecCopy(Vec x, Vec y);
PetscErrorCode VecAXPY(Vec y, PetscScalar alpha, Vec x);
PetscErrorCode VecScale(Vec x, PetscScalar alpha);
PetscErrorCode VecNorm(Vec x, NormType type, PetscReal *nrm);
/* host <-> device transfer of the whole vector (FP64 host arrays) */
PetscErrorCode VecSetValuesHost(Vec v, const PetscScalar *values);
PetscErrorCode VecGetValuesHost(Vec v, PetscScalar *values);
/* uniform(-1,1) values from the seeded counter generator, generated on the device */
PetscErrorCode VecSetRandomSeeded(Vec v, uint64_t seed);
/* device pointer of the vector storage */
PetscErrorCode VecHIPGetArray(Vec v, void **d_ptr);

/* ------------------------------------------------------------ Mat */
PetscErrorCode MatRegister(const char name[], PetscErrorCode (*create)(Mat));
PetscErrorCode MatCreate(Mat *A);
PetscErrorCode MatSetSizes(Mat A, PetscInt m, PetscInt n, PetscInt M, PetscInt N);
PetscErrorCode MatSetType(Mat A, MatType type);
PetscErrorCode MatGetType(Mat A, MatType *type);
PetscErrorCode MatGetSize(Mat A, PetscInt *M, PetscInt *N);
/* y = A^T x (for MATHIPCSR with an explicit transpose formed on the device on first use) */
PetscErrorCode MatMultTranspose(Mat A, Vec x, Vec y);
/* a new MATHIPCSR holding A^T, formed on the device (the left operator of two-sided Krylov-Schur) */
PetscErrorCode MatHIPCSRCreateTranspose(Mat A, Mat *AT);
PetscErrorCode MatMult(Mat A, Vec x, Vec y);
PetscErrorCode MatDestroy(Mat *A);
/* CSR arrays on the host (copied to the device); row pointers and column
   indices 32-bit, values FP64 */
PetscErrorCode MatSeqAIJSetPreallocationCSR(Mat A, const PetscInt i[], const PetscInt j[], const PetscScalar v[]);
/* same with FP32 values */
PetscErrorCode MatHIPCSRSetArraysF32(Mat A, const PetscInt i[], const PetscInt j[], const float v[]);
/* CSR arrays already on the device (copied) */
PetscErrorCode MatHIPCSRSetDeviceArrays(Mat A, PetscInt nnz, const PetscInt *d_i, const PetscInt *d_j,
                                        const void *d_v, SlepcAMDPrecision vprec);
PetscErrorCode MatHIPCSRGetInfo(Mat A, PetscInt *nnz, PetscInt *max_row, PetscInt *variant, PetscInt *nblocks);
PetscErrorCode MatCreateSeqDense(PetscInt m, PetscInt n, PetscScalar *data, Mat *A);
PetscErrorCode MatDenseGetArray(Mat A, PetscScalar **array);
PetscErrorCode MatDenseRestoreArray(Mat A, PetscScalar **array);
PetscErrorCode MatDenseGetLDA(Mat A, PetscInt *lda);
PetscErrorCode MatCreateHIPCSR(PetscInt n, const PetscInt i[], const PetscInt j[], const PetscScalar v[], Mat *A);
/* PETSc binary matrix files (MatLoad format, as read by src/eps/tutorials/ex4.c:43-50):
   host CSR with sorted rows (free with MatFreePETScBinary), or straight into a MATHIPCSR */
PetscErrorCode MatReadPETScBinary(const char *path, PetscInt *m, PetscInt *n, PetscInt **rowptr, PetscInt **colidx,
                                  PetscScalar **val);
void           MatFreePETScBinary(PetscInt *rowptr, PetscInt *colidx, PetscScalar *val);
PetscErrorCode MatLoadPETScBinary(const char *path, Mat *A);

/* ------------------------------------------------------------ BV */
typedef enum { BV_ORTHOG_CGS = 0, BV_ORTHOG_MGS = 1 } BVOrthogType;
typedef enum { BV_ORTHOG_REFINE_IFNEEDED = 0, BV_ORTHOG_REFINE_NEVER = 1, BV_ORTHOG_REFINE_ALWAYS = 2 } BVOrthogRefineType;
/* BVOrthogBlockType (include/slepcbv.h:84-88) */
typedef enum { BV_ORTHOG_BLOCK_GS = 0, BV_ORTHOG_BLOCK_CHOL = 1, BV_ORTHOG_BLOCK_TSQR = 2, BV_ORTHOG_BLOCK_TSQRCHOL = 3,
               BV_ORTHOG_BLOCK_SVQB = 4 } BVOrthogBlockType;
PetscErrorCode BVRegister(const char name[], PetscErrorCode (*create)(BV));
PetscErrorCode BVCreate(BV *bv);
PetscErrorCode BVSetSizes(BV bv, PetscInt n, PetscInt N, PetscInt m);
PetscErrorCode BVSetSizesFromVec(BV bv, Vec t, PetscInt m);
PetscErrorCode BVGetSizes(BV bv, PetscInt *n, PetscInt *N, PetscInt *m);
PetscErrorCode BVSetType(BV bv, BVType type);
PetscErrorCode BVGetType(BV bv, BVType *type);
PetscErrorCode BVHIPKSSetPrecision(BV bv, SlepcAMDPrecision prec);
PetscErrorCode BVHIPKSGetPrecision(BV bv, SlepcAMDPrecision *prec);
/* capture each Arnoldi step (SpMV, CGS/DGKS passes, scaling) of an expansion
   once per column as a hipGraph and replay it in later restarts; meant for
   small problems, where launch overhead rivals the kernels (no counterpart in
   the reference) */
PetscErrorCode BVHIPKSSetStepGraphs(BV bv, PetscBool flg);
PetscErrorCode BVSetOrthogonalization(BV bv, BVOrthogType type, BVOrthogRefineType refine, PetscReal eta,
                                      BVOrthogBlockType block);
PetscErrorCode BVGetOrthogonalization(BV bv, BVOrthogType *type, BVOrthogRefineType *refine, PetscReal *eta,
                                      BVOrthogBlockType *block);
/* BVSetVecType (bvbasic.c): VECHIP (FP64 columns, the default) or VECHIPF32 (FP32
   columns, the mixed-precision basis); before the storage is allocated */
PetscErrorCode BVSetVecType(BV bv, VecType vtype);
PetscErrorCode BVGetVecType(BV bv, VecType *vtype);
/* BVSetFromOptions (bvbasic.c): -bv_type, -bv_orthog_type cgs, -bv_orthog_refine
   never|ifneeded|always, -bv_orthog_eta, -bv_orthog_block gs|chol|tsqr|tsqrchol|svqb,
   and the type's own options (BVHIPKS: -bv_hipks_krylov 0|1) */
PetscErrorCode BVSetFromOptions(BV bv);
/* BVHIPKS: turn the device-resident Krylov expansion on (default) or off; off, the
   interface runs the reference's per-step loop over the gramschmidt op
   (also -bv_hipks_krylov 0 in the options database) */
PetscErrorCode BVHIPKSSetKrylovExtension(BV bv, PetscBool flg);
PetscErrorCode BVHIPKSGetKrylovExtension(BV bv, PetscBool *flg);
PetscErrorCode BVSetActiveColumns(BV bv, PetscInt l, PetscInt k);
PetscErrorCode BVGetActiveColumns(BV bv, PetscInt *l, PetscInt *k);
PetscErrorCode BVGetColumn(BV bv, PetscInt j, Vec *v);
PetscErrorCode BVRestoreColumn(BV bv, PetscInt j, Vec *v);
PetscErrorCode BVSetRandomColumn(BV bv, PetscInt j);
PetscErrorCode BVScaleColumn(BV bv, PetscInt j, PetscScalar alpha);
PetscErrorCode BVNormColumn(BV bv, PetscInt j, NormType type, PetscReal *norm);
PetscErrorCode BVCopyColumn(BV bv, PetscInt j, PetscInt i);
PetscErrorCode BVCopyVec(BV bv, PetscInt j, Vec w);
PetscErrorCode BVInsertVec(BV bv, PetscInt j, Vec w);
PetscErrorCode BVInsertVecs(BV bv, PetscInt s, PetscInt *m, Vec *W, PetscBool orth);
PetscErrorCode BVInsertConstraints(BV bv, PetscInt *nc, Vec *C);
PetscErrorCode BVGetNumConstraints(BV bv, PetscInt *nc);
PetscErrorCode BVDotVec(BV bv, Vec y, PetscScalar m[]);
PetscErrorCode BVMultVec(BV bv, PetscScalar alpha, PetscScalar beta, Vec y, PetscScalar q[]);
PetscErrorCode BVMultInPlace(BV bv, Mat Q, PetscInt s, PetscInt e);
/* column j = beta * column j + alpha * V(:, l:j) q (bvops.c:165-192); q has j - l entries */
PetscErrorCode BVMultColumn(BV bv, PetscScalar alpha, PetscScalar beta, PetscInt j, PetscScalar q[]);
/* scale the active columns to unit norm; columns i, i+1 with eigi[i] != 0 share one
   norm (bvglobal.c:829-849, bvlapack.c:88-130) */
PetscErrorCode BVNormalize(BV bv, PetscScalar eigi[]);
/* M = Y^T X over the active columns (bvglobal.c:86-120): rows ly..ky-1, columns lx..kx-1 of M */
PetscErrorCode BVDot(BV X, BV Y, Mat M);
/* QR of the active columns (bvorthog.c:726-767, methods bvorthog.c:489-663): leading columns 0..l-1 are kept
   (assumed orthonormal), columns l..k-1 become orthonormal, R (k x k or NULL) gets
   the columns l..k-1 of the factor; block types GS, CHOL and SVQB */
PetscErrorCode BVOrthogonalize(BV V, Mat R);
/* M = Y^T A X over the active columns (bvglobal.c:1093-1160, the non-symmetric path) */
PetscErrorCode BVMatProject(BV X, Mat A, BV Y, Mat M);
PetscErrorCode BVOrthogonalizeColumn(BV bv, PetscInt j, PetscScalar H[], PetscReal *norm, PetscBool *lindep);
PetscErrorCode BVOrthonormalizeColumn(BV bv, PetscInt j, PetscBool replace, PetscReal *norm, PetscBool *lindep);
/* v against the active columns and the constraints (bvorthog.c:236-266); H gets k-l coefficients */
PetscErrorCode BVOrthogonalizeVec(BV bv, Vec v, PetscScalar H[], PetscReal *norm, PetscBool *lindep);
PetscErrorCode BVMatMultColumn(BV bv, Mat A, PetscInt j);
PetscErrorCode BVMatArnoldi(BV bv, Mat A, Mat H, PetscInt k, PetscInt *m, PetscReal *beta, PetscBool *breakdown);
PetscErrorCode BVMatLanczos(BV bv, Mat A, Mat T, PetscInt k, PetscInt *m, PetscReal *beta, PetscBool *breakdown);
/* host copy of the coefficient buffer, (nc+m) x m column-major */
PetscErrorCode BVGetBufferHost(BV bv, const PetscScalar **a, PetscInt *ld);
PetscErrorCode BVGetColumnValuesHost(BV bv, PetscInt j, PetscScalar *values);
PetscErrorCode BVSetColumnValuesHost(BV bv, PetscInt j, const PetscScalar *values);
PetscErrorCode BVDestroy(BV *bv);
/* the plugin constructor of the device basis type */
PetscErrorCode BVCreate_HIPKS(BV bv);

/* ------------------------------------------------------------ DS */
typedef enum { DS_STATE_RAW = 0, DS_STATE_INTERMEDIATE = 1, DS_STATE_CONDENSED = 2, DS_STATE_TRUNCATED = 3 } DSStateType;
typedef enum { DS_MAT_A = 0, DS_MAT_B = 1, DS_MAT_T = 3, DS_MAT_Q = 5, DS_MAT_Z = 6, DS_MAT_X = 7, DS_MAT_Y = 8,
               DS_NUM_MAT = 12 } DSMatType;
PetscErrorCode DSRegister(const char name[], PetscErrorCode (*create)(DS));
PetscErrorCode DSCreate(DS *ds);
PetscErrorCode DSSetType(DS ds, DSType type);
PetscErrorCode DSAllocate(DS ds, PetscInt ld);
PetscErrorCode DSGetLeadingDimension(DS ds, PetscInt *ld);
PetscErrorCode DSSetDimensions(DS ds, PetscInt n, PetscInt l, PetscInt k);
PetscErrorCode DSGetDimensions(DS ds, PetscInt *n, PetscInt *l, PetscInt *k, PetscInt *t);
PetscErrorCode DSSetState(DS ds, DSStateType state);
PetscErrorCode DSGetState(DS ds, DSStateType *state);
PetscErrorCode DSSetCompact(DS ds, PetscBool comp);
PetscErrorCode DSSetExtraRow(DS ds, PetscBool ext);
PetscErrorCode DSGetArray(DS ds, DSMatType m, PetscScalar **a);
PetscErrorCode DSRestoreArray(DS ds, DSMatType m, PetscScalar **a);
PetscErrorCode DSGetMat(DS ds, DSMatType m, Mat *A);
PetscErrorCode DSRestoreMat(DS ds, DSMatType m, Mat *A);
PetscErrorCode DSSolve(DS ds, PetscScalar eigr[], PetscScalar eigi[]);
PetscErrorCode DSSort(DS ds, PetscScalar eigr[], PetscScalar eigi[], PetscScalar rr[], PetscScalar ri[], PetscInt *k);
PetscErrorCode DSUpdateExtraRow(DS ds);
PetscErrorCode DSVectors(DS ds, DSMatType mat, PetscInt *j, PetscReal *rnorm);
PetscErrorCode DSGetTruncateSize(DS ds, PetscInt l, PetscInt n, PetscInt *k);
PetscErrorCode DSTruncate(DS ds, PetscInt n, PetscBool trim);
/* A + g b^T with g = (A - tau I)^-T b, b = beta e_n (harmonic Rayleigh-Ritz), or with
   recover the translation undone on the restarted decomposition; gamma = sqrt(1 + g^T g)
   (dsops.c:1050-1066, DSNHEP: dsnhep.c:466-535) */
PetscErrorCode DSTranslateHarmonic(DS ds, PetscScalar tau, PetscReal beta, PetscBool recover, PetscScalar *g,
                                   PetscReal *gamma);
/* sorting criterion: which (EPSWhich values) and target */
PetscErrorCode DSSetSortCriterion(DS ds, int which, PetscScalar target);
PetscErrorCode DSDestroy(DS *ds);

/* ------------------------------------------------------------ ST */
typedef const char *STType;
typedef const char *KSPType;
typedef const char *PCType;
typedef struct _p_KSP *KSP;
typedef struct _p_PC *PC;
#define STSHIFT   "shift"
#define STSINVERT "sinvert"
#define KSPCG     "cg"
#define KSPBCGS   "bcgs"
#define KSPPREONLY "preonly"
#define PCJACOBI  "jacobi"
#define PCNONE    "none"
#define PCILU     "ilu"
#define PCLU      "lu"
PetscErrorCode STCreate(ST *st);
PetscErrorCode STSetType(ST st, STType type);
PetscErrorCode STGetType(ST st, STType *type);
PetscErrorCode STSetShift(ST st, PetscScalar shift);
PetscErrorCode STGetShift(ST st, PetscScalar *shift);
PetscErrorCode STSetDefaultShift(ST st, PetscScalar defaultshift);
/* STSINVERT's inner solver on the device: conjugate gradients (KSPCG, the
   default, for a positive definite A - sigma I) or BiCGStab (KSPBCGS, for a
   shift inside the spectrum), Jacobi or no preconditioner (the reference's
   KSP/PC objects, reduced to what runs here).  The option -st_ksp_type
   selects the solver, as KSPSetFromOptions does. */
PetscErrorCode STGetKSP(ST st, KSP *ksp);
PetscErrorCode KSPSetType(KSP ksp, KSPType type);
PetscErrorCode KSPGetPC(KSP ksp, PC *pc);
PetscErrorCode PCSetType(PC pc, PCType type);
PetscErrorCode KSPSetTolerances(KSP ksp, PetscReal rtol, PetscReal abstol, PetscReal dtol, PetscInt maxits);
PetscErrorCode KSPGetTotalIterations(KSP ksp, PetscInt *its);
PetscErrorCode STBackTransform(ST st, PetscInt n, PetscScalar *eigr, PetscScalar *eigi);
PetscErrorCode STSetMatrices(ST st, PetscInt n, Mat A[]);
PetscErrorCode STGetOperator(ST st, Mat *Op);
PetscErrorCode STRestoreOperator(ST st, Mat *Op);
PetscErrorCode STApply(ST st, Vec x, Vec y);
PetscErrorCode STDestroy(ST *st);

/* ------------------------------------------------------------ EPS */
typedef enum { EPS_HEP = 1, EPS_GHEP = 2, EPS_NHEP = 3, EPS_GNHEP = 4 } EPSProblemType;
typedef enum { EPS_LARGEST_MAGNITUDE = 1, EPS_SMALLEST_MAGNITUDE = 2, EPS_LARGEST_REAL = 3,
               EPS_SMALLEST_REAL = 4, EPS_LARGEST_IMAGINARY = 5, EPS_SMALLEST_IMAGINARY = 6,
               EPS_TARGET_MAGNITUDE = 7, EPS_TARGET_REAL = 8 } EPSWhich;
typedef enum { EPS_CONVERGED_TOL = 1, EPS_CONVERGED_USER = 3, EPS_DIVERGED_ITS = -1,
               EPS_DIVERGED_BREAKDOWN = -2, EPS_CONVERGED_ITERATING = 0 } EPSConvergedReason;
typedef enum { EPS_ERROR_ABSOLUTE = 0, EPS_ERROR_RELATIVE = 1, EPS_ERROR_BACKWARD = 2 } EPSErrorType;
/* include/slepceps.h:94-100; Krylov-Schur supports EPS_RITZ and EPS_HARMONIC */
typedef enum { EPS_RITZ = 0, EPS_HARMONIC = 1, EPS_HARMONIC_RELATIVE = 2, EPS_HARMONIC_RIGHT = 3,
               EPS_HARMONIC_LARGEST = 4, EPS_REFINED = 5, EPS_REFINED_HARMONIC = 6 } EPSExtraction;
typedef PetscErrorCode (*EPSMonitorFn)(EPS, PetscInt its, PetscInt nconv, PetscScalar *eigr,
                                       PetscScalar *eigi, PetscReal *errest, PetscInt nest, void *ctx);

PetscErrorCode EPSRegister(const char name[], PetscErrorCode (*create)(EPS));
PetscErrorCode EPSCreate(EPS *eps);
PetscErrorCode EPSSetType(EPS eps, EPSType type);
PetscErrorCode EPSGetType(EPS eps, EPSType *type);
PetscErrorCode EPSSetOperators(EPS eps, Mat A, Mat B);
PetscErrorCode EPSSetProblemType(EPS eps, EPSProblemType type);
PetscErrorCode EPSSetWhichEigenpairs(EPS eps, EPSWhich which);
PetscErrorCode EPSSetTarget(EPS eps, PetscScalar target);
PetscErrorCode EPSGetTarget(EPS eps, PetscScalar *target);
/* EPSSetExtraction (epsopts.c): harmonic extraction targets interior eigenvalues near the target */
PetscErrorCode EPSSetExtraction(EPS eps, EPSExtraction extr);
PetscErrorCode EPSGetExtraction(EPS eps, EPSExtraction *extr);
PetscErrorCode EPSSetDimensions(EPS eps, PetscInt nev, PetscInt ncv, PetscInt mpd);
PetscErrorCode EPSGetDimensions(EPS eps, PetscInt *nev, PetscInt *ncv, PetscInt *mpd);
PetscErrorCode EPSSetTolerances(EPS eps, PetscReal tol, PetscInt maxits);
PetscErrorCode EPSGetTolerances(EPS eps, PetscReal *tol, PetscInt *maxits);
PetscErrorCode EPSSetInitialSpace(EPS eps, PetscInt n, Vec is[]);
PetscErrorCode EPSSetDeflationSpace(EPS eps, PetscInt n, Vec v[]);
PetscErrorCode EPSKrylovSchurSetRestart(EPS eps, PetscReal keep);
PetscErrorCode EPSKrylovSchurSetLocking(EPS eps, PetscBool lock);
/* Mixed-precision form (no reference counterpart): an FP32-basis Krylov-Schur
   phase down to tol32 whose Ritz vectors seed the FP64 restarts; tol32 <= 0
   turns it off.  The stats report the FP32 phase. */
PetscErrorCode EPSKrylovSchurSetMixedPrecision(EPS eps, PetscReal tol32);
/* capture each Arnoldi step of the basis' expansions as a hipGraph and replay it
   (BVHIPKSSetStepGraphs on the solver's basis, applied at set-up) */
PetscErrorCode EPSKrylovSchurSetStepGraphs(EPS eps, PetscBool flg);
PetscErrorCode EPSKrylovSchurGetMixedStats(EPS eps, PetscInt *its32, PetscInt *nconv32, int64_t *nmatmult32,
                                           int64_t *passes32, PetscReal *t32);
PetscErrorCode EPSSetBV(EPS eps, BV bv);
PetscErrorCode EPSGetBV(EPS eps, BV *bv);
PetscErrorCode EPSGetDS(EPS eps, DS *ds);
PetscErrorCode EPSGetST(EPS eps, ST *st);
PetscErrorCode EPSMonitorSet(EPS eps, EPSMonitorFn fn, void *ctx);
PetscErrorCode EPSSetUp(EPS eps);
PetscErrorCode EPSSolve(EPS eps);
/* EPSSolve in parts, for callers that interleave solves (no reference
   counterpart; EPSSolve == Begin, Step until done, End).  EPSSolveReady
   reports whether the next Step would find the device work it collects done. */
PetscErrorCode EPSSolveBegin(EPS eps);
PetscErrorCode EPSSolveReady(EPS eps, PetscBool *ready);
PetscErrorCode EPSSolveStep(EPS eps, PetscBool *done);
PetscErrorCode EPSSolveEnd(EPS eps);
PetscErrorCode EPSGetConverged(EPS eps, PetscInt *nconv);
PetscErrorCode EPSGetIterationNumber(EPS eps, PetscInt *its);
PetscErrorCode EPSGetConvergedReason(EPS eps, EPSConvergedReason *reason);
PetscErrorCode EPSGetEigenvalue(EPS eps, PetscInt i, PetscScalar *eigr, PetscScalar *eigi);
PetscErrorCode EPSGetEigenvector(EPS eps, PetscInt i, Vec Vr, Vec Vi);
/* left eigenvector (two-sided solves, epssolve.c:567-611) */
PetscErrorCode EPSGetLeftEigenvector(EPS eps, PetscInt i, Vec Wr, Vec Wi);
/* compute left eigenvectors as well, by two-sided Krylov-Schur (ks-twosided.c) */
PetscErrorCode EPSSetTwoSided(EPS eps, PetscBool twosided);
PetscErrorCode EPSGetTwoSided(EPS eps, PetscBool *twosided);
PetscErrorCode EPSGetEigenpair(EPS eps, PetscInt i, PetscScalar *eigr, PetscScalar *eigi, Vec Vr, Vec Vi);
PetscErrorCode EPSGetErrorEstimate(EPS eps, PetscInt i, PetscReal *errest);
PetscErrorCode EPSComputeError(EPS eps, PetscInt i, EPSErrorType type, PetscReal *error);
PetscErrorCode EPSDestroy(EPS *eps);
/* solve statistics: SpMV count, executed CGS passes, restarts, host DS seconds, solve seconds */
PetscErrorCode EPSGetSolveStats(EPS eps, int64_t *nmatmult, int64_t *gs_passes, PetscInt *its,
                                PetscReal *t_ds, PetscReal *t_solve);

/* the plugin constructors of the solver and the device matrix type */
PetscErrorCode EPSCreate_KrylovSchur(EPS eps);
PetscErrorCode MatCreate_HIPCSR(Mat A);

/* the kernel-library context used by the objects (ks_ctx of ks_hip.h) */
void *SlepcAMDGetDefaultContext(void);

/* ---------------------------------------------------------------- serving
 * No reference counterpart (SLEPc runs one EPSSolve at a time).  A server
 * admits eigenproblems as they arrive, in arrival order, while a slot is free
 * and their HBM (CSR + basis + work vectors) fits the budget; each admitted
 * problem is an ordinary EPSKRYLOVSCHUR solve on its own stream, advanced one
 * restart at a time by SlepcServerProgress, which never waits on the device.
 * The CSR arrays and v0 must stay valid until the request finishes. */
typedef struct {
  PetscInt n;
  const PetscInt *rowptr, *colidx;
  const PetscScalar *val;        /* FP64 values, or NULL with val32 */
  const float *val32;            /* FP32 values (the operator is then stored in FP32) */
  PetscInt nev, ncv, max_it;     /* max_it <= 0: the solver's default */
  PetscReal tol;
  EPSWhich which;                /* 0: largest magnitude */
  PetscBool hermitian;
  const PetscScalar *v0;         /* start vector (length n) or NULL */
  PetscBool want_vectors;        /* keep the eigenvectors on the host */
  PetscReal mixed_tol;           /* > tol: FP32-basis first phase to this tolerance */
  double t_arrival;              /* on the SlepcServerClock clock; <= 0: submission time */
  PetscBool twosided;            /* two-sided Krylov-Schur (non-Hermitian, FP64 values, no FP32 phase) */
} SlepcRequest;
typedef struct {
  PetscInt nconv, nret, ngood, its; /* converged, returned (<= nev), returned with residual <= tol, restarts */
  EPSConvergedReason reason;
  int64_t nmatmult, gs_passes;
  double t_arrival, t_admit, t_done;
  double t_host_ds;              /* host seconds in the projected problems of this request */
  PetscErrorCode error;          /* nonzero: the request failed (at admission) with this code, no pairs */
} SlepcResult;
typedef struct _p_SlepcServer *SlepcServer;
PetscErrorCode SlepcServerCreate(size_t hbm_budget, PetscInt max_active, SlepcServer *srv);
PetscErrorCode SlepcServerSubmit(SlepcServer srv, const SlepcRequest *req, PetscInt *id);
PetscErrorCode SlepcServerProgress(SlepcServer srv, PetscBool *progressed, PetscInt *nactive, PetscInt *nqueued,
                                   PetscInt *nfinished);
PetscErrorCode SlepcServerGetResult(SlepcServer srv, PetscInt id, PetscBool *done, SlepcResult *res,
                                    PetscScalar *eigr, PetscScalar *eigi, PetscReal *err, PetscScalar *X);
PetscErrorCode SlepcServerGetStats(SlepcServer srv, size_t *peak_reserved, size_t *pool_peak, size_t *budget);
/* grouping (default on): the Arnoldi steps of all expanding requests go out as one
   launch per kernel kind (ks_group_step); off: each request launches its own */
PetscErrorCode SlepcServerSetGrouped(SlepcServer srv, PetscBool grouped);
PetscErrorCode SlepcServerGetGroupStats(SlepcServer srv, int64_t *steps, int64_t *member_steps);
/* host seconds spent admitting, issuing grouped steps, restarting, retiring, polling */
PetscErrorCode SlepcServerGetHostTimes(SlepcServer srv, double t[5]);
/* serve a whole stream from C: req[i].t_arrival is an offset (seconds, ascending)
   from the call; results into res[i] and eigr[i]/eigi[i]/err[i] (nev each), X[i]
   (n x nev or NULL); *wall = seconds until the last result */
PetscErrorCode SlepcServerRun(SlepcServer srv, PetscInt nreq, const SlepcRequest *req, SlepcResult *res,
                              PetscScalar **eigr, PetscScalar **eigi, PetscReal **err, PetscScalar **X, double *wall);
PetscErrorCode SlepcServerDestroy(SlepcServer *srv);
double SlepcServerClock(void);   /* monotonic seconds, the clock of t_arrival/t_admit/t_done */

#ifdef __cplusplus
}
#endif
#endif
