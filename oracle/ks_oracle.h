/*
 * ks_oracle.h -- CPU restatement of SLEPc 3.22's Krylov-Schur eigensolver path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the
 * smoke() entry point and bench.py's cpu_baseline leg may load it.  The
 * product (slepc_amd) never links or calls it.
 *
 * Every function follows the reference file:line given in ks_oracle.c.
 * Scalars are real double (PETSc real build), indices are 32-bit ints
 * (PetscInt default).
 */
#ifndef KS_ORACLE_H
#define KS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* which eigenpairs (EPSWhich, include/slepceps.h) */
enum { OR_LARGEST_MAGNITUDE = 1, OR_SMALLEST_MAGNITUDE = 2, OR_LARGEST_REAL = 3,
       OR_SMALLEST_REAL = 4, OR_LARGEST_IMAGINARY = 5, OR_SMALLEST_IMAGINARY = 6,
       OR_TARGET_MAGNITUDE = 7, OR_TARGET_REAL = 8 };

/* orthogonalization refinement (BVOrthogRefineType) */
enum { OR_REFINE_IFNEEDED = 0, OR_REFINE_NEVER = 1, OR_REFINE_ALWAYS = 2 };

/* converged reasons (EPSConvergedReason) */
enum { OR_CONVERGED_TOL = 1, OR_CONVERGED_ITERATING = 0, OR_DIVERGED_ITS = -1,
       OR_DIVERGED_BREAKDOWN = -2 };

typedef struct {
  int    nev, ncv, mpd, max_it;  /* ncv/mpd/max_it <= 0 means PETSC_DETERMINE */
  double tol;                    /* <= 0 means default 1e-8 */
  int    which;
  double target;
  int    hermitian;              /* 1: EPS_HEP (Lanczos, DSHEP), 0: EPS_NHEP (Arnoldi, DSNHEP) */
  double keep;                   /* <= 0 means 0.5 */
  int    lock;                   /* 1 = locking variant (default) */
  double eta;                    /* <= 0 means 0.7071 */
  int    refine;                 /* OR_REFINE_* */
  int    nthreads;               /* OpenMP threads for SpMV/BLAS; <= 0: leave as is */
  int    restart_cap;            /* > 0: stop after this many restarts (bounded timing sample) */
  int    extraction;             /* 0 EPS_RITZ, 1 EPS_HARMONIC (translation by target) */
  int    twosided;               /* 1: two-sided variant (left eigenvectors too) */
  int    st_sinvert;             /* 1: STSINVERT with shift = target (inner PCG unless set below) */
  int    pc_none;                /* 1: PCG without the Jacobi preconditioner */
  double ksp_rtol;               /* <= 0 means 1e-8 */
  int    ksp_bcgs;               /* 1: BiCGStab instead of CG for the inner solve */
  int    pc_ilu;                 /* 1: ILU(0) preconditioner (BiCGStab only) */
  int    ksp_lu;                 /* 1: KSPPREONLY + PCLU, the dense LU direct solve */
} or_opts;

typedef struct {
  int    nconv, its, reason;
  long   nmatmult;               /* SpMV count inside the restart loop */
  long   gs_passes;              /* executed CGS passes (DGKS included) */
  double t_solve;                /* wall seconds of the restart loop + vectors */
  double t_ds;                   /* wall seconds spent in DS (projected problem) */
  long   ksp_its;                /* inner PCG iterations (shift-and-invert) */
} or_info;

/* y = A x for CSR (PETSc MatMult_SeqAIJ semantics). */
int or_spmv(int n, const int *rowptr, const int *colidx, const double *val,
            const double *x, double *y);

/* Counter-based uniform(-1,1) generator used for seeded start vectors:
   u_i = 2*(splitmix64(seed*0x9E3779B97F4A7C15 + i) >> 11)*2^-53 - 1. */
void or_random_uniform(unsigned long long seed, int n, double *out);

/* BVOrthonormalizeColumn on column j of a column-major n x (nc+m) array V
   (leading dimension ld, nc constraint columns first).  H receives the
   nc+j+1 coefficients (norm last).  Returns 0 or a negative error code. */
int or_orthonormalize_column(int n, int ld, double *V, int nc, int j, int refine,
                             double eta, double *H, double *norm, int *lindep,
                             int *npasses);

/* BVMultInPlace: V(:,s:e-1) = V(:,l:k-1) * Q(l:k-1,s:e-1), columns counted
   after nc constraints. */
/* DSTranslateHarmonic (translate) on an n x n projected matrix with extra row */
int or_ds_translate_harmonic(int ld, int n, double *A, double tau, double beta, double *g, double *gamma);
/* M(ly:ky, lx:kx) = Y(:, ly:ky)^T X(:, lx:kx) (BVDot) */
int or_bv_dot(int n, const double *X, int ldx, int lx, int kx, const double *Y, int ldy, int ly, int ky,
              double *M, int ldm);
/* BVOrthogonalize of V(:, l:k) with block type 0 (GS), 1 (CHOL) or 4 (SVQB); R: columns l..k-1 */
int or_bv_orthogonalize(int n, int ld, double *V, int l, int k, int block, double *R, int ldr);
/* TSQR over P row blocks, the reference's multi-process TSQR with a flat reduction */
int or_bv_tsqr_split(int n, int ld, double *V, int k, int P, double *R, int ldr);
/* ILU(0) factors of A - sigma I on A's pattern (L below the diagonal, unit diagonal implied; U on and above) */
int or_ilu0(int n, const int *rowptr, const int *colidx, const double *val, double sigma, double *lu);
int or_lu(int n, const int *rowptr, const int *colidx, const double *val, double sigma, double *lu, int *ipiv);
int or_multinplace(int n, int ld, double *V, int nc, int l, int k,
                   const double *Q, int ldq, int s, int e);

/* Run the DS sequence used by the DS tests: DSSolve, DSSort (comparison
   `which`), DSUpdateExtraRow (if extrarow), DSVectors(X, jvec) for rnorm and
   DSVectors(X, all).  type 1 = DSHEP (T compact if compact!=0), 2 = DSNHEP.
   A is ld x ld (NHEP or non-compact HEP), T is ld x 3 (compact HEP).
   state: 0 RAW, 1 INTERMEDIATE.  Q, X are ld x ld outputs. */
int or_ds_run(int type, int ld, int n, int l, int k, int compact, int extrarow,
              int state, double *A, double *T, int which, double *wr, double *wi,
              double *Q, int jvec, double *rnorm, double *X);

/* Full EPSSolve with Krylov-Schur.  A given as CSR.  v0 (length n) is the
   initial space (EPSSetInitialSpace with one vector) or NULL for a random
   start from seed 0 of or_random_uniform.  defl (n x nds, column-major) is a
   deflation space (EPSSetDeflationSpace) or NULL.  Outputs, sized >= ncv+1:
   eigr, eigi, errest (Ritz estimates), err (EPSComputeError relative, for
   i < nconv).  X (n x (ncv+1), ld=n) receives eigenvectors in EPSGetEigenpair
   order (conjugate pairs as real/imaginary columns) when non-NULL. */
int or_eps_solve(int n, const int *rowptr, const int *colidx, const double *val,
                 const or_opts *opts, const double *v0, int nds, const double *defl,
                 double *eigr, double *eigi, double *errest, double *err, double *X,
                 or_info *info);
/* the same with the left eigenvectors of a two-sided solve (real parts) in Y (n x nconv) */
int or_eps_solve_twosided(int n, const int *rowptr, const int *colidx, const double *val,
                          const or_opts *opts, const double *v0, int nds, const double *defl,
                          double *eigr, double *eigi, double *errest, double *err, double *X, double *Y,
                          or_info *info);

#ifdef __cplusplus
}
#endif
#endif
