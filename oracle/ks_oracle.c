/*
 * ks_oracle.c -- CPU restatement of SLEPc 3.22's Krylov-Schur path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by smoke() and by
 * bench.py's cpu_baseline leg; never by the product.  It follows the reference
 * line by line so that its trajectory (restarts, DGKS passes, truncation sizes)
 * is the one SLEPc's CPU path takes on the same inputs:
 *
 *   CSR MatMult               PETSc MatMult_SeqAIJ row loop (called at
 *                             src/sys/classes/bv/interface/bvops.c:879)
 *   BVMAT ops                 src/sys/classes/bv/impls/mat/bvmat.c:38-71,110-176,250-260
 *                             src/sys/classes/bv/interface/bvblas.c:24-278
 *   CGS + DGKS                src/sys/classes/bv/interface/bvorthog.c:91-217,380-427
 *   coefficient buffer        include/slepc/private/bvimpl.h:121-141,289-415
 *   Arnoldi / Lanczos         src/sys/classes/bv/interface/bvkrylov.c:56-226
 *   DSHEP (compact)           src/sys/classes/ds/impls/hep/dshep.c:137-426,643-670
 *   DSNHEP                    src/sys/classes/ds/impls/dsutil.c:21-171,
 *                             src/sys/classes/ds/impls/nhep/dsnhep.c:101-226,318-425
 *   DS interface              src/sys/classes/ds/interface/dsops.c:63-160,232-275,329-344
 *                             src/sys/classes/ds/interface/dspriv.c:172-270
 *   sorting criteria          src/sys/slepcsc.c:40-190
 *   KS restart loop           src/eps/impls/krylov/krylovschur/krylovschur.c:93-194,227-337
 *   convergence test          src/eps/impls/krylov/epskrylov.c:207-295
 *   EPS driver                src/eps/interface/epssolve.c:119-208,666-873
 *                             src/eps/interface/epssetup.c:286-418,648-732
 *                             src/eps/interface/epsdefault.c:28-50,105-169,224-302
 *   shift-and-invert          src/sys/classes/st/impls/sinvert/sinvert.c:16-105,
 *                             src/sys/classes/st/interface/stsolve.c:244-262, stsles.c:47-54
 *   its inner solves          PETSc (not in the reference tree), restated from the
 *                             published methods: KSPCG, KSPBCGS, PCILU levels 0 (Saad,
 *                             Algorithm 10.4) and KSPPREONLY + PCLU as LAPACK's
 *                             dgetf2/dgetrs (the LU pinned to scipy's dgetrf: same
 *                             pivots and factors), the eigen-results pinned by goldens
 *
 * Third-party arithmetic at the boundary: BLAS/LAPACK (MKL via libmkl_rt,
 * LP64) for gemv/gemm/nrm2/lartg/rot/steqr/gehrd/orghr/hseqr/trexc/trevc/lascl,
 * exactly the routines the reference calls.
 */
#include "ks_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <omp.h>

/* ---------------------------------------------------------------- BLAS/LAPACK */
extern void   dgemv_(const char*, const int*, const int*, const double*, const double*, const int*,
                     const double*, const int*, const double*, double*, const int*);
extern void   dgemm_(const char*, const char*, const int*, const int*, const int*, const double*,
                     const double*, const int*, const double*, const int*, const double*, double*,
                     const int*);
extern double dnrm2_(const int*, const double*, const int*);
extern void   dpotrf_(const char*, const int*, double*, const int*, int*);
extern void   dtrtri_(const char*, const char*, const int*, double*, const int*, int*);
extern void   dgeqrf_(const int*, const int*, double*, const int*, double*, double*, const int*, int*);
extern void   dorgqr_(const int*, const int*, const int*, double*, const int*, const double*, double*, const int*, int*);
extern void   dsyev_(const char*, const char*, const int*, double*, const int*, double*, double*, const int*,
                     int*);
extern void   dgetrf_(const int*, const int*, double*, const int*, int*, int*);
extern void   dgetrs_(const char*, const int*, const int*, const double*, const int*, const int*, double*,
                      const int*, int*);
extern int    MKL_Set_Threading_Layer(int);     /* MKL service (mkl_rt), by value: 3 = GNU OpenMP */

/* MKL threads through the same GNU OpenMP runtime as the loops here; its Intel
   runtime nested under a GNU OpenMP region gives wrong panel products when a
   second libgomp (torch's) is the one in the process.  Set before the first
   BLAS call of every entry point. */
static void or_mkl_layer(void)
{
  static int done = 0;
  if (!done) { MKL_Set_Threading_Layer(3); done = 1; }
}
extern void   dlartg_(double*, double*, double*, double*, double*);
extern void   drot_(const int*, double*, const int*, double*, const int*, const double*, const double*);
extern void   dsteqr_(const char*, const int*, double*, double*, double*, const int*, double*, int*);
extern void   dgehrd_(const int*, const int*, const int*, double*, const int*, double*, double*,
                      const int*, int*);
extern void   dorghr_(const int*, const int*, const int*, double*, const int*, const double*, double*,
                      const int*, int*);
extern void   dhseqr_(const char*, const char*, const int*, const int*, const int*, double*,
                      const int*, double*, double*, double*, const int*, double*, const int*, int*);
extern void   dtrexc_(const char*, const int*, double*, const int*, double*, const int*, int*, int*,
                      double*, int*);
extern void   dtrevc_(const char*, const char*, int*, const int*, const double*, const int*, double*,
                      const int*, double*, const int*, const int*, int*, double*, int*);
extern void   dlascl_(const char*, const int*, const int*, const double*, const double*, const int*,
                      const int*, double*, const int*, int*);

#define OR_ERR(code, ...) do { fprintf(stderr, "ks_oracle: " __VA_ARGS__); fputc('\n', stderr); return (code); } while (0)
#define OR_CALL(x) do { int _e = (x); if (_e) return _e; } while (0)
#define MACH_EPS DBL_EPSILON

static double wtime(void) { return omp_get_wtime(); }

/* ------------------------------------------------------------------ RNG */
static unsigned long long splitmix64(unsigned long long z)
{
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void or_random_uniform(unsigned long long seed, int n, double *out)
{
  unsigned long long base = seed * 0x9E3779B97F4A7C15ULL;
  for (int i = 0; i < n; i++) {
    unsigned long long r = splitmix64(base + (unsigned long long)i);
    out[i] = 2.0 * ((double)(r >> 11) * 0x1.0p-53) - 1.0;
  }
}

/* ------------------------------------------------------------------ SpMV */
int or_spmv(int n, const int *rowptr, const int *colidx, const double *val,
            const double *x, double *y)
{
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    double s = 0.0;
    for (int p = rowptr[i]; p < rowptr[i + 1]; p++) s += val[p] * x[colidx[p]];
    y[i] = s;
  }
  return 0;
}

/* ------------------------------------------------------------------ sorting (slepcsc.c) */
typedef struct { int which; double target; int map; double sigma; } or_sc;   /* map: compare sigma + 1/theta */

/* SlepcCompare* (slepcsc.c:152-190): result < 0 means a goes first */
static int sc_compare(const or_sc *sc, double ar, double ai, double br, double bi)
{
  if (sc->map) {                                                   /* SlepcMap_ST (slepcsc.c:118-126) */
    double na = ar * ar + ai * ai, nb = br * br + bi * bi;
    if (!ai) ar = 1.0 / ar + sc->sigma; else { ar = ar / na + sc->sigma; ai = -ai / na; }
    if (!bi) br = 1.0 / br + sc->sigma; else { br = br / nb + sc->sigma; bi = -bi / nb; }
  }
  double a = 0, b = 0;
  switch (sc->which) {
  case OR_LARGEST_MAGNITUDE:  a = hypot(ar, ai); b = hypot(br, bi); return a < b ? 1 : (a > b ? -1 : 0);
  case OR_SMALLEST_MAGNITUDE: a = hypot(ar, ai); b = hypot(br, bi); return a > b ? 1 : (a < b ? -1 : 0);
  case OR_LARGEST_REAL:       a = ar; b = br; return a < b ? 1 : (a > b ? -1 : 0);
  case OR_SMALLEST_REAL:      a = ar; b = br; return a > b ? 1 : (a < b ? -1 : 0);
  case OR_LARGEST_IMAGINARY:  a = fabs(ai); b = fabs(bi); return a < b ? 1 : (a > b ? -1 : 0);
  case OR_SMALLEST_IMAGINARY: a = fabs(ai); b = fabs(bi); return a > b ? 1 : (a < b ? -1 : 0);
  case OR_TARGET_MAGNITUDE:   a = hypot(ar - sc->target, ai); b = hypot(br - sc->target, bi);
                              return a > b ? 1 : (a < b ? -1 : 0);
  case OR_TARGET_REAL:        a = fabs(ar - sc->target); b = fabs(br - sc->target);
                              return a > b ? 1 : (a < b ? -1 : 0);
  }
  return 0;
}

/* SlepcSortEigenvalues (slepcsc.c:89-150) */
static void sort_eigenvalues(const or_sc *sc, int n, const double *eigr, const double *eigi, int *perm)
{
  for (int i = n - 1; i >= 0; i--) {
    double re = eigr[perm[i]], im = eigi[perm[i]];
    int j = i + 1, tmp;
    if (im != 0) { i--; im = eigi[perm[i]]; }
    while (j < n) {
      int result = sc_compare(sc, re, im, eigr[perm[j]], eigi[perm[j]]);
      if (result <= 0) break;
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

/* ------------------------------------------------------------------ BV (BVMAT semantics) */
typedef struct {
  int n, m, nc, ld;         /* local rows, columns, constraints, leading dimension */
  int l, k;                 /* active columns (l may be -nc) */
  double *V;                /* ld x (nc+m), constraints first */
  double *buffer;           /* (nc+m) x m coefficient buffer (bvbasic.c:755-786) */
  int refine;
  double eta;
  long passes;
  unsigned long long seed, rcount;
} obv;

#define COL(bv, j) ((bv)->V + (size_t)((bv)->nc + (j)) * (size_t)(bv)->ld)

/* BVDotVec_Mat -> BVDotVec_BLAS_Private (bvmat.c:110-130, bvblas.c:240-261) */
static void bv_dotvec(obv *bv, const double *y, double *q)
{
  int n = bv->n, k = bv->k - bv->l, one = 1, ld = bv->ld;
  double done = 1.0, zero = 0.0;
  if (n && k) dgemv_("C", &n, &k, &done, COL(bv, bv->l), &ld, y, &one, &zero, q, &one);
  else for (int i = 0; i < k; i++) q[i] = 0.0;
}

/* BVMultVec_Mat -> BVMultVec_BLAS_Private (bvmat.c:38-53, bvblas.c:56-67) */
static void bv_multvec(obv *bv, double alpha, double beta, double *y, const double *q)
{
  int n = bv->n, k = bv->k - bv->l, one = 1, ld = bv->ld;
  if (n) dgemv_("N", &n, &k, &alpha, COL(bv, bv->l), &ld, q, &one, &beta, y, &one);
}

/* BVNorm_Mat for one column, NORM_2 (bvmat.c:165-176, bvlapack.c:49) */
static double bv_normcol(obv *bv, int j)
{
  int n = bv->n, one = 1;
  return dnrm2_(&n, COL(bv, j), &one);
}

/* BVScale_BLAS_Private (bvblas.c:266-278) */
static void bv_scalecol(obv *bv, int j, double alpha)
{
  double *x = COL(bv, j);
  if (alpha == 0.0) { memset(x, 0, sizeof(double) * (size_t)bv->n); return; }
  if (alpha == 1.0) return;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < bv->n; i++) x[i] *= alpha;
}

static void bv_copycol(obv *bv, int j, int i)
{
  memcpy(COL(bv, i), COL(bv, j), sizeof(double) * (size_t)bv->n);
}

/* BVSetRandomColumn: uniform(-1,1) from the counter generator */
static void bv_randomcol(obv *bv, int j)
{
  or_random_uniform(bv->seed + 1000003ULL * (++bv->rcount), bv->n, COL(bv, j));
}

/* BVMultInPlace_BLAS_Private (bvblas.c:74-106): row-blocked gemm through a work panel */
static void multinplace_blas(int m, int k, int s, int e, double *A, int lda, const double *B, int ldb)
{
  const int bs = 64;
  int n = e - s;
  if (m <= 0 || n <= 0) return;
  int nblk = (m + bs - 1) / bs;
#pragma omp parallel
  {
    double *work = (double*)malloc(sizeof(double) * (size_t)bs * (size_t)n);
    double one = 1.0, zero = 0.0;
    const double *pb = B + (size_t)s * ldb;
#pragma omp for schedule(static)
    for (int b = 0; b < nblk; b++) {
      int r0 = b * bs, rows = (r0 + bs <= m) ? bs : m - r0;
      dgemm_("N", "N", &rows, &n, &k, &one, A + r0, &lda, pb, &ldb, &zero, work, &rows);
      for (int j = 0; j < n; j++) memcpy(A + (size_t)(s + j) * lda + r0, work + (size_t)j * rows, sizeof(double) * rows);
    }
    free(work);
  }
}

/* BVMultInPlace_Mat (bvmat.c:55-71) */
static void bv_multinplace(obv *bv, const double *Q, int ldq, int s, int e)
{
  if (s >= e || !bv->n) return;
  multinplace_blas(bv->n, bv->k - bv->l, s - bv->l, e - bv->l, COL(bv, bv->l), bv->ld,
                   Q + (size_t)bv->l * ldq + bv->l, ldq);
}

int or_multinplace(int n, int ld, double *V, int nc, int l, int k,
                   const double *Q, int ldq, int s, int e)
{
  or_mkl_layer();
  obv bv = {0};
  bv.n = n; bv.ld = ld; bv.nc = nc; bv.V = V; bv.l = l; bv.k = k;
  bv_multinplace(&bv, Q, ldq, s, e);
  return 0;
}

/* BV_SafeSqrt (bvimpl.h:121-141), definite inner product */
static int safe_sqrt(double alpha, double *res)
{
  if (!(alpha > -10 * MACH_EPS)) OR_ERR(-62, "Invalid inner product: %g", alpha);
  *res = (alpha < 0.0) ? 0.0 : sqrt(alpha);
  return 0;
}

/* BVOrthogonalizeCGS1 (bvorthog.c:91-132) on column j, h = buffer column j, c = buffer column 0 */
static int cgs1(obv *bv, int j, double *onorm, double *norm)
{
  int nc = bv->nc, ldb = bv->nc + bv->m, ksave;
  double *c = bv->buffer, *h = bv->buffer + (size_t)j * ldb, beta = 0.0;
  bv->passes++;
  bv->k = j;
  if (onorm || norm) {
    ksave = bv->k; bv->k = j + 1;                  /* BVDotColumnInc (bvorthog.c:32-47) */
    bv_dotvec(bv, COL(bv, j), c);
    bv->k = ksave;
    OR_CALL(safe_sqrt(c[nc + j], &beta));          /* BV_SquareRoot */
  } else bv_dotvec(bv, COL(bv, j), c);
  bv_multvec(bv, -1.0, 1.0, COL(bv, j), c);        /* BVMultColumn */
  if (onorm) *onorm = beta;
  if (norm) {
    double sum = 0.0;
    for (int i = 0; i < nc + j; i++) sum += c[i] * c[i];   /* BV_SquareSum */
    *norm = beta * beta - sum;
    if (*norm <= 0.0) *norm = bv_normcol(bv, j);
    else *norm = sqrt(*norm);
  }
  for (int i = 0; i < nc + j; i++) h[i] += c[i];   /* BV_AddCoefficients */
  return 0;
}

/* BVOrthogonalizeGS (bvorthog.c:145-217), column mode (v == NULL) */
static int orthogonalize_gs(obv *bv, int j, double *norm, int *lindep)
{
  int nc = bv->nc, ldb = bv->nc + bv->m, l;
  double onrm = 0.0, nrm = 0.0, *h = bv->buffer + (size_t)j * ldb;
  int dolindep = lindep != NULL;
  for (int i = 0; i < nc + j; i++) h[i] = 0.0;     /* BV_CleanCoefficients */
  switch (bv->refine) {
  case OR_REFINE_IFNEEDED:
    OR_CALL(cgs1(bv, j, &onrm, &nrm));
    l = 1;
    while (l < 3 && nrm && fabs(nrm) < bv->eta * fabs(onrm)) {
      l++;
      OR_CALL(cgs1(bv, j, &onrm, &nrm));
    }
    if (dolindep) *lindep = !(nrm && fabs(nrm) >= bv->eta * fabs(onrm));
    break;
  case OR_REFINE_NEVER:
    OR_CALL(cgs1(bv, j, NULL, NULL));
    if (norm || dolindep) nrm = bv_normcol(bv, j);
    if (dolindep) *lindep = !nrm;
    break;
  case OR_REFINE_ALWAYS:
    OR_CALL(cgs1(bv, j, NULL, NULL));
    OR_CALL(cgs1(bv, j, dolindep ? &onrm : NULL, (norm || dolindep) ? &nrm : NULL));
    if (dolindep) *lindep = !(nrm && fabs(nrm) >= bv->eta * fabs(onrm));
    break;
  }
  if (norm) {
    *norm = nrm;
    h[nc + j] = (dolindep && *lindep) ? 0.0 : nrm;   /* BV_SetValue */
  }
  return 0;
}

/* BVOrthogonalizeColumn (bvorthog.c:285-320): l = -nc, k saved */
static int orthogonalize_column(obv *bv, int j, double *norm, int *lindep)
{
  int ksave = bv->k, lsave = bv->l;
  bv->l = -bv->nc;
  OR_CALL(orthogonalize_gs(bv, j, norm, lindep));
  bv->k = ksave; bv->l = lsave;
  return 0;
}

/* BVOrthonormalizeColumn (bvorthog.c:380-427), replace = FALSE */
static int orthonormalize_column(obv *bv, int j, double *norm, int *lindep)
{
  double nrm;
  int lnd = 0;
  OR_CALL(orthogonalize_column(bv, j, &nrm, &lnd));
  if (nrm != 1.0 && nrm != 0.0) bv_scalecol(bv, j, 1.0 / nrm);
  if (norm) *norm = nrm;
  if (lindep) *lindep = lnd;
  return 0;
}

int or_orthonormalize_column(int n, int ld, double *V, int nc, int j, int refine,
                             double eta, double *H, double *norm, int *lindep, int *npasses)
{
  or_mkl_layer();
  obv bv = {0};
  int m = j + 1;
  bv.n = n; bv.ld = ld; bv.nc = nc; bv.m = m; bv.V = V;
  bv.refine = refine; bv.eta = eta > 0 ? eta : 0.7071;
  bv.buffer = (double*)calloc((size_t)(nc + m) * m, sizeof(double));
  if (!bv.buffer) return -1;
  int e = orthonormalize_column(&bv, j, norm, lindep);
  if (!e && H) memcpy(H, bv.buffer + (size_t)j * (nc + m), sizeof(double) * (size_t)(nc + j + 1));
  if (npasses) *npasses = (int)bv.passes;
  free(bv.buffer);
  return e;
}

/* ------------------------------------------------------------------ block orthogonalization */
/* BVDot_BLAS_Private (bvblas.c:199-233): M(ly:ky, lx:kx) = Y(:, ly:ky)^T X(:, lx:kx) as one gemm */
static void bv_dot(const double *X, int ldx, int lx, int kx, const double *Y, int ldy, int ly, int ky, int n,
                   double *M, int ldm)
{
  int m = ky - ly, k = kx - lx;
  double one = 1.0, zero = 0.0;
  if (m <= 0 || k <= 0) return;
  if (n == 0) {
    for (int j = 0; j < k; j++) for (int i = 0; i < m; i++) M[(ly + i) + (size_t)(lx + j) * ldm] = 0.0;
    return;
  }
  dgemm_("C", "N", &m, &k, &n, &one, Y + (size_t)ly * ldy, &ldy, X + (size_t)lx * ldx, &ldx, &zero,
         M + (size_t)lx * ldm + ly, &ldm);
}

int or_bv_dot(int n, const double *X, int ldx, int lx, int kx, const double *Y, int ldy, int ly, int ky,
              double *M, int ldm)
{
  or_mkl_layer();
  bv_dot(X, ldx, lx, kx, Y, ldy, ly, ky, n, M, ldm);
  return 0;
}

/* BVOrthogonalize (bvorthog.c:726-767) on V (n x k, leading dimension ld, no
   constraints): columns 0..l-1 kept, l..k-1 orthonormalized; R (ldr x k) gets
   columns l..k-1 of the factor.  block: 0 GS (bvorthog.c:507-547), 1 CHOL
   (bvorthog.c:581-596 with BVMatCholInv_LAPACK_Private bvlapack.c:136-202),
   2 TSQR and 3 TSQRCHOL (bvorthog.c:615-663; on one process the TSQR of
   bvlapack.c:347-406 is dgeqrf + dorgqr of the panel), 4 SVQB (bvorthog.c:646-663 with BVMatSVQB_LAPACK_Private bvlapack.c:261-345);
   the leading columns by block Gram-Schmidt first (bvorthog.c:489-502). */
int or_bv_orthogonalize(int n, int ld, double *V, int l, int k, int block, double *R, int ldr)
{
  or_mkl_layer();
  if (l >= k) return 0;
  if (block == 0) {
    obv bv = {0};
    bv.n = n; bv.ld = ld; bv.nc = 0; bv.m = k; bv.V = V;
    bv.refine = 0; bv.eta = 0.7071;
    bv.buffer = (double*)calloc((size_t)k * k, sizeof(double));
    if (!bv.buffer) return -1;
    for (int j = l; j < k; j++) {
      double norm = 0.0;
      int e = orthogonalize_column(&bv, j, &norm, NULL);
      if (e || norm == 0.0) { free(bv.buffer); return e ? e : -2; }
      if (R) {
        for (int i = 0; i < j; i++) R[i + (size_t)j * ldr] = bv.buffer[(size_t)j * k + i];
        R[j + (size_t)j * ldr] = norm;
      }
      bv_scalecol(&bv, j, 1.0 / norm);
    }
    free(bv.buffer);
    return 0;
  }
  if (block < 1 || block > 4) return -3;
  const int nn = k - l;
  double *W = (double*)calloc((size_t)k * k, sizeof(double));     /* the buffer as a k x k matrix */
  double *S = (double*)calloc((size_t)k * k, sizeof(double));
  if (!W || !S) { free(W); free(S); return -1; }
  obv bv = {0};
  bv.n = n; bv.ld = ld; bv.nc = 0; bv.m = k; bv.V = V;
  if (l > 0) {   /* R12 = V1^T V2, V2 -= V1 R12 */
    bv_dot(V, ld, l, k, V, ld, 0, l, n, W, k);
    for (int j = l; j < k; j++) {
      for (int i = 0; i < l; i++) S[i + (size_t)j * k] = -W[i + (size_t)j * k];
      S[j + (size_t)j * k] = 1.0;
    }
    bv.l = 0; bv.k = k;
    bv_multinplace(&bv, S, k, l, k);
    memset(S, 0, sizeof(double) * (size_t)k * k);
  }
  int nn_ = nn, ldk = k, info = 0;
  if (block == 2 || block == 3) {
    /* TSQR on one process (bvorthog.c:615-663, bvlapack.c:347-406 and 510-560 with
       size 1): a Householder QR of the panel, dgeqrf; R is its upper triangle.
       TSQR forms Q with dorgqr, TSQRCHOL keeps V and applies R^-1 (BVMatTriInv). */
    int n_ = n, ld_ = ld, lwork = -1;
    double *A = block == 2 ? V + (size_t)l * ld : (double*)malloc(sizeof(double) * (size_t)ld * nn);
    double *tau = (double*)malloc(sizeof(double) * (nn > 0 ? nn : 1)), q = 0.0;
    if (!A || !tau) { free(tau); if (block == 3) free(A); free(W); free(S); return -1; }
    if (block == 3) memcpy(A, V + (size_t)l * ld, sizeof(double) * (size_t)ld * nn);
    dgeqrf_(&n_, &nn_, A, &ld_, tau, &q, &lwork, &info);
    lwork = (int)q;
    if (block == 2) {
      int lw2 = -1;
      dorgqr_(&n_, &nn_, &nn_, A, &ld_, tau, &q, &lw2, &info);
      if ((int)q > lwork) lwork = (int)q;
    }
    double *work = (double*)malloc(sizeof(double) * (size_t)(lwork > 1 ? lwork : 1));
    dgeqrf_(&n_, &nn_, A, &ld_, tau, work, &lwork, &info);
    for (int j = l; j < k; j++)
      for (int i = l; i < k; i++) W[i + (size_t)j * k] = i <= j ? A[(i - l) + (size_t)(j - l) * ld] : 0.0;
    if (block == 2) dorgqr_(&n_, &nn_, &nn_, A, &ld_, tau, work, &lwork, &info);
    free(work);
    free(tau);
    if (block == 3) free(A);
    if (info) { free(W); free(S); return -7; }
    if (block == 3) {   /* S(l:k, l:k) = R^-1, V2 = V2 S */
      for (int j = l; j < k; j++) memcpy(S + (size_t)j * k + l, W + (size_t)j * k + l, sizeof(double) * nn);
      dtrtri_("U", "N", &nn_, S + (size_t)l * k + l, &ldk, &info);
      if (info) { free(W); free(S); return -5; }
      bv.l = l; bv.k = k;
      bv_multinplace(&bv, S, k, l, k);
    }
    if (R)
      for (int j = l; j < k; j++)
        for (int i = 0; i <= j; i++) R[i + (size_t)j * ldr] = W[i + (size_t)j * k];
    free(W);
    free(S);
    return 0;
  }
  bv_dot(V, ld, l, k, V, ld, l, k, n, W, k);   /* W(l:k, l:k) = V2^T V2 */
  double *g = W + (size_t)l * k + l;
  if (block == 1) {
    for (int j = l; j < k; j++) memcpy(S + (size_t)j * k + l, W + (size_t)j * k + l, sizeof(double) * nn);
    dpotrf_("U", &nn_, g, &ldk, &info);
    if (info) {
      for (int j = l; j < k; j++) {
        memcpy(W + (size_t)j * k + l, S + (size_t)j * k + l, sizeof(double) * nn);
        W[j + (size_t)j * k] += 50.0 * DBL_EPSILON;
      }
      dpotrf_("U", &nn_, g, &ldk, &info);
    }
    if (info) { free(W); free(S); return -4; }
    memset(S, 0, sizeof(double) * (size_t)k * k);
    for (int j = l; j < k; j++) memcpy(S + (size_t)j * k + l, W + (size_t)j * k + l, sizeof(double) * nn);
    dtrtri_("U", "N", &nn_, S + (size_t)l * k + l, &ldk, &info);
    if (info) { free(W); free(S); return -5; }
    for (int j = l; j < k; j++)
      for (int i = j + 1; i < k; i++) { W[i + (size_t)j * k] = 0.0; S[i + (size_t)j * k] = 0.0; }
  } else {
    double *D = (double*)malloc(sizeof(double) * nn), *eig = (double*)malloc(sizeof(double) * nn), a = 0.0;
    int lwork = -1;
    for (int i = l; i < k; i++) D[i - l] = 1.0 / sqrt(W[i + (size_t)i * k]);
    for (int i = l; i < k; i++) for (int j = l; j < k; j++) S[i + (size_t)j * k] = W[i + (size_t)j * k] * D[i - l];
    for (int j = l; j < k; j++) for (int i = l; i < k; i++) S[i + (size_t)j * k] *= D[j - l];
    dsyev_("V", "L", &nn_, S + (size_t)l * k + l, &ldk, eig, &a, &lwork, &info);
    lwork = (int)a;
    double *work = (double*)malloc(sizeof(double) * (size_t)(lwork > 1 ? lwork : 1));
    dsyev_("V", "L", &nn_, S + (size_t)l * k + l, &ldk, eig, work, &lwork, &info);
    free(work);
    if (info) { free(D); free(eig); free(W); free(S); return -6; }
    for (int i = l; i < k; i++) for (int j = l; j < k; j++) W[i + (size_t)j * k] = S[j + (size_t)i * k];   /* R = U' */
    for (int i = l; i < k; i++) for (int j = l; j < k; j++) S[i + (size_t)j * k] *= D[i - l];
    for (int j = l; j < k; j++) for (int i = l; i < k; i++) S[i + (size_t)j * k] /= sqrt(eig[j - l]);
    for (int i = l; i < k; i++) for (int j = l; j < k; j++) W[i + (size_t)j * k] *= sqrt(eig[i - l]);
    for (int j = l; j < k; j++) for (int i = l; i < k; i++) W[i + (size_t)j * k] /= D[j - l];
    free(D);
    free(eig);
  }
  bv.l = l; bv.k = k;
  bv_multinplace(&bv, S, k, l, k);   /* V2 = V2 S */
  if (R)   /* BV_StoreCoeffsBlock_Default (bvorthog.c:567-579) */
    for (int j = l; j < k; j++) {
      const int rows = block == 1 ? j + 1 : k;
      for (int i = 0; i < rows; i++) R[i + (size_t)j * ldr] = W[i + (size_t)j * k];
    }
  free(W);
  free(S);
  return 0;
}

/* The reference's TSQR over P processes (bvlapack.c:347-466): each process
   factors its rows (PetscSplitOwnership: n/P rows, one more for the first n%P),
   the P triangular factors are reduced to one, and each process's Q block is
   its local Q times its slice of the reduction's Q.  Restated with a flat
   reduction (one QR of the P stacked factors) in place of the MPI butterfly:
   the same factorization up to the signs of R's rows, other rounding.  V is
   n x k (leading dimension ld), l = 0; R (ldr x k) gets the factor. */
int or_bv_tsqr_split(int n, int ld, double *V, int k, int P, double *R, int ldr)
{
  or_mkl_layer();
  if (P < 1 || k < 1 || n < P * k) return -3;
  const int m2 = P * k;
  double *S = (double*)calloc((size_t)m2 * k, sizeof(double));      /* stacked factors */
  double *tau = (double*)malloc(sizeof(double) * k);
  int *r0 = (int*)malloc(sizeof(int) * (P + 1));
  if (!S || !tau || !r0) { free(S); free(tau); free(r0); return -1; }
  r0[0] = 0;
  for (int p = 0; p < P; p++) r0[p + 1] = r0[p] + n / P + (p < n % P ? 1 : 0);
  int info = 0, lwork = 64 * k + 64;
  double *work = (double*)malloc(sizeof(double) * lwork);
  for (int p = 0; p < P && !info; p++) {
    int mp = r0[p + 1] - r0[p], kk = k;
    double *A = V + r0[p];
    dgeqrf_(&mp, &kk, A, &ld, tau, work, &lwork, &info);
    for (int j = 0; j < k; j++)
      for (int i = 0; i <= j; i++) S[(size_t)p * k + i + (size_t)j * m2] = A[i + (size_t)j * ld];
    if (!info) dorgqr_(&mp, &kk, &kk, A, &ld, tau, work, &lwork, &info);
  }
  int m2_ = m2, kk = k;
  if (!info) dgeqrf_(&m2_, &kk, S, &m2_, tau, work, &lwork, &info);
  if (!info && R)
    for (int j = 0; j < k; j++)
      for (int i = 0; i < k; i++) R[i + (size_t)j * ldr] = i <= j ? S[i + (size_t)j * m2] : 0.0;
  if (!info) dorgqr_(&m2_, &kk, &kk, S, &m2_, tau, work, &lwork, &info);
  /* V block p = Q_p S(p k .. p k + k - 1, :) */
  double *B = (double*)malloc(sizeof(double) * (size_t)(n / P + 1) * k);
  for (int p = 0; p < P && !info && B; p++) {
    int mp = r0[p + 1] - r0[p];
    for (int j = 0; j < k; j++)
      for (int i = 0; i < mp; i++) {
        double a = 0.0;
        for (int c = 0; c < k; c++) a += V[r0[p] + i + (size_t)c * ld] * S[(size_t)p * k + c + (size_t)j * m2];
        B[i + (size_t)j * mp] = a;
      }
    for (int j = 0; j < k; j++)
      for (int i = 0; i < mp; i++) V[r0[p] + i + (size_t)j * ld] = B[i + (size_t)j * mp];
  }
  int rc = info ? -7 : (B ? 0 : -1);
  free(B); free(work); free(S); free(tau); free(r0);
  return rc;
}

/* ------------------------------------------------------------------ DS */
enum { DS_RAW = 0, DS_INTERMEDIATE = 1, DS_CONDENSED = 2, DS_TRUNCATED = 3 };
enum { DS_HEP = 1, DS_NHEP = 2, DS_NHEPTS = 3 };

typedef struct {
  int type, ld, n, l, k, t, state, compact, extrarow;
  double *A, *T, *Q, *X, *work;
  double *B, *Z, *Y, *wr2, *wi2;        /* DSNHEPTS: the problem from the transpose */
  int *perm, *iwork;
  or_sc sc;
} ods;

/* DSNHEPTS runs the DSNHEP routines on (B, Z, Y) by exchanging them with (A, Q, X) */
static void ds_swap_sides(ods *ds)
{
  double *t;
  t = ds->A; ds->A = ds->B; ds->B = t;
  t = ds->Q; ds->Q = ds->Z; ds->Z = t;
  t = ds->X; ds->X = ds->Y; ds->Y = t;
}

static int ds_alloc(ods *ds, int type, int ld)
{
  memset(ds, 0, sizeof(*ds));
  ds->type = type; ds->ld = ld;
  ds->A = (double*)calloc((size_t)ld * ld, sizeof(double));
  ds->T = (double*)calloc((size_t)ld * 3, sizeof(double));
  ds->Q = (double*)calloc((size_t)ld * ld, sizeof(double));
  ds->X = (double*)calloc((size_t)ld * ld, sizeof(double));
  ds->work = (double*)calloc((size_t)ld * ld + 8 * (size_t)ld, sizeof(double));
  ds->perm = (int*)calloc((size_t)ld, sizeof(int));
  ds->iwork = (int*)calloc((size_t)ld, sizeof(int));
  if (!ds->A || !ds->T || !ds->Q || !ds->X || !ds->work || !ds->perm || !ds->iwork) return -1;
  if (type == DS_NHEPTS) {
    ds->B = (double*)calloc((size_t)ld * ld, sizeof(double));
    ds->Z = (double*)calloc((size_t)ld * ld, sizeof(double));
    ds->Y = (double*)calloc((size_t)ld * ld, sizeof(double));
    ds->wr2 = (double*)calloc((size_t)ld, sizeof(double));
    ds->wi2 = (double*)calloc((size_t)ld, sizeof(double));
    if (!ds->B || !ds->Z || !ds->Y || !ds->wr2 || !ds->wi2) return -1;
  }
  return 0;
}

static void ds_free(ods *ds)
{
  free(ds->A); free(ds->T); free(ds->Q); free(ds->X); free(ds->work); free(ds->perm); free(ds->iwork);
  free(ds->B); free(ds->Z); free(ds->Y); free(ds->wr2); free(ds->wi2);
}

/* DSSetDimensions (dsops.c:130-160) */
static void ds_setdims(ods *ds, int n, int l, int k)
{
  ds->n = n; ds->t = n; ds->l = l; ds->k = k;
}

/* DSSetIdentity (dspriv.c:377-395) */
static void ds_setidentity(ods *ds, double *M)
{
  memset(M + (size_t)ds->l * ds->ld, 0, sizeof(double) * (size_t)ds->ld * (ds->n - ds->l));
  for (int i = ds->l; i < ds->n; i++) M[i + (size_t)i * ds->ld] = 1.0;
}

/* DSArrowTridiag (dshep.c:221-262): Rutishauser reduction of an arrowhead */
static void arrow_tridiag(int n, double *d, double *e, double *Q, int ld)
{
  int one = 1;
  double c, s, p, off, temp;
  if (n <= 2) return;
  for (int j = 0; j < n - 2; j++) {
    temp = e[j + 1];
    dlartg_(&temp, &e[j], &c, &s, &e[j + 1]);
    s = -s;
    temp = d[j + 1];
    e[j] = c * s * (temp - d[j]);
    d[j + 1] = s * s * d[j] + c * c * temp;
    d[j] = c * c * d[j] + s * s * temp;
    int j2 = j + 2;
    drot_(&j2, Q + (size_t)j * ld, &one, Q + (size_t)(j + 1) * ld, &one, &c, &s);
    for (int i = j - 1; i >= 0; i--) {
      off = -s * e[i];
      e[i] = c * e[i];
      temp = e[i + 1];
      dlartg_(&temp, &off, &c, &s, &e[i + 1]);
      s = -s;
      temp = (d[i] - d[i + 1]) * s - 2.0 * c * e[i];
      p = s * temp;
      d[i + 1] += p;
      d[i] -= p;
      e[i] = -e[i] - c * temp;
      drot_(&j2, Q + (size_t)i * ld, &one, Q + (size_t)(i + 1) * ld, &one, &c, &s);
    }
  }
}

/* DSSolve_HEP_QR (dshep.c:383-426) with DSIntermediate_HEP compact (dshep.c:267-321) */
static int ds_solve_hep(ods *ds, double *wr, double *wi)
{
  int n = ds->n, l = ds->l, ld = ds->ld, n1, info;
  double *d = ds->T, *e = ds->T + ld;
  if (!ds->compact) OR_ERR(-56, "oracle DSHEP supports compact storage only");
  ds_setidentity(ds, ds->Q);
  n1 = ds->k - l + 1; if (n1 < 0) n1 = 0;
  if (ds->state < DS_INTERMEDIATE) arrow_tridiag(n1, d + l, e + l, ds->Q + l + (size_t)l * ld, ld);
  for (int i = 0; i < l; i++) wr[i] = d[i];
  n1 = n - l;
  dsteqr_("V", &n1, d + l, e + l, ds->Q + l + (size_t)l * ld, &ld, ds->work, &info);
  if (info) OR_ERR(-76, "steqr info %d", info);
  for (int i = l; i < n; i++) wr[i] = d[i];
  for (int i = 0; i < n - 1; i++) e[i] = 0.0;
  if (wi) for (int i = l; i < n; i++) wi[i] = 0.0;
  return 0;
}

/* DSSortEigenvaluesReal_Private (dspriv.c:224-246) */
static void ds_sort_real(ods *ds, const double *eig, int *perm)
{
  int n = ds->t, l = ds->l;
  for (int i = l + 1; i < n; i++) {
    double re = eig[perm[i]];
    int j = i - 1;
    int result = sc_compare(&ds->sc, re, 0.0, eig[perm[j]], 0.0);
    while (result < 0 && j >= l) {
      int tmp = perm[j]; perm[j] = perm[j + 1]; perm[j + 1] = tmp; j--;
      if (j >= l) result = sc_compare(&ds->sc, re, 0.0, eig[perm[j]], 0.0);
    }
  }
}

/* DSPermuteColumns_Private (dspriv.c:248-270) */
static void ds_permute_columns(ods *ds, int istart, int iend, int n, double *Q, int *perm)
{
  int ld = ds->ld;
  for (int i = istart; i < iend; i++) {
    int p = perm[i];
    if (p != i) {
      int j = i + 1;
      while (perm[j] != i) j++;
      perm[j] = p; perm[i] = i;
      for (int k = 0; k < n; k++) {
        double r = Q[k + (size_t)p * ld]; Q[k + (size_t)p * ld] = Q[k + (size_t)i * ld]; Q[k + (size_t)i * ld] = r;
      }
    }
  }
}

/* DSSort_HEP (dshep.c:323-347) */
static void ds_sort_hep(ods *ds, double *wr)
{
  int n = ds->n, l = ds->l;
  double *d = ds->T;
  for (int i = 0; i < n; i++) ds->perm[i] = i;
  ds_sort_real(ds, d, ds->perm);
  for (int i = l; i < n; i++) wr[i] = d[ds->perm[i]];
  ds_permute_columns(ds, l, n, n, ds->Q, ds->perm);
  for (int i = l; i < n; i++) d[i] = wr[i];
}

/* DSUpdateExtraRow_HEP, compact (dshep.c:349-381) */
static void ds_update_extra_row_hep(ods *ds)
{
  int n = ds->n, ld = ds->ld;
  double *e = ds->T + ld, beta = e[n - 1];
  for (int i = 0; i < n; i++) e[i] = beta * ds->Q[n - 1 + (size_t)i * ld];
  ds->k = n;
}

/* DSSolve_NHEP_Private (dsutil.c:21-87) */
static int ds_solve_nhep(ods *ds, double *wr, double *wi)
{
  int n = ds->n, ld = ds->ld, ilo = ds->l + 1, lwork = 6 * ld, info;
  double *A = ds->A, *Q = ds->Q, *tau = ds->work, *work = ds->work + ld;
  memset(Q, 0, sizeof(double) * (size_t)ld * ld);
  for (int i = 0; i < n; i++) Q[i + (size_t)i * ld] = 1.0;
  if (n == 1) { wr[0] = A[0]; if (wi) wi[0] = 0.0; return 0; }
  if (ds->state < DS_INTERMEDIATE) {
    dgehrd_(&n, &ilo, &n, A, &ld, tau, work, &lwork, &info);
    if (info) OR_ERR(-76, "gehrd info %d", info);
    for (int j = 0; j < n - 1; j++)
      for (int i = j + 2; i < n; i++) { Q[i + (size_t)j * ld] = A[i + (size_t)j * ld]; A[i + (size_t)j * ld] = 0.0; }
    dorghr_(&n, &ilo, &n, Q, &ld, tau, work, &lwork, &info);
    if (info) OR_ERR(-76, "orghr info %d", info);
  }
  dhseqr_("S", "V", &n, &ilo, &n, A, &ld, wr, wi, Q, &ld, work, &lwork, &info);
  for (int j = 0; j < ds->l; j++) {
    if (j == n - 1 || A[j + 1 + (size_t)j * ld] == 0.0) { wr[j] = A[j + (size_t)j * ld]; wi[j] = 0.0; }
    else {
      wr[j] = A[j + (size_t)j * ld]; wr[j + 1] = A[j + (size_t)j * ld];
      wi[j] = sqrt(fabs(A[j + 1 + (size_t)j * ld])) * sqrt(fabs(A[j + (size_t)(j + 1) * ld]));
      wi[j + 1] = -wi[j];
      j++;
    }
  }
  if (info) OR_ERR(-76, "hseqr info %d", info);
  return 0;
}

/* DSSort_NHEP_Total (dsutil.c:93-171) */
static int ds_sort_nhep(ods *ds, double *wr, double *wi)
{
  int n = ds->n, ld = ds->ld, info;
  double *T = ds->A, *Q = ds->Q;
  for (int i = ds->l; i < n - 1; i++) {
    double re = wr[i], im = wi[i];
    int pos = 0, j = i + 1;
    if (im != 0) j = i + 2;
    for (; j < n; j++) {
      if (sc_compare(&ds->sc, re, im, wr[j], wi[j]) > 0) { re = wr[j]; im = wi[j]; pos = j; }
      if (wi[j] != 0) j++;
    }
    if (pos) {
      int ifst = pos + 1, ilst = i + 1;
      dtrexc_("V", &n, T, &ld, Q, &ld, &ifst, &ilst, ds->work, &info);
      if (info) OR_ERR(-76, "trexc info %d", info);
      for (j = i; j < n; j++) {
        wr[j] = T[j + (size_t)j * ld];
        if (j < n - 1 && T[j + 1 + (size_t)j * ld] != 0.0) {
          wi[j] = sqrt(fabs(T[j + 1 + (size_t)j * ld])) * sqrt(fabs(T[j + (size_t)(j + 1) * ld]));
          wr[j + 1] = wr[j]; wi[j + 1] = -wi[j]; j++;
        } else wi[j] = 0.0;
      }
    }
    if (wi[i] != 0) i++;
  }
  return 0;
}

/* DSUpdateExtraRow_NHEP (dsnhep.c:318-340) */
static void ds_update_extra_row_nhep(ods *ds)
{
  int n = ds->n, ld = ds->ld, inc = 1;
  double one = 1.0, zero = 0.0, *x = ds->work, *y = ds->work + ld;
  for (int i = 0; i < n; i++) x[i] = ds->A[n + (size_t)i * ld];
  dgemv_("C", &n, &n, &one, ds->Q, &ld, x, &inc, &zero, y, &inc);
  for (int i = 0; i < n; i++) ds->A[n + (size_t)i * ld] = y[i];
  ds->k = n;
}

/* DSVectors_NHEP_Eigen_Some (dsnhep.c:101-167), right vectors */
static int ds_vectors_nhep_some(ods *ds, int *k, double *rnorm)
{
  int n = ds->n, ld = ds->ld, mm, mout, info, inc = 1, cols = 1, zero = 0, iscomplex = 0;
  double sone = 1.0, szero = 0.0, norm, done = 1.0;
  int *select = ds->iwork;
  double *Y = ds->X + (size_t)(*k) * ld;
  for (int i = 0; i < n; i++) select[i] = 0;
  select[*k] = 1;
  if (*k < n - 1 && ds->A[*k + 1 + (size_t)(*k) * ld] != 0.0) iscomplex = 1;
  mm = iscomplex ? 2 : 1;
  if (iscomplex) select[*k + 1] = 1;
  dtrevc_("R", "S", select, &n, ds->A, &ld, Y, &ld, Y, &ld, &mm, &mout, ds->work, &info);
  if (info) OR_ERR(-76, "trevc info %d", info);
  if (mout != mm) OR_ERR(-60, "Inconsistent arguments");
  if (ds->state >= DS_CONDENSED) {
    double *w = ds->work;
    memcpy(w, Y, sizeof(double) * (size_t)mout * ld);
    dgemv_("N", &n, &n, &sone, ds->Q, &ld, w, &inc, &szero, Y, &inc);
    if (iscomplex) dgemv_("N", &n, &n, &sone, ds->Q, &ld, w + ld, &inc, &szero, Y + ld, &inc);
    norm = dnrm2_(&n, Y, &inc);
    if (iscomplex) { norm = hypot(norm, dnrm2_(&n, Y + ld, &inc)); cols = 2; }
    dlascl_("G", &zero, &zero, &norm, &done, &n, &cols, Y, &ld, &info);
    if (info) OR_ERR(-76, "lascl info %d", info);
  }
  if (iscomplex) (*k)++;
  if (rnorm) *rnorm = iscomplex ? hypot(Y[n - 1], Y[n - 1 + ld]) : fabs(Y[n - 1]);
  return 0;
}

/* DSVectors_NHEP_Eigen_All (dsnhep.c:168-226), right vectors */
static int ds_vectors_nhep_all(ods *ds)
{
  int n = ds->n, ld = ds->ld, mout, info, inc = 1, cols, zero = 0;
  double norm, done = 1.0;
  const char *back = "A";
  if (ds->state >= DS_CONDENSED) { back = "B"; memcpy(ds->X, ds->Q, sizeof(double) * (size_t)ld * ld); }
  dtrevc_("R", back, NULL, &n, ds->A, &ld, NULL, &ld, ds->X, &ld, &n, &mout, ds->work, &info);
  if (info) OR_ERR(-76, "trevc info %d", info);
  for (int i = 0; i < n; i++) {
    int iscomplex = (i < n - 1 && ds->A[i + 1 + (size_t)i * ld] != 0.0);
    cols = 1;
    norm = dnrm2_(&n, ds->X + (size_t)i * ld, &inc);
    if (iscomplex) { norm = hypot(norm, dnrm2_(&n, ds->X + (size_t)(i + 1) * ld, &inc)); cols = 2; }
    dlascl_("G", &zero, &zero, &norm, &done, &n, &cols, ds->X + (size_t)i * ld, &ld, &info);
    if (info) OR_ERR(-76, "lascl info %d", info);
    if (iscomplex) i++;
  }
  return 0;
}

/* DSVectors_HEP (dshep.c:137-172) */
static void ds_vectors_hep(ods *ds, int *j, double *rnorm)
{
  int ld = ds->ld;
  if (j) {
    if (ds->state >= DS_CONDENSED) {
      memcpy(ds->X + (size_t)(*j) * ld, ds->Q + (size_t)(*j) * ld, sizeof(double) * ld);
      if (rnorm) *rnorm = fabs(ds->Q[ds->n - 1 + (size_t)(*j) * ld]);
    } else {
      memset(ds->X + (size_t)(*j) * ld, 0, sizeof(double) * ld);
      ds->X[*j + (size_t)(*j) * ld] = 1.0;
      if (rnorm) *rnorm = 0.0;
    }
  } else {
    if (ds->state >= DS_CONDENSED) memcpy(ds->X, ds->Q, sizeof(double) * (size_t)ld * ld);
    else ds_setidentity(ds, ds->X);
  }
}

/* DSSortWithPermutation_NHEP_Private (dsutil.c:177-236) on (A, Q) */
static int ds_sort_permutation_nhep(ods *ds, int *perm, double *wr, double *wi)
{
  int n = ds->n, ld = ds->ld, info;
  double *T = ds->A;
  for (int i = ds->l; i < n - 1; i++) {
    int pos = perm[i];
    int inc = (pos < n - 1 && T[pos + 1 + (size_t)pos * ld] != 0.0) ? 2 : 1;
    if (pos != i) {
      if (!((pos == 0 || T[pos + (size_t)(pos - 1) * ld] == 0.0 || perm[i + 1] == pos - 1) &&
            (T[pos + 1 + (size_t)pos * ld] == 0.0 || perm[i + 1] == pos + 1)))
        OR_ERR(-72, "Invalid permutation due to a 2x2 block at position %d", pos);
      int ifst = pos + 1, ilst = i + 1;
      dtrexc_("V", &n, T, &ld, ds->Q, &ld, &ifst, &ilst, ds->work, &info);
      if (info) OR_ERR(-76, "trexc info %d", info);
      for (int j = i + 1; j < n; j++) if (perm[j] >= i && perm[j] < pos) perm[j] += inc;
      perm[i] = i;
      if (inc == 2) perm[i + 1] = i + 1;
    }
    if (inc == 2) i++;
  }
  for (int j = ds->l; j < n; j++) {
    wr[j] = T[j + (size_t)j * ld];
    if (j < n - 1 && T[j + 1 + (size_t)j * ld] != 0.0) {
      wi[j] = sqrt(fabs(T[j + 1 + (size_t)j * ld])) * sqrt(fabs(T[j + (size_t)(j + 1) * ld]));
      wr[j + 1] = wr[j]; wi[j + 1] = -wi[j]; j++;
    } else wi[j] = 0.0;
  }
  return 0;
}

/* DSSort_NHEPTS (dsnhepts.c:188-224): both sorted by the criterion, then B's blocks
   permuted so that its eigenvalue i is the one nearest A's eigenvalue i */
static int ds_sort_nhepts(ods *ds, double *wr, double *wi)
{
  int n = ds->n, cont = 0, id = 0, e = 0;
  int *p = (int*)malloc(sizeof(int) * (size_t)ds->ld), *idx = (int*)malloc(sizeof(int) * (size_t)ds->ld),
      *idx2 = (int*)malloc(sizeof(int) * (size_t)ds->ld);
  double *wr2 = ds->wr2, *wi2 = ds->wi2;
  if (!p || !idx || !idx2) { free(p); free(idx); free(idx2); return -55; }
  e = ds_sort_nhep(ds, wr, wi);
  if (!e) { ds_swap_sides(ds); e = ds_sort_nhep(ds, wr2, wi2); ds_swap_sides(ds); }
  if (!e) {
    for (int i = 0; i < n; i++) {
      if (hypot(wr2[i] - wr[i], wi2[i] - wi[i]) > sqrt(DBL_EPSILON)) { idx2[cont] = i; idx[cont++] = i; }
      p[i] = -1;
    }
    if (cont) {
      for (int i = 0; i < cont; i++) {
        double t = DBL_MAX, sd;
        for (int j = 0; j < cont; j++)
          if (idx2[j] != -1 && (sd = hypot(wr2[idx[j]] - wr[idx[i]], wi2[idx[j]] - wi[idx[i]])) < t) { id = j; t = sd; }
        p[idx[i]] = idx[id];
        idx2[id] = -1;
      }
      for (int i = 0; i < n; i++) if (p[i] == -1) p[i] = i;
      ds_swap_sides(ds);
      e = ds_sort_permutation_nhep(ds, p, wr2, wi2);
      ds_swap_sides(ds);
    }
  }
  free(p); free(idx); free(idx2);
  return e;
}

static int ds_solve(ods *ds, double *wr, double *wi)
{
  if (ds->state >= DS_CONDENSED) return 0;
  if (ds->type == DS_HEP) OR_CALL(ds_solve_hep(ds, wr, wi));
  else OR_CALL(ds_solve_nhep(ds, wr, wi));
  if (ds->type == DS_NHEPTS) {                                     /* DSSolve_NHEPTS (dsnhepts.c:241-252) */
    int e;
    ds_swap_sides(ds);
    e = ds_solve_nhep(ds, ds->wr2, ds->wi2);
    ds_swap_sides(ds);
    if (e) return e;
  }
  ds->state = DS_CONDENSED;
  return 0;
}

static int ds_sort(ods *ds, double *wr, double *wi)
{
  for (int i = 0; i < ds->n; i++) ds->perm[i] = i;
  if (ds->type == DS_HEP) { ds_sort_hep(ds, wr); return 0; }
  if (ds->type == DS_NHEPTS) return ds_sort_nhepts(ds, wr, wi);
  return ds_sort_nhep(ds, wr, wi);
}

static void ds_update_extra_row(ods *ds)
{
  if (ds->type == DS_HEP) ds_update_extra_row_hep(ds);
  else ds_update_extra_row_nhep(ds);
  if (ds->type == DS_NHEPTS) { ds_swap_sides(ds); ds_update_extra_row_nhep(ds); ds_swap_sides(ds); }
}

static int ds_vectors(ods *ds, int *j, double *rnorm)
{
  if (ds->type == DS_HEP) { ds_vectors_hep(ds, j, rnorm); return 0; }
  return j ? ds_vectors_nhep_some(ds, j, rnorm) : ds_vectors_nhep_all(ds);
}

/* DS_MAT_Y of DSNHEPTS: right eigenvectors of B (dsnhepts.c:162-186) */
static int ds_vectors_left(ods *ds, int *j, double *rnorm)
{
  int e;
  ds_swap_sides(ds);
  e = j ? ds_vectors_nhep_some(ds, j, rnorm) : ds_vectors_nhep_all(ds);
  ds_swap_sides(ds);
  return e;
}

/* DSGetTruncateSize_Default (dsops.c:329-344); DSNHEPTS looks at B too (dsnhepts.c:349-367) */
static void ds_truncate_size(ods *ds, int l, int n, int *k)
{
  if (ds->type == DS_HEP) return;
  double val = ds->A[(l + *k) + (size_t)(l + *k - 1) * ds->ld];
  if (ds->type == DS_NHEPTS && val == 0.0) val = ds->B[(l + *k) + (size_t)(l + *k - 1) * ds->ld];
  if (val != 0.0) {
    if (l + *k < n - 1) (*k)++;
    else (*k)--;
  }
}

/* DSTruncate_HEP / DSTruncate_NHEP / DSTruncate_NHEPTS (dshep.c:643-670, dsnhep.c:394-425,
   dsnhepts.c:369-398) */
static void ds_truncate(ods *ds, int n, int trim)
{
  int ld = ds->ld, l = ds->l;
  int fullA = (ds->type != DS_HEP || !ds->compact) && ds->extrarow;
  double *B = ds->type == DS_NHEPTS ? ds->B : NULL;
  if (trim) {
    if (fullA)
      for (int i = l; i < ds->n; i++) { ds->A[ds->n + (size_t)i * ld] = 0.0; if (B) B[ds->n + (size_t)i * ld] = 0.0; }
    ds->l = 0; ds->k = 0; ds->n = n; ds->t = n;
  } else {
    if (fullA && ds->k == ds->n) {
      for (int i = l; i < n; i++) {
        ds->A[n + (size_t)i * ld] = ds->A[ds->n + (size_t)i * ld];
        if (B) B[n + (size_t)i * ld] = B[ds->n + (size_t)i * ld];
      }
      for (int i = l; i < ds->n; i++) { ds->A[ds->n + (size_t)i * ld] = 0.0; if (B) B[ds->n + (size_t)i * ld] = 0.0; }
    }
    ds->k = ds->extrarow ? n : 0;
    ds->t = ds->n;
    ds->n = n;
  }
  ds->state = trim ? DS_RAW : DS_TRUNCATED;
}

/* DSTranslateHarmonic_NHEP (dsnhep.c:466-535); g has ld entries */
static int ds_translate_harmonic(ods *ds, double tau, double beta, int recover, double *g, double *gammaout)
{
  int n = ds->n, ld = ds->ld, one = 1, info = 0;
  double *A = ds->A, gamma;
  if (!recover) {
    double *B = (double*)malloc(sizeof(double) * (size_t)ld * ld);
    int *ipiv = (int*)malloc(sizeof(int) * (size_t)ld);
    if (!B || !ipiv) { free(B); free(ipiv); return -55; }
    memcpy(B, A, sizeof(double) * (size_t)ld * ld);
    memset(g, 0, sizeof(double) * (size_t)n);
    g[n - 1] = beta;                                                /* b = beta e_n */
    for (int i = 0; i < n; i++) B[i + (size_t)i * ld] -= tau;
    dgetrf_(&n, &n, B, &ld, ipiv, &info);
    if (!info) dgetrs_("C", &n, &one, B, &ld, ipiv, g, &ld, &info);  /* g = (A - tau I)^-T b */
    free(B); free(ipiv);
    if (info) OR_ERR(-81, "harmonic translation: LAPACK info %d", info);
    for (int i = 0; i < n; i++) A[i + (size_t)(n - 1) * ld] += g[i] * beta;
  } else {
    int ncol = ds->l + ds->k;
    double *ghat = (double*)calloc((size_t)ld, sizeof(double));
    if (!ghat) return -55;
    for (int j = 0; j < ncol; j++) {                                 /* ghat = -Q(:,idx)^T g */
      double t = 0.0;
      for (int i = 0; i < n; i++) t += ds->Q[i + (size_t)j * ld] * g[i];
      ghat[j] = -t;
    }
    for (int i = 0; i < ncol; i++)                                  /* A += ghat b^T Q(:,l:l+k) */
      for (int j = ds->l; j < ds->l + ds->k; j++)
        A[i + (size_t)j * ld] += ghat[i] * ds->Q[n - 1 + (size_t)j * ld] * beta;
    for (int i = 0; i < n; i++) {                                   /* g~ = g + Q(:,idx) ghat */
      double t = 0.0;
      for (int j = 0; j < ncol; j++) t += ds->Q[i + (size_t)j * ld] * ghat[j];
      g[i] += t;
    }
    free(ghat);
  }
  gamma = hypot(1.0, dnrm2_(&n, g, &one));
  if (gammaout) *gammaout = gamma;
  if (recover && ds->extrarow)
    for (int j = ds->l; j < ds->l + ds->k; j++) A[ds->n + (size_t)j * ld] *= gamma;
  ds->state = DS_RAW;
  return 0;
}

/* the translation alone, on an (n+1) x n Hessenberg H with H(n, n-1) = beta (ld >= n+1):
   A gets H + g b^T, g and gamma are returned */
int or_ds_translate_harmonic(int ld, int n, double *A, double tau, double beta, double *g, double *gamma)
{
  or_mkl_layer();
  ods ds;
  int e;
  if (ds_alloc(&ds, DS_NHEP, ld)) return -55;
  ds.extrarow = 1;
  ds_setdims(&ds, n, 0, 0);
  memcpy(ds.A, A, sizeof(double) * (size_t)ld * ld);
  e = ds_translate_harmonic(&ds, tau, beta, 0, g, gamma);
  if (!e) memcpy(A, ds.A, sizeof(double) * (size_t)ld * ld);
  ds_free(&ds);
  return e;
}

int or_ds_run(int type, int ld, int n, int l, int k, int compact, int extrarow,
              int state, double *A, double *T, int which, double *wr, double *wi,
              double *Q, int jvec, double *rnorm, double *X)
{
  or_mkl_layer();
  ods ds;
  double *wi_local = NULL;
  if (ds_alloc(&ds, type, ld)) return -1;
  ds.compact = compact; ds.extrarow = extrarow; ds.state = state;
  ds.sc.which = which;
  ds_setdims(&ds, n, l, k);
  if (A) memcpy(ds.A, A, sizeof(double) * (size_t)ld * ld);
  if (T) memcpy(ds.T, T, sizeof(double) * (size_t)ld * 3);
  if (!wi) { wi_local = (double*)calloc((size_t)ld, sizeof(double)); wi = wi_local; }
  int e = ds_solve(&ds, wr, wi);
  if (!e) e = ds_sort(&ds, wr, wi);
  if (!e && extrarow) ds_update_extra_row(&ds);
  if (!e && jvec >= 0) { int j = jvec; e = ds_vectors(&ds, &j, rnorm); }
  if (!e) e = ds_vectors(&ds, NULL, NULL);
  if (!e) {
    if (A) memcpy(A, ds.A, sizeof(double) * (size_t)ld * ld);
    if (T) memcpy(T, ds.T, sizeof(double) * (size_t)ld * 3);
    if (Q) memcpy(Q, ds.Q, sizeof(double) * (size_t)ld * ld);
    if (X) memcpy(X, ds.X, sizeof(double) * (size_t)ld * ld);
  }
  free(wi_local);
  ds_free(&ds);
  return e;
}

/* ------------------------------------------------------------------ EPS + Krylov-Schur */
typedef struct {
  int n, nev, ncv, mpd, max_it, which, hermitian, lock, nconv, its, reason;
  int harmonic, dsherm;           /* dsherm: Lanczos + DSHEP (Hermitian and Ritz extraction) */
  int twosided;
  int sinvert, pc_jacobi, ksp_maxits;   /* shift-and-invert: (A - sigma I)^-1 by PCG or BiCGStab */
  int ksp_bcgs, pc_ilu, ksp_lu;
  double sigma, ksp_rtol, *dinv, *kwork;
  double *dlu;                    /* KSPPREONLY + PCLU: the dense LU factors of A - sigma I (n x n, by columns) */
  int *ipiv;                      /* KSPPREONLY + PCLU: row k swapped with row ipiv[k] at step k */
  int *lu_diag;                   /* PCILU: index of the diagonal entry of each row */
  double *lu;                     /* PCILU: the ILU(0) factors of A - sigma I on A's pattern */
  long ksp_its;
  int ksp_failed;
  double tol, target, keep;
  const int *rowptr, *colidx;
  const double *val;
  int *trowptr, *tcolidx;         /* A^T (two-sided) */
  double *tval, *M;               /* M = W^T V, ncv x ncv (two-sided) */
  obv V, W;
  ods ds;
  double *eigr, *eigi, *errest;
  int *perm;
  or_sc sc;
  long nmatmult;
  double t_ds;
} oeps;

/* The preconditioned conjugate gradient method in PETSc's KSPCG form (Hestenes-Stiefel
   recurrences; convergence on the preconditioned residual norm |B r| <= max(rtol |B b|,
   abstol), abstol 1e-50) with B = diag(A - sigma I)^-1 (PCJACOBI: zero entries -> 1) or
   B = I.  PETSc is not part of the reference tree: this restates its published
   algorithm; the reference calls it from STApply_Sinvert (sinvert.c:16-30). */
static int pcg_solve(oeps *eps, const double *b, double *x)
{
  int n = eps->n;
  double *r = eps->kwork, *z = r + n, *p = z + n, *q = p + n, rz = 0.0, rzold = 0.0, zz = 0.0;
  for (int i = 0; i < n; i++) { x[i] = 0.0; r[i] = b[i]; z[i] = eps->dinv[i] * r[i]; rz += z[i] * r[i]; zz += z[i] * z[i]; }
  double rnorm0 = sqrt(zz), ttol = fmax(eps->ksp_rtol * rnorm0, 1e-50);
  if (rnorm0 <= ttol) return 0;
  for (int it = 0; it < eps->ksp_maxits; it++) {
    double beta = it ? rz / rzold : 0.0, pq = 0.0, a;
    for (int i = 0; i < n; i++) p[i] = it ? z[i] + beta * p[i] : z[i];
    or_spmv(n, eps->rowptr, eps->colidx, eps->val, p, q);
    for (int i = 0; i < n; i++) { q[i] -= eps->sigma * p[i]; pq += p[i] * q[i]; }
    a = rz / pq; rzold = rz; rz = 0.0; zz = 0.0;
    for (int i = 0; i < n; i++) {
      x[i] += a * p[i]; r[i] -= a * q[i]; z[i] = eps->dinv[i] * r[i]; rz += z[i] * r[i]; zz += z[i] * z[i];
    }
    eps->ksp_its++;
    if (!(pq > 0.0)) OR_ERR(-91, "KSP did not converge: p^T (A - sigma I) p <= 0");
    if (sqrt(zz) <= ttol) return 0;
  }
  OR_ERR(-91, "KSP did not converge in %d iterations", eps->ksp_maxits);
}

/* ILU(0) of A - sigma I (Y. Saad, Iterative Methods for Sparse Linear Systems,
   2nd ed., Algorithm 10.4, the IKJ form): the factors L (unit lower) and U keep
   the pattern of A, row by row in the natural order, as PETSc's PCILU with
   levels 0 and the natural ordering (its defaults) on a sequential AIJ matrix.
   Each row's columns are in increasing order and include the diagonal.  A zero
   pivot fails, as PETSc's zero-pivot check does. */
static int ilu0_factor(oeps *eps)
{
  const int n = eps->n, *rp = eps->rowptr, *ci = eps->colidx;
  double *a = eps->lu;
  for (int i = 0; i < n; i++) {
    eps->lu_diag[i] = -1;
    for (int q = rp[i]; q < rp[i + 1]; q++) {
      a[q] = eps->val[q] - (ci[q] == i ? eps->sigma : 0.0);
      if (ci[q] == i) eps->lu_diag[i] = q;
      if (q > rp[i] && ci[q] <= ci[q - 1]) OR_ERR(-92, "ILU(0) needs increasing column indices in each row");
    }
    if (eps->lu_diag[i] < 0) OR_ERR(-92, "ILU(0) needs the diagonal entry of row %d", i);
  }
  for (int i = 1; i < n; i++)
    for (int q = rp[i]; q < eps->lu_diag[i]; q++) {     /* k = ci[q] < i, in increasing order */
      const int k = ci[q];
      const double ukk = a[eps->lu_diag[k]];
      if (ukk == 0.0) OR_ERR(-92, "ILU(0): zero pivot in row %d", k);
      const double lik = a[q] / ukk;
      a[q] = lik;
      /* a_ij -= l_ik u_kj for the j > k of row i that row k also holds */
      int t = eps->lu_diag[k] + 1;
      for (int r = q + 1; r < rp[i + 1]; r++) {
        const int j = ci[r];
        while (t < rp[k + 1] && ci[t] < j) t++;
        if (t < rp[k + 1] && ci[t] == j) a[r] -= lik * a[t];
      }
    }
  if (a[eps->lu_diag[n - 1]] == 0.0) OR_ERR(-92, "ILU(0): zero pivot in row %d", n - 1);
  return 0;
}

/* The factors alone, for the tests: lu (nnz, A's pattern) gets L below the
   diagonal and U on and above it */
int or_ilu0(int n, const int *rowptr, const int *colidx, const double *val, double sigma, double *lu)
{
  oeps eps;
  memset(&eps, 0, sizeof(eps));
  eps.n = n; eps.rowptr = rowptr; eps.colidx = colidx; eps.val = val; eps.sigma = sigma;
  eps.lu = lu;
  eps.lu_diag = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  if (!eps.lu_diag) return -55;
  int e = ilu0_factor(&eps);
  free(eps.lu_diag);
  return e;
}

/* The direct inner solve KSPPREONLY + PCLU, the reference's default for
   shift-and-invert (STSetDefaultKSP_Default, stsles.c:47-54): one application of
   the LU factors of A - sigma I.  PETSc (not part of the reference tree) factors
   the sparse matrix with a fill-reducing ordering; this restates the dense LU with
   partial pivoting of LAPACK's dgetf2 (right-looking, the pivot of column k is the
   first entry of largest magnitude on or below the diagonal, as idamax picks it;
   the column below the pivot is scaled by the pivot's reciprocal) and the solve of
   dgetrs (the row interchanges in order, then unit lower and upper triangular
   substitution).  In exact arithmetic both give (A - sigma I)^-1 b.  A pivot of
   magnitude below 100 machine epsilons fails as a zero pivot, the test and default
   tolerance of PETSc's factorizations (MatFactorInfo zeropivot). */
static int lu_factor(oeps *eps)
{
  const int n = eps->n;
  double *a = eps->dlu;
  memset(a, 0, sizeof(double) * (size_t)n * n);
  for (int i = 0; i < n; i++)
    for (int q = eps->rowptr[i]; q < eps->rowptr[i + 1]; q++) a[i + (size_t)eps->colidx[q] * n] += eps->val[q];
  for (int i = 0; i < n; i++) a[i + (size_t)i * n] -= eps->sigma;
  for (int k = 0; k < n; k++) {
    double *ck = a + (size_t)k * n;
    int p = k;
    double m = fabs(ck[k]);
    for (int i = k + 1; i < n; i++)
      if (fabs(ck[i]) > m) { m = fabs(ck[i]); p = i; }
    eps->ipiv[k] = p;
    if (!(m >= 100.0 * MACH_EPS)) OR_ERR(-92, "LU: zero pivot in column %d (|pivot| %g)", k, m);
    if (p != k)
      for (int j = 0; j < n; j++) {
        double t = a[k + (size_t)j * n];
        a[k + (size_t)j * n] = a[p + (size_t)j * n];
        a[p + (size_t)j * n] = t;
      }
    const double r = 1.0 / ck[k];
    for (int i = k + 1; i < n; i++) ck[i] *= r;
#pragma omp parallel for schedule(static) if (n - k > 256)
    for (int j = k + 1; j < n; j++) {               /* rank-1 update of the trailing matrix */
      double *cj = a + (size_t)j * n;
      const double akj = cj[k];
      if (akj != 0.0)
        for (int i = k + 1; i < n; i++) cj[i] -= ck[i] * akj;
    }
  }
  return 0;
}

static void lu_solve(const oeps *eps, const double *b, double *x)
{
  const int n = eps->n;
  const double *a = eps->dlu;
  memcpy(x, b, sizeof(double) * (size_t)n);
  for (int k = 0; k < n; k++)
    if (eps->ipiv[k] != k) { double t = x[k]; x[k] = x[eps->ipiv[k]]; x[eps->ipiv[k]] = t; }
  for (int j = 0; j < n; j++) {                     /* L y = P b, by columns */
    const double xj = x[j];
    if (xj != 0.0)
      for (int i = j + 1; i < n; i++) x[i] -= a[i + (size_t)j * n] * xj;
  }
  for (int j = n - 1; j >= 0; j--) {                /* U x = y, by columns */
    x[j] /= a[j + (size_t)j * n];
    const double xj = x[j];
    if (xj != 0.0)
      for (int i = 0; i < j; i++) x[i] -= a[i + (size_t)j * n] * xj;
  }
}

/* The dense factors alone, for the tests: lu (n x n, by columns) gets L below the
   diagonal and U on and above it, ipiv the interchanges (0-based) */
int or_lu(int n, const int *rowptr, const int *colidx, const double *val, double sigma, double *lu, int *ipiv)
{
  oeps eps;
  memset(&eps, 0, sizeof(eps));
  eps.n = n; eps.rowptr = rowptr; eps.colidx = colidx; eps.val = val; eps.sigma = sigma;
  eps.dlu = lu; eps.ipiv = ipiv;
  return lu_factor(&eps);
}

/* y = (L U)^-1 y in place: forward substitution with the unit lower factor, then
   backward substitution with U */
static void ilu0_apply(const oeps *eps, double *y)
{
  const int n = eps->n, *rp = eps->rowptr, *ci = eps->colidx;
  const double *a = eps->lu;
  for (int i = 0; i < n; i++) {
    double s = y[i];
    for (int q = rp[i]; q < eps->lu_diag[i]; q++) s -= a[q] * y[ci[q]];
    y[i] = s;
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = y[i];
    for (int q = eps->lu_diag[i] + 1; q < rp[i + 1]; q++) s -= a[q] * y[ci[q]];
    y[i] = s / a[eps->lu_diag[i]];
  }
}

/* y = B (A - sigma I) x, the left-preconditioned operator of the BiCGStab solve */
static void bcgs_op(oeps *eps, const double *x, double *y)
{
  or_spmv(eps->n, eps->rowptr, eps->colidx, eps->val, x, y);
  if (eps->pc_ilu) {
    for (int i = 0; i < eps->n; i++) y[i] -= eps->sigma * x[i];
    ilu0_apply(eps, y);
  } else
    for (int i = 0; i < eps->n; i++) y[i] = eps->dinv[i] * (y[i] - eps->sigma * x[i]);
}

/* BiCGStab (H. A. van der Vorst, SIAM J. Sci. Stat. Comput. 13 (1992) 631-644) on
   B (A - sigma I) x = B b, left preconditioned with B = diag(A - sigma I)^-1
   (PCJACOBI) or I, as PETSc's KSPBCGS with its default left preconditioning:
   x0 = 0, r = B b, shadow rhat = r; per iteration beta = (rho/rho_old)(alpha/omega),
   p = r + beta (p - omega v), v = K p, alpha = rho / (rhat, v), s = r - alpha v; if
   |s| <= ttol then x += alpha p and stop; t = K s, omega = (t, s)/(t, t),
   x += alpha p + omega s, r = s - omega t; converged when |r| <= max(rtol |B b|,
   abstol).  A zero rho, (rhat, v), (t, t) or omega is a breakdown.  PETSc is not
   part of the reference tree: this restates the published method; the
   reference's default inner solver for shift-and-invert is a direct solve
   (stsles.c:47-54), KSPBCGS being the iterative choice for an indefinite shifted
   matrix. */
static int bcgs_solve(oeps *eps, const double *b, double *x)
{
  int n = eps->n;
  double *r = eps->kwork, *rh = r + n, *p = rh + n, *v = p + n, *sv = v + n, *t = sv + n;
  double rho, rho_old = 1.0, alpha = 1.0, omega = 1.0, rr = 0.0;
  for (int i = 0; i < n; i++) r[i] = eps->pc_ilu ? b[i] : eps->dinv[i] * b[i];
  if (eps->pc_ilu) ilu0_apply(eps, r);
  for (int i = 0; i < n; i++) {
    x[i] = 0.0; rh[i] = r[i]; p[i] = 0.0; v[i] = 0.0; rr += r[i] * r[i];
  }
  double ttol = fmax(eps->ksp_rtol * sqrt(rr), 1e-50);
  if (sqrt(rr) <= ttol) return 0;
  rho = rr;
  for (int it = 0; it < eps->ksp_maxits; it++) {
    double beta = (rho / rho_old) * (alpha / omega), rv = 0.0, ss = 0.0, ts = 0.0, tt = 0.0, rhr = 0.0;
    for (int i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    bcgs_op(eps, p, v);
    for (int i = 0; i < n; i++) rv += rh[i] * v[i];
    if (rv == 0.0) OR_ERR(-91, "KSP breakdown: (rhat, v) = 0");
    alpha = rho / rv;
    for (int i = 0; i < n; i++) { sv[i] = r[i] - alpha * v[i]; ss += sv[i] * sv[i]; }
    eps->ksp_its++;
    if (sqrt(ss) <= ttol) {
      for (int i = 0; i < n; i++) x[i] += alpha * p[i];
      return 0;
    }
    bcgs_op(eps, sv, t);
    for (int i = 0; i < n; i++) { ts += t[i] * sv[i]; tt += t[i] * t[i]; }
    if (tt == 0.0) OR_ERR(-91, "KSP breakdown: (t, t) = 0");
    omega = ts / tt;
    rr = 0.0;
    for (int i = 0; i < n; i++) {
      x[i] += alpha * p[i] + omega * sv[i];
      r[i] = sv[i] - omega * t[i];
      rr += r[i] * r[i];
      rhr += rh[i] * r[i];
    }
    if (sqrt(rr) <= ttol) return 0;
    rho_old = rho;
    rho = rhr;
    if (rho == 0.0 || omega == 0.0) OR_ERR(-91, "KSP breakdown: rho or omega = 0");
  }
  OR_ERR(-91, "KSP did not converge in %d iterations", eps->ksp_maxits);
}

static void op_mult(oeps *eps, const double *x, double *y)
{
  if (eps->sinvert && eps->ksp_lu) {
    lu_solve(eps, x, y);
    eps->ksp_its++;
  } else if (eps->sinvert) {
    if ((eps->ksp_bcgs ? bcgs_solve(eps, x, y) : pcg_solve(eps, x, y))) eps->ksp_failed = 1;
  } else or_spmv(eps->n, eps->rowptr, eps->colidx, eps->val, x, y);
  eps->nmatmult++;
}

static void op_mult_transpose(oeps *eps, const double *x, double *y)
{
  or_spmv(eps->n, eps->trowptr, eps->tcolidx, eps->tval, x, y);
  eps->nmatmult++;
}

/* BVMatArnoldi / BVMatLanczos (bvkrylov.c:56-226) on basis bv with A (or A^T), the
   Hessenberg columns into H (or the tridiagonal of the DS) */
static int mat_krylov_on(oeps *eps, obv *bv, int transpose, double *H, int k, int *m, double *beta, int *breakdown)
{
  ods *ds = &eps->ds;
  int lindep = 0, ldb = bv->nc + bv->m, ld = ds->ld;
  const double *a = bv->buffer;
  for (int j = k; j < *m; j++) {
    if (transpose) op_mult_transpose(eps, COL(bv, j), COL(bv, j + 1));
    else op_mult(eps, COL(bv, j), COL(bv, j + 1));               /* BVMatMultColumn */
    if (eps->ksp_failed) OR_ERR(-91, "the inner solve of shift-and-invert did not converge");
    if (j == eps->n - 1) {                                         /* BV_OrthogonalizeColumn_Safe */
      int save = bv->refine;
      bv->refine = OR_REFINE_NEVER;
      OR_CALL(orthogonalize_column(bv, j + 1, NULL, NULL));
      bv->refine = save;
      *beta = 0.0; lindep = 1;
    } else OR_CALL(orthonormalize_column(bv, j + 1, beta, &lindep));
    if (lindep) { *m = j + 1; break; }
  }
  *breakdown = lindep;
  if (eps->dsherm) {
    double *alpha = ds->T, *betat = ds->T + ld;
    for (int j = k; j < *m; j++) {
      alpha[j] = a[bv->nc + j + (size_t)(j + 1) * ldb];
      betat[j] = a[bv->nc + j + 1 + (size_t)(j + 1) * ldb];
    }
  } else {
    double *h = H;
    for (int j = k; j < *m - 1; j++) memcpy(h + (size_t)j * ld, a + bv->nc + (size_t)(j + 1) * ldb, sizeof(double) * (j + 2));
    memcpy(h + (size_t)(*m - 1) * ld, a + bv->nc + (size_t)(*m) * ldb, sizeof(double) * (*m));
    if (ld > *m) h[*m + (size_t)(*m - 1) * ld] = a[bv->nc + *m + (size_t)(*m) * ldb];
  }
  return 0;
}

static int mat_krylov(oeps *eps, int k, int *m, double *beta, int *breakdown)
{
  return mat_krylov_on(eps, &eps->V, 0, eps->ds.A, k, m, beta, breakdown);
}

/* EPSGetStartVector (epssolve.c:841-873) */
static int get_start_vector(oeps *eps, int i, int have_initial, int *breakdown)
{
  double norm;
  int lindep = 0;
  if (i > 0 || !have_initial) bv_randomcol(&eps->V, i);
  OR_CALL(orthogonalize_column(&eps->V, i, &norm, &lindep));
  if (breakdown) *breakdown = lindep;
  else if (lindep || norm == 0.0) {
    if (!i) OR_ERR(-77, "Initial vector is zero or belongs to the deflation space");
    OR_ERR(-80, "Unable to generate more start vectors");
  }
  bv_scalecol(&eps->V, i, 1.0 / norm);
  return 0;
}

/* EPSKrylovConvergence (epskrylov.c:207-295), STSHIFT sigma=0, trivial RG, relative test */
static int krylov_convergence(oeps *eps, int kini, int nits, double beta, double betat, double corrf, int *kout)
{
  int k, marker = -1;
  double resnorm;
  for (k = kini; k < kini + nits; k++) {
    double re = eps->eigr[k], im = eps->eigi[k], w;
    int newk = k;
    OR_CALL(ds_vectors(&eps->ds, &newk, &resnorm));
    resnorm *= beta * corrf;                                       /* corrf: gamma of harmonic KS */
    w = hypot(re, im);                                             /* EPSConvergedRelative */
    eps->errest[k] = (w != 0.0) ? resnorm / w : DBL_MAX;
    if (marker == -1 && eps->errest[k] >= eps->tol) marker = k;
    if (eps->twosided) {                                           /* epskrylov.c:269-276 */
      int newk2 = k;
      double lerrest;
      OR_CALL(ds_vectors_left(&eps->ds, &newk2, &resnorm));
      resnorm *= betat;
      lerrest = (w != 0.0) ? resnorm / w : DBL_MAX;
      if (lerrest > eps->errest[k]) eps->errest[k] = lerrest;
      if (marker == -1 && lerrest >= eps->tol) marker = k;
    }
    if (newk == k + 1) { eps->errest[k + 1] = eps->errest[k]; k++; }
    if (marker != -1) break;
  }
  if (marker != -1) k = marker;
  *kout = k;
  return 0;
}

/* EPSSolve_KrylovSchur_Default (krylovschur.c:227-337), Ritz or harmonic extraction */
static int ks_loop(oeps *eps, int have_initial, int restart_cap, double *g)
{
  int k = 0, l = 0, nv, nconv, breakdown = 0;
  double beta = 0.0, t0, gamma = 1.0;
  obv *V = &eps->V;
  ods *ds = &eps->ds;

  OR_CALL(get_start_vector(eps, 0, have_initial, NULL));
  l = 0;
  while (eps->reason == OR_CONVERGED_ITERATING) {
    eps->its++;
    nv = eps->nconv + eps->mpd < eps->ncv ? eps->nconv + eps->mpd : eps->ncv;
    ds_setdims(ds, nv, eps->nconv, eps->nconv + l);
    OR_CALL(mat_krylov(eps, eps->nconv + l, &nv, &beta, &breakdown));
    ds_setdims(ds, nv, eps->nconv, eps->nconv + l);
    ds->state = l ? DS_RAW : DS_INTERMEDIATE;
    V->l = eps->nconv; V->k = nv;                                  /* BVSetActiveColumns */

    t0 = wtime();
    if (eps->harmonic) OR_CALL(ds_translate_harmonic(ds, eps->target, beta, 0, g, &gamma));
    OR_CALL(ds_solve(ds, eps->eigr, eps->eigi));
    OR_CALL(ds_sort(ds, eps->eigr, eps->eigi));
    ds_update_extra_row(ds);
    OR_CALL(krylov_convergence(eps, eps->nconv, nv - eps->nconv, beta, 0.0, gamma, &k));
    eps->t_ds += wtime() - t0;

    /* EPSStoppingBasic (epsdefault.c:290-302) */
    if (k >= eps->nev) eps->reason = OR_CONVERGED_TOL;
    else if (eps->its >= eps->max_it) eps->reason = OR_DIVERGED_ITS;
    if (restart_cap > 0 && eps->its >= restart_cap && eps->reason == OR_CONVERGED_ITERATING)
      eps->reason = OR_DIVERGED_ITS;
    nconv = k;

    if (eps->reason != OR_CONVERGED_ITERATING || breakdown || k == nv) l = 0;
    else {
      l = (int)((nv - k) * eps->keep); if (l < 1) l = 1;
      if (!eps->dsherm) ds_truncate_size(ds, k, nv, &l);
    }
    if (!eps->lock && l > 0) { l += k; k = 0; }

    if (eps->reason == OR_CONVERGED_ITERATING) {
      if (breakdown || k == nv) {
        if (k < eps->nev) {
          OR_CALL(get_start_vector(eps, k, have_initial, &breakdown));
          if (breakdown) eps->reason = OR_DIVERGED_BREAKDOWN;
        }
      } else {
        if (eps->harmonic) {                                       /* krylovschur.c:310-319 */
          ds_setdims(ds, nv, k, l);
          OR_CALL(ds_translate_harmonic(ds, 0.0, beta, 1, g, &gamma));
          V->l = 0; V->k = nv;                                     /* BVMultColumn(V,-1,1,nv,g) */
          bv_multvec(V, -1.0, 1.0, COL(V, nv), g);
          bv_scalecol(V, nv, 1.0 / gamma);
          V->l = eps->nconv; V->k = nv;
          ds_setdims(ds, nv, k, nv);
        }
        ds_truncate(ds, k + l, 0);
      }
    }
    /* BVMultInPlace(V, Q, nconv, k+l) */
    bv_multinplace(V, ds->Q, ds->ld, eps->nconv, k + l);
    if (eps->reason == OR_CONVERGED_ITERATING && !breakdown) bv_copycol(V, nv, k + l);
    eps->nconv = k;
    (void)nconv;
  }
  ds_truncate(ds, eps->nconv, 1);
  return 0;
}

/* ------------------------------------------------------------------ two-sided variant */
/* EPSGetLeftStartVector (epssolve.c:879-900): always a random column (no left initial space) */
static int get_left_start_vector(oeps *eps, int i, int *breakdown)
{
  double norm;
  int lindep = 0;
  bv_randomcol(&eps->W, i);
  OR_CALL(orthogonalize_column(&eps->W, i, &norm, &lindep));
  if (breakdown) *breakdown = lindep;
  else if (lindep || norm == 0.0) OR_ERR(-80, "Unable to generate more left start vectors");
  bv_scalecol(&eps->W, i, 1.0 / norm);
  return 0;
}

/* BVMatProject(V, NULL, W, M) (bvglobal.c:1093-1160): M(i,j) = W(:,i)^T V(:,j) for
   i, j < nv outside the leading (l x l) block, l = the active offset of both */
static void mat_project(oeps *eps, int l, int nv)
{
  int n = eps->n, ncv = eps->ncv;
  for (int j = 0; j < nv; j++)
    for (int i = 0; i < nv; i++) {
      if (i < l && j < l) continue;
      double t = 0.0;
      const double *w = COL(&eps->W, i), *v = COL(&eps->V, j);
      for (int r = 0; r < n; r++) t += w[r] * v[r];
      eps->M[i + (size_t)j * ncv] = t;
    }
}

/* EPSTwoSidedRQUpdate1 (ks-twosided.c:25-77) */
static int rq_update1(oeps *eps, int nv, double beta, double betat)
{
  int ncv = eps->ncv, n_ = nv, one = 1, info = 0, ld = eps->ds.ld;
  int *p = (int*)malloc(sizeof(int) * (size_t)(nv > 0 ? nv : 1));
  double *A = (double*)malloc(sizeof(double) * (size_t)ncv * ncv), *w = (double*)calloc((size_t)ncv, sizeof(double));
  obv *V = &eps->V, *W = &eps->W;
  int vl = V->l, vk = V->k;
  if (!p || !A || !w) { free(p); free(A); free(w); return -55; }
  V->l = 0; V->k = nv; W->l = 0; W->k = nv;
  bv_dotvec(W, COL(V, nv), w);                                     /* w = W^T u */
  memcpy(A, eps->M, sizeof(double) * (size_t)ncv * ncv);
  dgetrf_(&n_, &n_, A, &ncv, p, &info);
  if (!info) dgetrs_("N", &n_, &one, A, &ncv, p, w, &ncv, &info);
  if (!info) {
    bv_multvec(V, -1.0, 1.0, COL(V, nv), w);                       /* u -= V M^-1 W^T u */
    for (int i = 0; i < nv; i++) eps->ds.A[(size_t)(nv - 1) * ld + i] += beta * w[i];
    bv_dotvec(V, COL(W, nv), w);                                   /* w = V^T v */
    dgetrs_("C", &n_, &one, A, &ncv, p, w, &ncv, &info);
  }
  if (!info) {
    bv_multvec(W, -1.0, 1.0, COL(W, nv), w);
    for (int i = 0; i < nv; i++) eps->ds.B[(size_t)(nv - 1) * ld + i] += betat * w[i];
  }
  V->l = vl; V->k = vk; W->l = vl; W->k = vk;
  free(p); free(A); free(w);
  if (info) OR_ERR(-76, "getrf/getrs info %d", info);
  return 0;
}

/* EPSTwoSidedRQUpdate2 (ks-twosided.c:79-136) */
static int rq_update2(oeps *eps, int k)
{
  int ncv = eps->ncv, ld = eps->ds.ld, n_, ld_ = ld;
  obv *V = &eps->V, *W = &eps->W;
  int l = V->l, nv = V->k;
  double norm, zero = 0.0, sone = 1.0;
  double *w = (double*)calloc((size_t)ncv * ncv, sizeof(double));
  if (!w) return -55;
  for (int side = 0; side < 2; side++) {
    obv *X = side ? W : V;
    int ldb = X->nc + X->m;
    double *A = side ? eps->ds.B : eps->ds.A, *c;
    X->l = 0; X->k = nv;
    OR_CALL(orthogonalize_column(X, k, &norm, NULL));              /* coefficients in buffer column k */
    c = X->buffer + (size_t)k * ldb + X->nc;
    bv_scalecol(X, k, 1.0 / norm);
    for (int j = l; j < k; j++) {
      for (int i = 0; i < k; i++) A[i + (size_t)j * ld] += c[i] * A[k + (size_t)j * ld];
      A[k + (size_t)j * ld] *= norm;
    }
    X->l = l; X->k = nv;
  }
  n_ = nv;
  dgemm_("N", "N", &n_, &n_, &n_, &sone, eps->M, &ncv, eps->ds.Q, &ld_, &zero, w, &ncv);
  dgemm_("C", "N", &n_, &n_, &n_, &sone, eps->ds.Z, &ld_, w, &ncv, &zero, eps->M, &ncv);
  free(w);
  return 0;
}

/* EPSSolve_KrylovSchur_TwoSided (ks-twosided.c:138-241) */
static int ks_loop_twosided(oeps *eps, int have_initial, int restart_cap)
{
  int k = 0, l = 0, nv, nvt, breakdown = 0, breakdownt = 0, breakdownl = 0;
  double beta = 0.0, betat = 0.0, norm, norm2, t0;
  int one = 1;
  obv *V = &eps->V, *W = &eps->W;
  ods *ds = &eps->ds;
  OR_CALL(get_start_vector(eps, 0, have_initial, NULL));
  OR_CALL(get_left_start_vector(eps, 0, NULL));
  while (eps->reason == OR_CONVERGED_ITERATING) {
    eps->its++;
    nv = eps->nconv + eps->mpd < eps->ncv ? eps->nconv + eps->mpd : eps->ncv;
    ds_setdims(ds, nv, eps->nconv, eps->nconv + l);
    OR_CALL(mat_krylov_on(eps, V, 0, ds->A, eps->nconv + l, &nv, &beta, &breakdown));
    nvt = nv;
    ds_setdims(ds, nv, eps->nconv, eps->nconv + l);
    OR_CALL(mat_krylov_on(eps, W, 1, ds->B, eps->nconv + l, &nvt, &betat, &breakdownt));
    if (nvt < nv) nv = nvt;
    ds_setdims(ds, nv, eps->nconv, eps->nconv + l);
    ds->state = l == 0 ? DS_INTERMEDIATE : DS_RAW;
    breakdown = breakdown || breakdownt;

    V->l = eps->nconv + l; V->k = nv; W->l = eps->nconv + l; W->k = nv;
    mat_project(eps, eps->nconv + l, nv);
    OR_CALL(rq_update1(eps, nv, beta, betat));

    t0 = wtime();
    OR_CALL(ds_solve(ds, eps->eigr, eps->eigi));
    OR_CALL(ds_sort(ds, eps->eigr, eps->eigi));
    ds_update_extra_row(ds);
    norm = dnrm2_(&eps->n, COL(V, nv), &one);
    norm2 = dnrm2_(&eps->n, COL(W, nv), &one);
    OR_CALL(krylov_convergence(eps, eps->nconv, nv - eps->nconv, beta * norm, betat * norm2, 1.0, &k));
    eps->t_ds += wtime() - t0;
    if (k >= eps->nev) eps->reason = OR_CONVERGED_TOL;
    else if (eps->its >= eps->max_it) eps->reason = OR_DIVERGED_ITS;
    if (restart_cap > 0 && eps->its >= restart_cap && eps->reason == OR_CONVERGED_ITERATING)
      eps->reason = OR_DIVERGED_ITS;

    if (eps->reason != OR_CONVERGED_ITERATING || breakdown || k == nv) l = 0;
    else {
      l = (int)((nv - k) * eps->keep); if (l < 1) l = 1;
      ds_truncate_size(ds, k, nv, &l);
    }
    if (!eps->lock && l > 0) { l += k; k = 0; }

    V->l = eps->nconv; V->k = nv; W->l = eps->nconv; W->k = nv;
    bv_multinplace(V, ds->Q, ds->ld, eps->nconv, k + l);
    bv_multinplace(W, ds->Z, ds->ld, eps->nconv, k + l);
    if (eps->reason == OR_CONVERGED_ITERATING && !breakdown) { bv_copycol(V, nv, k + l); bv_copycol(W, nv, k + l); }
    if (eps->reason == OR_CONVERGED_ITERATING) {
      if (breakdown || k == nv) {
        if (k < eps->nev) {
          OR_CALL(get_start_vector(eps, k, have_initial, &breakdown));
          OR_CALL(get_left_start_vector(eps, k, &breakdownl));
          if (breakdown || breakdownl) eps->reason = OR_DIVERGED_BREAKDOWN;
        }
      } else {
        int dsn = ds->n, dsk = ds->k;
        ds_setdims(ds, dsn, k, dsk);
        ds_truncate(ds, k + l, 0);
      }
      OR_CALL(rq_update2(eps, k + l));
    }
    eps->nconv = k;
  }
  ds_truncate(ds, eps->nconv, 1);
  return 0;
}

static int ks_solve(oeps *eps, int have_initial, int restart_cap)
{
  double *g = NULL;                                                /* translation vector, harmonic only */
  int e;
  if (eps->twosided) return ks_loop_twosided(eps, have_initial, restart_cap);
  if (eps->harmonic && !(g = (double*)calloc((size_t)eps->ds.ld, sizeof(double)))) return -55;
  e = ks_loop(eps, have_initial, restart_cap, g);
  free(g);
  return e;
}

/* EPSComputeResidualNorm_Private (epssolve.c:666-718), standard problem */
/* y = A x for residuals (the matrix itself, also under shift-and-invert) */
static void a_mult(oeps *eps, const double *x, double *y)
{
  or_spmv(eps->n, eps->rowptr, eps->colidx, eps->val, x, y);
  eps->nmatmult++;
}

static double residual_norm(oeps *eps, double kr, double ki, const double *xr, const double *xi,
                            double *u, double *w)
{
  int n = eps->n, one = 1;
  if (ki == 0 || fabs(ki) < fabs(kr * MACH_EPS)) {
    a_mult(eps, xr, u);
    if (fabs(kr) > MACH_EPS) for (int i = 0; i < n; i++) u[i] -= kr * xr[i];
    return dnrm2_(&n, u, &one);
  } else {
    double nr, ni;
    a_mult(eps, xr, u);
    if (hypot(kr, ki) > MACH_EPS) for (int i = 0; i < n; i++) u[i] += -kr * xr[i] + ki * xi[i];
    nr = dnrm2_(&n, u, &one);
    a_mult(eps, xi, u);
    if (hypot(kr, ki) > MACH_EPS) for (int i = 0; i < n; i++) u[i] += -kr * xi[i] - ki * xr[i];
    ni = dnrm2_(&n, u, &one);
    (void)w;
    return hypot(nr, ni);
  }
}

/* A^T by a stable counting sort (rows of A^T in increasing column order) */
static int transpose_csr(int n, const int *rp, const int *ci, const double *v, int **trp, int **tci, double **tv)
{
  long nnz = rp[n];
  int *fill = (int*)calloc((size_t)n + 1, sizeof(int));
  *trp = (int*)calloc((size_t)n + 1, sizeof(int));
  *tci = (int*)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
  *tv = (double*)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
  if (!fill || !*trp || !*tci || !*tv) { free(fill); return -55; }
  for (long p = 0; p < nnz; p++) (*trp)[ci[p] + 1]++;
  for (int j = 0; j < n; j++) (*trp)[j + 1] += (*trp)[j];
  for (int i = 0; i < n; i++)
    for (int p = rp[i]; p < rp[i + 1]; p++) {
      int q = (*trp)[ci[p]] + fill[ci[p]]++;
      (*tci)[q] = i;
      (*tv)[q] = v[p];
    }
  free(fill);
  return 0;
}

int or_eps_solve(int n, const int *rowptr, const int *colidx, const double *val,
                 const or_opts *opts, const double *v0, int nds, const double *defl,
                 double *eigr, double *eigi, double *errest, double *err, double *X,
                 or_info *info)
{
  return or_eps_solve_twosided(n, rowptr, colidx, val, opts, v0, nds, defl, eigr, eigi, errest, err, X, NULL, info);
}

int or_eps_solve_twosided(int n, const int *rowptr, const int *colidx, const double *val,
                          const or_opts *opts, const double *v0, int nds, const double *defl,
                          double *eigr, double *eigi, double *errest, double *err, double *X, double *Y,
                          or_info *info)
{
  or_mkl_layer();
  oeps eps;
  int e = 0, ncv, mpd, requested;
  double t0;
  memset(&eps, 0, sizeof(eps));
  if (opts->nthreads > 0) omp_set_num_threads(opts->nthreads);

  eps.n = n; eps.rowptr = rowptr; eps.colidx = colidx; eps.val = val;
  eps.sinvert = opts->st_sinvert != 0;
  eps.nev = opts->nev;
  eps.which = opts->which ? opts->which : (eps.sinvert ? OR_TARGET_MAGNITUDE : OR_LARGEST_MAGNITUDE);
  eps.target = opts->target; eps.hermitian = opts->hermitian;
  eps.harmonic = opts->extraction == 1;
  if (opts->extraction != 0 && opts->extraction != 1) OR_ERR(-56, "Unsupported extraction type");
  eps.twosided = opts->twosided != 0;
  if (eps.twosided && eps.hermitian) OR_ERR(-56, "Two-sided methods are not intended for symmetric problems");
  if (eps.twosided) eps.harmonic = 0;                              /* the variant takes precedence */
  eps.dsherm = eps.hermitian && !eps.harmonic;
  eps.tol = opts->tol > 0 ? opts->tol : 1e-8;
  eps.keep = opts->keep > 0 ? opts->keep : 0.5;
  eps.lock = opts->lock;
  eps.sc.which = eps.which; eps.sc.target = eps.target;
  if (eps.sinvert) {                                               /* STSINVERT, shift = target */
    if (eps.which != OR_TARGET_MAGNITUDE && eps.which != OR_TARGET_REAL)
      OR_ERR(-90, "Shift-and-invert requires a target 'which'");
    if (opts->extraction != 0 || opts->twosided) OR_ERR(-56, "shift-and-invert runs with Ritz extraction, one-sided");
    eps.sigma = opts->target;
    eps.ksp_rtol = opts->ksp_rtol > 0 ? opts->ksp_rtol : 1e-8;
    eps.ksp_maxits = 10000;
    eps.pc_jacobi = !opts->pc_none;
    eps.ksp_bcgs = opts->ksp_bcgs != 0;
    eps.pc_ilu = opts->pc_ilu != 0;
    eps.ksp_lu = opts->ksp_lu != 0;
    if (eps.pc_ilu && !eps.ksp_bcgs) { e = -3; goto done; }   /* ILU(0) with BiCGStab only */
    if (eps.ksp_lu && (eps.pc_ilu || eps.ksp_bcgs)) { e = -3; goto done; }
  }

  /* EPSSetUp (epssetup.c:344-345) and EPSSetDimensions_Default (epssetup.c:648-671) */
  if (eps.nev > n) eps.nev = n;
  ncv = opts->ncv > 0 ? opts->ncv : -1;
  mpd = opts->mpd > 0 ? opts->mpd : -1;
  if (ncv > n) ncv = n;
  if (ncv > 0) {
    if (!(ncv >= eps.nev + 1 || (ncv == eps.nev && ncv == n))) OR_ERR(-62, "The value of ncv must be at least nev+1");
  } else if (mpd > 0) ncv = n < eps.nev + mpd ? n : eps.nev + mpd;
  else ncv = n < (2 * eps.nev > eps.nev + 15 ? 2 * eps.nev : eps.nev + 15) ? n : (2 * eps.nev > eps.nev + 15 ? 2 * eps.nev : eps.nev + 15);
  if (mpd <= 0) mpd = ncv;
  if (ncv > eps.nev + mpd) OR_ERR(-62, "The value of ncv must not be larger than nev+mpd");
  eps.ncv = ncv; eps.mpd = mpd;
  eps.max_it = opts->max_it > 0 ? opts->max_it : (100 > 2 * n / ncv ? 100 : 2 * n / ncv);

  /* EPSAllocateSolution(eps, 1) (epssetup.c:692-732) */
  requested = ncv + 1;
  eps.V.n = n; eps.V.m = requested; eps.V.nc = nds > 0 ? nds : 0; eps.V.ld = n;
  eps.V.refine = opts->refine; eps.V.eta = opts->eta > 0 ? opts->eta : 0.7071;
  eps.V.seed = 0;
  eps.V.V = (double*)calloc((size_t)n * (eps.V.nc + requested), sizeof(double));
  eps.V.buffer = (double*)calloc((size_t)(eps.V.nc + requested) * requested, sizeof(double));
  eps.eigr = (double*)calloc(requested, sizeof(double));
  eps.eigi = (double*)calloc(requested, sizeof(double));
  eps.errest = (double*)calloc(requested, sizeof(double));
  eps.perm = (int*)calloc(requested, sizeof(int));
  if (!eps.V.V || !eps.V.buffer || !eps.eigr || !eps.eigi || !eps.errest || !eps.perm) { e = -55; goto done; }
  if ((e = ds_alloc(&eps.ds, eps.twosided ? DS_NHEPTS : (eps.dsherm ? DS_HEP : DS_NHEP), ncv + 1))) goto done;
  if (eps.twosided) {
    eps.W = eps.V;                                                 /* same sizes and settings, own storage */
    eps.W.nc = 0; eps.W.seed = eps.V.seed + 1000000007ULL; eps.W.rcount = 0; eps.W.passes = 0;
    eps.W.V = (double*)calloc((size_t)n * requested, sizeof(double));
    eps.W.buffer = (double*)calloc((size_t)requested * requested, sizeof(double));
    eps.M = (double*)calloc((size_t)ncv * ncv, sizeof(double));
    if (!eps.W.V || !eps.W.buffer || !eps.M) { e = -55; goto done; }
    if ((e = transpose_csr(n, rowptr, colidx, val, &eps.trowptr, &eps.tcolidx, &eps.tval))) goto done;
  }
  eps.ds.compact = eps.dsherm; eps.ds.extrarow = 1; eps.ds.sc = eps.sc;
  if (eps.sinvert) {
    eps.ds.sc.map = 1; eps.ds.sc.sigma = eps.sigma;
    eps.dinv = (double*)malloc(sizeof(double) * (size_t)n);
    eps.kwork = (double*)malloc(sizeof(double) * 6 * (size_t)n);
    if (!eps.dinv || !eps.kwork) { e = -55; goto done; }
    if (eps.pc_ilu) {
      eps.lu = (double*)malloc(sizeof(double) * (size_t)rowptr[n]);
      eps.lu_diag = (int*)malloc(sizeof(int) * (size_t)n);
      if (!eps.lu || !eps.lu_diag) { e = -55; goto done; }
      if ((e = ilu0_factor(&eps))) goto done;
    }
    if (eps.ksp_lu) {
      eps.dlu = (double*)malloc(sizeof(double) * (size_t)n * n);
      eps.ipiv = (int*)malloc(sizeof(int) * (size_t)n);
      if (!eps.dlu || !eps.ipiv) { e = -55; goto done; }
      if ((e = lu_factor(&eps))) goto done;
    }
    for (int i = 0; i < n; i++) {                                  /* PCJACOBI on A - sigma I */
      double d = 0.0;
      if (eps.pc_jacobi) { for (int q = rowptr[i]; q < rowptr[i + 1]; q++) if (colidx[q] == i) d += val[q]; d -= eps.sigma; }
      eps.dinv[i] = d == 0.0 ? 1.0 : 1.0 / d;
    }
  }

  /* deflation space: BVInsertConstraints (orthonormalized, bvfunc.c:411) */
  if (eps.V.nc) {
    obv *bv = &eps.V;
    int nc = bv->nc;
    memcpy(bv->V, defl, sizeof(double) * (size_t)n * nc);
    bv->nc = 0; bv->m = nc;               /* orthonormalize constraints among themselves */
    for (int j = 0; j < nc; j++) {
      double nrm; int lnd;
      bv->l = 0; bv->k = j;
      if ((e = orthonormalize_column(bv, j, &nrm, &lnd))) goto done;
    }
    bv->nc = nc; bv->m = requested;
    memset(bv->buffer, 0, sizeof(double) * (size_t)(nc + requested) * requested);
  }
  /* initial space (one vector): BVInsertVecs(V,0,&k,IS,TRUE) (bvfunc.c:358-371) */
  if (v0) {
    double nrm; int lnd = 0;
    memcpy(COL(&eps.V, 0), v0, sizeof(double) * (size_t)n);
    eps.V.l = 0; eps.V.k = 0;
    if ((e = orthogonalize_column(&eps.V, 0, &nrm, &lnd))) goto done;
    if (nrm == 0.0 || lnd) { e = -77; fprintf(stderr, "ks_oracle: initial vector is zero or in the deflation space\n"); goto done; }
    bv_scalecol(&eps.V, 0, 1.0 / nrm);
  }

  for (int i = 0; i < ncv; i++) { eps.eigr[i] = eps.eigi[i] = eps.errest[i] = 0.0; eps.perm[i] = i; }
  eps.reason = OR_CONVERGED_ITERATING;
  eps.V.l = 0; eps.V.k = 0;

  t0 = wtime();
  if ((e = ks_solve(&eps, v0 != NULL, opts->restart_cap))) goto done;

  /* EPSComputeVectors_Schur (epsdefault.c:105-169) for NHEP; Hermitian with DSHEP: no-op;
     Hermitian with harmonic extraction: EPSComputeVectors_Hermitian (epsdefault.c:53-72) */
  if (!eps.dsherm && eps.nconv > 0) {
    if ((e = ds_vectors(&eps.ds, NULL, NULL))) goto done;
    eps.V.l = 0; eps.V.k = eps.nconv;
    bv_multinplace(&eps.V, eps.ds.X, eps.ds.ld, 0, eps.hermitian ? eps.ds.n : eps.nconv);
    if (eps.twosided) {                                            /* W = W * Y (epsdefault.c:141-148) */
      if ((e = ds_vectors_left(&eps.ds, NULL, NULL))) goto done;
      eps.W.l = 0; eps.W.k = eps.nconv;
      bv_multinplace(&eps.W, eps.ds.Y, eps.ds.ld, 0, eps.nconv);
    }
    if (eps.hermitian)                                             /* BVNormalize(V, eigi) */
      for (int j = 0; j < eps.nconv; j++) {
        int one = 1, pair = eps.eigi[j] != 0.0 && j + 1 < eps.nconv;
        double nr = dnrm2_(&n, COL(&eps.V, j), &one);
        if (pair) nr = hypot(nr, dnrm2_(&n, COL(&eps.V, j + 1), &one));
        bv_scalecol(&eps.V, j, 1.0 / nr);
        if (pair) bv_scalecol(&eps.V, ++j, 1.0 / nr);
      }
  }
  /* EPSComputeValues (epssolve.c:27-59): STBackTransform_Sinvert, lambda = sigma + 1/theta */
  if (eps.sinvert)
    for (int j = 0; j < eps.nconv; j++) {
      double re = eps.eigr[j], im = eps.eigi[j], nrm = re * re + im * im;
      if (!im) eps.eigr[j] = 1.0 / re + eps.sigma;
      else { eps.eigr[j] = re / nrm + eps.sigma; eps.eigi[j] = -im / nrm; }
    }
  /* reorder conjugate pairs (epssolve.c:161-175) */
  for (int i = 0; i < eps.nconv - 1; i++) {
    if (eps.eigi[i] != 0) {
      if (eps.eigi[i] < 0) {
        eps.eigi[i] = -eps.eigi[i]; eps.eigi[i + 1] = -eps.eigi[i + 1];
        bv_scalecol(&eps.V, i + 1, -1.0);
      }
      i++;
    }
  }
  sort_eigenvalues(&eps.sc, eps.nconv, eps.eigr, eps.eigi, eps.perm);
  if (info) info->t_solve = wtime() - t0;

  /* outputs in EPSGetEigenpair order: eigenvalue perm[i], vector BV_GetEigenvector (bvimpl.h:421-444) */
  {
    double *u = (double*)malloc(sizeof(double) * (size_t)n), *zero = NULL;
    for (int i = 0; i < eps.nconv; i++) {
      int k = eps.perm[i];
      double kr = eps.eigr[k], ki = eps.eigi[k];
      const double *xr, *xi = NULL;
      if (ki > 0) { xr = COL(&eps.V, k); xi = COL(&eps.V, k + 1); }
      else if (ki < 0) { xr = COL(&eps.V, k - 1); xi = COL(&eps.V, k); }
      else xr = COL(&eps.V, k);
      eigr[i] = kr; eigi[i] = ki; errest[i] = eps.errest[k];
      if (X) {
        if (ki < 0) for (int r = 0; r < n; r++) X[(size_t)i * n + r] = -xi[r];   /* imaginary part, negated */
        else if (ki > 0) memcpy(X + (size_t)i * n, xr, sizeof(double) * (size_t)n);
        else memcpy(X + (size_t)i * n, xr, sizeof(double) * (size_t)n);
      }
      if (Y && eps.twosided) {                                     /* EPSGetLeftEigenvector, real part */
        const double *yr = COL(&eps.W, ki < 0 ? k - 1 : k);
        memcpy(Y + (size_t)i * n, yr, sizeof(double) * (size_t)n);
      }
      if (err) {
        double r;
        if (ki == 0) r = residual_norm(&eps, kr, 0.0, xr, NULL, u, NULL);
        else if (ki > 0) r = residual_norm(&eps, kr, ki, xr, xi, u, NULL);
        else {   /* second of the pair: xi = -V(:,k) */
          double *nxi = (double*)malloc(sizeof(double) * (size_t)n);
          for (int q = 0; q < n; q++) nxi[q] = -xi[q];
          r = residual_norm(&eps, kr, ki, xr, nxi, u, NULL);
          free(nxi);
        }
        err[i] = r / hypot(kr, ki);
      }
    }
    free(u); (void)zero;
  }
  if (info) {
    info->ksp_its = eps.ksp_its;
    info->nconv = eps.nconv; info->its = eps.its; info->reason = eps.reason;
    info->nmatmult = eps.nmatmult; info->gs_passes = eps.V.passes; info->t_ds = eps.t_ds;
  }
done:
  free(eps.dinv); free(eps.kwork); free(eps.lu); free(eps.lu_diag); free(eps.dlu); free(eps.ipiv);
  if (eps.twosided) { free(eps.W.V); free(eps.W.buffer); free(eps.M); free(eps.trowptr); free(eps.tcolidx); free(eps.tval); }
  free(eps.V.V); free(eps.V.buffer); free(eps.eigr); free(eps.eigi); free(eps.errest); free(eps.perm);
  if (eps.ds.A) ds_free(&eps.ds);
  return e;
}
