/*
 * ks_hip.h -- C ABI of libks_hip.so, the MI355X (gfx950) kernels of the
 * Krylov-Schur hot path.  Plain pointers and sizes only; every call returns
 * 0 (KS_OK) or a negative KS_ERR_* code.  Device pointers are HIP device
 * memory obtained from ks_malloc (or any hipMalloc'd buffer).  All work is
 * stream-ordered on the context's HIP stream; calls that return host values
 * synchronize that stream.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the SLEPc 3.22 source tree; PETSc functions are outside it).
 */
#ifndef KS_HIP_H
#define KS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  KS_OK = 0,
  KS_ERR_ARG = -1,        /* bad argument (sizes, null pointers, out of range) */
  KS_ERR_HIP = -2,        /* HIP runtime error */
  KS_ERR_NOMEM = -3,      /* device allocation failed */
  KS_ERR_NODEV = -4,      /* no gfx950 device visible */
  KS_ERR_INNER = -5,      /* invalid inner product (negative self dot beyond tolerance) */
  KS_ERR_RANGE = -6,      /* size beyond a kernel's supported maximum */
  KS_ERR_PIVOT = -7,      /* zero pivot in a factorization (ILU(0), LU) */
  KS_ERR_TIMEOUT = -8     /* a bounded wait of a kernel on other workgroups ran out */
};

typedef enum { KS_F64 = 0, KS_F32 = 1 } ks_prec;

/* orthogonalization refinement, BVOrthogRefineType (include/slepcbv.h) */
typedef enum { KS_REFINE_IFNEEDED = 0, KS_REFINE_NEVER = 1, KS_REFINE_ALWAYS = 2 } ks_refine;

/* maximum number of columns one Gram-Schmidt pass or restart GEMM handles
   (ncv <= 256 plus constraints) */
#define KS_MAX_COLS 320

typedef struct ks_ctx_s *ks_ctx;
typedef struct ks_csr_s *ks_csr;
typedef struct ks_pool_s *ks_pool;
typedef struct ks_group_s *ks_group;
typedef struct ks_graph_s *ks_graph;
typedef struct ks_ilu_s *ks_ilu;
typedef struct ks_lu_s *ks_lu;

/* Device-side state of the fused CGS+DGKS orthonormalization (the scalars of
   BVOrthogonalizeGS, src/sys/classes/bv/interface/bvorthog.c:145-217, kept on
   the GPU so an Arnoldi expansion needs no host round trip). */
typedef struct {
  int32_t active;          /* 1 until a column is found linearly dependent */
  int32_t breakdown_col;   /* column index where lindep occurred, or -1 */
  int32_t col;             /* column being orthogonalized */
  int32_t pass_done;       /* CGS passes completed on that column */
  int32_t done;            /* 1 once the column is finalized */
  int32_t lindep;          /* lindep flag of the last finalized column */
  int32_t error;           /* KS_ERR_* raised on the device, 0 if none */
  int32_t need_norm;       /* estimate beta^2 - sum c^2 <= 0: use the exact norm */
  int64_t total_passes;    /* executed CGS passes since the last reset */
  double  onrm;            /* |v| before the current pass */
  double  nrm;             /* |q| after the current pass (estimate or exact) */
  double  beta_last;       /* norm of the last finalized column (H(j+1,j)) */
  double  scale;           /* 1/nrm to apply to the finalized column, or 1 */
  double  sumsq;           /* sum of squares of the current pass coefficients */
  double  exact_sq;        /* exact |q|^2 from the update pass */
  double  prescale;        /* 1/nrm written by the final pass's sweep itself ... */
  int32_t prescale_col;    /* ... for this column (-1: none) ... */
  int32_t prescale_pass;   /* ... and this pass */
  int32_t halt_col;        /* a lean orthonormalization stopped the expansion at this column (-1: none) ... */
  int32_t halt_need;       /* ... and needs 1: the third DGKS pass, 2: the scaling (ks_bv_orthonormalize_resume) */
  int32_t defer_col;       /* the first fused pass of this column did not store its vector (-1: none): */
  int32_t defer_pass;      /* ... the sweep of the pass after this one forms it again from the same loads */
} ks_status;

/* ---------------------------------------------------------------- runtime */
/* Selects the device; fails with KS_ERR_NODEV unless it is gfx950. */
int ks_init(int device);
/* Context = one HIP stream (default priority) + reduction workspace + status. */
int ks_ctx_create(ks_ctx *ctx);
int ks_ctx_destroy(ks_ctx ctx);
int ks_ctx_sync(ks_ctx ctx);
void *ks_ctx_stream(ks_ctx ctx);
int ks_malloc(void **ptr, size_t bytes);
int ks_free(void *ptr);
int ks_memcpy_htod(ks_ctx ctx, void *dst, const void *src, size_t bytes);
int ks_memcpy_dtoh(ks_ctx ctx, void *dst, const void *src, size_t bytes);
int ks_memcpy_dtod(ks_ctx ctx, void *dst, const void *src, size_t bytes);
int ks_memset_zero(ks_ctx ctx, void *dst, size_t bytes);   /* fill kernel, not hipMemset */
/* stream-ordered host->device copy that returns without waiting; src must stay
   untouched until the stream passes this point (use pinned memory) */
int ks_memcpy_htod_async(ks_ctx ctx, void *dst, const void *src, size_t bytes);
/* stream-ordered device-to-host copy without a wait (dst: pinned host memory) */
int ks_memcpy_dtoh_async(ks_ctx ctx, void *dst, const void *src, size_t bytes);
int ks_host_alloc(void **ptr, size_t bytes);   /* pinned host memory */
int ks_host_free(void *ptr);
const char *ks_error_string(int code);
int ks_mem_info(size_t *free_bytes, size_t *total_bytes);
/* 1 in *idle when every operation enqueued on the context's stream has
   finished (hipStreamQuery), 0 while work is pending; never waits. */
int ks_ctx_query(ks_ctx ctx, int *idle);

/* ---------------------------------------------------------------- pools
 * No reference counterpart (PETSc allocates per object).  A pool is one device
 * arena (plus an optional pinned host arena) with a first-fit allocator: while
 * a pool is current (ks_pool_use), ks_malloc/ks_host_alloc and the CSR
 * constructors carve from it, and ks_free/ks_host_free return pointers to the
 * pool that holds them.  Serving admits and retires solves against it without
 * hipMalloc/hipFree.  ks_pool_use returns the previously current pool (NULL =
 * plain hipMalloc).  Destroying a pool with live blocks fails. */
int ks_pool_create(ks_pool *pool, size_t dev_bytes, size_t host_bytes);
int ks_pool_destroy(ks_pool pool);
ks_pool ks_pool_use(ks_pool pool);
int ks_pool_info(ks_pool pool, size_t *used, size_t *peak, size_t *size, size_t *largest_free);

/* ---------------------------------------------------------------- CSR matrix
 * Replaces PETSc MatMult_SeqAIJHIPSPARSE / MatMult_SeqAIJ, called from
 * BVMatMultColumn (src/sys/classes/bv/interface/bvops.c:879), STApply_Generic
 * (src/sys/classes/st/interface/stsolve.c:22) and EPSComputeResidualNorm_Private
 * (src/eps/interface/epssolve.c:688).
 * Uploads the host CSR (int32 indices, values in vprec) and builds the
 * row-block schedule from the nnz/row histogram. */
int ks_csr_create(ks_csr *A, ks_ctx ctx, int nrows, int ncols, int64_t nnz,
                  const int32_t *rowptr, const int32_t *colidx, const void *val, ks_prec vprec);
/* Same, from arrays already resident on the device (copied, not adopted). */
int ks_csr_create_device(ks_csr *A, ks_ctx ctx, int nrows, int ncols, int64_t nnz,
                         const int32_t *d_rowptr, const int32_t *d_colidx, const void *d_val,
                         ks_prec vprec);
int ks_csr_destroy(ks_csr A);
/* New matrix with the same pattern and the values converted to `to` on the
   device (FP32 copy of an FP64 operator for the mixed-precision phase). */
int ks_csr_convert(ks_csr A, ks_ctx ctx, ks_prec to, ks_csr *out);
/* New matrix A^T (ncols x nrows), built on the device: each row lists its entries in
   increasing column order.  Replaces MatCreateHermitianTranspose of the operator in
   two-sided Krylov-Schur (src/eps/impls/krylov/krylovschur/ks-twosided.c:138) by an
   explicit matrix, so products with A^T use the same SpMV kernel. */
int ks_csr_transpose(ks_csr A, ks_ctx ctx, ks_csr *out);
/* Copies a device matrix back to host arrays (rowptr: nrows + 1 entries; colidx
   and val: nnz entries, val in the matrix's precision) and waits for the copy.
   No counterpart in the reference (whose matrices live on the host); for callers
   and tests that read back what the device built, such as the transpose above. */
int ks_csr_download(ks_csr A, ks_ctx ctx, int32_t *rowptr, int32_t *colidx, void *val);
/* The inner solve of shift-and-invert: (A - sigma I) x = b by preconditioned
   conjugate gradients (KSPCG's recurrences, preconditioned residual norm,
   rnorm <= max(rtol |B b|, abstol), at most maxits iterations) with
   B = diag(A - sigma I)^-1 from ks_pcg_jacobi (jacobi = 0: B = I).  Replaces
   STApply_Sinvert's KSPSolve (src/sys/classes/st/impls/sinvert/sinvert.c:16-30).
   The convergence tests run on the device; the host reads the solver state
   once per 8 iterations.  d_work holds 4n + 64 doubles; reason > 0 converged,
   -3 iteration limit, -8 p^T (A - sigma I) p <= 0 (not positive definite). */
int ks_pcg_jacobi(ks_ctx ctx, ks_csr A, double sigma, int jacobi, double *d_dinv);
int ks_pcg_solve(ks_ctx ctx, ks_csr A, double sigma, const double *d_dinv, const double *d_b, double *d_x,
                 double rtol, double abstol, int maxits, double *d_work, int *its, double *rnorm, int *reason);
/* The same solve by BiCGStab (KSPBCGS's recurrences with the left
   preconditioner B, PETSc src/ksp/ksp/impls/bcgs/bcgs.c), which needs no
   definiteness: the inner solver of shift-and-invert at a target inside the
   spectrum.  The iteration ends early when |s| <= the threshold.  d_work holds
   6n + 64 doubles; reason > 0 converged, -3 iteration limit, -5 breakdown
   ((rhat, v) = 0, (t, t) = 0, rho = 0 or omega = 0). */
int ks_bcgs_solve(ks_ctx ctx, ks_csr A, double sigma, const double *d_dinv, const double *d_b, double *d_x,
                  double rtol, double abstol, int maxits, double *d_work, int *its, double *rnorm, int *reason);
/* QR of the n x k FP64 panel V (leading dimension ld) by TSQR: Householder QR
   of row blocks of 2k rows (at least 64) in LDS, their triangular factors
   reduced by the same factorization up a tree.  R (k x k, host memory, leading
   dimension ldr) gets the factor; form_q != 0 overwrites V with Q, 0 leaves V
   as it is.  Replaces BVOrthogonalize_LAPACK_TSQR and _TSQR_OnlyR
   (src/sys/classes/bv/interface/bvlapack.c:347-560), which run LAPACK per
   process and reduce across processes.  Q and R are LAPACK's Householder QR up
   to the signs of R's rows.  Allocates its workspace (about 1.5 n k doubles)
   and waits for the stream. */
int ks_bv_tsqr(ks_ctx ctx, int n, double *d_V, int ld, int k, int form_q, double *R, int ldr);
/* ILU(0) of A - sigma I on A's pattern (Saad, Iterative Methods for Sparse Linear
   Systems, Algorithm 10.4; PETSc's PCILU with levels 0 and the natural
   ordering, its defaults), factored on the device one dependency level of rows
   at a time; A's rows must list increasing columns and hold their diagonal.
   Fails with KS_ERR_PIVOT on a zero pivot.  Applying it is one launch for each
   factor, rows pipelined through their published values (ks_ilu0_levels gives
   the dependency depth of each).  Replaces PCSetUp_ILU / PCApply_ILU as
   KSPSolve reaches them from STApply_Sinvert
   (src/sys/classes/st/impls/sinvert/sinvert.c:16-30). */
int ks_ilu0_create(ks_ctx ctx, ks_csr A, double sigma, ks_ilu *out);
int ks_ilu0_destroy(ks_ilu M);
int ks_ilu0_levels(ks_ilu M, int *nlower, int *nupper);
/* ks_bcgs_solve with the ILU(0) preconditioner B = (L U)^-1 */
int ks_bcgs_solve_ilu(ks_ctx ctx, ks_csr A, double sigma, ks_ilu M, const double *d_b, double *d_x, double rtol,
                      double abstol, int maxits, double *d_work, int *its, double *rnorm, int *reason);
/* The direct inner solve of shift-and-invert (KSPPREONLY with PCLU, the
   reference's default: src/sys/classes/st/interface/stsles.c:47-54,
   st/impls/sinvert/sinvert.c:131-154): LU with partial pivoting of A - sigma I
   held dense in HBM (n^2 doubles, n <= KS_LU_MAXN), factored on the device by
   LAPACK dgetrf's right-looking blocked algorithm (512-column panels of
   64-column sub-panels, pivots as idamax, updates on FP64 MFMA).  ks_lu_solve
   gives x = (A - sigma I)^-1 b (b and x device vectors of length n; they may be
   the same), stream-ordered, one launch per triangular solve.  A pivot below
   100 machine epsilons fails the factorization with KS_ERR_PIVOT. */
#define KS_LU_MAXN 32768
int ks_lu_create(ks_ctx ctx, ks_csr A, double sigma, ks_lu *out);
int ks_lu_solve(ks_ctx ctx, ks_lu F, const double *d_b, double *d_x);
int ks_lu_destroy(ks_lu F);
/* Waits for the context's stream and reports whether every solve so far ran
   clean (the triangular solves' bounded waits never ran out): KS_OK or
   KS_ERR_TIMEOUT.  No counterpart in the reference (its LU solves are synchronous
   host code); for callers that check the asynchronous solves, such as tests. */
int ks_lu_check(ks_ctx ctx, ks_lu F);
/* nrows, nnz, max row length and the kernel variant chosen (0 stream, 1 vector) */
int ks_csr_info(ks_csr A, int *nrows, int64_t *nnz, int *max_row, int *variant, int *nblocks);
/* y = A x.  x, y device vectors in xprec.  When gate != 0 the product is
   skipped once the context status says an expansion broke down. */
int ks_spmv(ks_ctx ctx, ks_csr A, const void *x, void *y, ks_prec xprec, int gate);

/* ---------------------------------------------------------------- BV kernels
 * V is column-major with leading dimension ld (elements), columns in prec. */

/* q[0..k) = V(:,0..k)^T y.  Replaces BVDotVec_BLAS_Private
   (src/sys/classes/bv/interface/bvblas.c:240-261) / BVDotVec_BLAS_HIP. */
int ks_bv_dotvec(ks_ctx ctx, ks_prec prec, int n, const void *V, int ld, int k,
                 const void *y, double *d_q);
/* y = beta*y + alpha*V(:,0..k) q.  Replaces BVMultVec_BLAS_Private (bvblas.c:56-67). */
int ks_bv_multvec(ks_ctx ctx, ks_prec prec, int n, double alpha, const void *V, int ld, int k,
                  const double *d_q, double beta, void *y);

/* Orthonormalize column w = V(:,k) against V(:,0..k) with CGS and the
   refinement policy (DGKS for IFNEEDED, eta), then scale it by 1/|q|:
   BVOrthonormalizeColumn (src/sys/classes/bv/interface/bvorthog.c:380-427)
   with BVOrthogonalizeGS/BVOrthogonalizeCGS1 (bvorthog.c:91-217) and the
   coefficient helpers (include/slepc/private/bvimpl.h:289-415) fused into
   stream-ordered kernels.  d_buf is the coefficient buffer with leading
   dimension ldbuf: column 0 receives the scratch coefficients c, column
   `hcol` accumulates h and gets |q| at row k (0 if lindep).  Fully
   asynchronous: the decision to run a second/third pass, lindep and the
   scaling factor stay on the device (ks_status).  `col` is the column index
   recorded in the status.  scale=0 skips the normalization
   (BVOrthogonalizeColumn semantics); safe=1 orthogonalizes with a single pass
   and reports lindep with norm 0 (BV_OrthogonalizeColumn_Safe, bvimpl.h:446-465). */
int ks_bv_orthonormalize_column(ks_ctx ctx, ks_prec prec, int n, void *V, int ld, int k,
                                double *d_buf, int ldbuf, int hcol, int col, ks_refine refine,
                                double eta, int scale, int safe);
/* The same orthonormalization without a context and without allocating: no
   counterpart in the reference.  For callers that cannot allocate device
   memory (code under HIP graph capture, serving loops, harnesses that forbid
   hipMalloc).  The status, the finishing counter and the partial sums live in
   one file-scope device array of the library (3.2 MB per GPU, held until the
   process ends), so k <= 128 (KS_ERR_RANGE beyond).  Work is enqueued on
   `stream` (a hipStream_t; NULL = the null stream); the status of the column
   (ks_status: norm, lindep, passes) is copied to d_status_out (device memory,
   may be NULL) in stream order.  Every call on a GPU shares that GPU's array:
   calls that overlap on one GPU (other streams, other host threads) race. */
int ks_bv_orthonormalize_column_noalloc(void *stream, ks_prec prec, int n, void *V, int ld, int k,
                                        double *d_buf, int ldbuf, int hcol, int col, ks_refine refine,
                                        double eta, int scale, int safe, ks_status *d_status_out);
/* The same orthonormalization with only the launches a column normally needs:
   the first-pass dots and the DGKS passes that IFNEEDED takes unless a third is
   due (two), or ALWAYS's two, each with its finish, and no separate scaling --
   the last pass's sweep stores the normalized column.  When the column needs
   more (a third pass, or a scaling the sweep could not apply) the finish stops
   the expansion on the device instead: status active = 0, halt_col = col,
   halt_need = 1 or 2, so every later gated launch does nothing.  The caller
   then runs ks_bv_orthonormalize_resume for that column and enqueues the rest
   of the expansion again; the result is bit for bit that of
   ks_bv_orthonormalize_column.  Other refinement types, safe mode, scale = 0
   and the LDS sweeps (k > 128) fall back to the full sequence. */
int ks_bv_orthonormalize_column_lean(ks_ctx ctx, ks_prec prec, int n, void *V, int ld, int k, double *d_buf,
                                     int ldbuf, int hcol, int col, ks_refine refine, double eta, int scale,
                                     int safe);
/* Completes a column stopped by ks_bv_orthonormalize_column_lean (the status
   names it): re-activates the status and enqueues the remaining pass and the
   scaling, each gated on the device as in ks_bv_orthonormalize_column. */
int ks_bv_orthonormalize_resume(ks_ctx ctx, ks_prec prec, int n, void *V, int ld, int k, double *d_buf, int ldbuf,
                                int hcol, int col, ks_refine refine, double eta);
/* Reset the status for a new expansion (active=1, breakdown_col=-1). */
int ks_status_reset(ks_ctx ctx);
/* Copy the status to the host (synchronizes the stream). */
int ks_status_get(ks_ctx ctx, ks_status *out);
/* post: enqueue a copy of the status to pinned host memory; read: wait for the
   stream, then return the posted copy (one wait instead of a copy round trip) */
int ks_status_post(ks_ctx ctx);
int ks_status_read(ks_ctx ctx, ks_status *out);

/* x *= alpha (alpha==0 zeroes).  Replaces BVScale_BLAS_Private (bvblas.c:266-278). */
int ks_bv_scale(ks_ctx ctx, ks_prec prec, int n, void *x, double alpha);
/* |x|_2 to the host.  Replaces BVNorm_LAPACK_Private NORM_2 (bvlapack.c:37-86). */
int ks_bv_norm(ks_ctx ctx, ks_prec prec, int n, const void *x, double *nrm);
/* y += alpha x */
int ks_bv_axpy(ks_ctx ctx, ks_prec prec, int n, double alpha, const void *x, void *y);
/* y = x (device copy).  Replaces BVCopyColumn_Mat (bvmat.c:250-260). */
int ks_bv_copy(ks_ctx ctx, ks_prec prec, int n, const void *x, void *y);
/* x = uniform(-1,1) from the counter generator
   u_i = 2*(splitmix64(seed*0x9E3779B97F4A7C15 + i) >> 11)*2^-53 - 1.
   Replaces BVSetRandomColumn (src/sys/classes/bv/interface/bvops.c:360-379). */
int ks_bv_random(ks_ctx ctx, ks_prec prec, int n, void *x, uint64_t seed);
/* convert between precisions (y in py, x in px) */
int ks_bv_convert(ks_ctx ctx, int n, ks_prec px, const void *x, ks_prec py, void *y);

/* In-place restart GEMM on the FP64 matrix cores:
   V(:,s:e-1) = V(:,0:k-1) * Q(0:k-1, s:e-1), Q device, column-major, ldq.
   Replaces BVMultInPlace_BLAS_Private (bvblas.c:74-106) / BVMultInPlace_BLAS_HIP
   (src/sys/classes/bv/impls/hip/bvhip.hip.cpp:65-112) without the n x (e-s)
   temporary.  Requires k <= KS_MAX_COLS and e <= KS_MAX_COLS. */
/* G = Y^T X for tall-skinny X (n x p, leading dimension ldx) and Y (n x q):
   G(i,j) = y_i . x_j into the device array d_G (q x p, leading dimension ldg),
   stream ordered.  The matrix form of BVDot (src/sys/classes/bv/interface/
   bvglobal.c:86-120, BVDot_BLAS_Private bvblas.c:199-233). */
int ks_bv_gram(ks_ctx ctx, ks_prec prec, int n, const void *X, int ldx, int p, const void *Y, int ldy, int q,
               double *d_G, int ldg);
int ks_bv_multinplace(ks_ctx ctx, ks_prec prec, int n, void *V, int ld, int k,
                      const double *d_Q, int ldq, int s, int e);

/* ---------------------------------------------------------------- profiling
 * When enabled, every kernel launch of the context is bracketed by HIP events
 * on the context stream; collect() accumulates per-class device time, the
 * algorithmic HBM bytes and flops of the launches that did work. */
enum { KS_PROF_SPMV = 0, KS_PROF_GS_DOT1, KS_PROF_GS_UPD1, KS_PROF_GS_FIN, KS_PROF_GS_UPDR,
       KS_PROF_SCALE, KS_PROF_NORM, KS_PROF_AXPY, KS_PROF_COPY, KS_PROF_MULTINPLACE,
       KS_PROF_DOTVEC, KS_PROF_MULTVEC, KS_PROF_OTHER, KS_PROF_NCLASS };
typedef struct {
  int64_t launches;      /* launches that did work */
  int64_t skipped;       /* launches gated off on the device */
  double  ms;            /* device time of the launches that did work */
  double  bytes;         /* algorithmic HBM bytes */
  double  flops;
} ks_prof_entry;
int ks_prof_enable(ks_ctx ctx, int on);
/* events only for the classes whose bit (1 << class) is set in mask; 0 turns profiling off */
int ks_prof_enable_classes(ks_ctx ctx, unsigned mask);
int ks_prof_collect(ks_ctx ctx);
int ks_prof_get(ks_ctx ctx, int cls, ks_prof_entry *out);
int ks_prof_reset(ks_ctx ctx);
const char *ks_prof_name(int cls);

/* ---------------------------------------------------------------- grouped steps (serving)
 * No reference counterpart (SLEPc runs one solve at a time).  One Arnoldi step
 * for each of up to 16 independent FP64 solves with one launch per kernel
 * kind on the group's own stream: member m computes column k of its basis as
 * A * column k-1 (gated on its status like ks_spmv with gate), then the whole
 * CGS+DGKS orthonormalization of column k against its first k columns and the
 * scaling (as ks_bv_orthonormalize_column with hcol, col, scale = 1, safe = 0).
 * Each member uses its own context's status and reduction workspace, so its
 * results are its own; the member's stream must not run work meanwhile
 * (order them with ks_group_wait_ctx / ks_ctx_wait_group).  Members need
 * k <= 128 and n * 8 < 4 GiB. */
typedef struct {
  ks_ctx ctx;       /* the member's context (status, reduction workspace) */
  ks_csr A;         /* FP64 values */
  void *V;          /* FP64 basis, leading dimension ld */
  int ld, n;
  int k;            /* column written by the SpMV and orthonormalized */
  double *buf;      /* coefficient buffer, leading dimension ldbuf */
  int ldbuf, hcol, col;
} ks_group_member;
int ks_group_create(ks_group *group);
int ks_group_destroy(ks_group group);
int ks_group_step(ks_group group, int nmem, const ks_group_member *members, ks_refine refine, double eta);
/* steps enqueued on the group that have not finished yet */
int ks_group_inflight(ks_group group, int *steps);
/* the group's stream waits for the work enqueued so far on ctx, and the reverse */
int ks_group_wait_ctx(ks_group group, ks_ctx ctx);
int ks_ctx_wait_group(ks_ctx ctx, ks_group group);

/* ---------------------------------------------------------------- captured launch sequences
 * Everything enqueued on the context's stream between ks_graph_begin and
 * ks_graph_end is recorded (hipStreamBeginCapture) instead of run;
 * ks_graph_launch replays it on the context stream with one call.  A graph
 * keeps the arguments it was recorded with, so it is valid only while every
 * buffer it touches stays allocated where it was; it records no memset nodes
 * (fills are kernels).  ks_graph_begin fails with KS_ERR_ARG while the
 * context's per-kernel events are on.  No counterpart in the reference: the
 * replay of one Arnoldi step (SpMV + CGS/DGKS + scaling) for small problems,
 * where launch overhead rivals the kernels. */
int ks_graph_begin(ks_ctx ctx);
int ks_graph_end(ks_ctx ctx, ks_graph *graph);   /* *graph = NULL when the capture failed */
int ks_graph_launch(ks_ctx ctx, ks_graph graph);
int ks_graph_destroy(ks_graph graph);

#ifdef __cplusplus
}
#endif
#endif
