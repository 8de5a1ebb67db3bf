// ks_ksp.hip -- the linear solves of shift-and-invert on the device.
//
// STSINVERT applies (A - sigma I)^-1 by an inner Krylov solve (the reference's
// STApply_Sinvert -> KSPSolve, src/sys/classes/st/impls/sinvert/sinvert.c:16-30,
// with KSPCG and PCJACOBI as in src/eps/tutorials/ex11.c's test 2).  PETSc is not
// part of the reference tree; the solver here is the preconditioned conjugate
// gradient method in PETSc's KSPCG form (Hestenes-Stiefel recurrences, the
// preconditioned residual norm |B r| for the test rnorm <= max(rtol |B b|, abstol))
// with B = diag(A - sigma I)^-1 (zero diagonal entries replaced by 1, as PCJACOBI
// does).  Every scalar of the recurrence stays on the device; the host reads one
// norm per iteration to test convergence.  Reductions add per-workgroup partial
// sums in a fixed order in a one-workgroup launch, so a solve is deterministic.
#include "ks_internal.hpp"

namespace {

#define KS_PCG_BLOCKS 512   // workgroups of the vector kernels (partials: 2 per workgroup)

// scalar slots in d_s: 0 rz, 1 previous rz, 2 p.q, 3 z.z
enum { S_RZ = 0, S_RZOLD = 1, S_PQ = 2, S_ZZ = 3 };

__device__ inline double block_sum(double v, double *lds) {
  // wave64 sum, then the four waves of the workgroup in order
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < KS_BLOCK / 64; w++) t += lds[w];
  __syncthreads();
  return t;
}

// dinv[i] = 1 / (a_ii - sigma), 1 where that is zero; all ones without a preconditioner
__global__ __launch_bounds__(KS_BLOCK) void k_jacobi(const int32_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ colidx,
                                                    const double *__restrict__ val, int n, double sigma, int jacobi,
                                                    double *__restrict__ dinv) {
  for (int i = blockIdx.x * KS_BLOCK + threadIdx.x; i < n; i += gridDim.x * KS_BLOCK) {
    double d = 0.0;
    if (jacobi) {
      for (int32_t e = rowptr[i]; e < rowptr[i + 1]; e++)
        if (colidx[e] == i) d += val[e];
      d -= sigma;
    }
    dinv[i] = (d == 0.0) ? 1.0 : 1.0 / d;
  }
}

// x = 0, r = b, z = B r; partials of z.r and z.z
__global__ __launch_bounds__(KS_BLOCK) void k_pcg_start(int n, const double *__restrict__ b, double *__restrict__ x,
                                                       double *__restrict__ r, double *__restrict__ z,
                                                       const double *__restrict__ dinv, double *__restrict__ part) {
  __shared__ double lds[KS_BLOCK / 64];
  double zr = 0.0, zz = 0.0;
  for (int i = blockIdx.x * KS_BLOCK + threadIdx.x; i < n; i += gridDim.x * KS_BLOCK) {
    const double ri = b[i], zi = dinv[i] * ri;
    x[i] = 0.0;
    r[i] = ri;
    z[i] = zi;
    zr += zi * ri;
    zz += zi * zi;
  }
  zr = block_sum(zr, lds);
  zz = block_sum(zz, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = zr;
    part[2 * blockIdx.x + 1] = zz;
  }
}

// the new z.r and z.z from the partials; the old z.r moves to S_RZOLD
__global__ __launch_bounds__(KS_BLOCK) void k_pcg_fin_rz(const double *__restrict__ part, int nb,
                                                        double *__restrict__ s) {
  __shared__ double lds[KS_BLOCK / 64];
  double zr = 0.0, zz = 0.0;
  for (int b = threadIdx.x; b < nb; b += KS_BLOCK) {
    zr += part[2 * b];
    zz += part[2 * b + 1];
  }
  zr = block_sum(zr, lds);
  zz = block_sum(zz, lds);
  if (threadIdx.x == 0) {
    s[S_RZOLD] = s[S_RZ];
    s[S_RZ] = zr;
    s[S_ZZ] = zz;
  }
}

// p = z (first iteration) or z + (rz / rzold) p
__global__ __launch_bounds__(KS_BLOCK) void k_pcg_p(int n, const double *__restrict__ z, double *__restrict__ p,
                                                   const double *__restrict__ s, int first) {
  const double beta = first ? 0.0 : s[S_RZ] / s[S_RZOLD];
  for (int i = blockIdx.x * KS_BLOCK + threadIdx.x; i < n; i += gridDim.x * KS_BLOCK)
    p[i] = first ? z[i] : z[i] + beta * p[i];
}

// q -= sigma p (q = A p on entry); partials of p.q
__global__ __launch_bounds__(KS_BLOCK) void k_pcg_q(int n, const double *__restrict__ p, double *__restrict__ q,
                                                   double sigma, double *__restrict__ part) {
  __shared__ double lds[KS_BLOCK / 64];
  double pq = 0.0;
  for (int i = blockIdx.x * KS_BLOCK + threadIdx.x; i < n; i += gridDim.x * KS_BLOCK) {
    const double qi = q[i] - sigma * p[i];
    q[i] = qi;
    pq += p[i] * qi;
  }
  pq = block_sum(pq, lds);
  if (threadIdx.x == 0) part[2 * blockIdx.x] = pq;
}

__global__ __launch_bounds__(KS_BLOCK) void k_pcg_fin_pq(const double *__restrict__ part, int nb,
                                                        double *__restrict__ s) {
  __shared__ double lds[KS_BLOCK / 64];
  double pq = 0.0;
  for (int b = threadIdx.x; b < nb; b += KS_BLOCK) pq += part[2 * b];
  pq = block_sum(pq, lds);
  if (threadIdx.x == 0) s[S_PQ] = pq;
}

// a = rz / pq; x += a p; r -= a q; z = B r; partials of z.r and z.z
__global__ __launch_bounds__(KS_BLOCK) void k_pcg_xr(int n, const double *__restrict__ p,
                                                    const double *__restrict__ q, double *__restrict__ x,
                                                    double *__restrict__ r, double *__restrict__ z,
                                                    const double *__restrict__ dinv, const double *__restrict__ s,
                                                    double *__restrict__ part) {
  __shared__ double lds[KS_BLOCK / 64];
  const double a = s[S_RZ] / s[S_PQ];
  double zr = 0.0, zz = 0.0;
  for (int i = blockIdx.x * KS_BLOCK + threadIdx.x; i < n; i += gridDim.x * KS_BLOCK) {
    x[i] += a * p[i];
    const double ri = r[i] - a * q[i], zi = dinv[i] * ri;
    r[i] = ri;
    z[i] = zi;
    zr += zi * ri;
    zz += zi * zi;
  }
  zr = block_sum(zr, lds);
  zz = block_sum(zz, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = zr;
    part[2 * blockIdx.x + 1] = zz;
  }
}

inline int pcg_blocks(int n) {
  int b = (n + KS_BLOCK - 1) / KS_BLOCK;
  if (b > KS_PCG_BLOCKS) b = KS_PCG_BLOCKS;
  return b < 1 ? 1 : b;
}

}  // namespace

int ks_pcg_jacobi(ks_ctx ctx, ks_csr A, double sigma, int jacobi, double *d_dinv) {
  KS_CHECK(ctx && A && d_dinv, KS_ERR_ARG, "null argument");
  KS_CHECK(A->vprec == KS_F64 && A->nrows == A->ncols, KS_ERR_ARG, "PCG needs a square FP64 matrix");
  hipLaunchKernelGGL(k_jacobi, dim3(pcg_blocks(A->nrows)), dim3(KS_BLOCK), 0, ctx->stream, A->d_rowptr, A->d_colidx,
                     (const double *)A->d_val, A->nrows, sigma, jacobi, d_dinv);
  KS_HIP(hipGetLastError());
  return KS_OK;
}

int ks_pcg_solve(ks_ctx ctx, ks_csr A, double sigma, const double *d_dinv, const double *d_b, double *d_x,
                 double rtol, double abstol, int maxits, double *d_work, int *its, double *rnorm, int *reason) {
  KS_CHECK(ctx && A && d_dinv && d_b && d_x && d_work && its && rnorm && reason, KS_ERR_ARG, "null PCG argument");
  KS_CHECK(A->vprec == KS_F64 && A->nrows == A->ncols, KS_ERR_ARG, "PCG needs a square FP64 matrix");
  static_assert(2 * KS_PCG_BLOCKS <= (size_t)KS_MAXB * KS_NCOLP, "PCG partials exceed the workspace");
  const int n = A->nrows, nb = pcg_blocks(n);
  double *r = d_work, *z = r + n, *p = z + n, *q = p + n, *s = q + n;   // s: 4 scalars
  double *part = ctx->d_partial, h[4];
  hipStream_t st = ctx->stream;
  *its = 0;
  *reason = 0;
  hipLaunchKernelGGL(k_pcg_start, dim3(nb), dim3(KS_BLOCK), 0, st, n, d_b, d_x, r, z, d_dinv, part);
  hipLaunchKernelGGL(k_pcg_fin_rz, dim3(1), dim3(KS_BLOCK), 0, st, (const double *)part, nb, s);
  KS_HIP(hipMemcpyAsync(h, s, sizeof(h), hipMemcpyDeviceToHost, st));
  KS_HIP(hipStreamSynchronize(st));
  const double rnorm0 = sqrt(h[S_ZZ]);
  const double ttol = fmax(rtol * rnorm0, abstol);
  *rnorm = rnorm0;
  if (rnorm0 <= ttol) { *reason = 1; return KS_OK; }
  for (int i = 0; i < maxits; i++) {
    hipLaunchKernelGGL(k_pcg_p, dim3(nb), dim3(KS_BLOCK), 0, st, n, (const double *)z, p, (const double *)s, i == 0);
    int e = ks_spmv(ctx, A, p, q, KS_F64, 0);
    if (e) return e;
    hipLaunchKernelGGL(k_pcg_q, dim3(nb), dim3(KS_BLOCK), 0, st, n, (const double *)p, q, sigma, part);
    hipLaunchKernelGGL(k_pcg_fin_pq, dim3(1), dim3(KS_BLOCK), 0, st, (const double *)part, nb, s);
    hipLaunchKernelGGL(k_pcg_xr, dim3(nb), dim3(KS_BLOCK), 0, st, n, (const double *)p, (const double *)q, d_x, r, z,
                       d_dinv, (const double *)s, part);
    hipLaunchKernelGGL(k_pcg_fin_rz, dim3(1), dim3(KS_BLOCK), 0, st, (const double *)part, nb, s);
    KS_HIP(hipMemcpyAsync(h, s, sizeof(h), hipMemcpyDeviceToHost, st));
    KS_HIP(hipStreamSynchronize(st));
    *its = i + 1;
    *rnorm = sqrt(h[S_ZZ]);
    if (!(h[S_PQ] > 0.0)) { *reason = -8; return KS_OK; }    // indefinite or breakdown (KSP_DIVERGED_INDEFINITE_MAT)
    if (*rnorm <= ttol) { *reason = 1; return KS_OK; }
  }
  *reason = -3;                                                // KSP_DIVERGED_ITS
  return KS_OK;
}
