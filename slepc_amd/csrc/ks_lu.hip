// ks_lu.hip -- the direct inner solve of shift-and-invert: LU factorization with
// partial pivoting of A - sigma I held dense in HBM, and the triangular solves.
//
// The reference's default inner solver for STSINVERT is KSPPREONLY with PCLU
// (src/sys/classes/st/interface/stsles.c:47-54, sinvert.c:131-154), a sparse LU.
// Here the shifted matrix is expanded to a dense ldn x ldn array (ldn = n
// rounded up to 64, the padding an identity block, so every tile is whole) and
// factored by the right-looking blocked algorithm of LAPACK's dgetrf with
// panels of 64 columns:
//   * panel: per column, one workgroup finds the pivot (largest magnitude,
//     lowest row on ties, as idamax) and swaps it in within the panel; a
//     row-parallel launch scales the column by the pivot's reciprocal and
//     updates the panel's remaining columns (dgetf2's dscal and dger);
//   * the panel's row interchanges applied to the columns outside it (dlaswp);
//   * U12 = L11^-1 A12, one wavefront per column with L11 in LDS (dtrsm);
//   * A22 -= L21 U12 on v_mfma_f64_16x16x4f64 in 64 x 64 tiles (dgemm).
// The solve applies the row permutation as a gather, then the unit-lower and
// upper triangular solves block by block: a 64 x 64 diagonal block by one
// wavefront, then a row-parallel update of the rest of the vector.
// Memory is n^2 doubles (for n = 32,768: 8.6 GB), so the factorization serves
// matrices of up to that size; larger ones take an iterative inner solver.
#include <vector>

#include "ks_internal.hpp"

namespace {

#define KS_LU_NB 64
#define KS_LU_PIV_NT 1024

// A = 0, then A(i, j) = a_ij - sigma [i == j] from the CSR, and the padding rows
// and columns an identity block
__global__ __launch_bounds__(KS_BLOCK) void k_lu_fill(const int32_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ colidx,
                                                     const double *__restrict__ val, int n, int ldn, double sigma,
                                                     double *__restrict__ A) {
  for (int i = blockIdx.x * KS_BLOCK + threadIdx.x; i < ldn; i += gridDim.x * KS_BLOCK) {
    if (i < n) {
      for (int32_t q = rowptr[i]; q < rowptr[i + 1]; q++) A[(size_t)i + (size_t)colidx[q] * ldn] += val[q];
      A[(size_t)i + (size_t)i * ldn] -= sigma;
    } else {
      A[(size_t)i + (size_t)i * ldn] = 1.0;
    }
  }
}

// pivot of column j (rows j..ldn-1), its row swapped with row j inside the panel
// columns j0..j0+KS_LU_NB-1; zero pivot -> *flag = 1
__global__ __launch_bounds__(KS_LU_PIV_NT) void k_lu_pivot(double *__restrict__ A, int ldn, int j, int j0,
                                                          int *__restrict__ ipiv, int *flag) {
  __shared__ double smax[KS_LU_PIV_NT / 64];
  __shared__ int sidx[KS_LU_PIV_NT / 64];
  __shared__ int sp;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double *col = A + (size_t)j * ldn;
  double best = -1.0;
  int bi = ldn;
  for (int i = j + tid; i < ldn; i += KS_LU_PIV_NT) {
    const double a = fabs(col[i]);
    if (a > best) {   // rows visited in increasing order: the first maximum stays
      best = a;
      bi = i;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double ob = __shfl_down(best, off, 64);
    const int oi = __shfl_down(bi, off, 64);
    if (ob > best || (ob == best && oi < bi)) {
      best = ob;
      bi = oi;
    }
  }
  if (lane == 0) {
    smax[wave] = best;
    sidx[wave] = bi;
  }
  __syncthreads();
  if (tid == 0) {
    double b = smax[0];
    int p = sidx[0];
    for (int w = 1; w < KS_LU_PIV_NT / 64; w++)
      if (smax[w] > b || (smax[w] == b && sidx[w] < p)) {
        b = smax[w];
        p = sidx[w];
      }
    if (!(b > 0.0)) {
      *flag = 1;
      p = j;
    }
    ipiv[j] = p;
    sp = p;
  }
  __syncthreads();
  const int p = sp;
  if (p != j && tid < KS_LU_NB) {
    double *a = A + (size_t)(j0 + tid) * ldn;
    const double t = a[j];
    a[j] = a[p];
    a[p] = t;
  }
}

// rows i > j: l_ij = a_ij / a_jj (as the reciprocal's product, dgetf2's dscal),
// then a_ic -= l_ij a_jc for the panel's columns c = j+1 .. j0+KS_LU_NB-1
__global__ __launch_bounds__(KS_BLOCK) void k_lu_panel_col(double *__restrict__ A, int ldn, int j, int j0) {
  const int i = j + 1 + blockIdx.x * KS_BLOCK + threadIdx.x;
  if (i >= ldn) return;
  const double r = 1.0 / A[(size_t)j + (size_t)j * ldn];
  const double l = A[(size_t)i + (size_t)j * ldn] * r;
  A[(size_t)i + (size_t)j * ldn] = l;
  for (int c = j + 1; c < j0 + KS_LU_NB; c++) A[(size_t)i + (size_t)c * ldn] -= l * A[(size_t)j + (size_t)c * ldn];
}

// the panel's interchanges (rows j0..j0+NB-1 with ipiv) on every column outside it
__global__ __launch_bounds__(KS_BLOCK) void k_lu_laswp(double *__restrict__ A, int ldn, int j0,
                                                      const int *__restrict__ ipiv) {
  const int t = blockIdx.x * KS_BLOCK + threadIdx.x;
  const int c = t < j0 ? t : t + KS_LU_NB;
  if (c >= ldn) return;
  double *a = A + (size_t)c * ldn;
  for (int j = j0; j < j0 + KS_LU_NB; j++) {
    const int p = ipiv[j];
    if (p != j) {
      const double x = a[j];
      a[j] = a[p];
      a[p] = x;
    }
  }
}

// U12 = L11^-1 A12 for the columns right of the panel: a wavefront per column,
// lane i holding row j0+i; L11 (unit lower) staged in LDS
__global__ __launch_bounds__(KS_BLOCK) void k_lu_trsm(double *__restrict__ A, int ldn, int j0) {
  __shared__ double L[KS_LU_NB][KS_LU_NB + 1];   // L[k][i] = A(j0+i, j0+k)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int e = tid; e < KS_LU_NB * KS_LU_NB; e += KS_BLOCK) {
    const int i = e % KS_LU_NB, k = e / KS_LU_NB;
    L[k][i] = A[(size_t)(j0 + i) + (size_t)(j0 + k) * ldn];
  }
  __syncthreads();
  const int ncols = ldn - j0 - KS_LU_NB;
  for (int cc = blockIdx.x * KS_WAVES + wave; cc < ncols; cc += gridDim.x * KS_WAVES) {
    double *a = A + (size_t)(j0 + KS_LU_NB + cc) * ldn + j0;
    double x = a[lane];
    for (int k = 0; k < KS_LU_NB - 1; k++) {
      const double xk = __shfl(x, k, 64);
      if (lane > k) x -= L[k][lane] * xk;
    }
    a[lane] = x;
  }
}

typedef double v4d __attribute__((ext_vector_type(4)));

// A22 -= L21 U12, one 64 x 64 tile of A22 per workgroup: the 64 x 64 blocks of
// L21 and U12 in LDS, each wavefront a 32 x 32 quarter (2 x 2 MFMA tiles over
// the 16 k-steps of 4)
__global__ __launch_bounds__(KS_BLOCK) void k_lu_gemm(double *__restrict__ A, int ldn, int j0) {
  __shared__ double sL[KS_LU_NB][KS_LU_NB + 1];   // sL[k][r] = L21(r, k)
  __shared__ double sU[KS_LU_NB][KS_LU_NB + 1];   // sU[k][c] = U12(k, c)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int base = j0 + KS_LU_NB;
  const int r0 = base + blockIdx.x * KS_LU_NB, c0 = base + blockIdx.y * KS_LU_NB;
  for (int e = tid; e < KS_LU_NB * KS_LU_NB; e += KS_BLOCK) {
    const int i = e % KS_LU_NB, k = e / KS_LU_NB;
    sL[k][i] = A[(size_t)(r0 + i) + (size_t)(j0 + k) * ldn];
    sU[i][k] = A[(size_t)(j0 + i) + (size_t)(c0 + k) * ldn];
  }
  __syncthreads();
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  v4d acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
  const int lr = lane & 15, lk = lane >> 4;
#pragma unroll 4
  for (int kk = 0; kk < KS_LU_NB / 4; kk++) {
    const int k = 4 * kk + lk;
    const double a0 = sL[k][wr + lr], a1 = sL[k][wr + 16 + lr];
    const double b0 = sU[k][wc + lr], b1 = sU[k][wc + 16 + lr];
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
  }
  // register q of lane l holds row l/16 + 4q, column l%16 of a 16 x 16 tile
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int r = r0 + wr + 16 * a + lk + 4 * q, c = c0 + wc + 16 * b + lr;
        A[(size_t)r + (size_t)c * ldn] -= acc[a][b][q];
      }
}

// x(i) = b(perm(i)) for i < n, 0 in the padding
__global__ __launch_bounds__(KS_BLOCK) void k_lu_permute(const int *__restrict__ perm, const double *__restrict__ b,
                                                        int n, int ldn, double *__restrict__ x) {
  for (int i = blockIdx.x * KS_BLOCK + threadIdx.x; i < ldn; i += gridDim.x * KS_BLOCK)
    x[i] = i < n ? b[perm[i]] : 0.0;
}

// the 64 x 64 diagonal block b0..b0+63 of L (unit lower, LOWER = 1) or U (upper)
// solved in place in x by one wavefront, lane i holding x(b0+i)
template <bool LOWER>
__global__ __launch_bounds__(64) void k_lu_diag(const double *__restrict__ A, int ldn, int b0, double *__restrict__ x) {
  const int lane = threadIdx.x;
  double xi = x[b0 + lane];
  if (LOWER) {
    for (int k = 0; k < KS_LU_NB - 1; k++) {
      const double xk = __shfl(xi, k, 64);
      if (lane > k) xi -= A[(size_t)(b0 + lane) + (size_t)(b0 + k) * ldn] * xk;
    }
  } else {
    for (int k = KS_LU_NB - 1; k >= 0; k--) {
      if (lane == k) xi /= A[(size_t)(b0 + k) + (size_t)(b0 + k) * ldn];
      const double xk = __shfl(xi, k, 64);
      if (lane < k) xi -= A[(size_t)(b0 + lane) + (size_t)(b0 + k) * ldn] * xk;
    }
  }
  x[b0 + lane] = xi;
}

// x(r) -= sum_k A(r, b0+k) x(b0+k) for the rows r in [r0, r1): the rest of the
// vector after the diagonal block b0 is solved
__global__ __launch_bounds__(KS_BLOCK) void k_lu_update(const double *__restrict__ A, int ldn, int b0, int r0, int r1,
                                                       double *__restrict__ x) {
  __shared__ double xb[KS_LU_NB];
  if (threadIdx.x < KS_LU_NB) xb[threadIdx.x] = x[b0 + threadIdx.x];
  __syncthreads();
  const int r = r0 + blockIdx.x * KS_BLOCK + threadIdx.x;
  if (r >= r1) return;
  double s = x[r];
#pragma unroll 8
  for (int k = 0; k < KS_LU_NB; k++) s -= A[(size_t)r + (size_t)(b0 + k) * ldn] * xb[k];
  x[r] = s;
}

}  // namespace

struct ks_lu_s {
  int n = 0, ldn = 0;
  double *d_A = nullptr, *d_x = nullptr;
  int *d_ipiv = nullptr, *d_perm = nullptr, *d_flag = nullptr;
};

extern "C" int ks_lu_destroy(ks_lu F) {
  if (!F) return KS_OK;
  if (F->d_A) (void)ks_free(F->d_A);
  if (F->d_x) (void)ks_free(F->d_x);
  if (F->d_ipiv) (void)ks_free(F->d_ipiv);
  if (F->d_perm) (void)ks_free(F->d_perm);
  if (F->d_flag) (void)ks_free(F->d_flag);
  delete F;
  return KS_OK;
}

extern "C" int ks_lu_create(ks_ctx ctx, ks_csr A, double sigma, ks_lu *out) {
  KS_CHECK(ctx && A && out, KS_ERR_ARG, "null LU argument");
  KS_CHECK(A->vprec == KS_F64 && A->nrows == A->ncols, KS_ERR_ARG, "the LU needs a square FP64 matrix");
  KS_CHECK(A->nrows <= KS_LU_MAXN, KS_ERR_RANGE, "the dense LU takes at most %d rows (%d given)", KS_LU_MAXN,
           A->nrows);
  *out = nullptr;
  ks_lu F = new ks_lu_s();
  const int n = A->nrows, ldn = (n + KS_LU_NB - 1) / KS_LU_NB * KS_LU_NB;
  F->n = n;
  F->ldn = ldn;
  int e = ks_malloc((void **)&F->d_A, sizeof(double) * (size_t)ldn * ldn);
  if (!e) e = ks_malloc((void **)&F->d_x, sizeof(double) * ldn);
  if (!e) e = ks_malloc((void **)&F->d_ipiv, sizeof(int) * ldn);
  if (!e) e = ks_malloc((void **)&F->d_perm, sizeof(int) * ldn);
  if (!e) e = ks_malloc((void **)&F->d_flag, sizeof(int));
  if (!e) e = ks_memset_zero(ctx, F->d_A, sizeof(double) * (size_t)ldn * ldn);
  if (!e) e = ks_memset_zero(ctx, F->d_flag, sizeof(int));
  if (e) {
    ks_lu_destroy(F);
    return e;
  }
  hipStream_t st = ctx->stream;
  const int fill_blocks = (ldn + KS_BLOCK - 1) / KS_BLOCK;
  hipLaunchKernelGGL(k_lu_fill, dim3(fill_blocks), dim3(KS_BLOCK), 0, st, A->d_rowptr, A->d_colidx,
                     (const double *)A->d_val, n, ldn, sigma, F->d_A);
  for (int j0 = 0; j0 < ldn; j0 += KS_LU_NB) {
    for (int j = j0; j < j0 + KS_LU_NB; j++) {
      hipLaunchKernelGGL(k_lu_pivot, dim3(1), dim3(KS_LU_PIV_NT), 0, st, F->d_A, ldn, j, j0, F->d_ipiv, F->d_flag);
      const int rows = ldn - j - 1;
      if (rows > 0)
        hipLaunchKernelGGL(k_lu_panel_col, dim3((rows + KS_BLOCK - 1) / KS_BLOCK), dim3(KS_BLOCK), 0, st, F->d_A, ldn,
                           j, j0);
    }
    const int other = ldn - KS_LU_NB;
    if (other > 0)
      hipLaunchKernelGGL(k_lu_laswp, dim3((other + KS_BLOCK - 1) / KS_BLOCK), dim3(KS_BLOCK), 0, st, F->d_A, ldn, j0,
                         (const int *)F->d_ipiv);
    const int rest = ldn - j0 - KS_LU_NB;
    if (rest > 0) {
      const int tb = (rest + KS_WAVES - 1) / KS_WAVES;
      hipLaunchKernelGGL(k_lu_trsm, dim3(tb < 1024 ? tb : 1024), dim3(KS_BLOCK), 0, st, F->d_A, ldn, j0);
      hipLaunchKernelGGL(k_lu_gemm, dim3(rest / KS_LU_NB, rest / KS_LU_NB), dim3(KS_BLOCK), 0, st, F->d_A, ldn, j0);
    }
  }
  KS_HIP(hipGetLastError());
  // the row permutation P (P A = L U): the interchanges applied in order to 0..ldn-1
  std::vector<int> ipiv(ldn), perm(ldn);
  int flag = 0;
  KS_HIP(hipMemcpyAsync(ipiv.data(), F->d_ipiv, sizeof(int) * ldn, hipMemcpyDeviceToHost, st));
  KS_HIP(hipMemcpyAsync(&flag, F->d_flag, sizeof(int), hipMemcpyDeviceToHost, st));
  KS_HIP(hipStreamSynchronize(st));
  if (flag) {
    fprintf(stderr, "ks_hip: the LU of A - sigma I met a zero pivot\n");
    ks_lu_destroy(F);
    return KS_ERR_PIVOT;
  }
  for (int i = 0; i < ldn; i++) perm[i] = i;
  for (int j = 0; j < ldn; j++) {
    const int p = ipiv[j];
    if (p != j) {
      const int t = perm[j];
      perm[j] = perm[p];
      perm[p] = t;
    }
  }
  for (int i = 0; i < n; i++)
    if (perm[i] >= n) {   // a padding row in the real part: only with a zero pivot, caught above
      ks_lu_destroy(F);
      return KS_ERR_PIVOT;
    }
  KS_HIP(hipMemcpyAsync(F->d_perm, perm.data(), sizeof(int) * ldn, hipMemcpyHostToDevice, st));
  KS_HIP(hipStreamSynchronize(st));
  *out = F;
  return KS_OK;
}

extern "C" int ks_lu_solve(ks_ctx ctx, ks_lu F, const double *d_b, double *d_x) {
  KS_CHECK(ctx && F && d_b && d_x, KS_ERR_ARG, "null LU solve argument");
  const int n = F->n, ldn = F->ldn;
  hipStream_t st = ctx->stream;
  double *x = F->d_x;
  hipLaunchKernelGGL(k_lu_permute, dim3((ldn + KS_BLOCK - 1) / KS_BLOCK), dim3(KS_BLOCK), 0, st,
                     (const int *)F->d_perm, d_b, n, ldn, x);
  for (int b0 = 0; b0 < ldn; b0 += KS_LU_NB) {   // L y = P b
    hipLaunchKernelGGL(k_lu_diag<true>, dim3(1), dim3(64), 0, st, (const double *)F->d_A, ldn, b0, x);
    const int r0 = b0 + KS_LU_NB, rows = ldn - r0;
    if (rows > 0)
      hipLaunchKernelGGL(k_lu_update, dim3((rows + KS_BLOCK - 1) / KS_BLOCK), dim3(KS_BLOCK), 0, st,
                         (const double *)F->d_A, ldn, b0, r0, ldn, x);
  }
  for (int b0 = ldn - KS_LU_NB; b0 >= 0; b0 -= KS_LU_NB) {   // U x = y
    hipLaunchKernelGGL(k_lu_diag<false>, dim3(1), dim3(64), 0, st, (const double *)F->d_A, ldn, b0, x);
    if (b0 > 0)
      hipLaunchKernelGGL(k_lu_update, dim3((b0 + KS_BLOCK - 1) / KS_BLOCK), dim3(KS_BLOCK), 0, st,
                         (const double *)F->d_A, ldn, b0, 0, b0, x);
  }
  KS_HIP(hipMemcpyAsync(d_x, x, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  KS_HIP(hipGetLastError());
  return KS_OK;
}
