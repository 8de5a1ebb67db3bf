// ks_lu.hip -- the direct inner solve of shift-and-invert: LU factorization with
// partial pivoting of A - sigma I held dense in HBM, and the triangular solves.
//
// The reference's default inner solver for STSINVERT is KSPPREONLY with PCLU
// (src/sys/classes/st/interface/stsles.c:47-54, sinvert.c:131-154), a sparse LU.
// Here the shifted matrix is expanded to a dense ldn x ldn array (ldn = n
// rounded up to 64, the padding an identity block, so every tile is whole) and
// factored by the right-looking blocked algorithm of LAPACK's dgetrf: panels
// of 512 columns, each factored as sub-panels of 64 (dgetrf2's recursion cut
// at 64 columns), the columns right of a panel updated once per panel:
//   * sub-panel: one launch per column.  The pivot rows are found and used in
//     place (implicit pivoting): every workgroup reduces the partial maxima of
//     column j the previous launch left (largest magnitude, lowest row on
//     ties), scales its rows of column j by the pivot's reciprocal, updates
//     their sub-panel columns right of j (dgetf2's dscal and dger) and leaves
//     its partial maximum of column j+1.  Rows already chosen as pivots are
//     skipped.  At the sub-panel's end one thread turns the pivot rows into
//     LAPACK's sequence of interchanges, which are then applied to every
//     column (dlaswp) -- the same rows and values dgetf2 arrives at by
//     swapping as it goes;
//   * U12 = L11^-1 A12 per 64-row block, one wavefront per column with L11 in
//     LDS (dtrsm);
//   * C -= L U on v_mfma_f64_16x16x4f64 in 64 x 64 tiles (dgemm), k staged
//     through LDS 32 at a time with the next slice loaded into registers
//     during the MFMAs, the operands ordered so that 16 lanes hold 16
//     consecutive rows of a column of C (128-byte reads and writes).
// The solve applies the row permutation as a gather, then the unit-lower and
// upper triangular solves block by block, one launch per 64-row block:
// wavefront 0 of every workgroup solves the diagonal block from registers
// while the others load their rows of the off-diagonal block, then they update
// those rows; the solved block goes to a second vector, so no workgroup reads
// what another writes in the launch.
// Memory is n^2 doubles (for n = 32,768: 8.6 GB), so the factorization serves
// matrices of up to that size; larger ones take an iterative inner solver.
#include <vector>

#include "ks_internal.hpp"

namespace {

#define KS_LU_NB 64
#define KS_LU_NB2 512       // columns of a panel of the trailing update (a multiple of 64)
#define KS_LU_PG_MAX 512   // most workgroups of a panel step (ldn <= 131,072)
// a pivot below this magnitude is a zero pivot (PETSc's default zeropivot, 100 eps)
#define KS_LU_ZEROPIVOT (100.0 * 2.220446049250313e-16)

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

// magnitude for the pivot search: |a|; a NaN row is still a candidate but any
// number wins over it
__device__ inline double piv_mag(double a) { return a == a ? fabs(a) : -0.5; }

__device__ inline bool piv_better(double m1, int i1, double m2, int i2) {
  return m1 > m2 || (m1 == m2 && i1 < i2);
}

// workgroup reduction of (magnitude, row) to thread 0; sm/si hold KS_WAVES entries
__device__ inline void piv_reduce(double &m, int &i, double *sm, int *si) {
  for (int off = 32; off > 0; off >>= 1) {
    const double om = __shfl_down(m, off, 64);
    const int oi = __shfl_down(i, off, 64);
    if (piv_better(om, oi, m, i)) {
      m = om;
      i = oi;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    sm[wave] = m;
    si[wave] = i;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < KS_WAVES; w++)
      if (piv_better(sm[w], si[w], m, i)) {
        m = sm[w];
        i = si[w];
      }
}

// Rows of a panel step: thread t of workgroup g owns row j0 + g * KS_BLOCK + t (the
// grid has G = ceil((ldn - j0) / KS_BLOCK) workgroups), the same row for every
// column of the panel.
// used[r] == j0 + 1: row r is the pivot of an earlier column of this panel.
// pm/pi[slot][g]: workgroup g's largest magnitude (and its row) of a column.

// the partial maxima of the panel's first column j0 (slot 0)
__global__ __launch_bounds__(KS_BLOCK) void k_lu_colmax(const double *__restrict__ A, int ldn, int j0,
                                                       double *__restrict__ pm, int *__restrict__ pi) {
  __shared__ double sm[KS_WAVES];
  __shared__ int si[KS_WAVES];
  double m = -1.0;
  int idx = INT_MAX;
  for (int r = j0 + blockIdx.x * KS_BLOCK + threadIdx.x; r < ldn; r += gridDim.x * KS_BLOCK) {
    const double a = piv_mag(A[(size_t)r + (size_t)j0 * ldn]);
    if (piv_better(a, r, m, idx)) {
      m = a;
      idx = r;
    }
  }
  piv_reduce(m, idx, sm, si);
  if (threadIdx.x == 0) {
    pm[blockIdx.x] = m;
    pi[blockIdx.x] = idx;
  }
}

// column j of the panel j0: the pivot row p from the partial maxima (slot j & 1),
// l_rj = a_rj / a_pj for the other candidate rows r, a_rc -= l_rj a_pc for the
// panel columns c > j, and the partial maxima of column j+1 (slot (j + 1) & 1).
// A pivot below KS_LU_ZEROPIVOT sets *flag.
__global__ __launch_bounds__(KS_BLOCK) void k_lu_step(double *__restrict__ A, int ldn, int j0, int j,
                                                     int *__restrict__ used, double *__restrict__ pm,
                                                     int *__restrict__ pi, int *__restrict__ prow, int *flag) {
  __shared__ double prv[KS_LU_NB];
  __shared__ double sm[KS_WAVES];
  __shared__ int si[KS_WAVES];
  __shared__ int sp;
  const int tid = threadIdx.x, G = gridDim.x, slot = j & 1;
  const int jend = j0 + KS_LU_NB, nc = jend - j, stamp = j0 + 1;
  // this thread's row (the grid covers rows j0 .. ldn-1 one row a thread): its flag
  // and its nc values (columns j .. jend-1) loaded in the same round trip as the
  // partial maxima
  const int r = j0 + blockIdx.x * KS_BLOCK + tid;
  const bool have = r < ldn;
  int u = stamp;
  double v[KS_LU_NB];
  if (have) {
    u = used[r];
#pragma unroll
    for (int c = 0; c < KS_LU_NB; c++)
      if (c < nc) v[c] = A[(size_t)r + (size_t)(j + c) * ldn];
  }
  double m = -1.0;
  int idx = INT_MAX;
  for (int g = tid; g < G; g += KS_BLOCK) {
    const double a = pm[slot * KS_LU_PG_MAX + g];
    const int i = pi[slot * KS_LU_PG_MAX + g];
    if (piv_better(a, i, m, idx)) {
      m = a;
      idx = i;
    }
  }
  piv_reduce(m, idx, sm, si);
  if (tid == 0) {
    sp = idx;
    if (blockIdx.x == 0) {
      prow[j - j0] = idx;
      if (!(m >= KS_LU_ZEROPIVOT)) *flag = 1;
    }
  }
  __syncthreads();
  const int p = sp;   // a candidate row always exists: rows j..ldn-1 outnumber the panel's pivots
  if (tid < nc) prv[tid] = A[(size_t)p + (size_t)(j + tid) * ldn];
  __syncthreads();
  const double rinv = 1.0 / prv[0];
  double m2 = -1.0;
  int i2 = INT_MAX;
  if (have && u != stamp) {
    if (r == p) {
      used[r] = stamp;
    } else {
      double *row = A + r;
      const double l = v[0] * rinv;
      row[(size_t)j * ldn] = l;
#pragma unroll
      for (int c = 1; c < KS_LU_NB; c++)
        if (c < nc) row[(size_t)(j + c) * ldn] = v[c] - l * prv[c];
      if (nc > 1) {
        m2 = piv_mag(v[1] - l * prv[1]);
        i2 = r;
      }
    }
  }
  if (j + 1 < jend) {
    piv_reduce(m2, i2, sm, si);
    if (tid == 0) {
      pm[(slot ^ 1) * KS_LU_PG_MAX + blockIdx.x] = m2;
      pi[(slot ^ 1) * KS_LU_PG_MAX + blockIdx.x] = i2;
    }
  }
}

// the pivot rows of the panel (physical rows, chosen in place) as LAPACK's
// interchanges: at step j rows j and ipiv[j] swap.  Only rows that started in
// the panel's own 64 positions can sit elsewhere when picked: lpos follows
// where those rows are, at which of the 64 positions holds which row.
__global__ __launch_bounds__(64) void k_lu_ipiv(const int *__restrict__ prow, int j0, int *__restrict__ ipiv) {
  __shared__ int lpos[KS_LU_NB], at[KS_LU_NB];
  lpos[threadIdx.x] = j0 + threadIdx.x;
  at[threadIdx.x] = j0 + threadIdx.x;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < KS_LU_NB; k++) {
      const int j = j0 + k, p = prow[k];
      const bool pin = p - j0 < KS_LU_NB;
      const int P = pin ? lpos[p - j0] : p;   // where row p stands now (>= j)
      ipiv[j] = P;
      const int x = at[k];                    // the row standing at j, one of the panel's own
      lpos[x - j0] = P;
      if (pin) lpos[p - j0] = j;
      at[k] = p;
      if (P - j0 < KS_LU_NB) at[P - j0] = x;
    }
  }
}

// the panel's interchanges (rows j0..j0+NB-1 with ipiv, in order) on the columns
// cb .. cb+nc-1, a thread per column
__global__ __launch_bounds__(KS_BLOCK) void k_lu_laswp(double *__restrict__ A, int ldn, int j0,
                                                      const int *__restrict__ ipiv, int cb, int nc) {
  const int t = blockIdx.x * KS_BLOCK + threadIdx.x;
  if (t >= nc) return;
  double *a = A + (size_t)(cb + t) * ldn;
  for (int j = j0; j < j0 + KS_LU_NB; j++) {
    const int p = ipiv[j];
    if (p != j) {
      const double x = a[j];
      a[j] = a[p];
      a[p] = x;
    }
  }
}

__device__ inline double readlane_f64(double v, int k) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)bits, k), hi = __builtin_amdgcn_readlane((int)(bits >> 32), k);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// U12 = L11^-1 A12 for the columns cb .. cb+ncols-1: a wavefront per column,
// lane i holding row j0+i; L11 (the unit lower block at (j0, j0)) staged in LDS
__global__ __launch_bounds__(KS_BLOCK) void k_lu_trsm(double *__restrict__ A, int ldn, int j0, int cb, int ncols) {
  __shared__ double L[KS_LU_NB][KS_LU_NB + 1];   // L[k][i] = A(j0+i, j0+k)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int e = tid; e < KS_LU_NB * KS_LU_NB; e += KS_BLOCK) {
    const int i = e % KS_LU_NB, k = e / KS_LU_NB;
    L[k][i] = A[(size_t)(j0 + i) + (size_t)(j0 + k) * ldn];
  }
  __syncthreads();
  for (int cc = blockIdx.x * KS_WAVES + wave; cc < ncols; cc += gridDim.x * KS_WAVES) {
    double *a = A + (size_t)(cb + cc) * ldn + j0;
    double x = a[lane];
    for (int k = 0; k < KS_LU_NB - 1; k++) {
      const double xk = readlane_f64(x, k);
      if (lane > k) x -= L[k][lane] * xk;
    }
    a[lane] = x;
  }
}

typedef double v4d __attribute__((ext_vector_type(4)));

// C -= L U for the 64 x 64 tile (rb + 64 bx, cb + 64 by) of A, with L the
// columns kb .. kb+K-1 of the tile's rows and U the rows kb .. kb+K-1 of its
// columns (the trailing update A22 -= L21 U12 of the factorization).  L and U
// are staged in LDS 32 along k at a time (33 KB, several workgroups a CU), each
// wavefront a 32 x 32 quarter (2 x 2 MFMA tiles).  The MFMA's first operand is
// U (its 16 rows the tile's columns), the second L (its 16 columns the tile's
// rows), so register q of lane l holds column l/16 + 4q, row l%16 of a 16 x 16
// tile: 16 lanes read and write 128 contiguous bytes.  K is a multiple of 32.
#define KS_LU_KH 32
__global__ __launch_bounds__(KS_BLOCK) void k_lu_gemm(double *__restrict__ A, int ldn, int kb, int K, int rb,
                                                     int cb) {
  __shared__ double sL[KS_LU_KH][KS_LU_NB + 1];   // sL[k][r] = L(r, h + k)
  __shared__ double sU[KS_LU_KH][KS_LU_NB + 1];   // sU[k][c] = U(h + k, c)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r0 = rb + blockIdx.x * KS_LU_NB, c0 = cb + blockIdx.y * KS_LU_NB;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  // this lane's 16 entries of C, read before the operands are staged
  double cv[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int q = 0; q < 4; q++)
        cv[a][b][q] = A[(size_t)(r0 + wr + 16 * b + lr) + (size_t)(c0 + wc + 16 * a + lk + 4 * q) * ldn];
  v4d acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
  // the operands of a 32-wide k slice: thread e-slots tid, tid + 256, ... (8 each);
  // the next slice is loaded into registers while the MFMAs of the current one run
  constexpr int PER = KS_LU_NB * KS_LU_KH / KS_BLOCK;
  double pl[PER], pu[PER];
  auto load = [&](int h) {
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int e = tid + u * KS_BLOCK;
      const int i = e % KS_LU_NB, k = e / KS_LU_NB;     // L: 64 rows of a column
      pl[u] = A[(size_t)(r0 + i) + (size_t)(kb + h + k) * ldn];
      const int kk = e % KS_LU_KH, c = e / KS_LU_KH;   // U: 32 rows of a column
      pu[u] = A[(size_t)(kb + h + kk) + (size_t)(c0 + c) * ldn];
    }
  };
  load(0);
  for (int h = 0; h < K; h += KS_LU_KH) {
    if (h) __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int e = tid + u * KS_BLOCK;
      sL[e / KS_LU_NB][e % KS_LU_NB] = pl[u];
      sU[e % KS_LU_KH][e / KS_LU_KH] = pu[u];
    }
    __syncthreads();
    if (h + KS_LU_KH < K) load(h + KS_LU_KH);
#pragma unroll
    for (int kk = 0; kk < KS_LU_KH / 4; kk++) {
      const int k = 4 * kk + lk;
      const double u0 = sU[k][wc + lr], u1 = sU[k][wc + 16 + lr];
      const double l0 = sL[k][wr + lr], l1 = sL[k][wr + 16 + lr];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(u0, l0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(u0, l1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(u1, l0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(u1, l1, acc[1][1], 0, 0, 0);
    }
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int q = 0; q < 4; q++)
        A[(size_t)(r0 + wr + 16 * b + lr) + (size_t)(c0 + wc + 16 * a + lk + 4 * q) * ldn] =
            cv[a][b][q] - acc[a][b][q];
}

// x(i) = b(perm(i)) for i < n, 0 in the padding
__global__ __launch_bounds__(KS_BLOCK) void k_lu_permute(const int *__restrict__ perm, const double *__restrict__ b,
                                                        int n, int ldn, double *__restrict__ x) {
  for (int i = blockIdx.x * KS_BLOCK + threadIdx.x; i < ldn; i += gridDim.x * KS_BLOCK)
    x[i] = i < n ? b[perm[i]] : 0.0;
}

#define KS_LU_TRI_ROWS (KS_BLOCK - 64)   // rows a workgroup updates (wavefronts 1..3)

// One block step of a triangular solve.  Wavefront 0 solves the 64 x 64
// diagonal block at b0 of L (unit lower, LOWER) or U from xin(b0..b0+63), lane
// i holding row i of the block in registers and the solved entries passed on
// by readlane (column-oriented substitution, the order of LAPACK's dtrsv);
// workgroup 0 writes the block to out(b0..) (its rows below n_out).  The other
// wavefronts load their rows of the rest, r >= b0 + 64 (LOWER) or r < b0,
// meanwhile and then take xin(r) -= A(r, b0:b0+64) x_block (in the order of k
// that dtrsv follows).  xin's block rows are read and never written in the
// launch, and out is not xin.
template <bool LOWER>
__global__ __launch_bounds__(KS_BLOCK) void k_lu_tri_step(const double *__restrict__ A, int ldn, int b0, int n_out,
                                                         double *__restrict__ xin, double *__restrict__ out) {
  __shared__ double xb[KS_LU_NB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r0 = LOWER ? b0 + KS_LU_NB : 0, r1 = LOWER ? ldn : b0;
  const int r = r0 + blockIdx.x * KS_LU_TRI_ROWS + tid - 64;   // this thread's row (wavefronts 1..3)
  double arow[KS_LU_NB], s = 0.0;
  if (wave == 0) {
    double d[KS_LU_NB];
#pragma unroll
    for (int k = 0; k < KS_LU_NB; k++) d[k] = A[(size_t)(b0 + lane) + (size_t)(b0 + k) * ldn];
    double xi = xin[b0 + lane];
    if (LOWER) {
#pragma unroll
      for (int k = 0; k < KS_LU_NB - 1; k++) {
        const double xk = readlane_f64(xi, k);
        if (lane > k) xi -= d[k] * xk;
      }
    } else {
#pragma unroll
      for (int k = KS_LU_NB - 1; k >= 0; k--) {
        if (lane == k) xi /= d[k];
        const double xk = readlane_f64(xi, k);
        if (lane < k) xi -= d[k] * xk;
      }
    }
    xb[lane] = xi;
    if (blockIdx.x == 0 && b0 + lane < n_out) out[b0 + lane] = xi;
  } else if (r < r1) {
    s = xin[r];
#pragma unroll
    for (int k = 0; k < KS_LU_NB; k++) arow[k] = A[(size_t)r + (size_t)(b0 + k) * ldn];
  }
  __syncthreads();
  if (wave == 0) return;
  if (r < r1) {
    if (LOWER) {
#pragma unroll
      for (int k = 0; k < KS_LU_NB; k++) s -= arow[k] * xb[k];
    } else {
#pragma unroll
      for (int k = KS_LU_NB - 1; k >= 0; k--) s -= arow[k] * xb[k];
    }
    xin[r] = s;
  }
  for (int rr = r + gridDim.x * KS_LU_TRI_ROWS; rr < r1; rr += gridDim.x * KS_LU_TRI_ROWS) {
    double t = xin[rr];
    if (LOWER) {
      for (int k = 0; k < KS_LU_NB; k++) t -= A[(size_t)rr + (size_t)(b0 + k) * ldn] * xb[k];
    } else {
      for (int k = KS_LU_NB - 1; k >= 0; k--) t -= A[(size_t)rr + (size_t)(b0 + k) * ldn] * xb[k];
    }
    xin[rr] = t;
  }
}

}  // namespace

struct ks_lu_s {
  int n = 0, ldn = 0;
  double *d_A = nullptr, *d_x = nullptr, *d_y = nullptr;
  int *d_ipiv = nullptr, *d_perm = nullptr;
};

extern "C" int ks_lu_destroy(ks_lu F) {
  if (!F) return KS_OK;
  if (F->d_A) (void)ks_free(F->d_A);
  if (F->d_x) (void)ks_free(F->d_x);
  if (F->d_y) (void)ks_free(F->d_y);
  if (F->d_ipiv) (void)ks_free(F->d_ipiv);
  if (F->d_perm) (void)ks_free(F->d_perm);
  delete F;
  return KS_OK;
}

extern "C" int ks_lu_create(ks_ctx ctx, ks_csr A, double sigma, ks_lu *out) {
  KS_CHECK(ctx && A && out, KS_ERR_ARG, "null LU argument");
  KS_CHECK(A->vprec == KS_F64 && A->nrows == A->ncols, KS_ERR_ARG, "the LU needs a square FP64 matrix");
  KS_CHECK(A->nrows <= KS_LU_MAXN, KS_ERR_RANGE, "the dense LU takes at most %d rows (%d given)", KS_LU_MAXN,
           A->nrows);
  static_assert((KS_LU_MAXN + KS_LU_NB - 1) / KS_LU_NB * KS_LU_NB <= KS_LU_PG_MAX * KS_BLOCK,
                "a panel step's workgroups must cover every row");
  *out = nullptr;
  ks_lu F = new ks_lu_s();
  const int n = A->nrows, ldn = (n + KS_LU_NB - 1) / KS_LU_NB * KS_LU_NB;
  F->n = n;
  F->ldn = ldn;
  // factorization scratch: used[ldn], prow[64], flag, pi[2][PG_MAX] (ints), pm[2][PG_MAX] (doubles)
  const size_t nint = (size_t)ldn + KS_LU_NB + 1 + 2 * KS_LU_PG_MAX;
  const size_t scratch_bytes = ((nint * sizeof(int) + 255) / 256) * 256 + 2 * KS_LU_PG_MAX * sizeof(double);
  void *scratch = nullptr;
  int e = ks_malloc((void **)&F->d_A, sizeof(double) * (size_t)ldn * ldn);
  if (!e) e = ks_malloc((void **)&F->d_x, sizeof(double) * ldn);
  if (!e) e = ks_malloc((void **)&F->d_y, sizeof(double) * ldn);
  if (!e) e = ks_malloc((void **)&F->d_ipiv, sizeof(int) * ldn);
  if (!e) e = ks_malloc((void **)&F->d_perm, sizeof(int) * ldn);
  if (!e) e = ks_malloc(&scratch, scratch_bytes);
  if (!e) e = ks_memset_zero(ctx, F->d_A, sizeof(double) * (size_t)ldn * ldn);
  if (!e) e = ks_memset_zero(ctx, scratch, scratch_bytes);
  if (e) {
    if (scratch) (void)ks_free(scratch);
    ks_lu_destroy(F);
    return e;
  }
  int *d_used = (int *)scratch, *d_prow = d_used + ldn, *d_flag = d_prow + KS_LU_NB, *d_pi = d_flag + 1;
  double *d_pm = (double *)((char *)scratch + ((nint * sizeof(int) + 255) / 256) * 256);
  hipStream_t st = ctx->stream;
  const int fill_blocks = (ldn + KS_BLOCK - 1) / KS_BLOCK;
  hipLaunchKernelGGL(k_lu_fill, dim3(fill_blocks), dim3(KS_BLOCK), 0, st, A->d_rowptr, A->d_colidx,
                     (const double *)A->d_val, n, ldn, sigma, F->d_A);
  // panels of KS_LU_NB2 columns, each factored as sub-panels of KS_LU_NB (dgetrf2's
  // recursion cut at 64): a sub-panel is first brought up to date with the earlier
  // sub-panels of its panel, and the columns right of the panel take the whole
  // panel's update at once (K = KS_LU_NB2), so A22 is read and written once per
  // KS_LU_NB2 columns
  auto trsm = [&](int j0, int cb, int ncols) {
    const int tb = (ncols + KS_WAVES - 1) / KS_WAVES;
    hipLaunchKernelGGL(k_lu_trsm, dim3(tb < 1024 ? tb : 1024), dim3(KS_BLOCK), 0, st, F->d_A, ldn, j0, cb, ncols);
  };
  auto gemm = [&](int kb, int K, int rb, int nrows, int cb, int ncols) {
    if (nrows > 0 && ncols > 0)
      hipLaunchKernelGGL(k_lu_gemm, dim3(nrows / KS_LU_NB, ncols / KS_LU_NB), dim3(KS_BLOCK), 0, st, F->d_A, ldn, kb,
                         K, rb, cb);
  };
  for (int J0 = 0; J0 < ldn; J0 += KS_LU_NB2) {
    const int W = ldn - J0 < KS_LU_NB2 ? ldn - J0 : KS_LU_NB2;
    for (int s = 0; s < W; s += KS_LU_NB) {
      const int j0 = J0 + s;
      if (s > 0) {   // the sub-panel's rows J0 .. j0-1 become U, its rows below take the update
        for (int t = 0; t < s; t += KS_LU_NB) {
          trsm(J0 + t, j0, KS_LU_NB);
          gemm(J0 + t, KS_LU_NB, J0 + t + KS_LU_NB, s - t - KS_LU_NB, j0, KS_LU_NB);
        }
        gemm(J0, s, j0, ldn - j0, j0, KS_LU_NB);
      }
      const int G = (ldn - j0 + KS_BLOCK - 1) / KS_BLOCK;
      hipLaunchKernelGGL(k_lu_colmax, dim3(G), dim3(KS_BLOCK), 0, st, (const double *)F->d_A, ldn, j0, d_pm, d_pi);
      for (int j = j0; j < j0 + KS_LU_NB; j++)
        hipLaunchKernelGGL(k_lu_step, dim3(G), dim3(KS_BLOCK), 0, st, F->d_A, ldn, j0, j, d_used, d_pm, d_pi, d_prow,
                           d_flag);
      hipLaunchKernelGGL(k_lu_ipiv, dim3(1), dim3(KS_LU_NB), 0, st, (const int *)d_prow, j0, F->d_ipiv);
      hipLaunchKernelGGL(k_lu_laswp, dim3(fill_blocks), dim3(KS_BLOCK), 0, st, F->d_A, ldn, j0,
                         (const int *)F->d_ipiv, 0, ldn);
    }
    const int cb = J0 + W, nc = ldn - cb;
    if (nc > 0) {   // U12 of the panel's rows, then A22 -= L21 U12 with K = W
      for (int t = 0; t < W; t += KS_LU_NB) {
        trsm(J0 + t, cb, nc);
        gemm(J0 + t, KS_LU_NB, J0 + t + KS_LU_NB, W - t - KS_LU_NB, cb, nc);
      }
      gemm(J0, W, cb, ldn - cb, cb, nc);
    }
  }
  hipError_t herr = hipSuccess;
  if (herr == hipSuccess) herr = hipGetLastError();
  // the row permutation P (P A = L U): the interchanges applied in order to 0..ldn-1
  std::vector<int> ipiv(ldn), perm(ldn);
  int flag = 0;
  if (herr == hipSuccess) herr = hipMemcpyAsync(ipiv.data(), F->d_ipiv, sizeof(int) * ldn, hipMemcpyDeviceToHost, st);
  if (herr == hipSuccess) herr = hipMemcpyAsync(&flag, d_flag, sizeof(int), hipMemcpyDeviceToHost, st);
  if (herr == hipSuccess) herr = hipStreamSynchronize(st);
  else (void)hipStreamSynchronize(st);   // nothing may still run on what is freed below
  (void)ks_free(scratch);
  if (herr != hipSuccess) {
    ks_lu_destroy(F);
    KS_HIP(herr);
  }
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
  herr = hipMemcpyAsync(F->d_perm, perm.data(), sizeof(int) * ldn, hipMemcpyHostToDevice, st);
  if (herr == hipSuccess) herr = hipStreamSynchronize(st);
  if (herr != hipSuccess) {
    ks_lu_destroy(F);
    KS_HIP(herr);
  }
  *out = F;
  return KS_OK;
}

extern "C" int ks_lu_solve(ks_ctx ctx, ks_lu F, const double *d_b, double *d_x) {
  KS_CHECK(ctx && F && d_b && d_x, KS_ERR_ARG, "null LU solve argument");
  const int n = F->n, ldn = F->ldn;
  hipStream_t st = ctx->stream;
  hipLaunchKernelGGL(k_lu_permute, dim3((ldn + KS_BLOCK - 1) / KS_BLOCK), dim3(KS_BLOCK), 0, st,
                     (const int *)F->d_perm, d_b, n, ldn, F->d_x);
  auto grid = [](int rows) {
    const int g = (rows + KS_LU_TRI_ROWS - 1) / KS_LU_TRI_ROWS;
    return g < 1 ? 1 : (g > 512 ? 512 : g);
  };
  for (int b0 = 0; b0 < ldn; b0 += KS_LU_NB)   // L y = P b: F->d_x updated in place, y to F->d_y
    hipLaunchKernelGGL(k_lu_tri_step<true>, dim3(grid(ldn - b0 - KS_LU_NB)), dim3(KS_BLOCK), 0, st,
                       (const double *)F->d_A, ldn, b0, ldn, F->d_x, F->d_y);
  for (int b0 = ldn - KS_LU_NB; b0 >= 0; b0 -= KS_LU_NB)   // U x = y: F->d_y updated in place, x to d_x
    hipLaunchKernelGGL(k_lu_tri_step<false>, dim3(grid(b0)), dim3(KS_BLOCK), 0, st, (const double *)F->d_A, ldn, b0,
                       n, F->d_y, d_x);
  KS_HIP(hipGetLastError());
  return KS_OK;
}
