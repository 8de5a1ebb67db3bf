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
//   * sub-panel: one cooperative launch (k_lu_panel), the 64 column steps
//     exchanging candidates through epoch-tagged words; or, should that launch
//     be refused or a wait run out, one launch per column (k_lu_step) -- the
//     same operations in the same order.  The pivot rows are found and used in
//     place (implicit pivoting): per column the best candidate over all rows
//     (largest magnitude, lowest row on ties) is the pivot, the other rows'
//     entries of column j are scaled by its reciprocal and their sub-panel
//     columns right of j updated (dgetf2's dscal and dger).  Rows already
//     chosen as pivots are skipped.  At the sub-panel's end one thread turns the pivot rows into
//     LAPACK's sequence of interchanges, which are then applied to every
//     column (dlaswp) -- the same rows and values dgetf2 arrives at by
//     swapping as it goes;
//   * U12 = L11^-1 A12 per 64-row block, one wavefront per column with L11 in
//     LDS (dtrsm);
//   * C -= L U on v_mfma_f64_16x16x4f64 in 64 x 64 tiles (dgemm), k staged
//     through LDS 32 at a time with the next slice loaded into registers
//     during the MFMAs, the operands ordered so that 16 lanes hold 16
//     consecutive rows of a column of C (128-byte reads and writes).
// The solve is two launches, the unit-lower and the upper triangular solve,
// each with one workgroup per 64-row block row, pipelined through the published
// block solutions (k_lu_tri_sf): a workgroup accumulates the blocks left
// (right) of its diagonal as their solutions are published, then solves its
// diagonal block and publishes it; the row permutation is applied as a gather
// on the way in.
// Memory is n^2 doubles (for n = 32,768: 8.6 GB), so the factorization serves
// matrices of up to that size; larger ones take an iterative inner solver.
#include <cstdlib>
#include <vector>

#include "ks_internal.hpp"

namespace {

#define KS_LU_NB 64
#define KS_LU_NB2 512       // columns of a panel of the trailing update (a multiple of 64)
#define KS_LU_PG_MAX 512
#define KS_LU_SPIN (1 << 26)   // bound of every wait: polls of about 50 ns, seconds in all   // most workgroups of a panel step (ldn <= 131,072)
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
template <int NW = KS_WAVES>
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
    for (int w = 1; w < NW; w++)
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

// Epoch-tagged publication: a 32-bit payload beside a 32-bit epoch in one 64-bit
// word, stored and polled with agent-scope atomics (coherent across the XCDs'
// caches); a double takes two words.  A reader that finds the current epoch in
// every word of a value has the value, with no fence.
__device__ inline void pub_store(unsigned long long *w, unsigned epoch, unsigned payload) {
  __hip_atomic_store(w, ((unsigned long long)epoch << 32) | payload, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void pub_store_f64(unsigned long long *w, unsigned epoch, double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  pub_store(w, epoch, (unsigned)(b & 0xffffffffull));
  pub_store(w + 1, epoch, (unsigned)(b >> 32));
}
// polls until the nw words at w carry epoch (bounded; sooner out once *err shows
// another wait ran out), then returns their payloads in p[]
template <int NWORD>
__device__ inline void pub_wait(const unsigned long long *w, unsigned epoch, unsigned *p, int *err) {
  int spins = 0;
  for (;;) {
    bool ok = true;
#pragma unroll
    for (int q = 0; q < NWORD; q++) {
      const unsigned long long v = __hip_atomic_load(w + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      p[q] = (unsigned)(v & 0xffffffffull);
      ok = ok && (unsigned)(v >> 32) == epoch;
    }
    if (ok) return;
    if (++spins > KS_LU_SPIN ||
        ((spins & 1023) == 0 && (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 1))) {
      atomicOr(err, 1);
      return;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// The whole sub-panel in one cooperative launch (all G workgroups resident):
// thread t of workgroup g keeps row j0 + g * KS_LU_PT + t of the 64 panel
// columns in LDS for the 64 column steps, which exchange their data through
// epoch-tagged words instead of launch boundaries.  Per column k (epoch
// e0 + k): each workgroup finds its best candidate (largest magnitude, lowest
// row) and publishes it -- its (magnitude, row) as three words, the row's
// values in columns k..63 as two words each, one value a thread -- into slot
// k & 1; every workgroup then waits for the G candidates, takes the best
// (ties to the lower row), waits for the winner's values and updates its rows
// as k_lu_step does (the same operations in the same order: the factors are
// those of the column-step launches).  A workgroup writes slot k & 1 again
// only at column k + 2, after every workgroup has published column k + 1,
// which each does after reading column k.
#define KS_LU_PT 128                                  // rows (threads) a workgroup
#define KS_LU_PW (KS_LU_PT / 64)
#define KS_LU_CO_MAX 256                              // workgroups at most (ldn <= 32,768)
#define KS_LU_CAND 3                                  // words of a candidate: magnitude (2), row
__global__ __launch_bounds__(KS_LU_PT) void k_lu_panel(double *__restrict__ A, int ldn, int j0,
                                                      unsigned long long *cand, unsigned long long *crow,
                                                      unsigned e0, int *__restrict__ prow, int *flag, int *err) {
  __shared__ double rowv[KS_LU_NB][KS_LU_PT];   // 64 KB: the workgroup's rows of the panel
  __shared__ double pv[KS_LU_NB];
  __shared__ double sm[KS_LU_PW];
  __shared__ int si[KS_LU_PW];
  __shared__ int s_best, s_p;
  const int tid = threadIdx.x, G = gridDim.x, g = blockIdx.x;
  const int r = j0 + g * KS_LU_PT + tid;
  const bool have = r < ldn;
#pragma unroll 8
  for (int c = 0; c < KS_LU_NB; c++) rowv[c][tid] = have ? A[(size_t)r + (size_t)(j0 + c) * ldn] : 0.0;
  bool used = !have;
  __syncthreads();
  for (int k = 0; k < KS_LU_NB; k++) {
    const unsigned ep = e0 + (unsigned)k;
    const int slot = k & 1;
    unsigned long long *mycand = cand + ((size_t)slot * KS_LU_CO_MAX + g) * KS_LU_CAND;
    unsigned long long *myrow = crow + ((size_t)slot * KS_LU_CO_MAX + g) * 2 * KS_LU_NB;
    double m = used ? -1.0 : piv_mag(rowv[k][tid]);
    int idx = used ? INT_MAX : r;
    piv_reduce<KS_LU_PW>(m, idx, sm, si);
    if (tid == 0) {
      s_best = idx;
      pub_store_f64(mycand, ep, m);
      pub_store(mycand + 2, ep, (unsigned)idx);
    }
    __syncthreads();
    {   // the candidate's values, columns k..63, one a thread (the row's index -> its thread)
      const int b = s_best;
      const int c = k + tid;
      if (c < KS_LU_NB) pub_store_f64(myrow + 2 * c, ep, b != INT_MAX ? rowv[c][b - (j0 + g * KS_LU_PT)] : 0.0);
    }
    // the best of the G candidates (each workgroup alike)
    double mb = -1.0;
    int ib = INT_MAX;
    for (int q = tid; q < G; q += KS_LU_PT) {
      unsigned w[KS_LU_CAND];
      pub_wait<KS_LU_CAND>(cand + ((size_t)slot * KS_LU_CO_MAX + q) * KS_LU_CAND, ep, w, err);
      const double a = __longlong_as_double((long long)(((unsigned long long)w[1] << 32) | w[0]));
      const int i = (int)w[2];
      if (piv_better(a, i, mb, ib)) {
        mb = a;
        ib = i;
      }
    }
    piv_reduce<KS_LU_PW>(mb, ib, sm, si);
    if (tid == 0) {
      s_p = ib;
      if (g == 0) {
        prow[k] = ib;
        if (!(mb >= KS_LU_ZEROPIVOT)) *flag = 1;
      }
    }
    __syncthreads();
    const int p = s_p;
    const int gw = p == INT_MAX ? g : (p - j0) / KS_LU_PT;   // a candidate row always exists
    if (tid >= k && tid < KS_LU_NB) {
      unsigned w[2];
      pub_wait<2>(crow + ((size_t)slot * KS_LU_CO_MAX + gw) * 2 * KS_LU_NB + 2 * tid, ep, w, err);
      pv[tid] = __longlong_as_double((long long)(((unsigned long long)w[1] << 32) | w[0]));
    }
    __syncthreads();
    if (!used) {
      if (r == p) {
        used = true;
      } else {
        const double l = rowv[k][tid] * (1.0 / pv[k]);
        rowv[k][tid] = l;
        for (int c = k + 1; c < KS_LU_NB; c++) rowv[c][tid] -= l * pv[c];
      }
    }
    __syncthreads();
  }
#pragma unroll 8
  for (int c = 0; c < KS_LU_NB; c++)
    if (have) A[(size_t)r + (size_t)(j0 + c) * ldn] = rowv[c][tid];
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
      const double xk = ks_readlane_f64(x, k);
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

// One launch for a whole triangular solve, block rows pipelined through their
// published solutions.  Workgroup g takes block row i = g (forward, L) or
// i = nblk-1-g (backward, U), so a workgroup only waits for lower-numbered ones
// (dispatched before it).  For each solved block j (j < i forward, j > i
// backward, in the order they are solved) it loads its 64 x 64 block L(i, j) or
// U(i, j) -- issued before the wait -- waits for x_j and accumulates the block
// times x_j: thread t takes row t%64 and 16 of the 64 columns.  Then wavefront 0
// solves the diagonal block from registers (column-oriented, in dtrsv's order
// within the block), writes x_i to xs (forward) or the output (backward, rows
// below n) and publishes it.  A value is published as two 64-bit words, each a
// 32-bit half of the double beside the solve's epoch, written and polled with
// agent-scope atomics (coherent across the XCDs' caches): a reader takes a value
// once both words carry the current epoch, so no fence is needed.  Waits are
// bounded (KS_LU_SPIN polls, or sooner once another wait has run out): then bit
// 0 of *err is set and the workgroup goes on, so every workgroup publishes and
// exits.  Forward: the right-hand side is b gathered through the row
// permutation (rows >= n zero); backward: the forward solve's xs.
template <bool LOWER>
__global__ __launch_bounds__(KS_BLOCK) void k_lu_tri_sf(const double *__restrict__ A, int ldn, int n,
                                                       const int *__restrict__ perm, const double *__restrict__ rhs,
                                                       double *__restrict__ xs, double *__restrict__ out,
                                                       unsigned long long *slot, unsigned epoch, int *err) {
  __shared__ double part[KS_WAVES][KS_LU_NB];
  __shared__ double xj[2][KS_LU_NB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nblk = ldn / KS_LU_NB;
  const int i = LOWER ? (int)blockIdx.x : nblk - 1 - (int)blockIdx.x;
  const int r0 = i * KS_LU_NB, row = r0 + lane, cq = wave * 16;   // this thread's row and 16 columns
  double acc = 0.0;
  const int jn = LOWER ? i : nblk - 1 - i;   // blocks to take, in solve order
  // wavefront 0's diagonal block, loaded before the waits
  double d[KS_LU_NB];
  if (wave == 0) {
#pragma unroll
    for (int k = 0; k < KS_LU_NB; k++) d[k] = A[(size_t)row + (size_t)(r0 + k) * ldn];
  }
  double a[16];
  auto load_block = [&](int j) {
#pragma unroll
    for (int c = 0; c < 16; c++) a[c] = A[(size_t)row + (size_t)(j * KS_LU_NB + cq + c) * ldn];
  };
  if (jn > 0) load_block(LOWER ? 0 : nblk - 1);
  for (int t = 0; t < jn; t++) {
    const int j = LOWER ? t : nblk - 1 - t, buf = t & 1;
    if (tid < KS_LU_NB) {   // x_j, element by element, as it is published
      const unsigned long long *sl = slot + 2 * (size_t)(j * KS_LU_NB + tid);
      unsigned long long lo, hi;
      int spins = 0;
      for (;;) {
        lo = __hip_atomic_load(sl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        hi = __hip_atomic_load(sl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((lo >> 32) == epoch && (hi >> 32) == epoch) break;
        if (++spins > KS_LU_SPIN || ((spins & 1023) == 0 && (__hip_atomic_load(err, __ATOMIC_RELAXED,
                                                                                  __HIP_MEMORY_SCOPE_AGENT) & 1))) {
          atomicOr(err, 1);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      xj[buf][tid] = __longlong_as_double((long long)(((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull)));
    }
    __syncthreads();   // (also: every thread is past the use of this buffer two blocks back)
    double cur[16];
#pragma unroll
    for (int c = 0; c < 16; c++) cur[c] = a[c];
    if (t + 1 < jn) load_block(LOWER ? t + 1 : nblk - 2 - t);   // the next block in flight meanwhile
    if (LOWER) {
#pragma unroll
      for (int c = 0; c < 16; c++) acc += cur[c] * xj[buf][cq + c];
    } else {
#pragma unroll
      for (int c = 15; c >= 0; c--) acc += cur[c] * xj[buf][cq + c];
    }
  }
  part[wave][lane] = acc;
  __syncthreads();
  if (wave == 0) {
    const double b = LOWER ? (row < n ? rhs[perm[row]] : 0.0) : rhs[row];
    const double sum = LOWER ? ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane]
                             : ((part[3][lane] + part[2][lane]) + part[1][lane]) + part[0][lane];
    double xi = b - sum;
    if (LOWER) {
#pragma unroll
      for (int k = 0; k < KS_LU_NB - 1; k++) {
        const double xk = ks_readlane_f64(xi, k);
        if (lane > k) xi -= d[k] * xk;
      }
    } else {
#pragma unroll
      for (int k = KS_LU_NB - 1; k >= 0; k--) {
        if (lane == k) xi /= d[k];
        const double xk = ks_readlane_f64(xi, k);
        if (lane < k) xi -= d[k] * xk;
      }
    }
    if (LOWER) xs[row] = xi;
    else if (row < n) out[row] = xi;
    const unsigned long long tag = (unsigned long long)epoch << 32;
    const unsigned long long bits = (unsigned long long)__double_as_longlong(xi);
    __hip_atomic_store(slot + 2 * (size_t)row, tag | (bits & 0xffffffffull), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(slot + 2 * (size_t)row + 1, tag | (bits >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

}  // namespace

struct ks_lu_s {
  int n = 0, ldn = 0;
  double *d_A = nullptr, *d_x = nullptr, *d_y = nullptr;
  int *d_ipiv = nullptr, *d_perm = nullptr;
  unsigned long long *d_slot = nullptr;   // [ldn][2] published values of the triangular solves
  int *d_err = nullptr;                    // bit 0: a solve's wait ran out
  unsigned epoch = 0;
};

extern "C" int ks_lu_destroy(ks_lu F) {
  if (!F) return KS_OK;
  if (F->d_A) (void)ks_free(F->d_A);
  if (F->d_x) (void)ks_free(F->d_x);
  if (F->d_y) (void)ks_free(F->d_y);
  if (F->d_ipiv) (void)ks_free(F->d_ipiv);
  if (F->d_perm) (void)ks_free(F->d_perm);
  if (F->d_slot) (void)ks_free(F->d_slot);
  if (F->d_err) (void)ks_free(F->d_err);
  delete F;
  return KS_OK;
}

namespace {
int lu_factor(ks_ctx ctx, ks_csr A, double sigma, ks_lu *out, bool coop);
}

// The cooperative sub-panels wait on each other; should another stream's work
// keep some of them from running until a wait runs out, the factorization is
// made again with the column-step launches, which wait on nothing.
extern "C" int ks_lu_create(ks_ctx ctx, ks_csr A, double sigma, ks_lu *out) {
  // KS_LU_COOPERATIVE=0 in the environment selects the column-step launches
  const char *env = getenv("KS_LU_COOPERATIVE");
  if (env && env[0] == '0') return lu_factor(ctx, A, sigma, out, false);
  const int e = lu_factor(ctx, A, sigma, out, true);
  if (e != KS_ERR_TIMEOUT) return e;
  return lu_factor(ctx, A, sigma, out, false);
}

namespace {
int lu_factor(ks_ctx ctx, ks_csr A, double sigma, ks_lu *out, bool coop) {
  KS_CHECK(ctx && A && out, KS_ERR_ARG, "null LU argument");
  KS_CHECK(A->vprec == KS_F64 && A->nrows == A->ncols, KS_ERR_ARG, "the LU needs a square FP64 matrix");
  KS_CHECK(A->nrows <= KS_LU_MAXN, KS_ERR_RANGE, "the dense LU takes at most %d rows (%d given)", KS_LU_MAXN,
           A->nrows);
  static_assert((KS_LU_MAXN + KS_LU_NB - 1) / KS_LU_NB * KS_LU_NB <= KS_LU_PG_MAX * KS_BLOCK,
                "a panel step's workgroups must cover every row");
  static_assert((KS_LU_MAXN + KS_LU_NB - 1) / KS_LU_NB * KS_LU_NB <= KS_LU_CO_MAX * KS_LU_PT,
                "a cooperative panel's workgroups must cover every row");
  *out = nullptr;
  ks_lu F = new ks_lu_s();
  const int n = A->nrows, ldn = (n + KS_LU_NB - 1) / KS_LU_NB * KS_LU_NB;
  F->n = n;
  F->ldn = ldn;
  // factorization scratch: used[ldn], prow[64], flag, err, pi[2][PG_MAX] (ints); pm[2][PG_MAX]
  // (doubles); the cooperative panels' published words cand[2][CO_MAX][3], crow[2][CO_MAX][2 * 64]
  const size_t nint = (size_t)ldn + KS_LU_NB + 2 + 2 * KS_LU_PG_MAX;
  const size_t off_dbl = ((nint * sizeof(int) + 255) / 256) * 256;
  const size_t nwords = 2 * KS_LU_CO_MAX * (KS_LU_CAND + 2 * KS_LU_NB);
  const size_t scratch_bytes = off_dbl + 2 * KS_LU_PG_MAX * sizeof(double) + nwords * sizeof(unsigned long long);
  void *scratch = nullptr;
  int e = ks_malloc((void **)&F->d_A, sizeof(double) * (size_t)ldn * ldn);
  if (!e) e = ks_malloc((void **)&F->d_x, sizeof(double) * ldn);
  if (!e) e = ks_malloc((void **)&F->d_y, sizeof(double) * ldn);
  if (!e) e = ks_malloc((void **)&F->d_ipiv, sizeof(int) * ldn);
  if (!e) e = ks_malloc((void **)&F->d_perm, sizeof(int) * ldn);
  if (!e) e = ks_malloc((void **)&F->d_slot, sizeof(unsigned long long) * 2 * (size_t)ldn);
  if (!e) e = ks_malloc((void **)&F->d_err, sizeof(int));
  if (!e) e = ks_memset_zero(ctx, F->d_slot, sizeof(unsigned long long) * 2 * (size_t)ldn);
  if (!e) e = ks_memset_zero(ctx, F->d_err, sizeof(int));
  if (!e) e = ks_malloc(&scratch, scratch_bytes);
  if (!e) e = ks_memset_zero(ctx, F->d_A, sizeof(double) * (size_t)ldn * ldn);
  if (!e) e = ks_memset_zero(ctx, scratch, scratch_bytes);
  if (e) {
    if (scratch) (void)ks_free(scratch);
    ks_lu_destroy(F);
    return e;
  }
  int *d_used = (int *)scratch, *d_prow = d_used + ldn, *d_flag = d_prow + KS_LU_NB, *d_perr = d_flag + 1,
      *d_pi = d_perr + 1;
  double *d_pm = (double *)((char *)scratch + off_dbl);
  unsigned long long *d_cand = (unsigned long long *)(d_pm + 2 * KS_LU_PG_MAX);
  unsigned long long *d_crow = d_cand + 2 * KS_LU_CO_MAX * KS_LU_CAND;
  unsigned e0 = 1;     // the next cooperative sub-panel's first epoch (the words start at 0)
  // coop: cooperative launches wanted and available (else the column-step launches)
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
      if (coop) {   // the sub-panel in one cooperative launch
        const int Gc = (ldn - j0 + KS_LU_PT - 1) / KS_LU_PT;
        double *a_A = F->d_A;
        int a_ldn = ldn, a_j0 = j0;
        unsigned long long *a_cand = d_cand, *a_crow = d_crow;
        unsigned a_e0 = e0;
        int *a_prow = d_prow, *a_flag = d_flag, *a_err = d_perr;
        void *args[] = {&a_A, &a_ldn, &a_j0, &a_cand, &a_crow, &a_e0, &a_prow, &a_flag, &a_err};
        if (hipLaunchCooperativeKernel((const void *)k_lu_panel, dim3(Gc), dim3(KS_LU_PT), args, 0, st) ==
            hipSuccess)
          e0 += KS_LU_NB;
        else {
          (void)hipGetLastError();
          coop = false;
        }
      }
      if (!coop) {
        const int G = (ldn - j0 + KS_BLOCK - 1) / KS_BLOCK;
        hipLaunchKernelGGL(k_lu_colmax, dim3(G), dim3(KS_BLOCK), 0, st, (const double *)F->d_A, ldn, j0, d_pm,
                           d_pi);
        for (int j = j0; j < j0 + KS_LU_NB; j++)
          hipLaunchKernelGGL(k_lu_step, dim3(G), dim3(KS_BLOCK), 0, st, F->d_A, ldn, j0, j, d_used, d_pm, d_pi,
                             d_prow, d_flag);
      }
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
  int perr = 0;
  if (herr == hipSuccess) herr = hipMemcpyAsync(&flag, d_flag, sizeof(int), hipMemcpyDeviceToHost, st);
  if (herr == hipSuccess) herr = hipMemcpyAsync(&perr, d_perr, sizeof(int), hipMemcpyDeviceToHost, st);
  if (herr == hipSuccess) herr = hipStreamSynchronize(st);
  else (void)hipStreamSynchronize(st);   // nothing may still run on what is freed below
  (void)ks_free(scratch);
  if (herr != hipSuccess) {
    ks_lu_destroy(F);
    KS_HIP(herr);
  }
  if (perr) {   // the caller makes it again without cooperative launches
    ks_lu_destroy(F);
    return KS_ERR_TIMEOUT;
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

}  // namespace

extern "C" int ks_lu_solve(ks_ctx ctx, ks_lu F, const double *d_b, double *d_x) {
  KS_CHECK(ctx && F && d_b && d_x, KS_ERR_ARG, "null LU solve argument");
  const int n = F->n, ldn = F->ldn, nblk = ldn / KS_LU_NB;
  hipStream_t st = ctx->stream;
  // L y = P b (y in d_y), then U x = y (x in the caller's vector), each with an
  // epoch of its own (the published values of earlier solves hold older ones)
  const unsigned ef = ++F->epoch, eb = ++F->epoch;
  hipLaunchKernelGGL(k_lu_tri_sf<true>, dim3(nblk), dim3(KS_BLOCK), 0, st, (const double *)F->d_A, ldn, n,
                     (const int *)F->d_perm, d_b, F->d_y, (double *)nullptr, F->d_slot, ef, F->d_err);
  hipLaunchKernelGGL(k_lu_tri_sf<false>, dim3(nblk), dim3(KS_BLOCK), 0, st, (const double *)F->d_A, ldn, n,
                     (const int *)F->d_perm, (const double *)F->d_y, (double *)nullptr, d_x, F->d_slot, eb,
                     F->d_err);
  KS_HIP(hipGetLastError());
  return KS_OK;
}

extern "C" int ks_lu_check(ks_ctx ctx, ks_lu F) {
  KS_CHECK(ctx && F, KS_ERR_ARG, "null LU argument");
  int err = 0;
  KS_HIP(hipMemcpyAsync(&err, F->d_err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  KS_HIP(hipStreamSynchronize(ctx->stream));
  KS_CHECK(!err, KS_ERR_TIMEOUT, "a triangular solve's wait for a block ran out of polls");
  return KS_OK;
}
