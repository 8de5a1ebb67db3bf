// ks_transpose.hip -- the transpose of a device CSR matrix, built on the device.
//
// Two-sided Krylov-Schur runs an Arnoldi process on A^T beside the one on A
// (src/eps/impls/krylov/krylovschur/ks-twosided.c:145-155, MatCreateHermitianTranspose
// of the operator); here A^T is formed once as its own CSR so that every product
// with it is the ordinary row-blocked SpMV.
//
// Construction: a stable least-significant-digit radix sort of the nonzeros by
// column (8-bit digits, as many passes as the column count needs), carrying each
// nonzero's position.  Stability keeps the entries of one column in row-major
// order, so each row of A^T lists its entries in increasing column order and the
// result does not depend on thread timing.  A pass:
//   * k_rs_hist: each tile of 4096 nonzeros counts its digits in LDS and writes
//     the 256 counts digit-major (count of digit d in tile t at d * ntiles + t);
//   * an exclusive scan of those counts gives every (digit, tile) its first
//     output position: all smaller digits first, then the same digit of the
//     earlier tiles;
//   * k_rs_scatter: each tile walks its nonzeros in 16 rounds of 256 (thread
//     order = input order); a nonzero's place among the equal digits of its round
//     comes from ballots within the wavefront and per-wavefront counts in LDS,
//     so equal digits leave the tile in input order.
// The column counts (integer atomics) and their exclusive scan give the row
// pointer of A^T, and a gather moves row indices and values.  The scan is the
// usual three phases: per-workgroup scans of 4096 entries, a scan of the
// workgroup totals (the same routine, recursively), and the add-back.
#include "ks_internal.hpp"

#include <vector>

namespace {

#define KS_RS_BITS 8
#define KS_RS_DIGITS (1 << KS_RS_BITS)
#define KS_RS_ITEMS 16                              // nonzeros a thread handles per tile
#define KS_RS_TILE (KS_BLOCK * KS_RS_ITEMS)         // 4096
static_assert(KS_BLOCK == KS_RS_DIGITS, "one thread per digit in the tile prefix");

// rows[e] = the row of nonzero e
__global__ __launch_bounds__(KS_BLOCK) void k_expand_rows(const int32_t *__restrict__ rowptr, int nrows,
                                                         int32_t *__restrict__ rows) {
  for (int i = blockIdx.x * KS_BLOCK + threadIdx.x; i < nrows; i += gridDim.x * KS_BLOCK)
    for (int32_t e = rowptr[i]; e < rowptr[i + 1]; e++) rows[e] = i;
}

__global__ __launch_bounds__(KS_BLOCK) void k_iota(int32_t *__restrict__ p, int n) {
  for (int i = blockIdx.x * KS_BLOCK + threadIdx.x; i < n; i += gridDim.x * KS_BLOCK) p[i] = i;
}

// cnt[c] += 1 for every nonzero of column c
__global__ __launch_bounds__(KS_BLOCK) void k_col_counts(const int32_t *__restrict__ colidx, int nnz,
                                                        int32_t *__restrict__ cnt) {
  for (int e = blockIdx.x * KS_BLOCK + threadIdx.x; e < nnz; e += gridDim.x * KS_BLOCK) atomicAdd(&cnt[colidx[e]], 1);
}

// digit counts of tile blockIdx.x, written digit-major
__global__ __launch_bounds__(KS_BLOCK) void k_rs_hist(const int32_t *__restrict__ keys, int nnz, int shift,
                                                     int ntiles, int32_t *__restrict__ hist) {
  __shared__ int32_t c[KS_RS_DIGITS];
  c[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * KS_RS_TILE;
  for (int r = 0; r < KS_RS_ITEMS; r++) {
    const int64_t e = base + r * KS_BLOCK + threadIdx.x;
    if (e < nnz) atomicAdd(&c[(keys[e] >> shift) & (KS_RS_DIGITS - 1)], 1);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = c[threadIdx.x];
}

// the stable scatter of tile blockIdx.x: keys/vals to the positions off[d * ntiles + t]
// + (equal digits before it in the tile)
__global__ __launch_bounds__(KS_BLOCK) void k_rs_scatter(const int32_t *__restrict__ keys,
                                                        const int32_t *__restrict__ vals, int nnz, int shift,
                                                        int ntiles, const int32_t *__restrict__ off,
                                                        int32_t *__restrict__ keys_out,
                                                        int32_t *__restrict__ vals_out) {
  __shared__ int32_t wcnt[KS_WAVES][KS_RS_DIGITS];   // this round's count per wavefront and digit
  __shared__ int32_t wbase[KS_WAVES][KS_RS_DIGITS];  // where each wavefront's run of a digit starts
  __shared__ int32_t tbase[KS_RS_DIGITS];            // the digit's output position for this round
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const unsigned long long lt = (1ull << lane) - 1ull;
  for (int w = 0; w < KS_WAVES; w++) wcnt[w][tid] = 0;
  tbase[tid] = off[(size_t)tid * ntiles + blockIdx.x];
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * KS_RS_TILE;
  for (int r = 0; r < KS_RS_ITEMS; r++) {
    const int64_t e = base + r * KS_BLOCK + tid;
    const bool valid = e < nnz;
    const int32_t key = valid ? keys[e] : 0;
    const int32_t val = valid ? vals[e] : 0;
    const int d = (key >> shift) & (KS_RS_DIGITS - 1);
    // lanes of this wavefront holding the same digit (valid ones only)
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < KS_RS_BITS; b++) {
      const unsigned long long bal = __ballot((d >> b) & 1);
      peers &= ((d >> b) & 1) ? bal : ~bal;
    }
    const int rank = __popcll(peers & lt);
    if (valid && rank == 0) wcnt[wave][d] = __popcll(peers);
    __syncthreads();
    {   // thread tid owns digit tid: the wavefronts' runs in order, then the next round's start
      int32_t run = tbase[tid];
      for (int w = 0; w < KS_WAVES; w++) {
        wbase[w][tid] = run;
        run += wcnt[w][tid];
        wcnt[w][tid] = 0;
      }
      tbase[tid] = run;
    }
    __syncthreads();
    if (valid) {
      const int32_t dst = wbase[wave][d] + rank;
      keys_out[dst] = key;
      vals_out[dst] = val;
    }
  }
}

// per workgroup: out = exclusive scan of its 4096 entries of in, bsum[b] = their total
__global__ __launch_bounds__(KS_BLOCK) void k_scan_local(const int32_t *__restrict__ in, int64_t n,
                                                        int32_t *__restrict__ out, int32_t *__restrict__ bsum) {
  __shared__ int32_t s[KS_BLOCK];
  const int tid = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * KS_RS_TILE + (int64_t)tid * KS_RS_ITEMS;
  int32_t v[KS_RS_ITEMS], t = 0;
#pragma unroll
  for (int i = 0; i < KS_RS_ITEMS; i++) {
    v[i] = base + i < n ? in[base + i] : 0;
    t += v[i];
  }
  s[tid] = t;
  __syncthreads();
  for (int o = 1; o < KS_BLOCK; o <<= 1) {   // inclusive scan of the thread totals
    const int32_t x = tid >= o ? s[tid - o] : 0;
    __syncthreads();
    s[tid] += x;
    __syncthreads();
  }
  int32_t run = tid ? s[tid - 1] : 0;
#pragma unroll
  for (int i = 0; i < KS_RS_ITEMS; i++) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
  if (tid == KS_BLOCK - 1) bsum[blockIdx.x] = s[tid];
}

__global__ __launch_bounds__(KS_BLOCK) void k_scan_add(int32_t *__restrict__ out, int64_t n,
                                                      const int32_t *__restrict__ boff) {
  const int32_t add = boff[blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * KS_RS_TILE;
  for (int i = threadIdx.x; i < KS_RS_TILE; i += KS_BLOCK)
    if (base + i < n) out[base + i] += add;
}

template <typename T>
__global__ __launch_bounds__(KS_BLOCK) void k_gather_t(const int32_t *__restrict__ perm,
                                                      const int32_t *__restrict__ rows, const T *__restrict__ val,
                                                      int nnz, int32_t *__restrict__ colT, T *__restrict__ valT) {
  for (int k = blockIdx.x * KS_BLOCK + threadIdx.x; k < nnz; k += gridDim.x * KS_BLOCK) {
    const int32_t e = perm[k];
    colT[k] = rows[e];
    valT[k] = val[e];
  }
}

inline int grid_for(int64_t n) {
  int64_t b = (n + KS_BLOCK - 1) / KS_BLOCK;
  if (b > 16384) b = 16384;
  return b < 1 ? 1 : (int)b;
}

// exclusive scan of n entries (n >= 1); scratch holds the per-level workgroup totals
int scan_exclusive(hipStream_t s, const int32_t *in, int64_t n, int32_t *out, std::vector<int32_t *> &scratch) {
  const int64_t nb = (n + KS_RS_TILE - 1) / KS_RS_TILE;
  int32_t *bsum = nullptr;
  int e = ks_malloc((void **)&bsum, sizeof(int32_t) * (size_t)(2 * nb));
  if (e) return e;
  scratch.push_back(bsum);
  hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(KS_BLOCK), 0, s, in, n, out, bsum);
  if (nb > 1) {
    int32_t *boff = bsum + nb;
    e = scan_exclusive(s, bsum, nb, boff, scratch);
    if (e) return e;
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(KS_BLOCK), 0, s, out, n, (const int32_t *)boff);
  }
  return KS_OK;
}

}  // namespace

int ks_csr_transpose(ks_csr A, ks_ctx ctx, ks_csr *out) {
  KS_CHECK(A && ctx && out, KS_ERR_ARG, "null argument");
  KS_CHECK(A->nnz < INT32_MAX, KS_ERR_RANGE, "transpose supports fewer than 2^31 nonzeros");
  const int nnz = (int)A->nnz, nr = A->nrows, nc = A->ncols;
  const size_t vs = A->vprec == KS_F64 ? 8 : 4, nz = nnz ? (size_t)nnz : 1;
  const int ntiles = nnz ? (nnz + KS_RS_TILE - 1) / KS_RS_TILE : 1;
  int bits = 0;
  while (bits < 31 && (int64_t)(nc - 1) >> bits) bits++;
  const int passes = (bits + KS_RS_BITS - 1) / KS_RS_BITS;
  int32_t *rows = nullptr, *kbuf[2] = {nullptr, nullptr}, *vbuf[2] = {nullptr, nullptr}, *hist = nullptr,
          *hoff = nullptr, *cnt = nullptr, *rowptrT = nullptr, *colT = nullptr;
  void *valT = nullptr;
  std::vector<int32_t *> scratch;
  int err = KS_OK;
  hipStream_t s = ctx->stream;
#define KS_TR(call)                  \
  do {                               \
    err = (call);                    \
    if (err) goto done;              \
  } while (0)
#define KS_TR_HIP(call)                                                               \
  do {                                                                                \
    hipError_t _e = (call);                                                           \
    if (_e != hipSuccess) {                                                           \
      fprintf(stderr, "ks_hip: %s failed: %s\n", #call, hipGetErrorString(_e));       \
      err = KS_ERR_HIP;                                                               \
      goto done;                                                                      \
    }                                                                                 \
  } while (0)
  KS_TR(ks_malloc((void **)&rows, sizeof(int32_t) * nz));
  KS_TR(ks_malloc((void **)&kbuf[0], sizeof(int32_t) * nz));
  KS_TR(ks_malloc((void **)&kbuf[1], sizeof(int32_t) * nz));
  KS_TR(ks_malloc((void **)&vbuf[0], sizeof(int32_t) * nz));
  KS_TR(ks_malloc((void **)&vbuf[1], sizeof(int32_t) * nz));
  KS_TR(ks_malloc((void **)&hist, sizeof(int32_t) * (size_t)KS_RS_DIGITS * ntiles));
  KS_TR(ks_malloc((void **)&hoff, sizeof(int32_t) * (size_t)KS_RS_DIGITS * ntiles));
  KS_TR(ks_malloc((void **)&cnt, sizeof(int32_t) * (size_t)(nc + 1)));
  KS_TR(ks_malloc((void **)&rowptrT, sizeof(int32_t) * (size_t)(nc + 1)));
  KS_TR(ks_malloc((void **)&colT, sizeof(int32_t) * nz));
  KS_TR(ks_malloc(&valT, vs * nz));
  KS_TR(ks_memset_zero(ctx, cnt, sizeof(int32_t) * (size_t)(nc + 1)));
  {
    const int32_t *perm = vbuf[0];
    if (nnz) {
      hipLaunchKernelGGL(k_expand_rows, dim3(grid_for(nr)), dim3(KS_BLOCK), 0, s, A->d_rowptr, nr, rows);
      hipLaunchKernelGGL(k_iota, dim3(grid_for(nnz)), dim3(KS_BLOCK), 0, s, vbuf[0], nnz);
      hipLaunchKernelGGL(k_col_counts, dim3(grid_for(nnz)), dim3(KS_BLOCK), 0, s, A->d_colidx, nnz, cnt);
      const int32_t *keys = A->d_colidx;
      int cur = 0;
      for (int p = 0; p < passes; p++) {
        const int shift = p * KS_RS_BITS;
        hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(KS_BLOCK), 0, s, keys, nnz, shift, ntiles, hist);
        KS_TR(scan_exclusive(s, hist, (int64_t)KS_RS_DIGITS * ntiles, hoff, scratch));
        hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(KS_BLOCK), 0, s, keys, (const int32_t *)vbuf[cur], nnz,
                           shift, ntiles, (const int32_t *)hoff, kbuf[cur ^ 1], vbuf[cur ^ 1]);
        cur ^= 1;
        keys = kbuf[cur];
      }
      perm = vbuf[cur];
      if (A->vprec == KS_F64)
        hipLaunchKernelGGL(k_gather_t<double>, dim3(grid_for(nnz)), dim3(KS_BLOCK), 0, s, perm, rows,
                           (const double *)A->d_val, nnz, colT, (double *)valT);
      else
        hipLaunchKernelGGL(k_gather_t<float>, dim3(grid_for(nnz)), dim3(KS_BLOCK), 0, s, perm, rows,
                           (const float *)A->d_val, nnz, colT, (float *)valT);
    }
  }
  // cnt[nc] = 0, so the exclusive scan's last entry is nnz
  KS_TR(scan_exclusive(s, cnt, (int64_t)nc + 1, rowptrT, scratch));
  KS_TR_HIP(hipGetLastError());
  KS_TR_HIP(hipStreamSynchronize(s));
  KS_TR(ks_csr_create_device(out, ctx, nc, nr, A->nnz, rowptrT, colT, valT, A->vprec));
  KS_TR_HIP(hipStreamSynchronize(s));
#undef KS_TR
#undef KS_TR_HIP
done:
  if (err) (void)hipStreamSynchronize(s);   // no kernel may still use what is freed below
  (void)ks_free(rows);
  for (int i = 0; i < 2; i++) {
    (void)ks_free(kbuf[i]);
    (void)ks_free(vbuf[i]);
  }
  (void)ks_free(hist);
  (void)ks_free(hoff);
  (void)ks_free(cnt);
  (void)ks_free(rowptrT);
  (void)ks_free(colT);
  (void)ks_free(valT);
  for (int32_t *p : scratch) (void)ks_free(p);
  return err;
}

int ks_csr_download(ks_csr A, ks_ctx ctx, int32_t *rowptr, int32_t *colidx, void *val) {
  KS_CHECK(A && ctx && rowptr && (A->nnz == 0 || (colidx && val)), KS_ERR_ARG, "null argument");
  const size_t vs = A->vprec == KS_F64 ? 8 : 4;
  KS_HIP(hipMemcpyAsync(rowptr, A->d_rowptr, sizeof(int32_t) * (size_t)(A->nrows + 1), hipMemcpyDeviceToHost,
                        ctx->stream));
  if (A->nnz) {
    KS_HIP(hipMemcpyAsync(colidx, A->d_colidx, sizeof(int32_t) * (size_t)A->nnz, hipMemcpyDeviceToHost, ctx->stream));
    KS_HIP(hipMemcpyAsync(val, A->d_val, vs * (size_t)A->nnz, hipMemcpyDeviceToHost, ctx->stream));
  }
  KS_HIP(hipStreamSynchronize(ctx->stream));
  return KS_OK;
}
