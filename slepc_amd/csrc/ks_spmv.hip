// ks_spmv.hip -- CSR SpMV y = A x for gfx950 (MatMult of the hipcsr matrix type).
//
// Row-block schedule built once per matrix from the nnz/row histogram:
//  * stream blocks: up to KS_SPMV_ROWS rows whose nonzeros fit KS_SPMV_NNZ.  The
//    block loads its contiguous val/colidx range with fully coalesced accesses,
//    gathers x, stages the products in LDS, then each row is summed by a group of
//    threads sized from the block's row count (one thread per row for short
//    stencil rows, up to a wavefront per row when few long rows share a block);
//  * long-row blocks: a single row with more than KS_SPMV_NNZ nonzeros, strided
//    over the whole workgroup and reduced across its four wavefronts.
// The variant recorded for the matrix is "stream" when every block is a stream
// block and "vector" when long rows are present.
#include "ks_internal.hpp"

#include <type_traits>

namespace {

template <typename XT> struct acc_of { using type = double; };
template <> struct acc_of<float> { using type = float; };

// one row block b of y = A x (the body of both the single and the grouped kernel)
template <typename VT, typename XT>
__device__ __forceinline__ void spmv_block(int b, const int32_t *__restrict__ blk, const int32_t *__restrict__ rowptr,
                                           const int32_t *__restrict__ colidx, const VT *__restrict__ val,
                                           const XT *__restrict__ x, XT *__restrict__ y) {
  using AT = typename acc_of<XT>::type;
  __shared__ AT prod[KS_SPMV_NNZ];
  __shared__ double red[KS_WAVES];
  const int tid = threadIdx.x;
  const int r0 = blk[b], r1 = blk[b + 1];
  const int p0 = rowptr[r0], p1 = rowptr[r1];
  const int nrows = r1 - r0;
  if (nrows == 1 && p1 - p0 > KS_SPMV_NNZ) {
    // long row: the whole workgroup strides over it
    AT s = 0;
    for (int p = p0 + tid; p < p1; p += KS_BLOCK) s += (AT)val[p] * (AT)x[colidx[p]];
    double tot = ks_block_sum((double)s, red);
    if (tid == 0) y[r0] = (XT)tot;
    return;
  }
  const int cnt = p1 - p0;
  if (cnt > 0) {
    // all of this thread's column indices and values are loaded first, then all
    // its x gathers: KS_SPMV_NNZ / 256 independent requests in flight per level.
    // No per-lane guards on the loads (a guarded load compiles to a branch that
    // waits for it): a lane past the block's last entry re-reads that entry,
    // which sits in a cache line the wave fetches anyway, and its product is
    // not read.
    constexpr int PER = KS_SPMV_NNZ / KS_BLOCK;
    const int last = p1 - 1;
    int ci[PER];
    VT va[PER];
    AT xv[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int p = min(p0 + tid + u * KS_BLOCK, last);
      ci[u] = colidx[p];
      va[u] = val[p];
    }
#pragma unroll
    for (int u = 0; u < PER; u++) xv[u] = (AT)x[ci[u]];
    // every slot is written (slots past the block's entries are never read)
#pragma unroll
    for (int u = 0; u < PER; u++) prod[tid + u * KS_BLOCK] = (AT)va[u] * xv[u];
  }
  __syncthreads();
  // threads per row: largest power of two with nrows * tpr <= KS_BLOCK, at most 64
  int tpr = 1;
  while (tpr < 64 && nrows * (tpr << 1) <= KS_BLOCK) tpr <<= 1;
  const int lr = tid / tpr, sub = tid & (tpr - 1);
  if (lr < nrows) {
    const int rs = rowptr[r0 + lr] - p0, re = rowptr[r0 + lr + 1] - p0;
    AT s = 0;
    for (int q = rs + sub; q < re; q += tpr) s += prod[q];
    for (int off = tpr >> 1; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (sub == 0) y[r0 + lr] = (XT)s;
  }
}

// Workgroups go to the 8 XCDs round-robin (workgroup i to XCD i mod 8), each
// XCD with its own L2.  Row block xcd_block(i) gives XCD x a contiguous eighth
// of the row blocks, so the x entries a block gathers from neighbouring rows
// (a stencil's +-1, +-nx, +-nx*ny) are mostly in the L2 its XCD already filled.
// A bijection on [0, nb) for any nb.
#ifndef KS_SPMV_XCD
#define KS_SPMV_XCD 1
#endif
__device__ __forceinline__ int xcd_block(int i, int nb) {
  if (!KS_SPMV_XCD) return i;
  const int q = nb >> 3, r = nb & 7, x = i & 7, j = i >> 3;
  return x < r ? x * (q + 1) + j : r * (q + 1) + (x - r) * q + j;
}

template <typename VT, typename XT>
__global__ __launch_bounds__(KS_BLOCK) void k_spmv_csr(const int32_t *__restrict__ blk,
                                                       const int32_t *__restrict__ rowptr,
                                                       const int32_t *__restrict__ colidx,
                                                       const VT *__restrict__ val,
                                                       const XT *__restrict__ x,
                                                       XT *__restrict__ y,
                                                       const ks_status *__restrict__ st) {
  if (st && !st->active) return;
  spmv_block<VT, XT>(xcd_block(blockIdx.x, gridDim.x), blk, rowptr, colidx, val, x, y);
}

// the SpMVs of several independent solves in one launch: workgroup b belongs to
// the member m with b0[m] <= b < b0[m+1] and runs that member's row block b - b0[m]
__global__ __launch_bounds__(KS_BLOCK) void k_grp_spmv(ks_grp_spmv g) {
  int m = 0;
  while (m + 1 < g.nmem && (int)blockIdx.x >= g.b0[m + 1]) m++;
  if (!g.st[m]->active) return;
  spmv_block<double, double>((int)blockIdx.x - g.b0[m], g.blk[m], g.rowptr[m], g.colidx[m], g.val[m], g.x[m],
                             g.y[m]);
}

// host: row blocks from the row lengths
std::vector<int32_t> build_blocks(int n, const int32_t *rowptr, int *max_row) {
  std::vector<int32_t> b;
  b.reserve((size_t)n / 16 + 2);
  b.push_back(0);
  int mr = 0;
  int r = 0;
  while (r < n) {
    int len = rowptr[r + 1] - rowptr[r];
    if (len > mr) mr = len;
    if (len > KS_SPMV_NNZ) {
      r++;
      b.push_back(r);
      continue;
    }
    int nnz = 0, rows = 0;
    while (r < n && rows < KS_SPMV_ROWS) {
      int l = rowptr[r + 1] - rowptr[r];
      if (l > KS_SPMV_NNZ || nnz + l > KS_SPMV_NNZ) break;
      if (l > mr) mr = l;
      nnz += l;
      rows++;
      r++;
    }
    b.push_back(r);
  }
  *max_row = mr;
  return b;
}

int csr_finish(ks_csr A, ks_ctx ctx, const int32_t *h_rowptr) {
  int mr = 0;
  std::vector<int32_t> blocks = build_blocks(A->nrows, h_rowptr, &mr);
  A->max_row = mr;
  A->nblocks = (int)blocks.size() - 1;
  A->variant = mr > KS_SPMV_NNZ ? 1 : 0;
  KS_TRY(ks_malloc((void **)&A->d_blk, sizeof(int32_t) * blocks.size()));
  KS_HIP(hipMemcpyAsync(A->d_blk, blocks.data(), sizeof(int32_t) * blocks.size(), hipMemcpyHostToDevice, ctx->stream));
  KS_HIP(hipStreamSynchronize(ctx->stream));
  return KS_OK;
}

int csr_check_host(int nrows, int ncols, int64_t nnz, const int32_t *rowptr, const int32_t *colidx) {
  KS_CHECK(rowptr[0] == 0 && rowptr[nrows] == nnz, KS_ERR_ARG, "rowptr[0]=%d rowptr[n]=%d nnz=%lld", rowptr[0],
           rowptr[nrows], (long long)nnz);
  for (int i = 0; i < nrows; i++)
    KS_CHECK(rowptr[i + 1] >= rowptr[i], KS_ERR_ARG, "rowptr decreases at row %d", i);
  for (int64_t p = 0; p < nnz; p++)
    KS_CHECK(colidx[p] >= 0 && colidx[p] < ncols, KS_ERR_ARG, "column index %d out of range at %lld", colidx[p],
             (long long)p);
  return KS_OK;
}

}  // namespace

void ks_group_launch_spmv(hipStream_t stream, const ks_grp_spmv &g) {
  if (g.nmem > 0 && g.b0[g.nmem] > 0)
    hipLaunchKernelGGL(k_grp_spmv, dim3(g.b0[g.nmem]), dim3(KS_BLOCK), 0, stream, g);
}

int ks_csr_blocks(ks_csr A, const int32_t **blk, const int32_t **rowptr, const int32_t **colidx, const void **val,
                  int *nblocks, ks_prec *vprec) {
  *blk = A->d_blk;
  *rowptr = A->d_rowptr;
  *colidx = A->d_colidx;
  *val = A->d_val;
  *nblocks = A->nblocks;
  *vprec = A->vprec;
  return KS_OK;
}

// allocation inside a CSR constructor: release what was built on failure
#define KS_TRY_CSR(M, call)                 \
  do {                                      \
    int _e = (call);                        \
    if (_e) { ks_csr_destroy(M); return _e; } \
  } while (0)

extern "C" {

int ks_csr_create(ks_csr *out, ks_ctx ctx, int nrows, int ncols, int64_t nnz, const int32_t *rowptr,
                  const int32_t *colidx, const void *val, ks_prec vprec) {
  KS_CHECK(out && ctx && rowptr && (nnz == 0 || (colidx && val)), KS_ERR_ARG, "null CSR argument");
  KS_CHECK(nrows > 0 && ncols > 0 && nnz >= 0 && nnz < INT32_MAX, KS_ERR_ARG, "bad CSR sizes");
  int e = csr_check_host(nrows, ncols, nnz, rowptr, colidx);
  if (e) return e;
  ks_csr A = new ks_csr_s();
  A->nrows = nrows;
  A->ncols = ncols;
  A->nnz = nnz;
  A->vprec = vprec;
  size_t vs = vprec == KS_F64 ? 8 : 4;
  KS_TRY_CSR(A, ks_malloc((void **)&A->d_rowptr, sizeof(int32_t) * (nrows + 1)));
  KS_TRY_CSR(A, ks_malloc((void **)&A->d_colidx, sizeof(int32_t) * (nnz ? nnz : 1)));
  KS_TRY_CSR(A, ks_malloc((void **)&A->d_val, vs * (nnz ? nnz : 1)));
  KS_HIP(hipMemcpyAsync(A->d_rowptr, rowptr, sizeof(int32_t) * (nrows + 1), hipMemcpyHostToDevice, ctx->stream));
  if (nnz) {
    KS_HIP(hipMemcpyAsync(A->d_colidx, colidx, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, ctx->stream));
    KS_HIP(hipMemcpyAsync(A->d_val, val, vs * nnz, hipMemcpyHostToDevice, ctx->stream));
  }
  e = csr_finish(A, ctx, rowptr);
  if (e) { ks_csr_destroy(A); return e; }
  *out = A;
  return KS_OK;
}

int ks_csr_create_device(ks_csr *out, ks_ctx ctx, int nrows, int ncols, int64_t nnz, const int32_t *d_rowptr,
                         const int32_t *d_colidx, const void *d_val, ks_prec vprec) {
  KS_CHECK(out && ctx && d_rowptr && (nnz == 0 || (d_colidx && d_val)), KS_ERR_ARG, "null CSR argument");
  KS_CHECK(nrows > 0 && ncols > 0 && nnz >= 0 && nnz < INT32_MAX, KS_ERR_ARG, "bad CSR sizes");
  std::vector<int32_t> h_rowptr(nrows + 1);
  std::vector<int32_t> h_col(nnz);
  KS_HIP(hipMemcpy(h_rowptr.data(), d_rowptr, sizeof(int32_t) * (nrows + 1), hipMemcpyDeviceToHost));
  if (nnz) KS_HIP(hipMemcpy(h_col.data(), d_colidx, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost));
  int e = csr_check_host(nrows, ncols, nnz, h_rowptr.data(), h_col.data());
  if (e) return e;
  ks_csr A = new ks_csr_s();
  A->nrows = nrows;
  A->ncols = ncols;
  A->nnz = nnz;
  A->vprec = vprec;
  size_t vs = vprec == KS_F64 ? 8 : 4;
  KS_TRY_CSR(A, ks_malloc((void **)&A->d_rowptr, sizeof(int32_t) * (nrows + 1)));
  KS_TRY_CSR(A, ks_malloc((void **)&A->d_colidx, sizeof(int32_t) * (nnz ? nnz : 1)));
  KS_TRY_CSR(A, ks_malloc((void **)&A->d_val, vs * (nnz ? nnz : 1)));
  KS_HIP(hipMemcpyAsync(A->d_rowptr, d_rowptr, sizeof(int32_t) * (nrows + 1), hipMemcpyDeviceToDevice, ctx->stream));
  if (nnz) {
    KS_HIP(hipMemcpyAsync(A->d_colidx, d_colidx, sizeof(int32_t) * nnz, hipMemcpyDeviceToDevice, ctx->stream));
    KS_HIP(hipMemcpyAsync(A->d_val, d_val, vs * nnz, hipMemcpyDeviceToDevice, ctx->stream));
  }
  e = csr_finish(A, ctx, h_rowptr.data());
  if (e) { ks_csr_destroy(A); return e; }
  *out = A;
  return KS_OK;
}

int ks_csr_convert(ks_csr A, ks_ctx ctx, ks_prec to, ks_csr *out) {
  KS_CHECK(A && ctx && out, KS_ERR_ARG, "null argument");
  ks_csr B = new ks_csr_s();
  *B = *A;
  B->vprec = to;
  B->d_rowptr = B->d_colidx = B->d_blk = nullptr;
  B->d_val = nullptr;
  const size_t nz = A->nnz ? (size_t)A->nnz : 1;
  KS_TRY_CSR(B, ks_malloc((void **)&B->d_rowptr, sizeof(int32_t) * (A->nrows + 1)));
  KS_TRY_CSR(B, ks_malloc((void **)&B->d_colidx, sizeof(int32_t) * nz));
  KS_TRY_CSR(B, ks_malloc((void **)&B->d_blk, sizeof(int32_t) * (A->nblocks + 1)));
  KS_TRY_CSR(B, ks_malloc((void **)&B->d_val, (to == KS_F64 ? 8 : 4) * nz));
  KS_HIP(hipMemcpyAsync(B->d_rowptr, A->d_rowptr, sizeof(int32_t) * (A->nrows + 1), hipMemcpyDeviceToDevice, ctx->stream));
  KS_HIP(hipMemcpyAsync(B->d_colidx, A->d_colidx, sizeof(int32_t) * A->nnz, hipMemcpyDeviceToDevice, ctx->stream));
  KS_HIP(hipMemcpyAsync(B->d_blk, A->d_blk, sizeof(int32_t) * (A->nblocks + 1), hipMemcpyDeviceToDevice, ctx->stream));
  int e = ks_bv_convert(ctx, (int)A->nnz, A->vprec, A->d_val, to, B->d_val);
  if (e) { ks_csr_destroy(B); return e; }
  KS_HIP(hipStreamSynchronize(ctx->stream));
  *out = B;
  return KS_OK;
}

int ks_csr_destroy(ks_csr A) {
  if (!A) return KS_OK;
  (void)ks_free(A->d_rowptr);
  (void)ks_free(A->d_colidx);
  (void)ks_free(A->d_val);
  (void)ks_free(A->d_blk);
  delete A;
  return KS_OK;
}

int ks_csr_info(ks_csr A, int *nrows, int64_t *nnz, int *max_row, int *variant, int *nblocks) {
  KS_CHECK(A, KS_ERR_ARG, "null matrix");
  if (nrows) *nrows = A->nrows;
  if (nnz) *nnz = A->nnz;
  if (max_row) *max_row = A->max_row;
  if (variant) *variant = A->variant;
  if (nblocks) *nblocks = A->nblocks;
  return KS_OK;
}

int ks_spmv(ks_ctx ctx, ks_csr A, const void *x, void *y, ks_prec xprec, int gate) {
  KS_CHECK(ctx && A && x && y, KS_ERR_ARG, "null SpMV argument");
  return ks_spmv_gated(ctx, A, x, y, xprec, gate ? ctx->d_status : nullptr);
}

}  // extern "C"

// y = A x, skipped on the device when st is given and st->active is 0
int ks_spmv_gated(ks_ctx ctx, ks_csr A, const void *x, void *y, ks_prec xprec, const ks_status *st) {
  KS_CHECK(ctx && A && x && y, KS_ERR_ARG, "null SpMV argument");
  KS_CHECK(A->nrows == A->ncols || x != y, KS_ERR_ARG, "in-place SpMV");
  const double vs = A->vprec == KS_F64 ? 8.0 : 4.0, xs = xprec == KS_F64 ? 8.0 : 4.0;
  const double bytes = (double)A->nnz * (vs + 4.0) + (A->nrows + 1.0) * 4.0 + A->ncols * xs + A->nrows * xs;
  ks_prof_begin(ctx, KS_PROF_SPMV, bytes, 2.0 * (double)A->nnz, false);
  dim3 grid(A->nblocks), block(KS_BLOCK);
  if (A->vprec == KS_F64 && xprec == KS_F64)
    hipLaunchKernelGGL((k_spmv_csr<double, double>), grid, block, 0, ctx->stream, A->d_blk, A->d_rowptr,
                       A->d_colidx, (const double *)A->d_val, (const double *)x, (double *)y, st);
  else if (A->vprec == KS_F32 && xprec == KS_F32)
    hipLaunchKernelGGL((k_spmv_csr<float, float>), grid, block, 0, ctx->stream, A->d_blk, A->d_rowptr,
                       A->d_colidx, (const float *)A->d_val, (const float *)x, (float *)y, st);
  else if (A->vprec == KS_F32 && xprec == KS_F64)
    hipLaunchKernelGGL((k_spmv_csr<float, double>), grid, block, 0, ctx->stream, A->d_blk, A->d_rowptr,
                       A->d_colidx, (const float *)A->d_val, (const double *)x, (double *)y, st);
  else
    hipLaunchKernelGGL((k_spmv_csr<double, float>), grid, block, 0, ctx->stream, A->d_blk, A->d_rowptr,
                       A->d_colidx, (const double *)A->d_val, (const float *)x, (float *)y, st);
  ks_prof_end(ctx);
  KS_HIP(hipGetLastError());
  return KS_OK;
}
