// ks_transpose.hip -- the transpose of a device CSR matrix, built on the device.
//
// Two-sided Krylov-Schur runs an Arnoldi process on A^T beside the one on A
// (src/eps/impls/krylov/krylovschur/ks-twosided.c:145-155, MatCreateHermitianTranspose
// of the operator); here A^T is formed once as its own CSR so that every product
// with it is the ordinary row-blocked SpMV.  Construction: expand the row index of
// every nonzero, stable radix sort of the (column, position) pairs by column, so
// that each row of A^T lists its entries in increasing column order (the result
// is deterministic and independent of thread timing), column counts and their
// exclusive scan give the row pointer of A^T, and a gather moves the values.
#include "ks_internal.hpp"

#include <hipcub/hipcub.hpp>

namespace {

// rows[e] = the row of nonzero e
__global__ __launch_bounds__(KS_BLOCK) void k_expand_rows(const int32_t *__restrict__ rowptr, int nrows,
                                                         int32_t *__restrict__ rows) {
  for (int i = blockIdx.x * KS_BLOCK + threadIdx.x; i < nrows; i += gridDim.x * KS_BLOCK)
    for (int32_t e = rowptr[i]; e < rowptr[i + 1]; e++) rows[e] = i;
}

__global__ __launch_bounds__(KS_BLOCK) void k_iota(int32_t *__restrict__ p, int n) {
  for (int i = blockIdx.x * KS_BLOCK + threadIdx.x; i < n; i += gridDim.x * KS_BLOCK) p[i] = i;
}

// cnt[c] = number of nonzeros in column c, from the column-sorted keys: the first
// entry of each run writes its run length (found by binary search for the run's end)
__global__ __launch_bounds__(KS_BLOCK) void k_run_lengths(const int32_t *__restrict__ keys, int nnz,
                                                         int32_t *__restrict__ cnt) {
  for (int e = blockIdx.x * KS_BLOCK + threadIdx.x; e < nnz; e += gridDim.x * KS_BLOCK) {
    const int32_t c = keys[e];
    if (e > 0 && keys[e - 1] == c) continue;
    int lo = e, hi = nnz;   // first index with key > c
    while (lo < hi) {
      const int mid = lo + (hi - lo) / 2;
      if (keys[mid] <= c) lo = mid + 1;
      else hi = mid;
    }
    cnt[c] = lo - e;
  }
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

}  // namespace

int ks_csr_transpose(ks_csr A, ks_ctx ctx, ks_csr *out) {
  KS_CHECK(A && ctx && out, KS_ERR_ARG, "null argument");
  KS_CHECK(A->nnz < INT32_MAX, KS_ERR_RANGE, "transpose supports fewer than 2^31 nonzeros");
  const int nnz = (int)A->nnz, nr = A->nrows, nc = A->ncols;
  const size_t vs = A->vprec == KS_F64 ? 8 : 4, nz = nnz ? (size_t)nnz : 1;
  int32_t *rows = nullptr, *idx = nullptr, *keys = nullptr, *perm = nullptr, *cnt = nullptr, *rowptrT = nullptr,
          *colT = nullptr;
  void *valT = nullptr, *tmp = nullptr;
  size_t tmp_sort = 0, tmp_scan = 0;
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
  KS_TR(ks_malloc((void **)&idx, sizeof(int32_t) * nz));
  KS_TR(ks_malloc((void **)&keys, sizeof(int32_t) * nz));
  KS_TR(ks_malloc((void **)&perm, sizeof(int32_t) * nz));
  KS_TR(ks_malloc((void **)&cnt, sizeof(int32_t) * (size_t)(nc + 1)));
  KS_TR(ks_malloc((void **)&rowptrT, sizeof(int32_t) * (size_t)(nc + 1)));
  KS_TR(ks_malloc((void **)&colT, sizeof(int32_t) * nz));
  KS_TR(ks_malloc(&valT, vs * nz));
  KS_TR(ks_memset_zero(ctx, cnt, sizeof(int32_t) * (size_t)(nc + 1)));
  if (nnz) {
    hipLaunchKernelGGL(k_expand_rows, dim3(grid_for(nr)), dim3(KS_BLOCK), 0, s, A->d_rowptr, nr, rows);
    hipLaunchKernelGGL(k_iota, dim3(grid_for(nnz)), dim3(KS_BLOCK), 0, s, idx, nnz);
    KS_TR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_sort, A->d_colidx, keys, idx, perm, nnz, 0, 32, s));
  }
  KS_TR_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_scan, cnt, rowptrT, nc + 1, s));
  KS_TR(ks_malloc(&tmp, tmp_sort > tmp_scan ? tmp_sort : tmp_scan));
  if (nnz) {
    // stable: entries of one column keep their row-major order, i.e. increasing row
    KS_TR_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_sort, A->d_colidx, keys, idx, perm, nnz, 0, 32, s));
    hipLaunchKernelGGL(k_run_lengths, dim3(grid_for(nnz)), dim3(KS_BLOCK), 0, s, keys, nnz, cnt);
    if (A->vprec == KS_F64)
      hipLaunchKernelGGL(k_gather_t<double>, dim3(grid_for(nnz)), dim3(KS_BLOCK), 0, s, perm, rows,
                         (const double *)A->d_val, nnz, colT, (double *)valT);
    else
      hipLaunchKernelGGL(k_gather_t<float>, dim3(grid_for(nnz)), dim3(KS_BLOCK), 0, s, perm, rows,
                         (const float *)A->d_val, nnz, colT, (float *)valT);
    KS_TR_HIP(hipGetLastError());
  }
  // cnt[nc] = 0, so the exclusive sum's last entry is nnz
  KS_TR_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_scan, cnt, rowptrT, nc + 1, s));
  KS_TR_HIP(hipStreamSynchronize(s));
  KS_TR(ks_csr_create_device(out, ctx, nc, nr, A->nnz, rowptrT, colT, valT, A->vprec));
  KS_TR_HIP(hipStreamSynchronize(s));
#undef KS_TR
#undef KS_TR_HIP
done:
  (void)ks_free(rows);
  (void)ks_free(idx);
  (void)ks_free(keys);
  (void)ks_free(perm);
  (void)ks_free(cnt);
  (void)ks_free(rowptrT);
  (void)ks_free(colT);
  (void)ks_free(valT);
  (void)ks_free(tmp);
  return err;
}
