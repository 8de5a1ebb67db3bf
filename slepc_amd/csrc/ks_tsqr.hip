// ks_tsqr.hip -- TSQR of a tall-skinny FP64 panel on one GPU (BVOrthogonalize
// with block types TSQR and TSQRCHOL; the reference's BVOrthogonalize_TSQR /
// _TSQRCHOL, src/sys/classes/bv/interface/bvorthog.c:615-663, factor the panel
// with LAPACK on each process and reduce the triangular factors across
// processes, bvlapack.c:347-560).
//
// Here the row blocks are workgroups.  The panel's rows are cut into leaves of
// rb rows (rb = 2k rounded up to 64, at least 64); each workgroup loads its
// leaf into LDS, factors it by Householder reflections (dgeqr2's steps) and,
// for TSQR, forms its explicit Q in place (dorg2r's steps) and writes it back
// over the leaf.  The leaves' k x k factors, stacked f = rb / k at a time, are
// the rows of the next level, factored by the same kernel, until one node
// remains: its R is the panel's.  For TSQR the Q of the panel is then the leaf
// Q times the product of the slices of the node Q's on the leaf's path to the
// root: a top-down pass forms that k x k product per leaf and a last pass
// multiplies each leaf by its own.
//
// Every level is one launch; the reductions inside a workgroup run in a fixed
// order, so a factorization is deterministic.  The workspace (the node levels
// and the per-leaf products, about 1.5 n k doubles in all) is allocated per call.
#include "ks_internal.hpp"

namespace {

#define KS_TSQR_NT 256
#define KS_TSQR_NW (KS_TSQR_NT / 64)

__device__ inline double tsqr_wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;   // every lane holds the sum
}

// Householder QR of the rb x k matrix A (LDS, column-major, leading dimension
// rb): R in the upper triangle, the reflectors' vectors below it (v_0 = 1
// implicit), their factors in tau.  dlarfg: beta = -sign(alpha) |(alpha, x)|,
// tau = (beta - alpha) / beta, v = x / (alpha - beta); tau = 0 when x = 0.
__device__ void tsqr_geqr2(double *A, int rb, int k, double *tau, double *red) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int j = 0; j < k; j++) {
    double *aj = A + (size_t)j * rb;
    double s = 0.0;
    for (int r = j + 1 + tid; r < rb; r += KS_TSQR_NT) s += aj[r] * aj[r];
    s = tsqr_wave_sum(s);
    if (lane == 0) red[wave] = s;
    __syncthreads();
    double xnorm2 = 0.0;
#pragma unroll
    for (int w = 0; w < KS_TSQR_NW; w++) xnorm2 += red[w];
    const double alpha = aj[j];
    double t = 0.0, scal = 0.0, beta = alpha;
    if (xnorm2 > 0.0) {
      beta = -copysign(sqrt(alpha * alpha + xnorm2), alpha);
      t = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    __syncthreads();   // every thread has read red and alpha
    if (t != 0.0)
      for (int r = j + 1 + tid; r < rb; r += KS_TSQR_NT) aj[r] *= scal;
    if (tid == 0) {
      aj[j] = beta;
      tau[j] = t;
    }
    __syncthreads();
    if (t != 0.0) {
      // H_j = I - t v v^T on the columns j+1..k-1, one wave per column
      for (int c = j + 1 + wave; c < k; c += KS_TSQR_NW) {
        double *ac = A + (size_t)c * rb;
        double d = lane == 0 ? ac[j] : 0.0;
        for (int r = j + 1 + lane; r < rb; r += 64) d += aj[r] * ac[r];
        d = t * tsqr_wave_sum(d);
        if (lane == 0) ac[j] -= d;
        for (int r = j + 1 + lane; r < rb; r += 64) ac[r] -= d * aj[r];
      }
    }
    __syncthreads();
  }
}

// dorg2r on the factored A: the first k columns of H_0 ... H_{k-1}, in place
__device__ void tsqr_org2r(double *A, int rb, int k, const double *tau) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int j = k - 1; j >= 0; j--) {
    double *aj = A + (size_t)j * rb;
    const double t = tau[j];
    if (t != 0.0)
      for (int c = j + 1 + wave; c < k; c += KS_TSQR_NW) {
        double *ac = A + (size_t)c * rb;
        double d = lane == 0 ? ac[j] : 0.0;
        for (int r = j + 1 + lane; r < rb; r += 64) d += aj[r] * ac[r];
        d = t * tsqr_wave_sum(d);
        if (lane == 0) ac[j] -= d;
        for (int r = j + 1 + lane; r < rb; r += 64) ac[r] -= d * aj[r];
      }
    __syncthreads();
    for (int r = tid; r < rb; r += KS_TSQR_NT) aj[r] = r > j ? -t * aj[r] : (r == j ? 1.0 - t : 0.0);
    __syncthreads();
  }
}

// One level: block b factors rows [b rb, min(b rb + rb, nrows)) of the level's
// matrix (block b's rows start at base + b bstride, leading dimension ld; rows
// past nrows count as zero).  Its R goes to rout: block b's k x k factor at
// rows (b % rf) k of the (rf k) x k matrix number b / rf, leading dimension
// rf k.  FORM_Q: its explicit Q overwrites its rows.
template <bool FORM_Q>
__global__ __launch_bounds__(KS_TSQR_NT) void k_tsqr_level(double *base, long ld, long bstride, int rb, long nrows,
                                                           int k, double *rout, int rf) {
  extern __shared__ double lds[];
  double *A = lds, *tau = A + (size_t)rb * k, *red = tau + k;
  const int b = blockIdx.x;
  double *in = base + (size_t)b * bstride;
  const long r0 = (long)b * rb;
  const int valid = (int)min((long)rb, nrows - r0);
  for (int c = 0; c < k; c++)
    for (int r = threadIdx.x; r < rb; r += KS_TSQR_NT) A[(size_t)c * rb + r] = r < valid ? in[r + (size_t)c * ld] : 0.0;
  __syncthreads();
  tsqr_geqr2(A, rb, k, tau, red);
  double *R = rout + (size_t)(b / rf) * ((size_t)rf * k * k) + (size_t)(b % rf) * k;
  const long ldr = (long)rf * k;
  for (int e = threadIdx.x; e < k * k; e += KS_TSQR_NT) {
    const int i = e % k, j = e / k;
    R[i + (size_t)j * ldr] = i <= j ? A[(size_t)j * rb + i] : 0.0;
  }
  if (FORM_Q) {
    __syncthreads();
    tsqr_org2r(A, rb, k, tau);
    for (int c = 0; c < k; c++)
      for (int r = threadIdx.x; r < valid; r += KS_TSQR_NT) in[r + (size_t)c * ld] = A[(size_t)c * rb + r];
  }
}

// Top-down: child b of the level below takes M_child[b] = S(b) M_parent[b / f],
// S(b) the k x k slice at rows (b % f) k of its parent's Q (leading dimension
// f k); M_parent = I at the root
__global__ __launch_bounds__(KS_TSQR_NT) void k_tsqr_down(const double *Q, int f, int k, const double *Mpar,
                                                          double *Mch) {
  const int b = blockIdx.x, p = b / f, t = b % f;
  const long ldq = (long)f * k;
  const double *S = Q + (size_t)p * ((size_t)f * k * k) + (size_t)t * k;
  const double *P = Mpar ? Mpar + (size_t)p * k * k : nullptr;
  double *out = Mch + (size_t)b * k * k;
  for (int e = threadIdx.x; e < k * k; e += KS_TSQR_NT) {
    const int i = e % k, j = e / k;
    double a;
    if (P) {
      a = 0.0;
      for (int c = 0; c < k; c++) a += S[i + (size_t)c * ldq] * P[c + (size_t)j * k];
    } else {
      a = S[i + (size_t)j * ldq];
    }
    out[i + (size_t)j * k] = a;
  }
}

// Leaf b: rows [b rb, ...) of V (its leaf Q) times M[b], in place through LDS
__global__ __launch_bounds__(KS_TSQR_NT) void k_tsqr_apply(double *V, long ld, long n, int rb, int k,
                                                           const double *M) {
  extern __shared__ double lds[];
  const int b = blockIdx.x;
  const long r0 = (long)b * rb;
  const int valid = (int)min((long)rb, n - r0);
  double *T = lds;
  double *Vb = V + r0;
  for (int c = 0; c < k; c++)
    for (int r = threadIdx.x; r < valid; r += KS_TSQR_NT) T[(size_t)c * rb + r] = Vb[r + (size_t)c * ld];
  __syncthreads();
  const double *Mb = M + (size_t)b * k * k;
  for (int r = threadIdx.x; r < valid; r += KS_TSQR_NT)
    for (int j = 0; j < k; j++) {
      double a = 0.0;
      for (int c = 0; c < k; c++) a += T[(size_t)c * rb + r] * Mb[c + (size_t)j * k];
      Vb[r + (size_t)j * ld] = a;
    }
}

}  // namespace

extern "C" int ks_bv_tsqr(ks_ctx ctx, int n, double *d_V, int ld, int k, int form_q, double *R, int ldr) {
  KS_CHECK(ctx && d_V && R, KS_ERR_ARG, "null TSQR argument");
  KS_CHECK(k >= 1 && n >= k && ld >= n && ldr >= k, KS_ERR_ARG, "TSQR needs 1 <= k <= n <= ld and ldr >= k");
  int dev = 0, maxlds = 0;
  KS_HIP(hipGetDevice(&dev));
  KS_HIP(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  const int rb = ((2 * k + 63) / 64) * 64 < 64 ? 64 : ((2 * k + 63) / 64) * 64;
  const size_t lds_level = sizeof(double) * ((size_t)rb * k + k + KS_TSQR_NW);
  const size_t lds_apply = sizeof(double) * (size_t)rb * k;
  KS_CHECK(lds_level <= (size_t)maxlds, KS_ERR_ARG,
           "TSQR of %d columns needs %zu bytes of LDS per workgroup (the device has %d)", k, lds_level, maxlds);
  const int f = rb / k;
  // level sizes: cnt[0] leaves, cnt[l] nodes of level l
  std::vector<long> cnt(1, ((long)n + rb - 1) / rb);
  while (cnt.back() > 1) cnt.push_back((cnt.back() + f - 1) / f);
  const int L = (int)cnt.size() - 1;
  KS_CHECK(cnt[0] <= 0x7fffffffL, KS_ERR_ARG, "TSQR panel too long");
  const size_t node = (size_t)f * k * k;
  std::vector<double *> buf(L + 1, nullptr), Mb(L + 1, nullptr);
  double *dR = nullptr;
  int rc = KS_OK;
  auto release = [&]() {
    for (double *p : buf)
      if (p) (void)ks_free(p);
    for (double *p : Mb)
      if (p) (void)ks_free(p);
    if (dR) (void)ks_free(dR);
  };
  rc = ks_malloc((void **)&dR, sizeof(double) * (size_t)k * k);
  for (int l = 1; l <= L && !rc; l++) rc = ks_malloc((void **)&buf[l], sizeof(double) * node * (size_t)cnt[l]);
  if (form_q)
    for (int l = 0; l < L && !rc; l++) rc = ks_malloc((void **)&Mb[l], sizeof(double) * (size_t)k * k * (size_t)cnt[l]);
  if (rc) {
    release();
    return rc;
  }
  hipStream_t st = ctx->stream;
  const dim3 blk(KS_TSQR_NT);
  // leaves
  {
    double *rout = L ? buf[1] : dR;
    const int rf = L ? f : 1;
    if (form_q)
      hipLaunchKernelGGL(k_tsqr_level<true>, dim3((unsigned)cnt[0]), blk, lds_level, st, d_V, (long)ld, (long)rb, rb,
                         (long)n, k, rout, rf);
    else
      hipLaunchKernelGGL(k_tsqr_level<false>, dim3((unsigned)cnt[0]), blk, lds_level, st, d_V, (long)ld, (long)rb, rb,
                         (long)n, k, rout, rf);
  }
  // node levels: level l factors the cnt[l-1] factors of the level below, f per node
  for (int l = 1; l <= L; l++) {
    double *rout = l < L ? buf[l + 1] : dR;
    const int rf = l < L ? f : 1;
    const long ldn = (long)f * k;
    if (form_q)
      hipLaunchKernelGGL(k_tsqr_level<true>, dim3((unsigned)cnt[l]), blk, lds_level, st, buf[l], ldn, (long)node,
                         f * k, cnt[l - 1] * k, k, rout, rf);
    else
      hipLaunchKernelGGL(k_tsqr_level<false>, dim3((unsigned)cnt[l]), blk, lds_level, st, buf[l], ldn, (long)node,
                         f * k, cnt[l - 1] * k, k, rout, rf);
  }
  if (form_q && L > 0) {
    for (int l = L; l >= 1; l--)
      hipLaunchKernelGGL(k_tsqr_down, dim3((unsigned)cnt[l - 1]), blk, 0, st, (const double *)buf[l], f, k,
                         (const double *)(l < L ? Mb[l] : nullptr), Mb[l - 1]);
    hipLaunchKernelGGL(k_tsqr_apply, dim3((unsigned)cnt[0]), blk, lds_apply, st, d_V, (long)ld, (long)n, rb, k,
                       (const double *)Mb[0]);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy2DAsync(R, sizeof(double) * (size_t)ldr, dR, sizeof(double) * (size_t)k,
                                            sizeof(double) * (size_t)k, (size_t)k, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  release();
  KS_CHECK(e == hipSuccess, KS_ERR_HIP, "TSQR launch failed: %s", hipGetErrorString(e));
  return KS_OK;
}
