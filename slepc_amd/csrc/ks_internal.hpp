// ks_internal.hpp -- shared pieces of libks_hip.so (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <vector>

#include "../../include/ks_hip.h"

#define KS_BLOCK 256                 // threads per block of the reduction kernels
#define KS_WAVES (KS_BLOCK / 64)
#define KS_MAXB 8192                 // maximum blocks of one row-partitioned reduction
#define KS_GROUP 32                  // blocks per first-level reduction group
#ifndef KS_GS_SPLITMAX
#define KS_GS_SPLITMAX 128           // Gram-Schmidt against <= this many columns: column-split kernel
#endif
// Loads of data a kernel streams exactly once (basis columns in a sweep): the
// nontemporal hint keeps them from displacing what is reused in L2 (the vector
// being orthogonalized, partial sums).  Measured on the split sweeps at
// n = 2^21: 5.0-5.2 -> 5.3-5.5 TB/s for k = 24..96.
template <typename T>
__device__ __forceinline__ T ks_stream_load(const T *p) {
  return __builtin_nontemporal_load(p);
}
typedef float ks_float2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ks_stream_load2(const float *p) {
  const ks_float2v v = __builtin_nontemporal_load((const ks_float2v *)p);
  return make_float2(v.x, v.y);
}

#ifndef KS_GS_SPLIT_BLOCKS
#define KS_GS_SPLIT_BLOCKS 1024      // workgroups of the column-split Gram-Schmidt (FP64 basis)
#endif
#ifndef KS_GS_SPLIT_BLOCKS32
#define KS_GS_SPLIT_BLOCKS32 2048    // the same for an FP32 basis
#endif
#ifndef KS_GS_RPL2_MAXCW
#define KS_GS_RPL2_MAXCW 12          // FP64 column-split sweeps with <= this many columns per wave: 2 rows per lane
#endif
#ifndef KS_GS_REGMAX
#define KS_GS_REGMAX 16              // Gram-Schmidt against <= this many columns: register-resident kernel
#endif
#ifndef KS_GS_REG_BLOCKS
#define KS_GS_REG_BLOCKS 512         // workgroups of the register-resident Gram-Schmidt
#endif
#ifndef KS_GS_BATCH
#define KS_GS_BATCH 8                // column loads in flight per thread in the Gram-Schmidt sweep
#endif
#define KS_MAXG (KS_MAXB / KS_GROUP)
#define KS_NCOLP (KS_MAX_COLS + 2)   // partial row width: columns + self dot + spare
#ifndef KS_SPMV_NNZ
#define KS_SPMV_NNZ 2048             // products staged in LDS per SpMV row block (a multiple of KS_BLOCK)
#endif
#define KS_SPMV_ROWS 256             // rows per SpMV row block (one per thread)

#define KS_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t _e = (call);                                                             \
    if (_e != hipSuccess) {                                                             \
      fprintf(stderr, "ks_hip: %s failed: %s (%s:%d)\n", #call, hipGetErrorString(_e),  \
              __FILE__, __LINE__);                                                      \
      return KS_ERR_HIP;                                                                \
    }                                                                                   \
  } while (0)

#define KS_TRY(call)          \
  do {                        \
    int _e = (call);          \
    if (_e) return _e;        \
  } while (0)

#define KS_CHECK(cond, code, ...)                                                       \
  do {                                                                                  \
    if (!(cond)) {                                                                      \
      fprintf(stderr, "ks_hip: " __VA_ARGS__);                                          \
      fputc('\n', stderr);                                                              \
      return (code);                                                                    \
    }                                                                                   \
  } while (0)

struct ks_prof_rec {
  int cls;
  int ran_slot;            // index into the device ran-flag array, -1 if always runs
  double bytes, flops;
  hipEvent_t ev0, ev1;
};

struct ks_ctx_s {
  hipStream_t stream = nullptr;
  ks_status *d_status = nullptr;
  double *d_partial = nullptr;     // [KS_MAXB][KS_NCOLP]
  double *d_gpartial = nullptr;    // [KS_MAXG][KS_NCOLP]
  unsigned *d_counters = nullptr;  // [KS_MAXG + 1], zero between launches
  unsigned long long *d_finpub = nullptr;   // k_gs_fin's epoch and epoch-tagged sums [1 + 2 * KS_NCOLP]
  double *d_coef1 = nullptr;                // first-pass coefficients kept for a deferred store [KS_NCOLP]
  double *d_scalar = nullptr;      // small device scratch (norm results)
  double *h_scalar = nullptr;      // pinned host scratch
  ks_status *h_status = nullptr;   // pinned mirror of d_status, filled by ks_status_post
  // profiling
  bool prof = false;
  unsigned prof_mask = ~0u;   // classes that get events while prof is on
  bool prof_open = false;     // the last ks_prof_begin recorded its start event
  bool capturing = false;     // between ks_graph_begin and ks_graph_end
  std::vector<ks_prof_rec> recs;
  std::vector<hipEvent_t> event_pool;
  int *d_ran = nullptr;            // per-launch "did work" flags
  int ran_cap = 0, ran_used = 0;
  ks_prof_entry acc[KS_PROF_NCLASS] = {};
};

// first-fit arena over [base, base+size), blocks aligned to KS_POOL_ALIGN
#define KS_POOL_ALIGN 256
struct ks_arena {
  char *base = nullptr;
  size_t size = 0, used = 0, peak = 0;
  std::map<size_t, size_t> free_;   // offset -> length, coalesced
  std::map<size_t, size_t> taken;   // offset -> length
  void init(char *b, size_t n) {
    base = b; size = n; used = peak = 0;
    free_.clear(); taken.clear();
    if (n) free_[0] = n;
  }
  bool owns(const void *p) const { return base && (const char *)p >= base && (const char *)p < base + size; }
  void *take(size_t n) {
    n = (n + KS_POOL_ALIGN - 1) / KS_POOL_ALIGN * KS_POOL_ALIGN;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      size_t off = it->first, len = it->second;
      free_.erase(it);
      if (len > n) free_[off + n] = len - n;
      taken[off] = n;
      used += n;
      if (used > peak) peak = used;
      return base + off;
    }
    return nullptr;
  }
  bool give(const void *p) {
    size_t off = (size_t)((const char *)p - base);
    auto t = taken.find(off);
    if (t == taken.end()) return false;
    size_t len = t->second;
    taken.erase(t);
    used -= len;
    auto nx = free_.lower_bound(off);
    if (nx != free_.end() && off + len == nx->first) { len += nx->second; free_.erase(nx); }
    auto nxt = free_.lower_bound(off);
    if (nxt != free_.begin()) {
      auto pv = std::prev(nxt);
      if (pv->first + pv->second == off) { pv->second += len; return true; }
    }
    free_[off] = len;
    return true;
  }
  size_t largest() const {
    size_t m = 0;
    for (auto &f : free_) m = f.second > m ? f.second : m;
    return m;
  }
};

struct ks_pool_s {
  ks_arena dev, host;
};

struct ks_csr_s {
  int nrows = 0, ncols = 0, max_row = 0, nblocks = 0, variant = 0;
  int64_t nnz = 0;
  ks_prec vprec = KS_F64;
  int32_t *d_rowptr = nullptr, *d_colidx = nullptr, *d_blk = nullptr;
  void *d_val = nullptr;
};

// ---------------------------------------------------------------- grouped launches (serving)
#define KS_GROUP_MAX 16   // solves per grouped launch
// members' SpMVs: workgroups b0[m] .. b0[m+1]-1 run member m's row blocks
struct ks_grp_spmv {
  int nmem;
  int b0[KS_GROUP_MAX + 1];
  const int32_t *blk[KS_GROUP_MAX], *rowptr[KS_GROUP_MAX], *colidx[KS_GROUP_MAX];
  const double *val[KS_GROUP_MAX];
  const double *x[KS_GROUP_MAX];
  double *y[KS_GROUP_MAX];
  const ks_status *st[KS_GROUP_MAX];
};
// members' Gram-Schmidt of one column each (FP64 basis)
struct ks_grp_gs_member {
  const double *V;
  double *w, *cbuf, *hcol, *partial, *sums;
  ks_status *st;
  unsigned *counter;
  unsigned long long *pub;         // the context's epoch-tagged finish area (k+2 finishing workgroups), or null
  double *coef1;                   // first-pass coefficient copy for a deferred store (null: none)
  long ld;
  int n, k, rpb, nb, col;
  int cw;                          // the member's own column split (as when it is grouped alone)
};
struct ks_grp_gs {
  int nmem;
  int b0_gs[KS_GROUP_MAX + 1];     // sweep workgroups
  int b0_fin[KS_GROUP_MAX + 1];    // finishing workgroups (k+2 per member, k+1 with a ticket)
  int b0_sc[KS_GROUP_MAX + 1];     // scaling workgroups
  ks_grp_gs_member m[KS_GROUP_MAX];
};
void ks_group_launch_spmv(hipStream_t stream, const ks_grp_spmv &g);
void ks_group_launch_gs(hipStream_t stream, const ks_grp_gs &g, int refine, double eta);
int ks_gs_split_geometry(int n, int k, int *cw, int *rpb, int *nb);   // KS_ERR_RANGE beyond the split kernel
int ks_csr_blocks(ks_csr A, const int32_t **blk, const int32_t **rowptr, const int32_t **colidx, const void **val,
                  int *nblocks, ks_prec *vprec);
int ks_ew_blocks(int n);
// y = A x, skipped on the device when st is given and st->active is 0 (the
// inner solvers gate their products on a status of their own)
int ks_spmv_gated(ks_ctx ctx, ks_csr A, const void *x, void *y, ks_prec xprec, const ks_status *st);

// ---------------------------------------------------------------- profiling hooks
// begin() returns a pointer to the device ran-flag for gated kernels (or null).
extern "C" int *ks_prof_begin(ks_ctx ctx, int cls, double bytes, double flops, bool gated);
extern "C" void ks_prof_end(ks_ctx ctx);

// ---------------------------------------------------------------- device helpers
// lane k's value of v, to every lane (k wave-uniform), as two 32-bit readlanes
__device__ __forceinline__ double ks_readlane_f64(double v, int k) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)bits, k), hi = __builtin_amdgcn_readlane((int)(bits >> 32), k);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Epoch-tagged publication: a 64-bit word holds a 32-bit payload (low half)
// and the epoch it was written in (high half), stored and loaded as one relaxed
// agent-scope atomic, so a reader that sees the current epoch sees the payload
// written with it, with no fence on either side.  A double takes two words.
__device__ __forceinline__ void ks_pub_store(unsigned long long *w, unsigned epoch, unsigned payload) {
  __hip_atomic_store(w, ((unsigned long long)epoch << 32) | payload, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ks_pub_store_f64(unsigned long long *w, unsigned epoch, double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  ks_pub_store(w, epoch, (unsigned)(b & 0xffffffffull));
  ks_pub_store(w + 1, epoch, (unsigned)(b >> 32));
}
__device__ __forceinline__ unsigned long long ks_pub_load(const unsigned long long *w) {
  return __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ks_pub_f64(unsigned long long lo, unsigned long long hi) {
  return __longlong_as_double((long long)(((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull)));
}

__device__ __forceinline__ double ks_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Wave sums of M values per lane by recursive halving: at each step the lanes
// of an xor pair keep opposite halves of their values and add the partner's
// copy of the half they keep, so M values cost about M + 3 log2(64) shuffles
// instead of 6 M.  While M stays even the halving continues (offsets 32, 16,
// ...); the values left are then summed over the remaining lane bits.  On
// return v[i] of lane l holds the wave total of value ks_multi_sum<M>::index(l, i).
// Fixed pairing and order, so the result does not depend on timing.
template <int M, int OFF>
struct ks_multi_halve {
  // every index below is a compile-time constant, so v stays in registers
  __device__ static __forceinline__ void run(double *v, int lane) {
    if constexpr (M % 2 == 0 && OFF >= 1) {
      constexpr int half = M / 2;
      const bool up = (lane & OFF) != 0;
#pragma unroll
      for (int i = 0; i < half; i++) {
        const double keep = up ? v[half + i] : v[i];
        const double give = up ? v[i] : v[half + i];
        v[i] = keep + __shfl_xor(give, OFF, 64);
      }
      ks_multi_halve<half, OFF / 2>::run(v, lane);
    } else {
#pragma unroll
      for (int off = OFF; off > 0; off >>= 1)
#pragma unroll
        for (int i = 0; i < M; i++) v[i] += __shfl_xor(v[i], off, 64);
    }
  }
};

template <int M>
struct ks_multi_sum {
  static constexpr int halvings(int m, int off) { return (m % 2 == 0 && off >= 1) ? 1 + halvings(m / 2, off / 2) : 0; }
  static constexpr int H = halvings(M, 32);    // halving steps
  static constexpr int R = M >> H;             // values left per lane
  __device__ static __forceinline__ void run(double *v, int lane) { ks_multi_halve<M, 32>::run(v, lane); }
  // which of the M values lane's v[i] holds after run()
  __device__ static __forceinline__ int index(int lane, int i) {
    int base = 0, m = M;
#pragma unroll
    for (int s = 0; s < H; s++) {
      const int half = m >> 1;
      if (lane & (32 >> s)) base += half;
      m = half;
    }
    return base + i;
  }
  // the lane that stores: lowest lane of its group of equal totals
  __device__ static __forceinline__ bool leader(int lane) { return (lane & ((64 >> H) - 1)) == 0; }
};

// Block-wide sum of one double; result valid in thread 0.  sh: >= KS_WAVES doubles.
__device__ __forceinline__ double ks_block_sum(double v, double *sh) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  v = ks_wave_sum(v);
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < KS_WAVES; w++) s += sh[w];
  }
  return s;
}

// Two-level "last block finishes" reduction of per-block partial rows.
// Every block has written partial[blockIdx.x * ncol + c] for c < ncol.  Returns
// true in exactly one block, whose threads then find the totals in out[0..ncol)
// (shared memory).  Summation order is fixed, so results are deterministic.
// Counters are returned to zero by the finishing blocks.
__device__ __forceinline__ bool ks_reduce_finish(double *partial, double *gpartial,
                                                 unsigned *counters, int nblocks, int ncol,
                                                 double *out) {
  __shared__ unsigned s_ticket;
  const int g = blockIdx.x / KS_GROUP;
  const int gbeg = g * KS_GROUP;
  const int gend = min(nblocks, gbeg + KS_GROUP);
  const int ngroups = (nblocks + KS_GROUP - 1) / KS_GROUP;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) s_ticket = atomicAdd(&counters[g], 1u);
  __syncthreads();
  if (s_ticket != (unsigned)(gend - gbeg - 1)) return false;
  __threadfence();
  for (int c = threadIdx.x; c < ncol; c += blockDim.x) {
    double s = 0.0;
    for (int b = gbeg; b < gend; b++) s += __builtin_nontemporal_load(&partial[(size_t)b * ncol + c]);
    gpartial[(size_t)g * ncol + c] = s;
  }
  if (threadIdx.x == 0) counters[g] = 0u;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) s_ticket = atomicAdd(&counters[KS_MAXG], 1u);
  __syncthreads();
  if (s_ticket != (unsigned)(ngroups - 1)) return false;
  __threadfence();
  for (int c = threadIdx.x; c < ncol; c += blockDim.x) {
    double s = 0.0;
    for (int q = 0; q < ngroups; q++) s += __builtin_nontemporal_load(&gpartial[(size_t)q * ncol + c]);
    out[c] = s;
  }
  if (threadIdx.x == 0) counters[KS_MAXG] = 0u;
  __syncthreads();
  return true;
}

// rows per thread (1,2,4,8,16) and block count for a row-partitioned reduction
inline void ks_rows_split(int n, int *rpt, int *nblocks) {
  int r = 1;
  while (r < 16 && (int64_t)(n + KS_BLOCK * r - 1) / (KS_BLOCK * r) > 1024) r <<= 1;
  int64_t nb = (n + (int64_t)KS_BLOCK * r - 1) / ((int64_t)KS_BLOCK * r);
  if (nb < 1) nb = 1;
  *rpt = r;
  *nblocks = (int)nb;
}

