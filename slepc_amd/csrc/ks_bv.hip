// ks_bv.hip -- basis-vector kernels for gfx950: fused CGS+DGKS Gram-Schmidt,
// tall-skinny V^T y / V q, FP64-MFMA in-place restart GEMM, norm/scale/axpy tails.
//
// Layout: V column-major, leading dimension ld (elements), FP64 or FP32.
// Reductions are row-partitioned: each workgroup owns RPT*256 consecutive rows,
// keeps its slice of the vector in registers, sweeps the columns in tiles of
// KS_CT, reduces each tile across its four wavefronts (DPP/shuffle + LDS) and
// writes one partial row; the last workgroup of each group of 32 and then the
// last group finish the sums in a fixed order (deterministic, no float atomics).
#include "ks_internal.hpp"

#include <cmath>
#include <mutex>

#define KS_CT 8          // columns per register tile in the dot kernels
#define KS_MIP_KC 16     // k-chunk of Q staged in LDS

namespace {

enum { GS_FIRST = 0, GS_REPEAT = 1, GS_PLAIN = 2 };

template <typename T> struct acc_t { using type = double; };
template <> struct acc_t<float> { using type = float; };

__device__ __forceinline__ bool gs_gate(const ks_status *st, int mode, int col, int pass) {
  if (mode == GS_FIRST) return st->active != 0;
  if (mode == GS_REPEAT) return st->active && st->col == col && !st->done && st->pass_done == pass - 1;
  return true;
}

// partial[b][0..ntot) = V(rows_b, 0..k)^T w(rows_b)  (+ w^T w at column k unless PLAIN)
template <typename T, int RPT, int MODE>
__global__ __launch_bounds__(KS_BLOCK) void k_gs_dot(const T *__restrict__ V, long ld, int k,
                                                    const T *__restrict__ w, int n, int nblocks,
                                                    double *partial, double *gpartial, unsigned *counters,
                                                    ks_status *st, double *cout, double *hcol, int col,
                                                    int pass, int *ran, double deftol) {
  using AT = typename acc_t<T>::type;
  if (MODE != GS_PLAIN && !gs_gate(st, MODE, col, pass)) return;
  if (ran && blockIdx.x == 0 && threadIdx.x == 0) *ran = 1;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ntot = MODE == GS_PLAIN ? k : k + 1;
  __shared__ double sh[KS_WAVES * KS_CT];
  __shared__ double fin[KS_NCOLP];
  const size_t row0 = (size_t)blockIdx.x * KS_BLOCK * RPT;
  T wv[RPT];
  bool ok[RPT];
#pragma unroll
  for (int r = 0; r < RPT; r++) {
    const size_t i = row0 + (size_t)r * KS_BLOCK + tid;
    ok[r] = i < (size_t)n;
    wv[r] = ok[r] ? w[i] : (T)0;
  }
  for (int c0 = 0; c0 < ntot; c0 += KS_CT) {
    AT acc[KS_CT];
#pragma unroll
    for (int t = 0; t < KS_CT; t++) acc[t] = 0;
#pragma unroll
    for (int r = 0; r < RPT; r++) {
      if (!ok[r]) continue;
      const size_t i = row0 + (size_t)r * KS_BLOCK + tid;
#pragma unroll
      for (int t = 0; t < KS_CT; t++) {
        const int c = c0 + t;
        if (c < k) acc[t] += (AT)V[i + (size_t)c * ld] * (AT)wv[r];
      }
    }
    double self = 0.0;
    if (MODE != GS_PLAIN && c0 + KS_CT > k) {   // this tile holds the self dot (w^T w, FP64)
#pragma unroll
      for (int r = 0; r < RPT; r++) self += (double)wv[r] * (double)wv[r];
    }
    // reduce the tile across the workgroup
#pragma unroll
    for (int t = 0; t < KS_CT; t++) {
      double v = (MODE != GS_PLAIN && c0 + t == k) ? self : (double)acc[t];
      v = ks_wave_sum(v);
      if (lane == 0) sh[wave * KS_CT + t] = v;
    }
    __syncthreads();
    if (tid < KS_CT && c0 + tid < ntot) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < KS_WAVES; q++) s += sh[q * KS_CT + tid];
      partial[(size_t)blockIdx.x * ntot + c0 + tid] = s;
    }
    __syncthreads();
  }
  if (!ks_reduce_finish(partial, gpartial, counters, nblocks, ntot, fin)) return;
  // finishing workgroup.  With constraints the accumulated column may be the
  // scratch column itself (column 0), so it is cleaned before c is written.
  if (MODE == GS_FIRST) {
    for (int c = tid; c <= k; c += KS_BLOCK) hcol[c] = 0.0;   // BV_CleanCoefficients
    __syncthreads();
  }
  for (int c = tid; c < k; c += KS_BLOCK) cout[c] = fin[c];
  if (MODE == GS_PLAIN) return;
  if (tid == 0) {
    const double self = fin[k];
    double sumsq = 0.0;
    for (int c = 0; c < k; c++) sumsq += fin[c] * fin[c];    // BV_SquareSum
    st->sumsq = sumsq;
    if (self < -deftol) {                                    // BV_SafeSqrt
      st->error = KS_ERR_INNER;
      st->active = 0;
    }
    st->onrm = self < 0.0 ? 0.0 : sqrt(self);
    if (MODE == GS_FIRST) {
      st->col = col;
      st->pass_done = 0;
      st->done = 0;
      st->lindep = 0;
    }
  }
}

// GS: w -= V(:,0..k) c, partial |w|^2, DGKS decision in the finishing workgroup.
// PLAIN: w = beta*w + alpha*V c (no reduction).
template <typename T, int RPT, int MODE>
__global__ __launch_bounds__(KS_BLOCK) void k_gs_update(const T *__restrict__ V, long ld, int k, T *__restrict__ w,
                                                       int n, int nblocks, const double *__restrict__ cin,
                                                       double alpha, double beta, double *partial,
                                                       double *gpartial, unsigned *counters, ks_status *st,
                                                       double *hcol, int col, int pass, int refine, double eta,
                                                       int do_scale, int safe, int *ran) {
  using AT = typename acc_t<T>::type;
  if (MODE == GS_FIRST) {
    if (!(st->active && st->col == col && st->pass_done == 0 && !st->done)) return;
  } else if (MODE == GS_REPEAT) {
    if (!gs_gate(st, MODE, col, pass)) return;
  }
  if (ran && blockIdx.x == 0 && threadIdx.x == 0) *ran = 1;
  const int tid = threadIdx.x;
  __shared__ AT sc[KS_MAX_COLS];
  __shared__ double red[KS_WAVES];
  __shared__ double fin[2];
  for (int c = tid; c < k; c += KS_BLOCK) sc[c] = (AT)(MODE == GS_PLAIN ? alpha * cin[c] : cin[c]);
  __syncthreads();
  const size_t row0 = (size_t)blockIdx.x * KS_BLOCK * RPT;
  double sq = 0.0;
#pragma unroll
  for (int r = 0; r < RPT; r++) {
    const size_t i = row0 + (size_t)r * KS_BLOCK + tid;
    if (i >= (size_t)n) continue;
    AT a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    int c = 0;
    for (; c + 4 <= k; c += 4) {
      a0 += (AT)V[i + (size_t)c * ld] * sc[c];
      a1 += (AT)V[i + (size_t)(c + 1) * ld] * sc[c + 1];
      a2 += (AT)V[i + (size_t)(c + 2) * ld] * sc[c + 2];
      a3 += (AT)V[i + (size_t)(c + 3) * ld] * sc[c + 3];
    }
    for (; c < k; c++) a0 += (AT)V[i + (size_t)c * ld] * sc[c];
    const AT dot = (a0 + a1) + (a2 + a3);
    AT res;
    if (MODE == GS_PLAIN) res = (beta == 0.0 ? (AT)0 : (AT)beta * (AT)w[i]) + dot;
    else res = (AT)w[i] - dot;
    w[i] = (T)res;
    sq += (double)res * (double)res;
  }
  if (MODE == GS_PLAIN) return;
  double bs = ks_block_sum(sq, red);
  if (tid == 0) partial[blockIdx.x] = bs;
  if (!ks_reduce_finish(partial, gpartial, counters, nblocks, 1, fin)) return;
  for (int c = tid; c < k; c += KS_BLOCK) hcol[c] += cin[c];   // BV_AddCoefficients
  if (tid == 0) {
    const double exact = sqrt(fin[0]);
    st->exact_sq = fin[0];
    st->total_passes += 1;
    const int l = ++st->pass_done;
    const double onrm = st->onrm;
    double nrm = 0.0;
    int lindep = 0, finalize = 1, have_norm = 1;
    if (safe) {                                    // BV_OrthogonalizeColumn_Safe
      lindep = 1;
      have_norm = 0;
    } else if (refine == KS_REFINE_NEVER) {
      nrm = exact;
      lindep = !(nrm != 0.0);
    } else {
      if (refine == KS_REFINE_ALWAYS && l < 2) {
        finalize = 0;
      } else {
        nrm = onrm * onrm - st->sumsq;             // estimate |v'| from |v|
        nrm = (nrm <= 0.0) ? exact : sqrt(nrm);
        if (refine == KS_REFINE_IFNEEDED && l < 3 && nrm != 0.0 && fabs(nrm) < eta * fabs(onrm)) finalize = 0;
        else lindep = !(nrm != 0.0 && fabs(nrm) >= eta * fabs(onrm));
      }
      st->nrm = nrm;
    }
    if (finalize) {
      st->done = 1;
      st->lindep = lindep;
      st->nrm = nrm;
      st->beta_last = nrm;
      if (have_norm) hcol[k] = lindep ? 0.0 : nrm;   // BV_SetValue
      st->scale = (do_scale && nrm != 1.0 && nrm != 0.0) ? 1.0 / nrm : 1.0;
      if (lindep) {
        st->active = 0;
        st->breakdown_col = col;
      }
    }
  }
}

static_assert(KS_GS_REGMAX <= 48, "the register-resident kernel is instantiated for at most 48 columns");

// which launches of a column's pass sequence run: the first pass unless a
// previous column broke down; a fused pass only while DGKS asks for it
template <bool FUSED>
__device__ __forceinline__ bool gs_pass_gate(const ks_status *st, int col, int pass) {
  if (FUSED) return st->active && st->col == col && !st->done && st->pass_done == pass - 1;
  return st->active != 0;
}

// Runs on the one workgroup that finished the k+1 sums in fin[]: the first pass
// stores c and |v| (BVDotColumnInc + BV_SquareRoot); a fused pass adds its c
// to h (BV_AddCoefficients) and takes the DGKS decision of bvorthog.c:91-217.
// The outcome of the next pass is fixed before it runs: its |v| and
// coefficients (st->onrm, st->sumsq) are known here, and its finish decides from
// them alone unless the estimate |v|^2 - sum c^2 is not positive.  When that
// pass will be the last (and not lindep) and its sweep can apply a factor
// (prescale_ok: the split sweeps), it writes the normalized vector itself, the
// same rounding as the separate scaling, and the scaling launch finds nothing
// to do.
__device__ inline void gs_predict_last(ks_status *st, int col, int refine, double eta, int do_scale, int safe,
                                       int prescale_ok) {
  st->prescale_col = -1;
  if (!prescale_ok || !do_scale || safe || refine == KS_REFINE_NEVER) return;
  const int lp = st->pass_done + 1;
  if (refine == KS_REFINE_ALWAYS && lp < 2) return;
  const double onrm = st->onrm;
  double nrm = onrm * onrm - st->sumsq;
  if (!(nrm > 0.0)) return;
  nrm = sqrt(nrm);
  if (refine == KS_REFINE_IFNEEDED && lp < 3 && fabs(nrm) < eta * fabs(onrm)) return;   // another pass follows
  if (!(fabs(nrm) >= eta * fabs(onrm)) || nrm == 1.0) return;                          // lindep, or nothing to scale
  st->prescale = 1.0 / nrm;
  st->prescale_col = col;
  st->prescale_pass = lp;
}

// halt_pass > 0 (the lean sequence): the finish of that pass stops the
// expansion when the column needs a launch the lean sequence left out -- a
// further DGKS pass, or a scaling its last sweep did not apply.
// coef1 (first pass, split sweeps only; null: never): when the first fused pass
// is certain not to be the last -- its finish will decide from |v| and the
// coefficients known here, as gs_predict_last reads them -- that pass leaves
// its vector unstored and the next sweep forms it again from the loads it
// makes anyway (the same operations, so the same bits), saving a write of the
// column.  c of the first pass is kept in coef1 for that.
template <bool FUSED>
__device__ void gs_finish(const double *fin, int k, double *cbuf, double *hcol, ks_status *st, int col,
                          int refine, double eta, int do_scale, int safe, double deftol, int prescale_ok,
                          int halt_pass, double *coef1 = nullptr) {
  __shared__ int s_continue;
  const int tid = threadIdx.x;
  if (!FUSED) {
    // first pass: BV_CleanCoefficients, c, |v| (BVDotColumnInc + BV_SquareRoot)
    for (int c = tid; c <= k; c += blockDim.x) hcol[c] = 0.0;
    __syncthreads();
    for (int c = tid; c < k; c += blockDim.x) {
      cbuf[c] = fin[c];
      if (coef1) coef1[c] = fin[c];
    }
    if (tid == 0) {
      const double self = fin[k];
      double sumsq = 0.0;
      for (int c = 0; c < k; c++) sumsq += fin[c] * fin[c];
      st->sumsq = sumsq;
      if (self < -deftol) {
        st->error = KS_ERR_INNER;
        st->active = 0;
      }
      st->onrm = self < 0.0 ? 0.0 : sqrt(self);
      st->col = col;
      st->pass_done = 0;
      st->done = 0;
      st->lindep = 0;
      if (!st->error) gs_predict_last(st, col, refine, eta, do_scale, safe, prescale_ok);
      else st->prescale_col = -1;
      st->defer_col = -1;
      if (coef1 && !st->error && !safe && refine != KS_REFINE_NEVER) {
        // the first fused pass's finish (l = 1) continues when ALWAYS (l < 2), or
        // when the estimate is positive and below eta |v| (IFNEEDED, l < 3)
        bool more = refine == KS_REFINE_ALWAYS;
        const double onrm = st->onrm;
        double e = onrm * onrm - st->sumsq;
        if (!more && e > 0.0) {
          e = sqrt(e);
          more = refine == KS_REFINE_IFNEEDED && e != 0.0 && fabs(e) < eta * fabs(onrm);
        }
        if (more) {
          st->defer_col = col;
          st->defer_pass = 1;
        }
      }
    }
    return;
  }
  // fused pass: h += c of this pass, DGKS decision, next pass coefficients
  for (int c = tid; c < k; c += blockDim.x) hcol[c] += cbuf[c];   // BV_AddCoefficients
  __syncthreads();
  if (tid == 0) {
    const double exact = sqrt(fin[k]);
    st->exact_sq = fin[k];
    st->total_passes += 1;
    const int l = ++st->pass_done;
    const double onrm = st->onrm;
    double nrm = 0.0;
    int lindep = 0, finalize = 1, have_norm = 1;
    if (safe) {
      lindep = 1;
      have_norm = 0;
    } else if (refine == KS_REFINE_NEVER) {
      nrm = exact;
      lindep = !(nrm != 0.0);
    } else {
      if (refine == KS_REFINE_ALWAYS && l < 2) {
        finalize = 0;
      } else {
        nrm = onrm * onrm - st->sumsq;             // estimate |v'| from |v| (bvorthog.c:124-127)
        nrm = (nrm <= 0.0) ? exact : sqrt(nrm);
        if (refine == KS_REFINE_IFNEEDED && l < 3 && nrm != 0.0 && fabs(nrm) < eta * fabs(onrm)) finalize = 0;
        else lindep = !(nrm != 0.0 && fabs(nrm) >= eta * fabs(onrm));
      }
      st->nrm = nrm;
    }
    if (finalize) {
      st->done = 1;
      st->lindep = lindep;
      st->nrm = nrm;
      st->beta_last = nrm;
      if (have_norm) hcol[k] = lindep ? 0.0 : nrm;   // BV_SetValue
      st->scale = (do_scale && nrm != 1.0 && nrm != 0.0) ? 1.0 / nrm : 1.0;
      if (st->prescale_col == col && st->prescale_pass == l) st->scale = 1.0;   // the sweep applied it
      if (lindep) {
        st->active = 0;
        st->breakdown_col = col;
      } else if (halt_pass > 0 && st->scale != 1.0) {
        st->active = 0;
        st->halt_col = col;
        st->halt_need = 2;
      }
      s_continue = 0;
    } else {
      // next pass: its |v| and coefficients are the dots of the updated vector
      double sumsq = 0.0;
      for (int c = 0; c < k; c++) sumsq += fin[c] * fin[c];
      st->sumsq = sumsq;
      st->onrm = exact;
      gs_predict_last(st, col, refine, eta, do_scale, safe, prescale_ok);
      if (halt_pass == l) {
        st->active = 0;
        st->halt_col = col;
        st->halt_need = 1;
      }
      s_continue = 1;
    }
  }
  __syncthreads();
  if (s_continue)
    for (int c = tid; c < k; c += blockDim.x) cbuf[c] = fin[c];
}

// ---------------------------------------------------------------- LDS-staged Gram-Schmidt
// One CGS pass for column w against V(:,0..k), streamed in sub-chunks of R rows:
//   phase A  the workgroup loads R rows x k columns with coalesced column
//            segments into LDS (column-major, stride S = R + T) and, in the
//            fused form, applies the update w <- w - V c of this pass; the
//            vector (updated or not) becomes LDS column k;
//   phase B  thread (column c, row class sub) accumulates column c . column k
//            over rows sub, sub+T, ...  (c = k gives |w|^2), kept in registers
//            across all sub-chunks of the workgroup.
// A workgroup therefore reduces once, at the end, and the two-level finish
// produces the k+1 dots.  FUSED=false is the first pass (c = V^T w, w^T w);
// FUSED=true updates with this pass's c and returns the next pass's dots from
// the updated vector, so a second DGKS pass costs one V read instead of two.
template <typename T, int R, bool FUSED>
__global__ __launch_bounds__(KS_BLOCK) void k_gs_lds(const T *__restrict__ V, long ld, int k, T *__restrict__ w,
                                                    int n, int cpb, int tsub, const double *__restrict__ cbuf,
                                                    double *__restrict__ partial, const ks_status *st, int col,
                                                    int pass, int *ran) {
  if (!gs_pass_gate<FUSED>(st, col, pass)) return;
  if (ran && blockIdx.x == 0 && threadIdx.x == 0) *ran = 1;
  extern __shared__ double smem[];
  const int tid = threadIdx.x;
  const int S = R + tsub;
  const int G = KS_BLOCK / R;
  const int r = tid % R, g = tid / R;
  double *Vs = smem;                       // (k+1) columns x S
  double *coef = Vs + (size_t)(k + 1) * S; // k
  double *red = coef + k;                  // G x R
  if (FUSED)
    for (int c = tid; c < k; c += KS_BLOCK) coef[c] = cbuf[c];
  const int cpt = KS_BLOCK / tsub;         // columns per phase-B sweep
  const int sub = tid % tsub;
  double dacc[2] = {0.0, 0.0};
  __syncthreads();
  for (int ch = 0; ch < cpb; ch++) {
    const long r0 = ((long)blockIdx.x * cpb + ch) * R;
    if (r0 >= n) break;
    const long i = r0 + r;
    const bool ok = i < n;
    double acc = 0.0;
    // KS_GS_BATCH independent column loads are issued before any is used, so
    // each thread keeps that many HBM requests in flight
    for (int c0 = g; c0 < k; c0 += G * KS_GS_BATCH) {
      double v[KS_GS_BATCH];
#pragma unroll
      for (int u = 0; u < KS_GS_BATCH; u++) {
        const int c = c0 + u * G;
        v[u] = (ok && c < k) ? (double)V[i + (size_t)c * ld] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < KS_GS_BATCH; u++) {
        const int c = c0 + u * G;
        if (c < k) {
          Vs[(size_t)c * S + r] = v[u];
          if (FUSED) acc += v[u] * coef[c];
        }
      }
    }
    if (FUSED) {
      red[g * R + r] = acc;
      __syncthreads();
      if (g == 0) {
        double s = 0.0;
        for (int q = 0; q < G; q++) s += red[q * R + r];
        double wn = 0.0;
        if (ok) {
          const T t = (T)((double)w[i] - s);
          w[i] = t;
          wn = (double)t;
        }
        Vs[(size_t)k * S + r] = wn;
      }
    } else {
      if (g == 0) Vs[(size_t)k * S + r] = ok ? (double)w[i] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int c = j * cpt + tid / tsub;
      if (c <= k) {
        const double *vc = Vs + (size_t)c * S, *vk = Vs + (size_t)k * S;
        double s0 = 0.0, s1 = 0.0;
        int q = sub;
        for (; q + tsub < R; q += 2 * tsub) {
          s0 += vc[q] * vk[q];
          s1 += vc[q + tsub] * vk[q + tsub];
        }
        if (q < R) s0 += vc[q] * vk[q];
        dacc[j] += s0 + s1;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 2; j++) {
    double v = dacc[j];
    for (int off = tsub >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int c = j * cpt + tid / tsub;
    if (sub == 0 && c <= k) partial[(size_t)c * gridDim.x + blockIdx.x] = v;
  }
}

// ---------------------------------------------------------------- register-resident Gram-Schmidt
// The same pass for k <= KC columns with no LDS staging: a thread owns rows
// of its workgroup's contiguous range and for each row issues all KC column
// loads at once (KC requests in flight per thread), forms the update
// w_i - V(i,:) c of the fused pass from those registers, then adds V(i,c) w_i
// into k+1 register accumulators.  Slots c >= k re-read column k-1 (a cached
// line; no per-slot branch, which would make each load wait for the one
// before it) and contribute nothing: their coefficients are zero and their
// sums are not stored.  FP32 data: FP32 products and per-thread sums.  The
// workgroup reduces its accumulators once, at the end (wave butterflies in
// FP64, then the four waves through LDS), into its partial row.
// the factor a fused sweep (split or register) applies to the vector it writes (see gs_predict_last)
template <bool FUSED>
__device__ __forceinline__ double gs_prescale(const ks_status *st, int col, int pass) {
  return (FUSED && st->prescale_col == col && st->prescale_pass == pass) ? st->prescale : 1.0;
}

// the deferred-store role of this fused pass's sweep (see gs_split_block)
template <bool FUSED>
__device__ __forceinline__ int gs_defer(const ks_status *st, int col, int pass) {
  if (!FUSED || st->defer_col != col) return 0;
  return pass == st->defer_pass ? 1 : (pass == st->defer_pass + 1 ? 2 : 0);
}

template <typename T, int KC, bool FUSED>
__global__ __launch_bounds__(KS_BLOCK) void k_gs_reg(const T *__restrict__ V, long ld, int k, T *__restrict__ w,
                                                    int n, int rpb, const double *__restrict__ cbuf,
                                                    double *__restrict__ partial, const ks_status *st, int col,
                                                    int pass, int *ran) {
  if (!gs_pass_gate<FUSED>(st, col, pass)) return;
  if (ran && blockIdx.x == 0 && threadIdx.x == 0) *ran = 1;
  const double wscale = gs_prescale<FUSED>(st, col, pass);
  using AT = typename acc_t<T>::type;
  __shared__ AT coef[KC];
  __shared__ double red[KS_WAVES][KC + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (FUSED)
    for (int c = tid; c < KC; c += KS_BLOCK) coef[c] = (AT)(c < k ? cbuf[c] : 0.0);
  AT acc[KC + 1];
#pragma unroll
  for (int c = 0; c <= KC; c++) acc[c] = (AT)0;
  __syncthreads();
  // each workgroup owns one contiguous row range (rpb rows, a multiple of 256),
  // so a column's accesses stay inside a few pages while the workgroup runs;
  // column c is read as (column base) + (32-bit row byte offset), n * sizeof(T) < 4 GiB
  const int klast = k > 0 ? k - 1 : 0;   // k = 0: column 0 is w itself, a valid address
  const unsigned r0 = blockIdx.x * (unsigned)rpb;
  const unsigned r1 = min((unsigned)n, r0 + (unsigned)rpb);
  for (unsigned i = r0 + tid; i < r1; i += KS_BLOCK) {
    const unsigned off = i * (unsigned)sizeof(T);
    T v[KC];
#pragma unroll
    for (int c = 0; c < KC; c++) {
      const char *colbase = (const char *)(V + (size_t)min(c, klast) * ld);
      v[c] = *(const T *)(colbase + off);
    }
    AT wi = (AT)w[i];
    // keep the KC + 1 loads ahead of every use: left alone the scheduler
    // interleaves them with the accumulations and only a few stay in flight
    __builtin_amdgcn_sched_barrier(0);
    if (FUSED) {
      AT s0 = (AT)0, s1 = (AT)0;
#pragma unroll
      for (int c = 0; c < KC; c += 2) {
        s0 += (AT)v[c] * coef[c];
        s1 += (AT)v[c + 1] * coef[c + 1];
      }
      const T t = (T)(wi - (s0 + s1));
      w[i] = wscale == 1.0 ? t : (T)((AT)t * (AT)wscale);   // the final pass stores it normalized
      wi = (AT)t;
    }
#pragma unroll
    for (int c = 0; c < KC; c++) acc[c] += (AT)v[c] * wi;
    acc[KC] += wi * wi;
  }
  // workgroup reduction of k+1 sums; column k's slot carries |w|^2
#pragma unroll
  for (int c = 0; c <= KC; c++) {
    double x = ks_wave_sum((double)acc[c]);
    if (lane == 0) red[wave][c] = x;
  }
  __syncthreads();
  const int ntot = k + 1;
  for (int c = tid; c < ntot; c += KS_BLOCK) {
    const int src = c == k ? KC : c;
    double x = 0.0;
#pragma unroll
    for (int q = 0; q < KS_WAVES; q++) x += red[q][src];
    partial[(size_t)c * gridDim.x + blockIdx.x] = x;
  }
}

// ---------------------------------------------------------------- column-split Gram-Schmidt
// The four wavefronts of a workgroup split the k columns (wave q owns columns
// q*CW .. q*CW+CW-1) and walk the same rows together, 64 rows per step (one per
// lane).  Each thread keeps only CW values and CW accumulators, so occupancy
// stays high while every wave still issues CW coalesced 512-byte column loads
// per step.  In the fused pass the four partial sums of V(i,:) c meet in LDS
// (double-buffered, one barrier per step) and are added in wave order, so
// every wave forms the same updated w_i; wave 0 stores it and accumulates
// |w|^2.  Each workgroup owns a contiguous range of rpb rows.
// workgroup b of nb: rows b*rpb .. (b+1)*rpb-1.  PAIR (FP32 data, even ld and
// 8-byte aligned columns): each lane takes two adjacent rows with one 8-byte
// load per column, so a wave still moves 512 contiguous bytes per column per
// step, as with FP64 data.
// defer (fused passes): 1 -- this pass leaves its vector unstored; 2 -- the
// previous pass did, so each step first forms that pass's vector from the
// loaded rows and coef1 (the previous coefficients), with the operations that
// pass used, then applies this pass's coefficients to it.
// REFORM: the sweep can take defer == 2 (its extra sums cost registers, so only
// the launches that may follow a deferred pass are built with it).
template <typename T, int CW, bool FUSED, int RPL, bool PAIR, bool REFORM = false>
__device__ __forceinline__ void gs_split_block(int b, int nb, const T *__restrict__ V, long ld, int k,
                                               T *__restrict__ w, int n, int rpb, const double *__restrict__ cbuf,
                                               double *__restrict__ partial, double wscale, int defer = 0,
                                               const double *__restrict__ coef1 = nullptr) {
  static_assert(!PAIR || (sizeof(T) == 4 && RPL == 2), "row pairs are two adjacent FP32 rows");
  using AT = typename acc_t<T>::type;   // FP32 data: FP32 products and per-thread sums
  constexpr int STEP = 64 * RPL;        // rows per step; each lane takes RPL of them
  __shared__ AT coef[KS_WAVES * CW];
  __shared__ AT red[2][KS_WAVES][STEP];
  __shared__ AT coefp[REFORM ? KS_WAVES * CW : 1];          // coef1, when defer == 2
  __shared__ AT redp[REFORM ? 2 : 1][KS_WAVES][REFORM ? STEP : 1];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform: column guards are scalar branches
  const int c0 = wave * CW;
  // this lane's q-th row of a step: adjacent pairs (one 8-byte load) or 64 apart
  auto loff = [&](int q) { return PAIR ? RPL * lane + q : 64 * q + lane; };
  if (FUSED)
    for (int c = tid; c < KS_WAVES * CW; c += KS_BLOCK) {
      coef[c] = (AT)(c < k ? cbuf[c] : 0.0);
      if (REFORM && defer == 2) coefp[c] = (AT)(c < k ? coef1[c] : 0.0);
    }
  __syncthreads();
  AT acc[CW];
#pragma unroll
  for (int u = 0; u < CW; u++) acc[u] = (AT)0;
  AT self = (AT)0;
  const unsigned r0 = (unsigned)b * (unsigned)rpb;
  const unsigned r1 = min((unsigned)n, r0 + (unsigned)rpb);
  int buf = 0;
  // the step's CW x RPL values of V and its RPL values of w
  auto load_step = [&](unsigned i0, T (&v)[RPL][CW], AT (&wi)[RPL]) {
    if (i0 + STEP <= r1) {
      // every lane's rows exist (all steps but the last workgroup's last): no
      // per-lane guards, so the CW * RPL loads are issued back to back
#pragma unroll
      for (int u = 0; u < CW; u++) {
        const T *col = V + (size_t)(c0 + u) * ld;
        if (c0 + u < k) {
          if (PAIR) {
            const float2 p2 = ks_stream_load2((const float *)(col + i0 + loff(0)));
            v[0][u] = p2.x;
            v[RPL - 1][u] = p2.y;
          } else {
#pragma unroll
            for (int q = 0; q < RPL; q++) v[q][u] = ks_stream_load(col + i0 + loff(q));
          }
        } else {
#pragma unroll
          for (int q = 0; q < RPL; q++) v[q][u] = (T)0;
        }
      }
#pragma unroll
      for (int q = 0; q < RPL; q++) wi[q] = (AT)w[i0 + loff(q)];
    } else {
#pragma unroll
      for (int u = 0; u < CW; u++) {
        const T *col = V + (size_t)(c0 + u) * ld;
        const bool have = c0 + u < k;
#pragma unroll
        for (int q = 0; q < RPL; q++) v[q][u] = (have && i0 + loff(q) < r1) ? col[i0 + loff(q)] : (T)0;
      }
#pragma unroll
      for (int q = 0; q < RPL; q++) wi[q] = i0 + loff(q) < r1 ? (AT)w[i0 + loff(q)] : (AT)0;
    }
  };
  // the step's work on loaded values: the fused update through LDS, then the dots
  auto work_step = [&](unsigned i0, T (&v)[RPL][CW], AT (&wi)[RPL]) {
    if (FUSED) {
#pragma unroll
      for (int q = 0; q < RPL; q++) {
        AT s = (AT)0;
#pragma unroll
        for (int u = 0; u < CW; u++) s += (AT)v[q][u] * coef[c0 + u];
        red[buf][wave][loff(q)] = s;
        if (REFORM && defer == 2) {
          AT sp = (AT)0;
#pragma unroll
          for (int u = 0; u < CW; u++) sp += (AT)v[q][u] * coefp[c0 + u];
          redp[buf][wave][loff(q)] = sp;
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < RPL; q++) {
        const int r = loff(q);
        const unsigned i = i0 + r;
        AT base = wi[q];
        if (REFORM && defer == 2)   // the previous pass's vector, as that pass formed it
          base = (AT)(T)(base - (((redp[buf][0][r] + redp[buf][1][r]) + redp[buf][2][r]) + redp[buf][3][r]));
        const AT tot = ((red[buf][0][r] + red[buf][1][r]) + red[buf][2][r]) + red[buf][3][r];
        const T t = (T)(base - tot);
        wi[q] = i < r1 ? (AT)t : (AT)0;
        // the final pass of a column stores it normalized, rounded as k_scale_status would
        if (defer != 1 && wave == 0 && i < r1) w[i] = wscale == 1.0 ? t : (T)((AT)t * (AT)wscale);
      }
      buf ^= 1;
    }
#pragma unroll
    for (int q = 0; q < RPL; q++) {
#pragma unroll
      for (int u = 0; u < CW; u++) acc[u] += (AT)v[q][u] * wi[q];
      if (wave == 0) self += wi[q] * wi[q];
    }
  };
  for (unsigned i0 = r0; i0 < r1; i0 += STEP) {
    T v[RPL][CW];
    AT wi[RPL];
    load_step(i0, v, wi);
    __builtin_amdgcn_sched_barrier(0);   // all loads of the step before any use
    work_step(i0, v, wi);
  }
  // lanes and workgroups combine in FP64; all CW sums of the wave by recursive halving
  double t[CW];
#pragma unroll
  for (int u = 0; u < CW; u++) t[u] = (double)acc[u];
  using MS = ks_multi_sum<CW>;
  MS::run(t, lane);
  if (MS::leader(lane)) {
#pragma unroll
    for (int i = 0; i < MS::R; i++) {
      const int c = c0 + MS::index(lane, i);
      if (c < k) partial[(size_t)c * nb + b] = t[i];
    }
  }
  if (wave == 0) {
    const double x = ks_wave_sum((double)self);
    if (lane == 0) partial[(size_t)k * nb + b] = x;
  }
}

template <typename T, int CW, bool FUSED, int RPL, bool PAIR, bool REFORM>
__global__ __launch_bounds__(KS_BLOCK) void k_gs_split(const T *__restrict__ V, long ld, int k, T *__restrict__ w,
                                                      int n, int rpb, const double *__restrict__ cbuf,
                                                      double *__restrict__ partial, const ks_status *st, int col,
                                                      int pass, int *ran, const double *__restrict__ coef1) {
  if (!gs_pass_gate<FUSED>(st, col, pass)) return;
  if (ran && blockIdx.x == 0 && threadIdx.x == 0) *ran = 1;
  gs_split_block<T, CW, FUSED, RPL, PAIR, REFORM>(blockIdx.x, gridDim.x, V, ld, k, w, n, rpb, cbuf, partial,
                                                  gs_prescale<FUSED>(st, col, pass), gs_defer<FUSED>(st, col, pass),
                                                  coef1);
}
// the re-forming sweep held to the occupancy of the plain fused sweep where that
// is four waves per SIMD (its extra sums would otherwise take it to three)
template <typename T, int CW, bool FUSED, int RPL, bool PAIR, bool REFORM>
__global__ __launch_bounds__(KS_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_gs_split_w4(
    const T *__restrict__ V, long ld, int k, T *__restrict__ w, int n, int rpb, const double *__restrict__ cbuf,
    double *__restrict__ partial, const ks_status *st, int col, int pass, int *ran, const double *__restrict__ coef1) {
  if (!gs_pass_gate<FUSED>(st, col, pass)) return;
  if (ran && blockIdx.x == 0 && threadIdx.x == 0) *ran = 1;
  gs_split_block<T, CW, FUSED, RPL, PAIR, REFORM>(blockIdx.x, gridDim.x, V, ld, k, w, n, rpb, cbuf, partial,
                                                  gs_prescale<FUSED>(st, col, pass), gs_defer<FUSED>(st, col, pass),
                                                  coef1);
}

// The finish of a Gram-Schmidt pass: a separate launch, ordered by the stream
// after every partial sum of the sweep is written (so the sweep kernels, which
// write most of the data, need no fences or atomics).  The sweeps store their
// partial sums column-major, partial[c * nb + b]; here each of the k+1
// columns gets a workgroup that adds its nb values in a fixed order
// (coalesced, eight loads in flight per thread, then the workgroup tree), and
// the last workgroup to finish -- counted with one device-scope atomic after
// a fence, cheap in this kernel since it has written only k+1 sums -- takes
// the DGKS step on them.  The counter resets itself.
// column b's nb partial sums added in a fixed order; the sum is valid in thread 0
template <int NT>
__device__ __forceinline__ double gs_fin_colsum(int b, const double *__restrict__ partial, int nb) {
  constexpr int NW = NT / 64;
  __shared__ double red[NW];
  const double *p = partial + (size_t)b * nb;
  // FIN_U predicated loads per thread issued together: one memory round trip
  // covers a column of up to FIN_U * NT partials (6144 with 256 threads)
  constexpr int FIN_U = 24;
  double a[FIN_U];
#pragma unroll
  for (int u = 0; u < FIN_U; u++) a[u] = 0.0;
  for (int q0 = threadIdx.x; q0 < nb; q0 += FIN_U * NT) {
#pragma unroll
    for (int u = 0; u < FIN_U; u++) {
      const int q = q0 + u * NT;
      a[u] += q < nb ? p[q] : 0.0;
    }
  }
#pragma unroll
  for (int h = FIN_U / 2; h >= 3; h /= 2)   // 24 -> 12 -> 6 -> 3, pairwise in a fixed order
#pragma unroll
    for (int u = 0; u < h; u++) a[u] += a[u + h];
  double v = ks_wave_sum((a[0] + a[1]) + a[2]);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < NW; w++) s += red[w];
  }
  return s;
}

template <bool FUSED, int NT>
__device__ __forceinline__ void gs_fin_block(int b, const double *__restrict__ partial, int nb, int k, double *sums,
                                             unsigned *counter, double *cbuf, double *hcol, ks_status *st, int col,
                                             int refine, double eta, int do_scale, int safe, double deftol,
                                             int prescale_ok, int halt_pass, double *coef1 = nullptr) {
  __shared__ double fin[KS_NCOLP];
  __shared__ unsigned ticket;
  const double s = gs_fin_colsum<NT>(b, partial, nb);
  if (threadIdx.x == 0) sums[b] = s;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) ticket = atomicAdd(counter, 1u);
  __syncthreads();
  if (ticket != (unsigned)k) return;   // k+1 workgroups: the last one continues
  __threadfence();
  for (int c = threadIdx.x; c <= k; c += blockDim.x) fin[c] = __builtin_nontemporal_load(&sums[c]);
  if (threadIdx.x == 0) *counter = 0u;
  __syncthreads();
  gs_finish<FUSED>(fin, k, cbuf, hcol, st, col, refine, eta, do_scale, safe, deftol, prescale_ok, halt_pass, coef1);
}

// one 256-thread workgroup per column (1024-thread workgroups, which read a
// column's partial sums in one round, measured 1.5-2.5x slower end to end)
#define KS_FIN_BLOCK 256
template <bool FUSED>
__global__ __launch_bounds__(KS_FIN_BLOCK) void k_gs_fin(const double *__restrict__ partial, int nb, int k,
                                                        double *sums, unsigned *counter, double *cbuf, double *hcol,
                                                        ks_status *st, int col, int pass, int refine, double eta,
                                                        int do_scale, int safe, double deftol, int prescale_ok,
                                                        int halt_pass, double *coef1) {
  if (!gs_pass_gate<FUSED>(st, col, pass)) return;
  gs_fin_block<FUSED, KS_FIN_BLOCK>(blockIdx.x, partial, nb, k, sums, counter, cbuf, hcol, st, col, refine, eta,
                                    do_scale, safe, deftol, prescale_ok, halt_pass, coef1);
}

// The same finish without the ticket and its fences: workgroups 0..k publish
// their column sums as epoch-tagged words, and workgroup k+1 -- dispatched
// after them -- waits for the k+1 sums (bounded; KS_ERR_TIMEOUT in the status
// if a sum never comes), runs gs_finish and advances the epoch kept in
// pub[0] for the next launch (every workgroup has read it by then).
#define KS_FIN_SPIN (1 << 22)
template <bool FUSED, int NT>
__device__ __forceinline__ void gs_fin_pub_block(int b, const double *__restrict__ partial, int nb, int k,
                                                 unsigned long long *pub, double *cbuf, double *hcol, ks_status *st,
                                                 int col, int refine, double eta, int do_scale, int safe,
                                                 double deftol, int prescale_ok, int halt_pass,
                                                 double *coef1 = nullptr) {
  unsigned *hdr = (unsigned *)pub;   // [0] the epoch of the last launch
  const unsigned e_last = __hip_atomic_load(hdr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned epoch = e_last + 1u == 0u ? 1u : e_last + 1u;   // 0 is the cleared words' epoch
  unsigned long long *sumw = pub + 1;
  const int tid = threadIdx.x;
  if (b <= k) {
    const double s = gs_fin_colsum<NT>(b, partial, nb);
    if (tid == 0) ks_pub_store_f64(sumw + 2 * b, epoch, s);
    return;
  }
  __shared__ double fin[KS_NCOLP];
  __shared__ int late;
  if (tid == 0) late = 0;
  __syncthreads();
  for (int j = tid; j <= k; j += NT) {
    unsigned long long lo, hi;
    int spins = 0;
    for (;;) {
      lo = ks_pub_load(sumw + 2 * j);
      hi = ks_pub_load(sumw + 2 * j + 1);
      if ((unsigned)(lo >> 32) == epoch && (unsigned)(hi >> 32) == epoch) break;
      if (++spins > KS_FIN_SPIN) {
        late = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    fin[j] = ks_pub_f64(lo, hi);
  }
  __syncthreads();
  gs_finish<FUSED>(fin, k, cbuf, hcol, st, col, refine, eta, do_scale, safe, deftol, prescale_ok, halt_pass, coef1);
  if (tid == 0) {
    __hip_atomic_store(hdr, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (late) {
      st->error = KS_ERR_TIMEOUT;
      st->active = 0;
    }
  }
}
template <bool FUSED>
__global__ __launch_bounds__(KS_FIN_BLOCK) void k_gs_fin_pub(const double *__restrict__ partial, int nb, int k,
                                                            unsigned long long *pub, double *cbuf, double *hcol,
                                                            ks_status *st, int col, int pass, int refine, double eta,
                                                            int do_scale, int safe, double deftol, int prescale_ok,
                                                            int halt_pass, double *coef1) {
  if (!gs_pass_gate<FUSED>(st, col, pass)) return;
  gs_fin_pub_block<FUSED, KS_FIN_BLOCK>(blockIdx.x, partial, nb, k, pub, cbuf, hcol, st, col, refine, eta, do_scale,
                                        safe, deftol, prescale_ok, halt_pass, coef1);
}

template <typename T>
__device__ __forceinline__ void scale_status_block(int b, int nblk, T *__restrict__ w, int n, double s) {
  size_t i = (size_t)b * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)nblk * blockDim.x;
  for (; i < (size_t)n; i += stride) w[i] = (T)((typename acc_t<T>::type)w[i] * (typename acc_t<T>::type)s);
}

// the expansion stopped by a lean orthonormalization goes on (ks_bv_orthonormalize_resume)
__global__ void k_status_resume(ks_status *st) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && st->halt_col >= 0) {
    st->active = 1;
    st->halt_col = -1;
    st->halt_need = 0;
  }
}

template <typename T>
__global__ void k_scale_status(T *__restrict__ w, int n, const ks_status *__restrict__ st, int col, int *ran) {
  if (!(st->col == col && st->done)) return;
  const double s = st->scale;
  if (s == 1.0) return;
  if (ran && blockIdx.x == 0 && threadIdx.x == 0) *ran = 1;
  scale_status_block<T>(blockIdx.x, gridDim.x, w, n, s);
}

// ---------------------------------------------------------------- grouped Gram-Schmidt (serving)
// The sweeps, finishes and scalings of several independent solves, one launch
// per kind: workgroup b belongs to the member whose prefix range holds it and
// runs that member's workgroup with the member's own status (gates), column
// count, rows and reduction workspace.  The member bodies are the single-solve
// kernels' own.
__device__ __forceinline__ int grp_member(const int *b0, int nmem, int b) {
  int m = 0;
  while (m + 1 < nmem && b >= b0[m + 1]) m++;
  return m;
}

// each workgroup runs its member's sweep with the member's own column split, so
// a member's arithmetic does not depend on which other solves share the launch
template <bool FUSED>
__global__ __launch_bounds__(KS_BLOCK) void k_grp_gs(ks_grp_gs g, int pass) {
  const int m = grp_member(g.b0_gs, g.nmem, blockIdx.x);
  const ks_grp_gs_member &M = g.m[m];
  if (!gs_pass_gate<FUSED>(M.st, M.col, pass)) return;
  const int b = (int)blockIdx.x - g.b0_gs[m];
  const double pre = gs_prescale<FUSED>(M.st, M.col, pass);
  const int defer = gs_defer<FUSED>(M.st, M.col, pass);
  switch (M.cw) {
#define KS_GRP_CW(C)                                                                                                 \
  case C:                                                                                                            \
    gs_split_block<double, C, FUSED, 1, false>(b, M.nb, M.V, M.ld, M.k, M.w, M.n, M.rpb, M.cbuf, M.partial, pre,     \
                                               defer, M.coef1);                                                      \
    break;
    KS_GRP_CW(2)
    KS_GRP_CW(4)
    KS_GRP_CW(8)
    KS_GRP_CW(12)
    KS_GRP_CW(16)
    KS_GRP_CW(24)
    default:
      gs_split_block<double, 32, FUSED, 1, false>(b, M.nb, M.V, M.ld, M.k, M.w, M.n, M.rpb, M.cbuf, M.partial, pre,
                                                  defer, M.coef1);
#undef KS_GRP_CW
  }
}

template <bool FUSED>
__global__ __launch_bounds__(KS_BLOCK) void k_grp_fin(ks_grp_gs g, int pass, int refine, double eta, double deftol,
                                                     int halt_pass) {
  const int m = grp_member(g.b0_fin, g.nmem, blockIdx.x);
  const ks_grp_gs_member &M = g.m[m];
  if (!gs_pass_gate<FUSED>(M.st, M.col, pass)) return;
  const int b = (int)blockIdx.x - g.b0_fin[m];
  if (M.pub)
    gs_fin_pub_block<FUSED, KS_BLOCK>(b, M.partial, M.nb, M.k, M.pub, M.cbuf, M.hcol, M.st, M.col, refine, eta, 1, 0,
                                      deftol, 1, halt_pass, M.coef1);
  else
    gs_fin_block<FUSED, KS_BLOCK>(b, M.partial, M.nb, M.k, M.sums, M.counter, M.cbuf, M.hcol, M.st, M.col, refine,
                                  eta, 1, 0, deftol, 1, halt_pass, M.coef1);
}

__global__ void k_grp_scale(ks_grp_gs g) {
  const int m = grp_member(g.b0_sc, g.nmem, blockIdx.x);
  const ks_grp_gs_member &M = g.m[m];
  if (!(M.st->col == M.col && M.st->done)) return;
  const double s = M.st->scale;
  if (s == 1.0) return;
  scale_status_block<double>((int)blockIdx.x - g.b0_sc[m], g.b0_sc[m + 1] - g.b0_sc[m], M.w, M.n, s);
}

template <typename T>
__global__ void k_scale(T *__restrict__ x, int n, double alpha) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  using AT = typename acc_t<T>::type;
  if (alpha == 0.0)
    for (; i < (size_t)n; i += stride) x[i] = (T)0;
  else
    for (; i < (size_t)n; i += stride) x[i] = (T)((AT)x[i] * (AT)alpha);
}

template <typename T>
__global__ void k_axpy(T *__restrict__ y, const T *__restrict__ x, int n, double alpha) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  using AT = typename acc_t<T>::type;
  for (; i < (size_t)n; i += stride) y[i] = (T)((AT)y[i] + (AT)alpha * (AT)x[i]);
}

template <typename TX, typename TY>
__global__ void k_convert(TY *__restrict__ y, const TX *__restrict__ x, int n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < (size_t)n; i += stride) y[i] = (TY)x[i];
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

template <typename T>
__global__ void k_random(T *__restrict__ x, int n, uint64_t seed) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const uint64_t base = seed * 0x9E3779B97F4A7C15ULL;
  for (; i < (size_t)n; i += stride) {
    const uint64_t r = splitmix64(base + (uint64_t)i);
    x[i] = (T)(2.0 * ((double)(r >> 11) * 0x1.0p-53) - 1.0);
  }
}

template <typename T, int RPT>
__global__ __launch_bounds__(KS_BLOCK) void k_norm(const T *__restrict__ x, int n, double *partial) {
  __shared__ double red[KS_WAVES];
  const size_t row0 = (size_t)blockIdx.x * KS_BLOCK * RPT;
  // rows of the last workgroup past n read row n-1 (a valid address) and add nothing
  T v[RPT];
#pragma unroll
  for (int r = 0; r < RPT; r++) {
    const size_t i = row0 + (size_t)r * KS_BLOCK + threadIdx.x;
    v[r] = x[i < (size_t)n ? i : (size_t)n - 1];
  }
  double s = 0.0;
#pragma unroll
  for (int r = 0; r < RPT; r++) {
    const size_t i = row0 + (size_t)r * KS_BLOCK + threadIdx.x;
    const double d = i < (size_t)n ? (double)v[r] : 0.0;
    s += d * d;
  }
  const double bs = ks_block_sum(s, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = bs;
}

// the norm's finish: one workgroup adds the per-workgroup sums in a fixed order
// (a separate stream-ordered launch: no fences or atomics in the sweep)
__global__ __launch_bounds__(KS_BLOCK) void k_norm_fin(const double *__restrict__ partial, int nb, double *out) {
  __shared__ double red[KS_WAVES];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += KS_BLOCK) s += partial[b];
  const double tot = ks_block_sum(s, red);
  if (threadIdx.x == 0) out[0] = sqrt(tot);
}

// ---------------------------------------------------------------- restart GEMM
// V(:, s:s+N) = V(:, 0:K) * Q(0:K, s:s+N) on v_mfma_f64_16x16x4f64.
// Each wavefront owns R tiles of 16 rows and all N output columns (R x NT
// 16-wide tiles kept in the accumulators).  A fragments come straight from HBM
// (lane l: row l%16, k = l/16), all R x KS_MIP_KC/4 of a k-chunk issued before
// the first MFMA that uses them; Q chunks are staged through LDS once per
// workgroup for its 64 R rows.  A wave reads every input column of its rows
// before it stores any output, and rows are disjoint across waves, so the
// update is in place without a temporary.  Lanes past the last row or column
// read row n-1 / column K-1 (valid addresses, no per-lane branch); their
// products meet zero rows of the staged Q or land in rows that are not stored.
typedef double v4d __attribute__((ext_vector_type(4)));

template <typename T, int NT, int R>
__global__ __launch_bounds__(KS_BLOCK) void k_multinplace_mfma(T *__restrict__ V, long ld, int n, int K,
                                                              const double *__restrict__ Q, int ldq, int s,
                                                              int N) {
  __shared__ double sQ[KS_MIP_KC][NT * 16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int rbase = (blockIdx.x * KS_WAVES + wave) * (16 * R);   // this wave's first row
  const int kq = lane >> 4;                                       // k offset of this lane inside a 4-step
  const int nlast = n - 1;
  v4d acc[R][NT];
#pragma unroll
  for (int r = 0; r < R; r++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[r][t] = v4d{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < K; k0 += KS_MIP_KC) {
    __syncthreads();
    for (int idx = tid; idx < KS_MIP_KC * NT * 16; idx += KS_BLOCK) {
      const int kk = idx / (NT * 16), j = idx - kk * (NT * 16);
      sQ[kk][j] = (k0 + kk < K && j < N) ? Q[(size_t)(k0 + kk) + (size_t)(s + j) * ldq] : 0.0;
    }
    __syncthreads();
    double a[R][KS_MIP_KC / 4];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int row = min(rbase + r * 16 + (lane & 15), nlast);
#pragma unroll
      for (int kk = 0; kk < KS_MIP_KC / 4; kk++) {
        const int kc = min(k0 + 4 * kk + kq, K - 1);
        a[r][kk] = (double)ks_stream_load(V + (size_t)row + (size_t)kc * ld);   // V is read once
      }
    }
#pragma unroll
    for (int kk = 0; kk < KS_MIP_KC / 4; kk++) {
#pragma unroll
      for (int t = 0; t < NT; t++) {
        const double b = sQ[4 * kk + kq][t * 16 + (lane & 15)];
#pragma unroll
        for (int r = 0; r < R; r++) acc[r][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r][kk], b, acc[r][t], 0, 0, 0);
      }
    }
  }
  // D layout of v_mfma_f64_16x16x4f64 (measured on gfx950): register q of lane l
  // holds row l/16 + 4q, column l%16
  const int cold = lane & 15;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int rowd = rbase + r * 16 + (lane >> 4);
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int c = t * 16 + cold;
      if (c < N) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int row = rowd + 4 * q;
          if (row < n) V[(size_t)row + (size_t)(s + c) * ld] = (T)acc[r][t][q];
        }
      }
    }
  }
}

// k-chunk of the persistent restart GEMM (a multiple of 4)
#ifndef KS_MIP_LDS_KC
#define KS_MIP_LDS_KC(R) 16
#endif
// The same product as a persistent kernel: each workgroup stages all of
// Q(0:K, s:s+N) in LDS once and then walks row tiles (4 waves x 16R rows)
// blockIdx.x, blockIdx.x + gridDim.x, ...; inside a tile the A fragments of
// the next k-chunk are loaded while the MFMAs of the current one run, the
// first chunk of the next tile is loaded before this tile's stores, and the
// output goes through a per-wave LDS tile so that each store instruction
// writes whole column runs (16R rows of 4/R columns: 512-byte runs for R = 4)
// instead of 32-byte pieces.  The MFMA sequence of a tile, and so every
// rounding, is the one of k_multinplace_mfma.
template <typename T, int NT, int R>
__global__ __launch_bounds__(KS_BLOCK) void k_multinplace_lds(T *__restrict__ V, long ld, int n, int K,
                                                             const double *__restrict__ Q, int ldq, int s, int N,
                                                             int Kp, int ntiles) {
  extern __shared__ double smem[];
  constexpr int WR = 16 * R;        // rows of a wave
  constexpr int SS = WR + 1;        // padded column stride of the output tile
  constexpr int JW = NT * 16;
  constexpr int KC = KS_MIP_LDS_KC(R);   // k-chunk: R x KC/4 A fragments per lane in flight
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double *sQ = smem;                                                // Kp x JW, [k][j]
  double *tile = smem + (size_t)Kp * JW + (size_t)wave * 16 * SS;   // 16 columns x SS
  const int kq = lane >> 4;
  const int nlast = n - 1;
  for (int idx = tid; idx < Kp * JW; idx += KS_BLOCK) {
    const int kk = idx / JW, j = idx - kk * JW;
    sQ[idx] = (kk < K && j < N) ? Q[(size_t)kk + (size_t)(s + j) * ldq] : 0.0;
  }
  __syncthreads();
  double a[R][KC / 4], an[R][KC / 4];
  auto load = [&](double (&dst)[R][KC / 4], int rbase, int k0) {
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int row = min(rbase + r * 16 + (lane & 15), nlast);
#pragma unroll
      for (int kk = 0; kk < KC / 4; kk++) {
        const int kc = min(k0 + 4 * kk + kq, K - 1);
        dst[r][kk] = (double)ks_stream_load(V + (size_t)row + (size_t)kc * ld);
      }
    }
  };
  constexpr int CPI = 64 / WR;       // columns one store instruction covers
  const int srow = lane % WR, scol = lane / WR;
  int t0 = blockIdx.x;
  if (t0 < ntiles && K > 0) load(a, (t0 * KS_WAVES + wave) * WR, 0);
  for (; t0 < ntiles; t0 += gridDim.x) {
    const int rbase = (t0 * KS_WAVES + wave) * WR;
    v4d acc[R][NT];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
      for (int t = 0; t < NT; t++) acc[r][t] = v4d{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += KC) {
      const bool more = k0 + KC < K;
      if (more) load(an, rbase, k0 + KC);
#pragma unroll
      for (int kk = 0; kk < KC / 4; kk++) {
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const double b = sQ[(size_t)(k0 + 4 * kk + kq) * JW + t * 16 + (lane & 15)];
#pragma unroll
          for (int r = 0; r < R; r++) acc[r][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r][kk], b, acc[r][t], 0, 0, 0);
        }
      }
      if (more) {
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
          for (int kk = 0; kk < KC / 4; kk++) a[r][kk] = an[r][kk];
      }
    }
    // the next tile's first chunk is in flight while this tile is stored; this
    // wave has read every input of its rows, and other waves own other rows
    const int tn = t0 + gridDim.x;
    if (tn < ntiles && K > 0) load(a, (tn * KS_WAVES + wave) * WR, 0);
    // register q of lane l holds row l/16 + 4q, column l%16 of a 16 x 16 tile
#pragma unroll
    for (int t = 0; t < NT; t++) {
#pragma unroll
      for (int r = 0; r < R; r++)
#pragma unroll
        for (int q = 0; q < 4; q++) tile[(lane & 15) * SS + r * 16 + (lane >> 4) + 4 * q] = acc[r][t][q];
      __syncthreads();
      const int row = rbase + srow;
#pragma unroll
      for (int c0 = 0; c0 < 16; c0 += CPI) {
        const int c = t * 16 + c0 + scol;
        if (row < n && c < N) V[(size_t)row + (size_t)(s + c) * ld] = (T)tile[(c0 + scol) * SS + srow];
      }
      __syncthreads();
    }
  }
}

// row tiles per wavefront: more rows per staged Q chunk while the accumulators
// (R * NT * 4 doubles per lane) stay within 64 registers' worth
// workgroups of the persistent restart GEMM (two per CU resident at a time)
#ifndef KS_MIP_GRID
#define KS_MIP_GRID 1024
#endif
// Q fully in LDS while it takes at most this many bytes (with the output tiles):
// two workgroups per CU (160 KB of LDS) at the largest
#ifndef KS_MIP_LDS_MAX
#define KS_MIP_LDS_MAX (80 * 1024)
#endif

template <typename T, int NT>
void launch_mip(ks_ctx ctx, T *V, long ld, int n, int K, const double *Q, int ldq, int s, int N) {
  constexpr int R = NT <= 2 ? 4 : (NT <= 4 ? 2 : 1);
  const int rows = KS_WAVES * 16 * R;
  dim3 grid((n + rows - 1) / rows), block(KS_BLOCK);
  constexpr int KC = KS_MIP_LDS_KC(R);
  const int Kp = (K + KC - 1) / KC * KC;
  const size_t lds = sizeof(double) * ((size_t)Kp * NT * 16 + (size_t)KS_WAVES * 16 * (16 * R + 1));
  const int ntiles = (int)grid.x;
  if (lds <= KS_MIP_LDS_MAX)
    hipLaunchKernelGGL((k_multinplace_lds<T, NT, R>), dim3(ntiles < KS_MIP_GRID ? ntiles : KS_MIP_GRID), block, lds,
                       ctx->stream, V, ld, n, K, Q, ldq, s, N, Kp, ntiles);
  else
    hipLaunchKernelGGL((k_multinplace_mfma<T, NT, R>), grid, block, 0, ctx->stream, V, ld, n, K, Q, ldq, s, N);
}

template <typename T>
int dispatch_mip(ks_ctx ctx, T *V, long ld, int n, int K, const double *Q, int ldq, int s, int N) {
  const int nt = (N + 15) / 16;
  switch (nt) {
#define KS_MIP_CASE(X) case X: launch_mip<T, X>(ctx, V, ld, n, K, Q, ldq, s, N); break;
    KS_MIP_CASE(1) KS_MIP_CASE(2) KS_MIP_CASE(3) KS_MIP_CASE(4) KS_MIP_CASE(5)
    KS_MIP_CASE(6) KS_MIP_CASE(7) KS_MIP_CASE(8) KS_MIP_CASE(9) KS_MIP_CASE(10)
    KS_MIP_CASE(11) KS_MIP_CASE(12) KS_MIP_CASE(13) KS_MIP_CASE(14) KS_MIP_CASE(15)
    KS_MIP_CASE(16) KS_MIP_CASE(17) KS_MIP_CASE(18) KS_MIP_CASE(19) KS_MIP_CASE(20)
#undef KS_MIP_CASE
    default: return KS_ERR_RANGE;
  }
  return KS_OK;
}

// ---------------------------------------------------------------- block dot (BVDot)
// G = Y^T X for tall-skinny X (n x p) and Y (n x q), one 64 x 64 block of G
// (rows i0.., columns j0..) per launch.  Workgroup b owns a contiguous range of
// rows; per 32-row chunk it stages the 32 x 64 slices of X and Y in LDS with
// coalesced column segments (rows past n and columns past p or q read a valid
// element and are zeroed by a select, not a branch), and thread (ti, tj) of a
// 16 x 16 grid accumulates the 4 x 4 entries G(i0+4ti.., j0+4tj..).  Partial
// blocks go to partial[b][64*64]; k_gram_fin adds them in workgroup order.
#define KS_GRAM_ROWS 32
#define KS_GRAM_MAXB 512
template <typename T>
__global__ __launch_bounds__(KS_BLOCK) void k_gram(const T *__restrict__ X, long ldx, int p, const T *__restrict__ Y,
                                                  long ldy, int q, int n, int rpb, int i0, int j0,
                                                  double *__restrict__ partial) {
  __shared__ double sX[KS_GRAM_ROWS][64];
  __shared__ double sY[KS_GRAM_ROWS][64];
  const int tid = threadIdx.x;
  const int ti = tid >> 4, tj = tid & 15;
  const int lr = tid & (KS_GRAM_ROWS - 1), lc = tid / KS_GRAM_ROWS;   // loader: row, first column
  const long r0 = (long)blockIdx.x * rpb;
  const long r1 = min((long)n, r0 + rpb);
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[a][c] = 0.0;
  for (long rb = r0; rb < r1; rb += KS_GRAM_ROWS) {
    const long row = rb + lr;
    const bool rok = row < r1;
    const long rowc = rok ? row : r1 - 1;
    double vx[8], vy[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int c = lc + 8 * u;
      const int cx = min(j0 + c, p - 1), cy = min(i0 + c, q - 1);
      vx[u] = (double)X[(size_t)cx * ldx + rowc];
      vy[u] = (double)Y[(size_t)cy * ldy + rowc];
    }
    __builtin_amdgcn_sched_barrier(0);
    __syncthreads();   // the previous chunk's reads are done
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int c = lc + 8 * u;
      sX[lr][c] = (rok && j0 + c < p) ? vx[u] : 0.0;
      sY[lr][c] = (rok && i0 + c < q) ? vy[u] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int r = 0; r < KS_GRAM_ROWS; r++) {
      double a[4], bb[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        a[u] = sY[r][4 * ti + u];
        bb[u] = sX[r][4 * tj + u];
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) acc[u][v] += a[u] * bb[v];
    }
  }
  double *out = partial + (size_t)blockIdx.x * 4096;
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int v = 0; v < 4; v++) out[(4 * ti + u) * 64 + 4 * tj + v] = acc[u][v];
}

// G(i0 + i, j0 + j) = sum over the nb partial blocks, in workgroup order
__global__ __launch_bounds__(KS_BLOCK) void k_gram_fin(const double *__restrict__ partial, int nb, int qb, int pb,
                                                      int i0, int j0, double *__restrict__ G, int ldg) {
  const int e = blockIdx.x * KS_BLOCK + threadIdx.x;   // entry of the 64 x 64 block
  if (e >= 4096) return;
  const int i = e >> 6, j = e & 63;
  double s = 0.0;
  for (int b = 0; b < nb; b++) s += partial[(size_t)b * 4096 + e];
  if (i < qb && j < pb) G[(size_t)(i0 + i) + (size_t)(j0 + j) * ldg] = s;
}

template <typename T>
int gram_t(ks_ctx ctx, int n, const T *X, long ldx, int p, const T *Y, long ldy, int q, double *Gdev, int ldg) {
  int nb = 0, rpb = 0;   // n == 0: no partial blocks, the finisher writes zeros
  if (n > 0) {
    nb = (n + KS_GRAM_ROWS - 1) / KS_GRAM_ROWS;
    if (nb > KS_GRAM_MAXB) nb = KS_GRAM_MAXB;
    rpb = ((n + nb - 1) / nb + KS_GRAM_ROWS - 1) / KS_GRAM_ROWS * KS_GRAM_ROWS;
    nb = (n + rpb - 1) / rpb;
  }
  static_assert((size_t)KS_GRAM_MAXB * 4096 <= (size_t)KS_MAXB * KS_NCOLP, "Gram partials exceed the workspace");
  for (int i0 = 0; i0 < q; i0 += 64)
    for (int j0 = 0; j0 < p; j0 += 64) {
      if (nb > 0)
        hipLaunchKernelGGL((k_gram<T>), dim3(nb), dim3(KS_BLOCK), 0, ctx->stream, X, ldx, p, Y, ldy, q, n, rpb, i0,
                           j0, ctx->d_partial);
      hipLaunchKernelGGL(k_gram_fin, dim3(4096 / KS_BLOCK), dim3(KS_BLOCK), 0, ctx->stream,
                         (const double *)ctx->d_partial, nb, min(64, q - i0), min(64, p - j0), i0, j0, Gdev, ldg);
    }
  KS_HIP(hipGetLastError());
  return KS_OK;
}

// ---------------------------------------------------------------- launch helpers
inline int ew_blocks(int n) {
  int b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return b;
}

template <typename T, int MODE>
void launch_dot(ks_ctx ctx, int rpt, int nb, const T *V, long ld, int k, const T *w, int n, double *cout,
                double *hcol, int col, int pass, int *ran) {
  dim3 grid(nb), block(KS_BLOCK);
  const double deftol = 10.0 * 2.220446049250313e-16;
#define KS_DOT_CASE(R)                                                                                   \
  case R:                                                                                                \
    hipLaunchKernelGGL((k_gs_dot<T, R, MODE>), grid, block, 0, ctx->stream, V, ld, k, w, n, nb,          \
                       ctx->d_partial, ctx->d_gpartial, ctx->d_counters, ctx->d_status, cout, hcol, col, \
                       pass, ran, deftol);                                                               \
    break;
  switch (rpt) { KS_DOT_CASE(1) KS_DOT_CASE(2) KS_DOT_CASE(4) KS_DOT_CASE(8) KS_DOT_CASE(16) }
#undef KS_DOT_CASE
}

template <typename T, int MODE>
void launch_upd(ks_ctx ctx, int rpt, int nb, const T *V, long ld, int k, T *w, int n, const double *cin,
                double alpha, double beta, double *hcol, int col, int pass, int refine, double eta, int do_scale,
                int safe, int *ran) {
  dim3 grid(nb), block(KS_BLOCK);
#define KS_UPD_CASE(R)                                                                                      \
  case R:                                                                                                   \
    hipLaunchKernelGGL((k_gs_update<T, R, MODE>), grid, block, 0, ctx->stream, V, ld, k, w, n, nb, cin,     \
                       alpha, beta, ctx->d_partial, ctx->d_gpartial, ctx->d_counters, ctx->d_status, hcol,  \
                       col, pass, refine, eta, do_scale, safe, ran);                                        \
    break;
  switch (rpt) { KS_UPD_CASE(1) KS_UPD_CASE(2) KS_UPD_CASE(4) KS_UPD_CASE(8) KS_UPD_CASE(16) }
#undef KS_UPD_CASE
}

int split_checked(int n, int *rpt, int *nb) {
  ks_rows_split(n, rpt, nb);
  KS_CHECK(*nb <= KS_MAXB, KS_ERR_RANGE, "vector length %d exceeds the reduction capacity", n);
  return KS_OK;
}

// geometry of the LDS-staged Gram-Schmidt kernels for k columns
struct gs_geom {
  int R, tsub, nb, cpb;   // cpb: LDS kernel sub-chunks per workgroup; register kernel rows per workgroup
  size_t lds;
  int kc;   // > 0: register-resident kernel for up to kc columns
  int cw;   // > 0: column-split kernel, cw columns per wavefront
  int pair; // column-split on FP32 data with two rows per lane (even ld, aligned)
  int rpl;  // column-split on FP64 data: rows per lane per step (2: two loads 64 rows apart)
};

int gs_geometry(int n, int k, size_t esize, gs_geom *g) {
  g->kc = 0;
  g->cw = 0;
  g->pair = 0;
  g->rpl = 1;
  if (k <= KS_GS_REGMAX && (uint64_t)n * esize < (1ull << 32)) {
    g->kc = k <= 8 ? 8 : (k + 7) / 8 * 8;
    const long tiles = (n + (long)KS_BLOCK - 1) / KS_BLOCK;            // 256-row tiles
    const long nb = tiles < KS_GS_REG_BLOCKS ? tiles : KS_GS_REG_BLOCKS;
    const long tpb = (tiles + nb - 1) / nb;                             // tiles per workgroup
    g->cpb = (int)(tpb * KS_BLOCK);                                     // rows per workgroup
    g->nb = (int)((n + (long)g->cpb - 1) / g->cpb);
    g->R = 0;
    g->tsub = 0;
    g->lds = 0;
    return KS_OK;
  }
  if (k <= KS_GS_SPLITMAX && (uint64_t)n * esize < (1ull << 32)) {
    const int per = (k + KS_WAVES - 1) / KS_WAVES;
    static const int cws[] = {2, 4, 8, 12, 16, 24, 32};
    for (int c : cws)
      if (per <= c) { g->cw = c; break; }
    // FP64 with few columns per wave: two rows per lane per step, so a wave
    // keeps 2 CW loads in flight (FP32 takes two rows per lane as a pair)
    if (esize == 8 && g->cw <= KS_GS_RPL2_MAXCW) g->rpl = 2;
    const long step = esize == 4 ? 128 : 64 * g->rpl;                   // rows per workgroup step, even
    const long tiles = (n + step - 1) / step;
    const long cap = esize == 4 ? KS_GS_SPLIT_BLOCKS32 : KS_GS_SPLIT_BLOCKS;
    const long nb = tiles < cap ? tiles : cap;
    const long tpb = (tiles + nb - 1) / nb;
    g->cpb = (int)(tpb * step);                                         // rows per workgroup
    g->nb = (int)((n + (long)g->cpb - 1) / g->cpb);
    g->R = 0;
    g->tsub = 0;
    g->lds = 0;
    return KS_OK;
  }
  int t = 8;
  while (t > 1 && (k + 1) * t > KS_BLOCK) t >>= 1;
  g->tsub = t;
  g->R = 0;
  for (int R = 256; R >= 16; R >>= 1) {
    const size_t lds = sizeof(double) * ((size_t)(k + 1) * (R + t) + k + KS_BLOCK);
    if (lds <= 60 * 1024) {
      g->R = R;
      g->lds = lds;
      break;
    }
  }
  KS_CHECK(g->R, KS_ERR_RANGE, "%d columns do not fit the LDS tile", k);
  const long chunks = (n + (long)g->R - 1) / g->R;
  long nb = chunks < 2048 ? chunks : 2048;
  g->cpb = (int)((chunks + nb - 1) / nb);
  g->nb = (int)((chunks + g->cpb - 1) / g->cpb);
  KS_CHECK(g->nb <= KS_MAXB, KS_ERR_RANGE, "vector length %d exceeds the reduction capacity", n);
  return KS_OK;
}

template <typename T, bool FUSED, bool RF>
void launch_split(ks_ctx ctx, const gs_geom &g, const T *V, long ld, int k, T *w, int n, const double *cbuf, int col,
                  int pass, int *ran) {
  dim3 grid(g.nb), block(KS_BLOCK);
#define KS_GSS_CASE(CC)                                                                                           \
  case CC:                                                                                                        \
    if (g.pair)                                                                                                   \
      hipLaunchKernelGGL((k_gs_split<T, CC, FUSED, 2, sizeof(T) == 4, RF>), grid, block, 0, ctx->stream, V, ld, k, w,  \
                         n, g.cpb, cbuf, ctx->d_partial, ctx->d_status, col, pass, ran, ctx->d_coef1);                          \
    else if (sizeof(T) == 8 && g.rpl == 2 && RF && CC == 12)                                                       \
      hipLaunchKernelGGL((k_gs_split_w4<T, CC, FUSED, 2, false, RF>), grid, block, 0, ctx->stream, V, ld, k, w, n,     \
                         g.cpb, cbuf, ctx->d_partial, ctx->d_status, col, pass, ran, ctx->d_coef1);                 \
    else if (sizeof(T) == 8 && g.rpl == 2)                                                                        \
      hipLaunchKernelGGL((k_gs_split<T, CC, FUSED, 2, false, RF>), grid, block, 0, ctx->stream, V, ld, k, w, n, g.cpb, \
                         cbuf, ctx->d_partial, ctx->d_status, col, pass, ran, ctx->d_coef1);                                    \
    else                                                                                                          \
      hipLaunchKernelGGL((k_gs_split<T, CC, FUSED, 1, false, RF>), grid, block, 0, ctx->stream, V, ld, k, w, n, g.cpb, \
                         cbuf, ctx->d_partial, ctx->d_status, col, pass, ran, ctx->d_coef1);                                    \
    break;
    switch (g.cw) { KS_GSS_CASE(2) KS_GSS_CASE(4) KS_GSS_CASE(8) KS_GSS_CASE(12) KS_GSS_CASE(16) KS_GSS_CASE(24)
                    KS_GSS_CASE(32) }
#undef KS_GSS_CASE
}

template <typename T, bool FUSED>
void launch_gs(ks_ctx ctx, const gs_geom &g, const T *V, long ld, int k, T *w, int n, const double *cbuf, int col,
               int pass, int *ran) {
  dim3 grid(g.nb), block(KS_BLOCK);
  if (g.cw) {
    if (FUSED && pass == 2)   // may follow a deferred first pass
      launch_split<T, FUSED, true>(ctx, g, V, ld, k, w, n, cbuf, col, pass, ran);
    else
      launch_split<T, FUSED, false>(ctx, g, V, ld, k, w, n, cbuf, col, pass, ran);
    return;
  }
  if (g.kc) {
#define KS_GSR_CASE(KK)                                                                                           \
  case KK:                                                                                                        \
    hipLaunchKernelGGL((k_gs_reg<T, KK, FUSED>), grid, block, 0, ctx->stream, V, ld, k, w, n, g.cpb, cbuf,         \
                       ctx->d_partial, ctx->d_status, col, pass, ran);                                             \
    break;
    switch (g.kc) { KS_GSR_CASE(8) KS_GSR_CASE(16) KS_GSR_CASE(24) KS_GSR_CASE(32) KS_GSR_CASE(40) KS_GSR_CASE(48) }
#undef KS_GSR_CASE
    return;
  }
#define KS_GS_CASE(RR)                                                                                            \
  case RR:                                                                                                        \
    hipLaunchKernelGGL((k_gs_lds<T, RR, FUSED>), grid, block, g.lds, ctx->stream, V, ld, k, w, n, g.cpb, g.tsub,  \
                       cbuf, ctx->d_partial, ctx->d_status, col, pass, ran);                                       \
    break;
  switch (g.R) { KS_GS_CASE(256) KS_GS_CASE(128) KS_GS_CASE(64) KS_GS_CASE(32) KS_GS_CASE(16) }
#undef KS_GS_CASE
}

template <bool FUSED>
void launch_gs_fin(ks_ctx ctx, const gs_geom &g, int k, double *cbuf, double *hcol, int col, int pass, int refine,
                   double eta, int do_scale, int safe, int halt_pass) {
  const double deftol = 10.0 * 2.220446049250313e-16;
  ks_prof_begin(ctx, KS_PROF_GS_FIN, (double)g.nb * (k + 1) * 8.0, (double)g.nb * (k + 1), false);
  double *coef1 = g.cw ? ctx->d_coef1 : nullptr;   // deferred first-pass store: split sweeps only
  if (ctx->d_finpub)
    hipLaunchKernelGGL((k_gs_fin_pub<FUSED>), dim3(k + 2), dim3(KS_FIN_BLOCK), 0, ctx->stream, ctx->d_partial, g.nb,
                       k, ctx->d_finpub, cbuf, hcol, ctx->d_status, col, pass, refine, eta, do_scale, safe, deftol,
                       g.cw != 0 || g.kc != 0, halt_pass, coef1);
  else
    hipLaunchKernelGGL((k_gs_fin<FUSED>), dim3(k + 1), dim3(KS_FIN_BLOCK), 0, ctx->stream, ctx->d_partial, g.nb, k,
                       ctx->d_gpartial, ctx->d_counters, cbuf, hcol, ctx->d_status, col, pass, refine, eta, do_scale,
                       safe, deftol, g.cw != 0 || g.kc != 0, halt_pass, coef1);
  ks_prof_end(ctx);
}

// mode 0: the whole sequence (first-pass dots, every DGKS pass the refinement
// allows, the scaling), each launch gated on the device status.  mode 1 (lean):
// the passes a column normally needs and no scaling launch; the finish stops
// the expansion when the column needs more (see gs_finish).  mode 2 (resume):
// the launches mode 1 left out, after re-activating the status.
template <typename T>
int orthonormalize_t(ks_ctx ctx, int n, T *V, long ld, int k, double *buf, int ldbuf, int hcol_idx, int col,
                     ks_refine refine, double eta, int do_scale, int safe, int mode = 0) {
  gs_geom g;
  int e;
  if ((e = gs_geometry(n, k, sizeof(T), &g))) return e;
  g.pair = g.cw && sizeof(T) == 4 && ld % 2 == 0 && ((uintptr_t)V % 8) == 0;
  const double s = sizeof(T);
  T *w = V + (size_t)k * ld;
  double *cbuf = buf;                              // buffer column 0: scratch coefficients
  double *hcol = buf + (size_t)hcol_idx * ldbuf;   // buffer column hcol_idx: accumulated h
  const int passes = safe || refine == KS_REFINE_NEVER ? 1 : (refine == KS_REFINE_ALWAYS ? 2 : 3);
  // the lean sequence needs the sweeps that store the normalized column
  const bool lean_ok = !safe && do_scale && refine != KS_REFINE_NEVER && (g.cw || g.kc);
  if (mode && !lean_ok) mode = mode == 1 ? 0 : -1;   // full sequence, or nothing left to resume
  if (mode < 0) return KS_OK;
  const int halt_pass = mode == 1 ? 2 : 0;
  const int p_first = mode == 2 ? 3 : 1, p_last = mode == 1 ? 2 : passes;
  int *ran;
  if (mode == 2) hipLaunchKernelGGL(k_status_resume, dim3(1), dim3(64), 0, ctx->stream, ctx->d_status);
  if (mode != 2) {
    // first pass dots: c = V^T w and w^T w
    ran = ks_prof_begin(ctx, KS_PROF_GS_DOT1, (k + 1.0) * n * s, 2.0 * (k + 1.0) * n, true);
    launch_gs<T, false>(ctx, g, V, ld, k, w, n, cbuf, col, 0, ran);
    ks_prof_end(ctx);
    launch_gs_fin<false>(ctx, g, k, cbuf, hcol, col, 0, refine, eta, do_scale, safe, 0);
  }
  // each pass: update with its c and the next pass's dots in the same sweep
  // with the first pass's store deferrable (split sweeps, a refining pass
  // follows) that pass is counted without its write -- it stores only in the
  // rare column it is also the last pass of -- and the next one with the
  // re-formed product's flops
  const bool deferrable = g.cw && ctx->d_coef1 && !safe && refine != KS_REFINE_NEVER;
  for (int p = p_first; p <= p_last; p++) {
    const double bytes = (p == 1 && deferrable ? k + 1.0 : k + 2.0) * n * s;
    const double flops = 2.0 * k * n + 2.0 * (k + 1.0) * n + (p == 2 && deferrable ? 2.0 * k * n : 0.0);
    ran = ks_prof_begin(ctx, p == 1 ? KS_PROF_GS_UPD1 : KS_PROF_GS_UPDR, bytes, flops, true);
    launch_gs<T, true>(ctx, g, V, ld, k, w, n, cbuf, col, p, ran);
    ks_prof_end(ctx);
    launch_gs_fin<true>(ctx, g, k, cbuf, hcol, col, p, refine, eta, do_scale, safe, halt_pass);
  }
  if (do_scale && mode != 1) {
    int *ran = ks_prof_begin(ctx, KS_PROF_SCALE, 2.0 * n * s, (double)n, true);
    hipLaunchKernelGGL((k_scale_status<T>), dim3(ew_blocks(n)), dim3(256), 0, ctx->stream, w, n, ctx->d_status, col,
                       ran);
    ks_prof_end(ctx);
  }
  KS_HIP(hipGetLastError());
  return KS_OK;
}

// ---------------------------------------------------------------- allocation-free orthonormalization
// ks_bv_orthonormalize_column_noalloc runs the kernels of
// ks_bv_orthonormalize_column with their workspace in the file-scope array
// `scratch` instead of a context's: the DGKS status, the finishing counter,
// the k+1 column sums and the (k+1) x nb partial sums of one sweep.
#define KS_NOALLOC_MAXK KS_GS_SPLITMAX                 // columns orthogonalized against, at most
#define KS_NOALLOC_STATUS 0                            // offsets in doubles, 256-byte aligned
#define KS_NOALLOC_COUNTER 32
#define KS_NOALLOC_SUMS 64
#define KS_NOALLOC_PARTIAL 448
#define KS_NOALLOC_PARTIAL_CAP ((size_t)(KS_NOALLOC_MAXK + 1) * 3072)   // partial sums of up to 3072 workgroups
#define KS_NOALLOC_DOUBLES (KS_NOALLOC_PARTIAL + KS_NOALLOC_PARTIAL_CAP)
static_assert(sizeof(ks_status) <= 32 * sizeof(double), "status slot too small");
static_assert(KS_NOALLOC_SUMS + KS_NCOLP <= KS_NOALLOC_PARTIAL, "column-sum slot too small");

// the status of a fresh column and a zero finishing counter (fill kernel, no memset node)
__global__ void k_noalloc_reset(ks_status *st, unsigned *counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->active = 1;
    st->breakdown_col = -1;
    st->col = -1;
    st->pass_done = 0;
    st->done = 0;
    st->lindep = 0;
    st->error = 0;
    st->need_norm = 0;
    st->total_passes = 0;
    st->onrm = 0.0;
    st->nrm = 0.0;
    st->beta_last = 0.0;
    st->scale = 1.0;
    st->sumsq = 0.0;
    st->exact_sq = 0.0;
    st->prescale = 1.0;
    st->prescale_col = -1;
    st->prescale_pass = 0;
    st->halt_col = -1;
    st->halt_need = 0;
    st->defer_col = -1;
    *counter = 0u;
  }
}

__global__ void k_noalloc_status_out(const ks_status *st, ks_status *out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = *st;
}

}  // namespace

// 3.2 MB per device, held until the process ends.  Host code never passes the
// array's name to a kernel: it uses the device address that hipGetSymbolAddress
// returns for the current device.
static __device__ __align__(256) double scratch[KS_NOALLOC_DOUBLES];

namespace {

#define KS_NOALLOC_MAXDEV 64
std::mutex g_noalloc_mu;
double *g_noalloc_ptr[KS_NOALLOC_MAXDEV];

// device address of `scratch` on the current device, looked up once per device
int noalloc_scratch(double **out) {
  int dev = -1;
  KS_HIP(hipGetDevice(&dev));
  KS_CHECK(dev >= 0 && dev < KS_NOALLOC_MAXDEV, KS_ERR_ARG, "device %d beyond the scratch table", dev);
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void *p = nullptr;
    KS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    KS_CHECK(p, KS_ERR_HIP, "no device address for the scratch array on device %d", dev);
    g_noalloc_ptr[dev] = (double *)p;
  }
  *out = g_noalloc_ptr[dev];
  return KS_OK;
}

}  // namespace

int ks_ew_blocks(int n) { return ew_blocks(n); }

int ks_gs_split_geometry(int n, int k, int *cw, int *rpb, int *nb) {
  gs_geom g;
  int e = gs_geometry(n, k, sizeof(double), &g);
  if (e) return e;
  if (!g.cw) {
    // the grouped launches run the column-split sweep only; its geometry for this k
    KS_CHECK(k <= KS_GS_SPLITMAX && (uint64_t)n * sizeof(double) < (1ull << 32), KS_ERR_RANGE,
             "grouped Gram-Schmidt supports up to %d columns (asked %d)", KS_GS_SPLITMAX, k);
    const int per = (k + KS_WAVES - 1) / KS_WAVES;
    static const int cws[] = {2, 4, 8, 12, 16, 24, 32};
    for (int c : cws)
      if (per <= c) { g.cw = c; break; }
    const long tiles = (n + 63L) / 64;
    const long nbl = tiles < KS_GS_SPLIT_BLOCKS ? tiles : KS_GS_SPLIT_BLOCKS;
    g.cpb = (int)(((tiles + nbl - 1) / nbl) * 64);
    g.nb = (int)((n + (long)g.cpb - 1) / g.cpb);
  }
  *cw = g.cw;
  *rpb = g.cpb;
  *nb = g.nb;
  return KS_OK;
}

// The lean sequence of the single-solve path (with refinement on): the first
// pass, then two fused passes whose last sweep stores the normalized column; a
// member that needs a third pass or a separate scaling stops its expansion there
// (the finish of pass 2 halts it, later gated launches exit), and BVMatKrylovEnd
// completes that column and runs the rest of its expansion on its own stream.
// Seven launches a grouped step with the SpMV instead of ten.
void ks_group_launch_gs(hipStream_t stream, const ks_grp_gs &g, int refine, double eta) {
  const double deftol = 10.0 * 2.220446049250313e-16;
  const bool lean = refine != KS_REFINE_NEVER;
  const int passes = refine == KS_REFINE_NEVER ? 1 : (refine == KS_REFINE_ALWAYS ? 2 : 3);
  const int last = lean ? 2 : passes, halt_pass = lean ? 2 : 0;
  const dim3 block(KS_BLOCK), ggs(g.b0_gs[g.nmem]), gfin(g.b0_fin[g.nmem]), gsc(g.b0_sc[g.nmem]);
  hipLaunchKernelGGL(k_grp_gs<false>, ggs, block, 0, stream, g, 0);
  hipLaunchKernelGGL((k_grp_fin<false>), gfin, block, 0, stream, g, 0, refine, eta, deftol, 0);
  for (int p = 1; p <= last; p++) {
    hipLaunchKernelGGL(k_grp_gs<true>, ggs, block, 0, stream, g, p);
    hipLaunchKernelGGL((k_grp_fin<true>), gfin, block, 0, stream, g, p, refine, eta, deftol, halt_pass);
  }
  if (!lean) hipLaunchKernelGGL(k_grp_scale, gsc, dim3(256), 0, stream, g);
}

extern "C" {

int ks_bv_orthonormalize_column(ks_ctx ctx, ks_prec prec, int n, void *V, int ld, int k, double *d_buf, int ldbuf,
                                int hcol, int col, ks_refine refine, double eta, int scale, int safe) {
  KS_CHECK(ctx && V && d_buf, KS_ERR_ARG, "null orthonormalize argument");
  KS_CHECK(n > 0 && ld >= n && k >= 0 && k < KS_MAX_COLS && ldbuf > k && hcol >= 0, KS_ERR_ARG,
           "bad orthonormalize sizes n=%d ld=%d k=%d ldbuf=%d", n, ld, k, ldbuf);
  KS_CHECK(refine >= KS_REFINE_IFNEEDED && refine <= KS_REFINE_ALWAYS, KS_ERR_ARG, "bad refinement type");
  if (prec == KS_F64)
    return orthonormalize_t<double>(ctx, n, (double *)V, ld, k, d_buf, ldbuf, hcol, col, refine, eta, scale, safe);
  return orthonormalize_t<float>(ctx, n, (float *)V, ld, k, d_buf, ldbuf, hcol, col, refine, eta, scale, safe);
}

int ks_bv_orthonormalize_column_lean(ks_ctx ctx, ks_prec prec, int n, void *V, int ld, int k, double *d_buf,
                                     int ldbuf, int hcol, int col, ks_refine refine, double eta, int scale, int safe) {
  KS_CHECK(ctx && V && d_buf, KS_ERR_ARG, "null orthonormalize argument");
  KS_CHECK(n > 0 && ld >= n && k >= 0 && k < KS_MAX_COLS && ldbuf > k && hcol >= 0, KS_ERR_ARG,
           "bad orthonormalize sizes n=%d ld=%d k=%d ldbuf=%d", n, ld, k, ldbuf);
  KS_CHECK(refine >= KS_REFINE_IFNEEDED && refine <= KS_REFINE_ALWAYS, KS_ERR_ARG, "bad refinement type");
  if (prec == KS_F64)
    return orthonormalize_t<double>(ctx, n, (double *)V, ld, k, d_buf, ldbuf, hcol, col, refine, eta, scale, safe, 1);
  return orthonormalize_t<float>(ctx, n, (float *)V, ld, k, d_buf, ldbuf, hcol, col, refine, eta, scale, safe, 1);
}

int ks_bv_orthonormalize_resume(ks_ctx ctx, ks_prec prec, int n, void *V, int ld, int k, double *d_buf, int ldbuf,
                                int hcol, int col, ks_refine refine, double eta) {
  KS_CHECK(ctx && V && d_buf, KS_ERR_ARG, "null orthonormalize argument");
  KS_CHECK(n > 0 && ld >= n && k >= 0 && k < KS_MAX_COLS && ldbuf > k && hcol >= 0, KS_ERR_ARG,
           "bad orthonormalize sizes n=%d ld=%d k=%d ldbuf=%d", n, ld, k, ldbuf);
  KS_CHECK(refine >= KS_REFINE_IFNEEDED && refine <= KS_REFINE_ALWAYS, KS_ERR_ARG, "bad refinement type");
  if (prec == KS_F64)
    return orthonormalize_t<double>(ctx, n, (double *)V, ld, k, d_buf, ldbuf, hcol, col, refine, eta, 1, 0, 2);
  return orthonormalize_t<float>(ctx, n, (float *)V, ld, k, d_buf, ldbuf, hcol, col, refine, eta, 1, 0, 2);
}

int ks_bv_dotvec(ks_ctx ctx, ks_prec prec, int n, const void *V, int ld, int k, const void *y, double *d_q) {
  KS_CHECK(ctx && y && d_q && (V || k == 0), KS_ERR_ARG, "null dotvec argument");
  KS_CHECK(n > 0 && ld >= n && k >= 0 && k <= KS_MAX_COLS, KS_ERR_ARG, "bad dotvec sizes");
  if (k == 0) return KS_OK;
  int rpt, nb, e;
  if ((e = split_checked(n, &rpt, &nb))) return e;
  const double s = prec == KS_F64 ? 8.0 : 4.0;
  ks_prof_begin(ctx, KS_PROF_DOTVEC, (k + 1.0) * n * s, 2.0 * k * n, false);
  if (prec == KS_F64)
    launch_dot<double, GS_PLAIN>(ctx, rpt, nb, (const double *)V, ld, k, (const double *)y, n, d_q, nullptr, 0, 0,
                                 nullptr);
  else
    launch_dot<float, GS_PLAIN>(ctx, rpt, nb, (const float *)V, ld, k, (const float *)y, n, d_q, nullptr, 0, 0,
                                nullptr);
  ks_prof_end(ctx);
  KS_HIP(hipGetLastError());
  return KS_OK;
}

int ks_bv_multvec(ks_ctx ctx, ks_prec prec, int n, double alpha, const void *V, int ld, int k, const double *d_q,
                  double beta, void *y) {
  KS_CHECK(ctx && y && (k == 0 || (V && d_q)), KS_ERR_ARG, "null multvec argument");
  KS_CHECK(n > 0 && ld >= n && k >= 0 && k <= KS_MAX_COLS, KS_ERR_ARG, "bad multvec sizes");
  int rpt, nb, e;
  if ((e = split_checked(n, &rpt, &nb))) return e;
  const double s = prec == KS_F64 ? 8.0 : 4.0;
  ks_prof_begin(ctx, KS_PROF_MULTVEC, (k + 2.0) * n * s, 2.0 * k * n, false);
  if (prec == KS_F64)
    launch_upd<double, GS_PLAIN>(ctx, rpt, nb, (const double *)V, ld, k, (double *)y, n, d_q, alpha, beta, nullptr,
                                 0, 0, 0, 0.0, 0, 0, nullptr);
  else
    launch_upd<float, GS_PLAIN>(ctx, rpt, nb, (const float *)V, ld, k, (float *)y, n, d_q, alpha, beta, nullptr, 0,
                                0, 0, 0.0, 0, 0, nullptr);
  ks_prof_end(ctx);
  KS_HIP(hipGetLastError());
  return KS_OK;
}

int ks_bv_scale(ks_ctx ctx, ks_prec prec, int n, void *x, double alpha) {
  KS_CHECK(ctx && x && n >= 0, KS_ERR_ARG, "bad scale arguments");
  if (n == 0 || alpha == 1.0) return KS_OK;
  const double s = prec == KS_F64 ? 8.0 : 4.0;
  ks_prof_begin(ctx, KS_PROF_SCALE, 2.0 * n * s, (double)n, false);
  if (prec == KS_F64)
    hipLaunchKernelGGL((k_scale<double>), dim3(ew_blocks(n)), dim3(256), 0, ctx->stream, (double *)x, n, alpha);
  else
    hipLaunchKernelGGL((k_scale<float>), dim3(ew_blocks(n)), dim3(256), 0, ctx->stream, (float *)x, n, alpha);
  ks_prof_end(ctx);
  KS_HIP(hipGetLastError());
  return KS_OK;
}

int ks_bv_norm(ks_ctx ctx, ks_prec prec, int n, const void *x, double *nrm) {
  KS_CHECK(ctx && x && nrm && n > 0, KS_ERR_ARG, "bad norm arguments");
  int rpt, nb, e;
  if ((e = split_checked(n, &rpt, &nb))) return e;
  const double s = prec == KS_F64 ? 8.0 : 4.0;
  ks_prof_begin(ctx, KS_PROF_NORM, (double)n * s, 2.0 * n, false);
  dim3 grid(nb), block(KS_BLOCK);
#define KS_NORM_CASE(T, R)                                                                                       \
  case R:                                                                                                        \
    hipLaunchKernelGGL((k_norm<T, R>), grid, block, 0, ctx->stream, (const T *)x, n, ctx->d_partial);            \
    break;
  if (prec == KS_F64) {
    switch (rpt) { KS_NORM_CASE(double, 1) KS_NORM_CASE(double, 2) KS_NORM_CASE(double, 4) KS_NORM_CASE(double, 8) KS_NORM_CASE(double, 16) }
  } else {
    switch (rpt) { KS_NORM_CASE(float, 1) KS_NORM_CASE(float, 2) KS_NORM_CASE(float, 4) KS_NORM_CASE(float, 8) KS_NORM_CASE(float, 16) }
  }
#undef KS_NORM_CASE
  hipLaunchKernelGGL(k_norm_fin, dim3(1), block, 0, ctx->stream, (const double *)ctx->d_partial, nb, ctx->d_scalar);
  ks_prof_end(ctx);
  KS_HIP(hipGetLastError());
  KS_HIP(hipMemcpyAsync(ctx->h_scalar, ctx->d_scalar, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  KS_HIP(hipStreamSynchronize(ctx->stream));
  *nrm = ctx->h_scalar[0];
  return KS_OK;
}

int ks_bv_axpy(ks_ctx ctx, ks_prec prec, int n, double alpha, const void *x, void *y) {
  KS_CHECK(ctx && x && y && n >= 0, KS_ERR_ARG, "bad axpy arguments");
  if (n == 0) return KS_OK;
  const double s = prec == KS_F64 ? 8.0 : 4.0;
  ks_prof_begin(ctx, KS_PROF_AXPY, 3.0 * n * s, 2.0 * n, false);
  if (prec == KS_F64)
    hipLaunchKernelGGL((k_axpy<double>), dim3(ew_blocks(n)), dim3(256), 0, ctx->stream, (double *)y,
                       (const double *)x, n, alpha);
  else
    hipLaunchKernelGGL((k_axpy<float>), dim3(ew_blocks(n)), dim3(256), 0, ctx->stream, (float *)y,
                       (const float *)x, n, alpha);
  ks_prof_end(ctx);
  KS_HIP(hipGetLastError());
  return KS_OK;
}

int ks_bv_copy(ks_ctx ctx, ks_prec prec, int n, const void *x, void *y) {
  KS_CHECK(ctx && x && y && n >= 0, KS_ERR_ARG, "bad copy arguments");
  if (n == 0 || x == y) return KS_OK;
  const size_t s = prec == KS_F64 ? 8 : 4;
  ks_prof_begin(ctx, KS_PROF_COPY, 2.0 * n * s, 0.0, false);
  KS_HIP(hipMemcpyAsync(y, x, s * n, hipMemcpyDeviceToDevice, ctx->stream));
  ks_prof_end(ctx);
  return KS_OK;
}

int ks_bv_random(ks_ctx ctx, ks_prec prec, int n, void *x, uint64_t seed) {
  KS_CHECK(ctx && x && n >= 0, KS_ERR_ARG, "bad random arguments");
  if (n == 0) return KS_OK;
  ks_prof_begin(ctx, KS_PROF_OTHER, (double)n * (prec == KS_F64 ? 8 : 4), 0.0, false);
  if (prec == KS_F64)
    hipLaunchKernelGGL((k_random<double>), dim3(ew_blocks(n)), dim3(256), 0, ctx->stream, (double *)x, n, seed);
  else
    hipLaunchKernelGGL((k_random<float>), dim3(ew_blocks(n)), dim3(256), 0, ctx->stream, (float *)x, n, seed);
  ks_prof_end(ctx);
  KS_HIP(hipGetLastError());
  return KS_OK;
}

int ks_bv_convert(ks_ctx ctx, int n, ks_prec px, const void *x, ks_prec py, void *y) {
  KS_CHECK(ctx && x && y && n >= 0, KS_ERR_ARG, "bad convert arguments");
  if (n == 0) return KS_OK;
  if (px == py) return ks_bv_copy(ctx, px, n, x, y);
  ks_prof_begin(ctx, KS_PROF_OTHER, 12.0 * n, 0.0, false);
  if (px == KS_F64)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(ew_blocks(n)), dim3(256), 0, ctx->stream, (float *)y,
                       (const double *)x, n);
  else
    hipLaunchKernelGGL((k_convert<float, double>), dim3(ew_blocks(n)), dim3(256), 0, ctx->stream, (double *)y,
                       (const float *)x, n);
  ks_prof_end(ctx);
  KS_HIP(hipGetLastError());
  return KS_OK;
}

int ks_bv_gram(ks_ctx ctx, ks_prec prec, int n, const void *X, int ldx, int p, const void *Y, int ldy, int q,
               double *d_G, int ldg) {
  KS_CHECK(ctx && X && Y && d_G, KS_ERR_ARG, "null gram argument");
  KS_CHECK(n > 0 && p >= 0 && q >= 0 && ldx >= n && ldy >= n && ldg >= q, KS_ERR_ARG,
           "bad gram sizes n=%d p=%d q=%d ldx=%d ldy=%d ldg=%d", n, p, q, ldx, ldy, ldg);
  KS_CHECK(p <= KS_MAX_COLS && q <= KS_MAX_COLS, KS_ERR_RANGE, "gram supports at most %d columns", KS_MAX_COLS);
  if (p == 0 || q == 0) return KS_OK;
  const double sz = prec == KS_F64 ? 8.0 : 4.0;
  ks_prof_begin(ctx, KS_PROF_OTHER, (double)n * (p + q) * sz, 2.0 * n * (double)p * q, false);
  int err = prec == KS_F64 ? gram_t<double>(ctx, n, (const double *)X, ldx, p, (const double *)Y, ldy, q, d_G, ldg)
                           : gram_t<float>(ctx, n, (const float *)X, ldx, p, (const float *)Y, ldy, q, d_G, ldg);
  ks_prof_end(ctx);
  return err;
}

int ks_bv_multinplace(ks_ctx ctx, ks_prec prec, int n, void *V, int ld, int k, const double *d_Q, int ldq, int s,
                      int e) {
  KS_CHECK(ctx && V && d_Q, KS_ERR_ARG, "null multinplace argument");
  KS_CHECK(n > 0 && ld >= n && k >= 0 && ldq >= k && s >= 0, KS_ERR_ARG, "bad multinplace sizes");
  KS_CHECK(k <= KS_MAX_COLS && e <= KS_MAX_COLS, KS_ERR_RANGE, "multinplace supports at most %d columns",
           KS_MAX_COLS);
  if (s >= e) return KS_OK;
  const int N = e - s;
  const double sz = prec == KS_F64 ? 8.0 : 4.0;
  ks_prof_begin(ctx, KS_PROF_MULTINPLACE, (double)n * k * sz + (double)n * N * sz, 2.0 * n * (double)k * N, false);
  int err = prec == KS_F64 ? dispatch_mip<double>(ctx, (double *)V, ld, n, k, d_Q, ldq, s, N)
                           : dispatch_mip<float>(ctx, (float *)V, ld, n, k, d_Q, ldq, s, N);
  ks_prof_end(ctx);
  if (err) return err;
  KS_HIP(hipGetLastError());
  return KS_OK;
}

int ks_bv_orthonormalize_column_noalloc(void *stream, ks_prec prec, int n, void *V, int ld, int k, double *d_buf,
                                        int ldbuf, int hcol, int col, ks_refine refine, double eta, int scale,
                                        int safe, ks_status *d_status_out) {
  KS_CHECK(V && d_buf, KS_ERR_ARG, "null orthonormalize argument");
  KS_CHECK(n > 0 && ld >= n && k >= 0 && ldbuf > k && hcol >= 0, KS_ERR_ARG,
           "bad orthonormalize sizes n=%d ld=%d k=%d ldbuf=%d", n, ld, k, ldbuf);
  KS_CHECK(refine >= KS_REFINE_IFNEEDED && refine <= KS_REFINE_ALWAYS, KS_ERR_ARG, "bad refinement type");
  KS_CHECK(k <= KS_NOALLOC_MAXK, KS_ERR_RANGE,
           "the allocation-free orthonormalization takes at most %d columns (asked %d); use "
           "ks_bv_orthonormalize_column", KS_NOALLOC_MAXK, k);
  gs_geom g;
  KS_TRY(gs_geometry(n, k, prec == KS_F64 ? sizeof(double) : sizeof(float), &g));
  KS_CHECK((size_t)(k + 1) * (size_t)g.nb <= KS_NOALLOC_PARTIAL_CAP, KS_ERR_RANGE,
           "%d x %d partial sums exceed the scratch array", k + 1, g.nb);
  double *base = nullptr;
  KS_TRY(noalloc_scratch(&base));
  // a context view of the scratch: no events, no allocation
  ks_ctx_s c;
  c.stream = (hipStream_t)stream;
  c.d_status = (ks_status *)(base + KS_NOALLOC_STATUS);
  c.d_counters = (unsigned *)(base + KS_NOALLOC_COUNTER);
  c.d_gpartial = base + KS_NOALLOC_SUMS;
  c.d_partial = base + KS_NOALLOC_PARTIAL;
  hipLaunchKernelGGL(k_noalloc_reset, dim3(1), dim3(64), 0, c.stream, c.d_status, c.d_counters);
  int err = prec == KS_F64 ? orthonormalize_t<double>(&c, n, (double *)V, ld, k, d_buf, ldbuf, hcol, col, refine,
                                                      eta, scale, safe)
                           : orthonormalize_t<float>(&c, n, (float *)V, ld, k, d_buf, ldbuf, hcol, col, refine,
                                                     eta, scale, safe);
  if (err) return err;
  if (d_status_out)
    hipLaunchKernelGGL(k_noalloc_status_out, dim3(1), dim3(64), 0, c.stream, c.d_status, d_status_out);
  KS_HIP(hipGetLastError());
  return KS_OK;
}

}  // extern "C"
