// ks_ksp.hip -- the linear solves of shift-and-invert on the device.
//
// STSINVERT applies (A - sigma I)^-1 by an inner Krylov solve (the reference's
// STApply_Sinvert -> KSPSolve, src/sys/classes/st/impls/sinvert/sinvert.c:16-30,
// with KSPCG and PCJACOBI as in src/eps/tutorials/ex11.c's test 2).  PETSc is not
// part of the reference tree; the solver here is the preconditioned conjugate
// gradient method in PETSc's KSPCG form (Hestenes-Stiefel recurrences, the
// preconditioned residual norm |B r| for the test rnorm <= max(rtol |B b|, abstol))
// with B = diag(A - sigma I)^-1 (zero diagonal entries replaced by 1, as PCJACOBI
// does), and for an indefinite shifted matrix BiCGStab (van der Vorst 1992) in
// PETSc's KSPBCGS form with the same left preconditioner.  Every scalar of a
// recurrence and the convergence test stay on the device: the finishing
// launch of an iteration decides, and once a solve has converged or failed the
// later launches of the chunk exit at their first instruction (the products
// through a status of the solver's own), so the host enqueues iterations in
// chunks and reads one small state block per chunk.  Reductions add
// per-workgroup partial sums in a fixed order in a one-workgroup launch, so a
// solve is deterministic.
#include <algorithm>
#include <cstring>
#include <vector>

#include "ks_internal.hpp"

namespace {

#define KS_PCG_BLOCKS 512   // workgroups of the vector kernels (partials: 3 per workgroup)

// the solver state after the vectors of the workspace: the gate of its
// products (gate.active = 1 while iterating), the scalars, the iteration count
// and the outcome (reason 0 while iterating, > 0 converged, < 0 failed)
struct ksp_state {
  ks_status gate;
  double sc[16];
  int32_t its, reason, sconv, maxits;
};
static_assert(sizeof(ksp_state) <= 64 * sizeof(double), "the solver state takes 64 doubles");
enum { S_RZ = 0, S_RZOLD, S_PQ, S_ZZ, S_TTOL, S_RNORM,                                  // CG
       S_RHO = 0, S_RHOOLD = 1, S_ALPHA = 6, S_OMEGA, S_RV, S_SS, S_TS, S_TT, S_RR, S_RHR };   // BiCGStab

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

// partial sums of up to three products, per workgroup, stored at part[3 b + j]
__device__ inline void store_partials(double *part, double a, double b, double c, double *lds) {
  a = block_sum(a, lds);
  b = block_sum(b, lds);
  c = block_sum(c, lds);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = a;
    part[3 * blockIdx.x + 1] = b;
    part[3 * blockIdx.x + 2] = c;
  }
}

// the three column sums of the partials, in workgroup order (thread 0)
__device__ inline void sum_partials(const double *part, int nb, double *out, double *lds) {
  double a = 0.0, b = 0.0, c = 0.0;
  for (int q = threadIdx.x; q < nb; q += KS_BLOCK) {
    a += part[3 * q];
    b += part[3 * q + 1];
    c += part[3 * q + 2];
  }
  out[0] = block_sum(a, lds);
  out[1] = block_sum(b, lds);
  out[2] = block_sum(c, lds);
}

#define KS_FOR_ROWS(i, n) for (int i = blockIdx.x * KS_BLOCK + threadIdx.x; i < n; i += gridDim.x * KS_BLOCK)

// dinv[i] = 1 / (a_ii - sigma), 1 where that is zero; all ones without a preconditioner
__global__ __launch_bounds__(KS_BLOCK) void k_jacobi(const int32_t *__restrict__ rowptr,
                                                    const int32_t *__restrict__ colidx,
                                                    const double *__restrict__ val, int n, double sigma, int jacobi,
                                                    double *__restrict__ dinv) {
  KS_FOR_ROWS(i, n) {
    double d = 0.0;
    if (jacobi) {
      for (int32_t e = rowptr[i]; e < rowptr[i + 1]; e++)
        if (colidx[e] == i) d += val[e];
      d -= sigma;
    }
    dinv[i] = (d == 0.0) ? 1.0 : 1.0 / d;
  }
}

// ---------------------------------------------------------------- CG
// x = 0, r = b, z = B r; partials of z.r and z.z
__global__ __launch_bounds__(KS_BLOCK) void k_pcg_start(int n, const double *__restrict__ b, double *__restrict__ x,
                                                       double *__restrict__ r, double *__restrict__ z,
                                                       const double *__restrict__ dinv, double *__restrict__ part) {
  __shared__ double lds[KS_BLOCK / 64];
  double zr = 0.0, zz = 0.0;
  KS_FOR_ROWS(i, n) {
    const double ri = b[i], zi = dinv[i] * ri;
    x[i] = 0.0;
    r[i] = ri;
    z[i] = zi;
    zr += zi * ri;
    zz += zi * zi;
  }
  store_partials(part, zr, zz, 0.0, lds);
}

// first z.r and z.z; the convergence threshold from |B b|
__global__ __launch_bounds__(KS_BLOCK) void k_pcg_fin_start(const double *__restrict__ part, int nb, ksp_state *S,
                                                           double rtol, double abstol, int maxits) {
  __shared__ double lds[KS_BLOCK / 64];
  double t[3];
  sum_partials(part, nb, t, lds);
  if (threadIdx.x == 0) {
    const double rnorm0 = sqrt(t[1]);
    S->sc[S_RZ] = t[0];
    S->sc[S_ZZ] = t[1];
    S->sc[S_RNORM] = rnorm0;
    S->sc[S_TTOL] = fmax(rtol * rnorm0, abstol);
    S->its = 0;
    S->maxits = maxits;
    S->reason = rnorm0 <= S->sc[S_TTOL] ? 1 : 0;
    S->gate.active = S->reason ? 0 : 1;
  }
}

// p = z (first iteration) or z + (rz / rzold) p
__global__ __launch_bounds__(KS_BLOCK) void k_pcg_p(int n, const double *__restrict__ z, double *__restrict__ p,
                                                   const ksp_state *S) {
  if (S->reason) return;
  const bool first = S->its == 0;
  const double beta = first ? 0.0 : S->sc[S_RZ] / S->sc[S_RZOLD];
  KS_FOR_ROWS(i, n) p[i] = first ? z[i] : z[i] + beta * p[i];
}

// q -= sigma p (q = A p on entry); partials of p.q
__global__ __launch_bounds__(KS_BLOCK) void k_pcg_q(int n, const double *__restrict__ p, double *__restrict__ q,
                                                   double sigma, const ksp_state *S, double *__restrict__ part) {
  __shared__ double lds[KS_BLOCK / 64];
  if (S->reason) return;
  double pq = 0.0;
  KS_FOR_ROWS(i, n) {
    const double qi = q[i] - sigma * p[i];
    q[i] = qi;
    pq += p[i] * qi;
  }
  store_partials(part, pq, 0.0, 0.0, lds);
}

__global__ __launch_bounds__(KS_BLOCK) void k_pcg_fin_pq(const double *__restrict__ part, int nb, ksp_state *S) {
  __shared__ double lds[KS_BLOCK / 64];
  if (S->reason) return;
  double t[3];
  sum_partials(part, nb, t, lds);
  if (threadIdx.x == 0) S->sc[S_PQ] = t[0];
}

// a = rz / pq; x += a p; r -= a q; z = B r; partials of z.r and z.z
__global__ __launch_bounds__(KS_BLOCK) void k_pcg_xr(int n, const double *__restrict__ p,
                                                    const double *__restrict__ q, double *__restrict__ x,
                                                    double *__restrict__ r, double *__restrict__ z,
                                                    const double *__restrict__ dinv, const ksp_state *S,
                                                    double *__restrict__ part) {
  __shared__ double lds[KS_BLOCK / 64];
  if (S->reason) return;
  const double a = S->sc[S_RZ] / S->sc[S_PQ];
  double zr = 0.0, zz = 0.0;
  KS_FOR_ROWS(i, n) {
    x[i] += a * p[i];
    const double ri = r[i] - a * q[i], zi = dinv[i] * ri;
    r[i] = ri;
    z[i] = zi;
    zr += zi * ri;
    zz += zi * zi;
  }
  store_partials(part, zr, zz, 0.0, lds);
}

// the new z.r and z.z, and the tests of the iteration: p^T (A - sigma I) p <= 0
// (KSP_DIVERGED_INDEFINITE_MAT, -8), |B r| <= ttol (1), the iteration limit (-3)
__global__ __launch_bounds__(KS_BLOCK) void k_pcg_fin_iter(const double *__restrict__ part, int nb, ksp_state *S) {
  __shared__ double lds[KS_BLOCK / 64];
  if (S->reason) return;
  double t[3];
  sum_partials(part, nb, t, lds);
  if (threadIdx.x == 0) {
    S->sc[S_RZOLD] = S->sc[S_RZ];
    S->sc[S_RZ] = t[0];
    S->sc[S_ZZ] = t[1];
    S->sc[S_RNORM] = sqrt(t[1]);
    S->its += 1;
    if (!(S->sc[S_PQ] > 0.0)) S->reason = -8;
    else if (S->sc[S_RNORM] <= S->sc[S_TTOL]) S->reason = 1;
    else if (S->its >= S->maxits) S->reason = -3;
    S->gate.active = S->reason ? 0 : 1;
  }
}

// ---------------------------------------------------------------- BiCGStab
// x = 0, r = B b, rhat = r, p = v = 0; partials of r.r
__global__ __launch_bounds__(KS_BLOCK) void k_bcgs_start(int n, const double *__restrict__ b, double *__restrict__ x,
                                                        double *__restrict__ r, double *__restrict__ rh,
                                                        double *__restrict__ p, double *__restrict__ v,
                                                        const double *__restrict__ dinv, double *__restrict__ part) {
  __shared__ double lds[KS_BLOCK / 64];
  double rr = 0.0;
  KS_FOR_ROWS(i, n) {
    const double ri = dinv ? dinv[i] * b[i] : b[i];   // no dinv: b holds B b already
    x[i] = 0.0;
    r[i] = ri;
    rh[i] = ri;
    p[i] = 0.0;
    v[i] = 0.0;
    rr += ri * ri;
  }
  store_partials(part, rr, 0.0, 0.0, lds);
}

__global__ __launch_bounds__(KS_BLOCK) void k_bcgs_fin_start(const double *__restrict__ part, int nb, ksp_state *S,
                                                            double rtol, double abstol, int maxits) {
  __shared__ double lds[KS_BLOCK / 64];
  double t[3];
  sum_partials(part, nb, t, lds);
  if (threadIdx.x == 0) {
    const double rnorm0 = sqrt(t[0]);
    S->sc[S_RHO] = t[0];   // (rhat, r) with rhat = r
    S->sc[S_RHOOLD] = 1.0;
    S->sc[S_ALPHA] = 1.0;
    S->sc[S_OMEGA] = 1.0;
    S->sc[S_RNORM] = rnorm0;
    S->sc[S_TTOL] = fmax(rtol * rnorm0, abstol);
    S->its = 0;
    S->sconv = 0;
    S->maxits = maxits;
    S->reason = rnorm0 <= S->sc[S_TTOL] ? 1 : 0;
    S->gate.active = S->reason ? 0 : 1;
  }
}

// p = r + beta (p - omega v), beta = (rho / rho_old) (alpha / omega)
__global__ __launch_bounds__(KS_BLOCK) void k_bcgs_p(int n, const double *__restrict__ r, double *__restrict__ p,
                                                    const double *__restrict__ v, const ksp_state *S) {
  if (S->reason) return;
  const double beta = (S->sc[S_RHO] / S->sc[S_RHOOLD]) * (S->sc[S_ALPHA] / S->sc[S_OMEGA]);
  const double om = S->sc[S_OMEGA];
  KS_FOR_ROWS(i, n) p[i] = r[i] + beta * (p[i] - om * v[i]);
}

// y = B (y - sigma x) (y = A x on entry) with B = diag(dinv), or y as it is
// when dinv is null (ILU: the shift and B applied by earlier launches);
// partials of (h, y), (x, y), (y, y)
__global__ __launch_bounds__(KS_BLOCK) void k_bcgs_apply(int n, const double *__restrict__ x, double *__restrict__ y,
                                                        const double *__restrict__ h, const double *__restrict__ dinv,
                                                        double sigma, const ksp_state *S, int skip_sconv,
                                                        double *__restrict__ part) {
  __shared__ double lds[KS_BLOCK / 64];
  if (S->reason || (skip_sconv && S->sconv)) return;
  double hy = 0.0, xy = 0.0, yy = 0.0;
  KS_FOR_ROWS(i, n) {
    const double yi = dinv ? dinv[i] * (y[i] - sigma * x[i]) : y[i];   // no dinv: y holds B (A - sigma I) x
    if (dinv) y[i] = yi;
    hy += h[i] * yi;
    xy += x[i] * yi;
    yy += yi * yi;
  }
  store_partials(part, hy, xy, yy, lds);
}

// alpha = rho / (rhat, v); a zero denominator is a breakdown (-5)
__global__ __launch_bounds__(KS_BLOCK) void k_bcgs_fin_alpha(const double *__restrict__ part, int nb, ksp_state *S) {
  __shared__ double lds[KS_BLOCK / 64];
  if (S->reason) return;
  double t[3];
  sum_partials(part, nb, t, lds);
  if (threadIdx.x == 0) {
    S->sc[S_RV] = t[0];
    if (t[0] == 0.0) {
      S->reason = -5;
      S->gate.active = 0;
    } else S->sc[S_ALPHA] = S->sc[S_RHO] / t[0];
  }
}

// s = r - alpha v; partials of (s, s)
__global__ __launch_bounds__(KS_BLOCK) void k_bcgs_s(int n, const double *__restrict__ r, const double *__restrict__ v,
                                                    double *__restrict__ sv, const ksp_state *S,
                                                    double *__restrict__ part) {
  __shared__ double lds[KS_BLOCK / 64];
  if (S->reason) return;
  const double a = S->sc[S_ALPHA];
  double ss = 0.0;
  KS_FOR_ROWS(i, n) {
    const double si = r[i] - a * v[i];
    sv[i] = si;
    ss += si * si;
  }
  store_partials(part, ss, 0.0, 0.0, lds);
}

// |s| <= ttol: the iteration ends with x += alpha p (no product with s)
__global__ __launch_bounds__(KS_BLOCK) void k_bcgs_fin_s(const double *__restrict__ part, int nb, ksp_state *S) {
  __shared__ double lds[KS_BLOCK / 64];
  if (S->reason) return;
  double t[3];
  sum_partials(part, nb, t, lds);
  if (threadIdx.x == 0) {
    S->sc[S_SS] = t[0];
    S->sconv = sqrt(t[0]) <= S->sc[S_TTOL] ? 1 : 0;
    S->gate.active = S->sconv ? 0 : 1;
  }
}

// omega = (t, s) / (t, t); a zero (t, t) is a breakdown
__global__ __launch_bounds__(KS_BLOCK) void k_bcgs_fin_omega(const double *__restrict__ part, int nb, ksp_state *S) {
  __shared__ double lds[KS_BLOCK / 64];
  if (S->reason || S->sconv) return;
  double t[3];
  sum_partials(part, nb, t, lds);
  if (threadIdx.x == 0) {
    // apply(x = s, y = t) returned (rhat, t), (s, t), (t, t)
    S->sc[S_TS] = t[1];
    S->sc[S_TT] = t[2];
    if (t[2] == 0.0) {
      S->reason = -5;
      S->gate.active = 0;
    } else S->sc[S_OMEGA] = t[1] / t[2];
  }
}

// x += alpha p + omega s, r = s - omega t (or x += alpha p after an early |s|);
// partials of (r, r) and (rhat, r)
__global__ __launch_bounds__(KS_BLOCK) void k_bcgs_x(int n, double *__restrict__ x, const double *__restrict__ p,
                                                    const double *__restrict__ sv, const double *__restrict__ t,
                                                    double *__restrict__ r, const double *__restrict__ rh,
                                                    const ksp_state *S, double *__restrict__ part) {
  __shared__ double lds[KS_BLOCK / 64];
  if (S->reason) return;
  const double a = S->sc[S_ALPHA];
  if (S->sconv) {
    KS_FOR_ROWS(i, n) x[i] += a * p[i];
    return;
  }
  const double om = S->sc[S_OMEGA];
  double rr = 0.0, rhr = 0.0;
  KS_FOR_ROWS(i, n) {
    x[i] += a * p[i] + om * sv[i];
    const double ri = sv[i] - om * t[i];
    r[i] = ri;
    rr += ri * ri;
    rhr += rh[i] * ri;
  }
  store_partials(part, rr, rhr, 0.0, lds);
}

// the tests of the iteration: converged (1), breakdown (-5), iteration limit (-3)
__global__ __launch_bounds__(KS_BLOCK) void k_bcgs_fin_iter(const double *__restrict__ part, int nb, ksp_state *S) {
  __shared__ double lds[KS_BLOCK / 64];
  if (S->reason) return;
  double t[3];
  const bool sconv = S->sconv != 0;
  if (!sconv) sum_partials(part, nb, t, lds);
  if (threadIdx.x == 0) {
    S->its += 1;
    if (sconv) {
      S->sc[S_RNORM] = sqrt(S->sc[S_SS]);
      S->reason = 1;
    } else {
      S->sc[S_RR] = t[0];
      S->sc[S_RNORM] = sqrt(t[0]);
      if (S->sc[S_RNORM] <= S->sc[S_TTOL]) S->reason = 1;
      else {
        S->sc[S_RHOOLD] = S->sc[S_RHO];
        S->sc[S_RHO] = t[1];
        if (t[1] == 0.0 || S->sc[S_OMEGA] == 0.0) S->reason = -5;
        else if (S->its >= S->maxits) S->reason = -3;
      }
    }
    S->gate.active = S->reason ? 0 : 1;
  }
}

// ---------------------------------------------------------------- ILU(0)
// lu = values of A - sigma I on A's pattern
__global__ __launch_bounds__(KS_BLOCK) void k_ilu0_init(const int32_t *__restrict__ rowptr,
                                                       const int32_t *__restrict__ colidx,
                                                       const double *__restrict__ val, int n, double sigma,
                                                       double *__restrict__ lu) {
  KS_FOR_ROWS(i, n)
    for (int32_t q = rowptr[i]; q < rowptr[i + 1]; q++) lu[q] = val[q] - (colidx[q] == i ? sigma : 0.0);
}

// one level of the IKJ factorization (Saad, Algorithm 10.4): a thread per row of
// the level; the rows it reads (its L part) belong to earlier levels.  The
// operations of a row are the oracle's, in the same order.
__global__ __launch_bounds__(64) void k_ilu0_factor_level(const int32_t *__restrict__ rows, int cnt,
                                                         const int32_t *__restrict__ rowptr,
                                                         const int32_t *__restrict__ colidx,
                                                         const int32_t *__restrict__ diag, double *__restrict__ lu,
                                                         int *zero_pivot) {
  const int t = blockIdx.x * 64 + threadIdx.x;
  if (t >= cnt) return;
  const int i = rows[t];
  const int32_t re = rowptr[i + 1];
  for (int32_t q = rowptr[i]; q < diag[i]; q++) {
    const int k = colidx[q];
    const double ukk = lu[diag[k]];
    if (ukk == 0.0) {
      atomicExch(zero_pivot, 1);
      return;
    }
    const double lik = lu[q] / ukk;
    lu[q] = lik;
    int32_t u = diag[k] + 1;
    const int32_t ke = rowptr[k + 1];
    for (int32_t r = q + 1; r < re; r++) {
      const int j = colidx[r];
      while (u < ke && colidx[u] < j) u++;
      if (u < ke && colidx[u] == j) lu[r] -= lik * lu[u];
    }
  }
}

// one level of y = L^-1 y (unit lower factor), in place; skipped once the solve
// One triangular solve of y = (L U)^-1 y in one launch (LOWER: the unit lower
// factor, forward; else the upper factor, backward), in place, one wavefront per
// row: row i goes to wavefront i (forward) or n-1-i (backward), so a row waits
// only for rows of lower-numbered wavefronts, dispatched before it.  The lanes
// take the row's entries of its factor part (64 at a time), wait until each
// entry's column is solved and load its value; lane order then subtracts them
// in column order, as the oracle's loop does.  A solved value is published as
// two 64-bit words, each a 32-bit half of the double beside the solve's epoch,
// written and polled with agent-scope atomics: a reader takes a value once both
// words carry the current epoch, so no fence is needed.  Waits are bounded
// (KS_ILU_SPIN polls, or sooner once another wait has run out): then bit 1 of
// *err is set and the row goes on, so every row publishes and every wavefront
// exits; the solve reports the failure.  Skipped like the products when the
// solver's state says so (the state is read once, at the start).
#define KS_ILU_SPIN (1 << 26)
template <bool LOWER>
__global__ __launch_bounds__(KS_BLOCK) void k_ilu0_sf(int n, const int32_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ colidx,
                                                     const int32_t *__restrict__ diag, const double *__restrict__ lu,
                                                     double *y, unsigned long long *slot, unsigned epoch, int *err,
                                                     const ksp_state *S, int skip_sconv) {
  if (S && (S->reason || (skip_sconv && S->sconv))) return;
  const int w = blockIdx.x * KS_WAVES + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (w >= n) return;
  const int i = LOWER ? w : n - 1 - w;
  const int q0 = LOWER ? rowptr[i] : diag[i] + 1, q1 = LOWER ? diag[i] : rowptr[i + 1];
  const unsigned long long tag = (unsigned long long)epoch << 32;
  double s = y[i];
  for (int qb = q0; qb < q1; qb += 64) {
    const int q = qb + lane;
    double a = 0.0, v = 0.0;
    if (q < q1) {
      a = lu[q];
      const unsigned long long *sl = slot + 2 * (size_t)colidx[q];
      unsigned long long lo, hi;
      int spins = 0;
      for (;;) {
        lo = __hip_atomic_load(sl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        hi = __hip_atomic_load(sl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((lo >> 32) == epoch && (hi >> 32) == epoch) break;
        // past the bound, or once another wait has run out, the row goes on (and publishes)
        if (++spins > KS_ILU_SPIN || ((spins & 1023) == 0 && (__hip_atomic_load(err, __ATOMIC_RELAXED,
                                                                                  __HIP_MEMORY_SCOPE_AGENT) & 2))) {
          atomicOr(err, 2);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      v = __longlong_as_double((long long)(((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull)));
    }
    const int cnt = q1 - qb < 64 ? q1 - qb : 64;
    for (int t = 0; t < cnt; t++) s -= ks_readlane_f64(a, t) * ks_readlane_f64(v, t);
  }
  const double x = LOWER ? s : s / lu[diag[i]];
  if (lane == 0) {
    y[i] = x;
    const unsigned long long bits = (unsigned long long)__double_as_longlong(x);
    __hip_atomic_store(slot + 2 * (size_t)i, tag | (bits & 0xffffffffull), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(slot + 2 * (size_t)i + 1, tag | (bits >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// y = y - sigma x, skipped like the products
__global__ __launch_bounds__(KS_BLOCK) void k_shift_gated(int n, const double *__restrict__ x, double *__restrict__ y,
                                                         double sigma, const ksp_state *S, int skip_sconv) {
  if (S->reason || (skip_sconv && S->sconv)) return;
  KS_FOR_ROWS(i, n) y[i] -= sigma * x[i];
}

inline int ksp_blocks(int n) {
  int b = (n + KS_BLOCK - 1) / KS_BLOCK;
  if (b > KS_PCG_BLOCKS) b = KS_PCG_BLOCKS;
  return b < 1 ? 1 : b;
}

#define KS_KSP_CHUNK 8   // iterations enqueued between two reads of the solver state

// waits for the stream, then the iteration count, the outcome and |B r|
static int read_state(ks_ctx ctx, const ksp_state *S, int *its, int *reason, double *rnorm) {
  ksp_state h;
  KS_HIP(hipMemcpyAsync(ctx->h_scalar, S, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  KS_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(&h, ctx->h_scalar, sizeof(h));
  *its = h.its;
  *reason = h.reason;
  *rnorm = h.sc[S_RNORM];
  return KS_OK;
}

}  // namespace

struct ks_ilu_s {
  int n = 0;
  std::vector<int> lptr, uptr;        // level starts (host): the factorization's schedule; the
                                      // backward solve's depth (ks_ilu0_levels)
  int32_t *d_lrows = nullptr, *d_diag = nullptr;
  double *d_lu = nullptr;
  int *d_flag = nullptr;              // bit 0: a zero pivot; bit 1: a solve's wait ran out
  unsigned long long *d_slot = nullptr;   // [n][2] published values of the triangular solves
  unsigned epoch = 0;
  const int32_t *d_rowptr = nullptr, *d_colidx = nullptr;   // the operator's pattern (borrowed)
};

namespace {

// y = (L U)^-1 y on the context stream: two launches, the forward and the
// backward solve, each a new epoch of the published values
void ilu0_apply(ks_ctx ctx, ks_ilu M, double *y, const ksp_state *S, int skip_sconv) {
  const int blocks = (M->n + KS_WAVES - 1) / KS_WAVES;
  M->epoch += 1;
  hipLaunchKernelGGL(k_ilu0_sf<true>, dim3(blocks), dim3(KS_BLOCK), 0, ctx->stream, M->n, M->d_rowptr,
                     M->d_colidx, (const int32_t *)M->d_diag, (const double *)M->d_lu, y, M->d_slot, M->epoch,
                     M->d_flag, S, skip_sconv);
  M->epoch += 1;
  hipLaunchKernelGGL(k_ilu0_sf<false>, dim3(blocks), dim3(KS_BLOCK), 0, ctx->stream, M->n, M->d_rowptr,
                     M->d_colidx, (const int32_t *)M->d_diag, (const double *)M->d_lu, y, M->d_slot, M->epoch,
                     M->d_flag, S, skip_sconv);
}

// rows grouped by level (stable in row order) and the level starts
void group_levels(const std::vector<int> &lev, std::vector<int32_t> &rows, std::vector<int> &ptr) {
  int nl = 0;
  for (int v : lev) nl = v + 1 > nl ? v + 1 : nl;
  ptr.assign(nl + 1, 0);
  for (int v : lev) ptr[v + 1]++;
  for (int l = 0; l < nl; l++) ptr[l + 1] += ptr[l];
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  rows.resize(lev.size());
  for (size_t i = 0; i < lev.size(); i++) rows[fill[lev[i]]++] = (int32_t)i;
}

}  // namespace

int ks_ilu0_create(ks_ctx ctx, ks_csr A, double sigma, ks_ilu *out) {
  KS_CHECK(ctx && A && out, KS_ERR_ARG, "null ILU argument");
  KS_CHECK(A->vprec == KS_F64 && A->nrows == A->ncols, KS_ERR_ARG, "ILU(0) needs a square FP64 matrix");
  *out = nullptr;
  const int n = A->nrows;
  const int64_t nnz = A->nnz;
  std::vector<int32_t> rp(n + 1), ci(nnz), diag(n, -1);
  KS_HIP(hipMemcpyAsync(rp.data(), A->d_rowptr, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost, ctx->stream));
  KS_HIP(hipMemcpyAsync(ci.data(), A->d_colidx, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, ctx->stream));
  KS_HIP(hipStreamSynchronize(ctx->stream));
  // the pattern's rows must list increasing columns and hold their diagonal
  for (int i = 0; i < n; i++) {
    for (int32_t q = rp[i]; q < rp[i + 1]; q++) {
      KS_CHECK(q == rp[i] || ci[q] > ci[q - 1], KS_ERR_ARG, "ILU(0) needs increasing column indices (row %d)", i);
      if (ci[q] == i) diag[i] = q;
    }
    KS_CHECK(diag[i] >= 0, KS_ERR_ARG, "ILU(0) needs the diagonal entry of row %d", i);
  }
  // levels: forward (row i after the rows of its L part), backward (after its U part)
  std::vector<int> lev(n, 0);
  for (int i = 0; i < n; i++)
    for (int32_t q = rp[i]; q < diag[i]; q++) lev[i] = lev[ci[q]] + 1 > lev[i] ? lev[ci[q]] + 1 : lev[i];
  ks_ilu M = new ks_ilu_s();
  M->n = n;
  M->d_rowptr = A->d_rowptr;
  M->d_colidx = A->d_colidx;
  std::vector<int32_t> lrows, urows;
  group_levels(lev, lrows, M->lptr);
  std::fill(lev.begin(), lev.end(), 0);
  for (int i = n - 1; i >= 0; i--)
    for (int32_t q = diag[i] + 1; q < rp[i + 1]; q++) lev[i] = lev[ci[q]] + 1 > lev[i] ? lev[ci[q]] + 1 : lev[i];
  group_levels(lev, urows, M->uptr);
  int e = KS_OK;
  if (!e) e = ks_malloc((void **)&M->d_lrows, sizeof(int32_t) * n);
  if (!e) e = ks_malloc((void **)&M->d_diag, sizeof(int32_t) * n);
  if (!e) e = ks_malloc((void **)&M->d_lu, sizeof(double) * (nnz > 0 ? nnz : 1));
  if (!e) e = ks_malloc((void **)&M->d_flag, sizeof(int));
  if (!e) e = ks_malloc((void **)&M->d_slot, sizeof(unsigned long long) * 2 * (size_t)n);
  if (!e) e = ks_memset_zero(ctx, M->d_slot, sizeof(unsigned long long) * 2 * (size_t)n);
  if (e) {
    ks_ilu0_destroy(M);
    return e;
  }
  hipStream_t st = ctx->stream;
  KS_HIP(hipMemcpyAsync(M->d_lrows, lrows.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
  KS_HIP(hipMemcpyAsync(M->d_diag, diag.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
  KS_TRY(ks_memset_zero(ctx, M->d_flag, sizeof(int)));
  hipLaunchKernelGGL(k_ilu0_init, dim3(ksp_blocks(n)), dim3(KS_BLOCK), 0, st, A->d_rowptr, A->d_colidx,
                     (const double *)A->d_val, n, sigma, M->d_lu);
  for (size_t l = 0; l + 1 < M->lptr.size(); l++) {
    const int cnt = M->lptr[l + 1] - M->lptr[l];
    hipLaunchKernelGGL(k_ilu0_factor_level, dim3((cnt + 63) / 64), dim3(64), 0, st,
                       (const int32_t *)M->d_lrows + M->lptr[l], cnt, A->d_rowptr, A->d_colidx,
                       (const int32_t *)M->d_diag, M->d_lu, M->d_flag);
  }
  KS_HIP(hipGetLastError());
  // the last pivot, and a zero pivot met on the way
  int flag = 0;
  double ulast = 0.0;
  KS_HIP(hipMemcpyAsync(&flag, M->d_flag, sizeof(int), hipMemcpyDeviceToHost, st));
  KS_HIP(hipMemcpyAsync(&ulast, M->d_lu + diag[n - 1], sizeof(double), hipMemcpyDeviceToHost, st));
  KS_HIP(hipStreamSynchronize(st));
  if (flag || ulast == 0.0) {
    fprintf(stderr, "ks_hip: ILU(0) met a zero pivot\n");
    ks_ilu0_destroy(M);
    return KS_ERR_PIVOT;
  }
  *out = M;
  return KS_OK;
}

int ks_ilu0_destroy(ks_ilu M) {
  if (!M) return KS_OK;
  if (M->d_lrows) (void)ks_free(M->d_lrows);
  if (M->d_slot) (void)ks_free(M->d_slot);
  if (M->d_diag) (void)ks_free(M->d_diag);
  if (M->d_lu) (void)ks_free(M->d_lu);
  if (M->d_flag) (void)ks_free(M->d_flag);
  delete M;
  return KS_OK;
}

int ks_ilu0_levels(ks_ilu M, int *nlower, int *nupper) {
  KS_CHECK(M && nlower && nupper, KS_ERR_ARG, "null argument");
  *nlower = (int)M->lptr.size() - 1;
  *nupper = (int)M->uptr.size() - 1;
  return KS_OK;
}

int ks_pcg_jacobi(ks_ctx ctx, ks_csr A, double sigma, int jacobi, double *d_dinv) {
  KS_CHECK(ctx && A && d_dinv, KS_ERR_ARG, "null argument");
  KS_CHECK(A->vprec == KS_F64 && A->nrows == A->ncols, KS_ERR_ARG, "the inner solve needs a square FP64 matrix");
  hipLaunchKernelGGL(k_jacobi, dim3(ksp_blocks(A->nrows)), dim3(KS_BLOCK), 0, ctx->stream, A->d_rowptr, A->d_colidx,
                     (const double *)A->d_val, A->nrows, sigma, jacobi, d_dinv);
  KS_HIP(hipGetLastError());
  return KS_OK;
}

int ks_pcg_solve(ks_ctx ctx, ks_csr A, double sigma, const double *d_dinv, const double *d_b, double *d_x,
                 double rtol, double abstol, int maxits, double *d_work, int *its, double *rnorm, int *reason) {
  KS_CHECK(ctx && A && d_dinv && d_b && d_x && d_work && its && rnorm && reason, KS_ERR_ARG, "null PCG argument");
  KS_CHECK(A->vprec == KS_F64 && A->nrows == A->ncols, KS_ERR_ARG, "PCG needs a square FP64 matrix");
  KS_CHECK(maxits > 0, KS_ERR_ARG, "PCG needs maxits > 0");
  static_assert(3 * KS_PCG_BLOCKS <= (size_t)KS_MAXB * KS_NCOLP, "partials exceed the workspace");
  static_assert(sizeof(ksp_state) <= 64 * sizeof(double), "state exceeds the pinned scratch");
  const int n = A->nrows, nb = ksp_blocks(n);
  double *r = d_work, *z = r + n, *p = z + n, *q = p + n;
  ksp_state *S = (ksp_state *)(q + n);
  double *part = ctx->d_partial;
  hipStream_t st = ctx->stream;
  hipLaunchKernelGGL(k_pcg_start, dim3(nb), dim3(KS_BLOCK), 0, st, n, d_b, d_x, r, z, d_dinv, part);
  hipLaunchKernelGGL(k_pcg_fin_start, dim3(1), dim3(KS_BLOCK), 0, st, (const double *)part, nb, S, rtol, abstol,
                     maxits);
  for (int done = 0;; done += KS_KSP_CHUNK) {
    KS_TRY(read_state(ctx, S, its, reason, rnorm));
    if (*reason || done >= maxits) break;
    for (int c = 0; c < KS_KSP_CHUNK; c++) {
      hipLaunchKernelGGL(k_pcg_p, dim3(nb), dim3(KS_BLOCK), 0, st, n, (const double *)z, p, (const ksp_state *)S);
      KS_TRY(ks_spmv_gated(ctx, A, p, q, KS_F64, &S->gate));
      hipLaunchKernelGGL(k_pcg_q, dim3(nb), dim3(KS_BLOCK), 0, st, n, (const double *)p, q, sigma,
                         (const ksp_state *)S, part);
      hipLaunchKernelGGL(k_pcg_fin_pq, dim3(1), dim3(KS_BLOCK), 0, st, (const double *)part, nb, S);
      hipLaunchKernelGGL(k_pcg_xr, dim3(nb), dim3(KS_BLOCK), 0, st, n, (const double *)p, (const double *)q, d_x, r,
                         z, d_dinv, (const ksp_state *)S, part);
      hipLaunchKernelGGL(k_pcg_fin_iter, dim3(1), dim3(KS_BLOCK), 0, st, (const double *)part, nb, S);
    }
    KS_HIP(hipGetLastError());
  }
  if (!*reason) *reason = -3;   // KSP_DIVERGED_ITS
  return KS_OK;
}

namespace {

// BiCGStab with B = diag(d_dinv) (ilu null) or B = (L U)^-1 (ilu)
int bcgs_solve(ks_ctx ctx, ks_csr A, double sigma, const double *d_dinv, const ks_ilu ilu, const double *d_b,
               double *d_x, double rtol, double abstol, int maxits, double *d_work, int *its, double *rnorm,
               int *reason) {
  KS_CHECK(maxits > 0, KS_ERR_ARG, "BiCGStab needs maxits > 0");
  const int n = A->nrows, nb = ksp_blocks(n);
  double *r = d_work, *rh = r + n, *p = rh + n, *v = p + n, *sv = v + n, *t = sv + n;
  ksp_state *S = (ksp_state *)(t + n);
  double *part = ctx->d_partial;
  hipStream_t st = ctx->stream;
  const dim3 g(nb), b1(KS_BLOCK), one(1);
  if (ilu) {
    KS_TRY(ks_memcpy_dtod(ctx, r, d_b, sizeof(double) * n));
    ilu0_apply(ctx, ilu, r, nullptr, 0);                                             // r = B b
    hipLaunchKernelGGL(k_bcgs_start, g, b1, 0, st, n, (const double *)r, d_x, r, rh, p, v, (const double *)nullptr,
                       part);
  } else {
    hipLaunchKernelGGL(k_bcgs_start, g, b1, 0, st, n, d_b, d_x, r, rh, p, v, d_dinv, part);
  }
  hipLaunchKernelGGL(k_bcgs_fin_start, one, b1, 0, st, (const double *)part, nb, S, rtol, abstol, maxits);
  const double *dv = ilu ? nullptr : d_dinv;
  for (int done = 0;; done += KS_KSP_CHUNK) {
    KS_TRY(read_state(ctx, S, its, reason, rnorm));
    if (*reason || done >= maxits) break;
    for (int c = 0; c < KS_KSP_CHUNK; c++) {
      hipLaunchKernelGGL(k_bcgs_p, g, b1, 0, st, n, (const double *)r, p, (const double *)v, (const ksp_state *)S);
      KS_TRY(ks_spmv_gated(ctx, A, p, v, KS_F64, &S->gate));                         // v = A p
      if (ilu) {
        hipLaunchKernelGGL(k_shift_gated, g, b1, 0, st, n, (const double *)p, v, sigma, (const ksp_state *)S, 0);
        ilu0_apply(ctx, ilu, v, S, 0);
      }
      hipLaunchKernelGGL(k_bcgs_apply, g, b1, 0, st, n, (const double *)p, v, (const double *)rh, dv, sigma,
                         (const ksp_state *)S, 0, part);                               // v = B (A - sigma I) p
      hipLaunchKernelGGL(k_bcgs_fin_alpha, one, b1, 0, st, (const double *)part, nb, S);
      hipLaunchKernelGGL(k_bcgs_s, g, b1, 0, st, n, (const double *)r, (const double *)v, sv, (const ksp_state *)S,
                         part);
      hipLaunchKernelGGL(k_bcgs_fin_s, one, b1, 0, st, (const double *)part, nb, S);
      KS_TRY(ks_spmv_gated(ctx, A, sv, t, KS_F64, &S->gate));                        // t = A s
      if (ilu) {
        hipLaunchKernelGGL(k_shift_gated, g, b1, 0, st, n, (const double *)sv, t, sigma, (const ksp_state *)S, 1);
        ilu0_apply(ctx, ilu, t, S, 1);
      }
      hipLaunchKernelGGL(k_bcgs_apply, g, b1, 0, st, n, (const double *)sv, t, (const double *)rh, dv, sigma,
                         (const ksp_state *)S, 1, part);                               // t = B (A - sigma I) s
      hipLaunchKernelGGL(k_bcgs_fin_omega, one, b1, 0, st, (const double *)part, nb, S);
      hipLaunchKernelGGL(k_bcgs_x, g, b1, 0, st, n, d_x, (const double *)p, (const double *)sv, (const double *)t, r,
                         (const double *)rh, (const ksp_state *)S, part);
      hipLaunchKernelGGL(k_bcgs_fin_iter, one, b1, 0, st, (const double *)part, nb, S);
    }
    KS_HIP(hipGetLastError());
  }
  if (!*reason) *reason = -3;
  return KS_OK;
}

}  // namespace

int ks_bcgs_solve(ks_ctx ctx, ks_csr A, double sigma, const double *d_dinv, const double *d_b, double *d_x,
                  double rtol, double abstol, int maxits, double *d_work, int *its, double *rnorm, int *reason) {
  KS_CHECK(ctx && A && d_dinv && d_b && d_x && d_work && its && rnorm && reason, KS_ERR_ARG,
           "null BiCGStab argument");
  KS_CHECK(A->vprec == KS_F64 && A->nrows == A->ncols, KS_ERR_ARG, "BiCGStab needs a square FP64 matrix");
  return bcgs_solve(ctx, A, sigma, d_dinv, nullptr, d_b, d_x, rtol, abstol, maxits, d_work, its, rnorm, reason);
}

int ks_bcgs_solve_ilu(ks_ctx ctx, ks_csr A, double sigma, ks_ilu M, const double *d_b, double *d_x, double rtol,
                      double abstol, int maxits, double *d_work, int *its, double *rnorm, int *reason) {
  KS_CHECK(ctx && A && M && d_b && d_x && d_work && its && rnorm && reason, KS_ERR_ARG,
           "null BiCGStab argument");
  KS_CHECK(A->vprec == KS_F64 && A->nrows == A->ncols && M->n == A->nrows, KS_ERR_ARG,
           "BiCGStab with ILU(0) needs the square FP64 matrix the factors were made from");
  KS_CHECK(M->d_rowptr == A->d_rowptr, KS_ERR_ARG, "the ILU(0) factors belong to another matrix");
  const int e = bcgs_solve(ctx, A, sigma, nullptr, M, d_b, d_x, rtol, abstol, maxits, d_work, its, rnorm, reason);
  if (e) return e;
  // the triangular solves' waits all met (bit 1 of the flag word); the solve has synchronized
  int flag = 0;
  KS_HIP(hipMemcpyAsync(&flag, M->d_flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  KS_HIP(hipStreamSynchronize(ctx->stream));
  KS_CHECK(!(flag & 2), KS_ERR_TIMEOUT, "an ILU(0) triangular solve's wait ran out of polls");
  return KS_OK;
}
