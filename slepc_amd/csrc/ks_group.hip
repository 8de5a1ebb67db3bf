// ks_group.hip -- grouped Arnoldi steps for serving: the SpMV and the whole
// CGS+DGKS orthonormalization of the current column of several independent
// solves, each kind of kernel launched once for all of them on the group's
// stream.  For small solves the cost of a step is launch overhead, not bytes;
// grouping divides it by the number of members.
#include "ks_internal.hpp"

#include <cstring>

struct ks_group_s {
  hipStream_t stream = nullptr;
  static constexpr int RING = 256;
  hipEvent_t ev[RING] = {};
  int head = 0, tail = 0;           // steps in flight: tail .. head-1 (mod RING)
  hipEvent_t xev[RING] = {};        // cross-stream markers, reused in turn
  int xnext = 0;
  ks_grp_spmv sp;
  ks_grp_gs gs;
};

extern "C" {

int ks_group_create(ks_group *out) {
  KS_CHECK(out, KS_ERR_ARG, "null group pointer");
  auto *g = new ks_group_s();
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    delete g;
    return KS_ERR_HIP;
  }
  for (int i = 0; i < ks_group_s::RING; i++) {
    if (hipEventCreateWithFlags(&g->ev[i], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&g->xev[i], hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      for (int j = 0; j <= i; j++) {
        if (g->ev[j]) (void)hipEventDestroy(g->ev[j]);
        if (g->xev[j]) (void)hipEventDestroy(g->xev[j]);
      }
      (void)hipStreamDestroy(g->stream);
      delete g;
      return KS_ERR_HIP;
    }
  }
  *out = g;
  return KS_OK;
}

int ks_group_destroy(ks_group g) {
  if (!g) return KS_OK;
  KS_HIP(hipStreamSynchronize(g->stream));
  for (int i = 0; i < ks_group_s::RING; i++) {
    (void)hipEventDestroy(g->ev[i]);
    (void)hipEventDestroy(g->xev[i]);
  }
  (void)hipStreamDestroy(g->stream);
  delete g;
  return KS_OK;
}

int ks_group_step(ks_group g, int nmem, const ks_group_member *mem, ks_refine refine, double eta) {
  KS_CHECK(g && (nmem == 0 || mem), KS_ERR_ARG, "null group argument");
  KS_CHECK(nmem >= 0 && nmem <= KS_GROUP_MAX, KS_ERR_RANGE, "%d members (at most %d per step)", nmem, KS_GROUP_MAX);
  KS_CHECK(refine >= KS_REFINE_IFNEEDED && refine <= KS_REFINE_ALWAYS, KS_ERR_ARG, "bad refinement type");
  if (nmem == 0) return KS_OK;
  const int slot = g->head % ks_group_s::RING;
  KS_CHECK(g->head - g->tail < ks_group_s::RING, KS_ERR_RANGE, "too many grouped steps in flight");
  ks_grp_spmv &sp = g->sp;
  ks_grp_gs &gs = g->gs;
  std::memset(&sp, 0, sizeof(sp));
  std::memset(&gs, 0, sizeof(gs));
  sp.nmem = gs.nmem = nmem;
  for (int i = 0; i < nmem; i++) {
    const ks_group_member &M = mem[i];
    KS_CHECK(M.ctx && M.A && M.V && M.buf, KS_ERR_ARG, "member %d: null pointer", i);
    KS_CHECK(M.n > 0 && M.ld >= M.n && M.k >= 1 && M.k < KS_MAX_COLS && M.ldbuf > M.hcol && M.hcol >= 0,
             KS_ERR_ARG, "member %d: bad sizes n=%d ld=%d k=%d", i, M.n, M.ld, M.k);
    const int32_t *blk, *rowptr, *colidx;
    const void *val;
    int nblocks;
    ks_prec vprec;
    ks_csr_blocks(M.A, &blk, &rowptr, &colidx, &val, &nblocks, &vprec);
    KS_CHECK(vprec == KS_F64, KS_ERR_ARG, "member %d: grouped steps take FP64 matrices", i);
    int mcw, rpb, nb;
    int e = ks_gs_split_geometry(M.n, M.k, &mcw, &rpb, &nb);
    if (e) return e;
    KS_CHECK(nb <= KS_MAXB, KS_ERR_RANGE, "member %d: vector length %d too long", i, M.n);
    const double *V = (const double *)M.V;
    double *x = (double *)M.V + (size_t)(M.k - 1) * (size_t)M.ld;
    double *w = (double *)M.V + (size_t)M.k * (size_t)M.ld;
    sp.b0[i + 1] = sp.b0[i] + nblocks;
    sp.blk[i] = blk;
    sp.rowptr[i] = rowptr;
    sp.colidx[i] = colidx;
    sp.val[i] = (const double *)val;
    sp.x[i] = x;
    sp.y[i] = w;
    sp.st[i] = M.ctx->d_status;
    ks_grp_gs_member &G = gs.m[i];
    G.V = V;
    G.w = w;
    G.cbuf = M.buf;                                        // buffer column 0: scratch coefficients
    G.hcol = M.buf + (size_t)M.hcol * (size_t)M.ldbuf;     // the column's h
    G.partial = M.ctx->d_partial;
    G.sums = M.ctx->d_gpartial;
    G.st = M.ctx->d_status;
    G.counter = M.ctx->d_counters;
    G.pub = M.ctx->d_finpub;
    G.coef1 = nullptr;   // grouped passes store every pass's vector (their sweep kernel takes no re-forming pass)
    G.ld = M.ld;
    G.n = M.n;
    G.k = M.k;
    G.rpb = rpb;
    G.nb = nb;
    G.col = M.col;
    G.cw = mcw;
    gs.b0_gs[i + 1] = gs.b0_gs[i] + nb;
    gs.b0_fin[i + 1] = gs.b0_fin[i] + M.k + (G.pub ? 2 : 1);
    gs.b0_sc[i + 1] = gs.b0_sc[i] + ks_ew_blocks(M.n);
  }
  ks_group_launch_spmv(g->stream, sp);
  ks_group_launch_gs(g->stream, gs, (int)refine, eta);
  KS_HIP(hipGetLastError());
  KS_HIP(hipEventRecord(g->ev[slot], g->stream));
  g->head++;
  return KS_OK;
}

int ks_group_inflight(ks_group g, int *steps) {
  KS_CHECK(g && steps, KS_ERR_ARG, "null argument");
  while (g->tail < g->head) {
    hipError_t e = hipEventQuery(g->ev[g->tail % ks_group_s::RING]);
    if (e == hipErrorNotReady) {
      (void)hipGetLastError();
      break;
    }
    KS_HIP(e);
    g->tail++;
  }
  *steps = g->head - g->tail;
  return KS_OK;
}

// a wait enqueued on a stream refers to the event's record at the time of the
// call, so the markers can be recorded again for later waits
int ks_group_wait_ctx(ks_group g, ks_ctx c) {
  KS_CHECK(g && c, KS_ERR_ARG, "null argument");
  hipEvent_t e = g->xev[g->xnext++ % ks_group_s::RING];
  KS_HIP(hipEventRecord(e, c->stream));
  KS_HIP(hipStreamWaitEvent(g->stream, e, 0));
  return KS_OK;
}

int ks_ctx_wait_group(ks_ctx c, ks_group g) {
  KS_CHECK(g && c, KS_ERR_ARG, "null argument");
  hipEvent_t e = g->xev[g->xnext++ % ks_group_s::RING];
  KS_HIP(hipEventRecord(e, g->stream));
  KS_HIP(hipStreamWaitEvent(c->stream, e, 0));
  return KS_OK;
}

}  // extern "C"
