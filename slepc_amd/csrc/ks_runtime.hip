// ks_runtime.hip -- device selection, contexts, memory, status and profiling
// of libks_hip.so.
#include "ks_internal.hpp"

#include <cstring>
#include <map>
#include <vector>

namespace {

__global__ void k_fill_zero(unsigned char *p, size_t bytes) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < bytes; i += stride) p[i] = 0;
}

__global__ void k_fill_zero8(uint64_t *p, size_t words) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < words; i += stride) p[i] = 0;
}

__global__ void k_status_reset(ks_status *st) {
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
    st->defer_col = -1;
    st->halt_need = 0;
  }
}

bool g_initialized = false;

// Pools: a device arena (and optionally a pinned host arena) carved by a
// first-fit allocator, so a serving loop admits and retires solves without
// hipMalloc/hipFree (hipFree waits for the whole device).  ks_malloc draws from
// the pool made current by ks_pool_use; ks_free returns a pointer to the pool
// whose range holds it.
std::vector<ks_pool> g_pools;
ks_pool g_pool_current = nullptr;

}  // namespace

extern "C" {

const char *ks_error_string(int code) {
  switch (code) {
    case KS_OK: return "ok";
    case KS_ERR_ARG: return "invalid argument";
    case KS_ERR_HIP: return "HIP runtime error";
    case KS_ERR_NOMEM: return "device out of memory";
    case KS_ERR_NODEV: return "no gfx950 device";
    case KS_ERR_INNER: return "invalid inner product";
    case KS_ERR_RANGE: return "size beyond supported maximum";
    case KS_ERR_PIVOT: return "zero pivot";
    case KS_ERR_TIMEOUT: return "a bounded device wait ran out";
  }
  return "unknown error";
}

int ks_init(int device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    fprintf(stderr, "ks_hip: no HIP device visible\n");
    return KS_ERR_NODEV;
  }
  KS_CHECK(device >= 0 && device < count, KS_ERR_ARG, "device %d out of range (%d devices)", device, count);
  KS_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  KS_HIP(hipGetDeviceProperties(&prop, device));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    fprintf(stderr, "ks_hip: device %d is %s, this library is built for gfx950\n", device, prop.gcnArchName);
    return KS_ERR_NODEV;
  }
  g_initialized = true;
  return KS_OK;
}

int ks_ctx_create(ks_ctx *out) {
  KS_CHECK(out, KS_ERR_ARG, "null context pointer");
  if (!g_initialized) {
    int e = ks_init(0);
    if (e) return e;
  }
  ks_ctx c = new ks_ctx_s();
  KS_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  KS_HIP(hipMalloc(&c->d_status, sizeof(ks_status)));
  KS_HIP(hipMalloc(&c->d_partial, sizeof(double) * (size_t)KS_MAXB * KS_NCOLP));
  KS_HIP(hipMalloc(&c->d_gpartial, sizeof(double) * (size_t)KS_MAXG * KS_NCOLP));
  KS_HIP(hipMalloc(&c->d_counters, sizeof(unsigned) * (KS_MAXG + 1)));
  if (!getenv("KS_GS_NODEFER")) KS_HIP(hipMalloc(&c->d_coef1, sizeof(double) * KS_NCOLP));
  {
    const char *t = getenv("KS_GS_FIN_TICKET");   // 1: finishing launches count with a ticket instead
    if (!(t && t[0] == '1')) KS_HIP(hipMalloc(&c->d_finpub, sizeof(unsigned long long) * (1 + 2 * KS_NCOLP)));
  }
  KS_HIP(hipMalloc(&c->d_scalar, sizeof(double) * 64));
  KS_HIP(hipHostMalloc(&c->h_scalar, sizeof(double) * 64, hipHostMallocDefault));
  KS_HIP(hipHostMalloc((void **)&c->h_status, sizeof(ks_status), hipHostMallocDefault));
  int e = ks_memset_zero(c, c->d_counters, sizeof(unsigned) * (KS_MAXG + 1));
  if (e) return e;
  if (c->d_finpub && (e = ks_memset_zero(c, c->d_finpub, sizeof(unsigned long long) * (1 + 2 * KS_NCOLP)))) return e;
  hipLaunchKernelGGL(k_status_reset, dim3(1), dim3(64), 0, c->stream, c->d_status);
  KS_HIP(hipGetLastError());
  KS_HIP(hipStreamSynchronize(c->stream));
  *out = c;
  return KS_OK;
}

int ks_ctx_destroy(ks_ctx c) {
  if (!c) return KS_OK;
  (void)hipStreamSynchronize(c->stream);
  for (auto &r : c->recs) { (void)hipEventDestroy(r.ev0); (void)hipEventDestroy(r.ev1); }
  for (auto ev : c->event_pool) (void)hipEventDestroy(ev);
  (void)hipFree(c->d_status);
  (void)hipFree(c->d_partial);
  (void)hipFree(c->d_gpartial);
  if (c->d_finpub) (void)hipFree(c->d_finpub);
  if (c->d_coef1) (void)hipFree(c->d_coef1);
  (void)hipFree(c->d_counters);
  (void)hipFree(c->d_scalar);
  if (c->d_ran) (void)hipFree(c->d_ran);
  (void)hipHostFree(c->h_scalar);
  (void)hipHostFree(c->h_status);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return KS_OK;
}

int ks_ctx_sync(ks_ctx c) {
  KS_CHECK(c, KS_ERR_ARG, "null context");
  KS_HIP(hipStreamSynchronize(c->stream));
  return KS_OK;
}

void *ks_ctx_stream(ks_ctx c) { return c ? (void *)c->stream : nullptr; }

int ks_malloc(void **ptr, size_t bytes) {
  KS_CHECK(ptr, KS_ERR_ARG, "null pointer");
  *ptr = nullptr;
  if (bytes == 0) bytes = 256;
  if (g_pool_current) {
    *ptr = g_pool_current->dev.take(bytes);
    if (!*ptr) {
      fprintf(stderr, "ks_hip: device pool exhausted (%zu bytes asked, %zu of %zu in use)\n", bytes,
              g_pool_current->dev.used, g_pool_current->dev.size);
      return KS_ERR_NOMEM;
    }
    return KS_OK;
  }
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fprintf(stderr, "ks_hip: hipMalloc(%zu) failed: %s\n", bytes, hipGetErrorString(e));
    return KS_ERR_NOMEM;
  }
  return KS_OK;
}

int ks_free(void *ptr) {
  if (!ptr) return KS_OK;
  for (ks_pool p : g_pools)
    if (p->dev.owns(ptr)) return p->dev.give(ptr) ? KS_OK : KS_ERR_ARG;
  KS_HIP(hipFree(ptr));
  return KS_OK;
}

int ks_memcpy_htod(ks_ctx c, void *dst, const void *src, size_t bytes) {
  KS_CHECK(c && (dst || !bytes) && (src || !bytes), KS_ERR_ARG, "bad memcpy arguments");
  if (!bytes) return KS_OK;
  KS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  KS_HIP(hipStreamSynchronize(c->stream));
  return KS_OK;
}

int ks_memcpy_htod_async(ks_ctx c, void *dst, const void *src, size_t bytes) {
  KS_CHECK(c && (dst || !bytes) && (src || !bytes), KS_ERR_ARG, "bad memcpy arguments");
  if (!bytes) return KS_OK;
  KS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  return KS_OK;
}

int ks_host_alloc(void **ptr, size_t bytes) {
  KS_CHECK(ptr, KS_ERR_ARG, "null pointer");
  *ptr = nullptr;
  if (g_pool_current && g_pool_current->host.size) {
    *ptr = g_pool_current->host.take(bytes ? bytes : 64);
    if (!*ptr) {
      fprintf(stderr, "ks_hip: pinned host pool exhausted (%zu bytes asked)\n", bytes);
      return KS_ERR_NOMEM;
    }
    return KS_OK;
  }
  KS_HIP(hipHostMalloc(ptr, bytes ? bytes : 64, hipHostMallocDefault));
  return KS_OK;
}

int ks_host_free(void *ptr) {
  if (!ptr) return KS_OK;
  for (ks_pool p : g_pools)
    if (p->host.owns(ptr)) return p->host.give(ptr) ? KS_OK : KS_ERR_ARG;
  KS_HIP(hipHostFree(ptr));
  return KS_OK;
}

/* ---------------------------------------------------------------- pools */
int ks_pool_create(ks_pool *out, size_t dev_bytes, size_t host_bytes) {
  KS_CHECK(out && dev_bytes > 0, KS_ERR_ARG, "bad pool arguments");
  *out = nullptr;
  auto *p = new ks_pool_s();
  void *d = nullptr, *h = nullptr;
  dev_bytes = (dev_bytes + KS_POOL_ALIGN - 1) / KS_POOL_ALIGN * KS_POOL_ALIGN;
  host_bytes = (host_bytes + KS_POOL_ALIGN - 1) / KS_POOL_ALIGN * KS_POOL_ALIGN;
  if (hipMalloc(&d, dev_bytes) != hipSuccess) {
    (void)hipGetLastError();
    delete p;
    fprintf(stderr, "ks_hip: pool hipMalloc(%zu) failed\n", dev_bytes);
    return KS_ERR_NOMEM;
  }
  if (host_bytes && hipHostMalloc(&h, host_bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(d);
    delete p;
    return KS_ERR_NOMEM;
  }
  p->dev.init((char *)d, dev_bytes);
  p->host.init((char *)h, host_bytes);
  g_pools.push_back(p);
  *out = p;
  return KS_OK;
}

int ks_pool_destroy(ks_pool p) {
  if (!p) return KS_OK;
  KS_CHECK(p->dev.used == 0 && p->host.used == 0, KS_ERR_ARG,
           "pool destroyed with %zu device / %zu host bytes still allocated", p->dev.used, p->host.used);
  if (g_pool_current == p) g_pool_current = nullptr;
  for (size_t i = 0; i < g_pools.size(); i++)
    if (g_pools[i] == p) { g_pools.erase(g_pools.begin() + (long)i); break; }
  KS_HIP(hipDeviceSynchronize());
  (void)hipFree(p->dev.base);
  if (p->host.base) (void)hipHostFree(p->host.base);
  delete p;
  return KS_OK;
}

ks_pool ks_pool_use(ks_pool p) {
  ks_pool prev = g_pool_current;
  g_pool_current = p;
  return prev;
}

int ks_pool_info(ks_pool p, size_t *used, size_t *peak, size_t *size, size_t *largest_free) {
  KS_CHECK(p, KS_ERR_ARG, "null pool");
  if (used) *used = p->dev.used;
  if (peak) *peak = p->dev.peak;
  if (size) *size = p->dev.size;
  if (largest_free) *largest_free = p->dev.largest();
  return KS_OK;
}

int ks_ctx_query(ks_ctx c, int *idle) {
  KS_CHECK(c && idle, KS_ERR_ARG, "bad query arguments");
  hipError_t e = hipStreamQuery(c->stream);
  if (e == hipSuccess) { *idle = 1; return KS_OK; }
  if (e == hipErrorNotReady) { (void)hipGetLastError(); *idle = 0; return KS_OK; }
  fprintf(stderr, "ks_hip: hipStreamQuery: %s\n", hipGetErrorString(e));
  return KS_ERR_HIP;
}

int ks_memcpy_dtoh(ks_ctx c, void *dst, const void *src, size_t bytes) {
  KS_CHECK(c && (dst || !bytes) && (src || !bytes), KS_ERR_ARG, "bad memcpy arguments");
  if (!bytes) return KS_OK;
  KS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  KS_HIP(hipStreamSynchronize(c->stream));
  return KS_OK;
}

int ks_memcpy_dtoh_async(ks_ctx c, void *dst, const void *src, size_t bytes) {
  KS_CHECK(c && (dst || !bytes) && (src || !bytes), KS_ERR_ARG, "bad memcpy arguments");
  if (!bytes) return KS_OK;
  KS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  return KS_OK;
}

int ks_memcpy_dtod(ks_ctx c, void *dst, const void *src, size_t bytes) {
  KS_CHECK(c && (dst || !bytes) && (src || !bytes), KS_ERR_ARG, "bad memcpy arguments");
  if (!bytes) return KS_OK;
  KS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
  return KS_OK;
}

int ks_memset_zero(ks_ctx c, void *dst, size_t bytes) {
  KS_CHECK(c && (dst || !bytes), KS_ERR_ARG, "bad memset arguments");
  if (!bytes) return KS_OK;
  if ((bytes % 8) == 0 && ((uintptr_t)dst % 8) == 0) {
    size_t words = bytes / 8;
    int blocks = (int)((words + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_zero8, dim3(blocks), dim3(256), 0, c->stream, (uint64_t *)dst, words);
  } else {
    int blocks = (int)((bytes + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(256), 0, c->stream, (unsigned char *)dst, bytes);
  }
  KS_HIP(hipGetLastError());
  return KS_OK;
}

int ks_mem_info(size_t *free_bytes, size_t *total_bytes) {
  size_t f = 0, t = 0;
  KS_HIP(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return KS_OK;
}

int ks_status_reset(ks_ctx c) {
  KS_CHECK(c, KS_ERR_ARG, "null context");
  hipLaunchKernelGGL(k_status_reset, dim3(1), dim3(64), 0, c->stream, c->d_status);
  KS_HIP(hipGetLastError());
  return KS_OK;
}

int ks_status_get(ks_ctx c, ks_status *out) {
  KS_CHECK(c && out, KS_ERR_ARG, "null argument");
  KS_HIP(hipMemcpyAsync(c->h_scalar, c->d_status, sizeof(ks_status), hipMemcpyDeviceToHost, c->stream));
  KS_HIP(hipStreamSynchronize(c->stream));
  memcpy(out, c->h_scalar, sizeof(ks_status));
  return KS_OK;
}

// the status copied to pinned host memory in stream order (no wait here); a
// later ks_status_read waits for the stream and returns that copy
int ks_status_post(ks_ctx c) {
  KS_CHECK(c, KS_ERR_ARG, "null context");
  KS_HIP(hipMemcpyAsync(c->h_status, c->d_status, sizeof(ks_status), hipMemcpyDeviceToHost, c->stream));
  return KS_OK;
}

int ks_status_read(ks_ctx c, ks_status *out) {
  KS_CHECK(c && out, KS_ERR_ARG, "null argument");
  KS_HIP(hipStreamSynchronize(c->stream));
  memcpy(out, c->h_status, sizeof(ks_status));
  return KS_OK;
}

// ---------------------------------------------------------------- captured launch sequences
struct ks_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

int ks_graph_begin(ks_ctx c) {
  KS_CHECK(c, KS_ERR_ARG, "null context");
  KS_CHECK(!c->capturing, KS_ERR_ARG, "capture already open on this context");
  KS_CHECK(!c->prof, KS_ERR_ARG, "no capture while per-kernel events are on");
  KS_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  return KS_OK;
}

int ks_graph_end(ks_ctx c, ks_graph *out) {
  KS_CHECK(c && out, KS_ERR_ARG, "null argument");
  KS_CHECK(c->capturing, KS_ERR_ARG, "no capture open on this context");
  *out = nullptr;
  c->capturing = false;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (e != hipSuccess || !g) {
    (void)hipGetLastError();
    if (g) (void)hipGraphDestroy(g);
    fprintf(stderr, "ks_hip: hipStreamEndCapture: %s\n", hipGetErrorString(e));
    return KS_ERR_HIP;
  }
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    (void)hipGraphDestroy(g);
    fprintf(stderr, "ks_hip: hipGraphInstantiate: %s\n", hipGetErrorString(e));
    return KS_ERR_HIP;
  }
  *out = new ks_graph_s{g, x};
  return KS_OK;
}

int ks_graph_launch(ks_ctx c, ks_graph g) {
  KS_CHECK(c && g, KS_ERR_ARG, "null argument");
  KS_CHECK(!c->capturing, KS_ERR_ARG, "graph launch inside a capture");
  KS_HIP(hipGraphLaunch(g->exec, c->stream));
  return KS_OK;
}

int ks_graph_destroy(ks_graph g) {
  if (!g) return KS_OK;
  (void)hipGraphExecDestroy(g->exec);
  (void)hipGraphDestroy(g->graph);
  delete g;
  return KS_OK;
}

// ---------------------------------------------------------------- profiling
static const char *g_prof_names[KS_PROF_NCLASS] = {
    "spmv", "gs_dot_first", "gs_update_first", "gs_finish", "gs_update_refine", "scale",
    "norm", "axpy", "copy", "multinplace", "dotvec", "multvec", "other"};

const char *ks_prof_name(int cls) {
  return (cls >= 0 && cls < KS_PROF_NCLASS) ? g_prof_names[cls] : "?";
}

static hipEvent_t pool_get(ks_ctx c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int *ks_prof_begin(ks_ctx c, int cls, double bytes, double flops, bool gated) {
  c->prof_open = false;
  if (!c->prof || c->capturing || !((c->prof_mask >> cls) & 1u)) return nullptr;
  if (c->recs.size() >= 65536 || (gated && c->ran_used >= c->ran_cap)) ks_prof_collect(c);
  ks_prof_rec r;
  r.cls = cls;
  r.bytes = bytes;
  r.flops = flops;
  r.ran_slot = gated ? c->ran_used++ : -1;
  r.ev0 = pool_get(c);
  r.ev1 = pool_get(c);
  (void)hipEventRecord(r.ev0, c->stream);
  c->recs.push_back(r);
  c->prof_open = true;
  return gated ? c->d_ran + r.ran_slot : nullptr;
}

void ks_prof_end(ks_ctx c) {
  if (!c->prof_open) return;
  (void)hipEventRecord(c->recs.back().ev1, c->stream);
  c->prof_open = false;
}

int ks_prof_enable(ks_ctx c, int on) {
  KS_CHECK(c, KS_ERR_ARG, "null context");
  KS_CHECK(!(on && c->capturing), KS_ERR_ARG, "no per-kernel events inside a capture");
  if (on && !c->d_ran) {
    c->ran_cap = 65536;
    KS_HIP(hipMalloc(&c->d_ran, sizeof(int) * c->ran_cap));
    int e = ks_memset_zero(c, c->d_ran, sizeof(int) * c->ran_cap);
    if (e) return e;
  }
  if (!on && c->prof) ks_prof_collect(c);
  c->prof = on != 0;
  c->prof_mask = ~0u;
  return KS_OK;
}

// events only around the launches of the classes in mask (bit cls): each
// event pair costs the stream a few microseconds, so a timed run can keep the
// one class it reports and leave the rest of the stream undisturbed
int ks_prof_enable_classes(ks_ctx c, unsigned mask) {
  int e = ks_prof_enable(c, mask != 0u);
  if (e) return e;
  c->prof_mask = mask;
  return KS_OK;
}

int ks_prof_collect(ks_ctx c) {
  KS_CHECK(c, KS_ERR_ARG, "null context");
  if (c->recs.empty()) return KS_OK;
  KS_HIP(hipStreamSynchronize(c->stream));
  std::vector<int> ran(c->ran_used > 0 ? c->ran_used : 1, 1);
  if (c->ran_used > 0)
    KS_HIP(hipMemcpy(ran.data(), c->d_ran, sizeof(int) * c->ran_used, hipMemcpyDeviceToHost));
  for (auto &r : c->recs) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r.ev0, r.ev1);
    bool did = r.ran_slot < 0 || ran[r.ran_slot] != 0;
    ks_prof_entry &a = c->acc[r.cls];
    if (did) {
      a.launches++;
      a.ms += ms;
      a.bytes += r.bytes;
      a.flops += r.flops;
    } else {
      a.skipped++;
    }
    c->event_pool.push_back(r.ev0);
    c->event_pool.push_back(r.ev1);
  }
  c->recs.clear();
  if (c->ran_used > 0) {
    int e = ks_memset_zero(c, c->d_ran, sizeof(int) * c->ran_used);
    if (e) return e;
  }
  c->ran_used = 0;
  return KS_OK;
}

int ks_prof_get(ks_ctx c, int cls, ks_prof_entry *out) {
  KS_CHECK(c && out && cls >= 0 && cls < KS_PROF_NCLASS, KS_ERR_ARG, "bad profiling query");
  *out = c->acc[cls];
  return KS_OK;
}

int ks_prof_reset(ks_ctx c) {
  KS_CHECK(c, KS_ERR_ARG, "null context");
  ks_prof_collect(c);
  for (auto &a : c->acc) a = ks_prof_entry{};
  return KS_OK;
}

}  // extern "C"
