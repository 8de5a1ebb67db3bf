/*
 * rmat.c -- workload generator for the at-scale graph eigenproblem: the
 * shifted normalized Laplacian M = 2I - (I - D^-1/2 A D^-1/2) = I + D^-1/2 A D^-1/2
 * of a seeded R-MAT graph, as CSR with FP32 values and int32 indices.
 *
 * Randomness is counter based, so the graph does not depend on the thread
 * count: edge e draws r_g = splitmix64(seed * 0x9E3779B97F4A7C15 + e * G + g)
 * for g = 0 .. G-1, G = ceil(scale / 4), and recursion level l = 4g + t takes
 * the 16-bit fraction x = (r_g >> 16t) & 0xffff: x < A picks quadrant (0,0),
 * x < A+B (0,1), x < A+B+C (1,0), else (1,1), with A = round(65536 a) etc.;
 * level l sets bit l of the source (row half) and destination (column half).
 * The edge list is symmetrized, self-loops and duplicates dropped; isolated
 * vertices get M_ii = 1.  slepc_amd/matgen.py restates the same steps in
 * numpy (the cross-check at small scales).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

static inline uint64_t splitmix64(uint64_t x)
{
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* LSD radix sort of 64-bit keys on their low `bits` bits, 8 bits per pass;
   tmp has n entries; the result ends in keys */
static void radix_sort(uint64_t *keys, uint64_t *tmp, int64_t n, int bits)
{
  const int passes = (bits + 7) / 8;
  const int nt = omp_get_max_threads();
  int64_t *cnt = (int64_t *)malloc(sizeof(int64_t) * 256 * (size_t)nt);
  uint64_t *src = keys, *dst = tmp;
  for (int p = 0; p < passes; p++) {
    const int sh = 8 * p;
    memset(cnt, 0, sizeof(int64_t) * 256 * (size_t)nt);
#pragma omp parallel num_threads(nt)
    {
      const int t = omp_get_thread_num();
      const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
      int64_t *c = cnt + 256 * (size_t)t;
      for (int64_t i = lo; i < hi; i++) c[(src[i] >> sh) & 255]++;
#pragma omp barrier
#pragma omp single
      {
        int64_t off = 0;
        for (int d = 0; d < 256; d++)
          for (int q = 0; q < nt; q++) {
            const int64_t v = cnt[256 * (size_t)q + d];
            cnt[256 * (size_t)q + d] = off;
            off += v;
          }
      }
      for (int64_t i = lo; i < hi; i++) dst[c[(src[i] >> sh) & 255]++] = src[i];
    }
    uint64_t *x = src;
    src = dst;
    dst = x;
  }
  if (src != keys) memcpy(keys, src, sizeof(uint64_t) * (size_t)n);
  free(cnt);
}

/* returns 0 on success, -1 on bad arguments, -2 when out of memory */
int ks_gen_rmat_laplacian(int scale, int edge_factor, uint64_t seed, double a, double b, double c,
                          int32_t **rowptr_out, int32_t **colidx_out, float **val_out, int64_t *nnz_out)
{
  if (scale < 1 || scale > 30 || edge_factor < 1 || !rowptr_out || !colidx_out || !val_out || !nnz_out) return -1;
  const int64_t n = (int64_t)1 << scale;
  const int64_t m = (int64_t)edge_factor << scale;
  const int G = (scale + 3) / 4;
  const uint32_t TA = (uint32_t)lrint(a * 65536.0), TB = (uint32_t)lrint((a + b) * 65536.0),
                 TC = (uint32_t)lrint((a + b + c) * 65536.0);
  const uint64_t base = seed * 0x9E3779B97F4A7C15ull;
  uint64_t *key = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)m);
  uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)m);
  if (!key || !tmp) { free(key); free(tmp); return -2; }
  /* edges as canonical keys lo * n + hi; self-loops become the sentinel n*n (sorted last) */
  const uint64_t SENT = (uint64_t)n * (uint64_t)n;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < m; e++) {
    uint64_t u = 0, v = 0;
    for (int g = 0; g < G; g++) {
      const uint64_t r = splitmix64(base + (uint64_t)e * (uint64_t)G + (uint64_t)g);
      for (int t = 0; t < 4; t++) {
        const int l = 4 * g + t;
        if (l >= scale) break;
        const uint32_t x = (uint32_t)((r >> (16 * t)) & 0xffffu);
        const uint64_t right = (x >= TA && x < TB) || x >= TC;
        const uint64_t down = x >= TB;
        u |= down << l;
        v |= right << l;
      }
    }
    const uint64_t lo = u < v ? u : v, hi = u < v ? v : u;
    key[e] = u == v ? SENT : lo * (uint64_t)n + hi;
  }
  radix_sort(key, tmp, m, 2 * scale + 1);
  free(tmp);
  /* unique edges */
  int64_t ne = 0;
  for (int64_t e = 0; e < m; e++) {
    if (key[e] == SENT) break;
    if (ne == 0 || key[e] != key[ne - 1]) key[ne++] = key[e];
  }
  int32_t *lower = (int32_t *)calloc((size_t)n, sizeof(int32_t));   /* entries j < i in row i */
  int32_t *upper = (int32_t *)calloc((size_t)n, sizeof(int32_t));   /* entries j > i in row i */
  int32_t *rowptr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
  double *dinv = (double *)malloc(sizeof(double) * (size_t)n);
  const int64_t nnz = 2 * ne + n;
  int32_t *colidx = (int32_t *)malloc(sizeof(int32_t) * (size_t)nnz);
  float *val = (float *)malloc(sizeof(float) * (size_t)nnz);
  if (!lower || !upper || !rowptr || !dinv || !colidx || !val || nnz > 0x7fffffffLL) {
    free(key); free(lower); free(upper); free(rowptr); free(dinv); free(colidx); free(val);
    return nnz > 0x7fffffffLL ? -1 : -2;
  }
  for (int64_t e = 0; e < ne; e++) {
    upper[key[e] / (uint64_t)n]++;
    lower[key[e] % (uint64_t)n]++;
  }
  rowptr[0] = 0;
  for (int64_t i = 0; i < n; i++) {
    const int64_t deg = (int64_t)lower[i] + upper[i];
    rowptr[i + 1] = rowptr[i] + (int32_t)(deg + 1);
    dinv[i] = deg > 0 ? 1.0 / sqrt((double)deg) : 0.0;
  }
  /* fill in key order: row i receives its j < i entries (ascending lo), then the
     diagonal, then its j > i entries (ascending hi), so every row is sorted */
  int32_t *curl = lower, *curu = upper;   /* reused as cursors */
  for (int64_t i = 0; i < n; i++) {
    const int32_t nl = lower[i];
    colidx[rowptr[i] + nl] = (int32_t)i;
    val[rowptr[i] + nl] = 1.0f;
    curu[i] = rowptr[i] + nl + 1;
    curl[i] = rowptr[i];
  }
  for (int64_t e = 0; e < ne; e++) {
    const int64_t lo = (int64_t)(key[e] / (uint64_t)n), hi = (int64_t)(key[e] % (uint64_t)n);
    const float w = (float)(dinv[lo] * dinv[hi]);
    colidx[curl[hi]] = (int32_t)lo;
    val[curl[hi]++] = w;
    colidx[curu[lo]] = (int32_t)hi;
    val[curu[lo]++] = w;
  }
  free(key); free(lower); free(upper); free(dinv);
  *rowptr_out = rowptr;
  *colidx_out = colidx;
  *val_out = val;
  *nnz_out = nnz;
  return 0;
}

void ks_gen_free(void *p) { free(p); }
