/*
 * tiling_linalg.c -- TEST INFRASTRUCTURE (oracle).
 *
 * Caterva two-level tiling and the reference's six matmul algorithms,
 * restated on the CPU from:
 *   - tiling: src/iarray_zproxy.c:46-85 (chunk/block coordinates, padding
 *     trims), src/linalg/matmul/gemm.c:25-69 (block number -> start),
 *     src/linalg/iarray_gemm.c:392-396 (ext shapes);
 *   - _iarray_get_slice_buffer: src/iarray_container.c:628-736 (transposed
 *     views swap start/stop, fetch, then transpose);
 *   - gemm_blosc: src/linalg/matmul/gemm.c:130-249 (column-major walk over C
 *     chunks, A row-panel and repartitioned B column-panel caches, one
 *     full-K GEMM per C block with beta 0);
 *   - gemv_blosc: src/linalg/matmul/gemv.c:104-190;
 *   - iarray_opt_gemm: src/linalg/iarray_gemm.c:91-432 (per C block, K-chunk
 *     and K-block loops, zero-chunk/zero-block skips, edge trimming only in
 *     padded chunks, beta 1);
 *   - iarray_opt_gemv: src/linalg/iarray_gemv.c:88-351;
 *   - iarray_opt_gemm_a / _b: src/linalg/iarray_gemm_a.c:81-358 and
 *     src/linalg/iarray_gemm_b.c:81-358 (one operand fetched a chunk at a
 *     time, per C block a K-block loop, zero-chunk/zero-block skips, beta 1).
 * The reference computes each C block inside a Blosc2 prefilter and then
 * compresses it; here the block is written straight into a dense C (the
 * values C decodes to).  Per-block BLAS runs on a pthread pool, as Blosc2's
 * prefilter threads do, with a single-threaded BLAS (gemm.c:261-264).
 */
#include <dlfcn.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* ---- tiny parallel-for ----------------------------------------------------- */
typedef void (*task_fn)(void *ctx, int64_t i, uint8_t *scratch);
typedef struct {
  task_fn fn;
  void *ctx;
  int64_t n;
  atomic_llong next;
  size_t scratch_bytes;
} pool_job_t;

static void *pool_worker(void *arg) {
  pool_job_t *j = (pool_job_t *)arg;
  uint8_t *scratch = j->scratch_bytes ? (uint8_t *)malloc(j->scratch_bytes) : NULL;
  for (;;) {
    int64_t i = atomic_fetch_add(&j->next, 1);
    if (i >= j->n) break;
    j->fn(j->ctx, i, scratch);
  }
  free(scratch);
  return NULL;
}

static void parallel_for(int nthreads, int64_t n, size_t scratch_bytes, task_fn fn, void *ctx) {
  pool_job_t j;
  j.fn = fn;
  j.ctx = ctx;
  j.n = n;
  atomic_init(&j.next, 0);
  j.scratch_bytes = scratch_bytes;
  if (nthreads <= 1 || n <= 1) { pool_worker(&j); return; }
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  int started = 0;
  for (int t = 0; t < nthreads - 1; t++)
    if (pthread_create(&th[t], NULL, pool_worker, &j) == 0) started++;
  pool_worker(&j);
  for (int t = 0; t < started; t++) pthread_join(th[t], NULL);
}

/* ---- BLAS backend ------------------------------------------------------------ */
typedef void (*dgemm_fn)(int, int, int, int, int, int, double, const double *, int, const double *, int,
                         double, double *, int);
typedef void (*sgemm_fn)(int, int, int, int, int, int, float, const float *, int, const float *, int,
                         float, float *, int);
typedef void (*dgemv_fn)(int, int, int, int, double, const double *, int, const double *, int, double,
                         double *, int);
typedef void (*sgemv_fn)(int, int, int, int, float, const float *, int, const float *, int, float, float *,
                         int);
typedef void (*setthreads_fn)(int);
static struct {
  int which;
  dgemm_fn dgemm;
  sgemm_fn sgemm;
  dgemv_fn dgemv;
  sgemv_fn sgemv;
} g_blas = {0, NULL, NULL, NULL, NULL};

enum { ROWMAJOR = 101, NOTRANS = 111 };

int orc_set_blas(int which) {
  if (which == 1 && !g_blas.dgemm) {
    const char *names[] = {"libmkl_rt.so", "libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so", NULL};
    for (int i = 0; names[i] && !g_blas.dgemm; i++) {
      void *h = dlopen(names[i], RTLD_NOW | RTLD_GLOBAL);
      if (!h) continue;
      g_blas.dgemm = (dgemm_fn)dlsym(h, "cblas_dgemm");
      g_blas.sgemm = (sgemm_fn)dlsym(h, "cblas_sgemm");
      g_blas.dgemv = (dgemv_fn)dlsym(h, "cblas_dgemv");
      g_blas.sgemv = (sgemv_fn)dlsym(h, "cblas_sgemv");
      setthreads_fn st = (setthreads_fn)dlsym(h, "MKL_Set_Num_Threads");
      /* the reference runs MKL single-threaded inside its block threads */
      if (st) st(1);
      if (!g_blas.dgemm || !g_blas.sgemm || !g_blas.dgemv || !g_blas.sgemv) g_blas.dgemm = NULL;
    }
  }
  g_blas.which = (which == 1 && g_blas.dgemm) ? 1 : 0;
  return g_blas.which;
}

const char *orc_blas_name(void) { return g_blas.which == 1 ? "mkl" : "plain-c"; }

/* C(MxN, ldc) = A(MxK, lda) * B(KxN, ldb) + beta * C ; row-major */
static void gemm_any(int itemsize, int M, int N, int K, const void *A, int lda, const void *B, int ldb,
                     int beta, void *C, int ldc) {
  if (M <= 0 || N <= 0) return;
  if (g_blas.which == 1) {
    if (itemsize == 8)
      g_blas.dgemm(ROWMAJOR, NOTRANS, NOTRANS, M, N, K, 1.0, (const double *)A, lda, (const double *)B, ldb,
                   (double)beta, (double *)C, ldc);
    else
      g_blas.sgemm(ROWMAJOR, NOTRANS, NOTRANS, M, N, K, 1.0f, (const float *)A, lda, (const float *)B, ldb,
                   (float)beta, (float *)C, ldc);
    return;
  }
  if (itemsize == 8) {
    const double *a = (const double *)A, *b = (const double *)B;
    double *c = (double *)C;
    for (int i = 0; i < M; i++)
      for (int j = 0; j < N; j++) {
        double s = 0.0;
        for (int k = 0; k < K; k++) s += a[(int64_t)i * lda + k] * b[(int64_t)k * ldb + j];
        c[(int64_t)i * ldc + j] = beta ? c[(int64_t)i * ldc + j] + s : s;
      }
  } else {
    const float *a = (const float *)A, *b = (const float *)B;
    float *c = (float *)C;
    for (int i = 0; i < M; i++)
      for (int j = 0; j < N; j++) {
        float s = 0.0f;
        for (int k = 0; k < K; k++) s += a[(int64_t)i * lda + k] * b[(int64_t)k * ldb + j];
        c[(int64_t)i * ldc + j] = beta ? c[(int64_t)i * ldc + j] + s : s;
      }
  }
}

static void gemv_any(int itemsize, int M, int N, const void *A, int lda, const void *x, int beta, void *y) {
  if (M <= 0) return;
  if (g_blas.which == 1) {
    if (itemsize == 8)
      g_blas.dgemv(ROWMAJOR, NOTRANS, M, N, 1.0, (const double *)A, lda, (const double *)x, 1, (double)beta,
                   (double *)y, 1);
    else
      g_blas.sgemv(ROWMAJOR, NOTRANS, M, N, 1.0f, (const float *)A, lda, (const float *)x, 1, (float)beta,
                   (float *)y, 1);
    return;
  }
  gemm_any(itemsize, M, 1, N, A, lda, x, 1, beta, y, 1);
}

/* ---- tiling ---------------------------------------------------------------------- */
typedef struct {
  int itemsize;
  int64_t shape[2], chunk[2], block[2], extshape[2], extchunk[2];
  int64_t nchunks_ax[2], nblocks_ax[2];
  int64_t blocksize; /* bytes */
} tiling_t;

static int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

/* A 1-D array is treated as shape (1, n). */
static void make_tiling(int ndim, int itemsize, const int64_t *shape, const int64_t *chunk,
                        const int64_t *block, tiling_t *t) {
  t->itemsize = itemsize;
  for (int i = 0; i < 2; i++) {
    int s = ndim == 2 ? i : 0;
    t->shape[i] = (ndim == 1 && i == 0) ? 1 : shape[s];
    t->chunk[i] = (ndim == 1 && i == 0) ? 1 : chunk[s];
    t->block[i] = (ndim == 1 && i == 0) ? 1 : block[s];
    t->extshape[i] = ceil_div(t->shape[i], t->chunk[i]) * t->chunk[i];
    t->extchunk[i] = ceil_div(t->chunk[i], t->block[i]) * t->block[i];
    t->nchunks_ax[i] = t->extshape[i] / t->chunk[i];
    t->nblocks_ax[i] = t->extchunk[i] / t->block[i];
  }
  t->blocksize = t->block[0] * t->block[1] * itemsize;
}

static void array_tiling(const orc_array_t *a, tiling_t *t) {
  make_tiling(a->ndim, a->itemsize, a->shape, a->chunkshape, a->blockshape, t);
}

/* Copy block (nc, nb) of an underlying array into `dst` (row-major, ld = ldd
 * elements) for the part that overlaps region [r0,r1) x [c0,c1); dst's
 * origin is the region origin. */
typedef struct {
  const orc_array_t *a;
  tiling_t t;
  int64_t r0, r1, c0, c1;
  uint8_t *dst;
  int64_t ldd;
  int64_t *tasks; /* pairs (nchunk, nblock) */
  atomic_int err;
} slice_ctx_t;

static void slice_task(void *vctx, int64_t i, uint8_t *blk) {
  slice_ctx_t *s = (slice_ctx_t *)vctx;
  const tiling_t *t = &s->t;
  const int64_t nc = s->tasks[2 * i], nb = s->tasks[2 * i + 1];
  const int64_t ci0 = nc / t->nchunks_ax[1], ci1 = nc % t->nchunks_ax[1];
  const int64_t bi0 = nb / t->nblocks_ax[1], bi1 = nb % t->nblocks_ax[1];
  if (orc_decode_block(s->a->chunks[nc], s->a->csizes[nc], (int)nb, blk) < 0) {
    atomic_store(&s->err, 1);
    return;
  }
  const int is = t->itemsize;
  for (int64_t r = 0; r < t->block[0]; r++) {
    const int64_t in_chunk_r = bi0 * t->block[0] + r;
    const int64_t gr = ci0 * t->chunk[0] + in_chunk_r;
    if (in_chunk_r >= t->chunk[0] || gr >= t->shape[0] || gr < s->r0 || gr >= s->r1) continue;
    for (int64_t c = 0; c < t->block[1]; c++) {
      const int64_t in_chunk_c = bi1 * t->block[1] + c;
      const int64_t gc = ci1 * t->chunk[1] + in_chunk_c;
      if (in_chunk_c >= t->chunk[1] || gc >= t->shape[1] || gc < s->c0 || gc >= s->c1) continue;
      memcpy(s->dst + ((gr - s->r0) * s->ldd + (gc - s->c0)) * is, blk + (r * t->block[1] + c) * is, is);
    }
  }
}

/* region of the underlying (non-transposed) array into a row-major buffer */
static int fetch_region(const orc_array_t *a, int64_t r0, int64_t r1, int64_t c0, int64_t c1, void *dst,
                        int64_t ldd, int nthreads) {
  slice_ctx_t s;
  s.a = a;
  array_tiling(a, &s.t);
  s.r0 = r0; s.r1 = r1; s.c0 = c0; s.c1 = c1;
  s.dst = (uint8_t *)dst;
  s.ldd = ldd;
  atomic_init(&s.err, 0);
  const tiling_t *t = &s.t;
  if (r1 <= r0 || c1 <= c0) return 0;
  int64_t cr0 = r0 / t->chunk[0], cr1 = (r1 - 1) / t->chunk[0];
  int64_t cc0 = c0 / t->chunk[1], cc1 = (c1 - 1) / t->chunk[1];
  int64_t ntask = 0, cap = (cr1 - cr0 + 1) * (cc1 - cc0 + 1) * t->nblocks_ax[0] * t->nblocks_ax[1];
  s.tasks = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)cap);
  for (int64_t ci0 = cr0; ci0 <= cr1; ci0++)
    for (int64_t ci1 = cc0; ci1 <= cc1; ci1++)
      for (int64_t bi0 = 0; bi0 < t->nblocks_ax[0]; bi0++)
        for (int64_t bi1 = 0; bi1 < t->nblocks_ax[1]; bi1++) {
          const int64_t br0 = ci0 * t->chunk[0] + bi0 * t->block[0];
          const int64_t bc0 = ci1 * t->chunk[1] + bi1 * t->block[1];
          if (br0 >= r1 || br0 + t->block[0] <= r0 || bc0 >= c1 || bc0 + t->block[1] <= c0) continue;
          if (bi0 * t->block[0] >= t->chunk[0] || bi1 * t->block[1] >= t->chunk[1]) continue;
          s.tasks[2 * ntask] = ci0 * t->nchunks_ax[1] + ci1;
          s.tasks[2 * ntask + 1] = bi0 * t->nblocks_ax[1] + bi1;
          ntask++;
        }
  parallel_for(nthreads, ntask, (size_t)t->blocksize, slice_task, &s);
  free(s.tasks);
  return atomic_load(&s.err) ? -1 : 0;
}

int orc_array_to_dense(const orc_array_t *a, void *dst, int nthreads) {
  tiling_t t;
  array_tiling(a, &t);
  if (a->transposed) {
    int64_t start[2] = {0, 0}, stop[2] = {t.shape[1], t.shape[0]};
    return orc_get_slice(a, start, stop, stop, dst, nthreads);
  }
  return fetch_region(a, 0, t.shape[0], 0, t.shape[1], dst, t.shape[1], nthreads);
}

int orc_get_slice(const orc_array_t *a, const int64_t *start, const int64_t *stop, const int64_t *bufshape,
                  void *buf, int nthreads) {
  const int is = a->itemsize;
  if (a->ndim == 1) {
    memset(buf, 0, (size_t)(bufshape[0] * is));
    return fetch_region(a, 0, 1, start[0], stop[0], buf, bufshape[0], nthreads);
  }
  memset(buf, 0, (size_t)(bufshape[0] * bufshape[1] * is));
  if (!a->transposed) return fetch_region(a, start[0], stop[0], start[1], stop[1], buf, bufshape[1], nthreads);
  /* transposed view: fetch the swapped region of the underlying array into a
   * buffer of the swapped shape, then transpose (mkl_?omatcopy 'R','T' in
   * src/iarray_container.c:702-728). */
  const int64_t rows = bufshape[1], cols = bufshape[0];
  uint8_t *aux = (uint8_t *)calloc((size_t)(rows * cols), (size_t)is);
  if (!aux) return -1;
  int rc = fetch_region(a, start[1], stop[1], start[0], stop[0], aux, cols, nthreads);
  uint8_t *b = (uint8_t *)buf;
  for (int64_t r = 0; r < rows; r++)
    for (int64_t c = 0; c < cols; c++) memcpy(b + (c * rows + r) * is, aux + (r * cols + c) * is, is);
  free(aux);
  return rc;
}

/* logical shape of a (possibly transposed) 2-D array */
static void logical_shape(const orc_array_t *a, int64_t *s) {
  if (a->ndim == 1) { s[0] = a->shape[0]; s[1] = 1; return; }
  s[0] = a->transposed ? a->shape[1] : a->shape[0];
  s[1] = a->transposed ? a->shape[0] : a->shape[1];
}

/* ---- gemm_blosc (src/linalg/matmul/gemm.c:130-249) ------------------------------- */
typedef struct {
  int is;
  int64_t M, N, K;
  tiling_t ct;
  int64_t c_i0, c_i1; /* chunk index */
  const uint8_t *cache_a, *cache_b;
  uint8_t *c;
  int gemv;
} cblk_ctx_t;

static void gemm_block_task(void *vctx, int64_t nb, uint8_t *out) {
  cblk_ctx_t *x = (cblk_ctx_t *)vctx;
  const tiling_t *t = &x->ct;
  /* _gemm_prefilter_block_info: block number -> start inside the chunk */
  const int64_t b0 = nb / t->nblocks_ax[1], b1 = nb % t->nblocks_ax[1];
  const int64_t s0 = b0 * t->block[0], s1 = b1 * t->block[1];
  const int64_t m = t->block[0], n = t->block[1], k = x->K;
  if (x->gemv) {
    gemv_any(x->is, (int)m, (int)k, x->cache_a + s0 * k * x->is, (int)k, x->cache_b, 0, out);
  } else {
    gemm_any(x->is, (int)m, (int)n, (int)k, x->cache_a + s0 * k * x->is, (int)k,
             x->cache_b + s1 * k * x->is, (int)n, 0, out, (int)n);
  }
  for (int64_t r = 0; r < m; r++) {
    const int64_t rin = s0 + r, gr = x->c_i0 * t->chunk[0] + rin;
    if (rin >= t->chunk[0] || gr >= x->M) continue;
    for (int64_t c = 0; c < n; c++) {
      const int64_t cin = s1 + c, gc = x->c_i1 * t->chunk[1] + cin;
      if (cin >= t->chunk[1] || gc >= x->N) continue;
      memcpy(x->c + (gr * x->N + gc) * x->is, out + (r * n + c) * x->is, x->is);
    }
  }
}

int orc_gemm_blosc(const orc_array_t *a, const orc_array_t *b, const int64_t *cchunk, const int64_t *cblock,
                   void *c, int nthreads) {
  int64_t as[2], bs[2];
  logical_shape(a, as);
  logical_shape(b, bs);
  if (a->ndim != 2 || b->ndim != 2 || as[1] != bs[0] || a->itemsize != b->itemsize) return -1;
  const int is = a->itemsize;
  cblk_ctx_t x;
  memset(&x, 0, sizeof(x));
  x.is = is;
  x.M = as[0]; x.K = as[1]; x.N = bs[1];
  int64_t cshape[2] = {x.M, x.N};
  make_tiling(2, is, cshape, cchunk, cblock, &x.ct);
  const tiling_t *t = &x.ct;
  /* cache_b: K x extchunk1 fetched, then repartitioned into extchunk1/block1
   * slabs of K x block1 (_gemm_repart_caches, gemm.c:108-127) */
  uint8_t *aux_b = (uint8_t *)malloc((size_t)(x.K * t->extchunk[1] * is));
  uint8_t *cache_b = (uint8_t *)malloc((size_t)(x.K * t->extchunk[1] * is));
  uint8_t *cache_a = (uint8_t *)malloc((size_t)(t->extchunk[0] * x.K * is));
  int rc = 0;
  const int64_t nrows = t->nchunks_ax[0], nchunks = t->nchunks_ax[0] * t->nchunks_ax[1];
  int64_t chunk_col = -1;
  for (int64_t nc = 0; nc < nchunks && rc == 0; nc++) {
    const int64_t i0 = nc % nrows, i1 = nc / nrows; /* column-major walk */
    const int64_t e0 = i0 * t->chunk[0], e1 = i1 * t->chunk[1];
    const int64_t cs0 = e0 + t->chunk[0] <= x.M ? t->chunk[0] : x.M - e0;
    const int64_t cs1 = e1 + t->chunk[1] <= x.N ? t->chunk[1] : x.N - e1;
    if (chunk_col != i1) {
      chunk_col = i1;
      int64_t st[2] = {0, e1}, sp[2] = {x.K, e1 + cs1}, sh[2] = {x.K, t->extchunk[1]};
      rc = orc_get_slice(b, st, sp, sh, aux_b, nthreads);
      int64_t p = 0;
      for (int64_t slab = 0; slab < t->extchunk[1] / t->block[1]; slab++)
        for (int64_t r = 0; r < x.K; r++, p += t->block[1])
          memcpy(cache_b + p * is, aux_b + (r * t->extchunk[1] + slab * t->block[1]) * is,
                 (size_t)(t->block[1] * is));
    }
    int64_t st[2] = {e0, 0}, sp[2] = {e0 + cs0, x.K}, sh[2] = {t->extchunk[0], x.K};
    if (rc == 0) rc = orc_get_slice(a, st, sp, sh, cache_a, nthreads);
    x.c_i0 = i0; x.c_i1 = i1;
    x.cache_a = cache_a; x.cache_b = cache_b;
    x.c = (uint8_t *)c;
    if (rc == 0) parallel_for(nthreads, t->nblocks_ax[0] * t->nblocks_ax[1], (size_t)t->blocksize, gemm_block_task, &x);
  }
  free(aux_b); free(cache_b); free(cache_a);
  return rc;
}

/* ---- gemv_blosc (src/linalg/matmul/gemv.c:104-190) ------------------------------ */
int orc_gemv_blosc(const orc_array_t *a, const orc_array_t *b, const int64_t *cchunk, const int64_t *cblock,
                   void *c, int nthreads) {
  int64_t as[2];
  logical_shape(a, as);
  if (a->ndim != 2 || b->ndim != 1 || as[1] != b->shape[0] || a->itemsize != b->itemsize) return -1;
  const int is = a->itemsize;
  cblk_ctx_t x;
  memset(&x, 0, sizeof(x));
  x.is = is; x.M = as[0]; x.K = as[1]; x.N = 1; x.gemv = 1;
  int64_t cshape[1] = {x.M};
  /* C is 1-D: tile it as rows (chunk x 1) so blocks are (block x 1) */
  int64_t cs2[2] = {x.M, 1}, cc2[2] = {cchunk[0], 1}, cb2[2] = {cblock[0], 1};
  (void)cshape;
  make_tiling(2, is, cs2, cc2, cb2, &x.ct);
  const tiling_t *t = &x.ct;
  tiling_t bt;
  array_tiling(b, &bt);
  uint8_t *cache_b = (uint8_t *)calloc((size_t)bt.extshape[1], (size_t)is);
  uint8_t *cache_a = (uint8_t *)malloc((size_t)(t->extchunk[0] * x.K * is));
  int64_t st1[1] = {0}, sp1[1] = {x.K}, sh1[1] = {bt.extshape[1]};
  int rc = orc_get_slice(b, st1, sp1, sh1, cache_b, nthreads);
  for (int64_t nc = 0; nc < t->nchunks_ax[0] && rc == 0; nc++) {
    const int64_t e0 = nc * t->chunk[0];
    const int64_t cs0 = e0 + t->chunk[0] <= x.M ? t->chunk[0] : x.M - e0;
    int64_t st[2] = {e0, 0}, sp[2] = {e0 + cs0, x.K}, sh[2] = {t->extchunk[0], x.K};
    rc = orc_get_slice(a, st, sp, sh, cache_a, nthreads);
    x.c_i0 = nc; x.c_i1 = 0;
    x.cache_a = cache_a; x.cache_b = cache_b; x.c = (uint8_t *)c;
    if (rc == 0) parallel_for(nthreads, t->nblocks_ax[0], (size_t)(t->block[0] * is), gemm_block_task, &x);
  }
  free(cache_b); free(cache_a);
  return rc;
}

/* ---- iarray_opt_gemm (src/linalg/iarray_gemm.c:91-432) --------------------------- */
typedef struct {
  const orc_array_t *a, *b;
  tiling_t at, bt, ct;
  int64_t c_i0, c_i1;
  uint8_t *c;
  int64_t M, N;
  int gemv;
  atomic_int err;
} opt_ctx_t;

static int64_t trim(int64_t shape, int64_t block_start, int64_t blk) {
  /* returns the valid length, or -1 when the block lies beyond the shape */
  if (shape <= block_start) return -1;
  if (shape < block_start + blk) return shape - block_start;
  return blk;
}

static void opt_block_task(void *vctx, int64_t c_nblock, uint8_t *scratch) {
  opt_ctx_t *x = (opt_ctx_t *)vctx;
  const tiling_t *at = &x->at, *bt = &x->bt, *ct = &x->ct;
  const int is = at->itemsize;
  uint8_t *out = scratch;
  uint8_t *a_block = scratch + ct->blocksize;
  uint8_t *b_block = a_block + at->blocksize;
  memset(out, 0, (size_t)ct->blocksize);
  const int64_t K_chunks = at->nchunks_ax[1], N_chunks = bt->nchunks_ax[1];
  const int64_t K_blocks = at->nblocks_ax[1], N_blocks = x->gemv ? 1 : bt->nblocks_ax[1];
  const int64_t cb0 = c_nblock / N_blocks, cb1 = c_nblock % N_blocks;
  for (int64_t kc = 0; kc < K_chunks; kc++) {
    const int64_t a_nc = x->c_i0 * K_chunks + kc;
    const int64_t b_nc = x->gemv ? kc : kc * N_chunks + x->c_i1;
    const uint8_t *ach = x->a->chunks[a_nc], *bch = x->b->chunks[b_nc];
    const int32_t acs = x->a->csizes[a_nc], bcs = x->b->csizes[b_nc];
    if (orc_chunk_is_zeros(bch) || orc_chunk_is_zeros(ach)) continue;
    const int pad_m = at->shape[0] != at->extshape[0] && x->c_i0 == at->nchunks_ax[0] - 1;
    const int pad_k = at->shape[1] != at->extshape[1] && kc == K_chunks - 1;
    const int pad_n = !x->gemv && bt->shape[1] != bt->extshape[1] && x->c_i1 == N_chunks - 1;
    for (int64_t kb = 0; kb < K_blocks; kb++) {
      const int64_t a_nb = cb0 * K_blocks + kb;
      const int64_t b_nb = x->gemv ? kb : kb * N_blocks + cb1;
      if (orc_block_is_zeros(ach, acs, (int)a_nb) || orc_block_is_zeros(bch, bcs, (int)b_nb)) continue;
      int64_t M = at->block[0], K = at->block[1], N = x->gemv ? 1 : bt->block[1];
      const int64_t lda = K, ldb = N, ldc = N;
      if (x->gemv) {
        /* opt_gemv trims every block against the shape (iarray_gemv.c:166-190) */
        M = trim(at->shape[0], at->chunk[0] * x->c_i0 + at->block[0] * cb0, at->block[0]);
        K = trim(at->shape[1], at->chunk[1] * kc + at->block[1] * kb, at->block[1]);
        if (M < 0 || K < 0) continue;
      } else {
        if (pad_m) { M = trim(at->shape[0], x->c_i0 * at->chunk[0] + cb0 * at->block[0], at->block[0]); if (M < 0) continue; }
        if (pad_k) { K = trim(at->shape[1], kc * at->chunk[1] + kb * at->block[1], at->block[1]); if (K < 0) continue; }
        if (pad_n) { N = trim(bt->shape[1], x->c_i1 * bt->chunk[1] + cb1 * bt->block[1], bt->block[1]); if (N < 0) continue; }
      }
      const int64_t a_items = at->block[0] * at->block[1], b_items = bt->block[0] * bt->block[1];
      if (orc_getitem(ach, acs, (int)(a_nb * a_items), (int)a_items, a_block) < 0 ||
          orc_getitem(bch, bcs, (int)(b_nb * b_items), (int)b_items, b_block) < 0) {
        atomic_store(&x->err, 1);
        return;
      }
      if (x->gemv) gemv_any(is, (int)M, (int)K, a_block, (int)lda, b_block, 1, out);
      else gemm_any(is, (int)M, (int)N, (int)K, a_block, (int)lda, b_block, (int)ldb, 1, out, (int)ldc);
    }
  }
  /* write the C block into dense C */
  const int64_t s0 = cb0 * ct->block[0], s1 = cb1 * ct->block[1];
  for (int64_t r = 0; r < ct->block[0]; r++) {
    const int64_t rin = s0 + r, gr = x->c_i0 * ct->chunk[0] + rin;
    if (rin >= ct->chunk[0] || gr >= x->M) continue;
    for (int64_t cc = 0; cc < ct->block[1]; cc++) {
      const int64_t cin = s1 + cc, gc = x->c_i1 * ct->chunk[1] + cin;
      if (cin >= ct->chunk[1] || gc >= x->N) continue;
      memcpy(x->c + (gr * x->N + gc) * is, out + (r * ct->block[1] + cc) * is, is);
    }
  }
}

static int opt_common(const orc_array_t *a, const orc_array_t *b, const int64_t *cchunk, const int64_t *cblock,
                      void *c, int nthreads, int gemv) {
  if (a->ndim != 2 || b->ndim != (gemv ? 1 : 2) || a->transposed || b->transposed) return -1;
  if (a->itemsize != b->itemsize || a->shape[1] != b->shape[0]) return -1;
  /* partition constraints (iarray_gemm.c:327-364, iarray_gemv.c:270-286) */
  if (a->chunkshape[0] != cchunk[0] || a->chunkshape[1] != b->chunkshape[0]) return -2;
  if (!gemv && b->chunkshape[1] != cchunk[1]) return -2;
  if (a->blockshape[0] != cblock[0] || a->blockshape[1] != b->blockshape[0]) return -3;
  if (!gemv && b->blockshape[1] != cblock[1]) return -3;
  opt_ctx_t x;
  memset(&x, 0, sizeof(x));
  x.a = a; x.b = b; x.gemv = gemv;
  array_tiling(a, &x.at);
  array_tiling(b, &x.bt);
  if (gemv) {
    /* view the 1-D b as (K x 1) with chunks (c x 1) and blocks (blk x 1) */
    int64_t s[2] = {b->shape[0], 1}, ch[2] = {b->chunkshape[0], 1}, bl[2] = {b->blockshape[0], 1};
    make_tiling(2, b->itemsize, s, ch, bl, &x.bt);
  }
  x.M = a->shape[0];
  x.N = gemv ? 1 : b->shape[1];
  int64_t cs[2] = {x.M, x.N}, cc[2] = {cchunk[0], gemv ? 1 : cchunk[1]}, cb[2] = {cblock[0], gemv ? 1 : cblock[1]};
  make_tiling(2, a->itemsize, cs, cc, cb, &x.ct);
  x.c = (uint8_t *)c;
  atomic_init(&x.err, 0);
  const size_t scratch = (size_t)(x.ct.blocksize + x.at.blocksize + x.bt.blocksize);
  for (int64_t nc = 0; nc < x.ct.nchunks_ax[0] * x.ct.nchunks_ax[1]; nc++) {
    x.c_i0 = nc / x.ct.nchunks_ax[1];
    x.c_i1 = nc % x.ct.nchunks_ax[1];
    parallel_for(nthreads, x.ct.nblocks_ax[0] * x.ct.nblocks_ax[1], scratch, opt_block_task, &x);
  }
  return atomic_load(&x.err) ? -4 : 0;
}

int orc_opt_gemm(const orc_array_t *a, const orc_array_t *b, const int64_t *cchunk, const int64_t *cblock,
                 void *c, int nthreads) {
  return opt_common(a, b, cchunk, cblock, c, nthreads, 0);
}

int orc_opt_gemv(const orc_array_t *a, const orc_array_t *b, const int64_t *cchunk, const int64_t *cblock,
                 void *c, int nthreads) {
  return opt_common(a, b, cchunk, cblock, c, nthreads, 1);
}

/* ---- iarray_opt_gemm_a / _b (src/linalg/iarray_gemm_a.c:81-358, iarray_gemm_b.c:81-358) ----
 * gemm_a: A has one chunk column; C chunk i is A's chunk row i, and block j of
 * it is B's chunk j (C blocks are A chunk rows x B chunk columns).  Each C
 * block starts at zero and adds, for every block k along K, A block k of A's
 * chunk i times B block k of B's chunk j (cblas_?gemm, beta 1), skipping
 * zero blocks, and stays zero when B's chunk is a zero chunk
 * (iarray_gemm_a.c:100-175).  gemm_b is the mirror: B has one chunk row, C
 * chunk j is B's chunk column j and block i of it is A's chunk i; the chunk
 * tested for zeros is A's (iarray_gemm_b.c:100-175). */
typedef struct {
  const orc_array_t *a, *b;
  tiling_t at, bt;
  int64_t c_chunk;
  int variant_b;
  uint8_t *c;
  int64_t M, N;
  atomic_int err;
} ab_ctx_t;

static void ab_block_task(void *vctx, int64_t nb, uint8_t *scratch) {
  ab_ctx_t *x = (ab_ctx_t *)vctx;
  const tiling_t *at = &x->at, *bt = &x->bt;
  const int is = at->itemsize;
  const int64_t a_nc = x->variant_b ? nb : x->c_chunk, b_nc = x->variant_b ? x->c_chunk : nb;
  const int64_t Mb = at->block[0], Nb = bt->block[1], Kb = at->block[1];
  uint8_t *out = scratch;
  uint8_t *a_block = out + Mb * Nb * is;
  uint8_t *b_block = a_block + at->blocksize;
  memset(out, 0, (size_t)(Mb * Nb * is));
  const uint8_t *ach = x->a->chunks[a_nc], *bch = x->b->chunks[b_nc];
  const int32_t acs = x->a->csizes[a_nc], bcs = x->b->csizes[b_nc];
  if (!orc_chunk_is_zeros(x->variant_b ? ach : bch)) {
    const int64_t nk = x->variant_b ? at->extchunk[1] / at->block[1] : bt->extchunk[0] / bt->block[0];
    for (int64_t k = 0; k < nk; k++) {
      if (orc_block_is_zeros(ach, acs, (int)k) || orc_block_is_zeros(bch, bcs, (int)k)) continue;
      const int64_t a_items = at->block[0] * at->block[1], b_items = bt->block[0] * bt->block[1];
      if (orc_getitem(ach, acs, (int)(k * a_items), (int)a_items, a_block) < 0 ||
          orc_getitem(bch, bcs, (int)(k * b_items), (int)b_items, b_block) < 0) {
        atomic_store(&x->err, 1);
        return;
      }
      gemm_any(is, (int)Mb, (int)Nb, (int)Kb, a_block, (int)Kb, b_block, (int)Nb, 1, out, (int)Nb);
    }
  }
  for (int64_t r = 0; r < Mb; r++) {
    const int64_t gr = a_nc * at->chunk[0] + r;
    if (gr >= x->M) break;
    for (int64_t cc = 0; cc < Nb; cc++) {
      const int64_t gc = b_nc * bt->chunk[1] + cc;
      if (gc >= x->N) break;
      memcpy(x->c + (gr * x->N + gc) * is, out + (r * Nb + cc) * is, is);
    }
  }
}

static int opt_ab(const orc_array_t *a, const orc_array_t *b, const int64_t *cchunk, const int64_t *cblock, void *c,
                  int nthreads, int variant_b) {
  if (a->ndim != 2 || b->ndim != 2 || a->transposed || b->transposed) return -1;
  if (a->itemsize != b->itemsize || a->shape[1] != b->shape[0]) return -1;
  /* partition constraints (iarray_gemm_a.c:213-262, iarray_gemm_b.c:213-263) */
  if (a->chunkshape[1] < a->shape[1] || b->chunkshape[0] < b->shape[0] || a->chunkshape[1] != b->chunkshape[0])
    return -2;
  if (!variant_b && (a->chunkshape[0] != cchunk[0] || b->shape[1] > cchunk[1])) return -2;
  if (variant_b && (a->shape[0] > cchunk[0] || b->chunkshape[1] != cchunk[1])) return -2;
  if (a->chunkshape[0] != cblock[0] || b->chunkshape[1] != cblock[1] || a->chunkshape[0] != a->blockshape[0] ||
      b->chunkshape[1] != b->blockshape[1] || a->blockshape[1] != b->blockshape[0])
    return -3;
  ab_ctx_t x;
  memset(&x, 0, sizeof(x));
  x.a = a;
  x.b = b;
  x.variant_b = variant_b;
  array_tiling(a, &x.at);
  array_tiling(b, &x.bt);
  x.M = a->shape[0];
  x.N = b->shape[1];
  x.c = (uint8_t *)c;
  atomic_init(&x.err, 0);
  const size_t scratch = (size_t)(x.at.block[0] * x.bt.block[1] * x.at.itemsize + x.at.blocksize + x.bt.blocksize);
  const int64_t n_c = variant_b ? x.bt.nchunks_ax[1] : x.at.nchunks_ax[0];
  const int64_t n_blocks = variant_b ? x.at.nchunks_ax[0] : x.bt.nchunks_ax[1];
  for (int64_t ci = 0; ci < n_c; ci++) {
    x.c_chunk = ci;
    parallel_for(nthreads, n_blocks, scratch, ab_block_task, &x);
  }
  return atomic_load(&x.err) ? -4 : 0;
}

int orc_opt_gemm_a(const orc_array_t *a, const orc_array_t *b, const int64_t *cchunk, const int64_t *cblock,
                   void *c, int nthreads) {
  return opt_ab(a, b, cchunk, cblock, c, nthreads, 0);
}

int orc_opt_gemm_b(const orc_array_t *a, const orc_array_t *b, const int64_t *cchunk, const int64_t *cblock,
                   void *c, int nthreads) {
  return opt_ab(a, b, cchunk, cblock, c, nthreads, 1);
}

/* ---- C write stage for the CPU baseline ------------------------------------------------ */
typedef int (*lz4_comp_fn)(const char *, char *, int, int, int);
static lz4_comp_fn g_lz4_comp = NULL;

typedef struct {
  const uint8_t *c;
  tiling_t t;
  int filter, accel;
  atomic_llong total;
} enc_ctx_t;

static void enc_task(void *vctx, int64_t i, uint8_t *scratch) {
  enc_ctx_t *x = (enc_ctx_t *)vctx;
  const tiling_t *t = &x->t;
  const int is = t->itemsize;
  const int64_t nblk = t->nblocks_ax[0] * t->nblocks_ax[1];
  const int64_t nc = i / nblk, nb = i % nblk;
  const int64_t ci0 = nc / t->nchunks_ax[1], ci1 = nc % t->nchunks_ax[1];
  const int64_t bi0 = nb / t->nblocks_ax[1], bi1 = nb % t->nblocks_ax[1];
  uint8_t *blk = scratch, *sh = scratch + t->blocksize, *out = scratch + 2 * t->blocksize;
  memset(blk, 0, (size_t)t->blocksize);
  for (int64_t r = 0; r < t->block[0]; r++) {
    const int64_t rin = bi0 * t->block[0] + r, gr = ci0 * t->chunk[0] + rin;
    if (rin >= t->chunk[0] || gr >= t->shape[0]) continue;
    for (int64_t cc = 0; cc < t->block[1]; cc++) {
      const int64_t cin = bi1 * t->block[1] + cc, gc = ci1 * t->chunk[1] + cin;
      if (cin >= t->chunk[1] || gc >= t->shape[1]) continue;
      memcpy(blk + (r * t->block[1] + cc) * is, x->c + (gr * t->shape[1] + gc) * is, is);
    }
  }
  int nstreams = 1;
  if (x->filter == 1) { orc_shuffle(is, (int32_t)t->blocksize, blk, sh); nstreams = is; }
  else if (x->filter == 2) orc_bitshuffle(is, (int32_t)t->blocksize, blk, sh, 5);
  else memcpy(sh, blk, (size_t)t->blocksize);
  const int neblock = (int)(t->blocksize / nstreams);
  int64_t tot = 0;
  for (int s = 0; s < nstreams; s++) {
    int r = g_lz4_comp((const char *)sh + (int64_t)s * neblock, (char *)out, neblock, neblock, x->accel);
    tot += 4 + ((r > 0 && r < neblock) ? r : neblock);
  }
  atomic_fetch_add(&x->total, tot);
}

int64_t orc_encode_dense(const void *c, int ndim, const int64_t *shape, int itemsize, const int64_t *chunkshape,
                         const int64_t *blockshape, int clevel, int filter, int nthreads) {
  if (!g_lz4_comp) {
    const char *names[] = {"liblz4.so.1", "/opt/conda/lib/liblz4.so.1", "/usr/lib/x86_64-linux-gnu/liblz4.so.1", NULL};
    for (int i = 0; names[i] && !g_lz4_comp; i++) {
      void *h = dlopen(names[i], RTLD_NOW | RTLD_LOCAL);
      if (h) g_lz4_comp = (lz4_comp_fn)dlsym(h, "LZ4_compress_fast");
    }
    if (!g_lz4_comp) return -1;
  }
  enc_ctx_t x;
  x.c = (const uint8_t *)c;
  make_tiling(ndim, itemsize, shape, chunkshape, blockshape, &x.t);
  x.filter = filter;
  x.accel = clevel >= 9 ? 1 : 10 - clevel; /* faster settings at lower clevel */
  atomic_init(&x.total, 0);
  const int64_t n = x.t.nchunks_ax[0] * x.t.nchunks_ax[1] * x.t.nblocks_ax[0] * x.t.nblocks_ax[1];
  parallel_for(nthreads, n, (size_t)(3 * x.t.blocksize + 64), enc_task, &x);
  return atomic_load(&x.total);
}

/* ---- compensated reference sums for error metrics ------------------------------------ */
void orc_dd_gemm_sample(const double *a, const double *b, int64_t M, int64_t N, int64_t K, const int64_t *rows,
                        const int64_t *cols, int64_t nsample, double *out) {
  (void)M;
  for (int64_t s = 0; s < nsample; s++) {
    const int64_t i = rows[s], j = cols[s];
    double hi = 0.0, lo = 0.0; /* double-double accumulator */
    for (int64_t k = 0; k < K; k++) {
      const double x = a[i * K + k], y = b[k * N + j];
      const double p = x * y;
      const double pe = fma(x, y, -p); /* exact product error */
      const double sh = hi + p;
      const double bb = sh - hi;
      const double se = (hi - (sh - bb)) + (p - bb); /* TwoSum */
      hi = sh;
      lo += se + pe;
    }
    out[s] = hi + lo;
  }
}
