/*
 * iarray_batch.c -- iarray_linalg_matmul_batch: many independent products in
 * one device pass, replayed from a HIP graph.
 *
 * The reference runs such a workload as one iarray_linalg_matmul call per
 * product (src/linalg/matmul.c:19-68), each re-decoding its A and B panels.
 * Here a batch plan is built once per batch signature (operand containers,
 * their device generations and the output storages):
 *   - every operand block of every product goes into ONE decode job table
 *     (LZ4 streams + un-filter jobs) writing into ONE dense buffer that holds
 *     all A images followed by all B images;
 *   - every 128x128 output tile of every product goes into ONE tile list,
 *     sorted by decreasing K (longest first), consumed by the persistent
 *     work-queue GEMM;
 *   - every C block goes into ONE pack table;
 * and the launch sequence (decode, queue reset, GEMM, pack) is captured into
 * a hipGraph.  A call with the same signature replays the graph, copies the
 * packed C blocks back and LZ4-encodes all C chunks on the host threads.
 * Plain precision only; the mixed forms run product by product.
 */
#include <stdlib.h>
#include <string.h>

#include "iarray_private.h"

typedef struct {
  int n;
  iarray_container_t **a, **b;
  uint64_t *gen;
  int64_t *store; /* chunk/block shapes of each output, 4 per product */
  int dt;
  void *graph;
  /* device buffers owned by the plan (their addresses are baked into the graph) */
  void *d_dense, *d_dense_c, *d_scratch, *d_packed;
  iamx_lz4_job_t *d_lz4;
  iamx_block_job_t *d_blocks;
  iamx_batch_prod_t *d_prods;
  int32_t *d_tiles, *d_pblocks;
  iamx_batch_pack_t *d_packs;
  int64_t packed_bytes;
  int64_t *out_off; /* per product: byte offset of its packed blocks */
  /* the captured launches' arguments, for an eager replay */
  int n_lz4, n_blk, ntiles, npb;
  int64_t b_base; /* byte offset of the B images in d_dense */
} batch_plan_t;

static batch_plan_t g_plan;

static void free_batch_plan(batch_plan_t *p) {
  iamx_graph_destroy(p->graph);
  void *bufs[] = {p->d_dense, p->d_dense_c, p->d_scratch, p->d_packed, p->d_lz4,
                  p->d_blocks, p->d_prods, p->d_tiles, p->d_pblocks, p->d_packs};
  for (size_t i = 0; i < sizeof(bufs) / sizeof(bufs[0]); i++) iamx_free(bufs[i]);
  free(p->a);
  free(p->b);
  free(p->gen);
  free(p->store);
  free(p->out_off);
  memset(p, 0, sizeof(*p));
}

static int same_signature(const batch_plan_t *p, int n, iarray_container_t **a, iarray_container_t **b,
                          const iarray_storage_t *st, int dt) {
  if (!p->graph || p->n != n || p->dt != dt) return 0;
  for (int i = 0; i < n; i++) {
    if (p->a[i] != a[i] || p->b[i] != b[i]) return 0;
    if (!a[i]->device_valid || !b[i]->device_valid) return 0;
    if (p->gen[2 * i] != a[i]->generation || p->gen[2 * i + 1] != b[i]->generation) return 0;
    const int64_t *s = p->store + 4 * i;
    if (s[0] != st[i].chunkshape[0] || s[1] != st[i].chunkshape[1] || s[2] != st[i].blockshape[0] ||
        s[3] != st[i].blockshape[1])
      return 0;
  }
  return 1;
}

static int64_t up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

typedef struct {
  int32_t p, tm, tn, k;
} tile_t;

static int by_k_desc(const void *x, const void *y) {
  const tile_t *a = (const tile_t *)x, *b = (const tile_t *)y;
  if (a->k != b->k) return a->k < b->k ? 1 : -1;
  if (a->p != b->p) return a->p - b->p;
  if (a->tm != b->tm) return a->tm - b->tm;
  return a->tn - b->tn;
}

/* Append operand o's decode jobs, rebased to `dense_off` elements of the
 * dense buffer and `scr_off` bytes of the scratch. */
static void append_jobs(const iarray_container_t *o, int64_t dense_off, int64_t scr_off, iamx_lz4_job_t *lz4,
                        int64_t *nl, iamx_block_job_t *blk, int64_t *nb) {
  const iax_plan_t *pl = &o->plan;
  for (int i = 0; i < pl->n_lz4; i++) {
    lz4[*nl] = pl->h_lz4[i];
    lz4[*nl].dst_off += scr_off;
    (*nl)++;
  }
  for (int i = 0; i < pl->n_blocks; i++) {
    iamx_block_job_t J = pl->h_blocks[i];
    for (int s = 0; s < J.nstreams; s++)
      if (J.kind[s] == IAMX_SRC_SCRATCH) J.src[s] += (uint64_t)scr_off;
    J.dst_off += dense_off;
    blk[(*nb)++] = J;
  }
}

static ina_rc_t upload(void **d, const void *h, size_t bytes) {
  *d = iamx_malloc(bytes ? bytes : 16);
  if (!*d) return iax_dev_error(IAMX_ENOMEM, "batch plan");
  int e = iamx_h2d(*d, h, bytes);
  return e ? iax_dev_error(e, "batch plan upload") : INA_SUCCESS;
}

static ina_rc_t build_plan(batch_plan_t *p, int n, iarray_container_t **a, iarray_container_t **b,
                           iarray_container_t **c, const iarray_storage_t *st, int dt) {
  const int is = dt == 0 ? 8 : 4;
  p->n = n;
  p->dt = dt;
  p->a = (iarray_container_t **)malloc(sizeof(void *) * (size_t)n);
  p->b = (iarray_container_t **)malloc(sizeof(void *) * (size_t)n);
  p->gen = (uint64_t *)malloc(sizeof(uint64_t) * 2 * (size_t)n);
  p->store = (int64_t *)malloc(sizeof(int64_t) * 4 * (size_t)n);
  p->out_off = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
  iamx_batch_prod_t *prods = (iamx_batch_prod_t *)calloc((size_t)n, sizeof(iamx_batch_prod_t));
  iamx_batch_pack_t *packs = (iamx_batch_pack_t *)calloc((size_t)n, sizeof(iamx_batch_pack_t));
  if (!p->a || !p->b || !p->gen || !p->store || !p->out_off || !prods || !packs) {
    free(prods);
    free(packs);
    return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  }
  /* sizes and offsets: A images first, then B images, in one dense buffer */
  int64_t na = 0, nbimg = 0, nc = 0, scr = 0, nlz = 0, nblk = 0, ntiles = 0, npb = 0, packed = 0;
  for (int i = 0; i < n; i++) {
    const int64_t M = a[i]->shape[0], K = a[i]->shape[1], N = b[i]->shape[1];
    prods[i] = (iamx_batch_prod_t){M, N, K, na, nbimg, nc};
    na += up(M * K, 64);
    nbimg += up(K * N, 64);
    nc += up(M * N, 64);
    scr += a[i]->plan.scratch_bytes + b[i]->plan.scratch_bytes;
    nlz += a[i]->plan.n_lz4 + b[i]->plan.n_lz4;
    nblk += a[i]->plan.n_blocks + b[i]->plan.n_blocks;
    ntiles += ((M + 127) / 128) * ((N + 127) / 128);
    const iarray_container_t *cc = c[i];
    iamx_tiling_t *t = &packs[i].tiling;
    for (int d = 0; d < 2; d++) {
      t->shape[d] = cc->shape[d];
      t->chunk[d] = cc->chunk[d];
      t->block[d] = cc->block[d];
      t->nchunks[d] = cc->nchunks_ax[d];
      t->nblocks[d] = cc->nblocks_ax[d];
    }
    t->typesize = cc->itemsize;
    t->filter = cc->cfg.compression_level > 0 ? iax_filter_code(&cc->cfg) : IAMX_FILTER_NONE;
    packs[i].c_off_bytes = prods[i].c_off * is;
    packs[i].out_off = packed;
    p->out_off[i] = packed;
    packed += cc->nchunks * cc->chunk_nbytes;
    npb += cc->nchunks * cc->nblocks_ax[0] * cc->nblocks_ax[1];
    p->a[i] = a[i];
    p->b[i] = b[i];
    p->gen[2 * i] = a[i]->generation;
    p->gen[2 * i + 1] = b[i]->generation;
    p->store[4 * i] = st[i].chunkshape[0];
    p->store[4 * i + 1] = st[i].chunkshape[1];
    p->store[4 * i + 2] = st[i].blockshape[0];
    p->store[4 * i + 3] = st[i].blockshape[1];
  }
  if (ntiles > INT32_MAX || npb > INT32_MAX || nlz > INT32_MAX || nblk > INT32_MAX) {
    free(prods);
    free(packs);
    return iax_set_error(INA_ERROR(INA_ERR_INVALID_ARGUMENT), "batch too large for one pass");
  }
  ina_rc_t rc = INA_SUCCESS;
  iamx_lz4_job_t *lz4 = (iamx_lz4_job_t *)malloc(sizeof(iamx_lz4_job_t) * (size_t)(nlz + 1));
  iamx_block_job_t *blk = (iamx_block_job_t *)malloc(sizeof(iamx_block_job_t) * (size_t)(nblk + 1));
  tile_t *tl = (tile_t *)malloc(sizeof(tile_t) * (size_t)(ntiles + 1));
  int32_t *tiles = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)(ntiles + 1));
  int32_t *pblk = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)(npb + 1));
  if (!lz4 || !blk || !tl || !tiles || !pblk) rc = INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  int64_t il = 0, ib = 0, it = 0, ip = 0, so = 0;
  for (int i = 0; i < n && INA_SUCCEED(rc); i++) {
    append_jobs(a[i], prods[i].a_off, so, lz4, &il, blk, &ib);
    so += a[i]->plan.scratch_bytes;
    append_jobs(b[i], na + prods[i].b_off, so, lz4, &il, blk, &ib);
    so += b[i]->plan.scratch_bytes;
    const int64_t M = prods[i].M, N = prods[i].N, K = prods[i].K;
    for (int32_t tm = 0; tm < (M + 127) / 128; tm++)
      for (int32_t tn = 0; tn < (N + 127) / 128; tn++) tl[it++] = (tile_t){i, tm, tn, (int32_t)K};
    const int64_t nblocks_c = c[i]->nchunks * c[i]->nblocks_ax[0] * c[i]->nblocks_ax[1];
    for (int64_t k = 0; k < nblocks_c; k++) {
      pblk[2 * ip] = i;
      pblk[2 * ip + 1] = (int32_t)k;
      ip++;
    }
  }
  if (INA_SUCCEED(rc)) {
    qsort(tl, (size_t)ntiles, sizeof(tile_t), by_k_desc);
    for (int64_t k = 0; k < ntiles; k++) {
      tiles[3 * k] = tl[k].p;
      tiles[3 * k + 1] = tl[k].tm;
      tiles[3 * k + 2] = tl[k].tn;
    }
  }
  iax_sort_lz4_jobs(lz4, il);
  if (INA_SUCCEED(rc)) rc = upload((void **)&p->d_lz4, lz4, sizeof(iamx_lz4_job_t) * (size_t)il);
  if (INA_SUCCEED(rc)) rc = upload((void **)&p->d_blocks, blk, sizeof(iamx_block_job_t) * (size_t)ib);
  if (INA_SUCCEED(rc)) rc = upload((void **)&p->d_prods, prods, sizeof(iamx_batch_prod_t) * (size_t)n);
  if (INA_SUCCEED(rc)) rc = upload((void **)&p->d_tiles, tiles, sizeof(int32_t) * 3 * (size_t)ntiles);
  if (INA_SUCCEED(rc)) rc = upload((void **)&p->d_pblocks, pblk, sizeof(int32_t) * 2 * (size_t)npb);
  if (INA_SUCCEED(rc)) rc = upload((void **)&p->d_packs, packs, sizeof(iamx_batch_pack_t) * (size_t)n);
  if (INA_SUCCEED(rc)) {
    p->d_dense = iamx_malloc((size_t)((na + nbimg) * is) + 256);
    p->d_dense_c = iamx_malloc((size_t)(nc * is) + 256);
    p->d_scratch = iamx_malloc((size_t)scr + 256);
    p->d_packed = iamx_malloc((size_t)packed + 256);
    p->packed_bytes = packed;
    if (!p->d_dense || !p->d_dense_c || !p->d_scratch || !p->d_packed)
      rc = iax_dev_error(IAMX_ENOMEM, "batch buffers");
  }
  if (INA_SUCCEED(rc)) {
    uint8_t *dense_b = (uint8_t *)p->d_dense + na * is;
    p->n_lz4 = (int)il;
    p->n_blk = (int)ib;
    p->ntiles = (int)ntiles;
    p->npb = (int)npb;
    p->b_base = na * is;
    int e = iamx_capture_begin();
    if (!e) e = iamx_decode(p->d_lz4, (int)il, p->d_blocks, (int)ib, p->d_scratch, p->d_dense);
    if (!e) e = iamx_gemm_batch(dt, p->d_prods, p->d_tiles, (int)ntiles, p->d_dense, dense_b, p->d_dense_c);
    if (!e) e = iamx_pack_batch(p->d_packs, p->d_pblocks, (int)npb, p->d_dense_c, p->d_packed);
    const int e2 = iamx_capture_end(&p->graph);
    if (e || e2) rc = iax_dev_error(e ? e : e2, "batch graph capture");
  }
  free(lz4);
  free(blk);
  free(tl);
  free(tiles);
  free(pblk);
  free(prods);
  free(packs);
  return rc;
}

typedef struct {
  iarray_container_t **c;
  const int64_t *first; /* prefix sums of chunk counts */
  int n;
  uint8_t *packed;
  const int64_t *out_off;
  int err;
} wctx_t;

static void write_task(void *vctx, int64_t g, int worker) {
  (void)worker;
  wctx_t *x = (wctx_t *)vctx;
  int lo = 0, hi = x->n - 1; /* product holding global chunk g */
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (x->first[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  iarray_container_t *c = x->c[lo];
  const int64_t nc = g - x->first[lo];
  const int filtered = c->cfg.compression_level > 0;
  free(c->chunks[nc]);
  c->chunks[nc] = NULL;
  if (iax_encode_chunk(x->packed + x->out_off[lo] + nc * c->chunk_nbytes, (int32_t)c->chunk_nbytes, c->blocksize,
                       c->itemsize, &c->cfg, filtered, &c->chunks[nc], &c->csizes[nc]) != 0)
    x->err = 1;
}

INA_API(ina_rc_t) iarray_linalg_matmul_batch(iarray_context_t *ctx, int n, iarray_container_t **a,
                                             iarray_container_t **b, iarray_storage_t *storage,
                                             const iarray_matmul_params_t *params, iarray_container_t **c) {
  if (!ctx || !a || !b || !storage || !c || n < 0) return INA_ERROR(INA_ERR_NULL_POINTER);
  if (n == 0) return INA_SUCCESS;
  const int plain = !params || params->precision == IARRAY_MATMUL_PLAIN;
  int dt = -1, eligible = plain;
  for (int i = 0; i < n; i++) {
    if (!a[i] || !b[i]) return INA_ERROR(INA_ERR_NULL_POINTER);
    const int d = a[i]->dtshape.dtype == IARRAY_DATA_TYPE_DOUBLE ? 0 : a[i]->dtshape.dtype == IARRAY_DATA_TYPE_FLOAT ? 1 : -1;
    if (i == 0) dt = d;
    if (d != dt || a[i]->viewed || b[i]->viewed || b[i]->dtshape.ndim != 2) eligible = 0;
  }
  if (!eligible) {
    /* views, GEMVs, mixed dtypes or mixed precision: product by product */
    for (int i = 0; i < n; i++) {
      ina_rc_t rc = iarray_linalg_matmul_params(ctx, a[i], b[i], &storage[i], params, &c[i]);
      if (INA_FAILED(rc)) return rc;
    }
    return INA_SUCCESS;
  }
  for (int i = 0; i < n; i++) c[i] = NULL;
  ina_rc_t rc = INA_SUCCESS;
  for (int i = 0; i < n && INA_SUCCEED(rc); i++) {
    /* the checks of iarray_linalg_matmul (src/linalg/matmul.c:31-55) */
    if (a[i]->dtshape.ndim != 2) rc = iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_NDIM), "product %d: ndim", i);
    else if (a[i]->dtshape.shape[1] != b[i]->dtshape.shape[0])
      rc = iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_SHAPE), "product %d: inner dimensions differ", i);
    else if (a[i]->dtshape.dtype != b[i]->dtshape.dtype)
      rc = iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_DTYPE), "product %d: dtypes differ", i);
    if (INA_FAILED(rc)) break;
    iarray_dtshape_t dts;
    memset(&dts, 0, sizeof(dts));
    dts.dtype = a[i]->dtshape.dtype;
    dts.ndim = 2;
    dts.shape[0] = a[i]->dtshape.shape[0];
    dts.shape[1] = b[i]->dtshape.shape[1];
    for (int d = 0; d < 2 && INA_SUCCEED(rc); d++)
      if (dts.shape[d] < storage[i].chunkshape[d])
        rc = iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_CHUNKSHAPE), "product %d: chunkshape", i);
    if (INA_SUCCEED(rc)) rc = iarray_empty(ctx, &dts, &storage[i], &c[i]);
    if (INA_SUCCEED(rc)) rc = iax_ensure_plan(a[i]);
    if (INA_SUCCEED(rc)) rc = iax_ensure_plan(b[i]);
  }
  if (INA_SUCCEED(rc) && !same_signature(&g_plan, n, a, b, storage, dt)) {
    free_batch_plan(&g_plan);
    rc = build_plan(&g_plan, n, a, b, c, storage, dt);
    if (INA_FAILED(rc)) free_batch_plan(&g_plan);
  }
  if (INA_FAILED(rc)) {
    for (int i = 0; i < n; i++)
      if (c[i]) iarray_container_free(ctx, &c[i]);
    return rc;
  }
  iamx_prof_reset();
  /* IARRAY_BATCH_EAGER=1 launches the captured stages one by one (same
   * kernels and arguments), so each can be timed on its own */
  const char *eager = getenv("IARRAY_BATCH_EAGER");
  int e;
  if (eager && atoi(eager) == 1) {
    e = iamx_decode(g_plan.d_lz4, g_plan.n_lz4, g_plan.d_blocks, g_plan.n_blk, g_plan.d_scratch, g_plan.d_dense);
    if (!e)
      e = iamx_gemm_batch(dt, g_plan.d_prods, g_plan.d_tiles, g_plan.ntiles, g_plan.d_dense,
                          (uint8_t *)g_plan.d_dense + g_plan.b_base, g_plan.d_dense_c);
    if (!e) e = iamx_pack_batch(g_plan.d_packs, g_plan.d_pblocks, g_plan.npb, g_plan.d_dense_c, g_plan.d_packed);
  } else {
    e = iamx_graph_launch(g_plan.graph);
  }
  if (e) return iax_dev_error(e, "batch replay");
  int st = iamx_decode_status();
  iax_last_device_ms = iamx_prof_span_ms();
  if (st)
    return st < 0 ? iax_dev_error(st, "batch status")
                  : iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "malformed LZ4 stream (0x%x)", st);
  int64_t *first = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
  if (!first) return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  first[0] = 0;
  for (int i = 0; i < n; i++) first[i + 1] = first[i] + c[i]->nchunks;
  int64_t *ends = (int64_t *)malloc(sizeof(int64_t) * (size_t)(first[n] + 1));
  if (!ends) {
    free(first);
    return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  }
  for (int i = 0; i < n; i++)
    for (int64_t k = 0; k < c[i]->nchunks; k++) ends[first[i] + k] = g_plan.out_off[i] + (k + 1) * c[i]->chunk_nbytes;
  wctx_t x = {c, first, n, NULL, g_plan.out_off, 0};
  const double t0 = iax_now();
  rc = iax_download_pipelined(g_plan.d_packed, (size_t)g_plan.packed_bytes, ends, first[n],
                                       iax_nthreads(&ctx->cfg), write_task, &x, &x.packed);
  iax_last_encode_s = iax_now() - t0;
  free(ends);
  free(first);
  if (INA_FAILED(rc)) return rc;
  if (x.err) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "chunk encode failed");
  return INA_SUCCESS;
}
