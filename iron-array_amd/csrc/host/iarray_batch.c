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
 * a hipGraph.  A call with the same signature replays the graph; then a
 * suffix of the products is LZ4-encoded and assembled into chunks on the GPU
 * (one encoder pass per group of products) while the host threads encode the
 * others from their downloaded packed blocks.
 * Plain precision only; the mixed forms run product by product.
 */
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#include "iarray_private.h"

/* Default share of a batch's output bytes encoded on the GPU (outputs of 128
 * MiB or more).  Measured on 512 ragged float32 products (1.3 GB of C, 16
 * host threads), ms per call at shares 0 / 0.3 / 0.5 / 0.6 / 0.7 / 0.8 / 1:
 * 99 / 73 / 63-72 / 62 / 60-61 / 65 / 70. */
#define IAX_BATCH_GPU_SHARE 0.65

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
  iamx_tiling_t *tilings; /* per product: C's tiling (for the GPU encoder) */
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
  free(p->tilings);
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
  p->tilings = (iamx_tiling_t *)calloc((size_t)n, sizeof(iamx_tiling_t));
  iamx_batch_prod_t *prods = (iamx_batch_prod_t *)calloc((size_t)n, sizeof(iamx_batch_prod_t));
  iamx_batch_pack_t *packs = (iamx_batch_pack_t *)calloc((size_t)n, sizeof(iamx_batch_pack_t));
  if (!p->a || !p->b || !p->gen || !p->store || !p->out_off || !p->tilings || !prods || !packs) {
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
    p->tilings[i] = *t;
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

/* The C write of a batch.  Products [0, hsplit) are LZ4-encoded on the host
 * threads from their packed blocks; products [hsplit, n) are encoded and
 * assembled into chunks on the GPU (on the side stream, one encoder pass per
 * group of products) and only copied into place on the host.  Every
 * download is queued up front on the copy stream, host pieces first; the
 * host threads take each chunk as soon as its piece or group has landed.
 * Staging: the host products' packed bytes at their packed offsets, then,
 * from gbase, the GPU products' chunks at gout_off[i] + chunk * cap[i]. */
typedef struct {
  iarray_container_t **c;
  const int64_t *first; /* prefix sums of chunk counts, n + 1 entries */
  int n, hsplit;
  const uint8_t *h;
  const int64_t *out_off, *gout_off, *cap;
  int64_t gbase;
  const int64_t *pfirst; /* piece k covers global chunks [pfirst[k], pfirst[k + 1]) */
  int npieces;
  int filtered;
  atomic_int landed[IAMX_COPY_EVENTS];
  atomic_int err;
} wctx_t;

static int find_le(const int64_t *v, int n, int64_t g) { /* last index i < n with v[i] <= g */
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (v[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

static void write_task(void *vctx, int64_t g, int worker) {
  (void)worker;
  wctx_t *x = (wctx_t *)vctx;
  const int k = find_le(x->pfirst, x->npieces, g);
  if (!atomic_load(&x->landed[k])) {
    if (iamx_event_sync(k) != 0) {
      atomic_store(&x->err, 1);
      return;
    }
    atomic_store(&x->landed[k], 1);
  }
  const int i = find_le(x->first, x->n, g);
  iarray_container_t *c = x->c[i];
  const int64_t nc = g - x->first[i];
  free(c->chunks[nc]);
  c->chunks[nc] = NULL;
  if (i >= x->hsplit) {
    const uint8_t *src = x->h + x->gbase + x->gout_off[i] + nc * x->cap[i];
    int32_t sz;
    memcpy(&sz, src + 12, 4); /* cbytes of the assembled chunk */
    uint8_t *q = sz > 32 && sz <= x->cap[i] ? (uint8_t *)malloc((size_t)sz) : NULL;
    if (!q) {
      atomic_store(&x->err, 1);
      return;
    }
    memcpy(q, src, (size_t)sz);
    c->chunks[nc] = q;
    c->csizes[nc] = sz;
  } else if (iax_encode_chunk(x->h + x->out_off[i] + nc * c->chunk_nbytes, (int32_t)c->chunk_nbytes, c->blocksize,
                              c->itemsize, &c->cfg, x->filtered, &c->chunks[nc], &c->csizes[nc]) != 0) {
    atomic_store(&x->err, 1);
  }
}

/* Share of the output bytes encoded on the GPU: IARRAY_GPU_ENCODE (0..1), or
 * the default below; 0 when the codec settings are not the GPU encoder's. */
static double batch_gpu_share(iarray_container_t **c, int n, int64_t total) {
  const iarray_container_t *c0 = c[0];
  const int ok = c0->cfg.compression_level > 0 &&
                 (c0->cfg.compression_codec == IARRAY_COMPRESSION_LZ4 ||
                  c0->cfg.compression_codec == IARRAY_COMPRESSION_LZ4HC) &&
                 c0->itemsize <= IAMX_MAX_STREAMS;
  if (!ok) return 0.0;
  for (int i = 0; i < n; i++)
    if (c[i]->chunk_nbytes % c[i]->blocksize != 0 || c[i]->chunk_nbytes > INT32_MAX / 2) return 0.0;
  const char *env = getenv("IARRAY_GPU_ENCODE");
  if (!env) return total >= ((int64_t)128 << 20) ? IAX_BATCH_GPU_SHARE : 0.0;
  const double v = atof(env);
  if (!(v > 0.0)) return 0.0;
  return v > 1.0 ? 1.0 : v;
}

static ina_rc_t batch_write(iarray_context_t *ctx, int n, iarray_container_t **c) {
  const batch_plan_t *p = &g_plan;
  const int64_t total = p->packed_bytes;
  const double share = batch_gpu_share(c, n, total);
  /* the GPU takes a suffix of the products holding about `share` of the bytes */
  int hsplit = n;
  if (share > 0.0) {
    int64_t acc = 0;
    while (hsplit > 0 && acc < share * (double)total) {
      hsplit--;
      acc += c[hsplit]->nchunks * c[hsplit]->chunk_nbytes;
    }
  }
  const int ngpu = n - hsplit;
  int64_t *first = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
  int64_t *gout = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  int64_t *cap = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  int *split = (int *)calloc((size_t)n + 1, sizeof(int));
  int64_t *pfirst = (int64_t *)malloc(sizeof(int64_t) * (IAMX_COPY_EVENTS + 1));
  int *gprod = (int *)malloc(sizeof(int) * (IAMX_COPY_EVENTS + 1)); /* first product of each GPU group */
  if (!first || !gout || !cap || !split || !pfirst || !gprod) {
    free(first); free(gout); free(cap); free(split); free(pfirst); free(gprod);
    return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  }
  first[0] = 0;
  for (int i = 0; i < n; i++) first[i + 1] = first[i] + c[i]->nchunks;
  ina_rc_t rc = INA_SUCCESS;
  int64_t gbytes = 0, sizes_bytes = 0;
  for (int i = hsplit; i < n && INA_SUCCEED(rc); i++) {
    split[i] = iax_split(&c[i]->cfg, c[i]->itemsize, c[i]->blocksize);
    cap[i] = iamx_chunk_capacity(&p->tilings[i], split[i]);
    if (cap[i] < c[i]->chunk_nbytes) rc = iax_dev_error(IAMX_EINVAL, "chunk capacity");
    gout[i] = gbytes;
    gbytes += c[i]->nchunks * cap[i];
    sizes_bytes += c[i]->nchunks * 8;
  }
  /* pieces: host products in downloads of about 4 MB or more (at most 40),
   * then the GPU products in groups of about equal bytes (at most 16) */
  const int64_t hbytes = hsplit < n ? p->out_off[hsplit] : total; /* packed bytes of the host products */
  int np = 0;
  pfirst[0] = 0;
  if (hsplit > 0) {
    int64_t piece = hbytes / 40 > ((int64_t)4 << 20) ? hbytes / 40 : ((int64_t)4 << 20);
    int64_t at = 0;
    for (int i = 0; i < hsplit; i++)
      for (int64_t k = 0; k < c[i]->nchunks; k++) {
        at += c[i]->chunk_nbytes;
        const int last = i == hsplit - 1 && k == c[i]->nchunks - 1;
        if ((at >= piece * (np + 1) && np < 39) || last) pfirst[++np] = first[i] + k + 1;
      }
  }
  const int nhp = np;
  if (ngpu > 0) {
    const int ng = ngpu < 16 ? ngpu : 16;
    int64_t acc = 0;
    int g = 0;
    gprod[0] = hsplit;
    for (int i = hsplit; i < n; i++) {
      acc += c[i]->nchunks * cap[i];
      if ((acc * ng >= gbytes * (g + 1) && g < ng - 1) || i == n - 1) {
        pfirst[++np] = first[i + 1];
        gprod[++g] = i + 1;
      }
    }
  }
  const int64_t gbase = (hbytes + 255) / 256 * 256;
  uint8_t *h = INA_SUCCEED(rc) ? (uint8_t *)iamx_host_buffer((size_t)(gbase + gbytes + 256)) : NULL;
  if (INA_SUCCEED(rc) && !h) rc = iax_dev_error(IAMX_ENOMEM, "pinned staging buffer");
  int e = 0;
  /* host pieces first on the copy stream */
  for (int k = 0; k < nhp && INA_SUCCEED(rc) && !e; k++) {
    const int i0 = find_le(first, n, pfirst[k]), i1 = find_le(first, n, pfirst[k + 1] - 1);
    const int64_t b0 = p->out_off[i0] + (pfirst[k] - first[i0]) * c[i0]->chunk_nbytes;
    const int64_t b1 = p->out_off[i1] + (pfirst[k + 1] - first[i1]) * c[i1]->chunk_nbytes;
    e = iamx_d2h_async(h + b0, (const uint8_t *)p->d_packed + b0, (size_t)(b1 - b0), k);
  }
  if (e) rc = iax_dev_error(e, "download");
  /* then the GPU groups: encode on the side stream, download behind each */
  if (INA_SUCCEED(rc) && ngpu > 0) {
    int64_t scr = 0;
    for (int g = 0; g < np - nhp && scr >= 0; g++) {
      const int64_t b = iamx_encode_batch_scratch_bytes(gprod[g + 1] - gprod[g], p->tilings + gprod[g], split + gprod[g]);
      scr = b < 0 ? -1 : scr + (b + 255) / 256 * 256;
    }
    uint8_t *d_scr = scr > 0 ? (uint8_t *)iamx_workspace(12, (size_t)scr) : NULL;
    uint8_t *d_out = (uint8_t *)iamx_workspace(14, (size_t)gbytes + 256);
    int64_t *d_sizes = (int64_t *)iamx_workspace(13, (size_t)sizes_bytes + 256);
    if (scr <= 0 || !d_scr || !d_out || !d_sizes) rc = iax_dev_error(IAMX_ENOMEM, "batch encode buffers");
    if (INA_SUCCEED(rc) && ((e = iamx_side_wait()) != 0 || (e = iamx_use_stream(1)) != 0))
      rc = iax_dev_error(e, "encode");
    int64_t so = 0, zo = 0;
    for (int g = 0; g < np - nhp && INA_SUCCEED(rc); g++) {
      const int i0 = gprod[g], m = gprod[g + 1] - gprod[g];
      e = iamx_encode_batch(m, p->tilings + i0, split + i0, p->d_packed, p->out_off + i0, d_scr + so, d_out,
                            gout + i0, d_sizes + zo);
      const int64_t b0 = gout[i0], b1 = g + 1 < np - nhp ? gout[gprod[g + 1]] : gbytes;
      if (!e) e = iamx_d2h_async(h + gbase + b0, d_out + b0, (size_t)(b1 - b0), nhp + g);
      if (e) rc = iax_dev_error(e, "encode");
      so += (iamx_encode_batch_scratch_bytes(m, p->tilings + i0, split + i0) + 255) / 256 * 256;
      zo += first[gprod[g + 1]] - first[i0];
    }
    iamx_use_stream(0);
  }
  if (INA_FAILED(rc)) {
    for (int k = 0; k < np; k++) (void)iamx_event_sync(k);
  } else {
    wctx_t x;
    x.c = c;
    x.first = first;
    x.n = n;
    x.hsplit = hsplit;
    x.h = h;
    x.out_off = p->out_off;
    x.gout_off = gout;
    x.cap = cap;
    x.gbase = gbase;
    x.pfirst = pfirst;
    x.npieces = np;
    x.filtered = c[0]->cfg.compression_level > 0;
    for (int k = 0; k < IAMX_COPY_EVENTS; k++) atomic_init(&x.landed[k], 0);
    atomic_init(&x.err, 0);
    iax_parallel_for(iax_nthreads(&ctx->cfg), first[n], write_task, &x);
    if (atomic_load(&x.err)) {
      for (int k = 0; k < np; k++) (void)iamx_event_sync(k);
      rc = iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "chunk encode failed");
    }
    for (int i = hsplit; i < n && INA_SUCCEED(rc); i++) {
      iax_last_gpu_enc_in += c[i]->nchunks * c[i]->chunk_nbytes;
      for (int64_t k = 0; k < c[i]->nchunks; k++) iax_last_gpu_enc_out += c[i]->csizes[k];
    }
  }
  free(first); free(gout); free(cap); free(split); free(pfirst); free(gprod);
  return rc;
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
  const double t0 = iax_now();
  iax_last_gpu_enc_in = iax_last_gpu_enc_out = 0;
  const ina_rc_t rc2 = batch_write(ctx, n, c);
  iax_last_encode_s = iax_now() - t0;
  return rc2;
}
