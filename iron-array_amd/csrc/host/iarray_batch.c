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
 * hipGraphs, one per phase of consecutive products (about equal output
 * bytes).  A call with the same signature replays the phases one after the
 * other, and behind each queues its C write: the phase's last products are
 * LZ4-encoded and assembled into chunks on the GPU (one encoder pass per group
 * of products) while the host threads encode the others from their
 * downloaded packed blocks; the downloads and GPU encoding of a phase overlap
 * the products of the next.
 * Plain precision only; the mixed forms run product by product.
 */
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#include "iarray_private.h"

/* Default share of a batch's output bytes encoded on the GPU (outputs of 128
 * MiB or more).  Measured on 512 ragged float32 products (2.1 GB of C, 16
 * host threads) in 2 phases, ms per call at shares 0.65 / 0.8 / 0.9 / 1:
 * 60 / 59 / 59 / 59 (one phase, share 0: 93-99). */
#define IAX_BATCH_GPU_SHARE 0.8

/* A batch runs in up to this many phases of consecutive products (about equal
 * output bytes each), one graph per phase: the C write of a phase (downloads
 * and GPU encoding) overlaps the products of the next.  Default 2: on 512
 * ragged products 1 / 2 / 4 phases took 62 / 59-60 / 59-64 ms. */
#define IAX_BATCH_MAX_PHASES 4
#define IAX_BATCH_PHASES 2
typedef struct {
  int p0, p1;         /* products [p0, p1) */
  int l0, nl, b0, nb; /* its LZ4 and block decode jobs */
  int t0, nt, q0, nq; /* its GEMM tiles and C blocks to pack */
  void *graph;        /* decode, GEMM and pack of the phase */
} batch_phase_t;

typedef struct {
  int n;
  iarray_container_t **a, **b;
  uint64_t *gen;
  int64_t *store; /* chunk/block shapes of each output, 4 per product */
  int dt;
  int nphase, nphase_asked; /* phases made, and the count asked for when the plan was built */
  batch_phase_t ph[IAX_BATCH_MAX_PHASES];
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
  int64_t b_base; /* byte offset of the B images in d_dense */
} batch_plan_t;

static batch_plan_t g_plan;

static void free_batch_plan(batch_plan_t *p) {
  for (int k = 0; k < p->nphase; k++) iamx_graph_destroy(p->ph[k].graph);
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

/* IARRAY_BATCH_PHASES (1..4) sets the number of phases */
static int batch_phases(int n) {
  const char *e = getenv("IARRAY_BATCH_PHASES");
  int k = e ? atoi(e) : IAX_BATCH_PHASES;
  if (k < 1) k = 1;
  if (k > IAX_BATCH_MAX_PHASES) k = IAX_BATCH_MAX_PHASES;
  return k < n ? k : n;
}

static int same_signature(const batch_plan_t *p, int n, iarray_container_t **a, iarray_container_t **b,
                          const iarray_storage_t *st, int dt) {
  if (p->nphase < 1 || p->n != n || p->dt != dt || p->nphase_asked != batch_phases(n)) return 0;
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

/* enqueue phase k's decode, GEMM and pack */
static int run_phase(const batch_plan_t *p, int k) {
  const batch_phase_t *q = &p->ph[k];
  int e = iamx_decode(p->d_lz4 + q->l0, q->nl, p->d_blocks + q->b0, q->nb, p->d_scratch, p->d_dense);
  if (!e)
    e = iamx_gemm_batch(p->dt, p->d_prods, p->d_tiles + 3 * (int64_t)q->t0, q->nt, p->d_dense,
                        (uint8_t *)p->d_dense + p->b_base, p->d_dense_c);
  if (!e) e = iamx_pack_batch(p->d_packs, p->d_pblocks + 2 * (int64_t)q->q0, q->nq, p->d_dense_c, p->d_packed);
  return e;
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
  /* phases of about equal output bytes */
  const int nph = batch_phases(n);
  p->nphase = nph;
  p->nphase_asked = nph;
  {
    int k = 0;
    p->ph[0].p0 = 0;
    for (int i = 0; i < n; i++) {
      const int64_t end = p->out_off[i] + c[i]->nchunks * c[i]->chunk_nbytes;
      if (k < nph - 1 && end * nph >= packed * (k + 1) && i < n - 1) {
        p->ph[k].p1 = i + 1;
        p->ph[++k].p0 = i + 1;
      }
    }
    p->ph[k].p1 = n;
    p->nphase = k + 1;
  }
  ina_rc_t rc = INA_SUCCESS;
  iamx_lz4_job_t *lz4 = (iamx_lz4_job_t *)malloc(sizeof(iamx_lz4_job_t) * (size_t)(nlz + 1));
  iamx_block_job_t *blk = (iamx_block_job_t *)malloc(sizeof(iamx_block_job_t) * (size_t)(nblk + 1));
  tile_t *tl = (tile_t *)malloc(sizeof(tile_t) * (size_t)(ntiles + 1));
  int32_t *tiles = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)(ntiles + 1));
  int32_t *pblk = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)(npb + 1));
  if (!lz4 || !blk || !tl || !tiles || !pblk) rc = INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  int64_t il = 0, ib = 0, it = 0, ip = 0, so = 0;
  int ph = 0;
  for (int i = 0; i < n && INA_SUCCEED(rc); i++) {
    if (i == p->ph[ph].p1) ph++;
    if (i == p->ph[ph].p0) {
      p->ph[ph].l0 = (int)il;
      p->ph[ph].b0 = (int)ib;
      p->ph[ph].t0 = (int)it;
      p->ph[ph].q0 = (int)ip;
    }
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
    if (i == p->ph[ph].p1 - 1) {
      p->ph[ph].nl = (int)il - p->ph[ph].l0;
      p->ph[ph].nb = (int)ib - p->ph[ph].b0;
      p->ph[ph].nt = (int)it - p->ph[ph].t0;
      p->ph[ph].nq = (int)ip - p->ph[ph].q0;
    }
  }
  if (INA_SUCCEED(rc)) {
    /* each phase's tiles longest K first, its LZ4 jobs largest first */
    for (int k = 0; k < p->nphase; k++) {
      qsort(tl + p->ph[k].t0, (size_t)p->ph[k].nt, sizeof(tile_t), by_k_desc);
      iax_sort_lz4_jobs(lz4 + p->ph[k].l0, p->ph[k].nl);
    }
    for (int64_t k = 0; k < ntiles; k++) {
      tiles[3 * k] = tl[k].p;
      tiles[3 * k + 1] = tl[k].tm;
      tiles[3 * k + 2] = tl[k].tn;
    }
  }
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
    p->b_base = na * is;
    for (int k = 0; k < p->nphase && INA_SUCCEED(rc); k++) {
      int e = iamx_capture_begin();
      if (!e) e = run_phase(p, k);
      const int e2 = iamx_capture_end(&p->ph[k].graph);
      if (e || e2) rc = iax_dev_error(e ? e : e2, "batch graph capture");
    }
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

/* The C write of a batch, phase by phase.  Each phase's graph is launched,
 * then its downloads and GPU encoding are queued behind it, so they overlap
 * the next phases' products.  In each phase the products holding about
 * `share` of its output bytes (the last ones) are LZ4-encoded and assembled
 * into chunks on the GPU (side stream, one encoder pass per group) and only
 * copied into place on the host; the others are encoded on the host threads
 * from their downloaded packed blocks.  Every chunk lands in the pinned
 * staging buffer at stage[i] + chunk * stride[i] (stride = chunk bytes for a
 * host product, the chunk capacity for a GPU one); the host threads take each
 * chunk as soon as its piece has landed. */
typedef struct {
  iarray_container_t **c;
  const int64_t *first; /* prefix sums of chunk counts, n + 1 entries */
  int n;
  const uint8_t *on_gpu;
  const uint8_t *h;
  const int64_t *stage, *cap;
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
  if (x->on_gpu[i]) {
    const uint8_t *src = x->h + x->stage[i] + nc * x->cap[i];
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
  } else if (iax_encode_chunk(x->h + x->stage[i] + nc * c->chunk_nbytes, (int32_t)c->chunk_nbytes, c->blocksize,
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

#define IAX_PIECES_PER_PHASE 10 /* host downloads per phase */
#define IAX_GROUPS_PER_PHASE 4  /* GPU encoder passes per phase */

static ina_rc_t batch_write(iarray_context_t *ctx, int n, iarray_container_t **c, int eager) {
  const batch_plan_t *p = &g_plan;
  const double share = batch_gpu_share(c, n, p->packed_bytes);
  int64_t *first = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
  int64_t *stage = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  int64_t *gout = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  int64_t *cap = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  int *split = (int *)calloc((size_t)n + 1, sizeof(int));
  uint8_t *on_gpu = (uint8_t *)calloc((size_t)n + 1, 1);
  int64_t *pfirst = (int64_t *)malloc(sizeof(int64_t) * (IAMX_COPY_EVENTS + 1));
  int *hsplit = (int *)malloc(sizeof(int) * (IAX_BATCH_MAX_PHASES + 1)); /* first GPU product of each phase */
  if (!first || !stage || !gout || !cap || !split || !on_gpu || !pfirst || !hsplit) {
    free(first); free(stage); free(gout); free(cap); free(split); free(on_gpu); free(pfirst); free(hsplit);
    return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  }
  ina_rc_t rc = INA_SUCCESS;
  first[0] = 0;
  for (int i = 0; i < n; i++) first[i + 1] = first[i] + c[i]->nchunks;
  /* plan: the GPU products of each phase, staging offsets (host products
   * first, then the GPU products' chunks), pieces and groups */
  int64_t hbytes = 0, gbytes = 0, sizes_bytes = 0;
  for (int k = 0; k < p->nphase; k++) {
    const int p0 = p->ph[k].p0, p1 = p->ph[k].p1;
    const int64_t pb = p->out_off[p1 - 1] + c[p1 - 1]->nchunks * c[p1 - 1]->chunk_nbytes - p->out_off[p0];
    int hs = p1;
    if (share > 0.0) {
      int64_t acc = 0;
      while (hs > p0 && acc < share * (double)pb) {
        hs--;
        acc += c[hs]->nchunks * c[hs]->chunk_nbytes;
      }
    }
    hsplit[k] = hs;
    for (int i = p0; i < hs; i++) {
      stage[i] = hbytes;
      hbytes += c[i]->nchunks * c[i]->chunk_nbytes;
    }
    for (int i = hs; i < p1 && INA_SUCCEED(rc); i++) {
      on_gpu[i] = 1;
      split[i] = iax_split(&c[i]->cfg, c[i]->itemsize, c[i]->blocksize);
      cap[i] = iamx_chunk_capacity(&p->tilings[i], split[i]);
      if (cap[i] < c[i]->chunk_nbytes) rc = iax_dev_error(IAMX_EINVAL, "chunk capacity");
      gout[i] = gbytes;
      gbytes += c[i]->nchunks * cap[i];
      sizes_bytes += c[i]->nchunks * 8;
    }
  }
  const int64_t gbase = (hbytes + 255) / 256 * 256;
  for (int i = 0; i < n; i++)
    if (on_gpu[i]) stage[i] = gbase + gout[i];
  int np = 0, ng = 0;
  pfirst[0] = 0;
  int *gphase = (int *)malloc(sizeof(int) * (IAMX_COPY_EVENTS + 1)); /* phase of each piece */
  uint8_t *is_group = (uint8_t *)calloc(IAMX_COPY_EVENTS + 1, 1);
  if (!gphase || !is_group) rc = INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  for (int k = 0; k < p->nphase && INA_SUCCEED(rc); k++) {
    const int p0 = p->ph[k].p0, p1 = p->ph[k].p1, hs = hsplit[k];
    if (hs > p0) { /* host pieces of about equal bytes */
      const int64_t hb = stage[hs - 1] + c[hs - 1]->nchunks * c[hs - 1]->chunk_nbytes - stage[p0];
      const int64_t piece = hb / IAX_PIECES_PER_PHASE > ((int64_t)4 << 20) ? hb / IAX_PIECES_PER_PHASE : ((int64_t)4 << 20);
      int64_t at = 0;
      int made = 0;
      for (int i = p0; i < hs; i++)
        for (int64_t q = 0; q < c[i]->nchunks; q++) {
          at += c[i]->chunk_nbytes;
          const int last = i == hs - 1 && q == c[i]->nchunks - 1;
          if ((at >= piece * (made + 1) && made < IAX_PIECES_PER_PHASE - 1) || last) {
            gphase[np] = k;
            pfirst[++np] = first[i] + q + 1;
            made++;
          }
        }
    }
    if (p1 > hs) { /* GPU groups of about equal bytes */
      const int m = p1 - hs < IAX_GROUPS_PER_PHASE ? p1 - hs : IAX_GROUPS_PER_PHASE;
      const int64_t gb = gout[p1 - 1] + c[p1 - 1]->nchunks * cap[p1 - 1] - gout[hs];
      int64_t acc = 0;
      int g = 0;
      for (int i = hs; i < p1; i++) {
        acc += c[i]->nchunks * cap[i];
        if ((acc * m >= gb * (g + 1) && g < m - 1) || i == p1 - 1) {
          g++;
          gphase[np] = k;
          is_group[np] = 1;
          pfirst[++np] = first[i + 1];
          ng++;
        }
      }
    }
  }
  uint8_t *h = INA_SUCCEED(rc) ? (uint8_t *)iamx_host_buffer((size_t)(gbase + gbytes + 256)) : NULL;
  if (INA_SUCCEED(rc) && !h) rc = iax_dev_error(IAMX_ENOMEM, "pinned staging buffer");
  /* GPU encoder buffers: one scratch region per group */
  int64_t *gscr = (int64_t *)calloc((size_t)ng + 1, sizeof(int64_t));
  uint8_t *d_scr = NULL, *d_out = NULL;
  int64_t *d_sizes = NULL;
  if (!gscr) rc = INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  if (INA_SUCCEED(rc) && ng > 0) {
    int64_t scr = 0;
    for (int g = 0, q = 0; g < ng && scr >= 0; q++) {
      if (!is_group[q]) continue;
      const int i0 = find_le(first, n, pfirst[q]), i1 = find_le(first, n, pfirst[q + 1] - 1) + 1;
      const int64_t b = iamx_encode_batch_scratch_bytes(i1 - i0, p->tilings + i0, split + i0);
      gscr[g++] = scr;
      scr = b < 0 ? -1 : scr + (b + 255) / 256 * 256;
    }
    d_scr = scr > 0 ? (uint8_t *)iamx_workspace(12, (size_t)scr) : NULL;
    d_out = (uint8_t *)iamx_workspace(14, (size_t)gbytes + 256);
    d_sizes = (int64_t *)iamx_workspace(13, (size_t)sizes_bytes + 256);
    if (scr <= 0 || !d_scr || !d_out || !d_sizes) rc = iax_dev_error(IAMX_ENOMEM, "batch encode buffers");
  }
  /* run: each phase's products, then its downloads and GPU groups */
  int queued = 0, g = 0;
  int64_t zo = 0;
  for (int k = 0; k < p->nphase && INA_SUCCEED(rc); k++) {
    int e = eager ? run_phase(p, k) : iamx_graph_launch(p->ph[k].graph);
    for (; queued < np && gphase[queued] == k && !e; queued++) {
      const int q = queued;
      const int i0 = find_le(first, n, pfirst[q]), i1 = find_le(first, n, pfirst[q + 1] - 1);
      if (!is_group[q]) {
        const int64_t b0 = p->out_off[i0] + (pfirst[q] - first[i0]) * c[i0]->chunk_nbytes;
        const int64_t b1 = p->out_off[i1] + (pfirst[q + 1] - first[i1]) * c[i1]->chunk_nbytes;
        const int64_t s0 = stage[i0] + (pfirst[q] - first[i0]) * c[i0]->chunk_nbytes;
        e = iamx_d2h_async(h + s0, (const uint8_t *)p->d_packed + b0, (size_t)(b1 - b0), q);
      } else {
        const int m = i1 + 1 - i0;
        if ((e = iamx_side_wait()) == 0 && (e = iamx_use_stream(1)) == 0) {
          e = iamx_encode_batch(m, p->tilings + i0, split + i0, p->d_packed, p->out_off + i0, d_scr + gscr[g], d_out,
                                gout + i0, d_sizes + zo);
          const int64_t b0 = gout[i0], b1 = gout[i1] + c[i1]->nchunks * cap[i1];
          if (!e) e = iamx_d2h_async(h + gbase + b0, d_out + b0, (size_t)(b1 - b0), q);
        }
        iamx_use_stream(0);
        zo += first[i1 + 1] - first[i0];
        g++;
      }
    }
    if (e) rc = iax_dev_error(e, "batch phase");
  }
  if (INA_FAILED(rc)) {
    for (int q = 0; q < queued; q++) (void)iamx_event_sync(q);
  } else {
    wctx_t x;
    x.c = c;
    x.first = first;
    x.n = n;
    x.on_gpu = on_gpu;
    x.h = h;
    x.stage = stage;
    x.cap = cap;
    x.pfirst = pfirst;
    x.npieces = np;
    x.filtered = c[0]->cfg.compression_level > 0;
    for (int q = 0; q < IAMX_COPY_EVENTS; q++) atomic_init(&x.landed[q], 0);
    atomic_init(&x.err, 0);
    iax_parallel_for(iax_nthreads(&ctx->cfg), first[n], write_task, &x);
    if (atomic_load(&x.err)) {
      for (int q = 0; q < np; q++) (void)iamx_event_sync(q);
      rc = iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "chunk encode failed");
    }
    for (int i = 0; i < n && INA_SUCCEED(rc); i++) {
      if (!on_gpu[i]) continue;
      iax_last_gpu_enc_in += c[i]->nchunks * c[i]->chunk_nbytes;
      for (int64_t q = 0; q < c[i]->nchunks; q++) iax_last_gpu_enc_out += c[i]->csizes[q];
    }
  }
  free(first); free(stage); free(gout); free(cap); free(split); free(on_gpu); free(pfirst);
  free(hsplit); free(gphase); free(is_group); free(gscr);
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
      if (INA_FAILED(rc)) {
        for (int j = 0; j < i; j++) iarray_container_free(ctx, &c[j]);
        return rc;
      }
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
    if (INA_SUCCEED(rc)) rc = iax_container_new(ctx, &dts, &storage[i], &c[i]);
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
  const double t0 = iax_now();
  iax_last_gpu_enc_in = iax_last_gpu_enc_out = 0;
  rc = batch_write(ctx, n, c, eager && atoi(eager) == 1);
  iax_last_encode_s = iax_now() - t0;
  /* the decode status of every phase (a malformed stream leaves its flag) */
  const ina_rc_t st = iax_check_decode_status();
  iax_last_device_ms = iamx_prof_span_ms();
  if (INA_SUCCEED(rc)) rc = st;
  for (int i = 0; i < n && INA_SUCCEED(rc); i++) rc = iax_persist(c[i]);
  /* a failed batch leaves no partly written outputs behind */
  if (INA_FAILED(rc))
    for (int i = 0; i < n; i++) iarray_container_free(ctx, &c[i]);
  return rc;
}
