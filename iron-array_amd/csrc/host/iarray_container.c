/*
 * iarray_container.c -- containers of libiarray.so: Caterva-style two-level
 * tiling, constructors, read-back, views, and the device mirror + decode plan
 * the GPU kernels consume.
 *
 * Tiling (restated from the reference's index math, src/iarray_zproxy.c:46-85,
 * src/linalg/matmul/gemm.c:25-69): extshape = ceil(shape/chunk)*chunk,
 * extchunkshape = ceil(chunk/block)*block; chunks are numbered row-major over
 * extshape/chunkshape, blocks row-major inside a chunk over
 * extchunkshape/blockshape, each block is a row-major blockshape tile and the
 * chunk's uncompressed content is its blocks back to back
 * (chunk bytes = prod(extchunkshape)*itemsize, blocksize =
 * prod(blockshape)*itemsize).  Element (c_nd*chunkshape + b_nd*blockshape + i)
 * is valid when it lies inside both the chunk's chunkshape and the array
 * shape; everything else is padding, written as zeros here and never read.
 *
 * Entry points mirror src/iarray_constructor.c:16-305 (empty, from_buffer,
 * to_buffer), src/constructors/iarray_constructor_arange.c:40 and
 * src/constructors/iarray_constructor_linspace.c:117 (value = start +
 * flat_index * step, step = (stop-start)/(n-1) for linspace),
 * src/iarray_container.c:500-541 / 628-736 (get_slice_buffer, transposed
 * views), src/iarray_container.c:1200-1287 (info, almost_equal) and
 * src/linalg/transpose.c:18-45.
 */
#include <math.h>
#include <pthread.h>
#include <sys/mman.h>
#include <stdlib.h>
#include <string.h>

#include <stdatomic.h>

#include "iarray_private.h"

#define FAIL_NULL(p) \
  do {               \
    if (!(p)) return iax_set_error(INA_ERROR(INA_ERR_NULL_POINTER), "null argument: %s", #p); \
  } while (0)

static inline int32_t rd32(const uint8_t *p) {
  int32_t v;
  memcpy(&v, p, 4);
  return v;
}

int64_t iax_dtype_size(iarray_data_type_t t) {
  switch (t) {
    case IARRAY_DATA_TYPE_DOUBLE: case IARRAY_DATA_TYPE_INT64: case IARRAY_DATA_TYPE_UINT64: return 8;
    case IARRAY_DATA_TYPE_FLOAT: case IARRAY_DATA_TYPE_INT32: case IARRAY_DATA_TYPE_UINT32: return 4;
    case IARRAY_DATA_TYPE_INT16: case IARRAY_DATA_TYPE_UINT16: return 2;
    case IARRAY_DATA_TYPE_INT8: case IARRAY_DATA_TYPE_UINT8: case IARRAY_DATA_TYPE_BOOL: return 1;
    default: return -1;
  }
}

static int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

static ina_rc_t check_config(const iarray_config_t *cfg) {
  if (cfg->compression_codec != IARRAY_COMPRESSION_LZ4 && cfg->compression_codec != IARRAY_COMPRESSION_LZ4HC)
    return iax_set_error(INA_ERROR(INA_ERR_NOT_SUPPORTED | IARRAY_ES_BLOSC),
                         "only the LZ4 codecs are on the GPU decode path (codec %d)", cfg->compression_codec);
  if (cfg->filter_flags & (IARRAY_COMP_DELTA | IARRAY_COMP_TRUNC_PREC))
    return iax_set_error(INA_ERROR(INA_ERR_NOT_SUPPORTED | IARRAY_ES_BLOSC), "DELTA/TRUNC_PREC filters unsupported");
  if (cfg->compression_level < 0 || cfg->compression_level > 9)
    return iax_set_error(INA_ERROR(INA_ERR_INVALID_ARGUMENT | IARRAY_ES_BLOSC), "clevel %d", cfg->compression_level);
  return INA_SUCCESS;
}

ina_rc_t iax_container_new(iarray_context_t *ctx, iarray_dtshape_t *dtshape, iarray_storage_t *storage,
                           iarray_container_t **cp) {
  FAIL_NULL(ctx);
  FAIL_NULL(dtshape);
  FAIL_NULL(storage);
  FAIL_NULL(cp);
  if (dtshape->ndim < 1 || dtshape->ndim > 2)
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_NDIM), "ndim %d: the matmul path handles 1-D and 2-D arrays",
                         dtshape->ndim);
  const int64_t is = iax_dtype_size(dtshape->dtype);
  if (is < 0) return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_DTYPE), "dtype %d", dtshape->dtype);
  if (storage->urlpath && !storage->urlpath[0])
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_STORAGE), "empty urlpath");
  ina_rc_t rc = check_config(&ctx->cfg);
  if (INA_FAILED(rc)) return rc;
  for (int i = 0; i < dtshape->ndim; i++) {
    if (dtshape->shape[i] < 0) return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_SHAPE), "negative shape");
    if (storage->chunkshape[i] < 1)
      return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_CHUNKSHAPE), "chunkshape[%d] = %lld", i,
                           (long long)storage->chunkshape[i]);
    if (storage->blockshape[i] < 1 || storage->blockshape[i] > storage->chunkshape[i])
      return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_BLOCKSHAPE), "blockshape[%d] = %lld", i,
                           (long long)storage->blockshape[i]);
  }
  iarray_container_t *c = (iarray_container_t *)calloc(1, sizeof(*c));
  if (!c) return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  c->dtshape = *dtshape;
  c->dtshape.dtype_size = (int32_t)is;
  c->storage = *storage;
  c->storage.urlpath = NULL;
  if (storage->urlpath) {
    /* stored on disk: the frame at urlpath follows the chunks (iax_persist) */
    if (!(c->storage.urlpath = strdup(storage->urlpath))) {
      free(c);
      return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
    }
    c->on_disk = true;
  }
  c->cfg = ctx->cfg;
  c->itemsize = (int)is;
  c->ndim = dtshape->ndim;
  for (int i = 0; i < 2; i++) {
    const int one = (c->ndim == 1 && i == 0);
    const int s = c->ndim == 2 ? i : 0;
    c->shape[i] = one ? 1 : dtshape->shape[s];
    c->chunk[i] = one ? 1 : storage->chunkshape[s];
    c->block[i] = one ? 1 : storage->blockshape[s];
    c->extshape[i] = cdiv(c->shape[i] > 0 ? c->shape[i] : 1, c->chunk[i]) * c->chunk[i];
    c->extchunk[i] = cdiv(c->chunk[i], c->block[i]) * c->block[i];
    c->nchunks_ax[i] = c->extshape[i] / c->chunk[i];
    c->nblocks_ax[i] = c->extchunk[i] / c->block[i];
  }
  c->nchunks = c->nchunks_ax[0] * c->nchunks_ax[1];
  c->chunk_nbytes = c->extchunk[0] * c->extchunk[1] * is;
  const int64_t bsz = c->block[0] * c->block[1] * is;
  if (c->chunk_nbytes > INT32_MAX || bsz > INT32_MAX)
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_CHUNKSHAPE), "chunks are limited to 2 GB (Blosc2 int32 sizes)");
  c->blocksize = (int32_t)bsz;
  c->chunks = (uint8_t **)calloc((size_t)c->nchunks, sizeof(uint8_t *));
  c->csizes = (int32_t *)calloc((size_t)c->nchunks, sizeof(int32_t));
  c->arena_off = (int64_t *)calloc((size_t)c->nchunks, sizeof(int64_t));
  if (!c->chunks || !c->csizes || !c->arena_off) {
    free(c->chunks); free(c->csizes); free(c->arena_off); free(c->storage.urlpath); free(c);
    return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  }
  *cp = c;
  return INA_SUCCESS;
}

static void free_plan(iax_plan_t *p) {
  iamx_free(p->d_lz4);
  iamx_free(p->d_blocks);
  iamx_free(p->d_blk_rc);
  free(p->h_lz4);
  free(p->h_blocks);
  memset(p, 0, sizeof(*p));
}

static void drop_device(iarray_container_t *c) {
  if (c->plan_valid) free_plan(&c->plan);
  c->plan_valid = 0;
  iamx_free(c->d_arena);
  c->d_arena = NULL;
  c->device_valid = 0;
}

INA_API(void) iarray_container_free(iarray_context_t *ctx, iarray_container_t **cp) {
  (void)ctx;
  if (!cp || !*cp) return;
  iarray_container_t *c = *cp;
  if (!c->viewed) {
    const uint8_t *m0 = c->lazy.map, *m1 = c->lazy.map + c->lazy.map_len;
    for (int64_t i = 0; i < c->nchunks; i++)
      if (!(m0 && c->chunks[i] >= m0 && c->chunks[i] < m1)) free(c->chunks[i]); /* mapped chunks are borrowed */
    if (c->lazy.map) munmap(c->lazy.map, c->lazy.map_len);
    free(c->lazy.off);
    drop_device(c);
    free(c->chunks);
    free(c->csizes);
    free(c->arena_off);
    free(c->storage.urlpath); /* the frame on disk stays */
  } else {
    iamx_free(c->d_tblocks);
  }
  free(c);
  *cp = NULL;
}

INA_API(ina_rc_t) iarray_empty(iarray_context_t *ctx, iarray_dtshape_t *dtshape, iarray_storage_t *storage,
                               iarray_container_t **container) {
  ina_rc_t rc = iax_container_new(ctx, dtshape, storage, container);
  return INA_SUCCEED(rc) ? iax_persist(*container) : rc;
}

/* A special-value chunk: a 32-byte header only, the value kind (1 zeros,
 * 2 NaNs, 4 uninitialized) in bits 4..6 of byte 31. */
uint8_t *iax_special_chunk(int kind, int32_t nbytes, int32_t blocksize, int typesize) {
  uint8_t *h = (uint8_t *)calloc(1, 32);
  if (!h) return NULL;
  const int32_t cb = 32;
  h[0] = 5; h[1] = 1; h[2] = 0x05 | 0x10 | (1 << 5); h[3] = (uint8_t)typesize;
  memcpy(h + 4, &nbytes, 4); memcpy(h + 8, &blocksize, 4); memcpy(h + 12, &cb, 4);
  h[31] = (uint8_t)((kind & 7) << 4);
  return h;
}

static uint8_t *zero_chunk(const iarray_container_t *c, int32_t *csize) {
  *csize = 32;
  return iax_special_chunk(1, (int32_t)c->chunk_nbytes, c->blocksize, c->itemsize);
}

INA_API(ina_rc_t) iarray_zeros(iarray_context_t *ctx, iarray_dtshape_t *dtshape, iarray_storage_t *storage,
                               iarray_container_t **container) {
  ina_rc_t rc = iax_container_new(ctx, dtshape, storage, container);
  if (INA_FAILED(rc)) return rc;
  iarray_container_t *c = *container;
  for (int64_t i = 0; i < c->nchunks; i++)
    if (!(c->chunks[i] = zero_chunk(c, &c->csizes[i]))) return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  return iax_persist(c);
}

/* ---- host encode of a row-major buffer --------------------------------------------------- */
typedef struct {
  iarray_container_t *c;
  const uint8_t *buf;
  uint8_t **scratch;
  int err;
} enc_ctx_t;

static void encode_chunk_task(void *vctx, int64_t nc, int worker) {
  enc_ctx_t *x = (enc_ctx_t *)vctx;
  iarray_container_t *c = x->c;
  const int is = c->itemsize;
  uint8_t *chunk = x->scratch[worker];
  memset(chunk, 0, (size_t)c->chunk_nbytes);
  const int64_t ci0 = nc / c->nchunks_ax[1], ci1 = nc % c->nchunks_ax[1];
  for (int64_t bi0 = 0; bi0 < c->nblocks_ax[0]; bi0++)
    for (int64_t bi1 = 0; bi1 < c->nblocks_ax[1]; bi1++) {
      uint8_t *blk = chunk + (bi0 * c->nblocks_ax[1] + bi1) * (int64_t)c->blocksize;
      const int64_t r_in = bi0 * c->block[0], c_in = bi1 * c->block[1];
      const int64_t row0 = ci0 * c->chunk[0] + r_in, col0 = ci1 * c->chunk[1] + c_in;
      int64_t vr = c->block[0], vc = c->block[1];
      if (c->chunk[0] - r_in < vr) vr = c->chunk[0] - r_in;
      if (c->shape[0] - row0 < vr) vr = c->shape[0] - row0;
      if (c->chunk[1] - c_in < vc) vc = c->chunk[1] - c_in;
      if (c->shape[1] - col0 < vc) vc = c->shape[1] - col0;
      if (vr <= 0 || vc <= 0) continue; /* all-padding block: stays zero */
      for (int64_t r = 0; r < vr; r++)
        memcpy(blk + r * c->block[1] * is, x->buf + ((row0 + r) * c->shape[1] + col0) * is, (size_t)(vc * is));
    }
  if (c->chunks[nc]) free(c->chunks[nc]);
  c->chunks[nc] = NULL;
  if (iax_encode_chunk(chunk, (int32_t)c->chunk_nbytes, c->blocksize, is, &c->cfg, 0, &c->chunks[nc],
                       &c->csizes[nc]) != 0)
    x->err = 1;
}

static ina_rc_t encode_buffer(iarray_container_t *c, const void *buf) {
  const int nt = iax_nthreads(&c->cfg);
  enc_ctx_t x = {c, (const uint8_t *)buf, NULL, 0};
  x.scratch = (uint8_t **)calloc((size_t)nt, sizeof(uint8_t *));
  if (!x.scratch) return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  for (int i = 0; i < nt; i++)
    if (!(x.scratch[i] = (uint8_t *)malloc((size_t)c->chunk_nbytes))) x.err = 1;
  if (!x.err) iax_parallel_for(nt, c->nchunks, encode_chunk_task, &x);
  for (int i = 0; i < nt; i++) free(x.scratch[i]);
  free(x.scratch);
  drop_device(c);
  if (x.err) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "chunk encode failed");
  return INA_SUCCESS;
}

INA_API(ina_rc_t) iarray_from_buffer(iarray_context_t *ctx, iarray_dtshape_t *dtshape, void *buffer, int64_t buflen,
                                     iarray_storage_t *storage, iarray_container_t **container) {
  FAIL_NULL(buffer);
  ina_rc_t rc = iax_container_new(ctx, dtshape, storage, container);
  if (INA_FAILED(rc)) return rc;
  iarray_container_t *c = *container;
  if (c->shape[0] * c->shape[1] * c->itemsize > buflen)
    return iax_set_error(INA_ERROR(IARRAY_ERR_TOO_SMALL_BUFFER), "The size of the buffer is not enough");
  rc = encode_buffer(c, buffer);
  return INA_SUCCEED(rc) ? iax_persist(c) : rc;
}

static ina_rc_t from_ramp(iarray_context_t *ctx, iarray_dtshape_t *dtshape, double start, double step,
                          iarray_storage_t *storage, iarray_container_t **container) {
  int64_t n = 1;
  for (int i = 0; i < dtshape->ndim; i++) n *= dtshape->shape[i];
  const int64_t is = iax_dtype_size(dtshape->dtype);
  if (is < 0) return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_DTYPE), "dtype %d", dtshape->dtype);
  void *buf = malloc((size_t)(n > 0 ? n : 1) * (size_t)is);
  if (!buf) return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  for (int64_t i = 0; i < n; i++) {
    const double v = start + (double)i * step;
    switch (dtshape->dtype) {
      case IARRAY_DATA_TYPE_DOUBLE: ((double *)buf)[i] = v; break;
      case IARRAY_DATA_TYPE_FLOAT: ((float *)buf)[i] = (float)v; break;
      case IARRAY_DATA_TYPE_INT64: ((int64_t *)buf)[i] = (int64_t)v; break;
      case IARRAY_DATA_TYPE_INT32: ((int32_t *)buf)[i] = (int32_t)v; break;
      case IARRAY_DATA_TYPE_INT16: ((int16_t *)buf)[i] = (int16_t)v; break;
      case IARRAY_DATA_TYPE_INT8: ((int8_t *)buf)[i] = (int8_t)v; break;
      case IARRAY_DATA_TYPE_UINT64: ((uint64_t *)buf)[i] = (uint64_t)v; break;
      case IARRAY_DATA_TYPE_UINT32: ((uint32_t *)buf)[i] = (uint32_t)v; break;
      case IARRAY_DATA_TYPE_UINT16: ((uint16_t *)buf)[i] = (uint16_t)v; break;
      case IARRAY_DATA_TYPE_UINT8: ((uint8_t *)buf)[i] = (uint8_t)v; break;
      case IARRAY_DATA_TYPE_BOOL: ((uint8_t *)buf)[i] = v != 0.0; break;
      default: break;
    }
  }
  ina_rc_t rc = iarray_from_buffer(ctx, dtshape, buf, n * is, storage, container);
  free(buf);
  return rc;
}

INA_API(ina_rc_t) iarray_arange(iarray_context_t *ctx, iarray_dtshape_t *dtshape, double start, double step,
                                iarray_storage_t *storage, iarray_container_t **container) {
  FAIL_NULL(dtshape);
  return from_ramp(ctx, dtshape, start, step, storage, container);
}

INA_API(ina_rc_t) iarray_linspace(iarray_context_t *ctx, iarray_dtshape_t *dtshape, double start, double stop,
                                  iarray_storage_t *storage, iarray_container_t **container) {
  FAIL_NULL(dtshape);
  int64_t n = 1;
  for (int i = 0; i < dtshape->ndim; i++) n *= dtshape->shape[i];
  return from_ramp(ctx, dtshape, start, (stop - start) / ((double)n - 1), storage, container);
}

/* ---- device mirror ------------------------------------------------------------------------ */
ina_rc_t iax_materialize(iarray_container_t *c) {
  if (c->viewed) c = c->viewed;
  if (!c->lazy.pending) return INA_SUCCESS;
  const int64_t hsize = c->lazy.hsize, opos = c->lazy.opos;
  for (int64_t i = 0; i < c->nchunks; i++) {
    const int64_t off = c->lazy.off[i];
    if (off < 0) continue;
    /* the offset was bounded at open (off <= opos - hsize - 32); the chunk's
     * size comes from its own header and must end inside the data section */
    const int64_t pos = hsize + off;
    int32_t cs;
    memcpy(&cs, c->lazy.map + pos + 12, 4);
    if (cs < 32 || cs > opos - pos)
      return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "frame: chunk %lld size", (long long)i);
    c->chunks[i] = c->lazy.map + pos;
    c->csizes[i] = cs;
    c->lazy.off[i] = -1;
  }
  c->lazy.pending = false;
  return INA_SUCCESS;
}

ina_rc_t iax_ensure_device(iarray_container_t *c) {
  if (c->viewed) return iax_ensure_device(c->viewed);
  if (c->device_valid) return INA_SUCCESS;
  ina_rc_t mrc = iax_materialize(c);
  if (INA_FAILED(mrc)) return mrc;
  int64_t total = 0;
  for (int64_t i = 0; i < c->nchunks; i++) {
    c->arena_off[i] = total;
    total += (c->chunks[i] ? ((c->csizes[i] + 15) & ~15) : 0);
  }
  iamx_free(c->d_arena);
  c->d_arena = (uint8_t *)iamx_malloc((size_t)(total > 0 ? total : 16));
  if (!c->d_arena) return iax_dev_error(iamx_init() ? iamx_init() : IAMX_ENOMEM, "device arena");
  uint8_t *host = (uint8_t *)malloc((size_t)(total > 0 ? total : 16));
  if (!host) return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  for (int64_t i = 0; i < c->nchunks; i++)
    if (c->chunks[i]) memcpy(host + c->arena_off[i], c->chunks[i], (size_t)c->csizes[i]);
  int rc = iamx_h2d(c->d_arena, host, (size_t)total);
  free(host);
  if (rc) return iax_dev_error(rc, "upload chunks");
  c->device_valid = 1;
  static uint64_t next_generation = 1;
  c->generation = next_generation++;
  return INA_SUCCESS;
}

/* ---- decode plan ------------------------------------------------------------------------------ */
typedef struct {
  iamx_lz4_job_t *lz4;
  int64_t n_lz4, cap_lz4;
  iamx_block_job_t *blk;
  int32_t *rc4;
  int64_t n_blk, cap_blk;
  int64_t scratch;
} jobs_t;

static int push_lz4(jobs_t *j, iamx_lz4_job_t v) {
  if (j->n_lz4 == j->cap_lz4) {
    j->cap_lz4 = j->cap_lz4 ? 2 * j->cap_lz4 : 1024;
    void *p = realloc(j->lz4, (size_t)j->cap_lz4 * sizeof(v));
    if (!p) return -1;
    j->lz4 = (iamx_lz4_job_t *)p;
  }
  j->lz4[j->n_lz4++] = v;
  return 0;
}

static int push_blk(jobs_t *j, const iamx_block_job_t *v, int32_t row0, int32_t col0, int32_t bcol, int32_t flags) {
  if (j->n_blk == j->cap_blk) {
    j->cap_blk = j->cap_blk ? 2 * j->cap_blk : 1024;
    void *p = realloc(j->blk, (size_t)j->cap_blk * sizeof(*v));
    void *q = realloc(j->rc4, (size_t)j->cap_blk * 4 * sizeof(int32_t));
    if (!p || !q) return -1;
    j->blk = (iamx_block_job_t *)p;
    j->rc4 = (int32_t *)q;
  }
  j->rc4[4 * j->n_blk] = row0;
  j->rc4[4 * j->n_blk + 1] = col0;
  j->rc4[4 * j->n_blk + 2] = bcol;
  j->rc4[4 * j->n_blk + 3] = flags;
  j->blk[j->n_blk++] = *v;
  return 0;
}

/* Stream table of block `nb` of one chunk (host copy `ch`, device copy at
 * `dev`).  Fills the stream fields of *J (typesize/filter/neblock/nstreams/
 * bsize/src/kind/fill); LZ4 streams get a scratch slot and an LZ4 job. */
static ina_rc_t block_streams(const uint8_t *ch, int32_t csize, uint64_t dev, const iax_header_t *h, int nb,
                              int itemsize, jobs_t *jobs, iamx_block_job_t *J) {
  const int ts = h->typesize;
  const int leftover = (nb == h->nblocks - 1) && h->leftover;
  const int32_t bsize = leftover ? h->leftover : h->blocksize;
  if (ts != itemsize) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "chunk typesize %d != %d", ts, itemsize);
  if (bsize % ts) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "block size not a multiple of typesize");
  J->typesize = (int16_t)ts;
  J->bsize = bsize;
  if (h->special) {
    uint8_t val[8] = {0};
    if (h->special == 2) {
      if (ts == 4) { const uint32_t v = 0x7fc00000u; memcpy(val, &v, 4); }
      else if (ts == 8) { const uint64_t v = 0x7ff8000000000000ull; memcpy(val, &v, 8); }
    } else if (h->special == 3) {
      if (csize < h->header_len + ts) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "short value chunk");
      memcpy(val, ch + h->header_len, (size_t)ts);
    }
    J->nstreams = (int16_t)ts;
    J->neblock = bsize / ts;
    J->filter = ts > 1 ? IAMX_FILTER_SHUFFLE : IAMX_FILTER_NONE;
    for (int s = 0; s < ts; s++) { J->kind[s] = IAMX_SRC_CONST; J->fill[s] = val[s]; }
    return INA_SUCCESS;
  }
  if (h->memcpyed) {
    if ((int64_t)h->header_len + (int64_t)nb * h->blocksize + bsize > csize)
      return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "short memcpyed chunk");
    J->nstreams = 1;
    J->neblock = bsize;
    J->filter = IAMX_FILTER_NONE;
    J->kind[0] = IAMX_SRC_ADDR;
    J->src[0] = dev + (uint64_t)h->header_len + (uint64_t)nb * (uint64_t)h->blocksize;
    return INA_SUCCESS;
  }
  /* filters: only (bit)shuffle in any slot; TRUNC_PREC (4) is a no-op on decode */
  int filter = IAMX_FILTER_NONE;
  for (int i = 0; i < 6; i++) {
    const int f = h->filters[i];
    if (f == 0 || f == 4) continue;
    if (filter != IAMX_FILTER_NONE) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "stacked filters");
    if (f == 1 && h->filters_meta[i] == 0) filter = IAMX_FILTER_SHUFFLE;
    else if (f == 2) filter = h->version >= 3 ? IAMX_FILTER_BITSHUFFLE : IAMX_FILTER_BITSHUFFLE_V2;
    else return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "unsupported filter %d", f);
  }
  if (filter == IAMX_FILTER_BITSHUFFLE_V2 && (bsize / ts) % 8) filter = IAMX_FILTER_NONE;
  if (h->blosc2_flags & 0x1) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "dictionaries unsupported");
  const int nstreams = (!h->dont_split && !leftover) ? ts : 1;
  if (nstreams > IAMX_MAX_STREAMS) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "too many streams");
  const int32_t neblock = bsize / nstreams;
  J->nstreams = (int16_t)nstreams;
  J->neblock = neblock;
  J->filter = filter;
  const int32_t bstart = rd32(ch + h->header_len + 4 * nb);
  if (bstart < h->header_len || bstart >= csize) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "bad bstart");
  int64_t p = bstart;
  for (int s = 0; s < nstreams; s++) {
    if (p + 4 > csize) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "truncated block");
    const int32_t cs = rd32(ch + p);
    p += 4;
    if (cs == 0) {
      J->kind[s] = IAMX_SRC_CONST;
      J->fill[s] = 0;
    } else if (cs < 0) {
      if (p + 1 > csize || !(ch[p] & 0x1) || cs < -255)
        return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "bad run stream");
      J->kind[s] = IAMX_SRC_CONST;
      J->fill[s] = (uint8_t)(-cs);
      p += 1;
    } else if (cs == neblock) {
      if (p + cs > csize) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "truncated raw stream");
      J->kind[s] = IAMX_SRC_ADDR;
      J->src[s] = dev + (uint64_t)p;
      p += cs;
    } else {
      if (p + cs > csize) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "truncated LZ4 stream");
      /* codec formats 1 (LZ4, LZ4HC), 0 (BloscLZ) and 4 (Zstd) decode on the GPU */
      if (h->compformat != 1 && h->compformat != 0 && h->compformat != 4)
        return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "codec format %d has no GPU decoder", h->compformat);
      const int codec = h->compformat == 0 ? IAMX_CODEC_BLOSCLZ : h->compformat == 4 ? IAMX_CODEC_ZSTD : IAMX_CODEC_LZ4;
      if (codec == IAMX_CODEC_ZSTD) iamx_enable_zstd(1);
      iamx_lz4_job_t lj = {dev + (uint64_t)p, jobs->scratch, cs, neblock, codec, 0};
      if (push_lz4(jobs, lj)) return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
      J->kind[s] = IAMX_SRC_SCRATCH;
      J->src[s] = (uint64_t)jobs->scratch;
      jobs->scratch += (neblock + 15) & ~15;
      p += cs;
    }
  }
  return INA_SUCCESS;
}

static int by_csize_desc(const void *x, const void *y) {
  const iamx_lz4_job_t *a = (const iamx_lz4_job_t *)x, *b = (const iamx_lz4_job_t *)y;
  if (a->csize != b->csize) return a->csize < b->csize ? 1 : -1;
  return a->dst_off < b->dst_off ? -1 : a->dst_off > b->dst_off;
}

/* LZ4 jobs longest first: a launch holds more streams than the GPU has
 * wavefront slots, and starting the long ones first shortens the tail (the
 * block jobs refer to scratch offsets, not to job positions) */
void iax_sort_lz4_jobs(iamx_lz4_job_t *jobs, int64_t n) {
  if (n > 1) qsort(jobs, (size_t)n, sizeof(*jobs), by_csize_desc);
}

static ina_rc_t upload_plan(jobs_t *j, iax_plan_t *plan) {
  memset(plan, 0, sizeof(*plan));
  iax_sort_lz4_jobs(j->lz4, j->n_lz4);
  plan->n_lz4 = (int)j->n_lz4;
  plan->n_blocks = (int)j->n_blk;
  plan->scratch_bytes = j->scratch;
  if (j->n_lz4) {
    plan->d_lz4 = (iamx_lz4_job_t *)iamx_malloc((size_t)j->n_lz4 * sizeof(iamx_lz4_job_t));
    if (!plan->d_lz4) return iax_dev_error(IAMX_ENOMEM, "plan");
    int rc = iamx_h2d(plan->d_lz4, j->lz4, (size_t)j->n_lz4 * sizeof(iamx_lz4_job_t));
    if (rc) return iax_dev_error(rc, "plan upload");
  }
  if (j->n_blk) {
    plan->d_blocks = (iamx_block_job_t *)iamx_malloc((size_t)j->n_blk * sizeof(iamx_block_job_t));
    plan->d_blk_rc = (int32_t *)iamx_malloc((size_t)j->n_blk * 4 * sizeof(int32_t));
    if (!plan->d_blocks || !plan->d_blk_rc) return iax_dev_error(IAMX_ENOMEM, "plan");
    int rc = iamx_h2d(plan->d_blocks, j->blk, (size_t)j->n_blk * sizeof(iamx_block_job_t));
    if (!rc) rc = iamx_h2d(plan->d_blk_rc, j->rc4, (size_t)j->n_blk * 4 * sizeof(int32_t));
    if (rc) return iax_dev_error(rc, "plan upload");
  }
  return INA_SUCCESS;
}

static void free_jobs(jobs_t *j) {
  free(j->lz4);
  free(j->blk);
  free(j->rc4);
}

/* Block jobs of chunk nc (host bytes ch of csize, or NULL for a
 * never-written chunk; its device copy at dev) appended to `jobs`. */
static ina_rc_t chunk_jobs(const iarray_container_t *c, int64_t nc, const uint8_t *ch, int32_t csize, uint64_t dev,
                           jobs_t *jobs) {
  ina_rc_t rc = INA_SUCCESS;
  const int64_t ld = c->shape[1];
  const int64_t ci0 = nc / c->nchunks_ax[1], ci1 = nc % c->nchunks_ax[1];
  iax_header_t h;
  if (ch) {
    if (iax_parse_header(ch, csize, &h) || h.nbytes != c->chunk_nbytes || h.blocksize != c->blocksize ||
        h.nblocks != c->nblocks_ax[0] * c->nblocks_ax[1])
      return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "chunk %lld does not match the tiling", (long long)nc);
  }
  for (int64_t bi0 = 0; bi0 < c->nblocks_ax[0] && INA_SUCCEED(rc); bi0++)
    for (int64_t bi1 = 0; bi1 < c->nblocks_ax[1] && INA_SUCCEED(rc); bi1++) {
      const int64_t r_in = bi0 * c->block[0], c_in = bi1 * c->block[1];
      const int64_t row0 = ci0 * c->chunk[0] + r_in, col0 = ci1 * c->chunk[1] + c_in;
      int64_t vr = c->block[0], vc = c->block[1];
      if (c->chunk[0] - r_in < vr) vr = c->chunk[0] - r_in;
      if (c->shape[0] - row0 < vr) vr = c->shape[0] - row0;
      if (c->chunk[1] - c_in < vc) vc = c->chunk[1] - c_in;
      if (c->shape[1] - col0 < vc) vc = c->shape[1] - col0;
      if (vr <= 0 || vc <= 0) continue; /* an all-padding block holds no data */
      iamx_block_job_t J;
      memset(&J, 0, sizeof(J));
      const int nb = (int)(bi0 * c->nblocks_ax[1] + bi1);
      if (ch) {
        rc = block_streams(ch, csize, dev, &h, nb, c->itemsize, jobs, &J);
        if (INA_FAILED(rc)) break;
      } else { /* never-written chunk of an empty container reads as zeros */
        J.typesize = (int16_t)c->itemsize;
        J.bsize = c->blocksize;
        J.nstreams = 1;
        J.neblock = c->blocksize;
        J.filter = IAMX_FILTER_NONE;
        J.kind[0] = IAMX_SRC_CONST;
      }
      J.rows = (int32_t)c->block[0];
      J.cols = (int32_t)c->block[1];
      J.vrows = (int32_t)vr;
      J.vcols = (int32_t)vc;
      J.dst_off = row0 * ld + col0;
      J.dst_ld = ld;
      /* bit 0: a zero block by the reference's test (special zero chunk, or
       * every stream of the block empty), skipped by opt_gemv */
      const int32_t zflag = ch && (iax_chunk_is_zeros(ch, csize) || iax_block_is_zeros(ch, csize, nb));
      if (push_blk(jobs, &J, (int32_t)row0, (int32_t)col0, (int32_t)(ci1 * c->nblocks_ax[1] + bi1), zflag))
        rc = INA_ERROR(INA_ERR_OUT_OF_MEMORY);
    }
  return rc;
}

ina_rc_t iax_ensure_plan(iarray_container_t *c) {
  if (c->viewed) return iax_ensure_plan(c->viewed);
  if (c->plan_valid) return INA_SUCCESS;
  ina_rc_t rc = iax_ensure_device(c);
  if (INA_FAILED(rc)) return rc;
  jobs_t jobs;
  memset(&jobs, 0, sizeof(jobs));
  for (int64_t nc = 0; nc < c->nchunks && INA_SUCCEED(rc); nc++)
    rc = chunk_jobs(c, nc, c->chunks[nc], c->csizes[nc], (uint64_t)(uintptr_t)(c->d_arena + c->arena_off[nc]),
                    &jobs);
  if (INA_SUCCEED(rc)) rc = upload_plan(&jobs, &c->plan);
  c->plan.nbcols = c->nchunks_ax[1] * c->nblocks_ax[1];
  c->plan.h_lz4 = jobs.lz4;
  c->plan.h_blocks = jobs.blk;
  free(jobs.rc4);
  if (INA_FAILED(rc)) {
    free_plan(&c->plan);
    return rc;
  }
  c->plan_valid = 1;
  return INA_SUCCESS;
}

/* ---- operands streamed through the device ----------------------------------------------------- *
 * A GEMV whose matrix does not fit the device budget (IARRAY_DEVICE_BUDGET
 * bytes, else half of the device's free memory) never holds it resident: its
 * chunk rows travel in panels through two upload stages.  For panel k the
 * host copies the panel's compressed chunks into stage k%2's pinned buffer
 * (from the container's chunks, which an opened frame pages in from its file
 * only now), appends the panel's decode jobs (device addresses inside the
 * stage's device buffer), and queues the upload; the library stream waits
 * for it, decodes and multiplies the panel's blocks into the partial sums and
 * releases the stage, so the upload of panel k+1 overlaps the GEMV of panel
 * k.  The blocks, their jobs and the order of every sum are those of the
 * resident GEMV, so y is the same to the bit. */
static int64_t stream_budget(void) {
  const char *e = getenv("IARRAY_DEVICE_BUDGET");
  if (e && atoll(e) > 0) return (int64_t)atoll(e);
  size_t f = 0, t = 0;
  if (iamx_mem_info(&f, &t) != 0) return INT64_MAX;
  return (int64_t)(f / 2);
}

/* bytes the resident GEMV keeps on the device for A: the compressed arena and
 * the decoded LZ4 streams (at most the decoded chunks) */
static int64_t resident_bytes(const iarray_container_t *a) {
  int64_t cb = 0;
  for (int64_t i = 0; i < a->nchunks; i++) cb += (a->csizes[i] + 15) & ~15;
  return cb + a->nchunks * a->chunk_nbytes;
}

int iax_gemv_should_stream(iarray_container_t *a) {
  if (a->viewed || a->device_valid) return 0;
  if (INA_FAILED(iax_materialize(a))) return 0;
  return resident_bytes(a) > stream_budget();
}

static size_t al16(size_t v) { return (v + 15) & ~(size_t)15; }

ina_rc_t iax_gemv_streamed(iarray_container_t *a, int dt, const void *dx, void *dy, const uint8_t *d_xzero) {
  ina_rc_t rc = iax_materialize(a);
  if (INA_FAILED(rc)) return rc;
  const int64_t M = a->shape[0];
  const int64_t ncr = a->nchunks_ax[0], ncc = a->nchunks_ax[1];
  const int64_t budget = stream_budget();
  /* panels of whole chunk rows: compressed bytes and decoded scratch each at
   * most a quarter of the budget (one chunk row at least) */
  int64_t *pstart = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ncr + 1));
  if (!pstart) return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  int np = 0;
  size_t max_stage = 0;
  int64_t max_scratch = 0;
  for (int64_t r = 0; r < ncr;) {
    pstart[np++] = r;
    int64_t cb = 0, db = 0;
    do {
      int64_t rcb = 0;
      for (int64_t j = 0; j < ncc; j++) rcb += al16((size_t)a->csizes[r * ncc + j]);
      const int64_t rdb = ncc * a->chunk_nbytes;
      if (cb > 0 && (cb + rcb > budget / 4 || db + rdb > budget / 4)) break;
      cb += rcb;
      db += rdb;
      r++;
    } while (r < ncr);
    if ((size_t)cb > max_stage) max_stage = (size_t)cb;
    if (db > max_scratch) max_scratch = db;
  }
  pstart[np] = ncr;
  iax_last_gemv_panels = np;
  const int64_t nbc = ncc * a->nblocks_ax[1];
  /* the stage's device buffer: compressed chunks, then the job tables (a
   * panel's blocks: at most its chunks' blocks; LZ4 jobs: at most 8 a block) */
  int64_t max_blocks = 0;
  for (int k = 0; k < np; k++) {
    const int64_t nb = (pstart[k + 1] - pstart[k]) * ncc * a->nblocks_ax[0] * a->nblocks_ax[1];
    if (nb > max_blocks) max_blocks = nb;
  }
  const size_t tables = al16((size_t)max_blocks * IAMX_MAX_STREAMS * sizeof(iamx_lz4_job_t)) +
                        al16((size_t)max_blocks * sizeof(iamx_block_job_t)) + al16((size_t)max_blocks * 16);
  const size_t stage_bytes = al16(max_stage) + tables;
  uint8_t *d_stage[2] = {(uint8_t *)iamx_workspace(0, 2 * stage_bytes + 16), NULL};
  void *scratch = iamx_workspace(7, (size_t)max_scratch + 16);
  iax_operand_forget(0);
  iax_operand_forget(7);
  iax_operand_forget(6);
  void *partial = iamx_workspace(6, (size_t)(nbc * M * a->itemsize) + 16);
  if (!d_stage[0] || !scratch || !partial) {
    free(pstart);
    return iax_dev_error(IAMX_ENOMEM, "streamed gemv buffers");
  }
  d_stage[1] = d_stage[0] + stage_bytes;
  int e = iamx_memset(partial, 0, (size_t)(nbc * M * a->itemsize));
  jobs_t jobs;
  memset(&jobs, 0, sizeof(jobs));
  int launched = 0;
  for (int k = 0; k < np && !e && INA_SUCCEED(rc); k++) {
    const int s = k & 1;
    uint8_t *h = (uint8_t *)iamx_upload_host(s, stage_bytes);
    if (!h) {
      e = IAMX_ENOMEM;
      break;
    }
    /* the panel's chunks, 16-byte aligned, and their jobs */
    jobs.n_lz4 = jobs.n_blk = 0;
    jobs.scratch = 0;
    size_t off = 0;
    for (int64_t nc = pstart[k] * ncc; nc < pstart[k + 1] * ncc && INA_SUCCEED(rc); nc++) {
      const uint8_t *ch = a->chunks[nc];
      if (ch) memcpy(h + off, ch, (size_t)a->csizes[nc]);
      rc = chunk_jobs(a, nc, ch, a->csizes[nc], (uint64_t)(uintptr_t)(d_stage[s] + off), &jobs);
      off += ch ? al16((size_t)a->csizes[nc]) : 0;
    }
    if (INA_FAILED(rc)) break;
    if ((size_t)jobs.n_lz4 > (size_t)max_blocks * IAMX_MAX_STREAMS || jobs.n_blk > max_blocks ||
        jobs.scratch > max_scratch) {
      rc = iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "streamed panel %d exceeds its buffers", k);
      break;
    }
    iax_sort_lz4_jobs(jobs.lz4, jobs.n_lz4);
    const size_t o_lz4 = al16(max_stage);
    const size_t o_blk = o_lz4 + al16((size_t)max_blocks * IAMX_MAX_STREAMS * sizeof(iamx_lz4_job_t));
    const size_t o_rc = o_blk + al16((size_t)max_blocks * sizeof(iamx_block_job_t));
    memcpy(h + o_lz4, jobs.lz4, (size_t)jobs.n_lz4 * sizeof(iamx_lz4_job_t));
    memcpy(h + o_blk, jobs.blk, (size_t)jobs.n_blk * sizeof(iamx_block_job_t));
    memcpy(h + o_rc, jobs.rc4, (size_t)jobs.n_blk * 16);
    e = iamx_upload(s, d_stage[s], o_rc + (size_t)jobs.n_blk * 16);
    if (e || jobs.n_blk == 0) {
      if (!e) e = iamx_upload_release(s);
      continue;
    }
    const iamx_lz4_job_t *dl = (const iamx_lz4_job_t *)(d_stage[s] + o_lz4);
    const iamx_block_job_t *db = (const iamx_block_job_t *)(d_stage[s] + o_blk);
    const int32_t *drc = (const int32_t *)(d_stage[s] + o_rc);
    e = d_xzero ? iamx_opt_gemv_fused(dt, dl, (int)jobs.n_lz4, db, drc, (int)jobs.n_blk, (int)nbc, (int)a->block[0],
                                      (int)a->block[1], dx, dy, M, scratch, partial, d_xzero)
                : iamx_gemv_fused(dt, dl, (int)jobs.n_lz4, db, drc, (int)jobs.n_blk, (int)nbc, (int)a->block[0],
                                  (int)a->block[1], dx, dy, M, scratch, partial);
    launched = 1;
    if (!e) e = iamx_upload_release(s);
  }
  if (!e && INA_SUCCEED(rc) && !launched) e = iamx_memset(dy, 0, (size_t)(M * a->itemsize));
  free_jobs(&jobs);
  free(pstart);
  if (INA_FAILED(rc)) {
    (void)iamx_sync();
    return rc;
  }
  if (e) return iax_dev_error(e, "streamed gemv");
  return INA_SUCCESS;
}

ina_rc_t iax_decode_dense(iarray_container_t *c, void *d_dst, int scratch_slot) {
  ina_rc_t rc = iax_ensure_plan(c);
  if (INA_FAILED(rc)) return rc;
  void *scratch = iamx_workspace(scratch_slot, (size_t)c->plan.scratch_bytes + 16);
  if (!scratch) return iax_dev_error(IAMX_ENOMEM, "decode scratch");
  int e = iamx_decode(c->plan.d_lz4, c->plan.n_lz4, c->plan.d_blocks, c->plan.n_blocks, scratch, d_dst);
  if (e) return iax_dev_error(e, "decode");
  return INA_SUCCESS;
}

iax_operand_key_t iax_operand_key(const iarray_container_t *c) {
  const iarray_container_t *base = c->viewed ? c->viewed : c;
  iax_operand_key_t k;
  memset(&k, 0, sizeof(k));
  k.gen = base->generation;
  k.transposed = c->transposed ? 1 : 0;
  if (c->viewed) {
    k.voff[0] = c->voff[0];
    k.voff[1] = c->voff[1];
  }
  k.shape[0] = c->dtshape.shape[0];
  k.shape[1] = c->dtshape.ndim == 2 ? c->dtshape.shape[1] : 1;
  return k;
}

/* Which container's dense image each workspace slot holds (its device
 * generation, unique across containers, and the buffer address): a plain
 * container asked for again in the same slot, unchanged, is not decoded again
 * (the GEMV's x every call).  Every other writer of a slot forgets it. */
#define IAX_IMG_SLOTS 16
static struct {
  uint64_t gen;
  const void *ptr;
} g_img[IAX_IMG_SLOTS];

void iax_operand_forget(int slot) {
  if (slot < 0) {
    memset(g_img, 0, sizeof(g_img));
  } else if (slot < IAX_IMG_SLOTS) {
    g_img[slot].gen = 0;
    g_img[slot].ptr = NULL;
  }
}

/* Decode the base of transposed view v into the view's row-major image
 * (base cols x base rows): the base's block jobs with their tiles written
 * transposed, so no separate transpose pass reads and writes the image. */
static ina_rc_t decode_transposed(iarray_container_t *v, iarray_container_t *base, void *dst) {
  ina_rc_t rc = iax_ensure_plan(base);
  if (INA_FAILED(rc)) return rc;
  const iax_plan_t *p = &base->plan;
  if (!v->d_tblocks || v->t_gen != base->generation) {
    iamx_free(v->d_tblocks);
    v->d_tblocks = NULL;
    const int n = p->n_blocks;
    iamx_block_job_t *h = (iamx_block_job_t *)malloc(sizeof(*h) * (size_t)(n > 0 ? n : 1));
    if (!h) return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
    const int64_t ld = base->shape[1], tld = base->shape[0];
    for (int i = 0; i < n; i++) {
      h[i] = p->h_blocks[i];
      const int64_t row0 = h[i].dst_off / ld, col0 = h[i].dst_off % ld;
      h[i].dst_off = col0 * tld + row0;
      h[i].dst_ld = tld;
      h[i].filter |= IAMX_JOB_TRANSPOSE;
    }
    v->d_tblocks = (iamx_block_job_t *)iamx_malloc(sizeof(*h) * (size_t)(n > 0 ? n : 1));
    int e = v->d_tblocks ? iamx_h2d(v->d_tblocks, h, sizeof(*h) * (size_t)n) : IAMX_ENOMEM;
    free(h);
    if (e) {
      iamx_free(v->d_tblocks);
      v->d_tblocks = NULL;
      return iax_dev_error(e, "transposed block jobs");
    }
    v->t_gen = base->generation;
  }
  void *scratch = iamx_workspace(0, (size_t)p->scratch_bytes + 16);
  if (!scratch) return iax_dev_error(IAMX_ENOMEM, "decode scratch");
  int e = iamx_decode(p->d_lz4, p->n_lz4, v->d_tblocks, p->n_blocks, scratch, dst);
  if (e) return iax_dev_error(e, "decode");
  return INA_SUCCESS;
}

ina_rc_t iax_operand(iarray_container_t *c, int slot, int tmp_slot, void **d_out) {
  iarray_container_t *base = c->viewed ? c->viewed : c;
  const size_t bytes = (size_t)(base->shape[0] * base->shape[1] * base->itemsize);
  void *dst = iamx_workspace(slot, bytes + 16);
  if (!dst) return iax_dev_error(iamx_init() ? iamx_init() : IAMX_ENOMEM, "operand buffer");
  if (!c->viewed) {
    if (slot >= 0 && slot < IAX_IMG_SLOTS && c->device_valid && g_img[slot].gen == c->generation &&
        g_img[slot].ptr == dst) {
      *d_out = dst;
      return INA_SUCCESS;
    }
    iax_operand_forget(slot);
    ina_rc_t rc = iax_decode_dense(base, dst, 0);
    if (INA_FAILED(rc)) return rc;
    if (slot >= 0 && slot < IAX_IMG_SLOTS && c->device_valid) {
      g_img[slot].gen = c->generation;
      g_img[slot].ptr = dst;
    }
    *d_out = dst;
    return INA_SUCCESS;
  }
  iax_operand_forget(slot);
  iax_operand_forget(tmp_slot);
  if (!c->transposed) { /* a slice: its rectangle out of the base's image */
    void *img = iamx_workspace(tmp_slot, bytes + 16);
    if (!img) return iax_dev_error(IAMX_ENOMEM, "slice buffer");
    ina_rc_t rc = iax_decode_dense(base, img, 0);
    if (INA_FAILED(rc)) return rc;
    const int64_t rows = c->dtshape.ndim == 2 ? c->dtshape.shape[0] : 1;
    const int64_t cols = c->dtshape.ndim == 2 ? c->dtshape.shape[1] : c->dtshape.shape[0];
    const size_t is = (size_t)base->itemsize;
    if (c->voff[0] < 0 || c->voff[1] < 0 || c->voff[0] + rows > base->shape[0] || c->voff[1] + cols > base->shape[1])
      return iax_set_error(INA_ERROR(IARRAY_ERR_ASSERTION_FAILED), "slice outside its container");
    const uint8_t *src = (const uint8_t *)img + ((size_t)c->voff[0] * (size_t)base->shape[1] + (size_t)c->voff[1]) * is;
    int e = iamx_copy2d(dst, (size_t)cols * is, src, (size_t)base->shape[1] * is, (size_t)cols * is, (size_t)rows);
    if (e) return iax_dev_error(e, "slice copy");
    *d_out = dst;
    return INA_SUCCESS;
  }
  /* the rectangle of base rows x cols the view transposes */
  const int64_t rows = c->dtshape.shape[1], cols = c->dtshape.shape[0];
  if (rows == base->shape[0] && cols == base->shape[1]) {
    /* the whole base: its blocks decoded straight into the transposed image */
    ina_rc_t rc = decode_transposed(c, base, dst);
    if (INA_FAILED(rc)) return rc;
    *d_out = dst;
    return INA_SUCCESS;
  }
  void *tmp = iamx_workspace(tmp_slot, bytes + 16);
  if (!tmp) return iax_dev_error(IAMX_ENOMEM, "transpose buffer");
  ina_rc_t rc = iax_decode_dense(base, tmp, 0);
  if (INA_FAILED(rc)) return rc;
  /* a slice of a transposed view: cut the rectangle into dst, transpose it back into tmp */
  const size_t is = (size_t)base->itemsize;
  if (c->voff[0] < 0 || c->voff[1] < 0 || c->voff[0] + rows > base->shape[0] || c->voff[1] + cols > base->shape[1])
    return iax_set_error(INA_ERROR(IARRAY_ERR_ASSERTION_FAILED), "slice outside its container");
  const uint8_t *src = (const uint8_t *)tmp + ((size_t)c->voff[0] * (size_t)base->shape[1] + (size_t)c->voff[1]) * is;
  int e = iamx_copy2d(dst, (size_t)cols * is, src, (size_t)base->shape[1] * is, (size_t)cols * is, (size_t)rows);
  if (!e) e = iamx_transpose(base->itemsize, rows, cols, dst, tmp);
  if (e) return iax_dev_error(e, "transposed slice");
  *d_out = tmp;
  return INA_SUCCESS;
}

/* The device status word of the decodes queued since the last check.  A
 * nonzero word means some image written since then may hold garbage, so every
 * remembered operand image is forgotten: the next call decodes again and
 * reports the same error. */
ina_rc_t iax_check_decode_status(void) {
  int st = iamx_decode_status();
  if (st) iax_operand_forget(-1);
  if (st < 0) return iax_dev_error(st, "decode status");
  if (st) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "malformed compressed stream (device flags 0x%x)", st);
  return INA_SUCCESS;
}

static ina_rc_t check_decode_status(void) { return iax_check_decode_status(); }

/* ---- write stage ------------------------------------------------------------------------------- */
typedef struct {
  iarray_container_t *c;
  uint8_t *packed;
  int filtered;
  int err;
} wctx_t;

static void write_chunk_task(void *vctx, int64_t nc, int worker) {
  (void)worker;
  wctx_t *x = (wctx_t *)vctx;
  iarray_container_t *c = x->c;
  free(c->chunks[nc]);
  c->chunks[nc] = NULL;
  if (iax_encode_chunk(x->packed + nc * c->chunk_nbytes, (int32_t)c->chunk_nbytes, c->blocksize, c->itemsize,
                       &c->cfg, x->filtered, &c->chunks[nc], &c->csizes[nc]) != 0)
    x->err = 1;
}

typedef struct {
  iax_task_fn fn;
  void *ctx;
  const int64_t *first; /* first task of each piece, np + 1 entries */
  int np;
  atomic_int landed[IAMX_COPY_EVENTS];
  atomic_int err;
} piped_t;

/* Task wrapper: wait (once per piece) until the piece holding the task's
 * bytes has landed, then run the task.  Tasks are handed out in order, so
 * workers block only on the piece being copied right now. */
static void piped_task(void *v, int64_t i, int worker) {
  piped_t *p = (piped_t *)v;
  int lo = 0, hi = p->np - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (p->first[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  if (!atomic_load(&p->landed[lo])) {
    if (iamx_event_sync(lo) != 0) {
      atomic_store(&p->err, 1);
      return;
    }
    atomic_store(&p->landed[lo], 1);
  }
  p->fn(p->ctx, i, worker);
}

/* Core of the pipelined downloads: the tasks are grouped into panels (panel k
 * holds tasks [panel_first[k], panel_first[k+1])); before the first copy of a
 * panel, produce(pctx, k, t0, t1) enqueues the device work that fills the
 * panel's bytes.  Each panel is copied in pieces of about total/32 bytes (at
 * least 4 MB), a piece never spanning two panels. */
static ina_rc_t download_panels(const void *d_src, size_t total, const int64_t *task_end, int64_t ntasks,
                                const int64_t *panel_first, int npanels, iax_panel_fn produce, void *pctx,
                                int nthreads, iax_task_fn fn, void *ctx, uint8_t **host) {
  uint8_t *h = (uint8_t *)iamx_host_buffer(total + 16);
  if (!h) return iax_dev_error(IAMX_ENOMEM, "pinned staging buffer");
  *host = h;
  const uint8_t *d = (const uint8_t *)d_src;
  size_t target = total / 32;
  if (target < ((size_t)4 << 20)) target = (size_t)4 << 20;
  int64_t first[IAMX_COPY_EVENTS + 1];
  int np = 0, k = 0;
  int64_t t = 0;
  size_t start = 0;
  while (t < ntasks) {
    /* enqueue the device work of every panel that starts at or before task t */
    while (k < npanels && panel_first[k] <= t) {
      if (produce) {
        const int e = produce(pctx, k, panel_first[k], panel_first[k + 1]);
        if (e) return iax_dev_error(e, "panel");
      }
      k++;
    }
    const int64_t pend = k < npanels ? panel_first[k] : ntasks; /* end of the current panel */
    int64_t t1 = t;
    while (t1 < pend && (size_t)task_end[t1] - start < target) t1++;
    if (t1 < pend) t1++; /* the task that reaches the target closes the piece */
    if (np == IAMX_COPY_EVENTS - 1) { /* last event: the rest in one piece */
      for (; k < npanels; k++)
        if (produce) {
          const int e = produce(pctx, k, panel_first[k], panel_first[k + 1]);
          if (e) return iax_dev_error(e, "panel");
        }
      t1 = ntasks;
    }
    const size_t end = (size_t)task_end[t1 - 1];
    const int e = iamx_d2h_async(h + start, d + start, end - start, np);
    if (e) return iax_dev_error(e, "download");
    first[np++] = t;
    t = t1;
    start = end;
  }
  first[np] = ntasks;
  piped_t p;
  p.fn = fn;
  p.ctx = ctx;
  p.first = first;
  p.np = np;
  for (int i = 0; i < IAMX_COPY_EVENTS; i++) atomic_init(&p.landed[i], 0);
  atomic_init(&p.err, 0);
  iax_parallel_for(nthreads, ntasks, piped_task, &p);
  if (atomic_load(&p.err)) {
    for (int i = 0; i < np; i++) (void)iamx_event_sync(i); /* no copy may still be writing the buffer */
    return iax_dev_error(IAMX_EINVAL, "download");
  }
  return INA_SUCCESS;
}

ina_rc_t iax_download_pipelined(const void *d_src, size_t total, const int64_t *task_end, int64_t ntasks,
                                int nthreads, iax_task_fn fn, void *ctx, uint8_t **host) {
  const int64_t pf[2] = {0, ntasks};
  return download_panels(d_src, total, task_end, ntasks, pf, 1, NULL, NULL, nthreads, fn, ctx, host);
}

static void tiling_of(const iarray_container_t *c, iamx_tiling_t *t) {
  memset(t, 0, sizeof(*t));
  for (int i = 0; i < 2; i++) {
    t->shape[i] = c->shape[i];
    t->chunk[i] = c->chunk[i];
    t->block[i] = c->block[i];
    t->nchunks[i] = c->nchunks_ax[i];
    t->nblocks[i] = c->nblocks_ax[i];
  }
  t->typesize = c->itemsize;
  t->filter = c->cfg.compression_level > 0 ? iax_filter_code(&c->cfg) : IAMX_FILTER_NONE;
}

typedef struct {
  iarray_container_t *c;
  const iamx_tiling_t *t;
  const void *d_dense;
  void *d_packed;
  iax_rows_fn rows;
  void *rctx;
} panel_ctx_t;

/* enqueue the rows of C behind chunks [t0, t1) (whole chunk rows), then their packing */
static int produce_panel(void *v, int k, int64_t t0, int64_t t1) {
  (void)k;
  panel_ctx_t *x = (panel_ctx_t *)v;
  const iarray_container_t *c = x->c;
  if (x->rows) {
    const int64_t r0 = (t0 / c->nchunks_ax[1]) * c->chunk[0];
    int64_t r1 = (t1 / c->nchunks_ax[1]) * c->chunk[0];
    if (r1 > c->shape[0]) r1 = c->shape[0];
    if (r1 > r0) {
      const int e = x->rows(x->rctx, r0, r1 - r0);
      if (e) return e;
    }
  }
  return iamx_pack_blocks_range(x->t, x->d_dense, c->shape[1], x->d_packed, t0, t1);
}

/* ---- C write stage shared between the host and the GPU encoder ---------------------------------- *
 * The last `share` of the panels are LZ4-encoded and assembled into chunks on
 * the GPU (on the side stream, overlapping the next panel's GEMM), the others
 * on the host as before.  Every chunk lands in the pinned staging buffer at
 * chunk * cap: packed bytes for a host panel, the finished chunk (its size in
 * its header) for a GPU panel; all of it is queued up front, and the host
 * threads take each chunk as soon as its panel has landed. */
typedef struct {
  iarray_container_t *c;
  const uint8_t *h;
  int64_t cap;
  const int64_t *first; /* first chunk of each panel, np + 1 entries */
  const uint8_t *on_gpu;
  int np, filtered;
  atomic_int landed[IAMX_COPY_EVENTS];
  atomic_int err;
} hyb_t;

static uint32_t rd32le(const uint8_t *p) {
  uint32_t v;
  memcpy(&v, p, 4);
  return v;
}

static void hybrid_task(void *v, int64_t nc, int worker) {
  (void)worker;
  hyb_t *x = (hyb_t *)v;
  int lo = 0, hi = x->np - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (x->first[mid] <= nc) lo = mid;
    else hi = mid - 1;
  }
  if (!atomic_load(&x->landed[lo])) {
    if (iamx_event_sync(lo) != 0) {
      atomic_store(&x->err, 1);
      return;
    }
    atomic_store(&x->landed[lo], 1);
  }
  iarray_container_t *c = x->c;
  const uint8_t *src = x->h + nc * x->cap;
  free(c->chunks[nc]);
  c->chunks[nc] = NULL;
  if (x->on_gpu[lo]) {
    const int32_t sz = (int32_t)rd32le(src + 12); /* cbytes of the assembled chunk */
    uint8_t *p = sz > 32 && sz <= x->cap ? (uint8_t *)malloc((size_t)sz) : NULL;
    if (!p) {
      atomic_store(&x->err, 1);
      return;
    }
    memcpy(p, src, (size_t)sz);
    c->chunks[nc] = p;
    c->csizes[nc] = sz;
  } else if (iax_encode_chunk(src, (int32_t)c->chunk_nbytes, c->blocksize, c->itemsize, &c->cfg, x->filtered,
                              &c->chunks[nc], &c->csizes[nc]) != 0) {
    atomic_store(&x->err, 1);
  }
}

static ina_rc_t write_hybrid(iarray_container_t *c, const iamx_tiling_t *t, const void *d_dense, void *d_packed,
                             const int64_t *pf, int npanels, iax_rows_fn rows, void *rctx, int ngpu) {
  const int split = iax_split(&c->cfg, c->itemsize, c->blocksize);
  const int64_t cap = iamx_chunk_capacity(t, split);
  if (cap < c->chunk_nbytes) return iax_dev_error(IAMX_EINVAL, "chunk capacity");
  if (ngpu > npanels) ngpu = npanels;
  if (ngpu < 0) ngpu = 0;
  uint8_t on_gpu[IAMX_COPY_EVENTS];
  for (int k = 0; k < npanels; k++) on_gpu[k] = k >= npanels - ngpu; /* the last panels: a shorter tail */
  void *d_scratch = NULL;
  int32_t *d_meta = NULL;
  uint8_t *d_out = NULL;
  int64_t *d_sizes = NULL;
  if (ngpu > 0) {
    const int64_t nstreams = split ? c->itemsize : 1;
    const int64_t nst = c->nchunks * c->nblocks_ax[0] * c->nblocks_ax[1] * nstreams;
    const int64_t scratch = iamx_encode_scratch_bytes(t, split);
    if (scratch <= 0) return iax_dev_error(IAMX_EINVAL, "encode scratch");
    d_scratch = iamx_workspace(12, (size_t)scratch);
    d_meta = (int32_t *)iamx_workspace(13, (size_t)(nst * 8 + c->nchunks * 8 + 16));
    d_out = (uint8_t *)iamx_workspace(14, (size_t)(c->nchunks * cap));
    if (!d_scratch || !d_meta || !d_out) return iax_dev_error(IAMX_ENOMEM, "encode buffers");
    d_sizes = (int64_t *)((uint8_t *)d_meta + (nst * 8 + 7) / 8 * 8);
  }
  uint8_t *h = (uint8_t *)iamx_host_buffer((size_t)(c->nchunks * cap));
  if (!h) return iax_dev_error(IAMX_ENOMEM, "pinned staging buffer");
  panel_ctx_t px = {c, t, d_dense, d_packed, rows, rctx};
  for (int k = 0; k < npanels; k++) {
    int e = produce_panel(&px, k, pf[k], pf[k + 1]);
    if (e) return iax_dev_error(e, "panel");
    if (on_gpu[k]) {
      if ((e = iamx_side_wait()) != 0 || (e = iamx_use_stream(1)) != 0) return iax_dev_error(e, "encode");
      e = iamx_encode_chunks(t, split, d_packed, pf[k], pf[k + 1], d_scratch, d_meta, d_out, d_sizes);
      for (int64_t nc = pf[k]; nc < pf[k + 1] && !e; nc++)
        e = nc == pf[k] ? iamx_d2h_async(h + nc * cap, d_out + nc * cap, (size_t)cap, k)
                        : iamx_d2h_chain(h + nc * cap, d_out + nc * cap, (size_t)cap, k);
      iamx_use_stream(0);
      if (e) return iax_dev_error(e, "encode");
    } else {
      const uint8_t *pk = (const uint8_t *)d_packed;
      for (int64_t nc = pf[k]; nc < pf[k + 1] && !e; nc++)
        e = nc == pf[k] ? iamx_d2h_async(h + nc * cap, pk + nc * c->chunk_nbytes, (size_t)c->chunk_nbytes, k)
                        : iamx_d2h_chain(h + nc * cap, pk + nc * c->chunk_nbytes, (size_t)c->chunk_nbytes, k);
      if (e) return iax_dev_error(e, "download");
    }
  }
  hyb_t x;
  x.c = c;
  x.h = h;
  x.cap = cap;
  x.first = pf;
  x.on_gpu = on_gpu;
  x.np = npanels;
  x.filtered = c->cfg.compression_level > 0;
  for (int k = 0; k < IAMX_COPY_EVENTS; k++) atomic_init(&x.landed[k], 0);
  atomic_init(&x.err, 0);
  iax_parallel_for(iax_nthreads(&c->cfg), c->nchunks, hybrid_task, &x);
  if (atomic_load(&x.err)) {
    for (int k = 0; k < npanels; k++) (void)iamx_event_sync(k);
    return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "chunk write failed");
  }
  for (int k = 0; k < npanels; k++)
    for (int64_t nc = pf[k]; nc < pf[k + 1] && on_gpu[k]; nc++) {
      iax_last_gpu_enc_in += c->chunk_nbytes;
      iax_last_gpu_enc_out += c->csizes[nc];
    }
  return INA_SUCCESS;
}

/* Share of the C panels encoded on the GPU.  IARRAY_GPU_ENCODE (0..1) sets it;
 * unset, it is 1/4 for outputs of at least 128 MiB in 4 or more panels and 0
 * otherwise.  The GPU encoder covers LZ4 chunks of whole blocks with up to 8
 * streams.  Measured with 16 host threads: the GPU encode of a 1024-row panel
 * of an 8192^2 float32 product takes 0.65 ms against about 2 ms of host
 * thread time, but it adds to the GPU's own work.  A quarter of the panels
 * on the GPU took 8192^2 float32 products from 20.5 to 18.7 ms (split-bf16)
 * and 6144^2 float64 ones from 24.4 to 21.7 ms; half or more gave less, and a
 * 2000^2 float64 product (32 MB, 2 panels) lost 10%. */
static double gpu_encode_share(const iarray_container_t *c, int npanels) {
  const int ok = c->cfg.compression_level > 0 &&
                 (c->cfg.compression_codec == IARRAY_COMPRESSION_LZ4 ||
                  c->cfg.compression_codec == IARRAY_COMPRESSION_LZ4HC) &&
                 c->itemsize <= IAMX_MAX_STREAMS && c->nchunks > 0 && c->chunk_nbytes % c->blocksize == 0 &&
                 c->chunk_nbytes <= INT32_MAX / 2;
  if (!ok) return 0.0;
  const char *env = getenv("IARRAY_GPU_ENCODE");
  if (!env) return (c->nchunks * c->chunk_nbytes >= ((int64_t)128 << 20) && npanels >= 4) ? 0.25 : 0.0;
  double v = atof(env);
  if (!(v > 0.0)) return 0.0;
  return v > 1.0 ? 1.0 : v;
}

/* ---- adaptive GPU share of the C write ------------------------------------------------------- *
 * How many panels the GPU should encode depends on how long the GPU's own
 * work takes against the host threads' encode, which differs by product form
 * (a split-bf16 product leaves the GPU idle sooner than a plain one) and by
 * machine.  Without IARRAY_GPU_ENCODE, each output class (element size, bytes,
 * panel count, and the product form the caller names in iax_write_form)
 * keeps a hill climb over the number of GPU panels: a call
 * runs with `cur` and times the whole panel stage; the next call tries one
 * panel more (or fewer) than the best count so far, and after a move in each
 * direction has failed to beat the best, the class stays at the best count.
 * The encoded bytes of a chunk differ between the two encoders (both decode
 * to the same C). */
#define IAX_SHARE_CLASSES 16
int iax_write_form = 0;

typedef struct {
  int64_t key[4];
  int used;
  int cur, best, dir, fails;
  double best_s;
} share_state_t;
static share_state_t g_share[IAX_SHARE_CLASSES];
static pthread_mutex_t g_share_mu = PTHREAD_MUTEX_INITIALIZER;

static share_state_t *share_class(const iarray_container_t *c, int npanels, int ngpu0) {
  const int64_t key[4] = {c->itemsize, c->nchunks * c->chunk_nbytes, npanels, iax_write_form};
  share_state_t *free_slot = NULL;
  for (int i = 0; i < IAX_SHARE_CLASSES; i++) {
    share_state_t *s = &g_share[i];
    if (s->used && !memcmp(s->key, key, sizeof key)) return s;
    if (!s->used && !free_slot) free_slot = s;
  }
  if (!free_slot) free_slot = &g_share[(c->nchunks + npanels) % IAX_SHARE_CLASSES]; /* evict one */
  memset(free_slot, 0, sizeof *free_slot);
  memcpy(free_slot->key, key, sizeof key);
  free_slot->used = 1;
  free_slot->cur = free_slot->best = ngpu0;
  free_slot->dir = 1;
  free_slot->best_s = -1.0;
  return free_slot;
}

/* GPU panels for this call (the class's current probe) */
static int adaptive_gpu_panels(const iarray_container_t *c, int npanels, double share0) {
  int n0 = (int)(share0 * npanels + 0.5);
  pthread_mutex_lock(&g_share_mu);
  const int n = share_class(c, npanels, n0)->cur;
  pthread_mutex_unlock(&g_share_mu);
  return n;
}

/* record the panel stage's time `s` of a call that ran `ngpu` GPU panels and
 * choose the next probe */
static void adaptive_record(const iarray_container_t *c, int npanels, int ngpu, double s) {
  pthread_mutex_lock(&g_share_mu);
  share_state_t *st = share_class(c, npanels, ngpu);
  if (st->cur == ngpu) {
    if (st->best_s < 0 || s < st->best_s) {
      if (st->best_s >= 0 && ngpu != st->best) st->fails = 0; /* a move that paid: keep going */
      st->best_s = s;
      st->best = ngpu;
    } else if (ngpu != st->best) {
      st->fails++;
      st->dir = -st->dir;
    }
    int next = st->best;
    if (st->fails < 2) {
      next = st->best + st->dir;
      if (next < 0 || next > npanels) { /* the edge: turn round */
        st->fails++;
        st->dir = -st->dir;
        next = st->best + st->dir;
        if (next < 0 || next > npanels || st->fails >= 2) next = st->best;
      }
    }
    st->cur = next;
  }
  pthread_mutex_unlock(&g_share_mu);
}

/* Outputs up to this size are packed straight into mapped host memory (the
 * pack kernels' byte stores then cross the host link); IARRAY_DIRECT_WRITE=0
 * sends every output through the pipelined downloads. */
#define IAX_DIRECT_WRITE_MAX ((size_t)2 << 20)
static int direct_write_on(void) {
  const char *e = getenv("IARRAY_DIRECT_WRITE");
  return !(e && e[0] == '0');
}

ina_rc_t iax_write_from_dense_rows(iarray_container_t *c, const void *d_dense, int64_t panel_rows,
                                   iax_rows_fn rows, void *rctx) {
  iamx_tiling_t t;
  tiling_of(c, &t);
  const int filtered = c->cfg.compression_level > 0;
  const size_t total = (size_t)(c->nchunks * c->chunk_nbytes);
  void *d_packed = iamx_workspace(5, total);
  if (!d_packed) return iax_dev_error(IAMX_ENOMEM, "pack buffer");
  /* panels of whole chunk rows, at most 16 of them */
  const int64_t ncr = c->nchunks_ax[0];
  int64_t cr = panel_rows > 0 ? (panel_rows + c->chunk[0] - 1) / c->chunk[0] : ncr;
  if (cr < 1) cr = 1;
  if ((ncr + cr - 1) / cr > 16) cr = (ncr + 15) / 16;
  const int npanels = (int)((ncr + cr - 1) / cr);
  int64_t *ends = (int64_t *)malloc(sizeof(int64_t) * (size_t)c->nchunks);
  int64_t *pf = (int64_t *)malloc(sizeof(int64_t) * (size_t)(npanels + 1));
  if (!ends || !pf) {
    free(ends);
    free(pf);
    return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  }
  for (int64_t i = 0; i < c->nchunks; i++) ends[i] = (i + 1) * c->chunk_nbytes;
  for (int k = 0; k <= npanels; k++) {
    int64_t r = (int64_t)k * cr;
    if (r > ncr) r = ncr;
    pf[k] = r * c->nchunks_ax[1];
  }
  const double share = gpu_encode_share(c, npanels);
  iax_last_gpu_enc_in = iax_last_gpu_enc_out = 0;
  if (share > 0.0) {
    const int adaptive = getenv("IARRAY_GPU_ENCODE") == NULL;
    const int ngpu = adaptive ? adaptive_gpu_panels(c, npanels, share) : (int)(share * npanels + 0.5);
    const double t0 = iax_now();
    ina_rc_t rc = write_hybrid(c, &t, d_dense, d_packed, pf, npanels, rows, rctx, ngpu);
    iax_last_encode_s = iax_now() - t0;
    if (adaptive && !INA_FAILED(rc)) adaptive_record(c, npanels, ngpu, iax_last_encode_s);
    free(ends);
    free(pf);
    if (INA_FAILED(rc)) return rc;
    drop_device(c);
    return INA_SUCCESS;
  }
  if (total <= IAX_DIRECT_WRITE_MAX && direct_write_on()) {
    /* small C: the pack kernels write straight into mapped host memory, so
     * the host encodes after one stream synchronisation and no copy */
    void *hdev = NULL;
    uint8_t *h = (uint8_t *)iamx_host_direct(total + 16, &hdev);
    if (h) {
      const double t0 = iax_now();
      panel_ctx_t pd = {c, &t, d_dense, hdev, rows, rctx};
      int e = 0;
      for (int k = 0; k < npanels && !e; k++) e = produce_panel(&pd, k, pf[k], pf[k + 1]);
      if (!e) e = iamx_status_snapshot(); /* read by the caller's status check without another launch */
      if (!e) e = iamx_sync();
      free(ends);
      free(pf);
      if (e) return iax_dev_error(e, "panel");
      wctx_t xd = {c, h, filtered, 0};
      iax_parallel_for(iax_nthreads(&c->cfg), c->nchunks, write_chunk_task, &xd);
      iax_last_encode_s = iax_now() - t0;
      drop_device(c);
      if (xd.err) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "chunk encode failed");
      return INA_SUCCESS;
    }
  }
  panel_ctx_t px = {c, &t, d_dense, d_packed, rows, rctx};
  const double t0 = iax_now();
  wctx_t x = {c, NULL, filtered, 0};
  ina_rc_t rc = download_panels(d_packed, total, ends, c->nchunks, pf, npanels, produce_panel, &px,
                                iax_nthreads(&c->cfg), write_chunk_task, &x, &x.packed);
  iax_last_encode_s = iax_now() - t0;
  free(ends);
  free(pf);
  if (INA_FAILED(rc)) return rc;
  drop_device(c);
  if (x.err) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "chunk encode failed");
  return INA_SUCCESS;
}

ina_rc_t iax_write_from_dense(iarray_container_t *c, const void *d_dense) {
  return iax_write_from_dense_rows(c, d_dense, 0, NULL, NULL);
}

/* ---- read back ----------------------------------------------------------------------------------- */
static int64_t nitems(const iarray_container_t *c) {
  int64_t n = 1;
  for (int i = 0; i < c->dtshape.ndim; i++) n *= c->dtshape.shape[i];
  return n;
}

INA_API(ina_rc_t) iarray_to_buffer(iarray_context_t *ctx, iarray_container_t *c, void *buffer, int64_t buflen) {
  FAIL_NULL(c);
  const int64_t n = nitems(c);
  if (n == 0) return INA_SUCCESS;
  FAIL_NULL(ctx);
  FAIL_NULL(buffer);
  if (n * c->itemsize > buflen)
    return iax_set_error(INA_ERROR(IARRAY_ERR_TOO_SMALL_BUFFER), "The buffer size is not enough");
  void *d = NULL;
  ina_rc_t rc = iax_operand(c, 2, 6, &d);
  if (INA_FAILED(rc)) return rc;
  int e = iamx_d2h(buffer, d, (size_t)(n * c->itemsize));
  if (e) return iax_dev_error(e, "download");
  return check_decode_status();
}

INA_API(ina_rc_t) iarray_get_slice_buffer(iarray_context_t *ctx, iarray_container_t *c, const int64_t *start,
                                          const int64_t *stop, void *buffer, int64_t buflen) {
  FAIL_NULL(ctx);
  FAIL_NULL(c);
  FAIL_NULL(start);
  FAIL_NULL(stop);
  FAIL_NULL(buffer);
  const int nd = c->dtshape.ndim;
  int64_t s[2] = {0, 0}, e[2] = {1, 1}, lshape[2] = {1, 1};
  for (int i = 0; i < nd; i++) {
    const int ax = nd == 2 ? i : 1;
    const int64_t dim = c->dtshape.shape[i];
    lshape[ax] = dim;
    s[ax] = start[i] < 0 ? start[i] + dim : start[i];
    e[ax] = stop[i] < 0 ? stop[i] + dim : stop[i];
    if (s[ax] < 0 || e[ax] > dim || s[ax] > e[ax])
      return iax_set_error(INA_ERROR(INA_ERR_INVALID_ARGUMENT), "Start is bigger than stop (axis %d)", i);
  }
  const int64_t rows = e[0] - s[0], cols = e[1] - s[1];
  if (rows * cols * c->itemsize > buflen)
    return iax_set_error(INA_ERROR(IARRAY_ERR_TOO_SMALL_BUFFER), "The buffer size is not enough");
  if (rows == 0 || cols == 0) return INA_SUCCESS;
  void *d = NULL;
  ina_rc_t rc = iax_operand(c, 2, 6, &d);
  if (INA_FAILED(rc)) return rc;
  const int is = c->itemsize;
  if (cols == lshape[1]) {
    int err = iamx_d2h(buffer, (uint8_t *)d + s[0] * lshape[1] * is, (size_t)(rows * cols * is));
    if (err) return iax_dev_error(err, "download");
  } else {
    uint8_t *tmp = (uint8_t *)malloc((size_t)(rows * lshape[1] * is));
    if (!tmp) return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
    int err = iamx_d2h(tmp, (uint8_t *)d + s[0] * lshape[1] * is, (size_t)(rows * lshape[1] * is));
    if (err) {
      free(tmp);
      return iax_dev_error(err, "download");
    }
    for (int64_t r = 0; r < rows; r++)
      memcpy((uint8_t *)buffer + r * cols * is, tmp + (r * lshape[1] + s[1]) * is, (size_t)(cols * is));
    free(tmp);
  }
  return check_decode_status();
}

/* ---- info ------------------------------------------------------------------------------------------ */
INA_API(ina_rc_t) iarray_container_info(iarray_container_t *c, int64_t *nbytes, int64_t *cbytes) {
  FAIL_NULL(c);
  iarray_container_t *b = c->viewed ? c->viewed : c;
  ina_rc_t rc = iax_materialize(b);
  if (INA_FAILED(rc)) return rc;
  int64_t cb = 0;
  for (int64_t i = 0; i < b->nchunks; i++) cb += b->csizes[i];
  if (nbytes) *nbytes = nitems(c) * c->itemsize;
  if (cbytes) *cbytes = cb;
  return INA_SUCCESS;
}

INA_API(ina_rc_t) iarray_get_dtshape(iarray_context_t *ctx, iarray_container_t *c, iarray_dtshape_t *dtshape) {
  (void)ctx;
  FAIL_NULL(c);
  FAIL_NULL(dtshape);
  *dtshape = c->dtshape;
  return INA_SUCCESS;
}

INA_API(ina_rc_t) iarray_get_storage(iarray_context_t *ctx, iarray_container_t *c, iarray_storage_t *storage) {
  (void)ctx;
  FAIL_NULL(c);
  FAIL_NULL(storage);
  *storage = c->storage;
  return INA_SUCCESS;
}

INA_API(ina_rc_t) iarray_is_view(iarray_context_t *ctx, iarray_container_t *c, bool *view) {
  (void)ctx;
  FAIL_NULL(c);
  FAIL_NULL(view);
  *view = c->viewed != NULL;
  return INA_SUCCESS;
}

INA_API(ina_rc_t) iarray_container_almost_equal(iarray_container_t *a, iarray_container_t *b, double tol) {
  FAIL_NULL(a);
  FAIL_NULL(b);
  if (a->dtshape.dtype != b->dtshape.dtype) return INA_ERROR(IARRAY_ERR_INVALID_DTYPE);
  if (a->dtshape.ndim != b->dtshape.ndim) return INA_ERROR(IARRAY_ERR_INVALID_NDIM);
  for (int i = 0; i < a->dtshape.ndim; i++)
    if (a->dtshape.shape[i] != b->dtshape.shape[i]) return INA_ERROR(IARRAY_ERR_INVALID_SHAPE);
  const int64_t n = nitems(a);
  const int is = a->itemsize;
  uint8_t *ba = (uint8_t *)malloc((size_t)(n * is + 1)), *bb = (uint8_t *)malloc((size_t)(n * is + 1));
  iarray_context_t ctx;
  ctx.cfg = a->cfg;
  ina_rc_t rc = (!ba || !bb) ? INA_ERROR(INA_ERR_OUT_OF_MEMORY) : iarray_to_buffer(&ctx, a, ba, n * is);
  if (INA_SUCCEED(rc)) rc = iarray_to_buffer(&ctx, b, bb, n * is);
  for (int64_t i = 0; i < n && INA_SUCCEED(rc); i++) {
    double x, y;
    if (a->dtshape.dtype == IARRAY_DATA_TYPE_DOUBLE) { x = ((double *)ba)[i]; y = ((double *)bb)[i]; }
    else if (a->dtshape.dtype == IARRAY_DATA_TYPE_FLOAT) { x = ((float *)ba)[i]; y = ((float *)bb)[i]; }
    else { rc = INA_ERROR(IARRAY_ERR_INVALID_DTYPE); break; }
    /* the reference divides by the signed value of a (src/iarray_container.c:1255-1257) */
    if ((x - y) / x > tol) rc = iax_set_error(INA_ERROR(IARRAY_ERR_ASSERTION_FAILED), "values differ at %lld", (long long)i);
  }
  free(ba);
  free(bb);
  return rc;
}

INA_API(ina_rc_t) iarray_linalg_transpose(iarray_context_t *ctx, iarray_container_t *a, iarray_container_t **b) {
  FAIL_NULL(ctx);
  FAIL_NULL(a);
  FAIL_NULL(b);
  if (a->viewed) return iax_set_error(INA_ERROR(INA_ERR_NOT_SUPPORTED), "Transposing views is not supported yet");
  if (a->dtshape.ndim != 2)
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_NDIM), "It is only supported for two dimensions");
  iarray_container_t *v = (iarray_container_t *)calloc(1, sizeof(*v));
  if (!v) return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  *v = *a;
  v->dtshape.shape[0] = a->dtshape.shape[1];
  v->dtshape.shape[1] = a->dtshape.shape[0];
  v->viewed = a;
  v->transposed = true;
  v->plan_valid = 0;
  memset(&v->plan, 0, sizeof(v->plan));
  v->frame.valid = 0;
  v->on_disk = false;
  v->storage.urlpath = NULL;
  v->d_tblocks = NULL;
  v->t_gen = 0;
  memset(v->tuned, 0, sizeof(v->tuned));
  *b = v;
  return INA_SUCCESS;
}

/* src/iarray_container.c:364-455.  Negative start / stop count from the end;
 * start > stop is INA_ERR_INVALID_ARGUMENT, a copy whose chunkshape exceeds the
 * slice IARRAY_ERR_INVALID_CHUNKSHAPE.  A view shares its container's chunks
 * (a slice of a slice starts at the sum of the offsets); a copy is decoded on
 * the GPU and written with `storage`.  A slice of a transposed view is the
 * transposed slice of its container. */
INA_API(ina_rc_t) iarray_get_slice(iarray_context_t *ctx, iarray_container_t *src, const int64_t *start,
                                   const int64_t *stop, bool view, iarray_storage_t *storage,
                                   iarray_container_t **container) {
  FAIL_NULL(ctx);
  FAIL_NULL(src);
  FAIL_NULL(start);
  FAIL_NULL(stop);
  FAIL_NULL(container);
  if (!view) FAIL_NULL(storage);
  const int nd = src->dtshape.ndim;
  iarray_dtshape_t dts = src->dtshape;
  int64_t s[2] = {0, 0};
  for (int i = 0; i < nd; i++) {
    const int64_t dim = src->dtshape.shape[i];
    const int64_t a = start[i] < 0 ? start[i] + dim : start[i];
    const int64_t b = stop[i] < 0 ? stop[i] + dim : stop[i];
    if (a > b) return iax_set_error(INA_ERROR(INA_ERR_INVALID_ARGUMENT), "Start is bigger than stop");
    if (a < 0 || b > dim) return iax_set_error(INA_ERROR(INA_ERR_INVALID_ARGUMENT), "slice outside the array (axis %d)", i);
    if (!view && storage->chunkshape[i] > b - a)
      return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_CHUNKSHAPE), "The chunkshape is bigger than shape");
    s[nd == 2 ? i : 1] = a;
    dts.shape[i] = b - a;
  }
  if (!view) {
    const int64_t n = (nd == 2 ? dts.shape[0] * dts.shape[1] : dts.shape[0]) * src->itemsize;
    void *buf = malloc((size_t)(n > 0 ? n : 1));
    if (!buf) return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
    ina_rc_t rc = iarray_get_slice_buffer(ctx, src, start, stop, buf, n);
    if (INA_SUCCEED(rc)) rc = iarray_from_buffer(ctx, &dts, buf, n, storage, container);
    free(buf);
    return rc;
  }
  iarray_container_t *base = src->viewed ? src->viewed : src;
  iarray_container_t *v = (iarray_container_t *)calloc(1, sizeof(*v));
  if (!v) return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  *v = *base;
  v->dtshape = dts;
  v->viewed = base;
  /* offsets are kept in the base's coordinates: a transposed view's rows are
   * the base's columns (a slice of a transposed view stays transposed) */
  v->transposed = src->transposed;
  v->voff[0] = src->voff[0] + (src->transposed ? s[1] : s[0]);
  v->voff[1] = src->voff[1] + (src->transposed ? s[0] : s[1]);
  v->plan_valid = 0;
  memset(&v->plan, 0, sizeof(v->plan));
  v->frame.valid = 0;
  v->on_disk = false;
  v->storage.urlpath = NULL;
  v->d_tblocks = NULL;
  v->t_gen = 0;
  memset(v->tuned, 0, sizeof(v->tuned));
  *container = v;
  return INA_SUCCESS;
}

INA_API(ina_rc_t) iarray_container_device_plan(iarray_container_t *c, iarray_device_plan_t *plan) {
  FAIL_NULL(c);
  FAIL_NULL(plan);
  if (c->viewed) return iax_set_error(INA_ERROR(INA_ERR_NOT_SUPPORTED), "views have no decode plan of their own");
  ina_rc_t rc = iax_ensure_plan(c);
  if (INA_FAILED(rc)) return rc;
  memset(plan, 0, sizeof(*plan));
  plan->lz4 = c->plan.d_lz4;
  plan->n_lz4 = c->plan.n_lz4;
  plan->blocks = c->plan.d_blocks;
  plan->blk_rc = c->plan.d_blk_rc;
  plan->n_blocks = c->plan.n_blocks;
  plan->nbcols = (int)c->plan.nbcols;
  plan->scratch_bytes = c->plan.scratch_bytes;
  plan->blockshape[0] = c->block[0];
  plan->blockshape[1] = c->block[1];
  return INA_SUCCESS;
}

/* ---- raw chunk access -------------------------------------------------------------------------------- */
INA_API(ina_rc_t) iarray_container_nchunks(iarray_container_t *c, int64_t *nchunks) {
  FAIL_NULL(c);
  FAIL_NULL(nchunks);
  *nchunks = (c->viewed ? c->viewed : c)->nchunks;
  return INA_SUCCESS;
}

INA_API(ina_rc_t) iarray_container_chunk(iarray_container_t *c, int64_t nchunk, const uint8_t **chunk,
                                         int32_t *csize) {
  FAIL_NULL(c);
  FAIL_NULL(chunk);
  FAIL_NULL(csize);
  iarray_container_t *b = c->viewed ? c->viewed : c;
  if (nchunk < 0 || nchunk >= b->nchunks) return INA_ERROR(INA_ERR_INVALID_ARGUMENT);
  ina_rc_t rc = iax_materialize(b);
  if (INA_FAILED(rc)) return rc;
  *chunk = b->chunks[nchunk];
  *csize = b->csizes[nchunk];
  return INA_SUCCESS;
}

INA_API(ina_rc_t) iarray_from_chunks(iarray_context_t *ctx, iarray_dtshape_t *dtshape, iarray_storage_t *storage,
                                     int64_t nchunks, const uint8_t *const *chunks, const int32_t *csizes,
                                     iarray_container_t **container) {
  FAIL_NULL(chunks);
  FAIL_NULL(csizes);
  ina_rc_t rc = iax_container_new(ctx, dtshape, storage, container);
  if (INA_FAILED(rc)) return rc;
  iarray_container_t *c = *container;
  if (nchunks != c->nchunks) {
    const long long want = (long long)c->nchunks;
    iarray_container_free(ctx, container);
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_CHUNKSHAPE), "expected %lld chunks", want);
  }
  for (int64_t i = 0; i < nchunks; i++) {
    c->chunks[i] = (uint8_t *)malloc((size_t)csizes[i]);
    if (!c->chunks[i]) {
      iarray_container_free(ctx, container);
      return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
    }
    memcpy(c->chunks[i], chunks[i], (size_t)csizes[i]);
    c->csizes[i] = csizes[i];
  }
  return iax_persist(c);
}

INA_API(ina_rc_t) iarray_decode_chunks_flat(iarray_context_t *ctx, int64_t nchunks, const uint8_t *const *chunks,
                                            const int32_t *csizes, void *out, int64_t outlen) {
  FAIL_NULL(ctx);
  FAIL_NULL(chunks);
  FAIL_NULL(csizes);
  FAIL_NULL(out);
  /* one arena with every chunk; each block becomes a one-row tile at its flat offset */
  int64_t total = 0, nbytes_total = 0;
  int64_t *off = (int64_t *)malloc((size_t)nchunks * sizeof(int64_t));
  if (!off) return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  int ts0 = 0;
  for (int64_t i = 0; i < nchunks; i++) {
    iax_header_t h;
    if (iax_parse_header(chunks[i], csizes[i], &h)) {
      free(off);
      return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "chunk %lld: bad header", (long long)i);
    }
    if (!ts0) ts0 = h.typesize;
    if (h.typesize != ts0 || h.nbytes % h.typesize) {
      free(off);
      return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "chunks must share one typesize");
    }
    off[i] = total;
    total += (csizes[i] + 15) & ~15;
    nbytes_total += h.nbytes;
  }
  if (nbytes_total > outlen) {
    free(off);
    return iax_set_error(INA_ERROR(IARRAY_ERR_TOO_SMALL_BUFFER), "output too small");
  }
  uint8_t *d_arena = (uint8_t *)iamx_malloc((size_t)(total + 16));
  uint8_t *d_out = (uint8_t *)iamx_malloc((size_t)(nbytes_total + 16));
  uint8_t *host = (uint8_t *)malloc((size_t)(total + 16));
  ina_rc_t rc = INA_SUCCESS;
  if (!d_arena || !d_out) rc = iax_dev_error(iamx_init() ? iamx_init() : IAMX_ENOMEM, "flat decode buffers");
  if (!host) rc = INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  jobs_t jobs;
  memset(&jobs, 0, sizeof(jobs));
  if (INA_SUCCEED(rc)) {
    for (int64_t i = 0; i < nchunks; i++) memcpy(host + off[i], chunks[i], (size_t)csizes[i]);
    int e = iamx_h2d(d_arena, host, (size_t)total);
    if (e) rc = iax_dev_error(e, "upload");
  }
  int64_t elem0 = 0;
  for (int64_t i = 0; i < nchunks && INA_SUCCEED(rc); i++) {
    iax_header_t h;
    iax_parse_header(chunks[i], csizes[i], &h);
    for (int nb = 0; nb < h.nblocks && INA_SUCCEED(rc); nb++) {
      iamx_block_job_t J;
      memset(&J, 0, sizeof(J));
      rc = block_streams(chunks[i], csizes[i], (uint64_t)(uintptr_t)(d_arena + off[i]), &h, nb, h.typesize, &jobs, &J);
      if (INA_FAILED(rc)) break;
      const int32_t n = J.bsize / J.typesize;
      J.rows = 1;
      J.cols = n;
      J.vrows = 1;
      J.vcols = n;
      J.dst_off = elem0 + (int64_t)nb * (h.blocksize / h.typesize);
      J.dst_ld = n;
      if (push_blk(&jobs, &J, 0, 0, 0, 0)) rc = INA_ERROR(INA_ERR_OUT_OF_MEMORY);
    }
    elem0 += h.nbytes / h.typesize;
  }
  iax_plan_t plan;
  memset(&plan, 0, sizeof(plan));
  if (INA_SUCCEED(rc)) rc = upload_plan(&jobs, &plan);
  if (INA_SUCCEED(rc)) {
    void *scratch = iamx_workspace(0, (size_t)plan.scratch_bytes + 16);
    int e = scratch ? iamx_decode(plan.d_lz4, plan.n_lz4, plan.d_blocks, plan.n_blocks, scratch, d_out) : IAMX_ENOMEM;
    if (!e) e = iamx_d2h(out, d_out, (size_t)nbytes_total);
    if (e) rc = iax_dev_error(e, "flat decode");
    if (INA_SUCCEED(rc)) rc = check_decode_status();
  }
  free_plan(&plan);
  free_jobs(&jobs);
  free(host);
  free(off);
  iamx_free(d_arena);
  iamx_free(d_out);
  return rc;
}
