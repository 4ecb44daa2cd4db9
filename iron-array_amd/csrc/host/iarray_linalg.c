/*
 * iarray_linalg.c -- iarray_linalg_matmul, iarray_opt_gemm, iarray_opt_gemm_a,
 * iarray_opt_gemm_b, iarray_opt_gemv on the GPU.
 *
 * Validation, error codes and output allocation follow the reference:
 *   iarray_linalg_matmul  src/linalg/matmul.c:19-68
 *   iarray_opt_gemm       src/linalg/iarray_gemm.c:315-364 (partition rules)
 *   iarray_opt_gemm_a/_b  src/linalg/iarray_gemm_a.c:191-262, src/linalg/iarray_gemm_b.c:191-263
 *   iarray_opt_gemv       src/linalg/iarray_gemv.c:249-296
 * The arithmetic is re-planned for one GPU instead of the reference's
 * per-C-chunk panel decode + per-C-block MKL call: every A and B block is
 * decoded exactly once into a dense operand in HBM, one MFMA GEMM launch
 * computes all of C, and one launch re-shuffles the C blocks for the Blosc2
 * write (LZ4 encode of C on the host threads, part of the panels of large
 * outputs on the GPU).  The GEMV branch never builds a
 * dense A: blocks are decoded and multiplied in one fused pass.
 */
#include <string.h>

#include <stdlib.h>

#include "iarray_private.h"

static int dtype_code(iarray_data_type_t t) {
  return t == IARRAY_DATA_TYPE_DOUBLE ? 0 : t == IARRAY_DATA_TYPE_FLOAT ? 1 : -1;
}

static ina_rc_t check_status(void) { return iax_check_decode_status(); }

/* A and B decoded by one launch pair into one dense buffer (A, then B at
 * element offset b_off), so the LZ4 kernel sees the streams of both operands
 * at once; the merged job tables are kept for the next product of the same
 * two containers (generations unchanged). */
static struct {
  uint64_t gen_a, gen_b;
  iamx_lz4_job_t *d_lz4;
  iamx_block_job_t *d_blocks;
  int n_lz4, n_blocks;
  int64_t scratch_bytes, b_off;
  int valid;
} g_pair;

static void free_pair(void) {
  iamx_free(g_pair.d_lz4);
  iamx_free(g_pair.d_blocks);
  memset(&g_pair, 0, sizeof(g_pair));
}

static ina_rc_t build_pair(iarray_container_t *a, iarray_container_t *b) {
  free_pair();
  const iax_plan_t *pa = &a->plan, *pb = &b->plan;
  const int nl = pa->n_lz4 + pb->n_lz4, nb = pa->n_blocks + pb->n_blocks;
  const int64_t scr_off = (pa->scratch_bytes + 255) / 256 * 256;
  const int64_t b_off = (a->shape[0] * a->shape[1] + 63) / 64 * 64;
  iamx_lz4_job_t *lz4 = (iamx_lz4_job_t *)malloc(sizeof(iamx_lz4_job_t) * (size_t)(nl ? nl : 1));
  iamx_block_job_t *blk = (iamx_block_job_t *)malloc(sizeof(iamx_block_job_t) * (size_t)(nb ? nb : 1));
  if (!lz4 || !blk) {
    free(lz4);
    free(blk);
    return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  }
  if (pa->n_lz4) memcpy(lz4, pa->h_lz4, sizeof(iamx_lz4_job_t) * (size_t)pa->n_lz4);
  for (int i = 0; i < pb->n_lz4; i++) {
    lz4[pa->n_lz4 + i] = pb->h_lz4[i];
    lz4[pa->n_lz4 + i].dst_off += scr_off;
  }
  if (pa->n_blocks) memcpy(blk, pa->h_blocks, sizeof(iamx_block_job_t) * (size_t)pa->n_blocks);
  for (int i = 0; i < pb->n_blocks; i++) {
    iamx_block_job_t J = pb->h_blocks[i];
    for (int k = 0; k < J.nstreams; k++)
      if (J.kind[k] == IAMX_SRC_SCRATCH) J.src[k] += (uint64_t)scr_off;
    J.dst_off += b_off;
    blk[pa->n_blocks + i] = J;
  }
  iax_sort_lz4_jobs(lz4, nl);
  ina_rc_t rc = INA_SUCCESS;
  g_pair.d_lz4 = (iamx_lz4_job_t *)iamx_malloc(sizeof(iamx_lz4_job_t) * (size_t)(nl ? nl : 1));
  g_pair.d_blocks = (iamx_block_job_t *)iamx_malloc(sizeof(iamx_block_job_t) * (size_t)(nb ? nb : 1));
  if (!g_pair.d_lz4 || !g_pair.d_blocks) rc = iax_dev_error(IAMX_ENOMEM, "pair plan");
  int e = 0;
  if (INA_SUCCEED(rc) && nl) e = iamx_h2d(g_pair.d_lz4, lz4, sizeof(iamx_lz4_job_t) * (size_t)nl);
  if (INA_SUCCEED(rc) && !e && nb) e = iamx_h2d(g_pair.d_blocks, blk, sizeof(iamx_block_job_t) * (size_t)nb);
  if (e) rc = iax_dev_error(e, "pair plan upload");
  free(lz4);
  free(blk);
  if (INA_FAILED(rc)) {
    free_pair();
    return rc;
  }
  g_pair.gen_a = a->generation;
  g_pair.gen_b = b->generation;
  g_pair.n_lz4 = nl;
  g_pair.n_blocks = nb;
  g_pair.scratch_bytes = scr_off + pb->scratch_bytes;
  g_pair.b_off = b_off;
  g_pair.valid = 1;
  return INA_SUCCESS;
}

static ina_rc_t decode_operands(iarray_container_t *a, iarray_container_t *b, void **dA, void **dB) {
  if (a->viewed || b->viewed || a == b) { /* views are transposed on the way; one container is decoded once */
    ina_rc_t rc = iax_operand(a, 2, 6, dA);
    if (INA_SUCCEED(rc)) rc = a == b ? (*dB = *dA, INA_SUCCESS) : iax_operand(b, 3, 7, dB);
    return rc;
  }
  ina_rc_t rc = iax_ensure_plan(a);
  if (INA_SUCCEED(rc)) rc = iax_ensure_plan(b);
  if (INA_FAILED(rc)) return rc;
  if (!g_pair.valid || g_pair.gen_a != a->generation || g_pair.gen_b != b->generation) {
    rc = build_pair(a, b);
    if (INA_FAILED(rc)) return rc;
  }
  const size_t bytes = (size_t)((g_pair.b_off + b->shape[0] * b->shape[1]) * a->itemsize);
  iax_operand_forget(2);
  uint8_t *dense = (uint8_t *)iamx_workspace(2, bytes + 16);
  void *scratch = iamx_workspace(0, (size_t)g_pair.scratch_bytes + 16);
  if (!dense || !scratch) return iax_dev_error(IAMX_ENOMEM, "operand buffers");
  int e = iamx_decode(g_pair.d_lz4, g_pair.n_lz4, g_pair.d_blocks, g_pair.n_blocks, scratch, dense);
  if (e) return iax_dev_error(e, "decode");
  *dA = dense;
  *dB = dense + g_pair.b_off * a->itemsize;
  return INA_SUCCESS;
}

typedef struct {
  const uint8_t *A;
  uint8_t *C;
  int64_t K, N, is;
} rows_ctx_t;

static int gemm_rows(void *v, int64_t row0, int64_t nrows) {
  const rows_ctx_t *x = (const rows_ctx_t *)v;
  return iamx_gemm_rows(nrows, x->A + row0 * x->K * x->is, x->K, x->C + row0 * x->N * x->is, x->N);
}

/* rows per C panel (rounded up to whole chunk rows): M/8, so the encode of
 * seven panels overlaps GEMM work (8192-row products measured 27.3 ms with
 * 1024-row panels against 27.6 ms with 2048 and 36.4 ms with one panel), but
 * at least 2^23/N rows, 512 tiles of 128x128, so a panel still fills the 256
 * CUs (2000^2 products: 4 panels of 500 rows took 0.375 ms each, one launch
 * for all of C 0.39 ms).  Split-bf16 and Ozaki-I launch one workgroup per
 * 256x256 tile, so their panels hold at least 2^24/N rows, 256 such tiles
 * (8192^2 split-bf16: 0.453 ms per 1024 rows with 2048-row panels against
 * 0.618 ms with 1024); the CRT form launches a workgroup per tile and
 * modulus and keeps the plain rule.  IARRAY_GEMM_PANEL_ROWS overrides (0 =
 * one panel) */
static int64_t panel_rows(int64_t M, int64_t N, int prec) {
  const char *env = getenv("IARRAY_GEMM_PANEL_ROWS");
  if (env) {
    const long long v = atoll(env);
    return v > 0 ? (int64_t)v : M;
  }
  int64_t r = M / 8;
  const int big = prec == IARRAY_MATMUL_SPLIT_BF16 || prec == IARRAY_MATMUL_OZAKI_INT8;
  const int64_t fill = ((int64_t)1 << (big ? 24 : 23)) / (N > 0 ? N : 1);
  if (r < fill) r = fill;
  return r;
}

/* C = A B with A, B (views allowed) already validated; c allocated.  C is
 * computed in row panels, each packed and downloaded as soon as it is done,
 * so the host LZ4 encode of a panel overlaps the GEMM of the next. */
/* ---- zero-block skipping of the opt paths ----------------------------------------------------
 * opt: 0 a plain matmul (nothing skipped, as gemm_blosc / gemv_blosc), 1
 * iarray_opt_gemm, 2 iarray_opt_gemm_a, 3 iarray_opt_gemm_b, 4
 * iarray_opt_gemv.  A pair (A block, B block) is skipped when either is a
 * zero block by the reference's block test, or lies in a zero chunk of the
 * operand whose chunks the variant tests: both for opt_gemm
 * (src/linalg/iarray_gemm.c:143-165), B's for opt_gemm_a and A's for
 * opt_gemm_b (iarray_gemm_a.c / _b.c:100-175), A's and x's for opt_gemv
 * (iarray_gemv.c:125-160). */
enum { OPT_NONE = 0, OPT_GEMM = 1, OPT_GEMM_A = 2, OPT_GEMM_B = 3, OPT_GEMV = 4 };

/* [first, end) of every block along one axis of a container's tiling,
 * trimmed at chunk and array edges (all-padding blocks are empty). */
static int32_t *axis_blocks(const iarray_container_t *c, int ax, int64_t *n) {
  const int64_t nc = c->nchunks_ax[ax], nb = c->nblocks_ax[ax];
  int32_t *r = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)(nc * nb > 0 ? nc * nb : 1));
  if (!r) return NULL;
  for (int64_t ci = 0; ci < nc; ci++)
    for (int64_t bi = 0; bi < nb; bi++) {
      const int64_t first = ci * c->chunk[ax] + bi * c->block[ax];
      int64_t end = first + c->block[ax];
      if (end > ci * c->chunk[ax] + c->chunk[ax]) end = ci * c->chunk[ax] + c->chunk[ax];
      if (end > c->shape[ax]) end = c->shape[ax];
      if (end < first) end = first;
      r[2 * (ci * nb + bi)] = (int32_t)first;
      r[2 * (ci * nb + bi) + 1] = (int32_t)end;
    }
  *n = nc * nb;
  return r;
}

/* zero flag of block (g0, g1) of the global block grid of c */
static uint8_t zero_block(const iarray_container_t *c, int64_t g0, int64_t g1, int chunk_rule) {
  const int64_t ci0 = g0 / c->nblocks_ax[0], bi0 = g0 % c->nblocks_ax[0];
  const int64_t ci1 = g1 / c->nblocks_ax[1], bi1 = g1 % c->nblocks_ax[1];
  const int64_t nc = ci0 * c->nchunks_ax[1] + ci1;
  const uint8_t *ch = c->chunks[nc];
  if (!ch) return 0;
  if (chunk_rule && iax_chunk_is_zeros(ch, c->csizes[nc])) return 1;
  return (uint8_t)iax_block_is_zeros(ch, c->csizes[nc], (int)(bi0 * c->nblocks_ax[1] + bi1));
}

typedef struct {
  int dt;
  const uint8_t *A, *B;
  uint8_t *C;
  int64_t K, N;
  int32_t *rows; /* host copy, for the block-row range of a panel */
  int64_t nbr, nbk, nbc;
  const int32_t *d_rows, *d_ks, *d_cols;
  const uint8_t *d_za, *d_zb;
  int max_rows, max_cols;
} bp_ctx_t;

static int blockpair_rows(void *v, int64_t row0, int64_t nrows) {
  const bp_ctx_t *x = (const bp_ctx_t *)v;
  int64_t br0 = 0, br1;
  while (br0 < x->nbr && x->rows[2 * br0] < row0) br0++;
  br1 = br0;
  while (br1 < x->nbr && x->rows[2 * br1] < row0 + nrows) br1++;
  return iamx_gemm_blockpairs(x->dt, x->A, x->K, x->B, x->N, x->C, x->N, x->d_rows, (int)x->nbr, x->d_ks,
                              (int)x->nbk, x->d_cols, (int)x->nbc, x->d_za, x->d_zb, (int)br0, (int)br1, x->max_rows,
                              x->max_cols);
}

/* Block-pair product when any pair is skipped; *used = 0 (and nothing
 * launched) when no zero block exists, so the plain GEMM gives the same C. */
static ina_rc_t run_blockpairs(iarray_container_t *a, iarray_container_t *b, iarray_container_t *c, int opt,
                               const void *dA, const void *dB, void *dC, int *used) {
  *used = 0;
  int64_t nbr, nbk, nbk2, nbc;
  int32_t *rows = axis_blocks(a, 0, &nbr), *ks = axis_blocks(a, 1, &nbk), *cols = axis_blocks(b, 1, &nbc);
  int32_t *ks2 = axis_blocks(b, 0, &nbk2);
  ina_rc_t rc = INA_SUCCESS;
  uint8_t *za = NULL, *zb = NULL;
  if (!rows || !ks || !cols || !ks2) rc = INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  else if (nbk2 != nbk || memcmp(ks, ks2, sizeof(int32_t) * 2 * (size_t)nbk) != 0)
    rc = iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_BLOCKSHAPE), "A's column blocks differ from B's row blocks");
  if (INA_SUCCEED(rc)) {
    za = (uint8_t *)calloc((size_t)(nbr * nbk > 0 ? nbr * nbk : 1), 1);
    zb = (uint8_t *)calloc((size_t)(nbk * nbc > 0 ? nbk * nbc : 1), 1);
    if (!za || !zb) rc = INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  }
  int any = 0;
  if (INA_SUCCEED(rc)) {
    const int rule_a = opt == OPT_GEMM || opt == OPT_GEMM_B, rule_b = opt == OPT_GEMM || opt == OPT_GEMM_A;
    for (int64_t i = 0; i < nbr; i++)
      for (int64_t k = 0; k < nbk; k++) any |= za[i * nbk + k] = zero_block(a, i, k, rule_a);
    for (int64_t k = 0; k < nbk; k++)
      for (int64_t j = 0; j < nbc; j++) any |= zb[k * nbc + j] = zero_block(b, k, j, rule_b);
  }
  if (INA_SUCCEED(rc) && any) {
    const size_t o_ks = (size_t)(2 * nbr) * 4, o_cols = o_ks + (size_t)(2 * nbk) * 4,
                 o_za = o_cols + (size_t)(2 * nbc) * 4, o_zb = o_za + (size_t)(nbr * nbk),
                 total = o_zb + (size_t)(nbk * nbc);
    uint8_t *h = (uint8_t *)malloc(total);
    uint8_t *d = (uint8_t *)iamx_workspace(15, total + 16);
    if (!h) rc = INA_ERROR(INA_ERR_OUT_OF_MEMORY);
    else if (!d) rc = iax_dev_error(IAMX_ENOMEM, "block-pair tables");
    if (INA_SUCCEED(rc)) {
      memcpy(h, rows, o_ks);
      memcpy(h + o_ks, ks, o_cols - o_ks);
      memcpy(h + o_cols, cols, o_za - o_cols);
      memcpy(h + o_za, za, o_zb - o_za);
      memcpy(h + o_zb, zb, total - o_zb);
      int e = iamx_h2d(d, h, total);
      if (e) rc = iax_dev_error(e, "block-pair tables");
    }
    free(h);
    if (INA_SUCCEED(rc)) {
      bp_ctx_t x = {dtype_code(a->dtshape.dtype), (const uint8_t *)dA, (const uint8_t *)dB, (uint8_t *)dC,
                    a->shape[1], b->shape[1], rows, nbr, nbk, nbc, (const int32_t *)d,
                    (const int32_t *)(d + o_ks), (const int32_t *)(d + o_cols), d + o_za, d + o_zb,
                    (int)a->block[0], (int)b->block[1]};
      *used = 1;
      rc = iax_write_from_dense_rows(c, dC, panel_rows(a->shape[0], b->shape[1], IARRAY_MATMUL_PLAIN), blockpair_rows, &x);
    }
  }
  free(rows);
  free(ks);
  free(ks2);
  free(cols);
  free(za);
  free(zb);
  return rc;
}

static ina_rc_t run_gemm(iarray_container_t *a, iarray_container_t *b, iarray_container_t *c,
                         const iarray_matmul_params_t *params, int opt) {
  const int dt = dtype_code(a->dtshape.dtype);
  const int64_t M = a->dtshape.shape[0], K = a->dtshape.shape[1], N = b->dtshape.shape[1];
  int prec = params ? (int)params->precision : 0;
  int slices = params ? params->ozaki_slices : 0;
  if (prec == IARRAY_MATMUL_SPLIT_BF16 && dt != 1)
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_DTYPE), "split-bf16 needs float32 operands");
  if ((prec == IARRAY_MATMUL_OZAKI_INT8 || prec == IARRAY_MATMUL_OZAKI_CRT) && dt != 0)
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_DTYPE), "ozaki forms need float64 operands");
  if (prec < IARRAY_MATMUL_PLAIN || prec > IARRAY_MATMUL_OZAKI_CRT)
    return iax_set_error(INA_ERROR(INA_ERR_INVALID_ARGUMENT), "unknown matmul precision %d", prec);
  void *dA = NULL, *dB = NULL;
  ina_rc_t rc = decode_operands(a, b, &dA, &dB);
  if (INA_FAILED(rc)) return rc;
  void *dC = iamx_workspace(4, (size_t)(M * N * a->itemsize) + 16);
  if (!dC) return iax_dev_error(IAMX_ENOMEM, "C buffer");
  if (opt != OPT_NONE && prec == IARRAY_MATMUL_PLAIN) {
    int used = 0;
    rc = run_blockpairs(a, b, c, opt, dA, dB, dC, &used);
    if (INA_FAILED(rc)) return rc;
    if (used) return check_status();
  }
  if (prec == IARRAY_MATMUL_OZAKI_INT8 || prec == IARRAY_MATMUL_OZAKI_CRT) {
    iax_last_slice_err = -1.0;
    if (slices == 0) {
      const iax_operand_key_t ka = iax_operand_key(a), kb = iax_operand_key(b);
      const int t = prec == IARRAY_MATMUL_OZAKI_CRT;
      if (a->tuned[t].b == b && !memcmp(&a->tuned[t].key_a, &ka, sizeof(ka)) &&
          !memcmp(&a->tuned[t].key_b, &kb, sizeof(kb)) && a->tuned[t].count > 0) {
        slices = a->tuned[t].count;
      } else {
        const char *env = getenv("IARRAY_OZAKI_TARGET");
        const double target = env && atof(env) > 0 ? atof(env) : 1e-13;
        double err = -1.0;
        int e = prec == IARRAY_MATMUL_OZAKI_CRT
                    ? iamx_ozaki_crt_pick_moduli(M, N, K, (const double *)dA, K, (const double *)dB, N, target, 8,
                                                 &slices, &err)
                    : iamx_ozaki_pick_slices(M, N, K, (const double *)dA, K, (const double *)dB, N, target, 4,
                                             &slices, &err);
        if (e) return iax_dev_error(e, "ozaki count tuning");
        a->tuned[t].b = b;
        a->tuned[t].key_a = ka;
        a->tuned[t].key_b = kb;
        a->tuned[t].count = slices;
        iax_last_slice_err = err;
      }
    }
    iax_last_slices = slices;
  }
  int e = iamx_gemm_prepare_b(dt, prec, slices, N, K, dB, N);
  if (e) return iax_dev_error(e, "gemm");
  rows_ctx_t rx = {(const uint8_t *)dA, (uint8_t *)dC, K, N, a->itemsize};
  iax_write_form = 1 + prec * 16 + slices;
  rc = iax_write_from_dense_rows(c, dC, panel_rows(M, N, prec), gemm_rows, &rx);
  iax_write_form = 0;
  if (INA_FAILED(rc)) return rc;
  return check_status();
}

/* y = A x ; fused block decode + GEMV when A is a plain container. */
static ina_rc_t run_gemv(iarray_container_t *a, iarray_container_t *x, iarray_container_t *y, int opt) {
  const int dt = dtype_code(a->dtshape.dtype);
  const int64_t M = a->dtshape.shape[0], K = a->dtshape.shape[1];
  void *dx = NULL;
  ina_rc_t rc = iax_operand(x, 3, 7, &dx);
  if (INA_FAILED(rc)) return rc;
  void *dy = iamx_workspace(4, (size_t)(M * a->itemsize) + 16);
  if (!dy) return iax_dev_error(IAMX_ENOMEM, "y buffer");
  int e = 0;
  iax_last_gemv_panels = 0;
  if (!a->viewed && a->block[1] <= 1024) {
    const int streamed = iax_gemv_should_stream(a);
    if (!streamed) {
      rc = iax_ensure_plan(a);
      if (INA_FAILED(rc)) return rc;
    }
    uint8_t *d_xzero = NULL;
    if (opt == OPT_GEMV) {
      /* zero flags of x's blocks (its chunk test and block test), one per
       * block column of A: the two tilings agree along K */
      const int64_t nbx = x->nchunks_ax[1] * x->nblocks_ax[1];
      if (nbx != a->nchunks_ax[1] * a->nblocks_ax[1])
        return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_BLOCKSHAPE), "x's blocks differ from A's block columns");
      uint8_t *hz = (uint8_t *)malloc((size_t)nbx);
      d_xzero = (uint8_t *)iamx_workspace(15, (size_t)nbx + 16);
      if (!hz) return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
      if (!d_xzero) {
        free(hz);
        return iax_dev_error(IAMX_ENOMEM, "x zero flags");
      }
      for (int64_t g = 0; g < nbx; g++) hz[g] = zero_block(x, 0, g, 1);
      e = iamx_h2d(d_xzero, hz, (size_t)nbx);
      free(hz);
    }
    if (streamed) {
      if (!e) {
        rc = iax_gemv_streamed(a, dt, dx, dy, d_xzero);
        if (INA_FAILED(rc)) return rc;
      }
    } else {
      const iax_plan_t *p = &a->plan;
      void *scratch = iamx_workspace(0, (size_t)p->scratch_bytes + 16);
      const size_t pbytes = (size_t)(p->nbcols * M * a->itemsize);
      iax_operand_forget(6);
      void *partial = iamx_workspace(6, pbytes + 16);
      if (!scratch || !partial) return iax_dev_error(IAMX_ENOMEM, "gemv scratch");
      if (!e) e = iamx_memset(partial, 0, pbytes);
      if (!e && p->n_blocks > 0 && d_xzero)
        e = iamx_opt_gemv_fused(dt, p->d_lz4, p->n_lz4, p->d_blocks, p->d_blk_rc, p->n_blocks, (int)p->nbcols,
                                (int)a->block[0], (int)a->block[1], dx, dy, M, scratch, partial, d_xzero);
      else if (!e && p->n_blocks > 0)
        e = iamx_gemv_fused(dt, p->d_lz4, p->n_lz4, p->d_blocks, p->d_blk_rc, p->n_blocks, (int)p->nbcols,
                            (int)a->block[0], (int)a->block[1], dx, dy, M, scratch, partial);
      else if (!e)
        e = iamx_memset(dy, 0, (size_t)(M * a->itemsize));
    }
  } else {
    void *dA = NULL;
    rc = iax_operand(a, 2, 6, &dA);
    if (INA_FAILED(rc)) return rc;
    e = iamx_gemv(dt, M, K, dA, K, dx, dy);
  }
  if (e) return iax_dev_error(e, "gemv");
  rc = iax_write_from_dense(y, dy);
  if (INA_FAILED(rc)) return rc;
  return check_status();
}

static ina_rc_t new_output(iarray_context_t *ctx, iarray_container_t *a, iarray_container_t *b,
                           iarray_storage_t *storage, iarray_container_t **c) {
  iarray_dtshape_t dts;
  memset(&dts, 0, sizeof(dts));
  dts.dtype = a->dtshape.dtype;
  dts.ndim = b->dtshape.ndim;
  dts.shape[0] = a->dtshape.shape[0];
  if (dts.ndim > 1) dts.shape[1] = b->dtshape.shape[1];
  for (int i = 0; i < dts.ndim; i++)
    if (dts.shape[i] < storage->chunkshape[i])
      return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_CHUNKSHAPE), "output chunkshape exceeds the output shape");
  return iax_container_new(ctx, &dts, storage, c);
}

static ina_rc_t matmul_impl(iarray_context_t *ctx, iarray_container_t *a, iarray_container_t *b,
                            iarray_storage_t *storage, const iarray_matmul_params_t *params, int opt,
                            iarray_container_t **c) {
  if (!ctx || !a || !b || !storage || !c) return INA_ERROR(INA_ERR_NULL_POINTER);
  if (a->dtshape.ndim != 2 || (b->dtshape.ndim != 1 && b->dtshape.ndim != 2))
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_NDIM), "matmul needs a 2-D and a 1-D or 2-D operand");
  if (a->dtshape.shape[1] != b->dtshape.shape[0])
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_SHAPE), "inner dimensions differ");
  if (a->dtshape.dtype != b->dtshape.dtype) return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_DTYPE), "dtypes differ");
  if (dtype_code(a->dtshape.dtype) < 0)
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_DTYPE), "matmul is defined for float32 and float64");
  ina_rc_t rc = new_output(ctx, a, b, storage, c);
  if (INA_FAILED(rc)) return rc;
  iamx_prof_reset();
  rc = b->dtshape.ndim == 2 ? run_gemm(a, b, *c, params, opt) : run_gemv(a, b, *c, opt);
  iax_last_device_ms = iamx_prof_span_ms();
  if (INA_SUCCEED(rc)) rc = iax_persist(*c);
  /* a failed product leaves no partly written output behind */
  if (INA_FAILED(rc)) iarray_container_free(ctx, c);
  return rc;
}

INA_API(ina_rc_t) iarray_linalg_matmul(iarray_context_t *ctx, iarray_container_t *a, iarray_container_t *b,
                                       iarray_storage_t *storage, iarray_container_t **c) {
  return matmul_impl(ctx, a, b, storage, NULL, OPT_NONE, c);
}

INA_API(ina_rc_t) iarray_linalg_matmul_params(iarray_context_t *ctx, iarray_container_t *a, iarray_container_t *b,
                                              iarray_storage_t *storage, const iarray_matmul_params_t *params,
                                              iarray_container_t **c) {
  return matmul_impl(ctx, a, b, storage, params, OPT_NONE, c);
}

INA_API(ina_rc_t) iarray_opt_gemm(iarray_context_t *ctx, iarray_container_t *a, iarray_container_t *b,
                                  iarray_storage_t *storage, iarray_container_t **c) {
  if (!ctx || !a || !b || !storage || !c) return INA_ERROR(INA_ERR_NULL_POINTER);
  if (a->viewed || b->viewed) return iax_set_error(INA_ERROR(INA_ERR_NOT_SUPPORTED), "opt_gemm reads chunks directly");
  if (a->ndim != 2) return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_NDIM), "The a dimension must be 2");
  if (b->ndim != 2) return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_NDIM), "The b dimension must be 2");
  if (a->chunk[0] != storage->chunkshape[0])
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_CHUNKSHAPE), "a->chunkshape[0] != c->chunkshape[0]");
  if (a->chunk[1] != b->chunk[0])
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_CHUNKSHAPE), "a->chunkshape[1] != b->chunkshape[0]");
  if (b->chunk[1] != storage->chunkshape[1])
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_CHUNKSHAPE), "b->chunkshape[1] != c->chunkshape[1]");
  if (a->block[0] != storage->blockshape[0])
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_BLOCKSHAPE), "a->blockshape[0] != c->blockshape[0]");
  if (a->block[1] != b->block[0])
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_BLOCKSHAPE), "a->blockshape[1] != b->blockshape[0]");
  if (b->block[1] != storage->blockshape[1])
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_BLOCKSHAPE), "b->blockshape[1] != c->blockshape[1]");
  return matmul_impl(ctx, a, b, storage, NULL, OPT_GEMM, c);
}

#define PART_ERR(code, msg) return iax_set_error(INA_ERROR(code), msg)

/* Checks shared by opt_gemm_a and opt_gemm_b, src/linalg/iarray_gemm_a.c:198-230
 * and src/linalg/iarray_gemm_b.c:198-230, in their order. */
static ina_rc_t opt_ab_head(iarray_context_t *ctx, iarray_container_t *a, iarray_container_t *b,
                            iarray_storage_t *storage, iarray_container_t **c, const char *name) {
  if (!ctx || !a || !b || !storage || !c) return INA_ERROR(INA_ERR_NULL_POINTER);
  if (a->viewed || b->viewed) return iax_set_error(INA_ERROR(INA_ERR_NOT_SUPPORTED), "%s reads chunks directly", name);
  if (a->ndim != 2) PART_ERR(IARRAY_ERR_INVALID_NDIM, "The a dimension must be 2");
  if (b->ndim != 2) PART_ERR(IARRAY_ERR_INVALID_NDIM, "The b dimension must be 2");
  if (a->shape[1] != b->shape[0]) PART_ERR(IARRAY_ERR_INVALID_SHAPE, "a->shape[1] != b->shape[0]");
  if (a->chunk[1] < a->shape[1]) PART_ERR(IARRAY_ERR_INVALID_CHUNKSHAPE, "a->chunkshape[1]  < a->shape[1]");
  if (b->chunk[0] < b->shape[0]) PART_ERR(IARRAY_ERR_INVALID_CHUNKSHAPE, "b->chunkshape[0] < c->chunkshape[0]");
  if (a->chunk[1] != b->chunk[0]) PART_ERR(IARRAY_ERR_INVALID_CHUNKSHAPE, "a->chunkshape[1] != b->chunkshape[0]");
  return INA_SUCCESS;
}

/* Block checks shared by both, src/linalg/iarray_gemm_a.c:241-262 and
 * src/linalg/iarray_gemm_b.c:241-263. */
static ina_rc_t opt_ab_blocks(const iarray_container_t *a, const iarray_container_t *b, const iarray_storage_t *storage) {
  if (a->chunk[0] != storage->blockshape[0]) PART_ERR(IARRAY_ERR_INVALID_BLOCKSHAPE, "a->chunkshape[0] != c->blockshape[0]");
  if (b->chunk[1] != storage->blockshape[1]) PART_ERR(IARRAY_ERR_INVALID_BLOCKSHAPE, "b->chunkshape[1] != c->blockshape[1]");
  if (a->chunk[0] != a->block[0]) PART_ERR(IARRAY_ERR_INVALID_BLOCKSHAPE, "a->chunkshape[0] != a->blockshape[0]");
  if (b->chunk[1] != b->block[1]) PART_ERR(IARRAY_ERR_INVALID_BLOCKSHAPE, "b->chunkshape[1] != b->blockshape[1]");
  if (a->block[1] != b->block[0]) PART_ERR(IARRAY_ERR_INVALID_BLOCKSHAPE, "a->blockshape[1] != b->blockshape[0]");
  return INA_SUCCESS;
}

/* The reference computes each C block as the sum over K blocks of A block x
 * B block (one cblas_?gemm each); here the whole product is the GPU GEMM. */
INA_API(ina_rc_t) iarray_opt_gemm_a(iarray_context_t *ctx, iarray_container_t *a, iarray_container_t *b,
                                    iarray_storage_t *storage, iarray_container_t **c) {
  ina_rc_t rc = opt_ab_head(ctx, a, b, storage, c, "opt_gemm_a");
  if (INA_FAILED(rc)) return rc;
  /* src/linalg/iarray_gemm_a.c:232-239 */
  if (a->chunk[0] != storage->chunkshape[0]) PART_ERR(IARRAY_ERR_INVALID_CHUNKSHAPE, "a->chunkshape[0] = c->chunkshape[0]");
  if (b->shape[1] > storage->chunkshape[1]) PART_ERR(IARRAY_ERR_INVALID_CHUNKSHAPE, "b->shape[1] > c->chunkshape[1]");
  rc = opt_ab_blocks(a, b, storage);
  if (INA_FAILED(rc)) return rc;
  return matmul_impl(ctx, a, b, storage, NULL, OPT_GEMM_A, c);
}

INA_API(ina_rc_t) iarray_opt_gemm_b(iarray_context_t *ctx, iarray_container_t *a, iarray_container_t *b,
                                    iarray_storage_t *storage, iarray_container_t **c) {
  ina_rc_t rc = opt_ab_head(ctx, a, b, storage, c, "opt_gemm_b");
  if (INA_FAILED(rc)) return rc;
  /* src/linalg/iarray_gemm_b.c:232-239 */
  if (a->shape[0] > storage->chunkshape[0]) PART_ERR(IARRAY_ERR_INVALID_CHUNKSHAPE, "a->shape[0] > c->chunkshape[0]");
  if (b->chunk[1] != storage->chunkshape[1]) PART_ERR(IARRAY_ERR_INVALID_CHUNKSHAPE, "b->chunkshape[1] != c->chunkshape[1]");
  rc = opt_ab_blocks(a, b, storage);
  if (INA_FAILED(rc)) return rc;
  return matmul_impl(ctx, a, b, storage, NULL, OPT_GEMM_B, c);
}

INA_API(ina_rc_t) iarray_opt_gemv(iarray_context_t *ctx, iarray_container_t *a, iarray_container_t *b,
                                  iarray_storage_t *storage, iarray_container_t **c) {
  if (!ctx || !a || !b || !storage || !c) return INA_ERROR(INA_ERR_NULL_POINTER);
  if (a->viewed || b->viewed) return iax_set_error(INA_ERROR(INA_ERR_NOT_SUPPORTED), "opt_gemv reads chunks directly");
  if (a->ndim != 2) return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_NDIM), "The a dimension must be 2");
  if (b->ndim != 1) return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_NDIM), "The b dimension must be 1");
  if (a->chunk[0] != storage->chunkshape[0])
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_CHUNKSHAPE), "a->chunkshape[0] != c->chunkshape[0]");
  if (a->chunk[1] != b->chunk[1])
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_CHUNKSHAPE), "a->chunkshape[1] != b->chunkshape[0]");
  if (a->block[0] != storage->blockshape[0])
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_BLOCKSHAPE), "a->blockshape[0] != c->blockshape[0]");
  if (a->block[1] != b->block[1])
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_BLOCKSHAPE), "a->blockshape[1] != b->blockshape[0]");
  return matmul_impl(ctx, a, b, storage, NULL, OPT_GEMV, c);
}
