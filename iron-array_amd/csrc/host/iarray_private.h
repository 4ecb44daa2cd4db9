/*
 * iarray_private.h -- host-side types of libiarray.so (not installed).
 *
 * A container is a Caterva-style two-level tiled array (chunks of blocks)
 * whose chunks are Blosc2 chunks (LZ4 + shuffle/bitshuffle).  The compressed
 * chunks live in host memory (the super-chunk) and are mirrored once into one
 * device arena; the decode plan (per-block stream table) is built from the
 * chunk headers and kept on the device next to it.
 * Mirrors the roles of iarray_context_s / iarray_container_s
 * (src/iarray_private.h:79-117).
 */
#ifndef IAMX_IARRAY_PRIVATE_H_
#define IAMX_IARRAY_PRIVATE_H_

#include <stdint.h>

#include "../../../include/iamx.h"
#include "../../../include/iarray.h"

struct iarray_context_s {
  iarray_config_t cfg;
};

/* Decode plan of a container: device job tables for iamx_decode. */
typedef struct iax_plan_s {
  iamx_lz4_job_t *d_lz4;
  int n_lz4;
  iamx_block_job_t *d_blocks;
  int n_blocks;
  int32_t *d_blk_rc; /* (row0, col0, block column, 0) per block job, for the fused GEMV */
  int64_t scratch_bytes;
  int64_t nbcols; /* block columns of the global block grid */
  /* host copies of the job tables (the batch path concatenates them) */
  iamx_lz4_job_t *h_lz4;
  iamx_block_job_t *h_blocks;
} iax_plan_t;

/* Blosc2 frame header fields that a container's chunks do not determine
 * (general flags, thread counts, the variable-length-metalayer flag, the
 * coding parameters).  A container read from a frame keeps them, so saving it
 * again writes the same header; other containers take them from their config. */
typedef struct {
  int valid;
  uint8_t flags[4];
  int64_t nthreads[2];
  uint8_t vlmeta; /* msgpack boolean byte */
  uint8_t coding_type;
  uint8_t coding[16];
} iax_frame_params_t;

/* What an operand reads: its base container's device generation (unique
 * across containers, renewed when the arena is rebuilt) and the view's
 * rectangle of that base.  Two operands with equal keys hold the same
 * elements. */
typedef struct {
  uint64_t gen;
  int64_t voff[2], shape[2];
  int transposed;
} iax_operand_key_t;

iax_operand_key_t iax_operand_key(const iarray_container_t *c);

struct iarray_container_s {
  iarray_dtshape_t dtshape; /* logical shape (a transposed view swaps it) */
  iarray_storage_t storage;
  iarray_config_t cfg;      /* codec settings the chunks are written with */
  int itemsize;
  int ndim;                 /* underlying ndim (1 or 2) */
  /* tiling, normalised to 2-D: a 1-D array is (1, n) */
  int64_t shape[2], chunk[2], block[2], extshape[2], extchunk[2];
  int64_t nchunks_ax[2], nblocks_ax[2];
  int64_t nchunks;
  int64_t chunk_nbytes; /* prod(extchunkshape) * itemsize */
  int32_t blocksize;    /* prod(blockshape) * itemsize */
  /* host super-chunk */
  uint8_t **chunks;
  int32_t *csizes;
  /* device mirror */
  uint8_t *d_arena;
  int64_t *arena_off;
  int device_valid;
  uint64_t generation; /* bumped whenever the device arena is (re)built */
  iax_plan_t plan;
  int plan_valid;
  /* views: a transposed view, or a slice of `viewed` starting at voff (2-D
   * normalised coordinates of the viewed container) */
  iarray_container_t *viewed;
  bool transposed;
  int64_t voff[2];
  /* a transposed view's block jobs: the base's, writing their tiles
   * transposed (IAMX_JOB_TRANSPOSE), valid while the base's generation is
   * t_gen */
  iamx_block_job_t *d_tblocks;
  uint64_t t_gen;
  /* frames: header fields of the frame it was read from; opened (rather than
   * loaded) containers count as stored on disk */
  iax_frame_params_t frame;
  bool on_disk;
  /* a contiguous frame opened with iarray_container_open: the file is mapped
   * read-only, and the data chunks are located in the mapping (their pages
   * read from the file) only when something first needs them
   * (iax_materialize); until then chunks[i] is NULL and off[i] >= 0 is the
   * chunk's offset in the data section [hsize, opos) of the file.  Chunks in
   * the mapping are borrowed, not owned. */
  struct {
    uint8_t *map;
    size_t map_len;
    int64_t hsize, opos;
    int64_t *off;
    bool pending;
  } lazy;
  /* Ozaki slice count [0] and CRT modulus count [1] picked for this
   * container as A against one B: valid while both operands read the same
   * elements (iax_operand_key) */
  struct {
    const iarray_container_t *b;
    iax_operand_key_t key_a, key_b;
    int count;
  } tuned[2];
};

/* ---- codec (blosc_codec.c) ---------------------------------------------------- */
typedef struct {
  int version, flags, typesize, header_len, memcpyed, dont_split, compformat, special;
  int32_t nbytes, blocksize, cbytes, leftover;
  int nblocks;
  uint8_t filters[6], filters_meta[6];
  int blosc2_flags;
} iax_header_t;

int iax_parse_header(const uint8_t *chunk, int32_t csize, iax_header_t *h);
int iax_chunk_is_zeros(const uint8_t *chunk, int32_t csize);
int iax_block_is_zeros(const uint8_t *chunk, int32_t csize, int nb);
int iax_lz4_compress(const uint8_t *src, int n, uint8_t *dst, int cap, int accel, uint32_t *hashtab);
void iax_shuffle(int ts, int32_t bsize, const uint8_t *src, uint8_t *dst);
void iax_bitshuffle(int ts, int32_t bsize, const uint8_t *src, uint8_t *dst);
/* Encode one chunk.  `data` holds nbytes of chunk content; when `filtered` is
 * non-zero every block of `data` already went through the chunk's filter (the
 * GPU pack stage did it).  Returns a malloc'ed chunk in *out and its size. */
int iax_encode_chunk(const uint8_t *data, int32_t nbytes, int32_t blocksize, int typesize,
                     const iarray_config_t *cfg, int filtered, uint8_t **out, int32_t *csize);
int iax_filter_code(const iarray_config_t *cfg);  /* IAMX_FILTER_* the config asks for */
int iax_split(const iarray_config_t *cfg, int typesize, int32_t blocksize);

/* ---- host parallel-for (threadpool.c) ----------------------------------------------- */
typedef void (*iax_task_fn)(void *ctx, int64_t i, int worker);
int iax_nthreads(const iarray_config_t *cfg);
void iax_parallel_for(int nthreads, int64_t n, iax_task_fn fn, void *ctx);
double iax_now(void);
/* Order LZ4 decode jobs by compressed size, largest first. */
void iax_sort_lz4_jobs(iamx_lz4_job_t *jobs, int64_t n);
/* Download `total` bytes of d_src into the pinned staging buffer in pieces
 * and run fn(ctx, task, worker) for each of `ntasks` tasks as soon as the
 * piece holding its bytes (up to task_end[task]) has landed, so the copy of
 * later pieces overlaps the work on earlier ones.  *host receives the
 * staging buffer (owned by the device runtime). */
ina_rc_t iax_download_pipelined(const void *d_src, size_t total, const int64_t *task_end, int64_t ntasks,
                                int nthreads, iax_task_fn fn, void *ctx, uint8_t **host);

/* Device work that fills part of a buffer before it is downloaded: enqueue
 * (asynchronously) what produces the bytes of tasks [t0, t1) of panel k;
 * returns an iamx error code. */
typedef int (*iax_panel_fn)(void *ctx, int k, int64_t t0, int64_t t1);
/* Enqueue the computation of rows [row0, row0 + nrows) of a dense result. */
typedef int (*iax_rows_fn)(void *ctx, int64_t row0, int64_t nrows);

/* ---- containers (iarray_container.c) ---------------------------------------------------- */
ina_rc_t iax_container_new(iarray_context_t *ctx, iarray_dtshape_t *dtshape, iarray_storage_t *storage,
                           iarray_container_t **c);
ina_rc_t iax_ensure_device(iarray_container_t *c);
ina_rc_t iax_ensure_plan(iarray_container_t *c);
/* Decode the whole (base) container into a dense row-major device buffer of
 * shape (shape0, shape1) with leading dimension shape1. */
ina_rc_t iax_decode_dense(iarray_container_t *c, void *d_dst, int scratch_slot);
/* Decode a container or transposed view as an operand: dense logical-shape
 * row-major device buffer in workspace slot `slot` (uses `tmp_slot` for views). */
ina_rc_t iax_operand(iarray_container_t *c, int slot, int tmp_slot, void **d_out);
/* The dense image in workspace `slot` is about to be overwritten by other
 * code (slot < 0: all slots, e.g. when the device runtime shuts down). */
void iax_operand_forget(int slot);
ina_rc_t iax_check_decode_status(void);
uint8_t *iax_special_chunk(int kind, int32_t nbytes, int32_t blocksize, int typesize);
/* writes an on-disk container (storage.urlpath) through to its frame */
ina_rc_t iax_persist(iarray_container_t *c);
/* Locate the chunks of a lazily opened frame (a no-op for every other
 * container); every reader of chunks[] or csizes[] calls it first. */
ina_rc_t iax_materialize(iarray_container_t *c);
/* GEMV y = A x with A streamed through the device in panels of chunk rows
 * (A does not fit the device budget; see iax_gemv_should_stream).  dx, dy:
 * device x and y; d_xzero: opt_gemv's zero flags of x's blocks, or NULL. */
int iax_gemv_should_stream(iarray_container_t *a);
extern int iax_last_gemv_panels; /* panels of the last GEMV's streamed matrix, 0 if resident */
ina_rc_t iax_gemv_streamed(iarray_container_t *a, int dt, const void *dx, void *dy, const uint8_t *d_xzero);
/* Write C from a dense device buffer (logical shape, ld = shape1): GPU pack +
 * host LZ4 encode; fills c->chunks. */
ina_rc_t iax_write_from_dense(iarray_container_t *c, const void *d_dense);
/* The same, with C produced panel by panel: rows(rctx, row0, nrows) enqueues
 * the computation of each panel of about `panel_rows` rows (whole chunk rows,
 * at most 16 panels) just before that panel is packed and downloaded, so the
 * host encode of earlier panels overlaps the computation of later ones. */
ina_rc_t iax_write_from_dense_rows(iarray_container_t *c, const void *d_dense, int64_t panel_rows,
                                   iax_rows_fn rows, void *rctx);
int64_t iax_dtype_size(iarray_data_type_t t);

ina_rc_t iax_set_error(ina_rc_t rc, const char *fmt, ...);
ina_rc_t iax_dev_error(int code, const char *what);

extern double iax_last_encode_s;
/* product form of the output being written (keys the adaptive GPU share of
 * the C write); 0 outside a product */
extern int iax_write_form;
extern double iax_last_device_ms;
extern int64_t iax_last_gpu_enc_in, iax_last_gpu_enc_out; /* bytes the GPU encoder read and wrote in the last write */
extern int iax_last_slices;
extern double iax_last_slice_err;

#endif
