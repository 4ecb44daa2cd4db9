/*
 * iamx.h -- C ABI of libiamx.so, the HIP/gfx950 kernels behind libiarray.so.
 *
 * Plain pointers and sizes only (no HIP or torch types in any signature).
 * The host C library (libiarray.so) is the only intended caller; each entry
 * names the reference routine whose work it replaces.  All device work is
 * issued on one library-owned HIP stream (default priority) and every entry
 * returns 0 on success or a negative IAMX_E* code; iamx_last_error() gives
 * the message.
 */
#ifndef IAMX_H_
#define IAMX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IAMX_API __attribute__((visibility("default")))

enum {
  IAMX_OK = 0,
  IAMX_ENODEV = -1,   /* no HIP device / runtime failure at init */
  IAMX_EHIP = -2,     /* a HIP call failed */
  IAMX_EINVAL = -3,   /* bad argument (shape, size, pointer) */
  IAMX_EDATA = -4,    /* a compressed stream is malformed */
  IAMX_ENOMEM = -5,
};

/* ---- device, memory, stream ---------------------------------------------- */
IAMX_API int iamx_init(void);
IAMX_API const char *iamx_last_error(void);
IAMX_API int iamx_device_info(char *name, int namelen, int *cus, char *arch, int archlen);
IAMX_API void *iamx_malloc(size_t bytes);
IAMX_API void iamx_free(void *dptr);
IAMX_API int iamx_h2d(void *dst, const void *src, size_t bytes);
IAMX_API int iamx_d2h(void *dst, const void *src, size_t bytes);
IAMX_API int iamx_d2d(void *dst, const void *src, size_t bytes);
IAMX_API int iamx_memset(void *dst, int value, size_t bytes);
/* height rows of width bytes, device to device, on the current stream (the
 * rectangle of a sliced view out of its container's dense image,
 * src/iarray_container.c:628-736) */
IAMX_API int iamx_copy2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height);
IAMX_API int iamx_sync(void);
/* Pinned host staging buffer (grows, kept until iamx_shutdown) and
 * downloads into it: iamx_d2h_async queues the copy on a copy stream, after
 * all work queued so far on the current compute stream, and records event `event`
 * (0..IAMX_COPY_EVENTS-1) after it, so kernels queued later overlap the copy;
 * iamx_event_sync blocks the host until that event has completed. */
#define IAMX_COPY_EVENTS 64
IAMX_API void *iamx_host_buffer(size_t bytes);
/* Pinned, mapped, fine-grained host buffer (grows, kept until iamx_shutdown)
 * that kernels write directly through *dev: small results skip the copy.
 * Its bytes are complete once the stream has been synchronised. */
IAMX_API void *iamx_host_direct(size_t bytes, void **dev);
IAMX_API int iamx_d2h_async(void *dst, const void *src, size_t bytes, int event);
IAMX_API int iamx_event_sync(int event);
/* A download queued on the copy stream after the ones before it, without
 * waiting for compute work queued since the last iamx_d2h_async (whose data
 * it may therefore not read); records `event`. */
IAMX_API int iamx_d2h_chain(void *dst, const void *src, size_t bytes, int event);
/* A second compute stream (default priority).  iamx_side_wait makes it wait
 * for the work queued so far on the library stream; between iamx_use_stream(1)
 * and iamx_use_stream(0) launches and the ready point of iamx_d2h_async go to
 * it, so that work overlaps what the library stream runs next. */
IAMX_API int iamx_side_wait(void);
IAMX_API int iamx_use_stream(int side);
/* Streamed uploads, for operands kept off the device between uses (larger
 * than the device budget): two stages, each a pinned host buffer.
 * iamx_upload_host waits for the stage's previous upload to finish and
 * returns its buffer (grown to `bytes`); iamx_upload copies `bytes` of it to
 * d_dst on an upload stream once the library stream has released the stage
 * (iamx_upload_release marks that point: after the work queued so far), and
 * makes the library stream wait for the copy.  No counterpart in the
 * reference, whose operands live in host memory or on disk. */
IAMX_API void *iamx_upload_host(int stage, size_t bytes);
IAMX_API int iamx_upload(int stage, void *d_dst, size_t bytes);
IAMX_API int iamx_upload_release(int stage);
/* hipMemGetInfo of the library's device */
IAMX_API int iamx_mem_info(size_t *free_bytes, size_t *total_bytes);
/* Scratch that lives until iamx_shutdown; grown (never shrunk) on demand.
 * Each slot is independent of the others. */
/* slots 0..7 and 12..14 belong to the host library, 8..9 to the Ozaki slice
 * tuner, 1 and 10..11 to the mixed-precision products (B planes of one-shot
 * products, B planes of the row-panel interface, planes of the current A
 * panel), 16 to the plain GEMM's split-K partial tiles, 17 to hipBLASLt */
#define IAMX_WORKSPACES 18
#define IAMX_BLT_WORKSPACE 17
IAMX_API void *iamx_workspace(int slot, size_t bytes);
IAMX_API void iamx_shutdown(void);

/* ---- profiling (HIP events around every launch on the library stream) ---- */
IAMX_API void iamx_prof_enable(int on);
IAMX_API void iamx_prof_reset(void);
IAMX_API int iamx_prof_entry(int i, const char **name, int64_t *launches, double *total_ms);
IAMX_API double iamx_prof_span_ms(void); /* first start .. last stop since reset */

/* ---- Blosc2 block decode --------------------------------------------------
 * Replaces the decode inside blosc2_getitem_ctx (src/linalg/iarray_gemm.c:259,
 * src/linalg/iarray_gemv.c:194) and caterva_get_slice_buffer
 * (src/iarray_container.c:704,730): LZ4 streams are decoded by one wavefront
 * each into scratch, then every block is un-shuffled / bit-un-shuffled and
 * written into its tile of a dense row-major destination (edge-trimmed). */
typedef struct iamx_lz4_job_s {
  uint64_t src;    /* device address of the compressed payload */
  int64_t dst_off; /* byte offset in scratch of the decoded stream */
  int32_t csize;   /* compressed bytes */
  int32_t nbytes;  /* decoded bytes (the stream's neblock) */
  int32_t codec;   /* IAMX_CODEC_LZ4 (LZ4 and LZ4HC streams), IAMX_CODEC_BLOSCLZ or IAMX_CODEC_ZSTD */
  int32_t reserved;
} iamx_lz4_job_t;
/* The stream formats of a job (a chunk's codec, bits 5-7 of its flags: 1, 0
 * or 4).  Zstd streams are decoded by a pass of their own that the decode
 * entry points run once iamx_enable_zstd(1) has been called (the host library
 * calls it when it plans a zstd stream); before that a zstd job flags
 * IAMX_STATUS_CODEC. */
enum { IAMX_CODEC_LZ4 = 0, IAMX_CODEC_BLOSCLZ = 1, IAMX_CODEC_ZSTD = 2 };
IAMX_API void iamx_enable_zstd(int on);

#define IAMX_MAX_STREAMS 8
enum { IAMX_SRC_CONST = 0, IAMX_SRC_ADDR = 1, IAMX_SRC_SCRATCH = 2 };
enum { IAMX_FILTER_NONE = 0, IAMX_FILTER_SHUFFLE = 1, IAMX_FILTER_BITSHUFFLE = 2, IAMX_FILTER_BITSHUFFLE_V2 = 3 };
/* OR'ed into a block job's filter: the tile is written transposed, element
 * (r, c) of the tile at dst_off + c * dst_ld + r (the decoded image of a
 * transposed view, written in place of the omatcopy of
 * src/iarray_container.c:718-723). */
#define IAMX_JOB_TRANSPOSE 0x100

typedef struct iamx_block_job_s {
  uint64_t src[IAMX_MAX_STREAMS]; /* device address or scratch offset of stream s */
  uint8_t kind[IAMX_MAX_STREAMS]; /* IAMX_SRC_* */
  uint8_t fill[IAMX_MAX_STREAMS]; /* byte of a constant stream */
  int32_t neblock;                /* bytes per stream */
  int16_t nstreams;
  int16_t typesize;
  int32_t filter;                 /* IAMX_FILTER_* */
  int32_t bsize;                  /* bytes of the block */
  int32_t rows, cols;             /* block tile shape in elements (row-major) */
  int32_t vrows, vcols;           /* valid extent inside the tile */
  int64_t dst_off;                /* element offset of the tile origin in dst */
  int64_t dst_ld;                 /* destination leading dimension (elements) */
} iamx_block_job_t;

IAMX_API int iamx_decode(const iamx_lz4_job_t *d_lz4, int n_lz4, const iamx_block_job_t *d_blocks, int n_blocks,
                         void *d_scratch, void *d_dst);

/* ---- GEMM / GEMV -------------------------------------------------------------
 * Replaces the per-C-block cblas_?gemm of the gemm branch
 * (src/linalg/matmul/gemm.c:96-102) and of iarray_opt_gemm
 * (src/linalg/iarray_gemm.c:276-294): one launch computes the whole dense
 * C(MxN) = A(MxK) B(KxN), row-major.  dtype 0 = float64, 1 = float32.
 * precision: 0 plain, 1 split-bf16 (float32), 2 ozaki-int8 (float64, `slices`
 * INT8 slices), 3 ozaki-crt (float64, `slices` = number of INT8 moduli,
 * 2..15). */
/* Smallest Ozaki slice count S in [smin, 8] whose product reaches a
 * componentwise error max |C_S - C| / (|A||B|) <= target on a sample of up
 * to 256 evenly spaced rows of A, measured against the plain FP64 MFMA product
 * of the same rows (S = 8 if none does).  *err receives the sampled error of
 * the chosen S.  Replaces nothing in the reference (its FP64 path is MKL). */
IAMX_API int iamx_ozaki_pick_slices(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B,
                                    int64_t ldb, double target, int smin, int *slices, double *err);
/* The same pick for the CRT form: the smallest number of moduli in
 * [nmin, 15] reaching `target` on the sampled rows (15 if none does). */
IAMX_API int iamx_ozaki_crt_pick_moduli(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda,
                                        const double *B, int64_t ldb, double target, int nmin, int *moduli,
                                        double *err);
IAMX_API int iamx_gemm(int dtype, int precision, int slices, int64_t M, int64_t N, int64_t K, const void *A,
                       int64_t lda, const void *B, int64_t ldb, void *C, int64_t ldc);

/* Row-panel form of iamx_gemm for callers that write C out panel by panel
 * while later panels are still being computed (the reference computes and
 * compresses C one chunk at a time, src/linalg/iarray_gemm.c:399-425).
 * iamx_gemm_prepare_b fixes B (K x N) and, for the mixed precisions, builds
 * its split or sliced planes once; each iamx_gemm_rows call then computes
 * C[0:M, :] = A[0:M, :] B, with A and C pointing at the panel's first row.
 * B must stay resident and unchanged until the last iamx_gemm_rows call. */
IAMX_API int iamx_gemm_prepare_b(int dtype, int precision, int slices, int64_t N, int64_t K, const void *B,
                                 int64_t ldb);
IAMX_API int iamx_gemm_rows(int64_t M, const void *A, int64_t lda, void *C, int64_t ldc);

/* The block-pair product of iarray_opt_gemm / _gemm_a / _gemm_b when a
 * zero block is present (src/linalg/iarray_gemm.c:143-165,
 * src/linalg/iarray_gemm_a.c:100-175): C block (gr, gc) = sum over K blocks
 * gk, ascending, of A block (gr, gk) x B block (gk, gc), skipping the pairs
 * with d_zero_a[gr * nbk + gk] or d_zero_b[gk * nbc + gc] set, so Inf/NaN
 * facing a zero block never reaches C.  d_rows / d_ks / d_cols hold [first,
 * end) of every block row, K block and block column (int32 pairs, device
 * memory); only block rows [br0, br1) are computed; max_rows / max_cols bound
 * one block's extent.  A, B, C dense row-major. */
IAMX_API int iamx_gemm_blockpairs(int dtype, const void *A, int64_t lda, const void *B, int64_t ldb, void *C,
                                  int64_t ldc, const int32_t *d_rows, int nbr, const int32_t *d_ks, int nbk,
                                  const int32_t *d_cols, int nbc, const uint8_t *d_zero_a, const uint8_t *d_zero_b,
                                  int br0, int br1, int max_rows, int max_cols);

/* Replaces the per-C-block cblas_?gemv of the gemv branch
 * (src/linalg/matmul/gemv.c:92-98) and of iarray_opt_gemv
 * (src/linalg/iarray_gemv.c:212-224): y(M) = A(MxK) x(K). */
IAMX_API int iamx_gemv(int dtype, int64_t M, int64_t K, const void *A, int64_t lda, const void *x, void *y);

/* GEMV fused with the block decode (the opt_gemv shape of the work,
 * src/linalg/iarray_gemv.c:156-229): A's LZ4 streams are decoded into
 * scratch, then each block is un-filtered on chip and multiplied with its x
 * segment without ever writing dense A; row sums per block column land in
 * `partial` (nbc x M) and are added in block-column order into y.
 * d_blk_rc holds 4 int32 per block job: first row and first column of the
 * tile in A, the tile's column index in the global block grid (partial row),
 * and flags (bit 0: a zero block, used by iamx_opt_gemv_fused only); nbc is the number of block columns; bs0/bs1 are the block shape
 * (bs1 <= 1024). */
IAMX_API int iamx_gemv_fused(int dtype, const iamx_lz4_job_t *d_lz4, int n_lz4, const iamx_block_job_t *d_blocks,
                             const int32_t *d_blk_rc, int n_blocks, int nbc, int bs0, int bs1, const void *x,
                             void *y, int64_t M, void *d_scratch, void *d_partial);

/* iamx_gemv_fused for iarray_opt_gemv, which skips zero blocks
 * (src/linalg/iarray_gemv.c:125-160): a block job whose d_blk_rc flag word
 * has bit 0 set (a zero block of A), or whose block column c has d_xzero[c]
 * set (a zero chunk or block of x), contributes nothing, so an Inf or NaN
 * facing a zero block never reaches y.  d_partial must start at zero. */
IAMX_API int iamx_opt_gemv_fused(int dtype, const iamx_lz4_job_t *d_lz4, int n_lz4, const iamx_block_job_t *d_blocks,
                                 const int32_t *d_blk_rc, int n_blocks, int nbc, int bs0, int bs1, const void *x,
                                 void *y, int64_t M, void *d_scratch, void *d_partial, const uint8_t *d_xzero);

/* iamx_gemv_fused without any allocation: the LZ4 scratch and the partial
 * sums live in a fixed device array of iamx_gemv_fused_noalloc_limit() bytes
 * (64 MiB; decoded stream bytes rounded up to 256, then nbc x M partial sums),
 * resolved once per device at iamx_init, which must have run.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  `scratch_bytes` is the extent the job table's decoded
 * streams need (max dst_off + nbytes, padded to 16); K is the length of x.
 * Requests above the limit fail with IAMX_EINVAL before any launch; job
 * entries outside these sizes are skipped on the device and flag
 * IAMX_STATUS_RANGE.  Every call on one GPU shares that GPU's array and
 * leaves its contents behind: calls that overlap on one GPU (other streams,
 * other threads) race. */
IAMX_API int64_t iamx_gemv_fused_noalloc_limit(void);
IAMX_API int iamx_gemv_fused_noalloc(int dtype, const iamx_lz4_job_t *d_lz4, int n_lz4,
                                     const iamx_block_job_t *d_blocks, const int32_t *d_blk_rc, int n_blocks, int nbc,
                                     int bs0, int bs1, const void *x, int64_t K, void *y, int64_t M,
                                     int64_t scratch_bytes);

/* ---- C write stage --------------------------------------------------------------
 * Replaces the shuffle/bitshuffle that blosc2_compress_ctx applies to every C
 * block before LZ4 (src/linalg/matmul/gemm.c:216-230): gathers each C block
 * tile from the dense C (zero padding outside the valid extent) and writes the
 * filtered block bytes chunk-major, block-major into `out`. */
typedef struct iamx_tiling_s {
  int64_t shape[2];
  int64_t chunk[2];
  int64_t block[2];
  int64_t nchunks[2]; /* chunks per axis (extshape / chunkshape) */
  int64_t nblocks[2]; /* blocks per chunk per axis (extchunkshape / blockshape) */
  int32_t typesize;
  int32_t filter; /* IAMX_FILTER_* */
} iamx_tiling_t;

IAMX_API int iamx_pack_blocks(const iamx_tiling_t *tiling, const void *d_src, int64_t src_ld, void *d_out);
/* The same for chunks [chunk_begin, chunk_end) of the chunk grid (row-major
 * chunk order), written at their place in `out`; d_src is still the whole
 * dense C, and only the rows of those chunks are read. */
IAMX_API int iamx_pack_blocks_range(const iamx_tiling_t *tiling, const void *d_src, int64_t src_ld, void *d_out,
                                    int64_t chunk_begin, int64_t chunk_end);

/* C write stage on the GPU: LZ4 encode and Blosc2 chunk assembly.  Replaces
 * the shuffle + LZ4 encode + chunk layout of blosc2_compress_ctx on every C
 * chunk (src/linalg/matmul/gemm.c:216-243, src/linalg/iarray_gemm.c:399-425);
 * the chunk layout is the one the host writer produces (format version 5,
 * 32-byte header with the filter in slot 5, bstarts, per block `nstreams`
 * streams of int32 csize + bytes: csize 0 = zero stream, = neblock = raw,
 * otherwise one LZ4 block), nstreams = typesize when `split`, else 1.
 * Input: the packed, filtered blocks of iamx_pack_blocks(_range).  Chunks
 * [chunk_begin, chunk_end) are encoded; chunk c is written at c * capacity
 * of d_out and its size stored in d_chunk_size[c].  d_scratch holds
 * iamx_encode_scratch_bytes bytes; d_meta 2 int32 per stream of the tiling
 * (chunk-major, block-major, stream-minor).  The LZ4 parse is this library's
 * own (encoded bytes are not those of c-blosc2, which nothing pins): every
 * 4 KiB of a stream is parsed on its own, matches staying inside it, and the
 * pieces are joined into one LZ4 block; every stream decodes to its input. */
IAMX_API int64_t iamx_chunk_capacity(const iamx_tiling_t *tiling, int split);
IAMX_API int64_t iamx_encode_scratch_bytes(const iamx_tiling_t *tiling, int split);
IAMX_API int iamx_encode_chunks(const iamx_tiling_t *tiling, int split, const void *d_packed, int64_t chunk_begin,
                                int64_t chunk_end, void *d_scratch, int32_t *d_meta, void *d_out,
                                int64_t *d_chunk_size);
/* The same stage for the outputs of many products at once (the C write of a
 * batch of iarray_linalg_matmul calls, src/linalg/matmul/gemm.c:216-243 per
 * product): one launch of each encoder kernel covers every chunk of the n
 * tilings.  Tiling i's packed blocks start at d_packed + packed_off[i]; its
 * chunk c is written at d_out + out_off[i] + c * iamx_chunk_capacity(tiling i)
 * (out_off 16-byte aligned) and its size stored in d_chunk_size[first + c],
 * first = the chunk count of tilings 0..i-1.  d_scratch holds
 * iamx_encode_batch_scratch_bytes bytes; the segment table is copied into it
 * synchronously before the kernels are queued on the current stream. */
IAMX_API int64_t iamx_encode_batch_scratch_bytes(int n, const iamx_tiling_t *tilings, const int *split);
IAMX_API int iamx_encode_batch(int n, const iamx_tiling_t *tilings, const int *split, const void *d_packed,
                               const int64_t *packed_off, void *d_scratch, void *d_out, const int64_t *out_off,
                               int64_t *d_chunk_size);

/* Transpose a dense row-major (rows x cols) matrix: the omatcopy of a
 * transposed view (src/iarray_container.c:718-723). */
IAMX_API int iamx_transpose(int typesize, int64_t rows, int64_t cols, const void *d_src, void *d_dst);

/* ---- many products in one pass ------------------------------------------------------
 * Replaces the reference's one-call-per-product loop (each call re-decoding its
 * panels, src/linalg/matmul.c:19-68): all tiles of all products are pulled
 * from one device work queue by a persistent kernel. Offsets are in elements
 * of the dtype from base_a / base_b / base_c (dense row-major, ld = K, N, N). */
typedef struct iamx_batch_prod_s {
  int64_t M, N, K;
  int64_t a_off, b_off, c_off;
} iamx_batch_prod_t;

/* d_tiles: 3 int32 per tile (product, tile row, tile column), 128x128 tiles. */
IAMX_API int iamx_gemm_batch(int dtype, const iamx_batch_prod_t *d_prods, const int32_t *d_tiles, int ntiles,
                             const void *base_a, const void *base_b, void *base_c);

typedef struct iamx_batch_pack_s {
  iamx_tiling_t tiling;
  int64_t c_off_bytes; /* product's dense C inside base_c */
  int64_t out_off;     /* product's packed blocks inside out */
} iamx_batch_pack_t;

/* d_blocks: 2 int32 per C block (product, block index inside the product). */
IAMX_API int iamx_pack_batch(const iamx_batch_pack_t *d_packs, const int32_t *d_blocks, int nblocks,
                             const void *base_c, void *d_out);

/* ---- HIP graphs of the library stream ------------------------------------------------
 * Everything the library launches between begin and end is recorded (not
 * run) and instantiated into an executable graph; launching it replays the
 * same kernels on the same buffers.  Profiling events are not recorded while
 * capturing; a replay is timed as one "graph_launch" entry. */
IAMX_API int iamx_capture_begin(void);
IAMX_API int iamx_capture_end(void **graph_exec);
IAMX_API int iamx_graph_launch(void *graph_exec);
IAMX_API void iamx_graph_destroy(void *graph_exec);

/* Error word written by the decode kernels (0 = no error since last reset).
 * iamx_status_snapshot queues its copy to the host behind the work queued so
 * far; iamx_decode_status waits for the stream and returns it (queueing the
 * copy itself when kernels were launched after the last snapshot) and resets
 * a non-zero word. */
IAMX_API int iamx_status_snapshot(void);
IAMX_API int iamx_decode_status(void);
/* Status bit of a job table entry that points outside the buffers it was
 * given (the job is skipped); bits 1..7 are malformed LZ4 streams. */
#define IAMX_STATUS_RANGE (1 << 8)
/* Status bit of a malformed BloscLZ stream. */
#define IAMX_STATUS_BLOSCLZ (1 << 9)
/* Status bit of a malformed zstd stream. */
#define IAMX_STATUS_ZSTD (1 << 10)
/* Status bit of a zstd job met while zstd decoding is off, or of a job whose
 * codec has no decoder here (a zlib stream, an unknown value). */
#define IAMX_STATUS_CODEC (1 << 11)

#ifdef __cplusplus
}
#endif
#endif
