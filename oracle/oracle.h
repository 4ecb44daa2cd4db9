/*
 * oracle.h -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A plain-C CPU restatement of ironArray's compressed chunk-blocked matmul path:
 * the Blosc2 chunk/block decode stage (LZ4 + byte/bit unshuffle), the Caterva
 * two-level tiling, and the four linalg algorithms (gemm_blosc, gemv_blosc,
 * iarray_opt_gemm, iarray_opt_gemv).  It exists only to CHECK the HIP product:
 * the only callers allowed are tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  The product libraries (libiarray.so,
 * libiamx.so) never link, load or call anything in this directory.
 *
 * Parity pins (each function below names the checks that pin it):
 *   - orc_lz4_decode ........ liblz4 1.9.3 LZ4_compress_default outputs (black box),
 *                              known-answer SHA-256s of the reference's frame fixtures.
 *   - orc_blosclz_decode .... c-blosc 1.21.0 chunks written with compressor "blosclz"
 *                              (black box): literal runs, short and long matches, the
 *                              far-distance escape, split and unsplit streams; c-blosc2's
 *                              own BloscLZ output is not in the image (parity unpinned
 *                              beyond the shared stream format).
 *   - orc_zstd_decode ....... libzstd 1.4.8 ZSTD_compress outputs (black box) at levels
 *                              -5..22 over random, low-entropy, float, run and structured
 *                              data (raw / RLE / compressed blocks, multi-block frames,
 *                              1- and 4-stream Huffman literals, all FSE table modes,
 *                              repeat offsets), and c-blosc 1.21.0 chunks written with
 *                              compressor "zstd" (black box).
 *   - orc_zlib_decode ....... zlib compress2 outputs (Python's zlib module, black box) at
 *                              levels 0-9 over random, low-entropy, float, run and
 *                              structured data (stored, fixed and dynamic blocks), and
 *                              c-blosc 1.21.0 chunks written with compressor "zlib".
 *   - orc_chunk_* / getitem .. the reference's tests/data/<name>.iarray frames (c-blosc2
 *                              writes them; SHA-256 of every decoded array is fixed),
 *                              c-blosc 1.21.0 chunks (blosc_compress, black box).
 *   - orc_unshuffle ......... c-blosc 1.21 byte shuffle (raw-stored streams), fixtures.
 *   - orc_bitunshuffle ...... c-blosc 1.21 bitshuffle for element counts that are a
 *                              multiple of 8 and the version-2 whole-block-copy rule;
 *                              the format>=3 leftover rule (copy the nelem%8 tail) is
 *                              restated from c-blosc2's published bitunshuffle() and is
 *                              pinned only by round trips ("parity unpinned" for
 *                              c-blosc2-written bitshuffled blocks with a ragged tail).
 *   - orc_stream run-length form (negative csize) ... restated from c-blosc2 blosc_d();
 *                              no fixture holds one: parity unpinned beyond round trips.
 *   - orc_array_to_dense .... closed-form tiling (Caterva, restated from
 *                              src/iarray_zproxy.c:46-85 and src/linalg/matmul/gemm.c:25-69),
 *                              brute-force checked on tiny shapes in tests.
 *   - orc_gemm_blosc / orc_gemv_blosc / orc_opt_gemm / orc_opt_gemv / orc_opt_gemm_a /
 *     orc_opt_gemm_b ......... integer-valued inputs (exact in any summation order)
 *                              against numpy's dense product: ragged tilings with
 *                              all-padding blocks, transposed operands, planted special
 *                              zero chunks and zero blocks; the opt paths' zero skipping
 *                              by an Inf that meets only those zeros (tests/
 *                              test_oracle_linalg.py, tests/test_host_cpu.py).
 *   - orc_dd_gemm_sample ..... exact rational dot products (fractions.Fraction) on
 *                              wide-range and cancelling data: within half an ulp of
 *                              the exact sum (tests/test_oracle_linalg.py).
 */
#ifndef IAMX_ORACLE_H
#define IAMX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- LZ4 block (raw, no frame) ---------------------------------------- */
int orc_lz4_decode(const uint8_t *src, int srclen, uint8_t *dst, int dstcap);
/* ---- BloscLZ stream (blosclz_block.c) ------------------------------------ */
int orc_blosclz_decode(const uint8_t *src, int srclen, uint8_t *dst, int dstcap);

/* ---- Zstd frames (zstd_block.c) -------------------------------------------- */
int orc_zstd_decode(const uint8_t *src, int srclen, uint8_t *dst, int dstcap);

/* ---- zlib streams (deflate_block.c) ------------------------------------------ */
int orc_zlib_decode(const uint8_t *src, int srclen, uint8_t *dst, int dstcap);

/* ---- Blosc chunk (blosc2 extended header v3..v5, and blosc1 v2) -------- */
typedef struct {
  int version;        /* byte 0 */
  int versionlz;      /* byte 1 */
  int flags;          /* byte 2 */
  int typesize;       /* byte 3 */
  int32_t nbytes;     /* bytes 4..7  */
  int32_t blocksize;  /* bytes 8..11 */
  int32_t cbytes;     /* bytes 12..15 */
  int header_len;     /* 32 (extended) or 16 (blosc1) */
  int extended;       /* flags & 0x5 == 0x5 */
  int memcpyed;       /* flags & 0x2 */
  int dont_split;     /* flags & 0x10 */
  int compformat;     /* flags >> 5 : 1 == LZ4 */
  uint8_t filters[6];
  uint8_t filters_meta[6];
  int blosc2_flags;   /* byte 31 (extended only) */
  int special;        /* (blosc2_flags >> 4) & 7 */
  int nblocks;
  int32_t leftover;
} orc_chunk_info_t;

int orc_chunk_info(const uint8_t *chunk, int32_t csize, orc_chunk_info_t *info);
/* Decode block `nblock` (bsize bytes, bsize = blocksize or the leftover) into dst. */
int orc_decode_block(const uint8_t *chunk, int32_t csize, int nblock, uint8_t *dst);
/* Whole chunk; returns nbytes or < 0 on error. */
int orc_decompress_chunk(const uint8_t *chunk, int32_t csize, uint8_t *dst, int64_t dstcap);
/* Restates blosc2_getitem_ctx: items [start, start+nitems) of the chunk. */
int orc_getitem(const uint8_t *chunk, int32_t csize, int start, int nitems, uint8_t *dst);
/* Reference's chunk_is_zeros/block_is_zeros (src/linalg/iarray_gemm.c:17-77),
 * restated with the correct split test (dont_split bit, not flags & 0x4). */
int orc_chunk_is_zeros(const uint8_t *chunk);
int orc_block_is_zeros(const uint8_t *chunk, int32_t csize, int nblock);

/* ---- filters ---------------------------------------------------------- */
void orc_shuffle(int typesize, int32_t bsize, const uint8_t *src, uint8_t *dst);
void orc_unshuffle(int typesize, int32_t bsize, const uint8_t *src, uint8_t *dst);
void orc_bitshuffle(int typesize, int32_t bsize, const uint8_t *src, uint8_t *dst, int format_version);
void orc_bitunshuffle(int typesize, int32_t bsize, const uint8_t *src, uint8_t *dst, int format_version);

/* ---- Caterva 2-level tiled array (ndim 1 or 2) ------------------------- */
typedef struct {
  int ndim;
  int itemsize;
  int64_t shape[2];
  int64_t chunkshape[2];
  int64_t blockshape[2];
  int64_t nchunks;
  const uint8_t *const *chunks;  /* nchunks compressed chunks, row-major chunk order */
  const int32_t *csizes;
  int transposed;                /* a transposed 2-D view (iarray_linalg_transpose) */
} orc_array_t;

int orc_array_to_dense(const orc_array_t *a, void *dst, int nthreads);
/* Restates _iarray_get_slice_buffer (src/iarray_container.c:628-736) for a
 * non-view or transposed 2-D container: region [start, stop) written into a
 * row-major buffer whose shape is bufshape (>= stop-start per axis). */
int orc_get_slice(const orc_array_t *a, const int64_t *start, const int64_t *stop,
                  const int64_t *bufshape, void *buf, int nthreads);

/* ---- BLAS / LZ4 backends ------------------------------------------------ */
/* 0 = plain C loops (deterministic, k-ascending), 1 = MKL via dlopen. Returns
 * the backend actually selected. */
int orc_set_blas(int which);
const char *orc_blas_name(void);
/* 0 = the restated decoder above, 1 = liblz4 LZ4_decompress_safe via dlopen. */
int orc_set_lz4(int which);

/* ---- linalg restatements (outputs: dense row-major C of the logical shape) */
int orc_gemm_blosc(const orc_array_t *a, const orc_array_t *b, const int64_t *c_chunkshape,
                   const int64_t *c_blockshape, void *c, int nthreads);
int orc_gemv_blosc(const orc_array_t *a, const orc_array_t *b, const int64_t *c_chunkshape,
                   const int64_t *c_blockshape, void *c, int nthreads);
int orc_opt_gemm_a(const orc_array_t *a, const orc_array_t *b, const int64_t *c_chunkshape,
                   const int64_t *c_blockshape, void *c, int nthreads);
int orc_opt_gemm_b(const orc_array_t *a, const orc_array_t *b, const int64_t *c_chunkshape,
                   const int64_t *c_blockshape, void *c, int nthreads);
int orc_opt_gemm(const orc_array_t *a, const orc_array_t *b, const int64_t *c_chunkshape,
                 const int64_t *c_blockshape, void *c, int nthreads);
int orc_opt_gemv(const orc_array_t *a, const orc_array_t *b, const int64_t *c_chunkshape,
                 const int64_t *c_blockshape, void *c, int nthreads);

/* The reference's C write stage for the CPU baseline: shuffle + LZ4-compress the
 * blocks of every C chunk (c-blosc2 calls liblz4 for this).  Returns total
 * compressed bytes, or -1 when liblz4 is not loadable. */
int64_t orc_encode_dense(const void *c, int ndim, const int64_t *shape, int itemsize,
                         const int64_t *chunkshape, const int64_t *blockshape,
                         int clevel, int filter, int nthreads);

/* Error-metric helpers in long double / compensated sums. */
void orc_dd_gemm_sample(const double *a, const double *b, int64_t M, int64_t N, int64_t K,
                        const int64_t *rows, const int64_t *cols, int64_t nsample, double *out);

#ifdef __cplusplus
}
#endif
#endif
