/*
 * iarray.h -- drop-in subset of the ironArray C API for the compressed,
 * chunk-blocked matmul path, implemented MI355X-native (libiarray.so, host C
 * over libiamx.so HIP kernels for gfx950).
 *
 * Every declaration mirrors the reference's public header
 * include/libiarray/iarray.h (cited per item) with the same names, argument
 * meaning, ownership and error behaviour.  libinac is not used, so the pieces
 * of its RC convention the API exposes (ina_rc_t, INA_ERROR,
 * INA_SUCCEED/INA_FAILED) are re-declared here; callers compare against the
 * IARRAY_ERR_* symbols, not bit patterns.
 */
#ifndef IAMX_IARRAY_H_
#define IAMX_IARRAY_H_

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define INA_API(x) __attribute__((visibility("default"))) x

/* ---- return codes (libinac convention, doc/inac/error.md:155-185) -------- */
typedef int64_t ina_rc_t;
#define INA_SUCCESS ((ina_rc_t)0)
#define INA_ERR_BIT ((uint64_t)1 << 63)
#define INA_ERROR(e) ((ina_rc_t)(INA_ERR_BIT | (uint64_t)(e)))
#define INA_FAILED(rc) (((uint64_t)(rc) & INA_ERR_BIT) != 0)
#define INA_SUCCEED(rc) (!INA_FAILED(rc))
#define INA_RC_ERROR(rc) ((uint64_t)(rc) & ~INA_ERR_BIT)

/* error kinds (high 32 bits) and subjects (low 32 bits) */
#define INA_ERR_INVALID ((uint64_t)0x01 << 32)
#define INA_ERR_FAILED ((uint64_t)0x02 << 32)
#define INA_ERR_EMPTY ((uint64_t)0x03 << 32)
#define INA_ERR_FULL ((uint64_t)0x04 << 32)
#define INA_ERR_TOO_SMALL ((uint64_t)0x05 << 32)
#define INA_ERR_NOT_SUPPORTED ((uint64_t)0x06 << 32)
#define INA_ERR_NULL_POINTER ((uint64_t)0x07 << 32)
#define INA_ERR_INVALID_ARGUMENT ((uint64_t)0x08 << 32)
#define INA_ERR_OUT_OF_MEMORY ((uint64_t)0x09 << 32)
#define INA_ES_USER_DEFINED ((uint64_t)0x100)
/* libinac's file subject and the error _iarray_container_load returns for a
 * path that does not exist (src/iarray_container.c:86-89) */
#define INA_ES_FILE ((uint64_t)0x05)
#define INA_ERR_FILE_OPEN (INA_ERR_FAILED | INA_ES_FILE)

/* subjects, include/libiarray/iarray.h:28-50 */
#define IARRAY_ES_CONTAINER (INA_ES_USER_DEFINED + 1)
#define IARRAY_ES_DTSHAPE (INA_ES_USER_DEFINED + 2)
#define IARRAY_ES_SHAPE (INA_ES_USER_DEFINED + 3)
#define IARRAY_ES_CHUNKSHAPE (INA_ES_USER_DEFINED + 4)
#define IARRAY_ES_NDIM (INA_ES_USER_DEFINED + 5)
#define IARRAY_ES_DTYPE (INA_ES_USER_DEFINED + 6)
#define IARRAY_ES_STORAGE (INA_ES_USER_DEFINED + 7)
#define IARRAY_ES_PERSISTENCY (INA_ES_USER_DEFINED + 8)
#define IARRAY_ES_BUFFER (INA_ES_USER_DEFINED + 9)
#define IARRAY_ES_CATERVA (INA_ES_USER_DEFINED + 10)
#define IARRAY_ES_BLOSC (INA_ES_USER_DEFINED + 11)
#define IARRAY_ES_ASSERTION (INA_ES_USER_DEFINED + 12)
#define IARRAY_ES_BLOCKSHAPE (INA_ES_USER_DEFINED + 13)
/* error subject of the GPU (HIP runtime / kernels) */
#define IARRAY_ES_DEVICE (INA_ES_USER_DEFINED + 64)

/* errors, include/libiarray/iarray.h:53-90 */
#define IARRAY_ERR_EMPTY_CONTAINER (INA_ERR_EMPTY | IARRAY_ES_CONTAINER)
#define IARRAY_ERR_INVALID_DTSHAPE (INA_ERR_INVALID | IARRAY_ES_DTSHAPE)
#define IARRAY_ERR_INVALID_DTYPE (INA_ERR_INVALID | IARRAY_ES_DTYPE)
#define IARRAY_ERR_INVALID_SHAPE (INA_ERR_INVALID | IARRAY_ES_SHAPE)
#define IARRAY_ERR_INVALID_CHUNKSHAPE (INA_ERR_INVALID | IARRAY_ES_CHUNKSHAPE)
#define IARRAY_ERR_INVALID_BLOCKSHAPE (INA_ERR_INVALID | IARRAY_ES_BLOCKSHAPE)
#define IARRAY_ERR_INVALID_NDIM (INA_ERR_INVALID | IARRAY_ES_NDIM)
#define IARRAY_ERR_INVALID_STORAGE (INA_ERR_INVALID | IARRAY_ES_STORAGE)
#define IARRAY_ERR_TOO_SMALL_BUFFER (INA_ERR_TOO_SMALL | IARRAY_ES_BUFFER)
#define IARRAY_ERR_CATERVA_FAILED (INA_ERR_FAILED | IARRAY_ES_CATERVA)
#define IARRAY_ERR_BLOSC_FAILED (INA_ERR_FAILED | IARRAY_ES_BLOSC)
#define IARRAY_ERR_ASSERTION_FAILED (INA_ERR_FAILED | IARRAY_ES_ASSERTION)
#define IARRAY_ERR_DEVICE_FAILED (INA_ERR_FAILED | IARRAY_ES_DEVICE)
#define IARRAY_ERR_NO_DEVICE (INA_ERR_NOT_SUPPORTED | IARRAY_ES_DEVICE)

#define IARRAY_DIMENSION_MAX 8 /* include/libiarray/iarray.h:20 */
#define IARRAY_METALAYER_VERSION 0

typedef struct iarray_context_s iarray_context_t;     /* iarray.h:108 */
typedef struct iarray_container_s iarray_container_t; /* iarray.h:109 */

/* include/libiarray/iarray.h:134-149 */
typedef enum iarray_data_type_e {
  IARRAY_DATA_TYPE_DOUBLE = 0,
  IARRAY_DATA_TYPE_FLOAT = 1,
  IARRAY_DATA_TYPE_INT64 = 10,
  IARRAY_DATA_TYPE_INT32 = 11,
  IARRAY_DATA_TYPE_INT16 = 12,
  IARRAY_DATA_TYPE_INT8 = 13,
  IARRAY_DATA_TYPE_UINT64 = 16,
  IARRAY_DATA_TYPE_UINT32 = 17,
  IARRAY_DATA_TYPE_UINT16 = 18,
  IARRAY_DATA_TYPE_UINT8 = 19,
  IARRAY_DATA_TYPE_BOOL = 24,
  IARRAY_DATA_TYPE_MAX
} iarray_data_type_t;

/* include/libiarray/iarray.h:164-169 */
typedef enum iarray_filter_flags_e {
  IARRAY_COMP_SHUFFLE = 0x1,
  IARRAY_COMP_BITSHUFFLE = 0x2,
  IARRAY_COMP_DELTA = 0x4,
  IARRAY_COMP_TRUNC_PREC = 0x8,
} iarray_filter_flags_t;

/* include/libiarray/iarray.h:177-187 */
typedef enum iarray_compression_codec_e {
  IARRAY_COMPRESSION_BLOSCLZ = 0,
  IARRAY_COMPRESSION_LZ4,
  IARRAY_COMPRESSION_LZ4HC,
  IARRAY_COMPRESSION_SNAPPY,
  IARRAY_COMPRESSION_ZLIB,
  IARRAY_COMPRESSION_ZSTD,
  IARRAY_COMPRESSION_ZFP_FIXED_ACCURACY,
  IARRAY_COMPRESSION_ZFP_FIXED_PRECISION,
  IARRAY_COMPRESSION_ZFP_FIXED_RATE
} iarray_compression_codec_t;

typedef enum iarray_compression_favor_e {
  IARRAY_COMPRESSION_FAVOR_BALANCE = 0,
  IARRAY_COMPRESSION_FAVOR_SPEED,
  IARRAY_COMPRESSION_FAVOR_CRATIO,
} iarray_compression_favor_t;

/* include/libiarray/iarray.h:195-200 */
typedef enum iarray_split_mode_e {
  IARRAY_ALWAYS_SPLIT = 1,
  IARRAY_NEVER_SPLIT = 2,
  IARRAY_AUTO_SPLIT = 3,
  IARRAY_FORWARD_COMPAT_SPLIT = 4,
} iarray_split_mode_t;

/* include/libiarray/iarray.h:214-226 */
typedef struct iarray_config_s {
  iarray_compression_codec_t compression_codec;
  int compression_level;
  iarray_compression_favor_t compression_favor;
  int use_dict;
  int splitmode;
  int filter_flags;
  unsigned int eval_method;
  int max_num_threads; /* host threads for the Blosc2 write stage */
  uint8_t fp_mantissa_bits;
  bool btune; /* accepted; codecs are never re-tuned (operands stay LZ4) */
  uint8_t compression_meta;
} iarray_config_t;

/* include/libiarray/iarray.h:228-233 */
typedef struct iarray_dtshape_s {
  iarray_data_type_t dtype;
  int32_t dtype_size;
  int8_t ndim;
  int64_t shape[IARRAY_DIMENSION_MAX];
} iarray_dtshape_t;

/* include/libiarray/iarray.h:235-240 */
typedef struct iarray_storage_s {
  char *urlpath; /* must be NULL: containers live in memory (frames are read by iarray_container_load) */
  bool contiguous;
  int64_t chunkshape[IARRAY_DIMENSION_MAX];
  int64_t blockshape[IARRAY_DIMENSION_MAX];
} iarray_storage_t;

/* include/libiarray/iarray.h:275-287 */
static const iarray_config_t IARRAY_CONFIG_DEFAULTS = {
    .compression_codec = IARRAY_COMPRESSION_LZ4,
    .compression_level = 5,
    .compression_favor = IARRAY_COMPRESSION_FAVOR_BALANCE,
    .use_dict = 0,
    .splitmode = IARRAY_AUTO_SPLIT,
    .filter_flags = IARRAY_COMP_SHUFFLE,
    .eval_method = 2,
    .max_num_threads = 1,
    .fp_mantissa_bits = 0,
    .btune = true,
    .compression_meta = 0,
};

/* include/libiarray/iarray.h:289-299 */
static const iarray_config_t IARRAY_CONFIG_NO_COMPRESSION = {
    .compression_codec = IARRAY_COMPRESSION_LZ4,
    .compression_level = 0,
    .compression_favor = IARRAY_COMPRESSION_FAVOR_BALANCE,
    .use_dict = 0,
    .splitmode = IARRAY_AUTO_SPLIT,
    .filter_flags = 0,
    .eval_method = 0,
    .max_num_threads = 1,
    .fp_mantissa_bits = 0,
    .btune = false,
    .compression_meta = 0,
};

/* ---- library and context: iarray.h:337-343 (src/iarray.c:77-112, 361-388) */
INA_API(ina_rc_t) iarray_init(void);
INA_API(void) iarray_destroy(void);
INA_API(const char *) iarray_err_strerror(ina_rc_t error);
INA_API(ina_rc_t) iarray_context_new(iarray_config_t *cfg, iarray_context_t **ctx);
INA_API(void) iarray_context_free(iarray_context_t **ctx);

/* ---- constructors: iarray.h:417-455, 664-674 (src/iarray_constructor.c) -- */
INA_API(ina_rc_t) iarray_empty(iarray_context_t *ctx, iarray_dtshape_t *dtshape, iarray_storage_t *storage,
                               iarray_container_t **container);
INA_API(ina_rc_t) iarray_zeros(iarray_context_t *ctx, iarray_dtshape_t *dtshape, iarray_storage_t *storage,
                               iarray_container_t **container);
INA_API(ina_rc_t) iarray_arange(iarray_context_t *ctx, iarray_dtshape_t *dtshape, double start, double step,
                                iarray_storage_t *storage, iarray_container_t **container);
INA_API(ina_rc_t) iarray_linspace(iarray_context_t *ctx, iarray_dtshape_t *dtshape, double start, double stop,
                                  iarray_storage_t *storage, iarray_container_t **container);
INA_API(ina_rc_t) iarray_from_buffer(iarray_context_t *ctx, iarray_dtshape_t *dtshape, void *buffer,
                                     int64_t buflen, iarray_storage_t *storage, iarray_container_t **container);
INA_API(ina_rc_t) iarray_to_buffer(iarray_context_t *ctx, iarray_container_t *container, void *buffer,
                                   int64_t buflen);
/* iarray.h:574 (src/iarray_container.c:500-541) */
INA_API(ina_rc_t) iarray_get_slice_buffer(iarray_context_t *ctx, iarray_container_t *container, const int64_t *start,
                                          const int64_t *stop, void *buffer, int64_t buflen);

/* ---- container info: iarray.h:630-681 ------------------------------------ */
INA_API(ina_rc_t) iarray_get_dtshape(iarray_context_t *ctx, iarray_container_t *c, iarray_dtshape_t *dtshape);
INA_API(ina_rc_t) iarray_get_storage(iarray_context_t *ctx, iarray_container_t *c, iarray_storage_t *storage);
INA_API(ina_rc_t) iarray_is_view(iarray_context_t *ctx, iarray_container_t *c, bool *view);
INA_API(ina_rc_t) iarray_container_info(iarray_container_t *container, int64_t *nbytes, int64_t *cbytes);
INA_API(void) iarray_container_free(iarray_context_t *ctx, iarray_container_t **container);
/* iarray.h:691 (src/iarray_container.c:1212-1287); |a-b|/a with a signed
 * denominator, exactly as the reference computes it */
INA_API(ina_rc_t) iarray_container_almost_equal(iarray_container_t *a, iarray_container_t *b, double tol);

/* ---- linear algebra: iarray.h:760-766 ---------------------------------------
 * iarray_linalg_matmul: src/linalg/matmul.c:19-68 (validation, allocation of
 * C, dispatch to the gemm branch src/linalg/matmul/gemm.c:130-267 or the gemv
 * branch src/linalg/matmul/gemv.c:104-208).  Here every C block is decoded,
 * multiplied and re-shuffled on the GPU; see iarray_amd.h for the precision
 * forms beside the plain one. */
INA_API(ina_rc_t) iarray_linalg_matmul(iarray_context_t *ctx, iarray_container_t *a, iarray_container_t *b,
                                       iarray_storage_t *storage, iarray_container_t **c);
/* iarray.h:560-566, src/iarray_container.c:364-455: a slice as a view of
 * src (sharing its chunks) or as a new container written with `storage`; a
 * slice of a transposed view is the transposed slice of its container. */
INA_API(ina_rc_t) iarray_get_slice(iarray_context_t *ctx, iarray_container_t *src, const int64_t *start,
                                   const int64_t *stop, bool view, iarray_storage_t *storage,
                                   iarray_container_t **container);
/* src/linalg/transpose.c:18-45 */
INA_API(ina_rc_t) iarray_linalg_transpose(iarray_context_t *ctx, iarray_container_t *a, iarray_container_t **b);

/* ---- frames: iarray.h:593-599 ------------------------------------------------
 * A container read from a contiguous Blosc2 frame file written by the
 * reference (src/iarray_container.c:79-218): shape, chunk and block shapes
 * from the "caterva" metalayer, dtype from the "iarray" one, the chunks copied
 * in.  iarray_container_open reads the chunks at open as well (the reference
 * reads them lazily).  A path that does not exist returns INA_ERR_FILE_OPEN;
 * frames whose offsets chunk is compressed, or that hold special-value
 * chunks, return INA_ERR_NOT_SUPPORTED. */
INA_API(ina_rc_t) iarray_container_load(iarray_context_t *ctx, char *urlpath, iarray_container_t **container);
INA_API(ina_rc_t) iarray_container_open(iarray_context_t *ctx, char *urlpath, iarray_container_t **container);

/* iarray.h:601-617, src/iarray_container.c:50-76,219-300.
 * iarray_container_save writes a base container as one contiguous frame file
 * (views and opened containers: IARRAY_ERR_INVALID_STORAGE); iarray_to_cframe
 * returns the same bytes in a malloc'ed buffer (*needs_free is always true);
 * iarray_from_cframe reads a frame from memory (the chunks are always
 * copied); iarray_container_remove deletes a frame file, or a frame directory
 * and the files in it (a path that does not exist is not an error).  A frame
 * read and saved again keeps its header fields, so the file is rewritten byte
 * for byte when it holds no variable-length metalayers. */
INA_API(ina_rc_t) iarray_container_save(iarray_context_t *ctx, iarray_container_t *container, char *urlpath);
INA_API(ina_rc_t) iarray_to_cframe(iarray_context_t *ctx, iarray_container_t *src, uint8_t **frame,
                                   int64_t *frame_len, bool *needs_free);
INA_API(ina_rc_t) iarray_from_cframe(iarray_context_t *ctx, uint8_t *frame, int64_t frame_len, bool copy,
                                     iarray_container_t **container);
INA_API(ina_rc_t) iarray_container_remove(char *urlpath);

/* ---- block-wise entry points: iarray.h:860-870 ------------------------------
 * iarray_opt_gemv: src/linalg/iarray_gemv.c:249-351; iarray_opt_gemm:
 * src/linalg/iarray_gemm.c:315-432 (same partition constraints and error
 * codes; zero chunks and zero blocks are skipped as in the reference). */
INA_API(ina_rc_t) iarray_opt_gemv(iarray_context_t *ctx, iarray_container_t *a, iarray_container_t *b,
                                  iarray_storage_t *storage, iarray_container_t **c);
INA_API(ina_rc_t) iarray_opt_gemm(iarray_context_t *ctx, iarray_container_t *a, iarray_container_t *b,
                                  iarray_storage_t *storage, iarray_container_t **c);
/* iarray.h:872-882.  iarray_opt_gemm_a: src/linalg/iarray_gemm_a.c:191-358 (A one
 * chunk column, fetched whole per C chunk); iarray_opt_gemm_b:
 * src/linalg/iarray_gemm_b.c:191-358 (B one chunk row, the mirror).  The same
 * partition checks in the same order with the same error codes; the product is
 * the GPU GEMM of iarray_linalg_matmul. */
INA_API(ina_rc_t) iarray_opt_gemm_a(iarray_context_t *ctx, iarray_container_t *a, iarray_container_t *b,
                                    iarray_storage_t *storage, iarray_container_t **c);
INA_API(ina_rc_t) iarray_opt_gemm_b(iarray_context_t *ctx, iarray_container_t *a, iarray_container_t *b,
                                    iarray_storage_t *storage, iarray_container_t **c);

#ifdef __cplusplus
}
#endif

#include "iarray_amd.h"

#endif
