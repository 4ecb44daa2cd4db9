/*
 * iarray_amd.h -- extensions that stand beside the drop-in iarray API.
 *
 * None of these has a counterpart in the reference's include/libiarray/iarray.h;
 * each says what it is for.  They use the same types, ownership and
 * ina_rc_t error convention as the drop-in entry points.
 */
#ifndef IAMX_IARRAY_AMD_H_
#define IAMX_IARRAY_AMD_H_

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic form of a matmul.  PLAIN is the drop-in (FP32 or FP64 products
 * and sums, like the reference's MKL blocked matmul).  The two mixed forms
 * trade exactness for MFMA throughput:
 *   SPLIT_BF16  (float32 only): a = hi + lo in bfloat16, C = hi*hi + hi*lo + lo*hi
 *               on BF16 MFMA with FP32 accumulation;
 *   OZAKI_INT8  (float64 only): per-row/column power-of-two scaling, operands
 *               cut into signed 7-bit INT8 slices, exact INT32 MFMA slice
 *               products, FP64 recombination and unscaling;
 *   OZAKI_CRT   (float64 only): per-row/column power-of-two scaling to
 *               integers, one exact INT8 MFMA product per modulus of their
 *               residues modulo n pairwise coprime moduli <= 256, exact
 *               Chinese-remainder recombination and FP64 unscaling (n INT8
 *               products where OZAKI_INT8 needs S(S+1)/2). */
typedef enum iarray_matmul_precision_e {
  IARRAY_MATMUL_PLAIN = 0,
  IARRAY_MATMUL_SPLIT_BF16 = 1,
  IARRAY_MATMUL_OZAKI_INT8 = 2,
  IARRAY_MATMUL_OZAKI_CRT = 3,
} iarray_matmul_precision_t;

typedef struct iarray_matmul_params_s {
  iarray_matmul_precision_t precision;
  /* slices (OZAKI_INT8, 2..8) or moduli (OZAKI_CRT, 2..15); 0 = pick the
   * smallest count that reaches ~1e-13 componentwise error */
  int ozaki_slices;
} iarray_matmul_params_t;

/* iarray_linalg_matmul with an explicit arithmetic form. */
INA_API(ina_rc_t) iarray_linalg_matmul_params(iarray_context_t *ctx, iarray_container_t *a, iarray_container_t *b,
                                              iarray_storage_t *storage, const iarray_matmul_params_t *params,
                                              iarray_container_t **c);

/* Many independent matmuls (any shapes) in one device pass: all operand blocks
 * are decoded by one launch, all C tiles are computed by one persistent
 * kernel pulling tiles from a device work queue, and all C blocks are
 * re-shuffled by one launch.  c[i] receives a new container per product. */
INA_API(ina_rc_t) iarray_linalg_matmul_batch(iarray_context_t *ctx, int n, iarray_container_t **a,
                                             iarray_container_t **b, iarray_storage_t *storage,
                                             const iarray_matmul_params_t *params, iarray_container_t **c);

/* Read-only view of compressed chunk `nchunk` (Blosc2 chunk bytes), the role
 * blosc2_schunk_get_lazychunk plays on catarr->sc in the reference. */
INA_API(ina_rc_t) iarray_container_chunk(iarray_container_t *c, int64_t nchunk, const uint8_t **chunk,
                                         int32_t *csize);
INA_API(ina_rc_t) iarray_container_nchunks(iarray_container_t *c, int64_t *nchunks);

/* Build a container from existing Blosc2 chunks (e.g. chunks of a frame
 * written by c-blosc2), copied.  Chunk i must be the i-th chunk of the
 * Caterva tiling implied by dtshape/storage. */
INA_API(ina_rc_t) iarray_from_chunks(iarray_context_t *ctx, iarray_dtshape_t *dtshape, iarray_storage_t *storage,
                                     int64_t nchunks, const uint8_t *const *chunks, const int32_t *csizes,
                                     iarray_container_t **container);

/* The device decode plan of a container (built and uploaded on first use,
 * owned by the container, valid until it is freed or changed): the job
 * tables that libiamx.so's iamx_decode / iamx_gemv_fused /
 * iamx_gemv_fused_noalloc consume, for callers that drive those entry points
 * themselves.  Views have no plan of their own (INA_ERR_NOT_SUPPORTED). */
typedef struct iarray_device_plan_s {
  const void *lz4;       /* iamx_lz4_job_t[n_lz4], device memory */
  int n_lz4;
  const void *blocks;    /* iamx_block_job_t[n_blocks], device memory */
  const int32_t *blk_rc; /* 4 int32 per block job (first row, first column, block column, flags: bit 0 a zero block), device memory */
  int n_blocks;
  int nbcols;            /* block columns of the global block grid */
  int64_t scratch_bytes; /* decoded-stream scratch the LZ4 jobs need */
  int64_t blockshape[2];
} iarray_device_plan_t;

INA_API(ina_rc_t) iarray_container_device_plan(iarray_container_t *c, iarray_device_plan_t *plan);

/* Decode chunk bytes back to their flat uncompressed content on the GPU
 * (no tiling applied): used to check the device decoder on arbitrary chunks. */
INA_API(ina_rc_t) iarray_decode_chunks_flat(iarray_context_t *ctx, int64_t nchunks, const uint8_t *const *chunks,
                                            const int32_t *csizes, void *out, int64_t outlen);

/* Where the last matmul spent device time.  When enabled, every kernel the
 * library launches is bracketed by HIP events on the library's stream. */
INA_API(void) iarray_prof_enable(bool on);
INA_API(void) iarray_prof_reset(void);
/* Entry i (0-based) of the per-kernel table; returns false past the end. */
INA_API(bool) iarray_prof_entry(int i, const char **name, int64_t *launches, double *total_ms);

/* Host seconds spent in the Blosc2 write stage (LZ4 encode of C streams) by
 * the last matmul, and device milliseconds from the first to the last kernel. */
INA_API(void) iarray_last_matmul_times(double *host_encode_s, double *device_ms);

/* Bytes of filtered C streams the GPU encoder read, and bytes of finished
 * chunks it wrote, in the last matmul's write stage (0 when every chunk was
 * encoded on the host). */
INA_API(void) iarray_last_gpu_encode_bytes(int64_t *stream_bytes, int64_t *chunk_bytes);

/* Slice (OZAKI_INT8) or modulus (OZAKI_CRT) count used by the last Ozaki
 * product (0 if none ran) and, when
 * it was picked automatically (ozaki_slices = 0), the componentwise error
 * max |C_S - C| / (|A||B|) measured on the sampled rows (else -1).  The
 * automatic pick is the smallest S >= 4 (moduli: >= 8) whose sampled error is at most
 * IARRAY_OZAKI_TARGET (environment, default 1e-13), cached per (A, B) pair
 * while neither container changes. */
INA_API(int) iarray_last_ozaki_slices(double *sampled_err);

/* Panels the last GEMV streamed its matrix in (0: the matrix was resident).
 * A matrix whose compressed chunks plus decoded streams exceed the device
 * budget (IARRAY_DEVICE_BUDGET bytes, environment, else half the device's
 * free memory) travels through the device in panels of chunk rows. */
INA_API(int) iarray_last_gemv_panels(void);

/* Device description (name, CU count, gfx arch string). */
INA_API(ina_rc_t) iarray_device_info(char *name, int namelen, int *cus, char *arch, int archlen);

#ifdef __cplusplus
}
#endif
#endif
