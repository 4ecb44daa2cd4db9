/*
 * iarray_core.c -- library and context lifetime, error strings.
 * Mirrors iarray_init/iarray_destroy (src/iarray.c:77-112),
 * iarray_context_new/free (src/iarray.c:361-388) and iarray_err_strerror.
 * The GPU is opened lazily by the first call that needs it, so the library
 * loads and the host-only entry points work on machines without a GPU; every
 * entry point that needs the device fails with IARRAY_ERR_NO_DEVICE there.
 */
#include <malloc.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "iarray_private.h"

static int g_initialized = 0;
static __thread char g_msg[512];
double iax_last_encode_s = 0.0;
double iax_last_device_ms = 0.0;
int64_t iax_last_gpu_enc_in = 0, iax_last_gpu_enc_out = 0;
int iax_last_slices = 0;
double iax_last_slice_err = -1.0;

ina_rc_t iax_set_error(ina_rc_t rc, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
  va_end(ap);
  if (getenv("IARRAY_TRACE")) fprintf(stderr, "iarray: %s\n", g_msg);
  return rc;
}

ina_rc_t iax_dev_error(int code, const char *what) {
  if (code == IAMX_ENODEV)
    return iax_set_error(INA_ERROR(IARRAY_ERR_NO_DEVICE), "%s: %s", what, iamx_last_error());
  if (code == IAMX_EDATA)
    return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "%s: %s", what, iamx_last_error());
  return iax_set_error(INA_ERROR(IARRAY_ERR_DEVICE_FAILED), "%s: %s", what, iamx_last_error());
}

/* Process-wide setup at init, as the reference's iarray_init sets the CPU
 * affinity of the process (src/iarray.c:94-102).  Chunk buffers are
 * megabytes each and a product frees and allocates thousands of them: glibc's
 * defaults hand such blocks back to the kernel on free (munmap, heap trim) and
 * fault them in again on the next product, which cost 5-7% of a large
 * product's time.  Blocks up to 32 MiB are therefore kept in the heap and the
 * heap is not trimmed, unless the process set glibc's own MALLOC_*_ variables
 * or IARRAY_MALLOC_TUNE=0. */
static void tune_malloc(void) {
  const char *off = getenv("IARRAY_MALLOC_TUNE");
  if (off && off[0] == '0') return;
  if (!getenv("MALLOC_MMAP_THRESHOLD_")) (void)mallopt(M_MMAP_THRESHOLD, 32 << 20);
  if (!getenv("MALLOC_TRIM_THRESHOLD_")) (void)mallopt(M_TRIM_THRESHOLD, INT32_MAX);
}

INA_API(ina_rc_t) iarray_init(void) {
  if (!g_initialized) tune_malloc();
  g_initialized = 1;
  return INA_SUCCESS;
}

INA_API(void) iarray_destroy(void) {
  iax_operand_forget(-1);
  if (g_initialized) iamx_shutdown();
  g_initialized = 0;
}

INA_API(const char *) iarray_err_strerror(ina_rc_t error) {
  if (INA_SUCCEED(error)) return "Success";
  const uint64_t e = INA_RC_ERROR(error);
  const char *what;
  switch (e) {
    case IARRAY_ERR_INVALID_DTSHAPE: what = "Invalid dtshape"; break;
    case IARRAY_ERR_INVALID_DTYPE: what = "Invalid dtype"; break;
    case IARRAY_ERR_INVALID_SHAPE: what = "Invalid shape"; break;
    case IARRAY_ERR_INVALID_CHUNKSHAPE: what = "Invalid chunkshape"; break;
    case IARRAY_ERR_INVALID_BLOCKSHAPE: what = "Invalid blockshape"; break;
    case IARRAY_ERR_INVALID_NDIM: what = "Invalid ndim"; break;
    case IARRAY_ERR_INVALID_STORAGE: what = "Invalid storage"; break;
    case IARRAY_ERR_TOO_SMALL_BUFFER: what = "Buffer too small"; break;
    case IARRAY_ERR_BLOSC_FAILED: what = "Blosc failed"; break;
    case IARRAY_ERR_EMPTY_CONTAINER: what = "Empty container"; break;
    case IARRAY_ERR_DEVICE_FAILED: what = "GPU operation failed"; break;
    case IARRAY_ERR_NO_DEVICE: what = "No GPU available"; break;
    default: what = "Error"; break;
  }
  static __thread char buf[640];
  snprintf(buf, sizeof(buf), "%s%s%s", what, g_msg[0] ? ": " : "", g_msg);
  return buf;
}

INA_API(ina_rc_t) iarray_context_new(iarray_config_t *cfg, iarray_context_t **ctx) {
  if (!cfg || !ctx) return INA_ERROR(INA_ERR_NULL_POINTER);
  if (!g_initialized) {
    /* the reference aborts the process here (src/iarray.c:363-366); report it */
    return iax_set_error(INA_ERROR(INA_ERR_FAILED), "iarray_init() must be called before iarray_context_new()");
  }
  *ctx = (iarray_context_t *)calloc(1, sizeof(iarray_context_t));
  if (!*ctx) return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  (*ctx)->cfg = *cfg;
  return INA_SUCCESS;
}

INA_API(void) iarray_context_free(iarray_context_t **ctx) {
  if (ctx && *ctx) {
    free(*ctx);
    *ctx = NULL;
  }
}

INA_API(void) iarray_prof_enable(bool on) { iamx_prof_enable(on ? 1 : 0); }
INA_API(void) iarray_prof_reset(void) { iamx_prof_reset(); }
INA_API(bool) iarray_prof_entry(int i, const char **name, int64_t *launches, double *total_ms) {
  return iamx_prof_entry(i, name, launches, total_ms) != 0;
}

int iax_last_gemv_panels = 0;
INA_API(int) iarray_last_gemv_panels(void) { return iax_last_gemv_panels; }

INA_API(int) iarray_last_ozaki_slices(double *sampled_err) {
  if (sampled_err) *sampled_err = iax_last_slice_err;
  return iax_last_slices;
}

INA_API(void) iarray_last_matmul_times(double *host_encode_s, double *device_ms) {
  if (host_encode_s) *host_encode_s = iax_last_encode_s;
  if (device_ms) *device_ms = iax_last_device_ms;
}

INA_API(void) iarray_last_gpu_encode_bytes(int64_t *stream_bytes, int64_t *chunk_bytes) {
  if (stream_bytes) *stream_bytes = iax_last_gpu_enc_in;
  if (chunk_bytes) *chunk_bytes = iax_last_gpu_enc_out;
}

INA_API(ina_rc_t) iarray_device_info(char *name, int namelen, int *cus, char *arch, int archlen) {
  int rc = iamx_device_info(name, namelen, cus, arch, archlen);
  if (rc) return iax_dev_error(rc, "iarray_device_info");
  return INA_SUCCESS;
}
