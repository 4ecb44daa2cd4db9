// iamx_internal.h -- shared pieces of the HIP translation units of libiamx.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/iamx.h"

namespace iamx {

extern hipStream_t g_stream;
// The stream launches go to: the library stream, or the side stream between
// iamx_use_stream(1) and iamx_use_stream(0).
extern hipStream_t g_cur;
// kernels and graph replays queued so far (a status snapshot is current while this is unchanged)
extern uint64_t g_launch_seq;
extern int g_num_cus;

int set_error(int code, const char *fmt, ...);
int check_hip(hipError_t e, const char *what);
int ensure_init();
// True once ensure_init has succeeded (the no-allocation entry point refuses
// to run before: init allocates the stream and status word).
bool initialized();

// True between iamx_capture_begin and iamx_capture_end.
bool capturing();

// Device word the decode kernels OR their error bits into.
int *status_word();

// Launch the LZ4 stream decoder over n_lz4 jobs into d_scratch, profiled as
// `name`; a job whose decoded bytes would not fit below byte `cap` of the
// scratch is skipped and flags IAMX_STATUS_RANGE.
int launch_lz4(const char *name, const iamx_lz4_job_t *d_lz4, int n_lz4, void *d_scratch,
               int64_t cap = INT64_MAX);

// Plain dense products through hipBLASLt (iamx_blaslt.hip).  Row-major
// C[b] (rows x n, ldc) = A[b] (rows x k, lda) B[b], B given as k x n (ldb;
// bt = 0) or as its n x k transpose (bt = 1); batch products with element
// strides sa, sb, sc.  blt_usable() is false during a graph capture or when
// the library cannot start; the autotuners then keep their own kernels.
enum { BLT_F64 = 0, BLT_F32 = 1, BLT_BF16_F32 = 2, BLT_I8_I32 = 3 };
bool blt_usable();
// `name` labels the launch in the profile (default: by type).
int blt_gemm(int type, int bt, int64_t rows, int64_t n, int64_t k, const void *A, int64_t lda, const void *B,
             int64_t ldb, void *C, int64_t ldc, int batch = 1, int64_t sa = 0, int64_t sb = 0, int64_t sc = 0,
             const char *name = nullptr);
void blt_shutdown();

// Autotuner of the dense products (iamx_mixed.hip): the first product of a
// (form, M, N, K) class runs every candidate in cands (the hand-written
// kernel's tile, 128 or 256, or 0 for hipBLASLt) through run(candidate,
// ctx), timed, and the fastest is used for that class from then on.
// Returns the candidate used; *rc is the status of the product.
int tune_run(int form, int64_t M, int64_t N, int64_t K, const int *cands, int ncand, int (*run)(int, void *),
             void *ctx, int *rc);

// Zstd streams: their decode pass (jobs of other codecs are skipped) and
// whether it runs (switched on once a zstd stream has been planned).
int launch_zstd(const char *name, const iamx_lz4_job_t *d_jobs, int njobs, void *d_scratch, int64_t cap);
bool zstd_enabled();

// Resolves this device's copy of the no-allocation scratch array (once per
// device, with the device current); called by the library's init.
int noalloc_setup();

// Profiling: brackets one launch with HIP events when enabled.
struct ProfScope {
  int slot = -1;
  explicit ProfScope(const char *name);
  ~ProfScope();
};

// Device addresses held in job tables are plain integers; reading them through
// a global-address-space pointer lets the compiler emit global_load/store
// instead of flat accesses.
#define IAMX_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ IAMX_GLOBAL T *gptr(uint64_t a) {
  return (IAMX_GLOBAL T *)a;
}

}  // namespace iamx

#define IAMX_HIP_CHECK(expr)                                   \
  do {                                                         \
    hipError_t _e = (expr);                                    \
    if (_e != hipSuccess) return iamx::check_hip(_e, #expr);   \
  } while (0)

// Launch `kernel` on the library stream inside a profiling scope and check
// the launch status.
#define IAMX_LAUNCH(name, kernel, grid, block, shmem, ...)                          \
  do {                                                                              \
    iamx::ProfScope _ps(name);                                                      \
    ++iamx::g_launch_seq;                                                           \
    hipLaunchKernelGGL(kernel, grid, block, shmem, iamx::g_cur, __VA_ARGS__);       \
    hipError_t _e = hipGetLastError();                                              \
    if (_e != hipSuccess) return iamx::check_hip(_e, name);                         \
  } while (0)
