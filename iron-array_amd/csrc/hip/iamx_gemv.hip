// iamx_gemv.hip -- GEMV fused with the Blosc2 block un-filter.
//
// After the LZ4 streams of A are decoded into scratch (k_lz4_streams), one
// workgroup (4 wavefronts) per compressed A block gathers each element of its
// tile straight from the byte planes / bit planes (scratch, raw streams read
// in place from the chunk arena, or constant streams), multiplies it by the x
// segment staged in LDS and reduces every row with a fixed xor tree.  Row sums
// land in partial[block column][row]; a second launch adds the block columns
// in order, so y is deterministic run to run.  Dense A never exists in HBM.
#include "iamx_internal.h"

namespace {

template <typename T>
__device__ __forceinline__ T bits_to(uint64_t v);
template <>
__device__ __forceinline__ float bits_to<float>(uint64_t v) {
  return __uint_as_float((uint32_t)v);
}
template <>
__device__ __forceinline__ double bits_to<double>(uint64_t v) {
  return __longlong_as_double((long long)v);
}

constexpr int kMaxCols = 1024;

template <typename T>
__global__ __launch_bounds__(256) void k_gemv_fused(const iamx_block_job_t *__restrict__ jobs,
                                                    const int32_t *__restrict__ blk_rc, int njobs,
                                                    const T *__restrict__ x, const uint8_t *__restrict__ scratch,
                                                    T *__restrict__ partial, int64_t M) {
  __shared__ iamx_block_job_t J;
  __shared__ const uint8_t *ptr[IAMX_MAX_STREAMS];
  __shared__ T xs[kMaxCols];
  const int b = blockIdx.x;
  if (b >= njobs) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) J = jobs[b];
  __syncthreads();
  const int64_t row0 = blk_rc[4 * b], col0 = blk_rc[4 * b + 1];
  const int bcol = blk_rc[4 * b + 2];
  if (tid < IAMX_MAX_STREAMS) {
    const int s = tid;
    ptr[s] = s >= J.nstreams                ? nullptr
             : J.kind[s] == IAMX_SRC_ADDR    ? reinterpret_cast<const uint8_t *>(J.src[s])
             : J.kind[s] == IAMX_SRC_SCRATCH ? scratch + J.src[s]
                                             : nullptr;
  }
  for (int c = tid; c < J.vcols; c += 256) xs[c] = x[col0 + c];
  __syncthreads();
  const int ts = J.typesize;
  const int32_t n = J.bsize / ts;
  const int32_t cols = J.cols;
  const int32_t neblock = J.neblock;
  const int filter = J.filter;
  const int32_t n8 = filter == IAMX_FILTER_BITSHUFFLE ? n - n % 8 : n;
  const int32_t rowlen = n8 / 8;
  auto byte_at = [&](int32_t q) -> uint32_t {
    const int s = q / neblock;
    const int32_t o = q - s * neblock;
    return ptr[s] ? ptr[s][o] : J.fill[s];
  };
  for (int32_t r = wave; r < J.vrows; r += 4) {
    T acc = 0;
    for (int32_t c = lane; c < J.vcols; c += 64) {
      const int32_t e = r * cols + c;
      uint64_t v = 0;
      if ((filter == IAMX_FILTER_BITSHUFFLE || filter == IAMX_FILTER_BITSHUFFLE_V2) && e < n8) {
        const int sh = e & 7;
        for (int p = 0; p < 8 * ts; p++) v |= (uint64_t)((byte_at(p * rowlen + (e >> 3)) >> sh) & 1u) << p;
      } else if (filter == IAMX_FILTER_SHUFFLE && ts > 1) {
        for (int j = 0; j < ts; j++) v |= (uint64_t)byte_at(j * n + e) << (8 * j);
      } else {
        for (int j = 0; j < ts; j++) v |= (uint64_t)byte_at(e * ts + j) << (8 * j);
      }
      acc += bits_to<T>(v) * xs[c];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    const int64_t row = row0 + r;
    if (lane == 0 && row < M) partial[(int64_t)bcol * M + row] = acc;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_reduce_cols(const T *__restrict__ partial, int nbc, int64_t M,
                                                     T *__restrict__ y) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= M) return;
  T s = 0;
  for (int c = 0; c < nbc; c++) s += partial[(int64_t)c * M + r];
  y[r] = s;
}

}  // namespace

extern "C" {

IAMX_API int iamx_gemv_fused(int dtype, const iamx_lz4_job_t *d_lz4, int n_lz4, const iamx_block_job_t *d_blocks,
                             const int32_t *d_blk_rc, int n_blocks, int nbc, int bs0, int bs1, const void *x,
                             void *y, int64_t M, void *d_scratch, void *d_partial) {
  int rc = iamx::ensure_init();
  if (rc) return rc;
  if (bs1 > kMaxCols) return iamx::set_error(IAMX_EINVAL, "iamx_gemv_fused: block width %d > %d", bs1, kMaxCols);
  (void)bs0;
  if (n_blocks <= 0 || nbc <= 0) return iamx::set_error(IAMX_EINVAL, "iamx_gemv_fused: no blocks");
  rc = iamx_decode(d_lz4, n_lz4, nullptr, 0, d_scratch, nullptr);
  if (rc) return rc;
  const dim3 rgrid((unsigned)((M + 255) / 256));
  if (dtype == 0) {
    IAMX_LAUNCH("gemv_fused_f64", k_gemv_fused<double>, dim3(n_blocks), dim3(256), 0, d_blocks, d_blk_rc, n_blocks,
                (const double *)x, (const uint8_t *)d_scratch, (double *)d_partial, M);
    IAMX_LAUNCH("gemv_reduce_f64", k_reduce_cols<double>, rgrid, dim3(256), 0, (const double *)d_partial, nbc, M,
                (double *)y);
  } else if (dtype == 1) {
    IAMX_LAUNCH("gemv_fused_f32", k_gemv_fused<float>, dim3(n_blocks), dim3(256), 0, d_blocks, d_blk_rc, n_blocks,
                (const float *)x, (const uint8_t *)d_scratch, (float *)d_partial, M);
    IAMX_LAUNCH("gemv_reduce_f32", k_reduce_cols<float>, rgrid, dim3(256), 0, (const float *)d_partial, nbc, M,
                (float *)y);
  } else {
    return iamx::set_error(IAMX_EINVAL, "iamx_gemv_fused: dtype %d", dtype);
  }
  return IAMX_OK;
}

}  // extern "C"
