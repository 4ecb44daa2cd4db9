// iamx_gemv.hip -- GEMV fused with the Blosc2 block un-filter.
//
// After the LZ4 streams of A are decoded into scratch (k_lz4_streams), one
// workgroup per compressed A block turns the block's streams straight into
// row sums of A_blk x_seg without writing A anywhere:
//   phase 1  every thread takes groups of 8 consecutive elements: for a
//            bitshuffled block the 8 elements' bits are byte g of each of the
//            8*typesize bit planes, and one 8x8 bit-matrix transpose per byte
//            position rebuilds them; for a byte-shuffled block byte j of the 8
//            elements is 8 consecutive bytes of plane j.  Each element is
//            multiplied by x[column] (x staged in LDS) and the group's products
//            are summed per row (a group touches at most two rows), into LDS;
//   phase 2  one thread per row adds its groups' partial sums in group order.
// Rows are processed in windows so the per-group partials fit in LDS.  Row
// sums land in partial[block column][row]; a second launch adds the block
// columns in order, so y is bit-identical run to run.
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

__device__ __forceinline__ uint64_t transpose8(uint64_t x) {
  uint64_t t;
  t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
  x = x ^ t ^ (t << 7);
  t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
  x = x ^ t ^ (t << 14);
  t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
  x = x ^ t ^ (t << 28);
  return x;
}

constexpr int kMaxCols = 1024;
constexpr int kGroups = 4096;  // per-window groups whose partial sums live in LDS

struct Src {
  const IAMX_GLOBAL uint8_t *ptr[IAMX_MAX_STREAMS];
  uint8_t fill[IAMX_MAX_STREAMS];
};

template <typename T>
__global__ __launch_bounds__(256) void k_gemv_fused(const iamx_block_job_t *__restrict__ jobs,
                                                    const int32_t *__restrict__ blk_rc, int njobs,
                                                    const T *__restrict__ x, const uint8_t *__restrict__ scratch,
                                                    T *__restrict__ partial, int64_t M) {
  __shared__ iamx_block_job_t J;
  __shared__ Src S;
  __shared__ T xs[kMaxCols];
  __shared__ T p0[kGroups], p1[kGroups];
  const int b = blockIdx.x;
  if (b >= njobs) return;
  const int tid = threadIdx.x;
  if (tid == 0) J = jobs[b];
  __syncthreads();
  const int64_t row0 = blk_rc[4 * b], col0 = blk_rc[4 * b + 1];
  const int bcol = blk_rc[4 * b + 2];
  if (tid < IAMX_MAX_STREAMS) {
    const int s = tid;
    S.ptr[s] = s >= J.nstreams                ? nullptr
               : J.kind[s] == IAMX_SRC_ADDR    ? iamx::gptr<const uint8_t>(J.src[s])
               : J.kind[s] == IAMX_SRC_SCRATCH ? iamx::gptr<const uint8_t>((uint64_t)(scratch + J.src[s]))
                                               : nullptr;
    S.fill[s] = s < J.nstreams ? J.fill[s] : 0;
  }
  for (int c = tid; c < J.vcols; c += 256) xs[c] = x[col0 + c];
  __syncthreads();

  const int ts = J.typesize;
  const int32_t n = J.bsize / ts;
  const int32_t cols = J.cols, vrows = J.vrows, vcols = J.vcols;
  const int32_t neblock = J.neblock;
  const int filter = J.filter;
  const bool bits = filter == IAMX_FILTER_BITSHUFFLE || filter == IAMX_FILTER_BITSHUFFLE_V2;
  const int32_t n8 = filter == IAMX_FILTER_BITSHUFFLE ? n - n % 8 : n;
  const int32_t rowlen = n8 / 8;
  const bool one_stream = J.nstreams == 1;
  const bool split_planes = filter == IAMX_FILTER_SHUFFLE && ts > 1;

  // byte q of the block's concatenated streams
  auto byte_at = [&](int32_t q) -> uint32_t {
    if (one_stream) return S.ptr[0] ? S.ptr[0][q] : S.fill[0];
    const int s = q / neblock;
    const int32_t o = q - s * neblock;
    return S.ptr[s] ? S.ptr[s][o] : S.fill[s];
  };
  // byte j of element e under a byte shuffle (plane j is n bytes long)
  auto plane_byte = [&](int j, int32_t e) -> uint32_t {
    if (J.nstreams == ts) return S.ptr[j] ? S.ptr[j][e] : S.fill[j];
    return byte_at(j * n + e);
  };

  // rows per window: the window's groups must fit kGroups
  const int32_t rows_per_win = cols >= 8 ? (int32_t)(((int64_t)(kGroups - 2) * 8) / cols) : 0;
  if (rows_per_win <= 0) {
    // tiny blocks (fewer than 8 columns): one wavefront per row, plain gather
    const int lane = tid & 63, wave = tid >> 6;
    for (int32_t r = wave; r < vrows; r += 4) {
      T acc = 0;
      for (int32_t c = lane; c < vcols; c += 64) {
        const int32_t e = r * cols + c;
        uint64_t v = 0;
        if (bits && e < n8) {
          for (int p = 0; p < 8 * ts; p++) v |= (uint64_t)((byte_at(p * rowlen + (e >> 3)) >> (e & 7)) & 1u) << p;
        } else if (split_planes) {
          for (int j = 0; j < ts; j++) v |= (uint64_t)plane_byte(j, e) << (8 * j);
        } else {
          for (int j = 0; j < ts; j++) v |= (uint64_t)byte_at(e * ts + j) << (8 * j);
        }
        acc += bits_to<T>(v) * xs[c];
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
      if (lane == 0 && row0 + r < M) partial[(int64_t)bcol * M + row0 + r] = acc;
    }
    return;
  }

  for (int32_t ra = 0; ra < vrows; ra += rows_per_win) {
    const int32_t rb = min(vrows, ra + rows_per_win);
    const int32_t g0 = (ra * cols) / 8, g1 = (rb * cols - 1) / 8;  // groups [g0, g1]
    // phase 1: per-group partial sums for rows in [ra, rb)
    for (int32_t g = g0 + tid; g <= g1; g += 256) {
      const int32_t e0 = 8 * g;
      uint64_t v[8];
      if (bits && e0 + 8 <= n8) {
#pragma unroll
        for (int i = 0; i < 8; i++) v[i] = 0;
        for (int j = 0; j < ts; j++) {
          uint64_t xw = 0;
#pragma unroll
          for (int k = 0; k < 8; k++) xw |= (uint64_t)byte_at((8 * j + k) * rowlen + g) << (8 * k);
          const uint64_t y = transpose8(xw);
#pragma unroll
          for (int i = 0; i < 8; i++) v[i] |= ((y >> (8 * i)) & 0xffull) << (8 * j);
        }
      } else {
#pragma unroll
        for (int i = 0; i < 8; i++) {
          const int32_t e = e0 + i;
          uint64_t w = 0;
          if (e < n) {
            if (split_planes) {
              for (int j = 0; j < ts; j++) w |= (uint64_t)plane_byte(j, e) << (8 * j);
            } else if (bits && e < n8) {
              for (int p = 0; p < 8 * ts; p++) w |= (uint64_t)((byte_at(p * rowlen + (e >> 3)) >> (e & 7)) & 1u) << p;
            } else {
              for (int j = 0; j < ts; j++) w |= (uint64_t)byte_at(e * ts + j) << (8 * j);
            }
          }
          v[i] = w;
        }
      }
      const int32_t rfirst = e0 / cols;
      int32_t r = rfirst, c = e0 - rfirst * cols;
      T s0 = 0, s1 = 0;
#pragma unroll
      for (int i = 0; i < 8; i++) {
        if (r >= ra && r < rb && c < vcols) {
          const T prod = bits_to<T>(v[i]) * xs[c];
          if (r == rfirst) s0 += prod;
          else s1 += prod;
        }
        if (++c == cols) {
          c = 0;
          ++r;
        }
      }
      p0[g - g0] = s0;
      p1[g - g0] = s1;
    }
    __syncthreads();
    // phase 2: one thread per row adds its groups in order
    for (int32_t r = ra + tid; r < rb; r += 256) {
      const int32_t ga = (r * cols) / 8, gb = (r * cols + cols - 1) / 8;
      T acc = 0;
      for (int32_t g = ga; g <= gb; g++) acc += ((8 * g) / cols == r) ? p0[g - g0] : p1[g - g0];
      if (row0 + r < M) partial[(int64_t)bcol * M + row0 + r] = acc;
    }
    __syncthreads();
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
