// iamx_batch.hip -- many independent matmuls in one device pass, and graph
// capture of the library stream.
//
// k_gemm_batch is a persistent kernel: every workgroup pulls output tiles
// from one device work queue (an atomic counter over a tile list that the host
// sorts by decreasing K, so the longest tiles start first) and computes each
// 128x128 tile of product p with the same MFMA tile code as k_gemm.  The grid
// is sized to the resident-workgroup capacity of the GPU, so every workgroup
// exits after the queue runs dry.  k_pack_batch re-shuffles every C block of
// every product in one launch.
#include "iamx_internal.h"

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int BM = 128, BN = 128, BK = 16, PAD = 4;

template <typename T>
struct MfmaB;
template <>
struct MfmaB<double> {
  typedef f64x4 acc_t;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * v + (lane >> 4); }
};
template <>
struct MfmaB<float> {
  typedef f32x4 acc_t;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

__global__ void k_zero_i32(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_gemm_batch(const iamx_batch_prod_t *__restrict__ prods,
                                                    const int32_t *__restrict__ tiles, int ntiles, int *counter,
                                                    const T *__restrict__ base_a, const T *__restrict__ base_b,
                                                    T *__restrict__ base_c) {
  __shared__ T As[2][BK][BM + PAD];
  __shared__ T Bs[2][BK][BN + PAD];
  __shared__ int s_tile;
  typedef typename MfmaB<T>::acc_t acc_t;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int a_r = tid >> 1, a_k = (tid & 1) * 8;
  const int b_k = tid >> 4, b_c = (tid & 15) * 8;
  for (;;) {
    if (tid == 0) s_tile = atomicAdd(counter, 1);
    __syncthreads();
    const int t = s_tile;
    __syncthreads();  // every thread has read s_tile before it is rewritten
    if (t >= ntiles) break;
    const int p = tiles[3 * t], tm = tiles[3 * t + 1], tn = tiles[3 * t + 2];
    const iamx_batch_prod_t P = prods[p];
    const T *A = base_a + P.a_off;
    const T *B = base_b + P.b_off;
    T *C = base_c + P.c_off;
    const int64_t M = P.M, N = P.N, K = P.K;
    const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
    T ra[8], rb[8];
    auto gload = [&](int64_t k0) {
      const int64_t gr = m0 + a_r;
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int64_t gk = k0 + a_k + i;
        ra[i] = (gr < M && gk < K) ? A[gr * K + gk] : (T)0;
      }
      const int64_t gk = k0 + b_k;
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int64_t gc = n0 + b_c + i;
        rb[i] = (gk < K && gc < N) ? B[gk * N + gc] : (T)0;
      }
    };
    auto sstore = [&](int buf) {
#pragma unroll
      for (int i = 0; i < 8; i++) As[buf][a_k + i][a_r] = ra[i];
#pragma unroll
      for (int i = 0; i < 8; i++) Bs[buf][b_k][b_c + i] = rb[i];
    };
    acc_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) acc[i][j] = acc_t{0, 0, 0, 0};
    const int64_t nk = (K + BK - 1) / BK;
    gload(0);
    sstore(0);
    __syncthreads();
    for (int64_t kt = 0; kt < nk; kt++) {
      const int cur = (int)(kt & 1);
      if (kt + 1 < nk) gload((kt + 1) * BK);
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        T a[4], b[4];
        const int k = kk + (lane >> 4);
#pragma unroll
        for (int i = 0; i < 4; i++) a[i] = As[cur][k][wr * 64 + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < 4; j++) b[j] = Bs[cur][k][wc * 64 + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) acc[i][j] = MfmaB<T>::mma(a[i], b[j], acc[i][j]);
      }
      if (kt + 1 < nk) sstore(cur ^ 1);
      __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int64_t col = n0 + wc * 64 + j * 16 + (lane & 15);
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int64_t row = m0 + wr * 64 + i * 16 + MfmaB<T>::row(lane, v);
          if (row < M && col < N) C[row * N + col] = acc[i][j][v];
        }
      }
  }
}

__device__ __forceinline__ uint64_t load_elem(const uint8_t *src, int ts, int64_t idx) {
  return ts == 8 ? reinterpret_cast<const uint64_t *>(src)[idx] : reinterpret_cast<const uint32_t *>(src)[idx];
}

__device__ __forceinline__ uint64_t t8(uint64_t x) {
  uint64_t t;
  t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
  x = x ^ t ^ (t << 7);
  t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
  x = x ^ t ^ (t << 14);
  t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
  x = x ^ t ^ (t << 28);
  return x;
}

// One workgroup per C block of any product: blk[2*i] = product, blk[2*i+1] =
// block index inside that product (chunk-major, block-major).
__global__ __launch_bounds__(256) void k_pack_batch(const iamx_batch_pack_t *__restrict__ packs,
                                                    const int32_t *__restrict__ blk, const uint8_t *__restrict__ base_c,
                                                    uint8_t *__restrict__ out) {
  const iamx_batch_pack_t Q = packs[blk[2 * blockIdx.x]];
  const iamx_tiling_t &t = Q.tiling;
  const int64_t id = blk[2 * blockIdx.x + 1];
  const int64_t nbpc = t.nblocks[0] * t.nblocks[1];
  const int64_t nc = id / nbpc, nb = id % nbpc;
  const int64_t ci0 = nc / t.nchunks[1], ci1 = nc % t.nchunks[1];
  const int64_t bi0 = nb / t.nblocks[1], bi1 = nb % t.nblocks[1];
  const int ts = t.typesize;
  const int32_t cols = (int32_t)t.block[1];
  const int32_t n = (int32_t)(t.block[0] * t.block[1]);
  const int64_t r_in = bi0 * t.block[0], c_in = bi1 * t.block[1];
  const int64_t gr0 = ci0 * t.chunk[0] + r_in, gc0 = ci1 * t.chunk[1] + c_in;
  int64_t vr = t.chunk[0] - r_in;
  if (t.shape[0] - gr0 < vr) vr = t.shape[0] - gr0;
  int64_t vc = t.chunk[1] - c_in;
  if (t.shape[1] - gc0 < vc) vc = t.shape[1] - gc0;
  const uint8_t *src = base_c + Q.c_off_bytes;
  const int64_t ld = t.shape[1];
  uint8_t *o = out + Q.out_off + id * (int64_t)n * ts;
  auto elem = [&](int32_t e) -> uint64_t {
    const int32_t r = e / cols, c = e - (e / cols) * cols;
    if (r >= vr || c >= vc) return 0;
    return load_elem(src, ts, (gr0 + r) * ld + gc0 + c);
  };
  if (t.filter == IAMX_FILTER_BITSHUFFLE) {
    const int32_t n8 = n - n % 8, rowlen = n8 / 8;
    for (int32_t g = threadIdx.x; g < rowlen; g += blockDim.x) {
      uint64_t v[8];
      for (int i = 0; i < 8; i++) v[i] = elem(8 * g + i);
      for (int j = 0; j < ts; j++) {
        uint64_t x = 0;
        for (int i = 0; i < 8; i++) x |= ((v[i] >> (8 * j)) & 0xffull) << (8 * i);
        const uint64_t y = t8(x);
        for (int k = 0; k < 8; k++) o[(int64_t)(8 * j + k) * rowlen + g] = (uint8_t)(y >> (8 * k));
      }
    }
    for (int32_t e = n8 + threadIdx.x; e < n; e += blockDim.x) {
      const uint64_t v = elem(e);
      for (int j = 0; j < ts; j++) o[(int64_t)e * ts + j] = (uint8_t)(v >> (8 * j));
    }
    return;
  }
  const bool shuffled = t.filter == IAMX_FILTER_SHUFFLE;
  for (int32_t e = threadIdx.x; e < n; e += blockDim.x) {
    const uint64_t v = elem(e);
    for (int j = 0; j < ts; j++) {
      if (shuffled) o[(int64_t)j * n + e] = (uint8_t)(v >> (8 * j));
      else o[(int64_t)e * ts + j] = (uint8_t)(v >> (8 * j));
    }
  }
}

}  // namespace

extern "C" {

IAMX_API int iamx_gemm_batch(int dtype, const iamx_batch_prod_t *d_prods, const int32_t *d_tiles, int ntiles,
                             const void *base_a, const void *base_b, void *base_c) {
  int rc = iamx::ensure_init();
  if (rc) return rc;
  if (ntiles <= 0) return IAMX_OK;
  int *counter = iamx::status_word() + 16;  // a spare word of the status page
  IAMX_LAUNCH("batch_queue_reset", k_zero_i32, dim3(1), dim3(64), 0, counter, 1);
  int per_cu = 0;
  const dim3 block(256);
  hipError_t e;
  if (dtype == 0) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gemm_batch<double>, 256, 0);
  else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gemm_batch<float>, 256, 0);
  if (e != hipSuccess || per_cu < 1) per_cu = 1;
  int grid = per_cu * iamx::g_num_cus;
  if (grid > ntiles) grid = ntiles;
  if (dtype == 0)
    IAMX_LAUNCH("gemm_batch_f64", k_gemm_batch<double>, dim3(grid), block, 0, d_prods, d_tiles, ntiles, counter,
                (const double *)base_a, (const double *)base_b, (double *)base_c);
  else if (dtype == 1)
    IAMX_LAUNCH("gemm_batch_f32", k_gemm_batch<float>, dim3(grid), block, 0, d_prods, d_tiles, ntiles, counter,
                (const float *)base_a, (const float *)base_b, (float *)base_c);
  else
    return iamx::set_error(IAMX_EINVAL, "iamx_gemm_batch: dtype %d", dtype);
  return IAMX_OK;
}

IAMX_API int iamx_pack_batch(const iamx_batch_pack_t *d_packs, const int32_t *d_blocks, int nblocks,
                             const void *base_c, void *d_out) {
  int rc = iamx::ensure_init();
  if (rc) return rc;
  if (nblocks <= 0) return IAMX_OK;
  IAMX_LAUNCH("pack_batch", k_pack_batch, dim3(nblocks), dim3(256), 0, d_packs, d_blocks, (const uint8_t *)base_c,
              (uint8_t *)d_out);
  return IAMX_OK;
}

}  // extern "C"
