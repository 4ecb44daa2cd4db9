// iamx_gemm.hip -- plain FP64 / FP32 GEMM on CDNA4 matrix cores, and GEMV.
//
// GEMM: C(MxN) = A(MxK) B(KxN), row-major, any M/N/K (edges predicated).
// 256 threads = 4 wavefronts in a 2x2 arrangement; each wavefront owns a
// 64x64 sub-tile made of 4x4 MFMA 16x16 tiles.  Operand tiles are staged in
// LDS k-major (A transposed on the way in) so that every MFMA operand load is
// 16 consecutive elements per k; the next K-slab is fetched into registers
// while the current one feeds the matrix cores (one barrier per slab).
//   float64: v_mfma_f64_16x16x4_f64   float32: v_mfma_f32_16x16x4_f32
// MFMA operand layout (16x16x4, one block): lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; result register v holds D[4*(l/16)+v][l%16] for float32 and
// D[4*v+l/16][l%16] for float64 (measured with a one-hot probe on gfx950).
#include <stdlib.h>

#include "iamx_internal.h"

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef f64x4 acc_t;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // result register v of lane l holds row 4*v + l/16 (column l%16)
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * v + (lane >> 4); }
};
template <>
struct Mfma<float> {
  typedef f32x4 acc_t;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // result register v of lane l holds row 4*(l/16) + v (column l%16)
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

constexpr int BM = 128, BN = 128, PAD = 4;

// Map the launch index onto output tiles so that the 8 XCDs (workgroups are
// dealt round-robin to XCDs) each get a contiguous band of tile rows: tiles
// that share A rows then share one L2.  XCD x receives nt / 8 workgroups, one
// more when x < nt % 8, and its band starts after the bands of XCDs 0..x-1, so
// the map is a bijection for every tile count.
__device__ __forceinline__ int band_id(int nt) {
  const int q = nt / 8, r = nt % 8, x = (int)blockIdx.x % 8;
  return x * q + min(x, r) + (int)blockIdx.x / 8;
}

// K slab depth: 32 for float32 (LDS 2 x 2 x 32 x 132 x 4 B = 68 KB), 16 for
// float64 (the same bytes).
template <typename T>
struct Slab {
  static constexpr int BK = sizeof(T) == 4 ? 32 : 16;
};

template <typename T>
__global__ __launch_bounds__(256) void k_gemm(int64_t M, int64_t N, int64_t K, const T *__restrict__ A, int64_t lda,
                                              const T *__restrict__ B, int64_t ldb, T *__restrict__ C, int64_t ldc,
                                              int tiles_n, int tile_base, int ntiles, int splits, int64_t kc,
                                              T *__restrict__ partial) {
  constexpr int BK = Slab<T>::BK;
  constexpr int PER = BK * BM / 256;  // elements per thread per operand and slab (16 f32, 8 f64)
  constexpr int VEC = 16 / sizeof(T);  // elements per 16-byte load
  __shared__ T As[2][BK][BM + PAD];
  __shared__ T Bs[2][BK][BN + PAD];
  typedef typename Mfma<T>::acc_t acc_t;
  typedef T vec_t __attribute__((ext_vector_type(VEC)));
  // work item id = (tile tile_base + t, K part s) for t < ntiles, s < splits
  const int id = band_id(ntiles * splits);
  const int tl = id / splits, part = id - tl * splits;
  const int tile = tile_base + tl;
  const int tm = tile / tiles_n, tn = tile % tiles_n;
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  const int64_t kb = (int64_t)part * kc, ke = min(K, kb + kc);  // this item's K range (kc: a multiple of BK)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;

  // global->LDS mapping: A: row tid/2, k (tid%2)*PER..+PER ; B: k tid/(BN/PER), cols ..+PER
  const int a_r = tid >> 1, a_k = (tid & 1) * PER;
  const int b_k = tid / (BN / PER), b_c = (tid % (BN / PER)) * PER;
  const bool rows_in = m0 + BM <= M && n0 + BN <= N;
  T ra[PER], rb[PER];
  auto gload = [&](int64_t k0) {
    if (rows_in && k0 + BK <= ke) {
      // interior slab: 16-byte loads
      const T *pa = A + (m0 + a_r) * lda + k0 + a_k;
      const T *pb = B + (k0 + b_k) * ldb + n0 + b_c;
#pragma unroll
      for (int v = 0; v < PER / VEC; v++) {
        vec_t x, y;
        __builtin_memcpy(&x, pa + v * VEC, sizeof(x));
        __builtin_memcpy(&y, pb + v * VEC, sizeof(y));
#pragma unroll
        for (int e = 0; e < VEC; e++) {
          ra[v * VEC + e] = x[e];
          rb[v * VEC + e] = y[e];
        }
      }
      return;
    }
    const int64_t gr = m0 + a_r;
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int64_t gk = k0 + a_k + i;
      ra[i] = (gr < M && gk < ke) ? A[gr * lda + gk] : (T)0;
    }
    const int64_t gk = k0 + b_k;
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int64_t gc = n0 + b_c + i;
      rb[i] = (gk < ke && gc < N) ? B[gk * ldb + gc] : (T)0;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < PER; i++) As[buf][a_k + i][a_r] = ra[i];
#pragma unroll
    for (int i = 0; i < PER; i++) Bs[buf][b_k][b_c + i] = rb[i];
  };

  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = acc_t{0, 0, 0, 0};

  const int64_t nk = (ke - kb + BK - 1) / BK;
  gload(kb);
  sstore(0);
  __syncthreads();
  for (int64_t kt = 0; kt < nk; kt++) {
    const int cur = (int)(kt & 1);
    if (kt + 1 < nk) gload(kb + (kt + 1) * BK);
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
        for (int j = 0; j < 4; j++) acc[i][j] = Mfma<T>::mma(a[i], b[j], acc[i][j]);
    }
    if (kt + 1 < nk) sstore(cur ^ 1);
    __syncthreads();
  }
  if (partial) {  // a K part of a split tile: the whole 128 x 128 partial tile
    T *pt = partial + ((int64_t)part * ntiles + tl) * (BM * BN);
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int col = wc * 64 + j * 16 + (lane & 15);
#pragma unroll
        for (int v = 0; v < 4; v++) pt[(wr * 64 + i * 16 + Mfma<T>::row(lane, v)) * BN + col] = acc[i][j][v];
      }
    return;
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int64_t col = n0 + wc * 64 + j * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int64_t row = m0 + wr * 64 + i * 16 + Mfma<T>::row(lane, v);
        if (row < M && col < N) C[row * ldc + col] = acc[i][j][v];
      }
    }
}

// C of the split tiles: the sum of their `splits` partial tiles, parts in
// ascending order (the same bits every run).
template <typename T>
__global__ __launch_bounds__(256) void k_sum_parts(const T *__restrict__ partial, int splits, int ntiles,
                                                   int tile_base, int tiles_n, int64_t M, int64_t N,
                                                   T *__restrict__ C, int64_t ldc) {
  const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (int64_t)ntiles * BM * BN) return;
  const int tl = (int)(idx / (BM * BN)), e = (int)(idx % (BM * BN));
  const int tile = tile_base + tl;
  const int64_t row = (int64_t)(tile / tiles_n) * BM + e / BN, col = (int64_t)(tile % tiles_n) * BN + e % BN;
  if (row >= M || col >= N) return;
  T sum = partial[(int64_t)tl * (BM * BN) + e];
  for (int p = 1; p < splits; p++) sum += partial[((int64_t)p * ntiles + tl) * (BM * BN) + e];
  C[row * ldc + col] = sum;
}

// y = A x : one wavefront per row, lanes stride over k, tree reduction.
template <typename T>
__global__ __launch_bounds__(256) void k_gemv(int64_t M, int64_t K, const T *__restrict__ A, int64_t lda,
                                              const T *__restrict__ x, T *__restrict__ y) {
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= M) return;
  const T *a = A + row * lda;
  T s = 0;
  for (int64_t k = lane; k < K; k += 64) s += a[k] * x[k];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) y[row] = s;
}

// ---- block-pair GEMM of the opt paths ------------------------------------------------------
// The reference's opt_gemm / opt_gemm_a / opt_gemm_b compute each C block as
// the sum, over the K blocks in ascending order, of A block x B block, and
// skip a pair when either block is a zero block (src/linalg/iarray_gemm.c:
// 143-165, iarray_gemm_a.c / _b.c:100-175): an Inf or NaN facing a zero
// block never reaches C.  Here each workgroup computes a 64 x 64 piece of one
// C block (block row gr, block column gc) from the dense operand images,
// skipping the flagged pairs; rows/ks/cols give each block row, K block and
// block column as [first, end) (trimmed at chunk and array edges).
template <typename T>
__global__ __launch_bounds__(256) void k_gemm_blockpairs(const T *__restrict__ A, int64_t lda,
                                                         const T *__restrict__ B, int64_t ldb, T *__restrict__ C,
                                                         int64_t ldc, const int32_t *__restrict__ rows,
                                                         const int32_t *__restrict__ ks, int nbk,
                                                         const int32_t *__restrict__ cols, int nbc,
                                                         const uint8_t *__restrict__ zero_a,
                                                         const uint8_t *__restrict__ zero_b, int br0, int sub_r,
                                                         int sub_c) {
  __shared__ T As[16][64 + 1];
  __shared__ T Bs[16][64];
  const int gc = (int)blockIdx.x / sub_c, sc = (int)blockIdx.x % sub_c;
  const int gr = br0 + (int)blockIdx.y / sub_r, sr = (int)blockIdx.y % sub_r;
  const int r0 = rows[2 * gr] + 64 * sr, r1 = min(r0 + 64, rows[2 * gr + 1]);
  const int c0 = cols[2 * gc] + 64 * sc, c1 = min(c0 + 64, cols[2 * gc + 1]);
  if (r0 >= r1 || c0 >= c1) return;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  T acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = (T)0;
  for (int gk = 0; gk < nbk; gk++) {
    if (zero_a[(int64_t)gr * nbk + gk] | zero_b[(int64_t)gk * nbc + gc]) continue;  // uniform
    const int k1 = ks[2 * gk + 1];
    for (int kk = ks[2 * gk]; kk < k1; kk += 16) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int e = tid + 256 * q;  // 1024 = 64 x 16
        const int ar = e >> 4, ak = e & 15;
        As[ak][ar] = (r0 + ar < r1 && kk + ak < k1) ? A[(int64_t)(r0 + ar) * lda + kk + ak] : (T)0;
        const int bk = e >> 6, bc = e & 63;
        Bs[bk][bc] = (kk + bk < k1 && c0 + bc < c1) ? B[(int64_t)(kk + bk) * ldb + c0 + bc] : (T)0;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < 16; k++) {
        T a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; i++) a[i] = As[k][ty * 4 + i];
#pragma unroll
        for (int j = 0; j < 4; j++) b[j] = Bs[k][tx * 4 + j];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) acc[i][j] = fma(a[i], b[j], acc[i][j]);
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int r = r0 + ty * 4 + i, c = c0 + tx * 4 + j;
      if (r < r1 && c < c1) C[(int64_t)r * ldc + c] = acc[i][j];
    }
}

}  // namespace

extern "C" {

IAMX_API int iamx_gemm_blockpairs(int dtype, const void *A, int64_t lda, const void *B, int64_t ldb, void *C,
                                  int64_t ldc, const int32_t *d_rows, int nbr, const int32_t *d_ks, int nbk,
                                  const int32_t *d_cols, int nbc, const uint8_t *d_zero_a, const uint8_t *d_zero_b,
                                  int br0, int br1, int max_rows, int max_cols) {
  int rc = iamx::ensure_init();
  if (rc) return rc;
  if (br0 < 0 || br1 > nbr || br0 > br1 || nbk < 0 || nbc <= 0 || max_rows <= 0 || max_cols <= 0)
    return iamx::set_error(IAMX_EINVAL, "iamx_gemm_blockpairs: bad block ranges");
  if (br0 == br1) return IAMX_OK;
  const int sub_r = (max_rows + 63) / 64, sub_c = (max_cols + 63) / 64;
  const dim3 grid((unsigned)(nbc * sub_c), (unsigned)((br1 - br0) * sub_r));
  if (dtype == 0)
    IAMX_LAUNCH("gemm_blockpairs_f64", k_gemm_blockpairs<double>, grid, dim3(256), 0, (const double *)A, lda,
                (const double *)B, ldb, (double *)C, ldc, d_rows, d_ks, nbk, d_cols, nbc, d_zero_a, d_zero_b, br0,
                sub_r, sub_c);
  else if (dtype == 1)
    IAMX_LAUNCH("gemm_blockpairs_f32", k_gemm_blockpairs<float>, grid, dim3(256), 0, (const float *)A, lda,
                (const float *)B, ldb, (float *)C, ldc, d_rows, d_ks, nbk, d_cols, nbc, d_zero_a, d_zero_b, br0,
                sub_r, sub_c);
  else
    return iamx::set_error(IAMX_EINVAL, "iamx_gemm_blockpairs: dtype %d", dtype);
  return IAMX_OK;
}

int iamx_gemm_split_bf16(int64_t M, int64_t N, int64_t K, const float *A, int64_t lda, const float *B, int64_t ldb,
                         float *C, int64_t ldc);
int iamx_gemm_ozaki_i8(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B, int64_t ldb,
                       double *C, int64_t ldc, int S);
int iamx_gemm_ozaki_crt(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B, int64_t ldb,
                        double *C, int64_t ldc, int n);
int iamx_mixed_prepare_b(int form, int64_t N, int64_t K, const void *B, int64_t ldb, int S);
int iamx_mixed_rows(int form, int64_t M, const void *A, int64_t lda, void *C, int64_t ldc);

IAMX_API int iamx_gemm(int dtype, int precision, int slices, int64_t M, int64_t N, int64_t K, const void *A,
                       int64_t lda, const void *B, int64_t ldb, void *C, int64_t ldc) {
  int rc = iamx::ensure_init();
  if (rc) return rc;
  (void)slices;
  if (M <= 0 || N <= 0 || K < 0) return iamx::set_error(IAMX_EINVAL, "iamx_gemm: bad shape %lld x %lld x %lld",
                                                        (long long)M, (long long)N, (long long)K);
  if (precision == 1) {
    if (dtype != 1) return iamx::set_error(IAMX_EINVAL, "iamx_gemm: split-bf16 needs float32");
    return iamx_gemm_split_bf16(M, N, K, (const float *)A, lda, (const float *)B, ldb, (float *)C, ldc);
  }
  if (precision == 2) {
    if (dtype != 0) return iamx::set_error(IAMX_EINVAL, "iamx_gemm: ozaki-int8 needs float64");
    return iamx_gemm_ozaki_i8(M, N, K, (const double *)A, lda, (const double *)B, ldb, (double *)C, ldc,
                              slices > 0 ? slices : 7);
  }
  if (precision == 3) {
    if (dtype != 0) return iamx::set_error(IAMX_EINVAL, "iamx_gemm: ozaki-crt needs float64");
    return iamx_gemm_ozaki_crt(M, N, K, (const double *)A, lda, (const double *)B, ldb, (double *)C, ldc,
                               slices > 0 ? slices : 14);
  }
  if (precision != 0) return iamx::set_error(IAMX_EINVAL, "iamx_gemm: precision %d not available", precision);
  if (dtype != 0 && dtype != 1) return iamx::set_error(IAMX_EINVAL, "iamx_gemm: dtype %d", dtype);
  // the hand-written MFMA kernel
  auto own = [&](int) -> int {
    const int64_t tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
    if (tiles_m * tiles_n > INT32_MAX / 8) return iamx::set_error(IAMX_EINVAL, "iamx_gemm: too many tiles");
    // Tiles run as rounds of `slots` workgroups (two per CU).  The tiles past
    // the last full round would leave most CUs idle for a whole round, so they
    // are split into K parts (enough to fill a round, at most 8) whose partial
    // tiles are summed in order by a second kernel; a product with fewer tiles
    // than one round is split the same way.  IAMX_GEMM_SPLITK=0 turns this off.
    static const bool splitk = [] {
      const char *e = getenv("IAMX_GEMM_SPLITK");
      return !(e && atoi(e) == 0);
    }();
    const int nt = (int)(tiles_m * tiles_n);
    const int slots = 2 * (iamx::g_num_cus > 0 ? iamx::g_num_cus : 256);
    const int full = nt / slots * slots, rem = nt - full;
    const int64_t bk = dtype == 0 ? Slab<double>::BK : Slab<float>::BK;
    int S = rem > 0 ? slots / rem : 1;
    if (S > 8) S = 8;
    int64_t kc = (K + bk - 1) / bk;  // K slabs
    kc = (kc + S - 1) / S * bk;       // slabs per part, as elements
    if (kc < 4 * bk) kc = 4 * bk;     // at least four slabs a part
    S = kc > 0 ? (int)((K + kc - 1) / kc) : 1;
    const bool split = splitk && rem > 0 && S >= 2 && rem * 4 <= slots * 3;
    const int head = split ? full : nt;
    const int64_t isz = dtype == 0 ? 8 : 4;
    if (head > 0) {
      const dim3 grid((unsigned)head);
      if (dtype == 0)
        IAMX_LAUNCH("gemm_f64_mfma", k_gemm<double>, grid, dim3(256), 0, M, N, K, (const double *)A, lda,
                    (const double *)B, ldb, (double *)C, ldc, (int)tiles_n, 0, head, 1, K, (double *)nullptr);
      else
        IAMX_LAUNCH("gemm_f32_mfma", k_gemm<float>, grid, dim3(256), 0, M, N, K, (const float *)A, lda,
                    (const float *)B, ldb, (float *)C, ldc, (int)tiles_n, 0, head, 1, K, (float *)nullptr);
    }
    if (!split) return IAMX_OK;
    void *part = iamx_workspace(16, (size_t)S * rem * BM * BN * isz + 256);
    if (!part) return iamx::set_error(IAMX_ENOMEM, "iamx_gemm: split-K partial tiles");
    const dim3 pgrid((unsigned)(rem * S)), rgrid((unsigned)(((int64_t)rem * BM * BN + 255) / 256));
    if (dtype == 0) {
      IAMX_LAUNCH("gemm_f64_mfma", k_gemm<double>, pgrid, dim3(256), 0, M, N, K, (const double *)A, lda,
                  (const double *)B, ldb, (double *)C, ldc, (int)tiles_n, full, rem, S, kc, (double *)part);
      IAMX_LAUNCH("gemm_sum_parts", k_sum_parts<double>, rgrid, dim3(256), 0, (const double *)part, S, rem, full,
                  (int)tiles_n, M, N, (double *)C, ldc);
    } else {
      IAMX_LAUNCH("gemm_f32_mfma", k_gemm<float>, pgrid, dim3(256), 0, M, N, K, (const float *)A, lda,
                  (const float *)B, ldb, (float *)C, ldc, (int)tiles_n, full, rem, S, kc, (float *)part);
      IAMX_LAUNCH("gemm_sum_parts", k_sum_parts<float>, rgrid, dim3(256), 0, (const float *)part, S, rem, full,
                  (int)tiles_n, M, N, (float *)C, ldc);
    }
    return IAMX_OK;
  };
  if (K == 0 || !iamx::blt_usable()) return own(256);
  // the autotuner weighs it against hipBLASLt once per shape class
  auto lib = [&]() -> int {
    return iamx::blt_gemm(dtype == 0 ? iamx::BLT_F64 : iamx::BLT_F32, 0, M, N, K, A, lda, B, ldb, C, ldc);
  };
  auto run = [&](int c) -> int { return c == 0 ? lib() : own(c); };
  const int cands[2] = {256, 0};
  int prc = IAMX_OK;
  iamx::tune_run(dtype == 0 ? 3 : 4, M, N, K, cands, 2, [](int c, void *x) { return (*(decltype(run) *)x)(c); },
                 &run, &prc);
  return prc;
}

// B operand fixed by iamx_gemm_prepare_b for the iamx_gemm_rows calls that follow
static struct {
  int dtype, precision;
  int64_t N, K, ldb;
  const void *B;
  bool valid;
} g_rows;

IAMX_API int iamx_gemm_prepare_b(int dtype, int precision, int slices, int64_t N, int64_t K, const void *B,
                                 int64_t ldb) {
  int rc = iamx::ensure_init();
  if (rc) return rc;
  g_rows.valid = false;
  if (N <= 0 || K < 0 || !B) return iamx::set_error(IAMX_EINVAL, "iamx_gemm_prepare_b: bad shape %lld x %lld",
                                                     (long long)K, (long long)N);
  if (dtype != 0 && dtype != 1) return iamx::set_error(IAMX_EINVAL, "iamx_gemm_prepare_b: dtype %d", dtype);
  if (precision == 1 && dtype != 1) return iamx::set_error(IAMX_EINVAL, "iamx_gemm: split-bf16 needs float32");
  if (precision == 2 && dtype != 0) return iamx::set_error(IAMX_EINVAL, "iamx_gemm: ozaki-int8 needs float64");
  if (precision == 3 && dtype != 0) return iamx::set_error(IAMX_EINVAL, "iamx_gemm: ozaki-crt needs float64");
  if (precision < 0 || precision > 3)
    return iamx::set_error(IAMX_EINVAL, "iamx_gemm: precision %d not available", precision);
  if (precision != 0) {
    const int count = precision == 2 ? (slices > 0 ? slices : 7) : precision == 3 ? (slices > 0 ? slices : 14) : 0;
    rc = iamx_mixed_prepare_b(precision, N, K, B, ldb, count);
    if (rc) return rc;
  }
  g_rows.dtype = dtype;
  g_rows.precision = precision;
  g_rows.N = N;
  g_rows.K = K;
  g_rows.ldb = ldb;
  g_rows.B = B;
  g_rows.valid = true;
  return IAMX_OK;
}

IAMX_API int iamx_gemm_rows(int64_t M, const void *A, int64_t lda, void *C, int64_t ldc) {
  int rc = iamx::ensure_init();
  if (rc) return rc;
  if (!g_rows.valid) return iamx::set_error(IAMX_EINVAL, "iamx_gemm_rows: no B operand prepared");
  if (M <= 0) return iamx::set_error(IAMX_EINVAL, "iamx_gemm_rows: bad row count %lld", (long long)M);
  if (g_rows.precision != 0) return iamx_mixed_rows(g_rows.precision, M, A, lda, C, ldc);
  return iamx_gemm(g_rows.dtype, 0, 0, M, g_rows.N, g_rows.K, A, lda, g_rows.B, g_rows.ldb, C, ldc);
}

IAMX_API int iamx_gemv(int dtype, int64_t M, int64_t K, const void *A, int64_t lda, const void *x, void *y) {
  int rc = iamx::ensure_init();
  if (rc) return rc;
  if (M <= 0 || K < 0) return iamx::set_error(IAMX_EINVAL, "iamx_gemv: bad shape");
  dim3 grid((unsigned)((M + 3) / 4));
  if (dtype == 0)
    IAMX_LAUNCH("gemv_f64", k_gemv<double>, grid, dim3(256), 0, M, K, (const double *)A, lda, (const double *)x,
                (double *)y);
  else if (dtype == 1)
    IAMX_LAUNCH("gemv_f32", k_gemv<float>, grid, dim3(256), 0, M, K, (const float *)A, lda, (const float *)x,
                (float *)y);
  else
    return iamx::set_error(IAMX_EINVAL, "iamx_gemv: dtype %d", dtype);
  return IAMX_OK;
}

}  // extern "C"
