// iamx_mixed.hip -- the two mixed-precision GEMM forms on CDNA4 matrix cores.
//
// Both run one "k-segmented" GEMM: C = sum over segments g of A_g . B_g^T,
// where every operand segment is a row-major byte matrix (rows x Kb bytes,
// K contiguous; B is stored transposed, N x Kb) and a segment table says
// which A and B planes each K-range pairs up.  128x128 output tile per
// workgroup, 4 wavefronts of 64x64 (4x4 MFMA 16x16 tiles), 128 operand bytes
// of K per LDS stage (two MFMA k-steps), next stage prefetched to registers.
//
// SPLIT_BF16 (float32 operands): x = hi + lo with hi = bf16(x), lo = bf16(x-hi)
//   (round to nearest even); segments (hi,hi), (hi,lo), (lo,hi) accumulate in
//   one FP32 accumulator on v_mfma_f32_16x16x32_bf16.
// OZAKI_INT8 (float64 operands): row i of A is scaled by 2^-e_i (e_i =
//   ilogb(max_k |A_ik|) + 2, so every scaled value lies in (-1/2, 1/2)), column j
//   of B by 2^-f_j, and each scaled value is cut into S balanced base-128 digits
//   d_s = rint(128 q), q <- 128 q - d_s (|d_s| <= 64, |q| <= 1/2, all exact in
//   FP64), so A_ik = 2^e_i (sum_s d_s 2^-7s + q_S 2^-7S) with |q_S| <= 1/2.
//   Rounding to nearest keeps the dropped residual unbiased.  Level L = s+t
//   collects the slice products D_s E_t (s,t <= S, L <= S+1) on
//   v_mfma_i32_16x16x64_i8 with exact INT32 sums (K*S*64^2 < 2^31 is checked);
//   at the end of each level
//   the INT32 tile is added into an FP64 accumulator with weight 2^-7L, and
//   the epilogue applies 2^(e_i+f_j).  FP64 arithmetic appears only in that
//   recombination and in the exponent scaling.
// OZAKI_CRT (float64 operands, the CRT scheme of Ozaki, Uchino and Imamura,
//   "Ozaki Scheme II", 2024): row i of A is scaled to integers A'_ik =
//   rint(A_ik 2^(p+1-e_i)) with |A'_ik| <= 2^p, column j of B likewise with q,
//   p + q chosen so that K 2^(p+q) < M/2 for M = m_1 ... m_n, n pairwise
//   coprime moduli <= 256.  Each modulus gets one INT8 product of the
//   symmetric residues (A' mod m_l)(B' mod m_l) on v_mfma_i32_32x32x32_i8 with
//   exact INT32 sums, reduced mod m_l in the epilogue to one byte.  A combine
//   kernel rebuilds C' = sum_l r_l W_l mod M exactly (W_l = (M/m_l)
//   ((M/m_l)^-1 mod m_l); 20-bit limbs), takes the representative in
//   [-M/2, M/2) and converts it to FP64 with the unscaling 2^(e_i+f_j-p-q-2):
//   n INT8 products instead of the S(S+1)/2 of the slicing form.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <initializer_list>

#include "iamx_internal.h"

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int TM = 128, TN = 128, KB = 128, ROWB = KB + 16;  // LDS row: 128 data bytes + 16 pad
constexpr int MAX_SEGS = 64;

struct Seg {
  int64_t a_off;  // byte offset of the A plane
  int64_t b_off;  // byte offset of the B plane
  int32_t level;  // Ozaki level (s+t); 0 for bf16
  int32_t level_end;
};

struct SegArgs {
  const uint8_t *A;  // planes: row-major rows x Kb bytes, row stride lda bytes
  const uint8_t *B;
  int64_t lda, ldb;  // bytes
  int64_t Kb;        // bytes of K per segment (multiple of KB)
  int64_t M, N;      // logical output size
  void *C;
  int64_t ldc;       // elements
  const int32_t *ea;  // Ozaki row exponents (M), column exponents (N)
  const int32_t *fb;
  int nseg;
  int tiles_n;
  Seg seg[MAX_SEGS];
  // OZAKI_CRT: one residue product per modulus, plane l of A at l * a_plane
  // bytes, of B at l * b_plane, residue bytes of C at l * c_plane (row
  // stride ldc bytes)
  int nmod;
  int64_t a_plane, b_plane, c_plane;
  int32_t mods[16];
};

// ---- bf16 split ------------------------------------------------------------------
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Writes hi/lo planes (rows_p x cols_p, zero padded, row stride ld), hi a
// second time into hi2.  trans = 0: plane[r][c] = X[r][c]; trans = 1:
// plane[r][c] = X[c][r] (B stored transposed).  The three planes sit side by
// side in rows of 3 cols_p, so the three products of the split are one
// product over K stacked three times.
__global__ __launch_bounds__(256) void k_split_bf16(const float *__restrict__ X, int64_t xr, int64_t xc, int64_t ldx,
                                                    int trans, uint16_t *__restrict__ hi, uint16_t *__restrict__ lo,
                                                    uint16_t *__restrict__ hi2, int64_t rows_p, int64_t cols_p,
                                                    int64_t ld) {
  __shared__ float t[32][33];
  const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  if (trans) {
    // plane rows are X columns: stage X[c0.., r0..] through LDS
    for (int i = ty; i < 32; i += 8) {
      const int64_t xrow = c0 + i, xcol = r0 + tx;
      t[i][tx] = (xrow < xr && xcol < xc) ? X[xrow * ldx + xcol] : 0.f;
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
      const int64_t r = r0 + i, c = c0 + tx;
      if (r < rows_p && c < cols_p) {
        const float v = t[tx][i];
        const uint16_t h = f2bf(v);
        hi[r * ld + c] = h;
        hi2[r * ld + c] = h;
        lo[r * ld + c] = f2bf(v - bf2f(h));
      }
    }
  } else {
    for (int i = ty; i < 32; i += 8) {
      const int64_t r = r0 + i, c = c0 + tx;
      if (r < rows_p && c < cols_p) {
        const float v = (r < xr && c < xc) ? X[r * ldx + c] : 0.f;
        const uint16_t h = f2bf(v);
        hi[r * ld + c] = h;
        hi2[r * ld + c] = h;
        lo[r * ld + c] = f2bf(v - bf2f(h));
      }
    }
  }
}

// ---- Ozaki prep -------------------------------------------------------------------
// exponent e with |x| * 2^-e < 1/2 for the largest |x| (0 for an all-zero line)
__device__ __forceinline__ int exp_of(double m) { return m > 0.0 ? ilogb(m) + 2 : 0; }

__global__ __launch_bounds__(256) void k_row_exp(const double *__restrict__ X, int64_t rows, int64_t cols,
                                                 int64_t ld, int32_t *__restrict__ e) {
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= rows) return;
  double m = 0.0;
  for (int64_t c = lane; c < cols; c += 64) m = fmax(m, fabs(X[r * ld + c]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if (lane == 0) e[r] = exp_of(m);
}

// Column exponents in three launches: fill with INT_MIN, a grid of 64-column
// x row-slab tiles (coalesced row reads) folding each slab's max exponent in
// with atomicMax (exp_of is monotonic in the magnitude), then INT_MIN (an
// all-zero column) -> 0.
__global__ __launch_bounds__(256) void k_fill_i32(int32_t *__restrict__ e, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) e[i] = v;
}

constexpr int kColSlab = 128;  // rows per tile of k_col_exp

__global__ __launch_bounds__(256) void k_col_exp(const double *__restrict__ X, int64_t rows, int64_t cols,
                                                 int64_t ld, int32_t *__restrict__ e) {
  __shared__ double red[4][64];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int64_t c = (int64_t)blockIdx.x * 64 + tx;
  const int64_t r0 = (int64_t)blockIdx.y * kColSlab;
  const int64_t r1 = r0 + kColSlab < rows ? r0 + kColSlab : rows;
  double m = 0.0;
  if (c < cols)
    for (int64_t r = r0 + ty; r < r1; r += 4) m = fmax(m, fabs(X[r * ld + c]));
  red[ty][tx] = m;
  __syncthreads();
  if (ty == 0 && c < cols) {
    m = fmax(fmax(red[0][tx], red[1][tx]), fmax(red[2][tx], red[3][tx]));
    if (m > 0.0) atomicMax(&e[c], exp_of(m));
  }
}

__global__ __launch_bounds__(256) void k_col_exp_zero(int32_t *__restrict__ e, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n && e[i] == INT32_MIN) e[i] = 0;
}

// S digit planes of rows_p x cols_p int8, side by side in rows of S cols_p:
// digit s (0 = most significant) at column block s, or S-1-s when rev (B,
// so that the slice pairs (s, L-s) of a level L sit at equal offsets of two
// contiguous column ranges: one product over K stacked n_L times).  trans as
// in k_split_bf16; the scale exponent is per plane row (e[r]).
__global__ __launch_bounds__(256) void k_slice_i8(const double *__restrict__ X, int64_t xr, int64_t xc, int64_t ldx,
                                                  int trans, const int32_t *__restrict__ e, int S,
                                                  int8_t *__restrict__ planes, int64_t rows_p, int64_t cols_p, int rev) {
  // a workgroup makes 32 plane rows x 128 plane columns (rows_p a multiple
  // of 32, cols_p of 128); a thread 4 consecutive columns of 4 rows, stored
  // as one dword per digit
  __shared__ double t[128][33];
  const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 128;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  if (trans) {
    for (int i = ty; i < 128; i += 8) {
      const int64_t xrow = c0 + i, xcol = r0 + tx;
      t[i][tx] = (xrow < xr && xcol < xc) ? X[xrow * ldx + xcol] : 0.0;
    }
    __syncthreads();
  }
  const int64_t c = c0 + 4 * tx;
  for (int i = ty; i < 32; i += 8) {
    const int64_t r = r0 + i;
    if (r >= rows_p || c >= cols_p) continue;
    const int er = r < (trans ? xc : xr) ? e[r] : 0;
    double q[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const double v = trans ? t[4 * tx + k][i] : ((r < xr && c + k < xc) ? X[r * ldx + c + k] : 0.0);
      q[k] = ldexp(v, -er);  // |q| < 1/2, exact
    }
    for (int s = 0; s < S; s++) {
      uint32_t word = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        q[k] *= 128.0;                // exact
        const double d = rint(q[k]);  // |d| <= 64
        q[k] -= d;                    // exact, |q| <= 1/2
        word |= (uint32_t)(uint8_t)(int8_t)(int)d << (8 * k);
      }
      *reinterpret_cast<uint32_t *>(planes + r * S * cols_p + (int64_t)(rev ? S - 1 - s : s) * cols_p + c) = word;
    }
  }
}

// One Ozaki level from its exact INT32 sums (the library's product): the
// epilogue of k_seg_gemm_big<1>, C (+)= 2^(e_i + f_j - 7 L) acc, so the same
// bits in the same level order.
__global__ __launch_bounds__(256) void k_ozaki_level(const int32_t *__restrict__ acc, int64_t lda, int64_t rows,
                                                     int64_t cols, const int32_t *__restrict__ ea,
                                                     const int32_t *__restrict__ fb, int level, int accumulate,
                                                     double *__restrict__ C, int64_t ldc) {
  const int64_t i = blockIdx.y, j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= rows || j >= cols) return;
  const double t = ldexp((double)acc[i * lda + j], ea[i] + fb[j] - 7 * level);
  double *c = C + i * ldc + j;
  *c = accumulate ? *c + t : t;
}

// ---- Ozaki CRT prep and combine ------------------------------------------------------------
constexpr int kMaxMods = 15;   // W_l < M < 2^118 fits 6 limbs of 20 bits; p, q <= 52
constexpr int kLimbs = 7;      // S = sum r_l W_l < 2^130 needs 7
constexpr uint32_t kLimbMask = (1u << 20) - 1;

struct CrtMods {
  int32_t n;
  int32_t m[kMaxMods];
  double inv[kMaxMods];  // 1 / m_l, rounded
};

// n residue planes, each rows_p x cols_p int8: the integer rint(X 2^(scale -
// e[r])) reduced to the symmetric range of m_l ([-128, 127] for 256,
// [-(m-1)/2, (m-1)/2] for odd m).  trans as in k_slice_i8 (B's planes are
// stored transposed); e[r] is the exponent of plane row r (ilogb(max) + 2, or
// 0 for an all-zero line).  The scaled integer has at most 52 bits, so
// a - m rint(a/m) is exact (the quotient may be off by one; one correction).
__global__ __launch_bounds__(256) void k_residues_i8(const double *__restrict__ X, int64_t xr, int64_t xc,
                                                     int64_t ldx, int trans, const int32_t *__restrict__ e, int scale,
                                                     const CrtMods mods, int8_t *__restrict__ planes, int64_t rows_p,
                                                     int64_t cols_p) {
  // a workgroup makes 32 plane rows x 128 plane columns (rows_p a multiple
  // of 32, cols_p of 128); a thread 4 consecutive columns of 4 rows, stored
  // as one dword per modulus
  __shared__ double t[128][33];
  const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 128;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  if (trans) {
    for (int i = ty; i < 128; i += 8) {
      const int64_t xrow = c0 + i, xcol = r0 + tx;
      t[i][tx] = (xrow < xr && xcol < xc) ? X[xrow * ldx + xcol] : 0.0;
    }
    __syncthreads();
  }
  const int64_t plane = rows_p * cols_p;
  const int64_t c = c0 + 4 * tx;
  for (int i = ty; i < 32; i += 8) {
    const int64_t r = r0 + i;
    if (r >= rows_p || c >= cols_p) continue;
    const int er = r < (trans ? xc : xr) ? e[r] : 0;
    double a[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const double v = trans ? t[4 * tx + k][i] : ((r < xr && c + k < xc) ? X[r * ldx + c + k] : 0.0);
      a[k] = rint(ldexp(v, scale - er));
    }
    for (int l = 0; l < mods.n; l++) {
      const double m = (double)mods.m[l];
      const double hi = (double)((mods.m[l] - 1) / 2), lo = -(double)(mods.m[l] / 2);
      uint32_t word = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        double res = fma(-m, rint(a[k] * mods.inv[l]), a[k]);
        res -= res > hi ? m : 0.0;
        res += res < lo ? m : 0.0;
        word |= (uint32_t)(uint8_t)(int8_t)(int)res << (8 * k);
      }
      *reinterpret_cast<uint32_t *>(planes + (int64_t)l * plane + r * cols_p + c) = word;
    }
  }
}

struct CrtCombine {
  int32_t n;
  int32_t m[kMaxMods];               // the moduli
  double minv[kMaxMods];             // 1 / m_l, rounded
  uint32_t w[kMaxMods][kLimbs - 1];  // W_l in 20-bit limbs, least significant first
  uint32_t mlimb[kLimbs];            // M
  uint32_t half[kLimbs];             // ceil(M / 2)
  double inv_m;                      // 1 / M, rounded
  int32_t p1, q1;                    // p + 1, q + 1
};

// r >= s over kLimbs limbs (most significant last)
__device__ __forceinline__ bool limbs_ge(const uint32_t *r, const uint32_t *s) {
#pragma unroll
  for (int k = kLimbs - 1; k >= 0; k--)
    if (r[k] != s[k]) return r[k] > s[k];
  return true;
}

// C[i][j] = 2^(e_i + f_j - p - q - 2) * C', C' the representative in
// [-M/2, M/2) of sum_l r_l W_l (mod M), r_l the residue byte of plane l.
// Everything up to the conversion to FP64 is exact integer arithmetic.
// r_l: a residue byte in [0, m_l), or an exact INT32 product sum reduced
// here: the quotient from the FP64 product with 1/m (exact v, so off by at
// most one) and one correction each way
__device__ __forceinline__ uint32_t residue(uint8_t v, int32_t, double) { return v; }
__device__ __forceinline__ uint32_t residue(int32_t v, int32_t m, double minv) {
  // (wrapping unsigned arithmetic: the true difference lies in [-m, 2m))
  int32_t r = (int32_t)((uint32_t)v - (uint32_t)(int32_t)floor((double)v * minv) * (uint32_t)m);
  r += r < 0 ? m : 0;
  r -= r >= m ? m : 0;
  return (uint32_t)r;
}

template <typename RT>
__global__ __launch_bounds__(256) void k_crt_combine(const RT *__restrict__ R, int64_t plane, int64_t ldr,
                                                     int64_t rows, int64_t cols, const int32_t *__restrict__ ea,
                                                     const int32_t *__restrict__ fb, const CrtCombine P,
                                                     double *__restrict__ C, int64_t ldc) {
  const int64_t i = blockIdx.y, j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= rows || j >= cols) return;
  uint32_t acc[kLimbs] = {0, 0, 0, 0, 0, 0, 0};
  const RT *rp = R + i * ldr + j;
  for (int l = 0; l < P.n; l++) {
    const uint32_t r = residue(rp[(int64_t)l * plane], P.m[l], P.minv[l]);
#pragma unroll
    for (int k = 0; k < kLimbs - 1; k++) acc[k] += __umul24(r, P.w[l][k]);  // < 15 * 255 * 2^20 < 2^32
  }
#pragma unroll
  for (int k = 0; k < kLimbs - 1; k++) {
    acc[k + 1] += acc[k] >> 20;
    acc[k] &= kLimbMask;
  }
  // q = floor(S / M) from the top limbs (may be one off), then S - q M exactly
  const double sd = ldexp((double)acc[6], 120) + ldexp((double)acc[5], 100) + ldexp((double)acc[4], 80) +
                    ldexp((double)acc[3], 60);
  const int64_t q = (int64_t)floor(sd * P.inv_m);
  int64_t carry = 0;
#pragma unroll
  for (int k = 0; k < kLimbs; k++) {
    const int64_t t = (int64_t)acc[k] - q * (int64_t)P.mlimb[k] + carry;
    acc[k] = (uint32_t)(t & kLimbMask);
    carry = t >> 20;  // arithmetic shift: floor division
  }
  if (carry < 0) {  // q one too large: add M back
    carry = 0;
#pragma unroll
    for (int k = 0; k < kLimbs; k++) {
      const int64_t t = (int64_t)acc[k] + P.mlimb[k] + carry;
      acc[k] = (uint32_t)(t & kLimbMask);
      carry = t >> 20;
    }
  } else if (limbs_ge(acc, P.mlimb)) {  // q one too small
    int64_t b = 0;
#pragma unroll
    for (int k = 0; k < kLimbs; k++) {
      const int64_t t = (int64_t)acc[k] - P.mlimb[k] + b;
      acc[k] = (uint32_t)(t & kLimbMask);
      b = t >> 20;
    }
  }
  // representative in [-M/2, M/2)
  const bool neg = limbs_ge(acc, P.half);
  if (neg) {
    int64_t b = 0;
#pragma unroll
    for (int k = 0; k < kLimbs; k++) {
      const int64_t t = (int64_t)P.mlimb[k] - acc[k] + b;
      acc[k] = (uint32_t)(t & kLimbMask);
      b = t >> 20;
    }
  }
  double v = (double)acc[kLimbs - 1];
#pragma unroll
  for (int k = kLimbs - 2; k >= 0; k--) v = fma(v, 1048576.0, (double)acc[k]);
  C[i * ldc + j] = ldexp(neg ? -v : v, ea[i] + fb[j] - P.p1 - P.q1);
}

// ---- the segmented GEMM -----------------------------------------------------------------
template <int MODE>
struct Acc;
template <>
struct Acc<0> {
  typedef f32x4 t;
};
template <>
struct Acc<1> {
  typedef i32x4 t;
};

template <int MODE>
__device__ __forceinline__ typename Acc<MODE>::t mma(const i32x4 &a, const i32x4 &b, typename Acc<MODE>::t c);
template <>
__device__ __forceinline__ f32x4 mma<0>(const i32x4 &a, const i32x4 &b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}
template <>
__device__ __forceinline__ i32x4 mma<1>(const i32x4 &a, const i32x4 &b, i32x4 c) {
  return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_seg_gemm(const SegArgs P) {
  __shared__ __attribute__((aligned(16))) uint8_t As[2][TM * ROWB];
  __shared__ __attribute__((aligned(16))) uint8_t Bs[2][TN * ROWB];
  typedef typename Acc<MODE>::t acc_t;
  const int tm = blockIdx.x / P.tiles_n, tn = blockIdx.x % P.tiles_n;
  const int64_t m0 = (int64_t)tm * TM, n0 = (int64_t)tn * TN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  // global -> LDS: 128 rows x 128 bytes per operand = 1024 x 16B; 4 per thread
  int lr[4], lc[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int idx = tid + 256 * i;
    lr[i] = idx >> 3;
    lc[i] = (idx & 7) * 16;
  }
  i32x4 ra[4], rb[4];
  const int ktiles = (int)(P.Kb / KB);
  const int total = P.nseg * ktiles;
  // stage fetched next: segment lseg, byte lk of K (advanced without divisions)
  int lseg = 0;
  int64_t lk = 0;
  auto gload = [&]() {
    const Seg &s = P.seg[lseg];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      ra[i] = *reinterpret_cast<const i32x4 *>(P.A + s.a_off + (m0 + lr[i]) * P.lda + lk + lc[i]);
      rb[i] = *reinterpret_cast<const i32x4 *>(P.B + s.b_off + (n0 + lr[i]) * P.ldb + lk + lc[i]);
    }
    lk += KB;
    if (lk == P.Kb) {
      lk = 0;
      lseg++;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      *reinterpret_cast<i32x4 *>(&As[buf][lr[i] * ROWB + lc[i]]) = ra[i];
      *reinterpret_cast<i32x4 *>(&Bs[buf][lr[i] * ROWB + lc[i]]) = rb[i];
    }
  };
  acc_t acc[4][4];
  double acc64[4][4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      acc[i][j] = acc_t{0, 0, 0, 0};
      if (MODE == 1)
#pragma unroll
        for (int v = 0; v < 4; v++) acc64[i][j][v] = 0.0;
    }
  gload();
  sstore(0);
  __syncthreads();
  int cseg = 0, ck = 0;  // segment and k-stage being multiplied
  for (int it = 0; it < total; it++) {
    const int cur = it & 1;
    if (it + 1 < total) gload();
#pragma unroll
    for (int kk = 0; kk < KB; kk += 64) {
      i32x4 a[4], b[4];
      const int kbyte = kk + (lane >> 4) * 16;
#pragma unroll
      for (int i = 0; i < 4; i++)
        a[i] = *reinterpret_cast<const i32x4 *>(&As[cur][(wr * 64 + i * 16 + (lane & 15)) * ROWB + kbyte]);
#pragma unroll
      for (int j = 0; j < 4; j++)
        b[j] = *reinterpret_cast<const i32x4 *>(&Bs[cur][(wc * 64 + j * 16 + (lane & 15)) * ROWB + kbyte]);
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = mma<MODE>(a[i], b[j], acc[i][j]);
    }
    if (MODE == 1) {
      const Seg &s = P.seg[cseg];
      if (s.level_end && ck == ktiles - 1) {
        const double w = ldexp(1.0, -7 * s.level);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) {
#pragma unroll
            for (int v = 0; v < 4; v++) acc64[i][j][v] += w * (double)acc[i][j][v];
            acc[i][j] = acc_t{0, 0, 0, 0};
          }
      }
    }
    if (++ck == ktiles) {
      ck = 0;
      cseg++;
    }
    if (it + 1 < total) sstore(cur ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int64_t col = n0 + wc * 64 + j * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int64_t row = m0 + wr * 64 + i * 16 + 4 * (lane >> 4) + v;
        if (row < P.M && col < P.N) {
          if (MODE == 0) reinterpret_cast<float *>(P.C)[row * P.ldc + col] = (float)acc[i][j][v];
          else reinterpret_cast<double *>(P.C)[row * P.ldc + col] = ldexp(acc64[i][j][v], P.ea[row] + P.fb[col]);
        }
      }
    }
}

// ---- the segmented GEMM, large tiles --------------------------------------------------------
// 256x256 output tile per workgroup, 4 wavefronts of 128x128 (4x4 MFMA 32x32
// tiles, 256 accumulators per lane in AGPRs).  K advances 128 operand bytes
// per stage (four 32-byte MFMA k-steps, 64 MFMAs per wavefront).  Operand
// rows reach LDS by direct global->LDS loads (global_load_lds_dwordx4: each
// lane's 16 bytes land at lane * 16 from a wave-uniform LDS address), so no
// register staging or LDS store instructions sit in the MFMA stream.  Two
// 64 KB buffers (A 256 rows | B 256 rows, 128 bytes a row, unpadded) alternate:
// the loads of stage s+1 are issued during the first two k-steps of stage s,
// two after each group of 4 MFMAs, into the buffer stage s-1 used; one wait
// for them and one barrier start each stage.  Rows are unpadded, so the
// 16-byte chunk c of row r is kept at chunk slot c ^ sw(r) (sw(r) = (r ^
// (r >> 3)) & 7): the eight rows one fragment read touches per bank cycle then
// fall on eight different slots (no bank conflicts for either 32 or 64 banks).
// On 8192^3 split-bf16 panels (random data) this ran at 1.19-1.27 PFLOP/s,
// against 0.85 with register staging and padded rows.  Tiles are grouped by
// XCD (consecutive workgroups go to different XCDs; each XCD gets a
// contiguous run of tiles) and, within that run, in 4-row column-major groups
// so concurrent tiles share A and B panels in L2.
// In OZAKI mode one launch handles one level (its segments share the level);
// the epilogue adds 2^(e_i+f_j-7L) * acc into C (or writes it, for the first
// level), so levels are still summed from the least significant one up.
constexpr int BM = 256, BN = 256, KB2 = 128;

// chunk slot of a row's 16-byte chunks in the unpadded LDS rows
__device__ __forceinline__ int chunk_swizzle(int r) { return (r ^ (r >> 3)) & 7; }

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef int i32x16 __attribute__((ext_vector_type(16)));

template <int MODE>
struct Acc32;
template <>
struct Acc32<0> {
  typedef f32x16 t;
};
template <>
struct Acc32<1> {
  typedef i32x16 t;
};

template <int MODE>
__device__ __forceinline__ typename Acc32<MODE>::t mma32(const i32x4 &a, const i32x4 &b, typename Acc32<MODE>::t c);
template <>
__device__ __forceinline__ f32x16 mma32<0>(const i32x4 &a, const i32x4 &b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}
template <>
__device__ __forceinline__ i32x16 mma32<1>(const i32x4 &a, const i32x4 &b, i32x16 c) {
  return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
}
template <>
struct Acc32<2> {
  typedef i32x16 t;
};
template <>
__device__ __forceinline__ i32x16 mma32<2>(const i32x4 &a, const i32x4 &b, i32x16 c) {
  return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_seg_gemm_big(const SegArgs P, int level, int accumulate, int tiles_m) {
  constexpr int TILE = BM * KB2, BUF = 2 * TILE;  // one operand's rows; one stage
  __shared__ __attribute__((aligned(1024))) uint8_t lds[2 * BUF];  // [2][A 256 rows | B 256 rows] x 128 bytes
  typedef typename Acc32<MODE>::t acc_t;
  // XCD x gets nt / 8 tiles (one more when x < nt % 8) in a contiguous band
  // (OZAKI_CRT: nt tiles per modulus, the modulus-major list banded the same
  // way, so an XCD works on one or two moduli and shares their planes in L2)
  const int nt = tiles_m * P.tiles_n;
  const int ntot = MODE == 2 ? nt * P.nmod : nt;
  const int xq = ntot / 8, xr = ntot % 8, xcd = (int)blockIdx.x % 8;
  int id = xcd * xq + min(xcd, xr) + (int)blockIdx.x / 8;
  int lmod = 0;
  if (MODE == 2) {
    lmod = id / nt;
    id -= lmod * nt;
  }
  // grouped order: consecutive ids walk 4 tile rows, then the next column, so
  // the tiles running together share few A and B panels
  constexpr int GM = 4;
  const int group = id / (GM * P.tiles_n);
  const int first_m = group * GM;
  const int gm = min(tiles_m - first_m, GM);
  const int within = id - group * GM * P.tiles_n;
  const int tm = first_m + within % gm, tn = within / gm;
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int ktiles = (int)(P.Kb / KB2);
  const int total = P.nseg * ktiles;
  // loads: wave w brings rows [64w, 64w + 64) of A and of B, 8 rows (1 KB) per
  // load; lane l takes row 8i + l/8 of the group and fills chunk slot l % 8
  // with the row's chunk (l % 8) ^ sw(row)
  int64_t goffA[8], goffB[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int row = wave * 64 + i * 8 + (lane >> 3);
    const int c = (lane & 7) ^ chunk_swizzle(row);
    goffA[i] = (m0 + row) * P.lda + c * 16;
    goffB[i] = (n0 + row) * P.ldb + c * 16;
  }
  // stage loaded next: segment lseg, byte lk of K (advanced without divisions)
  int lseg = 0;
  int64_t lk = 0;
  const uint8_t *ga = nullptr, *gb = nullptr;
  uint8_t *gd = nullptr;
  auto begin_stage = [&](int buf) {
    const int64_t a_off = MODE == 2 ? lmod * P.a_plane : P.seg[lseg].a_off;
    const int64_t b_off = MODE == 2 ? lmod * P.b_plane : P.seg[lseg].b_off;
    ga = P.A + a_off + lk;
    gb = P.B + b_off + lk;
    gd = lds + buf * BUF + wave * 64 * KB2;
    lk += KB2;
    if (lk == P.Kb) {
      lk = 0;
      lseg++;
    }
    // past the last stage the last one is loaded again into the free buffer
    // (never read): every stage issues the same loads, so the waits are uniform
    if (lseg == P.nseg) {
      lseg = P.nseg - 1;
      lk = P.Kb - KB2;
    }
  };
  auto load = [&](int q) {  // q in [0, 16): A rows q/2 (q even) or B rows q/2
    const int i = q >> 1;
    if (q & 1)
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(gb + goffB[i]),
                                       (__attribute__((address_space(3))) void *)(gd + TILE + i * 8 * KB2), 16, 0, 0);
    else
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(ga + goffA[i]),
                                       (__attribute__((address_space(3))) void *)(gd + i * 8 * KB2), 16, 0, 0);
  };
  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = acc_t{};
  const int rsub = lane & 31, half = lane >> 5;
  // fragment i of this lane: row 128 wr + 32 i + rsub (A) or 128 wc + 32 i +
  // rsub (B), bytes [32 step + 16 half, +16) of the stage: chunk 2 step + half
  // at slot (2 step + half) ^ sw(row), i.e. byte offset fa_off ^ (step << 5)
  int fa_off[4], fb_off[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int ra = wr * 128 + i * 32 + rsub, rb = wc * 128 + i * 32 + rsub;
    fa_off[i] = ra * KB2 + ((half ^ chunk_swizzle(ra)) << 4);
    fb_off[i] = TILE + rb * KB2 + ((half ^ chunk_swizzle(rb)) << 4);
  }
  // multiply the stage held in buffer `cur`; the loads of the next stage go
  // out two per group of 4 MFMAs during the first two k-steps
  auto compute = [&](int cur) {
    const uint8_t *base = lds + cur * BUF;
    // fragments of k-step ks+1 are read from LDS while the 16 MFMAs of k-step
    // ks run (two register sets), so LDS latency is hidden behind MFMA work
    i32x4 fa[2][4], fb[2][4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      fa[0][i] = *reinterpret_cast<const i32x4 *>(base + fa_off[i]);
      fb[0][i] = *reinterpret_cast<const i32x4 *>(base + fb_off[i]);
    }
#pragma unroll
    for (int step = 0; step < KB2 / 32; step++) {
      const int c = step & 1;
      if (step + 1 < KB2 / 32) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
          fa[c ^ 1][i] = *reinterpret_cast<const i32x4 *>(base + (fa_off[i] ^ ((step + 1) << 5)));
          fb[c ^ 1][i] = *reinterpret_cast<const i32x4 *>(base + (fb_off[i] ^ ((step + 1) << 5)));
        }
      }
      // keep the reads above the MFMAs: the scheduler would otherwise sink
      // them next to their use and expose LDS latency every 4 MFMAs
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < 4; i++) {
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = mma32<MODE>(fa[c][i], fb[c][j], acc[i][j]);
        if (step < 2) {
          load(step * 8 + 2 * i);
          load(step * 8 + 2 * i + 1);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  };
  begin_stage(0);
#pragma unroll
  for (int q = 0; q < 16; q++) load(q);
  for (int s = 0; s < total; s++) {
    // this wave's loads of stage s have landed (vmcnt 0); after the barrier
    // every wave's have, and every wave is done reading stage s-1's buffer
    __builtin_amdgcn_s_waitcnt(0x0f70);
    __syncthreads();
    begin_stage((s + 1) & 1);
    compute(s & 1);
  }
  __builtin_amdgcn_s_waitcnt(0x0f70);  // no load lands in LDS after the workgroup ends
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int64_t col = n0 + wc * 128 + j * 32 + rsub;
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const int64_t row = m0 + wr * 128 + i * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (row < P.M && col < P.N) {
          if (MODE == 0) {
            reinterpret_cast<float *>(P.C)[row * P.ldc + col] = (float)acc[i][j][v];
          } else if (MODE == 2) {
            // exact residue of the INT32 sum in [0, m)
            const int m = P.mods[lmod];
            const int a = (int)acc[i][j][v];
            int r = a - m * (int)floor((double)a / (double)m);
            r += r < 0 ? m : 0;
            r -= r >= m ? m : 0;
            reinterpret_cast<uint8_t *>(P.C)[lmod * P.c_plane + row * P.ldc + col] = (uint8_t)r;
          } else {
            double *c = reinterpret_cast<double *>(P.C) + row * P.ldc + col;
            const double t = ldexp((double)acc[i][j][v], P.ea[row] + P.fb[col] - 7 * level);
            *c = accumulate ? *c + t : t;
          }
        }
      }
      // one 32x32 tile at a time: keeps the scheduler from holding every
      // C value of the epilogue in registers at once
      __builtin_amdgcn_sched_barrier(0);
    }
}

// ---- Ozaki slice tuner helpers ---------------------------------------------------------------
// rows r*M/R of A (R x K, contiguous) and their absolute values
__global__ __launch_bounds__(256) void k_sample_rows(const double *__restrict__ A, int64_t lda, int64_t M, int64_t K,
                                                     int R, double *__restrict__ As, double *__restrict__ Aa) {
  const int r = blockIdx.y;
  const int64_t src = (int64_t)r * M / R;
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < K; k += (int64_t)gridDim.x * 256) {
    const double v = A[src * lda + k];
    As[(int64_t)r * K + k] = v;
    Aa[(int64_t)r * K + k] = fabs(v);
  }
}

__global__ __launch_bounds__(256) void k_abs_rows(const double *__restrict__ B, int64_t ldb, int64_t rows, int64_t cols,
                                                  double *__restrict__ Ba) {
  const int64_t n = rows * cols;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const int64_t r = i / cols, c = i - r * cols;
    Ba[i] = fabs(B[r * ldb + c]);
  }
}

// max over entries of |X - Ref| / den (den > 0), folded into *out as the bit
// pattern of a non-negative double (which orders like the unsigned integer)
__global__ __launch_bounds__(256) void k_err_max(const double *__restrict__ X, const double *__restrict__ Ref,
                                                 const double *__restrict__ den, int64_t n,
                                                 unsigned long long *__restrict__ out) {
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    if (den[i] > 0.0) m = fmax(m, fabs(X[i] - Ref[i]) / den[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }

// C[0:M, 0:N] = 0 (row stride ldc elements of T)
template <typename T>
__global__ __launch_bounds__(256) void k_zero_rows(T *__restrict__ C, int64_t M, int64_t N, int64_t ldc) {
  const int64_t n = M * N;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const int64_t r = i / N;
    C[r * ldc + (i - r * N)] = T(0);
  }
}

int64_t rup(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

// ---- autotuner ------------------------------------------------------------------------------
// The first product of a given (form, M, N, K) class (each dimension rounded
// up to a power of two) runs with every candidate -- the hand-written MFMA
// kernel with 128x128 or 256x256 tiles, or hipBLASLt (kLib) -- once untimed,
// then once timed with HIP events on the library stream.  The fastest is
// kept for that class and, so that every call of a class returns the same
// bits, re-run once if it was not the last one timed.  A candidate that
// fails to run (the library offering no algorithm for the shape) drops out.
// IAMX_TILE=128, 256 or lib pins the choice where it is a candidate.
constexpr int kLib = 0;
struct TuneEntry {
  int form;
  int64_t m, n, k;
  int pick;
};
static TuneEntry g_tune[128];
static int g_ntune = 0;

static int64_t pow2_up(int64_t v) {
  int64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

// read on every product, so a caller can pin a candidate for some products only
static int forced_pick() {
  const char *e = getenv("IAMX_TILE");
  if (!e) return -1;
  if (!strcmp(e, "lib")) return kLib;
  const int v = atoi(e);
  return (v == 128 || v == 256) ? v : -1;
}

}  // namespace

int iamx::tune_run(int form, int64_t M, int64_t N, int64_t K, const int *cands, int ncand, int (*run)(int, void *),
                   void *ctx, int *rc) {
  int c[4], nc = 0;
  for (int i = 0; i < ncand && nc < 4; i++)
    if (cands[i] != kLib || blt_usable()) c[nc++] = cands[i];
  if (nc == 0) {
    *rc = set_error(IAMX_EINVAL, "autotuner: no candidate");
    return -1;
  }
  const int f = forced_pick();
  for (int i = 0; i < nc; i++)
    if (c[i] == f) {
      *rc = run(f, ctx);
      return f;
    }
  if (nc == 1) {
    *rc = run(c[0], ctx);
    return c[0];
  }
  const int64_t m = pow2_up(M), n = pow2_up(N), k = pow2_up(K);
  for (int i = 0; i < g_ntune; i++)
    if (g_tune[i].form == form && g_tune[i].m == m && g_tune[i].n == n && g_tune[i].k == k) {
      // the class's pick among the candidates of this call (the library is
      // none under a capture; a class may hold picks of two candidate lists)
      for (int j = 0; j < nc; j++)
        if (c[j] == g_tune[i].pick) {
          *rc = run(c[j], ctx);
          return c[j];
        }
    }
  if (capturing()) {  // no timing inside a graph capture: the first hand-written candidate
    for (int i = 0; i < nc; i++)
      if (c[i] != kLib) {
        *rc = run(c[i], ctx);
        return c[i];
      }
    *rc = run(c[0], ctx);
    return c[0];
  }
  hipEvent_t ev[5];
  int made = 0;
  for (; made < nc + 1; made++)
    if (hipEventCreate(&ev[made]) != hipSuccess) break;
  if (made < nc + 1) {
    for (int i = 0; i < made; i++) (void)hipEventDestroy(ev[i]);
    *rc = run(c[0], ctx);
    return c[0];
  }
  // one untimed run each first: the library loads its kernels and picks its
  // algorithm on its first call of a shape, which must not count
  for (int i = 0; i < nc; i++) (void)run(c[i], ctx);
  int ok[4] = {0, 0, 0, 0}, last = -1;
  (void)hipEventRecord(ev[0], g_cur);
  for (int i = 0; i < nc; i++) {
    ok[i] = run(c[i], ctx) == IAMX_OK;
    if (ok[i]) last = i;
    (void)hipEventRecord(ev[i + 1], g_cur);
  }
  int best = -1;
  float best_ms = 0.f;
  if (last >= 0 && hipEventSynchronize(ev[nc]) == hipSuccess) {
    for (int i = 0; i < nc; i++) {
      float ms = 0.f;
      if (!ok[i] || hipEventElapsedTime(&ms, ev[i], ev[i + 1]) != hipSuccess) continue;
      if (best < 0 || ms < best_ms) {
        best = i;
        best_ms = ms;
      }
    }
  }
  for (int i = 0; i < nc + 1; i++) (void)hipEventDestroy(ev[i]);
  if (best < 0) {
    *rc = last < 0 ? run(c[0], ctx) : IAMX_OK;  // the error of the first candidate, or the last that ran
    return last < 0 ? c[0] : c[last];
  }
  *rc = best != last ? run(c[best], ctx) : IAMX_OK;
  if (!*rc && g_ntune < 128) g_tune[g_ntune++] = TuneEntry{form, m, n, k, c[best]};
  return c[best];
}

namespace {

// run(candidate) enqueues the product and returns an IAMX status; returns
// the candidate used.
template <typename F>
static int tuned(int form, int64_t M, int64_t N, int64_t K, std::initializer_list<int> cands, F run, int *rc) {
  const int list[4] = {cands.size() > 0 ? cands.begin()[0] : 0, cands.size() > 1 ? cands.begin()[1] : 0,
                       cands.size() > 2 ? cands.begin()[2] : 0, cands.size() > 3 ? cands.begin()[3] : 0};
  return iamx::tune_run(form, M, N, K, list, (int)cands.size(), [](int c, void *x) { return (*(F *)x)(c); },
                        &run, rc);
}

// ---- prepared B operand and row panels -------------------------------------------------------
// The B side of a mixed product (its bf16 hi/lo planes, or its column
// exponents and int8 digit planes) is built once by mixed_prepare_b into a
// workspace slot of its own; mixed_rows then multiplies any row panel of A by
// it (the panel's planes in slot 11), so C can be produced panel by panel.
// One-shot products keep their B in slot 1 and the row-panel interface in
// slot 10, so a one-shot product between two panels leaves the panels' B intact.
struct PrepB {
  int form;  // 1 split-bf16, 2 ozaki slices, 3 ozaki crt
  int64_t N, K, Np, Kp;
  int S;                  // slices (form 2) or moduli (form 3)
  const uint8_t *planes;  // B planes (transposed), then the column exponents (ozaki)
  const int32_t *fb;
  bool valid;
  int p, q;               // form 3: integer bits of the scaled A rows and B columns
};

// ---- CRT constants (host) ------------------------------------------------------------------
// The moduli, largest first: pairwise coprime (256 = 2^8, 255 = 3 5 17,
// 253 = 11 23, 247 = 13 19, 217 = 7 31, the rest prime).
static const int kMods[kMaxMods] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197};

// p and q for n moduli and depth K: the largest p + q with K 2^(p+q) < M/2
// (so sum_k |A'_ik B'_kj| < M/2 and the representative in [-M/2, M/2) is
// the product), split evenly, each at most 52 so the scaled integers and their
// residue arithmetic stay exact in FP64.
static int crt_bits(int n, int64_t K, int *p, int *q) {
  if (n < 2 || n > kMaxMods) return iamx::set_error(IAMX_EINVAL, "ozaki-crt: moduli %d outside 2..%d", n, kMaxMods);
  double log2m = 0.0;
  for (int l = 0; l < n; l++) log2m += log2((double)kMods[l]);
  const int budget = (int)floor(log2m - 1.0 - log2((double)K) - 1e-9);
  *p = budget / 2 < 52 ? budget / 2 : 52;
  *q = budget - *p < 52 ? budget - *p : 52;
  if (*p < 1 || *q < 1) return iamx::set_error(IAMX_EINVAL, "ozaki-crt: %d moduli too few for K=%lld", n, (long long)K);
  return IAMX_OK;
}

static CrtMods crt_mods(int n) {
  CrtMods m;
  memset(&m, 0, sizeof(m));
  m.n = n;
  for (int l = 0; l < n; l++) {
    m.m[l] = kMods[l];
    m.inv[l] = 1.0 / (double)kMods[l];
  }
  return m;
}

static void to_limbs(unsigned __int128 v, uint32_t *limb, int nl) {
  for (int k = 0; k < nl; k++) {
    limb[k] = (uint32_t)(v & kLimbMask);
    v >>= 20;
  }
}

static CrtCombine crt_combine_consts(int n, int p, int q) {
  CrtCombine c;
  memset(&c, 0, sizeof(c));
  c.n = n;
  unsigned __int128 M = 1;
  for (int l = 0; l < n; l++) M *= (unsigned)kMods[l];
  for (int l = 0; l < n; l++) {
    const unsigned m = (unsigned)kMods[l];
    const unsigned __int128 Ml = M / m;
    const unsigned r = (unsigned)(Ml % m);
    unsigned inv = 1;
    while ((r * inv) % m != 1) inv++;  // m <= 256: a direct search
    to_limbs((Ml * inv) % M, c.w[l], kLimbs - 1);
  }
  to_limbs(M, c.mlimb, kLimbs);
  to_limbs(M / 2, c.half, kLimbs);
  c.inv_m = 1.0 / (double)M;
  for (int l = 0; l < n; l++) {
    c.m[l] = kMods[l];
    c.minv[l] = 1.0 / (double)kMods[l];
  }
  c.p1 = p + 1;
  c.q1 = q + 1;
  return c;
}
static PrepB g_pb;  // the row-panel interface's B

static int mixed_prepare_b(PrepB &pb, int slot, int form, int64_t N, int64_t K, const void *B, int64_t ldb, int S) {
  pb.valid = false;
  const int64_t Np = rup(N, BN);
  if (form == 1) {
    const int64_t Kp = rup(K, KB / 2);
    const size_t b_bytes = (size_t)(Np * Kp * 2);
    uint8_t *ws = (uint8_t *)iamx_workspace(slot, 3 * b_bytes + 256);
    if (!ws) return iamx::set_error(IAMX_ENOMEM, "split-bf16 workspace");
    // rows of B^T: [hi | lo | hi]
    uint16_t *w16 = (uint16_t *)ws;
    IAMX_LAUNCH("split_bf16_B", k_split_bf16, dim3((unsigned)(Kp / 32), (unsigned)(Np / 32)), dim3(256), 0,
                (const float *)B, K, N, ldb, 1, w16, w16 + Kp, w16 + 2 * Kp, Np, Kp, 3 * Kp);
    pb = PrepB{1, N, K, Np, Kp, 0, ws, nullptr, true, 0, 0};
    return IAMX_OK;
  }
  if (form == 3) {
    int p, q;
    const int rc = crt_bits(S, K, &p, &q);
    if (rc) return rc;
    if (K > 131071) return iamx::set_error(IAMX_EINVAL, "ozaki-crt: K=%lld too large for exact INT32 sums", (long long)K);
    const int64_t Kp = rup(K, KB);
    const size_t b_bytes = (size_t)(S * Np * Kp);
    uint8_t *ws = (uint8_t *)iamx_workspace(slot, b_bytes + (size_t)Np * 4 + 256);
    if (!ws) return iamx::set_error(IAMX_ENOMEM, "ozaki-crt workspace");
    int32_t *fb = (int32_t *)(ws + b_bytes);
    const double *Bd = (const double *)B;
    IAMX_LAUNCH("ozaki_col_exp", k_fill_i32, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, fb, N, INT32_MIN);
    IAMX_LAUNCH("ozaki_col_exp", k_col_exp,
                dim3((unsigned)((N + 63) / 64), (unsigned)((K + kColSlab - 1) / kColSlab)), dim3(256), 0, Bd, K, N,
                ldb, fb);
    IAMX_LAUNCH("ozaki_col_exp", k_col_exp_zero, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, fb, N);
    IAMX_LAUNCH("ozaki_crt_residues_B", k_residues_i8, dim3((unsigned)(Kp / 128), (unsigned)(Np / 32)), dim3(256), 0,
                Bd, K, N, ldb, 1, fb, q + 1, crt_mods(S), (int8_t *)ws, Np, Kp);
    pb = PrepB{3, N, K, Np, Kp, S, ws, fb, true, p, q};
    return IAMX_OK;
  }
  if (S < 2 || S > 8) return iamx::set_error(IAMX_EINVAL, "ozaki: slices %d outside 2..8", S);
  if ((double)K * S * 64.0 * 64.0 >= 2147483647.0)
    return iamx::set_error(IAMX_EINVAL, "ozaki: K=%lld too large for exact INT32 level sums", (long long)K);
  const int64_t Kp = rup(K, KB);
  const size_t b_bytes = (size_t)(S * Np * Kp);
  uint8_t *ws = (uint8_t *)iamx_workspace(slot, b_bytes + (size_t)Np * 4 + 256);
  if (!ws) return iamx::set_error(IAMX_ENOMEM, "ozaki workspace");
  int32_t *fb = (int32_t *)(ws + b_bytes);
  const double *Bd = (const double *)B;
  IAMX_LAUNCH("ozaki_col_exp", k_fill_i32, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, fb, N, INT32_MIN);
  IAMX_LAUNCH("ozaki_col_exp", k_col_exp, dim3((unsigned)((N + 63) / 64), (unsigned)((K + kColSlab - 1) / kColSlab)),
              dim3(256), 0, Bd, K, N, ldb, fb);
  IAMX_LAUNCH("ozaki_col_exp", k_col_exp_zero, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, fb, N);
  IAMX_LAUNCH("ozaki_slice_B", k_slice_i8, dim3((unsigned)(Kp / 128), (unsigned)(Np / 32)), dim3(256), 0, Bd, K, N,
              ldb, 1, fb, S, (int8_t *)ws, Np, Kp, 1);
  pb = PrepB{2, N, K, Np, Kp, S, ws, fb, true, 0, 0};
  return IAMX_OK;
}

// C[0:M, :] = A[0:M, :] B for the prepared B
static int mixed_rows(const PrepB &pb, int form, int64_t M, const void *A, int64_t lda, void *C, int64_t ldc) {
  if (!pb.valid || pb.form != form) return iamx::set_error(IAMX_EINVAL, "mixed product: B not prepared");
  const int64_t N = pb.N, K = pb.K, Np = pb.Np, Kp = pb.Kp;
  const int64_t Mp = rup(M, BM);
  if (K == 0) {  // an empty sum: C is zero, and no product kernel runs with no K stage
    const int64_t blocks = (M * N + 255) / 256;
    const dim3 grid((unsigned)(blocks < 4096 ? blocks : 4096));
    if (form == 1)
      IAMX_LAUNCH("mixed_empty_k", k_zero_rows<float>, grid, dim3(256), 0, (float *)C, M, N, ldc);
    else
      IAMX_LAUNCH("mixed_empty_k", k_zero_rows<double>, grid, dim3(256), 0, (double *)C, M, N, ldc);
    return IAMX_OK;
  }
  SegArgs P;
  memset(&P, 0, sizeof(P));
  P.B = pb.planes;
  P.M = M;
  P.N = N;
  P.C = C;
  P.ldc = ldc;
  int rc = IAMX_OK;
  if (form == 1) {
    const size_t a_bytes = (size_t)(Mp * Kp * 2);
    uint8_t *ws = (uint8_t *)iamx_workspace(11, 3 * a_bytes + 256);
    if (!ws) return iamx::set_error(IAMX_ENOMEM, "split-bf16 workspace");
    // rows of A: [lo | hi | hi] against B^T's [hi | lo | hi]: the three
    // products, small ones first (lo*hi, hi*lo, then hi*hi), are one product
    // over K stacked three times
    uint16_t *w16 = (uint16_t *)ws;
    IAMX_LAUNCH("split_bf16_A", k_split_bf16, dim3((unsigned)(Kp / 32), (unsigned)(Mp / 32)), dim3(256), 0,
                (const float *)A, M, K, lda, 0, w16 + Kp, w16, w16 + 2 * Kp, Mp, Kp, 3 * Kp);
    P.A = ws;
    P.lda = 3 * Kp * 2;
    P.ldb = 3 * Kp * 2;
    P.Kb = Kp * 2;
    P.nseg = 3;
    P.seg[0] = Seg{0, 0, 0, 0};
    P.seg[1] = Seg{Kp * 2, Kp * 2, 0, 0};
    P.seg[2] = Seg{2 * Kp * 2, 2 * Kp * 2, 0, 0};
    const bool small = M < BM || N < BN;  // a 256-tile launch would be mostly padding
    tuned(0, M, N, K, small ? std::initializer_list<int>{128, kLib} : std::initializer_list<int>{256, 128, kLib},
          [&](int tile) -> int {
            if (tile == kLib)
              return iamx::blt_gemm(iamx::BLT_BF16_F32, 1, M, N, 3 * Kp, ws, 3 * Kp, pb.planes, 3 * Kp, C, ldc);
            if (tile == 256) {
              P.tiles_n = (int)(Np / BN);
              IAMX_LAUNCH("gemm_split_bf16_mfma", k_seg_gemm_big<0>, dim3((unsigned)((Mp / BM) * (Np / BN))),
                          dim3(256), 0, P, 0, 0, (int)(Mp / BM));
            } else {
              P.tiles_n = (int)(Np / TN);
              IAMX_LAUNCH("gemm_split_bf16_mfma", k_seg_gemm<0>, dim3((unsigned)((Mp / TM) * (Np / TN))),
                          dim3(256), 0, P);
            }
            return IAMX_OK;
          },
          &rc);
    return rc;
  }
  if (form == 3) {
    const int n = pb.S;
    // residue planes: bytes from the hand-written kernel, INT32 sums from hipBLASLt
    const size_t a_bytes = (size_t)(n * Mp * Kp), r_bytes = (size_t)(n * Mp * Np) * (iamx::blt_usable() ? 4 : 1);
    uint8_t *ws = (uint8_t *)iamx_workspace(11, a_bytes + (size_t)Mp * 4 + r_bytes + 512);
    if (!ws) return iamx::set_error(IAMX_ENOMEM, "ozaki-crt workspace");
    int32_t *ea = (int32_t *)(ws + a_bytes);
    uint8_t *R = ws + rup((int64_t)(a_bytes + (size_t)Mp * 4), 256);
    const double *Ad = (const double *)A;
    IAMX_LAUNCH("ozaki_row_exp", k_row_exp, dim3((unsigned)((M + 3) / 4)), dim3(256), 0, Ad, M, K, lda, ea);
    IAMX_LAUNCH("ozaki_crt_residues_A", k_residues_i8, dim3((unsigned)(Kp / 128), (unsigned)(Mp / 32)), dim3(256), 0,
                Ad, M, K, lda, 0, ea, pb.p + 1, crt_mods(n), (int8_t *)ws, Mp, Kp);
    P.A = ws;
    P.lda = Kp;
    P.ldb = Kp;
    P.Kb = Kp;
    P.nseg = 1;
    P.nmod = n;
    P.a_plane = Mp * Kp;
    P.b_plane = Np * Kp;
    P.c_plane = Mp * Np;
    for (int l = 0; l < n; l++) P.mods[l] = kMods[l];
    P.C = R;
    P.ldc = Np;
    P.tiles_n = (int)(Np / BN);
    const CrtCombine comb = crt_combine_consts(n, pb.p, pb.q);
    const dim3 cgrid((unsigned)((N + 255) / 256), (unsigned)M);
    // both candidates give the exact residues, so the same bits
    tuned(2, M, N, K * n, {256, kLib},
          [&](int tile) -> int {
            if (tile == kLib) {
              const int e = iamx::blt_gemm(iamx::BLT_I8_I32, 1, M, N, Kp, ws, Kp, pb.planes, Kp, R, Np, n, Mp * Kp,
                                           Np * Kp, Mp * Np, "gemm_ozaki_crt_i8_hipblaslt");
              if (e) return e;
              IAMX_LAUNCH("ozaki_crt_combine", k_crt_combine<int32_t>, cgrid, dim3(256), 0, (const int32_t *)R,
                          Mp * Np, Np, M, N, (const int32_t *)ea, pb.fb, comb, (double *)C, ldc);
              return IAMX_OK;
            }
            IAMX_LAUNCH("gemm_ozaki_crt_i8_mfma", k_seg_gemm_big<2>, dim3((unsigned)((Mp / BM) * (Np / BN) * n)),
                        dim3(256), 0, P, 0, 0, (int)(Mp / BM));
            IAMX_LAUNCH("ozaki_crt_combine", k_crt_combine<uint8_t>, cgrid, dim3(256), 0, (const uint8_t *)R, Mp * Np,
                        Np, M, N, (const int32_t *)ea, pb.fb, comb, (double *)C, ldc);
            return IAMX_OK;
          },
          &rc);
    return rc;
  }
  const int S = pb.S;
  const size_t a_bytes = (size_t)(S * Mp * Kp);
  // a level's INT32 sums when the library runs it
  const size_t l_bytes = iamx::blt_usable() ? (size_t)M * (size_t)N * 4 : 0;
  uint8_t *ws = (uint8_t *)iamx_workspace(11, a_bytes + (size_t)Mp * 4 + 256 + l_bytes + 256);
  if (!ws) return iamx::set_error(IAMX_ENOMEM, "ozaki workspace");
  int32_t *ea = (int32_t *)(ws + a_bytes);
  int32_t *lacc = (int32_t *)(ws + rup((int64_t)(a_bytes + (size_t)Mp * 4), 256));
  const double *Ad = (const double *)A;
  IAMX_LAUNCH("ozaki_row_exp", k_row_exp, dim3((unsigned)((M + 3) / 4)), dim3(256), 0, Ad, M, K, lda, ea);
  IAMX_LAUNCH("ozaki_slice_A", k_slice_i8, dim3((unsigned)(Kp / 128), (unsigned)(Mp / 32)), dim3(256), 0, Ad, M, K,
              lda, 0, ea, S, (int8_t *)ws, Mp, Kp, 0);
  // rows of A hold digits 1..S, rows of B^T digits S..1: the pair (s, t) of
  // a level sits at A column block s-1 and B column block S-t
  P.A = ws;
  P.lda = S * Kp;
  P.ldb = S * Kp;
  P.Kb = Kp;
  P.ea = ea;
  P.fb = pb.fb;
  // levels are added from the least significant (S+1) up to the most
  // significant (2), so small contributions are not lost to FP64 rounding first
  const bool small = M < BM || N < BN;
  const dim3 lgrid((unsigned)((N + 255) / 256), (unsigned)M);
  tuned(1, M, N, K * S, small ? std::initializer_list<int>{128, kLib} : std::initializer_list<int>{256, 128, kLib},
        [&](int tile) -> int {
          if (tile == kLib) {
            for (int L = S + 1; L >= 2; L--) {
              const int s_lo = L - S > 1 ? L - S : 1, s_hi = L - 1 < S ? L - 1 : S;
              const int e = iamx::blt_gemm(iamx::BLT_I8_I32, 1, M, N, (int64_t)(s_hi - s_lo + 1) * Kp,
                                           ws + (int64_t)(s_lo - 1) * Kp, S * Kp,
                                           pb.planes + (int64_t)(S - (L - s_lo)) * Kp, S * Kp, lacc, N, 1, 0, 0, 0,
                                           "gemm_ozaki_i8_hipblaslt");
              if (e) return e;
              IAMX_LAUNCH("ozaki_level", k_ozaki_level, lgrid, dim3(256), 0, (const int32_t *)lacc, N, M, N,
                          (const int32_t *)ea, pb.fb, L, L != S + 1 ? 1 : 0, (double *)C, ldc);
            }
            return IAMX_OK;
          }
          if (tile == 256) {
            P.tiles_n = (int)(Np / BN);
            for (int L = S + 1; L >= 2; L--) {
              int n = 0;
              for (int s = 1; s <= S; s++) {
                const int t = L - s;
                if (t < 1 || t > S) continue;
                P.seg[n++] = Seg{(int64_t)(s - 1) * Kp, (int64_t)(S - t) * Kp, L, 0};
              }
              P.nseg = n;
              IAMX_LAUNCH("gemm_ozaki_i8_mfma", k_seg_gemm_big<1>, dim3((unsigned)((Mp / BM) * (Np / BN))),
                          dim3(256), 0, P, L, L != S + 1, (int)(Mp / BM));
            }
          } else {
            P.tiles_n = (int)(Np / TN);
            int n = 0;
            for (int L = S + 1; L >= 2; L--) {
              for (int s = 1; s <= S; s++) {
                const int t = L - s;
                if (t < 1 || t > S) continue;
                P.seg[n++] = Seg{(int64_t)(s - 1) * Kp, (int64_t)(S - t) * Kp, L, 0};
              }
              P.seg[n - 1].level_end = 1;
            }
            P.nseg = n;
            IAMX_LAUNCH("gemm_ozaki_i8_mfma", k_seg_gemm<1>, dim3((unsigned)((Mp / TM) * (Np / TN))), dim3(256), 0,
                        P);
          }
          return IAMX_OK;
        },
        &rc);
  return rc;
}

}  // namespace

extern "C" {

// float32: C = A B through the bf16 hi/lo split (3 products).
int iamx_gemm_split_bf16(int64_t M, int64_t N, int64_t K, const float *A, int64_t lda, const float *B, int64_t ldb,
                         float *C, int64_t ldc) {
  PrepB pb;
  int rc = mixed_prepare_b(pb, 1, 1, N, K, B, ldb, 0);
  return rc ? rc : mixed_rows(pb, 1, M, A, lda, C, ldc);
}

// float64: C = A B through Ozaki INT8 slicing with S slices.
int iamx_gemm_ozaki_i8(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B, int64_t ldb,
                       double *C, int64_t ldc, int S) {
  PrepB pb;
  int rc = mixed_prepare_b(pb, 1, 2, N, K, B, ldb, S);
  return rc ? rc : mixed_rows(pb, 2, M, A, lda, C, ldc);
}

// float64: C = A B through the CRT form with n moduli.
int iamx_gemm_ozaki_crt(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B, int64_t ldb,
                        double *C, int64_t ldc, int n) {
  PrepB pb;
  int rc = mixed_prepare_b(pb, 1, 3, N, K, B, ldb, n);
  return rc ? rc : mixed_rows(pb, 3, M, A, lda, C, ldc);
}

int iamx_mixed_prepare_b(int form, int64_t N, int64_t K, const void *B, int64_t ldb, int S) {
  return mixed_prepare_b(g_pb, 10, form, N, K, B, ldb, S);
}

int iamx_mixed_rows(int form, int64_t M, const void *A, int64_t lda, void *C, int64_t ldc) {
  return mixed_rows(g_pb, form, M, A, lda, C, ldc);
}

}  // extern "C"

// The smallest count in [smin, smax] (Ozaki slices, form 2, or CRT moduli,
// form 3) whose product of up to 256 evenly spaced rows of A by B has a
// componentwise error max |C_S - C| / (|A||B|) <= target against the plain
// FP64 MFMA product of the same rows; smax if none does.
static int pick_count(int form, int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B,
                      int64_t ldb, double target, int smin, int smax, int *slices, double *err) {
  int rc = iamx::ensure_init();
  if (rc) return rc;
  if (!slices || M <= 0 || N <= 0 || K <= 0) return iamx::set_error(IAMX_EINVAL, "ozaki tuner: bad arguments");
  if (smin > smax) smin = smax;
  const int R = (int)(M < 256 ? M : 256);
  const size_t rk = (size_t)R * (size_t)K, rn = (size_t)R * (size_t)N;
  uint8_t *w = (uint8_t *)iamx_workspace(8, (2 * rk + 3 * rn) * sizeof(double) + 256);
  double *Ba = (double *)iamx_workspace(9, (size_t)K * (size_t)N * sizeof(double) + 256);
  if (!w || !Ba) return iamx::set_error(IAMX_ENOMEM, "ozaki tuner workspace");
  double *As = (double *)w, *Aa = As + rk, *Cref = Aa + rk, *den = Cref + rn, *Cs = den + rn;
  unsigned long long *d_err = (unsigned long long *)(Cs + rn);
  const unsigned gx = (unsigned)((K + 255) / 256 < 64 ? (K + 255) / 256 : 64);
  IAMX_LAUNCH("ozaki_tune_sample", k_sample_rows, dim3(gx, (unsigned)R), dim3(256), 0, A, lda, M, K, R, As, Aa);
  IAMX_LAUNCH("ozaki_tune_abs", k_abs_rows, dim3(2048), dim3(256), 0, B, ldb, K, N, Ba);
  rc = iamx_gemm(0, 0, 0, R, N, K, As, K, B, ldb, Cref, N);
  if (!rc) rc = iamx_gemm(0, 0, 0, R, N, K, Aa, K, Ba, N, den, N);
  if (rc) return rc;
  int chosen = smax;
  double chosen_err = -1.0;
  for (int S = smin; S <= smax; S++) {
    rc = form == 3 ? iamx_gemm_ozaki_crt(R, N, K, As, K, B, ldb, Cs, N, S)
                   : iamx_gemm_ozaki_i8(R, N, K, As, K, B, ldb, Cs, N, S);
    if (rc) return rc;
    IAMX_LAUNCH("ozaki_tune_err", k_zero_u64, dim3(1), dim3(1), 0, d_err);
    IAMX_LAUNCH("ozaki_tune_err", k_err_max, dim3(256), dim3(256), 0, Cs, Cref, den, (int64_t)rn, d_err);
    unsigned long long bits = 0;
    rc = iamx_d2h(&bits, d_err, sizeof(bits));
    if (rc) return rc;
    double e;
    memcpy(&e, &bits, sizeof(e));
    chosen = S;
    chosen_err = e;
    if (e <= target) break;
  }
  *slices = chosen;
  if (err) *err = chosen_err;
  return IAMX_OK;
}

extern "C" {

IAMX_API int iamx_ozaki_pick_slices(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda, const double *B,
                                    int64_t ldb, double target, int smin, int *slices, double *err) {
  if (smin < 2) smin = 2;
  return pick_count(2, M, N, K, A, lda, B, ldb, target, smin, 8, slices, err);
}

IAMX_API int iamx_ozaki_crt_pick_moduli(int64_t M, int64_t N, int64_t K, const double *A, int64_t lda,
                                        const double *B, int64_t ldb, double target, int nmin, int *moduli,
                                        double *err) {
  if (nmin < 8) nmin = 8;
  return pick_count(3, M, N, K, A, lda, B, ldb, target, nmin, kMaxMods, moduli, err);
}

}  // extern "C"
