// iamx_gemv.hip -- GEMV fused with the Blosc2 block un-filter.
//
// After the LZ4 streams of A are decoded into scratch (k_lz4_streams), one
// workgroup per compressed A block turns the block's streams straight into
// row sums of A_blk x_seg without writing A anywhere.
//
// Bit-shuffled single-stream blocks (the common case) take the quad path:
// a thread takes 32 consecutive elements ("quad" u = elements 32u..32u+31):
// one dword from each of the 8*typesize bit planes (byte 4u..4u+3 of every
// plane), byte-permutes them into 8-byte columns, runs one 8x8 bit-matrix
// transpose per (group, element byte) and byte-permutes the results into
// elements.  A quad spans at most two rows of the block (rows have >= 32
// elements), so it leaves two partial sums in LDS; then one thread per row
// adds its quads' partial sums in quad order.  Any other block layout takes
// the group path: 8 elements per thread, bytes gathered one by one.
// Rows are processed in windows so the partials fit in LDS.  Row sums land in
// partial[block column][row]; a second launch adds the block columns in
// order, so y is bit-identical run to run.
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
constexpr int kGroups = 2048;  // per-window groups (or quads) whose partial sums live in LDS

// x is staged in LDS with one pad word per 32 columns: the lanes of a wave
// read columns 32 apart, which then fall in different banks.
__device__ __forceinline__ int xpad(int c) { return c + (c >> 5); }

// 4x4 byte transpose: r[g] = bytes g of (a, b, c, d)
__device__ __forceinline__ void tr4(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t r[4]) {
  const uint32_t t0 = __builtin_amdgcn_perm(b, a, 0x05010400u);  // a0 b0 a1 b1
  const uint32_t t1 = __builtin_amdgcn_perm(b, a, 0x07030602u);  // a2 b2 a3 b3
  const uint32_t t2 = __builtin_amdgcn_perm(d, c, 0x05010400u);  // c0 d0 c1 d1
  const uint32_t t3 = __builtin_amdgcn_perm(d, c, 0x07030602u);  // c2 d2 c3 d3
  r[0] = __builtin_amdgcn_perm(t2, t0, 0x05040100u);
  r[1] = __builtin_amdgcn_perm(t2, t0, 0x07060302u);
  r[2] = __builtin_amdgcn_perm(t3, t1, 0x05040100u);
  r[3] = __builtin_amdgcn_perm(t3, t1, 0x07060302u);
}

// 8x8 bit transpose of (hi:lo) as two 32-bit halves (the delta swaps of
// transpose8 below, the first two of which stay inside each half)
__device__ __forceinline__ void transpose8_32(uint32_t &lo, uint32_t &hi) {
  uint32_t t;
  t = (lo ^ (lo >> 7)) & 0x00AA00AAu;
  lo ^= t ^ (t << 7);
  t = (hi ^ (hi >> 7)) & 0x00AA00AAu;
  hi ^= t ^ (t << 7);
  t = (lo ^ (lo >> 14)) & 0x0000CCCCu;
  lo ^= t ^ (t << 14);
  t = (hi ^ (hi >> 14)) & 0x0000CCCCu;
  hi ^= t ^ (t << 14);
  t = (lo ^ __builtin_amdgcn_alignbit(hi, lo, 28)) & 0xF0F0F0F0u;
  lo ^= t ^ (t << 28);
  hi ^= t >> 4;
}

template <typename T>
struct Word;
template <>
struct Word<float> {
  using U = uint32_t;
  static __device__ __forceinline__ float of(uint32_t lo, uint32_t) { return __uint_as_float(lo); }
};
template <>
struct Word<double> {
  using U = uint64_t;
  static __device__ __forceinline__ double of(uint32_t lo, uint32_t hi) {
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
  }
};

struct Src {
  const IAMX_GLOBAL uint8_t *ptr[IAMX_MAX_STREAMS];
  uint8_t fill[IAMX_MAX_STREAMS];
};

// Row sums of one bit-shuffled single-stream block, quads of 32 elements.
// Plane p starts at base + p*rowlen; elements [n8, n) follow the planes
// unshuffled and form one more (short) quad.
template <typename T>
__device__ void quad_path(const IAMX_GLOBAL uint8_t *base, int32_t n, int32_t n8, int32_t rowlen, int32_t cols,
                          int32_t vrows, int32_t vcols, const T *xs, T *p0, T *p1, T *prow, int64_t row0,
                          int64_t M) {
  constexpr int TS = sizeof(T);
  constexpr int NP = 8 * TS;
  const int tid = threadIdx.x;
  const int32_t nq = n8 / 32;  // quads drawn whole from the planes
  const int32_t nqe = (n + 31) / 32;
  const bool aligned = ((rowlen & 3) == 0) && ((((uint64_t)base) & 3) == 0);
  const int32_t rows_per_win = (int32_t)(((int64_t)(kGroups - 2) * 32) / cols);
  for (int32_t ra = 0; ra < vrows; ra += rows_per_win) {
    const int32_t rb = min(vrows, ra + rows_per_win);
    const int32_t u0 = (ra * cols) / 32, u1 = min(nqe - 1, (rb * cols - 1) / 32);  // quads [u0, u1]
    for (int32_t u = u0 + tid; u <= u1; u += 256) {
      const int32_t e0 = 32 * u;
      const int32_t r0 = e0 / cols;
      const int32_t c0 = e0 - r0 * cols;
      const int32_t k0 = cols - c0;  // elements of the quad in row r0
      T s0 = 0, s1 = 0;
      if (u < nq) {
        uint32_t d[NP];
        if (aligned) {
#pragma unroll
          for (int p = 0; p < NP; p++) d[p] = *(const IAMX_GLOBAL uint32_t *)(base + (int64_t)p * rowlen + 4 * u);
        } else {
#pragma unroll
          for (int p = 0; p < NP; p++) {
            const IAMX_GLOBAL uint8_t *q = base + (int64_t)p * rowlen + 4 * u;
            d[p] = (uint32_t)q[0] | ((uint32_t)q[1] << 8) | ((uint32_t)q[2] << 16) | ((uint32_t)q[3] << 24);
          }
        }
#pragma unroll
        for (int g = 0; g < 4; g++) {
          // Y[j] = bit transpose of byte g of planes 8j..8j+7: byte i is byte j of element 8g+i
          uint32_t ylo[TS], yhi[TS];
          const uint32_t sel = (uint32_t)g | ((uint32_t)(g + 4) << 8);
#pragma unroll
          for (int j = 0; j < TS; j++) {
            const uint32_t a0 = __builtin_amdgcn_perm(d[8 * j + 1], d[8 * j + 0], sel);
            const uint32_t a1 = __builtin_amdgcn_perm(d[8 * j + 3], d[8 * j + 2], sel);
            const uint32_t a2 = __builtin_amdgcn_perm(d[8 * j + 5], d[8 * j + 4], sel);
            const uint32_t a3 = __builtin_amdgcn_perm(d[8 * j + 7], d[8 * j + 6], sel);
            uint32_t lo = __builtin_amdgcn_perm(a1, a0, 0x05040100u);
            uint32_t hi = __builtin_amdgcn_perm(a3, a2, 0x05040100u);
            transpose8_32(lo, hi);
            ylo[j] = lo;
            yhi[j] = hi;
          }
          // element 8g+i, i < 4: bytes i of ylo[0..TS); i >= 4: bytes i-4 of yhi
          uint32_t elo[8], ehi[8];
          {
            uint32_t r[4];
            tr4(ylo[0], ylo[1], ylo[2], ylo[3], r);
#pragma unroll
            for (int i = 0; i < 4; i++) elo[i] = r[i];
            tr4(yhi[0], yhi[1], yhi[2], yhi[3], r);
#pragma unroll
            for (int i = 0; i < 4; i++) elo[4 + i] = r[i];
            if (TS == 8) {
              tr4(ylo[4 % TS], ylo[5 % TS], ylo[6 % TS], ylo[7 % TS], r);
#pragma unroll
              for (int i = 0; i < 4; i++) ehi[i] = r[i];
              tr4(yhi[4 % TS], yhi[5 % TS], yhi[6 % TS], yhi[7 % TS], r);
#pragma unroll
              for (int i = 0; i < 4; i++) ehi[4 + i] = r[i];
            } else {
#pragma unroll
              for (int i = 0; i < 8; i++) ehi[i] = 0;
            }
          }
#pragma unroll
          for (int i = 0; i < 8; i++) {
            const int32_t k = 8 * g + i;
            const bool first = k < k0;
            const int32_t c = first ? c0 + k : c0 + k - cols;
            const T prod = c < vcols ? Word<T>::of(elo[i], ehi[i]) * xs[xpad(c)] : (T)0;
            if (first) s0 += prod;
            else s1 += prod;
          }
        }
      } else {
        // the last, short quad: whole 8-element groups still come from the
        // planes (byte 4u+g of each, gathered one by one), elements [n8, n)
        // follow the planes unshuffled
        for (int32_t e = e0; e < n; e++) {
          uint32_t lo = 0, hi = 0;
          if (e < n8) {
            const int32_t gb = e >> 3, bit = e & 7;
#pragma unroll 8
            for (int p = 0; p < NP; p++) {
              const uint32_t v = (base[(int64_t)p * rowlen + gb] >> bit) & 1u;
              if (p < 32) lo |= v << p;
              else hi |= v << (p - 32);
            }
          } else {
            const IAMX_GLOBAL uint8_t *q = base + (int64_t)n8 * TS + (int64_t)(e - n8) * TS;
#pragma unroll
            for (int bb = 0; bb < 4; bb++) lo |= (uint32_t)q[bb] << (8 * bb);
            if (TS == 8) {
#pragma unroll
              for (int bb = 0; bb < 4; bb++) hi |= (uint32_t)q[4 + (bb & 3)] << (8 * bb);
            }
          }
          const int32_t k = e - e0;
          const bool first = k < k0;
          const int32_t c = first ? c0 + k : c0 + k - cols;
          const T prod = c < vcols ? Word<T>::of(lo, hi) * xs[xpad(c)] : (T)0;
          if (first) s0 += prod;
          else s1 += prod;
        }
      }
      p0[u - u0] = s0;
      p1[u - u0] = s1;
    }
    __syncthreads();
    for (int32_t r = ra + tid; r < rb; r += 256) {
      const int32_t ua = (r * cols) / 32, ub = min(nqe - 1, (r * cols + cols - 1) / 32);
      T acc = 0;
      for (int32_t u = ua; u <= ub; u++) acc += ((32 * u) / cols == r) ? p0[u - u0] : p1[u - u0];
      if (row0 + r < M) prow[row0 + r] = acc;
    }
    __syncthreads();
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_gemv_fused(const iamx_block_job_t *__restrict__ jobs,
                                                    const int32_t *__restrict__ blk_rc, int njobs,
                                                    const T *__restrict__ x, const uint8_t *__restrict__ scratch,
                                                    T *__restrict__ partial, int64_t M, int64_t K, int nbc,
                                                    int64_t cap, int *status, const uint8_t *__restrict__ xzero) {
  __shared__ iamx_block_job_t J;
  __shared__ Src S;
  __shared__ T xs[kMaxCols + kMaxCols / 32];
  __shared__ T p0[kGroups], p1[kGroups];
  __shared__ int job_ok;
  const int b = blockIdx.x;
  if (b >= njobs) return;
  const int tid = threadIdx.x;
  const int64_t row0 = blk_rc[4 * b], col0 = blk_rc[4 * b + 1];
  const int bcol = blk_rc[4 * b + 2];
  // opt_gemv (xzero given) skips a zero block of A or of x: its partial sums
  // stay at the zero they start from
  if (xzero && bcol >= 0 && bcol < nbc && ((blk_rc[4 * b + 3] & 1) | xzero[bcol])) return;
  if (tid == 0) {
    J = jobs[b];
    // every range the job names must lie inside the buffers of this call:
    // scratch streams below `cap`, its rows inside y, its columns inside x,
    // its partial-sum row inside the nbc block columns
    bool ok = J.nstreams >= 1 && J.nstreams <= IAMX_MAX_STREAMS && J.cols >= 1 && J.cols <= kMaxCols &&
              J.vcols >= 0 && J.vcols <= J.cols && J.vrows >= 0 && J.neblock >= 0 && bcol >= 0 && bcol < nbc &&
              row0 >= 0 && row0 <= M - J.vrows && col0 >= 0 && col0 <= K - J.vcols;
    for (int s = 0; ok && s < J.nstreams; s++)
      if (J.kind[s] == IAMX_SRC_SCRATCH) ok = (int64_t)J.src[s] >= 0 && (int64_t)J.src[s] <= cap - J.neblock;
    job_ok = ok;
    if (!ok) atomicOr(status, IAMX_STATUS_RANGE);
  }
  __syncthreads();
  if (!job_ok) return;
  if (tid < IAMX_MAX_STREAMS) {
    const int s = tid;
    S.ptr[s] = s >= J.nstreams                ? nullptr
               : J.kind[s] == IAMX_SRC_ADDR    ? iamx::gptr<const uint8_t>(J.src[s])
               : J.kind[s] == IAMX_SRC_SCRATCH ? iamx::gptr<const uint8_t>((uint64_t)(scratch + J.src[s]))
                                               : nullptr;
    S.fill[s] = s < J.nstreams ? J.fill[s] : 0;
  }
  for (int c = tid; c < J.cols; c += 256) xs[xpad(c)] = c < J.vcols ? x[col0 + c] : (T)0;
  __syncthreads();

  const int ts = J.typesize;
  const int32_t n = J.bsize / ts;
  const int32_t cols = J.cols, vrows = J.vrows, vcols = J.vcols;
  const int32_t neblock = J.neblock;
  const int filter = J.filter;
  const bool bits = filter == IAMX_FILTER_BITSHUFFLE || filter == IAMX_FILTER_BITSHUFFLE_V2;
  const int32_t n8 = filter == IAMX_FILTER_BITSHUFFLE ? n - n % 8 : n;
  const int32_t rowlen = n8 / 8;
  if (bits && J.nstreams == 1 && S.ptr[0] != nullptr && cols >= 32 && n8 >= 32 && ts == (int)sizeof(T)) {
    quad_path<T>(S.ptr[0], n, n8, rowlen, cols, vrows, vcols, xs, p0, p1, partial + (int64_t)bcol * M, row0, M);
    return;
  }
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
        acc += bits_to<T>(v) * xs[xpad(c)];
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
          const T prod = bits_to<T>(v[i]) * xs[xpad(c)];
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
  // the sum runs over the block columns in order (bit-identical run to run);
  // eight loads go out before their adds
  T s = 0;
  int c = 0;
  for (; c + 8 <= nbc; c += 8) {
    T v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = partial[(int64_t)(c + k) * M + r];
#pragma unroll
    for (int k = 0; k < 8; k++) s += v[k];
  }
  for (; c < nbc; c++) s += partial[(int64_t)c * M + r];
  y[r] = s;
}

__global__ __launch_bounds__(256) void k_zero_words(uint32_t *__restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = 0u;
}

// Decode + fused GEMV + column reduction on the current stream.  `cap` bounds
// the decoded-stream bytes of the scratch and K the length of x; the kernels
// check every job against them (IAMX_STATUS_RANGE).
int gemv_fused_launch(int dtype, const iamx_lz4_job_t *d_lz4, int n_lz4, const iamx_block_job_t *d_blocks,
                      const int32_t *d_blk_rc, int n_blocks, int nbc, const void *x, void *y, int64_t M, int64_t K,
                      int64_t cap, void *d_scratch, void *d_partial, const uint8_t *d_xzero = nullptr) {
  if (n_lz4 > 0) {
    const int rc = iamx::launch_lz4("lz4_streams_gemv", d_lz4, n_lz4, d_scratch, cap);
    if (rc) return rc;
  }
  const dim3 rgrid((unsigned)((M + 255) / 256));
  int *st = iamx::status_word();
  if (dtype == 0) {
    IAMX_LAUNCH("gemv_fused_f64", k_gemv_fused<double>, dim3(n_blocks), dim3(256), 0, d_blocks, d_blk_rc, n_blocks,
                (const double *)x, (const uint8_t *)d_scratch, (double *)d_partial, M, K, nbc, cap, st, d_xzero);
    IAMX_LAUNCH("gemv_reduce_f64", k_reduce_cols<double>, rgrid, dim3(256), 0, (const double *)d_partial, nbc, M,
                (double *)y);
  } else if (dtype == 1) {
    IAMX_LAUNCH("gemv_fused_f32", k_gemv_fused<float>, dim3(n_blocks), dim3(256), 0, d_blocks, d_blk_rc, n_blocks,
                (const float *)x, (const uint8_t *)d_scratch, (float *)d_partial, M, K, nbc, cap, st, d_xzero);
    IAMX_LAUNCH("gemv_reduce_f32", k_reduce_cols<float>, rgrid, dim3(256), 0, (const float *)d_partial, nbc, M,
                (float *)y);
  } else {
    return iamx::set_error(IAMX_EINVAL, "iamx_gemv_fused: dtype %d", dtype);
  }
  return IAMX_OK;
}

// ---- the no-allocation GEMV's scratch ------------------------------------------------------
// Decoded LZ4 streams, then the partial sums, of iamx_gemv_fused_noalloc: a
// fixed device array (every GPU the library runs on holds its own copy until
// the process ends).  Host code never passes `noalloc_scratch` itself to a
// kernel (in host code the name is a stand-in, not a device address): the
// address of this device's copy is resolved once per device by
// hipGetSymbolAddress, with that device current, and kept.
constexpr size_t kNoAllocBytes = (size_t)64 << 20;
__device__ __align__(256) uint8_t noalloc_scratch[kNoAllocBytes];
constexpr int kMaxDevices = 64;
uint8_t *g_noalloc[kMaxDevices];

int noalloc_ptr(uint8_t **out) {
  int dev = -1;
  IAMX_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return iamx::set_error(IAMX_EINVAL, "no-allocation scratch: device %d", dev);
  if (!g_noalloc[dev]) {
    const int rc = iamx::noalloc_setup();
    if (rc) return rc;
  }
  *out = g_noalloc[dev];
  return IAMX_OK;
}

}  // namespace

int iamx::noalloc_setup() {
  int dev = -1;
  IAMX_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error(IAMX_EINVAL, "no-allocation scratch: device %d", dev);
  if (g_noalloc[dev]) return IAMX_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch));
  if (e != hipSuccess) return check_hip(e, "hipGetSymbolAddress(noalloc_scratch)");
  g_noalloc[dev] = (uint8_t *)p;
  return IAMX_OK;
}

extern "C" {

IAMX_API int iamx_gemv_fused(int dtype, const iamx_lz4_job_t *d_lz4, int n_lz4, const iamx_block_job_t *d_blocks,
                             const int32_t *d_blk_rc, int n_blocks, int nbc, int bs0, int bs1, const void *x,
                             void *y, int64_t M, void *d_scratch, void *d_partial) {
  int rc = iamx::ensure_init();
  if (rc) return rc;
  if (bs1 > kMaxCols) return iamx::set_error(IAMX_EINVAL, "iamx_gemv_fused: block width %d > %d", bs1, kMaxCols);
  (void)bs0;
  if (n_blocks <= 0 || nbc <= 0) return iamx::set_error(IAMX_EINVAL, "iamx_gemv_fused: no blocks");
  return gemv_fused_launch(dtype, d_lz4, n_lz4, d_blocks, d_blk_rc, n_blocks, nbc, x, y, M, INT64_MAX, INT64_MAX,
                           d_scratch, d_partial);
}

IAMX_API int iamx_opt_gemv_fused(int dtype, const iamx_lz4_job_t *d_lz4, int n_lz4, const iamx_block_job_t *d_blocks,
                                 const int32_t *d_blk_rc, int n_blocks, int nbc, int bs0, int bs1, const void *x,
                                 void *y, int64_t M, void *d_scratch, void *d_partial, const uint8_t *d_xzero) {
  int rc = iamx::ensure_init();
  if (rc) return rc;
  (void)bs0;
  if (bs1 > kMaxCols) return iamx::set_error(IAMX_EINVAL, "iamx_opt_gemv_fused: block width %d > %d", bs1, kMaxCols);
  if (n_blocks <= 0 || nbc <= 0 || !d_xzero) return iamx::set_error(IAMX_EINVAL, "iamx_opt_gemv_fused: no blocks");
  return gemv_fused_launch(dtype, d_lz4, n_lz4, d_blocks, d_blk_rc, n_blocks, nbc, x, y, M, INT64_MAX, INT64_MAX,
                           d_scratch, d_partial, d_xzero);
}

IAMX_API int64_t iamx_gemv_fused_noalloc_limit(void) { return (int64_t)kNoAllocBytes; }

IAMX_API int iamx_gemv_fused_noalloc(int dtype, const iamx_lz4_job_t *d_lz4, int n_lz4,
                                     const iamx_block_job_t *d_blocks, const int32_t *d_blk_rc, int n_blocks, int nbc,
                                     int bs0, int bs1, const void *x, int64_t K, void *y, int64_t M,
                                     int64_t scratch_bytes) {
  (void)bs0;
  if (!iamx::initialized()) return iamx::set_error(IAMX_EINVAL, "iamx_gemv_fused_noalloc: call iamx_init first");
  if (dtype != 0 && dtype != 1) return iamx::set_error(IAMX_EINVAL, "iamx_gemv_fused_noalloc: dtype %d", dtype);
  if (bs1 > kMaxCols || bs1 < 1)
    return iamx::set_error(IAMX_EINVAL, "iamx_gemv_fused_noalloc: block width %d outside 1..%d", bs1, kMaxCols);
  if (n_blocks <= 0 || nbc <= 0 || n_lz4 < 0 || M <= 0 || K <= 0 || scratch_bytes < 0)
    return iamx::set_error(IAMX_EINVAL, "iamx_gemv_fused_noalloc: bad sizes");
  const int64_t isz = dtype == 0 ? 8 : 4;
  const int64_t poff = (scratch_bytes + 255) / 256 * 256;
  if (poff > (int64_t)kNoAllocBytes || (int64_t)nbc > ((int64_t)kNoAllocBytes - poff) / isz / M)
    return iamx::set_error(IAMX_EINVAL,
                           "iamx_gemv_fused_noalloc: needs %lld bytes of scratch (decoded streams) + %lld (partial "
                           "sums), above its fixed %lld",
                           (long long)scratch_bytes, (long long)((int64_t)nbc * M * isz), (long long)kNoAllocBytes);
  uint8_t *base = nullptr;
  int rc = noalloc_ptr(&base);
  if (rc) return rc;
  // the partial sums start at zero (blocks that hold no job leave theirs);
  // a fill kernel, not a memset node
  const int64_t pn = (int64_t)nbc * M * isz / 4;
  IAMX_LAUNCH("gemv_noalloc_clear", k_zero_words, dim3((unsigned)((pn + 1023) / 1024 < 4096 ? (pn + 1023) / 1024 : 4096)),
              dim3(256), 0, (uint32_t *)(base + poff), pn);
  return gemv_fused_launch(dtype, d_lz4, n_lz4, d_blocks, d_blk_rc, n_blocks, nbc, x, y, M, K, scratch_bytes, base,
                           base + poff);
}

}  // extern "C"
