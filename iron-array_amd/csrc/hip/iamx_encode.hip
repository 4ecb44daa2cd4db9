// iamx_encode.hip -- the C write stage on the GPU: LZ4 encode of the filtered
// block streams and Blosc2 chunk assembly, i.e. what blosc2_compress_ctx does
// for every C chunk of the gemm branch (src/linalg/matmul/gemm.c:216-243) and
// of iarray_opt_gemm (src/linalg/iarray_gemm.c:399-425).
//
// k_lz4_encode_sub one wavefront per 2 KiB sub-block of a stream, staged in
//                  LDS: a greedy LZ77 parse whose matches start and end inside
//                  the sub-block.  Candidates are found 64 positions at a
//                  time: each lane hashes the 4 bytes at its position and
//                  reads the hash table (LDS, 2^9 entries holding a position
//                  + 1); when the 4 bytes there are equal to its own it also
//                  compares the next 16 bytes at both places.
//                  Then every lane inserts its position with an atomic max,
//                  so the latest position of a hash wins whatever the lane
//                  order, and a candidate always comes from an earlier step.
//                  The parse walks the step's candidates in position order
//                  (the first at or after the parse position), extends a
//                  match that reached 20 bytes 256 bytes a step, and writes
//                  the sequences; a literal run that starts in the current
//                  step is written from the lanes' registers.  The leading
//                  literals (before the first match) and the trailing ones
//                  (after the last) are left to the merge.
// k_lz4_merge      one wavefront per stream: joins each sub-block's leading
//                  literals with the previous one's trailing literals into one
//                  run (token and length written here), copies the
//                  sub-blocks' sequences after it, and ends with the stream's
//                  last literals.  The format's end rules hold: the last match
//                  starts at least 12 bytes before the end of the stream and
//                  ends at least 5 before it.  A stream whose output would
//                  reach its own size is stored raw, a stream of zeros as a
//                  zero stream (csize 0).
// k_chunk_offsets  one workgroup per chunk: stream sizes -> stream offsets (a
//                  scan), the chunk header, bstarts and the chunk's size.
// k_chunk_copy     one workgroup per stream: its csize word and its bytes
//                  (compressed, or the raw filtered bytes) into the chunk.
// Chunk c is assembled at c * cap of the output; stream E (chunk-major,
// block-major, stream-minor) has an output slot at E * (neblock rounded up to
// 16) and two int32 of metadata (csize, then offset in its chunk).
#include <string.h>

#include <vector>

#include "iamx_internal.h"

namespace {

using iamx::gptr;


struct EncGeom {
  int64_t chunk_nbytes;  // bytes of one (extended) chunk in the packed layout
  int64_t cap;           // output stride of a chunk (its largest possible size, multiple of 16)
  int64_t slot;          // encode slot stride of a stream
  int32_t blocksize;
  int32_t nblocks;  // blocks per chunk
  int32_t nstreams;
  int32_t neblock;  // bytes per stream
  int32_t typesize;
  int32_t filter;
  int32_t split;
};

__device__ __forceinline__ uint32_t rdl(uint32_t v, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, l); }

// Group copy of len bytes: byte head up to the destination's dword alignment,
// dword stores realigned from the source, byte tail.
__device__ void wg_copy(IAMX_GLOBAL uint8_t *d, const IAMX_GLOBAL uint8_t *s, int32_t len, int tid, int nt) {
  const int32_t head = min((int32_t)((4 - ((uint64_t)d & 3)) & 3), len);
  if (tid < head) d[tid] = s[tid];
  const int32_t body = len - head;
  const int32_t nw = body >> 2;
  if (nw > 0) {
    const uint64_t sa = (uint64_t)(s + head);
    const uint32_t sh = (uint32_t)(sa & 3);
    const IAMX_GLOBAL uint32_t *sw = gptr<const uint32_t>(sa & ~3ull);
    const int32_t lastw = (int32_t)((sh + body - 1) >> 2);
    IAMX_GLOBAL uint32_t *dw = gptr<uint32_t>((uint64_t)(d + head));
    for (int32_t g = tid; g < nw; g += nt)
      dw[g] = __builtin_amdgcn_alignbyte(sw[min(g + 1, lastw)], sw[g], sh);
  }
  for (int32_t k = head + 4 * nw + tid; k < len; k += nt) d[k] = s[k];
}

// continuation bytes of an LZ4 length field for a count v (>= 0)
__device__ __forceinline__ int32_t ext_bytes(int32_t v) { return v / 255 + 1; }

// Sub-block parse.  Sub-block k of a stream covers bytes [k*kSub, (k+1)*kSub);
// its matches start and end inside it (and keep the stream's end rules) and
// refer to it only.  What it writes to its slot:
//   the rest of its first sequence after the literals (offset and length
//   continuation bytes), then its other sequences whole;
// and to its info record: first match position, offset and length, bytes
// written, where its trailing literals start, whether any byte is nonzero,
// and overflow.  The merge joins each sub-block's leading literals with the
// previous one's trailing literals into one run.
// 2 KiB sub-blocks: stage and hash table take 4.1 KB of LDS, so 8 wavefronts
// per SIMD are resident (4 KiB ones with 2^10 entries took 8.3 KB, 5 per
// SIMD, and each launch twice as long; their C was 1% smaller for float32
// products, 0.1% for float64)
constexpr int kSub = 2048;
constexpr int kSubHlog = 9;
constexpr int kSubSlot = kSub + 256;  // a sub-block's worst output, rounded
constexpr int kInfo = 8;              // int32 per info record
constexpr int kGiveUp = 1024;         // bytes scanned without a match before a sub-block stays literal
enum { I_HAS = 0, I_POS, I_OFF, I_ML, I_LEN, I_TRAIL, I_NZ, I_FAIL };

// Sub-block sbid of the call (stream e0 + sbid / K, sub-block sbid % K).
__device__ __forceinline__ void encode_sub(const EncGeom &G, const uint8_t *__restrict__ packed, int64_t e0,
                                           int64_t sbid, int32_t K, uint8_t *__restrict__ subslots,
                                           int32_t *__restrict__ info) {
  __shared__ uint32_t ht[1 << kSubHlog];  // latest position + 1 of each hash
  __shared__ __align__(16) uint32_t stage[kSub / 4 + 16];
  const int lane = threadIdx.x;
  const int64_t E = e0 + sbid / K;
  const int32_t k = (int32_t)(sbid % K);
  const int64_t per_chunk = (int64_t)G.nblocks * G.nstreams;
  const int64_t c = E / per_chunk, r = E % per_chunk;
  const int64_t b = r / G.nstreams, s = r % G.nstreams;
  const int32_t n = G.neblock;
  const int32_t S0 = k * kSub, m = min(n - S0, kSub);  // bytes of this sub-block
  const uint64_t src = (uint64_t)(packed + c * G.chunk_nbytes + b * G.blocksize + s * (int64_t)n + S0);
  IAMX_GLOBAL uint8_t *dst = gptr<uint8_t>((uint64_t)(subslots + sbid * kSubSlot));
  IAMX_GLOBAL int32_t *inf = gptr<int32_t>((uint64_t)(info + sbid * kInfo));
  const int32_t sh0 = (int32_t)(src & 3);
  const IAMX_GLOBAL uint32_t *__restrict__ sw = gptr<const uint32_t>(src & ~3ull);
  const int32_t lastw = (sh0 + m - 1) >> 2;
  // stage word g = sub-block bytes 4g..4g+3, zero past its end, plus 16 zero words
  const int32_t nw = (m + 3) >> 2;
  uint32_t nz = 0;
  for (int32_t g = lane; g < kSub / 4 + 16; g += 64) {
    uint32_t v = 0;
    if (g < nw) {
      v = __builtin_amdgcn_alignbyte(sw[min(g + 1, lastw)], sw[min(g, lastw)], (uint32_t)sh0);
      const int32_t valid = m - 4 * g;
      if (valid < 4) v &= (1u << (8 * valid)) - 1u;
    }
    nz |= v;
    stage[g] = v;
  }
  for (int i = lane; i < (1 << kSubHlog); i += 64) ht[i] = 0;
  __builtin_amdgcn_wave_barrier();
  auto word_at = [&](int32_t p) -> uint32_t {  // 4 bytes at local position p (p <= m + 40)
    const int32_t q = p >> 2;
    return __builtin_amdgcn_alignbyte(stage[q + 1], stage[q], (uint32_t)(p & 3));
  };
  auto byte_at = [&](int32_t p) -> uint32_t { return (stage[p >> 2] >> ((p & 3) << 3)) & 255u; };

  // local limits: a match starts before mflimit and ends at or before
  // matchlimit, so its 4 minimum bytes fit inside the sub-block
  const int32_t mflimit = min(m - 3, n - 12 - S0);
  const int32_t matchlimit = min(m, n - 5 - S0);
  const int32_t cap = kSubSlot - 32;
  int32_t anchor = 0, ip = 0, op = 0;
  bool fail = false, first = true;
  int32_t f_pos = 0, f_off = 0, f_ml = 0;

  auto put_ext = [&](int32_t v) {
    const int32_t nf = v / 255;
    for (int32_t q = lane; q <= nf; q += 64) dst[op + q] = (uint8_t)(q < nf ? 255 : v - 255 * nf);
    op += nf + 1;
  };
  auto put_stage = [&](int32_t from, int32_t len) {  // local bytes [from, from+len) to the output
    const int32_t head = min((4 - (op & 3)) & 3, len);
    if (lane < head) dst[op + lane] = (uint8_t)byte_at(from + lane);
    const int32_t nwd = (len - head) >> 2;
    IAMX_GLOBAL uint32_t *dw = gptr<uint32_t>((uint64_t)(dst + op + head));
    for (int32_t g = lane; g < nwd; g += 64) dw[g] = word_at(from + head + 4 * g);
    const int32_t t0 = head + 4 * nwd;
    if (lane < len - t0) dst[op + t0 + lane] = (uint8_t)byte_at(from + t0 + lane);
    op += len;
  };

  for (int32_t base = 0; base < mflimit; base += 64) {
    // A sub-block whose first kGiveUp bytes hold no match is left as
    // literals (random mantissa planes: LZ4 would store the stream raw), the
    // wavefront analogue of LZ4's search acceleration after misses.
    if (first && base >= kGiveUp) break;
    const int32_t pos = base + lane;
    const bool act = pos < mflimit;
    const int32_t lp = min(pos, mflimit - 1);
    uint32_t x[5];
#pragma unroll
    for (int t = 0; t < 5; t++) x[t] = word_at(lp + 4 * t);
    const uint32_t seq = x[0];
    const uint32_t h = (seq * 2654435761u) >> (32 - kSubHlog);
    const int32_t cand = act ? (int32_t)ht[h] - 1 : -1;
    __builtin_amdgcn_wave_barrier();
    const bool ok = cand >= 0 && word_at(cand) == seq;
    int32_t mlen = 0;
    if (ok) {
      int32_t ext = 16;
#pragma unroll
      for (int t = 1; t < 5; t++) {
        const uint32_t d = x[t] ^ word_at(cand + 4 * t);
        if (ext == 16 && d) ext = 4 * (t - 1) + (__builtin_ctz(d) >> 3);
      }
      mlen = 4 + ext;
    }
    if (act) atomicMax(&ht[h], (uint32_t)(pos + 1));
    __builtin_amdgcn_wave_barrier();
    uint64_t mask = __ballot(ok);
    if (ip > base) mask = ip - base >= 64 ? 0ull : mask & (~0ull << (ip - base));
    while (mask) {
      const int l = __builtin_ctzll(mask);
      const int32_t i = base + l;
      const int32_t off = i - (int32_t)rdl((uint32_t)cand, l);
      int32_t end = i + (int32_t)rdl((uint32_t)mlen, l);
      if (end - i == 20) {
        int32_t p = end;
        for (;;) {
          if (p >= matchlimit) break;
          const int32_t q = p + 4 * lane;
          const int32_t qc = min(q, matchlimit);
          const uint32_t d = q < matchlimit ? (word_at(qc) ^ word_at(qc - off)) : 0u;
          const uint64_t stop = __ballot(q >= matchlimit || d != 0);
          if (stop) {
            const int L = __builtin_ctzll(stop);
            const uint32_t dl = rdl(d, L);
            p += 4 * L + (dl ? (int32_t)(__builtin_ctz(dl) >> 3) : 0);
            break;
          }
          p += 256;
        }
        end = p;
      }
      if (end > matchlimit) end = matchlimit;
      const int32_t ml = end - i;
      if (first) {
        // leading literals are left to the merge: offset and length continuation only
        first = false;
        f_pos = i;
        f_off = off;
        f_ml = ml;
        if (lane < 2) dst[op + lane] = (uint8_t)(off >> (8 * lane));
        op += 2;
        if (ml - 4 >= 15) put_ext(ml - 19);
      } else {
        const int32_t lit = i - anchor;
        if (op + 1 + lit + ext_bytes(lit) + 2 + ext_bytes(ml) > cap) {
          fail = true;
          break;
        }
        if (lane == 0) dst[op] = (uint8_t)((min(lit, 15) << 4) | min(ml - 4, 15));
        op += 1;
        if (lit >= 15) put_ext(lit - 15);
        if (anchor >= base) {
          const int32_t l0 = anchor - base;
          if (lane >= l0 && lane < l0 + lit) dst[op + lane - l0] = (uint8_t)seq;
          op += lit;
        } else {
          put_stage(anchor, lit);
        }
        if (lane < 2) dst[op + lane] = (uint8_t)(off >> (8 * lane));
        op += 2;
        if (ml - 4 >= 15) put_ext(ml - 19);
      }
      anchor = ip = end;
      mask = ip - base >= 64 ? 0ull : mask & (~0ull << (ip - base));
    }
    if (fail) break;
  }
  const bool anynz = __ballot(nz != 0) != 0;
  if (lane == 0) {
    inf[I_HAS] = first ? 0 : 1;
    inf[I_POS] = S0 + f_pos;
    inf[I_OFF] = f_off;
    inf[I_ML] = f_ml;
    inf[I_LEN] = op;
    inf[I_TRAIL] = S0 + anchor;
    inf[I_NZ] = anynz ? 1 : 0;
    inf[I_FAIL] = fail ? 1 : 0;
  }
}

// Merge: one wavefront per stream joins its sub-blocks into one LZ4 block in
// the stream's slot (or records a zero / raw stream).  Stream j of the call is
// stream e0 + j.
__device__ __forceinline__ void merge_stream(const EncGeom &G, const uint8_t *__restrict__ packed, int64_t e0,
                                             int64_t j, int32_t K, const uint8_t *__restrict__ subslots,
                                             const int32_t *__restrict__ info, uint8_t *__restrict__ slots,
                                             int32_t *__restrict__ meta) {
  const int lane = threadIdx.x;
  const int64_t E = e0 + j;
  const int64_t sb0 = j * K;  // first sub-block record of this stream
  const int64_t per_chunk = (int64_t)G.nblocks * G.nstreams;
  const int64_t c = E / per_chunk, r = E % per_chunk;
  const int64_t b = r / G.nstreams, s = r % G.nstreams;
  const int32_t n = G.neblock;
  const IAMX_GLOBAL uint8_t *srcb =
      gptr<const uint8_t>((uint64_t)(packed + c * G.chunk_nbytes + b * G.blocksize + s * (int64_t)n));
  IAMX_GLOBAL uint8_t *dst = gptr<uint8_t>((uint64_t)(slots + E * G.slot));
  const IAMX_GLOBAL int32_t *inf = gptr<const int32_t>((uint64_t)(info + sb0 * kInfo));
  // sizes first (uniform scalar walk over the records)
  bool anynz = false, fail = false;
  int32_t run = 0, total = 0;
  for (int32_t k = 0; k < K; k++) {
    const IAMX_GLOBAL int32_t *q = inf + k * kInfo;
    anynz |= q[I_NZ] != 0;
    fail |= q[I_FAIL] != 0;
    if (q[I_HAS]) {
      const int32_t lit = q[I_POS] - run;
      total += 1 + (lit >= 15 ? ext_bytes(lit - 15) : 0) + lit + q[I_LEN];
      run = q[I_TRAIL];
    }
  }
  {
    const int32_t lit = n - run;
    total += 1 + (lit >= 15 ? ext_bytes(lit - 15) : 0) + lit;
  }
  int32_t cs;
  if (!anynz) {
    cs = 0;
  } else if (fail || total >= n) {
    cs = n;
  } else {
    cs = total;
    int32_t op = 0;
    run = 0;
    auto put_run = [&](int32_t lit, int32_t mlc) {  // token, literal continuation, literals [run, run+lit)
      if (lane == 0) dst[op] = (uint8_t)((min(lit, 15) << 4) | min(mlc, 15));
      op += 1;
      if (lit >= 15) {
        const int32_t v = lit - 15, nf = v / 255;
        for (int32_t q = lane; q <= nf; q += 64) dst[op + q] = (uint8_t)(q < nf ? 255 : v - 255 * nf);
        op += nf + 1;
      }
      wg_copy(dst + op, srcb + run, lit, lane, 64);
      op += lit;
    };
    for (int32_t k = 0; k < K; k++) {
      const IAMX_GLOBAL int32_t *q = inf + k * kInfo;
      if (!q[I_HAS]) continue;
      put_run(q[I_POS] - run, q[I_ML] - 4);
      const int32_t len = q[I_LEN];
      wg_copy(dst + op, gptr<const uint8_t>((uint64_t)(subslots + (sb0 + k) * kSubSlot)), len, lane, 64);
      op += len;
      run = q[I_TRAIL];
    }
    put_run(n - run, 0);
  }
  if (lane == 0) meta[2 * E] = cs;
}

// Chunk c: its streams' offsets, header and bstarts.
__device__ __forceinline__ void chunk_offsets(const EncGeom &G, int64_t c, int32_t *__restrict__ meta,
                                              uint8_t *__restrict__ out, int64_t *__restrict__ csizes) {
  __shared__ int32_t part[256];
  const int tid = threadIdx.x;
  const int32_t ne = G.nblocks * G.nstreams;
  const int64_t E0 = c * ne;
  const int32_t hdr = 32 + 4 * G.nblocks;
  IAMX_GLOBAL uint32_t *co = gptr<uint32_t>((uint64_t)(out + c * G.cap));
  int32_t carry = 0;
  for (int32_t t0 = 0; t0 < ne; t0 += 256) {
    const int32_t t = t0 + tid;
    const int32_t sz = t < ne ? 4 + meta[2 * (E0 + t)] : 0;
    part[tid] = sz;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
      const int32_t v = tid >= d ? part[tid - d] : 0;
      __syncthreads();
      part[tid] += v;
      __syncthreads();
    }
    const int32_t at = hdr + carry + part[tid] - sz;
    if (t < ne) {
      meta[2 * (E0 + t) + 1] = at;
      if (t % G.nstreams == 0) co[8 + t / G.nstreams] = (uint32_t)at;  // bstarts
    }
    carry += part[255];
    __syncthreads();
  }
  if (tid == 0) {
    const int32_t total = hdr + carry;
    const uint32_t flags = 0x05u | (G.split ? 0u : 0x10u) | 0x20u;
    co[0] = 5u | (1u << 8) | (flags << 16) | ((uint32_t)G.typesize << 24);
    co[1] = (uint32_t)(G.chunk_nbytes);
    co[2] = (uint32_t)G.blocksize;
    co[3] = (uint32_t)total;
    co[4] = 0u;
    const uint32_t fcode = G.filter == IAMX_FILTER_SHUFFLE ? 1u : G.filter == IAMX_FILTER_BITSHUFFLE ? 2u : 0u;
    co[5] = fcode << 8;  // byte 21: the filter in the last of the 6 filter slots
    co[6] = 0u;
    co[7] = 0u;
    csizes[c] = total;
  }
}

// Stream E: its csize word and bytes into its chunk.
__device__ __forceinline__ void chunk_copy(const EncGeom &G, int64_t E, const uint8_t *__restrict__ packed,
                                           const uint8_t *__restrict__ slots, const int32_t *__restrict__ meta,
                                           uint8_t *__restrict__ out) {
  const int64_t per_chunk = (int64_t)G.nblocks * G.nstreams;
  const int64_t c = E / per_chunk, r = E % per_chunk;
  const int64_t b = r / G.nstreams, s = r % G.nstreams;
  const int32_t cs = meta[2 * E], at = meta[2 * E + 1];
  IAMX_GLOBAL uint8_t *d = gptr<uint8_t>((uint64_t)(out + c * G.cap + at));
  if (threadIdx.x < 4) d[threadIdx.x] = (uint8_t)((uint32_t)cs >> (8 * threadIdx.x));
  if (cs == 0) return;
  const uint8_t *sp = cs == G.neblock ? packed + c * G.chunk_nbytes + b * G.blocksize + s * (int64_t)G.neblock
                                      : slots + E * G.slot;
  wg_copy(d + 4, gptr<const uint8_t>((uint64_t)sp), cs, threadIdx.x, blockDim.x);
}

// ---- one tiling per call --------------------------------------------------------
__global__ __launch_bounds__(64) void k_lz4_encode_sub(EncGeom G, const uint8_t *__restrict__ packed, int64_t e0,
                                                       int64_t nsub, int32_t K, uint8_t *__restrict__ subslots,
                                                       int32_t *__restrict__ info) {
  if ((int64_t)blockIdx.x >= nsub) return;
  encode_sub(G, packed, e0, blockIdx.x, K, subslots, info);
}

__global__ __launch_bounds__(64) void k_lz4_merge(EncGeom G, const uint8_t *__restrict__ packed, int64_t e0,
                                                  int njobs, int32_t K, const uint8_t *__restrict__ subslots,
                                                  const int32_t *__restrict__ info, uint8_t *__restrict__ slots,
                                                  int32_t *__restrict__ meta) {
  if ((int)blockIdx.x >= njobs) return;
  merge_stream(G, packed, e0, blockIdx.x, K, subslots, info, slots, meta);
}

__global__ __launch_bounds__(256) void k_chunk_offsets(EncGeom G, int64_t c0, int32_t *__restrict__ meta,
                                                       uint8_t *__restrict__ out, int64_t *__restrict__ csizes) {
  chunk_offsets(G, c0 + blockIdx.x, meta, out, csizes);
}

__global__ __launch_bounds__(256) void k_chunk_copy(EncGeom G, int64_t e0, const uint8_t *__restrict__ packed,
                                                    const uint8_t *__restrict__ slots,
                                                    const int32_t *__restrict__ meta, uint8_t *__restrict__ out) {
  chunk_copy(G, e0 + blockIdx.x, packed, slots, meta, out);
}

// ---- many tilings per call ------------------------------------------------------
// Segment i of a batch: one tiling's chunks.  Its sub-blocks, streams and
// chunks are numbered after those of segments 0..i-1 (sub0, st0, ch0); every
// buffer it touches starts at its own offset, so the one-tiling code above
// runs unchanged on segment-local indexes.
struct EncSeg {
  EncGeom G;
  int32_t K;           // sub-blocks per stream
  int32_t pad;
  int64_t sub0, st0, ch0;
  int64_t packed_off;  // bytes: its packed blocks in the packed buffer
  int64_t out_off;     // bytes: its chunks in the output (chunk c at out_off + c * G.cap)
  int64_t slot_off;    // bytes: its stream slots in the scratch
};

// the segment whose range of field F (0 sub-blocks, 1 streams, 2 chunks) holds i
template <int F>
__device__ __forceinline__ int seg_of(const EncSeg *__restrict__ S, int n, int64_t i) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int64_t v = F == 0 ? S[mid].sub0 : F == 1 ? S[mid].st0 : S[mid].ch0;
    if (v <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(64) void k_lz4_encode_sub_batch(const EncSeg *__restrict__ S, int n,
                                                             const uint8_t *__restrict__ packed,
                                                             uint8_t *__restrict__ subslots,
                                                             int32_t *__restrict__ info) {
  const EncSeg &q = S[seg_of<0>(S, n, blockIdx.x)];
  const EncGeom G = q.G;
  encode_sub(G, packed + q.packed_off, 0, (int64_t)blockIdx.x - q.sub0, q.K, subslots + q.sub0 * kSubSlot,
             info + q.sub0 * kInfo);
}

__global__ __launch_bounds__(64) void k_lz4_merge_batch(const EncSeg *__restrict__ S, int n,
                                                        const uint8_t *__restrict__ packed,
                                                        const uint8_t *__restrict__ subslots,
                                                        const int32_t *__restrict__ info,
                                                        uint8_t *__restrict__ slots, int32_t *__restrict__ meta) {
  const EncSeg &q = S[seg_of<1>(S, n, blockIdx.x)];
  const EncGeom G = q.G;
  merge_stream(G, packed + q.packed_off, 0, (int64_t)blockIdx.x - q.st0, q.K, subslots + q.sub0 * kSubSlot,
               info + q.sub0 * kInfo, slots + q.slot_off, meta + 2 * q.st0);
}

__global__ __launch_bounds__(256) void k_chunk_offsets_batch(const EncSeg *__restrict__ S, int n,
                                                             int32_t *__restrict__ meta, uint8_t *__restrict__ out,
                                                             int64_t *__restrict__ csizes) {
  const EncSeg &q = S[seg_of<2>(S, n, blockIdx.x)];
  const EncGeom G = q.G;
  chunk_offsets(G, (int64_t)blockIdx.x - q.ch0, meta + 2 * q.st0, out + q.out_off, csizes + q.ch0);
}

__global__ __launch_bounds__(256) void k_chunk_copy_batch(const EncSeg *__restrict__ S, int n,
                                                          const uint8_t *__restrict__ packed,
                                                          const uint8_t *__restrict__ slots,
                                                          const int32_t *__restrict__ meta,
                                                          uint8_t *__restrict__ out) {
  const EncSeg &q = S[seg_of<1>(S, n, blockIdx.x)];
  const EncGeom G = q.G;
  chunk_copy(G, (int64_t)blockIdx.x - q.st0, packed + q.packed_off, slots + q.slot_off, meta + 2 * q.st0,
             out + q.out_off);
}

bool geom_of(const iamx_tiling_t *t, int split, EncGeom *G) {
  if (!t || t->typesize < 1 || t->typesize > 8 || t->block[0] <= 0 || t->block[1] <= 0) return false;
  const int64_t bs = t->block[0] * t->block[1] * t->typesize;
  const int64_t nb = t->nblocks[0] * t->nblocks[1];
  if (bs <= 0 || bs > INT32_MAX / 2 || nb <= 0 || nb > (1 << 20)) return false;
  G->blocksize = (int32_t)bs;
  G->nblocks = (int32_t)nb;
  G->nstreams = split ? t->typesize : 1;
  G->neblock = (int32_t)(bs / G->nstreams);
  G->typesize = t->typesize;
  G->filter = t->filter;
  G->split = split ? 1 : 0;
  G->chunk_nbytes = bs * nb;
  G->cap = (32 + 4 * nb + 4 * nb * G->nstreams + G->chunk_nbytes + 15) & ~15ll;
  G->slot = (G->neblock + 15) & ~15ll;
  return G->chunk_nbytes <= INT32_MAX;
}

// scratch layout: stream slots (all streams of the tiling), then sub-block
// slots and info records (for the streams of one call)
struct ScratchLayout {
  int64_t nst_all, K, sub_off, info_off, bytes;
};

ScratchLayout layout_of(const iamx_tiling_t *t, const EncGeom &G) {
  ScratchLayout L;
  L.nst_all = t->nchunks[0] * t->nchunks[1] * (int64_t)G.nblocks * G.nstreams;
  L.K = (G.neblock + kSub - 1) / kSub;
  L.sub_off = L.nst_all * G.slot;
  L.info_off = L.sub_off + L.nst_all * L.K * kSubSlot;
  L.bytes = L.info_off + L.nst_all * L.K * kInfo * 4 + 256;
  return L;
}

// batch scratch: segment table, stream metadata, stream slots, sub-block slots, info records
struct BatchLayout {
  int64_t nsub, nst, nch, meta_off, slots_off, sub_off, info_off, bytes;
};

bool batch_layout(int n, const iamx_tiling_t *tilings, const int *split, EncSeg *segs, BatchLayout *B) {
  if (n < 1 || !tilings || !split) return false;
  int64_t sub = 0, st = 0, ch = 0, slot = 0;
  for (int i = 0; i < n; i++) {
    EncSeg q;
    memset(&q, 0, sizeof(q));
    if (!geom_of(&tilings[i], split[i], &q.G)) return false;
    const iamx_tiling_t *t = &tilings[i];
    const int64_t nch = t->nchunks[0] * t->nchunks[1];
    if (nch < 1) return false;
    const int64_t nst = nch * q.G.nblocks * q.G.nstreams;
    q.K = (q.G.neblock + kSub - 1) / kSub;
    q.sub0 = sub;
    q.st0 = st;
    q.ch0 = ch;
    q.slot_off = slot;
    if (segs) segs[i] = q;
    sub += nst * q.K;
    st += nst;
    ch += nch;
    slot += nst * q.G.slot;
  }
  if (sub > INT32_MAX || st > INT32_MAX || ch > INT32_MAX) return false;
  B->nsub = sub;
  B->nst = st;
  B->nch = ch;
  auto up = [](int64_t v) { return (v + 255) & ~255ll; };
  B->meta_off = up((int64_t)n * (int64_t)sizeof(EncSeg));
  B->slots_off = B->meta_off + up(st * 8);
  B->sub_off = B->slots_off + up(slot);
  B->info_off = B->sub_off + sub * kSubSlot;
  B->bytes = B->info_off + sub * kInfo * 4 + 256;
  return true;
}

}  // namespace

extern "C" {

IAMX_API int64_t iamx_encode_batch_scratch_bytes(int n, const iamx_tiling_t *tilings, const int *split) {
  BatchLayout B;
  return batch_layout(n, tilings, split, nullptr, &B) ? B.bytes : -1;
}

IAMX_API int iamx_encode_batch(int n, const iamx_tiling_t *tilings, const int *split, const void *d_packed,
                               const int64_t *packed_off, void *d_scratch, void *d_out, const int64_t *out_off,
                               int64_t *d_chunk_size) {
  int rc = iamx::ensure_init();
  if (rc) return rc;
  if (!d_packed || !packed_off || !d_scratch || !d_out || !out_off || !d_chunk_size)
    return iamx::set_error(IAMX_EINVAL, "iamx_encode_batch: null buffer");
  if (n < 1 || n > (1 << 20)) return iamx::set_error(IAMX_EINVAL, "iamx_encode_batch: %d tilings", n);
  std::vector<EncSeg> segs((size_t)n);
  BatchLayout B;
  if (!batch_layout(n, tilings, split, segs.data(), &B))
    return iamx::set_error(IAMX_EINVAL, "iamx_encode_batch: bad tiling or too many streams");
  for (int i = 0; i < n; i++) {
    if (packed_off[i] < 0 || out_off[i] < 0 || (out_off[i] & 15))
      return iamx::set_error(IAMX_EINVAL, "iamx_encode_batch: offsets of tiling %d", i);
    segs[(size_t)i].packed_off = packed_off[i];
    segs[(size_t)i].out_off = out_off[i];
  }
  uint8_t *scr = (uint8_t *)d_scratch;
  rc = iamx_h2d(scr, segs.data(), sizeof(EncSeg) * (size_t)n);
  if (rc) return rc;
  const EncSeg *d_segs = (const EncSeg *)scr;
  int32_t *meta = (int32_t *)(scr + B.meta_off);
  uint8_t *slots = scr + B.slots_off, *subslots = scr + B.sub_off;
  int32_t *info = (int32_t *)(scr + B.info_off);
  const uint8_t *pk = (const uint8_t *)d_packed;
  IAMX_LAUNCH("lz4_encode_batch", k_lz4_encode_sub_batch, dim3((unsigned)B.nsub), dim3(64), 0, d_segs, n, pk,
              subslots, info);
  IAMX_LAUNCH("lz4_merge_batch", k_lz4_merge_batch, dim3((unsigned)B.nst), dim3(64), 0, d_segs, n, pk,
              (const uint8_t *)subslots, (const int32_t *)info, slots, meta);
  IAMX_LAUNCH("chunk_offsets_batch", k_chunk_offsets_batch, dim3((unsigned)B.nch), dim3(256), 0, d_segs, n, meta,
              (uint8_t *)d_out, d_chunk_size);
  IAMX_LAUNCH("chunk_copy_batch", k_chunk_copy_batch, dim3((unsigned)B.nst), dim3(256), 0, d_segs, n, pk,
              (const uint8_t *)slots, (const int32_t *)meta, (uint8_t *)d_out);
  return IAMX_OK;
}

IAMX_API int64_t iamx_chunk_capacity(const iamx_tiling_t *t, int split) {
  EncGeom G;
  return geom_of(t, split, &G) ? G.cap : -1;
}

IAMX_API int64_t iamx_encode_scratch_bytes(const iamx_tiling_t *t, int split) {
  EncGeom G;
  return geom_of(t, split, &G) ? layout_of(t, G).bytes : -1;
}

IAMX_API int iamx_encode_chunks(const iamx_tiling_t *t, int split, const void *d_packed, int64_t chunk_begin,
                                int64_t chunk_end, void *d_scratch, int32_t *d_meta, void *d_out,
                                int64_t *d_chunk_size) {
  int rc = iamx::ensure_init();
  if (rc) return rc;
  EncGeom G;
  if (!geom_of(t, split, &G)) return iamx::set_error(IAMX_EINVAL, "iamx_encode_chunks: bad tiling");
  const int64_t nchunks = t->nchunks[0] * t->nchunks[1];
  if (chunk_begin < 0 || chunk_end > nchunks || chunk_begin > chunk_end)
    return iamx::set_error(IAMX_EINVAL, "iamx_encode_chunks: chunk range [%lld, %lld) outside [0, %lld)",
                           (long long)chunk_begin, (long long)chunk_end, (long long)nchunks);
  const ScratchLayout L = layout_of(t, G);
  const int64_t per_chunk = (int64_t)G.nblocks * G.nstreams;
  const int64_t nst = (chunk_end - chunk_begin) * per_chunk;
  if (nst == 0) return IAMX_OK;
  const int64_t nsub = nst * L.K;
  if (nsub > INT32_MAX || chunk_end - chunk_begin > INT32_MAX)
    return iamx::set_error(IAMX_EINVAL, "iamx_encode_chunks: too many streams");
  const int64_t e0 = chunk_begin * per_chunk;
  uint8_t *slots = (uint8_t *)d_scratch;
  uint8_t *subslots = slots + L.sub_off;
  int32_t *info = (int32_t *)(slots + L.info_off);
  IAMX_LAUNCH("lz4_encode", k_lz4_encode_sub, dim3((unsigned)nsub), dim3(64), 0, G, (const uint8_t *)d_packed, e0,
              nsub, (int32_t)L.K, subslots, info);
  IAMX_LAUNCH("lz4_merge", k_lz4_merge, dim3((unsigned)nst), dim3(64), 0, G, (const uint8_t *)d_packed, e0,
              (int)nst, (int32_t)L.K, (const uint8_t *)subslots, (const int32_t *)info, slots, d_meta);
  IAMX_LAUNCH("chunk_offsets", k_chunk_offsets, dim3((unsigned)(chunk_end - chunk_begin)), dim3(256), 0, G,
              chunk_begin, d_meta, (uint8_t *)d_out, d_chunk_size);
  IAMX_LAUNCH("chunk_copy", k_chunk_copy, dim3((unsigned)nst), dim3(256), 0, G, e0, (const uint8_t *)d_packed,
              (const uint8_t *)slots, (const int32_t *)d_meta, (uint8_t *)d_out);
  return IAMX_OK;
}

}  // extern "C"
