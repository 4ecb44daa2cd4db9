// iamx_decode.hip -- Blosc2 block decode (LZ4 + un-shuffle / bit-un-shuffle)
// and the C write-side block packing, for gfx950.
//
// Decode is two launches:
//   k_lz4_streams   one 64-lane wavefront per LZ4 stream.  The token stream is
//                   parsed uniformly by all lanes (every lane holds the same
//                   ip/op, loads of one address are broadcast), literal runs
//                   and matches are copied by the 64 lanes together.  A match
//                   copies out[op+i] = out[op-off + (i mod off)], so every
//                   source byte precedes op and one match never depends on
//                   itself; stores of earlier sequences are made visible with
//                   a workgroup-scope release/acquire only when the match
//                   source overlaps bytes stored since the last such fence.
//   k_unfilter      one workgroup per block: gathers each element's bytes from
//                   its byte planes (byte shuffle) or bit planes (bitshuffle,
//                   8x8 bit-matrix transposes), and writes the element into its
//                   tile of the dense destination, skipping the padding outside
//                   the valid extent.
// Stream sources are either the compressed chunk arena (raw streams, read in
// place), the LZ4 scratch, or a constant byte (zero / run streams, special
// chunks).
#include "iamx_internal.h"

#include <stdlib.h>

namespace {

// Delta-swap transpose of an 8x8 bit matrix held in a uint64 (row r = byte r,
// column c = bit c): bit 8r+c moves to bit 8c+r (Hacker's Delight, 7-3).
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

__device__ __forceinline__ uint32_t rdlane(uint32_t v, int32_t l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}

// One wavefront per LZ4 stream.
//   input   the wave keeps 1 KiB of the compressed stream in four VGPRs (lane
//           l of w[k] holds bytes wb+256k+4l..+3); sequence headers are parsed
//           uniformly with readlane, so the serial token chain does not wait
//           on memory.  The window slides by 256 bytes, loading w[3] two slides
//           before it is read, or jumps after a long literal run.
//   output  decoded bytes go to an LDS ring of RING bytes and reach the
//           stream's scratch slot in dword stores whenever RING/2 bytes are
//           pending.  Literal runs longer than 64 bytes are copied global ->
//           global in 16-byte groups per lane (dword loads realigned with
//           alignbyte) and also written to the ring in order.
//   matches with offset <= RING-64 are copied ring -> ring, 64 bytes per step,
//           each lane reading `off * ceil((lane+1)/off)` bytes back (`off`
//           back when off >= 64), so a step never reads what it writes.
//           Farther matches read the flushed stream back from global memory,
//           after a workgroup fence if it was stored since the last fence.
//   fast path  the headers of the next 64 input positions are decoded one
//           per lane; the walk over them reads each sequence's fields with
//           readlane and copies its literals and match with all 64 lanes.
// Ring steps store all 64 lanes: bytes past the sequence land on positions
// that are written again before anything reads them, and the slots they take
// belong to positions more than RING-64 bytes back, which no later match
// reads and which were flushed (at most RING/2 + 400 bytes are ever pending).
// With a 4 KiB ring the kernel fits 54 VGPRs: 8 wavefronts per SIMD, 32 per CU.
template <int RING>
__device__ __forceinline__ void lz4_stream(const iamx_lz4_job_t job, uint8_t *__restrict__ scratch, int *status,
                                           uint8_t *ring) {
  constexpr int32_t RM = RING - 1;
  constexpr int32_t NEAR = RING - 64;
  constexpr int32_t FLUSH_AT = RING / 2;
  constexpr int kRunUnroll = 4;  // 16-byte groups in flight per lane in a long literal run
  const int lane = threadIdx.x;
  const uint64_t abase = job.src & ~3ull;  // input positions are relative to this aligned address
  const IAMX_GLOBAL uint32_t *__restrict__ sw = iamx::gptr<const uint32_t>(abase);
  const IAMX_GLOBAL uint8_t *__restrict__ sb = iamx::gptr<const uint8_t>(abase);
  IAMX_GLOBAL uint8_t *dst = iamx::gptr<uint8_t>((uint64_t)(scratch + job.dst_off));
  IAMX_GLOBAL uint32_t *dstw = iamx::gptr<uint32_t>((uint64_t)(scratch + job.dst_off));
  const uint32_t *ring32 = reinterpret_cast<const uint32_t *>(ring);
  uint32_t *ring32w = reinterpret_cast<uint32_t *>(ring);
  const int32_t iend = (int32_t)(job.src & 3) + job.csize;
  const int32_t oend = job.nbytes;
  const int32_t lastw = (iend - 1) >> 2;  // last word holding stream bytes

  // window words past the stream repeat its last word (never parsed)
  auto ld_word = [&](int32_t pos) -> uint32_t { return sw[min(pos >> 2, lastw)]; };
  int32_t wb = 0;
  uint32_t w0 = ld_word(4 * lane), w1 = ld_word(256 + 4 * lane), w2 = ld_word(512 + 4 * lane),
           w3 = ld_word(768 + 4 * lane);
  auto jump = [&](int32_t pos) {
    wb = pos & ~3;
    w0 = ld_word(wb + 4 * lane);
    w1 = ld_word(wb + 256 + 4 * lane);
    w2 = ld_word(wb + 512 + 4 * lane);
    w3 = ld_word(wb + 768 + 4 * lane);
  };
  // make `pos - wb < 256`; positions only move forward
  auto refill = [&](int32_t pos) {
    const int32_t q = pos - wb;
    if (q >= 256) {
      if (q < 512) {
        w0 = w1;
        w1 = w2;
        w2 = w3;
        wb += 256;
        w3 = ld_word(wb + 768 + 4 * lane);
      } else {
        jump(pos);
      }
    }
  };
  // byte at input position pos, wb <= pos < wb + 512 (uniform)
  auto rbyte = [&](int32_t pos) -> uint32_t {
    const int32_t q = pos - wb;
    const uint32_t w = q < 256 ? rdlane(w0, q >> 2) : rdlane(w1, (q >> 2) - 64);
    return (w >> ((q & 3) << 3)) & 0xffu;
  };
  int32_t ip = (int32_t)(job.src & 3), op = 0, flushed = 0, fenced = 0, flush_at = FLUSH_AT;
  // store the whole dwords of ring bytes [flushed, cur) to the slot; a dword
  // is stored again only with the same bytes below `flushed`
  auto flush = [&](int32_t cur) {
    for (int32_t p = flushed + 4 * lane; p + 4 <= cur; p += 256) dstw[p >> 2] = ring32[(p & RM) >> 2];
    flushed = __builtin_amdgcn_readfirstlane(cur & ~3);  // uniform; keeps it in an SGPR
    flush_at = flushed + FLUSH_AT;
  };

  // Adds the length continuation bytes from ip on (255s, then one below 255)
  // to v, 64 input bytes a step (a ballot finds the last one); false when the
  // stream ends first or the length exceeds the block.
  auto read_len = [&](int32_t &v) -> bool {
    for (;;) {
      refill(ip);
      const int32_t q = ip - wb + lane;  // < 320
      const int32_t d = (q >> 2) & 63;
      const uint32_t a0 = (uint32_t)__shfl((int)w0, d, 64), a1 = (uint32_t)__shfl((int)w1, d, 64);
      const uint32_t b = ((q < 256 ? a0 : a1) >> ((q & 3) << 3)) & 255u;
      const uint64_t stop = __ballot(ip + lane < iend && b != 255u);
      if (stop) {
        const int32_t t = (int32_t)__builtin_ctzll(stop);
        v += 255 * t + (int32_t)rdlane(b, t);
        ip += t + 1;
        return true;
      }
      if (iend - ip < 64) return false;
      v += 255 * 64;
      ip += 64;
      if (v > oend) return false;
    }
  };

  int bad = 0;
  for (;;) {
    refill(ip);
    // Fast section (refill() above keeps ip - wb < 256).  Lane l decodes the
    // header of the sequence that would start at window position qs + l (the
    // window bytes are gathered with ds_bpermute): literal count, offset,
    // match length, and a control word: the header length plus literals
    // (`adv`) when the sequence can be decoded here, bit 31 when the walk
    // stops before it, bit 30 when the general path below must decode it.
    // Decoded here: up to 59 literals (read from w0/w1) and a match of
    // 4..273 bytes (at most one length byte each), an offset within the
    // output so far, the sequence inside the stream and at most 63 bytes
    // long, and the next header still among the 64 positions.  The walk
    // then costs one readlane per field, and it stops once the output
    // position passes `olim`: 333 bytes (the most one sequence here writes)
    // short of the slot end, and below the next flush.
    {
      const int32_t olim = min(oend - 333, flush_at - 1);
      while (op <= olim) {
        const int32_t qs = ip - wb;
        const int32_t r = qs + lane;  // <= 318
        auto gather = [&](int32_t d) -> uint32_t {  // window dword d < 128
          const uint32_t a0 = (uint32_t)__shfl((int)w0, d & 63, 64);
          const uint32_t a1 = (uint32_t)__shfl((int)w1, d & 63, 64);
          return d < 64 ? a0 : a1;
        };
        const uint32_t h = __builtin_amdgcn_alignbyte(gather((r >> 2) + 1), gather(r >> 2), (uint32_t)(r & 3));
        const uint32_t hl = (h >> 4) & 15u, hm = h & 15u, lb = (h >> 8) & 255u;
        const uint32_t lext = hl == 15 ? 1u : 0u;  // one literal length byte
        const uint32_t hlit = hl + (lext ? lb : 0u);
        const uint32_t hlsrc = 1 + lext;  // literals start this far after the token
        // offset field: <= 318 + 2 + 58 when the sequence is decoded here, and
        // only read through lane indexes taken mod 64 otherwise
        const int32_t r2 = r + (int32_t)(hlsrc + hlit);
        const uint32_t h2 = __builtin_amdgcn_alignbyte(gather((r2 >> 2) + 1), gather(r2 >> 2), (uint32_t)(r2 & 3));
        const uint32_t hoff = h2 & 0xffffu, he = (h2 >> 16) & 255u;
        const uint32_t ext = hm == 15 ? 1u : 0u;
        const uint32_t hml = hm + 4 + (ext ? he : 0u);
        const int32_t adv = 3 + (int32_t)(lext + hlit + ext);
        // (bitwise: one vector compare chain, no branches)
        const bool ok = (!(lext & (lb == 255)) & !(ext & (he == 255)) & (hoff >= 1) & ((int32_t)hoff <= op) &
                         (adv <= 63) & (adv <= iend - wb - r)) != 0;
        const uint32_t ctl = ok ? (lane + adv <= 63 ? (uint32_t)adv : 0x80000000u) : 0xC0000000u;
        // 1/off rounded up by 16 ulps: floor((lane + off) * rcp) is exact for
        // lane < 64, off < 64
        const uint32_t hrcp = __float_as_uint(__builtin_amdgcn_rcpf((float)hoff)) + 16u;
        int32_t l = 0;
        uint32_t c = rdlane(ctl, 0);
        while ((int32_t)(c | (uint32_t)(olim - op)) >= 0) {
          const int32_t lit = (int32_t)rdlane(hlit, l);
          const int32_t off = (int32_t)rdlane(hoff, l);
          const int32_t ml = (int32_t)rdlane(hml, l);
          const int32_t o2 = op + lit;
          if (lit) {
            const int32_t q = r + l + (int32_t)rdlane(hlsrc, l);  // window position of literal byte `lane`, <= 383
            const int32_t d = (q >> 2) & 63;
            const uint32_t a0 = (uint32_t)__shfl((int)w0, d, 64), a1 = (uint32_t)__shfl((int)w1, d, 64);
            ring[(op + lane) & RM] = (uint8_t)((q < 256 ? a0 : a1) >> ((q & 3) << 3));
            __builtin_amdgcn_wave_barrier();
          }
          // the first 64-byte step is peeled: most matches are one step
          if (off > NEAR) {
            // every source byte lies more than RING - 64 bytes back, and at
            // most RING/2 + 395 bytes are pending, so all of them were flushed
            const int32_t gend = o2 + ((ml + 63) & ~63) - off;
            if (gend > fenced) {
              __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
              __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
              fenced = flushed;
            }
            const int32_t d0 = o2 + lane;
            ring[d0 & RM] = dst[d0 - off];
            __builtin_amdgcn_wave_barrier();
            for (int32_t k = 64; k < ml; k += 64) {
              ring[(d0 + k) & RM] = dst[d0 + k - off];
              __builtin_amdgcn_wave_barrier();
            }
          }
          if (off <= NEAR) {
            int32_t back = off;  // off * ceil((lane + 1) / off)
            if (off < 64) {
              const float rc = __uint_as_float(rdlane(hrcp, l));
              back = (int32_t)__umul24((uint32_t)((float)(lane + off) * rc), (uint32_t)off);
            }
            const int32_t d0 = o2 + lane, s0 = d0 - back;
            const uint8_t v = ring[s0 & RM];
            __builtin_amdgcn_wave_barrier();
            ring[d0 & RM] = v;
            __builtin_amdgcn_wave_barrier();
            for (int32_t k = 64; k < ml; k += 64) {
              const uint8_t v2 = ring[(s0 + k) & RM];
              __builtin_amdgcn_wave_barrier();
              ring[(d0 + k) & RM] = v2;
              __builtin_amdgcn_wave_barrier();
            }
          }
          op = o2 + ml;
          l += (int32_t)c;
          c = rdlane(ctl, l);
        }
        ip = wb + qs + l;
        if (op > olim || (c & 0x40000000u)) break;  // a flush, or a sequence for the general path
        refill(ip);  // the walk left the 64 positions: the next ones
      }
    }
    if (op >= flush_at) {
      flush(op);
      continue;  // the fast loop again
    }
    refill(ip);
    if (ip >= iend) { bad = 1; break; }
    const uint32_t token = rbyte(ip);
    ++ip;
    int32_t lit = (int32_t)(token >> 4);
    if (lit == 15 && !read_len(lit)) { bad = 2; break; }
    if (lit > iend - ip || lit > oend - op) { bad = 3; break; }
    if (lit > 0) {
      if (lit <= 64) {
        // literals from the register window
        refill(ip);
        const int32_t q = ip - wb + lane;  // < 320
        const int sl = (q >> 2) & 63;
        const uint32_t a0 = (uint32_t)__shfl((int)w0, sl, 64), a1 = (uint32_t)__shfl((int)w1, sl, 64);
        const uint32_t v = ((q < 256 ? a0 : a1) >> ((q & 3) << 3)) & 0xffu;
        ring[(op + lane) & RM] = (uint8_t)v;
        __builtin_amdgcn_wave_barrier();
        ip += lit;
        op += lit;
        if (op >= flush_at) flush(op);
      } else {
        // long run: what is pending goes out first, then the run is copied
        // global -> global, 16 bytes a lane, the source realigned to the
        // destination's dwords with alignbyte; every run byte is also written
        // to the ring in position order, so the ring ends holding the run's
        // last RING bytes.  The window for the next header is requested now.
        flush(op);
        if (lane < op - flushed) dst[flushed + lane] = ring[(flushed + lane) & RM];
        const int32_t nip = ip + lit;
        if (nip - wb >= 256) jump(nip);
        const int32_t head = min((4 - (op & 3)) & 3, lit);  // bytes up to the first destination dword
        const int32_t ng = (lit - head) >> 4;                // whole 16-byte groups after the head
        const int32_t bend = head + (ng << 4);
        if (lane < head) {
          const uint8_t v = sb[ip + lane];
          dst[op + lane] = v;
          ring[(op + lane) & RM] = v;
        }
        const int32_t d0 = op + head;  // multiple of 4
        const int32_t sw0 = (ip + head) >> 2;
        const uint32_t sh = (uint32_t)((ip + head) & 3);  // the same for every group
        // one group: words k..k+3 hold run bytes, word k+4 only when sh > 0
        auto group_load = [&](int32_t g, uint32_t *w) {
          const int32_t k = sw0 + 4 * g;
          w[0] = sw[k];
          w[1] = sw[k + 1];
          w[2] = sw[k + 2];
          w[3] = sw[k + 3];
          w[4] = sw[min(k + 4, lastw)];
        };
        auto group_store = [&](int32_t g, const uint32_t *w) {
          const int32_t d = d0 + 16 * g;
          uint32_t o[4];
#pragma unroll
          for (int t = 0; t < 4; t++) o[t] = __builtin_amdgcn_alignbyte(w[t + 1], w[t], sh);
#pragma unroll
          for (int t = 0; t < 4; t++) dstw[(d >> 2) + t] = o[t];
#pragma unroll
          for (int t = 0; t < 4; t++) ring32w[((d + 4 * t) & RM) >> 2] = o[t];
        };
        int32_t g0 = 0;
        for (; g0 + 64 * kRunUnroll <= ng; g0 += 64 * kRunUnroll) {  // every lane has all groups
          uint32_t w[kRunUnroll][5];
#pragma unroll
          for (int u = 0; u < kRunUnroll; u++) group_load(g0 + 64 * u + lane, w[u]);
#pragma unroll
          for (int u = 0; u < kRunUnroll; u++) group_store(g0 + 64 * u + lane, w[u]);
        }
        for (; g0 < ng; g0 += 64) {
          if (g0 + lane < ng) {
            uint32_t w[5];
            group_load(g0 + lane, w);
            group_store(g0 + lane, w);
          }
        }
        __builtin_amdgcn_wave_barrier();
        if (lane < lit - bend) {  // fewer than 16 tail bytes, after the groups (ring order)
          const uint8_t v = sb[ip + bend + lane];
          dst[op + bend + lane] = v;
          ring[(op + bend + lane) & RM] = v;
        }
        __builtin_amdgcn_wave_barrier();
        ip = nip;
        op += lit;
        flushed = __builtin_amdgcn_readfirstlane(op & ~3);
        flush_at = flushed + FLUSH_AT;
      }
    }
    if (ip == iend) break;  // last sequence: literals only
    refill(ip);
    if (iend - ip < 2) { bad = 4; break; }
    const int32_t off = (int32_t)(rbyte(ip) | (rbyte(ip + 1) << 8));
    ip += 2;
    int32_t ml = (int32_t)(token & 15u);
    if (ml == 15 && !read_len(ml)) { bad = 5; break; }
    ml += 4;
    if (off == 0 || off > op || ml > oend - op) { bad = 6; break; }
    // back distance of this lane's source: off * ceil((lane + 1) / off) for off < 64
    int32_t back = off;
    if (off < 64) {
      const int32_t num = lane + off;
      int32_t qq = (int32_t)((float)num * __builtin_amdgcn_rcpf((float)off));
      const int32_t r = num - qq * off;
      qq += (r >= off) - (r < 0);
      back = off * qq;
    }
    if (off <= NEAR) {
      auto step = [&](int32_t k) {
        const uint8_t v = ring[(op + k + lane - back) & RM];
        __builtin_amdgcn_wave_barrier();
        ring[(op + k + lane) & RM] = v;
        __builtin_amdgcn_wave_barrier();
        const int32_t cur = op + min(ml, k + 64);
        if (cur >= flush_at) flush(cur);
      };
      // The match is periodic with period off, so position p may also copy
      // from p - backU, the first multiple of off at least 256 back (at most
      // 511 or off, both within the ring).  Once 256 bytes of that history
      // exist, four 64-byte steps read nothing the same four steps write and
      // go out together.
      constexpr int32_t U = 4;
      const int32_t backU = off >= 64 * U ? off : off * ((64 * U + off - 1) / off);
      int32_t k = 0;
      for (; k < ml && k < backU - off; k += 64) step(k);
      for (; k + 64 * U <= ml; k += 64 * U) {
        uint8_t v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = ring[(op + k + 64 * u + lane - backU) & RM];
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int u = 0; u < U; u++) ring[(op + k + 64 * u + lane) & RM] = v[u];
        __builtin_amdgcn_wave_barrier();
        if (op + k + 64 * U >= flush_at) flush(op + k + 64 * U);
      }
      for (; k < ml; k += 64) step(k);
    } else {
      for (int32_t k = 0; k < ml; k += 64) {
        // positions at or after `lo` are still in the ring; older ones were
        // flushed and are read back from global memory
        const int32_t lo = op + k + 64 - RING;
        const int32_t gmax = min(op + k + 64 - back, lo);  // end of the global positions this step reads
        if (gmax > fenced) {
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
          fenced = flushed;
        }
        const int32_t i = k + lane;
        const int32_t s = op + i - back;
        uint8_t v = 0;
        if (i < ml) v = s >= lo ? ring[s & RM] : dst[s];
        __builtin_amdgcn_wave_barrier();
        ring[(op + i) & RM] = v;
        __builtin_amdgcn_wave_barrier();
        const int32_t cur = op + min(ml, k + 64);
        if (cur >= flush_at) flush(cur);
      }
    }
    op += ml;
  }
  if (!bad && op != oend) bad = 7;
  if (!bad) flush((oend + 3) & ~3);  // the slot is padded to 16 bytes
  if (bad && lane == 0) atomicOr(status, 1 << bad);
}

// One BloscLZ stream by one wavefront (the stream format is restated in the
// oracle, oracle/blosclz_block.c, from c-blosc's compressor output): a
// control byte per token, literal runs of up to 32 bytes, matches of 3 or
// more bytes at distances up to 8191 + 1, or 8192 + 65535 through the
// far-distance escape.  The parse is uniform (a 256-byte register window of
// the input, bytes taken with readlane); every output byte is written to the
// scratch slot and to the LDS history ring.  Matches within RING - 64 bytes
// copy ring to ring in steps of min(distance, 64) bytes; farther ones read
// the slot in global memory after a release/acquire fence (again before
// every step that reads bytes of the same match).
template <int RING>
__device__ __forceinline__ void blosclz_stream(const iamx_lz4_job_t job, uint8_t *__restrict__ scratch,
                                               int *status, uint8_t *ring) {
  constexpr int32_t RM = RING - 1;
  const int lane = threadIdx.x;
  const uint64_t abase = job.src & ~3ull;
  const IAMX_GLOBAL uint32_t *__restrict__ sw = iamx::gptr<const uint32_t>(abase);
  const IAMX_GLOBAL uint8_t *__restrict__ sb = iamx::gptr<const uint8_t>(abase);
  IAMX_GLOBAL uint8_t *dst = iamx::gptr<uint8_t>((uint64_t)(scratch + job.dst_off));
  int32_t ip = (int32_t)(job.src & 3);
  const int32_t iend = ip + job.csize, oend = job.nbytes;
  const int32_t lastw = (iend - 1) >> 2;
  int32_t wb = 0;
  uint32_t w = sw[min(lane, lastw)];
  auto byte_at = [&](int32_t pos) -> uint32_t {  // pos uniform, pos < iend
    if (pos - wb >= 256) {
      wb = pos & ~3;
      w = sw[min((wb >> 2) + lane, lastw)];
    }
    const int32_t q = pos - wb;
    return (rdlane(w, q >> 2) >> ((q & 3) << 3)) & 0xffu;
  };
  auto fence = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  };
  int32_t op = 0;
  int bad = 0;
  uint32_t ctrl = byte_at(ip++) & 31u;
  for (;;) {
    if (ctrl >= 32) {
      int32_t len = (int32_t)(ctrl >> 5) - 1;
      int32_t dist = (int32_t)(ctrl & 31) << 8;
      if (len == 6) {
        uint32_t code;
        do {
          if (ip >= iend || len > oend) { bad = 1; break; }
          code = byte_at(ip++);
          len += (int32_t)code;
        } while (code == 255);
        if (bad) break;
      }
      if (ip >= iend) { bad = 1; break; }
      const uint32_t code = byte_at(ip++);
      len += 3;
      if (code == 255 && dist == (31 << 8)) {
        if (iend - ip < 2) { bad = 1; break; }
        dist = (int32_t)((byte_at(ip) << 8) | byte_at(ip + 1)) + 8192;
        ip += 2;
      } else {
        dist += (int32_t)code + 1;
      }
      if (dist > op || len > oend - op) { bad = 1; break; }
      if (dist <= RING - 64) {
        const int32_t step = min(dist, 64);
        for (int32_t k = 0; k < len; k += step) {
          const int32_t j = k + lane;
          const bool on = lane < step && j < len;
          uint8_t v = 0;
          if (on) v = ring[(op + j - dist) & RM];
          __builtin_amdgcn_wave_barrier();
          if (on) {
            ring[(op + j) & RM] = v;
            dst[op + j] = v;
          }
          __builtin_amdgcn_wave_barrier();
        }
      } else {
        fence();
        for (int32_t k = 0; k < len; k += 64) {
          if (k + 64 > dist) fence();  // this step reads bytes of this match
          const int32_t j = k + lane;
          if (j < len) {
            const uint8_t v = dst[op + j - dist];
            ring[(op + j) & RM] = v;
            dst[op + j] = v;
          }
          __builtin_amdgcn_wave_barrier();
        }
      }
      op += len;
    } else {
      const int32_t len = (int32_t)ctrl + 1;
      if (len > iend - ip || len > oend - op) { bad = 1; break; }
      if (lane < len) {
        const uint8_t v = sb[ip + lane];
        ring[(op + lane) & RM] = v;
        dst[op + lane] = v;
      }
      __builtin_amdgcn_wave_barrier();
      ip += len;
      op += len;
    }
    if (ip >= iend) break;
    ctrl = byte_at(ip++);
  }
  if (!bad && op != oend) bad = 1;
  if (bad && lane == 0) atomicOr(status, IAMX_STATUS_BLOSCLZ);
}

template <int RING>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RING <= 4096 ? 8 : 1, 8))) void k_lz4_streams(
    const iamx_lz4_job_t *__restrict__ jobs, int njobs, uint8_t *__restrict__ scratch, int64_t cap, int *status,
    int zstd_on) {
  __shared__ __align__(16) uint8_t ring[RING];
  const int j = blockIdx.x;
  if (j >= njobs) return;
  const iamx_lz4_job_t job = jobs[j];
  // the decoded stream, padded to 16 bytes, must lie inside the scratch
  if (job.dst_off < 0 || job.nbytes < 0 || job.dst_off > cap - (((int64_t)job.nbytes + 15) & ~(int64_t)15)) {
    if (threadIdx.x == 0) atomicOr(status, IAMX_STATUS_RANGE);
    return;
  }
  if (job.codec == IAMX_CODEC_BLOSCLZ) {
    blosclz_stream<RING>(job, scratch, status, ring);
    return;
  }
  if (job.codec == IAMX_CODEC_ZSTD) {  // decoded by a pass of its own
    if (!zstd_on && threadIdx.x == 0) atomicOr(status, IAMX_STATUS_CODEC);
    return;
  }
  if (job.codec != IAMX_CODEC_LZ4) {  // no decoder for this codec
    if (threadIdx.x == 0) atomicOr(status, IAMX_STATUS_CODEC);
    return;
  }
  lz4_stream<RING>(job, scratch, status, ring);
}

struct Planes {
  const IAMX_GLOBAL uint8_t *ptr[IAMX_MAX_STREAMS];
  uint8_t fill[IAMX_MAX_STREAMS];
  uint8_t is_const[IAMX_MAX_STREAMS];
};

// Byte q of the block's concatenated streams.
__device__ __forceinline__ uint8_t block_byte(const Planes &P, int32_t neblock, int32_t q) {
  const int s = q / neblock;
  const int32_t o = q - s * neblock;
  return P.is_const[s] ? P.fill[s] : P.ptr[s][o];
}

template <typename T>
__device__ __forceinline__ void put_elem(uint8_t *dst, int64_t idx, uint64_t v) {
  reinterpret_cast<T *>(dst)[idx] = (T)v;
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// 4x4 byte transpose: r[g] = bytes g of (a, b, c, d)
__device__ __forceinline__ void bytes4x4(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t r[4]) {
  const uint32_t t0 = __builtin_amdgcn_perm(b, a, 0x05010400u);  // a0 b0 a1 b1
  const uint32_t t1 = __builtin_amdgcn_perm(b, a, 0x07030602u);  // a2 b2 a3 b3
  const uint32_t t2 = __builtin_amdgcn_perm(d, c, 0x05010400u);  // c0 d0 c1 d1
  const uint32_t t3 = __builtin_amdgcn_perm(d, c, 0x07030602u);  // c2 d2 c3 d3
  r[0] = __builtin_amdgcn_perm(t2, t0, 0x05040100u);
  r[1] = __builtin_amdgcn_perm(t2, t0, 0x07060302u);
  r[2] = __builtin_amdgcn_perm(t3, t1, 0x05040100u);
  r[3] = __builtin_amdgcn_perm(t3, t1, 0x07060302u);
}

// The 4 bytes at p as a little-endian dword, p of any alignment: the aligned
// dword(s) holding them, funnel-shifted.  The second dword is read only when
// p is unaligned, and then it holds p[3], so no read leaves the page of the
// last byte asked for.
__device__ __forceinline__ uint32_t load_u32_any(const IAMX_GLOBAL uint8_t *p) {
  const uint64_t a = (uint64_t)p;
  const IAMX_GLOBAL uint32_t *w = iamx::gptr<const uint32_t>(a & ~3ull);
  const uint32_t sh = (uint32_t)(a & 3);
  if (sh == 0) return w[0];
  return __builtin_amdgcn_alignbyte(w[1], w[0], sh);
}

__device__ __forceinline__ void store_elem(uint8_t *dst, int ts, int64_t idx, uint64_t v) {
  switch (ts) {
    case 8: put_elem<uint64_t>(dst, idx, v); break;
    case 4: put_elem<uint32_t>(dst, idx, v); break;
    case 2: put_elem<uint16_t>(dst, idx, v); break;
    default: put_elem<uint8_t>(dst, idx, v); break;
  }
}

__global__ __launch_bounds__(256) void k_unfilter(const iamx_block_job_t *__restrict__ jobs, int njobs,
                                                  const uint8_t *__restrict__ scratch, uint8_t *__restrict__ dst) {
  __shared__ iamx_block_job_t J;
  __shared__ Planes P;
  const int bj = blockIdx.x;
  if (bj >= njobs) return;
  if (threadIdx.x == 0) {
    J = jobs[bj];
    for (int s = 0; s < IAMX_MAX_STREAMS; s++) {
      P.is_const[s] = (s >= J.nstreams) || J.kind[s] == IAMX_SRC_CONST;
      P.fill[s] = s < J.nstreams ? J.fill[s] : 0;
      P.ptr[s] = J.kind[s] == IAMX_SRC_ADDR      ? iamx::gptr<const uint8_t>(J.src[s])
                 : J.kind[s] == IAMX_SRC_SCRATCH ? iamx::gptr<const uint8_t>((uint64_t)(scratch + J.src[s]))
                                                 : nullptr;
    }
  }
  __syncthreads();
  const int ts = J.typesize;
  const int32_t n = J.bsize / ts;  // elements in the block
  const int32_t cols = J.cols;
  const int64_t ld = J.dst_ld;
  const int64_t base = J.dst_off;
  // a transposed destination takes tile element (r, c) at base + c * ld + r
  const bool tr = (J.filter & IAMX_JOB_TRANSPOSE) != 0;
  const int64_t sr = tr ? 1 : ld, sc = tr ? ld : 1;
  const int filter = J.filter & ~IAMX_JOB_TRANSPOSE;
  if (filter == IAMX_FILTER_BITSHUFFLE || filter == IAMX_FILTER_BITSHUFFLE_V2) {
    const int32_t n8 = filter == IAMX_FILTER_BITSHUFFLE ? n - n % 8 : n;
    const int32_t rowlen = n8 / 8;
    const int32_t ngroups = rowlen + (n - n8);  // 8-element groups, then tail elements one by one
    for (int32_t g = threadIdx.x; g < ngroups; g += blockDim.x) {
      uint64_t elems[8];
      int cnt;
      int32_t e0;
      if (g < rowlen) {
        for (int i = 0; i < 8; i++) elems[i] = 0;
        for (int j = 0; j < ts; j++) {
          uint64_t x = 0;
          for (int k = 0; k < 8; k++) {
            const int32_t q = (8 * j + k) * rowlen + g;
            const uint8_t b = J.nstreams == 1 ? (P.is_const[0] ? P.fill[0] : P.ptr[0][q]) : block_byte(P, J.neblock, q);
            x |= (uint64_t)b << (8 * k);
          }
          const uint64_t y = transpose8(x);
          for (int i = 0; i < 8; i++) elems[i] |= ((y >> (8 * i)) & 0xffull) << (8 * j);
        }
        cnt = 8;
        e0 = 8 * g;
      } else {
        const int32_t e = n8 + (g - rowlen);
        uint64_t v = 0;
        for (int j = 0; j < ts; j++) v |= (uint64_t)block_byte(P, J.neblock, e * ts + j) << (8 * j);
        elems[0] = v;
        cnt = 1;
        e0 = e;
      }
      int32_t r = e0 / cols, c = e0 - r * cols;
      for (int i = 0; i < cnt; i++) {
        if (r < J.vrows && c < J.vcols) store_elem(dst, ts, base + (int64_t)r * sr + (int64_t)c * sc, elems[i]);
        if (++c == cols) { c = 0; ++r; }
      }
    }
    return;
  }
  // byte shuffle (planes) or no filter
  const bool shuffled = filter == IAMX_FILTER_SHUFFLE && ts > 1;
  const IAMX_GLOBAL uint8_t *plane[8];
  uint8_t pconst[8], pfill[8];
  if (shuffled) {
    for (int j = 0; j < ts; j++) {
      if (J.nstreams == ts) {
        plane[j] = P.ptr[j]; pconst[j] = P.is_const[j]; pfill[j] = P.fill[j];
      } else {  // one stream holds all planes back to back
        plane[j] = P.ptr[0] ? P.ptr[0] + (int64_t)j * n : nullptr; pconst[j] = P.is_const[0]; pfill[j] = P.fill[0];
      }
    }
  }
  if (shuffled && (ts == 4 || ts == 8) && (J.nstreams == ts || J.nstreams == 1)) {
    // Four consecutive elements per thread: one dword from each plane (a last
    // group of fewer than 4 elements reads its bytes one by one, so nothing
    // past the planes is read) and a 4x4 byte transpose per 4 planes.  The
    // four are stored together when they share a row of an untransposed
    // destination, one by one otherwise.
    for (int32_t q = threadIdx.x; 4 * q < n; q += blockDim.x) {
      const int32_t e0 = 4 * q;
      const int32_t ne = min(4, n - e0);
      uint32_t w[8];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        if (j >= ts) continue;
        if (pconst[j]) {
          w[j] = 0x01010101u * pfill[j];
        } else if (ne == 4) {
          w[j] = load_u32_any(plane[j] + e0);
        } else {
          uint32_t v = 0;
          for (int i = 0; i < ne; i++) v |= (uint32_t)plane[j][e0 + i] << (8 * i);
          w[j] = v;
        }
      }
      uint32_t lo[4], hi[4] = {0, 0, 0, 0};
      bytes4x4(w[0], w[1], w[2], w[3], lo);
      if (ts == 8) bytes4x4(w[4], w[5], w[6], w[7], hi);
      int32_t r = e0 / cols, c = e0 - r * cols;
      if (!tr && ne == 4 && r < J.vrows && c + 4 <= J.vcols) {
        const int64_t idx = base + (int64_t)r * ld + c;
        if (((uint64_t)(dst + idx * ts) & 15) == 0) {
          if (ts == 4) {
            *iamx::gptr<u32x4>((uint64_t)(dst + idx * 4)) = u32x4{lo[0], lo[1], lo[2], lo[3]};
          } else {
            IAMX_GLOBAL u32x4 *o = iamx::gptr<u32x4>((uint64_t)(dst + idx * 8));
            o[0] = u32x4{lo[0], hi[0], lo[1], hi[1]};
            o[1] = u32x4{lo[2], hi[2], lo[3], hi[3]};
          }
        } else {
#pragma unroll
          for (int i = 0; i < 4; i++) store_elem(dst, ts, idx + i, ((uint64_t)hi[i] << 32) | lo[i]);
        }
      } else {
#pragma unroll
        for (int i = 0; i < 4; i++) {
          if (i < ne && r < J.vrows && c < J.vcols)
            store_elem(dst, ts, base + (int64_t)r * sr + (int64_t)c * sc, ((uint64_t)hi[i] << 32) | lo[i]);
          if (++c == cols) {
            c = 0;
            ++r;
          }
        }
      }
    }
    return;
  }
  for (int32_t r = threadIdx.x / 64; r < J.vrows; r += blockDim.x / 64) {
    for (int32_t c = threadIdx.x % 64; c < J.vcols; c += 64) {
      const int32_t e = r * cols + c;
      uint64_t v = 0;
      if (shuffled) {
        for (int j = 0; j < ts; j++) v |= (uint64_t)(pconst[j] ? pfill[j] : plane[j][e]) << (8 * j);
      } else {
        for (int j = 0; j < ts; j++) v |= (uint64_t)block_byte(P, J.neblock, e * ts + j) << (8 * j);
      }
      store_elem(dst, ts, base + (int64_t)r * sr + (int64_t)c * sc, v);
    }
  }
}

__device__ __forceinline__ uint64_t load_elem(const uint8_t *src, int ts, int64_t idx) {
  switch (ts) {
    case 8: return reinterpret_cast<const uint64_t *>(src)[idx];
    case 4: return reinterpret_cast<const uint32_t *>(src)[idx];
    case 2: return reinterpret_cast<const uint16_t *>(src)[idx];
    default: return src[idx];
  }
}

// One workgroup per C block: gather the tile (zeros outside the valid extent),
// apply the filter, write the block's bytes at (chunk * nblocks + block) * bsize.
__global__ __launch_bounds__(256) void k_pack(iamx_tiling_t t, const uint8_t *__restrict__ src, int64_t ld,
                                              uint8_t *__restrict__ out, int64_t id0) {
  const int64_t nbpc = t.nblocks[0] * t.nblocks[1];
  const int64_t id = id0 + blockIdx.x;
  const int64_t nc = id / nbpc, nb = id % nbpc;
  const int64_t ci0 = nc / t.nchunks[1], ci1 = nc % t.nchunks[1];
  const int64_t bi0 = nb / t.nblocks[1], bi1 = nb % t.nblocks[1];
  const int ts = t.typesize;
  const int32_t rows = (int32_t)t.block[0], cols = (int32_t)t.block[1];
  const int32_t n = rows * cols;
  const int64_t r_in = bi0 * t.block[0], c_in = bi1 * t.block[1];
  const int64_t gr0 = ci0 * t.chunk[0] + r_in, gc0 = ci1 * t.chunk[1] + c_in;
  int64_t vr = t.chunk[0] - r_in;
  if (t.shape[0] - gr0 < vr) vr = t.shape[0] - gr0;
  int64_t vc = t.chunk[1] - c_in;
  if (t.shape[1] - gc0 < vc) vc = t.shape[1] - gc0;
  uint8_t *o = out + id * (int64_t)n * ts;
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
        const uint64_t y = transpose8(x);
        for (int k = 0; k < 8; k++) o[(int64_t)(8 * j + k) * rowlen + g] = (uint8_t)(y >> (8 * k));
      }
    }
    for (int32_t e = n8 + threadIdx.x; e < n; e += blockDim.x) {
      const uint64_t v = elem(e);
      for (int j = 0; j < ts; j++) o[(int64_t)e * ts + j] = (uint8_t)(v >> (8 * j));
    }
    return;
  }
  const bool shuffled = t.filter == IAMX_FILTER_SHUFFLE && ts > 1;
  if (shuffled && (ts == 4 || ts == 8)) {
    // Four consecutive elements per thread, read together where they share a
    // row inside the valid extent, one by one otherwise; their bytes j form
    // 4 bytes of plane j (a 4x4 byte transpose per 4 planes), stored as one
    // dword where it is aligned.
    for (int32_t q = threadIdx.x; 4 * q < n; q += blockDim.x) {
      const int32_t e0 = 4 * q;
      const int32_t ne = min(4, n - e0);
      const int32_t r = e0 / cols, c = e0 - r * cols;
      uint32_t lo[4], hi[4] = {0, 0, 0, 0};
      const int64_t idx = (gr0 + r) * ld + gc0 + c;
      // vc counts valid columns from the block's first column to the chunk's
      // (or array's) edge and can pass the block width: the four elements
      // share a row only when they also end inside the block
      if (ne == 4 && r < vr && c + 4 <= vc && c + 4 <= cols && ((uint64_t)(src + idx * ts) & 15) == 0) {
        const IAMX_GLOBAL u32x4 *s = iamx::gptr<const u32x4>((uint64_t)(src + idx * ts));
        const u32x4 a = s[0];
        if (ts == 4) {
          lo[0] = a.x; lo[1] = a.y; lo[2] = a.z; lo[3] = a.w;
        } else {
          const u32x4 b = s[1];
          lo[0] = a.x; hi[0] = a.y; lo[1] = a.z; hi[1] = a.w;
          lo[2] = b.x; hi[2] = b.y; lo[3] = b.z; hi[3] = b.w;
        }
      } else {
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int32_t e = e0 + i, er = e / cols, ec = e - er * cols;
          uint64_t v = 0;
          if (i < ne && er < vr && ec < vc) v = load_elem(src, ts, (gr0 + er) * ld + gc0 + ec);
          lo[i] = (uint32_t)v;
          hi[i] = (uint32_t)(v >> 32);
        }
      }
      uint32_t w[8];
      bytes4x4(lo[0], lo[1], lo[2], lo[3], w);
      if (ts == 8) bytes4x4(hi[0], hi[1], hi[2], hi[3], w + 4);
#pragma unroll
      for (int j = 0; j < 8; j++) {
        if (j >= ts) continue;
        uint8_t *d = o + (int64_t)j * n + e0;
        if (ne == 4 && ((uint64_t)d & 3) == 0) {
          *iamx::gptr<uint32_t>((uint64_t)d) = w[j];
        } else {
          for (int i = 0; i < ne; i++) d[i] = (uint8_t)(w[j] >> (8 * i));
        }
      }
    }
    return;
  }
  for (int32_t e = threadIdx.x; e < n; e += blockDim.x) {
    const uint64_t v = elem(e);
    if (shuffled) {
      for (int j = 0; j < ts; j++) o[(int64_t)j * n + e] = (uint8_t)(v >> (8 * j));
    } else {
      for (int j = 0; j < ts; j++) o[(int64_t)e * ts + j] = (uint8_t)(v >> (8 * j));
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_transpose(int64_t rows, int64_t cols, const T *__restrict__ src,
                                                   T *__restrict__ dst) {
  __shared__ T tile[32][33];
  const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 32 x 8
  for (int i = ty; i < 32; i += 8) {
    const int64_t r = r0 + i, c = c0 + tx;
    if (r < rows && c < cols) tile[i][tx] = src[r * cols + c];
  }
  __syncthreads();
  for (int i = ty; i < 32; i += 8) {
    const int64_t c = c0 + i, r = r0 + tx;  // dst is cols x rows
    if (r < rows && c < cols) dst[c * rows + r] = tile[tx][i];
  }
}

}  // namespace

namespace iamx {

// History ring of the LZ4 kernel in KiB (4, 8 or 16): a bigger ring serves
// more matches from LDS, a smaller one fits more wavefronts per CU.
static int lz4_ring_kib() {
  static const int kib = [] {
    const char *e = getenv("IAMX_LZ4_RING_KIB");
    const int v = e ? atoi(e) : 4;
    return (v == 8 || v == 16) ? v : 4;
  }();
  return kib;
}

int launch_lz4(const char *name, const iamx_lz4_job_t *d_lz4, int n_lz4, void *d_scratch, int64_t cap) {
  const int z = zstd_enabled() ? 1 : 0;
  switch (lz4_ring_kib()) {
    case 8:
      IAMX_LAUNCH(name, k_lz4_streams<8192>, dim3(n_lz4), dim3(64), 0, d_lz4, n_lz4, (uint8_t *)d_scratch, cap,
                  status_word(), z);
      break;
    case 16:
      IAMX_LAUNCH(name, k_lz4_streams<16384>, dim3(n_lz4), dim3(64), 0, d_lz4, n_lz4, (uint8_t *)d_scratch, cap,
                  status_word(), z);
      break;
    default:
      IAMX_LAUNCH(name, k_lz4_streams<4096>, dim3(n_lz4), dim3(64), 0, d_lz4, n_lz4, (uint8_t *)d_scratch, cap,
                  status_word(), z);
      break;
  }
  if (z) return launch_zstd("zstd_streams", d_lz4, n_lz4, d_scratch, cap);
  return IAMX_OK;
}

}  // namespace iamx

extern "C" {

IAMX_API int iamx_decode(const iamx_lz4_job_t *d_lz4, int n_lz4, const iamx_block_job_t *d_blocks, int n_blocks,
                         void *d_scratch, void *d_dst) {
  int rc = iamx::ensure_init();
  if (rc) return rc;
  if (n_lz4 < 0 || n_blocks < 0) return iamx::set_error(IAMX_EINVAL, "iamx_decode: negative job count");
  if (n_lz4 > 0)
    rc = iamx::launch_lz4("lz4_streams", d_lz4, n_lz4, d_scratch);
  if (rc) return rc;
  if (n_blocks > 0)
    IAMX_LAUNCH("unfilter", k_unfilter, dim3(n_blocks), dim3(256), 0, d_blocks, n_blocks,
                (const uint8_t *)d_scratch, (uint8_t *)d_dst);
  return IAMX_OK;
}

IAMX_API int iamx_pack_blocks_range(const iamx_tiling_t *t, const void *d_src, int64_t src_ld, void *d_out,
                                    int64_t chunk_begin, int64_t chunk_end) {
  int rc = iamx::ensure_init();
  if (rc) return rc;
  if (!t || t->typesize < 1 || t->typesize > 8 || t->block[0] <= 0 || t->block[1] <= 0)
    return iamx::set_error(IAMX_EINVAL, "iamx_pack_blocks: bad tiling");
  const int64_t nchunks = t->nchunks[0] * t->nchunks[1], nbpc = t->nblocks[0] * t->nblocks[1];
  if (chunk_begin < 0 || chunk_end > nchunks || chunk_begin > chunk_end)
    return iamx::set_error(IAMX_EINVAL, "iamx_pack_blocks: chunk range [%lld, %lld) outside [0, %lld)",
                           (long long)chunk_begin, (long long)chunk_end, (long long)nchunks);
  const int64_t nblk = (chunk_end - chunk_begin) * nbpc;
  if (nblk == 0) return IAMX_OK;
  if (nblk < 0 || nblk > INT32_MAX) return iamx::set_error(IAMX_EINVAL, "iamx_pack_blocks: bad block count");
  IAMX_LAUNCH("pack_blocks", k_pack, dim3((unsigned)nblk), dim3(256), 0, *t, (const uint8_t *)d_src, src_ld,
              (uint8_t *)d_out, chunk_begin * nbpc);
  return IAMX_OK;
}

IAMX_API int iamx_pack_blocks(const iamx_tiling_t *t, const void *d_src, int64_t src_ld, void *d_out) {
  if (!t) return iamx::set_error(IAMX_EINVAL, "iamx_pack_blocks: bad tiling");
  return iamx_pack_blocks_range(t, d_src, src_ld, d_out, 0, t->nchunks[0] * t->nchunks[1]);
}

IAMX_API int iamx_transpose(int typesize, int64_t rows, int64_t cols, const void *d_src, void *d_dst) {
  int rc = iamx::ensure_init();
  if (rc) return rc;
  dim3 grid((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32));
  if (typesize == 8)
    IAMX_LAUNCH("transpose", k_transpose<uint64_t>, grid, dim3(256), 0, rows, cols, (const uint64_t *)d_src,
                (uint64_t *)d_dst);
  else if (typesize == 4)
    IAMX_LAUNCH("transpose", k_transpose<uint32_t>, grid, dim3(256), 0, rows, cols, (const uint32_t *)d_src,
                (uint32_t *)d_dst);
  else
    return iamx::set_error(IAMX_EINVAL, "iamx_transpose: typesize %d", typesize);
  return IAMX_OK;
}

}  // extern "C"
