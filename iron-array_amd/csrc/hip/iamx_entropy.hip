// iamx_entropy.hip -- decoder of the entropy-coded Blosc2 codec Zstandard
// (codec format 4: c-blosc2 compresses each stream of a block into one zstd
// frame).  zlib streams (codec format 3) have no GPU decoder: the host
// library refuses such chunks when it plans them.
//
// k_zstd_streams   one 64-lane wavefront per zstd stream (jobs of other codecs
//                  return at once).  The frame is parsed uniformly by all
//                  lanes (every lane holds the same state; loads of one
//                  address are broadcast): frame header, blocks (raw, RLE,
//                  compressed).  A compressed block's literals are decoded
//                  into the tail of the stream's output slot -- raw and RLE
//                  literals by all lanes, Huffman streams by one lane each
//                  (four streams: lanes 0-3) with the table in LDS -- and its
//                  sequences are decoded uniformly from the backward bit
//                  stream with the three FSE tables in LDS; each sequence's
//                  literals and match are then copied by the 64 lanes.
//                  Literals sit at [slot end - literal count, slot end), and
//                  a block's output never passes what is left of them (its
//                  output fits the slot), so output writes never reach an
//                  unread literal.  A match reads earlier output of this
//                  stream from global memory, after a workgroup-scope
//                  release/acquire when it reads bytes stored since the last
//                  one (once per 64-byte step for matches that overlap
//                  themselves).
// The format is RFC 8878's; the oracle's restatement (oracle/zstd_block.c) is
// pinned against libzstd and c-blosc 1.21's zstd chunks, and this kernel
// against the oracle.  Malformed streams flag IAMX_STATUS_ZSTD; every read
// and write is bounded by the stream and the slot.
#include "iamx_internal.h"

namespace {

using iamx::gptr;

// ---- tables of the format (RFC 8878, 3.1.1.3.2) ----
__device__ const int16_t kLLDef[36] = {4, 3, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 1, 1, 1, 2, 2,
                                       2, 2, 2, 2, 2, 2, 2, 3, 2, 1, 1, 1, 1, 1, -1, -1, -1, -1};
__device__ const int16_t kMLDef[53] = {1, 4, 3, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                       1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                       1, 1, 1, 1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1, -1, -1};
__device__ const int16_t kOFDef[29] = {1, 1, 1, 1, 1, 1, 2, 2, 2, 1, 1, 1, 1, 1, 1,
                                       1, 1, 1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1};
__device__ const uint32_t kLLBase[36] = {0,  1,  2,  3,  4,  5,  6,  7,  8,   9,   10,  11,   12,   13,   14,    15,    16,    18,
                                         20, 22, 24, 28, 32, 40, 48, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536};
__device__ const uint8_t kLLBits[36] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,  0,  0,  0,  1,  1,
                                        1, 1, 2, 2, 3, 3, 4, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16};
__device__ const uint32_t kMLBase[53] = {3,  4,  5,  6,  7,  8,  9,  10,  11,  12,  13,   14,   15,   16,   17,   18,    19,    20,
                                         21, 22, 23, 24, 25, 26, 27, 28,  29,  30,  31,   32,   33,   34,   35,   37,    39,    41,
                                         43, 47, 51, 59, 67, 83, 99, 131, 259, 515, 1027, 2051, 4099, 8195, 16387, 32771, 65539};
__device__ const uint8_t kMLBits[53] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,  0,  0,  0,  0,  0,  0,  0,  0,  0, 0,
                                        0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 3, 3, 4, 4, 5, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16};

constexpr int kHufMax = 11;   // longest Huffman code
constexpr int kFseMax = 9;    // largest FSE accuracy log (literal / match lengths)
constexpr int kBlockMax = 131072;

__device__ __forceinline__ int highbit32(uint32_t v) { return v ? 31 - __builtin_clz(v) : -1; }

// Backward bit stream over global bytes [0, len): bit i is bit (i & 7) of
// byte i >> 3; fields are taken from bit `pos` downwards, bits below the
// start read as 0.  A 64-bit window caches bytes [cb, cb + 8).
struct BitsB {
  const IAMX_GLOBAL uint8_t *s;
  int64_t len, pos, cb;
  uint64_t win;
  __device__ bool init(const IAMX_GLOBAL uint8_t *p, int64_t n) {
    s = p;
    len = n;
    cb = -64;
    if (n <= 0) return false;
    const uint32_t last = p[n - 1];
    if (!last) return false;
    pos = (n - 1) * 8 + highbit32(last);
    return true;
  }
  __device__ void fill(int64_t hi_byte) {
    cb = hi_byte - 7 > 0 ? hi_byte - 7 : 0;
    uint64_t w = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int64_t q = cb + i;
      const uint64_t b = q < len ? (uint64_t)s[q] : 0ull;
      w |= b << (8 * i);
    }
    win = w;
  }
  // the next n bits (n <= 32)
  __device__ uint32_t peek(int n) {
    if (n == 0) return 0u;
    const int64_t lo = pos - n;
    const int64_t hi_byte = (pos - 1) >> 3;
    const int64_t lo_byte = lo > 0 ? lo >> 3 : 0;
    if (pos <= 0) return 0u;
    if (lo_byte < cb || hi_byte >= cb + 8) fill(hi_byte);
    const uint64_t m = (n == 32) ? 0xffffffffull : ((1ull << n) - 1);
    if (lo >= 0) return (uint32_t)((win >> (lo - 8 * cb)) & m);
    // bits [0, pos) shifted up by -lo (cb is 0 here)
    const uint64_t part = win & ((1ull << pos) - 1);
    return (uint32_t)((part << (-lo)) & m);
  }
  __device__ uint32_t read(int n) {
    const uint32_t v = peek(n);
    pos -= n;
    return v;
  }
};

// Forward bit stream (FSE table descriptions) over bytes [0, len).
struct BitsF {
  const IAMX_GLOBAL uint8_t *s;
  int64_t len, pos;
  __device__ uint32_t read(int n) {
    uint32_t v = 0;
    for (int i = 0; i < n; i++, pos++) {
      const int64_t q = pos >> 3;
      const uint32_t bit = q < len ? ((uint32_t)s[q] >> (pos & 7)) & 1u : 0u;
      v |= bit << i;
    }
    return v;
  }
};

// FSE decoding tables live in LDS: entry = symbol | nbits << 8 | baseline << 16.
struct Lds {
  uint32_t fse[3][1 << kFseMax];  // literal lengths, offsets, match lengths
  uint32_t wfse[64];              // Huffman weights (accuracy log <= 6)
  uint16_t huf[1 << kHufMax];     // symbol | nbits << 8
  int16_t norm[256];
  uint16_t next[256];
  uint8_t w[256];
};

__device__ __forceinline__ void wave_sync() { __syncthreads(); }

__device__ __forceinline__ void fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Build table t (log) from L.norm[0..nsym); uniform, lane 0 writes.
__device__ bool fse_build(Lds &L, uint32_t *t, int nsym, int log) {
  const int lane = threadIdx.x;
  if (log < 1 || log > kFseMax || nsym < 1 || nsym > 256) return false;
  const int size = 1 << log;
  int high = size - 1;
  for (int s = 0; s < nsym; s++) {
    const int c = L.norm[s];
    if (c == -1) {
      if (high < 0) return false;
      if (lane == 0) t[high] = (uint32_t)s;
      high--;
      if (lane == 0) L.next[s] = 1;
    } else {
      if (lane == 0) L.next[s] = (uint16_t)c;
    }
  }
  if (high < 0) return false;
  const int step = (size >> 1) + (size >> 3) + 3, mask = size - 1;
  int p = 0, placed = 0;
  for (int s = 0; s < nsym; s++) {
    const int c = L.norm[s];
    for (int i = 0; i < c; i++) {
      if (placed++ >= size) return false;
      if (lane == 0) t[p] = (uint32_t)s;
      do p = (p + step) & mask;
      while (p > high);
    }
  }
  if (p != 0) return false;
  wave_sync();
  // every entry takes the running count of its symbol in index order (lane 0)
  if (lane == 0) {
    for (int i = 0; i < size; i++) {
      const int s = (int)(t[i] & 255u);
      const uint32_t ns = L.next[s]++;
      const int nb = log - highbit32(ns);
      t[i] = (uint32_t)s | ((uint32_t)nb << 8) | (((ns << nb) - (uint32_t)size) << 16);
    }
  }
  wave_sync();
  return true;
}

// Read a table description into L.norm and build t; returns bytes used or -1.
__device__ int fse_read(Lds &L, uint32_t *t, const IAMX_GLOBAL uint8_t *s, int64_t n, int maxlog, int maxsym,
                        int *log_out) {
  const int lane = threadIdx.x;
  BitsF f{s, n, 0};
  const int log = (int)f.read(4) + 5;
  if (log > maxlog) return -1;
  int remaining = (1 << log) + 1, sym = 0;
  while (remaining > 1 && sym <= maxsym) {
    const int nbits = highbit32((uint32_t)remaining) + 1;
    const int maxv = (1 << nbits) - 1 - remaining;
    int v = (int)f.read(nbits - 1);
    if (v >= maxv) {
      v += (int)f.read(1) << (nbits - 1);
      if (v >= (1 << (nbits - 1))) v -= maxv;
    }
    const int prob = v - 1;
    if (lane == 0) L.norm[sym] = (int16_t)prob;
    sym++;
    remaining -= prob < 0 ? -prob : prob;
    if (prob == 0) {
      int rep;
      do {
        rep = (int)f.read(2);
        for (int i = 0; i < rep && sym <= maxsym; i++) {
          if (lane == 0) L.norm[sym] = 0;
          sym++;
        }
        if (f.pos > 8 * n) return -1;
      } while (rep == 3);
    }
    if (f.pos > 8 * n) return -1;
  }
  if (remaining != 1 || sym > maxsym + 1) return -1;
  wave_sync();
  if (!fse_build(L, t, sym, log)) return -1;
  *log_out = log;
  return (int)((f.pos + 7) >> 3);
}

__device__ bool fse_default(Lds &L, uint32_t *t, const int16_t *def, int n, int log) {
  const int lane = threadIdx.x;
  for (int i = lane; i < n; i += 64) L.norm[i] = def[i];
  wave_sync();
  return fse_build(L, t, n, log);
}

// Huffman tree description at s (n bytes available); builds L.huf and
// returns bytes used (or -1); *maxbits receives the table's bit count.
__device__ int huf_read(Lds &L, const IAMX_GLOBAL uint8_t *s, int64_t n, int *maxbits) {
  const int lane = threadIdx.x;
  if (n < 1) return -1;
  int nw = 0, used;
  const int hb = s[0];
  if (hb >= 128) {
    nw = hb - 127;
    used = 1 + (nw + 1) / 2;
    if (used > n) return -1;
    for (int i = lane; i < nw; i += 64) {
      const uint32_t b = s[1 + i / 2];
      L.w[i] = (uint8_t)((i & 1) ? (b & 15u) : (b >> 4));
    }
  } else {
    used = 1 + hb;
    if (used > n || hb < 1) return -1;
    int log = 0;
    const int d = fse_read(L, L.wfse, s + 1, hb, 6, 255, &log);
    if (d < 0 || d > hb) return -1;
    BitsB b;
    if (!b.init(s + 1 + d, hb - d)) return -1;
    uint32_t s1 = b.read(log), s2 = b.read(log);
    for (;;) {
      if (nw >= 255) return -1;
      uint32_t e = L.wfse[s1];
      if (lane == 0) L.w[nw] = (uint8_t)(e & 255u);
      nw++;
      s1 = (e >> 16) + b.read((int)((e >> 8) & 255u));
      if (b.pos < 0) {
        if (lane == 0) L.w[nw] = (uint8_t)(L.wfse[s2] & 255u);
        nw++;
        break;
      }
      if (nw >= 255) return -1;
      e = L.wfse[s2];
      if (lane == 0) L.w[nw] = (uint8_t)(e & 255u);
      nw++;
      s2 = (e >> 16) + b.read((int)((e >> 8) & 255u));
      if (b.pos < 0) {
        if (lane == 0) L.w[nw] = (uint8_t)(L.wfse[s1] & 255u);
        nw++;
        break;
      }
    }
  }
  wave_sync();
  uint32_t total = 0;
  for (int i = 0; i < nw; i++) {
    const int wt = L.w[i];
    if (wt > kHufMax) return -1;
    if (wt) total += 1u << (wt - 1);
  }
  if (!total || nw >= 256) return -1;
  const int mb = highbit32(total) + 1;
  if (mb > kHufMax) return -1;
  const uint32_t rest = (1u << mb) - total;
  if (rest & (rest - 1)) return -1;
  if (lane == 0) L.w[nw] = (uint8_t)(highbit32(rest) + 1);
  nw++;
  wave_sync();
  // entries: weights ascending, symbols ascending within a weight, 2^(w-1)
  // entries each; lane 0 walks the symbols, all lanes fill each run
  int pos = 0;
  for (int wt = 1; wt <= mb; wt++)
    for (int sy = 0; sy < nw; sy++)
      if (L.w[sy] == wt) {
        const int cnt = 1 << (wt - 1);
        const uint16_t e = (uint16_t)(sy | ((mb + 1 - wt) << 8));
        if (pos + cnt > (1 << mb)) return -1;
        for (int k = lane; k < cnt; k += 64) L.huf[pos + k] = e;
        pos += cnt;
      }
  if (pos != 1 << mb) return -1;
  wave_sync();
  *maxbits = mb;
  return used;
}

// Decode `count` literals from one Huffman stream (this lane only).
__device__ bool huf_stream(const Lds &L, int mb, const IAMX_GLOBAL uint8_t *s, int64_t n, IAMX_GLOBAL uint8_t *out,
                           int64_t count) {
  BitsB b;
  if (!b.init(s, n)) return false;
  for (int64_t i = 0; i < count; i++) {
    const uint32_t e = L.huf[b.peek(mb)];
    out[i] = (uint8_t)(e & 255u);
    b.pos -= (int)(e >> 8);
  }
  return b.pos == 0;
}

// One compressed block: src[0, n) of the frame's block; output from op on
// into dst (slot of oend bytes).  Returns the new output position or -1.
__device__ int64_t zstd_block(Lds &L, int *huf_bits, int *flog, int *have, uint32_t *rep,
                              const IAMX_GLOBAL uint8_t *s, int64_t n, IAMX_GLOBAL uint8_t *dst, int64_t op,
                              int64_t oend, int64_t &fenced) {
  const int lane = threadIdx.x;
  if (n < 1) return -1;
  const int ltype = s[0] & 3, sf = (s[0] >> 2) & 3;
  int64_t regen, csize = 0, hdr;
  int nstreams = 1;
  if (ltype < 2) {
    if (sf == 0 || sf == 2) {
      regen = s[0] >> 3;
      hdr = 1;
    } else if (sf == 1) {
      if (n < 2) return -1;
      regen = (s[0] >> 4) + ((int64_t)s[1] << 4);
      hdr = 2;
    } else {
      if (n < 3) return -1;
      regen = (s[0] >> 4) + ((int64_t)s[1] << 4) + ((int64_t)s[2] << 12);
      hdr = 3;
    }
  } else {
    const int bits = sf < 2 ? 10 : sf == 2 ? 14 : 18;
    hdr = sf < 2 ? 3 : sf == 2 ? 4 : 5;
    if (n < hdr) return -1;
    uint64_t h = 0;
    for (int i = 0; i < hdr; i++) h |= (uint64_t)s[i] << (8 * i);
    nstreams = sf == 0 ? 1 : 4;
    regen = (int64_t)((h >> 4) & ((1ull << bits) - 1));
    csize = (int64_t)((h >> (4 + bits)) & ((1ull << bits) - 1));
  }
  if (regen > kBlockMax || regen > oend - op) return -1;
  // the block's literals go to the slot's tail
  IAMX_GLOBAL uint8_t *lit = dst + (oend - regen);
  int64_t p = hdr;
  if (ltype == 0) {
    if (p + regen > n) return -1;
    for (int64_t k = lane; k < regen; k += 64) lit[k] = s[p + k];
    p += regen;
  } else if (ltype == 1) {
    if (p + 1 > n) return -1;
    const uint8_t v = s[p];
    for (int64_t k = lane; k < regen; k += 64) lit[k] = v;
    p += 1;
  } else {
    if (p + csize > n) return -1;
    int64_t q = p;
    if (ltype == 2) {
      const int u = huf_read(L, s + q, csize, huf_bits);
      if (u < 0) return -1;
      q += u;
    } else if (*huf_bits == 0) {
      return -1;
    }
    const int64_t left = p + csize - q;
    bool ok = true;
    if (nstreams == 1) {
      if (lane == 0) ok = huf_stream(L, *huf_bits, s + q, left, lit, regen);
    } else {
      if (left < 6) return -1;
      const int64_t l1 = s[q] | (s[q + 1] << 8), l2 = s[q + 2] | (s[q + 3] << 8), l3 = s[q + 4] | (s[q + 5] << 8);
      const int64_t l4 = left - 6 - l1 - l2 - l3;
      const int64_t per = (regen + 3) / 4;
      if (l4 < 1 || 3 * per > regen) return -1;
      if (lane < 4) {
        const int64_t soff = 6 + (lane > 0 ? l1 : 0) + (lane > 1 ? l2 : 0) + (lane > 2 ? l3 : 0);
        const int64_t slen = lane == 0 ? l1 : lane == 1 ? l2 : lane == 2 ? l3 : l4;
        const int64_t cnt = lane < 3 ? per : regen - 3 * per;
        ok = huf_stream(L, *huf_bits, s + q + soff, slen, lit + lane * per, cnt);
      }
    }
    if (__ballot(!ok)) return -1;
    p += csize;
  }
  fence();  // the literals are visible to every lane
  // sequences
  if (p >= n) return -1;
  int64_t nseq = s[p++];
  if (nseq >= 128) {
    if (nseq == 255) {
      if (p + 2 > n) return -1;
      nseq = s[p] + ((int64_t)s[p + 1] << 8) + 0x7F00;
      p += 2;
    } else {
      if (p + 1 > n) return -1;
      nseq = ((nseq - 128) << 8) + s[p];
      p += 1;
    }
  }
  int64_t lp = 0;
  if (nseq > 0) {
    if (p >= n) return -1;
    const int modes = s[p++];
    if (modes & 3) return -1;
    const int16_t *defs[3] = {kLLDef, kOFDef, kMLDef};
    const int dn[3] = {36, 29, 53}, dlog[3] = {6, 5, 6}, maxlog[3] = {9, 8, 9}, maxsym[3] = {35, 31, 52};
    for (int k = 0; k < 3; k++) {
      const int mode = (modes >> (6 - 2 * k)) & 3;
      if (mode == 0) {
        if (!fse_default(L, L.fse[k], defs[k], dn[k], dlog[k])) return -1;
        flog[k] = dlog[k];
      } else if (mode == 1) {
        if (p >= n || s[p] > maxsym[k]) return -1;
        if (lane == 0) L.fse[k][0] = (uint32_t)s[p];
        flog[k] = 0;
        p++;
        wave_sync();
      } else if (mode == 2) {
        int lg = 0;
        const int u = fse_read(L, L.fse[k], s + p, n - p, maxlog[k], maxsym[k], &lg);
        if (u < 0) return -1;
        flog[k] = lg;
        p += u;
      } else if (!have[k]) {
        return -1;
      }
      have[k] = 1;
    }
    BitsB b;
    if (!b.init(s + p, n - p)) return -1;
    uint32_t sll = b.read(flog[0]), sof = b.read(flog[1]), sml = b.read(flog[2]);
    for (int64_t i = 0; i < nseq; i++) {
      const uint32_t ell = L.fse[0][sll], eof = L.fse[1][sof], eml = L.fse[2][sml];
      const int llc = (int)(ell & 255u), ofc = (int)(eof & 255u), mlc = (int)(eml & 255u);
      if (llc > 35 || mlc > 52 || ofc > 31) return -1;
      const uint32_t ofv = (1u << ofc) + b.read(ofc);
      const int64_t ml = kMLBase[mlc] + b.read(kMLBits[mlc]);
      const int64_t ll = kLLBase[llc] + b.read(kLLBits[llc]);
      uint32_t off;
      if (ofv > 3) {
        off = ofv - 3;
        rep[2] = rep[1];
        rep[1] = rep[0];
        rep[0] = off;
      } else {
        const int idx = (int)ofv - (ll == 0 ? 0 : 1);
        if (idx == 0) {
          off = rep[0];
        } else {
          off = idx == 3 ? rep[0] - 1 : rep[idx];
          if (idx != 1) rep[2] = rep[1];
          rep[1] = rep[0];
          rep[0] = off;
        }
      }
      if (i + 1 < nseq) {
        sll = (ell >> 16) + b.read((int)((ell >> 8) & 255u));
        sml = (eml >> 16) + b.read((int)((eml >> 8) & 255u));
        sof = (eof >> 16) + b.read((int)((eof >> 8) & 255u));
        if (sll >= (1u << flog[0]) && flog[0] > 0) return -1;
        if (sml >= (1u << flog[2]) && flog[2] > 0) return -1;
        if (sof >= (1u << flog[1]) && flog[1] > 0) return -1;
      }
      if (lp + ll > regen || ml > oend - op - ll || off == 0 || (int64_t)off > op + ll) return -1;
      // literals: tail -> output (the output stays below the unread tail)
      for (int64_t k = 0; k < ll; k += 64) {
        const int64_t j = k + lane;
        uint8_t v = 0;
        if (j < ll) v = lit[lp + j];
        __builtin_amdgcn_wave_barrier();
        if (j < ll) dst[op + j] = v;
      }
      op += ll;
      lp += ll;
      // match: out[op + j] = out[op + j - off], j < ml
      const int64_t src0 = op - (int64_t)off;
      if (src0 + ml > fenced) {  // the source holds bytes stored since the last fence
        fence();
        fenced = op;
      }
      if ((int64_t)off >= ml) {
        for (int64_t k = 0; k < ml; k += 64) {
          const int64_t j = k + lane;
          uint8_t v = 0;
          if (j < ml) v = dst[src0 + j];
          __builtin_amdgcn_wave_barrier();
          if (j < ml) dst[op + j] = v;
        }
      } else {
        // each lane reads off * ceil((lane + 1) / off) bytes back (or off back
        // once off >= 64): always output finished before this step
        const int64_t back = off >= 64 ? (int64_t)off : (int64_t)off * (((int64_t)lane + off) / off);
        for (int64_t k = 0; k < ml; k += 64) {
          if (k > 0) fence();
          const int64_t j = k + lane;
          uint8_t v = 0;
          if (j < ml) v = dst[op + j - back];
          __builtin_amdgcn_wave_barrier();
          if (j < ml) dst[op + j] = v;
        }
        fenced = op;  // the last step's stores are not fenced yet
      }
      op += ml;
    }
    if (b.pos != 0) return -1;
  } else if (p != n) {
    return -1;
  }
  const int64_t rest = regen - lp;
  if (rest > oend - op) return -1;
  for (int64_t k = 0; k < rest; k += 64) {
    const int64_t j = k + lane;
    uint8_t v = 0;
    if (j < rest) v = lit[lp + j];
    __builtin_amdgcn_wave_barrier();
    if (j < rest) dst[op + j] = v;
  }
  fence();
  fenced = op + rest;
  return op + rest;
}

__global__ __launch_bounds__(64) void k_zstd_streams(const iamx_lz4_job_t *__restrict__ jobs, int njobs,
                                                     uint8_t *__restrict__ scratch, int64_t cap, int *status) {
  __shared__ Lds L;
  const int j = blockIdx.x;
  if (j >= njobs) return;
  const iamx_lz4_job_t job = jobs[j];
  if (job.codec != IAMX_CODEC_ZSTD) return;
  const int lane = threadIdx.x;
  if (job.dst_off < 0 || job.nbytes < 0 || job.csize < 0 ||
      job.dst_off > cap - (((int64_t)job.nbytes + 15) & ~(int64_t)15)) {
    if (lane == 0) atomicOr(status, IAMX_STATUS_RANGE);
    return;
  }
  const IAMX_GLOBAL uint8_t *src = gptr<const uint8_t>(job.src);
  IAMX_GLOBAL uint8_t *dst = gptr<uint8_t>((uint64_t)(scratch + job.dst_off));
  const int64_t iend = job.csize, oend = job.nbytes;
  int64_t ip = 0, op = 0, fenced = 0;
  bool bad = false;
  while (ip < iend && !bad) {
    if (iend - ip < 4) { bad = true; break; }
    const uint32_t magic = (uint32_t)src[ip] | ((uint32_t)src[ip + 1] << 8) | ((uint32_t)src[ip + 2] << 16) |
                           ((uint32_t)src[ip + 3] << 24);
    ip += 4;
    if ((magic & 0xFFFFFFF0u) == 0x184D2A50u) {  // skippable frame
      if (iend - ip < 4) { bad = true; break; }
      const int64_t len = (int64_t)((uint32_t)src[ip] | ((uint32_t)src[ip + 1] << 8) | ((uint32_t)src[ip + 2] << 16) |
                                    ((uint32_t)src[ip + 3] << 24));
      ip += 4;
      if (len > iend - ip) { bad = true; break; }
      ip += len;
      continue;
    }
    if (magic != 0xFD2FB528u || ip >= iend) { bad = true; break; }
    const int fhd = src[ip++];
    const int fcs_flag = fhd >> 6, single = (fhd >> 5) & 1, checksum = (fhd >> 2) & 1, did = fhd & 3;
    if ((fhd & 0x08) || did) { bad = true; break; }
    if (!single) ip += 1;
    ip += fcs_flag == 0 ? (single ? 1 : 0) : fcs_flag == 1 ? 2 : fcs_flag == 2 ? 4 : 8;
    if (ip > iend) { bad = true; break; }
    int huf_bits = 0, flog[3] = {0, 0, 0}, have[3] = {0, 0, 0};
    uint32_t rep[3] = {1, 4, 8};
    for (;;) {
      if (iend - ip < 3) { bad = true; break; }
      const uint32_t bh = (uint32_t)src[ip] | ((uint32_t)src[ip + 1] << 8) | ((uint32_t)src[ip + 2] << 16);
      ip += 3;
      const int last = bh & 1, type = (bh >> 1) & 3;
      const int64_t bsize = bh >> 3;
      if (type == 0) {
        if (bsize > iend - ip || bsize > oend - op) { bad = true; break; }
        for (int64_t k = lane; k < bsize; k += 64) dst[op + k] = src[ip + k];
        ip += bsize;
        op += bsize;
      } else if (type == 1) {
        if (ip >= iend || bsize > oend - op) { bad = true; break; }
        const uint8_t v = src[ip];
        for (int64_t k = lane; k < bsize; k += 64) dst[op + k] = v;
        ip += 1;
        op += bsize;
      } else if (type == 2) {
        if (bsize > iend - ip || bsize > kBlockMax) { bad = true; break; }
        fence();
        fenced = op;
        const int64_t r = zstd_block(L, &huf_bits, flog, have, rep, src + ip, bsize, dst, op, oend, fenced);
        if (r < 0) { bad = true; break; }
        op = r;
        ip += bsize;
      } else {
        bad = true;
        break;
      }
      if (last) break;
    }
    if (bad) break;
    if (checksum) ip += 4;
    if (ip > iend) bad = true;
  }
  if (!bad && op != oend) bad = true;
  if (bad && lane == 0) atomicOr(status, IAMX_STATUS_ZSTD);
}

}  // namespace

namespace iamx {

// Zstd decoding is switched on (sticky) once a zstd stream has been planned;
// until then decode launches skip its pass, and the LZ4 kernel flags any zstd
// job it meets.
static bool g_zstd_on = false;
bool zstd_enabled() { return g_zstd_on; }

int launch_zstd(const char *name, const iamx_lz4_job_t *d_jobs, int njobs, void *d_scratch, int64_t cap) {
  if (njobs <= 0) return IAMX_OK;
  IAMX_LAUNCH(name, k_zstd_streams, dim3(njobs), dim3(64), 0, d_jobs, njobs, (uint8_t *)d_scratch, cap,
              status_word());
  return IAMX_OK;
}

}  // namespace iamx

extern "C" {

IAMX_API void iamx_enable_zstd(int on) { iamx::g_zstd_on = on != 0; }

}  // extern "C"
