/*
 * deflate_block.c -- TEST INFRASTRUCTURE (oracle).  zlib stream decoder.
 *
 * ZLIB is one of the codecs a Blosc2 chunk may carry (codec format 3 in the
 * chunk flags, bits 5-7) and one BTune may choose per chunk
 * (src/iarray.c:477-498, include/libiarray/iarray.h:177-187); c-blosc2
 * compresses each stream of a block with zlib's compress2, i.e. one zlib
 * stream (RFC 1950: a 2-byte header, a deflate stream, the Adler-32 of the
 * data).  This restates the published deflate format (RFC 1951) -- stored,
 * fixed-Huffman and dynamic-Huffman blocks, canonical codes read one bit at a
 * time, lengths 3..258 and distances up to 32768 -- and is pinned by
 * decoding the output of zlib's compress2 at every level (Python's zlib
 * module, black box) and c-blosc 1.21.0's zlib chunks
 * (tests/test_oracle_codec.py).  The Adler-32 is not checked.  Returns the
 * decoded size or a negative error.
 */
#include <string.h>

#include "oracle.h"

typedef struct {
  const uint8_t *s;
  int64_t n, pos; /* bit position, LSB first within each byte */
  int err;
} dbits_t;

static uint32_t dbits(dbits_t *b, int k) {
  uint32_t v = 0;
  for (int i = 0; i < k; i++, b->pos++) {
    if ((b->pos >> 3) >= b->n) {
      b->err = 1;
      return 0;
    }
    v |= (uint32_t)((b->s[b->pos >> 3] >> (b->pos & 7)) & 1) << i;
  }
  return v;
}

/* canonical Huffman code: count of codes per length and symbols by code */
typedef struct {
  int16_t count[16];
  int16_t sym[320];
} dhuf_t;

static int dhuf_build(dhuf_t *h, const uint8_t *len, int n) {
  int16_t offs[16];
  memset(h->count, 0, sizeof(h->count));
  for (int i = 0; i < n; i++) h->count[len[i]]++;
  if (h->count[0] == n) return 0; /* no codes (allowed for distances) */
  int left = 1;
  for (int l = 1; l < 16; l++) {
    left <<= 1;
    left -= h->count[l];
    if (left < 0) return -1; /* over-subscribed */
  }
  offs[1] = 0;
  for (int l = 1; l < 15; l++) offs[l + 1] = (int16_t)(offs[l] + h->count[l]);
  for (int i = 0; i < n; i++)
    if (len[i]) h->sym[offs[len[i]]++] = (int16_t)i;
  return left; /* > 0: incomplete */
}

static int dhuf_decode(dbits_t *b, const dhuf_t *h) {
  int code = 0, first = 0, index = 0;
  for (int l = 1; l < 16; l++) {
    code |= (int)dbits(b, 1);
    if (b->err) return -1;
    const int count = h->count[l];
    if (code - count < first) return h->sym[index + (code - first)];
    index += count;
    first += count;
    first <<= 1;
    code <<= 1;
  }
  return -1;
}

static const uint16_t LBASE[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                   31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
static const uint8_t LEXT[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
static const uint16_t DBASE[30] = {1,   2,   3,   4,   5,   7,    9,    13,   17,   25,   33,   49,   65,    97,    129,
                                   193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
static const uint8_t DEXT[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

static int codes(dbits_t *b, const dhuf_t *lit, const dhuf_t *dist, uint8_t *dst, int64_t *op, int64_t cap) {
  for (;;) {
    int s = dhuf_decode(b, lit);
    if (s < 0) return -1;
    if (s < 256) {
      if (*op >= cap) return -1;
      dst[(*op)++] = (uint8_t)s;
    } else if (s == 256) {
      return 0;
    } else {
      s -= 257;
      if (s >= 29) return -1;
      const int64_t len = LBASE[s] + dbits(b, LEXT[s]);
      const int d = dhuf_decode(b, dist);
      if (d < 0 || d >= 30) return -1;
      const int64_t dd = DBASE[d] + dbits(b, DEXT[d]);
      if (b->err || dd > *op || len > cap - *op) return -1;
      for (int64_t k = 0; k < len; k++) dst[*op + k] = dst[*op + k - dd];
      *op += len;
    }
  }
}

int orc_zlib_decode(const uint8_t *src, int srclen, uint8_t *dst, int dstcap) {
  if (srclen < 2) return -1;
  const int cmf = src[0], flg = src[1];
  if ((cmf & 15) != 8 || ((cmf << 8) | flg) % 31 || (flg & 0x20)) return -2; /* deflate, check bits, no dictionary */
  dbits_t b = {src + 2, srclen - 2, 0, 0};
  int64_t op = 0;
  int last;
  do {
    last = (int)dbits(&b, 1);
    const int type = (int)dbits(&b, 2);
    if (b.err) return -1;
    if (type == 0) {
      b.pos = (b.pos + 7) & ~(int64_t)7;
      const int64_t q = b.pos >> 3;
      if (q + 4 > b.n) return -1;
      const int len = b.s[q] | (b.s[q + 1] << 8), nlen = b.s[q + 2] | (b.s[q + 3] << 8);
      if (len != (~nlen & 0xffff) || q + 4 + len > b.n || op + len > dstcap) return -1;
      memcpy(dst + op, b.s + q + 4, (size_t)len);
      op += len;
      b.pos = (q + 4 + len) * 8;
    } else if (type == 1 || type == 2) {
      uint8_t lens[320];
      dhuf_t lit, dist;
      int nlen = 288, ndist = 30;
      if (type == 1) {
        for (int i = 0; i < 144; i++) lens[i] = 8;
        for (int i = 144; i < 256; i++) lens[i] = 9;
        for (int i = 256; i < 280; i++) lens[i] = 7;
        for (int i = 280; i < 288; i++) lens[i] = 8;
        for (int i = 0; i < 32; i++) lens[288 + i] = 5; /* codes 30, 31 take code space, never occur */
        ndist = 32;
      } else {
        static const uint8_t ORD[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
        nlen = (int)dbits(&b, 5) + 257;
        ndist = (int)dbits(&b, 5) + 1;
        const int ncode = (int)dbits(&b, 4) + 4;
        if (b.err || nlen > 286 || ndist > 30) return -1;
        uint8_t cl[19] = {0};
        for (int i = 0; i < ncode; i++) cl[ORD[i]] = (uint8_t)dbits(&b, 3);
        dhuf_t lencode;
        if (dhuf_build(&lencode, cl, 19) != 0) return -1; /* must be complete */
        int idx = 0;
        while (idx < nlen + ndist) {
          int s = dhuf_decode(&b, &lencode);
          if (s < 0) return -1;
          if (s < 16) {
            lens[idx++] = (uint8_t)s;
          } else {
            uint8_t v = 0;
            int rep;
            if (s == 16) {
              if (idx == 0) return -1;
              v = lens[idx - 1];
              rep = 3 + (int)dbits(&b, 2);
            } else if (s == 17) {
              rep = 3 + (int)dbits(&b, 3);
            } else {
              rep = 11 + (int)dbits(&b, 7);
            }
            if (b.err || idx + rep > nlen + ndist) return -1;
            while (rep--) lens[idx++] = v;
          }
        }
        if (lens[256] == 0) return -1; /* no end-of-block code */
        /* the distance lengths follow the literal/length ones */
        memmove(lens + 288, lens + nlen, (size_t)ndist);
      }
      const int el = dhuf_build(&lit, lens, nlen);
      if (el < 0 || (el > 0 && nlen - lit.count[0] != 1)) return -1;
      const int ed = dhuf_build(&dist, lens + 288, ndist);
      if (ed < 0 || (ed > 0 && ndist - dist.count[0] != 1)) return -1;
      if (codes(&b, &lit, &dist, dst, &op, dstcap)) return -1;
    } else {
      return -1;
    }
  } while (!last);
  return (int)op;
}
