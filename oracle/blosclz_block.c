/*
 * blosclz_block.c -- TEST INFRASTRUCTURE (oracle).  BloscLZ stream decoder.
 *
 * BloscLZ is the default codec of c-blosc (codec 0 in the chunk flags, bits
 * 5-7) and one of the codecs BTune may pick per chunk (src/iarray.c:477-498,
 * include/libiarray/iarray.h:177-187).  c-blosc2 / c-blosc are third-party
 * dependencies the reference does not vendor; this restates the stream
 * format (a FastLZ-style LZ77 token stream) as pinned by decoding the output
 * of c-blosc 1.21.0's blosc_compress with compressor "blosclz" (black box,
 * tests/test_oracle_codec.py):
 *   the first byte's low 5 bits are its control value (the top 3 bits mark
 *   the format); every later token starts with a control byte c:
 *     c < 32   literal run of c + 1 bytes follows;
 *     c >= 32  match: length L = (c >> 5) - 1, plus further bytes added while
 *              each is 255 (the last one is < 255) when L == 6; then one
 *              distance byte d; the match copies L + 3 bytes from
 *              (c & 31) * 256 + d + 1 bytes back, or, when d == 255 and
 *              (c & 31) == 31, from 8192 + the next two bytes (big-endian)
 *              bytes back.
 *   The stream ends with its last token.  Overlapping matches repeat the
 *   pattern (byte by byte copy).
 */
#include "oracle.h"

int orc_blosclz_decode(const uint8_t *src, int srclen, uint8_t *dst, int dstcap) {
  if (srclen <= 0 || dstcap < 0) return -1;
  const uint8_t *ip = src, *const iend = src + srclen;
  uint8_t *op = dst, *const oend = dst + dstcap;
  uint32_t ctrl = *ip++ & 31u;
  for (;;) {
    if (ctrl >= 32) {
      int64_t len = (int64_t)(ctrl >> 5) - 1;
      int64_t dist = (int64_t)(ctrl & 31) << 8;
      if (len == 6) {
        uint32_t code;
        do {
          if (ip >= iend) return -2;
          code = *ip++;
          len += code;
        } while (code == 255);
      }
      if (ip >= iend) return -3;
      const uint32_t code = *ip++;
      len += 3;
      if (code == 255 && dist == (31 << 8)) {
        if (iend - ip < 2) return -4;
        dist = ((int64_t)ip[0] << 8 | ip[1]) + 8192;
        ip += 2;
      } else {
        dist += code + 1;
      }
      if (dist > op - dst) return -5;
      if (len > oend - op) return -6;
      const uint8_t *ref = op - dist;
      for (int64_t k = 0; k < len; k++) op[k] = ref[k];
      op += len;
    } else {
      const int64_t len = (int64_t)ctrl + 1;
      if (len > iend - ip) return -7;
      if (len > oend - op) return -8;
      for (int64_t k = 0; k < len; k++) op[k] = ip[k];
      op += len;
      ip += len;
    }
    if (ip >= iend) break;
    ctrl = *ip++;
  }
  return (int)(op - dst);
}
