/*
 * lz4_block.c -- TEST INFRASTRUCTURE (oracle).  LZ4 raw-block decoder.
 *
 * The reference reaches LZ4 through c-blosc2 (blosc2_getitem_ctx at
 * src/linalg/iarray_gemm.c:259 and src/linalg/iarray_gemv.c:194,
 * caterva_get_slice_buffer at src/iarray_container.c:730), which calls
 * liblz4's LZ4_decompress_safe on every compressed stream.  liblz4 is a
 * third-party dependency not vendored in the reference (pinned here by the
 * system liblz4 1.9.3).  This file restates the published LZ4 Block Format
 * (lz4 doc/lz4_Block_format.md): a block is a series of sequences
 *   token | [literal-length bytes] | literals | offset(2, LE) | [match-length bytes]
 * where the high token nibble is the literal length (15 => continue with
 * bytes until one is != 255), the low nibble is match length - 4, and the
 * last sequence stops after its literals.  Matches copy byte-by-byte from
 * `offset` bytes back, so an offset smaller than the length repeats a pattern.
 *
 * Parity: pinned against liblz4-compressed buffers (tests/test_oracle_codec.py)
 * and through the fixture frames' SHA-256 known answers.
 */
#include <string.h>
#include "oracle.h"

int orc_lz4_decode(const uint8_t *src, int srclen, uint8_t *dst, int dstcap) {
  const uint8_t *ip = src;
  const uint8_t *const iend = src + srclen;
  uint8_t *op = dst;
  uint8_t *const oend = dst + dstcap;
  if (srclen <= 0 || dstcap < 0) return -1;
  for (;;) {
    if (ip >= iend) return -2;
    const unsigned token = *ip++;
    size_t lit = token >> 4;
    if (lit == 15) {
      unsigned b;
      do {
        if (ip >= iend) return -3;
        b = *ip++;
        lit += b;
      } while (b == 255);
    }
    if (lit > (size_t)(iend - ip) || lit > (size_t)(oend - op)) return -4;
    memcpy(op, ip, lit);
    op += lit;
    ip += lit;
    if (ip == iend) break; /* the last sequence carries literals only */
    if (iend - ip < 2) return -5;
    const size_t off = (size_t)ip[0] | ((size_t)ip[1] << 8);
    ip += 2;
    if (off == 0 || off > (size_t)(op - dst)) return -6;
    size_t ml = token & 15u;
    if (ml == 15) {
      unsigned b;
      do {
        if (ip >= iend) return -7;
        b = *ip++;
        ml += b;
      } while (b == 255);
    }
    ml += 4;
    if (ml > (size_t)(oend - op)) return -8;
    const uint8_t *m = op - off;
    for (size_t i = 0; i < ml; i++) op[i] = m[i];
    op += ml;
  }
  return (int)(op - dst);
}
