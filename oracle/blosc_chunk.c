/*
 * blosc_chunk.c -- TEST INFRASTRUCTURE (oracle).  Blosc chunk decode stage.
 *
 * c-blosc2 is an un-vendored dependency of the reference (the empty
 * contribs/caterva submodule, .gitmodules:1-3; chunk format version 5 in the
 * reference's own tests/data/<name>.iarray frames).  Its decode path is what the
 * reference calls on the hot path (blosc2_getitem_ctx,
 * src/linalg/iarray_gemm.c:259-274; src/linalg/iarray_gemv.c:194-209).  This
 * file restates c-blosc2's published chunk format and blosc_d() block decode:
 *
 *  header (little endian): version, versionlz, flags, typesize, nbytes,
 *    blocksize, cbytes; when (flags & 0x5) == 0x5 the 32-byte extended header
 *    follows with filters[6] @16, udcompcode @22, compcode_meta @23,
 *    filters_meta[6] @24 and blosc2 flags @31 (special value in bits 4..6);
 *    otherwise the 16-byte blosc1 header, whose filter is flags bit 0
 *    (byte shuffle) or bit 2 (bitshuffle).
 *  flags bit 1: memcpyed chunk (raw bytes follow the header);
 *  flags bit 4: blocks are not split into typesize streams;
 *  flags bits 5..7: codec format (1 = LZ4/LZ4HC).
 *  int32 bstarts[nblocks] follow the header; block j holds nstreams =
 *  (split && not the leftover block) ? typesize : 1 streams, each an int32
 *  csize then data: 0 = zeros, < 0 = a run of byte (-csize) (a token byte with
 *  bit 0 set follows), == neblock = raw, else one LZ4 raw block of neblock
 *  bytes.  The concatenated streams are then un-filtered, filters applied in
 *  reverse slot order (the reference puts SHUFFLE/BITSHUFFLE in slot 5,
 *  src/iarray.c:414-419).
 *
 * The reference's own chunk_is_zeros/block_is_zeros (src/linalg/iarray_gemm.c:
 * 17-77) are restated here too.
 */
#include <dlfcn.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static int32_t rd32(const uint8_t *p) {
  return (int32_t)((uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24));
}

/* ---- LZ4 backend: restated decoder or liblz4 (black box) ---------------- */
typedef int (*lz4_dec_fn)(const char *, char *, int, int);
static lz4_dec_fn g_liblz4_dec = NULL;
static int g_lz4_backend = 0;

int orc_set_lz4(int which) {
  if (which == 1 && !g_liblz4_dec) {
    const char *names[] = {"liblz4.so.1", "/opt/conda/lib/liblz4.so.1",
                           "/usr/lib/x86_64-linux-gnu/liblz4.so.1", NULL};
    for (int i = 0; names[i] && !g_liblz4_dec; i++) {
      void *h = dlopen(names[i], RTLD_NOW | RTLD_LOCAL);
      if (h) g_liblz4_dec = (lz4_dec_fn)dlsym(h, "LZ4_decompress_safe");
    }
  }
  g_lz4_backend = (which == 1 && g_liblz4_dec) ? 1 : 0;
  return g_lz4_backend;
}

static int lz4_dec(const uint8_t *src, int srclen, uint8_t *dst, int dstcap) {
  if (g_lz4_backend == 1) return g_liblz4_dec((const char *)src, (char *)dst, srclen, dstcap);
  return orc_lz4_decode(src, srclen, dst, dstcap);
}

/* ---- header --------------------------------------------------------------- */
int orc_chunk_info(const uint8_t *c, int32_t csize, orc_chunk_info_t *h) {
  memset(h, 0, sizeof(*h));
  if (csize < 16) return -1;
  h->version = c[0];
  h->versionlz = c[1];
  h->flags = c[2];
  h->typesize = c[3];
  h->nbytes = rd32(c + 4);
  h->blocksize = rd32(c + 8);
  h->cbytes = rd32(c + 12);
  h->extended = (h->flags & 0x5) == 0x5;
  h->header_len = h->extended ? 32 : 16;
  if (csize < h->header_len || h->cbytes > csize || h->nbytes < 0) return -1;
  h->memcpyed = (h->flags & 0x2) != 0;
  h->dont_split = (h->flags & 0x10) != 0;
  h->compformat = h->flags >> 5;
  if (h->extended) {
    memcpy(h->filters, c + 16, 6);
    memcpy(h->filters_meta, c + 24, 6);
    h->blosc2_flags = c[31];
    h->special = (h->blosc2_flags >> 4) & 0x7;
  } else {
    /* blosc1: the filter lives in the flags byte; map it onto slot 5 */
    if (h->flags & 0x1) h->filters[5] = 1;
    else if (h->flags & 0x4) h->filters[5] = 2;
  }
  if (h->typesize < 1) return -1;
  if (h->nbytes == 0) { h->nblocks = 0; return 0; }
  if (h->blocksize <= 0) return -1;
  h->nblocks = h->nbytes / h->blocksize + (h->nbytes % h->blocksize != 0);
  h->leftover = h->nbytes % h->blocksize;
  return 0;
}

/* ---- filters --------------------------------------------------------------- */
void orc_shuffle(int ts, int32_t bsize, const uint8_t *src, uint8_t *dst) {
  const int32_t n = bsize / ts;
  for (int j = 0; j < ts; j++)
    for (int32_t i = 0; i < n; i++) dst[(int64_t)j * n + i] = src[(int64_t)i * ts + j];
  memcpy(dst + (int64_t)n * ts, src + (int64_t)n * ts, (size_t)(bsize - n * ts));
}

void orc_unshuffle(int ts, int32_t bsize, const uint8_t *src, uint8_t *dst) {
  const int32_t n = bsize / ts;
  for (int j = 0; j < ts; j++)
    for (int32_t i = 0; i < n; i++) dst[(int64_t)i * ts + j] = src[(int64_t)j * n + i];
  memcpy(dst + (int64_t)n * ts, src + (int64_t)n * ts, (size_t)(bsize - n * ts));
}

/* Bit layout (bitshuffle's bshuf_trans_bit_elem): for n elements (n % 8 == 0)
 * of ts bytes, output row r = 8*j + k (byte j, bit k) holds n/8 bytes; bit i
 * of byte m in that row is bit k of byte j of element 8*m + i. */
static int bit_elems(int ts, int32_t bsize, int format_version) {
  int32_t n = bsize / ts;
  if (format_version <= 2) return (n % 8 == 0) ? n : -1; /* -1: whole-block copy */
  return n - n % 8;
}

void orc_bitshuffle(int ts, int32_t bsize, const uint8_t *src, uint8_t *dst, int fv) {
  int32_t n = bit_elems(ts, bsize, fv);
  if (n < 0) { memcpy(dst, src, (size_t)bsize); return; }
  const int32_t rowlen = n / 8;
  memset(dst, 0, (size_t)n * ts);
  for (int32_t e = 0; e < n; e++)
    for (int j = 0; j < ts; j++) {
      const uint8_t v = src[(int64_t)e * ts + j];
      for (int k = 0; k < 8; k++)
        if ((v >> k) & 1) dst[(int64_t)(8 * j + k) * rowlen + e / 8] |= (uint8_t)(1u << (e % 8));
    }
  memcpy(dst + (int64_t)n * ts, src + (int64_t)n * ts, (size_t)(bsize - n * ts));
}

/* 8x8 bit-matrix transpose (row r = byte r): bit 8r+c -> bit 8c+r. */
static uint64_t t8(uint64_t x) {
  uint64_t t;
  t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
  x = x ^ t ^ (t << 7);
  t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
  x = x ^ t ^ (t << 14);
  t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
  x = x ^ t ^ (t << 28);
  return x;
}

void orc_bitunshuffle(int ts, int32_t bsize, const uint8_t *src, uint8_t *dst, int fv) {
  int32_t n = bit_elems(ts, bsize, fv);
  if (n < 0) { memcpy(dst, src, (size_t)bsize); return; }
  const int32_t rowlen = n / 8;
  /* group g = elements 8g..8g+7: gathering byte g of the 8 bit planes of byte
   * j gives an 8x8 bit matrix whose transpose holds byte j of each element */
  for (int32_t g = 0; g < rowlen; g++)
    for (int j = 0; j < ts; j++) {
      uint64_t x = 0;
      for (int k = 0; k < 8; k++) x |= (uint64_t)src[(int64_t)(8 * j + k) * rowlen + g] << (8 * k);
      const uint64_t y = t8(x);
      for (int i = 0; i < 8; i++) dst[(int64_t)(8 * g + i) * ts + j] = (uint8_t)(y >> (8 * i));
    }
  memcpy(dst + (int64_t)n * ts, src + (int64_t)n * ts, (size_t)(bsize - n * ts));
}

/* ---- special-value chunks ---------------------------------------------------- */
static void fill_special(const orc_chunk_info_t *h, const uint8_t *c, uint8_t *dst, int64_t nbytes) {
  if (h->special == 2) { /* NaN */
    if (h->typesize == 4) {
      float v = NAN;
      for (int64_t i = 0; i + 4 <= nbytes; i += 4) memcpy(dst + i, &v, 4);
    } else {
      double v = NAN;
      for (int64_t i = 0; i + 8 <= nbytes; i += 8) memcpy(dst + i, &v, 8);
    }
  } else if (h->special == 3) { /* value: typesize bytes after the header */
    for (int64_t i = 0; i + h->typesize <= nbytes; i += h->typesize) memcpy(dst + i, c + h->header_len, h->typesize);
  } else { /* zero or uninitialised */
    memset(dst, 0, (size_t)nbytes);
  }
}

/* ---- one block ------------------------------------------------------------------ */
int orc_decode_block(const uint8_t *c, int32_t csize, int nblock, uint8_t *dst) {
  orc_chunk_info_t h;
  if (orc_chunk_info(c, csize, &h) < 0) return -1;
  if (nblock < 0 || nblock >= h.nblocks) return -2;
  const int leftoverblock = (nblock == h.nblocks - 1) && h.leftover > 0;
  const int32_t bsize = leftoverblock ? h.leftover : h.blocksize;
  if (h.special) { fill_special(&h, c, dst, bsize); return bsize; }
  if (h.memcpyed) {
    if ((int64_t)h.header_len + (int64_t)nblock * h.blocksize + bsize > csize) return -3;
    memcpy(dst, c + h.header_len + (int64_t)nblock * h.blocksize, (size_t)bsize);
    return bsize;
  }
  if (h.compformat != 0 && h.compformat != 1 && h.compformat != 3 && h.compformat != 4)
    return -4; /* BloscLZ, LZ4, zlib and Zstd streams */
  if (h.extended && (h.blosc2_flags & 0x1)) return -5; /* dictionaries unsupported */
  const int32_t bstart = rd32(c + h.header_len + 4 * nblock);
  if (bstart < h.header_len || bstart >= csize) return -6;
  const int nstreams = (!h.dont_split && !leftoverblock) ? h.typesize : 1;
  const int32_t neblock = bsize / nstreams;
  uint8_t *tmp = (uint8_t *)malloc((size_t)bsize);
  if (!tmp) return -7;
  const uint8_t *p = c + bstart;
  const uint8_t *end = c + csize;
  int rc = bsize;
  for (int s = 0; s < nstreams && rc > 0; s++) {
    uint8_t *out = tmp + (int64_t)s * neblock;
    if (end - p < 4) { rc = -8; break; }
    const int32_t cs = rd32(p);
    p += 4;
    if (cs == 0) {
      memset(out, 0, (size_t)neblock);
    } else if (cs < 0) {
      if (end - p < 1 || !(p[0] & 0x1) || cs < -255) { rc = -9; break; }
      p += 1;
      memset(out, (uint8_t)(-cs), (size_t)neblock);
    } else if (cs == neblock) {
      if (end - p < cs) { rc = -10; break; }
      memcpy(out, p, (size_t)neblock);
      p += cs;
    } else {
      if (end - p < cs) { rc = -11; break; }
      const int got = h.compformat == 0   ? orc_blosclz_decode(p, cs, out, neblock)
                      : h.compformat == 4 ? orc_zstd_decode(p, cs, out, neblock)
                      : h.compformat == 3 ? orc_zlib_decode(p, cs, out, neblock)
                                          : lz4_dec(p, cs, out, neblock);
      if (got != neblock) { rc = -12; break; }
      p += cs;
    }
  }
  if (rc > 0) {
    /* un-filter in reverse slot order; only (bit)shuffle is on this path */
    uint8_t *cur = tmp, *alt = dst;
    int applied = 0;
    for (int i = 5; i >= 0 && rc > 0; i--) {
      const int f = h.filters[i];
      if (f == 0) continue;
      if (f == 1) {
        for (int r = 0; r <= (h.extended ? h.filters_meta[i] : 0); r++) {
          orc_unshuffle(h.typesize, bsize, cur, alt);
          uint8_t *t = cur; cur = alt; alt = t; applied++;
        }
      } else if (f == 2) {
        orc_bitunshuffle(h.typesize, bsize, cur, alt, h.version);
        uint8_t *t = cur; cur = alt; alt = t; applied++;
      } else if (f == 4) {
        /* TRUNC_PREC is a no-op on decompression */
      } else {
        rc = -13; /* DELTA and user filters are not on this path */
      }
    }
    if (rc > 0 && cur != dst) memcpy(dst, cur, (size_t)bsize);
  }
  free(tmp);
  return rc;
}

int orc_decompress_chunk(const uint8_t *c, int32_t csize, uint8_t *dst, int64_t dstcap) {
  orc_chunk_info_t h;
  if (orc_chunk_info(c, csize, &h) < 0) return -1;
  if (h.nbytes > dstcap) return -2;
  if (h.special) { fill_special(&h, c, dst, h.nbytes); return h.nbytes; }
  if (h.memcpyed) {
    if (h.header_len + h.nbytes > csize) return -3;
    memcpy(dst, c + h.header_len, (size_t)h.nbytes);
    return h.nbytes;
  }
  for (int b = 0; b < h.nblocks; b++) {
    int r = orc_decode_block(c, csize, b, dst + (int64_t)b * h.blocksize);
    if (r < 0) return r;
  }
  return h.nbytes;
}

int orc_getitem(const uint8_t *c, int32_t csize, int start, int nitems, uint8_t *dst) {
  orc_chunk_info_t h;
  if (orc_chunk_info(c, csize, &h) < 0) return -1;
  const int ts = h.typesize;
  const int64_t b0 = (int64_t)start * ts, b1 = (int64_t)(start + nitems) * ts;
  if (start < 0 || nitems < 0 || b1 > h.nbytes) return -2;
  if (nitems == 0) return 0;
  uint8_t *blk = (uint8_t *)malloc((size_t)h.blocksize);
  if (!blk) return -3;
  int rc = nitems * ts;
  for (int b = (int)(b0 / h.blocksize); b < h.nblocks && (int64_t)b * h.blocksize < b1; b++) {
    int r = orc_decode_block(c, csize, b, blk);
    if (r < 0) { rc = r; break; }
    const int64_t lo = (int64_t)b * h.blocksize, hi = lo + r;
    const int64_t s = b0 > lo ? b0 : lo, e = b1 < hi ? b1 : hi;
    memcpy(dst + (s - b0), blk + (s - lo), (size_t)(e - s));
  }
  free(blk);
  return rc;
}

int orc_chunk_is_zeros(const uint8_t *c) {
  return ((c[31] >> 4) & 0x7) == 1;
}

int orc_block_is_zeros(const uint8_t *c, int32_t csize, int nblock) {
  orc_chunk_info_t h;
  if (orc_chunk_info(c, csize, &h) < 0) return 0;
  if (h.memcpyed || h.special) return 0;
  if (h.extended && (h.blosc2_flags & 0x08)) return 0; /* lazy chunks are not built here */
  const int leftoverblock = (nblock == h.nblocks - 1) && h.leftover > 0;
  const int nstreams = (!h.dont_split && !leftoverblock) ? h.typesize : 1;
  const uint8_t *p = c + rd32(c + h.header_len + 4 * nblock);
  for (int s = 0; s < nstreams; s++) {
    const int32_t cs = rd32(p);
    if (cs != 0) return 0;
    p += 4;
  }
  return 1;
}
