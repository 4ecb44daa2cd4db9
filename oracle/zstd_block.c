/*
 * zstd_block.c -- TEST INFRASTRUCTURE (oracle).  Zstandard frame decoder.
 *
 * Zstd is one of the codecs a Blosc2 chunk may carry (codec format 4 in the
 * chunk flags, bits 5-7) and one BTune may choose per chunk
 * (src/iarray.c:477-498, include/libiarray/iarray.h:177-187).  c-blosc2 is a
 * third-party dependency the reference does not vendor; it compresses every
 * stream of a block into one zstd frame (ZSTD_compressCCtx).  This restates
 * the published frame format (RFC 8878) and is pinned by decoding libzstd
 * 1.4.8's output (black box) and c-blosc 1.21.0's zstd chunks
 * (tests/test_oracle_codec.py):
 *   frame     magic 0xFD2FB528, frame header (descriptor, window, dictionary
 *             id, content size), blocks, optional 4-byte checksum (skipped);
 *             skippable frames (magic 0x184D2A5?) are skipped;
 *   block     3-byte header (last flag, type raw / RLE / compressed, size);
 *   literals  raw, RLE, or Huffman-coded in 1 or 4 streams (a tree
 *             description of FSE-coded or 4-bit weights, or the previous
 *             tree for "treeless" literals);
 *   sequences literal length, match length and offset codes, each coded by
 *             an FSE table (predefined, RLE, described, or the previous
 *             one), read from one backward bit stream; offsets 1-3 name the
 *             repeat offsets.
 * No dictionaries.  Returns the decoded size or a negative error.
 */
#include <string.h>

#include "oracle.h"

/* ---- backward bit stream: bit i of the stream is bit (i & 7) of byte i >> 3;
 * reading starts below the highest set bit of the last byte and takes
 * fields from the top down, each field's most significant bit first ---- */
typedef struct {
  const uint8_t *s;
  int64_t pos; /* bits left to read: the next field ends at bit pos */
} bbits_t;

static int bb_init(bbits_t *b, const uint8_t *s, int64_t n) {
  if (n <= 0 || s[n - 1] == 0) return -1;
  int hb = 7;
  while (!((s[n - 1] >> hb) & 1)) hb--;
  b->s = s;
  b->pos = (n - 1) * 8 + hb;
  return 0;
}

/* the next nb bits (nb <= 32); bits below the stream's start read as 0 */
static uint32_t bb_read(bbits_t *b, int nb) {
  uint32_t v = 0;
  for (int i = nb - 1; i >= 0; i--) {
    const int64_t p = b->pos - nb + i;
    const uint32_t bit = p >= 0 ? (b->s[p >> 3] >> (p & 7)) & 1u : 0u;
    v |= bit << i;
  }
  b->pos -= nb;
  return v;
}

static uint32_t bb_peek(const bbits_t *b, int nb) {
  bbits_t t = *b;
  return bb_read(&t, nb);
}

/* ---- forward bit stream (FSE table descriptions) ---- */
typedef struct {
  const uint8_t *s;
  int64_t n, pos; /* pos in bits */
} fbits_t;

static uint32_t fb_read(fbits_t *f, int nb) {
  uint32_t v = 0;
  for (int i = 0; i < nb; i++, f->pos++) {
    const uint32_t bit = (f->pos >> 3) < f->n ? (f->s[f->pos >> 3] >> (f->pos & 7)) & 1u : 0u;
    v |= bit << i;
  }
  return v;
}

static int highbit(uint32_t v) {
  int r = -1;
  while (v) {
    v >>= 1;
    r++;
  }
  return r;
}

/* ---- FSE decoding tables ---- */
#define FSE_MAX_LOG 9
typedef struct {
  uint8_t sym[1 << FSE_MAX_LOG];
  uint8_t nb[1 << FSE_MAX_LOG];
  uint16_t base[1 << FSE_MAX_LOG];
  int log;
} fse_t;

/* Build from normalized counts (-1 = "less than one") of nsym symbols. */
static int fse_build(fse_t *t, const int16_t *norm, int nsym, int log) {
  if (log > FSE_MAX_LOG || log < 1) return -1;
  const int size = 1 << log;
  int high = size - 1;
  uint16_t next[256];
  for (int s = 0; s < nsym; s++) {
    if (norm[s] == -1) {
      t->sym[high--] = (uint8_t)s;
      next[s] = 1;
    } else {
      next[s] = (uint16_t)norm[s];
    }
  }
  const int step = (size >> 1) + (size >> 3) + 3, mask = size - 1;
  int p = 0;
  for (int s = 0; s < nsym; s++)
    for (int i = 0; i < norm[s]; i++) {
      t->sym[p] = (uint8_t)s;
      do p = (p + step) & mask;
      while (p > high);
    }
  if (p != 0) return -1;
  for (int i = 0; i < size; i++) {
    const int s = t->sym[i];
    const uint32_t ns = next[s]++;
    const int nb = log - highbit(ns);
    t->nb[i] = (uint8_t)nb;
    t->base[i] = (uint16_t)((ns << nb) - (uint32_t)size);
  }
  t->log = log;
  return 0;
}

/* Read a table description (forward bits) at s; returns bytes used. */
static int fse_read(fse_t *t, const uint8_t *s, int64_t n, int maxlog, int maxsym) {
  fbits_t f = {s, n, 0};
  const int log = (int)fb_read(&f, 4) + 5;
  if (log > maxlog) return -1;
  int16_t norm[256];
  int remaining = (1 << log) + 1, sym = 0;
  while (remaining > 1 && sym <= maxsym) {
    const int nbits = highbit((uint32_t)remaining) + 1;
    const int maxv = (1 << nbits) - 1 - remaining; /* values below use nbits - 1 bits */
    int v = (int)fb_read(&f, nbits - 1);
    if (v >= maxv) {
      v += (int)fb_read(&f, 1) << (nbits - 1);
      if (v >= (1 << (nbits - 1))) v -= maxv;
    }
    const int prob = v - 1;
    norm[sym++] = (int16_t)prob;
    remaining -= prob < 0 ? -prob : prob;
    if (prob == 0) {
      int rep;
      do {
        rep = (int)fb_read(&f, 2);
        for (int i = 0; i < rep && sym <= maxsym; i++) norm[sym++] = 0;
      } while (rep == 3);
    }
  }
  if (remaining != 1 || sym > maxsym + 1) return -1;
  if (fse_build(t, norm, sym, log)) return -1;
  return (int)((f.pos + 7) >> 3);
}

static void fse_rle(fse_t *t, uint8_t sym) {
  t->log = 0;
  t->sym[0] = sym;
  t->nb[0] = 0;
  t->base[0] = 0;
}

/* predefined distributions (RFC 8878, 3.1.1.3.2.2) */
static const int16_t LL_DEF[36] = {4, 3, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 1, 1, 1, 2, 2,
                                   2, 2, 2, 2, 2, 2, 2, 3, 2, 1, 1, 1, 1, 1, -1, -1, -1, -1};
static const int16_t ML_DEF[53] = {1, 4, 3, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                   1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                   1, 1, 1, 1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1, -1, -1};
static const int16_t OF_DEF[29] = {1, 1, 1, 1, 1, 1, 2, 2, 2, 1, 1, 1, 1, 1, 1,
                                   1, 1, 1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1};
static const uint32_t LL_BASE[36] = {0,  1,  2,  3,  4,  5,  6,  7,  8,   9,   10,  11,   12,   13,   14,    15,    16,    18,
                                     20, 22, 24, 28, 32, 40, 48, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536};
static const uint8_t LL_BITS[36] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,  0,  0,  0,  1,  1,
                                    1, 1, 2, 2, 3, 3, 4, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16};
static const uint32_t ML_BASE[53] = {3,  4,  5,  6,  7,  8,  9,  10,  11,  12,  13,   14,   15,   16,   17,   18,    19,    20,
                                     21, 22, 23, 24, 25, 26, 27, 28,  29,  30,  31,   32,   33,   34,   35,   37,    39,    41,
                                     43, 47, 51, 59, 67, 83, 99, 131, 259, 515, 1027, 2051, 4099, 8195, 16387, 32771, 65539};
static const uint8_t ML_BITS[53] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,  0,  0,  0,  0,  0,  0,  0,  0,  0, 0,
                                    0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 3, 3, 4, 4, 5, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16};

/* ---- Huffman ---- */
#define HUF_MAX_BITS 11
typedef struct {
  uint8_t sym[1 << HUF_MAX_BITS];
  uint8_t nb[1 << HUF_MAX_BITS];
  int maxbits; /* 0: no table yet */
} huf_t;

static int huf_read(huf_t *h, const uint8_t *s, int64_t n) {
  if (n < 1) return -1;
  uint8_t w[256];
  int nw = 0, used;
  const int hb = s[0];
  if (hb >= 128) {
    nw = hb - 127;
    used = 1 + (nw + 1) / 2;
    if (used > n) return -1;
    for (int i = 0; i < nw; i++) w[i] = (uint8_t)((i & 1) ? s[1 + i / 2] & 15 : s[1 + i / 2] >> 4);
  } else {
    used = 1 + hb;
    if (used > n || hb < 1) return -1;
    fse_t t;
    const int d = fse_read(&t, s + 1, hb, 6, 255);
    if (d < 0 || d > hb) return -1;
    bbits_t b;
    if (bb_init(&b, s + 1 + d, hb - d)) return -1;
    /* two states, alternating, until the stream is exhausted */
    uint32_t s1 = bb_read(&b, t.log), s2 = bb_read(&b, t.log);
    for (;;) {
      if (nw >= 255) return -1;
      w[nw++] = t.sym[s1];
      s1 = t.base[s1] + bb_read(&b, t.nb[s1]);
      if (b.pos < 0) {
        w[nw++] = t.sym[s2];
        break;
      }
      if (nw >= 255) return -1;
      w[nw++] = t.sym[s2];
      s2 = t.base[s2] + bb_read(&b, t.nb[s2]);
      if (b.pos < 0) {
        w[nw++] = t.sym[s1];
        break;
      }
    }
  }
  /* the last weight completes the total to a power of two */
  uint32_t total = 0;
  for (int i = 0; i < nw; i++) {
    if (w[i] > HUF_MAX_BITS) return -1;
    if (w[i]) total += 1u << (w[i] - 1);
  }
  if (!total) return -1;
  const int maxbits = highbit(total) + 1;
  if (maxbits > HUF_MAX_BITS) return -1;
  const uint32_t rest = (1u << maxbits) - total;
  if (rest & (rest - 1)) return -1;
  w[nw++] = (uint8_t)(highbit(rest) + 1);
  /* entries: weights ascending, symbols ascending, 2^(w-1) entries each */
  int pos = 0;
  for (int wt = 1; wt <= maxbits; wt++)
    for (int sy = 0; sy < nw; sy++)
      if (w[sy] == wt)
        for (int k = 0; k < (1 << (wt - 1)); k++) {
          h->sym[pos] = (uint8_t)sy;
          h->nb[pos] = (uint8_t)(maxbits + 1 - wt);
          pos++;
        }
  if (pos != 1 << maxbits) return -1;
  h->maxbits = maxbits;
  return used;
}

static int huf_stream(const huf_t *h, const uint8_t *s, int64_t n, uint8_t *out, int64_t count) {
  bbits_t b;
  if (bb_init(&b, s, n)) return -1;
  for (int64_t i = 0; i < count; i++) {
    const uint32_t idx = bb_peek(&b, h->maxbits);
    out[i] = h->sym[idx];
    b.pos -= h->nb[idx];
  }
  return b.pos == 0 ? 0 : -1;
}

/* ---- frame decoder ---- */
typedef struct {
  huf_t huf;
  fse_t ll, of, ml;
  int have[3]; /* a table of LL / OF / ML was set in this frame */
  uint32_t rep[3];
} zstate_t;

static int64_t decode_block(zstate_t *z, const uint8_t *s, int64_t n, uint8_t *dst, int64_t op, int64_t cap,
                            uint8_t *lit) {
  /* literals section */
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
    if (n < 3) return -1;
    const uint64_t h = (uint64_t)s[0] | ((uint64_t)s[1] << 8) | ((uint64_t)s[2] << 16) |
                       (n > 3 ? (uint64_t)s[3] << 24 : 0) | (n > 4 ? (uint64_t)s[4] << 32 : 0);
    const int bits = sf < 2 ? 10 : sf == 2 ? 14 : 18;
    hdr = sf < 2 ? 3 : sf == 2 ? 4 : 5;
    nstreams = sf == 0 ? 1 : 4;
    regen = (int64_t)((h >> 4) & ((1u << bits) - 1));
    csize = (int64_t)((h >> (4 + bits)) & ((1u << bits) - 1));
    if (hdr > n) return -1;
  }
  if (regen > 131072) return -1;
  int64_t p = hdr;
  if (ltype == 0) {
    if (p + regen > n) return -1;
    memcpy(lit, s + p, (size_t)regen);
    p += regen;
  } else if (ltype == 1) {
    if (p + 1 > n) return -1;
    memset(lit, s[p], (size_t)regen);
    p += 1;
  } else {
    if (p + csize > n) return -1;
    int64_t q = p;
    if (ltype == 2) {
      const int u = huf_read(&z->huf, s + q, csize);
      if (u < 0) return -1;
      q += u;
    } else if (!z->huf.maxbits) {
      return -1;
    }
    const int64_t left = p + csize - q;
    if (nstreams == 1) {
      if (huf_stream(&z->huf, s + q, left, lit, regen)) return -1;
    } else {
      if (left < 6) return -1;
      const int64_t l1 = s[q] | (s[q + 1] << 8), l2 = s[q + 2] | (s[q + 3] << 8), l3 = s[q + 4] | (s[q + 5] << 8);
      const int64_t l4 = left - 6 - l1 - l2 - l3;
      if (l4 < 1) return -1;
      const int64_t per = (regen + 3) / 4;
      if (3 * per > regen) return -1;
      const uint8_t *b = s + q + 6;
      if (huf_stream(&z->huf, b, l1, lit, per) || huf_stream(&z->huf, b + l1, l2, lit + per, per) ||
          huf_stream(&z->huf, b + l1 + l2, l3, lit + 2 * per, per) ||
          huf_stream(&z->huf, b + l1 + l2 + l3, l4, lit + 3 * per, regen - 3 * per))
        return -1;
    }
    p += csize;
  }
  /* sequences section */
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
    fse_t *tabs[3] = {&z->ll, &z->of, &z->ml};
    const int16_t *defs[3] = {LL_DEF, OF_DEF, ML_DEF};
    const int dn[3] = {36, 29, 53}, dlog[3] = {6, 5, 6}, maxlog[3] = {9, 8, 9}, maxsym[3] = {35, 31, 52};
    for (int k = 0; k < 3; k++) {
      const int mode = (modes >> (6 - 2 * k)) & 3;
      if (mode == 0) {
        if (fse_build(tabs[k], defs[k], dn[k], dlog[k])) return -1;
      } else if (mode == 1) {
        if (p >= n || s[p] > maxsym[k]) return -1;
        fse_rle(tabs[k], s[p++]);
      } else if (mode == 2) {
        const int u = fse_read(tabs[k], s + p, n - p, maxlog[k], maxsym[k]);
        if (u < 0) return -1;
        p += u;
      } else if (!z->have[k]) { /* repeat mode: the previous block's table */
        return -1;
      }
      z->have[k] = 1;
    }
    bbits_t b;
    if (bb_init(&b, s + p, n - p)) return -1;
    uint32_t sll = bb_read(&b, z->ll.log), sof = bb_read(&b, z->of.log), sml = bb_read(&b, z->ml.log);
    for (int64_t i = 0; i < nseq; i++) {
      const int llc = z->ll.sym[sll], ofc = z->of.sym[sof], mlc = z->ml.sym[sml];
      if (llc > 35 || mlc > 52 || ofc > 31) return -1;
      const uint32_t ofv = (1u << ofc) + bb_read(&b, ofc);
      const int64_t ml = ML_BASE[mlc] + bb_read(&b, ML_BITS[mlc]);
      const int64_t ll = LL_BASE[llc] + bb_read(&b, LL_BITS[llc]);
      uint32_t off;
      if (ofv > 3) {
        off = ofv - 3;
        z->rep[2] = z->rep[1];
        z->rep[1] = z->rep[0];
        z->rep[0] = off;
      } else {
        const int idx = (int)ofv - (ll == 0 ? 0 : 1); /* 0..3 */
        if (idx == 0) {
          off = z->rep[0];
        } else {
          off = idx == 3 ? z->rep[0] - 1 : z->rep[idx];
          if (idx != 1) z->rep[2] = z->rep[1];
          z->rep[1] = z->rep[0];
          z->rep[0] = off;
        }
      }
      if (i + 1 < nseq) {
        sll = z->ll.base[sll] + bb_read(&b, z->ll.nb[sll]);
        sml = z->ml.base[sml] + bb_read(&b, z->ml.nb[sml]);
        sof = z->of.base[sof] + bb_read(&b, z->of.nb[sof]);
      }
      if (lp + ll > regen || op + ll + ml > cap || off == 0 || off > op + ll) return -1;
      memcpy(dst + op, lit + lp, (size_t)ll);
      op += ll;
      lp += ll;
      for (int64_t k = 0; k < ml; k++) dst[op + k] = dst[op + k - off];
      op += ml;
    }
    if (b.pos != 0) return -1;
  } else if (p != n) {
    return -1;
  }
  if (op + (regen - lp) > cap) return -1;
  memcpy(dst + op, lit + lp, (size_t)(regen - lp));
  return op + (regen - lp);
}

/* Static scratch for the literals of one block (at most 128 KiB) and the
 * decoder state: the oracle is single-threaded per call. */
int orc_zstd_decode(const uint8_t *src, int srclen, uint8_t *dst, int dstcap) {
  static __thread uint8_t lit[131072];
  static __thread zstate_t z;
  int64_t ip = 0, op = 0;
  while (ip < srclen) {
    if (srclen - ip < 4) return -1;
    const uint32_t magic = (uint32_t)src[ip] | ((uint32_t)src[ip + 1] << 8) | ((uint32_t)src[ip + 2] << 16) |
                           ((uint32_t)src[ip + 3] << 24);
    ip += 4;
    if ((magic & 0xFFFFFFF0u) == 0x184D2A50u) { /* skippable frame */
      if (srclen - ip < 4) return -1;
      const uint32_t len = (uint32_t)src[ip] | ((uint32_t)src[ip + 1] << 8) | ((uint32_t)src[ip + 2] << 16) |
                           ((uint32_t)src[ip + 3] << 24);
      ip += 4;
      if (len > (uint32_t)(srclen - ip)) return -1;
      ip += len;
      continue;
    }
    if (magic != 0xFD2FB528u) return -2;
    if (ip >= srclen) return -1;
    const int fhd = src[ip++];
    const int fcs_flag = fhd >> 6, single = (fhd >> 5) & 1, checksum = (fhd >> 2) & 1, did = fhd & 3;
    if (fhd & 0x08) return -1; /* reserved bit */
    if (did) return -3;         /* dictionaries are not used by c-blosc2 streams here */
    if (!single) ip += 1;       /* window descriptor */
    const int fcs_len = fcs_flag == 0 ? (single ? 1 : 0) : fcs_flag == 1 ? 2 : fcs_flag == 2 ? 4 : 8;
    ip += fcs_len;
    if (ip > srclen) return -1;
    memset(&z, 0, sizeof(z));
    z.rep[0] = 1;
    z.rep[1] = 4;
    z.rep[2] = 8;
    for (;;) {
      if (srclen - ip < 3) return -1;
      const uint32_t bh = (uint32_t)src[ip] | ((uint32_t)src[ip + 1] << 8) | ((uint32_t)src[ip + 2] << 16);
      ip += 3;
      const int last = bh & 1, type = (bh >> 1) & 3;
      const int64_t bsize = bh >> 3;
      if (type == 0) {
        if (bsize > srclen - ip || op + bsize > dstcap) return -1;
        memcpy(dst + op, src + ip, (size_t)bsize);
        ip += bsize;
        op += bsize;
      } else if (type == 1) {
        if (ip >= srclen || op + bsize > dstcap) return -1;
        memset(dst + op, src[ip], (size_t)bsize);
        ip += 1;
        op += bsize;
      } else if (type == 2) {
        if (bsize > srclen - ip || bsize > 131072) return -1;
        const int64_t r = decode_block(&z, src + ip, bsize, dst, op, dstcap, lit);
        if (r < 0) return -1;
        op = r;
        ip += bsize;
      } else {
        return -1;
      }
      if (last) break;
    }
    if (checksum) ip += 4;
    if (ip > srclen) return -1;
  }
  return (int)op;
}
