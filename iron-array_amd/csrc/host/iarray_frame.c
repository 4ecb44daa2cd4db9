/*
 * iarray_frame.c -- containers to and from contiguous Blosc2 frames (b2frame),
 * the on-disk form the reference reads and writes:
 *   iarray_container_load / _open (src/iarray_container.c:79-218: caterva_open
 *     on the file, the "iarray" metalayer for the dtype, the "caterva"
 *     metalayer for the shape, chunk shape and block shape),
 *   iarray_from_cframe / iarray_to_cframe (src/iarray_container.c:228-300),
 *   iarray_container_save (src/iarray_container.c:50-76: a copy of the
 *     super-chunk into a contiguous frame file),
 *   iarray_container_remove (src/iarray_container.c:219-225).
 *
 * The frame is read as c-blosc2 lays it out: a msgpack header (a 14-element
 * array: magic "b2frame\0", header size, frame size, flags, nbytes, cbytes,
 * typesize, blocksize, chunksize, two thread counts, a flag, the coding
 * parameters, then [index size, {metalayer name: header offset}, [contents]]),
 * the data chunks from the end of the header on, then the offsets chunk (one
 * int64 per chunk, relative to the end of the header) at header size +
 * cbytes, then a trailer that is not needed here.  The caterva metalayer is
 * [version, ndim, shape[ndim], chunkshape[ndim], blockshape[ndim]], the
 * iarray metalayer [version, dtype, flags] (src/iarray_private.h:364-419).
 *
 * The chunks are copied into an in-memory container; the GPU decodes them as
 * it does every other container.  Not read here (an error says so): an
 * offsets chunk that is compressed rather than stored, and special-value
 * chunks (negative offsets).
 */
#include <dirent.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "iarray_private.h"

/* ---- a minimal msgpack reader over the types c-blosc2's frame header uses ---- */
typedef struct {
  const uint8_t *p, *end;
  int bad;
} mp_t;

static uint64_t be(const uint8_t *p, int n) {
  uint64_t v = 0;
  for (int i = 0; i < n; i++) v = (v << 8) | p[i];
  return v;
}

static const uint8_t *mp_take(mp_t *m, size_t n) {
  if (m->bad || (size_t)(m->end - m->p) < n) {
    m->bad = 1;
    return NULL;
  }
  const uint8_t *q = m->p;
  m->p += n;
  return q;
}

static int64_t mp_int(mp_t *m) {
  const uint8_t *t = mp_take(m, 1);
  if (!t) return 0;
  const uint8_t c = *t;
  if (c <= 0x7f) return c;
  if (c >= 0xe0) return (int8_t)c;
  static const int width[8] = {1, 2, 4, 8, 1, 2, 4, 8}; /* cc..cf unsigned, d0..d3 signed */
  if (c < 0xcc || c > 0xd3) {
    m->bad = 1;
    return 0;
  }
  const int k = c - 0xcc, n = width[k];
  const uint8_t *q = mp_take(m, (size_t)n);
  if (!q) return 0;
  const uint64_t u = be(q, n);
  if (k < 4) return (int64_t)u;
  if (n == 1) return (int8_t)u;
  if (n == 2) return (int16_t)u;
  if (n == 4) return (int32_t)u;
  return (int64_t)u;
}

/* element count of an array (want_map = 0) or a map (1) */
static int64_t mp_len(mp_t *m, int want_map) {
  const uint8_t *t = mp_take(m, 1);
  if (!t) return 0;
  const uint8_t c = *t, fix = want_map ? 0x80 : 0x90, c16 = want_map ? 0xde : 0xdc;
  if ((c & 0xf0) == fix) return c & 0x0f;
  if (c == c16 || c == c16 + 1) {
    const int n = c == c16 ? 2 : 4;
    const uint8_t *q = mp_take(m, (size_t)n);
    return q ? (int64_t)be(q, n) : 0;
  }
  m->bad = 1;
  return 0;
}

/* a str (want_bin = 0) or bin (1) element: its bytes */
static const uint8_t *mp_bytes(mp_t *m, int want_bin, int64_t *len) {
  const uint8_t *t = mp_take(m, 1);
  *len = 0;
  if (!t) return NULL;
  const uint8_t c = *t;
  int n = -1;
  if (!want_bin && (c & 0xe0) == 0xa0) {
    *len = c & 0x1f;
  } else if (c == (want_bin ? 0xc4 : 0xd9)) {
    n = 1;
  } else if (c == (want_bin ? 0xc5 : 0xda)) {
    n = 2;
  } else if (c == (want_bin ? 0xc6 : 0xdb)) {
    n = 4;
  } else {
    m->bad = 1;
    return NULL;
  }
  if (n > 0) {
    const uint8_t *q = mp_take(m, (size_t)n);
    if (!q) return NULL;
    *len = (int64_t)be(q, n);
  }
  return mp_take(m, (size_t)*len);
}

/* skips one scalar element of the kinds the header holds between chunksize
 * and the metalayers (ints, booleans, nil, fixext) */
static void mp_skip(mp_t *m) {
  if (m->bad || m->p >= m->end) {
    m->bad = 1;
    return;
  }
  const uint8_t c = *m->p;
  if (c == 0xc0 || c == 0xc2 || c == 0xc3) {
    m->p++;
  } else if (c >= 0xd4 && c <= 0xd8) { /* fixext 1, 2, 4, 8, 16: type byte + data */
    static const int data[5] = {1, 2, 4, 8, 16};
    (void)mp_take(m, 2 + (size_t)data[c - 0xd4]);
  } else {
    (void)mp_int(m);
  }
}

/* ---- frame ------------------------------------------------------------------- */
/* a 14-element array whose first element is the 8-byte string "b2frame\0" */
static const uint8_t frame_magic[10] = {0x9e, 0xa8, 'b', '2', 'f', 'r', 'a', 'm', 'e', 0};

static ina_rc_t frame_error(const char *what) {
  return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "not a readable Blosc2 frame: %s", what);
}

static ina_rc_t read_meta(const uint8_t *d, int64_t hsize, int64_t off, const uint8_t **content, int64_t *len) {
  if (off <= 0 || off >= hsize) return frame_error("metalayer offset");
  mp_t m = {d + off, d + hsize, 0};
  *content = mp_bytes(&m, 1, len);
  return m.bad ? frame_error("metalayer") : INA_SUCCESS;
}

/* the header fields between chunksize and the metalayers, kept for a later save
 * when they have the form c-blosc2 writes (two int16 thread counts, a boolean,
 * a fixext16 of coding parameters) */
static void read_params(mp_t *m, iax_frame_params_t *fp) {
  memset(fp, 0, sizeof(*fp));
  const uint8_t *p0 = m->p;
  int form = 1;
  for (int i = 0; i < 2; i++) {
    form &= m->p < m->end && *m->p == 0xd1;
    fp->nthreads[i] = mp_int(m);
  }
  form &= m->p < m->end && (*m->p == 0xc2 || *m->p == 0xc3);
  if (m->p < m->end) fp->vlmeta = *m->p;
  mp_skip(m);
  form &= m->p + 18 <= m->end && *m->p == 0xd8;
  if (form) {
    fp->coding_type = m->p[1];
    memcpy(fp->coding, m->p + 2, 16);
  }
  mp_skip(m);
  fp->valid = form && !m->bad && m->p - p0 == 3 + 3 + 1 + 18;
}

/* Reads chunk file `id` of a sparse frame directory (%08X.chunk). */
static uint8_t *read_chunk_file(const char *dir, int64_t id, int32_t *csize) {
  char path[4096];
  if (snprintf(path, sizeof(path), "%s/%08X.chunk", dir, (unsigned)id) >= (int)sizeof(path)) return NULL;
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  uint8_t *b = NULL;
  long n = -1;
  if (fseek(f, 0, SEEK_END) == 0) n = ftell(f);
  if (n >= 32 && n <= INT32_MAX && fseek(f, 0, SEEK_SET) == 0 && (b = (uint8_t *)malloc((size_t)n)) != NULL &&
      fread(b, 1, (size_t)n, f) != (size_t)n) {
    free(b);
    b = NULL;
  }
  fclose(f);
  if (b) {
    int32_t cs;
    memcpy(&cs, b + 12, 4);
    if (cs < 32 || cs > n) {  /* the chunk's own size must fit the file */
      free(b);
      return NULL;
    }
    *csize = cs;
  }
  return b;
}

/* `dir` non-NULL: d is the index frame (chunks.b2frame) of a sparse frame
 * directory; its offsets chunk holds chunk-file numbers instead of byte
 * offsets and follows the header directly (the index has no data section). */
/* `map` non-NULL: d is that read-only mapping of the frame file (size
 * bytes), and the container is made lazy: its data chunks stay in the file
 * until iax_materialize locates them. */
static ina_rc_t container_from_frame(iarray_context_t *ctx, const uint8_t *d, int64_t size, const char *dir,
                                     uint8_t *map, iarray_container_t **container) {
  if (size < 64 || memcmp(d, frame_magic, sizeof(frame_magic)) != 0) return frame_error("no b2frame header");
  mp_t m = {d + 10, d + size, 0};
  const int64_t hsize = mp_int(&m);
  const int64_t fsize = mp_int(&m);
  int64_t flen;
  const uint8_t *gflags = mp_bytes(&m, 0, &flen); /* general flags */
  const int64_t nbytes = mp_int(&m), cbytes = mp_int(&m);
  const int64_t typesize = mp_int(&m), blocksize = mp_int(&m), chunksize = mp_int(&m);
  iax_frame_params_t fp;
  read_params(&m, &fp); /* thread counts, a flag, the coding parameters */
  if (gflags && flen == 4) memcpy(fp.flags, gflags, 4);
  else fp.valid = 0;
  if (m.bad || hsize <= 0 || hsize > size || fsize > size || nbytes < 0 || cbytes < 0 || typesize <= 0 ||
      blocksize <= 0 || chunksize < 0)
    return frame_error("header fields");
  m.end = d + hsize;
  if (mp_len(&m, 0) < 2) return frame_error("metalayer section");
  (void)mp_int(&m); /* index size */
  const int64_t nmeta = mp_len(&m, 1);
  int64_t cat_off = -1, ia_off = -1;
  for (int64_t i = 0; i < nmeta && !m.bad; i++) {
    int64_t nl;
    const uint8_t *name = mp_bytes(&m, 0, &nl);
    const int64_t off = mp_int(&m);
    if (!name) break;
    if (nl == 7 && memcmp(name, "caterva", 7) == 0) cat_off = off;
    if (nl == 6 && memcmp(name, "iarray", 6) == 0) ia_off = off;
  }
  if (m.bad) return frame_error("metalayer map");
  if (cat_off < 0) return iax_set_error(INA_ERROR(IARRAY_ERR_CATERVA_FAILED), "frame has no caterva metalayer");
  if (ia_off < 0) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "frame has no iarray metalayer");

  /* caterva: [version, ndim, shape, chunkshape, blockshape] */
  const uint8_t *cm;
  int64_t cl;
  ina_rc_t rc = read_meta(d, hsize, cat_off, &cm, &cl);
  if (INA_FAILED(rc)) return rc;
  mp_t c = {cm, cm + cl, 0};
  iarray_dtshape_t dts;
  iarray_storage_t st;
  memset(&dts, 0, sizeof(dts));
  memset(&st, 0, sizeof(st));
  if (mp_len(&c, 0) != 5) return frame_error("caterva metalayer");
  (void)mp_int(&c);
  const int64_t ndim = mp_int(&c);
  if (c.bad || ndim < 1 || ndim > IARRAY_DIMENSION_MAX) return frame_error("caterva ndim");
  int64_t *dst[3] = {dts.shape, st.chunkshape, st.blockshape};
  for (int k = 0; k < 3; k++) {
    if (mp_len(&c, 0) != ndim) return frame_error("caterva shapes");
    for (int64_t i = 0; i < ndim; i++) dst[k][i] = mp_int(&c);
  }
  if (c.bad) return frame_error("caterva shapes");
  dts.ndim = (int8_t)ndim;

  /* iarray: [version, dtype, flags] */
  const uint8_t *im;
  int64_t il;
  rc = read_meta(d, hsize, ia_off, &im, &il);
  if (INA_FAILED(rc)) return rc;
  if (il < 3 || im[0] != 0x93) return frame_error("iarray metalayer");
  if (im[2] >= IARRAY_DATA_TYPE_MAX) return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_DTYPE), "The data type is invalid");
  dts.dtype = (iarray_data_type_t)im[2];
  if (iax_dtype_size(dts.dtype) != typesize) return frame_error("typesize disagrees with the dtype");

  /* the metalayers must describe the frame's own chunks: a Caterva chunk is
   * its chunk shape rounded up to whole blocks, and there is one per cell of
   * the chunk grid */
  int64_t ext = typesize, ncells = 1;
  for (int64_t i = 0; i < ndim; i++) {
    const int64_t s = dts.shape[i], ch = st.chunkshape[i], bl = st.blockshape[i];
    if (s < 0 || ch < 1 || bl < 1 || bl > ch || ch > ((int64_t)1 << 31) || s > ((int64_t)1 << 48))
      return frame_error("caterva shapes");
    ext *= (ch + bl - 1) / bl * bl;
    ncells *= s == 0 ? 0 : (s + ch - 1) / ch;
    if (ext > ((int64_t)1 << 31) || ncells > ((int64_t)1 << 31)) return frame_error("caterva shapes");
  }
  const int64_t nchunks = chunksize > 0 ? (nbytes + chunksize - 1) / chunksize : 0;
  if (ext != chunksize || ncells != nchunks) return frame_error("caterva metalayer disagrees with the chunks");

  /* chunk offsets: the chunk after the data, stored (memcpyed) */
  /* bounded before adding: hsize and cbytes come from the file */
  const int64_t dbytes = dir ? 0 : cbytes;  /* bytes of the data section */
  if (dbytes > size - hsize - 32) return frame_error("offsets chunk position");
  const int64_t opos = hsize + dbytes;
  const uint8_t *oc = d + opos;
  int32_t onb, ocb;
  memcpy(&onb, oc + 4, 4);
  memcpy(&ocb, oc + 12, 4);
  if (oc[0] < 1 || onb != 8 * nchunks || ocb < 32 || opos + ocb > size) return frame_error("offsets chunk");
  if (!(oc[2] & 0x02) || ocb < 32 + onb)
    return iax_set_error(INA_ERROR(INA_ERR_NOT_SUPPORTED), "frames with a compressed offsets chunk are not read");
  const uint8_t **ptrs = (const uint8_t **)malloc(sizeof(*ptrs) * (size_t)(nchunks > 0 ? nchunks : 1));
  int32_t *csz = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nchunks > 0 ? nchunks : 1));
  uint8_t **special = (uint8_t **)calloc((size_t)(nchunks > 0 ? nchunks : 1), sizeof(uint8_t *));
  int64_t *loff = map ? (int64_t *)malloc(sizeof(int64_t) * (size_t)(nchunks > 0 ? nchunks : 1)) : NULL;
  if (!ptrs || !csz || !special || (map && !loff)) {
    free(ptrs);
    free(csz);
    free(special);
    free(loff);
    return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  }
  rc = INA_SUCCESS;
  for (int64_t i = 0; i < nchunks && INA_SUCCEED(rc); i++) {
    int64_t off;
    memcpy(&off, oc + 32 + 8 * i, 8);
    if (off < 0) {
      /* a special chunk has no bytes in the frame: bit 63 of its offset is
       * set and the top byte's low 7 bits name the value (1 zeros, 2 NaNs,
       * 4 uninitialized) */
      const int kind = (int)(((uint64_t)off >> 56) & 0x7f);
      if (kind != 1 && kind != 2 && kind != 4) {
        rc = frame_error("special chunk kind");
        break;
      }
      const int64_t cn = (i == nchunks - 1 && nbytes % chunksize) ? nbytes % chunksize : chunksize;
      int64_t bsz = typesize;
      for (int64_t k = 0; k < ndim; k++) bsz *= st.blockshape[k];
      if (!(special[i] = iax_special_chunk(kind, (int32_t)cn, (int32_t)bsz, (int)typesize))) {
        rc = INA_ERROR(INA_ERR_OUT_OF_MEMORY);
        break;
      }
      ptrs[i] = special[i];
      csz[i] = 32;
      if (loff) loff[i] = -1;
      continue;
    }
    if (dir) {
      if (!(special[i] = read_chunk_file(dir, off, &csz[i]))) {
        rc = frame_error("chunk file");
        break;
      }
      ptrs[i] = special[i];
      continue;
    }
    if (off > opos - hsize - 32) {
      rc = frame_error("chunk offset");
      break;
    }
    if (loff) { /* located in the mapping on first use */
      loff[i] = off;
      ptrs[i] = NULL;
      csz[i] = 0;
      continue;
    }
    const int64_t pos = hsize + off;
    int32_t cs;
    memcpy(&cs, d + pos + 12, 4);
    if (cs < 32 || pos + cs > opos) {
      rc = frame_error("chunk size");
      break;
    }
    ptrs[i] = d + pos;
    csz[i] = cs;
  }
  if (INA_SUCCEED(rc) && map) {
    rc = iax_container_new(ctx, &dts, &st, container);
    if (INA_SUCCEED(rc) && (*container)->nchunks != nchunks) {
      iarray_container_free(ctx, container);
      rc = frame_error("chunk count");
    }
    if (INA_SUCCEED(rc)) {
      iarray_container_t *c = *container;
      for (int64_t i = 0; i < nchunks; i++)
        if (special[i]) { /* special chunks are made here and owned by the container */
          c->chunks[i] = special[i];
          c->csizes[i] = csz[i];
          special[i] = NULL;
        }
      c->lazy.map = map;
      c->lazy.map_len = (size_t)size;
      c->lazy.hsize = hsize;
      c->lazy.opos = opos;
      c->lazy.off = loff;
      c->lazy.pending = true;
      loff = NULL;
    }
  } else if (INA_SUCCEED(rc)) {
    rc = iarray_from_chunks(ctx, &dts, &st, nchunks, ptrs, csz, container);
  }
  if (INA_SUCCEED(rc)) (*container)->frame = fp;
  for (int64_t i = 0; i < nchunks; i++) free(special[i]);
  free(special);
  free(ptrs);
  free(csz);
  free(loff);
  return rc;
}

/* iarray_container_open of a contiguous frame file: the file is mapped
 * read-only and the container reads its chunks from the mapping on first use
 * (the reference's Caterva opens the frame and reads chunks lazily); the
 * mapping lives until the container is freed. */
static ina_rc_t open_mapped(iarray_context_t *ctx, const char *urlpath, iarray_container_t **container) {
  const int fd = open(urlpath, O_RDONLY);
  if (fd < 0) return iax_set_error(INA_ERROR(IARRAY_ERR_CATERVA_FAILED), "cannot open %s", urlpath);
  struct stat sb;
  void *m = MAP_FAILED;
  if (fstat(fd, &sb) == 0 && sb.st_size > 0) m = mmap(NULL, (size_t)sb.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
  close(fd);
  if (m == MAP_FAILED) return iax_set_error(INA_ERROR(IARRAY_ERR_CATERVA_FAILED), "cannot map %s", urlpath);
  const ina_rc_t rc = container_from_frame(ctx, (const uint8_t *)m, (int64_t)sb.st_size, NULL, (uint8_t *)m, container);
  if (INA_FAILED(rc)) {
    munmap(m, (size_t)sb.st_size);
    return rc;
  }
  (*container)->on_disk = true;
  return INA_SUCCESS;
}

static ina_rc_t load(iarray_context_t *ctx, const char *urlpath, bool on_disk, iarray_container_t **container) {
  if (!ctx || !urlpath || !container) return INA_ERROR(INA_ERR_NULL_POINTER);
  if (access(urlpath, F_OK) != 0) return iax_set_error(INA_ERROR(INA_ERR_FILE_OPEN), "%s does not exist", urlpath);
  struct stat sb;
  const bool sparse = stat(urlpath, &sb) == 0 && S_ISDIR(sb.st_mode);
  if (on_disk && !sparse) return open_mapped(ctx, urlpath, container);
  char ipath[4096];
  if (sparse && snprintf(ipath, sizeof(ipath), "%s/chunks.b2frame", urlpath) >= (int)sizeof(ipath))
    return iax_set_error(INA_ERROR(INA_ERR_FILE_OPEN), "path too long");
  FILE *f = fopen(sparse ? ipath : urlpath, "rb");
  if (!f) return iax_set_error(INA_ERROR(IARRAY_ERR_CATERVA_FAILED), "cannot open %s", urlpath);
  uint8_t *d = NULL;
  int64_t size = -1;
  if (fseek(f, 0, SEEK_END) == 0) size = ftell(f);
  if (size > 0 && fseek(f, 0, SEEK_SET) == 0 && (d = (uint8_t *)malloc((size_t)size)) != NULL &&
      fread(d, 1, (size_t)size, f) != (size_t)size) {
    free(d);
    d = NULL;
  }
  fclose(f);
  if (!d) return iax_set_error(INA_ERROR(IARRAY_ERR_CATERVA_FAILED), "cannot read %s", urlpath);
  const ina_rc_t rc = container_from_frame(ctx, d, size, sparse ? urlpath : NULL, NULL, container);
  free(d);
  if (INA_SUCCEED(rc)) (*container)->on_disk = on_disk;
  return rc;
}

INA_API(ina_rc_t) iarray_container_load(iarray_context_t *ctx, char *urlpath, iarray_container_t **container) {
  return load(ctx, urlpath, false, container);
}

/* The reference keeps an opened frame on disk and reads chunks lazily; so
 * does this one for a frame file (a frame directory's chunk files are read at
 * open), and the container counts as stored on disk (it cannot be saved). */
INA_API(ina_rc_t) iarray_container_open(iarray_context_t *ctx, char *urlpath, iarray_container_t **container) {
  return load(ctx, urlpath, true, container);
}

INA_API(ina_rc_t) iarray_from_cframe(iarray_context_t *ctx, uint8_t *frame, int64_t frame_len, bool copy,
                                     iarray_container_t **container) {
  (void)copy; /* the chunks are always copied into the container */
  if (!ctx || !frame || !container) return INA_ERROR(INA_ERR_NULL_POINTER);
  return container_from_frame(ctx, frame, frame_len, NULL, NULL, container);
}

/* ---- writing a frame ------------------------------------------------------------
 * The layout the reader above takes, written in c-blosc2's encodings (checked
 * against the reference's frame files, which a load and a save reproduce byte
 * for byte):
 *   header   14-element array: "b2frame\0", int32 header size, uint64 frame
 *            size, the 4 flag bytes (format version with the 64-bit-offsets
 *            bit, frame type, codec | clevel << 4, 0), int64 nbytes and
 *            cbytes, int32 typesize, blocksize and chunksize, int16 thread
 *            counts, a boolean (no variable-length metalayers), a fixext16 of
 *            coding parameters (filters[6], codec, codec meta, filters_meta[6],
 *            2 reserved), then [uint16 index size, map16 {name: int32 offset},
 *            array16 [bin32 content]] with the caterva metalayer before the
 *            iarray one; the index size counts from the array's first byte to
 *            the end of the map
 *   chunks   back to back from the end of the header
 *   offsets  a memcpyed Blosc2 chunk of int64 chunk offsets (from the end of
 *            the header)
 *   trailer  [1, [uint16 6, {}, []], uint32 trailer size, fixext16 of zeros]
 */
static uint8_t *put(uint8_t *p, uint64_t v, int n) {
  for (int i = n - 1; i >= 0; i--) *p++ = (uint8_t)(v >> (8 * i));
  return p;
}

static uint8_t *put_le(uint8_t *p, uint64_t v, int n) {
  for (int i = 0; i < n; i++) *p++ = (uint8_t)(v >> (8 * i));
  return p;
}

static uint8_t *put_str(uint8_t *p, const char *s) {
  const size_t n = strlen(s);
  *p++ = (uint8_t)(0xa0 | n);
  memcpy(p, s, n);
  return p + n;
}

/* sparse: the index frame of a sparse frame directory (frame type 1 in the
 * second flag byte, no data section, chunk-file numbers as offsets); the
 * chunks go to their own files (sparse_write). */
static ina_rc_t frame_write(const iarray_container_t *c, bool sparse, uint8_t **out, int64_t *out_len) {
  if (c->viewed)
    return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_STORAGE),
                         "Container must be stored on a blosc schunk and must not be a view");
  const int64_t n = c->nchunks;
  int64_t cbytes = 0;
  for (int64_t i = 0; i < n; i++) {
    if (!c->chunks[i] || c->csizes[i] < 32) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "chunk %lld is missing", (long long)i);
    cbytes += c->csizes[i];
  }
  if (8 * n > INT32_MAX - 32) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "too many chunks for one frame");

  /* metalayer contents */
  const int nd = c->dtshape.ndim;
  uint8_t cat[3 + 3 * (1 + 9 * IARRAY_DIMENSION_MAX)], *q = cat;
  *q++ = 0x95;
  *q++ = 0; /* version */
  *q++ = (uint8_t)nd;
  *q++ = (uint8_t)(0x90 | nd);
  for (int i = 0; i < nd; i++) {
    *q++ = 0xd3;
    q = put(q, (uint64_t)c->dtshape.shape[i], 8);
  }
  const int64_t *tiles[2] = {c->storage.chunkshape, c->storage.blockshape};
  for (int k = 0; k < 2; k++) {
    *q++ = (uint8_t)(0x90 | nd);
    for (int i = 0; i < nd; i++) {
      *q++ = 0xd2;
      q = put(q, (uint64_t)(uint32_t)tiles[k][i], 4);
    }
  }
  const int64_t catlen = q - cat;
  const uint8_t ia[4] = {0x93, IARRAY_METALAYER_VERSION, (uint8_t)c->dtshape.dtype, 0};

  /* header fields the chunks do not fix */
  iax_frame_params_t fp = c->frame;
  if (!fp.valid) {
    memset(&fp, 0, sizeof(fp));
    const int codec = (int)c->cfg.compression_codec, clevel = c->cfg.compression_level;
    fp.flags[0] = 0x12; /* frame format 2, 64-bit offsets */
    fp.flags[2] = (uint8_t)((codec & 0x0f) | (clevel << 4));
    fp.nthreads[0] = fp.nthreads[1] = c->cfg.max_num_threads;
    fp.vlmeta = 0xc2;
    fp.coding_type = 6; /* filter slots */
    const int f = iax_filter_code(&c->cfg);
    fp.coding[5] = (uint8_t)(f == IAMX_FILTER_SHUFFLE ? 1 : f == IAMX_FILTER_BITSHUFFLE ? 2 : 0);
    fp.coding[6] = (uint8_t)codec;
  }
  fp.flags[1] = sparse ? 1 : 0;
  const int64_t dbytes = sparse ? 0 : cbytes;

  const int64_t meta_at = 87, cat_off = meta_at + 32 + 3, ia_off = cat_off + 5 + catlen;
  const int64_t hsize = ia_off + 5 + (int64_t)sizeof(ia);
  const int64_t olen = 32 + 8 * n, tlen = 35;
  const int64_t fsize = hsize + dbytes + olen + tlen;
  uint8_t *d = (uint8_t *)malloc((size_t)fsize);
  if (!d) return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
  uint8_t *p = d;
  memcpy(p, frame_magic, sizeof(frame_magic));
  p += sizeof(frame_magic);
  *p++ = 0xd2;
  p = put(p, (uint64_t)hsize, 4);
  *p++ = 0xcf;
  p = put(p, (uint64_t)fsize, 8);
  *p++ = 0xa4;
  memcpy(p, fp.flags, 4);
  p += 4;
  *p++ = 0xd3;
  p = put(p, (uint64_t)(n * c->chunk_nbytes), 8);
  *p++ = 0xd3;
  p = put(p, (uint64_t)dbytes, 8);
  const int64_t i32s[3] = {c->itemsize, c->blocksize, c->chunk_nbytes};
  for (int i = 0; i < 3; i++) {
    *p++ = 0xd2;
    p = put(p, (uint64_t)(uint32_t)i32s[i], 4);
  }
  for (int i = 0; i < 2; i++) {
    *p++ = 0xd1;
    p = put(p, (uint64_t)(uint16_t)fp.nthreads[i], 2);
  }
  *p++ = fp.vlmeta;
  *p++ = 0xd8;
  *p++ = fp.coding_type;
  memcpy(p, fp.coding, 16);
  p += 16;
  /* metalayers */
  *p++ = 0x93;
  *p++ = 0xcd;
  p = put(p, 32, 2);
  *p++ = 0xde;
  p = put(p, 2, 2);
  p = put_str(p, "caterva");
  *p++ = 0xd2;
  p = put(p, (uint64_t)cat_off, 4);
  p = put_str(p, "iarray");
  *p++ = 0xd2;
  p = put(p, (uint64_t)ia_off, 4);
  *p++ = 0xdc;
  p = put(p, 2, 2);
  *p++ = 0xc6;
  p = put(p, (uint64_t)catlen, 4);
  memcpy(p, cat, (size_t)catlen);
  p += catlen;
  *p++ = 0xc6;
  p = put(p, sizeof(ia), 4);
  memcpy(p, ia, sizeof(ia));
  p += sizeof(ia);
  /* chunks */
  uint8_t *offs = d + hsize + dbytes + 32;
  int64_t at = 0;
  for (int64_t i = 0; i < n; i++) {
    if (sparse) {
      put_le(offs + 8 * i, (uint64_t)i, 8);
      continue;
    }
    memcpy(p, c->chunks[i], (size_t)c->csizes[i]);
    p += c->csizes[i];
    put_le(offs + 8 * i, (uint64_t)at, 8);
    at += c->csizes[i];
  }
  /* offsets chunk: version 5, LZ4 format 1, memcpyed + shuffle + extended header flags, typesize 8 */
  memset(p, 0, 32);
  p[0] = 5;
  p[1] = 1;
  p[2] = 0x07;
  p[3] = 8;
  put_le(p + 4, (uint64_t)(8 * n), 4);
  put_le(p + 8, (uint64_t)(8 * n), 4);
  put_le(p + 12, (uint64_t)olen, 4);
  p[21] = 1;
  p += olen;
  /* trailer */
  *p++ = 0x94;
  *p++ = 0x01;
  *p++ = 0x93;
  *p++ = 0xcd;
  p = put(p, 6, 2);
  *p++ = 0xde;
  p = put(p, 0, 2);
  *p++ = 0xdc;
  p = put(p, 0, 2);
  *p++ = 0xce;
  p = put(p, (uint64_t)tlen, 4);
  *p++ = 0xd8;
  *p++ = 0;
  memset(p, 0, 16);
  p += 16;
  if (p - d != fsize) {
    free(d);
    return iax_set_error(INA_ERROR(IARRAY_ERR_ASSERTION_FAILED), "frame layout");
  }
  *out = d;
  *out_len = fsize;
  return INA_SUCCESS;
}

static int write_file(const char *path, const uint8_t *d, int64_t len) {
  FILE *f = fopen(path, "wb");
  int ok = f && fwrite(d, 1, (size_t)len, f) == (size_t)len;
  if (f && fclose(f) != 0) ok = 0;
  return ok;
}

static ina_rc_t write_contiguous(const iarray_container_t *c, const char *urlpath) {
  uint8_t *d;
  int64_t len;
  ina_rc_t rc = frame_write(c, false, &d, &len);
  if (INA_FAILED(rc)) return rc;
  const int ok = write_file(urlpath, d, len);
  free(d);
  if (!ok) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "cannot write %s", urlpath);
  return INA_SUCCESS;
}

/* A sparse frame directory: chunks.b2frame (the index frame) and one
 * %08X.chunk file per chunk; whatever the path held before is replaced. */
static ina_rc_t write_sparse(const iarray_container_t *c, const char *urlpath) {
  ina_rc_t rc = iarray_container_remove((char *)urlpath);
  if (INA_FAILED(rc)) return rc;
  if (mkdir(urlpath, 0755) != 0) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "cannot create %s", urlpath);
  uint8_t *d;
  int64_t len;
  rc = frame_write(c, true, &d, &len);
  if (INA_FAILED(rc)) return rc;
  char path[4096];
  int ok = snprintf(path, sizeof(path), "%s/chunks.b2frame", urlpath) < (int)sizeof(path) && write_file(path, d, len);
  free(d);
  for (int64_t i = 0; ok && i < c->nchunks; i++)
    ok = snprintf(path, sizeof(path), "%s/%08X.chunk", urlpath, (unsigned)i) < (int)sizeof(path) &&
         write_file(path, c->chunks[i], c->csizes[i]);
  if (!ok) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "cannot write %s", urlpath);
  return INA_SUCCESS;
}

/* A container created with storage.urlpath keeps its chunks in memory (and
 * HBM) and writes them through to its frame once they are complete: a
 * contiguous frame file, or a sparse frame directory when
 * storage.contiguous is false (src/iarray_constructor.c:16-45 hands the
 * storage to Caterva, which writes the super-chunk to that path).  Chunks
 * not written yet are stored as uninitialized special chunks. */
ina_rc_t iax_persist(iarray_container_t *c) {
  if (!c || c->viewed || !c->storage.urlpath) return INA_SUCCESS;
  for (int64_t i = 0; i < c->nchunks; i++)
    if (!c->chunks[i]) {
      if (!(c->chunks[i] = iax_special_chunk(4, (int32_t)c->chunk_nbytes, c->blocksize, c->itemsize)))
        return INA_ERROR(INA_ERR_OUT_OF_MEMORY);
      c->csizes[i] = 32;
    }
  return c->storage.contiguous ? write_contiguous(c, c->storage.urlpath) : write_sparse(c, c->storage.urlpath);
}

INA_API(ina_rc_t) iarray_to_cframe(iarray_context_t *ctx, iarray_container_t *src, uint8_t **frame,
                                   int64_t *frame_len, bool *needs_free) {
  if (!ctx || !src || !frame || !frame_len || !needs_free) return INA_ERROR(INA_ERR_NULL_POINTER);
  ina_rc_t rc = iax_materialize(src);
  if (INA_FAILED(rc)) return rc;
  rc = frame_write(src, false, frame, frame_len);
  *needs_free = INA_SUCCEED(rc);
  return rc;
}

INA_API(ina_rc_t) iarray_container_save(iarray_context_t *ctx, iarray_container_t *container, char *urlpath) {
  if (!ctx || !container || !urlpath) return INA_ERROR(INA_ERR_NULL_POINTER);
  if (container->on_disk && !container->viewed) return iax_set_error(INA_ERROR(IARRAY_ERR_INVALID_STORAGE), "Container is already on disk");
  return write_contiguous(container, urlpath);
}

INA_API(ina_rc_t) iarray_container_remove(char *urlpath) {
  if (!urlpath) return INA_SUCCESS;
  struct stat st;
  if (stat(urlpath, &st) != 0)
    return errno == ENOENT ? INA_SUCCESS : iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "cannot access %s", urlpath);
  if (S_ISDIR(st.st_mode)) {
    /* a frame directory holds its index and chunk files, no subdirectories */
    DIR *dir = opendir(urlpath);
    if (!dir) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "cannot open %s", urlpath);
    struct dirent *e;
    int ok = 1;
    char path[4096];
    while ((e = readdir(dir)) != NULL) {
      if (strcmp(e->d_name, ".") == 0 || strcmp(e->d_name, "..") == 0) continue;
      if (snprintf(path, sizeof(path), "%s/%s", urlpath, e->d_name) >= (int)sizeof(path) || remove(path) != 0) ok = 0;
    }
    closedir(dir);
    if (!ok || rmdir(urlpath) != 0) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "cannot remove %s", urlpath);
    return INA_SUCCESS;
  }
  if (remove(urlpath) != 0) return iax_set_error(INA_ERROR(IARRAY_ERR_BLOSC_FAILED), "cannot remove %s", urlpath);
  return INA_SUCCESS;
}
