"""GPU parity of the block decode stage (bit-exact).

* the reference's Blosc2 frame fixtures (written by c-blosc2) decode on the GPU
  to the SHA-256s of tests/golden/fixture_hashes.json, as raw chunks and as
  containers read by iarray_container_load;
* chunks from c-blosc 1.21 (byte shuffle, bitshuffle, ragged element counts,
  format-2 rules) decode on the GPU to their inputs, when that library exists;
* every tiling / dtype / filter / split mode the host writes round-trips
  through the GPU decoder bit-exactly, checked against the oracle decoder;
* a corrupted LZ4 stream is reported as IARRAY_ERR_BLOSC_FAILED (the kernel
  bounds-checks every read and write, so no fault is possible);
* LZ4 blocks built sequence by sequence (literal and match lengths with zero,
  one and many length bytes, offsets from 1 to 65535, periodic matches) decode
  to the bytes they were built from, and the oracle's decoder agrees.
"""
import ctypes
import glob
import hashlib
import json
import os
import struct
import zlib

import numpy as np
import pytest

import pyiarray as ia

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
HASHES = json.load(open(os.path.join(GOLD, "fixture_hashes.json")))


@pytest.fixture(scope="module")
def ctx():
    return ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "frames", "*.iarray"))), ids=os.path.basename)
def test_fixture_frames_decode_on_gpu(ctx, orc, path):
    print(f"case: fixture {os.path.basename(path)}", flush=True)
    chunks, nbytes = orc.frame_chunks(open(path, "rb").read())
    out = ia.decode_chunks_flat(ctx, chunks, nbytes)
    assert hashlib.sha256(out).hexdigest() == HASHES[os.path.basename(path)]["sha256"]


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "frames", "*.iarray"))), ids=os.path.basename)
def test_loaded_frame_decodes_to_known_answer(ctx, path):
    """iarray_container_load reads the reference's frame file; the container
    decodes on the GPU to the fixture's SHA-256."""
    print(f"case: load {os.path.basename(path)}", flush=True)
    c = ia.Container.load(ctx, path)
    assert hashlib.sha256(c.to_numpy().tobytes()).hexdigest() == HASHES[os.path.basename(path)]["sha256"]


def test_fixture_frame_as_container(ctx, orc):
    """The same frame, wrapped as a 1-D container with the generator's tiling
    (shape 1e6, chunks 1e5, blocks 1e4; tests/generator.py:23-38)."""
    path = os.path.join(GOLD, "frames", "test_rand_float32.iarray")
    chunks, nbytes = orc.frame_chunks(open(path, "rb").read())
    c = ia.Container.from_chunks(ctx, np.float32, (1000000,), (100000,), (10000,), chunks)
    x = c.to_numpy()
    assert hashlib.sha256(x.tobytes()).hexdigest() == HASHES["test_rand_float32.iarray"]["sha256"]
    assert np.array_equal(c.slice((123456,), (223457,)), x[123456:223457])


def _libblosc():
    try:
        b = ctypes.CDLL("/opt/conda/lib/libblosc.so.1")
    except OSError:
        return None
    b.blosc_init()
    b.blosc_compress.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_size_t]
    b.blosc_set_blocksize.argtypes = [ctypes.c_size_t]
    b.blosc_set_compressor.argtypes = [ctypes.c_char_p]
    b.blosc_set_compressor(b"lz4")
    return b


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shuf", [1, 2], ids=["shuffle", "bitshuffle"])
def test_cblosc1_chunks_decode_on_gpu(ctx, dtype, shuf):
    b = _libblosc()
    if b is None:
        pytest.skip("c-blosc 1.21 not present")
    rng = np.random.default_rng(5)
    for n in (4096, 4100, 62500):
        for x in (rng.random(n), np.sin(np.arange(n) * 1e-3) + rng.normal(0, 1e-3, n), np.arange(n) % 9):
            x = x.astype(dtype)
            for bs in (0, 2048, 8000):
                print(f"case: cblosc1 {np.dtype(dtype).name} shuf={shuf} n={n} bs={bs}", flush=True)
                b.blosc_set_blocksize(bs)
                dst = ctypes.create_string_buffer(x.nbytes + 16)
                cs = b.blosc_compress(5, shuf, x.itemsize, x.nbytes, x.ctypes.data, dst, x.nbytes + 16)
                assert cs > 0
                assert ia.decode_chunks_flat(ctx, [dst.raw[:cs]], x.nbytes) == x.tobytes()


CASES = [
    (np.float64, (50, 37), (20, 15), (7, 6), ia.COMP_SHUFFLE, 5, ia.AUTO_SPLIT),
    (np.float32, (100, 170), (27, 30), (12, 7), ia.COMP_SHUFFLE, 5, ia.AUTO_SPLIT),
    (np.float32, (64, 64), (64, 64), (16, 16), ia.COMP_BITSHUFFLE, 5, ia.AUTO_SPLIT),
    (np.float32, (61, 47), (30, 20), (9, 13), ia.COMP_BITSHUFFLE, 9, ia.AUTO_SPLIT),
    (np.float64, (33, 65), (33, 65), (11, 13), ia.COMP_SHUFFLE, 1, ia.NEVER_SPLIT),
    (np.float64, (40, 40), (20, 20), (10, 10), 0, 5, ia.ALWAYS_SPLIT),
    (np.float32, (40, 40), (20, 20), (10, 10), ia.COMP_SHUFFLE, 0, ia.AUTO_SPLIT),
    (np.int32, (1000,), (300,), (64,), ia.COMP_SHUFFLE, 5, ia.AUTO_SPLIT),
    (np.float64, (777,), (256,), (100,), ia.COMP_BITSHUFFLE, 3, ia.AUTO_SPLIT),
    (np.float32, (500, 500), (250, 250), (250, 250), ia.COMP_BITSHUFFLE, 5, ia.AUTO_SPLIT),
    (np.float64, (300, 260), (128, 128), (64, 64), ia.COMP_SHUFFLE, 5, ia.AUTO_SPLIT),
    # rows of a multiple of 4 elements with ragged edges, planes in one stream
    # or one stream per plane
    (np.float32, (90, 130), (40, 60), (20, 20), ia.COMP_SHUFFLE, 5, ia.NEVER_SPLIT),
    (np.float64, (90, 130), (40, 60), (20, 12), ia.COMP_SHUFFLE, 5, ia.ALWAYS_SPLIT),
]


def _data(dtype, shape, kind):
    rng = np.random.default_rng(sum(shape) * 31 + len(kind))
    n = int(np.prod(shape))
    if kind == "smooth":
        x = np.sin(np.arange(n) * 1e-2) + rng.normal(0, 1e-3, n)
    elif kind == "random":
        x = rng.standard_normal(n) * 100
    elif kind == "zeros":
        x = np.zeros(n)
    else:
        x = np.arange(n) % 7
    return x.astype(dtype).reshape(shape)


@pytest.mark.parametrize("case", CASES, ids=[f"{np.dtype(c[0]).name}-{c[1]}-{c[2]}-{c[3]}-f{c[4]}-l{c[5]}-s{c[6]}"
                                              for c in CASES])
@pytest.mark.parametrize("kind", ["smooth", "random", "zeros", "ramp"])
def test_roundtrip_gpu_decode_bit_exact(orc, case, kind):
    dtype, shape, chunks, blocks, filt, clevel, split = case
    print(f"case: roundtrip {case} {kind}", flush=True)
    ctx = ia.Context(clevel=clevel, filters=filt, splitmode=split)
    x = _data(dtype, shape, kind)
    c = ia.Container.from_numpy(ctx, x, chunks, blocks)
    got = c.to_numpy()
    assert got.tobytes() == x.tobytes()
    ref = orc.Array(c.chunks(), shape, chunks, blocks, x.itemsize).to_dense(dtype=dtype)
    assert ref.tobytes() == got.tobytes()


def test_zeros_and_empty_read_as_zero(ctx):
    z = ia.Container.zeros(ctx, np.float64, (70, 90), (32, 40), (16, 20))
    assert not z.to_numpy().any()
    e = ia.Container.empty(ctx, np.float32, (70, 90), (32, 40), (16, 20))
    assert not e.to_numpy().any()


def test_transposed_view_reads_transposed(ctx):
    x = np.arange(170 * 100, dtype=np.float64).reshape(170, 100)
    c = ia.Container.from_numpy(ctx, x, (20, 30), (7, 11))
    t = c.transpose()
    assert np.array_equal(t.to_numpy(), x.T)
    assert np.array_equal(t.slice((3, 5), (50, 160)), x.T[3:50, 5:160])


def test_get_slice_buffer(ctx):
    x = np.random.default_rng(3).standard_normal((257, 129)).astype(np.float32)
    c = ia.Container.from_numpy(ctx, x, (100, 64), (25, 16))
    for st, sp in [((0, 0), (257, 129)), ((10, 5), (200, 100)), ((-50, -20), (-1, -1)), ((5, 0), (6, 129))]:
        s = [(a if a >= 0 else a + d) for a, d in zip(st, x.shape)]
        e = [(b if b >= 0 else b + d) for b, d in zip(sp, x.shape)]
        assert np.array_equal(c.slice(st, sp), x[s[0]:e[0], s[1]:e[1]])


def _corrupt_offset(chunk: bytes) -> bytes:
    """Point the first LZ4 match of the first LZ4 stream far before the
    stream start (an offset larger than the bytes produced so far)."""
    b = bytearray(chunk)
    nblocks = -(-struct.unpack("<i", b[4:8])[0] // struct.unpack("<i", b[8:12])[0])
    ts = b[3]
    neb = struct.unpack("<i", b[8:12])[0] // ts
    p = struct.unpack("<i", b[32:36])[0]
    for _ in range(ts * nblocks):
        cs = struct.unpack("<i", b[p:p + 4])[0]
        if 0 < cs < neb:
            q = p + 4
            tok = b[q]
            lit = tok >> 4
            q += 1
            if lit == 15:
                while b[q] == 255:
                    lit += 255
                    q += 1
                lit += b[q]
                q += 1
            q += lit
            b[q] = 0xFF
            b[q + 1] = 0xFF
            return bytes(b)
        p += 4 + max(cs, 0)
    raise AssertionError("no LZ4 stream found")


def test_corrupted_stream_reports_error(ctx):
    x = (np.arange(4096) % 5).astype(np.float32)
    c = ia.Container.from_numpy(ctx, x, (4096,), (1024,))
    bad = _corrupt_offset(c.chunks()[0])
    with pytest.raises(ia.IArrayError) as e:
        ia.decode_chunks_flat(ctx, [bad], x.nbytes)
    assert e.value.rc == ia.ERR_BLOSC_FAILED
    # the library keeps working afterwards
    assert ia.decode_chunks_flat(ctx, c.chunks(), x.nbytes) == x.tobytes()


def test_corrupted_container_errors_on_every_call(ctx):
    """A container holding a malformed LZ4 stream reports the error on every
    read, not only the first: a failed decode leaves no remembered image."""
    x = (np.arange(4096) % 5).astype(np.float32)
    c = ia.Container.from_numpy(ctx, x, (4096,), (1024,))
    bad = ia.Container.from_chunks(ctx, np.float32, (4096,), (4096,), (1024,), [_corrupt_offset(c.chunks()[0])])
    for _ in range(2):
        with pytest.raises(ia.IArrayError) as e:
            bad.to_numpy()
        assert e.value.rc == ia.ERR_BLOSC_FAILED
    # as the x of a GEMV, twice
    A = ia.Container.from_numpy(ctx, np.ones((64, 4096), np.float32), (64, 4096), (64, 1024))
    for _ in range(2):
        with pytest.raises(ia.IArrayError) as e:
            ia.matmul(A, bad, (64,), (64,))
        assert e.value.rc == ia.ERR_BLOSC_FAILED
    assert np.array_equal(c.to_numpy(), x)
    assert np.array_equal(ia.matmul(A, c, (64,), (64,)).to_numpy(), np.full(64, x.sum(), np.float32))


def _lenbytes(v):
    """LZ4 length continuation bytes for a 4-bit field that holds 15."""
    v -= 15
    b = bytearray()
    while v >= 255:
        b.append(255)
        v -= 255
    b.append(v)
    return b


def _craft_lz4(rng, n, pick):
    """An LZ4 block of n decoded bytes whose sequences come from pick(rng) ->
    (literals, offset, match length), kept within the format's end rules.
    Returns (block bytes, decoded bytes)."""
    out = bytearray()
    s = bytearray()
    misses = 0
    while misses < 50:
        lit, off, ml = pick(rng)
        if not out and lit == 0:
            lit = 1
        if len(out) + lit + ml > n - 12:
            misses += 1
            continue
        off = max(1, min(off, len(out) + lit, 65535))
        s.append((min(lit, 15) << 4) | min(ml - 4, 15))
        if lit >= 15:
            s += _lenbytes(lit)
        lits = rng.integers(0, 256, lit, dtype=np.uint8).tobytes()
        s += lits
        out += lits
        s += struct.pack("<H", off)
        if ml - 4 >= 15:
            s += _lenbytes(ml - 4)
        for _ in range(ml):
            out.append(out[-off])
    lit = n - len(out)
    s.append(min(lit, 15) << 4)
    if lit >= 15:
        s += _lenbytes(lit)
    lits = rng.integers(0, 256, lit, dtype=np.uint8).tobytes()
    s += lits
    out += lits
    return bytes(s), bytes(out)


def _one_stream_chunk(ctx, stream, n):
    """A Blosc2 chunk of one uint8 block whose single stream is `stream`."""
    x = (np.arange(n) % 7).astype(np.uint8)
    ch = bytearray(ia.Container.from_numpy(ctx, x, (n,), (n,)).chunks()[0][:32])
    assert ch[3] == 1 and struct.unpack("<i", ch[8:12])[0] == n
    body = struct.pack("<i", 36) + struct.pack("<i", len(stream)) + stream
    ch[12:16] = struct.pack("<i", 32 + len(body))
    return bytes(ch) + body


def _choice(rng, ranges):
    lo, hi = ranges[rng.integers(0, len(ranges))]
    return int(rng.integers(lo, hi + 1))


# (literals, offset, match length) ranges per plan; every path of the LZ4
# kernel: the per-lane header walk (literals up to 59 with one length byte,
# matches up to 273 with one, offsets within and beyond the LDS history ring,
# periodic matches with offsets below 64), and the general path (longer
# literal runs copied global to global, matches with two or more length bytes
# copied four 64-byte steps at a time, offsets beyond the ring read back from
# global memory).
LZ4_PLANS = {
    "short": ([(0, 0), (1, 14)], [(1, 63), (64, 4000)], [(4, 18)]),
    "literal-byte": ([(15, 59), (0, 3)], [(1, 3000)], [(4, 40)]),
    "long-literals": ([(60, 300), (300, 3000), (0, 5)], [(1, 3000)], [(4, 30)]),
    "long-matches": ([(0, 3)], [(1, 8), (9, 63), (64, 255), (256, 300), (1000, 4000)], [(274, 6000), (19, 273)]),
    "far": ([(0, 20)], [(4033, 8127), (8128, 65535), (100, 200)], [(4, 300)]),
    "far-long": ([(0, 3)], [(4033, 65535)], [(274, 2000)]),
    "huge-lengths": ([(0, 3), (17000, 20000)], [(1, 300), (5000, 9000)], [(20000, 60000), (4, 30)]),
    "mixed": ([(0, 0), (1, 14), (15, 59), (60, 700)], [(1, 7), (8, 63), (64, 4031), (4032, 65535)],
              [(4, 18), (19, 273), (274, 3000)]),
    # the kernel's limits, each value exactly: literals around one length byte
    # and the walk's 59, offsets around 64 (periodic copies), the 4 KiB ring
    # (4032 = RING - 64) and its size, match lengths around 64-byte steps and
    # the walk's 273
    "edges": ([(v, v) for v in (0, 1, 14, 15, 16, 58, 59, 60, 63, 64, 65)],
              [(v, v) for v in (1, 2, 3, 4, 31, 32, 33, 63, 64, 65, 255, 256, 4031, 4032, 4033, 4095, 4096, 4097,
                                8191, 8192, 65535)],
              [(v, v) for v in (4, 5, 18, 19, 20, 63, 64, 65, 128, 272, 273, 274, 275)]),
}


@pytest.mark.parametrize("n", [200_000, 70_000, 3000])
@pytest.mark.parametrize("plan", sorted(LZ4_PLANS))
def test_crafted_lz4_streams(ctx, orc, plan, n):
    print(f"case: crafted lz4 {plan} n={n}", flush=True)
    lits, offs, mls = LZ4_PLANS[plan]
    rng = np.random.default_rng(zlib.crc32(plan.encode()) + n)
    stream, want = _craft_lz4(rng, n, lambda r: (_choice(r, lits), _choice(r, offs), _choice(r, mls)))
    assert len(stream) < n
    assert orc.lz4_decode(stream, n) == want  # the generator itself, against the oracle decoder
    c = ia.Context(clevel=5, filters=0)
    chunk = _one_stream_chunk(c, stream, n)
    assert orc.decompress_chunk(chunk, n) == want
    got = ia.decode_chunks_flat(c, [chunk], n)
    assert got == want


# ---- BloscLZ streams and chunks of mixed codecs ------------------------------------------------
def _cblosc_compress(b, compressor, raw, clevel, shuf, ts, bs):
    b.blosc_set_compressor(compressor)
    b.blosc_set_blocksize(bs)
    dst = ctypes.create_string_buffer(len(raw) + 64)
    cs = b.blosc_compress(clevel, shuf, ts, len(raw), raw, dst, len(raw) + 64)
    b.blosc_set_compressor(b"lz4")
    b.blosc_set_blocksize(0)
    assert cs > 0
    return dst.raw[:cs]


def _lz77_structured(rng, n):
    out = bytearray(rng.integers(0, 256, 64, dtype=np.uint8).tobytes())
    while len(out) < n:
        if rng.random() < 0.4:
            out += rng.integers(0, 256, int(rng.integers(1, 80)), dtype=np.uint8).tobytes()
        else:
            d = int(rng.choice([rng.integers(1, 300), rng.integers(300, 8200), rng.integers(8000, 8500),
                                rng.integers(8192, 74000)]))
            d = min(d, len(out))
            for _ in range(int(rng.integers(3, 1200))):
                out.append(out[len(out) - d])
    return bytes(out[:n])


@pytest.mark.parametrize("seed", range(3))
def test_blosclz_chunks_decode_on_gpu(ctx, orc, seed):
    """BloscLZ streams (codec 0) written by c-blosc 1.21.0 (black box) decode
    on the GPU to their input, bit for bit: structured data with far
    matches and long lengths, random bytes and smooth floats, unfiltered,
    shuffled and bitshuffled, clevels 1/5/9, forced block sizes."""
    b = _libblosc()
    if b is None:
        pytest.skip("c-blosc 1.21 not present")
    rng = np.random.default_rng(200 + seed)
    n = 160000
    datas = [_lz77_structured(rng, n), rng.integers(0, 256, n, dtype=np.uint8).tobytes(),
             (np.sin(np.arange(n // 8) * 1e-3) + rng.normal(0, 1e-4, n // 8)).tobytes()]
    for di, raw in enumerate(datas):
        for clevel in (1, 5, 9):
            for ts, shuf in ((1, 0), (4, 1), (8, 1), (4, 2)):
                bs = int(rng.choice([0, 16384, 65536]))
                print(f"case: blosclz seed={seed} data={di} clevel={clevel} ts={ts} shuf={shuf} bs={bs}", flush=True)
                chunk = _cblosc_compress(b, b"blosclz", raw, clevel, shuf, ts, bs)
                assert chunk[2] >> 5 == 0
                assert ia.decode_chunks_flat(ctx, [chunk], len(raw)) == raw


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_mixed_codec_operands_matmul(orc, dt):
    """Operands whose chunks change codec from chunk to chunk, as BTune may
    leave them (src/iarray.c:477-498): each chunk re-compressed by c-blosc
    1.21 with BloscLZ or LZ4 (alternating).  c-blosc picks its own block
    size, which for chunks this small is the whole chunk, so the tilings use
    one block per chunk.  The GPU decodes them to the data and the product
    equals the oracle's on the same chunks (integer data: exact)."""
    b = _libblosc()
    if b is None:
        pytest.skip("c-blosc 1.21 not present")
    ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
    rng = np.random.default_rng(9)
    isz = np.dtype(dt).itemsize

    def mixed(x, cs, bs):
        c = ia.Container.from_numpy(ctx, x, cs, bs)
        chunks = []
        for i, ch in enumerate(c.chunks()):
            raw = orc.decompress_chunk(ch, struct.unpack("<i", ch[4:8])[0])
            comp = b"blosclz" if i % 2 == 0 else b"lz4"
            cc = _cblosc_compress(b, comp, raw, 5, 1, isz, 0)
            assert struct.unpack("<i", cc[8:12])[0] == bs[0] * bs[1] * isz
            chunks.append(cc)
        out = ia.Container.from_chunks(ctx, dt, x.shape, cs, bs, chunks)
        return out, chunks

    a = rng.integers(-20, 21, (300, 260)).astype(dt)
    w = rng.integers(-20, 21, (260, 180)).astype(dt)
    A, ach = mixed(a, (40, 50), (40, 50))
    W, wch = mixed(w, (50, 40), (50, 40))
    assert {c[2] >> 5 for c in ach} == {0, 1}
    assert np.array_equal(A.to_numpy(), a)
    z = ia.matmul(A, W, (100, 90), (50, 45)).to_numpy()
    o = orc.gemm_blosc(orc.Array(ach, a.shape, (40, 50), (40, 50), isz),
                       orc.Array(wch, w.shape, (50, 40), (50, 40), isz), (100, 90), (50, 45))
    assert np.array_equal(z, o) and np.array_equal(z, a @ w)


@pytest.mark.parametrize("seed", range(3))
def test_zstd_chunks_decode_on_gpu(ctx, orc, seed):
    """Zstd streams (codec 4) written by c-blosc 1.21.0 (black box, libzstd
    inside) decode on the GPU to their input, bit for bit, and to the
    oracle's decode: structured data with far matches and long lengths
    (compressed blocks, repeat offsets), random bytes (raw blocks or raw
    streams), smooth floats, runs (RLE blocks), unfiltered, shuffled and
    bitshuffled, clevels 1/5/9, forced block sizes."""
    b = _libblosc()
    if b is None:
        pytest.skip("c-blosc 1.21 not present")
    rng = np.random.default_rng(500 + seed)
    n = 160000
    datas = [_lz77_structured(rng, n), rng.integers(0, 256, n, dtype=np.uint8).tobytes(),
             (np.sin(np.arange(n // 8) * 1e-3) + rng.normal(0, 1e-4, n // 8)).tobytes(),
             np.repeat(rng.integers(0, 256, n // 4000, dtype=np.uint8), 4000).tobytes()]
    for di, raw in enumerate(datas):
        for clevel in (1, 5, 9):
            for ts, shuf in ((1, 0), (4, 1), (8, 1), (4, 2)):
                bs = int(rng.choice([0, 16384, 65536]))
                print(f"case: zstd seed={seed} data={di} clevel={clevel} ts={ts} shuf={shuf} bs={bs}", flush=True)
                chunk = _cblosc_compress(b, b"zstd", raw, clevel, shuf, ts, bs)
                assert chunk[2] >> 5 == 4
                assert orc.decompress_chunk(chunk, len(raw)) == raw
                assert ia.decode_chunks_flat(ctx, [chunk], len(raw)) == raw


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_three_codec_operands_gemm_and_gemv(orc, dt):
    """Operands whose chunks cycle through Zstd, BloscLZ and LZ4 (as BTune
    may leave them, src/iarray.c:477-498), re-compressed by c-blosc 1.21:
    the GEMM and the fused GEMV on the GPU equal the oracle's products on
    the same chunks (integer data: exact)."""
    b = _libblosc()
    if b is None:
        pytest.skip("c-blosc 1.21 not present")
    ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
    rng = np.random.default_rng(19)
    isz = np.dtype(dt).itemsize
    comps = (b"zstd", b"blosclz", b"lz4")

    def mixed(x, cs, bs, shuf=1):
        c = ia.Container.from_numpy(ctx, x, cs, bs)
        chunks = []
        for i, ch in enumerate(c.chunks()):
            raw = orc.decompress_chunk(ch, struct.unpack("<i", ch[4:8])[0])
            cc = _cblosc_compress(b, comps[i % 3], raw, 5, shuf, isz, 0)
            assert struct.unpack("<i", cc[8:12])[0] == (bs[0] * bs[1] if len(bs) == 2 else bs[0]) * isz
            chunks.append(cc)
        return ia.Container.from_chunks(ctx, dt, x.shape, cs, bs, chunks), chunks

    a = rng.integers(-20, 21, (300, 260)).astype(dt)
    w = rng.integers(-20, 21, (260, 180)).astype(dt)
    A, ach = mixed(a, (40, 50), (40, 50))
    W, wch = mixed(w, (50, 40), (50, 40))
    assert {c[2] >> 5 for c in ach} == {0, 1, 4}
    assert np.array_equal(A.to_numpy(), a)
    z = ia.matmul(A, W, (100, 90), (50, 45)).to_numpy()
    o = orc.gemm_blosc(orc.Array(ach, a.shape, (40, 50), (40, 50), isz),
                       orc.Array(wch, w.shape, (50, 40), (50, 40), isz), (100, 90), (50, 45))
    assert np.array_equal(z, o) and np.array_equal(z, a @ w)
    # GEMV with bit-shuffled Zstd / BloscLZ / LZ4 chunks of A
    x = rng.integers(-5, 6, 260).astype(dt)
    A2, a2ch = mixed(a, (40, 50), (40, 50), shuf=2)
    X = ia.Container.from_numpy(ctx, x, (50,), (50,))
    y = ia.matmul(A2, X, (40,), (40,)).to_numpy()
    assert np.array_equal(y, a @ x)


# A GPU zlib decoder (one wavefront per stream, canonical Huffman codes read
# bit by bit) stopped with a memory aperture violation on its first run over
# c-blosc 1.21 zlib chunks; a CPU emulation of the kernel over the same
# streams stays in bounds and decodes them exactly, so the cause was not
# found and the decoder was removed.  These cases stay here, skipped.
ZLIB_SKIP = ("no GPU zlib decoder: the one written faulted on the GPU with its cause not found, "
             "and zlib chunks are refused")


@pytest.mark.skip(reason=ZLIB_SKIP)
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_four_codec_operands_gemm_and_gemv(orc, dt):
    """Operands whose chunks cycle through Zstd, BloscLZ, zlib and LZ4 (as
    BTune may leave them, src/iarray.c:477-498), re-compressed by c-blosc
    1.21: the GEMM and the fused GEMV on the GPU equal the oracle's products
    on the same chunks (integer data: exact)."""
    b = _libblosc()
    if b is None:
        pytest.skip("c-blosc 1.21 not present")
    ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
    rng = np.random.default_rng(19)
    isz = np.dtype(dt).itemsize
    comps = (b"zstd", b"blosclz", b"zlib", b"lz4")

    def mixed(x, cs, bs, shuf=1):
        c = ia.Container.from_numpy(ctx, x, cs, bs)
        chunks = []
        for i, ch in enumerate(c.chunks()):
            raw = orc.decompress_chunk(ch, struct.unpack("<i", ch[4:8])[0])
            cc = _cblosc_compress(b, comps[i % 4], raw, 5, shuf, isz, 0)
            assert struct.unpack("<i", cc[8:12])[0] == (bs[0] * bs[1] if len(bs) == 2 else bs[0]) * isz
            chunks.append(cc)
        return ia.Container.from_chunks(ctx, dt, x.shape, cs, bs, chunks), chunks

    a = rng.integers(-20, 21, (300, 260)).astype(dt)
    w = rng.integers(-20, 21, (260, 180)).astype(dt)
    A, ach = mixed(a, (40, 50), (40, 50))
    W, wch = mixed(w, (50, 40), (50, 40))
    assert {c[2] >> 5 for c in ach} == {0, 1, 3, 4}
    assert np.array_equal(A.to_numpy(), a)
    z = ia.matmul(A, W, (100, 90), (50, 45)).to_numpy()
    o = orc.gemm_blosc(orc.Array(ach, a.shape, (40, 50), (40, 50), isz),
                       orc.Array(wch, w.shape, (50, 40), (50, 40), isz), (100, 90), (50, 45))
    assert np.array_equal(z, o) and np.array_equal(z, a @ w)
    # GEMV with bit-shuffled Zstd / BloscLZ / LZ4 chunks of A
    x = rng.integers(-5, 6, 260).astype(dt)
    A2, a2ch = mixed(a, (40, 50), (40, 50), shuf=2)
    X = ia.Container.from_numpy(ctx, x, (50,), (50,))
    y = ia.matmul(A2, X, (40,), (40,)).to_numpy()
    assert np.array_equal(y, a @ x)


@pytest.mark.skip(reason=ZLIB_SKIP)
@pytest.mark.parametrize("seed", range(2))
def test_zlib_chunks_decode_on_gpu(ctx, orc, seed):
    """zlib streams (codec 3) written by c-blosc 1.21.0 (black box) decode on
    the GPU to their input, bit for bit, and to the oracle's decode: stored,
    fixed and dynamic blocks from structured, random, smooth and run data,
    unfiltered, shuffled and bitshuffled, clevels 1/5/9."""
    b = _libblosc()
    if b is None:
        pytest.skip("c-blosc 1.21 not present")
    rng = np.random.default_rng(800 + seed)
    n = 120000
    datas = [_lz77_structured(rng, n), rng.integers(0, 256, n, dtype=np.uint8).tobytes(),
             (np.sin(np.arange(n // 8) * 1e-3) + rng.normal(0, 1e-4, n // 8)).tobytes(),
             np.repeat(rng.integers(0, 256, n // 3000, dtype=np.uint8), 3000).tobytes()]
    for di, raw in enumerate(datas):
        for clevel in (1, 5, 9):
            for ts, shuf in ((1, 0), (4, 1), (8, 2)):
                bs = int(rng.choice([0, 16384]))
                print(f"case: zlib seed={seed} data={di} clevel={clevel} ts={ts} shuf={shuf} bs={bs}", flush=True)
                chunk = _cblosc_compress(b, b"zlib", raw, clevel, shuf, ts, bs)
                assert chunk[2] >> 5 == 3
                assert orc.decompress_chunk(chunk, len(raw)) == raw
                assert ia.decode_chunks_flat(ctx, [chunk], len(raw)) == raw


def test_zlib_chunks_are_refused(ctx):
    """zlib streams (codec 3) have no GPU decoder: a chunk holding them is
    refused when its decode is planned, with an error that names the codec,
    and nothing is decoded from it."""
    b = _libblosc()
    if b is None:
        pytest.skip("c-blosc 1.21 not present")
    raw = np.arange(20000, dtype=np.float32).tobytes()
    chunk = _cblosc_compress(b, b"zlib", raw, 5, 1, 4, 0)
    assert chunk[2] >> 5 == 3
    with pytest.raises(ia.IArrayError, match="codec format 3 has no GPU decoder"):
        ia.decode_chunks_flat(ctx, [chunk], len(raw))
