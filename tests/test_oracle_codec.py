"""CPU pins of the oracle's decode stage (the checker the GPU decoder is held to).

Pins, each chosen so that a plausible decoder mistake fails one of them:
  * the reference's own Blosc2 frame fixtures (tests/data/<name>.iarray, written by
    c-blosc2) decode to SHA-256s computed independently with liblz4 + numpy;
  * liblz4 1.9.3 (black box) compressed buffers decode to their inputs;
  * c-blosc 1.21 (black box) chunks -- byte shuffle, bitshuffle, ragged element
    counts, forced block sizes -- decode to their inputs;
  * the bitshuffle permutation equals a numpy bit-matrix restatement.
"""
import ctypes
import glob
import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
HASHES = json.load(open(os.path.join(GOLD, "fixture_hashes.json")))


def _frames():
    names = sorted(glob.glob(os.path.join(GOLD, "frames", "*.iarray")))
    ref = "/root/reference/tests/data"
    if os.path.isdir(ref):
        have = {os.path.basename(n) for n in names}
        names += [f for f in sorted(glob.glob(os.path.join(ref, "*.iarray"))) if os.path.basename(f) not in have]
    return names


@pytest.mark.parametrize("path", _frames(), ids=os.path.basename)
def test_fixture_frames_known_answer(orc, path):
    data = open(path, "rb").read()
    chunks, nbytes = orc.frame_chunks(data)
    out = b"".join(orc.decompress_chunk(c, 1 << 22) for c in chunks)
    assert len(out) == nbytes
    assert hashlib.sha256(out).hexdigest() == HASHES[os.path.basename(path)]["sha256"]


def _liblz4():
    for name in ("liblz4.so.1", "/opt/conda/lib/liblz4.so.1", "/usr/lib/x86_64-linux-gnu/liblz4.so.1"):
        try:
            L = ctypes.CDLL(name)
            L.LZ4_compress_default.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
            L.LZ4_compress_HC.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
            return L
        except OSError:
            continue
    return None


def _corpus():
    rng = np.random.default_rng(7)
    yield "zeros", bytes(5000)
    yield "ramp", np.arange(30000, dtype=np.int32).tobytes()
    yield "random", rng.integers(0, 256, 20000, dtype=np.uint8).tobytes()
    yield "lowentropy", rng.integers(0, 4, 65536, dtype=np.uint8).tobytes()
    yield "runs", bytes(np.repeat(rng.integers(0, 256, 300, dtype=np.uint8), 217))
    yield "text", (b"the quick brown fox jumps over the lazy dog " * 1500)[:65000]
    yield "tiny", b"abc"
    yield "f32exp", (np.frombuffer(rng.normal(size=16384).astype(np.float32).tobytes(), np.uint8)[3::4]).tobytes()


@pytest.mark.parametrize("name,data", list(_corpus()), ids=[c[0] for c in _corpus()])
def test_lz4_against_liblz4(orc, name, data):
    L = _liblz4()
    if L is None:
        pytest.skip("liblz4 not present")
    for hc in (False, True):
        dst = ctypes.create_string_buffer(len(data) + 1024)
        n = (L.LZ4_compress_HC(data, dst, len(data), len(dst), 9) if hc
             else L.LZ4_compress_default(data, dst, len(data), len(dst)))
        assert n > 0
        assert orc.lz4_decode(dst.raw[:n], len(data)) == data


def test_lz4_rejects_malformed(orc):
    lib = orc.lib()
    out = ctypes.create_string_buffer(64)
    # offset 0 is illegal; offset beyond the output start is illegal
    assert lib.orc_lz4_decode(bytes([0x14, 0x41, 0x00, 0x00]), 4, out, 64) < 0
    assert lib.orc_lz4_decode(bytes([0x14, 0x41, 0x05, 0x00]), 4, out, 64) < 0
    # a literal run longer than the input
    assert lib.orc_lz4_decode(bytes([0xF0, 0x20]), 2, out, 64) < 0


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


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32])
@pytest.mark.parametrize("shuf", [1, 2], ids=["shuffle", "bitshuffle"])
@pytest.mark.parametrize("n", [4096, 4100, 1003, 62500])
def test_against_cblosc1(orc, dtype, shuf, n):
    b = _libblosc()
    if b is None:
        pytest.skip("c-blosc 1.21 not present")
    rng = np.random.default_rng(n)
    for kind in range(3):
        x = [rng.random(n), np.sin(np.arange(n) * 1e-3) + rng.normal(0, 1e-3, n), np.arange(n)][kind].astype(dtype)
        for bs in (0, 2048, 8000):
            b.blosc_set_blocksize(bs)
            dst = ctypes.create_string_buffer(x.nbytes + 16)
            cs = b.blosc_compress(5, shuf, x.itemsize, x.nbytes, x.ctypes.data, dst, x.nbytes + 16)
            assert cs > 0
            assert orc.decompress_chunk(dst.raw[:cs], x.nbytes) == x.tobytes()


def _lz77_structured(rng, n):
    """Bytes built of literal pieces and copies from earlier output at
    distances up to 74000 (BloscLZ's short, 13-bit and far-distance forms)
    with lengths up to 1200 (its long-length form)."""
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


@pytest.mark.parametrize("seed", range(6))
def test_blosclz_against_cblosc1(orc, seed):
    """BloscLZ streams (codec 0) written by c-blosc 1.21.0's blosc_compress
    with compressor "blosclz" (black box) decode through the oracle to the
    input: structured LZ77 data with far matches, random bytes, smooth
    floats, every clevel band, unfiltered, shuffled and bitshuffled, forced
    block sizes."""
    b = _libblosc()
    if b is None:
        pytest.skip("c-blosc 1.21 not present")
    b.blosc_set_compressor(b"blosclz")
    try:
        rng = np.random.default_rng(100 + seed)
        n = int(rng.integers(4000, 160000)) // 8 * 8
        datas = [np.frombuffer(_lz77_structured(rng, n), np.uint8),
                 rng.integers(0, 256, n, dtype=np.uint8),
                 (np.sin(np.arange(n // 8) * 1e-3) + rng.normal(0, 1e-4, n // 8))]
        seen_lz = 0
        for x in datas:
            for clevel in (1, 5, 9):
                for ts, shuf in ((1, 0), (4, 1), (8, 1), (4, 2)):
                    bs = int(rng.choice([0, 16384, 65536]))
                    b.blosc_set_blocksize(bs)
                    raw = x.tobytes()
                    dst = ctypes.create_string_buffer(len(raw) + 64)
                    cs = b.blosc_compress(clevel, shuf, ts, len(raw), raw, dst, len(raw) + 64)
                    assert cs > 0
                    chunk = dst.raw[:cs]
                    assert chunk[2] >> 5 == 0   # codec 0: BloscLZ
                    seen_lz += not (chunk[2] & 0x2)
                    assert orc.decompress_chunk(chunk, len(raw)) == raw
        assert seen_lz > 0
    finally:
        b.blosc_set_compressor(b"lz4")
        b.blosc_set_blocksize(0)


@pytest.mark.parametrize("ts,n", [(4, 64), (8, 40), (4, 62500), (2, 24), (4, 13)])
def test_bitshuffle_permutation(orc, ts, n):
    rng = np.random.default_rng(ts * 1000 + n)
    x = rng.integers(0, 256, ts * n, dtype=np.uint8)
    m = n - n % 8
    # numpy restatement: bits of byte j of element e -> plane 8j+k, position e (LSB-first packing)
    bits = np.unpackbits(x[:m * ts].reshape(m, ts), axis=1, bitorder="little")  # (m, ts*8), col = 8j+k
    want = np.packbits(bits.T.reshape(ts * 8, m), axis=1, bitorder="little").tobytes() + x[m * ts:].tobytes()
    got = orc.bitshuffle(ts, x.tobytes(), version=5)
    assert got == want
    assert orc.bitunshuffle(ts, got, version=5) == x.tobytes()
    if n % 8:
        # format version 2 copies the whole block verbatim when n % 8 != 0
        assert orc.bitshuffle(ts, x.tobytes(), version=2) == x.tobytes()


@pytest.mark.parametrize("ts", [1, 2, 4, 8])
def test_shuffle_is_byte_transpose(orc, ts):
    x = np.random.default_rng(ts).integers(0, 256, ts * 1001 + 3, dtype=np.uint8)
    n = (len(x) // ts) * ts
    want = x[:n].reshape(-1, ts).T.tobytes() + x[n:].tobytes()
    got = orc.shuffle(ts, x.tobytes())
    assert got == want
    assert orc.unshuffle(ts, got) == x.tobytes()


def _libzstd():
    for name in ("/usr/lib/x86_64-linux-gnu/libzstd.so.1", "/opt/conda/lib/libzstd.so.1"):
        try:
            z = ctypes.CDLL(name)
        except OSError:
            continue
        z.ZSTD_compress.restype = ctypes.c_size_t
        z.ZSTD_compress.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        z.ZSTD_compressBound.restype = ctypes.c_size_t
        z.ZSTD_compressBound.argtypes = [ctypes.c_size_t]
        z.ZSTD_isError.argtypes = [ctypes.c_size_t]
        return z
    return None


def _zstd(z, data, level):
    cap = z.ZSTD_compressBound(len(data))
    dst = ctypes.create_string_buffer(cap)
    n = z.ZSTD_compress(dst, cap, data, len(data), level)
    assert not z.ZSTD_isError(n)
    return dst.raw[:n]


@pytest.mark.parametrize("seed", range(12))
def test_zstd_against_libzstd(orc, seed):
    """Zstd frames written by libzstd 1.4.8's ZSTD_compress (black box) decode
    through the oracle to their input: raw, RLE and compressed blocks,
    frames of several 128 KiB blocks (repeat tables and treeless literals),
    single- and four-stream Huffman literals, predefined / RLE / described FSE
    tables, repeat offsets; levels from the fast -5 to 22."""
    z = _libzstd()
    if z is None:
        pytest.skip("libzstd not present")
    rng = np.random.default_rng(300 + seed)
    size = int(rng.integers(1, 330000))
    kind = seed % 6
    if kind == 0:
        d = rng.integers(0, 256, size, dtype=np.uint8).tobytes()
    elif kind == 1:
        d = rng.integers(0, 1 + seed, size, dtype=np.uint8).tobytes()
    elif kind == 2:
        d = (rng.standard_normal(size // 4 + 1).astype(np.float32) * 1000).tobytes()[:size]
    elif kind == 3:
        d = np.repeat(rng.integers(0, 256, size // 50 + 1, dtype=np.uint8), 50).tobytes()[:size]
    elif kind == 4:
        d = _lz77_structured(rng, size)
    else:
        d = bytes(size)
    for level in (-5, -1, 1, 3, 6, 9, 13, 17, 19, 22):
        assert orc.zstd_decode(_zstd(z, d, level), len(d)) == d, f"level {level}"


def test_zstd_rejects_malformed(orc):
    z = _libzstd()
    if z is None:
        pytest.skip("libzstd not present")
    d = _lz77_structured(np.random.default_rng(5), 50000)
    c = bytearray(_zstd(z, d, 5))
    with pytest.raises(ValueError):
        orc.zstd_decode(bytes(c[:len(c) // 2]), len(d))     # truncated
    with pytest.raises(ValueError):
        orc.zstd_decode(b"\x00" * 4 + bytes(c[4:]), len(d))  # bad magic
    with pytest.raises(ValueError):
        orc.zstd_decode(bytes(c), len(d) - 1)                 # output larger than the slot


@pytest.mark.parametrize("seed", range(4))
def test_zstd_chunks_against_cblosc1(orc, seed):
    """Zstd streams (codec 4) written by c-blosc 1.21.0's blosc_compress with
    compressor "zstd" (black box) decode through the oracle's chunk decoder:
    unfiltered, shuffled and bitshuffled, clevels 1/5/9, forced block sizes."""
    b = _libblosc()
    if b is None:
        pytest.skip("c-blosc 1.21 not present")
    b.blosc_set_compressor(b"zstd")
    try:
        rng = np.random.default_rng(400 + seed)
        n = int(rng.integers(4000, 200000)) // 8 * 8
        datas = [np.frombuffer(_lz77_structured(rng, n), np.uint8), rng.integers(0, 256, n, dtype=np.uint8),
                 (np.sin(np.arange(n // 8) * 1e-3) + rng.normal(0, 1e-4, n // 8))]
        seen = 0
        for x in datas:
            for clevel in (1, 5, 9):
                for ts, shuf in ((1, 0), (4, 1), (8, 1), (4, 2)):
                    b.blosc_set_blocksize(int(rng.choice([0, 16384, 65536])))
                    raw = x.tobytes()
                    dst = ctypes.create_string_buffer(len(raw) + 64)
                    cs = b.blosc_compress(clevel, shuf, ts, len(raw), raw, dst, len(raw) + 64)
                    assert cs > 0
                    chunk = dst.raw[:cs]
                    assert chunk[2] >> 5 == 4  # codec 4: Zstd
                    seen += not (chunk[2] & 0x2)
                    assert orc.decompress_chunk(chunk, len(raw)) == raw
        assert seen > 0
    finally:
        b.blosc_set_compressor(b"lz4")
        b.blosc_set_blocksize(0)


@pytest.mark.parametrize("seed", range(8))
def test_zlib_against_python_zlib(orc, seed):
    """zlib streams written by zlib's compress2 (Python's zlib module, black
    box) at every level decode through the oracle to their input: stored
    blocks (level 0 and random data), fixed and dynamic Huffman blocks, long
    and overlapping matches."""
    import zlib
    rng = np.random.default_rng(600 + seed)
    size = int(rng.integers(1, 200000))
    kind = seed % 4
    if kind == 0:
        d = rng.integers(0, 256, size, dtype=np.uint8).tobytes()
    elif kind == 1:
        d = rng.integers(0, 1 + seed, size, dtype=np.uint8).tobytes()
    elif kind == 2:
        d = _lz77_structured(rng, size)
    else:
        d = np.repeat(rng.integers(0, 256, size // 300 + 1, dtype=np.uint8), 300).tobytes()[:size]
    for level in range(10):
        assert orc.zlib_decode(zlib.compress(d, level), len(d)) == d, f"level {level}"
    with pytest.raises(ValueError):
        orc.zlib_decode(zlib.compress(d, 6)[: max(3, size // 100)], len(d))


@pytest.mark.parametrize("seed", range(3))
def test_zlib_chunks_against_cblosc1(orc, seed):
    """zlib streams (codec 3) written by c-blosc 1.21.0's blosc_compress with
    compressor "zlib" (black box) decode through the oracle's chunk decoder."""
    b = _libblosc()
    if b is None:
        pytest.skip("c-blosc 1.21 not present")
    b.blosc_set_compressor(b"zlib")
    try:
        rng = np.random.default_rng(700 + seed)
        n = int(rng.integers(4000, 120000)) // 8 * 8
        datas = [np.frombuffer(_lz77_structured(rng, n), np.uint8), rng.integers(0, 256, n, dtype=np.uint8),
                 (np.sin(np.arange(n // 8) * 1e-3) + rng.normal(0, 1e-4, n // 8))]
        for x in datas:
            for clevel in (1, 5, 9):
                for ts, shuf in ((1, 0), (4, 1), (8, 2)):
                    b.blosc_set_blocksize(int(rng.choice([0, 16384])))
                    raw = x.tobytes()
                    dst = ctypes.create_string_buffer(len(raw) + 64)
                    cs = b.blosc_compress(clevel, shuf, ts, len(raw), raw, dst, len(raw) + 64)
                    assert cs > 0
                    chunk = dst.raw[:cs]
                    assert chunk[2] >> 5 == 3  # codec 3: zlib
                    assert orc.decompress_chunk(chunk, len(raw)) == raw
    finally:
        b.blosc_set_compressor(b"lz4")
        b.blosc_set_blocksize(0)
