"""CPU tests of the product's host side (no GPU calls).

* libiarray.so / libiamx.so load and export every symbol include/*.h declares;
* the host write stage (tiling + shuffle/bitshuffle + LZ4 encode + Blosc2
  chunk assembly) round-trips through the oracle decoder AND through liblz4
  (a black box), for every dtype, filter, split mode and ragged tiling;
* the validation and error codes of iarray_linalg_matmul / opt_gemm / opt_gemv
  (src/linalg/matmul.c:31-55, src/linalg/iarray_gemm.c:327-364,
  src/linalg/iarray_gemv.c:261-286) are returned before any device work;
* entry points that need the GPU fail loudly (IARRAY_ERR_NO_DEVICE) when no
  GPU is present -- there is no CPU fallback.
"""
import ctypes
import glob
import os
import re
import struct

import numpy as np
import pytest

import pyiarray as ia

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header, macro):
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = "\n".join(l for l in text.splitlines() if not l.lstrip().startswith("#"))
    if macro == "INA_API":
        return set(re.findall(r"INA_API\([^)]*\)\s*(\w+)\s*\(", text))
    return set(re.findall(r"IAMX_API\s+[\w\s\*]+?\b(\w+)\s*\(", text))


def test_libraries_export_every_declared_symbol():
    L = ia.lib()
    H = ctypes.CDLL(ia.HIP_LIB_PATH)
    api = _declared("iarray.h", "INA_API") | _declared("iarray_amd.h", "INA_API")
    hip = _declared("iamx.h", "IAMX_API")
    assert len(api) > 30 and len(hip) > 20
    for name in sorted(api):
        assert hasattr(L, name), name
    for name in sorted(hip):
        assert hasattr(H, name), name


def _has_gpu():
    h = ctypes.CDLL(ia.HIP_LIB_PATH)
    return h.iamx_init() == 0


CASES = [
    # dtype, shape, chunks, blocks, filters, clevel, splitmode
    (np.float64, (50, 37), (20, 15), (7, 6), ia.COMP_SHUFFLE, 5, ia.AUTO_SPLIT),
    (np.float32, (100, 170), (27, 30), (12, 7), ia.COMP_SHUFFLE, 5, ia.AUTO_SPLIT),
    (np.float32, (64, 64), (64, 64), (16, 16), ia.COMP_BITSHUFFLE, 5, ia.AUTO_SPLIT),
    (np.float32, (61, 47), (30, 20), (9, 13), ia.COMP_BITSHUFFLE, 9, ia.AUTO_SPLIT),
    (np.float64, (33, 65), (33, 65), (11, 13), ia.COMP_SHUFFLE, 1, ia.NEVER_SPLIT),
    (np.float64, (40, 40), (20, 20), (10, 10), 0, 5, ia.ALWAYS_SPLIT),
    (np.float32, (40, 40), (20, 20), (10, 10), ia.COMP_SHUFFLE, 0, ia.AUTO_SPLIT),
    (np.int32, (1000,), (300,), (64,), ia.COMP_SHUFFLE, 5, ia.AUTO_SPLIT),
    (np.float64, (777,), (256,), (100,), ia.COMP_BITSHUFFLE, 3, ia.AUTO_SPLIT),
]


def _data(dtype, shape, kind):
    rng = np.random.default_rng(abs(hash((shape, kind))) % (2 ** 32))
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
def test_write_stage_roundtrips_through_oracle_and_liblz4(orc, case, kind):
    dtype, shape, chunks, blocks, filt, clevel, split = case
    ctx = ia.Context(clevel=clevel, filters=filt, splitmode=split)
    x = _data(dtype, shape, kind)
    c = ia.Container.from_numpy(ctx, x, chunks, blocks)
    chunk_bytes = c.chunks()
    assert len(chunk_bytes) == int(np.prod([-(-s // k) for s, k in zip(shape, chunks)]))
    oa = orc.Array(chunk_bytes, shape, chunks, blocks, x.itemsize)
    for lz4 in (0, 1):
        if orc.set_lz4(lz4) != lz4:
            continue
        got = np.zeros(shape, dtype=dtype)
        assert orc.lib().orc_array_to_dense(ctypes.byref(oa.s), got.ctypes.data, 2) == 0
        assert got.tobytes() == x.tobytes()
    orc.set_lz4(0)
    nb, cb = c.info()
    assert nb == x.nbytes and cb == sum(len(b) for b in chunk_bytes)
    if kind == "zeros" and clevel > 0:
        assert cb < nb / 3


def test_lz4_encoder_respects_format_end_rules():
    """liblz4's safe decoder rejects blocks that break the end-of-block rules,
    so decoding every stream we write with it (above) pins them; here a direct
    probe of tiny and long streams through the write stage."""
    L = ia.lib()
    for n in (1, 2, 12, 13, 14, 100, 4096, 65536):
        ctx = ia.Context(clevel=9, filters=0, splitmode=ia.NEVER_SPLIT)
        x = (np.arange(n) % 3).astype(np.uint8)
        c = ia.Container.from_numpy(ctx, x, (n,), (n,))
        assert len(c.chunks()) == 1
    assert L is not None


def _ctx():
    return ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)


def _err(fn):
    with pytest.raises(ia.IArrayError) as e:
        fn()
    return e.value.rc


def test_matmul_validation_errors():
    ctx = _ctx()
    a = ia.Container.arange(ctx, np.float64, (30, 20), (10, 10), (5, 5))
    b = ia.Container.arange(ctx, np.float64, (20, 12), (10, 6), (5, 3))
    bf = ia.Container.arange(ctx, np.float32, (20, 12), (10, 6), (5, 3))
    v3 = ia.Container.arange(ctx, np.float64, (7,), (7,), (7,))
    bad = ia.Container.arange(ctx, np.float64, (21, 12), (10, 6), (5, 3))
    assert _err(lambda: ia.matmul(v3, b, (7,), (7,))) == ia.ERR_INVALID_NDIM
    assert _err(lambda: ia.matmul(a, bad, (10, 6), (5, 3))) == ia.ERR_INVALID_SHAPE
    assert _err(lambda: ia.matmul(a, bf, (10, 6), (5, 3))) == ia.ERR_INVALID_DTYPE
    # output chunkshape larger than the output shape (matmul.c:50-55)
    assert _err(lambda: ia.matmul(a, b, (31, 6), (5, 3))) == ia.ERR_INVALID_CHUNKSHAPE


def test_opt_partition_errors():
    ctx = _ctx()
    a = ia.Container.arange(ctx, np.float64, (40, 30), (20, 10), (10, 5))
    b = ia.Container.arange(ctx, np.float64, (30, 24), (10, 12), (5, 6))
    x = ia.Container.arange(ctx, np.float64, (30,), (10,), (5,))
    assert _err(lambda: ia.opt_gemm(a, b, (10, 12), (10, 6))) == ia.ERR_INVALID_CHUNKSHAPE
    assert _err(lambda: ia.opt_gemm(a, b, (20, 6), (10, 6))) == ia.ERR_INVALID_CHUNKSHAPE
    assert _err(lambda: ia.opt_gemm(a, b, (20, 12), (5, 6))) == ia.ERR_INVALID_BLOCKSHAPE
    assert _err(lambda: ia.opt_gemm(a, b, (20, 12), (10, 3))) == ia.ERR_INVALID_BLOCKSHAPE
    b2 = ia.Container.arange(ctx, np.float64, (30, 24), (15, 12), (5, 6))
    assert _err(lambda: ia.opt_gemm(a, b2, (20, 12), (10, 6))) == ia.ERR_INVALID_CHUNKSHAPE
    assert _err(lambda: ia.opt_gemv(a, b, (20,), (10,))) == ia.ERR_INVALID_NDIM
    assert _err(lambda: ia.opt_gemv(a, x, (10,), (10,))) == ia.ERR_INVALID_CHUNKSHAPE
    assert _err(lambda: ia.opt_gemv(a, x, (20,), (5,))) == ia.ERR_INVALID_BLOCKSHAPE
    x2 = ia.Container.arange(ctx, np.float64, (30,), (10,), (10,))
    assert _err(lambda: ia.opt_gemv(a, x2, (20,), (10,))) == ia.ERR_INVALID_BLOCKSHAPE


def test_opt_gemm_a_partition_errors():
    """The checks of src/linalg/iarray_gemm_a.c:198-262, one at a time in their
    order; a valid partition gets past them (and then needs the GPU)."""
    ctx = _ctx()
    f = np.float64
    A = ia.Container.arange(ctx, f, (40, 30), (20, 30), (20, 10))
    B = ia.Container.arange(ctx, f, (30, 24), (30, 12), (10, 12))
    ok_c, ok_b = (20, 24), (20, 12)
    v1 = ia.Container.arange(ctx, f, (30,), (30,), (10,))
    assert _err(lambda: ia.opt_gemm_a(v1, B, ok_c, ok_b)) == ia.ERR_INVALID_NDIM
    assert _err(lambda: ia.opt_gemm_a(A, v1, ok_c, ok_b)) == ia.ERR_INVALID_NDIM
    B31 = ia.Container.arange(ctx, f, (31, 24), (31, 12), (10, 12))
    assert _err(lambda: ia.opt_gemm_a(A, B31, ok_c, ok_b)) == ia.ERR_INVALID_SHAPE
    A_c = ia.Container.arange(ctx, f, (40, 30), (20, 10), (20, 10))  # a chunk narrower than a
    assert _err(lambda: ia.opt_gemm_a(A_c, B, ok_c, ok_b)) == ia.ERR_INVALID_CHUNKSHAPE
    B_c = ia.Container.arange(ctx, f, (30, 24), (10, 12), (10, 12))  # b chunk shorter than b
    assert _err(lambda: ia.opt_gemm_a(A, B_c, ok_c, ok_b)) == ia.ERR_INVALID_CHUNKSHAPE
    B_k = ia.Container.arange(ctx, f, (30, 24), (40, 12), (10, 12))  # b chunk rows != a chunk cols
    assert _err(lambda: ia.opt_gemm_a(A, B_k, ok_c, ok_b)) == ia.ERR_INVALID_CHUNKSHAPE
    assert _err(lambda: ia.opt_gemm_a(A, B, (10, 24), ok_b)) == ia.ERR_INVALID_CHUNKSHAPE  # c chunk rows
    assert _err(lambda: ia.opt_gemm_a(A, B, (20, 12), ok_b)) == ia.ERR_INVALID_CHUNKSHAPE  # c chunk too narrow
    assert _err(lambda: ia.opt_gemm_a(A, B, ok_c, (10, 12))) == ia.ERR_INVALID_BLOCKSHAPE
    assert _err(lambda: ia.opt_gemm_a(A, B, ok_c, (20, 6))) == ia.ERR_INVALID_BLOCKSHAPE
    A_b = ia.Container.arange(ctx, f, (40, 30), (20, 30), (10, 10))  # a block rows != a chunk rows
    assert _err(lambda: ia.opt_gemm_a(A_b, B, ok_c, ok_b)) == ia.ERR_INVALID_BLOCKSHAPE
    B_b = ia.Container.arange(ctx, f, (30, 24), (30, 12), (10, 6))  # b block cols != b chunk cols
    assert _err(lambda: ia.opt_gemm_a(A, B_b, ok_c, ok_b)) == ia.ERR_INVALID_BLOCKSHAPE
    A_k = ia.Container.arange(ctx, f, (40, 30), (20, 30), (20, 15))  # a block cols != b block rows
    assert _err(lambda: ia.opt_gemm_a(A_k, B, ok_c, ok_b)) == ia.ERR_INVALID_BLOCKSHAPE
    if not _has_gpu():
        assert _err(lambda: ia.opt_gemm_a(A, B, ok_c, ok_b)) == ia.ERR_NO_DEVICE


def test_opt_gemm_b_partition_errors():
    """The checks of src/linalg/iarray_gemm_b.c:198-263 in their order."""
    ctx = _ctx()
    f = np.float64
    A = ia.Container.arange(ctx, f, (40, 30), (10, 30), (10, 10))
    B = ia.Container.arange(ctx, f, (30, 24), (30, 12), (10, 12))
    ok_c, ok_b = (40, 12), (10, 12)
    A_c = ia.Container.arange(ctx, f, (40, 30), (10, 15), (10, 15))
    assert _err(lambda: ia.opt_gemm_b(A_c, B, ok_c, ok_b)) == ia.ERR_INVALID_CHUNKSHAPE
    assert _err(lambda: ia.opt_gemm_b(A, B, (20, 12), ok_b)) == ia.ERR_INVALID_CHUNKSHAPE  # c chunk shorter than a
    assert _err(lambda: ia.opt_gemm_b(A, B, (40, 24), ok_b)) == ia.ERR_INVALID_CHUNKSHAPE  # c chunk cols != b's
    assert _err(lambda: ia.opt_gemm_b(A, B, ok_c, (20, 12))) == ia.ERR_INVALID_BLOCKSHAPE
    assert _err(lambda: ia.opt_gemm_b(A, B, ok_c, (10, 6))) == ia.ERR_INVALID_BLOCKSHAPE
    A_b = ia.Container.arange(ctx, f, (40, 30), (10, 30), (5, 10))
    assert _err(lambda: ia.opt_gemm_b(A_b, B, ok_c, ok_b)) == ia.ERR_INVALID_BLOCKSHAPE
    if not _has_gpu():
        assert _err(lambda: ia.opt_gemm_b(A, B, ok_c, ok_b)) == ia.ERR_NO_DEVICE


# (variant, M, K, N, A chunk/block, B chunk/block, C chunk/block): gemm_a wants A
# in one chunk column with block rows = chunk rows, B in one chunk row with
# block columns = chunk columns; gemm_b the mirror of the C tiling.  Padded
# edges on every axis.
AB_CASES = [
    ("a", 70, 45, 50, ((30, 45), (30, 20)), ((45, 20), (20, 20)), ((30, 50), (30, 20))),
    ("a", 64, 64, 64, ((32, 64), (32, 16)), ((64, 32), (16, 32)), ((32, 64), (32, 32))),
    ("b", 70, 45, 50, ((30, 45), (30, 20)), ((45, 20), (20, 20)), ((70, 20), (30, 20))),
    ("b", 64, 64, 96, ((16, 64), (16, 32)), ((64, 48), (32, 48)), ((64, 48), (16, 48))),
]


@pytest.mark.parametrize("case", AB_CASES, ids=[f"gemm_{c[0]}-{c[1]}x{c[2]}x{c[3]}" for c in AB_CASES])
def test_oracle_opt_gemm_ab_exact(orc, case):
    """The oracle's restatement of opt_gemm_a / opt_gemm_b on integer-valued
    operands (exact in any summation order) equals the dense product, with
    a zero chunk and zero blocks planted in the operand each variant fetches
    block by block; a transposed operand would not."""
    var, M, K, N, (acs, abs_), (bcs, bbs), (ccs, cbs) = case
    ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
    rng = np.random.default_rng(M + 3 * K + 7 * N)
    a = rng.integers(-8, 9, (M, K)).astype(np.float64)
    b = rng.integers(-8, 9, (K, N)).astype(np.float64)
    if var == "a":
        b[:, :bcs[1]] = 0                      # B's first chunk is all zeros
        b[bbs[0]:2 * bbs[0], bcs[1]:] = 0      # a zero block row inside the others
    else:
        a[:acs[0], :] = 0
        a[acs[0]:, :abs_[1]] = 0
    ca = ia.Container.from_numpy(ctx, a, acs, abs_)
    cb = ia.Container.from_numpy(ctx, b, bcs, bbs)
    oa = orc.Array(ca.chunks(), a.shape, acs, abs_, 8)
    ob = orc.Array(cb.chunks(), b.shape, bcs, bbs, 8)
    fn = orc.opt_gemm_a if var == "a" else orc.opt_gemm_b
    got = fn(oa, ob, ccs, cbs)
    assert np.array_equal(got, a @ b)
    if M == K:
        assert not np.array_equal(got, a.T @ b)


FRAMES = sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "frames", "*.iarray")))


@pytest.mark.parametrize("path", FRAMES, ids=os.path.basename)
def test_frame_load_reads_fixture(orc, path):
    """iarray_container_load on the reference's frame fixtures: dtype and shape
    from the metalayers (tests/generator.py:23-38 writes 1e6 elements), and
    the same chunk bytes the oracle's independent frame parser finds."""
    ctx = _ctx()
    c = ia.Container.load(ctx, path)
    assert c.shape == (1000000,)
    assert c.dtype == (np.int32 if "int32" in path else np.float32)
    want, nbytes = orc.frame_chunks(open(path, "rb").read())
    assert c.chunks() == want
    nb, cb = c.info()
    assert nb == nbytes and cb == sum(len(w) for w in want)
    assert ia.Container.load(ctx, path, lazy=True).chunks() == want


def test_frame_load_errors(tmp_path):
    ctx = _ctx()
    assert _err(lambda: ia.Container.load(ctx, str(tmp_path / "missing.iarray"))) == ia.ERR_FILE_OPEN
    bad = tmp_path / "bad.iarray"
    bad.write_bytes(b"not a frame" * 20)
    assert _err(lambda: ia.Container.load(ctx, str(bad))) == ia.ERR_BLOSC_FAILED
    data = open(FRAMES[0], "rb").read()
    short = tmp_path / "short.iarray"
    short.write_bytes(data[: len(data) // 2])  # the offsets chunk is gone
    assert _err(lambda: ia.Container.load(ctx, str(short))) == ia.ERR_BLOSC_FAILED


def test_frame_load_survives_damage(tmp_path):
    """Truncated frames and frames with flipped header or offset bytes load
    or fail with an error; the reader bounds-checks every field."""
    ctx = _ctx()
    data = bytearray(open(FRAMES[0], "rb").read())
    hl = int.from_bytes(data[11:15], "big")
    rng = np.random.default_rng(3)
    f = tmp_path / "damaged.iarray"
    outcomes = set()
    for i in range(150):
        d = bytearray(data)
        if i < 30:
            d = d[: int(rng.integers(1, len(d)))]
        else:
            lo, hi = (0, hl) if i < 110 else (len(d) - 120, len(d))
            for _ in range(int(rng.integers(1, 4))):
                d[int(rng.integers(lo, hi))] = int(rng.integers(0, 256))
        f.write_bytes(bytes(d))
        try:
            c = ia.Container.load(ctx, str(f))
            outcomes.add("loaded")
            del c
        except ia.IArrayError:
            outcomes.add("error")
    assert "error" in outcomes


def _offsets_pos(frame: bytes):
    """Byte position of the int64 chunk offsets in a contiguous frame: header
    size (int32 BE at byte 11) + data cbytes (the sixth msgpack field, an
    int64 BE after its 0xd3 marker), then the offsets chunk's 32-byte header."""
    hsize = int.from_bytes(frame[11:15], "big")
    # fields: magic (str8), hsize (int32), fsize (uint64), flags, nbytes, cbytes
    p = frame.index(b"\xd3", 16 + 8 + 5)      # nbytes marker after flags
    cbytes = int.from_bytes(frame[p + 10:p + 18], "big", signed=True)
    assert frame[p + 9] == 0xd3
    return hsize + cbytes + 32, hsize, cbytes


def test_frame_load_rejects_overflowing_offsets(tmp_path):
    """A frame whose data cbytes or chunk offset is near INT64_MAX is refused
    with an error: the reader bounds each field before adding it to a
    position."""
    ctx = _ctx()
    data = bytearray(open(FRAMES[0], "rb").read())
    opos, hsize, cbytes = _offsets_pos(bytes(data))
    f = tmp_path / "o.iarray"
    for off in (2 ** 63 - 1 - hsize, 2 ** 63 - 40, 2 ** 62):
        d = bytearray(data)
        d[opos:opos + 8] = int(off).to_bytes(8, "little")
        f.write_bytes(bytes(d))
        assert _err(lambda: ia.Container.load(ctx, str(f))) == ia.ERR_BLOSC_FAILED
    p = data.index(b"\xd3", 16 + 8 + 5)
    for cb in (2 ** 63 - 1, 2 ** 63 - 1 - hsize + 8, 2 ** 63 - 20):
        d = bytearray(data)
        d[p + 10:p + 18] = int(cb).to_bytes(8, "big")
        f.write_bytes(bytes(d))
        assert _err(lambda: ia.Container.load(ctx, str(f))) == ia.ERR_BLOSC_FAILED


@pytest.mark.parametrize("kind,fill", [(1, 0), (2, None), (4, 0)])
def test_frame_load_special_value_chunks(orc, tmp_path, kind, fill):
    """A chunk whose frame offset is negative holds no bytes: bit 63 is set
    and the top byte's low bits name the value (1 zeros, 2 NaNs, 4
    uninitialized, read as zeros).  The container gets a 32-byte special
    chunk that the oracle decodes to that value, the other chunks as stored."""
    ctx = _ctx()
    x = np.arange(3 * 40 * 50, dtype=np.float64).reshape(120, 50) + 1.0
    c = ia.Container.from_numpy(ctx, x, (40, 50), (20, 25))
    f = tmp_path / "s.iarray"
    c.save(str(f))
    data = bytearray(f.read_bytes())
    opos, _, _ = _offsets_pos(bytes(data))
    data[opos + 8:opos + 16] = ((1 << 63) | (kind << 56)).to_bytes(8, "little")
    f.write_bytes(bytes(data))
    d = ia.Container.load(ctx, str(f))
    ch = d.chunks()
    assert len(ch[1]) == 32 and (ch[1][31] >> 4) & 7 == kind
    assert ch[0] == c.chunks()[0] and ch[2] == c.chunks()[2]
    got = orc.Array(ch, x.shape, (40, 50), (20, 25), 8).to_dense()
    want = x.copy()
    want[40:80] = np.nan if fill is None else fill
    assert np.array_equal(got, want, equal_nan=True)
    for bad in (3, 5, 0x7f):
        data[opos + 8:opos + 16] = ((1 << 63) | (bad << 56)).to_bytes(8, "little")
        f.write_bytes(bytes(data))
        assert _err(lambda: ia.Container.load(ctx, str(f))) == ia.ERR_BLOSC_FAILED


@pytest.mark.parametrize("path", FRAMES, ids=os.path.basename)
def test_frame_save_rewrites_fixture(path, tmp_path):
    """A frame the reference wrote, loaded and saved again, is the same file
    byte for byte (header, metalayers, chunks, offsets chunk, trailer); so is
    the frame of a container built from the same chunks with the writer's
    settings (LZ4 clevel 5, shuffle, 8 threads), which pins the header fields
    a new container takes from its config."""
    data = open(path, "rb").read()
    c = ia.Container.load(_ctx(), path)
    out = tmp_path / "saved.iarray"
    c.save(str(out))
    assert out.read_bytes() == data
    assert c.to_cframe() == data
    ctx8 = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE, nthreads=8)
    fresh = ia.Container.from_chunks(ctx8, c.dtype, c.shape, (100000,), (10000,), c.chunks())
    assert fresh.to_cframe() == data
    back = ia.Container.from_cframe(_ctx(), data)
    assert back.shape == c.shape and back.dtype == c.dtype and back.chunks() == c.chunks()


def test_frame_save_two_dimensional(tmp_path):
    """A 2-D container goes through a frame and back with its shape, tiling
    and chunks; saving the loaded copy writes the same bytes again."""
    ctx = _ctx()
    src = ia.Container.load(ctx, FRAMES[0])
    c = ia.Container.from_chunks(ctx, src.dtype, (10, 100000), (1, 100000), (1, 10000), src.chunks())
    f1 = tmp_path / "a.iarray"
    c.save(str(f1))
    d = ia.Container.load(ctx, str(f1))
    assert d.shape == (10, 100000) and d.dtype == src.dtype and d.chunks() == src.chunks()
    f2 = tmp_path / "b.iarray"
    d.save(str(f2))
    assert f2.read_bytes() == f1.read_bytes()
    # the caterva metalayer: [version, ndim, shape (int64), chunkshape, blockshape (int32)]
    raw = f1.read_bytes()
    cat = bytes([0x95, 0, 2, 0x92, 0xd3]) + (10).to_bytes(8, "big") + b"\xd3" + (100000).to_bytes(8, "big")
    cat += bytes([0x92, 0xd2]) + (1).to_bytes(4, "big") + b"\xd2" + (100000).to_bytes(4, "big")
    cat += bytes([0x92, 0xd2]) + (1).to_bytes(4, "big") + b"\xd2" + (10000).to_bytes(4, "big")
    assert raw.find(cat) > 0


def test_frame_save_errors_and_remove(tmp_path):
    """iarray_container_save refuses views and opened containers
    (src/iarray_container.c:57-66); iarray_container_remove deletes files and
    frame directories and accepts a path that does not exist."""
    ctx = _ctx()
    src = ia.Container.load(ctx, FRAMES[0])
    c = ia.Container.from_chunks(ctx, src.dtype, (10, 100000), (1, 100000), (1, 10000), src.chunks())
    t = c.transpose()
    assert _err(lambda: t.save(str(tmp_path / "t.iarray"))) == ia.ERR_INVALID_STORAGE
    assert _err(lambda: t.to_cframe()) == ia.ERR_INVALID_STORAGE
    opened = ia.Container.load(ctx, FRAMES[0], lazy=True)
    assert _err(lambda: opened.save(str(tmp_path / "o.iarray"))) == ia.ERR_INVALID_STORAGE
    f = tmp_path / "s.iarray"
    src.save(str(f))
    assert f.exists()
    ia.remove(str(f))
    assert not f.exists()
    ia.remove(str(f))
    d = tmp_path / "dir.iarr"
    d.mkdir()
    (d / "chunks.b2frame").write_bytes(b"x")
    (d / "00000000.chunk").write_bytes(b"y")
    ia.remove(str(d))
    assert not d.exists()
    assert _err(lambda: src.save(str(tmp_path / "no" / "such" / "dir.iarray"))) == ia.ERR_BLOSC_FAILED


def test_get_slice_views_and_errors():
    """iarray_get_slice (src/iarray_container.c:364-455): a view's shape,
    negative bounds, slices of slices, and the argument errors, all before any
    decode (no GPU needed)."""
    ctx = _ctx()
    src = ia.Container.load(ctx, FRAMES[0])
    c = ia.Container.from_chunks(ctx, src.dtype, (10, 100000), (1, 100000), (1, 10000), src.chunks())
    v = c.get_slice((2, 100), (-1, 5000))
    assert v.shape == (7, 4900)
    is_view = ctypes.c_bool()
    ia._check(ia.lib().iarray_is_view(ctx.ptr, v.ptr, ctypes.byref(is_view)))
    assert is_view.value
    w = v.get_slice((1, 10), (3, 20))
    assert w.shape == (2, 10)
    assert _err(lambda: c.get_slice((5, 0), (2, 10))) == ia.ERR_INVALID_ARGUMENT
    assert _err(lambda: c.get_slice((0, 0), (11, 10))) == ia.ERR_INVALID_ARGUMENT
    assert _err(lambda: c.get_slice((0, 0), (4, 100), view=False, chunks=(5, 10), blocks=(5, 10))) == \
        ia.ERR_INVALID_CHUNKSHAPE
    t = c.transpose().get_slice((100, 2), (-5, 9))
    assert t.shape == (99895, 7)
    assert _err(lambda: v.transpose()) == ia.ERR_NOT_SUPPORTED  # transposing views (src/linalg/transpose.c:23-26)
    one = src.get_slice((999000,), (-1,))
    assert one.shape == (999,)


def test_constructor_errors():
    ctx = _ctx()
    small = np.zeros(10)
    d = ia._dtshape(np.float64, (4, 4))
    s = ia._storage((2, 2), (2, 2))
    p = ctypes.c_void_p()
    rc = ia.lib().iarray_from_buffer(ctx.ptr, ctypes.byref(d), small.ctypes.data_as(ctypes.c_void_p), small.nbytes,
                                     ctypes.byref(s), ctypes.byref(p))
    rc %= 1 << 64
    assert rc & ia.ERR_BIT and rc - ia.ERR_BIT == ia.ERR_TOO_SMALL_BUFFER
    assert _err(lambda: ia.Container.arange(ctx, np.float64, (4, 4, 4), (2, 2, 2), (1, 1, 1))) == ia.ERR_INVALID_NDIM
    assert _err(lambda: ia.Container.arange(ctx, np.float64, (8, 8), (4, 4), (5, 2))) == ia.ERR_INVALID_BLOCKSHAPE
    s = ia._storage((4, 4), (2, 2))
    s.urlpath = b""
    rc = ia.lib().iarray_empty(ctx.ptr, ctypes.byref(d), ctypes.byref(s), ctypes.byref(p))
    assert rc % (1 << 64) - ia.ERR_BIT == ia.ERR_INVALID_STORAGE


@pytest.mark.parametrize("contiguous", [True, False], ids=["frame-file", "sparse-dir"])
def test_storage_urlpath_persists(orc, tmp_path, contiguous):
    """A container created with storage.urlpath (src/iarray_constructor.c:
    16-45) is written through to a contiguous frame file, or to a sparse frame
    directory (chunks.b2frame index + one %08X.chunk file per chunk) when
    contiguous is false; loading the path gives the same shape, dtype and
    chunks, the oracle decodes them to the data, it counts as stored on disk
    (save refuses, as in src/iarray_container.c:57-66), and creating at the
    same path again replaces it."""
    ctx = _ctx()
    rng = np.random.default_rng(4)
    x = rng.integers(-50, 50, (130, 70)).astype(np.float64)
    path = str(tmp_path / "x.iarr")
    c = ia.Container.from_numpy(ctx, x, (40, 30), (20, 15), urlpath=path, contiguous=contiguous)
    if contiguous:
        assert os.path.isfile(path)
    else:
        names = sorted(os.listdir(path))
        assert names[-1] == "chunks.b2frame" and len(names) == 1 + 4 * 3
        assert names[0] == "00000000.chunk" and "0000000B.chunk" in names
    d = ia.Container.load(ctx, path)
    assert d.shape == x.shape and d.dtype == x.dtype and d.chunks() == c.chunks()
    assert np.array_equal(orc.Array(d.chunks(), x.shape, (40, 30), (20, 15), 8).to_dense(), x)
    assert _err(lambda: c.save(str(tmp_path / "copy.iarr"))) == ia.ERR_INVALID_STORAGE
    # the other constructors and a second container at the same path
    r = ia.Container.arange(ctx, np.float32, (50,), (20,), (5,), urlpath=path, contiguous=contiguous)
    e = ia.Container.load(ctx, path)
    assert e.shape == (50,) and e.dtype == np.float32 and e.chunks() == r.chunks()
    assert np.array_equal(orc.Array(e.chunks(), (50,), (20,), (5,), 4).to_dense(), np.arange(50, dtype=np.float32))
    z = ia.Container.zeros(ctx, np.float64, (30, 30), (10, 10), (5, 5), urlpath=path, contiguous=contiguous)
    zz = ia.Container.load(ctx, path)
    assert zz.chunks() == z.chunks() and not orc.Array(zz.chunks(), (30, 30), (10, 10), (5, 5), 8).to_dense().any()
    ia.Container.empty(ctx, np.float64, (30, 30), (10, 10), (5, 5), urlpath=path, contiguous=contiguous)
    assert ia.Container.load(ctx, path).shape == (30, 30)
    ia.remove(path)
    assert not os.path.exists(path)


def test_arange_linspace_values(orc):
    """iarray_arange: start + flat_index*step (arange.c:40); linspace step =
    (stop-start)/(n-1) (linspace.c:117); checked through the oracle decoder."""
    ctx = _ctx()
    a = ia.Container.arange(ctx, np.float32, (13, 9), (5, 4), (2, 3), start=3.0, step=0.5)
    oa = orc.Array(a.chunks(), (13, 9), (5, 4), (2, 3), 4)
    assert np.array_equal(oa.to_dense(), (3.0 + 0.5 * np.arange(117)).astype(np.float32).reshape(13, 9))
    b = ia.Container.linspace(ctx, np.float64, (100,), (30,), (7,), 0.0, 10.0)
    ob = orc.Array(b.chunks(), (100,), (30,), (7,), 8)
    step = 10.0 / 99
    assert np.array_equal(ob.to_dense(), 0.0 + np.arange(100) * step)


@pytest.mark.skipif(_has_gpu(), reason="checks the no-GPU behaviour")
def test_device_paths_fail_loudly_without_gpu():
    ctx = _ctx()
    a = ia.Container.arange(ctx, np.float64, (20, 20), (10, 10), (5, 5))
    assert _err(lambda: a.to_numpy()) == ia.ERR_NO_DEVICE
    assert _err(lambda: ia.matmul(a, a, (10, 10), (5, 5))) == ia.ERR_NO_DEVICE


@pytest.mark.parametrize("path", FRAMES[:3], ids=os.path.basename)
def test_open_reads_chunks_lazily(tmp_path, path):
    """iarray_container_open maps the frame file and locates a chunk only when
    it is first needed (the reference's Caterva reads an opened frame's chunks
    lazily, src/iarray_container.c:79-218): the chunks and sizes equal those
    of iarray_container_load, and a chunk whose header is damaged is reported
    on first use, not at open."""
    ctx = _ctx()
    full = ia.Container.load(ctx, path)
    lazy = ia.Container.load(ctx, path, lazy=True)
    assert lazy.shape == full.shape and lazy.dtype == full.dtype
    assert lazy.info() == full.info()
    assert lazy.chunks() == full.chunks()
    # damage the size field of the last chunk of a copy: open succeeds, the
    # first reader of the chunks fails
    raw = bytearray(open(path, "rb").read())
    chunks = full.chunks()
    last = bytes(chunks[-1])
    pos = raw.rfind(last)
    assert pos > 0
    raw[pos + 12:pos + 16] = struct.pack("<i", len(raw))  # ends past the data section
    bad = tmp_path / "damaged.iarray"
    bad.write_bytes(bytes(raw))
    opened = ia.Container.load(ctx, str(bad), lazy=True)
    assert _err(lambda: opened.chunks()) == ia.ERR_BLOSC_FAILED
    assert _err(lambda: opened.info()) == ia.ERR_BLOSC_FAILED
    assert _err(lambda: ia.Container.load(ctx, str(bad))) != 0
