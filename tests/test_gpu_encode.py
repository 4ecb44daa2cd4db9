"""GPU C write stage: LZ4 encode + Blosc2 chunk assembly on the device.

Bars:
  * every chunk the GPU writes decodes, through the oracle's decoder
    (oracle/blosc_chunk.c, oracle/lz4_block.c), to exactly the values of the
    product, and to the same values as the host-encoded chunks of the same
    product (IARRAY_GPU_ENCODE=0);
  * the chunk layout is the host writer's: header fields, bstarts, stream
    csize semantics (0 = zeros, neblock = raw, otherwise LZ4 and smaller);
  * every LZ4 stream keeps the format's end rules (last 5 bytes literals, last
    match starting at least 12 bytes before the end, offsets within the
    output so far and at most 65535);
  * the encoder is deterministic.
The encoded bytes themselves are this library's own parse; c-blosc2's bytes
are not pinned by any reference test (parity unpinned for the encoder's byte
output, pinned for what it decodes to).
"""
import struct

import numpy as np
import pytest

import pyiarray as ia

pytestmark = pytest.mark.gpu


def _lz4_rules(s, n):
    """Parse one LZ4 block; check the end rules; return the decoded length."""
    ip, op = 0, 0
    while True:
        t = s[ip]
        ip += 1
        lit = t >> 4
        if lit == 15:
            while True:
                b = s[ip]
                ip += 1
                lit += b
                if b != 255:
                    break
        ip += lit
        op += lit
        if ip == len(s):
            assert n < 13 or lit >= 5, "last sequence must hold the last 5 bytes"
            return op
        off = s[ip] | (s[ip + 1] << 8)
        ip += 2
        ml = t & 15
        if ml == 15:
            while True:
                b = s[ip]
                ip += 1
                ml += b
                if b != 255:
                    break
        ml += 4
        assert 1 <= off <= op and off <= 65535
        assert op < n - 12, f"match at {op} starts within the last 12 bytes of {n}"
        op += ml
        assert op <= n - 5, "match ends within the last 5 bytes"


def _check_chunk(ch, nbytes_expected):
    """Header, bstarts and per-stream csize semantics of one chunk."""
    assert ch[0] == 5 and ch[1] == 1
    flags, ts = ch[2], ch[3]
    nbytes, blocksize, cbytes = struct.unpack("<iii", ch[4:16])
    assert nbytes == nbytes_expected and cbytes == len(ch)
    nblocks = nbytes // blocksize
    nstreams = 1 if flags & 0x10 else ts
    neblock = blocksize // nstreams
    kinds = {"zero": 0, "raw": 0, "lz4": 0}
    for b in range(nblocks):
        p = struct.unpack("<i", ch[32 + 4 * b:36 + 4 * b])[0]
        for _ in range(nstreams):
            cs = struct.unpack("<i", ch[p:p + 4])[0]
            p += 4
            if cs == 0:
                kinds["zero"] += 1
            elif cs == neblock:
                kinds["raw"] += 1
                p += cs
            else:
                assert 0 < cs < neblock
                assert _lz4_rules(ch[p:p + cs], neblock) == neblock
                kinds["lz4"] += 1
                p += cs
        if b + 1 < nblocks:
            assert p == struct.unpack("<i", ch[36 + 4 * b:40 + 4 * b])[0]
    assert p == len(ch)
    return kinds


def _product(monkeypatch, ctx, a, b, acs, abs_, bcs, bbs, zc, zb, gpu_encode=True):
    monkeypatch.setenv("IARRAY_GPU_ENCODE", "1" if gpu_encode else "0")
    ca = ia.Container.from_numpy(ctx, a, acs, abs_)
    cb = ia.Container.from_numpy(ctx, b, bcs, bbs)
    return ia.matmul(ca, cb, zc, zb)


CASES = [
    # name, dtype, (M, K, N), filters, splitmode, C chunk, C block, data
    ("f32-shuffle-random", np.float32, (300, 257, 190), ia.COMP_SHUFFLE, ia.AUTO_SPLIT, (128, 96), (64, 32), "randn"),
    ("f64-shuffle-odd-streams", np.float64, (250, 200, 250), ia.COMP_SHUFFLE, ia.AUTO_SPLIT, (125, 125), (125, 125),
     "randn"),
    ("f32-bitshuffle-smooth", np.float32, (256, 300, 512), ia.COMP_BITSHUFFLE, ia.AUTO_SPLIT, (128, 256), (64, 64),
     "smooth"),
    ("f64-nosplit-int", np.float64, (200, 150, 170), ia.COMP_SHUFFLE, ia.NEVER_SPLIT, (100, 85), (50, 17), "int"),
    ("f32-zeros", np.float32, (128, 64, 200), ia.COMP_SHUFFLE, ia.AUTO_SPLIT, (64, 100), (32, 50), "zeros"),
    ("f64-constant", np.float64, (96, 80, 96), ia.COMP_SHUFFLE, ia.AUTO_SPLIT, (96, 96), (48, 48), "ones"),
    ("f32-tiny-blocks", np.float32, (40, 30, 20), ia.COMP_SHUFFLE, ia.AUTO_SPLIT, (8, 8), (2, 2), "randn"),
]


def _data(kind, rng, shape, dt):
    if kind == "randn":
        return rng.standard_normal(shape).astype(dt)
    if kind == "int":
        return rng.integers(-8, 9, shape).astype(dt)
    if kind == "zeros":
        return np.zeros(shape, dt)
    if kind == "ones":
        return np.ones(shape, dt)
    i = np.arange(shape[0])[:, None]
    j = np.arange(shape[1])[None, :]
    return (np.sin(0.01 * i) + np.cos(0.007 * j)).astype(dt)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_gpu_encoded_chunks(monkeypatch, orc, case):
    name, dt, (M, K, N), filters, split, zc, zb, kind = case
    print(f"case: gpu encode {name}", flush=True)
    ctx = ia.Context(clevel=5, filters=filters, splitmode=split)
    rng = np.random.default_rng(len(name))
    a = _data(kind, rng, (M, K), dt)
    b = _data(kind, rng, (K, N), dt)
    acs, abs_ = (min(M, 128), min(K, 128)), (min(M, 32), min(K, 64))
    bcs, bbs = (min(K, 128), min(N, 128)), (min(K, 64), min(N, 32))
    z = _product(monkeypatch, ctx, a, b, acs, abs_, bcs, bbs, zc, zb)
    zh = _product(monkeypatch, ctx, a, b, acs, abs_, bcs, bbs, zc, zb, gpu_encode=False)
    got = z.to_numpy()
    assert got.tobytes() == zh.to_numpy().tobytes()
    chunks = z.chunks()
    dense = orc.Array(chunks, z.shape, zc, zb, z.dtype.itemsize).to_dense()
    assert dense.tobytes() == got.tobytes()
    ext = [-(-s // c) * c for s, c in zip(z.shape, zc)]
    nbytes = zc[0] * zc[1] * np.dtype(dt).itemsize
    assert len(chunks) == (ext[0] // zc[0]) * (ext[1] // zc[1])
    kinds = {"zero": 0, "raw": 0, "lz4": 0}
    for ch in chunks:
        for k, v in _check_chunk(ch, nbytes).items():
            kinds[k] += v
    print(name, kinds)
    if kind == "zeros":
        assert kinds["raw"] == kinds["lz4"] == 0
    if kind in ("ones", "smooth"):
        assert kinds["lz4"] > 0
    # deterministic
    z2 = _product(monkeypatch, ctx, a, b, acs, abs_, bcs, bbs, zc, zb)
    assert z2.chunks() == chunks


@pytest.mark.parametrize("share", ["0.5", "0.25"])
def test_mixed_host_and_gpu_panels(monkeypatch, orc, share):
    """Some panels encoded on the GPU, the others on the host, in one product:
    every chunk valid, values equal to the all-host product."""
    print(f"case: gpu encode share {share}", flush=True)
    ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
    rng = np.random.default_rng(5)
    a = rng.standard_normal((520, 200)).astype(np.float32)
    b = rng.standard_normal((200, 300)).astype(np.float32)
    ca = ia.Container.from_numpy(ctx, a, (128, 100), (64, 50))
    cb = ia.Container.from_numpy(ctx, b, (100, 150), (50, 75))
    monkeypatch.setenv("IARRAY_GEMM_PANEL_ROWS", "1")  # one chunk row a panel: 5 panels
    monkeypatch.setenv("IARRAY_GPU_ENCODE", "0")
    ref = ia.matmul(ca, cb, (128, 150), (64, 50)).to_numpy()
    monkeypatch.setenv("IARRAY_GPU_ENCODE", share)
    z = ia.matmul(ca, cb, (128, 150), (64, 50))
    got = z.to_numpy()
    assert got.tobytes() == ref.tobytes()
    chunks = z.chunks()
    nbytes = 128 * 150 * 4
    for ch in chunks:
        _check_chunk(ch, nbytes)
    dense = orc.Array(chunks, z.shape, (128, 150), (64, 50), 4).to_dense()
    assert dense.tobytes() == got.tobytes()


@pytest.mark.parametrize("share,phases,skew", [("0.5", "2", False), ("1", "1", False), ("1", "4", False),
                                               ("0.7", "4", True)])
def test_batch_outputs_encoded_on_the_gpu(monkeypatch, orc, share, phases, skew):
    """A batch run in phases whose later products are encoded by one GPU
    encoder pass per group (ragged f32 products, several chunks each, edge
    chunks; `skew`: one product holds most of the output, so fewer phases are
    made than asked): every chunk valid, every output equal to the all-host
    one-phase batch and decoding to the same values through the oracle, also
    when the plan is replayed."""
    print(f"case: batch gpu encode share {share} phases {phases} skew {skew}", flush=True)
    ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
    rng = np.random.default_rng(11)
    pairs, cs_l, bs_l = [], [], []
    for j in range(24):
        M, N, K = (int(v) for v in rng.integers(40, 400, 3))
        if skew:
            M, N = (600, 500) if j == 0 else (int(M) // 8 + 1, int(N) // 8 + 1)
        a = rng.standard_normal((M, K)).astype(np.float32)
        b = rng.standard_normal((K, N)).astype(np.float32)
        ca = ia.Container.from_numpy(ctx, a, (min(M, 128), min(K, 128)), (min(M, 64), min(K, 64)))
        cb = ia.Container.from_numpy(ctx, b, (min(K, 128), min(N, 128)), (min(K, 64), min(N, 64)))
        pairs.append((ca, cb))
        cs_l.append((min(M, 150), min(N, 120)))
        bs_l.append((min(M, 50), min(N, 40)))
    monkeypatch.setenv("IARRAY_GPU_ENCODE", "0")
    monkeypatch.setenv("IARRAY_BATCH_PHASES", "1")
    ref = [z.to_numpy() for z in ia.matmul_batch(pairs, cs_l, bs_l)]
    monkeypatch.setenv("IARRAY_GPU_ENCODE", share)
    monkeypatch.setenv("IARRAY_BATCH_PHASES", phases)
    for rep in range(2):  # the second call replays the graphs
        zs = ia.matmul_batch(pairs, cs_l, bs_l)
        for z, r, cs, bs in zip(zs, ref, cs_l, bs_l):
            got = z.to_numpy()
            assert got.tobytes() == r.tobytes()
            chunks = z.chunks()
            ext = [(-(-cs[d] // bs[d])) * bs[d] for d in range(2)]
            for ch in chunks:
                _check_chunk(ch, ext[0] * ext[1] * 4)
            if rep == 1:
                dense = orc.Array(chunks, z.shape, cs, bs, 4).to_dense()
                assert dense.tobytes() == got.tobytes()

