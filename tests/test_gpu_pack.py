"""The C-write pack stage on the GPU (iamx_pack_blocks): every C block tile
gathered from the dense C (zeros outside the valid extent), byte-shuffled or
bit-shuffled, written chunk-major and block-major -- the bytes
blosc2_compress_ctx filters before LZ4 (src/linalg/matmul/gemm.c:216-230).
Checked byte for byte against numpy: the closed forms of Caterva's tiling
(block r-major inside a chunk, chunk r-major in the grid) and of the shuffle
(byte plane j holds byte j of every element)."""
import ctypes

import numpy as np
import pytest

import pyiarray as ia

pytestmark = pytest.mark.gpu


class Tiling(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int64 * 2), ("chunk", ctypes.c_int64 * 2), ("block", ctypes.c_int64 * 2),
                ("nchunks", ctypes.c_int64 * 2), ("nblocks", ctypes.c_int64 * 2), ("typesize", ctypes.c_int32),
                ("filter", ctypes.c_int32)]


def _cdiv(a, b):
    return -(-a // b)


def _expected(C, chunk, block, filt):
    """Packed bytes of every block: numpy restatement."""
    ts = C.itemsize
    ext = [_cdiv(c, b) * b for c, b in zip(chunk, block)]
    nch = [_cdiv(s, c) for s, c in zip(C.shape, chunk)]
    nbl = [e // b for e, b in zip(ext, block)]
    out = []
    for ci in range(nch[0]):
        for cj in range(nch[1]):
            for bi in range(nbl[0]):
                for bj in range(nbl[1]):
                    r0, c0 = ci * chunk[0] + bi * block[0], cj * chunk[1] + bj * block[1]
                    tile = np.zeros(block, C.dtype)
                    vr = max(0, min(block[0], chunk[0] - bi * block[0], C.shape[0] - r0))
                    vc = max(0, min(block[1], chunk[1] - bj * block[1], C.shape[1] - c0))
                    tile[:vr, :vc] = C[r0:r0 + vr, c0:c0 + vc]
                    b = tile.ravel().view(np.uint8).reshape(-1, ts)
                    if filt == 1:
                        out.append(b.T.ravel())
                    else:
                        out.append(b.ravel())
    return np.concatenate(out)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape,chunk,block", [
    ((500, 500), (250, 250), (125, 125)),   # odd elements per block, rows not a multiple of 4
    ((300, 260), (128, 128), (64, 64)),     # ragged edges, rows of a multiple of 4
    ((90, 130), (40, 60), (20, 12)),
    ((37, 51), (37, 51), (11, 13)),
    ((64, 64), (64, 64), (8, 9)),
])
@pytest.mark.parametrize("filt", [1, 0], ids=["shuffle", "none"])
def test_pack_blocks_matches_numpy(dtype, shape, chunk, block, filt):
    print(f"case: pack {np.dtype(dtype).name} {shape} {chunk} {block} filter {filt}", flush=True)
    rng = np.random.default_rng(sum(shape) + sum(block))
    C = rng.standard_normal(shape).astype(dtype)
    want = _expected(C, chunk, block, filt)
    h = ctypes.CDLL(ia.HIP_LIB_PATH)
    h.iamx_malloc.restype = ctypes.c_void_p
    h.iamx_malloc.argtypes = [ctypes.c_size_t]
    h.iamx_free.argtypes = [ctypes.c_void_p]
    h.iamx_h2d.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    h.iamx_d2h.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    h.iamx_pack_blocks.argtypes = [ctypes.POINTER(Tiling), ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    assert h.iamx_init() == 0
    t = Tiling()
    for i in range(2):
        t.shape[i], t.chunk[i], t.block[i] = shape[i], chunk[i], block[i]
        t.nchunks[i] = _cdiv(shape[i], chunk[i])
        t.nblocks[i] = _cdiv(chunk[i], block[i])
    t.typesize, t.filter = C.itemsize, filt
    d_src = h.iamx_malloc(C.nbytes + 64)
    d_out = h.iamx_malloc(want.nbytes + 64)
    try:
        assert d_src and d_out
        assert h.iamx_h2d(d_src, C.ctypes.data, C.nbytes) == 0
        assert h.iamx_pack_blocks(ctypes.byref(t), d_src, shape[1], d_out) == 0
        got = np.empty_like(want)
        assert h.iamx_d2h(got.ctypes.data, d_out, got.nbytes) == 0
    finally:
        h.iamx_free(d_src)
        h.iamx_free(d_out)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} bytes differ, first at {bad[:8].tolist()}"
