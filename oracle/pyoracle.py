"""ctypes binding of oracle/liboracle.so -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, always as the checker (or the timed CPU restatement of the
reference), never as part of the product path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


class OrcArray(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int),
        ("itemsize", ctypes.c_int),
        ("shape", ctypes.c_int64 * 2),
        ("chunkshape", ctypes.c_int64 * 2),
        ("blockshape", ctypes.c_int64 * 2),
        ("nchunks", ctypes.c_int64),
        ("chunks", ctypes.POINTER(ctypes.c_void_p)),
        ("csizes", ctypes.POINTER(ctypes.c_int32)),
        ("transposed", ctypes.c_int),
    ]


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        L.orc_lz4_decode.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int]
        L.orc_zstd_decode.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int]
        L.orc_zlib_decode.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int]
        L.orc_decode_block.argtypes = [vp, i32, ctypes.c_int, vp]
        L.orc_decompress_chunk.argtypes = [vp, i32, vp, i64]
        L.orc_getitem.argtypes = [vp, i32, ctypes.c_int, ctypes.c_int, vp]
        L.orc_chunk_is_zeros.argtypes = [vp]
        L.orc_block_is_zeros.argtypes = [vp, i32, ctypes.c_int]
        for f in (L.orc_shuffle, L.orc_unshuffle):
            f.argtypes = [ctypes.c_int, i32, vp, vp]
            f.restype = None
        for f in (L.orc_bitshuffle, L.orc_bitunshuffle):
            f.argtypes = [ctypes.c_int, i32, vp, vp, ctypes.c_int]
            f.restype = None
        L.orc_array_to_dense.argtypes = [ctypes.POINTER(OrcArray), vp, ctypes.c_int]
        L.orc_get_slice.argtypes = [ctypes.POINTER(OrcArray), vp, vp, vp, vp, ctypes.c_int]
        for f in (L.orc_gemm_blosc, L.orc_gemv_blosc, L.orc_opt_gemm, L.orc_opt_gemv, L.orc_opt_gemm_a, L.orc_opt_gemm_b):
            f.argtypes = [ctypes.POINTER(OrcArray), ctypes.POINTER(OrcArray), vp, vp, vp, ctypes.c_int]
        L.orc_set_blas.argtypes = [ctypes.c_int]
        L.orc_blas_name.restype = ctypes.c_char_p
        L.orc_set_lz4.argtypes = [ctypes.c_int]
        L.orc_encode_dense.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int]
        L.orc_encode_dense.restype = i64
        L.orc_dd_gemm_sample.argtypes = [vp, vp, i64, i64, i64, vp, vp, i64, vp]
        L.orc_dd_gemm_sample.restype = None
        _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def zstd_decode(src: bytes, n: int) -> bytes:
    """Decode zstd frames (orc_zstd_decode) into at most n bytes."""
    out = ctypes.create_string_buffer(max(n, 1))
    r = lib().orc_zstd_decode(src, len(src), out, n)
    if r < 0:
        raise ValueError(f"zstd decode error {r}")
    return out.raw[:r]


def zlib_decode(src: bytes, n: int) -> bytes:
    """Decode a zlib stream (orc_zlib_decode) into at most n bytes."""
    out = ctypes.create_string_buffer(max(n, 1))
    r = lib().orc_zlib_decode(src, len(src), out, n)
    if r < 0:
        raise ValueError(f"zlib decode error {r}")
    return out.raw[:r]


def lz4_decode(src: bytes, n: int) -> bytes:
    out = ctypes.create_string_buffer(max(n, 1))
    r = lib().orc_lz4_decode(src, len(src), out, n)
    if r < 0:
        raise ValueError(f"lz4 decode error {r}")
    return out.raw[:r]


def decompress_chunk(chunk: bytes, nbytes: int) -> bytes:
    out = ctypes.create_string_buffer(max(nbytes, 1))
    r = lib().orc_decompress_chunk(chunk, len(chunk), out, nbytes)
    if r < 0:
        raise ValueError(f"chunk decode error {r}")
    return out.raw[:r]


def shuffle(ts, data: bytes) -> bytes:
    out = ctypes.create_string_buffer(len(data))
    lib().orc_shuffle(ts, len(data), data, out)
    return out.raw


def unshuffle(ts, data: bytes) -> bytes:
    out = ctypes.create_string_buffer(len(data))
    lib().orc_unshuffle(ts, len(data), data, out)
    return out.raw


def bitshuffle(ts, data: bytes, version=5) -> bytes:
    out = ctypes.create_string_buffer(len(data))
    lib().orc_bitshuffle(ts, len(data), data, out, version)
    return out.raw


def bitunshuffle(ts, data: bytes, version=5) -> bytes:
    out = ctypes.create_string_buffer(len(data))
    lib().orc_bitunshuffle(ts, len(data), data, out, version)
    return out.raw


class Array:
    """An oracle view of a chunked array: compressed chunk bytes plus tiling."""

    def __init__(self, chunks, shape, chunkshape, blockshape, itemsize, transposed=False):
        self.chunks = [bytes(c) for c in chunks]
        self._bufs = [ctypes.create_string_buffer(c, len(c)) for c in self.chunks]
        self.ndim = len(shape)
        s = OrcArray()
        s.ndim = self.ndim
        s.itemsize = itemsize
        for i in range(self.ndim):
            s.shape[i] = shape[i]
            s.chunkshape[i] = chunkshape[i]
            s.blockshape[i] = blockshape[i]
        s.nchunks = len(chunks)
        self._ptrs = (ctypes.c_void_p * len(chunks))(*[ctypes.addressof(b) for b in self._bufs])
        self._sizes = (ctypes.c_int32 * len(chunks))(*[len(c) for c in self.chunks])
        s.chunks = ctypes.cast(self._ptrs, ctypes.POINTER(ctypes.c_void_p))
        s.csizes = ctypes.cast(self._sizes, ctypes.POINTER(ctypes.c_int32))
        s.transposed = 1 if transposed else 0
        self.s = s
        self.shape = tuple(shape)
        self.itemsize = itemsize

    @property
    def dtype(self):
        return np.float64 if self.itemsize == 8 else np.float32

    def logical_shape(self):
        if self.ndim == 2 and self.s.transposed:
            return (self.shape[1], self.shape[0])
        return self.shape

    def to_dense(self, dtype=None, nthreads=4):
        out = np.zeros(self.logical_shape(), dtype=dtype or self.dtype)
        r = lib().orc_array_to_dense(ctypes.byref(self.s), _ptr(out), nthreads)
        if r != 0:
            raise ValueError(f"oracle to_dense error {r}")
        return out


def _matmul(fn, a: Array, b: Array, cchunk, cblock, nthreads):
    M = a.logical_shape()[0]
    out = np.zeros((M,) if b.ndim == 1 else (M, b.logical_shape()[1]), dtype=a.dtype)
    cc = (ctypes.c_int64 * 2)(*list(cchunk) + [1] * (2 - len(cchunk)))
    cb = (ctypes.c_int64 * 2)(*list(cblock) + [1] * (2 - len(cblock)))
    r = fn(ctypes.byref(a.s), ctypes.byref(b.s), cc, cb, _ptr(out), nthreads)
    if r != 0:
        raise ValueError(f"oracle matmul error {r}")
    return out


def gemm_blosc(a, b, cchunk, cblock, nthreads=4):
    return _matmul(lib().orc_gemm_blosc, a, b, cchunk, cblock, nthreads)


def gemv_blosc(a, b, cchunk, cblock, nthreads=4):
    return _matmul(lib().orc_gemv_blosc, a, b, cchunk, cblock, nthreads)


def opt_gemm(a, b, cchunk, cblock, nthreads=4):
    return _matmul(lib().orc_opt_gemm, a, b, cchunk, cblock, nthreads)


def opt_gemm_a(a, b, cchunk, cblock, nthreads=4):
    return _matmul(lib().orc_opt_gemm_a, a, b, cchunk, cblock, nthreads)


def opt_gemm_b(a, b, cchunk, cblock, nthreads=4):
    return _matmul(lib().orc_opt_gemm_b, a, b, cchunk, cblock, nthreads)


def opt_gemv(a, b, cchunk, cblock, nthreads=4):
    return _matmul(lib().orc_opt_gemv, a, b, cchunk, cblock, nthreads)


def set_blas(which):
    return lib().orc_set_blas(which)


def blas_name():
    return lib().orc_blas_name().decode()


def set_lz4(which):
    return lib().orc_set_lz4(which)


def encode_dense(c: np.ndarray, chunkshape, blockshape, clevel=5, filt=1, nthreads=4):
    shape = (ctypes.c_int64 * 2)(*c.shape)
    cc = (ctypes.c_int64 * 2)(*chunkshape)
    cb = (ctypes.c_int64 * 2)(*blockshape)
    return lib().orc_encode_dense(_ptr(c), c.ndim, shape, c.itemsize, cc, cb, clevel, filt, nthreads)


def dd_gemm_sample(a: np.ndarray, b: np.ndarray, rows, cols):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    out = np.empty(len(rows), dtype=np.float64)
    lib().orc_dd_gemm_sample(_ptr(a), _ptr(b), a.shape[0], b.shape[1], a.shape[1], _ptr(rows), _ptr(cols),
                             len(rows), _ptr(out))
    return out


def frame_chunks(data: bytes):
    """Chunks of a contiguous Blosc2 frame (b2frame): header length is the
    int32 big-endian at byte 11, the schunk nbytes/chunksize are msgpack
    fields at bytes 30 and 58, data chunks follow the header back to back."""
    import struct
    hl = struct.unpack(">i", data[11:15])[0]
    nbytes = struct.unpack(">q", data[30:38])[0]
    chunksize = struct.unpack(">i", data[58:62])[0]
    n = -(-nbytes // chunksize)
    pos, out = hl, []
    for _ in range(n):
        cb = struct.unpack("<i", data[pos + 12:pos + 16])[0]
        out.append(data[pos:pos + cb])
        pos += cb
    return out, nbytes
