"""ctypes binding of libiarray.so (the iarray C API of include/iarray.h).

Glue for tests and benchmarks: every call goes through the C ABI exactly as a
C caller (or python-iarray's Cython layer) would make it.  Importing this
module loads the in-tree libraries and fails loudly when they are missing;
there is no fallback path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libiarray.so")
HIP_LIB_PATH = os.path.join(HERE, "libiamx.so")

DOUBLE, FLOAT = 0, 1
COMP_SHUFFLE, COMP_BITSHUFFLE = 0x1, 0x2
BLOSCLZ, LZ4, LZ4HC = 0, 1, 2
ALWAYS_SPLIT, NEVER_SPLIT, AUTO_SPLIT = 1, 2, 3
PLAIN, SPLIT_BF16, OZAKI_INT8, OZAKI_CRT = 0, 1, 2, 3

_DT = {np.dtype(np.float64): 0, np.dtype(np.float32): 1, np.dtype(np.int64): 10, np.dtype(np.int32): 11,
       np.dtype(np.int16): 12, np.dtype(np.int8): 13, np.dtype(np.uint64): 16, np.dtype(np.uint32): 17,
       np.dtype(np.uint16): 18, np.dtype(np.uint8): 19}
_NP = {v: k for k, v in _DT.items()}


class Config(ctypes.Structure):
    _fields_ = [("compression_codec", ctypes.c_int), ("compression_level", ctypes.c_int),
                ("compression_favor", ctypes.c_int), ("use_dict", ctypes.c_int), ("splitmode", ctypes.c_int),
                ("filter_flags", ctypes.c_int), ("eval_method", ctypes.c_uint), ("max_num_threads", ctypes.c_int),
                ("fp_mantissa_bits", ctypes.c_uint8), ("btune", ctypes.c_bool), ("compression_meta", ctypes.c_uint8)]


class DTShape(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int), ("dtype_size", ctypes.c_int32), ("ndim", ctypes.c_int8),
                ("shape", ctypes.c_int64 * 8)]


class Storage(ctypes.Structure):
    _fields_ = [("urlpath", ctypes.c_char_p), ("contiguous", ctypes.c_bool), ("chunkshape", ctypes.c_int64 * 8),
                ("blockshape", ctypes.c_int64 * 8)]


class MatmulParams(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("ozaki_slices", ctypes.c_int)]


class IArrayError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(msg)
        self.rc = rc


# error symbols (values of include/iarray.h)
ERR_BIT = 1 << 63
_K = {"INVALID": 0x01, "FAILED": 0x02, "EMPTY": 0x03, "TOO_SMALL": 0x05, "NOT_SUPPORTED": 0x06}
_S = {"DTSHAPE": 2, "SHAPE": 3, "CHUNKSHAPE": 4, "NDIM": 5, "DTYPE": 6, "STORAGE": 7, "BUFFER": 9, "CATERVA": 10, "BLOSC": 11,
      "BLOCKSHAPE": 13, "DEVICE": 64}


def err_code(kind, subject):
    return (_K[kind] << 32) | (0x100 + _S[subject])


ERR_INVALID_NDIM = err_code("INVALID", "NDIM")
ERR_INVALID_SHAPE = err_code("INVALID", "SHAPE")
ERR_INVALID_DTYPE = err_code("INVALID", "DTYPE")
ERR_INVALID_CHUNKSHAPE = err_code("INVALID", "CHUNKSHAPE")
ERR_INVALID_BLOCKSHAPE = err_code("INVALID", "BLOCKSHAPE")
ERR_INVALID_STORAGE = err_code("INVALID", "STORAGE")
ERR_TOO_SMALL_BUFFER = err_code("TOO_SMALL", "BUFFER")
ERR_BLOSC_FAILED = err_code("FAILED", "BLOSC")
ERR_CATERVA_FAILED = err_code("FAILED", "CATERVA")
ERR_NO_DEVICE = err_code("NOT_SUPPORTED", "DEVICE")
ERR_FILE_OPEN = (_K["FAILED"] << 32) | 0x05  # libinac's file subject, not a user-defined one
ERR_NOT_SUPPORTED = 0x06 << 32  # INA_ERR_NOT_SUPPORTED (no subject)
ERR_INVALID_ARGUMENT = 0x08 << 32  # INA_ERR_INVALID_ARGUMENT (no subject)

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not (os.path.exists(LIB_PATH) and os.path.exists(HIP_LIB_PATH)):
        raise ImportError(f"{LIB_PATH} / {HIP_LIB_PATH} not built: run `make -C {HERE}` "
                          "(or __graft_entry__.build()) first")
    L = ctypes.CDLL(LIB_PATH)
    rc_t, vp, i64, P = ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER
    sig = {
        "iarray_init": ([], rc_t), "iarray_destroy": ([], None),
        "iarray_err_strerror": ([rc_t], ctypes.c_char_p),
        "iarray_context_new": ([P(Config), P(vp)], rc_t), "iarray_context_free": ([P(vp)], None),
        "iarray_empty": ([vp, P(DTShape), P(Storage), P(vp)], rc_t),
        "iarray_zeros": ([vp, P(DTShape), P(Storage), P(vp)], rc_t),
        "iarray_arange": ([vp, P(DTShape), ctypes.c_double, ctypes.c_double, P(Storage), P(vp)], rc_t),
        "iarray_linspace": ([vp, P(DTShape), ctypes.c_double, ctypes.c_double, P(Storage), P(vp)], rc_t),
        "iarray_from_buffer": ([vp, P(DTShape), vp, i64, P(Storage), P(vp)], rc_t),
        "iarray_to_buffer": ([vp, vp, vp, i64], rc_t),
        "iarray_get_slice_buffer": ([vp, vp, P(i64), P(i64), vp, i64], rc_t),
        "iarray_get_dtshape": ([vp, vp, P(DTShape)], rc_t),
        "iarray_container_info": ([vp, P(i64), P(i64)], rc_t),
        "iarray_container_free": ([vp, P(vp)], None),
        "iarray_container_almost_equal": ([vp, vp, ctypes.c_double], rc_t),
        "iarray_linalg_matmul": ([vp, vp, vp, P(Storage), P(vp)], rc_t),
        "iarray_linalg_matmul_params": ([vp, vp, vp, P(Storage), P(MatmulParams), P(vp)], rc_t),
        "iarray_linalg_matmul_batch": ([vp, ctypes.c_int, P(vp), P(vp), P(Storage), P(MatmulParams), P(vp)], rc_t),
        "iarray_linalg_transpose": ([vp, vp, P(vp)], rc_t),
        "iarray_get_slice": ([vp, vp, P(i64), P(i64), ctypes.c_bool, P(Storage), P(vp)], rc_t),
        "iarray_opt_gemv": ([vp, vp, vp, P(Storage), P(vp)], rc_t),
        "iarray_opt_gemm": ([vp, vp, vp, P(Storage), P(vp)], rc_t),
        "iarray_opt_gemm_a": ([vp, vp, vp, P(Storage), P(vp)], rc_t),
        "iarray_container_load": ([vp, ctypes.c_char_p, P(vp)], rc_t),
        "iarray_container_open": ([vp, ctypes.c_char_p, P(vp)], rc_t),
        "iarray_container_save": ([vp, vp, ctypes.c_char_p], rc_t),
        "iarray_to_cframe": ([vp, vp, P(ctypes.POINTER(ctypes.c_uint8)), P(i64), P(ctypes.c_bool)], rc_t),
        "iarray_from_cframe": ([vp, vp, i64, ctypes.c_bool, P(vp)], rc_t),
        "iarray_container_remove": ([ctypes.c_char_p], rc_t),
        "iarray_opt_gemm_b": ([vp, vp, vp, P(Storage), P(vp)], rc_t),
        "iarray_container_chunk": ([vp, i64, P(ctypes.POINTER(ctypes.c_uint8)), P(ctypes.c_int32)], rc_t),
        "iarray_container_nchunks": ([vp, P(i64)], rc_t),
        "iarray_from_chunks": ([vp, P(DTShape), P(Storage), i64, P(vp), P(ctypes.c_int32), P(vp)], rc_t),
        "iarray_decode_chunks_flat": ([vp, i64, P(vp), P(ctypes.c_int32), vp, i64], rc_t),
        "iarray_prof_enable": ([ctypes.c_bool], None), "iarray_prof_reset": ([], None),
        "iarray_prof_entry": ([ctypes.c_int, P(ctypes.c_char_p), P(i64), P(ctypes.c_double)], ctypes.c_bool),
        "iarray_last_matmul_times": ([P(ctypes.c_double), P(ctypes.c_double)], None),
        "iarray_last_gpu_encode_bytes": ([P(i64), P(i64)], None),
        "iarray_last_ozaki_slices": ([P(ctypes.c_double)], ctypes.c_int),
        "iarray_last_gemv_panels": ([], ctypes.c_int),
        "iarray_device_info": ([ctypes.c_char_p, ctypes.c_int, P(ctypes.c_int), ctypes.c_char_p, ctypes.c_int], rc_t),
        "iarray_container_device_plan": ([vp, vp], rc_t),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    _check(L.iarray_init())
    return L


def _libc_free(ptr):
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    libc.free(ctypes.cast(ptr, ctypes.c_void_p))


def last_gpu_encode_bytes():
    """(filtered stream bytes read, chunk bytes written) by the GPU encoder in the last matmul."""
    a, b = ctypes.c_int64(), ctypes.c_int64()
    lib().iarray_last_gpu_encode_bytes(ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def remove(path):
    """Deletes a frame file or frame directory (iarray_container_remove)."""
    _check(lib().iarray_container_remove(os.fsencode(path)))


def exported_symbols():
    return ["iarray_init", "iarray_destroy", "iarray_err_strerror", "iarray_context_new", "iarray_context_free",
            "iarray_empty", "iarray_zeros", "iarray_arange", "iarray_linspace", "iarray_from_buffer",
            "iarray_to_buffer", "iarray_get_slice_buffer", "iarray_get_dtshape", "iarray_get_storage",
            "iarray_is_view", "iarray_container_info", "iarray_container_free", "iarray_container_almost_equal",
            "iarray_linalg_matmul", "iarray_linalg_transpose", "iarray_opt_gemv", "iarray_opt_gemm",
            "iarray_opt_gemm_a", "iarray_opt_gemm_b"]


def _check(rc):
    if rc & ERR_BIT or rc < 0:
        code = rc & ~ERR_BIT if rc >= 0 else (rc + (1 << 64)) & ~ERR_BIT
        raise IArrayError(code, _lib.iarray_err_strerror(rc).decode() if _lib else f"rc={rc:#x}")


def _dtshape(dtype, shape):
    d = DTShape()
    d.dtype = _DT[np.dtype(dtype)]
    d.ndim = len(shape)
    for i, s in enumerate(shape):
        d.shape[i] = s
    return d


def _storage(chunks, blocks, urlpath=None, contiguous=False):
    """iarray_storage_t: with a urlpath the container is also kept on disk, as
    one frame file (contiguous) or a sparse frame directory."""
    s = Storage()
    s.urlpath = os.fsencode(urlpath) if urlpath is not None else None
    s.contiguous = bool(contiguous)
    for i, c in enumerate(chunks):
        s.chunkshape[i] = c
    for i, b in enumerate(blocks):
        s.blockshape[i] = b
    return s


def make_config(clevel=5, codec=LZ4, filters=COMP_SHUFFLE, splitmode=AUTO_SPLIT, nthreads=1, btune=False):
    c = Config()
    c.compression_codec = codec
    c.compression_level = clevel
    c.compression_favor = 0
    c.use_dict = 0
    c.splitmode = splitmode
    c.filter_flags = filters
    c.eval_method = 2
    c.max_num_threads = nthreads
    c.fp_mantissa_bits = 0
    c.btune = btune
    c.compression_meta = 0
    return c


class Context:
    def __init__(self, **cfg):
        L = lib()
        self.cfg = make_config(**cfg)
        self.ptr = ctypes.c_void_p()
        _check(L.iarray_context_new(ctypes.byref(self.cfg), ctypes.byref(self.ptr)))

    def __del__(self):
        if _lib is not None and getattr(self, "ptr", None):
            _lib.iarray_context_free(ctypes.byref(self.ptr))


class Container:
    def __init__(self, ptr, ctx, dtype, shape, base=None):
        self.ptr = ptr
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.shape = tuple(shape)
        self._base = base  # a view keeps its viewed container alive

    def __del__(self):
        if _lib is not None and getattr(self, "ptr", None):
            _lib.iarray_container_free(self.ctx.ptr, ctypes.byref(self.ptr))

    @classmethod
    def _new(cls, fn, ctx, dtype, shape, chunks, blocks, *extra, urlpath=None, contiguous=False):
        p = ctypes.c_void_p()
        d = _dtshape(dtype, shape)
        s = _storage(chunks, blocks, urlpath, contiguous)
        _check(fn(ctx.ptr, ctypes.byref(d), *extra, ctypes.byref(s), ctypes.byref(p)))
        return cls(p, ctx, dtype, shape)

    @classmethod
    def from_numpy(cls, ctx, arr, chunks, blocks, urlpath=None, contiguous=False):
        arr = np.ascontiguousarray(arr)
        p = ctypes.c_void_p()
        d = _dtshape(arr.dtype, arr.shape)
        s = _storage(chunks, blocks, urlpath, contiguous)
        _check(lib().iarray_from_buffer(ctx.ptr, ctypes.byref(d), arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes,
                                        ctypes.byref(s), ctypes.byref(p)))
        return cls(p, ctx, arr.dtype, arr.shape)

    @classmethod
    def arange(cls, ctx, dtype, shape, chunks, blocks, start=0.0, step=1.0, urlpath=None, contiguous=False):
        return cls._new(lib().iarray_arange, ctx, dtype, shape, chunks, blocks, ctypes.c_double(start),
                        ctypes.c_double(step), urlpath=urlpath, contiguous=contiguous)

    @classmethod
    def linspace(cls, ctx, dtype, shape, chunks, blocks, start, stop, urlpath=None, contiguous=False):
        return cls._new(lib().iarray_linspace, ctx, dtype, shape, chunks, blocks, ctypes.c_double(start),
                        ctypes.c_double(stop), urlpath=urlpath, contiguous=contiguous)

    @classmethod
    def zeros(cls, ctx, dtype, shape, chunks, blocks, urlpath=None, contiguous=False):
        return cls._new(lib().iarray_zeros, ctx, dtype, shape, chunks, blocks, urlpath=urlpath,
                        contiguous=contiguous)

    @classmethod
    def empty(cls, ctx, dtype, shape, chunks, blocks, urlpath=None, contiguous=False):
        return cls._new(lib().iarray_empty, ctx, dtype, shape, chunks, blocks, urlpath=urlpath,
                        contiguous=contiguous)

    @classmethod
    def from_chunks(cls, ctx, dtype, shape, chunks_shape, blocks, chunk_bytes):
        p = ctypes.c_void_p()
        d = _dtshape(dtype, shape)
        s = _storage(chunks_shape, blocks)
        bufs = [ctypes.create_string_buffer(c, len(c)) for c in chunk_bytes]
        ptrs = (ctypes.c_void_p * len(bufs))(*[ctypes.addressof(b) for b in bufs])
        sizes = (ctypes.c_int32 * len(bufs))(*[len(c) for c in chunk_bytes])
        _check(lib().iarray_from_chunks(ctx.ptr, ctypes.byref(d), ctypes.byref(s), len(bufs), ptrs, sizes,
                                        ctypes.byref(p)))
        return cls(p, ctx, dtype, shape)

    @classmethod
    def load(cls, ctx, path, lazy=False):
        """A container read from a Blosc2 frame file (iarray_container_load /
        iarray_container_open)."""
        p = ctypes.c_void_p()
        fn = lib().iarray_container_open if lazy else lib().iarray_container_load
        _check(fn(ctx.ptr, os.fsencode(path), ctypes.byref(p)))
        d = DTShape()
        _check(lib().iarray_get_dtshape(ctx.ptr, p, ctypes.byref(d)))
        inv = {v: k for k, v in _DT.items()}
        return cls(p, ctx, inv[d.dtype], tuple(d.shape[i] for i in range(d.ndim)))

    @classmethod
    def from_cframe(cls, ctx, frame):
        """A container read from the bytes of a Blosc2 frame (iarray_from_cframe)."""
        buf = ctypes.create_string_buffer(bytes(frame), len(frame))
        p = ctypes.c_void_p()
        _check(lib().iarray_from_cframe(ctx.ptr, buf, len(frame), True, ctypes.byref(p)))
        d = DTShape()
        _check(lib().iarray_get_dtshape(ctx.ptr, p, ctypes.byref(d)))
        inv = {v: k for k, v in _DT.items()}
        return cls(p, ctx, inv[d.dtype], tuple(d.shape[i] for i in range(d.ndim)))

    def save(self, path):
        """Writes the container as a contiguous frame file (iarray_container_save)."""
        _check(lib().iarray_container_save(self.ctx.ptr, self.ptr, os.fsencode(path)))

    def to_cframe(self):
        """The container's frame as bytes (iarray_to_cframe)."""
        buf = ctypes.POINTER(ctypes.c_uint8)()
        n = ctypes.c_int64()
        needs_free = ctypes.c_bool()
        _check(lib().iarray_to_cframe(self.ctx.ptr, self.ptr, ctypes.byref(buf), ctypes.byref(n),
                                      ctypes.byref(needs_free)))
        out = ctypes.string_at(buf, n.value)
        if needs_free.value:
            _libc_free(buf)
        return out

    def to_numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        _check(lib().iarray_to_buffer(self.ctx.ptr, self.ptr, out.ctypes.data_as(ctypes.c_void_p), out.nbytes))
        return out

    def slice(self, start, stop):
        st = (ctypes.c_int64 * len(start))(*start)
        sp = (ctypes.c_int64 * len(stop))(*stop)
        shape = [((b if b >= 0 else b + d) - (a if a >= 0 else a + d)) for a, b, d in zip(start, stop, self.shape)]
        out = np.empty(shape, dtype=self.dtype)
        _check(lib().iarray_get_slice_buffer(self.ctx.ptr, self.ptr, st, sp, out.ctypes.data_as(ctypes.c_void_p),
                                             out.nbytes))
        return out

    def info(self):
        nb, cb = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().iarray_container_info(self.ptr, ctypes.byref(nb), ctypes.byref(cb)))
        return nb.value, cb.value

    def chunks(self):
        n = ctypes.c_int64()
        _check(lib().iarray_container_nchunks(self.ptr, ctypes.byref(n)))
        out = []
        for i in range(n.value):
            p = ctypes.POINTER(ctypes.c_uint8)()
            cs = ctypes.c_int32()
            _check(lib().iarray_container_chunk(self.ptr, i, ctypes.byref(p), ctypes.byref(cs)))
            out.append(ctypes.string_at(p, cs.value) if p else b"")
        return out

    def transpose(self):
        p = ctypes.c_void_p()
        _check(lib().iarray_linalg_transpose(self.ctx.ptr, self.ptr, ctypes.byref(p)))
        return Container(p, self.ctx, self.dtype, self.shape[::-1], base=self)

    def get_slice(self, start, stop, view=True, chunks=None, blocks=None):
        """iarray_get_slice: a view of [start, stop), or a copy written with chunks / blocks."""
        st = (ctypes.c_int64 * len(start))(*start)
        sp = (ctypes.c_int64 * len(stop))(*stop)
        s = _storage(chunks or [], blocks or [])
        p = ctypes.c_void_p()
        _check(lib().iarray_get_slice(self.ctx.ptr, self.ptr, st, sp, view, ctypes.byref(s), ctypes.byref(p)))
        d = DTShape()
        _check(lib().iarray_get_dtshape(self.ctx.ptr, p, ctypes.byref(d)))
        shape = tuple(d.shape[i] for i in range(d.ndim))
        out = Container(p, self.ctx, self.dtype, shape, base=self if view else None)
        if not view:
            out._cs, out._bs = tuple(chunks), tuple(blocks)
        return out


def _out_shape(a, b):
    return (a.shape[0],) if len(b.shape) == 1 else (a.shape[0], b.shape[1])


def matmul(a, b, chunks, blocks, precision=PLAIN, slices=0, urlpath=None, contiguous=False):
    p = ctypes.c_void_p()
    s = _storage(chunks, blocks, urlpath, contiguous)
    if precision == PLAIN and slices == 0:
        _check(lib().iarray_linalg_matmul(a.ctx.ptr, a.ptr, b.ptr, ctypes.byref(s), ctypes.byref(p)))
    else:
        prm = MatmulParams(precision, slices)
        _check(lib().iarray_linalg_matmul_params(a.ctx.ptr, a.ptr, b.ptr, ctypes.byref(s), ctypes.byref(prm),
                                                 ctypes.byref(p)))
    return Container(p, a.ctx, a.dtype, _out_shape(a, b))


def opt_gemm(a, b, chunks, blocks, urlpath=None, contiguous=False):
    p = ctypes.c_void_p()
    s = _storage(chunks, blocks, urlpath, contiguous)
    _check(lib().iarray_opt_gemm(a.ctx.ptr, a.ptr, b.ptr, ctypes.byref(s), ctypes.byref(p)))
    return Container(p, a.ctx, a.dtype, _out_shape(a, b))


def opt_gemm_a(a, b, chunks, blocks, urlpath=None, contiguous=False):
    p = ctypes.c_void_p()
    s = _storage(chunks, blocks, urlpath, contiguous)
    _check(lib().iarray_opt_gemm_a(a.ctx.ptr, a.ptr, b.ptr, ctypes.byref(s), ctypes.byref(p)))
    return Container(p, a.ctx, a.dtype, _out_shape(a, b))


def opt_gemm_b(a, b, chunks, blocks, urlpath=None, contiguous=False):
    p = ctypes.c_void_p()
    s = _storage(chunks, blocks, urlpath, contiguous)
    _check(lib().iarray_opt_gemm_b(a.ctx.ptr, a.ptr, b.ptr, ctypes.byref(s), ctypes.byref(p)))
    return Container(p, a.ctx, a.dtype, _out_shape(a, b))


def opt_gemv(a, b, chunks, blocks, urlpath=None, contiguous=False):
    p = ctypes.c_void_p()
    s = _storage(chunks, blocks, urlpath, contiguous)
    _check(lib().iarray_opt_gemv(a.ctx.ptr, a.ptr, b.ptr, ctypes.byref(s), ctypes.byref(p)))
    return Container(p, a.ctx, a.dtype, _out_shape(a, b))


def matmul_batch(pairs, chunks_list, blocks_list, precision=PLAIN, slices=0):
    n = len(pairs)
    A = (ctypes.c_void_p * n)(*[a.ptr.value for a, _ in pairs])
    B = (ctypes.c_void_p * n)(*[b.ptr.value for _, b in pairs])
    S = (Storage * n)(*[_storage(c, b) for c, b in zip(chunks_list, blocks_list)])
    C = (ctypes.c_void_p * n)()
    prm = MatmulParams(precision, slices)
    ctx = pairs[0][0].ctx
    _check(lib().iarray_linalg_matmul_batch(ctx.ptr, n, A, B, S, ctypes.byref(prm), C))
    return [Container(ctypes.c_void_p(C[i]), ctx, a.dtype, _out_shape(a, b)) for i, (a, b) in enumerate(pairs)]


def decode_chunks_flat(ctx, chunk_bytes, nbytes):
    bufs = [ctypes.create_string_buffer(c, len(c)) for c in chunk_bytes]
    ptrs = (ctypes.c_void_p * len(bufs))(*[ctypes.addressof(b) for b in bufs])
    sizes = (ctypes.c_int32 * len(bufs))(*[len(c) for c in chunk_bytes])
    out = ctypes.create_string_buffer(max(nbytes, 1))
    _check(lib().iarray_decode_chunks_flat(ctx.ptr, len(bufs), ptrs, sizes, out, nbytes))
    return out.raw[:nbytes]


def prof_enable(on=True):
    lib().iarray_prof_enable(on)


def prof_reset():
    lib().iarray_prof_reset()


def prof_entries():
    out = {}
    i = 0
    while True:
        name, n, ms = ctypes.c_char_p(), ctypes.c_int64(), ctypes.c_double()
        if not lib().iarray_prof_entry(i, ctypes.byref(name), ctypes.byref(n), ctypes.byref(ms)):
            break
        out[name.value.decode()] = (n.value, ms.value)
        i += 1
    return out


def last_gemv_panels():
    """Panels the last GEMV streamed its matrix in (0: resident)."""
    return lib().iarray_last_gemv_panels()


def last_ozaki_slices():
    """(slices used by the last OZAKI_INT8 product, sampled componentwise error
    of an automatic pick or -1)."""
    e = ctypes.c_double()
    s = lib().iarray_last_ozaki_slices(ctypes.byref(e))
    return s, e.value


def last_matmul_times():
    e, d = ctypes.c_double(), ctypes.c_double()
    lib().iarray_last_matmul_times(ctypes.byref(e), ctypes.byref(d))
    return e.value, d.value


def device_info():
    name = ctypes.create_string_buffer(256)
    arch = ctypes.create_string_buffer(64)
    cus = ctypes.c_int()
    _check(lib().iarray_device_info(name, 256, ctypes.byref(cus), arch, 64))
    return name.value.decode(), cus.value, arch.value.decode()


def device_sync():
    """Wait for all work on the library's HIP stream."""
    h = ctypes.CDLL(HIP_LIB_PATH)
    rc = h.iamx_sync()
    if rc:
        raise IArrayError(ERR_NO_DEVICE, "iamx_sync failed")
