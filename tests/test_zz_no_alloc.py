"""iamx_gemv_fused_noalloc (libiamx.so C ABI) against iamx_gemv_fused.

The no-allocation entry point runs the same kernels as iamx_gemv_fused with
its LZ4 scratch and partial sums in a fixed device array, so on the same job
tables and x it must give the same bits.  Checked on two calls in a row, the
second on a smaller container (the array keeps the first call's contents:
the partial sums must be cleared, and stale decoded bytes never read), on
float32 bitshuffled and float64 shuffled operands, and a request above the
array's size is refused before any launch.
"""
import ctypes

import numpy as np
import pytest

import pyiarray as ia

pytestmark = pytest.mark.gpu


class DevicePlan(ctypes.Structure):
    _fields_ = [("lz4", ctypes.c_void_p), ("n_lz4", ctypes.c_int), ("blocks", ctypes.c_void_p),
                ("blk_rc", ctypes.c_void_p), ("n_blocks", ctypes.c_int), ("nbcols", ctypes.c_int),
                ("scratch_bytes", ctypes.c_int64), ("blockshape", ctypes.c_int64 * 2)]


def _hip():
    h = ctypes.CDLL(ia.HIP_LIB_PATH)
    h.iamx_malloc.restype = ctypes.c_void_p
    h.iamx_malloc.argtypes = [ctypes.c_size_t]
    h.iamx_free.argtypes = [ctypes.c_void_p]
    h.iamx_workspace.restype = ctypes.c_void_p
    h.iamx_workspace.argtypes = [ctypes.c_int, ctypes.c_size_t]
    for f in ("iamx_h2d", "iamx_d2h"):
        getattr(h, f).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    h.iamx_memset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    h.iamx_last_error.restype = ctypes.c_char_p
    h.iamx_gemv_fused_noalloc_limit.restype = ctypes.c_int64
    h.iamx_gemv_fused.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
    h.iamx_gemv_fused_noalloc.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                          ctypes.c_int64]
    return h


def _plan(c):
    p = DevicePlan()
    rc = ia.lib().iarray_container_device_plan(c.ptr, ctypes.cast(ctypes.byref(p), ctypes.c_void_p))
    assert rc == 0, ia.lib().iarray_err_strerror(rc)
    return p


def _run_both(h, ctx, A, x, chunks, blocks):
    """y from iamx_gemv_fused (library workspaces) and from
    iamx_gemv_fused_noalloc, same plan and x."""
    dt = 0 if A.dtype == np.float64 else 1
    M, K = A.shape
    cA = ia.Container.from_numpy(ctx, A, chunks, blocks)
    p = _plan(cA)
    isz = A.itemsize
    dx = h.iamx_malloc(K * isz)
    dy1 = h.iamx_malloc(M * isz)
    dy2 = h.iamx_malloc(M * isz)
    try:
        assert h.iamx_h2d(dx, x.ctypes.data, K * isz) == 0
        scratch = h.iamx_workspace(0, p.scratch_bytes + 16)
        partial = h.iamx_workspace(6, p.nbcols * M * isz + 16)
        assert h.iamx_memset(partial, 0, p.nbcols * M * isz) == 0
        rc = h.iamx_gemv_fused(dt, p.lz4, p.n_lz4, p.blocks, p.blk_rc, p.n_blocks, p.nbcols, p.blockshape[0],
                               p.blockshape[1], dx, dy1, M, scratch, partial)
        assert rc == 0, h.iamx_last_error()
        rc = h.iamx_gemv_fused_noalloc(dt, p.lz4, p.n_lz4, p.blocks, p.blk_rc, p.n_blocks, p.nbcols,
                                       p.blockshape[0], p.blockshape[1], dx, K, dy2, M, p.scratch_bytes)
        assert rc == 0, h.iamx_last_error()
        assert h.iamx_sync() == 0
        assert h.iamx_decode_status() == 0
        y1 = np.empty(M, A.dtype)
        y2 = np.empty(M, A.dtype)
        assert h.iamx_d2h(y1.ctypes.data, dy1, M * isz) == 0
        assert h.iamx_d2h(y2.ctypes.data, dy2, M * isz) == 0
        return y1, y2
    finally:
        for d in (dx, dy1, dy2):
            h.iamx_free(d)


def test_noalloc_gemv_matches_fused_twice():
    h = _hip()
    assert h.iamx_init() == 0
    rng = np.random.default_rng(12)
    cases = [
        (np.float32, ia.COMP_BITSHUFFLE, (1500, 1300), (500, 650), (250, 130)),
        (np.float32, ia.COMP_BITSHUFFLE, (700, 400), (350, 400), (70, 100)),   # smaller second call
        (np.float64, ia.COMP_SHUFFLE, (900, 800), (300, 400), (100, 80)),
        (np.float64, ia.COMP_SHUFFLE, (130, 90), (64, 45), (16, 15)),          # smaller again, ragged
    ]
    for dt, filt, shape, chunks, blocks in cases:
        print(f"case: noalloc gemv {np.dtype(dt).name} {shape} chunks {chunks} blocks {blocks}", flush=True)
        ctx = ia.Context(clevel=5, filters=filt)
        A = (rng.integers(-8, 9, shape) + 0.25 * rng.integers(0, 4, shape)).astype(dt)
        x = rng.integers(-4, 5, shape[1]).astype(dt)
        y1, y2 = _run_both(h, ctx, A, x, chunks, blocks)
        assert y2.tobytes() == y1.tobytes()
        assert np.array_equal(y2, A @ x)  # quarter-integer data: every order is exact


def test_noalloc_gemv_refuses_oversize():
    h = _hip()
    assert h.iamx_init() == 0
    limit = h.iamx_gemv_fused_noalloc_limit()
    assert limit == 64 << 20
    ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
    A = np.ones((256, 256), np.float32)
    cA = ia.Container.from_numpy(ctx, A, (256, 256), (64, 64))
    p = _plan(cA)
    dx = h.iamx_malloc(256 * 4)
    dy = h.iamx_malloc(256 * 4)
    try:
        # decoded streams alone above the limit
        rc = h.iamx_gemv_fused_noalloc(1, p.lz4, p.n_lz4, p.blocks, p.blk_rc, p.n_blocks, p.nbcols, 64, 64, dx, 256,
                                       dy, 256, limit + 1)
        assert rc == -3 and b"scratch" in h.iamx_last_error()
        # partial sums pushing past it
        rc = h.iamx_gemv_fused_noalloc(1, p.lz4, p.n_lz4, p.blocks, p.blk_rc, p.n_blocks, 1 << 20, 64, 64, dx, 256,
                                       dy, 256, 1 << 20)
        assert rc == -3
        assert h.iamx_sync() == 0
        assert h.iamx_decode_status() == 0
    finally:
        h.iamx_free(dx)
        h.iamx_free(dy)
