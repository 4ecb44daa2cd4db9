"""The reference's own linalg fixtures restated with their storage: operands
and results that the reference keeps on disk (storage.urlpath; a sparse frame
directory when contiguous is false, a frame file when true) are kept on disk
here too, and the result is read back from its path.

Fixtures (shapes, chunk/block shapes, generators, storage, tolerance):
  tests/test_linalg_gemm.c:176-271        arange operands, 4 cases
  tests/test_linalg_gemv.c:176-246        linspace(0, 10) operands, 3 cases
  tests/test_linalg_gemv_transpose.c:184-206  x transposed, z on disk
  tests/test_opt_gemv.c:179-225           linspace(0, 10), opt_gemv
  examples/example_matmul.c:25-50,114-144 opt_gemv 2000^2 float64, clevel 9,
                                          x = linspace(3, 2), y = linspace(2, 10),
                                          tolerance 1e-8 against dgemv
Each result is checked against the oracle's restatement of the reference
path on the same compressed operands, elementwise |z - o| / |z| at the
reference's tolerance (1e-14 float64, 1e-5 float32,
tests/test_linalg_gemm.c:128-147), and the file or directory at the result's
urlpath loads back to the same bits.
"""
import os

import numpy as np
import pytest

import pyiarray as ia

pytestmark = pytest.mark.gpu

TOL = {np.dtype(np.float64): 1e-14, np.dtype(np.float32): 1e-5}


def _rel_ok(z, o, tol):
    z = z.astype(np.float64)
    o = o.astype(np.float64)
    nz = z != 0
    assert np.array_equal(z[~nz], o[~nz])
    rel = np.abs((z[nz] - o[nz]) / z[nz])
    assert rel.max(initial=0) <= tol, f"max rel err {rel.max():.3e} > {tol}"


def _path(tmp, name):
    return None if name is None else str(tmp / name)


def _oracle(orc, c, cs, bs, transposed=False, shape=None):
    return orc.Array(c.chunks(), shape or c.shape, cs, bs, c.dtype.itemsize, transposed=transposed)


def _check_stored(ctx, z, path, got):
    if path is None:
        return
    assert os.path.exists(path)
    back = ia.Container.load(ctx, path)
    assert back.shape == z.shape and back.chunks() == z.chunks()
    assert back.to_numpy().tobytes() == got.tobytes()


GEMM = [
    # name, dtype, x (shape, cs, bs), y, z (cs, bs), (x, y, z) storage: (contiguous, urlpath)
    ("f_schunk_schunk", np.float32, ((100, 170), (27, 30), (12, 7)), ((170, 21), (20, 11), (10, 5)),
     ((40, 3), (20, 3)), ((False, "xarr.iarr"), (False, "yarr.iarr"), (False, "zarr.iarr"))),
    ("d_schunk_schunk", np.float64, ((300, 670), (28, 300), (11, 31)), ((670, 210), (20, 45), (7, 12)),
     ((43, 11), (13, 5)), ((True, None), (True, None), (True, None))),
    ("f_schunk_schunk2", np.float32, ((100, 200), (10, 30), (10, 10)), ((200, 150), (25, 30), (15, 12)),
     ((20, 30), (11, 6)), ((False, "xarr.iarr"), (True, "yarr.iarr"), (False, "zarr.iarr"))),
    ("d_schunk_schunk2", np.float64, ((100, 200), (10, 30), (7, 7)), ((200, 150), (25, 30), (5, 11)),
     ((20, 30), (17, 7)), ((True, None), (True, "yarr.iarr"), (False, "zarr.iarr"))),
]


@pytest.mark.parametrize("fx", GEMM, ids=[f[0] for f in GEMM])
def test_linalg_gemm_fixtures_with_storage(orc, tmp_path, fx):
    name, dt, (xs, xc, xb), (ys, yc, yb), (zc, zb), ((xct, xp), (yct, yp), (zct, zp)) = fx
    print(f"case: test_linalg_gemm {name}", flush=True)
    ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
    xp, yp, zp = _path(tmp_path, xp), _path(tmp_path, yp), _path(tmp_path, zp)
    x = ia.Container.arange(ctx, dt, xs, xc, xb, urlpath=xp, contiguous=xct)
    y = ia.Container.arange(ctx, dt, ys, yc, yb, urlpath=yp, contiguous=yct)
    z = ia.matmul(x, y, zc, zb, urlpath=zp, contiguous=zct)
    got = z.to_numpy()
    o = orc.gemm_blosc(_oracle(orc, x, xc, xb), _oracle(orc, y, yc, yb), zc, zb)
    _rel_ok(got, o, TOL[np.dtype(dt)])
    _check_stored(ctx, z, zp, got)
    # operands kept on disk load back to the containers used
    for c, p in ((x, xp), (y, yp)):
        if p:
            assert ia.Container.load(ctx, p).chunks() == c.chunks()


GEMV = [
    ("d_schunk", np.float32, ((100, 200), (30, 30), (10, 15)), ((200,), (125,), (45,)), ((80,), (37,)),
     ((False, "xarr.iarr"), (False, None), (True, None))),
    ("f_schunk_schunk", np.float64, ((650, 900), (200, 140), (20, 14)), ((900,), (421,), (55,)), ((155,), (47,)),
     ((True, None), (True, "yarr.iarr"), (False, "zarr.iarr"))),
    ("d_schunk_schunk", np.float64, ((60, 300), (14, 200), (5, 14)), ((300,), (41,), (17,)), ((18,), (18,)),
     ((True, None), (False, None), (True, "zarr.iarr"))),
]


@pytest.mark.parametrize("fx", GEMV, ids=[f[0] for f in GEMV])
def test_linalg_gemv_fixtures_with_storage(orc, tmp_path, fx):
    name, dt, (xs, xc, xb), (ys, yc, yb), (zc, zb), ((xct, xp), (yct, yp), (zct, zp)) = fx
    print(f"case: test_linalg_gemv {name}", flush=True)
    ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
    xp, yp, zp = _path(tmp_path, xp), _path(tmp_path, yp), _path(tmp_path, zp)
    x = ia.Container.linspace(ctx, dt, xs, xc, xb, 0.0, 10.0, urlpath=xp, contiguous=xct)
    y = ia.Container.linspace(ctx, dt, ys, yc, yb, 0.0, 10.0, urlpath=yp, contiguous=yct)
    z = ia.matmul(x, y, zc, zb, urlpath=zp, contiguous=zct)
    got = z.to_numpy()
    o = orc.gemv_blosc(_oracle(orc, x, xc, xb), _oracle(orc, y, yc, yb), zc, zb)
    _rel_ok(got, o, TOL[np.dtype(dt)])
    _check_stored(ctx, z, zp, got)


def test_linalg_gemv_transpose_fixture(orc, tmp_path):
    """tests/test_linalg_gemv_transpose.c:184-206: x (200 x 100 float32,
    linspace(0, 10)) is transposed before the product; z on disk."""
    ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
    xt = ia.Container.linspace(ctx, np.float32, (200, 100), (30, 30), (15, 10), 0.0, 10.0)
    y = ia.Container.linspace(ctx, np.float32, (200,), (125,), (45,), 0.0, 10.0)
    zp = str(tmp_path / "zarr.iarr")
    z = ia.matmul(xt.transpose(), y, (80,), (37,), urlpath=zp, contiguous=True)
    got = z.to_numpy()
    o = orc.gemv_blosc(_oracle(orc, xt, (30, 30), (15, 10), transposed=True, shape=(200, 100)),
                       _oracle(orc, y, (125,), (45,)), (80,), (37,))
    _rel_ok(got, o, 1e-5)
    _check_stored(ctx, z, zp, got)


OPT_GEMV = [
    ("f_schunk", np.float32, ((6000, 4000), (2000, 1000), (200, 200)), ((4000,), (1000,), (200,)),
     ((2000,), (200,)), ((False, None), (True, None), (True, "zarr.iarr"))),
    ("d_schunk", np.float64, ((8000, 2000), (400, 300), (200, 150)), ((2000,), (300,), (150,)),
     ((400,), (200,)), ((True, "xarr.iarr"), (False, "yarr.iarr"), (False, "zarr.iarr"))),
]


@pytest.mark.parametrize("fx", OPT_GEMV, ids=[f[0] for f in OPT_GEMV])
def test_opt_gemv_fixtures_with_storage(orc, tmp_path, fx):
    name, dt, (xs, xc, xb), (ys, yc, yb), (zc, zb), ((xct, xp), (yct, yp), (zct, zp)) = fx
    print(f"case: test_opt_gemv {name}", flush=True)
    ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
    xp, yp, zp = _path(tmp_path, xp), _path(tmp_path, yp), _path(tmp_path, zp)
    x = ia.Container.linspace(ctx, dt, xs, xc, xb, 0.0, 10.0, urlpath=xp, contiguous=xct)
    y = ia.Container.linspace(ctx, dt, ys, yc, yb, 0.0, 10.0, urlpath=yp, contiguous=yct)
    z = ia.opt_gemv(x, y, zc, zb, urlpath=zp, contiguous=zct)
    got = z.to_numpy()
    o = orc.opt_gemv(_oracle(orc, x, xc, xb), _oracle(orc, y, yc, yb), zc, zb)
    _rel_ok(got, o, TOL[np.dtype(dt)])
    _check_stored(ctx, z, zp, got)


def test_example_matmul_opt_gemv(orc):
    """examples/example_matmul.c: opt_gemv of a 2000 x 2000 float64 linspace(3, 2)
    by a linspace(2, 10) vector, chunks 1000, blocks 200, LZ4 clevel 9; each
    entry within 1e-8 (relative) of dgemv on the decoded buffers."""
    ctx = ia.Context(clevel=9, filters=ia.COMP_SHUFFLE)
    x = ia.Container.linspace(ctx, np.float64, (2000, 2000), (1000, 1000), (200, 200), 3.0, 2.0)
    y = ia.Container.linspace(ctx, np.float64, (2000,), (1000,), (200,), 2.0, 10.0)
    z = ia.opt_gemv(x, y, (1000,), (200,)).to_numpy()
    bx, by = x.to_numpy(), y.to_numpy()
    assert np.array_equal(bx.ravel(), np.linspace(3.0, 2.0, 2000 * 2000)) or np.allclose(
        bx.ravel(), np.linspace(3.0, 2.0, 2000 * 2000), rtol=0, atol=4e-16)
    want = bx @ by
    assert np.max(np.abs((z - want) / z)) <= 1e-8
    o = orc.opt_gemv(_oracle(orc, x, (1000, 1000), (200, 200)), _oracle(orc, y, (1000,), (200,)), (1000,), (200,))
    _rel_ok(z, o, 1e-14)
