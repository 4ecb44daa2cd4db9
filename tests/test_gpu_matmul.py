"""GPU parity of iarray_linalg_matmul / iarray_opt_gemm / iarray_opt_gemm_a /
iarray_opt_gemm_b / iarray_opt_gemv.

Oracle: oracle/ (gemm_blosc, gemv_blosc, opt_gemm, opt_gemv restated on the
CPU, plain-C BLAS), run on the same compressed containers.  Bars:
  * the reference's own test fixtures (tests/test_linalg_gemm.c:180-271,
    test_linalg_gemv.c:200-222, test_opt_gemv.c:179-225,
    test_linalg_gemm_transpose.c:193-215) with their generators (arange /
    linspace) and tolerance, elementwise |z-o|/|z| <= 1e-14 (float64) or
    1e-5 (float32) (tests/test_linalg_gemm.c:128-147);
  * integer-valued operands, where every summation order is exact: bit-exact;
  * determinism: the same call twice gives the same bits.
"""
import numpy as np
import pytest

import pyiarray as ia

pytestmark = pytest.mark.gpu

TOL = {np.dtype(np.float64): 1e-14, np.dtype(np.float32): 1e-5}


@pytest.fixture(scope="module")
def ctx():
    return ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)


def _oracle(orc, c):
    return orc.Array(c.chunks(), c.shape, c._cs, c._bs, c.dtype.itemsize)


def _mk(ctx, gen, dtype, shape, cs, bs, **kw):
    if gen == "arange":
        c = ia.Container.arange(ctx, dtype, shape, cs, bs, **kw)
    else:
        c = ia.Container.linspace(ctx, dtype, shape, cs, bs, **kw)
    c._cs, c._bs = cs, bs
    return c


def _rel_ok(z, o, tol):
    z = z.astype(np.float64)
    o = o.astype(np.float64)
    nz = z != 0
    assert np.array_equal(z[~nz], o[~nz]) or np.allclose(z[~nz], o[~nz], atol=0)
    rel = np.abs((z[nz] - o[nz]) / z[nz])
    assert rel.max(initial=0) <= tol, f"max rel err {rel.max():.3e} > {tol}"


GEMM_FIXTURES = [
    # name, dtype, x(shape, cs, bs), y(shape, cs, bs), z(cs, bs)   -- test_linalg_gemm.c
    ("f_schunk_schunk", np.float32, ((100, 170), (27, 30), (12, 7)), ((170, 21), (20, 11), (10, 5)), ((40, 3), (20, 3))),
    ("d_schunk_schunk", np.float64, ((300, 670), (28, 300), (11, 31)), ((670, 210), (20, 45), (7, 12)),
     ((43, 11), (13, 5))),
    ("f_schunk_schunk2", np.float32, ((100, 200), (10, 30), (10, 10)), ((200, 150), (25, 30), (15, 12)),
     ((20, 30), (11, 6))),
    ("d_schunk_schunk2", np.float64, ((100, 200), (10, 30), (7, 7)), ((200, 150), (25, 30), (5, 11)),
     ((20, 30), (17, 7))),
]


@pytest.mark.parametrize("fx", GEMM_FIXTURES, ids=[f[0] for f in GEMM_FIXTURES])
def test_matmul_reference_gemm_fixtures(ctx, orc, fx):
    name, dt, (xs, xc, xb), (ys, yc, yb), (zc, zb) = fx
    print(f"case: gemm fixture {name}", flush=True)
    x = _mk(ctx, "arange", dt, xs, xc, xb)
    y = _mk(ctx, "arange", dt, ys, yc, yb)
    z = ia.matmul(x, y, zc, zb)
    got = z.to_numpy()
    o = orc.gemm_blosc(_oracle(orc, x), _oracle(orc, y), zc, zb)
    _rel_ok(got, o, TOL[np.dtype(dt)])
    # the oracle decoder reads the C chunks the GPU path wrote
    oz = orc.Array(z.chunks(), z.shape, zc, zb, z.dtype.itemsize)
    assert oz.to_dense().tobytes() == got.tobytes()


def test_matmul_transposed_fixture(ctx, orc):
    """test_linalg_gemm_transpose.c:193-215: x and y are transposed views."""
    xt = _mk(ctx, "arange", np.float64, (170, 100), (20, 30), (10, 12))
    yt = _mk(ctx, "arange", np.float64, (21, 170), (11, 20), (5, 10))
    x, y = xt.transpose(), yt.transpose()
    z = ia.matmul(x, y, (40, 3), (20, 3))
    xd = orc.Array(xt.chunks(), (170, 100), (20, 30), (10, 12), 8, transposed=True)
    yd = orc.Array(yt.chunks(), (21, 170), (11, 20), (5, 10), 8, transposed=True)
    o = orc.gemm_blosc(xd, yd, (40, 3), (20, 3))
    _rel_ok(z.to_numpy(), o, 1e-14)


GEMV_FIXTURES = [
    # test_linalg_gemv.c:200-222 style: x(shape, cs, bs), y(shape, cs, bs), z(cs, bs)
    ("gemv_f", np.float32, ((650, 900), (200, 140), (20, 14)), ((900,), (421,), (55,)), ((155,), (47,))),
    ("gemv_d", np.float64, ((650, 900), (200, 140), (20, 14)), ((900,), (421,), (55,)), ((155,), (47,))),
]


@pytest.mark.parametrize("fx", GEMV_FIXTURES, ids=[f[0] for f in GEMV_FIXTURES])
def test_matmul_gemv_fixtures(ctx, orc, fx):
    name, dt, (xs, xc, xb), (ys, yc, yb), (zc, zb) = fx
    print(f"case: gemv fixture {name}", flush=True)
    x = _mk(ctx, "linspace", dt, xs, xc, xb, start=0.0, stop=10.0)
    y = _mk(ctx, "linspace", dt, ys, yc, yb, start=0.0, stop=10.0)
    z = ia.matmul(x, y, zc, zb)
    o = orc.gemv_blosc(_oracle(orc, x), _oracle(orc, y), zc, zb)
    _rel_ok(z.to_numpy(), o, TOL[np.dtype(dt)])


OPT_GEMV = [
    # test_opt_gemv.c:179-225
    ("opt_gemv_f", np.float32, ((6000, 4000), (2000, 1000), (200, 200)), ((4000,), (1000,), (200,)),
     ((2000,), (200,))),
    ("opt_gemv_d", np.float64, ((8000, 2000), (400, 300), (200, 150)), ((2000,), (300,), (150,)),
     ((400,), (200,))),
]


@pytest.mark.parametrize("fx", OPT_GEMV, ids=[f[0] for f in OPT_GEMV])
def test_opt_gemv_fixtures(ctx, orc, fx):
    name, dt, (xs, xc, xb), (ys, yc, yb), (zc, zb) = fx
    print(f"case: opt_gemv fixture {name}", flush=True)
    x = _mk(ctx, "linspace", dt, xs, xc, xb, start=0.0, stop=10.0)
    y = _mk(ctx, "linspace", dt, ys, yc, yb, start=0.0, stop=10.0)
    z = ia.opt_gemv(x, y, zc, zb)
    o = orc.opt_gemv(_oracle(orc, x), _oracle(orc, y), zc, zb, nthreads=8)
    _rel_ok(z.to_numpy(), o, TOL[np.dtype(dt)])


def test_opt_gemm_matches_oracle(ctx, orc):
    x = _mk(ctx, "arange", np.float64, (300, 250), (100, 90), (50, 30))
    y = _mk(ctx, "arange", np.float64, (250, 170), (90, 60), (30, 20))
    z = ia.opt_gemm(x, y, (100, 60), (50, 20))
    o = orc.opt_gemm(_oracle(orc, x), _oracle(orc, y), (100, 60), (50, 20))
    _rel_ok(z.to_numpy(), o, 1e-14)


# (variant, M, K, N, A chunk/block, B chunk/block, C chunk/block); padded edges
AB_CASES = [
    ("a", 300, 250, 170, ((100, 250), (100, 50)), ((250, 60), (50, 60)), ((100, 170), (100, 60))),
    ("a", 128, 96, 200, ((64, 96), (64, 32)), ((96, 100), (32, 100)), ((64, 200), (64, 100))),
    ("b", 300, 250, 170, ((100, 250), (100, 50)), ((250, 60), (50, 60)), ((300, 60), (100, 60))),
    ("b", 128, 96, 200, ((32, 96), (32, 48)), ((96, 80), (48, 80)), ((128, 80), (32, 80))),
]


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("case", AB_CASES, ids=[f"gemm_{c[0]}-{c[1]}x{c[2]}x{c[3]}" for c in AB_CASES])
def test_opt_gemm_ab_matches_oracle(orc, dt, case):
    """iarray_opt_gemm_a / _b against the oracle's restatement of the
    reference's block loops: linspace operands at the reference tests'
    tolerances, integer-valued operands exactly."""
    var, M, K, N, (acs, abs_), (bcs, bbs), (ccs, cbs) = case
    print(f"case: opt_gemm_{var} {M}x{K}x{N} {np.dtype(dt).name}", flush=True)
    ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
    x = _mk(ctx, "linspace", dt, (M, K), acs, abs_, start=-1.0, stop=2.0)
    y = _mk(ctx, "linspace", dt, (K, N), bcs, bbs, start=0.5, stop=3.0)
    f = ia.opt_gemm_a if var == "a" else ia.opt_gemm_b
    of = orc.opt_gemm_a if var == "a" else orc.opt_gemm_b
    z = f(x, y, ccs, cbs)
    _rel_ok(z.to_numpy(), of(_oracle(orc, x), _oracle(orc, y), ccs, cbs), TOL[np.dtype(dt)])
    rng = np.random.default_rng(M * 5 + N)
    a = rng.integers(-8, 9, (M, K)).astype(dt)
    b = rng.integers(-8, 9, (K, N)).astype(dt)
    ca = ia.Container.from_numpy(ctx, a, acs, abs_)
    cb = ia.Container.from_numpy(ctx, b, bcs, bbs)
    got = f(ca, cb, ccs, cbs).to_numpy()
    assert np.array_equal(got, (a.astype(np.float64) @ b.astype(np.float64)).astype(dt))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(129, 257, 65), (300, 200, 500), (64, 64, 64), (1, 300, 7)])
def test_integer_operands_bit_exact(ctx, dt, shape):
    M, K, N = shape
    rng = np.random.default_rng(M * 7 + K)
    a = rng.integers(-8, 9, (M, K)).astype(dt)
    b = rng.integers(-8, 9, (K, N)).astype(dt)
    ca = ia.Container.from_numpy(ctx, a, (min(M, 100), min(K, 96)), (min(M, 50), min(K, 32)))
    cb = ia.Container.from_numpy(ctx, b, (min(K, 96), min(N, 80)), (min(K, 32), min(N, 40)))
    cc = (min(M, 64), min(N, 64))
    z = ia.matmul(ca, cb, cc, (min(M, 32), min(N, 16)))
    want = (a.astype(np.float64) @ b.astype(np.float64)).astype(dt)
    assert np.array_equal(z.to_numpy(), want)
    x = rng.integers(-8, 9, (K,)).astype(dt)
    cx = ia.Container.from_numpy(ctx, x, (min(K, 96),), (min(K, 32),))
    y = ia.matmul(ca, cx, (min(M, 64),), (min(M, 32),))
    assert np.array_equal(y.to_numpy(), (a.astype(np.float64) @ x.astype(np.float64)).astype(dt))


@pytest.mark.parametrize("form", ["f32", "f64", "split_bf16", "ozaki_int8"])
@pytest.mark.parametrize("panel_rows", ["1", "150"])
def test_row_panels_bit_exact(ctx, monkeypatch, form, panel_rows):
    """C computed in row panels (one chunk row each, or three), ragged M and
    chunk rows, equals the one-panel product bit for bit; integer operands
    make every form exact, so both also equal the exact product."""
    dt = np.float32 if form in ("f32", "split_bf16") else np.float64
    M, K, N = 333, 190, 170
    print(f"case: row panels {form} panel_rows={panel_rows}", flush=True)
    rng = np.random.default_rng(len(form) * 31 + int(panel_rows))
    a = rng.integers(-8, 9, (M, K)).astype(dt)
    b = rng.integers(-8, 9, (K, N)).astype(dt)
    ca = ia.Container.from_numpy(ctx, a, (64, 96), (32, 32))
    cb = ia.Container.from_numpy(ctx, b, (96, 80), (32, 40))
    kw = {"split_bf16": {"precision": ia.SPLIT_BF16}, "ozaki_int8": {"precision": ia.OZAKI_INT8, "slices": 3}}
    kw = kw.get(form, {})
    monkeypatch.setenv("IARRAY_GEMM_PANEL_ROWS", "0")
    one = ia.matmul(ca, cb, (64, 64), (32, 16), **kw).to_numpy()
    monkeypatch.setenv("IARRAY_GEMM_PANEL_ROWS", panel_rows)
    z = ia.matmul(ca, cb, (64, 64), (32, 16), **kw)
    got = z.to_numpy()
    want = (a.astype(np.float64) @ b.astype(np.float64)).astype(dt)
    assert np.array_equal(one, want)
    assert got.tobytes() == one.tobytes()


@pytest.mark.parametrize("kind", ["gemv", "gemm-f32", "gemm-f64-panels"])
def test_direct_and_pipelined_writes_agree(monkeypatch, kind):
    """Small C packed straight into mapped host memory and C downloaded
    through the copy stream give the same chunks, byte for byte."""
    print(f"case: direct vs pipelined write {kind}", flush=True)
    rng = np.random.default_rng(11)
    ctx = ia.Context(clevel=5, filters=ia.COMP_BITSHUFFLE if kind == "gemv" else ia.COMP_SHUFFLE)
    if kind == "gemv":
        a = rng.standard_normal((700, 500)).astype(np.float32)
        x = rng.standard_normal(500).astype(np.float32)
        ca = ia.Container.from_numpy(ctx, a, (256, 250), (128, 125))
        cx = ia.Container.from_numpy(ctx, x, (250,), (125,))
        run = lambda: ia.matmul(ca, cx, (256,), (64,))
    else:
        dt = np.float32 if kind == "gemm-f32" else np.float64
        a = rng.standard_normal((300, 200)).astype(dt)
        b = rng.standard_normal((200, 260)).astype(dt)
        ca = ia.Container.from_numpy(ctx, a, (128, 100), (64, 50))
        cb = ia.Container.from_numpy(ctx, b, (100, 130), (50, 65))
        if kind == "gemm-f64-panels":
            monkeypatch.setenv("IARRAY_GEMM_PANEL_ROWS", "1")
        run = lambda: ia.matmul(ca, cb, (128, 130), (64, 65))
    monkeypatch.setenv("IARRAY_DIRECT_WRITE", "0")
    piped = run()
    monkeypatch.setenv("IARRAY_DIRECT_WRITE", "1")
    direct = run()
    assert direct.chunks() == piped.chunks()
    want = a.astype(np.float64) @ (x if kind == "gemv" else b).astype(np.float64)
    np.testing.assert_allclose(direct.to_numpy(), want, rtol=1e-4 if a.dtype == np.float32 else 1e-12,
                               atol=1e-3 if a.dtype == np.float32 else 1e-10)


def test_bitshuffled_gemv_bit_exact(ctx):
    """The headline GEMV layout (float32, LZ4 + bitshuffle, 250x250 blocks,
    62500 elements per block so every block has a 4-element tail)."""
    bctx = ia.Context(clevel=5, filters=ia.COMP_BITSHUFFLE)
    rng = np.random.default_rng(11)
    a = rng.integers(-4, 5, (1000, 750)).astype(np.float32)
    x = rng.integers(-4, 5, (750,)).astype(np.float32)
    ca = ia.Container.from_numpy(bctx, a, (500, 500), (250, 250))
    cx = ia.Container.from_numpy(bctx, x, (500,), (250,))
    y = ia.matmul(ca, cx, (500,), (250,))
    assert np.array_equal(y.to_numpy(), (a.astype(np.float64) @ x.astype(np.float64)).astype(np.float32))


BITS_GEMV = [
    # dtype, A shape, A chunks, A blocks: the quad path (whole quads, short last
    # quads with plane groups and unshuffled tails, unaligned planes) and the
    # group path (rows shorter than 32 elements)
    (np.float32, (300, 410), (150, 205), (33, 40)),
    (np.float32, (200, 150), (100, 150), (37, 50)),
    (np.float64, (500, 600), (250, 300), (125, 125)),
    (np.float64, (400, 256), (200, 256), (64, 64)),
    (np.float32, (100, 64), (100, 64), (100, 20)),
    (np.float32, (12, 3000), (8, 2000), (4, 1000)),
]


@pytest.mark.parametrize("dt,shape,cs,bs", BITS_GEMV, ids=[f"{np.dtype(d).name}-{b[0]}x{b[1]}" for d, _, _, b in BITS_GEMV])
def test_bitshuffled_gemv_block_shapes(dt, shape, cs, bs):
    print(f"case: bitshuffle gemv {np.dtype(dt).name} {shape} {cs} {bs}", flush=True)
    bctx = ia.Context(clevel=5, filters=ia.COMP_BITSHUFFLE)
    rng = np.random.default_rng(shape[0] + bs[1])
    a = rng.integers(-6, 7, shape).astype(dt)
    x = rng.integers(-6, 7, (shape[1],)).astype(dt)
    ca = ia.Container.from_numpy(bctx, a, cs, bs)
    cx = ia.Container.from_numpy(bctx, x, (cs[1],), (bs[1],))
    y = ia.matmul(ca, cx, (cs[0],), (bs[0],))
    assert np.array_equal(y.to_numpy(), (a.astype(np.float64) @ x.astype(np.float64)).astype(dt))


@pytest.mark.parametrize("share", ["0", "1"])
def test_matmul_through_frames(ctx, tmp_path, monkeypatch, share):
    """The reference's persistent gemm fixtures (test_linalg_gemm.c:227-248,
    x, y and z stored as frames): the operands are saved and loaded back,
    multiplied, and the product goes through a frame file and an in-memory
    frame; every copy decodes to the product of the in-memory operands, with
    C's chunks written by the host encoder (share 0) or the GPU encoder
    (share 1)."""
    monkeypatch.setenv("IARRAY_GPU_ENCODE", share)
    name, dt, (xs, xcs, xbs), (ys, ycs, ybs), (zcs, zbs) = GEMM_FIXTURES[2]
    print(f"case: frames {name} share {share}", flush=True)
    x = _mk(ctx, "arange", dt, xs, xcs, xbs)
    y = _mk(ctx, "arange", dt, ys, ycs, ybs)
    x.save(str(tmp_path / "xarr.iarray"))
    y.save(str(tmp_path / "yarr.iarray"))
    xl = ia.Container.load(ctx, str(tmp_path / "xarr.iarray"))
    yl = ia.Container.load(ctx, str(tmp_path / "yarr.iarray"), lazy=True)
    assert xl.chunks() == x.chunks() and yl.chunks() == y.chunks()
    z = ia.matmul(x, y, zcs, zbs).to_numpy()
    zc = ia.matmul(xl, yl, zcs, zbs)
    zc.save(str(tmp_path / "zarr.iarray"))
    zl = ia.Container.load(ctx, str(tmp_path / "zarr.iarray"))
    zf = ia.Container.from_cframe(ctx, zc.to_cframe())
    for got in (zc, zl, zf):
        assert got.to_numpy().tobytes() == z.tobytes()
    o = np.arange(np.prod(xs), dtype=np.float64).reshape(xs) @ np.arange(np.prod(ys), dtype=np.float64).reshape(ys)
    _rel_ok(z, o, TOL[np.dtype(dt)])
    ia.remove(str(tmp_path / "zarr.iarray"))
    assert not (tmp_path / "zarr.iarray").exists()


def test_slice_views_decode_and_multiply(ctx):
    """iarray_get_slice (src/iarray_container.c:364-455) on the GPU: slice
    views (negative bounds, a slice of a slice, a 1-D slice) decode to the
    numpy slices and multiply as operands of the gemm and gemv branches; a
    copy is a new container with the storage asked for.  Integer-valued
    float64 operands make every product exact."""
    rng = np.random.default_rng(21)
    a = rng.integers(-8, 9, (300, 260)).astype(np.float64)
    b = rng.integers(-8, 9, (260, 170)).astype(np.float64)
    x = rng.integers(-8, 9, 260).astype(np.float64)
    ca = ia.Container.from_numpy(ctx, a, (100, 90), (25, 30))
    cb = ia.Container.from_numpy(ctx, b, (90, 80), (30, 20))
    cx = ia.Container.from_numpy(ctx, x, (90,), (30,))
    print("case: slice views", flush=True)
    sa, sb, sx = a[13:-21, 7:207], b[7:207, 3:-5], x[7:207]
    va = ca.get_slice((13, 7), (-21, 207))
    vb = cb.get_slice((7, 3), (207, -5))
    vx = cx.get_slice((7,), (207,))
    assert va.shape == sa.shape and vb.shape == sb.shape and vx.shape == sx.shape
    assert va.to_numpy().tobytes() == np.ascontiguousarray(sa).tobytes()
    assert vx.to_numpy().tobytes() == np.ascontiguousarray(sx).tobytes()
    assert np.array_equal(vb.slice((5, 9), (50, 100)), sb[5:50, 9:100])
    w = va.get_slice((10, 20), (60, 90))
    assert w.to_numpy().tobytes() == np.ascontiguousarray(sa[10:60, 20:90]).tobytes()
    z = ia.matmul(va, vb, (100, 60), (50, 30)).to_numpy()
    assert np.array_equal(z, sa @ sb)
    y = ia.matmul(va, vx, (100,), (50,)).to_numpy()
    assert np.array_equal(y, sa @ sx)
    st = cb.transpose().get_slice((3, 7), (-5, 207))  # (b.T)[3:-5, 7:207] = (b[7:207, 3:-5]).T
    assert st.shape == sb.T.shape
    assert st.to_numpy().tobytes() == np.ascontiguousarray(sb.T).tobytes()
    st2 = st.get_slice((4, 10), (60, 150))
    assert st2.to_numpy().tobytes() == np.ascontiguousarray(sb.T[4:60, 10:150]).tobytes()
    zt = ia.matmul(st, va.get_slice((0, 0), (200, 100)), (60, 50), (30, 25)).to_numpy()
    assert np.array_equal(zt, sb.T @ sa[:200, :100])
    cp = ca.get_slice((13, 7), (-21, 207), view=False, chunks=(64, 64), blocks=(32, 32))
    assert cp.to_numpy().tobytes() == np.ascontiguousarray(sa).tobytes()
    z2 = ia.matmul(cp, vb, (100, 60), (50, 30)).to_numpy()
    assert z2.tobytes() == z.tobytes()


def test_operand_image_cache_follows_containers(ctx):
    """A plain operand decoded into a workspace is reused only while the
    same, unchanged container is asked for again in the same slot: other
    vectors, a freed and reallocated container, a view through the same slot
    and a gemm in between all give exact results."""
    rng = np.random.default_rng(31)
    a = rng.integers(-8, 9, (300, 260)).astype(np.float64)
    b = rng.integers(-8, 9, (260, 120)).astype(np.float64)
    ca = ia.Container.from_numpy(ctx, a, (100, 90), (25, 30))
    cb = ia.Container.from_numpy(ctx, b, (90, 60), (30, 20))
    print("case: operand cache", flush=True)
    xs = [rng.integers(-8, 9, 260).astype(np.float64) for _ in range(3)]
    cx = [ia.Container.from_numpy(ctx, x, (90,), (30,)) for x in xs]
    for k in (0, 1, 0, 0):
        assert np.array_equal(ia.matmul(ca, cx[k], (100,), (50,)).to_numpy(), a @ xs[k])
    del cx[1]
    x3 = rng.integers(-8, 9, 260).astype(np.float64)
    c3 = ia.Container.from_numpy(ctx, x3, (90,), (30,))
    assert np.array_equal(ia.matmul(ca, c3, (100,), (50,)).to_numpy(), a @ x3)
    assert np.array_equal(ia.matmul(ca, cb, (100, 60), (50, 30)).to_numpy(), a @ b)
    v = cx[0].get_slice((0,), (260,))
    assert np.array_equal(ia.matmul(ca, v, (100,), (50,)).to_numpy(), a @ xs[0])
    assert np.array_equal(ia.matmul(ca, cx[0], (100,), (50,)).to_numpy(), a @ xs[0])
    assert np.array_equal(ia.matmul(ca, c3, (100,), (50,)).to_numpy(), a @ x3)


def test_matmul_deterministic(ctx):
    rng = np.random.default_rng(2)
    a = rng.standard_normal((400, 300))
    b = rng.standard_normal((300, 350))
    ca = ia.Container.from_numpy(ctx, a, (200, 150), (50, 50))
    cb = ia.Container.from_numpy(ctx, b, (150, 175), (50, 25))
    z1 = ia.matmul(ca, cb, (200, 175), (50, 25)).to_numpy()
    z2 = ia.matmul(ca, cb, (200, 175), (50, 25)).to_numpy()
    assert z1.tobytes() == z2.tobytes()
    assert np.allclose(z1, a @ b, rtol=1e-12, atol=1e-12)


def test_batch_ragged_matches_per_product(ctx):
    """iarray_linalg_matmul_batch (one decode launch, one work-queue GEMM,
    one pack launch, replayed from a HIP graph) against numpy on integer
    data (exact), with ragged shapes, ragged edge blocks and all-padding
    blocks; the second call replays the captured graph."""
    rng = np.random.default_rng(55)
    pairs, outs, cs_list, bs_list = [], [], [], []
    for _ in range(12):
        M, K, N = (int(v) for v in rng.integers(64, 700, 3))
        a = rng.integers(-6, 7, (M, K)).astype(np.float32)
        b = rng.integers(-6, 7, (K, N)).astype(np.float32)
        ca = ia.Container.from_numpy(ctx, a, (min(M, 512), min(K, 512)), (min(M, 128), min(K, 128)))
        cb = ia.Container.from_numpy(ctx, b, (min(K, 512), min(N, 512)), (min(K, 128), min(N, 128)))
        pairs.append((ca, cb))
        outs.append(a.astype(np.float64) @ b.astype(np.float64))
        cs_list.append((min(M, 512), min(N, 512)))
        bs_list.append((min(M, 128), min(N, 128)))
    for rep in range(3):
        cs = ia.matmul_batch(pairs, cs_list, bs_list)
        for c, want in zip(cs, outs):
            assert np.array_equal(c.to_numpy(), want.astype(np.float32)), f"replay {rep}"


# ---- zero-block skipping of the opt paths ------------------------------------------------------
def _special_zero_chunk(like: bytes) -> bytes:
    import struct
    h = bytearray(like[:32])
    h[2] = 0x05 | 0x10 | (h[2] & 0xE0)
    h[12:16] = struct.pack("<i", 32)
    h[16:31] = bytes(15)
    h[31] = 1 << 4
    return bytes(h)


def _with_zero_chunks(ctx, x, cs, bs, zero_chunks):
    c = ia.Container.from_numpy(ctx, x, cs, bs)
    ch = list(c.chunks())
    for i in zero_chunks:
        ch[i] = _special_zero_chunk(ch[i])
    out = ia.Container.from_chunks(ctx, x.dtype, x.shape, cs, bs, ch)
    out._cs, out._bs = cs, bs
    return out


SKIP_CASES = [
    # variant, M, K, N, A (chunk, block), B (chunk, block), C (chunk, block)
    ("gemm", 100, 90, 70, ((40, 30), (20, 10)), ((30, 35), (10, 15)), ((40, 35), (20, 15))),
    ("gemm", 64, 128, 96, ((32, 64), (16, 32)), ((64, 48), (32, 16)), ((32, 48), (16, 16))),
    ("a", 70, 45, 50, ((30, 45), (30, 20)), ((45, 20), (20, 20)), ((30, 50), (30, 20))),
    ("b", 70, 45, 50, ((30, 45), (30, 20)), ((45, 20), (20, 20)), ((70, 20), (30, 20))),
]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("case", SKIP_CASES, ids=[f"{c[0]}-{c[1]}x{c[2]}x{c[3]}" for c in SKIP_CASES])
def test_opt_gemm_skips_zero_blocks_on_gpu(ctx, orc, case, dt):
    """iarray_opt_gemm / _a / _b on the GPU skip the same (A block, B block)
    pairs as the reference (src/linalg/iarray_gemm.c:143-165,
    iarray_gemm_a.c / _b.c:100-175): with a special zero chunk and zero
    blocks planted and Inf wherever the other operand faces them, the result
    equals the oracle's bit for bit, NaN and Inf included; the plain matmul
    skips nothing and gets NaN there, as gemm_blosc does."""
    var, M, K, N, (acs, abs_), (bcs, bbs), (ccs, cbs) = case
    print(f"case: opt skip {var} {M}x{K}x{N} {np.dtype(dt).name}", flush=True)
    rng = np.random.default_rng(M + 2 * K + 3 * N)
    a = rng.integers(-8, 9, (M, K)).astype(dt)
    b = rng.integers(-8, 9, (K, N)).astype(dt)
    za, zb = [], []
    if var in ("gemm", "b"):
        a[:acs[0], :acs[1]] = 0              # A chunk (0, 0): a special zero chunk
        za = [0]
    if var in ("gemm", "a"):
        b[:bcs[0], :bcs[1]] = 0              # B chunk (0, 0)
        zb = [0]
    a[acs[0]:acs[0] + abs_[0], :abs_[1]] = 0  # a zero block (streams all empty) of A's second chunk row
    b[:bbs[0], bcs[1]:bcs[1] + bbs[1]] = 0    # and of B's second chunk column
    a[-1, -1] = np.inf
    b[bbs[0]:2 * bbs[0], :] = np.inf          # B rows that meet A's zero blocks
    b[:, -1] = -np.inf
    A = _with_zero_chunks(ctx, a, acs, abs_, za)
    B = _with_zero_chunks(ctx, b, bcs, bbs, zb)
    fn = {"gemm": ia.opt_gemm, "a": ia.opt_gemm_a, "b": ia.opt_gemm_b}[var]
    ofn = {"gemm": orc.opt_gemm, "a": orc.opt_gemm_a, "b": orc.opt_gemm_b}[var]
    with np.errstate(invalid="ignore"):
        got = fn(A, B, ccs, cbs).to_numpy()
        want = ofn(_oracle(orc, A), _oracle(orc, B), ccs, cbs)
        plain_want = orc.gemm_blosc(_oracle(orc, A), _oracle(orc, B), ccs, cbs)
    assert np.isnan(plain_want).sum() > np.isnan(want).sum()   # the planted cases do differ
    assert np.array_equal(got, want, equal_nan=True)
    plain = ia.matmul(A, B, ccs, cbs).to_numpy()
    assert np.array_equal(np.isnan(plain), np.isnan(plain_want))


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_opt_gemv_skips_zero_blocks_on_gpu(ctx, orc, dt):
    """iarray_opt_gemv skips zero chunks and zero blocks of A and of x
    (src/linalg/iarray_gemv.c:125-160): an Inf in x facing only A's zero
    blocks, and an Inf in A facing only x's zero block, never reach y."""
    M, K = 130, 90
    acs, abs_, xcs, xbs = (64, 45), (16, 15), (45,), (15,)
    rng = np.random.default_rng(33)
    a = rng.integers(-8, 9, (M, K)).astype(dt)
    x = rng.integers(-4, 5, K).astype(dt)
    a[:acs[0], :acs[1]] = 0                       # A chunk (0, 0): a special zero chunk
    a[acs[0]:acs[0] + abs_[0], :abs_[1]] = 0      # a zero block of chunk (1, 0)
    x[:abs_[1]] = np.inf                          # meets only those zeros for rows < 80
    x[xcs[0] + xbs[0]:xcs[0] + 2 * xbs[0]] = 0    # a zero block of x's second chunk
    a[70, xcs[0] + xbs[0]] = np.inf               # faces only that zero block (row 70 also meets
    #                                               x's Inf only through A's zero block)
    A = _with_zero_chunks(ctx, a, acs, abs_, [0])
    X = _with_zero_chunks(ctx, x, xcs, xbs, [])
    with np.errstate(invalid="ignore"):
        got = ia.opt_gemv(A, X, (64,), (16,)).to_numpy()
        want = orc.opt_gemv(_oracle(orc, A), _oracle(orc, X), (64,), (16,))
    assert np.isfinite(want[:acs[0] + abs_[0]]).all() and not np.isfinite(want[acs[0] + abs_[0]:]).all()
    assert np.array_equal(got, want, equal_nan=True)


@pytest.mark.parametrize("dt,filt", [(np.float32, ia.COMP_BITSHUFFLE), (np.float64, ia.COMP_SHUFFLE),
                                     (np.float32, 0)])
def test_transposed_view_decodes_in_place(orc, dt, filt):
    """A transposed view's image is decoded with every tile written transposed
    (no separate transpose pass): ragged chunks and blocks, bitshuffled,
    shuffled and unfiltered streams; the image equals x.T and a product with
    it matches the oracle's transposed read."""
    c2 = ia.Context(clevel=5, filters=filt)
    rng = np.random.default_rng(17)
    x = rng.integers(-100, 100, (173, 91)).astype(dt)
    cx = ia.Container.from_numpy(c2, x, (64, 40), (24, 16))
    t = cx.transpose()
    assert np.array_equal(t.to_numpy(), x.T)
    y = rng.integers(-10, 10, (173, 30)).astype(dt)
    cy = ia.Container.from_numpy(c2, y, (50, 30), (25, 10))
    z = ia.matmul(t, cy, (45, 30), (15, 10)).to_numpy()
    assert np.array_equal(z, x.T @ y)
    o = orc.gemm_blosc(orc.Array(cx.chunks(), x.shape, (64, 40), (24, 16), x.itemsize, transposed=True),
                       orc.Array(cy.chunks(), y.shape, (50, 30), (25, 10), y.itemsize), (45, 30), (15, 10))
    assert np.array_equal(z, o)
    # the table is kept for the next call, and follows a rebuilt base
    assert np.array_equal(t.to_numpy(), x.T)
