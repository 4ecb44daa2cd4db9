"""GEMV with the matrix streamed through the device (a matrix larger than the
device budget is never held resident: its chunk rows travel in panels
through two pinned upload stages).  The budget is forced down with
IARRAY_DEVICE_BUDGET so the panels are small.  The streamed product runs the
resident product's blocks, jobs and sums, so y must equal the resident y to
the bit; both are checked against the oracle's gemv_blosc
(src/linalg/matmul/gemv.c:104-190) and opt_gemv (src/linalg/iarray_gemv.c:
249-351) on the same containers.  A frame opened with iarray_container_open
(mapped, chunks paged in on first use) streams the same way."""
import os

import numpy as np
import pytest

import pyiarray as ia

pytestmark = pytest.mark.gpu


def _containers(dtype, filt, seed):
    rng = np.random.default_rng(seed)
    M, K = 2300, 2600
    A = (np.sin(0.003 * np.arange(M))[:, None] + np.cos(0.002 * np.arange(K))[None, :]
         + rng.normal(0, 1e-3, (M, K))).astype(dtype)
    x = rng.uniform(-1, 1, K).astype(dtype)
    ctx = ia.Context(clevel=5, filters=filt)
    tiling = ((500, 600), (100, 150))
    return ctx, A, x, tiling


def _budget_for(c, tiling):
    """A budget that leaves one chunk row of decoded bytes per panel."""
    ncc = -(-c.shape[1] // tiling[0][1])
    row = ncc * tiling[0][0] * tiling[0][1] * c.dtype.itemsize
    return 4 * row


def _with_budget(budget, fn):
    os.environ["IARRAY_DEVICE_BUDGET"] = str(budget)
    try:
        return fn()
    finally:
        del os.environ["IARRAY_DEVICE_BUDGET"]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("filt", [ia.COMP_BITSHUFFLE, ia.COMP_SHUFFLE], ids=["bitshuffle", "shuffle"])
def test_streamed_gemv_equals_resident(orc, dtype, filt):
    print(f"case: streamed gemv {np.dtype(dtype).name} filter {filt}", flush=True)
    ctx, A, x, (cs, bs) = _containers(dtype, filt, 5)
    cx = ia.Container.from_numpy(ctx, x, (cs[1],), (bs[1],))
    resident = ia.Container.from_numpy(ctx, A, cs, bs)
    streamed = ia.Container.from_numpy(ctx, A, cs, bs)
    y0 = ia.matmul(resident, cx, (cs[0],), (bs[0],)).to_numpy()
    assert ia.last_gemv_panels() == 0
    budget = _budget_for(streamed, (cs, bs))
    y1 = _with_budget(budget, lambda: ia.matmul(streamed, cx, (cs[0],), (bs[0],)).to_numpy())
    assert ia.last_gemv_panels() == 5  # one panel per chunk row
    assert y1.tobytes() == y0.tobytes()
    # and again (the upload stages are reused)
    y2 = _with_budget(budget, lambda: ia.matmul(streamed, cx, (cs[0],), (bs[0],)).to_numpy())
    assert y2.tobytes() == y0.tobytes()
    isz = A.itemsize
    want = orc.gemv_blosc(orc.Array(streamed.chunks(), A.shape, cs, bs, isz), orc.Array(cx.chunks(), x.shape,
                          (cs[1],), (bs[1],), isz), (cs[0],), (bs[0],))
    den = np.abs(A.astype(np.float64)) @ np.abs(x.astype(np.float64))
    tol = 1e-5 if dtype == np.float32 else 1e-14
    assert float(np.max(np.abs(y1 - want) / den)) <= tol


def test_streamed_opt_gemv_skips_zero_blocks(orc):
    print("case: streamed opt_gemv with zero blocks", flush=True)
    ctx, A, x, (cs, bs) = _containers(np.float64, ia.COMP_SHUFFLE, 9)
    A[100:300, :] = 0.0   # zero blocks of A
    x[150:450] = 0.0      # zero blocks of x
    cx = ia.Container.from_numpy(ctx, x, (cs[1],), (bs[1],))
    resident = ia.Container.from_numpy(ctx, A, cs, bs)
    streamed = ia.Container.from_numpy(ctx, A, cs, bs)
    y0 = ia.opt_gemv(resident, cx, (cs[0],), (bs[0],)).to_numpy()
    y1 = _with_budget(_budget_for(streamed, (cs, bs)),
                      lambda: ia.opt_gemv(streamed, cx, (cs[0],), (bs[0],)).to_numpy())
    assert ia.last_gemv_panels() == 5
    assert y1.tobytes() == y0.tobytes()
    want = orc.opt_gemv(orc.Array(streamed.chunks(), A.shape, cs, bs, 8),
                        orc.Array(cx.chunks(), x.shape, (cs[1],), (bs[1],), 8), (cs[0],), (bs[0],))
    den = np.abs(A) @ np.abs(x)
    assert float(np.max(np.abs(y1 - want) / np.where(den > 0, den, 1))) <= 1e-14


def test_streamed_gemv_from_opened_frame(tmp_path):
    print("case: streamed gemv of an opened (mapped) frame", flush=True)
    ctx, A, x, (cs, bs) = _containers(np.float32, ia.COMP_BITSHUFFLE, 11)
    cx = ia.Container.from_numpy(ctx, x, (cs[1],), (bs[1],))
    src = ia.Container.from_numpy(ctx, A, cs, bs)
    y0 = ia.matmul(src, cx, (cs[0],), (bs[0],)).to_numpy()
    path = str(tmp_path / "a.iarray")
    src.save(path)
    opened = ia.Container.load(ctx, path, lazy=True)
    y1 = _with_budget(_budget_for(opened, (cs, bs)), lambda: ia.matmul(opened, cx, (cs[0],), (bs[0],)).to_numpy())
    assert ia.last_gemv_panels() == 5
    assert y1.tobytes() == y0.tobytes()
