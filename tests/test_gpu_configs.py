"""GPU parity at the five benchmark workloads' full sizes.

The operands are exactly the ones bench.py measures (bench.make_workload:
same generators, seeds, tilings and codecs).  Oracles:
  * orc.gemm_blosc / orc.gemv_blosc: the reference's CPU algorithm
    (src/linalg/matmul/gemm.c:130-249, gemv.c:104-190) restated on the same
    compressed containers;
  * orc.dd_gemm_sample: the exact dot product rounded once (double-double),
    on seeded samples of entries.
Metrics: componentwise |C - C*| / (|A||B|) (the meaningful bar for random data,
whose entries can cancel to ~0), normwise ||C - C*|| / ||C*|| and elementwise
|C - C*| / |C*|, the last two printed.  Bars, written here:
  2000^2 float64 GEMM ........ componentwise <= 1e-14 against gemm_blosc (every
                               entry) and against the double-double sample
                               (the reference's float64 tolerance,
                               tests/test_linalg_gemm.c:128-147);
  20000^2 float32 GEMV ....... componentwise <= 1e-5 against gemv_blosc (every
                               row) and <= 1e-6 against the double-double sample;
  8192^3 float32 GEMM ........ plain and split-BF16: componentwise <= 1e-6
                               against the double-double sample (4096 entries);
  6000^3 float64 GEMM ........ plain and both Ozaki forms (count autotuned):
                               componentwise <= 1e-13 (4096 entries);
  512 ragged float32 GEMMs ... componentwise <= 1e-6 on 32 sampled entries of
                               every product.
"""
import os
import sys

import numpy as np
import pytest

import pyiarray as ia

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

pytestmark = pytest.mark.gpu


def _componentwise(got, want, a, b):
    den = np.abs(a.astype(np.float64)) @ np.abs(b.astype(np.float64))
    err = np.abs(got.astype(np.float64) - want.astype(np.float64))
    return float(np.max(err / np.where(den > 0, den, 1.0)))


def _oracle_array(orc, c, tiling):
    return orc.Array(c.chunks(), c.shape, tiling[0], tiling[1], c.dtype.itemsize)


def test_gemm2k_f64_full_product(orc):
    print("case: gemm 2000^3 f64, chunks 500^2 blocks 125^2, lz4 c5 shuffle", flush=True)
    orc.set_blas(1)
    w = bench.make_workload(ia, "gemm2k_f64")
    c = ia.matmul(w["a"], w["b"], w["c_chunks"], w["c_blocks"])
    got = c.to_numpy()
    want = orc.gemm_blosc(_oracle_array(orc, w["a"], w["a_tiling"]), _oracle_array(orc, w["b"], w["b_tiling"]),
                          w["c_chunks"], w["c_blocks"], nthreads=16)
    cw = _componentwise(got, want, w["a_dense"], w["b_dense"])
    e = bench.error_metrics(got, w["a_dense"], w["b_dense"])
    print(f"gemm2k f64: componentwise vs oracle {cw:.3e}; vs double-double {e}", flush=True)
    assert cw <= 1e-14
    assert e["componentwise_max"] <= 1e-14
    # the C chunks the GPU wrote decode (oracle decoder) to the same bits
    oc = orc.Array(c.chunks(), c.shape, w["c_chunks"], w["c_blocks"], 8)
    assert oc.to_dense().tobytes() == got.tobytes()


def test_gemv20k_f32_full_product(orc):
    print("case: gemv 20000^2 f32, chunks 2000^2 blocks 250^2, lz4 c5 bitshuffle", flush=True)
    orc.set_blas(1)
    orc.set_lz4(1)
    w = bench.make_workload(ia, "gemv20k")
    y = ia.matmul(w["a"], w["b"], w["c_chunks"], w["c_blocks"]).to_numpy()
    want = orc.gemv_blosc(_oracle_array(orc, w["a"], w["a_tiling"]), _oracle_array(orc, w["b"], w["b_tiling"]),
                          w["c_chunks"], w["c_blocks"], nthreads=16)
    A, x = w["a_dense"], w["b_dense"]
    den = np.abs(A) @ np.abs(x)  # float32 is enough for a denominator
    cw = float(np.max(np.abs(y.astype(np.float64) - want) / den))
    e = bench.error_metrics(y, A, x)
    print(f"gemv20k f32: componentwise vs oracle {cw:.3e}; vs double-double {e}", flush=True)
    assert cw <= 1e-5
    assert e["componentwise_max"] <= 1e-6
    # the same call again gives the same bits
    y2 = ia.matmul(w["a"], w["b"], w["c_chunks"], w["c_blocks"]).to_numpy()
    assert y2.tobytes() == y.tobytes()


@pytest.mark.parametrize("precision", ["plain", "split_bf16"])
def test_gemm8k_f32(orc, precision):
    print(f"case: gemm 8192^3 f32 {precision}", flush=True)
    w = _cached("gemm8k_f32")
    pc = ia.PLAIN if precision == "plain" else ia.SPLIT_BF16
    c = ia.matmul(w["a"], w["b"], w["c_chunks"], w["c_blocks"], precision=pc).to_numpy()
    e = bench.error_metrics(c, w["a_dense"], w["b_dense"])
    print(f"gemm8k f32 {precision}: {e}", flush=True)
    assert e["componentwise_max"] <= 1e-6


@pytest.mark.parametrize("precision", ["plain", "ozaki_int8", "ozaki_crt"])
def test_gemm6k_f64(orc, precision):
    print(f"case: gemm 6000^3 f64 wide-range {precision}", flush=True)
    w = _cached("gemm6k_f64")
    pc = bench.PRECISIONS[precision]
    c = ia.matmul(w["a"], w["b"], w["c_chunks"], w["c_blocks"], precision=pc, slices=0).to_numpy()
    e = bench.error_metrics(c, w["a_dense"], w["b_dense"])
    extra = ""
    if precision != "plain":
        extra = f" slices {ia.last_ozaki_slices()}"
    print(f"gemm6k f64 {precision}: {e}{extra}", flush=True)
    assert e["componentwise_max"] <= 1e-13


def test_batch512_every_product(orc):
    print("case: batch of 512 ragged f32 products", flush=True)
    w = bench.make_workload(ia, "batch512", keep_dense=True)
    outs = ia.matmul_batch(w["pairs"], w["c_chunks_list"], w["c_blocks_list"])
    rng = np.random.default_rng(3)
    worst = 0.0
    for i, (c, (a, b)) in enumerate(zip(outs, w["dense"])):
        got = c.to_numpy()
        assert got.shape == (a.shape[0], b.shape[1])
        rows = rng.integers(0, a.shape[0], 32)
        cols = rng.integers(0, b.shape[1], 32)
        a64, b64 = a.astype(np.float64), b.astype(np.float64)
        ref = orc.dd_gemm_sample(a64, b64, rows, cols)
        den = np.einsum("ij,ji->i", np.abs(a64[rows]), np.abs(b64[:, cols]))
        err = float(np.max(np.abs(got[rows, cols] - ref) / den))
        worst = max(worst, err)
        assert err <= 1e-6, f"product {i} {a.shape}x{b.shape}: componentwise {err:.3e}"
    print(f"batch512: worst sampled componentwise {worst:.3e}", flush=True)


_CACHE = {}


def _cached(name):
    """One workload kept for the parametrised cases of a test (the operands
    take seconds to generate and encode)."""
    if name not in _CACHE:
        _CACHE.clear()
        _CACHE[name] = bench.make_workload(ia, name)
    return _CACHE[name]
