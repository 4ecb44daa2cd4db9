"""Pins of the oracle's linalg restatements, on the CPU.

The oracle (oracle/tiling_linalg.c) restates
  gemm_blosc        src/linalg/matmul/gemm.c:130-249
  gemv_blosc        src/linalg/matmul/gemv.c:104-190
  iarray_opt_gemm   src/linalg/iarray_gemm.c:91-312
  iarray_opt_gemv   src/linalg/iarray_gemv.c:88-246
and a double-double dot product (orc_dd_gemm_sample) that is the FP64
reference of every mixed-precision check.  Each is checked here against what
the mathematics fixes, chosen so that a plausible slip fails:
  * integer-valued operands make every summation order exact, so the blocked
    products must equal numpy's a @ b bit for bit, on ragged tilings (edge
    chunks and blocks, all-padding blocks), on transposed operands, and with
    zero chunks (special-value chunks) and zero blocks (all streams empty)
    planted; a dropped block, a wrong block origin or a transposed operand
    changes the result;
  * the opt paths skip zero chunks and zero blocks (src/linalg/iarray_gemm.c:
    17-77, 143-165; src/linalg/iarray_gemv.c:125-176): an Inf planted in the
    other operand where it only meets those zeros must not reach C (0 * Inf
    would be NaN);
  * dd_gemm_sample against exact rational dot products (fractions.Fraction)
    rounded once.
"""
import struct
from fractions import Fraction

import numpy as np
import pytest

import pyiarray as ia


def _ctx():
    return ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)


def _special_zero_chunk(like: bytes) -> bytes:
    """A 32-byte special chunk of zeros with the header fields of `like`."""
    h = bytearray(like[:32])
    h[2] = 0x05 | 0x10 | (h[2] & 0xE0)
    h[12:16] = struct.pack("<i", 32)
    h[16:31] = bytes(15)
    h[31] = 1 << 4
    return bytes(h)


def _arr(orc, x, cs, bs, zero_chunks=(), transposed=False):
    c = ia.Container.from_numpy(_ctx(), x, cs, bs)
    ch = list(c.chunks())
    for i in zero_chunks:
        ch[i] = _special_zero_chunk(ch[i])
    return orc.Array(ch, x.shape, cs, bs, x.itemsize, transposed=transposed)


def _ints(rng, shape, dt=np.float64, lo=-8, hi=9):
    return rng.integers(lo, hi, shape).astype(dt)


GEMM_CASES = [
    # M, K, N, A (chunk, block), B (chunk, block), C (chunk, block)
    (100, 170, 21, ((27, 30), (12, 7)), ((20, 11), (10, 5)), ((40, 3), (20, 3))),
    (300, 670, 210, ((28, 300), (11, 31)), ((20, 45), (7, 12)), ((43, 11), (13, 5))),
    (130, 97, 75, ((64, 40), (16, 16)), ((40, 32), (16, 16)), ((64, 32), (32, 16))),
]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("case", GEMM_CASES, ids=[f"{c[0]}x{c[1]}x{c[2]}" for c in GEMM_CASES])
def test_gemm_blosc_exact_on_integers(orc, case, dt):
    M, K, N, (acs, abs_), (bcs, bbs), (ccs, cbs) = case
    rng = np.random.default_rng(M * 7 + K)
    a, b = _ints(rng, (M, K), dt), _ints(rng, (K, N), dt)
    got = orc.gemm_blosc(_arr(orc, a, acs, abs_), _arr(orc, b, bcs, bbs), ccs, cbs)
    assert got.dtype == dt
    assert np.array_equal(got, a @ b)


def test_gemm_blosc_transposed_operands(orc):
    """Both operands as transposed views of stored (K, M) and (N, K) arrays
    (src/linalg/transpose.c:18-45); the stored arrays are not square, so an
    ignored transpose flag cannot give the same shape, let alone values."""
    rng = np.random.default_rng(11)
    at, bt = _ints(rng, (170, 100)), _ints(rng, (21, 170))
    A = _arr(orc, at, (20, 30), (10, 12), transposed=True)
    B = _arr(orc, bt, (11, 20), (5, 10), transposed=True)
    got = orc.gemm_blosc(A, B, (40, 3), (20, 3))
    assert np.array_equal(got, at.T @ bt.T)
    # square case: the transposed and plain products differ
    sq = _ints(rng, (96, 96))
    S, St = _arr(orc, sq, (32, 48), (16, 16)), _arr(orc, sq, (32, 48), (16, 16), transposed=True)
    r = orc.gemm_blosc(St, S, (32, 32), (16, 16))
    assert np.array_equal(r, sq.T @ sq) and not np.array_equal(r, sq @ sq)


def test_gemm_blosc_zero_chunks_and_blocks(orc):
    rng = np.random.default_rng(5)
    a, b = _ints(rng, (120, 90)), _ints(rng, (90, 80))
    a[40:80, 0:45] = 0          # one whole chunk of A
    a[0:20, 45:60] = 0          # one block of A
    A = _arr(orc, a, (40, 45), (20, 15), zero_chunks=[2])
    B = _arr(orc, b, (45, 40), (15, 20))
    assert np.array_equal(orc.gemm_blosc(A, B, (40, 40), (20, 20)), a @ b)


GEMV_CASES = [
    # M, K, A (chunk, block), x (chunk, block), y (chunk, block)
    (650, 900, ((200, 140), (20, 14)), ((421,), (55,)), ((155,), (47,))),
    (333, 257, ((100, 64), (25, 16)), ((64,), (16,)), ((100,), (25,))),
]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("case", GEMV_CASES, ids=[f"{c[0]}x{c[1]}" for c in GEMV_CASES])
def test_gemv_blosc_exact_on_integers(orc, case, dt):
    M, K, (acs, abs_), (xcs, xbs), (ycs, ybs) = case
    rng = np.random.default_rng(M + K)
    a, x = _ints(rng, (M, K), dt), _ints(rng, (K,), dt)
    got = orc.gemv_blosc(_arr(orc, a, acs, abs_), _arr(orc, x, xcs, xbs), ycs, ybs)
    assert np.array_equal(got, a @ x)
    # a transposed A (stored (K, M)) reads the other way
    at = _ints(rng, (K, M), dt)
    got_t = orc.gemv_blosc(_arr(orc, at, (acs[1], acs[0]), (abs_[1], abs_[0]), transposed=True),
                           _arr(orc, x, xcs, xbs), ycs, ybs)
    assert np.array_equal(got_t, at.T @ x)


OPT_GEMM_CASES = [
    # M, K, N, A (chunk, block), B (chunk, block), C (chunk, block): the partition
    # rules of src/linalg/iarray_gemm.c:315-364 (A chunk rows = C chunk rows,
    # A chunk cols = B chunk rows, B chunk cols = C chunk cols, blocks alike)
    (100, 90, 70, ((40, 30), (20, 10)), ((30, 35), (10, 15)), ((40, 35), (20, 15))),
    (64, 128, 96, ((32, 64), (16, 32)), ((64, 48), (32, 16)), ((32, 48), (16, 16))),
]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("case", OPT_GEMM_CASES, ids=[f"{c[0]}x{c[1]}x{c[2]}" for c in OPT_GEMM_CASES])
def test_opt_gemm_exact_and_skips_zeros(orc, case, dt):
    M, K, N, (acs, abs_), (bcs, bbs), (ccs, cbs) = case
    rng = np.random.default_rng(M + 2 * K + 3 * N)
    a, b = _ints(rng, (M, K), dt), _ints(rng, (K, N), dt)
    got = orc.opt_gemm(_arr(orc, a, acs, abs_), _arr(orc, b, bcs, bbs), ccs, cbs)
    assert np.array_equal(got, a @ b)
    # A's first chunk a special zero chunk and one block of its second chunk
    # zero; B holds Inf exactly where only those zeros multiply it
    a[:acs[0], :acs[1]] = 0
    a[:abs_[0], acs[1]:acs[1] + abs_[1]] = 0
    b_inf = b.copy()
    b_inf[:acs[1], :] = np.inf                                   # meets A chunk (0, 0) only for rows < acs[0]
    b_inf[acs[1]:acs[1] + abs_[1], :] = np.inf                   # meets the zero block only for rows < abs_[0]
    exp = a @ b
    A = _arr(orc, a, acs, abs_, zero_chunks=[0])
    got = orc.opt_gemm(A, _arr(orc, b, bcs, bbs), ccs, cbs)
    assert np.array_equal(got, exp)
    got_inf = orc.opt_gemm(A, _arr(orc, b_inf, bcs, bbs), ccs, cbs)
    top = got_inf[:abs_[0]]
    # the first block row never sees the Inf rows of B: K range [0, acs[1])
    # is a zero chunk and [acs[1], acs[1] + abs_[1]) a zero block of A
    k0 = acs[1] + abs_[1]
    want_top = a[:abs_[0], k0:] @ b[k0:, :]
    assert np.array_equal(top, want_top)
    # rows of later chunks do meet the Inf (nonzero entries there)
    assert not np.all(np.isfinite(got_inf[acs[0]:]))


OPT_GEMV_CASES = [
    # M, K, A (chunk, block), x (chunk, block), y (chunk, block): the rules of
    # src/linalg/iarray_gemv.c:249-296
    (600, 400, ((200, 100), (20, 20)), ((100,), (20,)), ((200,), (20,))),
    (130, 90, ((64, 45), (16, 15)), ((45,), (15,)), ((64,), (16,))),
]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("case", OPT_GEMV_CASES, ids=[f"{c[0]}x{c[1]}" for c in OPT_GEMV_CASES])
def test_opt_gemv_exact_and_skips_zeros(orc, case, dt):
    M, K, (acs, abs_), (xcs, xbs), (ycs, ybs) = case
    rng = np.random.default_rng(M * 5 + K)
    a, x = _ints(rng, (M, K), dt), _ints(rng, (K,), dt)
    got = orc.opt_gemv(_arr(orc, a, acs, abs_), _arr(orc, x, xcs, xbs), ycs, ybs)
    assert np.array_equal(got, a @ x)
    a[:acs[0], :acs[1]] = 0                                     # chunk (0, 0)
    a[acs[0]:acs[0] + abs_[0], :abs_[1]] = 0                    # first block of chunk (1, 0)
    x_inf = x.copy()
    x_inf[:abs_[1]] = np.inf
    A = _arr(orc, a, acs, abs_, zero_chunks=[0])
    got = orc.opt_gemv(A, _arr(orc, x_inf, xcs, xbs), ycs, ybs)
    # rows [0, acs[0] + abs_[0]) meet the Inf only through zeros
    lim = acs[0] + abs_[0]
    want = a[:lim, abs_[1]:] @ x[abs_[1]:]
    assert np.array_equal(got[:lim], want)
    assert not np.all(np.isfinite(got[lim:]))


def _exact_dot(a_row, b_col):
    s = Fraction(0)
    for u, v in zip(a_row, b_col):
        s += Fraction(float(u)) * Fraction(float(v))
    return s


def test_dd_gemm_sample_matches_exact_rationals(orc):
    """Wide-range data (sign * 10^U[-3,3]) and cancelling sums: the
    double-double dot product rounded once is within half an ulp (plus a
    2^-100 relative margin for the double-double tail) of the exact rational
    sum; a plain float64 dot product is not, on the same entries."""
    rng = np.random.default_rng(21)
    M, K, N = 12, 64, 10
    a = np.where(rng.random((M, K)) < 0.5, -1.0, 1.0) * 10.0 ** rng.uniform(-3, 3, (M, K))
    b = np.where(rng.random((K, N)) < 0.5, -1.0, 1.0) * 10.0 ** rng.uniform(-3, 3, (K, N))
    # make some entries cancel: b's last column is minus a permutation that
    # pairs large terms
    b[:, -1] = -b[:, 0]
    b[0, -1] += 1e-9
    rows, cols = np.meshgrid(np.arange(M), np.arange(N), indexing="ij")
    got = orc.dd_gemm_sample(a, b, rows.ravel(), cols.ravel())
    worse = 0
    for idx, (i, j) in enumerate(zip(rows.ravel(), cols.ravel())):
        ex = _exact_dot(a[i], b[:, j])
        g = Fraction(float(got[idx]))
        ulp = Fraction(np.spacing(abs(float(ex)))) if ex != 0 else Fraction(0)
        assert abs(g - ex) <= ulp / 2 + abs(ex) * Fraction(1, 2 ** 100), (i, j, float(ex), got[idx])
        naive = Fraction(float(np.dot(a[i], b[:, j])))
        if abs(naive - ex) > ulp / 2:
            worse += 1
    assert worse > 0, "the check must be able to tell a plain float64 dot product from the exact sum"
