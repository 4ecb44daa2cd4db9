"""GPU accuracy of the mixed-precision matmul forms against an FP64 oracle.

The oracle is a double-double (compensated) dot product per output element
(oracle/tiling_linalg.c orc_dd_gemm_sample), i.e. the exact product rounded
once.  Metrics per result C against C*:
  componentwise  max_ij |C - C*|_ij / (|A||B|)_ij
  normwise       ||C - C*||_F / ||C*||_F
Bars (written here):
  split-bf16 (float32): componentwise <= 6e-5 (3 bf16 products drop lo*lo ~
    2^-18 and accumulate in FP32 over K), and never worse than 4x the plain
    FP32 GPU result on the same inputs plus that bound;
  ozaki-int8 (float64, values +-10^U[-3,3]): every entry within the a-priori
    bound of the slicing -- residual |a - a~| <= 2^(e_i-7S-1) per operand
    (balanced digits, e_i = ilogb(max_k|A_ik|)+2), dropped levels s+t > S+1
    at most 1.01 (S-1) 2^(e_i+f_j-7S-2) per product, plus 2S ulp of FP64
    recombination -- and componentwise <= 1e-12 (S=7), <= 1e-14 (S=8).
"""
import numpy as np
import pytest

import pyiarray as ia

pytestmark = pytest.mark.gpu


def _metrics(orc, a, b, c):
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    M, N = a.shape[0], b.shape[1]
    rows, cols = np.meshgrid(np.arange(M), np.arange(N), indexing="ij")
    ref = orc.dd_gemm_sample(a64, b64, rows.ravel(), cols.ravel()).reshape(M, N)
    den = np.abs(a64) @ np.abs(b64)
    err = np.abs(c.astype(np.float64) - ref)
    comp = float(np.max(err / np.where(den > 0, den, 1)))
    norm = float(np.linalg.norm(err) / np.linalg.norm(ref))
    return comp, norm


@pytest.mark.parametrize("shape", [(256, 384, 320), (300, 700, 129), (128, 1024, 128)])
def test_split_bf16_accuracy(orc, shape):
    M, K, N = shape
    print(f"case: split-bf16 {shape}", flush=True)
    ctx = ia.Context(clevel=3, filters=ia.COMP_SHUFFLE)
    rng = np.random.default_rng(M + K + N)
    a = rng.standard_normal((M, K)).astype(np.float32)
    b = rng.standard_normal((K, N)).astype(np.float32)
    ca = ia.Container.from_numpy(ctx, a, (min(M, 128), min(K, 256)), (64, 64))
    cb = ia.Container.from_numpy(ctx, b, (min(K, 256), min(N, 128)), (64, 64))
    cs, bs = (min(M, 128), min(N, 128)), (64, 64)
    mixed = ia.matmul(ca, cb, cs, bs, precision=ia.SPLIT_BF16).to_numpy()
    plain = ia.matmul(ca, cb, cs, bs).to_numpy()
    cm, nm = _metrics(orc, a, b, mixed)
    cp, npl = _metrics(orc, a, b, plain)
    print(f"split-bf16 {shape}: componentwise {cm:.3e} normwise {nm:.3e} | plain fp32 {cp:.3e} {npl:.3e}")
    assert cm <= 6e-5
    assert cm <= 4 * cp + 6e-5


def _wide(rng, shape):
    return np.where(rng.random(shape) < 0.5, -1.0, 1.0) * 10.0 ** rng.uniform(-3, 3, shape)


def _ozaki_bound(a, b, S):
    ea = np.frexp(np.abs(a).max(axis=1))[1] + 1  # ilogb(max)+2
    fb = np.frexp(np.abs(b).max(axis=0))[1] + 1
    K = a.shape[1]
    ra = np.ldexp(1.0, ea - 7 * S - 1)[:, None]
    rb = np.ldexp(1.0, fb - 7 * S - 1)[None, :]
    trunc = ra * np.abs(b).sum(axis=0)[None, :] + (np.abs(a).sum(axis=1)[:, None] + K * ra) * rb
    dropped = 1.01 * (S - 1) * K * np.ldexp(1.0, (ea[:, None] + fb[None, :]) - 7 * S - 2)
    return trunc + dropped + 2 * S * 2.0 ** -52 * (np.abs(a) @ np.abs(b))


@pytest.mark.parametrize("slices,bar", [(7, 1e-12), (8, 1e-14)])
@pytest.mark.parametrize("shape", [(256, 500, 192), (129, 260, 300)])
def test_ozaki_int8_accuracy(orc, shape, slices, bar):
    M, K, N = shape
    print(f"case: ozaki {shape} S={slices}", flush=True)
    ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
    rng = np.random.default_rng(M * 3 + K)
    a = _wide(rng, (M, K))
    b = _wide(rng, (K, N))
    ca = ia.Container.from_numpy(ctx, a, (min(M, 128), min(K, 250)), (64, 50))
    cb = ia.Container.from_numpy(ctx, b, (min(K, 250), min(N, 128)), (50, 64))
    c = ia.matmul(ca, cb, (min(M, 128), min(N, 128)), (64, 64), precision=ia.OZAKI_INT8, slices=slices).to_numpy()
    cm, nm = _metrics(orc, a, b, c)
    print(f"ozaki S={slices} {shape}: componentwise {cm:.3e} normwise {nm:.3e}")
    rows, cols = np.meshgrid(np.arange(M), np.arange(N), indexing="ij")
    ref = orc.dd_gemm_sample(a, b, rows.ravel(), cols.ravel()).reshape(M, N)
    assert np.all(np.abs(c - ref) <= _ozaki_bound(a, b, slices))
    assert cm <= bar


def test_ozaki_auto_slices_reach_target(orc):
    """slices=0 picks the smallest S whose sampled componentwise error is at
    most 1e-13; the full product must then meet the bar on every entry."""
    M, K, N = 300, 700, 260
    ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
    rng = np.random.default_rng(77)
    a = _wide(rng, (M, K))
    b = _wide(rng, (K, N))
    ca = ia.Container.from_numpy(ctx, a, (150, 350), (50, 70))
    cb = ia.Container.from_numpy(ctx, b, (350, 130), (70, 65))
    c = ia.matmul(ca, cb, (150, 130), (50, 65), precision=ia.OZAKI_INT8, slices=0).to_numpy()
    S, err = ia.last_ozaki_slices()
    cm, nm = _metrics(orc, a, b, c)
    print(f"auto slices: S={S} sampled {err:.3e} full componentwise {cm:.3e}")
    assert 4 <= S <= 8
    assert 0 <= err <= 1e-13 or S == 8
    assert cm <= 1e-13
    rows, cols = np.meshgrid(np.arange(M), np.arange(N), indexing="ij")
    ref = orc.dd_gemm_sample(a, b, rows.ravel(), cols.ravel()).reshape(M, N)
    assert np.all(np.abs(c - ref) <= _ozaki_bound(a, b, S))
    # the pick is cached for the pair: the second call reuses it, no sample error
    c2 = ia.matmul(ca, cb, (150, 130), (50, 65), precision=ia.OZAKI_INT8, slices=0).to_numpy()
    assert ia.last_ozaki_slices() == (S, -1.0)
    assert c2.tobytes() == c.tobytes()


MODS = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197]


def _crt_bits(n, K):
    """p, q of the CRT form: the largest p + q with K 2^(p+q) < M/2, split
    evenly, each at most 52."""
    import math
    log2m = sum(math.log2(m) for m in MODS[:n])
    budget = math.floor(log2m - 1 - math.log2(K) - 1e-9)
    p = min(budget // 2, 52)
    return p, min(budget - p, 52)


def _crt_bound(a, b, n):
    """A-priori bound of the CRT form: A'_ik = rint(A_ik 2^(p+1-e_i)) is off
    by at most 2^(e_i-p-2) in A's units (B alike with q), the modular
    products and the recombination are exact, and the final FP64 conversion
    and unscaling round at most 8 times."""
    K = a.shape[1]
    p, q = _crt_bits(n, K)
    ea = np.frexp(np.abs(a).max(axis=1))[1] + 1
    fb = np.frexp(np.abs(b).max(axis=0))[1] + 1
    ra = np.ldexp(1.0, ea - p - 2)[:, None]
    rb = np.ldexp(1.0, fb - q - 2)[None, :]
    trunc = ra * np.abs(b).sum(axis=0)[None, :] + (np.abs(a).sum(axis=1)[:, None] + K * ra) * rb
    return trunc + 8 * 2.0 ** -53 * (np.abs(a) @ np.abs(b))


# bars: the rounding of the scaled integers is 2^-p of a row's largest entry;
# on +-10^U[-3,3] data the largest entry is ~14x the mean magnitude, so the
# componentwise error is about 28 * 2^-min(p, q) (n = 12: p = 42 at K = 500;
# n = 14: p >= 49 up to K = 2100)
@pytest.mark.parametrize("moduli,bar", [(12, 1e-11), (14, 1e-13)])
@pytest.mark.parametrize("shape", [(256, 500, 192), (129, 260, 300), (300, 2100, 260)])
def test_ozaki_crt_accuracy(orc, shape, moduli, bar):
    M, K, N = shape
    print(f"case: ozaki-crt {shape} n={moduli}", flush=True)
    ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
    rng = np.random.default_rng(M * 5 + K)
    a = _wide(rng, (M, K))
    b = _wide(rng, (K, N))
    ca = ia.Container.from_numpy(ctx, a, (min(M, 128), min(K, 250)), (64, 50))
    cb = ia.Container.from_numpy(ctx, b, (min(K, 250), min(N, 128)), (50, 64))
    c = ia.matmul(ca, cb, (min(M, 128), min(N, 128)), (64, 64), precision=ia.OZAKI_CRT, slices=moduli).to_numpy()
    cm, nm = _metrics(orc, a, b, c)
    print(f"ozaki-crt n={moduli} {shape}: componentwise {cm:.3e} normwise {nm:.3e}")
    rows, cols = np.meshgrid(np.arange(M), np.arange(N), indexing="ij")
    ref = orc.dd_gemm_sample(a, b, rows.ravel(), cols.ravel()).reshape(M, N)
    assert np.all(np.abs(c - ref) <= _crt_bound(a, b, moduli))
    assert cm <= bar
    assert ia.last_ozaki_slices()[0] == moduli


def test_ozaki_crt_auto_moduli_reach_target(orc):
    M, K, N = 300, 700, 260
    ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
    rng = np.random.default_rng(78)
    a = _wide(rng, (M, K))
    b = _wide(rng, (K, N))
    ca = ia.Container.from_numpy(ctx, a, (150, 350), (50, 70))
    cb = ia.Container.from_numpy(ctx, b, (350, 130), (70, 65))
    c = ia.matmul(ca, cb, (150, 130), (50, 65), precision=ia.OZAKI_CRT, slices=0).to_numpy()
    n, err = ia.last_ozaki_slices()
    cm, nm = _metrics(orc, a, b, c)
    print(f"auto moduli: n={n} sampled {err:.3e} full componentwise {cm:.3e}")
    assert 8 <= n <= 15
    assert 0 <= err <= 1e-13 or n == 15
    assert cm <= 1e-13
    # the Ozaki-slice pick for the same pair is kept apart from this one
    ia.matmul(ca, cb, (150, 130), (50, 65), precision=ia.OZAKI_INT8, slices=0)
    S, serr = ia.last_ozaki_slices()
    assert 4 <= S <= 8 and serr >= 0
    c2 = ia.matmul(ca, cb, (150, 130), (50, 65), precision=ia.OZAKI_CRT, slices=0).to_numpy()
    assert ia.last_ozaki_slices() == (n, -1.0)
    assert c2.tobytes() == c.tobytes()


def test_ozaki_crt_exact_on_integers_and_signs(orc):
    """Integer operands are scaled by powers of two only, so every modular
    product and the recombination are exact: the result equals a @ b, negative
    entries (representatives above M/2) included."""
    ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
    rng = np.random.default_rng(10)
    a = rng.integers(-1000, 1001, (200, 300)).astype(np.float64)
    b = rng.integers(-1000, 1001, (300, 150)).astype(np.float64)
    a[5] = 0.0
    ca = ia.Container.from_numpy(ctx, a, (100, 100), (50, 50))
    cb = ia.Container.from_numpy(ctx, b, (100, 75), (50, 25))
    c = ia.matmul(ca, cb, (100, 75), (50, 25), precision=ia.OZAKI_CRT, slices=10).to_numpy()
    want = a @ b
    assert (want < 0).any() and (want > 0).any()
    assert np.array_equal(c, want)


def test_ozaki_exact_on_small_integers(orc):
    """Integers below 2^7 are one slice: every form must be exact."""
    ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
    rng = np.random.default_rng(9)
    a = rng.integers(-100, 101, (200, 300)).astype(np.float64)
    b = rng.integers(-100, 101, (300, 150)).astype(np.float64)
    ca = ia.Container.from_numpy(ctx, a, (100, 100), (50, 50))
    cb = ia.Container.from_numpy(ctx, b, (100, 75), (50, 25))
    c = ia.matmul(ca, cb, (100, 75), (50, 25), precision=ia.OZAKI_INT8, slices=3).to_numpy()
    assert np.array_equal(c, a @ b)


def test_mixed_dtype_errors():
    ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
    a = ia.Container.arange(ctx, np.float64, (64, 64), (32, 32), (16, 16))
    f = ia.Container.arange(ctx, np.float32, (64, 64), (32, 32), (16, 16))
    with pytest.raises(ia.IArrayError) as e:
        ia.matmul(a, a, (32, 32), (16, 16), precision=ia.SPLIT_BF16)
    assert e.value.rc == ia.ERR_INVALID_DTYPE
    with pytest.raises(ia.IArrayError) as e:
        ia.matmul(f, f, (32, 32), (16, 16), precision=ia.OZAKI_INT8)
    assert e.value.rc == ia.ERR_INVALID_DTYPE
