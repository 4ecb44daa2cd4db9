"""The dense products behind a matmul run either through a hand-written MFMA
kernel or through hipBLASLt, whichever the autotuner timed faster for the
shape class; IAMX_TILE (environment, read on every product: 128, 256 or lib)
pins one.  Every candidate of every form is checked here against the FP64
double-double oracle (oracle/tiling_linalg.c orc_dd_gemm_sample), and the
profile names the path that ran.  Bars (componentwise max |C - C*| / (|A||B|)):
plain float32 1e-5, split-BF16 2e-6, plain float64 and Ozaki-II (14 moduli)
1e-13, Ozaki-I (7 slices) 1e-12.  The Ozaki products are exact integers (the
residues, or each level's digit products) recombined the same way, so the
hand-written kernels and the library give the same bits."""
import os

import numpy as np
import pytest

import pyiarray as ia

pytestmark = pytest.mark.gpu

FORMS = {
    "f32": (np.float32, ia.PLAIN, 0, 1e-5, ["256", "lib"]),
    "split_bf16": (np.float32, ia.SPLIT_BF16, 0, 2e-6, ["256", "128", "lib"]),
    "f64": (np.float64, ia.PLAIN, 0, 1e-13, ["256", "lib"]),
    "ozaki_crt": (np.float64, ia.OZAKI_CRT, 14, 1e-13, ["256", "lib"]),
    "ozaki_int8": (np.float64, ia.OZAKI_INT8, 7, 1e-12, ["256", "128", "lib"]),
}


def _componentwise(orc, a, b, c):
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    M, N = a.shape[0], b.shape[1]
    rows, cols = np.meshgrid(np.arange(M), np.arange(N), indexing="ij")
    ref = orc.dd_gemm_sample(a64, b64, rows.ravel(), cols.ravel()).reshape(M, N)
    den = np.abs(a64) @ np.abs(b64)
    return float(np.max(np.abs(c.astype(np.float64) - ref) / np.where(den > 0, den, 1)))


def _product(form, pick, a, b):
    dt, prec, slices, _, _ = FORMS[form]
    ctx = ia.Context(clevel=3, filters=ia.COMP_SHUFFLE)
    ca = ia.Container.from_numpy(ctx, a, (200, 300), (100, 100))
    cb = ia.Container.from_numpy(ctx, b, (300, 260), (100, 130))
    os.environ["IAMX_TILE"] = pick
    try:
        ia.prof_enable(True)
        ia.prof_reset()
        c = ia.matmul(ca, cb, (200, 260), (100, 130), precision=prec, slices=slices).to_numpy()
        ia.device_sync()
        names = set(ia.prof_entries())
        ia.prof_enable(False)
    finally:
        del os.environ["IAMX_TILE"]
    return c, names


@pytest.mark.parametrize("form", list(FORMS))
def test_every_candidate_matches_the_oracle(orc, form):
    dt, _, _, bar, picks = FORMS[form]
    rng = np.random.default_rng(len(form))
    M, K, N = 600, 900, 520
    if form.startswith("ozaki"):
        a = np.where(rng.random((M, K)) < 0.5, -1.0, 1.0) * 10.0 ** rng.uniform(-3, 3, (M, K))
        b = np.where(rng.random((K, N)) < 0.5, -1.0, 1.0) * 10.0 ** rng.uniform(-3, 3, (K, N))
    else:
        a = rng.standard_normal((M, K)).astype(dt)
        b = rng.standard_normal((K, N)).astype(dt)
    out = {}
    for pick in picks:
        print(f"case: tuner {form} pick {pick}", flush=True)
        c, names = _product(form, pick, a, b)
        lib = any(n.endswith("_hipblaslt") for n in names)
        assert lib == (pick == "lib"), (pick, sorted(names))
        cw = _componentwise(orc, a, b, c)
        print(f"{form} {pick}: componentwise {cw:.3e}", flush=True)
        assert cw <= bar
        out[pick] = c
    if form.startswith("ozaki"):
        for pick in picks:
            assert out[pick].tobytes() == out["256"].tobytes(), pick


def test_library_product_repeats_its_bits():
    """A product class keeps its pick: the same call twice gives the same
    bits, whichever path the autotuner chose."""
    rng = np.random.default_rng(5)
    a = rng.standard_normal((1200, 1000))
    b = rng.standard_normal((1000, 1100))
    ctx = ia.Context(clevel=3, filters=ia.COMP_SHUFFLE)
    ca = ia.Container.from_numpy(ctx, a, (400, 500), (100, 100))
    cb = ia.Container.from_numpy(ctx, b, (500, 550), (100, 110))
    z = [ia.matmul(ca, cb, (400, 550), (100, 110)).to_numpy() for _ in range(3)]
    assert z[1].tobytes() == z[0].tobytes() and z[2].tobytes() == z[0].tobytes()
    assert np.allclose(z[0], a @ b, rtol=1e-11, atol=1e-11)
