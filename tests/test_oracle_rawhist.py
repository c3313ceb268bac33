"""Pins of the oracle's rawhist/nprof (noise profiling: per-brightness
moments of the raw greens and the noise-model fit) against what the
statistics fix, and the product's host-side fit against the oracle's.

  * collect reads the greens only: with the greens of a block set to one
    value, red and blue are free and every block lands in the bin
    int(bins * log2(v + 1)) with its exact moments (Bayer (1,0),(0,1);
    X-Trans corners and centre of each 3x3 block);
  * the dispatch's 8x8 rounding: blocks past the node extent are collected
    while their first site lies inside the mosaic; a non-mosaic input adds
    every invocation into bin 0;
  * the statistic of a bin is the within-block spread of the greens: for
    Poisson-Gaussian noise var = a*x + b it is var/2 for the two Bayer greens
    and 4/5 var for the five X-Trans greens, so the fitted line recovers
    (b/2, a/2) and (4b/5, 4a/5);
  * the fit falls back to the reference's placeholder (100, 1) without data.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_py as op
from vkdt_amd import PKG_DIR, synth

BAYER = synth.BAYER_FILTERS


def norm(v, black, white):
    """collect.comp:28-35 in float32: UNORM read, levels over 65536"""
    v = np.float32(np.float32(v) / np.float32(65535.0))
    b, w = np.float32(black) / np.float32(65536.0), np.float32(white) / np.float32(65536.0)
    return np.clip((v - b) / (w - b), 0, 1).astype(np.float32)


@pytest.mark.parametrize("kind", ["bayer", "xtrans"])
def test_greens_only_and_exact_moments(kind):
    W, H = (64, 48) if kind == "bayer" else (66, 48)
    rng = np.random.default_rng(1)
    cm = synth.channel_map(W, H, kind)
    g = 9000
    raw = np.where(cm == 1, g, rng.integers(600, 16000, size=(H, W))).astype(np.uint16)
    filters = 9 if kind == "xtrans" else BAYER
    acc = op.rawhist_collect(32, raw, filters, 512, 16383)
    v = float(norm(g, 512, 16383))
    b = int(np.float32(1000) * np.log2(np.float32(v) + np.float32(1)))
    blk = 3 if kind == "xtrans" else 2
    n = (W // blk) * (H // blk)
    assert acc[0, b] == n and acc[0].sum() == n
    assert np.isclose(acc[1, b], n * v, rtol=1e-6) and np.isclose(acc[2, b], n * v * v, rtol=1e-6)
    assert not acc[3:].any()


def test_dispatch_rounding_and_non_mosaic_input():
    W, H = 131, 99                 # 65 x 49 blocks: groups of 8 reach 72 x 56
    raw = np.full((H, W), 8000, np.uint16)
    acc = op.rawhist_collect(64, raw, BAYER, 512, 16383)
    # blocks whose first site (2x, 2y) is inside: x <= 65, y <= 49
    assert acc[0].sum() == 66 * 50
    acc0 = op.rawhist_collect(64, raw, 0, 512, 16383)
    assert acc0[0, 0] == 72 * 56 and acc0[0].sum() == 72 * 56 and not acc0[1:].any()


def noisy_ramp(W, H, kind, a, b, seed):
    """a smooth ramp over the whole range with Poisson-Gaussian noise var = a*x + b
    on the white-normalised signal x, quantised to 14 bits (black 512)"""
    rng = np.random.default_rng(seed)
    x = np.linspace(0.01, 0.8, W)[None, :] * np.ones((H, 1))
    x = x + rng.normal(size=x.shape) * np.sqrt(a * x + b)
    return np.clip(np.round(512 + x * (16383 - 512)), 0, 16383).astype(np.uint16)


@pytest.mark.parametrize("kind,factor", [("bayer", 0.5), ("xtrans", 0.8)])
def test_fit_recovers_the_noise_model(kind, factor):
    a, b = 4e-4, 2e-5
    W, H = (1024, 768) if kind == "bayer" else (1026, 768)
    raw = noisy_ramp(W, H, kind, a, b, 5)
    acc = op.rawhist_collect(64, raw, 9 if kind == "xtrans" else BAYER, 512, 16383)
    fa, fb, n = op.nprof_fit(64, acc)
    print(f"[nprof] {kind}: intercept {fa:.4g} (want {factor * b:.4g}), slope {fb:.4g} (want {factor * a:.4g}), "
          f"{n} bins", flush=True)
    assert n > 20
    assert abs(fb - factor * a) < 0.1 * factor * a
    assert abs(fa - factor * b) < 0.3 * factor * b


def test_fit_placeholder_without_data():
    assert op.nprof_fit(64, np.zeros((5, 1000))) == (100.0, 1.0, 0)


def product_fit():
    L = C.CDLL(os.path.join(PKG_DIR, "modules", "nprof", "libnprof.so"))
    L.nprof_fit.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    return L.nprof_fit


def test_module_fit_equals_the_oracle():
    fit = product_fit()
    for kind, seed in (("bayer", 6), ("xtrans", 7)):
        raw = noisy_ramp(514 if kind == "xtrans" else 512, 384, kind, 3e-4, 1e-5, seed)
        acc = np.ascontiguousarray(op.rawhist_collect(64, raw, 9 if kind == "xtrans" else BAYER, 512, 16383))
        a, b = C.c_float(), C.c_float()
        n = fit(acc.ctypes.data, acc.shape[1], C.byref(a), C.byref(b))
        assert (a.value, b.value, n) == op.nprof_fit(64, acc)
