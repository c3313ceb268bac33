"""filmcurv's colour modes on the GPU through the C ABI (filmcurv/main.comp:
68-167): every mode on one image against the oracle's float build, and a
whole graph in the UCS and lum-ratio modes end to end.

The per-node bound is the twin count: texels more than parity.MAX_NODE_ULP
binary16 ulps from the float oracle number at most twice those of the
oracle's perturbed float twins (every transcendental moved by one ulp) plus
two.  The modes branch on their data (the Munsell chart search and its
give-up value, the lum-ratio mode's skin-hue window and its hue wrap), so a
last-bit difference can take the other side there; the twins show how many
texels of the image sit that close.

Image: 192x128 rgba f16 of colours as the colour module hands them over
(random rgb in [0, 1.5], dark and bright grays, near-grays and a few
slightly negative channels)."""
import ctypes as C

import numpy as np
import pytest

import graphspec as gs
import oracle_py as op
import parity
from vkdt_amd import synth

pytestmark = pytest.mark.gpu

W, H = 192, 128
LUM_RATIO = dict(chroma=1.3, rolloff=0.3, red=0.2, yellow=-0.1, blue=0.1, shadows=0.4)


def image(seed=5):
    rng = np.random.default_rng(seed)
    img = rng.uniform(0.0, 1.5, size=(H, W, 4))
    g = rng.uniform(0.0, 2.0, size=(H // 4, W))
    img[: H // 4, :, :3] = g[..., None]                                   # grays
    img[H // 4: H // 2, :, :3] = g[..., None] * (1 + rng.uniform(-0.02, 0.02, size=(H // 4, W, 3)))
    img[H // 2: H // 2 + 8, :, :3] -= 0.05                                # slightly negative channels
    img[..., 3] = 1.0
    return img.astype(np.float16)


def params(mode, chroma=1.0, rolloff=0.0, red=0.0, yellow=0.0, blue=0.0, shadows=0.0):
    return gs.mixed_blob([("f", 3.0), ("f", 1.2), ("f", 0.0), ("i", mode), ("f", chroma), ("f", rolloff),
                          ("f", red), ("f", yellow), ("f", blue), ("f", shadows)])


def oracle(bits, img, par, seed=None):
    if seed is not None:
        op.lib(bits).or_set_perturb_seed(C.c_uint32(seed))
    src = np.ascontiguousarray(img, np.float64)
    out = np.zeros((H, W, 4), np.float64)
    op.node(bits, "filmcurv", "main", None, par, [op.orbuf(src, op.OR_F16, 4), op.orbuf(out, op.OR_F16, 4)], W, H)
    return out


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4, 5])
def test_mode_vs_oracle32(hip, mode):
    img = image()
    par = params(mode, **(LUM_RATIO if mode == 5 else {}))
    print(f"[gpu_filmcurv] mode {mode} {W}x{H}", flush=True)
    bufs = {"raw": (W, H, 4, gs.F16, 1), "fc": (W, H, 4, gs.F16, 1)}
    nodes = [dict(module="filmcurv", kernel="main", pc=None, params=par, conn=["raw", "fc"], wd=W, ht=H, dp=1)]
    gpu = parity.run_gpu_graph(hip, bufs, nodes, img)["fc"].astype(np.float64)
    ref = oracle(32, img, par)
    d = parity.f16_ulps(gpu[..., :3], ref[..., :3]).max(-1)
    n_gpu = int((d > parity.MAX_NODE_ULP).sum())
    n_twin = [int((parity.f16_ulps(oracle("32p", img, par, s)[..., :3], ref[..., :3]).max(-1)
                   > parity.MAX_NODE_ULP).sum()) for s in parity.TWIN_SEEDS]
    fin = np.isfinite(ref[..., :3]).all(-1)
    print(f"[gpu_filmcurv] mode {mode}: {n_gpu} texels beyond {parity.MAX_NODE_ULP} ulp (twins {n_twin}), "
          f"median {float(np.median(d)):.3g} ulp, {int((~fin).sum())} non-finite in the oracle", flush=True)
    assert np.isfinite(gpu[..., :3][fin]).all()
    assert (gpu[..., 3] == 1).all()
    assert n_gpu <= 2 * max(n_twin) + 2
    if mode in (1, 3):
        assert n_gpu == 0


@pytest.mark.parametrize("mode", [0, 5])
def test_graph_end_to_end(hip, mode):
    """demosaic -> colour -> filmcurv (colour mode) -> llap on a Bayer
    mosaic against the oracle's whole-graph run"""
    Wg, Hg = 256, 192
    cfa, _ = synth.mosaic(Wg, Hg, 41 + mode, "bayer")
    na, nb = synth.noise_params(2e-4, 1e-6)
    p = gs.DevelopParams(kind="bayer", noise_a=na, noise_b=nb, fc_colour=mode)
    bufs, nodes = gs.build(Wg, Hg, p)
    print(f"[gpu_filmcurv] graph {Wg}x{Hg} colour mode {mode}: {len(nodes)} nodes", flush=True)
    host = parity.run_gpu_graph(hip, bufs, nodes, cfa)
    parity.e2e_check(cfa, p, host["final"], log=lambda s: print(s, flush=True), label=f" mode {mode}")
