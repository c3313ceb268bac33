"""Pins of the oracle's demosaic/halfsize and shared/resample (the half-size
and scaled-export paths of demosaic/main.c):

  * halfsize of per-colour constants returns those constants on the block
    grid (Bayer gather order, X-Trans block parity);
  * resample at the input's own extent is the identity;
  * its nine-tap form equals the sixteen-tap Catmull-Rom cubic evaluated
    directly (brute force) away from the mirrored border, for down- and
    up-scaling -- the merge of the middle taps into one bilinear tap is exact
    while both middle weights are positive;
  * it reproduces linear functions there (Catmull-Rom's partition of unity
    and first moment).
"""
import numpy as np
import pytest

import graphspec as gs
import oracle_py as op
from vkdt_amd import synth


@pytest.mark.parametrize("kind", ["bayer", "xtrans"])
def test_halfsize_per_colour_constants(kind):
    W, H = (32, 24) if kind == "bayer" else (33, 24)
    cm = synth.channel_map(W, H, kind)
    rgb = (0.625, 0.25, 0.125)
    mos = np.where(cm == 0, rgb[0], np.where(cm == 1, rgb[1], rgb[2])).astype(np.float64)
    blk = 3 if kind == "xtrans" else 2
    out = np.zeros((H // blk, W // blk, 4))
    pc = gs.mixed_blob([("f", 1.0)] * 4 + [("i", 9 if kind == "xtrans" else synth.BAYER_FILTERS)])
    op.node(64, "demosaic", "halfsize", pc, None, [op.orbuf(mos, op.OR_F16, 1), op.orbuf(out, op.OR_F16, 4)],
            W // blk, H // blk)
    np.testing.assert_allclose(out[..., :3], np.broadcast_to(rgb, out[..., :3].shape), atol=1e-12)
    np.testing.assert_array_equal(out[..., 3], 1.0)


def resample(img, ow, oh, bits=64):
    out = np.zeros((oh, ow, 4))
    op.node(bits, "shared", "resample", None, None,
            [op.orbuf(np.ascontiguousarray(img), op.OR_F32, 4), op.orbuf(out, op.OR_F32, 4)], ow, oh,
            storage_f16=False)
    return out


def catmull_rom_16(img, ow, oh, xs, ys):
    """direct evaluation at output texels xs x ys: four taps per axis at the
    sample position with the Catmull-Rom cubic's weights (taps must be inside)"""
    H, W = img.shape[:2]
    out = np.zeros((len(ys), len(xs), 4))

    def taps(s, n):
        t1 = np.floor(s - 0.5) + 0.5
        f = s - t1
        w = [f * (-0.5 + f * (1 - 0.5 * f)), 1 + f * f * (-2.5 + 1.5 * f), f * (0.5 + f * (2 - 1.5 * f)),
             f * f * (-0.5 + 0.5 * f)]
        i1 = int(t1 - 0.5)
        idx = [i1 - 1, i1, i1 + 1, i1 + 2]
        assert 0 <= idx[0] and idx[3] < n
        return idx, w

    for a, y in enumerate(ys):
        ty, wy = taps((y + 0.5) / oh * H, H)
        for b, x in enumerate(xs):
            tx, wx = taps((x + 0.5) / ow * W, W)
            acc = np.zeros(4)
            for j in range(4):
                for i in range(4):
                    acc += img[ty[j], tx[i]] * wx[i] * wy[j]
            out[a, b] = acc
    return out


def test_resample_identity_at_equal_extent():
    rng = np.random.default_rng(2)
    img = rng.uniform(0, 1, size=(20, 28, 4))
    np.testing.assert_allclose(resample(img, 28, 20), img, atol=1e-12)


@pytest.mark.parametrize("ow,oh", [(17, 11), (45, 37)])
def test_resample_equals_sixteen_tap_catmull_rom(ow, oh):
    rng = np.random.default_rng(3)
    img = rng.uniform(0, 1, size=(24, 30, 4))
    # interior: every tap inside the image
    ys = [y for y in range(oh) if 2 <= (y + 0.5) / oh * 24 <= 24 - 3]
    xs = [x for x in range(ow) if 2 <= (x + 0.5) / ow * 30 <= 30 - 3]
    got, ref = resample(img, ow, oh), catmull_rom_16(img, ow, oh, xs, ys)
    np.testing.assert_allclose(got[np.ix_(ys, xs)], ref, atol=1e-12)


@pytest.mark.parametrize("ow,oh", [(13, 9), (50, 41)])
def test_resample_reproduces_linear_functions(ow, oh):
    H, W = 24, 30
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    img = np.stack([0.1 + 0.01 * xx, 0.2 + 0.02 * yy, 0.3 + 0.005 * xx - 0.003 * yy, np.ones_like(xx)], -1)
    got = resample(img, ow, oh)
    oy, ox = np.mgrid[0:oh, 0:ow].astype(np.float64)
    sx, sy = (ox + 0.5) / ow * W - 0.5, (oy + 0.5) / oh * H - 0.5      # texel coordinates of the samples
    want = np.stack([0.1 + 0.01 * sx, 0.2 + 0.02 * sy, 0.3 + 0.005 * sx - 0.003 * sy, np.ones_like(sx)], -1)
    inner = (sx >= 1) & (sx <= W - 3) & (sy >= 1) & (sy <= H - 3)
    np.testing.assert_allclose(got[inner], want[inner], atol=1e-12)
