"""Pins of the oracle's hist and autoexp restatements (src/pipe/modules/hist,
autoexp) against their closed forms:

  * hist/collect puts each pixel's r, g and b into row
    clamp(floor((1 - v) * 600 + 0.5), 0, 599) of column floor(x * 1000 / W),
    counting 1, 1024 and 1048576 -- checked against a numpy histogram of a
    random image (every count field below 1024, so no field carries);
  * hist/map is pow(field / 255, 0.3) per 10-bit field;
  * autoexp's bins are 1 + floor((log2 l + 15) / 20 * 254) for the rec2020
    luminance l >= 2^-15 (numpy over a random image, thinned by the radial
    mask restated with numpy);
  * autoexp/lum of a histogram with every count in bin k is
    2^((k - 1) / 254 * 20 - 15) on frame 0 or with reset, and moves towards
    a new value by 1 - exp(-timediff * v) (v = vup brighter, vdown darker);
  * autoexp/exp scales rgb by clamp(0.18 / lum, 1e-3, 1e3) and sets alpha 1.
"""
import numpy as np
import pytest

import oracle_py as op


def collect(img, ow=1000, oh=600, bits=64):
    out = np.zeros((oh, ow))
    op.node(bits, "hist", "collect", None, None, [op.orbuf(img, op.OR_F16, 4), op.orbuf(out, op.OR_F32, 1)],
            img.shape[1], img.shape[0], storage_f16=False)
    return out


def test_hist_collect_closed_form():
    rng = np.random.default_rng(1)
    H, W = 40, 200
    img = rng.uniform(-0.2, 1.2, size=(H, W, 4))
    got = collect(img).astype(np.uint64)
    want = np.zeros((600, 1000), np.uint64)
    col = np.clip((np.arange(W) * (1000.0 / W)).astype(int), 0, 999)
    for c, w in enumerate((1, 1024, 1048576)):
        rows = np.clip(np.floor((1.0 - img[..., c]) * 600 + 0.5), 0, 599).astype(int)
        np.add.at(want, (rows, np.broadcast_to(col, (H, W))), w)
    assert np.array_equal(got, want)
    for s in (0, 10, 20):
        assert ((got >> s) & 1023).sum() == H * W


def test_hist_map_closed_form():
    cnt = np.array([[0, 1, 255 | (7 << 10) | (1000 << 20), 1023 * (1 + 1024 + 1048576)]], np.float64)
    out = np.zeros((1, 4, 4))
    op.node(64, "hist", "map", None, None, [op.orbuf(cnt, op.OR_F32, 1), op.orbuf(out, op.OR_F16, 4)], 4, 1,
            storage_f16=False)
    c = cnt.astype(np.uint64)[0]
    for k, s in enumerate((0, 10, 20)):
        np.testing.assert_allclose(out[0, :, k], (((c >> s) & 1023) / 255.0) ** 0.3, rtol=1e-12)
    assert (out[0, :, 3] == 1.0).all()


def ae_hist(img, bits=64):
    h = np.zeros((16, 17))
    op.node(bits, "autoexp", "hist", None, None, [op.orbuf(img, op.OR_F16, 4), op.orbuf(h, op.OR_F32, 1)],
            (img.shape[1] + 15) // 16 * 8, (img.shape[0] + 15) // 16 * 8, storage_f16=False)
    return h.ravel()


def mask_np(W, H):
    y, x = np.mgrid[0:H, 0:W]
    d = np.sqrt((W // 2 - x) ** 2.0 + (H // 2 - y) ** 2.0)
    e0, e1 = 0.1 * W / 2, W // 2
    t = np.clip((d - e0) / (e1 - e0), 0, 1)
    s = t * t * (3 - 2 * t)
    mx = np.maximum((s * 13).astype(int), 1)
    my = np.maximum((s * 7).astype(int), 1)
    return (x % mx == 0) & (y % my == 0)


def test_autoexp_hist_closed_form():
    rng = np.random.default_rng(2)
    H, W = 90, 160
    img = np.exp2(rng.uniform(-17, 4, size=(H, W, 4)))
    got = ae_hist(img)
    lum = 0.262700212 * img[..., 0] + 0.677998072 * img[..., 1] + 0.0593017165 * img[..., 2]
    m = mask_np(W, H) & (lum >= 2.0 ** -15)
    bins = np.clip(((np.log2(lum[m]) + 15) / 20 * 254 + 1).astype(int), 0, 255)
    want = np.bincount(bins, minlength=256).astype(float)
    assert np.array_equal(got[:256], want)
    assert got[256] == m.sum() and got[257:].sum() == 0


def lum_node(h, prev, frame, reset=0, vup=3.0, vdown=5.0, timediff=0.016, bits=64):
    par = np.zeros(5, np.uint32)
    par[0:3] = np.array([vup, vdown, timediff], np.float32).view(np.uint32)
    par[3:5] = [reset, frame]
    lum = np.full((8, 8), prev)
    op.node(bits, "autoexp", "lum", None, par, [op.orbuf(np.zeros((1, 1, 4)), op.OR_F16, 4),
                                                op.orbuf(np.ascontiguousarray(np.reshape(h, (16, 17))), op.OR_F32, 1),
                                                op.orbuf(lum, op.OR_F32, 1)], 1, 1,
            storage_f16=False)
    return lum[0, 0]


@pytest.mark.parametrize("k", [1, 100, 254, 255])
def test_autoexp_lum_closed_form(k):
    h = np.zeros(17 * 16)
    h[k] = 5000.0
    h[256] = 5000
    target = 2.0 ** ((k - 1) / 254 * 20 - 15)
    assert abs(lum_node(h, 7.0, 0) - target) <= 1e-12 * target         # frame 0: no history
    assert abs(lum_node(h, 7.0, 3, reset=1) - target) <= 1e-12 * target
    for prev in (target * 0.5, target * 2.0):                             # towards target at vup / vdown
        v = 3.0 if target > prev else 5.0
        want = prev + (target - prev) * (1 - np.exp(-float(np.float32(0.016)) * v))
        assert abs(lum_node(h, prev, 3) - want) <= 1e-9 * target
    assert abs(lum_node(h, np.nan, 3) - target) <= 1e-12 * target        # NaN history: restart


def test_autoexp_exp_closed_form():
    img = np.random.default_rng(3).uniform(0, 2, size=(6, 8, 4))
    for lum, e in ((0.36, 0.5), (1e-9, 1e3), (1e9, 1e-3)):
        out = np.zeros_like(img)
        op.node(64, "autoexp", "exp", None, None,
                [op.orbuf(img, op.OR_F16, 4), op.orbuf(np.zeros((16, 17)), op.OR_F32, 1),
                 op.orbuf(np.full((8, 8), lum), op.OR_F32, 1), op.orbuf(out, op.OR_F16, 4)], 8, 6,
                storage_f16=False)
        np.testing.assert_allclose(out[..., :3], img[..., :3] * e, rtol=1e-12)
        assert (out[..., 3] == 1.0).all()
