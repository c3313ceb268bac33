"""Pins of the oracle's RCD demosaic (ratio corrected demosaicing, demosaic
method 1: rcd_conv.comp, rcd_fill.comp) on RGGB mosaics:

  * per-colour constants come back as those constants (to binary16 and the
    1e-5 regularisers) away from the image border, across tile seams -- the
    colour of each site, the white balance in and out, and the signs of the
    colour-difference estimates;
  * a planar grey field comes back exactly (to binary16) with a white
    balance applied and removed: the ratio-corrected green estimate is exact
    on a plane, and the diagonal and axial colour differences average to the
    centre's -- a misplaced low-pass texel, a wrong ratio or a one-sided
    difference fails this.
"""
import numpy as np

import graphspec as gs
import oracle_py as op
from vkdt_amd import synth


def rcd(mos, wb=(2.0, 1.0, 1.5, 1.0), bits=64):
    H, W = mos.shape
    p = gs.DevelopParams(kind="bayer", wb=wb, dm_method=1)
    dmpc = gs.mixed_blob([("f", v) for v in p.wb] + [("i", p.filters)])
    bufs = {"in": (W, H, 1, op.OR_F16, 1), "vh": (W, H, 1, op.OR_F16, 1), "pq": (W // 2, H, 1, op.OR_F16, 1),
            "lp": (W // 2, H, 1, op.OR_F16, 1), "out": (W, H, 4, op.OR_F16, 1)}
    data = {k: np.zeros((h, w, c) if c == 4 else (h, w)) for k, (w, h, c, f, pl) in bufs.items()}
    data["in"][...] = mos
    ob = lambda names: [op.orbuf(data[n], bufs[n][3], bufs[n][2]) for n in names]
    op.node(bits, "demosaic", "rcd_conv", None, None, ob(["in", "vh", "pq", "lp"]), W, H)
    op.node(bits, "demosaic", "rcd_fill", dmpc, None, ob(["in", "vh", "pq", "lp", "out"]),
            (W + 57) // 58 * 8, (H + 25) // 26 * 8)
    return data["out"]


def test_per_colour_constants():
    W, H = 128, 64
    cm = synth.channel_map(W, H, "bayer")
    rgb = (0.5, 0.25, 0.125)
    mos = np.where(cm == 0, rgb[0], np.where(cm == 1, rgb[1], rgb[2]))
    out = rcd(mos)
    # the fill reads the mosaic with texelFetch (zero outside), so the sites
    # within 5 of the image border see a dark frame; inside, every tile seam
    # included, the constants come back
    inner = out[5:-5, 5:-5, :3]
    np.testing.assert_allclose(inner, np.broadcast_to(rgb, inner.shape), rtol=2e-3)
    np.testing.assert_array_equal(out[..., 3], 1.0)


def test_planar_grey_is_exact():
    """ratio correction makes the green estimate of a plane exact: for
    v = a + k x + l y the ratio 2 lp(x)/(lp(x) + lp(x-2)) times the
    neighbour gives the centre value, and equal gradients on both sides of
    each diagonal make the colour differences average to the centre's"""
    W, H = 128, 64
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    mos = 0.2 + 0.002 * xx + 0.003 * yy
    out = rcd(mos)
    # away from the image border and from the first and last two written rows
    # and columns of each 58x26 tile: those read tile entries computed from
    # reads past the tile's array (0 here, undefined in the reference)
    keep = np.zeros((H, W), bool)
    keep[5:-5, 5:-5] = True
    keep &= ~np.isin(yy % 26, [0, 1, 24, 25]) & ~np.isin(xx % 58, [0, 1, 56, 57])
    for ch in range(3):
        np.testing.assert_allclose(out[..., ch][keep], mos[keep], rtol=2e-3)
