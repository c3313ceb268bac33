"""filmcurv's colour modes on the oracle (filmcurv/main.comp:68-167): the
colour spaces they use pinned by published values and by what the spaces are
built to do, and each mode by the property it exists for.

Pins of the spaces (oracle or_fc_space):
* oklab (colourspaces.glsl:21-49): Ottosson's published oklab of the D65
  white and of XYZ (1,0,0) and (0,1,0) (3 decimals) reached through the
  reference's xyz_to_rec2020 (matrices.h:50); XYZ (0,0,1) is left out, its
  long-wave cone response is negative and the reference clamps it to 0
  (colourspaces.glsl:28) where the published transform takes a signed cube
  root; forward then inverse is the identity;
* darktable UCS 22 (dtucs.glsl): the D65 white has chroma 0; xyY -> JCh ->
  xyY is the identity; and, what a uniform colour space is fit for, one
  chroma step of the Munsell chart (the chart of munsell.glsl) has nearly
  the same UCS chroma at every hue (coefficient of variation 8 to 19 % over
  the steps 2, 4 and 8, bound 25 %) and the chart's hues step around the UCS
  hue circle in one direction -- a transposed 2x2 mix gives about 55 % and
  hue steps of both signs;
* the Munsell chart (munsell.glsl:4-6, data): chroma 0 is the white of every
  hue, chroma grows outwards along each hue, and the hues go round the white
  once at every chroma; xy -> (hue, chroma) -> xy is the identity, and so is
  (hue, chroma) -> xy -> (hue, chroma) past the first chroma step.

Pins of the modes (on the node):
* 0 (UCS): the output keeps the input's UCS hue, takes the per-channel
  curve's UCS chroma at the input's lightness and the per-channel curve's
  luminance Y; gray stays gray at the weibull cdf;
* 2 (Munsell): the output keeps the input's Munsell hue, takes the
  per-channel curve's Munsell chroma and luminance Y;
* 4 (AgX): gray stays gray at the weibull cdf (inset and outset keep
  gray); before the outset the output has the per-channel curve's oklab
  lightness and chroma and the hue 0.4 of the way from the input's to the
  curve's;
* 5 (lum ratio): gray v goes to the tone-mapped luminance (weibull cdf, the
  rolloff and the shadow toe) on all channels; the oklab lightness is the
  input's times the cube root of the luminance ratio (BT.2020 luma weights);
  chroma 0 gives oklab's gray axis; at chroma 1 away from the skin-hue window the output
  has the per-channel curve's oklab a and b; the red trim only raises
  chroma, by at most its amount, and not at hues a radian from its centre.

Parity unpinned: the lum-ratio mode hands radians to a hue lerp that wraps
modulo 1 (main.comp:150 with main.comp:17-26), so hues between 1 and 1.66
radians come out below 1; the oracle and the kernel restate that as the
reference has it, and only their agreement holds it.
"""
import os
import re

import numpy as np
import pytest
from scipy import stats

import graphspec as gs
import oracle_py as op

ROOT = op.ROOT
X2R = np.array([1.71665119, -0.35567078, -0.25336628, -0.66668435, 1.61648124, 0.01576855, 0.01763986,
                -0.04277061, 0.94210312]).reshape(3, 3)           # matrices.h:50
R2X = np.array([0.636958048301290991, 0.144616903586208406, 0.168880975164172054, 0.26270021201126692,
                0.677998071518871148, 0.0593017164698619384, 4.9999999999999999e-17, 0.0280726930490874452,
                1.06098505771079066]).reshape(3, 3)               # matrices.h:49
LUMA = np.array([0.2627, 0.6780, 0.0593])                         # BT.2020 luma weights
# AgX inset of main.comp:33-37, written as rows of the product it computes
AGX = np.array([[0.856627153315983, 0.0951212405381588, 0.0482516061458583],
                [0.137318972929847, 0.761241990602591, 0.101439036467562],
                [0.11189821299995, 0.0767994186031903, 0.811302368396859]])
# and its outset (main.comp:38-42), not the inverse of the inset
AGX_OUT = np.array([[1.1271005818144368, -0.11060664309660323, -0.016493938717834573],
                    [-0.1413297634984383, 1.157823702216272, -0.016493938717834257],
                    [-0.14132976349843826, -0.11060664309660294, 1.2519364065950405]])
IL, K = 1.0 / 3.0, float(np.float32(1.2))


def chart():
    """the Munsell chart (40 hues x 21 chroma steps, x and y) from munsell.h"""
    txt = open(os.path.join(ROOT, "vkdt_amd", "csrc", "munsell.h")).read()
    vals = np.array([int(v) for v in re.findall(r"(\d+)u", txt[txt.index("#define MUNSELL_XY_TABLE"):])], np.uint32)
    assert vals.size == 40 * 21
    x = (vals & 0xffff).astype(np.uint16).view(np.float16).astype(np.float64).reshape(40, 21)
    y = (vals >> 16).astype(np.uint16).view(np.float16).astype(np.float64).reshape(40, 21)
    return x, y


def node(img, mode, light=3.0, contrast=1.2, bias=0.0, chroma=1.0, rolloff=0.0, red=0.0, yellow=0.0, blue=0.0,
         shadows=0.0, bits=64):
    """filmcurv/main on the oracle, rgb rows n x 3 in, n x 3 out"""
    img = np.asarray(img, np.float64).reshape(-1, 3)
    n = img.shape[0]
    rgba = np.ascontiguousarray(np.concatenate([img, np.ones((n, 1))], 1).reshape(1, n, 4))
    out = np.zeros_like(rgba)
    par = gs.mixed_blob([("f", light), ("f", contrast), ("f", bias), ("i", mode), ("f", chroma), ("f", rolloff),
                         ("f", red), ("f", yellow), ("f", blue), ("f", shadows)])
    op.node(bits, "filmcurv", "main", None, par, [op.orbuf(rgba, op.OR_F32, 4), op.orbuf(out, op.OR_F32, 4)], n, 1,
            storage_f16=False)
    return out[0, :, :3]


def weibull(v):
    return stats.weibull_min.cdf(np.maximum(v, 1e-7), K, scale=1.0 / 3.0)


def xyY(rgb):
    xyz = np.asarray(rgb) @ R2X.T
    return np.stack([xyz[:, 0] / xyz.sum(1), xyz[:, 1] / xyz.sum(1), xyz[:, 1]], 1)


def colours(seed, n=200, lo=0.02, hi=1.2):
    return np.random.default_rng(seed).uniform(lo, hi, size=(n, 3))


# ---- colour spaces -----------------------------------------------------------

def check_oklab_published_values():
    xyz = np.array([[0.950, 1.0, 1.089], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    lab = op.fc_space(64, 0, xyz @ X2R.T)
    ref = np.array([[1.000, 0.000, 0.000], [0.450, 1.236, -0.019], [0.922, -0.671, 0.263]])
    assert np.abs(lab - ref).max() < 1e-3, lab


def check_oklab_round_trip():
    c = colours(1, lo=0.0, hi=2.0)
    assert np.abs(op.fc_space(64, 1, op.fc_space(64, 0, c)) - c).max() < 1e-6


def check_ucs_white_and_round_trip():
    jch = op.fc_space(64, 2, np.array([[0.31271, 0.32902, 1.0]]))
    assert jch[0, 1] < 1e-5, "the D65 white has chroma"
    rng = np.random.default_rng(2)
    v = np.concatenate([rng.uniform(0.2, 0.45, (200, 2)), rng.uniform(0.01, 2.0, (200, 1))], 1)
    assert np.abs(op.fc_space(64, 3, op.fc_space(64, 2, v)) - v).max() < 1e-9


def check_ucs_is_uniform_over_the_munsell_chart():
    x, y = chart()
    for c in (2, 4, 8):
        v = np.stack([x[:, c], y[:, c], np.full(40, 0.2)], 1)
        jch = op.fc_space(64, 2, v)
        cv = jch[:, 1].std() / jch[:, 1].mean()
        assert cv < 0.25, f"UCS chroma of Munsell chroma step {c} varies by {cv:.2f}"
        d = np.diff(np.unwrap(np.concatenate([jch[:, 2], jch[:1, 2]])))
        assert (d < 0).all() or (d > 0).all(), "the chart's hues do not go round the UCS hue circle in order"


def check_munsell_round_trips():
    """(hue, chroma) -> xy -> (hue, chroma) outside the first chroma step,
    whose cells are triangles at the white (the two directions split them
    differently); xy -> (hue, chroma) -> xy everywhere"""
    rng = np.random.default_rng(3)
    hc = np.stack([rng.uniform(0, 1, 300), rng.uniform(1.02 / 21, 0.9, 300), np.zeros(300)], 1)
    xy = op.fc_space(64, 5, hc)
    back = op.fc_space(64, 4, xy)
    dh = (back[:, 0] - hc[:, 0] + 0.5) % 1.0 - 0.5
    assert np.abs(dh).max() < 1e-9 and np.abs(back[:, 1] - hc[:, 1]).max() < 1e-9
    hc[:100, 1] = rng.uniform(0.005, 1.0 / 21, 100)   # at the white itself the walk can give up (vec2(1))
    xy = op.fc_space(64, 5, hc)
    again = op.fc_space(64, 5, op.fc_space(64, 4, xy))
    assert np.abs(again[:, :2] - xy[:, :2]).max() < 1e-12


# ---- modes -------------------------------------------------------------------

def check_ucs_mode_keeps_hue_takes_curve_chroma_and_luminance():
    c0 = colours(4, hi=0.8)
    out = node(c0, 0)
    pc = weibull(c0)
    a0, a1, ao = xyY(c0), xyY(pc), xyY(out)
    assert np.allclose(ao[:, 2], a1[:, 2], rtol=1e-10)
    j0 = op.fc_space(64, 2, a0)
    jo = op.fc_space(64, 2, np.stack([ao[:, 0], ao[:, 1], a0[:, 2]], 1))
    j1 = op.fc_space(64, 2, np.stack([a1[:, 0], a1[:, 1], a0[:, 2]], 1))
    assert np.abs((jo[:, 2] - j0[:, 2] + np.pi) % (2 * np.pi) - np.pi).max() < 1e-7
    assert np.allclose(jo[:, 1], j1[:, 1], rtol=1e-7, atol=1e-9)


def check_ucs_mode_gray():
    v = np.linspace(0.01, 2.0, 50)
    out = node(np.stack([v, v, v], 1), 0)
    assert np.abs(out / weibull(v)[:, None] - 1).max() < 5e-4    # UCS white: chroma 2.5e-6


def check_munsell_mode_keeps_hue_takes_curve_chroma_and_luminance():
    c0 = colours(5, lo=0.05, hi=0.9)
    out = node(c0, 2)
    pc = weibull(c0)
    a0, a1, ao = xyY(c0), xyY(pc), xyY(out)
    assert np.allclose(ao[:, 2], a1[:, 2], rtol=1e-10)
    z = np.zeros((len(c0), 1))
    m0 = op.fc_space(64, 4, np.concatenate([a0[:, :2], z], 1))
    m1 = op.fc_space(64, 4, np.concatenate([a1[:, :2], z], 1))
    mo = op.fc_space(64, 4, np.concatenate([ao[:, :2], z], 1))
    past = m1[:, 1] > 1.0 / 21          # see check_munsell_round_trips
    assert past.sum() > 100
    assert np.abs((mo[past, 0] - m0[past, 0] + 0.5) % 1.0 - 0.5).max() < 1e-7
    assert np.abs(mo[:, 1] - m1[:, 1]).max() < 1e-7


def check_agx_mode_gray():
    v = np.linspace(0.01, 2.0, 50)
    out = node(np.stack([v, v, v], 1), 4)
    assert np.abs(out - weibull(v)[:, None]).max() < 1e-4


def check_agx_mode_lightness_chroma_and_hue_lerp():
    c0 = colours(6, hi=1.0)
    out = node(c0, 4)
    ins = c0 @ AGX.T
    lab0 = op.fc_space(64, 0, ins)
    lab1 = op.fc_space(64, 0, weibull(ins))
    labo = op.fc_space(64, 0, out @ np.linalg.inv(AGX_OUT).T)
    assert np.allclose(labo[:, 0], lab1[:, 0], atol=1e-7)
    C1, Co = np.hypot(lab1[:, 1], lab1[:, 2]), np.hypot(labo[:, 1], labo[:, 2])
    assert np.allclose(Co, C1, atol=1e-7)
    h0, h1, ho = (np.arctan2(l[:, 2], l[:, 1]) / (2 * np.pi) for l in (lab0, lab1, labo))
    want = h0 + 0.4 * ((h1 - h0 + 0.5) % 1.0 - 0.5)
    assert np.abs((ho - want + 0.5) % 1.0 - 0.5).max() < 1e-5   # oklab matrices invert to 8 digits


def lum_ratio_lum1(lum0, rolloff, shadows):
    w = weibull(lum0)
    w = w + rolloff * w * w * (1 - w)
    if shadows:
        t = np.clip((w - 0.3) / (0.0 - 0.3), 0, 1)
        w = w + (np.maximum(w, 1e-7) ** (1 - 0.5 * shadows) - w) * t * t * (3 - 2 * t)
    return w


def check_lum_ratio_gray_and_lightness():
    v = np.linspace(0.01, 2.0, 50)
    out = node(np.stack([v, v, v], 1), 5, rolloff=0.4, shadows=0.5)
    # oklab's white has a and b of 3e-4 (its matrices to 8 digits), which the
    # chroma path keeps: a tint of that order
    assert np.abs(out / lum_ratio_lum1(v, 0.4, 0.5)[:, None] - 1).max() < 5e-4
    c0 = colours(7, hi=1.0)
    out = node(c0, 5, rolloff=0.4, shadows=0.5, chroma=1.3)
    lum0 = c0 @ LUMA
    L0 = op.fc_space(64, 0, c0)[:, 0]
    Lo = op.fc_space(64, 0, out)[:, 0]
    assert np.allclose(Lo, L0 * np.cbrt(lum_ratio_lum1(lum0, 0.4, 0.5) / lum0), rtol=1e-7)


def check_lum_ratio_chroma_zero_is_gray():
    """chroma 0: the output lies on oklab's lightness axis (a = b = 0), which
    its 8-digit inverse maps to rgb in the ratio of (1, 0, 0)'s image"""
    out = node(colours(8), 5, chroma=0.0, rolloff=0.3)
    axis = op.fc_space(64, 1, np.array([[1.0, 0.0, 0.0]]))[0]
    r = out / axis
    assert np.abs(r / r[:, :1] - 1).max() < 1e-9


def check_lum_ratio_keeps_curve_ab_off_the_skin_window():
    c0 = colours(9, hi=1.0)
    out = node(c0, 5)
    lab1 = op.fc_space(64, 0, weibull(c0))
    h = np.arctan2(lab1[:, 2], lab1[:, 1])
    off = np.abs((h - 0.96 + np.pi) % (2 * np.pi) - np.pi) >= 0.7
    assert off.sum() > 50
    labo = op.fc_space(64, 0, out)
    assert np.abs(labo[off, 1:] - lab1[off, 1:]).max() < 1e-7


def check_lum_ratio_red_trim():
    c0 = colours(10, hi=1.0)
    base, trim = node(c0, 5), node(c0, 5, red=0.5)
    lb, lt = op.fc_space(64, 0, base), op.fc_space(64, 0, trim)
    Cb, Ct = np.hypot(lb[:, 1], lb[:, 2]), np.hypot(lt[:, 1], lt[:, 2])
    ok = Cb > 1e-3
    r = Ct[ok] / Cb[ok]
    assert (r > 1 - 1e-7).all() and (r < 1.5 + 1e-7).all()
    h = np.arctan2(lb[ok, 2], lb[ok, 1])
    far = np.abs((h - 0.7 + np.pi) % (2 * np.pi) - np.pi) >= 1.0
    assert far.sum() > 10 and np.abs(r[far] - 1).max() < 1e-7
    assert r.max() > 1.3


# the checks the planted-mistake test runs beside the stencil pins
CHECKS = [check_oklab_published_values, check_oklab_round_trip, check_ucs_white_and_round_trip,
          check_ucs_is_uniform_over_the_munsell_chart, check_munsell_round_trips,
          check_ucs_mode_keeps_hue_takes_curve_chroma_and_luminance, check_ucs_mode_gray,
          check_munsell_mode_keeps_hue_takes_curve_chroma_and_luminance, check_agx_mode_gray,
          check_agx_mode_lightness_chroma_and_hue_lerp, check_lum_ratio_gray_and_lightness,
          check_lum_ratio_chroma_zero_is_gray, check_lum_ratio_keeps_curve_ab_off_the_skin_window,
          check_lum_ratio_red_trim]


@pytest.mark.parametrize("check", CHECKS, ids=[c.__name__[6:] for c in CHECKS])
def test_filmcurv_mode_pin(check):
    check()


def test_munsell_chart_shape():
    x, y = chart()
    assert np.abs(x[:, 0] - x[0, 0]).max() == 0 and np.abs(y[:, 0] - y[0, 0]).max() == 0
    assert abs(x[0, 0] - 0.31271) < 1e-3 and abs(y[0, 0] - 0.32902) < 1e-3, "chroma 0 is not the D65 white"
    r = np.hypot(x - x[0, 0], y - y[0, 0])
    assert (np.diff(r, axis=1) > 0).all(), "chroma does not grow outwards along a hue"
    ang = np.mod(2 * np.pi - 2.52 - np.arctan2(y - y[0, 0], x - x[0, 0]), 2 * np.pi)
    for c in range(1, 21):
        d = np.diff(np.concatenate([ang[:, c], ang[:1, c]]))
        assert (d < 0).sum() == 1, f"chroma step {c}: the hues do not go round the white once"


def test_mode_out_of_range_is_refused():
    img = np.ascontiguousarray(np.ones((1, 2, 4)))
    par = gs.mixed_blob([("f", 3), ("f", 1.2), ("f", 0), ("i", 6)])
    with pytest.raises(RuntimeError):
        op.node(64, "filmcurv", "main", None, par, [op.orbuf(img, op.OR_F32, 4), op.orbuf(img.copy(), op.OR_F32, 4)],
                2, 1, storage_f16=False)
