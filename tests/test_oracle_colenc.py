"""Pins of the oracle's colenc (export colour encoding, colenc/main.comp)
against the published standards it implements:

  * sRGB (IEC 61966-2-1): 0 -> 0, 1 -> 1, the linear segment 12.92 v below
    0.0031308, 0.5 -> 0.7353569;
  * the 709 curve as the reference writes it (exponent 1/2.2): 1 -> 1 and
    4.5 v below 0.018053968510807;
  * PQ (SMPTE ST 2084): equals the textbook ((c1 + c2 p)/(1 + c3 p))^m2 with
    p = v^m1, and 1 -> 1;
  * HLG (BT.2100): 1/12 -> 0.5, 1 -> 1;
  * primaries: rec2020 white stays white (to the 8-digit constants) in
    rec709, adobe rgb and P3, and
    maps to D65 (x 0.3127, y 0.3290) in XYZ; the rec2020 -> XYZ matrix times
    the reference's XYZ -> rec2020 matrix is the identity.
"""
import numpy as np
import pytest

import graphspec as gs
import oracle_py as op


def enc(vals, prim, trc, bits=64):
    v = np.asarray(vals, np.float64)
    img = np.zeros((1, len(v), 4))
    img[0, :, :3] = v[:, None] if v.ndim == 1 else v
    out = np.zeros_like(img)
    op.node(bits, "colenc", "main", None, gs.mixed_blob([("i", prim), ("i", trc)]),
            [op.orbuf(img, op.OR_F32, 4), op.orbuf(out, op.OR_F32, 4)], len(v), 1, storage_f16=False)
    return out[0, :, :3]


def test_srgb_curve():
    v = [0.0, 0.001, 0.0031308 * 0.999, 0.5, 1.0]
    got = enc(v, 2, 2)[:, 0]
    assert got[0] == 0.0 and abs(got[4] - 1.0) < 1e-12
    assert abs(got[1] - 0.01292) < 1e-12 and abs(got[2] - 12.92 * 0.0031308 * 0.999) < 1e-12
    assert abs(got[3] - 0.7353569830524495) < 1e-9


def test_709_curve_as_written():
    got = enc([0.01, 1.0], 2, 1)[:, 0]
    assert abs(got[0] - 0.045) < 1e-12 and abs(got[1] - 1.0) < 1e-12


def test_pq_is_st2084():
    v = np.array([0.0, 1e-4, 0.01, 0.18, 0.5, 1.0])
    got = enc(v, 2, 3)[:, 0]
    m1, m2 = 2610 / 16384, 2523 / 4096 * 128
    c1, c2, c3 = 3424 / 4096, 2413 / 4096 * 32, 2392 / 4096 * 32
    p = v ** m1
    np.testing.assert_allclose(got, ((c1 + c2 * p) / (1 + c3 * p)) ** m2, rtol=1e-12)
    assert abs(got[-1] - 1.0) < 1e-12


def test_hlg_bt2100():
    got = enc([1.0 / 12.0, 1.0], 2, 5)[:, 0]
    assert abs(got[0] - 0.5) < 1e-12 and abs(got[1] - 1.0) < 1e-6


@pytest.mark.parametrize("prim", [1, 3, 4])
def test_white_stays_white(prim):
    # the reference's constants carry 8 digits; its rec709 rows sum to 1 within 2.4e-4
    np.testing.assert_allclose(enc([1.0], prim, 0)[0], 1.0, atol=5e-4)


def test_xyz_white_and_inverse():
    # D65 from its chromaticity (0.3127, 0.3290), as rec2020 defines its white
    np.testing.assert_allclose(enc([1.0], 5, 0)[0], [0.3127 / 0.329, 1.0, (1 - 0.3127 - 0.329) / 0.329], atol=1e-5)
    # the rec2020 -> XYZ rows applied to the columns of the reference's XYZ -> rec2020 matrix
    cols = np.asarray(gs.XYZ_TO_REC2020, np.float64).T
    np.testing.assert_allclose(enc(cols, 5, 0), np.eye(3), atol=2e-6)
