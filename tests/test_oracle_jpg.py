"""The oracle's baseline JPEG encoder (oracle/jpeg.c) against libjpeg itself:
PIL's JPEG writer links libjpeg(-turbo), the library the reference's o-jpg
calls (o-jpg/main.c:101-160, optimised Huffman coding, islow DCT), so the
same image, quality and sampling must give byte-identical quantisation
tables, Huffman tables, frame header and entropy-coded data.  Covers 4:4:4,
4:2:2 and 4:2:0, extents that end inside an MCU (edge replication and the
dummy blocks of the last MCUs), flat, smooth and noisy content.

The settings o-jpg takes outside quality 80..95 (main.c:142-146: the float
DCT above 95, the fast integer DCT below 50, smoothing 20/40/60 below
80/60/40), which PIL cannot select, are checked against the same library
through its C API (tests/libjpeg_ref.py): the whole file byte for byte,
including flat images whose float coefficients fall exactly halfway between
two integers (libjpeg-turbo's float quantiser rounds them to even)."""
import ctypes as C
import io
import os

import numpy as np
import pytest

PIL = pytest.importorskip("PIL.Image")
LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_build", "liboracle64.so")


def oracle_jpg(img, q, hs, vs, dct=0, smooth=0):
    """dct 0 islow, 1 ifast, 2 float (libjpeg's dct_method); smooth: its smoothing_factor"""
    L = C.CDLL(LIB)
    L.or_jpg_encode_ex.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]
    L.or_jpg_free.argtypes = [C.c_void_p]
    H, W = img.shape[:2]
    rgba = np.ascontiguousarray(np.dstack([img, np.full((H, W), 255, np.uint8)]))
    p, n = C.c_void_p(), C.c_size_t()
    assert L.or_jpg_encode_ex(rgba.ctypes.data, W, H, W * 4, q, hs, vs, dct, smooth, C.byref(p), C.byref(n)) == 0
    out = C.string_at(p, n.value)
    L.or_jpg_free(p)
    return out


def pil_jpg(img, q, sub):
    f = io.BytesIO()
    PIL.fromarray(img).save(f, "JPEG", quality=q, subsampling=sub, optimize=True)
    return f.getvalue()


def segments(b):
    """marker -> list of payloads; the scan's payload includes its entropy-coded data"""
    out, i = {}, 2
    assert b[:2] == b"\xff\xd8" and b[-2:] == b"\xff\xd9"
    while i < len(b) - 2:
        assert b[i] == 0xff
        m, n = b[i + 1], b[i + 2] << 8 | b[i + 3]
        seg = b[i + 4:i + 2 + n]
        i += 2 + n
        if m == 0xda:
            j = b.index(b"\xff\xd9", i) if b"\xff\xd9" in b[i:] else len(b)
            seg += b[i:len(b) - 2]
            i = len(b) - 2
        out.setdefault(m, []).append(seg)
    return out


def images():
    rng = np.random.default_rng(0)
    yy, xx = np.mgrid[0:75, 0:133]
    yield "smooth-133x75", np.dstack([(xx * 1.9) % 256, (yy * 3.1) % 256, ((xx + yy) * 1.3) % 256]).astype(np.uint8)
    yield "noisy-90x61", rng.integers(0, 256, size=(61, 90, 3)).astype(np.uint8)
    yield "flat-17x9", np.full((9, 17, 3), (200, 30, 90), np.uint8)
    yield "tiny-1x1", np.array([[[250, 4, 128]]], np.uint8)
    g = np.clip(rng.normal(0, 1, size=(97, 161, 3)).cumsum(0).cumsum(1) * 2 + 128, 0, 255)
    yield "walk-161x97", g.astype(np.uint8)


@pytest.mark.parametrize("q,sub,hs,vs", [(95, 0, 1, 1), (91, 1, 2, 1), (85, 2, 2, 2), (50, 2, 2, 2), (100, 0, 1, 1)])
def test_matches_libjpeg(q, sub, hs, vs):
    for name, img in images():
        a, b = segments(oracle_jpg(img, q, hs, vs)), segments(pil_jpg(img, q, sub))
        for m in (0xdb, 0xc0, 0xc4, 0xda):
            assert a[m] == b[m], (name, hex(m))


import libjpeg_ref  # noqa: E402

# (quality, hs, vs, dct, smoothing) as o-jpg/main.c:139-146 derives them, and
# the combinations libjpeg allows beyond them (smoothing at 4:4:4 and 4:2:2)
SETTINGS = [(97, 1, 1, 2, 0), (100, 1, 1, 2, 0), (75, 2, 2, 0, 20), (55, 2, 2, 0, 40), (45, 2, 2, 1, 40),
            (30, 2, 2, 1, 60), (20, 2, 2, 1, 60), (70, 1, 1, 0, 20), (45, 2, 1, 1, 40), (85, 1, 1, 2, 30)]


@pytest.mark.parametrize("q,hs,vs,dct,smooth", SETTINGS)
def test_dct_methods_and_smoothing_match_libjpeg(q, hs, vs, dct, smooth):
    if not libjpeg_ref.available():
        pytest.skip("libjpeg (pillow.libs) not loadable with the version-62 layout")
    for name, img in images():
        a, b = oracle_jpg(img, q, hs, vs, dct, smooth), libjpeg_ref.encode(img, q, hs, vs, dct, smooth)
        assert a == b, (name, len(a), len(b))


def test_float_ties_round_to_even():
    """flat blocks whose DC is exactly k + 1/2 quantiser steps (64 (v - 128)
    = 128 (k + 1/2) at quantiser 16): libjpeg-turbo rounds them to even"""
    if not libjpeg_ref.available():
        pytest.skip("libjpeg (pillow.libs) not loadable with the version-62 layout")
    for v in (129, 131, 133, 127, 125):
        img = np.full((16, 24, 3), v, np.uint8)
        assert oracle_jpg(img, 50, 1, 1, 2, 0) == libjpeg_ref.encode(img, 50, 1, 1, 2, 0), v
