"""The oracle's baseline JPEG encoder (oracle/jpeg.c) against libjpeg itself:
PIL's JPEG writer links libjpeg(-turbo), the library the reference's o-jpg
calls (o-jpg/main.c:101-160, optimised Huffman coding, islow DCT), so the
same image, quality and sampling must give byte-identical quantisation
tables, Huffman tables, frame header and entropy-coded data.  Covers 4:4:4,
4:2:2 and 4:2:0, extents that end inside an MCU (edge replication and the
dummy blocks of the last MCUs), flat, smooth and noisy content."""
import ctypes as C
import io
import os

import numpy as np
import pytest

PIL = pytest.importorskip("PIL.Image")
LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_build", "liboracle64.so")


def oracle_jpg(img, q, hs, vs):
    L = C.CDLL(LIB)
    L.or_jpg_encode.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]
    L.or_jpg_free.argtypes = [C.c_void_p]
    H, W = img.shape[:2]
    rgba = np.ascontiguousarray(np.dstack([img, np.full((H, W), 255, np.uint8)]))
    p, n = C.c_void_p(), C.c_size_t()
    assert L.or_jpg_encode(rgba.ctypes.data, W, H, W * 4, q, hs, vs, C.byref(p), C.byref(n)) == 0
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
