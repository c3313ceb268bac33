"""libjpeg itself as the checker of the oracle's JPEG encoder, for the
settings PIL's writer cannot select: libjpeg-turbo, the build that Pillow
ships (pillow.libs/libjpeg-*.so, JPEG_LIB_VERSION 62), driven through its C
API with ctypes -- the calls the reference's o-jpg makes
(o-jpg/main.c:120-166: set_defaults, set_quality, luma sampling, dct_method,
smoothing_factor, optimize_coding, 300 dpi, rgb scanlines).

The compress struct is addressed by the field offsets of the version-62
layout on x86-64 (520 bytes); `available()` checks them on the defaults
jpeg_set_defaults writes before any test relies on them, in a child
process first (a struct size the library rejects ends the process)."""
import ctypes as C
import functools
import glob
import os
import struct
import subprocess
import sys

import numpy as np

SIZE = 520
OFF_WIDTH, OFF_GAMMA, OFF_PREC, OFF_COMP = 48, 64, 72, 88
OFF_OPTIMIZE, OFF_SMOOTH, OFF_DCT, OFF_JFIF = 264, 272, 276, 288
JDCT_ISLOW, JDCT_IFAST, JDCT_FLOAT = 0, 1, 2


def _path():
    import PIL
    libs = glob.glob(os.path.join(os.path.dirname(os.path.dirname(PIL.__file__)), "pillow.libs", "libjpeg-*.so*"))
    return libs[0] if libs else None


@functools.lru_cache(None)
def _lib():
    L = C.CDLL(_path())
    L.jpeg_std_error.restype = C.c_void_p
    return L


def _create():
    L = _lib()
    err = C.create_string_buffer(1024)
    cinfo = C.create_string_buffer(SIZE)
    C.memmove(cinfo, C.byref(C.c_void_p(L.jpeg_std_error(err))), 8)
    L.jpeg_CreateCompress(cinfo, C.c_int(62), C.c_size_t(SIZE))
    return cinfo, err


def _layout_ok(cinfo):
    """the defaults of jpeg_set_defaults at the offsets used here"""
    gamma, = struct.unpack_from("d", cinfo, OFF_GAMMA)
    prec, ncomp = struct.unpack_from("ii", cinfo, OFF_PREC)
    smooth, dct = struct.unpack_from("ii", cinfo, OFF_SMOOTH)
    jfif = struct.unpack_from("iBBBxHH", cinfo, OFF_JFIF)
    return gamma == 1.0 and prec == 8 and ncomp == 3 and smooth == 0 and dct == 0 and jfif == (1, 1, 1, 0, 1, 1)


@functools.lru_cache(None)
def available():
    """libjpeg found and its struct layout as expected (checked in a child)"""
    if _path() is None:
        return False
    probe = ("import sys; sys.path.insert(0, %r); import libjpeg_ref as r, struct; c, e = r._create(); "
             "struct.pack_into('IIii', c, r.OFF_WIDTH, 8, 8, 3, 2); r._lib().jpeg_set_defaults(c); "
             "print('ok' if r._layout_ok(c) else 'layout'); r._lib().jpeg_destroy_compress(c)") % os.path.dirname(
                 os.path.abspath(__file__))
    r = subprocess.run([sys.executable, "-c", probe], capture_output=True, text=True, timeout=120)
    return r.returncode == 0 and r.stdout.strip() == "ok"


def encode(img, quality, hs, vs, dct=JDCT_ISLOW, smooth=0):
    """the JPEG file libjpeg writes for rgb `img` (H x W x 3 uint8)"""
    assert available()
    L = _lib()
    H, W = img.shape[:2]
    cinfo, err = _create()
    buf, n = C.c_void_p(), C.c_ulong()
    L.jpeg_mem_dest(cinfo, C.byref(buf), C.byref(n))
    struct.pack_into("IIii", cinfo, OFF_WIDTH, W, H, 3, 2)          # JCS_RGB
    L.jpeg_set_defaults(cinfo)
    assert _layout_ok(cinfo)
    L.jpeg_set_quality(cinfo, C.c_int(int(quality)), C.c_int(1))
    comp, = struct.unpack_from("Q", cinfo, OFF_COMP)
    struct.pack_into("ii", (C.c_char * 16).from_address(comp + 8), 0, hs, vs)   # luma h, v sampling
    struct.pack_into("i", cinfo, OFF_OPTIMIZE, 1)
    struct.pack_into("ii", cinfo, OFF_SMOOTH, smooth, dct)
    struct.pack_into("B", cinfo, OFF_JFIF + 6, 1)                     # density in dots per inch
    struct.pack_into("HH", cinfo, OFF_JFIF + 8, 300, 300)
    L.jpeg_start_compress(cinfo, C.c_int(1))
    a = np.ascontiguousarray(img, np.uint8)
    rows = (C.c_void_p * H)(*[a.ctypes.data + y * W * 3 for y in range(H)])
    L.jpeg_write_scanlines(cinfo, rows, C.c_uint(H))
    L.jpeg_finish_compress(cinfo)
    out = C.string_at(buf, n.value)
    L.jpeg_destroy_compress(cinfo)
    C.CDLL(None).free(buf)
    return out
