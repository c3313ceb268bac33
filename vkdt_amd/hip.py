"""ctypes binding of the C ABI in include/vkdt_hip.h (libvkdt_hip.so).

This is the same boundary the host runtime (libvkdt_pipe.so) links against;
Python uses it for node-level tests and the benchmark.  Loading fails loudly
when the library has not been built.
"""
import ctypes as C
import os

import numpy as np

from . import lib_path

OK, INCOMPLETE, OOM, DEVICE = 0, -1, -2, -3
FMT_UI16, FMT_F16, FMT_F32, FMT_ATOM, FMT_UI32, FMT_UI8 = 1, 2, 3, 4, 5, 6   # atom: FP64 accumulators; ui32: counters


class Buf(C.Structure):
    """vkdt_buf_t"""
    _fields_ = [
        ("ptr", C.c_void_p),
        ("wd", C.c_uint32), ("ht", C.c_uint32),
        ("pitch", C.c_uint32), ("chan", C.c_uint32),
        ("format", C.c_uint32), ("array_length", C.c_uint32),
        ("plane_stride", C.c_uint64),
    ]


class Lj92Tile(C.Structure):
    """vkdt_lj92_tile_t"""
    _fields_ = [("offset", C.c_uint64), ("length", C.c_uint64), ("x0", C.c_uint32), ("y0", C.c_uint32),
                ("tile_wd", C.c_uint32), ("tile_ht", C.c_uint32), ("image", C.c_uint32)]


class Lj92Image(C.Structure):
    """vkdt_lj92_image_t"""
    _fields_ = [("out", C.c_void_p), ("wd", C.c_uint32), ("ht", C.c_uint32), ("pitch", C.c_uint32),
                ("x0", C.c_uint32), ("y0", C.c_uint32)]


def token(s):
    """dt_token(): up to 8 chars packed little endian into a uint64."""
    b = s.encode()[:8]
    return int.from_bytes(b.ljust(8, b"\0"), "little")


class HipError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = lib_path("libvkdt_hip.so")
        if not os.path.exists(path):
            raise ImportError(f"libvkdt_hip.so not built ({path}); run __graft_entry__.build()")
        L = C.CDLL(path)
        vp, sz, u32, u64, i32 = C.c_void_p, C.c_size_t, C.c_uint32, C.c_uint64, C.c_int
        L.vkdt_hip_abi_version.restype = i32
        L.vkdt_hip_init.argtypes = [i32]
        L.vkdt_hip_node_launch.argtypes = [u64, u64, vp, sz, vp, sz, C.POINTER(Buf), i32, u32, u32, u32, vp]
        L.vkdt_hip_node_supported.argtypes = [u64, u64]
        L.vkdt_hip_alloc.argtypes = [C.POINTER(vp), sz]
        L.vkdt_hip_free.argtypes = [vp]
        L.vkdt_hip_host_alloc.argtypes = [C.POINTER(vp), sz]
        L.vkdt_hip_host_free.argtypes = [vp]
        L.vkdt_hip_upload.argtypes = [vp, vp, sz, vp]
        L.vkdt_hip_download.argtypes = [vp, vp, sz, vp]
        L.vkdt_hip_buf_clear.argtypes = [C.POINTER(Buf), vp]
        L.vkdt_hip_stream_create.argtypes = [C.POINTER(vp)]
        L.vkdt_hip_stream_destroy.argtypes = [vp]
        L.vkdt_hip_stream_sync.argtypes = [vp]
        L.vkdt_hip_event_create.argtypes = [C.POINTER(vp)]
        L.vkdt_hip_event_destroy.argtypes = [vp]
        L.vkdt_hip_event_record.argtypes = [vp, vp]
        L.vkdt_hip_event_elapsed_ms.argtypes = [C.POINTER(C.c_float), vp, vp]
        L.vkdt_hip_event_query.argtypes = [vp]
        L.vkdt_lj92_info.argtypes = [vp, sz] + [C.POINTER(C.c_uint32)] * 4
        L.vkdt_lj92_decode_host.argtypes = [vp, C.POINTER(Lj92Tile), vp, C.c_uint32, C.c_uint32, C.c_uint32]
        L.vkdt_hip_lj92_decode.argtypes = [vp, sz, C.POINTER(Lj92Tile), C.c_int, C.POINTER(Lj92Image), C.c_int, vp]
        L.vkdt_hip_unpack_msb16.argtypes = [vp, sz, vp, C.c_uint32, C.c_int, vp]
        L.vkdt_hip_jpg_encode.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_int,
                                          C.POINTER(C.c_void_p), C.POINTER(sz), vp]
        L.vkdt_hip_jpg_encode_ex.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_int,
                                             C.c_int, C.c_int, vp, sz, C.POINTER(C.c_void_p), C.POINTER(sz), vp]
        L.vkdt_jpg_free.argtypes = [vp]
        L.vkdt_hip_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != OK:
        msg = lib().vkdt_hip_last_error().decode(errors="replace")
        raise HipError(f"{what} failed ({rc}): {msg}")
    return rc


def init(dev=0):
    check(lib().vkdt_hip_init(dev), "vkdt_hip_init")


class DeviceBuffer:
    """A device allocation with the texel layout of one connector."""

    def __init__(self, wd, ht, chan, fmt, planes=1):
        self.wd, self.ht, self.chan, self.fmt, self.planes = wd, ht, chan, fmt, planes
        self.esize = {FMT_ATOM: 8, FMT_F32: 4, FMT_UI32: 4, FMT_UI8: 1}.get(fmt, 2) * chan
        self.plane_bytes = wd * ht * self.esize
        self.nbytes = self.plane_bytes * planes
        p = C.c_void_p()
        check(lib().vkdt_hip_alloc(C.byref(p), self.nbytes), "alloc")
        self.ptr = p.value

    def buf(self):
        return Buf(self.ptr, self.wd, self.ht, self.wd, self.chan, self.fmt, self.planes, self.plane_bytes)

    def dtype(self):
        return {FMT_UI16: np.uint16, FMT_F16: np.float16, FMT_F32: np.float32, FMT_ATOM: np.float64,
                FMT_UI32: np.uint32, FMT_UI8: np.uint8}[self.fmt]

    def shape(self):
        s = (self.planes, self.ht, self.wd) if self.planes > 1 else (self.ht, self.wd)
        return s + ((self.chan,) if self.chan > 1 else ())

    def clear(self, stream=None):
        b = self.buf()
        check(lib().vkdt_hip_buf_clear(C.byref(b), stream), "clear")
        sync(stream)

    def upload(self, arr, stream=None):
        a = np.ascontiguousarray(arr, dtype=self.dtype())
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        check(lib().vkdt_hip_upload(self.ptr, a.ctypes.data, a.nbytes, stream), "upload")
        sync(stream)

    def download(self, stream=None):
        a = np.empty(self.shape(), dtype=self.dtype())
        check(lib().vkdt_hip_download(a.ctypes.data, self.ptr, a.nbytes, stream), "download")
        sync(stream)
        return a

    def free(self):
        if self.ptr:
            check(lib().vkdt_hip_free(self.ptr), "free")
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def sync(stream=None):
    check(lib().vkdt_hip_stream_sync(stream), "sync")


def launch(module, kernel, pc, params, bufs, wd, ht, dp=1, stream=None):
    """vkdt_hip_node_launch with numpy blobs and DeviceBuffer/Buf connectors."""
    conn = (Buf * max(1, len(bufs)))()
    for i, b in enumerate(bufs):
        conn[i] = b.buf() if isinstance(b, DeviceBuffer) else (b if b is not None else Buf())
    pcb = np.ascontiguousarray(pc) if pc is not None else None
    pab = np.ascontiguousarray(params) if params is not None else None
    rc = lib().vkdt_hip_node_launch(
        token(module), token(kernel),
        pcb.ctypes.data if pcb is not None else None, pcb.nbytes if pcb is not None else 0,
        pab.ctypes.data if pab is not None else None, pab.nbytes if pab is not None else 0,
        conn, len(bufs), wd, ht, dp, stream)
    check(rc, f"{module}/{kernel}")


class Event:
    def __init__(self):
        p = C.c_void_p()
        check(lib().vkdt_hip_event_create(C.byref(p)), "event")
        self.ptr = p.value

    def record(self, stream=None):
        check(lib().vkdt_hip_event_record(self.ptr, stream), "record")

    def done(self):
        """True once the work recorded before this event has finished"""
        rc = lib().vkdt_hip_event_query(self.ptr)
        if rc < 0:
            check(rc, "event query")
        return rc == 1

    def elapsed_ms(self, end):
        ms = C.c_float()
        check(lib().vkdt_hip_event_elapsed_ms(C.byref(ms), self.ptr, end.ptr), "elapsed")
        return ms.value


class Stream:
    def __init__(self):
        p = C.c_void_p()
        check(lib().vkdt_hip_stream_create(C.byref(p)), "stream")
        self.ptr = p.value

    def sync(self):
        sync(self.ptr)


def lj92_tiles(streams, tw, th, W, H, image=0):
    """(bytes, [Lj92Tile]) for row-major tile streams covering a W x H image"""
    ntx = -(-W // tw)
    tiles, blob, off = [], bytearray(), 0
    for k, s in enumerate(streams):
        tiles.append(Lj92Tile(off, len(s), (k % ntx) * tw, (k // ntx) * th, tw, th, image))
        blob += s
        off += len(s)
    return bytes(blob), tiles


def lj92_decode_host(blob, tile, W, H):
    """one stream on the host: uint16 H x W (zero where the tile does not reach)"""
    out = np.zeros((H, W), np.uint16)
    buf = (C.c_uint8 * len(blob)).from_buffer_copy(blob)
    check(lib().vkdt_lj92_decode_host(buf, C.byref(tile), out.ctypes.data, W, H, W), "lj92 host decode")
    return out


def unpack_msb16(packed, out, n, bits, stream=None):
    """bit-packed samples (bytes, Magic Lantern layout) into a uint16 device buffer"""
    buf = (C.c_uint8 * max(1, len(packed))).from_buffer_copy(packed or b"\0")
    check(lib().vkdt_hip_unpack_msb16(buf, len(packed), out, n, bits, stream), "unpack")


JPG_DCT_ISLOW, JPG_DCT_IFAST, JPG_DCT_FLOAT = 0, 1, 2


def jpg_encode(rgba_dev, wd, ht, pitch, quality=95, hs=1, vs=1, stream=None, dct=None, smooth=0, icc=None):
    """baseline JPEG (bytes) of an 8-bit rgba device image; dct (JPG_DCT_*),
    smooth (0..100) and icc (bytes) select vkdt_hip_jpg_encode_ex"""
    p, n = C.c_void_p(), C.c_size_t()
    if dct is None and not smooth and icc is None:
        rc = lib().vkdt_hip_jpg_encode(rgba_dev, wd, ht, pitch, quality, hs, vs, C.byref(p), C.byref(n), stream)
    else:
        ib = C.create_string_buffer(icc, len(icc)) if icc else None
        rc = lib().vkdt_hip_jpg_encode_ex(rgba_dev, wd, ht, pitch, quality, hs, vs, dct or 0, smooth, ib,
                                          len(icc) if icc else 0, C.byref(p), C.byref(n), stream)
    check(rc, "jpg encode")
    out = C.string_at(p, n.value)
    lib().vkdt_jpg_free(p)
    return out


def lj92_decode(blob, tiles, images, stream=None):
    """tiles of several images on the GPU; images: [(device ptr, W, H, pitch[, x0, y0])], each a
    window of W x H sites from (x0, y0) on of its tiles' image"""
    buf = (C.c_uint8 * len(blob)).from_buffer_copy(blob)
    t = (Lj92Tile * len(tiles))(*tiles)
    im = (Lj92Image * len(images))(*[Lj92Image(*(tuple(i) + (0, 0))[:6]) for i in images])
    check(lib().vkdt_hip_lj92_decode(buf, len(blob), t, len(tiles), im, len(images), stream), "lj92 decode")
