"""Lossless JPEG (T.81 process 14) raw data, no device work:

  * the stream writer round-trips through the reference's own decoder
    (liblj92 from src/pipe/modules/i-mlv, built into oracle/_ref where the
    reference tree is present) for one component with every predictor and
    for two components with predictor 1 -- liblj92 keeps its previous line
    at an offset of one frame width, not width x components, so it mispredicts
    multi-component lines for predictors 2..7 (its i-mlv streams use 1);
  * the product's host decoder (i-raw's path) inverts streams of 1..4
    components, every predictor and 2..16 bits, and streams the reference's
    encoder wrote (predictor 6);
  * a compressed DNG reads back as the same mosaic as the uncompressed one
    (tiles ragged at the right and bottom edges), and corrupt or restart-coded
    streams are refused.
"""
import ctypes as C
import os

import numpy as np
import pytest

from vkdt_amd import hip, lj92, pipe, synth

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "liblj92.so")
needs_ref = pytest.mark.skipif(not os.path.exists(REF), reason="reference liblj92 not built (no reference tree)")


def ref_lib():
    L = C.CDLL(REF)
    L.lj92_encode.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                              C.POINTER(C.c_void_p), C.POINTER(C.c_int)]
    return L


def ref_decode(stream):
    L = ref_lib()
    h, w, hh, bd, comp = C.c_void_p(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
    data = (C.c_uint8 * len(stream)).from_buffer_copy(stream)
    assert L.lj92_open(C.byref(h), data, len(stream), C.byref(w), C.byref(hh), C.byref(bd), C.byref(comp)) == 0
    out = np.zeros(w.value * hh.value * comp.value, np.uint16)
    rc = L.lj92_decode(h, out.ctypes.data_as(C.POINTER(C.c_uint16)), w.value * comp.value, 0, None, 0)
    L.lj92_close(h)
    assert rc == 0
    return out.reshape(hh.value, w.value * comp.value)


def tile(seed, H=48, W=96, bits=14):
    rng = np.random.default_rng(seed)
    t = rng.normal(0.4, 0.15, size=(H, W)).clip(0, 1) * ((1 << bits) - 1)
    t[5:15, 20:50] = (1 << bits) - 1                                  # flat clipped block
    return t.astype(np.uint16)


def host_decode(stream, W, H):
    blob, tiles = hip.lj92_tiles([stream], W, H, W, H)
    return hip.lj92_decode_host(blob, tiles[0], W, H)


@needs_ref
@pytest.mark.parametrize("comp,pred", [(1, p) for p in range(1, 8)] + [(2, 1)])
def test_writer_round_trips_through_the_reference_decoder(comp, pred):
    t = tile(pred + 10 * comp)
    np.testing.assert_array_equal(ref_decode(lj92.encode(t, 14, pred, comp)), t)


@pytest.mark.parametrize("comp", [1, 2, 3, 4])
@pytest.mark.parametrize("pred", range(1, 8))
def test_host_decoder_inverts_the_writer(comp, pred):
    W = 96 if comp != 3 else 90
    t = tile(100 + pred, W=W)
    np.testing.assert_array_equal(host_decode(lj92.encode(t, 14, pred, comp), W, 48), t)


@pytest.mark.parametrize("bits", [2, 8, 12, 16])
def test_host_decoder_bit_depths(bits):
    t = tile(7, bits=bits)
    if bits == 16:
        t[0, :8] = [0, 65535, 0, 65535, 32768, 32767, 1, 65534]     # differences of +-32768 (category 16)
    np.testing.assert_array_equal(host_decode(lj92.encode(t, bits, 1, 2), 96, 48), t)


@needs_ref
def test_host_decoder_reads_the_reference_encoder():
    t = tile(3)
    enc, n = C.c_void_p(), C.c_int()
    assert ref_lib().lj92_encode(t.ctypes.data, 96, 48, 14, 96, 0, None, 0, C.byref(enc), C.byref(n)) == 0
    stream = C.string_at(enc, n.value)
    np.testing.assert_array_equal(host_decode(stream, 96, 48), t)


def test_corrupt_and_restart_streams_are_refused():
    s = bytearray(lj92.encode(tile(1), 14, 1, 1))
    with pytest.raises(hip.HipError, match="SOI"):
        host_decode(bytes(s[1:]), 96, 48)
    dri = s[:2] + b"\xff\xdd\x00\x04\x00\x10" + s[2:]
    with pytest.raises(hip.HipError, match="restart"):
        host_decode(bytes(dri), 96, 48)
    # a Huffman table whose codes overflow their lengths
    i = s.index(b"\xff\xc4") + 5
    bad = s[:i] + bytes([3] + [0] * 15) + s[i + 16:]
    with pytest.raises(hip.HipError):
        host_decode(bytes(bad), 96, 48)


@pytest.mark.parametrize("kind,W,H,comp", [("bayer", 300, 200, 2), ("xtrans", 330, 246, 1)])
def test_compressed_dng_reads_as_the_uncompressed(tmp_path, kind, W, H, comp):
    cfa, _ = synth.mosaic(W, H, 3, kind)
    a, b = str(tmp_path / "u.dng"), str(tmp_path / "c.dng")
    synth.write_dng(a, cfa, kind)
    synth.write_dng(b, cfa, kind, compress=dict(tile=(128, 96), components=comp, predictor=1, bits=14))
    assert os.path.getsize(b) < os.path.getsize(a)
    ra = pipe.Graph(lines=pipe.noise_profile_cfg(a)).read_source("i-raw", "main", W * H * 2)
    rb = pipe.Graph(lines=pipe.noise_profile_cfg(b)).read_source("i-raw", "main", W * H * 2)
    np.testing.assert_array_equal(ra, rb)
