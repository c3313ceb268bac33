"""Lossless JPEG raw data decoded on the GPU (vkdt_hip_lj92_decode): tiles of
several images with different components, predictors and tile shapes in one
launch equal the host decoder (i-raw's path) and the original mosaics bit
for bit -- streams with one table and predictor 1 take the many-lanes
decode (1..4 components, scan lines of one or two tile lines, whole-image
strips), the others one lane each -- corrupt entropy data is refused by
both decoders, and a graph reading a compressed DNG develops to the same
output as the uncompressed one."""
import os
import time

import numpy as np
import pytest

from vkdt_amd import hip as H, lj92, pfm, pipe, synth

pytestmark = pytest.mark.gpu


def encode_tiles(cfa, tw, th, comp, pred, bits=14, fold=1):
    """row-major tile streams; fold > 1 frames each tile as th/fold lines of
    fold tile lines (the scan fills the tile row-major either way)"""
    Hh, W = cfa.shape
    ntx, nty = -(-W // tw), -(-Hh // th)
    pad = np.pad(cfa, ((0, nty * th - Hh), (0, ntx * tw - W)), mode="edge")
    return [lj92.encode(pad[ty * th:(ty + 1) * th, tx * tw:(tx + 1) * tw].reshape(th // fold, tw * fold),
                        bits, pred, comp)
            for ty in range(nty) for tx in range(ntx)]


SPECS = [  # kind, W, H, tile w, tile h, components, predictor, tile lines per scan line
    ("bayer", 1000, 700, 256, 256, 2, 1, 1), ("xtrans", 990, 660, 240, 128, 1, 6, 1),
    ("bayer", 640, 480, 640, 16, 4, 7, 1), ("bayer", 512, 384, 132, 96, 3, 4, 1),
    ("bayer", 700, 500, 128, 128, 1, 1, 1), ("bayer", 520, 390, 132, 64, 3, 1, 1),
    ("bayer", 610, 300, 256, 64, 4, 1, 1), ("bayer", 600, 400, 128, 128, 2, 1, 2),
    ("xtrans", 606, 402, 606, 402, 2, 1, 1)]


def test_many_images_one_launch(hip):
    blob, tiles, images, bufs, cfas = b"", [], [], [], []
    for i, (kind, W, Hh, tw, th, comp, pred, fold) in enumerate(SPECS):
        cfa, _ = synth.mosaic(W, Hh, 20 + i, kind)
        streams = encode_tiles(cfa, tw, th, comp, pred, fold=fold)
        b, t = H.lj92_tiles(streams, tw, th, W, Hh, image=i)
        for x in t:
            x.offset += len(blob)
        blob += b
        tiles += t
        buf = hip.DeviceBuffer(W, Hh, 1, H.FMT_UI16)
        bufs.append(buf)
        images.append((buf.ptr, W, Hh, W))
        cfas.append(cfa)
    print(f"[gpu_lj92] {len(tiles)} tiles of {len(SPECS)} images, {len(blob)} bytes", flush=True)
    H.lj92_decode(blob, tiles, images)
    for i, (buf, cfa) in enumerate(zip(bufs, cfas)):
        got = buf.download()
        assert np.array_equal(got, cfa), i
        # the host decoder on the same tiles
        host = np.zeros_like(cfa)
        for t in tiles:
            if t.image == i:
                part = H.lj92_decode_host(blob, t, cfa.shape[1], cfa.shape[0])
                m = np.zeros_like(cfa, bool)
                m[t.y0:t.y0 + t.tile_ht, t.x0:t.x0 + t.tile_wd] = True
                host[m] = part[m]
        assert np.array_equal(host, cfa), i
        buf.free()


@pytest.mark.parametrize("pred", [1, 6])
def test_corrupt_entropy_data_refused(hip, pred):
    """entropy data of all ones (stuffed 0xff bytes): the tables reserve that
    code, so the first sample has no valid code on either path"""
    W, Hh = 64, 32
    cfa, _ = synth.mosaic(W, Hh, 5, "bayer")
    good = lj92.encode(cfa, 14, pred, 2)
    scan = good.index(b"\xff\xda")
    scan += 2 + int.from_bytes(good[scan + 2:scan + 4], "big")
    bad = good[:scan] + b"\xff\x00" * ((len(good) - scan - 2) // 2) + b"\xff\xd9"
    blob, tiles = H.lj92_tiles([bad], W, Hh, W, Hh)
    with pytest.raises(RuntimeError):
        H.lj92_decode_host(blob, tiles[0], W, Hh)
    buf = hip.DeviceBuffer(W, Hh, 1, H.FMT_UI16)
    with pytest.raises(RuntimeError):
        H.lj92_decode(blob, tiles, [(buf.ptr, W, Hh, W)])
    # and the good stream still decodes on both
    blob, tiles = H.lj92_tiles([good], W, Hh, W, Hh)
    H.lj92_decode(blob, tiles, [(buf.ptr, W, Hh, W)])
    assert np.array_equal(buf.download(), cfa)
    buf.free()


def test_throughput_24mp_tiles(hip):
    W, Hh, tw, th = 6000, 4000, 256, 256
    cfa, _ = synth.mosaic(1500, 1000, 30, "bayer")
    cfa = np.tile(cfa, (4, 4))                        # 24 MP without generating a 24 MP scene
    streams = encode_tiles(cfa, tw, th, 2, 1)
    per = [H.lj92_tiles(streams, tw, th, W, Hh)]
    out = hip.DeviceBuffer(W, Hh, 1, H.FMT_UI16)
    blob, tiles = per[0]
    H.lj92_decode(blob, tiles, [(out.ptr, W, Hh, W)])
    assert np.array_equal(out.download(), cfa)
    n = 16   # a batch: 16 images' tiles in one call
    big, all_tiles = b"", []
    for i in range(n):
        for t in tiles:
            all_tiles.append(H.Lj92Tile(t.offset + len(big), t.length, t.x0, t.y0, t.tile_wd, t.tile_ht, 0))
        big += blob
    t0 = time.perf_counter()
    H.lj92_decode(big, all_tiles, [(out.ptr, W, Hh, W)])
    dt = time.perf_counter() - t0
    t1 = time.perf_counter()
    H.lj92_decode(blob, tiles, [(out.ptr, W, Hh, W)])
    d1 = time.perf_counter() - t1
    print(f"[gpu_lj92] 24 MP, {len(tiles)} tiles of 256x256, {len(blob) / 48e6:.2f} of raw size: one image "
          f"{d1 * 1e3:.1f} ms ({24 / d1:.0f} Mpix/s); {n} images in one call {dt * 1e3:.1f} ms "
          f"({24 * n / dt:.0f} Mpix/s, including the {len(big) / 1e6:.0f} MB upload)", flush=True)
    out.free()


@pytest.mark.parametrize("pred,comp", [(1, 2), (6, 1)])
def test_window(hip, pred, comp):
    """a window of the tiles' image (x0, y0 and extent) lands at the out
    pointer, nothing outside it is written"""
    W, Hh, tw, th = 300, 200, 64, 48
    cfa, _ = synth.mosaic(W, Hh, 8, "bayer")
    blob, tiles = H.lj92_tiles(encode_tiles(cfa, tw, th, comp, pred), tw, th, W, Hh)
    x0, y0, w, h = 3, 5, W - 7, Hh - 9
    buf = hip.DeviceBuffer(w + 8, h + 2, 1, H.FMT_UI16)
    buf.upload(np.full((h + 2, w + 8), 7, np.uint16))
    H.lj92_decode(blob, tiles, [(buf.ptr, w, h, w + 8, x0, y0)])
    got = buf.download()
    assert np.array_equal(got[:h, :w], cfa[y0:y0 + h, x0:x0 + w])
    assert (got[:h, w:] == 7).all() and (got[h:] == 7).all()
    buf.free()


@pytest.mark.parametrize("kind,W,Hh,pattern", [("bayer", 320, 240, None), ("bayer", 331, 247, ((2, 1), (1, 0))),
                                               ("xtrans", 330, 246, None)], ids=["rggb", "bggr-ragged", "xtrans"])
def test_graph_from_compressed_dng(tmp_path, hip, kind, W, Hh, pattern):
    """i-raw decodes a compressed DNG on the GPU (read_source_hip) into the
    CFA-aligned window; the developed output equals the uncompressed DNG's"""
    cfa, _ = synth.mosaic(W, Hh, 31, kind)
    outs = []
    for comp in (None, dict(tile=(128, 96), components=2, predictor=1, bits=14)):
        dng = str(tmp_path / f"{'c' if comp else 'u'}.dng")
        synth.write_dng(dng, cfa, kind, compress=comp, pattern=pattern)
        na, nb = synth.noise_params(2e-4, 1e-6)
        g = pipe.Graph(lines=pipe.raw_graph_cfg(dng, noise_a=na, noise_b=nb))
        out = str(tmp_path / f"o{len(outs)}")
        g.export("o-pfm", out)
        g.run()
        outs.append(pfm.read_pfm(out + ".pfm")[0])
        g.close()
    assert np.array_equal(outs[0], outs[1])
