"""The o-jpg encoder on the GPU (vkdt_hip_jpg_encode, _ex) against the
oracle's encoder (oracle/jpeg.c, itself byte-identical to libjpeg): the
whole file is byte-identical for 4:4:4, 4:2:2 and 4:2:0, at qualities from
20 to 100, with the islow, ifast and float DCTs and input smoothing, for
smooth, noisy, flat and tiny images with extents that end inside an MCU
(edge replication, dummy blocks), and for an image read through a row pitch
larger than its width; ICC profiles land in APP2 markers after the JFIF
header.  Graph exports through colenc and o-jpg write the oracle's file
with the settings and profile o-jpg derives.  A 24 MP encode is timed."""
import time

import numpy as np
import pytest

from vkdt_amd import hip as H
from test_oracle_jpg import images, oracle_jpg

pytestmark = pytest.mark.gpu


def gpu_jpg(hip, img, q, hs, vs, pad=0, **ex):
    Hh, W = img.shape[:2]
    rgba = np.zeros((Hh, W + pad, 4), np.uint8)
    rgba[:, :W, :3] = img
    rgba[..., 3] = 255
    buf = hip.DeviceBuffer(W + pad, Hh, 4, H.FMT_UI8)
    buf.upload(rgba)
    out = H.jpg_encode(buf.ptr, W, Hh, W + pad, q, hs, vs, **ex)
    buf.free()
    return out


def icc_profiles():
    """the profiles o-jpg embeds, from vkdt_amd/modules/o-jpg/icc.h"""
    import os
    import re
    txt = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "vkdt_amd", "modules",
                            "o-jpg", "icc.h")).read()
    out = {}
    for name in ("ojpg_icc_adobe", "ojpg_icc_rec2020"):
        body = re.search(r"%s\[(\d+)\] = \{(.*?)\};" % name, txt, re.S)
        data = bytes(int(v, 16) for v in re.findall(r"0x([0-9a-f]{2})", body.group(2)))
        assert len(data) == int(body.group(1))
        out[name] = data
    return out


def with_icc(jpg, icc):
    """`jpg` with `icc` in APP2 markers after its JFIF header (ICC.1 B.4)"""
    assert jpg[2:4] == b"\xff\xe0"
    head = 4 + (jpg[4] << 8 | jpg[5])
    chunks = [icc[i:i + 65519] for i in range(0, len(icc), 65519)]
    app2 = b"".join(b"\xff\xe2" + (16 + len(c)).to_bytes(2, "big") + b"ICC_PROFILE\0" + bytes([k + 1, len(chunks)]) + c
                    for k, c in enumerate(chunks))
    return jpg[:head] + app2 + jpg[head:]


# o-jpg/main.c:142-146 (dct, smoothing by quality) and libjpeg's other
# combinations: (quality, hs, vs, dct, smoothing)
SETTINGS = [(97, 1, 1, 2, 0), (100, 1, 1, 2, 0), (75, 2, 2, 0, 20), (55, 2, 2, 0, 40), (45, 2, 2, 1, 40),
            (30, 2, 2, 1, 60), (70, 1, 1, 0, 20), (45, 2, 1, 1, 40), (85, 1, 1, 2, 30)]


@pytest.mark.parametrize("q,hs,vs,dct,smooth", SETTINGS)
def test_dct_methods_and_smoothing_match_oracle(hip, q, hs, vs, dct, smooth):
    for name, img in images():
        print(f"[gpu_jpg] q{q} {hs}x{vs} dct {dct} smoothing {smooth} {name}", flush=True)
        a = gpu_jpg(hip, img, q, hs, vs, dct=dct, smooth=smooth)
        assert a == oracle_jpg(img, q, hs, vs, dct, smooth), name


def test_icc_markers(hip):
    """one profile in one marker, and a 140000-byte one in three (65519
    bytes a marker); the rest of the file is the file without them"""
    img = next(images())[1]
    plain = gpu_jpg(hip, img, 95, 1, 1)
    big = bytes(np.random.default_rng(3).integers(0, 256, 140000, dtype=np.uint8))
    for icc in list(icc_profiles().values()) + [big]:
        out = gpu_jpg(hip, img, 95, 1, 1, dct=0, icc=icc)
        assert out == with_icc(plain, icc)


@pytest.mark.parametrize("q,hs,vs", [(95, 1, 1), (91, 2, 1), (85, 2, 2), (50, 2, 2), (100, 1, 1), (75, 2, 2)])
def test_matches_oracle(hip, q, hs, vs):
    for name, img in images():
        print(f"[gpu_jpg] q{q} {hs}x{vs} {name}", flush=True)
        a, b = gpu_jpg(hip, img, q, hs, vs), oracle_jpg(img, q, hs, vs)
        assert a == b, (name, len(a), len(b))


def test_pitch(hip):
    img = next(images())[1]
    assert gpu_jpg(hip, img, 95, 1, 1, pad=37) == oracle_jpg(img, 95, 1, 1)


def test_24mp_timing(hip):
    rng = np.random.default_rng(5)
    W, Hh = 6000, 4000
    yy, xx = np.mgrid[0:Hh, 0:W]
    img = np.dstack([(xx // 7 + yy // 5) % 256, (xx // 3) % 256, (yy // 2) % 256]).astype(np.uint8)
    img = np.clip(img.astype(np.int16) + rng.integers(-6, 7, size=img.shape), 0, 255).astype(np.uint8)
    rgba = np.dstack([img, np.full((Hh, W), 255, np.uint8)])
    buf = hip.DeviceBuffer(W, Hh, 4, H.FMT_UI8)
    buf.upload(rgba)
    H.jpg_encode(buf.ptr, W, Hh, W, 95, 1, 1)          # warm up
    t0 = time.perf_counter()
    n = 3
    for _ in range(n):
        out = H.jpg_encode(buf.ptr, W, Hh, W, 95, 1, 1)
    dt = (time.perf_counter() - t0) / n
    print(f"[gpu_jpg] 24 MP q95 4:4:4: {len(out) / 1e6:.1f} MB in {dt * 1e3:.1f} ms ({24 / dt:.0f} Mpix/s)", flush=True)
    buf.free()


def colenc_out(g):
    for n in g.nodes():
        if n["module"] == "colenc" and not n["culled"]:
            return n["conns"][1]["buf"]
    raise AssertionError("no colenc node")


@pytest.mark.parametrize("quality,hs,vs,dct,smooth,prim,trc,icc",
                         [(95, 1, 1, 0, 0, 1, 2, None), (91, 2, 1, 0, 0, 1, 2, None), (85, 2, 2, 0, 0, 1, 2, None),
                          (97, 1, 1, 2, 0, 3, 6, "ojpg_icc_adobe"), (45, 2, 2, 1, 40, 2, 1, "ojpg_icc_rec2020"),
                          (70, 2, 2, 0, 20, 1, 2, None)],
                         ids=["q95", "q91", "q85", "q97-adobe-gamma", "q45-rec2020-709", "q70"])
def test_graph_export(tmp_path, hip, quality, hs, vs, dct, smooth, prim, trc, icc):
    """raw > ... > llap > colenc (8 bit) > o-jpg: the written file is the
    oracle's encoding of colenc's 8-bit output with the sampling, DCT and
    smoothing the quality selects (o-jpg/main.c:139-146), with the ICC
    profile the encoding selects (main.c:152-156); that output is the 8-bit
    rounding of colenc's float output; PIL reads the file back"""
    from vkdt_amd import pipe, synth
    W, Hh = 320, 213
    cfa, _ = synth.mosaic(W, Hh, 12, "bayer")
    dng = str(tmp_path / "a.dng")
    synth.write_dng(dng, cfa, "bayer")
    na, nb = synth.noise_params(2e-4, 1e-6)
    outs = {}
    for fmt in ("o-jpg", "o-pfm"):
        g = pipe.Graph(lines=pipe.raw_graph_cfg(dng, noise_a=na, noise_b=nb))
        g.export(fmt, str(tmp_path / "o"), prim=prim, trc=trc)
        if fmt == "o-jpg":
            g.set_param("o-jpg", "main", "quality", quality)
        g.run()
        b = colenc_out(g)
        dt = np.uint8 if fmt == "o-jpg" else np.float32
        a = np.empty((b.ht, b.wd, 4), dt)
        H.check(H.lib().vkdt_hip_download(a.ctypes.data, b.ptr, a.nbytes, g.stream()), "download")
        H.sync(g.stream())
        outs[fmt] = a
        g.close()
    ui8, f32 = outs["o-jpg"], outs["o-pfm"]
    want = np.rint(np.clip(np.nan_to_num(f32), 0, 1) * 255).astype(np.uint8)
    assert np.array_equal(ui8, want)
    data = open(str(tmp_path / "o") + ".jpg", "rb").read()
    want = oracle_jpg(np.ascontiguousarray(ui8[..., :3]), quality, hs, vs, dct, smooth)
    assert data == (with_icc(want, icc_profiles()[icc]) if icc else want)
    from PIL import Image
    im = np.asarray(Image.open(str(tmp_path / "o") + ".jpg").convert("RGB")).astype(np.float64)
    mse = ((im - ui8[..., :3]) ** 2).mean()
    psnr = 10 * np.log10(255.0 ** 2 / max(mse, 1e-12))
    print(f"[gpu_jpg] graph export q{quality}: {len(data)} bytes, PSNR {psnr:.1f} dB vs the 8-bit image", flush=True)
    # chroma subsampling of saturated synthetic edges costs dBs, and so do
    # the coarse quantisers and the smoothing below quality 80
    assert psnr > (35 if hs * vs == 1 else 28) - (8 if quality < 80 else 0)
