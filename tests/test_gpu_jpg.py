"""The o-jpg encoder on the GPU (vkdt_hip_jpg_encode) against the oracle's
encoder (oracle/jpeg.c, itself byte-identical to libjpeg's segments): the
whole file is byte-identical for 4:4:4, 4:2:2 and 4:2:0, at qualities from
50 to 100, for smooth, noisy, flat and tiny images with extents that end
inside an MCU (edge replication, dummy blocks), and for an image read
through a row pitch larger than its width.  A 24 MP encode is timed."""
import time

import numpy as np
import pytest

from vkdt_amd import hip as H
from test_oracle_jpg import images, oracle_jpg

pytestmark = pytest.mark.gpu


def gpu_jpg(hip, img, q, hs, vs, pad=0):
    Hh, W = img.shape[:2]
    rgba = np.zeros((Hh, W + pad, 4), np.uint8)
    rgba[:, :W, :3] = img
    rgba[..., 3] = 255
    buf = hip.DeviceBuffer(W + pad, Hh, 4, H.FMT_UI8)
    buf.upload(rgba)
    out = H.jpg_encode(buf.ptr, W, Hh, W + pad, q, hs, vs)
    buf.free()
    return out


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


@pytest.mark.parametrize("quality,hs,vs", [(95, 1, 1), (91, 2, 1), (85, 2, 2)])
def test_graph_export(tmp_path, hip, quality, hs, vs):
    """raw > ... > llap > colenc (sRGB, 8 bit) > o-jpg: the written file is the
    oracle's encoding of colenc's 8-bit output with the sampling the quality
    selects (o-jpg/main.c:139-141); that output is the 8-bit rounding of
    colenc's float output; PIL reads the file back"""
    from vkdt_amd import pipe, synth
    W, Hh = 320, 213
    cfa, _ = synth.mosaic(W, Hh, 12, "bayer")
    dng = str(tmp_path / "a.dng")
    synth.write_dng(dng, cfa, "bayer")
    na, nb = synth.noise_params(2e-4, 1e-6)
    outs = {}
    for fmt in ("o-jpg", "o-pfm"):
        g = pipe.Graph(lines=pipe.raw_graph_cfg(dng, noise_a=na, noise_b=nb))
        g.export(fmt, str(tmp_path / "o"), prim=1, trc=2)
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
    assert data == oracle_jpg(np.ascontiguousarray(ui8[..., :3]), quality, hs, vs)
    from PIL import Image
    im = np.asarray(Image.open(str(tmp_path / "o") + ".jpg").convert("RGB")).astype(np.float64)
    mse = ((im - ui8[..., :3]) ** 2).mean()
    psnr = 10 * np.log10(255.0 ** 2 / max(mse, 1e-12))
    print(f"[gpu_jpg] graph export q{quality}: {len(data)} bytes, PSNR {psnr:.1f} dB vs the 8-bit image", flush=True)
    assert psnr > (35 if hs * vs == 1 else 28)   # chroma subsampling of saturated synthetic edges costs dBs
