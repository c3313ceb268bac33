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
