"""hist and autoexp on the GPU (src/pipe/modules/hist, autoexp) against the
float oracle: the u32 waveform counts of hist/collect equal bit for bit
(out-of-range and NaN values included), hist/map within one binary16 step
(pow), autoexp/hist's 256 bins equal up to the rare sample whose log2 sits
on a bin boundary, autoexp/lum to 1e-6 relative, autoexp/exp bit for bit.
Through a graph, autoexp keeps its luminance across frames in the protected
buffer: frame 1 moves from frame 0's value as the oracle says."""
import numpy as np
import pytest

import oracle_py as op
from vkdt_amd import hip as H, pipe, synth
from test_oracle_hist import ae_hist, collect, lum_node

pytestmark = pytest.mark.gpu


def f16img(W, Hh, seed, lo=-0.2, hi=1.2, log=False):
    rng = np.random.default_rng(seed)
    img = np.exp2(rng.uniform(-17, 4, size=(Hh, W, 4))) if log else rng.uniform(lo, hi, size=(Hh, W, 4))
    img = img.astype(np.float16)
    img.reshape(-1, 4)[::997, 0] = np.nan
    img.reshape(-1, 4)[::991, 1] = np.inf
    img.reshape(-1, 4)[::983, 2] = -np.inf
    return img


def dev(hip, arr, fmt, chan):
    Hh, W = arr.shape[:2]
    b = hip.DeviceBuffer(W, Hh, chan, fmt)
    b.upload(arr)
    return b


@pytest.mark.parametrize("W,Hh", [(1920, 1080), (333, 77)])
def test_hist_collect_and_map(hip, W, Hh):
    img = f16img(W, Hh, 4)
    src = dev(hip, img, H.FMT_F16, 4)
    cnt = hip.DeviceBuffer(1000, 600, 1, H.FMT_UI32)
    cnt.clear()
    hip.launch("hist", "collect", None, None, [src, cnt], W, Hh)
    hip.sync()
    got = cnt.download()
    ref = collect(img.astype(np.float64), bits=32).astype(np.uint64).astype(np.uint32)
    print(f"[gpu_hist] collect {W}x{Hh}: {int((got != ref).sum())} bins differ", flush=True)
    assert np.array_equal(got, ref)
    out = hip.DeviceBuffer(1000, 600, 4, H.FMT_F16)
    hip.launch("hist", "map", None, None, [cnt, out], 1000, 600)
    hip.sync()
    o = out.download().astype(np.float64)
    want = np.zeros((600, 1000, 4))
    op.node(32, "hist", "map", None, None, [op.orbuf(ref.astype(np.float64), op.OR_F32, 1),
                                            op.orbuf(want, op.OR_F16, 4)], 1000, 600)
    d = np.abs(o - want)
    if d.max() > 2.0 ** -11 * max(1.0, np.abs(want).max()):
        y, x, ch = np.unravel_index(np.argmax(d), d.shape)
        print(f"[gpu_hist] map differs at {y},{x},{ch}: gpu {o[y, x]} oracle {want[y, x]} count {ref[y, x]:#x} "
              f"(gpu count {got[y, x]:#x}); {int((d > 1e-3).sum())} values", flush=True)
    assert d.max() <= 2.0 ** -11 * max(1.0, np.abs(want).max())
    for b in (src, cnt, out):
        b.free()


@pytest.mark.parametrize("W,Hh", [(1920, 1080), (250, 130)])
def test_autoexp_nodes(hip, W, Hh):
    img = f16img(W, Hh, 5, log=True)
    src = dev(hip, img, H.FMT_F16, 4)
    hist = hip.DeviceBuffer(17, 16, 1, H.FMT_UI32)
    hist.clear()
    hip.launch("autoexp", "hist", None, None, [src, hist], (W + 15) // 16 * 8, (Hh + 15) // 16 * 8)
    hip.sync()
    got = hist.download().ravel().astype(np.float64)
    ref = ae_hist(img.astype(np.float64), bits=32)
    moved = np.abs(got[:256] - ref[:256]).sum() / 2
    print(f"[gpu_hist] autoexp/hist {W}x{Hh}: {int(ref[256])} samples, {moved:.0f} moved", flush=True)
    assert got[256] == ref[256] and moved <= 2 and not got[257:].any()
    # lum from the GPU histogram: frame 0, then a later frame from a history
    par = np.zeros(5, np.uint32)
    par[0:3] = np.array([3.0, 5.0, 0.016], np.float32).view(np.uint32)
    lum = hip.DeviceBuffer(8, 8, 1, H.FMT_F32)
    for frame, prev in ((0, 123.0), (4, 0.05), (4, 1e-6)):
        lum.upload(np.full((8, 8), prev, np.float32))
        par[4] = frame
        hip.launch("autoexp", "lum", None, par, [src, hist, lum], 1, 1)
        hip.sync()
        l = float(lum.download()[0, 0])
        want = lum_node(got, prev, frame, bits=32)
        assert abs(l - want) <= 1e-6 * abs(want), (frame, prev, l, want)
    # exp with the last luminance
    out = hip.DeviceBuffer(W, Hh, 4, H.FMT_F16)
    hip.launch("autoexp", "exp", None, None, [src, hist, lum, out], W, Hh)
    hip.sync()
    o = out.download().astype(np.float64)
    want = np.zeros((Hh, W, 4))
    op.node(32, "autoexp", "exp", None, None,
            [op.orbuf(img.astype(np.float64), op.OR_F16, 4), op.orbuf(np.zeros((16, 17)), op.OR_F32, 1),
             op.orbuf(lum.download().astype(np.float64), op.OR_F32, 1), op.orbuf(want, op.OR_F16, 4)], W, Hh)
    fin = np.isfinite(want)
    assert np.array_equal(np.isnan(o), np.isnan(want)) and np.array_equal(o[fin], want[fin])
    for b in (src, hist, lum, out):
        b.free()


def test_autoexp_across_frames(tmp_path, hip):
    """two frames of a clip, the second two stops brighter: autoexp's
    luminance after frame 1 is frame 0's moved by 1 - exp(-timediff vup)
    towards frame 1's mean (the protected buffer holds frame 0's value)"""
    W, Hh = 256, 192
    base = synth.mosaic(W, Hh, 9, "bayer", clip_frac=0.01)[0].astype(np.float64)
    frames = [np.clip(base * s, 0, 15000).astype(np.uint16) for s in (0.25, 1.0)]
    path = str(tmp_path / "clip.mlv")
    synth.write_mlv(path, frames, bits=14)
    lines = [l for l in pipe.mlv_graph_cfg(path) if l != "connect:llap:01:output:display:main:input"]
    lines += ["module:autoexp:01", "connect:llap:01:output:autoexp:01:input",
              "connect:autoexp:01:output:display:main:input"]
    g = pipe.Graph(lines=lines)
    g.export("o-pfm", str(tmp_path / "o"))
    lums, hists = [], []
    for f in range(2):
        g.set_frame(f)
        g.run()
        nodes = {n["kernel"]: n for n in g.nodes() if n["module"] == "autoexp" and not n["culled"]}
        hb, lb = nodes["hist"]["conns"][1]["buf"], nodes["lum"]["conns"][2]["buf"]
        h = np.empty(272, np.uint32)
        l = np.empty(1, np.float32)
        H.check(H.lib().vkdt_hip_download(h.ctypes.data, hb.ptr, h.nbytes, g.stream()), "download")
        H.check(H.lib().vkdt_hip_download(l.ctypes.data, lb.ptr, 4, g.stream()), "download")
        H.sync(g.stream())
        hists.append(h.astype(np.float64))
        lums.append(float(l[0]))
    g.close()
    want0 = lum_node(hists[0], np.nan, 0, bits=32)
    want1 = lum_node(hists[1], lums[0], 1, bits=32)
    print(f"[gpu_hist] autoexp over frames: {lums} want {want0:.6g} {want1:.6g}", flush=True)
    assert abs(lums[0] - want0) <= 1e-6 * want0
    assert abs(lums[1] - want1) <= 1e-6 * want1
    assert lums[1] > lums[0]
