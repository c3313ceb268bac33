"""The batched develop path on the GPU.  A batch of stills of mixed sizes,
CFAs and module parameters runs every stage in one launch over all images;
each image's output must equal -- bit for bit -- its own per-node graph run
(the product drop-in path), whatever it is batched with, in eager and in
hipGraph mode, with one tile per work group and with persistent work groups.
The per-node graph itself is held to the oracle by test_gpu_graph.py; one
image is also compared with the FP64 oracle here."""
import numpy as np
import pytest

import graphspec as gs
import oracle_py as op
from vkdt_amd import batch, hip as H, pfm, pipe, synth

pytestmark = pytest.mark.gpu

# kind, W, H, seed, module overrides
IMAGES = [
    ("bayer", 320, 240, 41, dict()),
    ("xtrans", 330, 246, 42, dict()),
    ("bayer", 258, 194, 43, dict(strength=0.8, clarity=0.5, numgamma=6, levels=5)),
    ("xtrans", 198, 132, 44, dict(extra=["param:demosaic:01:colour:1", "param:llap:01:shadows:0.7"])),
    ("bayer", 64, 48, 45, dict(extra=["param:colour:01:exposure:0.5", "param:filmcurv:01:colour:1"])),
    # hilite between denoise and demosaic (the stock darkroom chain)
    ("bayer", 290, 210, 46, dict(hilite=True)),
    ("xtrans", 246, 180, 47, dict(hilite=True, extra=["param:hilite:01:soft:0.3", "param:hilite:01:desat:0.6"])),
    ("bayer", 64, 48, 48, dict(hilite=True, levels=4)),
]


def plain_and_planned(tmp_path, kind, W, H, seed, kw):
    cfa, _ = synth.mosaic(W, H, seed, kind)
    dng = str(tmp_path / f"in{seed}.dng")
    synth.write_dng(dng, cfa, kind)
    na, nb = synth.noise_params(2e-4 * (1 + seed % 3), 1e-6)
    kw = dict(kw)
    extra = kw.pop("extra", [])
    g = pipe.Graph(lines=pipe.raw_graph_cfg(dng, noise_a=na, noise_b=nb, extra=extra, **kw))
    out = str(tmp_path / f"out{seed}")
    g.export("o-pfm", out)
    g.run()
    img, _ = pfm.read_pfm(out + ".pfm")
    return g, cfa, img


@pytest.fixture(scope="module")
def cases(tmp_path_factory, hip):
    tmp = tmp_path_factory.mktemp("batch")
    res = []
    for kind, W, Hh, seed, kw in IMAGES:
        print(f"[batch] plain {kind} {W}x{Hh} seed {seed} {kw}", flush=True)
        res.append(plain_and_planned(tmp, kind, W, Hh, seed, kw))
    yield res
    for g, _, _ in res:
        g.close()


def check_equal(outs, cases, tag):
    for i, (o, (g, _, ref)) in enumerate(zip(outs, cases)):
        rgb = o[..., :3]
        assert rgb.shape == ref.shape, (tag, i, rgb.shape, ref.shape)
        nd = int(np.count_nonzero(rgb != ref))
        print(f"[batch] {tag} image {i}: {nd} values differ from the plain graph, max abs "
              f"{float(np.abs(rgb.astype(np.float64) - ref).max()):.3e}", flush=True)
        assert nd == 0, (tag, i)
        assert np.all(o[..., 3] == 1.0)


@pytest.mark.parametrize("tune", [0, 2], ids=["tile_per_group", "persistent2"])
def test_batch_equals_plain_graphs(cases, tune):
    graphs = [c[0] for c in cases]
    cfas = [c[1] for c in cases]
    outs = batch.develop(graphs, cfas, tune=[tune] * batch.STAGES)
    check_equal(outs, cases, f"eager tune {tune}")


def test_alone_equals_in_the_mix(cases):
    """each image developed in a batch of one equals its plain run too"""
    for i, c in enumerate(cases):
        outs = batch.develop([c[0]], [c[1]], tune=[1] * batch.STAGES)
        check_equal(outs, [c], f"alone {i}")


def test_graph_replays_match_eager(cases):
    """the captured hipGraph, replayed with eager work and syncs between the
    replays, gives the eager result every time"""
    descs, bufs = [], []
    for g, cfa, _ in cases:
        d = batch.describe(g)
        src = H.DeviceBuffer(d.cfa_wd, d.cfa_ht, 1, H.FMT_UI16)
        src.upload(cfa)
        dst = H.DeviceBuffer(d.wd, d.ht, 4, H.FMT_F32)
        d.cfa, d.out = src.ptr, dst.ptr
        descs.append(d)
        bufs.append((src, dst))
    nb = batch.scratch_bytes(descs)
    scratch = H.DeviceBuffer((nb + 1) // 2, 1, 1, H.FMT_UI16)
    stream = H.Stream()
    b = batch.Batch(descs, scratch.ptr, scratch.nbytes, [2] * batch.STAGES)
    try:
        b.run(stream.ptr, graph=False)
        stream.sync()
        eager = [dst.download(stream.ptr) for _, dst in bufs]
        check_equal(eager, cases, "eager")
        nans = [np.full(dst.nbytes // 4, np.nan, np.float32) for _, dst in bufs]
        for rep in range(3):
            for (_, dst), nan in zip(bufs, nans):   # clobber the outputs with eager work in between
                H.check(H.lib().vkdt_hip_upload(dst.ptr, nan.ctypes.data, dst.nbytes, stream.ptr), "upload")
            stream.sync()
            b.run(stream.ptr, graph=True)
            stream.sync()
            outs = [dst.download(stream.ptr) for _, dst in bufs]
            for o, e in zip(outs, eager):
                assert np.array_equal(o, e), rep
        print("[batch] 3 graph replays equal the eager run", flush=True)
    finally:
        b.close()
        scratch.free()
        for s, d in bufs:
            s.free()
            d.free()


def test_autotune_keeps_results(cases):
    graphs = [c[0] for c in cases]
    descs, bufs = [], []
    for g, cfa, _ in cases:
        d = batch.describe(g)
        src = H.DeviceBuffer(d.cfa_wd, d.cfa_ht, 1, H.FMT_UI16)
        src.upload(cfa)
        dst = H.DeviceBuffer(d.wd, d.ht, 4, H.FMT_F32)
        d.cfa, d.out = src.ptr, dst.ptr
        descs.append(d)
        bufs.append((src, dst))
    nb = batch.scratch_bytes(descs)
    scratch = H.DeviceBuffer((nb + 1) // 2, 1, 1, H.FMT_UI16)
    b = batch.Batch(descs, scratch.ptr, scratch.nbytes)
    try:
        tune = b.autotune()
        print(f"[batch] autotune picks {dict(zip(batch.stage_names(), tune))}", flush=True)
        # launch shape per stage group: persistent work groups per CU | rows per thread << 8
        assert all((t & 0xff) in (0, 1, 2, 4, 8) and (t >> 8) in (1, 2, 4, 8) for t in tune)
        b.set_timing(True)
        b.run()
        ms = b.stage_ms()
        assert all(m >= 0 for m in ms)
        b.set_timing(False)
        b.run(graph=False)
        H.sync()
        check_equal([dst.download() for _, dst in bufs], cases, "autotuned")
    finally:
        b.close()
        scratch.free()
        for s, d in bufs:
            s.free()
            d.free()


def test_batch_vs_oracle64(cases):
    g, cfa, ref_plain = cases[0]
    out = batch.develop([g], [cfa])[0][..., :3]
    ip = g.img_param("i-raw", "main")
    na, nb = synth.noise_params(2e-4 * (1 + 41 % 3), 1e-6)
    p = gs.DevelopParams(kind="bayer", noise_a=np.float32(na), noise_b=np.float32(nb),
                         cam=np.array(ip["cam"], np.float32).reshape(3, 3), wb=tuple(ip["wb"][:3]) + (1.0,))
    ref = op.develop(64, cfa[:out.shape[0], :out.shape[1]], gs.oracle_params(p))[..., :3]
    d = np.abs(out.astype(np.float64) - ref)
    psnr = 10 * np.log10(1.0 / max(float(np.mean(d ** 2)), 1e-30))
    frac = float(np.mean(d > 1e-3))
    print(f"[batch] vs oracle64: max abs {d.max():.3e} frac>1e-3 {frac:.2e} psnr {psnr:.1f} dB", flush=True)
    assert psnr >= 60.0 and frac <= 1e-3
