"""rawhist/nprof on the GPU (noise profiling, src/pipe/modules/rawhist and
nprof): the FP64 accumulators of collect against the float oracle (same
per-block float values; counts equal up to the rare block whose bin index
sits on a log2 rounding boundary, moments to 1e-9 in the bins whose counts
agree), the plot of map, and the noise-profile graph end to end: its
written profile equals the oracle's fit, and i-raw then reads it back."""
import os

import numpy as np
import pytest

import graphspec as gs
import oracle_py as op
from vkdt_amd import hip as H, pipe, synth

pytestmark = pytest.mark.gpu

BAYER = synth.BAYER_FILTERS


def noisy(W, Hh, kind, seed):
    cfa, _ = synth.mosaic(W, Hh, seed, kind, a=4e-4, b=2e-5)
    return cfa


def params(filters, black=512, white=16383, noise=(0.0, 0.0)):
    p = np.zeros(5, np.uint32)
    p[0:2] = np.array(noise, np.float32).view(np.uint32)
    p[2:5] = [filters, black, white]
    return p


def gpu_collect(hip, raw, filters, node_wd, node_ht, bins=1000, runs=1):
    Hh, W = raw.shape
    src = hip.DeviceBuffer(W, Hh, 1, H.FMT_UI16)
    src.upload(raw)
    acc = hip.DeviceBuffer(bins, 5, 1, H.FMT_ATOM)
    outs = []
    for _ in range(runs):
        acc.clear()
        hip.launch("rawhist", "collect", None, params(filters), [src, acc], node_wd, node_ht)
        hip.sync()
        outs.append(acc.download())
    src.free()
    acc.free()
    return outs


@pytest.mark.parametrize("kind,W,Hh", [("bayer", 1024, 768), ("xtrans", 1026, 768), ("bayer", 131, 99),
                                       ("xtrans", 133, 100)])
def test_collect_vs_oracle32(hip, kind, W, Hh):
    raw = noisy(W, Hh, kind, 51)
    filters = 9 if kind == "xtrans" else BAYER
    blk = 3 if kind == "xtrans" else 2
    print(f"[gpu_rawhist] collect {kind} {W}x{Hh}", flush=True)
    g1, g2 = gpu_collect(hip, raw, filters, W // blk, Hh // blk, runs=2)
    ref = op.rawhist_collect(32, raw, filters, 512, 16383)
    # the second run starts from a cleared buffer: counts identical
    assert np.array_equal(g1[0], g2[0])
    assert g1[0].sum() == ref[0].sum()
    moved = np.abs(g1[0] - ref[0]).sum() / 2
    same = g1[0] == ref[0]
    rel1 = np.abs(g1[1] - ref[1])[same].max() / max(ref[1].max(), 1e-30)
    rel2 = np.abs(g1[2] - ref[2])[same].max() / max(ref[2].max(), 1e-30)
    print(f"[gpu_rawhist] {kind}: {int(ref[0].sum())} blocks, {moved:.0f} moved bins, "
          f"moment rel diff {rel1:.2e} {rel2:.2e}", flush=True)
    assert moved <= max(2, 1e-4 * ref[0].sum())
    assert rel1 < 1e-9 and rel2 < 1e-9
    assert not g1[3:].any()


def test_map_vs_oracle32(hip):
    raw = noisy(512, 384, "bayer", 52)
    acc = op.rawhist_collect(64, raw, BAYER, 512, 16383)
    h = hip.DeviceBuffer(1000, 5, 1, H.FMT_ATOM)
    h.upload(acc)
    out = hip.DeviceBuffer(1000, 600, 4, H.FMT_F16)
    noise = (1e-5, 2e-4)
    hip.launch("rawhist", "map", None, params(BAYER, noise=noise), [h, out], 1000, 600)
    hip.sync()
    gpu = out.download().astype(np.float64)
    ref = np.zeros((600, 1000, 4))
    op.node(32, "rawhist", "map", None, params(BAYER, noise=noise),
            [op.orbuf(acc, op.OR_F32, 1), op.orbuf(ref, op.OR_F16, 4)], 1000, 600)
    bad = np.mean(np.abs(gpu - ref) > 1e-3)
    print(f"[gpu_rawhist] map: frac differing {bad:.2e}, lit {np.mean(ref[..., 0] > 0):.3f}", flush=True)
    assert np.mean(ref[..., 0] > 0) > 0.01
    assert bad < 1e-3
    h.free()
    out.free()


def test_noise_profile_graph_writes_the_oracle_fit_and_i_raw_reads_it(tmp_path, hip, monkeypatch):
    monkeypatch.chdir(tmp_path)
    W, Hh = 1024, 768
    raw = noisy(W, Hh, "bayer", 53)
    dng = str(tmp_path / "np.dng")
    synth.write_dng(dng, raw, "bayer")
    g = pipe.Graph(lines=pipe.noise_profile_cfg(dng))
    g.run()
    names = {(n["module"], n["kernel"]) for n in g.nodes() if not n["culled"]}
    assert ("rawhist", "collect") in names and ("rawhist", "map") not in names
    g.close()
    fname = "vkdt-synth-synthetic-0.nprof"
    a, b = (float(v) for v in open(tmp_path / fname).read().split())
    ref = op.rawhist_collect(32, raw, BAYER, 512, 16383)
    ra, rb, n = op.nprof_fit(64, ref)
    print(f"[gpu_rawhist] profile a {a:g} b {b:g}; oracle {ra:g} {rb:g} ({n} bins)", flush=True)
    assert np.isclose(a, ra, rtol=1e-4) and np.isclose(b, rb, rtol=1e-4)
    # i-raw picks the profile up when its noise parameters are zero
    os.makedirs(tmp_path / "nprof")
    os.replace(tmp_path / fname, tmp_path / "nprof" / fname)
    g = pipe.Graph(lines=pipe.raw_graph_cfg(dng))
    g.export("o-pfm", str(tmp_path / "out"))
    g.plan()
    na, nb = g.img_param("i-raw", "main")["noise"]
    assert np.isclose(na, a, rtol=1e-6) and np.isclose(nb, b, rtol=1e-6)
    g.close()
