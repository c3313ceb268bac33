"""The half-size and scaled-export paths on the GPU (demosaic/main.c:92-110,
144-152, 195-203): demosaic/halfsize and shared/resample per node against
the float oracle on the GPU's own inputs, end to end against the FP64
oracle running the same node list, and the product graph exported with a
size limit (vkdt-cli --width/--height) against that oracle chain."""
import numpy as np
import pytest

import graphspec as gs
import oracle_py as op
from parity import f16_ulps, run_gpu_graph, to_or
from vkdt_amd import pfm, pipe, synth

pytestmark = pytest.mark.gpu

# (kind, W, H, seed, demosaic method, export width, export height)
CASES = [("bayer", 256, 192, 61, 0, 64, 0),      # scale 4: half size + resample
         ("bayer", 256, 192, 62, 0, 200, 0),     # scale 1.28: full demosaic + resample
         ("xtrans", 258, 198, 63, 2, 0, 0),      # method 2: half size, resampled to (W+1)/2
         ("bayer", 256, 192, 64, 2, 0, 0)]       # method 2 on the block grid: no resample
IDS = [f"{c[0]}-m{c[4]}-w{c[5]}" for c in CASES]


@pytest.mark.parametrize("kind,W,H,seed,method,max_wd,max_ht", CASES, ids=IDS)
def test_nodes_and_end_to_end(hip, kind, W, H, seed, method, max_wd, max_ht):
    cfa, _ = synth.mosaic(W, H, seed, kind)
    na, nb = synth.noise_params(2e-4, 1e-6)
    p = gs.DevelopParams(kind=kind, noise_a=na, noise_b=nb, dm_method=method)
    bufs, nodes = gs.build(W, H, p, max_wd, max_ht)
    kinds = [n["module"] + "/" + n["kernel"] for n in nodes]
    print(f"[gpu_lod] {kind} {W}x{H} method {method} max {max_wd}x{max_ht}: output {bufs['final'][:2]}, "
          f"{'halfsize' if 'demosaic/halfsize' in kinds else 'full'}"
          f"{' + resample' if 'shared/resample' in kinds else ''}", flush=True)
    host = run_gpu_graph(hip, bufs, nodes, cfa)
    for n in nodes:
        if n["module"] not in ("shared",) and n["kernel"] != "halfsize":
            continue
        keep, ins = [], []
        for c in n["conn"][:-1]:
            a, b = to_or(host[c], bufs[c])
            keep.append(a)
            ins.append(b)
        name = n["conn"][-1]
        spec = bufs[name]
        out = np.zeros(host[name].shape, np.float64)
        op.node(32, n["module"], n["kernel"], n["pc"], n["params"], ins + [op.orbuf(out, op.OR_F16, spec[2])],
                n["wd"], n["ht"], n["dp"])
        d = f16_ulps(host[name].astype(np.float64), out)
        print(f"  {n['module']}/{n['kernel']}: max {d.max():.3g} ulp, frac>4ulp {np.mean(d > 4):.2e}", flush=True)
        assert np.mean(d > 4) <= 1e-3
    ref = op.run_nodes(64, bufs, nodes, cfa)["final"]
    gpu = host["final"].astype(np.float64)
    d = np.abs(gpu[..., :3] - ref[..., :3])
    psnr = 10 * np.log10(1.0 / max(float(np.mean(d ** 2)), 1e-30))
    print(f"[gpu_lod] e2e vs FP64: max abs {d.max():.3e}, frac>1e-3 {np.mean(d > 1e-3):.2e}, psnr {psnr:.1f} dB",
          flush=True)
    assert psnr >= 60.0 and np.mean(d > 1e-3) <= 1e-3


def test_product_graph_export_size(tmp_path, hip):
    W, H = 320, 240
    cfa, _ = synth.mosaic(W, H, 65, "bayer")
    dng = str(tmp_path / "in.dng")
    synth.write_dng(dng, cfa, "bayer")
    na, nb = synth.noise_params(2e-4, 1e-6)
    g = pipe.Graph(lines=pipe.raw_graph_cfg(dng, noise_a=na, noise_b=nb))
    out = str(tmp_path / "out")
    g.export("o-pfm", out, max_wd=100)
    g.run()
    img, _ = pfm.read_pfm(out + ".pfm")
    ow, oh = gs.export_extent(W, H, 100, 0)
    assert img.shape == (oh, ow, 3)
    ip = g.img_param("i-raw", "main")
    p = gs.DevelopParams(kind="bayer", noise_a=np.float32(na), noise_b=np.float32(nb),
                         cam=np.array(ip["cam"], np.float32).reshape(3, 3), wb=tuple(ip["wb"][:3]) + (1.0,))
    bufs, nodes = gs.build(W, H, p, 100, 0)
    ref = op.run_nodes(64, bufs, nodes, cfa)["final"][..., :3]
    d = np.abs(img.astype(np.float64) - ref)
    psnr = 10 * np.log10(1.0 / max(float(np.mean(d ** 2)), 1e-30))
    print(f"[gpu_lod] export 100 px wide: {ow}x{oh}, max abs {d.max():.3e}, psnr {psnr:.1f} dB", flush=True)
    assert psnr >= 60.0 and np.mean(d > 1e-3) <= 1e-3
    g.close()
