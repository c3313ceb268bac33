"""The half-size and scaled-export paths on the GPU (demosaic/main.c:92-110,
144-152, 195-203): every node against the float oracle on the GPU's own
inputs (at most 4 binary16 ulps outside near ties), end to end against the
oracle running the same node list (its float build, perturbed twins and
FP64 build, tests/parity.py), and the product graph exported with a size
limit (vkdt-cli --width/--height) against that oracle chain."""
import numpy as np
import pytest

import graphspec as gs
import parity
from parity import run_gpu_graph
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
    parity.check_nodes(bufs, nodes, host, log=lambda m: print(m, flush=True))
    parity.e2e_check_nodes(bufs, nodes, cfa, host["final"], log=lambda m: print(m, flush=True), label=" lod")


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
    print(f"[gpu_lod] export 100 px wide: {ow}x{oh}", flush=True)
    parity.e2e_check_nodes(bufs, nodes, cfa, img, log=lambda m: print(m, flush=True), label=" export")
    g.close()
