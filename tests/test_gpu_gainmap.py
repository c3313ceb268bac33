"""DNG gain maps on the GPU (noop.comp:45-55, doub.comp:106-114): denoise's
noop and doub with the four-site gain map bound, per node against the float
oracle on the GPU's own inputs (at most 4 binary16 ulps outside near ties,
as every node), and the product graph from a DNG whose OpcodeList2 holds
the maps (i-raw decode, the gain-map source upload) against the oracle
running the same node list (float build, perturbed twins, FP64 build)."""
import numpy as np
import pytest

import graphspec as gs
import oracle_py as op
import parity
from vkdt_amd import pfm, pipe, synth

pytestmark = pytest.mark.gpu


def maps(mh, mw, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(0.7, 1.4, size=(mh, mw)).astype(np.float32) for _ in range(4)]


# (strength, map rows, map columns, origin (v, h), spacing (v, h))
CASES = [(0.5, 5, 7, (0.0, 0.0), None), (0.0, 5, 7, (0.0, 0.0), None),
         (0.5, 4, 3, (0.05, 0.1), (0.4, 0.6))]
IDS = ["wavelet", "noop", "wavelet-offset-grid"]


@pytest.mark.parametrize("strength,mh,mw,origin,spacing", CASES, ids=IDS)
def test_nodes_with_gain_map(hip, strength, mh, mw, origin, spacing):
    W, H = 192, 144
    cfa, _ = synth.mosaic(W, H, 71, "bayer")
    na, nb = synth.noise_params(2e-4, 1e-6)
    spacing = spacing or (1.0 / (mh - 1), 1.0 / (mw - 1))
    gmd = dict(gains=np.stack(maps(mh, mw, 5), -1), origin=origin, spacing=spacing)
    p = gs.DevelopParams(kind="bayer", noise_a=na, noise_b=nb, strength=strength, gain_map=gmd)
    bufs, nodes = gs.build(W, H, p)
    print(f"[gpu_gainmap] strength {strength}, map {mh}x{mw}, origin {origin}, spacing {spacing}", flush=True)
    host = parity.run_gpu_graph(hip, bufs, nodes, cfa, sources={"gm": gmd["gains"]})
    parity.check_nodes(bufs, nodes, host, log=lambda s: print(s, flush=True))
    # the map moved the mosaic: against the same graph without it
    p0 = gs.DevelopParams(kind="bayer", noise_a=na, noise_b=nb, strength=strength)
    b0, n0 = gs.build(W, H, p0)
    h0 = parity.run_gpu_graph(hip, b0, n0, cfa)
    ratio = host["dn"].astype(np.float64)[8:-8, 8:-8] / np.maximum(h0["dn"].astype(np.float64)[8:-8, 8:-8], 1e-3)
    print(f"[gpu_gainmap] gain applied: ratio {ratio.min():.3f}..{ratio.max():.3f}", flush=True)
    assert ratio.min() < 0.9 and ratio.max() > 1.1


def test_product_graph_with_opcode_list(tmp_path, hip):
    W, H = 256, 192
    cfa, _ = synth.mosaic(W, H, 72, "bayer")
    m = maps(6, 8, 9)
    dng = str(tmp_path / "gm.dng")
    lst = synth.opcode_list(synth.bayer_gain_maps(W, H, m))
    synth.write_dng(dng, cfa, "bayer", extra_tags=((51009, 7, list(lst)),))
    na, nb = synth.noise_params(2e-4, 1e-6)
    g = pipe.Graph(lines=pipe.raw_graph_cfg(dng, noise_a=na, noise_b=nb))
    out = str(tmp_path / "out")
    g.export("o-pfm", out)
    g.run()
    img, _ = pfm.read_pfm(out + ".pfm")
    assert any(n["kernel"] == "gainmap" for n in g.nodes())
    ip = g.img_param("i-raw", "main")
    gmd = dict(gains=np.stack(m, -1), origin=(0.0, 0.0), spacing=(1.0 / 5, 1.0 / 7))
    p = gs.DevelopParams(kind="bayer", noise_a=np.float32(na), noise_b=np.float32(nb),
                         cam=np.array(ip["cam"], np.float32).reshape(3, 3), wb=tuple(ip["wb"][:3]) + (1.0,),
                         gain_map=gmd)
    bufs, nodes = gs.build(W, H, p)
    parity.e2e_check_nodes(bufs, nodes, cfa, img, sources={"gm": gmd["gains"]}, log=lambda m: print(m, flush=True),
                           label=" gain map")
    p0 = gs.DevelopParams(kind="bayer", noise_a=np.float32(na), noise_b=np.float32(nb),
                          cam=np.array(ip["cam"], np.float32).reshape(3, 3), wb=tuple(ip["wb"][:3]) + (1.0,))
    b0, n0 = gs.build(W, H, p0)
    ref0 = op.run_nodes(64, b0, n0, cfa)["final"][..., :3]
    psnr0 = 10 * np.log10(1.0 / max(float(np.mean((img.astype(np.float64) - ref0) ** 2)), 1e-30))
    print(f"[gpu_gainmap] the product graph against the oracle without the maps: psnr {psnr0:.1f} dB", flush=True)
    assert psnr0 < 45.0
    g.close()
