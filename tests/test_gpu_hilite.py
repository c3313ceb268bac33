"""hilite on the GPU (src/pipe/modules/hilite) beyond the graph cases of
test_gpu_nodes.py:

  * the module wired straight behind the raw input: UNORM16 mosaic in and
    out, then demosaic reading the UNORM16 mosaic -- per node against the
    float oracle on the GPU's own inputs (UNORM16 outputs within one count);
  * ragged mosaics (extent not a multiple of the block): doub's trailing
    sites come from the threads past the node extent;
  * the product graph (cfg -> module DSOs -> nodes -> HIP) with hilite
    between denoise and demosaic against the oracle's whole graph (float
    build, perturbed twins, FP64 build: tests/parity.py).
"""
import numpy as np
import pytest

import graphspec as gs
import oracle_py as op
import parity
from vkdt_amd import pfm, pipe, synth

pytestmark = pytest.mark.gpu

ORF = {gs.UI16: op.OR_UI16, gs.F16: op.OR_F16, gs.F32: op.OR_F32}


def hilite_nodes(W, H, kind, fmt, wb=(1.8, 1.0, 1.4, 1.0), demosaic=False):
    p = gs.DevelopParams(kind=kind, wb=wb)
    bufs, nodes = {"raw": (W, H, 1, gs.UI16, 1)}, []
    buf = lambda n, w, h, c, f, pl=1: bufs.__setitem__(n, (w, h, c, f, pl))
    add = lambda m, k, pc, par, conn, w, h, dp=1: nodes.append(
        dict(module=m, kernel=k, pc=pc, params=par, conn=conn, wd=w, ht=h, dp=dp))
    gs.add_hilite(W, H, p, buf, add, "raw", "hl", fmt_out=fmt)
    if demosaic:
        blk = 3 if kind == "xtrans" else 2
        dmpc = gs.mixed_blob([("f", v) for v in wb] + [("i", p.filters)])
        dmpar = gs.mixed_blob([("i", 0), ("i", 0)])
        buf("cov", W // blk, H // blk, 4, gs.F16)
        add("demosaic", "gauss", dmpc, dmpar, ["hl", "hl", "cov"], W // blk, H // blk)
        buf("green", W, H, 1, gs.F16)
        add("demosaic", "splat", dmpc, dmpar, ["hl", "cov", "green"], W, H)
        buf("dm", W, H, 4, gs.F16)
        add("demosaic", "fix", dmpc, dmpar, ["hl", "green", "cov", "dm"], W, H)
    return bufs, nodes


def run_gpu(hip, bufs, nodes, raw):
    dev = {k: hip.DeviceBuffer(*v) for k, v in bufs.items()}
    dev["raw"].upload(raw)
    for n in nodes:
        hip.launch(n["module"], n["kernel"], n["pc"], n["params"], [dev[c] for c in n["conn"]],
                   n["wd"], n["ht"], n["dp"])
    hip.sync()
    host = {k: b.download() for k, b in dev.items()}
    for b in dev.values():
        b.free()
    return host


def as_or(arr, spec):
    wd, ht, chan, fmt, planes = spec
    a = np.ascontiguousarray(arr, np.uint16 if fmt == gs.UI16 else np.float64)
    return a, op.orbuf(a, ORF[fmt], chan, planes)


def node_diffs(bufs, nodes, host):
    """per node: the GPU output against the float oracle on the GPU's inputs"""
    rep = []
    for n in nodes:
        keep, ins = [], []
        for c in n["conn"][:-1]:
            a, b = as_or(host[c], bufs[c])
            keep.append(a)
            ins.append(b)
        name = n["conn"][-1]
        spec = bufs[name]
        # doub writes only the sites its 8x8 groups reach: start from the
        # GPU's buffer so that sites neither side writes compare equal
        out, ob = as_or(host[name], spec)
        op.node(32, n["module"], n["kernel"], n["pc"], n["params"], ins + [ob], n["wd"], n["ht"], n["dp"])
        gpu = host[name]
        if spec[3] == gs.UI16:
            d = np.abs(gpu.astype(np.int64) - out.astype(np.int64)).astype(np.float64)
            bad = float(np.mean(d > 1))
        else:
            g64 = gpu.astype(np.float64)
            d = np.abs(g64 - out) / np.maximum(np.abs(out), 2.0 ** -14) * 1024
            d[np.isnan(g64) & np.isnan(out)] = 0
            d[np.isnan(d)] = np.inf
            bad = float(np.mean(d > 4))
        rep.append((n["module"], n["kernel"], name, float(d.max()), bad))
        print(f"  {n['module']}/{n['kernel']} -> {name}: max {d.max():.3g}, frac bad {bad:.2e}", flush=True)
    return rep


def clipped_raw(W, H, kind, seed):
    """a mosaic whose brightest 4% sit at the UNORM16 ceiling region used as
    the clipping point (white param 0.985 of full scale)"""
    cfa, _ = synth.mosaic(W, H, seed, kind, black=0, white=65535, clip_frac=0.04)
    return cfa


@pytest.mark.parametrize("kind,W,H", [("bayer", 256, 192), ("xtrans", 258, 198)])
def test_unorm16_chain_and_demosaic_vs_oracle32(hip, kind, W, H):
    raw = clipped_raw(W, H, kind, 31)
    print(f"[gpu_hilite] ui16 {kind} {W}x{H}: clipped {(raw >= 0.985 * 65535).mean():.3f}", flush=True)
    bufs, nodes = hilite_nodes(W, H, kind, gs.UI16, demosaic=True)
    host = run_gpu(hip, bufs, nodes, raw)
    assert np.mean(host["hl"] != raw) > 0.01      # inpainting changed the clipped sites
    rep = node_diffs(bufs, nodes, host)
    worst = [r for r in rep if r[4] > 1e-3]
    assert not worst, worst


# extents whose block count is not a multiple of 8 in either axis, so the
# trailing sites are written by threads past the node extent
@pytest.mark.parametrize("kind,W,H", [("bayer", 131, 99), ("xtrans", 133, 100)])
def test_ragged_mosaic_trailing_sites(hip, kind, W, H):
    raw = clipped_raw(W, H, kind, 32)
    print(f"[gpu_hilite] ragged {kind} {W}x{H}", flush=True)
    for fmt in (gs.UI16, gs.F16):
        bufs, nodes = hilite_nodes(W, H, kind, fmt)
        host = run_gpu(hip, bufs, nodes, raw)
        rep = node_diffs(bufs, nodes, host)
        assert not [r for r in rep if r[4] > 1e-3], rep
    # on these extents the oracle's doub reaches every site, the trailing
    # column and row too, so the comparison above covered all of them
    n = nodes[-1]
    ins = [as_or(host[c], bufs[c]) for c in n["conn"][:-1]]
    sentinel = np.full((H, W), -7.0)
    op.node(32, n["module"], n["kernel"], n["pc"], n["params"],
            [b for _, b in ins] + [op.orbuf(sentinel, op.OR_F16, 1)], n["wd"], n["ht"], n["dp"])
    assert not (sentinel == -7.0).any()


@pytest.mark.parametrize("kind,W,H,seed", [("bayer", 320, 240, 41), ("xtrans", 330, 246, 42)])
def test_product_graph_with_hilite_vs_oracle(tmp_path, hip, kind, W, H, seed):
    cfa, _ = synth.mosaic(W, H, seed, kind, clip_frac=0.04)
    dng = str(tmp_path / "in.dng")
    synth.write_dng(dng, cfa, kind, as_shot_neutral=(0.5, 1.0, 0.7))
    na, nb = synth.noise_params(2e-4, 1e-6)
    print(f"[gpu_hilite] graph {kind} {W}x{H} seed {seed}", flush=True)
    g = pipe.Graph(lines=pipe.raw_graph_cfg(dng, noise_a=na, noise_b=nb, hilite=True))
    out = str(tmp_path / "out")
    g.export("o-pfm", out)
    g.run()
    assert any(n["module"] == "hilite" for n in g.nodes())
    img, _ = pfm.read_pfm(out + ".pfm")
    ip = g.img_param("i-raw", "main")
    p = gs.DevelopParams(kind=kind, noise_a=np.float32(na), noise_b=np.float32(nb), hilite=True,
                         cam=np.array(ip["cam"], np.float32).reshape(3, 3), wb=tuple(ip["wb"][:3]) + (1.0,))
    mos = cfa[:img.shape[0], :img.shape[1]]
    parity.e2e_check(mos, p, img, log=lambda m: print(m, flush=True), label=f" hilite {kind}")
    ref = op.develop(32, mos, gs.oracle_params(p))[..., :3]
    p.hilite = False
    ref0 = op.develop(32, mos, gs.oracle_params(p))[..., :3]
    moved = float(np.mean(np.abs(ref - ref0) > 1e-3))
    print(f"[gpu_hilite] {kind}: hilite changes {moved:.3f} of the output", flush=True)
    assert moved > 0.001
    g.close()
