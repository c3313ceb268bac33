"""HIP node kernels against the oracle, through the C ABI (vkdt_hip_node_launch).

Every node of the raw-development graph runs on the GPU in graph order.
Each node's output is compared with the oracle's float build run on the
*same* inputs (the GPU's own input buffers), which isolates per-node error:
outside the near-tie texels the oracle flags, at most
parity.MAX_NODE_ULP binary16 ulps.  The final rgba is compared with the
oracle's whole-graph run under the end-to-end bound of tests/parity.py.
"""
import pytest

import graphspec as gs
import parity
from vkdt_amd import synth

pytestmark = pytest.mark.gpu

# (kind, W, H, seed, hilite, denoise strength): hilite inserts hilite between
# denoise and demosaic (bin/default-darkroom.i-raw:17-18); strength 0 (the
# param's default) makes denoise one noop node
CASES = [("bayer", 256, 192, 7, False, 0.5), ("xtrans", 258, 198, 11, False, 0.5),
         ("bayer", 256, 192, 17, True, 0.5), ("xtrans", 258, 198, 19, True, 0.5),
         ("bayer", 256, 192, 23, False, 0.0), ("xtrans", 258, 198, 29, False, 0.0)]
IDS = [c[0] + ("-hilite" if c[4] else "") + ("-noop" if c[5] <= 0 else "") for c in CASES]


@pytest.fixture(scope="module", params=CASES, ids=IDS)
def graph_run(request, hip):
    kind, W, H, seed, hilite, strength = request.param
    cfa, _ = synth.mosaic(W, H, seed, kind)
    na, nb = synth.noise_params(2e-4, 1e-6)
    wb = (1.8, 1.0, 1.4, 1.0) if hilite else (1.0, 1.0, 1.0, 1.0)
    p = gs.DevelopParams(kind=kind, noise_a=na, noise_b=nb, hilite=hilite, wb=wb, strength=strength)
    bufs, nodes = gs.build(W, H, p)
    print(f"[gpu_nodes] {kind} {W}x{H} seed {seed} hilite {hilite}: {len(nodes)} nodes", flush=True)
    host = parity.run_gpu_graph(hip, bufs, nodes, cfa)
    return kind, W, H, cfa, p, bufs, nodes, host


def test_nodes_vs_oracle32(graph_run):
    kind, W, H, cfa, p, bufs, nodes, host = graph_run
    parity.check_nodes(bufs, nodes, host, log=lambda s: print(s, flush=True))


def test_end_to_end(graph_run):
    kind, W, H, cfa, p, bufs, nodes, host = graph_run
    parity.e2e_check(cfa, p, host["final"], log=lambda s: print(s, flush=True), label=f" {kind}")


@pytest.mark.parametrize("kind", ["bayer", "xtrans"])
def test_demosaic_down_node(hip, kind):
    """demosaic/down (down.comp:18-47), the block-mean luminance the module
    wires but gauss never samples (so the graph culls it): run on its own
    through the C ABI against the oracle"""
    import numpy as np
    import oracle_py as op
    blk = 3 if kind == "xtrans" else 2
    W, H = 96 * blk // 2, 66
    W -= W % blk
    rng = np.random.default_rng(31)
    mos = rng.uniform(0, 1, size=(H, W)).astype(np.float16)
    p = gs.DevelopParams(kind=kind)
    pc = gs.mixed_blob([("f", 1.0)] * 4 + [("i", p.filters)])
    bufs = {"raw": (W, H, 1, gs.F16, 1), "out": (W // blk, H // blk, 1, gs.F16, 1)}
    nodes = [dict(module="demosaic", kernel="down", pc=pc, params=gs.mixed_blob([("i", 0), ("i", 0)]),
                  conn=["raw", "out"], wd=W // blk, ht=H // blk, dp=1)]
    host = parity.run_gpu_graph(hip, bufs, nodes, mos)
    rep = parity.check_nodes(bufs, nodes, host, log=lambda s: print(s, flush=True))
    assert rep[0][2] <= 1
