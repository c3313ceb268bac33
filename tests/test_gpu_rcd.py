"""RCD demosaic (method 1) on the GPU: rcd_conv's three outputs and
rcd_fill's tiles against the float oracle on the GPU's own inputs, and the
graph end to end against the oracle running the same node list (its float
build, perturbed twins and FP64 build, tests/parity.py).

Per node the RCD kernels keep a fraction bound (at most 0.1% of values
beyond 4 binary16 ulps): rcd_fill picks directions by comparing
discriminators (rcd_fill.comp), and the RCD restatement does not flag its
near ties, so a flipped choice cannot be told from a mistake by the ulp
count alone.  Measured: rcd_conv 0 ulps, rcd_fill 1.7e-4 of its sites."""
import numpy as np
import pytest

import graphspec as gs
import oracle_py as op
import parity
from parity import f16_ulps, run_gpu_graph, to_or
from vkdt_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("W,H,seed,hilite,max_wd", [(256, 192, 81, False, 0), (262, 190, 82, True, 200)],
                         ids=["plain", "hilite-scaled"])
def test_rcd_nodes_and_end_to_end(hip, W, H, seed, hilite, max_wd):
    cfa, _ = synth.mosaic(W, H, seed, "bayer")
    na, nb = synth.noise_params(2e-4, 1e-6)
    p = gs.DevelopParams(kind="bayer", noise_a=na, noise_b=nb, dm_method=1, hilite=hilite,
                         wb=(1.8, 1.0, 1.4, 1.0))
    bufs, nodes = gs.build(W, H, p, max_wd, 0)
    print(f"[gpu_rcd] {W}x{H} seed {seed} hilite {hilite} max_wd {max_wd}: {len(nodes)} nodes", flush=True)
    host = run_gpu_graph(hip, bufs, nodes, cfa)
    for n in nodes:
        if not n["kernel"].startswith("rcd_"):
            continue
        nin = 1 if n["kernel"] == "rcd_conv" else 4
        keep, ins = [], []
        for cn in n["conn"][:nin]:
            a, b = to_or(host[cn], bufs[cn])
            keep.append(a)
            ins.append(b)
        outs = []
        for cn in n["conn"][nin:]:
            o = np.zeros(host[cn].shape, np.float64)
            outs.append((cn, o))
            ins.append(op.orbuf(o, op.OR_F16, bufs[cn][2]))
        op.node(32, n["module"], n["kernel"], n["pc"], n["params"], ins, n["wd"], n["ht"], n["dp"])
        for cn, o in outs:
            d = f16_ulps(host[cn].astype(np.float64), o)
            print(f"  {n['kernel']} -> {cn}: max {d.max():.3g} ulp, frac>4ulp {np.mean(d > 4):.2e}", flush=True)
            assert np.mean(d > 4) <= 1e-3, (n["kernel"], cn)
    parity.e2e_check_nodes(bufs, nodes, cfa, host["final"], log=lambda m: print(m, flush=True), label=" rcd")
