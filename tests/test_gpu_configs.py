"""The workloads of the benchmark configurations, on the GPU under checks.

C1  2048x1536 RGGB Bayer (the CPU-runnable case)         per node and end to end
C2  6000x4000 RGGB Bayer (24 MP, the plain-path still)     per node and end to end
C3  6240x4160 X-Trans (6240x4158 processed), llap with 8 curves and 6 levels
C4  the 64-still batch (16x 4000x3000, 24x 6000x4000, 16x 8256x5504,
    8x 9504x6336, a quarter X-Trans) in one batched launch: every image
    equals its own per-node graph bit for bit, is finite, and has alpha 1
C5  serving 4096x2160 12-bit frames from two streams and a 61 MP still:
    every request's result in the mix equals its result served alone

C1-C3 run the per-node graph through the C ABI and hold every node to the
oracle's float build on the GPU's own inputs, and the result to the
oracle's whole-graph run (tests/parity.py).
"""
import ctypes as C
import os

import numpy as np
import pytest

import graphspec as gs
import parity
from vkdt_amd import synth

pytestmark = pytest.mark.gpu

CONFIGS = {
    "C1": dict(kind="bayer", W=2048, H=1536, seed=1),
    "C2": dict(kind="bayer", W=6000, H=4000, seed=2),
    "C3": dict(kind="xtrans", W=6240, H=4160, seed=3, num_gamma=8, max_levels=6),
}


def log(s):
    print(s, flush=True)


@pytest.fixture(scope="module", params=sorted(CONFIGS), ids=sorted(CONFIGS))
def config_run(request, hip):
    c = dict(CONFIGS[request.param])
    kind, W, H, seed = c.pop("kind"), c.pop("W"), c.pop("H"), c.pop("seed")
    blk = 3 if kind == "xtrans" else 2
    W, H = W - W % blk, H - H % blk          # i-raw's block rounding
    cfa, _ = synth.mosaic(W, H, seed, kind)
    na, nb = synth.noise_params(2e-4, 1e-6)
    p = gs.DevelopParams(kind=kind, noise_a=na, noise_b=nb, **c)
    bufs, nodes = gs.build(W, H, p)
    log(f"[config] {request.param}: {kind} {W}x{H}, {len(nodes)} nodes")
    host = parity.run_gpu_graph(hip, bufs, nodes, cfa)
    return request.param, cfa, p, bufs, nodes, host


def test_config_nodes(config_run):
    name, cfa, p, bufs, nodes, host = config_run
    log(f"[config] {name} per node")
    parity.check_nodes(bufs, nodes, host, log=log)


def test_config_end_to_end(config_run):
    name, cfa, p, bufs, nodes, host = config_run
    parity.e2e_check(cfa, p, host["final"], log=log, label=f" {name}")


# ---------------------------------------------------------------------------
# C4: the 64-still batch
# ---------------------------------------------------------------------------

class _DevArray:
    """a device allocation seen through __cuda_array_interface__ (no ownership)"""

    def __init__(self, ptr, shape, typestr):
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": typestr, "data": (ptr, False),
                                         "version": 2, "strides": None}


def devview(ptr, shape, typestr):
    import torch
    return torch.as_tensor(_DevArray(ptr, shape, typestr), device="cuda")


def final_buffer(g):
    """device pointer and extent of llap/colour's rgba f32 output in a run graph"""
    for n in g.nodes():
        if n["module"] == "llap" and n["kernel"] == "colour" and not n["culled"]:
            b = n["conns"][-1]["buf"]
            assert b.pitch == b.wd and b.chan == 4
            return b.ptr, b.wd, b.ht
    raise AssertionError("no llap/colour node")


def test_c4_batch64_equals_per_node_graphs(tmp_path, hip):
    import torch
    import bench
    from vkdt_amd import batch, pipe, synth_gpu
    torch.cuda.init()
    work = bench.batch_workload()
    combos = sorted(set((W, H, k) for W, H, k, _ in work))
    dngs, mos = {}, {}
    for ci, (W, H, k) in enumerate(combos):
        cfa, _ = synth_gpu.mosaic(W, H, 900 + ci, k)
        dngs[(W, H, k)] = str(tmp_path / f"s{ci}.dng")
        synth.write_dng(dngs[(W, H, k)], cfa.cpu().numpy().view(np.uint16), k)
        mos[(W, H, k)] = cfa
    log(f"[c4] {len(combos)} shapes written")

    def graph(i):
        W, H, k, a = work[i]
        na, nb = synth.noise_params(a, 1e-6)
        g = pipe.Graph(lines=pipe.raw_graph_cfg(dngs[(W, H, k)], noise_a=na, noise_b=nb))
        g.export("o-pfm", str(tmp_path / f"o{i}"))
        return g

    descs, outs = [], []
    for i, (W, H, k, a) in enumerate(work):
        g = graph(i)
        g.plan()
        d = batch.describe(g)
        g.close()
        src = mos[(W, H, k)][:d.cfa_ht, :d.cfa_wd].contiguous()   # the mosaic as i-raw stages it
        mos[(W, H, k, i)] = src
        out = torch.empty((d.ht, d.wd, 4), dtype=torch.float32, device="cuda")
        d.cfa, d.out = src.data_ptr(), out.data_ptr()
        descs.append(d)
        outs.append(out)
    nbytes = batch.scratch_bytes(descs)
    scratch = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    P = sum(d.wd * d.ht for d in descs)
    log(f"[c4] 64 images, {P / 1e6:.0f} Mpix, scratch {nbytes / 2**30:.1f} GiB")
    b = batch.Batch(descs, scratch.data_ptr(), nbytes)
    try:
        b.run(None, graph=True)
        hip.sync()
    finally:
        b.close()
    del scratch
    for i in range(len(work)):
        o = outs[i]
        assert bool(torch.isfinite(o).all()), i
        assert bool((o[..., 3] == 1).all()), i
        g = graph(i)
        g.run(pipe.RUN_ALL & ~pipe.RUN_DOWNLOAD)
        hip.sync(g.stream())
        ptr, wd, ht = final_buffer(g)
        assert (ht, wd) == tuple(o.shape[:2])
        ref = devview(ptr, (ht, wd, 4), "<f4")
        same = bool(torch.equal(ref, o))
        g.close()
        if (i + 1) % 16 == 0 or not same:
            log(f"[c4] image {i} {work[i][0]}x{work[i][1]} {work[i][2]}: batch == per-node graph: {same}")
        assert same, i


# ---------------------------------------------------------------------------
# C5: serving 4096x2160 frames and a 61 MP still
# ---------------------------------------------------------------------------

def test_c5_serving_mix_equals_alone(tmp_path, hip):
    from test_gpu_serve import Pinned
    from vkdt_amd import batch, hip as H, pipe, serve, synth_gpu

    def profile(name, W, Hh, a, black, white):
        cfa, _ = synth_gpu.mosaic(W, Hh, 3, "bayer", black=black, white=white)
        dng = str(tmp_path / f"{name}.dng")
        synth.write_dng(dng, cfa.cpu().numpy().view(np.uint16), "bayer", black=black, white=white)
        na, nb = synth.noise_params(a, 1e-6, black, white)
        g = pipe.Graph(lines=pipe.raw_graph_cfg(dng, noise_a=na, noise_b=nb))
        g.export("o-pfm", str(tmp_path / name))
        g.plan()
        d = batch.describe(g)
        g.close()
        return serve.Profile(d, name)

    profs = {"f0": profile("f0", 4096, 2160, 2e-4, 256, 4095), "f1": profile("f1", 4096, 2160, 7e-4, 256, 4095),
             "st": profile("st", 9504, 6336, 3e-4, 512, 16383)}
    mos = {}
    for s, seed in (("f0", 21), ("f1", 22)):
        sc = None
        for f in range(3):    # frames drifting 2 px per frame
            cfa, sc = synth_gpu.mosaic(4096, 2160, seed, "bayer", black=256, white=4095, origin=(2 * f, 0),
                                       canvas=(4096 + 8, 2160), scale=sc)
            mos[(s, f)] = Pinned(cfa.cpu().numpy().view(np.uint16))
    d = profs["st"].desc
    mos[("st", 0)] = Pinned(synth_gpu.mosaic(d.cfa_wd, d.cfa_ht, 23, "bayer")[0].cpu().numpy().view(np.uint16))
    rng = np.random.default_rng(5)
    t, reqs = 0.0, []
    for i in range(12):
        t += rng.exponential(1.0 / 600.0)
        key = ("st", 0) if i in (3, 8) else (("f0", "f1")[i % 2], (i // 2) % 3)
        reqs.append((key, t))
    log(f"[c5] {len(reqs)} requests over {t * 1e3:.1f} ms")

    def serve_all(rs, slots, max_batch):
        outs = {}

        def done(r):
            o = np.empty((r.profile.desc.ht, r.profile.desc.wd, 4), np.float32)
            H.check(H.lib().vkdt_hip_download(o.ctypes.data, r.out, o.nbytes, None), "download")
            outs[r.id] = o

        rq = [serve.Request(i, profs[k[0]], mos[k].ptr, t, tag=k) for i, (k, t) in enumerate(rs)]
        srv = serve.Server(64 << 30, slots=slots, max_batch=max_batch, on_done=done)
        try:
            srv.run(rq)
        finally:
            srv.close()
        return rq, outs, srv

    try:
        rq, outs, _ = serve_all([(k, 0.0) for k in sorted(set(k for k, _ in reqs))], 1, 1)
        alone = {r.tag: outs[r.id] for r in rq}
        rq, outs, srv = serve_all(reqs, 3, 8)
        sizes = [b[0] for b in srv.batches]
        log(f"[c5] batch sizes {sizes}, peak {srv.arena.peak / 2**30:.1f} GiB")
        for r in rq:
            o = outs[r.id]
            assert np.isfinite(o).all() and (o[..., 3] == 1).all(), r.tag
            assert np.array_equal(o, alone[r.tag]), (r.id, r.tag)
        assert max(sizes) > 1
    finally:
        for m in mos.values():
            m.free()
