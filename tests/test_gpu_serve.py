"""The serving loop on the GPU: requests from two raw-video streams and three
stills (one of them with hilite between denoise and demosaic) arrive as a Poisson process; the server batches whatever has arrived
whenever a slot frees, keeps several batches in flight, and admits within
its memory budget.  Every request's result must equal -- bit for bit -- the
same request served alone (a batch of one), which is itself the per-node
graph's result (test_gpu_batch.py)."""
import ctypes as C

import numpy as np
import pytest

from vkdt_amd import batch, hip as H, pipe, serve, synth

pytestmark = pytest.mark.gpu


class Pinned:
    """a mosaic in pinned host memory"""

    def __init__(self, arr):
        self.nbytes = arr.nbytes
        p = C.c_void_p()
        H.check(H.lib().vkdt_hip_host_alloc(C.byref(p), arr.nbytes), "host alloc")
        self.ptr = p.value
        C.memmove(self.ptr, np.ascontiguousarray(arr).ctypes.data, arr.nbytes)

    def free(self):
        H.lib().vkdt_hip_host_free(self.ptr)


def profile(tmp, name, kind, W, Hh, a, hilite=False):
    cfa, _ = synth.mosaic(W, Hh, 3, kind)
    dng = str(tmp / f"{name}.dng")
    synth.write_dng(dng, cfa, kind)
    na, nb = synth.noise_params(a, 1e-6)
    g = pipe.Graph(lines=pipe.raw_graph_cfg(dng, noise_a=na, noise_b=nb, hilite=hilite))
    g.export("o-pfm", str(tmp / name))
    g.plan()
    d = batch.describe(g)
    g.close()
    return serve.Profile(d, name)


@pytest.fixture(scope="module")
def world(tmp_path_factory, hip):
    tmp = tmp_path_factory.mktemp("serve")
    profs = {"s0": profile(tmp, "s0", "bayer", 320, 240, 2e-4),
             "s1": profile(tmp, "s1", "bayer", 320, 240, 6e-4),
             "x": profile(tmp, "x", "xtrans", 330, 246, 4e-4),
             "b": profile(tmp, "b", "bayer", 400, 300, 3e-4),
             "h": profile(tmp, "h", "bayer", 360, 270, 3e-4, hilite=True)}
    mos = {}
    for s, seed in (("s0", 11), ("s1", 12)):   # four frames per stream, drifting 2 px per frame
        sc = None
        for f in range(4):
            cfa, sc = synth.mosaic(320, 240, seed, "bayer", origin=(2 * f, 0), canvas=(400, 240), scale=sc)
            mos[(s, f)] = Pinned(cfa)
    mos[("x", 0)] = Pinned(synth.mosaic(330, 246, 13, "xtrans")[0])
    mos[("b", 0)] = Pinned(synth.mosaic(400, 300, 14, "bayer")[0])
    mos[("h", 0)] = Pinned(synth.mosaic(360, 270, 15, "bayer")[0])
    rng = np.random.default_rng(7)
    t, reqs = 0.0, []
    for i in range(48):
        t += rng.exponential(1.0 / 3000.0)      # Poisson arrivals, 3000 / s
        if rng.random() < 0.2:
            key = (("x", "b", "h")[int(rng.integers(3))], 0)
        else:
            s = "s0" if rng.random() < 0.5 else "s1"
            key = (s, i % 4)
        reqs.append((key, t))
    yield profs, mos, reqs
    for m in mos.values():
        m.free()


def serve_all(profs, mos, reqs, budget, slots, max_batch):
    outs = {}

    def done(r):
        o = np.empty((r.profile.desc.ht, r.profile.desc.wd, 4), np.float32)
        H.check(H.lib().vkdt_hip_download(o.ctypes.data, r.out, o.nbytes, None), "download")
        outs[r.id] = o

    rq = [serve.Request(i, profs[k[0]], mos[k].ptr, t, tag=k) for i, (k, t) in enumerate(reqs)]
    srv = serve.Server(budget, slots=slots, max_batch=max_batch, on_done=done)
    try:
        wall = srv.run(rq)
    finally:
        srv.close()
    return rq, outs, srv, wall


def test_mix_equals_alone_and_batches_overlap(world):
    profs, mos, reqs = world
    budget = 1 << 30
    # alone: one request per batch, one batch in flight, all arrivals at once
    alone_reqs = [(k, 0.0) for k in sorted(set(k for k, _ in reqs))]
    rq, outs, _, _ = serve_all(profs, mos, alone_reqs, budget, slots=1, max_batch=1)
    alone = {r.tag: outs[r.id] for r in rq}
    rq, outs, srv, wall = serve_all(profs, mos, reqs, budget, slots=3, max_batch=8)
    sizes = [b[0] for b in srv.batches]
    overlap = sum(1 for i, a in enumerate(srv.batches) for b in srv.batches[i + 1:] if b[1] < a[2])
    s = serve.summary(rq, wall)
    print(f"[serve] {len(rq)} requests in {len(sizes)} batches (sizes {sizes}), {overlap} overlapping launches, "
          f"{s}", flush=True)
    assert all(r.done is not None for r in rq)
    for r in rq:
        assert np.array_equal(outs[r.id], alone[r.tag]), (r.id, r.tag)
    assert max(sizes) > 1          # requests were batched
    assert overlap > 0             # a batch was admitted while another was on the device
    assert srv.arena.in_use() == 0


def test_admission_keeps_within_the_budget(world):
    profs, mos, reqs = world
    budget = int(2.2 * max(p.bytes for p in profs.values()))
    rq, outs, srv, _ = serve_all(profs, mos, reqs[:16], budget, slots=3, max_batch=8)
    sizes = [b[0] for b in srv.batches]
    print(f"[serve] budget {budget} B: batch sizes {sizes}, peak {srv.arena.peak} B", flush=True)
    assert all(r.done is not None for r in rq) and len(outs) == 16
    assert srv.arena.peak <= budget
    assert max(sizes) <= budget // min(p.bytes for p in profs.values())
