"""vkdt_hip_batch_develop_static, the batch entry point that allocates
nothing: its intermediates live in the library's static device scratch.
Its results must equal the caller-scratch batch (vkdt_hip_batch_create +
vkdt_hip_batch_run) bit for bit, on two calls in a row with the second
smaller than the first, and images that need more scratch than the static
array holds are refused."""
import ctypes as C

import numpy as np
import pytest

from vkdt_amd import batch, hip as H, pipe, synth

pytestmark = pytest.mark.gpu

STATIC_BYTES = 1 << 30


def planned(tmp, name, kind, W, Hh, seed, a):
    cfa, _ = synth.mosaic(W, Hh, seed, kind)
    dng = str(tmp / f"{name}.dng")
    synth.write_dng(dng, cfa, kind)
    na, nb = synth.noise_params(a, 1e-6)
    g = pipe.Graph(lines=pipe.raw_graph_cfg(dng, noise_a=na, noise_b=nb))
    g.export("o-pfm", str(tmp / name))
    g.plan()
    d = batch.describe(g)
    g.close()
    return d, cfa[:d.cfa_ht, :d.cfa_wd]


def develop(descs, cfas, static):
    """outputs of one batch over `descs` (mosaics uploaded, outputs allocated here)"""
    bufs = []
    for d, cfa in zip(descs, cfas):
        src = H.DeviceBuffer(d.cfa_wd, d.cfa_ht, 1, H.FMT_UI16)
        src.upload(cfa)
        dst = H.DeviceBuffer(d.wd, d.ht, 4, H.FMT_F32)
        dst.clear()
        d.cfa, d.out = src.ptr, dst.ptr
        bufs.append((src, dst))
    arr = (batch.DevelopImage * len(descs))(*descs)
    L = H.lib()
    try:
        if static:
            L.vkdt_hip_batch_develop_static.argtypes = [C.POINTER(batch.DevelopImage), C.c_int, C.c_void_p]
            H.check(L.vkdt_hip_batch_develop_static(arr, len(descs), None), "develop static")
        else:
            nbytes = batch.scratch_bytes(list(descs))
            scratch = H.DeviceBuffer(max(1, (nbytes + 1) // 2), 1, 1, H.FMT_UI16)
            try:
                b = batch.Batch(list(descs), scratch.ptr, scratch.nbytes)
                b.run(None, graph=False)
                H.sync()
                b.close()
            finally:
                scratch.free()
        H.sync()
        return [dst.download() for _, dst in bufs]
    finally:
        for s, d in bufs:
            s.free()
            d.free()


def test_static_scratch_equals_caller_scratch(tmp_path, hip):
    first = [planned(tmp_path, "a", "bayer", 640, 480, 51, 3e-4), planned(tmp_path, "b", "xtrans", 558, 420, 52, 6e-4),
             planned(tmp_path, "c", "bayer", 1024, 768, 53, 2e-4)]
    second = [planned(tmp_path, "d", "bayer", 320, 240, 54, 5e-4)]
    for tag, imgs in (("first", first), ("second, smaller", second)):
        descs = [d for d, _ in imgs]
        cfas = [c for _, c in imgs]
        need = batch.scratch_bytes(descs)
        print(f"[no_alloc] {tag}: {len(descs)} images, scratch {need / 2**20:.1f} MiB", flush=True)
        ref = develop(descs, cfas, static=False)
        got = develop(descs, cfas, static=True)
        for i, (r, o) in enumerate(zip(ref, got)):
            assert np.isfinite(o).all() and (o[..., 3] == 1).all(), (tag, i)
            assert np.array_equal(r, o), (tag, i)


def test_static_scratch_refuses_oversized_batches(tmp_path, hip):
    d, cfa = planned(tmp_path, "big", "bayer", 1024, 768, 55, 3e-4)
    per = batch.scratch_bytes([d])
    n = STATIC_BYTES // per + 2
    descs = [d] * n
    assert batch.scratch_bytes(descs) > STATIC_BYTES
    arr = (batch.DevelopImage * n)(*descs)
    L = H.lib()
    L.vkdt_hip_batch_develop_static.argtypes = [C.POINTER(batch.DevelopImage), C.c_int, C.c_void_p]
    rc = L.vkdt_hip_batch_develop_static(arr, n, None)     # refused before any device work
    msg = L.vkdt_hip_last_error()
    print(f"[no_alloc] {n} images: rc {rc}, {msg}", flush=True)
    assert rc != 0 and b"static scratch" in msg
