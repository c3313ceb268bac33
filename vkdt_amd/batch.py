"""ctypes binding of the batched develop path (vkdt_hip_batch_* in
libvkdt_hip.so) and the descriptor filler of the host runtime
(dt_graph_develop_image in libvkdt_pipe.so).

A batch develops many stills or frames -- mixed sizes, Bayer and X-Trans --
with one launch per stage over all of them.  Each image is described by the
values its own planned per-node graph would use, so its output equals that
graph's output.
"""
import ctypes as C

import numpy as np

from . import hip, pipe

STAGES = 10


class DevelopImage(C.Structure):
    """vkdt_develop_image_t (include/vkdt_hip.h)"""
    _fields_ = [
        ("cfa", C.c_void_p), ("cfa_wd", C.c_uint32), ("cfa_ht", C.c_uint32), ("cfa_pitch", C.c_uint32),
        ("out", C.c_void_p), ("wd", C.c_uint32), ("ht", C.c_uint32), ("filters", C.c_uint32),
        ("wb", C.c_float * 4), ("black", C.c_float * 4), ("white", C.c_float * 4), ("crop", C.c_int32 * 4),
        ("noise_a", C.c_float), ("noise_b", C.c_float), ("dn_params", C.c_float * 8), ("dm_fixup", C.c_int32),
        ("co_params", C.c_float * 242), ("fc_params", C.c_float * 4), ("ll_params", C.c_float * 4),
        ("ll_numgamma", C.c_int32), ("ll_levels", C.c_int32),
        ("hl_on", C.c_int32), ("hl_wb", C.c_float * 4), ("hl_params", C.c_float * 3), ("hl_levels", C.c_int32),
    ]


_bound = False


def _bind():
    global _bound
    if _bound:
        return
    L, P = hip.lib(), pipe.lib()
    vp, i32, sz = C.c_void_p, C.c_int, C.c_size_t
    DI = C.POINTER(DevelopImage)
    L.vkdt_hip_batch_scratch_bytes.argtypes = [DI, i32, C.POINTER(sz)]
    L.vkdt_hip_batch_create.argtypes = [C.POINTER(vp), DI, i32, vp, sz, C.POINTER(C.c_int32)]
    L.vkdt_hip_batch_run.argtypes = [vp, vp, i32]
    L.vkdt_hip_batch_autotune.argtypes = [vp, vp, C.POINTER(C.c_int32)]
    L.vkdt_hip_batch_set_timing.argtypes = [vp, i32]
    L.vkdt_hip_batch_stage_ms.argtypes = [vp, C.POINTER(C.c_float)]
    L.vkdt_hip_batch_stage_name.argtypes = [i32]
    L.vkdt_hip_batch_stage_name.restype = C.c_char_p
    L.vkdt_hip_batch_destroy.argtypes = [vp]
    P.dt_graph_develop_image.argtypes = [vp, DI]
    _bound = True


def describe(graph):
    """the descriptor of a planned pipe.Graph (device pointers left NULL)"""
    _bind()
    d = DevelopImage()
    if pipe.lib().dt_graph_develop_image(graph.g, C.byref(d)):
        raise pipe.PipeError(graph.error())
    return d


def stage_names():
    _bind()
    return [hip.lib().vkdt_hip_batch_stage_name(k).decode() for k in range(STAGES)]


def scratch_bytes(images):
    _bind()
    arr = (DevelopImage * len(images))(*images)
    n = C.c_size_t()
    hip.check(hip.lib().vkdt_hip_batch_scratch_bytes(arr, len(images), C.byref(n)), "batch scratch")
    return n.value


class Batch:
    """A planned batch over caller-owned mosaics, outputs and scratch."""

    def __init__(self, images, scratch_ptr, scratch_nbytes, tune=None):
        _bind()
        self.L = hip.lib()
        self.n = len(images)
        self._arr = (DevelopImage * self.n)(*images)
        t = (C.c_int32 * STAGES)(*tune) if tune is not None else None
        p = C.c_void_p()
        hip.check(self.L.vkdt_hip_batch_create(C.byref(p), self._arr, self.n, scratch_ptr, scratch_nbytes, t),
                  "batch create")
        self.ptr = p.value

    def run(self, stream=None, graph=False):
        hip.check(self.L.vkdt_hip_batch_run(self.ptr, stream, 1 if graph else 0), "batch run")

    def autotune(self, stream=None):
        """time the launch shapes of each stage group and keep the fastest;
        returns them as ints: persistent work groups per CU (0: one tile per
        work group) | rows per thread of the per-pixel stages << 8"""
        t = (C.c_int32 * STAGES)()
        hip.check(self.L.vkdt_hip_batch_autotune(self.ptr, stream, t), "batch autotune")
        return list(t)

    def set_timing(self, on=True):
        hip.check(self.L.vkdt_hip_batch_set_timing(self.ptr, 1 if on else 0), "batch timing")

    def stage_ms(self):
        ms = (C.c_float * STAGES)()
        hip.check(self.L.vkdt_hip_batch_stage_ms(self.ptr, ms), "batch stage ms")
        return list(ms)

    def close(self):
        if getattr(self, "ptr", None):
            self.L.vkdt_hip_batch_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def develop(graphs, cfas, stream=None, graph=False, tune=None):
    """Develop planned graphs as one batch.  `cfas[i]` is the ui16 mosaic
    of graph i as i-raw uploads it (cfa_ht x cfa_wd).  Returns the rgba f32
    outputs (ht x wd x 4) as numpy arrays."""
    descs, bufs = [], []
    for g, cfa in zip(graphs, cfas):
        d = describe(g)
        assert cfa.shape == (d.cfa_ht, d.cfa_wd), (cfa.shape, d.cfa_ht, d.cfa_wd)
        src = hip.DeviceBuffer(d.cfa_wd, d.cfa_ht, 1, hip.FMT_UI16)
        src.upload(cfa)
        dst = hip.DeviceBuffer(d.wd, d.ht, 4, hip.FMT_F32)
        d.cfa, d.out = src.ptr, dst.ptr
        descs.append(d)
        bufs.append((src, dst))
    nbytes = scratch_bytes(descs)
    scratch = hip.DeviceBuffer(max(1, (nbytes + 1) // 2), 1, 1, hip.FMT_UI16)   # byte arena
    try:
        b = Batch(descs, scratch.ptr, scratch.nbytes, tune)
        try:
            b.run(stream, graph)
            hip.sync(stream)
            return [dst.download() for _, dst in bufs]
        finally:
            b.close()
    finally:
        scratch.free()
        for s, d in bufs:
            s.free()
            d.free()
