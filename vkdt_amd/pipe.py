"""ctypes binding of the host graph runtime (libvkdt_pipe.so).

A Graph reads a vkdt .cfg, replaces its display by an output module and
runs it: the same code path as ``vkdt_amd/bin/vkdt-cli``.  Node-level
accessors expose the DAG the modules' create_nodes built (module/kernel
tokens, push constants, connectors) for tests and the benchmark.
"""
import ctypes as C
import os

from . import PKG_DIR, lib_path
from .hip import Buf, check as hip_check, lib as hip_lib

RUN_ROI, RUN_CREATE, RUN_ALLOC, RUN_RECORD, RUN_UPLOAD, RUN_DOWNLOAD, RUN_WAIT = (1 << i for i in range(7))
RUN_ALL = 0xffffffff

_lib = None


def lib():
    global _lib
    if _lib is None:
        path = lib_path("libvkdt_pipe.so")
        if not os.path.exists(path):
            raise ImportError(f"libvkdt_pipe.so not built ({path}); run __graft_entry__.build()")
        hip_lib()   # the runtime links the C ABI library
        L = C.CDLL(path)
        vp, i32, u32, u64, cp = C.c_void_p, C.c_int, C.c_uint32, C.c_uint64, C.c_char_p
        L.dt_pipe_global_init.argtypes = [cp]
        L.dt_graph_new.restype = vp
        L.dt_graph_free.argtypes = [vp]
        L.dt_graph_read_config_ascii.argtypes = [vp, cp]
        L.dt_graph_read_config_line.argtypes = [vp, cp]
        L.dt_graph_run.argtypes = [vp, u32]
        L.dt_graph_plan.argtypes = [vp, u32]
        L.dt_graph_replace_display.argtypes = [vp, u64, u64]
        L.dt_graph_disconnect_display_modules.argtypes = [vp]
        L.dt_graph_set_export_size.argtypes = [vp, i32, i32, i32]
        L.dt_graph_replace_display_colour.argtypes = [vp, u64, u64, i32, i32]
        L.dt_graph_last_error.argtypes = [vp]
        L.dt_graph_last_error.restype = cp
        L.dt_graph_set_perf.argtypes = [vp, i32]
        L.dt_graph_set_skip_dead.argtypes = [vp, i32]
        L.dt_graph_set_replay.argtypes = [vp, i32]
        L.dt_graph_set_frame.argtypes = [vp, i32]
        L.dt_graph_frame_cnt.argtypes = [vp]
        L.dt_graph_frame_rate.argtypes = [vp]
        L.dt_graph_frame_rate.restype = C.c_float
        L.dt_graph_num_nodes.argtypes = [vp]
        L.dt_graph_num_exec.argtypes = [vp]
        L.dt_graph_exec_node.argtypes = [vp, i32]
        L.dt_graph_exec_ms.argtypes = [vp, i32]
        L.dt_graph_exec_ms.restype = C.c_float
        L.dt_graph_stream.argtypes = [vp]
        L.dt_graph_stream.restype = vp
        L.dt_graph_arena_bytes.argtypes = [vp]
        L.dt_graph_arena_bytes.restype = u64
        L.dt_graph_node_info.argtypes = [vp, i32, C.POINTER(u64), C.POINTER(u64), C.POINTER(u32),
                                         C.POINTER(vp), C.POINTER(u64), C.POINTER(i32), C.POINTER(i32)]
        L.dt_graph_node_params.argtypes = [vp, i32, C.POINTER(u64)]
        L.dt_graph_node_params.restype = vp
        L.dt_graph_node_conn.argtypes = [vp, i32, i32, C.POINTER(u64), C.POINTER(i32), C.POINTER(Buf)]
        L.dt_graph_module.argtypes = [vp, cp, cp]
        L.dt_graph_set_param.argtypes = [vp, i32, cp, cp]
        L.dt_graph_node_buffer.argtypes = [vp, i32, i32]
        L.dt_graph_node_buffer.restype = vp
        f4, u4 = C.c_float * 4, C.c_uint32 * 4
        L.dt_graph_img_meta_get.argtypes = [vp, i32, f4, cp, cp, cp]
        L.dt_graph_img_param_get.argtypes = [vp, i32, f4, f4, f4, C.c_float * 9, C.POINTER(u32), u4,
                                             C.c_float * 2, C.c_uint32 * 2]
        L.dt_graph_commit.argtypes = [vp]
        L.dt_graph_read_source.argtypes = [vp, i32, vp]
        L.dt_graph_node_read_source.argtypes = [vp, i32, vp]
        L.dt_graph_module_audio.argtypes = [vp, i32, u64, u32, C.POINTER(C.c_void_p)]
        L.dt_graph_img_snd_get.argtypes = [vp, i32, C.c_int * 3]
        rc = L.dt_pipe_global_init(PKG_DIR.encode())
        if rc != 0:
            raise ImportError(f"no module classes found under {PKG_DIR}/modules")
        _lib = L
    return _lib


def tok_str(t):
    return int(t).to_bytes(8, "little").rstrip(b"\0").decode(errors="replace")


def token(s):
    return int.from_bytes(s.encode()[:8].ljust(8, b"\0"), "little")


class PipeError(RuntimeError):
    pass


class Graph:
    def __init__(self, cfg=None, lines=None):
        self.L = lib()
        self.g = self.L.dt_graph_new()
        if not self.g:
            raise PipeError("dt_graph_new failed")
        if cfg is not None and self.L.dt_graph_read_config_ascii(self.g, cfg.encode()):
            raise PipeError(f"reading {cfg}: {self.error()}")
        for l in lines or []:
            if self.L.dt_graph_read_config_line(self.g, l.encode()) < 0:
                raise PipeError(f"cfg line '{l}': {self.error()}")

    def error(self):
        return self.L.dt_graph_last_error(self.g).decode(errors="replace")

    def export(self, output_module="o-pfm", filename=None, inst="main", max_wd=0, max_ht=0, prim=2, trc=0):
        """replace display:inst by an output module; max_wd/max_ht scale the
        image to fit (vkdt-cli --width/--height); prim/trc other than linear
        rec2020 (2, 0) put a colenc module in between (--colour-prim/-trc)"""
        m = self.L.dt_graph_replace_display_colour(self.g, token(inst), token(output_module), prim, trc)
        if m < 0:
            raise PipeError(f"replace display:{inst} failed ({m}): {self.error()}")
        self.L.dt_graph_disconnect_display_modules(self.g)
        if filename is not None:
            self.set_param(output_module, inst, "filename", filename)
        if max_wd or max_ht:
            self.L.dt_graph_set_export_size(self.g, m, int(max_wd), int(max_ht))
        return m

    def set_param(self, module, inst, param, value):
        m = self.L.dt_graph_module(self.g, module.encode(), inst.encode())
        if m < 0:
            raise PipeError(f"no module {module}:{inst}")
        if isinstance(value, (list, tuple)):
            value = ":".join(str(v) for v in value)
        rc = self.L.dt_graph_set_param(self.g, m, param.encode(), str(value).encode())
        if rc != 0:
            raise PipeError(f"set {module}:{inst}:{param}={value} failed ({rc})")

    def img_param(self, module, inst):
        """metadata the module publishes after the roi pass"""
        m = self.L.dt_graph_module(self.g, module.encode(), inst.encode())
        if m < 0:
            raise PipeError(f"no module {module}:{inst}")
        f4 = C.c_float * 4
        black, white, wb, cam = f4(), f4(), f4(), (C.c_float * 9)()
        filters, crop, noise, roi = C.c_uint32(), (C.c_uint32 * 4)(), (C.c_float * 2)(), (C.c_uint32 * 2)()
        self.L.dt_graph_img_param_get(self.g, m, black, white, wb, cam, C.byref(filters), crop, noise, roi)
        meta, maker, model, dtime = f4(), C.create_string_buffer(32), C.create_string_buffer(32), C.create_string_buffer(20)
        self.L.dt_graph_img_meta_get(self.g, m, meta, maker, model, dtime)
        snd = (C.c_int * 3)()
        self.L.dt_graph_img_snd_get(self.g, m, snd)
        return dict(black=list(black), white=list(white), wb=list(wb), cam=list(cam), filters=filters.value,
                    snd_format=snd[0], snd_channels=snd[1], snd_samplerate=snd[2],
                    crop=list(crop), noise=list(noise), full=(roi[0], roi[1]), exposure=meta[0], aperture=meta[1],
                    iso=meta[2], focal_length=meta[3], maker=maker.value.decode(errors="replace"),
                    model=model.value.decode(errors="replace"), datetime=dtime.value.decode(errors="replace"))

    def commit(self):
        self.L.dt_graph_commit(self.g)

    def read_source(self, module, inst, nbytes):
        """host side of a source upload: the bytes read_source stages"""
        import numpy as np
        m = self.L.dt_graph_module(self.g, module.encode(), inst.encode())
        buf = np.zeros(nbytes, np.uint8)
        if self.L.dt_graph_read_source(self.g, m, buf.ctypes.data):
            raise PipeError(f"read_source {module}:{inst} failed: {self.error()}")
        return buf

    def audio(self, module, inst, beg, cnt, bytes_per_sample):
        """a module's audio callback: (samples returned, their bytes)"""
        m = self.L.dt_graph_module(self.g, module.encode(), inst.encode())
        if m < 0:
            raise PipeError(f"no module {module}:{inst}")
        p = C.c_void_p()
        k = self.L.dt_graph_module_audio(self.g, m, beg, cnt, C.byref(p))
        return k, (C.string_at(p, k * bytes_per_sample) if k > 0 else b"")

    def plan(self, run=RUN_ALL):
        if self.L.dt_graph_plan(self.g, run):
            raise PipeError(f"plan failed: {self.error()}")

    def run(self, run=RUN_ALL):
        rc = self.L.dt_graph_run(self.g, run)
        if rc:
            raise PipeError(f"run failed ({rc}): {self.error()}")

    def set_perf(self, on=True):
        self.L.dt_graph_set_perf(self.g, 1 if on else 0)

    def set_skip_dead(self, on=True):
        self.L.dt_graph_set_skip_dead(self.g, 1 if on else 0)

    def set_replay(self, on=True):
        """record repeated runs' node launches once and replay them as one hipGraph"""
        self.L.dt_graph_set_replay(self.g, 1 if on else 0)

    def set_frame(self, frame):
        """the frame the next run delivers (video sources)"""
        self.L.dt_graph_set_frame(self.g, frame)

    def frame_cnt(self):
        return self.L.dt_graph_frame_cnt(self.g)

    def frame_rate(self):
        return self.L.dt_graph_frame_rate(self.g)

    def stream(self):
        return self.L.dt_graph_stream(self.g)

    def arena_bytes(self):
        return self.L.dt_graph_arena_bytes(self.g)

    def exec_nodes(self):
        return [self.L.dt_graph_exec_node(self.g, k) for k in range(self.L.dt_graph_num_exec(self.g))]

    def exec_ms(self):
        return [self.L.dt_graph_exec_ms(self.g, k) for k in range(self.L.dt_graph_num_exec(self.g))]

    def node(self, n):
        name, kern = C.c_uint64(), C.c_uint64()
        dim = (C.c_uint32 * 3)()
        pc, pcs = C.c_void_p(), C.c_uint64()
        nconn, culled = C.c_int(), C.c_int()
        if self.L.dt_graph_node_info(self.g, n, C.byref(name), C.byref(kern), dim, C.byref(pc), C.byref(pcs),
                                     C.byref(nconn), C.byref(culled)):
            raise PipeError(f"no node {n}")
        psz = C.c_uint64()
        pp = self.L.dt_graph_node_params(self.g, n, C.byref(psz))
        conns = []
        for c in range(nconn.value):
            tk = (C.c_uint64 * 4)()
            link = (C.c_int * 2)()
            b = Buf()
            self.L.dt_graph_node_conn(self.g, n, c, tk, link, C.byref(b))
            conns.append(dict(name=tok_str(tk[0]), type=tok_str(tk[1]), chan=tok_str(tk[2]),
                              format=tok_str(tk[3]), link=(link[0], link[1]), buf=b))
        return dict(id=n, module=tok_str(name.value), kernel=tok_str(kern.value), wd=dim[0], ht=dim[1], dp=dim[2],
                    pc=C.string_at(pc.value, pcs.value) if pcs.value else b"",
                    params=C.string_at(pp, psz.value) if pp and psz.value else b"",
                    conns=conns, culled=bool(culled.value))

    def nodes(self):
        return [self.node(n) for n in range(self.L.dt_graph_num_nodes(self.g))]

    def close(self):
        if self.g:
            self.L.dt_graph_free(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def noise_profile_cfg(dng):
    """cfg lines of the noise-profiling graph i-raw -> rawhist -> nprof
    (bin/presets/noise-profile.pst without its display): nprof writes
    <maker>-<model>-<iso>.nprof into the working directory"""
    return ["module:i-raw:main", "module:rawhist:np", "module:nprof:np",
            "connect:i-raw:main:output:rawhist:np:input",
            "connect:rawhist:np:data:nprof:np:input",
            f"param:i-raw:main:filename:{dng}"]


def raw_graph_cfg(dng, strength=0.5, noise_a=0.0, noise_b=0.0, clarity=0.2, numgamma=10, levels=12,
                  extra=(), hilite=False):
    """cfg lines of the raw-development graph i-raw -> denoise -> [hilite ->]
    demosaic -> colour -> filmcurv -> llap -> display (with hilite: the
    chain of bin/default-darkroom.i-raw up to llap)."""
    mid = (["connect:denoise:01:output:hilite:01:input", "connect:hilite:01:output:demosaic:01:input"]
           if hilite else ["connect:denoise:01:output:demosaic:01:input"])
    lines = [
        "module:i-raw:main", "module:denoise:01", "module:demosaic:01", "module:colour:01",
        "module:filmcurv:01", "module:llap:01", "module:display:main",
    ] + (["module:hilite:01"] if hilite else []) + [
        "connect:i-raw:main:output:denoise:01:input",
    ] + mid + [
        "connect:demosaic:01:output:colour:01:input",
        "connect:colour:01:output:filmcurv:01:input",
        "connect:filmcurv:01:output:llap:01:input",
        "connect:llap:01:output:display:main:input",
        f"param:i-raw:main:filename:{dng}",
        f"param:i-raw:main:noise a:{noise_a!r}",
        f"param:i-raw:main:noise b:{noise_b!r}",
        f"param:denoise:01:strength:{strength}",
        f"param:llap:01:clarity:{clarity}",
        f"param:llap:01:numgamma:{numgamma}",
        f"param:llap:01:levels:{levels}",
    ]
    return lines + list(extra)


def mlv_graph_cfg(mlv, strength=0.5, clarity=0.2, numgamma=10, levels=12, extra=()):
    """cfg lines of the raw-development graph with a raw video source:
    i-mlv -> denoise -> demosaic -> colour -> filmcurv -> llap -> display"""
    lines = [l.replace("i-raw:main", "i-mlv:main") for l in
             raw_graph_cfg(mlv, strength=strength, clarity=clarity, numgamma=numgamma, levels=levels)
             if "noise a" not in l and "noise b" not in l]
    return lines + list(extra)
