"""ctypes binding of the oracle (oracle/oracle.h) for the tests.

The oracle is the checker: tests compare the HIP path against it, nothing
else in the repository calls it.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle", "_build")

OR_UI16, OR_F16, OR_F32 = 1, 2, 3


class OrBuf(C.Structure):
    _fields_ = [("data", C.c_void_p), ("wd", C.c_int32), ("ht", C.c_int32), ("chan", C.c_int32),
                ("format", C.c_int32), ("array_length", C.c_int32)]


class OrDevelop(C.Structure):
    _fields_ = [
        ("filters", C.c_int32),
        ("black", C.c_float * 4), ("white", C.c_float * 4), ("wb", C.c_float * 4),
        ("cam_to_rec2020", C.c_float * 9),
        ("noise_a", C.c_float), ("noise_b", C.c_float),
        ("dn_strength", C.c_float), ("dn_luma", C.c_float), ("dn_detail", C.c_float), ("dn_edges", C.c_float * 4),
        ("dm_colour", C.c_int32),
        ("co_exposure", C.c_float), ("co_white", C.c_float * 3),
        ("fc_light", C.c_float), ("fc_contrast", C.c_float), ("fc_bias", C.c_float), ("fc_colour", C.c_int32),
        ("ll_sigma", C.c_float), ("ll_shadows", C.c_float), ("ll_hilights", C.c_float), ("ll_clarity", C.c_float),
        ("ll_num_gamma", C.c_int32), ("ll_max_levels", C.c_int32),
        ("hl_on", C.c_int32), ("hl_white", C.c_float), ("hl_desat", C.c_float), ("hl_soft", C.c_float),
    ]


_libs = {}


def lib(bits=64):
    """64: the double build, 32: the float build, "32p": the float build
    with its transcendentals moved by one ulp"""
    if bits not in _libs:
        path = os.path.join(ORACLE_DIR, f"liboracle{bits}.so")
        if not os.path.exists(path):
            raise ImportError(f"oracle not built: {path}")
        _libs[bits] = load(path)
    return _libs[bits]


def load(path):
    """an oracle build at `path` with its entry points typed"""
    L = C.CDLL(path)
    L.or_node.argtypes = [C.c_char_p, C.c_char_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                          C.POINTER(OrBuf), C.c_int, C.c_int, C.c_int, C.c_int]
    L.or_develop.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(OrDevelop), C.c_void_p,
                             C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.or_f16_round.argtypes = [C.c_double]
    L.or_f16_round.restype = C.c_double
    L.or_mirror.argtypes = [C.c_int, C.c_int]
    L.or_bilinear.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double]
    L.or_bilinear.restype = C.c_double
    L.or_gather.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p]
    L.or_evd2x2.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.or_set_storage_f16.argtypes = [C.c_int]
    L.or_nprof_fit.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.or_set_ties.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double]
    L.or_set_develop_ties.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int]
    L.or_set_tie_effect.argtypes = [C.c_double]
    L.or_set_threads.argtypes = [C.c_int]
    L.or_fc_space.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int]
    L.or_mcraw_prefix.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
    L.or_mcraw_block.argtypes = [C.c_void_p, C.c_size_t, C.c_uint32, C.c_uint32, C.c_void_p]
    L.or_mcraw_blocks.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
    L.or_mcraw_decode.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_int,
                                  C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int]
    return L


def as_double(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def orbuf(arr, fmt, chan, planes=1):
    """OrBuf over a numpy array (uint16 for OR_UI16, float64 otherwise)."""
    if fmt == OR_UI16:
        assert arr.dtype == np.uint16 and arr.flags.c_contiguous
    else:
        assert arr.dtype == np.float64 and arr.flags.c_contiguous
    if planes > 1:
        ht, wd = arr.shape[1], arr.shape[2]
    else:
        ht, wd = arr.shape[0], arr.shape[1]
    b = OrBuf(arr.ctypes.data, wd, ht, chan, fmt, planes)
    b._keep = arr   # the pointer is only valid while the array lives (callers pass temporaries)
    return b


def node(bits, module, kernel, pc, params, bufs, wd, ht, dp=1, storage_f16=True, ties=None, tie_eps=1e-5):
    """one node on the oracle; `ties` (uint8 array of the output extent, or
    None) receives the node's near-tie flags at margin `tie_eps`"""
    L = lib(bits)
    L.or_set_storage_f16(1 if storage_f16 else 0)
    if ties is not None:
        assert ties.dtype == np.uint8 and ties.flags.c_contiguous
        L.or_set_ties(ties.ctypes.data, ties.shape[1], ties.shape[0], tie_eps)
    conn = (OrBuf * len(bufs))(*bufs)
    pcb = None if pc is None else np.ascontiguousarray(pc)
    pab = None if params is None else np.ascontiguousarray(params)
    rc = L.or_node(module.encode(), kernel.encode(),
                   None if pcb is None else pcb.ctypes.data, 0 if pcb is None else pcb.nbytes,
                   None if pab is None else pab.ctypes.data, 0 if pab is None else pab.nbytes,
                   conn, len(bufs), wd, ht, dp)
    L.or_set_ties(None, 0, 0, 0.0)
    if rc != 0:
        raise RuntimeError(f"oracle {module}/{kernel} failed: {rc}")


def develop(bits, cfa, p, stages=False, storage_f16=True, ties=None, tie_eps=2e-3, tie_radii=(16, 4),
            tie_effect=2.5e-4):
    """Whole graph on the CPU; returns rgba float64 H x W x 4 (and stages).
    `ties` (uint8 H x W, or None) receives the graph's near-tie mask: margins
    within `tie_eps` of the operands' magnitude whose other branch moves the
    node's result by more than `tie_effect`, denoise flags dilated by
    tie_radii[0] block texels and demosaic flags by tie_radii[1] pixels."""
    L = lib(bits)
    L.or_set_storage_f16(1 if storage_f16 else 0)
    H, W = cfa.shape
    if ties is not None:
        assert ties.dtype == np.uint8 and ties.shape == (H, W) and ties.flags.c_contiguous
        L.or_set_develop_ties(ties.ctypes.data, tie_eps, tie_radii[0], tie_radii[1])
        L.or_set_tie_effect(tie_effect)
    cfa = np.ascontiguousarray(cfa, dtype=np.uint16)
    out = np.zeros((H, W, 4), np.float64)
    st = {}
    if stages:
        st = {"denoise": np.zeros((H, W), np.float64), "demosaic": np.zeros((H, W, 4), np.float64),
              "colour": np.zeros((H, W, 4), np.float64), "filmcurv": np.zeros((H, W, 4), np.float64),
              "llap": np.zeros((H, W), np.float64)}
    ptr = lambda k: st[k].ctypes.data if k in st else None
    rc = L.or_develop(cfa.ctypes.data, W, H, C.byref(p), out.ctypes.data,
                      ptr("denoise"), ptr("demosaic"), ptr("colour"), ptr("filmcurv"), ptr("llap"))
    L.or_set_develop_ties(None, 0.0, 0, 0)
    L.or_set_tie_effect(0.0)
    if rc != 0:
        raise RuntimeError(f"oracle develop failed: {rc}")
    return (out, st) if stages else out


def rawhist_collect(bits, raw, filters, black, white, bins=1000, rows=5, noise=(0.0, 0.0), node_wd=None,
                    node_ht=None):
    """rawhist/collect on the oracle: accumulators rows x bins (float64)"""
    H, W = raw.shape
    blk = 3 if filters == 9 else 2
    raw = np.ascontiguousarray(raw, np.uint16)
    acc = np.zeros((rows, bins), np.float64)
    par = np.zeros(5, np.uint32)
    par[0:2] = np.array(noise, np.float32).view(np.uint32)
    par[2:5] = [filters, black, white]
    node(bits, "rawhist", "collect", None, par, [orbuf(raw, OR_UI16, 1), orbuf(acc, OR_F32, 1)],
         W // blk if node_wd is None else node_wd, H // blk if node_ht is None else node_ht)
    return acc


def nprof_fit(bits, acc):
    """nprof's fit on accumulators (rows x bins): (a, b, bins kept)"""
    acc = np.ascontiguousarray(acc, np.float64)
    a, b = C.c_float(), C.c_float()
    n = lib(bits).or_nprof_fit(acc.ctypes.data, acc.shape[1], C.byref(a), C.byref(b))
    return a.value, b.value, n


def run_nodes(bits, bufs, nodes, raw, storage_f16=True, sources=None):
    """a restated node list (graphspec) on the oracle from the ui16 mosaic
    `raw` (and further source buffers, name -> array); returns every buffer
    (float64, uint16 for ui16 buffers)"""
    data = {}
    for name, (wd, ht, chan, fmt, planes) in bufs.items():
        shape = ((planes,) if planes > 1 else ()) + (ht, wd) + ((chan,) if chan > 1 else ())
        data[name] = np.zeros(shape, np.uint16 if fmt == OR_UI16 else np.float64)
    data["raw"][...] = raw
    for name, arr in (sources or {}).items():
        data[name][...] = arr
    for n in nodes:
        ob = [orbuf(data[c], bufs[c][3], bufs[c][2], bufs[c][4]) for c in n["conn"]]
        node(bits, n["module"], n["kernel"], n["pc"], n["params"], ob, n["wd"], n["ht"], n["dp"],
             storage_f16=storage_f16)
    return data


def fc_space(bits, which, v):
    """filmcurv's colour space `which` (or_fc_space) on triples v (... x 3)"""
    v = np.ascontiguousarray(v, np.float64)
    out = np.zeros_like(v)
    n = v.size // 3
    if lib(bits).or_fc_space(which, v.ctypes.data, out.ctypes.data, n) != 0:
        raise ValueError(f"no colour space {which}")
    return out
