"""GPU-against-oracle parity checks shared by the GPU tests and smoke().

Per node (the oracle's float build run on the GPU's own inputs): outside
the texels the oracle flags as near ties (tests/errbar.py), every output
texel is within MAX_NODE_ULP binary16 ulps of the oracle.  gauss's
covariance is compared as splat reads it: eigenvalues clamped to
[0.01, 25], the axis up to its sign.

End to end (the oracle's whole-graph float run on the same raw): PSNR of
at least 60 dB on linear rec2020, and the pixels beyond the per-pixel bound
of tests/errbar.py (1e-3, or three f16 storage quanta) number no more than
twice those of the oracle's perturbed float build (liboracle32p, another
float implementation that differs from it only in the last place of its
transcendentals) on the same raw, plus two: the GPU is as close to the
oracle as a second float implementation of the reference is.  The FP64
oracle's error is reported beside it.
"""
import ctypes as C

import numpy as np

import errbar
import graphspec as gs
import oracle_py as op

MAX_NODE_ULP = 4
NODE_TIE_EPS = 1e-5      # identical inputs: the two sides differ in their own rounding only
TWIN_SEEDS = (1, 2)


def run_gpu_graph(hip, bufs, nodes, raw, sources=None):
    """every node on the GPU through the C ABI, in graph order; returns the
    host copy of every buffer.  `sources` maps further source buffers (the
    gain map) to their host arrays."""
    dev = {k: hip.DeviceBuffer(*v) for k, v in bufs.items()}
    dev["raw"].upload(raw)
    for k, v in (sources or {}).items():
        dev[k].upload(v)
    for n in nodes:
        hip.launch(n["module"], n["kernel"], n["pc"], n["params"], [dev[c] for c in n["conn"]],
                   n["wd"], n["ht"], n["dp"])
    hip.sync()
    host = {k: b.download() for k, b in dev.items()}
    for b in dev.values():
        b.free()
    return host


def f16_ulps(a, b):
    """distance in binary16 ulps of the larger magnitude (nan and inf aware)"""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    same_inf = np.isinf(a) & np.isinf(b) & (np.sign(a) == np.sign(b))
    with np.errstate(invalid="ignore"):
        d = np.abs(a - b) / errbar.f16_quantum(np.maximum(np.abs(a), np.abs(b)))
    d[both_nan | same_inf] = 0
    d[np.isnan(d)] = np.inf
    return d


def node_distance(n, gpu, ref, spec):
    """per-texel distance (max over channels and planes) of one node's output"""
    gpu = np.asarray(gpu, np.float64)
    if n["module"] == "demosaic" and n["kernel"] == "gauss":
        ge, re = np.clip(gpu[..., :2], 0.01, 25), np.clip(ref[..., :2], 0.01, 25)
        de = f16_ulps(ge, re).max(-1)
        da = np.minimum(f16_ulps(gpu[..., 2:], ref[..., 2:]).max(-1), f16_ulps(gpu[..., 2:], -ref[..., 2:]).max(-1))
        return np.maximum(de, da)
    if spec[3] == gs.F16:
        d = f16_ulps(gpu, ref)
    else:   # f32: relative error in units of 2^-10
        with np.errstate(invalid="ignore", divide="ignore"):
            d = np.abs(gpu - ref) / np.maximum(np.abs(ref), 1e-3) * 1024
    wd, ht, chan, fmt, planes = spec
    if chan > 1:
        d = d.max(-1)
    if planes > 1:
        d = d.max(0)
    return d


def to_or(arr, spec):
    wd, ht, chan, fmt, planes = spec
    if fmt == gs.UI16:
        a = np.ascontiguousarray(arr, np.uint16)
        return a, op.orbuf(a, op.OR_UI16, chan, planes)
    a = np.ascontiguousarray(arr, np.float64)
    return a, op.orbuf(a, op.OR_F16 if fmt == gs.F16 else op.OR_F32, chan, planes)


def check_nodes(bufs, nodes, host, log=print, max_ulp=MAX_NODE_ULP):
    """each node against the float oracle on the GPU's own inputs; returns
    [(index, module/kernel, max ulp outside near ties, near-tie texels)]"""
    report = []
    for i, n in enumerate(nodes):
        oi = n.get("out", len(n["conn"]) - 1)
        ins = [to_or(host[c], bufs[c])[1] for c in n["conn"]]
        outname = n["conn"][oi]
        spec = bufs[outname]
        out = np.zeros(host[outname].shape, np.float64)
        fmt = op.OR_F16 if spec[3] == gs.F16 else (op.OR_UI16 if spec[3] == gs.UI16 else op.OR_F32)
        if fmt == op.OR_UI16:
            out = np.zeros(host[outname].shape, np.uint16)
        ties = np.zeros((spec[1], spec[0]), np.uint8)
        ins[oi] = op.orbuf(out, fmt, spec[2], spec[4])
        op.node(32, n["module"], n["kernel"], n["pc"], n["params"], ins,
                n["wd"], n["ht"], n["dp"], ties=ties, tie_eps=NODE_TIE_EPS)
        d = node_distance(n, host[outname], out, spec)
        flagged = ties.astype(bool)
        worst = float(d[~flagged].max()) if (~flagged).any() else 0.0
        report.append((i, f"{n['module']}/{n['kernel']}", worst, int(flagged.sum())))
        log(f"  node {i:2d} {n['module']}/{n['kernel']}: max {worst:.3g} ulp outside {int(flagged.sum())} near-tie "
            f"texels (max {float(d.max()):.3g} with them)")
    bad = [r for r in report if not r[2] <= max_ulp]
    assert not bad, f"nodes beyond {max_ulp} ulp outside near ties: {bad}"
    return report


def e2e_check(cfa, p, gpu, log=print, label=""):
    """the GPU's rgba against the oracle's whole-graph float run (and its
    perturbed float twins, and the FP64 oracle); returns the statistics"""
    o = gs.oracle_params(p)
    ref, st = op.develop(32, cfa, o, stages=True)
    tol = errbar.tolerance(st["filmcurv"], st["llap"])
    gpu = np.asarray(gpu, np.float64)[..., :3]
    ref = ref[..., :3]
    d = np.abs(gpu - ref).max(-1)
    n_gpu = int((d > tol).sum())
    twin = op.lib("32p")
    n_twin = []
    for s in TWIN_SEEDS:
        twin.or_set_perturb_seed(C.c_uint32(s))
        t = op.develop("32p", cfa, o)[..., :3]
        n_twin.append(int((np.abs(t - ref).max(-1) > tol).sum()))
    psnr = 10 * np.log10(1.0 / max(float(np.mean((gpu - ref) ** 2)), 1e-30))
    ref64 = op.develop(64, cfa, o)[..., :3]
    d64 = np.abs(gpu - ref64)
    psnr64 = 10 * np.log10(1.0 / max(float(np.mean(d64 ** 2)), 1e-30))
    stats = dict(max_abs=float(d.max()), n_over_1e3=int((d > 1e-3).sum()), n_over_tol=n_gpu, n_twin=n_twin,
                 storage_limited=float((tol > 1e-3).mean()), psnr=psnr, max_abs64=float(d64.max()), psnr64=psnr64,
                 pixels=int(d.size))
    log(f"[e2e{label}] vs float oracle: max abs {stats['max_abs']:.3e}, >1e-3 {stats['n_over_1e3']}, "
        f">bound {n_gpu} (perturbed float twins: {n_twin}), storage-limited {stats['storage_limited']:.3f}, "
        f"psnr {psnr:.1f} dB; vs FP64: max abs {stats['max_abs64']:.3e}, psnr {psnr64:.1f} dB")
    assert np.isfinite(gpu).all()
    assert psnr >= 60.0 and psnr64 >= 60.0, stats
    assert n_gpu <= 2 * max(n_twin) + 2, stats
    return stats


def e2e_check_nodes(bufs, nodes, cfa, gpu, sources=None, log=print, label=""):
    """e2e_check for graphs the whole-graph oracle does not build (half size,
    scaled exports, RCD, gain maps): the restated node list (graphspec) run on
    the oracle's float build, its perturbed twins and the FP64 build, against
    the GPU's rgba, by the same three criteria"""
    ref_all = op.run_nodes(32, bufs, nodes, cfa, sources=sources)
    tol = errbar.tolerance(ref_all["fc"], ref_all["out0"])
    if tol.shape != ref_all["final"].shape[:2]:          # a scaled export: the bound of the output extent
        tol = np.full(ref_all["final"].shape[:2], 1e-3)
    ref = ref_all["final"][..., :3]
    gpu = np.asarray(gpu, np.float64)[..., :3]
    d = np.abs(gpu - ref).max(-1)
    n_gpu = int((d > tol).sum())
    twin = op.lib("32p")
    n_twin = []
    for s in TWIN_SEEDS:
        twin.or_set_perturb_seed(C.c_uint32(s))
        t = op.run_nodes("32p", bufs, nodes, cfa, sources=sources)["final"][..., :3]
        n_twin.append(int((np.abs(t - ref).max(-1) > tol).sum()))
    psnr = 10 * np.log10(1.0 / max(float(np.mean((gpu - ref) ** 2)), 1e-30))
    ref64 = op.run_nodes(64, bufs, nodes, cfa, sources=sources)["final"][..., :3]
    d64 = np.abs(gpu - ref64)
    psnr64 = 10 * np.log10(1.0 / max(float(np.mean(d64 ** 2)), 1e-30))
    stats = dict(max_abs=float(d.max()), n_over_1e3=int((d > 1e-3).sum()), n_over_tol=n_gpu, n_twin=n_twin,
                 psnr=psnr, max_abs64=float(d64.max()), psnr64=psnr64, pixels=int(d.size))
    log(f"[e2e{label}] vs float oracle (node list): max abs {stats['max_abs']:.3e}, >1e-3 {stats['n_over_1e3']}, "
        f">bound {n_gpu} (perturbed float twins: {n_twin}), psnr {psnr:.1f} dB; vs FP64: max abs "
        f"{stats['max_abs64']:.3e}, psnr {psnr64:.1f} dB")
    assert np.isfinite(gpu).all()
    assert psnr >= 60.0 and psnr64 >= 60.0, stats
    assert n_gpu <= 2 * max(n_twin) + 2, stats
    return stats
