"""The node DAG of the raw-development graph, restated for node-level tests.

Mirrors what the modules' create_nodes build for i-raw -> denoise ->
demosaic -> colour -> filmcurv -> llap with strength > 0 and the splat
demosaic (denoise/main.c:133-333, demosaic/main.c:59-203,
colour/main.c:219-365, filmcurv/main.c:25-51, llap/main.c:14-114), so a test
can run every node on the GPU through the C ABI and on the oracle with the
same push-constant and params blobs.
"""
from dataclasses import dataclass, field

import numpy as np

from vkdt_amd import synth

UI16, F16, F32 = 1, 2, 3

# matrices.h: matrix_xyz_to_rec2020 (row major)
XYZ_TO_REC2020 = np.array([
    [1.71665119, -0.35567078, -0.25336628],
    [-0.66668435, 1.61648124, 0.01576855],
    [0.01763986, -0.04277061, 0.94210312]], dtype=np.float32)


def mat3inv_f32(m):
    """core/mat3.h:21-45 in float32."""
    A = lambda y, x: np.float32(m[y - 1][x - 1])
    det = (A(1, 1) * (A(3, 3) * A(2, 2) - A(3, 2) * A(2, 3))
           - A(2, 1) * (A(3, 3) * A(1, 2) - A(3, 2) * A(1, 3))
           + A(3, 1) * (A(2, 3) * A(1, 2) - A(2, 2) * A(1, 3)))
    idet = np.float32(1.0) / det
    B = np.zeros((3, 3), np.float32)
    B[0, 0] = idet * (A(3, 3) * A(2, 2) - A(3, 2) * A(2, 3))
    B[0, 1] = -idet * (A(3, 3) * A(1, 2) - A(3, 2) * A(1, 3))
    B[0, 2] = idet * (A(2, 3) * A(1, 2) - A(2, 2) * A(1, 3))
    B[1, 0] = -idet * (A(3, 3) * A(2, 1) - A(3, 1) * A(2, 3))
    B[1, 1] = idet * (A(3, 3) * A(1, 1) - A(3, 1) * A(1, 3))
    B[1, 2] = -idet * (A(2, 3) * A(1, 1) - A(2, 1) * A(1, 3))
    B[2, 0] = idet * (A(3, 2) * A(2, 1) - A(3, 1) * A(2, 2))
    B[2, 1] = -idet * (A(3, 2) * A(1, 1) - A(3, 1) * A(1, 2))
    B[2, 2] = idet * (A(2, 2) * A(1, 1) - A(2, 1) * A(1, 2))
    return B


def mat3mul_f32(a, b):
    out = np.zeros((3, 3), np.float32)
    for k in range(3):
        for i in range(3):
            s = np.float32(0)
            for j in range(3):
                s = np.float32(s + np.float32(a[k][j]) * np.float32(b[j][i]))
            out[k, i] = s
    return out


def cam_to_rec2020(xyz_to_cam):
    """i-raw/main.c:228-258 for a D65 matrix."""
    return mat3mul_f32(XYZ_TO_REC2020, mat3inv_f32(np.asarray(xyz_to_cam, np.float32)))


@dataclass
class DevelopParams:
    kind: str = "bayer"
    black: float = 512.0
    white: float = 16383.0
    wb: tuple = (1.0, 1.0, 1.0, 1.0)
    noise_a: float = 0.0
    noise_b: float = 0.0
    strength: float = 0.5
    luma: float = 0.6
    detail: float = 1.0
    edges: tuple = (0.0, 0.0, 0.0, 0.0)
    dm_colour: int = 0
    dm_method: int = 0          # 0 gaussian splats, 1 RCD, 2 half size
    exposure: float = 0.0
    co_white: tuple = (0.0, 0.0, 0.0)
    light: float = 3.0
    contrast: float = 1.2
    bias: float = 0.0
    fc_colour: int = 3
    sigma: float = 0.12
    shadows: float = 1.0
    hilights: float = 1.0
    clarity: float = 0.2
    num_gamma: int = 10
    max_levels: int = 12
    hilite: bool = False
    hl_white: float = 0.985
    hl_desat: float = 0.3
    hl_soft: float = 0.6
    cam: np.ndarray = field(default_factory=lambda: cam_to_rec2020(synth.xyz_to_cam()))
    # DNG gain maps (Bayer): dict(gains=(map_ht, map_wd, 4) float32 with one
    # channel per site of the 2x2 block, origin=(v, h), spacing=(v, h))
    gain_map: dict = None

    @property
    def filters(self):
        return synth.XTRANS_FILTERS if self.kind == "xtrans" else synth.BAYER_FILTERS


def f32blob(vals):
    return np.asarray(vals, dtype=np.float32)


def mixed_blob(items):
    """std140 scalar blob from (kind, value) pairs, kind 'f' or 'i'."""
    out = np.zeros(len(items), np.uint32)
    for i, (k, v) in enumerate(items):
        out[i] = np.float32(v).view(np.uint32) if k == "f" else np.uint32(np.int64(v) & 0xffffffff)
    return out


def llap_levels(W, H, max_levels):
    lw, lh = [W], [H]
    nl = max_levels
    for l in range(1, max_levels):
        lw.append((lw[-1] - 1) // 2 + 1)
        lh.append((lh[-1] - 1) // 2 + 1)
        nw, nh = (lw[-1] - 1) // 2 + 1, (lh[-1] - 1) // 2 + 1
        if nw <= 1 or nh <= 1:
            nl = l + 1
            break
    return nl, lw[:nl], lh[:nl]


def colour_params(p):
    """colour/main.c:219-365 commit_params for matrix=1, mode=0."""
    cp = np.zeros(242, np.float32)
    cu = cp.view(np.uint32)
    pwb = list(np.float32(v) for v in p.co_white)
    cam = np.asarray(p.cam, np.float32)
    if pwb[0] == 0 and pwb[1] == 0 and pwb[2] == 0:
        w = [np.float32(v) for v in p.wb[:3]]
        w0 = [np.float32(0)] * 3
        for j in range(3):
            for i in range(3):
                w0[j] = np.float32(w0[j] + cam[j, i] / w[i])
        w0[0] = np.float32(w0[0] / w0[1])
        w0[2] = np.float32(w0[2] / w0[1])
        w0[1] = np.float32(1)
        pwb = [np.float32(1) / w0[0], np.float32(1), np.float32(1) / w0[2]]
    if not (pwb[0] == pwb[0]) or pwb[0] == 0 or pwb[1] == 0 or pwb[2] == 0:
        pwb = [np.float32(1)] * 3
    cp[0] = pwb[0] / pwb[1]
    cp[1] = 1
    cp[2] = pwb[2] / pwb[1]
    cp[3] = np.float32(2.0) ** np.float32(p.exposure)
    for j in range(3):
        for i in range(3):
            cp[4 + 4 * i + j] = cam[j, i]
    cp[224] = 0.0          # clut anchor from the Kelvin slider: unused by the matrix path
    cp[226] = 1.0          # saturation
    cu[229] = 0            # primaries: custom (camera matrix from i-raw)
    cu[230] = 0            # trc: linear
    w = [np.float32(v) for v in p.wb[:3]]
    if not (w[0] > 0 and w[1] > 0 and w[2] > 0):
        w = [np.float32(1)] * 3
    cp[232:236] = [w[0] / w[1], 1.0, w[2] / w[1], 1.0]   # as-shot wb, g-normalised
    return cp


def export_extent(W, H, max_wd=0, max_ht=0):
    """the extent a sink asks for when the export is limited to max_wd x
    max_ht (graph-run-modules.h:466-470, float32 as the runtime computes it)"""
    f32 = np.float32
    sx = f32(W) / f32(max_wd) if max_wd > 0 else f32(1)
    sy = f32(H) / f32(max_ht) if max_ht > 0 else f32(1)
    s = max(sx, sy)
    return int(f32(W) / s + f32(0.5)), int(f32(H) / s + f32(0.5))


def gain_map_os(gm):
    """the map origin and inverse extent the denoise module pushes
    (denoise/main.c:187-196): origin as float, 1/(spacing*(points-1)) in
    double rounded to float"""
    mh, mw = gm["gains"].shape[:2]
    (ov, oh), (sv, sh) = gm["origin"], gm["spacing"]
    return [np.float32(oh), np.float32(ov), np.float32(1.0 / (sh * (mw - 1))), np.float32(1.0 / (sv * (mh - 1)))]


def build(W, H, p, max_wd=0, max_ht=0, prim=2, trc=0):
    """(buffers, nodes): buffers name -> (wd, ht, chan, fmt, planes);
    nodes are dicts with module, kernel, pc, params, conn (buffer names), wd, ht, dp.
    max_wd/max_ht: the export size limit (vkdt-cli --width/--height);
    prim/trc: the export encoding (a colenc node unless linear rec2020)."""
    xt = p.kind == "xtrans"
    block = 3 if xt else 2
    Qw, Qh = W // block, H // block
    black = [np.float32(p.black) / np.float32(65535.0)] * 4
    white = [np.float32(p.white) / np.float32(65535.0)] * 4
    wb = list(p.wb)
    bufs, nodes = {}, []

    def buf(name, wd, ht, chan, fmt, planes=1):
        bufs[name] = (wd, ht, chan, fmt, planes)

    def add(module, kernel, pc, params, conn, wd, ht, dp=1, out=None):
        # out: the index of the written connector (the last one unless an
        # input binding follows it, as the gain map of noop and doub)
        nodes.append(dict(module=module, kernel=kernel, pc=pc, params=params, conn=conn, wd=wd, ht=ht, dp=dp,
                          out=len(conn) - 1 if out is None else out))

    buf("raw", W, H, 1, UI16)
    dpar = mixed_blob([("f", p.strength), ("f", p.luma), ("f", p.detail), ("f", 0.0)]
                      + [("f", e) for e in p.edges] + [("i", 1)])
    rgbc = lambda crop: [("f", v) for v in wb] + [("f", v) for v in black] + [("f", v) for v in white] + \
        [("i", v) for v in crop]
    gm = p.gain_map
    gm_os = [("f", v) for v in gain_map_os(gm)] if gm is not None else [("i", 0)] * 4
    if gm is not None:   # the gain-map source (denoise/main.c:182-188): rgba f32 on the map grid
        buf("gm", gm["gains"].shape[1], gm["gains"].shape[0], 4, F32)
    gmc = ["gm"] if gm is not None else []
    if p.strength <= 0:   # denoise/main.c:202-225: one noop node, crop + black/white scaling
        buf("dn", W, H, 1, F16)
        pc = mixed_blob([("i", v) for v in (0, 0, W, H)] + [("f", v) for v in black] + [("f", v) for v in white]
                        + gm_os + [("i", p.filters), ("i", int(gm is not None))])
        add("denoise", "noop", pc, dpar, ["raw", "dn"] + gmc, W, H, out=1)
        return _build_tail(W, H, p, bufs, nodes, buf, add, wb, max_wd, max_ht, prim, trc)
    buf("half", Qw, Qh, 4, F16)
    add("denoise", "half", mixed_blob(rgbc([0, 0, W, H]) + [("i", p.filters)]), dpar, ["raw", "half"], Qw, Qh)
    for l in range(4):
        buf(f"s{l}", Qw, Qh, 4, F16)
        pc = mixed_blob(rgbc([0, 0, 0, 0]) + [("f", p.noise_a), ("f", p.noise_b), ("i", l), ("i", block)])
        add("denoise", "downcov" if l == 0 else "down", pc, dpar, ["half" if l == 0 else f"s{l-1}", f"s{l}"], Qw, Qh)
    buf("asm", Qw, Qh, 4, F16)
    pc = mixed_blob(rgbc([0, 0, 0, 0]) + [("f", p.noise_a), ("f", p.noise_b), ("i", p.filters)])
    add("denoise", "assemble", pc, dpar, ["half", "s0", "s1", "s2", "s3", "asm"], Qw, Qh)
    buf("dn", W, H, 1, F16)
    pc = mixed_blob(rgbc([0, 0, W, H]) + [("i", p.filters), ("f", p.noise_a), ("f", p.noise_b),
                                           ("i", int(gm is not None))] + gm_os)
    add("denoise", "doub", pc, dpar, ["raw", "asm", "half", "dn"] + gmc, W, H, out=3)
    return _build_tail(W, H, p, bufs, nodes, buf, add, wb, max_wd, max_ht, prim, trc)


def _build_tail(W, H, p, bufs, nodes, buf, add, wb, max_wd, max_ht, prim, trc):
    """the nodes after denoise (its output in buffer "dn")"""
    xt = p.kind == "xtrans"
    block = 3 if xt else 2
    Qw, Qh = W // block, H // block
    mos = "dn"
    if p.hilite:
        mos = "hl"
        add_hilite(W, H, p, buf, add, "dn", "hl")

    dmpc = mixed_blob([("f", v) for v in wb] + [("i", p.filters)])
    dmpar = mixed_blob([("i", p.dm_colour), ("i", p.dm_method)])
    # demosaic's output: half size for method 2 (demosaic/main.c:28-33), then
    # scaled to what the sink asks for; half size again once the scale
    # reaches 1.5 blocks, and a resample to the requested extent
    fw, fh = ((W + 1) // 2, (H + 1) // 2) if p.dm_method == 2 else (W, H)
    ow, oh = export_extent(fw, fh, max_wd, max_ht)
    scale = float(np.float32(W) / np.float32(ow))
    if scale >= 1.5 * block or p.dm_method == 2:
        buf("dmh", Qw, Qh, 4, F16)
        add("demosaic", "halfsize", dmpc, dmpar, [mos, "dmh"], Qw, Qh)
        if float(block) != scale:
            buf("dm", ow, oh, 4, F16)
            add("shared", "resample", None, None, ["dmh", "dm"], ow, oh)
        else:
            bufs["dm"] = bufs.pop("dmh")
            nodes[-1]["conn"] = [mos, "dm"]
    elif p.dm_method == 1:   # RCD (demosaic/main.c:112-155)
        buf("vh", W, H, 1, F16)
        buf("pq", W // 2, H, 1, F16)
        buf("lp", W // 2, H, 1, F16)
        add("demosaic", "rcd_conv", None, dmpar, [mos, "vh", "pq", "lp"], W, H)
        fill_out = "dm" if (ow, oh) == (W, H) else "dmf"
        buf(fill_out, W, H, 4, F16)
        nx, ny = (W + 57) // 58 * 8, (H + 25) // 26 * 8
        add("demosaic", "rcd_fill", dmpc, dmpar, [mos, "vh", "pq", "lp", fill_out], nx, ny)
        if fill_out != "dm":
            buf("dm", ow, oh, 4, F16)
            add("shared", "resample", None, None, ["dmf", "dm"], ow, oh)
    else:
        buf("cov", Qw, Qh, 4, F16)
        add("demosaic", "gauss", dmpc, dmpar, [mos, mos, "cov"], Qw, Qh)
        buf("green", W, H, 1, F16)
        add("demosaic", "splat", dmpc, dmpar, [mos, "cov", "green"], W, H)
        fix_out = "dm" if (ow, oh) == (W, H) else "dmf"
        buf(fix_out, W, H, 4, F16)
        add("demosaic", "fix", dmpc, dmpar, [mos, "green", "cov", fix_out], W, H)
        if fix_out != "dm":
            buf("dm", ow, oh, 4, F16)
            add("shared", "resample", None, None, ["dmf", "dm"], ow, oh)
    W, H = ow, oh      # colour, filmcurv and llap run at the output extent

    buf("co", W, H, 4, F16)
    add("colour", "main", mixed_blob([("i", 0)] * 3), colour_params(p), ["dm", "co"], W, H)
    buf("fc", W, H, 4, F16)
    fpar = mixed_blob([("f", p.light), ("f", p.contrast), ("f", p.bias), ("i", p.fc_colour), ("f", 1.0),
                       ("f", 0), ("f", 0), ("f", 0), ("f", 0), ("f", 0)])
    add("filmcurv", "main", None, fpar, ["co", "fc"], W, H)

    ng = p.num_gamma
    lpar = f32blob([p.sigma, p.shadows, p.hilights, p.clarity])
    nl, lw, lh = llap_levels(W, H, p.max_levels)
    for l in range(nl):
        buf(f"pyr{l}", lw[l], lh[l], 1, F16, ng + 1)
    add("llap", "curve", mixed_blob([("i", ng)]), lpar, ["fc", "pyr0"], W, H)
    for l in range(1, nl):
        add("llap", "reduce", None, lpar, [f"pyr{l-1}", f"pyr{l}"], lw[l], lh[l], ng + 1)
    for l in range(nl - 1, 0, -1):
        buf(f"out{l-1}", lw[l - 1], lh[l - 1], 1, F16)
        coarse = "pyr0" if l == nl - 1 else f"out{l}"
        add("llap", "assemble", mixed_blob([("i", ng), ("i", 1 if l == nl - 1 else 0)]), lpar,
            [coarse, f"pyr{l-1}", f"pyr{l}", f"out{l-1}"], lw[l - 1], lh[l - 1])
    if (prim, trc) == (2, 0):
        buf("final", W, H, 4, F32)
        add("llap", "colour", None, lpar, ["out0", "fc", "final"], W, H)
    else:   # graph-export.c:66-86: llap keeps f16, colenc writes the sink's f32
        buf("lin", W, H, 4, F16)
        add("llap", "colour", None, lpar, ["out0", "fc", "lin"], W, H)
        buf("final", W, H, 4, F32)
        add("colenc", "main", None, mixed_blob([("i", prim), ("i", trc)]), ["lin", "final"], W, H)
    return bufs, nodes


def hilite_levels(Qw, Qh):
    """extents of the hilite pyramid: level 0 the block grid, then one per
    reduce node (hilite/main.c:43-85)"""
    ew, eh = [Qw], [Qh]
    for l in range(1, 15):
        ew.append((ew[-1] - 1) // 2 + 1)
        eh.append((eh[-1] - 1) // 2 + 1)
        if (ew[-1] - 1) // 2 + 1 <= 1 or (eh[-1] - 1) // 2 + 1 <= 1 or l + 1 == 15:
            break
    return ew, eh


def add_hilite(W, H, p, buf, add, src, dst, fmt_out=F16):
    """hilite's nodes from mosaic buffer `src` to `dst` (hilite/main.c:6-86)"""
    block = 3 if p.kind == "xtrans" else 2
    Qw, Qh = W // block, H // block
    pc = mixed_blob([("f", v) for v in p.wb] + [("i", p.filters)])
    par = f32blob([p.hl_white, p.hl_desat, p.hl_soft])
    ew, eh = hilite_levels(Qw, Qh)
    L = len(ew) - 1
    for l in range(L + 1):
        buf(f"hl_l{l}", ew[l], eh[l], 4, F16)
    add("hilite", "half", pc, par, [src, "hl_l0"], Qw, Qh)
    for l in range(1, L + 1):
        add("hilite", "reduce", pc, par, [f"hl_l{l-1}", f"hl_l{l}"], ew[l], eh[l])
    coarse = f"hl_l{L}"
    for l in range(L, 0, -1):
        buf(f"hl_a{l-1}", ew[l - 1], eh[l - 1], 4, F16)
        add("hilite", "assemble", pc, par, [f"hl_l{l-1}", coarse, f"hl_a{l-1}"], ew[l - 1], eh[l - 1])
        coarse = f"hl_a{l-1}"
    buf(dst, W, H, 1, fmt_out)
    add("hilite", "doub", pc, par, [src, coarse, dst], Qw, Qh)


def oracle_params(p):
    """tests.oracle_py.OrDevelop for the same graph."""
    from oracle_py import OrDevelop
    o = OrDevelop()
    o.filters = p.filters
    for k in range(4):
        o.black[k] = p.black
        o.white[k] = p.white
        o.wb[k] = p.wb[k]
        o.dn_edges[k] = p.edges[k]
    cam = np.asarray(p.cam, np.float32).ravel()
    for k in range(9):
        o.cam_to_rec2020[k] = float(cam[k])
    o.noise_a, o.noise_b = p.noise_a, p.noise_b
    o.dn_strength, o.dn_luma, o.dn_detail = p.strength, p.luma, p.detail
    o.dm_colour = p.dm_colour
    o.co_exposure = p.exposure
    for k in range(3):
        o.co_white[k] = p.co_white[k]
    o.fc_light, o.fc_contrast, o.fc_bias, o.fc_colour = p.light, p.contrast, p.bias, p.fc_colour
    o.ll_sigma, o.ll_shadows, o.ll_hilights, o.ll_clarity = p.sigma, p.shadows, p.hilights, p.clarity
    o.ll_num_gamma, o.ll_max_levels = p.num_gamma, p.max_levels
    o.hl_on = 1 if p.hilite else 0
    o.hl_white, o.hl_desat, o.hl_soft = p.hl_white, p.hl_desat, p.hl_soft
    return o
