"""The reference's own module sources, built against this runtime's module API.

vkdt's module sources (src/pipe/modules/<name>/main.c) include
"modules/api.h".  Here that include resolves to a one-line header that
includes vkdt_amd/pipe/api.h, so the reference's main.c builds, unchanged and
from where it lies, into a module DSO of this runtime.  Each such DSO
replaces this repository's build of the module in a copy of the module tree
(the params and connectors files stay the same), and the graph it plans must
equal the graph this repository's module source plans: the same nodes in the
same order, extents, push-constant and uniform bytes, connector names, types,
channels, formats and links, and the same culling.

This checks the drop-in claim of the module API on the reference's code rather
than on this repository's restatement of it.  Only the host side is compared
(dt_graph_plan: no device).  Skipped where the reference tree is absent.
"""
import os
import subprocess

import numpy as np
import pytest

from vkdt_amd import pipe, synth

REF = "/root/reference/src"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "vkdt_amd")

# the modules of the raw-development graph (and of its exports and noise
# profile) whose reference sources build against this API; i-raw needs the
# reference's rawloader (rawler), nprof its core/ log and file helpers
MODULES = ["denoise", "demosaic", "hilite", "colour", "filmcurv", "llap", "o-pfm", "colenc", "rawhist"]

pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "pipe", "modules")),
                                reason="reference tree absent")


def build_tree(tmp):
    """a module tree with the reference's main.c built for MODULES and this
    repository's DSOs elsewhere; returns its base directory"""
    inc = os.path.join(tmp, "inc", "modules")
    os.makedirs(inc)
    with open(os.path.join(inc, "api.h"), "w") as f:
        f.write('#include "pipe/api.h"\n')
    base = os.path.join(tmp, "base")
    for name in sorted(os.listdir(os.path.join(PKG, "modules"))):
        src = os.path.join(PKG, "modules", name)
        if not os.path.isdir(src) or name == "shared":
            continue
        dst = os.path.join(base, "modules", name)
        os.makedirs(dst)
        for fn in ("params", "connectors"):
            if os.path.exists(os.path.join(src, fn)):
                os.symlink(os.path.join(src, fn), os.path.join(dst, fn))
        so = f"lib{name}.so"
        if name in MODULES:
            cmd = ["gcc", "-std=gnu11", "-O1", "-fPIC", "-shared", "-w",
                   "-I" + os.path.join(tmp, "inc"), "-I" + os.path.join(ROOT, "include"), "-I" + PKG,
                   "-idirafter", REF, "-I" + os.path.join(REF, "pipe", "modules", name),
                   os.path.join(REF, "pipe", "modules", name, "main.c"), "-o", os.path.join(dst, so),
                   "-L" + os.path.join(PKG, "lib"), "-lvkdt_pipe", "-Wl,-rpath," + os.path.join(PKG, "lib"), "-lm"]
            r = subprocess.run(cmd, capture_output=True, text=True)
            assert r.returncode == 0, f"{name}: {r.stderr[-2000:]}"
        elif os.path.exists(os.path.join(src, so)):
            os.symlink(os.path.join(src, so), os.path.join(dst, so))
    return base


def signature(g):
    out = []
    for n in g.nodes():
        # a reading connector links to its writer (node, connector); a writing
        # one keeps its readers' bookkeeping there, which is not compared
        conns = [(c["name"], c["type"], c["chan"], c["format"],
                  c["link"] if c["type"] in ("read", "sink") else None,
                  (c["buf"].wd, c["buf"].ht, c["buf"].chan, c["buf"].format, c["buf"].array_length))
                 for c in n["conns"]]
        out.append((n["module"], n["kernel"], n["wd"], n["ht"], n["dp"], n["pc"], n["params"], n["culled"], conns))
    return out


def drop_identity_resample(sig):
    """the graph without the reference demosaic's identity resample.  The
    reference's demosaic ends its full-resolution paths in shared/resample
    whenever its output roi carries a hard marker (demosaic/main.c:145,193),
    which this chain's roi pass sets also at equal extents; this repository's
    demosaic leaves that identity out (weights 0 1 0 0, a full-resolution
    pass for nothing).  Readers of the resample read its input instead and
    later node ids shift down by one."""
    drop = [i for i, n in enumerate(sig) if n[:2] == ("shared", "resample")
            and n[8][0][5][:2] == n[8][1][5][:2] and n[8][0][4][0] >= 0
            and sig[n[8][0][4][0]][0] == "demosaic"]
    if not drop:
        return sig
    assert len(drop) == 1
    r = drop[0]
    src = sig[r][8][0][4]

    def relink(link):
        if link[0] == r:
            link = src
        return (link[0] - (link[0] > r), link[1])

    out = []
    for i, n in enumerate(sig):
        if i == r:
            continue
        conns = [c[:4] + (relink(c[4]) if c[4] is not None and c[4][0] >= 0 else c[4],) + c[5:] for c in n[8]]
        out.append(n[:8] + (conns,))
    return out


def plan_cases(tmp_path):
    cases = {}
    for kind, W, H in (("bayer", 200, 150), ("xtrans", 201, 150)):
        cfa, _ = synth.mosaic(W, H, 1, kind)
        path = str(tmp_path / f"{kind}.dng")
        synth.write_dng(path, cfa, kind, as_shot_neutral=(0.45, 1.0, 0.7))
        na, nb = synth.noise_params(2e-4, 1e-6)
        # (hilite, demosaic method, export width, primaries, transfer, extra cfg lines)
        for hil, method, mw, prim, trc, extra in (
                (False, 0, 0, 2, 0, []), (True, 0, 0, 2, 0, []), (False, 2, 0, 2, 0, []),
                (False, 0, 64, 2, 0, []), (False, 0, 150, 1, 2, []), (True, 0, 80, 4, 3, []),
                (False, 1, 0, 2, 0, []), (False, 0, 0, 2, 0, ["param:llap:01:numgamma:8", "param:llap:01:levels:6"]),
                (False, 0, 0, 2, 0, ["param:colour:01:exposure:0.7", "param:filmcurv:01:colour:3"])):
            if kind == "xtrans" and method == 1:
                continue
            lines = pipe.raw_graph_cfg(path, noise_a=na, noise_b=nb, hilite=hil,
                                       extra=extra + ([f"param:demosaic:01:method:{method}"] if method else []))
            cases[f"{kind}-h{int(hil)}-m{method}-w{mw}-p{prim}t{trc}-{len(extra)}"] = (lines, mw, prim, trc)
        cases[f"{kind}-noise-profile"] = (pipe.noise_profile_cfg(path), None, 2, 0)
        # a DNG with four GainMap opcodes in OpcodeList2: denoise's gain-map source
        rng = np.random.default_rng(2)
        maps = [rng.uniform(0.8, 1.2, size=(4, 5)).astype(np.float32) for _ in range(4)]
        gpath = str(tmp_path / f"{kind}-gm.dng")
        lst = synth.opcode_list(synth.bayer_gain_maps(W, H, maps, spacing=(1 / 3.0, 0.25), origin=(0.0, 0.0)))
        synth.write_dng(gpath, cfa, kind, extra_tags=((51009, 7, list(lst)),))
        for hil in (False, True):
            cases[f"{kind}-gainmap-h{int(hil)}"] = (pipe.raw_graph_cfg(gpath, noise_a=na, noise_b=nb, hilite=hil),
                                                  0, 2, 0)
    return cases


def plan_all(cases, tmp_path):
    res = {}
    for key, (lines, mw, prim, trc) in cases.items():
        g = pipe.Graph(lines=lines)
        try:
            if mw is not None:
                g.export("o-pfm", str(tmp_path / "out"), max_wd=mw, prim=prim, trc=trc)
            g.plan()
            g.commit()
            res[key] = signature(g)
        finally:
            g.close()
    return res


def test_reference_module_sources_plan_the_same_graphs(tmp_path):
    cases = plan_cases(tmp_path)
    ours = plan_all(cases, tmp_path)
    L = pipe.lib()
    base = build_tree(str(tmp_path / "tree"))
    assert L.dt_pipe_global_init(base.encode()) == 0
    try:
        theirs = plan_all(cases, tmp_path)
    finally:
        assert L.dt_pipe_global_init(pipe.PKG_DIR.encode()) == 0
    for key in cases:
        a, b = ours[key], drop_identity_resample(theirs[key])
        assert drop_identity_resample(a) == a, key
        if "gainmap" in key:
            assert any(n[:2] == ("denoise", "gainmap") for n in a), key
        assert [n[:2] for n in a] == [n[:2] for n in b], key
        for na, nb in zip(a, b):
            what = f"{key}: {na[0]}/{na[1]}"
            assert na[2:5] == nb[2:5], what
            assert na[5] == nb[5], what + " push constants"
            if na[0] == "colour":
                # word 224 is the clut anchor of the Kelvin slider, from tanf/asinhf
                # of the temperature: the same formula, the host's libm rounding
                assert na[6][:224 * 4] == nb[6][:224 * 4] and na[6][225 * 4:] == nb[6][225 * 4:], what
            else:
                assert na[6] == nb[6], what + " uniforms"
            assert na[7:] == nb[7:], what + " culling / connectors"
