"""The host runtime on the CPU: cfg parsing, module -> node expansion, roi
passes, culling and the push-constant / uniform blobs every node hands to
the HIP C ABI, compared with an independent restatement of the modules'
create_nodes (graphspec).  No device is touched (dt_graph_plan)."""
import numpy as np
import pytest

import graphspec as gs
from vkdt_amd import pipe, synth


def planned(tmp_path, kind, W, H, extra=(), hilite=False, max_wd=0, max_ht=0, prim=2, trc=0):
    cfa, _ = synth.mosaic(W, H, 1, kind)
    path = str(tmp_path / "p.dng")
    synth.write_dng(path, cfa, kind, as_shot_neutral=(0.45, 1.0, 0.7))
    na, nb = synth.noise_params(2e-4, 1e-6)
    g = pipe.Graph(lines=pipe.raw_graph_cfg(path, noise_a=na, noise_b=nb, extra=extra, hilite=hilite))
    g.export("o-pfm", str(tmp_path / "out"), max_wd=max_wd, max_ht=max_ht, prim=prim, trc=trc)
    g.plan()
    g.commit()
    return g, np.float32(na), np.float32(nb)


# (kind, W, H, hilite, demosaic method, export width, export height)
PLAN_CASES = [("bayer", 200, 150, False, 0, 0, 0), ("xtrans", 201, 150, False, 0, 0, 0),
              ("bayer", 200, 150, True, 0, 0, 0), ("xtrans", 201, 150, True, 0, 0, 0),
              ("bayer", 200, 150, False, 0, 64, 0),     # scale 3.1: half size + resample
              ("bayer", 200, 150, False, 0, 100, 100),  # scale 2: half size on the block grid
              ("bayer", 200, 150, False, 0, 150, 0),    # scale 1.3: full demosaic + resample
              ("xtrans", 201, 150, False, 0, 0, 40),    # scale 3.7 < 4.5: full + resample
              ("xtrans", 201, 150, False, 2, 0, 0),     # method 2: half size + resample
              ("bayer", 200, 150, False, 2, 0, 0),      # method 2 on the block grid
              ("bayer", 200, 150, False, 0, 0, 0, 1, 2),    # sRGB export: colenc
              ("bayer", 200, 150, False, 1, 0, 0),      # RCD
              ("bayer", 200, 150, True, 1, 150, 0),     # RCD behind hilite, scaled: + resample
              ("xtrans", 201, 150, True, 0, 80, 0, 4, 3)]   # P3 + PQ, hilite, scaled


@pytest.mark.parametrize("case", PLAN_CASES, ids=[str(i) for i in range(len(PLAN_CASES))])
def test_nodes_match_restatement(tmp_path, case):
    kind, W, H, hilite, method, max_wd, max_ht = case[:7]
    prim, trc = case[7:] if len(case) > 7 else (2, 0)
    extra = [f"param:demosaic:01:method:{method}"] if method else []
    g, na, nb = planned(tmp_path, kind, W, H, hilite=hilite, extra=extra, max_wd=max_wd, max_ht=max_ht,
                        prim=prim, trc=trc)
    ip = g.img_param("i-raw", "main")
    W, H = ip["full"]
    p = gs.DevelopParams(kind=kind, noise_a=na, noise_b=nb, cam=np.array(ip["cam"], np.float32).reshape(3, 3),
                         wb=tuple(ip["wb"][:3]) + (ip["wb"][3],), hilite=hilite, dm_method=method)
    _, ref = gs.build(W, H, p, max_wd, max_ht, prim, trc)
    nodes = {n["id"]: n for n in g.nodes()}
    live = [nodes[k] for k in g.exec_nodes() if nodes[k]["module"] not in ("i-raw", "o-pfm")]
    assert [(n["module"], n["kernel"]) for n in live] == [(r["module"], r["kernel"]) for r in ref]
    for n, r in zip(live, ref):
        what = f"{n['module']}/{n['kernel']}"
        assert (n["wd"], n["ht"], n["dp"]) == (r["wd"], r["ht"], r["dp"]), what
        if r["pc"] is not None:
            rb = np.ascontiguousarray(r["pc"]).tobytes()
            assert n["pc"][:len(rb)] == rb, what
        if r["params"] is not None:
            rb = np.ascontiguousarray(r["params"]).tobytes()
            nb = n["params"][:len(rb)]
            if n["module"] == "colour":
                # word 224 is the clut anchor of the Kelvin slider (tanf/asinhf of the
                # temp param), unused by the matrix path: not restated
                rb, nb = rb[:224 * 4] + rb[225 * 4:], nb[:224 * 4] + nb[225 * 4:]
            assert nb == rb, what


def test_culling_and_dead_stores(tmp_path):
    g, _, _ = planned(tmp_path, "bayer", 64, 48)
    nodes = g.nodes()
    culled = {(n["module"], n["kernel"]) for n in nodes if n["culled"]}
    # filmcurv's histogram and curve display feed nothing; demosaic/down feeds an unsampled binding
    assert culled == {("OpenDRT", "hist"), ("filmcurv", "dspy"), ("demosaic", "down")}
    dc = [n for n in nodes if n["kernel"] == "downcov"][0]
    assert dc["conns"][2]["name"] == "cov"
    # llap pyramid (llap/main.c:57-97): a level is added while the one after it
    # stays wider and higher than 1 texel
    red = [n for n in nodes if n["kernel"] == "reduce"]
    assert [(n["wd"], n["ht"]) for n in red] == [(32, 24), (16, 12), (8, 6), (4, 3), (2, 2)]
    assert all(n["dp"] == 11 for n in red)


def test_llap_parameters_change_the_pyramid(tmp_path):
    g, _, _ = planned(tmp_path, "xtrans", 300, 240, extra=["param:llap:01:numgamma:8", "param:llap:01:levels:6"])
    red = [n for n in g.nodes() if n["kernel"] == "reduce"]
    assert len(red) == 5 and all(n["dp"] == 9 for n in red)
    asm = [n for n in g.nodes() if n["module"] == "llap" and n["kernel"] == "assemble"]
    first = [np.frombuffer(n["pc"], np.int32)[1] for n in asm]
    assert sum(first) == 1 and np.frombuffer(asm[-1]["pc"], np.int32)[1] == 1


def test_sink_is_f32_and_llap_output_follows(tmp_path):
    g, _, _ = planned(tmp_path, "bayer", 64, 48)
    col = [n for n in g.nodes() if n["module"] == "llap" and n["kernel"] == "colour"][0]
    assert col["conns"][2]["format"] == "f32"


def test_errors_are_reported(tmp_path):
    with pytest.raises(pipe.PipeError):
        pipe.Graph(lines=["module:nosuchmod:01"])
    # missing raw file: the sink has no size
    g = pipe.Graph(lines=pipe.raw_graph_cfg(str(tmp_path / "missing.dng")))
    g.export("o-pfm", str(tmp_path / "o"))
    with pytest.raises(pipe.PipeError, match="uninited size"):
        g.plan()
    # unconnected input
    cfa = np.zeros((32, 32), np.uint16)
    path = str(tmp_path / "z.dng")
    synth.write_dng(path, cfa)
    lines = [l for l in pipe.raw_graph_cfg(path) if "demosaic:01:output" not in l]
    lines.append("connect:demosaic:01:output:display:main:input")   # colour... left dangling
    g2 = pipe.Graph(lines=lines)
    g2.export("o-pfm", str(tmp_path / "o2"))
    g2.plan()   # colour/filmcurv/llap are not reachable from the sink: fine
    names = {n["module"] for n in g2.nodes()}
    assert "llap" not in names
    # a format mismatch is a connect error (11)
    g3 = pipe.Graph(lines=["module:i-raw:main", "module:llap:01",
                           "connect:i-raw:main:output:llap:01:input"])
    assert "11" in g3.error() or "failed" in g3.error()


def test_cfg_file_round(tmp_path):
    """a .cfg on disk with comments, the default-darkroom style module lines
    with gui coordinates, and params with spaces in their names"""
    cfa, _ = synth.mosaic(64, 48, 2, "bayer")
    dng = str(tmp_path / "c.dng")
    synth.write_dng(dng, cfa)
    lines = pipe.raw_graph_cfg("c.dng", noise_a=1e-7, noise_b=2e-5)
    lines = [l + ":43:400" if l.startswith("module:") else l for l in lines]
    cfg = tmp_path / "g.cfg"
    cfg.write_text("# a comment\n" + "\n".join(lines) + "\n")
    g = pipe.Graph(cfg=str(cfg))
    g.export("o-pfm", str(tmp_path / "o"))
    g.plan()
    ip = g.img_param("i-raw", "main")
    assert ip["full"] == (64, 48)
    assert np.allclose(ip["noise"], [1e-7, 2e-5], rtol=1e-6)


def test_hilite_formats_follow_the_module_connectors(tmp_path):
    """hilite's half/doub declare ui16 mosaics but dt_connector_copy gives them
    the module's negotiated format: f16 behind denoise (api.h:88-121)"""
    g, _, _ = planned(tmp_path, "bayer", 96, 64, hilite=True)
    nodes = g.nodes()
    half = [n for n in nodes if n["module"] == "hilite" and n["kernel"] == "half"][0]
    doub = [n for n in nodes if n["module"] == "hilite" and n["kernel"] == "doub"][0]
    assert half["conns"][0]["format"] == "f16"
    assert doub["conns"][0]["format"] == "f16" and doub["conns"][2]["format"] == "f16"
    red = [n for n in nodes if n["module"] == "hilite" and n["kernel"] == "reduce"]
    # block grid 48x32: levels 24x16 ... until the next would be 1 texel
    assert [(n["wd"], n["ht"]) for n in red] == [(24, 16), (12, 8), (6, 4), (3, 2)]


def test_noise_profile_graph_plan(tmp_path):
    """i-raw -> rawhist -> nprof (bin/presets/noise-profile.pst): collect runs on
    the block grid into a cleared 1000x5 accumulator; the plot feeds nothing"""
    cfa, _ = synth.mosaic(200, 150, 1, "bayer")
    path = str(tmp_path / "n.dng")
    synth.write_dng(path, cfa, "bayer")
    g = pipe.Graph(lines=pipe.noise_profile_cfg(path))
    g.plan()
    nodes = g.nodes()
    col = [n for n in nodes if n["module"] == "rawhist" and n["kernel"] == "collect"][0]
    assert (col["wd"], col["ht"]) == (100, 75) and not col["culled"]
    assert col["conns"][1]["format"] == "atom"
    assert [n["culled"] for n in nodes if n["kernel"] == "map"] == [True]
    g.commit()
    col = [n for n in g.nodes() if n["module"] == "rawhist" and n["kernel"] == "collect"][0]
    par = np.frombuffer(col["params"], np.uint32)
    assert par[2] == synth.BAYER_FILTERS and par[3] == 512 and par[4] == 16383


@pytest.mark.parametrize("kind,W,H", [("bayer", 200, 150), ("xtrans", 201, 150)])
def test_strength_zero_is_one_noop(tmp_path, kind, W, H):
    """denoise strength 0 (the param's default) plans one noop node
    (denoise/main.c:202-225) writing the one-channel f16 mosaic demosaic reads"""
    g, na, nb = planned(tmp_path, kind, W, H, extra=["param:denoise:01:strength:0"])
    ip = g.img_param("i-raw", "main")
    W, H = ip["full"]
    p = gs.DevelopParams(kind=kind, noise_a=na, noise_b=nb, cam=np.array(ip["cam"], np.float32).reshape(3, 3),
                         wb=tuple(ip["wb"][:3]) + (ip["wb"][3],), strength=0.0)
    _, ref = gs.build(W, H, p)
    nodes = {n["id"]: n for n in g.nodes()}
    live = [nodes[k] for k in g.exec_nodes() if nodes[k]["module"] not in ("i-raw", "o-pfm")]
    assert [(n["module"], n["kernel"]) for n in live] == [(r["module"], r["kernel"]) for r in ref]
    noop = live[0]
    assert (noop["module"], noop["kernel"]) == ("denoise", "noop")
    assert noop["pc"][:72] == np.ascontiguousarray(ref[0]["pc"]).tobytes()
    out = noop["conns"][1]["buf"]
    assert (out.chan, out.format, out.wd, out.ht) == (1, gs.F16, W, H)
