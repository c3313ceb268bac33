"""Planning (no device) of the graphs the export, video and reduction modules
build, against the reference's main.c node layouts: o-jpg behind a colenc
writing 8-bit UNORM; autoexp's three nodes with a cleared u32 histogram of
17 x 16 and a protected f32 luminance; hist's collect over the input and
its 1000 x 600 map; i-mlv publishing the clip's frame count and rate."""
import numpy as np

from vkdt_amd import pipe, synth


def dng(tmp_path, W=96, H=64):
    cfa, _ = synth.mosaic(W, H, 3, "bayer")
    p = str(tmp_path / "a.dng")
    synth.write_dng(p, cfa, "bayer")
    return p


def nodes(g, module):
    return [n for n in g.nodes() if n["module"] == module and not n["culled"]]


def test_o_jpg_behind_8bit_colenc(tmp_path):
    g = pipe.Graph(lines=pipe.raw_graph_cfg(dng(tmp_path)))
    g.export("o-jpg", str(tmp_path / "o"), prim=1, trc=2)
    g.plan()
    (ce,) = nodes(g, "colenc")
    assert ce["conns"][1]["format"] == "ui8" and ce["conns"][1]["buf"].format == 6
    (sink,) = nodes(g, "o-jpg")
    assert sink["conns"][0]["buf"].chan == 4
    g.close()


def with_module(path, mod):
    lines = [l for l in pipe.raw_graph_cfg(path) if l != "connect:llap:01:output:display:main:input"]
    return lines + [f"module:{mod}:01", f"connect:llap:01:output:{mod}:01:input",
                    f"connect:{mod}:01:output:display:main:input"]


def test_autoexp_nodes(tmp_path):
    g = pipe.Graph(lines=with_module(dng(tmp_path), "autoexp"))
    g.export("o-pfm", str(tmp_path / "o"))
    g.plan()
    ks = {n["kernel"]: n for n in nodes(g, "autoexp")}
    assert set(ks) == {"hist", "lum", "exp"}
    assert (ks["hist"]["wd"], ks["hist"]["ht"]) == ((96 + 15) // 16 * 8, (64 + 15) // 16 * 8)
    assert (ks["lum"]["wd"], ks["lum"]["ht"]) == (1, 1)
    h, l = ks["hist"]["conns"][1]["buf"], ks["lum"]["conns"][2]["buf"]
    assert (h.wd, h.ht, h.format) == (17, 16, 5) and (l.wd, l.ht, l.format) == (8, 8, 3)
    g.close()


def test_hist_nodes(tmp_path):
    g = pipe.Graph(lines=with_module(dng(tmp_path), "hist"))
    g.export("o-pfm", str(tmp_path / "o"))
    g.plan()
    ks = {n["kernel"]: n for n in nodes(g, "hist")}
    assert (ks["collect"]["wd"], ks["collect"]["ht"]) == (96, 64)
    c = ks["collect"]["conns"][1]["buf"]
    assert (c.wd, c.ht, c.format) == (1000, 600, 5)
    assert (ks["map"]["wd"], ks["map"]["ht"]) == (1000, 600)
    g.close()


def test_mlv_frames(tmp_path):
    frames = [np.full((32, 48), 1000 * (k + 1), np.uint16) for k in range(5)]
    p = str(tmp_path / "c.MLV")
    synth.write_mlv(p, frames, fps=(30000, 1001))
    g = pipe.Graph(lines=pipe.mlv_graph_cfg(p))
    g.export("o-pfm", str(tmp_path / "o"))
    g.plan()
    assert g.frame_cnt() == 5 and abs(g.frame_rate() - 29.97) < 1e-2
    g.close()
