"""The batched-develop descriptor read off a planned graph (no device work):
every field equals what the per-node graph's nodes receive, graphs outside
the raw-development chain are refused, and the scratch size follows the
buffer layout."""
import os

import numpy as np
import pytest

from vkdt_amd import batch, hip, pipe, synth


def planned(tmp_path, kind, W, H, extra=(), **kw):
    cfa, _ = synth.mosaic(W, H, 7, kind)
    dng = str(tmp_path / f"{kind}_{W}x{H}.dng")
    synth.write_dng(dng, cfa, kind)
    na, nb = synth.noise_params(2e-4, 1e-6)
    g = pipe.Graph(lines=pipe.raw_graph_cfg(dng, noise_a=na, noise_b=nb, extra=list(extra), **kw))
    g.export("o-pfm", str(tmp_path / "out"))
    g.plan()
    return g, cfa


def node(g, module, kernel):
    return [n for n in g.nodes() if n["module"] == module and n["kernel"] == kernel]


@pytest.mark.parametrize("kind,W,H", [("bayer", 320, 240), ("xtrans", 330, 246)])
def test_descriptor_matches_nodes(tmp_path, kind, W, H):
    g, cfa = planned(tmp_path, kind, W, H, extra=["param:demosaic:01:colour:1", "param:llap:01:shadows:0.8"],
                     strength=0.7, clarity=0.35, numgamma=7, levels=5)
    d = batch.describe(g)
    doub = node(g, "denoise", "doub")[0]
    pc = np.frombuffer(doub["pc"], np.uint32)
    pf = np.frombuffer(doub["pc"], np.float32)
    assert (d.wd, d.ht) == (doub["wd"], doub["ht"])
    assert (d.cfa_wd, d.cfa_ht) == (cfa.shape[1], cfa.shape[0])
    assert d.filters == pc[16] and (d.filters == 9) == (kind == "xtrans")
    np.testing.assert_array_equal(np.array(d.black, np.float32), pf[4:8])
    np.testing.assert_array_equal(np.array(d.white, np.float32), pf[8:12])
    assert list(d.crop) == list(pc[12:16].astype(np.int32))
    assert np.float32(d.noise_a) == pf[17] and np.float32(d.noise_b) == pf[18]
    assert d.dn_params[0] == np.float32(0.7)
    assert d.dm_fixup == 1
    col = node(g, "colour", "main")[0]
    np.testing.assert_array_equal(np.array(d.co_params, np.float32), np.frombuffer(col["params"], np.float32)[:242])
    assert d.ll_numgamma == 7
    assert d.ll_levels == len(node(g, "llap", "reduce")) + 1 == 5
    assert np.isclose(d.ll_params[3], 0.35) and np.isclose(d.ll_params[1], 0.8)
    g.close()


def test_scratch_bytes_follow_the_layout(tmp_path):
    g, _ = planned(tmp_path, "bayer", 320, 240)
    d = batch.describe(g)
    a = lambda v: (v + 255) // 256 * 256
    p8 = lambda w: (w + 7) // 8 * 8
    W, H, ng, nl = 320, 240, d.ll_numgamma, d.ll_levels
    qw, qh = W // 2, H // 2
    lw, lh = [W], [H]
    for _ in range(1, nl):
        lw.append((lw[-1] - 1) // 2 + 1)
        lh.append((lh[-1] - 1) // 2 + 1)
    want = 6 * a(qw * qh * 8) + a(p8(W) * H * 2) + a(qw * qh * 8) + a(W * H * 8)
    want += sum(a((ng + 1) * p8(lw[l]) * lh[l] * 2) for l in range(1, nl))
    want += sum(a(p8(lw[l]) * lh[l] * 2) for l in range(1, nl - 1))
    # plus the device tables (image records, job lists, tile counters)
    one, two = batch.scratch_bytes([d]), batch.scratch_bytes([d, d])
    assert want < one <= want + 64 * 1024
    assert 2 * want < two <= 2 * want + 128 * 1024
    g.close()


def test_refuses_graphs_the_batch_does_not_run(tmp_path):
    g, _ = planned(tmp_path, "bayer", 128, 96, strength=0.0)
    with pytest.raises(pipe.PipeError, match="mosaic wavelet"):
        batch.describe(g)
    g.close()
    # a second colour module inserted before filmcurv breaks the chain
    cfa, _ = synth.mosaic(128, 96, 7, "bayer")
    dng = str(tmp_path / "c.dng")
    synth.write_dng(dng, cfa, "bayer")
    lines = [l for l in pipe.raw_graph_cfg(dng) if not l.startswith("connect:colour:01:output:filmcurv")]
    lines += ["module:colour:02", "connect:colour:01:output:colour:02:input",
              "connect:colour:02:output:filmcurv:01:input"]
    g = pipe.Graph(lines=lines)
    g.export("o-pfm", str(tmp_path / "out2"))
    g.plan()
    with pytest.raises(pipe.PipeError, match="chain"):
        batch.describe(g)
    g.close()
    # filmcurv's colour modes other than 1 and 3 run on the per-node graph only
    for mode in (0, 2, 4, 5):
        g = pipe.Graph(lines=pipe.raw_graph_cfg(dng) + [f"param:filmcurv:01:colour:{mode}"])
        g.export("o-pfm", str(tmp_path / "out3"))
        g.plan()
        with pytest.raises(pipe.PipeError, match="colour modes 1 and 3"):
            batch.describe(g)
        g.close()


@pytest.mark.parametrize("kind,W,H", [("bayer", 128, 96), ("xtrans", 198, 132)])
def test_hilite_between_denoise_and_demosaic(tmp_path, kind, W, H):
    """hilite's push constants, params and coarse levels are read off its
    nodes (hilite/main.c); the scratch grows by its buffers"""
    g, _ = planned(tmp_path, kind, W, H, hilite=True, extra=["param:hilite:01:desat:0.5"])
    d = batch.describe(g)
    half = node(g, "hilite", "half")[0]
    pf = np.frombuffer(half["pc"], np.float32)
    assert d.hl_on == 1
    np.testing.assert_array_equal(np.array(d.hl_wb, np.float32), pf[:4])
    np.testing.assert_array_equal(np.array(d.hl_params, np.float32), np.float32([0.985, 0.5, 0.6]))
    assert d.hl_levels == len(node(g, "hilite", "reduce")) == len(node(g, "hilite", "assemble"))
    b = 3 if kind == "xtrans" else 2
    assert (half["wd"], half["ht"]) == (W // b, H // b)
    a = lambda v: (v + 255) // 256 * 256
    hw, hh = [W // b], [H // b]
    for _ in range(d.hl_levels):
        hw.append((hw[-1] - 1) // 2 + 1)
        hh.append((hh[-1] - 1) // 2 + 1)
    extra = a((W + 7) // 8 * 8 * H * 2) + sum(a(hw[l] * hh[l] * 8) for l in range(d.hl_levels + 1))
    extra += sum(a(hw[l] * hh[l] * 8) for l in range(d.hl_levels))
    with_hl = batch.scratch_bytes([d])
    d.hl_on = 0
    assert batch.scratch_bytes([d]) + extra == with_hl
    d.hl_on, d.hl_levels = 1, d.hl_levels + 1
    with pytest.raises(hip.HipError, match="hilite levels"):
        batch.scratch_bytes([d])
    g.close()


def test_create_rejects_bad_input_without_touching_the_device(tmp_path):
    g, _ = planned(tmp_path, "bayer", 128, 96)
    d = batch.describe(g)
    with pytest.raises(hip.HipError, match="no mosaic or output"):
        batch.Batch([d], None, 0)
    d.cfa, d.out = 256, 512                      # never dereferenced: the scratch check fails first
    with pytest.raises(hip.HipError, match="scratch"):
        batch.Batch([d], 4096, 16)
    d.ll_numgamma = 40
    with pytest.raises(hip.HipError, match="num_gamma"):
        batch.scratch_bytes([d])
    g.close()
