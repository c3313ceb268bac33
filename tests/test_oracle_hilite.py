"""Pins of the oracle's hilite (highlight inpainting on the mosaic,
src/pipe/modules/hilite) against what the algorithm fixes in closed form:

  * below the soft threshold 0.97*white nothing changes: the module is the
    identity on the mosaic, for f16 and for UNORM16 buffers (doub.comp:105-113
    only mixes when the block maximum exceeds 0.97*white);
  * a constant colour with one clipped channel is reconstructed from the
    unclipped ratios: every clipped coarse level propagates 1.0
    (reduce.comp:56), the finest assemble rescales that grey to the mean of
    the fine texel's channels, and doub writes it back -- Bayer: all four
    sites of a block (one smoothstep of the block maximum), X-Trans: the
    red/blue sites only (greens use the smoothstep of the green minimum);
  * half's channel assignment (gather order, x-trans block parity);
  * reduce of a grey ramp: the binomial mean is the centre value and the
    edge magnitude is 8 * slope * (sum of the rec2020 luminance weights);
  * assemble's expansion reproduces linear functions, so a fine level that
    equals the expanded coarse level is left unchanged.
"""
import numpy as np
import pytest

import graphspec as gs
import oracle_py as op
from vkdt_amd import synth

LUMA_SUM = 2.62700212e-01 + 6.77998072e-01 + 5.93017165e-02


def hl_pc(filters, wb=(1.0, 1.0, 1.0, 1.0)):
    return gs.mixed_blob([("f", v) for v in wb] + [("i", filters)])


def hl_par(white=0.985, desat=0.3, soft=0.6):
    return gs.f32blob([white, desat, soft])


def run_chain(bits, mos, kind, fmt=gs.F16, white=0.985, desat=0.3, soft=0.6, wb=(1.0, 1.0, 1.0, 1.0)):
    """the whole hilite node chain on the oracle; mos is float (f16 values)
    or uint16 (UNORM16); returns the output mosaic in the same format"""
    H, W = mos.shape
    p = gs.DevelopParams(kind=kind, wb=wb, hl_white=white, hl_desat=desat, hl_soft=soft)
    bufs, nodes = {}, []
    bufs["in"] = (W, H, 1, fmt, 1)
    gs.add_hilite(W, H, p, lambda n, w, h, c, f, pl=1: bufs.__setitem__(n, (w, h, c, f, pl)),
                  lambda m, k, pc, par, conn, w, h, dp=1: nodes.append((m, k, pc, par, conn, w, h, dp)),
                  "in", "out", fmt_out=fmt)
    data = {}
    for name, (w, h, c, f, pl) in bufs.items():
        if f == gs.UI16:
            data[name] = np.zeros((h, w), np.uint16)
        else:
            data[name] = np.zeros((h, w, c) if c == 4 else (h, w), np.float64)
    data["in"][...] = mos
    orf = {gs.UI16: op.OR_UI16, gs.F16: op.OR_F16, gs.F32: op.OR_F32}
    for m, k, pc, par, conn, w, h, dp in nodes:
        ob = [op.orbuf(data[c], orf[bufs[c][3]], bufs[c][2]) for c in conn]
        op.node(bits, m, k, pc, par, ob, w, h, dp)
    return data["out"], data


def f16(x):
    return np.float16(x).astype(np.float64)


@pytest.mark.parametrize("kind,W,H", [("bayer", 64, 48), ("xtrans", 66, 48)])
def test_identity_below_the_soft_threshold(kind, W, H):
    rng = np.random.default_rng(3)
    mos = f16(rng.uniform(0.0, 0.9, size=(H, W)))
    out, _ = run_chain(64, mos, kind)
    np.testing.assert_array_equal(out, mos)
    raw = rng.integers(0, int(0.9 * 65535), size=(H, W), dtype=np.uint16)
    out, _ = run_chain(64, raw, kind, fmt=gs.UI16)
    np.testing.assert_array_equal(out, raw)


@pytest.mark.parametrize("kind,W,H", [("bayer", 64, 48), ("xtrans", 66, 48)])
def test_clipped_red_reconstructed_from_ratios(kind, W, H):
    r, g, b = 1.0, 0.5, 0.25
    cm = synth.channel_map(W, H, kind)
    mos = np.where(cm == 0, r, np.where(cm == 1, g, b)).astype(np.float64)
    out, data = run_chain(64, mos, kind)
    # every reduce level is fully clipped: rgb 1, no edge
    for name, arr in data.items():
        if name.startswith("hl_l") and name != "hl_l0":
            np.testing.assert_array_equal(arr[..., :3], 1.0)
            np.testing.assert_array_equal(arr[..., 3], 0.0)
    grey = f16((r + g + b) / 3.0)
    np.testing.assert_allclose(data["hl_a0"][..., :3], grey, atol=1e-3)
    if kind == "bayer":
        np.testing.assert_allclose(out, grey, atol=1e-3)
    else:
        np.testing.assert_allclose(out[cm != 1], grey, atol=1e-3)
        np.testing.assert_array_equal(out[cm == 1], g)


def test_half_channels_bayer_and_xtrans_parity():
    rng = np.random.default_rng(8)
    for kind, W, H in (("bayer", 16, 12), ("xtrans", 18, 12)):
        cm = synth.channel_map(W, H, kind)
        mos = f16(rng.uniform(0.05, 0.9, size=(H, W)))
        blk = 3 if kind == "xtrans" else 2
        out = np.zeros((H // blk, W // blk, 4))
        op.node(64, "hilite", "half", hl_pc(9 if kind == "xtrans" else gs.synth.BAYER_FILTERS), hl_par(),
                [op.orbuf(mos, op.OR_F16, 1), op.orbuf(out, op.OR_F16, 4)], W // blk, H // blk)
        for by in range(H // blk):
            for bx in range(W // blk):
                v = mos[blk * by:blk * by + blk, blk * bx:blk * bx + blk]
                c = cm[blk * by:blk * by + blk, blk * bx:blk * bx + blk]
                ref = [v[c == 0].mean(), v[c == 1].mean(), v[c == 2].mean(), 1.0]
                np.testing.assert_allclose(out[by, bx], f16(np.array(ref)), atol=1e-3, err_msg=kind)


def test_half_bayer_clipped_green_takes_the_other():
    W, H = 4, 4
    mos = np.full((H, W), 0.25)
    mos[1, 0] = 0.99       # green of block (0,0) at (0,1): clipped
    mos[0, 1] = 0.5
    out = np.zeros((2, 2, 4))
    op.node(64, "hilite", "half", hl_pc(gs.synth.BAYER_FILTERS), hl_par(),
            [op.orbuf(mos, op.OR_F16, 1), op.orbuf(out, op.OR_F16, 4)], 2, 2)
    assert out[0, 0, 1] == 0.5


def test_reduce_grey_ramp_mean_and_edge():
    W, H, k = 32, 16, 1.0 / 64.0
    ramp = np.arange(W) * k
    fine = np.zeros((H, W, 4))
    fine[..., 0] = fine[..., 1] = fine[..., 2] = ramp[None, :]
    fine[..., 3] = 1.0
    cw, ch = (W - 1) // 2 + 1, (H - 1) // 2 + 1
    out = np.zeros((ch, cw, 4))
    op.node(64, "hilite", "reduce", hl_pc(gs.synth.BAYER_FILTERS), hl_par(white=2.0),
            [op.orbuf(fine, op.OR_F32, 4), op.orbuf(out, op.OR_F32, 4)], cw, ch, storage_f16=False)
    inner = slice(1, cw - 1)
    np.testing.assert_allclose(out[:, inner, 0], np.broadcast_to((2 * np.arange(cw) * k)[inner], (ch, cw - 2)),
                               atol=1e-12)
    np.testing.assert_allclose(out[:, inner, 3], 8 * k * LUMA_SUM, rtol=1e-6)


def test_assemble_keeps_a_fine_level_equal_to_the_expanded_coarse():
    cw, ch, k = 16, 10, 1.0 / 32.0
    W, H = 2 * cw - 1, 2 * ch - 1          # the coarse level is the ceil-half of the fine one
    coarse = np.zeros((ch, cw, 4))
    coarse[..., 0] = (0.1 + k * np.arange(cw))[None, :]
    coarse[..., 1] = 0.5
    coarse[..., 2] = 0.25
    fine = np.zeros((H, W, 4))
    fine[..., 0] = (0.1 + k * np.arange(W) / 2.0)[None, :]
    fine[..., 1] = 0.5
    fine[..., 2] = 0.25
    out = np.zeros((H, W, 4))
    op.node(64, "hilite", "assemble", hl_pc(gs.synth.BAYER_FILTERS), hl_par(),
            [op.orbuf(fine, op.OR_F32, 4), op.orbuf(coarse, op.OR_F32, 4), op.orbuf(out, op.OR_F32, 4)],
            W, H, storage_f16=False)
    # away from the mirrored border the expansion is exact linear interpolation
    np.testing.assert_allclose(out[1:-1, 2:-2, :3], fine[1:-1, 2:-2, :3], atol=1e-12)
    np.testing.assert_array_equal(out[..., 3], 1.0)


def test_doub_only_writes_inside_the_output():
    """a ragged mosaic (width not a multiple of the block): the trailing
    column is written by the threads of the last 8x8 group past the node
    extent, reading zero coarse texels (doub.comp:24-25 bounds by the output)"""
    W, H = 9, 6
    mos = np.full((H, W), 0.3)
    coarse = np.zeros((H // 2, W // 2, 4))
    out = np.full((H, W), -7.0)
    op.node(64, "hilite", "doub", hl_pc(gs.synth.BAYER_FILTERS), hl_par(),
            [op.orbuf(mos, op.OR_F16, 1), op.orbuf(coarse, op.OR_F16, 4), op.orbuf(out, op.OR_F16, 1)],
            W // 2, H // 2)
    np.testing.assert_array_equal(out, f16(0.3))


def test_float_and_double_oracles_agree_on_the_chain():
    W, H = 96, 64
    cfa, _ = synth.mosaic(W, H, 4, "bayer")
    mos = f16(np.clip((cfa.astype(np.float64) - 512) / (16383 - 512), 0, None))
    o64, _ = run_chain(64, mos, "bayer", wb=(2.0, 1.0, 1.5, 1.0))
    o32, _ = run_chain(32, mos, "bayer", wb=(2.0, 1.0, 1.5, 1.0))
    assert (mos >= 0.985).mean() > 0.005        # the scene clips
    changed = np.mean(o64 != mos)
    assert changed > 0.005
    d = np.abs(o64 - o32)
    assert np.mean(d > 1e-3) < 1e-3 and d.max() < 0.05
