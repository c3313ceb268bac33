"""DNG gain maps (OpcodeList2 GainMap opcodes) on the CPU: the raw input's
opcode-list decode, the denoise module's gain-map source node and push
constants, its upload, the acceptance rules, and the oracle's gain
application pinned by closed forms.

Reference: i-raw/main.c:60-92 and dng_opcode_decode.c (decode),
denoise/main.c:12-78,172-200 (acceptance, source node, upload),
noop.comp:45-55 and doub.comp:106-114 (application).  The reference ships
no DNG with opcode lists, so the files here are written by
synth.gain_map_opcode from the DNG specification's record layout.
"""
import ctypes as C
import struct

import numpy as np
import pytest

import graphspec as gs
import oracle_py as op
from vkdt_amd import pipe, synth

BAYER = synth.BAYER_FILTERS


def ramp_maps(mh, mw, seed=3):
    rng = np.random.default_rng(seed)
    return [rng.uniform(0.8, 1.3, size=(mh, mw)).astype(np.float32) for _ in range(4)]


def write(tmp_path, W=64, H=48, maps=None, ops=None, pattern=None, kind="bayer", name="g.dng", **kw):
    cfa, _ = synth.mosaic(W, H, 1, kind)
    if ops is None:
        ops = synth.bayer_gain_maps(W, H, maps, **kw)
    lst = synth.opcode_list(ops)
    path = str(tmp_path / name)
    synth.write_dng(path, cfa, kind, pattern=pattern, extra_tags=((51009, 7, list(lst)),))
    return path, cfa


def plan(path, strength=0.5, extra=()):
    g = pipe.Graph(lines=pipe.raw_graph_cfg(path, strength=strength, noise_a=1e-4, noise_b=1e-6, extra=list(extra)))
    g.export("o-pfm", "/tmp/unused")
    g.plan()
    g.commit()
    return g


def node(g, module, kernel):
    ns = [n for n in g.nodes() if n["module"] == module and n["kernel"] == kernel]
    assert len(ns) == 1, (module, kernel, len(ns))
    return ns[0]


def words(n):
    return np.frombuffer(n["pc"], np.uint32)


def f32(u):
    return np.uint32(u).view(np.float32)


@pytest.mark.parametrize("strength", [0.5, 0.0], ids=["wavelet", "noop"])
def test_gain_map_node_and_push_constants(tmp_path, strength):
    maps = ramp_maps(5, 7)
    path, _ = write(tmp_path, maps=maps, spacing=(0.25, 1 / 6.0), origin=(0.0, 0.0))
    g = plan(path, strength)
    gm = node(g, "denoise", "gainmap")
    assert (gm["wd"], gm["ht"]) == (7, 5)
    assert gm["conns"][0]["type"] == "source" and gm["conns"][0]["format"] == "f32"
    k = node(g, "denoise", "doub" if strength > 0 else "noop")
    w = words(k)
    ref = gs.gain_map_os(dict(gains=np.zeros((5, 7, 4)), origin=(0.0, 0.0), spacing=(0.25, 1 / 6.0)))
    if strength > 0:
        assert w[19] == 1 and [f32(v) for v in w[20:24]] == ref
        assert k["conns"][4]["link"][0] == gm["id"]
    else:
        assert w[17] == 1 and w[16] == BAYER and [f32(v) for v in w[12:16]] == ref
        assert k["conns"][2]["link"][0] == gm["id"]
    g.close()


def test_push_constants_match_the_restatement(tmp_path):
    maps = ramp_maps(4, 6)
    path, cfa = write(tmp_path, W=80, H=60, maps=maps, spacing=(1 / 3.0, 0.2), origin=(0.0, 0.0))
    g = plan(path)
    ip = g.img_param("i-raw", "main")
    gmd = dict(gains=np.stack(maps, -1), origin=(0.0, 0.0), spacing=(1 / 3.0, 0.2))
    p = gs.DevelopParams(kind="bayer", noise_a=np.float32(1e-4), noise_b=np.float32(1e-6),
                         cam=np.array(ip["cam"], np.float32).reshape(3, 3),
                         wb=tuple(ip["wb"][:3]) + (ip["wb"][3],), gain_map=gmd)
    _, ref = gs.build(80, 60, p)
    live = {(n["module"], n["kernel"]): n for n in g.nodes() if not n["culled"]}
    r = [x for x in ref if x["kernel"] == "doub"][0]
    assert live[("denoise", "doub")]["pc"] == np.ascontiguousarray(r["pc"]).tobytes()
    g.close()


@pytest.mark.parametrize("pattern,site_of", [(None, [0, 1, 2, 3]), ([[1, 0], [2, 1]], [1, 0, 3, 2]),
                                             ([[2, 1], [1, 0]], [3, 2, 1, 0])],
                         ids=["rggb", "grbg", "bggr"])
def test_upload_interleaves_the_sites_of_the_emitted_buffer(tmp_path, pattern, site_of):
    """map k covers raw site (k//2, k%2); after the CFA offset it lands on
    the emitted buffer's site ((top+oy)&1)*2 + ((left+ox)&1), its channel"""
    maps = ramp_maps(3, 5, seed=7)
    path, _ = write(tmp_path, maps=maps, pattern=pattern)
    g = plan(path)
    gm = node(g, "denoise", "gainmap")
    buf = np.zeros((3, 5, 4), np.float32)
    assert g.L.dt_graph_node_read_source(g.g, gm["id"], buf.ctypes.data) == 0
    for k in range(4):
        assert np.array_equal(buf[..., site_of[k]], maps[k]), k
    g.close()


def opcode(oid, params):
    return struct.pack(">4I", oid, 0x01030000, 0, len(params)) + params


@pytest.mark.parametrize("case", ["three_maps", "pitch_1", "different_grids", "same_site_twice", "one_point",
                                  "truncated_list", "polynomial_first", "map_polynomial_dng_size"])
def test_gain_maps_the_module_does_not_take(tmp_path, case):
    W, H = 64, 48
    maps = ramp_maps(4, 4)
    ops = synth.bayer_gain_maps(W, H, maps)
    poly = lambda d: opcode(8, struct.pack(">8I", 0, 0, H, W, 0, 1, 1, 1) + struct.pack(">I", d)
                            + struct.pack(f">{d}d", *([0.0] * d)))
    if case == "three_maps":
        ops = ops[:3]
    elif case == "pitch_1":
        ops[2] = synth.gain_map_opcode(maps[2], top=1, left=0, height=H, width=W, row_pitch=1)
    elif case == "different_grids":
        ops[3] = synth.gain_map_opcode(ramp_maps(5, 4)[0], top=1, left=1, height=H, width=W)
    elif case == "same_site_twice":
        ops[3] = synth.gain_map_opcode(maps[3], top=0, left=0, height=H, width=W)
    elif case == "one_point":
        ops = [synth.gain_map_opcode(np.ones((1, 1), np.float32), top=k // 2, left=k % 2, height=H, width=W)
               for k in range(4)]
    elif case == "polynomial_first":   # the first four opcodes must be the maps
        ops = [poly(1)] + ops
    elif case == "map_polynomial_dng_size":
        # MapPolynomial with degree+1 coefficients (the specification's size):
        # the reference's size rule (36 + 8*degree) drops the whole list
        ops = ops + [opcode(8, struct.pack(">8I", 0, 0, H, W, 0, 1, 1, 1) + struct.pack(">I", 2)
                            + struct.pack(">3d", 0.0, 1.0, 0.0))]
    if case == "truncated_list":
        lst = synth.opcode_list(ops)[:-4]
        cfa, _ = synth.mosaic(W, H, 1, "bayer")
        path = str(tmp_path / "t.dng")
        synth.write_dng(path, cfa, "bayer", extra_tags=((51009, 7, list(lst)),))
    else:
        path, _ = write(tmp_path, W, H, ops=ops)
    g = plan(path)
    assert not [n for n in g.nodes() if n["kernel"] == "gainmap"]
    assert words(node(g, "denoise", "doub"))[19] == 0
    g.close()


def test_map_polynomial_of_the_reference_size_keeps_the_maps(tmp_path):
    W, H = 64, 48
    ops = synth.bayer_gain_maps(W, H, ramp_maps(4, 4)) + [
        opcode(8, struct.pack(">8I", 0, 0, H, W, 0, 1, 1, 1) + struct.pack(">I", 2) + struct.pack(">2d", 0.0, 1.0))]
    path, _ = write(tmp_path, W, H, ops=ops)
    g = plan(path)
    assert words(node(g, "denoise", "doub"))[19] == 1
    g.close()


def test_batched_path_refuses_applied_gain_maps(tmp_path):
    from vkdt_amd import batch
    path, _ = write(tmp_path, maps=ramp_maps(4, 4))
    g = plan(path)
    with pytest.raises(Exception, match="gain maps"):
        batch.describe(g)
    g.close()
    g = plan(path, extra=["param:denoise:01:gainmap:0"])   # switched off: the batch takes it
    batch.describe(g)
    g.close()


# ---------------------------------------------------------------------------
# oracle: the gain application
# ---------------------------------------------------------------------------

def dn_params(gainmap=1, strength=0.5):
    return gs.mixed_blob([("f", strength), ("f", 0.6), ("f", 1.0), ("f", 0)] + [("f", 0)] * 4 + [("i", gainmap)])


def noop_pc(W, H, map_os, gainmap=1, filters=BAYER, crop=(0, 0)):
    black, white = np.float32(512 / 65535.0), np.float32(16383 / 65535.0)
    return gs.mixed_blob([("i", crop[0]), ("i", crop[1]), ("i", W), ("i", H)] + [("f", black)] * 4
                         + [("f", white)] * 4 + [("f", v) for v in map_os] + [("i", filters), ("i", gainmap)])


def run_noop(raw, gm, pc, params, wd, ht, bits=64):
    out = np.zeros((ht, wd, 4))
    bufs = [op.orbuf(raw, op.OR_UI16, 1), op.orbuf(out, op.OR_F32, 4)]
    if gm is not None:
        bufs.append(op.orbuf(np.ascontiguousarray(gm, np.float64), op.OR_F32, 4))
    op.node(bits, "denoise", "noop", pc, params, bufs, wd, ht, storage_f16=False)
    return out


def check_noop_constant_gains_scale_each_site(bits=64):
    rng = np.random.default_rng(11)
    W, H = 40, 30
    raw = rng.integers(0, 16384, size=(H, W), dtype=np.uint16)
    k = np.array([0.7, 1.1, 1.3, 0.9])
    gm = np.broadcast_to(k, (3, 4, 4)).copy()
    os_ = [0.0, 0.0, 1.0, 1.0]
    base = run_noop(raw, None, noop_pc(W, H, os_, gainmap=0), dn_params(), W, H, bits=bits)
    with_g = run_noop(raw, gm, noop_pc(W, H, os_), dn_params(), W, H, bits=bits)
    site = (np.arange(W)[None, :] & 1) + 2 * (np.arange(H)[:, None] & 1)
    kk = k.astype(np.float32) if bits == 32 else k     # the f32 build reads the grid as float
    want = (base[..., 0].astype(kk.dtype) * kk[site]).astype(np.float64)
    # texture() of a constant grid is the constant up to the lerp's rounding
    assert np.allclose(with_g[..., 0], want, rtol=1e-6 if bits == 32 else 1e-14, atol=0)
    # the module parameter off, or an x-trans mosaic: no gain
    assert np.array_equal(run_noop(raw, gm, noop_pc(W, H, os_), dn_params(gainmap=0), W, H, bits=bits), base)
    xt = run_noop(raw, gm, noop_pc(W, H, os_, filters=9), dn_params(), W, H, bits=bits)
    assert np.array_equal(xt, run_noop(raw, None, noop_pc(W, H, os_, gainmap=0, filters=9), dn_params(), W, H,
                                       bits=bits))


def linear_grid(mh, mw):
    a = np.array([1.0, 0.9, 1.1, 0.95]); bx = np.array([0.2, -0.1, 0.05, 0.3]); by = np.array([-0.15, 0.1, 0.2, 0.0])
    ii, jj = np.meshgrid(np.arange(mw), np.arange(mh))
    gm = a + bx * ii[..., None] / (mw - 1) + by * jj[..., None] / (mh - 1)

    def at(px, py, site):
        """the grid's value at texture position (px, py): linear in the texel
        coordinate p*n - 0.5, clamped to the grid (mirrored edges)"""
        u = np.clip(px * mw - 0.5, 0, mw - 1) / (mw - 1)
        v = np.clip(py * mh - 0.5, 0, mh - 1) / (mh - 1)
        return a[site] + bx[site] * u + by[site] * v
    return gm, at


def check_noop_linear_map(origin=(0.1, 0.05), scale=(1.25, 0.8)):
    """gains linear in the map's columns and rows come back as the same
    linear function of the texture coordinate, at pos = clamp(((x + 0.5 +
    crop)/W) * sx - ox, 0, 1)"""
    W, H, cx, cy = 48, 36, 4, 2
    raw = np.full((H, W), 8000, np.uint16)
    gm, at = linear_grid(5, 6)
    os_ = [float(np.float32(v)) for v in (origin[1], origin[0], scale[1], scale[0])]   # as pushed
    wd, ht = W - 2 * cx, H - 2 * cy
    base = run_noop(raw, None, noop_pc(W, H, os_, gainmap=0, crop=(cx, cy)), dn_params(), wd, ht)
    out = run_noop(raw, gm, noop_pc(W, H, os_, crop=(cx, cy)), dn_params(), wd, ht)
    x = np.arange(wd)[None, :]
    y = np.arange(ht)[:, None]
    px = np.clip((x + 0.5 + cx) / W * os_[2] - os_[0], 0, 1)
    py = np.clip((y + 0.5 + cy) / H * os_[3] - os_[1], 0, 1)
    want = at(px, py, (x & 1) + 2 * (y & 1))
    assert np.allclose(out[..., 0] / base[..., 0], want, rtol=0, atol=1e-12)


def run_doub(raw, crs0, crs1, gm, gainmap, map_os, W, H, bits=64):
    black, white = np.float32(512 / 65535.0), np.float32(16383 / 65535.0)
    pc = gs.mixed_blob([("f", 1.0)] * 4 + [("f", black)] * 4 + [("f", white)] * 4
                       + [("i", 0), ("i", 0), ("i", W), ("i", H)]
                       + [("i", BAYER), ("f", 1e-4), ("f", 1e-6), ("i", gainmap)] + [("f", v) for v in map_os])
    out = np.zeros((H, W))
    bufs = [op.orbuf(raw, op.OR_UI16, 1), op.orbuf(crs0, op.OR_F32, 4), op.orbuf(crs1, op.OR_F32, 4),
            op.orbuf(out, op.OR_F32, 1)]
    if gm is not None:
        bufs.append(op.orbuf(np.ascontiguousarray(gm, np.float64), op.OR_F32, 4))
    op.node(bits, "denoise", "doub", pc, dn_params(), bufs, W, H, storage_f16=False)
    return out


def doub_inputs(W, H, seed=12):
    rng = np.random.default_rng(seed)
    raw = rng.integers(600, 15000, size=(H, W), dtype=np.uint16)
    crs0 = rng.uniform(0.05, 0.8, size=(H // 2, W // 2, 4))
    crs1 = rng.uniform(0.05, 0.8, size=(H // 2, W // 2, 4))
    return raw, crs0, crs1


def check_doub_constant_gains_scale_each_site(bits=64):
    """doub.comp:106-114 on a constant map: the denoised value times the
    gain of its CFA site, and nothing else changes"""
    W, H = 32, 24
    raw, crs0, crs1 = doub_inputs(W, H)
    k = np.array([1.2, 0.85, 1.05, 0.75])
    gm = np.broadcast_to(k, (4, 5, 4)).copy()
    base = run_doub(raw, crs0, crs1, None, 0, [0, 0, 1, 1], W, H, bits)
    out = run_doub(raw, crs0, crs1, gm, 1, [0, 0, 1, 1], W, H, bits)
    site = (np.arange(W)[None, :] & 1) + 2 * (np.arange(H)[:, None] & 1)
    kk = k.astype(np.float32) if bits == 32 else k
    assert np.allclose(out, (base.astype(kk.dtype) * kk[site]).astype(np.float64), rtol=1e-6 if bits == 32 else 1e-14,
                       atol=0)


def check_doub_linear_map_at_the_block_centre():
    """doub samples the grid at its 2x2 block's centre on the half-resolution
    grid: pos = (0.5 + (x + crop)//2) / (W//2), the same for the block's four
    sites (which then read their own channels)"""
    W, H = 36, 28
    raw, crs0, crs1 = doub_inputs(W, H, seed=13)
    gm, at = linear_grid(4, 5)
    os_ = [float(np.float32(v)) for v in (0.02, 0.04, 1.1, 0.9)]
    base = run_doub(raw, crs0, crs1, None, 0, os_, W, H)
    out = run_doub(raw, crs0, crs1, gm, 1, os_, W, H)
    x = np.arange(W)[None, :]
    y = np.arange(H)[:, None]
    px = np.clip((0.5 + x // 2) / (W // 2) * os_[2] - os_[0], 0, 1)
    py = np.clip((0.5 + y // 2) / (H // 2) * os_[3] - os_[1], 0, 1)
    want = at(px, py, (x & 1) + 2 * (y & 1))
    ok = base > 1e-6
    assert np.allclose(out[ok] / base[ok], np.broadcast_to(want, out.shape)[ok], rtol=0, atol=1e-12)


# the gain checks the planted-mistake test runs beside the stencil pins
CHECKS = [check_noop_constant_gains_scale_each_site, check_noop_linear_map, check_doub_constant_gains_scale_each_site,
          check_doub_linear_map_at_the_block_centre]


@pytest.mark.parametrize("bits", [64, 32])
def test_noop_constant_gains_scale_each_site(bits):
    check_noop_constant_gains_scale_each_site(bits)


@pytest.mark.parametrize("origin,scale", [((0.0, 0.0), (1.0, 1.0)), ((0.1, 0.05), (1.25, 0.8))])
def test_noop_linear_map_is_interpolated_at_the_texel_position(origin, scale):
    check_noop_linear_map(origin, scale)


@pytest.mark.parametrize("bits", [64, 32])
def test_doub_constant_gains_scale_each_site(bits):
    check_doub_constant_gains_scale_each_site(bits)


def test_doub_linear_map_at_the_block_centre():
    check_doub_linear_map_at_the_block_centre()
