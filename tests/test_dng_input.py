"""Raw input on the CPU: the DNG reader of the i-raw module and i-raw's
integer steps (CFA offset normalisation, block rounding, row copy), which
must be bit exact (i-raw/main.c:138-299, rawloader-c/lib.rs:200-283)."""
import ctypes as C
import os

import numpy as np
import pytest

from vkdt_amd import PKG_DIR, pipe, synth


class DngImage(C.Structure):
    _fields_ = [("maker", C.c_char * 32), ("model", C.c_char * 32), ("width", C.c_uint64), ("height", C.c_uint64),
                ("stride", C.c_uint32), ("cpp", C.c_uint32), ("wb_coeffs", C.c_float * 4),
                ("whitelevels", C.c_float * 4), ("blacklevels", C.c_float * 4), ("xyz_to_cam", C.c_float * 12),
                ("illuminant", C.c_uint32), ("cfa_w", C.c_uint32), ("cfa_h", C.c_uint32), ("cfa", C.c_uint8 * 36),
                ("crop_aabb", C.c_uint64 * 4), ("iso", C.c_float), ("data", C.POINTER(C.c_uint16)),
                ("packed", C.c_void_p), ("packed_len", C.c_size_t), ("tiles", C.c_void_p), ("ntiles", C.c_uint32)]


@pytest.fixture(scope="module")
def iraw():
    pipe.lib()
    L = C.CDLL(os.path.join(PKG_DIR, "modules", "i-raw", "libi-raw.so"))
    L.dng_load.argtypes = [C.c_char_p, C.POINTER(DngImage), C.c_char_p, C.c_size_t]
    L.dng_free.argtypes = [C.POINTER(DngImage)]
    return L


def load(L, path):
    img, err = DngImage(), C.create_string_buffer(512)
    rc = L.dng_load(path.encode(), C.byref(img), err, 512)
    return rc, img, err.value.decode()


@pytest.mark.parametrize("kind,W,H", [("bayer", 64, 48), ("xtrans", 66, 48)])
def test_round_trip_bit_exact(tmp_path, iraw, kind, W, H):
    rng = np.random.default_rng(11)
    cfa = rng.integers(0, 65536, size=(H, W), dtype=np.uint16)
    path = str(tmp_path / "a.dng")
    synth.write_dng(path, cfa, kind, black=600, white=15000, as_shot_neutral=(0.5, 1.0, 0.8))
    rc, img, err = load(iraw, path)
    assert rc == 0, err
    assert (img.width, img.height, img.stride, img.cpp) == (W, H, W, 1)
    data = np.ctypeslib.as_array(img.data, shape=(H * W,)).reshape(H, W).copy()
    assert np.array_equal(data, cfa)
    assert list(img.blacklevels) == [600.0] * 4 and list(img.whitelevels) == [15000.0] * 4
    n = synth.cfa_pattern(kind).shape[0]
    assert (img.cfa_w, img.cfa_h) == (n, n)
    assert np.array_equal(np.array(img.cfa[:n * n]).reshape(n, n), synth.cfa_pattern(kind))
    assert np.allclose(list(img.wb_coeffs)[:3], [2.0, 1.0, 1.25], rtol=1e-6)
    assert np.allclose(np.array(img.xyz_to_cam[:9]).reshape(3, 3), synth.xyz_to_cam(), atol=2e-6)
    assert img.illuminant == 4   # D65
    assert list(img.crop_aabb) == [0, 0, W, H]
    iraw.dng_free(C.byref(img))


def test_rejects_bad_files(tmp_path, iraw):
    cfa = np.zeros((8, 8), np.uint16)
    good = str(tmp_path / "g.dng")
    synth.write_dng(good, cfa)
    raw = open(good, "rb").read()
    trunc = str(tmp_path / "t.dng")
    open(trunc, "wb").write(raw[:len(raw) - 20])
    rc, _, err = load(iraw, trunc)
    assert rc != 0 and "truncated" in err
    notdng = str(tmp_path / "n.dng")
    open(notdng, "wb").write(b"P5\n8 8\n255\n" + bytes(64))
    rc, _, err = load(iraw, notdng)
    assert rc != 0 and "TIFF" in err
    comp = str(tmp_path / "c.dng")
    synth.write_dng(comp, cfa, extra_tags=[(259, 3, [7])])   # lossless JPEG tag over plain samples
    rc, _, err = load(iraw, comp)
    assert rc != 0 and "lj92" in err
    lossy = str(tmp_path / "l.dng")
    synth.write_dng(lossy, cfa, extra_tags=[(259, 3, [34892])])   # lossy JPEG
    rc, _, err = load(iraw, lossy)
    assert rc != 0 and "compression" in err
    rc, _, err = load(iraw, str(tmp_path / "missing.dng"))
    assert rc != 0


def plan_raw(tmp_path, cfa, kind, pattern=None, name="p.dng"):
    path = str(tmp_path / name)
    synth.write_dng(path, cfa, kind, pattern=pattern)
    g = pipe.Graph(lines=pipe.raw_graph_cfg(path, noise_a=1e-7, noise_b=1e-5))
    g.export("o-pfm", str(tmp_path / "o"))
    g.plan()
    return g


BAYER_SHIFTS = {  # pattern -> (ox, oy) that makes it RGGB (lib.rs:245-257)
    ((0, 1), (1, 2)): (0, 0), ((1, 0), (2, 1)): (1, 0), ((1, 2), (0, 1)): (0, 1), ((2, 1), (1, 0)): (1, 1)}


@pytest.mark.parametrize("pat", list(BAYER_SHIFTS), ids=["rggb", "grbg", "gbrg", "bggr"])
def test_bayer_offset_normalisation(tmp_path, pat):
    W, H = 34, 26
    cm = np.tile(np.array(pat, np.uint8), (H // 2, W // 2))
    cfa = (1000 * (cm.astype(np.uint16) + 1) + np.arange(W, dtype=np.uint16)[None, :] % 7).astype(np.uint16)
    g = plan_raw(tmp_path, cfa, "bayer", pattern=pat)
    ip = g.img_param("i-raw", "main")
    ox, oy = BAYER_SHIFTS[pat]
    w, h = ((W - ox) // 2) * 2, ((H - oy) // 2) * 2
    assert ip["full"] == (w, h)
    assert ip["filters"] == synth.BAYER_FILTERS
    buf = g.read_source("i-raw", "main", w * h * 2).view(np.uint16).reshape(h, w)
    assert np.array_equal(buf, cfa[oy:oy + h, ox:ox + w])
    assert np.array_equal((buf // 1000 - 1).astype(np.uint8), synth.channel_map(w, h, "bayer"))


@pytest.mark.parametrize("shift", [(0, 0), (1, 0), (2, 0), (0, 1), (3, 2), (5, 5), (4, 3), (2, 4)])
def test_xtrans_offset_normalisation(tmp_path, shift):
    """any 6x6 phase of the x-trans layout is moved to the canonical one, and
    the extent is rounded down to whole 3x3 blocks"""
    dx, dy = shift
    W, H = 48, 42
    canon = synth.XTRANS
    pat = np.roll(np.roll(canon, -dy, axis=0), -dx, axis=1)
    cm = np.tile(pat, (H // 6 + 1, W // 6 + 1))[:H, :W]
    cfa = (1000 * (cm.astype(np.uint16) + 1)).astype(np.uint16)
    g = plan_raw(tmp_path, cfa, "xtrans", pattern=pat)
    ip = g.img_param("i-raw", "main")
    w, h = ip["full"]
    # the offset is below 6, block rounding drops at most 2 more
    assert ip["filters"] == 9 and w % 3 == 0 and h % 3 == 0 and W - w < 8 and H - h < 8
    buf = g.read_source("i-raw", "main", w * h * 2).view(np.uint16).reshape(h, w)
    assert np.array_equal((buf // 1000 - 1).astype(np.uint8), synth.channel_map(w, h, "xtrans"))


def test_xtrans_6240x4160_rounds_to_4158(tmp_path):
    W, H = 6240, 4160
    path = str(tmp_path / "x.dng")
    synth.write_dng(path, np.zeros((H, W), np.uint16), "xtrans")
    g = pipe.Graph(lines=pipe.raw_graph_cfg(path))
    g.export("o-pfm", str(tmp_path / "o"))
    g.plan()
    assert g.img_param("i-raw", "main")["full"] == (6240, 4158)


def test_camera_matrix_and_white_balance(tmp_path):
    """cam_to_rec2020 = xyz_to_rec2020 . inverse(xyz_to_cam); wb = 1/neutral, g-normalised"""
    cfa = np.zeros((16, 16), np.uint16)
    path = str(tmp_path / "m.dng")
    synth.write_dng(path, cfa, "bayer", as_shot_neutral=(0.4, 0.8, 0.5))
    g = pipe.Graph(lines=pipe.raw_graph_cfg(path))
    g.export("o-pfm", str(tmp_path / "o"))
    g.plan()
    ip = g.img_param("i-raw", "main")
    assert np.allclose(ip["wb"][:3], [2.0, 1.0, 1.6], rtol=1e-6)
    xyz_to_2020 = np.array([[1.71665119, -0.35567078, -0.25336628], [-0.66668435, 1.61648124, 0.01576855],
                            [0.01763986, -0.04277061, 0.94210312]])
    ref = xyz_to_2020 @ np.linalg.inv(synth.xyz_to_cam())
    assert np.allclose(np.array(ip["cam"]).reshape(3, 3), ref, atol=1e-5)


def test_rejects_crop_outside_the_raw(tmp_path, iraw):
    """DefaultCropOrigin + DefaultCropSize beyond the raw buffer is refused at load"""
    cfa = np.full((24, 32), 1000, np.uint16)
    path = str(tmp_path / "c.dng")
    synth.write_dng(path, cfa, "bayer", extra_tags=[(50719, 4, [0, 0]), (50720, 4, [40, 24])])
    rc, img, err = load(iraw, path)
    assert rc != 0 and "outside the 32x24 raw" in err


@pytest.mark.parametrize("W,H,crop,what", [(2, 4, None, "smaller than its 2x2 CFA"),
                                           (32, 24, (2, 24), "smaller than one CFA block")])
def test_rejects_extents_below_the_cfa_offset(tmp_path, W, H, crop, what):
    """a GRBG raw (CFA offset 1,0) narrower than the offset plus one CFA, or
    cropped narrower than the offset plus one block, is refused with a
    message instead of wrapping its extent"""
    cfa = np.full((H, W), 1000, np.uint16)
    path = str(tmp_path / "o.dng")
    extra = [] if crop is None else [(50719, 4, [0, 0]), (50720, 4, list(crop))]
    synth.write_dng(path, cfa, "bayer", pattern=[[1, 0], [2, 1]], extra_tags=extra)
    g = pipe.Graph(lines=pipe.raw_graph_cfg(path))
    g.export("o-pfm", str(tmp_path / "out"))
    with pytest.raises(pipe.PipeError):
        g.plan()
    assert what in g.error()


def test_rejects_lj92_streams_taller_than_their_tile(tmp_path, iraw):
    """a lossless-JPEG stream with more rows than its tile would write into
    the next tile's rows: refused at load"""
    rng = np.random.default_rng(3)
    cfa = rng.integers(0, 16384, size=(64, 64), dtype=np.uint16)
    path = str(tmp_path / "t.dng")
    synth.write_dng(path, cfa, "bayer", compress=dict(tile=(32, 32), extra_rows=2))
    rc, img, err = load(iraw, path)
    assert rc != 0 and "does not fill its 32x32 tile" in err
    synth.write_dng(path, cfa, "bayer", compress=dict(tile=(32, 32)))
    rc, img, err = load(iraw, path)
    assert rc == 0, err
    iraw.dng_free(C.byref(img))
