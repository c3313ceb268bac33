"""Pins of the oracle against what the mathematics and the reference's
formats fix, so that a plausible restatement mistake (dropped half-texel
offset, transposed matrix, wrong gather order, wrong CFA parity, wrong
eigenvector) fails here and not only against the GPU.

Each test names the reference behaviour it pins.  The reference ships no
golden vectors or known-answer tests for this path, so these closed forms,
invariants, library routines (numpy float16, numpy eigh, scipy Weibull)
and brute-force checks on tiny inputs are the pins.
"""
import numpy as np
import pytest
from scipy import stats

import graphspec as gs
import oracle_py as op
from vkdt_amd import synth

BAYER = synth.BAYER_FILTERS


def blob_f(*vals):
    return np.asarray(vals, np.float32)


# ---------------------------------------------------------------------------
# storage and sampler semantics (graph-run-nodes-allocate.h:39-96, qvk.c:596-611)
# ---------------------------------------------------------------------------

def test_f16_round_matches_numpy_binary16():
    rng = np.random.default_rng(1)
    xs = np.concatenate([
        rng.normal(size=20000) * np.exp2(rng.integers(-30, 18, size=20000)),
        [0.0, -0.0, 65504.0, 65519.99, 65520.0, 65536.0, 1e9, -1e9, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26,
         2.0 ** -14, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, np.inf, -np.inf]])
    L = op.lib(64)
    ours = np.array([L.or_f16_round(float(x)) for x in xs])
    with np.errstate(over="ignore"):
        ref = xs.astype(np.float16).astype(np.float64)
    assert np.array_equal(ours, ref)
    assert np.isnan(L.or_f16_round(float("nan")))


def test_mirrored_repeat_addressing():
    L = op.lib(64)
    for n in (1, 2, 3, 7):
        idx = np.arange(-3 * n, 3 * n)
        got = np.array([L.or_mirror(int(i), n) for i in idx])
        assert got.min() >= 0 and got.max() < n
        # the first period equals numpy's symmetric padding (edge repeated once)
        ref = np.pad(np.arange(n), n, mode="symmetric")
        assert np.array_equal(got[2 * n:5 * n], ref)
        # period 2n, and the reflections about -1/2 and n-1/2
        assert np.array_equal(got[:2 * n], got[2 * n:4 * n])
        for i in range(n):
            assert L.or_mirror(-1 - i, n) == i and L.or_mirror(2 * n - 1 - i, n) == i


def test_bilinear_reproduces_affine_and_half_texel_offset():
    L = op.lib(64)
    W, H = 9, 7
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    img = np.ascontiguousarray(0.3 + 0.07 * x - 0.05 * y)
    rng = np.random.default_rng(2)
    for _ in range(200):
        s, t = rng.uniform(0.5, W - 0.5), rng.uniform(0.5, H - 0.5)
        v = L.or_bilinear(img.ctypes.data, W, H, s, t)
        assert abs(v - (0.3 + 0.07 * (s - 0.5) - 0.05 * (t - 0.5))) < 1e-12
    # left of the first texel centre the mirrored tap repeats texel 0
    assert abs(L.or_bilinear(img.ctypes.data, W, H, 0.2, 3.5) - img[3, 0]) < 1e-12


def test_gather_component_order():
    """textureGather: x=(i0,j1) y=(i1,j1) z=(i1,j0) w=(i0,j0) (Vulkan spec 16.x)"""
    L = op.lib(64)
    img = np.array([[11, 22], [33, 44]], np.uint16)   # img[y][x]
    out = np.zeros(4)
    L.or_gather(img.ctypes.data, 2, 2, 1.0, 1.0, out.ctypes.data)
    expect = np.float32(np.array([33, 44, 22, 11]) / 65535.0)
    assert np.array_equal(out, expect.astype(np.float64))


def test_evd2x2_against_numpy_eigh():
    L = op.lib(64)
    rng = np.random.default_rng(3)
    for _ in range(500):
        a, c = rng.uniform(0.01, 5, 2)
        b = rng.uniform(-2, 2)
        M = np.array([a, b, b, c], np.float64)     # column major, symmetric
        ev, e0, e1 = np.zeros(2), np.zeros(2), np.zeros(2)
        L.or_evd2x2(M.ctypes.data, ev.ctypes.data, e0.ctypes.data, e1.ctypes.data)
        A = np.array([[a, b], [b, c]])
        w = np.linalg.eigh(A)[0]
        assert abs(ev[0] - w[1]) < 1e-9 and abs(ev[1] - w[0]) < 1e-9
        assert np.allclose(A @ e0, ev[0] * e0, atol=1e-9)   # evec0: major axis
        assert np.allclose(A @ e1, ev[1] * e1, atol=1e-9)
        assert abs(np.linalg.norm(e0) - 1) < 1e-12 and abs(e0 @ e1) < 1e-12


# ---------------------------------------------------------------------------
# denoise (src/pipe/modules/denoise)
# ---------------------------------------------------------------------------

def rgbc_pc(W, H, black=512.0, white=16383.0, wb=(1, 1, 1, 1), crop=(0, 0, 0, 0), tail=()):
    b = np.float32(black) / np.float32(65535)
    w = np.float32(white) / np.float32(65535)
    return gs.mixed_blob([("f", v) for v in wb] + [("f", b)] * 4 + [("f", w)] * 4 + [("i", c) for c in crop] + list(tail))


def dn_params(strength=0.5, luma=0.6, detail=1.0):
    return gs.mixed_blob([("f", strength), ("f", luma), ("f", detail), ("f", 0)] + [("f", 0)] * 4 + [("i", 1)])


def test_noop_is_black_white_scaling():
    """noop.comp:43-44: max(0, (v/65535 - black)/(white - black)) on the crop"""
    rng = np.random.default_rng(4)
    W, H = 40, 30
    raw = rng.integers(0, 16384, size=(H, W), dtype=np.uint16)
    out = np.zeros((H - 4, W - 6, 4))
    black, white = np.float32(512 / 65535.0), np.float32(16383 / 65535.0)
    pc = gs.mixed_blob([("i", 6), ("i", 4), ("i", W), ("i", H)] + [("f", black)] * 4 + [("f", white)] * 4
                       + [("f", 0)] * 4 + [("i", BAYER), ("i", 0)])
    op.node(64, "denoise", "noop", pc, dn_params(0.0), [op.orbuf(raw, op.OR_UI16, 1), op.orbuf(out, op.OR_F32, 4)],
            W - 6, H - 4, storage_f16=False)
    v = np.float32(raw[4:, 6:] / 65535.0).astype(np.float64)
    ref = np.maximum(0, (v - float(black)) / (float(white) - float(black)))
    assert np.allclose(out[..., 0], ref, rtol=0, atol=1e-15)
    assert np.all(out[..., 3] == 1)


def test_half_bayer_block_average_with_clipped_green_swap():
    """half.comp:65-72: r=(0,0), g=mean of the greens, b=(1,1); a green at the
    white level is replaced by the other green"""
    rng = np.random.default_rng(5)
    W, H = 16, 12
    raw = rng.integers(600, 16000, size=(H, W), dtype=np.uint16)
    raw[2, 3] = 16383        # clipped green at (x=3,y=2) of block (1,1)
    out = np.zeros((H // 2, W // 2, 4))
    op.node(64, "denoise", "half", rgbc_pc(W, H, tail=[("i", BAYER)]), None,
            [op.orbuf(raw, op.OR_UI16, 1), op.orbuf(out, op.OR_F32, 4)], W // 2, H // 2, storage_f16=False)
    u = np.float32(raw / 65535.0).astype(np.float64)
    white = float(np.float32(16383 / 65535.0))
    for by in range(H // 2):
        for bx in range(W // 2):
            r, g1, g2, b = u[2 * by, 2 * bx], u[2 * by, 2 * bx + 1], u[2 * by + 1, 2 * bx], u[2 * by + 1, 2 * bx + 1]
            if g2 >= white:
                g2 = g1
            if g1 >= white:
                g1 = g2
            assert np.allclose(out[by, bx], [r, (g1 + g2) / 2, b, 1.0], atol=1e-15)


def test_half_xtrans_uses_the_canonical_pattern():
    """half.comp:32-64: 5 greens averaged, the two reds / two blues of each 3x3
    block averaged; checked against the canonical x-trans layout"""
    rng = np.random.default_rng(6)
    W, H = 18, 12
    raw = rng.integers(600, 16000, size=(H, W), dtype=np.uint16)
    out = np.zeros((H // 3, W // 3, 4))
    op.node(64, "denoise", "half", rgbc_pc(W, H, tail=[("i", 9)]), None,
            [op.orbuf(raw, op.OR_UI16, 1), op.orbuf(out, op.OR_F32, 4)], W // 3, H // 3, storage_f16=False)
    u = np.float32(raw / 65535.0).astype(np.float64)
    cm = synth.channel_map(W, H, "xtrans")
    for by in range(H // 3):
        for bx in range(W // 3):
            blk, c = u[3 * by:3 * by + 3, 3 * bx:3 * bx + 3], cm[3 * by:3 * by + 3, 3 * bx:3 * bx + 3]
            assert (c == 1).sum() == 5 and (c == 0).sum() == 2 and (c == 2).sum() == 2
            ref = [blk[c == 0].mean(), blk[c == 1].mean(), blk[c == 2].mean()]
            assert np.allclose(out[by, bx, :3], ref, atol=1e-15)


def develop_params(kind, **kw):
    na, nb = synth.noise_params(2e-4, 1e-6)
    return gs.DevelopParams(kind=kind, noise_a=na, noise_b=nb, **kw)


@pytest.mark.parametrize("kind", ["bayer", "xtrans"])
def test_denoise_constant_field(kind):
    """a flat raw has zero wavelet bands: denoise returns (c-black)/(white-black)"""
    W, H = 48, 36
    c = 5000
    cfa = np.full((H, W), c, np.uint16)
    _, st = op.develop(64, cfa, gs.oracle_params(develop_params(kind)), stages=True, storage_f16=False)
    ref = (np.float32(c / 65535.0) - np.float32(512 / 65535.0)) / (np.float32(16383 / 65535.0) - np.float32(512 / 65535.0))
    assert np.allclose(st["denoise"], float(ref), rtol=2e-6)


@pytest.mark.parametrize("kind", ["bayer", "xtrans"])
def test_demosaic_per_channel_constant(kind):
    """gaussian splats normalise their weights: a CFA holding (a,b,c) per
    colour gives (a,b,c) away from the borders"""
    W, H = 48, 36
    vals = np.array([3000, 7000, 2000], np.uint16)
    cfa = vals[synth.channel_map(W, H, kind)]
    p = develop_params(kind, strength=1e-6)
    _, st = op.develop(64, cfa, gs.oracle_params(p), stages=True, storage_f16=False)
    b, w = np.float32(512 / 65535.0), np.float32(16383 / 65535.0)
    ref = [(float(np.float32(v / 65535.0)) - float(b)) / (float(w) - float(b)) for v in vals]
    inner = st["demosaic"][6:-6, 6:-6, :3]
    assert np.allclose(inner, ref, rtol=1e-4)


# ---------------------------------------------------------------------------
# colour (colour/main-impl.glsl:200-341) and filmcurv (filmcurv/main.comp)
# ---------------------------------------------------------------------------

M_CAT16 = np.array([[0.401288, 0.650173, -0.051461], [-0.250268, 1.204414, 0.045854], [-0.002079, 0.048952, 0.953127]])
M_2020_XYZ = np.array([[0.636958048301290991, 0.144616903586208406, 0.168880975164172054],
                       [0.26270021201126692, 0.677998071518871148, 0.0593017164698619384],
                       [4.9999999999999999e-17, 0.0280726930490874452, 1.06098505771079066]])


def colour_node(img, cam, mul, expo):
    H, W = img.shape[:2]
    cp = np.zeros(242, np.float32)
    cp[0:3] = mul
    cp[3] = 2.0 ** expo
    for j in range(3):
        for i in range(3):
            cp[4 + 4 * i + j] = cam[j][i]
    cp[226] = 1.0
    out = np.zeros_like(img)
    op.node(64, "colour", "main", gs.mixed_blob([("i", 0)] * 3), cp,
            [op.orbuf(img, op.OR_F32, 4), op.orbuf(out, op.OR_F32, 4)], W, H, storage_f16=False)
    return out


def test_colour_identity():
    rng = np.random.default_rng(7)
    img = np.ascontiguousarray(rng.uniform(0, 2, size=(5, 7, 4)))
    out = colour_node(img, np.eye(3), (1, 1, 1), 0.0)
    assert np.allclose(out[..., :3], img[..., :3], atol=1e-6)


def test_colour_matrix_cat16_exposure_closed_form():
    """rgb' = 2^e * X2R M16^-1 diag(M16 R2X mul / M16 R2X 1) M16 R2X (cam . rgb),
    clamped to +-65535 (von Kries in CAT16 cone space)"""
    rng = np.random.default_rng(8)
    img = np.ascontiguousarray(rng.uniform(-0.1, 3, size=(6, 5, 4)))
    cam = np.float32(rng.uniform(-0.3, 1.2, size=(3, 3)) + np.eye(3))
    mul = np.float32([1.7, 1.0, 0.6])
    expo = 0.7
    out = colour_node(img, cam, mul, expo)
    T = M_CAT16 @ M_2020_XYZ
    D = np.diag((T @ mul.astype(np.float64)) / (T @ np.ones(3)))
    A = (2 ** expo) * np.linalg.inv(T) @ D @ T @ cam.astype(np.float64)
    ref = np.clip(img[..., :3] @ A.T, -65535, 65535)
    # the reference's xyz_to_rec2020 is rec2020_to_xyz^-1 to 8 digits
    assert np.allclose(out[..., :3], ref, rtol=2e-6, atol=2e-6)


def filmcurv_node(img, light=3.0, contrast=1.2, bias=0.0, mode=3):
    H, W = img.shape[:2]
    out = np.zeros_like(img)
    par = gs.mixed_blob([("f", light), ("f", contrast), ("f", bias), ("i", mode), ("f", 1)] + [("f", 0)] * 5)
    op.node(64, "filmcurv", "main", None, par, [op.orbuf(img, op.OR_F32, 4), op.orbuf(out, op.OR_F32, 4)], W, H,
            storage_f16=False)
    return out


def test_filmcurv_gray_is_weibull_cdf():
    x = np.linspace(1e-4, 2.0, 64)
    img = np.ascontiguousarray(np.stack([x, x, x, np.ones_like(x)], -1)[None])
    out = filmcurv_node(img)
    ref = stats.weibull_min.cdf(x, float(np.float32(1.2)), scale=1.0 / 3.0)
    assert np.allclose(out[0, :, 0], ref, atol=1e-12)
    assert np.allclose(out[0, :, 1], ref, atol=1e-12) and np.allclose(out[0, :, 2], ref, atol=1e-12)


def test_filmcurv_dng_hue_constancy():
    """adjust_colour_dng (shared.glsl:371-387): per-channel curve on max and
    min, the middle channel keeps its relative position"""
    rng = np.random.default_rng(9)
    img = np.ascontiguousarray(np.concatenate([rng.uniform(0.001, 1.5, size=(8, 8, 3)), np.ones((8, 8, 1))], -1))
    out = filmcurv_node(img)
    c0, c1 = img[..., :3].reshape(-1, 3), out[..., :3].reshape(-1, 3)
    cdf = lambda v: stats.weibull_min.cdf(v, float(np.float32(1.2)), scale=1.0 / 3.0)
    for a, b in zip(c0, c1):
        lo, mid, hi = np.argsort(a, kind="stable")
        assert abs(b[hi] - cdf(a[hi])) < 1e-12 and abs(b[lo] - cdf(a[lo])) < 1e-12
        t = (a[mid] - a[lo] + 1e-6) / (a[hi] - a[lo] + 1e-6)
        assert abs(b[mid] - (b[lo] + t * (b[hi] - b[lo]))) < 1e-9


# ---------------------------------------------------------------------------
# llap (llap/curve.comp, reduce.comp, assemble.comp, colour.comp)
# ---------------------------------------------------------------------------

def llap_colour_node(yv, rgb):
    n = len(yv)
    lum = np.ascontiguousarray(np.asarray(yv, np.float64).reshape(1, n))
    org = np.ascontiguousarray(np.concatenate([rgb, np.ones((n, 1))], 1).reshape(1, n, 4))
    out = np.zeros((1, n, 4))
    op.node(64, "llap", "colour", None, None, [op.orbuf(lum, op.OR_F32, 1), op.orbuf(org, op.OR_F32, 4),
                                                op.orbuf(out, op.OR_F32, 4)], n, 1, storage_f16=False)
    return out[0, :, :3]


LUMA2020 = np.array([0.26270021201126692, 0.677998071518871148, 0.0593017164698619384])   # Y of rec2020 (matrices.h:49)


def check_llap_colour_transfer_and_black_rolloff():
    """colour.comp:221-240: the output is the input colour scaled to the new
    luminance yv where yv is at least the input's luminance yo; below it the
    luminance rolls off smoothly instead of reaching 0: it rises with yv,
    meets yo at yv = yo, stays above 0 for yv far below, and the colour keeps
    its channel ratios"""
    rng = np.random.default_rng(12)
    rgb = rng.uniform(0.05, 1.0, size=(1, 3)).repeat(201, 0)
    yo = float(rgb[0] @ LUMA2020)
    yv = yo + np.linspace(-3.0, 1.0, 201)
    out = llap_colour_node(yv, rgb)
    ly = out @ LUMA2020
    above = yv >= yo
    assert np.allclose(ly[above], yv[above], rtol=1e-7)
    assert np.all(np.diff(ly) > 0) and np.all(ly > 0)
    below = ~above
    assert np.abs(ly[below].max() - yo) < 0.03 * yo and np.all(ly[below] <= yo + 1e-12)
    assert np.allclose(out / out[:, :1], rgb / rgb[:, :1], rtol=1e-9)


def test_llap_colour_transfer_and_black_rolloff():
    check_llap_colour_transfer_and_black_rolloff()


def test_llap_reduce_is_binomial_with_mirrored_border():
    """sample_semisoft at 2*o+.5 equals [1 2 1]/4 x [1 2 1]/4 centred on 2*o"""
    rng = np.random.default_rng(10)
    W, H = 13, 10
    inp = np.ascontiguousarray(rng.uniform(size=(2, H, W)))
    ow, oh = (W - 1) // 2 + 1, (H - 1) // 2 + 1
    out = np.zeros((2, oh, ow))
    op.node(64, "llap", "reduce", None, blob_f(0.12, 1, 1, 0), [op.orbuf(inp, op.OR_F32, 1, 2), op.orbuf(out, op.OR_F32, 1, 2)],
            ow, oh, 2, storage_f16=False)
    k = np.array([1, 2, 1]) / 4.0
    pad = np.pad(inp, ((0, 0), (1, 1), (1, 1)), mode="symmetric")
    for z in range(2):
        for y in range(oh):
            for x in range(ow):
                win = pad[z, 2 * y:2 * y + 3, 2 * x:2 * x + 3]
                assert abs(out[z, y, x] - k @ win @ k) < 1e-12


def test_llap_expand_reproduces_linear_functions():
    """sample_soft at o/2+.5: (1/2,1/2,1,1/2,1/2)/3 (even o) and (1,1/2,1/2,1)/3
    (odd o) per axis; both reproduce f(o/2) for linear f in the interior"""
    W, H = 12, 10           # coarse level
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    coarse = np.ascontiguousarray(0.2 + 0.03 * x + 0.05 * y)
    fw, fh = 2 * W - 1, 2 * H - 1
    ng = 2
    # l0 planes: zero curves (v = 0 selects planes 0/1 with weight a = 0)
    l0 = np.zeros((ng + 1, fh, fw))
    l1 = np.zeros((ng + 1, H, W))
    out = np.zeros((fh, fw))
    op.node(64, "llap", "assemble", gs.mixed_blob([("i", ng), ("i", 0)]), None,
            [op.orbuf(coarse, op.OR_F32, 1), op.orbuf(l0, op.OR_F32, 1, ng + 1), op.orbuf(l1, op.OR_F32, 1, ng + 1),
             op.orbuf(out, op.OR_F32, 1)], fw, fh, storage_f16=False)
    oy, ox = np.mgrid[0:fh, 0:fw].astype(np.float64)
    ref = 0.2 + 0.03 * ox / 2 + 0.05 * oy / 2
    assert np.allclose(out[4:-4, 4:-4], ref[4:-4, 4:-4], atol=1e-12)


def test_llap_curve_closed_form():
    """curve.comp:39-63 at the break points: identity at x = g, linear part
    beyond 2 sigma, plus the clarity bump c*exp(-3c^2/(2 sigma^2))"""
    s, sh, hi, cl = (float(np.float32(v)) for v in (0.12, 0.7, 1.3, 0.2))   # the params are float32
    ng = 10
    xs = np.array([0.5, 0.5 + 0.3, 0.5 - 0.3, 0.5 + 0.1])
    img = np.ascontiguousarray(np.stack([xs, xs, xs, np.ones_like(xs)], -1)[None])
    out = np.zeros((ng + 1, 1, 4))
    op.node(64, "llap", "curve", gs.mixed_blob([("i", ng)]), blob_f(s, sh, hi, cl),
            [op.orbuf(img, op.OR_F32, 4), op.orbuf(out, op.OR_F32, 1, ng + 1)], 4, 1, storage_f16=False)
    y = 0.2627002120 * xs + 0.6779980720 * xs + 0.0593017165 * xs   # luminance_rec2020 of gray
    assert np.allclose(out[ng, 0], y, atol=1e-8)
    g = 5 / 9.0
    bump = lambda c: cl * c * np.exp(-c * c / (2 * s * s / 3))
    for k, x in enumerate(y):
        c = x - g
        if abs(c) > 2 * s:
            lin = g + np.sign(c) * s + (sh if c > 0 else hi) * (c - np.sign(c) * s)
            assert abs(out[5, 0, k] - (lin + bump(c))) < 1e-9
    # at x == g the curve is the identity
    img2 = np.ascontiguousarray(np.array([[[g, g, g, 1.0]]]) / (0.2627002120 + 0.6779980720 + 0.0593017165))
    out2 = np.zeros((ng + 1, 1, 1))
    op.node(64, "llap", "curve", gs.mixed_blob([("i", ng)]), blob_f(s, sh, hi, cl),
            [op.orbuf(img2, op.OR_F32, 4), op.orbuf(out2, op.OR_F32, 1, ng + 1)], 1, 1, storage_f16=False)
    assert abs(out2[5, 0, 0] - out2[ng, 0, 0]) < 1e-12


@pytest.mark.parametrize("kind", ["bayer", "xtrans"])
def test_llap_identity_curves_telescope(kind):
    """clarity 0, shadows = hilights = 1: every curve is the identity, the
    laplacians telescope and llap returns its input (max(rgb, 0))"""
    W, H = 60, 42
    cfa, _ = synth.mosaic(W, H, 12, kind)
    p = develop_params(kind, clarity=0.0, shadows=1.0, hilights=1.0)
    out, st = op.develop(64, cfa, gs.oracle_params(p), stages=True, storage_f16=False)
    assert np.allclose(out[..., :3], np.maximum(st["filmcurv"][..., :3], 0), atol=1e-12)
    # a non-identity clarity must change the image (the test is not vacuous)
    p2 = develop_params(kind, clarity=0.5)
    out2 = op.develop(64, cfa, gs.oracle_params(p2), storage_f16=False)
    assert np.abs(out2 - out).max() > 1e-3


def test_float_and_double_oracles_agree():
    W, H = 96, 72
    cfa, _ = synth.mosaic(W, H, 13, "bayer")
    p = gs.oracle_params(develop_params("bayer"))
    a = op.develop(64, cfa, p)
    b = op.develop(32, cfa, p)
    mse = np.mean((a[..., :3] - b[..., :3]) ** 2)
    assert 10 * np.log10(1.0 / max(mse, 1e-30)) > 60
