"""Pins of the oracle's denoise wavelet and demosaic weights by what the
mathematics fixes for hand-built inputs: impulse responses, linear limits,
closed-form single-band shrinkage, permutation round trips, flat-field
tensors and library routines (numpy's weighted covariance and eigh).

Each check feeds one node of the oracle a synthetic input whose output has
a closed form written here from the GLSL's definitions, independent of the
oracle's loops.  The checks are plain functions so that
tests/test_oracle_mutations.py can run them against oracle builds with a
planted mistake and show that each mistake fails one of them.
"""
import numpy as np
import pytest

import graphspec as gs
import oracle_py as op

BAYER = 0x94949494
XTRANS = 9
BLOCK2 = 1.414213          # down.comp:84 / assemble.comp:89, Bayer block factor
FLOWER = ((1.2, 0.4), (-1.2, -0.4), (0.4, -1.2), (-0.4, 1.2))   # down.comp:96-99


def blob(items):
    return gs.mixed_blob(items)


def rgbc(wb=(1, 1, 1, 1), black=0.0, white=1.0, crop=(0, 0, 0, 0)):
    return [("f", v) for v in wb] + [("f", black)] * 4 + [("f", white)] * 4 + [("i", c) for c in crop]


def dn_par(strength=0.5, luma=1.0, detail=1.0, edges=(0, 0, 0, 0)):
    return blob([("f", strength), ("f", luma), ("f", detail), ("f", 0)] + [("f", e) for e in edges] + [("i", 1)])


def hat(t):
    return np.maximum(0.0, 1.0 - np.abs(t))


def sorted_index(n):
    """down.comp:102-104: texel i is stored at i/2 + (i&1)*(n+1)/2"""
    i = np.arange(n)
    return i // 2 + (i & 1) * ((n + 1) // 2)


def polyphase_sort(img):
    """the store permutation of down/downcov applied to an (H, W, ...) image"""
    H, W = img.shape[:2]
    out = np.zeros_like(img)
    out[np.ix_(sorted_index(H), sorted_index(W))] = img
    return out


def run_down(img, level, noise_a, block=2, noise_b=0.0):
    H, W = img.shape[:2]
    img = np.ascontiguousarray(img, np.float64)
    out = np.zeros((H, W, 4))
    pc = blob(rgbc() + [("f", noise_a), ("f", noise_b), ("i", level), ("i", block)])
    op.node(64, "denoise", "down", pc, dn_par(), [op.orbuf(img, op.OR_F32, 4), op.orbuf(out, op.OR_F32, 4)],
            W, H, storage_f16=False)
    return out


# ---------------------------------------------------------------------------
# denoise/down.comp:59-108
# ---------------------------------------------------------------------------

def check_down_impulse_response():
    """With sigma at its cap (noise_sigma clamps to 1e3) and values of 4e-3,
    every edge weight is 1 - O(1e-5): down is the linear filter
    0.2 * centre + 0.2 * sum over the four flower taps of the bilinear
    sample.  The impulse response is then a sum of bilinear hats centred on
    the tap offsets, stored polyphase-sorted.  Fixes the tap positions
    (including their handedness), the centre and tap weights and the
    sorted store."""
    W, H = 21, 17
    q = (10, 8)
    h = 0.004
    img = np.zeros((H, W, 4))
    img[..., 3] = 1
    img[q[1], q[0], :3] = h
    out = run_down(img, 0, 1e6)
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    resp = 0.2 * ((x == q[0]) & (y == q[1]))
    for ox, oy in FLOWER:
        resp = resp + 0.2 * hat(q[0] - (x + ox)) * hat(q[1] - (y + oy))
    ref = polyphase_sort(h * resp)
    assert np.abs(out[..., :3] - ref[..., None]).max() < 2e-5 * h
    assert np.all(out[..., 3] == 1)


def check_down_edge_stopping_levels():
    """A column of value B in a field of A, noise sigma sqrt(noise_a) (noise_b
    0, edges 0).  Left of the column the taps at +1.2 and +0.4 see
    0.8B+0.2A and 0.4B+0.6A; their weights are
    clamp(1 - 0.5|c^0.8 - A^0.8| / s_l), s_l = 0.7^level sqrt(noise_a)/sqrt(2),
    the other taps see A.  Checked at levels 0, 1 and 2."""
    W, H, x0 = 16, 12, 8
    A, B = 0.2, 0.26
    img = np.zeros((H, W, 4))
    img[..., :3] = A
    img[:, x0, :3] = B
    img[..., 3] = 1
    g = lambda v: v ** 0.8
    for level in (0, 1, 2):
        s = 0.7 ** level * np.sqrt(0.01) / BLOCK2
        out = run_down(img, level, 0.01)
        col = np.full(W, A)
        col[x0] = B

        def bil(xf):           # rows are identical: bilinear along x only
            i0 = int(np.floor(xf))
            f = xf - i0
            return col[i0] * (1 - f) + col[i0 + 1] * f

        ref = np.zeros(W)
        for x in range(3, W - 3):
            c0 = col[x]
            num, den = 0.2 * c0, 0.2
            for ox, _ in FLOWER:
                c = bil(x + ox)
                e = min(1.0, max(0.0, 1.0 - 0.5 * abs(g(c) - g(c0)) / s))
                num += 0.2 * e * c
                den += 0.2 * e
            ref[x] = num / den
        sx = sorted_index(W)
        for y in (3, 6):
            sy = sorted_index(H)[y]
            got = out[sy, sx[3:W - 3], 0]
            assert np.abs(got - ref[3:W - 3]).max() < 1e-9, (level, got, ref[3:W - 3])
        # the test is not vacuous: the column's neighbours are partly stopped
        assert ref[x0 - 1] != pytest.approx((0.6 * A + 0.2 * (0.8 * B + 0.2 * A) + 0.2 * (0.4 * B + 0.6 * A)), abs=1e-6)


# ---------------------------------------------------------------------------
# denoise/downcov.comp + cov.glsl:21-139
# ---------------------------------------------------------------------------

def run_downcov(img):
    H, W = img.shape[:2]
    img = np.ascontiguousarray(img, np.float64)
    out, cov = np.zeros((H, W, 4)), np.zeros((H, W, 4))
    pc = blob(rgbc() + [("f", 1e-4), ("f", 0.0), ("i", 0), ("i", 2)])
    op.node(64, "denoise", "downcov", pc, dn_par(),
            [op.orbuf(img, op.OR_F32, 4), op.orbuf(out, op.OR_F32, 4), op.orbuf(cov, op.OR_F32, 4)],
            W, H, storage_f16=False)
    return out, cov


def check_downcov_flat_field_and_hot_pixel():
    """Flat field: both tensors are the 5x5 grid's covariance 2*I, so the
    eigenvalues are (2, 2) * (1, 0.05); the edge mask
    clamp(0.02 + smoothstep(.4, .75, clamp(75 (0.1 - 0.09)))) is 1 and the
    output equals the input.  A hot pixel (above twice the 5x5 mean) is left
    out of every sum that sees it: its neighbours and itself return the
    flat value exactly."""
    W, H = 13, 11
    img = np.zeros((H, W, 4))
    img[..., :3] = 0.3
    img[..., 3] = 1
    out, cov = run_downcov(img)
    assert np.allclose(out[..., :3], 0.3, rtol=0, atol=1e-15)
    assert np.allclose(out[..., 3], 1.0, atol=1e-15)
    assert np.allclose(cov[..., 0], 2.0, atol=1e-12) and np.allclose(cov[..., 1], 0.1, atol=1e-12)
    hot = img.copy()
    hot[5, 6, :3] = 0.9
    out, _ = run_downcov(hot)
    assert np.allclose(out[..., :3], 0.3, rtol=0, atol=1e-14)


def check_downcov_tensor_against_numpy():
    """Random luminance patches: the bright tensor is the px^2-weighted
    covariance of the 5x5 offsets, the dark one the px^-2-weighted one
    (means weighted by px and 1/px), the smaller determinant wins, and the
    output is the anisotropic gaussian average with the eigen decomposition
    from numpy (eigenvalues scaled by (1, 0.05), clamped to [0.01, 25]);
    the edge mask follows the minor eigenvalue."""
    rng = np.random.default_rng(41)
    W, H = 9, 9
    img = np.zeros((H, W, 4))
    img[..., :3] = rng.uniform(0.2, 0.6, size=(H, W, 3))
    img[..., 3] = 1
    out, cov = run_downcov(img)
    L = img[..., 0] * 2.62700212e-01 + img[..., 1] * 6.77998072e-01 + img[..., 2] * 5.93017165e-02
    jj, ii = np.mgrid[-2:3, -2:3]
    P = np.stack([ii.ravel(), jj.ravel()]).astype(np.float64)
    for y in range(2, H - 2):
        for x in range(2, W - 2):
            px = L[y - 2:y + 3, x - 2:x + 3].ravel()
            mw = (P * px).sum(1) / px.sum()
            mb = (P / px).sum(1) / (1 / px).sum()
            Sw = np.cov(P, aweights=px * px, bias=True) + np.outer(P @ (px * px) / (px * px).sum() - mw,
                                                                    P @ (px * px) / (px * px).sum() - mw)
            Sb = np.cov(P, aweights=1 / px ** 2, bias=True) + np.outer(
                P @ px ** -2 / (px ** -2).sum() - mb, P @ px ** -2 / (px ** -2).sum() - mb)
            S = Sw if np.linalg.det(Sw) < np.linalg.det(Sb) else Sb
            ev, V = np.linalg.eigh(S)          # ascending
            e_maj, e_min = ev[1], ev[0]
            v_maj, v_min = V[:, 1], V[:, 0]
            lam = np.clip([e_maj * 1.0, e_min * 0.05], 0.01, 25.0)
            t = P.T
            q = (t @ v_maj) ** 2 / lam[0] + (t @ v_min) ** 2 / lam[1]
            w = np.maximum(1e-9, np.exp(-0.5 * q))
            keep = img[y - 2:y + 3, x - 2:x + 3, 0].ravel() <= 2 * px.mean()
            rgb = (w[keep, None] * img[y - 2:y + 3, x - 2:x + 3, :3].reshape(-1, 3)[keep]).sum(0) / w[keep].sum()
            sy, sx = sorted_index(H)[y], sorted_index(W)[x]
            assert np.allclose(out[sy, sx, :3], rgb, rtol=1e-9, atol=1e-12), (x, y)
            assert np.allclose(cov[y, x, :2], lam, rtol=1e-9)
            edge = np.clip(75 * max(0.0, lam[1] - 0.09), 0, 1)
            t2 = np.clip((edge - 0.4) / 0.35, 0, 1)
            edge = np.clip(0.02 + t2 * t2 * (3 - 2 * t2), 0, 1)
            assert abs(out[sy, sx, 3] - edge) < 1e-9


# ---------------------------------------------------------------------------
# denoise/assemble.comp:43-165
# ---------------------------------------------------------------------------

def run_assemble(s, strength, noise_a, luma=1.0, filters=BAYER):
    H, W = s[0].shape[:2]
    bufs = [np.ascontiguousarray(a, np.float64) for a in s]
    out = np.zeros((H, W, 4))
    pc = blob(rgbc() + [("f", noise_a), ("f", 0.0), ("i", filters)])
    op.node(64, "denoise", "assemble", pc, dn_par(strength, luma),
            [op.orbuf(b, op.OR_F32, 4) for b in bufs] + [op.orbuf(out, op.OR_F32, 4)], W, H, storage_f16=False)
    return out


def shrink(d, thr):
    """assemble.comp:134-150: the per-band soft/hard shrinkage"""
    tt = min(1.0, abs(d) / (2 * thr))
    return np.sign(d) * ((1 - tt) * max(abs(d) - thr, 0.0) + tt * abs(d))


def check_assemble_single_band_shrinkage():
    """Scales s0..s_l equal to A and s_(l+1)..s4 equal to C make band l the
    only non-zero one: d_l = (A - C)/(sigma b_l), b_l = 0.7^(l+1)/sqrt(2)
    (Bayer).  The result is C + sigma b_l shrink(d_l), shrink with
    thr = strength^4 (or 10000 once |d_l.x| > 10), back through the YUV
    round trip (luma 1, unit white balance and range); the detail mask is
    the slope (11/36)(L3 - L0) + (1/12)(L2 - L1) of the band lengths."""
    W, H = 7, 5
    sigma = 0.1                    # sqrt(noise_a)
    strength = 1.2
    thr = strength ** 4
    for l in range(4):
        b = 0.7 ** (l + 1) / BLOCK2
        for dmag in (0.5 * thr, 1.5 * thr, 7.0, 12.0):    # soft part, linear part, below/above the hard threshold
            C = 0.4
            A = C + dmag * sigma * b
            s = []
            for k in range(5):
                a = np.zeros((H, W, 4))
                a[..., :3] = A if k <= l else C
                a[..., 3] = 1
                s.append(a)
            out = run_assemble(s, strength, sigma * sigma)
            t = 10000.0 if dmag > 10 else thr
            ref = C + sigma * b * shrink(dmag, t)
            assert np.allclose(out[..., :3], ref, rtol=0, atol=3e-6), (l, dmag, out[0, 0, :3], ref)
            Ls = np.zeros(4)
            Ls[l] = np.sqrt(3) * dmag
            slope = 11 / 36 * (Ls[3] - Ls[0]) + 1 / 12 * (Ls[2] - Ls[1])
            test = np.clip(1.5 * max(0.0, 1 - max(0.0, -slope)) ** 16, 0, 1)
            assert np.allclose(out[..., 3], test, atol=1e-9), (l, dmag, out[0, 0, 3], test)


def check_assemble_walk_inverts_the_sort():
    """If every down level were the identity filter, s_(k+1) would be the
    polyphase sort of s_k (down.comp:102-104).  assemble walks each texel to
    where the sorts put it (assemble.comp:55-70), so all bands vanish and it
    returns s0 (up to the 1e-6 of the YUV round trip) with the detail mask
    at 1.  Odd extents, where (n+1)/2 differs from n/2."""
    rng = np.random.default_rng(43)
    W, H = 23, 19
    s0 = np.zeros((H, W, 4))
    s0[..., :3] = rng.uniform(0.2, 0.8, size=(H, W, 3))
    s0[..., 3] = 1
    s = [s0]
    for _ in range(4):
        s.append(polyphase_sort(s[-1]))
    out = run_assemble(s, 1.0, 1e-4)
    assert np.abs(out[..., :3] - s0[..., :3]).max() < 3e-6
    assert np.allclose(out[..., 3], 1.0)
    # a walk that misses (here: an unsorted s2) leaves bands behind
    s_bad = list(s)
    s_bad[2] = s[1]
    bad = run_assemble(s_bad, 1.0, 1e-4)
    assert np.abs(bad[..., :3] - s0[..., :3]).max() > 1e-2


def run_doub(raw, up, dn, filters, strength, luma, detail, noise_a, edges):
    H, W = raw.shape
    blk = 3 if filters == XTRANS else 2
    r = np.ascontiguousarray(raw, np.uint16)
    c0 = np.ascontiguousarray(np.broadcast_to(up, (H // blk, W // blk, 4)), np.float64)
    c1 = np.ascontiguousarray(np.broadcast_to(dn, (H // blk, W // blk, 4)), np.float64)
    out = np.zeros((H, W, 1))
    pc = blob(rgbc(black=512 / 65535.0, white=16383 / 65535.0, crop=(0, 0, W, H))
              + [("i", filters), ("f", noise_a), ("f", 0.0), ("i", 0)] + [("f", 0.0)] * 4)
    op.node(64, "denoise", "doub", pc, dn_par(strength, luma, detail, edges),
            [op.orbuf(r, op.OR_UI16, 1), op.orbuf(c0, op.OR_F32, 4), op.orbuf(c1, op.OR_F32, 4),
             op.orbuf(out, op.OR_F32, 1)], W, H, storage_f16=False)
    return out[..., 0]


def check_doub_closed_form():
    """Constant coarse images make the bilinear upsampling exact, so every
    mosaic site v (below white) follows doub.comp:78-104 in closed form with
    its own CFA colour k (red, green or blue of the canonical layouts):
    wav = (v - down_k)/(sigma_r + sigma_b), T = strength/2 * w (X-Trans red/blue:
    divided by max(1e-4, w)), tt = min(1, wav/(2T)), protection
    p = 1 - (1 - min(1, w)^4) detail, blend b = (1 + (luma on green, else 1))/2,
    v' = mix(v, max(0, up_k + sigma_g sign(wav) mix(max(0, |wav| - T), |wav|, tt)), p b),
    then max(0, (v' - black)/(white - black)).  Sites at or above white skip
    the shrink.  The noise sigma per channel is sqrt(noise_a) times the edge
    gain exp2(12 edges_c + edges_a) (noise.glsl:9)."""
    from vkdt_amd import synth
    rng = np.random.default_rng(47)
    black, white = np.float32(512 / 65535.0), np.float32(16383 / 65535.0)
    noise_a = 1e-4
    edges = (0.05, -0.05, 0.1, 0.02)
    e32 = [float(np.float32(e)) for e in edges]     # the push constants and params are float32
    sig = [2.0 ** (12 * e32[c] + e32[3]) * float(np.sqrt(np.float32(noise_a))) for c in range(3)]
    up = np.array([0.10, 0.12, 0.08, 0.7])
    dn = np.array([0.105, 0.118, 0.083, 1.0])
    strength, luma, detail = (float(np.float32(v)) for v in (0.6, 0.3, 0.8))
    for kind, filters in (("bayer", BAYER), ("xtrans", XTRANS)):
        W, H = 24, 18
        raw = rng.integers(5000, 9000, size=(H, W), dtype=np.uint16)
        raw[3, 4] = 16383
        raw[7, 9] = 17000
        out = run_doub(raw, up, dn, filters, strength, luma, detail, noise_a, edges)
        cm = synth.channel_map(W, H, kind)
        v = np.float32(raw / 65535.0).astype(np.float64)
        w = up[3]
        for y in range(H):
            for x in range(W):
                k = cm[y, x]
                val = v[y, x]
                if val < white:
                    T = 0.5 * strength * w
                    if kind == "xtrans" and k != 1:
                        T /= max(1e-4, w)
                    wav = (val - dn[k]) / (sig[0] + sig[2])
                    tt = min(1.0, wav / max(2 * T, 1e-8))
                    p = 1 - (1 - min(1.0, w) ** 4) * detail
                    b = 0.5 * ((luma if k == 1 else 1.0) + 1)
                    sh = (1 - tt) * max(0.0, abs(wav) - T) + tt * abs(wav)
                    val = val * (1 - p * b) + max(0.0, up[k] + sig[1] * np.sign(wav) * sh) * p * b
                ref = max(0.0, (val - float(black)) / (float(white) - float(black)))
                assert abs(out[y, x] - ref) < 1e-9, (kind, x, y, k, out[y, x], ref)


# ---------------------------------------------------------------------------
# demosaic/gauss.comp:19-126
# ---------------------------------------------------------------------------

def run_gauss(cfa, filters):
    H, W = cfa.shape
    blk = 3 if filters == XTRANS else 2
    inp = np.ascontiguousarray(cfa, np.float64)
    out = np.zeros((H // blk, W // blk, 4))
    pc = blob([("f", 1.0)] * 4 + [("i", filters)])
    op.node(64, "demosaic", "gauss", pc, blob([("i", 0), ("i", 0)]),
            [op.orbuf(inp, op.OR_F32, 1), op.orbuf(inp, op.OR_F32, 1), op.orbuf(out, op.OR_F32, 4)],
            W // blk, H // blk, storage_f16=False)
    return out


def check_gauss_flat_field():
    """Flat greens: both tensors are the plain covariance of the sample
    positions.  Bayer: the 8 greens of the 4x4 window (-1..2) have
    covariance [[1.25, -.25], [-.25, 1.25]], eigenvalues 1.5 and 1, scaled by
    0.2 -> (0.3, 0.2); the major axis is the diagonal, where the h/v snap
    is a tie that rounding decides, so only the eigenvalues are fixed.
    X-Trans: the 5 greens of the 3x3 block have covariance 0.8*I, scaled by
    0.4 -> (0.32, 0.32), axis (0, 1)."""
    out = run_gauss(np.full((16, 20), 0.4), BAYER)
    assert np.allclose(out[1:-1, 1:-1, :2], [0.3, 0.2], atol=1e-12)
    out = run_gauss(np.full((18, 21), 0.4), XTRANS)
    assert np.allclose(out, [0.32, 0.32, 0.0, 1.0], atol=1e-12)


def check_gauss_tensor_against_numpy():
    """Random greens: weighted covariances (px^2 and px^-2) of the green
    positions from numpy, smaller determinant, eigen decomposition from
    numpy, eigenvalues x0.2 (Bayer) / x0.4 (X-Trans), major axis snapped to
    h/v (X-Trans: to h/v when one component is twice the other, else
    eigenvalues (.1, .1))."""
    rng = np.random.default_rng(44)
    for filters, blk, rng_ij, par, scale in ((BAYER, 2, range(-1, 3), 1, 0.2), (XTRANS, 3, range(0, 3), 0, 0.4)):
        W, H = 8 * blk, 6 * blk
        cfa = rng.uniform(0.1, 0.9, size=(H, W))
        out = run_gauss(cfa, filters)
        pos = [(i, j) for j in rng_ij for i in rng_ij if ((i + j) & 1) == par]
        P = np.array(pos, np.float64).T
        for y in range(1, H // blk - 1):
            for x in range(1, W // blk - 1):
                px = np.array([cfa[blk * y + j, blk * x + i] for i, j in pos])
                Sw = np.cov(P, aweights=px * px, bias=True)
                Sb = np.cov(P, aweights=px ** -2, bias=True)
                mw, mb = (P * px).sum(1) / px.sum(), (P / px).sum(1) / (1 / px).sum()
                cw, cb = (P * px * px).sum(1) / (px * px).sum(), (P / px ** 2).sum(1) / (px ** -2).sum()
                Sw = Sw + np.outer(cw - mw, cw - mw)
                Sb = Sb + np.outer(cb - mb, cb - mb)
                S = Sw if np.linalg.det(Sw) < np.linalg.det(Sb) else Sb
                ev, V = np.linalg.eigh(S)
                lam = np.array([ev[1], ev[0]]) * scale
                # evec0 is the major axis (evd2x2: evec1 spans the minor one, evec0 is its normal)
                a = np.abs(V[:, 1])
                if filters == BAYER:
                    ax = [1.0, 0.0] if a[0] > a[1] else [0.0, 1.0]
                elif a[0] > 2 * a[1]:
                    ax = [1.0, 0.0]
                elif a[1] > 2 * a[0]:
                    ax = [0.0, 1.0]
                else:
                    ax, lam = out[y, x, 2:4], np.array([0.1, 0.1])
                assert np.allclose(out[y, x, :2], lam, rtol=1e-9), (filters, x, y, out[y, x], lam)
                assert np.allclose(out[y, x, 2:4], ax), (filters, x, y)


# ---------------------------------------------------------------------------
# demosaic/splat.comp:17-140 and fix.comp:36-135 (Bayer)
# ---------------------------------------------------------------------------

def rot_weight(o, lam, theta, lo, scale=1.0):
    """exp(-1/2 |E o|^2_lam): E = mat2(c, -s, s, c) (column major), so
    (E o) = (c ox + s oy, -s ox + c oy); clamped below by `lo`, above by 1"""
    c, s = np.cos(theta), np.sin(theta)
    fx, fy = c * o[0] + s * o[1], -s * o[0] + c * o[1]
    return min(1.0, max(lo, np.exp(-0.5 * (fx * fx / (scale * lam[0]) + fy * fy / (scale * lam[1])))))


def run_splat(cfa, cov):
    H, W = cfa.shape
    inp = np.ascontiguousarray(cfa, np.float64)
    cv = np.ascontiguousarray(np.broadcast_to(cov, (H // 2 + 1, W // 2 + 1, 4)), np.float64)
    out = np.zeros((H, W, 1))
    op.node(64, "demosaic", "splat", blob([("f", 1.0)] * 4 + [("i", BAYER)]), blob([("i", 0), ("i", 0)]),
            [op.orbuf(inp, op.OR_F32, 1), op.orbuf(cv, op.OR_F32, 4), op.orbuf(out, op.OR_F32, 1)], W, H,
            storage_f16=False)
    return out[..., 0]


def check_splat_bayer_weights():
    """Random greens.  Red/blue sites average their four edge neighbours, a
    green site weighs itself 666 and its four diagonal greens; every other
    tap weighs exp(-1/2 |E o|^2_lam) (lam clamped to [0.01, 25], clamped
    below by 1e-4).  A linear ramp of green is reproduced at every interior
    red and blue site."""
    rng = np.random.default_rng(46)
    W, H = 12, 10
    yy, xx = np.mgrid[0:H, 0:W]
    green = ((xx + yy) & 1) == 1
    cfa = np.where(green, rng.uniform(0.2, 0.8, size=(H, W)), 0.9)
    for lam, theta in (((0.3, 0.2), 0.0), ((40.0, 0.5), 0.5236), ((0.005, 0.7), 1.1)):
        lc = np.clip(lam, 0.01, 25)
        out = run_splat(cfa, [lam[0], lam[1], np.cos(theta), np.sin(theta)])
        w = lambda i, j: rot_weight((i, j), lc, theta, 1e-4)
        for y in range(2, H - 2):
            for x in range(2, W - 2):
                taps = [(i, j) for j in (-1, 0, 1) for i in (-1, 0, 1) if green[y + j, x + i]]
                ws = [666.0 if (i, j) == (0, 0) else w(i, j) for i, j in taps]
                ref = sum(wt * cfa[y + j, x + i] for wt, (i, j) in zip(ws, taps)) / sum(ws)
                assert abs(out[y, x] - ref) < 1e-12, (lam, theta, x, y, out[y, x], ref)
    ramp = np.where(green, 0.2 + 0.01 * xx + 0.02 * yy, 0.9)
    out = run_splat(ramp, [0.3, 0.2, 1.0, 0.0])
    ref = 0.2 + 0.01 * xx + 0.02 * yy
    assert np.allclose(out[2:-2, 2:-2][~green[2:-2, 2:-2]], ref[2:-2, 2:-2][~green[2:-2, 2:-2]], atol=1e-12)


def run_fix(cfa, green, cov, colour=0):
    H, W = cfa.shape
    inp = np.ascontiguousarray(cfa, np.float64)
    g = np.ascontiguousarray(green, np.float64)
    cv = np.ascontiguousarray(np.broadcast_to(cov, (H // 2 + 1, W // 2 + 1, 4)), np.float64)
    out = np.zeros((H, W, 4))
    op.node(64, "demosaic", "fix", blob([("f", 1.0)] * 4 + [("i", BAYER)]), blob([("i", colour), ("i", 0)]),
            [op.orbuf(inp, op.OR_F32, 1), op.orbuf(g, op.OR_F32, 1), op.orbuf(cv, op.OR_F32, 4),
             op.orbuf(out, op.OR_F32, 4)], W, H, storage_f16=False)
    return out


def check_fix_bayer_weights_and_ratio():
    """Constant green: red at a blue site is the average of the four diagonal
    reds weighted by exp(-1/2 |E o|^2 / (2 lam)), lam clamped to [1, 49];
    at green sites the two reds of the row (or column) are averaged.  With
    reds k (1e-4 + g) over any green field the ratio interpolation returns
    k (1e-4 + g_centre) exactly."""
    rng = np.random.default_rng(45)
    W, H = 12, 10
    yy, xx = np.mgrid[0:H, 0:W]
    red = (xx % 2 == 0) & (yy % 2 == 0)
    blue = (xx % 2 == 1) & (yy % 2 == 1)
    r = rng.uniform(0.2, 0.8, size=(H, W))
    cfa = np.where(red, r, 0.5)
    g = np.full((H, W), 0.5)
    for lam, theta in (((3.0, 1.5), 0.5236), ((0.5, 60.0), 0.9)):
        lc = np.clip(lam, 1.0, 49.0)
        out = run_fix(cfa, g, [lam[0], lam[1], np.cos(theta), np.sin(theta)])
        w = lambda i, j: rot_weight((i, j), lc, theta, 1e-3, scale=2.0)
        for y in range(2, H - 2):
            for x in range(2, W - 2):
                if blue[y, x]:
                    nb = [(1, 1), (-1, -1), (1, -1), (-1, 1)]
                    ref = sum(w(i, j) * r[y + j, x + i] for i, j in nb) / sum(w(i, j) for i, j in nb)
                elif red[y, x]:
                    ref = r[y, x]
                elif y % 2 == 0:
                    ref = 0.5 * (r[y, x - 1] + r[y, x + 1])
                else:
                    ref = 0.5 * (r[y - 1, x] + r[y + 1, x])
                assert abs(out[y, x, 0] - ref) < 1e-12, (lam, x, y, out[y, x, 0], ref)
                assert out[y, x, 1] == 0.5
    k = 0.7
    g = rng.uniform(0.1, 0.9, size=(H, W))
    cfa = np.where(red | blue, k * (1e-4 + g), g)
    out = run_fix(cfa, g, [2.0, 1.0, 1.0, 0.0])
    ref = k * (1e-4 + g)
    assert np.allclose(out[1:-1, 1:-1, 0], ref[1:-1, 1:-1], rtol=1e-12)
    assert np.allclose(out[1:-1, 1:-1, 2], ref[1:-1, 1:-1], rtol=1e-12)


CHECKS = [
    check_down_impulse_response,
    check_down_edge_stopping_levels,
    check_downcov_flat_field_and_hot_pixel,
    check_downcov_tensor_against_numpy,
    check_assemble_single_band_shrinkage,
    check_assemble_walk_inverts_the_sort,
    check_doub_closed_form,
    check_gauss_flat_field,
    check_gauss_tensor_against_numpy,
    check_splat_bayer_weights,
    check_fix_bayer_weights_and_ratio,
]


@pytest.mark.parametrize("check", CHECKS, ids=[c.__name__[6:] for c in CHECKS])
def test_stencil_pin(check):
    check()
