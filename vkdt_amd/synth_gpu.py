"""The seeded synthetic raw scene of synth.py, generated on the GPU with torch
(same scene model: smooth log-uniform background over `stops` stops with
smooth per-channel gains, `nrect` rectangles of log-uniform radiance and
random colour, exposure set so `clip_frac` of the sites clip, Poisson-
Gaussian noise var = a*x + b on white-normalised x).  Used where gigapixels
of input are needed (batch and serving workloads); the random streams are
torch's, so the pixels differ from synth.mosaic for the same seed."""
import math

import numpy as np

from .synth import XTRANS


def channel_map(W, H, kind, device):
    import torch
    if kind == "xtrans":
        p = torch.as_tensor(XTRANS.astype(np.int64), device=device)
        ys = torch.arange(H, device=device) % 6
        xs = torch.arange(W, device=device) % 6
        return p[ys[:, None], xs[None, :]]
    ys = (torch.arange(H, device=device) & 1)[:, None]
    xs = (torch.arange(W, device=device) & 1)[None, :]
    return ys + xs   # RGGB: 0 r, 1 g, 2 b


def radiance(W, H, seed, kind="bayer", nrect=200, stops=12.0, origin=(0, 0), canvas=None, device="cuda"):
    """noise-free CFA radiance (float32 H x W) of the frame at `origin`
    inside a scene of extent `canvas` (default: the frame itself)"""
    import torch
    rng = np.random.default_rng(seed)          # scene layout: host, seeded
    CW, CH = canvas if canvas is not None else (W, H)
    ox, oy = origin
    xs = (torch.arange(W, device=device, dtype=torch.float32) + ox) / CW
    ys = (torch.arange(H, device=device, dtype=torch.float32) + oy) / CH
    ph = rng.uniform(0, 2 * np.pi, size=6)
    X, Y = xs[None, :], ys[:, None]
    # the background's range is taken over the whole canvas so that frames of
    # one scene share an exposure
    u = 0.45 * X + 0.35 * Y + 0.1 * torch.sin(3.1 * X + ph[0]) * torch.cos(2.3 * Y + ph[1])
    lo, hi = -0.1, 0.9
    u = ((u - lo) / (hi - lo)).clamp(0.0, 1.0)
    lum = torch.exp2(-stops * (1.0 - u))
    chan = channel_map(W, H, kind, device)
    g0 = 0.8 + 0.3 * torch.sin(2.0 * X + ph[2]) * torch.cos(1.7 * Y + ph[3])
    g2 = 0.8 + 0.3 * torch.cos(1.3 * X + ph[4]) * torch.sin(2.9 * Y + ph[5])
    gain = torch.where(chan == 0, g0, torch.where(chan == 2, g2, torch.ones_like(g0)))
    rad = lum * gain
    del u, lum, gain, g0, g2
    for _ in range(nrect):
        rw = int(rng.uniform(0.03, 0.25) * CW)
        rh = int(rng.uniform(0.03, 0.25) * CH)
        x0 = int(rng.uniform(0, CW - rw)) - ox
        y0 = int(rng.uniform(0, CH - rh)) - oy
        level = float(np.exp2(-stops * rng.uniform(0.0, 1.0)))
        col = rng.uniform(0.25, 1.0, size=3)
        col = col / col.max()
        xa, xb = max(x0, 0), min(x0 + rw, W)
        ya, yb = max(y0, 0), min(y0 + rh, H)
        if xa >= xb or ya >= yb:
            continue
        c = chan[ya:yb, xa:xb]
        rad[ya:yb, xa:xb] = level * torch.where(c == 0, col[0], torch.where(c == 1, col[1], col[2]))
    return rad


def clip_scale(rad, clip_frac=0.02, sample=1 << 20, seed=0):
    """radiance at the (1 - clip_frac) quantile, from a random subsample"""
    import torch
    g = torch.Generator(device=rad.device).manual_seed(seed)
    idx = torch.randint(0, rad.numel(), (min(sample, rad.numel()),), device=rad.device, generator=g)
    v = rad.reshape(-1)[idx]
    k = max(1, int(math.ceil((1.0 - clip_frac) * v.numel())))
    return float(torch.kthvalue(v.float().cpu(), k).values)


def mosaic(W, H, seed, kind="bayer", a=2e-4, b=1e-6, black=512, white=16383, nrect=200, stops=12.0,
           clip_frac=0.02, origin=(0, 0), canvas=None, scale=None, device="cuda"):
    """(cfa, scale): the uint16 CFA as a torch int16 tensor (H x W, values in
    [0, white] < 2^15, so its bits are the uint16 mosaic) on `device`"""
    import torch
    rad = radiance(W, H, seed, kind, nrect, stops, origin, canvas, device)
    if scale is None:
        scale = clip_scale(rad, clip_frac, seed=seed)
    x = rad / scale
    del rad
    g = torch.Generator(device=device).manual_seed(seed ^ 0x5EED5EED)
    n = torch.randn(x.shape, device=device, generator=g, dtype=torch.float32)
    x = x + n * torch.sqrt(a * x.clamp(min=0.0) + b)
    del n
    raw = torch.round(black + x * float(white - black)).clamp(0, white)
    return raw.to(torch.int16), scale
