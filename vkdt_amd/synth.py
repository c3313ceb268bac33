"""Seeded synthetic raw scenes and the uncompressed DNGs that carry them.

Scene model (the workloads of the raw-development benchmark):
  * a smooth background whose luminance spans ``stops`` stops log-uniformly,
    with smoothly varying per-channel gains,
  * ``nrect`` axis-aligned rectangles of log-uniform radiance over the same
    range and random colour,
  * scaled so that ``clip_frac`` of the CFA sites exceed the white level,
  * Poisson-Gaussian noise on the white-normalised signal x:
    var = a * x + b, then quantised to the raw integer range and clipped
    at [0, white].

vkdt's noise model evaluates sigma = sqrt(noise_a + v * noise_b) on
black-subtracted UNORM16 values (denoise/noise.glsl:9) with noise_a the
gaussian and noise_b the poissonian term (i-raw/readme.md:9-10), so the
matching graph parameters are ``noise_params(a, b, black, white)``.
"""
import numpy as np

# canonical x-trans layout after i-raw's offset normalisation
# (demosaic/splat.comp:103-110): 0 red, 1 green, 2 blue
XTRANS = np.array([
    [1, 0, 1, 1, 2, 1],
    [2, 1, 2, 0, 1, 0],
    [1, 0, 1, 1, 2, 1],
    [1, 2, 1, 1, 0, 1],
    [0, 1, 0, 2, 1, 2],
    [1, 2, 1, 1, 0, 1],
], dtype=np.uint8)
BAYER_RGGB = np.array([[0, 1], [1, 2]], dtype=np.uint8)

BAYER_FILTERS = 0x49494949   # i-raw's "any bayer" marker after RGGB normalisation
XTRANS_FILTERS = 9

# fixed camera -> XYZ(D65) matrix of the synthetic camera (row major); the
# DNG stores its inverse as ColorMatrix1 with CalibrationIlluminant1 = D65
CAM_TO_XYZ = np.array([
    [0.6280, 0.2310, 0.1020],
    [0.2650, 0.6850, 0.0500],
    [0.0230, 0.1180, 0.9550],
], dtype=np.float64)


def xyz_to_cam():
    return np.linalg.inv(CAM_TO_XYZ)


def cfa_pattern(kind):
    return XTRANS if kind == "xtrans" else BAYER_RGGB


def channel_map(W, H, kind):
    """CFA colour (0 r, 1 g, 2 b) of every site, uint8 H x W."""
    p = cfa_pattern(kind)
    n = p.shape[0]
    reps = (-(-H // n), -(-W // n))
    return np.tile(p, reps)[:H, :W]


def noise_params(a, b, black=512.0, white=16383.0):
    """(noise_a, noise_b) in vkdt units for var = a*x + b on white-normalised x."""
    s = (white - black) / 65535.0
    return b * s * s, a * s * s


def radiance(W, H, seed, kind="bayer", nrect=200, stops=12.0, origin=(0, 0), canvas=None):
    """Noise-free CFA radiance (float32 H x W, arbitrary scale).

    ``origin``/``canvas`` place the W x H frame inside a larger scene of
    extent ``canvas`` (for video streams that drift across one scene)."""
    rng = np.random.default_rng(seed)
    CW, CH = canvas if canvas is not None else (W, H)
    ox, oy = origin
    xs = (np.arange(W, dtype=np.float32) + ox) / np.float32(CW)
    ys = (np.arange(H, dtype=np.float32) + oy) / np.float32(CH)
    ph = rng.uniform(0, 2 * np.pi, size=6).astype(np.float32)
    u = (0.45 * xs[None, :] + 0.35 * ys[:, None]
         + 0.1 * np.sin(3.1 * xs[None, :] + ph[0]) * np.cos(2.3 * ys[:, None] + ph[1]))
    u = (u - u.min()) / max(float(u.max() - u.min()), 1e-6)
    lum = np.exp2(-stops * (1.0 - u)).astype(np.float32)
    chan = channel_map(W, H, kind)
    gains = np.stack([
        0.8 + 0.3 * np.sin(2.0 * xs[None, :] + ph[2]) * np.cos(1.7 * ys[:, None] + ph[3]),
        np.ones((1, 1), np.float32) + 0 * xs[None, :] + 0 * ys[:, None],
        0.8 + 0.3 * np.cos(1.3 * xs[None, :] + ph[4]) * np.sin(2.9 * ys[:, None] + ph[5]),
    ]).astype(np.float32)
    rad = lum * np.take_along_axis(gains, chan[None].astype(np.intp), 0)[0]
    for _ in range(nrect):
        rw = int(rng.uniform(0.03, 0.25) * CW)
        rh = int(rng.uniform(0.03, 0.25) * CH)
        x0 = int(rng.uniform(0, CW - rw)) - ox
        y0 = int(rng.uniform(0, CH - rh)) - oy
        level = np.float32(np.exp2(-stops * rng.uniform(0.0, 1.0)))
        col = rng.uniform(0.25, 1.0, size=3).astype(np.float32)
        col /= col.max()
        xa, xb = max(x0, 0), min(x0 + rw, W)
        ya, yb = max(y0, 0), min(y0 + rh, H)
        if xa >= xb or ya >= yb:
            continue
        rad[ya:yb, xa:xb] = level * col[chan[ya:yb, xa:xb]]
    return rad


def mosaic(W, H, seed, kind="bayer", a=2e-4, b=1e-6, black=512, white=16383,
           nrect=200, stops=12.0, clip_frac=0.02, origin=(0, 0), canvas=None, scale=None):
    """uint16 CFA of a seeded scene with Poisson-Gaussian noise.

    Returns (cfa, scale) where ``scale`` is the radiance that maps to the
    white level (pass it back in to keep exposure fixed across video frames)."""
    rad = radiance(W, H, seed, kind, nrect, stops, origin, canvas)
    if scale is None:
        k = int((1.0 - clip_frac) * rad.size)
        scale = float(np.partition(rad.ravel(), k)[k])
    x = rad / np.float32(scale)
    rng = np.random.default_rng(seed ^ 0x5EED5EED)
    noise = rng.standard_normal(size=x.shape, dtype=np.float32)
    x = x + noise * np.sqrt(np.float32(a) * np.maximum(x, 0) + np.float32(b))
    raw = np.rint(black + x * np.float32(white - black))
    return np.clip(raw, 0, white).astype(np.uint16), scale


# ---------------------------------------------------------------------------
# uncompressed DNG writer (little endian TIFF, one raw IFD)
# ---------------------------------------------------------------------------

def _rational(v, signed):
    from fractions import Fraction
    f = Fraction(float(v)).limit_denominator(1_000_000)
    if not signed and f < 0:
        raise ValueError("negative unsigned rational")
    return f.numerator, f.denominator


def write_dng(path, cfa, kind="bayer", black=512, white=16383, as_shot_neutral=(1.0, 1.0, 1.0),
              color_matrix=None, make="vkdt-synth", model="synthetic", pattern=None, illuminant=21,
              extra_tags=(), compress=None):
    """Write ``cfa`` (uint16 H x W) as a CFA DNG, uncompressed or (``compress``
    a dict with tile=(wd, ht), components, predictor, bits) as lossless-JPEG
    tiles (Compression 7), edge tiles padded by repeating the last row/column.

    ``pattern`` overrides the CFA layout (a 2x2 or 6x6 array of 0/1/2),
    ``illuminant`` is CalibrationIlluminant1 (21 = D65), ``extra_tags`` are
    (tag, type, values) triples appended to the IFD (for negative tests)."""
    import struct
    H, W = cfa.shape
    chunks = None
    if compress is not None:
        from . import lj92
        tw, th = compress.get("tile", (256, 256))
        ntx, nty = -(-W // tw), -(-H // th)
        padded = np.pad(np.asarray(cfa, np.uint16), ((0, nty * th - H), (0, ntx * tw - W)), mode="edge")
        # extra_rows > 0 writes streams taller than their tile (negative tests)
        ex = compress.get("extra_rows", 0)
        padded = np.pad(padded, ((0, ex), (0, 0)), mode="edge")
        chunks = [lj92.encode(padded[ty * th:(ty + 1) * th + ex, tx * tw:(tx + 1) * tw], compress.get("bits", 16),
                              compress.get("predictor", 1), compress.get("components", 2))
                  for ty in range(nty) for tx in range(ntx)]
    pat = cfa_pattern(kind) if pattern is None else np.asarray(pattern, np.uint8)
    n = pat.shape[0]
    cm = xyz_to_cam() if color_matrix is None else np.asarray(color_matrix, np.float64)
    entries = []   # (tag, type, count, payload bytes)

    def add(tag, typ, values):
        fmt = {1: "B", 2: "s", 3: "H", 4: "I", 5: "II", 7: "B", 10: "ii", 12: "d"}[typ]
        if typ == 2:
            data = values.encode() + b"\0"
            cnt = len(data)
        elif typ in (5, 10):
            data = b"".join(struct.pack("<" + fmt, *_rational(v, typ == 10)) for v in values)
            cnt = len(values)
        else:
            data = struct.pack("<%d%s" % (len(values), fmt), *values)
            cnt = len(values)
        entries.append((tag, typ, cnt, data))

    strip_bytes = W * H * 2
    add(254, 4, [0])                      # NewSubFileType: main image
    add(256, 4, [W])                      # ImageWidth
    add(257, 4, [H])                      # ImageLength
    add(258, 3, [compress.get("bits", 16) if chunks else 16])   # BitsPerSample
    add(259, 3, [7 if chunks else 1])     # Compression: lossless JPEG / none
    add(262, 3, [32803])                  # PhotometricInterpretation: CFA
    add(271, 2, make)                     # Make
    add(272, 2, model)                    # Model
    if chunks:
        add(322, 4, [tw])                 # TileWidth
        add(323, 4, [th])                 # TileLength
        add(324, 4, [0] * len(chunks))    # TileOffsets (patched)
        add(325, 4, [len(c) for c in chunks])   # TileByteCounts
    else:
        add(273, 4, [0])                  # StripOffsets (patched)
        add(278, 4, [H])                  # RowsPerStrip
        add(279, 4, [strip_bytes])        # StripByteCounts
    add(277, 3, [1])                      # SamplesPerPixel
    add(284, 3, [1])                      # PlanarConfiguration
    add(33421, 3, [n, n])                 # CFARepeatPatternDim
    add(33422, 1, [int(v) for v in pat.ravel()])  # CFAPattern
    add(50706, 1, [1, 4, 0, 0])           # DNGVersion
    add(50708, 2, f"{make} {model}")      # UniqueCameraModel
    add(50714, 4, [int(black)])           # BlackLevel
    add(50717, 4, [int(white)])           # WhiteLevel
    add(50721, 10, [float(v) for v in cm.ravel()])  # ColorMatrix1
    add(50728, 5, [float(v) for v in as_shot_neutral])  # AsShotNeutral
    add(50778, 3, [int(illuminant)])      # CalibrationIlluminant1 (21 = D65)
    for tag, typ, vals in extra_tags:
        entries[:] = [e for e in entries if e[0] != tag]
        add(tag, typ, vals)
    entries.sort(key=lambda e: e[0])

    ifd_off = 8
    ifd_size = 2 + 12 * len(entries) + 4
    data_off = ifd_off + ifd_size
    blobs = []
    cur = data_off
    fixed = []
    for tag, typ, cnt, data in entries:
        if len(data) > 4:
            pad = cur & 1
            cur += pad
            blobs.append((cur, data))
            fixed.append((tag, typ, cnt, None, cur))
            cur += len(data)
        else:
            fixed.append((tag, typ, cnt, data.ljust(4, b"\0"), None))
    strip_off = (cur + 15) & ~15
    if chunks:   # tile offsets, laid out after the IFD data
        offs, o = [], strip_off
        for c in chunks:
            offs.append(o)
            o += len(c)
        tile_data = struct.pack("<%dI" % len(offs), *offs)
        if len(tile_data) <= 4:
            fixed = [(tg, ty, cn, tile_data.ljust(4, b"\0") if tg == 324 else il, of) for tg, ty, cn, il, of in fixed]
        else:
            at = [of for tg, ty, cn, il, of in fixed if tg == 324][0]
            blobs = [(of, tile_data if of == at else d) for of, d in blobs]
    out = bytearray(b"II*\0" + struct.pack("<I", ifd_off))
    out += struct.pack("<H", len(fixed))
    for tag, typ, cnt, inline, off in fixed:
        if tag == 273:
            inline = struct.pack("<I", strip_off)
        out += struct.pack("<HHI", tag, typ, cnt)
        out += inline if inline is not None else struct.pack("<I", off)
    out += struct.pack("<I", 0)
    for off, data in blobs:
        out += b"\0" * (off - len(out))
        out += data
    out += b"\0" * (strip_off - len(out))
    with open(path, "wb") as f:
        f.write(out)
        if chunks:
            for c in chunks:
                f.write(c)
        else:
            f.write(np.ascontiguousarray(cfa, dtype="<u2").tobytes())
    return path


def gain_map_opcode(gains, top=0, left=0, bottom=None, right=None, spacing=None, origin=(0.0, 0.0),
                    plane=0, planes=1, row_pitch=2, col_pitch=2, flags=0, height=None, width=None):
    """one GainMap opcode (id 9, DNG specification 1.4 chapter 7) as stored in
    an opcode list, big endian: the region, the map grid of ``gains`` (an
    array of map_points_v x map_points_h [x planes] float32), its spacing and
    origin (v, h) in relative image coordinates.  ``spacing`` defaults to a
    grid spanning the unit square."""
    import struct
    g = np.asarray(gains, np.float32)
    if g.ndim == 2:
        g = g[..., None]
    mv, mh, mp = g.shape
    if spacing is None:
        spacing = (1.0 / max(mv - 1, 1), 1.0 / max(mh - 1, 1))
    bottom = height if bottom is None else bottom
    right = width if right is None else right
    params = struct.pack(">8I", top, left, bottom, right, plane, planes, row_pitch, col_pitch)
    params += struct.pack(">2I4d", mv, mh, spacing[0], spacing[1], origin[0], origin[1])
    params += struct.pack(">I", mp) + g.astype(">f4").tobytes()
    return struct.pack(">4I", 9, 0x01030000, flags, len(params)) + params


def opcode_list(ops):
    """an opcode list (OpcodeList1/2/3 tag value): big-endian count, then the
    opcodes' records (from gain_map_opcode, or raw (id, params) pairs)"""
    import struct
    recs = []
    for op in ops:
        if isinstance(op, tuple):
            oid, params = op
            op = struct.pack(">4I", oid, 0x01030000, 0, len(params)) + params
        recs.append(op)
    return struct.pack(">I", len(recs)) + b"".join(recs)


def bayer_gain_maps(W, H, gains4, **kw):
    """the four GainMap opcodes of a Bayer raw, one per site of the 2x2 block
    (top, left) = (0,0) (0,1) (1,0) (1,1), gains4[k] the k-th site's map"""
    return [gain_map_opcode(gains4[k], top=k // 2, left=k % 2, height=H, width=W, **kw) for k in range(4)]


def pack_msb16(samples, bits):
    """samples (uint16) as the bit stream of Magic Lantern raw frames: each
    sample's low `bits` bits most significant first, cut into 16-bit words
    stored little endian"""
    s = np.ascontiguousarray(samples, np.uint16).ravel()
    b = np.unpackbits(s.astype(">u2").view(np.uint8).reshape(-1, 2), axis=1)[:, 16 - bits:].ravel()
    b = np.concatenate([b, np.zeros((-b.size) % 16, np.uint8)])
    words = np.packbits(b.reshape(-1, 16), axis=1)             # [hi, lo] per word
    return words[:, ::-1].tobytes()                           # little endian


def write_mlv(path, frames, bits=14, lj92=None, black=2048, white=15000, fps=(25000, 1000),
              camera="Canon EOS 5D Mark III", iso=800, shutter_us=20000, aperture=280, focal=50,
              order=None, junk_before=(), chunk_split=None, frame_space=0, header_frames=None, audio=None):
    """Write a Magic Lantern raw video clip of uint16 CFA frames (H x W, rggb).

    ``lj92`` None packs ``bits`` per sample; a dict (components, predictor)
    stores each frame as one lossless-JPEG stream (class flag 0x20).
    ``order`` permutes the frames' positions in the file (timestamps keep
    the frame index, so a reader must sort); ``junk_before`` puts a block of
    unknown type before those frame indices; ``chunk_split`` k writes frames
    from index k on into ``<name>.M00``; ``header_frames`` overrides the
    frame count in the file header.  ``audio`` (dict: data bytes, channels,
    rate, bits, blocks: [(timestamp_us, nbytes)] splitting the data in file
    order, space: padding before each block's samples) adds a WAVI block and
    AUDF blocks after the metadata, and sets the header's audio class."""
    import struct
    from vkdt_amd import lj92 as L
    H, W = frames[0].shape
    n = len(frames)
    vclass = 0x01 | (0x20 if lj92 else 0)

    def mlvi():
        cnt = n if header_frames is None else header_frames
        aclass = 1 if audio else 0
        nblk = len(audio["blocks"]) if audio else 0
        return struct.pack("<4sI8sQHHIHHIIII", b"MLVI", 52, b"v2.0\0\0\0\0", 0x1234, 0, 1, 0, vclass, aclass,
                           cnt, nblk, fps[0], fps[1])

    raw_info = struct.pack("<II7i4i4i2i2i18ii", 1, 0, H, W, W * bits // 8, W * H * bits // 8, bits, black, white,
                           0, 0, W, H, 0, 0, H, W, 0, 0, 0x02010100, 0, *([0] * 18), 1100)
    meta = [struct.pack("<4sIQHH", b"RAWI", 180, 0, W, H) + raw_info,
            struct.pack("<4sIQ32sI32s", b"IDNT", 84, 0, camera.encode(), 0x80000285, b"123"),
            struct.pack("<4sIQIIIIQ", b"EXPO", 40, 0, 0, iso, iso, 0, shutter_us),
            struct.pack("<4sIQHHHBBII32s32s", b"LENS", 96, 0, focal, 65535, aperture, 0, 1, 0, 0, b"EF50mm f/1.8", b""),
            struct.pack("<4sIQ10H8s", b"RTCI", 44, 0, 12, 34, 10, 2, 11, 126, 3, 0, 0, 0, b"UTC")]

    def vidf(k):
        f = np.ascontiguousarray(frames[k], np.uint16)
        if lj92:
            data = L.encode(f, 16 if bits > 14 else 14, lj92.get("predictor", 1), lj92.get("components", 1))
        else:
            data = pack_msb16(f, bits)
        pad = b"\0" * frame_space
        body = struct.pack("<IHHHHI", k, 0, 0, 0, 0, frame_space) + pad + data
        return struct.pack("<4sIQ", b"VIDF", 16 + len(body), 1000 + 40000 * k) + body

    if audio:
        ch, rate, ab = audio["channels"], audio["rate"], audio.get("bits", 16)
        meta.append(struct.pack("<4sIQHHIIHH", b"WAVI", 32, 0, 1, ch, rate, rate * ch * ab // 8, ch * ab // 8, ab))
        space, at = audio.get("space", 0), 0
        for i, (ts, nb) in enumerate(audio["blocks"]):
            body = struct.pack("<II", i, space) + b"\0" * space + audio["data"][at:at + nb]
            at += nb
            meta.append(struct.pack("<4sIQ", b"AUDF", 16 + len(body), ts) + body)
    pos = list(order) if order is not None else list(range(n))
    split = chunk_split if chunk_split is not None else n + 1
    chunks = [bytearray(mlvi() + b"".join(meta)), bytearray(mlvi())]
    for i, k in enumerate(pos):
        c = chunks[0] if i < split else chunks[1]
        if k in junk_before:
            c += struct.pack("<4sIQ", b"ZZZZ", 40, 0) + b"\x5a" * 24
        c += vidf(k)
    open(path, "wb").write(bytes(chunks[0]))
    if chunk_split is not None:
        open(path[:-2] + "00", "wb").write(bytes(chunks[1]))
