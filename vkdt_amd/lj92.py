"""Lossless JPEG (ITU-T T.81 process 14, "LJ92") encoding of uint16 tiles, as
compressed DNGs carry their raw data (Compression = 7, one stream per tile).

Used to write synthetic compressed DNGs for tests and benchmarks.  Samples
are predicted per T.81 H.1.2.1 (first sample 2^(P-1), first row from the
left, first column from above, elsewhere the chosen predictor 1..7), the
difference modulo 2^16 is coded as its bit length category (one Huffman
table built from the tile's category histogram) plus that many raw bits
(negative differences as diff - 1), with 0xFF bytes stuffed by 0x00.
A tile of width W is framed as W/C columns of C interleaved components."""
import heapq
import struct

import numpy as np


def _predict(x, pred, bits):
    """prediction of every sample of x (H x W' x C) per T.81 H.1.2.1"""
    x = x.astype(np.int64)
    p = np.zeros_like(x)
    a = np.zeros_like(x)   # left
    b = np.zeros_like(x)   # above
    c = np.zeros_like(x)   # above left
    a[:, 1:] = x[:, :-1]
    b[1:, :] = x[:-1, :]
    c[1:, 1:] = x[:-1, :-1]
    p = {1: a, 2: b, 3: c, 4: a + b - c, 5: a + ((b - c) >> 1), 6: b + ((a - c) >> 1), 7: (a + b) >> 1}[pred]
    p = p.copy()
    p[0, :] = a[0, :]            # first row: from the left
    p[:, 0] = b[:, 0]            # first column: from above
    p[0, 0] = 1 << (bits - 1)    # first sample
    return p


def _code_lengths(freq):
    """Huffman code lengths for the categories with non-zero counts; one
    reserved pseudo-symbol keeps any code from being all ones (T.81 K.2)"""
    heap = [(int(f), i, (i,)) for i, f in enumerate(freq) if f > 0] + [(1, 99, (99,))]
    heapq.heapify(heap)
    depth = {s: 0 for _, _, ss in heap for s in ss}
    k = 100
    while len(heap) > 1:
        f1, _, s1 = heapq.heappop(heap)
        f2, _, s2 = heapq.heappop(heap)
        for s in s1 + s2:
            depth[s] += 1
        heapq.heappush(heap, (f1 + f2, k, s1 + s2))
        k += 1
    depth.pop(99)
    assert max(depth.values(), default=1) <= 16
    return depth


def _canonical(lengths):
    """(BITS[1..16], HUFFVAL, code per symbol) of canonical codes (T.81 C.2)"""
    syms = sorted(lengths, key=lambda s: (lengths[s], s))
    bits = [0] * 17
    for s in syms:
        bits[lengths[s]] += 1
    codes, code, prev = {}, 0, 0
    for s in syms:
        code <<= lengths[s] - prev
        prev = lengths[s]
        codes[s] = code
        code += 1
    return bits[1:], syms, codes


def encode(tile, bits=16, pred=1, components=1):
    """LJ92 stream (bytes) of a uint16 tile (H x W, W a multiple of components)"""
    tile = np.ascontiguousarray(tile, np.uint16)
    H, W = tile.shape
    assert W % components == 0 and 1 <= pred <= 7 and 2 <= bits <= 16
    x = tile.reshape(H, W // components, components)
    d = (x.astype(np.int64) - _predict(x, pred, bits)) & 0xffff
    d = np.where(d >= 0x8000, d - 0x10000, d).ravel()       # differences in [-32768, 32767]
    mag = np.abs(d)
    ssss = np.zeros(d.shape, np.int64)
    nz = mag > 0
    ssss[nz] = np.floor(np.log2(mag[nz])).astype(np.int64) + 1
    ssss[d == -32768] = 16
    lengths = _code_lengths(np.bincount(ssss, minlength=17))
    bits_cnt, huffval, codes = _canonical(lengths)
    code = np.array([codes.get(i, 0) for i in range(17)], np.int64)[ssss]
    clen = np.array([lengths.get(i, 0) for i in range(17)], np.int64)[ssss]
    extra = np.where(d >= 0, d, d - 1) & ((1 << ssss) - 1)
    extra[ssss == 16] = 0
    elen = np.where(ssss == 16, 0, ssss)
    val = (code << elen) | extra
    n = clen + elen
    # the bitstream: every sample's n bits, most significant first
    sh = np.arange(31, -1, -1)
    allbits = ((val[:, None] >> sh[None, :]) & 1).astype(np.uint8)
    keep = sh[None, :] < n[:, None]
    stream = allbits[keep]
    pad = (-stream.size) % 8
    stream = np.concatenate([stream, np.ones(pad, np.uint8)])
    body = np.packbits(stream)
    ff = np.nonzero(body == 0xff)[0]
    body = np.insert(body, ff + 1, 0)                         # byte stuffing
    out = bytearray(b"\xff\xd8")
    out += b"\xff\xc4" + struct.pack(">HB", 2 + 1 + 16 + len(huffval), 0x00)
    out += bytes(bits_cnt) + bytes(huffval)
    out += b"\xff\xc3" + struct.pack(">HBHHB", 8 + 3 * components, bits, H, W // components, components)
    for c in range(components):
        out += bytes([c, 0x11, 0])
    out += b"\xff\xda" + struct.pack(">HB", 6 + 2 * components, components)
    for c in range(components):
        out += bytes([c, 0x00])
    out += bytes([pred, 0, 0])
    out += body.tobytes()
    out += b"\xff\xd9"
    return bytes(out)
