"""Test-side encoder of the i-mcraw block format (i-mcraw/decode.comp).

The decoder is the format's only definition here (motioncam-decoder, which
writes the files, is not in the reference tree).  Its bit layout is read off
the oracle's restatement of decode.comp by probing: each input bit of a block
is set alone and the one output bit it lands on is recorded.  The encoder
inverts that map; tests/test_mcraw.py checks that the map is a bijection for
every count, which a mistyped layout would break.
"""
import ctypes as C
from functools import lru_cache

import numpy as np

import oracle_py as op

WIDTH = [0, 1, 2, 3, 4, 5, 6, 8, 8, 10, 10, 16, 16, 16, 16, 16]      # bits stored per value, by count
BYTES = [w * 8 for w in WIDTH]                                        # data bytes per block


def decode_block(data, cnt, off=0):
    """64 offsets of one block (lane order) on the oracle"""
    data = np.ascontiguousarray(data, np.uint8)
    out = np.zeros(64, np.uint32)
    op.lib(64).or_mcraw_block(data.ctypes.data, data.nbytes, off, cnt, out.ctypes.data)
    return out


@lru_cache(maxsize=None)
def bit_map(cnt):
    """{(byte, bit): (value, value bit)} of the decoder for count cnt: one
    probe block per input bit, that bit alone set, decoded in one call"""
    nb = BYTES[cnt]
    n = nb * 8
    dat = np.zeros((max(n, 1), max(nb, 1)), np.uint8)
    for k in range(nb):
        for j in range(8):
            dat[8 * k + j, k] = 1 << j
    off = (np.arange(max(n, 1), dtype=np.uint32) * max(nb, 1)).astype(np.uint32)
    cnts = np.full(max(n, 1), cnt, np.uint32)
    out = np.zeros((max(n, 1), 64), np.uint32)
    op.lib(64).or_mcraw_blocks(dat.ctypes.data, dat.nbytes, off.ctypes.data, cnts.ctypes.data, n, out.ctypes.data)
    m = {}
    for k in range(nb):
        for j in range(8):
            o = out[8 * k + j]
            hits = [(v, b) for v in np.nonzero(o)[0].tolist() for b in range(32) if (int(o[v]) >> b) & 1]
            m[(k, j)] = hits[0] if len(hits) == 1 else tuple(hits)
    return m


def pack(deltas, cnt):
    """the data bytes of blocks (n x BYTES[cnt]) whose offsets (n x 64, lane
    order) are `deltas`, all with count cnt"""
    deltas = np.asarray(deltas, np.int64).reshape(-1, 64)
    d = np.zeros((deltas.shape[0], BYTES[cnt]), np.uint8)
    for (k, j), (v, b) in bit_map(cnt).items():
        d[:, k] |= (((deltas[:, v] >> b) & 1) << j).astype(np.uint8)
    return d


def block_sites(idx, bpr):
    """coded-frame (x, y) of the 64 values of block(s) idx, lane order
    (decode.comp:314-337)"""
    idx = np.asarray(idx)[..., None]
    k, s = idx // 4, np.arange(32)
    x = 64 * (k % bpr) - np.where(s >= 16, 64, 0) + (idx & 1) + 4 * s
    y = 4 * (k // bpr) + np.where(s >= 16, 2, 0) + ((idx >> 1) & 1)
    xs = np.stack([x, x + 2], -1).reshape(x.shape[:-1] + (64,))
    ys = np.stack([y, y], -1).reshape(y.shape[:-1] + (64,))
    return xs, ys


def encode(coded, counts=None):
    """bit counts (packed nibbles), reference values and data of an owd x oht
    coded frame (uint16, owd a multiple of 64, oht a multiple of 4).
    counts(idx, need) (arrays) may pick the stored count among the aliases
    of the width each block needs."""
    oht, owd = coded.shape
    bpr, nb = owd // 64, owd * oht // 64
    xs, ys = block_sites(np.arange(nb), bpr)
    v = coded[ys, xs].astype(np.int64)
    refs = v.min(1)
    dl = v - refs[:, None]
    need = np.ceil(np.log2(dl.max(1) + 1)).astype(np.int64)
    least = np.array([next(c for c in range(16) if WIDTH[c] >= w) for w in range(17)])
    cnts = least[need]
    if counts is not None:
        cnts = np.asarray(counts(np.arange(nb), cnts))
    sizes = np.array(BYTES)[cnts]
    ends = np.cumsum(sizes)
    data = np.zeros(int(ends[-1]) if nb else 0, np.uint8)
    for c in np.unique(cnts):
        sel = np.nonzero(cnts == c)[0]
        if BYTES[c] == 0:
            continue
        pk = pack(dl[sel], int(c))
        pos = (ends[sel] - BYTES[c])[:, None] + np.arange(BYTES[c])
        data[pos] = pk
    bits = np.zeros((nb + 1) // 2, np.uint8)
    bits[:] = cnts[0::2]
    bits[:len(cnts[1::2])] |= (cnts[1::2] << 4).astype(np.uint8)
    return bits, refs.astype(np.uint16), data


def unorm_store(v):
    """the R16 UNORM store of v/65536: v * 65535/65536 rounded to nearest,
    ties to even, in exact (integer) arithmetic, saturating at 65535"""
    v = np.asarray(v, np.int64)
    num = v * 65535                          # / 65536, rounded half to even
    q, r = num >> 16, num & 0xffff
    up = (r > 0x8000) | ((r == 0x8000) & (q & 1 == 1))
    return np.minimum(q + up, 65535).astype(np.uint16)


def oracle_prefix(bits, n):
    off = np.zeros(n, np.uint32)
    op.lib(64).or_mcraw_prefix(np.ascontiguousarray(bits).ctypes.data, n, off.ctypes.data)
    return off


def oracle_decode(bits, off, refs, data, ox, oy, owd, oht, wd, ht):
    out = np.zeros((ht, wd), np.uint16)
    data = np.ascontiguousarray(data, np.uint8)
    op.lib(64).or_mcraw_decode(np.ascontiguousarray(bits).ctypes.data, off.ctypes.data,
                               np.ascontiguousarray(refs).ctypes.data, data.ctypes.data, data.nbytes,
                               ox, oy, owd, oht, out.ctypes.data, wd, ht)
    return out
