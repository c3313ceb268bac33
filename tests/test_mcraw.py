"""i-mcraw's frame codec on the CPU: the oracle's restatement of
i-mcraw/decode.comp and prefix.comp, pinned by properties the format must
have, and a frame round trip through the test-side encoder.

The reference tree holds the decoder only (the files come from
motioncam-decoder, hanatos fork d4f1f7d, absent), so there are no golden
frames.  What pins the restatement:

* for every count the decoder's map from input bits to output bits is a
  bijection between the block's 8*width bytes and the low `width` bits of its
  64 values (the stored widths are dense, so a mistyped register, offset or
  shift would map two input bits to one output bit or leave one unreached);
* counts 7, 9 and 11..15 decode as 8, 10 and 16 bits (decode.comp:306-313;
  the counts are 4 bits, so its `case 16` is the default's)
  and take 64, 80 and 128 bytes (prefix.comp:40-55);
* a frame encoded with that map and decoded comes back, through the R16
  UNORM store of v/65536 (samples of 32769 and more one less).

Parity unpinned: which lane of a block a group of bits feeds (a permutation
of the values within one count) keeps the bijection and the round trip; it
is held only by the HIP kernel and the oracle, two transcriptions of
decode.comp, agreeing (tests/test_gpu_mcraw.py).
"""
import numpy as np
import pytest

import mcraw_codec as mc


def check_bit_layouts_are_bijections():
    """every count's bit map, probed afresh (the planted-mistake test swaps
    the oracle build under it)"""
    for cnt in range(16):
        check_bijection(mc.bit_map.__wrapped__(cnt), cnt)


def check_bijection(m, cnt):
    assert len(m) == mc.BYTES[cnt] * 8
    outs = list(m.values())
    assert all(isinstance(o, tuple) and len(o) == 2 and isinstance(o[0], int) for o in outs), \
        f"count {cnt}: an input bit reaches no value bit or more than one"
    assert sorted(outs) == sorted((v, b) for v in range(64) for b in range(mc.WIDTH[cnt])), cnt


# the codec checks the planted-mistake test runs beside the stencil pins
CHECKS = [check_bit_layouts_are_bijections]


@pytest.mark.parametrize("cnt", range(16))
def test_bit_layout_is_a_bijection(cnt):
    m = mc.bit_map(cnt)
    assert len(m) == mc.BYTES[cnt] * 8
    outs = list(m.values())
    assert all(isinstance(o, tuple) and len(o) == 2 and isinstance(o[0], int) for o in outs), \
        "an input bit reaches no value bit or more than one"
    assert sorted(outs) == sorted((v, b) for v in range(64) for b in range(mc.WIDTH[cnt]))


def test_counts_alias_to_stored_widths():
    rng = np.random.default_rng(4)
    for alias, base in ((7, 8), (9, 10), (12, 11), (13, 11), (14, 11), (15, 11)):   # 11..15: 16-bit values
        d = rng.integers(0, 256, size=mc.BYTES[base], dtype=np.uint8)
        assert np.array_equal(mc.decode_block(d, alias), mc.decode_block(d, base))
    bits = np.array([0x10, 0x32, 0x54, 0x76, 0x98, 0xba, 0xdc, 0xfe], np.uint8)   # counts 0..15
    off = mc.oracle_prefix(bits, 16)
    assert list(np.diff(np.concatenate([[0], off]))) == mc.BYTES


def test_unorm_store_of_v_over_65536():
    v = np.array([0, 1, 1000, 32767, 32768, 32769, 40000, 65535, 70000])
    assert list(mc.unorm_store(v)) == [0, 1, 1000, 32767, 32768, 32768, 39999, 65534, 65535]


@pytest.mark.parametrize("wd,ht,ox,oy", [(128, 16, 0, 0), (190, 19, 1, 1), (250, 8, 1, 0)])
def test_frame_round_trip(wd, ht, ox, oy):
    """a frame with a CFA offset (the coded frame is the mosaic shifted by
    (ox, oy), width rounded up to 64) in blocks of every count"""
    rng = np.random.default_rng(wd + ht)
    owd, oht = -(-(wd + ox) // 64) * 64, ht + oy
    bpr = owd // 64
    coded = np.zeros((oht, owd), np.uint16)
    for idx in range(owd * oht // 64):            # per block a base and a delta width
        xs, ys = mc.block_sites(idx, bpr)
        xs, ys = xs.ravel(), ys.ravel()
        w = [0, 1, 2, 3, 4, 5, 6, 8, 10, 16][idx % 10]
        base = int(rng.integers(0, max(1, 65536 - (1 << w))))
        coded[ys, xs] = base + rng.integers(0, 1 << w, size=64)
    alias = np.arange(16)
    alias[[8, 10, 11]] = [7, 9, 13]
    bits, refs, data = mc.encode(coded, counts=lambda i, c: np.where(i % 3 == 0, alias[c], c))
    nb = owd * oht // 64
    off = mc.oracle_prefix(bits, nb)
    assert off[-1] == data.nbytes
    out = mc.oracle_decode(bits, off, refs, data, ox, oy, owd, oht, wd, ht)
    assert np.array_equal(out, mc.unorm_store(coded[oy:oy + ht, ox:ox + wd]))
