"""i-mcraw's frame codec on the GPU through the C ABI: the prefix node's
offsets and the decode node's ui16 frame equal the oracle's bit for bit
(two transcriptions of prefix.comp and decode.comp), and the frame equals
the one the test-side encoder coded (through the R16 UNORM store of v/65536).

Frames: 4096x2160 (the serving workload's video frames) with a BGGR offset
(1, 1) and without one (the decode's 8-byte row stores), blocks of every
count including the aliases 7, 9 and 11..15, small frames whose blocks all
use one count, and counts that start off the allocation's 16-byte alignment
(the scan's byte loads)."""
import numpy as np
import pytest

import mcraw_codec as mc
from vkdt_amd import hip as H

pytestmark = pytest.mark.gpu


def run_gpu(hip, bits, refs, data, ox, oy, owd, oht, wd, ht, shift=0):
    """prefix and decode through the C ABI; `shift` bytes move the counts
    off the allocation's alignment (the scan's byte-load path)"""
    nb = owd * oht // 64
    rblock = -(-nb // 64) * 64                          # i-mcraw/main.cc:371-377
    nbits = (rblock + 1) // 2
    bb = np.zeros(nbits + shift, np.uint8)
    bb[shift:shift + bits.size] = bits
    rr = np.zeros(rblock, np.uint16)
    rr[:refs.size] = refs
    dd = np.zeros(max(2 * owd * oht, data.size), np.uint8)
    dd[:data.size] = data
    b_bits = hip.DeviceBuffer(nbits + shift, 1, 1, H.FMT_UI8)
    b_ref = hip.DeviceBuffer(rblock, 1, 1, H.FMT_UI16)
    b_dat = hip.DeviceBuffer(dd.size, 1, 1, H.FMT_UI8)
    b_off = hip.DeviceBuffer(rblock, 1, 1, H.FMT_UI32)
    b_scr = hip.DeviceBuffer((rblock + 511) // 512 + 1, 4, 1, H.FMT_UI32)
    b_out = hip.DeviceBuffer(wd, ht, 1, H.FMT_UI16)
    try:
        b_bits.upload(bb)
        b_ref.upload(rr)
        b_dat.upload(dd)
        b_scr.clear()
        cnt = H.Buf(b_bits.ptr + shift, nbits, 1, nbits, 1, H.FMT_UI8, 1, nbits)
        hip.launch("i-mcraw", "prefix", None, None, [cnt, b_off, b_scr], (nb + 511) // 512 * 8, 1)
        pc = np.array([ox, oy, owd, oht], np.int32)
        hip.launch("i-mcraw", "decode", pc, None, [b_out, cnt, b_off, b_ref, b_dat], nb * 8, 1)
        hip.sync()
        return b_off.download()[0], b_out.download()
    finally:
        for b in (b_bits, b_ref, b_dat, b_off, b_scr, b_out):
            b.free()


@pytest.mark.parametrize("owd,oht,ox,oy,single,shift",
                         [(4096, 2160, 1, 1, None, 0), (4096, 2160, 0, 0, None, 0), (256, 32, 0, 1, 3, 0),
                          (192, 16, 1, 0, 10, 0), (1024, 72, 0, 0, None, 4)],
                         ids=["4096x2160-bggr", "4096x2160-rggb", "256x32-count3", "192x16-count10",
                              "1024x72-unaligned-counts"])
def test_decode_equals_oracle_and_the_coded_frame(hip, owd, oht, ox, oy, single, shift):
    rng = np.random.default_rng(owd + oht)
    nb = owd * oht // 64
    widths = np.array([0, 1, 2, 3, 4, 5, 6, 8, 10, 16])
    w = widths[rng.integers(0, widths.size, size=nb)] if single is None else np.full(nb, mc.WIDTH[single])
    base = rng.integers(0, np.maximum(1, 65536 - (1 << w)))
    xs, ys = mc.block_sites(np.arange(nb), owd // 64)
    coded = np.zeros((oht, owd), np.uint16)
    coded[ys, xs] = (base[:, None] + rng.integers(0, 1 << w[:, None], size=(nb, 64))).astype(np.uint16)
    alias = np.arange(16)
    alias[[8, 10, 11]] = [7, 9, 14]
    bits, refs, data = mc.encode(coded, counts=lambda i, c: np.where(i % 4 == 1, alias[c], c))
    wd, ht = ((owd - ox) // 2) * 2, ((oht - oy) // 2) * 2
    print(f"[gpu_mcraw] {owd}x{oht} offset ({ox},{oy}) -> {wd}x{ht}: {nb} blocks, {data.nbytes} bytes", flush=True)
    off_gpu, out_gpu = run_gpu(hip, bits, refs, data, ox, oy, owd, oht, wd, ht, shift)
    off = mc.oracle_prefix(bits, nb)
    assert np.array_equal(off_gpu[:nb], off)
    ref = mc.oracle_decode(bits, off, refs, data, ox, oy, owd, oht, wd, ht)
    assert np.array_equal(out_gpu, ref)
    assert np.array_equal(out_gpu, mc.unorm_store(coded[oy:oy + ht, ox:ox + wd]))
