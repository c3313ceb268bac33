"""Host logic of the serving loop (no device): the first-fit pool the
server admits requests against."""
import random

from vkdt_amd.serve import FirstFit


def test_first_fit_alignment_reuse_and_coalescing():
    ff = FirstFit(1 << 20)
    a = ff.alloc(1000)
    b = ff.alloc(300)
    c = ff.alloc(256)
    assert (a, b, c) == (0, 1024, 1536)          # 256-byte aligned, first fit
    ff.free(b)
    assert ff.alloc(200) == 1024                 # the hole is reused
    ff.free(a)
    ff.free(c)
    ff.free(1024)
    assert ff.free_blocks == [(0, 1 << 20)]      # everything merged back
    assert ff.in_use() == 0 and ff.peak == 1792


def test_first_fit_never_overlaps_and_refuses_when_full():
    rng = random.Random(5)
    ff = FirstFit(64 << 10)
    live = {}
    for _ in range(2000):
        if live and rng.random() < 0.45:
            o = rng.choice(list(live))
            ff.free(o)
            del live[o]
        else:
            n = rng.randint(1, 6000)
            o = ff.alloc(n)
            if o is None:
                assert ff.largest_free() < (n + 255) // 256 * 256
                continue
            live[o] = (n + 255) // 256 * 256
        spans = sorted(live.items())
        for (o1, n1), (o2, _) in zip(spans, spans[1:]):
            assert o1 + n1 <= o2
        assert all(o % 256 == 0 and o + n <= ff.size for o, n in spans)
        assert ff.in_use() == sum(live.values()) <= ff.size
