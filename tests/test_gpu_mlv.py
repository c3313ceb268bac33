"""Magic Lantern raw video on the GPU: i-mlv's read_source_hip uploads each
frame's stored bytes and unpacks (bit-packed frames, vkdt_hip_unpack_msb16)
or decodes (lossless JPEG, vkdt_hip_lj92_decode) them on the device into
the source buffer.  Every frame of every clip kind equals the frame written,
bit for bit -- the same frames the host path (read_source) reads back in the
CPU tests -- and the unpack entry point alone matches the reference's unpack
formula for 10..16 bits and ragged counts.  A whole clip develops through
the graph frame by frame."""
import numpy as np
import pytest

from vkdt_amd import hip as H, pfm, pipe, synth
from test_mlv import CASES, clip_frames, ref_unpack

pytestmark = pytest.mark.gpu


def source_buffer(g):
    for n in g.nodes():
        if n["module"] == "i-mlv" and not n["culled"]:
            return n["conns"][0]["buf"]
    raise AssertionError("no i-mlv node")


@pytest.mark.parametrize("case", list(CASES))
def test_frames_on_the_gpu(tmp_path, hip, case):
    kw = CASES[case]
    W, Hh = 96, 64
    frames = clip_frames(W=W, H=Hh, bits=kw["bits"])
    path = str(tmp_path / "clip.MLV")
    synth.write_mlv(path, frames, **kw)
    g = pipe.Graph(lines=pipe.mlv_graph_cfg(path))
    g.export("o-pfm", str(tmp_path / "o"))
    print(f"[gpu_mlv] {case}", flush=True)
    for f in range(len(frames)):
        g.set_frame(f)
        g.run()
        b = source_buffer(g)
        got = np.empty((Hh, W), np.uint16)
        H.check(H.lib().vkdt_hip_download(got.ctypes.data, b.ptr, got.nbytes, g.stream()), "download")
        H.sync(g.stream())
        assert np.array_equal(got, frames[f]), (case, f)
        out = pfm.read_pfm(str(tmp_path / "o") + ".pfm")[0]
        assert np.isfinite(out).all()
    g.close()


@pytest.mark.parametrize("bits,n", [(10, 4096), (12, 1001), (14, 24 * 1024 + 3), (16, 777), (11, 333)])
def test_unpack_entry_point(hip, bits, n):
    rng = np.random.default_rng(bits + n)
    x = rng.integers(0, 1 << bits, n).astype(np.uint16)
    packed = synth.pack_msb16(x, bits)
    out = hip.DeviceBuffer(n, 1, 1, H.FMT_UI16)
    out.upload(np.full(n, 0xdead, np.uint16))
    H.unpack_msb16(packed, out.ptr, n, bits)
    got = out.download().ravel()
    assert np.array_equal(got, x)
    assert np.array_equal(got[:200], ref_unpack(packed, 200, bits))
    out.free()

