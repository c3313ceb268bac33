"""Magic Lantern raw video (i-mlv) on the CPU: the clip reader, the metadata
it publishes and the host unpack/decode of read_source (reference
i-mlv/main.c, video_mlv.c).

  * the synthetic writer's bit packing is checked against the reference's
    unpack formula restated (a 32-bit window of two little-endian words
    rotated right by 16 + (32 - bits) - shift, the count taken mod 32 as
    x86 does, video_mlv.c:262-276) for 10..16 bits -- this pins the format
    both the host and the GPU unpack read;
  * every frame of clips with packed (10, 12, 14 bit) and lossless-JPEG
    frames reads back bit for bit, in timestamp order when the file holds
    them out of order, across an unknown block (resynchronised on the next
    known block type), across a second chunk file (.M00) and past frame
    padding (frameSpace);
  * the frame index is held at the header's last frame;
  * extent, levels, filters, crop, exposure, aperture, ISO, focal length,
    maker/model, frame count and rate are the clip's;
  * the audio callback hands back the clip's AUDF bytes synced to the first
    frame (cut before it, silence after it, in whole samples), limited to
    the frames' duration in 1024-sample blocks, with the reference's
    sample/byte count (main.c:214-230, video_mlv.c:18-87).
"""
import numpy as np
import pytest

from vkdt_amd import pipe, synth


def ref_unpack(packed, n, bits):
    w = np.frombuffer(packed + b"\0\0\0\0", "<u2")
    out = np.zeros(n, np.uint32)
    for i in range(n):
        off = i * bits
        a, s = off // 16, off % 16
        v = int(w[a]) | (int(w[a + 1]) << 16)
        r = (16 + ((32 - bits) - s)) % 32
        v = ((v >> r) | (v << (32 - r))) & 0xffffffff
        out[i] = v & ((1 << bits) - 1)
    return out


@pytest.mark.parametrize("bits", [10, 11, 12, 14, 16])
def test_packing_matches_the_reference_unpack(bits):
    rng = np.random.default_rng(bits)
    x = rng.integers(0, 1 << bits, 777).astype(np.uint16)
    assert np.array_equal(ref_unpack(synth.pack_msb16(x, bits), x.size, bits), x)


def clip_frames(n=4, W=96, H=64, bits=14, seed=3):
    rng = np.random.default_rng(seed)
    return [rng.integers(0, 1 << bits, size=(H, W)).astype(np.uint16) for _ in range(n)]


def read_frames(path, n, W, H):
    g = pipe.Graph(lines=pipe.mlv_graph_cfg(path))
    g.export("o-pfm", path + ".out")
    out = []
    for f in range(n):
        g.set_frame(f)
        g.plan()
        out.append(g.read_source("i-mlv", "main", W * H * 2).view(np.uint16).reshape(H, W).copy())
    return g, out


CASES = {
    "packed14": dict(bits=14),
    "packed12": dict(bits=12),
    "packed10": dict(bits=10),
    "lj92-1": dict(bits=14, lj92=dict(components=1, predictor=1)),
    "lj92-2": dict(bits=14, lj92=dict(components=2, predictor=1)),
    "lj92-pred6": dict(bits=14, lj92=dict(components=1, predictor=6)),
    "shuffled": dict(bits=14, order=[2, 0, 3, 1]),
    "junk": dict(bits=12, junk_before=(1, 3)),
    "chunks": dict(bits=14, chunk_split=2),
    "padding": dict(bits=14, frame_space=20),
}


@pytest.mark.parametrize("case", list(CASES))
def test_frames_read_back(tmp_path, case):
    kw = CASES[case]
    frames = clip_frames(bits=kw["bits"])
    path = str(tmp_path / "clip.MLV")
    synth.write_mlv(path, frames, **kw)
    g, got = read_frames(path, len(frames), 96, 64)
    for f, (a, b) in enumerate(zip(got, frames)):
        assert np.array_equal(a, b), f
    assert g.frame_cnt() == len(frames)
    g.close()


def test_frame_index_held_at_the_last(tmp_path):
    frames = clip_frames(n=3)
    path = str(tmp_path / "c.mlv")
    synth.write_mlv(path, frames, header_frames=2)
    g = pipe.Graph(lines=pipe.mlv_graph_cfg(path))
    g.export("o-pfm", path + ".out")
    g.set_frame(7)
    g.plan()
    got = g.read_source("i-mlv", "main", 96 * 64 * 2).view(np.uint16).reshape(64, 96)
    assert np.array_equal(got, frames[1])       # min(frame, header count - 1)
    assert g.frame_cnt() == 2
    g.close()


def test_metadata(tmp_path):
    frames = clip_frames(n=2)
    path = str(tmp_path / "m.mlv")
    synth.write_mlv(path, frames, black=2047, white=15200, fps=(24000, 1001), camera="Canon EOS 5D Mark III",
                    iso=1600, shutter_us=10000, aperture=560, focal=35)
    g = pipe.Graph(lines=pipe.mlv_graph_cfg(path))
    g.export("o-pfm", path + ".out")
    g.plan()
    ip = g.img_param("i-mlv", "main")
    assert ip["full"] == (96, 64)
    assert ip["black"] == [2047.0] * 4 and ip["white"] == [15200.0] * 4
    assert ip["filters"] == 0x5d5d5d5d
    assert ip["crop"] == [0, 0, 96, 64]
    assert ip["iso"] == 1600.0
    assert abs(ip["exposure"] - 0.01) < 1e-9 and abs(ip["aperture"] - 5.6) < 1e-6 and ip["focal_length"] == 35.0
    assert ip["maker"] == "Canon" and ip["model"] == "Canon EOS 5D Mark III"
    assert ip["noise"] == [1.0, 1.0]
    assert g.frame_cnt() == 2 and abs(g.frame_rate() - 24000 / 1001) < 1e-4
    # the body's D65 XYZ -> camera matrix (Adobe DNG Converter ColorMatrix2 of
    # the EOS 5D Mark III, x 1e4): cam -> rec2020 = xyz -> rec2020 . inverse,
    # the white balance is the row sums of the inverse, green-normalised
    # (i-mlv/main.c:172-201)
    M = np.array([[6722, -635, -963], [-4287, 12460, 2028], [-908, 2162, 5668]]) / 10000.0
    cam_to_xyz = np.linalg.inv(M)
    xyz_to_2020 = np.array([[1.7166511880, -0.3556707838, -0.2533662814], [-0.6666843518, 1.6164812366, 0.0157685458],
                            [0.0176398574, -0.0427706133, 0.9421031212]])
    assert np.allclose(np.array(ip["cam"]).reshape(3, 3), xyz_to_2020 @ cam_to_xyz, rtol=1e-5, atol=1e-6)
    wb = cam_to_xyz.sum(1)
    assert np.allclose(ip["wb"][:3], [wb[0] / wb[1], 1.0, wb[2] / wb[1]], rtol=1e-5)
    g.close()


def test_unknown_camera_takes_the_identity(tmp_path):
    """a body missing from the coefficient table: camera rgb is taken as xyz,
    with unit white balance (the reference's fallback)"""
    path = str(tmp_path / "u.mlv")
    synth.write_mlv(path, clip_frames(n=1), camera="Canon EOS 9000X")
    g = pipe.Graph(lines=pipe.mlv_graph_cfg(path))
    g.export("o-pfm", path + ".out")
    g.plan()
    ip = g.img_param("i-mlv", "main")
    xyz_to_2020 = [1.7166511880, -0.3556707838, -0.2533662814, -0.6666843518, 1.6164812366, 0.0157685458,
                   0.0176398574, -0.0427706133, 0.9421031212]
    assert np.allclose(ip["cam"], xyz_to_2020, rtol=1e-6)
    assert ip["wb"][:3] == [1.0, 1.0, 1.0]
    g.close()


def test_rejects_bad_clips(tmp_path):
    bad = str(tmp_path / "x.mlv")
    open(bad, "wb").write(b"RIFF" + bytes(60))
    g = pipe.Graph(lines=pipe.mlv_graph_cfg(bad))
    g.export("o-pfm", bad + ".out")
    with pytest.raises(pipe.PipeError):
        g.plan()
    g.close()


# ---- audio (i-mlv/main.c:214-230, video_mlv.c:18-87) -------------------------
# Video frames are stamped at 1000 + 40000 k us (25 fps).  The audio blocks
# of these clips are 16-bit stereo at 48 kHz (4 bytes a sample of both
# channels, 192 bytes a millisecond) with distinct byte values, so every
# check below compares bytes the module hands back with bytes of the file.

def audio_clip(tmp_path, start_us, nframes=4, nbytes=96000, blocks=3, space=0, fps=(25000, 1000)):
    rng = np.random.default_rng(nbytes + start_us)
    data = rng.integers(0, 256, nbytes, dtype=np.uint8).tobytes()
    cuts = np.linspace(0, nbytes, blocks + 1).astype(int) // 4 * 4
    bl = [(start_us + int(cuts[i]) // 192 * 1000, int(cuts[i + 1] - cuts[i])) for i in range(blocks)]
    path = str(tmp_path / f"a{start_us}.mlv")
    synth.write_mlv(path, clip_frames(n=nframes), fps=fps,
                    audio=dict(data=data, channels=2, rate=48000, bits=16, blocks=bl, space=space))
    g = pipe.Graph(lines=pipe.mlv_graph_cfg(path))
    g.export("o-pfm", path + ".out")
    g.plan()
    return g, data


def all_audio(g):
    out, beg = b"", 0
    while True:
        k, b = g.audio("i-mlv", "main", beg, 4096, 4)
        if k <= 0:
            return out
        out += b
        beg += k


def test_audio_in_step_with_the_first_frame(tmp_path):
    """audio stamped with the first frame (1000 us) comes back as written,
    past per-block padding, described as S16_LE stereo at 48 kHz"""
    g, data = audio_clip(tmp_path, 1000, space=12)
    ip = g.img_param("i-mlv", "main")
    assert (ip["snd_format"], ip["snd_channels"], ip["snd_samplerate"]) == (2, 2, 48000)
    got = all_audio(g)
    assert len(got) == 32768 and got == data[:32768]
    g.close()


def test_audio_before_the_first_frame_is_cut(tmp_path):
    """audio stamped from 0 us, 1 ms (48 samples, 192 bytes) before the first
    frame, loses those samples"""
    g, data = audio_clip(tmp_path, 0, nbytes=64000)
    got = all_audio(g)
    assert got == data[192:192 + len(got)] and len(got) == 32768
    g.close()


def test_audio_after_the_first_frame_gets_silence(tmp_path):
    """audio starting 5 ms after the first frame is preceded by 240 silent
    samples (960 bytes)"""
    g, data = audio_clip(tmp_path, 6000, nbytes=64000)
    got = all_audio(g)
    assert len(got) == 32768 and got[:960] == b"\0" * 960 and got[960:] == data[:32768 - 960]
    g.close()


def test_audio_length_is_the_frames_duration(tmp_path):
    """4 frames at 25 fps are 160 ms = 30720 bytes, rounded up to whole
    1024-sample blocks (4096 bytes): 32768 of the 96000 bytes play; a
    clip with less audio than that plays all of it (rounded up the same way,
    the rest silent)"""
    g, data = audio_clip(tmp_path, 1000, nframes=4, nbytes=96000)
    got = all_audio(g)
    assert len(got) == 32768 and got == data[:32768]
    assert g.audio("i-mlv", "main", 32768 // 4, 16, 4)[0] == 0
    g.close()
    g, data = audio_clip(tmp_path, 1000, nframes=40, nbytes=10000, blocks=2)
    got = all_audio(g)
    assert len(got) == 12288 and got[:10000] == data and got[10000:] == b"\0" * 2288
    g.close()


def test_audio_count_mixes_samples_and_bytes(tmp_path):
    """the callback returns min(cnt, bytes left) / bytes a sample, as the
    reference computes it"""
    g, data = audio_clip(tmp_path, 1000, nframes=4, nbytes=96000)
    k, b = g.audio("i-mlv", "main", 0, 1000, 4)
    assert k == 250 and b == data[:1000]
    k, b = g.audio("i-mlv", "main", 8000, 4096, 4)       # 768 bytes left of 32768
    assert k == 192 and b == data[32000:32768]
    g.close()


def test_no_audio_without_wavi_or_audf(tmp_path):
    path = str(tmp_path / "n.mlv")
    synth.write_mlv(path, clip_frames(n=2))
    g = pipe.Graph(lines=pipe.mlv_graph_cfg(path))
    g.export("o-pfm", path + ".out")
    g.plan()
    assert g.audio("i-mlv", "main", 0, 100, 4) == (0, b"")
    assert g.img_param("i-mlv", "main")["snd_channels"] == 0
    g.close()
