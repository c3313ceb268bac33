"""The product path end to end on the GPU: DNG -> i-raw -> denoise ->
demosaic -> colour -> filmcurv -> llap -> o-pfm through the host runtime
(module DSOs, cfg, roi passes, node creation, arena allocation) and the HIP
C ABI, compared with the oracle's whole-graph restatement on the same raw
under the end-to-end bound of tests/parity.py.
"""
import os

import numpy as np
import pytest

import graphspec as gs
import parity
from vkdt_amd import pfm, pipe, synth

pytestmark = pytest.mark.gpu

# (kind, W, H, seed, denoise strength): strength 0 is the param's default (one noop node)
CASES = [("bayer", 320, 240, 21, 0.5), ("xtrans", 330, 246, 22, 0.5), ("bayer", 320, 240, 24, 0.0)]


@pytest.mark.parametrize("kind,W,H,seed,strength", CASES, ids=[f"{c[0]}-{c[4]}" for c in CASES])
def test_graph_vs_oracle(tmp_path, hip, kind, W, H, seed, strength):
    cfa, _ = synth.mosaic(W, H, seed, kind)
    dng = str(tmp_path / "in.dng")
    synth.write_dng(dng, cfa, kind)
    na, nb = synth.noise_params(2e-4, 1e-6)
    print(f"[gpu_graph] {kind} {W}x{H} seed {seed} strength {strength}", flush=True)
    g = pipe.Graph(lines=pipe.raw_graph_cfg(dng, noise_a=na, noise_b=nb, strength=strength))
    out = str(tmp_path / "out")
    g.export("o-pfm", out)
    g.run()
    img, hdr = pfm.read_pfm(out + ".pfm")
    assert hdr % 16 == 0
    assert img.shape == (H - H % (3 if kind == "xtrans" else 2), W - W % (3 if kind == "xtrans" else 2), 3)
    ip = g.img_param("i-raw", "main")
    p = gs.DevelopParams(kind=kind, noise_a=np.float32(na), noise_b=np.float32(nb),
                         cam=np.array(ip["cam"], np.float32).reshape(3, 3), wb=tuple(ip["wb"][:3]) + (1.0,),
                         strength=strength)
    parity.e2e_check(cfa[:img.shape[0], :img.shape[1]], p, img, log=lambda s: print(s, flush=True), label=f" {kind}")
    g.close()


def test_cli_writes_pfm(tmp_path, hip):
    """vkdt-cli -g graph.cfg --format o-pfm --filename out"""
    import subprocess
    from vkdt_amd import PKG_DIR
    W, H = 256, 160
    cfa, _ = synth.mosaic(W, H, 5, "bayer")
    dng = str(tmp_path / "c.dng")
    synth.write_dng(dng, cfa, "bayer")
    cfg = tmp_path / "g.cfg"
    cfg.write_text("\n".join(pipe.raw_graph_cfg(dng, noise_a=1e-7, noise_b=1e-5)) + "\n")
    out = str(tmp_path / "cli")
    r = subprocess.run([os.path.join(PKG_DIR, "bin", "vkdt-cli"), "-g", str(cfg), "--format", "o-pfm",
                        "--filename", out, "--colour-prim", "bt2020", "--colour-trc", "linear", "-d", "perf"],
                       capture_output=True, text=True, timeout=120)
    print(r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    img, _ = pfm.read_pfm(out + ".pfm")
    assert img.shape == (H, W, 3) and np.isfinite(img).all()


def test_replay_equals_eager_and_follows_params(tmp_path, hip):
    """dt_graph_set_replay: repeated runs replay the recorded launches as one
    hipGraph; their results equal the eager run's, and a changed parameter
    records again (the replayed result follows it)"""
    W, H = 256, 192
    cfa, _ = synth.mosaic(W, H, 9, "bayer")
    dng = str(tmp_path / "r.dng")
    synth.write_dng(dng, cfa, "bayer")
    na, nb = synth.noise_params(2e-4, 1e-6)

    def develop(clarity, replay, runs):
        g = pipe.Graph(lines=pipe.raw_graph_cfg(dng, noise_a=na, noise_b=nb, clarity=clarity))
        out = str(tmp_path / f"o{clarity}{replay}")
        g.export("o-pfm", out)
        g.set_replay(replay)
        imgs = []
        for _ in range(runs):
            g.run()
            imgs.append(pfm.read_pfm(out + ".pfm")[0].copy())
        return g, out, imgs

    _, _, eager = develop(0.2, False, 1)
    g, out, rep = develop(0.2, True, 3)
    for k, img in enumerate(rep):
        assert np.array_equal(img, eager[0]), k
    g.set_param("llap", "01", "clarity", 0.5)
    g.run()
    changed = pfm.read_pfm(out + ".pfm")[0]
    g.close()
    _, _, eager5 = develop(0.5, False, 1)
    assert np.array_equal(changed, eager5[0])
    assert not np.array_equal(changed, eager[0])
