"""The product path end to end on the GPU: DNG -> i-raw -> denoise ->
demosaic -> colour -> filmcurv -> llap -> o-pfm through the host runtime
(module DSOs, cfg, roi passes, node creation, arena allocation) and the HIP
C ABI, compared with the oracle's whole-graph restatement on the same raw.

Tolerance vs the FP64 oracle (f16 storage between nodes, as the reference
stores): PSNR >= 60 dB on linear rec2020 and at most 0.1% of values off by
more than 1e-3 (branch flips between float and double arithmetic).
"""
import os

import numpy as np
import pytest

import graphspec as gs
import oracle_py as op
from vkdt_amd import pfm, pipe, synth

pytestmark = pytest.mark.gpu

CASES = [("bayer", 320, 240, 21), ("xtrans", 330, 246, 22)]


@pytest.mark.parametrize("kind,W,H,seed", CASES, ids=[c[0] for c in CASES])
def test_graph_vs_oracle(tmp_path, hip, kind, W, H, seed):
    cfa, _ = synth.mosaic(W, H, seed, kind)
    dng = str(tmp_path / "in.dng")
    synth.write_dng(dng, cfa, kind)
    na, nb = synth.noise_params(2e-4, 1e-6)
    print(f"[gpu_graph] {kind} {W}x{H} seed {seed}", flush=True)
    g = pipe.Graph(lines=pipe.raw_graph_cfg(dng, noise_a=na, noise_b=nb))
    out = str(tmp_path / "out")
    g.export("o-pfm", out)
    g.run()
    img, hdr = pfm.read_pfm(out + ".pfm")
    assert hdr % 16 == 0
    assert img.shape == (H - H % (3 if kind == "xtrans" else 2), W - W % (3 if kind == "xtrans" else 2), 3)
    ip = g.img_param("i-raw", "main")
    p = gs.DevelopParams(kind=kind, noise_a=np.float32(na), noise_b=np.float32(nb),
                         cam=np.array(ip["cam"], np.float32).reshape(3, 3), wb=tuple(ip["wb"][:3]) + (1.0,))
    ref = op.develop(64, cfa[:img.shape[0], :img.shape[1]], gs.oracle_params(p))[..., :3]
    d = np.abs(img.astype(np.float64) - ref)
    psnr = 10 * np.log10(1.0 / max(float(np.mean(d ** 2)), 1e-30))
    frac = float(np.mean(d > 1e-3))
    print(f"[gpu_graph] {kind}: max abs {d.max():.3e} frac>1e-3 {frac:.2e} psnr {psnr:.1f} dB", flush=True)
    assert np.isfinite(img).all()
    assert psnr >= 60.0
    assert frac <= 1e-3
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
